// Host-side structures of the SCS library (not part of the C ABI).
#pragma once
#include <cstdint>
#include <vector>

namespace scs {

struct HostGraph {
  uint32_t n = 0;
  uint32_t dmax = 0;
  std::vector<int64_t> off;  // n+1
  std::vector<int32_t> adj;  // off[n]
  std::vector<uint32_t> core;
};

// heu3's state of one vertex, in one 32-byte record so that a visit of a
// neighbour touches one cache line: the largest score pushed for it, and its
// H-neighbours in ascending order (the first three in place; a longer list
// at offset ext of the workspace's pool, in room for max(8, the next power of
// two) entries).
struct Heu3Rec {
  double key;
  uint64_t ext;
  uint32_t len;
  uint32_t inl[3];
};

// Per-thread scratch of the heuristic; n-sized arrays are reused across
// queries and only touched entries are cleared.
struct HeuWork {
  std::vector<uint8_t> sta;  // bit 7 = touched, low bits = reference status
  std::vector<uint32_t> deg, cnt;
  std::vector<Heu3Rec> rec3;
  std::vector<uint32_t> pool3;  // heu3: the longer H-neighbour lists of one run
  size_t top3 = 0;              // first free entry of pool3
  std::vector<uint8_t> inS;
  std::vector<uint32_t> touched;
  void reset(uint32_t n);
  // Zeroes sta and the arrays named in `what` at every touched vertex; each
  // construction names only the arrays it wrote at vertices outside its H
  // list (heu2 and heu3 zero deg over their H list themselves).
  enum : unsigned { CLEAR_CNT = 1, CLEAR_REC3 = 2, CLEAR_DEG_INS = 4 };
  void clear_touched(unsigned what);
  bool mark_touched(uint32_t v);
};

uint32_t cssc_heu(const HostGraph& g, uint32_t l, uint32_t h, uint32_t q, HeuWork& W,
                  std::vector<uint32_t>& H);

// Diameter bound used at admission (main.cu:154-172).
inline uint32_t ubd_host(uint32_t kl, uint32_t h) {
  if (kl <= 1) return h - 1;
  for (uint32_t d = 1; d <= h; d++) {
    if (d <= 2) {
      if (kl + d > h) return d - 1;
    } else {
      uint32_t min_n = kl + d + 1 + (d / 3) * (kl - 2);
      if (h < min_n) return d - 1;
    }
  }
  return h - 1;
}

}  // namespace scs
