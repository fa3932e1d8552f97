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

// Per-thread scratch of the heuristic; n-sized arrays are reused across
// queries and only touched entries are cleared.
struct HeuWork {
  std::vector<uint8_t> sta;  // bit 7 = touched, low bits = reference status
  std::vector<uint32_t> deg, cnt;
  std::vector<std::vector<uint32_t>> thn;
  std::vector<double> key3;  // heu3: largest score pushed for each vertex
  std::vector<uint8_t> inS;
  std::vector<uint32_t> touched;
  void reset(uint32_t n);
  void clear_touched();
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
