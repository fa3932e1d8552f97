// First-fit allocator over the device arena that holds the compacted local
// graphs of resident queries (256-byte granules).  Host-only bookkeeping:
// it hands out offsets into one device allocation made at engine creation.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace scs {

struct Arena {
  uint8_t* base = nullptr;
  uint64_t size = 0;
  uint64_t used = 0;
  uint64_t bad_releases = 0;  // releases refused as never handed out
  std::map<uint64_t, uint64_t> free_;  // offset -> length, disjoint, coalesced

  static uint64_t granule(uint64_t bytes) {
    bytes = (bytes + 255) & ~255ull;
    return bytes ? bytes : 256;
  }
  void init(uint8_t* b, uint64_t s) {
    base = b;
    size = s & ~255ull;
    used = 0;
    bad_releases = 0;
    free_.clear();
    if (size) free_[0] = size;
  }
  bool empty() const { return used == 0; }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = granule(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes;
        const uint64_t noff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[noff] = rest;
        used += bytes;
        return true;
      }
    }
    return false;
  }
  // Returns a range to the free list.  A range that is out of bounds or
  // overlaps free memory was never handed out: it is refused and counted.
  bool release(uint64_t off, uint64_t bytes) {
    bytes = granule(bytes);
    if (off + bytes > size || off + bytes < off) {
      bad_releases++;
      return false;
    }
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && nx->first < off + bytes) {
      bad_releases++;
      return false;
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second > off) {
        bad_releases++;
        return false;
      }
    }
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    return true;
  }
};

}  // namespace scs
