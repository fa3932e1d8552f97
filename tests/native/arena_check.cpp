// Randomised check of the device-arena allocator against a byte map:
// every allocation is in bounds and disjoint from every live one, releases
// coalesce back to one free range, and releases of ranges that were never
// handed out (stale offsets, double releases) are refused without changing
// the free list.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "scs_arena.h"

#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
      return 1;                                               \
    }                                                         \
  } while (0)

int main() {
  const uint64_t size = 1 << 20;
  scs::Arena a;
  a.init(nullptr, size);
  std::vector<uint8_t> owner(size / 256, 0);  // 0 free, 1 live
  struct Blk {
    uint64_t off, len;
  };
  std::vector<Blk> live;
  std::mt19937_64 rng(7);
  uint64_t fails = 0;
  for (int it = 0; it < 200000; it++) {
    const int op = (int)(rng() % 10);
    if (op < 5) {
      const uint64_t len = 1 + rng() % 40000;
      uint64_t off = ~0ull;
      if (a.alloc(len, &off)) {
        const uint64_t g = scs::Arena::granule(len);
        CHECK(off % 256 == 0 && off + g <= size);
        for (uint64_t b = off / 256; b < (off + g) / 256; b++) {
          CHECK(owner[b] == 0);
          owner[b] = 1;
        }
        live.push_back({off, len});
      } else {
        fails++;
      }
    } else if (op < 8 && !live.empty()) {
      const size_t k = rng() % live.size();
      const Blk b = live[k];
      live[k] = live.back();
      live.pop_back();
      CHECK(a.release(b.off, b.len));
      for (uint64_t x = b.off / 256; x < (b.off + scs::Arena::granule(b.len)) / 256; x++) owner[x] = 0;
      // a second release of the same range is refused
      const uint64_t bad = a.bad_releases;
      bool overlaps_live = false;
      for (uint64_t x = b.off / 256; x < (b.off + scs::Arena::granule(b.len)) / 256; x++)
        overlaps_live |= owner[x] != 0;
      if (!overlaps_live) {
        CHECK(!a.release(b.off, b.len));
        CHECK(a.bad_releases == bad + 1);
      }
    } else {
      // a release of a stale offset that is free refuses; the free list and
      // the live blocks stay intact (checked by the byte map on later allocs)
      const uint64_t off = (rng() % (size / 256)) * 256;
      if (owner[off / 256] == 0) CHECK(!a.release(off, 256));
    }
    uint64_t used = 0;
    for (auto& b : live) used += scs::Arena::granule(b.len);
    CHECK(a.used == used);
  }
  for (auto& b : live) CHECK(a.release(b.off, b.len));
  CHECK(a.empty());
  CHECK(a.free_.size() == 1 && a.free_.begin()->first == 0 && a.free_.begin()->second == size);
  CHECK(!a.release(size - 256, 512));  // out of bounds
  printf("arena ok: %llu failed allocations\n", (unsigned long long)fails);
  return 0;
}
