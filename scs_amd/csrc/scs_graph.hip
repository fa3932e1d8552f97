// Graph-level and per-query preprocessing kernels for gfx950.
//
// Core decomposition (once per graph) replaces the reference's host-side
// linear-heap peeling (src/Graph.h:82-109).  Per-query preprocessing
// replaces cal_query_dist (src/Graph.h:111-142), initialReductionRules
// (src/helpers.cc:207-248), CompressTask (src/helpers.cc:250-313), the
// thrust::inclusive_scan of main.cu:208-211 and NeighborUpdate
// (src/helpers.cc:317-368): a BFS from q bounded by depth ubD and by
// core > kl0, a compaction to local ids in ascending original-id order, and
// a compacted per-query CSR whose adjacency is ascending and deterministic.
#include "scs_internal.h"

#define BLK 256

// ---------------------------------------------------------------------------
// Core decomposition: level-synchronous peeling.  At level k the vertices
// still alive with degree <= k are peeled (core = k); every decrement that
// takes a neighbour from k+1 to k appends it to the next frontier of the
// same level.  Core numbers are unique, so the result equals any exact
// sequential peeling.
// ---------------------------------------------------------------------------
__global__ void core_init(const i64* __restrict__ off, u32 n, u32* __restrict__ deg,
                          u32* __restrict__ core) {
  for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    deg[v] = (u32)(off[v + 1] - off[v]);
    core[v] = 0xffffffffu;
  }
}

// Appends alive vertices with deg <= k to F (sets core = k) and computes the
// minimum degree among the remaining alive vertices.
__global__ void core_scan(u32 n, u32 k, const u32* __restrict__ deg, u32* __restrict__ core,
                          u32* __restrict__ F, u32* __restrict__ fcount, u32* __restrict__ minrest) {
  u32 mn = 0xffffffffu;
  for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    if (core[v] != 0xffffffffu) continue;
    u32 d = deg[v];
    if (d <= k) {
      core[v] = k;
      F[atomicAdd(fcount, 1u)] = v;
    } else {
      mn = min(mn, d);
    }
  }
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, (u32)__shfl_xor((int)mn, o));
  if ((threadIdx.x & 63) == 0 && mn != 0xffffffffu) atomicMin(minrest, mn);
}

// One wave per frontier vertex; decrements alive neighbours.
__global__ void core_peel(const i64* __restrict__ off, const int32_t* __restrict__ adj, u32 k,
                          const u32* __restrict__ F, u32 fn, u32* __restrict__ deg,
                          u32* __restrict__ core, u32* __restrict__ Fn, u32* __restrict__ fncount) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  for (u32 i = wid; i < fn; i += nw) {
    u32 v = F[i];
    i64 b = off[v], e = off[v + 1];
    for (i64 j = b + lane; j < e; j += 64) {
      u32 w = (u32)adj[j];
      if (core[w] != 0xffffffffu) continue;
      u32 old = atomicSub(&deg[w], 1u);
      if (old == k + 1) {
        core[w] = k;
        Fn[atomicAdd(fncount, 1u)] = w;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of u32 values produced by a functor, in three passes.
// ---------------------------------------------------------------------------
#define SCAN_T 256
#define SCAN_ITEMS 8
#define SCAN_TILE (SCAN_T * SCAN_ITEMS)

template <typename F>
__device__ inline u32 block_scan_tile(F f, u32 base, u32 n, u32* vals, u32* total) {
  __shared__ u32 wsum[SCAN_T / 64];
  u32 local[SCAN_ITEMS];
  u32 s = 0;
  const u32 t = threadIdx.x;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    u32 idx = base + t * SCAN_ITEMS + i;
    u32 x = idx < n ? f(idx) : 0u;
    local[i] = s;
    s += x;
  }
  // inclusive scan of s across the block
  u32 lane = t & 63, w = t >> 6;
  u32 incl = s;
  for (int o = 1; o < 64; o <<= 1) {
    u32 y = __shfl_up(incl, o);
    if (lane >= (u32)o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  u32 woff = 0, tot = 0;
  for (u32 i = 0; i < SCAN_T / 64; i++) {
    if (i < w) woff += wsum[i];
    tot += wsum[i];
  }
  u32 excl = woff + incl - s;
  for (int i = 0; i < SCAN_ITEMS; i++) vals[i] = excl + local[i];
  *total = tot;
  __syncthreads();
  return excl;
}

struct KeptFlag {
  const u32* stamp;
  u32 id;
  __device__ u32 operator()(u32 v) const { return stamp[v] == id ? 1u : 0u; }
};
struct LoadU32 {
  const u32* a;
  __device__ u32 operator()(u32 i) const { return a[i]; }
};

template <typename F>
__global__ void scan_reduce(F f, u32 n, u32* __restrict__ bsum) {
  u32 vals[SCAN_ITEMS], tot;
  block_scan_tile(f, blockIdx.x * SCAN_TILE, n, vals, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place; total to *out_total
__global__ void scan_blocksums(u32* __restrict__ bsum, u32 nb, u32* __restrict__ out_total) {
  __shared__ u32 carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (u32 base = 0; base < nb; base += SCAN_TILE) {
    u32 vals[SCAN_ITEMS], tot;
    LoadU32 ld{bsum};
    u32 c = carry;
    block_scan_tile(ld, base, nb, vals, &tot);
    for (int i = 0; i < SCAN_ITEMS; i++) {
      u32 idx = base + threadIdx.x * SCAN_ITEMS + i;
      if (idx < nb) bsum[idx] = vals[i] + c;
    }
    __syncthreads();
    if (threadIdx.x == 0) carry = c + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_total = carry;
}

template <typename F>
__global__ void scan_apply(F f, u32 n, const u32* __restrict__ bsum, u32* __restrict__ out) {
  u32 vals[SCAN_ITEMS], tot;
  u32 base = blockIdx.x * SCAN_TILE;
  block_scan_tile(f, base, n, vals, &tot);
  u32 add = bsum[blockIdx.x];
  for (int i = 0; i < SCAN_ITEMS; i++) {
    u32 idx = base + threadIdx.x * SCAN_ITEMS + i;
    if (idx < n) out[idx] = vals[i] + add;
  }
}

// ---------------------------------------------------------------------------
// Per-query BFS.  stamp[v] == id marks v visited by the current query (no
// per-query clearing of n-sized arrays).  Only vertices with core > kl0 are
// entered; vertices at depth ubD are not expanded.
// ---------------------------------------------------------------------------
__global__ void bfs_seed(u32 q, u32 id, u32* __restrict__ stamp, u32* __restrict__ dist,
                         u32* __restrict__ F, u32* __restrict__ fcount) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    stamp[q] = id;
    dist[q] = 0;
    F[0] = q;
    *fcount = 1;
  }
}

// Small-degree vertices: one lane each; large ones: the whole wave.
__global__ void bfs_level(const i64* __restrict__ off, const int32_t* __restrict__ adj,
                          const u32* __restrict__ core, u32 kl0, u32 id, u32 level,
                          const u32* __restrict__ F, u32 fn, u32* __restrict__ stamp,
                          u32* __restrict__ dist, u32* __restrict__ Fn, u32* __restrict__ fncount) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  for (u32 i0 = wid * 64; i0 < fn; i0 += nw * 64) {
    // each lane owns one frontier vertex; big ones are shared by the wave
    u32 i = i0 + lane;
    u32 v = i < fn ? F[i] : 0xffffffffu;
    i64 b = 0, e = 0;
    if (v != 0xffffffffu) {
      b = off[v];
      e = off[v + 1];
    }
    bool big = (e - b) > 64;
    if (!big) {
      for (i64 j = b; j < e; j++) {
        u32 w = (u32)adj[j];
        if (core[w] <= kl0 || stamp[w] == id) continue;
        if (atomicExch(&stamp[w], id) != id) {
          dist[w] = level + 1;
          Fn[atomicAdd(fncount, 1u)] = w;
        }
      }
    }
    unsigned long long bm = __ballot(big);
    while (bm) {
      int src = __ffsll((long long)bm) - 1;
      bm &= bm - 1;
      u32 bv = __shfl(v, src);
      i64 bb = off[bv], be = off[bv + 1];
      for (i64 j = bb + lane; j < be; j += 64) {
        u32 w = (u32)adj[j];
        if (core[w] <= kl0 || stamp[w] == id) continue;
        if (atomicExch(&stamp[w], id) != id) {
          dist[w] = level + 1;
          Fn[atomicAdd(fncount, 1u)] = w;
        }
      }
    }
  }
}

// Kept vertices in local order: korder[locid[v]] = v.
__global__ void kept_order(u32 n, const u32* __restrict__ stamp, u32 id,
                           const u32* __restrict__ locid, u32* __restrict__ korder) {
  for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (stamp[v] == id) korder[locid[v]] = v;
}

// Local degree (neighbours inside K), one wave per kept vertex.
__global__ void local_degree(const i64* __restrict__ off, const int32_t* __restrict__ adj,
                             const u32* __restrict__ stamp, u32 id, const u32* __restrict__ korder,
                             u32 K, u32* __restrict__ ldeg) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  for (u32 i = wid; i < K; i += nw) {
    u32 v = korder[i];
    i64 b = off[v], e = off[v + 1];
    u32 c = 0;
    for (i64 j = b + lane; j < e; j += 64) c += (stamp[(u32)adj[j]] == id);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) ldeg[i] = c;
  }
}

// Local adjacency in ascending order (ballot compaction keeps the order),
// plus distance, core and original id per local vertex.
__global__ void local_fill(const i64* __restrict__ off, const int32_t* __restrict__ adj,
                           const u32* __restrict__ stamp, u32 id, const u32* __restrict__ locid,
                           const u32* __restrict__ korder, u32 K, const u32* __restrict__ loff,
                           const u32* __restrict__ dist, const u32* __restrict__ core,
                           u32* __restrict__ ladj, u32* __restrict__ ldist, u32* __restrict__ lcore,
                           u32* __restrict__ lorig) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (u32 i = wid; i < K; i += nw) {
    u32 v = korder[i];
    i64 b = off[v], e = off[v + 1];
    u32 wpos = loff[i];
    for (i64 j0 = b; j0 < e; j0 += 64) {
      i64 j = j0 + lane;
      u32 w = 0;
      bool keep = false;
      if (j < e) {
        w = (u32)adj[j];
        keep = stamp[w] == id;
      }
      unsigned long long bm = __ballot(keep);
      if (keep) ladj[wpos + __popcll(bm & below)] = locid[w];
      wpos += __popcll(bm);
    }
    if (lane == 0) {
      ldist[i] = dist[v];
      lcore[i] = core[v];
      lorig[i] = v;
    }
  }
}

// Algorithmic bytes of one query's preprocessing (for the roofline figure):
// visited vertices' offsets, adjacency and per-neighbour core/stamp gathers,
// plus the kept vertices' second and third adjacency passes and writes.
__global__ void prep_bytes(const i64* __restrict__ off, const u32* __restrict__ korder, u32 K,
                           unsigned long long* __restrict__ out) {
  unsigned long long s = 0;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < K; i += gridDim.x * blockDim.x) {
    u32 v = korder[i];
    unsigned long long d = (unsigned long long)(off[v + 1] - off[v]);
    // BFS: 16 B offsets + 4 B adj + 4 B core + 4 B stamp per neighbour + 8 B dist/stamp write
    // degree pass: 16 B offsets + 8 B per neighbour (adj + stamp)
    // fill pass:   16 B offsets + 8 B per neighbour + 4 B locid per kept neighbour (bounded by d)
    s += 16 + 12 * d + 8 + 16 + 8 * d + 16 + 12 * d + 16;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

// ---------------------------------------------------------------------------
// Host-side launch helpers (C linkage inside the library).
// ---------------------------------------------------------------------------
static inline u32 nblocks(u64 work, u32 per, u32 cap) {
  u64 b = (work + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (u32)b;
}

extern "C" hipError_t scs_launch_core_init(const i64* off, u32 n, u32* deg, u32* core,
                                           hipStream_t s) {
  core_init<<<nblocks(n, BLK, 4096), BLK, 0, s>>>(off, n, deg, core);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_core_scan(u32 n, u32 k, const u32* deg, u32* core, u32* F,
                                           u32* fcount, u32* minrest, hipStream_t s) {
  core_scan<<<nblocks(n, BLK, 4096), BLK, 0, s>>>(n, k, deg, core, F, fcount, minrest);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_core_peel(const i64* off, const int32_t* adj, u32 k, const u32* F,
                                           u32 fn, u32* deg, u32* core, u32* Fn, u32* fncount,
                                           hipStream_t s) {
  core_peel<<<nblocks((u64)fn * 64, BLK, 8192), BLK, 0, s>>>(off, adj, k, F, fn, deg, core, Fn,
                                                             fncount);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_bfs_seed(u32 q, u32 id, u32* stamp, u32* dist, u32* F,
                                          u32* fcount, hipStream_t s) {
  bfs_seed<<<1, 64, 0, s>>>(q, id, stamp, dist, F, fcount);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_bfs_level(const i64* off, const int32_t* adj, const u32* core,
                                           u32 kl0, u32 id, u32 level, const u32* F, u32 fn,
                                           u32* stamp, u32* dist, u32* Fn, u32* fncount,
                                           hipStream_t s) {
  bfs_level<<<nblocks(fn, BLK, 8192), BLK, 0, s>>>(off, adj, core, kl0, id, level, F, fn, stamp,
                                                    dist, Fn, fncount);
  return hipGetLastError();
}
// exclusive scan of kept flags over n -> locid; total -> *K
extern "C" hipError_t scs_launch_kept_scan(u32 n, const u32* stamp, u32 id, u32* bsum, u32* locid,
                                           u32* total, hipStream_t s) {
  u32 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0) nb = 1;
  KeptFlag f{stamp, id};
  scan_reduce<<<nb, SCAN_T, 0, s>>>(f, n, bsum);
  scan_blocksums<<<1, SCAN_T, 0, s>>>(bsum, nb, total);
  scan_apply<<<nb, SCAN_T, 0, s>>>(f, n, bsum, locid);
  return hipGetLastError();
}
// exclusive scan of a u32 array of length n -> out (may alias); total -> *total
extern "C" hipError_t scs_launch_scan_u32(const u32* in, u32 n, u32* bsum, u32* out, u32* total,
                                          hipStream_t s) {
  u32 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0) nb = 1;
  LoadU32 f{in};
  scan_reduce<<<nb, SCAN_T, 0, s>>>(f, n, bsum);
  scan_blocksums<<<1, SCAN_T, 0, s>>>(bsum, nb, total);
  scan_apply<<<nb, SCAN_T, 0, s>>>(f, n, bsum, out);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_kept_order(u32 n, const u32* stamp, u32 id, const u32* locid,
                                            u32* korder, hipStream_t s) {
  kept_order<<<nblocks(n, BLK, 4096), BLK, 0, s>>>(n, stamp, id, locid, korder);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_local_degree(const i64* off, const int32_t* adj, const u32* stamp,
                                              u32 id, const u32* korder, u32 K, u32* ldeg,
                                              hipStream_t s) {
  local_degree<<<nblocks((u64)K * 64, BLK, 8192), BLK, 0, s>>>(off, adj, stamp, id, korder, K,
                                                               ldeg);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_local_fill(const i64* off, const int32_t* adj, const u32* stamp,
                                            u32 id, const u32* locid, const u32* korder, u32 K,
                                            const u32* loff, const u32* dist, const u32* core,
                                            u32* ladj, u32* ldist, u32* lcore, u32* lorig,
                                            hipStream_t s) {
  local_fill<<<nblocks((u64)K * 64, BLK, 8192), BLK, 0, s>>>(off, adj, stamp, id, locid, korder, K,
                                                             loff, dist, core, ladj, ldist, lcore,
                                                             lorig);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_prep_bytes(const i64* off, const u32* korder, u32 K,
                                            unsigned long long* out, hipStream_t s) {
  prep_bytes<<<nblocks(K, BLK, 1024), BLK, 0, s>>>(off, korder, K, out);
  return hipGetLastError();
}
