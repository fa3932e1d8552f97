// Graph-level kernels for gfx950: core decomposition, once per graph,
// replacing the reference's host-side linear-heap peeling
// (src/Graph.h:82-109).  Per-query preprocessing is batched in
// scs_prep.hip.
#include "scs_internal.h"

#include <mutex>

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

// ---------------------------------------------------------------------------
// Core numbers without allocation, for callers that cannot allocate device
// memory.  The working memory (degrees, two frontiers, three counters) is
// this fixed device array; the host finds its per-device address once with
// hipGetSymbolAddress and passes that address to the kernels.
// Layout (words): deg[CORE_NA_MAX_N] | F[..] | Fn[..] | counters[64].
// ---------------------------------------------------------------------------
#define CORE_NA_MAX_N (1u << 20)
static __device__ __align__(256) u32 scratch[3u * CORE_NA_MAX_N + 64];

#define CORE_NA_DEVICES 64
static u32* g_scratch_dev[CORE_NA_DEVICES];
static std::mutex g_scratch_mu;

// cnt[0..3) = a, b, c (a kernel, not a memset: callers may capture this)
__global__ void core_counters(u32* cnt, u32 a, u32 b, u32 c) {
  if (threadIdx.x == 0) {
    cnt[0] = a;
    cnt[1] = b;
    cnt[2] = c;
  }
}

static int scratch_address(u32** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= CORE_NA_DEVICES) return SCS_EHIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SCS_EHIP;
    g_scratch_dev[dev] = (u32*)p;
  }
  *out = g_scratch_dev[dev];
  return SCS_OK;
}

#define NA_CHK(x)                          \
  do {                                     \
    if ((x) != hipSuccess) return SCS_EHIP; \
  } while (0)

extern "C" int scs_core_numbers_noalloc(const int64_t* d_off, const int32_t* d_adj, uint32_t n,
                                        uint32_t* d_core, void* stream) {
  if (!d_off || !d_core || (n && !d_adj)) return SCS_EINVAL;
  if (n > CORE_NA_MAX_N) return SCS_EBIG;
  if (n == 0) return SCS_OK;
  u32* S = nullptr;
  const int rc = scratch_address(&S);
  if (rc) return rc;
  u32* deg = S;
  u32* F = S + CORE_NA_MAX_N;
  u32* Fn = S + 2u * CORE_NA_MAX_N;
  u32* cnt = S + 3u * CORE_NA_MAX_N;
  hipStream_t s = (hipStream_t)stream;
  NA_CHK(scs_launch_core_init(d_off, n, deg, d_core, s));
  u32 peeled = 0, k = 0;
  while (peeled < n) {
    // cnt[0] = frontier size, cnt[1] = min remaining degree
    core_counters<<<1, 64, 0, s>>>(cnt, 0u, 0xffffffffu, 0u);
    NA_CHK(hipGetLastError());
    NA_CHK(scs_launch_core_scan(n, k, deg, d_core, F, cnt, cnt + 1, s));
    u32 h[2];
    NA_CHK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    NA_CHK(hipStreamSynchronize(s));
    u32 fn = h[0];
    if (fn == 0) {
      if (h[1] == 0xffffffffu) break;
      k = h[1] > k + 1 ? h[1] : k + 1;
      continue;
    }
    peeled += fn;
    u32 *cur = F, *nxt = Fn;
    while (fn) {
      core_counters<<<1, 64, 0, s>>>(cnt, 0u, 0xffffffffu, 0u);
      NA_CHK(hipGetLastError());
      NA_CHK(scs_launch_core_peel(d_off, d_adj, k, cur, fn, deg, d_core, nxt, cnt + 2, s));
      u32 f2 = 0;
      NA_CHK(hipMemcpyAsync(&f2, cnt + 2, sizeof(u32), hipMemcpyDeviceToHost, s));
      NA_CHK(hipStreamSynchronize(s));
      fn = f2;
      peeled += fn;
      u32* t = cur;
      cur = nxt;
      nxt = t;
    }
    k++;
  }
  return SCS_OK;
}
