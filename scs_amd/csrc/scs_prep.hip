// Batched per-query preprocessing on gfx950: up to 64 queries per pass.
//
// Replaces, for a whole batch at once, the per-query host BFS and the
// admission kernels of the reference: cal_query_dist (src/Graph.h:111-142),
// initialReductionRules (src/helpers.cc:207-248), CompressTask
// (src/helpers.cc:250-313), the Thrust offset scan (main.cu:208-211) and
// NeighborUpdate (src/helpers.cc:317-368).
//
// Query b of the batch owns bit b of a 64-bit mask per vertex:
//   elig[v]  queries whose BFS may enter v (core[v] > kl0_b).  The host
//            numbers the queries by ascending kl0, so elig[v] is the mask of
//            the first cnt(v) bits: one compare per query, no table.
//   vis[v]   queries whose BFS reached v (= v is kept for them).
//   fr / nf  queries for which v is in the current / next frontier.
// A level is bottom-up (pull): every vertex that some active query still
// wants ORs the frontier masks of its neighbours, stopping once it has every
// bit it wants.  No atomics, and the adjacency of a vertex is read once per
// level for all 64 queries.  Kept vertex v of query b gets local id
//   base[g][b] + rank of v among group g's vertices kept by b
// (g = v / 64, ranks by ballot), so local ids ascend with original ids as
// the reference's compaction order does.  The local degree and adjacency
// passes read the adjacency of each kept vertex once for every query of
// the batch that kept it, with one ballot per query per 64-entry chunk.
//
// Outputs per query b (device pointers in MsQuery): loff[K+1] (exclusive
// scan of local degrees, loff[K] = nnz), ldist[K], lcore[K], lorig[K] and
// ladj[nnz] (local ids, ascending per vertex).
//
// Every kernel adds the bytes its loads and stores name (each element once)
// to a counter: the algorithmic traffic of the batched pass.
#include "scs_internal.h"

#define MS_BITS 64

__device__ __forceinline__ u64 ms_or(u64 x) {
  u32 lo = (u32)x, hi = (u32)(x >> 32);
  for (int o = 32; o > 0; o >>= 1) {
    lo |= __shfl_xor(lo, o);
    hi |= __shfl_xor(hi, o);
  }
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 ms_sum(u64 x) {
  u32 lo = (u32)x, hi = (u32)(x >> 32);
  // 64-bit add through two 32-bit halves with carry
  for (int o = 32; o > 0; o >>= 1) {
    const u32 olo = __shfl_xor(lo, o), ohi = __shfl_xor(hi, o);
    const u32 nlo = lo + olo;
    hi = hi + ohi + (nlo < lo ? 1u : 0u);
    lo = nlo;
  }
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 ms_shfl64(u64 x, int src) {
  const u32 lo = __shfl((u32)x, src), hi = __shfl((u32)(x >> 32), src);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ void ms_count_bytes(unsigned long long* ctr, u64 b) {
  b = ms_sum(b);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(ctr, (unsigned long long)b);
}
__device__ __forceinline__ u64 ms_below() {
  const u32 lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// elig/vis/fr for every vertex; level flags cleared.
__global__ void __launch_bounds__(256) ms_init(u32 n, const u32* __restrict__ core,
                                               const MsBatch* __restrict__ B,
                                               u64* __restrict__ elig, u64* __restrict__ vis,
                                               u64* __restrict__ fr, u32* __restrict__ flags,
                                               u32 nflags, unsigned long long* bytes) {
  const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 nt = gridDim.x * blockDim.x;
  u64 nb = 0;
  for (u32 v = tid; v < n; v += nt) {
    const u32 c = core[v];
    u32 cnt = 0;
    const u32 nbq = B->nb;
    for (u32 b = 0; b < nbq; b++) cnt += B->kl0[b] < c ? 1u : 0u;
    elig[v] = cnt >= 64 ? ~0ull : ((1ull << cnt) - 1ull);
    vis[v] = 0;
    fr[v] = 0;
    nb += 4 + 24;
  }
  for (u32 i = tid; i < nflags; i += nt) flags[i] = 0;
  ms_count_bytes(bytes, nb);
}

// Query vertices: visited and in the first frontier, distance 0.
__global__ void ms_seed(const MsBatch* __restrict__ B, u64* __restrict__ vis,
                        u64* __restrict__ fr, u16* __restrict__ dist16) {
  const u32 b = threadIdx.x;
  if (b < B->nb) {
    const u32 q = B->q[b];
    atomicOr((unsigned long long*)&vis[q], (unsigned long long)(1ull << b));
    atomicOr((unsigned long long*)&fr[q], (unsigned long long)(1ull << b));
    dist16[(u64)q * MS_BITS + b] = 0;
  }
}

// One BFS level for every query whose depth bound exceeds `level` (act).
// Runs only if the previous level found something (flag_prev).  A 16-lane
// segment takes one vertex (four per wave) and reads 16 adjacency entries
// per step, stopping once it holds every bit it wants; vertices of degree
// above MS_HUB are taken by the whole wave.
#define MS_HUB 512

__device__ __forceinline__ u64 ms_seg_or(u64 x) {
  u32 lo = (u32)x, hi = (u32)(x >> 32);
  for (int o = 8; o > 0; o >>= 1) {
    lo |= __shfl_xor(lo, o, 16);
    hi |= __shfl_xor(hi, o, 16);
  }
  return ((u64)hi << 32) | lo;
}

__global__ void __launch_bounds__(256) ms_level(const i64* __restrict__ off,
                                                const int32_t* __restrict__ adj, u32 n, u32 level,
                                                u64 act, const u64* __restrict__ elig,
                                                u64* __restrict__ vis, const u64* __restrict__ fr,
                                                u64* __restrict__ nf, u16* __restrict__ dist16,
                                                const u32* __restrict__ flag_prev,
                                                u32* __restrict__ flag_cur,
                                                unsigned long long* bytes) {
  if (level > 0 && *flag_prev == 0) return;
  const u32 lane = threadIdx.x & 63, seg = lane >> 4, sl = lane & 15;
  const u64 wid = (blockIdx.x * (u64)blockDim.x + threadIdx.x) >> 6;
  const u64 nw = ((u64)gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  bool found = false;
  for (u64 base = wid * 4; base < n; base += nw * 4) {
    const u64 vv = base + seg;
    const bool in = vv < n;
    const u32 v = (u32)vv;
    u64 want = 0, visv = 0;
    i64 b0 = 0, e0 = 0;
    if (in) {
      visv = vis[v];
      want = elig[v] & act & ~visv;
      if (sl == 0) nbytes += 16;
    }
    if (want) {
      b0 = off[v];
      e0 = off[v + 1];
      if (sl == 0) nbytes += 16;
    }
    const bool big = want && (e0 - b0) > MS_HUB;
    bool active = want && !big;
    u64 acc = 0;
    i64 j = b0 + sl;
    while (__ballot(active)) {
      u64 x = 0;
      if (active && j < e0) {
        x = fr[(u32)adj[j]] & want;
        nbytes += 12;
      }
      acc |= ms_seg_or(x);
      j += 16;
      if (active && (acc == want || j - (i64)sl >= e0)) active = false;
    }
    // hubs: the whole wave, 64 entries per step
    u64 bm = __ballot(big && sl == 0);
    while (bm) {
      const int src = __ffsll((unsigned long long)bm) - 1;
      bm &= bm - 1;
      const u64 bwant = ms_shfl64(want, src);
      const i64 bb = (i64)ms_shfl64((u64)b0, src), be = (i64)ms_shfl64((u64)e0, src);
      u64 bacc = 0;
      for (i64 jj = bb; jj < be; jj += 64) {
        u64 x = 0;
        if (jj + lane < be) {
          x = fr[(u32)adj[jj + lane]] & bwant;
          nbytes += 12;
        }
        bacc |= ms_or(x);
        if (bacc == bwant) break;
      }
      // hand the result to the hub's segment
      if ((lane >> 4) == ((u32)src >> 4)) acc = bacc;
    }
    if (in) {
      if (sl == 0) {
        nf[v] = acc;
        nbytes += 8;
      }
      if (acc) {
        if (sl == 0) {
          vis[v] = visv | acc;
          nbytes += 8;
        }
        found = true;
        for (u32 b = sl; b < 64; b += 16)
          if ((acc >> b) & 1ull) {
            dist16[(u64)v * MS_BITS + b] = (u16)(level + 1);
            nbytes += 2;
          }
      }
    }
  }
  if (__ballot(found) && lane == 0) atomicOr(flag_cur, 1u);
  ms_count_bytes(bytes, nbytes);
}

// Per 64-vertex group g and query b: number of group vertices kept by b,
// stored at gcnt[g * 64 + b].
__global__ void __launch_bounds__(256) ms_count(u32 n, u32 G, const u64* __restrict__ vis,
                                                u64 valid, u32* __restrict__ gcnt,
                                                unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  for (u32 g = wid; g < G; g += nw) {
    const u64 v = (u64)g * 64 + lane;
    const u64 m = v < n ? vis[v] & valid : 0;
    if (v < n) nbytes += 8;
    u64 any = ms_or(m);
    u32 mine = 0;
    while (any) {
      const int b = __ffsll((unsigned long long)any) - 1;
      any &= any - 1;
      const u64 bal = __ballot((m >> b) & 1ull);
      if ((int)lane == b) mine = (u32)__popcll(bal);
    }
    gcnt[(u64)g * 64 + lane] = mine;
    nbytes += 4;
  }
  ms_count_bytes(bytes, nbytes);
}

// Exclusive block-wide scan of one value per thread (256 threads).
__device__ __forceinline__ u32 ms_block_excl(u32 s, u32* sh, u32* total) {
  const u32 t = threadIdx.x;
  sh[t] = s;
  __syncthreads();
  for (u32 o = 1; o < 256; o <<= 1) {
    const u32 y = t >= o ? sh[t - o] : 0u;
    __syncthreads();
    sh[t] += y;
    __syncthreads();
  }
  const u32 incl = sh[t];
  *total = sh[255];
  __syncthreads();
  return incl - s;
}

// One workgroup per query b: group counts -> exclusive bases in place;
// K_b -> tot[b].
__global__ void __launch_bounds__(256) ms_bscan(u32 G, u32* __restrict__ gcnt, u32* __restrict__ tot,
                                                unsigned long long* bytes) {
  __shared__ u32 sh[256];
  const u32 b = blockIdx.x, t = threadIdx.x;
  const u32 per = (G + 255) / 256;
  const u32 g0 = min(G, t * per), g1 = min(G, g0 + per);
  u32 s = 0;
  for (u32 g = g0; g < g1; g++) s += gcnt[(u64)g * 64 + b];
  u32 total;
  u32 run = ms_block_excl(s, sh, &total);
  for (u32 g = g0; g < g1; g++) {
    const u32 c = gcnt[(u64)g * 64 + b];
    gcnt[(u64)g * 64 + b] = run;
    run += c;
  }
  if (t == 0) tot[b] = total;
  ms_count_bytes(bytes, (u64)(g1 - g0) * 12);
}

// Local ids, and per kept vertex: orig id, distance, core number.
__global__ void __launch_bounds__(256) ms_scatter(u32 n, u32 G, const u64* __restrict__ vis,
                                                  u64 valid, const u32* __restrict__ gcnt,
                                                  const u32* __restrict__ core,
                                                  const u16* __restrict__ dist16,
                                                  const MsBatch* __restrict__ B,
                                                  u32* __restrict__ locid, u32* __restrict__ qloc,
                                                  unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  const u64 below = ms_below();
  u64 nbytes = 0;
  for (u32 g = wid; g < G; g += nw) {
    const u64 v = (u64)g * 64 + lane;
    const u64 m = v < n ? vis[v] & valid : 0;
    u32 cv = 0;
    if (m) {
      cv = core[v];
      nbytes += 12;
    }
    u64 any = ms_or(m);
    while (any) {
      const int b = __ffsll((unsigned long long)any) - 1;
      any &= any - 1;
      const u64 bal = __ballot((m >> b) & 1ull);
      if ((m >> b) & 1ull) {
        const u32 id = gcnt[(u64)g * 64 + b] + (u32)__popcll(bal & below);
        const MsQuery Q = B->P[b];
        Q.lorig[id] = (u32)v;
        Q.ldist[id] = dist16[v * MS_BITS + b];
        Q.lcore[id] = cv;
        locid[v * MS_BITS + b] = id;
        if ((u32)v == B->q[b]) qloc[b] = id;
        nbytes += 4 + 2 + 12 + 4;
      }
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// Local degree of every (query, kept vertex) into loff[id] (scanned later).
// One wave per kept vertex; lane b counts query b's kept neighbours.
__global__ void __launch_bounds__(256) ms_degree(const i64* __restrict__ off,
                                                 const int32_t* __restrict__ adj, u32 n,
                                                 const u64* __restrict__ vis, u64 valid,
                                                 const u32* __restrict__ locid,
                                                 const MsBatch* __restrict__ B,
                                                 unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  for (u32 u = wid; u < n; u += nw) {
    const u64 mu = vis[u] & valid;
    if (lane == 0) nbytes += 8;
    if (!mu) continue;
    const i64 b0 = off[u], e0 = off[u + 1];
    if (lane == 0) nbytes += 16;
    u32 cnt = 0;
    for (i64 j = b0; j < e0; j += 64) {
      u64 mw = 0;
      if (j + lane < e0) {
        mw = vis[(u32)adj[j + lane]] & mu;
        nbytes += 12;
      }
      u64 any = ms_or(mw);
      while (any) {
        const int b = __ffsll((unsigned long long)any) - 1;
        any &= any - 1;
        const u64 bal = __ballot((mw >> b) & 1ull);
        if ((int)lane == b) cnt += (u32)__popcll(bal);
      }
    }
    if ((mu >> lane) & 1ull) {
      const u32 id = locid[(u64)u * MS_BITS + lane];
      B->P[lane].loff[id] = cnt;
      nbytes += 8;
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// One workgroup per query: loff[0..K) exclusive scan in place,
// loff[K] = nnz -> tot[64 + b].
__global__ void __launch_bounds__(256) ms_scan_loff(const MsBatch* __restrict__ B,
                                                    u32* __restrict__ tot, u64 valid,
                                                    unsigned long long* bytes) {
  __shared__ u32 sh[256];
  const u32 b = blockIdx.x, t = threadIdx.x;
  if (!((valid >> b) & 1ull)) return;
  u32* a = B->P[b].loff;
  const u32 K = tot[b];
  u32 carry = 0;
  for (u32 base = 0; base < K; base += 1024) {
    const u32 i0 = base + t * 4;
    u32 x[4];
    u32 s = 0;
    for (int k = 0; k < 4; k++) {
      x[k] = i0 + k < K ? a[i0 + k] : 0u;
      s += x[k];
    }
    u32 tile;
    u32 run = carry + ms_block_excl(s, sh, &tile);
    for (int k = 0; k < 4; k++) {
      if (i0 + k < K) a[i0 + k] = run;
      run += x[k];
    }
    carry += tile;
  }
  if (t == 0) {
    a[K] = carry;
    tot[64 + b] = carry;
  }
  ms_count_bytes(bytes, t == 0 ? (u64)K * 8 + 4 : 0);
}

// Local adjacency: for every (query, kept vertex u), the local ids of u's
// kept neighbours in ascending order at ladj[loff[id(u)] ..].
__global__ void __launch_bounds__(256) ms_fill(const i64* __restrict__ off,
                                               const int32_t* __restrict__ adj, u32 n,
                                               const u64* __restrict__ vis, u64 valid,
                                               const u32* __restrict__ locid,
                                               const MsBatch* __restrict__ B,
                                               unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  const u64 below = ms_below();
  u64 nbytes = 0;
  for (u32 u = wid; u < n; u += nw) {
    const u64 mu = vis[u] & valid;
    if (lane == 0) nbytes += 8;
    if (!mu) continue;
    const i64 b0 = off[u], e0 = off[u + 1];
    if (lane == 0) nbytes += 16;
    u32 start = 0;  // lane b: where query b's list of u starts
    if ((mu >> lane) & 1ull) {
      start = B->P[lane].loff[locid[(u64)u * MS_BITS + lane]];
      nbytes += 8;
    }
    u32 run = 0;  // lane b: entries written so far for query b
    for (i64 j = b0; j < e0; j += 64) {
      u64 mw = 0;
      u32 w = 0;
      if (j + lane < e0) {
        w = (u32)adj[j + lane];
        mw = vis[w] & mu;
        nbytes += 12;
      }
      u64 any = ms_or(mw);
      while (any) {
        const int b = __ffsll((unsigned long long)any) - 1;
        any &= any - 1;
        const u64 bal = __ballot((mw >> b) & 1ull);
        const u32 pos0 = __shfl(start + run, b);
        if ((mw >> b) & 1ull) {
          B->P[b].ladj[pos0 + (u32)__popcll(bal & below)] = locid[(u64)w * MS_BITS + b];
          nbytes += 8;
        }
        if ((int)lane == b) run += (u32)__popcll(bal);
      }
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// ---------------------------------------------------------------------------
// Launchers.  B points to the batch header in device memory.
// ---------------------------------------------------------------------------
static u32 ms_blocks(u64 threads, u32 cap) {
  u64 b = (threads + 255) / 256;
  if (b == 0) b = 1;
  return (u32)(b < cap ? b : cap);
}

extern "C" hipError_t scs_launch_ms_init(u32 n, const u32* core, const MsBatch* B, u64* elig,
                                         u64* vis, u64* fr, u32* flags, u32 nflags,
                                         unsigned long long* bytes, hipStream_t s) {
  ms_init<<<ms_blocks(n, 8192), 256, 0, s>>>(n, core, B, elig, vis, fr, flags, nflags, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_seed(const MsBatch* B, u64* vis, u64* fr, u16* dist16,
                                         hipStream_t s) {
  ms_seed<<<1, 64, 0, s>>>(B, vis, fr, dist16);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_level(const i64* off, const int32_t* adj, u32 n, u32 level,
                                          u64 act, const u64* elig, u64* vis, const u64* fr,
                                          u64* nf, u16* dist16, const u32* flag_prev,
                                          u32* flag_cur, unsigned long long* bytes,
                                          hipStream_t s) {
  ms_level<<<ms_blocks(((u64)n + 3) / 4 * 64, 16384), 256, 0, s>>>(
      off, adj, n, level, act, elig, vis, fr, nf, dist16, flag_prev, flag_cur, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_count(u32 n, u32 G, const u64* vis, u64 valid, u32* gcnt,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_count<<<ms_blocks((u64)G * 64, 8192), 256, 0, s>>>(n, G, vis, valid, gcnt, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_bscan(u32 nb, u32 G, u32* gcnt, u32* tot,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_bscan<<<nb, 256, 0, s>>>(G, gcnt, tot, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_scatter(u32 n, u32 G, const u64* vis, u64 valid,
                                            const u32* gcnt, const u32* core, const u16* dist16,
                                            const MsBatch* B, u32* locid, u32* qloc,
                                            unsigned long long* bytes, hipStream_t s) {
  ms_scatter<<<ms_blocks((u64)G * 64, 8192), 256, 0, s>>>(n, G, vis, valid, gcnt, core, dist16, B,
                                                          locid, qloc, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_degree(const i64* off, const int32_t* adj, u32 n,
                                           const u64* vis, u64 valid, const u32* locid,
                                           const MsBatch* B, unsigned long long* bytes,
                                           hipStream_t s) {
  ms_degree<<<ms_blocks((u64)n * 64, 16384), 256, 0, s>>>(off, adj, n, vis, valid, locid, B, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_scan_loff(u32 nb, const MsBatch* B, u32* tot, u64 valid,
                                              unsigned long long* bytes, hipStream_t s) {
  ms_scan_loff<<<nb, 256, 0, s>>>(B, tot, valid, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_fill(const i64* off, const int32_t* adj, u32 n, const u64* vis,
                                         u64 valid, const u32* locid, const MsBatch* B,
                                         unsigned long long* bytes, hipStream_t s) {
  ms_fill<<<ms_blocks((u64)n * 64, 16384), 256, 0, s>>>(off, adj, n, vis, valid, locid, B, bytes);
  return hipGetLastError();
}
