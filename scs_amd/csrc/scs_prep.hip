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
// Level 0 pushes from the query vertices (top-down).  Every later level is
// bottom-up (pull): every vertex that some active query still wants ORs the
// frontier masks of its neighbours, stopping once it has every bit it
// wants.  No atomics, and the adjacency of a vertex is read once per level
// for all 64 queries.  Kept vertex v of query b gets local id
//   base[g][b] + rank of v among group g's vertices kept by b
// (g = v / 64, ranks by ballot), so local ids ascend with original ids as
// the reference's compaction order does.  The local degree and adjacency
// passes read the adjacency of each kept vertex once for every query of
// the batch that kept it; a 64 x 64 bit transpose of the per-neighbour keep
// masks gives every query its own bits.
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
// One add per wave into one of MS_BYTE_SLOTS counters, each on its own
// 128-byte line: adds to a single address serialise at the L2 (about 12 ns
// each), which had made the count, not the work, bound the wide launches.
__device__ __forceinline__ void ms_count_bytes(unsigned long long* ctr, u64 b) {
  b = ms_sum(b);
  const u32 slot = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (MS_BYTE_SLOTS - 1);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(ctr + slot * MS_BYTE_STRIDE, (unsigned long long)b);
}
__device__ __forceinline__ u64 ms_below() {
  const u32 lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}
__device__ __forceinline__ u64 ms_shfl_xor64(u64 x, int m) {
  const u32 lo = __shfl_xor((u32)x, m), hi = __shfl_xor((u32)(x >> 32), m);
  return ((u64)hi << 32) | lo;
}
// 64 x 64 bit transpose across the wave: on entry lane j holds row j, on
// exit lane b holds column b (bit j of lane b = bit b of row j).  Six
// butterfly stages; at stage j the pair (lane, lane ^ j) swaps the j x j
// blocks off the diagonal.
__device__ __forceinline__ u64 ms_transpose(u64 a) {
  const u32 lane = threadIdx.x & 63;
  const u64 M[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                    0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int j = 32 >> s;
    const u64 p = ms_shfl_xor64(a, j);
    const u64 up = ((p >> j) ^ a) & M[s];           // lane has bit j set
    const u64 dn = (((a >> j) ^ p) & M[s]) << j;    // lane has bit j clear
    a ^= (lane & (u32)j) ? up : dn;
  }
  return a;
}
// Inclusive sum over the wave's lanes.
__device__ __forceinline__ u32 ms_wave_incl(u32 x) {
  const u32 lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u32 y = __shfl_up(x, o);
    if (lane >= (u32)o) x += y;
  }
  return x;
}
// The wave number as a wave-uniform value.
__device__ __forceinline__ u32 ms_wave_id() {
  return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
}

// elig/vis/fr for every vertex; level flags cleared.
__global__ void __launch_bounds__(256) ms_init(u32 n, const u32* __restrict__ core,
                                               const MsBatch* __restrict__ B,
                                               u64* __restrict__ elig, u64* __restrict__ vis,
                                               u64* __restrict__ fr, u64* __restrict__ nf,
                                               u32* __restrict__ flags, u32 nflags,
                                               unsigned long long* bytes) {
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
    nf[v] = 0;
    nb += 4 + 32;
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

// Level 0, top-down: the frontier is the query vertices themselves, so
// rather than every eligible vertex scanning its adjacency for them (the
// bottom-up form of the later levels), the wave of query b pushes bit b to
// the eligible neighbours of q_b.  The atomicOr on vis decides which wave
// reached a vertex first for a bit, so each (vertex, query) is taken once.
// Sets flags[0] if anything was reached.
__global__ void __launch_bounds__(64) ms_push0(const i64* __restrict__ off,
                                               const int32_t* __restrict__ adj,
                                               const MsBatch* __restrict__ B, u64 act,
                                               const u64* __restrict__ elig, u64* __restrict__ vis,
                                               u64* __restrict__ nf, u16* __restrict__ dist16,
                                               u32* __restrict__ flag0, unsigned long long* bytes) {
  const u32 b = blockIdx.x, lane = threadIdx.x;
  if (!((act >> b) & 1ull)) return;
  const u32 q = B->q[b];
  const u64 bit = 1ull << b;
  const i64 b0 = off[q], e0 = off[q + 1];
  bool found = false;
  u64 nbytes = lane == 0 ? 20 : 0;
  for (i64 j = b0 + lane; j < e0; j += 64) {
    const u32 w = (u32)adj[j];
    nbytes += 12;
    if (elig[w] & bit) {
      const u64 old = atomicOr((unsigned long long*)&vis[w], (unsigned long long)bit);
      if (!(old & bit)) {
        atomicOr((unsigned long long*)&nf[w], (unsigned long long)bit);
        dist16[(u64)w * MS_BITS + b] = 1;
        found = true;
        nbytes += 18;
      }
    }
  }
  if (__ballot(found) && lane == 0) atomicOr(flag0, 1u);
  ms_count_bytes(bytes, nbytes);
}

// One BFS level for every query whose depth bound exceeds `level` (act).
// Runs only if the previous level found something (flag_prev).  A 16-lane
// segment takes one vertex (four per wave) and reads 64 adjacency entries
// per step, stopping once it holds every bit it wants; vertices of degree
// above MS_HUB are taken by the whole wave, 512 entries per step.
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
  const u64 wid = ms_wave_id();
  const u64 nw = ((u64)gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  bool found = false;
  // segment s of wave w takes vertex (4*it + s)*nw + w at step it:
  // neighbouring ids go to different waves, so a run of hubs (generators
  // often number them first) is spread over the grid, not queued in one wave
  for (u64 it = 0; it * 4 * nw < n; it++) {
    const u64 vv = (4 * it + seg) * nw + wid;
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
    // 64 entries per segment step: each lane issues its four adjacency
    // loads, then its four frontier gathers, so a step costs two memory
    // round trips
    u64 acc = 0;
    i64 j = b0 + sl;
    while (__ballot(active)) {
      u64 x = 0;
      if (active) {
        u32 w[4];
#pragma unroll
        for (int t = 0; t < 4; t++) w[t] = j + 16 * t < e0 ? (u32)adj[j + 16 * t] : 0u;
#pragma unroll
        for (int t = 0; t < 4; t++)
          if (j + 16 * t < e0) {
            x |= fr[w[t]];
            nbytes += 12;
          }
        x &= want;
      }
      acc |= ms_seg_or(x);
      j += 64;
      if (active && (acc == want || j - (i64)sl >= e0)) active = false;
    }
    // hubs: the whole wave, 512 entries per step
    u64 bm = __ballot(big && sl == 0);
    while (bm) {
      const int src = __ffsll((unsigned long long)bm) - 1;
      bm &= bm - 1;
      const u64 bwant = ms_shfl64(want, src);
      const i64 bb = (i64)ms_shfl64((u64)b0, src), be = (i64)ms_shfl64((u64)e0, src);
      u64 bacc = 0;
      for (i64 jj = bb + lane; jj - (i64)lane < be; jj += 512) {
        u32 w[8];
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = jj + 64 * t < be ? (u32)adj[jj + 64 * t] : 0u;
        u64 x = 0;
#pragma unroll
        for (int t = 0; t < 8; t++)
          if (jj + 64 * t < be) {
            x |= fr[w[t]];
            nbytes += 12;
          }
        bacc |= ms_or(x & bwant);
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
  // set the level flag once: later waves read it from the L2 and skip the
  // atomic, so the finders do not queue on one address
  if (__ballot(found) && lane == 0 &&
      __hip_atomic_load(flag_cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    atomicOr(flag_cur, 1u);
  ms_count_bytes(bytes, nbytes);
}

// Per 64-vertex group g and query b: number of group vertices kept by b,
// stored at gcnt[g * 64 + b].
__global__ void __launch_bounds__(256) ms_count(u32 n, u32 G, const u64* __restrict__ vis,
                                                u64 valid, u32* __restrict__ gcnt,
                                                unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  for (u32 g = wid; g < G; g += nw) {
    const u64 v = (u64)g * 64 + lane;
    const u64 m = v < n ? vis[v] & valid : 0;
    if (v < n) nbytes += 8;
    // lane b: the group's vertices kept by query b
    gcnt[(u64)g * 64 + lane] = (u32)__popcll(ms_transpose(m));
    nbytes += 4;
  }
  ms_count_bytes(bytes, nbytes);
}

// One workgroup per query b: group counts -> exclusive bases in place;
// K_b -> tot[b].
// One wave per query and no LDS, so the pass also finds room on CUs whose
// LDS the search kernel holds.
__global__ void __launch_bounds__(64) ms_bscan(u32 G, u32* __restrict__ gcnt, u32* __restrict__ tot,
                                               unsigned long long* bytes) {
  const u32 b = blockIdx.x, t = threadIdx.x;
  const u32 per = (G + 63) / 64;
  const u32 g0 = min(G, t * per), g1 = min(G, g0 + per);
  u32 s = 0;
  for (u32 g = g0; g < g1; g++) s += gcnt[(u64)g * 64 + b];
  const u32 incl = ms_wave_incl(s);
  u32 run = incl - s;
  for (u32 g = g0; g < g1; g++) {
    const u32 c = gcnt[(u64)g * 64 + b];
    gcnt[(u64)g * 64 + b] = run;
    run += c;
  }
  if (t == 63) tot[b] = incl;
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
  // lane b holds query b's output pointers and vertex; the per-query loop
  // below takes them by shuffles, so its stores wait on no memory load
  const MsQuery Qb = B->P[lane];
  const u32 qb = B->q[lane];
  u64 nbytes = 0;
  for (u32 g = wid; g < G; g += nw) {
    const u64 v = (u64)g * 64 + lane;
    const u64 m = v < n ? vis[v] & valid : 0;
    const u32 gb = gcnt[(u64)g * 64 + lane];  // lane b: query b's first id in the group
    u32 cv = 0;
    if (m) {
      cv = core[v];
      nbytes += 12;
    }
    u64 any = ms_or(m);
    // eight queries at a time: their distance loads are in flight together
    // before the stores that need them
    while (any) {
      int bb[8];
      u32 dv[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        bb[t] = any ? __ffsll((unsigned long long)any) - 1 : -1;
        any &= any - 1;  // 0 stays 0
      }
#pragma unroll
      for (int t = 0; t < 8; t++)
        dv[t] = (bb[t] >= 0 && ((m >> bb[t]) & 1ull)) ? (u32)dist16[v * MS_BITS + bb[t]] : 0u;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int b = bb[t];
        if (b < 0) break;
        const u64 bal = __ballot((m >> b) & 1ull);
        const u32 base = (u32)__shfl((int)gb, b), qv = (u32)__shfl((int)qb, b);
        u32* const lorig = (u32*)ms_shfl64((u64)Qb.lorig, b);
        u32* const ldist = (u32*)ms_shfl64((u64)Qb.ldist, b);
        u32* const lcore = (u32*)ms_shfl64((u64)Qb.lcore, b);
        if ((m >> b) & 1ull) {
          const u32 id = base + (u32)__popcll(bal & below);
          lorig[id] = (u32)v;
          ldist[id] = dv[t];
          lcore[id] = cv;
          locid[v * MS_BITS + b] = id;
          if ((u32)v == qv) qloc[b] = id;
          nbytes += 4 + 2 + 12 + 4;
        }
      }
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// Whole-graph queries (gmask): query b's distance array over every vertex of
// the graph, SCS_DIST_NONE where b does not keep the vertex.  Lane = vertex,
// so each query's 64 values of a group are one coalesced store.  Queries are
// taken eight at a time: their eight distance loads are in flight together
// before the eight stores.
__global__ void __launch_bounds__(256) ms_gdist(u32 n, u32 G, const u64* __restrict__ vis,
                                                u64 gmask, const u16* __restrict__ dist16,
                                                const MsBatch* __restrict__ B,
                                                unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  const u64 dp = (u64)B->P[lane].ldist;
  u64 nbytes = 0;
  for (u32 g = wid; g < G; g += nw) {
    const u64 v = (u64)g * 64 + lane;
    const bool in = v < n;
    const u64 m = in ? vis[v] & gmask : 0ull;
    if (in) nbytes += 8;
    u64 left = gmask;
    while (left) {
      int bb[8];
      u32 dv[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        bb[t] = left ? __ffsll((unsigned long long)left) - 1 : -1;
        left &= left - 1;  // 0 stays 0
      }
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const bool kept = bb[t] >= 0 && ((m >> bb[t]) & 1ull);
        dv[t] = kept ? (u32)dist16[v * MS_BITS + bb[t]] : SCS_DIST_NONE;
        if (in && bb[t] >= 0) nbytes += kept ? 6 : 4;
      }
#pragma unroll
      for (int t = 0; t < 8; t++) {
        if (bb[t] < 0) break;
        u32* const d = (u32*)ms_shfl64(dp, bb[t]);
        if (in) d[v] = dv[t];
      }
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// One adjacency unit (at most MS_UNIT entries): lane i loads entries i,
// i+64, i+128, i+192, then their vis words, and keeps w[t] = the neighbour
// id (0 past the unit's end) and mw[t] = the queries that keep that
// neighbour.  (Unconditional loads from clamped indices, which let all
// eight be in flight, measured no faster: most units have one chunk.)
__device__ __forceinline__ void ms_unit_load(const int32_t* __restrict__ adj,
                                             const u64* __restrict__ vis, i64 b0, u32 len,
                                             u32 w[4], u64 mw[4], u64& nbytes) {
  const u32 lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 4; t++) w[t] = lane + 64u * t < len ? (u32)adj[b0 + lane + 64 * t] : 0u;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    mw[t] = 0;
    if (lane + 64u * t < len) {
      mw[t] = vis[w[t]];
      nbytes += 12;
    }
  }
}
__device__ __forceinline__ i64 ms_unit_b0(const MsUnit& U) {
  return (i64)(U.pos & 0xFFFFFFFFFFFFull);
}
__device__ __forceinline__ u32 ms_unit_len(const MsUnit& U) { return (u32)(U.pos >> 48); }

// Local degree of every (query, kept vertex).  One wave per adjacency unit;
// lane b counts query b's kept neighbours in the unit as the popcount of
// column b of the transposed keep masks.  A single-unit vertex writes
// loff[id] directly; a hub's units write partial counts to hubcnt, which
// ms_hubsum turns into unit offsets and the degree.  The unit's vis word,
// locid row and adjacency are loaded together, before the keep test.
__global__ void __launch_bounds__(256) ms_degree(const int32_t* __restrict__ adj,
                                                 const MsUnit* __restrict__ units, u32 nunits,
                                                 const u64* __restrict__ vis, u64 valid,
                                                 const u32* __restrict__ locid,
                                                 const MsBatch* __restrict__ B,
                                                 u32* __restrict__ hubcnt,
                                                 unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  u32* const loff = B->P[lane].loff;
  u64 nbytes = 0;
  for (u32 k = wid; k < nunits; k += nw) {
    const MsUnit U = units[k];
    const i64 b0 = ms_unit_b0(U);
    const u32 len = ms_unit_len(U);
    const u64 mu = vis[U.u] & valid;
    const u32 id = locid[(u64)U.u * MS_BITS + lane];
    u32 w[4] = {0u, 0u, 0u, 0u};
    u64 mw[4] = {0ull, 0ull, 0ull, 0ull};
    if (len) ms_unit_load(adj, vis, b0, len, w, mw, nbytes);
    if (lane == 0) nbytes += 16 + 8;
    if (!mu) continue;
    u32 cnt = 0;
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (64u * t < len) cnt += (u32)__popcll(ms_transpose(mw[t] & mu));
    if (U.h == MS_NOHUB) {
      if ((mu >> lane) & 1ull) {
        loff[id] = cnt;
        nbytes += 8;
      }
    } else {
      hubcnt[(u64)U.h * 64 + lane] = cnt;
      nbytes += 4;
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// One wave per hub vertex: lane b turns the hub's per-unit counts into
// exclusive unit offsets (in place) and writes the degree to loff[id].
__global__ void __launch_bounds__(256) ms_hubsum(const MsHub* __restrict__ hubs, u32 nhubs,
                                                 const u64* __restrict__ vis, u64 valid,
                                                 const u32* __restrict__ locid,
                                                 const MsBatch* __restrict__ B,
                                                 u32* __restrict__ hubcnt,
                                                 unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  u64 nbytes = 0;
  for (u32 k = wid; k < nhubs; k += nw) {
    const MsHub H = hubs[k];
    const u64 mu = vis[H.u] & valid;
    if (lane == 0) nbytes += 16 + 8;
    if (!mu) continue;
    u32 run = 0;
    for (u32 t = 0; t < H.nu; t++) {
      u32* p = &hubcnt[(u64)(H.h0 + t) * 64 + lane];
      const u32 c = *p;
      *p = run;
      run += c;
    }
    nbytes += (u64)H.nu * 8;
    if ((mu >> lane) & 1ull) {
      const u32 id = locid[(u64)H.u * MS_BITS + lane];
      B->P[lane].loff[id] = run;
      nbytes += 8;
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// Exclusive scan of every valid query's loff[0..K) in place, loff[K] = nnz
// -> tot[64 + b], in three passes with no LDS: chunk sums, chunk bases per
// query, then each chunk rescanned from its base.  A chunk is MS_SCAN_CHUNK
// entries, one wave, 64 contiguous entries per lane (16 aligned 16-byte
// loads: loff starts on an arena granule).  Wave w of the first and last
// passes takes query w / cmax, chunk w % cmax.
#define MS_SCAN_CHUNK 4096

__device__ __forceinline__ u32 ms_lane_load(const u32* __restrict__ a, u64 i0, u32 K, uint4 v[16]) {
  u32 s = 0;
  if (i0 + 64 <= K) {
    const uint4* p = (const uint4*)(a + i0);
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = p[k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const u64 i = i0 + 4 * k;
      v[k].x = i < K ? a[i] : 0u;
      v[k].y = i + 1 < K ? a[i + 1] : 0u;
      v[k].z = i + 2 < K ? a[i + 2] : 0u;
      v[k].w = i + 3 < K ? a[i + 3] : 0u;
    }
  }
#pragma unroll
  for (int k = 0; k < 16; k++) s += v[k].x + v[k].y + v[k].z + v[k].w;
  return s;
}

__global__ void __launch_bounds__(256) ms_loff_sum(const MsBatch* __restrict__ B,
                                                   const u32* __restrict__ tot, u64 valid,
                                                   u32 cmax, u32* __restrict__ csum) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 b = wid / cmax, c = wid % cmax;
  if (b >= 64 || !((valid >> b) & 1ull)) return;
  const u32 K = tot[b];
  const u64 c0 = (u64)c * MS_SCAN_CHUNK;
  if (c0 >= K) return;
  const u64 i0 = c0 + 64ull * lane;
  uint4 v[16];
  const u32 s = i0 < K ? ms_lane_load(B->P[b].loff, i0, K, v) : 0u;
  const u32 t = __shfl(ms_wave_incl(s), 63);
  if (lane == 0) csum[(u64)b * cmax + c] = t;
}

// One wave per query: chunk sums -> exclusive chunk bases; loff[K] and
// tot[64 + b] get the total.
__global__ void __launch_bounds__(64) ms_loff_base(const MsBatch* __restrict__ B,
                                                   u32* __restrict__ tot, u64 valid, u32 cmax,
                                                   u32* __restrict__ csum) {
  const u32 b = blockIdx.x, lane = threadIdx.x;
  if (!((valid >> b) & 1ull)) return;
  const u32 K = tot[b];
  const u32 nc = (K + MS_SCAN_CHUNK - 1) / MS_SCAN_CHUNK;
  const u32 per = (nc + 63) / 64;
  const u32 c0 = min(nc, lane * per), c1 = min(nc, c0 + per);
  u32* cs = csum + (u64)b * cmax;
  u32 s = 0;
  for (u32 c = c0; c < c1; c++) s += cs[c];
  const u32 incl = ms_wave_incl(s);
  u32 run = incl - s;
  for (u32 c = c0; c < c1; c++) {
    const u32 x = cs[c];
    cs[c] = run;
    run += x;
  }
  if (lane == 63) {
    B->P[b].loff[K] = incl;
    tot[64 + b] = incl;
  }
}

__global__ void __launch_bounds__(256) ms_loff_apply(const MsBatch* __restrict__ B,
                                                     const u32* __restrict__ tot, u64 valid,
                                                     u32 cmax, const u32* __restrict__ csum,
                                                     unsigned long long* bytes) {
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 b = wid / cmax, c = wid % cmax;
  if (b >= 64 || !((valid >> b) & 1ull)) return;
  const u32 K = tot[b];
  const u64 c0 = (u64)c * MS_SCAN_CHUNK;
  if (c0 >= K) return;
  u32* a = B->P[b].loff;
  const u64 i0 = c0 + 64ull * lane;
  uint4 v[16];
  const u32 s = i0 < K ? ms_lane_load(a, i0, K, v) : 0u;
  u32 run = csum[(u64)b * cmax + c] + ms_wave_incl(s) - s;
  u64 nbytes = 0;
  if (i0 < K) {
    uint4 o[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      o[k].x = run; run += v[k].x;
      o[k].y = run; run += v[k].y;
      o[k].z = run; run += v[k].z;
      o[k].w = run; run += v[k].w;
    }
    if (i0 + 64 <= K) {
      uint4* p = (uint4*)(a + i0);
#pragma unroll
      for (int k = 0; k < 16; k++) p[k] = o[k];
      nbytes = 64 * 8;
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const u64 i = i0 + 4 * k;
        if (i < K) a[i] = o[k].x;
        if (i + 1 < K) a[i + 1] = o[k].y;
        if (i + 2 < K) a[i + 2] = o[k].z;
        if (i + 3 < K) a[i + 3] = o[k].w;
      }
      nbytes = (K - i0) * 8;
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// Local adjacency: for every (query, kept vertex u), the local ids of u's
// kept neighbours in ascending order at ladj[loff[id(u)] ..].
// One wave per adjacency unit.  Query b's entries of the unit go to its
// write position dst_b, in neighbour order; column b of the transposed keep
// masks says which neighbours query b keeps.  The work is cut into tiles of
// MS_FILL_QT queries x (64 / MS_FILL_QT) neighbours: lane (ji, qi) =
// (lane / QT, lane % QT) takes query qb*QT+qi and neighbour jb*JT+ji.  A
// store instruction then writes QT runs of up to JT consecutive entries and
// a load instruction reads JT pieces of QT entries of the locid rows, where
// a whole-row or whole-column form touches 64 lines with one of the two.
// On C2 (Chung-Lu, 100K vertices) 8 x 8 tiles with 32 row loads in flight
// were fastest; 4 x 16 and 2 x 32 were 1 % and 22 % slower, non-temporal
// stores (no write-combining in the L2) 2.7x slower, and runs of 2 to 8
// consecutive units per wave 3-6 % slower.
#ifndef MS_FILL_QT
#define MS_FILL_QT 8
#endif
#ifndef MS_FILL_LOADS
#define MS_FILL_LOADS 32  // row loads in flight per wave
#endif
__global__ void __launch_bounds__(256) ms_fill(const int32_t* __restrict__ adj,
                                               const MsUnit* __restrict__ units, u32 nunits,
                                               const u64* __restrict__ vis, u64 valid,
                                               const u32* __restrict__ locid,
                                               const MsBatch* __restrict__ B,
                                               const u32* __restrict__ hubcnt,
                                               unsigned long long* bytes) {
  constexpr u32 QT = MS_FILL_QT, JT = 64 / QT, QG = MS_FILL_LOADS / QT;
  constexpr u64 GMASK = QG * QT == 64 ? ~0ull : ((1ull << (QG * QT)) - 1ull);
  constexpr u64 JMASK = JT == 64 ? ~0ull : ((1ull << JT) - 1ull);
  const u32 lane = threadIdx.x & 63;
  const u32 wid = ms_wave_id();
  const u32 nw = (gridDim.x * blockDim.x) >> 6;
  const u32 qi = lane % QT, ji = lane / QT;
  const u64 below = ji ? (~0ull >> (64 - ji)) : 0ull;
  const u32* const loff = B->P[lane].loff;
  u32* const ladj = B->P[lane].ladj;
  u64 nbytes = 0;
  for (u32 k = wid; k < nunits; k += nw) {
    const MsUnit U = units[k];
    const i64 b0 = ms_unit_b0(U);
    const u32 len = ms_unit_len(U);
    const u64 mu = vis[U.u] & valid;
    const u32 id = locid[(u64)U.u * MS_BITS + lane];
    const u32 part = hubcnt[(U.h != MS_NOHUB ? (u64)U.h * 64 : 0ull) + lane] &
                     (U.h != MS_NOHUB ? ~0u : 0u);
    u32 w[4] = {0u, 0u, 0u, 0u};
    u64 mw[4] = {0ull, 0ull, 0ull, 0ull};
    if (len) ms_unit_load(adj, vis, b0, len, w, mw, nbytes);
    if (lane == 0) nbytes += 16 + 8;
    if (!mu) continue;
    // lane b: where query b's entries of this unit go
    u64 dst = 0;
    if ((mu >> lane) & 1ull) {
      dst = (u64)(ladj + loff[id] + part);
      nbytes += 8;
    }
    const u32 nt = (len + 63) / 64;
    u64 col[4];  // lane b: bit j = query b keeps entry 64t + j
#pragma unroll
    for (int t = 0; t < 4; t++) col[t] = (u32)t < nt ? ms_transpose(mw[t] & mu) : 0ull;
#pragma unroll
    for (int t = 0; t < 4; t++) nbytes += 8ull * (u32)__popcll(col[t]);
    // QG query groups at a time, so QG * QT row loads are in flight
    for (u32 qb0 = 0; qb0 < 64 / QT; qb0 += QG) {
      if (!((mu >> (qb0 * QT)) & GMASK)) continue;  // none of these queries keeps U.u
      __attribute__((address_space(1))) u32* dq[QG];
#pragma unroll
      for (u32 x = 0; x < QG; x++)
        dq[x] = (__attribute__((address_space(1))) u32*)ms_shfl64(dst, (int)((qb0 + x) * QT + qi));
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if ((u32)t >= nt) break;
        u32 wj[QT];  // w[t] is 0 past the unit's end, so every row index is a vertex
#pragma unroll
        for (u32 jb = 0; jb < QT; jb++) wj[jb] = (u32)__shfl(w[t], (int)(jb * JT + ji));
        u32 r[QG][QT];
#pragma unroll
        for (u32 x = 0; x < QG; x++)
#pragma unroll
          for (u32 jb = 0; jb < QT; jb++)
            r[x][jb] = locid[(u64)wj[jb] * MS_BITS + (qb0 + x) * QT + qi];
#pragma unroll
        for (u32 x = 0; x < QG; x++) {
          const u64 colq = ms_shfl64(col[t], (int)((qb0 + x) * QT + qi));
#pragma unroll
          for (u32 jb = 0; jb < QT; jb++) {
            const u64 bits = (colq >> (jb * JT)) & JMASK;
            if ((bits >> ji) & 1ull) dq[x][__popcll(bits & below)] = r[x][jb];
            dq[x] += __popcll(bits);
          }
        }
      }
    }
  }
  ms_count_bytes(bytes, nbytes);
}

// ---------------------------------------------------------------------------
// Launchers.  B points to the batch header in device memory.
// ---------------------------------------------------------------------------
// Grid-stride kernels launch at most one generation of resident waves
// (2048 blocks of four waves = 8 waves on each SIMD of 256 CUs).
#define MS_GRID 2048
static u32 ms_blocks(u64 threads, u32 cap) {
  u64 b = (threads + 255) / 256;
  if (b == 0) b = 1;
  return (u32)(b < cap ? b : cap);
}

extern "C" hipError_t scs_launch_ms_init(u32 n, const u32* core, const MsBatch* B, u64* elig,
                                         u64* vis, u64* fr, u64* nf, u32* flags, u32 nflags,
                                         unsigned long long* bytes, hipStream_t s) {
  ms_init<<<ms_blocks(n, MS_GRID), 256, 0, s>>>(n, core, B, elig, vis, fr, nf, flags, nflags,
                                                bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_push0(const i64* off, const int32_t* adj, u32 nb,
                                          const MsBatch* B, u64 act, const u64* elig, u64* vis,
                                          u64* nf, u16* dist16, u32* flag0,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_push0<<<nb, 64, 0, s>>>(off, adj, B, act, elig, vis, nf, dist16, flag0, bytes);
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
  ms_level<<<ms_blocks(((u64)n + 3) / 4 * 64, MS_GRID), 256, 0, s>>>(
      off, adj, n, level, act, elig, vis, fr, nf, dist16, flag_prev, flag_cur, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_count(u32 n, u32 G, const u64* vis, u64 valid, u32* gcnt,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_count<<<ms_blocks((u64)G * 64, MS_GRID), 256, 0, s>>>(n, G, vis, valid, gcnt, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_bscan(u32 nb, u32 G, u32* gcnt, u32* tot,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_bscan<<<nb, 64, 0, s>>>(G, gcnt, tot, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_gdist(u32 n, u32 G, const u64* vis, u64 gmask,
                                          const u16* dist16, const MsBatch* B,
                                          unsigned long long* bytes, hipStream_t s) {
  ms_gdist<<<ms_blocks((u64)G * 64, MS_GRID), 256, 0, s>>>(n, G, vis, gmask, dist16, B, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_scatter(u32 n, u32 G, const u64* vis, u64 valid,
                                            const u32* gcnt, const u32* core, const u16* dist16,
                                            const MsBatch* B, u32* locid, u32* qloc,
                                            unsigned long long* bytes, hipStream_t s) {
  ms_scatter<<<ms_blocks((u64)G * 64, MS_GRID), 256, 0, s>>>(n, G, vis, valid, gcnt, core, dist16, B,
                                                          locid, qloc, bytes);
  return hipGetLastError();
}
// Local degrees, then the hubs' unit offsets (nhubs may be 0).
extern "C" hipError_t scs_launch_ms_degree(const int32_t* adj, const MsUnit* units, u32 nunits,
                                           const MsHub* hubs, u32 nhubs, const u64* vis,
                                           u64 valid, const u32* locid, const MsBatch* B,
                                           u32* hubcnt, unsigned long long* bytes,
                                           hipStream_t s) {
  ms_degree<<<ms_blocks((u64)nunits * 64, MS_GRID), 256, 0, s>>>(adj, units, nunits, vis, valid,
                                                                locid, B, hubcnt, bytes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nhubs == 0) return e;
  ms_hubsum<<<ms_blocks((u64)nhubs * 64, MS_GRID), 256, 0, s>>>(hubs, nhubs, vis, valid, locid, B,
                                                             hubcnt, bytes);
  return hipGetLastError();
}
// cmax = chunks of the largest valid K (>= 1); csum holds nb * cmax words.
extern "C" hipError_t scs_launch_ms_scan_loff(u32 nb, const MsBatch* B, u32* tot, u64 valid,
                                              u32 cmax, u32* csum, unsigned long long* bytes,
                                              hipStream_t s) {
  const u32 blocks = (u32)(((u64)nb * cmax + 3) / 4);
  ms_loff_sum<<<blocks, 256, 0, s>>>(B, tot, valid, cmax, csum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ms_loff_base<<<nb, 64, 0, s>>>(B, tot, valid, cmax, csum);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  ms_loff_apply<<<blocks, 256, 0, s>>>(B, tot, valid, cmax, csum, bytes);
  return hipGetLastError();
}
extern "C" hipError_t scs_launch_ms_fill(const int32_t* adj, const MsUnit* units, u32 nunits,
                                         const u64* vis, u64 valid, const u32* locid,
                                         const MsBatch* B, const u32* hubcnt,
                                         unsigned long long* bytes, hipStream_t s) {
  ms_fill<<<ms_blocks((u64)nunits * 64, MS_GRID), 256, 0, s>>>(adj, units, nunits, vis, valid,
                                                              locid, B, hubcnt, bytes);
  return hipGetLastError();
}
