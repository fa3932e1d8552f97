// Greedy lower bounds on the GPU: the reference's CSSC_heu
// (inc/heuristic.h:321-348), the best of its three greedy constructions
// heu2 (:8-120), heu3 (:125-239) and heu4 (:242-317), one wave per
// construction of each query.
//
// The reference drives each construction with a std::priority_queue that
// holds stale duplicates.  Its pop order reduces to a deterministic choice
// per step, which a wave makes with a reduction instead of a heap:
//   heu2  add the frontier vertex with the largest (count of H-neighbours,
//         id); every increment of a count pushes, so the largest pushed key
//         of a vertex is its current count;
//   heu3  add the frontier vertex with the largest (key, id), where key is
//         the largest score ever pushed for it (scores can fall, and a
//         stale larger entry is popped first);
//   heu4  peel the vertex of S with the smallest (degree in S, id); degrees
//         only fall, so the smallest pushed key of a vertex is its current
//         degree.
// heu3's score of v sums 1/deg(w) over v's neighbours w in H in ascending
// id order, then adds deg(v)/dmax, as the reference's adjacency scan does,
// so the doubles are bit-identical.  heu4 drops the reference's +1 on q when
// q is popped alone (inc/heuristic.h:259), as the host version does.
//
// Candidates sit in an append-only list.  Each block of 64 list positions,
// and each group of 64 blocks, keeps the largest key of its members; keys
// only rise while a vertex is a candidate, so a rise is an atomic max on its
// block and group, and only the block and group of the vertex chosen in a
// step are recomputed.  A step costs the chosen vertex's adjacency plus one
// pass over the group keys (n / 4096 of them).
#include "scs_internal.h"

#define HEU_SMALL 64  // heu3: neighbours of at most this degree are scored by one lane

struct HeuArgs {
  const u32* off;  // n+1, 32-bit offsets of the resident graph
  const u32* adj;
  u32 n, dmax, nq, hcap;
  const u32* ql;
  const u32* qh;
  const u32* qq;
  u8* ws;
  HeuLayout lay;
  u32* kl;     // nq
  u32* hsize;  // nq (sizes above hcap are reported, H is cut at hcap)
  u32* H;      // nq * hcap
  u32* kl3;    // 3 * nq: heu2, heu3, heu4 of each query
  u32* hs3;    // 3 * nq
  u32* H3;     // 3 * nq * hcap
};

#define LANE (threadIdx.x & 63u)

__device__ __forceinline__ void hsync() { __syncthreads(); }
// the counters are changed by atomics (performed at the L2), so they are
// read past the L1
__device__ __forceinline__ u32 hld(u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 h_sum(u32 x) {
  for (int o = 32; o > 0; o >>= 1) x += (u32)__shfl_xor((int)x, o);
  return x;
}
__device__ __forceinline__ u32 h_min(u32 x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, (u32)__shfl_xor((int)x, o));
  return x;
}
__device__ __forceinline__ u64 h_shfl64(u64 x, int src) {
  const u32 lo = (u32)__shfl((int)(u32)x, src), hi = (u32)__shfl((int)(u32)(x >> 32), src);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 h_max64(u64 x) {
  for (int o = 32; o > 0; o >>= 1) {
    const u64 y = h_shfl64(x, (int)(LANE ^ (u32)o));
    x = y > x ? y : x;
  }
  return x;
}
__device__ __forceinline__ double h_shfld(double x, int src) {
  return __longlong_as_double((long long)h_shfl64((u64)__double_as_longlong(x), src));
}
struct HW {
  const u32* off;
  const u32* adj;
  u32 n, dmax;
  u8* sta;
  u8* ins;
  u32* cnt;
  u32* deg;
  double* key;
  u32* pos;
  u32* list;
  u32* rmv;
  u32* th;
  u64* bkey;  // per block of 64 list positions: the largest member key (0 = none)
  u64* gkey;  // per group of 64 blocks: the largest block key
  HeuLds* L;
};

// Candidate keys, larger is chosen first, 0 = not a candidate.
//   heu2: (count of H-neighbours + 1, id)
//   heu4: the complement of (degree in S, id), so the smallest pair is largest
//   heu3: the bits of the (non-negative) double key, plus one; ties are broken
//         by id when the chosen vertex is looked up
__device__ __forceinline__ u64 kv2(const HW& w, u32 x) {
  return w.sta[x] == 1 ? (((u64)(w.cnt[x] + 1) << 32) | x) : 0ull;
}
__device__ __forceinline__ u64 kv4(const HW& w, u32 x) {
  return w.ins[x] ? ~(((u64)w.deg[x] << 32) | x) : 0ull;
}
__device__ __forceinline__ u64 kv3(const HW& w, u32 x) {
  return w.sta[x] == 1 ? (u64)__double_as_longlong(w.key[x]) + 1ull : 0ull;
}
template <int MODE>
__device__ __forceinline__ u64 kv(const HW& w, u32 x) {
  return MODE == 2 ? kv2(w, x) : (MODE == 4 ? kv4(w, x) : kv3(w, x));
}
// block and group keys are raised by atomics at the L2; reads go past the L1
__device__ __forceinline__ u64 hld64(u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void hst64(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a member's key rose (or it joined): raise its block's and group's keys
__device__ __forceinline__ void bump(HW& w, u32 p, u64 k) {
  atomicMax((unsigned long long*)&w.bkey[p >> 6], (unsigned long long)k);
  atomicMax((unsigned long long*)&w.gkey[p >> 12], (unsigned long long)k);
}

// Exact keys of blocks b0 .. b0+cnt (eight per pass, so eight chains of
// dependent loads are in flight), then of the groups they fall in.
template <int MODE>
__device__ void blocks_exact(HW& w, u32 b0, u32 cnt, u32 m) {
  constexpr int U = 8;
  for (u32 i0 = 0; i0 < cnt; i0 += U) {
    u64 k[U];
#pragma unroll
    for (int t = 0; t < U; t++) {
      const u32 p = (b0 + i0 + t) * 64 + LANE;
      const bool in = i0 + t < cnt && p < m;
      k[t] = in ? kv<MODE>(w, w.list[p]) : 0ull;
    }
#pragma unroll
    for (int t = 0; t < U; t++) {
      const u64 r = h_max64(k[t]);
      if (LANE == 0 && i0 + t < cnt) hst64(&w.bkey[b0 + i0 + t], r);
    }
  }
  __threadfence();
  hsync();
  const u32 nblk = (m + 63) / 64;
  for (u32 g = (b0 >> 6); g <= ((b0 + cnt - 1) >> 6); g++) {
    const u32 b = g * 64 + LANE;
    const u64 r = h_max64(b < nblk ? hld64(&w.bkey[b]) : 0ull);
    if (LANE == 0) hst64(&w.gkey[g], r);
  }
  __threadfence();
  hsync();
}

// The largest key over all candidates (0 when there is none).
__device__ __forceinline__ u64 top_key(HW& w, u32 m) {
  const u32 ngrp = ((m + 63) / 64 + 63) / 64;
  u64 best = 0;
  for (u32 g = LANE; g < ngrp; g += 64) best = max(best, hld64(&w.gkey[g]));
  return h_max64(best);
}

// heu3: the largest id among the candidates whose key is M.
__device__ u32 find3(HW& w, u32 m, u64 M) {
  const u32 nblk = (m + 63) / 64, ngrp = (nblk + 63) / 64;
  u64 best = 0;
  for (u32 g0 = 0; g0 < ngrp; g0 += 64) {
    u64 gm = __ballot(g0 + LANE < ngrp && hld64(&w.gkey[g0 + LANE]) == M);
    while (gm) {
      const u32 g = g0 + (u32)(__ffsll((unsigned long long)gm) - 1);
      gm &= gm - 1;
      const u32 b1 = g * 64 + LANE;
      u64 bm = __ballot(b1 < nblk && hld64(&w.bkey[b1]) == M);
      while (bm) {
        const u32 b = g * 64 + (u32)(__ffsll((unsigned long long)bm) - 1);
        bm &= bm - 1;
        const u32 p = b * 64 + LANE;
        u64 c = 0;
        if (p < m) {
          const u32 x = w.list[p];
          if (kv3(w, x) == M) c = (u64)x + 1;
        }
        best = max(best, h_max64(c));
      }
    }
  }
  return best ? (u32)(best - 1) : 0xffffffffu;
}

// Appends v to the candidate list (one lane per call site); its key is
// raised by the caller.
__device__ __forceinline__ u32 push_cand(HW& w, u32 v) {
  const u32 p = atomicAdd(&w.L->m, 1u);
  w.list[p] = v;
  w.pos[v] = p;
  return p;
}

// min over H = th[0..nH) of deg
__device__ __forceinline__ u32 min_deg_h(HW& w, u32 nH) {
  u32 mn = 0xffffffffu;
  for (u32 i = LANE; i < nH; i += 64) mn = min(mn, w.deg[w.th[i]]);
  return h_min(mn);
}

// Every array the run touched back to zero: the candidates' state and the
// block and group keys.
__device__ __forceinline__ void cleanup(HW& w) {
  hsync();
  const u32 m = hld(&w.L->m);
  for (u32 i = LANE; i < m; i += 64) {
    const u32 x = w.list[i];
    w.sta[x] = 0;
    w.ins[x] = 0;
    w.cnt[x] = 0;
    w.deg[x] = 0;
    w.key[x] = 0.0;
  }
  const u32 nblk = (m + 63) / 64, ngrp = (nblk + 63) / 64;
  for (u32 b = LANE; b < nblk; b += 64) hst64(&w.bkey[b], 0ull);
  for (u32 g = LANE; g < ngrp; g += 64) hst64(&w.gkey[g], 0ull);
  __threadfence();
  hsync();
  if (LANE == 0) __hip_atomic_store(&w.L->m, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __threadfence();
  hsync();
}

// heu2 (inc/heuristic.h:8-120): H grows from q by the frontier vertex with
// the most H-neighbours (ties: larger id); the best prefix of size in
// [N1, N2] by minimum degree (ties: the later one) is kept.
__device__ __forceinline__ void heu2(HW& w, u32 N1, u32 N2, u32 q, u32* outH, u32 hcap,
                                     u32& osize, u32& okl) {
  if (LANE == 0) {
    w.sta[q] = 1;
    bump(w, push_cand(w, q), kv2(w, q));
  }
  hsync();
  u32 nH = 0, tkl = 0, tsize = 0;
  while (true) {
    const u64 M = top_key(w, hld(&w.L->m));
    if (M == 0) break;
    const u32 v = (u32)M;
    if (LANE == 0) {
      w.th[nH] = v;
      w.sta[v] = 2;
    }
    nH++;
    hsync();
    u32 dv = 0;
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 nb = w.adj[j];
      const u32 c = w.cnt[nb] + 1;
      w.cnt[nb] = c;
      const u8 s = w.sta[nb];
      const u64 k = ((u64)(c + 1) << 32) | nb;
      if (s == 2) {
        w.deg[nb] += 1;
        dv++;
      } else if (s == 0) {
        w.sta[nb] = 1;
        bump(w, push_cand(w, nb), k);
      } else {
        bump(w, w.pos[nb], k);
      }
    }
    dv = h_sum(dv);
    if (LANE == 0) w.deg[v] += dv;
    hsync();
    // v left the frontier: its block and group are recomputed exactly
    blocks_exact<2>(w, w.pos[v] >> 6, 1, hld(&w.L->m));
    if (nH >= N1) {
      const u32 mn = min_deg_h(w, nH);
      if (mn >= tkl) {
        tkl = mn;
        tsize = nH;
      }
    }
    if (nH == N2) break;
  }
  for (u32 i = LANE; i < tsize && i < hcap; i += 64) outH[i] = w.th[i];
  osize = tsize;
  okl = tkl;
  cleanup(w);
}

// Score of a frontier vertex nb whose degree is at most HEU_SMALL, by one
// lane: 1/deg(w) over nb's H-neighbours w with deg(w) != 0 in ascending id
// order, then deg(nb)/dmax.  nb has `target` H-neighbours (the count kept in
// cnt during heu3), so the walk stops at the last of them.
__device__ __forceinline__ double score_lane(const HW& w, u32 nb, u32 target) {
  const u32 b = w.off[nb], e = w.off[nb + 1];
  double sc = 0.0;
  u32 found = 0;
  for (u32 k = b; k < e && found < target; k++) {
    const u32 x = w.adj[k];
    if (w.sta[x] == 2) {
      found++;
      const u32 d = w.deg[x];
      if (d != 0) sc += 1.0 / (double)d;
    }
  }
  return sc + (double)(e - b) / (double)w.dmax;
}

// The same sum by the whole wave for a vertex of large degree: chunks of 64
// entries, the chunk's terms added one after another in lane order, until
// the last H-neighbour has been met.
__device__ __forceinline__ double score_wave(const HW& w, u32 nb, u32 target) {
  const u32 b = w.off[nb], e = w.off[nb + 1];
  double sc = 0.0;
  u32 found = 0;
  for (u32 c = b; c < e && found < target; c += 64) {
    const u32 k = c + LANE;
    double t = 0.0;
    bool inh = false, on = false;
    if (k < e) {
      const u32 x = w.adj[k];
      if (w.sta[x] == 2) {
        inh = true;
        const u32 d = w.deg[x];
        if (d != 0) {
          t = 1.0 / (double)d;
          on = true;
        }
      }
    }
    found += (u32)__popcll(__ballot(inh));
    u64 msk = __ballot(on);
    while (msk) {
      const int i = __ffsll((unsigned long long)msk) - 1;
      msk &= msk - 1;
      sc += h_shfld(t, i);
    }
  }
  return sc + (double)(e - b) / (double)w.dmax;
}

__device__ __forceinline__ void set_key3(HW& w, u32 nb, double sc) {
  const u8 s = w.sta[nb];
  if (s == 0) {
    w.key[nb] = sc;
    w.sta[nb] = 1;
    bump(w, push_cand(w, nb), (u64)__double_as_longlong(sc) + 1ull);
  } else if (sc > w.key[nb]) {
    w.key[nb] = sc;
    bump(w, w.pos[nb], (u64)__double_as_longlong(sc) + 1ull);
  }
}

// heu3 (inc/heuristic.h:125-239): H grows from q by the frontier vertex of
// largest connection score (the largest score pushed for it).
__device__ __forceinline__ void heu3(HW& w, u32 N1, u32 N2, u32 q, u32* outH, u32 hcap,
                                     u32& osize, u32& okl) {
  if (LANE == 0) {
    w.sta[q] = 1;
    w.key[q] = 0.0;
    bump(w, push_cand(w, q), kv3(w, q));
  }
  hsync();
  u32 nH = 0, tkl = 0, tsize = 0;
  while (true) {
    const u32 m = hld(&w.L->m);
    const u64 M = top_key(w, m);
    if (M == 0) break;
    const u32 v = find3(w, m, M);
    if (v == 0xffffffffu) break;  // keys and members disagree: stop (never seen)
    if (LANE == 0) {
      w.th[nH] = v;
      w.sta[v] = 2;
    }
    nH++;
    hsync();
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    u32 dv = 0;
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 nb = w.adj[j];
      w.cnt[nb] += 1;  // heu3 keeps each vertex's number of H-neighbours in cnt
      if (w.sta[nb] == 2) {
        w.deg[nb] += 1;
        dv++;
      }
    }
    dv = h_sum(dv);
    if (LANE == 0) w.deg[v] += dv;
    hsync();
    // scores of v's neighbours outside H, against the degrees after v joined;
    // a vertex whose only H-neighbour is v needs no walk: its sum is v's term
    const u32 degv = w.deg[v];
    const double termv = degv != 0 ? 1.0 / (double)degv : 0.0;
    for (u32 c = b0; c < e0; c += 64) {
      const u32 j = c + LANE;
      u32 nb = 0, target = 0;
      bool one = false, small = false, big = false;
      if (j < e0) {
        nb = w.adj[j];
        if (w.sta[nb] != 2) {
          target = w.cnt[nb];
          const u32 dg = w.off[nb + 1] - w.off[nb];
          one = target == 1;
          small = !one && dg <= HEU_SMALL;
          big = !one && !small;
        }
      }
      // a score reads only H's degrees, which no key update changes
      if (one) {
        const u32 dg = w.off[nb + 1] - w.off[nb];
        set_key3(w, nb, (0.0 + termv) + (double)dg / (double)w.dmax);
      }
      if (small) set_key3(w, nb, score_lane(w, nb, target));
      u64 bm = __ballot(big);
      while (bm) {
        const int src = __ffsll((unsigned long long)bm) - 1;
        bm &= bm - 1;
        const u32 x = (u32)__shfl((int)nb, src);
        const u32 tg = (u32)__shfl((int)target, src);
        const double s2 = score_wave(w, x, tg);
        if (LANE == 0) set_key3(w, x, s2);
      }
    }
    hsync();
    blocks_exact<3>(w, w.pos[v] >> 6, 1, hld(&w.L->m));
    if (nH >= N1) {
      const u32 mn = min_deg_h(w, nH);
      if (mn >= tkl) {
        tkl = mn;
        tsize = nH;
      }
    }
    if (nH == N2) break;
  }
  for (u32 i = LANE; i < tsize && i < hcap; i += 64) outH[i] = w.th[i];
  osize = tsize;
  okl = tkl;
  cleanup(w);
}

// heu4 (inc/heuristic.h:242-317): S = q and its neighbours; peel the vertex
// of smallest (degree in S, id) until N1 remain, keeping the peel prefix
// whose popped degree is the largest with |S left| in [N1, N2].
__device__ __forceinline__ void heu4(HW& w, u32 N1, u32 N2, u32 q, u32* outH, u32 hcap, u32& osize, u32& okl) {
  osize = 0;
  okl = 0;
  const u32 qb = w.off[q], qe = w.off[q + 1];
  if (qe - qb < N1 - 1u) return;
  const u32 m = 1 + (qe - qb);
  if (LANE == 0) {
    w.list[0] = q;
    w.pos[q] = 0;
    w.ins[q] = 1;
    __hip_atomic_store(&w.L->m, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (u32 j = qb + LANE; j < qe; j += 64) {
    const u32 x = w.adj[j];
    w.list[1 + (j - qb)] = x;
    w.pos[x] = 1 + (j - qb);
    w.ins[x] = 1;
  }
  hsync();
  // degree of each member inside S (q: one more, as the reference counts it)
  for (u32 i0 = 0; i0 < m; i0 += 64) {
    const u32 i = i0 + LANE;
    u32 x = 0, b = 0, e = 0;
    if (i < m) {
      x = w.list[i];
      b = w.off[x];
      e = w.off[x + 1];
    }
    const bool big = i < m && (e - b) > HEU_SMALL;
    if (i < m && !big) {
      u32 c = 0;
      for (u32 k = b; k < e; k++) c += w.ins[w.adj[k]];
      w.deg[x] = c + (x == q ? 1u : 0u);
    }
    u64 bm = __ballot(big);
    while (bm) {
      const int src = __ffsll((unsigned long long)bm) - 1;
      bm &= bm - 1;
      const u32 bx = (u32)__shfl((int)x, src), bb = (u32)__shfl((int)b, src),
                be = (u32)__shfl((int)e, src);
      u32 c = 0;
      for (u32 k = bb + LANE; k < be; k += 64) c += w.ins[w.adj[k]];
      c = h_sum(c);
      if (LANE == 0) w.deg[bx] = c + (bx == q ? 1u : 0u);
    }
  }
  hsync();
  blocks_exact<4>(w, 0, (m + 63) / 64, m);
  u32 nr = 0, rmv_idx = 0, mindeg = 0;
  while (true) {
    const u64 M = top_key(w, m);
    if (M == 0) break;
    const u64 best = ~M;
    const u32 v = (u32)best;
    u32 d = (u32)(best >> 32);
    const u32 remain = m - nr;
    if (v == q) d -= 1;
    if (remain >= N1 && remain <= N2 && d > mindeg) {
      mindeg = d;
      rmv_idx = nr;
    }
    if (remain == N1) break;
    if (LANE == 0) {
      w.ins[v] = 0;
      w.rmv[nr] = v;
    }
    nr++;
    hsync();
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 x = w.adj[j];
      if (w.ins[x]) {
        const u32 d2 = w.deg[x] - 1;
        w.deg[x] = d2;
        bump(w, w.pos[x], ~(((u64)d2 << 32) | x));
      }
    }
    hsync();
    blocks_exact<4>(w, w.pos[v] >> 6, 1, m);
  }
  // H = S without the first rmv_idx peeled vertices, in S order
  for (u32 i = LANE; i < m; i += 64) w.ins[w.list[i]] = 1;
  hsync();
  for (u32 i = LANE; i < rmv_idx; i += 64) w.ins[w.rmv[i]] = 0;
  hsync();
  u32 base = 0;
  for (u32 i0 = 0; i0 < m; i0 += 64) {
    const u32 i = i0 + LANE;
    const bool keep = i < m && w.ins[w.list[i]];
    const u64 bal = __ballot(keep);
    if (keep) {
      const u32 p = base + (u32)__popcll(bal & ((1ull << LANE) - 1ull));
      if (p < hcap) outH[p] = w.list[i];
    }
    base += (u32)__popcll(bal);
  }
  osize = base;
  okl = mindeg;
  cleanup(w);
}

// Three waves per query, one per construction (block 3i + c runs heu2,
// heu3 or heu4 of query i in its own workspace).  No LDS: the wave's
// counter lives in its workspace, so the kernel also finds room on CUs
// whose LDS the search kernel holds.
__global__ void __launch_bounds__(64) __attribute__((flatten)) scs_heu_kernel(HeuArgs A) {
  const u32 bi = blockIdx.x, qi = bi / 3, which = bi % 3;
  if (qi >= A.nq) return;
  u8* base = A.ws + (u64)bi * A.lay.stride;
  HW w;
  w.off = A.off;
  w.adj = A.adj;
  w.n = A.n;
  w.dmax = A.dmax;
  w.sta = base + A.lay.sta;
  w.ins = base + A.lay.ins;
  w.cnt = (u32*)(base + A.lay.cnt);
  w.deg = (u32*)(base + A.lay.deg);
  w.key = (double*)(base + A.lay.key);
  w.pos = (u32*)(base + A.lay.pos);
  w.list = (u32*)(base + A.lay.list);
  w.rmv = (u32*)(base + A.lay.rmv);
  w.th = (u32*)(base + A.lay.th);
  w.bkey = (u64*)(base + A.lay.bkey);
  w.gkey = (u64*)(base + A.lay.gkey);
  w.L = (HeuLds*)(base + A.lay.lds);
  if (LANE == 0) __hip_atomic_store(&w.L->m, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __threadfence();
  hsync();
  const u32 l = A.ql[qi], h = A.qh[qi], q = A.qq[qi];
  u32* out = A.H3 + (u64)bi * A.hcap;
  u32 sz = 0, kk = 0;
  if (which == 0)
    heu2(w, l, h, q, out, A.hcap, sz, kk);
  else if (which == 1)
    heu3(w, l, h, q, out, A.hcap, sz, kk);
  else
    heu4(w, l, h, q, out, A.hcap, sz, kk);
  if (LANE == 0) {
    A.kl3[bi] = kk;
    A.hs3[bi] = sz;
  }
}

// CSSC_heu's pick (inc/heuristic.h:321-348): heu4 when it ties with both,
// else heu3 when it ties with heu2, else heu2.
__global__ void __launch_bounds__(64) scs_heu_pick(HeuArgs A) {
  const u32 qi = blockIdx.x;
  if (qi >= A.nq) return;
  const u32 k2 = A.kl3[3 * qi], k3 = A.kl3[3 * qi + 1], k4 = A.kl3[3 * qi + 2];
  const u32 c = (k4 >= k3 && k4 >= k2) ? 2u : (k3 >= k2 ? 1u : 0u);
  const u32 sz = A.hs3[3 * qi + c];
  const u32* src = A.H3 + (u64)(3 * qi + c) * A.hcap;
  u32* out = A.H + (u64)qi * A.hcap;
  for (u32 i = LANE; i < sz && i < A.hcap; i += 64) out[i] = src[i];
  if (LANE == 0) {
    A.kl[qi] = c == 2 ? k4 : (c == 1 ? k3 : k2);
    A.hsize[qi] = sz;
  }
}

static u64 heu_al(u64 x) { return (x + 255) & ~255ull; }

extern "C" HeuLayout scs_heu_layout(u32 n, u32 hcap) {
  HeuLayout L{};
  const u64 nblk = (n + 63) / 64 + 1;
  u64 o = 0;
  L.sta = o; o = heu_al(o + n);
  L.ins = o; o = heu_al(o + n);
  L.cnt = o; o = heu_al(o + 4ull * n);
  L.deg = o; o = heu_al(o + 4ull * n);
  L.key = o; o = heu_al(o + 8ull * n);
  L.pos = o; o = heu_al(o + 4ull * n);
  L.list = o; o = heu_al(o + 4ull * (n + 64));
  L.rmv = o; o = heu_al(o + 4ull * (n + 64));
  L.th = o; o = heu_al(o + 4ull * (4ull * hcap + 64));
  L.bkey = o; o = heu_al(o + 8 * nblk);
  L.gkey = o; o = heu_al(o + 8 * (nblk / 64 + 2));
  L.lds = o; o = heu_al(o + sizeof(HeuLds));
  L.stride = o;
  return L;
}

// ws holds 3 * nq workspaces; kl3/hs3/H3 hold the three constructions'
// results of each query, kl/hsize/H the picked one.
extern "C" hipError_t scs_launch_heu(const u32* off, const u32* adj, u32 n, u32 dmax, u32 nq,
                                     const u32* ql, const u32* qh, const u32* qq, u8* ws,
                                     HeuLayout lay, u32 hcap, u32* kl, u32* hsize, u32* H,
                                     u32* kl3, u32* hs3, u32* H3, hipStream_t s) {
  HeuArgs A;
  A.off = off;
  A.adj = adj;
  A.n = n;
  A.dmax = dmax;
  A.nq = nq;
  A.hcap = hcap;
  A.ql = ql;
  A.qh = qh;
  A.qq = qq;
  A.ws = ws;
  A.lay = lay;
  A.kl = kl;
  A.hsize = hsize;
  A.H = H;
  A.kl3 = kl3;
  A.hs3 = hs3;
  A.H3 = H3;
  scs_heu_kernel<<<3 * nq, 64, 0, s>>>(A);
  scs_heu_pick<<<nq, 64, 0, s>>>(A);
  return hipGetLastError();
}
