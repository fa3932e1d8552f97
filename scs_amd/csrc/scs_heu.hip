// Greedy lower bounds on the GPU: the reference's CSSC_heu
// (inc/heuristic.h:321-348), the best of its three greedy constructions
// heu2 (:8-120), heu3 (:125-239) and heu4 (:242-317), one wave per query.
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
// Candidates sit in an append-only list; each block of 64 list positions
// keeps its best key, and a step recomputes only the blocks whose members
// changed (all of them when most did), so a step costs the changed
// vertices plus one pass over the block keys.
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
__device__ __forceinline__ u64 h_min64(u64 x) {
  for (int o = 32; o > 0; o >>= 1) {
    const u64 y = h_shfl64(x, (int)(LANE ^ (u32)o));
    x = y < x ? y : x;
  }
  return x;
}
__device__ __forceinline__ double h_shfld(double x, int src) {
  return __longlong_as_double((long long)h_shfl64((u64)__double_as_longlong(x), src));
}
// (key, id) pairs ordered as std::pair<double, uint32_t>: the larger key,
// then the larger id; invalid entries carry id INV and key -1
__device__ __forceinline__ bool pair_gt(double a, u32 ai, double b, u32 bi) {
  return a > b || (a == b && ai > bi && ai != 0xffffffffu) || (bi == 0xffffffffu && ai != 0xffffffffu);
}
__device__ __forceinline__ void h_maxpair(double& k, u32& id) {
  for (int o = 32; o > 0; o >>= 1) {
    const double k2 = h_shfld(k, (int)(LANE ^ (u32)o));
    const u32 i2 = (u32)__shfl((int)id, (int)(LANE ^ (u32)o));
    if (pair_gt(k2, i2, k, id)) {
      k = k2;
      id = i2;
    }
  }
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
  u64* bkey;
  double* bkd;
  u32* bki;
  u32* dbits;
  u32* dlist;
  HeuLds* L;
};

__device__ __forceinline__ void mark_dirty(HW& w, u32 b) {
  const u32 bit = 1u << (b & 31);
  if (!(atomicOr(&w.dbits[b >> 5], bit) & bit)) w.dlist[atomicAdd(&w.L->ndirty, 1u)] = b;
}

// Block keys.  MODE 2: max of (cnt+1, id) over frontier members; MODE 4: min
// of (deg, id) over members of S; MODE 3: max of (key, id) over frontier
// members.  A block whose members all left holds 0 / ~0 / (-1, INV).
// Eight blocks per pass (blocks blk[0..cnt), or 0..cnt-1 when blk is null),
// so eight chains of dependent loads are in flight together.
template <int MODE>
__device__ __forceinline__ void block_keys(HW& w, const u32* blk, u32 cnt, u32 m) {
  constexpr int U = 8;
  for (u32 i0 = 0; i0 < cnt; i0 += U) {
    u32 b[U], x[U];
    bool in[U];
#pragma unroll
    for (int t = 0; t < U; t++) {
      const u32 i = i0 + t;
      b[t] = i < cnt ? (blk ? blk[i] : i) : 0u;
      const u32 p = b[t] * 64 + LANE;
      in[t] = i < cnt && p < m;
      x[t] = in[t] ? w.list[p] : 0u;
    }
    if (MODE == 2) {
      u64 k[U];
#pragma unroll
      for (int t = 0; t < U; t++)
        k[t] = (in[t] && w.sta[x[t]] == 1) ? (((u64)(w.cnt[x[t]] + 1) << 32) | x[t]) : 0ull;
#pragma unroll
      for (int t = 0; t < U; t++) {
        const u64 r = h_max64(k[t]);
        if (LANE == 0 && i0 + t < cnt) w.bkey[b[t]] = r;
      }
    } else if (MODE == 4) {
      u64 k[U];
#pragma unroll
      for (int t = 0; t < U; t++)
        k[t] = (in[t] && w.ins[x[t]]) ? (((u64)w.deg[x[t]] << 32) | x[t]) : ~0ull;
#pragma unroll
      for (int t = 0; t < U; t++) {
        const u64 r = h_min64(k[t]);
        if (LANE == 0 && i0 + t < cnt) w.bkey[b[t]] = r;
      }
    } else {
      double k[U];
      u32 id[U];
#pragma unroll
      for (int t = 0; t < U; t++) {
        const bool f = in[t] && w.sta[x[t]] == 1;
        k[t] = f ? w.key[x[t]] : -1.0;
        id[t] = f ? x[t] : 0xffffffffu;
      }
#pragma unroll
      for (int t = 0; t < U; t++) {
        h_maxpair(k[t], id[t]);
        if (LANE == 0 && i0 + t < cnt) {
          w.bkd[b[t]] = k[t];
          w.bki[b[t]] = id[t];
        }
      }
    }
  }
}

// Recomputes the dirty blocks (every block when most are dirty) and clears
// the dirty marks.
template <int MODE>
__device__ __forceinline__ void refresh(HW& w) {
  hsync();
  const u32 m = hld(&w.L->m), nblk = (m + 63) / 64, nd = hld(&w.L->ndirty);
  if (nd == 0) return;
  if (nd * 4 >= nblk)
    block_keys<MODE>(w, nullptr, nblk, m);
  else
    block_keys<MODE>(w, w.dlist, nd, m);
  for (u32 i = LANE; i < nd; i += 64) w.dbits[w.dlist[i] >> 5] = 0u;
  hsync();
  if (LANE == 0) w.L->ndirty = 0;
  hsync();
}

// Appends v to the candidate list (one lane per call site).
__device__ __forceinline__ void push_cand(HW& w, u32 v) {
  const u32 p = atomicAdd(&w.L->m, 1u);
  w.list[p] = v;
  w.pos[v] = p;
  mark_dirty(w, p >> 6);
}

// min over H = th[0..nH) of deg
__device__ __forceinline__ u32 min_deg_h(HW& w, u32 nH) {
  u32 mn = 0xffffffffu;
  for (u32 i = LANE; i < nH; i += 64) mn = min(mn, w.deg[w.th[i]]);
  return h_min(mn);
}

__device__ __forceinline__ void cleanup(HW& w) {
  hsync();
  const u32 m = hld(&w.L->m), nd = hld(&w.L->ndirty);
  // dirty marks left by the last step (never refreshed)
  for (u32 i = LANE; i < nd; i += 64) w.dbits[w.dlist[i] >> 5] = 0u;
  for (u32 i = LANE; i < m; i += 64) {
    const u32 x = w.list[i];
    w.sta[x] = 0;
    w.ins[x] = 0;
    w.cnt[x] = 0;
    w.deg[x] = 0;
    w.key[x] = 0.0;
  }
  hsync();
  if (LANE == 0) {
    w.L->m = 0;
    w.L->ndirty = 0;
  }
  hsync();
}

// heu2 (inc/heuristic.h:8-120): H grows from q by the frontier vertex with
// the most H-neighbours (ties: larger id); the best prefix of size in
// [N1, N2] by minimum degree (ties: the later one) is kept.
__device__ __forceinline__ void heu2(HW& w, u32 N1, u32 N2, u32 q, u32* outH, u32 hcap, u32& osize, u32& okl) {
  if (LANE == 0) {
    w.sta[q] = 1;
    push_cand(w, q);
  }
  refresh<2>(w);
  u32 nH = 0, tkl = 0, tsize = 0;
  while (true) {
    const u32 nblk = (hld(&w.L->m) + 63) / 64;
    u64 best = 0;
    for (u32 b = LANE; b < nblk; b += 64) best = max(best, w.bkey[b]);
    best = h_max64(best);
    if (best == 0) break;
    const u32 v = (u32)best;
    if (LANE == 0) {
      w.th[nH] = v;
      w.sta[v] = 2;
      mark_dirty(w, w.pos[v] >> 6);
    }
    nH++;
    hsync();
    u32 dv = 0;
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 nb = w.adj[j];
      w.cnt[nb] += 1;
      const u8 s = w.sta[nb];
      if (s == 2) {
        w.deg[nb] += 1;
        dv++;
      } else if (s == 0) {
        w.sta[nb] = 1;
        push_cand(w, nb);
      } else {
        mark_dirty(w, w.pos[nb] >> 6);
      }
    }
    dv = h_sum(dv);
    if (LANE == 0) w.deg[v] += dv;
    hsync();
    if (nH >= N1) {
      const u32 mn = min_deg_h(w, nH);
      if (mn >= tkl) {
        tkl = mn;
        tsize = nH;
      }
    }
    if (nH == N2) break;
    refresh<2>(w);
  }
  for (u32 i = LANE; i < tsize && i < hcap; i += 64) outH[i] = w.th[i];
  osize = tsize;
  okl = tkl;
  cleanup(w);
}

// Score of a frontier vertex nb whose degree is at most HEU_SMALL, by one
// lane: 1/deg(w) over nb's H-neighbours w with deg(w) != 0 in ascending id
// order, then deg(nb)/dmax.
__device__ __forceinline__ double score_lane(const HW& w, u32 nb) {
  const u32 b = w.off[nb], e = w.off[nb + 1];
  double sc = 0.0;
  for (u32 k = b; k < e; k++) {
    const u32 x = w.adj[k];
    if (w.sta[x] == 2) {
      const u32 d = w.deg[x];
      if (d != 0) sc += 1.0 / (double)d;
    }
  }
  return sc + (double)(e - b) / (double)w.dmax;
}

// The same sum by the whole wave for a vertex of large degree: chunks of 64
// entries, the chunk's terms added one after another in lane order.
__device__ __forceinline__ double score_wave(const HW& w, u32 nb) {
  const u32 b = w.off[nb], e = w.off[nb + 1];
  double sc = 0.0;
  for (u32 c = b; c < e; c += 64) {
    const u32 k = c + LANE;
    double t = 0.0;
    bool on = false;
    if (k < e) {
      const u32 x = w.adj[k];
      if (w.sta[x] == 2) {
        const u32 d = w.deg[x];
        if (d != 0) {
          t = 1.0 / (double)d;
          on = true;
        }
      }
    }
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
    push_cand(w, nb);
  } else {
    if (sc > w.key[nb]) w.key[nb] = sc;
    mark_dirty(w, w.pos[nb] >> 6);
  }
}

// heu3 (inc/heuristic.h:125-239): H grows from q by the frontier vertex of
// largest connection score (the largest score pushed for it).
__device__ __forceinline__ void heu3(HW& w, u32 N1, u32 N2, u32 q, u32* outH, u32 hcap, u32& osize, u32& okl) {
  if (LANE == 0) {
    w.sta[q] = 1;
    w.key[q] = 0.0;
    push_cand(w, q);
  }
  refresh<3>(w);
  u32 nH = 0, tkl = 0, tsize = 0;
  while (true) {
    const u32 nblk = (hld(&w.L->m) + 63) / 64;
    double bk = -1.0;
    u32 bi = 0xffffffffu;
    for (u32 b = LANE; b < nblk; b += 64) {
      const double k = w.bkd[b];
      const u32 i = w.bki[b];
      if (pair_gt(k, i, bk, bi)) {
        bk = k;
        bi = i;
      }
    }
    h_maxpair(bk, bi);
    if (bi == 0xffffffffu) break;
    const u32 v = bi;
    if (LANE == 0) {
      w.th[nH] = v;
      w.sta[v] = 2;
      mark_dirty(w, w.pos[v] >> 6);
    }
    nH++;
    hsync();
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    u32 dv = 0;
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 nb = w.adj[j];
      if (w.sta[nb] == 2) {
        w.deg[nb] += 1;
        dv++;
      }
    }
    dv = h_sum(dv);
    if (LANE == 0) w.deg[v] += dv;
    hsync();
    // scores of v's neighbours outside H, against the degrees after v joined
    for (u32 c = b0; c < e0; c += 64) {
      const u32 j = c + LANE;
      u32 nb = 0;
      bool small = false, big = false;
      if (j < e0) {
        nb = w.adj[j];
        if (w.sta[nb] != 2) {
          const u32 dg = w.off[nb + 1] - w.off[nb];
          small = dg <= HEU_SMALL;
          big = !small;
        }
      }
      double sc = 0.0;
      if (small) sc = score_lane(w, nb);
      // keys are written after every score of the chunk is known: a score
      // reads only H's degrees, which no key update changes
      if (small) set_key3(w, nb, sc);
      u64 bm = __ballot(big);
      while (bm) {
        const int src = __ffsll((unsigned long long)bm) - 1;
        bm &= bm - 1;
        const u32 x = (u32)__shfl((int)nb, src);
        const double s2 = score_wave(w, x);
        if (LANE == 0) set_key3(w, x, s2);
      }
    }
    hsync();
    if (nH >= N1) {
      const u32 mn = min_deg_h(w, nH);
      if (mn >= tkl) {
        tkl = mn;
        tsize = nH;
      }
    }
    if (nH == N2) break;
    refresh<3>(w);
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
    w.L->m = m;
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
  const u32 nblk = (m + 63) / 64;
  block_keys<4>(w, nullptr, nblk, m);
  hsync();
  u32 nr = 0, rmv_idx = 0, mindeg = 0;
  while (true) {
    u64 best = ~0ull;
    for (u32 b = LANE; b < nblk; b += 64) best = min(best, w.bkey[b]);
    best = h_min64(best);
    if (best == ~0ull) break;
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
      mark_dirty(w, w.pos[v] >> 6);
    }
    nr++;
    hsync();
    const u32 b0 = w.off[v], e0 = w.off[v + 1];
    for (u32 j = b0 + LANE; j < e0; j += 64) {
      const u32 x = w.adj[j];
      if (w.ins[x]) {
        w.deg[x] -= 1;
        mark_dirty(w, w.pos[x] >> 6);
      }
    }
    refresh<4>(w);
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

// No LDS: the wave's two counters live in its workspace, so the kernel also
// finds room on CUs whose LDS the search kernel holds.
__global__ void __launch_bounds__(64) __attribute__((flatten)) scs_heu_kernel(HeuArgs A) {
  const u32 qi = blockIdx.x;
  if (qi >= A.nq) return;
  u8* base = A.ws + (u64)blockIdx.x * A.lay.stride;
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
  w.bkd = (double*)(base + A.lay.bkd);
  w.bki = (u32*)(base + A.lay.bki);
  w.dbits = (u32*)(base + A.lay.dbits);
  w.dlist = (u32*)(base + A.lay.dlist);
  w.L = (HeuLds*)(base + A.lay.lds);
  if (LANE == 0) {
    w.L->m = 0;
    w.L->ndirty = 0;
  }
  hsync();
  const u32 l = A.ql[qi], h = A.qh[qi], q = A.qq[qi];
  u32* out = A.H + (u64)qi * A.hcap;
  // each construction writes its H to its own third of th's tail, then the
  // best is copied out (CSSC_heu: heu4 on ties with both, heu3 over heu2)
  u32* h2 = w.th + A.hcap;
  u32* h3 = h2 + A.hcap;
  u32* h4 = h3 + A.hcap;
  u32 s2, k2, s3, k3, s4, k4;
  heu2(w, l, h, q, h2, A.hcap, s2, k2);
  heu3(w, l, h, q, h3, A.hcap, s3, k3);
  heu4(w, l, h, q, h4, A.hcap, s4, k4);
  hsync();
  const u32* src;
  u32 sz, kk;
  if (k4 >= k3 && k4 >= k2) {
    src = h4;
    sz = s4;
    kk = k4;
  } else if (k3 >= k2) {
    src = h3;
    sz = s3;
    kk = k3;
  } else {
    src = h2;
    sz = s2;
    kk = k2;
  }
  for (u32 i = LANE; i < sz && i < A.hcap; i += 64) out[i] = src[i];
  if (LANE == 0) {
    A.kl[qi] = kk;
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
  L.bkd = o; o = heu_al(o + 8 * nblk);
  L.bki = o; o = heu_al(o + 4 * nblk);
  L.dbits = o; o = heu_al(o + 4 * (nblk / 32 + 1));
  L.dlist = o; o = heu_al(o + 4 * nblk);
  L.lds = o; o = heu_al(o + sizeof(HeuLds));
  L.stride = o;
  return L;
}

extern "C" hipError_t scs_launch_heu(const u32* off, const u32* adj, u32 n, u32 dmax, u32 nq,
                                     const u32* ql, const u32* qh, const u32* qq, u8* ws,
                                     HeuLayout lay, u32 hcap, u32* kl, u32* hsize, u32* H,
                                     hipStream_t s) {
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
  scs_heu_kernel<<<nq, 64, 0, s>>>(A);
  return hipGetLastError();
}
