// Branch-and-bound search engine for size-bounded community search, gfx950.
//
// Replaces ProcessTask / FindDoms / Expand / RemoveCompletedTasks
// (src/helpers.cc:371-889) and the host level loop (main.cu:261-442).
//
// Design: one wavefront (64 lanes, one workgroup) is one depth-first
// worker.  It keeps, for the query it is bound to, the full search state in
// its own HBM region: vertex status (R = candidate, C = chosen, X = out),
// dC = |N(v) & C|, dR = |N(v) & R|, an explicit list of R, a histogram of
// dC over R and an undo trail.  A node costs O(changes x degree) instead of
// the reference's O(|T| x deg x |T|) rescans (findIndexKernel,
// src/helpers.cc:23-47).  Waves that run out of work take tasks from a
// bounded lock-free pool that busy waves fill by donating the unexplored
// "exclude u*" branch of their shallowest frame.  The kernel runs for a
// bounded time slice and returns; all state persists in HBM between slices,
// so the host never has to wait on a spinning kernel.
//
// Node semantics (identical to the CPU oracle, so a single wave expands
// exactly the oracle's nodes): min degree of C when l <= |C| <= h, rule 3,
// rules 1+2 to a fixpoint, U_d and U_nr, ustar by fixed-point connection
// score, the set dominated by ustar, then children A, B_0.., Z.
#include "scs_internal.h"

#define LANE (threadIdx.x)
#define INV 0xffffffffu

// ---------------------------------------------------------------------------
// Wave primitives (64 lanes; one wave per workgroup).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wsync() { __syncthreads(); }
__device__ __forceinline__ u32 wmin(u32 x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, (u32)__shfl_xor((int)x, o));
  return x;
}
__device__ __forceinline__ u32 wsum(u32 x) {
  for (int o = 32; o > 0; o >>= 1) x += (u32)__shfl_xor((int)x, o);
  return x;
}
__device__ __forceinline__ u64 wmax64(u64 x) {
  for (int o = 32; o > 0; o >>= 1) {
    u64 y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return x;
}
__device__ __forceinline__ u32 bcast(u32 x) { return __shfl(x, 0); }
__device__ __forceinline__ u64 lanes_below() {
  return LANE == 0 ? 0ull : (~0ull >> (64 - LANE));
}
__device__ __forceinline__ u32 aload(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 aload64(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void astore64(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u64 recip32(u32 d) {
  return d == 0 ? (1ull << 32) : ((1ull << 32) / d);
}
__device__ __forceinline__ u64 rterm(u32 dR, u32 dmax) {
  return dmax ? (((u64)dR << 32) / dmax) : 0ull;
}

// Diameter bound per node (src/helpers.cc:974-990), 32-bit arithmetic.
__device__ u32 ubd_node(u32 kl, u32 h) {
  u32 ub = h - 1;
  for (u32 d = 1; d <= h; d++) {
    if (d == 1 || d == 2) {
      if (kl + d > h) return d - 1;
    } else {
      u32 min_n = kl + d + 1 + (d / 3) * (kl - 2);
      if (h < min_n) return d - 1;
    }
  }
  return ub;
}

// ---------------------------------------------------------------------------
// Per-wave state view.
// ---------------------------------------------------------------------------
struct W {
  u8* st;
  u8* nmark;
  u8* queued;
  u16* dC;
  u32* dR;
  u32* pos;
  u32* rlist;
  u64* acc;
  u32* work;
  u32* front;
  u32* trail;
  Frame* frames;
  u32* clist;
  u32* hist;   // h+2 counters of dC over R
  WaveHdr* hdr;
  // query
  QuerySlot* Q;
  const u32* off;
  const u32* adj;
  const u32* dist;
  u32 K, l, h, flags, limit_doms, dmax, ku;
  // cached header fields (wave-uniform)
  u32 nC, nR, ntrail, nframes;
};

__device__ void bind(W& w, const EngineArgs& A, u32 wave) {
  u8* base = A.wave_mem + (u64)wave * A.lay.stride;
  w.st = base + A.lay.st;
  w.nmark = base + A.lay.nmark;
  w.queued = base + A.lay.queued;
  w.dC = (u16*)(base + A.lay.dC);
  w.dR = (u32*)(base + A.lay.dR);
  w.pos = (u32*)(base + A.lay.pos);
  w.rlist = (u32*)(base + A.lay.rlist);
  w.acc = (u64*)(base + A.lay.acc);
  w.work = (u32*)(base + A.lay.work);
  w.front = (u32*)(base + A.lay.front);
  w.trail = (u32*)(base + A.lay.trail);
  w.frames = (Frame*)(base + A.lay.frames);
  w.clist = (u32*)(base + A.lay.clist);
  w.hist = (u32*)(base + A.lay.hist);
  w.hdr = (WaveHdr*)(base + A.lay.hdr);
}

__device__ void bind_query(W& w, QuerySlot* Q) {
  w.Q = Q;
  w.off = Q->off;
  w.adj = Q->adj;
  w.dist = Q->dist;
  w.K = Q->K;
  w.l = Q->l;
  w.h = Q->h;
  w.flags = Q->flags;
  w.limit_doms = Q->limit_doms;
  w.dmax = Q->dmax;
  w.ku = Q->ku;
}

// ---------------------------------------------------------------------------
// State updates.  Counters of X vertices are frozen; moves and removals only
// touch neighbours that are still live, and undo applies the exact inverse
// (the trail restores statuses in reverse order, so the same neighbours are
// live again when an entry is undone).
// ---------------------------------------------------------------------------
__device__ void rlist_remove(W& w, u32 v) {  // lane 0 only
  u32 p = w.pos[v];
  u32 last = w.rlist[w.nR - 1];
  w.rlist[p] = last;
  w.pos[last] = p;
  w.nR--;
}
__device__ void rlist_add(W& w, u32 v) {  // lane 0 only
  w.rlist[w.nR] = v;
  w.pos[v] = w.nR;
  w.nR++;
}

__device__ void move_c(W& w, u32 v) {
  const u32 b = w.off[v], e = w.off[v + 1];
  for (u32 j = b + LANE; j < e; j += 64) {
    u32 x = w.adj[j];
    u8 s = w.st[x];
    if (s == ST_X) continue;
    u32 dc = w.dC[x];
    if (s == ST_R) {
      atomicSub(&w.hist[dc], 1u);
      atomicAdd(&w.hist[dc + 1], 1u);
    }
    w.dC[x] = (u16)(dc + 1);
    w.dR[x] -= 1;
  }
  wsync();
  if (LANE == 0) {
    w.st[v] = ST_C;
    w.hist[w.dC[v]] -= 1;
    rlist_remove(w, v);
    w.clist[w.nC] = v;
    w.nC++;
    w.trail[w.ntrail] = v | 0x80000000u;
    w.ntrail++;
  }
  w.nR = bcast(w.nR);
  w.nC = bcast(w.nC);
  w.ntrail = bcast(w.ntrail);
  wsync();
}

__device__ void remove_r(W& w, u32 v) {
  const u32 b = w.off[v], e = w.off[v + 1];
  for (u32 j = b + LANE; j < e; j += 64) {
    u32 x = w.adj[j];
    if (w.st[x] != ST_X) w.dR[x] -= 1;
  }
  wsync();
  if (LANE == 0) {
    w.st[v] = ST_X;
    w.hist[w.dC[v]] -= 1;
    rlist_remove(w, v);
    w.trail[w.ntrail] = v;
    w.ntrail++;
  }
  w.nR = bcast(w.nR);
  w.ntrail = bcast(w.ntrail);
  wsync();
}

__device__ void undo_to(W& w, u32 mark) {
  while (w.ntrail > mark) {
    u32 e = w.trail[w.ntrail - 1];
    u32 v = e & 0x7fffffffu;
    bool wasC = (e & 0x80000000u) != 0;
    // restore v first so that its own status is R while neighbours update
    if (LANE == 0) {
      w.st[v] = ST_R;
      w.hist[w.dC[v]] += 1;
      rlist_add(w, v);
      if (wasC) w.nC--;
      w.ntrail--;
    }
    wsync();
    const u32 b = w.off[v], en = w.off[v + 1];
    for (u32 j = b + LANE; j < en; j += 64) {
      u32 x = w.adj[j];
      u8 s = w.st[x];
      if (s == ST_X) continue;
      if (wasC) {
        u32 dc = w.dC[x];
        if (s == ST_R) {
          atomicSub(&w.hist[dc], 1u);
          atomicAdd(&w.hist[dc - 1], 1u);
        }
        w.dC[x] = (u16)(dc - 1);
      }
      w.dR[x] += 1;
    }
    w.nR = bcast(w.nR);
    w.nC = bcast(w.nC);
    w.ntrail = bcast(w.ntrail);
    wsync();
  }
}

// ---------------------------------------------------------------------------
// Witness and kl.
// ---------------------------------------------------------------------------
__device__ void offer(W& w, u32 mn) {
  QuerySlot* Q = w.Q;
  u32 improved = 0;
  if (LANE == 0) {
    u32 old = atomicMax(&Q->kl, mn);
    if (old < mn) improved = 1;
    else if (old == mn && aload(&Q->wsize) == 0) improved = 1;
  }
  improved = bcast(improved);
  if (!improved) return;
  u32 got = 0;
  if (LANE == 0) {
    for (u32 spin = 0; spin < (1u << 20); spin++) {
      if (atomicCAS(&Q->wlock, 0u, 1u) == 0u) {
        got = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  got = bcast(got);
  if (!got) return;
  __threadfence();
  u32 ok = 0;
  if (LANE == 0) ok = (Q->wsize == 0 || Q->wkl < mn) ? 1u : 0u;
  ok = bcast(ok);
  if (ok) {
    for (u32 i = LANE; i < w.nC; i += 64) Q->witness[i] = Q->orig[w.clist[i]];
    __threadfence();
    if (LANE == 0) {
      Q->wkl = mn;
      Q->wsize = w.nC;
    }
  }
  __threadfence();
  if (LANE == 0) atomicExch(&Q->wlock, 0u);
  wsync();
}

__device__ __forceinline__ u32 cur_kl(const W& w) { return aload(&w.Q->kl); }

// ---------------------------------------------------------------------------
// Node processing.
// ---------------------------------------------------------------------------
__device__ bool cond12(const W& w, u32 v, u32 Cs, u32 ubD, u32 kl) {
  u32 dc = w.dC[v], dr = w.dR[v];
  if ((w.flags & QF_RED1) && min(dr + dc, dc + w.h - Cs - 1) <= kl) return true;
  if ((w.flags & QF_RED2) && ubD < w.dist[v]) return true;
  return false;
}

// Worklist helpers: work[] holds queued vertices; count kept in *cnt (LDS).
__device__ __forceinline__ void try_push(W& w, u32 x, u32 Cs, u32 ubD, u32 kl, u32* cnt) {
  if (w.st[x] == ST_R && !w.queued[x] && cond12(w, x, Cs, ubD, kl)) {
    // queued[] is a byte array: claim with a 32-bit CAS on the containing word
    u32* wp = (u32*)((uintptr_t)(w.queued + x) & ~(uintptr_t)3);
    u32 sh = ((uintptr_t)(w.queued + x) & 3) * 8;
    u32 old = atomicOr(wp, 1u << sh);
    if (!((old >> sh) & 0xffu)) w.work[atomicAdd(cnt, 1u)] = x;
  }
}

// Rule 1/2 closure (see the CPU oracle's TrailSolver::r12 for the seeding
// argument).  Returns true if anything was removed.
__device__ bool r12(W& w, u32 from, bool full, u32 kl, u32* lds_cnt) {
  const u32 Cs = w.nC;
  const u32 ubD = ubd_node(kl, w.h);
  if (LANE == 0) *lds_cnt = 0;
  wsync();
  if (!full && (w.flags & QF_RED1)) {
    long T = (long)kl + (long)Cs + 1 - (long)w.h;
    if (T >= 0) {
      u32 c = 0;
      for (u32 d = LANE; d <= (u32)T && d < w.h + 2; d += 64) c += w.hist[d];
      if (wsum(c)) full = true;
    }
  }
  if (full) {
    for (u32 i = LANE; i < w.nR; i += 64) try_push(w, w.rlist[i], Cs, ubD, kl, lds_cnt);
  } else {
    for (u32 t = from; t < w.ntrail; t++) {
      u32 e = w.trail[t];
      if (e & 0x80000000u) continue;
      const u32 b = w.off[e], en = w.off[e + 1];
      for (u32 j = b + LANE; j < en; j += 64) try_push(w, w.adj[j], Cs, ubD, kl, lds_cnt);
    }
  }
  wsync();
  u32 head = 0;
  u32 cnt = bcast(*lds_cnt);
  bool removed = cnt > 0;
  while (head < cnt) {
    u32 v = w.work[head];
    head++;
    if (LANE == 0) w.queued[v] = 0;
    remove_r(w, v);
    const u32 b = w.off[v], en = w.off[v + 1];
    for (u32 j = b + LANE; j < en; j += 64) try_push(w, w.adj[j], Cs, ubD, kl, lds_cnt);
    wsync();
    cnt = bcast(*lds_cnt);
  }
  return removed;
}

__device__ bool tight_exists(const W& w, u32 kl) {
  u32 f = 0;
  for (u32 i = LANE; i < w.nC; i += 64) {
    u32 u = w.clist[i];
    if (w.dC[u] + w.dR[u] == kl + 1 && w.dR[u] > 0) f = 1;
  }
  return __ballot(f) != 0;
}

// U_nr water-filling on one lane: C degrees (histogram) get +1 on the
// smallest entries for each of the k largest R->C degrees (src/helpers.cc:
// 1066-1101 with a correct sort).  Values are capped at h+1: a minimum of
// h+1 or more can never be <= kl, so the prune decision is unchanged.
__device__ u32 unr_bound(W& w, u32 k, u32* c) {
  const u32 h = w.h;
  const u32 top = h + 1;
  for (u32 i = LANE; i <= top; i += 64) c[i] = 0;
  wsync();
  for (u32 i = LANE; i < w.nC; i += 64) atomicAdd(&c[min((u32)w.dC[w.clist[i]], top)], 1u);
  wsync();
  u32 res = INV;
  if (LANE == 0) {
    u32 lo = 0;
    while (lo < top && c[lo] == 0) lo++;
    const u32 m = w.nC;
    int d = (int)top;
    u32 left = w.hist[d];
    for (u32 i = 0; i < k; i++) {
      while (d > 0 && left == 0) left = w.hist[--d];
      if (d <= 0) break;  // the remaining R vertices have dC == 0
      left--;
      const u32 x = min((u32)d, m);
      u32 acc = 0, b = lo;
      while (acc + c[b] < x) {
        acc += c[b];
        b++;
      }
      const u32 take_top = x - acc;
      c[b] -= take_top;
      c[min(b + 1, top)] += take_top;
      for (int bb = (int)b - 1; bb >= (int)lo; bb--) {
        const u32 t = c[bb];
        c[bb] = 0;
        c[bb + 1] += t;
      }
      while (lo < top && c[lo] == 0) lo++;
    }
    res = lo;
  }
  res = bcast(res);
  wsync();
  return res;
}

struct NodeOut {
  u32 branch;
  u32 ustar;
  u32 nD;
  u32 klc;  // kl at the end of the closure
};

// lds scratch layout per workgroup
struct Lds {
  u32 cnt;
  u32 nfront;
  u32 ncand;
  u32 pad;
  u32 D[SCS_FRAME_DOMS];
  u32 chist[SCS_HCAP + 2];
};

__device__ NodeOut process(W& w, u32 from, u32 kl_seen, Lds* L) {
  NodeOut o;
  o.branch = 0;
  o.nD = 0;
  o.ustar = INV;
  const u32 h = w.h;
  u32 kl = cur_kl(w);
  bool full = (kl != kl_seen);
  while (true) {
    const u32 Cs = w.nC;
    if (Cs >= w.l && Cs <= h) {
      u32 mn = INV;
      for (u32 i = LANE; i < Cs; i += 64) mn = min(mn, (u32)w.dC[w.clist[i]]);
      mn = wmin(mn);
      if (mn != INV && mn >= kl) {
        offer(w, mn);
      }
      u32 nk = cur_kl(w);
      if (nk != kl) full = true;
      kl = nk;
    }
    if (Cs >= h) {
      o.klc = kl;
      return o;
    }
    if ((w.flags & QF_RED3) && tight_exists(w, kl)) {
      // collect tight C vertices first (their dC+dR is invariant under moves)
      if (LANE == 0) L->cnt = 0;
      wsync();
      for (u32 i = LANE; i < Cs; i += 64) {
        u32 u = w.clist[i];
        if (w.dC[u] + w.dR[u] == kl + 1 && w.dR[u] > 0) w.work[atomicAdd(&L->cnt, 1u)] = u;
      }
      wsync();
      u32 nt = bcast(L->cnt);
      bool over = false;
      for (u32 t = 0; t < nt && !over; t++) {
        u32 u = w.work[t];
        const u32 b = w.off[u], e = w.off[u + 1];
        for (u32 j = b; j < e; j++) {
          u32 x = w.adj[j];
          if (w.st[x] == ST_R) {
            move_c(w, x);
            if (w.nC > h) {
              over = true;
              break;
            }
          }
        }
      }
      if (over) {
        o.klc = kl;
        return o;
      }
      continue;
    }
    bool removed = r12(w, from, full, kl, &L->cnt);
    from = w.ntrail;
    full = false;
    if (removed && (w.flags & QF_RED3) && tight_exists(w, kl)) continue;
    break;
  }
  o.klc = kl;
  const u32 Cs = w.nC;
  if (Cs + w.nR < w.l || w.nR == 0) return o;
  // bounds
  u32 ub = INV;
  if (w.flags & QF_PRUN1) {
    u32 b = INV;
    for (u32 i = LANE; i < Cs; i += 64) {
      u32 u = w.clist[i];
      u32 dc = w.dC[u], dr = w.dR[u];
      b = min(b, min(dc + h - Cs, dc + dr));
    }
    ub = wmin(b);
  }
  if ((w.flags & QF_PRUN2) && ub > kl) {
    u32 k = min(h - Cs, w.nR);
    ub = min(ub, unr_bound(w, k, L->chist));
  }
  if (ub <= kl) return o;
  // connection scores by scattering from C into acc[]; frontier = R with dC>0
  if (LANE == 0) L->nfront = 0;
  wsync();
  const bool single0 = (Cs == 1 && w.dC[w.clist[0]] == 0);
  for (u32 i = 0; i < Cs; i++) {
    u32 u = w.clist[i];
    u64 r = recip32(w.dC[u]);
    const u32 b = w.off[u], e = w.off[u + 1];
    for (u32 j = b + LANE; j < e; j += 64) {
      u32 x = w.adj[j];
      if (w.st[x] != ST_R) continue;
      u64 old = atomicAdd((unsigned long long*)&w.acc[x], (unsigned long long)r);
      if (old == 0) w.front[atomicAdd(&L->nfront, 1u)] = x;
    }
  }
  wsync();
  const u32 nf = bcast(L->nfront);
  // argmax score, ties to the lowest id
  u64 best = 0;
  u32 bid = INV;
  for (u32 i = LANE; i < nf; i += 64) {
    u32 x = w.front[i];
    u64 s = w.acc[x] + rterm(w.dR[x], w.dmax);
    if (s > best || (s == best && x < bid)) {
      best = s;
      bid = x;
    }
  }
  u64 gbest = wmax64(best);
  u32 cand = (best == gbest) ? bid : INV;
  const u32 us = wmin(cand);
  if (nf == 0 || us == INV) {
    for (u32 i = LANE; i < nf; i += 64) w.acc[w.front[i]] = 0;
    wsync();
    return o;
  }
  // dominated set of us
  u32 nD = 0;
  if ((w.flags & QF_DOMS) && w.limit_doms > 0) {
    const u32 lu = w.dC[us] + w.dR[us];
    const u32 ub0 = w.off[us], ue = w.off[us + 1];
    for (u32 j = ub0 + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 1;
    wsync();
    if (LANE == 0) L->ncand = 0;
    wsync();
    // each lane tests one frontier vertex
    for (u32 i = LANE; i < nf; i += 64) {
      u32 x = w.front[i];
      if (x == us || w.dC[x] + w.dR[x] > lu + 1) continue;
      bool dom = true;
      const u32 b = w.off[x], e = w.off[x + 1];
      for (u32 j = b; j < e; j++) {
        u32 y = w.adj[j];
        if (w.st[y] == ST_X) continue;
        if (y != us && !w.nmark[y]) {
          dom = false;
          break;
        }
      }
      if (dom) w.work[atomicAdd(&L->ncand, 1u)] = x;
    }
    wsync();
    for (u32 j = ub0 + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 0;
    const u32 nc = bcast(L->ncand);
    // select the top limit_doms by (score desc, id asc); score uses the
    // C neighbours with dC != 0 (only C = {q} with dC(q) = 0 differs)
    const u32 want = min(nc, w.limit_doms);
    for (u32 t = 0; t < want; t++) {
      u64 bs = 0;
      u32 bx = INV;
      for (u32 i = LANE; i < nc; i += 64) {
        u32 x = w.work[i];
        if (x == INV) continue;
        u64 s = (single0 ? w.acc[x] - (1ull << 32) : w.acc[x]) + rterm(w.dR[x], w.dmax);
        if (bx == INV || s > bs || (s == bs && x < bx)) {
          bs = s;
          bx = x;
        }
      }
      // pack: compare score first, then smaller id
      u64 gb = 0;
      {
        u64 key = (bx == INV) ? 0ull : bs + 1;  // +1 keeps valid keys above 0
        gb = wmax64(key);
      }
      u32 c2 = (bx != INV && bs + 1 == gb) ? bx : INV;
      u32 pick = wmin(c2);
      if (LANE == 0) L->D[t] = pick;
      for (u32 i = LANE; i < nc; i += 64)
        if (w.work[i] == pick) w.work[i] = INV;
      wsync();
    }
    nD = want;
  }
  for (u32 i = LANE; i < nf; i += 64) w.acc[w.front[i]] = 0;
  wsync();
  o.branch = 1;
  o.ustar = us;
  o.nD = nD;
  return o;
}

// ---------------------------------------------------------------------------
// Task pool (Vyukov bounded MPMC ring).  Lane 0 claims; the wave copies.
// ---------------------------------------------------------------------------
__device__ int pool_try_enqueue(const Pool& P, u64* out_pos) {
  u64 pos = aload64(P.enq);
  for (int tries = 0; tries < 64; tries++) {
    u64 s = aload64(&P.seq[pos % P.nslots]);
    long long dif = (long long)s - (long long)pos;
    if (dif == 0) {
      u64 exp = pos;
      if (__hip_atomic_compare_exchange_strong(P.enq, &exp, pos + 1, __ATOMIC_ACQ_REL,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        *out_pos = pos;
        return 1;
      }
      pos = exp;
    } else if (dif < 0) {
      return 0;
    } else {
      pos = aload64(P.enq);
    }
  }
  return 0;
}
__device__ int pool_try_dequeue(const Pool& P, u64* out_pos) {
  u64 pos = aload64(P.deq);
  for (int tries = 0; tries < 64; tries++) {
    u64 s = aload64(&P.seq[pos % P.nslots]);
    long long dif = (long long)s - (long long)(pos + 1);
    if (dif == 0) {
      u64 exp = pos;
      if (__hip_atomic_compare_exchange_strong(P.deq, &exp, pos + 1, __ATOMIC_ACQ_REL,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        *out_pos = pos;
        return 1;
      }
      pos = exp;
    } else if (dif < 0) {
      return 0;
    } else {
      pos = aload64(P.deq);
    }
  }
  return 0;
}

// Rebuild the wave's state from a task payload.
__device__ void rebuild(W& w, const u32* task, Lds* L) {
  const u32 nC = task[1], nL = task[2], form = task[3];
  const u32* C = task + 4;
  const u32* Lst = task + 4 + nC;
  const u32 K = w.K;
  const u8 fill = form ? ST_X : ST_R;
  for (u32 v = LANE; v < K; v += 64) {
    w.st[v] = fill;
    w.queued[v] = 0;
    w.nmark[v] = 0;
    w.acc[v] = 0;
  }
  wsync();
  for (u32 i = LANE; i < nL; i += 64) w.st[Lst[i]] = form ? ST_R : ST_X;
  wsync();
  for (u32 i = LANE; i < nC; i += 64) {
    w.st[C[i]] = ST_C;
    w.clist[i] = C[i];
  }
  for (u32 i = LANE; i < w.h + 2; i += 64) w.hist[i] = 0;
  wsync();
  // R list by ballot compaction (ascending ids)
  u32 base = 0;
  for (u32 v0 = 0; v0 < K; v0 += 64) {
    u32 v = v0 + LANE;
    bool r = v < K && w.st[v] == ST_R;
    u64 bm = __ballot(r);
    if (r) {
      u32 p = base + __popcll(bm & lanes_below());
      w.rlist[p] = v;
      w.pos[v] = p;
    }
    base += __popcll(bm);
  }
  w.nR = base;
  w.nC = nC;
  wsync();
  // degrees of live vertices, one lane per vertex
  for (u32 i = LANE; i < w.nR + nC; i += 64) {
    u32 v = i < w.nR ? w.rlist[i] : w.clist[i - w.nR];
    u32 c = 0, r = 0;
    for (u32 j = w.off[v]; j < w.off[v + 1]; j++) {
      u8 s = w.st[w.adj[j]];
      c += (s == ST_C);
      r += (s == ST_R);
    }
    w.dC[v] = (u16)c;
    w.dR[v] = r;
    if (i < w.nR) atomicAdd(&w.hist[min(c, w.h + 1)], 1u);
  }
  w.ntrail = 0;
  w.nframes = 0;
  wsync();
}

// Serialise the Z child of frame fi (the unexplored "exclude u*" branch)
// into dst; returns the payload length in words.
__device__ u32 serialize_z(W& w, u32 fi, u32* dst, u32 qslot, u32 cap, Lds* L) {
  const Frame& f = w.frames[fi];
  // mark vertices changed after f.mark (they were R at the frame)
  for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) w.nmark[w.trail[t] & 0x7fffffffu] = 1;
  wsync();
  for (u32 i = LANE; i < f.nD; i += 64) w.nmark[f.D[i]] = 2;
  if (LANE == 0) w.nmark[f.ustar] = 2;
  wsync();
  // R at Z = (current R) + (changed after mark) - ustar - D
  const u32 nRz = w.nR + (w.ntrail - f.mark) - 1 - f.nD;
  const u32 nXz = w.K - f.csize - nRz;
  const u32 form = nRz <= nXz ? 1u : 0u;
  const u32 nL = form ? nRz : nXz;
  if (4 + f.csize + nL > cap) {
    for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) w.nmark[w.trail[t] & 0x7fffffffu] = 0;
    wsync();
    for (u32 i = LANE; i < f.nD; i += 64) w.nmark[f.D[i]] = 0;
    if (LANE == 0) w.nmark[f.ustar] = 0;
    wsync();
    return 0;
  }
  for (u32 i = LANE; i < f.csize; i += 64) dst[4 + i] = w.clist[i];
  u32* out = dst + 4 + f.csize;
  u32 base = 0;
  for (u32 v0 = 0; v0 < w.K; v0 += 64) {
    u32 v = v0 + LANE;
    bool take = false;
    if (v < w.K) {
      // at the frame: nmark 2 = ustar or D (out in Z), 1 = changed after the
      // frame's mark (was R), 0 = unchanged (current status holds)
      u8 s = w.st[v], m = w.nmark[v];
      bool inRz = (m == 1) || (m == 0 && s == ST_R);
      bool inC = (m == 0 && s == ST_C);
      bool inXz = !inRz && !inC;
      take = form ? inRz : inXz;
    }
    u64 bm = __ballot(take);
    if (take) out[base + __popcll(bm & lanes_below())] = v;
    base += __popcll(bm);
  }
  if (LANE == 0) {
    dst[0] = qslot;
    dst[1] = f.csize;
    dst[2] = base;
    dst[3] = form;
  }
  wsync();
  for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) w.nmark[w.trail[t] & 0x7fffffffu] = 0;
  wsync();
  for (u32 i = LANE; i < f.nD; i += 64) w.nmark[f.D[i]] = 0;
  if (LANE == 0) w.nmark[f.ustar] = 0;
  wsync();
  return 4 + f.csize + base;
}

__device__ void save_hdr(W& w, u32 active, u32 qslot, u32 from, u32 kl_seen, u32 pend) {
  if (LANE == 0) {
    WaveHdr* H = w.hdr;
    H->active = active;
    H->qslot = qslot;
    H->nframes = w.nframes;
    H->trail = w.ntrail;
    H->nC = w.nC;
    H->nR = w.nR;
    H->from = from;
    H->kl_seen = kl_seen;
    H->pend_nodes = pend;
  }
  wsync();
}

__device__ void finish_task(W& w, const EngineArgs& A, u32 qslot) {
  if (LANE == 0) {
    QuerySlot* Q = &A.slots[qslot];
    u32 left = atomicSub(&Q->pending, 1u) - 1;
    if (left == 0) {
      __threadfence();
      atomicExch(&Q->state, (u32)QS_DONE);
      atomicSub(&A.glob->active_q, 1u);
    }
  }
  wsync();
}

// ---------------------------------------------------------------------------
// The persistent epoch kernel.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) scs_epoch(EngineArgs A) {
  __shared__ Lds L;
  const u32 wave = blockIdx.x;
  if (wave >= A.nwaves) return;
  const u64 t0 = wall_clock64();
  const u64 deadline = t0 + A.slice_ticks;
  W w;
  bind(w, A, wave);
  WaveHdr* H = w.hdr;
  u32 active = H->active, qslot = H->qslot;
  u32 from = H->from, kl_seen = H->kl_seen, pend = H->pend_nodes;
  if (active) {
    bind_query(w, &A.slots[qslot]);
    w.nframes = H->nframes;
    w.ntrail = H->trail;
    w.nC = H->nC;
    w.nR = H->nR;
  }
  bool idle_counted = false;
  u32 idle_spins = 0;
  while (true) {
    if (wall_clock64() > deadline) break;
    if (!active) {
      // take a task from the pool
      u64 pos = 0;
      u32 got = 0;
      if (LANE == 0) got = pool_try_dequeue(A.pool, &pos);
      got = bcast(got);
      if (!got) {
        if (!idle_counted) {
          if (LANE == 0) atomicAdd(&A.glob->idle, 1u);
          idle_counted = true;
        }
        if (aload(&A.glob->active_q) == 0) break;
        if (++idle_spins > (1u << 22)) break;
        __builtin_amdgcn_s_sleep(20);
        continue;
      }
      if (idle_counted) {
        if (LANE == 0) atomicSub(&A.glob->idle, 1u);
        idle_counted = false;
      }
      u32 plo = __shfl((u32)pos, 0), phi = __shfl((u32)(pos >> 32), 0);
      pos = ((u64)phi << 32) | plo;
      const u32* task = A.pool.data + (pos % A.pool.nslots) * (u64)A.pool.slot_words;
      __threadfence();
      qslot = task[0];
      bind_query(w, &A.slots[qslot]);
      if (aload(&w.Q->stop)) {
        wsync();
        if (LANE == 0) astore64(&A.pool.seq[pos % A.pool.nslots], pos + A.pool.nslots);
        finish_task(w, A, qslot);
        continue;
      }
      rebuild(w, task, &L);
      __threadfence();
      if (LANE == 0) {
        astore64(&A.pool.seq[pos % A.pool.nslots], pos + A.pool.nslots);
        atomicAdd((unsigned long long*)&A.glob->rebuilds, 1ull);
      }
      active = 1;
      from = 0;
      kl_seen = INV;
      pend = 0;
      wsync();
    }
    // one node: stop / optimality / budget checks first (as the oracle)
    QuerySlot* Q = w.Q;
    u32 quit = 0;
    if (LANE == 0) {
      if (aload(&Q->stop)) {
        quit = 1;
      } else if (aload(&Q->kl) >= Q->ku) {
        atomicExch(&Q->stop, 1u);
        quit = 1;
      } else if (Q->node_budget && aload64(&Q->nodes) + pend >= Q->node_budget) {
        atomicExch(&Q->capped, 1u);
        atomicExch(&Q->stop, 1u);
        quit = 1;
      }
    }
    quit = bcast(quit);
    if (quit) {
      active = 0;
      if (pend) {
        if (LANE == 0) {
          atomicAdd((unsigned long long*)&Q->nodes, (unsigned long long)pend);
          atomicAdd((unsigned long long*)&A.glob->nodes, (unsigned long long)pend);
        }
        pend = 0;
      }
      finish_task(w, A, qslot);
      continue;
    }
    NodeOut o = process(w, from, kl_seen, &L);
    pend++;
    if (pend >= A.node_flush) {
      if (LANE == 0) {
        atomicAdd((unsigned long long*)&Q->nodes, (unsigned long long)pend);
        atomicAdd((unsigned long long*)&A.glob->nodes, (unsigned long long)pend);
      }
      pend = 0;
      wsync();
    }
    if (o.branch) {
      u32 fi = w.nframes;
      if (LANE == 0) {
        Frame& f = w.frames[fi];
        f.mark = w.ntrail;
        f.ustar = o.ustar;
        f.nD = o.nD;
        f.next = 0;
        f.kl = o.klc;
        f.csize = w.nC;
        f.zgone = 0;
      }
      for (u32 i = LANE; i < o.nD; i += 64) w.frames[fi].D[i] = L.D[i];
      wsync();
      w.nframes = fi + 1;
      from = w.ntrail;
      kl_seen = o.klc;
      for (u32 i = 0; i < o.nD; i++) remove_r(w, L.D[i]);
      move_c(w, o.ustar);
    } else {
      // backtrack to the next pending child
      bool cont = false;
      while (w.nframes > 0) {
        Frame* f = &w.frames[w.nframes - 1];
        undo_to(w, f->mark);
        from = f->mark;
        kl_seen = f->kl;
        u32 next = f->next + 1;
        if (LANE == 0) f->next = next;
        wsync();
        if (next <= f->nD) {
          u32 i = next - 1;
          for (u32 j = 0; j < i; j++) remove_r(w, f->D[j]);
          move_c(w, f->ustar);
          move_c(w, f->D[i]);
          cont = true;
          break;
        }
        if (f->zgone) {
          w.nframes--;
          continue;
        }
        Frame z = *f;
        w.nframes--;
        remove_r(w, z.ustar);
        for (u32 j = 0; j < z.nD; j++) remove_r(w, z.D[j]);
        cont = true;
        break;
      }
      if (!cont) {
        active = 0;
        if (pend) {
          if (LANE == 0) {
            atomicAdd((unsigned long long*)&Q->nodes, (unsigned long long)pend);
            atomicAdd((unsigned long long*)&A.glob->nodes, (unsigned long long)pend);
          }
          pend = 0;
          wsync();
        }
        finish_task(w, A, qslot);
        continue;
      }
    }
    // donate the shallowest pending Z while other waves are idle
    if (!A.deterministic && w.nframes > 0 && aload(&A.glob->idle) > 0) {
      u64 queued_tasks = aload64(A.pool.enq) - aload64(A.pool.deq);
      if (queued_tasks < (u64)A.donate_below) {
        u32 fi = INV;
        for (u32 i = 0; i < w.nframes; i++) {
          if (!w.frames[i].zgone) {
            fi = i;
            break;
          }
        }
        if (fi != INV) {
          const Frame& f = w.frames[fi];
          const u32 nRz = w.nR + (w.ntrail - f.mark) - 1 - f.nD;
          const u32 nXz = w.K - f.csize - nRz;
          const u32 need = 4 + f.csize + min(nRz, nXz);
          u64 pos = 0;
          u32 got = 0;
          if (need <= A.pool.slot_words && LANE == 0) got = pool_try_enqueue(A.pool, &pos);
          got = bcast(got);
          if (got) {
            u32 plo = __shfl((u32)pos, 0), phi = __shfl((u32)(pos >> 32), 0);
            pos = ((u64)phi << 32) | plo;
            u32* dst = A.pool.data + (pos % A.pool.nslots) * (u64)A.pool.slot_words;
            serialize_z(w, fi, dst, qslot, A.pool.slot_words, &L);
            if (LANE == 0) {
              w.frames[fi].zgone = 1;
              atomicAdd(&Q->pending, 1u);
            }
            __threadfence();
            if (LANE == 0) {
              astore64(&A.pool.seq[pos % A.pool.nslots], pos + 1);
              atomicAdd((unsigned long long*)&A.glob->donations, 1ull);
            }
            wsync();
          }
        }
      }
    }
  }
  if (idle_counted && LANE == 0) atomicSub(&A.glob->idle, 1u);
  if (active) {
    save_hdr(w, 1, qslot, from, kl_seen, pend);
  } else {
    save_hdr(w, 0, 0, 0, INV, 0);
  }
}

extern "C" hipError_t scs_launch_epoch(const EngineArgs* A, hipStream_t s) {
  scs_epoch<<<A->nwaves, 64, 0, s>>>(*A);
  return hipGetLastError();
}

// Pushes the root task (C = {q}, nothing removed) of each admitted slot.
__global__ void __launch_bounds__(64) scs_admit(EngineArgs A, const u32* slots_in, u32 n,
                                                u32* failed) {
  for (u32 i = 0; i < n; i++) {
    const u32 sl = slots_in[i];
    u64 pos = 0;
    u32 got = 0;
    if (LANE == 0) got = pool_try_enqueue(A.pool, &pos);
    got = bcast(got);
    if (!got) {
      if (LANE == 0) failed[i] = 1;
      continue;
    }
    if (LANE == 0) {
      u32* dst = A.pool.data + (pos % A.pool.nslots) * (u64)A.pool.slot_words;
      dst[0] = sl;
      dst[1] = 1;
      dst[2] = 0;
      dst[3] = 0;
      dst[4] = A.slots[sl].qloc;
      failed[i] = 0;
      __threadfence();
      astore64(&A.pool.seq[pos % A.pool.nslots], pos + 1);
    }
    wsync();
  }
}

extern "C" hipError_t scs_launch_admit(const EngineArgs* A, const u32* slots, u32 n, u32* failed,
                                       hipStream_t s) {
  scs_admit<<<1, 64, 0, s>>>(*A, slots, n, failed);
  return hipGetLastError();
}

// Initialises the pool sequence numbers (seq[i] = i).
__global__ void scs_pool_init(u64* seq, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) seq[i] = i;
}
extern "C" hipError_t scs_launch_pool_init(u64* seq, u32 n, hipStream_t s) {
  scs_pool_init<<<(n + 255) / 256, 256, 0, s>>>(seq, n);
  return hipGetLastError();
}
