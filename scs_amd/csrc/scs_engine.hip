// Branch-and-bound search engine for size-bounded community search, gfx950.
//
// Replaces ProcessTask / FindDoms / Expand / RemoveCompletedTasks
// (src/helpers.cc:371-889) and the host level loop (main.cu:261-442).
//
// Design: one wavefront (64 lanes, one workgroup) is one depth-first
// worker.  It keeps, for the query it is bound to, the full search state in
// its own HBM region: vertex status (R = candidate, C = chosen, X = out),
// dC = |N(v) & C|, dR = |N(v) & R|, an explicit list of R, an undo trail,
// and (in LDS) a histogram of dC over R.  A node costs O(changes x degree)
// instead of the reference's O(|T| x deg x |T|) rescans (findIndexKernel,
// src/helpers.cc:23-47).  Removals are applied in batches: every lane takes
// removed vertices and decrements their neighbours' dR with atomics, and an
// undo restores a whole batch the same way, so the large rule 1/2 closures
// deep in the tree are parallel instead of vertex by vertex.  Waves that
// run out of work take tasks from a bounded lock-free pool that busy waves
// fill by donating the unexplored "exclude u*" branch of their shallowest
// frame.  The kernel runs for a bounded time slice and returns; all state
// persists in HBM between slices, so the host never waits on a spinning
// kernel.
//
// Node semantics (identical to the CPU oracle, so a single wave expands
// exactly the oracle's nodes): min degree of C when l <= |C| <= h, rule 3,
// rules 1+2 to a fixpoint, U_d and U_nr, ustar by fixed-point connection
// score, the set dominated by ustar, then children A, B_0.., Z.
#include "scs_internal.h"
#include "scs_mlp.h"

#define LANE (threadIdx.x)
#define INV 0xffffffffu
// trail entries: v | TR_MOVE (R -> C), v (R -> X, part of a batch), and a
// batch marker TR_MARK | m after the m vertices of a batch; v < 2^30
#define TR_MOVE 0x80000000u
#define TR_MARK 0x40000000u
#define TR_RECOUNT 0x20000000u  // marker flag: batch applied by recounting
#define TR_VMASK 0x3fffffffu
#define TR_MMASK 0x1fffffffu

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
__device__ __forceinline__ u64 wsum64(u64 x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
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

// LDS of one wave (one workgroup).
struct Lds {
  u32 cnt, cnt2, nfront, ncand, nh, nk, pad0, pad1;
  u32 root[8];              // payload of a level root this wave starts itself
  u64 tph[SCS_PHASES];      // wall-clock ticks per phase (lane 0 adds)
  u32 D[SCS_FRAME_DOMS];
  u32 hist[SCS_HCAP + 4];   // dC histogram of R (live copy of the wave's state)
  union {                   // never live at the same time within a node
    u32 chist[SCS_HCAP + 4];  // scratch histogram of C degrees (U_nr)
    u16 h1s[16 * MLP_H];      // MLP hidden-layer tile (bf16)
  };
};

// ---------------------------------------------------------------------------
// Per-wave state view.
// Candidates the MLP scores per node (learned order).
#define LEARN_SHORTLIST 128u

// Phases of the search whose wall-clock time is accounted separately.
#define PH_REBUILD 0  // state rebuilt from a pooled task
#define PH_RULES 1    // rule 3 and rule 1/2 closure, kl offers
#define PH_BOUNDS 2   // U_d and U_nr
#define PH_SCORE 3    // connection scores, argmax, learned order
#define PH_DOMS 4     // dominated set of ustar
#define PH_MOVE 5     // applying children, backtracking (undo)
#define PH_DONATE 6   // serialising a donated task
#define PH_IDLE 7     // waiting for a task
#define PH_CSCAN 8    // rule 1/2 closure: finding the vertices to remove
#define PH_CREMOVE 9  // rule 1/2 closure: applying the removal batches
#define PH_ARGMAX 10  // argmax of the connection score over the frontier
#define PH_CLEAR 11   // clearing the frontier's scores
#define PH_PAIR 12    // pair rule (common-neighbour counts of C vertices)

__device__ __forceinline__ void ph_switch(Lds* L, u64& tp, u32& ph, u32 next) {
  const u64 t = wall_clock64();
  if ((threadIdx.x & 63) == 0) L->tph[ph] += t - tp;
  tp = t;
  ph = next;
}

// ---------------------------------------------------------------------------
struct W {
  VState* vs;  // per vertex: dR, dC, status, queued flag (one 8-byte record)
  u8* nmark;
  u32* pos;
  u32* rlist;
  u32* work;
  u32* work2;
  u32* front;
  u32* aux;
  u32* trail;
  u32* sav;    // saved (vertex, dR) pairs of recount batches, as a stack
  Frame* frames;
  u32* clist;
  u64* gb;     // R/C bitmaps of the last rebuild (WaveLayout::gbits)
  u32* ghist;  // HBM copy of the histogram between epochs
  WaveHdr* hdr;
  Lds* L;
  // query
  QuerySlot* Q;
  const u32* off;
  const u32* adj;
  const u32* dist;
  const u32* core;
  u32 K, l, h, flags, limit_doms, dmax, ku, tag;
  // learned ordering and search traces
  const MlpDev* model;
  u32* trace;
  u32* trace_cnt;
  u32 trace_cap, trace_stride;
  EngineGlobals* gl;
  // LDS bitmaps over the bound query's local ids (when they fit): xb = the
  // vertices with status X, tb = scratch set (C during scoring, N[ustar]
  // during the domination test), all-zero between uses
  u32* xb;
  u32* tb;
  u32 bm;         // 1 = the bitmaps are in use for the bound query
  u32 lds_words;  // dynamic LDS words of this wave
  // wave-uniform counters
  u32 nC, nR, ntrail, nframes, nsav;
  u64 edges;
};

__device__ void bind(W& w, const EngineArgs& A, u32 wave, Lds* L) {
  u8* base = A.wave_mem + (u64)wave * A.lay.stride;
  w.vs = (VState*)(base + A.lay.vs);
  w.nmark = base + A.lay.nmark;
  w.pos = (u32*)(base + A.lay.pos);
  w.rlist = (u32*)(base + A.lay.rlist);
  w.work = (u32*)(base + A.lay.work);
  w.work2 = (u32*)(base + A.lay.work2);
  w.front = (u32*)(base + A.lay.front);
  w.aux = (u32*)(base + A.lay.aux);
  w.trail = (u32*)(base + A.lay.trail);
  w.sav = (u32*)(base + A.lay.sav);
  w.frames = (Frame*)(base + A.lay.frames);
  w.clist = (u32*)(base + A.lay.clist);
  w.gb = (u64*)(base + A.lay.gbits);
  w.ghist = (u32*)(base + A.lay.hist);
  w.hdr = (WaveHdr*)(base + A.lay.hdr);
  w.L = L;
  w.edges = 0;
  w.nsav = 0;
  w.model = A.model;
  w.gl = A.glob;
  extern __shared__ u32 lds_dyn[];
  w.lds_words = A.lds_words;
  w.xb = lds_dyn;
  w.tb = lds_dyn;
  w.bm = 0;
  w.trace = A.trace;
  w.trace_cnt = A.trace_cnt;
  w.trace_cap = A.trace_cap;
  w.trace_stride = A.trace_stride ? A.trace_stride : 1;
}

__device__ void bind_query(W& w, QuerySlot* Q) {
  w.Q = Q;
  w.off = Q->off;
  w.adj = Q->adj;
  w.dist = Q->dist;
  w.core = Q->core;
  w.tag = Q->tag;
  w.K = Q->K;
  w.l = Q->l;
  w.h = Q->h;
  w.flags = Q->flags;
  w.limit_doms = Q->limit_doms;
  w.dmax = Q->dmax;
  w.ku = Q->ku;
  const u32 nw = (w.K + 31) / 32;
  w.bm = 2 * nw <= w.lds_words ? 1u : 0u;
  w.tb = w.xb + nw;
}

__device__ bool xbit(const W& w, u32 v) { return (w.xb[v >> 5] >> (v & 31)) & 1u; }
__device__ bool tbit(const W& w, u32 v) { return (w.tb[v >> 5] >> (v & 31)) & 1u; }
__device__ void tset(W& w, u32 v) { atomicOr(&w.tb[v >> 5], 1u << (v & 31)); }
__device__ void tclr(W& w, u32 v) { atomicAnd(&w.tb[v >> 5], ~(1u << (v & 31))); }

// Builds xb from the status array and clears tb (bitmap mode only).
__device__ void build_bitmaps(W& w) {
  if (!w.bm) return;
  const u32 K = w.K;
  const u32 nw = (K + 31) / 32;
  for (u32 v0 = 0; v0 < K; v0 += 64) {
    const u32 v = v0 + LANE;
    const u64 m = __ballot(v < K && w.vs[v].st == ST_X);
    if (LANE == 0) w.xb[v0 >> 5] = (u32)m;
    if (LANE == 1 && (v0 >> 5) + 1 < nw) w.xb[(v0 >> 5) + 1] = (u32)(m >> 32);
  }
  for (u32 i = LANE; i < nw; i += 64) w.tb[i] = 0;
  wsync();
}

// ---------------------------------------------------------------------------
// Load-balanced adjacency scans over a list of vertices.  vertex(i) gives
// the i-th vertex (i < m).  Vertices of degree <= SEQ_DEG are taken one per
// lane; larger ones (power-law hubs) by the whole wave, 64 entries at a
// time, so no lane walks a long list alone.  edge(x) returns a
// contribution (AccT) of adjacency entry x; done(i, v, total) runs once
// per vertex with the sum of its contributions, on one lane.
// ---------------------------------------------------------------------------
#define SEQ_DEG 32u

template <typename AccT>
__device__ __forceinline__ AccT wsum_t(AccT x);
template <>
__device__ __forceinline__ u32 wsum_t<u32>(u32 x) { return wsum(x); }
template <>
__device__ __forceinline__ u64 wsum_t<u64>(u64 x) { return wsum64(x); }

template <typename AccT, typename VertexF, typename EdgeF, typename DoneF>
__device__ void scan_adj(W& w, u32 m, VertexF vertex, EdgeF edge, DoneF done) {
  u32 ed = 0;
  for (u32 c0 = 0; c0 < m; c0 += 64) {
    const u32 i = c0 + LANE;
    u32 v = 0, b = 0, e = 0;
    if (i < m) {
      v = vertex(i);
      b = w.off[v];
      e = w.off[v + 1];
    }
    const bool big = i < m && (e - b) > SEQ_DEG;
    if (i < m && !big) {
      AccT acc = 0;
      u32 j = b;
      // four entries per step: their gathers are in flight together
      for (; j + 4 <= e; j += 4) {
        const u32 x0 = w.adj[j], x1 = w.adj[j + 1], x2 = w.adj[j + 2], x3 = w.adj[j + 3];
        acc += edge(x0);
        acc += edge(x1);
        acc += edge(x2);
        acc += edge(x3);
      }
      for (; j < e; j++) acc += edge(w.adj[j]);
      done(i, v, acc);
      ed += e - b;
    }
    u64 bm = __ballot(big);
    while (bm) {
      const int src = __ffsll((unsigned long long)bm) - 1;
      bm &= bm - 1;
      const u32 bv = __shfl(v, src), bb = __shfl(b, src), be = __shfl(e, src);
      AccT acc = 0;
      for (u32 j = bb + LANE; j < be; j += 64) acc += edge(w.adj[j]);
      acc = wsum_t<AccT>(acc);
      if (LANE == 0) {
        done(c0 + (u32)src, bv, acc);
        ed += be - bb;
      }
    }
  }
  w.edges += wsum(ed);
}

// Adds delta to hist[d] for every lane with `on` (call on all lanes).  Most
// updates hit dC = 0 or 1, so those two are summed over the wave first: one
// LDS atomic instead of up to 64 colliding ones.
__device__ __forceinline__ void hist_add(u32* hist, bool on, u32 d, int delta) {
  const u64 m0 = __ballot(on && d == 0), m1 = __ballot(on && d == 1);
  if (LANE == 0) {
    if (m0) atomicAdd(&hist[0], (u32)(delta * (int)__popcll(m0)));
    if (m1) atomicAdd(&hist[1], (u32)(delta * (int)__popcll(m1)));
  }
  if (on && d > 1) atomicAdd(&hist[d], (u32)delta);
}

// ---------------------------------------------------------------------------
// State updates.  Counters of X vertices are frozen: a move or a removal
// updates only the neighbours that are live at that moment, and an undo,
// which runs in exact reverse order, finds the same neighbours live again.
// A removal batch decrements first (all batch members still R) and then
// marks the batch X; its undo marks the batch R first and then increments,
// so batch members see each other both ways.
// ---------------------------------------------------------------------------
// Adds delta (+1 or -1) to dC and -delta to dR of every live neighbour of
// v, keeping the dC histogram of R.  U chunks of 64 entries per step,
// software-pipelined: a step issues its state gathers and the next step's
// adjacency loads together, then writes {dR, dC, status, queued} as one
// 64-bit word (v's neighbours are distinct, so no lane shares a record).
template <int U>
__device__ __forceinline__ void shift_dc(W& w, u32 v, int delta) {
  const u32 b = w.off[v], e = w.off[v + 1];
  u32* hist = w.L->hist;
  if (e > b) {
    u32 x[U];
#pragma unroll
    for (int k = 0; k < U; k++) x[k] = w.adj[min(b + LANE + 64 * k, e - 1)];
    for (u32 j = b; j < e; j += 64 * U) {
      u64 wd[U];  // {dR, dC, status, queued} of each neighbour
      u32 xn[U];
#pragma unroll
      for (int k = 0; k < U; k++) wd[k] = *(const u64*)&w.vs[x[k]];
#pragma unroll
      for (int k = 0; k < U; k++) xn[k] = w.adj[min(j + 64 * U + LANE + 64 * k, e - 1)];
#pragma unroll
      for (int k = 0; k < U; k++) {
        const u32 st = (u32)(wd[k] >> 48) & 0xffu;
        const bool live = j + LANE + 64 * k < e && st != ST_X;
        const u32 dc = live ? (u32)(wd[k] >> 32) & 0xffffu : (delta > 0 ? 0u : 1u);
        const bool r = live && st == ST_R;
        hist_add(hist, r, dc, -1);
        hist_add(hist, r, dc + delta, 1);
        if (live)
          *(u64*)&w.vs[x[k]] = (wd[k] & 0xFFFF000000000000ull) | ((u64)(u16)(dc + delta) << 32) |
                               (u32)((u32)wd[k] - delta);
      }
#pragma unroll
      for (int k = 0; k < U; k++) x[k] = xn[k];
    }
  }
  w.edges += e - b;
  wsync();
}

// U: adjacency chunks in flight per step.
template <int U = 4>
__device__ void move_c(W& w, u32 v) {
  shift_dc<U>(w, v, 1);
  u32* hist = w.L->hist;
  if (LANE == 0) {
    w.vs[v].st = ST_C;
    hist[w.vs[v].dC] -= 1;
    u32 p = w.pos[v];
    u32 last = w.rlist[w.nR - 1];
    w.rlist[p] = last;
    w.pos[last] = p;
    w.clist[w.nC] = v;
    w.trail[w.ntrail] = v | TR_MOVE;
  }
  w.nR--;
  w.nC++;
  w.ntrail++;
  wsync();
}

template <int U = 4>
__device__ void undo_move(W& w, u32 v) {
  u32* hist = w.L->hist;
  if (LANE == 0) {
    w.vs[v].st = ST_R;
    hist[w.vs[v].dC] += 1;
    w.rlist[w.nR] = v;
    w.pos[v] = w.nR;
  }
  w.nR++;
  w.nC--;
  wsync();
  shift_dc<U>(w, v, -1);
}

// Closure parameters for crossing detection during a batch (rule 1).
struct Clo {
  u32 on;   // push vertices whose dC + dR falls to kl
  u32 kl;
  u32* next;
  u32* cnt;  // LDS counter of next
};

__device__ __forceinline__ bool dec_live(const W& w, u32 x) {
  return w.bm ? !xbit(w, x) : w.vs[x].st != ST_X;
}
// old: the record's first word before the decrement of a live neighbour x
__device__ __forceinline__ void dec_check(W& w, u32 x, u64 old, const Clo& c) {
  const u32 s = (u32)(old >> 48) & 0xffu;
  if (c.on && s == ST_R) {
    const u32 dc = (u32)(old >> 32) & 0xffffu;
    if (dc <= c.kl && (u32)old + dc == c.kl + 1) {
      u32* wp = (u32*)((uintptr_t)(&w.vs[x].queued) & ~(uintptr_t)3);
      u32 sh = ((uintptr_t)(&w.vs[x].queued) & 3) * 8;
      u32 o2 = atomicOr(wp, 1u << sh);
      if (!((o2 >> sh) & 0xffu)) c.next[atomicAdd(c.cnt, 1u)] = x;
    }
  }
}
// Live check first (an X vertex's counters are frozen and may be 0), then
// one 64-bit atomic on the record's first word {dR, dC, status, queued}: it
// decrements dR (>= 1: a live neighbour counts the removed R vertex) and
// returns dC and the status in the same round trip.
__device__ __forceinline__ void dec_neighbour(W& w, u32 x, const Clo& c) {
  if (!dec_live(w, x)) return;
  const u64 old = atomicSub((unsigned long long*)&w.vs[x], 1ull);
  dec_check(w, x, old, c);
}

// Decrements dR of every live neighbour of list[0..m) (load-balanced like
// scan_adj).  Four entries per lane per step, software-pipelined: the next
// step's adjacency loads are issued before this step's decrements, so one
// memory round trip covers both (the memory counter is in order).  Loads
// past a list's end read its last entry (or entry 0 of an empty graph) and
// are ignored.
__device__ __forceinline__ void dec4(W& w, const u32 x[4], const bool in[4], const Clo& c) {
  bool lv[4];
  u64 od[4];
#pragma unroll
  for (int k = 0; k < 4; k++) lv[k] = in[k] && dec_live(w, x[k]);
#pragma unroll
  for (int k = 0; k < 4; k++)
    od[k] = lv[k] ? atomicSub((unsigned long long*)&w.vs[x[k]], 1ull) : 0ull;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (lv[k]) dec_check(w, x[k], od[k], c);
}
__device__ __forceinline__ void dec_neighbours(W& w, const u32* list, u32 m, const Clo& c) {
  u32 ed = 0;
  for (u32 c0 = 0; c0 < m; c0 += 64) {
    const u32 i = c0 + LANE;
    u32 v = 0, b = 0, e = 0;
    if (i < m) {
      v = list[i];
      b = w.off[v];
      e = w.off[v + 1];
    }
    const bool big = i < m && (e - b) > SEQ_DEG;
    if (i < m && !big && e > b) {
      u32 x[4];
#pragma unroll
      for (int k = 0; k < 4; k++) x[k] = w.adj[min(b + k, e - 1)];
      for (u32 j = b; j < e; j += 4) {
        u32 xn[4];
#pragma unroll
        for (int k = 0; k < 4; k++) xn[k] = w.adj[min(j + 4 + k, e - 1)];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; k++) in[k] = j + k < e;
        dec4(w, x, in, c);
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = xn[k];
      }
      ed += e - b;
    }
    u64 bm = __ballot(big);
    while (bm) {
      const int src = __ffsll((unsigned long long)bm) - 1;
      bm &= bm - 1;
      const u32 bb = __shfl(b, src), be = __shfl(e, src);
      // the whole wave, 256 entries per step
      u32 x[4];
#pragma unroll
      for (int k = 0; k < 4; k++) x[k] = w.adj[min(bb + LANE + 64 * k, be - 1)];
      for (u32 j = bb; j < be; j += 256) {
        u32 xn[4];
#pragma unroll
        for (int k = 0; k < 4; k++) xn[k] = w.adj[min(j + 256 + LANE + 64 * k, be - 1)];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; k++) in[k] = j + LANE + 64 * k < be;
        dec4(w, x, in, c);
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = xn[k];
      }
      if (LANE == 0) ed += be - bb;
    }
  }
  w.edges += wsum(ed);
}

// Removes the m vertices of list (all currently R) as one trail batch.
//
// Two equivalent ways to apply it.  Exact mode decrements dR of every live
// neighbour (cost: the batch's total degree).  Recount mode, used when the
// batch is much larger than what stays live, marks the batch X without
// touching its neighbours and then recounts dR of the surviving live
// vertices (cost: their total degree), pushing the old values on the save
// stack for the undo.  Both leave identical dC/dR on every live vertex.
__device__ __forceinline__ void remove_batch(W& w, const u32* list, u32 m, const Clo& c) {
  if (m == 0) return;
  u32* hist = w.L->hist;
  const u32 live_after = w.nC + (w.nR - m);
  const bool recount = m >= 256 && m > 2 * live_after &&
                       (u64)w.nsav + 2ull * live_after + 1 <= 2ull * w.K + 64;
  // phase 1 (exact mode): decrement the live neighbours (batch included)
  if (!recount) {
    if (m >= 8) {
      dec_neighbours(w, list, m, c);
    } else {
      for (u32 i = 0; i < m; i++) {
        u32 v = list[i];
        const u32 b = w.off[v], e = w.off[v + 1];
        for (u32 j = b + LANE; j < e; j += 64) dec_neighbour(w, w.adj[j], c);
        w.edges += e - b;
      }
    }
  }
  if (LANE == 0) {
    w.L->nh = 0;
    w.L->nk = 0;
  }
  wsync();
  // phase 2: mark X, histogram, trail, queued flags
  for (u32 i0 = 0; i0 < m; i0 += 64) {
    const u32 i = i0 + LANE;
    const bool on = i < m;
    u32 d = 0;
    if (on) {
      const u32 v = list[i];
      d = w.vs[v].dC;
      w.vs[v].st = ST_X;
      if (w.bm) atomicOr(&w.xb[v >> 5], 1u << (v & 31));
      w.vs[v].queued = 0;
      w.trail[w.ntrail + i] = v;
    }
    hist_add(hist, on, d, -1);
  }
  wsync();
  // phase 3: fill the holes the batch leaves in R's list with the surviving
  // entries of the list's last m positions
  const u32 tail = w.nR - m;
  for (u32 i = LANE; i < m; i += 64) {
    u32 p = w.pos[list[i]];
    if (p < tail) w.aux[atomicAdd(&w.L->nh, 1u)] = p;
  }
  for (u32 t = tail + LANE; t < w.nR; t += 64) {
    u32 x = w.rlist[t];
    if (w.vs[x].st == ST_R) w.front[atomicAdd(&w.L->nk, 1u)] = x;
  }
  wsync();
  const u32 nh = bcast(w.L->nh);
  for (u32 r = LANE; r < nh; r += 64) {
    u32 p = w.aux[r], x = w.front[r];
    w.rlist[p] = x;
    w.pos[x] = p;
  }
  if (LANE == 0) w.trail[w.ntrail + m] = TR_MARK | (recount ? TR_RECOUNT : 0u) | m;
  w.ntrail += m + 1;
  w.nR = tail;
  wsync();
  if (recount) {
    // phase 4: save and recount dR of the live vertices (R list, then C)
    u32* S = w.sav + w.nsav;
    const u32 nR = w.nR;
    const bool bm = w.bm;
    if (bm) {
      // tb = C: "x is R" is two LDS bit tests
      for (u32 i = LANE; i < w.nC; i += 64) tset(w, w.clist[i]);
      wsync();
    }
    scan_adj<u32>(
        w, live_after, [&](u32 i) { return i < nR ? w.rlist[i] : w.clist[i - nR]; },
        [&](u32 x) { return (u32)(bm ? !(xbit(w, x) || tbit(w, x)) : w.vs[x].st == ST_R); },
        [&](u32 i, u32 v, u32 r) {
          S[2 * i] = v;
          S[2 * i + 1] = w.vs[v].dR;
          w.vs[v].dR = r;
          if (c.on && i < nR && w.vs[v].dC + r <= c.kl) {
            u32* wp = (u32*)((uintptr_t)(&w.vs[v].queued) & ~(uintptr_t)3);
            u32 sh = ((uintptr_t)(&w.vs[v].queued) & 3) * 8;
            u32 o2 = atomicOr(wp, 1u << sh);
            if (!((o2 >> sh) & 0xffu)) c.next[atomicAdd(c.cnt, 1u)] = v;
          }
        });
    if (bm) {
      wsync();
      for (u32 i = LANE; i < w.nC; i += 64) tclr(w, w.clist[i]);
    }
    if (LANE == 0) S[2 * live_after] = live_after;
    w.nsav += 2 * live_after + 1;
    wsync();
  }
}

__device__ void undo_batch(W& w, u32 start, u32 m, bool recount) {
  u32* hist = w.L->hist;
  const u32* E = w.trail + start;
  if (recount) {
    // restore the recounted dR values first (same live set as at the batch)
    const u32 nl = w.sav[w.nsav - 1];
    const u32* S = w.sav + w.nsav - 1 - 2 * nl;
    for (u32 i = LANE; i < nl; i += 64) w.vs[S[2 * i]].dR = S[2 * i + 1];
    w.nsav -= 2 * nl + 1;
    wsync();
  }
  for (u32 i0 = 0; i0 < m; i0 += 64) {
    const u32 i = i0 + LANE;
    const bool on = i < m;
    u32 d = 0;
    if (on) {
      const u32 v = E[i];
      w.vs[v].st = ST_R;
      if (w.bm) atomicAnd(&w.xb[v >> 5], ~(1u << (v & 31)));
      d = w.vs[v].dC;
      w.rlist[w.nR + i] = v;
      w.pos[v] = w.nR + i;
    }
    hist_add(hist, on, d, 1);
  }
  w.nR += m;
  wsync();
  if (recount) return;
  if (m >= 8) {
    scan_adj<u32>(
        w, m, [&](u32 i) { return E[i]; },
        [&](u32 x) {
          if (w.bm ? !xbit(w, x) : w.vs[x].st != ST_X) atomicAdd(&w.vs[x].dR, 1u);
          return 0u;
        },
        [&](u32, u32, u32) {});
  } else {
    for (u32 i = 0; i < m; i++) {
      u32 v = E[i];
      const u32 b = w.off[v], e = w.off[v + 1];
      for (u32 j = b + LANE; j < e; j += 64) {
        u32 x = w.adj[j];
        if (w.bm ? !xbit(w, x) : w.vs[x].st != ST_X) atomicAdd(&w.vs[x].dR, 1u);
      }
      w.edges += e - b;
    }
  }
  wsync();
}

template <int U = 4>
__device__ void undo_to(W& w, u32 mark) {
  while (w.ntrail > mark) {
    const u32 e = w.trail[w.ntrail - 1];
    if (e & TR_MOVE) {
      w.ntrail--;
      undo_move<U>(w, e & TR_VMASK);
    } else {
      const u32 m = e & TR_MMASK;  // batch marker
      const u32 start = w.ntrail - 1 - m;
      w.ntrail = start;
      undo_batch(w, start, m, (e & TR_RECOUNT) != 0);
    }
  }
}

// ---------------------------------------------------------------------------
// Witness and kl.
// ---------------------------------------------------------------------------
__device__ void offer(W& w, u32 mn) {
  // The witness is stored under the slot's lock before kl is raised, so a
  // reader that sees kl >= x always finds a stored witness of min degree >= x.
  QuerySlot* Q = w.Q;
  u32 want = 0;
  if (LANE == 0) {
    const u32 cur = aload(&Q->kl);
    want = (mn > cur || (mn == cur && aload(&Q->wsize) == 0)) ? 1u : 0u;
  }
  want = bcast(want);
  if (!want) return;
  u32 got = 0;
  if (LANE == 0) {
    // The holder's critical section is at most h stores; the bound only
    // guards against a lost lock word.
    for (u32 spin = 0; spin < (1u << 26); spin++) {
      if (atomicCAS(&Q->wlock, 0u, 1u) == 0u) {
        got = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (!got) {
      atomicExch(&Q->wlost, 1u);
      atomicAdd(&w.gl->lock_fail, 1u);
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
      __threadfence();
      atomicMax(&Q->kl, mn);
    }
  }
  __threadfence();
  if (LANE == 0) atomicExch(&Q->wlock, 0u);
  wsync();
}

// The bound the rules work against: the best witness's min degree, or the
// current level's floor when that is higher (no task of the query runs while
// the floor changes).
__device__ u32 cur_kl(const W& w) { return max(aload(&w.Q->kl), aload(&w.Q->floor)); }

// ---------------------------------------------------------------------------
// Node processing.
// ---------------------------------------------------------------------------
__device__ bool cond12(const W& w, u32 v, u32 Cs, u32 ubD, u32 kl) {
  u32 dc = w.vs[v].dC, dr = w.vs[v].dR;
  if ((w.flags & QF_RED1) && min(dr + dc, dc + w.h - Cs - 1) <= kl) return true;
  if ((w.flags & QF_RED2) && ubD < w.dist[v]) return true;
  return false;
}

__device__ void try_push(W& w, u32 x, u32 Cs, u32 ubD, u32 kl, u32* list,
                                         u32* cnt) {
  if (w.vs[x].st == ST_R && !w.vs[x].queued && cond12(w, x, Cs, ubD, kl)) {
    u32* wp = (u32*)((uintptr_t)(&w.vs[x].queued) & ~(uintptr_t)3);
    u32 sh = ((uintptr_t)(&w.vs[x].queued) & 3) * 8;
    u32 old = atomicOr(wp, 1u << sh);
    if (!((old >> sh) & 0xffu)) list[atomicAdd(cnt, 1u)] = x;
  }
}

// Rule 1/2 closure.  At a parent node no R vertex met the condition after
// its closure; in a child only these can: neighbours of vertices removed
// since `from` (their dR dropped), vertices with dC <= kl + |C| + 1 - h
// (|C| grew), or everything when kl changed.  Each round removes all
// queued vertices as one batch and queues the neighbours it pushes below
// the rule-1 threshold.  Returns true if anything was removed.
__device__ __forceinline__ bool r12(W& w, u32 from, bool full, u32 kl, u64& tp, u32& ph) {
  Lds* L = w.L;
  const u32 Cs = w.nC;
  const u32 ubD = ubd_node(kl, w.h);
  const u32 ph0 = ph;
  ph_switch(L, tp, ph, PH_CSCAN);
  if (LANE == 0) L->cnt = 0;
  wsync();
  if (!full && (w.flags & QF_RED1)) {
    long T = (long)kl + (long)Cs + 1 - (long)w.h;
    if (T >= 0) {
      u32 c = 0;
      for (u32 d = LANE; d <= (u32)T && d < w.h + 3; d += 64) c += L->hist[d];
      if (wsum(c)) full = true;
    }
  }
  if (full) {
    if (LANE == 0) {
      atomicAdd((unsigned long long*)&w.gl->full_scans, 1ull);
      atomicAdd((unsigned long long*)&w.gl->full_rows, (unsigned long long)w.nR);
    }
    // eight R vertices per lane per step: the list loads, then their
    // state words and distances, are in flight together (indices past the
    // list's end read its last entry and are ignored)
    const u32 nR = w.nR;
    const bool r1 = w.flags & QF_RED1, r2 = w.flags & QF_RED2;
    for (u32 i0 = 0; i0 < nR; i0 += 512) {
      u32 x[8];
#pragma unroll
      for (int k = 0; k < 8; k++) x[k] = w.rlist[min(i0 + LANE + 64u * k, nR - 1)];
      u64 wd[8];
      u32 ds[8];
#pragma unroll
      for (int k = 0; k < 8; k++) wd[k] = *(const u64*)&w.vs[x[k]];
#pragma unroll
      for (int k = 0; k < 8; k++) ds[k] = r2 ? w.dist[x[k]] : 0u;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 dr = (u32)wd[k], dc = (u32)(wd[k] >> 32) & 0xffffu;
        const u32 st = (u32)(wd[k] >> 48) & 0xffu, qd = (u32)(wd[k] >> 56);
        const bool hit = (r1 && min(dr + dc, dc + w.h - Cs - 1) <= kl) || (r2 && ubD < ds[k]);
        if (i0 + LANE + 64u * k < nR && st == ST_R && !qd && hit) {
          u32* wp = (u32*)((uintptr_t)(&w.vs[x[k]].queued) & ~(uintptr_t)3);
          const u32 sh = ((uintptr_t)(&w.vs[x[k]].queued) & 3) * 8;
          const u32 old = atomicOr(wp, 1u << sh);
          if (!((old >> sh) & 0xffu)) w.work[atomicAdd(&L->cnt, 1u)] = x[k];
        }
      }
    }
  } else {
    // vertices removed since `from` (trail entries without the move or
    // batch-marker flags), compacted into work2, then their neighbours
    // checked with the load-balanced scan
    u32 nrm = 0;
    for (u32 t0 = from; t0 < w.ntrail; t0 += 64) {
      const u32 t = t0 + LANE;
      const u32 e = t < w.ntrail ? w.trail[t] : TR_MARK;
      const bool plain = !(e & (TR_MOVE | TR_MARK));
      const u64 bm = __ballot(plain);
      if (plain) w.work2[nrm + __popcll(bm & lanes_below())] = e;
      nrm += __popcll(bm);
    }
    wsync();
    if (nrm) {
      const u32* rm = w.work2;
      scan_adj<u32>(
          w, nrm, [&](u32 i) { return rm[i]; },
          [&](u32 x) {
            try_push(w, x, Cs, ubD, kl, w.work, &L->cnt);
            return 0u;
          },
          [&](u32, u32, u32) {});
    }
  }
  wsync();
  u32 cnt = bcast(L->cnt);
  const bool removed = cnt > 0;
  ph_switch(L, tp, ph, PH_CREMOVE);
  u32* cur = w.work;
  u32* nxt = w.work2;
  while (cnt) {
    if (LANE == 0) L->cnt2 = 0;
    wsync();
    Clo c;
    c.on = (w.flags & QF_RED1) ? 1u : 0u;
    c.kl = kl;
    c.next = nxt;
    c.cnt = &L->cnt2;
    remove_batch(w, cur, cnt, c);
    cnt = bcast(L->cnt2);
    u32* t = cur;
    cur = nxt;
    nxt = t;
  }
  ph_switch(L, tp, ph, ph0);
  return removed;
}

__device__ bool tight_exists(const W& w, u32 kl) {
  u32 f = 0;
  for (u32 i = LANE; i < w.nC; i += 64) {
    u32 u = w.clist[i];
    if (w.vs[u].dC + w.vs[u].dR == kl + 1 && w.vs[u].dR > 0) f = 1;
  }
  return __ballot(f) != 0;
}

// Pair rule.  Two vertices u, v of an H with min degree >= k = kl+1 and
// |H| <= h each have >= k - [u~v] neighbours in H minus {u, v}, a set of at
// most h-2 vertices, so they share >= 2k-h (adjacent) or 2k-h+2 (not
// adjacent) neighbours in H, all live (C or R).  For each C vertex u in C
// order: the live common neighbours of u with every live vertex are counted
// by a two-hop scatter (low word of acc, which is zero between uses; first
// touches listed in front, N(u) marked in nmark); the node is pruned when
// another C vertex falls short, else every R vertex that falls short is
// removed as one batch before the next u.  Returns -1 (prune), 1 (removed
// something) or 0.  A C vertex whose live neighbours' lists hold more than
// PAIR_MAX_SCAN entries, or more than what is left of the node's
// PAIR_NODE_SCAN entries (*left), is skipped: a node whose C holds hubs of a
// large kept set stays cheap (a skipped check only weakens the pruning).
#define PAIR_MAX_SCAN (1ull << 20)
#define PAIR_NODE_SCAN (1ull << 24)
__device__ __forceinline__ int pair_step(W& w, u32 kl, u64& left) {
  Lds* L = w.L;
  const long tadj = 2L * (long)(kl + 1) - (long)w.h;
  if (tadj + 2 <= 0) return 0;
  const Clo none{0, 0, nullptr, nullptr};
  int any = 0;
  for (u32 ci = 0; ci < w.nC; ci++) {
    const u32 u = w.clist[ci];
    const u32 ub = w.off[u], ue = w.off[u + 1];
    if (LANE == 0) {
      L->nfront = 0;
      L->cnt = 0;
    }
    u32 nl = 0;
    u64 scan = 0;
    for (u32 j0 = ub; j0 < ue; j0 += 64) {
      const u32 j = j0 + LANE;
      u32 x = 0;
      bool live = false;
      if (j < ue) {
        x = w.adj[j];
        live = dec_live(w, x);
      }
      const u64 bmk = __ballot(live);
      if (live) {
        w.work2[nl + __popcll(bmk & lanes_below())] = x;
        scan += w.off[x + 1] - w.off[x];
      }
      nl += __popcll(bmk);
    }
    w.edges += ue - ub;
    scan = wsum64(scan);
    if (scan > PAIR_MAX_SCAN || scan > left) continue;
    left -= scan;
    for (u32 j = ub + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 1;
    wsync();
    const u32* lv = w.work2;
    scan_adj<u32>(
        w, nl, [&](u32 i) { return lv[i]; },
        [&](u32 x) {
          if (dec_live(w, x)) {
            const u32 old = atomicAdd((u32*)&w.vs[x].acc, 1u);
            if (old == 0) w.front[atomicAdd(&L->nfront, 1u)] = x;
          }
          return 0u;
        },
        [&](u32, u32, u32) {});
    // the barrier's fence waits for the count atomics
    wsync();
    auto short_of = [&](u32 v) -> bool {
      const long c = (long)(u32)w.vs[v].acc;
      return c < (w.nmark[v] ? tadj : tadj + 2);
    };
    u32 pr = 0;
    for (u32 i = LANE; i < w.nC; i += 64) {
      const u32 c = w.clist[i];
      if (c != u && short_of(c)) pr = 1;
    }
    const bool prune = __ballot(pr) != 0;
    if (!prune) {
      const u32 nR = w.nR;
      for (u32 i = LANE; i < nR; i += 64) {
        const u32 v = w.rlist[i];
        if (short_of(v)) w.work[atomicAdd(&L->cnt, 1u)] = v;
      }
    }
    wsync();
    const u32 nt = bcast(L->nfront);
    for (u32 i = LANE; i < nt; i += 64) w.vs[w.front[i]].acc = 0;
    for (u32 j = ub + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 0;
    wsync();
    if (prune) return -1;
    const u32 m = bcast(L->cnt);
    if (m) {
      remove_batch(w, w.work, m, none);
      any = 1;
    }
  }
  return any;
}

// U_nr water-filling on one lane: C degrees (histogram) get +1 on the
// smallest entries for each of the k largest R->C degrees (src/helpers.cc:
// 1066-1101 with a correct sort).  Values are capped at h+1: a minimum of
// h+1 or more can never be <= kl, so the prune decision is unchanged.
__device__ u32 unr_bound(W& w, u32 k) {
  u32* c = w.L->chist;
  const u32* hist = w.L->hist;
  const u32 h = w.h;
  const u32 top = h + 1;
  for (u32 i = LANE; i <= top; i += 64) c[i] = 0;
  wsync();
  for (u32 i = LANE; i < w.nC; i += 64) atomicAdd(&c[min((u32)w.vs[w.clist[i]].dC, top)], 1u);
  wsync();
  u32 res = INV;
  if (LANE == 0) {
    u32 lo = 0;
    while (lo < top && c[lo] == 0) lo++;
    const u32 m = w.nC;
    int d = (int)top;
    u32 left = hist[d];
    for (u32 i = 0; i < k; i++) {
      while (d > 0 && left == 0) left = hist[--d];
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

__device__ __forceinline__ MlpFeat cand_features(const W& w, u32 x, u32 Cs, u32 kl) {
  MlpFeat f;
  f.dC = w.vs[x].dC;
  f.dR = w.vs[x].dR;
  f.dist = w.dist[x];
  f.core = w.core[x];
  f.slack = (int)(f.dC + min(f.dR, w.h - Cs - 1)) - (int)kl;
  f.cscore = (float)((double)(w.vs[x].acc + rterm(f.dR, w.dmax)) * (1.0 / 4294967296.0));
  return f;
}

// Highest MLP score over the frontier, ties to the lowest id (INV if no
// finite score).
__device__ u32 learned_ustar(W& w, const u32* cand, u32 nf, u32 Cs, u32 kl) {
  float bestf = -INFINITY;
  u32 bestx = INV;
#pragma unroll 1
  for (u32 t0 = 0; t0 < nf; t0 += 16) {
    auto feat = [&](u32 row) {
      const u32 i = t0 + row;
      if (i < nf) return cand_features(w, cand[i], Cs, kl);
      return MlpFeat{0, 0, 0, 0, 0, 0.f};
    };
    const float sc = mlp_tile(w.model, feat, w.L->h1s);
    const u32 i = t0 + (LANE & 15u);
    if (LANE < 16 && i < nf) {
      const u32 x = cand[i];
      if (sc > bestf || (sc == bestf && x < bestx)) {
        bestf = sc;
        bestx = x;
      }
    }
  }
  float gb = bestf;
  for (int o = 32; o > 0; o >>= 1) gb = fmaxf(gb, __shfl_xor(gb, o));
  return wmin((bestx != INV && bestf == gb) ? bestx : INV);
}

// Records up to 64 frontier candidates of this node as trace rows.
__device__ void trace_node(W& w, u32 nf, u32 Cs, u32 kl) {
  const u32 k = min(nf, 64u);
  u32 base = 0;
  if (LANE == 0) base = atomicAdd(w.trace_cnt, k);
  base = bcast(base);
  if (LANE < k && base + LANE < w.trace_cap) {
    const u32 x = w.front[LANE];
    const MlpFeat f = cand_features(w, x, Cs, kl);
    u32* r = w.trace + (u64)(base + LANE) * 8;
    r[0] = w.tag;
    r[1] = w.Q->orig[x];
    r[2] = f.dC;
    r[3] = f.dR;
    r[4] = f.dist;
    r[5] = f.core;
    r[6] = (u32)f.slack;
    r[7] = __float_as_uint(f.cscore);
  }
}

// Zeroes acc[] of the frontier (four independent entries per lane per step).
__device__ void clear_acc(W& w, u32 nf) {
  u32 i = LANE;
  for (; i + 192 < nf; i += 256) {
    const u32 x0 = w.front[i], x1 = w.front[i + 64], x2 = w.front[i + 128], x3 = w.front[i + 192];
    w.vs[x0].acc = 0;
    w.vs[x1].acc = 0;
    w.vs[x2].acc = 0;
    w.vs[x3].acc = 0;
  }
  for (; i < nf; i += 64) w.vs[w.front[i]].acc = 0;
  wsync();
}

// LEARN selects the kernel variant that can score candidates with the MLP;
// the plain variant leaves that path out and needs half the registers.
template <bool LEARN>
__device__ NodeOut process(W& w, u32 from, u32 kl_seen, u64 nodeno, u64& tp, u32& ph) {
  Lds* L = w.L;
  NodeOut o;
  o.branch = 0;
  o.nD = 0;
  o.ustar = INV;
  const u32 h = w.h;
  u32 kl = cur_kl(w);
  bool full = (kl != kl_seen);
  u64 pair_left = PAIR_NODE_SCAN;
  while (true) {
    const u32 Cs = w.nC;
    if (Cs >= w.l && Cs <= h) {
      u32 mn = INV;
      for (u32 i = LANE; i < Cs; i += 64) mn = min(mn, (u32)w.vs[w.clist[i]].dC);
      mn = wmin(mn);
      if (mn != INV && mn >= kl) offer(w, mn);
      u32 nk = cur_kl(w);
      if (nk != kl) full = true;
      kl = nk;
    }
    if (Cs >= h) {
      o.klc = kl;
      return o;
    }
    if ((w.flags & QF_RED3) && tight_exists(w, kl)) {
      // tight C vertices first (their dC+dR is invariant under moves)
      if (LANE == 0) L->cnt = 0;
      wsync();
      for (u32 i = LANE; i < Cs; i += 64) {
        u32 u = w.clist[i];
        if (w.vs[u].dC + w.vs[u].dR == kl + 1 && w.vs[u].dR > 0) w.work[atomicAdd(&L->cnt, 1u)] = u;
      }
      wsync();
      const u32 nt = bcast(L->cnt);
      bool over = false;
      for (u32 t = 0; t < nt && !over; t++) {
        u32 u = w.work[t];
        const u32 b = w.off[u], e = w.off[u + 1];
        for (u32 j = b; j < e; j++) {
          u32 x = w.adj[j];
          if (w.vs[x].st == ST_R) {
            move_c<4>(w, x);
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
    const bool removed = r12(w, from, full, kl, tp, ph);
    from = w.ntrail;
    full = false;
    if (removed && (w.flags & QF_RED3) && tight_exists(w, kl)) continue;
    if (w.flags & QF_PAIR) {
      ph_switch(L, tp, ph, PH_PAIR);
      const int pr = pair_step(w, kl, pair_left);
      ph_switch(L, tp, ph, PH_RULES);
      if (pr < 0) {
        o.klc = kl;
        return o;
      }
      if (pr > 0) continue;
    }
    break;
  }
  o.klc = kl;
  const u32 Cs = w.nC;
  if (Cs + w.nR < w.l || w.nR == 0) return o;
  // bounds
  ph_switch(L, tp, ph, PH_BOUNDS);
  u32 ub = INV;
  if (w.flags & QF_PRUN1) {
    u32 b = INV;
    for (u32 i = LANE; i < Cs; i += 64) {
      u32 u = w.clist[i];
      u32 dc = w.vs[u].dC, dr = w.vs[u].dR;
      b = min(b, min(dc + h - Cs, dc + dr));
    }
    ub = wmin(b);
  }
  if ((w.flags & QF_PRUN2) && ub > kl) ub = min(ub, unr_bound(w, min(h - Cs, w.nR)));
  if (ub <= kl) return o;
  // connection scores by scattering from C into acc[]; frontier = R with dC>0
  ph_switch(L, tp, ph, PH_SCORE);
  if (LANE == 0) L->nfront = 0;
  wsync();
  const bool single0 = (Cs == 1 && w.vs[w.clist[0]].dC == 0);
  if (w.bm) {
    // C is marked in xb for the scatter, so "x is R" is one LDS bit test
    // (C and X neighbours are skipped alike); tb records which R vertices
    // the scatter has reached
    for (u32 i = LANE; i < Cs; i += 64) atomicOr(&w.xb[w.clist[i] >> 5], 1u << (w.clist[i] & 31));
    wsync();
  }
  // The adjacency lists of C, concatenated: 64 C vertices at a time, each
  // lane loads one vertex's setup (score term, list bounds) and the wave
  // walks the concatenated lists 64 entries per chunk; the owner of an
  // entry is found by binary search over the lanes' exclusive degree sums.
  // U chunks per step keep their loads, tests and atomics in flight
  // together.
#ifndef SCS_SCORE_U
#define SCS_SCORE_U 4
#endif
  constexpr int U = SCS_SCORE_U;
  for (u32 c0 = 0; c0 < Cs; c0 += 64) {
    const u32 ci = c0 + LANE;
    u32 ub = 0, ud = 0;
    u64 ur = 0;
    if (ci < Cs) {
      const u32 u = w.clist[ci];
      ur = recip32(w.vs[u].dC);
      ub = w.off[u];
      ud = w.off[u + 1] - ub;
    }
    u32 inc = ud;
    for (int o = 1; o < 64; o <<= 1) {
      const u32 t = __shfl_up(inc, o);
      if (LANE >= (u32)o) inc += t;
    }
    const u32 tot = __shfl(inc, 63);
    const u32 exc = inc - ud;
    // entry kk of the concatenated lists (clamped to the last one): its
    // neighbour id and its owner's score term
    auto fetch = [&](u32 k0, u32* x, u64* rr) {
#pragma unroll
      for (int k = 0; k < U; k++) {
        const u32 kk = k0 + 64 * k + LANE;
        const u32 q = kk < tot ? kk : tot - 1;
        u32 lo = 0;
        for (u32 step = 32; step > 0; step >>= 1) {
          const u32 probe = lo + step;
          const u32 pe = __shfl(exc, (int)(probe < 64 ? probe : 63));
          if (probe < 64 && pe <= q) lo = probe;
        }
        const u32 ob = __shfl(ub, (int)lo), oe = __shfl(exc, (int)lo);
        const u32 rlo = __shfl((u32)ur, (int)lo), rhi = __shfl((u32)(ur >> 32), (int)lo);
        rr[k] = ((u64)rhi << 32) | rlo;
        x[k] = w.adj[ob + (q - oe)];
      }
    };
    if (w.bm && tot > 0) {
      // Software-pipelined: the next step's adjacency loads are issued
      // before this step's score atomics, which return nothing, so no step
      // waits for an atomic (the memory counter is in order: a load issued
      // after an atomic would wait for it).  First touch of an R vertex is
      // an LDS test-and-set on tb.
      u32 xa[U], xn[U];
      u64 ra[U], rn[U];
      fetch(0, xa, ra);
      for (u32 k0 = 0; k0 < tot; k0 += 64 * U) {
        fetch(k0 + 64 * U < tot ? k0 + 64 * U : k0, xn, rn);
        bool ok[U];
#pragma unroll
        for (int k = 0; k < U; k++) ok[k] = k0 + 64 * k + LANE < tot && !xbit(w, xa[k]);
#pragma unroll
        for (int k = 0; k < U; k++)
          if (ok[k]) atomicAdd((unsigned long long*)&w.vs[xa[k]].acc, (unsigned long long)ra[k]);
#pragma unroll
        for (int k = 0; k < U; k++)
          if (ok[k]) {
            const u32 bit = 1u << (xa[k] & 31);
            if (!(atomicOr(&w.tb[xa[k] >> 5], bit) & bit)) w.front[atomicAdd(&L->nfront, 1u)] = xa[k];
          }
#pragma unroll
        for (int k = 0; k < U; k++) {
          xa[k] = xn[k];
          ra[k] = rn[k];
        }
      }
    } else if (tot > 0) {
      // Status from device memory, first touch from the returned old score.
      // Pipelined over two steps: a step issues the next step's adjacency
      // loads with its status gathers, then its score atomics, and only then
      // appends the previous step's first touches, whose returned words
      // have arrived by the time this step's gathers have.
      u32 xa[U], xp[U];
      u64 ra[U], op[U];
      bool okp[U];
      fetch(0, xa, ra);
#pragma unroll
      for (int k = 0; k < U; k++) okp[k] = false;
      for (u32 k0 = 0; k0 < tot; k0 += 64 * U) {
        u32 xn[U];
        u64 rn[U];
        fetch(k0 + 64 * U < tot ? k0 + 64 * U : k0, xn, rn);
        bool ok[U];
#pragma unroll
        for (int k = 0; k < U; k++) ok[k] = k0 + 64 * k + LANE < tot && w.vs[xa[k]].st == ST_R;
        u64 old[U];
#pragma unroll
        for (int k = 0; k < U; k++)
          old[k] = ok[k] ? atomicAdd((unsigned long long*)&w.vs[xa[k]].acc, (unsigned long long)ra[k])
                         : 1ull;
#pragma unroll
        for (int k = 0; k < U; k++)
          if (okp[k] && op[k] == 0) w.front[atomicAdd(&L->nfront, 1u)] = xp[k];
#pragma unroll
        for (int k = 0; k < U; k++) {
          xp[k] = xa[k];
          op[k] = old[k];
          okp[k] = ok[k];
          xa[k] = xn[k];
          ra[k] = rn[k];
        }
      }
#pragma unroll
      for (int k = 0; k < U; k++)
        if (okp[k] && op[k] == 0) w.front[atomicAdd(&L->nfront, 1u)] = xp[k];
    }
    w.edges += tot;
  }
  // the barrier's fence waits for the score atomics
  wsync();
  if (w.bm) {
    for (u32 i = LANE; i < Cs; i += 64) atomicAnd(&w.xb[w.clist[i] >> 5], ~(1u << (w.clist[i] & 31)));
    const u32 nfr = L->nfront;
    for (u32 i = LANE; i < nfr; i += 64) w.tb[w.front[i] >> 5] = 0u;  // tb was all zero before
    wsync();
  }
  const u32 nf = bcast(L->nfront);
  // argmax score, ties to the lowest id
  ph_switch(L, tp, ph, PH_ARGMAX);
  u64 best = 0;
  u32 bid = INV;
  auto consider = [&](u32 x, u64 s) {
    if (s > best || (s == best && x < bid)) {
      best = s;
      bid = x;
    }
  };
  {
    u32 i = LANE;
    // eight candidates per lane per step: their gathers overlap
    for (; i + 448 < nf; i += 512) {
      u32 xs[8];
#pragma unroll
      for (int k = 0; k < 8; k++) xs[k] = w.front[i + 64 * k];
      VState vv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) vv[k] = w.vs[xs[k]];
#pragma unroll
      for (int k = 0; k < 8; k++) consider(xs[k], vv[k].acc + rterm(vv[k].dR, w.dmax));
    }
    for (; i < nf; i += 64) {
      const u32 x = w.front[i];
      consider(x, w.vs[x].acc + rterm(w.vs[x].dR, w.dmax));
    }
  }
  const u64 gbest = wmax64(best);
  u32 us = wmin(best == gbest ? bid : INV);
  ph_switch(L, tp, ph, PH_SCORE);
  if (nf > 0 && w.trace && (nodeno % w.trace_stride) == 0) trace_node(w, nf, Cs, kl);
  if (LEARN && nf > 0 && us != INV && (w.flags & QF_LEARNED) && w.model) {
    // learned order: any frontier vertex is a valid branching vertex.  The
    // MLP re-ranks a shortlist (plain score at least half the best, at most
    // LEARN_SHORTLIST vertices) so that its cost does not grow with the
    // frontier, which hubs make tens of thousands of vertices wide.
    const u64 half = gbest >> 1;
    if (LANE == 0) L->nk = 0;
    wsync();
    for (u32 i = LANE; i < nf; i += 64) {
      const u32 x = w.front[i];
      const VState v = w.vs[x];
      if (v.acc + rterm(v.dR, w.dmax) >= half) {
        const u32 p = atomicAdd(&L->nk, 1u);
        if (p < LEARN_SHORTLIST) w.aux[p] = x;
      }
    }
    wsync();
    const u32 ns = min(bcast(L->nk), (u32)LEARN_SHORTLIST);
    const u32 ul = learned_ustar(w, w.aux, ns, Cs, kl);
    if (ul != INV) us = ul;
  }
  if (nf == 0 || us == INV) {
    ph_switch(L, tp, ph, PH_CLEAR);
    clear_acc(w, nf);
    return o;
  }
  // dominated set of us
  ph_switch(L, tp, ph, PH_DOMS);
  u32 nD = 0;
  if ((w.flags & QF_DOMS) && w.limit_doms > 0) {
    const u32 lu = w.vs[us].dC + w.vs[us].dR;
    const u32 ub0 = w.off[us], ue = w.off[us + 1];
    const bool bm = w.bm;
    if (bm) {
      // tb = N[us]: the test becomes two LDS bit lookups per entry
      for (u32 j = ub0 + LANE; j < ue; j += 64) tset(w, w.adj[j]);
      if (LANE == 0) tset(w, us);
    } else {
      for (u32 j = ub0 + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 1;
    }
    if (LANE == 0) L->ncand = 0;
    wsync();
    // candidates: frontier vertices (R with dC > 0) adjacent to us, a sound
    // subset of the dominated set; x is dominated when every live neighbour
    // is us or a neighbour of us.  One candidate per lane: the scans mostly
    // stop after a few entries.
    u32 ed = 0, ntest = 0;
    for (u32 jj = ub0 + LANE; jj < ue; jj += 64) {
      const u32 x = w.adj[jj];
      const VState xs = w.vs[x];
      if (xs.st != ST_R || xs.dC == 0 || xs.dC + xs.dR > lu + 1) continue;
      ntest++;
      bool dom = true;
      const u32 b = w.off[x], e = w.off[x + 1];
      if (bm) {
        // eight entries per step (their loads in flight together); us is
        // in tb, so a padding entry of us always passes
        for (u32 j = b; j < e && dom; j += 8) {
          u32 y[8];
#pragma unroll
          for (int t = 0; t < 8; t++) y[t] = j + t < e ? w.adj[j + t] : us;
          ed += min(8u, e - j);
          bool ok = true;
#pragma unroll
          for (int t = 0; t < 8; t++) ok = ok && (tbit(w, y[t]) || xbit(w, y[t]));
          if (!ok) dom = false;
        }
      } else {
        // four entries per step with their status and mark gathers in
        // flight together; a padding entry of us always passes
        for (u32 j = b; j < e && dom; j += 4) {
          u32 y[4];
#pragma unroll
          for (int t = 0; t < 4; t++) y[t] = j + t < e ? w.adj[j + t] : us;
          u8 st[4], nm[4];
#pragma unroll
          for (int t = 0; t < 4; t++) st[t] = w.vs[y[t]].st;
#pragma unroll
          for (int t = 0; t < 4; t++) nm[t] = w.nmark[y[t]];
          ed += min(4u, e - j);
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (st[t] != ST_X && y[t] != us && !nm[t]) dom = false;
        }
      }
      if (dom) w.work[atomicAdd(&L->ncand, 1u)] = x;
    }
    w.edges += wsum(ed) + (ue - ub0);
    wsync();
    if (bm) {
      for (u32 j = ub0 + LANE; j < ue; j += 64) tclr(w, w.adj[j]);
      if (LANE == 0) tclr(w, us);
    } else {
      for (u32 j = ub0 + LANE; j < ue; j += 64) w.nmark[w.adj[j]] = 0;
    }
    const u32 nc = bcast(L->ncand);
    ntest = wsum(ntest);
    if (LANE == 0) {
      atomicAdd((unsigned long long*)&w.gl->dom_front, (unsigned long long)nf);
      atomicAdd((unsigned long long*)&w.gl->dom_cand, (unsigned long long)ntest);
      atomicAdd((unsigned long long*)&w.gl->dom_found, (unsigned long long)nc);
    }
    // top limit_doms by (score desc, id asc); the score counts C neighbours
    // with dC != 0 (only C = {q} with dC(q) = 0 differs from acc)
    const u32 want = min(nc, w.limit_doms);
    for (u32 t = 0; t < want; t++) {
      u64 bs = 0;
      u32 bx = INV;
      for (u32 i = LANE; i < nc; i += 64) {
        u32 x = w.work[i];
        if (x == INV) continue;
        u64 s = (single0 ? w.vs[x].acc - (1ull << 32) : w.vs[x].acc) + rterm(w.vs[x].dR, w.dmax);
        if (bx == INV || s > bs || (s == bs && x < bx)) {
          bs = s;
          bx = x;
        }
      }
      const u64 key = (bx == INV) ? 0ull : bs + 1;
      const u64 gb = wmax64(key);
      const u32 pick = wmin((bx != INV && key == gb) ? bx : INV);
      if (LANE == 0) L->D[t] = pick;
      for (u32 i = LANE; i < nc; i += 64)
        if (w.work[i] == pick) w.work[i] = INV;
      wsync();
    }
    nD = want;
  }
  ph_switch(L, tp, ph, PH_CLEAR);
  clear_acc(w, nf);
  o.branch = 1;
  o.ustar = us;
  o.nD = nD;
  return o;
}

// ---------------------------------------------------------------------------
// Task pool (Vyukov bounded MPMC ring).  Lane 0 claims; the wave copies.
// ---------------------------------------------------------------------------
// limit: most tasks the pool may hold after this enqueue.  deq only grows,
// so pos - (deq read before the attempt) bounds the occupancy strictly.
__device__ int pool_try_enqueue(const Pool& P, u64* out_pos, u32 limit) {
  const u64 deq0 = aload64(P.deq);
  u64 pos = aload64(P.enq);
  for (int tries = 0; tries < 64; tries++) {
    if (pos >= deq0 + limit) return 0;
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
__device__ __forceinline__ void rebuild(W& w, const u32* task) {
  const u32 nC = task[1], nL = task[2], form = task[3];
  const u32* C = task + 4;
  const u32* Lst = task + 4 + nC;
  const u32 K = w.K;
  // queued[], nmark[] and acc[] are all-zero between nodes (every node
  // clears what it set), so only the status needs initialising.  The O(K)
  // passes below handle eight 64-vertex chunks per step, so each lane keeps
  // eight loads in flight instead of waiting on one at a time.
  const u32 nw = (K + 31) / 32;
  u32 base = 0;
  // one chunk of the pass that writes the X bitmap (when in use), the R and
  // C bitmaps of gb and the R list (ascending ids) from the chunk's status
  // values
  auto compact = [&](u32 c0, bool in, bool isx, bool isr, bool isc) {
    if (w.bm) {
      const u64 m = __ballot(in && isx);
      if (LANE == 0) w.xb[c0 >> 5] = (u32)m;
      if (LANE == 1 && (c0 >> 5) + 1 < nw) w.xb[(c0 >> 5) + 1] = (u32)(m >> 32);
    }
    const u64 bmk = __ballot(in && isr);
    const u64 cmk = __ballot(in && isc);
    if (LANE == 0) w.gb[c0 >> 5] = (bmk & 0xFFFFFFFFull) | (cmk << 32);
    if (LANE == 1 && (c0 >> 5) + 1 < nw)
      w.gb[(c0 >> 5) + 1] = (bmk >> 32) | (cmk & 0xFFFFFFFF00000000ull);
    if (in && isr) {
      const u32 p = base + __popcll(bmk & lanes_below());
      w.rlist[p] = c0 + LANE;
      w.pos[c0 + LANE] = p;
    }
    base += __popcll(bmk);
  };
  if (form >= 2) {
    // a level's root (C = {q}): R = the vertices that can lie in an H with
    // min degree > kl and |H| <= h (core > kl, within ubD(kl, h) of q); the
    // root's rule-1/2 closure would remove every other one.  A plain root of
    // a whole-graph query (form 3): R = every other kept vertex (the ones
    // with a distance)
    const u32 kl = cur_kl(w), ub = form == 2 ? ubd_node(kl, w.h) : SCS_DIST_NONE - 1, q0 = C[0];
    for (u32 v0 = 0; v0 < K; v0 += 512) {
      u32 cr[8], ds[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 v = min(v0 + 64u * k + LANE, K - 1);
        cr[k] = w.core[v];
        ds[k] = w.dist[v];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 c0 = v0 + 64u * k;
        if (c0 >= K) break;
        const u32 v = c0 + LANE;
        const bool in = v < K, isq = v == q0;
        const bool keep = !isq && (form == 3 || cr[k] > kl) && ds[k] <= ub;
        if (in) w.vs[v].st = isq ? ST_C : (keep ? ST_R : ST_X);
        compact(c0, in, !isq && !keep, keep, in && isq);
      }
    }
    if (LANE == 0) w.clist[0] = q0;
  } else {
    const u8 fill = form ? ST_X : ST_R;
    for (u32 v = LANE; v < K; v += 64) w.vs[v].st = fill;
    wsync();
    for (u32 i = LANE; i < nL; i += 64) w.vs[Lst[i]].st = form ? ST_R : ST_X;
    wsync();
    for (u32 i = LANE; i < nC; i += 64) {
      w.vs[C[i]].st = ST_C;
      w.clist[i] = C[i];
    }
    wsync();
    for (u32 v0 = 0; v0 < K; v0 += 512) {
      u8 st[8];
#pragma unroll
      for (int k = 0; k < 8; k++) st[k] = w.vs[min(v0 + 64u * k + LANE, K - 1)].st;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 c0 = v0 + 64u * k;
        if (c0 >= K) break;
        compact(c0, c0 + LANE < K, st[k] == ST_X, st[k] == ST_R, st[k] == ST_C);
      }
    }
  }
  if (w.bm)
    for (u32 i = LANE; i < nw; i += 64) w.tb[i] = 0;
  for (u32 i = LANE; i < SCS_HCAP + 4; i += 64) w.L->hist[i] = 0;
  wsync();
  w.nR = base;
  w.nC = nC;
  wsync();
  const u32 nR = w.nR;
  if (!form) {
    // the list holds X (smaller than the live set): start every counter
    // from the local degree (one coalesced pass) and subtract what X and C
    // take away.  dR(v) = deg(v) - |N(v) & X| - |N(v) & C|, dC(v) = |N(v) & C|;
    // counters of X vertices are never read, so their values do not matter.
    for (u32 v0 = 0; v0 < K; v0 += 512) {
      u32 a[8], b[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 v = min(v0 + 64u * k + LANE, K - 1);
        a[k] = w.off[v];
        b[k] = w.off[v + 1];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const u32 v = v0 + 64u * k + LANE;
        if (v < K) {
          w.vs[v].dC = 0;
          w.vs[v].dR = b[k] - a[k];
        }
      }
    }
    wsync();
    if (nL >= 8) {
      scan_adj<u32>(
          w, nL, [&](u32 i) { return Lst[i]; },
          [&](u32 y) {
            atomicSub(&w.vs[y].dR, 1u);
            return 0u;
          },
          [&](u32, u32, u32) {});
    } else {
      for (u32 i = 0; i < nL; i++) {
        const u32 x = Lst[i], b = w.off[x], e = w.off[x + 1];
        for (u32 j = b + LANE; j < e; j += 64) atomicSub(&w.vs[w.adj[j]].dR, 1u);
        w.edges += e - b;
      }
    }
    wsync();
    // C one vertex at a time: its neighbours are distinct, so plain updates
    for (u32 i = 0; i < nC; i++) {
      const u32 c = C[i], b = w.off[c], e = w.off[c + 1];
      for (u32 j = b + LANE; j < e; j += 64) {
        const u32 y = w.adj[j];
        w.vs[y].dC = (u16)(w.vs[y].dC + 1);
        w.vs[y].dR -= 1;
      }
      w.edges += e - b;
      wsync();
    }
    for (u32 i0 = 0; i0 < nR; i0 += 64) {
      const u32 i = i0 + LANE;
      const bool on = i < nR;
      const u32 d = on ? min((u32)w.vs[w.rlist[i]].dC, w.h + 1) : 0u;
      hist_add(w.L->hist, on, d, 1);
    }
    w.ntrail = 0;
    w.nframes = 0;
    w.nsav = 0;
    wsync();
    return;
  }
  // the list holds R: count the live vertices' neighbours directly
  // (dC in the high word, dR in the low word).  Membership comes from the
  // gb bitmaps (8 bytes per 32 vertices) rather than from the 16-byte state
  // records, so the gathers land in a small array that stays in the caches.
  scan_adj<u64>(
      w, nR + nC, [&](u32 i) { return i < nR ? w.rlist[i] : w.clist[i - nR]; },
      [&](u32 x) {
        const u64 g = w.gb[x >> 5];
        const u32 b = x & 31;
        return (((g >> (32 + b)) & 1ull) << 32) | ((g >> b) & 1ull);
      },
      [&](u32 i, u32 v, u64 cr) {
        const u32 c = (u32)(cr >> 32), r = (u32)cr;
        w.vs[v].dC = (u16)c;
        w.vs[v].dR = r;
        if (i < nR) atomicAdd(&w.L->hist[min(c, w.h + 1)], 1u);
      });
  w.ntrail = 0;
  w.nframes = 0;
  w.nsav = 0;
  wsync();
}

// Serialises the Z child of frame fi (the unexplored "exclude u*" branch)
// into dst (dst[3] = form set by the caller: 1 = list R, 0 = list X);
// returns the payload length in words.
__device__ __forceinline__ u32 serialize_z(W& w, u32 fi, u32* dst, u32 qslot) {
  const Frame& f = w.frames[fi];
  // mark vertices changed after f.mark (they were R at the frame)
  for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) {
    const u32 e = w.trail[t];
    if (!(e & TR_MARK)) w.nmark[e & TR_VMASK] = 1;
  }
  wsync();
  for (u32 i = LANE; i < f.nD; i += 64) w.nmark[f.D[i]] = 2;
  if (LANE == 0) w.nmark[f.ustar] = 2;
  wsync();
  for (u32 i = LANE; i < f.csize; i += 64) dst[4 + i] = w.clist[i];
  const u32 form = dst[3];
  u32* out = dst + 4 + f.csize;
  u32 base = 0;
  for (u32 v0 = 0; v0 < w.K; v0 += 64) {
    u32 v = v0 + LANE;
    bool take = false;
    if (v < w.K) {
      // at the frame: nmark 2 = ustar or D (out in Z), 1 = changed after the
      // frame's mark (was R), 0 = unchanged (current status holds)
      u8 s = w.vs[v].st, m = w.nmark[v];
      bool inRz = (m == 1) || (m == 0 && s == ST_R);
      bool inC = (m == 0 && s == ST_C);
      take = form ? inRz : (!inRz && !inC);
    }
    u64 bm = __ballot(take);
    if (take) out[base + __popcll(bm & lanes_below())] = v;
    base += __popcll(bm);
  }
  if (LANE == 0) {
    dst[0] = qslot;
    dst[1] = f.csize;
    dst[2] = base;
  }
  wsync();
  for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) {
    const u32 e = w.trail[t];
    if (!(e & TR_MARK)) w.nmark[e & TR_VMASK] = 0;
  }
  wsync();
  for (u32 i = LANE; i < f.nD; i += 64) w.nmark[f.D[i]] = 0;
  if (LANE == 0) w.nmark[f.ustar] = 0;
  wsync();
  return 4 + f.csize + base;
}

__device__ __forceinline__ void save_state(W& w, u32 active, u32 qslot, u32 from, u32 kl_seen) {
  if (active)
    for (u32 i = LANE; i < SCS_HCAP + 4; i += 64) w.ghist[i] = w.L->hist[i];
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
    H->pend_nodes = 0;
    H->nsav = w.nsav;
  }
  wsync();
}

// The wave that ends a query's last task either starts the query's next
// level or marks the query done.  A level that ends with no witness above
// its floor proves opt <= floor: the stopping bound ku drops to the floor,
// the floor goes one lower, and this wave itself searches the new level's
// root (return value true; its payload is left in the wave's LDS, and the
// pool is not involved, so a full ring cannot stop a level).  Every task of the query has ended and flushed its node
// count, so the slot's fields are this wave's alone here.
__device__ bool finish_task(W& w, const EngineArgs& A, u32 qslot) {
  u32 again = 0;
  if (LANE == 0) {
    QuerySlot* Q = &A.slots[qslot];
    u32 left = atomicSub(&Q->pending, 1u) - 1;
    if (left == 0) {
      __threadfence();
      const u32 kl = aload(&Q->kl), fl = aload(&Q->floor);
      if ((Q->flags & QF_LEVELS) && !aload(&Q->stop) && kl < fl) {
        Q->ku = min(aload(&Q->ku), fl);
        Q->floor = fl - 1;
        Q->lvl_nodes0 = aload64(&Q->nodes);
        Q->pending = 1;
        u32* rt = w.L->root;
        rt[0] = qslot;
        rt[1] = 1;
        rt[2] = 0;
        rt[3] = 2;  // level root
        rt[4] = Q->qloc;
        __threadfence();
        again = 1;
      } else {
        atomicExch(&Q->state, (u32)QS_DONE);
        atomicSub(&A.glob->active_q, 1u);
      }
    }
  }
  again = bcast(again);
  wsync();
  return again != 0;
}

__device__ void flush_counters(W& w, const EngineArgs& A, u32& pend) {
  if (LANE == 0) {
    if (pend) {
      atomicAdd((unsigned long long*)&w.Q->nodes, (unsigned long long)pend);
      atomicAdd((unsigned long long*)&A.glob->nodes, (unsigned long long)pend);
    }
    if (w.edges) atomicAdd((unsigned long long*)&A.glob->edges, (unsigned long long)w.edges);
  }
  pend = 0;
  w.edges = 0;
}

// ---------------------------------------------------------------------------
// The persistent epoch kernel.
// ---------------------------------------------------------------------------
// flatten: every call is inlined, so the wave state W never needs an address
// in private (scratch) memory.  Launch bounds: at least 2 waves per SIMD,
// i.e. up to 256 registers: the pipelined walks keep two steps of loads and
// atomics in registers (at 128 they spilled), and the engine runs at most 8
// waves per CU.
template <bool LEARN>
__global__ void __launch_bounds__(64, 2) __attribute__((flatten)) scs_epoch(EngineArgs A) {
  __shared__ Lds L;
  const u32 wave = blockIdx.x;
  if (wave >= A.nwaves) return;
  const u64 t0 = wall_clock64();
  const u64 deadline = t0 + A.slice_ticks;
  W w;
  bind(w, A, wave, &L);
  if (LANE < SCS_PHASES) L.tph[LANE] = 0;
  wsync();
  u64 tp = t0;
  u32 ph = PH_IDLE;
  WaveHdr* H = w.hdr;
  u32 active = H->active, qslot = H->qslot;
  u32 from = H->from, kl_seen = H->kl_seen, pend = 0;
  if (active) {
    bind_query(w, &A.slots[qslot]);
    w.nframes = H->nframes;
    w.ntrail = H->trail;
    w.nC = H->nC;
    w.nR = H->nR;
    w.nsav = H->nsav;
    for (u32 i = LANE; i < SCS_HCAP + 4; i += 64) L.hist[i] = w.ghist[i];
    wsync();
    build_bitmaps(w);
  }
  bool idle_counted = false;
  u32 idle_spins = 0;
  while (true) {
    if (wall_clock64() > deadline) break;
    if (!active) {
      // take a task from the pool
      if (ph != PH_IDLE) ph_switch(&L, tp, ph, PH_IDLE);
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
        // back off: idle waves polling the pool contend with busy waves'
        // atomics on the same cache lines
        if (idle_spins < 8)
          __builtin_amdgcn_s_sleep(8);
        else
          __builtin_amdgcn_s_sleep(127);
        continue;
      }
      if (idle_counted) {
        if (LANE == 0) atomicSub(&A.glob->idle, 1u);
        idle_counted = false;
      }
      idle_spins = 0;
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
      ph_switch(&L, tp, ph, PH_REBUILD);
      rebuild(w, task);
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
      } else if (aload(&Q->kl) >= aload(&Q->ku)) {
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
      flush_counters(w, A, pend);
      finish_task(w, A, qslot);
      continue;
    }
    ph_switch(&L, tp, ph, PH_RULES);
    NodeOut o = process<LEARN>(w, from, kl_seen, w.trace ? aload64(&Q->nodes) + pend : 0ull, tp, ph);
    ph_switch(&L, tp, ph, PH_MOVE);
    pend++;
    if (pend >= A.node_flush) flush_counters(w, A, pend);
    if (o.branch) {
      const u32 fi = w.nframes;
      Frame* f = &w.frames[fi];
      if (LANE == 0) {
        f->mark = w.ntrail;
        f->ustar = o.ustar;
        f->nD = o.nD;
        f->next = 0;
        f->kl = o.klc;
        f->csize = w.nC;
        f->zgone = 0;
      }
      for (u32 i = LANE; i < o.nD; i += 64) f->D[i] = L.D[i];
      wsync();
      w.nframes = fi + 1;
      from = w.ntrail;
      kl_seen = o.klc;
      const Clo none{0, 0, nullptr, nullptr};
      remove_batch(w, f->D, o.nD, none);
      move_c<4>(w, o.ustar);
    } else {
      // backtrack to the next pending child
      bool cont = false;
      while (w.nframes > 0) {
        Frame* f = &w.frames[w.nframes - 1];
        undo_to<4>(w, f->mark);
        from = f->mark;
        kl_seen = f->kl;
        const u32 next = f->next + 1;
        if (LANE == 0) f->next = next;
        wsync();
        const u32 nD = f->nD;
        const Clo none{0, 0, nullptr, nullptr};
        if (next <= nD) {
          const u32 i = next - 1;
          remove_batch(w, f->D, i, none);
          move_c<4>(w, f->ustar);
          move_c<4>(w, f->D[i]);
          cont = true;
          break;
        }
        if (f->zgone) {
          w.nframes--;
          continue;
        }
        // Z: drop ustar and D
        if (LANE == 0) w.work[0] = f->ustar;
        for (u32 i = LANE; i < nD; i += 64) w.work[1 + i] = f->D[i];
        wsync();
        w.nframes--;
        remove_batch(w, w.work, nD + 1, none);
        cont = true;
        break;
      }
      if (!cont) {
        active = 0;
        flush_counters(w, A, pend);
        if (finish_task(w, A, qslot)) {
          // the query's next level: its root, searched by this wave
          ph_switch(&L, tp, ph, PH_REBUILD);
          rebuild(w, L.root);
          active = 1;
          from = 0;
          kl_seen = INV;
          wsync();
        }
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
          // R at Z = current R + every vertex changed after the frame's
          // mark (markers excluded) - ustar - D
          u32 chg = 0;
          for (u32 t = f.mark + LANE; t < w.ntrail; t += 64) chg += !(w.trail[t] & TR_MARK);
          chg = wsum(chg);
          const u32 nRz = w.nR + chg - 1 - f.nD;
          const u32 nXz = w.K - f.csize - nRz;
          const u32 form = nRz <= nXz ? 1u : 0u;
          const u32 need = 4 + f.csize + (form ? nRz : nXz);
          u64 pos = 0;
          u32 got = 0;
          if (need <= A.pool.slot_words && LANE == 0) got = pool_try_enqueue(A.pool, &pos, A.donor_cap);
          got = bcast(got);
          if (got) {
            u32 plo = __shfl((u32)pos, 0), phi = __shfl((u32)(pos >> 32), 0);
            pos = ((u64)phi << 32) | plo;
            u32* dst = A.pool.data + (pos % A.pool.nslots) * (u64)A.pool.slot_words;
            if (LANE == 0) dst[3] = form;
            wsync();
            ph_switch(&L, tp, ph, PH_DONATE);
            serialize_z(w, fi, dst, qslot);
            ph_switch(&L, tp, ph, PH_MOVE);
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
  ph_switch(&L, tp, ph, ph);
  wsync();
  if (LANE < SCS_PHASES && L.tph[LANE])
    atomicAdd((unsigned long long*)&A.glob->tph[LANE], (unsigned long long)L.tph[LANE]);
  if (active) {
    flush_counters(w, A, pend);
    save_state(w, 1, qslot, from, kl_seen);
  } else {
    save_state(w, 0, 0, 0, INV);
  }
}

// Dynamic LDS of one wave: its share of the CU's 160 KB at the requested
// waves per CU, minus the static part, capped at 48 KB.
static u32 epoch_lds_bytes(u32 waves_per_cu) {
  const u32 wpc = waves_per_cu ? waves_per_cu : 16;
  const i64 share = (i64)(160 * 1024 / wpc) - (i64)sizeof(Lds);
  i64 b = share < (i64)(48 * 1024) ? share : (i64)(48 * 1024);
  b = b > 0 ? (b & ~(i64)511) : 0;
  return (u32)b;
}

extern "C" hipError_t scs_launch_epoch(const EngineArgs* A, hipStream_t s) {
  EngineArgs a = *A;
  const u32 dyn = epoch_lds_bytes(a.waves_per_cu);
  a.lds_words = a.no_bitmaps ? 0u : dyn / 4;
  if (a.model)
    scs_epoch<true><<<a.nwaves, 64, dyn, s>>>(a);
  else
    scs_epoch<false><<<a.nwaves, 64, dyn, s>>>(a);
  return hipGetLastError();
}

// Pushes the root task (C = {q}, nothing removed) of each admitted slot.
__global__ void __launch_bounds__(64) scs_admit(EngineArgs A, const u32* slots_in, u32 n,
                                                u32* failed) {
  for (u32 i = 0; i < n; i++) {
    const u32 sl = slots_in[i];
    u64 pos = 0;
    u32 got = 0;
    if (LANE == 0) got = pool_try_enqueue(A.pool, &pos, A.pool.nslots);
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
      // level root, plain root of a whole-graph query, or plain root
      const u32 fl = A.slots[sl].flags;
      dst[3] = (fl & QF_LEVELS) ? 2u : ((fl & QF_WHOLE) ? 3u : 0u);
      dst[4] = A.slots[sl].qloc;
      failed[i] = 0;
      __threadfence();
      astore64(&A.pool.seq[pos % A.pool.nslots], pos + 1);
    }
    wsync();
  }
}

// Admitted slots, staged by the host: block i writes slot idx[2i] from
// stage[i] and its idx[2i+1] witness seed words.  Runs between epochs.
__global__ void __launch_bounds__(64) scs_slot_write(const QuerySlot* __restrict__ stage,
                                                     const u32* __restrict__ idx,
                                                     const u32* __restrict__ wit, u32 n,
                                                     QuerySlot* __restrict__ slots,
                                                     u32* __restrict__ witness) {
  const u32 i = blockIdx.x;
  if (i >= n) return;
  const u32 si = idx[2 * i], ws = idx[2 * i + 1];
  const u32* src = (const u32*)&stage[i];
  u32* dst = (u32*)&slots[si];
  for (u32 k = LANE; k < sizeof(QuerySlot) / 4; k += 64) dst[k] = src[k];
  for (u32 k = LANE; k < ws && k < SCS_MAX_H; k += 64)
    witness[(u64)si * SCS_MAX_H + k] = wit[(u64)i * SCS_MAX_H + k];
}
extern "C" hipError_t scs_launch_slot_write(const QuerySlot* stage, const u32* idx, const u32* wit,
                                            u32 n, QuerySlot* slots, u32* witness,
                                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  scs_slot_write<<<n, 64, 0, s>>>(stage, idx, wit, n, slots, witness);
  return hipGetLastError();
}

// Queries past their deadline: capped and stopped, so their waves abandon
// them at the next node.  Runs on the engine stream before an epoch.
__global__ void scs_expire(QuerySlot* slots, const u32* idx, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    atomicExch(&slots[idx[i]].capped, 1u);
    atomicExch(&slots[idx[i]].stop, 1u);
  }
}
extern "C" hipError_t scs_launch_expire(QuerySlot* slots, const u32* idx, u32 n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  scs_expire<<<(n + 63) / 64, 64, 0, s>>>(slots, idx, n);
  return hipGetLastError();
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
