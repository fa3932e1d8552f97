// Internal layouts shared by the host driver and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/scs.h"

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

#define SCS_WAVE 64
#define SCS_FRAME_DOMS SCS_MAX_DOMS
#define SCS_HCAP (SCS_MAX_H + 2)
#define SCS_PHASES 13

// vertex status inside a search state
#define ST_R 0
#define ST_C 1
#define ST_X 2

// query slot states
#define QS_FREE 0
#define QS_ACTIVE 1
#define QS_DONE 2

// flags of a query (bit layout shared with the CPU oracle)
#define QF_RED1 1u
#define QF_RED2 2u
#define QF_RED3 4u
#define QF_PRUN1 8u
#define QF_PRUN2 16u
#define QF_DOMS 32u
#define QF_LEARNED 64u
#define QF_LEVELS 128u  // descending level floors
#define QF_PAIR 256u    // common-neighbour pair rule
// engine only: the query searches the resident graph's CSR (global ids), and
// its distance array holds SCS_DIST_NONE for the vertices it does not keep
#define QF_WHOLE 512u
#define SCS_DIST_NONE 0xFFFFFFFFu

// Per-query slot, device resident.  Written by the host at admission (while
// no kernel touches the slot) and by search waves through atomics.
struct QuerySlot {
  u32 state;       // QS_*
  u32 l, h, limit_doms;
  u32 flags;       // QF_*
  u32 ku, K, dmax, qloc;
  u32 kl;          // best min degree so far (atomicMax)
  u32 stop;        // 1 = abandon remaining work (kl >= ku or budget)
  u32 pending;     // tasks queued or being searched
  u32 capped;      // node budget hit
  u32 wlock;       // witness spin lock
  u32 wkl;         // min degree of the stored witness
  u32 wsize;       // |witness| (0 = none)
  u64 nodes;       // nodes expanded (atomicAdd)
  u64 node_budget; // 0 = unlimited
  const u32* off;  // local CSR offsets (K+1)
  const u32* adj;  // local CSR adjacency (local ids, ascending)
  const u32* dist; // BFS distance from q in the core > kl0 subgraph
  const u32* orig; // local -> original vertex id
  const u32* core; // core numbers (MLP feature)
  u32* witness;    // h entries, original ids
  u32 tag;         // low 32 bits of the caller's tag (search traces)
  u32 wlost;       // 1 = a better witness could not be stored (kl not raised for it)
  u32 floor;       // level floor: the search looks for min degree > max(kl, floor)
  u32 pad2;
  u64 lvl_nodes0;  // nodes when this level started
};

// Depth-first frame: one per branching node on the current path.
struct Frame {
  u32 mark;    // trail length after the node's own rule changes
  u32 ustar;
  u32 nD;      // kept dominated vertices
  u32 next;    // child being explored: 0 = A, i = B_{i-1}
  u32 kl;      // kl at the end of the node's rule closure
  u32 csize;   // |C| at the node
  u32 zgone;   // 1 = Z child donated to the pool
  u32 pad;
  u32 D[SCS_FRAME_DOMS];
};

// Header of a wave's persistent depth-first state.
struct WaveHdr {
  u32 active;    // a task is being searched
  u32 qslot;
  u32 nframes;
  u32 trail;     // trail length
  u32 nC, nR;
  u32 from;      // trail position where the pending node's delta starts
  u32 kl_seen;   // kl at the parent's closure (0xffffffff = rescan all)
  u32 K;         // kept-set size of the bound query
  u32 pend_nodes;// nodes not yet added to the query counter
  u32 pend_edges;// adjacency entries not yet added to the global counter
  u32 nsav;      // words on the recount save stack
  u32 pad[4];
};

// Search state of one vertex, packed so that the fields a rule or counter
// update reads together come from one cache line.
struct __attribute__((aligned(16))) VState {
  u32 dR;     // |N(v) & R|
  u16 dC;     // |N(v) & C|
  u8 st;      // ST_R / ST_C / ST_X
  u8 queued;  // queued for removal in the current closure round
  u64 acc;    // connection score being accumulated at the current node (0 between nodes)
};
// The engine updates {dR, dC, st, queued} as one little-endian 64-bit word
// (dR in bits 0-31, dC in 32-47, st in 48-55, queued in 56-63).
static_assert(sizeof(VState) == 16, "VState is 16 bytes");
static_assert(__builtin_offsetof(VState, dR) == 0 && __builtin_offsetof(VState, dC) == 4 &&
                  __builtin_offsetof(VState, st) == 6 && __builtin_offsetof(VState, queued) == 7,
              "VState word layout");

// Offsets (bytes) of each per-wave array inside one wave's region.
// gbits: one u64 per 32 local ids, R bits in the low word and C bits in the
// high word, written by a task rebuild for its neighbour-count scan.
struct WaveLayout {
  u64 vs, nmark, pos, rlist, work, work2, front, aux, trail, sav, frames, clist, hist, hdr, gbits;
  u64 stride;
};

// Pool of donated tasks: bounded MPMC ring (per-slot sequence numbers).
// Task payload: [qslot, nC, nL, form(0 = list is X, 1 = list is R, 2 = a level root:
// C = {q}, R = the kept vertices with core > kl within ubD(kl, h), 3 = a plain root of a
// whole-graph query: C = {q}, R = the other kept vertices)], C[nC], L[nL].
struct Pool {
  u64* seq;        // nslots
  u32* data;       // nslots * slot_words
  u64* enq;        // enqueue position
  u64* deq;        // dequeue position
  u32 nslots;
  u32 slot_words;
};

struct EngineGlobals {
  u32 idle;        // waves currently without work
  u32 active_q;    // queries with pending work
  u64 donations;
  u64 rebuilds;
  u64 nodes;
  u64 edges;       // adjacency entries scanned by search waves
  u32 lock_fail;
  u32 pad[5];
  u64 dom_front;  // frontier vertices over all nodes that branched
  u64 dom_cand;   // of them, candidates whose adjacency was tested
  u64 dom_found;  // of them, dominated by ustar
  u64 full_scans; // rule 1/2 closures that rescanned all of R
  u64 full_rows;  // R vertices those rescans visited
  u64 tph[SCS_PHASES];  // wall-clock ticks (100 MHz) per search phase, all waves
};

struct MlpDev;

// Batched preprocessing: device pointers of one query's local graph, and
// the batch header (query b of the batch owns bit b; queries ordered by
// ascending kl0).
struct MsQuery {
  u32* loff;   // K+1
  u32* ldist;  // K
  u32* lcore;  // K
  u32* lorig;  // K
  u32* ladj;   // nnz
};
struct MsBatch {
  u32 nb;
  u32 pad;
  u32 q[64];
  u32 kl0[64];
  MsQuery P[64];
};

// The adjacency of every vertex cut into units of at most MS_UNIT entries,
// once per graph: the local degree and adjacency passes take one unit per
// wave, so a hub's list is spread over many waves.  A vertex with more than
// one unit (a hub) numbers its units h0 .. h0+nu-1 for the per-unit partial
// counts; a single-unit vertex has h == MS_NOHUB.
#define MS_UNIT 256
#define MS_NOHUB 0xFFFFFFFFu
struct MsUnit {
  u32 u;    // vertex
  u32 h;    // hub unit number, or MS_NOHUB
  u64 pos;  // first adjacency entry (bits 0-47) | entries, at most MS_UNIT (bits 48-63)
};
struct MsHub {
  u32 u;
  u32 h0;  // first hub unit number
  u32 nu;  // units
  u32 pad;
};

// Algorithmic-byte counter of the preprocessing kernels: MS_BYTE_SLOTS
// u64 counters MS_BYTE_STRIDE words apart, summed by the host.
#define MS_BYTE_SLOTS 64
#define MS_BYTE_STRIDE 16

struct EngineArgs {
  const MlpDev* model;  // branching MLP (queries with QF_LEARNED)
  u32* trace;           // search-trace rows (8 words each), or null
  u32* trace_cnt;
  u32 trace_cap;
  u32 trace_stride;
  QuerySlot* slots;
  u32 nslots;
  u8* wave_mem;
  WaveLayout lay;
  u32 kcap;
  u32 nwaves;
  Pool pool;
  EngineGlobals* glob;
  u64 slice_ticks;     // wall_clock64 ticks per epoch
  u32 donate_below;    // donate while fewer tasks are pooled
  u32 donor_cap;       // pooled tasks above which a donation is refused (root tasks keep the rest)
  u32 deterministic;
  u32 node_flush;      // flush node counters every this many nodes
  u32 waves_per_cu;    // resident waves per CU the launch is sized for
  u32 lds_words;       // dynamic LDS words per wave (set by the launcher)
  u32 no_bitmaps;      // 1 = no LDS bitmaps (status from device memory only)
};

// GPU greedy heuristic (scs_heu.hip): byte offsets of one wave's workspace
// (every n-sized array is all-zero between queries) and its two counters.
struct HeuLayout {
  u64 sta, ins, cnt, deg, key, pos, list, rmv, th, bkey, gkey, lds, stride;
};
struct HeuLds {
  u32 m;  // candidates in the list
  u32 pad[3];
};
