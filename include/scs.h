/*
 * scs.h - C ABI of the MI355X-native size-bounded community search engine.
 *
 * The reference (wajidmanzoor/SCS) has no plugin or FFI layer: its host
 * server calls its CUDA kernels directly from main.cu.  This header is the
 * seam that replaces that host->device boundary:
 *
 *   reference                                  this ABI
 *   ---------------------------------------    --------------------------
 *   load_graph + core_decomposition_linear_    scs_graph_load /
 *     list (src/Graph.h:5-109),                scs_graph_create
 *     memoryAllocationGenGraph (main.cu:535)
 *   memeoryAllocationGraph/Task/Buffer         scs_engine_create
 *     (main.cu:536-563), CLI args 2-12
 *     (main.cu:481-502)
 *   processMessages admission: CSSC_heu,       scs_submit
 *     cal_query_dist, initialReductionRules,
 *     CompressTask, NeighborUpdate
 *     (main.cu:84-240)
 *   one level of ProcessTask/FindDoms/Expand/  scs_step
 *     RemoveCompletedTasks (main.cu:261-442)
 *   per-slot completion test and results       scs_poll
 *     line (main.cu:317-360, 444-463)
 *   freeGenGraph/freeGraph/... (main.cu:588)   scs_engine_destroy,
 *                                              scs_graph_destroy
 *
 * All functions return 0 on success or a negative SCS_E* code; none exits
 * the process.  Pointers are plain host pointers; no torch types appear.
 * One host thread drives an engine; scs_submit may be called from any
 * thread (it only appends to a locked host queue).
 */
#ifndef SCS_H_
#define SCS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCS_OK 0
#define SCS_EINVAL (-1)   /* bad argument (malformed query, h == 0, q >= n) */
#define SCS_ENOMEM (-2)   /* device or host allocation failed */
#define SCS_EAGAIN (-3)   /* admission refused for now: slots or memory full */
#define SCS_EHIP (-4)     /* a HIP runtime call failed */
#define SCS_EIO (-5)      /* graph file missing or unreadable */
#define SCS_EBIG (-6)     /* query needs more than the engine's per-query limits */

#define SCS_MAX_H 1000    /* largest upper size bound h accepted */
#define SCS_MAX_DOMS 32   /* largest limit_dominating_set accepted */
#define SCS_MS_LEVELS 1024 /* BFS levels of one preprocessing pass (ubD < h <= 1000) */

/* Result status (reference "Overtime"/"Heu" columns, main.cu:123-128,328-333,
 * 351, 454, extended with explicit per-query outcomes). */
enum scs_status {
  SCS_ST_OK = 0,         /* search complete: kl is the optimum */
  SCS_ST_HEU_OPT = 1,    /* heuristic reached ku = min(core[q], h-1): optimal */
  SCS_ST_CAPPED = 2,     /* node budget exhausted: kl is a lower bound */
  SCS_ST_INFEASIBLE = 3, /* no connected H with q and l <= |H| <= h exists */
  SCS_ST_REJECTED = 4    /* kept set larger than the engine's per-query limit */
};

typedef struct scs_graph scs_graph;
typedef struct scs_engine scs_engine;

/* Engine parameters.  The first twelve mirror the reference's positional
 * command line (main.cu:481-502) with the same meaning where the design
 * keeps the concept; see each comment. */
typedef struct scs_params {
  uint32_t partition_size; /* argv[2]: words of depth-first state per wave
                              (0 = sized from the largest kept set) */
  uint32_t buffer_size;    /* argv[3]: words of the shared task pool
                              (0 = derived from limit_tasks) */
  double copy_limit;       /* argv[4]: a wave donates work while fewer than
                              copy_limit * waves tasks are pooled */
  uint32_t read_limit;     /* argv[5]: kept for the command line; one task is
                              taken per visit to the pool */
  uint32_t limit_queries;  /* argv[6]: resident query slots */
  uint32_t factor;         /* argv[7]: accepted for the command line; task
                              offsets are not compressed here */
  uint32_t red1, red2, red3;  /* argv[8..10]: reduction rules 1-3 */
  uint32_t prun1, prun2;      /* argv[11..12]: bounds U_d and U_nr */
  /* named extensions */
  uint32_t waves_per_cu;   /* resident search waves per CU (0 = 4, max 8: the search
                              kernel's registers allow 8 resident waves per CU); at 4
                              each wave has room for LDS bitmaps over 130K kept vertices */
  uint32_t limit_tasks;    /* task pool slots (0 = 4 * waves) */
  uint64_t node_budget;    /* per-query node cap before CAPPED (0 = none) */
  uint64_t mem_budget;     /* bytes of device memory for per-query graphs */
  uint32_t max_kept;       /* per-query kept-set cap (0 = graph size) */
  uint32_t time_slice_us;  /* length of one scs_step kernel epoch (0 = 200 ms; a node
                              in progress finishes after the deadline) */
  uint32_t deterministic;  /* 1 = a single wave and no work sharing: node
                              counts equal the CPU oracle's */
  uint32_t learned;        /* 1 = MLP scores the branching vertex */
  const void* model;       /* const scs_model* used when learned */
  uint32_t heuristic_threads; /* host threads for CSSC_heu (0 = 8) */
  uint32_t admit_per_step; /* queries admitted per scs_step (0 = 256): the
                              search starts while later arrivals are prepared */
  uint32_t no_lds_bitmaps; /* 1 = keep vertex status in device memory only, even for
                              kept sets whose bitmaps fit in LDS (the path large kept
                              sets take; for testing it on small graphs) */
  uint32_t levels;         /* 1 = descending level floors: search for min degree
                              >= ku first, then lower bounds down to the heuristic's
                              kl (same optimum, no counterpart in the reference) */
  uint32_t pair;           /* 1 = common-neighbour pair rule: two vertices of an H
                              with min degree > kl and |H| <= h share at least
                              2(kl+1)-h neighbours in H (adjacent) or 2(kl+1)-h+2
                              (not adjacent); no counterpart in the reference */
  uint32_t deadline_ms;    /* per-query wall-clock bound from submission (0 = none): a
                              query still searching when it passes is stopped at its
                              next node and reported CAPPED with its best witness */
  uint32_t whole_graph;    /* how a query's graph reaches the search: 0 = always its own
                              compacted CSR (the reference's CompressTask/NeighborUpdate
                              form), 1 = the resident graph's CSR with the vertices
                              outside the kept set marked removed, when the kept set
                              holds at least half the graph (default), 2 = the resident
                              CSR whenever the search state fits; same optimum, node
                              counts and witness in every mode */
  uint32_t gpu_heuristic;  /* 1 = admission computes CSSC_heu (inc/heuristic.h:321-348) on
                              the GPU, one wave per query, instead of on host threads;
                              same kl and H (needs fewer than 2^32 adjacency entries) */
} scs_params;

typedef struct scs_result {
  uint64_t tag;          /* caller's tag from scs_submit */
  uint32_t l, h, q;      /* the query */
  uint32_t kl;           /* maximum minimum degree found */
  uint32_t heu_kl;       /* initial lower bound from the heuristic */
  uint32_t ku;           /* min(core[q], h-1) */
  int32_t status;        /* enum scs_status */
  uint32_t kept;         /* |K| after per-query preprocessing */
  uint64_t nodes;        /* branch-and-bound nodes expanded */
  double t_submit_us;    /* host clock at scs_submit */
  double t_done_us;      /* host clock when the result was collected */
  uint32_t wsize;        /* |H|, 0 when no witness */
  uint32_t witness[SCS_MAX_H]; /* H (original vertex ids) */
} scs_result;

/* Graph: builds the CSR on the device and computes core numbers there.
 * off: n+1 offsets (int64), adj: off[n] neighbour ids (int32), symmetric,
 * no self-loops or duplicates.  flags: reserved, pass 0. */
int scs_graph_create(const int64_t* off, const int32_t* adj, uint32_t n, uint64_t nnz,
                     uint32_t flags, scs_graph** out);
/* Reference edge-list loader semantics (src/Graph.h:5-67): header "n m",
 * whitespace-separated pairs, self-loops dropped, duplicates merged,
 * adjacency ascending; ids >= n are SCS_EINVAL. */
int scs_graph_load(const char* path, scs_graph** out);
int scs_graph_info(const scs_graph* g, uint32_t* n, uint64_t* nnz, uint32_t* dmax);
/* Copies core numbers (n values) and the CSR (off n+1, adj nnz) to the host. */
int scs_graph_core(const scs_graph* g, uint32_t* core);
int scs_graph_csr(const scs_graph* g, int64_t* off, int32_t* adj);
/* Device time of the core decomposition in microseconds. */
double scs_graph_core_time_us(const scs_graph* g);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the core numbers scs_graph_create computes, for a CSR
 * already in device memory (d_off: n+1 int64, d_adj: int32, both device
 * pointers, a simple symmetric graph), written to d_core (n words, device),
 * on `stream` (a hipStream_t; NULL = the default stream).  Working memory is
 * a fixed device array inside the library (3 * 2^20 + 64 words, 12 MB per
 * GPU, held until the process ends), so n is at most 2^20 (SCS_EBIG beyond).
 * Allocates nothing.  Every call on one GPU uses that GPU's array: calls
 * that overlap on one GPU race. */
#define SCS_CORE_NOALLOC_MAX_N (1u << 20)
int scs_core_numbers_noalloc(const int64_t* d_off, const int32_t* d_adj, uint32_t n,
                             uint32_t* d_core, void* stream);
void scs_graph_destroy(scs_graph* g);

void scs_params_default(scs_params* p);
int scs_engine_create(scs_graph* g, const scs_params* p, scs_engine** out);
/* Queues a query "<l> <h> <q> <heuristic_flag> <limit_dominating_set>".
 * Returns SCS_OK (queued), SCS_EINVAL. */
int scs_submit(scs_engine* e, uint32_t l, uint32_t h, uint32_t q, uint32_t heu,
               uint32_t limit_doms, uint64_t tag);
/* Parses the reference's query text (main.cu:92-109) and submits it. */
int scs_submit_text(scs_engine* e, const char* text, uint64_t tag);
/* Admits queued queries while slots and memory allow, runs one kernel
 * epoch and collects finished queries.  Returns the number of queries
 * still queued or in flight (>= 0) or a negative error. */
int scs_step(scs_engine* e);
/* Copies up to max finished results; returns how many were copied. */
int scs_poll(scs_engine* e, scs_result* out, int max);
/* Runs scs_step until nothing is queued or in flight. */
int scs_drain(scs_engine* e);
void scs_engine_destroy(scs_engine* e);

/* Engine counters. */
typedef struct scs_stats {
  uint64_t nodes;          /* total nodes expanded */
  uint64_t donations;      /* tasks handed to the shared pool */
  uint64_t rebuilds;       /* wave state rebuilds from pooled tasks */
  double engine_us;        /* device time inside search epochs */
  double prep_us;          /* device time inside batched preprocessing */
  double heu_us;           /* host time inside CSSC_heu */
  uint64_t prep_bytes;     /* algorithmic bytes of preprocessing */
  uint64_t epochs;
  uint32_t waves;
  uint32_t kcap;
  uint64_t edges;          /* adjacency entries scanned by the search */
  double core_us;          /* device time of the core decomposition */
  uint64_t prep_batches;   /* batched preprocessing passes (up to 64 queries each) */
  uint64_t phase_ticks[13]; /* search wall-clock ticks (100 MHz) summed over waves:
                              rebuild, rules (other), bounds, score (scatter and
                              shortlist), doms, move, donate, idle, closure scan,
                              closure removals, score argmax, score clearing,
                              pair rule */
  uint64_t dom_front;      /* frontier vertices summed over branching nodes */
  uint64_t dom_cand;       /* frontier vertices tested for domination */
  uint64_t dom_found;      /* tested vertices found dominated by ustar */
  uint64_t full_scans;     /* rule 1/2 closures that rescanned all of R */
  uint64_t full_rows;      /* R vertices visited by those rescans */
} scs_stats;
int scs_engine_stats(const scs_engine* e, scs_stats* s);

/* Host-only: the greedy lower bound of the reference's CSSC_heu
 * (inc/heuristic.h:321-348) on a CSR; writes kl and H (up to cap ids).
 * SCS_EINVAL for q >= n, h == 0, l > h, offsets that do not start at 0 or
 * decrease, or adjacency missing or holding ids outside [0, n). */
int scs_heuristic_csr(const int64_t* off, const int32_t* adj, uint32_t n, uint32_t l,
                      uint32_t h, uint32_t q, uint32_t* kl, uint32_t* H, uint32_t cap,
                      uint32_t* hsize);

/* The same greedy lower bound (inc/heuristic.h:8-348) for `count` queries on
 * the GPU, one wave per construction of each query, on a resident graph: kl[i], hsize[i] (the size
 * of H, also when it exceeds cap) and H row i (up to cap ids).  Equal to
 * scs_heuristic_csr value for value.  SCS_EBIG when the graph has 2^32
 * adjacency entries or more. */
int scs_heuristic_gpu(const scs_graph* g, uint32_t count, const uint32_t* l,
                      const uint32_t* h, const uint32_t* q, uint32_t* kl, uint32_t* hsize,
                      uint32_t* H, uint32_t cap);

/* Runs the GPU per-query preprocessing (bounded BFS, compaction, local CSR)
 * for (q, kl0, ubD) and copies the result out: K kept vertices (original
 * ids, ascending), local offsets (K+1), adjacency (local ids) and BFS
 * distances.  *K and *nnz are always set; arrays are written only when
 * they fit (kept_cap, adj_cap).  Replaces the reference's per-query
 * initialReductionRules/CompressTask/NeighborUpdate (src/helpers.cc:207-368). */
int scs_prep_probe(scs_engine* e, uint32_t q, uint32_t kl0, uint32_t ubD, uint32_t* K,
                   uint64_t* nnz, uint32_t* kept, uint32_t* loff, uint32_t* dist,
                   uint32_t kept_cap, uint32_t* ladj, uint64_t adj_cap);
/* The engine's own admission bound for one query: the CSSC_heu lower bound
 * kl (inc/heuristic.h:321-348) on the engine's graph and the diameter bound
 * ubD(kl, h) of main.cu:154-172 that limits the BFS. */
int scs_engine_heuristic(scs_engine* e, uint32_t l, uint32_t h, uint32_t q, uint32_t* kl,
                         uint32_t* ubD);
/* The same for count <= 64 queries preprocessed together in one batched
 * pass.  Query i writes K[i], nnz[i], and (when they fit) kept/dist at
 * i*kept_cap, loff at i*(kept_cap+1) and ladj at i*adj_cap. */
int scs_prep_probe_many(scs_engine* e, uint32_t count, const uint32_t* q, const uint32_t* kl0,
                        const uint32_t* ubD, uint32_t* K, uint64_t* nnz, uint32_t* kept,
                        uint32_t* loff, uint32_t* dist, uint32_t kept_cap, uint32_t* ladj,
                        uint64_t adj_cap);

/* Branching MLP (6 features -> 64 -> 64 -> 1, ReLU).  No counterpart in
 * the reference: it replaces the connection score of selectUstar
 * (src/helpers.cc:1104-1168) as the order of exploration only.  Features
 * of a candidate v: dC(v), dR(v), dist(q, v), core(v), slack =
 * dC + min(dR, h-|C|-1) - kl, and its plain connection score (sum of
 * 1/dC(u) over C neighbours u plus dR/dmax); inputs are normalised as
 * (x - mean) * scale.  FP32 master weights; the engine runs them in BF16 on
 * the MFMA units with FP32 accumulation. */
typedef struct scs_model {
  float w1[8][64]; /* rows 6 and 7 unused (zero) */
  float b1[64];
  float w2[64][64];
  float b2[64];
  float w3[64];
  float b3;
  float mean[8];
  float scale[8];
  float pad[7];
} scs_model;

int scs_model_init(scs_model* m, uint64_t seed);
/* Mean weighted binary cross-entropy gradient over n rows (X: n x 6 raw
 * features, y: 0/1 labels), computed by the GPU training kernels. */
int scs_mlp_grad(const scs_model* m, const float* X, const float* y, uint32_t n,
                 float pos_weight, scs_model* grad, float* loss);
/* Mini-batch Adam training on the GPU; sets mean/scale from the data. */
int scs_mlp_train(scs_model* m, const float* X, const float* y, uint32_t n, uint32_t epochs,
                  uint32_t batch, float lr, float pos_weight, uint64_t seed, float* final_loss);
/* Scores n rows with the engine's BF16 MFMA inference routine. */
int scs_mlp_infer(const scs_model* m, const float* X, uint32_t n, float* scores);
/* Installs (or with NULL removes) the model used by queries of this engine
 * when scs_params.learned is set. */
int scs_engine_set_model(scs_engine* e, const scs_model* m, uint32_t learned);
/* Search traces: every stride-th node of a query records up to 64 of its
 * branching candidates as rows {tag, vertex (original id), dC, dR, dist,
 * core, slack, connection score (float bits)}.  capacity_rows = 0 disables tracing. */
int scs_engine_trace(scs_engine* e, uint32_t capacity_rows, uint32_t stride);
/* Copies up to max_rows rows (8 uint32 each) and clears the trace. */
int scs_engine_trace_read(scs_engine* e, uint32_t* rows, uint32_t max_rows, uint32_t* n);

const char* scs_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif /* SCS_H_ */
