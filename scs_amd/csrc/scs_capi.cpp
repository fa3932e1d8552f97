// C ABI of the SCS engine: graph upload and core decomposition, per-query
// admission (heuristic, GPU preprocessing), search epochs and results.
// Mirrors the reference host server's responsibilities (main.cu:80-597)
// without its global state, busy loops or process exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "scs_arena.h"
#include "scs_host.h"
#include "scs_internal.h"
#include "scs_mlp.h"

void scs_pack_model(const scs_model* m, MlpDev* d);

extern "C" {
hipError_t scs_launch_core_init(const i64* off, u32 n, u32* deg, u32* core, hipStream_t s);
hipError_t scs_launch_core_scan(u32 n, u32 k, const u32* deg, u32* core, u32* F, u32* fcount,
                                u32* minrest, hipStream_t s);
hipError_t scs_launch_core_peel(const i64* off, const int32_t* adj, u32 k, const u32* F, u32 fn,
                                u32* deg, u32* core, u32* Fn, u32* fncount, hipStream_t s);
hipError_t scs_launch_ms_init(u32 n, const u32* core, const MsBatch* B, u64* elig, u64* vis,
                              u64* fr, u64* nf, u32* flags, u32 nflags, unsigned long long* bytes,
                              hipStream_t s);
hipError_t scs_launch_ms_push0(const i64* off, const int32_t* adj, u32 nb, const MsBatch* B,
                               u64 act, const u64* elig, u64* vis, u64* nf, u16* dist16,
                               u32* flag0, unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_seed(const MsBatch* B, u64* vis, u64* fr, u16* dist16, hipStream_t s);
hipError_t scs_launch_ms_level(const i64* off, const int32_t* adj, u32 n, u32 level, u64 act,
                               const u64* elig, u64* vis, const u64* fr, u64* nf, u16* dist16,
                               const u32* flag_prev, u32* flag_cur, unsigned long long* bytes,
                               hipStream_t s);
hipError_t scs_launch_ms_count(u32 n, u32 G, const u64* vis, u64 valid, u32* gcnt,
                               unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_bscan(u32 nb, u32 G, u32* gcnt, u32* tot, unsigned long long* bytes,
                               hipStream_t s);
hipError_t scs_launch_ms_gdist(u32 n, u32 G, const u64* vis, u64 gmask, const u16* dist16,
                               const MsBatch* B, unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_scatter(u32 n, u32 G, const u64* vis, u64 valid, const u32* gcnt,
                                 const u32* core, const u16* dist16, const MsBatch* B, u32* locid,
                                 u32* qloc, unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_degree(const int32_t* adj, const MsUnit* units, u32 nunits,
                                const MsHub* hubs, u32 nhubs, const u64* vis, u64 valid,
                                const u32* locid, const MsBatch* B, u32* hubcnt,
                                unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_scan_loff(u32 nb, const MsBatch* B, u32* tot, u64 valid, u32 cmax,
                                   u32* csum, unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_ms_fill(const int32_t* adj, const MsUnit* units, u32 nunits,
                              const u64* vis, u64 valid, const u32* locid, const MsBatch* B,
                              const u32* hubcnt, unsigned long long* bytes, hipStream_t s);
hipError_t scs_launch_epoch(const EngineArgs* A, hipStream_t s);
hipError_t scs_launch_slot_write(const QuerySlot* stage, const u32* idx, const u32* wit, u32 n,
                                 QuerySlot* slots, u32* witness, hipStream_t s);
hipError_t scs_launch_expire(QuerySlot* slots, const u32* idx, u32 n, hipStream_t s);
hipError_t scs_launch_admit(const EngineArgs* A, const u32* slots, u32 n, u32* failed,
                            hipStream_t s);
hipError_t scs_launch_pool_init(u64* seq, u32 n, hipStream_t s);
HeuLayout scs_heu_layout(u32 n, u32 hcap);
hipError_t scs_launch_heu(const u32* off, const u32* adj, u32 n, u32 dmax, u32 nq, const u32* ql,
                          const u32* qh, const u32* qq, u8* ws, HeuLayout lay, u32 hcap, u32* kl,
                          u32* hsize, u32* H, u32* kl3, u32* hs3, u32* H3, hipStream_t s);
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "scs: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
              __LINE__);                                                            \
      return SCS_EHIP;                                                              \
    }                                                                               \
  } while (0)

static double now_us() {
  using namespace std::chrono;
  return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Graph
// ---------------------------------------------------------------------------
struct scs_graph {
  scs::HostGraph hg;
  i64* d_off = nullptr;
  int32_t* d_adj = nullptr;
  u32* d_core = nullptr;
  // the search's view of the whole graph (queries that keep most of it
  // search it in place): 32-bit offsets and the identity id map; null when
  // the adjacency has 2^32 entries or more
  u32* d_off32 = nullptr;
  u32* d_ident = nullptr;
  double core_us = 0;
};

static int core_decompose(scs_graph* G) {
  const u32 n = G->hg.n;
  hipStream_t s;
  HIPCHK(hipStreamCreate(&s));
  u32 *deg, *F, *Fn, *cnt;
  HIPCHK(hipMalloc(&deg, sizeof(u32) * std::max<u32>(n, 1)));
  HIPCHK(hipMalloc(&F, sizeof(u32) * std::max<u32>(n, 1)));
  HIPCHK(hipMalloc(&Fn, sizeof(u32) * std::max<u32>(n, 1)));
  HIPCHK(hipMalloc(&cnt, sizeof(u32) * 4));
  u32* h_cnt;
  HIPCHK(hipHostMalloc(&h_cnt, sizeof(u32) * 4));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  HIPCHK(scs_launch_core_init(G->d_off, n, deg, G->d_core, s));
  u32 peeled = 0, k = 0;
  while (peeled < n) {
    // cnt[0] = frontier size, cnt[1] = min remaining degree
    u32 init[2] = {0, 0xffffffffu};
    HIPCHK(hipMemcpyAsync(cnt, init, sizeof(init), hipMemcpyHostToDevice, s));
    HIPCHK(scs_launch_core_scan(n, k, deg, G->d_core, F, cnt, cnt + 1, s));
    HIPCHK(hipMemcpyAsync(h_cnt, cnt, 2 * sizeof(u32), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    u32 fn = h_cnt[0], minrest = h_cnt[1];
    if (fn == 0) {
      if (minrest == 0xffffffffu) break;
      k = std::max(k + 1, minrest);
      continue;
    }
    peeled += fn;
    u32 *cur = F, *nxt = Fn;
    while (fn) {
      HIPCHK(hipMemsetAsync(cnt + 2, 0, sizeof(u32), s));
      HIPCHK(scs_launch_core_peel(G->d_off, G->d_adj, k, cur, fn, deg, G->d_core, nxt, cnt + 2, s));
      HIPCHK(hipMemcpyAsync(h_cnt + 2, cnt + 2, sizeof(u32), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      fn = h_cnt[2];
      peeled += fn;
      std::swap(cur, nxt);
    }
    k++;
  }
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  G->core_us = ms * 1000.0;
  G->hg.core.resize(n);
  HIPCHK(hipMemcpy(G->hg.core.data(), G->d_core, sizeof(u32) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(deg));
  HIPCHK(hipFree(F));
  HIPCHK(hipFree(Fn));
  HIPCHK(hipFree(cnt));
  HIPCHK(hipHostFree(h_cnt));
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipStreamDestroy(s));
  return SCS_OK;
}

static int graph_upload(scs_graph* G) {
  const u32 n = G->hg.n;
  const u64 nnz = (u64)G->hg.off[n];
  HIPCHK(hipMalloc(&G->d_off, sizeof(i64) * (n + 1)));
  HIPCHK(hipMalloc(&G->d_adj, sizeof(int32_t) * std::max<u64>(nnz, 1)));
  HIPCHK(hipMalloc(&G->d_core, sizeof(u32) * std::max<u32>(n, 1)));
  HIPCHK(hipMemcpy(G->d_off, G->hg.off.data(), sizeof(i64) * (n + 1), hipMemcpyHostToDevice));
  if (nnz)
    HIPCHK(hipMemcpy(G->d_adj, G->hg.adj.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  G->hg.dmax = 0;
  for (u32 v = 0; v < n; v++) G->hg.dmax = std::max<u32>(G->hg.dmax, (u32)(G->hg.off[v + 1] - G->hg.off[v]));
  if (nnz <= 0xffffffffull) {
    std::vector<u32> o32(n + 1), id(std::max<u32>(n, 1));
    for (u32 v = 0; v <= n; v++) o32[v] = (u32)G->hg.off[v];
    for (u32 v = 0; v < n; v++) id[v] = v;
    HIPCHK(hipMalloc(&G->d_off32, sizeof(u32) * (n + 1)));
    HIPCHK(hipMalloc(&G->d_ident, sizeof(u32) * id.size()));
    HIPCHK(hipMemcpy(G->d_off32, o32.data(), sizeof(u32) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(G->d_ident, id.data(), sizeof(u32) * id.size(), hipMemcpyHostToDevice));
  }
  return core_decompose(G);
}

extern "C" int scs_graph_create(const int64_t* off, const int32_t* adj, uint32_t n, uint64_t nnz,
                                uint32_t flags, scs_graph** out) {
  (void)flags;
  if (!off || !out || (nnz && !adj)) return SCS_EINVAL;
  if (off[0] != 0 || (u64)off[n] != nnz) return SCS_EINVAL;
  for (u32 v = 0; v < n; v++)
    if (off[v + 1] < off[v]) return SCS_EINVAL;
  for (u64 j = 0; j < nnz; j++)
    if (adj[j] < 0 || (u32)adj[j] >= n) return SCS_EINVAL;
  // The search kernels need a simple undirected graph: every row strictly
  // ascending (sorted, no duplicates), no self-loops, and w in row v exactly
  // when v is in row w.  Rows are visited in ascending v, so the entries
  // below w in row w must match, in order, the rows v < w that name w.
  {
    std::vector<i64> ptr(off, off + n);
    for (u32 v = 0; v < n; v++) {
      for (i64 j = off[v]; j < off[v + 1]; j++) {
        const u32 w = (u32)adj[j];
        if (w == v) return SCS_EINVAL;
        if (j > off[v] && adj[j - 1] >= adj[j]) return SCS_EINVAL;
        if (w > v) {
          if (ptr[w] >= off[w + 1] || (u32)adj[ptr[w]] != v) return SCS_EINVAL;
          ptr[w]++;
        }
      }
      // every entry of row v below v must have been matched by now
      if (ptr[v] < off[v + 1] && (u32)adj[ptr[v]] < v) return SCS_EINVAL;
    }
  }
  scs_graph* G = new scs_graph;
  G->hg.n = n;
  G->hg.off.assign(off, off + n + 1);
  G->hg.adj.assign(adj, adj + nnz);
  int rc = graph_upload(G);
  if (rc) {
    scs_graph_destroy(G);
    return rc;
  }
  *out = G;
  return SCS_OK;
}

// Reference loader semantics (src/Graph.h:5-67).
extern "C" int scs_graph_load(const char* path, scs_graph** out) {
  if (!path || !out) return SCS_EINVAL;
  FILE* f = fopen(path, "rb");
  if (!f) return SCS_EIO;
  std::string buf;
  {
    char tmp[1 << 16];
    size_t r;
    while ((r = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.append(tmp, r);
  }
  fclose(f);
  const char* p = buf.c_str();
  const char* end = p + buf.size();
  auto next = [&](u64& x) -> bool {
    while (p < end && (*p < '0' || *p > '9')) {
      if (*p == '-') return false;
      p++;
    }
    if (p >= end) return false;
    x = 0;
    while (p < end && *p >= '0' && *p <= '9') x = x * 10 + (u64)(*p++ - '0');
    return true;
  };
  u64 n, m;
  if (!next(n) || !next(m) || n > 0xfffffffeull) return SCS_EIO;
  std::vector<u64> keys;
  keys.reserve(2 * m + 16);
  u64 a, b;
  while (next(a)) {
    if (!next(b)) break;
    if (a >= n || b >= n) return SCS_EINVAL;
    if (a == b) continue;
    keys.push_back((a << 32) | b);
    keys.push_back((b << 32) | a);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  scs_graph* G = new scs_graph;
  G->hg.n = (u32)n;
  G->hg.off.assign(n + 1, 0);
  G->hg.adj.resize(keys.size());
  for (size_t i = 0; i < keys.size(); i++) {
    G->hg.off[(keys[i] >> 32) + 1]++;
    G->hg.adj[i] = (int32_t)(keys[i] & 0xffffffffu);
  }
  for (u64 i = 0; i < n; i++) G->hg.off[i + 1] += G->hg.off[i];
  int rc = graph_upload(G);
  if (rc) {
    scs_graph_destroy(G);
    return rc;
  }
  *out = G;
  return SCS_OK;
}

extern "C" int scs_graph_info(const scs_graph* g, uint32_t* n, uint64_t* nnz, uint32_t* dmax) {
  if (!g) return SCS_EINVAL;
  if (n) *n = g->hg.n;
  if (nnz) *nnz = (u64)g->hg.off[g->hg.n];
  if (dmax) *dmax = g->hg.dmax;
  return SCS_OK;
}
extern "C" int scs_graph_core(const scs_graph* g, uint32_t* core) {
  if (!g || !core) return SCS_EINVAL;
  std::copy(g->hg.core.begin(), g->hg.core.end(), core);
  return SCS_OK;
}
extern "C" int scs_graph_csr(const scs_graph* g, int64_t* off, int32_t* adj) {
  if (!g) return SCS_EINVAL;
  if (off) std::copy(g->hg.off.begin(), g->hg.off.end(), off);
  if (adj) std::copy(g->hg.adj.begin(), g->hg.adj.end(), adj);
  return SCS_OK;
}
extern "C" double scs_graph_core_time_us(const scs_graph* g) { return g ? g->core_us : 0.0; }
extern "C" void scs_graph_destroy(scs_graph* g) {
  if (!g) return;
  if (g->d_off) (void)hipFree(g->d_off);
  if (g->d_adj) (void)hipFree(g->d_adj);
  if (g->d_core) (void)hipFree(g->d_core);
  if (g->d_off32) (void)hipFree(g->d_off32);
  if (g->d_ident) (void)hipFree(g->d_ident);
  delete g;
}

// ---------------------------------------------------------------------------
// GPU greedy heuristic: up to `waves` queries per launch, three waves each
// (one per construction), on the graph's 32-bit CSR; the workspace's n-sized
// arrays start zeroed and every wave leaves its own zeroed again.
// ---------------------------------------------------------------------------
struct HeuGpu {
  u8* ws = nullptr;
  HeuLayout lay{};
  u32 waves = 0;
  u32* d_in = nullptr;   // l, h, q (3 * waves)
  u32* d_out = nullptr;  // kl, hsize (2 * waves)
  u32* d_H = nullptr;    // waves * SCS_MAX_H
  u32* d_out3 = nullptr;  // the three constructions: kl, hsize (2 * 3 * waves)
  u32* d_H3 = nullptr;    // 3 * waves * SCS_MAX_H
  std::vector<u32> h_out, h_H;

  int init(u32 n, u32 w) {
    waves = w;
    lay = scs_heu_layout(n, SCS_MAX_H);
    HIPCHK(hipMalloc(&ws, lay.stride * 3 * waves));
    HIPCHK(hipMemset(ws, 0, lay.stride * 3 * waves));
    HIPCHK(hipMalloc(&d_in, sizeof(u32) * 3 * waves));
    HIPCHK(hipMalloc(&d_out, sizeof(u32) * 2 * waves));
    HIPCHK(hipMalloc(&d_H, sizeof(u32) * (u64)SCS_MAX_H * waves));
    HIPCHK(hipMalloc(&d_out3, sizeof(u32) * 6 * waves));
    HIPCHK(hipMalloc(&d_H3, sizeof(u32) * (u64)SCS_MAX_H * 3 * waves));
    h_out.resize(2 * waves);
    h_H.resize((size_t)SCS_MAX_H * waves);
    return SCS_OK;
  }
  void release() {
    for (void* p : {(void*)ws, (void*)d_in, (void*)d_out, (void*)d_H, (void*)d_out3, (void*)d_H3})
      if (p) (void)hipFree(p);
    ws = nullptr;
    d_in = d_out = d_H = d_out3 = d_H3 = nullptr;
  }
  // queries i0..i0+cnt (cnt <= waves): kl, and H as vectors
  int run(const scs_graph* g, u32 cnt, const u32* l, const u32* h, const u32* q, u32* kl,
          std::vector<u32>* H, hipStream_t s) {
    std::vector<u32> in(3 * (size_t)cnt);
    for (u32 i = 0; i < cnt; i++) {
      in[i] = l[i];
      in[cnt + i] = h[i];
      in[2 * cnt + i] = q[i];
    }
    HIPCHK(hipMemcpyAsync(d_in, in.data(), sizeof(u32) * 3 * cnt, hipMemcpyHostToDevice, s));
    HIPCHK(scs_launch_heu(g->d_off32, (const u32*)g->d_adj, g->hg.n, g->hg.dmax, cnt, d_in,
                          d_in + cnt, d_in + 2 * cnt, ws, lay, SCS_MAX_H, d_out, d_out + cnt, d_H,
                          d_out3, d_out3 + 3 * cnt, d_H3, s));
    HIPCHK(hipMemcpyAsync(h_out.data(), d_out, sizeof(u32) * 2 * cnt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(h_H.data(), d_H, sizeof(u32) * (u64)SCS_MAX_H * cnt,
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (u32 i = 0; i < cnt; i++) {
      kl[i] = h_out[i];
      const u32 sz = std::min<u32>(h_out[cnt + i], SCS_MAX_H);
      H[i].assign(h_H.begin() + (size_t)i * SCS_MAX_H, h_H.begin() + (size_t)i * SCS_MAX_H + sz);
    }
    return SCS_OK;
  }
};

extern "C" int scs_heuristic_gpu(const scs_graph* g, uint32_t count, const uint32_t* l,
                                 const uint32_t* h, const uint32_t* q, uint32_t* kl,
                                 uint32_t* hsize, uint32_t* H, uint32_t cap) {
  if (!g || !l || !h || !q || !kl || !hsize || (cap && !H)) return SCS_EINVAL;
  if (!g->d_off32) return SCS_EBIG;
  for (u32 i = 0; i < count; i++)
    if (q[i] >= g->hg.n || h[i] == 0 || h[i] > SCS_MAX_H || l[i] > h[i]) return SCS_EINVAL;
  if (count == 0) return SCS_OK;
  HeuGpu hg;
  int rc = hg.init(g->hg.n, std::min<u32>(count, 64));
  hipStream_t s = nullptr;
  if (rc == SCS_OK && hipStreamCreate(&s) != hipSuccess) rc = SCS_EHIP;
  std::vector<std::vector<u32>> HH(hg.waves);
  for (u32 i0 = 0; i0 < count && rc == SCS_OK; i0 += hg.waves) {
    const u32 cnt = std::min<u32>(hg.waves, count - i0);
    rc = hg.run(g, cnt, l + i0, h + i0, q + i0, kl + i0, HH.data(), s);
    for (u32 i = 0; i < cnt && rc == SCS_OK; i++) {
      hsize[i0 + i] = hg.h_out[cnt + i];
      for (u32 k = 0; k < HH[i].size() && k < cap; k++) H[(u64)(i0 + i) * cap + k] = HH[i][k];
    }
  }
  if (s) (void)hipStreamDestroy(s);
  hg.release();
  return rc;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
struct PendingQuery {
  uint64_t tag;
  u32 l, h, q, heu, ld;
  double t_submit;
};

struct HostSlot {
  bool used = false;
  PendingQuery pq;
  std::vector<u32> seed_witness;  // heuristic H, copied into the slot at admission
  u32 kl0 = 0, ku = 0, K = 0;
  u64 a_off = 0, a_len = 0, b_off = 0, b_len = 0;
  u64 need = 0;  // arena bytes known to be needed when admission was refused
  bool whole = false;     // searches the resident graph's CSR (no local graph)
  bool admitted = false;  // root task pushed: the query is on the device
  bool expired = false;   // stopped at its deadline
};

struct scs_engine {
  scs_graph* g = nullptr;
  scs_params p{};
  hipStream_t s = nullptr;      // search epochs, slot writes, admission
  hipStream_t ps = nullptr;     // per-query preprocessing (overlaps epochs)
  u32 nwaves = 0, kcap = 0;
  // device
  QuerySlot* d_slots = nullptr;
  u32* d_witness = nullptr;
  u8* d_wave = nullptr;
  u64* d_pool_seq = nullptr;
  u32* d_pool_data = nullptr;
  u64* d_pool_pos = nullptr;
  EngineGlobals* d_glob = nullptr;
  EngineArgs args{};
  // batched preprocessing (bit b of a mask = query b of the batch)
  u64 *d_elig = nullptr, *d_vis = nullptr, *d_fr = nullptr, *d_nf = nullptr;
  u16* d_dist16 = nullptr;   // [n][64] BFS distance per query
  u32* d_locid64 = nullptr;  // [n][64] local id per query
  u32* d_gcnt = nullptr;     // [ceil(n/64)][64] group counts, then bases
  MsUnit* d_units = nullptr;  // adjacency units of every vertex
  MsHub* d_hubs = nullptr;    // vertices with more than one unit
  u32 nunits = 0, nhubs = 0;
  u32* d_hubcnt = nullptr;   // [hub units][64] partial counts, then unit offsets
  u32* d_csum = nullptr;     // [64][cmax] loff chunk sums, then chunk bases
  u32* d_flags = nullptr;    // one per BFS level: something was found
  u32* d_tot = nullptr;      // K[64] | nnz[64] | qloc[64]
  MsBatch* d_batch = nullptr;
  MsBatch* h_batch = nullptr;  // pinned
  u32* d_admit = nullptr;
  unsigned long long* d_prepbytes = nullptr;
  u8* d_arena = nullptr;
  MlpDev* d_model = nullptr;
  bool have_model = false;
  u32* d_trace = nullptr;
  u32* d_trace_cnt = nullptr;
  u32 trace_cap = 0;
  scs::Arena arena;
  u32* h_tot = nullptr;  // pinned, 192 words
  // slot hand-over between epochs: admitted slots and their witness seeds
  // are staged in pinned memory and scattered by one kernel; finished
  // slots' witnesses come back in one copy
  QuerySlot* h_stage = nullptr;  // pinned, limit_queries slots
  u32* h_stage_wit = nullptr;    // pinned, limit_queries x SCS_MAX_H
  u32* h_stage_idx = nullptr;    // pinned, 2 x limit_queries (slot, witness words)
  u32* h_wit = nullptr;          // pinned, limit_queries x SCS_MAX_H
  QuerySlot* d_stage = nullptr;
  u32* d_stage_wit = nullptr;
  u32* d_stage_idx = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, pe0 = nullptr, pe1 = nullptr;
  // host
  std::mutex qmu;
  std::deque<PendingQuery> queue;
  std::vector<HostSlot> slots;
  std::vector<QuerySlot> h_slots;
  std::deque<scs_result> results;
  std::vector<scs::HeuWork> heu_work;
  HeuGpu heu_gpu;  // used when p.gpu_heuristic (waves == 0 otherwise)
  u32 active = 0;
  scs_stats st{};
};

extern "C" void scs_params_default(scs_params* p) {
  memset(p, 0, sizeof(*p));
  p->partition_size = 0;
  p->buffer_size = 0;
  p->copy_limit = 0.5;
  p->read_limit = 1;
  p->limit_queries = 64;
  p->factor = 1;
  p->red1 = p->red2 = p->red3 = 1;
  p->prun1 = p->prun2 = 1;
  p->waves_per_cu = 4;
  p->limit_tasks = 0;
  p->node_budget = 0;
  p->mem_budget = 1ull << 30;
  p->max_kept = 0;
  p->time_slice_us = 200000;
  p->deterministic = 0;
  p->learned = 0;
  p->model = nullptr;
  p->heuristic_threads = 8;
  p->admit_per_step = 256;
  p->levels = 1;
  p->pair = 1;
  p->whole_graph = 1;
}

static u64 align256(u64 x) { return (x + 255) & ~255ull; }

static WaveLayout make_layout(u32 kcap) {
  WaveLayout L{};
  u64 o = 0;
  const u64 k4 = (u64)((kcap + 3) & ~3u);
  L.vs = o; o = align256(o + sizeof(VState) * k4);
  L.nmark = o; o = align256(o + k4);
  L.pos = o; o = align256(o + 4 * k4);
  L.rlist = o; o = align256(o + 4 * k4);
  L.work = o; o = align256(o + 4 * (k4 + 64));
  L.work2 = o; o = align256(o + 4 * (k4 + 64));
  L.front = o; o = align256(o + 4 * (k4 + 64));
  L.aux = o; o = align256(o + 4 * (k4 + 64));
  // each vertex changes once per path; each removal batch adds one marker
  L.trail = o; o = align256(o + 4 * (2 * k4 + 64));
  L.sav = o; o = align256(o + 4 * (2 * k4 + 64));
  L.frames = o; o = align256(o + sizeof(Frame) * (SCS_HCAP + 1));
  L.clist = o; o = align256(o + 4 * (SCS_HCAP + 64));
  L.hist = o; o = align256(o + 4 * (SCS_HCAP + 8));
  L.hdr = o; o = align256(o + sizeof(WaveHdr));
  L.gbits = o; o = align256(o + 8 * ((k4 + 63) / 32));
  L.stride = o;
  return L;
}

extern "C" int scs_engine_create(scs_graph* g, const scs_params* pin, scs_engine** out) {
  if (!g || !out) return SCS_EINVAL;
  scs_params p;
  if (pin)
    p = *pin;
  else
    scs_params_default(&p);
  if (p.limit_queries == 0) p.limit_queries = 1;
  if (p.waves_per_cu == 0) p.waves_per_cu = 4;
  // the search kernel is built for at least 2 waves per SIMD (up to 256
  // registers): 8 per CU are resident, and every wave of the persistent
  // grid must be, or the states of the others would never run
  if (p.waves_per_cu > 8) p.waves_per_cu = 8;
  if (p.time_slice_us == 0) p.time_slice_us = 200000;
  if (p.heuristic_threads == 0) p.heuristic_threads = 8;
  if (p.admit_per_step == 0) p.admit_per_step = 256;
  scs_engine* E = new scs_engine;
  E->g = g;
  E->p = p;
  auto fail = [&](int rc) {
    scs_engine_destroy(E);
    return rc;
  };
  if (hipStreamCreate(&E->s) != hipSuccess) return fail(SCS_EHIP);
  if (hipStreamCreate(&E->ps) != hipSuccess) return fail(SCS_EHIP);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(SCS_EHIP);
  const u32 n = g->hg.n;
  E->nwaves = p.deterministic ? 1u : (u32)prop.multiProcessorCount * p.waves_per_cu;
  E->kcap = p.max_kept ? std::min(p.max_kept, n) : n;
  if (E->kcap == 0) E->kcap = 1;
  if (p.partition_size && p.partition_size < E->kcap) E->kcap = std::max<u32>(p.partition_size, 1);
  if (!p.max_kept && !p.partition_size) {
    // per-wave depth-first state grows with kcap: keep all waves' state
    // within 60% of the device memory left after the arena and the batched
    // preprocessing buffers; larger kept sets are rejected per query
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return fail(SCS_EHIP);
    const u64 fixed = p.mem_budget + 420ull * n + (64ull << 20);
    const u64 room = freeb > fixed ? (u64)((freeb - fixed) * 0.6) : 0;
    const u64 pool_slots = std::max<u64>(p.limit_tasks ? p.limit_tasks : 2ull * E->nwaves,
                                         (u64)E->nwaves + p.limit_queries + 64);
    auto need = [&](u32 kc) {
      return (u64)E->nwaves * make_layout(kc).stride + pool_slots * 4ull * ((kc + 1) / 2 + 1100);
    };
    while (E->kcap > 4096 && need(E->kcap) > room) E->kcap = E->kcap / 4 * 3;
  }
  const WaveLayout lay = make_layout(E->kcap);
  // pool
  u32 donate_below = p.deterministic ? 0u : (u32)std::max(1.0, p.copy_limit * E->nwaves);
  u32 nslots = p.limit_tasks ? p.limit_tasks : 2 * E->nwaves;
  nslots = std::max<u32>(nslots, donate_below + p.limit_queries + 64);
  u32 slot_words = 4 + SCS_HCAP + (E->kcap + 1) / 2 + 64;
  if (p.buffer_size) nslots = std::max<u32>(p.limit_queries + 64, p.buffer_size / slot_words);
  // limit_queries slots stay reserved for root tasks: a donation is refused
  // once donor_cap tasks are pooled, so a root push never finds the pool full
  const u32 donor_cap = nslots - p.limit_queries;
  donate_below = std::min(donate_below, donor_cap);
  const u32 nq = p.limit_queries;
  // allocations
  if (hipMalloc(&E->d_slots, sizeof(QuerySlot) * nq) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_witness, sizeof(u32) * SCS_MAX_H * nq) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_wave, lay.stride * E->nwaves) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_pool_seq, sizeof(u64) * nslots) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_pool_data, sizeof(u32) * (u64)slot_words * nslots) != hipSuccess)
    return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_pool_pos, sizeof(u64) * 2) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_glob, sizeof(EngineGlobals)) != hipSuccess) return fail(SCS_ENOMEM);
  const u32 n1 = std::max<u32>(n, 1);
  const u64 G = (n1 + 63) / 64;
  for (u64** b : {&E->d_elig, &E->d_vis, &E->d_fr, &E->d_nf})
    if (hipMalloc(b, sizeof(u64) * n1) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_dist16, sizeof(u16) * 64 * (u64)n1) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_locid64, sizeof(u32) * 64 * (u64)n1) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_gcnt, sizeof(u32) * 64 * G) != hipSuccess) return fail(SCS_ENOMEM);
  {
    // adjacency units: one per vertex, or ceil(deg / MS_UNIT) for a hub
    std::vector<MsUnit> units;
    std::vector<MsHub> hubs;
    units.reserve(n1);
    u32 hub_units = 0;
    for (u32 v = 0; v < n; v++) {
      const u64 deg = (u64)(E->g->hg.off[v + 1] - E->g->hg.off[v]);
      const u64 nu = (deg + MS_UNIT - 1) / MS_UNIT;
      const u64 b0 = (u64)E->g->hg.off[v];
      if (nu <= 1) {
        units.push_back(MsUnit{v, MS_NOHUB, b0 | (deg << 48)});
        continue;
      }
      hubs.push_back(MsHub{v, hub_units, (u32)nu, 0});
      for (u32 si = 0; si < nu; si++) {
        const u64 ulen = std::min<u64>(MS_UNIT, deg - (u64)si * MS_UNIT);
        units.push_back(MsUnit{v, hub_units + si, (b0 + (u64)si * MS_UNIT) | (ulen << 48)});
      }
      hub_units += (u32)nu;
    }
    E->nunits = (u32)units.size();
    E->nhubs = (u32)hubs.size();
    if (hipMalloc(&E->d_units, sizeof(MsUnit) * std::max<size_t>(units.size(), 1)) != hipSuccess ||
        hipMalloc(&E->d_hubs, sizeof(MsHub) * std::max<size_t>(hubs.size(), 1)) != hipSuccess ||
        hipMalloc(&E->d_hubcnt, sizeof(u32) * 64 * std::max<u64>(hub_units, 1)) != hipSuccess ||
        hipMalloc(&E->d_csum, sizeof(u32) * 64 * ((u64)E->kcap / 4096 + 2)) != hipSuccess)
      return fail(SCS_ENOMEM);
    if (!units.empty() && hipMemcpy(E->d_units, units.data(), sizeof(MsUnit) * units.size(),
                                    hipMemcpyHostToDevice) != hipSuccess)
      return fail(SCS_EHIP);
    if (!hubs.empty() && hipMemcpy(E->d_hubs, hubs.data(), sizeof(MsHub) * hubs.size(),
                                   hipMemcpyHostToDevice) != hipSuccess)
      return fail(SCS_EHIP);
  }
  if (hipMalloc(&E->d_flags, sizeof(u32) * SCS_MS_LEVELS) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_tot, sizeof(u32) * 192) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_batch, sizeof(MsBatch)) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipHostMalloc(&E->h_batch, sizeof(MsBatch)) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_admit, sizeof(u32) * 2 * nq) != hipSuccess) return fail(SCS_ENOMEM);
  const size_t prepbytes_len = sizeof(unsigned long long) * MS_BYTE_SLOTS * MS_BYTE_STRIDE;
  if (hipMalloc(&E->d_prepbytes, prepbytes_len) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMalloc(&E->d_arena, p.mem_budget) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipHostMalloc(&E->h_tot, sizeof(u32) * 192) != hipSuccess) return fail(SCS_ENOMEM);
  {
    const u64 wit_words = (u64)nq * SCS_MAX_H;
    if (hipHostMalloc(&E->h_stage, sizeof(QuerySlot) * nq) != hipSuccess ||
        hipHostMalloc(&E->h_stage_wit, sizeof(u32) * wit_words) != hipSuccess ||
        hipHostMalloc(&E->h_stage_idx, sizeof(u32) * 2 * nq) != hipSuccess ||
        hipHostMalloc(&E->h_wit, sizeof(u32) * wit_words) != hipSuccess ||
        hipMalloc(&E->d_stage, sizeof(QuerySlot) * nq) != hipSuccess ||
        hipMalloc(&E->d_stage_wit, sizeof(u32) * wit_words) != hipSuccess ||
        hipMalloc(&E->d_stage_idx, sizeof(u32) * 2 * nq) != hipSuccess)
      return fail(SCS_ENOMEM);
  }
  E->arena.init(E->d_arena, p.mem_budget);
  if (hipMemset(E->d_slots, 0, sizeof(QuerySlot) * nq) != hipSuccess) return fail(SCS_EHIP);
  if (hipMemset(E->d_wave, 0, lay.stride * E->nwaves) != hipSuccess) return fail(SCS_EHIP);
  if (hipMemset(E->d_pool_pos, 0, sizeof(u64) * 2) != hipSuccess) return fail(SCS_EHIP);
  if (hipMemset(E->d_glob, 0, sizeof(EngineGlobals)) != hipSuccess) return fail(SCS_EHIP);
  if (hipMemset(E->d_prepbytes, 0, prepbytes_len) != hipSuccess) return fail(SCS_EHIP);
  if (scs_launch_pool_init(E->d_pool_seq, nslots, E->s) != hipSuccess) return fail(SCS_EHIP);
  if (hipStreamSynchronize(E->s) != hipSuccess) return fail(SCS_EHIP);
  for (hipEvent_t* ev : {&E->ev0, &E->ev1, &E->pe0, &E->pe1})
    if (hipEventCreate(ev) != hipSuccess) return fail(SCS_EHIP);
  EngineArgs& A = E->args;
  A.slots = E->d_slots;
  A.nslots = nq;
  A.wave_mem = E->d_wave;
  A.lay = lay;
  A.kcap = E->kcap;
  A.nwaves = E->nwaves;
  A.pool.seq = E->d_pool_seq;
  A.pool.data = E->d_pool_data;
  A.pool.enq = E->d_pool_pos;
  A.pool.deq = E->d_pool_pos + 1;
  A.pool.nslots = nslots;
  A.pool.slot_words = slot_words;
  A.glob = E->d_glob;
  A.slice_ticks = (u64)p.time_slice_us * 100ull;  // wall_clock64 runs at 100 MHz
  A.donate_below = donate_below;
  A.donor_cap = donor_cap;
  A.deterministic = p.deterministic;
  A.node_flush = p.deterministic ? 1 : 16;
  A.waves_per_cu = p.deterministic ? 1 : p.waves_per_cu;
  A.no_bitmaps = p.no_lds_bitmaps ? 1u : 0u;
  E->slots.resize(nq);
  E->h_slots.resize(nq);
  E->heu_work.resize(p.heuristic_threads);
  if (p.gpu_heuristic && g->d_off32) {
    // as many queries per launch as an admission step takes, within 16 GiB
    // of workspace (a launch lasts as long as its slowest query)
    const u64 stride = 3 * scs_heu_layout(g->hg.n, SCS_MAX_H).stride;
    const u32 w = (u32)std::max<u64>(16, std::min<u64>(std::max<u32>(p.admit_per_step, 64),
                                                       (16ull << 30) / std::max<u64>(stride, 1)));
    if (E->heu_gpu.init(g->hg.n, w) != SCS_OK) return fail(SCS_ENOMEM);
  }
  E->st.waves = E->nwaves;
  E->st.kcap = E->kcap;
  if (hipMalloc(&E->d_model, sizeof(MlpDev)) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMemset(E->d_model, 0, sizeof(MlpDev)) != hipSuccess) return fail(SCS_EHIP);
  if (hipMalloc(&E->d_trace_cnt, sizeof(u32)) != hipSuccess) return fail(SCS_ENOMEM);
  if (hipMemset(E->d_trace_cnt, 0, sizeof(u32)) != hipSuccess) return fail(SCS_EHIP);
  A.model = nullptr;
  A.trace = nullptr;
  A.trace_cnt = E->d_trace_cnt;
  A.trace_cap = 0;
  A.trace_stride = 1;
  if (p.learned && p.model) {
    int rc = scs_engine_set_model(E, (const scs_model*)p.model, 1);
    if (rc) return fail(rc);
  }
  *out = E;
  return SCS_OK;
}

extern "C" int scs_engine_set_model(scs_engine* E, const scs_model* m, uint32_t learned) {
  if (!E) return SCS_EINVAL;
  if (m) {
    MlpDev h;
    scs_pack_model(m, &h);
    HIPCHK(hipMemcpy(E->d_model, &h, sizeof(MlpDev), hipMemcpyHostToDevice));
    E->have_model = true;
  }
  E->p.learned = learned && E->have_model ? 1u : 0u;
  E->args.model = E->p.learned ? E->d_model : nullptr;
  return SCS_OK;
}

extern "C" int scs_engine_trace(scs_engine* E, uint32_t capacity_rows, uint32_t stride) {
  if (!E) return SCS_EINVAL;
  if (E->d_trace) {
    HIPCHK(hipFree(E->d_trace));
    E->d_trace = nullptr;
  }
  E->trace_cap = 0;
  if (capacity_rows) {
    HIPCHK(hipMalloc(&E->d_trace, sizeof(u32) * 8ull * capacity_rows));
    E->trace_cap = capacity_rows;
  }
  HIPCHK(hipMemset(E->d_trace_cnt, 0, sizeof(u32)));
  E->args.trace = E->d_trace;
  E->args.trace_cap = E->trace_cap;
  E->args.trace_stride = stride ? stride : 1;
  return SCS_OK;
}

extern "C" int scs_engine_trace_read(scs_engine* E, uint32_t* rows, uint32_t max_rows, uint32_t* n) {
  if (!E || !n) return SCS_EINVAL;
  u32 cnt = 0;
  HIPCHK(hipMemcpy(&cnt, E->d_trace_cnt, sizeof(u32), hipMemcpyDeviceToHost));
  cnt = std::min(cnt, E->trace_cap);
  const u32 k = std::min(cnt, max_rows);
  if (k && rows) HIPCHK(hipMemcpy(rows, E->d_trace, sizeof(u32) * 8ull * k, hipMemcpyDeviceToHost));
  *n = k;
  HIPCHK(hipMemset(E->d_trace_cnt, 0, sizeof(u32)));
  return SCS_OK;
}

extern "C" void scs_engine_destroy(scs_engine* E) {
  if (!E) return;
  if (E->s) (void)hipStreamSynchronize(E->s);
  if (E->ps) (void)hipStreamSynchronize(E->ps);
  E->heu_gpu.release();
  void* ptrs[] = {E->d_slots, E->d_witness, E->d_wave, E->d_pool_seq, E->d_pool_data,
                  E->d_pool_pos, E->d_glob, E->d_elig, E->d_vis, E->d_fr, E->d_nf, E->d_dist16,
                  E->d_locid64, E->d_gcnt, E->d_units, E->d_hubs, E->d_hubcnt, E->d_csum,
                  E->d_flags, E->d_tot, E->d_batch, E->d_admit,
                  E->d_prepbytes, E->d_arena, E->d_model, E->d_trace, E->d_trace_cnt};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (E->h_tot) (void)hipHostFree(E->h_tot);
  for (void* q : {(void*)E->h_stage, (void*)E->h_stage_wit, (void*)E->h_stage_idx, (void*)E->h_wit})
    if (q) (void)hipHostFree(q);
  for (void* q : {(void*)E->d_stage, (void*)E->d_stage_wit, (void*)E->d_stage_idx})
    if (q) (void)hipFree(q);
  if (E->h_batch) (void)hipHostFree(E->h_batch);
  for (hipEvent_t ev : {E->ev0, E->ev1, E->pe0, E->pe1})
    if (ev) (void)hipEventDestroy(ev);
  if (E->s) (void)hipStreamDestroy(E->s);
  if (E->ps) (void)hipStreamDestroy(E->ps);
  delete E;
}

extern "C" int scs_submit(scs_engine* E, uint32_t l, uint32_t h, uint32_t q, uint32_t heu,
                          uint32_t limit_doms, uint64_t tag) {
  if (!E) return SCS_EINVAL;
  if (h == 0 || h > SCS_MAX_H || q >= E->g->hg.n || limit_doms > SCS_MAX_DOMS) return SCS_EINVAL;
  std::lock_guard<std::mutex> lk(E->qmu);
  E->queue.push_back(PendingQuery{tag, l, h, q, heu, limit_doms, now_us()});
  return SCS_OK;
}

// Query text "<l> <h> <q> <heu> <limitDoms>" (main.cu:92-109).
extern "C" int scs_submit_text(scs_engine* E, const char* text, uint64_t tag) {
  if (!E || !text) return SCS_EINVAL;
  u64 v[6];
  int c = 0;
  const char* p = text;
  while (*p && c < 6) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (!*p) break;
    if (*p < '0' || *p > '9') return SCS_EINVAL;
    u64 x = 0;
    while (*p >= '0' && *p <= '9') {
      x = x * 10 + (u64)(*p - '0');
      if (x > 0xffffffffull) return SCS_EINVAL;
      p++;
    }
    if (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') return SCS_EINVAL;
    v[c++] = x;
  }
  while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
  if (c != 5 || *p) return SCS_EINVAL;
  return scs_submit(E, (u32)v[0], (u32)v[1], (u32)v[2], (u32)v[3], (u32)v[4], tag);
}

static void push_result(scs_engine* E, const PendingQuery& pq, u32 kl, u32 heu_kl, u32 ku,
                        int status, u32 K, u64 nodes, const u32* wit, u32 wsize) {
  scs_result r;
  memset(&r, 0, sizeof(r) - sizeof(r.witness));
  r.tag = pq.tag;
  r.l = pq.l;
  r.h = pq.h;
  r.q = pq.q;
  r.kl = kl;
  r.heu_kl = heu_kl;
  r.ku = ku;
  r.status = status;
  r.kept = K;
  r.nodes = nodes;
  r.t_submit_us = pq.t_submit;
  r.t_done_us = now_us();
  r.wsize = std::min<u32>(wsize, SCS_MAX_H);
  if (wit && r.wsize) memcpy(r.witness, wit, sizeof(u32) * r.wsize);
  E->results.push_back(r);
}

// Writes the device record of an admitted query (local graph in the arena).
static void fill_slot(scs_engine* E, u32 si, const MsQuery& P, u32 qloc) {
  HostSlot& hs = E->slots[si];
  QuerySlot& Q = E->h_slots[si];
  memset(&Q, 0, sizeof(Q));
  Q.state = QS_ACTIVE;
  Q.l = hs.pq.l;
  Q.h = hs.pq.h;
  Q.limit_doms = hs.pq.ld;
  const scs_params& p = E->p;
  Q.flags = (p.red1 ? QF_RED1 : 0) | (p.red2 ? QF_RED2 : 0) | (p.red3 ? QF_RED3 : 0) |
            (p.prun1 ? QF_PRUN1 : 0) | (p.prun2 ? QF_PRUN2 : 0) | QF_DOMS |
            (p.learned ? QF_LEARNED : 0) | (p.levels ? QF_LEVELS : 0) | (p.pair ? QF_PAIR : 0);
  Q.tag = (u32)hs.pq.tag;
  Q.ku = hs.ku;
  Q.K = hs.K;
  Q.dmax = E->g->hg.dmax;
  Q.qloc = qloc;
  Q.kl = hs.kl0;
  // first level: min degree >= ku (floor ku-1), when that is above kl0 + 1
  Q.floor = (p.levels && hs.ku > hs.kl0 + 1) ? hs.ku - 1 : 0;
  Q.lvl_nodes0 = 0;
  Q.pending = 1;
  Q.node_budget = p.node_budget;
  Q.dist = P.ldist;
  if (hs.whole) {
    // the resident graph with global ids: vertices outside the kept set
    // have no distance and start removed
    Q.flags |= QF_WHOLE;
    Q.K = E->g->hg.n;
    Q.off = E->g->d_off32;
    Q.adj = (const u32*)E->g->d_adj;
    Q.orig = E->g->d_ident;
    Q.core = E->g->d_core;
  } else {
    Q.off = P.loff;
    Q.adj = P.ladj;
    Q.orig = P.lorig;
    Q.core = P.lcore;
  }
  Q.witness = E->d_witness + (u64)si * SCS_MAX_H;
}

// GPU preprocessing of up to 64 queries at once (slots sis[0..cnt), each
// with its heuristic kl0 and depth bound ubD).  rcs[i] becomes SCS_OK,
// SCS_EBIG (kept set above kcap) or SCS_EAGAIN (arena full; hs.need holds
// the bytes known to be needed).  Successful slots get their QuerySlot.
static int prep_batch(scs_engine* E, const u32* sis, const u32* kl0s, const u32* ubds, u32 cnt,
                      int* rcs, u32 whole) {
  if (cnt == 0 || cnt > 64) return SCS_EINVAL;
  const scs::HostGraph& hg = E->g->hg;
  const u32 n = hg.n;
  const u32 G = (n + 63) / 64;
  hipStream_t s = E->ps;
  // bit b = the b-th query by ascending kl0 (elig masks are prefixes)
  std::vector<u32> ord(cnt);
  for (u32 i = 0; i < cnt; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](u32 x, u32 y) { return kl0s[x] < kl0s[y]; });
  MsBatch& B = *E->h_batch;
  memset(&B, 0, sizeof(B));
  B.nb = cnt;
  std::vector<u32> ubd(cnt);
  u32 maxU = 0;
  for (u32 b = 0; b < cnt; b++) {
    const u32 i = ord[b];
    const HostSlot& hs = E->slots[sis[i]];
    if (hs.pq.q >= n) return SCS_EINVAL;
    B.q[b] = hs.pq.q;
    B.kl0[b] = kl0s[i];
    ubd[b] = std::min<u32>(ubds[i], SCS_MS_LEVELS);
    maxU = std::max(maxU, ubd[b]);
    rcs[i] = SCS_OK;
  }
  const u64 all = cnt == 64 ? ~0ull : ((1ull << cnt) - 1ull);
  const i64* off = E->g->d_off;
  const int32_t* adj = E->g->d_adj;
  HIPCHK(hipEventRecord(E->pe0, s));
  HIPCHK(hipMemcpyAsync(E->d_batch, &B, sizeof(MsBatch), hipMemcpyHostToDevice, s));
  HIPCHK(scs_launch_ms_init(n, E->g->d_core, E->d_batch, E->d_elig, E->d_vis, E->d_fr, E->d_nf,
                            E->d_flags, SCS_MS_LEVELS, E->d_prepbytes, s));
  HIPCHK(scs_launch_ms_seed(E->d_batch, E->d_vis, E->d_fr, E->d_dist16, s));
  // BFS levels; a level that finds nothing ends the later ones on the
  // device, and the host checks every 8 levels
  u64 *fr = E->d_fr, *nf = E->d_nf;
  u32 L = 0;
  if (maxU > 0) {
    // level 0 pushes from the query vertices (top-down)
    u64 act = 0;
    for (u32 b = 0; b < cnt; b++)
      if (ubd[b] > 0) act |= 1ull << b;
    HIPCHK(scs_launch_ms_push0(off, adj, cnt, E->d_batch, act, E->d_elig, E->d_vis, nf,
                               E->d_dist16, E->d_flags, E->d_prepbytes, s));
    std::swap(fr, nf);
    L = 1;
  }
  while (L < maxU) {
    const u32 stop = std::min(maxU, L + 8);
    for (; L < stop; L++) {
      u64 act = 0;
      for (u32 b = 0; b < cnt; b++)
        if (ubd[b] > L) act |= 1ull << b;
      HIPCHK(scs_launch_ms_level(off, adj, n, L, act, E->d_elig, E->d_vis, fr, nf, E->d_dist16,
                                 E->d_flags + (L ? L - 1 : 0), E->d_flags + L, E->d_prepbytes, s));
      std::swap(fr, nf);
    }
    if (L >= maxU) break;
    HIPCHK(hipMemcpyAsync(E->h_tot, E->d_flags + (L - 1), sizeof(u32), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (E->h_tot[0] == 0) break;
  }
  // kept sets: group counts, per-query bases and sizes
  HIPCHK(scs_launch_ms_count(n, G, E->d_vis, all, E->d_gcnt, E->d_prepbytes, s));
  HIPCHK(scs_launch_ms_bscan(cnt, G, E->d_gcnt, E->d_tot, E->d_prepbytes, s));
  HIPCHK(hipMemcpyAsync(E->h_tot, E->d_tot, sizeof(u32) * 64, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  u64 valid = 0, gmask = 0;
  u32 maxK = 0;
  // whole-graph queries search the resident CSR in place: their search
  // state covers every vertex, so the graph must fit the per-wave capacity
  const bool can_whole = whole && E->g->d_off32 && n <= E->kcap;
  for (u32 b = 0; b < cnt; b++) {
    const u32 i = ord[b];
    HostSlot& hs = E->slots[sis[i]];
    const u32 K = E->h_tot[b];
    hs.K = K;
    hs.whole = false;
    if (K > E->kcap || K == 0) {
      rcs[i] = SCS_EBIG;
      continue;
    }
    if (can_whole && (whole == 2 || 2ull * K >= n)) {
      // one distance array over the whole graph
      const u64 a_len = sizeof(u32) * (u64)n;
      if (!E->arena.alloc(a_len, &hs.a_off)) {
        hs.need = scs::Arena::granule(a_len);
        rcs[i] = SCS_EAGAIN;
        continue;
      }
      hs.a_len = a_len;
      hs.whole = true;
      B.P[b] = MsQuery{nullptr, (u32*)(E->d_arena + hs.a_off), nullptr, nullptr, nullptr};
      gmask |= 1ull << b;
      continue;
    }
    // block A: loff (K+1) | ldist (K) | lcore (K) | lorig (K).  a_len/b_len
    // are set only once the allocation succeeded: release_slot frees what
    // they describe.
    const u64 a_len = sizeof(u32) * (4ull * K + 1);
    if (!E->arena.alloc(a_len, &hs.a_off)) {
      hs.need = scs::Arena::granule(a_len);
      rcs[i] = SCS_EAGAIN;
      continue;
    }
    hs.a_len = a_len;
    u32* A = (u32*)(E->d_arena + hs.a_off);
    B.P[b].loff = A;
    B.P[b].ldist = A + K + 1;
    B.P[b].lcore = A + 2ull * K + 1;
    B.P[b].lorig = A + 3ull * K + 1;
    valid |= 1ull << b;
    maxK = std::max(maxK, K);
  }
  if (valid | gmask) HIPCHK(hipMemcpyAsync(E->d_batch, &B, sizeof(MsBatch), hipMemcpyHostToDevice, s));
  if (gmask)
    HIPCHK(scs_launch_ms_gdist(n, G, E->d_vis, gmask, E->d_dist16, E->d_batch, E->d_prepbytes, s));
  if (valid) {
    HIPCHK(scs_launch_ms_scatter(n, G, E->d_vis, valid, E->d_gcnt, E->g->d_core, E->d_dist16,
                                 E->d_batch, E->d_locid64, E->d_tot + 128, E->d_prepbytes, s));
    HIPCHK(scs_launch_ms_degree(adj, E->d_units, E->nunits, E->d_hubs, E->nhubs, E->d_vis,
                                valid, E->d_locid64, E->d_batch, E->d_hubcnt, E->d_prepbytes, s));
    HIPCHK(scs_launch_ms_scan_loff(cnt, E->d_batch, E->d_tot, valid,
                                   std::max<u32>(1, (maxK + 4095) / 4096), E->d_csum,
                                   E->d_prepbytes, s));
    HIPCHK(hipMemcpyAsync(E->h_tot + 64, E->d_tot + 64, sizeof(u32) * 128, hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    for (u32 b = 0; b < cnt; b++) {
      if (!((valid >> b) & 1ull)) continue;
      const u32 i = ord[b];
      HostSlot& hs = E->slots[sis[i]];
      const u32 nnz = E->h_tot[64 + b];
      const u64 b_len = sizeof(u32) * (u64)std::max<u32>(nnz, 1);
      if (!E->arena.alloc(b_len, &hs.b_off)) {
        E->arena.release(hs.a_off, hs.a_len);
        hs.a_len = 0;
        hs.need = scs::Arena::granule(sizeof(u32) * (4ull * hs.K + 1)) + scs::Arena::granule(b_len);
        rcs[i] = SCS_EAGAIN;
        valid &= ~(1ull << b);
        continue;
      }
      hs.b_len = b_len;
      B.P[b].ladj = (u32*)(E->d_arena + hs.b_off);
    }
    if (valid) {
      HIPCHK(hipMemcpyAsync(E->d_batch, &B, sizeof(MsBatch), hipMemcpyHostToDevice, s));
      HIPCHK(scs_launch_ms_fill(adj, E->d_units, E->nunits, E->d_vis, valid, E->d_locid64,
                                E->d_batch, E->d_hubcnt, E->d_prepbytes, s));
    }
  }
  HIPCHK(hipEventRecord(E->pe1, s));
  HIPCHK(hipEventSynchronize(E->pe1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, E->pe0, E->pe1));
  E->st.prep_us += ms * 1000.0;
  E->st.prep_batches++;
  for (u32 b = 0; b < cnt; b++) {
    if ((valid >> b) & 1ull) fill_slot(E, sis[ord[b]], B.P[b], E->h_tot[128 + b]);
    if ((gmask >> b) & 1ull) fill_slot(E, sis[ord[b]], B.P[b], B.q[b]);
  }
  return SCS_OK;
}

static void release_slot(scs_engine* E, u32 si) {
  HostSlot& hs = E->slots[si];
  bool ok = true;
  if (hs.a_len) ok &= E->arena.release(hs.a_off, hs.a_len);
  if (hs.b_len) ok &= E->arena.release(hs.b_off, hs.b_len);
  if (!ok) fprintf(stderr, "scs: arena release of slot %u refused (range not allocated)\n", si);
  hs.a_len = hs.b_len = 0;
  hs.whole = false;
  hs.used = false;
  hs.admitted = false;
  hs.expired = false;
}

// Results of the queries whose slots the last epoch marked done.  Runs
// between epochs (no kernel touches the slots): one copy brings the slots,
// one the witness rows of the finished ones, and one writes the slot array
// back with those slots marked free.
static int collect(scs_engine* E) {
  const u32 nq = E->p.limit_queries;
  std::vector<QuerySlot> dev(nq);
  HIPCHK(hipMemcpy(dev.data(), E->d_slots, sizeof(QuerySlot) * nq, hipMemcpyDeviceToHost));
  std::vector<u32> fin;
  for (u32 si = 0; si < nq; si++)
    if (E->slots[si].used && dev[si].state == QS_DONE) fin.push_back(si);
  if (fin.empty()) return SCS_OK;
  const u64 r0 = fin.front(), r1 = (u64)fin.back() + 1;  // witness rows to bring back
  HIPCHK(hipMemcpy(E->h_wit + r0 * SCS_MAX_H, E->d_witness + r0 * SCS_MAX_H,
                   sizeof(u32) * SCS_MAX_H * (r1 - r0), hipMemcpyDeviceToHost));
  std::vector<u32> wit(SCS_MAX_H);
  for (u32 si : fin) {
    HostSlot& hs = E->slots[si];
    QuerySlot& Q = dev[si];
    u32 ws = std::min<u32>(Q.wsize, SCS_MAX_H);
    if (ws) memcpy(wit.data(), E->h_wit + (u64)si * SCS_MAX_H, sizeof(u32) * ws);
    int status;
    // kl is raised only after its witness is stored, so kl == wkl whenever
    // a witness exists; a lost witness (lock never acquired) leaves kl at the
    // last stored one and the result is only a lower bound.  A query stopped
    // by its budget or deadline is still proven when kl reached the slot's
    // ku (min(core[q], h-1), lowered only by levels that ended unstopped).
    const bool wit_ok = !Q.wlost && !(Q.wsize && Q.wkl != Q.kl);
    if (!wit_ok || (Q.capped && Q.kl < Q.ku))
      status = SCS_ST_CAPPED;
    else if (Q.kl == 0 && hs.pq.l >= 2)
      status = SCS_ST_INFEASIBLE;
    else
      status = SCS_ST_OK;
    if (Q.kl == 0) {
      // min degree 0 is only feasible as H = {q}
      ws = hs.pq.l <= 1 ? 1 : 0;
      wit[0] = hs.pq.q;
    }
    push_result(E, hs.pq, Q.kl, hs.kl0, hs.ku, status, hs.K, Q.nodes, wit.data(), ws);
    E->st.nodes += Q.nodes;
    release_slot(E, si);
    Q.state = QS_FREE;
    E->active--;
  }
  HIPCHK(hipMemcpy(E->d_slots, dev.data(), sizeof(QuerySlot) * nq, hipMemcpyHostToDevice));
  return SCS_OK;
}

// Stops the admitted queries that have passed the per-query deadline.
static int expire_late(scs_engine* E) {
  if (!E->p.deadline_ms) return SCS_OK;
  const double now = now_us(), lim = 1000.0 * E->p.deadline_ms;
  std::vector<u32> late;
  for (u32 si = 0; si < E->p.limit_queries; si++) {
    HostSlot& hs = E->slots[si];
    if (hs.used && hs.admitted && !hs.expired && now - hs.pq.t_submit > lim) {
      hs.expired = true;
      late.push_back(si);
    }
  }
  if (late.empty()) return SCS_OK;
  HIPCHK(hipMemcpyAsync(E->d_admit, late.data(), sizeof(u32) * late.size(), hipMemcpyHostToDevice,
                        E->s));
  HIPCHK(scs_launch_expire(E->d_slots, E->d_admit, (u32)late.size(), E->s));
  HIPCHK(hipStreamSynchronize(E->s));
  return SCS_OK;
}

static int launch_epoch(scs_engine* E) {
  int rc = expire_late(E);
  if (rc) return rc;
  EngineGlobals gl;
  HIPCHK(hipMemcpy(&gl, E->d_glob, sizeof(gl), hipMemcpyDeviceToHost));
  gl.idle = 0;
  gl.active_q = E->active;
  HIPCHK(hipMemcpyAsync(E->d_glob, &gl, sizeof(gl), hipMemcpyHostToDevice, E->s));
  HIPCHK(hipEventRecord(E->ev0, E->s));
  HIPCHK(scs_launch_epoch(&E->args, E->s));
  HIPCHK(hipEventRecord(E->ev1, E->s));
  return SCS_OK;
}

static int finish_epoch(scs_engine* E) {
  HIPCHK(hipEventSynchronize(E->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
  E->st.engine_us += ms * 1000.0;
  E->st.epochs++;
  return collect(E);
}

// One serving step.  The search epoch of the queries already admitted runs
// on the engine stream while newly queued queries get their heuristic (host
// threads) and GPU preprocessing (second stream); they are written into
// their slots and their root tasks pushed once the epoch has ended, so they
// search from the next step on.
extern "C" int scs_step(scs_engine* E) {
  if (!E) return SCS_EINVAL;
  const scs::HostGraph& hg = E->g->hg;
  const bool running = E->active > 0;
  if (running) {
    int rc = launch_epoch(E);
    if (rc) return rc;
  }
  // 1. take queued queries for free slots
  std::vector<u32> free_slots;
  for (u32 si = 0; si < E->p.limit_queries; si++)
    if (!E->slots[si].used) free_slots.push_back(si);
  std::vector<PendingQuery> batch;
  {
    std::lock_guard<std::mutex> lk(E->qmu);
    while (!E->queue.empty() && batch.size() < free_slots.size() &&
           batch.size() < E->p.admit_per_step) {
      batch.push_back(E->queue.front());
      E->queue.pop_front();
    }
  }
  // 2. heuristic lower bounds on host threads
  std::vector<u32> kl0(batch.size(), 0);
  std::vector<std::vector<u32>> H(batch.size());
  if (!batch.empty() && E->heu_gpu.waves) {
    // on the GPU, on the preprocessing stream, up to heu_gpu.waves queries
    // per launch
    double t0 = now_us();
    std::vector<u32> idx, gl, gh, gq;
    for (size_t i = 0; i < batch.size(); i++)
      if (batch[i].heu && batch[i].l <= batch[i].h) {
        idx.push_back((u32)i);
        gl.push_back(batch[i].l);
        gh.push_back(batch[i].h);
        gq.push_back(batch[i].q);
      }
    std::vector<u32> gk(idx.size());
    std::vector<std::vector<u32>> GH(E->heu_gpu.waves);
    for (size_t i0 = 0; i0 < idx.size(); i0 += E->heu_gpu.waves) {
      const u32 cnt = (u32)std::min<size_t>(E->heu_gpu.waves, idx.size() - i0);
      int rc = E->heu_gpu.run(E->g, cnt, &gl[i0], &gh[i0], &gq[i0], &gk[i0], GH.data(), E->ps);
      if (rc) return rc;
      for (u32 i = 0; i < cnt; i++) {
        kl0[idx[i0 + i]] = gk[i0 + i];
        H[idx[i0 + i]].swap(GH[i]);
      }
    }
    E->st.heu_us += now_us() - t0;
  } else if (!batch.empty()) {
    double t0 = now_us();
    std::atomic<size_t> next{0};
    auto worker = [&](u32 t) {
      size_t i;
      while ((i = next.fetch_add(1)) < batch.size()) {
        const PendingQuery& pq = batch[i];
        if (pq.heu && pq.l <= pq.h) kl0[i] = scs::cssc_heu(hg, pq.l, pq.h, pq.q, E->heu_work[t], H[i]);
      }
    };
    u32 nt = std::min<u32>(E->p.heuristic_threads, (u32)batch.size());
    std::vector<std::thread> th;
    for (u32 t = 1; t < nt; t++) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    E->st.heu_us += now_us() - t0;
  }
  // 3. immediate answers; the rest get batched GPU preprocessing (second
  // stream, 64 queries per pass)
  std::vector<u32> admitted;
  std::vector<u32> cand_i, cand_si, cand_kl0, cand_ubd;
  size_t fs = 0;
  for (size_t i = 0; i < batch.size(); i++) {
    const PendingQuery& pq = batch[i];
    const u32 ku = std::min(hg.core[pq.q], pq.h - 1);
    if (pq.l > pq.h) {
      push_result(E, pq, 0, 0, ku, SCS_ST_INFEASIBLE, 0, 0, nullptr, 0);
      continue;
    }
    if (pq.heu && kl0[i] == ku) {
      if (ku == 0) {
        // a connected H of two or more vertices has min degree >= 1
        const u32 one = pq.q;
        if (pq.l >= 2)
          push_result(E, pq, 0, 0, ku, SCS_ST_INFEASIBLE, 0, 0, nullptr, 0);
        else
          push_result(E, pq, 0, 0, ku, SCS_ST_HEU_OPT, 0, 0, &one, 1);
        continue;
      }
      push_result(E, pq, kl0[i], kl0[i], ku, SCS_ST_HEU_OPT, 0, 0, H[i].data(), (u32)H[i].size());
      continue;
    }
    const u32 si = free_slots[fs++];
    HostSlot& hs = E->slots[si];
    hs.used = true;
    hs.a_len = hs.b_len = 0;
    hs.need = 0;
    hs.pq = pq;
    hs.kl0 = kl0[i];
    hs.ku = ku;
    hs.K = 0;
    cand_i.push_back((u32)i);
    cand_si.push_back(si);
    cand_kl0.push_back(kl0[i]);
    cand_ubd.push_back(scs::ubd_host(kl0[i], pq.h));
  }
  std::vector<int> rcs(cand_i.size(), SCS_OK);
  for (size_t c0 = 0; c0 < cand_i.size(); c0 += 64) {
    const u32 cnt = (u32)std::min<size_t>(64, cand_i.size() - c0);
    int rc = prep_batch(E, &cand_si[c0], &cand_kl0[c0], &cand_ubd[c0], cnt, &rcs[c0],
                        E->p.whole_graph);
    if (rc) return rc;
  }
  std::vector<PendingQuery> requeue;
  for (size_t c = 0; c < cand_i.size(); c++) {
    const size_t i = cand_i[c];
    const u32 si = cand_si[c];
    HostSlot& hs = E->slots[si];
    const PendingQuery& pq = batch[i];
    if (rcs[c] == SCS_EBIG || (rcs[c] == SCS_EAGAIN && hs.need > E->arena.size)) {
      // kept set above the per-query limit, or a local graph larger than
      // the whole arena
      push_result(E, pq, kl0[i], kl0[i], hs.ku, SCS_ST_REJECTED, hs.K, 0, H[i].data(),
                  (u32)H[i].size());
      release_slot(E, si);
      continue;
    }
    if (rcs[c] == SCS_EAGAIN) {
      release_slot(E, si);
      requeue.push_back(pq);
      continue;
    }
    QuerySlot& Q = E->h_slots[si];
    if (kl0[i] > 0 && !H[i].empty() && H[i].size() <= SCS_MAX_H) {
      Q.wsize = (u32)H[i].size();
      Q.wkl = kl0[i];
      hs.seed_witness = H[i];
    } else {
      hs.seed_witness.clear();
    }
    admitted.push_back(si);
  }
  if (!requeue.empty()) {
    std::lock_guard<std::mutex> lk(E->qmu);
    for (auto it = requeue.rbegin(); it != requeue.rend(); ++it) E->queue.push_front(*it);
  }
  // 4. end of the epoch: results of finished queries
  if (running) {
    int rc = finish_epoch(E);
    if (rc) return rc;
  }
  // 5. admitted queries become visible to the search (engine stream)
  if (!admitted.empty()) {
    const u32 na = (u32)admitted.size();
    for (u32 i = 0; i < na; i++) {
      const u32 si = admitted[i];
      const HostSlot& hs = E->slots[si];
      const u32 ws = (u32)std::min<size_t>(hs.seed_witness.size(), SCS_MAX_H);
      E->h_stage[i] = E->h_slots[si];
      E->h_stage_idx[2 * i] = si;
      E->h_stage_idx[2 * i + 1] = ws;
      if (ws) memcpy(E->h_stage_wit + (u64)i * SCS_MAX_H, hs.seed_witness.data(), sizeof(u32) * ws);
    }
    HIPCHK(hipMemcpyAsync(E->d_stage, E->h_stage, sizeof(QuerySlot) * na, hipMemcpyHostToDevice,
                          E->s));
    HIPCHK(hipMemcpyAsync(E->d_stage_idx, E->h_stage_idx, sizeof(u32) * 2 * na,
                          hipMemcpyHostToDevice, E->s));
    HIPCHK(hipMemcpyAsync(E->d_stage_wit, E->h_stage_wit, sizeof(u32) * SCS_MAX_H * (u64)na,
                          hipMemcpyHostToDevice, E->s));
    HIPCHK(scs_launch_slot_write(E->d_stage, E->d_stage_idx, E->d_stage_wit, na, E->d_slots,
                                 E->d_witness, E->s));
    HIPCHK(hipMemcpyAsync(E->d_admit, admitted.data(), sizeof(u32) * admitted.size(),
                          hipMemcpyHostToDevice, E->s));
    HIPCHK(scs_launch_admit(&E->args, E->d_admit, (u32)admitted.size(),
                            E->d_admit + E->p.limit_queries, E->s));
    std::vector<u32> failed(admitted.size());
    HIPCHK(hipMemcpyAsync(failed.data(), E->d_admit + E->p.limit_queries,
                          sizeof(u32) * admitted.size(), hipMemcpyDeviceToHost, E->s));
    HIPCHK(hipStreamSynchronize(E->s));
    // a root task that found the pool full (not expected: donations leave
    // limit_queries slots free) frees its slot and goes back to the queue
    std::vector<PendingQuery> back;
    u32 nok = 0;
    for (u32 i = 0; i < na; i++) {
      const u32 si = admitted[i];
      if (!failed[i]) {
        E->slots[si].admitted = true;
        E->slots[si].expired = false;
        nok++;
        continue;
      }
      QuerySlot& Q = E->h_slots[si];
      memset(&Q, 0, sizeof(Q));
      Q.state = QS_FREE;
      HIPCHK(hipMemcpy(E->d_slots + si, &Q, sizeof(QuerySlot), hipMemcpyHostToDevice));
      back.push_back(E->slots[si].pq);
      release_slot(E, si);
    }
    if (!back.empty()) {
      std::lock_guard<std::mutex> lk(E->qmu);
      for (auto it = back.rbegin(); it != back.rend(); ++it) E->queue.push_front(*it);
    }
    E->active += nok;
  }
  size_t queued;
  {
    std::lock_guard<std::mutex> lk(E->qmu);
    queued = E->queue.size();
  }
  return (int)(queued + E->active);
}

extern "C" int scs_poll(scs_engine* E, scs_result* out, int max) {
  if (!E || (max > 0 && !out)) return SCS_EINVAL;
  int c = 0;
  while (c < max && !E->results.empty()) {
    out[c++] = E->results.front();
    E->results.pop_front();
  }
  return c;
}

extern "C" int scs_drain(scs_engine* E) {
  while (true) {
    int r = scs_step(E);
    if (r <= 0) return r;
  }
}

extern "C" int scs_engine_stats(const scs_engine* E, scs_stats* s) {
  if (!E || !s) return SCS_EINVAL;
  *s = E->st;
  EngineGlobals gl;
  HIPCHK(hipMemcpy(&gl, E->d_glob, sizeof(gl), hipMemcpyDeviceToHost));
  s->donations = gl.donations;
  s->rebuilds = gl.rebuilds;
  s->edges = gl.edges;
  for (int i = 0; i < SCS_PHASES; i++) s->phase_ticks[i] = gl.tph[i];
  s->dom_front = gl.dom_front;
  s->dom_cand = gl.dom_cand;
  s->dom_found = gl.dom_found;
  s->full_scans = gl.full_scans;
  s->full_rows = gl.full_rows;
  s->core_us = E->g->core_us;
  std::vector<unsigned long long> pb((size_t)MS_BYTE_SLOTS * MS_BYTE_STRIDE);
  HIPCHK(hipMemcpy(pb.data(), E->d_prepbytes, sizeof(unsigned long long) * pb.size(),
                   hipMemcpyDeviceToHost));
  s->prep_bytes = 0;
  for (u32 i = 0; i < MS_BYTE_SLOTS; i++) s->prep_bytes += pb[(size_t)i * MS_BYTE_STRIDE];
  return SCS_OK;
}

extern "C" int scs_heuristic_csr(const int64_t* off, const int32_t* adj, uint32_t n, uint32_t l,
                                 uint32_t h, uint32_t q, uint32_t* kl, uint32_t* H, uint32_t cap,
                                 uint32_t* hsize) {
  if (!off || !kl || q >= n || h == 0 || l > h) return SCS_EINVAL;
  if (off[0] != 0 || off[n] < 0 || (off[n] && !adj)) return SCS_EINVAL;
  for (u32 v = 0; v < n; v++)
    if (off[v + 1] < off[v]) return SCS_EINVAL;
  for (int64_t j = 0; j < off[n]; j++)
    if (adj[j] < 0 || (u32)adj[j] >= n) return SCS_EINVAL;
  scs::HostGraph g;
  g.n = n;
  g.off.assign(off, off + n + 1);
  g.adj.assign(adj, adj + off[n]);
  g.dmax = 0;
  for (u32 v = 0; v < n; v++) g.dmax = std::max<u32>(g.dmax, (u32)(off[v + 1] - off[v]));
  scs::HeuWork W;
  std::vector<u32> HH;
  *kl = scs::cssc_heu(g, l, h, q, W, HH);
  if (hsize) *hsize = (u32)HH.size();
  if (H)
    for (u32 i = 0; i < HH.size() && i < cap; i++) H[i] = HH[i];
  return SCS_OK;
}

extern "C" int scs_engine_heuristic(scs_engine* E, uint32_t l, uint32_t h, uint32_t q,
                                    uint32_t* kl, uint32_t* ubD) {
  if (!E || !kl || h == 0 || l > h || q >= E->g->hg.n) return SCS_EINVAL;
  std::vector<u32> HH;
  *kl = scs::cssc_heu(E->g->hg, l, h, q, E->heu_work[0], HH);
  if (ubD) *ubD = scs::ubd_host(*kl, h);
  return SCS_OK;
}

extern "C" int scs_prep_probe_many(scs_engine* E, uint32_t count, const uint32_t* q,
                                   const uint32_t* kl0, const uint32_t* ubD, uint32_t* K,
                                   uint64_t* nnz, uint32_t* kept, uint32_t* loff, uint32_t* dist,
                                   uint32_t kept_cap, uint32_t* ladj, uint64_t adj_cap) {
  if (!E || !q || !kl0 || !ubD || !K || !nnz || count == 0 || count > 64) return SCS_EINVAL;
  for (u32 i = 0; i < count; i++)
    if (q[i] >= E->g->hg.n) return SCS_EINVAL;
  // borrow free slots for the duration of the call
  std::vector<u32> sis;
  for (u32 si = 0; si < E->p.limit_queries && sis.size() < count; si++)
    if (!E->slots[si].used) sis.push_back(si);
  if (sis.size() < count) return SCS_EAGAIN;
  for (u32 i = 0; i < count; i++) {
    HostSlot& hs = E->slots[sis[i]];
    hs.pq = PendingQuery{0, 1, 1, q[i], 0, 0, now_us()};
    hs.ku = 0;
    hs.kl0 = kl0[i];
    hs.K = 0;
    hs.a_len = hs.b_len = 0;
  }
  std::vector<int> rcs(count, SCS_OK);
  // local graphs when the caller asks for them; otherwise the batch is
  // prepared as admission prepares it (whole-graph queries included)
  const bool copy = kept || loff || dist || ladj;
  const int brc = prep_batch(E, sis.data(), kl0, ubD, count, rcs.data(),
                             copy ? 0u : E->p.whole_graph);
  int out = brc;
  for (u32 i = 0; i < count && brc == SCS_OK; i++) {
    HostSlot& hs = E->slots[sis[i]];
    K[i] = hs.K;
    nnz[i] = 0;
    if (rcs[i] == SCS_EAGAIN) out = SCS_EAGAIN;
    if (rcs[i] != SCS_OK) continue;
    if (hs.whole) continue;  // no local graph (nnz stays 0)
    const QuerySlot& Q = E->h_slots[sis[i]];
    const u32 k = hs.K;
    u32 nz = 0;
    HIPCHK(hipMemcpy(&nz, Q.off + k, sizeof(u32), hipMemcpyDeviceToHost));
    nnz[i] = nz;
    if (k <= kept_cap) {
      if (kept)
        HIPCHK(hipMemcpy(kept + (u64)i * kept_cap, Q.orig, sizeof(u32) * k, hipMemcpyDeviceToHost));
      if (loff)
        HIPCHK(hipMemcpy(loff + (u64)i * (kept_cap + 1ull), Q.off, sizeof(u32) * (k + 1),
                         hipMemcpyDeviceToHost));
      if (dist)
        HIPCHK(hipMemcpy(dist + (u64)i * kept_cap, Q.dist, sizeof(u32) * k, hipMemcpyDeviceToHost));
    }
    if (ladj && nz <= adj_cap)
      HIPCHK(hipMemcpy(ladj + (u64)i * adj_cap, Q.adj, sizeof(u32) * (u64)nz,
                       hipMemcpyDeviceToHost));
  }
  for (u32 si : sis) release_slot(E, si);
  return out;
}

extern "C" int scs_prep_probe(scs_engine* E, uint32_t q, uint32_t kl0, uint32_t ubD, uint32_t* K,
                              uint64_t* nnz, uint32_t* kept, uint32_t* loff, uint32_t* dist,
                              uint32_t kept_cap, uint32_t* ladj, uint64_t adj_cap) {
  if (!E || !K || !nnz) return SCS_EINVAL;
  return scs_prep_probe_many(E, 1, &q, &kl0, &ubD, K, nnz, kept, loff, dist, kept_cap, ladj,
                             adj_cap);
}

extern "C" const char* scs_strerror(int code) {
  switch (code) {
    case SCS_OK: return "ok";
    case SCS_EINVAL: return "invalid argument";
    case SCS_ENOMEM: return "out of memory";
    case SCS_EAGAIN: return "admission refused for now";
    case SCS_EHIP: return "HIP runtime error";
    case SCS_EIO: return "graph file unreadable";
    case SCS_EBIG: return "query exceeds per-query limits";
    default: return "unknown error";
  }
}
