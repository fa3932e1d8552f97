// CPU oracle for size-bounded community search (SCS).
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU and in plain
// C++, the algorithm of the reference's SCS hot path so that the HIP path
// can be checked against it.  Only tests/, the smoke check and bench.py's
// CPU-baseline leg may load it; the product never links or calls it.
//
// How each part is pinned (tests/test_oracle.py, tests/test_workloads.py):
//   loader, core numbers, BFS: networkx on every shipped and generated
//     graph, closed forms (K_n, cycles, trees), the shipped data files'
//     edge counts;
//   optimum: brute force over connected sets on random small graphs, the
//     known answer of data/graph.txt (kl 3), closed forms, monotonicity in
//     l and h, and two independent exact solvers (full rules vs no rules)
//     agreeing on every C1 query;
//   U_nr: equal to the brute-force optimum of its relaxation;
//   diameter bound: hand-checked small cases;
//   heuristic kl values: invariants only (heu kl <= optimum <= ku); their
//     exact values are "parity unpinned" (the reference's priority-queue
//     tie order is not reproducible without its own build), and they never
//     change the optimum.
//
// Reference citations are file:line in wajidmanzoor/SCS.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <queue>
#include <set>
#include <string>
#include <vector>

typedef uint32_t u32;
typedef int64_t i64;
typedef uint64_t u64;

static const u32 INF_DIST = 1000000000u;  // inc/Utility.h:33 (INF)

struct Graph {
  u32 n = 0;
  std::vector<i64> off;  // n+1
  std::vector<u32> adj;  // off[n]
  u32 dmax = 0;
  u32 deg(u32 v) const { return (u32)(off[v + 1] - off[v]); }
};

// ---------------------------------------------------------------------------
// Loader: header "n m", then whitespace-separated pairs; self-loops dropped,
// duplicates merged, adjacency ascending (src/Graph.h:5-67).  Vertex ids >= n
// are rejected (the reference would index out of bounds).
// ---------------------------------------------------------------------------
static int load_edge_list(const char* path, Graph& g) {
  std::ifstream in(path);
  if (!in.is_open()) return -1;
  u64 n, m;
  if (!(in >> n >> m)) return -2;
  std::vector<std::pair<u32, u32>> e;
  u64 a, b;
  while (in >> a >> b) {
    if (a >= n || b >= n) return -3;
    if (a == b) continue;
    e.push_back({(u32)a, (u32)b});
    e.push_back({(u32)b, (u32)a});
  }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  g.n = (u32)n;
  g.off.assign(n + 1, 0);
  g.adj.resize(e.size());
  for (size_t i = 0; i < e.size(); i++) {
    g.off[e[i].first + 1]++;
    g.adj[i] = e[i].second;
  }
  for (u64 i = 0; i < n; i++) g.off[i + 1] += g.off[i];
  g.dmax = 0;
  for (u32 v = 0; v < g.n; v++) g.dmax = std::max(g.dmax, g.deg(v));
  return 0;
}

// ---------------------------------------------------------------------------
// Core numbers: Batagelj-Zaversnik bucket peeling.  Same result as the
// reference's linear-heap peeling (src/Graph.h:82-109, inc/ListLinearHeap.h);
// core numbers are unique.
// ---------------------------------------------------------------------------
static void core_numbers(const Graph& g, std::vector<u32>& core) {
  u32 n = g.n;
  std::vector<u32> deg(n), pos(n), vert(n);
  u32 md = 0;
  for (u32 v = 0; v < n; v++) {
    deg[v] = g.deg(v);
    md = std::max(md, deg[v]);
  }
  std::vector<u32> bin(md + 2, 0);
  for (u32 v = 0; v < n; v++) bin[deg[v]]++;
  u32 s = 0;
  for (u32 d = 0; d <= md; d++) {
    u32 c = bin[d];
    bin[d] = s;
    s += c;
  }
  for (u32 v = 0; v < n; v++) {
    pos[v] = bin[deg[v]]++;
    vert[pos[v]] = v;
  }
  for (u32 d = md + 1; d > 0; d--) bin[d] = bin[d - 1];
  bin[0] = 0;
  for (u32 i = 0; i < n; i++) {
    u32 v = vert[i];
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 u = g.adj[j];
      if (deg[u] > deg[v]) {
        u32 du = deg[u], pu = pos[u], pw = bin[du], w = vert[pw];
        if (u != w) {
          pos[u] = pw;
          vert[pu] = w;
          pos[w] = pu;
          vert[pw] = u;
        }
        bin[du]++;
        deg[u]--;
      }
    }
  }
  core = deg;
}

// BFS over the whole graph (src/Graph.h:111-142).
static void bfs_full(const Graph& g, u32 q, std::vector<u32>& dist) {
  dist.assign(g.n, INF_DIST);
  std::vector<u32> fr{q};
  dist[q] = 0;
  for (size_t h = 0; h < fr.size(); h++) {
    u32 v = fr[h];
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 w = g.adj[j];
      if (dist[w] == INF_DIST) {
        dist[w] = dist[v] + 1;
        fr.push_back(w);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Heuristics (inc/heuristic.h).  kl values and witnesses follow the
// reference's greedy rules, including std::priority_queue tie order (larger
// vertex id first on equal keys).
// ---------------------------------------------------------------------------

// heu2: greedy growth by degree-into-H (inc/heuristic.h:8-120).
static void heu2(const Graph& g, u32 N1, u32 N2, u32 q, std::vector<u32>& H, u32& kl) {
  H.clear();
  kl = 0;
  std::vector<u32> tH;
  u32 tkl = 0, tsize = 0;
  std::vector<unsigned char> sta(g.n, 0);
  std::vector<u32> deg(g.n, 0);
  std::priority_queue<std::pair<u32, u32>> Q;
  Q.push({0, q});
  sta[q] = 1;
  while (!Q.empty()) {
    u32 v = Q.top().second;
    Q.pop();
    if (sta[v] == 2) continue;
    tH.push_back(v);
    sta[v] = 2;
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 nb = g.adj[j];
      if (sta[nb] == 0 || sta[nb] == 1) {
        u32 d = 0;
        for (i64 w = g.off[nb]; w < g.off[nb + 1]; w++)
          if (sta[g.adj[w]] == 2) ++d;
        Q.push({d, nb});
        sta[nb] = 1;
      } else {
        ++deg[nb];
        ++deg[v];
      }
    }
    if (tH.size() >= N1) {
      u32 mindeg = INF_DIST;
      for (u32 x : tH) mindeg = std::min(mindeg, deg[x]);
      if (mindeg >= tkl) {
        tkl = mindeg;
        tsize = (u32)tH.size();
      }
    }
    if (tH.size() == N2) break;
  }
  H.assign(tH.begin(), tH.begin() + tsize);
  kl = tkl;
}

// heu3: greedy growth by connection score (inc/heuristic.h:125-239).
static void heu3(const Graph& g, const std::vector<u32>& degree, u32 dMAX, u32 N1, u32 N2,
                 u32 q, std::vector<u32>& H, u32& kl) {
  H.clear();
  kl = 0;
  std::vector<u32> tH;
  u32 tkl = 0, tsize = 0;
  std::vector<unsigned char> sta(g.n, 0);
  std::vector<u32> deg(g.n, 0);
  std::priority_queue<std::pair<double, u32>> Q;
  Q.push({0.0, q});
  sta[q] = 1;
  auto score_of = [&](u32 x) {
    double s = 0;
    for (i64 w = g.off[x]; w < g.off[x + 1]; w++) {
      u32 y = g.adj[w];
      if (sta[y] == 2 && deg[y] != 0) s += (double)1 / deg[y];
    }
    s += (double)degree[x] / dMAX;
    return s;
  };
  while (!Q.empty()) {
    u32 v = Q.top().second;
    Q.pop();
    if (sta[v] == 2) continue;
    tH.push_back(v);
    sta[v] = 2;
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++)
      if (sta[g.adj[j]] == 2) {
        ++deg[g.adj[j]];
        ++deg[v];
      }
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 nb = g.adj[j];
      if (!sta[nb]) {
        Q.push({score_of(nb), nb});
        sta[nb] = 1;
      } else if (sta[nb] == 1) {
        Q.push({score_of(nb), nb});
      }
    }
    if (tH.size() >= N1) {
      u32 mindeg = INF_DIST;
      for (u32 x : tH) mindeg = std::min(mindeg, deg[x]);
      if (mindeg >= tkl) {
        tkl = mindeg;
        tsize = (u32)tH.size();
      }
    }
    if (tH.size() == N2) break;
  }
  H.assign(tH.begin(), tH.begin() + tsize);
  kl = tkl;
}

// heu4: peel q plus its neighbourhood (inc/heuristic.h:242-317), including
// the +1 on q's own degree (inc/heuristic.h:259) and the unsigned N1-1 test.
static void heu4(const Graph& g, u32 N1, u32 N2, u32 q, std::vector<u32>& H, u32& kl) {
  H.clear();
  kl = 0;
  if (g.deg(q) < N1 - 1u) return;
  std::vector<u32> S;
  std::vector<unsigned char> inS(g.n, 0);
  std::vector<u32> deg(g.n, 0);
  S.push_back(q);
  inS[q] = 1;
  ++deg[q];
  for (i64 j = g.off[q]; j < g.off[q + 1]; j++) {
    u32 x = g.adj[j];
    S.push_back(x);
    inS[x] = 1;
    for (i64 w = g.off[x]; w < g.off[x + 1]; w++)
      if (inS[g.adj[w]]) {
        ++deg[g.adj[w]];
        ++deg[x];
      }
  }
  std::priority_queue<std::pair<u32, u32>, std::vector<std::pair<u32, u32>>,
                      std::greater<std::pair<u32, u32>>>
      Q;
  for (u32 e : S) Q.push({deg[e], e});
  std::vector<u32> rmv;
  u32 rmv_idx = 0, mindeg = 0;
  while (!Q.empty()) {
    u32 d = Q.top().first, v = Q.top().second;
    Q.pop();
    if (!inS[v]) continue;
    u32 remain = (u32)S.size() - (u32)rmv.size();
    // q carries a +1 (inc/heuristic.h:259) that keeps it from being peeled
    // before the others; it is popped only when alone, where its true
    // degree is 0, so the +1 is taken back for the min-degree test
    if (v == q) d -= 1;
    if (remain >= N1 && remain <= N2 && d > mindeg) {
      mindeg = d;
      rmv_idx = (u32)rmv.size();
    }
    if (remain == N1) break;
    inS[v] = 0;
    rmv.push_back(v);
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 x = g.adj[j];
      if (inS[x]) {
        --deg[x];
        Q.push({deg[x], x});
      }
    }
  }
  std::fill(inS.begin(), inS.end(), 0);
  for (u32 e : S) inS[e] = 1;
  for (u32 i = 0; i < rmv_idx; i++) inS[rmv[i]] = 0;
  for (u32 e : S)
    if (inS[e]) H.push_back(e);
  kl = mindeg;
}

// CSSC_heu: best of heu2/heu3/heu4 with the reference's tie order
// (inc/heuristic.h:321-348).
static u32 cssc_heu(const Graph& g, const std::vector<u32>& degree, u32 N1, u32 N2, u32 q,
                    std::vector<u32>& H) {
  std::vector<u32> H2, H3, H4;
  u32 k2, k3, k4;
  heu2(g, N1, N2, q, H2, k2);
  heu3(g, degree, g.dmax, N1, N2, q, H3, k3);
  heu4(g, N1, N2, q, H4, k4);
  if (k4 >= k3 && k4 >= k2) {
    H = H4;
    return k4;
  }
  if (k3 >= k2) {
    H = H3;
    return k3;
  }
  H = H2;
  return k2;
}

// ---------------------------------------------------------------------------
// Diameter bounds.
// ---------------------------------------------------------------------------
// Host form (main.cu:154-172): kl <= 1 gives h-1.
static u32 ubd_host(u32 kl, u32 h) {
  if (kl <= 1) return h - 1;
  for (u32 d = 1; d <= h; d++) {
    if (d <= 2) {
      if (kl + d > h) return d - 1;
    } else {
      u32 min_n = kl + d + 1 + (d / 3) * (kl - 2);
      if (h < min_n) return d - 1;
    }
  }
  return h - 1;  // loop never broke: the reference leaves ubD at 0 here
}
// Per-node form (src/helpers.cc:974-990), 32-bit unsigned arithmetic.
static u32 ubd_node(u32 kl, u32 h) {
  u32 ub = h - 1;
  for (u32 d = 1; d <= h; d++) {
    if (d == 1 || d == 2) {
      if (kl + d > h) {
        ub = d - 1;
        break;
      }
    } else {
      u32 min_n = kl + d + 1 + (d / 3) * (kl - 2);
      if (h < min_n) {
        ub = d - 1;
        break;
      }
    }
  }
  return ub;
}

// ---------------------------------------------------------------------------
// Per-query preprocessing.
//   K = { v : core[v] > kl0 and dist(q, v) <= ubD } where dist is measured in
//   the subgraph induced by { core > kl0 }.  Every H with min degree > kl0
//   lies inside { core > kl0 } and has diameter <= ubD, so this is the
//   reference's initial reduction (src/helpers.cc:232-233) restricted to the
//   part a better solution can reach.  Local ids are ranks of the original
//   ids; the local CSR keeps edges inside K, ascending.
// ---------------------------------------------------------------------------
struct Local {
  u32 K = 0;
  u32 qloc = 0;
  std::vector<u32> orig;   // local -> original
  std::vector<u32> off;    // K+1
  std::vector<u32> adj;    // local ids
  std::vector<u32> dist;   // restricted BFS distance
  std::vector<u32> core;
};

static void preprocess(const Graph& g, const std::vector<u32>& core, u32 q, u32 kl0, u32 ubD,
                       Local& L) {
  std::vector<u32> dist(g.n, INF_DIST);
  std::vector<u32> fr{q};
  dist[q] = 0;
  for (size_t i = 0; i < fr.size(); i++) {
    u32 v = fr[i];
    if (dist[v] >= ubD) continue;
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++) {
      u32 w = g.adj[j];
      if (core[w] > kl0 && dist[w] == INF_DIST) {
        dist[w] = dist[v] + 1;
        fr.push_back(w);
      }
    }
  }
  std::sort(fr.begin(), fr.end());
  L.K = (u32)fr.size();
  L.orig = fr;
  std::vector<u32> loc(g.n, UINT32_MAX);
  for (u32 i = 0; i < L.K; i++) loc[fr[i]] = i;
  L.qloc = loc[q];
  L.off.assign(L.K + 1, 0);
  L.adj.clear();
  L.dist.resize(L.K);
  L.core.resize(L.K);
  for (u32 i = 0; i < L.K; i++) {
    u32 v = fr[i];
    for (i64 j = g.off[v]; j < g.off[v + 1]; j++)
      if (loc[g.adj[j]] != UINT32_MAX) L.adj.push_back(loc[g.adj[j]]);
    L.off[i + 1] = (u32)L.adj.size();
    L.dist[i] = dist[v];
    L.core[i] = core[v];
  }
}

// ---------------------------------------------------------------------------
// U_nr (prun2, src/helpers.cc:1066-1101) with a correct sort: water-filling
// of the k = min(h-|C|, |R|) largest R->C degrees onto the smallest C degrees.
// cvals: dC of C vertices.  rhist[d]: number of R vertices with dC == d.
// ---------------------------------------------------------------------------
static u32 unr_bound(const std::vector<u32>& cvals, const std::vector<u32>& rhist, u32 k) {
  if (cvals.empty()) return UINT32_MAX;
  u32 mx = 0, lo = UINT32_MAX;
  for (u32 x : cvals) {
    mx = std::max(mx, x);
    lo = std::min(lo, x);
  }
  std::vector<u32> cnt(mx + k + 2, 0);
  for (u32 x : cvals) cnt[x]++;
  const u32 m = (u32)cvals.size();
  int d = (int)rhist.size() - 1;
  u32 left_in_bucket = d >= 0 ? rhist[d] : 0;
  std::vector<std::pair<u32, u32>> take;
  for (u32 i = 0; i < k; i++) {
    while (d > 0 && left_in_bucket == 0) left_in_bucket = rhist[--d];
    if (d <= 0) break;  // remaining R vertices have dC == 0
    left_in_bucket--;
    u32 x = std::min((u32)d, m);
    u32 need = x, v = lo;
    take.clear();
    while (need) {
      u32 t = std::min(need, cnt[v]);
      if (t) take.push_back({v, t});
      need -= t;
      v++;
    }
    for (int j = (int)take.size() - 1; j >= 0; j--) {
      cnt[take[j].first] -= take[j].second;
      cnt[take[j].first + 1] += take[j].second;
    }
    while (cnt[lo] == 0) lo++;
  }
  return lo;
}

// Fixed-point connection score (src/helpers.cc:1111-1134 in double): each C
// neighbour u adds 2^32/dC(u) (2^32 when dC(u)==0), plus dR(v)*2^32/dmax.
// Integer sums are independent of summation order, so every processor picks
// the same ustar.  The dominated-set score (src/helpers.cc:605-642) skips
// C neighbours with dC == 0.
static inline u64 recip32(u32 d) { return d == 0 ? (1ull << 32) : ((1ull << 32) / d); }
static inline u64 rterm(u32 dR, u32 dmax) { return dmax ? (((u64)dR << 32) / dmax) : 0; }

// ---------------------------------------------------------------------------
// Exact branch and bound.  Node order of checks follows ProcessTask
// (src/helpers.cc:443-512): min degree of C when l<=|C|<=h, rule 3, rules
// 1+2, upper bounds U_d/U_nr, ustar (src/helpers.cc:1104-1168), the set
// dominated by ustar (FindDoms, src/helpers.cc:545-670) and the branches
// written by Expand (src/helpers.cc:1171-1391), explored in the order
//   A   = (C+u*,      R-u*-D)
//   B_i = (C+u*+d_i,  R-u*-d_0..d_{i-1})   i = 0..|D|-1
//   Z   = (C,         R-u*-D)
// Rules 3 and 1/2 run to a fixpoint inside a node; domination is tested
// against the neighbours still in C or R, for the frontier vertices (R with
// dC > 0) adjacent to ustar: a sound subset of FindDoms' candidates (any
// vertex whose live neighbours all lie in N[ustar] may be dropped from the
// "exclude ustar" branch); the search stops as soon as
// kl == ku = min(core[q], h-1), which no solution can exceed.
// ---------------------------------------------------------------------------
struct Params {
  u32 l, h, limit_doms;
  u32 red1, red2, red3, prun1, prun2;
  u32 use_doms;
  u32 pair = 0;
  u32 root_filter = 0;  // level roots start with R = vertices that can matter at kl
  double deadline = 0;  // steady-clock seconds after which the search stops (0 = none)
  u64 node_cap;
  u32 dmax;
  u32 ku;
};

enum { ST_R = 0, ST_C = 1, ST_X = 2 };
enum { NODE_PRUNE = 0, NODE_BRANCH = 1 };

static double steady_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static bool past(const Params& P) { return P.deadline > 0 && steady_s() > P.deadline; }

// Level root filter: with levels on, a level's root starts with R = the
// kept vertices other than q with core > kl and dist <= ubd_node(kl, h);
// the rest start removed.  None of them can lie in an H with min degree >
// kl and |H| <= h (core(v) <= kl, or farther from q than such an H's
// diameter), and the root's rule-1/2 closure would remove every one of them.
static bool root_keeps(const Local& L, const Params& P, u32 kl, u32 v) {
  return L.core[v] > kl && L.dist[v] <= ubd_node(kl, P.h);
}

// (a) recompute solver: every node rebuilds dC/dR from the explicit (C, R)
// lists, as the reference's kernels read them from the task arrays.
struct RecomputeSolver {
  const Local& L;
  Params P;
  u32 kl;
  std::vector<u32> witness;
  bool have_witness = false;
  u64 nodes = 0;
  std::vector<unsigned char> st;
  std::vector<u32> dC, dR;
  std::vector<unsigned char> mark;
  struct Task {
    std::vector<u32> C, R;
  };

  u32 real = 0;  // min degree of the best witness (or the heuristic's kl0)
  u32 flr = 0;   // level floor: rules and offers use kl = max(real, flr)
  void set_floor(u32 f) {
    flr = f;
    kl = std::max(real, flr);
  }
  RecomputeSolver(const Local& L_, const Params& P_, u32 kl0) : L(L_), P(P_), kl(kl0), real(kl0) {
    st.assign(L.K, ST_X);
    dC.assign(L.K, 0);
    dR.assign(L.K, 0);
    mark.assign(L.K, 0);
  }
  void degrees(const std::vector<u32>& C, const std::vector<u32>& R) {
    auto f = [&](u32 v) {
      u32 c = 0, r = 0;
      for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
        unsigned char s = st[L.adj[j]];
        c += (s == ST_C);
        r += (s == ST_R);
      }
      dC[v] = c;
      dR[v] = r;
    };
    for (u32 v : C) f(v);
    for (u32 v : R) f(v);
  }
  bool cond(u32 v, u32 Cs, u32 ubD) const {
    if (P.red1 && std::min(dR[v] + dC[v], dC[v] + P.h - Cs - 1) <= kl) return true;
    if (P.red2 && ubD < L.dist[v]) return true;
    return false;
  }
  bool tight_exists(const std::vector<u32>& C) const {
    for (u32 u : C)
      if (dC[u] + dR[u] == kl + 1 && dR[u] > 0) return true;
    return false;
  }

  void process(Task& t, std::vector<Task>& out) {
    nodes++;
    std::vector<u32>& C = t.C;
    std::vector<u32>& R = t.R;
    for (u32 v : C) st[v] = ST_C;
    for (u32 v : R) st[v] = ST_R;
    auto cleanup = [&]() {
      for (u32 v : C) st[v] = ST_X;
      for (u32 v : R) st[v] = ST_X;
    };
    const u32 h = P.h;
    while (true) {
      degrees(C, R);
      if (C.size() >= P.l && C.size() <= h) {
        u32 mn = UINT32_MAX;
        for (u32 u : C) mn = std::min(mn, dC[u]);
        if (mn >= kl && (mn > real || (!have_witness && mn >= real))) {
          real = mn;
          kl = std::max(real, flr);
          witness = C;
          have_witness = true;
        }
      }
      if (C.size() >= h) return cleanup();
      if (P.red3 && tight_exists(C)) {
        std::vector<u32> add;
        for (u32 u : C)
          if (dC[u] + dR[u] == kl + 1)
            for (u32 j = L.off[u]; j < L.off[u + 1]; j++) {
              u32 w = L.adj[j];
              if (st[w] == ST_R) {
                st[w] = ST_C;
                add.push_back(w);
              }
            }
        std::vector<u32> R2;
        for (u32 v : R)
          if (st[v] == ST_R) R2.push_back(v);
        R.swap(R2);
        for (u32 w : add) C.push_back(w);
        if (C.size() > h) return cleanup();
        continue;
      }
      u32 ubD = ubd_node(kl, h);
      u32 Cs = (u32)C.size();
      std::vector<u32> stack;
      for (u32 v : R)
        if (cond(v, Cs, ubD)) {
          mark[v] = 1;
          stack.push_back(v);
        }
      bool removed = !stack.empty();
      while (!stack.empty()) {
        u32 v = stack.back();
        stack.pop_back();
        st[v] = ST_X;
        mark[v] = 0;
        for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
          u32 w = L.adj[j];
          if (st[w] != ST_X) {
            dR[w]--;
            if (st[w] == ST_R && !mark[w] && cond(w, Cs, ubD)) {
              mark[w] = 1;
              stack.push_back(w);
            }
          }
        }
      }
      if (removed) {
        std::vector<u32> R2;
        for (u32 v : R)
          if (st[v] == ST_R) R2.push_back(v);
        R.swap(R2);
        if (P.red3 && tight_exists(C)) continue;
      }
      break;
    }
    const u32 Cs = (u32)C.size();
    if (Cs + R.size() < P.l || R.empty()) return cleanup();
    u32 ub = UINT32_MAX;
    if (P.prun1)
      for (u32 u : C) ub = std::min(ub, std::min(dC[u] + h - Cs, dC[u] + dR[u]));
    if (P.prun2 && ub > kl) {
      std::vector<u32> cv, rh(h + 2, 0);
      for (u32 u : C) cv.push_back(dC[u]);
      for (u32 v : R) rh[dC[v]]++;
      u32 k = std::min<u32>(h - Cs, (u32)R.size());
      ub = std::min(ub, unr_bound(cv, rh, k));
    }
    if (ub <= kl) return cleanup();
    int us = -1;
    u64 best = 0;
    for (u32 v : R) {  // R is ascending: strict > keeps the lowest id on ties
      if (dC[v] == 0) continue;
      u64 s = 0;
      for (u32 j = L.off[v]; j < L.off[v + 1]; j++)
        if (st[L.adj[j]] == ST_C) s += recip32(dC[L.adj[j]]);
      s += rterm(dR[v], P.dmax);
      if (s > best) {
        best = s;
        us = (int)v;
      }
    }
    if (us < 0) return cleanup();
    const u32 u = (u32)us;
    std::vector<std::pair<u64, u32>> D;
    if (P.use_doms && P.limit_doms > 0) {
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) mark[L.adj[j]] = 1;
      for (u32 v : R) {
        // candidates: frontier vertices adjacent to ustar
        if (v == u || dC[v] == 0 || !mark[v]) continue;
        bool dom = true;
        u64 s = 0;
        for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
          u32 w = L.adj[j];
          if (st[w] == ST_X) continue;
          if (w != u && !mark[w]) {
            dom = false;
            break;
          }
          if (st[w] == ST_C && dC[w] != 0) s += recip32(dC[w]);
        }
        if (dom) D.push_back({s + rterm(dR[v], P.dmax), v});
      }
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) mark[L.adj[j]] = 0;
      std::sort(D.begin(), D.end(), [](const std::pair<u64, u32>& a, const std::pair<u64, u32>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
      });
      if (D.size() > P.limit_doms) D.resize(P.limit_doms);
    }
    cleanup();
    std::vector<u32> dset;
    for (auto& x : D) dset.push_back(x.second);
    std::sort(dset.begin(), dset.end());
    std::vector<u32> Rp;
    for (u32 v : R)
      if (v != u && !std::binary_search(dset.begin(), dset.end(), v)) Rp.push_back(v);
    Task A;
    A.C = C;
    A.C.push_back(u);
    A.R = Rp;
    out.push_back(std::move(A));
    for (size_t i = 0; i < D.size(); i++) {
      Task B;
      B.C = C;
      B.C.push_back(u);
      B.C.push_back(D[i].second);
      B.R = Rp;
      for (size_t j = i + 1; j < D.size(); j++) B.R.push_back(D[j].second);
      std::sort(B.R.begin(), B.R.end());
      out.push_back(std::move(B));
    }
    Task Z;
    Z.C = C;
    Z.R = Rp;
    out.push_back(std::move(Z));
  }

  int run(u32 qloc) {
    std::vector<Task> stack;
    Task root;
    root.C.push_back(qloc);
    for (u32 v = 0; v < L.K; v++)
      if (v != qloc && (!P.root_filter || root_keeps(L, P, kl, v))) root.R.push_back(v);
    stack.push_back(std::move(root));
    std::vector<Task> kids;
    while (!stack.empty()) {
      if (real >= P.ku) return 0;
      if (nodes >= P.node_cap || past(P)) return 1;
      Task t = std::move(stack.back());
      stack.pop_back();
      kids.clear();
      process(t, kids);
      for (int i = (int)kids.size() - 1; i >= 0; i--) stack.push_back(std::move(kids[i]));
    }
    return 0;
  }
};

// (b) trail solver: one depth-first walker with incremental dC/dR and an
// undo trail, the layout the GPU engine keeps per wavefront.  Must expand
// exactly the same nodes as (a).
struct TrailSolver {
  const Local& L;
  Params P;
  u32 kl;
  std::vector<u32> witness;
  bool have_witness = false;
  u64 nodes = 0;
  std::vector<unsigned char> st, queued, nmark;
  std::vector<u32> dC, dR, hist, Clist, trail, work;
  u32 nR = 0;
  struct Frame {
    u32 mark, ustar, nD, next, kl;
    u32 D[32];
  };
  std::vector<Frame> frames;
  u32 ustar = 0;
  std::vector<u32> pcnt;            // common-neighbour counts (pair rule)
  std::vector<unsigned char> padj;  // 1 = neighbour of the C vertex being checked
  std::vector<u32> ptouch, prl, prm;
  u32 closed_kl = 0;
  std::vector<std::pair<u64, u32>> Dv;
  std::vector<u32> front;
  std::vector<unsigned char> inF;
  std::vector<u64> acc, accd;

  u32 real = 0;  // min degree of the best witness (or the heuristic's kl0)
  u32 flr = 0;   // level floor: rules and offers use kl = max(real, flr)
  void set_floor(u32 f) {
    flr = f;
    kl = std::max(real, flr);
  }
  TrailSolver(const Local& L_, const Params& P_, u32 kl0) : L(L_), P(P_), kl(kl0), real(kl0) {}

  void moveC(u32 v) {
    st[v] = ST_C;
    hist[dC[v]]--;
    nR--;
    Clist.push_back(v);
    for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
      u32 w = L.adj[j];
      if (st[w] == ST_R) {
        hist[dC[w]]--;
        hist[dC[w] + 1]++;
      }
      dC[w]++;
      dR[w]--;
    }
    trail.push_back(v | 0x80000000u);
  }
  void removeR(u32 v) {
    st[v] = ST_X;
    hist[dC[v]]--;
    nR--;
    for (u32 j = L.off[v]; j < L.off[v + 1]; j++) dR[L.adj[j]]--;
    trail.push_back(v);
  }
  void undo(u32 m) {
    while (trail.size() > m) {
      u32 e = trail.back();
      trail.pop_back();
      u32 v = e & 0x7fffffffu;
      if (e & 0x80000000u) {
        Clist.pop_back();
        st[v] = ST_R;
        hist[dC[v]]++;
        nR++;
        for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
          u32 w = L.adj[j];
          if (st[w] == ST_R) {
            hist[dC[w]]--;
            hist[dC[w] - 1]++;
          }
          dC[w]--;
          dR[w]++;
        }
      } else {
        st[v] = ST_R;
        hist[dC[v]]++;
        nR++;
        for (u32 j = L.off[v]; j < L.off[v + 1]; j++) dR[L.adj[j]]++;
      }
    }
  }
  bool cond(u32 v, u32 Cs, u32 ubD) const {
    if (P.red1 && std::min(dR[v] + dC[v], dC[v] + P.h - Cs - 1) <= kl) return true;
    if (P.red2 && ubD < L.dist[v]) return true;
    return false;
  }
  bool tight_exists() const {
    for (u32 u : Clist)
      if (dC[u] + dR[u] == kl + 1 && dR[u] > 0) return true;
    return false;
  }

  // Rule 1/2 closure.  At a parent node no R vertex met the condition
  // after its closure; in a child only these can: neighbours of vertices
  // removed since `from` (their dR dropped), vertices with
  // dC <= kl + |C| + 1 - h (|C| grew), or everything when kl changed.
  bool r12(u32 from, bool full) {
    const u32 Cs = (u32)Clist.size();
    const u32 ubD = ubd_node(kl, P.h);
    work.clear();
    auto push = [&](u32 w) {
      if (st[w] == ST_R && !queued[w] && cond(w, Cs, ubD)) {
        queued[w] = 1;
        work.push_back(w);
      }
    };
    i64 T = (i64)kl + Cs + 1 - (i64)P.h;
    if (!full && P.red1 && T >= 0) {
      u32 c = 0;
      for (i64 d = 0; d <= T && d < (i64)hist.size(); d++) c += hist[d];
      if (c) full = true;
    }
    if (full) {
      for (u32 v = 0; v < L.K; v++) push(v);
    } else {
      for (size_t t = from; t < trail.size(); t++)
        if (!(trail[t] & 0x80000000u)) {
          u32 v = trail[t];
          for (u32 j = L.off[v]; j < L.off[v + 1]; j++) push(L.adj[j]);
        }
    }
    bool removed = !work.empty();
    while (!work.empty()) {
      u32 v = work.back();
      work.pop_back();
      queued[v] = 0;
      removeR(v);
      for (u32 j = L.off[v]; j < L.off[v + 1]; j++) push(L.adj[j]);
    }
    return removed;
  }

  // Bounds on the pair rule's work, so that a node whose C holds hubs of a
  // large kept set stays cheap: a C vertex whose live neighbours' lists hold
  // more than PAIR_MAX_SCAN entries, or more than what is left of the node's
  // PAIR_NODE_SCAN entries, is skipped (a skipped check only weakens the
  // pruning).
  static constexpr u64 PAIR_MAX_SCAN = 1u << 20;
  static constexpr u64 PAIR_NODE_SCAN = 1u << 24;
  u64 pair_left = 0;

  // Pair rule.  Two vertices u, v of an H with min degree >= k = kl+1 and
  // |H| <= h each have >= k - [u~v] neighbours in H \ {u, v}, a set of at
  // most h-2 vertices, so they share >= 2k - h (adjacent) or 2k - h + 2 (not
  // adjacent) neighbours in H, all of them live (C or R).  For each C vertex
  // u in C order: count the live common neighbours of u with every live
  // vertex (two-hop scatter over the live neighbours of u), prune the node
  // if another C vertex falls short, else remove every R vertex that falls
  // short before the next u.  Returns -1 prune, 1 when something was
  // removed, 0 otherwise.
  int pair_step() {
    const long tadj = 2L * (long)(kl + 1) - (long)P.h;
    if (tadj + 2 <= 0) return 0;
    if (pcnt.size() != L.K) {
      pcnt.assign(L.K, 0);
      padj.assign(L.K, 0);
    }
    int any = 0;
    prl.clear();
    for (u32 v = 0; v < L.K; v++)
      if (st[v] == ST_R) prl.push_back(v);
    for (size_t ci = 0; ci < Clist.size(); ci++) {
      const u32 u = Clist[ci];
      u64 scan = 0;
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) {
        const u32 w = L.adj[j];
        if (st[w] != ST_X) scan += L.off[w + 1] - L.off[w];
      }
      if (scan > PAIR_MAX_SCAN || scan > pair_left) continue;
      pair_left -= scan;
      ptouch.clear();
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) padj[L.adj[j]] = 1;
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) {
        const u32 w = L.adj[j];
        if (st[w] == ST_X) continue;
        for (u32 t = L.off[w]; t < L.off[w + 1]; t++) {
          const u32 x = L.adj[t];
          if (st[x] == ST_X) continue;
          if (pcnt[x]++ == 0) ptouch.push_back(x);
        }
      }
      auto short_of = [&](u32 v) { return (long)pcnt[v] < (padj[v] ? tadj : tadj + 2); };
      bool prune = false;
      for (u32 c : Clist)
        if (c != u && short_of(c)) prune = true;
      prm.clear();
      if (!prune) {
        // R vertices in ascending order; the survivors stay in prl
        size_t k = 0;
        for (u32 v : prl) {
          if (short_of(v))
            prm.push_back(v);
          else
            prl[k++] = v;
        }
        prl.resize(k);
      }
      for (u32 x : ptouch) pcnt[x] = 0;
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) padj[L.adj[j]] = 0;
      if (prune) return -1;
      for (u32 v : prm) removeR(v);
      if (!prm.empty()) any = 1;
    }
    return any;
  }

  // kl_seen: kl at the parent's closure (UINT32_MAX at the root)
  int process(u32 from, u32 kl_seen) {
    nodes++;
    const u32 h = P.h;
    bool full = (kl != kl_seen);
    pair_left = PAIR_NODE_SCAN;
    while (true) {
      u32 Cs = (u32)Clist.size();
      if (Cs >= P.l && Cs <= h) {
        u32 mn = UINT32_MAX;
        for (u32 u : Clist) mn = std::min(mn, dC[u]);
        if (mn >= kl && (mn > real || (!have_witness && mn >= real))) {
          real = mn;
          const u32 ne = std::max(real, flr);
          if (ne != kl) full = true;
          kl = ne;
          witness = Clist;
          have_witness = true;
        }
      }
      if (Cs >= h) return NODE_PRUNE;
      if (P.red3 && tight_exists()) {
        work.clear();
        for (u32 u : Clist)
          if (dC[u] + dR[u] == kl + 1 && dR[u] > 0) work.push_back(u);
        // stop as soon as |C| > h: the node is pruned either way, and dC of R
        // vertices then stays <= h + 1 (histogram bound)
        for (u32 u : work) {
          for (u32 j = L.off[u]; j < L.off[u + 1] && Clist.size() <= h; j++)
            if (st[L.adj[j]] == ST_R) moveC(L.adj[j]);
          if (Clist.size() > h) return NODE_PRUNE;
        }
        continue;
      }
      bool removed = r12(from, full);
      from = (u32)trail.size();
      full = false;
      if (removed && P.red3 && tight_exists()) continue;
      if (P.pair) {
        const int pr = pair_step();
        if (pr < 0) return NODE_PRUNE;
        if (pr > 0) continue;
      }
      break;
    }
    closed_kl = kl;
    const u32 Cs = (u32)Clist.size();
    if (Cs + nR < P.l || nR == 0) return NODE_PRUNE;
    u32 ub = UINT32_MAX;
    if (P.prun1)
      for (u32 u : Clist) ub = std::min(ub, std::min(dC[u] + h - Cs, dC[u] + dR[u]));
    if (P.prun2 && ub > kl) {
      std::vector<u32> cv;
      for (u32 u : Clist) cv.push_back(dC[u]);
      u32 k = std::min<u32>(h - Cs, nR);
      ub = std::min(ub, unr_bound(cv, hist, k));
    }
    if (ub <= kl) return NODE_PRUNE;
    // scores by scattering from C: frontier = R vertices with dC > 0
    front.clear();
    for (u32 u : Clist) {
      u64 r = recip32(dC[u]);
      u64 rd = dC[u] ? r : 0;
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++) {
        u32 w = L.adj[j];
        if (st[w] != ST_R) continue;
        if (!inF[w]) {
          inF[w] = 1;
          acc[w] = 0;
          accd[w] = 0;
          front.push_back(w);
        }
        acc[w] += r;
        accd[w] += rd;
      }
    }
    int us = -1;
    u64 best = 0;
    for (u32 v : front) {
      u64 sc = acc[v] + rterm(dR[v], P.dmax);
      if (sc > best || (sc == best && (u32)us > v)) {
        best = sc;
        us = (int)v;
      }
    }
    if (us < 0) {
      for (u32 v : front) inF[v] = 0;
      return NODE_PRUNE;
    }
    ustar = (u32)us;
    Dv.clear();
    if (P.use_doms && P.limit_doms > 0) {
      const u32 lu = dC[ustar] + dR[ustar];
      for (u32 j = L.off[ustar]; j < L.off[ustar + 1]; j++) nmark[L.adj[j]] = 1;
      for (u32 v : front) {
        // candidates: frontier vertices adjacent to ustar
        if (v == ustar || !nmark[v] || dC[v] + dR[v] > lu + 1) continue;
        bool dom = true;
        for (u32 j = L.off[v]; j < L.off[v + 1]; j++) {
          u32 w = L.adj[j];
          if (st[w] == ST_X) continue;
          if (w != ustar && !nmark[w]) {
            dom = false;
            break;
          }
        }
        if (dom) Dv.push_back({accd[v] + rterm(dR[v], P.dmax), v});
      }
      for (u32 j = L.off[ustar]; j < L.off[ustar + 1]; j++) nmark[L.adj[j]] = 0;
      std::sort(Dv.begin(), Dv.end(), [](const std::pair<u64, u32>& a, const std::pair<u64, u32>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
      });
      if (Dv.size() > P.limit_doms) Dv.resize(P.limit_doms);
    }
    for (u32 v : front) inF[v] = 0;
    return NODE_BRANCH;
  }

  int run(u32 qloc) {
    const u32 K = L.K;
    st.assign(K, ST_R);
    queued.assign(K, 0);
    nmark.assign(K, 0);
    inF.assign(K, 0);
    acc.assign(K, 0);
    accd.assign(K, 0);
    dC.assign(K, 0);
    dR.assign(K, 0);
    hist.assign(P.h + 3, 0);
    if (P.root_filter)
      for (u32 v = 0; v < K; v++)
        if (v != qloc && !root_keeps(L, P, kl, v)) st[v] = ST_X;
    nR = 0;
    for (u32 v = 0; v < K; v++) {
      u32 d = 0;
      for (u32 j = L.off[v]; j < L.off[v + 1]; j++) d += st[L.adj[j]] != ST_X;
      dR[v] = d;
      nR += st[v] == ST_R;
    }
    hist[0] = nR;
    moveC(qloc);
    trail.clear();
    u32 from = 0, kl_seen = UINT32_MAX;
    while (true) {
      if (real >= P.ku) return 0;
      if (nodes >= P.node_cap || past(P)) return 1;
      int r = process(from, kl_seen);
      if (r == NODE_BRANCH) {
        Frame f;
        f.mark = (u32)trail.size();
        f.ustar = ustar;
        f.nD = (u32)Dv.size();
        f.next = 0;
        f.kl = closed_kl;

        for (u32 i = 0; i < f.nD; i++) f.D[i] = Dv[i].second;
        frames.push_back(f);
        from = f.mark;
        kl_seen = f.kl;
        for (u32 i = 0; i < f.nD; i++) removeR(f.D[i]);
        moveC(f.ustar);
        continue;
      }
      // backtrack to the next pending child
      if (frames.empty()) return 0;
      Frame& f = frames.back();
      undo(f.mark);
      from = f.mark;
      kl_seen = f.kl;
      f.next++;
      if (f.next <= f.nD) {
        u32 i = f.next - 1;
        for (u32 j = 0; j < i; j++) removeR(f.D[j]);
        moveC(f.ustar);
        moveC(f.D[i]);
        continue;
      }
      Frame z = f;
      frames.pop_back();
      removeR(z.ustar);
      for (u32 j = 0; j < z.nD; j++) removeR(z.D[j]);
    }
  }
};

// ---------------------------------------------------------------------------
// Independent second solver: plain enumeration of connected sets containing
// q (include/exclude on a frontier vertex), pruned only by size and by the
// degree bound min over C of degC + (h-|C|).  No rules, no domination.
// ---------------------------------------------------------------------------
struct PlainSolver {
  const Local& L;
  u32 l, h;
  u32 best;
  bool found = false;
  u64 nodes = 0, cap;
  std::vector<unsigned char> st;  // 0 avail, 1 in C, 2 excluded
  std::vector<u32> C;
  PlainSolver(const Local& L_, u32 l_, u32 h_, u32 kl0, u64 cap_)
      : L(L_), l(l_), h(h_), best(kl0), cap(cap_) {
    st.assign(L.K, 0);
  }
  u32 degc(u32 v) {
    u32 c = 0;
    for (u32 j = L.off[v]; j < L.off[v + 1]; j++) c += st[L.adj[j]] == 1;
    return c;
  }
  bool rec() {
    if (++nodes > cap) return false;
    u32 mn = UINT32_MAX;
    for (u32 u : C) mn = std::min(mn, degc(u));
    if (C.size() >= l && C.size() <= h && (mn > best || (!found && mn >= best))) {
      best = std::max(best, mn);
      found = true;
    }
    if (C.size() >= h) return true;
    // bound: every C vertex can gain at most h-|C| more neighbours
    if (mn + (h - (u32)C.size()) <= best && found) return true;
    // pick the lowest-id available vertex adjacent to C
    int pick = -1;
    for (u32 u : C) {
      for (u32 j = L.off[u]; j < L.off[u + 1]; j++)
        if (st[L.adj[j]] == 0 && (pick < 0 || L.adj[j] < (u32)pick)) pick = (int)L.adj[j];
    }
    if (pick < 0) return true;
    u32 p = (u32)pick;
    st[p] = 1;
    C.push_back(p);
    bool ok = rec();
    C.pop_back();
    if (!ok) {
      st[p] = 0;
      return false;
    }
    st[p] = 2;
    ok = rec();
    st[p] = 0;
    return ok;
  }
};

// ---------------------------------------------------------------------------
// C ABI for the tests (ctypes).
// ---------------------------------------------------------------------------
extern "C" {

struct orc_graph {
  Graph g;
  std::vector<u32> core;
  std::vector<u32> degree;
};

orc_graph* orc_graph_from_csr(u32 n, const i64* off, const int32_t* adj) {
  orc_graph* o = new orc_graph;
  o->g.n = n;
  o->g.off.assign(off, off + n + 1);
  o->g.adj.assign((const u32*)adj, (const u32*)adj + off[n]);
  o->g.dmax = 0;
  o->degree.resize(n);
  for (u32 v = 0; v < n; v++) {
    o->degree[v] = o->g.deg(v);
    o->g.dmax = std::max(o->g.dmax, o->degree[v]);
  }
  core_numbers(o->g, o->core);
  return o;
}

orc_graph* orc_graph_load(const char* path, int* err) {
  orc_graph* o = new orc_graph;
  int rc = load_edge_list(path, o->g);
  if (err) *err = rc;
  if (rc) {
    delete o;
    return nullptr;
  }
  o->degree.resize(o->g.n);
  for (u32 v = 0; v < o->g.n; v++) o->degree[v] = o->g.deg(v);
  core_numbers(o->g, o->core);
  return o;
}

void orc_graph_free(orc_graph* o) { delete o; }
u32 orc_graph_n(orc_graph* o) { return o->g.n; }
i64 orc_graph_nnz(orc_graph* o) { return o->g.off[o->g.n]; }
u32 orc_graph_dmax(orc_graph* o) { return o->g.dmax; }
void orc_graph_csr(orc_graph* o, i64* off, int32_t* adj) {
  std::copy(o->g.off.begin(), o->g.off.end(), off);
  std::copy(o->g.adj.begin(), o->g.adj.end(), (u32*)adj);
}
void orc_graph_core(orc_graph* o, u32* core) { std::copy(o->core.begin(), o->core.end(), core); }
void orc_bfs(orc_graph* o, u32 q, u32* dist) {
  std::vector<u32> d;
  bfs_full(o->g, q, d);
  std::copy(d.begin(), d.end(), dist);
}
u32 orc_ubd_host(u32 kl, u32 h) { return ubd_host(kl, h); }
// U_nr water-filling on explicit values (for tests): cvals = dC over C,
// rvals = dC over R, k rounds.
u32 orc_unr_bound(const u32* cvals, u32 m, const u32* rvals, u32 nr, u32 k) {
  std::vector<u32> cv(cvals, cvals + m);
  u32 mx = 0;
  for (u32 i = 0; i < nr; i++) mx = std::max(mx, rvals[i]);
  std::vector<u32> rh(mx + 2, 0);
  for (u32 i = 0; i < nr; i++) rh[rvals[i]]++;
  return unr_bound(cv, rh, k);
}
u32 orc_ubd_node(u32 kl, u32 h) { return ubd_node(kl, h); }

// heuristic: which = 2, 3, 4 or 0 (CSSC_heu).  Writes up to cap witness
// vertices and returns kl; *wsize gets the witness size.
u32 orc_heuristic(orc_graph* o, u32 which, u32 l, u32 h, u32 q, u32* wit, u32 cap, u32* wsize) {
  std::vector<u32> H;
  u32 kl = 0;
  if (which == 2)
    heu2(o->g, l, h, q, H, kl);
  else if (which == 3)
    heu3(o->g, o->degree, o->g.dmax, l, h, q, H, kl);
  else if (which == 4)
    heu4(o->g, l, h, q, H, kl);
  else
    kl = cssc_heu(o->g, o->degree, l, h, q, H);
  *wsize = (u32)H.size();
  for (u32 i = 0; i < H.size() && i < cap; i++) wit[i] = H[i];
  return kl;
}

// Preprocessing summary: returns K; fills up to cap kept original ids
// (ascending), local offsets (K+1) and adjacency (up to adj_cap entries).
u32 orc_preprocess(orc_graph* o, u32 q, u32 kl0, u32 ubD, u32* kept, u32 cap, u32* loff,
                   u32* ladj, u64 adj_cap, u32* ldist, u64* nnz_out) {
  Local L;
  preprocess(o->g, o->core, q, kl0, ubD, L);
  if (L.K <= cap) {
    std::copy(L.orig.begin(), L.orig.end(), kept);
    std::copy(L.off.begin(), L.off.end(), loff);
    std::copy(L.dist.begin(), L.dist.end(), ldist);
    if (L.adj.size() <= adj_cap) std::copy(L.adj.begin(), L.adj.end(), ladj);
  }
  *nnz_out = L.adj.size();
  return L.K;
}

struct orc_result {
  u32 kl;       // final optimum (lower bound if capped)
  u32 heu_kl;   // heuristic kl (0 when the heuristic is off)
  u32 ku;       // min(core[q], h-1)
  int32_t status;  // 0 ok, 1 heuristic optimal, 2 node cap, 3 infeasible, -1 invalid
  u64 nodes;
  u32 K;        // kept vertices
  u32 ubD;
  u32 wsize;
  u32 levels;   // level searches run (1 without levels)
};

static int solve_impl(orc_graph* o, u32 l, u32 h, u32 q, u32 heu, u32 limit_doms, u32 flags,
                      u64 node_cap, double seconds, orc_result* res, u32* wit, u32 wcap);

// flags bit0 red1, bit1 red2, bit2 red3, bit3 prun1, bit4 prun2, bit5 doms,
// bit6 selects the recompute solver instead of the trail solver, bit7 levels,
// bit8 pair rule (trail solver only)
int orc_solve(orc_graph* o, u32 l, u32 h, u32 q, u32 heu, u32 limit_doms, u32 flags,
              u64 node_cap, orc_result* res, u32* wit, u32 wcap) {
  return solve_impl(o, l, h, q, heu, limit_doms, flags, node_cap, 0.0, res, wit, wcap);
}
// The same with a wall-clock limit (seconds, 0 = none): a search still
// running when it passes ends CAPPED like one that reaches node_cap.
int orc_solve_timed(orc_graph* o, u32 l, u32 h, u32 q, u32 heu, u32 limit_doms, u32 flags,
                    u64 node_cap, double seconds, orc_result* res, u32* wit, u32 wcap) {
  return solve_impl(o, l, h, q, heu, limit_doms, flags, node_cap, seconds, res, wit, wcap);
}
}  // extern "C"

static int solve_impl(orc_graph* o, u32 l, u32 h, u32 q, u32 heu, u32 limit_doms, u32 flags,
                      u64 node_cap, double seconds, orc_result* res, u32* wit, u32 wcap) {
  memset(res, 0, sizeof(*res));
  const Graph& g = o->g;
  if (h == 0 || q >= g.n || limit_doms > 32) {
    res->status = -1;
    return -1;
  }
  u32 ku = std::min(o->core[q], h - 1);
  res->ku = ku;
  std::vector<u32> H;
  u32 kl = 0;
  if (heu) kl = cssc_heu(g, o->degree, l, h, q, H);
  res->heu_kl = kl;
  if (l > h) {
    res->kl = 0;
    res->status = 3;
    return 0;
  }
  if (heu && kl == ku) {
    res->kl = kl;
    res->status = 1;
    if (kl == 0) {
      // a connected H of two or more vertices has min degree >= 1
      res->status = l >= 2 ? 3 : 1;
      H.assign(1, q);
      if (l >= 2) H.clear();
    }
    res->wsize = (u32)H.size();
    for (u32 i = 0; i < H.size() && i < wcap; i++) wit[i] = H[i];
    return 0;
  }
  u32 ubD = ubd_host(kl, h);
  res->ubD = ubD;
  Local L;
  preprocess(g, o->core, q, kl, ubD, L);
  res->K = L.K;
  Params P;
  P.l = l;
  P.h = h;
  P.limit_doms = limit_doms;
  P.red1 = flags & 1;
  P.red2 = (flags >> 1) & 1;
  P.red3 = (flags >> 2) & 1;
  P.prun1 = (flags >> 3) & 1;
  P.prun2 = (flags >> 4) & 1;
  P.use_doms = (flags >> 5) & 1;
  P.pair = (flags >> 8) & 1;
  P.root_filter = (flags >> 7) & 1;
  P.node_cap = node_cap;
  P.dmax = g.dmax;
  P.ku = ku;
  P.deadline = seconds > 0 ? steady_s() + seconds : 0;
  // Levels (flag bit 7): search for min degree > flr with flr = ku-1 first,
  // then lower floors, down to the heuristic's kl.  A level whose search
  // ends without a solution above its floor proves opt <= flr; the first
  // level that finds one returns the optimum, as its search was exact for
  // every solution above its floor.  A level that ends empty also lowers the
  // stopping bound ku to its floor (opt <= flr), so the next level stops as
  // soon as it finds a solution at that bound.
  const bool levels = (flags >> 7) & 1;
  u32 flr = (levels && ku > kl + 1) ? ku - 1 : 0;
  u32 real = kl;
  bool have_w = kl > 0;
  std::vector<u32> wloc;
  int capped = 0;
  u64 nodes = 0;
  while (true) {
    P.node_cap = node_cap > nodes ? node_cap - nodes : 0;
    if ((flags >> 6) & 1) {
      RecomputeSolver S(L, P, real);
      S.have_witness = have_w;
      S.set_floor(flr);
      capped = S.run(L.qloc);
      if (S.real > real || (!have_w && S.have_witness)) wloc = S.witness;
      real = S.real;
      have_w = S.have_witness;
      nodes += S.nodes;
    } else {
      TrailSolver S(L, P, real);
      S.have_witness = have_w;
      S.set_floor(flr);
      capped = S.run(L.qloc);
      if (S.real > real || (!have_w && S.have_witness)) wloc = S.witness;
      real = S.real;
      have_w = S.have_witness;
      nodes += S.nodes;
    }
    res->levels++;
    if (capped || real >= flr) break;
    P.ku = std::min(P.ku, flr);
    flr--;
  }
  const u32 final_kl = real;
  res->kl = final_kl;
  res->nodes = nodes;
  if (!wloc.empty()) {
    H.clear();
    for (u32 v : wloc) H.push_back(L.orig[v]);
  }
  if (final_kl == 0) {
    H.clear();
    if (l <= 1) H.push_back(q);
  }
  res->wsize = (u32)H.size();
  for (u32 i = 0; i < H.size() && i < wcap; i++) wit[i] = H[i];
  if (capped)
    res->status = 2;
  else if (final_kl == 0 && l >= 2)
    res->status = 3;
  else
    res->status = 0;
  return 0;
}

extern "C" {
// Independent plain solver (no rules, no domination); returns the optimum
// over the same kept set K (kl0 = 0, ubD = h-1 so K is q's component
// restricted to core > 0), or UINT32_MAX when the node cap is hit.
u32 orc_solve_plain(orc_graph* o, u32 l, u32 h, u32 q, u64 node_cap, u64* nodes) {
  if (h == 0 || l > h) return 0;
  Local L;
  preprocess(o->g, o->core, q, 0, h - 1, L);
  PlainSolver S(L, l, h, 0, node_cap);
  S.st[L.qloc] = 1;
  S.C.push_back(L.qloc);
  bool ok = S.rec();
  *nodes = S.nodes;
  if (!ok) return UINT32_MAX;
  return S.found ? S.best : 0;
}

}  // extern "C"
