// Host greedy lower bound kl for one query: the reference's CSSC_heu
// (inc/heuristic.h:321-348), the best of three greedy constructions.
//
// The values (kl and H) are those of the reference's heu2/heu3/heu4,
// including the std::priority_queue tie order; the work per step is
// reduced without changing any compared quantity:
//   heu2 (inc/heuristic.h:8-120): a vertex's degree into H is kept as a
//     running count instead of rescanning its adjacency;
//   heu3 (inc/heuristic.h:125-239): the connection score sums 1/deg over the
//     vertex's neighbours in H; each vertex keeps its H-neighbours in
//     ascending order (in a 32-byte record with its largest pushed key), so
//     the sum has the same terms in the same order as the reference's
//     adjacency scan (bit-identical doubles), and a joining vertex's degree
//     into H is its list's length; a push no larger than the vertex's
//     largest earlier key is dropped, since it could never be popped before
//     that one;
//   heu4 (inc/heuristic.h:242-317): unchanged except that q's +1
//     (inc/heuristic.h:259) is not counted in the min-degree test when q is
//     popped alone (l <= 1); the reference then reports 1 for an isolated q.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "scs_host.h"

namespace scs {

static const uint32_t INF_DEG = 1000000000u;
// The neighbour loops of heu2 and heu3 are bound by cache misses on the
// n-sized per-vertex arrays; they prefetch the entries of the neighbour this
// many places ahead in the adjacency list.
static const int64_t PF_AHEAD = 16;

void HeuWork::reset(uint32_t n) {
  if (sta.size() != n) {
    sta.assign(n, 0);
    deg.assign(n, 0);
    cnt.assign(n, 0);
    rec3.assign(n, Heu3Rec{});
    inS.assign(n, 0);
  }
}

void HeuWork::clear_touched(unsigned what) {
  // every entry cleared here is a random access into an n-sized array, so
  // only the arrays the finished construction wrote are visited
  for (uint32_t v : touched) {
    sta[v] = 0;
    if (what & CLEAR_CNT) cnt[v] = 0;
    if (what & CLEAR_REC3) {
      rec3[v].key = 0.0;
      rec3[v].len = 0;
    }
    if (what & CLEAR_DEG_INS) {
      deg[v] = 0;
      inS[v] = 0;
    }
  }
  touched.clear();
}

static inline void touch(HeuWork& W, uint32_t v) {
  if (!W.mark_touched(v)) W.touched.push_back(v);
}

bool HeuWork::mark_touched(uint32_t v) {
  // a vertex is recorded once: touched vertices always carry sta != 0 or a
  // count; use a dedicated bit in sta (bit 7)
  if (sta[v] & 0x80) return true;
  sta[v] |= 0x80;
  return false;
}

static void heu2(const HostGraph& g, uint32_t N1, uint32_t N2, uint32_t q, HeuWork& W,
                 std::vector<uint32_t>& H, uint32_t& kl) {
  H.clear();
  kl = 0;
  std::vector<uint32_t> tH;
  uint32_t tkl = 0, tsize = 0;
  std::priority_queue<std::pair<uint32_t, uint32_t>> Q;
  auto S = [&](uint32_t v) { return W.sta[v] & 0x7f; };
  auto setS = [&](uint32_t v, uint8_t s) { W.sta[v] = (W.sta[v] & 0x80) | s; };
  Q.push({0, q});
  touch(W, q);
  setS(q, 1);
  while (!Q.empty()) {
    uint32_t v = Q.top().second;
    Q.pop();
    if (S(v) == 2) continue;
    tH.push_back(v);
    setS(v, 2);
    const int64_t b = g.off[v], e = g.off[v + 1];
    // neighbours are distinct, so each count is final once incremented
    for (int64_t j = b; j < e; j++) {
      uint32_t nb = (uint32_t)g.adj[j];
      if (j + PF_AHEAD < e) __builtin_prefetch(&W.cnt[(uint32_t)g.adj[j + PF_AHEAD]], 1);
      touch(W, nb);
      W.cnt[nb]++;
      if (S(nb) != 2) {
        Q.push({W.cnt[nb], nb});
        setS(nb, 1);
      } else {
        ++W.deg[nb];
        ++W.deg[v];
      }
    }
    if (tH.size() >= N1) {
      uint32_t mindeg = INF_DEG;
      for (uint32_t x : tH) mindeg = std::min(mindeg, W.deg[x]);
      if (mindeg >= tkl) {
        tkl = mindeg;
        tsize = (uint32_t)tH.size();
      }
    }
    if (tH.size() == N2) break;
  }
  H.assign(tH.begin(), tH.begin() + tsize);
  kl = tkl;
  for (uint32_t x : tH) W.deg[x] = 0;  // deg is written only at H members
  W.clear_touched(HeuWork::CLEAR_CNT);
}

// Inserts v into r's ascending list and returns the list.  A list longer
// than three lives in W.pool3 and moves to twice the room when its length
// reaches a power of two >= 8; the pool is emptied at the start of each heu3
// run.
static inline uint32_t* insert3(HeuWork& W, Heu3Rec& r, uint32_t v) {
  uint32_t* L = r.len <= 3 ? r.inl : W.pool3.data() + r.ext;
  if (r.len == 3 || (r.len >= 8 && (r.len & (r.len - 1)) == 0)) {
    const size_t cap = r.len == 3 ? 8 : 2 * (size_t)r.len;
    if (W.top3 + cap > W.pool3.size()) {
      W.pool3.resize(std::max(2 * W.pool3.size(), W.top3 + cap));
      if (r.len > 3) L = W.pool3.data() + r.ext;
    }
    uint32_t* M = W.pool3.data() + W.top3;
    std::memcpy(M, L, r.len * sizeof(uint32_t));
    r.ext = W.top3;
    W.top3 += cap;
    L = M;
  }
  uint32_t i = r.len;
  while (i > 0 && L[i - 1] > v) {
    L[i] = L[i - 1];
    i--;
  }
  L[i] = v;
  r.len++;
  return L;
}

static void heu3(const HostGraph& g, uint32_t N1, uint32_t N2, uint32_t q, HeuWork& W,
                 std::vector<uint32_t>& H, uint32_t& kl) {
  H.clear();
  kl = 0;
  W.top3 = 0;
  std::vector<uint32_t> tH;
  uint32_t tkl = 0, tsize = 0;
  std::priority_queue<std::pair<double, uint32_t>> Q;
  auto S = [&](uint32_t v) { return W.sta[v] & 0x7f; };
  auto setS = [&](uint32_t v, uint8_t s) { W.sta[v] = (W.sta[v] & 0x80) | s; };
  const double dMAX = (double)g.dmax;
  Q.push({0.0, q});
  touch(W, q);
  setS(q, 1);
  while (!Q.empty()) {
    uint32_t v = Q.top().second;
    Q.pop();
    if (S(v) == 2) continue;
    tH.push_back(v);
    setS(v, 2);
    const int64_t b = g.off[v], e = g.off[v + 1];
    {
      // v's list holds exactly its neighbours already in H (each joined H
      // while v was outside it), so the degree updates need no scan of v's
      // adjacency
      const Heu3Rec& rv = W.rec3[v];
      const uint32_t* Lv = rv.len <= 3 ? rv.inl : W.pool3.data() + rv.ext;
      W.deg[v] += rv.len;
      for (uint32_t i = 0; i < rv.len; i++) ++W.deg[Lv[i]];
    }
    // v joins the ascending H-neighbour list of each neighbour outside H
    // (lists of H members are never read again); all degree updates are
    // done above, so each neighbour's score is final once v is in its list
    for (int64_t j = b; j < e; j++) {
      uint32_t nb = (uint32_t)g.adj[j];
      if (j + PF_AHEAD < e) {
        const uint32_t f = (uint32_t)g.adj[j + PF_AHEAD];
        __builtin_prefetch(&W.rec3[f], 1);
        __builtin_prefetch(&g.off[f], 0);
      }
      touch(W, nb);
      if (S(nb) == 2) continue;
      Heu3Rec& r = W.rec3[nb];
      uint32_t* L = insert3(W, r, v);
      double score = 0;
      for (uint32_t i = 0; i < r.len; i++)
        if (W.deg[L[i]] != 0) score += (double)1 / W.deg[L[i]];
      score += (double)(g.off[nb + 1] - g.off[nb]) / dMAX;
      // a push no larger than nb's largest earlier key is never the first
      // of nb's entries popped, so it cannot change the pop sequence
      if (S(nb) == 0 || score > r.key) {
        r.key = score;
        Q.push({score, nb});
      }
      if (S(nb) == 0) setS(nb, 1);
    }
    if (tH.size() >= N1) {
      uint32_t mindeg = INF_DEG;
      for (uint32_t x : tH) mindeg = std::min(mindeg, W.deg[x]);
      if (mindeg >= tkl) {
        tkl = mindeg;
        tsize = (uint32_t)tH.size();
      }
    }
    if (tH.size() == N2) break;
  }
  H.assign(tH.begin(), tH.begin() + tsize);
  kl = tkl;
  for (uint32_t x : tH) W.deg[x] = 0;  // deg is written only at H members
  W.clear_touched(HeuWork::CLEAR_REC3);
}

static void heu4(const HostGraph& g, uint32_t N1, uint32_t N2, uint32_t q, HeuWork& W,
                 std::vector<uint32_t>& H, uint32_t& kl) {
  H.clear();
  kl = 0;
  if ((uint32_t)(g.off[q + 1] - g.off[q]) < N1 - 1u) return;
  std::vector<uint32_t> S;
  S.push_back(q);
  touch(W, q);
  W.inS[q] = 1;
  ++W.deg[q];
  for (int64_t j = g.off[q]; j < g.off[q + 1]; j++) {
    uint32_t x = (uint32_t)g.adj[j];
    S.push_back(x);
    touch(W, x);
    W.inS[x] = 1;
    for (int64_t k = g.off[x]; k < g.off[x + 1]; k++) {
      uint32_t y = (uint32_t)g.adj[k];
      if (W.inS[y]) {
        ++W.deg[y];
        ++W.deg[x];
      }
    }
  }
  std::priority_queue<std::pair<uint32_t, uint32_t>, std::vector<std::pair<uint32_t, uint32_t>>,
                      std::greater<std::pair<uint32_t, uint32_t>>>
      Q;
  for (uint32_t x : S) Q.push({W.deg[x], x});
  std::vector<uint32_t> rmv;
  uint32_t rmv_idx = 0, mindeg = 0;
  while (!Q.empty()) {
    uint32_t d = Q.top().first, v = Q.top().second;
    Q.pop();
    if (!W.inS[v]) continue;
    uint32_t remain = (uint32_t)S.size() - (uint32_t)rmv.size();
    // q's +1 (inc/heuristic.h:259) only keeps it from being peeled early;
    // it is popped only when alone, where its true degree is 0
    if (v == q) d -= 1;
    if (remain >= N1 && remain <= N2 && d > mindeg) {
      mindeg = d;
      rmv_idx = (uint32_t)rmv.size();
    }
    if (remain == N1) break;
    W.inS[v] = 0;
    rmv.push_back(v);
    for (int64_t j = g.off[v]; j < g.off[v + 1]; j++) {
      uint32_t x = (uint32_t)g.adj[j];
      if (W.inS[x]) {
        --W.deg[x];
        Q.push({W.deg[x], x});
      }
    }
  }
  for (uint32_t x : S) W.inS[x] = 1;
  for (uint32_t i = 0; i < rmv_idx; i++) W.inS[rmv[i]] = 0;
  for (uint32_t x : S)
    if (W.inS[x]) H.push_back(x);
  kl = mindeg;
  W.clear_touched(HeuWork::CLEAR_DEG_INS);
}

uint32_t cssc_heu(const HostGraph& g, uint32_t l, uint32_t h, uint32_t q, HeuWork& W,
                  std::vector<uint32_t>& H) {
  W.reset(g.n);
  std::vector<uint32_t> H2, H3, H4;
  uint32_t k2, k3, k4;
  heu2(g, l, h, q, W, H2, k2);
  heu3(g, l, h, q, W, H3, k3);
  heu4(g, l, h, q, W, H4, k4);
  if (k4 >= k3 && k4 >= k2) {
    H.swap(H4);
    return k4;
  }
  if (k3 >= k2) {
    H.swap(H3);
    return k3;
  }
  H.swap(H2);
  return k2;
}

}  // namespace scs
