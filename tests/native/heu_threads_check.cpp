// The admission path's host threads: several threads take queries from one
// atomic counter and run the greedy heuristic (cssc_heu) on the shared,
// read-only graph, each with its own scratch, writing their answers into
// per-query slots, as scs_step does.  Built with -fsanitize=thread (or
// address,undefined) by tests/test_host_sanitizers_cpu.py; the answers must
// equal a single-threaded pass, which in turn must equal runs with a fresh
// scratch per query.
//
// Input (argv[1]): a binary CSR, n (u32), nnz (u64), off[n+1] (i64),
// adj[nnz] (i32).  argv[2]: number of queries, argv[3]: threads.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "scs_host.h"

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  scs::HostGraph g;
  uint64_t nnz = 0;
  if (fread(&g.n, 4, 1, f) != 1 || fread(&nnz, 8, 1, f) != 1) return 2;
  g.off.resize(g.n + 1);
  g.adj.resize(nnz);
  if (fread(g.off.data(), 8, g.n + 1, f) != g.n + 1) return 2;
  if (nnz && fread(g.adj.data(), 4, nnz, f) != nnz) return 2;
  fclose(f);
  for (uint32_t v = 0; v < g.n; v++)
    g.dmax = std::max<uint32_t>(g.dmax, (uint32_t)(g.off[v + 1] - g.off[v]));
  const uint32_t nq = (uint32_t)atoi(argv[2]), nt = (uint32_t)atoi(argv[3]);
  struct Q {
    uint32_t l, h, q;
  };
  std::vector<Q> qs(nq);
  std::mt19937 rng(11);
  for (auto& x : qs) {
    x.l = 3 + rng() % 12;
    x.h = x.l + 2 + rng() % 20;
    x.q = rng() % g.n;
  }
  // single-threaded answers, one scratch reused across the queries
  std::vector<uint32_t> kl1(nq);
  std::vector<std::vector<uint32_t>> H1(nq);
  {
    scs::HeuWork W;
    for (uint32_t i = 0; i < nq; i++) kl1[i] = scs::cssc_heu(g, qs[i].l, qs[i].h, qs[i].q, W, H1[i]);
  }
  // a fresh scratch per query gives the same answers: clearing the touched
  // entries leaves nothing of the previous query behind
  for (uint32_t i = 0; i < nq; i += 7) {
    scs::HeuWork W;
    std::vector<uint32_t> H0;
    if (scs::cssc_heu(g, qs[i].l, qs[i].h, qs[i].q, W, H0) != kl1[i] || H0 != H1[i]) {
      fprintf(stderr, "query %u differs with a fresh scratch\n", i);
      return 1;
    }
  }
  // threads claiming queries from one counter, as the admission step does
  std::vector<uint32_t> kl(nq);
  std::vector<std::vector<uint32_t>> H(nq);
  std::vector<scs::HeuWork> work(nt);
  std::atomic<uint32_t> next{0};
  auto worker = [&](uint32_t t) {
    uint32_t i;
    while ((i = next.fetch_add(1)) < nq) kl[i] = scs::cssc_heu(g, qs[i].l, qs[i].h, qs[i].q, work[t], H[i]);
  };
  std::vector<std::thread> th;
  for (uint32_t t = 1; t < nt; t++) th.emplace_back(worker, t);
  worker(0);
  for (auto& t : th) t.join();
  for (uint32_t i = 0; i < nq; i++)
    if (kl[i] != kl1[i] || H[i] != H1[i]) {
      fprintf(stderr, "query %u differs: %u vs %u\n", i, kl[i], kl1[i]);
      return 1;
    }
  printf("threads ok %u queries %u threads\n", nq, nt);
  return 0;
}
