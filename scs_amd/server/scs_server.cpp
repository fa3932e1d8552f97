// SCS query server: the reference's process surface (main.cu:475-597) over
// the MI355X engine.
//
//   scs_server <graph> <partitionSize> <bufferSize> <copyLimit> <readLimit>
//              <limitQueries> <factor> <red1> <red2> <red3> <prun1> <prun2>
//              [--wave-per-cu N] [--limit-tasks N] [--node-budget N]
//              [--mem-budget BYTES] [--learned MODEL.bin] [--time-slice-us N]
//              [--results DIR] [--queue-key C] [--levels 0|1] [--pair 0|1]
//              [--whole-graph 0|1|2]
//
// The twelve positional arguments are those of the reference (main.cu:
// 481-502).  A listener thread receives "<l> <h> <q> <heu> <limitDoms>"
// messages (or "server_exit") from the SysV queue and queues them in the
// engine; the main thread runs continuous-batching steps and appends one
// line per finished query to results/results_<graph basename>:
//   N1|N2|QID|Time|Degree|Overtime|Heu|0
// (Time in microseconds from arrival, comma-grouped; Overtime 1 = node
// budget reached, 2 = query rejected by the kept-set limit), and at exit
//   -1|-1|-1|-1|-1|-1|-1|<total microseconds>.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "include/scs.h"
#include "scs_amd/server/scs_ipc.h"

static std::string commas(long long x) {
  std::string s = std::to_string(x < 0 ? 0 : x);
  std::string out;
  int c = 0;
  for (int i = (int)s.size() - 1; i >= 0; i--) {
    out.insert(out.begin(), s[i]);
    if (++c % 3 == 0 && i > 0) out.insert(out.begin(), ',');
  }
  return out;
}

static bool is_exit(std::string s) {
  auto issp = [](char ch) { return ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t'; };
  while (!s.empty() && issp(s.back())) s.pop_back();
  size_t i = 0;
  while (i < s.size() && issp(s[i])) i++;
  s = s.substr(i);
  for (char& ch : s) ch = (char)tolower((unsigned char)ch);
  return s == "server_exit";
}

static void usage() {
  fprintf(stderr,
          "Server wrong input parameters!\n"
          "usage: scs_server <graph> <partitionSize> <bufferSize> <copyLimit> <readLimit> "
          "<limitQueries> <factor> <red1> <red2> <red3> <prun1> <prun2> [options]\n");
}

int main(int argc, char** argv) {
  if (argc < 13) {
    usage();
    return 1;
  }
  scs_params p;
  scs_params_default(&p);
  const char* graph_path = argv[1];
  p.partition_size = (uint32_t)strtoul(argv[2], nullptr, 10);
  p.buffer_size = (uint32_t)strtoul(argv[3], nullptr, 10);
  p.copy_limit = strtod(argv[4], nullptr);
  p.read_limit = (uint32_t)strtoul(argv[5], nullptr, 10);
  p.limit_queries = (uint32_t)strtoul(argv[6], nullptr, 10);
  p.factor = (uint32_t)strtoul(argv[7], nullptr, 10);
  p.red1 = (uint32_t)atoi(argv[8]);
  p.red2 = (uint32_t)atoi(argv[9]);
  p.red3 = (uint32_t)atoi(argv[10]);
  p.prun1 = (uint32_t)atoi(argv[11]);
  p.prun2 = (uint32_t)atoi(argv[12]);
  std::string results_dir = "results";
  std::string model_path;
  char key = 'g';
  for (int i = 13; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&]() -> const char* {
      if (i + 1 >= argc) {
        usage();
        exit(1);
      }
      return argv[++i];
    };
    if (a == "--wave-per-cu") p.waves_per_cu = (uint32_t)atoi(val());
    else if (a == "--limit-tasks") p.limit_tasks = (uint32_t)atoi(val());
    else if (a == "--node-budget") p.node_budget = strtoull(val(), nullptr, 10);
    else if (a == "--mem-budget") p.mem_budget = strtoull(val(), nullptr, 10);
    else if (a == "--time-slice-us") p.time_slice_us = (uint32_t)atoi(val());
    else if (a == "--levels") p.levels = (uint32_t)atoi(val());
    else if (a == "--pair") p.pair = (uint32_t)atoi(val());
    else if (a == "--whole-graph") p.whole_graph = (uint32_t)atoi(val());
    else if (a == "--learned") model_path = val();
    else if (a == "--results") results_dir = val();
    else if (a == "--queue-key") key = val()[0];
    else {
      usage();
      return 1;
    }
  }
  scs_model model;
  if (!model_path.empty()) {
    FILE* f = fopen(model_path.c_str(), "rb");
    if (!f || fread(&model, sizeof(model), 1, f) != 1) {
      fprintf(stderr, "cannot read model %s\n", model_path.c_str());
      return 1;
    }
    fclose(f);
    p.learned = 1;
    p.model = &model;
  }
  scs_graph* g = nullptr;
  int rc = scs_graph_load(graph_path, &g);
  if (rc) {
    fprintf(stderr, "Graph file Open Failed: %s (%s)\n", graph_path, scs_strerror(rc));
    return 1;
  }
  uint32_t n;
  uint64_t nnz;
  uint32_t dmax;
  scs_graph_info(g, &n, &nnz, &dmax);
  printf("n=%u,m=%llu,dMAX=%u\n", n, (unsigned long long)(nnz / 2), dmax);
  scs_engine* e = nullptr;
  rc = scs_engine_create(g, &p, &e);
  if (rc) {
    fprintf(stderr, "engine: %s\n", scs_strerror(rc));
    return 1;
  }
  std::string base = graph_path;
  size_t pos = base.find_last_of("/\\");
  if (pos != std::string::npos) base = base.substr(pos + 1);
  mkdir(results_dir.c_str(), 0755);
  const std::string fname = results_dir + "/results_" + base;
  {
    struct stat sb;
    if (stat(fname.c_str(), &sb) != 0) {
      std::ofstream f(fname, std::ios::app);
      f << "N1|N2|QID|Time|Degree|Overtime|Heu|TotalTime\n";
    }
  }
  const int qid = scs_ipc::open_queue(key, true);
  if (qid < 0) {
    perror("msgget");
    return 1;
  }
  std::atomic<bool> quit{false};
  std::atomic<uint64_t> submitted{0};
  std::thread listener([&]() {
    std::string msg;
    uint64_t arrival = 0;
    while (!quit.load()) {
      int r = scs_ipc::try_recv(qid, msg);
      if (r < 0) {
        perror("msgrcv");
        quit = true;
        break;
      }
      if (r == 0) {
        usleep(50);
        continue;
      }
      if (is_exit(msg)) {
        quit = true;
        break;
      }
      if (scs_submit_text(e, msg.c_str(), arrival) == SCS_OK) {
        submitted++;
      } else {
        printf("Client wrong input parameters! %s\n", msg.c_str());
        fflush(stdout);
      }
      arrival++;
    }
  });
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<scs_result> res(64);
  uint64_t done = 0;
  while (true) {
    int left = scs_step(e);
    if (left < 0) {
      fprintf(stderr, "scs_step: %s\n", scs_strerror(left));
      break;
    }
    int k;
    while ((k = scs_poll(e, res.data(), (int)res.size())) > 0) {
      std::ofstream f(fname, std::ios::app);
      for (int i = 0; i < k; i++) {
        const scs_result& r = res[i];
        int over = r.status == SCS_ST_CAPPED ? 1 : r.status == SCS_ST_REJECTED ? 2 : 0;
        int heu = r.status == SCS_ST_HEU_OPT ? 1 : 0;
        f << r.l << "|" << r.h << "|" << r.q << "|"
          << commas((long long)(r.t_done_us - r.t_submit_us)) << "|" << r.kl << "|" << over
          << "|" << heu << "|0\n";
        printf("Found Solution : N1 = %u, N2 = %u, QID = %u, kl = %u, ku = %u, status = %d\n", r.l,
               r.h, r.q, r.kl, r.ku, r.status);
        done++;
      }
      fflush(stdout);
    }
    if (quit.load() && left == 0 && done >= submitted.load()) break;
    if (left == 0) usleep(100);
  }
  quit = true;
  listener.join();
  const double total =
      std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  {
    std::ofstream f(fname, std::ios::app);
    f << "-1|-1|-1|-1|-1|-1|-1|" << commas((long long)total) << "\n";
  }
  scs_ipc::remove_queue(qid);
  scs_engine_destroy(e);
  scs_graph_destroy(g);
  printf("End\n");
  return 0;
}
