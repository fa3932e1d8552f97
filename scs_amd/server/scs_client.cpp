// Batch client: sends each line of a query file to the server's queue
// (reference client/batchClient.cpp:17-53), retrying while the queue is
// full, then "server_exit" if the file contains it.
//   scs_client <query file> [queue key char]
#include <unistd.h>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "scs_amd/server/scs_ipc.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: scs_client <query file> [queue key]\n");
    return 1;
  }
  std::ifstream in(argv[1]);
  if (!in) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  const char key = argc > 2 ? argv[2][0] : 'g';
  std::vector<std::string> lines;
  bool server_exit = false;
  std::string s;
  while (std::getline(in, s)) {
    if (s == "server_exit") {
      server_exit = true;
      break;
    }
    if (!s.empty()) lines.push_back(s);
  }
  int qid = -1;
  for (int tries = 0; tries < 2000 && qid < 0; tries++) {
    qid = scs_ipc::open_queue(key, false);
    if (qid < 0) usleep(5000);
  }
  if (qid < 0) {
    perror("msgget (is the server running?)");
    return 1;
  }
  for (const std::string& l : lines)
    while (!scs_ipc::try_send(qid, l)) usleep(1);
  if (server_exit)
    while (!scs_ipc::try_send(qid, "server_exit")) usleep(1);
  return 0;
}
