// SysV message-queue transport of the query protocol.  Same wire format as
// the reference (ipc/msgtool.h:16-24): key ftok("/tmp", c), message type 1,
// a NUL-terminated text of at most 80 bytes per message.
#pragma once
#include <errno.h>
#include <string.h>
#include <sys/ipc.h>
#include <sys/msg.h>
#include <sys/types.h>

#include <string>

namespace scs_ipc {

constexpr int kMaxText = 80;

struct Msg {
  long mtype;
  char mtext[kMaxText];
};

// Opens (creating if needed) the queue for service character c; -1 on error.
inline int open_queue(char c, bool create) {
  key_t key = ftok("/tmp", c);
  if (key == (key_t)-1) return -1;
  return msgget(key, create ? (IPC_CREAT | 0660) : 0);
}

// Non-blocking receive of one type-1 message.  1 = got one, 0 = none,
// -1 = error.
inline int try_recv(int qid, std::string& out) {
  Msg m;
  ssize_t r = msgrcv(qid, &m, kMaxText, 1, IPC_NOWAIT);
  if (r < 0) return errno == ENOMSG ? 0 : -1;
  m.mtext[kMaxText - 1] = 0;
  out.assign(m.mtext, strnlen(m.mtext, (size_t)r < (size_t)kMaxText ? (size_t)r : kMaxText));
  return 1;
}

// Non-blocking send; false when the queue is full or on error.
inline bool try_send(int qid, const std::string& text) {
  Msg m;
  m.mtype = 1;
  size_t n = text.size() < (size_t)kMaxText - 1 ? text.size() : (size_t)kMaxText - 1;
  memcpy(m.mtext, text.data(), n);
  m.mtext[n] = 0;
  return msgsnd(qid, &m, n + 1, IPC_NOWAIT) == 0;
}

inline void remove_queue(int qid) { msgctl(qid, IPC_RMID, nullptr); }

}  // namespace scs_ipc
