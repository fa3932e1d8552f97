"""The batch client speaks the reference's wire format (SysV queue key
ftok("/tmp", c), type 1, NUL-terminated text <= 80 bytes,
ipc/msgtool.h:16-24): read back with plain libc calls, no GPU needed."""

import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "scs_amd", "server", "scs_client")


class Msg(C.Structure):
    _fields_ = [("mtype", C.c_long), ("mtext", C.c_char * 80)]


def test_client_wire_format(tmp_path):
    if not os.path.exists(CLIENT):
        pytest.fail("scs_client not built: run python scs_amd/build.py")
    libc = C.CDLL(None, use_errno=True)
    libc.ftok.restype = C.c_int
    libc.ftok.argtypes = [C.c_char_p, C.c_int]
    key = libc.ftok(b"/tmp", ord("y"))
    qid = libc.msgget(key, 0o1000 | 0o660)  # IPC_CREAT
    assert qid >= 0
    try:
        qf = tmp_path / "q.txt"
        qf.write_text("3 6 10 1 1\n6 9 54 1 1\n\nserver_exit\n9 12 10 1 1\n")
        subprocess.run([CLIENT, str(qf), "y"], check=True, timeout=30)
        got = []
        m = Msg()
        while libc.msgrcv(qid, C.byref(m), 80, 1, 0o4000) >= 0:  # IPC_NOWAIT
            got.append(m.mtext.decode())
        assert got == ["3 6 10 1 1", "6 9 54 1 1", "server_exit"]
    finally:
        libc.msgctl(qid, 0, None)  # IPC_RMID
