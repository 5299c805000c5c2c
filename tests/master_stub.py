"""A stand-in for the Satsuma2 master's side of the slave exchange (test infrastructure).

One exchange per connection, host byte order, no framing (analysis/HomologyByXCorrSlave.cc:
453-517 for the slave, analysis/WorkQueue.cc:71-160 for the master):
    slave -> master   int32 slave_id
    slave -> master   uint32 n, then n t_result records (72 bytes each)
    master -> slave   int32 m (-1 = stop), then m t_pair records (28 bytes each)
The stub hands out its blocks at most `per_exchange` at a time (the master sends at most
2 * threads), collects every record, and answers -1 once `expect` records have arrived
(or after `deadline` seconds).
"""
from __future__ import annotations

import socket
import struct
import threading
import time

import numpy as np

RESULT_DTYPE = np.dtype([("query_id", "<u8"), ("target_id", "<u8"), ("query_size", "<u8"), ("qstart", "<u8"),
                         ("tstart", "<u8"), ("len", "<u8"), ("reverse", "u1"), ("pad_", "u1", (7,)),
                         ("prob", "<f8"), ("ident", "<f8")])
assert RESULT_DTYPE.itemsize == 72


def pack_pair(target_from, target_to, query_from, query_to, fast, slave_id=0, status=1) -> bytes:
    """t_pair (analysis/WorkQueue.h:17-22): four int32, bool (+3 pad), int32, uchar (+3 pad)."""
    return struct.pack("<4iB3xiB3x", target_from, target_to, query_from, query_to, 1 if fast else 0, slave_id,
                       status)


def _read_exact(conn, n):
    buf = bytearray()
    while len(buf) < n:
        chunk = conn.recv(n - len(buf))
        if not chunk:
            raise ConnectionError(f"slave closed after {len(buf)} of {n} bytes")
        buf.extend(chunk)
    return bytes(buf)


class MasterStub:
    def __init__(self, blocks, expect, per_exchange=2, deadline=120.0):
        self.blocks = list(blocks)
        self.expect = expect
        self.per_exchange = per_exchange
        self.deadline = deadline
        self.records = []
        self.slave_ids = []
        self.exchanges = 0
        self.error = None
        self.sock = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.sock.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.sock.bind(("127.0.0.1", 0))
        self.sock.listen(8)
        self.sock.settimeout(1.0)
        self.port = self.sock.getsockname()[1]
        self.thread = threading.Thread(target=self._serve, daemon=True)
        self.done = threading.Event()

    def start(self):
        self.thread.start()
        return self

    def results(self) -> np.ndarray:
        return np.frombuffer(b"".join(self.records), RESULT_DTYPE) if self.records else np.zeros(0, RESULT_DTYPE)

    def _serve(self):
        t0 = time.time()
        try:
            while not self.done.is_set():
                try:
                    conn, _ = self.sock.accept()
                except socket.timeout:
                    if time.time() - t0 > self.deadline:
                        break
                    continue
                with conn:
                    conn.settimeout(30.0)
                    (sid,) = struct.unpack("<i", _read_exact(conn, 4))
                    self.slave_ids.append(sid)
                    (n,) = struct.unpack("<I", _read_exact(conn, 4))
                    if n:
                        self.records.append(_read_exact(conn, 72 * n))
                    self.exchanges += 1
                    got = sum(len(r) for r in self.records) // 72
                    if not self.blocks and (got >= self.expect or time.time() - t0 > self.deadline):
                        conn.sendall(struct.pack("<i", -1))
                        self.done.set()
                        break
                    give, self.blocks = self.blocks[:self.per_exchange], self.blocks[self.per_exchange:]
                    conn.sendall(struct.pack("<i", len(give)) + b"".join(pack_pair(*b) for b in give))
        except Exception as e:  # reported by the test
            self.error = e
        finally:
            self.sock.close()
            self.done.set()
