"""Independent Python reader/writer for the canonical LigeSIS proof encoding.

The reference has no proof serializer (pcs/src/ligesis.rs:124-138 only sizes it), so the
encoding is fixed by this build: the fields of LigesisProof (ligesis.rs:99-119) and BatchProof
(basefold.rs:151-162) in declaration order; u64 little-endian integers and lengths; a base
element is a canonical u64, an extension element (c0, c1); [u8; 32] raw; Vec<T> = u64 length +
elements; BTreeMap<usize, Vec<T>> = u64 count + (u64 key, Vec<T>) in ascending key order.
"""
from __future__ import annotations

import struct


class _R:
    def __init__(self, b: bytes):
        self.b, self.c = b, 0

    def u64(self):
        v = struct.unpack_from("<Q", self.b, self.c)[0]
        self.c += 8
        return v

    def ef(self):
        return (self.u64(), self.u64())

    def h32(self):
        v = self.b[self.c:self.c + 32]
        self.c += 32
        return v

    def bytes_(self):
        n = self.u64()
        v = self.b[self.c:self.c + n]
        self.c += n
        return v

    def vec(self, f):
        return [f() for _ in range(self.u64())]


def _sc(r):
    return {"round_polys": r.vec(lambda: r.vec(r.ef)), "final_evals": r.vec(r.ef)}


SUMCHECKS = ["sc_binary", "sc_hash", "sc_recon", "sc_table", "sc_rslink", "sc_bi_batch", "sc_a_batch"]


def decode(b: bytes) -> dict:
    r = _R(b)
    d = {"y": r.ef(), "com_a_vec": r.h32(), "com_i": r.h32(), "y4": r.ef(), "y5": r.ef()}
    d["h_cols"] = r.vec(lambda: r.vec(r.u64))
    d["h_open"] = r.bytes_()
    for name in SUMCHECKS:
        d[name] = _sc(r)

    def opening(val):
        return {"proof_bytes": r.bytes_(), "values": r.vec(lambda: (r.u64(), r.vec(val)))}

    bp = {"reduction_msgs": r.vec(lambda: r.vec(r.ef)), "fold_roots": r.vec(r.h32),
          "final_codeword": r.vec(r.ef)}
    bp["fold_openings"] = r.vec(lambda: opening(r.ef))
    bp["base_bf"] = r.vec(lambda: opening(r.u64))
    bp["base_ef"] = r.vec(lambda: opening(r.ef))
    bp["code_rate"] = r.u64()
    bp["log_len"] = r.u64()
    d["batch"] = bp
    assert r.c == len(b), "trailing bytes"
    return d


def _u(v):
    return struct.pack("<Q", v)


def _efb(e):
    return _u(e[0]) + _u(e[1])


def _vec(xs, f):
    return _u(len(xs)) + b"".join(f(x) for x in xs)


def encode(d: dict) -> bytes:
    out = _efb(d["y"]) + d["com_a_vec"] + d["com_i"] + _efb(d["y4"]) + _efb(d["y5"])
    out += _vec(d["h_cols"], lambda c: _vec(c, _u))
    out += _u(len(d["h_open"])) + d["h_open"]
    for name in SUMCHECKS:
        s = d[name]
        out += _vec(s["round_polys"], lambda rp: _vec(rp, _efb)) + _vec(s["final_evals"], _efb)
    bp = d["batch"]

    def opening(o, val):
        return (_u(len(o["proof_bytes"])) + o["proof_bytes"] +
                _vec(o["values"], lambda kv: _u(kv[0]) + _vec(kv[1], val)))

    out += _vec(bp["reduction_msgs"], lambda c: _vec(c, _efb))
    out += _vec(bp["fold_roots"], lambda h: h)
    out += _vec(bp["final_codeword"], _efb)
    out += _vec(bp["fold_openings"], lambda o: opening(o, _efb))
    out += _vec(bp["base_bf"], lambda o: opening(o, _u))
    out += _vec(bp["base_ef"], lambda o: opening(o, _efb))
    out += _u(bp["code_rate"]) + _u(bp["log_len"])
    return out


def size_bytes(d: dict) -> int:
    """LigesisProof::size_bytes (ligesis.rs:124-138) of the decoded proof."""
    ef, bf = 16, 8
    s = ef * 3 + 64
    s += len(d["h_open"]) + sum(len(c) * bf for c in d["h_cols"])
    for name in SUMCHECKS:
        sc = d[name]
        s += sum(len(r) * ef for r in sc["round_polys"]) + len(sc["final_evals"]) * ef
    bp = d["batch"]
    s += sum(len(c) * ef for c in bp["reduction_msgs"]) + len(bp["fold_roots"]) * 32
    s += len(bp["final_codeword"]) * ef
    for o in bp["fold_openings"]:
        s += len(o["proof_bytes"]) + sum(len(v) * ef for _, v in o["values"])
    for o in bp["base_bf"]:
        s += len(o["proof_bytes"]) + sum(len(v) * bf for _, v in o["values"])
    for o in bp["base_ef"]:
        s += len(o["proof_bytes"]) + sum(len(v) * ef for _, v in o["values"])
    return s
