"""Oracle: a CPU restatement of the reference LigeSIS path, used only as a checker.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package. The product library (``ligesis_amd``) never imports it and never falls back to it.

``oracle.lib()`` loads ``oracle/liblgs_oracle.so`` (built by ``make -C oracle`` or
``__graft_entry__.build()``); the thin wrappers below convert numpy arrays to the C entry
points declared in ``oracle/capi.cpp``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liblgs_oracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> str:
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        i64 = ctypes.c_int64
        vp = ctypes.c_void_p
        u64 = ctypes.c_uint64
        L.lgso_last_error.restype = ctypes.c_char_p
        L.lgso_two_adic_generator.restype = u64
        L.lgso_two_adic_generator.argtypes = [u64]
        for name in [
            "lgso_merkle_root", "lgso_merkle_open", "lgso_merkle_verify", "lgso_sumcheck_prove",
            "lgso_sumcheck_verify", "lgso_basefold_batch_roundtrip", "lgso_shape_ok",
            "lgso_pk_com_a", "lgso_cd_enc", "lgso_cd_h", "lgso_prove", "lgso_verify",
            "lgso_proof_reencode", "lgso_basefold_commit", "lgso_basefold_prove", "lgso_basefold_verify",
        ]:
            getattr(L, name).restype = i64
        L.lgso_setup.restype = vp
        L.lgso_setup.argtypes = [u64, u64, vp, ctypes.c_int]
        L.lgso_commit.restype = vp
        L.lgso_commit.argtypes = [vp, vp, vp, ctypes.c_int]
        L.lgso_free_pk.argtypes = [vp]
        L.lgso_free_cd.argtypes = [vp]
        L.lgso_pk_com_a.argtypes = [vp, vp]
        L.lgso_cd_enc.argtypes = [vp, vp]
        L.lgso_cd_h.argtypes = [vp, vp]
        L.lgso_prove.argtypes = [vp, vp, vp, vp, u64, vp, u64, vp, u64, vp, vp]
        L.lgso_verify.argtypes = [u64, u64, vp, vp, vp, vp, u64, vp, u64, vp, u64]
        L.lgso_proof_reencode.argtypes = [vp, u64, vp]
        L.lgso_shape_ok.argtypes = [u64, u64]
        L.lgso_basefold_commit.argtypes = [vp, u64, i64, u64, u64, vp]
        L.lgso_basefold_prove.argtypes = [vp, u64, i64, u64, u64, vp, vp, u64, vp, u64, vp, vp, u64, vp]
        L.lgso_basefold_verify.argtypes = [vp, vp, u64, vp, u64, i64, vp, u64, vp, u64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def err() -> str:
    return lib().lgso_last_error().decode()


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# ---- field ---------------------------------------------------------------------------------
def field_ops(a, b):
    a, b = _u64(a), _u64(b)
    n = a.size
    add, sub, mul = (np.zeros(n, np.uint64) for _ in range(3))
    lib().lgso_field_ops(_p(a), _p(b), ctypes.c_uint64(n), _p(add), _p(sub), _p(mul))
    return add, sub, mul


def reduce128(lo, hi):
    lo, hi = _u64(lo), _u64(hi)
    out = np.zeros(lo.size, np.uint64)
    lib().lgso_reduce128(_p(lo), _p(hi), ctypes.c_uint64(lo.size), _p(out))
    return out


def ef_ops(a, b):
    """a, b: (n, 2) uint64 arrays. Returns (a*b, a^-1)."""
    a, b = _u64(a), _u64(b)
    n = a.shape[0]
    mul, inv = np.zeros((n, 2), np.uint64), np.zeros((n, 2), np.uint64)
    lib().lgso_ef_ops(_p(a), _p(b), ctypes.c_uint64(n), _p(mul), _p(inv))
    return mul, inv


def two_adic_generator(bits: int) -> int:
    return int(lib().lgso_two_adic_generator(bits))


# ---- dft -----------------------------------------------------------------------------------
def dft(a):
    a = _u64(a).copy()
    lib().lgso_dft(_p(a), ctypes.c_uint64(a.size))
    return a


def dft_ext(a):
    a = _u64(a).copy()
    lib().lgso_dft_ext(_p(a), ctypes.c_uint64(a.shape[0]))
    return a


def rs_encode_rows(f, rows: int, n: int):
    f = _u64(f)
    out = np.zeros(rows * 2 * n, np.uint64)
    lib().lgso_rs_encode_rows(_p(f), ctypes.c_uint64(rows), ctypes.c_uint64(n), _p(out))
    return out.reshape(rows, 2 * n)


# ---- blake3 / merkle -----------------------------------------------------------------------
def blake3(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().lgso_blake3(_p(buf), ctypes.c_uint64(len(data)), _p(out))
    return out.tobytes()


def merkle_root(leaves: np.ndarray) -> bytes:
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8)
    out = np.zeros(32, np.uint8)
    rc = lib().lgso_merkle_root(_p(leaves), ctypes.c_uint64(leaves.shape[0]),
                                ctypes.c_uint64(leaves.shape[1]), _p(out))
    assert rc == 0, err()
    return out.tobytes()


def merkle_open(leaves: np.ndarray, idx) -> bytes:
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8)
    idx = _u64(idx)
    n, ll = leaves.shape
    need = lib().lgso_merkle_open(_p(leaves), ctypes.c_uint64(n), ctypes.c_uint64(ll), _p(idx),
                                  ctypes.c_uint64(idx.size), None, ctypes.c_uint64(0))
    assert need >= 0, err()
    out = np.zeros(max(need, 1), np.uint8)
    lib().lgso_merkle_open(_p(leaves), ctypes.c_uint64(n), ctypes.c_uint64(ll), _p(idx),
                           ctypes.c_uint64(idx.size), _p(out), ctypes.c_uint64(need))
    return out[:need].tobytes()


def merkle_verify(root: bytes, n_leaves: int, proof: bytes, idx, leaves: np.ndarray) -> bool:
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8)
    idx = _u64(idx)
    r = np.frombuffer(root, np.uint8).copy()
    pf = np.frombuffer(proof, np.uint8).copy() if proof else np.zeros(1, np.uint8)
    rc = lib().lgso_merkle_verify(_p(r), ctypes.c_uint64(n_leaves), _p(pf), ctypes.c_uint64(len(proof)),
                                  _p(idx), ctypes.c_uint64(idx.size), _p(leaves),
                                  ctypes.c_uint64(leaves.shape[1]))
    assert rc >= 0, err()
    return bool(rc)


# ---- multilinear / sumcheck ----------------------------------------------------------------
def mle_eval(v, point):
    v, point = _u64(v), _u64(point)
    out = np.zeros(2, np.uint64)
    lib().lgso_mle_eval(_p(v), ctypes.c_uint64(point.shape[0]), _p(point), _p(out))
    return out


def new_eq(point):
    point = _u64(point)
    nv = point.shape[0]
    out = np.zeros((1 << nv, 2), np.uint64)
    lib().lgso_new_eq(_p(point), ctypes.c_uint64(nv), _p(out))
    return out


def sumcheck_prove(polys, fields):
    """polys: (d, 2^nv, 2) EF. Returns (rounds (nv, d+1, 2), finals (d, 2), fields_used)."""
    polys, fields = _u64(polys), _u64(fields)
    d, L = polys.shape[0], polys.shape[1]
    nv = L.bit_length() - 1
    rounds = np.zeros((nv, d + 1, 2), np.uint64)
    finals = np.zeros((d, 2), np.uint64)
    used = lib().lgso_sumcheck_prove(ctypes.c_uint64(d), ctypes.c_uint64(nv), _p(polys), _p(fields),
                                     ctypes.c_uint64(fields.shape[0]), _p(rounds), _p(finals))
    assert used >= 0, err()
    return rounds, finals, used


def sumcheck_verify(claim, rounds, finals, fields) -> bool:
    claim, rounds, finals, fields = _u64(claim), _u64(rounds), _u64(finals), _u64(fields)
    nv, d1 = rounds.shape[0], rounds.shape[1]
    rc = lib().lgso_sumcheck_verify(ctypes.c_uint64(d1 - 1), ctypes.c_uint64(nv), _p(claim),
                                    _p(rounds), _p(finals), _p(fields), ctypes.c_uint64(fields.shape[0]))
    assert rc >= 0, err()
    return bool(rc)


def basefold_commit(poly, li: int, code_rate: int) -> bytes:
    """Root of Basefold::commit_{base,ext}_on_domain (basefold.rs:341-382): ``poly`` is 2^nv base
    values or (2^nv, 2) extension pairs."""
    a = _u64(poly)
    ext = a.ndim == 2
    nv = a.shape[0].bit_length() - 1
    root = np.zeros(32, np.uint8)
    flat = np.ascontiguousarray(a).reshape(-1)
    rc = lib().lgso_basefold_commit(_p(flat), nv, int(ext), li, code_rate, _p(root))
    assert rc >= 0, err()
    return root.tobytes()


def basefold_prove(poly, li: int, code_rate: int, point, fields, ints):
    """Basefold::commit_{base,ext}_on_domain then Basefold::prove (basefold.rs:341-516) at
    ``point`` (nv extension elements). Returns (root, proof bytes, (fields used, ints used))."""
    a = _u64(poly)
    ext = a.ndim == 2
    nv = a.shape[0].bit_length() - 1
    flat = np.ascontiguousarray(a).reshape(-1)
    point, fields, ints = _u64(point).reshape(-1, 2), _u64(fields).reshape(-1, 2), _u64(ints)
    args = (_p(flat), nv, int(ext), li, code_rate, _p(point), _p(fields), fields.shape[0], _p(ints), ints.size)
    root = np.zeros(32, np.uint8)
    need = lib().lgso_basefold_prove(*args, _p(root), None, 0, None)
    if need < 0:
        raise RuntimeError(err())
    out = np.zeros(need, np.uint8)
    used = np.zeros(2, np.uint64)
    got = lib().lgso_basefold_prove(*args, _p(root), _p(out), need, _p(used))
    assert got == need, err()
    return root.tobytes(), out.tobytes(), (int(used[0]), int(used[1]))


def basefold_verify(root: bytes, point, proof: bytes, ext: bool, fields, ints) -> bool:
    """Basefold::verify (basefold.rs:518-642) of a proof in the canonical encoding."""
    point, fields, ints = _u64(point).reshape(-1, 2), _u64(fields).reshape(-1, 2), _u64(ints)
    rt = np.frombuffer(root, np.uint8).copy()
    pf = np.frombuffer(proof, np.uint8).copy() if proof else np.zeros(1, np.uint8)
    rc = lib().lgso_basefold_verify(_p(rt), _p(point), point.shape[0], _p(pf), len(proof), int(ext),
                                    _p(fields), fields.shape[0], _p(ints), ints.size)
    if rc < 0:
        raise RuntimeError(err())
    return bool(rc)


def basefold_batch_roundtrip(bf_polys, ef_polys, points, rounds, code_rate, fields, ints, tamper=0):
    bf_nv = _u64([p.size.bit_length() - 1 for p in bf_polys])
    ef_nv = _u64([p.shape[0].bit_length() - 1 for p in ef_polys])
    bfp = _u64(np.concatenate([_u64(p) for p in bf_polys])) if bf_polys else np.zeros(1, np.uint64)
    efp = _u64(np.concatenate([_u64(p).reshape(-1) for p in ef_polys])) if ef_polys else np.zeros(1, np.uint64)
    pts = _u64(np.concatenate([_u64(p).reshape(-1) for p in points]))
    fields, ints = _u64(fields), _u64(ints)
    rc = lib().lgso_basefold_batch_roundtrip(
        _p(bf_nv), ctypes.c_uint64(len(bf_polys)), _p(bfp), _p(ef_nv), ctypes.c_uint64(len(ef_polys)),
        _p(efp), _p(pts), ctypes.c_uint64(rounds), ctypes.c_uint64(code_rate), _p(fields),
        ctypes.c_uint64(fields.shape[0]), _p(ints), ctypes.c_uint64(ints.size), ctypes.c_int64(tamper))
    assert rc >= 0, err()
    return bool(rc)


def hash_columns(a, bits, ring64: bool = False):
    """a: (32, W) u64; bits: (ncols, W) u8 (column-major bit matrix). Returns H (32, ncols), sums
    mod p, or mod 2^64 with ``ring64``."""
    a = _u64(a)
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    ncols, W = bits.shape
    out = np.zeros((32, ncols), np.uint64)
    lib().lgso_hash_columns(_p(a), ctypes.c_uint64(W), _p(bits), ctypes.c_uint64(ncols), _p(out), int(ring64))
    return out


# ---- LigeSIS -------------------------------------------------------------------------------
def shape_ok(log_m: int, log_n: int) -> bool:
    return bool(lib().lgso_shape_ok(log_m, log_n))


class ProvingKey:
    def __init__(self, log_m: int, log_n: int, a: np.ndarray, commit_a: bool = True, ring64: bool = False):
        """``ring64``: the Z_2^64 subset-sum hash (commit only; no A-bar commitment)."""
        self.a = _u64(a)
        self.log_m, self.log_n = log_m, log_n
        self.ring64 = ring64
        self.ptr = lib().lgso_setup(log_m, log_n, _p(self.a), int(commit_a) | (2 if ring64 else 0))
        if not self.ptr:
            raise RuntimeError(err())

    def com_a(self) -> bytes:
        out = np.zeros(32, np.uint8)
        assert lib().lgso_pk_com_a(self.ptr, _p(out)) == 0
        return out.tobytes()

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().lgso_free_pk(self.ptr)
            self.ptr = None


class CommitData:
    def __init__(self, pk: ProvingKey, f: np.ndarray, keep_enc: bool = False):
        self.pk = pk
        f = _u64(f)
        root = np.zeros(32, np.uint8)
        self.ptr = lib().lgso_commit(pk.ptr, _p(f), _p(root), int(keep_enc))
        if not self.ptr:
            raise RuntimeError(err())
        self.root = root.tobytes()

    def enc(self) -> np.ndarray:
        m, n = 1 << self.pk.log_m, 1 << self.pk.log_n
        out = np.zeros((m, 2 * n), np.uint64)
        assert lib().lgso_cd_enc(self.ptr, _p(out)) == 0
        return out

    def h(self) -> np.ndarray:
        n = 1 << self.pk.log_n
        out = np.zeros((32, 2 * n), np.uint64)
        lib().lgso_cd_h(self.ptr, _p(out))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().lgso_free_cd(self.ptr)
            self.ptr = None


def prove(pk: ProvingKey, cd: CommitData, z, fields, ints, want_trace: bool = False):
    """Returns (proof bytes, (fields used, ints used)[, trace dict])."""
    z, fields, ints = _u64(z), _u64(fields), _u64(ints)
    used = np.zeros(2, np.uint64)
    n = 1 << pk.log_n
    trace = np.zeros(2 * n + 4 * n + 128, np.uint64) if want_trace else None
    need = lib().lgso_prove(pk.ptr, cd.ptr, _p(z), _p(fields), fields.shape[0], _p(ints), ints.size,
                            None, 0, None, None)
    if need < 0:
        raise RuntimeError(err())
    out = np.zeros(need, np.uint8)
    got = lib().lgso_prove(pk.ptr, cd.ptr, _p(z), _p(fields), fields.shape[0], _p(ints), ints.size,
                           _p(out), need, _p(used), _p(trace) if want_trace else None)
    assert got == need, err()
    res = (out.tobytes(), (int(used[0]), int(used[1])))
    if want_trace:
        tr = {
            "a_vec": trace[: 2 * n].reshape(n, 2),
            "ra": trace[2 * n: 6 * n].reshape(2 * n, 2),
            "idx": trace[6 * n:].astype(np.int64),
        }
        res = res + (tr,)
    return res


def verify(log_m: int, log_n: int, com_a: bytes, root: bytes, z, proof: bytes, fields, ints) -> bool:
    z, fields, ints = _u64(z), _u64(fields), _u64(ints)
    ca = np.frombuffer(com_a, np.uint8).copy()
    rt = np.frombuffer(root, np.uint8).copy()
    pf = np.frombuffer(proof, np.uint8).copy()
    rc = lib().lgso_verify(log_m, log_n, _p(ca), _p(rt), _p(z), _p(pf), len(proof), _p(fields),
                           fields.shape[0], _p(ints), ints.size)
    if rc < 0:
        raise RuntimeError(err())
    return bool(rc)


def proof_reencode(proof: bytes) -> bytes:
    pf = np.frombuffer(proof, np.uint8).copy()
    out = np.zeros(len(proof) + 64, np.uint8)
    n = lib().lgso_proof_reencode(_p(pf), len(proof), _p(out))
    if n < 0:
        raise ValueError("proof does not decode")
    return out[:n].tobytes()
