"""MI355X LigeSIS: host-side mirror of the reference API over libligesis_hip.so.

The reference API (pcs/src/ligesis.rs:205-756) is::

    Ligesis::setup(log_m, log_n, rng) -> (ProvingKey, VerifyingKey)        ligesis.rs:239
    Ligesis::commit(&pk, MlPoly f)   -> (LigesisCommit, CommitData)        ligesis.rs:277
    Ligesis::prove(&pk, &data, z, &mut RandomOracle) -> LigesisProof       ligesis.rs:347

This module keeps those names and argument meanings; the randomness that the reference draws
inside (`A` in setup, the oracle streams) is passed in explicitly. Shape and length violations
raise ``LigesisError`` where the reference panics. There is no CPU fallback: every operation
runs on the GPU through the C ABI in ``include/ligesis_hip.h`` and fails loudly when the
library or the GPU is missing.
"""
from __future__ import annotations

import sys as _sys
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libligesis_hip.so")

ETA = 64
C = 32
NUM_COLS_OPENED = 128
LOG_INTERLEAVE = 6

LGS_OK = 0
LGS_ERR_ARG = 1
LGS_ERR_SHAPE = 2
LGS_ERR_DEVICE = 3
LGS_ERR_NOMEM = 4
LGS_ERR_ORACLE = 5
LGS_ERR_BUFFER = 6
LGS_ERR_STATE = 7

LGS_SETUP_COMMIT_ONLY = 1
LGS_SETUP_RING64 = 2
LGS_INPUT_DEVICE = 1
LGS_BASEFOLD_EXT = 2
LGS_READ_ENCODED, LGS_READ_DIGESTS, LGS_READ_TREE = 0, 1, 2

_lib = None


class LigesisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libligesis_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        # PyTorch (device tensors) ships its own HIP runtime under the same soname; loaded after
        # this library, it finds the system runtime already in the process and sees no devices.
        # Loading it first makes both use one runtime.
        if "torch" not in _sys.modules:
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.lgs_init.argtypes = [ctypes.c_int, pp]
        L.lgs_shutdown.argtypes = [vp]
        L.lgs_last_error.argtypes = [vp]
        L.lgs_last_error.restype = ctypes.c_char_p
        L.lgs_set_option.argtypes = [vp, ctypes.c_char_p, i64]
        L.lgs_stage_times.argtypes = [vp, vp, u32]
        L.lgs_setup.argtypes = [vp, u32, u32, vp, u32, pp]
        L.lgs_key_com_a.argtypes = [vp, vp]
        L.lgs_basefold_commit.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, vp]
        L.lgs_key_release.argtypes = [vp]
        L.lgs_basefold_commit_state.argtypes = [vp, vp, u32, u32, u32, u32, vp, pp]
        L.lgs_basefold_prove.argtypes = [vp, vp, vp, u64, vp, u64, vp, u64, vp, ctypes.POINTER(ctypes.c_uint64), vp]
        L.lgs_basefold_release.argtypes = [vp]
        L.lgs_prove_batch.argtypes = [vp, u32, vp, vp, vp, vp, vp, u64, vp, u64, vp, vp, vp]
        L.lgs_commit.argtypes = [vp, vp, vp, u64, u32, vp, pp]
        L.lgs_commit_batch.argtypes = [vp, u32, vp, vp, vp, u32, vp, vp]
        L.lgs_commitment_release.argtypes = [vp]
        L.lgs_commit_noalloc.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp, vp]
        L.lgs_commitment_read.argtypes = [vp, u32, vp, u64]
        L.lgs_prove.argtypes = [vp, vp, vp, vp, u64, vp, u64, vp, u64, vp,
                                ctypes.POINTER(ctypes.c_uint64), vp]
        L.lgs_autotune.argtypes = [vp, vp, u32, vp, ctypes.POINTER(ctypes.c_double)]
        L.lgs_tune_get.argtypes = [vp, u32, u32, vp]
        L.lgs_tune_save.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        L.lgs_tune_load.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        L.lgs_tune_clear.argtypes = []
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


class Context:
    """One HIP device + stream (``lgs_init``)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().lgs_init(device, ctypes.byref(h))
        if rc != LGS_OK:
            raise LigesisError(rc, f"lgs_init(device={device}) failed")
        self.h = h

    def check(self, rc: int):
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_last_error(self.h).decode())

    def set_option(self, name: str, value: int):
        self.check(lib().lgs_set_option(self.h, name.encode(), int(value)))

    def stage_times(self) -> dict:
        out = np.zeros(5, np.float64)
        self.check(lib().lgs_stage_times(self.h, _ptr(out), 5))
        return {"ntt_ms": out[0], "hash_ms": out[1], "hash_finalize_ms": out[2],
                "merkle_ms": out[3], "commit_ms": out[4]}

    def autotune(self, key: "ProvingKey", reps: int = 5) -> dict:
        """Time the commit schedules of ``key``'s shape and keep the fastest (``lgs_autotune``)."""
        cfg = np.zeros(TUNE_FIELDS_COUNT, np.int32)
        ms = ctypes.c_double()
        self.check(lib().lgs_autotune(self.h, key.h, reps, _ptr(cfg), ctypes.byref(ms)))
        out = dict(zip(TUNE_FIELDS, (int(v) for v in cfg)))
        out["ms"] = ms.value
        return out

    def tuning(self, log_m: int, log_n: int) -> dict:
        """Schedule a commit of this shape would use now (``lgs_tune_get``)."""
        cfg = np.zeros(TUNE_FIELDS_COUNT, np.int32)
        self.check(lib().lgs_tune_get(self.h, log_m, log_n, _ptr(cfg)))
        return dict(zip(TUNE_FIELDS, (int(v) for v in cfg)))

    def close(self):
        if getattr(self, "h", None):
            lib().lgs_shutdown(self.h)
            self.h = None

    def __del__(self, _finalizing=_sys.is_finalizing):  # bound now: module globals go at exit
        if not _finalizing():
            self.close()


TUNE_FIELDS = ("hash_variant", "hash_impl", "hash_ksplit", "ntt_impl", "merkle_lg", "pipeline")
TUNE_FIELDS_COUNT = len(TUNE_FIELDS)


def tune_save(path: str) -> int:
    """Write the process tuning table (``lgs_tune_save``); returns the entry count."""
    n = ctypes.c_uint64()
    rc = lib().lgs_tune_save(None, os.fsencode(path), ctypes.byref(n))
    if rc != LGS_OK:
        raise LigesisError(rc, f"cannot write tuning table {path}")
    return n.value


def tune_load(path: str) -> int:
    """Merge a tuning table file into the process table (``lgs_tune_load``)."""
    n = ctypes.c_uint64()
    rc = lib().lgs_tune_load(None, os.fsencode(path), ctypes.byref(n))
    if rc != LGS_OK:
        raise LigesisError(rc, f"cannot load tuning table {path} (missing or malformed)")
    return n.value


def tune_clear():
    lib().lgs_tune_clear()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class VerifyingKey:
    """``VerifyingKey`` (ligesis.rs:77-83)."""

    def __init__(self, com_a: bytes, log_m: int, log_n: int):
        self.com_a, self.log_m, self.log_n = com_a, log_m, log_n
        self.width = ETA << log_m


class ProvingKey:
    """``ProvingKey`` (ligesis.rs:67-74): the hash matrix A and the A-bar Basefold state, in HBM."""

    def __init__(self, ctx: Context, log_m: int, log_n: int, a: np.ndarray, commit_only: bool = False,
                 ring64: bool = False):
        """``ring64``: the Z_2^64 subset-sum hash (sums mod 2^64, any 64-bit A), a commit-only
        variant the reference leaves out (subset_sum.rs:10-12)."""
        a = _u64(a)
        if a.size != C * (ETA << log_m):
            raise LigesisError(LGS_ERR_ARG, f"A must have {C} x {ETA << log_m} entries")
        h = ctypes.c_void_p()
        flags = (LGS_SETUP_COMMIT_ONLY if commit_only else 0) | (LGS_SETUP_RING64 if ring64 else 0)
        ctx.check(lib().lgs_setup(ctx.h, log_m, log_n, _ptr(a), flags, ctypes.byref(h)))
        self.ctx, self.h = ctx, h
        self.log_m, self.log_n = log_m, log_n
        self.commit_only = commit_only or ring64
        self.ring64 = ring64

    def com_a(self) -> bytes:
        out = np.zeros(32, np.uint8)
        self.ctx.check(lib().lgs_key_com_a(self.h, _ptr(out)))
        return out.tobytes()

    def release(self):
        if getattr(self, "h", None):
            lib().lgs_key_release(self.h)
            self.h = None

    def __del__(self, _finalizing=_sys.is_finalizing):  # bound now: module globals go at exit
        if not _finalizing():
            self.release()


class CommitData:
    """``CommitData`` (ligesis.rs:88-97): f, F', H and the Merkle tree, resident in HBM."""

    def __init__(self, ctx: Context, key: ProvingKey, h: ctypes.c_void_p, root: bytes, keepalive=None):
        self.ctx, self.key, self.h, self.root = ctx, key, h, root
        self._keepalive = keepalive

    def _read(self, what: int, shape) -> np.ndarray:
        out = np.zeros(shape, np.uint64)
        self.ctx.check(lib().lgs_commitment_read(self.h, what, _ptr(out), out.nbytes))
        return out

    def encoded(self) -> np.ndarray:
        m, n = 1 << self.key.log_m, 1 << self.key.log_n
        return self._read(LGS_READ_ENCODED, (m, 2 * n))

    def digests(self) -> np.ndarray:
        return self._read(LGS_READ_DIGESTS, (C, 2 << self.key.log_n))

    def tree(self) -> bytes:
        n2 = 2 << self.key.log_n
        out = np.zeros((2 * n2 - 1) * 32, np.uint8)
        self.ctx.check(lib().lgs_commitment_read(self.h, LGS_READ_TREE, _ptr(out), out.nbytes))
        return out.tobytes()

    def release(self):
        if getattr(self, "h", None):
            lib().lgs_commitment_release(self.h)
            self.h = None

    def __del__(self, _finalizing=_sys.is_finalizing):  # bound now: module globals go at exit
        if not _finalizing():
            self.release()


class RandomOracle:
    """``RandomOracle`` (utils/src/oracle.rs:10-64) with its pre-drawn streams given explicitly:
    ``fields`` is (k, 2) uint64 extension elements, ``ints`` is uint64."""

    def __init__(self, fields: np.ndarray, ints: np.ndarray):
        self.fields = _u64(fields).reshape(-1, 2)
        self.ints = _u64(ints)
        self.fields_idx = 0
        self.ints_idx = 0

    def restart(self):
        self.fields_idx = self.ints_idx = 0

    def drawn(self):
        return (self.fields_idx, self.ints_idx)


class Ligesis:
    """The reference's ``Ligesis<Goldilocks, BinomialExtensionField<Goldilocks, 2>>``."""

    @staticmethod
    def setup(log_m: int, log_n: int, a: np.ndarray, ctx: Context | None = None,
              commit_only: bool = False, ring64: bool = False):
        ctx = ctx or default_context()
        pk = ProvingKey(ctx, log_m, log_n, a, commit_only=commit_only, ring64=ring64)
        vk = VerifyingKey(b"" if pk.commit_only else pk.com_a(), log_m, log_n)
        return pk, vk

    @staticmethod
    def commit(pk: ProvingKey, f, device: bool = False):
        """Returns (root bytes, CommitData). With ``device=True``, ``f`` is a torch tensor on the
        GPU (uint64/int64 storage) that the commitment borrows."""
        ctx = pk.ctx
        n = 1 << (pk.log_m + pk.log_n)
        h = ctypes.c_void_p()
        root = np.zeros(32, np.uint8)
        if device:
            ptr, length, keep = ctypes.c_void_p(f.data_ptr()), f.numel(), f
            flags = LGS_INPUT_DEVICE
        else:
            fa = _u64(f)
            ptr, length, keep, flags = _ptr(fa), fa.size, None, 0
        if length != n:
            raise LigesisError(LGS_ERR_ARG, f"commit: expected {n} evaluations, got {length}")
        ctx.check(lib().lgs_commit(ctx.h, pk.h, ptr, length, flags, _ptr(root), ctypes.byref(h)))
        return root.tobytes(), CommitData(ctx, pk, h, root.tobytes(), keep)

    @staticmethod
    def commit_batch(keys, fs, ctx: Context | None = None, device: bool = False):
        """Commit many polynomials (any mix of shapes) in one batched pass; returns
        [(root, CommitData)], identical to per-polynomial ``commit``. With ``device=True`` the
        ``fs`` are GPU tensors that the commitments borrow."""
        ctx = ctx or (keys[0].ctx if keys else default_context())
        count = len(keys)
        if device:
            ptrs = [f.data_ptr() for f in fs]
            lens = np.array([f.numel() for f in fs], np.uint64)
            keep = list(fs)
        else:
            arrs = [_u64(f) for f in fs]
            ptrs = [a.ctypes.data for a in arrs]
            lens = np.array([a.size for a in arrs], np.uint64)
            keep = [None] * count
        kp = (ctypes.c_void_p * count)(*[k.h.value for k in keys])
        fp = (ctypes.c_void_p * count)(*ptrs)
        roots = np.zeros((count, 32), np.uint8)
        outs = (ctypes.c_void_p * count)()
        ctx.check(lib().lgs_commit_batch(ctx.h, count, kp, fp, _ptr(lens),
                                         LGS_INPUT_DEVICE if device else 0, _ptr(roots), outs))
        res = []
        for i in range(count):
            r = roots[i].tobytes()
            res.append((r, CommitData(ctx, keys[i], ctypes.c_void_p(outs[i]), r, keep[i])))
        return res

    @staticmethod
    def commit_noalloc(pk: ProvingKey, f, encoded, digests, tree, stream=None, want_root: bool = True):
        """``lgs_commit_noalloc``: the commit of ``commit`` with f and every output in
        caller-owned GPU tensors (F' m x 2n words, H 32 x 2n words, the tree (4n - 1) x 32 bytes);
        the library allocates nothing. ``stream``: a torch stream to enqueue on (None: the
        context's). Returns the root (waiting for the stream) or None with ``want_root=False``."""
        m, n = 1 << pk.log_m, 1 << pk.log_n
        need = {"f": (f, m * n * 8), "encoded": (encoded, m * 2 * n * 8), "digests": (digests, C * 2 * n * 8),
                "tree": (tree, (4 * n - 1) * 32)}
        for name, (t, nbytes) in need.items():
            if not t.is_cuda or not t.is_contiguous() or t.numel() * t.element_size() != nbytes:
                raise LigesisError(LGS_ERR_ARG, f"commit_noalloc: {name} must be a contiguous GPU tensor "
                                                f"of {nbytes} bytes")
        root = np.zeros(32, np.uint8)
        st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
        pk.ctx.check(lib().lgs_commit_noalloc(pk.ctx.h, pk.h, ctypes.c_void_p(f.data_ptr()), m * n,
                                              ctypes.c_void_p(encoded.data_ptr()), ctypes.c_void_p(digests.data_ptr()),
                                              ctypes.c_void_p(tree.data_ptr()), st,
                                              _ptr(root) if want_root else None))
        return root.tobytes() if want_root else None

    @staticmethod
    def prove(pk: ProvingKey, data: CommitData, z, oracle: RandomOracle) -> bytes:
        """Eval-prove (ligesis.rs:347): returns the canonical proof bytes and advances the
        oracle's cursors by the draws the proof consumed (the reference's RandomOracle is read
        from its start, i.e. call ``oracle.restart()`` first for a fresh transcript)."""
        ctx = pk.ctx
        z = _u64(z).reshape(-1, 2)
        if oracle.fields_idx or oracle.ints_idx:
            raise LigesisError(LGS_ERR_ARG, "prove reads the oracle streams from the start; restart() it")
        used = np.zeros(2, np.uint64)
        args = (ctx.h, pk.h, data.h, _ptr(z), z.shape[0], _ptr(oracle.fields), oracle.fields.shape[0],
                _ptr(oracle.ints), oracle.ints.size)
        # proofs of one key have one size: keep the buffer the last proof needed
        out = getattr(pk, "_proof_buf", None)
        if out is None:
            out = np.empty(4 << 20, np.uint8)
        cap = ctypes.c_uint64(out.size)
        rc = lib().lgs_prove(*args, _ptr(out), ctypes.byref(cap), _ptr(used))
        if rc == LGS_ERR_BUFFER:  # larger than the first guess: retry with the reported size
            out = np.empty(cap.value, np.uint8)
            rc = lib().lgs_prove(*args, _ptr(out), ctypes.byref(cap), _ptr(used))
        if rc == LGS_OK:
            pk._proof_buf = out
        ctx.check(rc)
        oracle.fields_idx, oracle.ints_idx = int(used[0]), int(used[1])
        return out[: cap.value].tobytes()

    @staticmethod
    def prove_batch(keys, datas, points, oracle: RandomOracle) -> list:
        """Consecutive eval-proofs on one RandomOracle (``lgs_prove_batch``): proof i proves
        ``datas[i]`` (committed with ``keys[i]``) at ``points[i]``, reading the oracle from where
        proof i - 1 stopped, exactly as sequential ``prove`` calls continuing one oracle's cursors
        would; their device work overlaps. Starts at the oracle's current cursors and advances
        them past every proof."""
        if not (len(keys) == len(datas) == len(points)):
            raise LigesisError(LGS_ERR_ARG, "prove_batch: keys, datas and points differ in length")
        return _prove_batch(keys, datas, points, oracle)


def prove_draws(log_m: int, log_n: int) -> tuple[int, int]:
    """(field, integer) draws one eval-proof at 2^log_m x 2^log_n takes from the RandomOracle: the
    schedule of pcs/src/ligesis.rs:390-530 and basefold.rs:732, 761, 788."""
    log_w, log_s = 6 + log_m, 7
    lb = log_w + log_s
    r = max(11 + log_m, 13 + log_m, log_n) - LOG_INTERLEAVE
    return (lb + 5 + lb + log_s + 2 * log_w + (log_n + 1) + log_n + 1 + lb + 1 + log_n + 3 + r,
            4 * NUM_COLS_OPENED + 100)


def _prove_batch(keys, datas, points, oracle: RandomOracle) -> list:
    """``lgs_prove_batch``: consecutive proofs on one oracle, read from its current cursors, which
    it then advances past every proof's draws."""
    count = len(keys)
    if count == 0:
        return []
    ctx = keys[0].ctx
    zs = [_u64(z).reshape(-1, 2) for z in points]
    fi, ii = oracle.drawn()
    fields, ints = oracle.fields[fi:], oracle.ints[ii:]
    # proofs of one key have one size: start from the size the key's last proof needed
    bufs = [np.empty(getattr(k, "_proof_buf", np.empty(4 << 20, np.uint8)).size, np.uint8) for k in keys]
    kp = (ctypes.c_void_p * count)(*[k.h.value for k in keys])
    cp = (ctypes.c_void_p * count)(*[d.h.value for d in datas])
    zp = (ctypes.c_void_p * count)(*[z.ctypes.data for z in zs])
    zl = np.array([z.shape[0] for z in zs], np.uint64)
    lens = np.array([b.size for b in bufs], np.uint64)
    used = np.zeros(2, np.uint64)
    fields_c, ints_c = np.ascontiguousarray(fields), np.ascontiguousarray(ints)

    def call():
        pp = (ctypes.c_void_p * count)(*[b.ctypes.data for b in bufs])
        return lib().lgs_prove_batch(ctx.h, count, kp, cp, zp, _ptr(zl), _ptr(fields_c), fields_c.shape[0],
                                     _ptr(ints_c), ints_c.size, pp, _ptr(lens), _ptr(used))

    rc = call()
    if rc == LGS_ERR_BUFFER:
        bufs = [np.empty(int(n), np.uint8) for n in lens]
        rc = call()
    ctx.check(rc)
    oracle.fields_idx, oracle.ints_idx = fi + int(used[0]), ii + int(used[1])
    return [b[: int(n)].tobytes() for b, n in zip(bufs, lens)]


class Basefold:
    """The reference's ``Basefold<Goldilocks, BinomialExtensionField<Goldilocks, 2>>`` commitment
    (pcs/src/basefold.rs:309-382): the root only, computed on the GPU."""

    LOG_INTERLEAVE = 6  # basefold.rs:16

    @staticmethod
    def _words(poly, ext: bool | None, device: bool):
        """Validates a polynomial for the C ABI: returns (pointer, num_vars, ext, keepalive).
        Base-field input is 2^nv canonical words; extension input is (2^nv, 2) (c0, c1) pairs.
        ``ext`` None takes the field from the shape; True / False demands that field (a 1-D
        base polynomial given as extension is upcast to (c0, 0) pairs, as MlPoly<EF> holds it)."""
        if device:
            import torch

            if poly.dtype not in (torch.int64, torch.uint64):
                raise LigesisError(LGS_ERR_ARG, "basefold commit: device polynomial must hold 8-byte words")
            if not poly.is_contiguous():
                raise LigesisError(LGS_ERR_ARG, "basefold commit: device polynomial must be contiguous")
            if poly.dim() == 2 and poly.shape[1] != 2:
                raise LigesisError(LGS_ERR_ARG, "basefold commit: extension values are (c0, c1) pairs")
            if poly.dim() not in (1, 2):
                raise LigesisError(LGS_ERR_ARG, "basefold commit: expected 1-D or (N, 2) input")
            is_ext = poly.dim() == 2
            if ext is not None and ext != is_ext:
                if ext and not is_ext:
                    poly = torch.stack([poly, torch.zeros_like(poly)], dim=1).contiguous()
                    is_ext = True
                else:
                    raise LigesisError(LGS_ERR_ARG, "basefold commit_base: polynomial must be base-field (1-D)")
            count = poly.shape[0]
            ptr, keep = ctypes.c_void_p(poly.data_ptr()), poly
        else:
            a = np.ascontiguousarray(_u64(poly))
            if a.ndim == 2 and a.shape[1] != 2:
                raise LigesisError(LGS_ERR_ARG, "basefold commit: extension values are (c0, c1) pairs")
            if a.ndim not in (1, 2):
                raise LigesisError(LGS_ERR_ARG, "basefold commit: expected 1-D or (N, 2) input")
            is_ext = a.ndim == 2
            if ext is not None and ext != is_ext:
                if ext:
                    a = np.ascontiguousarray(np.stack([a, np.zeros_like(a)], axis=1))
                    is_ext = True
                else:
                    raise LigesisError(LGS_ERR_ARG, "basefold commit_base: polynomial must be base-field (1-D)")
            count = a.shape[0]
            ptr, keep = _ptr(a), a
        nv = count.bit_length() - 1
        if count != 1 << nv:
            raise LigesisError(LGS_ERR_ARG, "basefold commit: length must be a power of two")
        return ptr, nv, is_ext, keep

    @staticmethod
    def commit_root(poly, code_rate: int, log_interleave: int, ctx: Context | None = None,
                    device: bool = False, ext: bool | None = None) -> bytes:
        """The root of ``commit_{base,ext}_on_domain`` only (``lgs_basefold_commit``): ``poly`` is
        2^nv base values, or (2^nv, 2) extension pairs (on the host, or with ``device=True`` a
        contiguous GPU tensor of 8-byte words); the prover state is not kept."""
        ptr, nv, is_ext, keep = Basefold._words(poly, ext, device)
        ctx = ctx or default_context()
        root = np.zeros(32, np.uint8)
        flags = (LGS_BASEFOLD_EXT if is_ext else 0) | (LGS_INPUT_DEVICE if device else 0)
        ctx.check(lib().lgs_basefold_commit(ctx.h, ptr, nv, flags, log_interleave, code_rate, _ptr(root)))
        del keep
        return root.tobytes()

    @staticmethod
    def _commit(poly, code_rate: int, log_interleave: int, ctx, device: bool, ext):
        ptr, nv, is_ext, keep = Basefold._words(poly, ext, device)
        ctx = ctx or default_context()
        root = np.zeros(32, np.uint8)
        h = ctypes.c_void_p()
        flags = (LGS_BASEFOLD_EXT if is_ext else 0) | (LGS_INPUT_DEVICE if device else 0)
        ctx.check(lib().lgs_basefold_commit_state(ctx.h, ptr, nv, flags, log_interleave, code_rate, _ptr(root),
                                                  ctypes.byref(h)))
        state = BasefoldProverState(ctx, h, nv, log_interleave, code_rate, is_ext, keep if device else None)
        return state, root.tobytes()

    @staticmethod
    def commit_base_on_domain(poly, code_rate: int, log_interleave: int, ctx: Context | None = None,
                              device: bool = False):
        """``commit_base_on_domain`` (basefold.rs:341-362): returns (prover state, root)."""
        return Basefold._commit(poly, code_rate, log_interleave, ctx, device, False)

    @staticmethod
    def commit_ext_on_domain(poly, code_rate: int, log_interleave: int, ctx: Context | None = None,
                             device: bool = False):
        """``commit_ext_on_domain`` (basefold.rs:364-382): returns (prover state, root); a 1-D base
        polynomial is taken as its upcast (c0, 0) pairs."""
        return Basefold._commit(poly, code_rate, log_interleave, ctx, device, True)

    @staticmethod
    def commit_base(poly, code_rate: int, ctx: Context | None = None, device: bool = False):
        """``commit_base`` (basefold.rs:309): a base-field polynomial, log_interleave 6."""
        return Basefold._commit(poly, code_rate, Basefold.LOG_INTERLEAVE, ctx, device, False)

    @staticmethod
    def commit_ext(poly, code_rate: int, ctx: Context | None = None, device: bool = False):
        """``commit_ext`` (basefold.rs:330): an extension-field polynomial, log_interleave 6."""
        return Basefold._commit(poly, code_rate, Basefold.LOG_INTERLEAVE, ctx, device, True)

    @staticmethod
    def prove(state: "BasefoldProverState", point, oracle: RandomOracle) -> bytes:
        """``Basefold::prove`` (basefold.rs:384-516): opens the committed polynomial at ``point``
        (num_vars extension elements). Reads the oracle from its start and advances its cursors
        by the draws used (one field per reduction round, then ceil(100 / code_rate) integers)."""
        ctx = state.ctx
        point = _u64(point).reshape(-1, 2)
        if oracle.fields_idx or oracle.ints_idx:
            raise LigesisError(LGS_ERR_ARG, "prove reads the oracle streams from the start; restart() it")
        used = np.zeros(2, np.uint64)
        args = (ctx.h, state.h, _ptr(point), point.shape[0], _ptr(oracle.fields), oracle.fields.shape[0],
                _ptr(oracle.ints), oracle.ints.size)
        out = np.empty(1 << 20, np.uint8)
        cap = ctypes.c_uint64(out.size)
        rc = lib().lgs_basefold_prove(*args, _ptr(out), ctypes.byref(cap), _ptr(used))
        if rc == LGS_ERR_BUFFER:
            out = np.empty(cap.value, np.uint8)
            rc = lib().lgs_basefold_prove(*args, _ptr(out), ctypes.byref(cap), _ptr(used))
        ctx.check(rc)
        oracle.fields_idx, oracle.ints_idx = int(used[0]), int(used[1])
        return out[: cap.value].tobytes()


class BasefoldProverState:
    """``BasefoldProverState`` (basefold.rs:36-54) resident in HBM: codewords, Merkle tree and the
    polynomial (``lgs_basefold_commit_state``)."""

    def __init__(self, ctx: Context, h, num_vars: int, log_interleave: int, code_rate: int, ext: bool, keep):
        self.ctx, self.h = ctx, h
        self.num_vars, self.log_interleave, self.code_rate, self.ext = num_vars, log_interleave, code_rate, ext
        self._keep = keep

    def release(self):
        if getattr(self, "h", None):
            lib().lgs_basefold_release(self.h)
            self.h = None

    def __del__(self, _finalizing=_sys.is_finalizing):  # bound now: module globals go at exit
        if not _finalizing():
            self.release()


LGS_PENDING = 8


class Server:
    """Serving loop (``lgs_server_*``): commit-then-open requests of mixed sizes, continuously
    batched under an HBM budget; each result is identical to ``commit`` + ``prove`` alone."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, prove_workers: int = 3,
                 max_batch: int = 16):
        L = lib()
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.lgs_server_start.argtypes = [ctypes.c_int, u64, u32, u32, ctypes.POINTER(ctypes.c_void_p)]
        L.lgs_server_submit.argtypes = [vp, vp, vp, u64, u32, vp, u64, vp, u64, vp, u64,
                                        ctypes.POINTER(ctypes.c_uint64)]
        L.lgs_server_poll.argtypes = [vp, u64, ctypes.c_int, vp, vp, ctypes.POINTER(ctypes.c_uint64), vp]
        L.lgs_server_stats.argtypes = [vp, vp]
        L.lgs_server_submit_commit.argtypes = [vp, vp, vp, u64, u32, ctypes.POINTER(ctypes.c_uint64)]
        L.lgs_server_submit_open.argtypes = [vp, u64, vp, u64, vp, u64, vp, u64, ctypes.POINTER(ctypes.c_uint64)]
        L.lgs_server_release.argtypes = [vp, u64]
        L.lgs_server_stop.argtypes = [vp]
        L.lgs_server_last_error.restype = ctypes.c_char_p
        h = ctypes.c_void_p()
        rc = L.lgs_server_start(device, hbm_budget, prove_workers, max_batch, ctypes.byref(h))
        if rc != LGS_OK:
            raise LigesisError(rc, L.lgs_server_last_error().decode())
        self.h = h
        self._keep = {}
        self._resident_keep = {}  # inputs of resident commitments (a device f is borrowed)

    def submit(self, key: ProvingKey, f, z, oracle: RandomOracle, device: bool = False) -> int:
        z = _u64(z).reshape(-1, 2)
        if device:
            fptr, flen, keep = ctypes.c_void_p(f.data_ptr()), f.numel(), f
        else:
            fa = _u64(f)
            fptr, flen, keep = _ptr(fa), fa.size, fa
        t = ctypes.c_uint64(0)
        rc = lib().lgs_server_submit(self.h, key.h, fptr, flen, LGS_INPUT_DEVICE if device else 0, _ptr(z),
                                     z.shape[0], _ptr(oracle.fields), oracle.fields.shape[0],
                                     _ptr(oracle.ints), oracle.ints.size, ctypes.byref(t))
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_server_last_error().decode())
        self._keep[t.value] = (keep, key)
        return t.value

    def submit_commit(self, key: ProvingKey, f, device: bool = False) -> int:
        """Commit-only request: the commitment stays resident in the server until ``release``;
        ``poll`` returns (root, b"", times)."""
        if device:
            fptr, flen, keep = ctypes.c_void_p(f.data_ptr()), f.numel(), f
        else:
            fa = _u64(f)
            fptr, flen, keep = _ptr(fa), fa.size, fa
        t = ctypes.c_uint64(0)
        rc = lib().lgs_server_submit_commit(self.h, key.h, fptr, flen, LGS_INPUT_DEVICE if device else 0,
                                            ctypes.byref(t))
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_server_last_error().decode())
        self._keep[t.value] = (keep, key)
        self._resident_keep[t.value] = (keep, key)
        return t.value

    def submit_open(self, commitment: int, z, oracle: RandomOracle) -> int:
        """Open request (eval-prove) of the resident commitment of a commit-only ticket."""
        z = _u64(z).reshape(-1, 2)
        t = ctypes.c_uint64(0)
        rc = lib().lgs_server_submit_open(self.h, commitment, _ptr(z), z.shape[0], _ptr(oracle.fields),
                                          oracle.fields.shape[0], _ptr(oracle.ints), oracle.ints.size,
                                          ctypes.byref(t))
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_server_last_error().decode())
        return t.value

    def release(self, commitment: int):
        """Releases a resident commitment (its memory returns once no open of it runs)."""
        rc = lib().lgs_server_release(self.h, commitment)
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_server_last_error().decode())
        self._resident_keep.pop(commitment, None)

    def poll(self, ticket: int, wait: bool = True):
        """Returns None while running, else (root, proof bytes, {commit_ms, prove_ms, total_ms})."""
        root = np.zeros(32, np.uint8)
        buf = np.zeros(4 << 20, np.uint8)
        ln = ctypes.c_uint64(buf.size)
        times = np.zeros(3, np.float64)
        rc = lib().lgs_server_poll(self.h, ticket, int(wait), _ptr(root), _ptr(buf), ctypes.byref(ln), _ptr(times))
        if rc == LGS_ERR_BUFFER:
            buf = np.zeros(ln.value, np.uint8)
            rc = lib().lgs_server_poll(self.h, ticket, int(wait), _ptr(root), _ptr(buf), ctypes.byref(ln),
                                       _ptr(times))
        if rc == LGS_PENDING:
            return None
        self._keep.pop(ticket, None)
        if rc != LGS_OK:
            raise LigesisError(rc, lib().lgs_server_last_error().decode())
        return root.tobytes(), buf[: ln.value].tobytes(), {
            "commit_ms": times[0], "prove_ms": times[1], "total_ms": times[2]}

    def stats(self) -> dict:
        out = np.zeros(4, np.uint64)
        lib().lgs_server_stats(self.h, _ptr(out))
        return {"batches": int(out[0]), "requests": int(out[1]), "reserved_bytes": int(out[2]),
                "queued": int(out[3])}

    def stop(self):
        if getattr(self, "h", None):
            lib().lgs_server_stop(self.h)
            self.h = None
            self._keep.clear()
            self._resident_keep.clear()

    def __del__(self, _finalizing=_sys.is_finalizing):  # bound now: module globals go at exit
        if not _finalizing():
            self.stop()
