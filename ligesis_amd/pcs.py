"""``PolyCommitmentScheme`` (pcs/src/noop.rs:23-48) over the GPU LigeSIS commit and eval-prove.

The trait is batched: ``commit(poly) -> (prover_data, commitment)``, ``batch_prove(tasks, oracle)``
with ``tasks[i] = (prover_data, point)``, ``batch_verify(claims, proof, oracle)``. Its callers
(proofsys/src/commit.rs:24 binds ``Pcs = NoopPcs<EF>``) commit extension-field polynomials, while
LigeSIS commits base-field ones (pcs/src/ligesis.rs:277) and opens them at extension points
(pcs/src/ligesis.rs:347). The adapter covers both:

- a base-field polynomial (N uint64 values) is one LigeSIS commitment;
- an extension-field polynomial (N x 2 uint64, c0 + c1 X with X^2 = 7) is committed as its two
  base-field components f0, f1 in one batched commit. Multilinear evaluation is linear over
  F_p^2, so f(z) = f0(z) + X f1(z): the opening of f at z is the two component openings, and
  the verifier combines the two claimed values.

The trait's ``commit`` takes no key, so the scheme object carries one proving key (one shape,
pcs/src/ligesis.rs:239). ``batch_prove`` opens every component of every task in task order as
consecutive proofs on the caller's oracle: proof k reads the oracle's streams from where proof
k - 1 stopped, which is the transcript that sequential ``Ligesis::prove`` calls on one
``RandomOracle`` (utils/src/oracle.rs:10-64) give; the proofs' device work overlaps
(``lgs_prove_batch``). ``batch_verify`` replays the same transcript through a verifier the caller
supplies (the reference's CPU ``Ligesis::verify``, pcs/src/ligesis.rs:569-756): it checks that
each claimed value is the combination of its component proofs' values and that every component
proof verifies, and returns False (it never raises) on any mismatch.
"""
from __future__ import annotations

import abc
from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import (CommitData, Ligesis, LigesisError, ProvingKey, RandomOracle, LGS_ERR_ARG, prove_draws)

P = 0xFFFFFFFF00000001


class PolyCommitmentScheme(abc.ABC):
    """The reference trait (pcs/src/noop.rs:23-48): batched commitments to multilinear
    polynomials given by their evaluations, low-bit-first points."""

    @abc.abstractmethod
    def commit(self, poly):
        """-> (prover_data, commitment)"""

    @abc.abstractmethod
    def batch_prove(self, tasks, oracle):
        """tasks[i] = (prover_data, point) -> proof"""

    @abc.abstractmethod
    def batch_verify(self, claims, proof, oracle) -> bool:
        """claims[i] = (commitment, point, claimed value)"""


@dataclass(frozen=True)
class LigesisPcsCommitment:
    """Merkle roots of the committed components: one for a base-field polynomial, two (f0, f1)
    for an extension-field one."""

    roots: tuple


@dataclass
class LigesisPcsData:
    """The prover's side of one commitment: the resident LigeSIS commitments of its components."""

    parts: tuple
    extension: bool

    def release(self):
        for p in self.parts:
            p.release()


# verifier(log_m, log_n, com_a, root, z, proof, fields, ints) -> bool: the reference's
# Ligesis::verify with the RandomOracle given as its two streams read from the start
Verifier = Callable[[int, int, bytes, bytes, np.ndarray, bytes, np.ndarray, np.ndarray], bool]


def proof_value(proof: bytes) -> tuple[int, int]:
    """The claimed value y = f(z) of a LigeSIS proof: the first field of LigesisProof
    (ligesis.rs:99-119), an extension element (c0, c1) in the canonical encoding."""
    if len(proof) < 16:
        raise ValueError("proof too short")
    w = np.frombuffer(proof[:16], np.uint64)
    return int(w[0]), int(w[1])


def _combine(y0, y1):
    """y0 + X y1 in F_p[X]/(X^2 - 7): (a0 + a1 X) + X (b0 + b1 X) = a0 + 7 b1 + (a1 + b0) X."""
    return ((y0[0] + 7 * y1[1]) % P, (y0[1] + y1[0]) % P)


class LigesisPcs(PolyCommitmentScheme):
    """LigeSIS as a ``PolyCommitmentScheme`` for polynomials of 2^(log_m + log_n) evaluations.
    ``com_a`` (the verifying key's A-bar root) and ``verifier`` are what ``batch_verify`` needs."""

    def __init__(self, pk: ProvingKey, com_a: bytes | None = None, verifier: Verifier | None = None):
        self.pk = pk
        self.size = 1 << (pk.log_m + pk.log_n)
        self.com_a = com_a
        self.verifier = verifier

    def _components(self, poly) -> tuple[list[np.ndarray], bool]:
        a = np.asarray(poly)
        if a.dtype != np.uint64:
            a = a.astype(np.uint64)
        if a.ndim == 1 and a.size == self.size:
            return [np.ascontiguousarray(a)], False
        if a.ndim == 2 and a.shape == (self.size, 2):
            return [np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1])], True
        raise LigesisError(LGS_ERR_ARG, f"pcs commit: expected {self.size} base-field values or "
                                        f"({self.size}, 2) extension values, got shape {a.shape}")

    def commit(self, poly):
        comps, ext = self._components(poly)
        if ext:
            res = Ligesis.commit_batch([self.pk, self.pk], comps, ctx=self.pk.ctx)
        else:
            res = [Ligesis.commit(self.pk, comps[0])]
        parts = tuple(cd for _, cd in res)
        return LigesisPcsData(parts, ext), LigesisPcsCommitment(tuple(r for r, _ in res))

    def batch_prove(self, tasks, oracle: RandomOracle) -> list:
        """One list of proof bytes per task (one proof per component), consecutive on the oracle
        from its current cursors, which it leaves after the last proof's draws."""
        nvars = self.pk.log_m + self.pk.log_n
        for data, point in tasks:
            if not isinstance(data, LigesisPcsData) or any(not isinstance(p, CommitData) for p in data.parts):
                raise LigesisError(LGS_ERR_ARG, "pcs batch_prove: prover data must come from this scheme's commit")
            if np.asarray(point).reshape(-1, 2).shape[0] != nvars:
                raise LigesisError(LGS_ERR_ARG, f"pcs batch_prove: points have {nvars} extension coordinates")
        flat = [(cd, point) for data, point in tasks for cd in data.parts]
        proofs = Ligesis.prove_batch([self.pk] * len(flat), [cd for cd, _ in flat], [z for _, z in flat], oracle)
        out, k = [], 0
        for data, _ in tasks:
            out.append(proofs[k:k + len(data.parts)])
            k += len(data.parts)
        return out

    def batch_verify(self, claims, proof, oracle: RandomOracle) -> bool:
        """``claims[i] = (commitment, point, claimed value)`` against ``proof`` from batch_prove:
        every component proof must verify on the consecutive transcript (the verifier takes the
        draws the prover took, ``prove_draws``) and the claimed value must equal the component
        proofs' values combined (y0 + X y1 for an extension polynomial). Advances the oracle past
        every proof; returns False on any mismatch or malformed input."""
        if self.verifier is None or self.com_a is None:
            return False
        try:
            if len(claims) != len(proof):
                return False
            nf, ni = prove_draws(self.pk.log_m, self.pk.log_n)
            nvars = self.pk.log_m + self.pk.log_n
            fi, ii = oracle.drawn()
            ok = True
            for (com, point, value), proofs in zip(claims, proof):
                roots = tuple(getattr(com, "roots", ()))
                if len(roots) not in (1, 2) or len(proofs) != len(roots):
                    return False
                z = np.ascontiguousarray(np.asarray(point, dtype=np.uint64).reshape(-1, 2))
                if z.shape[0] != nvars:
                    return False
                ys = []
                for root, pf in zip(roots, proofs):
                    if fi + nf > oracle.fields.shape[0] or ii + ni > oracle.ints.size:
                        return False
                    ok = ok and bool(self.verifier(self.pk.log_m, self.pk.log_n, self.com_a, root, z, pf,
                                                   oracle.fields[fi:], oracle.ints[ii:]))
                    ys.append(proof_value(pf))
                    fi, ii = fi + nf, ii + ni
                y = ys[0] if len(ys) == 1 else _combine(ys[0], ys[1])
                v = np.asarray(value, dtype=np.uint64).reshape(-1)
                want = (int(v[0]), int(v[1]) if v.size > 1 else 0)
                ok = ok and y == want
            oracle.fields_idx, oracle.ints_idx = fi, ii
            return ok
        except (ValueError, TypeError, IndexError):
            return False
