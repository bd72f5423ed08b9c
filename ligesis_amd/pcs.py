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
pcs/src/ligesis.rs:239). ``batch_prove`` proves the tasks in order, each component after the
previous one, all drawing from the caller's oracle: every proof reads the oracle's streams from
its current cursors and advances them by what it consumed, which is the transcript that
sequential ``Ligesis::prove`` calls on one ``RandomOracle`` (utils/src/oracle.rs:10-64) give.

Verification is the reference's CPU verifier (pcs/src/ligesis.rs:569-756) and is not part of
this path; ``batch_verify`` says so instead of accepting.
"""
from __future__ import annotations

import abc
from dataclasses import dataclass

import numpy as np

from . import CommitData, Ligesis, LigesisError, ProvingKey, RandomOracle, LGS_ERR_ARG


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


class LigesisPcs(PolyCommitmentScheme):
    """LigeSIS as a ``PolyCommitmentScheme`` for polynomials of 2^(log_m + log_n) evaluations."""

    def __init__(self, pk: ProvingKey):
        self.pk = pk
        self.size = 1 << (pk.log_m + pk.log_n)

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
        """One list of proof bytes per task (one proof per component), proved in task order
        from the oracle's current cursors, which it leaves after the last proof's draws."""
        nvars = self.pk.log_m + self.pk.log_n
        for data, point in tasks:
            if not isinstance(data, LigesisPcsData) or any(not isinstance(p, CommitData) for p in data.parts):
                raise LigesisError(LGS_ERR_ARG, "pcs batch_prove: prover data must come from this scheme's commit")
            if np.asarray(point).reshape(-1, 2).shape[0] != nvars:
                raise LigesisError(LGS_ERR_ARG, f"pcs batch_prove: points have {nvars} extension coordinates")
        out = []
        for data, point in tasks:
            proofs = []
            for cd in data.parts:
                fi, ii = oracle.drawn()
                sub = RandomOracle(oracle.fields[fi:], oracle.ints[ii:])
                proofs.append(Ligesis.prove(self.pk, cd, point, sub))
                oracle.fields_idx, oracle.ints_idx = fi + sub.fields_idx, ii + sub.ints_idx
            out.append(proofs)
        return out

    def batch_verify(self, claims, proof, oracle) -> bool:
        raise NotImplementedError("LigeSIS verification is the reference's CPU verifier "
                                  "(pcs/src/ligesis.rs:569-756); the GPU path proves only")
