"""Host-side checks of the PolyCommitmentScheme adapter (pcs/src/noop.rs:23-48): argument
validation happens before any device call, and batch_verify replays the consecutive transcript
through the caller's verifier, combines extension claims and never raises."""
from types import SimpleNamespace

import numpy as np
import pytest

from ligesis_amd import LigesisError, RandomOracle
from ligesis_amd.pcs import P, LigesisPcs, LigesisPcsData, PolyCommitmentScheme


def _scheme():
    return LigesisPcs(SimpleNamespace(log_m=2, log_n=6, ctx=None))


def test_is_a_poly_commitment_scheme():
    assert isinstance(_scheme(), PolyCommitmentScheme)
    assert _scheme().size == 256


@pytest.mark.parametrize("shape", [(255,), (256, 3), (128, 2), (2, 256)])
def test_commit_rejects_wrong_shapes(shape):
    with pytest.raises(LigesisError):
        _scheme().commit(np.zeros(shape, np.uint64))


def test_components_split_extension_values():
    s = _scheme()
    base, ext = s._components(np.arange(256, dtype=np.uint64))
    assert not ext and len(base) == 1
    poly = np.arange(512, dtype=np.uint64).reshape(256, 2)
    comps, ext = s._components(poly)
    assert ext
    np.testing.assert_array_equal(comps[0], poly[:, 0])
    np.testing.assert_array_equal(comps[1], poly[:, 1])


def test_batch_prove_rejects_foreign_data_and_bad_points():
    s = _scheme()
    ro = RandomOracle(np.zeros((8, 2), np.uint64), np.zeros(8, np.uint64))
    with pytest.raises(LigesisError):
        s.batch_prove([(object(), np.zeros((8, 2), np.uint64))], ro)
    with pytest.raises(LigesisError):
        s.batch_prove([(LigesisPcsData((), False), np.zeros((7, 2), np.uint64))], ro)
    assert ro.drawn() == (0, 0)


def test_batch_verify_without_a_verifier_returns_false():
    ro = RandomOracle(np.zeros((8, 2), np.uint64), np.zeros(8, np.uint64))
    assert _scheme().batch_verify([], [], ro) is False


def test_batch_verify_combines_claims_and_replays_the_transcript():
    # a fake verifier that records the transcript positions it is handed; the claims combine as
    # y = y0 + X y1 for extension polynomials (X^2 = 7)
    from ligesis_amd import prove_draws
    from ligesis_amd.pcs import LigesisPcsCommitment

    seen = []

    def verifier(log_m, log_n, com_a, root, z, proof, fields, ints):
        seen.append((root, fields.shape[0], ints.size))
        return True

    s = LigesisPcs(SimpleNamespace(log_m=2, log_n=6, ctx=None), com_a=bytes(32), verifier=verifier)
    nf, ni = prove_draws(2, 6)
    ro = RandomOracle(np.zeros((3 * nf + 5, 2), np.uint64), np.zeros(3 * ni + 5, np.uint64))
    y0, y1, yb = (3, 4), (5, 6), (9, 10)
    pf = lambda y: np.array(y, np.uint64).tobytes() + bytes(40)  # noqa: E731
    claims = [(LigesisPcsCommitment((b"a",)), np.zeros((8, 2), np.uint64), np.array(yb, np.uint64)),
              (LigesisPcsCommitment((b"b", b"c")), np.zeros((8, 2), np.uint64),
               np.array([(3 + 7 * 6) % P, (4 + 5) % P], np.uint64))]
    proof = [[pf(yb)], [pf(y0), pf(y1)]]
    assert s.batch_verify(claims, proof, ro)
    assert [r for r, _, _ in seen] == [b"a", b"b", b"c"]
    assert [f for _, f, _ in seen] == [3 * nf + 5, 2 * nf + 5, nf + 5]
    assert ro.drawn() == (3 * nf, 3 * ni)
    # a wrong combined value, a missing component, a malformed point: rejected, never raised
    ro.restart()
    bad = [claims[0], (claims[1][0], claims[1][1], np.array([1, 2], np.uint64))]
    assert not s.batch_verify(bad, proof, ro)
    ro.restart()
    assert not s.batch_verify(claims, [proof[0], proof[1][:1]], ro)
    ro.restart()
    assert not s.batch_verify([(claims[0][0], np.zeros(3, np.uint64), claims[0][2])], [proof[0]], ro)
    ro.restart()
    assert not s.batch_verify(claims, [[b"x"], proof[1]], ro)
