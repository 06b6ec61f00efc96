"""The allocation-free pairwise narrowphase (axm_collide_pairs_noalloc) against axm_collide_pairs,
through the C ABI: two calls in a row, the second with fewer pairs (the shared staging buffer keeps
the first call's contents), each bit-identical to the allocating entry point on the same input; a
call above the staging capacity is refused with OutOfRange."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(rng, n):
    prm = {0: lambda: [rng.uniform(0.25, 0.75), 0, 0], 1: lambda: list(rng.uniform(0.25, 0.75, 3)),
           2: lambda: [rng.uniform(0.2, 0.5), rng.uniform(0.2, 1.0), 0]}

    def quats(k):
        q = rng.normal(size=(k, 4))
        return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)

    ta = rng.integers(0, 3, n).astype(np.uint32)
    tb = rng.integers(0, 4, n).astype(np.uint32)
    pa = np.array([prm[int(t)]() for t in ta], np.float32)
    xa = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    pb = np.zeros((n, 3), np.float32)
    xb = np.zeros((n, 3), np.float32)
    for i in range(n):
        if tb[i] == 3:
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            pb[i] = nrm
            xb[i, 0] = nrm @ xa[i] + rng.uniform(-1.5, 0.3)
        else:
            pb[i] = prm[int(tb[i])]()
            xb[i] = xa[i] + rng.uniform(-1.6, 1.6, 3)
    return ta, pa, xa, quats(n), tb, pb, xb, quats(n)


def test_noalloc_matches_allocating_entry_point(axm, gpu):
    rng = np.random.default_rng(77)
    for n in (6000, 1500):
        args = _pairs(rng, n)
        ref = axm.collide_pairs(*args, margin=0.02)
        got = axm.collide_pairs(*args, margin=0.02, noalloc=True)
        assert got["status"] == ref["status"]
        for k in ("npts", "pos", "normal", "depth", "feature"):
            assert np.array_equal(got[k], ref[k]), (n, k)
        assert ref["npts"].sum() > n // 4


def test_noalloc_refuses_more_than_its_capacity(axm, gpu):
    rng = np.random.default_rng(78)
    args = _pairs(rng, 16385)
    got = axm.collide_pairs(*args, margin=0.02, noalloc=True)
    assert got["status"] == axm.OUT_OF_RANGE
