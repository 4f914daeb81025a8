"""CPU: seeded batch shapes (uneven lines) and the ragged-line mesh builder."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import closed_forms as cf
import synth_np


def test_shapes_meet_the_configuration(tbs):
    sh = tbs.batch_shapes(3, nsys=64, total=384000, lmin=96, lmax=1280, nlmin=8, nlmax=40)
    assert len(sh) == 64
    nl = np.array([len(s) for s in sh])
    L = np.concatenate(sh)
    assert nl.min() >= 8 and nl.max() <= 40
    assert L.sum() == 384000 and L.min() >= 96 and L.max() <= 1280


def test_shapes_are_seeded(tbs):
    a, b, c = tbs.batch_shapes(7), tbs.batch_shapes(7), tbs.batch_shapes(8)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not all(len(x) == len(y) and np.array_equal(x, y) for x, y in zip(a, c))


def test_unscaled_shapes_are_uniform_draws(tbs):
    sh = tbs.batch_shapes(11, nsys=400, total=0, lmin=96, lmax=1280, nlmin=8, nlmax=40)
    L = np.concatenate(sh)
    assert L.min() >= 96 and L.max() <= 1280 and abs(L.mean() - 688) < 25
    nl = np.array([len(s) for s in sh])
    assert abs(nl.mean() - 24) < 1.5


def test_infeasible_total_is_an_error(tbs):
    with pytest.raises(tbs.TbsError):
        tbs.batch_shapes(1, nsys=2, total=10, lmin=96, lmax=1280, nlmin=8, nlmax=8)


def test_ragged_equal_lengths_is_structured(tbs):
    a, b = tbs.ragged_mesh([37] * 11), tbs.structured_mesh(11, 37)
    for f in ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface"):
        assert np.array_equal(getattr(a, f), getattr(b, f))


@pytest.mark.parametrize("lens", [[5, 1, 9, 3], [96, 130, 97, 200, 96], [1], [3, 3, 1, 1, 7]])
def test_ragged_mesh_is_a_consistent_system(tbs, oracle, lens):
    m = tbs.ragged_mesh(lens, nvar=3)
    jac = synth_np.generate(5, 3, m.nelem, m.nintface, 8.0)
    Dia, DinvC = oracle.lineLU(m, jac["jacD"], jac["jacO1"], jac["jacO2"])
    A = oracle.assembleTriBlocks(m, jac["jacO1"], jac["jacO2"])
    R = np.random.default_rng(1).uniform(-1, 1, 3 * m.nelem)
    dU = oracle.solveDU(m, Dia, DinvC, A, R)
    assert np.allclose(dU, -spla.splu(cf.assemble_M(m, jac).tocsc()).solve(R), rtol=1e-12, atol=1e-14)
    U = np.random.default_rng(2).uniform(-1, 1, 3 * m.nelem)
    Rr = oracle.lineRes(m, jac["jacD"], jac["jacO1"], jac["jacO2"], U, jac["rhs"])
    assert np.allclose(Rr, jac["rhs"] + cf.assemble_J(m, jac) @ U, rtol=1e-13)
    # every face appears in the ef lists of both its elements
    seen = np.zeros(m.nintface, int)
    for f in m.ef[m.ef >= 0]:
        seen[f] += 1
    assert np.all(seen == 2)
