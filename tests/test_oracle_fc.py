"""Cohesion / friction state (src/Development/cohesion_friction.cpp) in the
oracle (CPU): match_f_c and update_mrd pinned bitwise against the reference
build on random states that reach every branch, and against hand-worked
values of the cohesion and friction laws."""
import numpy as np
import pytest

import oracle_lib

O = oracle_lib.oracle()
REF = oracle_lib.reference_or_none()

# cr_open, peak_ts, cr_slip, peak_sc, peak_sf, res_sf
PAR = np.array([[1.0, 2.0, 0.5, 3.0, 0.6, 0.2],
                [0.5, 1.5, 1.0, 0.5, 0.8, 0.1],
                [2.0, 1.0, 2.0, 1.0, 0.7, 0.3],
                [1.5, 0.7, 0.8, 2.5, 0.9, 0.4]])
CMID = np.array([[0, 0], [1, 1], [2, 2], [7, 1]], dtype=np.int32)   # id 7: unknown model


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    mat = rng.integers(0, len(PAR), n)
    dd = rng.normal(scale=0.8, size=(n, 3))
    dd[::7, 2] = -np.abs(dd[::7, 2])           # closed nodes
    dd[1::7, 2] = 3.0 * PAR[mat[1::7], 0]      # fully open nodes
    mr_open = rng.uniform(0.0, 1.2, n)
    mr_slip = rng.uniform(0.0, 1.2, n)
    return dd, mat, mr_open, mr_slip


@pytest.mark.skipif(REF is None, reason="reference build (oracle/_ref) not available")
def test_match_f_c_matches_reference_build_bitwise():
    dd, mat, mo, ms = random_state(20000, 1)
    a = O.match_f_c(dd, mat, CMID, PAR, mo, ms)
    b = REF.match_f_c(dd, mat, CMID, PAR, mo, ms)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    fc, sc, oc = a
    # every branch was reached: cohesion on and off, all friction levels
    assert (oc > 0).any() and (oc == 0).any() and (fc == 0).any()
    assert len(np.unique(np.round(fc, 12))) > 10


@pytest.mark.skipif(REF is None, reason="reference build (oracle/_ref) not available")
def test_update_mrd_matches_reference_build_bitwise():
    dd, mat, mo, ms = random_state(20000, 2)
    a = O.update_mrd(dd, mat, PAR, mo, ms)
    b = REF.update_mrd(dd, mat, PAR, mo, ms)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert (a[0] != mo).any() and (a[1] != ms).any()


def node(mat, shear, opening, mr_open=0.0, mr_slip=0.0):
    dd = np.array([[shear, 0.0, opening]])
    r = O.match_f_c(dd, [mat], CMID, PAR, [mr_open], [mr_slip])
    return tuple(float(v[0]) for v in r)     # friction, shear cohesion, opening cohesion


def test_cohesion_and_friction_laws_by_hand():
    # Dugdale box (material 0): full cohesion while partly open
    fc, sc, oc = node(0, 0.0, 0.5)
    assert (oc, sc) == (2.0, 3.0) and fc == 0.6
    # unloading below the largest opening reached scales it linearly
    fc, sc, oc = node(0, 0.0, 0.25, mr_open=0.5)
    assert oc == 2.0 * 0.5
    # linear softening (material 1, cr_open 0.5): opening 0.125 -> 1/4 open
    fc, sc, oc = node(1, 0.0, 0.125)
    assert abs(oc - 1.5 * 0.75) < 1e-15 and abs(sc - 0.5 * 0.75) < 1e-15
    # linear slip weakening from the larger of slip and reached slip
    fc, _, _ = node(1, 0.4, 0.1, mr_slip=0.6)
    assert abs(fc - (0.1 + (0.8 - 0.1) * 0.4)) < 1e-15
    fc, _, _ = node(1, 0.6, 0.1, mr_slip=0.2)
    assert abs(fc - (0.1 + (0.8 - 0.1) * 0.4)) < 1e-15
    # beyond the critical slip: residual friction; fully open: none
    assert node(1, 1.5, 0.1)[0] == 0.1
    assert node(1, 0.0, 0.6) == (0.0, 0.0, 0.0)
    # closed node: no cohesion, friction still acts
    assert node(0, 0.0, -0.1) == (0.6, 0.0, 0.0)
    # model 2 and unknown ids: no cohesion, residual friction
    assert node(2, 0.1, 0.5) == (0.3, 0.0, 0.0)
    fc, sc, oc = node(3, 0.2, 0.3)
    assert oc == 0.0 and abs(fc - (0.4 + (0.9 - 0.4) * (1 - 0.25))) < 1e-15


def test_update_mrd_by_hand():
    dd = np.array([[0.3, 0.4, 0.5], [0.0, 0.0, 5.0], [0.0, 0.0, 0.1]])
    mo, ms = O.update_mrd(dd, [0, 0, 0], PAR, [0.2, 0.9, 0.7], [0.5, 0.0, 2.0])
    # opening 0.5 > 0.2; a full opening (5) beats the previous 0.9; 0.1 < 0.7 keeps it
    assert list(mo) == [0.5, 5.0, 0.7]
    # slip |(0.3, 0.4)| / 0.5 = 1 > 0.5; zero slip keeps the reached values
    assert abs(ms[0] - 1.0) < 1e-15 and list(ms[1:]) == [0.0, 2.0]
