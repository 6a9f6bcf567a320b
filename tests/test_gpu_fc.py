"""Cohesion / friction state on the device (hfpx_match_f_c,
hfpx_update_mrd) against the oracle, which is pinned bitwise to the
reference build's match_f_c / update_mrd (GPU)."""
import numpy as np
import pytest

import oracle_lib
from test_oracle_fc import CMID, PAR, random_state

pytestmark = pytest.mark.gpu

O = oracle_lib.oracle()


@pytest.fixture(scope="module")
def hx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return hfp3d


def props(hx):
    return hx.Properties_T(n_surface=len(PAR), c_model_id=CMID,
                           f_c_param=[hx.F_C_Param_T(*row) for row in PAR])


@pytest.mark.parametrize("n", [1, 1000, 300001])
def test_match_f_c_and_update_mrd_match_oracle(hx, n):
    print(f"case: friction/cohesion n_nod {n}", flush=True)
    dd, mat, mo, ms = random_state(n, 3 + n)
    prop = props(hx)
    fs = hx.Frac_State_T.zeros(n)
    fs.mr_open, fs.mr_slip = mo.copy(), ms.copy()
    hx.match_f_c(prop, mat, dd, False, fs)
    fc, sc, oc = O.match_f_c(dd, mat, CMID, PAR, mo, ms)
    # the device may contract a * b + c into one rounding: 1-ulp differences
    for a, b in ((fs.friction_coef, fc), (fs.slip_cohesion, sc), (fs.open_cohesion, oc)):
        assert np.abs(a - b).max() <= 4e-16 * max(1.0, np.abs(b).max())
    hx.update_mrd(prop, mat, dd, False, fs)
    mo2, ms2 = O.update_mrd(dd, mat, PAR, mo, ms)
    assert np.abs(fs.mr_open - mo2).max() <= 4e-16 * max(1.0, np.abs(mo2).max())
    assert np.abs(fs.mr_slip - ms2).max() <= 4e-16 * max(1.0, np.abs(ms2).max())


def test_bad_material_id_and_zero_critical_values_are_rejected(hx):
    from hfpx3d_amd import _lib
    print("case: friction/cohesion argument checks", flush=True)
    dd, mat, mo, ms = random_state(100, 9)
    prop = props(hx)
    fs = hx.Frac_State_T.zeros(100)
    bad = mat.copy()
    bad[17] = len(PAR)
    with pytest.raises(Exception):
        hx.match_f_c(prop, bad, dd, False, fs)
    bad[17] = -1
    with pytest.raises(Exception):
        hx.update_mrd(prop, bad, dd, False, fs)
    prop.f_c_param[2].cr_slip = 0.0
    with pytest.raises(Exception):
        hx.update_mrd(prop, mat, dd, False, fs)
    # the context is still usable afterwards
    prop = props(hx)
    hx.match_f_c(prop, mat, dd, False, fs)
    fc, _, _ = O.match_f_c(dd, mat, CMID, PAR, fs.mr_open, fs.mr_slip)
    assert np.abs(fs.friction_coef - fc).max() < 1e-15
    assert _lib.HFPX_ERR_INVALID == 1


def test_cpp_host_mirror_matches_oracle():
    """hfp3d::match_f_c / update_mrd of the C++ host mirror, through the
    hfpx3d_fc_check program, against the oracle on the inputs it prints."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "native", "hfpx3d_fc_check")
    print("case: C++ host mirror of the friction/cohesion laws, 2000 nodes", flush=True)
    r = subprocess.run([exe, "2000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    n = len(d["mat"])
    dd = np.array(d["dd"]).reshape(n, 3, order="F")
    mat = np.array(d["mat"], dtype=np.int32)
    par = np.array([[1.0, 2.0, 0.5, 3.0, 0.6, 0.2], [0.5, 1.5, 1.0, 0.5, 0.8, 0.1], [2.0, 1.0, 2.0, 1.0, 0.7, 0.3]])
    cmid = np.array([[0, 0], [1, 1], [2, 1]], dtype=np.int32)
    fc, sc, oc = O.match_f_c(dd, mat, cmid, par, d["mr_open0"], d["mr_slip0"])
    mo, ms = O.update_mrd(dd, mat, par, d["mr_open0"], d["mr_slip0"])
    for got, want in ((d["friction_coef"], fc), (d["slip_cohesion"], sc), (d["open_cohesion"], oc),
                      (d["mr_open"], mo), (d["mr_slip"], ms)):
        assert np.abs(np.array(got) - want).max() <= 4e-16 * max(1.0, np.abs(want).max())
    assert (oc > 0).any() and (fc != fc[0]).any()
