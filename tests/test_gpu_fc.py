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
