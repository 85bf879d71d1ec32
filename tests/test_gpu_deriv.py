"""GPU parity of the Jastrow parameter derivatives and nuclear forces: every getter, called
through the C-ABI, against the oracle (oracle/qmckl_oracle_deriv.c, pinned by finite
differences in tests/test_oracle_deriv.py).

Tolerance: max |gpu - oracle| <= 1e-11 * max |oracle| per output (FP64; the two sum the
same terms in different orders).
"""
import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu

GETTERS = {
    "asymp_jasb_pderiv": "qmckl_get_jastrow_champ_asymp_jasb_pderiv",
    "asymp_jasa_pderiv": "qmckl_get_jastrow_champ_asymp_jasa_pderiv",
    "factor_ee_pderiv": "qmckl_get_jastrow_champ_factor_ee_pderiv",
    "factor_ee_gl_pderiv": "qmckl_get_jastrow_champ_factor_ee_gl_pderiv",
    "factor_en_pderiv": "qmckl_get_jastrow_champ_factor_en_pderiv",
    "factor_en_gl_pderiv": "qmckl_get_jastrow_champ_factor_en_gl_pderiv",
    "factor_een_pderiv": "qmckl_get_jastrow_champ_factor_een_pderiv",
    "factor_een_gl_pderiv": "qmckl_get_jastrow_champ_factor_een_gl_pderiv",
    "forces_jastrow_en": "qmckl_get_forces_jastrow_en",
    "forces_jastrow_en_g": "qmckl_get_forces_jastrow_en_g",
    "forces_jastrow_en_l": "qmckl_get_forces_jastrow_en_l",
    "forces_jastrow_een": "qmckl_get_forces_jastrow_een",
    "forces_jastrow_een_g": "qmckl_get_forces_jastrow_een_g",
    "forces_jastrow_een_l": "qmckl_get_forces_jastrow_een_l",
}


def _check_all(ctx, s, x, W, tag):
    ref = oracle.jastrow_derivatives(s, x, W)
    for key, fn in GETTERS.items():
        r = ref[key]
        got = q.get_array(ctx, fn, r.size).reshape(r.shape)
        scale = max(np.abs(r).max(), 1e-300)
        err = np.abs(got - r).max()
        assert err <= 1e-11 * scale, f"{tag} {key}: {err:.3e} vs scale {scale:.3e}"


@pytest.mark.parametrize("cfg", [(2, 7, 7, 5, 5, 5, 16), (3, 3, 2, 5, 5, 5, 8), (5, 9, 8, 3, 4, 3, 4)])
def test_derivatives_vs_oracle(cfg):
    *sys_cfg, W = cfg
    s, x, W = S.jastrow_workload(*sys_cfg, walk_num=W)
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    _check_all(ctx, s, x, W, s.name)
    q.context_destroy(ctx)


def test_derivatives_small_workload():
    """The 14-electron N2-like system of the differentiation benchmark, 64 walkers."""
    s, x, W = S.workload("small", walk_num=64)
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    _check_all(ctx, s, x, W, "small")
    q.context_destroy(ctx)


def test_derivatives_follow_coordinate_updates():
    s, x, W = S.jastrow_workload(3, 3, 2, 5, 5, 5, walk_num=4)
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    f1 = q.get_array(ctx, "qmckl_get_forces_jastrow_een", W * s.nucl_num * 3)
    x2 = x + 0.05
    S.set_coords(ctx, s, x2, W)
    f2 = q.get_array(ctx, "qmckl_get_forces_jastrow_een", W * s.nucl_num * 3)
    assert not np.array_equal(f1, f2)
    ref = oracle.jastrow_derivatives(s, x2, W)["forces_jastrow_een"].ravel()
    assert np.abs(f2 - ref).max() <= 1e-11 * np.abs(ref).max()
    q.context_destroy(ctx)


def test_derivatives_in_walker_ranges(monkeypatch):
    """With a 1 MB budget (QMCKL_AMD_DERIV_BUDGET_MB) every getter materialises the
    intermediates one walker range at a time (here one walker: about 1.5 MB each); the
    results equal the whole-batch evaluation bit for bit (the walker sums run once, in walker
    order) and the oracle."""
    s, x, W = S.jastrow_workload(5, 20, 20, 5, 5, 5, walk_num=24)
    ctx = q.context_create()
    try:
        S.load(ctx, s, mo=False)
        S.set_coords(ctx, s, x, W)
        ref = oracle.jastrow_derivatives(s, x, W)
        monkeypatch.delenv("QMCKL_AMD_DERIV_BUDGET_MB", raising=False)
        whole = {k: q.get_array(ctx, fn, ref[k].size) for k, fn in GETTERS.items()}
        monkeypatch.setenv("QMCKL_AMD_DERIV_BUDGET_MB", "1")
        S.set_coords(ctx, s, x, W)  # a new context date: nothing cached
        for k, fn in GETTERS.items():
            got = q.get_array(ctx, fn, ref[k].size)
            assert np.array_equal(got, whole[k]), f"{k}: walker ranges change the result"
            r = ref[k].ravel()
            assert np.abs(got - r).max() <= 1e-11 * max(np.abs(r).max(), 1e-300), k
    finally:
        monkeypatch.delenv("QMCKL_AMD_DERIV_BUDGET_MB", raising=False)
        q.context_destroy(ctx)
