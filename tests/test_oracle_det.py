"""The determinant / local-energy / Sherman-Morrison-Woodbury oracle (oracle/determinant.py)
against what pins it:

* the reference's SMW tests (org/under_development/qmckl_sherman_morrison_woodbury.org:
  697-755, 2017-2046, 2709-2738, 3158-3188) on their data (share/qmckl/test_data/sm_test.h,
  tests/golden/sm_test.npz) at every leading dimension: the asserted determinants (1e-15) and
  Slater * Slater_inv = 1 for the updated inverse, tighter than the reference's 1e-3;
* the Slater gather against explicit picks, determinant and inverse against 2x2 and 3x3
  closed forms;
* kinetic energy and drift against central finite differences of det M(r) built from the C
  oracle's MO values, which a transposed inverse or a dropped spin block fails;
* the potential against a hand-computed two-electron, two-nucleus configuration.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import determinant as OD
from qmckl_amd import system as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIM = 21


@pytest.fixture(scope="module")
def sm():
    return np.load(os.path.join(GOLDEN, "sm_test.npz"))


def _identity_err(Slater, S_inv, LDS):
    res = Slater.reshape(DIM, DIM) @ S_inv.reshape(DIM, LDS)[:, :DIM]
    return np.abs(res - np.eye(DIM)).max()


def input_err(g, k, inv_name, LDS):
    """|S_before S_inv_before - 1|: the precision the data set's inverse comes with (its
    QMC run's), with S_before the Slater matrix after the updates less the update columns."""
    Sb = g(f"Slater{k}").reshape(DIM, DIM).copy()
    U = g(f"Updates{k}").reshape(-1, LDS)[:, :DIM]
    for l, i in enumerate(g(f"Updates_index{k}")):
        Sb[:, int(i) - 1] -= U[l]
    return _identity_err(Sb, g(inv_name), LDS)


# (data set, kernel, inverse array, determinant before, determinant after) of the reference tests
REF_CASES = [
    (1, "naive", "Slater_inv1", 3.407025646103221e-10, -4.120398385068217e-10),
    (2, "wb2", "Slater_inv2", -1.4432116661319376e-11, 2.367058141251457e-10),
    (3, "wb3", "Slater_inv3_1", -1.23743195512859e-09, 1.602708950725074e-10),
    (3, "splitting", "Slater_inv3_2", -1.23743195512859e-09, 1.602708950725074e-10),
]


def run_kernel(kind, LDS, U, idx, S_inv, det, breakdown=1e-3):
    if kind == "naive":
        return OD.sm_naive(LDS, DIM, U, idx, breakdown, S_inv, det)
    if kind == "splitting":
        return OD.sm_splitting(LDS, DIM, U, idx, breakdown, S_inv, det)
    return OD.woodbury(2 if kind == "wb2" else 3, LDS, DIM, U, idx, breakdown, S_inv, det)


@pytest.mark.parametrize("LDS", [24, 22, 21])
@pytest.mark.parametrize("case", REF_CASES, ids=lambda c: f"{c[1]}{c[0]}")
def test_smw_reference_cases(sm, LDS, case):
    k, kind, inv_name, det0, det1 = case
    g = lambda n: sm[f"lds{LDS}/{n}"]
    ok, S_new, det = run_kernel(kind, LDS, g(f"Updates{k}"), g(f"Updates_index{k}"), g(inv_name), det0)
    assert ok
    assert abs(det - det1) < 1e-15, (det, det1)
    # the reference's bound is 1e-3; the updated inverse is as good as the data's own
    err = _identity_err(g(f"Slater{k}"), S_new, LDS)
    assert err < min(1e-3, max(1e-9, 10 * input_err(g, k, inv_name, LDS))), err
    # the inverse before the updates does not invert the updated matrix
    assert _identity_err(g(f"Slater{k}"), g(inv_name), LDS) > 1e-3


@pytest.mark.parametrize("LDS", [24, 22, 21])
def test_smw_kernels_agree_on_every_data_set(sm, LDS):
    """Sequential rank-1 updates, their splitting variant and the Woodbury blocks describe the
    same matrix: they agree wherever none breaks down, also on the commented-out set 5.  On
    set 3 the first sequential denominator is 1.6e-14: naive reports the breakdown and
    splitting defers that update."""
    g = lambda n: sm[f"lds{LDS}/{n}"]
    for k, inv_name in ((1, "Slater_inv1"), (2, "Slater_inv2"), (3, "Slater_inv3_1"), (5, "Slater_inv5")):
        U, idx, S0 = g(f"Updates{k}"), g(f"Updates_index{k}"), g(inv_name)
        ok, A, da = OD.sm_splitting(LDS, DIM, U, idx, 1e-3, S0, 1.0)
        assert ok
        okn, B, db = OD.sm_naive(LDS, DIM, U, idx, 1e-3, S0, 1.0)
        assert okn == (k != 3)
        if okn:
            assert np.abs(A - B).max() <= 1e-10 * np.abs(A).max() and abs(da - db) <= 1e-12 * abs(da)
        else:
            assert np.array_equal(B, S0.reshape(DIM, LDS)) and db == 1.0
        assert _identity_err(g(f"Slater{k}"), A, LDS) < max(1e-9, 10 * input_err(g, k, inv_name, LDS))
        if len(idx) in (2, 3):
            ok, C, dc = OD.woodbury(len(idx), LDS, DIM, U, idx, 1e-3, S0, 1.0)
            assert ok and np.abs(A - C).max() <= 1e-8 * np.abs(A).max() and abs(da - dc) <= 1e-10 * abs(da)
        # the determinant ratio is det(S after) / det(S before) = det(S after) * det(S_inv before)
        # (to the data's precision)
        ratio = np.linalg.det(g(f"Slater{k}").reshape(DIM, DIM)) * np.linalg.det(S0.reshape(DIM, LDS)[:, :DIM])
        assert abs(da - ratio) <= max(1e-9, 100 * input_err(g, k, inv_name, LDS)) * abs(ratio)


def test_smw_breakdown_and_splitting():
    """A rank-1 update that makes the matrix singular: naive reports the breakdown and leaves
    its inputs; splitting applies it in halves (then fails only if the result is singular)."""
    rng = np.random.default_rng(0)
    n = 6
    A = rng.normal(size=(n, n))
    Ai = np.linalg.inv(A)
    # column 2 replaced by a combination of the others: singular
    u = A[:, [0, 1]] @ np.array([0.7, -1.1]) - A[:, 2]
    ok, S1, d = OD.sm_naive(n, n, u, np.array([3], np.uint64), 1e-6, Ai, 2.0)
    assert not ok and np.array_equal(S1, Ai) and d == 2.0
    # a near-breakdown update that splitting can apply: denominator 1 + e_2^T A^-1 u = 1e-4
    u2 = -A[:, 2] * (1 - 1e-4)
    ok, S2, d2 = OD.sm_splitting(n, n, u2, np.array([3], np.uint64), 1e-3, Ai, 1.0)
    A2 = A.copy()
    A2[:, 2] += u2
    assert ok and np.allclose(S2, np.linalg.inv(A2), rtol=1e-8, atol=1e-8 * np.abs(S2).max())
    assert abs(d2 - np.linalg.det(A2) / np.linalg.det(A)) <= 1e-8 * abs(d2)
    later = OD.sm_splitting_core(n, n, u2, np.array([3], np.uint64), 1e-3, Ai, 1.0)
    assert len(later[3]) == 1 and later[3][0] == 3 and np.allclose(later[2][0], u2 / 2)


def test_slater_gather_and_small_closed_forms():
    rng = np.random.default_rng(1)
    W, Ne, up, mo = 2, 5, 3, 7
    mv = rng.normal(size=(W * Ne, 5, mo))
    idx = np.array([[[2, 5, 7], [1, 3, 4]]])  # [det][W][up], 1-based
    v = OD.slater_vgl(mv, idx, W, Ne, 0, up)
    for w in range(W):
        for e in range(up):
            for m in range(up):
                for c in range(5):
                    assert v[0, w, c, e, m] == mv[w * Ne + e, c, idx[0, w, m] - 1]
    vb = OD.slater_vgl(mv, idx[:, :, :2], W, Ne, up, 2)
    assert vb[0, 1, 2, 1, 0] == mv[1 * Ne + up + 1, 2, idx[0, 1, 0] - 1]
    # 2x2 and 3x3 closed forms
    M2 = v[:, :, :, :2, :2].copy()
    det, inv, adj = OD.det_inverse(M2)
    a = M2[0, 0, 0]
    assert abs(det[0, 0] - (a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0])) < 1e-14
    assert np.allclose(adj[0, 0], [[a[1, 1], -a[0, 1]], [-a[1, 0], a[0, 0]]], atol=1e-14)
    b = v[0, 1, 0]
    cof = np.array([[b[1, 1] * b[2, 2] - b[1, 2] * b[2, 1], b[0, 2] * b[2, 1] - b[0, 1] * b[2, 2],
                     b[0, 1] * b[1, 2] - b[0, 2] * b[1, 1]],
                    [b[1, 2] * b[2, 0] - b[1, 0] * b[2, 2], b[0, 0] * b[2, 2] - b[0, 2] * b[2, 0],
                     b[0, 2] * b[1, 0] - b[0, 0] * b[1, 2]],
                    [b[1, 0] * b[2, 1] - b[1, 1] * b[2, 0], b[0, 1] * b[2, 0] - b[0, 0] * b[2, 1],
                     b[0, 0] * b[1, 1] - b[0, 1] * b[1, 0]]])
    det, inv, adj = OD.det_inverse(v)
    assert np.allclose(adj[0, 1], cof, atol=1e-12)
    assert abs(det[0, 1] - b[0] @ cof[:, 0]) < 1e-12


def _mo(s, x):
    return oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))


def test_kinetic_and_drift_match_finite_differences():
    s, x, W = S.workload("small", walk_num=2)
    Ne, up, dn = s.elec_num, s.up, s.dn
    rng = np.random.default_rng(2)
    mo_n = s.mo_num_kept
    ia = np.array([[rng.permutation(mo_n)[:up] + 1 for _ in range(W)]])
    ib = np.array([[rng.permutation(mo_n)[:dn] + 1 for _ in range(W)]])

    def dets(xx):
        mv = _mo(s, xx)
        va, vb = OD.slater_vgl(mv, ia, W, Ne, 0, up), OD.slater_vgl(mv, ib, W, Ne, up, dn)
        return va, vb, OD.det_inverse(va), OD.det_inverse(vb)

    va, vb, (da, ia_inv, _), (db, ib_inv, _) = dets(x)
    kin, drift = OD.kinetic_drift(va, ia_inv, vb, ib_inv)
    h = 1e-4
    D0 = da[0] * db[0]
    lap = np.zeros(W)
    fd_drift = np.zeros((W, Ne, 3))
    for e in range(Ne):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            for w in range(W):
                xp[d, w * Ne + e] += h
                xm[d, w * Ne + e] -= h
            _, _, (pa, _, _), (pb, _, _) = dets(xp)
            _, _, (ma, _, _), (mb, _, _) = dets(xm)
            Dp, Dm = pa[0] * pb[0], ma[0] * mb[0]
            fd_drift[:, e, d] = 2.0 * (Dp - Dm) / (2 * h * D0)
            lap += (Dp - 2 * D0 + Dm) / (h * h * D0)
    assert np.abs(drift - fd_drift).max() <= 1e-6 * max(1.0, np.abs(drift).max())
    assert np.abs(kin - (-0.5 * lap)).max() <= 1e-5 * max(1.0, np.abs(kin).max())
    # the transposed contraction (the under-development reference's loop) is not the derivative
    lap_t = -0.5 * np.einsum("wem,wem->w", va[0, :, 4], ia_inv[0]) - 0.5 * np.einsum("wem,wem->w", vb[0, :, 4], ib_inv[0])
    assert np.abs(lap_t - kin).max() > 1e-3 * np.abs(kin).max()


def test_potential_hand_example():
    # electrons at (0,0,0) and (0,0,2); nuclei Z=1 at (0,0,1) and Z=2 at (3,0,0)
    x = np.array([[0.0, 0.0], [0.0, 0.0], [0.0, 2.0]])
    Z = np.array([1.0, 2.0])
    R = np.array([[0.0, 3.0], [0.0, 0.0], [1.0, 0.0]])
    vee = 0.5
    ven = -(1 / 1.0 + 1 / 1.0 + 2 / 3.0 + 2 / np.sqrt(13.0))
    vnn = 2.0 / np.sqrt(10.0)
    assert abs(OD.potential(x, 1, 2, Z, R)[0] - (vee + ven + vnn)) < 1e-15
    # an electron on a nucleus and two coincident electrons are left out
    x2 = np.array([[0.0, 0.0], [0.0, 0.0], [1.0, 1.0]])
    assert abs(OD.potential(x2, 1, 2, Z, R)[0] - (-2 * 2 / np.sqrt(10.0) + vnn)) < 1e-15
