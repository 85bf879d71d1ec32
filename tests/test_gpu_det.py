"""GPU: Slater determinants, the local energy and the Sherman-Morrison-Woodbury updates through
the C-ABI, against oracle/determinant.py.

* det_vgl is the gather of the context's own MO VGL (bit for bit); determinant and inverse
  of those matrices against LAPACK (numpy) within the matrices' condition number, on up to
  150 electrons per spin (past 128 the matrices are worked on in global memory), with
  several determinants per walker; the adjugate is det times the inverse;
* kinetic energy and drift (one determinant per spin) against the oracle's Jacobi formula on
  the same MO values, potential and local energy against the oracle's potential;
* the SMW entry points on the reference's test data (sm_test.h): the determinants its tests
  assert (1e-15) and the oracle's updated inverse; a breakdown returns QMCKL_FAILURE and
  leaves the inverse; the splitting core returns the oracle's deferred halves;
* the batched device entry point on 48 random inverses of order 21 and 150 per kind, with a
  planted breakdown per batch (status 1, matrix untouched).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
import qmckl_amd as q
from oracle import determinant as OD
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _system(name, W, extra_mo):
    if name == "mol300":
        s, x, W = S.molecule_workload(300, walk_num=W)
    else:
        s, x, W = S.workload(name, walk_num=W)
    s.keep_mo = None
    if extra_mo:
        rng = np.random.default_rng(4)
        s.mo_coef = S._orthonormal_rows(rng, s.mo_num + extra_mo, s.ao_num)
    return s, x, W


def _ctx(s, x, W, det_a, det_b, seed=0):
    ctx = q.context_create()
    S.load(ctx, s, jastrow=False)
    S.set_coords(ctx, s, x, W)
    L = q.lib()
    rng = np.random.default_rng(seed)
    ia = np.array([[rng.choice(s.mo_num, s.up, replace=False) + 1 for _ in range(W)] for _ in range(det_a)])
    ib = np.array([[rng.choice(s.mo_num, s.dn, replace=False) + 1 for _ in range(W)] for _ in range(det_b)])
    q.check(ctx, L.qmckl_set_determinant_type(ctx, b"G"))
    q.check(ctx, L.qmckl_set_determinant_det_num_alpha(ctx, det_a))
    q.check(ctx, L.qmckl_set_determinant_det_num_beta(ctx, det_b))
    a, b = np.ascontiguousarray(ia, np.int64), np.ascontiguousarray(ib, np.int64)
    q.check(ctx, L.qmckl_set_determinant_mo_index_alpha(ctx, q.ptr(a), a.size))
    q.check(ctx, L.qmckl_set_determinant_mo_index_beta(ctx, q.ptr(b), b.size))
    assert L.qmckl_determinant_provided(ctx)
    return ctx, ia, ib


def _get(ctx, fn, n):
    out = np.empty(int(n))
    q.check(ctx, getattr(q.lib(), fn)(ctx, q.ptr(out)), fn)
    return out


@pytest.mark.parametrize("name,W,extra,det_a,det_b", [("small", 16, 0, 3, 2), ("medium", 4, 20, 2, 1),
                                                      ("mol300", 2, 20, 2, 1)])
def test_determinants_and_inverses(name, W, extra, det_a, det_b):
    print(f"det {name} W={W} dets {det_a}/{det_b}", flush=True)
    s, x, W = _system(name, W, extra)
    ctx, ia, ib = _ctx(s, x, W, det_a, det_b)
    Ne, mo = s.elec_num, s.mo_num
    mv = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * mo).reshape(W * Ne, 5, mo)
    for spin, (idx, n, e0, nd) in enumerate(((ia, s.up, 0, det_a), (ib, s.dn, s.up, det_b))):
        sfx = "alpha" if spin == 0 else "beta"
        vgl = _get(ctx, f"qmckl_get_det_vgl_{sfx}", nd * W * 5 * n * n).reshape(nd, W, 5, n, n)
        ref_vgl = OD.slater_vgl(mv, idx, W, Ne, e0, n)
        np.testing.assert_array_equal(vgl, ref_vgl)
        det = _get(ctx, f"qmckl_get_det_{sfx}", nd * W).reshape(nd, W)
        inv = _get(ctx, f"qmckl_get_det_inv_matrix_{sfx}", nd * W * n * n).reshape(nd, W, n, n)
        adj = _get(ctx, f"qmckl_get_det_adj_matrix_{sfx}", nd * W * n * n).reshape(nd, W, n, n)
        rdet, rinv, _ = OD.det_inverse(ref_vgl)
        for d in range(nd):
            for w in range(W):
                cond = np.linalg.cond(ref_vgl[d, w, 0])
                tol = max(1e-13, 1e-15 * n * cond)
                assert abs(det[d, w] - rdet[d, w]) <= tol * abs(rdet[d, w]), (name, spin, d, w, det[d, w], rdet[d, w])
                err = np.abs(inv[d, w] - rinv[d, w]).max() / np.abs(rinv[d, w]).max()
                assert err <= tol, (name, spin, d, w, err, cond)
                np.testing.assert_allclose(adj[d, w], det[d, w] * inv[d, w], rtol=1e-15, atol=0)
    q.context_destroy(ctx)


@pytest.mark.parametrize("name,W,extra", [("small", 32, 0), ("medium", 8, 0), ("mol300", 2, 0)])
def test_local_energy(name, W, extra):
    print(f"local energy {name} W={W}", flush=True)
    s, x, W = _system(name, W, extra)
    ctx, ia, ib = _ctx(s, x, W, 1, 1, seed=1)
    Ne, mo = s.elec_num, s.mo_num
    mv = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * mo).reshape(W * Ne, 5, mo)
    va, vb = OD.slater_vgl(mv, ia, W, Ne, 0, s.up), OD.slater_vgl(mv, ib, W, Ne, s.up, s.dn)
    (_, inva, _), (_, invb, _) = OD.det_inverse(va), OD.det_inverse(vb)
    rkin, rdrift = OD.kinetic_drift(va, inva, vb, invb)
    rpot = OD.potential(x, W, Ne, s.nucl_charge, s.nucl_coord)
    kin = _get(ctx, "qmckl_get_kinetic_energy", W)
    pot = _get(ctx, "qmckl_get_potential_energy", W)
    drift = _get(ctx, "qmckl_get_drift_vector", W * Ne * 3).reshape(W, Ne, 3)
    eloc = np.empty(W)
    q.check(ctx, q.lib().qmckl_get_local_energy(ctx, q.ptr(eloc), W))
    conds = max(np.linalg.cond(va[0, w, 0]) for w in range(W))
    tol = max(1e-11, 1e-15 * s.up * conds)
    assert np.abs(kin - rkin).max() <= tol * max(1.0, np.abs(rkin).max()), (np.abs(kin - rkin).max(), conds)
    assert np.abs(drift - rdrift).max() <= tol * max(1.0, np.abs(rdrift).max())
    assert np.abs(pot - rpot).max() <= 1e-12 * np.abs(rpot).max()
    np.testing.assert_allclose(eloc, kin + pot, rtol=1e-15, atol=1e-15 * np.abs(pot).max())
    # the potential's parts through the electron getters
    vee = _get(ctx, "qmckl_get_electron_ee_potential", W)
    ven = _get(ctx, "qmckl_get_electron_en_potential", W)
    vnn = ctypes.c_double()
    q.check(ctx, q.lib().qmckl_get_nucleus_repulsion(ctx, ctypes.byref(vnn)))
    np.testing.assert_allclose(vee + ven + vnn.value, pot, rtol=1e-14)
    q.context_destroy(ctx)


def test_determinant_errors():
    s, x, W = _system("small", 4, 0)
    ctx, ia, ib = _ctx(s, x, W, 2, 1)
    out = np.empty(W)
    L = q.lib()
    assert L.qmckl_get_kinetic_energy(ctx, q.ptr(out)) == q.QMCKL_FAILURE  # two alpha determinants
    bad = np.ascontiguousarray(ia, np.int64).copy()
    bad[0, 0, 0] = s.mo_num + 1
    q.check(ctx, L.qmckl_set_determinant_mo_index_alpha(ctx, q.ptr(bad), bad.size))
    assert L.qmckl_get_det_alpha(ctx, q.ptr(np.empty(2 * W))) == q.QMCKL_FAILURE
    assert "outside" in q.last_error(ctx)
    # other walkers than the indices were set for
    S.set_coords(ctx, s, S.electron_coords(np.random.default_rng(0), s, W + 1), W + 1)
    assert L.qmckl_get_det_beta(ctx, q.ptr(np.empty(W + 1))) == q.QMCKL_FAILURE
    assert L.qmckl_set_determinant_type(ctx, b"X") == 2
    q.context_destroy(ctx)
    # a singular Slater matrix (an MO that is zero everywhere): determinant 0, NaN inverse
    s.mo_coef = np.vstack([s.mo_coef, np.zeros((1, s.ao_num))])
    ctx, ia, ib = _ctx(s, x, W, 1, 1)
    ia2 = np.ascontiguousarray(ia, np.int64).copy()
    ia2[0, 1, 1] = s.mo_num  # the zero MO, 1-based
    q.check(ctx, L.qmckl_set_determinant_mo_index_alpha(ctx, q.ptr(ia2), ia2.size))
    det = _get(ctx, "qmckl_get_det_alpha", W)
    assert det[1] == 0.0 and np.all(det[[0, 2, 3]] != 0.0)
    inv = _get(ctx, "qmckl_get_det_inv_matrix_alpha", W * s.up * s.up).reshape(W, -1)
    assert np.all(np.isnan(inv[1])) and np.all(np.isfinite(inv[0]))
    q.context_destroy(ctx)


# ---- Sherman-Morrison-Woodbury ----------------------------------------------------------
REF_CASES = [
    (1, "naive", "Slater_inv1", 3.407025646103221e-10, -4.120398385068217e-10),
    (2, "wb2", "Slater_inv2", -1.4432116661319376e-11, 2.367058141251457e-10),
    (3, "wb3", "Slater_inv3_1", -1.23743195512859e-09, 1.602708950725074e-10),
    (3, "splitting", "Slater_inv3_2", -1.23743195512859e-09, 1.602708950725074e-10),
]
DIM = 21


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _host_smw(ctx, kind, LDS, U, idx, S_inv, det, breakdown=1e-3):
    L = q.lib()
    Sx = np.ascontiguousarray(S_inv, np.float64).copy()
    d = ctypes.c_double(det)
    U = np.ascontiguousarray(U, np.float64)
    idx = _u64(idx)
    if kind == "naive":
        rc = L.qmckl_sm_naive(ctx, LDS, DIM, len(idx), q.ptr(U), q.ptr(idx), breakdown, q.ptr(Sx), ctypes.byref(d))
    elif kind == "splitting":
        rc = L.qmckl_sm_splitting(ctx, LDS, DIM, len(idx), q.ptr(U), q.ptr(idx), breakdown, q.ptr(Sx),
                                  ctypes.byref(d))
    elif kind == "wb2":
        rc = L.qmckl_woodbury_2x2(ctx, LDS, DIM, q.ptr(U), q.ptr(idx), breakdown, q.ptr(Sx), ctypes.byref(d))
    else:
        rc = L.qmckl_woodbury_3x3(ctx, LDS, DIM, q.ptr(U), q.ptr(idx), breakdown, q.ptr(Sx), ctypes.byref(d))
    return rc, Sx, d.value


@pytest.mark.parametrize("LDS", [24, 22, 21])
def test_smw_reference_cases(LDS):
    z = np.load(os.path.join(GOLDEN, "sm_test.npz"))
    g = lambda n: z[f"lds{LDS}/{n}"]
    ctx = q.context_create()
    for k, kind, inv_name, det0, det1 in REF_CASES:
        rc, Sx, det = _host_smw(ctx, kind, LDS, g(f"Updates{k}"), g(f"Updates_index{k}"), g(inv_name), det0)
        assert rc == 0, (kind, q.last_error(ctx))
        assert abs(det - det1) < 1e-15, (kind, det, det1)
        if kind == "naive":
            ok, ref, _ = OD.sm_naive(LDS, DIM, g(f"Updates{k}"), g(f"Updates_index{k}"), 1e-3, g(inv_name), det0)
        elif kind == "splitting":
            ok, ref, _ = OD.sm_splitting(LDS, DIM, g(f"Updates{k}"), g(f"Updates_index{k}"), 1e-3, g(inv_name), det0)
        else:
            ok, ref, _ = OD.woodbury(2 if kind == "wb2" else 3, LDS, DIM, g(f"Updates{k}"), g(f"Updates_index{k}"),
                                     1e-3, g(inv_name), det0)
        assert ok
        assert np.abs(Sx.reshape(DIM, LDS) - ref).max() <= 1e-12 * np.abs(ref).max(), kind
    # naive on set 3 breaks down at its first update: QMCKL_FAILURE, inputs untouched
    S0 = np.ascontiguousarray(g("Slater_inv3_1"))
    rc, Sx, det = _host_smw(ctx, "naive", LDS, g("Updates3"), g("Updates_index3"), S0, 5.0)
    assert rc == q.QMCKL_FAILURE and np.array_equal(Sx, S0) and det == 5.0
    # the splitting core defers that update's second half
    L = q.lib()
    Sx = S0.copy()
    later_u = np.zeros(3 * LDS)
    later_i = np.zeros(3, np.uint64)
    later = ctypes.c_uint64(0)
    d = ctypes.c_double(1.0)
    U, idx = np.ascontiguousarray(g("Updates3")), _u64(g("Updates_index3"))
    q.check(ctx, L.qmckl_sm_splitting_core(ctx, LDS, DIM, 3, q.ptr(U), q.ptr(idx), 1e-3, q.ptr(Sx), q.ptr(later_u),
                                           q.ptr(later_i), ctypes.byref(later), ctypes.byref(d)))
    rS, rdet, rlu, rli = OD.sm_splitting_core(LDS, DIM, U, idx, 1e-3, S0, 1.0)
    assert later.value == len(rli) == 1 and later_i[0] == rli[0]
    np.testing.assert_array_equal(later_u[:LDS], rlu[0])
    assert np.abs(Sx.reshape(DIM, LDS) - rS).max() <= 1e-12 * np.abs(rS).max()
    assert abs(d.value - rdet) <= 1e-13 * abs(rdet)
    # an index outside 1..Dim is an argument error
    bad = _u64([22])
    assert L.qmckl_sm_naive(ctx, LDS, DIM, 1, q.ptr(U), q.ptr(bad), 1e-3, q.ptr(Sx), None) == 6
    q.context_destroy(ctx)


@pytest.mark.parametrize("dim", [21, 150])
@pytest.mark.parametrize("kind", ["naive", "splitting", "wb2", "wb3"])
def test_smw_batched_device(dim, kind):
    print(f"smw device {kind} dim={dim}", flush=True)
    rng = np.random.default_rng(dim)
    n_mat, lds = 48, dim + 3
    N = {"wb2": 2, "wb3": 3}.get(kind, 4)
    A = rng.normal(size=(n_mat, dim, dim)) / np.sqrt(dim) + 2 * np.eye(dim)  # condition number of a few
    Sinv = np.zeros((n_mat, dim, lds))
    Sinv[:, :, :dim] = np.linalg.inv(A)
    U = np.zeros((n_mat, N, lds))
    U[:, :, :dim] = 0.3 * rng.normal(size=(n_mat, N, dim))
    idx = np.stack([rng.choice(dim, N, replace=False) + 1 for _ in range(n_mat)]).astype(np.uint64)
    # matrix 5: its first update makes the matrix singular (column replaced by the others' sum);
    # the rank-1 and Woodbury kinds report the breakdown.  Splitting halves that update again and
    # again while the rounding of its denominator doubles each pass, so whether it ends in a
    # breakdown or a (meaningless) near-singular inverse is up to rounding: not compared.
    r = int(idx[5, 0]) - 1
    U[5, 0, :dim] = A[5][:, [c for c in range(dim) if c != r][:2]].sum(1) - A[5][:, r]
    det0 = rng.uniform(0.5, 2.0, n_mat)
    dS, dU, dI = q.DeviceBuffer.from_array(Sinv), q.DeviceBuffer.from_array(U), q.DeviceBuffer.from_array(idx)
    dD, dSt = q.DeviceBuffer.from_array(det0), q.DeviceBuffer(4 * n_mat)
    kinds = {"naive": 0, "splitting": 1, "wb2": 3, "wb3": 4}
    ctx = q.context_create()
    q.check(ctx, q.lib().qmckl_amd_sm_update_device(ctx, kinds[kind], n_mat, lds, dim, N, dU.ptr, dI.ptr, 1e-6,
                                                    dS.ptr, dD.ptr, dSt.ptr))
    q.synchronize(ctx)
    got = dS.to_array((n_mat, dim, lds))
    det = dD.to_array((n_mat,))
    st = dSt.to_array((n_mat,), np.int32)
    for m in range(n_mat):
        if kind == "splitting" and m == 5:
            continue
        if kind == "naive":
            ok, ref, rd = OD.sm_naive(lds, dim, U[m], idx[m], 1e-6, Sinv[m], det0[m])
        elif kind == "splitting":
            ok, ref, rd = OD.sm_splitting(lds, dim, U[m], idx[m], 1e-6, Sinv[m], det0[m])
        else:
            ok, ref, rd = OD.woodbury(N, lds, dim, U[m], idx[m], 1e-6, Sinv[m], det0[m])
        assert st[m] == (0 if ok else 1), (m, st[m], ok)
        if not ok:
            assert np.array_equal(got[m], Sinv[m]) and det[m] == det0[m]
            continue
        assert np.abs(got[m] - ref).max() <= 1e-11 * np.abs(ref).max(), (m, np.abs(got[m] - ref).max())
        assert abs(det[m] - rd) <= 1e-11 * abs(rd)
    if kind != "splitting":
        assert st[5] == 1
    q.context_destroy(ctx)
