"""GPU: single-electron moves of the CHAMP Jastrow (org/qmckl_jastrow_champ_single.org)
through the C-ABI, against the oracle's full Jastrow before and after the move.

The reference's own tests (the N2 system, walkers 2, each electron moved in turn to
(1, 2, 3) and (4, 5, 6)) check every single-move getter against the difference, or the
row, of the full quantities before and after the move, at 1e-12; the same checks run here
with the oracle giving the full quantities, on N2 and on two synthetic systems (orders and
spins off the N2 values, up to 240 electrons).  Acceptance must leave the context as if
the coordinates had been set anew.
"""
import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import system as S
from test_gpu_parity import _n2_ctx, make_ctx

pytestmark = pytest.mark.gpu

ALL = ("factor_ee", "factor_en", "factor_een", "factor_ee_gl", "factor_en_gl", "factor_een_gl", "ee_rescaled",
       "ee_rescaled_gl", "en_rescaled", "en_rescaled_gl", "een_e", "een_e_gl", "een_n", "een_n_gl")


def _moved(x, W, Ne, elec, new):
    """coord_T [3][W*Ne] with electron elec of walker w at new[w]."""
    y = x.reshape(3, W, Ne).copy()
    y[:, :, elec] = new.T
    return np.ascontiguousarray(y.reshape(3, W * Ne))


def _check_move(ctx, s, x, W, elec, new, tol):
    Ne, Nn, C = s.elec_num, s.nucl_num, s.cord
    L = q.lib()
    nc = np.ascontiguousarray(new)  # [W][3]
    q.check(ctx, L.qmckl_set_single_point(ctx, b"N", elec, q.ptr(nc), nc.size))
    g = lambda fn, n: q.get_array(ctx, fn, n)
    old = oracle.jastrow(s, x, W, want=ALL)
    y = _moved(x, W, Ne, elec, new)
    nw = oracle.jastrow(s, y, W, want=ALL)

    def close(got, ref, what):
        err = np.abs(got - ref).max()
        assert err <= tol * max(1.0, np.abs(ref).max()), f"electron {elec} {what}: {err:.3e}"

    for k in ("ee", "en", "een"):
        close(g(f"qmckl_get_jastrow_champ_single_{k}", W), nw[f"factor_{k}"] - old[f"factor_{k}"], k)
    d_ee_gl = g("qmckl_get_jastrow_champ_single_ee_gl", W * Ne * 4).reshape(W, Ne, 4)
    close(d_ee_gl.transpose(0, 2, 1), nw["factor_ee_gl"] - old["factor_ee_gl"], "ee_gl")
    d_en_gl = g("qmckl_get_jastrow_champ_single_en_gl", W * Ne * 4).reshape(W, Ne, 4)
    close(d_en_gl.transpose(0, 2, 1), nw["factor_en_gl"] - old["factor_en_gl"], "en_gl")
    d_een_gl = g("qmckl_get_jastrow_champ_single_een_gl", W * 4 * Ne).reshape(W, 4, Ne)
    close(d_een_gl, nw["factor_een_gl"] - old["factor_een_gl"], "een_gl")
    d_een_g = g("qmckl_get_jastrow_champ_single_een_g", W * 4 * Ne).reshape(W, 4, Ne)
    close(d_een_g[:, :3, elec], (nw["factor_een_gl"] - old["factor_een_gl"])[:, :3, elec], "een_g")
    rest = d_een_g.copy()
    rest[:, :3, elec] = 0.0
    assert np.all(rest == 0.0), "een_g holds the moved electron's gradient alone"
    # intermediates of the new position, against the rows of the new full quantities
    others = np.arange(Ne) != elec
    xw = y.reshape(3, W, Ne)
    ref_d = np.linalg.norm(xw[:, :, :] - xw[:, :, elec][:, :, None], axis=0)  # [W][Ne]
    close(g("qmckl_get_single_electron_ee_distance", W * Ne).reshape(W, Ne)[:, others], ref_d[:, others], "ee distance")
    ref_dn = np.linalg.norm(xw[:, :, elec][:, :, None] - s.nucl_coord[:, None, :], axis=0)  # [W][Nn]
    close(g("qmckl_get_single_electron_en_distance", W * Nn).reshape(W, Nn), ref_dn, "en distance")
    close(g("qmckl_get_ee_rescaled_single", W * Ne).reshape(W, Ne)[:, others], nw["ee_rescaled"][:, elec, others],
          "ee rescaled")
    close(g("qmckl_get_ee_rescaled_single_gl", W * Ne * 4).reshape(W, Ne, 4)[:, others],
          nw["ee_rescaled_gl"][:, elec, others], "ee rescaled gl")
    close(g("qmckl_get_en_rescaled_single", W * Nn).reshape(W, Nn), nw["en_rescaled"][:, :, elec], "en rescaled")
    close(g("qmckl_get_en_rescaled_single_gl", W * Nn * 4).reshape(W, Nn, 4), nw["en_rescaled_gl"][:, :, elec],
          "en rescaled gl")
    close(g("qmckl_get_een_rescaled_single_e", W * (C + 1) * Ne).reshape(W, C + 1, Ne), nw["een_e"][:, :, elec],
          "een rescaled e")
    se_gl = g("qmckl_get_een_rescaled_single_e_gl", W * (C + 1) * 4 * Ne).reshape(W, C + 1, 4, Ne)
    close(se_gl, nw["een_e_gl"][:, :, :, :, elec].transpose(0, 1, 3, 2), "een rescaled e gl")
    close(g("qmckl_get_een_rescaled_single_n", W * (C + 1) * Nn).reshape(W, C + 1, Nn), nw["een_n"][:, :, :, elec],
          "een rescaled n")
    close(g("qmckl_get_een_rescaled_single_n_gl", W * (C + 1) * Nn * 4).reshape(W, C + 1, Nn, 4),
          nw["een_n_gl"][:, :, :, :, elec], "een rescaled n gl")
    return y, nw


def test_single_moves_n2_reference_protocol():
    """Each electron of the two N2 walkers moved to (1, 2, 3) and (4, 5, 6) from the
    original coordinates, the reference's test protocol (1e-12)."""
    ctx, s, x, W = _n2_ctx()
    new = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    for elec in range(s.elec_num):
        S.set_coords(ctx, s, x, W)
        _check_move(ctx, s, x, W, elec, new, 1e-12)
    q.context_destroy(ctx)


@pytest.mark.parametrize("nn,up,dn,cord,spin_indep,W", [(20, 19, 18, 5, 0, 8), (9, 12, 11, 4, 1, 5),
                                                         (40, 120, 120, 5, 0, 2)])
def test_single_moves_synthetic(nn, up, dn, cord, spin_indep, W):
    s, x, W = S.jastrow_workload(nn, up, dn, 5, 5, cord, walk_num=W)
    s.spin_independent = spin_indep
    ctx = make_ctx(s, x, W, mo=False)
    rng = np.random.default_rng(nn)
    Ne = s.elec_num
    for elec in (0, up - 1, up, Ne - 1):
        S.set_coords(ctx, s, x, W)
        new = x.reshape(3, W, Ne)[:, :, elec].T + rng.normal(0.0, 0.3, (W, 3))
        _check_move(ctx, s, x, W, elec, new, 1e-11)
    q.context_destroy(ctx)


@pytest.mark.parametrize("device_coords", [False, True])
def test_accepted_moves_match_a_fresh_evaluation(device_coords):
    """A Metropolis-like sequence of accepted moves: after each accept the context's factors,
    their GLs, e^J and its GL equal the oracle's at the moved coordinates, the coordinate
    getter returns them, and the MO VGL is recomputed there."""
    s, x, W = S.workload("small", walk_num=16)
    ctx = make_ctx(s, x, W)
    L = q.lib()
    Ne = s.elec_num
    if device_coords:
        xd = q.DeviceBuffer.from_array(x)
        q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
    q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
    rng = np.random.default_rng(3)
    y = x
    for step, elec in enumerate((0, 5, 13, 5)):
        new = np.ascontiguousarray(y.reshape(3, W, Ne)[:, :, elec].T + rng.normal(0.0, 0.4, (W, 3)))
        # 'T' layout [3][walk_num] on every other move
        if step % 2:
            nt = np.ascontiguousarray(new.T)
            q.check(ctx, L.qmckl_set_single_point(ctx, b"T", elec, q.ptr(nt), nt.size))
        else:
            q.check(ctx, L.qmckl_set_single_point(ctx, b"N", elec, q.ptr(new), new.size))
        q.check(ctx, L.qmckl_get_jastrow_champ_single_accept(ctx))
        y = _moved(y, W, Ne, elec, new)
        ref = oracle.jastrow(s, y, W)
        for k in ("ee", "en", "een"):
            got = q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}", W)
            assert np.abs(got - ref[f"factor_{k}"]).max() <= 1e-11 * max(1.0, np.abs(ref[f"factor_{k}"]).max()), k
            gl = q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}_gl", W * 4 * Ne).reshape(W, 4, Ne)
            assert np.abs(gl - ref[f"factor_{k}_gl"]).max() <= 1e-11 * max(1.0, np.abs(ref[f"factor_{k}_gl"]).max())
        val = q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
        J = ref["factor_ee"] + ref["factor_en"] + ref["factor_een"]
        live = np.abs(J + 100.0) > 1e-6  # e^J is flushed to 0 below J = -100
        np.testing.assert_allclose(val[live], ref["value"][live], rtol=1e-11, atol=0)
        back = np.empty(3 * W * Ne)
        q.check(ctx, L.qmckl_get_electron_coord(ctx, b"T", q.ptr(back), back.size))
        np.testing.assert_array_equal(back.reshape(3, W * Ne), y)
        mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * s.mo_num_kept).reshape(W * Ne, 5, -1)
        ref_mo = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, y))
        assert np.abs(mo - ref_mo).max() <= 1e-12 * np.abs(ref_mo).max()
    q.context_destroy(ctx)
