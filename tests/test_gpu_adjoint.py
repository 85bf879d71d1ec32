"""GPU: the adjoint of L = sum_w [wJ ln Psi_J + wL E_L] (E_L = -1/2 sum_i (lap_i J +
|grad_i J|^2) + V_ee + V_en + V_nn), qmckl_amd_get_jastrow_champ_adjoint, on the 14-electron
system (2 nuclei, Jastrow orders 5/5/5):

* L against the oracle's J and GL with numpy potentials (1e-11 relative);
* every gradient component (6 + 6 + 23 parameters, 6 nuclear coordinates) against central
  finite differences (h = 1e-5) of that oracle L, for wJ = wL = 1 and for each of the two
  parts alone;
* with wL = 0 (the gradient of sum_w J), against the walker sums of the reference-layout
  derivative getters (factor_{ee,en,een}_pderiv, forces_jastrow_{en,een}), apart from the
  three components where the reference departs from the exact derivative (see
  tests/test_oracle_deriv.py);
* through torch.autograd: JastrowSurrogate's gradients equal the C-ABI's.
"""
import copy

import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import autograd as AG
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu
H = 1e-5


def oracle_L(s, x, W, wJ, wL):
    """L from the oracle's Jastrow (factors and factor GLs) and numpy potentials."""
    r = oracle.jastrow(s, x, W, want=("factor_ee", "factor_en", "factor_een", "factor_ee_gl", "factor_en_gl",
                                      "factor_een_gl"))
    J = r["factor_ee"] + r["factor_en"] + r["factor_een"]
    g = r["factor_ee_gl"] + r["factor_en_gl"] + r["factor_een_gl"]  # [W][4][Ne]
    kin = -0.5 * (g[:, 3] + (g[:, :3] ** 2).sum(1)).sum(1)
    Ne = s.elec_num
    xw = x.reshape(3, W, Ne)
    vee = np.zeros(W)
    ven = np.zeros(W)
    for w in range(W):
        rr = np.linalg.norm(xw[:, w, :, None] - xw[:, w, None, :], axis=0)
        vee[w] = (1.0 / rr[np.triu_indices(Ne, 1)]).sum()
        R = np.linalg.norm(xw[:, w, :, None] - s.nucl_coord[:, None, :], axis=0)
        ven[w] = -(s.nucl_charge[None, :] / R).sum()
    Z, Rn = s.nucl_charge, s.nucl_coord
    vnn = sum(Z[p] * Z[q] / np.linalg.norm(Rn[:, p] - Rn[:, q]) for p in range(s.nucl_num) for q in range(p))
    return float((wJ * J + wL * (kin + vee + ven + vnn)).sum())


def _fd(s, x, W, wJ, wL, attr, idx):
    sp, sm = copy.deepcopy(s), copy.deepcopy(s)
    getattr(sp, attr)[idx] += H
    getattr(sm, attr)[idx] -= H
    return (oracle_L(sp, x, W, wJ, wL) - oracle_L(sm, x, W, wJ, wL)) / (2 * H)


@pytest.fixture(scope="module")
def case():
    s, x, W = S.workload("small", walk_num=32)
    return s, x, W


@pytest.mark.parametrize("wJ,wL", [(1.0, 1.0), (1.0, 0.0), (0.0, 1.0)])
def test_adjoint_matches_finite_differences(case, wJ, wL):
    s, x, W = case
    val, ga, gb, gc, gr = AG.surrogate(s, x, W, wJ, wL)
    ref = oracle_L(s, x, W, wJ, wL)
    assert abs(val - ref) <= 1e-11 * max(1.0, abs(ref)), (val, ref)
    got, fd = [], []
    for t in range(s.type_nucl_num):
        for p in range(s.aord + 1):
            got.append(ga[t, p]), fd.append(_fd(s, x, W, wJ, wL, "a_vector", (t, p)))
    for p in range(s.bord + 1):
        got.append(gb[p]), fd.append(_fd(s, x, W, wJ, wL, "b_vector", (p,)))
    for t in range(s.type_nucl_num):
        for n in range(gc.shape[1]):
            got.append(gc[t, n]), fd.append(_fd(s, x, W, wJ, wL, "c_vector", (t, n)))
    for a in range(s.nucl_num):
        for d in range(3):
            got.append(gr[a, d]), fd.append(_fd(s, x, W, wJ, wL, "nucl_coord", (d, a)))
    got, fd = np.array(got), np.array(fd)
    scale = max(1.0, np.abs(fd).max())
    err = np.abs(got - fd) / scale
    print(f"wJ={wJ} wL={wL}: L = {val:.6e}, gradient scale {scale:.3e}, max |adjoint - FD| / scale {err.max():.2e}",
          flush=True)
    assert err.max() <= 2e-7, (np.argmax(err), got[np.argmax(err)], fd[np.argmax(err)])


def test_adjoint_of_J_matches_derivative_getters(case):
    """wL = 0: d(sum_w J)/dtheta against the walker sums of the reference-layout getters."""
    s, x, W = case
    _, ga, gb, gc, gr = AG.surrogate(s, x, W, 1.0, 0.0)
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    T, Nn = s.type_nucl_num, s.nucl_num
    dc = gc.shape[1]
    ee = q.get_array(ctx, "qmckl_get_jastrow_champ_factor_ee_pderiv", s.bord + 1) * W  # walker mean
    en = q.get_array(ctx, "qmckl_get_jastrow_champ_factor_en_pderiv", T * (s.aord + 1)).reshape(T, -1) * W
    een = q.get_array(ctx, "qmckl_get_jastrow_champ_factor_een_pderiv", T * dc).reshape(T, dc)
    fen = q.get_array(ctx, "qmckl_get_forces_jastrow_en", W * Nn * 3).reshape(W, Nn, 3).sum(0)
    feen = q.get_array(ctx, "qmckl_get_forces_jastrow_een", W * Nn * 3).reshape(W, Nn, 3).sum(0)
    q.context_destroy(ctx)

    def close(a, b, what):
        assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(b).max()), (what, a, b)

    close(gb[1:], ee[1:], "b (B0 departs in the reference)")
    close(np.delete(ga, 1, axis=1), np.delete(en, 1, axis=1), "a (A1 departs in the reference)")
    close(gc, een, "c")
    close(gr, fen + feen, "forces")


def test_autograd_function_matches_the_c_abi(case):
    # the library (and its HIP runtime) is loaded before torch, whose bundled runtime has the
    # same soname: torch's CPU tensors only are used here
    q.lib()
    import torch

    s, x, W = case
    fn = AG.JastrowSurrogate()
    a = torch.tensor(s.a_vector, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(s.b_vector, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(s.c_vector, dtype=torch.float64, requires_grad=True)
    R = torch.tensor(s.nucl_coord, dtype=torch.float64, requires_grad=True)
    Lt = fn.apply(a, b, c, R, s, x, W, 1.0, 1.0)
    (2.0 * Lt).backward()
    val, ga, gb, gc, gr = AG.surrogate(s, x, W, 1.0, 1.0)
    assert float(Lt.detach()) == val
    np.testing.assert_array_equal(a.grad.numpy(), 2.0 * ga)
    np.testing.assert_array_equal(b.grad.numpy(), 2.0 * gb)
    np.testing.assert_array_equal(c.grad.numpy(), 2.0 * gc)
    np.testing.assert_array_equal(R.grad.numpy(), 2.0 * gr.T)
    # a small step along the gradient changes L by |grad|^2 step to first order
    eps = 1e-10
    with torch.no_grad():
        L2 = fn.apply(a + eps * a.grad, b + eps * b.grad, c + eps * c.grad, R + eps * R.grad, s, x, W, 1.0, 1.0)
    g2 = sum(float((t.grad ** 2).sum()) for t in (a, b, c, R)) / 2.0
    assert abs((float(L2) - val) - eps * g2) <= 1e-3 * eps * g2
