"""CPU oracle for Slater determinants, the local energy and the Sherman-Morrison-Woodbury
inverse updates -- TEST INFRASTRUCTURE ONLY (tests, smoke() and bench.py's CPU baseline use
it as the checker; the product never does).

numpy restatement of org/under_development/qmckl_determinant.org,
qmckl_local_energy.org and qmckl_sherman_morrison_woodbury.org.  The first two are not
built by the reference and their Fortran does not run as written (mo_vgl without a walker
index, the inverse contracted transposed, determinants summed without coefficients); the
quantities here are the ones the modules define in their text, in C layouts:

  Slater matrix  M[d][w][e][m] = phi_{mo_index[d][w][m] - 1}(r_e), e the electrons of one
                 spin (1-based MO indices, as the reference's test sets them, :1249-1258)
  inverse        M^{-1} of that C-layout matrix; adjugate det(M) M^{-1} (org/qmckl_blas.org:1675)
  kinetic        -1/2 sum_e sum_m lap phi_m(r_e) M^{-1}[m][e]   (Jacobi: lap_e D / D)
  drift          2 sum_m grad phi_m(r_e) M^{-1}[m][e]             ([W][Ne][3])
  potential      V_ee + V_en + V_nn (qmckl_local_energy.org:987-1046)

How each function is pinned (tests/test_oracle_det.py): the SMW updates by the golden
vectors and determinants of the reference's tests (share/qmckl/test_data/sm_test.h,
qmckl_sherman_morrison_woodbury.org:697-755, 2017-2046, 2709-2738, 3158-3188) and by the
exact inverse of the updated Slater matrix; determinant and inverse by LAPACK through numpy
and 2x2 / 3x3 closed forms; kinetic energy and drift by central finite differences of the
determinant of the oracle's own MO values in the electron coordinates.
"""

import numpy as np


def slater_vgl(mo_vgl, mo_index, walk_num, elec_num, e0, n):
    """[det][W][5][n][n]: component c of MO mo_index[d][w][m] - 1 at electron e0 + e
    (qmckl_determinant.org:842-990).  mo_vgl is [W*Ne][5][mo]."""
    mo = np.asarray(mo_vgl).reshape(walk_num, elec_num, 5, -1)[:, e0:e0 + n]  # [W][n][5][mo]
    idx = np.asarray(mo_index).reshape(-1, walk_num, n) - 1
    out = np.empty((idx.shape[0], walk_num, 5, n, n))
    for d in range(idx.shape[0]):
        for w in range(walk_num):
            out[d, w] = mo[w][:, :, idx[d, w]].transpose(1, 0, 2)
    return out


def det_inverse(vgl):
    """(det [d][W], inverse [d][W][n][n], adjugate) of the value planes of slater_vgl
    (qmckl_determinant.org:1761-1900 through qmckl_adjugate)."""
    M = vgl[:, :, 0]
    det = np.linalg.det(M)
    inv = np.linalg.inv(M)
    return det, inv, det[..., None, None] * inv


def kinetic_drift(vgl_a, inv_a, vgl_b, inv_b):
    """Per walker -1/2 sum_e lap_e D/D over both spins, and the drift 2 grad_e D / D as
    [W][Ne][3] (alpha electrons first), for one determinant per spin
    (qmckl_local_energy.org:423-528, 1524-1640)."""
    kin = 0.0
    drift = []
    for vgl, inv in ((vgl_a, inv_a), (vgl_b, inv_b)):
        if vgl.shape[-1] == 0:
            continue
        # sum_m comp[w][e][m] * inv[w][m][e]
        lap = np.einsum("wem,wme->we", vgl[0, :, 4], inv[0])
        kin = kin - 0.5 * lap.sum(1)
        g = np.einsum("wcem,wme->wec", vgl[0, :, 1:4], inv[0])
        drift.append(2.0 * g)
    return kin, np.concatenate(drift, axis=1)


def potential(coord_T, walk_num, elec_num, nucl_charge, nucl_coord):
    """V_ee + V_en + V_nn per walker, with the reference's cut-offs on coincident particles
    (org/qmckl_electron.org:1022-1192, 1477-1658; org/qmckl_nucleus.org:946-1090)."""
    x = np.asarray(coord_T).reshape(3, walk_num, elec_num)
    out = np.zeros(walk_num)
    Z, R = np.asarray(nucl_charge), np.asarray(nucl_coord)
    vnn = 0.0
    for a in range(len(Z)):
        for b in range(a):
            vnn += Z[a] * Z[b] / np.linalg.norm(R[:, a] - R[:, b])
    for w in range(walk_num):
        r = np.linalg.norm(x[:, w, :, None] - x[:, w, None, :], axis=0)
        iu = np.triu_indices(elec_num, 1)
        rr = r[iu]
        vee = (1.0 / rr[np.abs(rr) > 1e-5]).sum()
        ra = np.linalg.norm(x[:, w, :, None] - R[:, None, :], axis=0)  # [Ne][Nn]
        m = np.abs(ra) > 1e-6
        ven = -(np.broadcast_to(Z[None, :], ra.shape)[m] / ra[m]).sum()
        out[w] = vee + ven + vnn
    return out


# ---- Sherman-Morrison-Woodbury (qmckl_sherman_morrison_woodbury.org) --------------------
# Slater_inv is [Dim][LDS] row-major, Updates [N][LDS], Updates_index 1-based: update l adds
# Updates[l][:Dim] to column Updates_index[l] - 1 of the Slater matrix.  Each function returns
# (ok, inverse, determinant); on a breakdown ok is False and the inputs are returned unchanged.

def sm_naive(LDS, Dim, updates, index, breakdown, S_inv, det):
    """qmckl_sm_naive_doc_f (:158-262): rank-1 updates one by one."""
    S = np.array(S_inv, dtype=np.float64).reshape(Dim, LDS).copy()
    U = np.asarray(updates).reshape(-1, LDS)
    for l in range(len(index)):
        C = S[:, :Dim] @ U[l, :Dim]
        row = int(index[l]) - 1
        den = 1.0 + C[row]
        if abs(den) < breakdown:
            return False, np.array(S_inv, dtype=np.float64).reshape(Dim, LDS), det
        det = det * den
        D = S[row, :Dim].copy()
        S[:, :Dim] -= np.outer(C, D) / den
    return True, S, det


def sm_splitting_core(LDS, Dim, updates, index, breakdown, S_inv, det):
    """qmckl_sm_splitting_core_doc_f (:809-918): an update whose denominator is below the
    break-down is halved, half applied now and half returned for later.  Returns
    (inverse, determinant, later_updates [k][LDS], later_index [k])."""
    S = np.array(S_inv, dtype=np.float64).reshape(Dim, LDS).copy()
    U = np.asarray(updates).reshape(-1, LDS)
    later_u, later_i = [], []
    for l in range(len(index)):
        u = U[l, :Dim].copy()
        C = S[:, :Dim] @ u
        row = int(index[l]) - 1
        den = 1.0 + C[row]
        if abs(den) < breakdown:
            later_u.append(U[l].copy() / 2.0)  # the whole row, padding included (:1085-1090)
            later_i.append(int(index[l]))
            C = C / 2.0
            den = 1.0 + C[row]
        det = det * den
        D = S[row, :Dim].copy()
        S[:, :Dim] -= np.outer(C, D) / den
    return S, det, np.array(later_u).reshape(-1, LDS), np.array(later_i, dtype=np.uint64)


def sm_splitting(LDS, Dim, updates, index, breakdown, S_inv, det, depth=64):
    """qmckl_sm_splitting_doc_f (:2787-2858): the core, then the split halves recursively, at
    most `depth` passes in all (the reference recurses without a bound)."""
    S, det, lu, li = sm_splitting_core(LDS, Dim, updates, index, breakdown, S_inv, det)
    for _ in range(depth - 1):
        if len(li) == 0:
            return True, S, det
        S, det, lu, li = sm_splitting_core(LDS, Dim, lu, li, breakdown, S, det)
    if len(li) == 0:
        return True, S, det
    return False, np.array(S_inv, dtype=np.float64).reshape(Dim, LDS), det


def woodbury(k, LDS, Dim, updates, index, breakdown, S_inv, det):
    """qmckl_woodbury_2x2_doc_f (:1458-1587) and _3x3 (:2095-2233): k updates at once,
    S^-1 - C B^-1 V S^-1 with C = S^-1 U and B = 1 + V C (V selects the k updated rows)."""
    S = np.array(S_inv, dtype=np.float64).reshape(Dim, LDS).copy()
    U = np.asarray(updates).reshape(-1, LDS)[:k, :Dim]
    rows = [int(i) - 1 for i in index[:k]]
    C = S[:, :Dim] @ U.T  # [Dim][k]
    B = np.eye(k) + C[rows, :]
    detB = np.linalg.det(B)
    if abs(detB) < breakdown:
        return False, np.array(S_inv, dtype=np.float64).reshape(Dim, LDS), det
    F = np.linalg.inv(B) @ S[rows, :Dim]
    S[:, :Dim] -= C @ F
    return True, S, det * detB
