"""CPU oracle for the orbital Hessians, the MO cusp correction and the forces on the orbitals --
TEST INFRASTRUCTURE ONLY (tests and smoke() use it as the checker; the product never does).

  shell_hessian, ao_hessian   the C oracle (qmckl_oracle.c), org/qmckl_ao.org:5257-5585,
                              9397-9718: [point][3][4][shell] and [3][point][4][ao]
  mo_hessian                  org/qmckl_mo.org:2362-2776: [3][point][3][mo]
  cusp_params                 qmckl_set_mo_basis_r_cusp, org/qmckl_mo.org:479-700:
                              cusp_param [nucl][4][mo]
  mo_vgl_cusp                 qmckl_compute_mo_basis_mo_vgl_cusp_doc, org/qmckl_mo.org:3070-3212
  forces_ao_value             org/qmckl_forces.org:5122-5466: [nucl][point][3][ao]
  forces_mo_value             :5467-5910: [nucl][point][3][mo]
  forces_mo_g                 :5911-6490: [nucl][3][point][3][mo]
  forces_mo_l                 :6491-6878: [nucl][3][point][mo]

Pinned in tests/test_oracle_orb.py: the Hessians by central finite differences of the C
oracle's AO and shell VGL in the electron coordinates (and by the reference's own protocol,
finite differences in the nuclear coordinates); the forces by finite differences of the MO
VGL in the nuclear coordinates; the cusp fit by the conditions it imposes (value, gradient
and Laplacian continuous at r_cusp, the cusp condition d phi / dr = -Z phi at the nucleus).
"""
import numpy as np

from . import _basis_args, _c, _p, kept_mo_coef, lib, ao_vgl


def _bind():
    L = lib()
    P, I, I32 = __import__("ctypes").c_void_p, __import__("ctypes").c_int64, __import__("ctypes").c_int32
    if not getattr(L, "_orb_bound", False):
        L.oracle_shell_hessian.argtypes = [I, I, I] + [P] * 11 + [I32, I, P, P]
        L.oracle_ao_hessian.argtypes = [I, I, I, I] + [P] * 12 + [I32, I, P, P]
        L._orb_bound = True
    return L


def shell_hessian(s, coord_T):
    x = _c(coord_T)
    n = x.shape[1]
    out = np.empty((n, 3, 4, s.shell_num))
    args = _basis_args(s)
    _bind().oracle_shell_hessian(s.nucl_num, s.shell_num, s.prim_num, *[_p(a) for a in args], s.precision, n,
                                 _p(x), _p(out))
    return out


def ao_hessian(s, coord_T):
    x = _c(coord_T)
    n = x.shape[1]
    out = np.empty((3, n, 4, s.ao_num))
    args = _basis_args(s)
    af = _c(s.ao_factor)
    _bind().oracle_ao_hessian(s.nucl_num, s.shell_num, s.prim_num, s.ao_num, *[_p(a) for a in args], _p(af),
                              s.precision, n, _p(x), _p(out))
    return out


def ao_centres(s):
    """(nucleus, angular momentum) of every AO, shells in nucleus order."""
    nucl, ang = [], []
    for a in range(s.nucl_num):
        for sh in range(s.nucleus_index[a], s.nucleus_index[a] + s.nucleus_shell_num[a]):
            l = int(s.shell_ang_mom[sh])
            nucl += [a] * ((l + 1) * (l + 2) // 2)
            ang += [l] * ((l + 1) * (l + 2) // 2)
    return np.array(nucl), np.array(ang)


def mo_hessian(s, coord_T, r_cusp=None, cusp_param=None):
    """[3][point][3][mo]; with r_cusp, the cusp-corrected Hessian (:2595-2700)."""
    C = kept_mo_coef(s)
    H = ao_hessian(s, coord_T)[:, :, :3, :]  # [3][P][3][ao]
    if r_cusp is None:
        return np.einsum("npmk,ik->npmi", H, C)
    nucl, ang = ao_centres(s)
    x = _c(coord_T)
    out = np.zeros((3, x.shape[1], 3, C.shape[0]))
    for p in range(x.shape[1]):
        r_en = np.linalg.norm(x[:, p, None] - s.nucl_coord, axis=0)
        keep = ~((r_en[nucl] <= r_cusp[nucl]) & (ang == 0))
        out[:, p] = np.einsum("nmk,ik->nmi", H[:, p][:, :, keep], C[:, keep])
        for a in range(s.nucl_num):
            r = r_en[a]
            if r < r_cusp[a]:
                rv = x[:, p] - s.nucl_coord[:, a]
                c = cusp_param[a]  # [4][mo]
                # d_n d_m of c0 + c1 r + c2 r^2 + c3 r^3
                for n in range(3):
                    for m in range(3):
                        d = 1.0 if n == m else 0.0
                        out[n, p, m] += (c[1] * (d / r - rv[n] * rv[m] / r ** 3) + c[2] * 2 * d
                                         + c[3] * 3 * (rv[n] * rv[m] / r + d * r))
    return out


def mo_vgl_cusp(s, coord_T, r_cusp, cusp_param, ao=None):
    """qmckl_compute_mo_basis_mo_vgl_cusp_doc: the s AOs of a nucleus closer than its r_cusp are
    left out and c0 + c1 r + c2 r^2 + c3 r^3 of that nucleus added.  [point][5][mo]."""
    C = kept_mo_coef(s)
    x = _c(coord_T)
    ao = ao_vgl(s, x) if ao is None else ao
    nucl, ang = ao_centres(s)
    out = np.zeros((x.shape[1], 5, C.shape[0]))
    for p in range(x.shape[1]):
        r_en = np.linalg.norm(x[:, p, None] - s.nucl_coord, axis=0)
        keep = (r_en[nucl] > r_cusp[nucl]) | (ang > 0)
        out[p] = ao[p][:, keep] @ C[:, keep].T
        for a in range(s.nucl_num):
            r = r_en[a]
            if r < r_cusp[a]:
                rv = x[:, p] - s.nucl_coord[:, a]
                c = cusp_param[a]
                out[p, 0] += c[0] + r * (c[1] + r * (c[2] + r * c[3]))
                # at the nucleus itself (r = 0) the gradient and Laplacian are undefined (NaN),
                # as in the reference; the value is what the fit uses there
                with np.errstate(divide="ignore", invalid="ignore"):
                    c1 = c[1] / r + 2.0 * c[2] + 3.0 * r * c[3]
                    for d in range(3):
                        out[p, 1 + d] += rv[d] * c1
                    out[p, 4] += 2.0 * c[1] / r + 6.0 * c[2] + 12.0 * c[3] * r
    return out


def cusp_params(s, r_cusp):
    """cusp_param [nucl][4][mo] of qmckl_set_mo_basis_r_cusp (org/qmckl_mo.org:610-667): eta is
    the MO value at the nucleus without the s AOs inside their nucleus' r_cusp; phi, grad, lap
    the value, d/dz and Laplacian of the nucleus' own s AOs at R + r_cusp z."""
    C = kept_mo_coef(s)
    mo = C.shape[0]
    zero = np.zeros((s.nucl_num, 4, mo))
    eta = mo_vgl_cusp(s, s.nucl_coord, r_cusp, zero)[:, 0, :]  # [nucl][mo]
    pts = s.nucl_coord.copy()
    pts[2] += r_cusp
    ao = ao_vgl(s, pts)  # [nucl][5][ao]
    nucl, ang = ao_centres(s)
    out = np.zeros((s.nucl_num, 4, mo))
    for a in range(s.nucl_num):
        Z = s.nucl_charge[a]
        if Z < 0.1:
            continue
        R = r_cusp[a]
        sel = (nucl == a) & (ang == 0)
        phi = C[:, sel] @ ao[a, 0, sel]
        grad = C[:, sel] @ ao[a, 3, sel]
        lap = C[:, sel] @ ao[a, 4, sel]
        e = eta[a]
        out[a, 0] = -(R * (2.0 * e * Z - 6.0 * grad) + lap * R * R + 6.0 * phi) / (2.0 * R * Z - 6.0)
        out[a, 1] = (lap * R * R * Z - 6.0 * grad * R * Z + 6.0 * phi * Z + 6.0 * e * Z) / (2.0 * R * Z - 6.0)
        out[a, 2] = -(R * (-5.0 * grad * Z - 1.5 * lap) + lap * R * R * Z + 3.0 * phi * Z + 3.0 * e * Z
                      + 6.0 * grad) / (R * R * Z - 3.0 * R)
        out[a, 3] = (R * (-2.0 * grad * Z - lap) + 0.5 * lap * R * R * Z + phi * Z + e * Z
                     + 3.0 * grad) / (R * R * R * Z - 3.0 * R * R)
    return out


def forces_ao_value(s, coord_T, ao=None):
    """[nucl][point][3][ao]: -grad AO_k for the AOs on the nucleus, 0 for the others."""
    x = _c(coord_T)
    ao = ao_vgl(s, x) if ao is None else ao
    nucl, _ = ao_centres(s)
    out = np.zeros((s.nucl_num, x.shape[1], 3, s.ao_num))
    for a in range(s.nucl_num):
        sel = nucl == a
        out[a][:, :, sel] = -ao[:, 1:4, :][:, :, sel]
    return out


def forces_mo_value(s, coord_T, ao=None):
    """[nucl][point][3][mo] = -sum_{k on nucleus} C_ik grad AO_k."""
    C = kept_mo_coef(s)
    fa = forces_ao_value(s, coord_T, ao)
    return np.einsum("apdk,ik->apdi", fa, C)


def forces_mo_g(s, coord_T, H=None):
    """[nucl][3][point][3][mo] = -sum_{k on nucleus} C_ik d_n d_m AO_k."""
    C = kept_mo_coef(s)
    H = ao_hessian(s, coord_T) if H is None else H
    nucl, _ = ao_centres(s)
    out = np.zeros((s.nucl_num, 3, H.shape[1], 3, C.shape[0]))
    for a in range(s.nucl_num):
        sel = nucl == a
        out[a] = -np.einsum("npmk,ik->npmi", H[:, :, :3, sel], C[:, sel])
    return out


def forces_mo_l(s, coord_T, H=None):
    """[nucl][3][point][mo] = -sum_{k on nucleus} C_ik d_n lap AO_k."""
    C = kept_mo_coef(s)
    H = ao_hessian(s, coord_T) if H is None else H
    nucl, _ = ao_centres(s)
    out = np.zeros((s.nucl_num, 3, H.shape[1], C.shape[0]))
    for a in range(s.nucl_num):
        sel = nucl == a
        out[a] = -np.einsum("npk,ik->npi", H[:, :, 3, sel], C[:, sel])
    return out
