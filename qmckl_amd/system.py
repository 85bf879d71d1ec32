"""Wavefunction descriptions and the synthetic workloads.

``System`` holds everything the qmckl_context setters take (nuclei, electrons, AO basis,
MO coefficients, CHAMP Jastrow parameters, numerical precision) and ``load`` pushes it
through the C-ABI in the order the reference's own tests use.  ``workload(name)`` builds
the seeded synthetic systems the parity tests and bench.py run on:

  small      2 N (Z=7) 2.07 bohr apart, 14 e-, per atom 3s+2p+1d Cartesian shells of 6
             primitives (30 AOs), 7 occupied MOs, Jastrow orders 5/5/5, 256 walkers
  medium     50 nuclei (20 C, 30 H) in a 20-bohr cube (min separation 1.8 bohr), 150 e-,
             C 4s3p2d1f / H 3s2p1d (1150 AOs), 75 occupied MOs, 4096 walkers, p = 41
  large      200 nuclei (80 C, 100 H, 20 O) in a 40-bohr cube, 1000 e-, same shells
             (5000 AOs), 500 occupied MOs, 1024 walkers, p = 35
  batched    64 molecules of the small recipe scaled to 10..400 electrons, 256 walkers each

Exponents are log-uniform in [0.1, 1e4] (6 per shell, sorted descending), contraction
coefficients uniform in (0, 1], primitives normalised through ``prim_factor`` and each
contraction through ``shell_factor``.  MO coefficients are orthonormal (QR of a Gaussian
matrix with sign fix).  Electron i of every walker is drawn from N(R_(i mod Nn), 1 bohr^2).
"""

from dataclasses import dataclass, field
import math

import numpy as np

from . import check, lib, ptr

L_CHAR = "spdfghik"


@dataclass
class System:
    nucl_charge: np.ndarray
    nucl_coord: np.ndarray  # [3][Nn]
    up: int
    dn: int
    shell_ang_mom: np.ndarray  # int32 [shell]
    shell_prim_num: np.ndarray  # int64 [shell]
    shell_prim_index: np.ndarray  # int64 [shell]
    nucleus_index: np.ndarray  # int64 [Nn]
    nucleus_shell_num: np.ndarray  # int64 [Nn]
    shell_factor: np.ndarray
    exponent: np.ndarray
    coefficient: np.ndarray
    prim_factor: np.ndarray
    ao_factor: np.ndarray
    mo_coef: np.ndarray  # [mo][ao]
    type_nucl_vector: np.ndarray  # int64 [Nn]
    a_vector: np.ndarray  # [type][aord+1]
    b_vector: np.ndarray  # [bord+1]
    c_vector: np.ndarray  # [type][dim_c]
    rescale_en: np.ndarray  # [type]
    rescale_ee: float = 0.6
    aord: int = 5
    bord: int = 5
    cord: int = 5
    spin_independent: int = 0
    precision: int = 53
    keep_mo: np.ndarray = None  # optional select_mo mask
    r_power: np.ndarray = None  # optional int32 [shell] (qmckl_set_ao_basis_r_power); default 0
    basis_type: str = "G"  # 'G' Gaussian or 'S' Slater shells (qmckl_set_ao_basis_type)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nucl_num(self):
        return int(self.nucl_charge.size)

    @property
    def elec_num(self):
        return self.up + self.dn

    @property
    def shell_num(self):
        return int(self.shell_ang_mom.size)

    @property
    def prim_num(self):
        return int(self.exponent.size)

    @property
    def ao_num(self):
        return int(self.ao_factor.size)

    @property
    def mo_num(self):
        return int(self.mo_coef.shape[0])

    @property
    def mo_num_kept(self):
        return self.mo_num if self.keep_mo is None else int(np.count_nonzero(self.keep_mo))

    @property
    def type_nucl_num(self):
        return int(self.rescale_en.size)


def dim_c_vector(cord):
    n = 0
    for p in range(2, cord + 1):
        for k in range(p - 1, -1, -1):
            lmax = p - k if k != 0 else p - k - 2
            for l in range(lmax, -1, -1):
                if (p - k - l) % 2 == 0:
                    n += 1
    return n


def load(ctx, s: System, jastrow=True, mo=True):
    """Pushes `s` into the context through the C-ABI setters (raises on error)."""
    L = lib()

    def c(rc, fn):
        check(ctx, rc, fn)

    def f64(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def i64(a):
        return np.ascontiguousarray(a, dtype=np.int64)

    def i32(a):
        return np.ascontiguousarray(a, dtype=np.int32)

    keep = []  # keep numpy temporaries alive until the calls return
    Nn = s.nucl_num
    c(L.qmckl_set_numprec_precision(ctx, int(s.precision)), "set_numprec_precision")
    c(L.qmckl_set_nucleus_num(ctx, Nn), "set_nucleus_num")
    a = f64(s.nucl_coord); keep.append(a)
    c(L.qmckl_set_nucleus_coord(ctx, b"T", ptr(a), a.size), "set_nucleus_coord")
    a = f64(s.nucl_charge); keep.append(a)
    c(L.qmckl_set_nucleus_charge(ctx, ptr(a), a.size), "set_nucleus_charge")
    c(L.qmckl_set_electron_num(ctx, s.up, s.dn), "set_electron_num")
    c(L.qmckl_set_ao_basis_type(ctx, s.basis_type.encode()), "set_ao_basis_type")
    c(L.qmckl_set_ao_basis_shell_num(ctx, s.shell_num), "set_ao_basis_shell_num")
    c(L.qmckl_set_ao_basis_prim_num(ctx, s.prim_num), "set_ao_basis_prim_num")
    for name, arr, conv in [
        ("nucleus_index", s.nucleus_index, i64), ("nucleus_shell_num", s.nucleus_shell_num, i64),
        ("shell_ang_mom", s.shell_ang_mom, i32), ("shell_prim_num", s.shell_prim_num, i64),
        ("shell_prim_index", s.shell_prim_index, i64), ("shell_factor", s.shell_factor, f64),
        ("exponent", s.exponent, f64), ("coefficient", s.coefficient, f64),
        ("prim_factor", s.prim_factor, f64),
    ]:
        a = conv(arr); keep.append(a)
        c(getattr(L, "qmckl_set_ao_basis_" + name)(ctx, ptr(a), a.size), "set_ao_basis_" + name)
    if s.r_power is not None:
        a = i32(s.r_power); keep.append(a)
        c(L.qmckl_set_ao_basis_r_power(ctx, ptr(a), a.size), "set_ao_basis_r_power")
    c(L.qmckl_set_ao_basis_ao_num(ctx, s.ao_num), "set_ao_basis_ao_num")
    a = f64(s.ao_factor); keep.append(a)
    c(L.qmckl_set_ao_basis_ao_factor(ctx, ptr(a), a.size), "set_ao_basis_ao_factor")
    if mo:
        c(L.qmckl_set_mo_basis_mo_num(ctx, s.mo_num), "set_mo_basis_mo_num")
        a = f64(s.mo_coef); keep.append(a)
        c(L.qmckl_set_mo_basis_coefficient(ctx, ptr(a), a.size), "set_mo_basis_coefficient")
        if s.keep_mo is not None:
            a = i32(s.keep_mo); keep.append(a)
            c(L.qmckl_mo_basis_select_mo(ctx, ptr(a), a.size), "mo_basis_select_mo")
    if jastrow:
        c(L.qmckl_set_jastrow_champ_spin_independent(ctx, int(s.spin_independent)), "spin_independent")
        c(L.qmckl_set_jastrow_champ_aord_num(ctx, s.aord), "aord_num")
        c(L.qmckl_set_jastrow_champ_bord_num(ctx, s.bord), "bord_num")
        c(L.qmckl_set_jastrow_champ_cord_num(ctx, s.cord), "cord_num")
        c(L.qmckl_set_jastrow_champ_type_nucl_num(ctx, s.type_nucl_num), "type_nucl_num")
        a = i64(s.type_nucl_vector); keep.append(a)
        c(L.qmckl_set_jastrow_champ_type_nucl_vector(ctx, ptr(a), a.size), "type_nucl_vector")
        a = f64(s.a_vector); keep.append(a)
        c(L.qmckl_set_jastrow_champ_a_vector(ctx, ptr(a), a.size), "a_vector")
        a = f64(s.b_vector); keep.append(a)
        c(L.qmckl_set_jastrow_champ_b_vector(ctx, ptr(a), a.size), "b_vector")
        if s.cord > 1:
            a = f64(s.c_vector); keep.append(a)
            c(L.qmckl_set_jastrow_champ_c_vector(ctx, ptr(a), a.size), "c_vector")
        a = f64(s.rescale_en); keep.append(a)
        c(L.qmckl_set_jastrow_champ_rescale_factor_en(ctx, ptr(a), a.size), "rescale_factor_en")
        c(L.qmckl_set_jastrow_champ_rescale_factor_ee(ctx, float(s.rescale_ee)), "rescale_factor_ee")


def from_context(ctx, name=""):
    """The System held by a context whose nucleus, electron, AO-basis and MO-basis groups are set
    (for instance by qmckl_trexio_read), read back through the getters.  The Jastrow fields are
    zero parameters of orders 5/5/5 over one nucleus type."""
    import ctypes

    from . import get_array, get_scalar

    i64 = lambda fn, n: get_array(ctx, "qmckl_get_" + fn, n, np.int64)
    f64 = lambda fn, n: get_array(ctx, "qmckl_get_" + fn, n)
    Nn = get_scalar(ctx, "qmckl_get_nucleus_num")
    shell = get_scalar(ctx, "qmckl_get_ao_basis_shell_num")
    prim = get_scalar(ctx, "qmckl_get_ao_basis_prim_num")
    ao = get_scalar(ctx, "qmckl_get_ao_basis_ao_num")
    mo = get_scalar(ctx, "qmckl_get_mo_basis_mo_num")
    coord = np.empty(3 * Nn)
    check(ctx, lib().qmckl_get_nucleus_coord(ctx, b"T", ptr(coord), coord.size), "get_nucleus_coord")
    return System(
        nucl_charge=f64("nucleus_charge", Nn), nucl_coord=coord.reshape(3, Nn),
        up=get_scalar(ctx, "qmckl_get_electron_up_num"), dn=get_scalar(ctx, "qmckl_get_electron_down_num"),
        shell_ang_mom=get_array(ctx, "qmckl_get_ao_basis_shell_ang_mom", shell, np.int32),
        shell_prim_num=i64("ao_basis_shell_prim_num", shell), shell_prim_index=i64("ao_basis_shell_prim_index", shell),
        nucleus_index=i64("ao_basis_nucleus_index", Nn), nucleus_shell_num=i64("ao_basis_nucleus_shell_num", Nn),
        shell_factor=f64("ao_basis_shell_factor", shell), exponent=f64("ao_basis_exponent", prim),
        coefficient=f64("ao_basis_coefficient", prim), prim_factor=f64("ao_basis_prim_factor", prim),
        ao_factor=f64("ao_basis_ao_factor", ao), mo_coef=f64("mo_basis_coefficient", mo * ao).reshape(mo, ao),
        r_power=get_array(ctx, "qmckl_get_ao_basis_r_power", shell, np.int32),
        basis_type=get_scalar(ctx, "qmckl_get_ao_basis_type", ctypes.c_char).decode(),
        type_nucl_vector=np.zeros(Nn, np.int64), a_vector=np.zeros((1, 6)), b_vector=np.zeros(6),
        c_vector=np.zeros((1, dim_c_vector(5))), rescale_en=np.ones(1), name=name)


def set_coords(ctx, s: System, coord_T, walk_num):
    """qmckl_set_electron_coord with coordinates given as [3][walk_num*elec_num]."""
    a = np.ascontiguousarray(coord_T, dtype=np.float64)
    check(ctx, lib().qmckl_set_electron_coord(ctx, b"T", walk_num, ptr(a), a.size), "set_electron_coord")


# ---------------------------------------------------------------------------------
# synthetic workloads
# ---------------------------------------------------------------------------------

def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def _prim_norm(alpha, l):
    """Normalisation of x^l exp(-alpha r^2)."""
    return (2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (l / 2.0) / math.sqrt(_dfact(2 * l - 1))


def _shell_norm(alphas, coefs, pf, l):
    """1/sqrt of the self-overlap of the contracted x^l component."""
    s = 0.0
    for ai, ci, ni in zip(alphas, coefs, pf):
        for aj, cj, nj in zip(alphas, coefs, pf):
            p = ai + aj
            s += ci * cj * ni * nj * (math.pi / p) ** 1.5 * _dfact(2 * l - 1) / (2.0 * p) ** l
    return 1.0 / math.sqrt(s)


def _basis(rng, elements, shells_of, nprim=6):
    ang, pnum, pidx, nidx, nsh, sf, ex, co, pfac = [], [], [], [], [], [], [], [], []
    ao = 0
    for el in elements:
        nidx.append(len(ang))
        recipe = shells_of[el]
        nsh.append(sum(recipe))
        for l, count in enumerate(recipe):
            for _ in range(count):
                al = np.sort(10.0 ** rng.uniform(-1.0, 4.0, nprim))[::-1]
                cf = rng.uniform(0.0, 1.0, nprim)
                cf = np.where(cf == 0.0, 1.0, cf)
                pf = np.array([_prim_norm(a, l) for a in al])
                ang.append(l)
                pnum.append(nprim)
                pidx.append(len(ex))
                ex.extend(al.tolist())
                co.extend(cf.tolist())
                pfac.extend(pf.tolist())
                sf.append(_shell_norm(al, cf, pf, l))
                ao += (l + 1) * (l + 2) // 2
    return dict(
        shell_ang_mom=np.array(ang, np.int32), shell_prim_num=np.array(pnum, np.int64),
        shell_prim_index=np.array(pidx, np.int64), nucleus_index=np.array(nidx, np.int64),
        nucleus_shell_num=np.array(nsh, np.int64), shell_factor=np.array(sf),
        exponent=np.array(ex), coefficient=np.array(co), prim_factor=np.array(pfac),
        ao_factor=np.ones(ao),
    )


def _orthonormal_rows(rng, n_keep, n_ao):
    g = rng.standard_normal((n_ao, n_keep))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))[None, :]
    return np.ascontiguousarray(q.T)  # [mo][ao]


def _place_nuclei(rng, n, box, min_sep):
    pts = []
    while len(pts) < n:
        p = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    return np.array(pts).T.copy()  # [3][n]


def _jastrow(rng, types, n_types, aord=5, bord=5, cord=5):
    dim_c = dim_c_vector(cord)
    return dict(
        type_nucl_vector=np.array(types, np.int64),
        a_vector=rng.uniform(-0.1, 0.1, (n_types, aord + 1)),
        b_vector=rng.uniform(-0.1, 0.1, bord + 1),
        c_vector=rng.uniform(-0.1, 0.1, (n_types, max(dim_c, 1))),
        rescale_en=np.ones(n_types),
        rescale_ee=0.6, aord=aord, bord=bord, cord=cord, spin_independent=0,
    )


def electron_coords(rng, s: System, walk_num):
    """[3][walk_num*elec_num]: electron i of each walker ~ N(R_(i mod Nn), 1)."""
    Ne, Nn = s.elec_num, s.nucl_num
    centers = s.nucl_coord[:, np.arange(Ne) % Nn]  # [3][Ne]
    x = centers[:, None, :] + rng.standard_normal((3, walk_num, Ne))
    return np.ascontiguousarray(x.reshape(3, walk_num * Ne))


CONFIG_SEED = 20261203


def workload(name, walk_num=None, seed=None):
    """Returns (System, coord_T[3][W*Ne], walk_num) for a named synthetic workload."""
    cid = {"small": 1, "medium": 2, "large": 3}[name]
    rng = np.random.Generator(np.random.PCG64(CONFIG_SEED + cid if seed is None else seed))
    if name == "small":
        coord = np.array([[0.0, 0.0], [0.0, 0.0], [0.0, 2.07]])
        charge = np.array([7.0, 7.0])
        elements = ["N", "N"]
        shells = {"N": (3, 2, 1)}
        up = dn = 7
        W = 256
        prec = 53
        types = [0, 0]
        ntypes = 1
    elif name == "medium":
        coord = _place_nuclei(rng, 50, 20.0, 1.8)
        elements = ["C"] * 20 + ["H"] * 30
        charge = np.array([6.0] * 20 + [1.0] * 30)
        shells = {"C": (4, 3, 2, 1), "H": (3, 2, 1)}
        up = dn = 75
        W = 4096
        prec = 41
        types = [0] * 20 + [1] * 30
        ntypes = 2
    else:
        coord = _place_nuclei(rng, 200, 40.0, 1.8)
        elements = ["C"] * 80 + ["H"] * 100 + ["O"] * 20
        charge = np.array([6.0] * 80 + [1.0] * 100 + [8.0] * 20)
        shells = {"C": (4, 3, 2, 1), "H": (3, 2, 1), "O": (4, 3, 2, 1)}
        up = dn = 500
        W = 1024
        prec = 35
        types = [0] * 80 + [1] * 100 + [2] * 20
        ntypes = 3
    basis = _basis(rng, elements, shells)
    nao = basis["ao_factor"].size
    nocc = max(up, dn)
    if name == "small":
        mo = _orthonormal_rows(rng, nao, nao)
        keep = np.zeros(nao, np.int32)
        keep[:nocc] = 1
    else:
        mo = _orthonormal_rows(rng, nocc, nao)
        keep = None
    s = System(nucl_charge=charge, nucl_coord=coord, up=up, dn=dn, mo_coef=mo, keep_mo=keep,
               precision=prec, name=name, **basis, **_jastrow(rng, types, ntypes))
    W = W if walk_num is None else walk_num
    x = electron_coords(rng, s, W)
    return s, x, W


def jastrow_workload(nucl_num, up, dn, aord=5, bord=5, cord=5, walk_num=16, seed=7):
    """A synthetic system of C and H nuclei (alternating, two Jastrow types) in a cube of
    side 2 Nn^(1/3) bohr, with Jastrow orders aord/bord/cord and an H-type basis on every
    nucleus; returns (System, coord_T[3][W*Ne], walk_num).  Used to vary the shapes the
    Jastrow kernels see (electron and nucleus counts off the tile sizes, every order)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    coord = _place_nuclei(rng, nucl_num, 2.0 * nucl_num ** (1.0 / 3.0), 1.4)
    elements = ["C" if a % 2 == 0 else "H" for a in range(nucl_num)]
    charge = np.array([6.0 if e == "C" else 1.0 for e in elements])
    types = [0 if e == "C" else 1 for e in elements]
    basis = _basis(rng, elements, {"C": (3, 2, 1), "H": (3, 2, 1)})
    nao = basis["ao_factor"].size
    mo = _orthonormal_rows(rng, max(up, dn), nao)
    s = System(nucl_charge=charge, nucl_coord=coord, up=up, dn=dn, mo_coef=mo, precision=53,
               name=f"jastrow-{nucl_num}n-{up + dn}e-c{cord}", **basis,
               **_jastrow(rng, types, 2, aord, bord, cord))
    x = electron_coords(rng, s, walk_num)
    return s, x, walk_num


def molecule_workload(elec_num, walk_num=256, seed=11):
    """A molecule of the batched workload: ceil(Ne/7) nitrogen-like nuclei (Z = 7) with
    4s3p2d1f cartesian shells (35 AOs per nucleus, 5 AOs per electron) at random positions
    1.8 bohr apart or more in a cube of side 2.2 (number of nuclei)^(1/3) bohr; Ne electrons
    (up-spin first, up = ceil(Ne/2)); max(up, dn) occupied MOs as random orthonormal rows;
    Jastrow orders 5/5/5; screening precision 41.  Returns (System, coord_T, walk_num)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nn = max(1, -(-elec_num // 7))
    coord = _place_nuclei(rng, nn, 2.2 * nn ** (1.0 / 3.0) + 1.0, 1.8)
    elements = ["N"] * nn
    basis = _basis(rng, elements, {"N": (4, 3, 2, 1)})
    up = -(-elec_num // 2)
    dn = elec_num - up
    nao = basis["ao_factor"].size
    mo = _orthonormal_rows(rng, max(up, dn), nao)
    s = System(nucl_charge=np.full(nn, 7.0), nucl_coord=coord, up=up, dn=dn, mo_coef=mo, precision=41,
               name=f"molecule-{elec_num}e", **basis, **_jastrow(rng, [0] * nn, 1))
    x = electron_coords(rng, s, walk_num)
    return s, x, walk_num


def slater_workload(walk_num=4, seed=17):
    """A system of Slater ('S') shells: 3 nuclei (Z = 6, 1, 8) in a 6-bohr box; per nucleus an
    s shell with r_power 0, an s shell with r_power 1, a p shell with r_power 0 and a d shell
    with r_power 1 (cartesian, 16 AOs per nucleus), 2 primitives each with exponents in
    [0.5, 3] and the normalisation (2 alpha)^(n + 3/2) / sqrt((2n + 2)!) of r^n e^(-alpha r) as
    primitive factors; 12 electrons (6 up), all 48 MOs as random orthonormal rows; Jastrow
    orders 5/5/5; precision 53.  Returns (System, coord_T, walk_num)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nn = 3
    coord = _place_nuclei(rng, nn, 6.0, 1.5)
    shells = [(0, 0), (0, 1), (1, 0), (2, 1)]  # (l, r_power)
    ang, rpow, pnum, pidx, nidx, nsn, expo, coef, pf = [], [], [], [], [], [], [], [], []
    for a in range(nn):
        nidx.append(len(ang))
        nsn.append(len(shells))
        for l, n in shells:
            ang.append(l)
            rpow.append(n)
            pidx.append(len(expo))
            pnum.append(2)
            for al in sorted(rng.uniform(0.5, 3.0, 2), reverse=True):
                expo.append(al)
                coef.append(rng.uniform(0.2, 1.0))
                pf.append((2.0 * al) ** (n + 1.5) / math.sqrt(math.factorial(2 * n + 2)))
    nao = sum((l + 1) * (l + 2) // 2 for l in ang)
    basis = dict(shell_ang_mom=np.array(ang, np.int32), shell_prim_num=np.array(pnum, np.int64),
                 shell_prim_index=np.array(pidx, np.int64), nucleus_index=np.array(nidx, np.int64),
                 nucleus_shell_num=np.array(nsn, np.int64), shell_factor=np.ones(len(ang)),
                 exponent=np.array(expo), coefficient=np.array(coef), prim_factor=np.array(pf),
                 ao_factor=rng.uniform(0.8, 1.2, nao))
    up, dn = 6, 6
    mo = _orthonormal_rows(rng, nao, nao)
    types = [0, 1, 2]
    s = System(nucl_charge=np.array([6.0, 1.0, 8.0]), nucl_coord=coord, up=up, dn=dn, mo_coef=mo, precision=53,
               r_power=np.array(rpow, np.int32), basis_type="S", name="slater-3n-12e", **basis,
               **_jastrow(rng, types, 3))
    x = electron_coords(rng, s, walk_num)
    return s, x, walk_num


def batched_workload(n_mol=64, walk_num=256, seed=20261207):
    """The batched workload: n_mol molecules with electron counts uniform in [10, 400]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = rng.integers(10, 401, n_mol)
    return [molecule_workload(int(ne), walk_num, seed=int(seed + 1 + i)) for i, ne in enumerate(counts)]
