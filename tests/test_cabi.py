"""C-ABI checks that need no GPU: the library loads, exports every entry point of
include/qmckl.h, and the host-side API behaves like the reference's (return codes of
org/qmckl_error.org:152-181, set-once rules, NOT_PROVIDED before the data is complete,
INVALID_ARG_n on short arrays) -- the misuse tests of org/qmckl_context.org:733-760 and
org/qmckl_ao.org:2457-2675.  No call here reaches a kernel.
"""
import ctypes
import subprocess

import numpy as np
import pytest

import qmckl_amd as q
import systems
from qmckl_amd import system as S


def test_every_declared_symbol_is_exported():
    declared = set(q.parse_header())
    out = subprocess.run(["nm", "-D", "--defined-only", q.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = sorted(declared - exported)
    assert not missing, missing
    assert len(declared) > 140


def test_context_lifecycle():
    L = q.lib()
    ctx = L.qmckl_context_create()
    assert ctx != 0 and L.qmckl_context_check(ctx) == ctx
    c2 = L.qmckl_context_copy(ctx)
    assert c2 not in (0, ctx)
    assert L.qmckl_context_destroy(c2) == 0
    assert L.qmckl_context_destroy(ctx) == 0
    assert L.qmckl_context_check(ctx) == 0
    assert L.qmckl_context_destroy(ctx) == q.QMCKL_INVALID_CONTEXT
    assert L.qmckl_set_nucleus_num(0, 2) == q.QMCKL_INVALID_CONTEXT


def test_error_strings_and_last_error():
    L = q.lib()
    assert L.qmckl_string_of_error(0) == b"Success"
    assert L.qmckl_string_of_error(3) == b"Invalid argument 3"
    assert L.qmckl_string_of_error(106) == b"Not provided"
    ctx = L.qmckl_context_create()
    rc = L.qmckl_failwith(ctx, 101, b"somewhere", b"what happened")
    assert rc == 101
    code = ctypes.c_int32()
    fn = ctypes.create_string_buffer(256)
    msg = ctypes.create_string_buffer(1024)
    assert L.qmckl_get_error(ctx, ctypes.byref(code), fn, msg) == 0
    assert (code.value, fn.value, msg.value) == (101, b"somewhere", b"what happened")
    assert "Failure" in q.last_error(ctx)
    L.qmckl_context_destroy(ctx)


def test_numprec():
    L = q.lib()
    ctx = L.qmckl_context_create()
    assert L.qmckl_get_numprec_precision(ctx) == 53
    assert L.qmckl_set_numprec_precision(ctx, 1) == 2
    assert L.qmckl_set_numprec_precision(ctx, 54) == 2
    assert L.qmckl_set_numprec_precision(ctx, 41) == 0
    L.qmckl_get_numprec_epsilon.restype = ctypes.c_double
    assert L.qmckl_get_numprec_epsilon(ctx) == 2.0 ** -40
    assert L.qmckl_test_precision_64(1.0, 1.0) == 53
    assert L.qmckl_test_precision_64(1.0, np.nextafter(1.0, 2.0)) == 52
    L.qmckl_context_destroy(ctx)


def test_set_once_and_not_provided_rules():
    L = q.lib()
    s, x, W = systems.chbrclf()
    ctx = L.qmckl_context_create()
    out = np.zeros(10)
    assert L.qmckl_get_ao_basis_ao_vgl(ctx, q.ptr(out), 10) == q.QMCKL_NOT_PROVIDED
    S.load(ctx, s, jastrow=False)
    assert L.qmckl_ao_basis_provided(ctx) and L.qmckl_mo_basis_provided(ctx)
    idx = np.ascontiguousarray(s.nucleus_index)
    assert L.qmckl_set_ao_basis_nucleus_index(ctx, q.ptr(idx), idx.size) == q.QMCKL_ALREADY_SET
    assert L.qmckl_set_nucleus_num(ctx, 3) == q.QMCKL_ALREADY_SET
    assert L.qmckl_set_electron_num(ctx, 1, 1) == q.QMCKL_ALREADY_SET
    S.set_coords(ctx, s, x, W)
    assert L.qmckl_get_ao_basis_ao_vgl(ctx, q.ptr(out), 10) == 3  # INVALID_ARG_3: array too small
    assert L.qmckl_get_mo_basis_mo_vgl(ctx, q.ptr(out), 10) == 3
    assert L.qmckl_get_jastrow_champ_value(ctx, q.ptr(out), 10) == q.QMCKL_NOT_PROVIDED
    n = ctypes.c_int64()
    assert L.qmckl_get_electron_walk_num(ctx, ctypes.byref(n)) == 0 and n.value == W
    assert L.qmckl_get_mo_basis_mo_num(ctx, ctypes.byref(n)) == 0 and n.value == 224
    back = np.zeros(3 * W * s.elec_num)
    assert L.qmckl_get_electron_coord(ctx, b"N", q.ptr(back), back.size) == 0
    np.testing.assert_array_equal(back.reshape(-1, 3).T, x)
    keep = np.zeros(224, np.int32)
    keep[:34] = 1
    assert L.qmckl_mo_basis_select_mo(ctx, q.ptr(keep), keep.size) == 0
    assert L.qmckl_get_mo_basis_mo_num(ctx, ctypes.byref(n)) == 0 and n.value == 34
    L.qmckl_context_destroy(ctx)


def test_jastrow_setter_rules():
    L = q.lib()
    s, x, W = systems.n2()
    ctx = L.qmckl_context_create()
    assert L.qmckl_set_nucleus_num(ctx, 2) == 0
    nc = np.ascontiguousarray(s.nucl_coord)
    assert L.qmckl_set_nucleus_coord(ctx, b"T", q.ptr(nc), nc.size) == 0
    ch = np.ascontiguousarray(s.nucl_charge)
    assert L.qmckl_set_nucleus_charge(ctx, q.ptr(ch), ch.size) == 0
    assert L.qmckl_set_electron_num(ctx, 5, 5) == 0
    a = np.ascontiguousarray(s.a_vector)
    assert L.qmckl_set_jastrow_champ_a_vector(ctx, q.ptr(a), a.size) == q.QMCKL_NOT_PROVIDED
    assert L.qmckl_set_jastrow_champ_aord_num(ctx, -1) == 2
    assert L.qmckl_set_jastrow_champ_aord_num(ctx, 5) == 0
    assert L.qmckl_set_jastrow_champ_aord_num(ctx, 5) == q.QMCKL_ALREADY_SET
    assert L.qmckl_set_jastrow_champ_type_nucl_num(ctx, 1) == 0
    assert L.qmckl_set_jastrow_champ_a_vector(ctx, q.ptr(a), 3) == 3
    assert L.qmckl_set_jastrow_champ_a_vector(ctx, q.ptr(a), a.size) == 0
    bad = np.array([0, 1], np.int64)
    assert L.qmckl_set_jastrow_champ_type_nucl_vector(ctx, q.ptr(bad), 2) == 2
    assert L.qmckl_set_jastrow_champ_rescale_factor_ee(ctx, 0.0) == 2
    assert L.qmckl_set_jastrow_champ_cord_num(ctx, 5) == 0
    d = ctypes.c_int64()
    assert L.qmckl_get_jastrow_champ_dim_c_vector(ctx, ctypes.byref(d)) == 0 and d.value == 23
    assert not L.qmckl_jastrow_champ_provided(ctx)
    L.qmckl_context_destroy(ctx)


def test_multi_gpu_request_is_refused():
    r = subprocess.run(["python", "bench.py", "--gpus", "2"], capture_output=True, text=True,
                       cwd=q.HERE + "/..")
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_c99_program_against_the_header(tmp_path):
    """A C99 caller (tests/c_abi_host.c) compiles against include/qmckl.h with
    `gcc -std=c99 -pedantic -Werror`, links libqmckl_amd.so and passes its host-side checks."""
    import os

    here = os.path.dirname(os.path.abspath(__file__))
    libdir = os.path.dirname(q.LIB_PATH)
    exe = str(tmp_path / "c_abi_host")
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(here, "..", "include"), os.path.join(here, "c_abi_host.c"),
                    "-o", exe, "-L", libdir, "-lqmckl_amd", f"-Wl,-rpath,{libdir}", "-lm"],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def _mo_tile(ctx):
    import ctypes

    g, w, p = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    q.check(ctx, q.lib().qmckl_amd_get_mo_tile(ctx, ctypes.byref(g), ctypes.byref(w), ctypes.byref(p)))
    return g.value, w.value, p.value


def test_tile_setter_refuses_unbuilt_shapes():
    """Host-side rules of qmckl_amd_set_mo_tile / qmckl_amd_get_mo_tile (no GPU call)."""
    from qmckl_amd import system as S

    s, x, W = S.workload("small", walk_num=1)
    ctx = q.context_create()
    try:
        S.load(ctx, s)
        L = q.lib()
        assert L.qmckl_amd_set_mo_tile(ctx, 17, 8) == 2
        assert L.qmckl_amd_set_mo_tile(ctx, 9, 4) == 3
        assert L.qmckl_amd_set_mo_tile(ctx, 4, 8) == 3
        assert L.qmckl_amd_set_mo_tile(ctx, -1, 4) == 2
        assert L.qmckl_amd_set_mo_tile(ctx, 6, 8) == 0
        assert _mo_tile(ctx) == (1, 4, 1)  # 7 MOs: one column tile, waves adjusted to the cap
    finally:
        q.context_destroy(ctx)


def test_mo_precision_follows_the_numerical_precision_until_set():
    """Unset, the MO product precision follows the numerical precision: FP64 above 26 bits,
    3xBF16 at 26 or fewer (where the reference's optimised build switches the product to
    single precision); once set it stays."""
    import ctypes

    s, x, W = S.workload("small", walk_num=1)
    ctx = q.context_create()
    try:
        S.load(ctx, s)
        L = q.lib()
        m = ctypes.c_int32()
        q.check(ctx, L.qmckl_amd_get_mo_precision(ctx, ctypes.byref(m)))
        assert m.value == q.MO_FP64
        q.check(ctx, L.qmckl_set_numprec_precision(ctx, 26))
        q.check(ctx, L.qmckl_amd_get_mo_precision(ctx, ctypes.byref(m)))
        assert m.value == q.MO_BF16X3
        q.check(ctx, L.qmckl_set_numprec_precision(ctx, 27))
        q.check(ctx, L.qmckl_amd_get_mo_precision(ctx, ctypes.byref(m)))
        assert m.value == q.MO_FP64
        q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, q.MO_FP32))
        q.check(ctx, L.qmckl_set_numprec_precision(ctx, 20))
        q.check(ctx, L.qmckl_amd_get_mo_precision(ctx, ctypes.byref(m)))
        assert m.value == q.MO_FP32
    finally:
        q.context_destroy(ctx)
