"""CPU tests of the C-ABI library: it loads, exports every entry point the
header declares, and rejects bad parameters before touching the GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mglc_lbm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_binding_list():
    from mglc_amd import lbm
    assert declared_functions() == sorted(lbm.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol(hip_lib):
    for name in declared_functions():
        assert hasattr(hip_lib, name), name


def test_create_rejects_bad_tau_and_grid(hip_lib):
    from mglc_amd import lbm
    ctx = ctypes.c_void_p()
    p = lbm.CavityConfig(8, 8, 8, tauf=0.5).params()
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 4  # LBM_ERR_TAU
    assert b"tau" in hip_lib.lbm_last_error()
    p = lbm.CavityConfig(1, 8, 8).params()
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 1  # LBM_ERR_ARG
    p = lbm.CavityConfig(8, 8, 8, block_x=48).params()
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 1
    p = lbm.ThermalConfig(8, 8, 8).params()
    p.collision = 1  # BGK: the thermal path's flow collision is MRT
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 7  # LBM_ERR_UNSUPPORTED
    p = lbm.ThermalConfig(8, 8, 8).params()
    p.u_lid = 0.1
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 1
    p = lbm.CavityConfig(8, 8, 8, halo_faces=64).params()
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 1
    assert not ctx.value
    assert hip_lib.lbm_step(None, 1) == 1
    assert hip_lib.lbm_status_string(3) == b"residual is NaN"


def test_struct_layout_matches_header():
    """ctypes mirror of lbm_params / lbm_info has the C sizes (ints then doubles)."""
    from mglc_amd import lbm
    assert ctypes.sizeof(lbm.LbmParams) == 16 * 4 + 11 * 8
    assert lbm.LbmParams.tau.offset == 64 and lbm.LbmParams.t_ref.offset == 64 + 8 * 8
    assert ctypes.sizeof(lbm.LbmInfo) == 7 * 8 + 8 + 2 * 4


def _driver():
    from mglc_amd import build as B
    exe = os.path.join(B.BIN_DIR, "mglc_cavity")
    if not os.path.exists(exe):
        B.build_fortran_driver()
    return exe


def test_fortran_driver_abi_matches():
    """The Fortran bind(C) mirrors of lbm_params, lbm_particle_params and
    lbm_info have the C structs' sizes."""
    import subprocess
    from mglc_amd import lbm
    out = subprocess.run([_driver(), "--abi"], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(lbm.LbmParams)
    assert int(out[1]) == ctypes.sizeof(lbm.LbmParticleParams)
    assert int(out[2]) == ctypes.sizeof(lbm.LbmInfo)


def test_fortran_module_binds_every_single_context_entry_point():
    """mglc_lbm_capi.f90 has a bind(C) interface for every entry point of a
    single context (the batch API and the bandwidth probe are host-language
    conveniences; lbm_status_string is not needed by the drivers)."""
    text = open(os.path.join(ROOT, "mglc_amd", "fortran", "mglc_lbm_capi.f90")).read()
    bound = set(re.findall(r"bind\(C, name='(lbm_\w+)'\)", text))
    want = {f for f in declared_functions()
            if not f.startswith("lbm_batch_") and f not in ("lbm_bandwidth_probe", "lbm_status_string")}
    assert want <= bound, sorted(want - bound)


def test_fortran_driver_binds_only_declared_symbols():
    import subprocess
    out = subprocess.run(["nm", "-D", "--undefined-only", _driver()], capture_output=True, text=True,
                         check=True).stdout
    used = sorted({l.split()[-1] for l in out.splitlines() if " lbm_" in l})
    assert used and set(used) <= set(declared_functions())


def test_bandwidth_probe_rejects_bad_arguments(hip_lib):
    """lbm_bandwidth_probe validates before any device call."""
    g = ctypes.c_double()
    assert hip_lib.lbm_bandwidth_probe(3, 0, 64, 1, ctypes.byref(g)) == 1
    assert hip_lib.lbm_bandwidth_probe(2, 1, 64, 1, ctypes.byref(g)) == 1
    assert hip_lib.lbm_bandwidth_probe(0, 3, 64, 1, ctypes.byref(g)) == 1
    assert hip_lib.lbm_bandwidth_probe(1, 0, 8, 1, ctypes.byref(g)) == 1
    assert hip_lib.lbm_bandwidth_probe(1, 0, 64, 0, ctypes.byref(g)) == 1
    assert hip_lib.lbm_bandwidth_probe(1, 0, 64, 1, None) == 1
    assert b"lbm_bandwidth_probe" in hip_lib.lbm_last_error()


def test_batch_create_rejects_more_tiles_than_int_range(hip_lib):
    """300 instances of 1024 x 1024 x 512 cells are 2.5e9 tiles of 64 cells:
    refused while the tile table is built, before any device allocation."""
    from mglc_amd import lbm
    one = lbm.CavityConfig(1024, 1024, 512, storage="f16").params()
    arr = (lbm.LbmParams * 300)(*([one] * 300))
    b = ctypes.c_void_p()
    assert hip_lib.lbm_batch_create(arr, 300, ctypes.byref(b)) == 1
    assert b"tiles" in hip_lib.lbm_last_error()
    assert not b.value


def test_create_validates_the_split_step(hip_lib):
    """halo_overlap: y/z halo faces only, isothermal, with halo faces, and a
    looped-back split needs both faces of each halo axis."""
    from mglc_amd import lbm
    ctx = ctypes.c_void_p()
    cases = [
        (dict(halo_faces=0x3C, halo_overlap=2), 1),
        (dict(halo_faces=0x01, halo_overlap=1), 7),    # x face
        (dict(halo_faces=0x00, halo_overlap=1), 1),    # nothing to overlap
        (dict(halo_faces=0x04, halo_overlap=1, halo_loopback=1), 1),  # y- without y+
        (dict(halo_faces=0x10, halo_overlap=1, halo_loopback=1), 1),  # z- without z+
    ]
    for kw, status in cases:
        p = lbm.CavityConfig(16, 16, 16, **kw).params()
        assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == status, kw
        assert not ctx.value
    p = lbm.ThermalConfig(8, 8, 8).params()
    p.halo_overlap = 1
    assert hip_lib.lbm_create(ctypes.byref(p), ctypes.byref(ctx)) == 7
    assert hip_lib.lbm_step_begin(None) == 1 and hip_lib.lbm_step_end(None) == 1


def test_batch_create_checks_the_fp16_scale_and_density(hip_lib):
    """f16_scale_log2 within [0, 20] (2^k must stay a finite float scale), and
    the same scale and rho0 in every instance: the batch stores all instances
    with instance 0's.  Refused before any device allocation."""
    from mglc_amd import lbm
    b = ctypes.c_void_p()
    bad = lbm.CavityConfig(16, 16, 16, storage="f16", f16_scale_log2=130).params()
    assert hip_lib.lbm_batch_create((lbm.LbmParams * 1)(bad), 1, ctypes.byref(b)) == 1
    assert b"f16_scale_log2" in hip_lib.lbm_last_error() and not b.value
    p0 = lbm.CavityConfig(16, 16, 16, storage="f16", f16_scale_log2=12).params()
    p1 = lbm.CavityConfig(16, 16, 16, storage="f16", f16_scale_log2=10).params()
    assert hip_lib.lbm_batch_create((lbm.LbmParams * 2)(p0, p1), 2, ctypes.byref(b)) == 1
    assert b"f16_scale_log2" in hip_lib.lbm_last_error() and not b.value
    p2 = lbm.CavityConfig(16, 16, 16, storage="f16", f16_scale_log2=12, rho0=1.5).params()
    assert hip_lib.lbm_batch_create((lbm.LbmParams * 2)(p0, p2), 2, ctypes.byref(b)) == 1
    assert b"rho0" in hip_lib.lbm_last_error() and not b.value


def test_check_sums_noalloc_rejects_bad_arguments(hip_lib):
    """lbm_check_sums_noalloc validates its pointers and cell count before any
    device call."""
    p = ctypes.c_void_p(4096)
    assert hip_lib.lbm_check_sums_noalloc(None, p, p, p, p, 8, p, None) == 1
    assert hip_lib.lbm_check_sums_noalloc(p, p, p, p, p, 8, None, None) == 1
    assert hip_lib.lbm_check_sums_noalloc(p, p, p, p, p, 0, p, None) == 1
    assert hip_lib.lbm_check_sums_noalloc(p, p, p, p, p, (1 << 40) + 1, p, None) == 1
    assert b"outside [1, 2^40]" in hip_lib.lbm_last_error()


def test_batch_cells_per_thread_rejects_bad_arguments(hip_lib):
    assert hip_lib.lbm_batch_set_cells_per_thread(None, 2) == 1
