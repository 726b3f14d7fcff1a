"""ion_update_e_b_noalloc (the allocation-free field step) against
ion_compute_fields through the C ABI: E_dyn and B_dyn bit-identical on the
same inputs, on calls in a row whose lattices shrink (the static source
array keeps the previous call's longer source list), in both field modes,
through the LOD-tiled kernel (64^3, depth 3) and the generic kernel with a
near field (32^3 and 16^3, depth 2). Device buffers for the new entry point
come from hipMalloc of the HIP runtime the library itself links."""
import ctypes

import numpy as np
import pytest

import ionsolver_amd as I
from harness import make_config
from ionsolver_amd import Field, FieldMode

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _FieldBuffers(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("flags", "q", "u", "qu_lod", "e_stat", "b_stat", "e_dyn", "b_dyn")]


def hip():
    I.load_library()  # the library's own libamdhip64 is then the one dlopen returns
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return h


def inputs(cfg, rng):
    n = cfg.n_x * cfg.n_y * cfg.n_z
    n_lod = cfg.kernel_constants()["n_lod"]
    flags = np.where(rng.uniform(size=n) < 0.1, 1, 0).astype(np.uint8)
    flags[rng.uniform(size=n) < 0.05] = 0x11  # magnet cells are fluid for the field step
    lod = (0.05 * rng.standard_normal((n_lod, 4))).astype(np.float32)
    lod[rng.uniform(size=n_lod) < 0.3, 0] = 0.0  # zero-charge LODs are dropped by the compaction
    return dict(flags=flags, q=(0.05 * rng.standard_normal(n)).astype(np.float32),
                u=(0.02 * rng.standard_normal(3 * n)).astype(np.float32),
                qu_lod=lod.reshape(-1),
                e_stat=(0.05 * rng.standard_normal(3 * n)).astype(np.float32),
                b_stat=(0.05 * rng.standard_normal(3 * n)).astype(np.float32))


def with_handle(cfg, arr, mode):
    lbm = I.Lbm(cfg)
    try:
        for k, f in (("flags", Field.FLAGS), ("q", Field.Q), ("u", Field.U), ("qu_lod", Field.QU_LOD),
                     ("e_stat", Field.E_STAT), ("b_stat", Field.B_STAT)):
            lbm.write_field(f, arr[k])
        lbm.set_field_mode(mode)
        lbm.compute_fields()
        return lbm.read_field(Field.E_DYN), lbm.read_field(Field.B_DYN)
    finally:
        lbm.close()


def no_alloc(h, cfg, arr, mode):
    n = cfg.n_x * cfg.n_y * cfg.n_z
    ptrs = {}
    try:
        for k in ("flags", "q", "u", "qu_lod", "e_stat", "b_stat"):
            p = ctypes.c_void_p()
            assert h.hipMalloc(ctypes.byref(p), arr[k].nbytes) == 0
            ptrs[k] = p
            assert h.hipMemcpy(p, arr[k].ctypes.data, arr[k].nbytes, H2D) == 0
        zeros = np.zeros(3 * n, np.float32)  # as ion_create leaves E_dyn / B_dyn; solid cells are not written
        for k in ("e_dyn", "b_dyn"):
            p = ctypes.c_void_p()
            assert h.hipMalloc(ctypes.byref(p), 12 * n) == 0
            ptrs[k] = p
            assert h.hipMemcpy(p, zeros.ctypes.data, zeros.nbytes, H2D) == 0
        buf = _FieldBuffers(**{k: v.value for k, v in ptrs.items()})
        c = cfg._to_c()
        rc = I.load_library().ion_update_e_b_noalloc(ctypes.byref(c), ctypes.byref(buf), int(mode), None)
        assert rc == 0, I.load_library().ion_last_error().decode()
        assert h.hipDeviceSynchronize() == 0
        out = []
        for k in ("e_dyn", "b_dyn"):
            o = np.empty(3 * n, np.float32)
            assert h.hipMemcpy(o.ctypes.data, ptrs[k], o.nbytes, D2H) == 0
            out.append(o)
        return out
    finally:
        for p in ptrs.values():
            h.hipFree(p)


@pytest.mark.parametrize("mode", [FieldMode.FAST, FieldMode.EXACT])
def test_no_alloc_field_step_matches_compute_fields(mode):
    h = hip()
    rng = np.random.default_rng(31)
    # depth 5 first (up to 8^5 LOD sources, the static array's capacity; at
    # 32^3 its near-field block is the whole lattice, as in the reference),
    # then smaller calls over what the one before left in the array
    for n, D in ((32, 5), (64, 3), (32, 2), (16, 2)):
        cfg = make_config(nx=n, ny=n, nz=n, mhd=True, lod_depth=D, float_type=2)
        arr = inputs(cfg, rng)
        print(f"no-alloc field step: {n}^3, depth {D}, mode {mode.name}", flush=True)
        e_ref, b_ref = with_handle(cfg, arr, mode)
        e_new, b_new = no_alloc(h, cfg, arr, mode)
        assert np.isfinite(e_ref).all() and np.abs(e_ref - arr["e_stat"]).max() > 0
        np.testing.assert_array_equal(e_new, e_ref)
        np.testing.assert_array_equal(b_new, b_ref)

