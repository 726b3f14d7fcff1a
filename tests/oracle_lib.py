"""ctypes binding of the CPU oracle (oracle/liblbm_oracle.so).

The oracle is the checker: tests, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` load it to compute expected results. It is
never part of the product path.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liblbm_oracle.so")

EXT_EQUILIBRIUM_BOUNDARIES = 1 << 0
EXT_VOLUME_FORCE = 1 << 1
EXT_FORCE_FIELD = 1 << 2
EXT_MAGNETO_HYDRO = 1 << 3
EXT_UPDATE_FIELDS = 1 << 4
EXT_CHARGE_INLET = 1 << 5
EXT_SUBGRID_ECR = 1 << 6

FP16S, FP16C, FP32, FP64 = 0, 1, 2, 3
D2Q9, D3Q15, D3Q19, D3Q27 = 0, 1, 2, 3
Q_OF = {D2Q9: 9, D3Q15: 15, D3Q19: 19}

BUF = dict(fi=0, rho=1, u=2, flags=3, F=4, e_stat=5, b_stat=6, e_dyn=7, b_dyn=8,
           fqi=9, q=10, qu_lod=11, ei=12, e_var=13, eti=14, et=15)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("velocity_set", ctypes.c_int32), ("relaxation_time", ctypes.c_int32),
        ("float_type", ctypes.c_int32),
        ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
        ("lod_depth", ctypes.c_uint32), ("ext", ctypes.c_uint32),
        ("w", ctypes.c_float), ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("fz", ctypes.c_float),
        ("ke", ctypes.c_float), ("kmu", ctypes.c_float), ("kkge", ctypes.c_float), ("wq", ctypes.c_float),
        ("kkbme", ctypes.c_float), ("keabs", ctypes.c_float), ("ei_rho0", ctypes.c_float),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        vp, u64, f32, f64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_char_p]
        L.orc_destroy.argtypes = [vp]
        L.orc_buffer.restype = vp
        L.orc_buffer.argtypes = [vp, i32, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_uint32)]
        for name in ["orc_initialize", "orc_do_time_step", "orc_kernel_initialize", "orc_kernel_stream_collide",
                     "orc_kernel_update_fields", "orc_kernel_clear_qu_lod", "orc_kernel_update_e_b_dynamic"]:
            getattr(L, name).argtypes = [vp]
        L.orc_run.argtypes = [vp, u64]
        L.orc_time_step.restype = u64
        L.orc_time_step.argtypes = [vp]
        L.orc_set_time_step.argtypes = [vp, u64]
        L.orc_n_lod.restype = u64
        L.orc_n_lod.argtypes = [vp]
        L.orc_sum_rho.restype = f64
        L.orc_sum_rho.argtypes = [vp]
        L.orc_sum_q.restype = f64
        L.orc_sum_q.argtypes = [vp]
        for name in ["orc_fp16c_encode", "orc_half_encode_rte", "orc_fp16s_store"]:
            getattr(L, name).argtypes = [vp, vp, u64]
        for name in ["orc_fp16c_decode", "orc_half_decode", "orc_fp16s_load"]:
            getattr(L, name).argtypes = [vp, vp, u64]
        L.orc_f_eq.argtypes = [i32, f32, f32, f32, f32, vp]
        L.orc_f_eq_f64.argtypes = [i32, f64, f64, f64, f64, vp]
        L.orc_rho_u.argtypes = [i32, vp, vp]
        L.orc_rho_u_f64.argtypes = [i32, vp, vp]
        L.orc_forcing.argtypes = [i32, f32, f32, f32, f32, f32, f32, vp]
        L.orc_a_eq.argtypes = [f32, f32, f32, f32, vp]
        L.orc_num_threads.restype = i32
        L.orc_set_num_threads.argtypes = [i32]
        u32, u8 = ctypes.c_uint32, ctypes.c_uint8
        L.orc_voxelize.argtypes = [u32, u32, u32, vp, vp, vp, u32, vp, vp, u8, f32, f32, f32]
        L.orc_precompute_b.argtypes = [u32, u32, u32, vp, vp, vp, vp, f32]
        L.orc_precompute_e.argtypes = [u32, u32, u32, vp, vp, vp, f32]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _codec(fn, src, out_dtype):
    src = np.ascontiguousarray(src)
    out = np.empty(src.shape, dtype=out_dtype)
    getattr(lib(), fn)(_ptr(src), _ptr(out), src.size)
    return out


def fp16c_encode(x):
    return _codec("orc_fp16c_encode", np.asarray(x, np.float32), np.uint16)


def fp16c_decode(h):
    return _codec("orc_fp16c_decode", np.asarray(h, np.uint16), np.float32)


def half_encode_rte(x):
    return _codec("orc_half_encode_rte", np.asarray(x, np.float32), np.uint16)


def half_decode(h):
    return _codec("orc_half_decode", np.asarray(h, np.uint16), np.float32)


def fp16s_store(x):
    return _codec("orc_fp16s_store", np.asarray(x, np.float32), np.uint16)


def fp16s_load(h):
    return _codec("orc_fp16s_load", np.asarray(h, np.uint16), np.float32)


def f_eq(vs, rho, ux, uy, uz, f64=False):
    q = Q_OF[vs]
    if f64:
        out = np.zeros(q, np.float64)
        lib().orc_f_eq_f64(vs, rho, ux, uy, uz, _ptr(out))
    else:
        out = np.zeros(q, np.float32)
        lib().orc_f_eq(vs, rho, ux, uy, uz, _ptr(out))
    return out


def rho_u(vs, f, f64=False):
    if f64:
        f = np.ascontiguousarray(f, np.float64)
        out = np.zeros(4, np.float64)
        lib().orc_rho_u_f64(vs, _ptr(f), _ptr(out))
    else:
        f = np.ascontiguousarray(f, np.float32)
        out = np.zeros(4, np.float32)
        lib().orc_rho_u(vs, _ptr(f), _ptr(out))
    return out


def forcing(vs, u, force):
    out = np.zeros(Q_OF[vs], np.float32)
    lib().orc_forcing(vs, *[float(v) for v in u], *[float(v) for v in force], _ptr(out))
    return out


def a_eq(qv, ux, uy, uz):
    out = np.zeros(7, np.float32)
    lib().orc_a_eq(qv, ux, uy, uz, _ptr(out))
    return out


_DT = {(4, False): np.float32, (8, False): np.float64, (2, False): np.uint16, (1, False): np.uint8}


class OracleSim:
    """One oracle simulation (mirrors the reference Lbm of one domain)."""

    def __init__(self, nx, ny, nz, velocity_set=D3Q19, relaxation_time=0, float_type=FP32, lod_depth=4,
                 ext=0, w=1.0, fx=0.0, fy=0.0, fz=0.0, ke=0.0, kmu=0.0, kkge=0.0, wq=0.4, kkbme=0.0, keabs=0.0,
                 ei_rho0=0.0):
        p = OrcParams(velocity_set, relaxation_time, float_type, nx, ny, nz, lod_depth, ext,
                      w, fx, fy, fz, ke, kmu, kkge, wq, kkbme, keabs, ei_rho0)
        err = ctypes.create_string_buffer(256)
        h = lib().orc_create(ctypes.byref(p), err)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        self.params = p
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.float_type = float_type
        self.q = Q_OF[velocity_set]

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buf(self, name):
        """numpy view onto an oracle buffer (writes go to the oracle state)."""
        cnt = ctypes.c_uint64()
        esz = ctypes.c_uint32()
        p = lib().orc_buffer(self.h, BUF[name], ctypes.byref(cnt), ctypes.byref(esz))
        if not p:
            return None
        dt = {4: np.float32, 8: np.float64, 2: np.uint16, 1: np.uint8}[esz.value]
        if name == "flags":
            dt = np.uint8
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                    shape=(cnt.value,))
        return arr

    @property
    def t(self):
        return lib().orc_time_step(self.h)

    @t.setter
    def t(self, v):
        lib().orc_set_time_step(self.h, v)

    @property
    def n_lod(self):
        return lib().orc_n_lod(self.h)

    def initialize(self):
        lib().orc_initialize(self.h)

    def run(self, steps):
        lib().orc_run(self.h, steps)

    def do_time_step(self):
        lib().orc_do_time_step(self.h)

    def kernel(self, name):
        getattr(lib(), "orc_kernel_" + name)(self.h)

    def sum_rho(self):
        return lib().orc_sum_rho(self.h)

    def sum_q(self):
        return lib().orc_sum_q(self.h)

    def ddfs(self, name="fi"):
        """Decoded DDF values as float (shape [q, N])."""
        raw = self.buf(name)
        nq = raw.size // self.N
        raw = raw.reshape(nq, self.N)
        if self.float_type == FP16C:
            return fp16c_decode(raw)
        if self.float_type == FP16S:
            return fp16s_load(raw)
        return raw.copy()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


# mesh voxelisation and static fields (oracle/mesh_oracle.cpp), in place on numpy buffers
MODEL_FLAG = {0: 0x01, 1: 0x11, 2: 0x09, 3: 0x05}  # src/mesh.rs:309-314


def voxelize(shape, flags, b_dyn, tri, p_min, p_max, flag, m=(0.0, 0.0, 0.0)):
    """voxelize_mesh on flags (uint8 [N]) and b_dyn (float32 [3N] or None)."""
    nx, ny, nz = shape
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1)
    lo = np.ascontiguousarray(p_min, np.float32)
    hi = np.ascontiguousarray(p_max, np.float32)
    assert flags.dtype == np.uint8 and flags.flags.c_contiguous and flags.size == nx * ny * nz
    bd = None
    if b_dyn is not None:
        assert b_dyn.dtype == np.float32 and b_dyn.flags.c_contiguous and b_dyn.size == 3 * flags.size
        bd = _ptr(b_dyn)
    lib().orc_voxelize(nx, ny, nz, _ptr(flags), bd, _ptr(tri), tri.size // 9, _ptr(lo), _ptr(hi), flag,
                       float(m[0]), float(m[1]), float(m[2]))


def precompute_b(shape, flags, b_dyn, b_stat, kmu0):
    """psi_from_mesh + static_b_from_mesh; returns psi on the padded lattice."""
    nx, ny, nz = shape
    psi = np.zeros((nx + 2) * (ny + 2) * (nz + 2), np.float32)
    for a in (b_dyn, b_stat):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.size == 3 * flags.size
    lib().orc_precompute_b(nx, ny, nz, _ptr(flags), _ptr(b_dyn), _ptr(psi), _ptr(b_stat), float(kmu0))
    return psi


def precompute_e(shape, flags, b_dyn, e_stat, ke):
    """static_e_from_mesh (charges read from the x block of b_dyn)."""
    nx, ny, nz = shape
    for a in (b_dyn, e_stat):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.size == 3 * flags.size
    lib().orc_precompute_e(nx, ny, nz, _ptr(flags), _ptr(b_dyn), _ptr(e_stat), float(ke))
