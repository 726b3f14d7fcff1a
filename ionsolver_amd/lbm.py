"""Python mirror of the reference's ``Lbm`` / ``LbmConfig`` API over the C ABI
of ``libionsolver_hip.so`` (include/ionsolver.h).

Names, defaults and argument meanings follow the reference's Rust host:
``LbmConfig`` (src/lbm/mod.rs:45-135), ``Units`` (src/lbm/units.rs),
``Lbm::{new, initialize, run, do_time_step, get_time_step}``
(src/lbm/mod.rs:166-272, 492) and ``file::{read, write}`` (src/file.rs:12-30).
The reference's ``bwrite!``/``bread!`` on the public domain buffers become
``Lbm.write_field`` / ``Lbm.read_field``. Errors raise ``IonError`` carrying the
library's message; there is no CPU fallback: without the built library every
call fails.
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libionsolver_hip.so")


class IonError(RuntimeError):
    pass


class VelocitySet(enum.IntEnum):  # src/lbm/types.rs:16-26
    D2Q9 = 0
    D3Q15 = 1
    D3Q19 = 2
    D3Q27 = 3

    def get_set_values(self):  # types.rs:37-46 (dimensions, velocities, transfers)
        return {0: (2, 9, 3), 1: (3, 15, 5), 2: (3, 19, 5), 3: (3, 27, 9)}[int(self)]


class RelaxationTime(enum.IntEnum):  # types.rs:55-61
    Srt = 0
    Trt = 1


class FloatType(enum.IntEnum):  # types.rs:75-83
    FP16S = 0
    FP16C = 1
    FP32 = 2
    FP64 = 3   # build-side FP64 instantiation (error reference of the storage formats), D3Q19 only

    def size_of(self):
        return {FloatType.FP32: 4, FloatType.FP64: 8}.get(self, 2)


class Propellant(enum.IntEnum):  # units.rs:214-222
    H = 0
    He = 1
    Ne = 2
    Ar = 3
    Kr = 4
    Xe = 5


class Field(enum.IntEnum):
    FI = 0
    RHO = 1
    U = 2
    FLAGS = 3
    F = 4
    E_STAT = 5
    B_STAT = 6
    E_DYN = 7
    B_DYN = 8
    FQI = 9
    Q = 10
    QU_LOD = 11
    EI = 12
    FIELD_SUMS = 13   # raw direct source sums (E then B, 6 x N), once a surrogate exists
    E_VAR = 14        # ext_subgrid_ecr: oscillating electric field (3 x N)
    ETI = 15          # ext_subgrid_ecr: electron temperature DDFs (7 x N storage codes)
    ET = 16           # ext_subgrid_ecr: electron temperature (N)


class FieldMode(enum.IntEnum):
    EXACT = 0
    FAST = 1


# --------------------------------------------------------------------------
# ctypes structures (layout of include/ionsolver.h)
# --------------------------------------------------------------------------
class _Units(ctypes.Structure):
    _fields_ = [("m", ctypes.c_float), ("kg", ctypes.c_float), ("s", ctypes.c_float), ("a", ctypes.c_float),
                ("k", ctypes.c_float), ("prop", ctypes.c_int32)]


class _Config(ctypes.Structure):
    _fields_ = [
        ("velocity_set", ctypes.c_int32), ("relaxation_time", ctypes.c_int32), ("float_type", ctypes.c_int32),
        ("units", _Units),
        ("n_x", ctypes.c_uint32), ("n_y", ctypes.c_uint32), ("n_z", ctypes.c_uint32),
        ("d_x", ctypes.c_uint32), ("d_y", ctypes.c_uint32), ("d_z", ctypes.c_uint32),
        ("nu", ctypes.c_float), ("f_x", ctypes.c_float), ("f_y", ctypes.c_float), ("f_z", ctypes.c_float),
        ("ext_equilibrium_boudaries", ctypes.c_uint8), ("ext_volume_force", ctypes.c_uint8),
        ("ext_force_field", ctypes.c_uint8), ("ext_magneto_hydro", ctypes.c_uint8),
        ("ext_subgrid_ecr", ctypes.c_uint8), ("mhd_lod_depth", ctypes.c_uint8),
        ("graphics_active", ctypes.c_uint8), ("ext_charge_inlet", ctypes.c_uint8),
        ("ecr_freq", ctypes.c_float), ("ecr_field_strength", ctypes.c_float),
        ("run_steps", ctypes.c_uint64), ("electron_density0", ctypes.c_float),
    ]


class _KernelConstants(ctypes.Structure):
    _fields_ = [("w", ctypes.c_float), ("c", ctypes.c_float), ("ke", ctypes.c_float), ("kmu", ctypes.c_float),
                ("kmu0", ctypes.c_float), ("kkge", ctypes.c_float), ("wq", ctypes.c_float),
                ("n_lod", ctypes.c_uint32), ("kkbme", ctypes.c_float), ("keabs", ctypes.c_float)]


class _Timing(ctypes.Structure):
    _fields_ = [("ms_stream_collide", ctypes.c_double), ("ms_lod_reduce", ctypes.c_double),
                ("ms_field_prep", ctypes.c_double), ("ms_update_e_b", ctypes.c_double),
                ("ms_other", ctypes.c_double), ("n_steps", ctypes.c_uint64)]


class _SurConfig(ctypes.Structure):
    _fields_ = [("channels", ctypes.c_int32), ("seed", ctypes.c_uint64), ("lr", ctypes.c_float),
                ("lr_head", ctypes.c_float)]


class _SurStats(ctypes.Structure):
    _fields_ = [("train_steps", ctypes.c_uint64), ("direct_steps", ctypes.c_uint64),
                ("surrogate_steps", ctypes.c_uint64), ("fallbacks", ctypes.c_uint64),
                ("last_residual_e", ctypes.c_double), ("last_residual_b", ctypes.c_double)]


class FieldSolver(enum.IntEnum):
    DIRECT = 0
    SURROGATE = 1


class ModelType(enum.IntEnum):
    """ModelType, src/mesh.rs:9-14."""
    Solid = 0
    Magnet = 1      # magnetisation (x, y, z) in A/m
    Charged = 2     # charge per cell in C
    ChargedECR = 3  # charge per cell in C


# Every entry point of include/ionsolver.h: name -> (restype, argtypes)
_VP, _I, _U64, _F, _D, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
ABI = {
    "ion_config_default": (None, [ctypes.POINTER(_Config)]),
    "ion_units_default": (None, [ctypes.POINTER(_Units)]),
    "ion_units_set": (None, [ctypes.POINTER(_Units)] + [_F] * 10),
    "ion_kernel_constants_of": (_I, [ctypes.POINTER(_Config), ctypes.POINTER(_KernelConstants)]),
    "ion_create": (_I, [ctypes.POINTER(_Config), ctypes.POINTER(_VP)]),
    "ion_destroy": (None, [_VP]),
    "ion_get_config": (_I, [_VP, ctypes.POINTER(_Config)]),
    "ion_initialize": (_I, [_VP]),
    "ion_run": (_I, [_VP, _U64]),
    "ion_do_time_step": (_I, [_VP]),
    "ion_get_time_step": (_U64, [_VP]),
    "ion_update_fields": (_I, [_VP]),
    "ion_synchronize": (_I, [_VP]),
    "ion_field_size": (_I, [_VP, _I, ctypes.POINTER(_SZ), ctypes.POINTER(ctypes.c_uint32)]),
    "ion_write_field": (_I, [_VP, _I, _VP, _SZ]),
    "ion_read_field": (_I, [_VP, _I, _VP, _SZ]),
    "ion_set_volume_force": (_I, [_VP, _F, _F, _F]),
    "ion_save": (_I, [_VP, ctypes.c_char_p]),
    "ion_load": (_I, [ctypes.c_char_p, ctypes.POINTER(_Config), ctypes.POINTER(_VP)]),
    "ion_conservation": (_I, [_VP, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_set_field_mode": (_I, [_VP, _I]),
    "ion_set_timing": (_I, [_VP, _I]),
    "ion_get_timing": (_I, [_VP, ctypes.POINTER(_Timing)]),
    "ion_device_codec": (_I, [_I, _I, _VP, _VP, _U64]),
    "ion_update_e_b_noalloc": (_I, [_VP, _VP, _I, _VP]),
    "ion_set_watchdog": (_I, [_VP, ctypes.c_uint32]),
    "ion_watchdog_status": (_I, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_last_error": (ctypes.c_char_p, []),
    "ion_surrogate_config_default": (None, [ctypes.POINTER(_SurConfig)]),
    "ion_surrogate_create": (_I, [_VP, ctypes.POINTER(_SurConfig)]),
    "ion_compute_fields": (_I, [_VP]),
    "ion_surrogate_infer": (_I, [_VP]),
    "ion_surrogate_train_step": (_I, [_VP, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_surrogate_residual": (_I, [_VP, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_set_field_solver": (_I, [_VP, _I, ctypes.c_uint32, _D]),
    "ion_surrogate_set_learning_rate": (_I, [_VP, ctypes.c_float, ctypes.c_float]),
    "ion_hbm_copy_probe": (_I, [_U64, ctypes.c_uint32, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_hbm_stream_probe": (_I, [ctypes.c_uint32] * 4 + [ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_hbm_pull_probe": (_I, [ctypes.c_uint32] * 6 + [ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ion_import_mesh": (_I, [_VP, ctypes.c_char_p] + [_F] * 7 + [ctypes.POINTER(ctypes.c_uint32)]),
    "ion_import_mesh_reposition": (_I, [_VP, ctypes.c_char_p] + [_F] * 7 + [ctypes.POINTER(ctypes.c_uint32)]),
    "ion_mesh_get": (_I, [_VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), _VP, _VP]),
    "ion_voxelise_mesh": (_I, [_VP, ctypes.c_uint32, _I, _F, _F, _F]),
    "ion_precompute_b": (_I, [_VP]),
    "ion_precompute_e": (_I, [_VP]),
    "ion_precompute_e_ecr": (_I, [_VP]),
    "ion_psi_math_check": (_I, [ctypes.POINTER(_U64)] * 3),
    "ion_surrogate_get_stats": (_I, [_VP, ctypes.POINTER(_SurStats)]),
    "ion_surrogate_param_count": (_I, [_VP, ctypes.POINTER(_U64)]),
    "ion_surrogate_read_params": (_I, [_VP, _I, _VP, _U64]),
    "ion_surrogate_write_params": (_I, [_VP, _VP, _U64]),
}

_lib = None


def load_library():
    """Load libionsolver_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IonError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise IonError(f"[{rc}] {load_library().ion_last_error().decode()}")


# --------------------------------------------------------------------------
# Config mirror
# --------------------------------------------------------------------------
class Units:
    """Unit conversion factors (src/lbm/units.rs:13-59)."""

    def __init__(self):
        u = _Units()
        load_library().ion_units_default(ctypes.byref(u))
        self._u = u

    def __getattr__(self, name):
        if name in ("m", "kg", "s", "a", "k", "prop"):
            return getattr(self.__dict__["_u"], name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in ("m", "kg", "s", "a", "k", "prop"):
            setattr(self._u, name, value)
        else:
            super().__setattr__(name, value)

    def set(self, lbm_length, lbm_velocity, lbm_rho, lbm_charge, lbm_temp, si_length, si_velocity, si_rho,
            si_charge, si_temp):
        load_library().ion_units_set(ctypes.byref(self._u), lbm_length, lbm_velocity, lbm_rho, lbm_charge, lbm_temp,
                                     si_length, si_velocity, si_rho, si_charge, si_temp)

    # conversions used by setups (units.rs:62-147), f32 arithmetic
    def nu_si_lu(self, nu):
        f = np.float32
        return float(f(nu) / (f(self.m) * f(self.m) / f(self.s)))

    def e_field_si_lu(self, e):
        f = np.float32
        return float(f(e) / ((f(self.kg) * f(self.m)) / (f(self.a) * (f(self.s) * f(self.s) * f(self.s)))))

    def mag_flux_si_lu(self, b):
        f = np.float32
        return float(f(b) / (f(self.kg) / (f(self.a) * (f(self.s) * f(self.s)))))


class GraphicsConfig:
    """Only ``graphics_active`` reaches the kernels (UPDATE_FIELDS, domain.rs:856)."""

    def __init__(self):
        self.graphics_active = True  # src/lbm/graphics.rs:160


class LbmConfig:
    """LbmConfig with the reference defaults (src/lbm/mod.rs:105-135)."""

    def __init__(self):
        self.velocity_set = VelocitySet.D2Q9
        self.relaxation_time = RelaxationTime.Srt
        self.float_type = FloatType.FP16S
        self.units = Units()
        self.n_x = self.n_y = self.n_z = 1
        self.d_x = self.d_y = self.d_z = 1
        self.nu = float(np.float32(1.0) / np.float32(6.0))
        self.f_x = self.f_y = self.f_z = 0.0
        self.ext_equilibrium_boudaries = False
        self.ext_volume_force = False
        self.ext_force_field = False
        self.ext_magneto_hydro = False
        self.ext_subgrid_ecr = False
        # no reference counterpart: TYPE_E cells also hold their charge (Q) fixed
        self.ext_charge_inlet = False
        self.mhd_lod_depth = 4
        self.ecr_freq = 0.0
        self.ecr_field_strength = 0.0
        self.graphics_config = GraphicsConfig()
        self.run_steps = 0
        # no reference counterpart: initial electron-gas density of non-solid
        # cells (0 = the reference's initialize)
        self.electron_density0 = 0.0

    new = classmethod(lambda cls: cls())

    def _to_c(self):
        c = _Config()
        c.velocity_set = int(self.velocity_set)
        c.relaxation_time = int(self.relaxation_time)
        c.float_type = int(self.float_type)
        c.units = self.units._u
        c.n_x, c.n_y, c.n_z = self.n_x, self.n_y, self.n_z
        c.d_x, c.d_y, c.d_z = self.d_x, self.d_y, self.d_z
        c.nu = self.nu
        c.f_x, c.f_y, c.f_z = self.f_x, self.f_y, self.f_z
        c.ext_equilibrium_boudaries = int(bool(self.ext_equilibrium_boudaries))
        c.ext_volume_force = int(bool(self.ext_volume_force))
        c.ext_force_field = int(bool(self.ext_force_field))
        c.ext_magneto_hydro = int(bool(self.ext_magneto_hydro))
        c.ext_subgrid_ecr = int(bool(self.ext_subgrid_ecr))
        c.ext_charge_inlet = int(bool(self.ext_charge_inlet))
        c.mhd_lod_depth = self.mhd_lod_depth
        c.graphics_active = int(bool(self.graphics_config.graphics_active))
        c.ecr_freq = self.ecr_freq
        c.ecr_field_strength = self.ecr_field_strength
        c.run_steps = self.run_steps
        c.electron_density0 = self.electron_density0
        return c

    @classmethod
    def _from_c(cls, c):
        cfg = cls()
        cfg.velocity_set = VelocitySet(c.velocity_set)
        cfg.relaxation_time = RelaxationTime(c.relaxation_time)
        cfg.float_type = FloatType(c.float_type)
        for k in ("m", "kg", "s", "a", "k", "prop"):
            setattr(cfg.units, k, getattr(c.units, k))
        cfg.n_x, cfg.n_y, cfg.n_z = c.n_x, c.n_y, c.n_z
        cfg.d_x, cfg.d_y, cfg.d_z = c.d_x, c.d_y, c.d_z
        cfg.nu = c.nu
        cfg.f_x, cfg.f_y, cfg.f_z = c.f_x, c.f_y, c.f_z
        cfg.ext_equilibrium_boudaries = bool(c.ext_equilibrium_boudaries)
        cfg.ext_volume_force = bool(c.ext_volume_force)
        cfg.ext_force_field = bool(c.ext_force_field)
        cfg.ext_magneto_hydro = bool(c.ext_magneto_hydro)
        cfg.ext_subgrid_ecr = bool(c.ext_subgrid_ecr)
        cfg.ext_charge_inlet = bool(c.ext_charge_inlet)
        cfg.mhd_lod_depth = c.mhd_lod_depth
        cfg.graphics_config.graphics_active = bool(c.graphics_active)
        cfg.ecr_freq, cfg.ecr_field_strength, cfg.run_steps = c.ecr_freq, c.ecr_field_strength, c.run_steps
        cfg.electron_density0 = c.electron_density0
        return cfg

    def kernel_constants(self):
        """The f32 constants the reference bakes into its kernels (domain.rs:808, 836-849)."""
        k = _KernelConstants()
        _check(load_library().ion_kernel_constants_of(ctypes.byref(self._to_c()), ctypes.byref(k)))
        return {f: getattr(k, f) for f, _ in k._fields_}


_DTYPES = {8: np.float64, 4: np.float32, 2: np.uint16, 1: np.uint8}


class Lbm:
    """One simulation on one GPU (reference ``Lbm`` with a single domain)."""

    def __init__(self, config, _handle=None):
        self._lib = load_library()
        if _handle is None:
            h = ctypes.c_void_p()
            _check(self._lib.ion_create(ctypes.byref(config._to_c()), ctypes.byref(h)))
            _handle = h
        self._h = _handle
        c = _Config()
        _check(self._lib.ion_get_config(self._h, ctypes.byref(c)))
        self.config = LbmConfig._from_c(c)
        self.n = self.config.n_x * self.config.n_y * self.config.n_z

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ion_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # Lbm::initialize / run / do_time_step / get_time_step
    def initialize(self):
        _check(self._lib.ion_initialize(self._h))

    def run(self, steps):
        _check(self._lib.ion_run(self._h, int(steps)))

    def do_time_step(self):
        _check(self._lib.ion_do_time_step(self._h))

    def get_time_step(self):
        return int(self._lib.ion_get_time_step(self._h))

    def update_fields(self):
        _check(self._lib.ion_update_fields(self._h))

    def synchronize(self):
        _check(self._lib.ion_synchronize(self._h))

    # buffers
    def field_size(self, field):
        b = ctypes.c_size_t()
        e = ctypes.c_uint32()
        _check(self._lib.ion_field_size(self._h, int(field), ctypes.byref(b), ctypes.byref(e)))
        return b.value, e.value

    def read_field(self, field):
        nbytes, esz = self.field_size(field)
        if nbytes == 0:
            raise IonError(f"field {Field(field).name} does not exist in this configuration")
        out = np.empty(nbytes // esz, dtype=_DTYPES[esz])
        _check(self._lib.ion_read_field(self._h, int(field), out.ctypes.data_as(ctypes.c_void_p), nbytes))
        return out

    def write_field(self, field, data):
        nbytes, esz = self.field_size(field)
        arr = np.ascontiguousarray(data, dtype=_DTYPES[esz])
        _check(self._lib.ion_write_field(self._h, int(field), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))

    def set_volume_force(self, fx, fy, fz):
        _check(self._lib.ion_set_volume_force(self._h, fx, fy, fz))

    # persistence (src/file.rs)
    def save(self, path):
        _check(self._lib.ion_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path, config=None):
        c = (config or LbmConfig())._to_c()
        h = ctypes.c_void_p()
        _check(load_library().ion_load(os.fsencode(path), ctypes.byref(c), ctypes.byref(h)))
        return cls(None, _handle=h)

    # diagnostics / measurement
    def conservation(self):
        m = ctypes.c_double()
        q = ctypes.c_double()
        _check(self._lib.ion_conservation(self._h, ctypes.byref(m), ctypes.byref(q)))
        return m.value, q.value

    def set_watchdog(self, every):
        """Check the float64 mass and charge sums every `every` steps; a
        non-finite one stops run/do_time_step with IonError (0: off)."""
        _check(self._lib.ion_set_watchdog(self._h, every))

    def watchdog_status(self):
        """(first step with a non-finite sum or None, mass, charge of the last check)"""
        t, m, q = _U64(), _D(), _D()
        _check(self._lib.ion_watchdog_status(self._h, ctypes.byref(t), ctypes.byref(m), ctypes.byref(q)))
        return (None if t.value == 2 ** 64 - 1 else t.value), m.value, q.value

    def set_field_mode(self, mode):
        _check(self._lib.ion_set_field_mode(self._h, int(mode)))

    def set_timing(self, enable):
        _check(self._lib.ion_set_timing(self._h, int(bool(enable))))

    def get_timing(self):
        t = _Timing()
        _check(self._lib.ion_get_timing(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}


    # learned field surrogate
    def surrogate_create(self, channels=16, seed=1, lr=1e-3, lr_head=1e-4):
        c = _SurConfig()
        self._lib.ion_surrogate_config_default(ctypes.byref(c))
        c.channels, c.seed, c.lr, c.lr_head = channels, seed, lr, lr_head
        _check(self._lib.ion_surrogate_create(self._h, ctypes.byref(c)))

    def compute_fields(self):
        """Direct field update from the current LOD sums (enqueue_update_e_b_dyn)."""
        _check(self._lib.ion_compute_fields(self._h))

    def surrogate_infer(self):
        _check(self._lib.ion_surrogate_infer(self._h))

    def surrogate_train_step(self, update=True):
        e, b = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ion_surrogate_train_step(self._h, int(bool(update)), ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def surrogate_residual(self):
        e, b = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ion_surrogate_residual(self._h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def surrogate_set_learning_rate(self, lr, lr_head):
        _check(self._lib.ion_surrogate_set_learning_rate(self._h, float(lr), float(lr_head)))

    def set_field_solver(self, solver, judge_every=50, tolerance=0.01):
        _check(self._lib.ion_set_field_solver(self._h, int(solver), int(judge_every), float(tolerance)))

    def surrogate_stats(self):
        st = _SurStats()
        _check(self._lib.ion_surrogate_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def surrogate_params(self, grads=False):
        n = ctypes.c_uint64()
        _check(self._lib.ion_surrogate_param_count(self._h, ctypes.byref(n)))
        out = np.empty(n.value, np.float32)
        _check(self._lib.ion_surrogate_read_params(self._h, int(bool(grads)), out.ctypes.data_as(ctypes.c_void_p),
                                                   n.value))
        return out

    def set_surrogate_params(self, params):
        arr = np.ascontiguousarray(params, np.float32)
        _check(self._lib.ion_surrogate_write_params(self._h, arr.ctypes.data_as(ctypes.c_void_p), arr.size))

    # meshes and static fields (src/mesh.rs, src/lbm/mod.rs:284-314)
    def import_mesh(self, path, scale, ox, oy, oz, rx=0.0, ry=0.0, rz=0.0):
        """Lbm::import_mesh (src/mesh.rs:234); returns the mesh index."""
        i = ctypes.c_uint32()
        _check(self._lib.ion_import_mesh(self._h, os.fsencode(path), scale, ox, oy, oz, rx, ry, rz, ctypes.byref(i)))
        return i.value

    def import_mesh_reposition(self, path, cx, cy, cz, rx, ry, rz, size):
        """Lbm::import_mesh_reposition (src/mesh.rs:248); returns the mesh index."""
        i = ctypes.c_uint32()
        _check(self._lib.ion_import_mesh_reposition(self._h, os.fsencode(path), cx, cy, cz, rx, ry, rz, size,
                                                    ctypes.byref(i)))
        return i.value

    def mesh(self, index):
        """(vertices [T, 3, 3] f32, bounds dict) of an imported mesh."""
        t = ctypes.c_uint32()
        _check(self._lib.ion_mesh_get(self._h, index, ctypes.byref(t), None, None))
        v = np.empty((t.value, 3, 3), np.float32)
        b = np.empty(9, np.float32)
        _check(self._lib.ion_mesh_get(self._h, index, ctypes.byref(t), v.ctypes.data_as(ctypes.c_void_p),
                                      b.ctypes.data_as(ctypes.c_void_p)))
        return v, dict(p_min=b[0:3], p_max=b[3:6], center=b[6:9])

    def voxelise_mesh(self, index, model_type, value=(0.0, 0.0, 0.0)):
        """Lbm::voxelise_mesh (src/mesh.rs:256); value: magnetisation (A/m) or (charge, 0, 0) in C."""
        v = tuple(value) if np.ndim(value) else (value, 0.0, 0.0)
        _check(self._lib.ion_voxelise_mesh(self._h, index, int(model_type), *[float(x) for x in v]))

    def precompute_b(self):
        """Lbm::precompute_B (src/lbm/mod.rs:284)."""
        _check(self._lib.ion_precompute_b(self._h))

    def precompute_e(self):
        """Lbm::precompute_E (src/lbm/mod.rs:302)."""
        _check(self._lib.ion_precompute_e(self._h))

    def precompute_e_ecr(self):
        """Lbm::precompute_E_ECR (src/lbm/mod.rs:319): the static E of charged meshes into E_var."""
        _check(self._lib.ion_precompute_e_ecr(self._h))


def hbm_copy_probe(nbytes=1 << 30, reps=10):
    """Best (ms, GB/s read + write) of a plain 16-byte streaming copy of nbytes."""
    ms, gbs = ctypes.c_double(), ctypes.c_double()
    _check(load_library().ion_hbm_copy_probe(int(nbytes), int(reps), ctypes.byref(ms), ctypes.byref(gbs)))
    return ms.value, gbs.value


def hbm_stream_probe(elem_bytes, streams, cells, reps=10):
    """Best (ms, GB/s read + write) of stream_collide's access shape without arithmetic."""
    ms, gbs = ctypes.c_double(), ctypes.c_double()
    _check(load_library().ion_hbm_stream_probe(int(elem_bytes), int(streams), int(cells), int(reps),
                                               ctypes.byref(ms), ctypes.byref(gbs)))
    return ms.value, gbs.value


def hbm_pull_probe(elem_bytes, streams, nx, ny, nz, reps=10):
    """Best (ms, GB/s read + write) of stream_collide's in-place esoteric-pull
    addressing without arithmetic on an nx x ny x nz lattice."""
    ms, gbs = ctypes.c_double(), ctypes.c_double()
    _check(load_library().ion_hbm_pull_probe(int(elem_bytes), int(streams), int(nx), int(ny), int(nz), int(reps),
                                             ctypes.byref(ms), ctypes.byref(gbs)))
    return ms.value, gbs.value


def psi_math_check():
    """(sqrt mismatches, division mismatches, division samples) of the static-field kernel's operators."""
    a, b, c = _U64(), _U64(), _U64()
    _check(load_library().ion_psi_math_check(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value


def device_codec(float_type, direction, data):
    """Run the device storage codec on a host array (0: encode, 1: decode)."""
    L = load_library()
    if direction == 0:
        src = np.ascontiguousarray(data, np.float32)
        out = np.empty(src.shape, np.uint16)
    else:
        src = np.ascontiguousarray(data, np.uint16)
        out = np.empty(src.shape, np.float32)
    _check(L.ion_device_codec(int(float_type), direction, src.ctypes.data_as(ctypes.c_void_p),
                              out.ctypes.data_as(ctypes.c_void_p), src.size))
    return out
