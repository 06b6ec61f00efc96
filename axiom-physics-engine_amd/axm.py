"""Python host binding of the C ABI (include/axiom/axm.h), via ctypes.

Mirrors the C++20 API in include/axiom/dynamics/physics_world.hpp
(axiom::dynamics::PhysicsWorld): create a world from a description, upload
bodies, step, read back bodies and stage outputs.  Failures raise AxmError
carrying the axiom::core::ErrorCode value (include/axiom/core/error_code.hpp:21-57).

The GPU path is the only compute path: if libaxm.so is missing this module
raises on import of the library, it never falls back to a CPU implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaxm.so")

SUCCESS = 0
INVALID_PARAMETER = 600
OUT_OF_RANGE = 601
ERROR_NAMES = {
    0: "Success", 200: "OutOfMemory", 202: "NullPointer", 300: "InvalidShape", 301: "GJKFailedToConverge",
    302: "EPAFailedToConverge", 400: "InvalidRigidBody", 401: "ConstraintSolverFailure",
    502: "BufferAllocationFailed", 503: "GPU_INVALID_OPERATION", 504: "GPU_TIMEOUT", 505: "GPU_OPERATION_FAILED",
    600: "InvalidParameter", 601: "OutOfRange",
}
SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_PLANE = 0, 1, 2, 3
PLANE_BIT = 0x80000000
WARM_NONE, WARM_CACHE, WARM_MLP = 0, 1, 2
STEP_TIMINGS, STEP_RESIDUAL, STEP_SYNC, STEP_KEEP_STAGES = 1, 2, 4, 8
MAX_PLANES = 8


class AxmError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = int(code)
        self.name = ERROR_NAMES.get(self.code, "Unknown")
        super().__init__("%s: %s (%d)" % (where, self.name, self.code))


class WorldDesc(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3), ("friction", C.c_float), ("restitution", C.c_float),
        ("rest_threshold", C.c_float), ("baumgarte", C.c_float), ("slop", C.c_float), ("max_bias", C.c_float),
        ("margin", C.c_float), ("lin_damp", C.c_float), ("ang_damp", C.c_float), ("cell_size", C.c_float),
        ("origin", C.c_float * 3), ("num_planes", C.c_uint32), ("planes", (C.c_float * 4) * MAX_PLANES),
        ("warm_start", C.c_uint32), ("record_residual", C.c_uint32), ("softness", C.c_float),
        ("allow_sleep", C.c_uint32), ("sleep_linear", C.c_float), ("sleep_angular", C.c_float),
        ("sleep_time", C.c_float), ("island_solve", C.c_uint32), ("max_bodies", C.c_uint32),
        ("max_pairs", C.c_uint32), ("max_manifolds", C.c_uint32), ("auto_grow", C.c_uint32),
        ("residual_tol", C.c_float), ("residual_every", C.c_uint32), ("max_iterations", C.c_uint32),
        ("device", C.c_int32), ("solver_mode", C.c_uint32), ("strict", C.c_uint32),
    ]


class BodySoA(C.Structure):
    _fields_ = [
        ("pos", C.POINTER(C.c_float)), ("quat", C.POINTER(C.c_float)), ("lin_vel", C.POINTER(C.c_float)),
        ("ang_vel", C.POINTER(C.c_float)), ("inv_mass", C.POINTER(C.c_float)),
        ("inv_inertia", C.POINTER(C.c_float)), ("shape_type", C.POINTER(C.c_uint32)),
        ("shape_params", C.POINTER(C.c_float)), ("world", C.POINTER(C.c_uint32)),
        ("material", C.POINTER(C.c_float)), ("flags", C.POINTER(C.c_uint32)), ("filter", C.POINTER(C.c_uint32)),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("pairs", C.c_uint32), ("manifolds", C.c_uint32), ("contacts", C.c_uint32), ("colours", C.c_uint32),
        ("colour_rounds", C.c_uint32), ("iterations", C.c_uint32), ("residual", C.c_float),
        ("residual_first", C.c_float), ("gjk_fail", C.c_uint32), ("epa_fail", C.c_uint32),
        ("warm_hits", C.c_uint32), ("overflow", C.c_uint32), ("broadphase_ms", C.c_float),
        ("narrowphase_ms", C.c_float), ("solver_ms", C.c_float), ("integration_ms", C.c_float),
        ("total_ms", C.c_float), ("islands", C.c_uint32), ("sleeping", C.c_uint32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelProfile(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("pgs_launches", C.c_uint64), ("pgs_ms", C.c_double),
                ("pgs_bytes", C.c_double), ("sweeps", C.c_uint64), ("bundle_launches", C.c_uint64),
                ("bundle_ms", C.c_double), ("bundle_point_sweeps", C.c_uint64)]


class WorldTotals(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("steps_at_tol", C.c_uint64), ("sweeps", C.c_uint64),
                ("contact_sweeps", C.c_uint64), ("max_sweeps", C.c_uint32), ("pad", C.c_uint32),
                ("local_contact_sweeps", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class SweepConfig(C.Structure):
    _fields_ = [("groups", C.c_uint32), ("lds_kib", C.c_uint32), ("policy", C.c_uint32), ("sweep_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Exported C-ABI symbols (the header's declarations).
EXPORTS = [
    "axm_device_count", "axm_status_string", "axm_world_desc_default", "axm_world_create", "axm_world_destroy", "axm_world_set_bodies",
    "axm_world_get_bodies", "axm_world_body_count", "axm_world_step", "axm_world_run", "axm_world_sync", "axm_world_profile", "axm_world_get_totals", "axm_world_get_islands", "axm_world_configure", "axm_world_get_desc", "axm_world_save", "axm_world_load", "axm_mlp_save", "axm_mlp_load",
    "axm_world_get_pairs", "axm_world_get_manifolds", "axm_world_get_colours", "axm_world_get_impulses", "axm_world_get_friction_bounds",
    "axm_world_get_residuals", "axm_world_get_pre_solve_velocities", "axm_broadphase", "axm_collide",
    "axm_collide_pairs", "axm_collide_pairs_noalloc", "axm_gjk_epa_pairs", "axm_gjk_epa_pairs_capped", "axm_scene_generate", "axm_mass_properties", "axm_rng_stream",
    "axm_mlp_create", "axm_mlp_destroy", "axm_mlp_init", "axm_mlp_set_params", "axm_mlp_get_params",
    "axm_mlp_forward", "axm_mlp_loss_grad", "axm_mlp_adam_step", "axm_mlp_train", "axm_world_set_mlp",
    "axm_world_set_feature_logging", "axm_world_get_features", "axm_world_set_sweep_config", "axm_world_autotune",
]
MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 24, 3, 20099

_lib = None


def lib():
    """Loads libaxm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libaxm.so not built (run __graft_entry__.build()): " + LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32p, f32p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_int32)
    L.axm_device_count.restype = C.c_int32
    L.axm_device_count.argtypes = []
    L.axm_status_string.restype = C.c_char_p
    L.axm_status_string.argtypes = [C.c_int32]
    L.axm_world_desc_default.argtypes = [C.POINTER(WorldDesc)]
    L.axm_world_create.argtypes = [C.POINTER(WorldDesc), C.POINTER(P)]
    L.axm_world_destroy.argtypes = [P]
    L.axm_world_set_bodies.argtypes = [P, C.POINTER(BodySoA), C.c_uint32]
    L.axm_world_get_bodies.argtypes = [P, C.POINTER(BodySoA)]
    L.axm_world_body_count.argtypes = [P, u32p]
    L.axm_world_step.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(StepStats)]
    L.axm_world_run.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(StepStats)]
    L.axm_world_sync.argtypes = [P]
    L.axm_world_profile.argtypes = [P, C.c_int, C.POINTER(KernelProfile)]
    L.axm_world_get_totals.argtypes = [P, C.POINTER(WorldTotals), C.c_int]
    L.axm_world_get_islands.argtypes = [P, u32p, u32p, f32p]
    L.axm_world_configure.argtypes = [P, C.POINTER(WorldDesc)]
    L.axm_world_get_desc.argtypes = [P, C.POINTER(WorldDesc)]
    L.axm_world_save.argtypes = [P, C.c_char_p]
    L.axm_world_load.argtypes = [P, C.c_char_p]
    L.axm_mlp_save.argtypes = [P, C.c_char_p]
    L.axm_mlp_load.argtypes = [P, C.c_char_p]
    L.axm_world_get_pairs.argtypes = [P, u32p, u32p, C.c_uint32, u32p]
    L.axm_world_get_manifolds.argtypes = [P, u32p, u32p, u32p, f32p, f32p, f32p, u32p, C.c_uint32, u32p]
    L.axm_world_get_colours.argtypes = [P, u32p, u32p, C.c_uint32, u32p, u32p]
    L.axm_world_get_impulses.argtypes = [P, f32p, f32p, C.c_uint32, u32p]
    L.axm_world_get_friction_bounds.argtypes = [P, f32p, C.c_uint32, u32p]
    L.axm_world_get_residuals.argtypes = [P, f32p, C.c_uint32, u32p]
    L.axm_world_get_pre_solve_velocities.argtypes = [P, f32p, f32p]
    L.axm_broadphase.argtypes = [P, u32p]
    L.axm_collide.argtypes = [P, u32p]
    L.axm_collide_pairs.argtypes = [C.c_int32, C.c_uint32, u32p, f32p, f32p, f32p, u32p, f32p, f32p, f32p, C.c_float,
                                    u32p, f32p, f32p, f32p, u32p]
    L.axm_collide_pairs_noalloc.argtypes = [C.c_int32, C.c_uint32, u32p, f32p, f32p, f32p, u32p, f32p, f32p, f32p, C.c_float,
                                    u32p, f32p, f32p, f32p, u32p]
    L.axm_gjk_epa_pairs.argtypes = [C.c_int32, C.c_uint32, u32p, f32p, f32p, f32p, u32p, f32p, f32p, f32p, f32p, i32p]
    L.axm_gjk_epa_pairs_capped.argtypes = [C.c_int32, C.c_uint32, u32p, f32p, f32p, f32p, u32p, f32p, f32p, f32p,
                                           C.c_uint32, C.c_uint32, f32p, i32p]
    L.axm_scene_generate.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(BodySoA), u32p,
                                     C.POINTER(WorldDesc)]
    L.axm_mass_properties.argtypes = [C.c_uint32, f32p, C.c_float, f32p, f32p]
    f64p = C.POINTER(C.c_double)
    L.axm_mlp_create.argtypes = [C.c_int32, C.POINTER(P)]
    L.axm_mlp_destroy.argtypes = [P]
    L.axm_mlp_init.argtypes = [P, C.c_uint64]
    L.axm_mlp_set_params.argtypes = [P, f32p]
    L.axm_mlp_get_params.argtypes = [P, f32p]
    L.axm_mlp_forward.argtypes = [P, f32p, C.c_uint32, f32p]
    L.axm_mlp_loss_grad.argtypes = [P, f32p, f32p, C.c_uint32, f32p, f64p]
    L.axm_mlp_adam_step.argtypes = [P, f32p, C.c_float, C.c_float, C.c_float, C.c_float]
    L.axm_mlp_train.argtypes = [P, f32p, f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, f32p]
    L.axm_world_set_mlp.argtypes = [P, P]
    L.axm_world_set_sweep_config.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
    L.axm_world_autotune.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(SweepConfig)]
    L.axm_world_set_feature_logging.argtypes = [P, C.c_int]
    L.axm_world_get_features.argtypes = [P, f32p, f32p, C.c_uint32, u32p]
    L.axm_rng_stream.argtypes = [C.c_uint64, C.c_uint32, u32p, C.c_float, C.c_float, f32p]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("axm_status_string", "axm_world_desc_default", "axm_rng_stream"):
            f.restype = C.c_int32
    L.axm_rng_stream.restype = None
    L.axm_world_desc_default.restype = None
    _lib = L
    return L


def device_count():
    return int(lib().axm_device_count())


def _p(a, ct=C.c_float):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def _check(code, where):
    if code != SUCCESS:
        raise AxmError(code, where)


def default_desc(**kw):
    d = WorldDesc()
    lib().axm_world_desc_default(C.byref(d))
    for k, v in kw.items():
        setattr(d, k, v)
    return d


class Bodies:
    """Host structure-of-arrays body set (gui::RigidBodyData field set)."""

    FIELDS = ("pos", "quat", "lin_vel", "ang_vel", "inv_mass", "inv_inertia", "shape_type", "shape_params", "world")

    def __init__(self, n):
        self.n = n
        self.pos = np.zeros((n, 3), np.float32)
        self.quat = np.zeros((n, 4), np.float32)
        self.quat[:, 3] = 1.0
        self.lin_vel = np.zeros((n, 3), np.float32)
        self.ang_vel = np.zeros((n, 3), np.float32)
        self.inv_mass = np.zeros(n, np.float32)
        self.inv_inertia = np.zeros((n, 3), np.float32)
        self.shape_type = np.zeros(n, np.uint32)
        self.shape_params = np.zeros((n, 3), np.float32)
        self.world = np.zeros(n, np.uint32)
        # optional extras (None: the world description's values): material [n, 4] (friction,
        # restitution, linear damping, angular damping), flags [n] (bit 0 kinematic), filter [n, 3]
        # (category, mask, group as int32 bits)
        self.material = None
        self.flags = None
        self.filter = None

    def copy(self):
        b = Bodies(self.n)
        for f in self.FIELDS:
            setattr(b, f, getattr(self, f).copy())
        for f in ("material", "flags", "filter"):
            a = getattr(self, f)
            setattr(b, f, None if a is None else a.copy())
        return b

    def soa(self):
        for f in self.FIELDS:
            a = getattr(self, f)
            if not a.flags["C_CONTIGUOUS"]:
                setattr(self, f, np.ascontiguousarray(a))
        s = BodySoA()
        s.pos, s.quat, s.lin_vel, s.ang_vel = _p(self.pos), _p(self.quat), _p(self.lin_vel), _p(self.ang_vel)
        s.inv_mass, s.inv_inertia = _p(self.inv_mass), _p(self.inv_inertia)
        s.shape_type = _p(self.shape_type, C.c_uint32)
        s.shape_params = _p(self.shape_params)
        s.world = _p(self.world, C.c_uint32)
        for f, ct, dt in (("material", C.c_float, np.float32), ("flags", C.c_uint32, np.uint32),
                          ("filter", C.c_uint32, np.uint32)):
            a = getattr(self, f)
            if a is not None:
                a = np.ascontiguousarray(a, dt)
                setattr(self, f, a)
                setattr(s, f, _p(a, ct))
        return s


def scene(config, seed, n=0, n_worlds=0):
    """Seeded synthetic scene: returns (Bodies, WorldDesc)."""
    L = lib()
    cnt = C.c_uint32(0)
    d = WorldDesc()
    empty = BodySoA()
    _check(L.axm_scene_generate(config, seed, n, n_worlds, C.byref(empty), C.byref(cnt), C.byref(d)), "scene")
    b = Bodies(cnt.value)
    s = b.soa()
    _check(L.axm_scene_generate(config, seed, n, n_worlds, C.byref(s), C.byref(cnt), C.byref(d)), "scene")
    return b, d


def mass_properties(shape_type, params, density=1000.0):
    """(inverse mass, inverse principal inertia[3]) of a solid shape."""
    prm = np.zeros(3, np.float32)
    prm[:len(params)] = params
    im = C.c_float()
    ii = np.zeros(3, np.float32)
    _check(lib().axm_mass_properties(shape_type, _p(prm), density, C.byref(im), _p(ii)), "axm_mass_properties")
    return im.value, ii


class Mlp:
    """Learned warm-start network on one GPU (axm_mlp_*)."""

    def __init__(self, device=0, seed=None):
        self._L = lib()
        h = C.c_void_p()
        _check(self._L.axm_mlp_create(device, C.byref(h)), "axm_mlp_create")
        self._h = h
        if seed is not None:
            self.init(seed)

    def close(self):
        if getattr(self, "_h", None):
            self._L.axm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed):
        _check(self._L.axm_mlp_init(self._h, seed), "axm_mlp_init")

    @property
    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        _check(self._L.axm_mlp_get_params(self._h, _p(p)), "axm_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        _check(self._L.axm_mlp_set_params(self._h, _p(p)), "axm_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], MLP_OUTPUTS), np.float32)
        _check(self._L.axm_mlp_forward(self._h, _p(X), X.shape[0], _p(Y)), "axm_mlp_forward")
        return Y

    def loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(self._L.axm_mlp_loss_grad(self._h, _p(X), _p(Y), X.shape[0], _p(g), C.byref(loss)), "axm_mlp_loss_grad")
        return loss.value, g

    def adam_step(self, grad=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        g = None if grad is None else np.ascontiguousarray(grad, np.float32)
        _check(self._L.axm_mlp_adam_step(self._h, _p(g), lr, beta1, beta2, eps), "axm_mlp_adam_step")

    def train(self, X, Y, epochs=1, batch=4096, lr=1e-3):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        losses = np.zeros(epochs, np.float32)
        _check(self._L.axm_mlp_train(self._h, _p(X), _p(Y), X.shape[0], epochs, batch, lr, _p(losses)),
               "axm_mlp_train")
        return losses

    def save(self, path):
        """Writes the parameters and the Adam state to `path`."""
        _check(self._L.axm_mlp_save(self._h, os.fsencode(str(path))), "axm_mlp_save")

    def load(self, path):
        """Restores parameters and Adam state from a file written by save()."""
        _check(self._L.axm_mlp_load(self._h, os.fsencode(str(path))), "axm_mlp_load")


def rng_stream(seed, n, lo=0.0, hi=1.0):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    lib().axm_rng_stream(seed, n, _p(u, C.c_uint32), lo, hi, _p(f))
    return u, f


class World:
    """A simulation world on one GPU (axiom::dynamics::PhysicsWorld)."""

    def __init__(self, desc=None, **kw):
        self._L = lib()
        self.desc = desc if desc is not None else default_desc()
        for k, v in kw.items():
            setattr(self.desc, k, v)
        h = C.c_void_p()
        _check(self._L.axm_world_create(C.byref(self.desc), C.byref(h)), "axm_world_create")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.axm_world_destroy(self._h)
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

    def set_bodies(self, b):
        self._keep = b
        _check(self._L.axm_world_set_bodies(self._h, C.byref(b.soa()), b.n), "axm_world_set_bodies")
        self.n = b.n

    def get_bodies(self, into=None):
        b = into if into is not None else Bodies(self.n)
        s = BodySoA()
        s.pos, s.quat, s.lin_vel, s.ang_vel = _p(b.pos), _p(b.quat), _p(b.lin_vel), _p(b.ang_vel)
        _check(self._L.axm_world_get_bodies(self._h, C.byref(s)), "axm_world_get_bodies")
        return b

    def step(self, dt, iters, flags=0):
        st = StepStats()
        _check(self._L.axm_world_step(self._h, dt, iters, flags, C.byref(st)), "axm_world_step")
        return st

    def run(self, dt, iters, steps, graph=False):
        st = StepStats()
        _check(self._L.axm_world_run(self._h, dt, iters, steps, 1 if graph else 0, C.byref(st)), "axm_world_run")
        return st

    def sync(self):
        _check(self._L.axm_world_sync(self._h), "axm_world_sync")

    def profile(self, enable):
        """Enables/disables sweep-kernel event timing; returns the profile accumulated so far."""
        p = KernelProfile()
        _check(self._L.axm_world_profile(self._h, 1 if enable else 0, C.byref(p)), "axm_world_profile")
        return p

    def configure(self, desc):
        """Replaces the solver settings of the live world (capacities, grid and device stay)."""
        _check(self._L.axm_world_configure(self._h, C.byref(desc)), "axm_world_configure")
        self.desc = desc

    def get_desc(self):
        """The world's description as the library holds it."""
        d = WorldDesc()
        _check(self._L.axm_world_get_desc(self._h, C.byref(d)), "axm_world_get_desc")
        return d

    def save(self, path):
        """Writes a snapshot (description, bodies, last step's warm-start cache) to `path`."""
        _check(self._L.axm_world_save(self._h, os.fsencode(str(path))), "axm_world_save")

    def load(self, path):
        """Replaces this world's description (not its device), bodies and cache with a snapshot's."""
        _check(self._L.axm_world_load(self._h, os.fsencode(str(path))), "axm_world_load")
        self.desc = self.get_desc()
        cnt = C.c_uint32(0)
        _check(self._L.axm_world_body_count(self._h, C.byref(cnt)), "axm_world_body_count")
        self.n = int(cnt.value)

    def islands(self):
        """(label, asleep, rest time) of every body after the last step."""
        n = self.n
        lab = np.zeros(n, np.uint32)
        sl = np.zeros(n, np.uint32)
        rest = np.zeros(n, np.float32)
        _check(self._L.axm_world_get_islands(self._h, _p(lab, C.c_uint32), _p(sl, C.c_uint32), _p(rest)),
               "axm_world_get_islands")
        return lab, sl, rest

    def totals(self, reset=False):
        """Device-counted totals since the last reset (steps, steps at tolerance, sweeps, contact x sweeps,
        max sweeps of one step)."""
        t = WorldTotals()
        _check(self._L.axm_world_get_totals(self._h, C.byref(t), 1 if reset else 0), "axm_world_get_totals")
        return t

    def set_mlp(self, mlp):
        """Attaches a warm-start network (used when desc.warm_start == WARM_MLP); None detaches."""
        self._mlp = mlp
        _check(self._L.axm_world_set_mlp(self._h, mlp._h if mlp is not None else None), "axm_world_set_mlp")

    def set_sweep_config(self, groups=0, lds_kib=0, policy=0x12):
        _check(self._L.axm_world_set_sweep_config(self._h, groups, lds_kib, policy), "axm_world_set_sweep_config")

    def autotune(self, dt, iters, steps=5):
        """Picks the fastest sweep launch shape on this world (advances it by a few steps per
        candidate); returns (groups, lds_kib, policy, sweep_ms)."""
        c = SweepConfig()
        _check(self._L.axm_world_autotune(self._h, dt, iters, steps, C.byref(c)), "axm_world_autotune")
        return c.groups, c.lds_kib, c.policy, c.sweep_ms

    def set_feature_logging(self, enable=True):
        _check(self._L.axm_world_set_feature_logging(self._h, 1 if enable else 0), "axm_world_set_feature_logging")

    def features(self):
        """(X [M,4,24], Y [M,4,3]) network inputs and targets ((final - previous impulse) / u) of the
        last step, manifold order."""
        cnt = C.c_uint32()
        code = self._L.axm_world_get_features(self._h, None, None, 0, C.byref(cnt))
        if code not in (0, 601):
            _check(code, "axm_world_get_features")
        M = cnt.value
        X = np.zeros((M, 4, MLP_INPUTS), np.float32)
        Y = np.zeros((M, 4, MLP_OUTPUTS), np.float32)
        _check(self._L.axm_world_get_features(self._h, _p(X), _p(Y), M, C.byref(cnt)), "axm_world_get_features")
        return X, Y

    def broadphase(self):
        c = C.c_uint32()
        _check(self._L.axm_broadphase(self._h, C.byref(c)), "axm_broadphase")
        return c.value

    def collide(self):
        c = C.c_uint32()
        _check(self._L.axm_collide(self._h, C.byref(c)), "axm_collide")
        return c.value

    def pairs(self):
        c = C.c_uint32()
        self._L.axm_world_get_pairs(self._h, None, None, 0, C.byref(c))
        a = np.zeros(c.value, np.uint32)
        b = np.zeros(c.value, np.uint32)
        _check(self._L.axm_world_get_pairs(self._h, _p(a, C.c_uint32), _p(b, C.c_uint32), c.value, C.byref(c)),
               "axm_world_get_pairs")
        return a, b

    def manifolds(self):
        c = C.c_uint32()
        self._L.axm_world_get_manifolds(self._h, None, None, None, None, None, None, None, 0, C.byref(c))
        M = c.value
        out = dict(a=np.zeros(M, np.uint32), b=np.zeros(M, np.uint32), npts=np.zeros(M, np.uint32),
                   pos=np.zeros((M, 4, 3), np.float32), normal=np.zeros((M, 4, 3), np.float32),
                   depth=np.zeros((M, 4), np.float32), feature=np.zeros((M, 4), np.uint32))
        _check(self._L.axm_world_get_manifolds(
            self._h, _p(out["a"], C.c_uint32), _p(out["b"], C.c_uint32), _p(out["npts"], C.c_uint32),
            _p(out["pos"]), _p(out["normal"]), _p(out["depth"]), _p(out["feature"], C.c_uint32), M, C.byref(c)),
            "axm_world_get_manifolds")
        return out

    def colours(self):
        c = C.c_uint32()
        k = C.c_uint32()
        self._L.axm_world_get_colours(self._h, None, None, 0, C.byref(c), C.byref(k))
        col = np.zeros(c.value, np.uint32)
        order = np.zeros(c.value, np.uint32)
        _check(self._L.axm_world_get_colours(self._h, _p(col, C.c_uint32), _p(order, C.c_uint32), c.value,
                                             C.byref(c), C.byref(k)), "axm_world_get_colours")
        return col, order, k.value

    def friction_bounds(self):
        """Friction bound mu * B of every contact point of the last step, [M, 4] in manifold order."""
        c = C.c_uint32()
        self._L.axm_world_get_friction_bounds(self._h, None, 0, C.byref(c))
        out = np.zeros((c.value, 4), np.float32)
        _check(self._L.axm_world_get_friction_bounds(self._h, _p(out), c.value, C.byref(c)),
               "axm_world_get_friction_bounds")
        return out

    def impulses(self):
        c = C.c_uint32()
        self._L.axm_world_get_impulses(self._h, None, None, 0, C.byref(c))
        i0 = np.zeros((c.value, 4, 3), np.float32)
        i1 = np.zeros((c.value, 4, 3), np.float32)
        _check(self._L.axm_world_get_impulses(self._h, _p(i0), _p(i1), c.value, C.byref(c)), "axm_world_get_impulses")
        return i0, i1

    def residuals(self):
        c = C.c_uint32()
        self._L.axm_world_get_residuals(self._h, None, 0, C.byref(c))
        r = np.zeros(c.value, np.float32)
        _check(self._L.axm_world_get_residuals(self._h, _p(r), c.value, C.byref(c)), "axm_world_get_residuals")
        return r

    def pre_solve_velocities(self):
        v = np.zeros((self.n, 3), np.float32)
        w = np.zeros((self.n, 3), np.float32)
        _check(self._L.axm_world_get_pre_solve_velocities(self._h, _p(v), _p(w)), "get_pre_solve_velocities")
        return v, w


def collide_pairs(ta, pa, xa, qa, tb, pb, xb, qb, margin=0.02, device=0, noalloc=False):
    """Batched narrowphase on the GPU; returns dict of per-pair manifolds.  noalloc: the entry point
    that stages through the library's fixed device buffer (at most 16384 pairs, no allocation)."""
    n = len(ta)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in
            ((ta, np.uint32), (pa, np.float32), (xa, np.float32), (qa, np.float32),
             (tb, np.uint32), (pb, np.float32), (xb, np.float32), (qb, np.float32))]
    out = dict(npts=np.zeros(n, np.uint32), pos=np.zeros((n, 4, 3), np.float32), normal=np.zeros((n, 4, 3), np.float32),
               depth=np.zeros((n, 4), np.float32), feature=np.zeros((n, 4), np.uint32))
    fn = lib().axm_collide_pairs_noalloc if noalloc else lib().axm_collide_pairs
    code = fn(
        device, n, _p(arrs[0], C.c_uint32), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), _p(arrs[4], C.c_uint32),
        _p(arrs[5]), _p(arrs[6]), _p(arrs[7]), margin, _p(out["npts"], C.c_uint32), _p(out["pos"]),
        _p(out["normal"]), _p(out["depth"]), _p(out["feature"], C.c_uint32))
    out["status"] = code
    return out


def gjk_epa_pairs(ta, pa, xa, qa, tb, pb, xb, qb, device=0, gjk_iters=0, epa_iters=0):
    """Batched GJK (+ EPA) on shape cores: (out [n, 10], status [n]); gjk_iters / epa_iters cap the
    iterations (0: the defaults)."""
    n = len(ta)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in
            ((ta, np.uint32), (pa, np.float32), (xa, np.float32), (qa, np.float32),
             (tb, np.uint32), (pb, np.float32), (xb, np.float32), (qb, np.float32))]
    out = np.zeros((n, 10), np.float32)
    st = np.zeros(n, np.int32)
    _check(lib().axm_gjk_epa_pairs_capped(device, n, _p(arrs[0], C.c_uint32), _p(arrs[1]), _p(arrs[2]),
                                          _p(arrs[3]), _p(arrs[4], C.c_uint32), _p(arrs[5]), _p(arrs[6]),
                                          _p(arrs[7]), gjk_iters, epa_iters, _p(out), _p(st, C.c_int32)),
           "axm_gjk_epa_pairs_capped")
    return out, st
