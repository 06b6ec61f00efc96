"""ctypes binding of the CPU oracle (oracle/liborc.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() (as the checker) and by the CPU
baseline leg of bench.py.  The product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
MAX_PLANES = 8


class Params(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3), ("friction", C.c_float), ("restitution", C.c_float),
        ("rest_threshold", C.c_float), ("baumgarte", C.c_float), ("slop", C.c_float), ("max_bias", C.c_float),
        ("margin", C.c_float), ("lin_damp", C.c_float), ("ang_damp", C.c_float), ("cell_size", C.c_float),
        ("origin", C.c_float * 3), ("num_planes", C.c_uint32), ("planes", (C.c_float * 4) * MAX_PLANES),
        ("warm_start", C.c_uint32), ("record_residual", C.c_uint32), ("softness", C.c_float),
        ("allow_sleep", C.c_uint32), ("sleep_linear", C.c_float), ("sleep_angular", C.c_float),
        ("sleep_time", C.c_float), ("island_solve", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("pairs", C.c_uint32), ("manifolds", C.c_uint32), ("contacts", C.c_uint32), ("colours", C.c_uint32),
        ("rounds", C.c_uint32), ("iterations", C.c_uint32), ("residual", C.c_float), ("residual_first", C.c_float),
        ("kinetic", C.c_double), ("potential", C.c_double), ("gjk_fail", C.c_uint32), ("epa_fail", C.c_uint32),
        ("warm_hits", C.c_uint32), ("islands", C.c_uint32), ("sleeping", C.c_uint32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32p, f32p, f64p, u64p = (C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_double),
                              C.POINTER(C.c_uint64))
    L.orc_params_size.restype = C.c_uint32
    L.orc_stats_size.restype = C.c_uint32
    assert L.orc_params_size() == C.sizeof(Params), "oracle Params layout mismatch"
    assert L.orc_stats_size() == C.sizeof(Stats), "oracle Stats layout mismatch"
    L.orc_create.restype = P
    L.orc_create.argtypes = [C.POINTER(Params), C.c_int]
    L.orc_destroy.argtypes = [P]
    L.orc_set_threads.argtypes = [P, C.c_uint32]
    L.orc_set_body_extras.argtypes = [P, f32p, u32p, u32p]
    L.orc_get_islands.argtypes = [P, u32p, u32p, f32p]
    L.orc_set_tolerance.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32]
    L.orc_set_params.argtypes = [P, C.POINTER(Params)]
    L.orc_set_bodies.argtypes = [P, C.c_uint32, f32p, f32p, f32p, f32p, f32p, f32p, u32p, f32p, u32p]
    L.orc_get_bodies.argtypes = [P, f32p, f32p, f32p, f32p]
    L.orc_get_pre_solve_velocities.argtypes = [P, f32p, f32p]
    L.orc_step.argtypes = [P, C.c_float, C.c_uint32, C.POINTER(Stats)]
    L.orc_run.restype = C.c_double
    L.orc_run.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32]
    L.orc_run_many.restype = C.c_double
    L.orc_run_many.argtypes = [C.POINTER(P), C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_broadphase.restype = C.c_uint32
    L.orc_broadphase.argtypes = [P]
    L.orc_collide_world.restype = C.c_uint32
    L.orc_collide_world.argtypes = [P]
    L.orc_get_pairs.restype = C.c_uint32
    L.orc_get_pairs.argtypes = [P, u32p, u32p, C.c_uint32]
    L.orc_get_keys.argtypes = [P, u64p, u32p]
    L.orc_brute_pairs.restype = C.c_uint32
    L.orc_brute_pairs.argtypes = [P, u32p, u32p, C.c_uint32]
    L.orc_get_manifolds.restype = C.c_uint32
    L.orc_get_manifolds.argtypes = [P, u32p, u32p, u32p, f32p, f32p, f32p, u32p, C.c_uint32]
    L.orc_get_colours.restype = C.c_uint32
    L.orc_get_colours.argtypes = [P, u32p, u32p, C.c_uint32]
    L.orc_get_lambda.argtypes = [P, f32p, C.c_uint32]
    L.orc_get_residuals.restype = C.c_uint32
    L.orc_get_residuals.argtypes = [P, f32p, C.c_uint32]
    L.orc_step_given.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, u32p, u32p, u32p, f32p, f32p, f32p, u32p, f32p,
                                 u32p, f32p, C.POINTER(Stats)]
    L.orc_get_friction_bounds.restype = C.c_uint32
    L.orc_get_friction_bounds.argtypes = [P, f32p, C.c_uint32]
    L.orc_collide.restype = C.c_int
    L.orc_collide.argtypes = [C.c_int, C.c_uint32, f32p, f32p, f32p, C.c_uint32, f32p, f32p, f32p, C.c_float, u32p,
                              f32p, f32p, f32p, u32p]
    L.orc_collide_plane.argtypes = [C.c_int, C.c_uint32, f32p, f32p, f32p, f32p, C.c_float, u32p, f32p, f32p, f32p,
                                    u32p]
    L.orc_gjk_epa.restype = C.c_int
    L.orc_gjk_epa.argtypes = [C.c_int, C.c_uint32, f32p, f32p, f32p, C.c_uint32, f32p, f32p, f32p, f32p]
    L.orc_gjk_epa_capped.restype = C.c_int
    L.orc_gjk_epa_capped.argtypes = [C.c_int, C.c_uint32, f32p, f32p, f32p, C.c_uint32, f32p, f32p, f32p, C.c_int,
                                     C.c_int, f32p]
    L.orc_mlp_nparam.restype = C.c_uint32
    L.orc_mlp_forward.argtypes = [f32p, f32p, C.c_int, C.c_int, f32p]
    L.orc_mlp_grad.restype = C.c_double
    L.orc_mlp_grad.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, f64p]
    L.orc_mlp_loss.restype = C.c_double
    L.orc_mlp_loss.argtypes = [f32p, f32p, f32p, C.c_int]
    L.orc_adam.argtypes = [f32p, f32p, f32p, f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_bf16_round.argtypes = [f32p, f32p, C.c_int]
    L.orc_get_features.restype = C.c_uint32
    L.orc_get_features.argtypes = [P, f32p, C.c_uint32]
    L.orc_quat_mul.argtypes = [f32p, f32p, f32p]
    L.orc_quat_rotate.argtypes = [f32p, f32p, f32p]
    L.orc_world_inv_inertia.argtypes = [f32p, f32p, f32p]
    L.orc_tangents.argtypes = [f32p, f32p, f32p]
    L.orc_morton3.restype = C.c_uint32
    L.orc_morton3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_pair_hash24.restype = C.c_uint32
    L.orc_pair_hash24.argtypes = [C.c_uint32, C.c_uint32]
    _lib = L
    return L


def _p(a, ct=C.c_float):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def params_from_desc(desc):
    """Copies the shared leading fields of a WorldDesc into oracle Params."""
    p = Params()
    for name, _ in Params._fields_:
        src = getattr(desc, name)
        if name == "planes":
            for k in range(MAX_PLANES):
                for c in range(4):
                    p.planes[k][c] = src[k][c]
        elif name in ("gravity", "origin"):
            for k in range(3):
                getattr(p, name)[k] = src[k]
        else:
            setattr(p, name, src)
    return p


class OracleWorld:
    """CPU oracle world (float or double)."""

    def __init__(self, params, fp64=False):
        self.L = lib()
        self.params = params
        self.h = self.L.orc_create(C.byref(params), 1 if fp64 else 0)
        self.n = 0

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def islands(self):
        """(label, asleep, rest time) of every body after the last step."""
        lab = np.zeros(self.n, np.uint32)
        sl = np.zeros(self.n, np.uint32)
        rest = np.zeros(self.n, np.float32)
        self.L.orc_get_islands(self.h, _p(lab, C.c_uint32), _p(sl, C.c_uint32), _p(rest))
        return lab, sl, rest

    def set_threads(self, threads):
        self.L.orc_set_threads(self.h, int(threads))

    def set_tolerance(self, tol, max_iters, every=1):
        self.L.orc_set_tolerance(self.h, float(tol), int(max_iters), int(every))

    def set_params(self, params):
        self.params = params
        self.L.orc_set_params(self.h, C.byref(params))

    def set_bodies(self, b):
        self.n = b.n
        self._keep = b
        self.L.orc_set_bodies(self.h, b.n, _p(b.pos), _p(b.quat), _p(b.lin_vel), _p(b.ang_vel), _p(b.inv_mass),
                              _p(b.inv_inertia), _p(b.shape_type, C.c_uint32), _p(b.shape_params),
                              _p(b.world, C.c_uint32))
        mat, fl, fi = getattr(b, "material", None), getattr(b, "flags", None), getattr(b, "filter", None)
        if mat is not None or fl is not None or fi is not None:
            self._extras = [None if a is None else np.ascontiguousarray(a) for a in (mat, fl, fi)]
            self.L.orc_set_body_extras(self.h, _p(self._extras[0]), _p(self._extras[1], C.c_uint32),
                                       _p(self._extras[2], C.c_uint32))

    def get_state(self):
        pos = np.zeros((self.n, 3), np.float32)
        quat = np.zeros((self.n, 4), np.float32)
        v = np.zeros((self.n, 3), np.float32)
        w = np.zeros((self.n, 3), np.float32)
        self.L.orc_get_bodies(self.h, _p(pos), _p(quat), _p(v), _p(w))
        return pos, quat, v, w

    def pre_solve_velocities(self):
        v = np.zeros((self.n, 3), np.float32)
        w = np.zeros((self.n, 3), np.float32)
        self.L.orc_get_pre_solve_velocities(self.h, _p(v), _p(w))
        return v, w

    def step(self, dt, iters):
        st = Stats()
        self.L.orc_step(self.h, dt, iters, C.byref(st))
        return st

    def run(self, dt, iters, steps):
        return self.L.orc_run(self.h, dt, iters, steps)

    def broadphase(self):
        return self.L.orc_broadphase(self.h)

    def collide(self):
        return self.L.orc_collide_world(self.h)

    def pairs(self):
        n = self.L.orc_get_pairs(self.h, None, None, 0)
        a = np.zeros(n, np.uint32)
        b = np.zeros(n, np.uint32)
        self.L.orc_get_pairs(self.h, _p(a, C.c_uint32), _p(b, C.c_uint32), n)
        return a, b

    def keys(self):
        k = np.zeros(self.n, np.uint64)
        o = np.zeros(self.n, np.uint32)
        self.L.orc_get_keys(self.h, _p(k, C.c_uint64), _p(o, C.c_uint32))
        return k, o

    def brute_pairs(self):
        n = self.L.orc_brute_pairs(self.h, None, None, 0)
        a = np.zeros(n, np.uint32)
        b = np.zeros(n, np.uint32)
        self.L.orc_brute_pairs(self.h, _p(a, C.c_uint32), _p(b, C.c_uint32), n)
        return a, b

    def manifolds(self):
        M = self.L.orc_get_manifolds(self.h, None, None, None, None, None, None, None, 0)
        out = dict(a=np.zeros(M, np.uint32), b=np.zeros(M, np.uint32), npts=np.zeros(M, np.uint32),
                   pos=np.zeros((M, 4, 3), np.float32), normal=np.zeros((M, 4, 3), np.float32),
                   depth=np.zeros((M, 4), np.float32), feature=np.zeros((M, 4), np.uint32))
        self.L.orc_get_manifolds(self.h, _p(out["a"], C.c_uint32), _p(out["b"], C.c_uint32),
                                 _p(out["npts"], C.c_uint32), _p(out["pos"]), _p(out["normal"]), _p(out["depth"]),
                                 _p(out["feature"], C.c_uint32), M)
        return out

    def colours(self):
        M = self.L.orc_get_manifolds(self.h, None, None, None, None, None, None, None, 0)
        col = np.zeros(M, np.uint32)
        order = np.zeros(M, np.uint32)
        K = self.L.orc_get_colours(self.h, _p(col, C.c_uint32), _p(order, C.c_uint32), M)
        return col, order, K

    def impulses(self):
        M = self.L.orc_get_manifolds(self.h, None, None, None, None, None, None, None, 0)
        lam = np.zeros((M, 4, 3), np.float32)
        self.L.orc_get_lambda(self.h, _p(lam), M)
        return lam

    def features(self):
        """Warm-start network inputs of the last step, [M, 4, 24] in manifold order."""
        M = self.L.orc_get_features(self.h, None, 0)
        X = np.zeros((M, 4, 24), np.float32)
        self.L.orc_get_features(self.h, _p(X), M)
        return X

    def residuals(self):
        n = self.L.orc_get_residuals(self.h, None, 0)
        r = np.zeros(n, np.float32)
        self.L.orc_get_residuals(self.h, _p(r), n)
        return r

    def step_given(self, dt, iters, man, warm=None, colours=None, bounds=None):
        """One step on given manifolds (optionally with a given warm start, colouring and friction
        bounds [M, 4])."""
        st = Stats()
        M = len(man["a"])
        arr = {k: np.ascontiguousarray(v) for k, v in man.items()}
        w = np.ascontiguousarray(warm, np.float32) if warm is not None else None
        c = np.ascontiguousarray(colours, np.uint32) if colours is not None else None
        fb = np.ascontiguousarray(bounds, np.float32) if bounds is not None else None
        self.L.orc_step_given(self.h, dt, iters, M, _p(arr["a"], C.c_uint32), _p(arr["b"], C.c_uint32),
                              _p(arr["npts"], C.c_uint32), _p(arr["pos"]), _p(arr["normal"]), _p(arr["depth"]),
                              _p(arr["feature"], C.c_uint32), _p(w), _p(c, C.c_uint32), _p(fb), C.byref(st))
        return st

    def friction_bounds(self):
        """Friction bound of every contact point of the last step, [M, 4] in manifold order."""
        M = self.L.orc_get_friction_bounds(self.h, None, 0)
        out = np.zeros((M, 4), np.float32)
        self.L.orc_get_friction_bounds(self.h, _p(out), M)
        return out


def collide(ta, pa, xa, qa, tb, pb, xb, qb, margin=0.02, fp64=False):
    """One shape pair on the CPU oracle. tb == 3 (plane): pb = normal, xb[0] = offset."""
    L = lib()
    f = lambda x: np.ascontiguousarray(x, np.float32)
    np_ = C.c_uint32()
    pos = np.zeros((4, 3), np.float32)
    nrm = np.zeros((4, 3), np.float32)
    dep = np.zeros(4, np.float32)
    fid = np.zeros(4, np.uint32)
    pa, xa, qa = f(pa), f(xa), f(qa)
    if tb == 3:
        plane = np.array([pb[0], pb[1], pb[2], xb[0]], np.float32)
        L.orc_collide_plane(1 if fp64 else 0, ta, _p(pa), _p(xa), _p(qa), _p(plane), margin, C.byref(np_), _p(pos),
                            _p(nrm), _p(dep), _p(fid, C.c_uint32))
        status = 0
    else:
        pb, xb, qb = f(pb), f(xb), f(qb)
        status = L.orc_collide(1 if fp64 else 0, ta, _p(pa), _p(xa), _p(qa), tb, _p(pb), _p(xb), _p(qb), margin,
                               C.byref(np_), _p(pos), _p(nrm), _p(dep), _p(fid, C.c_uint32))
    k = np_.value
    return dict(npts=k, pos=pos[:k], normal=nrm[:k], depth=dep[:k], feature=fid[:k], status=status)


def gjk_epa(ta, pa, xa, qa, tb, pb, xb, qb, fp64=False, gjk_iters=32, epa_iters=48):
    L = lib()
    f = lambda x: np.ascontiguousarray(x, np.float32)
    out = np.zeros(10, np.float32)
    st = L.orc_gjk_epa_capped(1 if fp64 else 0, ta, _p(f(pa)), _p(f(xa)), _p(f(qa)), tb, _p(f(pb)), _p(f(xb)),
                              _p(f(qb)), gjk_iters, epa_iters, _p(out))
    return out, st


def mlp_forward(prm, X, emu=True):
    X = np.ascontiguousarray(X, np.float32)
    Y = np.zeros((X.shape[0], 3), np.float32)
    lib().orc_mlp_forward(_p(np.ascontiguousarray(prm, np.float32)), _p(X), X.shape[0], 1 if emu else 0, _p(Y))
    return Y


def mlp_grad(prm, X, Y, emu=True, fp64=False):
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    g = np.zeros(lib().orc_mlp_nparam(), np.float64)
    loss = lib().orc_mlp_grad(_p(np.ascontiguousarray(prm, np.float32)), _p(X), _p(Y), X.shape[0], 1 if emu else 0,
                              1 if fp64 else 0, _p(g, C.c_double))
    return loss, g


def mlp_loss(prm, X, Y):
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    return lib().orc_mlp_loss(_p(np.ascontiguousarray(prm, np.float32)), _p(X), _p(Y), X.shape[0])


def adam(prm, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    lib().orc_adam(_p(prm), _p(m), _p(v), _p(np.ascontiguousarray(g, np.float32)), prm.size, t, lr, b1, b2, eps)


def _f32(x):
    return np.ascontiguousarray(x, np.float32)


def quat_mul(a, b):
    out = np.zeros(4, np.float32)
    lib().orc_quat_mul(_p(_f32(a)), _p(_f32(b)), _p(out))
    return out


def quat_rotate(q, v):
    out = np.zeros(3, np.float32)
    lib().orc_quat_rotate(_p(_f32(q)), _p(_f32(v)), _p(out))
    return out


def world_inv_inertia(q, d):
    out = np.zeros(6, np.float32)
    lib().orc_world_inv_inertia(_p(_f32(q)), _p(_f32(d)), _p(out))
    return out


def tangents(n):
    t1 = np.zeros(3, np.float32)
    t2 = np.zeros(3, np.float32)
    lib().orc_tangents(_p(_f32(n)), _p(t1), _p(t2))
    return t1, t2


def morton3(x, y, z):
    return lib().orc_morton3(x, y, z)


def pair_hash24(a, b):
    return lib().orc_pair_hash24(a, b)


def bf16_round(x):
    x = _f32(x)
    y = np.zeros_like(x)
    lib().orc_bf16_round(_p(x), _p(y), x.size)
    return y
