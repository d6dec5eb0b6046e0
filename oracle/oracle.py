"""ctypes wrapper of liboracle.so (sph_oracle.c), the CPU FP64 restatement of the reference
step.  TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline, never as the
product.  All arrays are numpy float64 (N, 3) / (N,) in particle order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

__all__ = [
    "build", "lib", "OracleConsts", "OracleParams", "consts_of", "derived", "step", "naive_density",
    "naive_pressure", "naive_viscosity", "grid_naive_terms", "voxel_keys", "space_dims", "sort_map",
    "voxel_begin", "voxel_neighbours", "rng_init", "rng_next", "rng_jump", "rng_uniform", "splitmix64",
    "find_segment", "x_at_segment_beginning", "radius_in_position", "is_out_of_pipe", "vector_length",
    "distance_between_points", "solve_collision", "collide_box", "collide_pipe", "integrate", "w", "grad_w",
    "lap_w", "set_threads", "num_threads", "voxel_forces", "rollout", "rollout_adjoint", "adjoint_step",
]


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "sph_oracle.c")):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    return LIB


class OracleConsts(ctypes.Structure):
    _fields_ = [("mass", c_double), ("rho0", c_double), ("h", c_double), ("visc", c_double), ("k", c_double),
                ("damp", c_double)]


class OracleParams(ctypes.Structure):
    _fields_ = [("n", c_int64), ("ext", c_double * 3), ("dt", c_double), ("space", c_double * 3),
                ("voxel", c_double * 3), ("mode", c_int32), ("formulation", c_int32), ("pipe_rows", c_int32),
                ("use_grid", c_int32), ("pipe", POINTER(c_double))]


_lib = None
D = POINTER(c_double)
I32 = POINTER(c_int32)
U64 = POINTER(c_uint64)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "oc_derived_constants": ([POINTER(OracleConsts), D], None),
            "oc_vector_length": ([D], c_double),
            "oc_distance_between_points": ([D, D], c_double),
            "oc_x_at_segment_beginning": ([D, c_int], c_double),
            "oc_radius_in_position": ([D, D, c_int], c_double),
            "oc_is_out_of_pipe": ([D, D, c_int], c_int),
            "oc_find_segment": ([D, D, c_int], c_int),
            "oc_solve_collision": ([D, D, D, c_int], None),
            "oc_splitmix64": ([U64], c_uint64),
            "oc_rng_next": ([U64], c_uint64),
            "oc_rng_jump": ([U64], None),
            "oc_rng_uniform": ([U64], c_double),
            "oc_rng_init": ([c_uint64, c_int64, U64], None),
            "oc_collide_pipe": ([D, D, c_int64, D, c_int, U64], None),
            "oc_collide_box": ([D, D, c_int64, D, c_double], None),
            "oc_integrate": ([D, D, D, D, D, D, D, c_int64, c_double], None),
            "oc_w": ([POINTER(OracleConsts), D, D], c_double),
            "oc_grad_w": ([POINTER(OracleConsts), D, D, D], None),
            "oc_lap_w": ([POINTER(OracleConsts), D, D], c_double),
            "oc_naive_density": ([D, c_int64, POINTER(OracleConsts), D], None),
            "oc_naive_pressure": ([D, D, c_int64, POINTER(OracleConsts), D], None),
            "oc_naive_viscosity": ([D, D, D, c_int64, POINTER(OracleConsts), D], None),
            "oc_space_dims": ([D, D, c_int, I32], None),
            "oc_voxel_keys": ([D, c_int64, D, I32, I32], None),
            "oc_sort_map": ([I32, c_int64, I32], None),
            "oc_voxel_begin": ([I32, c_int64, c_int64, I32], None),
            "oc_voxel_neighbours": ([D, c_int64, I32, c_int64, I32, D, I32, c_double, c_int64, I32], c_int),
            "oc_voxel_forces": ([D, D, c_int64, I32, I32, POINTER(OracleConsts), D, D, D], None),
            "oc_step": ([POINTER(OracleParams), POINTER(OracleConsts), D, D, D, D, U64], c_int),
            "oc_grid_naive_terms": ([D, D, c_int64, POINTER(OracleConsts), D, D, D], None),
            "oc_rollout": ([POINTER(OracleParams), POINTER(OracleConsts), D, D, c_int, D, D, c_void_p, c_int], c_int),
            "oc_adjoint_step": ([POINTER(OracleParams), POINTER(OracleConsts), D, D, D, D, D, D, D], c_int),
            "oc_rollout_adjoint": ([POINTER(OracleParams), POINTER(OracleConsts), D, D, c_int, D, c_int, D, D, D],
                                   c_int),
            "oc_num_threads": ([], c_int),
            "oc_set_num_threads": ([c_int], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I32)


def _u(a):
    return a.ctypes.data_as(U64)


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a.reshape(shape) if shape is not None else a


def consts_of(c) -> OracleConsts:
    o = OracleConsts()
    o.mass, o.rho0, o.h, o.visc, o.k, o.damp = (float(c.mass), float(c.rho0), float(c.h), float(c.visc),
                                                float(c.k), float(c.damp))
    return o


def derived(c) -> np.ndarray:
    out = np.zeros(4)
    lib().oc_derived_constants(ctypes.byref(consts_of(c)), _d(out))
    return out


def set_threads(t: int) -> None:
    lib().oc_set_num_threads(int(t))


def num_threads() -> int:
    return int(lib().oc_num_threads())


# -- step -------------------------------------------------------------------------------
def step(pos, vel, *, consts, ext, dt, space, voxel, mode="BOX", formulation="NAIVE", pipe=None, rng=None,
         use_grid=False):
    """One reference step.  Returns (pos, vel, rho, force, rng) as new arrays."""
    pos = _f64(pos, (-1, 3)).copy()
    vel = _f64(vel, (-1, 3)).copy()
    n = pos.shape[0]
    p = OracleParams()
    p.n = n
    p.ext[:] = [float(v) for v in ext]
    p.dt = float(dt)
    p.space[:] = [float(v) for v in space]
    p.voxel[:] = [float(v) for v in voxel]
    p.mode = 1 if mode == "PIPE" else 0
    p.formulation = 1 if formulation == "VOXEL" else 0
    p.use_grid = 1 if use_grid else 0
    tbl = None
    if mode == "PIPE":
        tbl = _f64(pipe, (-1, 5))
        p.pipe_rows = tbl.shape[0]
        p.pipe = _d(tbl)
        rng = np.ascontiguousarray(np.asarray(rng, dtype=np.uint64)).copy()
    else:
        p.pipe_rows = 0
        p.pipe = None
        rng = np.zeros(2, dtype=np.uint64) if rng is None else np.ascontiguousarray(rng, dtype=np.uint64).copy()
    rho = np.zeros(n)
    force = np.zeros((n, 3))
    rc = lib().oc_step(ctypes.byref(p), ctypes.byref(consts_of(consts)), _d(pos), _d(vel), _d(rho), _d(force),
                       _u(rng))
    if rc != 0:
        raise RuntimeError("oracle step failed")
    return pos, vel, rho, force, rng


# -- stages -----------------------------------------------------------------------------
def naive_density(pos, consts):
    pos = _f64(pos, (-1, 3))
    out = np.zeros(pos.shape[0])
    lib().oc_naive_density(_d(pos), pos.shape[0], ctypes.byref(consts_of(consts)), _d(out))
    return out


def naive_pressure(pos, rho, consts):
    pos, rho = _f64(pos, (-1, 3)), _f64(rho)
    out = np.zeros_like(pos)
    lib().oc_naive_pressure(_d(pos), _d(rho), pos.shape[0], ctypes.byref(consts_of(consts)), _d(out))
    return out


def naive_viscosity(pos, vel, rho, consts):
    pos, vel, rho = _f64(pos, (-1, 3)), _f64(vel, (-1, 3)), _f64(rho)
    out = np.zeros_like(pos)
    lib().oc_naive_viscosity(_d(pos), _d(vel), _d(rho), pos.shape[0], ctypes.byref(consts_of(consts)), _d(out))
    return out


def grid_naive_terms(pos, vel, consts):
    """rho, P, V of the NAIVE formulation through the grid search (same neighbour sets)."""
    pos, vel = _f64(pos, (-1, 3)), _f64(vel, (-1, 3))
    n = pos.shape[0]
    rho, P, V = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
    lib().oc_grid_naive_terms(_d(pos), _d(vel), n, ctypes.byref(consts_of(consts)), _d(rho), _d(P), _d(V))
    return rho, P, V


def integrate(pos, vel, ext, rho, P, V, dt):
    pos, vel = _f64(pos, (-1, 3)).copy(), _f64(vel, (-1, 3)).copy()
    ext, rho, P, V = _f64(ext), _f64(rho), _f64(P, (-1, 3)), _f64(V, (-1, 3))
    force = np.zeros_like(pos)
    lib().oc_integrate(_d(force), _d(pos), _d(vel), _d(ext), _d(rho), _d(P), _d(V), pos.shape[0], float(dt))
    return pos, vel, force


def collide_box(pos, vel, space, damp):
    pos, vel = _f64(pos, (-1, 3)).copy(), _f64(vel, (-1, 3)).copy()
    lib().oc_collide_box(_d(pos), _d(vel), pos.shape[0], _d(_f64(space)), float(damp))
    return pos, vel


def collide_pipe(pos, vel, pipe, rng):
    pos, vel = _f64(pos, (-1, 3)).copy(), _f64(vel, (-1, 3)).copy()
    tbl = _f64(pipe, (-1, 5))
    rng = np.ascontiguousarray(np.asarray(rng, dtype=np.uint64)).copy()
    lib().oc_collide_pipe(_d(pos), _d(vel), pos.shape[0], _d(tbl), tbl.shape[0], _u(rng))
    return pos, vel, rng


# -- voxel keys, sort, tables -----------------------------------------------------------
def space_dims(space, voxel, use_ceil=True):
    out = np.zeros(3, dtype=np.int32)
    lib().oc_space_dims(_d(_f64(space)), _d(_f64(voxel)), 1 if use_ceil else 0, _i(out))
    return out


def voxel_keys(pos, voxel, dims):
    pos = _f64(pos, (-1, 3))
    out = np.zeros(pos.shape[0], dtype=np.int32)
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    lib().oc_voxel_keys(_d(pos), pos.shape[0], _d(_f64(voxel)), _i(dims), _i(out))
    return out


def sort_map(keys):
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    out = np.zeros(keys.shape[0], dtype=np.int32)
    lib().oc_sort_map(_i(keys), keys.shape[0], _i(out))
    return out


def voxel_begin(sorted_keys, n_voxels):
    sk = np.ascontiguousarray(sorted_keys, dtype=np.int32)
    out = np.zeros(int(n_voxels), dtype=np.int32)
    lib().oc_voxel_begin(_i(sk), sk.shape[0], int(n_voxels), _i(out))
    return out


def voxel_neighbours(pos, begin, map_pid, voxel, tdims, h, p):
    pos = _f64(pos, (-1, 3))
    begin = np.ascontiguousarray(begin, dtype=np.int32)
    mp = np.ascontiguousarray(map_pid, dtype=np.int32)
    out = np.zeros(32, dtype=np.int32)
    tdims = np.ascontiguousarray(tdims, dtype=np.int32)
    k = lib().oc_voxel_neighbours(_d(pos), pos.shape[0], _i(begin), begin.shape[0], _i(mp), _d(_f64(voxel)),
                                  _i(tdims), float(h), int(p), _i(out))
    return out[:k].copy()


def voxel_forces(pos, vel, nb, nn, consts):
    pos, vel = _f64(pos, (-1, 3)), _f64(vel, (-1, 3))
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    nn = np.ascontiguousarray(nn, dtype=np.int32)
    n = pos.shape[0]
    rho, P, V = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
    lib().oc_voxel_forces(_d(pos), _d(vel), n, _i(nb), _i(nn), ctypes.byref(consts_of(consts)), _d(rho), _d(P), _d(V))
    return rho, P, V


# -- RNG --------------------------------------------------------------------------------
def splitmix64(x: int):
    st = (c_uint64 * 1)(x)
    v = lib().oc_splitmix64(st)
    return int(v), int(st[0])


def rng_init(seed: int, n: int) -> np.ndarray:
    out = np.zeros(2 * max(n, 1), dtype=np.uint64)
    lib().oc_rng_init(int(seed), int(n), _u(out))
    return out[: 2 * n]


def rng_next(state: np.ndarray) -> int:
    return int(lib().oc_rng_next(_u(state)))


def rng_jump(state: np.ndarray) -> None:
    lib().oc_rng_jump(_u(state))


def rng_uniform(state: np.ndarray) -> float:
    return float(lib().oc_rng_uniform(_u(state)))


# -- pipe geometry ----------------------------------------------------------------------
def find_segment(pos3, pipe) -> int:
    tbl = _f64(pipe, (-1, 5))
    return int(lib().oc_find_segment(_d(_f64(pos3)), _d(tbl), tbl.shape[0]))


def x_at_segment_beginning(pipe, seg) -> float:
    return float(lib().oc_x_at_segment_beginning(_d(_f64(pipe, (-1, 5))), int(seg)))


def radius_in_position(pos3, pipe, seg) -> float:
    return float(lib().oc_radius_in_position(_d(_f64(pos3)), _d(_f64(pipe, (-1, 5))), int(seg)))


def is_out_of_pipe(pos3, pipe, seg) -> bool:
    return bool(lib().oc_is_out_of_pipe(_d(_f64(pos3)), _d(_f64(pipe, (-1, 5))), int(seg)))


def vector_length(v) -> float:
    return float(lib().oc_vector_length(_d(_f64(v))))


def distance_between_points(a, b) -> float:
    return float(lib().oc_distance_between_points(_d(_f64(a)), _d(_f64(b))))


def solve_collision(pos3, vel3, pipe, seg):
    p, v = _f64(pos3).copy(), _f64(vel3).copy()
    lib().oc_solve_collision(_d(p), _d(v), _d(_f64(pipe, (-1, 5))), int(seg))
    return p, v


def w(consts, a, b) -> float:
    return float(lib().oc_w(ctypes.byref(consts_of(consts)), _d(_f64(a)), _d(_f64(b))))


def grad_w(consts, a, b):
    out = np.zeros(3)
    lib().oc_grad_w(ctypes.byref(consts_of(consts)), _d(_f64(a)), _d(_f64(b)), _d(out))
    return out


def lap_w(consts, a, b) -> float:
    return float(lib().oc_lap_w(ctypes.byref(consts_of(consts)), _d(_f64(a)), _d(_f64(b))))


# -- rollout and its adjoint (NAIVE, box walls) ------------------------------------------
def _box_params(n, ext, dt, space) -> OracleParams:
    p = OracleParams()
    p.n = n
    p.ext[:] = [float(v) for v in ext]
    p.dt = float(dt)
    p.space[:] = [float(v) for v in space]
    p.voxel[:] = [1.0, 1.0, 1.0]
    p.mode, p.formulation, p.pipe_rows, p.use_grid = 0, 0, 0, 1
    return p


def rollout(pos, vel, *, consts, ext, dt, space, steps, branches=None, branch_mode=0):
    """Positions and velocities after `steps` NAIVE box steps (FP64).  `branches`, a uint8
    array of steps * 3N, records (branch_mode 1) or replays (2) the wall-bounce decisions."""
    pos = _f64(pos, (-1, 3))
    vel = _f64(vel, (-1, 3))
    n = pos.shape[0]
    xT = np.zeros_like(pos)
    vT = np.zeros_like(vel)
    c = consts_of(consts)
    bp = None
    if branches is not None:
        assert branches.dtype == np.uint8 and branches.flags.c_contiguous and branches.size >= steps * 3 * n
        bp = branches.ctypes.data_as(c_void_p)
    lib().oc_rollout(ctypes.byref(_box_params(n, ext, dt, space)), ctypes.byref(c), _d(pos), _d(vel), int(steps),
                     _d(xT), _d(vT), bp, int(branch_mode) if bp is not None else 0)
    return xT, vT


def adjoint_step(pos, vel, gx1, gv1, *, consts, ext, dt, space):
    """Adjoint of one NAIVE box step at (pos, vel): returns (gx0, gv0, dL/dmu) for the
    cotangents (gx1, gv1) of the step's output."""
    pos = _f64(pos, (-1, 3))
    vel = _f64(vel, (-1, 3))
    gx1 = _f64(gx1, (-1, 3))
    gv1 = _f64(gv1, (-1, 3))
    n = pos.shape[0]
    gx0 = np.zeros_like(pos)
    gv0 = np.zeros_like(vel)
    gmu = c_double(0.0)
    lib().oc_adjoint_step(ctypes.byref(_box_params(n, ext, dt, space)), ctypes.byref(consts_of(consts)), _d(pos),
                          _d(vel), _d(gx1), _d(gv1), _d(gx0), _d(gv0), ctypes.byref(gmu))
    return gx0, gv0, gmu.value


def rollout_adjoint(pos, vel, *, consts, ext, dt, space, steps, target, every=20):
    """Loss (1/N) sum |x_T - target|^2 and its gradients with respect to the initial
    velocities (N, 3) and the viscosity, by the FP64 reverse pass with checkpoints every
    `every` steps.  Returns (loss, grad_v0, grad_mu)."""
    pos = _f64(pos, (-1, 3))
    vel = _f64(vel, (-1, 3))
    target = _f64(target, (-1, 3))
    n = pos.shape[0]
    g = np.zeros_like(vel)
    loss = c_double(0.0)
    gmu = c_double(0.0)
    c = consts_of(consts)
    lib().oc_rollout_adjoint(ctypes.byref(_box_params(n, ext, dt, space)), ctypes.byref(c), _d(pos), _d(vel),
                             int(steps), _d(target), int(every), ctypes.byref(loss), _d(g), ctypes.byref(gmu))
    return loss.value, g, gmu.value
