"""ctypes binding of libsph_amd.so (declarations in include/sph_amd.h).

torch is imported before the library is opened: libsph_amd.so needs libamdhip64.so.7 and
torch has already loaded its own copy under that soname, so both share one HIP runtime
(streams and device pointers pass between them).  There is no fallback: if the library
is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np
import torch

from ._build import LIB_PATH

SPH_MODE_BOX, SPH_MODE_PIPE = 0, 1
SPH_FORM_NAIVE, SPH_FORM_VOXEL = 0, 1
SPH_FLAG_KEEP_TERMS = 1
SPH_FLAG_ADJOINT = 2
SPH_FLAG_CHECK_FINITE = 4
SPH_E_NONFINITE = -4
SPH_ABI_VERSION = 1  # include/sph_amd.h; the loaded library must report the same

STAGE_SORT, STAGE_TABLE, STAGE_REORDER, STAGE_DENSITY, STAGE_FORCE = 1, 2, 3, 4, 5

BUF_KEYS, BUF_PERM, BUF_SORTED_KEYS, BUF_CELL_BEGIN, BUF_DENSITY = 1, 2, 3, 4, 5
BUF_PRESSURE, BUF_VISCOSITY, BUF_FORCE, BUF_RNG, BUF_NEIGHBOURS = 6, 7, 8, 9, 10

_BUF_DTYPE = {
    BUF_KEYS: torch.int32, BUF_PERM: torch.int32, BUF_SORTED_KEYS: torch.int32, BUF_CELL_BEGIN: torch.int32,
    BUF_DENSITY: torch.float32, BUF_PRESSURE: torch.float32, BUF_VISCOSITY: torch.float32,
    BUF_FORCE: torch.float32, BUF_RNG: torch.int64, BUF_NEIGHBOURS: torch.int32,
}


class SphConstantsC(ctypes.Structure):
    _fields_ = [("mass", c_double), ("rho0", c_double), ("h", c_double), ("visc", c_double), ("k", c_double),
                ("damp", c_double)]


class SphParamsC(ctypes.Structure):
    _fields_ = [
        ("particle_count", c_int64),
        ("external_force", c_double * 3),
        ("dt", c_double),
        ("space_size", c_double * 3),
        ("voxel_size", c_double * 3),
        ("mode", c_int32),
        ("formulation", c_int32),
        ("pipe_rows", c_int32),
        ("flags", c_int32),
        ("pipe", POINTER(c_double)),
        ("seed", c_uint64),
    ]


EXPORTED = [
    "sph_abi_version", "sph_last_error", "sph_create", "sph_destroy", "sph_workspace_bytes", "sph_bind_workspace",
    "sph_upload_state", "sph_upload_state_f32", "sph_step", "sph_download_state", "sph_download_state_f32",
    "sph_run_stage", "sph_buffer_elems", "sph_read_buffer", "sph_diagnostics", "sph_grid_info", "sph_set_launch",
    "sph_create_batch", "sph_batch_info", "sph_set_constants", "sph_set_lanes", "sph_adjoint_seed", "sph_adjoint_step",
    "sph_adjoint_read", "sph_set_batch_lanes", "sph_adjoint_snapshot_bytes", "sph_adjoint_save", "sph_step_save", "sph_adjoint_step_from",
    "sph_check_finite", "sph_noalloc_scratch_bytes", "sph_step_noalloc", "sph_step_timed", "sph_set_slabs",
]

_lib = None


def load_library() -> ctypes.CDLL:
    """Opens libsph_amd.so (once) and declares its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback for the SPH step)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = c_void_p, c_int32, c_int64
    sig = {
        "sph_abi_version": ([], i32),
        "sph_last_error": ([ctypes.c_char_p, c_size_t], i32),
        "sph_create": ([POINTER(SphParamsC), POINTER(SphConstantsC), POINTER(vp)], i32),
        "sph_destroy": ([vp], i32),
        "sph_workspace_bytes": ([vp], c_size_t),
        "sph_bind_workspace": ([vp, vp, c_size_t, vp], i32),
        "sph_upload_state": ([vp, vp, vp, vp], i32),
        "sph_upload_state_f32": ([vp] + [vp] * 6 + [vp], i32),
        "sph_step": ([vp, i32, vp], i32),
        "sph_step_timed": ([vp, i32, POINTER(ctypes.c_double), vp], i32),
        "sph_download_state": ([vp, vp, vp, vp, vp, vp], i32),
        "sph_download_state_f32": ([vp] + [vp] * 7 + [vp], i32),
        "sph_run_stage": ([vp, i32, vp], i32),
        "sph_buffer_elems": ([vp, i32], i64),
        "sph_read_buffer": ([vp, i32, vp, vp], i32),
        "sph_diagnostics": ([vp, vp, vp], i32),
        "sph_grid_info": ([vp, POINTER(i32), POINTER(i32), POINTER(i32), POINTER(i32)], i32),
        "sph_set_launch": ([vp, i32, i32], i32),
        "sph_create_batch": ([POINTER(SphParamsC), POINTER(SphConstantsC), i32, POINTER(vp)], i32),
        "sph_batch_info": ([vp, POINTER(i32), POINTER(i32)], i32),
        "sph_set_constants": ([vp, POINTER(SphConstantsC)], i32),
        "sph_set_lanes": ([vp, i32], i32),
        "sph_set_slabs": ([vp, i32, vp], i32),
        "sph_set_batch_lanes": ([vp, vp, i32, vp], i32),
        "sph_adjoint_seed": ([vp, vp, vp, vp], i32),
        "sph_adjoint_step": ([vp, vp], i32),
        "sph_adjoint_read": ([vp, vp, vp, vp, vp], i32),
        "sph_adjoint_snapshot_bytes": ([vp], i64),
        "sph_adjoint_save": ([vp, vp, c_size_t, vp], i32),
        "sph_step_save": ([vp, vp, c_size_t, vp], i32),
        "sph_adjoint_step_from": ([vp, vp, c_size_t, vp], i32),
        "sph_check_finite": ([vp, POINTER(i64), vp], i32),
        "sph_noalloc_scratch_bytes": ([], c_size_t),
        "sph_step_noalloc": ([POINTER(SphParamsC), POINTER(SphConstantsC), vp, vp, vp, vp, vp, i32, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    version = lib.sph_abi_version()
    if version != SPH_ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} reports ABI version {version}, this binding expects {SPH_ABI_VERSION}: "
                           "rebuild it with __graft_entry__.build()")
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load_library().sph_last_error(buf, 512)
    return buf.value.decode(errors="replace")


class NonFiniteStateError(RuntimeError):
    """A particle position became NaN or infinite (SPH_E_NONFINITE)."""


def check(rc: int, what: str) -> None:
    if rc == SPH_E_NONFINITE:
        raise NonFiniteStateError(f"{what} failed ({rc}): {last_error()}")
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def _ptr(t) -> c_void_p:
    return c_void_p(0 if t is None else t.data_ptr())


def _stream(device) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


class NativeContext:
    """One SPH context on one device: owns the workspace tensor and the C handle."""

    def __init__(self, params, consts, device="cuda:0"):
        """params/consts: one SphParamsC/SphConstantsC for a plain step, or equal-length
        sequences of them for the batched form (sph_create_batch)."""
        if not torch.cuda.is_available():
            raise RuntimeError("the SPH step runs on a ROCm GPU; none is available")
        self.lib = load_library()
        self.device = torch.device(device)
        h = c_void_p()
        if isinstance(params, SphParamsC):
            check(self.lib.sph_create(ctypes.byref(params), ctypes.byref(consts), ctypes.byref(h)), "sph_create")
        else:
            count = len(params)
            parr = (SphParamsC * count)(*params)
            carr = (SphConstantsC * count)(*consts)
            check(self.lib.sph_create_batch(parr, carr, count, ctypes.byref(h)), "sph_create_batch")
        self.handle = h
        # bumped by every call that changes the resident state (upload, steps, stages), so a
        # holder of a downloaded copy can tell whether the device still holds that state
        self.generation = 0
        cnt = c_int32()
        self.lib.sph_batch_info(h, ctypes.byref(cnt), None)
        first = (c_int32 * (cnt.value + 1))()
        self.lib.sph_batch_info(h, None, first)
        self.first = list(first)
        self.n = self.first[-1]
        nbytes = int(self.lib.sph_workspace_bytes(h))
        # 256-byte alignment for the slices: over-allocate and offset
        self._ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.device)
        off = (-self._ws.data_ptr()) % 256
        self.workspace_bytes = nbytes
        with torch.cuda.device(self.device):
            check(self.lib.sph_bind_workspace(h, c_void_p(self._ws.data_ptr() + off), nbytes, _stream(self.device)),
                  "sph_bind_workspace")

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.sph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------------------
    def upload(self, pos: torch.Tensor, vel: torch.Tensor):
        """pos, vel: float64 (N, 3) contiguous on the device."""
        assert pos.dtype == torch.float64 and vel.dtype == torch.float64
        assert pos.shape == (self.n, 3) and vel.shape == (self.n, 3)
        assert pos.is_contiguous() and vel.is_contiguous()
        self.generation += 1
        check(self.lib.sph_upload_state(self.handle, _ptr(pos), _ptr(vel), _stream(self.device)), "sph_upload_state")

    def upload_f32(self, x, y, z, vx, vy, vz):
        for t in (x, y, z, vx, vy, vz):
            assert t.dtype == torch.float32 and t.shape == (self.n,) and t.is_contiguous()
        self.generation += 1
        check(self.lib.sph_upload_state_f32(self.handle, *[_ptr(t) for t in (x, y, z, vx, vy, vz)],
                                            _stream(self.device)), "sph_upload_state_f32")

    def step(self, nsteps: int = 1):
        self.generation += 1
        check(self.lib.sph_step(self.handle, int(nsteps), _stream(self.device)), "sph_step")

    def step_timed(self, nsteps: int) -> dict:
        """step(nsteps) with HIP events between the stages (waits for them): mean ms per step
        of the sort (with the gather into cell order), the density pass and the force pass."""
        self.generation += 1
        out = (ctypes.c_double * 3)()
        check(self.lib.sph_step_timed(self.handle, int(nsteps), out, _stream(self.device)), "sph_step_timed")
        return {"sort": out[0], "density": out[1], "force": out[2]}

    def run_stage(self, stage: int):
        self.generation += 1
        check(self.lib.sph_run_stage(self.handle, int(stage), _stream(self.device)), "sph_run_stage")

    def check_finite(self) -> int:
        """Waits for the stream; returns the number of positions found non-finite when hashed
        since the last check (raises NonFiniteStateError if it is not zero)."""
        cnt = c_int64(0)
        rc = self.lib.sph_check_finite(self.handle, ctypes.byref(cnt), _stream(self.device))
        check(rc, "sph_check_finite")
        return int(cnt.value)

    def nonfinite_count(self) -> int:
        """Like check_finite, but returns the count instead of raising."""
        cnt = c_int64(0)
        rc = self.lib.sph_check_finite(self.handle, ctypes.byref(cnt), _stream(self.device))
        if rc not in (0, SPH_E_NONFINITE):
            check(rc, "sph_check_finite")
        return int(cnt.value)

    def download(self, force: bool = False, out=None):
        """Returns float64 device tensors pos (N,3), vel (N,3), rho (N,) [, force (N,3)];
        `out` may give (pos, vel, rho) destinations (contiguous float64 on the device)."""
        kw = dict(dtype=torch.float64, device=self.device)
        if out is not None:
            pos, vel, rho = out
            for t, shape in ((pos, (self.n, 3)), (vel, (self.n, 3)), (rho, (self.n,))):
                assert t.dtype == torch.float64 and tuple(t.shape) == shape and t.is_contiguous()
                assert t.device == self.device
        else:
            pos = torch.empty((self.n, 3), **kw)
            vel = torch.empty((self.n, 3), **kw)
            rho = torch.empty((self.n,), **kw)
        f = torch.empty((self.n, 3), **kw) if force else None
        check(self.lib.sph_download_state(self.handle, _ptr(pos), _ptr(vel), _ptr(rho), _ptr(f),
                                          _stream(self.device)), "sph_download_state")
        return (pos, vel, rho, f) if force else (pos, vel, rho)

    def download_f32(self, out=None):
        """Returns seven float32 device tensors x, y, z, vx, vy, vz, rho; `out` may give the
        destination tensors (each (N,) contiguous, or None to skip that field)."""
        if out is None:
            out = [torch.empty((self.n,), dtype=torch.float32, device=self.device) for _ in range(7)]
        for t in out:
            assert t is None or (t.dtype == torch.float32 and t.shape == (self.n,) and t.is_contiguous())
        check(self.lib.sph_download_state_f32(self.handle, *[_ptr(t) for t in out], _stream(self.device)),
              "sph_download_state_f32")
        return out

    def read_buffer(self, which: int) -> torch.Tensor:
        n = int(self.lib.sph_buffer_elems(self.handle, which))
        if n < 0:
            raise ValueError(f"unknown buffer {which}")
        out = torch.empty((max(n, 1),), dtype=_BUF_DTYPE[which], device=self.device)
        check(self.lib.sph_read_buffer(self.handle, which, _ptr(out), _stream(self.device)), "sph_read_buffer")
        return out[:n]

    def diagnostics(self) -> np.ndarray:
        out = torch.empty((8,), dtype=torch.float64, device=self.device)
        check(self.lib.sph_diagnostics(self.handle, _ptr(out), _stream(self.device)), "sph_diagnostics")
        return out.cpu().numpy()

    def grid_info(self):
        dims = (c_int32 * 3)()
        sten = (c_int32 * 3)()
        bits, passes = c_int32(), c_int32()
        check(self.lib.sph_grid_info(self.handle, dims, sten, ctypes.byref(bits), ctypes.byref(passes)),
              "sph_grid_info")
        return dict(dims=tuple(dims), stencil=tuple(sten), key_bits=bits.value, passes=passes.value)

    def set_launch(self, density_block: int, force_block: int):
        check(self.lib.sph_set_launch(self.handle, density_block, force_block), "sph_set_launch")

    def set_lanes(self, lanes: int):
        check(self.lib.sph_set_lanes(self.handle, int(lanes)), "sph_set_lanes")

    def set_slabs(self, enable: bool):
        self.generation += 1
        check(self.lib.sph_set_slabs(self.handle, 1 if enable else 0, _stream(self.device)), "sph_set_slabs")

    def set_batch_lanes(self, lanes):
        """Lanes per particle of each instance of a batched context."""
        arr = (ctypes.c_int32 * len(lanes))(*[int(x) for x in lanes])
        check(self.lib.sph_set_batch_lanes(self.handle, arr, len(lanes), _stream(self.device)),
              "sph_set_batch_lanes")

    def set_constants(self, consts: SphConstantsC):
        check(self.lib.sph_set_constants(self.handle, ctypes.byref(consts)), "sph_set_constants")

    # -- reverse mode (SPH_FLAG_ADJOINT) --------------------------------------------------
    def adjoint_seed(self, gx: torch.Tensor = None, gv: torch.Tensor = None):
        """gx, gv: float32 (N, 3) contiguous device tensors in particle order, or None (zero)."""
        for t in (gx, gv):
            assert t is None or (t.dtype == torch.float32 and t.shape == (self.n, 3) and t.is_contiguous())
        check(self.lib.sph_adjoint_seed(self.handle, _ptr(gx), _ptr(gv), _stream(self.device)), "sph_adjoint_seed")

    def adjoint_step(self):
        self.generation += 1
        check(self.lib.sph_adjoint_step(self.handle, _stream(self.device)), "sph_adjoint_step")

    def adjoint_snapshot_bytes(self) -> int:
        return int(self.lib.sph_adjoint_snapshot_bytes(self.handle))

    def _snapshot(self, snapshot: torch.Tensor) -> int:
        if snapshot.dtype != torch.uint8 or snapshot.device != self.device or not snapshot.is_contiguous():
            raise ValueError("a snapshot is a contiguous uint8 tensor on the context's device")
        return snapshot.numel()

    def adjoint_save(self, snapshot: torch.Tensor):
        """The last step's snapshot into `snapshot` (uint8 device tensor of at least
        adjoint_snapshot_bytes, 256-byte aligned)."""
        nbytes = self._snapshot(snapshot)
        check(self.lib.sph_adjoint_save(self.handle, _ptr(snapshot), nbytes, _stream(self.device)),
              "sph_adjoint_save")

    def step_save(self, snapshot: torch.Tensor):
        """One step whose snapshot is written straight into `snapshot`."""
        nbytes = self._snapshot(snapshot)
        self.generation += 1
        check(self.lib.sph_step_save(self.handle, _ptr(snapshot), nbytes, _stream(self.device)), "sph_step_save")

    def adjoint_step_from(self, snapshot: torch.Tensor):
        nbytes = self._snapshot(snapshot)
        check(self.lib.sph_adjoint_step_from(self.handle, _ptr(snapshot), nbytes, _stream(self.device)),
              "sph_adjoint_step_from")

    def adjoint_read(self):
        """Returns float32 device tensors gx (N, 3), gv (N, 3) and dL/dmu (1,)."""
        kw = dict(dtype=torch.float32, device=self.device)
        gx = torch.empty((self.n, 3), **kw)
        gv = torch.empty((self.n, 3), **kw)
        gmu = torch.empty((1,), **kw)
        check(self.lib.sph_adjoint_read(self.handle, _ptr(gx), _ptr(gv), _ptr(gmu), _stream(self.device)),
              "sph_adjoint_read")
        return gx, gv, gmu


def make_params(n, external_force, dt, space_size, voxel_size, mode, formulation, pipe_table=None, flags=0,
                seed=16435234):
    """Builds the C parameter struct; keeps the pipe table alive on the returned object."""
    p = SphParamsC()
    p.particle_count = int(n)
    p.external_force[:] = [float(v) for v in external_force]
    p.dt = float(dt)
    p.space_size[:] = [float(v) for v in space_size]
    p.voxel_size[:] = [float(v) for v in voxel_size]
    p.mode = int(mode)
    p.formulation = int(formulation)
    p.flags = int(flags)
    p.seed = int(seed)
    if pipe_table is not None and np.size(pipe_table) > 0:
        tbl = np.ascontiguousarray(np.asarray(pipe_table, dtype=np.float64).reshape(-1, 5))
        p.pipe_rows = tbl.shape[0]
        p.pipe = tbl.ctypes.data_as(POINTER(c_double))
        p._keep = tbl
    else:
        p.pipe_rows = 0
        p.pipe = None
    return p


def make_consts(c) -> SphConstantsC:
    k = SphConstantsC()
    k.mass, k.rho0, k.h, k.visc, k.k, k.damp = (float(c.mass), float(c.rho0), float(c.h), float(c.visc),
                                                float(c.k), float(c.damp))
    return k
