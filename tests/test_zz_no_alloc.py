"""sph_step_noalloc (the step with its workspace in a static device array) against the same
step through a context with a caller-owned workspace, through the C ABI.

Both run the same kernels with the same launch shapes, so the bar is bitwise equality:
two calls in a row (the second with fewer particles, on a smaller grid), the NAIVE and the
VOXEL formulation, and the rejection of problems that do not fit the array.
"""
import ctypes

import numpy as np
import pytest
import torch

from sph_amd import _native as N
from sph_amd.workloads import box_uniform

pytestmark = pytest.mark.gpu


def _lib():
    return N.load_library()  # declares sph_step_noalloc (include/sph_amd.h)


def _case(n, seed, formulation, edge=1.0):
    w = box_uniform(n, seed=seed, edge=edge)
    # the VOXEL formulation's 27 voxels reach h only with voxels of at least h (0.0457)
    voxel = w.params.voxel_size if formulation == N.SPH_FORM_NAIVE else np.full(3, 0.0625)
    p = N.make_params(n, w.params.external_force, 1.0 / w.params.fps, w.params.space_size, voxel,
                      N.SPH_MODE_BOX, formulation)
    rng = np.random.default_rng(seed)
    pos = torch.from_numpy(w.state.position.astype(np.float32).astype(np.float64)).cuda()
    vel = torch.from_numpy(rng.normal(0.0, 0.3, size=(n, 3)).astype(np.float32).astype(np.float64)).cuda()
    return p, N.make_consts(w.constants), pos, vel


def _with_context(p, k, pos, vel, steps):
    ctx = N.NativeContext(p, k, "cuda:0")
    ctx.upload(pos, vel)
    ctx.step(steps)
    out = ctx.download()
    torch.cuda.synchronize()
    return [t.cpu() for t in out]


def _noalloc(lib, p, k, pos, vel, steps):
    n = pos.shape[0]
    outs = [torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda:0"),
            torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda:0"),
            torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0")]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.sph_step_noalloc(ctypes.byref(p), ctypes.byref(k), pos.data_ptr(), vel.data_ptr(),
                              *[t.data_ptr() for t in outs], steps, stream)
    N.check(rc, "sph_step_noalloc")
    torch.cuda.synchronize()
    return [t.cpu() for t in outs]


@pytest.mark.parametrize("formulation", [N.SPH_FORM_NAIVE, N.SPH_FORM_VOXEL])
def test_noalloc_equals_context_step(gpu, formulation):
    lib = _lib()
    assert lib.sph_noalloc_scratch_bytes() == 64 << 20
    # first call: 20,000 particles; second call: fewer particles on a smaller box (fewer
    # cells), so stale counts, lists and tables of the first call would show.  The VOXEL box
    # is a whole number of voxels (0.625 = 10 x 0.0625): otherwise the reference's lookups
    # name other voxels than its keys (voxel_sph_strategy.py:70-73 against :109-116), some
    # particles find no neighbour and the state turns NaN, as it does in the reference
    small = 0.6 if formulation == N.SPH_FORM_NAIVE else 0.625
    for n, seed, edge, steps in ((20000, 61, 1.0, 4), (3000, 62, small, 3)):
        p, k, pos, vel = _case(n, seed, formulation, edge)
        ref = _with_context(p, k, pos, vel, steps)
        got = _noalloc(lib, p, k, pos, vel, steps)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (n, formulation)
        assert torch.isfinite(got[0]).all() and torch.isfinite(got[2]).all()


def test_noalloc_rejects_what_does_not_fit(gpu):
    lib = _lib()
    p, k, pos, vel = _case(300_000, 63, N.SPH_FORM_NAIVE)  # about 90 MB of workspace
    rc = lib.sph_step_noalloc(ctypes.byref(p), ctypes.byref(k), pos.data_ptr(), vel.data_ptr(), pos.data_ptr(),
                              vel.data_ptr(), None, 1, None)
    assert rc == -1 and "bytes of scratch" in N.last_error()
    p2 = N.make_params(10, (0, 0, 0), 1e-3, (1, 1, 1), (0.0625,) * 3, N.SPH_MODE_PIPE, N.SPH_FORM_NAIVE,
                       np.array([[0, 0.5, 0.5, 0.1, 1.0], [1, 0.5, 0.5, 0.1, 1.0]]))
    rc = lib.sph_step_noalloc(ctypes.byref(p2), ctypes.byref(k), pos.data_ptr(), vel.data_ptr(), pos.data_ptr(),
                              vel.data_ptr(), None, 1, None)
    assert rc == -1 and "BOX mode" in N.last_error()
