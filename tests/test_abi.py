"""The C ABI library loads without a GPU and exports every function include/sph_amd.h declares;
host-only entry points (planning, validation) behave as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

import torch  # noqa: F401  (binds the HIP runtime soname before the library loads)
from sph_amd import _native as N
from sph_amd._build import LIB_PATH

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sph_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|int64_t)\s+(sph_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_functions():
    names = declared()
    assert len(names) >= 15
    assert set(names) == set(N.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB_PATH):
        pytest.skip("libsph_amd.so not built")
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def declared_arity():
    """Number of parameters of each function include/sph_amd.h declares."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"^\s*(?:int|size_t|int64_t)\s+(sph_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        args = args.strip()
        out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_ctypes_signatures_match_the_header():
    """Every binding passes as many arguments as the C declaration takes (a stale ctypes
    signature would shift pointers and sizes)."""
    if not os.path.exists(LIB_PATH):
        pytest.skip("libsph_amd.so not built")
    lib = N.load_library()
    arity = declared_arity()
    assert set(arity) == set(declared())
    for name, count in arity.items():
        assert len(getattr(lib, name).argtypes) == count, name


def _params(**over):
    kw = dict(n=1000, external_force=(0, -9.81, 0), dt=1e-3, space_size=(1, 1, 1), voxel_size=(0.0625,) * 3,
              mode=N.SPH_MODE_BOX, formulation=N.SPH_FORM_NAIVE)
    kw.update(over)
    return N.make_params(**kw)


def _consts(h=0.0457):
    from sph_amd.constants import SPHConstants

    return N.make_consts(SPHConstants(mass=0.02, rho0=1000, h=h, visc=3.5, k=3, damp=0.5))


def test_abi_version_agrees_with_header():
    import os
    import re

    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "sph_amd.h")).read()
    assert int(re.search(r"#define SPH_ABI_VERSION (\d+)", hdr).group(1)) == N.SPH_ABI_VERSION
    assert N.load_library().sph_abi_version() == N.SPH_ABI_VERSION  # load_library checks it too


def test_create_plans_grid_and_sort():
    lib = N.load_library()
    assert lib.sph_abi_version() == 1
    h = ctypes.c_void_p()
    assert lib.sph_create(ctypes.byref(_params()), ctypes.byref(_consts()), ctypes.byref(h)) == 0
    dims = (ctypes.c_int32 * 3)()
    sten = (ctypes.c_int32 * 3)()
    bits, passes = ctypes.c_int32(), ctypes.c_int32()
    assert lib.sph_grid_info(h, dims, sten, ctypes.byref(bits), ctypes.byref(passes)) == 0
    assert tuple(dims) == (16, 16, 16) and tuple(sten) == (1, 1, 1)
    assert bits.value == 12 and passes.value == 2
    assert lib.sph_workspace_bytes(h) > 1000 * 16 * 5
    # calls needing device memory fail cleanly before a workspace is bound
    assert lib.sph_step(h, 1, None) == -2
    assert "workspace" in N.last_error()
    assert lib.sph_destroy(h) == 0


@pytest.mark.parametrize("over,msg", [
    (dict(n=-1), "particle_count"),
    (dict(dt=0.0), "dt"),
    (dict(space_size=(1, 0, 1)), "space_size"),
    (dict(mode=7), "mode"),
    (dict(mode=N.SPH_MODE_PIPE), "pipe table"),
    (dict(voxel_size=(1e-7, 1e-7, 1e-7)), "too many"),
])
def test_create_rejects_bad_parameters(over, msg):
    lib = N.load_library()
    h = ctypes.c_void_p()
    rc = lib.sph_create(ctypes.byref(_params(**over)), ctypes.byref(_consts()), ctypes.byref(h))
    assert rc == -1
    assert msg in N.last_error()


def test_voxel_formulation_sorts_full_signed_keys():
    lib = N.load_library()
    h = ctypes.c_void_p()
    p = _params(formulation=N.SPH_FORM_VOXEL, voxel_size=(0.0457,) * 3)
    assert lib.sph_create(ctypes.byref(p), ctypes.byref(_consts()), ctypes.byref(h)) == 0
    dims = (ctypes.c_int32 * 3)()
    bits, passes = ctypes.c_int32(), ctypes.c_int32()
    lib.sph_grid_info(h, dims, None, ctypes.byref(bits), ctypes.byref(passes))
    assert tuple(dims) == (22, 22, 22)  # ceil(1 / 0.0457), voxel_sph_strategy.py:70-73
    assert bits.value == 32 and passes.value == 4
    lib.sph_destroy(h)


def test_pipe_table_passes_through():
    from sph_amd.data_classes import Pipe, Segment

    tbl = Pipe([Segment(), Segment((1, 0, 0), 1, 2, 1)]).to_numpy()
    lib = N.load_library()
    h = ctypes.c_void_p()
    p = _params(mode=N.SPH_MODE_PIPE, pipe_table=tbl)
    assert lib.sph_create(ctypes.byref(p), ctypes.byref(_consts()), ctypes.byref(h)) == 0
    lib.sph_destroy(h)
    assert np.asarray(tbl).shape == (3, 5)
