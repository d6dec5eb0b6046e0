"""Saver / Loader in the reference's format (common/serializer/saver.py, loader.py),
following the reference's own tests (common/tests/test_saving_loading.py) plus the exact
file layout: params.json text, .npy names and dtypes."""
import json
import os

import numpy as np
import pytest

from sph_amd.data_classes import Pipe, Segment, SimulationParameters, SimulationState
from sph_amd.serializer import PARAMS_FILENAME, Loader, Saver


def test_settings_loading(tmp_path):
    """test_saving_loading.py:11-31"""
    params = SimulationParameters(pipe=Pipe([Segment(), Segment((1.0, 0.0, 0.0), 1, 4, 5)]))
    Saver("data", params, root=str(tmp_path))
    loaded = Loader("data", root=str(tmp_path)).load_simulation_parameters()
    for name in vars(params):
        a, b = getattr(params, name), getattr(loaded, name)
        assert type(a) is type(b)
        if type(a) is np.ndarray:
            assert np.all(a == b)
        else:
            assert a == b


def test_states_round_trip(tmp_path):
    """test_saving_loading.py:33-55"""
    saver = Saver("data", SimulationParameters(), root=str(tmp_path))
    first = SimulationState(position=np.asarray([0, 1, 0]))
    second = SimulationState(position=np.asarray([0, 2, 0]))
    saver.save_next_state(first)
    saver.save_next_state(second)
    loader = Loader("data", root=str(tmp_path))
    for k, true in enumerate([first, second]):
        got = loader.load_simulation_state(k)
        for name in vars(first):
            assert np.all(getattr(true, name) == getattr(got, name))


def test_file_layout(tmp_path):
    params = SimulationParameters(particle_count=3, external_force=np.array([0.0, -2.0, 0.0]), duration=15, fps=20,
                                  space_size=np.array([40.0, 6.0, 6.0]), voxel_size=np.array([2.0, 2.0, 2.0]),
                                  pipe=Pipe([Segment((0.0, 3.0, 3.0), 1.0, 1.2, 2.5)]))
    saver = Saver("out", params, root=str(tmp_path))
    rng = np.random.default_rng(0)
    st = SimulationState(rng.normal(size=(3, 3)), rng.normal(size=(3, 3)), rng.normal(size=3))
    saver.save_next_state(st)
    folder = tmp_path / "out"
    assert sorted(os.listdir(folder)) == ["density_0.npy", PARAMS_FILENAME, "position_0.npy", "velocity_0.npy"]
    expected = {"duration": 15, "external_force": [0.0, -2.0, 0.0], "fps": 20, "particle_count": 3,
                "pipe": {"segments": [{"end_radius": 1.2, "length": 2.5, "start_point": [0.0, 3.0, 3.0],
                                       "start_radius": 1.0}]},
                "space_size": [40.0, 6.0, 6.0], "voxel_size": [2.0, 2.0, 2.0]}
    assert (folder / PARAMS_FILENAME).read_text() == json.dumps(expected, sort_keys=True, indent=4)
    pos = np.load(folder / "position_0.npy")
    assert pos.dtype == np.float64 and np.array_equal(pos, st.position)
    assert np.array_equal(np.load(folder / "density_0.npy"), st.density)


def test_numpy_scalar_fields_are_written(tmp_path):
    Saver("d", SimulationParameters(particle_count=np.int64(7), fps=np.int32(30)), root=str(tmp_path))
    p = Loader("d", root=str(tmp_path)).load_simulation_parameters()
    assert p.particle_count == 7 and p.fps == 30


def test_loader_missing_folder(tmp_path):
    with pytest.raises(Exception, match="does not exists"):
        Loader("nowhere", root=str(tmp_path))
