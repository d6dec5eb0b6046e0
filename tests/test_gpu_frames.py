"""Asynchronous frame output against the synchronous path: the same run written by
AsyncFrameWriter and by StateGenerator + Saver (one compute_next_state per frame, as
sim/src/main.py does) must give identical files, frame for frame."""
from itertools import islice

import numpy as np
import pytest

from sph_amd.serializer import Loader, Saver, run_to_disk
from sph_amd.state_generator import StateGenerator
from sph_amd.strategy import NaiveSPHStrategy
from sph_amd.workloads import box_uniform

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slots", [1, 3])
def test_async_frames_equal_sync_frames(gpu, tmp_path, slots):
    w = box_uniform(3000, seed=71)
    params = w.params  # dt = 1 / fps = 1e-3; the first 25 of duration * fps frames
    n_frames = 25
    strat = NaiveSPHStrategy(params, constants=w.constants, device=gpu)
    assert run_to_disk(strat, w.state, params, "async", root=str(tmp_path), frames=n_frames, slots=slots) == n_frames
    saver = Saver("sync", params, root=str(tmp_path))
    for state in islice(StateGenerator(w.state, params, constants=w.constants, device=gpu), n_frames):
        assert np.isfinite(state.position).all()
        saver.save_next_state(state)
    assert saver.current_epoch == n_frames
    a, s = Loader("async", root=str(tmp_path)), Loader("sync", root=str(tmp_path))
    assert (tmp_path / "async" / "params.json").read_text() == (tmp_path / "sync" / "params.json").read_text()
    for k in range(n_frames):
        fa, fs = a.load_simulation_state(k), s.load_simulation_state(k)
        for name in ("position", "velocity", "density"):
            assert np.array_equal(getattr(fa, name), getattr(fs, name)), (k, name)
