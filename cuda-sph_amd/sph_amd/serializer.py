"""Frame output in the reference's on-disk format, and an asynchronous writer for it.

Format (common/serializer/saver.py:8-46, loader.py:9-64), one folder per run:
  params.json          every SimulationParameters field (the pipe as {"segments": [...]})
                       with arrays as lists, sorted keys, indent 4
  position_<k>.npy     float64 (N, 3) state after frame k (k = 0, 1, ...)
  velocity_<k>.npy     float64 (N, 3)
  density_<k>.npy      float64 (N,)
Saver and Loader mirror the reference classes (same methods, same files), so the
reference's visualiser and its Loader read what these write and the other way round.
The reference resolves the folder against its project root (config.ROOT_PROJ_DIRNAME);
here `root` defaults to the current directory.

AsyncFrameWriter is the MI355X-side path for the same files: the state stays on the GPU,
each frame is gathered into particle order by a kernel on the compute stream, copied to
pinned host memory on a second stream and written by a thread, so the GPU steps on while
earlier frames are copied and written.  A ring of `slots` frames bounds the memory; when
the disk is slower than the GPU the ring fills and the step loop waits for a free slot.
"""
from __future__ import annotations

import dataclasses
import json
import os
import queue
import threading
from dataclasses import fields
from typing import Optional

import numpy as np

from .data_classes import Pipe, Segment, SimulationParameters, SimulationState

PARAMS_FILENAME = "params.json"  # config.py:43

# One .npy file per quantity and frame: the fields of SimulationState, in order
FRAME_QUANTITIES = tuple(f.name for f in fields(SimulationState))


def frame_file(folder: str, quantity: str, frame: int) -> str:
    """<folder>/<quantity>_<frame>.npy, the name both classes below use for one quantity of
    one frame (saver.py:28-36, loader.py:52-64)."""
    return os.path.join(folder, f"{quantity}_{frame}.npy")


def _json_form(value):
    """A parameter value as params.json holds it: a dataclass as the object of its fields
    (the pipe as {"segments": [...]}), arrays and tuples as lists, numpy scalars as plain
    numbers (json cannot encode a numpy integer)."""
    if dataclasses.is_dataclass(value):
        return {f.name: _json_form(getattr(value, f.name)) for f in fields(value)}
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, (list, tuple)):
        return [_json_form(v) for v in value]
    if isinstance(value, np.generic):
        return value.item()
    return value


def params_json(params: SimulationParameters) -> str:
    """The text of params.json: every parameter in _json_form, keys sorted, indent 4."""
    return json.dumps(_json_form(params), sort_keys=True, indent=4)


def _default(f: dataclasses.Field):
    return f.default_factory() if f.default is dataclasses.MISSING else f.default


def params_from_json(obj: dict) -> SimulationParameters:
    """The inverse of params_json: array parameters (those whose default is an array) back
    to arrays, the pipe back to Pipe and Segment objects, the rest as read."""
    values = {}
    for f in fields(SimulationParameters):
        v = obj[f.name]
        if f.name == "pipe":
            v = Pipe([Segment(start_point=tuple(s["start_point"]), start_radius=s["start_radius"],
                              end_radius=s["end_radius"], length=s["length"]) for s in v["segments"]])
        elif isinstance(_default(f), np.ndarray):
            v = np.asarray(v)
        values[f.name] = v
    return SimulationParameters(**values)


class Saver:
    """common/serializer/saver.py:8-46: the run's folder under `root` (created if missing),
    params.json written at construction, then one frame per save_next_state call."""

    def __init__(self, out_dirname: str, params: SimulationParameters, root: Optional[str] = None) -> None:
        self.out_folder_path = os.path.join(os.getcwd() if root is None else root, out_dirname)
        if not os.path.isdir(self.out_folder_path):
            os.mkdir(self.out_folder_path)
        with open(os.path.join(self.out_folder_path, PARAMS_FILENAME), "w") as fh:
            fh.write(params_json(params))
        self.current_epoch = 0

    def save_next_state(self, state: SimulationState) -> None:
        """Writes the state as frame current_epoch, then advances the frame counter."""
        for q in FRAME_QUANTITIES:
            np.save(frame_file(self.out_folder_path, q, self.current_epoch), getattr(state, q))
        self.current_epoch += 1


class Loader:
    """common/serializer/loader.py:9-64: reads a folder Saver wrote; a missing folder is an
    error at construction (with the reference's message)."""

    def __init__(self, out_dirname: str, root: Optional[str] = None) -> None:
        self.out_folder_path = os.path.join(os.getcwd() if root is None else root, out_dirname)
        if not os.path.exists(self.out_folder_path):
            raise Exception(f"Directory ({self.out_folder_path}) does not exists! Could not load simulation!")
        with open(os.path.join(self.out_folder_path, PARAMS_FILENAME)) as fh:
            self._params = json.load(fh)

    def load_simulation_parameters(self) -> SimulationParameters:
        return params_from_json(self._params)

    def load_simulation_state(self, epoch: int) -> SimulationState:
        return SimulationState(**{q: np.load(frame_file(self.out_folder_path, q, epoch)) for q in FRAME_QUANTITIES})


class AsyncFrameWriter:
    """Saver-format frames of a strategy's resident state, written without stopping the
    GPU.  capture() enqueues the current state as the next frame; close() waits for every
    file.  Calls come from one thread (the step loop)."""

    def __init__(self, strategy, out_dirname: str, params: SimulationParameters, root: Optional[str] = None,
                 slots: int = 3) -> None:
        import torch

        if slots < 1:
            raise ValueError("slots must be >= 1")
        self._torch = torch
        self.saver = Saver(out_dirname, params, root)
        self.ctx = strategy.ctx
        self.device = self.ctx.device
        n = self.ctx.n
        f64 = dict(dtype=torch.float64)
        self._dev = [(torch.empty((n, 3), device=self.device, **f64), torch.empty((n, 3), device=self.device, **f64),
                      torch.empty((n,), device=self.device, **f64)) for _ in range(slots)]
        self._host = [(torch.empty((n, 3), pin_memory=True, **f64), torch.empty((n, 3), pin_memory=True, **f64),
                       torch.empty((n,), pin_memory=True, **f64)) for _ in range(slots)]
        self._copied = [torch.cuda.Event() for _ in range(slots)]
        self._stream = torch.cuda.Stream(self.device)  # default priority
        self._free = threading.Semaphore(slots)
        self._queue: "queue.Queue" = queue.Queue()
        self._slot = 0
        self._error: Optional[BaseException] = None
        self._thread = threading.Thread(target=self._write_loop, name="sph-frame-writer", daemon=True)
        self._thread.start()

    @property
    def frames(self) -> int:
        return self.saver.current_epoch

    def capture(self) -> None:
        torch = self._torch
        if self._error is not None:
            raise RuntimeError("frame writer failed") from self._error
        self._free.acquire()  # the slot's previous frame is on disk: both buffers are reusable
        s = self._slot
        self._slot = (s + 1) % len(self._dev)
        pos, vel, rho = self._dev[s]
        self.ctx.download(out=(pos, vel, rho))  # gather to particle order, compute stream
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self._stream):
            self._stream.wait_event(ready)
            for h, d in zip(self._host[s], self._dev[s]):
                h.copy_(d, non_blocking=True)
            self._copied[s].record(self._stream)
        epoch = self.saver.current_epoch
        self.saver.current_epoch += 1
        self._queue.put((s, epoch))

    def _write_loop(self) -> None:
        while True:
            item = self._queue.get()
            if item is None:
                return
            s, epoch = item
            try:
                self._copied[s].synchronize()
                for q, t in zip(FRAME_QUANTITIES, self._host[s]):
                    np.save(frame_file(self.saver.out_folder_path, q, epoch), t.numpy())
            except BaseException as e:  # reported by the next capture() or close()
                self._error = e
            finally:
                self._free.release()

    def close(self) -> None:
        self._queue.put(None)
        self._thread.join()
        if self._error is not None:
            raise RuntimeError("frame writer failed") from self._error

    def __enter__(self) -> "AsyncFrameWriter":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def run_to_disk(strategy, start_state: SimulationState, params: SimulationParameters, out_dirname: str,
                frames: Optional[int] = None, steps_per_frame: int = 1, root: Optional[str] = None,
                slots: int = 3) -> int:
    """sim/src/main.py:10-26 on the GPU: duration * fps frames (or `frames`), each the state
    after `steps_per_frame` steps (the reference saves every step), written in the Saver
    format while the GPU steps on.  Returns the number of frames written."""
    import torch

    n_frames = int(params.duration * params.fps) if frames is None else int(frames)
    pos = torch.from_numpy(np.ascontiguousarray(start_state.position, dtype=np.float64)).to(strategy.device)
    vel = torch.from_numpy(np.ascontiguousarray(start_state.velocity, dtype=np.float64)).to(strategy.device)
    strategy.ctx.upload(pos, vel)
    with AsyncFrameWriter(strategy, out_dirname, params, root=root, slots=slots) as w:
        for _ in range(n_frames):
            strategy.ctx.step(steps_per_frame)
            w.capture()
    return n_frames
