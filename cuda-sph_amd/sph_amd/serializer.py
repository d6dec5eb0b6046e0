"""Frame output in the reference's on-disk format, and an asynchronous writer for it.

Format (common/serializer/saver.py:8-46, loader.py:9-64), one folder per run:
  params.json          dataclasses.asdict(SimulationParameters) with arrays as lists,
                       sorted keys, indent 4
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


def _plain(o):
    """JSON form of values json cannot encode itself: numpy scalars and arrays (a numpy
    integer field would make the reference's json.dump fail), and objects by __dict__
    (saver.py:45)."""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    return o.__dict__


class Saver:
    """common/serializer/saver.py:8-46: writes params.json at construction and one set of
    .npy files per save_next_state call."""

    def __init__(self, out_dirname: str, params: SimulationParameters, root: Optional[str] = None) -> None:
        self.out_folder_path = os.path.join(root if root is not None else os.getcwd(), out_dirname)
        if not os.path.exists(self.out_folder_path):
            os.mkdir(self.out_folder_path)
        self._save_parameters(params)
        self.current_epoch = 0

    def save_next_state(self, state: SimulationState) -> None:
        """Saves the state as frame current_epoch and advances the counter (saver.py:28-36)."""
        for name, value in vars(state).items():
            np.save(os.path.join(self.out_folder_path, name + "_" + str(self.current_epoch)), value)
        self.current_epoch += 1

    def _save_parameters(self, params: SimulationParameters) -> None:
        d = dataclasses.asdict(params)
        for key, value in d.items():
            if type(value) is np.ndarray:
                d[key] = value.tolist()
        with open(os.path.join(self.out_folder_path, PARAMS_FILENAME), "w") as f:
            json.dump(d, f, default=_plain, sort_keys=True, indent=4)


class Loader:
    """common/serializer/loader.py:9-64."""

    def __init__(self, out_dirname: str, root: Optional[str] = None) -> None:
        self.out_folder_path = os.path.join(root if root is not None else os.getcwd(), out_dirname)
        if not os.path.exists(self.out_folder_path):
            raise Exception(f"Directory ({self.out_folder_path}) does not exists! Could not load simulation!")
        with open(os.path.join(self.out_folder_path, PARAMS_FILENAME)) as f:
            self._json = json.load(f)

    def load_simulation_parameters(self) -> SimulationParameters:
        d = {}
        for f in fields(SimulationParameters):
            if f.name == "pipe":
                d["pipe"] = self._load_pipe()
            elif f.type == "np.ndarray":
                d[f.name] = np.asarray(self._json[f.name])
            else:
                d[f.name] = self._json[f.name]
        return SimulationParameters(**d)

    def _load_pipe(self) -> Pipe:
        segs = []
        for s in self._json["pipe"]["segments"]:
            segs.append(Segment(start_point=tuple(s["start_point"]), start_radius=s["start_radius"],
                                end_radius=s["end_radius"], length=s["length"]))
        return Pipe(segs)

    def load_simulation_state(self, epoch: int) -> SimulationState:
        d = {}
        for f in fields(SimulationState):
            d[f.name] = np.load(os.path.join(self.out_folder_path, f.name + "_" + str(epoch)) + ".npy")
        return SimulationState(**d)


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
                path = self.saver.out_folder_path
                for name, t in zip(("position", "velocity", "density"), self._host[s]):
                    np.save(os.path.join(path, f"{name}_{epoch}"), t.numpy())
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
