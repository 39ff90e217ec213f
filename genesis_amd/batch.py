"""Batched replicas: many independent universes stepped by one set of
launches (``gs_batch_*`` of include/genesis_gpu.h).  The caller it replaces
is the reference's temperature sweep (examples/temperature_sweep.rs:70-113),
which runs one Universe after another."""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .inputs import State
from .params import GsParams, Physics
from .universe import _dp, _f64


class GpuBatch:
    def __init__(self, physics: Sequence[Physics], states: Sequence[State]):
        assert len(physics) == len(states) and len(states) > 0
        self.lib = _lib.load()
        self.R = len(states)
        self.sizes = [s.n for s in states]
        params = (GsParams * self.R)(*[p.to_c() for p in physics])
        n = (ctypes.c_int64 * self.R)(*self.sizes)
        h = ctypes.c_void_p()
        _lib.check(self.lib.gs_batch_create(self.R, params, n, ctypes.byref(h)))
        self.h = h
        self.species = [np.ascontiguousarray(s.species, dtype=np.uint8) for s in states]
        self.charge = [_f64(s.charge) for s in states]
        for r, s in enumerate(states):
            self.upload(r, s)

    def upload(self, r: int, s: State) -> None:
        arrs = [_f64(a) for a in (s.x, s.y, s.z, s.vx, s.vy, s.vz)]
        acc = (None, None, None) if s.ax is None else tuple(_f64(a) for a in (s.ax, s.ay, s.az))
        sp = np.ascontiguousarray(s.species, dtype=np.uint8)
        q = _f64(s.charge)
        _lib.check(self.lib.gs_batch_upload(self.h, r, *[_dp(a) for a in arrs], *[_dp(a) for a in acc],
                                            sp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(q)))

    def download(self, r: int) -> State:
        n = self.sizes[r]
        out = [np.empty(n) for _ in range(9)]
        _lib.check(self.lib.gs_batch_download(self.h, r, *[_dp(a) for a in out]))
        return State(out[0], out[1], out[2], out[3], out[4], out[5], self.species[r].copy(), self.charge[r].copy(),
                     out[6], out[7], out[8])

    def step(self, n_steps: int, graph_segment: int = 0, observe: bool = False) -> Optional[np.ndarray]:
        """n_steps ticks of all replicas; observables (n_steps, R, 6) if observe."""
        if not observe:
            _lib.check(self.lib.gs_batch_step(self.h, n_steps, graph_segment, None))
            return None
        buf = (_lib.Observables * max(1, n_steps * self.R))()
        _lib.check(self.lib.gs_batch_step(self.h, n_steps, graph_segment, buf))
        a = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_double)),
                                  shape=(max(1, n_steps * self.R), 6))[:n_steps * self.R]
        return a.reshape(n_steps, self.R, 6).copy()

    def autotune(self, trial_steps: int = 20):
        """Times every candidate setting (cell tiles at three cell sizes,
        Verlet lists at three skins and two lane counts) and keeps the
        fastest."""
        s, ms = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_batch_autotune(self.h, trial_steps, ctypes.byref(s), ctypes.byref(ms)))
        return {"cell_scale": s.value, "ms_per_tick": ms.value, **self.get_tuning()}

    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.check(self.lib.gs_batch_get_tuning(self.h, ctypes.byref(t)))
        return {"pair_mode": t.pair_mode, "lanes_per_atom": t.lanes_per_atom, "skin": t.skin,
                "sort_every": t.sort_every, "list_builds": t.list_builds, "list_replays": t.list_replays}

    def set_tuning(self, pair_mode: int, lanes_per_atom: int = 2, skin: Optional[float] = None,
                   sort_every: int = 0) -> None:
        """pair_mode 0: cell tiles every tick; 1: Verlet lists (skin in length
        units, default 0.12 rc)."""
        t = _lib.Tuning()
        t.pair_mode = pair_mode
        t.lanes_per_atom = lanes_per_atom
        t.skin = self.get_tuning()["skin"] if skin is None else skin
        t.sort_every = sort_every
        _lib.check(self.lib.gs_batch_set_tuning(self.h, ctypes.byref(t)))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gs_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
