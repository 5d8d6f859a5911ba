"""Served stream driver: replays a trace of "advance scene s by k steps"
requests in real time against the library's request loop
(prtcl_hip_serve_submit / _tick / _poll).

Every request arrives at its scheduled time; between ticks the driver submits
the requests that have arrived, and each tick advances every scene with
pending work by one step in one batched step (continuous batching: requests
join and leave at step granularity). Scenes are created on their first
request and admitted against the context's HBM budget, evicting idle scenes
(least recently stepped first) when it is short. Latency of a request =
completion time - scheduled arrival time, so queueing counts.
"""

from __future__ import annotations

import time
from typing import Callable, Dict, List, Sequence, Tuple

import numpy as np

from . import hip


class SceneSet:
    """Scenes of a served workload, created on first use."""

    def __init__(self, ctx: hip.Context, make: Callable[[int], object], prepare=True,
                 warm_steps: int = 0):
        self.ctx = ctx
        self.make = make
        self.prepare = prepare
        self.warm_steps = warm_steps  # steps run at admission, before any request
        self.live: Dict[int, hip.HipScene] = {}
        self.nf: Dict[int, int] = {}

    def get(self, index: int) -> hip.HipScene:
        h = self.live.get(index)
        if h is None:
            sc = self.make(index)
            desc = hip.desc_for(sc)
            need = int(self.ctx.L.prtcl_hip_scene_footprint(hip.C.byref(desc)))
            # admission: wait, letting the admitted scenes progress, until
            # idle scenes can be evicted to make room
            while True:
                try:
                    self.ctx.make_room(need)
                    break
                except hip.HipError:
                    if self.ctx.pending()[0] == 0:
                        raise
                    self.ctx.tick()
            h = hip.HipScene(self.ctx, sc)
            if self.prepare:
                h.prepare()
            if self.warm_steps:
                h.step(self.warm_steps)
            self.live[index] = h
            self.nf[index] = sc.n_fluid
        return h


def replay(ctx: hip.Context, scenes: SceneSet, arrivals: Sequence[Tuple[float, int]],
           k: int, on_tick: Callable[[], None] = None) -> Dict[str, object]:
    """Run the trace in real time; returns latencies (s), particle-steps and
    wall time of the run."""
    t0 = time.perf_counter()
    inflight: Dict[int, Tuple[float, int]] = {}
    lat: List[float] = []
    lat_scene: List[int] = []
    tick_ms: List[float] = []
    active: List[int] = []
    iters: List[Tuple[int, int]] = []
    psteps = 0
    ticks = 0
    i = 0
    n = len(arrivals)
    while i < n or inflight:
        now = time.perf_counter() - t0
        while i < n and arrivals[i][0] <= now:
            t_arr, idx = arrivals[i]
            i += 1
            tk = ctx.submit(scenes.get(idx), k)
            inflight[tk] = (t_arr, idx)
        if inflight:
            tt = time.perf_counter()
            nstep = ctx.tick()
            tick_ms.append(1e3 * (time.perf_counter() - tt))
            active.append(nstep)
            ticks += 1
            if on_tick:
                on_tick()
            t_done = time.perf_counter() - t0
            for tk in ctx.poll():
                t_arr, idx = inflight.pop(tk)
                lat.append(t_done - t_arr)
                lat_scene.append(idx)
                psteps += k * scenes.nf[idx]
                info = scenes.live[idx].step_info()  # the request's last step
                iters.append((info["iisph_iterations"], info["wkbb18_iterations"]))
        elif i < n:
            time.sleep(max(0.0, min(arrivals[i][0] - (time.perf_counter() - t0), 0.01)))
    wall = time.perf_counter() - t0
    return {"latencies": np.asarray(lat), "latency_scene": np.asarray(lat_scene),
            "particle_steps": psteps, "wall": wall, "ticks": ticks, "requests": len(lat),
            "tick_ms": np.asarray(tick_ms), "active": np.asarray(active),
            "last_step_iterations": np.asarray(iters, np.int64).reshape(-1, 2)}
