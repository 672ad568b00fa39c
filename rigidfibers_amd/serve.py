"""Serving: independent suspensions stepped together on one GPU.

Host side of the rf_server_* entry points of librigidfibers.so.  The reference
runs one suspension per process (src/main.cu:83-99); here a native worker
thread keeps every admitted suspension resident and advances all of them one
timestep per tick (continuous batching: requests that arrive while others run
are admitted as slots and device memory free up).

``request_stream`` builds the serving workload: suspensions with M drawn
uniformly from a range, each from the cube generator with its side scaled to
a constant number density and its own seed, arriving as a Poisson process.
``serve_stream`` submits them at their arrival times and collects the results;
``serial_stream`` runs the same requests one at a time through Simulation for
comparison.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .config import Parameters
from .scenes import cube
from .simulation import Simulation, make_config


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Server:
    def __init__(self, params: Parameters, integration="numerical", memory_budget=0, max_active=64, eager=False,
                 **kw):
        self._lib = _lib.lib()
        cfg = _lib.RfServerConfig()
        cfg.base = make_config(params, "gmres", integration, **kw)
        cfg.memory_budget = int(memory_budget)
        cfg.max_active = int(max_active)
        cfg.eager = 1 if eager else 0
        handle = C.c_void_p()
        _lib.check(self._lib.rf_server_create(C.byref(cfg), C.byref(handle)))
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, positions, orientations, steps: int) -> int:
        pos = np.ascontiguousarray(positions, dtype=np.float64).ravel()
        ori = np.ascontiguousarray(orientations, dtype=np.float64).ravel()
        if pos.size != ori.size or pos.size % 3:
            raise ValueError("positions and orientations must both hold 3*M values")
        rid = C.c_int64()
        _lib.check(self._lib.rf_server_submit(self._h, pos.size // 3, _dptr(pos), _dptr(ori), int(steps),
                                              C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout: float | None = None) -> bool:
        done = C.c_int32()
        _lib.check(self._lib.rf_server_wait(self._h, rid, -1.0 if timeout is None else float(timeout),
                                            C.byref(done)))
        return bool(done.value)

    def collect(self, rid: int, fibers: int, steps: int) -> dict:
        pos, ori, ms = np.empty(3 * fibers), np.empty(3 * fibers), np.empty(steps)
        info = _lib.RfRequestInfo()
        _lib.check(self._lib.rf_server_collect(self._h, rid, _dptr(pos), _dptr(ori), _dptr(ms), C.byref(info)))
        return {"positions": pos, "orientations": ori, "step_ms": ms,
                "info": {k: getattr(info, k) for k, _ in _lib.RfRequestInfo._fields_}}

    def stats(self) -> dict:
        st = _lib.RfServerStats()
        _lib.check(self._lib.rf_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.RfServerStats._fields_}

    def clock(self) -> float:
        return float(self._lib.rf_server_clock(self._h))


def request_stream(count=512, m_range=(16, 400), density=0.1, rate=50.0, steps=20, seed=0, min_dist=0.4):
    """[(arrival_s, M, positions, orientations, steps)]: Poisson arrivals at `rate`
    per second, M uniform in m_range, cube side (M / density)^(1/3), seed + k per request."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        M = int(rng.integers(m_range[0], m_range[1] + 1))
        pos, ori = cube(M, (M / density) ** (1.0 / 3.0), min_dist=min_dist, seed=seed + 1000 + k)
        out.append((t, M, pos, ori, steps))
    return out


def _pct(values, q):
    return float(np.percentile(np.asarray(values, dtype=np.float64), q)) if len(values) else float("nan")


def summarize(results, wall_s):
    step = [v for r in results for v in r["step_ms"]]
    lat = [1e3 * (r["finished"] - r["arrival"]) for r in results]
    wait = [1e3 * (r["admitted"] - r["arrival"]) for r in results]
    fs = sum(r["fibers"] * r["steps"] for r in results)
    return {"requests": len(results), "fiber_steps": fs, "wall_s": round(wall_s, 3),
            "fiber_steps_per_s": round(fs / wall_s, 1) if wall_s > 0 else None,
            "step_ms_p50": round(_pct(step, 50), 3), "step_ms_p99": round(_pct(step, 99), 3),
            "request_ms_p50": round(_pct(lat, 50), 2), "request_ms_p99": round(_pct(lat, 99), 2),
            "queue_ms_p50": round(_pct(wait, 50), 2), "queue_ms_p99": round(_pct(wait, 99), 2)}


def serve_stream(server: Server, requests, time_scale=1.0):
    """Submits every request at its arrival time (seconds from now, times
    time_scale) and waits for all; returns (results, wall seconds)."""
    start = server.clock()
    t0 = time.perf_counter()
    ids = []
    for arrival, M, pos, ori, steps in requests:
        delay = arrival * time_scale - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append((server.submit(pos, ori, steps), M, steps))
    results = []
    for rid, M, steps in ids:
        server.wait(rid)
        r = server.collect(rid, M, steps)
        info = r["info"]
        results.append({"id": rid, "fibers": M, "steps": steps, "status": info["status"],
                        "arrival": info["submitted"] - start, "admitted": info["admitted"] - start,
                        "finished": info["finished"] - start, "step_ms": list(r["step_ms"]),
                        "gmres_iterations": info["gmres_iterations"],
                        "positions": r["positions"], "orientations": r["orientations"]})
    wall = max(r["finished"] for r in results) if results else 0.0
    return results, wall


def serial_stream(params: Parameters, requests, integration="numerical", time_scale=1.0, **kw):
    """The same requests one at a time (FIFO) through Simulation: a request
    starts at max(arrival, previous finish); step times are wall-clock."""
    t0 = time.perf_counter()
    results = []
    for arrival, M, pos, ori, steps in requests:
        delay = arrival * time_scale - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        start = time.perf_counter() - t0
        p = Parameters(**{**params.__dict__, "NUMBER_OF_FIBERS": M})
        step_ms = []
        with Simulation(p, pos, ori, solver="gmres", integration=integration, **kw) as sim:
            for t in range(steps):
                a = time.perf_counter()
                sim.step(t)
                step_ms.append(1e3 * (time.perf_counter() - a))
            pos_end, ori_end = sim.state()
        results.append({"fibers": M, "steps": steps, "arrival": arrival * time_scale, "admitted": start,
                        "finished": time.perf_counter() - t0, "step_ms": step_ms,
                        "positions": pos_end, "orientations": ori_end})
    wall = results[-1]["finished"] if results else 0.0
    return results, wall
