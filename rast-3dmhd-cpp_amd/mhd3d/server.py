"""Request server (mhd3d_server_* of include/mhd3d.h) and the serving loop.

The server admits runs first-come-first-served while their device memory fits
an HBM budget, keeps their state on the device and advances all admitted runs
together, a quantum of steps per tick, each run stopping on the device at its
own step count (continuous batching); runs whose state turns non-finite are
evicted.  `serve_trace`
replays a timed arrival trace against it and reports completion latencies and
aggregate throughput; `sequential_trace` runs the same trace one run at a
time in the plain form, the baseline.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from .desc import Descriptor
from .domain import Domain
from .initial import stratified_profiles


class Server:
    def __init__(self, precision=0, hbm_budget_bytes=64 << 30, quantum_steps=10):
        h = C.c_void_p()
        rc = L.lib().mhd3d_server_create(int(precision), int(hbm_budget_bytes), int(quantum_steps), C.byref(h))
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_server_last_error(None).decode())
        self.h = h
        self.precision = precision
        self._descs = {}

    def close(self):
        if getattr(self, "h", None):
            L.lib().mhd3d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_server_last_error(self.h).decode())

    def now(self):
        return float(L.lib().mhd3d_server_now(self.h))

    @property
    def bytes_in_use(self):
        return int(L.lib().mhd3d_server_bytes_in_use(self.h))

    @property
    def cell_steps(self):
        return int(L.lib().mhd3d_server_cell_steps(self.h))

    def submit(self, desc: Descriptor, nsteps, seed, amp=1e-4, b0=1e-3, b_amp=1e-5):
        T, rho = stratified_profiles(desc)
        T = np.ascontiguousarray(T)
        rho = np.ascontiguousarray(rho)
        rid = C.c_int64()
        dp = C.POINTER(C.c_double)
        self._check(L.lib().mhd3d_server_submit(self.h, C.cast(desc.ptr(), C.c_void_p), T.ctypes.data_as(dp),
                                                rho.ctypes.data_as(dp), int(seed), float(amp), float(b0),
                                                float(b_amp), int(nsteps), C.byref(rid)))
        self._descs[rid.value] = desc
        return rid.value

    def tick(self):
        na, nq = C.c_int(), C.c_int()
        self._check(L.lib().mhd3d_server_tick(self.h, C.byref(na), C.byref(nq)))
        return na.value, nq.value

    def poll(self, rid):
        info = L.RequestInfo()
        self._check(L.lib().mhd3d_server_poll(self.h, int(rid), C.byref(info)))
        return {"status": info.status, "steps_done": info.steps_done, "t_submit": info.t_submit,
                "t_admit": info.t_admit, "t_finish": info.t_finish, "dt_last": info.dt_last}

    def result(self, rid, fetch=True):
        desc = self._descs[rid]
        nf = 8 if desc.cfg.lMag else 5
        if not fetch:
            self._check(L.lib().mhd3d_server_result(self.h, int(rid), None))
            return None
        arrs = [np.empty(desc.shape) for _ in range(nf)]
        self._check(L.lib().mhd3d_server_result(self.h, int(rid), L._fieldptrs(arrs)))
        return dict(zip(L.FIELDS[:nf], arrs))


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else float("nan")


def serve_trace(specs, precision=0, hbm_budget_bytes=64 << 30, quantum_steps=10, fetch=False,
                time_limit=600.0):
    """Replays `specs` (RunSpec with .arrival in seconds and .steps) against a
    Server.  Latency of a run = its finish time - its arrival time."""
    srv = Server(precision, hbm_budget_bytes, quantum_steps)
    descs = [Descriptor(s.config(precision)) for s in specs]
    order = sorted(range(len(specs)), key=lambda i: specs[i].arrival)
    t0 = srv.now()
    nxt = 0
    ids = {}
    open_ids = set()
    lat = {}
    results = {}
    failed = []
    evicted = []
    exited = []
    peak_bytes = 0
    while len(lat) < len(specs):
        now = srv.now() - t0
        if now > time_limit:
            raise TimeoutError("serving trace exceeded its time limit")
        while nxt < len(order) and specs[order[nxt]].arrival <= now:
            i = order[nxt]
            rid = srv.submit(descs[i], specs[i].steps, specs[i].seed)
            ids[rid] = i
            open_ids.add(rid)
            nxt += 1
        na, nq = srv.tick()
        peak_bytes = max(peak_bytes, srv.bytes_in_use)
        for rid in list(open_ids):
            info = srv.poll(rid)
            if info["status"] in (L.REQ_DONE, L.REQ_FAILED, L.REQ_NONFINITE, L.REQ_EXITED):
                i = ids[rid]
                lat[i] = info["t_finish"] - t0 - specs[i].arrival
                results[i] = srv.result(rid, fetch=fetch and info["status"] == L.REQ_DONE)
                open_ids.discard(rid)
                if info["status"] == L.REQ_FAILED:
                    failed.append(i)
                elif info["status"] == L.REQ_NONFINITE:
                    evicted.append(i)
                elif info["status"] == L.REQ_EXITED:
                    exited.append(i)
        if na == 0 and nq == 0 and nxt < len(order):
            wait = specs[order[nxt]].arrival - (srv.now() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.05))
    t_end = max(specs[i].arrival + lat[i] for i in lat)
    t_begin = min(s.arrival for s in specs)
    cell_steps = srv.cell_steps   # steps actually taken (halted runs stop early)
    srv.close()
    v = [lat[i] for i in range(len(specs))]
    return {"p50_s": _percentile(v, 50), "p99_s": _percentile(v, 99), "max_s": max(v),
            "mcell_steps_per_s": cell_steps / (t_end - t_begin) / 1e6, "makespan_s": t_end - t_begin,
            "cell_steps": cell_steps, "peak_device_bytes": peak_bytes, "failed": sorted(failed),
            "evicted_nonfinite": sorted(evicted), "exited": sorted(exited), "latencies": v,
            "results": results}


def sequential_trace(specs, precision=0):
    """The same runs one at a time in arrival order (plain domains, device
    initial state): per-run times (context creation, initial state and
    steps), then the FIFO latencies those imply for the trace's arrival
    times."""
    order = sorted(range(len(specs)), key=lambda i: specs[i].arrival)
    busy_until = 0.0
    failed = []
    cell_steps = 0
    lat = [0.0] * len(specs)
    total = 0.0
    for i in order:
        s = specs[i]
        desc = Descriptor(s.config(precision))
        t = time.perf_counter()
        dom = Domain(desc)
        dom.init_stratified(seed=s.seed)
        dom.set_dt_guard()   # the reference driver's guard, as the server applies
        try:
            dom.step(s.steps)
        except L.MhdError as err:
            if err.code == L.EHALTED:
                failed.append(i)
            elif err.code != L.ENONFINITE:
                raise
        dom.synchronize()
        cell_steps += s.cells * dom.clock()[0]
        dur = time.perf_counter() - t
        dom.close()
        total += dur
        start = max(s.arrival, busy_until)
        busy_until = start + dur
        lat[i] = busy_until - s.arrival
    return {"p50_s": _percentile(lat, 50), "p99_s": _percentile(lat, 99), "max_s": max(lat),
            "mcell_steps_per_s": cell_steps / total / 1e6, "compute_s": total, "cell_steps": cell_steps,
            "failed": sorted(failed), "latencies": lat}
