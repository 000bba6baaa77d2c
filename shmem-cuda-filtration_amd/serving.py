"""Serving workload: Poisson arrivals of simulation requests driven through
the continuous-batching server of the C ABI (filt_serve_*).

Request stream of the serving configuration: arrivals at rate `rate` per
second for `duration` seconds; per request each grid edge drawn log-uniformly
in [32, 192] (round(32 * 6**U)), the step count uniformly in [100, 2000], the
parameters and fields as in the batched ensemble (workloads.ensemble_member).
Time-to-solution of a request = time its result is back on the host minus its
arrival time.
"""
import ctypes
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import abi, workloads


class Server:
    def __init__(self, backend, precision, reserve_bytes=8 << 30, max_active=256,
                 max_round_steps=8):
        self.lib = backend.lib
        self.h = ctypes.c_void_p()
        abi.check(self.lib, self.lib.filt_serve_create(ctypes.byref(self.h), precision,
                                                       reserve_bytes, max_active,
                                                       max_round_steps), "filt_serve_create")
        self.shapes = {}

    def submit(self, params, P1, S2, perm, poro, nsteps):
        rid = ctypes.c_int()
        abi.check(self.lib, self.lib.filt_serve_submit(
            self.h, ctypes.byref(params), P1.ctypes.data, S2.ctypes.data,
            None if perm is None else perm.ctypes.data,
            None if poro is None else poro.ctypes.data, None, nsteps, ctypes.byref(rid)),
            "filt_serve_submit")
        self.shapes[rid.value] = (params.Nz, params.Ny, params.NX)
        return rid.value

    def round(self):
        a, r = ctypes.c_int(), ctypes.c_int()
        abi.check(self.lib, self.lib.filt_serve_round(self.h, ctypes.byref(a), ctypes.byref(r)),
                  "filt_serve_round")
        return a.value, r.value

    def status(self):
        q, a, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ac, cu = ctypes.c_longlong(), ctypes.c_longlong()
        self.lib.filt_serve_status(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(d),
                                   ctypes.byref(ac), ctypes.byref(cu))
        return {"queued": q.value, "active": a.value, "done": d.value,
                "active_cells": ac.value, "cell_updates": cu.value}

    def is_done(self, rid):
        return bool(self.lib.filt_serve_is_done(self.h, rid))

    def take(self, rid, want_arrays=True):
        shape = self.shapes.pop(rid)
        P = np.empty(shape) if want_arrays else None
        S = np.empty(shape) if want_arrays else None
        m = (ctypes.c_double * 2)()
        abi.check(self.lib, self.lib.filt_serve_take(
            self.h, rid, None if P is None else P.ctypes.data,
            None if S is None else S.ctypes.data, m), "filt_serve_take")
        return P, S, (m[0], m[1])

    def close(self):
        if self.h:
            self.lib.filt_serve_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_stream(rate=20.0, duration=60.0, seed=5005, edge_lo=32, edge_hi=192,
              steps_lo=100, steps_hi=2000):
    """[(arrival_s, shape, nsteps, member_seed)] of the serving configuration."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    ratio = edge_hi / edge_lo
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        shape = tuple(int(round(edge_lo * ratio ** rng.uniform())) for _ in range(3))
        nsteps = int(rng.integers(steps_lo, steps_hi + 1))
        out.append((t, shape, nsteps, int(rng.integers(1, 2 ** 62))))
    return out


def make_request(shape, member_seed, threads=2):
    rng = np.random.default_rng(member_seed)
    m = workloads.ensemble_member(rng, seed=member_seed, shape=shape, threads=threads)
    g = workloads.initial_grid(m["params"], m["perm"], m["poro"])
    return m["params"], g["P1"], g["S2"], m["perm"], m["poro"]


def run(backend, precision, stream, workers=8, lookahead_s=3.0, keep=None, max_round_steps=8):
    """Replay `stream` in real time against a server.  Returns statistics;
    results of request indices in `keep` are returned too."""
    keep = set(keep or ())
    srv = Server(backend, precision, max_round_steps=max_round_steps)
    pool = ThreadPoolExecutor(max_workers=workers)
    futures = {}
    nxt_gen = 0
    t0 = time.perf_counter()
    nxt = 0
    active = {}  # rid -> request index
    arrival, finish, cells_steps = {}, {}, {}
    results = {}
    max_active = max_queue = 0
    lag = 0.0
    try:
        while nxt < len(stream) or active:
            now = time.perf_counter() - t0
            while nxt_gen < len(stream) and stream[nxt_gen][0] <= now + lookahead_s:
                _, shape, _, mseed = stream[nxt_gen]
                futures[nxt_gen] = pool.submit(make_request, shape, mseed)
                nxt_gen += 1
            while nxt < len(stream) and stream[nxt][0] <= now:
                if not futures[nxt].done():
                    break  # input of the next arrival not generated yet
                p, P1, S2, perm, poro = futures.pop(nxt).result()
                lag = max(lag, now - stream[nxt][0])
                rid = srv.submit(p, P1, S2, perm, poro, stream[nxt][2])
                active[rid] = nxt
                arrival[nxt] = stream[nxt][0]
                cells_steps[nxt] = p.NX * p.Ny * p.Nz * stream[nxt][2]
                nxt += 1
            st = srv.status()
            max_active = max(max_active, st["active"])
            max_queue = max(max_queue, st["queued"])
            if st["active"] or st["queued"]:
                srv.round()
                now = time.perf_counter() - t0
                for rid in [r for r in active if srv.is_done(r)]:
                    q = active.pop(rid)
                    P, S, m = srv.take(rid, want_arrays=q in keep)
                    finish[q] = time.perf_counter() - t0
                    if q in keep:
                        results[q] = (P, S, m)
            elif nxt < len(stream):
                time.sleep(min(0.002, max(0.0, stream[nxt][0] - now)))
        total = sum(cells_steps.values())
        tts = np.array([finish[q] - arrival[q] for q in finish])
        span = max(finish.values()) - min(arrival.values())
        return {
            "requests": len(finish),
            "cell_updates": total,
            "makespan_s": span,
            "throughput": total / span,
            "tts_p50_s": float(np.percentile(tts, 50)),
            "tts_p99_s": float(np.percentile(tts, 99)),
            "tts_mean_s": float(tts.mean()),
            "max_active": max_active,
            "max_queued": max_queue,
            "max_input_lag_s": lag,
        }, results
    finally:
        pool.shutdown(wait=True, cancel_futures=True)
        srv.close()
