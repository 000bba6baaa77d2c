"""Benchmark of the filtration time step on one MI355X.

Default workload (the configuration the headline metric is quoted on that
fits one GPU): the plain FP64 path on a 256^3 grid with per-cell log-normal
permeability (sigma 1, correlation length 8 cells) and porosity U[0.35, 0.45],
200 steps of time_step_function as filt_time_step runs it: stage 1, stages
2-4 (u, roS, Newton) in one z-march, the boundary conditions, every array of
ptr_Arrays materialised with the bits the five stages give.  Inputs are
resident in HBM before the timed region.  The five defines.h stage functions
called one by one (the reference driver's calls) are reported beside it.

Prints one JSON line (rank 0):
  value      cell-updates/s of the whole job (Nx*Ny*Nz*steps / wall time)
  roofline   the dominant kernel's algorithmic bytes per launch / its mean
             duration (HIP events around every launch inside the timed region)
             against MI355X's 8 TB/s; traffic from rocprofv3 PMC counters when
             a summary for this workload is committed under profiles/
  cpu_baseline  the reference CPU path (oracle/_ref, compiled from the
             reference sources) or the oracle port, on a bounded sample
"""
import argparse
import ctypes
import importlib
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = importlib.import_module("shmem-cuda-filtration_amd")

HBM_PEAK = 8.0e12
PROFILES = os.path.join(ROOT, "profiles")

# Algorithmic bytes per cell of each plain-path kernel for this workload
# (one medium: the media array is not read; per-cell perm / poro are):
#   ro_P2_Xi: S2, P1, perm in (24); P2, Xi1, Xi2, ro1, ro2 out (40)     = 64
#   u:        Xi1, Xi2, ro1, ro2, P1, P2 in (48); six u out (48)        = 96
#   roS:      S2, P1, P2, Xi1, Xi2, ro1, ro2, poro in (64); 4 out (32)  = 96
#   Newton:   S2, P1, roS1, roS2 in (32); P1, S2 out (16)               = 48
#   u + roS + Newton in one z-march (filt_u_roS_P1_S2_calculation;
#             ro1, ro2 follow from P2 as stage 1 computed them):
#             P2, Xi1, Xi2, S2, poro in (40); six u, four roS,
#             P1, S2 out (96)                                           = 136
PLAIN_BYTES = {"ro_P2_Xi": 64, "u": 96, "roS": 96, "newton": 48, "u_roS_newton": 136,
               "boundary": 0}
STEP_STAGES = ("ro_P2_Xi", "u_roS_newton", "boundary")     # time_step_function
SPLIT_STAGES = ("ro_P2_Xi", "u", "roS", "newton", "boundary")  # the five stage calls
KERNEL_NAMES = {"ro_P2_Xi": "k_ro_P2_Xi", "u": "k_u_zm", "roS": "k_roS_zm", "newton": "k_newton",
                "u_roS_newton": "k_step_zm", "boundary": "k_bc_a+k_bc_b"}


def cpu_baseline(args):
    """Reference CPU path on a bounded sample: two time steps of a
    256 x 256 x 128 slab of the same grid (the reference's rate is the same
    per cell at 256^3; its i-outer / k-inner loop order strides whole
    xy-planes), one thread, about 15-20 s.  Beside it, the oracle port
    (loop-interchanged, OpenMP on up to 16 host threads, bitwise equal to the
    reference) on the same sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    nx, ny, nz = args.n, args.n, max(3, args.n // 2)
    p = PKG.abi.default_params_py()
    p.h1, p.h2, p.h3 = PKG.workloads.INI_H
    p.dt, p.beta1, p.beta2 = 1e-4, 1e-4, 1e-8
    p.NX, p.Ny, p.Nz, p.source = nx, ny, nz, 2
    PKG.abi.finalize_patch_py(p)
    steps = 2
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass

    def timed(impl):
        g = PKG.grid.HostGrid(p)
        impl.initial_data(g, p)
        t0 = time.perf_counter()
        impl.run(g, p, steps)
        return time.perf_counter() - t0, g

    if oracle.ref_available():
        impl, kind = oracle.RefCPU(), "reference"
    else:
        impl, kind = oracle.Oracle(), "port"
        impl.set_loop_order(1)
    dt, g_ref = timed(impl)
    out = {"value": nx * ny * nz * steps / dt, "unit": "cell-updates/s", "cores": 1,
           "kind": kind,
           "sample": f"{steps} steps of a {nx}x{ny}x{nz} slab (media table, same h/dt/beta), "
                     f"{dt:.1f} s on 1 thread of {cpu}"}
    port = oracle.Oracle()
    threads = max(1, min(16, os.cpu_count() or 1))
    port.set_threads(threads)
    port.set_loop_order(0)
    dtp, g_port = timed(port)
    import numpy as np
    out["best_effort_port"] = {
        "value": nx * ny * nz * steps / dtp, "cores": threads,
        "sample": f"same sample, the oracle port with x-fastest loops and OpenMP, {dtp:.2f} s",
        "bitwise_equal_to_reference": bool(np.array_equal(g_port["P1"], g_ref["P1"]) and
                                           np.array_equal(g_port["S2"], g_ref["S2"]))}
    return out


def load_traffic(workload_tag, kernel):
    path = os.path.join(PROFILES, f"pmc_{kernel}_{workload_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def mixed_errors(P64, S64, m64, Pm, Sm, mm):
    """Mixed-form errors against the FP64 form after the same steps."""
    import numpy as np
    return {
        "S2_linf": float(np.max(np.abs(Sm - S64))),
        "P1_rel_linf": float(np.max(np.abs(Pm - P64)) / np.max(np.abs(P64))),
        "water_mass_rel": abs(mm[0] - m64[0]) / m64[0],
        "napl_mass_rel": abs(mm[1] - m64[1]) / m64[1] if m64[1] > 0 else 0.0,
        "total_mass_rel": abs(mm[0] + mm[1] - m64[0] - m64[1]) / (m64[0] + m64[1]),
        "reference": "fused FP64 (bit-identical to the plain path) after the same steps",
    }


def run_slab(args, nx, ny, nz):
    """One large grid, un-batched: a 512 x 512 x 256 slab (sigma 1.5,
    correlation 16 cells, ln K clipped at +-2 sigma), args.slab_steps steps
    through the plain five-stage path, the fused FP64 form and the mixed
    form, so that the fusion and fast-form gain shows apart from batching."""
    import types
    W = PKG.workloads
    t0 = time.perf_counter()
    p, perm, poro = W.c3_workload(nx, ny, nz)
    gen_s = time.perf_counter() - t0
    sub = types.SimpleNamespace(steps=args.slab_steps, warmup=3, autotune=False)
    p, wall, stage_ms, r, finite, tuned = run_plain(sub, p, perm, poro)
    cells = nx * ny * nz
    plain = cells * sub.steps / wall
    out = {"grid": [nx, ny, nz], "steps": sub.steps, "field_generation_s": gen_s,
           "stability_r": r, "state_finite": finite,
           "plain": {"value": plain, "ms_per_step": wall * 1e3 / sub.steps,
                     "stages_ms": stage_ms, "stencil_tiles": tuned}}
    f64, (P64, S64, m64) = run_fused(sub, p, perm, poro, 0)
    mix, (Pm, Sm, mm) = run_fused(sub, p, perm, poro, 1)
    mix["error_vs_fp64"] = mixed_errors(P64, S64, m64, Pm, Sm, mm)
    out["fused_fp64"] = f64
    out["mixed"] = mix
    out["speedup_fused_fp64_vs_plain"] = f64["value"] / plain
    out["speedup_mixed_vs_plain"] = mix["value"] / plain
    return out


def run_plain(args, p, perm, poro, split=False, steps=None):
    """The plain FP64 path on one grid, inputs resident.  Default: the time
    step as filt_time_step runs it (stage 1, stages 2-4 in one z-march, the
    boundary conditions), three launches with HIP events between them on the
    library stream.  split=True: the five defines.h stage functions called
    one after another, as the reference's own driver calls them."""
    from importlib import import_module
    hiprt = import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    r = PKG.workloads.stability(p, perm, poro)
    if r > 2.0:
        raise SystemExit(f"stability guard: r = {r:.3g} > 2 for this workload")
    g = PKG.workloads.initial_grid(p, perm, poro)
    d = be.grid(p, perm=True, poro=True)
    d.upload(g)
    lib = d.lib
    stream = lib.filt_get_stream()
    fref, pref = ctypes.byref(d.f), ctypes.byref(d.params)
    nx = d.nx
    stage1 = lambda t: lib.filt_ro_P2_Xi_calculation(fref, pref, nx, 0, 1)
    bc = lambda t: lib.filt_boundary_conditions(fref, pref, nx, 0, 1)
    if split:
        names = SPLIT_STAGES
        calls = (stage1,
                 lambda t: lib.filt_u_calculation(fref, pref, nx, 0, 1),
                 lambda t: lib.filt_roS_calculation(fref, pref, t, nx, 0, 1),
                 lambda t: lib.filt_P1_S2_calculation(fref, pref, nx, 0, 1),
                 bc)
    else:
        names = STEP_STAGES
        calls = (stage1,
                 lambda t: lib.filt_u_roS_P1_S2_calculation(fref, pref, t, nx, 0, 1, 1),
                 bc)
    best = (ctypes.c_double * 2)()
    if args.autotune and split:
        # stage 1 fills the inputs the stencil kernels read, then the
        # library's autotuner picks tile shapes for this grid
        PKG.abi.check(lib, stage1(0.0), "stage 1")
        PKG.abi.check(lib, lib.filt_autotune_stencil(fref, pref, nx, 5, best), "autotune")
    tuned = {"autotuned": bool(args.autotune and split)}
    for kern, name in ((0, "roS"), (1, "u"), (2, "u_roS_newton")):
        v = [ctypes.c_int() for _ in range(4)]
        lib.filt_get_tile_config(kern, nx, p.Ny, p.Nz, *[ctypes.byref(x) for x in v])
        tuned[name] = {"tile": [v[0].value, v[1].value], "kz": v[2].value, "order": v[3].value}
        if args.autotune and split and kern < 2:
            tuned[name]["tuned_ms"] = best[kern]
    for s in range(1, args.warmup + 1):
        for c in calls:
            PKG.abi.check(lib, c(s * p.dt), "warmup")
    d.sync()
    K = steps or args.steps
    ev = [[hiprt.Event() for _ in range(len(calls) + 1)] for _ in range(K)]
    hiprt.device_synchronize()
    t0 = time.perf_counter()
    for s in range(K):
        t = (args.warmup + s + 1) * p.dt
        e = ev[s]
        e[0].record(stream)
        for q, c in enumerate(calls):
            PKG.abi.check(lib, c(t), "step")
            e[q + 1].record(stream)
    hiprt.device_synchronize()
    wall = time.perf_counter() - t0
    stage_ms = {name: sum(ev[s][q].elapsed_ms(ev[s][q + 1]) for s in range(K)) / K
                for q, name in enumerate(names)}
    P1, S2 = d.get("P1"), d.get("S2")
    d.free()
    import numpy as np
    finite = bool(np.isfinite(P1).all() and np.isfinite(S2).all())
    return p, wall, stage_ms, r, finite, tuned


def run_shim(args, p, perm, poro, steps, fuse):
    """The reference driver's backend calls through the defines.h shims
    (lib/libfiltration_defines.so) in time_step_function's order
    (main.cpp:86-92), each complete on return as in gpu.cu.  fuse: the shims
    defer u and roS and run stages 2-4 as one launch (their default);
    otherwise FILT_FUSE_STAGES=0 and five launches.  The defines.h interface
    has no per-cell K or porosity: the same initial state runs on the first
    medium of the table, so the kernels read the media array instead."""
    from importlib import import_module
    import numpy as np
    shim_mod = import_module(PKG.__name__ + ".defines_shim")
    shim = shim_mod.Shim()
    PKG.device.Backend()
    old = os.environ.get("FILT_FUSE_STAGES")
    os.environ["FILT_FUSE_STAGES"] = "1" if fuse else "0"
    g = PKG.workloads.initial_grid(p, perm, poro)
    c = shim_mod.consts_from_params(p)
    nx = p.NX
    dev, buf = shim_mod.PtrArrays(), ctypes.c_void_p()
    shim.device_memory_alloc(ctypes.byref(dev), ctypes.byref(buf), nx, c)
    if not dev.P1:
        raise SystemExit("defines_shim: device_memory_alloc failed")
    try:
        for n in PKG.abi.FIELD_NAMES:
            a = np.ascontiguousarray(g.a[n])
            shim.load_data_to_device(a.ctypes.data, getattr(dev, n), nx, c)
        media = np.zeros((p.Nz, p.Ny, nx), np.int32)
        shim.load_data_to_device_int(media.ctypes.data, dev.media, nx, c)
        for s in range(1, args.warmup + 1):
            shim.time_step(dev, c, s * p.dt, nx)
        t0 = time.perf_counter()
        for s in range(steps):
            shim.time_step(dev, c, (args.warmup + s + 1) * p.dt, nx)
        wall = time.perf_counter() - t0
        P1 = np.empty((p.Nz, p.Ny, nx))
        S2 = np.empty_like(P1)
        shim.load_data_to_host(P1.ctypes.data, dev.P1, nx, c)
        shim.load_data_to_host(S2.ctypes.data, dev.S2, nx, c)
    finally:
        shim.device_memory_free(dev, buf)
        if old is None:
            os.environ.pop("FILT_FUSE_STAGES", None)
        else:
            os.environ["FILT_FUSE_STAGES"] = old
    cells = p.NX * p.Ny * p.Nz
    return {"value": cells * steps / wall, "ms_per_step": wall * 1e3 / steps,
            "steps": steps, "stages_2_4_fused": fuse,
            "state_finite": bool(np.isfinite(P1).all() and np.isfinite(S2).all())}


FUSED_BYTES = {0: 48, 1: 48, 2: 48}  # P, S in + out (32) + perm, poro in FP64 (16)


def time_ensemble(e, steps, hiprt, lib):
    """Capture `steps` fused steps as one graph, replay it once untimed and
    once timed (HIP events on the library stream).  Returns ms per step."""
    e.capture(steps)
    e.replay()
    stream = lib.filt_get_stream()
    a, b = hiprt.Event(), hiprt.Event()
    hiprt.device_synchronize()
    t0 = time.perf_counter()
    a.record(stream)
    e.replay()
    b.record(stream)
    b.sync()
    wall = time.perf_counter() - t0
    return a.elapsed_ms(b) / steps, wall


def run_fused(args, p, perm, poro, precision):
    """The same workload through the fused one-launch-per-step form."""
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    g = PKG.workloads.initial_grid(p, perm, poro)
    e = be.ensemble(precision)
    try:
        e.add(p, g["P1"], g["S2"], perm, poro)
        tiling = e.autotune(3)
        ms, wall = time_ensemble(e, args.steps, hiprt, be.lib)
    finally:
        e.close()
    # accuracy run: exactly args.steps steps from the initial state
    e = be.ensemble(precision)
    try:
        mid = e.add(p, g["P1"], g["S2"], perm, poro)
        e.apply_tiling(tiling)
        e.step(args.steps)
        P, S = e.get(mid)
        m = e.mass(mid)
    finally:
        e.close()
    cells = p.NX * p.Ny * p.Nz
    return {"value": cells / (ms * 1e-3), "ms_per_step": ms, "tiling": tiling,
            "bytes_per_cell_update": FUSED_BYTES[precision],
            "hbm_GBps": FUSED_BYTES[precision] * cells / (ms * 1e-3) / 1e9,
            "wall_s": wall}, (P, S, m)


def run_c4(args):
    """Batched ensemble: args.c4 realizations with uneven grids, args.c4_steps
    steps.  Plain FP64 path run realization by realization (all 5 stages,
    all arrays, default tiles) against the batched forms (every member in one
    launch per step, the whole run one hipGraph)."""
    import numpy as np
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    lib = be.lib
    stream = lib.filt_get_stream()
    W = PKG.workloads
    t0 = time.perf_counter()
    members = W.c4_members(args.c4)
    gen_s = time.perf_counter() - t0
    cells = sum(m["params"].NX * m["params"].Ny * m["params"].Nz for m in members)
    steps = args.c4_steps
    out = {"realizations": len(members), "cells": cells, "steps": steps,
           "field_generation_s": gen_s,
           "sigma_range": [min(m["sigma"] for m in members), max(m["sigma"] for m in members)],
           "max_stability_r": max(m["stability_r"] for m in members)}
    # plain, one realization after another (inputs resident first)
    grids = []
    for m in members:
        g = W.initial_grid(m["params"], m["perm"], m["poro"])
        d = be.grid(m["params"], perm=True, poro=True)
        d.upload(g)
        grids.append(d)
    be.sync()
    a, b = hiprt.Event(), hiprt.Event()
    a.record(stream)
    for d in grids:
        for s in range(1, steps + 1):
            PKG.abi.check(lib, lib.filt_time_step(ctypes.byref(d.f), ctypes.byref(d.params),
                                                  s * d.params.dt, d.nx, 0, 1, 1), "plain")
    b.record(stream)
    b.sync()
    plain_ms = a.elapsed_ms(b)
    for d in grids:
        d.free()
    out["plain_per_realization"] = {"value": cells * steps / (plain_ms * 1e-3),
                                    "total_ms": plain_ms}
    results = {}
    for prec, name in ((0, "batched_fp64"), (1, "batched_fast"), (2, "batched_mixed")):
        e = be.ensemble(prec)
        ids = []
        for m in members:
            g = W.initial_grid(m["params"], m["perm"], m["poro"])
            ids.append(e.add(m["params"], g["P1"], g["S2"], m["perm"], m["poro"]))
        tiling = e.autotune(2)
        e.close()
        # fresh ensemble: the timed run starts from the initial states
        e = be.ensemble(prec)
        ids = []
        for m in members:
            g = W.initial_grid(m["params"], m["perm"], m["poro"])
            ids.append(e.add(m["params"], g["P1"], g["S2"], m["perm"], m["poro"]))
        e.apply_tiling(tiling)
        e.capture(steps)
        be.sync()
        a.record(stream)
        e.replay()
        b.record(stream)
        b.sync()
        ms = a.elapsed_ms(b)
        out[name] = {"value": cells * steps / (ms * 1e-3), "total_ms": ms, "tiling": tiling,
                     "bytes_per_cell_update": FUSED_BYTES[prec],
                     "hbm_GBps": FUSED_BYTES[prec] * cells * steps / (ms * 1e-3) / 1e9}
        results[prec] = [(e.get(mid), e.mass(mid)) for mid in ids]
        e.close()
    errs = {"S2_linf": 0.0, "P1_rel_linf": 0.0, "water_mass_rel": 0.0, "napl_mass_rel": 0.0,
            "total_mass_rel": 0.0}
    for ((P64, S64), m64), ((Pm, Sm), mm) in zip(results[0], results[1]):
        errs["S2_linf"] = max(errs["S2_linf"], float(np.max(np.abs(Sm - S64))))
        errs["P1_rel_linf"] = max(errs["P1_rel_linf"],
                                  float(np.max(np.abs(Pm - P64)) / np.max(np.abs(P64))))
        errs["water_mass_rel"] = max(errs["water_mass_rel"], abs(mm[0] - m64[0]) / m64[0])
        if m64[1] > 0:
            errs["napl_mass_rel"] = max(errs["napl_mass_rel"], abs(mm[1] - m64[1]) / m64[1])
        errs["total_mass_rel"] = max(errs["total_mass_rel"],
                                     abs(sum(mm) - sum(m64)) / sum(m64))
    out["batched_mixed"]["error_vs_fp64_max_over_members"] = errs
    out["speedup_batched_mixed_vs_plain"] = (out["batched_mixed"]["value"] /
                                             out["plain_per_realization"]["value"])
    out["speedup_batched_fp64_vs_plain"] = (out["batched_fp64"]["value"] /
                                            out["plain_per_realization"]["value"])
    return out


def run_serving(args):
    """Serving configuration: Poisson arrivals (rate args.serve_rate per s for
    args.serve_duration s), edges log-uniform in [32, 192], steps U[100, 2000],
    continuous batching through filt_serve_*.  FP64 (the fused form, whose
    results equal the plain path's bit for bit) against mixed precision on the
    same request stream."""
    serving = importlib.import_module(PKG.__name__ + ".serving")
    be = PKG.device.Backend()
    stream = serving.c5_stream(rate=args.serve_rate, duration=args.serve_duration)
    out = {"rate_per_s": args.serve_rate, "window_s": args.serve_duration,
           "requests": len(stream),
           "offered_cell_updates_per_s": sum(s[1][0] * s[1][1] * s[1][2] * s[2] for s in stream)
           / args.serve_duration}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        stats, _ = serving.run(be, prec, stream)
        out[name] = stats
    return out


def stream_mix_ceiling(kernel, cells):
    """Time of a pure stream with the kernel's mix of FP64 arrays
    (tests/hip/bin/stream_mix, built from tests/hip/stream_mix.hip): the fastest any kernel reading and
    writing those arrays once can run on this GPU.  Best of plain and
    non-temporal stores; None when the tool is missing or has no such mix."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "hip", "bin", "stream_mix")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe, str(cells), kernel], capture_output=True, text=True,
                         timeout=120, check=True).stdout
    rows = [json.loads(x) for x in out.splitlines() if x.startswith("{")]
    if not rows:
        return None
    best = min(rows, key=lambda r: r["ms"])
    return {"mix": f"{best['in']} in / {best['out']} out FP64 arrays",
            "bytes_per_cell": best["bytes_per_cell"], "ms": best["ms"],
            "GBps": best["TBps"] * 1e3, "nontemporal_stores": best["nontemporal_stores"]}


def copy_calibration(n_bytes=1 << 30, reps=5):
    """The library's streaming copy kernel over n_bytes: the achievable HBM
    ceiling, and the reference point that scales the PMC byte counters."""
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    PKG.device.Backend()
    lib = PKG.abi.load_library()
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    PKG.abi.check(lib, lib.filt_device_malloc(ctypes.byref(src), n_bytes), "malloc")
    PKG.abi.check(lib, lib.filt_device_malloc(ctypes.byref(dst), n_bytes), "malloc")
    stream = lib.filt_get_stream()
    PKG.abi.check(lib, lib.filt_device_copy(dst, src, n_bytes), "copy")
    times = []
    for _ in range(reps):
        a, b = hiprt.Event(), hiprt.Event()
        a.record(stream)
        PKG.abi.check(lib, lib.filt_device_copy(dst, src, n_bytes), "copy")
        b.record(stream)
        b.sync()
        times.append(a.elapsed_ms(b))
    lib.filt_device_free_ptr(src)
    lib.filt_device_free_ptr(dst)
    best = min(times)
    return {"copy_bytes_moved": 2 * n_bytes, "copy_ms": best,
            "copy_GBps": 2 * n_bytes / (best * 1e-3) / 1e9,
            "copy_frac_of_peak": 2 * n_bytes / (best * 1e-3) / HBM_PEAK}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="grid edge")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--c4", type=int, default=96,
                    help="realizations of the batched-ensemble leg (0: skip)")
    ap.add_argument("--c4-steps", type=int, default=300)
    ap.add_argument("--slab-steps", type=int, default=100,
                    help="steps of the 512x512x256 and 256x512x512 slab legs (0: skip)")
    ap.add_argument("--serve-duration", type=float, default=20.0,
                    help="arrival window of the serving leg in seconds (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--plain-only", action="store_true",
                    help="skip the fused FP64 / mixed forms and the defines.h shim legs")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the stencil tile shapes of the plain path on this grid "
                         "(default: the library's built-in shapes)")
    ap.add_argument("--copy", action="store_true",
                    help="only run the copy-kernel calibration (1 GiB) and print it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        raise SystemExit(2)
    if args.copy:
        print(json.dumps(copy_calibration()), flush=True)
        return

    p, perm, poro = PKG.workloads.c2_workload(args.n)
    p, wall, stage_ms, r, finite, tuned = run_plain(args, p, perm, poro)
    forms = {}
    if not args.plain_only:
        import numpy as np
        f64, (P64, S64, m64) = run_fused(args, p, perm, poro, 0)
        fast, (Pf, Sf, mf) = run_fused(args, p, perm, poro, 1)
        fast["error_vs_fp64"] = mixed_errors(P64, S64, m64, Pf, Sf, mf)
        mix, (Pm, Sm, mm) = run_fused(args, p, perm, poro, 2)
        mix["error_vs_fp64"] = mixed_errors(P64, S64, m64, Pm, Sm, mm)
        forms = {"fused_fp64": f64, "fast": fast, "mixed": mix}
        if args.slab_steps > 0:
            for shape in ((512, 512, 256), (256, 512, 512)):
                forms["slab_%dx%dx%d" % shape] = run_slab(args, *shape)
        if args.c4 > 0:
            forms["batched_ensemble"] = run_c4(args)
        if args.serve_duration > 0:
            forms["serving"] = run_serving(args)
    cells = p.NX * p.Ny * p.Nz
    value = cells * args.steps / wall
    dom = max((s for s in STEP_STAGES if PLAIN_BYTES[s]), key=lambda s: stage_ms[s])
    dom_bytes = PLAIN_BYTES[dom] * cells
    achieved = dom_bytes / (stage_ms[dom] * 1e-3)
    tag = f"plain_{args.n}cube"
    traffic = load_traffic(tag, KERNEL_NAMES[dom])
    mix_ceiling = stream_mix_ceiling(KERNEL_NAMES[dom], cells)
    if mix_ceiling:
        mix_ceiling["kernel_frac_of_mix"] = mix_ceiling["ms"] / stage_ms[dom]
    step_bytes = sum(PLAIN_BYTES[s] for s in STEP_STAGES) * cells
    split_bytes = sum(PLAIN_BYTES[s] for s in SPLIT_STAGES) * cells
    # the five stage functions one by one (the reference driver's calls)
    _, wall5, stage5, _, finite5, _ = run_plain(args, p, perm, poro, split=True,
                                                  steps=max(1, args.steps // 2))
    split = {"value": cells * max(1, args.steps // 2) / wall5,
             "ms_per_step": wall5 * 1e3 / max(1, args.steps // 2),
             "stages_ms": stage5, "state_finite": finite5,
             "bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in SPLIT_STAGES),
             "hbm_frac": split_bytes * max(1, args.steps // 2) / wall5 / HBM_PEAK}
    # the same calls through the defines.h shims, as the reference driver makes
    # them (another workload: media table instead of per-cell K and porosity,
    # so not under --plain-only, whose launches the kernel profiles average)
    shim_steps = max(1, args.steps // 2)
    shims = None
    if not args.plain_only:
        shims = {"fused": run_shim(args, p, perm, poro, shim_steps, True),
                 "separate": run_shim(args, p, perm, poro, shim_steps, False)}
    line = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP64, {args.n}^3, time_step_function per step (stage 1, "
                        "stages 2-4 fused, boundary; every ptr_Arrays array materialised), "
                        "per-cell log-normal K (sigma 1, corr 8 cells) and porosity",
            "grid": [p.NX, p.Ny, p.Nz],
            "form": "plain",
            "deviations": "beta1=1e-4, beta2=1e-8 (stated 4.4e-10 is unstable); "
                          "defines.ini spacing",
            "stability_r": r,
            "stencil_tiles": tuned,
            "state_finite": finite,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL_NAMES[dom],
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes,
            "stream_mix_ceiling": mix_ceiling,
        },
        "stages_ms": stage_ms,
        "step_roofline": {
            "bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in STEP_STAGES),
            "achieved_GBps": step_bytes * args.steps / wall / 1e9,
            "frac": step_bytes * args.steps / wall / HBM_PEAK,
            "vs_split_model": {"bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in SPLIT_STAGES),
                               "frac": split_bytes * args.steps / wall / HBM_PEAK},
        },
        "stage_calls": split,
        "defines_shims": shims,
        "forms": forms,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
