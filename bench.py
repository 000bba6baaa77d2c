"""Benchmark of the filtration time step on one MI355X.

Default workload (the configuration the headline metric is quoted on that
fits one GPU): the plain FP64 path on a 256^3 grid with per-cell log-normal
permeability (sigma 1, correlation length 8 cells) and porosity U[0.35, 0.45],
200 steps of time_step_function as filt_time_step runs it, every array of
ptr_Arrays materialised with the bits the five stages give.  Inputs are
resident in HBM before the timed region.

Beside the headline the JSON line carries the other configurations:
  stage_calls / defines_shims  the five defines.h functions called one by one,
                 and the reference driver's calls through the defines.h shims;
  forms          the fused FP64, fast and mixed forms on the same grid;
  oracle_check   the plain path and every form against the oracle port run
                 on all host threads this process may use (cell for cell);
  slab_*         one rank's share (512 x 512 x 256 and its transpose): plain
                 vs fused FP64 / fast / mixed, un-batched; the 512 x 512 x 256
                 slab also against the oracle with its cumulative mass-balance
                 error every 10 steps;
  batched_ensemble  96 uneven realizations, 300 steps, every form batched in
                 one graph against the plain path run realization by
                 realization;
  serving        Poisson arrivals at 20 requests/s for 60 s through the
                 continuous-batching server, FP64 vs mixed;
  divgrad        the reference's tau-relaxed divgrad scheme in place of the
                 upwind roS stage: the five stage calls per step, the divgrad
                 kernel's roofline, 3 steps against the oracle port;
  io             checkpoint (save/save.dat) and Tecplot output from a 128^3
                 grid on the device, and a snapshot written in the background
                 while the steps go on;
  cpu_baseline   the reference's own CPU path (oracle/_ref) on one thread on
                 a bounded sample of the headline workload, and on
                 the homogeneous 32^3 cube in full; the oracle port on all threads.
Progress goes to stderr; the one JSON line to stdout.
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
T0 = time.perf_counter()

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
#   the whole step but the boundary in one z-march (k_step_full):
#             P1, S2, perm, poro in (32); P2, Xi1, Xi2, ro1, ro2, six u,
#             four roS, P1, S2 out (136)                                = 168
#   the stash copy (k_unstash) moves no algorithmic bytes: it is what the
#             one-kernel form pays for its in-place writes
PLAIN_BYTES = {"ro_P2_Xi": 64, "u": 96, "roS": 96, "newton": 48, "u_roS_newton": 136,
               "step_full": 168, "unstash": 0, "boundary": 0}
STEP_STAGES = ("step_full", "unstash", "boundary")   # filt_time_step (lambda = 2)
TWO_KERNEL_STAGES = ("ro_P2_Xi", "u_roS_newton", "boundary")  # filt_time_step otherwise
SPLIT_STAGES = ("ro_P2_Xi", "u", "roS", "newton", "boundary")  # the five stage calls
KERNEL_NAMES = {"ro_P2_Xi": "k_ro_P2_Xi", "u": "k_u_zm", "roS": "k_roS_zm", "newton": "k_newton",
                "u_roS_newton": "k_step_zm", "step_full": "k_step_full", "unstash": "k_unstash",
                "boundary": "k_bc_a+k_bc_b"}
# k_roS_divgrad_zm: ro1, ro2, S2, poro, u1x, u1y, u2x, u2y, roS1_old, roS2_old in;
# roS1_old, roS2_old, roS1, roS2 out
DIVGRAD_BYTES = 112
FORMS = {0: "fused_fp64", 1: "fast", 2: "mixed"}
# P, S in + out (32) + perm, poro in FP64 (16), every form
FUSED_BYTES = {0: 48, 1: 48, 2: 48}


def log(msg):
    print(f"[bench {time.perf_counter() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def host_info():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"cpu_model": cpu, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def host_threads():
    """Threads for the CPU legs: the CPUs this process may use, at most 16
    (the host share of one GPU)."""
    return max(1, min(16, host_info()["affinity_cpus"]))


def form_errors(P, S, m, Pr, Sr, mr, reference):
    """Per-field errors of a result (P, S, masses m) against a reference
    state: S2 L-inf, P1 and S2 relative L-inf, water / NAPL / total mass."""
    import numpy as np
    return {
        "S2_linf": float(np.max(np.abs(S - Sr))),
        "S2_rel_linf": float(np.max(np.abs(S - Sr)) / np.max(np.abs(Sr))),
        "P1_rel_linf": float(np.max(np.abs(P - Pr)) / np.max(np.abs(Pr))),
        "water_mass_rel": abs(m[0] - mr[0]) / mr[0],
        "napl_mass_rel": abs(m[1] - mr[1]) / mr[1] if mr[1] > 0 else abs(m[1]),
        "total_mass_rel": abs(m[0] + m[1] - mr[0] - mr[1]) / (mr[0] + mr[1]),
        "reference": reference,
    }


def mixed_errors(P64, S64, m64, Pm, Sm, mm):
    """A form's errors against the fused FP64 form after the same steps."""
    return form_errors(Pm, Sm, mm, P64, S64, m64,
                       "fused FP64 (bit-identical to the plain path) after the same steps")


def mass_series_errors(series, ref_series, every):
    return [{"step": every * (q + 1),
             "water": abs(m[0] - r[0]) / r[0],
             "napl": abs(m[1] - r[1]) / r[1] if r[1] else abs(m[1]),
             "total": abs(m[0] + m[1] - r[0] - r[1]) / (r[0] + r[1])}
            for q, (m, r) in enumerate(zip(series, ref_series))]


def load_traffic(workload_tag, kernel):
    path = os.path.join(PROFILES, f"pmc_{kernel}_{workload_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


# ---------------------------------------------------------------- plain --

def run_plain(args, p, perm, poro, split=False, steps=None):
    """The plain FP64 path on one grid, inputs resident, timed.  Default: the
    time step as filt_time_step runs it: stage 1, then stages 2-4 in one
    z-march, with the boundary conditions folded into those two launches
    where the grid allows it (one rank, every extent >= 3; the third stage
    interval then holds no launch), the stages 2-4 kernel's tile chosen for
    the grid by the library's autotuner before the warmup.  split=True: the
    five defines.h stage
    functions called one after another, as the reference's own driver calls
    them."""
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
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
    path = ctypes.c_int(-1)
    if split:
        names = SPLIT_STAGES
        calls = (stage1,
                 lambda t: lib.filt_u_calculation(fref, pref, nx, 0, 1),
                 lambda t: lib.filt_roS_calculation(fref, pref, t, nx, 0, 1),
                 lambda t: lib.filt_P1_S2_calculation(fref, pref, nx, 0, 1),
                 bc)
    else:
        # filt_time_step, with HIP events recorded between its launches
        PKG.abi.check(lib, lib.filt_time_step_timed(fref, pref, 0.0, nx, 0, 1, 1, None,
                                                    ctypes.byref(path)), "path probe")
        names = STEP_STAGES if path.value == 1 else TWO_KERNEL_STAGES
        calls = None
        step_best = ctypes.c_double()
        if path.value == 0 and args.step_tile:
            t, kz, order = args.step_tile.split(",")
            tx, ty = (int(v) for v in t.split("x"))
            PKG.abi.check(lib, lib.filt_set_tile_config(2, tx, ty, int(kz), int(order)),
                          "step tile")
        elif path.value == 0:
            # the library's autotuner picks the stages 2-4 kernel's tile for
            # this grid (it times the kernel's variant without the folded
            # boundary, so the time step's own kernel launches only here)
            PKG.abi.check(lib, lib.filt_autotune_step(fref, pref, nx, 5, ctypes.byref(step_best)),
                          "step autotune")
    best = (ctypes.c_double * 2)()
    if split and getattr(args, "stencil_tile", None):
        t, kz, order = args.stencil_tile.split(",")
        tx, ty = (int(v) for v in t.split("x"))
        for kern in (0, 1):
            PKG.abi.check(lib, lib.filt_set_tile_config(kern, tx, ty, int(kz), int(order)),
                          "stencil tile")
    elif args.autotune and split:
        # stage 1 fills the inputs the stencil kernels read, then the
        # library's autotuner picks tile shapes for this grid
        PKG.abi.check(lib, stage1(0.0), "stage 1")
        PKG.abi.check(lib, lib.filt_autotune_stencil(fref, pref, nx, 5, best), "autotune")
    tuned = {"autotuned": bool(args.autotune and split),
             "step_autotuned": bool(not split and path.value == 0 and not args.step_tile)}
    for kern, name in ((0, "roS"), (1, "u"), (2, "u_roS_newton"), (3, "step_full")):
        v = [ctypes.c_int() for _ in range(4)]
        lib.filt_get_tile_config(kern, nx, p.Ny, p.Nz, *[ctypes.byref(x) for x in v])
        tuned[name] = {"tile": [v[0].value, v[1].value], "kz": v[2].value, "order": v[3].value}
        if args.autotune and split and kern < 2:
            tuned[name]["tuned_ms"] = best[kern]
        if kern == 2 and not split and path.value == 0 and not args.step_tile:
            tuned[name]["tuned_ms"] = step_best.value
    def step_timed(t, ev):
        arr = (ctypes.c_void_p * 4)(*[e.h.value for e in ev])
        PKG.abi.check(lib, lib.filt_time_step_timed(fref, pref, t, nx, 0, 1, 1, arr, None),
                      "step")

    for s in range(1, args.warmup + 1):
        if calls:
            for c in calls:
                PKG.abi.check(lib, c(s * p.dt), "warmup")
        else:
            PKG.abi.check(lib, lib.filt_time_step(fref, pref, s * p.dt, nx, 0, 1, 1), "warmup")
    d.sync()
    K = steps or args.steps
    # The timed region: K steps, wall clock.  A recorded HIP event costs the
    # stream a few microseconds, so inside it only the dominant kernel of
    # filt_time_step is bracketed (two events per step: its measured
    # duration for the roofline); the stage calls record none.  The per-stage
    # breakdown comes from a separate pass after the timed region, with an
    # event at every stage boundary.
    dom_q = None if calls else (0 if path.value == 1 else 1)
    dev = [[hiprt.Event() for _ in range(2)] for _ in range(K)] if dom_q is not None else None
    hiprt.device_synchronize()
    t0 = time.perf_counter()
    for s in range(K):
        t = (args.warmup + s + 1) * p.dt
        if calls:
            for c in calls:
                PKG.abi.check(lib, c(t), "step")
        else:
            arr = (ctypes.c_void_p * 4)()
            arr[dom_q], arr[dom_q + 1] = dev[s][0].h.value, dev[s][1].h.value
            PKG.abi.check(lib, lib.filt_time_step_timed(fref, pref, t, nx, 0, 1, 1, arr, None),
                          "step")
    hiprt.device_synchronize()
    wall = time.perf_counter() - t0
    dominant = None
    if dom_q is not None:
        dominant = {"stage": names[dom_q],
                    "ms": sum(dev[s][0].elapsed_ms(dev[s][1]) for s in range(K)) / K}
    # per-stage breakdown, after the timed region
    B = min(K, 50)
    nev = len(calls) + 1 if calls else 4
    ev = [[hiprt.Event() for _ in range(nev)] for _ in range(B)]
    for s in range(B):
        t = (args.warmup + K + s + 1) * p.dt
        e = ev[s]
        if calls:
            e[0].record(stream)
            for q, c in enumerate(calls):
                PKG.abi.check(lib, c(t), "step")
                e[q + 1].record(stream)
        else:
            step_timed(t, e)
    hiprt.device_synchronize()
    stage_ms = {name: sum(ev[s][q].elapsed_ms(ev[s][q + 1]) for s in range(B)) / B
                for q, name in enumerate(names)}
    P1, S2 = d.get("P1"), d.get("S2")
    d.free()
    import numpy as np
    finite = bool(np.isfinite(P1).all() and np.isfinite(S2).all())
    return {"wall": wall, "steps": K, "stage_ms": stage_ms, "stability_r": r, "finite": finite,
            "tiles": tuned, "names": names, "dominant": dominant, "breakdown_steps": B}


def plain_state(p, perm, poro, steps, every=None):
    """The plain path's state after exactly `steps` steps from the initial
    state (u included), with its device mass; with `every`, the masses after
    every `every` steps."""
    be = PKG.device.Backend()
    d = be.grid(p, perm=True, poro=True)
    try:
        d.upload(PKG.workloads.initial_grid(p, perm, poro))
        series, done = [], 0
        while done < steps:
            k = min(every or steps, steps - done)
            d.run(k, step0=done, with_u=1)
            done += k
            if every:
                series.append(d.mass())
        return d.get("P1"), d.get("S2"), d.mass(), series
    finally:
        d.free()


def run_shim(args, p, perm, poro, steps, fuse):
    """The reference driver's backend calls through the defines.h shims
    (lib/libfiltration_defines.so) in time_step_function's order
    (main.cpp:86-92), each complete on return as in gpu.cu.  fuse: the shims
    defer u and roS and run stages 2-4 as one launch (their default);
    otherwise FILT_FUSE_STAGES=0 and five launches.  The defines.h interface
    has no per-cell K or porosity: the same initial state runs on the first
    medium of the table, so the kernels read the media array instead."""
    import numpy as np
    shim_mod = importlib.import_module(PKG.__name__ + ".defines_shim")
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


# ---------------------------------------------------------------- fused --

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


def run_fused(args, p, perm, poro, precision, steps=None, every=None):
    """One grid through a fused one-launch-per-step form: the autotuner picks
    the tiling, a graph of `steps` steps is timed; then an accuracy run of
    exactly `steps` steps from the initial state (masses every `every`
    steps when given)."""
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    steps = steps or args.steps
    g = PKG.workloads.initial_grid(p, perm, poro)
    e = be.ensemble(precision)
    try:
        e.add(p, g["P1"], g["S2"], perm, poro)
        tiling = e.autotune(7)  # the medians of 7 launches: close tilings differ by ~2%
        ms, wall = time_ensemble(e, steps, hiprt, be.lib)
    finally:
        e.close()
    e = be.ensemble(precision)
    series = []
    try:
        mid = e.add(p, g["P1"], g["S2"], perm, poro)
        e.apply_tiling(tiling)
        done = 0
        while done < steps:
            k = min(every or steps, steps - done)
            e.step(k)
            done += k
            if every:
                series.append(e.mass(mid))
        P, S = e.get(mid)
        m = e.mass(mid)
    finally:
        e.close()
    cells = p.NX * p.Ny * p.Nz
    info = {"value": cells / (ms * 1e-3), "ms_per_step": ms, "tiling": tiling,
            "bytes_per_cell_update": FUSED_BYTES[precision],
            "hbm_GBps": FUSED_BYTES[precision] * cells / (ms * 1e-3) / 1e9,
            "hbm_frac": FUSED_BYTES[precision] * cells / (ms * 1e-3) / HBM_PEAK,
            "tiled_mix_ceiling": tiled_mix_ceiling("k_fused", (p.NX, p.Ny, p.Nz),
                                                   tiling["tile"][:2], tiling["kz"], ms)}
    return info, (P, S, m, series)


def run_one_form(args):
    """One fused form on the headline grid: args.warmup eager steps, then
    args.steps steps as one graph, timed; the tiling from --tiling or the
    autotuner."""
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    prec = {"fp64": 0, "fast": 1, "mixed": 2}[args.form]
    p, perm, poro = PKG.workloads.lognormal_workload(args.n)
    g = PKG.workloads.initial_grid(p, perm, poro)
    e = be.ensemble(prec)
    try:
        e.add(p, g["P1"], g["S2"], perm, poro)
        if args.tiling:
            t, kz, order = args.tiling.split(",")
            e.set_tiling(int(kz), int(order), tuple(int(v) for v in t.split("x")))
        else:
            e.autotune(7)
        tiling = e.tiling()
        e.step(args.warmup)
        ms, _ = time_ensemble(e, args.steps, hiprt, be.lib)
    finally:
        e.close()
    cells = p.NX * p.Ny * p.Nz
    return {"form": args.form, "value": cells / (ms * 1e-3), "unit": "cell-updates/s",
            "ms_per_step": ms, "tiling": tiling, "grid": [p.NX, p.Ny, p.Nz],
            "hbm_frac": FUSED_BYTES[prec] * cells / (ms * 1e-3) / HBM_PEAK}


def oracle_run(p, perm, poro, steps, every=None, with_u=1):
    """The oracle port (loop-interchanged, OpenMP) on all host threads this
    process may use: (state, seconds, masses every `every` steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    orc = oracle.Oracle()
    orc.set_threads(host_threads())
    orc.set_loop_order(0)
    g = PKG.workloads.initial_grid(p, perm, poro)
    series, done, secs = [], 0, 0.0
    while done < steps:
        k = min(every or steps, steps - done)
        t0 = time.perf_counter()
        orc.run(g, p, k, step0=done, with_u=with_u)
        secs += time.perf_counter() - t0
        done += k
        if every:
            series.append(orc.mass(g, p))
            log(f"oracle {p.NX}x{p.Ny}x{p.Nz}: step {done}/{steps}")
    return g, secs, orc.mass(g, p), series


def run_slab(args, nx, ny, nz, with_oracle):
    """One large grid, un-batched: one rank's share (sigma 1.5, correlation
    16 cells, ln K clipped at +-2 sigma), args.slab_steps steps through the
    plain path and the fused FP64, fast and mixed forms, so that the fusion
    and precision gains show apart from batching.  with_oracle: every form
    against the oracle, final per-field errors and the cumulative
    mass-balance error every 10 steps."""
    import types
    W = PKG.workloads
    t0 = time.perf_counter()
    p, perm, poro = W.slab_workload(nx, ny, nz)
    gen_s = time.perf_counter() - t0
    steps = args.slab_steps
    sub = types.SimpleNamespace(steps=steps, warmup=3, autotune=False, step_tile=None)
    pr = run_plain(sub, p, perm, poro)
    cells = nx * ny * nz
    plain = cells * steps / pr["wall"]
    out = {"grid": [nx, ny, nz], "steps": steps, "field_generation_s": gen_s,
           "stability_r": pr["stability_r"], "state_finite": pr["finite"],
           "plain": {"value": plain, "ms_per_step": pr["wall"] * 1e3 / steps,
                     "stages_ms": pr["stage_ms"], "stencil_tiles": pr["tiles"]}}
    every = 10 if with_oracle else None
    states = {"plain": plain_state(p, perm, poro, steps, every)}
    for prec, name in FORMS.items():
        info, st = run_fused(sub, p, perm, poro, prec, steps, every)
        out[name] = info
        out[f"speedup_{name}_vs_plain"] = info["value"] / plain
        states[name] = st
    if with_oracle:
        log(f"slab {nx}x{ny}x{nz}: oracle, {steps} steps on {host_threads()} threads")
        ref, secs, mref, rseries = oracle_run(p, perm, poro, steps, every, with_u=0)
        out["oracle"] = {"seconds": secs, "threads": host_threads(),
                         "cell_updates_per_s": cells * steps / secs}
        for name, (P, S, m, series) in states.items():
            tgt = out[name] if name in out else out["plain"]
            tgt["error_vs_oracle"] = form_errors(P, S, m, ref["P1"], ref["S2"], mref,
                                                 f"oracle after {steps} steps")
            tgt["mass_balance_error_every_10_steps"] = mass_series_errors(series, rseries, 10)
    else:
        P64, S64, m64, _ = states["fused_fp64"]
        for name in ("fast", "mixed"):
            P, S, m, _ = states[name]
            out[name]["error_vs_fp64"] = mixed_errors(P64, S64, m64, P, S, m)
    return out


# ---------------------------------------------------------- ensembles --

def run_ensemble(args):
    """Batched ensemble: args.ensemble realizations with uneven grids, args.ensemble_steps
    steps.  Plain FP64 path run realization by realization (all stages, all
    arrays, default tiles) against the batched forms (every member in one
    launch per step, the whole run one hipGraph, tiling from the autotuner)."""
    import numpy as np
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    lib = be.lib
    stream = lib.filt_get_stream()
    W = PKG.workloads
    t0 = time.perf_counter()
    members = W.ensemble_members(args.ensemble)
    gen_s = time.perf_counter() - t0
    cells = sum(m["params"].NX * m["params"].Ny * m["params"].Nz for m in members)
    steps = args.ensemble_steps
    base_dt = W.base_params(4, 4, 4).dt
    lowered = [m for m in members if m["dt"] < base_dt]
    out = {"realizations": len(members), "cells": cells, "steps": steps,
           "field_generation_s": gen_s,
           "sigma_range": [min(m["sigma"] for m in members), max(m["sigma"] for m in members)],
           "lnK_clip_sigma": sorted({m["clip"] for m in members}),
           "max_stability_r": max(m["stability_r"] for m in members),
           "dt_lowered_for_stability": {"members": len(lowered),
                                        "min_dt": min((m["dt"] for m in members)),
                                        "sigmas": [round(m["sigma"], 3) for m in lowered]},
           "members": [{"shape": list(m["shape"]), "sigma": round(m["sigma"], 4),
                        "clip": m["clip"], "dt": m["dt"], "r": round(m["stability_r"], 4)}
                       for m in members]}
    log(f"ensemble: {len(members)} members generated ({gen_s:.1f} s); plain per realization")
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
    for prec, form in FORMS.items():
        name = "batched_" + form.replace("fused_", "")
        log(f"ensemble: {name}")
        e = be.ensemble(prec)
        for m in members:
            g = W.initial_grid(m["params"], m["perm"], m["poro"])
            e.add(m["params"], g["P1"], g["S2"], m["perm"], m["poro"])
        tiling = e.autotune(5)
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
        out[f"speedup_{name}_vs_plain"] = out[name]["value"] / out["plain_per_realization"]["value"]
    keys = ("S2_linf", "P1_rel_linf", "water_mass_rel", "napl_mass_rel", "total_mass_rel")
    for prec in (1, 2):
        errs = {k: 0.0 for k in keys}
        for ((P64, S64), m64), ((Pm, Sm), mm) in zip(results[0], results[prec]):
            e = mixed_errors(P64, S64, m64, Pm, Sm, mm)
            for k in keys:
                errs[k] = max(errs[k], e[k])
        out["batched_" + FORMS[prec]]["error_vs_fp64_max_over_members"] = errs
    # the three smallest members against the oracle (every form)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        orc = oracle.Oracle()
        orc.set_threads(host_threads())
        small = sorted(range(len(members)), key=lambda q: out["members"][q]["shape"][0] *
                       out["members"][q]["shape"][1] * out["members"][q]["shape"][2])[:3]
        vs = {}
        for q in small:
            m = members[q]
            ref = W.initial_grid(m["params"], m["perm"], m["poro"])
            orc.run(ref, m["params"], steps, with_u=0)
            mr = orc.mass(ref, m["params"])
            vs[str(q)] = {FORMS[prec]: form_errors(results[prec][q][0][0], results[prec][q][0][1],
                                                   results[prec][q][1], ref["P1"], ref["S2"], mr,
                                                   f"oracle after {steps} steps")
                          for prec in FORMS}
        out["smallest_members_vs_oracle"] = vs
    return out


def auxiliary(name, fn, *a):
    """A leg beside the headline whose failure (a file system the box does
    not let us write, say) is reported in its entry instead of ending the
    run."""
    try:
        return fn(*a)
    except Exception as err:  # noqa: BLE001 -- recorded, the headline goes on
        log(f"{name}: {type(err).__name__}: {err}")
        return {"error": f"{type(err).__name__}: {err}"}


def run_divgrad(args, p, perm, poro):
    """The reference's tau-relaxed divgrad scheme (assign_rS, cpu.cpp:162-216)
    in place of the upwind roS stage, on the headline grid: the five stage
    calls per step (the stage path; the one-kernel fusion is upwind only),
    timed like the stage-call leg, with the divgrad kernel's roofline, the u
    and divgrad kernels' tiles chosen by the library's stencil autotuner;
    then 3 steps from the initial state against the oracle port, cell for
    cell.  The tuned tiles are dropped afterwards, so later legs run the
    built-in ones."""
    import copy
    import numpy as np
    q = copy.copy(p)
    q.roS_scheme = PKG.abi.FILT_SCHEME_DIVGRAD
    K = max(1, args.steps // 4)
    sub = copy.copy(args)
    sub.autotune = not args.stencil_tile
    r = run_plain(sub, q, perm, poro, split=True, steps=K)
    cells = q.NX * q.Ny * q.Nz
    ms = r["stage_ms"]["roS"]
    out = {"value": cells * K / r["wall"], "unit": "cell-updates/s",
           "ms_per_step": r["wall"] * 1e3 / K, "steps": K, "stages_ms": r["stage_ms"],
           "state_finite": r["finite"],
           "bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in SPLIT_STAGES) -
           PLAIN_BYTES["roS"] + DIVGRAD_BYTES,
           "k_roS_divgrad_zm": {"bytes_per_cell": DIVGRAD_BYTES, "ms": ms,
                                "GBps": DIVGRAD_BYTES * cells / (ms * 1e-3) / 1e9,
                                "frac": DIVGRAD_BYTES * cells / (ms * 1e-3) / HBM_PEAK,
                                "tiles": r["tiles"]["roS"]}}
    flat = stream_mix_ceiling("k_roS_divgrad_zm", cells)
    if flat:
        flat["kernel_frac_of_mix"] = flat["ms"] / ms
    out["k_roS_divgrad_zm"]["stream_mix_ceiling"] = flat
    if not args.no_cpu:
        n = 3
        be = PKG.device.Backend()
        d = be.grid(q, perm=True, poro=True)
        try:
            d.upload(PKG.workloads.initial_grid(q, perm, poro))
            d.run(n)
            d.sync()
            P1, S2 = d.get("P1"), d.get("S2")
        finally:
            d.free()
        ref, secs, _, _ = oracle_run(q, perm, poro, n)
        out["oracle_check"] = {
            "steps": n, "oracle_seconds": secs,
            "P1_bit_equal": bool(np.array_equal(P1.view(np.uint64), ref["P1"].view(np.uint64))),
            "S2_bit_equal": bool(np.array_equal(S2.view(np.uint64), ref["S2"].view(np.uint64))),
            "P1_rel_linf": float(np.max(np.abs(P1 - ref["P1"])) / np.max(np.abs(ref["P1"]))),
            "S2_linf": float(np.max(np.abs(S2 - ref["S2"])))}
    lib = PKG.abi.load_library()
    for kern in (0, 1):  # back to the built-in tiles for the legs after this one
        PKG.abi.check(lib, lib.filt_set_tile_config(kern, 0, 0, 0, 0), "tile reset")
    return out


def run_io(args, n=128, steps=20):
    """Checkpoint and plot output (save/save.dat and the Tecplot files of
    main.cpp:319-569) from a grid resident on the device, n^3 cells: the
    synchronous writes, and a snapshot written by its background thread
    while the time steps go on.  Files go to a temporary folder that is
    removed afterwards."""
    import shutil
    import tempfile
    hiprt = importlib.import_module(PKG.__name__ + ".hiprt")
    be = PKG.device.Backend()
    p, perm, poro = PKG.workloads.lognormal_workload(n)
    d = be.grid(p, perm=True, poro=True)
    tmp = tempfile.mkdtemp(prefix="filt_io_")
    out = {"grid": [p.NX, p.Ny, p.Nz]}
    try:
        d.upload(PKG.workloads.initial_grid(p, perm, poro))
        d.run(5)
        d.sync()

        def files_bytes(path):
            if os.path.isfile(path):
                return os.path.getsize(path)
            return sum(os.path.getsize(os.path.join(a, f)) for a, _, fs in os.walk(path)
                       for f in fs)

        save = os.path.join(tmp, "save.dat")
        t0 = time.perf_counter()
        d.save(save, 5)
        ts = time.perf_counter() - t0
        b = files_bytes(save)
        out["save"] = {"seconds": ts, "bytes": b, "MBps": b / ts / 1e6}
        plots = os.path.join(tmp, "plots")
        t0 = time.perf_counter()
        d.save_plots(plots, 5 * p.dt)
        tp = time.perf_counter() - t0
        b = files_bytes(plots)
        out["plots"] = {"seconds": tp, "bytes": b, "MBps": b / tp / 1e6}
        # time steps alone, then with a snapshot written in the background
        stream = be.lib.filt_get_stream()
        a, c = hiprt.Event(), hiprt.Event()
        a.record(stream)
        d.run(steps, step0=5)
        c.record(stream)
        c.sync()
        alone = a.elapsed_ms(c) / steps
        snap = PKG.device.Snapshot(d)
        try:
            t0 = time.perf_counter()
            snap.take(d, 25, 25 * p.dt)
            d.sync()
            take_ms = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            snap.write_async(os.path.join(tmp, "save_async.dat"), os.path.join(tmp, "plots_async"))
            a.record(stream)
            d.run(steps, step0=25)
            c.record(stream)
            c.sync()
            during = a.elapsed_ms(c) / steps
            snap.wait()
            tw = time.perf_counter() - t0
        finally:
            snap.close()
        out["snapshot"] = {"take_ms": take_ms, "write_seconds": tw,
                           "step_ms_alone": alone, "step_ms_during_write": during}
    finally:
        d.free()
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def run_serving(args):
    """Serving configuration: Poisson arrivals (rate args.serve_rate per s for
    args.serve_duration s), edges log-uniform in [32, 192], steps U[100, 2000],
    continuous batching through filt_serve_*.  FP64 (the fused form, whose
    results equal the plain path's bit for bit) against the mixed form on the
    same request stream; the fast form too with --serve-fast."""
    serving = importlib.import_module(PKG.__name__ + ".serving")
    be = PKG.device.Backend()
    stream = serving.request_stream(rate=args.serve_rate, duration=args.serve_duration)
    out = {"rate_per_s": args.serve_rate, "window_s": args.serve_duration,
           "requests": len(stream),
           "offered_cell_updates_per_s": sum(s[1][0] * s[1][1] * s[1][2] * s[2] for s in stream)
           / args.serve_duration}
    legs = [(0, "fp64"), (2, "mixed")] + ([(1, "fast")] if args.serve_fast else [])
    for prec, name in legs:
        log(f"serving: {name}, {len(stream)} requests over {args.serve_duration:.0f} s")
        stats, _ = serving.run(be, prec, stream)
        out[name] = stats
    return out


# ----------------------------------------------------------------- CPU --

def cpu_baseline(args, port_full=None):
    """The reference's own CPU path (oracle/_ref, compiled from its sources;
    i-outer / k-inner loops, one thread, as its 1-rank CPU build runs):
      value: two time steps of the headline 256^3 grid (its rate per cell
             does not depend on the step count), about 20 s;
      cube32_full: the homogeneous 32^3 cube in full (500 steps).
    The compiled reference cannot express per-cell fields, a 4 x 4 patch or
    Sn = 0.5: its samples use the configurations' grid, spacing, time step,
    fluids and Brooks-Corey values through its media table, and its own
    5 x 5 patch with Sn = 0.4 (the cost per cell is the same).  Beside it the
    oracle port (loop-interchanged, OpenMP) on all threads this process may
    use: on the 32^3 cube in full (bitwise equal to the reference there)
    and on the headline workload in full (port_full, from the oracle check)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    W = PKG.workloads
    info = host_info()

    def ref_expressible(nx, ny, nz, h):
        p = PKG.abi.default_params_py()
        p.h1, p.h2, p.h3 = h
        p.dt, p.beta1, p.beta2 = 1e-4, W.STABLE_BETA[0], W.STABLE_BETA[1]
        p.NX, p.Ny, p.Nz, p.source = nx, ny, nz, 2
        for q in range(2):
            p.K[q], p.lambda_[q], p.S_wr[q], p.m[q], p.P_d[q] = W.K_MEAN, 2.0, 0.1, 0.4, 0.0
        return PKG.abi.finalize_patch_py(p)

    def timed(impl, p, steps):
        g = PKG.grid.HostGrid(p)
        impl.initial_data(g, p)
        t0 = time.perf_counter()
        impl.run(g, p, steps)
        return time.perf_counter() - t0, g

    have_ref = oracle.ref_available()
    if have_ref:
        impl, kind = oracle.RefCPU(), "reference"
    else:
        impl, kind = oracle.Oracle(), "port"
        impl.set_loop_order(1)
    n = args.n
    p2 = ref_expressible(n, n, n, W.INI_H)
    log(f"cpu_baseline: {kind} on 2 steps of {n}^3, 1 thread")
    dt2, _ = timed(impl, p2, 2)
    p1 = ref_expressible(32, 32, 32, (1 / 31, 1 / 31, 1 / 31))
    log(f"cpu_baseline: {kind} on the homogeneous 32^3 cube (500 steps), 1 thread")
    dt1, g_ref1 = timed(impl, p1, 500)
    port = oracle.Oracle()
    port.set_threads(host_threads())
    port.set_loop_order(0)
    dtp1, g_port1 = timed(port, p1, 500)
    out = {"value": n ** 3 * 2 / dt2, "unit": "cell-updates/s", "cores": 1, "kind": kind,
           "sample": f"2 steps of the {n}^3 grid (media table with the configuration's "
                     f"fluids, reference 5x5 patch), {dt2:.1f} s on 1 thread of "
                     f"{info['cpu_model']}; the reference's rate per cell does not depend on "
                     f"the step count",
           "host": info,
           "cube32_full": {"value": 32 ** 3 * 500 / dt1, "cores": 1, "kind": kind,
                       "sample": f"the homogeneous 32^3 cube in full: 500 steps, {dt1:.1f} s "
                                 "(reference 5x5 patch, Sn 0.4)"},
           "best_effort_port": {
               "cube32_full": {"value": 32 ** 3 * 500 / dtp1, "cores": host_threads(),
                           "seconds": dtp1,
                           "bitwise_equal_to_reference": bool(
                               np.array_equal(g_port1["P1"], g_ref1["P1"]) and
                               np.array_equal(g_port1["S2"], g_ref1["S2"]))},
               "lognormal_full": port_full}}
    return out


def stream_mix_ceiling(kernel, cells):
    """Time of a pure stream with the kernel's mix of FP64 arrays
    (tests/hip/bin/stream_mix, built from tests/hip/stream_mix.hip): the
    fastest any kernel reading and writing those arrays once can run on this
    GPU.  Best of plain and non-temporal stores; None when the tool is
    missing or has no such mix."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "hip", "bin", "stream_mix")
    if not os.path.exists(exe):
        return None
    try:
        out = subprocess.run([exe, str(cells), kernel], capture_output=True, text=True,
                             timeout=120, check=True).stdout
    except (subprocess.SubprocessError, OSError) as err:  # context only: never fatal
        log(f"stream_mix {kernel}: {err}")
        return None
    rows = [json.loads(x) for x in out.splitlines() if x.startswith("{")]
    if not rows:
        return None
    best = min(rows, key=lambda r: r["ms"])
    return {"mix": f"{best['in']} in / {best['out']} out FP64 arrays",
            "bytes_per_cell": best["bytes_per_cell"], "ms": best["ms"],
            "GBps": best["TBps"] * 1e3, "nontemporal_stores": best["nontemporal_stores"]}


def tiled_mix_ceiling(kernel, grid, tile, kz, ms=None):
    """Time of a pure stream with the kernel's mix of FP64 arrays and the
    z-march's access pattern: tiles of tile[0] x tile[1] columns from x = 0,
    kz planes per block, no halo, no arithmetic (stream_mix's tiled mode).
    With the kernel's own time ms, also the kernel's share of it."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "hip", "bin", "stream_mix")
    if not os.path.exists(exe):
        return None
    args = [exe, "tiled", *[str(v) for v in grid], str(tile[0]), str(tile[1]), str(kz), kernel]
    try:
        out = subprocess.run(args, capture_output=True, text=True, timeout=120,
                             check=True).stdout
    except (subprocess.SubprocessError, OSError) as err:  # context only: never fatal
        log(f"stream_mix tiled {kernel}: {err}")
        return None
    rows = [json.loads(x) for x in out.splitlines() if x.startswith("{")]
    if not rows:
        return None
    r = rows[0]
    res = {"mix": f"{r['in']} in / {r['out']} out FP64 arrays", "tile": list(tile), "kz": kz,
           "bytes_per_cell": r["bytes_per_cell"], "ms": r["ms"], "GBps": r["TBps"] * 1e3}
    if ms:
        res["kernel_frac_of_mix"] = r["ms"] / ms
    return res


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
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU legs: the baseline and the oracle checks")
    ap.add_argument("--ensemble", type=int, default=96,
                    help="realizations of the batched-ensemble leg (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--slab-steps", type=int, default=100,
                    help="steps of the 512x512x256 and 256x512x512 slab legs (0: skip)")
    ap.add_argument("--serve-duration", type=float, default=60.0,
                    help="arrival window of the serving leg in seconds (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-fast", action="store_true",
                    help="also serve the request stream with the fast form")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the headline plain path (the profiled workload)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the stencil tile shapes of the plain path on this grid "
                         "(default: the library's built-in shapes)")
    ap.add_argument("--copy", action="store_true",
                    help="only run the copy-kernel calibration (1 GiB) and print it")
    ap.add_argument("--form", choices=("fp64", "fast", "mixed"),
                    help="only run this fused form on the headline grid and print its line "
                         "(for profiles)")
    ap.add_argument("--leg", choices=("ensemble", "serving", "divgrad", "io"),
                    help="only run this leg (the batched ensemble, the serving loop, the "
                         "divgrad scheme or the checkpoint and plot output) and print its "
                         "entry (for profiles)")
    ap.add_argument("--tiling", default=None,
                    help="with --form: fixed tiling TXxTY[xCY],kz,order (e.g. 64x8x2,64,1) "
                         "instead of the autotuner")
    ap.add_argument("--stencil-tile", default=None,
                    help="fixed tile of the u and roS stencil kernels for the divgrad leg, "
                         "TXxTY,kz,order, instead of their autotuner: for profiler runs")
    ap.add_argument("--step-tile", default=None,
                    help="fixed tile of the plain path's stages 2-4 kernel, TXxTY,kz,order "
                         "(e.g. 128x8,64,0), instead of its autotuner: for profiler runs "
                         "that must launch the same kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        raise SystemExit(2)
    if args.copy:
        print(json.dumps(copy_calibration()), flush=True)
        return
    if args.form:
        print(json.dumps(run_one_form(args)), flush=True)
        return
    if args.leg:
        if args.leg == "divgrad":
            p, perm, poro = PKG.workloads.lognormal_workload(args.n)
            res = run_divgrad(args, p, perm, poro)
        else:
            res = {"ensemble": run_ensemble, "serving": run_serving, "io": run_io}[args.leg](args)
        print(json.dumps(res), flush=True)
        return

    p, perm, poro = PKG.workloads.lognormal_workload(args.n)
    log(f"plain {args.n}^3: {args.steps} timed steps")
    pr = run_plain(args, p, perm, poro)
    wall, stage_ms = pr["wall"], pr["stage_ms"]
    cells = p.NX * p.Ny * p.Nz
    value = cells * args.steps / wall
    forms, states, split, shims, oracle_check, port_full = {}, {}, None, None, None, None
    one_kernel = None
    if not args.plain_only:
        states["plain"] = plain_state(p, perm, poro, args.steps)
        for prec, name in FORMS.items():
            log(f"{name} {args.n}^3")
            info, st = run_fused(args, p, perm, poro, prec)
            forms[name] = info
            states[name] = st
        P64, S64, m64, _ = states["fused_fp64"]
        for name in ("fast", "mixed"):
            P, S, m, _ = states[name]
            forms[name]["error_vs_fp64"] = mixed_errors(P64, S64, m64, P, S, m)
        # the five stage functions one by one (the reference driver's calls)
        log("stage calls and defines.h shims")
        k5 = max(1, args.steps // 2)
        r5 = run_plain(args, p, perm, poro, split=True, steps=k5)
        split = {"value": cells * k5 / r5["wall"], "ms_per_step": r5["wall"] * 1e3 / k5,
                 "stages_ms": r5["stage_ms"], "state_finite": r5["finite"],
                 "bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in SPLIT_STAGES),
                 "hbm_frac": sum(PLAIN_BYTES[s] for s in SPLIT_STAGES) * cells * k5 /
                 r5["wall"] / HBM_PEAK}
        shims = {"fused": run_shim(args, p, perm, poro, k5, True),
                 "separate": run_shim(args, p, perm, poro, k5, False)}
        # the opt-in whole-step kernel (stage 1 with stages 2-4, stash copy)
        lib = PKG.abi.load_library()
        PKG.abi.check(lib, lib.filt_set_tile_config(3, 64, 16, 64, 0), "whole-step kernel on")
        try:
            r1 = run_plain(args, p, perm, poro, steps=k5)
        finally:
            lib.filt_set_tile_config(3, 0, 0, 0, 0)
        full_ms = r1["dominant"]["ms"] if r1["dominant"]["stage"] == "step_full" else None
        one_kernel = {"value": cells * k5 / r1["wall"], "ms_per_step": r1["wall"] * 1e3 / k5,
                      "stages_ms": r1["stage_ms"], "tile": [64, 16], "kz": 64, "order": 0,
                      "bytes_per_cell_update": PLAIN_BYTES["step_full"],
                      "k_step_full_frac": (PLAIN_BYTES["step_full"] * cells /
                                           (full_ms * 1e-3) / HBM_PEAK) if full_ms else None,
                      "stream_mix_ceiling": stream_mix_ceiling("k_step_full", cells)}
        if not args.no_cpu:
            log(f"oracle port on {args.n}^3, {args.steps} steps, {host_threads()} threads")
            ref, secs, mref, _ = oracle_run(p, perm, poro, args.steps)
            port_full = {"value": cells * args.steps / secs, "cores": host_threads(),
                       "seconds": secs, "sample": f"the headline workload in full: {args.n}^3, "
                                                  f"{args.steps} steps, per-cell K and porosity"}
            oracle_check = {name: form_errors(P, S, m, ref["P1"], ref["S2"], mref,
                                           f"oracle port after {args.steps} steps")
                         for name, (P, S, m, _) in states.items()}
        log("divgrad scheme")
        forms["divgrad"] = auxiliary("divgrad", run_divgrad, args, p, perm, poro)
        log("checkpoint and plot output")
        forms["io"] = auxiliary("io", run_io, args)
        if args.slab_steps > 0:
            for shape, with_oracle in (((512, 512, 256), not args.no_cpu),
                                       ((256, 512, 512), False)):
                log("slab %dx%dx%d" % shape)
                forms["slab_%dx%dx%d" % shape] = run_slab(args, *shape, with_oracle)
        if args.ensemble > 0:
            forms["batched_ensemble"] = run_ensemble(args)
        if args.serve_duration > 0:
            forms["serving"] = run_serving(args)
    step_names = pr["names"]
    dom, dom_ms = pr["dominant"]["stage"], pr["dominant"]["ms"]
    dom_bytes = PLAIN_BYTES[dom] * cells
    achieved = dom_bytes / (dom_ms * 1e-3)
    tag = f"plain_{args.n}cube"
    traffic = load_traffic(tag, KERNEL_NAMES[dom])
    mix_ceiling = stream_mix_ceiling(KERNEL_NAMES[dom], cells)
    if mix_ceiling:
        mix_ceiling["kernel_frac_of_mix"] = mix_ceiling["ms"] / dom_ms
    step_bytes = sum(PLAIN_BYTES[s] for s in step_names) * cells
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
            "workload": f"plain FP64, {args.n}^3, time_step_function per step (stage 1 and "
                        "stages 2-4 in one kernel, the boundary; every ptr_Arrays array "
                        "materialised), per-cell log-normal K (sigma 1, corr 8 cells) and "
                        "porosity",
            "grid": [p.NX, p.Ny, p.Nz],
            "form": "plain",
            "deviations": "beta1=1e-4, beta2=1e-8 (stated 4.4e-10 is unstable); "
                          "defines.ini spacing; ln K clipped at +-2.5 sigma",
            "stability_r": pr["stability_r"],
            "stencil_tiles": pr["tiles"],
            "state_finite": pr["finite"],
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
            "kernel_ms": dom_ms,
            "stream_mix_ceiling": mix_ceiling,
        },
        "stages_ms": stage_ms,
        "stages_ms_from": f"{pr['breakdown_steps']} steps after the timed region, with a HIP "
                          "event at every stage boundary",
        "step_roofline": {
            "bytes_per_cell_update": sum(PLAIN_BYTES[s] for s in step_names),
            "achieved_GBps": step_bytes * args.steps / wall / 1e9,
            "frac": step_bytes * args.steps / wall / HBM_PEAK,
        },
        "stage_calls": split,
        "plain_one_kernel": one_kernel,
        "defines_shims": shims,
        "oracle_check": oracle_check,
        "forms": forms,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, port_full)
    log("done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
