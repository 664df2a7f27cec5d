#!/usr/bin/env python3
"""Throughput of the CGFD3D time-stepping hot path on one MI355X.

Workload: the large single-GPU heterogeneous model -- 600x600x300 cells,
dh = 50 m, random von Karman surface (a = 5 km, RMS 300 m), 1-D gradient
Vp 3500 -> 7000 m/s with a 5 % von Karman perturbation, Gardner density,
20-cell ADE CFS-PML on five faces, traction-image free surface, a finite
fault of 200 double-couple subfaults, 400 surface receivers and PGV.  Inputs
are synthetic (seeded) and resident in HBM before timing starts.

A step is one full RK4 time step (4 stages) of the whole grid.  The metric
is Gcell-RKstages/s = ni*nj*nk*4*steps / t.  The roofline entry divides the
algorithmic bytes of one stage -- the compulsory traffic of the 3-register
RK4 with 9 wavefield, 9 metric and 3 media values: 192 B per cell + 144 B per
PML cell-axis, FP32 -- by the mean duration of the fused stage kernel
measured with HIP events on the engine's stream over the timed steps.  When
the horizontal metrics are the constants the mesh implies, the kernel loads
only the 3 zeta metrics (168 B per cell); `config.bytes_per_stage_loaded`
reports that figure.  `traffic` is the measured HBM bytes per launch from the
committed counter summary.  The CPU baseline is the oracle
(a C/OpenMP restatement of the reference kernels) timed on this host on three
steps (about 15 s of CPU work) of the same problem; the GPU then runs the
same three steps from the same initial state and `config.oracle_parity`
says whether every wavefield value, trace sample and PGV value equals the
oracle's.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "copara-cgfd3d_amd"))

HBM_PEAK_GBS = 8000.0
PROFILE_STEPS = 8


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name, nt):
    import scenario

    if name == "vonkarman":
        return scenario.config_vonkarman(nt=nt)
    if name == "hill":
        return scenario.config_hill(nt=nt)
    if name == "flat":
        return scenario.config_flat(nt=nt)
    raise SystemExit(f"unknown workload {name}")


def describe(sc):
    return {
        "workload": sc["name"],
        "cells": [sc["ni"], sc["nj"], sc["nk"]],
        "dh_m": sc["dh"],
        "dt_s": sc["dt"],
        "pml": list(map(int, sc["nabs"])),
        "sources": int(len(sc.get("moment_locs", [])) + len(sc.get("force_locs", []) if sc.get("force_locs") is not None else [])),
        "receivers": int(len(sc["recv_locs"])),
        "free_surface": "traction image",
        "storage": "f32",
        "stencil_accumulation": "f64",
    }


def batch_bench(args):
    """Batched FD: `--batch` independent hill-medium scenarios of uneven size
    (96-192 x 96-192 x 64-128 cells), advanced together with cgfd_run_batch
    (one launch per kernel and stage for all of them), against the same
    scenarios run one after another with per-grid autotuned z-chunks.
    Prints one JSON line."""
    import cgfd3d
    import scenario

    if args.full:
        return batch_full_bench(args)
    nt = args.warmup + args.steps
    t0 = time.perf_counter()
    scs = scenario.config_scenario_batch(args.batch, nt_range=(nt, nt), seed=11)
    log(f"{len(scs)} scenarios built in {time.perf_counter() - t0:.1f} s")
    sims = []
    for sc in scs:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.setup(sc)
        g.set_graph(args.graph)
        sims.append(g)
    zch = [g.autotune(trial_steps=2) for g in sims]
    cells = sum(sc["ni"] * sc["nj"] * sc["nk"] for sc in scs)

    def timed(run):
        for g in sims:
            g.init()
        run(args.warmup)
        for g in sims:
            g.sync()
        t = time.perf_counter()
        run(args.steps)
        for g in sims:
            g.sync()
        return time.perf_counter() - t

    def one_by_one(n):
        for g in sims:  # each scenario finishes before the next starts
            g.run(n)
            g.sync()

    t_seq = timed(one_by_one)
    t_bat = timed(lambda n: cgfd3d.run_batch(sims, n))
    fused = cgfd3d.batch_fusable(sims)
    for g in sims:
        g.close()
    value = cells * 4 * args.steps / t_bat / 1e9
    line = {
        "metric": "Gcell-RKstages/s", "value": round(value, 4), "unit": "Gcell-RKstages/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_bat / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch-{len(scs)}-hill", "graph_replay": bool(args.graph), "fused": fused,
                   "scenarios": [sc["name"] for sc in scs],
                   "zchunks": zch, "cells_total": cells,
                   "scenario_steps_per_s_batched": round(len(scs) * args.steps / t_bat, 2),
                   "scenario_steps_per_s_sequential": round(len(scs) * args.steps / t_seq, 2),
                   "speedup_vs_sequential": round(t_seq / t_bat, 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def batch_full_bench(args):
    """The uneven-scenario batch run to completion: `--batch` hill scenarios
    with their own sizes (96-192 x 96-192 x 64-128 cells) and lengths
    (800-2500 steps), 32 stations and PGV each.  Batched: every call of
    cgfd_run_batch advances all unfinished scenarios (one launch per kernel
    and stage) up to the next scenario's end.  Sequential: each scenario
    alone, to its end.  The value is whole-job cell-stages per second of the
    batched run; scenarios/s of both runs are in config."""
    import numpy as np

    import cgfd3d
    import scenario

    t0 = time.perf_counter()
    scs = scenario.config_scenario_batch(args.batch, seed=11)
    log(f"{len(scs)} scenarios built in {time.perf_counter() - t0:.1f} s")
    sims = []
    for sc in scs:
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
        g.setup(sc)
        sims.append(g)
    nts = [sc["nt"] for sc in scs]
    work = sum(sc["ni"] * sc["nj"] * sc["nk"] * 4 * sc["nt"] for sc in scs)

    def reset_and_warm():
        for g in sims:
            g.init()
        cgfd3d.run_batch(sims, 2)   # first launches of every kernel instance
        for g in sims:
            g.init()
            g.sync()

    reset_and_warm()
    t = time.perf_counter()
    for g, nt in zip(sims, nts):
        g.run(nt)
        g.sync()
    t_seq = time.perf_counter() - t
    log(f"sequential: {t_seq:.1f} s")
    pgv_seq = [g.get_pgv() for g in sims]

    reset_and_warm()
    t = time.perf_counter()
    done, calls = 0, 0
    active = list(range(len(sims)))
    while active:
        n = min(nts[i] - done for i in active)
        cgfd3d.run_batch([sims[i] for i in active], n)
        calls += 1
        done += n
        active = [i for i in active if nts[i] > done]
    for g in sims:
        g.sync()
    t_bat = time.perf_counter() - t
    log(f"batched: {t_bat:.1f} s in {calls} calls")
    same = all(np.array_equal(g.get_pgv(), p) for g, p in zip(sims, pgv_seq))
    for g in sims:
        g.close()
    line = {
        "metric": "Gcell-RKstages/s", "value": round(work / t_bat / 1e9, 4), "unit": "Gcell-RKstages/s",
        "n_gpus": 1, "steps": int(max(nts)), "warmup": 2, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch-{len(scs)}-hill-full", "steps_per_scenario": [int(n) for n in nts],
                   "cells_total": int(sum(sc["ni"] * sc["nj"] * sc["nk"] for sc in scs)),
                   "batch_calls": calls,
                   "scenarios_per_s_batched": round(len(scs) / t_bat, 4),
                   "scenarios_per_s_sequential": round(len(scs) / t_seq, 4),
                   "speedup_vs_sequential": round(t_seq / t_bat, 3),
                   "seconds_batched": round(t_bat, 2), "seconds_sequential": round(t_seq, 2),
                   "pgv_equal_to_sequential": bool(same)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


SURROGATE_MEDIA = {
    # the 192x192x128 Gaussian-hill medium: 128x128 PGV window, 32 stations'
    # traces of 1024 samples, sources 2-10 km deep
    "192x192x128": dict(ni=192, nj=192, nk=128, nt=1024, dt=0.006, npml=10, window=128, stations=32,
                        depth=(20, 101)),
    # the reduced medium of the longer training studies
    "80x80x48": dict(ni=80, nj=80, nk=48, nt=600, dt=0.006, npml=10),
}


def surrogate_study(medium, ntrain, ncal, ntest, epochs, width, modes, batch, augment, nseq, log_fn=None):
    """FD data from the batched engine on a fixed Gaussian-hill medium, two
    FNOs (4 layers) trained on the project's kernels (with the seismogram
    head where the medium has stations), the gate calibrated for a 5 %
    median PGV error of the accepted scenarios, and the test scenarios
    answered by the gated surrogate.  Returns the measurements as a dict."""
    import numpy as np

    import surrogate

    base = surrogate.base_medium(**SURROGATE_MEDIA[medium])
    traces = len(base["stations"]) > 0
    ptr = surrogate.draw_sources(ntrain, base, seed=1)
    pcal = surrogate.draw_sources(ncal, base, seed=2)
    pte = surrogate.draw_sources(ntest, base, seed=3)
    nseq = min(nseq, len(pte))
    _, t_seq = surrogate.timed(lambda: surrogate.run_fd(base, pte[:nseq], group=1))
    log(f"sequential FD: {nseq} scenarios in {t_seq:.1f} s")
    out, t_fd = surrogate.timed(lambda: surrogate.run_fd(base, ptr, group=32, log=log_fn, traces=traces))
    ytr, ttr = out if traces else (out, None)
    log(f"batched FD training data: {len(ptr)} scenarios in {t_fd:.1f} s")
    ycal = surrogate.run_fd(base, pcal, group=32)
    out = surrogate.run_fd(base, pte, group=32, traces=traces)
    yte, tte = out if traces else (out, None)
    s = surrogate.Surrogate(base, width=width, modes=modes, layers=4)
    losses, t_train = surrogate.timed(lambda: s.fit(ptr, ytr, epochs=epochs, batch=batch, lr=3e-3, log=log_fn,
                                                    augment=augment, traces=ttr))
    log(f"training: {t_train:.1f} s")
    rank = lambda v: np.argsort(np.argsort(v))
    q = lambda v, p: round(float(np.quantile(v, p)), 4) if len(v) else None
    res = {}
    for tta in (False, True):  # plain, and averaged over the symmetric images of each query
        s.tta = tta
        thr = s.calibrate(pcal, ycal, target_err=0.05)
        out, t_sur = surrogate.timed(lambda: s.predict(pte, traces=traces))
        pred, ind = out[:2]
        if tta:  # the reported configuration: answered end to end, FD for the rejected
            (ans, ok), t_gate = surrogate.timed(lambda: s.answer(pte))
        else:
            ans, ok, t_gate = None, ind <= thr, None
        err = surrogate.pgv_error(pred, yte)
        terr = surrogate.trace_error(out[2], tte) if traces else None
        res[tta] = {"threshold": thr, "t_sur": t_sur, "t_gate": t_gate, "ok": ok, "err": err, "terr": terr,
                    "err_pw": surrogate.pgv_pointwise_error(pred, yte),
                    "err_gate": None if ans is None else surrogate.pgv_error(ans, yte),
                    "rho": float(np.corrcoef(rank(ind), rank(err))[0, 1]) if len(err) > 1 else None}
        log(f"tta={tta}: median PGV error {np.median(err):.4f}, accepted {ok.mean():.3f}")
    r = res[True]
    gated = len(pte) / r["t_gate"]
    seq_rate = nseq / t_seq
    acc = r["err"][r["ok"]]
    return {
        "workload": f"surrogate-hill-{medium}-nt{base['nt']}", "pgv_window": base["window"][2],
        "stations": len(base["stations"]), "train_scenarios": len(ptr), "calibration_scenarios": len(pcal),
        "test_scenarios": len(pte), "epochs": epochs,
        "fno": {"width": width, "modes": modes, "layers": 4, "ensemble": 2},
        "symmetry_augmentation": bool(augment), "train_batch": batch,
        "fd_sequential_scenarios_per_s": round(seq_rate, 3),
        "fd_batched_scenarios_per_s": round(len(ptr) / t_fd, 3),
        "surrogate_only_scenarios_per_s": round(len(pte) / r["t_sur"], 1),
        "gated_scenarios_per_s": round(gated, 3),
        "gated_speedup_vs_sequential_fd": round(gated / seq_rate, 2),
        "batched_fd_speedup_vs_sequential_fd": round(len(ptr) / t_fd / seq_rate, 2),
        "train_seconds": round(t_train, 1), "final_train_loss": round(losses[-1], 5),
        "test_time_symmetry_averaging": True,
        "gate_target": "median PGV error of the accepted scenarios <= 0.05 on the calibration set",
        "gate_threshold": r["threshold"],
        "accepted_fraction": round(float(r["ok"].mean()), 3),
        "median_pgv_error_accepted": q(acc, 0.5), "p90_pgv_error_accepted": q(acc, 0.9),
        "median_pgv_error_gated_all": round(float(np.median(r["err_gate"])), 4),
        "median_pgv_error_surrogate_all": round(float(np.median(r["err"])), 4),
        "p90_pgv_error_surrogate_all": q(r["err"], 0.9),
        "median_pointwise_pgv_error_surrogate_all": round(float(np.median(r["err_pw"])), 4),
        "median_trace_error_accepted": q(r["terr"][r["ok"]], 0.5) if traces else None,
        "median_trace_error_surrogate_all": q(r["terr"], 0.5) if traces else None,
        "indicator_error_rank_correlation": None if r["rho"] is None else round(r["rho"], 3),
        "without_averaging": {"median_pgv_error": round(float(np.median(res[False]["err"])), 4),
                              "accepted_fraction": round(float(res[False]["ok"].mean()), 3),
                              "surrogate_only_scenarios_per_s": round(len(pte) / res[False]["t_sur"], 1)},
        "pgv_error": "relative L2 of the three surface PGV maps per scenario vs full FD",
        "trace_error": "relative L2 of all station traces per scenario vs full FD",
    }


def surrogate_bench(args):
    """Learned surrogate with FD fallback (surrogate_study) at the options'
    sizes; prints one JSON line whose value is the gated scenarios/s, with
    sequential plain FD beside it."""
    r = surrogate_study(args.sur_medium, args.train, args.calib, args.test, args.epochs, args.fno_width,
                        args.fno_modes, args.sur_batch, args.augment, 8, log)
    line = {
        "metric": "scenarios/s", "value": r["gated_scenarios_per_s"], "unit": "scenarios/s", "n_gpus": 1,
        "steps": SURROGATE_MEDIA[args.sur_medium]["nt"], "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (surrogate, fp32 accumulation); f32 (FD)", "data": "synthetic", "config": r,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def fp64_error(sc, steps):
    """The FP32 run against the FP64 build of the same kernels on the same
    inputs: relative L2 of each of the 9 wavefield components after `steps`
    steps (largest and mean), and the FP64 build's Gcell-RKstages/s over the
    steps after the first two."""
    import numpy as np

    import cgfd3d

    fields = {}
    rate = None
    for prec in (cgfd3d.FP32, cgfd3d.FP64):
        g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"], precision=prec)
        if prec == cgfd3d.FP64:
            sc = dict(sc, **{k: sc[k].astype(np.float64) for k in ("alpha", "beta", "rho")})
        g.setup(sc)
        head = min(2, steps)
        g.run(head)
        g.sync()
        t = time.perf_counter()
        g.run(steps - head)
        g.sync()
        if prec == cgfd3d.FP64 and steps > head:
            cells = sc["ni"] * sc["nj"] * sc["nk"]
            rate = round(cells * 4 * (steps - head) / (time.perf_counter() - t) / 1e9, 3)
        fields[prec] = [g.get_field(c) for c in range(9)]
        g.close()
    errs = []
    for a, b in zip(fields[cgfd3d.FP32], fields[cgfd3d.FP64]):
        d = np.linalg.norm((a.astype(np.float64) - b).ravel())
        n = np.linalg.norm(b.ravel())
        errs.append(d / n if n > 0 else 0.0)
    return {"steps": steps, "rel_l2_max": float(f"{max(errs):.3e}"), "rel_l2_mean": float(f"{np.mean(errs):.3e}"),
            "fp64_gcell_rkstages_per_s": rate,
            "what": f"FP32 vs FP64 GPU build, the 9 wavefield components after {steps} steps of the same problem"}


def host_cpu():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(sc, steps):
    """Time the oracle on `steps` steps of the same problem, then run the same
    steps on the GPU from the same initial state and compare every wavefield
    component, the traces and the PGV maps value for value.  Returns the
    cpu_baseline record and the parity record."""
    import numpy as np

    import cgfd3d

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    import scenario

    # both start from the same random wavefield (the sources of the first
    # steps alone would leave most of the grid at zero)
    o = oracle.Oracle(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    o.setup(sc)
    g = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    g.setup(sc)
    for c, a in enumerate(scenario.random_state(g.shape, np.random.default_rng(600))):
        o.set_field(c, a)
        g.set_field(c, a)
        del a
    t0 = time.perf_counter()
    o.run(steps)
    t = time.perf_counter() - t0
    nv = sc["ni"] * sc["nj"] * sc["nk"]
    val = nv * 4 * steps / t / 1e9
    threads = oracle.num_threads()
    model, nproc = host_cpu()

    g.run(steps)
    g.sync()
    differ = []
    for c in range(9):
        a, b = g.get_field(c), o.get_field(c)
        if not np.array_equal(a, b):
            differ.append(f"{cgfd3d.COMPONENTS[c]}: {int(np.count_nonzero(a != b))} values")
        del a, b
    for name, a, b in (("traces", g.get_seismo(), o.get_seismo()), ("pgv", g.get_pgv(), o.get_pgv())):
        if not np.array_equal(a, b):
            differ.append(f"{name}: {int(np.count_nonzero(a != b))} values")
    g.close()
    o.close()
    parity = {"steps": steps, "identical": not differ, "differences": differ,
              "what": f"GPU FP32 engine vs the CPU oracle after {steps} step(s) of the full "
                      f"{sc['ni']}x{sc['nj']}x{sc['nk']} problem from the same random initial wavefield: the 9 "
                      f"wavefield components, the receiver traces and the PGV maps compared value for value"}
    cpu = {"value": round(val, 6), "unit": "Gcell-RKstages/s", "cores": threads, "kind": "port",
           "nproc": nproc, "cpu_model": model,
           "sample": f"{steps} step(s) = {4 * steps} RK stages of the full {sc['ni']}x{sc['nj']}x{sc['nk']} "
                     f"problem (random initial wavefield) on the C/OpenMP oracle, {threads} threads, {t:.1f} s",
           "port": "a C/OpenMP restatement of the reference kernels that fuses each stage into one pass over the "
                   "grid; the reference launches 4+ kernels per stage with 27 derivative arrays and state copies "
                   "in between, so the port moves fewer bytes and is, if anything, faster than a host build of "
                   "the reference would be"}
    return cpu, parity


def copy_calibration(mib, reps=3):
    """A plain device-to-device copy of `mib` MiB, `reps` times: under a
    counter profiler its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the
    counters against a known byte count (mib MiB read and written).  The copy
    goes through the HIP runtime the engine library already loaded (a second
    runtime in the same process, such as torch's bundled one, is avoided)."""
    import ctypes

    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    if not paths:
        raise SystemExit("copy calibration: the HIP runtime is not loaded")
    hip = ctypes.CDLL(sorted(paths)[0])
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()

    def ok(rc, what):
        if rc != 0:
            raise SystemExit(f"copy calibration: {what} returned {rc}")

    ok(hip.hipMalloc(ctypes.byref(a), n), "hipMalloc")
    ok(hip.hipMalloc(ctypes.byref(b), n), "hipMalloc")
    ok(hip.hipMemset(a, 1, n), "hipMemset")
    for _ in range(reps):
        ok(hip.hipMemcpy(b, a, n, 3), "hipMemcpy")  # 3 = hipMemcpyDeviceToDevice
    ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    ok(hip.hipFree(a), "hipFree")
    ok(hip.hipFree(b), "hipFree")
    log(f"copy calibration: {reps} x {mib} MiB device copy")


def algorithmic_bytes(sc):
    """Bytes of one RK stage by the per-unit model: 192 B per valid cell
    (mean over the 4 stages of 39/57/57/39 floats) + 144 B per cell and per
    active PML axis (mean of 27/45/45/27 floats), FP32."""
    ni, nj, nk = sc["ni"], sc["nj"], sc["nk"]
    nab = [int(v) for v in sc["nabs"]]
    npml = (nab[0] + nab[1]) * nj * nk + (nab[2] + nab[3]) * ni * nk + (nab[4] + nab[5]) * ni * nj
    return 192.0 * ni * nj * nk + 144.0 * npml


def measured_traffic(workload, variant):
    """HBM bytes per stage-kernel launch for this workload, from the committed
    counter summary (profiles/stage_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "stage_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    entry = table.get(f"{workload}/variant{variant}")
    return None if entry is None else entry.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="vonkarman", choices=["vonkarman", "hill", "flat", "batch", "surrogate"])
    ap.add_argument("--batch", type=int, default=16, help="scenarios of the batch workload")
    ap.add_argument("--full", action="store_true",
                    help="batch workload: run every scenario to its own end (800-2500 steps) instead of --steps")
    ap.add_argument("--train", type=int, default=512, help="training scenarios of the surrogate workload")
    ap.add_argument("--calib", type=int, default=32, help="gate-calibration scenarios of the surrogate workload")
    ap.add_argument("--test", type=int, default=64, help="test scenarios of the surrogate workload")
    ap.add_argument("--epochs", type=int, default=70, help="training epochs of the surrogate workload")
    ap.add_argument("--fno-width", type=int, default=64, help="FNO channel width of the surrogate workload")
    ap.add_argument("--fno-modes", type=int, default=16, help="retained Fourier modes per axis")
    ap.add_argument("--augment", action=argparse.BooleanOptionalAction, default=True,
                    help="also train on the symmetric images of every FD run")
    ap.add_argument("--sur-batch", type=int, default=32, help="training batch size of the surrogate workload")
    ap.add_argument("--sur-medium", default="192x192x128", choices=sorted(SURROGATE_MEDIA),
                    help="medium of the surrogate workload")
    ap.add_argument("--sur-leg", type=int, default=48,
                    help="training scenarios of the bounded surrogate leg of the default run (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--fp64-steps", type=int, default=20, help="steps of the FP64-build error check (0 = skip)")
    ap.add_argument("--zchunks", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--graph", action="store_true", help="replay captured 4-step cycles (cgfd_set_graph)")
    ap.add_argument("--variant", type=int, default=0,
                    help="stage kernel: 0 = LDS-staged planes, 1 = direct loads, 2 = split velocity/stress")
    ap.add_argument("--tile-width", type=int, default=0, choices=[0, 32, 64],
                    help="tile width of the LDS stage kernel (0: the one the grid's rows pick)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed steps, run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU decomposition is out of scope).", file=sys.stderr)
        sys.exit(2)

    # torch before the engine: torch and libcgfd3d.so both bind
    # libamdhip64.so.7, the first one loaded serves the whole process, and
    # torch (the surrogate leg) only initialises on the runtime it ships with
    import torch  # noqa: F401

    if args.workload == "batch":
        batch_bench(args)
        return
    if args.workload == "surrogate":
        surrogate_bench(args)
        return

    import cgfd3d

    nt = args.warmup + args.steps + (PROFILE_STEPS if args.graph else 0)
    t0 = time.perf_counter()
    sc = build_workload(args.workload, nt)
    log(f"workload {sc['name']} built in {time.perf_counter() - t0:.1f} s")
    sim = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
    sim.set_variant(args.variant)
    sim.set_tile_width(args.tile_width)
    if args.graph:
        sim.set_graph(True)
    if args.zchunks or args.rows:
        sim.set_tuning(args.zchunks, args.rows)
    t0 = time.perf_counter()
    sim.setup(sc)
    sim.sync()
    log(f"inputs uploaded in {time.perf_counter() - t0:.1f} s")

    sim.run(args.warmup)
    sim.sync()
    log(f"{args.warmup} warmup steps done")
    # kernel durations come from HIP events around each launch; graph replay
    # is inactive while they are recorded, so a graph run is timed without
    # them and the kernel durations come from PROFILE_STEPS eager steps after it
    sim.profile(not args.graph)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    t = time.perf_counter() - t0
    if args.graph:
        sim.profile(True)
        sim.run(PROFILE_STEPS)
        sim.sync()
    prof = sim.profile_read()
    sim.profile(False)
    log(f"{args.steps} timed steps in {t:.3f} s")

    nv = sc["ni"] * sc["nj"] * sc["nk"]
    value = nv * 4 * args.steps / t / 1e9
    stage_ms = prof["stage_ms"] / max(prof["stage_n"], 1)
    bytes_stage = algorithmic_bytes(sc)
    bytes_loaded = sim.stage_bytes()
    tile_width = sim.tile_width() if args.variant == 0 else None
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    seis = sim.get_seismo()
    finite = bool(abs(seis).max() < float("inf"))
    sim.close()
    del sim
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)

    fp64 = None
    if args.fp64_steps > 0:
        log(f"FP64 reference run ({args.fp64_steps} steps)")
        fp64 = fp64_error(sc, args.fp64_steps)
        log(f"FP32 vs FP64: max relative L2 {fp64['rel_l2_max']:.2e}")

    cpu = parity = None
    if args.cpu_steps > 0:
        log("timing the CPU baseline")
        cpu, parity = cpu_baseline(sc, args.cpu_steps)
        log(f"CPU baseline {cpu['value']:.4f} Gcell-RKstages/s on {cpu['cores']} threads")
        log(f"GPU vs oracle after {args.cpu_steps} steps: {'identical' if parity['identical'] else parity['differences']}")

    leg = None
    if args.sur_leg > 0 and args.workload == "vonkarman":
        # the other half of the metric, bounded (about two minutes):
        # sequential vs batched FD vs the gated surrogate on the 192x192x128
        # medium, with a training set of --sur-leg runs (its errors are those
        # of so few runs; the long studies use thousands)
        log(f"surrogate leg: {args.sur_leg} training runs on the 192x192x128 medium")
        leg = surrogate_study("192x192x128", args.sur_leg, 16, 16, 120, 64, 16, 16, True, 2)

    cfg = describe(sc)
    cfg.update({"graph_replay": bool(args.graph), "tile_width": tile_width,
                "stage_kernel_ms": round(stage_ms, 4), "timg_kernel_ms": round(prof["timg_ms"] / max(prof["timg_n"], 1), 4),
                "bytes_per_stage_model": bytes_stage, "bytes_per_stage_loaded": bytes_loaded, "fp64_oracle": fp64,
                "oracle_parity": parity, "surrogate_leg": leg,
                "hbm_frac_of_loaded_bytes": round(bytes_loaded / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "finite_output": finite})
    line = {
        "metric": "Gcell-RKstages/s",
        "value": round(value, 4),
        "unit": "Gcell-RKstages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(sc["name"], args.variant)},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["vs_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
