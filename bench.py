#!/usr/bin/env python3
"""Benchmark of the single-GPU FFT path: 512^3 f32 R2C + normalized C2R
round trips per second on one MI355X.

One step = one 3D R2C forward of a 512^3 real N(0,1) field (seed 2) followed
by the normalized C2R inverse back to real space, inputs resident in HBM,
through the C ABI (pencil_fft_hip). Prints ONE JSON line.

roofline: the dominant kernels are the line-stage launches (k_line_fast
<float, 16, 16|32, 0, 16> at 512^3: every stage of both directions, 6 per
round trip). achieved = algorithmic bytes per launch (one read of the stage's
input array + one write of its output array) / average launch duration in the
back-to-back timed loop (HIP events around the loop on the stream the kernels
run on, divided by the launches of a round trip). traffic = HBM bytes per
launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
(profiles/line_stages_512cube_f32_pmc.md).
accuracy: forward, normalized inverse and raw backward against an FP64
scipy.fft reference of the same input, per dtype (after the timed region).
cpu_baseline: the oracle's f32 restatement of the reference route (per-line
FFT with copy-in/out, pack / self-exchange / unpack transitions) timed on the
host: one thread (single rank) and 16 threads, with numpy / scipy lines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pencil-rs_amd"))

METRIC = "512³ f32 R2C+C2R round-trips/s, % of HBM roofline; batched & adjoint vs plain"
PEAK_GBPS = 8000.0


def route_model_bytes(n: int, real_bytes: int) -> int:
    """Route-model bytes of one round trip: (R + 5C) per direction."""
    R = n ** 3 * real_bytes
    C = n * n * (n // 2 + 1) * 2 * real_bytes
    return 2 * (R + 5 * C)


def traffic_from_profiles():
    """HBM bytes per line-stage launch of the round trip from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration): its line "hbm_bytes_per_launch = N", or None."""
    path = os.path.join(ROOT, "profiles", "line_stages_512cube_f32_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("hbm_bytes_per_launch"):
                    return float(line.split("=")[1].strip().split()[0])
    except OSError:
        pass
    return None


def copy_calibration(torch, reps: int = 5):
    """A plain device copy of a known size (1 GiB read + 1 GiB write per
    launch): the reference point that scales FETCH_SIZE / WRITE_SIZE."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": reps}))


def _cpu_info():
    """nproc and the CPU model line of lscpu (the host the baseline ran on)."""
    import subprocess

    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count(), model


def _reps(fn, reps: int, warmup: int):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts


def cpu_baseline(n: int, threads: int, reps: int = 3):
    """The oracle's f32 restatement of the reference route (per-line FFT with
    copy-in / copy-out, pack / self-exchange / unpack transitions): one
    512^3 R2C forward + normalized C2R inverse on 1 thread (the single-rank
    run, one repetition) and on `threads` threads (1 warm-up, then `reps`
    repetitions: median and min), plus third-party library lines
    (numpy.fft, one thread; scipy.fft with workers) on the same input."""
    import hashlib

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rng = np.random.default_rng(2)
    x = rng.standard_normal((n, n, n), dtype=np.float32)
    sha = hashlib.sha256(x.tobytes()).hexdigest()
    errs = []

    def restated(th):
        def run():
            spec = O.transform(O.KIND_R2C, x, O.FORWARD, shape=(n, n, n), threads=th)
            back = O.transform(O.KIND_R2C, spec, O.INVERSE, shape=(n, n, n), threads=th)
            errs.append(float(np.linalg.norm(back - x.reshape(-1)) / np.linalg.norm(x)))
        return run

    single = _reps(restated(1), 1, 0)
    multi = _reps(restated(threads), reps, 1)

    def np_rt():
        s = np.fft.rfftn(x)
        np.fft.irfftn(s, s=x.shape, axes=(0, 1, 2))

    npt = _reps(np_rt, reps, 1)
    third = {"numpy.fft rfftn+irfftn (1 thread)": {"median_s": npt[len(npt) // 2], "min_s": npt[0]}}
    try:
        import scipy.fft as sf

        def sp_rt():
            s = sf.rfftn(x, workers=threads)
            sf.irfftn(s, s=x.shape, axes=(0, 1, 2), workers=threads)

        spt = _reps(sp_rt, reps, 1)
        third[f"scipy.fft rfftn+irfftn (workers={threads})"] = {"median_s": spt[len(spt) // 2], "min_s": spt[0]}
    except ImportError:
        pass
    nproc, model = _cpu_info()
    med = multi[len(multi) // 2]
    return {
        "value": 1.0 / med,
        "unit": "round_trips/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full {n}^3 f32 R2C forward + normalized C2R inverse of the oracle's restatement of the "
                  f"reference route (restatement, not the reference: the Rust crates cannot be built here) on "
                  f"{threads} threads, median of {reps} after 1 warm-up; value = 1 / median",
        "median_s": med,
        "min_s": multi[0],
        "reps": reps,
        "single_rank": {"threads": 1, "s": single[0], "round_trips_per_s": 1.0 / single[0], "reps": 1},
        "round_trip_rel_l2": max(errs),
        "third_party": third,
        "nproc": nproc,
        "cpu_model": model,
        "threads_note": "threads = min(16, nproc): the host cores one GPU's job may use",
        "input": f"numpy default_rng(2) standard_normal f32 {n}^3, sha256 {sha}",
    }


def fp64_accuracy(torch, pf, plan, x, n: int, threads: int):
    """Relative L2 error of the device forward, normalized inverse and raw
    backward against an FP64 reference (scipy.fft in f64 on the host, after
    the timed region): forward vs rfftn(x); inverse of the reference spectrum
    vs x; raw backward vs n^3 x."""
    import numpy as np
    import scipy.fft as sf

    x64 = x.double().cpu().numpy().reshape(n, n, n)
    ref = sf.rfftn(x64, workers=threads)                 # [x][y][kz], complex128
    ref_mem = np.ascontiguousarray(ref.transpose(2, 1, 0)).reshape(-1)  # plan output order [kz][ky][kx]
    cdt = torch.complex64 if x.dtype == torch.float32 else torch.complex128
    ws = plan.allocate_workspace()
    spec = plan.forward(x, workspace=ws)
    got = spec.cpu().numpy().astype(np.complex128)
    fwd = float(np.linalg.norm(got - ref_mem) / np.linalg.norm(ref_mem))
    del got, spec
    s_ref = torch.from_numpy(ref_mem).to(cdt).cuda()
    del ref, ref_mem
    inv = plan.inverse(s_ref, workspace=ws).double().cpu().numpy()
    xv = x64.reshape(-1)
    inv_err = float(np.linalg.norm(inv - xv) / np.linalg.norm(xv))
    del inv
    bwd = plan.backward(s_ref, workspace=ws).double().cpu().numpy()
    scale = float(n) ** 3
    bwd_err = float(np.linalg.norm(bwd - scale * xv) / (scale * np.linalg.norm(xv)))
    return {"forward_rel_l2": fwd, "inverse_rel_l2": inv_err, "backward_rel_l2": bwd_err,
            "bound": 2e-6 if x.dtype == torch.float32 else 1e-12,
            "reference": "scipy.fft rfftn in f64 (pocketfft) of the same input"}


def batched_vs_plain(torch, pf, reps: int = 5):
    """96 independent 3D C2C f32 problems, edges drawn per axis from
    {48,60,64,84,96,112,128,160}, extra dim [3], inputs N(0,1): one
    autotuned persistent launch per stage replayed from a hipGraph, against
    one C2cPlan launch sequence per problem."""
    import numpy as np

    rng = np.random.default_rng(20261202)
    edges = (48, 60, 64, 84, 96, 112, 128, 160)
    shapes = [tuple(int(e) for e in rng.choice(edges, size=3)) for _ in range(96)]
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    srcs = [torch.randn(int(np.prod(s)) * 3, generator=g, dtype=torch.complex64, device="cuda") for s in shapes]
    dsts = [torch.empty_like(t) for t in srcs]
    # the tunings go to a file of this run (unless PFH_TUNE_CACHE names one),
    # so the first creation measures; a second creation, after the process
    # map is cleared as in a new process, reads the file
    tune_dir = None
    if "PFH_TUNE_CACHE" not in os.environ:
        import tempfile

        tune_dir = tempfile.mkdtemp(prefix="pfh_tune_")
        os.environ["PFH_TUNE_CACHE"] = os.path.join(tune_dir, "batch_tuning.txt")
    pf.clear_tuning_cache()
    t0 = time.perf_counter()
    batch = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=[(3,)] * 96)
    tune_s = time.perf_counter() - t0
    first_source = batch.tuning_source
    pf.clear_tuning_cache()
    t0 = time.perf_counter()
    again = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=[(3,)] * 96)
    cached_s = time.perf_counter() - t0
    cached = {"create_s": cached_s, "source": again.tuning_source, "same_tuning": again.tuning() == batch.tuning()}
    del again
    if tune_dir is not None:
        import shutil

        del os.environ["PFH_TUNE_CACHE"]
        shutil.rmtree(tune_dir, ignore_errors=True)
    bws = batch.allocate_workspace()
    plans = [pf.C2cPlan(s, extra_shape=(3,)) for s in shapes]
    pws = [p.allocate_workspace() for p in plans]

    def run_batch():
        batch.forward(srcs, dsts, bws, use_graph=True)

    def run_plain():
        for p, a, b, w in zip(plans, srcs, dsts, pws):
            p.forward(a, b, w)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    tb = timed(run_batch)
    tp = timed(run_plain)
    elems = sum(int(np.prod(s)) * 3 for s in shapes)
    return {
        "workload": "96 x 3D C2C f32 forward, edges from {48,60,64,84,96,112,128,160}, extra [3]",
        "batched_ms": tb * 1e3,
        "plain_ms": tp * 1e3,
        "speedup": tp / tb,
        "batched_points_per_s": elems / tb,
        "tuning": list(batch.tuning()),
        "autotune_s": tune_s,
        "autotune_source": first_source,
        "cached_creation": cached,
    }


def adjoint_vs_plain(torch, pf, plan, x, spec, back, ws, reps: int = 5):
    """Adjoint round trip (adjoint of the C2R inverse, then adjoint of the R2C
    forward) against the plain forward + inverse on the same plan."""

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    def plain():
        plan.forward(x, spec, ws)
        plan.inverse(spec, back, ws, flags=pf.EXEC_DEFER_STATUS)

    def adjoint():
        plan.adjoint_inverse(x, spec, ws)
        plan.adjoint_forward(spec, back, ws)

    tp = timed(plain)
    ta = timed(adjoint)
    return {"plain_ms": tp * 1e3, "adjoint_ms": ta * 1e3, "adjoint_over_plain": ta / tp}


def f64_round_trips(torch, pf, n: int, steps: int = 20, threads: int = 16):
    """The same round trip in f64 (the reference's other precision at this
    size): rate, route-model bytes and roofline fraction, accuracy."""
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randn(n ** 3, generator=g, device="cuda", dtype=torch.float64)
    plan = pf.R2cPlan((n, n, n), dtype=torch.float64)
    spec = torch.empty(plan.output_elements(), dtype=torch.complex128, device="cuda")
    back = torch.empty_like(x)
    ws = plan.allocate_workspace()

    def step():
        plan.forward(x, spec, ws)
        plan.inverse(spec, back, ws, flags=pf.EXEC_DEFER_STATUS)

    step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    plan.take_status(ws)
    rm = route_model_bytes(n, 8)
    err = float(torch.linalg.norm(back - x) / torch.linalg.norm(x))
    del spec, back
    acc = fp64_accuracy(torch, pf, plan, x, n, threads)
    del x, ws
    return {"round_trips_per_s": 1.0 / dt, "ms_per_round_trip": dt * 1e3, "route_model_bytes": rm,
            "route_model_frac_of_peak": rm / dt / 1e9 / PEAK_GBPS, "round_trip_rel_l2": err,
            "accuracy_vs_fp64": acc}


def r2r_dht_round_trips(torch, pf, shape=(256, 384, 240), steps: int = 20):
    """Mixed-kind R2R and DHT on a 256 x 384 x 240 real grid (lengths with
    factors 2, 3, 5; the DST-I axis runs a 385 = 5 x 7 x 11 point FFT):
    forward + normalized inverse per dtype, the route model's bytes (three
    stages, each reading and writing the field once, per operation) and the
    round-trip error."""
    K = pf.R2rKind
    n = shape[0] * shape[1] * shape[2]
    out = {"workload": "256x384x240 real: R2R (DCT-II, DST-I, DCT-IV) and DHT, forward + normalized inverse"}
    for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = torch.Generator(device="cuda")
        g.manual_seed(3)
        x = torch.rand(n, generator=g, device="cuda", dtype=dt) * 2 - 1
        y = torch.empty_like(x)
        back = torch.empty_like(x)
        res = {}
        for tag, plan in (("r2r", pf.R2rPlan(shape, [K.DctII, K.DstI, K.DctIV], dtype=dt)),
                          ("dht", pf.DhtPlan(shape, dtype=dt))):
            ws = plan.allocate_workspace()

            def step():
                plan.forward(x, y, ws)
                plan.inverse(y, back, ws)

            step()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            dt_s = (time.perf_counter() - t) / steps
            rm = 2 * 6 * n * x.element_size()
            res[tag + "_pair_ms"] = dt_s * 1e3
            res[tag + "_route_model_frac_of_peak"] = rm / dt_s / 1e9 / PEAK_GBPS
            res[tag + "_round_trip_rel_l2"] = float(torch.linalg.norm(back - x) / torch.linalg.norm(x))
            del ws
        out[name] = res
        del x, y, back
    return out


def long_lines(torch, pf, reps: int = 20):
    """Local C2C f32 forward on lines that run as pieces: 2^16 and 2^20
    (four-step split, two line-stage launches), 10007 (Bluestein, four
    launches of length 20480 split into 128 x 160), against 8192-point lines
    (one launch), about 8.4M points each. GB/s counts one read and one write
    of the array (what a one-launch stage moves)."""
    out = []
    for n, batch in ((8192, 1024), (65536, 128), (1 << 20, 8), (10007, 838)):
        x = torch.randn(batch * n, dtype=torch.complex64, device="cuda")
        y = torch.empty_like(x)
        plan = pf.LocalC2cPlan(n, dtype=torch.complex64)
        ws = plan.allocate_workspace()
        plan.forward(x, y, ws)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.forward(x, y, ws)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        times.sort()
        ms = times[len(times) // 2]
        nbytes = 2 * x.numel() * x.element_size()
        out.append({"n": n, "lines": batch, "ms": ms, "GBps_one_pass": nbytes / (ms * 1e-3) / 1e9,
                    "workspace_MiB": ws.numel() / 2 ** 20})
    return {"workload": "LocalC2cPlan f32 forward, ~8.4M points per call, median of 20 (HIP events)",
            "cases": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-reps", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and adjoint comparisons")
    ap.add_argument("--copy-calibration", action="store_true", help="only run the 1 GiB copy kernel")
    ap.add_argument("--batched-only", action="store_true", help="only run the batched comparison")
    ap.add_argument("--r2r-only", action="store_true", help="only run the 256x384x240 R2R / DHT pairs")
    ap.add_argument("--long-only", action="store_true", help="only run the long-line (split / Bluestein) timings")
    ap.add_argument("--loop-only", action="store_true",
                    help="only the warm-up and the timed loop (for a kernel-trace profile of the timed region)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (this build has no "
              "multi-GPU path: each route transition would become an exchange within a process-grid row "
              "or column over RCCL)", file=sys.stderr)
        sys.exit(2)

    import torch

    import pencil_fft_hip as pf

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.batched_only:
        print(json.dumps(batched_vs_plain(torch, pf)))
        return
    if args.r2r_only:
        print(json.dumps(r2r_dht_round_trips(torch, pf, steps=min(args.steps, 20))))
        return
    if args.long_only:
        print(json.dumps(long_lines(torch, pf)))
        return
    threads = min(16, os.cpu_count() or 1)
    n = args.size
    shape = (n, n, n)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randn(n ** 3, generator=g, device="cuda", dtype=torch.float32)
    plan = pf.R2cPlan(shape, dtype=torch.float32)
    spec = torch.empty(plan.output_elements(), dtype=torch.complex64, device="cuda")
    back = torch.empty(n ** 3, dtype=torch.float32, device="cuda")
    ws = plan.allocate_workspace()

    def step():
        plan.forward(x, spec, ws)
        plan.inverse(spec, back, ws, flags=pf.EXEC_DEFER_STATUS)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # HIP events on the stream the library launches on (torch's current
    # stream) bracket the same timed loop as the host clock
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan.take_status(ws)  # raises InvalidSpectrum if any inverse was rejected
    rt_per_s = args.steps / elapsed
    ms = elapsed / args.steps * 1e3
    loop_dev_ms = ev0.elapsed_time(ev1) / args.steps
    if args.loop_only:
        print(json.dumps({"loop_only": True, "steps": args.steps, "ms_per_step": ms,
                          "device_ms_per_step": loop_dev_ms, "round_trips_per_s": rt_per_s}))
        return

    # accuracy of the round trip (f64 on device)
    rt_err = float(torch.linalg.norm((back - x).double()) / torch.linalg.norm(x.double()))

    # per-kernel HIP-event timing on the same stream
    # (a stage run slab by slab is one entry with several launches)
    stage_ms, stage_bytes, check_ms, launches = [], [], [], 0
    for _ in range(args.timing_reps):
        _, tf = plan.execute_with_timing(pf.FORWARD, x, spec, ws)
        _, ti = plan.execute_with_timing(pf.INVERSE, spec, back, ws)
        for t in (tf, ti):
            for (axis, t_ms, b), nl in zip(t.entries(), t.launches()):
                if axis < 0:
                    check_ms.append(t_ms)
                else:
                    stage_ms.append(t_ms)
                    stage_bytes.append(b)
                    launches += nl
    iso_ms = sum(stage_ms) / launches
    avg_bytes = sum(stage_bytes) / launches
    lpr = launches / args.timing_reps
    # average line-stage launch in the back-to-back timed loop: the loop's
    # device time per round trip over the launches of one round trip; the
    # endpoint check (two small launches) stays inside it, so this is an
    # upper bound on the line-stage launch time
    avg_ms = loop_dev_ms / lpr
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    stage_avg_ms = sum(stage_ms) / len(stage_ms)
    traffic = traffic_from_profiles()
    rm = route_model_bytes(n, 4)
    route_gbps = rm / (ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": rt_per_s,
        "unit": "round_trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic i.i.d. N(0,1), seed 2 (torch cuda generator)",
        "config": {
            "workload": f"3D R2C forward + normalized C2R inverse, {n}x{n}x{n} f32 real -> {n}x{n}x{n // 2 + 1} "
                        "complex, 1x1 pencil grid, permute_dims",
            "shape": list(shape),
            "output_layout": "[kz][ky][kx]",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_GBPS,
            "traffic": traffic,
            "kernel": "line stages (k_line_fast<float,16,16|32,0,16,NT>)",
            "avg_launch_ms": avg_ms,
            "avg_launch_ms_method": "HIP events around the timed loop on the launch stream: device time per "
                                    "round trip / line-stage launches per round trip (endpoint check included)",
            "isolated_launch_ms": iso_ms,
            "algorithmic_bytes_per_launch": avg_bytes,
            "launches_per_round_trip": lpr,
            "avg_stage_ms": stage_avg_ms,
            "loop_device_ms_per_round_trip": loop_dev_ms,
        },
        "route_model": {
            "bytes_per_round_trip": rm,
            "achieved_GBps": route_gbps,
            "frac_of_peak": route_gbps / PEAK_GBPS,
            "endpoint_check_ms": (sum(check_ms) / len(check_ms)) if check_ms else 0.0,
        },
        "accuracy": {"round_trip_rel_l2": rt_err},
    }
    out["accuracy"].update(fp64_accuracy(torch, pf, plan, x, n, threads))
    if not args.no_extras:
        out["adjoint"] = adjoint_vs_plain(torch, pf, plan, x, spec, back, ws)
        plan.take_status(ws)
        out["batched"] = batched_vs_plain(torch, pf)
        out["f64"] = f64_round_trips(torch, pf, n, threads=threads)
        out["r2r_dht"] = r2r_dht_round_trips(torch, pf)
        out["long_lines"] = long_lines(torch, pf)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
