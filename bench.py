#!/usr/bin/env python3
"""Frames/s of the calcIR spectroscopy hot path on one MI355X.

Workload (the 2048-OH one-box configuration): 1024 rigid TIP4P/2005 waters
(2048 OH chromophores) in a cubic box L = 3.128 nm, 2000 seeded OU frames
10 fs apart, O-H field cutoff 0.7831 nm, 150-point correlation functions
(1.5 ps), time origins every 20 frames (93 samples), T1 = 260 fs, IR + VV/VH
Raman, omega 2800-4000 cm^-1, C2R to 2048 points.

One bench step is the whole job over that trajectory (every sample, every
(origin, tcf point) step, lineshapes) with the trajectory resident in HBM.
"frames/s" is (origin, tcf point) steps per second (nsamples * ntcfpoints /
wall time), the trip count of the reference's sample x frame loop.

Timed: the batched mixed-precision path (each trajectory frame's field,
Hamiltonian and eigensolve computed once; all active origins propagated
together with split-BF16x3 MFMA products).  Reported beside it: the batched
FP32 path, the plain FP32 path with the reference control flow (timed on a
bounded subset), the spectrum error of each form against the FP64 pipeline,
the roofline of the dominant kernel and the CPU oracle baseline.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s at 2048 OH: mixed-prec batched vs plain FP32, ≤1e-3 rel spectrum err"
NWATER, NFRAMES, SEED, BOX = 1024, 2000, 1002, 3.128
WORK = dict(species="H2O", model="tip4p2005", map=0, ntcfpoints=150, sampleEvery=0.2, avef=3400.0,
            omegaStart=2800.0, omegaStop=4000.0, omegaStep=5, nzeros=1898, t1=0.26, dt=0.010,
            cutoff=0.7831)
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def nsamples_for(nframes):
    # origins every 20 frames, each needing 150 frames
    return (nframes - 150) // 20 + 1


def bulge_bt_flops(n, nb=32):
    """Algorithmic flops per matrix of applying the bulge-chasing reflectors
    to the n eigenvectors: sweep s, task k has a reflector of length
    min(nb, n - r0), r0 = s + 1 + k nb, and H = I - tau v v^T costs a dot and
    an update (4 len flops) per column."""
    tot = 0
    for s in range(n - 2):
        for r0 in range(s + 1, n - 1, nb):
            tot += min(nb, n - r0)
    return 4.0 * n * tot


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "q2_apply_kernel_2048oh_traffic.json")


def traffic_per_launch(matrices_per_launch):
    """HBM bytes per q2_apply_kernel launch from the committed counter
    measurement (FETCH_SIZE + WRITE_SIZE per matrix, each scaled by the same
    counter's reading for a 1 GiB copy); None when that file is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return float(t["hbm_bytes_per_matrix"]) * matrices_per_launch
    except (OSError, KeyError, ValueError):
        return None


def traffic_from(fetch_db, write_db, frames, kernel="q2_apply_kernel"):
    """The committed traffic file from two rocprofv3 counter runs of
    `bench.py --frames F --steps 1 --warmup 0 --no-extras --no-cpu
    --copy-calibration` (FETCH_SIZE in one, WRITE_SIZE in the other; their
    rocpd databases): each counter's reading for the kernel's dispatch,
    scaled by what the same counter reads for the 1 GiB calibration copy, per
    matrix of the dispatch (all distinct frames of the run, one chunk)."""
    import sqlite3

    def per_dispatch(db, counter):
        con = sqlite3.connect("file:%s?mode=ro" % db, uri=True)
        rows = con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                           "where counter_name = ? group by dispatch_id, kernel_name", (counter,)).fetchall()
        con.close()
        # The largest dispatch of the kernel is the run's one chunk; smaller
        # ones time its launch shapes over a few matrices on first use.
        k = [v for _, name, v in rows if kernel in name]
        cp = [v for _, name, v in rows if "copyBuffer" in name]
        if not k or not cp:
            raise SystemExit("%s: expected a %s dispatch and the calibration copy, found %d and %d"
                             % (db, kernel, len(k), len(cp)))
        return float(max(k)), float(max(cp))

    nmat = 150 + (nsamples_for(frames) - 1) * 20
    fk, fc = per_dispatch(fetch_db, "FETCH_SIZE")
    wk, wc = per_dispatch(write_db, "WRITE_SIZE")
    gib_kib = 1024.0 * 1024.0
    fs, ws = gib_kib / fc, gib_kib / wc
    hbm = (fk * fs + wk * ws) * 1024.0 / nmat
    n = 2 * NWATER
    alg = 2.0 * 4.0 * n * n  # Z read and written once
    out = {"kernel": kernel, "workload": "2048-OH box, N=2048, ld=2048, %d frames" % frames,
           "matrices_per_dispatch": nmat, "fetch_kib_raw": fk, "write_kib_raw": wk,
           "fetch_scale": fs, "write_scale": ws, "copy_fetch_kib_per_gib": fc, "copy_write_kib_per_gib": wc,
           "hbm_bytes_per_matrix": hbm, "eigenvector_bytes_per_matrix": alg, "traffic_over_eigenvectors": hbm / alg}
    return out


def lineshape(p, ft):
    """(freq, value) rows as written to <outf>_irls.dat (calcIR.cu:893-919)."""
    nt, nz = p.ntcfpoints, p.nzeros
    n = nt + nz
    factor = np.float32(2 * 3.14159265359 * 5.308837367 / (np.float32(p.dt) * np.float32(n)))
    fr, va = [], []
    for i in range(n // 2, n):
        f = np.float32(-(i - nt - nz)) * factor + np.float32(p.avef)
        if f <= p.omegaStop:
            fr.append(f)
            va.append(ft[i] / (factor * n))
    for i in range(0, nt + nz // 2):
        f = np.float32(-i) * factor + np.float32(p.avef)
        if f >= p.omegaStart:
            fr.append(f)
            va.append(ft[i] / (factor * n))
    return np.array(fr, np.float64), np.array(va, np.float64)


def spectrum_error(p, res, ref):
    out = {}
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        f, a = lineshape(p, res[k])
        _, b = lineshape(p, ref[k])
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        m1a = float(np.sum(f * a) / np.sum(a))
        m1b = float(np.sum(f * b) / np.sum(b))
        out[k[5:]] = {"rel_l2": l2, "first_moment_shift_cm-1": m1a - m1b}
    return out


CPU_POINTS = 8


def host_info():
    """Threads the CPU legs use and what the host is: a shared GPU host may
    allot each GPU a share of its cores and say so in OMP_NUM_THREADS, which
    the CPU legs follow (the process's CPU affinity where it is unset)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or affinity
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "nproc": nproc, "affinity": affinity, "model": model}


def cpu_c1(C, nthreads):
    """The 256-OH workload (128 waters, L = 1.564 nm, 200 frames, 6 origins x
    150 points, cutoff 0.75 nm) in full: the FP32 CPU oracle with the
    reference-literal propagation beside the GPU plain FP32 and batched BF16x3
    paths on the same frames."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    xyz, box, times = C.synth_trajectory(128, 200, 1001, box_nm=1.564)
    kw = dict(WORK, cutoff=0.75, sampleEvery=0.1, nsamples=6)
    p = C.default_params(precision="fp32", mode="plain", **kw)
    cfg = O.config_from_params(p)
    cfg.propagation = O.PROP_LITERAL
    cfg.nthreads = nthreads
    print("[bench] 256-OH workload on the CPU oracle (%d threads)" % nthreads, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = O.run(O.F32, cfg, xyz, box, times)
    tcpu = time.perf_counter() - t0
    out = {"workload": "128 waters (256 OH), L=1.564 nm, 200 frames, 6 origins x 150 points, cutoff 0.75 nm",
           "cpu_oracle_fp32": {"seconds": tcpu, "frames_per_s": r["nsteps"] / tcpu, "threads": nthreads,
                               "out_of_window_modes": int(r["hist_skipped"])}}
    for name, prec, mode in (("gpu_plain_fp32", "fp32", "plain"), ("gpu_batched_bf16x3", "bf16x3", "batched")):
        pg = C.default_params(precision=prec, mode=mode, **kw)
        C.run(pg, xyz, box, times)  # warm
        t0 = time.perf_counter()
        g = C.run(pg, xyz, box, times)
        tg = time.perf_counter() - t0
        ft = g["ftcf_ir"]
        err = float(np.linalg.norm(ft - r["ftcf_ir"]) / np.linalg.norm(r["ftcf_ir"]))
        out[name] = {"seconds": tg, "frames_per_s": g["nsteps"] / tg, "ir_c2r_rel_l2_vs_cpu_fp32": err}
    return out


def cpu_baseline(xyz, box, times, nthreads):
    """The CPU oracle (FP32, reference-literal three-complex-product
    propagation) on a bounded sample of the workload: one origin x 8 tcf
    points (one F = I step and seven propagated steps, ~10-20 s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    cfg = O.config(imodel=1, ispecies=0, imap=0, cutoff=WORK["cutoff"], avef=WORK["avef"], dt=WORK["dt"],
                   t1=WORK["t1"], ntcfpoints=CPU_POINTS, nsamples=1, nzeros=WORK["nzeros"], sampleEvery=0.2,
                   omegaStart=WORK["omegaStart"], omegaStop=WORK["omegaStop"], omegaStep=5,
                   propagation=O.PROP_LITERAL, nthreads=nthreads)
    t0 = time.perf_counter()
    k = CPU_POINTS
    r = O.run(O.F32, cfg, xyz[:k], box[:k], times[:k])
    dt = time.perf_counter() - t0
    return {"value": r["nsteps"] / dt, "unit": "frames/s", "cores": nthreads, "kind": "port",
            "sample": "%d-OH box, 1 origin x %d tcf points (1 F=I step + %d propagated steps), FP32 oracle, "
                      "literal F-matrix propagation (3 complex products per step), %.1f s"
                      % (2 * (xyz.shape[1] // 4), k, k - 1, dt),
            "phases_s": {"field_hamiltonian": r["t_hamiltonian"], "eigh": r["t_eigh"],
                         "ipr_spectral_density": r["t_hist"], "propagate_correlate": r["t_propagate"],
                         "finalize": r["t_finalize"]}}


def adjoint_bench(C):
    """Lineshape adjoint on the 1024-OH differentiation workload: 512 waters
    (L = 2.483 nm), 400 frames, origins every 10 frames (26 samples x 150
    points); forward + backward time per precision and the FP32 gradient's
    deviation from FP64 (relative sensitivities coef * dL/dcoef)."""
    from calcir_amd.adjoint import LineshapeSession, lineshape_map

    xyz, box, times = C.synth_trajectory(512, 400, 1004, box_nm=2.483)
    kw = dict(WORK, ntcfpoints=150, nsamples=26, sampleEvery=0.1)
    coef = C.map_coefs(0, 0)
    pert = coef * (1 + 0.02 * np.random.default_rng(1004).choice([-1.0, 1.0], coef.size))
    out, grads, target = {}, {}, None
    for prec in ("fp64", "fp32"):
        print("[bench] adjoint %s" % prec, file=sys.stderr, flush=True)
        p = C.default_params(precision=prec, **kw)
        idx, _, scale = lineshape_map(p)
        sess = LineshapeSession(p, xyz, box, times)
        if target is None:
            target = sess.forward(pert)[idx] * scale
        sess.forward(coef)  # warm
        t0 = time.perf_counter()
        ft = sess.forward(coef)
        t1 = time.perf_counter()
        r = ft[idx] * scale - target
        dL = np.zeros_like(ft)
        np.add.at(dL, idx, 2.0 * r * scale)
        grads[prec] = sess.backward(dL)
        t2 = time.perf_counter()
        sess.close()
        out[prec] = {"forward_s": t1 - t0, "backward_s": t2 - t1}
    s64, s32 = coef * grads["fp64"], coef * grads["fp32"]
    out["fp32_vs_fp64_sensitivity_err"] = float(np.abs(s32 - s64).max() / np.abs(s64).max())
    out["workload"] = "512 waters, 400 frames, 26 samples x 150 points, 13 map coefficients"
    return out


def uneven_boxes_bench(C, nbox=64, nframes=500):
    """Independent boxes of uneven size (water counts drawn from
    {64,128,256,512,768,1024}, seeded), one after another and at once
    (cir_run_many); BF16x3 batched path, origins every 10 frames."""
    rng = np.random.default_rng(3000)
    counts = rng.choice([64, 128, 256, 512, 768, 1024], nbox)
    kw = dict(WORK, precision="bf16x3", sampleEvery=0.1, nsamples=(nframes - 150) // 10 + 1)
    jobs = []
    for b, nw in enumerate(counts):
        xyz, box, times = C.synth_trajectory(int(nw), nframes, 3000 + b, box_nm=C.synth_box_nm(int(nw)))
        jobs.append((C.default_params(**kw), xyz, box, times))
    print("[bench] uneven boxes (%d x %d frames)" % (nbox, nframes), file=sys.stderr, flush=True)
    C.run(*jobs[-1])  # warm
    t0 = time.perf_counter()
    for j in jobs[:16]:
        C.run(*j)
    tseq16 = time.perf_counter() - t0
    t0 = time.perf_counter()
    res16 = C.run_many(jobs[:16], nthreads=8)
    tmany16 = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = C.run_many(jobs, nthreads=8)
    tmany = time.perf_counter() - t0
    steps16 = sum(int(r["nsteps"]) for r in res16)
    steps = sum(int(r["nsteps"]) for r in res)
    return {"boxes": nbox, "frames_per_box": nframes, "waters": [int(c) for c in counts],
            "frames_per_s": steps / tmany, "trajectory_frames_per_s": sum(int(r["nframes_used"]) for r in res) / tmany,
            "seconds": tmany, "steps": steps,
            "first16_sequential_frames_per_s": steps16 / tseq16, "first16_concurrent_frames_per_s": steps16 / tmany16,
            "first16_concurrent_speedup": tseq16 / tmany16,
            "sample": "%d boxes x %d frames, origins every 10 frames, bf16x3 (host arrays, transfers included)"
                      % (nbox, nframes)}


def guarded(name, fn):
    """One of the reported extras: a failure is recorded in its place (and
    on stderr), so that the timed result's line is still printed."""
    try:
        return fn()
    except Exception as e:
        print("[bench] %s failed: %r" % (name, e), file=sys.stderr, flush=True)
        return {"error": repr(e)}


def large_box_bench(C):
    """The 8192-OH box: 4096 TIP4P/2005 waters, L = 4.965 nm, 1000 seeded
    frames, origins every 38 frames (4 concurrent, 23 samples x 150 points),
    BF16x3 batched path with the propagated states re-unitarised every 25
    points; error against the FP64 pipeline on a 40-frame subset (one origin
    x 40 points, re-unitarised at point 25)."""
    xyz, box, times = C.synth_trajectory(4096, 1000, 1005, box_nm=4.965)
    kw = dict(WORK, ntcfpoints=150, sampleEvery=0.38, nsamples=23)
    print("[bench] 8192-OH box, 1000 frames", file=sys.stderr, flush=True)
    p = C.default_params(precision="bf16x3", reunitarise_every=25, **kw)
    t0 = time.perf_counter()
    r = C.run(p, xyz, box, times, profile=True)
    dt = time.perf_counter() - t0
    out = {"metric": "frames/s at 8192 OH (large single box)", "value": r["nsteps"] / dt, "unit": "frames/s",
           "trajectory_frames_per_s": r["nframes_used"] / dt, "seconds": dt, "distinct_frames": r["nframes_used"],
           "steps": r["nsteps"], "phases_s": {k: v[0] for k, v in r["phases"].items()},
           "workload": "4096 waters, 1000 frames, 23 origins x 150 points every 38 frames (4 concurrent), bf16x3, "
                       "re-unitarised every 25 points"}
    sub = dict(WORK, ntcfpoints=40, sampleEvery=0.01, nsamples=1)
    print("[bench] 8192-OH subset, fp64 reference", file=sys.stderr, flush=True)
    a = C.run(C.default_params(precision="fp64", **sub), xyz[:40], box[:40], times[:40])
    pb = C.default_params(precision="bf16x3", reunitarise_every=25, **sub)
    b = C.run(pb, xyz[:40], box[:40], times[:40])
    out["subset_error_vs_fp64"] = spectrum_error(pb, b, a)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=NFRAMES, help="trajectory frames (default 2000)")
    ap.add_argument("--plain-points", type=int, default=20, help="tcf points of the plain-FP32 subset")
    ap.add_argument("--no-extras", action="store_true", help="only the timed path (for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-large-box", action="store_true", help="skip the 8192-OH box among the extras")
    ap.add_argument("--large-box", action="store_true",
                    help="only the 8192-OH box (1000 frames) and its FP64 subset check, as one JSON line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="write the traffic file of the dominant kernel from two counter runs (see traffic_from) "
                         "and exit; needs no GPU")
    args = ap.parse_args()
    if args.traffic_from:
        t = traffic_from(args.traffic_from[0], args.traffic_from[1], args.frames)
        with open(TRAFFIC_FILE, "w") as f:
            json.dump(t, f, indent=1)
        print(json.dumps(t, indent=1))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    import torch

    import calcir_amd as C

    dev = torch.device("cuda:0")
    if args.large_box:
        print(json.dumps(large_box_bench(C)), flush=True)
        return 0
    if args.copy_calibration:
        src = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print("[bench] 1 GiB calibration copy done", file=sys.stderr, flush=True)
    nframes = args.frames
    ns = nsamples_for(nframes)
    xyz, box, times = C.synth_trajectory(NWATER, nframes, SEED, box_nm=BOX)
    xyz_d = torch.from_numpy(xyz).to(dev)
    box_d = torch.from_numpy(box).to(dev)
    torch.cuda.synchronize()

    p = C.default_params(precision="bf16x3", mode="batched", nsamples=ns, **WORK)
    print("[bench] warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        C.run_device(p, xyz_d, box_d, times)
    torch.cuda.synchronize()
    phases = {k: [0.0, 0] for k in C.PHASES}
    t0 = time.perf_counter()
    res = None
    for i in range(args.steps):
        res = C.run_device(p, xyz_d, box_d, times, profile=True)
        for k, (s, c) in res["phases"].items():
            phases[k][0] += s
            phases[k][1] += c
        print("[bench] step %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps_per_run = res["nsteps"]
    value = steps_per_run * args.steps / elapsed

    n = 2 * NWATER
    nmat = res["nframes_used"] * args.steps
    # the dominant kernel: q2_apply_kernel (the bulge-chasing reflectors
    # applied to the eigenvectors on FP32 MFMA), alone in its phase, one
    # launch per chunk of matrices
    bt_s, bt_n = phases["bulge_backtransform"]
    per_launch = bt_s / max(bt_n, 1)
    mats_launch = nmat / max(bt_n, 1)
    fl_bt = bulge_bt_flops(n)
    achieved = fl_bt * mats_launch / max(per_launch, 1e-12) / 1e12
    # the whole eigensolver on its algorithmic flops: 4/3 N^3 for the
    # reduction to tridiagonal form and 2 N^3 for the back-transform
    eig_keys = ("tridiag", "bulge", "bisect", "invit", "bulge_backtransform", "backtransform")
    eig_s = sum(phases[k][0] for k in eig_keys)
    red_s = phases["tridiag"][0] + phases["bulge"][0]
    fl_red, fl_eig = 4.0 / 3.0 * n ** 3, (4.0 / 3.0 + 2.0) * n ** 3
    out = {
        "metric": METRIC, "value": value, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2048-OH box: 1024 TIP4P/2005 waters, L=3.128 nm, %d seeded OU frames at 10 fs, "
                               "cutoff 0.7831 nm, 150-point TCF, origins every 20 frames (%d samples)" % (nframes, ns),
                   "path": "batched", "precision": "f32 field/H/reduction to tridiagonal form, f64 bisection + "
                                                   "inverse iteration, split-bf16x3 MFMA propagator",
                   "nsteps_per_run": steps_per_run, "distinct_frames": res["nframes_used"]},
        "roofline": {"kernel": "q2_apply_kernel", "bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic_per_launch(mats_launch), "avg_launch_ms": 1e3 * per_launch,
                     "matrices_per_launch": mats_launch, "flops_per_matrix": fl_bt},
        "phases_s_per_run": {k: v[0] / args.steps for k, v in phases.items()},
        # launch shapes the library chose by timing on first use (warmup)
        "tuning": {" ".join(str(x) for x in k): {"choice": v[0], "ms": v[1]} for k, v in C.tuning_report().items()},
        "trajectory_frames_per_s": nmat / elapsed,
        "eigensolver": {"matrices": nmat, "n": n, "seconds": eig_s, "ms_per_matrix": 1e3 * eig_s / nmat,
                        "reduction_flops_per_matrix": fl_red, "reduction_s": red_s,
                        "reduction_tflops": fl_red * nmat / max(red_s, 1e-12) / 1e12,
                        "reduction_frac_fp32_peak": fl_red * nmat / max(red_s, 1e-12) / 1e12 / FP32_PEAK_TFLOPS,
                        "eigensolve_flops_per_matrix": fl_eig,
                        "eigensolve_tflops": fl_eig * nmat / max(eig_s, 1e-12) / 1e12,
                        "eigensolve_frac_fp32_peak": fl_eig * nmat / max(eig_s, 1e-12) / 1e12 / FP32_PEAK_TFLOPS},
    }
    if not args.no_extras:
        print("[bench] fp32 batched", file=sys.stderr, flush=True)
        p32 = C.default_params(precision="fp32", mode="batched", nsamples=ns, **WORK)
        # (each form's first run times its launch shapes: warm it first, on
        # 600 frames, whose two chunks of 295 matrices fall in the same size
        # buckets as the full run's chunks)
        def warm(prec):
            pw = C.default_params(precision=prec, mode="batched", nsamples=nsamples_for(600), **WORK)
            C.run_device(pw, xyz_d[:600], box_d[:600], times[:600])
            torch.cuda.synchronize()

        warm("fp32")
        t1 = time.perf_counter()
        r32 = C.run_device(p32, xyz_d, box_d, times)
        torch.cuda.synchronize()
        t32 = time.perf_counter() - t1
        print("[bench] fp64 batched (error reference)", file=sys.stderr, flush=True)
        p64 = C.default_params(precision="fp64", mode="batched", nsamples=ns, **WORK)
        warm("fp64")
        t1 = time.perf_counter()
        r64 = C.run_device(p64, xyz_d, box_d, times)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t1
        print("[bench] plain fp32 subset", file=sys.stderr, flush=True)
        pp = C.default_params(precision="fp32", mode="plain", **dict(WORK, ntcfpoints=args.plain_points, nsamples=1))
        C.run_device(pp, xyz_d, box_d, times)  # warm
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rp = C.run_device(pp, xyz_d, box_d, times)
        torch.cuda.synchronize()
        tplain = time.perf_counter() - t1
        plain_fs = rp["nsteps"] / tplain
        # the same run with per-phase events (launch by launch, no graph)
        rpp = C.run_device(pp, xyz_d, box_d, times, profile=True)
        plain_phases = {k: 1e3 * v[0] / rpp["nsteps"] for k, v in rpp["phases"].items() if v[1] > 0}
        out["batched_fp32"] = {"value": steps_per_run / t32, "unit": "frames/s",
                               "error_vs_fp64": spectrum_error(p32, r32, r64)}
        out["batched_bf16x3_error_vs_fp64"] = spectrum_error(p, res, r64)
        print("[bench] bf16x3 with re-unitarisation every 25 points", file=sys.stderr, flush=True)
        pr = C.default_params(precision="bf16x3", mode="batched", nsamples=ns, reunitarise_every=25, **WORK)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rr = C.run_device(pr, xyz_d, box_d, times)
        torch.cuda.synchronize()
        out["batched_bf16x3_reunit25"] = {"value": steps_per_run / (time.perf_counter() - t1), "unit": "frames/s",
                                          "error_vs_fp64": spectrum_error(pr, rr, r64)}
        del rr
        out["batched_fp64"] = {"value": steps_per_run / t64, "unit": "frames/s"}
        out["plain_fp32"] = {"value": plain_fs, "unit": "frames/s",
                             "sample": "1 origin x %d tcf points, reference control flow" % args.plain_points,
                             "phases_ms_per_step": plain_phases,
                             "note": "value with each step's eigensolve replayed from a captured graph; the phases "
                                     "from a second run timed launch by launch"}
        out["speedup_vs_plain_fp32"] = value / plain_fs
        out["speedup_vs_batched_fp32"] = value * t32 / steps_per_run
        del r32, r64, rp
        out["adjoint_1024oh"] = guarded("adjoint", lambda: adjoint_bench(C))
        out["uneven_boxes"] = guarded("uneven boxes", lambda: uneven_boxes_bench(C))
        if not args.no_large_box:
            out["large_box_8192oh"] = guarded("8192-OH box", lambda: large_box_bench(C))
    if not args.no_cpu:
        host = host_info()
        print("[bench] cpu baseline (%d threads)" % host["threads"], file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(xyz, box, times, host["threads"])
        out["cpu_baseline"]["host"] = host
        if not args.no_extras:
            out["workload_256oh_full"] = guarded("256-OH workload", lambda: cpu_c1(C, host["threads"]))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
