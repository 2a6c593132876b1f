"""Benchmark of the coherent scattering hot path on one MI355X.

Workload: 20,000 atoms in a 60 A box, Gaussian random walk (sigma 0.1 A per
frame) over 2,000 frames, float32; neutron coherent lengths H/C/N/O
(0.5/0.3/0.1/0.1); 40 |q| shells over 0.1-4.0 A^-1 with 200 seeded random
orientations each; coherent "all" scattering with FFT time autocorrelation
(pad 4096).  One step = all 40 shells (3.2e11 atom*frame*vector terms) through
libsassena_amd.so, coordinates already staged in HBM; outputs (fqt, fq, fq2)
are copied back to the host inside the step.

Prints one JSON line.  The CPU baseline is the FP64 oracle (a restatement of
the reference's arithmetic) timed on this host on a bounded sample: one
shell, a few vectors, all atoms and frames.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs, 2.4 GHz, 64 lanes/CU; v_sin_f32 / v_cos_f32 issue at a
# quarter of the full VALU rate.
TRANS_PEAK = 256 * 64 * 2.4e9 / 4.0   # transcendental ops/s
# Published vector peaks of MI355X.  The FP32 figure counts packed FP32 (two
# FMAs per lane and instruction); this build keeps FP32 code off the packed
# instructions, so its reachable FP32 FMA rate is half of it.
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12
# Algorithmic cost of one (atom, frame, q vector) term of the amplitude sum:
# the phase q.r (3 mul + 2 add) and its reduction mod 2 pi (1 mul + 1
# round/sub) in FP64, sin and cos, and b cos, b sin accumulated (2 FMA = 4
# FP32 flops).  Kahan or other compensation is implementation overhead.
ALG_FP64_FLOPS = 7.0
ALG_FP32_FLOPS = 4.0
ALG_TRANS = 2.0
# Executed per term on linearly spaced shells (the stepped-phasor path): two
# FMAs step cos and sin to the next shell and two add b cos, b sin to the
# sums; shells of a reference q scan (residual-corrected steps) add two FMAs
# of the correction sums, except on the shells that lie on their tile's line
# (sx_timing.shell_exact_runs: the middle shells and the walks from them that
# stay in an exact run of the scan).  The FP64 phase and four hardware
# sin/cos per (orientation, atom, frame) are shared by a tile's shells.
EXEC_FLOPS_STEPPED = 8.0
EXEC_FLOPS_RESIDUAL = 12.0
# HBM bytes per amplitude launch on this workload (tile 1 x 20, 4 atom
# ranges), from rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0), the
# factors a 1 GiB device copy reads for the same counters on MI355X.
AMP_TRAFFIC_BYTES = 2.467e9 + 0.524e9


# FP64 flops per (atom, frame, moment, shell) term of the multipole amplitude
# kernels on the mp / mpc workloads: rocprofv3 counts of FP64 FMA (2 flops),
# MUL and ADD wave instructions x 64 lanes over the terms of one launch
# (sphere 4.84e10 terms, cylinder 1.64e10), plus those of the pre-pass that
# tabulates the Bessel functions of every (shell, atom, frame), both inside
# the timed multipole stage: sphere 18.9 + 1.1 (0.92e12 + 0.05e12 flops),
# cylinder 27.9 + 15.5 (0.46e12 + 0.25e12; the cylinder's pre-pass also
# forms the two phases e^{i sgn z q_z} and e^{i psi}).
MP_FP64_PER_TERM = {"sphere": 20.0, "cylinder": 43.4}


def roofline_legs(terms_per_launch, launch_s):
    """The algorithmic legs of the amplitude kernel (per-term unit above)
    against the published peaks; the binding leg is the largest fraction."""
    rate = terms_per_launch / launch_s
    legs = {
        "fp64": {"achieved": ALG_FP64_FLOPS * rate / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                 "frac": ALG_FP64_FLOPS * rate / FP64_PEAK, "per_term": ALG_FP64_FLOPS},
        "fp32": {"achieved": ALG_FP32_FLOPS * rate / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                 "frac": ALG_FP32_FLOPS * rate / FP32_PEAK, "per_term": ALG_FP32_FLOPS},
        "transcendental": {"achieved": ALG_TRANS * rate / 1e12, "peak": TRANS_PEAK / 1e12, "unit": "T/s",
                           "frac": ALG_TRANS * rate / TRANS_PEAK, "per_term": ALG_TRANS,
                           "note": "peak assumes quarter-rate v_sin/v_cos; the stepped path issues about 0.2 "
                                   "per term (sin/cos per orientation, atom and frame only), so this leg can "
                                   "exceed 1"},
    }
    return legs


def log(msg):
    """Progress of the long workloads, on stderr."""
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_threads():
    """Threads of the CPU baseline: the CPUs this process may run on, capped
    at 16, the host-CPU share of one GPU on the measurement hosts."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": aff,
            "threads_used": host_threads(),
            "note": "the baseline runs on the host-CPU share of one GPU (16 threads at most)"}


def cpu_baseline(T, threads):
    """The FP64 oracle on the host: one shell (the last) x 8 vectors per
    thread over all atoms and frames.  Returns the record and the oracle's
    results (for the GPU's error on the same vectors)."""
    import oracle
    frames = np.ascontiguousarray(np.transpose(T.xyz, (2, 1, 0)))
    k = min(T.NM, 8 * threads)
    qv = T.qvectors(T.nshells - 1)[:k]
    t0 = time.perf_counter()
    ref = oracle.compute("all", frames, T.b, qv, "autocorrelate", "fftw", nthreads=threads)
    dt = time.perf_counter() - t0
    terms = k * T.NA * T.NF
    rec = {"value": terms / dt, "unit": "terms/s", "cores": threads, "kind": "port",
           "sample": f"1 shell x {k} vectors x {T.NA} atoms x {T.NF} frames ({terms:.2e} terms) in {dt:.1f} s, "
                     f"FP64 libm sin/cos, {threads} threads on {cpu_model()}",
           "host": host_info()}
    return rec, ref, k


def fft_flops(L):
    """Algorithmic flops of one zero-padded series: two complex FFTs of L/2
    points (the even and odd bins of the length-L transform), 5 n log2 n each."""
    import math
    n = L // 2
    return 2 * 5.0 * n * math.log2(n)


def bench_c3(args):
    """Incoherent "self": 50,000 atoms (60% H, 25.27 fm; 40% C/O, zero
    incoherent length), Ornstein-Uhlenbeck walks (sigma 0.2 A, tau 50
    frames) over 5,000 frames, 20 shells over 0.2-3.0 A^-1 x 100 vectors,
    FP32 FFT autocorrelation of length 10,240 per (atom, vector), FP64 shell
    averaging.  One step = all 20 shells."""
    from sassena_amd import ScatterDevice, synth
    T = synth.case("c3")
    dev = ScatterDevice("self", "autocorrelate", "fftw", profile=True)
    dev.stage(T.xyz, layout="soa")
    dev.set_factors(T.b)
    step = lambda: dev.compute_oriented(T.qlen, T.orient)
    for _ in range(args.warmup):
        step()
    k_ms = dev_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        tm = dev.timing()
        k_ms += tm["amplitude_ms"]
        dev_ms += tm["total_ms"]
        launches += tm["correlate_launches"]
        L = tm["pad_length"]
    wall = time.perf_counter() - t0
    dev.close()
    terms = T.terms() * args.steps
    # zero-length atoms are skipped by the kernel; count the series it transforms
    nz = int(np.count_nonzero(T.b))
    series_per_launch = nz * T.NM * T.nshells * args.steps / launches
    k_avg_s = k_ms / launches * 1e-3
    achieved = fft_flops(L) * series_per_launch / k_avg_s
    out = {
        "metric": "atom*frame*q-vector terms/s (incoherent self F(q,t), per-pair FFT autocorrelation)",
        "value": terms / wall, "unit": "terms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 FFT / f64 phase and shell sums", "data": "synthetic",
        "config": {"workload": "incoherent self, 50000 atoms x 5000 frames, 20 shells x 100 vectors, pad 10240",
                   "NA": T.NA, "NF": T.NF, "shells": T.nshells, "vectors": T.NM, "pad": L,
                   "nonzero_atoms": nz, "terms_per_step": T.terms()},
        "kernel_ms_per_step": {"corr_self": k_ms / args.steps, "device_total": dev_ms / args.steps},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": achieved / 157.3e12, "traffic": None, "kernel": "corr_self_kernel",
                     "note": "FP32 FFT flops 2 x 5 (L/2) log2(L/2) per (atom, vector) series over the fused "
                             "amplitude + FFT kernel time; peak is the published FP32 vector rate"},
    }
    if not args.no_cpu:
        import oracle
        threads = host_threads()
        # about 10 s of host work: 300 atoms per thread x 16 vectors of the
        # last shell over all frames
        idx = np.flatnonzero(T.b)[:threads * 300]
        atoms = np.ascontiguousarray(np.transpose(T.xyz[:, idx, :], (1, 2, 0)))
        qv = T.qvectors(T.nshells - 1)[:16]
        t1 = time.perf_counter()
        oracle.compute("self", atoms, T.b[idx], qv, "autocorrelate", "fftw", nthreads=threads)
        dt = time.perf_counter() - t1
        n = len(idx) * len(qv) * T.NF
        out["cpu_baseline"] = {"value": n / dt, "unit": "terms/s", "cores": threads, "kind": "port",
                               "sample": f"1 shell x {len(qv)} vectors x {len(idx)} atoms x {T.NF} frames "
                                         f"({n:.2e} terms) in {dt:.1f} s, FP64 FFT of {L}, {threads} threads "
                                         f"on {cpu_model()}", "host": host_info()}
    print(json.dumps(out), flush=True)


def bench_mp(args):
    """Multipole averaging on the headline trajectory (20,000 atoms x 2,000
    frames), 10 shells over 0.2-2.0 A^-1, FFT autocorrelation.  Sphere
    (MPSphereScatterDevice, workload mp): moments (0,0) and l = 1..10,
    m = -l..l (121, the reference's list for resolution 10), q along z.
    Cylinder (MPCylinderScatterDevice, workload mpc): (0,0) and l = 1..10,
    m = 0..3 (41), axis z, q at 30 degrees from the axis.  One term = one
    (atom, frame, moment, shell)."""
    from sassena_amd import ScatterDevice, synth
    T = synth.case("c2")
    lmax = 10
    kind = "cylinder" if args.workload == "mpc" else "sphere"
    if kind == "sphere":
        moments = [(0, 0)] + [(l, m) for l in range(1, lmax + 1) for m in range(-l, l + 1)]
        qv = np.array([[0.0, 0.0, q] for q in np.linspace(0.2, 2.0, 10)])
    else:
        moments = [(0, 0)] + [(l, m) for l in range(1, lmax + 1) for m in range(4)]
        qv = np.array([[0.5 * q, 0.0, np.sqrt(0.75) * q] for q in np.linspace(0.2, 2.0, 10)])
    dev = ScatterDevice("all", "autocorrelate", "fftw", profile=True)
    dev.stage(T.xyz, layout="soa")
    dev.set_factors(T.b)
    step = lambda: dev.compute_multipole(kind, qv, moments)
    for _ in range(args.warmup):
        step()
    amp_ms = dev_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        tm = dev.timing()
        amp_ms += tm["amplitude_ms"]
        dev_ms += tm["total_ms"]
    wall = time.perf_counter() - t0
    dev.close()
    terms = T.NA * T.NF * len(moments) * len(qv)
    out = {
        "metric": f"atom*frame*moment*shell terms/s (multipole {kind} average, FFT autocorrelation)",
        "value": terms * args.steps / wall, "unit": "terms/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 special functions and sums / f32 amplitudes",
        "data": "synthetic",
        "config": {"workload": f"multipole {kind}, {T.NA} atoms x {T.NF} frames, l <= {lmax} "
                               f"({len(moments)} moments) x {len(qv)} shells",
                   "NA": T.NA, "NF": T.NF, "moments": len(moments), "shells": len(qv), "terms_per_step": terms},
        "kernel_ms_per_step": {"multipole": amp_ms / args.steps, "device_total": dev_ms / args.steps},
    }
    per = MP_FP64_PER_TERM[kind]
    if per is not None:
        amp_s = amp_ms / args.steps * 1e-3
        out["roofline"] = {"bound": "valu", "achieved": per * terms / amp_s / 1e12, "peak": FP64_PEAK / 1e12,
                           "unit": "TFLOP/s", "frac": per * terms / amp_s / FP64_PEAK, "traffic": None,
                           "kernel": "multipole_kernel", "fp64_flops_per_term": per,
                           "note": "FP64 flops per (atom, frame, moment, shell) from rocprofv3 counts of "
                                   "FP64 FMA (2 flops), MUL and ADD instructions over the multipole kernel, "
                                   "against the published FP64 vector peak"}
    if not args.no_cpu:
        # the FP64 oracle on the host: one shell x one moment per thread
        # (the highest degree of the list) over all atoms and frames
        import oracle
        from concurrent.futures import ThreadPoolExecutor
        threads = host_threads()
        frames = synth.to_frames_aos(T.xyz)
        sample = moments[-threads:]
        t1 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda mo: oracle.compute_multipole(kind, frames, T.b, qv[-1], [mo]), sample))
        dt = time.perf_counter() - t1
        n = len(sample) * T.NA * T.NF
        out["cpu_baseline"] = {"value": n / dt, "unit": "terms/s", "cores": threads, "kind": "port",
                               "sample": f"1 shell x {len(sample)} moments (l = {lmax}) x {T.NA} atoms x {T.NF} "
                                         f"frames ({n:.2e} terms) in {dt:.1f} s, FP64, one moment per thread, "
                                         f"{threads} threads on {cpu_model()}", "host": host_info()}
    print(json.dumps(out), flush=True)


def bench_c4(args):
    """64 trajectories (atoms log-uniform 500-50,000, frames 256-4,096, box
    (NA/0.1)^(1/3) A, sigma 0.1 walk, 10-60 shells over 0.1-4.0 A^-1 with
    50-300 orientations): one captured graph per step; per-trajectory
    latency from each trajectory run alone through the same kernels."""
    from sassena_amd import Batch, synth
    t0 = time.perf_counter()
    trajs = synth.c4_batch(scale=args.c4_scale)
    gen_s = time.perf_counter() - t0
    log(f"c4: generated {len(trajs)} trajectories in {gen_s:.1f} s")
    bt = Batch()
    bt.stage([(T.xyz, T.b, T.qlen, T.orient) for T in trajs])
    log("c4: staged")
    for _ in range(args.warmup):
        bt.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bt.run()
    wall = time.perf_counter() - t0
    amp_ms = bt.time_amplitude()
    tm = bt.timing()
    log(f"c4: {args.steps} graph steps in {wall:.2f} s")
    lat = []
    for k in range(len(trajs)):
        bt.run_one(k)                       # builds and warms this trajectory's graph
        lat.append(bt.run_one(k)[1])
    bt.close()
    terms = tm["terms"]
    amp_s = amp_ms * 1e-3
    legs = roofline_legs(terms, amp_s)
    ex_flops = (EXEC_FLOPS_STEPPED * (tm["terms_stepped"] + tm["terms_exact_runs"])
                + EXEC_FLOPS_RESIDUAL * (tm["terms_residual"] - tm["terms_exact_runs"])
                + ALG_FP32_FLOPS * (terms - tm["terms_stepped"] - tm["terms_residual"]))
    out = {
        "metric": "atom*frame*q-vector terms/s (64 uneven coherent trajectories, one graph launch)",
        "value": terms * args.steps / wall,
        "unit": "terms/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 sincos / f64 phase", "data": "synthetic",
        "config": {"workload": "64 uneven coherent trajectories (batched graph, persistent amplitude kernel)",
                   "trajectories": len(trajs), "terms_per_step": terms, "tiles": tm["tiles"],
                   "tile_shape": tm["tile_shape"], "scale": args.c4_scale, "generation_s": gen_s,
                   "terms_stepped_exact": tm["terms_stepped"], "terms_stepped_residual": tm["terms_residual"],
                   "terms_residual_in_exact_runs": tm["terms_exact_runs"]},
        "graph_ms": tm["total_ms"],
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "min": float(np.min(lat)), "max": float(np.max(lat))},
        "roofline": {"bound": "valu", "achieved": legs["fp64"]["achieved"], "peak": legs["fp64"]["peak"],
                     "unit": "TFLOP/s", "frac": legs["fp64"]["frac"], "traffic": None,
                     "kernel": "amp_persistent_kernel", "legs": legs,
                     "executed": {"flops_per_term": ex_flops / terms, "achieved": ex_flops / amp_s / 1e12,
                                  "unit": "TFLOP/s", "frac_of_unpacked_fp32": ex_flops / amp_s / (FP32_PEAK / 2),
                                  "note": "8 FP32 flops per term of exactly stepped shells, 12 with residual "
                                          "corrections, 4 (plus sin/cos) on the direct path"}},
    }
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """100,000 atoms in a 100 A box, sigma 0.1 walk over 10,000 frames, 60
    shells over 0.05-5.0 A^-1, 1,000 orientations: the network is trained on
    frames [0, 2000) (1,000- and 16-orientation estimates there), validated
    out of sample per shell, and corrects the 16-orientation estimate on
    frames [2000, 10000) where validation meets the tolerance; other shells
    fall back to the plain 1,000-orientation average.  Judged against the
    plain 1,000-orientation average on those frames."""
    from sassena_amd import cv, synth
    rng = np.random.default_rng(505)
    NA = max(1000, int(100000 * args.c5_scale))
    NF = max(500, int(10000 * args.c5_scale))
    box = (NA / 0.1) ** (1.0 / 3.0)
    t0 = time.perf_counter()
    b = synth.scattering_lengths(rng, NA)
    xyz = synth.random_walk(rng, NA, NF, box, 0.1)
    orient = synth.orientations(rng, 1000)
    qlen = synth.linear_shells(0.05, 5.0, 60)
    gen_s = time.perf_counter() - t0
    log(f"c5: generated {NA} atoms x {NF} frames in {gen_s:.1f} s")
    r = cv.run(xyz, b, qlen, orient, ntrain_frames=NF // 5, tol=args.c5_tol, steps=3000, batch=4096,
               progress=lambda stage, sec: log(f"c5: {stage}: {sec:.2f} s"))
    tm = r["timing_s"]
    cv_eval = tm["small_eval"] + tm["predict"] + tm["fallback"]
    cv_total = cv_eval + tm["many_train"] + tm["small_train"] + tm["train"]
    out = {
        "metric": "wall-time ratio of the plain 1000-orientation average to the learned control variate "
                  "(with fallback) on the evaluated frames, at the stated shell error",
        "value": tm["many_eval"] / cv_eval, "unit": "x less wall time on the evaluated frames",
        "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
        "config": {"workload": "learned control variate, 100000 atoms x 10000 frames, 60 shells, "
                               "1000 vs 16 orientations", "NA": NA, "NF": NF, "shells": 60, "orientations": 1000,
                   "small_orientations": 16, "train_frames": NF // 5, "tol": args.c5_tol, "generation_s": gen_s},
        "cv_shells": r["cv_shells"], "timing_s": tm, "speedup_including_training": tm["many_eval"] / cv_total,
        "error_small": r["error_small"], "error_net_all_shells": r["error_net"], "error_cv": r["error_cv"],
        "bias_net": r["bias_net"], "train_loss": r["train_loss"],
        "speedup_net_all_shells": tm["many_eval"] / (tm["small_eval"] + tm["predict"]),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--vector-block", type=int, default=0)
    ap.add_argument("--atom-split", type=int, default=0)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5", "mp", "mpc"],
                    help="c2: the headline coherent case; c3: incoherent self with per-pair FFTs; "
                         "c4: 64 uneven trajectories in one batched graph; "
                         "c5: learned control variate against the 1,000-orientation average; "
                         "mp: multipole sphere average (121 moments x 10 shells) on the c2 trajectory; "
                         "mpc: multipole cylinder average (41 moments x 10 shells) on the same trajectory")
    ap.add_argument("--c5-scale", type=float, default=1.0, help="shrink the c5 atoms and frames (tests)")
    ap.add_argument("--c4-scale", type=float, default=1.0, help="shrink the c4 trajectories (tests)")
    ap.add_argument("--c5-tol", type=float, default=1e-3,
                    help="c5: shell error (mean over lags, relative to F(q,0)) a shell must meet in validation "
                         "to use the control variate")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--generic", action="store_true",
                    help="pass the 8000 expanded q vectors (sx_compute_shells) instead of |q| x orientations")
    args = ap.parse_args()
    if args.calibrate_copy:
        # a plain copy kernel of known size: the FETCH_SIZE / WRITE_SIZE
        # counters of the hot-path kernels are scaled by what they read here
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: the hot path shards by q shell with "
                                   "no exchange, see the design notes; run with --gpus 1"}))
        sys.exit(2)

    from sassena_amd import ScatterDevice, synth

    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload in ("mp", "mpc"):
        return bench_mp(args)

    T = synth.case("c2")
    qv = np.array([T.qvectors(s) for s in range(T.nshells)])
    dev = ScatterDevice("all", "autocorrelate", "fftw", profile=True, vector_block=args.vector_block,
                        atom_split=args.atom_split)
    t_stage = time.perf_counter()
    dev.stage(T.xyz, layout="soa")   # host (pageable) -> HBM, SoA; not part of value
    stage_s = time.perf_counter() - t_stage
    dev.set_factors(T.b)
    if args.generic:
        step = lambda: dev.compute_shells(qv)
    else:
        step = lambda: dev.compute_oriented(T.qlen, T.orient)
    for _ in range(args.warmup):
        step()
    amp_ms = corr_ms = fin_ms = dev_ms = 0.0
    amp_launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()   # returns after the device stream has drained
        tm = dev.timing()
        amp_ms += tm["amplitude_ms"]
        corr_ms += tm["correlate_ms"]
        fin_ms += tm["finalize_ms"]
        dev_ms += tm["total_ms"]
        amp_launches += tm["amplitude_launches"]
    wall = time.perf_counter() - t0
    terms = T.terms() * args.steps
    value = terms / wall
    amp_avg_s = amp_ms / amp_launches * 1e-3
    terms_per_launch = T.terms() * args.steps / amp_launches
    stepped = bool(tm.get("shell_progression", 0))
    residual = bool(tm.get("shell_residuals", 0))
    legs = roofline_legs(terms_per_launch, amp_avg_s)
    # the binding algorithmic leg: FP64 (7 flops per term against 78.6 TFLOP/s)
    roof = {"bound": "valu", "achieved": legs["fp64"]["achieved"], "peak": legs["fp64"]["peak"],
            "unit": "TFLOP/s", "frac": legs["fp64"]["frac"],
            "unit_of_work": "one (atom, frame, q vector) term: 7 FP64 flops (phase and its reduction), 2 "
                            "transcendentals, 4 FP32 flops; the FP64 leg binds",
            "legs": legs}
    if stepped:
        exact_shells = tm.get("shell_exact_runs", 0) if residual else T.nshells
        ex = (EXEC_FLOPS_STEPPED * exact_shells + EXEC_FLOPS_RESIDUAL * (T.nshells - exact_shells)) / T.nshells
        roof["executed"] = {
            "algorithm": "stepped phasors across linearly spaced shells" +
                         (" with first-order corrections of the scan's single-precision residuals, except "
                          "on the shells that lie on their tile's line" if residual else ""),
            "shells_exact_runs": exact_shells if residual else None,
            "flops_per_term": ex, "achieved": ex * terms_per_launch / amp_avg_s / 1e12, "unit": "TFLOP/s",
            "frac_of_unpacked_fp32": ex * terms_per_launch / amp_avg_s / (FP32_PEAK / 2),
            "tile_shape": tm.get("tile_shape"),
            "note": "FP32 FMAs executed per term; the FP64 phase and 4 hardware sin/cos per (orientation, "
                    "atom, frame) are shared by the tile's shells and not counted; packed FP32 is not used, so "
                    "the reachable rate is half the published 157.3 TFLOP/s"}
    out = {
        "metric": "atom*frame*q-vector terms/s (coherent F(q,t), FFT autocorrelation)",
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sincos / f64 phase",
        "data": "synthetic",
        "config": {"workload": "coherent all, 20000 atoms x 2000 frames, 40 shells x 200 vectors, pad 4096",
                   "entry": "sx_compute_shells" if args.generic else "sx_compute_shells_oriented",
                   "shells": "q scan 0.1-4.0 A^-1, 40 points, interior points rounded by single-precision "
                             "powf as the reference's scan builder does",
                   "NA": T.NA, "NF": T.NF, "nshells": T.nshells, "vectors": T.NM, "terms_per_step": T.terms(),
                   "stepped": stepped, "residual_corrected": residual},
        "kernel_ms_per_step": {"amplitude": amp_ms / args.steps, "correlate": corr_ms / args.steps,
                               "finalize": fin_ms / args.steps, "device_total": dev_ms / args.steps},
        "staging_s": stage_s,
        "staging_GBps": 12.0 * T.NA * T.NF / stage_s / 1e9,
        "pcie_inclusive_value": T.terms() / (wall / args.steps + stage_s),
        "roofline": dict(roof, traffic=AMP_TRAFFIC_BYTES if not (args.generic or args.vector_block or
                                                                 args.atom_split) else None,
                         traffic_unit="bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
                         kernel="amp_all_kernel" if args.generic else
                         ("amp_oriented_res_kernel" if residual else "amp_oriented_kernel")),
    }
    if not args.no_cpu:
        threads = host_threads()
        rec, ref, k = cpu_baseline(T, threads)
        out["cpu_baseline"] = rec
        # the GPU on the same vectors of the same shell (all 40 shells with the
        # baseline's orientations, so the same stepped path): error against the FP64 oracle
        gfqt, gfq, _ = dev.compute_oriented(T.qlen, T.orient[:k])
        err = float(np.max(np.abs(gfqt[-1] - ref[0])) / abs(ref[0][0]))
        out["error_vs_fp64"] = {"shell_error": err, "fq_error": float(abs(gfq[-1] - ref[1]) / abs(ref[0][0])),
                                "definition": "max_tau |F - F_oracle| / |F_oracle(0)| over the cpu_baseline's "
                                              f"{k} vectors of the last shell",
                                "stepped": bool(dev.timing().get("shell_progression", 0))}
    dev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
