"""Benchmark: non-negative CPD-ALS on one MI355X, chosen form vs the plain FP64 path.

Workload (the 800^3 configuration the metric is quoted on): dense 3-way
800 x 800 x 800 tensor, true rank 32 with Uniform[0,1) factors (seed 0) plus
Gaussian noise at SNR 20 dB, decomposed at R = 32 from Uniform[0,1) initial
factors (seed 1) with the reference's default options (nonnegativity,
acceleration, normalization, Nesterov tolerances 1e-2), threshold 0 so every
run does exactly W + K sweeps (30 by default, the configuration's count).

A step is one ALS sweep.  The timed region is sweeps W+1 .. W+K of one
decomposition (per-sweep device-synchronised timestamps from the library);
inputs are resident in HBM before it starts.  `value` is the chosen form's
sweeps/s, the median of three runs of the timed region.  Also reported: the
plain FP64 GPU paths (per-mode MTTKRP as cpd, and the dimension tree as
cpdDimTree; FP64 tensor and arithmetic) on the same tensor and initial factors,
the final relative errors measured against the FP64 tensor, the chosen factors'
relative difference from the plain ones, the dominant contraction's HBM
roofline (HIP events on the library stream), and the FP64 CPU oracle (OpenMP
restatement of cpd_omp) timed on a bounded sample on the host cores, whose
four-sweep factors the device's four-sweep runs are also compared with.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "partensor-toolbox_amd"))

import numpy as np  # noqa: E402

import ptn_hip as P  # noqa: E402

METRIC = "ALS sweeps/s + MTTKRP GB/s (800³,R=32): chosen vs plain GPU @ matched rel. error"
DIMS = [800, 800, 800]
R = 32
HBM_PEAK_GBS = 8000.0
REPEATS = 3          # the timed region is run this many times; value is the median


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed_sweeps(status, sweep_log, W, K):
    t_end = [rec["t_end"] for rec in sweep_log]
    t0 = t_end[W - 1] if W > 0 else 0.0
    return t_end[W + K - 1] - t0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_tensor_passes_800cubed_R32.md")


def measured_traffic(key="left_pair_traffic_bytes"):
    """HBM bytes per launch of a left-node contraction from the committed counter
    passes (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's readings)."""
    try:
        vals = {}
        for line in open(TRAFFIC_FILE):
            k, sep, v = line.partition(":")
            if sep and k.strip().endswith("_traffic_bytes"):
                vals[k.strip()] = float(v)
        return vals.get(key)
    except OSError:
        return None


def pass_times(sweep_log, W, K):
    """Average device time (ms) of the two tensor passes over the timed sweeps, from
    the HIP events the library records around them (time_passes)."""
    out = []
    for q in range(2):
        v = [rec["t_pass_ms"][q] for rec in sweep_log[W:W + K] if rec["t_pass_ms"][q] > 0]
        out.append((sum(v) / len(v), len(v)) if v else (None, 0))
    return out


def cpu_info():
    """CPU model, logical CPUs this process may run on, and the physical cores among them."""
    model, phys = "unknown", {}
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            k, sep, v = line.partition(":")
            if not sep:
                if "processor" in cur:
                    phys[int(cur["processor"])] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
                cur = {}
                continue
            cur[k.strip()] = v.strip()
            if k.strip() == "model name":
                model = v.strip()
        if "processor" in cur:
            phys[int(cur["processor"])] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
    except OSError:
        pass
    avail = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = len({phys[c] for c in avail if c in phys}) or len(avail)
    return model, len(avail), cores


def cpu_baseline(device_runs):
    """The FP64 OpenMP restatement of cpd_omp (CpdOpenMP.hpp:266-329: blocked MTTKRP
    MTTKRP.hpp:102-225, explicit cost CpdOpenMP.hpp:121-125) on the host cores, bounded:
    the 800^3 workload for 1 warm-up + 3 timed sweeps, three times (median), the same with
    the Gram-form cost once, and the 100^3 R=10 configuration for 1 + 49 sweeps.  The
    four-sweep runs also give the oracle's factors that the device's four-sweep runs
    (device_runs: name -> factors) are compared with."""
    os.environ.setdefault("OMP_PROC_BIND", "close")      # read when the OpenMP runtime loads
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    model, logical, physical = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or physical
    t = time.time()
    X, _ = o.make_tensor(DIMS, R, seed=0, snr_db=20.0)
    init = o.gen_factors(DIMS, R, seed=1)
    frob = float(np.sum(X * X))
    log("cpu baseline: %s, %d logical CPUs / %d physical cores available, %d threads, tensor in %.1f s"
        % (model, logical, physical, threads, time.time() - t))

    def sweeps_per_s(cost_form):
        F, st, lg = o.cpd(X, R, init, max_iter=4, threshold_error=0.0, mttkrp_form=o.MTTKRP_BLOCKED,
                          cost_form=cost_form, threads=threads)
        return 3.0 / (lg[3]["t_end"] - lg[0]["t_end"]), F, st
    reps = []
    Fo = sto = None
    for _ in range(3):
        v, Fo, sto = sweeps_per_s(o.COST_EXPLICIT)
        reps.append(v)
    gram_v, _, _ = sweeps_per_s(o.COST_GRAM)
    vs_oracle = {}
    for name, F in device_runs.items():
        vs_oracle[name] = {
            "factor_rel_err": max(P.rel_err(a, b) for a, b in zip(F, Fo)),
            "rel_error_gap": abs(o.cost_explicit(X, F) - o.cost_explicit(X, Fo)) / math.sqrt(frob)}
    del X
    X1, _ = o.make_tensor([100, 100, 100], 10, seed=0, snr_db=30.0)
    init1 = o.gen_factors([100, 100, 100], 10, seed=1)
    _, _, lg1 = o.cpd(X1, 10, init1, max_iter=50, threshold_error=0.0, mttkrp_form=o.MTTKRP_BLOCKED,
                      cost_form=o.COST_EXPLICIT, threads=threads)
    c1 = 49.0 / (lg1[49]["t_end"] - lg1[0]["t_end"])
    med = sorted(reps)[1]
    return {"value": med, "unit": "sweeps/s", "cores": threads, "kind": "port",
            "sample": "800^3 R=32 FP64 workload, oracle cpd_omp restatement (blocked MTTKRP + explicit cost): "
                      "1 warm-up + 3 timed sweeps, median of 3 runs (%s sweeps/s); Gram-form cost variant %.4f "
                      "sweeps/s; 100^3 R=10 (50 sweeps, 49 timed) %.2f sweeps/s"
                      % (", ".join("%.4f" % v for v in reps), gram_v, c1),
            "cpu_model": model, "logical_cpus": logical, "physical_cores": physical,
            "omp": "OMP_NUM_THREADS=%d OMP_PROC_BIND=%s OMP_PLACES=%s"
                   % (threads, os.environ["OMP_PROC_BIND"], os.environ["OMP_PLACES"]),
            "runs": reps, "gram_cost_variant": gram_v, "c1_100cubed_R10": c1,
            "device_vs_oracle_after_4_sweeps": vs_oracle}


def copy_calibration(nbytes):
    """Device-to-device hipMemcpy of nbytes, three times: the plain copy kernel whose
    FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic of the contraction."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:    # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: 3 device copies of %d bytes" % nbytes)


def bench_gemm_node():
    """The order-4 dimension tree's first contraction, X_(0,1) (A_3 (.) A_2), on a dense
    200^4 FP32 tensor (true rank 64, Exponential(1) factor columns from seed 2, whose
    expected pairwise cosine is 1/2, SNR 20 dB), R = 64: FP32 MFMA vs bf16x3 split
    products (FP32 accumulate), with the FP64 MFMA node on the FP64 copy as the
    reference; errors are those of the tree's four MTTKRPs against FP64."""
    dims, Rn = [200, 200, 200, 200], 64
    h = P.Handle(0)
    t = time.time()
    h.generate_tensor(dims, Rn, seed=2, snr_db=20.0, dist=1, storage=P.TENSOR_BOTH)
    A = h.generate_factors(dims, Rn, seed=1)
    log("200^4 tensor generated in %.2f s" % (time.time() - t))
    M64 = h.mttkrp_dimtree(A, arith=P.ARITH_F64)
    flops = 2.0 * Rn * float(np.prod(dims))
    res = {}
    for name, ar in (("fp32_mfma", P.ARITH_F32), ("bf16x3", P.ARITH_BF16X3), ("fp64_mfma", P.ARITH_F64)):
        ms, b = h.time_contraction(0, A, arith=ar, iters=10)
        entry = {"ms": ms, "gbs": b / (ms * 1e-3) / 1e9, "tflops": flops / (ms * 1e-3) / 1e12,
                 "algorithmic_bytes": b}
        if ar != P.ARITH_F64:
            M = h.mttkrp_dimtree(A, arith=ar)
            entry["mttkrp_rel_err_vs_fp64"] = [P.rel_err(M[k], M64[k]) for k in range(4)]
        res[name] = entry
        log("%s: %.3f ms, %.0f GB/s, %.1f TFLOP/s" % (name, ms, entry["gbs"], entry["tflops"]))
    h.close()
    best = min(("fp32_mfma", "bf16x3"), key=lambda k: res[k]["ms"])
    out = {
        "metric": "order-4 GEMM node X_(0,1)(A3 (.) A2) time, 200^4 FP32, R=64: FP32 MFMA vs bf16x3",
        "value": res[best]["ms"], "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic (device Philox4x32-10: Exponential(1) factors seed 2, noise at SNR 20 dB; "
                "evaluation factors U[0,1) seed 1)",
        "config": {"workload": "dense 4-way 200x200x200x200, true rank 64, R=64, fused GEMM+TTV node "
                               "(no intermediate), split over k and reduced", "fastest": best, **res},
        "roofline": {"fp32_mfma": {"bound": "mfma", "peak": 157.3, "unit": "TFLOP/s",
                                   "achieved": res["fp32_mfma"]["tflops"],
                                   "frac": res["fp32_mfma"]["tflops"] / 157.3},
                     "bf16x3": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "achieved": res["bf16x3"]["gbs"], "frac": res["bf16x3"]["gbs"] / HBM_PEAK_GBS}},
    }
    print(json.dumps(out))


BATCH_P, BATCH_SWEEPS = 512, 100
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def batch_problems():
    """512 problems: mode sizes i.i.d. uniform integers in [16, 128] and ranks uniform in
    [4, 16], drawn with numpy's default_rng(3)."""
    rng = np.random.default_rng(3)
    dims = rng.integers(16, 129, (BATCH_P, 3))
    ranks = rng.integers(4, 17, BATCH_P)
    return [tuple(int(v) for v in d) for d in dims], [int(r) for r in ranks]


def batch_cpu_baseline(dims, ranks, sweeps=10, every=1):
    """The CPU oracle's cpd() (standard MTTKRP, OpenMP over output rows, Gram-form cost)
    on every `every`-th problem of the batch, one problem after another, on the FP32-
    rounded tensors the device holds; scaled to whole-batch sweeps per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    secs, n = 0.0, 0
    for p in range(0, len(dims), every):
        X, _ = o.make_tensor(list(dims[p]), ranks[p], seed=BATCH_SEED_X + p, snr_db=25.0)
        X = np.asfortranarray(X.astype(np.float32).astype(np.float64))
        init = o.gen_factors(list(dims[p]), ranks[p], seed=BATCH_SEED_INIT + p)
        _, _, lg = o.cpd(X, ranks[p], init, max_iter=sweeps, threshold_error=0.0, cost_form=o.COST_GRAM)
        secs += lg[-1]["t_end"] - lg[0]["t_start"]
        n += 1
    per_batch_sweep = secs / sweeps * (len(dims) / n)
    return {"value": 1.0 / per_batch_sweep, "unit": "batch sweeps/s", "cores": threads, "kind": "port",
            "sample": "%d sweeps of every %dth problem (%d problems), oracle cpd() one problem at a time "
                      "(OpenMP over MTTKRP rows), %.1f s, scaled by %d/%d to the whole batch"
                      % (sweeps, every, n, secs, len(dims), n)}


BATCH_SEED_X, BATCH_SEED_INIT = 3000, 1000


def bench_batch(args):
    """The batch: 512 independent non-negative order-3 CPDs of uneven sizes and ranks,
    all advanced by one persistent kernel per sweep (a captured graph replayed 100
    times), launch shape chosen by the library's autotuner.  The chosen form runs the
    MTTKRPs on the FP32 matrix cores (FP32 tensors, FP64 totals; everything else FP64);
    the same run on the FP64 matrix cores is reported beside it with the largest gap in
    the final relative costs."""
    dims, ranks = batch_problems()
    sweeps = args.batch_sweeps
    h = P.Handle(0)
    t = time.time()
    h.batch_generate(dims, ranks, seed=BATCH_SEED_X, snr_db=25.0)
    init = h.batch_init_factors(seed=BATCH_SEED_INIT)
    log("batch generated in %.2f s" % (time.time() - t))
    F, st, _, rep = h.cpd_batch(init, max_iter=sweeps, threshold_error=0.0, block=args.batch_block,
                                grid=args.batch_grid, with_log=False, arith=P.ARITH_F32)
    F64, st64, _, rep64 = h.cpd_batch(init, max_iter=sweeps, threshold_error=0.0, block=args.batch_block,
                                      grid=args.batch_grid, with_log=False, arith=P.ARITH_F64)
    h.close()
    secs = rep["seconds"]
    gap = [abs(a["rel_cost"] - b["rel_cost"]) for a, b in zip(st, st64)]
    log("fp64 matrix cores: %.3f ms/sweep; final rel cost gap fp32 vs fp64: max %.3e median %.3e"
        % (1e3 * rep64["seconds"] / sweeps, max(gap), float(np.median(gap))))
    nel = [d[0] * d[1] * d[2] for d in dims]
    # algorithmic traffic of one sweep: the FP32 tensor read by the three MTTKRPs and, from
    # the second sweep on, by the line search's MTTKRP; plus every factor row read and written
    acc_sweeps = max(sweeps - 2, 0)            # line searches run in sweeps 2 .. max_iter - 1
    tensor_bytes = 4.0 * sum(nel) * (3 * sweeps + acc_sweeps)
    rows = sum(sum(d) for d in dims)
    factor_bytes = 8.0 * 16 * rows * (6 * sweeps)
    flops = sum(2.0 * ((r + 3) // 4 * 4) * n for r, n in zip(ranks, nel)) * (3 * sweeps + acc_sweeps)
    gbs = (tensor_bytes + factor_bytes) / secs / 1e9
    tfl = flops / secs / 1e12
    rel = [s["rel_cost"] for s in st]
    log("batch: %d sweeps in %.3f s (block %d grid %d, tuned %d in %.2f s): %.0f GB/s, %.2f TFLOP/s"
        % (sweeps, secs, rep["block"], rep["grid"], rep["tuned"], rep["tune_seconds"], gbs, tfl))
    cpu = None if args.no_cpu else batch_cpu_baseline(dims, ranks)
    if cpu:
        log("cpu baseline %.4f batch sweeps/s" % cpu["value"])
    value = sweeps / secs
    frac_h, frac_f = gbs / HBM_PEAK_GBS, tfl / FP32_PEAK_TFLOPS     # the MTTKRPs run on the FP32 matrix cores
    out = {
        "metric": "batched CPD-ALS sweeps/s (512 uneven order-3 problems, I~U{16..128}, R~U{4..16})",
        "value": value, "unit": "batch sweeps/s", "n_gpus": 1, "steps": sweeps, "warmup": 0,
        "ms_per_step": 1e3 * secs / sweeps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (device Philox4x32-10: problem p factors U[0,1) seed %d+p, noise "
                                "SNR 25 dB, init U[0,1) seed %d+p; sizes numpy default_rng(3))"
                                % (BATCH_SEED_X, BATCH_SEED_INIT),
        "config": {"workload": "512 independent non-negative CPDs, FP32 tensors, %d sweeps, one persistent "
                               "kernel per sweep replayed from a captured graph" % sweeps,
                   "chosen_form": "MTTKRPs on the FP32 matrix cores with FP64 totals; Grams, eigenvalues, "
                                  "Nesterov, cost, Normalize and line search in FP64",
                   "fp64_matrix_cores": {"batch_sweeps_per_s": sweeps / rep64["seconds"],
                                         "ms_per_sweep": 1e3 * rep64["seconds"] / sweeps},
                   "final_rel_cost_gap_vs_fp64": {"max": max(gap), "median": float(np.median(gap))},
                   "problem_sweeps_per_s": value * len(dims), "block": rep["block"], "grid": rep["grid"],
                   "autotuned": bool(rep["tuned"]), "tune_seconds": rep["tune_seconds"],
                   "total_elements": sum(nel), "median_final_rel_cost": float(np.median(rel)),
                   "max_final_rel_cost": float(np.max(rel))},
        "roofline": {"bound": "hbm" if frac_h >= frac_f else "mfma", "achieved": gbs if frac_h >= frac_f else tfl,
                     "peak": HBM_PEAK_GBS if frac_h >= frac_f else FP32_PEAK_TFLOPS,
                     "unit": "GB/s" if frac_h >= frac_f else "TFLOP/s", "frac": max(frac_h, frac_f),
                     "traffic": None, "kernel": "k_batch_sweep (whole sweep)",
                     "hbm_gbs": gbs, "mfma_tflops_fp32": tfl},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _time_to(log, target):
    """Seconds from the start of the run to the end of the first sweep whose relative
    cost is <= target (None if never)."""
    for rec in log:
        if rec["rel_cost"] <= target:
            return rec["t_end"]
    return None


def bench_warm(args):
    """The learned NNLS warm start.  The network is trained on (subproblem, Nesterov
    solution) pairs harvested on the device at the evaluation's rank: two 400^3 R=64 runs
    (true rank 64, U[0,1) factors, SNR 15 dB, 30 sweeps of the chosen form), then used on
    the 1000^3 FP32 tensor (true rank 64, U[0,1) factors seed 4, SNR 15 dB) at R = 64.
    Reported for the plain start, the network's start under the 1e-6 KKT gate, the same
    after 10 refinement steps from every proposal (counted as NNLS steps), and the refined
    proposals under a loose 5e-2 gate: NNLS steps, rows that took the network's start, and
    the time each run needs to reach the plain FP64 GPU path's final relative error (+1e-6)."""
    h = P.Handle(0)
    sweeps = args.warm_sweeps
    t = time.time()
    h.mlp_data_clear()
    for seed in (6, 8):
        h.generate_tensor([400] * 3, 64, seed=seed, snr_db=15.0, storage=P.TENSOR_F32)
        h.cpd(64, h.generate_factors([400] * 3, 64, seed=seed + 1), max_iter=30, threshold_error=0.0,
              mttkrp_form=P.MTTKRP_DIMTREE, arith=P.ARITH_BF16X3, fuse_accel=1, mlp_harvest=1)
    rows = h.mlp_data_size()
    log("harvested %d pairs in %.1f s" % (rows, time.time() - t))
    t = time.time()
    h.mlp_init(4)
    losses = h.mlp_train(args.warm_epochs, batch=1024, lr=1e-3, seed=4)
    train_s = time.time() - t
    log("trained %d epochs in %.2f s: loss %.3e -> %.3e" % (args.warm_epochs, train_s, losses[0], losses[-1]))
    dims5, R5 = [1000] * 3, 64
    t = time.time()
    h.generate_tensor(dims5, R5, seed=4, snr_db=15.0, storage=P.TENSOR_BOTH)
    init = h.generate_factors(dims5, R5, seed=5)
    frob = h.norm_sq()
    log("1000^3 tensor generated in %.1f s" % (time.time() - t))
    common = dict(max_iter=sweeps, threshold_error=0.0)
    Fr, sr, lr_ = h.cpd(R5, init, mttkrp_form=P.MTTKRP_STANDARD, arith=P.ARITH_F64, **common)
    e_ref = h.cost_explicit(Fr) / math.sqrt(frob)
    chosen = dict(mttkrp_form=P.MTTKRP_DIMTREE, arith=P.ARITH_BF16X3, fuse_accel=1, **common)
    Fp, sp, lp = h.cpd(R5, init, **chosen)
    Fw, sw, lw = h.cpd(R5, init, warm_start=1, **chosen)                       # gate: KKT residual <= 1e-6
    Fr, sr, lr = h.cpd(R5, init, warm_start=1, warm_refine=10, **chosen)       # 10 steps from each proposal first
    Fl, sl, ll = h.cpd(R5, init, warm_start=1, warm_refine=10, warm_kkt_tol=5e-2, **chosen)   # loose gate
    h.close()
    target = e_ref + 1e-6
    out = {}
    for name, F, st, lg in (("plain_start", Fp, sp, lp), ("warm_start", Fw, sw, lw),
                            ("warm_start_refine_10", Fr, sr, lr),
                            ("warm_start_refine_10_loose_gate_5e-2", Fl, sl, ll)):
        out[name] = {"nnls_steps": int(sum(sum(r["nnls_iters"]) for r in lg)),
                     "warm_rows": int(sum(sum(r["warm_rows"]) for r in lg)),
                     "refine_steps": int(sum(sum(r["warm_steps"]) for r in lg)),
                     "seconds": lg[-1]["t_end"], "sweeps_per_s": len(lg) / lg[-1]["t_end"],
                     "final_rel_cost": st["rel_cost"],
                     "seconds_to_fp64_final_error": _time_to(lg, target)}
    ref_t = lr_[-1]["t_end"]
    saved = out["plain_start"]["nnls_steps"] - out["warm_start"]["nnls_steps"]
    tp, tw = out["plain_start"]["seconds_to_fp64_final_error"], out["warm_start"]["seconds_to_fp64_final_error"]
    log("fp64 ref %.9f in %.2f s; plain start %s s, warm start %s s; NNLS steps saved %d"
        % (e_ref, ref_t, tp, tw, saved))
    res = {
        "metric": "learned NNLS warm start (1000^3, R=64): NNLS steps saved and time to the FP64 path's final error",
        "value": (tp / tw) if (tp and tw) else None, "unit": "x (time to target, plain start / warm start)",
        "n_gpus": 1, "steps": sweeps, "warmup": 0, "ms_per_step": 1e3 * out["warm_start"]["seconds"] / sweeps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (device Philox4x32-10): training pairs from two 400^3 true-rank-64 SNR 15 dB runs at "
                "R=64 (seeds 6, 8); evaluation tensor 1000^3 true rank 64 U[0,1) seed 4, SNR 15 dB, init seed 5",
        "config": {"workload": "dense 3-way 1000x1000x1000 FP32 (+FP64 copy), R=64, %d sweeps, chosen form "
                               "(dimension tree, bf16x3, fused line search) with and without the warm start"
                               % sweeps,
                   "training_pairs": rows, "epochs": args.warm_epochs, "train_seconds": train_s,
                   "train_loss_first_last": [float(losses[0]), float(losses[-1])],
                   "fp64_plain_final_rel_error": e_ref, "fp64_plain_seconds": ref_t, "target": target,
                   "nnls_steps_saved": saved, **out},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=28)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--arith", default="bf16x3", choices=["bf16x3", "f32", "f32_f64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--batch", action="store_true",
                    help="only run the batch of 512 uneven order-3 CPDs and exit")
    ap.add_argument("--batch-sweeps", type=int, default=BATCH_SWEEPS)
    ap.add_argument("--batch-block", type=int, default=0, help="sweep kernel block size (0: autotune)")
    ap.add_argument("--batch-grid", type=int, default=0, help="sweep kernel grid (0: autotune)")
    ap.add_argument("--warm", action="store_true",
                    help="only run the learned warm-start experiment (train, then 1000^3 R=64 with/without) and exit")
    ap.add_argument("--warm-sweeps", type=int, default=30)
    ap.add_argument("--warm-epochs", type=int, default=8)
    ap.add_argument("--gemm-node", action="store_true",
                    help="only time the order-4 GEMM node (200^4, R=64, FP32 MFMA vs bf16x3) and exit")
    args = ap.parse_args()
    if args.gemm_node:
        bench_gemm_node()
        return
    if args.batch:
        bench_batch(args)
        return
    if args.warm:
        bench_warm(args)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    W, K = args.warmup, args.steps
    if W < 1 or K < 1:
        print("need --warmup >= 1 and --steps >= 1", file=sys.stderr)
        sys.exit(2)
    arith = {"bf16x3": P.ARITH_BF16X3, "f32": P.ARITH_F32, "f32_f64": P.ARITH_F32_F64}[args.arith]

    h = P.Handle(0)
    t = time.time()
    h.generate_tensor(DIMS, R, seed=0, snr_db=20.0, storage=P.TENSOR_BOTH)
    init = h.generate_factors(DIMS, R, seed=1)
    frob64 = h.norm_sq()
    log("tensor generated on device in %.2f s" % (time.time() - t))
    total = W + K
    common = dict(max_iter=total, threshold_error=0.0)

    chosen_kw = dict(mttkrp_form=P.MTTKRP_DIMTREE, arith=arith, fuse_accel=1, **common)
    try:
        Fc, stc, lgc = h.cpd(R, init, **chosen_kw)
    except P.PtnError as e:
        log("chosen arith %s unavailable (%s); falling back to f32_f64 contractions" % (args.arith, e))
        arith = P.ARITH_F32_F64
        chosen_kw.update(arith=arith, fuse_accel=0)
        Fc, stc, lgc = h.cpd(R, init, **chosen_kw)
    # the timed region (sweeps W+1 .. W+K) three times; `value` is the median
    runs = [K / timed_sweeps(stc, lgc, W, K)]
    for _ in range(REPEATS - 1):
        _, st_r, lg_r = h.cpd(R, init, **chosen_kw)
        runs.append(K / timed_sweeps(st_r, lg_r, W, K))
    value = sorted(runs)[len(runs) // 2]
    err_c = h.cost_explicit(Fc) / math.sqrt(frob64)
    log("chosen: %s sweeps/s (median %.1f), rel err vs FP64 X %.9e"
        % (", ".join("%.1f" % v for v in runs), value, err_c))

    plain = plain_tree = None
    if not args.no_plain:
        def plain_run(form, label):
            Fp, stp, lgp = h.cpd(R, init, mttkrp_form=form, arith=P.ARITH_F64, **common)
            tp = timed_sweeps(stp, lgp, W, K)
            err_p = h.cost_explicit(Fp) / math.sqrt(frob64)
            same = [l["accel_accepted"] for l in lgp] == [l["accel_accepted"] for l in lgc]
            log("%s: %.1f sweeps/s, rel err %.9e" % (label, K / tp, err_p))
            return Fp, {"sweeps_per_s": K / tp, "ms_per_sweep": 1e3 * tp / K, "final_rel_error": err_p,
                        "chosen_factor_rel_err": max(P.rel_err(a, b) for a, b in zip(Fc, Fp)),
                        "same_line_search_path_as_chosen": same}
        Fp, plain = plain_run(P.MTTKRP_STANDARD, "plain per-mode")
        plain["form"] = "per-mode MTTKRP (cpd), FP64 tensor, FP64 matrix cores, Gram-form FP64 cost"
        _, plain_tree = plain_run(P.MTTKRP_DIMTREE, "plain tree")
        plain_tree["form"] = "dimension tree (cpdDimTree), FP64 tensor, FP64 matrix cores, Gram-form FP64 cost"

    # four-sweep runs for the comparison with the CPU oracle's four sweeps (CPU leg)
    short = dict(common, max_iter=4)
    dev4 = {"chosen": h.cpd(R, init, **dict(chosen_kw, max_iter=4))[0]}
    if not args.no_plain:
        dev4["plain_per_mode"] = h.cpd(R, init, mttkrp_form=P.MTTKRP_STANDARD, arith=P.ARITH_F64, **short)[0]

    # roofline of the dominant kernel, measured live: the same run again with HIP events
    # around the two tensor passes of every sweep (on the streams they are launched on).
    # In steady state the left-node pass is the paired one, X x_2 [A_2 | Ahat_2] (64
    # output columns), and the right node is X x_0 A_0 written out for the mode-2 leaf.
    Ft, stt, lgt = h.cpd(R, init, time_passes=1, **chosen_kw)
    (ms_p, n_p), (ms_rn, n_rn) = pass_times(lgt, W, K)
    I0, I1, I2 = DIMS
    sx = 4.0
    pair_cols = 2 * R if chosen_kw.get("fuse_accel") else R
    bytes_p = sx * I0 * I1 * I2 + 4.0 * pair_cols * I0 * I1 + 8.0 * pair_cols * I2
    bytes_rn = sx * I0 * I1 * I2 + 4.0 * R * I1 * I2 + 8.0 * R * I0
    gbs_p = bytes_p / (ms_p * 1e-3) / 1e9
    gbs_rn = bytes_rn / (ms_rn * 1e-3) / 1e9 if ms_rn else None
    log("live: left-node pass %.3f ms %.0f GB/s (%d sweeps); right-node pass %s" %
        (ms_p, gbs_p, n_p, "%.3f ms %.0f GB/s" % (ms_rn, gbs_rn) if ms_rn else "n/a"))
    # the two contractions alone (repeated launches of the kernel, operands packed once)
    ms_l, bytes_l = h.time_contraction(0, Fc, arith=arith, iters=20)
    ms_r, bytes_r = h.time_contraction(1, Fc, arith=arith, iters=20)
    gbs_l = bytes_l / (ms_l * 1e-3) / 1e9
    gbs_r = bytes_r / (ms_r * 1e-3) / 1e9
    log("alone: left node %.3f ms %.0f GB/s; right node (fused with its TTV) %.3f ms %.0f GB/s"
        % (ms_l, gbs_l, ms_r, gbs_r))
    h.close()

    cpu = None if args.no_cpu else cpu_baseline(dev4)
    if cpu:
        log("cpu baseline %.4f sweeps/s on %d threads; device vs oracle after 4 sweeps: %s"
            % (cpu["value"], cpu["cores"], cpu["device_vs_oracle_after_4_sweeps"]))

    tc = K / value
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * tc / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device Philox4x32-10 generator: factors U[0,1) seed 0, noise seed 0 at SNR 20 dB)",
        "config": {
            "workload": "dense 3-way 800x800x800 non-negative CPD-ALS, true rank 32, SNR 20 dB, R=32, "
                        "init U[0,1) seed 1, %d sweeps, reference default options" % total,
            "chosen_form": "dimension tree, FP32 tensor stream, contractions %s, FP64 Gram/eigen/NNLS/cost"
                           % {P.ARITH_BF16X3: "bf16x3 MFMA", P.ARITH_F32: "FP32 MFMA",
                              P.ARITH_F32_F64: "FP64 MFMA on FP32 data"}[arith],
            "chosen_final_rel_error": err_c,
            "chosen_runs_sweeps_per_s": runs,
            "plain_gpu": plain,
            "plain_gpu_tree": plain_tree,
            "speedup_vs_plain": (value / plain["sweeps_per_s"]) if plain else None,
            "speedup_vs_plain_tree": (value / plain_tree["sweeps_per_s"]) if plain_tree else None,
            "rel_error_gap": abs(err_c - plain["final_rel_error"]) if plain else None,
            "mttkrp_gbs": {"left_node_pass_in_sweep": gbs_p, "right_node_pass_in_sweep": gbs_rn,
                           "left_node_alone": gbs_l, "right_node_fused_alone": gbs_r},
        },
        "roofline": {"bound": "hbm", "achieved": gbs_p, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs_p / HBM_PEAK_GBS,
                     "traffic": measured_traffic("left_pair_traffic_bytes" if pair_cols == 2 * R
                                                 else "left_node_traffic_bytes"),
                     "traffic_source": "calibration committed in profiles/pmc_traffic_tensor_passes_800cubed_R32.md "
                                       "(rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel scaled by a plain copy's "
                                       "readings), not measured in this run",
                     "kernel": "k_gemm_cm_f32, the sweep's left-node pass X x_2 [A_2 | Ahat_2] (paired with the "
                               "line search)" if pair_cols == 2 * R else "k_gemm_cm_f32, left-node pass X x_2 A_2",
                     "avg_ms": ms_p, "launches_timed": n_p, "algorithmic_bytes": bytes_p,
                     "timing": "HIP events around the kernel on its stream in every timed sweep (time_passes run)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
