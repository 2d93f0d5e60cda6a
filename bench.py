"""Benchmark: seconds to the effective stiffness C_eff of a 256^3 two-phase
elastic image within 1e-4 of the FP64 result, on one MI355X.

Workload: overlapping spheres (radius 6-12 voxels, inclusion fraction 0.3,
seed 3) in a matrix; matrix E=1, nu=0.3, inclusion E=10, nu=0.2; trilinear
hexahedra, periodic boundaries, all 6 load cases.  A step is one full
homogenization (6 x [RHS on the device -> Jacobi-preconditioned Krylov solve ->
C column on the device]) with the model already resident in HBM.

Methods timed in the same run, all at --tol and zero initial guess unless said:
  plain   : FP64 Jacobi-PCG (the reference method)
  mixed   : FP32 operator and CG vectors, FP64 dots/alpha/beta, FP64 residual
            replacement (reliable updates)
  learned : mixed, seeded by the learned initial guess (ramp + network
            prediction of every load case's fluctuation field; the network
            is trained before the timed region on --train-samples seeded
            64^3 images of the same recipe, solved by the library)
            (reported as `value`)
Each is checked against the CPU FP64 oracle's columns for this image
(tests/golden/config_headline.json, matched by the image's SHA-256) and
against C_ref, FP64 PCG at tolerance 1e-9 on the same image (untimed); a
result over the metric's 1e-4 is marked invalid and exits non-zero.  The
cold end-to-end time (hmgInit to C, autotuner included) is reported beside
the resident-model step time.  The operator kernel of the timed method is timed with HIP
events on the solver's stream (roofline); the CPU FP64 oracle is timed on a
bounded sample of the same workload (cpu_baseline).
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to C_eff @1e-4 err, 256³ elastic: plain vs learned+mixed PCG; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_image(n, seed):
    from chfem_amd import synthetic
    return synthetic.spheres((n, n, n), rmin=6, rmax=12, fraction=0.3, seed=seed)


def hip_sync():
    try:
        C.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def run_method(L, precision, tol, steps, warmup, label, net=None):
    """Prepares the resident solver and times `steps` homogenizations."""
    from chfem_amd.lib import SolveStats
    L.hmgSetInitialGuessNetwork(net.h if net is not None else None)
    L.hmgSetPrecisionFlag(precision)
    L.hmgSetSolverTolerance(tol)
    if L.hmgPrepare() != 1:
        raise RuntimeError("hmgPrepare failed")
    Cm = np.zeros(36)
    try:
        for w in range(warmup):
            log(f"[bench] {label}: warmup {w + 1}/{warmup}")
            L.hmgSolveResident(Cm.ctypes.data_as(C.c_void_p))
        hip_sync()
        times = []
        for s in range(steps):
            t0 = time.perf_counter()
            ok = L.hmgSolveResident(Cm.ctypes.data_as(C.c_void_p))
            hip_sync()
            times.append(time.perf_counter() - t0)
            log(f"[bench] {label}: step {s + 1}/{steps} {times[-1]:.3f} s")
            if ok != 1:
                raise RuntimeError("hmgSolveResident failed")
        st = (SolveStats * 6)()
        cnt = C.c_uint(0)
        L.hmgGetSolveStats(st, C.byref(cnt))
        its = [st[i].iterations for i in range(cnt.value)]
        sweeps = [st[i].sweeps for i in range(cnt.value)]
        tile = np.zeros(3, dtype=np.int32)
        L.cudapcgGetOperatorTile(precision, tile.ctypes.data_as(C.c_void_p))
        form = int(L.cudapcgGetLoopForm(precision))
        x0res = [st[i].x0_residual for i in range(cnt.value)]
        x0used = [st[i].x0_used for i in range(cnt.value)]
        op_ms = C.c_double(0)
        op_n = C.c_uint(0)
        L.hmgGetKernelTiming(C.byref(op_ms), C.byref(op_n))
    finally:
        L.hmgRelease()
        L.hmgSetInitialGuessNetwork(None)
    return {"seconds": float(np.mean(times)) if times else None, "times": times, "C": Cm.reshape(6, 6).copy(),
            "iterations": its, "sweeps": sweeps, "op_ms_total": op_ms.value, "op_launches": op_n.value,
            "x0_residual": x0res, "x0_used": x0used, "tile": "x".join(str(int(t)) for t in tile),
            "form": form}


def run_batched(L, nf, flat, tol, steps, warmup, net):
    """The 6 load cases as one batch of systems (chfemBatch*: every step of
    all systems is one launch of the single-kernel CG step over all their
    tiles), seeded by the learned guess, mixed precision.  A step is one
    chfemBatchSolve on the sample added once (its maps resident): right-hand
    sides, the network's prediction, the solve with its vector allocation,
    and the C columns."""
    from chfem_amd.lib import SolveStats
    h = L.chfemBatchCreate()
    try:
        if L.chfemBatchAddSample(h, nf.encode(), flat.ctypes.data_as(C.c_void_p)) != 1:
            raise RuntimeError("chfemBatchAddSample failed")
        if L.chfemBatchSetInitialGuessNetwork(h, net.h) != 1:
            raise RuntimeError("chfemBatchSetInitialGuessNetwork failed")
        for w in range(warmup):
            log(f"[bench] learned+mixed+batched: warmup {w + 1}/{warmup}")
            L.chfemBatchSolve(h, tol, 20000)
        hip_sync()
        times, dev = [], []
        for k in range(steps):
            t0 = time.perf_counter()
            rc = L.chfemBatchSolve(h, tol, 20000)
            hip_sync()
            times.append(time.perf_counter() - t0)
            dev.append(L.chfemBatchSeconds(h))
            log(f"[bench] learned+mixed+batched: step {k + 1}/{steps} {times[-1]:.3f} s (device {dev[-1]:.3f} s)")
            if rc != 1:
                raise RuntimeError(f"chfemBatchSolve returned {rc}")
        Cm = np.zeros(36)
        L.chfemBatchGetC(h, 0, Cm.ctypes.data_as(C.c_void_p))
        st = (SolveStats * 6)()
        cnt = C.c_uint(0)
        L.chfemBatchGetStats(h, 0, st, C.byref(cnt))
    finally:
        L.chfemBatchDestroy(h)
    return {"seconds": float(np.mean(times)), "device_seconds": float(np.mean(dev)), "C": Cm.reshape(6, 6),
            "iterations": [st[i].iterations for i in range(cnt.value)],
            "sweeps": [st[i].sweeps for i in range(cnt.value)],
            "x0_residual": [st[i].x0_residual for i in range(cnt.value)],
            "x0_used": [st[i].x0_used for i in range(cnt.value)]}


def train_guess(n_samples, size, epochs, seed):
    """Learned initial guess trained on library solutions of seeded images of
    the headline recipe (untimed)."""
    from chfem_amd import guess
    mp = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    t0 = time.perf_counter()
    samples = []
    for k in range(n_samples):
        img = build_image(size, 1000 + seed + k)
        samples.append(guess.solve_fluctuations(img, mp, "elastic", tol=1e-6))
    t1 = time.perf_counter()
    net = guess.LearnedGuess("elastic", seed=seed)
    hist = net.train(samples, epochs=epochs, lr=2e-3, batch=4, log=lambda m: log(f"[bench] guess {m}"))
    t2 = time.perf_counter()
    info = {"samples": n_samples, "size": size, "epochs": epochs, "params": net.num_params,
            "data_seconds": t1 - t0, "train_seconds": t2 - t1, "loss_first": hist[0], "loss_last": hist[-1]}
    return net, info


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family and model
    numbers (a virtualised host may report only the brand) and the logical
    CPU count."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model") and k not in fields:
                    fields[k] = v.strip()
                if not line.strip() and fields:
                    break
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    if "cpu family" in fields:
        name += f" (family {fields['cpu family']}, model {fields.get('model', '?')}, {os.cpu_count()} logical CPUs)"
    return name


def cpu_baseline(img, n, gpu_iterations, iters_sample=50):
    """CPU FP64 oracle (OpenMP, all host threads) on the same image: a
    bounded sample of PCG iterations of load case X timed, scaled by the FP64
    GPU iteration count of the full homogenization; and the 32^3 conductivity
    workload solved fully on the CPU."""
    from oracle import oracle as O
    from chfem_amd import synthetic
    em = synthetic.elem_map(img, [0, 255])
    t0 = time.perf_counter()
    m = O.Model(O.ELASTIC, (n, n, n), em, [1.0, 0.3, 10.0, 0.2])
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.solve(0, tol=1e-30, maxit=1)
    t1 = time.perf_counter()
    m.solve(0, tol=1e-30, maxit=1 + iters_sample)
    t2 = time.perf_counter()
    per_iter = ((t2 - t1) - (t1 - t0)) / iters_sample
    total_it = int(sum(gpu_iterations))
    threads = int(O.lib().orc_num_threads())
    # the 32^3 thermal GRF workload (k 1 / 10, 3 load cases, tol 1e-6), whole
    g = synthetic.grf((32, 32, 32), sigma=2.0, vf=0.4, seed=1)
    t3 = time.perf_counter()
    mc = O.Model(O.THERMAL, (32, 32, 32), synthetic.elem_map(g, [0, 255]), [1.0, 10.0])
    Cc, info = mc.homogenize(tol=1e-6)
    t4 = time.perf_counter()
    return {"value": per_iter * total_it, "unit": "s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": (f"oracle FP64 Jacobi-PCG (NBN) on the same {n}^3 image, load case X, {iters_sample} "
                       f"iterations timed ({per_iter:.4f} s/iteration on {threads} threads, model setup "
                       f"{setup:.1f} s), x {total_it} iterations = FP64 GPU iteration count of the 6 load "
                       f"cases at the same tolerance"),
            "per_iteration_s": per_iter, "iterations_timed": iters_sample,
            "c1_thermal_32": {"seconds": t4 - t3, "iterations": [info[d][1] for d in sorted(info)],
                              "C": Cc.tolist(),
                              "workload": "32^3 GRF (sigma 2, vf 0.4, seed 1), k 1/10, FP64 Jacobi-PCG NBN, "
                                          "tol 1e-6, 3 load cases, solved fully on the CPU"}}


def oracle_columns(n, seed, img_sha):
    """Columns of C the CPU FP64 oracle solved for this image
    (tests/golden/config_headline.json, written by
    tests/golden/make_configs.py), if its image hash matches."""
    p = os.path.join(ROOT, "tests", "golden", "config_headline.json")
    if n != 256 or seed != 3 or not os.path.exists(p):
        return {}, "no oracle fixture for this image"
    with open(p) as f:
        fx = json.load(f)
    if fx.get("image_sha256") != img_sha:
        return {}, "oracle fixture image hash differs: not used"
    cols = {int(k): np.array(v["column"]) for k, v in fx.get("columns", {}).items()}
    return cols, (f"oracle FP64 PCG at tol {fx.get('tol')}, load cases {sorted(cols)}, image hash matches")


def read_traffic(kernel_key):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg (tuning sweeps)")
    ap.add_argument("--train-samples", type=int, default=24)
    ap.add_argument("--train-size", type=int, default=64)
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--only-learned", action="store_true", help="time only the learned method (tuning sweeps)")
    ap.add_argument("--skip-batched", action="store_true", help="omit the batched learned+mixed method")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only; --gpus {args.gpus} is not supported (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)
    # the library prints its solver reports on stdout: send them to stderr so
    # that stdout carries only the JSON line
    json_fd = os.dup(1)
    os.dup2(2, 1)
    n = args.size
    from chfem_amd import io as cio
    from chfem_amd.lib import lib, ptr
    L = lib()
    log(f"[bench] training the learned initial guess ({args.train_samples} x {args.train_size}^3)")
    net, net_info = train_guess(args.train_samples, args.train_size, args.train_epochs, args.seed)
    log(f"[bench] generating {n}^3 sphere image (seed {args.seed})")
    img = build_image(n, args.seed)
    img_sha = hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()
    oracle_cols, oracle_note = oracle_columns(n, args.seed, img_sha)
    tmpd = tempfile.mkdtemp(prefix="chfem_bench_")
    nf = os.path.join(tmpd, f"spheres_{n}.raw")
    cio.export_for_chfem(nf, img, 1, {0: (1.0, 0.3), 255: (10.0, 0.2)}, voxel_size=1.0, export_raw=False,
                         solver_tolerance=args.tol, solver_maxiter=20000)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)

    def init():
        if L.hmgInit(nf.encode(), None, None, ptr(flat)) != 1:
            raise RuntimeError("hmgInit failed")
        L.hmgSetSolverFlag(0)
        L.hmgSetPreConditionerFlag(1)
        L.hmgSetParallelStrategyFlag(0)
        L.hmgSetHomogenizationFlag(6)

    # cold end to end: hmgInit (maps, local matrices) -> upload and solver
    # allocation -> first homogenization (tile autotuner included) -> C,
    # with the network already trained
    cold = {}
    for label, prec, nt in (("learned_mixed", 1, net), ("plain_fp64", 0, None)):
        hip_sync()
        t0 = time.perf_counter()
        init()
        try:
            r = run_method(L, prec, args.tol, 1, 0, f"cold {label}", net=nt)
        finally:
            L.hmgEnd()
        cold[label] = time.perf_counter() - t0
        log(f"[bench] cold {label}: {cold[label]:.3f} s")
    init()
    try:
        log("[bench] reference C: FP64 PCG, tol 1e-9")
        ref = run_method(L, 0, 1e-9, 1, 0, "reference")
        Cref = ref["C"]
        plain = mixed = None
        if not args.only_learned:
            plain = run_method(L, 0, args.tol, args.steps, args.warmup, "plain-fp64")
            mixed = run_method(L, 1, args.tol, args.steps, args.warmup, "mixed")
        fast = run_method(L, 1, args.tol, args.steps, args.warmup, "learned+mixed", net=net)
    finally:
        L.hmgEnd()
    batched = None
    if not args.skip_batched:
        batched = run_batched(L, nf, flat, args.tol, args.steps, args.warmup, net)

    def err(Cm):
        return float(np.abs(Cm - Cref).max() / np.abs(Cref).max())

    def err_oracle(Cm):
        """max over the oracle's solved columns of max |C - C*| / max |C*|"""
        if not oracle_cols:
            return None
        return float(max(np.abs(Cm[:, d] - ref).max() / np.abs(ref).max() for d, ref in oracle_cols.items()))

    N = n ** 3
    op_avg_ms = fast["op_ms_total"] / max(fast["op_launches"], 1)
    # Algorithmic bytes per node of the operator launch inside the CG loop,
    # by the loop form the library used: 1 B image + the 3-vectors streamed
    # per node (12 B in FP32, 24 B in FP64).  Plain apply: d read, q written;
    # form 1: z, d_old read + d, q written; form 2 (single-kernel iteration,
    # z recovered from the two previous directions): d_old2, qt_old, d_old
    # read + d, qt written.  The iterate is updated per window of iterations
    # by the ring flush, outside these launches.
    form = fast["form"]
    nvec = {2: 5, 1: 4}.get(form, 2)
    per_node32 = 1.0 + nvec * 12.0
    plain_form = plain.get("form", form) if plain is not None else form
    per_node64 = 1.0 + {2: 5, 1: 4}.get(plain_form, 2) * 24.0
    bytes_per_launch = per_node32 * N
    achieved = bytes_per_launch / (op_avg_ms * 1e-3) / 1e9
    key = {2: "cg1_", 1: "dir_"}.get(form, "")
    traffic = read_traffic(f"elastic_walsh_{key}fp32_{n}_{fast['tile']}")
    tr_, tc_, lz_ = fast["tile"].split("x")
    # tile rows 65: the kernel's 64-row tiles with the row above on an extra
    # wavefront (its last template argument)
    halo = tr_ == "65"
    kt = f"k_walsh<float,3,{64 if halo else tr_},{tc_},true,{{}},{'true' if halo else 'false'}>"
    geom = f"LZ {lz_}" + (", 64 owned rows per tile + the row above on an extra wavefront" if halo else "")
    kernel = {
        2: kt.format(2) + " (FP32 elastic NBN single-kernel CG iteration: z = d_old - beta' d_old2 - "
           f"alpha M^-1 q, d = z + beta d_old, q = A d, d.q z.q q.M^-1q z.Dz), {geom}",
        1: kt.format(1) + " (FP32 elastic NBN apply with the CG direction update "
           f"d = z + beta d (z = M^-1 r) and d.Ad fused), {geom}",
    }.get(form, kt.format(0) + f" (FP32 elastic NBN apply + fused d.Ad), {geom}")
    cpu = None
    if not args.skip_cpu and plain is not None:
        log(f"[bench] cpu baseline (oracle)")
        cpu = cpu_baseline(img, n, plain["iterations"])
    out = {
        "metric": METRIC,
        "value": fast["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["seconds"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"elastic_{n}^3_spheres_r6-12_f0.3_seed{args.seed}_E1-10_nu0.3-0.2_6loads",
                   "method": "learned initial guess + mixed (FP32 CG, FP64 reductions and reliable updates), "
                             "Jacobi, NBN",
                   "solver_tolerance": args.tol,
                   "tolerance_note": ("PCG stops at sqrt(r.M^-1r / r0.M^-1r0) <= tol; the metric bounds the error "
                                      "of C at 1e-4, which tol 1e-5 meets with a measured margin (err_vs_oracle, "
                                      "err_vs_fp64); the plain and mixed methods run at the same tolerance"),
                   "image_sha256": img_sha,
                   "reference": "CPU FP64 oracle fixture (tests/golden/config_headline.json, tol 1e-8) on its "
                                "columns; GPU FP64 PCG tol 1e-9 on all 36 entries"},
        "err_vs_oracle": err_oracle(fast["C"]),
        "oracle_fixture": oracle_note,
        "err_vs_fp64": err(fast["C"]),
        "cold_seconds": cold,
        "plain_fp64": None if plain is None else {
            "seconds": plain["seconds"], "err_vs_fp64": err(plain["C"]), "err_vs_oracle": err_oracle(plain["C"]),
            "iterations": plain["iterations"],
            "op_avg_ms": plain["op_ms_total"] / max(plain["op_launches"], 1),
            "op_GBps": per_node64 * N / (plain["op_ms_total"] / max(plain["op_launches"], 1) * 1e-3) / 1e9,
            "loop_form": plain_form, "tile": plain["tile"]},
        "mixed": None if mixed is None else {
            "seconds": mixed["seconds"], "err_vs_fp64": err(mixed["C"]), "err_vs_oracle": err_oracle(mixed["C"]),
            "iterations": mixed["iterations"],
            "sweeps": mixed["sweeps"]},
        "learned_mixed": {"seconds": fast["seconds"], "err_vs_fp64": err(fast["C"]), "iterations": fast["iterations"],
                          "sweeps": fast["sweeps"], "x0_residual": fast["x0_residual"], "x0_used": fast["x0_used"],
                          "network": net_info},
        "learned_mixed_batched": None if batched is None else {
            "seconds": batched["seconds"], "device_seconds": batched["device_seconds"],
            "err_vs_fp64": err(batched["C"]), "err_vs_oracle": err_oracle(batched["C"]),
            "iterations": batched["iterations"], "sweeps": batched["sweeps"],
            "x0_residual": batched["x0_residual"], "x0_used": batched["x0_used"],
            "method": "the 6 load cases as one batch of systems: each CG step of all of them is one "
                      "single-kernel launch over their tiles (64x7 halo tiles, z recovered from the direction "
                      "ring) and one finishing launch; FP64 residual replacement per system; the step "
                      "includes vector allocation and the guess's prediction"},
        "speedup_vs_plain": None if plain is None else plain["seconds"] / fast["seconds"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "loop_form": form, "bytes_per_node": per_node32,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": op_avg_ms,
                     "launches": fast["op_launches"]},
        "cpu_baseline": cpu,
    }
    # a result whose C misses the metric's bound is not a valid measurement
    errs = [e for e in (out["err_vs_oracle"], out["err_vs_fp64"]) if e is not None]
    if batched is not None:
        errs += [e for e in (out["learned_mixed_batched"]["err_vs_fp64"],
                             out["learned_mixed_batched"]["err_vs_oracle"]) if e is not None]
    out["valid"] = bool(errs) and max(errs) < 1e-4
    C.CDLL(None).fflush(None)
    sys.stdout.flush()
    os.dup2(json_fd, 1)
    print(json.dumps(out), flush=True)
    if not out["valid"]:
        print(f"bench.py: C misses the 1e-4 bound (errors {errs}); the measurement is invalid", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
