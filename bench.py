"""XC-build benchmark on one MI355X.

Workload (default): B88+LYP XC build for a seeded (H2O)_128 cluster,
def2-SVP (spherical, 3072 AO), 75 radial (Euler-Maclaurin) x 302 Lebedev
points per atom = 8,697,600 grid points, D = C C^T with C a seeded Gaussian
3072 x 640 matrix orthonormalised by QR -- the largest single-GPU
configuration of the benchmark set.  A step is one XC build (D in HBM -> Vxc
in HBM, Exc, N_elec) on a context whose grid, leaves and AO values are
resident, as in an SCF iteration.  The default run also reports, under
"extras", the (H2O)_32 / PBE build rate, one serving point (p50/p99 latency)
and the batched 512-molecule rate against the plain FP64 path.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "fp32cache": 78.6}
# the matrix-core precision each stage runs in, per mode (the FP32 mode keeps
# the density contraction in FP64: D's rounding would repeat at every point)
STAGE_PEAK = {("fp32", "density"): 78.6, ("fp32", "vxc"): 157.3}
PEAK_HBM_GBS = 8000.0
DTYPE = {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3-f32acc", "fp32cache": "f64 (f32 AO cache)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, mode, stage):
    """HBM bytes per launch of a build stage, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/ (None when not measured)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.txt")
    try:
        for line in open(path):
            f = line.split("#")[0].split()
            if len(f) == 4 and f[:3] == [workload, mode, stage]:
                return float(f[3])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def workload(name):
    from occ_amd import molecules
    if name == "h2o32":
        return (molecules.water_cluster(32, seed=0), "def2-svp", True, "pbe", 160,
                "PBE XC build, (H2O)32 cluster seed 0, def2-SVP spherical (768 AO), "
                "75 radial x 302 Lebedev per atom, Becke/TA partition")
    if name == "h2o":
        return (molecules.water(), "6-31g", False, "svwn5", 5,
                "SVWN5 XC build, H2O, 6-31G (13 AO), 75 radial x 302 Lebedev per atom")
    if name == "tpss32":
        return (molecules.water_cluster(32, seed=0), "def2-svp", True, "tpss", 160,
                "TPSS meta-GGA XC build, (H2O)32 cluster seed 0, def2-SVP spherical (768 AO), "
                "75 radial x 302 Lebedev per atom, Becke/TA partition")
    if name == "h2o128":
        return (molecules.water_cluster(128, seed=0), "def2-svp", True, "blyp", 640,
                "B88+LYP XC build, (H2O)128 cluster seed 0, def2-SVP spherical (3072 AO), "
                "75 radial x 302 Lebedev per atom")
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="h2o128")
    ap.add_argument("--modes", default="fp64,fp32cache,fp32,bf16x3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle (no accuracy, "
                    "no cpu_baseline)")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="CPU oracle builds (median)")
    ap.add_argument("--batch-budget-gb", type=float, default=0.0,
                    help="device memory per batch context (default: half of free memory)")
    ap.add_argument("--serve-lambdas", default="50,200,1000,2000")
    ap.add_argument("--serve-sessions", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-pool", type=int, default=12)
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="molecules built one context at a time for the FP64 reference rate")
    ap.add_argument("--cpu-single", type=int, default=1,
                    help="also time one 1-thread build (skipped when it would take over 60 s)")
    ap.add_argument("--extras", type=int, default=-1,
                    help="add the (H2O)32 rate, a serving point and the batched rate "
                         "(default: on for the default workload)")
    ap.add_argument("--serve-extra-lambda", type=float, default=200.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (a known-size kernel against "
                         "which profiler byte counters are scaled)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the XC build shards "
              "by grid leaves across GPUs by design but this build measures one GPU.",
              file=sys.stderr)
        sys.exit(2)

    from occ_amd import AOBasis
    from occ_amd.device import DeviceArray, copy_device, synchronize
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        a, b = DeviceArray(n), DeviceArray(n)
        copy_device(b.ptr, a.ptr, 8 * n)
        synchronize()
        log(f"copy calibration: {8 * n} bytes device to device")
        a.free()
        b.free()
    from occ_amd.dft import XCContext, functional_ids
    from occ_amd.grid import GridSettings, MolecularGrid

    if args.workload == "batch512":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload in ("grad32", "tpssgrad32", "ugrad32", "utpssgrad32"):
        if args.workload in ("tpssgrad32", "utpssgrad32", "ugrad32"):
            args.modes = ",".join(m for m in args.modes.split(",") if m in ("fp64", "fp32cache"))
        return run_gradient(args)
    if args.workload in ("uks32", "utpss32"):
        return run_unrestricted(args)
    if args.workload == "serve":
        print(json.dumps(run_serving(args)), flush=True)
        return
    if args.workload == "vv10_32":
        return run_vv10(args)
    # meta-GGA: FP64 and FP32-cache contexts only; each stage is four GEMMs
    # (rho and the three tau directions; the GGA and three tau Vxc terms)
    mgga = args.workload == "tpss32"
    if mgga:
        args.modes = ",".join(m for m in args.modes.split(",") if m in ("fp64", "fp32cache"))
    atoms, bname, pure, method, nocc, desc = workload(args.workload)
    basis = AOBasis.load(atoms, bname, pure=pure)
    nbf = basis.nbf
    t0 = time.time()
    grid = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302))
    pts, wts = grid.build(device=0)
    log(f"grid: {len(wts)} points, {nbf} AO, built in {time.time() - t0:.2f}s")
    rng = np.random.default_rng(0)
    C, _ = np.linalg.qr(rng.normal(size=(nbf, nocc)))
    D = np.asfortranarray(C @ C.T)
    ids = functional_ids(method)

    ref = None
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ob = oracle.OracleBasis.from_flat(basis.flat())
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        times = []
        for _ in range(args.cpu_repeats):
            t0 = time.time()
            K, exc, nel = oracle.xc_build(ob, pts, wts, ids, D, nthreads=threads)
            times.append(time.time() - t0)
        dt = float(np.median(times))
        ref = (K, exc, nel)
        t1 = None
        if args.cpu_single and dt * threads <= 60.0:
            t0 = time.time()
            oracle.xc_build(ob, pts, wts, ids, D, nthreads=1)
            t1 = time.time() - t0
        cpu = {"value": 1.0 / dt, "unit": "XC builds/s", "cores": threads, "kind": "port",
               "sample": f"full FP64 screened builds of the same workload on the CPU oracle "
                         f"(oracle/occxc_oracle.c: same leaves, screening and per-leaf GEMMs, "
                         f"OpenMP over leaves, hand-blocked loops, no BLAS); median of "
                         f"{len(times)} builds on {threads} threads = {dt:.2f} s"
                         + (f"; one build on 1 thread = {t1:.2f} s" if t1 else ""),
               "grid_pts_per_s": len(wts) / dt,
               "single_thread_builds_per_s": (1.0 / t1) if t1 else None,
               "cpu_model": cpu_model()}
        log(f"cpu oracle: {dt:.2f}s on {threads} threads (1 thread: {t1}), Exc={exc:.10f} "
            f"N={nel:.10f}")

    d_D = DeviceArray(nbf * nbf)
    d_D.upload(D)
    d_V = DeviceArray(nbf * nbf)
    d_r = DeviceArray(2)
    results = {}
    for mode in [m for m in args.modes.split(",") if m]:
        ctx = XCContext(basis, pts, wts, ids, precision=mode)
        info = ctx.info()
        for _ in range(args.warmup):
            ctx.build_device(d_D.ptr, d_V.ptr, d_r.ptr)
        ctx.synchronize()
        ctx.set_timing(True)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.build_device(d_D.ptr, d_V.ptr, d_r.ptr)
        ctx.synchronize()
        synchronize()
        wall = time.perf_counter() - t0
        tim = ctx.timing()
        V = d_V.download((nbf, nbf))
        exc, nel = d_r.download()
        # cold build: AO evaluation + build
        ctx.set_timing(True)
        t1 = time.perf_counter()
        ctx.evaluate_ao()
        ctx.build_device(d_D.ptr, d_V.ptr, d_r.ptr)
        ctx.synchronize()
        cold = time.perf_counter() - t1
        tcold = ctx.timing()
        ms = wall / args.steps * 1e3
        r = {"ms_per_step": ms, "builds_per_s": 1e3 / ms, "info": info, "timing": tim,
             "cold_ms": cold * 1e3, "ao_ms": tcold["ao"][0], "exc": exc, "nelec": nel}
        if ref is not None:
            r["dexc"] = abs(exc - ref[1])
            r["dvxc"] = float(np.abs(V - ref[0]).max())
            r["dnelec"] = abs(nel - ref[2])
        log(f"{mode}: {ms:.3f} ms/build, stages {tim}, cold {cold * 1e3:.2f} ms, "
            f"dExc={r.get('dexc')}, dV={r.get('dvxc')}, dN={r.get('dnelec')}")
        results[mode] = r
        ctx.close()

    # headline: the fastest mode within every stated tolerance of the FP64
    # oracle (|dExc| <= 1e-6 Ha, max|dVxc| <= 1e-7, |dN_elec| <= 1e-8);
    # without the oracle only FP64 (the reference precision) qualifies
    if ref is None:
        ok = [m for m in results if m == "fp64"]
    else:
        ok = [m for m in results if results[m]["dexc"] <= 1e-6 and results[m]["dvxc"] <= 1e-7
              and results[m]["dnelec"] <= 1e-8]
    head = min(ok or list(results), key=lambda m: results[m]["ms_per_step"])
    R = results[head]
    tim = R["timing"]
    dens_ms = tim["density"][0] / max(tim["density"][1], 1)
    vxc_ms = tim["vxc"][0] / max(tim["vxc"][1], 1)
    flops_half = R["info"]["contraction_flops"] / 2.0   # 2 n m^2 per GEMM
    if mgga:
        flops_half *= 4.0
    kname, kms = ("density", dens_ms) if dens_ms >= vxc_ms else ("vxc", vxc_ms)
    achieved = flops_half / (kms * 1e-3) / 1e12
    peak = STAGE_PEAK.get((head, kname), PEAK_TFLOPS[head])
    roof = {"bound": "mfma", "kernel": kname, "achieved": achieved,
            "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": measured_traffic(args.workload, head, kname), "kernel_ms": kms,
            "work": "sum over leaves of 2 n_b m_b^2 (one GEMM of the build, screened m_b)"
                    + (" x 4 (rho and three tau GEMMs per stage)" if mgga else "")}
    # the other contraction stage against its own peak
    oname, oms = ("vxc", vxc_ms) if kname == "density" else ("density", dens_ms)
    o_ach = flops_half / (oms * 1e-3) / 1e12
    o_peak = STAGE_PEAK.get((head, oname), PEAK_TFLOPS[head])
    other = {"bound": "mfma", "kernel": oname, "achieved": o_ach, "peak": o_peak,
             "unit": "TFLOP/s", "frac": o_ach / o_peak,
             "traffic": measured_traffic(args.workload, head, oname), "kernel_ms": oms}
    ao_gbs = R["info"]["ao_write_bytes"] / (R["ao_ms"] * 1e-3) / 1e9 if R["ao_ms"] else None
    out = {
        "metric": "XC builds/s + grid pts/s at |dExc|<=1e-6 Ha vs FP64; p50/p99 request latency",
        "value": R["builds_per_s"],
        "unit": "XC builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic",
        "config": {"workload": desc, "npts": len(wts), "nbf": nbf,
                   "leaves": R["info"]["nleaves"], "mean_local_nbf": R["info"]["mean_local_nbf"],
                   "precision": head},
        "grid_pts_per_s": len(wts) * R["builds_per_s"],
        "roofline": roof,
        "roofline_other_stage": other,
        "ao_eval": {"ms": R["ao_ms"], "algorithmic_bytes": R["info"]["ao_write_bytes"],
                    "achieved_GBs": ao_gbs, "peak_GBs": PEAK_HBM_GBS,
                    "frac": (ao_gbs / PEAK_HBM_GBS) if ao_gbs else None},
        "cold_build_ms": R["cold_ms"],
        "accuracy": {m: {k: results[m].get(k) for k in ("dexc", "dvxc", "dnelec")}
                     for m in results},
        "modes_ms": {m: results[m]["ms_per_step"] for m in results},
        "cpu_baseline": cpu,
    }
    if args.extras == 1 or (args.extras == -1 and args.workload == "h2o128"):
        out["extras"] = run_extras(args)
    print(json.dumps(out), flush=True)


def run_rate(name, mode, steps, warmup):
    """Warm FP64 build rate of a single-molecule workload (no CPU oracle)."""
    from occ_amd import AOBasis
    from occ_amd.device import DeviceArray, synchronize
    from occ_amd.dft import XCContext, functional_ids
    from occ_amd.grid import GridSettings, MolecularGrid
    atoms, bname, pure, method, nocc, desc = workload(name)
    basis = AOBasis.load(atoms, bname, pure=pure)
    nbf = basis.nbf
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=76,
                                                 angular_points=302)).build(device=0)
    D = seeded_density(nbf, nocc, 0)
    d_D, d_V, d_r = DeviceArray(nbf * nbf), DeviceArray(nbf * nbf), DeviceArray(2)
    d_D.upload(D)
    ctx = XCContext(basis, pts, wts, functional_ids(method), precision=mode)
    for _ in range(warmup):
        ctx.build_device(d_D.ptr, d_V.ptr, d_r.ptr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_device(d_D.ptr, d_V.ptr, d_r.ptr)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    ctx.close()
    for a in (d_D, d_V, d_r):
        a.free()
    log(f"{name} {mode}: {ms:.3f} ms/build")
    return {"workload": desc, "precision": mode, "npts": len(wts), "nbf": nbf,
            "ms_per_build": ms, "builds_per_s": 1e3 / ms, "grid_pts_per_s": len(wts) * 1e3 / ms}


def run_extras(args):
    """The other benchmark configurations, one line each: the (H2O)32 / PBE
    build rate, one serving point and the batched launch."""
    import copy
    ex = {}
    ex["h2o32_pbe"] = run_rate("h2o32", "fp64", args.steps, args.warmup)
    a = copy.copy(args)
    a.serve_lambdas = f"{args.serve_extra_lambda:g}"
    srv = run_serving(a)
    row = srv["sweep"][0]
    ex["serving"] = {"workload": srv["config"]["workload"], "lambda_req_per_s": row["lambda_req_per_s"],
                     "p50_ms": row["p50_ms"], "p99_ms": row["p99_ms"],
                     "throughput_builds_per_s": row["throughput_builds_per_s"],
                     "builds": row["builds"], "errors": row["errors"]}
    a = copy.copy(args)
    a.modes = "fp64,fp32"  # FP64, and the fastest mode within the tolerances
    b = run_batch(a)
    ex["batch512"] = {"workload": b["config"]["workload"], "precision": b["config"]["precision"],
                      "molecules_per_s": b["value"],
                      "cold_molecules_per_s": b["cold_molecules_per_s"],
                      "plain_fp64_molecules_per_s": b["plain_fp64_molecules_per_s"],
                      "speedup_vs_plain_fp64": b["speedup_vs_plain_fp64"],
                      "modes": b["modes"]}
    return ex


def run_gradient(args):
    """XC nuclear gradient (compute_xc_gradient) of the (H2O)32 / def2-SVP /
    PBE workload: a step is one gradient from a resident context (D in ->
    3 x 96 gradient out: D2, the density stage and the gradient kernels); the
    first call, which evaluates the AO Hessian once, is a warm-up.  ugrad32 /
    utpssgrad32: the unrestricted gradient (PBE / TPSS) of the doublet cation
    of the unrestricted build (D_s = C_s C_s^T / 2, 160 alpha, 159 beta)."""
    from occ_amd import AOBasis
    from occ_amd.dft import XCContext, functional_ids
    from occ_amd.device import synchronize
    from occ_amd.grid import GridSettings, MolecularGrid
    uks = args.workload in ("ugrad32", "utpssgrad32")
    mg = args.workload in ("tpssgrad32", "utpssgrad32")
    atoms, bname, pure, method, nocc, desc = workload("tpss32" if mg else "h2o32")
    basis = AOBasis.load(atoms, bname, pure=pure)
    nbf = basis.nbf
    grid = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302))
    pts, wts = grid.build(device=0)
    if uks:
        D = np.asfortranarray(np.vstack([seeded_density(nbf, nocc, 0) / 2,
                                         seeded_density(nbf, nocc - 1, 1) / 2]))
        desc = "unrestricted " + desc + ", doublet cation (160 alpha, 159 beta)"
    else:
        rng = np.random.default_rng(0)
        C, _ = np.linalg.qr(rng.normal(size=(nbf, nocc)))
        D = np.asfortranarray(C @ C.T)
    ids = functional_ids(method)
    sa, nat = basis.shell_atom, len(atoms)
    ref, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ob = oracle.OracleBasis.from_flat(basis.flat())
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        times = []
        for _ in range(max(1, min(args.cpu_repeats, 3))):
            t0 = time.time()
            ref = (oracle.xc_gradient_unrestricted if uks else oracle.xc_gradient)(
                ob, sa, nat, pts, wts, ids, D, nthreads=threads)
            times.append(time.time() - t0)
        dt = float(np.median(times))
        cpu = {"value": 1.0 / dt, "unit": "XC gradients/s", "cores": threads, "kind": "port",
               "sample": f"full FP64 screened XC gradients of the same workload on the CPU "
                         f"oracle (oracle/occxc_oracle.c: the reference's per-leaf V_x, V_y, "
                         f"V_z matrices and distribution, OpenMP over leaves); median of "
                         f"{len(times)} on {threads} threads = {dt:.2f} s",
               "cpu_model": cpu_model()}
        log(f"cpu oracle gradient: {dt:.2f}s on {threads} threads")
    from occ_amd.device import DeviceArray
    d_D = DeviceArray(D.size)
    d_D.upload(D)
    d_G = DeviceArray(3 * nat)
    results = {}
    for mode in [m for m in args.modes.split(",") if m]:
        ctx = XCContext(basis, pts, wts, ids, precision=mode)
        info = ctx.info()
        grad = ctx.xc_gradient_unrestricted_device if uks else ctx.xc_gradient_device
        for _ in range(max(args.warmup, 1)):
            grad(d_D.ptr, sa, nat, d_G.ptr)
        ctx.synchronize()
        ctx.set_timing(True)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            grad(d_D.ptr, sa, nat, d_G.ptr)
        ctx.synchronize()
        synchronize()
        wall = time.perf_counter() - t0
        G = d_G.download((3, nat))
        tim = ctx.timing()
        ms = wall / args.steps * 1e3
        r = {"ms_per_step": ms, "per_s": 1e3 / ms, "info": info, "timing": tim}
        if ref is not None:
            r["dgrad"] = float(np.abs(G - ref).max())
        log(f"{mode}: {ms:.3f} ms/gradient, stages {tim}, max|dG|={r.get('dgrad')}")
        results[mode] = r
        ctx.close()
    # headline: FP64, the reference's precision (the other modes are reported
    # with their accuracy)
    head = "fp64" if "fp64" in results else min(results, key=lambda m: results[m]["ms_per_step"])
    R = results[head]
    tim = R["timing"]
    gms = tim["gradient"][0] / max(tim["gradient"][1], 1)
    flops = R["info"]["contraction_flops"]   # T = Phi P and U = Aow P: 2 x 2 n m^2
    if mg:
        flops *= 2.5   # + T_x, T_y, T_z of the tau terms: 5 x 2 n m^2
    if uks:
        flops *= 2     # once per spin
    achieved = flops / (gms * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "gradient", "achieved": achieved,
            "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[head],
            "traffic": measured_traffic(args.workload, head, "gradient"), "kernel_ms": gms,
            "work": "sum over leaves of 4 n_b m_b^2 (T = Phi P and U = Aow P, screened m_b)"
                    + (" + 6 n_b m_b^2 (the three tau GEMMs)" if mg else "")
                    + (", per spin" if uks else "")}
    out = {
        "metric": ("unrestricted " if uks else "") + "XC nuclear gradients/s (FP64)",
        "value": R["per_s"], "unit": "XC gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[head], "data": "synthetic",
        "config": {"workload": "XC gradient of the " + desc, "npts": len(wts), "nbf": nbf,
                   "natoms": nat, "leaves": R["info"]["nleaves"], "precision": head},
        "roofline": roof,
        "accuracy": {m: results[m].get("dgrad") for m in results},
        "modes_ms": {m: results[m]["ms_per_step"] for m in results},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_unrestricted(args):
    """Unrestricted XC build of the (H2O)32 / def2-SVP / PBE workload for a
    doublet cation: alpha C (768 x 160) and beta C (768 x 159) seeded Gaussian,
    orthonormalised, D_s = C_s C_s^T / 2 (OCC's density_matrix_unrestricted).
    A step is one build from D (2 nbf x nbf) in HBM to both spins' Vxc in HBM,
    Exc, N_alpha, N_beta.  utpss32 is the same doublet with the TPSS meta-GGA
    (compute_K_dft<2, Unrestricted>: per spin, rho and the three tau
    directions in the density stage, the GGA and three tau terms in Vxc)."""
    from occ_amd import AOBasis
    from occ_amd.dft import XCContext, functional_ids
    from occ_amd.device import DeviceArray, synchronize
    from occ_amd.grid import GridSettings, MolecularGrid
    mgga = args.workload == "utpss32"
    atoms, bname, pure, method, nocc, desc = workload("tpss32" if mgga else "h2o32")
    basis = AOBasis.load(atoms, bname, pure=pure)
    nbf = basis.nbf
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302)).build(
        device=0)
    Du = np.vstack([seeded_density(nbf, nocc, 0) / 2, seeded_density(nbf, nocc - 1, 1) / 2])
    ids = functional_ids(method)
    ref, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ob = oracle.OracleBasis.from_flat(basis.flat())
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        times = []
        for _ in range(max(1, min(args.cpu_repeats, 3))):
            t0 = time.time()
            ref = oracle.xc_build_unrestricted(ob, pts, wts, ids, Du, nthreads=threads)
            times.append(time.time() - t0)
        dt = float(np.median(times))
        cpu = {"value": 1.0 / dt, "unit": "UKS XC builds/s", "cores": threads, "kind": "port",
               "sample": f"full FP64 screened unrestricted builds of the same workload on the "
                         f"CPU oracle (oracle/occxc_oracle.c, OpenMP over leaves); median of "
                         f"{len(times)} on {threads} threads = {dt:.2f} s",
               "cpu_model": cpu_model()}
        log(f"cpu oracle unrestricted build: {dt:.2f}s on {threads} threads")
    d_D, d_V, d_r = DeviceArray(2 * nbf * nbf), DeviceArray(2 * nbf * nbf), DeviceArray(3)
    d_D.upload(np.asfortranarray(Du))
    results = {}
    for mode in [m for m in args.modes.split(",") if m in ("fp64", "fp32cache")]:
        ctx = XCContext(basis, pts, wts, ids, precision=mode)
        info = ctx.info()
        for _ in range(max(args.warmup, 1)):
            ctx.build_unrestricted_device(d_D.ptr, d_V.ptr, d_r.ptr)
        ctx.synchronize()
        ctx.set_timing(True)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.build_unrestricted_device(d_D.ptr, d_V.ptr, d_r.ptr)
        ctx.synchronize()
        synchronize()
        wall = time.perf_counter() - t0
        tim = ctx.timing()
        V = d_V.download((2 * nbf, nbf))
        exc, na, nb = d_r.download()
        ms = wall / args.steps * 1e3
        r = {"ms_per_step": ms, "per_s": 1e3 / ms, "info": info, "timing": tim, "exc": exc}
        if ref is not None:
            r["dexc"] = abs(exc - ref[1])
            r["dvxc"] = float(np.abs(V - ref[0]).max())
            r["dnelec"] = max(abs(na - ref[2][0]), abs(nb - ref[2][1]))
        log(f"{mode}: {ms:.3f} ms/UKS build, stages {tim}, dExc={r.get('dexc')}, "
            f"dV={r.get('dvxc')}, dN={r.get('dnelec')}")
        results[mode] = r
        ctx.close()
    head = "fp64" if "fp64" in results else next(iter(results))
    R = results[head]
    tim = R["timing"]
    dens_ms = tim["density"][0] / max(tim["density"][1], 1)
    vxc_ms = tim["vxc"][0] / max(tim["vxc"][1], 1)
    kname, kms = ("density", dens_ms) if dens_ms >= vxc_ms else ("vxc", vxc_ms)
    flops = R["info"]["contraction_flops"]   # per stage: one 2 n m^2 GEMM per spin
    if mgga:
        flops *= 4
    achieved = flops / (kms * 1e-3) / 1e12
    out = {
        "metric": ("unrestricted meta-GGA" if mgga else "unrestricted") + " XC builds/s (FP64)",
        "value": R["per_s"], "unit": "UKS XC builds/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[head], "data": "synthetic",
        "config": {"workload": "unrestricted " + desc + ", doublet cation (160 alpha, 159 beta)",
                   "npts": len(wts), "nbf": nbf, "leaves": R["info"]["nleaves"],
                   "precision": head},
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved,
                     "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS[head],
                     "traffic": measured_traffic(args.workload, head, kname), "kernel_ms": kms,
                     "work": "sum over leaves of 2 n_b m_b^2 per spin (two GEMMs per stage)"
                             + (", times four for rho and the three tau directions"
                                if mgga else "")},
        "accuracy": {m: {k: results[m].get(k) for k in ("dexc", "dvxc", "dnelec")}
                     for m in results},
        "modes_ms": {m: results[m]["ms_per_step"] for m in results},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_vv10(args):
    """VV10 non-local correlation (NonLocalCorrelationFunctional::vv10) of the
    (H2O)32 / def2-SVP system on an NLC grid of 50 radial x 110 Lebedev points
    per atom (the reference's NLC grid has at most 110 angular points):
    a step is one evaluation from D in HBM to E_nlc and Vxc in HBM (D2, rho and
    grad rho of every point, the O(N^2) pair sums over earlier points, the
    coefficients and the Vxc contraction)."""
    from occ_amd import AOBasis, molecules
    from occ_amd.dft import XCContext
    from occ_amd.device import DeviceArray, synchronize
    from occ_amd.grid import GridSettings, MolecularGrid
    atoms = molecules.water_cluster(32, seed=0)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    nbf = basis.nbf
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=50, angular_points=110)).build(
        device=0)
    npts = len(wts)
    D = seeded_density(nbf, 160, 0)
    pairs = npts * (npts - 1) / 2.0
    flops_per_pair = 27.0  # algorithmic FP64 operations of one pair (one reciprocal counted once)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        # bounded sample: the oracle's pair sums over the first m points (the
        # pair loop is O(N^2); the AO and Vxc parts are O(N) and not timed)
        m = min(npts, 360000)
        rng = np.random.default_rng(1)
        rho = 10 ** rng.uniform(-6, 0, m)
        sig = 10 ** rng.uniform(-2, 0, m) * rho ** (8 / 3)
        t0 = time.time()
        oracle.vv10_kernel(rho, sig, pts[:m], wts[:m], 6.0, 0.01, nthreads=threads)
        dt = time.time() - t0
        pps = m * (m - 1) / 2.0 / dt
        cpu = {"value": pps / pairs, "unit": "VV10 evaluations/s", "cores": threads,
               "kind": "port",
               "sample": f"the oracle's VV10 pair sums (oracle/occxc_oracle.c, OpenMP over "
                         f"points) over the first {m} of the {npts} points, {dt:.2f} s on "
                         f"{threads} threads = {pps:.3e} pairs/s, scaled to the workload's "
                         f"{pairs:.3e} pairs",
               "cpu_model": cpu_model()}
        log(f"cpu oracle VV10 sample: {m} points in {dt:.2f}s ({pps:.3e} pairs/s)")
    ctx = XCContext(basis, pts, wts, [101])
    d_D, d_V, d_r = DeviceArray(nbf * nbf), DeviceArray(nbf * nbf), DeviceArray(2)
    d_D.upload(D)
    for _ in range(max(args.warmup, 1)):
        ctx.nlc_vv10_device(d_D.ptr, d_V.ptr, d_r.ptr)
    ctx.synchronize()
    ctx.set_timing(True)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.nlc_vv10_device(d_D.ptr, d_V.ptr, d_r.ptr)
    ctx.synchronize()
    synchronize()
    wall = time.perf_counter() - t0
    tim = ctx.timing()
    e_nlc, nel = d_r.download()
    ms = wall / args.steps * 1e3
    nlc_ms = tim["nlc"][0] / max(tim["nlc"][1], 1)
    achieved = pairs * flops_per_pair / (nlc_ms * 1e-3) / 1e12
    log(f"VV10: {ms:.3f} ms/evaluation, stages {tim}, E_nlc={e_nlc:.10f}, N={nel:.6f}")
    ctx.close()
    out = {
        "metric": "VV10 evaluations/s (FP64)",
        "value": 1e3 / ms, "unit": "VV10 evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "VV10 (b = 6.0, C = 0.01), (H2O)32 cluster seed 0, def2-SVP "
                               "spherical (768 AO), 50 radial x 110 Lebedev per atom",
                   "npts": npts, "nbf": nbf, "pairs": pairs},
        "roofline": {"bound": "fp64-valu", "kernel": "k_vv10_pairs (nlc stage)",
                     "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": achieved / 78.6, "traffic": None, "kernel_ms": nlc_ms,
                     "work": f"N(N-1)/2 pairs x {flops_per_pair:.0f} FP64 operations"},
        "pairs_per_s": pairs / (nlc_ms * 1e-3),
        "e_nlc": e_nlc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def seeded_density(nbf, nocc, seed):
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.normal(size=(nbf, min(nocc, nbf))))
    return np.asfortranarray(C @ C.T)


def run_batch(args):
    """512 seeded molecules (half water clusters n ~ U{1..16}, half alkanes
    n ~ U{1..24}; 6-31G or def2-SVP each), PBE, 50 radial x 194 Lebedev per
    atom.  The molecules are packed into batch contexts that fit a device-memory
    budget (their AO values together do not fit at once); per context: create
    (grid leaves + AO values), one warm-up build, timed builds replayed from a
    captured graph.  Reported: molecules/s of warm builds (AO values resident,
    as across SCF iterations) and of cold builds (context creation included),
    against the same molecules built one context at a time in FP64."""
    from occ_amd import AOBasis, molecules
    from occ_amd.device import DeviceArray, synchronize
    from occ_amd.dft import XCBatch, XCContext, estimate_batch_bytes, functional_ids
    from occ_amd.grid import GridSettings, MolecularGrid
    mols = molecules.batch_set(512, seed=0)
    ids = functional_ids("pbe")
    settings = GridSettings(radial_points=51, angular_points=194)
    t0 = time.time()
    grids = [MolecularGrid(atoms, settings).build(device=0) for _, atoms, _ in mols]
    t_grid = time.time() - t0
    bases = [AOBasis.load(atoms, bn, pure=(bn == "def2-svp")) for _, atoms, bn in mols]
    Ds = [seeded_density(b.nbf, max(1, molecules.nuclear_charge(a) // 2), 1000 + i)
          for i, (b, (_, a, _)) in enumerate(zip(bases, mols))]
    npts = sum(len(w) for _, w in grids)
    log(f"batch: {len(mols)} molecules, {npts} points, grids in {t_grid:.1f}s")
    from occ_amd.device import memory_info
    budget = args.batch_budget_gb * 1e9 if args.batch_budget_gb else 0.5 * memory_info()[0]
    results = {}
    for mode in [m for m in args.modes.split(",") if m]:
        t0 = time.time()
        est = list(estimate_batch_bytes(bases, grids, ids, precision=mode))
        t_est = time.time() - t0
        chunks, cur, cur_b = [], [], 0.0
        for i, e in enumerate(est):
            if cur and cur_b + e > budget:
                chunks.append(cur)
                cur, cur_b = [], 0.0
            cur.append(i)
            cur_b += e
        if cur:
            chunks.append(cur)
        warm = cold = 0.0
        tuned = []
        exc = np.zeros(len(mols))
        nel = np.zeros(len(mols))
        for ch in chunks:
            synchronize()
            t1 = time.perf_counter()
            batch = XCBatch([bases[i] for i in ch], [grids[i] for i in ch], ids, precision=mode)
            binfo = batch.info()
            tuned.append({"molecules": len(ch), "tile_ao": binfo["tile_ao"],
                          "persistent": binfo["persistent"], "tuning_row": binfo["tuning_row"]})
            off = batch.offsets
            tot = int(off[-1] + batch.nbf[-1] ** 2)
            hD = np.zeros(tot)
            for k, i in enumerate(ch):
                hD[off[k]:off[k] + batch.nbf[k] ** 2] = Ds[i].ravel(order="F")
            dD, dV, dr = DeviceArray(tot), DeviceArray(tot), DeviceArray(2 * len(ch))
            dD.upload(hD)
            batch.set_graph(True)
            batch.build_device(dD.ptr, dV.ptr, dr.ptr)
            batch.synchronize()
            cold += time.perf_counter() - t1
            t2 = time.perf_counter()
            for _ in range(args.steps):
                batch.build_device(dD.ptr, dV.ptr, dr.ptr)
            batch.synchronize()
            warm += (time.perf_counter() - t2) / args.steps
            r = dr.download()
            exc[ch] = r[0::2]
            nel[ch] = r[1::2]
            batch.close()
            for a in (dD, dV, dr):
                a.free()
        results[mode] = {"warm_s": warm, "cold_s": cold, "chunks": len(chunks),
                         "estimate_s": t_est, "exc": exc, "nel": nel,
                         "bytes": float(sum(est)), "tuning": tuned}
        log(f"batch {mode}: {len(chunks)} contexts, warm {warm * 1e3:.1f} ms "
            f"({len(mols) / warm:.0f} mol/s), cold {cold:.2f} s, estimate {t_est:.1f} s, "
            f"{sum(est) / 1e9:.1f} GB")
    # plain FP64: one context per molecule, built one after another
    sample = list(range(0, len(mols), max(1, len(mols) // args.plain_sample)))
    t_plain = 0.0
    plain_exc = {}
    for i in sample:
        ctx = XCContext(bases[i], grids[i][0], grids[i][1], ids, precision="fp64")
        n = bases[i].nbf
        dD, dV, dr = DeviceArray(n * n), DeviceArray(n * n), DeviceArray(2)
        dD.upload(Ds[i])
        ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
        ctx.synchronize()
        t_plain += (time.perf_counter() - t1) / args.steps
        plain_exc[i] = dr.download()[0]
        ctx.close()
        for a in (dD, dV, dr):
            a.free()
    plain_rate = len(sample) / t_plain
    # accuracy against the FP64 oracle on a sample of molecules
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    acc_sample = sample[:: max(1, len(sample) // 8)][:8]
    worst = {m: 0.0 for m in results}
    worst_n = {m: 0.0 for m in results}
    for i in acc_sample:
        _, e_ref, n_ref = oracle.xc_build(oracle.OracleBasis.from_flat(bases[i].flat()),
                                          grids[i][0], grids[i][1], ids, Ds[i])
        for m in results:
            worst[m] = max(worst[m], abs(results[m]["exc"][i] - e_ref))
            worst_n[m] = max(worst_n[m], abs(results[m]["nel"][i] - n_ref))
    ok = [m for m in results if worst[m] <= 1e-6 and worst_n[m] <= 1e-8]
    head = min(ok or list(results), key=lambda m: results[m]["warm_s"])
    R = results[head]
    out = {
        "metric": "XC builds/s (molecules/s) for the batched launch at |dExc|<=1e-6 Ha vs FP64",
        "value": len(mols) / R["warm_s"],
        "unit": "molecules/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": R["warm_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[head], "data": "synthetic",
        "config": {"workload": "512 seeded molecules (water clusters n~U{1..16}, alkanes "
                               "n~U{1..24}; 6-31G/def2-SVP), PBE, 50x194 grid",
                   "npts": npts, "contexts": R["chunks"], "precision": head},
        "cold_molecules_per_s": len(mols) / R["cold_s"],
        "plain_fp64_molecules_per_s": plain_rate,
        "speedup_vs_plain_fp64": (len(mols) / R["warm_s"]) / plain_rate,
        "modes": {m: {"warm_molecules_per_s": len(mols) / results[m]["warm_s"],
                      "cold_molecules_per_s": len(mols) / results[m]["cold_s"],
                      "contexts": results[m]["chunks"], "device_gb": results[m]["bytes"] / 1e9,
                      "max_dexc_sample": worst[m], "max_dnelec_sample": worst_n[m],
                      "launch_settings": results[m]["tuning"]}
                  for m in results},
        "grid_build_s": t_grid,
    }
    return out


def run_serving(args):
    """Serving: SCF sessions arrive as a Poisson process; each runs U{8..20}
    XC-build iterations back to back (the next D is submitted when the previous
    build returns, as an SCF loop does).  Molecule sizes are log-uniform in
    3..300 atoms (water clusters or alkanes, def2-SVP, PBE, 50x194 grid).
    Session arrival rate = lambda / 14 so that the offered request rate is
    lambda.  Every session is served by the library's native loop
    (occxc_server_*): admission by device-memory footprint, contexts resident
    across a session's iterations, builds of different sessions concurrent on
    several HIP streams.  Reported per lambda: p50/p99 latency (submit to
    result) and throughput (completed builds / makespan)."""
    import threading
    from occ_amd import AOBasis, molecules
    from occ_amd.dft import functional_ids
    from occ_amd.grid import GridSettings, MolecularGrid
    from occ_amd.serving import XCServer
    rng = np.random.default_rng(7)
    ids = functional_ids("pbe")
    settings = GridSettings(radial_points=51, angular_points=194)
    pool = []
    t0 = time.time()
    for k in range(args.serve_pool):
        natoms = int(round(np.exp(rng.uniform(np.log(3), np.log(300)))))
        name, atoms = molecules.sized_molecule(natoms, rng)
        basis = AOBasis.load(atoms, "def2-svp", pure=True)
        pts, w = MolecularGrid(atoms, settings).build(device=0)
        nocc = max(1, molecules.nuclear_charge(atoms) // 2)
        Ds = [seeded_density(basis.nbf, nocc, 5000 + 10 * k + j) for j in range(2)]
        pool.append((name, len(atoms), basis, pts, w, Ds))
    log(f"serving pool: {[p[0] for p in pool]} prepared in {time.time() - t0:.1f}s")
    sweep = []
    for lam in [float(x) for x in args.serve_lambdas.split(",") if x]:
        lat = []
        lock = threading.Lock()
        errors = []
        with XCServer(slots=args.serve_slots) as srv:
            sessions = []
            t_next = 0.0
            for s in range(args.serve_sessions):
                t_next += rng.exponential(14.0 / lam)
                sessions.append((t_next, pool[int(rng.integers(len(pool)))],
                                 int(rng.integers(8, 21))))

            def run_session(start, item, iters, t_origin):
                name, na, basis, pts, w, Ds = item
                delay = t_origin + start - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                try:
                    scf = srv.open(basis, pts, w, ids)
                    for it in range(iters):
                        _, _, _, ms = srv.wait(srv.submit(scf, Ds[it % 2]))
                        with lock:
                            lat.append(ms)
                    srv.close_scf(scf)
                except Exception as e:  # reported, not hidden
                    with lock:
                        errors.append(repr(e))

            t_origin = time.perf_counter()
            threads = [threading.Thread(target=run_session, args=(st, it, n, t_origin))
                       for st, it, n in sessions]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
            span = time.perf_counter() - t_origin
            stats = srv.stats()
        lat = np.array(lat)
        row = {"lambda_req_per_s": lam, "sessions": len(sessions), "builds": int(lat.size),
               "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
               "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
               "throughput_builds_per_s": lat.size / span, "makespan_s": span,
               "evictions": stats["evictions"], "deferred": stats["admissions_deferred"],
               "errors": errors[:3]}
        log(f"lambda {lam:g}: {row}")
        sweep.append(row)
    best = max(sweep, key=lambda r: r["throughput_builds_per_s"])
    out = {
        "metric": "serving: XC builds/s with p50/p99 request latency",
        "value": best["throughput_builds_per_s"], "unit": "XC builds/s",
        "n_gpus": 1, "steps": sum(r["builds"] for r in sweep), "warmup": 0,
        "ms_per_step": 1e3 / best["throughput_builds_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Poisson SCF sessions, U{8..20} iterations, log-uniform 3-300 "
                               "atoms, def2-SVP, PBE, 50x194 grid", "slots": args.serve_slots,
                   "pool": [p[0] for p in pool]},
        "sweep": sweep,
    }
    return out


if __name__ == "__main__":
    main()
