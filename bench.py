#!/usr/bin/env python3
"""CG-recon solves/s of the MI355X Toeplitz subspace reconstruction.

Workload (one step = one solve): single 2D subspace reconstruction, image
256 x 256 (Toeplitz grid 512 x 512), 20 coils, golden-angle radial trajectory
with 1000 frames x 1 spoke x 512 samples, K = 5 temporal basis (SVD of a T2
dictionary), ComplexF32 data.  A solve is the backprojection (right-hand side)
plus 25 CG iterations with the Toeplitz normal operator; the kernel, coil maps
and sorted trajectory stay resident on the device between solves.

Prints one JSON line with solves/s, p50/p99 latency, the plain
one-recon-at-a-time GPU path (the reference's GPU call sequence on this
library's kernels: one single-channel type-1 NUFFT per coefficient and coil for
the backprojection, per-coil applies that re-read the kernel,
IterativeSolvers-style host-synchronised CG with ComplexF32 vectors) timed
beside it, the HBM roofline of the Toeplitz apply, the CPU
baseline (FP32 port of the reference CPU path) and the relative error of the
solve against an FP64 oracle.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mrfingerprintingrecon.jl_amd")

N_IMG = (256, 256)
COILS = 20
K = 5
NT = 1000
NR = 512
CG_ITERS = 25
PRECISION = "f64c"         # CG precision of the measured solve (FP64 vectors and apply arithmetic, 8-byte apply intermediates)
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "mrisubspacerecon_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(msr, seed=1002):
    import numpy as np
    synth = msr.synth
    trj = synth.traj_2d_radial_goldenratio(NR, 1, NT)              # (Nt, Nsamp, 2)
    U = synth.t2_basis(NT, K, seed=seed)                           # (K, Nt) complex64
    cm = synth.gaussian_coil_maps(N_IMG, COILS, seed=seed)         # (C, Ny, Nx)
    data = synth.gaussian_blob_kspace(trj, None, N_IMG, U, COILS, seed=seed, n_blobs=2)
    return trj, U, cm, np.ascontiguousarray(data)


def apply_bytes(nkm, lam_real, precision="f32"):
    """Bytes of one fused apply: s_v*2VK + s_t*8CKV + 8*2CV + Nkm*P*s_lam (b
    read and x written once, coil maps read twice, the x-FFT and y-FFT
    intermediates written and read once each, the kernel read once on its
    mask).  precision "f32" gives the algorithmic bytes of the roofline
    (ComplexF32 vectors and intermediates, s_v = s_t = 8); "f64" the bytes the
    ComplexF64 apply stores (s_v = s_t = 16); "f64c" those of the apply with
    compact intermediates (s_v = 16, s_t = 8).  Coil maps and kernel are
    ComplexF32 in every case."""
    V = N_IMG[0] * N_IMG[1]
    P = K * (K + 1) // 2
    s_lam = 4 if lam_real else 8
    s_v = 8 if precision == "f32" else 16
    s_t = 16 if precision == "f64" else 8
    return s_v * 2 * V * K + s_t * 8 * COILS * K * V + 8 * 2 * COILS * V + nkm * P * s_lam


def read_traffic(key):
    """HBM bytes per apply from the committed rocprofv3 counter summary, if any
    (line '<key> = <int>', key HBM_BYTES_PER_APPLY_F64, _F64C or _F32)."""
    p = os.path.join(ROOT, "profiles", "apply2d_256_c20_k5_counters.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.strip().startswith(key + " "):
            try:
                return int(line.split("=")[1].strip().split()[0])
            except ValueError:
                return None
    return None


def percentile(v, q):
    s = sorted(v)
    if not s:
        return float("nan")
    i = min(len(s) - 1, max(0, int(math.ceil(q / 100.0 * len(s))) - 1))
    return s[i]


def main():
    global PRECISION
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 check")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end FP64 pipeline check")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (64 recons) leg")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--volume", action="store_true",
                    help="3D case instead: 128^3 image, 16 coils, K=4, koosh-ball trajectory "
                         "(2000 frames x 4 spokes x 256 samples), backprojection + 15-iteration "
                         "CG; prints its own JSON line")
    ap.add_argument("--rows", action="store_true",
                    help="the operators beside the solve instead (forward model, Cartesian "
                         "operator, coil-map calibration) at the same size; prints its own "
                         "JSON line")
    ap.add_argument("--precision", default=PRECISION, choices=["f64", "f64c"],
                    help="CG precision of the measured solve (FP64 vectors and apply; f64c "
                         "stores the apply's FFT intermediates in 8 bytes)")
    ap.add_argument("--tuning", metavar="FILE",
                    help="tuned apply choices: load FILE if it exists (no tuning runs), else "
                         "tune and write it; without the option the benchmark tunes in-process")
    ap.add_argument("--apply-only", type=int, default=0, metavar="N",
                    help="profiling mode: build the operator, run N fused applies in each "
                         "precision (ComplexF64, 8-byte and ComplexF32 intermediates) and a "
                         "256 MiB calibration copy, then exit")
    args = ap.parse_args()
    PRECISION = args.precision
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU: this build has no multi-GPU path "
                 "(a reconstruction shards by coil or by job; see DESIGN.md)")

    import numpy as np
    import torch
    msr = load_pkg()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.volume:
        volume_leg(msr, torch, dev, args)
        return

    log("generating inputs")
    trj, U, cm, data = make_inputs(msr)
    trj_d = torch.from_numpy(trj).to(dev)
    U_d = torch.from_numpy(U).to(dev)
    cml = [torch.from_numpy(c).to(dev) for c in cm]
    data_d = torch.from_numpy(data).to(dev)

    loaded = 0
    if args.tuning and os.path.exists(args.tuning):
        loaded = msr.load_tuning(args.tuning)
        log(f"tuning: {loaded} entries from {args.tuning}")
    log("trajectory plan + kernel build")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    traj = msr.plan_trajectory(trj_d)
    A = msr.NFFTNormalOp(N_IMG, traj, U_d, cmaps=cml)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    # measured launch choices of the apply (the counterpart of the reference's
    # FFTW.MEASURE planning at operator construction), outside the timed solves
    t0 = time.perf_counter()
    tuning = {}
    for p in ("f64c", "f64", "f32"):
        tuning[p] = A.A.tune_(p, nrep=0 if loaded else 5)["chosen"]
    tune_ms = (time.perf_counter() - t0) * 1e3
    if args.tuning and not loaded:
        msr.save_tuning(args.tuning)
    log(f"tuning ({'loaded' if loaded else f'{tune_ms:.0f} ms'}): {tuning}")
    nvec = K * N_IMG[0] * N_IMG[1]
    if args.rows:
        rows_leg(msr, torch, dev, trj_d, traj, U_d, cml, data_d)
        return
    if args.apply_only:
        xin = torch.randn(nvec, dtype=torch.complex64, device=dev)
        y = torch.empty_like(xin)
        x64 = xin.to(torch.complex128)
        y64 = torch.empty_like(x64)
        for _ in range(args.apply_only):
            A.A.mul_f64_(y64, x64)
        for _ in range(args.apply_only):
            A.A.mul_f64_(y64, x64, compact=True)
        for _ in range(args.apply_only):
            A.A.mul_(y, xin)
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"apply_only": args.apply_only, "copy_bytes": 2 * src.numel() * 4,
                          **{f"apply_bytes_{p}": apply_bytes(A.A.nkm, A.A.lambda_real, p)
                             for p in ("f64", "f64c", "f32")},
                          "nkm": A.A.nkm}), flush=True)
        return
    b = torch.empty((K,) + N_IMG[::-1], dtype=torch.complex64, device=dev)
    x = torch.empty(nvec, dtype=torch.complex64, device=dev)
    bflat = b.reshape(-1)

    def solve():
        msr.calculate_backprojection(data_d, traj, cml, U=U_d, out=b)
        msr.cg_(x, A, bflat, maxiter=CG_ITERS, reltol=0.0, initially_zero=True, mode="graph",
                precision=PRECISION)

    def solve_plain():
        msr.calculate_backprojection(data_d, traj, cml, U=U_d, out=b, channelwise=True)
        msr.cg_(x, A, bflat, maxiter=CG_ITERS, reltol=0.0, initially_zero=True, mode="host",
                precision="f32")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for s, e in evs:
            s.record()
            fn()
            e.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [s.elapsed_time(e) for s, e in evs]
        return wall, lat

    log(f"timing {args.steps} solves (warmup {args.warmup})")
    wall, lat = timed(solve, args.steps, args.warmup)
    solves_per_s = args.steps / wall
    log(f"fused/graph: {1e3 * wall / args.steps:.3f} ms/solve")
    x_gpu = x.cpu().numpy()
    b_gpu = b.cpu().numpy()

    log(f"timing {args.plain_steps} plain solves")
    pwall, plat = timed(solve_plain, args.plain_steps, 1)
    plain_sps = args.plain_steps / pwall
    log(f"plain: {1e3 * pwall / args.plain_steps:.3f} ms/solve")

    # ---- roofline of the Toeplitz apply (the solve's dominant work), timed
    # with events on the stream the applies are launched on (the current one)
    xin = torch.randn(nvec, dtype=torch.complex64, device=dev)
    s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def time_apply(fn, napp=50):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        s_ev.record()
        for _ in range(napp):
            fn()
        e_ev.record()
        torch.cuda.synchronize()
        return s_ev.elapsed_time(e_ev) / napp

    names = {"f64": "Toeplitz apply, ComplexF64 arithmetic and intermediates (3 fused passes "
                    "+ epilogue)",
             "f64c": "Toeplitz apply, ComplexF64 arithmetic, intermediates in 8 bytes (3 fused "
                     "passes + epilogue)",
             "f32": "Toeplitz apply, ComplexF32 (3 fused passes + epilogue)"}

    def roofline_of(t_ms, precision):
        # achieved = the algorithmic bytes of the apply (ComplexF32 vectors and
        # intermediates, kernel on its kmask) per launch time; the bytes the
        # precision actually stores are reported beside it
        B = apply_bytes(A.A.nkm, A.A.lambda_real, "f32")
        achieved = B / (t_ms * 1e-3) / 1e9
        Bs = apply_bytes(A.A.nkm, A.A.lambda_real, precision)
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic("HBM_BYTES_PER_APPLY_" + precision.upper()),
                "kernel": names[precision], "bytes_per_launch": B, "ms_per_launch": round(t_ms, 4),
                "stored_bytes_per_launch": Bs,
                "stored_frac": round(Bs / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    x64 = xin.to(torch.complex128)
    y64 = torch.empty_like(x64)
    y = torch.empty_like(xin)
    t_app = {"f64": time_apply(lambda: A.A.mul_f64_(y64, x64)),
             "f64c": time_apply(lambda: A.A.mul_f64_(y64, x64, compact=True)),
             "f32": time_apply(lambda: A.A.mul_(y, xin))}
    roofline = roofline_of(t_app[PRECISION], PRECISION)
    roofline_other = {p: roofline_of(t, p) for p, t in t_app.items() if p != PRECISION}
    # plain (per-coil) apply for comparison
    s_ev.record()
    for _ in range(5):
        A.A.mul_(y, xin, percoil=True)
    e_ev.record()
    torch.cuda.synchronize()
    t_apply_plain = s_ev.elapsed_time(e_ev) / 5

    result = {
        "metric": "CG recon solves/s + p50/p99 latency vs plain GPU path @<=1e-4 rel err vs FP64",
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 apply and CG vectors (complex128 FFTs / kernel multiply / intermediates); "
                  "f32 kernel, coil maps, gridding and backprojection") if PRECISION == "f64" else
                 ("f64 apply arithmetic and CG vectors, apply intermediates stored in 8 bytes "
                  "(shared exponent, 28-bit mantissas); f32 kernel, coil maps, gridding and "
                  "backprojection"),
        "data": "synthetic",
        "config": {"workload": "single 2D recon 256x256 (grid 512x512), 20 coils, 1000 frames x "
                               "1 golden-angle spoke x 512 samples, K=5, BP + 25-iteration CG",
                   "img": list(N_IMG), "coils": COILS, "K": K, "frames": NT,
                   "samples_per_frame": NR, "cg_iters": CG_ITERS},
        "latency_ms": {"p50": round(percentile(lat, 50), 4), "p99": round(percentile(lat, 99), 4)},
        "plain_gpu": {"value": round(plain_sps, 3), "unit": "solves/s",
                      "path": "reference GPU call sequence: K*C single-channel NUFFT backprojection, "
                              "per-coil applies re-reading the kernel, host-synchronised ComplexF32 CG",
                      "ms_per_solve": round(1e3 * pwall / args.plain_steps, 3),
                      "p50_ms": round(percentile(plat, 50), 3),
                      "speedup": round(solves_per_s / plain_sps, 3),
                      "apply_ms": round(t_apply_plain, 4)},
        "kernel_build_ms": round(build_ms, 1),
        "tuning": {"chosen_Gb_orderA_orderC": tuning, "source": "file" if loaded else "measured",
                   "ms": None if loaded else round(tune_ms, 1)},
        "roofline": roofline,
        "roofline_other_applies": roofline_other,
        "cpu_baseline": None,
        "rel_err_vs_fp64": None,
    }

    result["kernel_cache"] = cache_leg(msr, torch, traj, U_d, cml)
    if not args.no_batched:
        result["batched"] = batched_leg(msr)
    if not args.no_served:
        result["served"] = served_leg(msr)
    if not args.no_cpu:
        result["cpu_baseline"], result["rel_err_vs_fp64"] = cpu_leg(
            msr, trj, U, cm, data, A, x_gpu, b_gpu, e2e=not args.no_e2e)
    print(json.dumps(result), flush=True)


def volume_leg(msr, torch, dev, args, N=128, C=16, K3=4, nt=2000, spokes=4, iters=15):
    """3D workload (the reference's koosh-ball case at 128^3): trajectory plan,
    kernel build, then `steps` solves (backprojection + FP64-apply graph CG)."""
    trj = msr.synth.traj_kooshball_goldenratio(2 * N, spokes, nt).astype("float32")
    U = msr.synth.t2_basis(nt, K3, seed=5)
    cm = msr.synth.gaussian_coil_maps((N, N, N), C, seed=5)
    trj_d = torch.from_numpy(trj).to(dev)
    U_d = torch.from_numpy(U).to(dev)
    cml = [torch.from_numpy(c).to(dev) for c in cm]
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    data = torch.randn((C,) + tuple(trj.shape[:2]), dtype=torch.complex64, device=dev, generator=g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    traj = msr.plan_trajectory(trj_d)
    A = msr.NFFTNormalOp((N, N, N), traj, U_d, cmaps=cml)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    b = torch.empty((K3, N, N, N), dtype=torch.complex64, device=dev)
    x = torch.empty(b.numel(), dtype=torch.complex64, device=dev)

    def solve():
        msr.calculate_backprojection(data, traj, cml, U=U_d, out=b)
        msr.cg_(x, A, b.reshape(-1), maxiter=iters, reltol=0.0, initially_zero=True,
                mode="graph", precision=PRECISION)

    steps, warm = max(1, min(args.steps, 5)), max(1, min(args.warmup, 2))
    for _ in range(warm):
        solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solve()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    print(json.dumps({"workload": f"3D recon {N}^3 (grid {2 * N}^3), {C} coils, K={K3}, koosh-ball "
                                  f"{nt} frames x {spokes} spokes x {2 * N} samples, BP + {iters}-"
                                  f"iteration CG (FP64 arithmetic, {'8-byte' if PRECISION == 'f64c' else 'ComplexF64'} "
                                  f"intermediates)", "precision": PRECISION,
                                  "solves_per_s": round(1.0 / wall, 3),
                      "s_per_solve": round(wall, 4), "steps": steps,
                      "kernel_build_s": round(build_s, 3),
                      "device_state_bytes": list(msr.footprint(A, PRECISION))}), flush=True)


def rows_leg(msr, torch, dev, trj_d, traj, U_d, cml, data_d, lines=8, calib_iters=100, reps=20):
    """The operators either side of the solve, at the same 256x256 / 20-coil /
    K=5 / 1000-frame size: the forward model A x, the Cartesian operator
    (8 random phase-encode lines x 256 readout samples per frame: kernel
    build, applies, backprojection + 25-iteration FP64 CG) and the
    calibration front end of calculate_coil_maps (calibration region,
    coil-wise K=1 reconstruction, calibration k-space).  Wall-clock times of
    `reps` calls between device synchronisations."""
    import numpy as np
    nx, ny = N_IMG
    nvec = K * nx * ny

    def wall(fn, n=reps, warm=2):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    g = torch.Generator(device=dev)
    g.manual_seed(11)
    x = torch.randn((K, ny, nx), dtype=torch.complex64, device=dev, generator=g)
    y_fwd = torch.empty((COILS, NT, NR), dtype=torch.complex64, device=dev)
    fwd_ms = wall(lambda: msr.forward(x, traj, cmaps=cml, U=U_d, out=y_fwd))
    fwd = {"ms": round(fwd_ms, 4), "samples_per_s": round(COILS * NT * NR / fwd_ms * 1e3, 1)}

    rng = np.random.default_rng(11)
    pe = np.stack([rng.choice(ny, lines, replace=False) for _ in range(NT)]) + 1  # 1-based
    trj_c = np.empty((NT, lines * nx, 2), dtype=np.int32)
    trj_c[:, :, 0] = np.tile(np.arange(1, nx + 1), lines)
    trj_c[:, :, 1] = np.repeat(pe, nx, axis=1)
    trj_cd = torch.from_numpy(trj_c).to(dev)
    dc = torch.randn((COILS, NT, lines * nx), dtype=torch.complex64, device=dev, generator=g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    traj_c = msr.plan_trajectory(trj_cd)
    Ac = msr.FFTNormalOp(N_IMG, traj_c, U_d, cmaps=cml)
    torch.cuda.synchronize()
    cart_build = (time.perf_counter() - t0) * 1e3
    v32 = torch.randn(nvec, dtype=torch.complex64, device=dev, generator=g)
    w32 = torch.empty_like(v32)
    v64 = v32.to(torch.complex128)
    w64 = torch.empty_like(v64)
    a64 = wall(lambda: Ac.A.mul_f64_(w64, v64), 50)
    a32 = wall(lambda: Ac.A.mul_(w32, v32), 50)
    bc = torch.empty((K, ny, nx), dtype=torch.complex64, device=dev)
    xc = torch.empty(nvec, dtype=torch.complex64, device=dev)
    bp_ms = wall(lambda: msr.calculate_backprojection(dc, traj_c, cml, U=U_d, out=bc))

    def solve():
        msr.calculate_backprojection(dc, traj_c, cml, U=U_d, out=bc)
        msr.cg_(xc, Ac, bc.reshape(-1), maxiter=CG_ITERS, reltol=0.0, initially_zero=True,
                mode="graph", precision=PRECISION)

    solve_ms = wall(solve, 10)
    cart = {"kernel_build_ms": round(cart_build, 3), "apply_f64_ms": round(a64, 4),
            "apply_f32_ms": round(a32, 4), "backprojection_ms": round(bp_ms, 4),
            "solve_ms": round(solve_ms, 4), "solves_per_s": round(1e3 / solve_ms, 2),
            "samples": int(NT * lines * nx)}

    calib = msr.default_calib_size(N_IMG)
    out = {}

    def calibrate():
        tc, mc = msr.calibration_region(trj_d, N_IMG, calib)
        xcoil = msr.reconstruct_coilwise(data_d, tc, calib, sample_mask=mc, Niter_cg=calib_iters)
        out["kbp"] = msr.calibration_kspace(xcoil)

    cal_ms = wall(calibrate, 3, 1)
    print(json.dumps({"workload": f"operators beside the solve at {nx}x{ny}, {COILS} coils, K={K}, "
                                  f"{NT} frames", "forward": fwd,
                      "cartesian": dict(cart, workload=f"{lines} phase-encode lines x {nx} samples "
                                                       f"per frame, BP + {CG_ITERS}-iteration "
                                                       f"{PRECISION.upper()} CG"),
                      "calibration": {"calib_size": list(calib), "cg_iters": calib_iters,
                                      "ms": round(cal_ms, 3),
                                      "workload": "calibration region + coil-wise K=1 "
                                                  "reconstruction + calibration k-space "
                                                  "(ESPIRiT itself excluded)"}}), flush=True)


def cache_leg(msr, torch, traj, U_d, cml):
    """Operator set-up for a protocol seen before: kernel build (miss) against
    a device-memory hit and a kernel-file read (new cache on the same
    directory), for the benchmark's protocol."""
    import tempfile
    out = {}
    with tempfile.TemporaryDirectory() as d:
        def timed_get(cache):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            op = cache.get(N_IMG, traj, U_d, cmaps=cml)
            torch.cuda.synchronize()
            return op, (time.perf_counter() - t0) * 1e3
        c1 = msr.KernelCache(directory=d)
        op, out["build_ms"] = timed_get(c1)
        src = [c1.last_source]
        del op
        op, out["memory_hit_ms"] = timed_get(c1)
        src.append(c1.last_source)
        del op
        c2 = msr.KernelCache(directory=d)
        op, out["file_read_ms"] = timed_get(c2)
        src.append(c2.last_source)
        out["file_bytes"] = os.path.getsize(os.path.join(d, c1.last_key + ".msrk"))
        del op, c1, c2
    out = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in out.items()}
    out["sources"] = src
    log(f"kernel cache: {out}")
    return out


def _job_inputs(msr, torch, dev, rng, N, C, K, spokes, nt, seed):
    """Synthetic 2D job: golden-angle radial trajectory (2N samples per spoke),
    T2 basis, Gaussian coil maps, complex Gaussian k-space data."""
    trj = msr.synth.traj_2d_radial_goldenratio(2 * N, spokes, nt)
    U = msr.synth.t2_basis(nt, K, seed=seed)
    cm = msr.synth.gaussian_coil_maps((N, N), C, seed=seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    data = torch.randn((C, nt, spokes * 2 * N), dtype=torch.complex64, device=dev, generator=g)
    return (torch.from_numpy(trj).to(dev), torch.from_numpy(U).to(dev),
            [torch.from_numpy(c).to(dev) for c in cm], data)


def batched_leg(msr, njobs=64, iters=CG_ITERS, seed=1004):
    """Batched 2D reconstructions (64 jobs, N in {128,160,192,256}, 8-32 coils,
    K 3-6, 1-4 spokes per frame, 500-1000 frames): backprojections, then one
    lockstep CG over all jobs (graph), against the one-at-a-time loop (each
    job's backprojection and single-operator graph CG in turn)."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    jobs = []
    log(f"batched: building {njobs} jobs")
    tuned = set()
    for i in range(njobs):
        N = int(rng.choice([128, 160, 192, 256]))
        C = int(rng.integers(8, 33))
        K = int(rng.integers(3, 7))
        spokes = int(rng.integers(1, 5))
        nt = int(rng.integers(500, 1001))
        trj, U, cml, data = _job_inputs(msr, torch, dev, rng, N, C, K, spokes, nt, seed * 100 + i)
        traj = msr.plan_trajectory(trj)
        A = msr.NFFTNormalOp((N, N), traj, U, cmaps=cml)
        if (N, C, K) not in tuned:  # later operators of the shape start with the choice
            A.A.tune_(PRECISION, nrep=3)
            tuned.add((N, C, K))
        b = torch.empty((K, N, N), dtype=torch.complex64, device=dev)
        jobs.append(dict(N=N, C=C, K=K, M=traj.M, traj=traj, U=U, cml=cml, data=data, A=A, b=b,
                         x1=torch.empty(K * N * N, dtype=torch.complex64, device=dev),
                         xb=torch.empty(K * N * N, dtype=torch.complex64, device=dev)))
    B = msr.NFFTNormalOpBatch([j["A"] for j in jobs])

    def one_at_a_time():
        for j in jobs:
            msr.calculate_backprojection(j["data"], j["traj"], j["cml"], U=j["U"], out=j["b"])
            msr.cg_(j["x1"], j["A"], j["b"].reshape(-1), maxiter=iters, reltol=0.0,
                    initially_zero=True, mode="graph", precision=PRECISION)

    def batched():
        for j in jobs:
            msr.calculate_backprojection(j["data"], j["traj"], j["cml"], U=j["U"], out=j["b"])
        B.cg_([j["xb"] for j in jobs], [j["b"].reshape(-1) for j in jobs], maxiter=iters,
              reltol=0.0, initially_zero=True, mode="graph", precision=PRECISION)

    out = {}
    for name, fn in (("one_at_a_time", one_at_a_time), ("batched", batched)):
        fn()  # warm-up: graph captures
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out[name] = time.perf_counter() - t0
    err = max(float(torch.linalg.norm(j["xb"] - j["x1"]) / torch.linalg.norm(j["x1"])) for j in jobs)
    res = {"workload": f"{njobs} 2D recons: N in {{128,160,192,256}}, 8-32 coils, K 3-6, 1-4 spokes x "
                       f"500-1000 frames (radial golden angle), BP + {iters}-iteration CG each",
           "solves_per_s": round(njobs / out["batched"], 2),
           "one_at_a_time_solves_per_s": round(njobs / out["one_at_a_time"], 2),
           "speedup": round(out["one_at_a_time"] / out["batched"], 3),
           "samples_total": int(sum(j["M"] for j in jobs)),
           "max_rel_diff_batched_vs_single": err}
    log(f"batched: {res}")
    del B, jobs
    torch.cuda.synchronize()
    return res


SERVER_CONFIGS = (  # (policy, max jobs in flight, CG iterations per server step)
    ("fifo", 64, 5), ("shortest_first", 1, CG_ITERS), ("shortest_first", 2, 5),
    ("shortest_first", 4, 5), ("shortest_first", 8, 5))


def served_leg(msr, njobs=80, load=1.25, iters=CG_ITERS, iters3d=15, seed=1005, nproto=5,
               configs=SERVER_CONFIGS, frac3d=0.15, n3d=96):
    """Served stream at saturating load: Poisson arrivals of jobs drawn over
    resident protocols whose kernels stay on the device: 2D protocols as in
    the batched leg (N in {128,160,192,256}, K 3-6) and a 3D koosh-ball
    protocol at 96^3 (K 4, 4 spokes x 500 frames) that draws frac3d of the
    jobs.  Each job brings its own coil maps (8-32 coils in 2D, 12-16 in 3D;
    an operator sharing its protocol's kernel) and data.  The arrival rate is
    `load` times the one-at-a-time capacity measured on the same jobs (the
    mean of their backprojection + graph CG service times), so the queue
    grows under the one-at-a-time baseline.  On arrival a job's
    backprojection runs and its CG (25 iterations in 2D, 15 in 3D) is
    submitted to the server (continuous batching, admission by device
    memory); latency is arrival to solution.  The same arrivals served one at
    a time (FIFO) are the baseline; the server also runs them under
    shortest-work-first admission with one job in flight."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    log(f"served: {nproto} 2D protocols + one 3D protocol at {n3d}^3, {njobs} jobs")
    protos = []
    for p in range(nproto):
        N = int(rng.choice([128, 160, 192, 256]))
        K = int(rng.integers(3, 7))
        spokes = int(rng.integers(1, 5))
        nt = int(rng.integers(500, 1001))
        trj, U, cml, _ = _job_inputs(msr, torch, dev, rng, N, 8, K, spokes, nt, seed * 100 + p)
        traj = msr.plan_trajectory(trj)
        A = msr.NFFTNormalOp((N, N), traj, U, cmaps=cml)
        protos.append(dict(shape=(N, N), K=K, nsamp=spokes * 2 * N, nt=nt, traj=traj, U=U, A=A,
                           iters=iters, coils=(8, 33)))
    nt3, sp3, K3 = 500, 4, 4
    trj3 = msr.synth.traj_kooshball_goldenratio(2 * n3d, sp3, nt3).astype("float32")
    U3 = torch.from_numpy(msr.synth.t2_basis(nt3, K3, seed=seed)).to(dev)
    traj3 = msr.plan_trajectory(torch.from_numpy(trj3).to(dev))
    cm3 = [torch.from_numpy(c).to(dev) for c in msr.synth.gaussian_coil_maps((n3d,) * 3, 2, seed=seed)]
    A3 = msr.NFFTNormalOp((n3d,) * 3, traj3, U3, cmaps=cm3)
    p3 = dict(shape=(n3d,) * 3, K=K3, nsamp=trj3.shape[1], nt=nt3, traj=traj3, U=U3, A=A3,
              iters=iters3d, coils=(12, 17))
    jobs = []
    for i in range(njobs):
        pr = p3 if rng.uniform() < frac3d else protos[int(rng.integers(0, nproto))]
        C = int(rng.integers(*pr["coils"]))
        shape, K = pr["shape"], pr["K"]
        cml = [torch.from_numpy(c).to(dev) for c in msr.synth.gaussian_coil_maps(shape, C, seed=i)]
        g = torch.Generator(device=dev)
        g.manual_seed(seed * 1000 + i)
        data = torch.randn((C, pr["nt"], pr["nsamp"]), dtype=torch.complex64, device=dev,
                           generator=g)
        A = msr.NFFTNormalOp_shared(pr["A"], cmaps=cml)
        n = K * int(np.prod(shape))
        jobs.append(dict(pr=pr, cml=cml, data=data, A=A, three_d=len(shape) == 3,
                         b=torch.empty((K,) + shape[::-1], dtype=torch.complex64, device=dev),
                         x=torch.empty(n, dtype=torch.complex64, device=dev),
                         x1=torch.empty(n, dtype=torch.complex64, device=dev)))
    def bp(j):
        msr.calculate_backprojection(j["data"], j["pr"]["traj"], j["cml"], U=j["pr"]["U"], out=j["b"])

    def solve_alone(j):
        bp(j)
        msr.cg_(j["x1"], j["A"], j["b"].reshape(-1), maxiter=j["pr"]["iters"], reltol=0.0,
                initially_zero=True, mode="graph", precision=PRECISION)

    # warm-up (graph captures of the single-operator baseline, first launches),
    # then each job's one-at-a-time service time
    for j in jobs:
        solve_alone(j)
    torch.cuda.synchronize()
    service = []
    for j in jobs:
        t0 = time.perf_counter()
        solve_alone(j)
        torch.cuda.synchronize()
        service.append(time.perf_counter() - t0)
    capacity = len(jobs) / sum(service)
    rate = load * capacity
    t = 0.0
    for j in jobs:
        t += float(rng.exponential(1.0 / rate))
        j["t"] = t

    def run_fifo():
        lat = []
        t0 = time.perf_counter()
        for j in jobs:
            while time.perf_counter() - t0 < j["t"]:
                pass
            solve_alone(j)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0 - j["t"])
        return time.perf_counter() - t0, lat

    def run_server(policy, max_active, niter):
        srv = msr.NFFTServer(max_active=max_active, precision=PRECISION, policy=policy)
        lat = [None] * len(jobs)
        ids = {}
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(jobs) or ids:
            now = time.perf_counter() - t0
            while nxt < len(jobs) and jobs[nxt]["t"] <= now:
                j = jobs[nxt]
                bp(j)
                ids[srv.submit(j["A"], j["x"], j["b"].reshape(-1), maxiter=j["pr"]["iters"])] = nxt
                nxt += 1
            if ids:
                srv.step(niter)
                tdone = time.perf_counter() - t0
                for jid in [k for k in ids if srv.poll(k)["state"] == srv.DONE]:
                    i = ids.pop(jid)
                    lat[i] = tdone - jobs[i]["t"]
                    srv.forget(jid)
        return time.perf_counter() - t0, lat

    t_fifo, lat_fifo = run_fifo()
    ms = lambda v: round(1e3 * v, 3)
    n3 = sum(j["three_d"] for j in jobs)
    lat2 = lambda lat: [v for v, j in zip(lat, jobs) if not j["three_d"]]
    summary = lambda t, lat: {"throughput_solves_per_s": round(njobs / t, 2),
                              "latency_ms": {"p50": ms(percentile(lat, 50)),
                                             "p99": ms(percentile(lat, 99)),
                                             "p50_2d": ms(percentile(lat2(lat), 50))}}
    runs = []
    err = 0.0
    for policy, max_active, niter in configs:
        t_srv, lat_srv = run_server(policy, max_active, niter)
        err = max(err, max(float(torch.linalg.norm(j["x"] - j["x1"]) / torch.linalg.norm(j["x1"]))
                           for j in jobs))
        runs.append({"policy": policy, "max_active": max_active, "iterations_per_step": niter,
                     **summary(t_srv, lat_srv)})
        log(f"served {policy}/{max_active}/{niter}: {runs[-1]}")
    res = {"workload": f"{njobs} recons ({njobs - n3} 2D, {n3} 3D at {n3d}^3), Poisson arrivals "
                       f"at {load} x the one-at-a-time capacity over {nproto} resident 2D "
                       f"protocols (N in {{128,160,192,256}}, K 3-6, 1-4 spokes x 500-1000 "
                       f"frames) and one 3D koosh-ball protocol (K 4, {sp3} spokes x {nt3} "
                       f"frames); per-job coil maps (2D 8-32, 3D 12-16 coils) and data; BP + "
                       f"CG ({iters} iterations 2D, {iters3d} 3D) each",
           "arrival_rate_per_s": round(rate, 2),
           "one_at_a_time_capacity_per_s": round(capacity, 2),
           "arrival_span_s": round(jobs[-1]["t"], 3),
           "one_at_a_time": summary(t_fifo, lat_fifo),
           "server": runs,
           "max_rel_diff_served_vs_single": err}
    log(f"served: {res}")
    del jobs, protos
    torch.cuda.synchronize()
    return res


def cpu_leg(msr, trj, U, cm, data, A, x_gpu, b_gpu, e2e=True, nrep=5):
    """CPU baseline (FP32 port of the reference CPU path, bounded sample) and
    the FP64 oracle check of the GPU solve (test infrastructure from oracle/,
    used here only as the baseline and the checker)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import cg as ocg
    from oracle import cpu_port, native, toeplitz
    native.build()
    cores = os.cpu_count()
    lam = A.A.Λ.cpu().numpy()
    km = A.A.kmask_indcs.cpu().numpy() - 1
    # --- CPU baseline: medians of nrep applies and backprojections, and whole
    # solves (backprojection + 25 FP32 CG iterations on the port's apply) timed
    log("cpu baseline: port apply")
    op = cpu_port.CpuNormalOp(N_IMG, lam, km, cm, workers=cores)
    v = np.asarray(b_gpu, dtype=np.complex64).reshape(-1)
    op.apply(v)
    t_apply = []
    for _ in range(nrep):
        t0 = time.perf_counter()
        op.apply(v)
        t_apply.append(time.perf_counter() - t0)
    log(f"cpu apply median {np.median(t_apply):.3f} s; port backprojection")
    t_bp = []
    for _ in range(nrep):
        t0 = time.perf_counter()
        cpu_port.backprojection_port(data, trj, cm, U, workers=cores)
        t_bp.append(time.perf_counter() - t0)
    log(f"cpu backprojection median {np.median(t_bp):.3f} s; whole solves")
    nsolve = 2
    t_solve = []
    x_cpu = None
    for _ in range(nsolve):
        t0 = time.perf_counter()
        bc = cpu_port.backprojection_port(data, trj, cm, U, workers=cores).reshape(-1)
        x_cpu = ocg.cg(op.apply, bc, maxiter=CG_ITERS, reltol=0.0)
        t_solve.append(time.perf_counter() - t0)
    ts = float(np.median(t_solve))
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                 if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    cpu = {"value": round(1.0 / ts, 5), "unit": "solves/s", "cores": cores, "kind": "port",
           "sample": f"{nsolve} whole solves of the same 256x256/20-coil/K=5 workload (port "
                     f"backprojection + 25 FP32 CG iterations on the port apply, median "
                     f"{ts:.2f} s, not extrapolated); beside them the medians of {nrep} applies "
                     f"({np.median(t_apply):.3f} s) and {nrep} backprojections "
                     f"({np.median(t_bp):.3f} s); {cores} threads on {model}",
           "apply_s": round(float(np.median(t_apply)), 4),
           "backprojection_s": round(float(np.median(t_bp)), 4),
           "solve_s": [round(t, 3) for t in t_solve]}
    # --- FP64 checks
    log("fp64 oracle: solve on the device-built operator")
    cm2 = cm.astype(np.complex128)
    Ad = toeplitz.NormalOp(N_IMG, lam.astype(np.complex128), km, cm2)
    ref_solve = ocg.cg(Ad, np.asarray(b_gpu, dtype=np.complex128).reshape(-1),
                       maxiter=CG_ITERS, reltol=0.0)
    err = {"solve_given_operator": float(np.linalg.norm(x_gpu - ref_solve) / np.linalg.norm(ref_solve))}
    if e2e:
        log("fp64 oracle: end-to-end pipeline (backprojection, kernel, CG)")
        b64 = toeplitz.backprojection(data, trj, cm, U=U)
        err["backprojection"] = float(np.linalg.norm(b_gpu - b64) / np.linalg.norm(b64))
        lam64, km64 = toeplitz.toeplitz_kernel(N_IMG, trj, U)
        err["kmask_identical"] = bool(np.array_equal(km64, km))
        err["kernel"] = float(np.linalg.norm(lam - lam64) / np.linalg.norm(lam64))
        A64 = toeplitz.NormalOp(N_IMG, lam64, km64, cm2)
        ref = ocg.cg(A64, b64.reshape(-1), maxiter=CG_ITERS, reltol=0.0)
        err["end_to_end"] = float(np.linalg.norm(x_gpu - ref) / np.linalg.norm(ref))
        # the CPU baseline's own result (FP32 port pipeline) against the same FP64 pipeline
        err["cpu_port_end_to_end"] = float(np.linalg.norm(x_cpu - ref) / np.linalg.norm(ref))
    log(f"errors {err}")
    return cpu, err


if __name__ == "__main__":
    main()
