"""GPU tests at the benchmark workloads, each through the C ABI against the
FP64 oracle or, where the oracle cannot afford the full size, through
size-independent properties plus an oracle comparison on a part it can:

  C2  single 2D recon 256^2 (grid 512^2), 20 coils, 1000 frames x 1 spoke x
      512 samples, K = 5: device backprojection + kernel build + 25-iteration
      FP64-apply CG against the FP64 pipeline (oracle backprojection, oracle
      kernel, oracle CG) end to end, <= 1e-4;
  C3  3D koosh-ball 128^3 (grid 256^3), 16 coils, 2000 frames x 4 spokes x 256
      samples, K = 4: kmask against the oracle's rule, the apply of a 2-coil
      operator on the full kernel and CG on it against the oracle, FP64 vs FP32
      apply, linearity, backprojection adjoint to the forward model,
      reproducible 15-iteration CG; and the same pipeline at 48^3 end to end
      against the oracle;
  C4  64 mixed 2D recons (N in {128,160,192,256}, 8-32 coils, K 3-6, 1-4
      spokes x 500-1000 frames) solved in one lockstep batch: every result
      bit-identical to its single-operator solve, three against the oracle;
  C5  a served stream with 2D jobs drawn as in C4 and 3D jobs at 96^3 (12-16
      coils, K = 4): every result bit-identical alone and in the mix.

Data are the device forward model of ellipse phantoms plus complex Gaussian
noise at 30 dB SNR (the recipe of test/reconstruct_radial.jl:62-74) or, for
C4/C5, complex Gaussian k-space; every result is compared in relative L2
norm (Julia isapprox form)."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import toeplitz

pytestmark = pytest.mark.gpu

E2E_TOL = 1e-4          # every result vs the FP64 oracle
KERNEL_TOL = 3e-7       # FP64-built kernel rounded once to ComplexF32
BP_TOL = 1e-5           # ComplexF32 backprojection
APPLY32_TOL = 2e-6      # one ComplexF32 apply
SOLVE_F64_TOL = 1e-7    # FP64 CG on the same operator (result rounded to ComplexF32)


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def msr():
    return load_package()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def simulate(torch, msr, x, traj, cml, U_d, seed, snr_db=30.0):
    """data = forward model of x plus complex Gaussian noise at snr_db."""
    y = msr.forward(dev(torch, x), traj, cmaps=cml, U=U_d)
    rms = torch.sqrt(torch.mean(torch.abs(y) ** 2))
    g = torch.Generator(device=y.device)
    g.manual_seed(seed)
    noise = torch.randn(y.shape, dtype=torch.complex64, device=y.device, generator=g)
    return y + (rms * 10 ** (-snr_db / 20)) * noise


def vdot(a, b):
    return complex(np.vdot(np.asarray(a, np.complex128).reshape(-1),
                           np.asarray(b, np.complex128).reshape(-1)))


# ------------------------------------------------------------------------ C2
@pytest.mark.timeout(900)
def test_c2_end_to_end_vs_fp64_pipeline(torch, msr):
    N, C, K, nt, nr, iters = (256, 256), 20, 5, 1000, 512, 25
    print(f"C2: {N} x {C} coils, K={K}, {nt} frames x {nr} samples, {iters} CG iterations",
          flush=True)
    synth = msr.synth
    trj = synth.traj_2d_radial_goldenratio(nr, 1, nt)
    U = synth.t2_basis(nt, K, seed=1002)
    cm = synth.gaussian_coil_maps(N, C, seed=1002)
    xt = synth.ellipse_phantom(N, K, seed=1002)
    cml = [dev(torch, c) for c in cm]
    U_d = dev(torch, U)
    traj = msr.plan_trajectory(dev(torch, trj))
    data = simulate(torch, msr, xt, traj, cml, U_d, seed=1002)
    A = msr.NFFTNormalOp(N, traj, U_d, cmaps=cml)
    b = msr.calculate_backprojection(data, traj, cml, U=U_d)
    x = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="graph").cpu().numpy()
    xc = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="graph",
                precision="f64c").cpu().numpy()
    lam = A.prod.A.Λ.cpu().numpy()
    km = A.prod.A.kmask_indcs.cpu().numpy() - 1
    b = b.cpu().numpy()
    data = data.cpu().numpy()
    print("C2: device done; FP64 oracle pipeline", flush=True)
    lam64, km64 = toeplitz.toeplitz_kernel(N, trj, U)
    assert np.array_equal(km, km64)
    e_lam = rel_err(lam, lam64)
    b64 = toeplitz.backprojection(data, trj, cm, U=U)
    e_b = rel_err(b, b64)
    cm2 = cm.astype(np.complex128)
    ref_given = ocg.cg(toeplitz.NormalOp(N, lam.astype(np.complex128), km, cm2),
                       b.astype(np.complex128).reshape(-1), maxiter=iters, reltol=0.0)
    e_given = rel_err(x, ref_given)
    ref = ocg.cg(toeplitz.NormalOp(N, lam64, km64, cm2), b64.reshape(-1), maxiter=iters,
                 reltol=0.0)
    e2e = rel_err(x, ref)
    e2e_c = rel_err(xc, ref)
    print(f"C2: kernel {e_lam:.2e}, backprojection {e_b:.2e}, solve on the device operator "
          f"{e_given:.2e} (compact intermediates {rel_err(xc, ref_given):.2e}), end to end "
          f"{e2e:.2e} (compact intermediates {e2e_c:.2e})", flush=True)
    assert e_lam < KERNEL_TOL
    assert e_b < BP_TOL
    assert e_given < SOLVE_F64_TOL
    assert e2e < E2E_TOL
    assert e2e_c < E2E_TOL


# ------------------------------------------------------------------------ C3
def kooshball_inputs(msr, N, C, K, nt, spokes, seed):
    synth = msr.synth
    trj = synth.traj_kooshball_goldenratio(2 * N[0], spokes, nt).astype(np.float32)
    U = synth.t2_basis(nt, K, seed=seed)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    xt = synth.ellipse_phantom(N, K, seed=seed)
    return trj, U, cm, xt


@pytest.mark.timeout(900)
def test_c3_full_size_properties_and_two_coil_oracle(torch, msr):
    N, C, K, nt, spokes, iters = (128, 128, 128), 16, 4, 2000, 4, 15
    print(f"C3: {N} x {C} coils, K={K}, koosh-ball {nt} x {spokes} x {2 * N[0]}", flush=True)
    trj, U, cm, xt = kooshball_inputs(msr, N, C, K, nt, spokes, seed=1003)
    cml = [dev(torch, c) for c in cm]
    U_d = dev(torch, U)
    traj = msr.plan_trajectory(dev(torch, trj))
    assert traj.M == nt * spokes * 2 * N[0]
    A = msr.NFFTNormalOp(N, traj, U_d, cmaps=cml)
    # kmask: the oracle's rule on the full 256^3 grid, bit-exact
    km = A.prod.A.kmask_indcs.cpu().numpy() - 1
    kc, _, _ = toeplitz.select_samples(trj)
    km_ref = toeplitz.kmask_indcs(tuple(2 * v for v in N), kc)
    print(f"C3: kmask {km.size} cells ({km.size / 256 ** 3:.3f} of the grid)", flush=True)
    assert np.array_equal(km, km_ref)
    lam = A.prod.A.Λ
    assert tuple(lam.shape) == (km.size, K * (K + 1) // 2)

    n = K * int(np.prod(N))
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    v = torch.randn(n, dtype=torch.complex64, device="cuda", generator=g)
    w = torch.randn(n, dtype=torch.complex64, device="cuda", generator=g)
    # FP64 apply against the FP32 apply, and linearity of the FP64 apply
    v64, w64 = v.to(torch.complex128), w.to(torch.complex128)
    y64 = torch.empty_like(v64)
    A.A.mul_f64_(y64, v64)
    y32 = A @ v
    e_32 = rel_err(y32.cpu().numpy(), y64.cpu().numpy())
    z64 = torch.empty_like(v64)
    A.A.mul_f64_(z64, w64)
    lin = torch.empty_like(v64)
    A.A.mul_f64_(lin, v64 + (0.5 - 2j) * w64)
    e_lin = rel_err(lin.cpu().numpy(), (y64 + (0.5 - 2j) * z64).cpu().numpy())
    print(f"C3: FP32 vs FP64 apply {e_32:.2e}, FP64 linearity {e_lin:.2e}", flush=True)
    assert e_32 < APPLY32_TOL
    assert e_lin < 1e-13

    # backprojection adjoint to the forward model: <BP(y), x> = <y, F x>
    data = simulate(torch, msr, xt, traj, cml, U_d, seed=1003)
    b = msr.calculate_backprojection(data, traj, cml, U=U_d)
    xr = v.reshape((K,) + N[::-1])
    Fx = msr.forward(xr, traj, cmaps=cml, U=U_d)
    lhs = complex(torch.vdot(xr.reshape(-1).to(torch.complex128),
                             b.reshape(-1).to(torch.complex128)))
    rhs = complex(torch.vdot(Fx.reshape(-1).to(torch.complex128),
                             data.reshape(-1).to(torch.complex128)))
    e_adj = abs(lhs - rhs) / abs(rhs)
    print(f"C3: <x, BP y> vs <F x, y> {e_adj:.2e}", flush=True)
    assert e_adj < 1e-5

    # 15-iteration FP64 CG: graph and device modes and a re-run agree bit for bit
    x1 = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="graph")
    x2 = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="device")
    x3 = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="graph")
    assert torch.equal(x1, x3) and torch.equal(x1, x2)
    assert torch.isfinite(torch.view_as_real(x1)).all()

    # the full kernel with two of the coils: apply and 3 CG iterations vs the oracle
    A2 = msr.NFFTNormalOp_shared(A, cmaps=cml[:2])
    y2 = (A2 @ v).cpu().numpy()
    lam_h = lam.cpu().numpy().astype(np.complex128)
    del A, y64, z64, lin, v64, w64, x1, x2, x3
    torch.cuda.empty_cache()
    cm2 = cm[:2].astype(np.complex128)
    op2 = toeplitz.NormalOp(N, lam_h, km, cm2)
    vh = v.cpu().numpy()
    e_a2 = rel_err(y2, op2(vh))
    b2 = (A2 @ w).cpu().numpy()   # any right-hand side
    xc = msr.cg(A2, dev(torch, b2), maxiter=3, reltol=0.0, mode="device").cpu().numpy()
    e_cg2 = rel_err(xc, ocg.cg(op2, b2.astype(np.complex128), maxiter=3, reltol=0.0))
    print(f"C3 (2 coils, full kernel): apply {e_a2:.2e}, 3-iteration CG {e_cg2:.2e}", flush=True)
    assert e_a2 < APPLY32_TOL
    assert e_cg2 < SOLVE_F64_TOL


@pytest.mark.timeout(600)
def test_c3_geometry_48_end_to_end_vs_oracle(torch, msr):
    """C3's pipeline (koosh-ball, K = 4, 15 FP64 CG iterations) at 48^3 with
    4 coils, 100 frames x 4 spokes x 96 samples, where the FP64 oracle runs the
    whole pipeline."""
    N, C, K, nt, spokes, iters = (48, 48, 48), 4, 4, 100, 4, 15
    trj, U, cm, xt = kooshball_inputs(msr, N, C, K, nt, spokes, seed=1013)
    cml = [dev(torch, c) for c in cm]
    U_d = dev(torch, U)
    traj = msr.plan_trajectory(dev(torch, trj))
    data = simulate(torch, msr, xt, traj, cml, U_d, seed=1013)
    A = msr.NFFTNormalOp(N, traj, U_d, cmaps=cml)
    b = msr.calculate_backprojection(data, traj, cml, U=U_d)
    x = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode="graph").cpu().numpy()
    lam = A.prod.A.Λ.cpu().numpy()
    km = A.prod.A.kmask_indcs.cpu().numpy() - 1
    lam64, km64 = toeplitz.toeplitz_kernel(N, trj, U)
    b64 = toeplitz.backprojection(data.cpu().numpy(), trj, cm, U=U)
    ref = ocg.cg(toeplitz.NormalOp(N, lam64, km64, cm.astype(np.complex128)), b64.reshape(-1),
                 maxiter=iters, reltol=0.0)
    e_lam, e_b, e2e = rel_err(lam, lam64), rel_err(b.cpu().numpy(), b64), rel_err(x, ref)
    print(f"C3 at 48^3: kernel {e_lam:.2e}, backprojection {e_b:.2e}, end to end {e2e:.2e}",
          flush=True)
    assert np.array_equal(km, km64)
    assert e_lam < KERNEL_TOL and e_b < BP_TOL and e2e < E2E_TOL


# ------------------------------------------------------------------------ C4
def mixed_2d_job(torch, msr, rng, seed):
    N = int(rng.choice([128, 160, 192, 256]))
    C = int(rng.integers(8, 33))
    K = int(rng.integers(3, 7))
    spokes = int(rng.integers(1, 5))
    nt = int(rng.integers(500, 1001))
    synth = msr.synth
    trj = synth.traj_2d_radial_goldenratio(2 * N, spokes, nt)
    U = synth.t2_basis(nt, K, seed=seed)
    cm = synth.gaussian_coil_maps((N, N), C, seed=seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    data = torch.randn((C, nt, spokes * 2 * N), dtype=torch.complex64, device="cuda",
                       generator=g)
    return dict(N=(N, N), C=C, K=K, trj=trj, U=U, cm=cm, data=data)


@pytest.mark.timeout(900)
def test_c4_batch_of_64_mixed_recons(torch, msr):
    rng = np.random.default_rng(1004)
    jobs = []
    for i in range(64):
        j = mixed_2d_job(torch, msr, rng, seed=1004 * 100 + i)
        j["cml"] = [dev(torch, c) for c in j["cm"]]
        j["U_d"] = dev(torch, j["U"])
        j["traj"] = msr.plan_trajectory(dev(torch, j["trj"]))
        j["A"] = msr.NFFTNormalOp(j["N"], j["traj"], j["U_d"], cmaps=j["cml"])
        j["b"] = msr.calculate_backprojection(j["data"], j["traj"], j["cml"], U=j["U_d"]).reshape(-1)
        jobs.append(j)
    Ns = sorted({j["N"][0] for j in jobs})
    print(f"C4: 64 jobs, N {Ns}, coils {min(j['C'] for j in jobs)}-"
          f"{max(j['C'] for j in jobs)}, K {min(j['K'] for j in jobs)}-"
          f"{max(j['K'] for j in jobs)}", flush=True)
    B = msr.NFFTNormalOpBatch([j["A"] for j in jobs])
    xs = [torch.empty_like(j["b"]) for j in jobs]
    B.cg_(xs, [j["b"] for j in jobs], maxiter=25, reltol=0.0, initially_zero=True, mode="graph")
    same = 0
    for j, xb in zip(jobs, xs):
        x1 = msr.cg(j["A"], j["b"], maxiter=25, reltol=0.0, mode="graph")
        same += int(torch.equal(xb, x1))
    print(f"C4: {same}/64 batched results bit-identical to their single-operator solves",
          flush=True)
    assert same == 64
    # three of the smallest jobs against the FP64 oracle: the solve on the
    # device operator, and one end to end
    order = sorted(range(64), key=lambda i: jobs[i]["C"] * jobs[i]["K"] * jobs[i]["N"][0] ** 2)
    for rank, i in enumerate(order[:3]):
        j = jobs[i]
        lam = j["A"].prod.A.Λ.cpu().numpy().astype(np.complex128)
        km = j["A"].prod.A.kmask_indcs.cpu().numpy() - 1
        cm2 = j["cm"].astype(np.complex128)
        bh = j["b"].cpu().numpy()
        ref = ocg.cg(toeplitz.NormalOp(j["N"], lam, km, cm2), bh.astype(np.complex128),
                     maxiter=25, reltol=0.0)
        e = rel_err(xs[i].cpu().numpy(), ref)
        msg = f"C4 job {i} (N={j['N'][0]}, C={j['C']}, K={j['K']}): solve on the device operator {e:.2e}"
        assert e < SOLVE_F64_TOL
        if rank == 0:
            b64 = toeplitz.backprojection(j["data"].cpu().numpy(), j["trj"], j["cm"], U=j["U"])
            lam64, km64 = toeplitz.toeplitz_kernel(j["N"], j["trj"], j["U"])
            assert np.array_equal(km64, km)
            ref = ocg.cg(toeplitz.NormalOp(j["N"], lam64, km64, cm2), b64.reshape(-1),
                         maxiter=25, reltol=0.0)
            e2e = rel_err(xs[i].cpu().numpy(), ref)
            msg += f", end to end {e2e:.2e}"
            assert e2e < E2E_TOL
        print(msg, flush=True)


# ------------------------------------------------------------------------ C5
@pytest.mark.timeout(900)
def test_c5_served_stream_with_3d_jobs(torch, msr):
    """Jobs join a running server (2D jobs drawn as in C4 on shared protocol
    kernels with per-job coil maps, and two 3D koosh-ball jobs at 96^3): each
    result is bit-identical to the same job served alone."""
    rng = np.random.default_rng(1005)
    protos = []
    for p in range(3):
        j = mixed_2d_job(torch, msr, rng, seed=1005 * 100 + p)
        traj = msr.plan_trajectory(dev(torch, j["trj"]))
        U_d = dev(torch, j["U"])
        protos.append(dict(j, traj=traj, U_d=U_d,
                           A=msr.NFFTNormalOp(j["N"], traj, U_d, cmaps=[dev(torch, c) for c in j["cm"]])))
    jobs = []
    for i in range(8):   # 2D jobs: a protocol's kernel, their own coil maps and data
        pr = protos[int(rng.integers(0, len(protos)))]
        C = int(rng.integers(8, 33))
        cml = [dev(torch, c) for c in msr.synth.gaussian_coil_maps(pr["N"], C, seed=i)]
        g = torch.Generator(device="cuda")
        g.manual_seed(5000 + i)
        data = torch.randn((C,) + tuple(pr["data"].shape[1:]), dtype=torch.complex64,
                           device="cuda", generator=g)
        A = msr.NFFTNormalOp_shared(pr["A"], cmaps=cml)
        b = msr.calculate_backprojection(data, pr["traj"], cml, U=pr["U_d"]).reshape(-1)
        jobs.append(dict(kind=f"2D {pr['N'][0]}^2 C={C} K={pr['K']}", A=A, b=b, iters=25))
    for i, C in enumerate((12, 16)):   # 3D jobs at 96^3, K = 4, ~1.15 M samples
        N = (96, 96, 96)
        trj, U, cm, xt = kooshball_inputs(msr, N, C, 4, 1500, 4, seed=7000 + i)
        cml = [dev(torch, c) for c in cm]
        U_d = dev(torch, U)
        traj = msr.plan_trajectory(dev(torch, trj))
        data = simulate(torch, msr, xt, traj, cml, U_d, seed=7000 + i)
        A = msr.NFFTNormalOp(N, traj, U_d, cmaps=cml)
        b = msr.calculate_backprojection(data, traj, cml, U=U_d).reshape(-1)
        jobs.insert(3 + 3 * i, dict(kind=f"3D 96^3 C={C} K=4", A=A, b=b, iters=15))

    def alone(j):
        srv = msr.NFFTServer(max_active=1)
        x = torch.empty_like(j["b"])
        jid = srv.submit(j["A"], x, j["b"], maxiter=j["iters"])
        srv.run(niter=5)
        assert srv.poll(jid)["iters"] == j["iters"]
        return x

    solo = [alone(j) for j in jobs]
    srv = msr.NFFTServer(max_active=16)
    xs = [torch.empty_like(j["b"]) for j in jobs]
    ids = [srv.submit(jobs[k]["A"], xs[k], jobs[k]["b"], maxiter=jobs[k]["iters"]) for k in (0, 1)]
    in_flight = []
    for k in range(2, len(jobs)):   # one arrival per step while the others run
        na, _ = srv.step(niter=3)
        in_flight.append(na)
        ids.append(srv.submit(jobs[k]["A"], xs[k], jobs[k]["b"], maxiter=jobs[k]["iters"]))
    srv.run(niter=3)
    print(f"C5: jobs in flight per step {in_flight}", flush=True)
    assert max(in_flight) >= 3
    for j, x, s, jid in zip(jobs, xs, solo, ids):
        info = srv.poll(jid)
        ok = bool(torch.equal(x, s))
        print(f"  {j['kind']}: {info['iters']} iterations, identical alone and in the mix {ok}",
              flush=True)
        assert info["state"] == srv.DONE and info["iters"] == j["iters"] and ok
