"""The reconstruction server (msr_server_*): continuous batching with admission
by device-memory footprint.  Each job's result is bit-identical alone and in
a mix of jobs that join and leave while others run (the batch engine runs
every operator's blocks exactly as it does alone), within 1e-6 of the
single-operator msr_cg, and the number of jobs in flight never exceeds what
the budget admits.  Operators sharing one kernel (msr_normalop_create_shared)
give the results of operators that own a copy of it."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from test_gpu_apply import make_case

pytestmark = pytest.mark.gpu

MIX = [  # (N, K, C, real, identity)
    ((32, 32), 3, 4, False, False),
    ((48, 32), 5, 3, False, False),
    ((32, 32), 3, 6, True, False),
    ((64, 40), 4, 2, False, True),
    ((16, 16, 16), 2, 3, False, False),
    ((32, 32), 6, 5, False, False),
]
ITERS = [10, 7, 12, 9, 6, 11]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def jobs(torch):
    msr = load_package()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(91)
    out = []
    for (N, K, C, real, ident), it in zip(MIX, ITERS):
        lam, km, cm, x = make_case(rng, N, K, C, real, ident)
        cml = None if cm is None else [torch.from_numpy(c).to(dev) for c in cm]
        A = msr.NFFTNormalOp(N, torch.from_numpy(lam).to(dev), torch.from_numpy(km + 1).to(dev),
                             cmaps=cml)
        out.append(dict(N=N, A=A, b=torch.from_numpy(x).to(dev).reshape(-1), iters=it,
                        lam=lam, km=km, cml=cml))
    return out


def run_alone(msr, torch, j):
    srv = msr.NFFTServer(max_active=1)
    x = torch.empty_like(j["b"])
    jid = srv.submit(j["A"], x, j["b"], maxiter=j["iters"])
    srv.run(niter=3)
    info = srv.poll(jid)
    assert info["state"] == srv.DONE and info["iters"] == j["iters"]
    return x.cpu().numpy()


def test_server_results_alone_and_in_a_mix_are_identical(torch, jobs):
    msr = load_package()
    alone = [run_alone(msr, torch, j) for j in jobs]
    srv = msr.NFFTServer(max_active=8)
    xs = [torch.empty_like(j["b"]) for j in jobs]
    ids = []
    # jobs join while others run: two, then one per step, staggered iteration counts
    for k in (0, 1):
        ids.append(srv.submit(jobs[k]["A"], xs[k], jobs[k]["b"], maxiter=jobs[k]["iters"]))
    active_seen = []
    for k in range(2, len(jobs)):
        na, nq = srv.step(niter=2)
        active_seen.append(na)
        ids.append(srv.submit(jobs[k]["A"], xs[k], jobs[k]["b"], maxiter=jobs[k]["iters"]))
    srv.run(niter=2)
    print(f"jobs in flight per step: {active_seen}", flush=True)
    assert max(active_seen) >= 2  # they did overlap
    for j, x, a, jid in zip(jobs, xs, alone, ids):
        info = srv.poll(jid)
        assert info["state"] == srv.DONE and info["iters"] == j["iters"]
        xm = x.cpu().numpy()
        single = msr.cg(j["A"], j["b"], maxiter=j["iters"], reltol=0.0, mode="device").cpu().numpy()
        print(f"  N={j['N']} it={j['iters']}: mix vs alone bit-exact {np.array_equal(xm, a)}, "
              f"vs msr_cg {rel_err(xm, single):.2e}, times {info['queued_ms']:.1f}/"
              f"{info['running_ms']:.1f} ms", flush=True)
        assert np.array_equal(xm, a)
        assert rel_err(xm, single) <= 1e-6
        srv.forget(jid)


def test_server_admission_respects_the_budget(torch, jobs):
    msr = load_package()
    fp = [msr.footprint(j["A"], "f64")[0] for j in jobs]
    budget = sorted(fp)[-1] + sorted(fp)[-2]  # the two largest fit together, not three
    srv = msr.NFFTServer(budget_bytes=budget, max_active=64)
    xs = [torch.empty_like(j["b"]) for j in jobs]
    ids = [srv.submit(j["A"], x, j["b"], maxiter=j["iters"]) for j, x in zip(jobs, xs)]
    while True:
        na, nq = srv.step(niter=3)
        running = [i for i, jid in enumerate(ids) if srv.poll(jid)["state"] == srv.RUNNING]
        assert sum(fp[i] for i in running) <= budget
        if na == 0 and nq == 0:
            break
    for j, x in zip(jobs, xs):
        single = msr.cg(j["A"], j["b"], maxiter=j["iters"], reltol=0.0, mode="device").cpu().numpy()
        assert rel_err(x.cpu().numpy(), single) <= 1e-6
    with pytest.raises(msr.ArgumentError):  # a job above the budget is refused
        small = msr.NFFTServer(budget_bytes=min(fp) // 2)
        small.submit(jobs[0]["A"], xs[0], jobs[0]["b"], maxiter=3)
    srv2 = msr.NFFTServer()
    srv2.submit(jobs[0]["A"], xs[0], jobs[0]["b"], maxiter=3)
    with pytest.raises(msr.ArgumentError):  # one job per operator at a time
        srv2.submit(jobs[0]["A"], xs[1], jobs[0]["b"], maxiter=3)
    srv2.run()


def test_shared_kernel_operators_match_owned_copies(torch, jobs):
    msr = load_package()
    j = jobs[1]
    rng = np.random.default_rng(5)
    dev = j["b"].device
    C = len(j["cml"])
    other = [torch.from_numpy((rng.normal(size=c.shape) + 1j * rng.normal(size=c.shape))
                              .astype(np.complex64)).to(dev) for c in j["cml"]]
    shared = msr.NFFTNormalOp_shared(j["A"], cmaps=other)
    owned = msr.NFFTNormalOp(j["N"], torch.from_numpy(j["lam"]).to(dev),
                             torch.from_numpy(j["km"] + 1).to(dev), cmaps=other)
    own_b, ker_b = msr.footprint(shared, "f64")
    assert ker_b > 0 and own_b > 0 and C == len(other)
    y1 = (shared @ j["b"]).cpu().numpy()
    y2 = (owned @ j["b"]).cpu().numpy()
    assert np.array_equal(y1, y2)
    # both operators of one kernel solve at once in a server
    srv = msr.NFFTServer()
    xa, xb = torch.empty_like(j["b"]), torch.empty_like(j["b"])
    srv.submit(j["A"], xa, j["b"], maxiter=8)
    srv.submit(shared, xb, j["b"], maxiter=8)
    srv.run(niter=4)
    ra = msr.cg(j["A"], j["b"], maxiter=8, reltol=0.0, mode="device").cpu().numpy()
    rb = msr.cg(owned, j["b"], maxiter=8, reltol=0.0, mode="device").cpu().numpy()
    assert rel_err(xa.cpu().numpy(), ra) <= 1e-6
    assert rel_err(xb.cpu().numpy(), rb) <= 1e-6
    del shared  # the kernel stays with the prototype
    y3 = (j["A"] @ j["b"]).cpu().numpy()
    assert np.isfinite(y3).all()


def test_shortest_first_policy_admits_least_work_first(torch, jobs):
    """MSR_SERVER_SHORTEST_FIRST with one job in flight: jobs finish in order
    of estimated work (iterations x coils x K x grid cells), ties in arrival
    order; results as alone."""
    msr = load_package()
    srv = msr.NFFTServer(max_active=1, policy="shortest_first")
    xs = [torch.empty_like(j["b"]) for j in jobs]
    ids = [srv.submit(j["A"], x, j["b"], maxiter=j["iters"]) for j, x in zip(jobs, xs)]
    order = []
    while len(order) < len(ids):
        srv.step(niter=50)
        for i, jid in enumerate(ids):
            if i not in order and srv.poll(jid)["state"] == srv.DONE:
                order.append(i)

    def work(i):
        j = jobs[i]
        cells = 1
        for n in j["N"]:
            cells *= 2 * n
        C = 1 if j["cml"] is None else len(j["cml"])
        K = int(j["lam"].shape[1] if j["lam"].ndim == 2 else 1)
        K = int((np.sqrt(8 * K + 1) - 1) // 2)
        return j["iters"] * C * K * cells

    expect = sorted(range(len(jobs)), key=lambda i: (work(i), i))
    print(f"completion order {order}, expected {expect}", flush=True)
    assert order == expect
    for j, x in zip(jobs, xs):
        single = msr.cg(j["A"], j["b"], maxiter=j["iters"], reltol=0.0, mode="device").cpu().numpy()
        assert rel_err(x.cpu().numpy(), single) <= 1e-6
    with pytest.raises(msr.ArgumentError):
        msr.NFFTServer(policy="random")
