// Batched applies and lockstep CG over many normal operators (see batch.hpp).
#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

#include "batch.hpp"

namespace msr {

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

BatchTables::~BatchTables() {
    if (ev) (void)hipEventDestroy(ev);
    if (host) (void)hipHostFree(host);
}

void BatchTables::ensure(size_t n) {
    if (bytes == n && host) return;
    if (ev) MSR_HIP(hipEventSynchronize(ev));
    if (host) (void)hipHostFree(host);
    host = nullptr;
    MSR_HIP(hipHostMalloc(&host, n, hipHostMallocDefault));
    dev.alloc(n);
    bytes = n;
    if (!ev) MSR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
}

// ------------------------------------------------------------------ construction
msr_batch* batch_create(msr_ctx* ctx, int nops, msr_normalop* const* ops) {
    MSR_REQUIRE(nops >= 1, MSR_ERR_ARGUMENT, "a batch needs at least one operator (got %d)", nops);
    MSR_REQUIRE(ops != nullptr, MSR_ERR_ARGUMENT, "ops is NULL");
    auto bt = std::make_unique<msr_batch>();
    bt->ctx = ctx;
    std::set<const msr_normalop*> seen;
    for (int j = 0; j < nops; ++j) {
        MSR_REQUIRE(ops[j] != nullptr, MSR_ERR_ARGUMENT, "ops[%d] is NULL", j);
        MSR_REQUIRE(ops[j]->ctx == ctx, MSR_ERR_ARGUMENT, "ops[%d] belongs to another context", j);
        MSR_REQUIRE(seen.insert(ops[j]).second, MSR_ERR_ARGUMENT,
                    "ops[%d] appears twice (each operator owns one scratch and CG state)", j);
        bt->ops.push_back(ops[j]);
        bt->knobs.push_back({ops[j]->knobs[0], ops[j]->knobs[1], ops[j]->knobs[2]});
        // fused 2D operators share grouped launches; the others (3D, Cartesian,
        // line-kernel apply) run their own passes in the batch's sequence
        (ops[j]->d == 2 && !ops[j]->cart && ops[j]->fused ? bt->j2d : bt->j3d).push_back(j);
    }
    // launch groups of the 2D passes
    std::map<int, int> ia;
    std::map<std::tuple<int, int, int>, int> ib;
    for (int j : bt->j2d) {
        const msr_normalop* op = bt->ops[j];
        const int nx2 = (int)(2 * op->N[0]), ny2 = (int)(2 * op->N[1]);
        if (!ia.count(nx2)) {
            ia[nx2] = (int)bt->ga.size();
            bt->ga.emplace_back();
            bt->ga.back().key[0] = nx2;
            bt->gc.emplace_back();
            bt->gc.back().key[0] = nx2;
        }
        bt->ga[ia[nx2]].jobs.push_back(j);
        bt->gc[ia[nx2]].jobs.push_back(j);
        const auto kb = std::make_tuple(ny2, op->K, op->lam_real);
        if (!ib.count(kb)) {
            ib[kb] = (int)bt->gb.size();
            bt->gb.emplace_back();
            bt->gb.back().key[0] = ny2;
            bt->gb.back().key[1] = op->K;
            bt->gb.back().key[2] = op->lam_real;
        }
        bt->gb[ib[kb]].jobs.push_back(j);
    }
    auto prefix = [&](BatchGroup& g, int64_t (*count)(const Apply2DArgs&)) {
        for (int f = 0; f < 3; ++f) {  // ApplyMode
            g.pre[f].assign(1, 0);
            int64_t tot = 0;
            for (int j : g.jobs) {
                Apply2DArgs a{};
                const msr_normalop* op = bt->ops[j];
                a.nx = (int)op->N[0];
                a.ny = (int)op->N[1];
                a.K = op->K;
                a.C = op->C;
                a.G = op->G;
                a.Gb = bt->knobs[j][f].Gb;
                a.f64 = f != AM_F32;
                a.compact = f == AM_F64C;
                tot += count(a);
                MSR_REQUIRE(tot < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED,
                            "batched launch above 2^31 blocks");
                g.pre[f].push_back((int)tot);
            }
            g.nblocks[f] = tot;
        }
    };
    for (auto& g : bt->ga) prefix(g, blocks_a2d);
    for (auto& g : bt->gb) prefix(g, blocks_b2d);
    for (auto& g : bt->gc) prefix(g, blocks_c2d);
    bt->vpre.assign(1, 0);
    for (auto* op : bt->ops) bt->vpre.push_back(bt->vpre.back() + red_grid(op->V * op->K));
    bt->vblocks = bt->vpre.back();
    // table layout (the CG job region is sized for the largest job record)
    size_t off = 0;
    for (auto* gv : {&bt->ga, &bt->gb, &bt->gc}) {
        for (auto& g : *gv) {
            g.off_tab = off;
            off = align256(off + sizeof(Apply2DArgs) * g.jobs.size());
            g.off_pre = off;
            off = align256(off + sizeof(int) * (g.jobs.size() + 1));
        }
    }
    constexpr size_t job_bytes =
        std::max({sizeof(CgJob<float2, float2>), sizeof(CgJob<double2, float2>),
                  sizeof(CgJob<double2, double2>)});
    bt->off_cgjobs = off;
    off = align256(off + job_bytes * nops);
    bt->off_vpre = off;
    off = align256(off + sizeof(int) * bt->vpre.size());
    bt->table_bytes = off;
    return bt.release();
}

// ------------------------------------------------------------------ tables
// Apply input of every operator for one use of the tables.
struct TableUse {
    const void* const* b = nullptr;  // per operator
    int am = AM_F32;                 // apply mode (ApplyMode)
    bool skip_done = false;          // gate the passes by the CG done flags
};

template <class JT>
static void fill_tables(msr_batch* bt, BatchTables& T, const TableUse& use,
                        const std::vector<JT>& cj, hipStream_t st) {
    T.ensure(bt->table_bytes);
    if (T.ev) MSR_HIP(hipEventSynchronize(T.ev));  // the previous upload has left the staging
    char* h = static_cast<char*>(T.host);
    const int f = use.am;
    auto put_group = [&](const BatchGroup& g) {
        auto* tab = reinterpret_cast<Apply2DArgs*>(h + g.off_tab);
        for (size_t i = 0; i < g.jobs.size(); ++i) {
            const int j = g.jobs[i];
            Apply2DArgs a = args2d(bt->ops[j], use.am);
            const ApplyKnobs& kn = bt->knobs[j][use.am];
            a.Gb = kn.Gb;
            a.order_a = kn.order_a;
            a.order_c = kn.order_c;
            a.b = use.b[j];
            a.skip = use.skip_done ? &bt->ops[j]->cg_scal.as<CgScalars>()->done : nullptr;
            tab[i] = a;
        }
        memcpy(h + g.off_pre, g.pre[f].data(), sizeof(int) * g.pre[f].size());
    };
    for (auto& g : bt->ga) put_group(g);
    for (auto& g : bt->gb) put_group(g);
    for (auto& g : bt->gc) put_group(g);
    memcpy(h + bt->off_cgjobs, cj.data(), sizeof(JT) * cj.size());
    memcpy(h + bt->off_vpre, bt->vpre.data(), sizeof(int) * bt->vpre.size());
    MSR_HIP(hipMemcpyAsync(T.dev.p, T.host, bt->table_bytes, hipMemcpyHostToDevice, st));
    MSR_HIP(hipEventRecord(T.ev, st));
}

template <class VT, class AT>
static CgJob<VT, AT>* dev_cgjobs(msr_batch* bt, BatchTables& T) {
    return reinterpret_cast<CgJob<VT, AT>*>(T.dev.as<char>() + bt->off_cgjobs);
}
static const int* dev_vpre(msr_batch* bt, BatchTables& T) {
    return reinterpret_cast<const int*>(T.dev.as<char>() + bt->off_vpre);
}

template <class VT, class AT>
static CgJob<VT, AT> base_job(msr_normalop* op, float2* x) {
    CgJob<VT, AT> J{};
    J.v = op_cg_vecs<VT, AT>(op, x);
    J.partial = op->partial.as<AT>();
    J.G = op->G;
    J.scale = op_scale(op);
    J.n = op->V * op->K;
    J.S = op->cg_scal.as<CgScalars>();
    J.part = op->cg_part.as<double>();
    J.alpha = CTraits<AT>::make(1.0, 0.0);
    J.beta = cz<AT>();
    J.beta_zero = 1;
    return J;
}

// the three pass sequences of every operator, in pass order across the batch
static void batch_passes(msr_batch* bt, BatchTables& T, const TableUse& use, hipStream_t st) {
    char* base = T.dev.as<char>();
    const int f = use.am;
    auto tab = [&](const BatchGroup& g) { return reinterpret_cast<const Apply2DArgs*>(base + g.off_tab); };
    auto pre = [&](const BatchGroup& g) { return reinterpret_cast<const int*>(base + g.off_pre); };
    for (auto& g : bt->ga)
        launch_pass_a_2d_batch(g.key[0], f, tab(g), pre(g), (int)g.jobs.size(), g.nblocks[f], st);
    for (auto& g : bt->gb)
        launch_pass_b_2d_batch(g.key[0], g.key[1], g.key[2], f, tab(g), pre(g),
                               (int)g.jobs.size(), g.nblocks[f], st);
    for (auto& g : bt->gc)
        launch_pass_c_2d_batch(g.key[0], f, tab(g), pre(g), (int)g.jobs.size(), g.nblocks[f], st);
    for (int j : bt->j3d) {
        msr_normalop* op = bt->ops[j];
        op_passes(op, use.b[j], use.am, st,
                  use.skip_done ? &op->cg_scal.as<CgScalars>()->done : nullptr);
    }
}

// ------------------------------------------------------------------ apply
void batch_apply(msr_batch* bt, float2* const* x, const float2* const* b, float2 alpha,
                 float2 beta, bool beta_zero, hipStream_t st) {
    const int nj = (int)bt->ops.size();
    MSR_REQUIRE(x && b, MSR_ERR_ARGUMENT, "x or b is NULL");
    std::vector<CgJob<float2, float2>> cj(nj);
    for (int j = 0; j < nj; ++j) {
        MSR_REQUIRE(x[j] && b[j], MSR_ERR_ARGUMENT, "x[%d] or b[%d] is NULL", j, j);
        MSR_REQUIRE((const void*)x[j] != (const void*)b[j], MSR_ERR_ARGUMENT,
                    "x[%d] and b[%d] must not alias", j, j);
        msr_normalop* op = bt->ops[j];
        op_ensure_scratch(op, false);
        CgJob<float2, float2> J{};
        J.partial = op->partial.as<float2>();
        J.G = op->G;
        J.scale = op_scale(op);
        J.n = op->V * op->K;
        J.y = x[j];
        J.alpha = alpha;
        J.beta = beta;
        J.beta_zero = beta_zero ? 1 : 0;
        cj[j] = J;
    }
    std::vector<const void*> bin(b, b + nj);
    TableUse use;
    use.b = bin.data();
    fill_tables(bt, bt->t_apply, use, cj, st);
    batch_passes(bt, bt->t_apply, use, st);
    launch_epilogue_batch(dev_cgjobs<float2, float2>(bt, bt->t_apply), dev_vpre(bt, bt->t_apply),
                          nj, bt->vblocks, st);
}

// ------------------------------------------------------------------ CG
// Per-operator inputs of the CG passes: each operator's search direction in
// the apply type.
template <class VT, class AT>
static std::vector<const void*> cg_inputs(msr_batch* bt) {
    std::vector<const void*> u(bt->ops.size());
    for (size_t j = 0; j < bt->ops.size(); ++j)
        u[j] = op_cg_vecs<VT, AT>(bt->ops[j], bt->ops[j]->cg_x_user).ua;
    return u;
}

template <class VT, class AT>
static TableUse cg_use(std::vector<const void*>& u, int am) {
    TableUse use;
    use.b = u.data();
    use.am = am;
    use.skip_done = true;
    return use;
}

template <class VT, class AT>
static void build_cg_tables(msr_batch* bt, BatchTables& T, const float2* const* b, int am,
                            hipStream_t st) {
    const int nj = (int)bt->ops.size();
    std::vector<CgJob<VT, AT>> cj(nj);
    for (int j = 0; j < nj; ++j) {
        msr_normalop* op = bt->ops[j];
        CgJob<VT, AT> J = base_job<VT, AT>(op, op->cg_x_user);
        J.b = b ? b[j] : nullptr;
        J.x_in = op->cg_x_user;
        J.x_out = op->cg_x_user;
        cj[j] = J;
    }
    auto u = cg_inputs<VT, AT>(bt);
    fill_tables(bt, T, cg_use<VT, AT>(u, am), cj, st);
}

// The x0 tables: inputs x_j in the apply type (converted into each operator's
// apply-input buffer first when that is FP64), epilogue target c_j.
template <class VT, class AT>
static void build_x0_tables(msr_batch* bt, BatchTables& T, float2* const* x, int am,
                            hipStream_t st) {
    const int nj = (int)bt->ops.size();
    std::vector<CgJob<VT, AT>> cj(nj);
    std::vector<const void*> xin(nj);
    for (int j = 0; j < nj; ++j) {
        CgJob<VT, AT> J = base_job<VT, AT>(bt->ops[j], x[j]);
        J.y = J.v.c;
        cj[j] = J;
        xin[j] = std::is_same<AT, double2>::value ? (const void*)J.v.ua : (const void*)x[j];
    }
    TableUse use;
    use.b = xin.data();
    use.am = am;
    fill_tables(bt, T, use, cj, st);
}

__global__ void batch_f32_to_f64_kernel(const float2* __restrict__ a, double2* __restrict__ b,
                                        int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = make_double2(a[i].x, a[i].y);
}

template <class VT, class AT>
static void start_seq(msr_batch* bt, BatchTables& Tcg, BatchTables& Tx0, float2* const* x,
                      int maxiter, double reltol, double abstol, bool initially_zero, int am,
                      hipStream_t st) {
    const int nj = (int)bt->ops.size();
    if (!initially_zero) {
        // c = A x0 (the same apply and epilogue as a single operator's prologue)
        std::vector<const void*> xin(nj);
        for (int j = 0; j < nj; ++j) {
            msr_normalop* op = bt->ops[j];
            if constexpr (std::is_same<AT, double2>::value) {
                const CgVecs<VT, AT> v = op_cg_vecs<VT, AT>(op, x[j]);
                const int64_t n = op->V * op->K;
                int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
                hipLaunchKernelGGL(batch_f32_to_f64_kernel, dim3(grid), dim3(256), 0, st, x[j],
                                   v.ua, n);
                MSR_CHECK_LAUNCH();
                xin[j] = v.ua;
            } else {
                xin[j] = x[j];
            }
        }
        TableUse use;
        use.b = xin.data();
        use.am = am;
        batch_passes(bt, Tx0, use, st);
        launch_epilogue_batch(dev_cgjobs<VT, AT>(bt, Tx0), dev_vpre(bt, Tx0), nj, bt->vblocks, st);
    }
    launch_cg_start_batch<VT, AT>(dev_cgjobs<VT, AT>(bt, Tcg), dev_vpre(bt, Tcg), nj, bt->vblocks,
                                  initially_zero, !initially_zero, reltol, abstol, maxiter, st);
}

template <class VT, class AT>
static void iterate_seq(msr_batch* bt, BatchTables& T, int niter, int am, hipStream_t st) {
    const int nj = (int)bt->ops.size();
    auto u = cg_inputs<VT, AT>(bt);
    const TableUse use = cg_use<VT, AT>(u, am);
    CgJob<VT, AT>* cj = dev_cgjobs<VT, AT>(bt, T);
    const int* vp = dev_vpre(bt, T);
    for (int it = 0; it < niter; ++it) {
        launch_cg_u_batch<VT, AT>(cj, vp, nj, bt->vblocks, st);
        batch_passes(bt, T, use, st);
        launch_cg_reduce_dot_batch<VT, AT>(cj, vp, nj, bt->vblocks, st);
        launch_cg_xr_batch<VT, AT>(cj, vp, nj, bt->vblocks, st);
    }
}

template <class VT, class AT>
static void finish_seq(msr_batch* bt, BatchTables& T, hipStream_t st) {
    if constexpr (std::is_same<VT, double2>::value)
        launch_to_f32_batch<AT>(dev_cgjobs<double2, AT>(bt, T), dev_vpre(bt, T),
                                (int)bt->ops.size(), bt->vblocks, st);
}

static void check_vectors(msr_batch* bt, float2* const* x, const float2* const* b) {
    MSR_REQUIRE(x && b, MSR_ERR_ARGUMENT, "x or b is NULL");
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        MSR_REQUIRE(x[j] && b[j], MSR_ERR_ARGUMENT, "x[%zu] or b[%zu] is NULL", j, j);
        MSR_REQUIRE((const void*)x[j] != (const void*)b[j], MSR_ERR_ARGUMENT,
                    "x[%zu] and b[%zu] must not alias", j, j);
    }
}

static void check_prec(int prec) {
    MSR_REQUIRE(prec_valid(prec), MSR_ERR_ARGUMENT, "unknown CG precision %d", prec);
}

static void record_start(msr_batch* bt, float2* const* x, int prec) {
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        msr_normalop* op = bt->ops[j];
        op_cg_buffers(op, prec);
        op->cg_fp64 = prec;
        op->cg_x_user = x[j];
    }
}

// dispatch a templated step on the CG precision
#define MSR_PREC_DISPATCH(prec, CALL)                          \
    do {                                                       \
        if (apply_mode(prec) != AM_F32) CALL(double2, double2); \
        else if ((prec) == MSR_PREC_MIXED) CALL(double2, float2); \
        else CALL(float2, float2);                             \
    } while (0)

template <class VT, class AT>
static void start_eager(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter,
                        double reltol, double abstol, bool initially_zero, int am,
                        hipStream_t st) {
    build_cg_tables<VT, AT>(bt, bt->t_cg, b, am, st);
    if (!initially_zero) build_x0_tables<VT, AT>(bt, bt->t_x0, x, am, st);
    start_seq<VT, AT>(bt, bt->t_cg, bt->t_x0, x, maxiter, reltol, abstol, initially_zero, am, st);
}

void batch_cg_start(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter,
                    double reltol, double abstol, bool initially_zero, int prec, hipStream_t st) {
    check_vectors(bt, x, b);
    check_prec(prec);
    MSR_REQUIRE(maxiter >= 0, MSR_ERR_ARGUMENT, "maxiter = %d", maxiter);
    record_start(bt, x, prec);
#define MSR_START(VT, AT) \
    start_eager<VT, AT>(bt, x, b, maxiter, reltol, abstol, initially_zero, apply_mode(prec), st)
    MSR_PREC_DISPATCH(prec, MSR_START);
#undef MSR_START
    bt->cg_ready = true;
}

static int batch_prec(msr_batch* bt) {
    const int f = bt->ops[0]->cg_fp64;
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        const msr_normalop* op = bt->ops[j];
        MSR_REQUIRE(op->cg_fp64 >= 0 && op->cg_x_user, MSR_ERR_ARGUMENT,
                    "ops[%zu] has no started CG (call msr_batch_cg_start first)", j);
        MSR_REQUIRE(op->cg_fp64 == f, MSR_ERR_ARGUMENT,
                    "the operators of a batch were started with different precisions");
    }
    return f;
}

static void ensure_cg_tables(msr_batch* bt, int prec, hipStream_t st) {
    if (bt->cg_ready) return;
#define MSR_BUILD(VT, AT) build_cg_tables<VT, AT>(bt, bt->t_cg, nullptr, apply_mode(prec), st)
    MSR_PREC_DISPATCH(prec, MSR_BUILD);
#undef MSR_BUILD
    bt->cg_ready = true;
}

void batch_cg_iterate(msr_batch* bt, int niter, hipStream_t st) {
    MSR_REQUIRE(niter >= 0, MSR_ERR_ARGUMENT, "niter = %d", niter);
    const int prec = batch_prec(bt);
    ensure_cg_tables(bt, prec, st);
#define MSR_ITER(VT, AT) iterate_seq<VT, AT>(bt, bt->t_cg, niter, apply_mode(prec), st)
    MSR_PREC_DISPATCH(prec, MSR_ITER);
#undef MSR_ITER
}

void batch_cg_finish(msr_batch* bt, hipStream_t st) {
    const int prec = batch_prec(bt);
    ensure_cg_tables(bt, prec, st);
#define MSR_FIN(VT, AT) finish_seq<VT, AT>(bt, bt->t_cg, st)
    MSR_PREC_DISPATCH(prec, MSR_FIN);
#undef MSR_FIN
}

template <class VT, class AT>
static void cg_graph_t(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter,
                       double reltol, double abstol, bool initially_zero, int am, hipStream_t st,
                       const std::vector<const void*>& key) {
    if (bt->graph && bt->graph_key == key) {
        MSR_HIP(hipGraphLaunch(bt->graph, st));
        return;
    }
    if (bt->graph) {
        MSR_HIP(hipStreamSynchronize(st));  // the old graph may still read its tables
        (void)hipGraphExecDestroy(bt->graph);
        bt->graph = nullptr;
    }
    // tables of the graph, uploaded before capture and left untouched afterwards
    build_cg_tables<VT, AT>(bt, bt->g_cg, b, am, st);
    if (!initially_zero) build_x0_tables<VT, AT>(bt, bt->g_x0, x, am, st);
    hipGraph_t g = nullptr;
    MSR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        start_seq<VT, AT>(bt, bt->g_cg, bt->g_x0, x, maxiter, reltol, abstol, initially_zero, am,
                          st);
        iterate_seq<VT, AT>(bt, bt->g_cg, maxiter, am, st);
        finish_seq<VT, AT>(bt, bt->g_cg, st);
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    MSR_HIP(hipStreamEndCapture(st, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MSR_HIP(e);
    bt->graph = ge;
    bt->graph_key = key;
    MSR_HIP(hipGraphLaunch(bt->graph, st));
}

void batch_cg(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter, double reltol,
              double abstol, bool initially_zero, int mode, int prec, hipStream_t st) {
    check_vectors(bt, x, b);
    check_prec(prec);
    MSR_REQUIRE(maxiter >= 0, MSR_ERR_ARGUMENT, "maxiter = %d", maxiter);
    if (mode == MSR_CG_DEVICE) {
        batch_cg_start(bt, x, b, maxiter, reltol, abstol, initially_zero, prec, st);
        batch_cg_iterate(bt, maxiter, st);
        batch_cg_finish(bt, st);
        return;
    }
    MSR_REQUIRE(mode == MSR_CG_GRAPH, MSR_ERR_UNSUPPORTED,
                "batched CG runs in MSR_CG_DEVICE or MSR_CG_GRAPH mode (got %d)", mode);
    MSR_REQUIRE(st != nullptr, MSR_ERR_ARGUMENT, "MSR_CG_GRAPH needs an explicit (non-default) stream");
    record_start(bt, x, prec);
    // the captured kernels hold every operator's buffer addresses
    std::vector<const void*> key;
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        key.push_back(x[j]);
        key.push_back(b[j]);
        const std::vector<const void*> k = op_graph_key(bt->ops[j]);
        key.insert(key.end(), k.begin(), k.end());
    }
    uint64_t bits[2];
    memcpy(&bits[0], &reltol, 8);
    memcpy(&bits[1], &abstol, 8);
    key.push_back((const void*)(uintptr_t)bits[0]);
    key.push_back((const void*)(uintptr_t)bits[1]);
    key.push_back((const void*)(intptr_t)maxiter);
    key.push_back((const void*)(intptr_t)((initially_zero ? 1 : 0) + 2 * prec));
    key.push_back(st);
#define MSR_GRAPH(VT, AT) \
    cg_graph_t<VT, AT>(bt, x, b, maxiter, reltol, abstol, initially_zero, apply_mode(prec), st, key)
    MSR_PREC_DISPATCH(prec, MSR_GRAPH);
#undef MSR_GRAPH
    bt->cg_ready = false;  // eager calls rebuild t_cg from the operators' state
}

void batch_cg_status(msr_batch* bt, int* iters, double* res, int* done, hipStream_t st) {
    std::vector<CgScalars> hs(bt->ops.size());
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        MSR_REQUIRE(bt->ops[j]->cg_scal.p, MSR_ERR_ARGUMENT, "ops[%zu] has no CG state", j);
        MSR_HIP(hipMemcpyAsync(&hs[j], bt->ops[j]->cg_scal.p, sizeof(CgScalars),
                               hipMemcpyDeviceToHost, st));
    }
    MSR_HIP(hipStreamSynchronize(st));
    for (size_t j = 0; j < bt->ops.size(); ++j) {
        if (iters) iters[j] = hs[j].iter;
        if (res) res[j] = hs[j].res;
        if (done) done[j] = hs[j].done;
    }
}

}  // namespace msr

msr_batch::~msr_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
}
