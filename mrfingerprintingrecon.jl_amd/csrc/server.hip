// Reconstruction server: continuous batching of CG solves with admission by
// device-memory footprint, on the stepwise batch engine (batch.hip).
//
// Jobs (an operator, its right-hand side and solution vectors, an iteration
// budget and tolerances) wait in a FIFO queue.  Each step admits queued jobs
// while the device state of the jobs in flight (coil maps, apply scratch and
// CG vectors of their operators; kernels are resident and shared between
// operators built with msr_normalop_create_shared) stays within the budget,
// starts their CG, runs a number of lockstep CG iterations over every job in
// flight (one batched launch sequence), and retires the jobs whose CG has
// stopped, writing their solutions.  A job that joins while others run
// overlaps with them on the device; every job's result is the one its
// operator gives alone (the batch engine runs each operator's blocks exactly
// as a single-operator launch does).
#include <algorithm>

#include "server.hpp"

namespace msr {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

msr_server* server_create(msr_ctx* ctx, size_t budget, int max_active, int prec) {
    MSR_REQUIRE(max_active >= 1, MSR_ERR_ARGUMENT, "max_active = %d", max_active);
    MSR_REQUIRE(prec_valid(prec), MSR_ERR_ARGUMENT, "precision = %d", prec);
    auto* s = new msr_server;
    s->ctx = ctx;
    s->budget = budget;
    s->max_active = max_active;
    s->prec = prec;
    return s;
}

int64_t server_submit(msr_server* s, msr_normalop* op, float2* x, const float2* b, int maxiter,
                      double reltol, double abstol) {
    MSR_REQUIRE(op->ctx == s->ctx, MSR_ERR_ARGUMENT, "the operator belongs to another context");
    MSR_REQUIRE(maxiter >= 0, MSR_ERR_ARGUMENT, "maxiter = %d", maxiter);
    for (const auto& kv : s->jobs)
        MSR_REQUIRE(kv.second.state == MSR_JOB_DONE || kv.second.op != op, MSR_ERR_ARGUMENT,
                    "the operator already has a job in this server (one solve per operator at "
                    "a time; msr_normalop_create_shared gives more operators on one kernel)");
    ServerJob j;
    j.op = op;
    j.x = x;
    j.b = b;
    j.maxiter = maxiter;
    j.reltol = reltol;
    j.abstol = abstol;
    j.bytes = op_footprint(op, s->prec);
    int64_t cells = 1;
    for (int a = 0; a < op->d; ++a) cells *= (op->cart ? 1 : 2) * op->N[a];
    j.work = (double)maxiter * op->C * op->K * (double)cells;
    MSR_REQUIRE(j.bytes <= s->budget, MSR_ERR_ARGUMENT,
                "the job needs %zu bytes of device state, above the server budget of %zu", j.bytes,
                s->budget);
    j.t_submit = now_ms();
    const int64_t id = s->next_id++;
    s->jobs[id] = j;
    s->queue.push_back(id);
    return id;
}

static void rebuild_batch(msr_server* s) {
    delete s->batch;
    s->batch = nullptr;
    if (s->active.empty()) return;
    std::vector<msr_normalop*> ops;
    for (int64_t id : s->active) ops.push_back(s->jobs[id].op);
    s->batch = batch_create(s->ctx, (int)ops.size(), ops.data());
}

void server_set_policy(msr_server* s, int policy) {
    MSR_REQUIRE(policy == MSR_SERVER_FIFO || policy == MSR_SERVER_SHORTEST_FIRST,
                MSR_ERR_ARGUMENT, "server policy %d", policy);
    s->policy = policy;
}

void server_step(msr_server* s, int niter, int* n_active, int* n_queued, hipStream_t st) {
    MSR_REQUIRE(niter >= 1, MSR_ERR_ARGUMENT, "niter = %d", niter);
    if (s->policy == MSR_SERVER_SHORTEST_FIRST)  // least work first, ties in arrival order
        std::stable_sort(s->queue.begin(), s->queue.end(), [&](int64_t a, int64_t b) {
            return s->jobs[a].work < s->jobs[b].work;
        });
    // admit (queue order) while the jobs in flight fit the budget; a job alone
    // always fits (submit refuses jobs above the budget)
    bool admitted = false;
    try {
        admitted = admit_jobs(s->queue, s->active, s->active_bytes, s->budget, s->max_active,
                              [&](int64_t id) -> size_t { return s->jobs[id].bytes; },
                              [&](int64_t id) {
                                  ServerJob& j = s->jobs[id];
                                  std::unique_ptr<msr_batch> one(batch_create(s->ctx, 1, &j.op));
                                  batch_cg_start(one.get(), &j.x, &j.b, j.maxiter, j.reltol,
                                                 j.abstol, true, s->prec, st);
                                  // the start kernels read the batch's tables
                                  MSR_HIP(hipStreamSynchronize(st));
                                  j.state = MSR_JOB_RUNNING;
                                  j.t_start = now_ms();
                              });
    } catch (...) {
        // the jobs started before the failure are in flight: batch them
        if (!s->active.empty()) rebuild_batch(s);
        throw;
    }
    if (admitted) rebuild_batch(s);
    if (s->batch) {
        batch_cg_iterate(s->batch, niter, st);
        // retire the jobs whose CG has stopped
        const size_t na = s->active.size();
        std::vector<int> iters(na), done(na);
        std::vector<double> res(na);
        batch_cg_status(s->batch, iters.data(), res.data(), done.data(), st);
        std::vector<msr_normalop*> fin;
        std::vector<int64_t> keep;
        for (size_t i = 0; i < na; ++i) {
            ServerJob& j = s->jobs[s->active[i]];
            j.iters = iters[i];
            j.res = res[i];
            if (done[i]) fin.push_back(j.op);
            else keep.push_back(s->active[i]);
        }
        if (!fin.empty()) {
            std::unique_ptr<msr_batch> fb(batch_create(s->ctx, (int)fin.size(), fin.data()));
            batch_cg_finish(fb.get(), st);
            MSR_HIP(hipStreamSynchronize(st));
            const double t = now_ms();
            for (size_t i = 0; i < na; ++i) {
                ServerJob& j = s->jobs[s->active[i]];
                if (!done[i]) continue;
                j.state = MSR_JOB_DONE;
                j.t_done = t;
                s->active_bytes -= j.bytes;
            }
            s->active = keep;
            rebuild_batch(s);
        }
    }
    if (n_active) *n_active = (int)s->active.size();
    if (n_queued) *n_queued = (int)s->queue.size();
}

void server_poll(msr_server* s, int64_t id, int* state, int* iters, double* res, double* times) {
    auto it = s->jobs.find(id);
    MSR_REQUIRE(it != s->jobs.end(), MSR_ERR_ARGUMENT, "no job %lld in this server", (long long)id);
    const ServerJob& j = it->second;
    if (state) *state = j.state;
    if (iters) *iters = j.iters;
    if (res) *res = j.res;
    if (times) {
        const double t = now_ms();
        const double start = j.state == MSR_JOB_QUEUED ? t : j.t_start;
        const double end = j.state == MSR_JOB_DONE ? j.t_done : t;
        times[0] = start - j.t_submit;  // queued
        times[1] = end - start;         // in flight
        times[2] = end - j.t_submit;    // total
    }
}

void server_forget(msr_server* s, int64_t id) {
    auto it = s->jobs.find(id);
    MSR_REQUIRE(it != s->jobs.end(), MSR_ERR_ARGUMENT, "no job %lld in this server", (long long)id);
    MSR_REQUIRE(it->second.state == MSR_JOB_DONE, MSR_ERR_ARGUMENT,
                "job %lld has not finished", (long long)id);
    s->jobs.erase(it);
}

}  // namespace msr
