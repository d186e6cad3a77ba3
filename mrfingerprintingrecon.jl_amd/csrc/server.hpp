// Server state (server.hip): the job table and queue of the continuous
// batching loop.
#pragma once

#include <chrono>
#include <deque>
#include <map>
#include <vector>

#include "batch.hpp"
#include "server_admit.hpp"

namespace msr {

struct ServerJob {
    msr_normalop* op = nullptr;
    float2* x = nullptr;
    const float2* b = nullptr;
    int maxiter = 0;
    double reltol = 0.0, abstol = 0.0;
    int state = MSR_JOB_QUEUED;
    int iters = 0;
    double res = 0.0;
    size_t bytes = 0;
    double work = 0.0;        // estimated work (admission order of MSR_SERVER_SHORTEST_FIRST)
    double t_submit = 0.0, t_start = 0.0, t_done = 0.0;  // host clock, ms
};

}  // namespace msr

struct msr_server {
    msr_ctx* ctx = nullptr;
    size_t budget = 0;
    int max_active = 0;
    int prec = MSR_PREC_F64;
    int policy = MSR_SERVER_FIFO;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    std::map<int64_t, msr::ServerJob> jobs;
    int64_t next_id = 1;
    size_t active_bytes = 0;
    msr_batch* batch = nullptr;  // over the operators of the jobs in flight
    ~msr_server() { delete batch; }
};
