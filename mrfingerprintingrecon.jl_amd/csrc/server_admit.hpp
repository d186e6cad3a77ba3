// Admission step of the reconstruction server (server.hip), kept free of HIP
// so that its bookkeeping can be tested on the host.
#pragma once

#include <cstddef>
#include <cstdint>
#include <deque>
#include <vector>

namespace msr {

// Admits jobs from the front of `queue` while fewer than max_active jobs are
// in flight and their bytes fit the budget: start(id) starts a job's solve,
// and only after it returns does the job leave the queue, join `active` and
// count its bytes.  If start throws, that job stays at the front of the queue
// and is not counted, the jobs started before it stay admitted, and the
// exception propagates.  Returns whether any job was admitted.
template <class BytesFn, class StartFn>
bool admit_jobs(std::deque<int64_t>& queue, std::vector<int64_t>& active, size_t& active_bytes,
                size_t budget, int max_active, BytesFn bytes, StartFn start) {
    bool admitted = false;
    while (!queue.empty() && (int)active.size() < max_active) {
        const int64_t id = queue.front();
        const size_t b = bytes(id);
        if (active_bytes + b > budget) break;
        start(id);
        queue.pop_front();
        active_bytes += b;
        active.push_back(id);
        admitted = true;
    }
    return admitted;
}

}  // namespace msr
