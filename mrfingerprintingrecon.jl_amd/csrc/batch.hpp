// Batched reconstructions: many normal operators (uneven image size, coil
// count and K) applied and solved together.
//
// Each pass of the fused 2D apply is one launch per launch group, not one per
// operator: operators sharing 2Nx form one group of passes A and C, operators
// sharing (2Ny, K, real/complex kernel) one group of pass B, and the blocks of
// all operators of a group are concatenated in that launch (each block finds
// its operator in a job table).  3D operators, which fill the device on their
// own, keep their per-operator five-pass launches inside the same sequence.
// The CG vector steps run as one launch over all operators.  A CG iteration of
// a whole batch is therefore a fixed, small number of launches, captured once
// into a hipGraph for all iterations.
//
// Each operator keeps its own scratch and CG state, and its blocks do exactly
// the work they do in a single-operator launch, so every result is the same
// alone and in any batch.  Operators that converge early skip their blocks.
#pragma once

#include <vector>

#include "normalop.hpp"

namespace msr {

struct BatchGroup {
    int key[3] = {0, 0, 0};       // A/C: {2Nx}; B: {2Ny, K, lam_real}
    std::vector<int> jobs;        // batch indices of the group's operators
    std::vector<int> pre[3];      // block prefix (jobs.size() + 1) per ApplyMode
    int64_t nblocks[3] = {0, 0, 0};
    size_t off_tab = 0, off_pre = 0;  // byte offsets in a table set
};

// Device copy of the job tables of one use (which input the passes read,
// whether CG done flags gate them), with pinned host staging.
struct BatchTables {
    DevBuf dev;
    void* host = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;  // last upload
    ~BatchTables();
    void ensure(size_t n);
};

}  // namespace msr

struct msr_batch {
    msr_ctx* ctx = nullptr;
    std::vector<msr_normalop*> ops;
    std::vector<int> j2d, j3d;
    // the fused 2D operators' knobs per apply mode when the batch was made:
    // the block prefixes count blocks with them, so later tuning of an
    // operator does not reach this batch
    std::vector<std::array<msr::ApplyKnobs, 3>> knobs;
    std::vector<msr::BatchGroup> ga, gb, gc;
    std::vector<int> vpre;       // vector-launch block prefix over all operators
    int vblocks = 0;
    size_t off_cgjobs = 0, off_vpre = 0, table_bytes = 0;
    msr::BatchTables t_apply, t_x0, t_cg;      // eager calls
    msr::BatchTables g_x0, g_cg;               // owned by the captured graph
    bool cg_ready = false;                     // t_cg built from the operators' CG state
    hipGraphExec_t graph = nullptr;
    std::vector<const void*> graph_key;
    ~msr_batch();
};

namespace msr {
msr_batch* batch_create(msr_ctx* ctx, int nops, msr_normalop* const* ops);
void batch_apply(msr_batch* bt, float2* const* x, const float2* const* b, float2 alpha,
                 float2 beta, bool beta_zero, hipStream_t st);
void batch_cg_start(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter,
                    double reltol, double abstol, bool initially_zero, int prec, hipStream_t st);
void batch_cg_iterate(msr_batch* bt, int niter, hipStream_t st);
void batch_cg_finish(msr_batch* bt, hipStream_t st);
void batch_cg(msr_batch* bt, float2* const* x, const float2* const* b, int maxiter, double reltol,
              double abstol, bool initially_zero, int mode, int prec, hipStream_t st);
void batch_cg_status(msr_batch* bt, int* iters, double* res, int* done, hipStream_t st);
}  // namespace msr
