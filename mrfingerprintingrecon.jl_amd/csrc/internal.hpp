// Internal object definitions behind the opaque C handles.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"

#include <tuple>

namespace msr {
// Launch choices of the fused 2D passes that leave the result unchanged bit
// for bit: pass B coil groups and the pass A / pass C block orders (apply2d.hip).
struct ApplyKnobs {
    int Gb = 1;
    int order_a = 1;
    int order_c = 1;
};
}  // namespace msr

struct msr_ctx {
    int device = 0;
    // tuned knobs per (nx, ny, K, C, kernel real, identity maps, apply mode)
    std::map<std::array<int64_t, 7>, msr::ApplyKnobs> tuned;
    std::map<int, float2*> twiddles;     // N -> device table exp(-2 pi i m/N), ComplexF32
    std::map<int, double2*> twiddles64;  // the same in ComplexF64
    std::map<std::tuple<int, int, int>, float*> deapod;  // (n_os, n_out, sel) -> 1/phihat
    std::map<std::tuple<int, int, int>, double*> deapod64;  // the same in FP64
    std::mutex lock;  // guards the two lazily filled tables (calls may come from several threads)
    ~msr_ctx();
};

namespace msr {

// Device buffer with RAII free.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n == bytes && p) return;
        if (p) MSR_HIP(hipDeviceSynchronize());  // in-flight kernels may still use p
        release();
        if (n == 0) return;
        MSR_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

const float2* twiddle_table(msr_ctx* ctx, int n);
const double2* twiddle_table64(msr_ctx* ctx, int n);
void ensure_device(const msr_ctx* ctx);

// Tile sort of a trajectory's points for one oversampled grid (spread.hip).
struct SpreadPlan {
    int d = 2;
    int64_t n[3] = {1, 1, 1};   // grid size
    int tb[3] = {1, 1, 1};      // tile size
    int nb[3] = {1, 1, 1};      // tiles per dim
    int64_t nbins = 0;
    DevBuf perm;                // (M) int32: points sorted by tile
    DevBuf offs;                // (nbins + 1) int32
    // output-block streams of the ComplexF32 backprojection spreading (built
    // on first use; 8 x 8 blocks in 2D, 4 x 4 x 4 in 3D): per sorted point its
    // cell, frame and data index; per block the sorted points whose support
    // reaches it (pair list) with their kernel weights over the block's
    // columns, rows (and planes), cut into bounded work items
    bool streams_ready = false;
    int64_t npairs = 0, nitems = 0, nsplit = 0, nslots = 0;
    DevBuf pw;                  // (npairs, 16) float: weights over the block's 8 columns and 8 rows
                                // (2D); (npairs, 12): 4 columns, rows and planes (3D)
    DevBuf cell;                // (M) int2 (2D) or int4 (3D): floor cell
    DevBuf frame_s;             // (M) int32
    DevBuf jsel_s;              // (M) int64
    DevBuf kc_s;                // (M) float2 (2D)
    DevBuf r2_s;                // (M) float: |k|^2 of the sample (3D)
    DevBuf pairs;               // (npairs) int32: sorted point index, grouped by output block
    DevBuf items;               // (nitems) int4 {tile, q0, q1, slab slot or -1}
    DevBuf split_tiles;         // (nsplit) int2 {tile, first item}; items of a split tile are consecutive
    DevBuf split_nitems;        // (nsplit) int32
    DevBuf order;               // (8, order_len) int32: dispatch table of the gathers (item
                                // per XCD and position, -1 past an XCD's last item)
    int64_t order_len = 0;
};

// Cell sort of an integer (Cartesian) trajectory's samples for an N grid
// (cartesian.hip): masked samples ordered by grid cell, ties in sample order.
struct CartPlan {
    int64_t n[3] = {1, 1, 1};
    int64_t G = 1;
    DevBuf perm;                // (M) int32: masked sample indices sorted by cell
    DevBuf offs;                // (G + 1) int32
};

// CG scalar state kept on the device (FP64).
struct CgScalars {
    double res, prev, tol;
    double alpha_re, alpha_im;
    double dot_re, dot_im;
    int iter, done, maxiter, pad;
};

}  // namespace msr

struct msr_traj {
    msr_ctx* ctx = nullptr;
    int d = 0;
    int64_t nsamp = 0, nt = 0;
    int64_t M = 0;                    // masked samples
    msr::DevBuf kc;                   // (M, d) float32, masked, frame-major
    msr::DevBuf frame;                // (M) int32 frame index
    msr::DevBuf sel;                  // (M) int64 index of the masked sample in nsamp*nt, or empty if unmasked
    std::vector<int64_t> nsamp_t;     // host copy of the per-frame counts
    mutable std::map<std::array<int64_t, 3>, std::unique_ptr<msr::SpreadPlan>> plans;
    mutable msr::DevBuf bp_work;      // backprojection workspace (grow-only)
    bool integer = false;             // created from integer (Cartesian) k-space indices
    mutable std::map<std::array<int64_t, 3>, std::unique_ptr<msr::CartPlan>> cart_plans;
    // The plan maps and the workspace are shared by every call on the handle:
    // `mu` serialises the host side of those calls (plan creation, workspace
    // growth), and a backprojection on another stream than the workspace's
    // last user waits for `work_event` (recorded after that user's kernels),
    // so calls from several threads or streams on one handle do not race.
    mutable std::recursive_mutex mu;
    mutable hipEvent_t work_event = nullptr;
    mutable hipStream_t work_stream = nullptr;
    ~msr_traj() {
        if (work_event) (void)hipEventDestroy(work_event);
    }
};

struct msr_normalop {
    msr_ctx* ctx = nullptr;
    int d = 2;
    int64_t N[3] = {1, 1, 1};         // image size (x first)
    int64_t V = 1;                    // voxels
    int K = 1, P = 1, C = 1;
    int lam_real = 0;
    int cart = 0;                     // Cartesian FFTNormalOp: N grid, full K x K kernel per cell
    bool identity_cmaps = true;
    int64_t nkm = 0;
    msr::DevBuf cmaps;                // (C, V) complex
    std::shared_ptr<msr::DevBuf> lam = std::make_shared<msr::DevBuf>();    // dense kernel (shared by msr_normalop_create_shared)
    std::shared_ptr<msr::DevBuf> kmask = std::make_shared<msr::DevBuf>();  // (nkm) int64, 1-based
    msr::DevBuf T;                    // FFT intermediate
    msr::DevBuf T2;                   // second intermediate (3D)
    msr::DevBuf partial;              // (G, K, V)
    size_t scratch_es = 0;            // element size T / T2 / partial are allocated for
    int G = 1, Gb = 1;                // coil groups of pass C / of pass B (3D)
    msr::ApplyKnobs knobs[3];         // fused 2D launch choices per apply mode
    bool fused = true;                // fused Toeplitz passes (else the line-kernel apply)
    int Cb = 1;                       // coils per chunk of the line-kernel apply
    // CG state (allocated on first use)
    msr::DevBuf cg_r, cg_u, cg_c, cg_scal, cg_part, cg_x64, cg_u32;
    int cg_fp64 = -1;                 // precision (MSR_PREC_*) of the last started CG
    float2* cg_x_user = nullptr;      // its caller-side FP32 solution
    hipGraphExec_t graph = nullptr;
    std::vector<const void*> graph_key;
    int graph_maxiter = -1;
    double graph_tol[2] = {0.0, 0.0};
    int graph_fp64 = -1;              // precision of the captured CG
    hipStream_t graph_stream = nullptr;
    ~msr_normalop();
};
