// Non-uniform spreading onto an oversampled grid (the spreading half of a
// type-1 NUFFT) with an exponential-of-semicircle kernel
//   phi(t) = exp(beta (sqrt(1 - (2t/W)^2) - 1)),  |t| < W/2,
// and its Fourier transform phihat(r) for deapodisation.
//
// Points are stably sorted into grid tiles once per (trajectory, grid); one
// workgroup accumulates one tile (plus a 4-cell halo) for a group of value
// channels in LDS, one wavefront per channel walking the tile's points in
// order (one lane per kernel cell).  2D tiles are then summed into the grid by
// a gather in a fixed order (bit-reproducible); 3D tiles are added with global
// atomics.  Grids are channel-major, x fastest: g[ch][z][y][x].
#pragma once

#include "internal.hpp"

namespace msr {

constexpr int SPREAD_W = 8;                 // kernel width (cells): ~1.2e-7 at sigma = 2
constexpr float SPREAD_BETA = 2.30f * SPREAD_W;
constexpr int SPREAD_HALO = 4;              // cells of halo on each tile side

// Build (or fetch from the trajectory's cache) the tile sort for grid n.
const SpreadPlan& spread_plan(const msr_traj* traj, const int64_t* n, hipStream_t st);

// Values of the spread channels for one point: BP or kernel-build sources.
struct ValueSrc {
    int kind;             // 0: backprojection data, 1: basis pair products
    // backprojection
    const float2* data;   // (nsamp*nt, C) column-major (sample fastest)
    const int64_t* sel;   // masked sample -> index in nsamp*nt (nullptr: identity)
    int64_t nsel;         // nsamp*nt
    const float* dcf;     // per masked sample weight or nullptr
    int dcf_mode;         // 0 none, 1 radial 2D, 2 radial 3D, 3 array
    int K;                // coefficients
    // basis (both kinds): U (nt, K) column-major, complex or real, nullptr = identity
    const void* U;
    int u_real;
    int64_t nt;
    const int* pair_j;    // kind 1: channel -> (j, k) with j <= k
    const int* pair_k;
};

// grid (nch, prod n) complex is overwritten with the spread values of all
// channels.  2D needs a workspace of spread_workspace_bytes(plan, nch) (the
// padded tiles, summed into the grid in a fixed order); 3D accumulates with
// global atomics and takes none.
// f64: accumulate, and store the grid and workspace, in ComplexF64 (the
// values and kernel weights are then evaluated in FP64 too).
size_t spread_workspace_bytes(const SpreadPlan& plan, int nch, bool f64 = false);
void spread(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
            void* grid, void* slab, bool f64, hipStream_t st);

// 2D ComplexF32 backprojection spreading onto output-owned tiles
// (spread_bp.hip): builds the plan's point streams and pair lists on first
// use; the workspace holds the partial tiles of split tiles and the coil data
// in sorted point order.  grid (nch, n1, n0) is overwritten.
size_t spread_bp_workspace_bytes(const msr_traj* traj, const SpreadPlan& plan, int nch, int ncoils,
                                 hipStream_t st);
void spread_bp(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
               float2* grid, float2* work, hipStream_t st);
// The 3D form on 4 x 4 x 4 output blocks (spread_bp3.hip); grid (nch, n2,
// n1, n0) is overwritten.  Supported up to 2^31 / 64 samples (int pair
// offsets); larger trajectories take the tile spreading.
bool spread_bp3_supported(const msr_traj* traj);
size_t spread_bp3_workspace_bytes(const msr_traj* traj, const SpreadPlan& plan, int nch,
                                  hipStream_t st);
void spread_bp3(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
                float2* grid, float2* work, hipStream_t st);

// Device table 1/phihat(r) for the output frequencies of a type-1 NUFFT of
// size n_out (sel: 1 centred r = j - n_out/2, 2 FFT order) on an n_os grid.
const float* deapod_table(msr_ctx* ctx, int n_os, int n_out, int sel);
const double* deapod_table64(msr_ctx* ctx, int n_os, int n_out, int sel);

}  // namespace msr
