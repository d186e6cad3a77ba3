/*
 * Oracle support code (test infrastructure only; never on the product path).
 *
 * 1. FP64 "exponential of semicircle" (ES) spreading and interpolation used by
 *    the oracle's high-accuracy NUFFT (kernel width w ~ 14, error ~1e-13).  It
 *    is the checker for backprojection and Toeplitz-kernel results at sizes
 *    where the exact NUDFT is too slow.
 * 2. An FP32 spreading routine with one value channel per call, used by the
 *    CPU-baseline port of the reference's backprojection
 *    (src/BackProjection.jl:62-102 calls one type-1 NUFFT per (coefficient,
 *    coil); NonuniformFFTs.jl 0.9 spreads with half-support m = 4, i.e. 8 grid
 *    points per dimension, on a sigma = 2 grid).
 *
 * Grid conventions: a grid has d dims n[0..d-1], n[0] fastest (Julia
 * column-major).  Point coordinates s are in grid units in [0, n_i); the ES
 * kernel phi(t) = exp(beta*(sqrt(1-(2t/w)^2)-1)) covers cells
 * l = ceil(s - w/2) ... ceil(s - w/2) + w - 1, wrapped periodically.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double es_kernel(double t, int w, double beta) {
    double z = 2.0 * t / (double)w;
    double a = 1.0 - z * z;
    if (a <= 0.0) return 0.0;
    return exp(beta * (sqrt(a) - 1.0));
}

/* weights for one point along one dim: base cell (unwrapped) and w values */
static inline int64_t es_weights(double s, int w, double beta, double *wt) {
    int64_t l0 = (int64_t)ceil(s - 0.5 * (double)w);
    for (int i = 0; i < w; ++i) wt[i] = es_kernel((double)(l0 + i) - s, w, beta);
    return l0;
}

static inline int64_t wrap(int64_t l, int64_t n) {
    l %= n;
    return l < 0 ? l + n : l;
}

/*
 * FP64 spreading of nch complex channels.
 *   s     : M*d point coordinates (point-major: s[m*d + i])
 *   vals  : complex values, channel-major: vals[2*(ch*M + m) + {0,1}]
 *   grid  : output, zeroed here, channel-major grids of prod(n) complex cells
 * Parallel over channels (each thread owns whole channel grids).
 */
void es_spread_f64(int d, int64_t M, const double *s, int nch, const double *vals,
                   const int64_t *n, int w, double beta, double *grid) {
    int64_t G = 1;
    for (int i = 0; i < d; ++i) G *= n[i];
    memset(grid, 0, sizeof(double) * 2 * (size_t)G * (size_t)nch);
    /* precompute weights and base cells */
    double *wt = (double *)malloc(sizeof(double) * (size_t)M * d * w);
    int64_t *base = (int64_t *)malloc(sizeof(int64_t) * (size_t)M * d);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < M; ++m)
        for (int i = 0; i < d; ++i)
            base[m * d + i] = es_weights(s[m * d + i], w, beta, wt + ((size_t)m * d + i) * w);

#pragma omp parallel for schedule(dynamic, 1)
    for (int ch = 0; ch < nch; ++ch) {
        double *g = grid + 2 * (size_t)G * ch;
        const double *v = vals + 2 * (size_t)M * ch;
        int64_t idx1[64], idx2[64], idx0[64];
        for (int64_t m = 0; m < M; ++m) {
            double vr = v[2 * m], vi = v[2 * m + 1];
            if (vr == 0.0 && vi == 0.0) continue;
            const double *w0 = wt + ((size_t)m * d + 0) * w;
            for (int a = 0; a < w; ++a) idx0[a] = wrap(base[m * d] + a, n[0]);
            if (d == 1) {
                for (int a = 0; a < w; ++a) {
                    g[2 * idx0[a]] += vr * w0[a];
                    g[2 * idx0[a] + 1] += vi * w0[a];
                }
                continue;
            }
            const double *w1 = wt + ((size_t)m * d + 1) * w;
            for (int b = 0; b < w; ++b) idx1[b] = wrap(base[m * d + 1] + b, n[1]);
            if (d == 2) {
                for (int b = 0; b < w; ++b) {
                    double fr = vr * w1[b], fi = vi * w1[b];
                    double *row = g + 2 * (idx1[b] * n[0]);
                    for (int a = 0; a < w; ++a) {
                        row[2 * idx0[a]] += fr * w0[a];
                        row[2 * idx0[a] + 1] += fi * w0[a];
                    }
                }
                continue;
            }
            const double *w2 = wt + ((size_t)m * d + 2) * w;
            for (int c = 0; c < w; ++c) idx2[c] = wrap(base[m * d + 2] + c, n[2]);
            for (int c = 0; c < w; ++c) {
                for (int b = 0; b < w; ++b) {
                    double f = w2[c] * w1[b];
                    double fr = vr * f, fi = vi * f;
                    double *row = g + 2 * ((idx2[c] * n[1] + idx1[b]) * n[0]);
                    for (int a = 0; a < w; ++a) {
                        row[2 * idx0[a]] += fr * w0[a];
                        row[2 * idx0[a] + 1] += fi * w0[a];
                    }
                }
            }
        }
    }
    free(wt);
    free(base);
}

/*
 * FP64 interpolation (adjoint of spreading) of nch complex channels:
 *   out[ch][m] = sum_l grid[ch][l] * phi(l - s_m).  Parallel over points.
 */
void es_interp_f64(int d, int64_t M, const double *s, int nch, const double *grid,
                   const int64_t *n, int w, double beta, double *out) {
    int64_t G = 1;
    for (int i = 0; i < d; ++i) G *= n[i];
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < M; ++m) {
        double w0[64], w1[64], w2[64];
        int64_t idx0[64], idx1[64], idx2[64];
        int64_t b0 = es_weights(s[m * d], w, beta, w0);
        for (int a = 0; a < w; ++a) idx0[a] = wrap(b0 + a, n[0]);
        if (d >= 2) {
            int64_t b1 = es_weights(s[m * d + 1], w, beta, w1);
            for (int a = 0; a < w; ++a) idx1[a] = wrap(b1 + a, n[1]);
        }
        if (d >= 3) {
            int64_t b2 = es_weights(s[m * d + 2], w, beta, w2);
            for (int a = 0; a < w; ++a) idx2[a] = wrap(b2 + a, n[2]);
        }
        for (int ch = 0; ch < nch; ++ch) {
            const double *g = grid + 2 * (size_t)G * ch;
            double ar = 0.0, ai = 0.0;
            if (d == 1) {
                for (int a = 0; a < w; ++a) {
                    ar += g[2 * idx0[a]] * w0[a];
                    ai += g[2 * idx0[a] + 1] * w0[a];
                }
            } else if (d == 2) {
                for (int b = 0; b < w; ++b) {
                    const double *row = g + 2 * (idx1[b] * n[0]);
                    double rr = 0.0, ri = 0.0;
                    for (int a = 0; a < w; ++a) {
                        rr += row[2 * idx0[a]] * w0[a];
                        ri += row[2 * idx0[a] + 1] * w0[a];
                    }
                    ar += rr * w1[b];
                    ai += ri * w1[b];
                }
            } else {
                for (int c = 0; c < w; ++c)
                    for (int b = 0; b < w; ++b) {
                        const double *row = g + 2 * ((idx2[c] * n[1] + idx1[b]) * n[0]);
                        double rr = 0.0, ri = 0.0;
                        for (int a = 0; a < w; ++a) {
                            rr += row[2 * idx0[a]] * w0[a];
                            ri += row[2 * idx0[a] + 1] * w0[a];
                        }
                        ar += rr * w1[b] * w2[c];
                        ai += ri * w1[b] * w2[c];
                    }
            }
            out[2 * ((size_t)ch * M + m)] = ar;
            out[2 * ((size_t)ch * M + m) + 1] = ai;
        }
    }
}

/*
 * FP32 single-channel spreading for the CPU-baseline port: one call per
 * type-1 NUFFT, as the reference issues them.  Threads spread disjoint point
 * ranges into private grids that are then summed (the grid is small next to
 * the point set at the benchmark sizes).
 *   s    : M*d float32 coordinates in grid units
 *   vals : M complex64 (interleaved)
 *   grid : prod(n) complex64, overwritten
 *   scratch: nthreads * prod(n) complex64 caller-provided
 */
void spread_f32_1ch(int d, int64_t M, const float *s, const float *vals, const int64_t *n,
                    int w, float beta, float *grid, float *scratch, int nthreads) {
    int64_t G = 1;
    for (int i = 0; i < d; ++i) G *= n[i];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        float *g = scratch + 2 * (size_t)G * tid;
        memset(g, 0, sizeof(float) * 2 * (size_t)G);
        int64_t lo = M * tid / nth, hi = M * (tid + 1) / nth;
        float w0[64], w1[64], w2[64];
        int64_t i0[64], i1[64], i2[64];
        for (int64_t m = lo; m < hi; ++m) {
            float vr = vals[2 * m], vi = vals[2 * m + 1];
            for (int dim = 0; dim < d; ++dim) {
                float sc = s[m * d + dim];
                int64_t l0 = (int64_t)ceilf(sc - 0.5f * (float)w);
                float *wt = dim == 0 ? w0 : (dim == 1 ? w1 : w2);
                int64_t *ix = dim == 0 ? i0 : (dim == 1 ? i1 : i2);
                for (int a = 0; a < w; ++a) {
                    float z = 2.0f * ((float)(l0 + a) - sc) / (float)w;
                    float q = 1.0f - z * z;
                    wt[a] = q > 0.f ? expf(beta * (sqrtf(q) - 1.0f)) : 0.f;
                    ix[a] = wrap(l0 + a, n[dim]);
                }
            }
            if (d == 2) {
                for (int b = 0; b < w; ++b) {
                    float fr = vr * w1[b], fi = vi * w1[b];
                    float *row = g + 2 * (i1[b] * n[0]);
                    for (int a = 0; a < w; ++a) {
                        row[2 * i0[a]] += fr * w0[a];
                        row[2 * i0[a] + 1] += fi * w0[a];
                    }
                }
            } else {
                for (int c = 0; c < w; ++c)
                    for (int b = 0; b < w; ++b) {
                        float f = w1[b] * w2[c];
                        float fr = vr * f, fi = vi * f;
                        float *row = g + 2 * ((i2[c] * n[1] + i1[b]) * n[0]);
                        for (int a = 0; a < w; ++a) {
                            row[2 * i0[a]] += fr * w0[a];
                            row[2 * i0[a] + 1] += fi * w0[a];
                        }
                    }
            }
        }
#pragma omp barrier
        /* reduce private grids into the output, split over cells */
        int64_t c0 = G * tid / nth, c1 = G * (tid + 1) / nth;
        for (int64_t c = c0; c < c1; ++c) {
            float ar = 0.f, ai = 0.f;
            for (int t = 0; t < nth; ++t) {
                ar += scratch[2 * ((size_t)G * t + c)];
                ai += scratch[2 * ((size_t)G * t + c) + 1];
            }
            grid[2 * c] = ar;
            grid[2 * c + 1] = ai;
        }
    }
}
