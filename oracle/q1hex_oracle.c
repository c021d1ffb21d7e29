/*
 * ORACLE -- test infrastructure only.  Nothing in the product links or calls
 * this file: tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it as the checker.
 *
 * A CPU restatement, in plain C with OpenMP, of the reference's native
 * hexahedral Q1 conduction path:
 *   - lane_sum / ordered_sum      cpp/include/pcbcore/lane_sum.hpp:18-41
 *   - node-owned line gather      src/thermal/matrix_free_mpir_fem/native/hex_q1.cpp:90-193
 *     (generic gather for the two end nodes of a line, the two-element
 *      x-gather with a free-mask multiply for the interior)
 *   - static line partition       hex_q1.cpp:195-205
 *   - FP32 Jacobi PCG in one SPMD region with per-line FP64 lane-sum
 *     partials, FTZ/DAZ           hex_q1.cpp:323-499 (Jacobi branch)
 *   - FP64 MPIR outer loop        hex_q1.cpp:837-925
 *   - plain FP64 Jacobi PCG       the same PCG in double precision (the
 *                                 "plain FP64" solve of the 64x64x8 case)
 *
 * Every result is independent of the thread count: lines are split
 * statically and every inner product is a per-line lane sum followed by an
 * ordered sum over lines.  Build with -ffp-contract=off (no FMA contraction),
 * as the reference's CPU kernels are built.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>
#include <pmmintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t idx_t;

/* ---------------------------------------------------------------- sums */

/* Eight index lanes, combined in a fixed tree (lane_sum.hpp:18-31). */
#define LANE_SUM(result, count, TERM)                                          \
    do {                                                                       \
        double acc_[8] = {0, 0, 0, 0, 0, 0, 0, 0};                             \
        idx_t i_ = 0;                                                          \
        for (; i_ + 8 <= (count); i_ += 8) {                                   \
            for (int k_ = 0; k_ < 8; ++k_) {                                   \
                const idx_t i = i_ + k_;                                       \
                acc_[k_] += (TERM);                                            \
            }                                                                  \
        }                                                                      \
        for (int k_ = 0; i_ < (count); ++i_, ++k_) {                           \
            const idx_t i = i_;                                                \
            acc_[k_] += (TERM);                                                \
        }                                                                      \
        (result) = ((acc_[0] + acc_[1]) + (acc_[2] + acc_[3])) +               \
                   ((acc_[4] + acc_[5]) + (acc_[6] + acc_[7]));                \
    } while (0)

double oracle_lane_sum(const double *v, idx_t n) {
    double s;
    LANE_SUM(s, n, v[i]);
    return s;
}

/* ---------------------------------------------------------- operator */

typedef struct {
    int slabs, rows, cols;
    int threads;
} dims_t;

static inline int node_rows(const dims_t *d) { return d->rows + 1; }
static inline int node_cols(const dims_t *d) { return d->cols + 1; }
static inline int line_count(const dims_t *d) { return (d->slabs + 1) * (d->rows + 1); }

static void thread_lines(int lines, int *b, int *e) {
#ifdef _OPENMP
    const int nt = omp_get_num_threads();
    const int t = omp_get_thread_num();
#else
    const int nt = 1, t = 0;
#endif
    *b = (int)(((long long)lines * t) / nt);
    *e = (int)(((long long)lines * (t + 1)) / nt);
}

/* The operator body, written once for both precisions. */
#define DEFINE_OPERATOR(T, SUF)                                                                      \
    static T gather_##SUF(const dims_t *d, const T *coef, const T *unit, const T *robin,            \
                          const uint8_t *free_nodes, const T *x, int z, int y, int xi) {             \
        const int nr = node_rows(d), nc = node_cols(d), plane = nr * nc;                             \
        const idx_t node = ((idx_t)z * nr + y) * nc + xi;                                            \
        if (free_nodes[node] == 0) return x[node];                                                   \
        const int z0 = z > 0 ? z - 1 : 0, z1 = z < d->slabs ? z : d->slabs - 1;                      \
        const int y0 = y > 0 ? y - 1 : 0, y1 = y < d->rows ? y : d->rows - 1;                        \
        const int x0 = xi > 0 ? xi - 1 : 0, x1 = xi < d->cols ? xi : d->cols - 1;                    \
        const idx_t ne = (idx_t)d->slabs * d->rows * d->cols;                                        \
        T acc = 0;                                                                                   \
        for (int ez = z0; ez <= z1; ++ez)                                                            \
            for (int ey = y0; ey <= y1; ++ey)                                                        \
                for (int ex = x0; ex <= x1; ++ex) {                                                  \
                    const int lr = 4 * (z - ez) + 2 * (y - ey) + (xi - ex);                          \
                    const idx_t e = ((idx_t)ez * d->rows + ey) * d->cols + ex;                       \
                    const T a = coef[e], b = coef[ne + e], c = coef[2 * ne + e];                     \
                    const idx_t corner = ((idx_t)ez * nr + ey) * nc + ex;                            \
                    for (int k = 0; k < 8; ++k) {                                                    \
                        const idx_t cn = corner + (idx_t)(k >> 2) * plane + ((k >> 1) & 1) * nc + (k & 1); \
                        if (free_nodes[cn] == 0) continue;                                           \
                        acc += ((a * unit[8 * lr + k]) + (b * unit[64 + 8 * lr + k]) +               \
                                (c * unit[128 + 8 * lr + k])) * x[cn];                               \
                    }                                                                                \
                }                                                                                    \
        return acc + robin[node] * x[node];                                                          \
    }                                                                                                \
                                                                                                     \
    static void apply_lines_##SUF(const dims_t *d, const T *coef, const T *unit, const T *robin,    \
                                  const uint8_t *free_nodes, const T *free_mask, const T *x, T *out, \
                                  int lb, int le) {                                                  \
        const int nr = node_rows(d), nc = node_cols(d), plane = nr * nc;                             \
        const idx_t ne = (idx_t)d->slabs * d->rows * d->cols;                                        \
        for (int line = lb; line < le; ++line) {                                                     \
            const int z = line / nr, y = line - z * nr;                                              \
            const idx_t base = (idx_t)line * nc;                                                     \
            out[base] = gather_##SUF(d, coef, unit, robin, free_nodes, x, z, y, 0);                  \
            out[base + d->cols] = gather_##SUF(d, coef, unit, robin, free_nodes, x, z, y, d->cols);  \
            if (d->cols < 2) continue;                                                               \
            const int z0 = z > 0 ? z - 1 : 0, z1 = z < d->slabs ? z : d->slabs - 1;                  \
            const int y0 = y > 0 ? y - 1 : 0, y1 = y < d->rows ? y : d->rows - 1;                    \
            T *o = out + base;                                                                       \
            for (int xi = 1; xi < d->cols; ++xi) o[xi] = 0;                                          \
            for (int ez = z0; ez <= z1; ++ez)                                                        \
                for (int ey = y0; ey <= y1; ++ey) {                                                  \
                    const int lr0 = 4 * (z - ez) + 2 * (y - ey);                                     \
                    const idx_t erow = ((idx_t)ez * d->rows + ey) * d->cols;                         \
                    const idx_t crow = ((idx_t)ez * nr + ey) * nc;                                   \
                    for (int xi = 1; xi < d->cols; ++xi) {                                           \
                        T acc = 0;                                                                   \
                        for (int dx = 0; dx < 2; ++dx) {                                             \
                            const int ex = xi - 1 + dx, lr = lr0 + (1 - dx);                         \
                            const T a = coef[erow + ex], b = coef[ne + erow + ex],                   \
                                    c = coef[2 * ne + erow + ex];                                    \
                            const idx_t corner = crow + ex;                                          \
                            for (int k = 0; k < 8; ++k) {                                            \
                                const idx_t cn = corner + (idx_t)(k >> 2) * plane + ((k >> 1) & 1) * nc + (k & 1); \
                                const T w = (a * unit[8 * lr + k]) + (b * unit[64 + 8 * lr + k]) +   \
                                            (c * unit[128 + 8 * lr + k]);                            \
                                acc += w * (x[cn] * free_mask[cn]);                                  \
                            }                                                                        \
                        }                                                                            \
                        o[xi] += acc;                                                                \
                    }                                                                                \
                }                                                                                    \
            for (int xi = 1; xi < d->cols; ++xi) {                                                   \
                const idx_t n = base + xi;                                                           \
                const T f = free_mask[n];                                                            \
                o[xi] = (f * (o[xi] + robin[n] * x[n])) + (((T)1 - f) * x[n]);                       \
            }                                                                                        \
        }                                                                                            \
    }

DEFINE_OPERATOR(float, f32)
DEFINE_OPERATOR(double, f64)

#define DEFINE_APPLY(T, SUF)                                                                         \
    void oracle_apply_##SUF(const T *coef, const T *unit, const T *robin, const uint8_t *free_nodes, \
                            const T *free_mask, int slabs, int rows, int cols, const T *x, T *y,     \
                            int threads) {                                                           \
        dims_t d = {slabs, rows, cols, threads < 1 ? 1 : threads};                                   \
        const int lines = line_count(&d);                                                            \
        const int team = d.threads < lines ? d.threads : lines;                                      \
        _Pragma("omp parallel num_threads(team) if (team > 1)")                                      \
        {                                                                                            \
            int b, e;                                                                                \
            thread_lines(lines, &b, &e);                                                             \
            apply_lines_##SUF(&d, coef, unit, robin, free_nodes, free_mask, x, y, b, e);             \
        }                                                                                            \
    }

DEFINE_APPLY(float, f32)
DEFINE_APPLY(double, f64)

/* ----------------------------------------------------------------- PCG */

/* Inner Jacobi PCG of hex_q1.cpp:323-499, for T = float (MPIR inner solve,
 * FTZ/DAZ on) or T = double (the plain FP64 solve).  ``rhs`` is FP64 and is
 * rounded to T; the solution ``xsol`` starts at zero.  Returns 0, or 1 when
 * the curvature was not finite and positive. */
#define DEFINE_PCG(T, SUF, FLUSH)                                                                     \
    int oracle_pcg_##SUF(const T *coef, const T *unit, const T *robin, const uint8_t *free_nodes,     \
                         const T *free_mask, int slabs, int rows, int cols, const double *rhs_in,     \
                         const T *diag, double tol, int max_its, T *xsol, int *iterations_out,        \
                         int *applications_out, double *relative_out, int threads) {                  \
        dims_t d = {slabs, rows, cols, threads < 1 ? 1 : threads};                                    \
        const int nc = node_cols(&d);                                                                 \
        const int lines = line_count(&d);                                                             \
        const idx_t n = (idx_t)lines * nc;                                                            \
        const int team = d.threads < lines ? d.threads : lines;                                       \
        T *r = (T *)calloc((size_t)n, sizeof(T));                                                     \
        T *z = (T *)calloc((size_t)n, sizeof(T));                                                     \
        T *p = (T *)calloc((size_t)n, sizeof(T));                                                     \
        T *q = (T *)calloc((size_t)n, sizeof(T));                                                     \
        double *part = (double *)calloc(3 * (size_t)lines, sizeof(double));                          \
        int its = 0, apps = 0, bad = 0;                                                               \
        double rel = 1.0;                                                                             \
        _Pragma("omp parallel num_threads(team) if (team > 1)")                                       \
        {                                                                                             \
            unsigned int csr_saved = _mm_getcsr();                                                    \
            if (FLUSH) _mm_setcsr(csr_saved | 0x8040u);                                               \
            int lb, le;                                                                               \
            thread_lines(lines, &lb, &le);                                                            \
            const idx_t lo = (idx_t)lb * nc, hi = (idx_t)le * nc;                                     \
            int l_its = 0, l_apps = 0, l_bad = 0;                                                     \
            double l_rel = 1.0;                                                                       \
            for (idx_t i = lo; i < hi; ++i) {                                                         \
                r[i] = (T)rhs_in[i];                                                                  \
                xsol[i] = 0;                                                                          \
            }                                                                                         \
            /* store: per-line lane sums of a*b into slot s */                                        \
            _Pragma("omp barrier")                                                                    \
            for (int line = lb; line < le; ++line) {                                                  \
                const T *al = r + (idx_t)line * nc;                                                   \
                double s;                                                                             \
                LANE_SUM(s, nc, (double)al[i] * (double)al[i]);                                       \
                part[line] = s;                                                                       \
            }                                                                                         \
            _Pragma("omp barrier")                                                                    \
            double rhs_norm = 0;                                                                      \
            for (int l = 0; l < lines; ++l) rhs_norm += part[l];                                      \
            rhs_norm = sqrt(rhs_norm);                                                                \
            if (rhs_norm == 0.0) {                                                                    \
                l_rel = 0.0;                                                                          \
            } else {                                                                                  \
                for (idx_t i = lo; i < hi; ++i) z[i] = r[i] / diag[i];                                \
                for (idx_t i = lo; i < hi; ++i) p[i] = z[i];                                          \
                for (int line = lb; line < le; ++line) {                                              \
                    const T *al = r + (idx_t)line * nc, *bl = z + (idx_t)line * nc;                   \
                    double s;                                                                         \
                    LANE_SUM(s, nc, (double)al[i] * (double)bl[i]);                                   \
                    part[2 * lines + line] = s;                                                       \
                }                                                                                     \
                _Pragma("omp barrier")                                                                \
                double rz = 0;                                                                        \
                for (int l = 0; l < lines; ++l) rz += part[2 * lines + l];                            \
                while (l_its < max_its) {                                                             \
                    _Pragma("omp barrier")                                                            \
                    apply_lines_##SUF(&d, coef, unit, robin, free_nodes, free_mask, p, q, lb, le);    \
                    ++l_apps;                                                                         \
                    for (int line = lb; line < le; ++line) {                                          \
                        const T *al = p + (idx_t)line * nc, *bl = q + (idx_t)line * nc;               \
                        double s;                                                                     \
                        LANE_SUM(s, nc, (double)al[i] * (double)bl[i]);                               \
                        part[lines + line] = s;                                                       \
                    }                                                                                 \
                    _Pragma("omp barrier")                                                            \
                    double curv = 0;                                                                  \
                    for (int l = 0; l < lines; ++l) curv += part[lines + l];                          \
                    if (!isfinite(curv) || curv <= 0.0) {                                             \
                        l_bad = 1;                                                                    \
                        break;                                                                        \
                    }                                                                                 \
                    const T alpha = (T)(rz / curv);                                                   \
                    for (idx_t i = lo; i < hi; ++i) {                                                 \
                        xsol[i] += alpha * p[i];                                                      \
                        r[i] -= alpha * q[i];                                                         \
                    }                                                                                 \
                    for (int line = lb; line < le; ++line) {                                          \
                        const T *al = r + (idx_t)line * nc;                                           \
                        double s;                                                                     \
                        LANE_SUM(s, nc, (double)al[i] * (double)al[i]);                               \
                        part[line] = s;                                                               \
                    }                                                                                 \
                    _Pragma("omp barrier")                                                            \
                    double rr = 0;                                                                    \
                    for (int l = 0; l < lines; ++l) rr += part[l];                                    \
                    l_rel = sqrt(rr) / rhs_norm;                                                      \
                    ++l_its;                                                                          \
                    if (l_rel <= tol) break;                                                          \
                    for (idx_t i = lo; i < hi; ++i) z[i] = r[i] / diag[i];                            \
                    for (int line = lb; line < le; ++line) {                                          \
                        const T *al = r + (idx_t)line * nc, *bl = z + (idx_t)line * nc;               \
                        double s;                                                                     \
                        LANE_SUM(s, nc, (double)al[i] * (double)bl[i]);                               \
                        part[2 * lines + line] = s;                                                   \
                    }                                                                                 \
                    _Pragma("omp barrier")                                                            \
                    double next_rz = 0;                                                               \
                    for (int l = 0; l < lines; ++l) next_rz += part[2 * lines + l];                   \
                    if (!isfinite(next_rz) || rz == 0.0) break;                                       \
                    const T beta = (T)(next_rz / rz);                                                 \
                    for (idx_t i = lo; i < hi; ++i) p[i] = z[i] + beta * p[i];                        \
                    rz = next_rz;                                                                     \
                }                                                                                     \
            }                                                                                         \
            _Pragma("omp master")                                                                     \
            {                                                                                         \
                its = l_its;                                                                          \
                apps = l_apps;                                                                        \
                rel = l_rel;                                                                          \
                bad = l_bad;                                                                          \
            }                                                                                         \
            _mm_setcsr(csr_saved);                                                                    \
        }                                                                                             \
        free(r);                                                                                      \
        free(z);                                                                                      \
        free(p);                                                                                      \
        free(q);                                                                                      \
        free(part);                                                                                   \
        *iterations_out = its;                                                                        \
        *applications_out = apps;                                                                     \
        *relative_out = rel;                                                                          \
        return bad;                                                                                   \
    }

DEFINE_PCG(float, f32, 1)
DEFINE_PCG(double, f64, 0)

/* ---------------------------------------------------------------- MPIR */

/* The FP64 outer loop of hex_q1.cpp:837-925: r = b - A64 x, lane-sum norm,
 * stop at ||r|| <= atol + rtol ||b||, else an FP32 (inner_f64 = 0) or FP64
 * (inner_f64 = 1) Jacobi PCG correction and x += correction.  ``history``
 * receives (outer, fp64 relative residual before, inner iterations, inner
 * relative residual) per correction, 4 doubles each, at most max_outer rows.
 * ``stats`` receives converged, outer, inner, high applications, low
 * applications, not_spd.  Returns the final relative residual. */
double oracle_solve_mpir(const float *coef32, const float *unit32, const float *robin32,
                         const float *mask32, const float *diag32, const double *coef64,
                         const double *unit64, const double *robin64, const double *mask64,
                         const double *diag64, const uint8_t *free_nodes, int slabs, int rows,
                         int cols, const double *rhs, double *x, double rtol, double atol,
                         double inner_tol, int max_outer, int max_inner, int inner_f64,
                         int threads, int *stats, double *history) {
    dims_t d = {slabs, rows, cols, threads < 1 ? 1 : threads};
    const idx_t n = (idx_t)line_count(&d) * node_cols(&d);
    double bn;
    LANE_SUM(bn, n, rhs[i] * rhs[i]);
    const double scale = sqrt(bn) > 0.0 ? sqrt(bn) : 1.0;
    const double target = atol + rtol * scale;
    double *ax = (double *)calloc((size_t)n, sizeof(double));
    double *res = (double *)calloc((size_t)n, sizeof(double));
    float *corr32 = (float *)calloc((size_t)n, sizeof(float));
    double *corr64 = (double *)calloc((size_t)n, sizeof(double));
    int converged = 0, outer_done = 0, inner_total = 0, high = 0, low = 0, not_spd = 0;
    double rel = 1.0;
    for (int outer = 0; outer <= max_outer; ++outer) {
        outer_done = outer;
        oracle_apply_f64(coef64, unit64, robin64, free_nodes, mask64, slabs, rows, cols, x, ax,
                         threads);
        ++high;
        for (idx_t i = 0; i < n; ++i) res[i] = rhs[i] - ax[i];
        double rn;
        LANE_SUM(rn, n, res[i] * res[i]);
        rn = sqrt(rn);
        rel = rn / scale;
        if (rn <= target) {
            converged = 1;
            break;
        }
        if (outer == max_outer) break;
        int its = 0, apps = 0;
        double irel = 1.0;
        int bad;
        if (inner_f64) {
            bad = oracle_pcg_f64(coef64, unit64, robin64, free_nodes, mask64, slabs, rows, cols,
                                 res, diag64, inner_tol, max_inner, corr64, &its, &apps, &irel,
                                 threads);
        } else {
            bad = oracle_pcg_f32(coef32, unit32, robin32, free_nodes, mask32, slabs, rows, cols,
                                 res, diag32, inner_tol, max_inner, corr32, &its, &apps, &irel,
                                 threads);
        }
        if (bad) {
            not_spd = 1;
            break;
        }
        inner_total += its;
        low += apps;
        history[4 * outer + 0] = outer + 1;
        history[4 * outer + 1] = rel;
        history[4 * outer + 2] = its;
        history[4 * outer + 3] = irel;
        if (inner_f64) {
            for (idx_t i = 0; i < n; ++i) x[i] += corr64[i];
        } else {
            for (idx_t i = 0; i < n; ++i) x[i] += (double)corr32[i];
        }
    }
    free(ax);
    free(res);
    free(corr32);
    free(corr64);
    stats[0] = converged;
    stats[1] = outer_done;
    stats[2] = inner_total;
    stats[3] = high;
    stats[4] = low;
    stats[5] = not_spd;
    return rel;
}

int oracle_openmp(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}
