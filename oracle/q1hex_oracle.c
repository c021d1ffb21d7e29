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
 *   - FP32 PCG in one SPMD region with per-line FP64 lane-sum
 *     partials, FTZ/DAZ           hex_q1.cpp:323-499
 *   - its two-level preconditioner (Jacobi plus a patch-constant coarse
 *     correction through a dense FP32 inverse)   hex_q1.cpp:388-432
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
                    /* x lanes are independent: vectorising them keeps every sum's order */          \
                    _Pragma("omp simd")                                                              \
                    for (int xi = 1; xi < d->cols; ++xi) {                                           \
                        T acc = 0;                                                                   \
                        _Pragma("GCC unroll 2")                                                      \
                        for (int dx = 0; dx < 2; ++dx) {                                             \
                            const int ex = xi - 1 + dx, lr = lr0 + (1 - dx);                         \
                            const T a = coef[erow + ex], b = coef[ne + erow + ex],                   \
                                    c = coef[2 * ne + erow + ex];                                    \
                            const idx_t corner = crow + ex;                                          \
                            _Pragma("GCC unroll 8")                                                  \
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
            _Pragma("omp simd")                                                                      \
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

/* ----------------------------------------------------- preconditioners */

/* The coarse stage of the two-level preconditioner (hex_q1.cpp:388-432):
 * patch-constant aggregates of ``block`` x ``block`` nodes in each node
 * layer, and the dense FP32 inverse of their Galerkin matrix, row-major. */
#define DEFINE_COARSE(T, SUF)                                                                        \
    typedef struct {                                                                                 \
        int block, coarse_rows, coarse_cols;                                                         \
        idx_t size;                                                                                  \
        const T *inverse;                                                                            \
        T *restricted, *corrected;                                                                   \
    } coarse_##SUF;                                                                                  \
                                                                                                     \
    /* z = r / diag on this thread's lines, then, with a coarse stage,         \
     * z += P C^-1 P^T r: each coarse value is summed by the thread that owns  \
     * its index (patch rows, then columns, in node order, accumulated in     \
     * FP64 from FP32 products) so the result is the same at any thread       \
     * count.  Called by every thread of the team: the barriers are the       \
     * team's. */                                                                                    \
    static void precondition_##SUF(const dims_t *d, const coarse_##SUF *cs, const T *r, const T *diag, \
                                   const T *free_mask, T *z, int lb, int le) {                       \
        const int nr = node_rows(d), nc = node_cols(d);                                              \
        for (idx_t i = (idx_t)lb * nc; i < (idx_t)le * nc; ++i) z[i] = r[i] / diag[i];               \
        if (cs == NULL) return;                                                                      \
        int nt = 1, tid = 0;                                                                         \
        OMP_TEAM(nt, tid);                                                                           \
        const idx_t cb = (cs->size * tid) / nt, ce = (cs->size * (tid + 1)) / nt;                    \
        const idx_t per_layer = (idx_t)cs->coarse_rows * cs->coarse_cols;                            \
        const int b = cs->block;                                                                     \
        for (idx_t c = cb; c < ce; ++c) {                                                            \
            const int layer = (int)(c / per_layer);                                                  \
            const int yc = (int)((c - layer * per_layer) / cs->coarse_cols);                         \
            const int xc = (int)(c - layer * per_layer - (idx_t)yc * cs->coarse_cols);               \
            const int y_end = (yc + 1) * b < nr ? (yc + 1) * b : nr;                                 \
            const int x_end = (xc + 1) * b < nc ? (xc + 1) * b : nc;                                 \
            double acc = 0.0;                                                                        \
            for (int y = yc * b; y < y_end; ++y) {                                                   \
                const idx_t row = ((idx_t)layer * nr + y) * nc;                                      \
                for (int x = xc * b; x < x_end; ++x) acc += (double)(r[row + x] * free_mask[row + x]); \
            }                                                                                        \
            cs->restricted[c] = (T)acc;                                                              \
        }                                                                                            \
        _Pragma("omp barrier")                                                                       \
        for (idx_t c = cb; c < ce; ++c) {                                                            \
            const T *row = cs->inverse + c * cs->size;                                               \
            const T *v = cs->restricted;                                                             \
            double s;                                                                                \
            LANE_SUM(s, cs->size, (double)row[i] * (double)v[i]);                                    \
            cs->corrected[c] = (T)s;                                                                 \
        }                                                                                            \
        _Pragma("omp barrier")                                                                       \
        for (int line = lb; line < le; ++line) {                                                     \
            const int layer = line / nr, y = line - layer * nr;                                      \
            const idx_t cbase = ((idx_t)layer * cs->coarse_rows + y / b) * cs->coarse_cols;          \
            T *zl = z + (idx_t)line * nc;                                                            \
            const T *ml = free_mask + (idx_t)line * nc;                                              \
            for (int x = 0; x < nc; ++x) zl[x] += ml[x] * cs->corrected[cbase + x / b];              \
        }                                                                                            \
    }                                                                                                \
                                                                                                     \
    static int coarse_init_##SUF(coarse_##SUF *cs, const dims_t *d, int block, const T *inverse) {   \
        if (inverse == NULL) return 0;                                                               \
        if (block < 1) return -1;                                                                    \
        cs->block = block;                                                                           \
        cs->coarse_rows = (node_rows(d) + block - 1) / block;                                        \
        cs->coarse_cols = (node_cols(d) + block - 1) / block;                                        \
        cs->size = (idx_t)(d->slabs + 1) * cs->coarse_rows * cs->coarse_cols;                        \
        cs->inverse = inverse;                                                                       \
        cs->restricted = (T *)calloc((size_t)cs->size, sizeof(T));                                   \
        cs->corrected = (T *)calloc((size_t)cs->size, sizeof(T));                                    \
        return 1;                                                                                    \
    }                                                                                                \
                                                                                                     \
    static void coarse_free_##SUF(coarse_##SUF *cs) {                                                \
        free(cs->restricted);                                                                        \
        free(cs->corrected);                                                                         \
    }

#ifdef _OPENMP
#define OMP_TEAM(nt, tid)             \
    do {                              \
        (nt) = omp_get_num_threads(); \
        (tid) = omp_get_thread_num(); \
    } while (0)
#else
#define OMP_TEAM(nt, tid) \
    do {                  \
    } while (0)
#endif

DEFINE_COARSE(float, f32)
DEFINE_COARSE(double, f64)

/* ----------------------------------------------------------------- PCG */

/* Inner PCG of hex_q1.cpp:323-499, for T = float (MPIR inner solve, FTZ/DAZ
 * on) or T = double (the plain FP64 solve), with the Jacobi preconditioner
 * or, given a coarse stage, the two-level one.  ``rhs`` is FP64 and is
 * rounded to T; the solution ``xsol`` starts at zero.  Returns 0, or 1 when
 * the curvature was not finite and positive. */
#define DEFINE_PCG(T, SUF, FLUSH)                                                                     \
    static int pcg_core_##SUF(const T *coef, const T *unit, const T *robin, const uint8_t *free_nodes, \
                              const T *free_mask, int slabs, int rows, int cols, const double *rhs_in, \
                              const T *diag, const coarse_##SUF *cs, double tol, int max_its, T *xsol, \
                              int *iterations_out, int *applications_out, double *relative_out,       \
                              int threads) {                                                          \
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
                precondition_##SUF(&d, cs, r, diag, free_mask, z, lb, le);                            \
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
                    precondition_##SUF(&d, cs, r, diag, free_mask, z, lb, le);                        \
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

#define DEFINE_PCG_EXPORTS(T, SUF)                                                                    \
    int oracle_pcg_##SUF(const T *coef, const T *unit, const T *robin, const uint8_t *free_nodes,     \
                         const T *free_mask, int slabs, int rows, int cols, const double *rhs_in,     \
                         const T *diag, double tol, int max_its, T *xsol, int *iterations_out,        \
                         int *applications_out, double *relative_out, int threads) {                  \
        return pcg_core_##SUF(coef, unit, robin, free_nodes, free_mask, slabs, rows, cols, rhs_in,    \
                              diag, NULL, tol, max_its, xsol, iterations_out, applications_out,       \
                              relative_out, threads);                                                 \
    }                                                                                                 \
                                                                                                      \
    /* The same PCG with the two-level preconditioner: ``inverse`` is the     \
     * dense inverse of the Galerkin matrix of ``block`` x ``block`` patches  \
     * (row-major, (slabs + 1) * ceil(rows + 1, block) * ceil(cols + 1, block) \
     * squared).  Returns -1 for a block below one. */                                                \
    int oracle_pcg_two_level_##SUF(const T *coef, const T *unit, const T *robin,                      \
                                   const uint8_t *free_nodes, const T *free_mask, int slabs,          \
                                   int rows, int cols, const double *rhs_in, const T *diag,           \
                                   int block, const T *inverse, double tol, int max_its, T *xsol,     \
                                   int *iterations_out, int *applications_out,                        \
                                   double *relative_out, int threads) {                               \
        dims_t d = {slabs, rows, cols, threads};                                                      \
        coarse_##SUF cs;                                                                              \
        if (coarse_init_##SUF(&cs, &d, block, inverse) != 1) return -1;                               \
        const int bad = pcg_core_##SUF(coef, unit, robin, free_nodes, free_mask, slabs, rows, cols,   \
                                       rhs_in, diag, &cs, tol, max_its, xsol, iterations_out,         \
                                       applications_out, relative_out, threads);                      \
        coarse_free_##SUF(&cs);                                                                       \
        return bad;                                                                                   \
    }

DEFINE_PCG_EXPORTS(float, f32)
DEFINE_PCG_EXPORTS(double, f64)

/* ---------------------------------------------------------------- MPIR */

/* The FP64 outer loop of hex_q1.cpp:837-925: r = b - A64 x, lane-sum norm,
 * stop at ||r|| <= atol + rtol ||b||, else an FP32 (inner_f64 = 0) or FP64
 * (inner_f64 = 1) Jacobi PCG correction and x += correction.  ``history``
 * receives (outer, fp64 relative residual before, inner iterations, inner
 * relative residual) per correction, 4 doubles each, at most max_outer rows.
 * ``stats`` receives converged, outer, inner, high applications, low
 * applications, not_spd.  Returns the final relative residual. */
static double solve_mpir_core(const float *coef32, const float *unit32, const float *robin32,
                              const float *mask32, const float *diag32, const double *coef64,
                              const double *unit64, const double *robin64, const double *mask64,
                              const double *diag64, const uint8_t *free_nodes, int slabs, int rows,
                              int cols, const coarse_f32 *cs, const double *rhs, double *x,
                              double rtol, double atol, double inner_tol, int max_outer,
                              int max_inner, int inner_f64, int threads, int *stats,
                              double *history) {
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
            bad = pcg_core_f32(coef32, unit32, robin32, free_nodes, mask32, slabs, rows, cols, res,
                               diag32, cs, inner_tol, max_inner, corr32, &its, &apps, &irel, threads);
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

/* The FP64 outer loop of hex_q1.cpp:837-925 over Jacobi-preconditioned
 * inner solves (see solve_mpir_core for the arguments and outputs). */
double oracle_solve_mpir(const float *coef32, const float *unit32, const float *robin32,
                         const float *mask32, const float *diag32, const double *coef64,
                         const double *unit64, const double *robin64, const double *mask64,
                         const double *diag64, const uint8_t *free_nodes, int slabs, int rows,
                         int cols, const double *rhs, double *x, double rtol, double atol,
                         double inner_tol, int max_outer, int max_inner, int inner_f64,
                         int threads, int *stats, double *history) {
    return solve_mpir_core(coef32, unit32, robin32, mask32, diag32, coef64, unit64, robin64, mask64,
                           diag64, free_nodes, slabs, rows, cols, NULL, rhs, x, rtol, atol, inner_tol,
                           max_outer, max_inner, inner_f64, threads, stats, history);
}

/* The same outer loop with the FP32 two-level inner PCG (hex_q1.cpp:837-925
 * with the preconditioner of :388-432); ``inverse`` as for
 * oracle_pcg_two_level_f32.  Returns -1.0 for a block below one. */
double oracle_solve_mpir_two_level(const float *coef32, const float *unit32, const float *robin32,
                                   const float *mask32, const float *diag32, const double *coef64,
                                   const double *unit64, const double *robin64,
                                   const double *mask64, const double *diag64,
                                   const uint8_t *free_nodes, int slabs, int rows, int cols,
                                   int block, const float *inverse, const double *rhs, double *x,
                                   double rtol, double atol, double inner_tol, int max_outer,
                                   int max_inner, int threads, int *stats, double *history) {
    dims_t d = {slabs, rows, cols, threads};
    coarse_f32 cs;
    if (coarse_init_f32(&cs, &d, block, inverse) != 1) return -1.0;
    const double rel = solve_mpir_core(coef32, unit32, robin32, mask32, diag32, coef64, unit64,
                                       robin64, mask64, diag64, free_nodes, slabs, rows, cols, &cs,
                                       rhs, x, rtol, atol, inner_tol, max_outer, max_inner, 0,
                                       threads, stats, history);
    coarse_free_f32(&cs);
    return rel;
}

int oracle_openmp(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}
