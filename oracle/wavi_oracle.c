/*
 * oracle/wavi_oracle.c - CPU oracle of WAVI.jl's velocity solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, single-threaded, the
 * reference algorithm of update_velocities! and of the update_state! /
 * update_thickness! steps around it.  Only the tests, the package's smoke
 * check and the CPU-baseline leg of the benchmark load it, and only as the
 * checker; the GPU library never links or calls it.
 *
 * Reference files followed (paths relative to the WAVI.jl repository):
 *   src/Models/update_velocities.jl      (all)
 *   src/utilities.jl                     1-241, 255-396, 436-466, 475-507, 537-552
 *   src/Wavelets/update_preconditioners.jl, update_wavelets.jl
 *   src/KroneckerProduct.jl              57-73 (evaluation order of A (x) B)
 *   src/Fields/ (all grid files)         operator construction
 *   src/Models/update_state.jl           7-199
 *   src/Simulations/run_simulation.jl    24-42
 *   IterativeSolvers.jl cg! (not vendored; version unpinned by the
 *   reference's Project.toml; restated from its documented v0.9 algorithm:
 *   abstol/reltol keywords, reltol = sqrt(eps), stop when ||r|| <= max(reltol
 *   ||r0||, abstol), prev_residual initialised to one).
 *
 * Floating-point order.  Every sum is written in the order in which the
 * reference's sparse / Kronecker products accumulate it (KroneckerProduct.jl
 * picks the factor applied first from the operand shapes; those rules are
 * reproduced).  Norms and dot products, which the reference hands to BLAS,
 * use one fixed blocked tree (canonical_dot) that the GPU library uses too,
 * and x^y uses det_pow below, a double-double log/exp built only from IEEE
 * +,-,*,/,sqrt and fma, so that CPU and GPU results can agree bit for bit.
 * Build with -ffp-contract=off (see oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/wavi_hip.h"

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Deterministic power function.                                       */
/* x^y = exp(y log x) with log and the product carried in double-double.
 * log: x = 2^k m, m in [1/sqrt2, sqrt2), log m = 2 atanh(s), s = (m-1)/(m+1),
 * atanh by its Taylor series (coefficients 1/(2n+1)); exp: Cody-Waite
 * reduction by ln 2 and the Taylor series of e^r on |r| <= ln2/2.
 * Accuracy is about one ulp; tests bound it against the C library's pow. */

static const double LN2_HI = 0x1.62e42fefa39efp-1;
static const double LN2_LO = 0x1.abc9e3b39803fp-56;
static const double INV_LN2 = 0x1.71547652b82fep+0;

static double dd_log_hi(double x, double *lo)
{
    int k;
    double m = frexp(x, &k); /* x = m 2^k, m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        k = k - 1;
    }
    double f = m - 1.0;                 /* exact (Sterbenz) */
    double d_hi = 2.0 + f;
    double d_lo = f - (d_hi - 2.0);     /* Fast2Sum, |2| >= |f| */
    double s_hi = f / d_hi;
    double e = fma(-s_hi, d_hi, f);
    e = e - s_hi * d_lo;
    double s_lo = e / d_hi;
    double t = s_hi * s_hi;
    double p = 1.0 / 23.0;
    p = p * t + 1.0 / 21.0;
    p = p * t + 1.0 / 19.0;
    p = p * t + 1.0 / 17.0;
    p = p * t + 1.0 / 15.0;
    p = p * t + 1.0 / 13.0;
    p = p * t + 1.0 / 11.0;
    p = p * t + 1.0 / 9.0;
    p = p * t + 1.0 / 7.0;
    p = p * t + 1.0 / 5.0;
    p = p * t + 1.0 / 3.0;
    p = p * t;                          /* s^2/3 + s^4/5 + ... */
    /* log m = 2 s_hi + (2 s_lo + 2 s_hi p) */
    double lm_hi = 2.0 * s_hi;
    double lm_lo = 2.0 * s_lo + 2.0 * (s_hi * p);
    /* k ln2 in double-double */
    double dk = (double)k;
    double kh = dk * LN2_HI;
    double kl = fma(dk, LN2_HI, -kh) + dk * LN2_LO;
    /* (kh + kl) + (lm_hi + lm_lo), TwoSum on the high parts */
    double sh = kh + lm_hi;
    double bb = sh - kh;
    double sl = (kh - (sh - bb)) + (lm_hi - bb);
    sl = sl + (kl + lm_lo);
    double rh = sh + sl;
    *lo = sl - (rh - sh);
    return rh;
}

static double dd_exp(double zh, double zl)
{
    if (zh > 709.78) return INFINITY;
    if (zh < -745.2) return 0.0;
    double n = rint(zh * INV_LN2);
    double r = fma(-n, LN2_HI, zh);
    r = r - n * LN2_LO;
    r = r + zl;
    double q = 1.0 / 6227020800.0;      /* 1/13! */
    q = q * r + 1.0 / 479001600.0;
    q = q * r + 1.0 / 39916800.0;
    q = q * r + 1.0 / 3628800.0;
    q = q * r + 1.0 / 362880.0;
    q = q * r + 1.0 / 40320.0;
    q = q * r + 1.0 / 5040.0;
    q = q * r + 1.0 / 720.0;
    q = q * r + 1.0 / 120.0;
    q = q * r + 1.0 / 24.0;
    q = q * r + 1.0 / 6.0;
    q = q * r + 0.5;
    double em1 = r + (r * r) * q;       /* e^r - 1 */
    double e = 1.0 + em1;
    return ldexp(e, (int)n);
}

EXPORT double om_det_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
    if (x < 0.0) return NAN;
    if (isinf(x)) return y > 0.0 ? INFINITY : 0.0;
    double ll;
    double lh = dd_log_hi(x, &ll);
    double zh = y * lh;
    double zl = fma(y, lh, -zh) + y * ll;
    return dd_exp(zh, zl);
}

/* ------------------------------------------------------------------ */
/* Model.                                                              */

typedef struct om_model {
    int nx, ny, ns;
    int nh, nu, nv, nc;
    double dx, dy, rdx, rdy;
    wavi_phys_params P;
    wavi_solver_params S;
    uint8_t *h_mask, *h_isfixed, *u_mask, *v_mask, *c_mask, *u_inner, *v_inner;
    uint8_t *u_isfixed, *v_isfixed;
    double *sigma, *zeta, *qw;
    double *f[WAVI_NFIELDS];
    /* rheological coefficients (update_rheological_operators!) */
    double *heta_h, *heta_c, *dimp_h;
    /* operator work arrays */
    double *w_rxx, *w_ryy, *w_rxy, *w_extra;
    /* Haar: per block length m (1..B), dense m x m row-major matrices */
    int B;
    double *Wfwd[65], *Winv[65];
    int last_cg_iters;
} om_model;

static int field_len(const om_model *m, int id)
{
    if (id < 0 || id >= WAVI_NFIELDS) return -1;
    if (id <= WAVI_P_ACCUMULATION_RATE) return m->nh;
    if (id <= WAVI_U_RESIDUAL) return m->nu;
    if (id <= WAVI_V_RESIDUAL) return m->nv;
    if (id <= WAVI_S_GLEN_B) return m->nh * m->ns;
    if (id <= WAVI_WU_CORRECTION_COARSE) return m->nu;
    return m->nv;
}

#define HI(i, j) ((i) + m->nx * (j))
#define UI(i, j) ((i) + (m->nx + 1) * (j))
#define VI(i, j) ((i) + m->nx * (j))
#define CI(i, j) ((i) + (m->nx - 1) * (j))

/* wavelet_matrix(n, levels, direction), src/Wavelets/update_wavelets.jl:38-89,
 * as a dense n x n row-major matrix.  Entries are sums of products of +-1 and
 * 0.5 over disjoint supports, so they are exact. */
static void wavelet_matrix(int n, int levels, int forward, double *W)
{
    memset(W, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) W[i * n + i] = 1.0;
    int step = 1;
    double *tmp = (double *)malloc(sizeof(double) * n);
    for (int jl = levels; jl >= 0; --jl) {
        step *= 2;
        int h = step / 2;
        /* ixodd = 1:step:(n-h) (1-based) -> o = 0, step, ... with o+1 <= n-h */
        for (int o = 0; o + 1 <= n - h; o += step) {
            int e = o + h;
            if (forward) {
                /* fwav = (others'others + 0.5(evens'(evens-odds) + odds'(odds+evens))) * fwav
                 * : row o <- 0.5(row_o + row_e), row e <- 0.5(row_e - row_o) */
                for (int c = 0; c < n; ++c) {
                    double ro = W[o * n + c], re = W[e * n + c];
                    tmp[c] = 0.5 * (ro + re);
                    W[e * n + c] = 0.5 * (re - ro);
                }
                for (int c = 0; c < n; ++c) W[o * n + c] = tmp[c];
            } else {
                /* rwav = rwav * (others'others + evens'(evens+odds) + odds'(odds-evens))
                 * : col o <- col_e + col_o, col e <- col_e - col_o */
                for (int r = 0; r < n; ++r) {
                    double co = W[r * n + o], ce = W[r * n + e];
                    W[r * n + o] = ce + co;
                    W[r * n + e] = ce - co;
                }
            }
        }
    }
    free(tmp);
}

EXPORT void om_wavelet_matrix(int n, int levels, int forward, double *W)
{
    wavelet_matrix(n, levels, forward, W);
}

/* 1-D transform of `len` entries with stride `st`: out = M in (trans=0) or
 * M^T in (trans=1), M block diagonal with B-blocks.  Each output sums its
 * terms in increasing source index, as SparseArrays' products do inside
 * KroneckerProduct mul! (src/KroneckerProduct.jl:57-73). */
static void haar_1d(const om_model *m, double *const *Wtab, int trans, const double *in,
                    double *out, int len, int st)
{
    int B = m->B;
    for (int b0 = 0; b0 < len; b0 += B) {
        int bl = len - b0 < B ? len - b0 : B;
        const double *W = Wtab[bl];
        for (int r = 0; r < bl; ++r) {
            double acc = 0.0;
            for (int c = 0; c < bl; ++c) {
                double w = trans ? W[c * bl + r] : W[r * bl + c];
                if (w != 0.0) acc += w * in[(b0 + c) * st];
            }
            out[(b0 + r) * st] = acc;
        }
    }
}

/* 2-D transform (Wy (x) Wx) vec(X) on an nxw x nyw grid: the y factor first,
 * then x, which is the branch KroneckerProduct mul! takes for square
 * factors. */
static void haar_2d(const om_model *m, int fwd, int trans, const double *in, double *out,
                    int nxw, int nyw)
{
    double *const *Wtab = fwd ? m->Wfwd : m->Winv;
    double *tmp = (double *)malloc(sizeof(double) * nxw * nyw);
    for (int i = 0; i < nxw; ++i) haar_1d(m, Wtab, trans, in + i, tmp + i, nyw, nxw);
    for (int j = 0; j < nyw; ++j) haar_1d(m, Wtab, trans, tmp + j * nxw, out + j * nxw, nxw, 1);
    free(tmp);
}

EXPORT void om_haar(om_model *m, int comp, int dir, const double *in, double *out)
{
    int nxw = comp == 0 ? m->nx + 1 : m->nx;
    int nyw = comp == 0 ? m->ny : m->ny + 1;
    haar_2d(m, dir == 0, dir == 2, in, out, nxw, nyw);
}

/* ------------------------------------------------------------------ */
/* Canonical reduction: sum over the stacked [u; v] vector of a_k * b_k.
 * u then v grids are cut into 16 x 16 blocks (x fastest); each block's 256
 * products (zero outside the grid) are summed by tree256: within each run of
 * 64 elements t[e] += t[e+s] for s = 32, 16, ..., 1, then the four run sums
 * combine as (r0 + r2) + (r1 + r3).  Block partials P[0..np) are folded into
 * 256 lanes (lane l sums P[l], P[l+256], ... in order) and the lanes are
 * summed by the same tree.  The GPU kernels use this tree (a 64-wide
 * wavefront reduces one run with shuffles). */

static double tree256(double *t)
{
    for (int c = 0; c < 4; ++c)
        for (int s = 32; s > 0; s >>= 1)
            for (int e = 0; e < s; ++e) t[64 * c + e] = t[64 * c + e] + t[64 * c + e + s];
    return (t[0] + t[128]) + (t[64] + t[192]);
}

static double block_tree(const double *a, const double *b, int nxg, int nyg, int bx, int by)
{
    double t[256];
    for (int ly = 0; ly < 16; ++ly)
        for (int lx = 0; lx < 16; ++lx) {
            int i = bx * 16 + lx, j = by * 16 + ly;
            double v = 0.0;
            if (i < nxg && j < nyg) {
                int k = i + nxg * j;
                v = b ? a[k] * b[k] : a[k] * a[k];
            }
            t[lx + 16 * ly] = v;
        }
    return tree256(t);
}

static double lanes_tree(const double *P, int np)
{
    double t[256];
    for (int l = 0; l < 256; ++l) {
        double acc = 0.0;
        for (int q = l; q < np; q += 256) acc = acc + P[q];
        t[l] = acc;
    }
    return tree256(t);
}

static double canonical_dot(const om_model *m, const double *au, const double *av,
                            const double *bu, const double *bv)
{
    int nxu = m->nx + 1, nyu = m->ny, nxv = m->nx, nyv = m->ny + 1;
    int bxu = (nxu + 15) / 16, byu = (nyu + 15) / 16;
    int bxv = (nxv + 15) / 16, byv = (nyv + 15) / 16;
    int np = bxu * byu + bxv * byv;
    double *P = (double *)malloc(sizeof(double) * np);
    int q = 0;
    for (int by = 0; by < byu; ++by)
        for (int bx = 0; bx < bxu; ++bx) P[q++] = block_tree(au, bu, nxu, nyu, bx, by);
    for (int by = 0; by < byv; ++by)
        for (int bx = 0; bx < bxv; ++bx) P[q++] = block_tree(av, bv, nxv, nyv, bx, by);
    double s = lanes_tree(P, np);
    free(P);
    return s;
}

EXPORT double om_canonical_dot(om_model *m, const double *au, const double *av,
                               const double *bu, const double *bv)
{
    return canonical_dot(m, au, av, bu, bv);
}

static double canonical_norm(const om_model *m, const double *au, const double *av)
{
    return sqrt(canonical_dot(m, au, av, NULL, NULL));
}

/* ------------------------------------------------------------------ */
/* Construction: masks of setup_fields (src/Fields/Fields.jl:26-35,
 * utilities.jl:475-507) and default field values of the grid constructors. */

EXPORT om_model *om_create(const wavi_grid_desc *g, const wavi_phys_params *P,
                           const wavi_solver_params *S)
{
    om_model *m = (om_model *)calloc(1, sizeof(om_model));
    int nx = g->nx, ny = g->ny;
    m->nx = nx; m->ny = ny; m->ns = g->nsigma;
    m->nh = nx * ny; m->nu = (nx + 1) * ny; m->nv = nx * (ny + 1); m->nc = (nx - 1) * (ny - 1);
    m->dx = g->dx; m->dy = g->dy;
    m->rdx = 1.0 / g->dx; m->rdy = 1.0 / g->dy; /* spdiagm(...)/dx entries */
    m->P = *P; m->S = *S;
    m->h_mask = (uint8_t *)calloc(m->nh, 1);
    m->h_isfixed = (uint8_t *)calloc(m->nh, 1);
    m->u_mask = (uint8_t *)calloc(m->nu, 1);
    m->v_mask = (uint8_t *)calloc(m->nv, 1);
    m->c_mask = (uint8_t *)calloc(m->nc > 0 ? m->nc : 1, 1);
    m->u_inner = (uint8_t *)calloc(m->nu, 1);
    m->v_inner = (uint8_t *)calloc(m->nv, 1);
    m->u_isfixed = (uint8_t *)calloc(m->nu, 1);
    m->v_isfixed = (uint8_t *)calloc(m->nv, 1);
    for (int k = 0; k < m->nh; ++k) {
        m->h_mask[k] = g->h_mask[k] ? 1 : 0;
        m->h_isfixed[k] = g->h_isfixed && g->h_isfixed[k] ? 1 : 0;
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int um = (i < nx && m->h_mask[HI(i, j)]) || (i > 0 && m->h_mask[HI(i - 1, j)]);
            if (g->u_iszero && g->u_iszero[UI(i, j)]) um = 0;
            m->u_mask[UI(i, j)] = (uint8_t)um;
            m->u_isfixed[UI(i, j)] = g->u_isfixed && g->u_isfixed[UI(i, j)] ? 1 : 0;
            m->u_inner[UI(i, j)] = (uint8_t)(um && !m->u_isfixed[UI(i, j)]);
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int vm = (j < ny && m->h_mask[HI(i, j)]) || (j > 0 && m->h_mask[HI(i, j - 1)]);
            if (g->v_iszero && g->v_iszero[VI(i, j)]) vm = 0;
            m->v_mask[VI(i, j)] = (uint8_t)vm;
            m->v_isfixed[VI(i, j)] = g->v_isfixed && g->v_isfixed[VI(i, j)] ? 1 : 0;
            m->v_inner[VI(i, j)] = (uint8_t)(vm && !m->v_isfixed[VI(i, j)]);
        }
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i)
            m->c_mask[CI(i, j)] = m->h_mask[HI(i, j)] && m->h_mask[HI(i, j + 1)] &&
                                  m->h_mask[HI(i + 1, j)] && m->h_mask[HI(i + 1, j + 1)];
    m->sigma = (double *)malloc(sizeof(double) * m->ns);
    m->zeta = (double *)malloc(sizeof(double) * m->ns);
    m->qw = (double *)malloc(sizeof(double) * m->ns);
    for (int k = 0; k < m->ns; ++k) {
        m->sigma[k] = g->sigma[k];
        m->zeta[k] = 1.0 - g->sigma[k]; /* Grid.jl:170 */
        m->qw[k] = g->quadrature_weights[k];
    }
    for (int id = 0; id < WAVI_NFIELDS; ++id) m->f[id] = (double *)calloc(field_len(m, id), sizeof(double));
    /* defaults: dsdh = ones (HGrid.jl:213), gu/gv grounded_fraction = ones
     * (UGrid.jl:94), wavelet masks all true (UWavelets.jl:38) */
    for (int k = 0; k < m->nh; ++k) m->f[WAVI_H_DSDH][k] = 1.0;
    for (int k = 0; k < m->nu; ++k) m->f[WAVI_U_GROUNDED_FRACTION][k] = 1.0;
    for (int k = 0; k < m->nv; ++k) m->f[WAVI_V_GROUNDED_FRACTION][k] = 1.0;
    for (int k = 0; k < m->nu; ++k) m->f[WAVI_WU_MASK][k] = 1.0;
    for (int k = 0; k < m->nv; ++k) m->f[WAVI_WV_MASK][k] = 1.0;
    m->heta_h = (double *)calloc(m->nh, sizeof(double));
    m->heta_c = (double *)calloc(m->nc > 0 ? m->nc : 1, sizeof(double));
    m->dimp_h = (double *)calloc(m->nh, sizeof(double));
    m->w_rxx = (double *)calloc(m->nh, sizeof(double));
    m->w_ryy = (double *)calloc(m->nh, sizeof(double));
    m->w_rxy = (double *)calloc(m->nc > 0 ? m->nc : 1, sizeof(double));
    m->w_extra = (double *)calloc(m->nh, sizeof(double));
    m->B = 1 << (S->levels + 1);
    for (int bl = 1; bl <= m->B; ++bl) {
        m->Wfwd[bl] = (double *)malloc(sizeof(double) * bl * bl);
        m->Winv[bl] = (double *)malloc(sizeof(double) * bl * bl);
        wavelet_matrix(bl, S->levels, 1, m->Wfwd[bl]);
        wavelet_matrix(bl, S->levels, 0, m->Winv[bl]);
    }
    return m;
}

EXPORT void om_destroy(om_model *m)
{
    if (!m) return;
    free(m->h_mask); free(m->h_isfixed); free(m->u_mask); free(m->v_mask); free(m->c_mask);
    free(m->u_inner); free(m->v_inner); free(m->u_isfixed); free(m->v_isfixed);
    free(m->sigma); free(m->zeta); free(m->qw);
    for (int id = 0; id < WAVI_NFIELDS; ++id) free(m->f[id]);
    free(m->heta_h); free(m->heta_c); free(m->dimp_h);
    free(m->w_rxx); free(m->w_ryy); free(m->w_rxy); free(m->w_extra);
    for (int bl = 1; bl <= m->B; ++bl) { free(m->Wfwd[bl]); free(m->Winv[bl]); }
    free(m);
}

EXPORT double *om_field(om_model *m, int id, int *n)
{
    int len = field_len(m, id);
    if (n) *n = len;
    return len < 0 ? NULL : m->f[id];
}

EXPORT void om_set_dt(om_model *m, double dt) { m->P.dt = dt; }

EXPORT void om_masks(om_model *m, uint8_t *u_mask, uint8_t *v_mask, uint8_t *c_mask,
                     uint8_t *u_inner, uint8_t *v_inner)
{
    memcpy(u_mask, m->u_mask, m->nu);
    memcpy(v_mask, m->v_mask, m->nv);
    if (m->nc > 0) memcpy(c_mask, m->c_mask, m->nc);
    memcpy(u_inner, m->u_inner, m->nu);
    memcpy(v_inner, m->v_inner, m->nv);
}

EXPORT void om_init_quad_f2(om_model *m)
{
    /* quad_f2[mask] = h[mask]./(3*ηav[mask]) (HGrid.jl:229) */
    for (int k = 0; k < m->nh; ++k)
        if (m->h_mask[k]) m->f[WAVI_H_QUAD_F2][k] = m->f[WAVI_H_H][k] / (3.0 * m->f[WAVI_H_ETA_AV][k]);
}

/* ------------------------------------------------------------------ */
/* Averages between staggered grids with the reference's rounding.      */

/* h (nx x ny) -> u face (i in 0..nx): (c(nx)^T f)[i] = 0.5 f[i-1] + 0.5 f[i]
 * in increasing source order, terms outside the grid absent (UGrid.jl:84). */
static double avg_h_to_u(const om_model *m, const double *f, int i, int j)
{
    double a = 0.0;
    if (i >= 1) a = a + 0.5 * f[HI(i - 1, j)];
    if (i <= m->nx - 1) a = a + 0.5 * f[HI(i, j)];
    return a;
}

static double avg_h_to_v(const om_model *m, const double *f, int i, int j)
{
    double a = 0.0;
    if (j >= 1) a = a + 0.5 * f[HI(i, j - 1)];
    if (j <= m->ny - 1) a = a + 0.5 * f[HI(i, j)];
    return a;
}

/* masked h -> u average of a field times the h mask (crop), as used by
 * update_βeff_on_uv_grids! and update_geometry_on_uv_grids!. */
static double avgm_h_to_u(const om_model *m, const double *f, int i, int j)
{
    double a = 0.0;
    if (i >= 1) a = a + 0.5 * (m->h_mask[HI(i - 1, j)] ? f[HI(i - 1, j)] : 0.0);
    if (i <= m->nx - 1) a = a + 0.5 * (m->h_mask[HI(i, j)] ? f[HI(i, j)] : 0.0);
    return a;
}

static double avgm_h_to_v(const om_model *m, const double *f, int i, int j)
{
    double a = 0.0;
    if (j >= 1) a = a + 0.5 * (m->h_mask[HI(i, j - 1)] ? f[HI(i, j - 1)] : 0.0);
    if (j <= m->ny - 1) a = a + 0.5 * (m->h_mask[HI(i, j)] ? f[HI(i, j)] : 0.0);
    return a;
}

static double mask_avg_u(const om_model *m, int i, int j)
{
    double a = 0.0;
    if (i >= 1) a = a + 0.5 * (double)m->h_mask[HI(i - 1, j)];
    if (i <= m->nx - 1) a = a + 0.5 * (double)m->h_mask[HI(i, j)];
    return a;
}

static double mask_avg_v(const om_model *m, int i, int j)
{
    double a = 0.0;
    if (j >= 1) a = a + 0.5 * (double)m->h_mask[HI(i, j - 1)];
    if (j <= m->ny - 1) a = a + 0.5 * (double)m->h_mask[HI(i, j)];
    return a;
}

/* ------------------------------------------------------------------ */
/* The momentum operator op_fun! (src/utilities.jl:61-145) on full-grid
 * inputs; outputs at inner points, zero elsewhere. */

EXPORT void om_op_apply(om_model *m, const double *U, const double *V, double *fu, double *fv)
{
    const int nx = m->nx, ny = m->ny;
    const double rdx = m->rdx, rdy = m->rdy;
    const double *hu = m->f[WAVI_U_H], *hv = m->f[WAVI_V_H];
    const double *bu = m->f[WAVI_U_BETA_EFF], *bv = m->f[WAVI_V_BETA_EFF];
    double *rxx = m->w_rxx, *ryy = m->w_ryy, *rxy = m->w_rxy, *extra = m->w_extra;
    /* extensional stresses, divergence term on h (utilities.jl:83-90,118-125) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            double dudx = rdx * U[UI(i + 1, j)] - rdx * U[UI(i, j)];
            double dvdy = rdy * V[VI(i, j + 1)] - rdy * V[VI(i, j)];
            if (m->h_mask[k]) {
                rxx[k] = 2.0 * (m->heta_h[k] * (2.0 * dudx + dvdy));
                ryy[k] = 2.0 * (m->heta_h[k] * (2.0 * dvdy + dudx));
            } else {
                rxx[k] = 0.0;
                ryy[k] = 0.0;
            }
            double qx1 = m->u_mask[UI(i + 1, j)] ? hu[UI(i + 1, j)] * U[UI(i + 1, j)] : 0.0;
            double qx0 = m->u_mask[UI(i, j)] ? hu[UI(i, j)] * U[UI(i, j)] : 0.0;
            double qy1 = m->v_mask[VI(i, j + 1)] ? hv[VI(i, j + 1)] * V[VI(i, j + 1)] : 0.0;
            double qy0 = m->v_mask[VI(i, j)] ? hv[VI(i, j)] * V[VI(i, j)] : 0.0;
            double divq = (rdx * qx1 - rdx * qx0) + (rdy * qy1 - rdy * qy0);
            extra[k] = m->dimp_h[k] * divq;
        }
    /* shear stress on c (utilities.jl:93-99) */
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            int k = CI(i, j);
            if (m->c_mask[k]) {
                double dudy = rdy * U[UI(i + 1, j + 1)] - rdy * U[UI(i + 1, j)];
                double dvdx = rdx * V[VI(i + 1, j + 1)] - rdx * V[VI(i, j + 1)];
                rxy[k] = m->heta_c[k] * (dudy + dvdx);
            } else {
                rxy[k] = 0.0;
            }
        }
    /* forces on u (utilities.jl:102-104,112-113,126-134,138) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = UI(i, j);
            if (!m->u_inner[k]) { fu[k] = 0.0; continue; }
            double a1 = i < nx ? rxx[HI(i, j)] : 0.0;
            double a0 = i > 0 ? rxx[HI(i - 1, j)] : 0.0;
            double d_rxx_dx = rdx * a1 - rdx * a0;
            double c1 = (i >= 1 && i <= nx - 1 && j <= ny - 2) ? rxy[CI(i - 1, j)] : 0.0;
            double c0 = (i >= 1 && i <= nx - 1 && j >= 1) ? rxy[CI(i - 1, j - 1)] : 0.0;
            double d_rxy_dy = rdy * c1 - rdy * c0;
            double taubx = m->u_mask[k] ? bu[k] * U[k] : 0.0;
            double e1 = i < nx ? extra[HI(i, j)] : 0.0;
            double e0 = i > 0 ? extra[HI(i - 1, j)] : 0.0;
            double d_extra_dx = rdx * e1 - rdx * e0;
            fu[k] = ((d_rxx_dx + d_rxy_dy) - taubx) - hu[k] * d_extra_dx;
        }
    /* forces on v (utilities.jl:106-109,114-115,128-135,139) */
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = VI(i, j);
            if (!m->v_inner[k]) { fv[k] = 0.0; continue; }
            double a1 = j < ny ? ryy[HI(i, j)] : 0.0;
            double a0 = j > 0 ? ryy[HI(i, j - 1)] : 0.0;
            double d_ryy_dy = rdy * a1 - rdy * a0;
            double c1 = (j >= 1 && j <= ny - 1 && i <= nx - 2) ? rxy[CI(i, j - 1)] : 0.0;
            double c0 = (j >= 1 && j <= ny - 1 && i >= 1) ? rxy[CI(i - 1, j - 1)] : 0.0;
            double d_rxy_dx = rdx * c1 - rdx * c0;
            double tauby = m->v_mask[k] ? bv[k] * V[k] : 0.0;
            double e1 = j < ny ? extra[HI(i, j)] : 0.0;
            double e0 = j > 0 ? extra[HI(i, j - 1)] : 0.0;
            double d_extra_dy = rdy * e1 - rdy * e0;
            fv[k] = ((d_ryy_dy + d_rxy_dx) - tauby) - hv[k] * d_extra_dy;
        }
}

/* ------------------------------------------------------------------ */
/* inner_update! pieces (src/Models/update_velocities.jl:32-45,202-383). */

/* gc.cent (c -> h), CGrid.jl:45, in the factor order KroneckerProduct picks:
 * x first when nx > ny, else y first.  f is a c-grid field. */
static double cent_c_to_h(const om_model *m, const double *f, int i, int j)
{
    int nxc = m->nx - 1, nyc = m->ny - 1;
#define FC(a, b) (((a) >= 0 && (a) < nxc && (b) >= 0 && (b) < nyc) ? f[CI(a, b)] : 0.0)
#define HAS(a, b) ((a) >= 0 && (a) < nxc && (b) >= 0 && (b) < nyc)
    double out = 0.0;
    if (m->nx > m->ny) {
        /* temp = cx^T X (x), then Y = temp cy (y) */
        double t0 = 0.0, t1 = 0.0;
        int l0 = j - 1, l1 = j;
        if (l0 >= 0 && l0 < nyc) {
            if (HAS(i - 1, l0)) t0 = t0 + 0.5 * FC(i - 1, l0);
            if (HAS(i, l0)) t0 = t0 + 0.5 * FC(i, l0);
            out = out + 0.5 * t0;
        }
        if (l1 >= 0 && l1 < nyc) {
            if (HAS(i - 1, l1)) t1 = t1 + 0.5 * FC(i - 1, l1);
            if (HAS(i, l1)) t1 = t1 + 0.5 * FC(i, l1);
            out = out + 0.5 * t1;
        }
    } else {
        double t0 = 0.0, t1 = 0.0;
        int r0 = i - 1, r1 = i;
        if (r0 >= 0 && r0 < nxc) {
            if (HAS(r0, j - 1)) t0 = t0 + 0.5 * FC(r0, j - 1);
            if (HAS(r0, j)) t0 = t0 + 0.5 * FC(r0, j);
            out = out + 0.5 * t0;
        }
        if (r1 >= 0 && r1 < nxc) {
            if (HAS(r1, j - 1)) t1 = t1 + 0.5 * FC(r1, j - 1);
            if (HAS(r1, j)) t1 = t1 + 0.5 * FC(r1, j);
            out = out + 0.5 * t1;
        }
    }
#undef FC
#undef HAS
    return out;
}

/* gh.cent_xy (h -> c), HGrid.jl:203: y first when ny <= nx, else x first. */
static double cent_h_to_c(const om_model *m, const double *f, int i, int j)
{
    if (m->ny <= m->nx) {
        double t0 = 0.5 * f[HI(i, j)] + 0.5 * f[HI(i, j + 1)];
        double t1 = 0.5 * f[HI(i + 1, j)] + 0.5 * f[HI(i + 1, j + 1)];
        return 0.5 * t0 + 0.5 * t1;
    } else {
        double t0 = 0.5 * f[HI(i, j)] + 0.5 * f[HI(i + 1, j)];
        double t1 = 0.5 * f[HI(i, j + 1)] + 0.5 * f[HI(i + 1, j + 1)];
        return 0.5 * t0 + 0.5 * t1;
    }
}

EXPORT void om_inner_update(om_model *m)
{
    const int nx = m->nx, ny = m->ny, ns = m->ns, nh = m->nh;
    const double rdx = m->rdx, rdy = m->rdy;
    const double *U = m->f[WAVI_U_U], *V = m->f[WAVI_V_V];
    double *sr = m->f[WAVI_H_SHELF_STRAIN_RATE], *avs = m->f[WAVI_H_AV_SPEED];
    double *bs = m->f[WAVI_H_BED_SPEED], *beta = m->f[WAVI_H_BETA], *tau = m->f[WAVI_H_TAU_BED];
    double *f1 = m->f[WAVI_H_QUAD_F1], *f2 = m->f[WAVI_H_QUAD_F2];
    double *etaav = m->f[WAVI_H_ETA_AV], *beff = m->f[WAVI_H_BETA_EFF];
    double *eta = m->f[WAVI_S_ETA];
    const double *glen_b = m->f[WAVI_S_GLEN_B], *h = m->f[WAVI_H_H];
    const double *wc = m->f[WAVI_H_WEERTMAN_C];

    /* update_shelf_strain_rate! (:202-209): shear on c first */
    double *sc = (double *)calloc(m->nc > 0 ? m->nc : 1, sizeof(double));
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            double u1 = m->u_mask[UI(i + 1, j + 1)] ? U[UI(i + 1, j + 1)] : 0.0;
            double u0 = m->u_mask[UI(i + 1, j)] ? U[UI(i + 1, j)] : 0.0;
            double v1 = m->v_mask[VI(i + 1, j + 1)] ? V[VI(i + 1, j + 1)] : 0.0;
            double v0 = m->v_mask[VI(i, j + 1)] ? V[VI(i, j + 1)] : 0.0;
            double dudy = rdy * u1 - rdy * u0;
            double dvdx = rdx * v1 - rdx * v0;
            sc[CI(i, j)] = m->c_mask[CI(i, j)] ? dudy + dvdx : 0.0;
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            double ux = 0.0, vy = 0.0, sh = 0.0;
            if (m->h_mask[k]) {
                double u1 = m->u_mask[UI(i + 1, j)] ? U[UI(i + 1, j)] : 0.0;
                double u0 = m->u_mask[UI(i, j)] ? U[UI(i, j)] : 0.0;
                double v1 = m->v_mask[VI(i, j + 1)] ? V[VI(i, j + 1)] : 0.0;
                double v0 = m->v_mask[VI(i, j)] ? V[VI(i, j)] : 0.0;
                ux = rdx * u1 - rdx * u0;
                vy = rdy * v1 - rdy * v0;
                sh = cent_c_to_h(m, sc, i, j);
            }
            sr[k] = sqrt(((ux * ux + vy * vy) + ux * vy) + 0.25 * (sh * sh));
        }
    free(sc);
    /* update_av_speed! (:216-220) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            double ub = 0.0, vb = 0.0;
            if (m->h_mask[k]) {
                double u0 = m->u_mask[UI(i, j)] ? U[UI(i, j)] : 0.0;
                double u1 = m->u_mask[UI(i + 1, j)] ? U[UI(i + 1, j)] : 0.0;
                double v0 = m->v_mask[VI(i, j)] ? V[VI(i, j)] : 0.0;
                double v1 = m->v_mask[VI(i, j + 1)] ? V[VI(i, j + 1)] : 0.0;
                ub = 0.5 * u0 + 0.5 * u1;
                vb = 0.5 * v0 + 0.5 * v1;
            }
            avs[k] = sqrt(ub * ub + vb * vb);
        }
    /* update_bed_speed!, update_β!, update_basal_drag! (:227-255) */
    const double wexp = 1.0 / m->P.weertman_m - 1.0;
    const double wreg2 = m->P.weertman_reg_speed * m->P.weertman_reg_speed;
    for (int k = 0; k < nh; ++k) {
        bs[k] = avs[k] / (1.0 + f2[k] * beta[k]);
        beta[k] = wc[k] * om_det_pow(sqrt(bs[k] * bs[k] + wreg2), wexp);
        tau[k] = beta[k] * bs[k];
    }
    /* inner_update_viscosity! (:264-283) */
    const double gexp = 1.0 / m->P.glen_n - 1.0;
    const double greg2 = m->P.glen_reg_strain_rate * m->P.glen_reg_strain_rate;
    for (int ks = 0; ks < ns; ++ks)
        for (int k = 0; k < nh; ++k) {
            if (!m->h_mask[k]) continue;
            int k3 = k + nh * ks;
            for (int it = 0; it < m->S.n_iter_viscosity; ++it) {
                double q = (tau[k] * m->zeta[ks]) / eta[k3];
                double arg = sqrt((sr[k] * sr[k] + 0.25 * (q * q)) + greg2);
                eta[k3] = (0.5 * glen_b[k3]) * om_det_pow(arg, gexp);
            }
        }
    /* update_av_viscosity!, update_quadrature_falpha! (:292-328) */
    for (int k = 0; k < nh; ++k) {
        double av = 0.0, q1 = 0.0, q2 = 0.0;
        if (m->h_mask[k])
            for (int ks = 0; ks < ns; ++ks) {
                double e = eta[k + nh * ks];
                av = av + m->qw[ks] * e;
                q1 = q1 + ((m->qw[ks] * h[k]) * m->zeta[ks]) / e;
                q2 = q2 + ((m->qw[ks] * h[k]) * (m->zeta[ks] * m->zeta[ks])) / e;
            }
        etaav[k] = av;
        f1[k] = q1;
        f2[k] = q2;
    }
    /* update_βeff! (:335-339) */
    for (int k = 0; k < nh; ++k) beff[k] = beta[k] / (1.0 + f2[k] * beta[k]);
    /* update_βeff_on_uv_grids! (:348-366) */
    const double *gfh = m->f[WAVI_H_GROUNDED_FRACTION];
    double *bu = m->f[WAVI_U_BETA_EFF], *bv = m->f[WAVI_V_BETA_EFF];
    const double *gfu = m->f[WAVI_U_GROUNDED_FRACTION], *gfv = m->f[WAVI_V_GROUNDED_FRACTION];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = UI(i, j);
            if (!m->u_mask[k]) continue;
            double den = mask_avg_u(m, i, j);
            double b = avgm_h_to_u(m, beff, i, j) / den;
            double ip = avgm_h_to_u(m, gfh, i, j) / den;
            if (ip > 0.0) b = (b * gfu[k]) / ip;
            bu[k] = b;
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = VI(i, j);
            if (!m->v_mask[k]) continue;
            double den = mask_avg_v(m, i, j);
            double b = avgm_h_to_v(m, beff, i, j) / den;
            double ip = avgm_h_to_v(m, gfh, i, j) / den;
            if (ip > 0.0) b = (b * gfv[k]) / ip;
            bv[k] = b;
        }
    /* update_rheological_operators! (:374-383) */
    double *hq = (double *)malloc(sizeof(double) * nh);
    for (int k = 0; k < nh; ++k) hq[k] = h[k] * etaav[k];
    for (int k = 0; k < nh; ++k) m->heta_h[k] = m->h_mask[k] ? hq[k] : 0.0;
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i)
            m->heta_c[CI(i, j)] = m->c_mask[CI(i, j)] ? cent_h_to_c(m, hq, i, j) : 0.0;
    free(hq);
    const double c0 = (((-m->P.density_ice) * m->P.g) * m->S.super_implicitness) * m->P.dt;
    const double *dsdh = m->f[WAVI_H_DSDH];
    for (int k = 0; k < nh; ++k) m->dimp_h[k] = m->h_mask[k] ? c0 * dsdh[k] : 0.0;
}

/* ------------------------------------------------------------------ */
/* Right-hand side get_rhs (update_velocities.jl:106-182,406-421).      */

EXPORT void om_rhs(om_model *m, double *bu, double *bv)
{
    const int nx = m->nx, ny = m->ny, nh = m->nh;
    const double rdx = m->rdx, rdy = m->rdy;
    const double g = m->P.g, ri = m->P.density_ice, rw = m->P.density_ocean, sl = m->P.sea_level_wrt_geoid;
    const double *s = m->f[WAVI_H_S], *dsdh = m->f[WAVI_H_DSDH];
    const double *acc = m->f[WAVI_H_ACCUMULATION], *mel = m->f[WAVI_H_BASAL_MELT];
    double *sa = (double *)malloc(sizeof(double) * nh);
    double *mh = (double *)malloc(sizeof(double) * nh);
    const double tdt = m->S.super_implicitness * m->P.dt;
    for (int k = 0; k < nh; ++k) {
        double v = s[k] + (tdt * dsdh[k]) * (acc[k] - mel[k]);
        sa[k] = m->h_mask[k] ? v : 0.0;
        mh[k] = m->h_mask[k] ? 1.0 : 0.0;
    }
    /* f3 = -A [u .* u_isfixed ; v .* v_isfixed] (vecSampled=false) */
    double *xu = (double *)malloc(sizeof(double) * m->nu);
    double *xv = (double *)malloc(sizeof(double) * m->nv);
    double *f3u = (double *)malloc(sizeof(double) * m->nu);
    double *f3v = (double *)malloc(sizeof(double) * m->nv);
    for (int k = 0; k < m->nu; ++k) xu[k] = m->u_isfixed[k] ? m->f[WAVI_U_U][k] : 0.0;
    for (int k = 0; k < m->nv; ++k) xv[k] = m->v_isfixed[k] ? m->f[WAVI_V_V][k] : 0.0;
    om_op_apply(m, xu, xv, f3u, f3v);
    const double rg = ri * g;
    const double hri = 0.5 * ri, hrw = 0.5 * rw;
    const double *hu = m->f[WAVI_U_H], *su = m->f[WAVI_U_S];
    const double *hv = m->f[WAVI_V_H], *sv = m->f[WAVI_V_S];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = UI(i, j);
            if (!m->u_inner[k]) { bu[k] = 0.0; continue; }
            double a1 = i < nx ? sa[HI(i, j)] : 0.0, a0 = i > 0 ? sa[HI(i - 1, j)] : 0.0;
            double t = rdx * a1 - rdx * a0;
            double f1 = (rg * hu[k]) * t;
            double m1 = i < nx ? mh[HI(i, j)] : 0.0, m0 = i > 0 ? mh[HI(i - 1, j)] : 0.0;
            double tm = rdx * m1 - rdx * m0;
            double d = hu[k] - (su[k] - sl);
            if (!(d > 0.0)) d = 0.0;
            double X = (hri * (hu[k] * hu[k]) - hrw * (d * d)) - (ri * hu[k]) * su[k];
            double f2 = (tm * g) * X;
            bu[k] = (f1 + f2) + (-f3u[k]);
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = VI(i, j);
            if (!m->v_inner[k]) { bv[k] = 0.0; continue; }
            double a1 = j < ny ? sa[HI(i, j)] : 0.0, a0 = j > 0 ? sa[HI(i, j - 1)] : 0.0;
            double t = rdy * a1 - rdy * a0;
            double f1 = (rg * hv[k]) * t;
            double m1 = j < ny ? mh[HI(i, j)] : 0.0, m0 = j > 0 ? mh[HI(i, j - 1)] : 0.0;
            double tm = rdy * m1 - rdy * m0;
            double d = hv[k] - (sv[k] - sl);
            if (!(d > 0.0)) d = 0.0;
            double X = (hri * (hv[k] * hv[k]) - hrw * (d * d)) - (ri * hv[k]) * sv[k];
            double f2 = (tm * g) * X;
            bv[k] = (f1 + f2) + (-f3v[k]);
        }
    free(sa); free(mh); free(xu); free(xv); free(f3u); free(f3v);
}

/* ------------------------------------------------------------------ */
/* Preconditioner (src/Wavelets/update_preconditioners.jl).            */

/* get_op_diag (:92-109): probe with 2*stencil_margin^2 colour vectors. */
EXPORT void om_op_diag(om_model *m, double *du, double *dv)
{
    int sm = m->S.stencil_margin;
    int nx = m->nx, ny = m->ny;
    double *eu = (double *)malloc(sizeof(double) * m->nu), *ev = (double *)malloc(sizeof(double) * m->nv);
    double *ou = (double *)malloc(sizeof(double) * m->nu), *ov = (double *)malloc(sizeof(double) * m->nv);
    for (int k = 0; k < m->nu; ++k) du[k] = 0.0;
    for (int k = 0; k < m->nv; ++k) dv[k] = 0.0;
    for (int col = 1; col <= 2 * sm * sm; ++col) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int c = 1 + (i % sm) + sm * (j % sm);
                eu[UI(i, j)] = (m->u_inner[UI(i, j)] && c == col) ? 1.0 : 0.0;
            }
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = 1 + sm * sm + (i % sm) + sm * (j % sm);
                ev[VI(i, j)] = (m->v_inner[VI(i, j)] && c == col) ? 1.0 : 0.0;
            }
        om_op_apply(m, eu, ev, ou, ov);
        for (int k = 0; k < m->nu; ++k) if (eu[k] != 0.0) du[k] = ou[k];
        for (int k = 0; k < m->nv; ++k) if (ev[k] != 0.0) dv[k] = ov[k];
    }
    free(eu); free(ev); free(ou); free(ov);
}

static void resid(om_model *m, const double *xu, const double *xv, const double *bu, const double *bv,
                  double *ru, double *rv)
{
    /* get_resid! (utilities.jl:549-552): r = b - A x */
    om_op_apply(m, xu, xv, ru, rv);
    for (int k = 0; k < m->nu; ++k) ru[k] = m->u_inner[k] ? bu[k] - ru[k] : 0.0;
    for (int k = 0; k < m->nv; ++k) rv[k] = m->v_inner[k] ? bv[k] - rv[k] : 0.0;
}

/* gauss_seidel_smoother! (:133-150) with the four-colour sweep of
 * get_preconditioner (:26-32): colours 1+mod(i-j,2) (u), 3+mod(i-j,2) (v). */
static void smoother(om_model *m, double *xu, double *xv, const double *bu, const double *bv,
                     const double *du, const double *dv, const int order[4])
{
    int nx = m->nx, ny = m->ny;
    double *ru = (double *)malloc(sizeof(double) * m->nu), *rv = (double *)malloc(sizeof(double) * m->nv);
    double om = m->S.smoother_omega;
    resid(m, xu, xv, bu, bv, ru, rv);
    for (int it = 0; it < m->S.nsmooth; ++it)
        for (int q = 0; q < 4; ++q) {
            int col = order[q];
            if (col <= 2) {
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i <= nx; ++i) {
                        int k = UI(i, j);
                        if (m->u_inner[k] && 1 + ((i - j) & 1) == col) xu[k] = xu[k] + (om * ru[k]) / du[k];
                    }
            } else {
                for (int j = 0; j <= ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        int k = VI(i, j);
                        if (m->v_inner[k] && 3 + ((i - j) & 1) == col) xv[k] = xv[k] + (om * rv[k]) / dv[k];
                    }
            }
            resid(m, xu, xv, bu, bv, ru, rv);
        }
    free(ru); free(rv);
}

/* restrict R = samp_w idwt^T spread_inner and prolong P = samp_inner idwt
 * spread_w (utilities.jl:157-241) in full layout. */
static void restrict_op(om_model *m, const double *ru, const double *rv, double *cu, double *cv)
{
    double *tu = (double *)malloc(sizeof(double) * m->nu), *tv = (double *)malloc(sizeof(double) * m->nv);
    for (int k = 0; k < m->nu; ++k) tu[k] = m->u_inner[k] ? ru[k] : 0.0;
    for (int k = 0; k < m->nv; ++k) tv[k] = m->v_inner[k] ? rv[k] : 0.0;
    haar_2d(m, 0, 1, tu, cu, m->nx + 1, m->ny);
    haar_2d(m, 0, 1, tv, cv, m->nx, m->ny + 1);
    for (int k = 0; k < m->nu; ++k) if (!(m->f[WAVI_WU_MASK][k] > 0.5)) cu[k] = 0.0;
    for (int k = 0; k < m->nv; ++k) if (!(m->f[WAVI_WV_MASK][k] > 0.5)) cv[k] = 0.0;
    free(tu); free(tv);
}

static void prolong_op(om_model *m, const double *cu, const double *cv, double *xu, double *xv)
{
    double *tu = (double *)malloc(sizeof(double) * m->nu), *tv = (double *)malloc(sizeof(double) * m->nv);
    for (int k = 0; k < m->nu; ++k) tu[k] = m->f[WAVI_WU_MASK][k] > 0.5 ? cu[k] : 0.0;
    for (int k = 0; k < m->nv; ++k) tv[k] = m->f[WAVI_WV_MASK][k] > 0.5 ? cv[k] : 0.0;
    haar_2d(m, 0, 0, tu, xu, m->nx + 1, m->ny);
    haar_2d(m, 0, 0, tv, xv, m->nx, m->ny + 1);
    for (int k = 0; k < m->nu; ++k) if (!m->u_inner[k]) xu[k] = 0.0;
    for (int k = 0; k < m->nv; ++k) if (!m->v_inner[k]) xv[k] = 0.0;
    free(tu); free(tv);
}

/* A_c = R A P (get_op_coarse_fun, :176-193) */
static void coarse_op(om_model *m, const double *yu, const double *yv, double *ou, double *ov)
{
    double *pu = (double *)malloc(sizeof(double) * m->nu), *pv = (double *)malloc(sizeof(double) * m->nv);
    double *au = (double *)malloc(sizeof(double) * m->nu), *av = (double *)malloc(sizeof(double) * m->nv);
    prolong_op(m, yu, yv, pu, pv);
    om_op_apply(m, pu, pv, au, av);
    restrict_op(m, au, av, ou, ov);
    free(pu); free(pv); free(au); free(av);
}

/* IterativeSolvers.cg! on A_c (call site update_preconditioners.jl:73). */
static int coarse_cg(om_model *m, double *xu, double *xv, const double *bu, const double *bv)
{
    int nu = m->nu, nv = m->nv;
    double *uu = (double *)calloc(nu, sizeof(double)), *uv = (double *)calloc(nv, sizeof(double));
    double *ru = (double *)malloc(sizeof(double) * nu), *rv = (double *)malloc(sizeof(double) * nv);
    double *cu = (double *)malloc(sizeof(double) * nu), *cv = (double *)malloc(sizeof(double) * nv);
    double abstol = m->S.tol_coarse * canonical_norm(m, bu, bv);
    double reltol = sqrt(2.220446049250313e-16);
    /* r = b - A x0 */
    coarse_op(m, xu, xv, cu, cv);
    for (int k = 0; k < nu; ++k) ru[k] = bu[k] - cu[k];
    for (int k = 0; k < nv; ++k) rv[k] = bv[k] - cv[k];
    double residual = canonical_norm(m, ru, rv);
    double tol = reltol * residual;
    if (abstol > tol) tol = abstol;
    double prev = 1.0;
    int it = 0;
    while (!(it >= m->S.maxiter_coarse || residual <= tol)) {
        double beta = (residual * residual) / (prev * prev);
        for (int k = 0; k < nu; ++k) uu[k] = ru[k] + beta * uu[k];
        for (int k = 0; k < nv; ++k) uv[k] = rv[k] + beta * uv[k];
        coarse_op(m, uu, uv, cu, cv);
        double alpha = (residual * residual) / canonical_dot(m, uu, uv, cu, cv);
        for (int k = 0; k < nu; ++k) xu[k] = xu[k] + alpha * uu[k];
        for (int k = 0; k < nv; ++k) xv[k] = xv[k] + alpha * uv[k];
        for (int k = 0; k < nu; ++k) ru[k] = ru[k] - alpha * cu[k];
        for (int k = 0; k < nv; ++k) rv[k] = rv[k] - alpha * cv[k];
        prev = residual;
        residual = canonical_norm(m, ru, rv);
        ++it;
    }
    free(uu); free(uv); free(ru); free(rv); free(cu); free(cv);
    return it;
}

/* Test hooks: the coarse operator A_c, and the coarse CG alone with a given
 * tolerance and iteration cap (x: the warm start on entry, the solution on
 * return; returns the iterations). */
EXPORT void om_coarse_op(om_model *m, const double *yu, const double *yv, double *ou, double *ov)
{
    coarse_op(m, yu, yv, ou, ov);
}

EXPORT int om_coarse_cg(om_model *m, const double *bu, const double *bv, double *xu, double *xv, double tol_coarse,
                        int maxiter)
{
    const double t = m->S.tol_coarse;
    const int mi = m->S.maxiter_coarse;
    m->S.tol_coarse = tol_coarse;
    m->S.maxiter_coarse = maxiter;
    const int it = coarse_cg(m, xu, xv, bu, bv);
    m->S.tol_coarse = t;
    m->S.maxiter_coarse = mi;
    return it;
}

/* precondition!(x, p, b) (:54-83) with x = 0 on entry. */
EXPORT int om_vcycle(om_model *m, const double *bu, const double *bv, double *xu, double *xv)
{
    int nu = m->nu, nv = m->nv;
    double *du = (double *)malloc(sizeof(double) * nu), *dv = (double *)malloc(sizeof(double) * nv);
    om_op_diag(m, du, dv);
    for (int k = 0; k < nu; ++k) xu[k] = 0.0;
    for (int k = 0; k < nv; ++k) xv[k] = 0.0;
    static const int fwd[4] = {1, 3, 2, 4}, rev[4] = {4, 2, 3, 1};
    smoother(m, xu, xv, bu, bv, du, dv, fwd);
    double *ru = (double *)malloc(sizeof(double) * nu), *rv = (double *)malloc(sizeof(double) * nv);
    double *cu = (double *)malloc(sizeof(double) * nu), *cv = (double *)malloc(sizeof(double) * nv);
    resid(m, xu, xv, bu, bv, ru, rv);
    restrict_op(m, ru, rv, cu, cv);
    double *wu = m->f[WAVI_WU_CORRECTION_COARSE], *wv = m->f[WAVI_WV_CORRECTION_COARSE];
    int iters = coarse_cg(m, wu, wv, cu, cv);
    prolong_op(m, wu, wv, ru, rv);
    for (int k = 0; k < nu; ++k) xu[k] = xu[k] + ru[k];
    for (int k = 0; k < nv; ++k) xv[k] = xv[k] + rv[k];
    smoother(m, xu, xv, bu, bv, du, dv, rev);
    free(du); free(dv); free(ru); free(rv); free(cu); free(cv);
    m->last_cg_iters = iters;
    return iters;
}

/* precondition!(model, ::BasicParallelSpec) (update_velocities.jl:50-83) */
static int precondition_model(om_model *m, double *rel_out, int *cg_iters)
{
    int nu = m->nu, nv = m->nv;
    double *xu = (double *)malloc(sizeof(double) * nu), *xv = (double *)malloc(sizeof(double) * nv);
    double *bu = (double *)malloc(sizeof(double) * nu), *bv = (double *)malloc(sizeof(double) * nv);
    double *ru = (double *)malloc(sizeof(double) * nu), *rv = (double *)malloc(sizeof(double) * nv);
    double *eu = (double *)calloc(nu, sizeof(double)), *ev = (double *)calloc(nv, sizeof(double));
    for (int k = 0; k < nu; ++k) xu[k] = m->u_inner[k] ? m->f[WAVI_U_U][k] : 0.0;
    for (int k = 0; k < nv; ++k) xv[k] = m->v_inner[k] ? m->f[WAVI_V_V][k] : 0.0;
    om_rhs(m, bu, bv);
    resid(m, xu, xv, bu, bv, ru, rv);
    for (int k = 0; k < nu; ++k) if (m->u_inner[k]) m->f[WAVI_U_RESIDUAL][k] = ru[k];
    for (int k = 0; k < nv; ++k) if (m->v_inner[k]) m->f[WAVI_V_RESIDUAL][k] = rv[k];
    double rel = canonical_norm(m, ru, rv) / canonical_norm(m, bu, bv);
    int converged = rel < m->S.tol_picard;
    *cg_iters = 0;
    if (!converged) *cg_iters = om_vcycle(m, ru, rv, eu, ev);
    for (int k = 0; k < nu; ++k) if (m->u_inner[k]) m->f[WAVI_U_U][k] = xu[k] + eu[k];
    for (int k = 0; k < nv; ++k) if (m->v_inner[k]) m->f[WAVI_V_V][k] = xv[k] + ev[k];
    free(xu); free(xv); free(bu); free(bv); free(ru); free(rv); free(eu); free(ev);
    *rel_out = rel;
    return converged;
}

EXPORT int om_precondition(om_model *m, double *rel, int *cg_iters)
{
    return precondition_model(m, rel, cg_iters);
}

/* update_velocities! (update_velocities.jl:11-29) */
EXPORT int om_update_velocities(om_model *m, wavi_solve_report *rep)
{
    int converged = 0, it = 0, total = 0;
    double rel = INFINITY;
    if (rep) memset(rep, 0, sizeof(*rep));
    while (!converged && it < m->S.maxiter_picard) {
        it = it + 1;
        om_inner_update(m);
        int cg = 0;
        converged = precondition_model(m, &rel, &cg);
        total += cg;
        if (rep && it <= WAVI_REPORT_MAX_PICARD) {
            rep->coarse_cg_iters[it - 1] = cg;
            rep->rel_resid_hist[it - 1] = rel;
        }
    }
    if (rep) {
        rep->picard_iters = it;
        rep->converged = converged;
        rep->rel_resid = rel;
        rep->coarse_cg_total = total;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* update_state! pieces (src/Models/update_state.jl) and update_wavelets!. */

/* pos_fraction(z1; mask) (src/utilities.jl:255-396). Outputs m x n, (m+1) x n,
 * m x (n+1) in column-major order. */
EXPORT void om_pos_fraction(int M, int N, const double *z1, const uint8_t *mask,
                            double *af, double *afu, double *afv)
{
#define Z(i, j) z1[(i) + M * (j)]
#define MK(i, j) mask[(i) + M * (j)]
    double *q = (double *)malloc(sizeof(double) * M * N);
    for (int k = 0; k < M * N; ++k) af[k] = 0.0;
    for (int k = 0; k < (M + 1) * N; ++k) afu[k] = 0.0;
    for (int k = 0; k < M * (N + 1); ++k) afv[k] = 0.0;
    for (int quad = 1; quad <= 4; ++quad) {
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int i2, j2, i3, j3;
                if (quad == 1) { i2 = i + 1 < M ? i + 1 : M - 1; j2 = j; i3 = i; j3 = j + 1 < N ? j + 1 : N - 1; }
                else if (quad == 2) { i2 = i; j2 = j + 1 < N ? j + 1 : N - 1; i3 = i > 0 ? i - 1 : 0; j3 = j; }
                else if (quad == 3) { i2 = i > 0 ? i - 1 : 0; j2 = j; i3 = i; j3 = j > 0 ? j - 1 : 0; }
                else { i2 = i; j2 = j > 0 ? j - 1 : 0; i3 = i + 1 < M ? i + 1 : M - 1; j3 = j; }
                double v1 = Z(i, j);
                double v2 = MK(i2, j2) ? Z(i2, j2) : v1;
                double v3 = MK(i3, j3) ? Z(i3, j3) : v1;
                double s1 = (v1 > 0) - (v1 < 0), s2 = (v2 > 0) - (v2 < 0), s3 = (v3 > 0) - (v3 < 0);
                double area = 0.0;
                if (!((s1 == s2) && (s1 == s3))) {
                    int flip = fabs(v3 - v1) < fabs(v2 - v1);
                    double a, b;
                    if (!flip) {
                        a = -(v2 - v1) / (v3 - v1);
                        b = -2.0 * v1 / (v3 - v1);
                    } else {
                        a = -(v3 - v1) / (v2 - v1);
                        b = -2.0 * v1 / (v2 - v1);
                    }
                    double a1 = 0.5 * (b * b) / a;
                    double a2 = 0.5 * ((1.0 - b) * (1.0 - b)) / a;
                    double a3 = 0.5 * a + b;
                    int t1 = (b > 0.0) + (b > 1.0);
                    double ab = a + b;
                    int t2 = (ab > 0.0) + (ab > 1.0);
                    int ix = 1 + t1 + 3 * t2;
                    switch (ix) {
                    case 1: area = 0.0; break;
                    case 2: area = -a1; break;
                    case 3: area = a2 - a1; break;
                    case 4: area = a1 + a3; break;
                    case 5: area = a3; break;
                    case 6: area = a2 + a3; break;
                    case 7: area = (1.0 - a2) + a1; break;
                    case 8: area = 1.0 - a2; break;
                    default: area = 1.0; break;
                    }
                    if (b < 0.0) area = 1.0 - area;
                    if (v1 < 0.0) area = 1.0 - area;
                }
                if (v1 > 0.0 && v2 > 0.0 && v3 > 0.0) area = 1.0;
                if (v1 <= 0.0 && v2 <= 0.0 && v3 <= 0.0) area = 0.0;
                if (!MK(i, j)) area = 0.0;
                q[i + M * j] = area;
            }
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) af[i + M * j] = af[i + M * j] + 0.25 * q[i + M * j];
        for (int j = 0; j < N; ++j)
            for (int i = 0; i <= M; ++i) {
                double add;
                if (quad == 1 || quad == 4) add = i >= 1 ? q[(i - 1) + M * j] : 0.0;
                else add = i < M ? q[i + M * j] : 0.0;
                afu[i + (M + 1) * j] = afu[i + (M + 1) * j] + 0.25 * add;
            }
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i < M; ++i) {
                double add;
                if (quad == 1 || quad == 2) add = j >= 1 ? q[i + M * (j - 1)] : 0.0;
                else add = j < N ? q[i + M * j] : 0.0;
                afv[i + M * j] = afv[i + M * j] + 0.25 * add;
            }
    }
    free(q);
#undef Z
#undef MK
}

/* update_wavelets! (src/Wavelets/update_wavelets.jl:6-32) */
EXPORT void om_update_wavelets(om_model *m)
{
    double *tu = (double *)malloc(sizeof(double) * m->nu), *tv = (double *)malloc(sizeof(double) * m->nv);
    for (int k = 0; k < m->nu; ++k) tu[k] = m->u_mask[k] ? m->f[WAVI_U_U][k] : 0.0;
    for (int k = 0; k < m->nv; ++k) tv[k] = m->v_mask[k] ? m->f[WAVI_V_V][k] : 0.0;
    haar_2d(m, 1, 0, tu, m->f[WAVI_WU_WAVELETS], m->nx + 1, m->ny);
    haar_2d(m, 1, 0, tv, m->f[WAVI_WV_WAVELETS], m->nx, m->ny + 1);
    double thr = m->S.wavelet_threshold;
    for (int k = 0; k < m->nu; ++k) {
        m->f[WAVI_WU_MASK][k] = fabs(m->f[WAVI_WU_WAVELETS][k]) >= thr ? 1.0 : 0.0;
        m->f[WAVI_WU_CORRECTION_COARSE][k] = 0.0;
    }
    for (int k = 0; k < m->nv; ++k) {
        m->f[WAVI_WV_MASK][k] = fabs(m->f[WAVI_WV_WAVELETS][k]) >= thr ? 1.0 : 0.0;
        m->f[WAVI_WV_CORRECTION_COARSE][k] = 0.0;
    }
    free(tu); free(tv);
}

/* update_state! steps before the velocity solve (update_state.jl:8-15). */
EXPORT void om_update_geometry(om_model *m)
{
    const int nx = m->nx, ny = m->ny, nh = m->nh;
    const double ri = m->P.density_ice, rw = m->P.density_ocean, sl = m->P.sea_level_wrt_geoid;
    double *h = m->f[WAVI_H_H], *b = m->f[WAVI_H_B], *s = m->f[WAVI_H_S];
    /* update_surface_elevation! (:51-57) */
    for (int k = 0; k < nh; ++k)
        if (m->h_mask[k]) {
            double s1 = b[k] + h[k], s2 = sl + h[k] * (1.0 - ri / rw);
            s[k] = s1 > s2 ? s1 : s2;
            if (isnan(s1) || isnan(s2)) s[k] = NAN;
        }
    /* update_geometry_on_uv_grids! (:65-73) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = UI(i, j);
            if (!m->u_mask[k]) continue;
            double den = mask_avg_u(m, i, j);
            m->f[WAVI_U_H][k] = avgm_h_to_u(m, h, i, j) / den;
            m->f[WAVI_U_S][k] = avgm_h_to_u(m, s, i, j) / den;
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = VI(i, j);
            if (!m->v_mask[k]) continue;
            double den = mask_avg_v(m, i, j);
            m->f[WAVI_V_H][k] = avgm_h_to_v(m, h, i, j) / den;
            m->f[WAVI_V_S][k] = avgm_h_to_v(m, s, i, j) / den;
        }
    /* update_height_above_floatation! (:80-85) */
    double *haf = m->f[WAVI_H_HAF];
    for (int k = 0; k < nh; ++k) haf[k] = h[k] - (rw / ri) * (sl - b[k]);
    /* update_grounded_fraction_on_huv_grids! (:92-99) */
    om_pos_fraction(nx, ny, haf, m->h_mask, m->f[WAVI_H_GROUNDED_FRACTION],
                    m->f[WAVI_U_GROUNDED_FRACTION], m->f[WAVI_V_GROUNDED_FRACTION]);
    /* accumulation, basal melt, weertman_c, dsdh (:106-147) */
    for (int k = 0; k < nh; ++k) {
        m->f[WAVI_H_ACCUMULATION][k] = m->f[WAVI_P_ACCUMULATION_RATE][k];
        m->f[WAVI_H_BASAL_MELT][k] = m->P.uniform_melt_rate;
        m->f[WAVI_H_WEERTMAN_C][k] = m->f[WAVI_P_WEERTMAN_C][k] * m->f[WAVI_H_GROUNDED_FRACTION][k];
        m->f[WAVI_H_DSDH][k] = (1.0 - ri / rw) + (ri / rw) * m->f[WAVI_H_GROUNDED_FRACTION][k];
    }
}

/* update_state!(model, clock) (update_state.jl:7-21) with UniformMeltRate. */
EXPORT void om_update_state(om_model *m, wavi_solve_report *rep)
{
    const int nx = m->nx, ny = m->ny;
    om_update_geometry(m);
    om_update_velocities(m, rep);
    /* update_velocities_on_h_grid! (:164-178) */
    const double *U = m->f[WAVI_U_U], *V = m->f[WAVI_V_V];
    const double *beta = m->f[WAVI_H_BETA], *f1 = m->f[WAVI_H_QUAD_F1], *f2 = m->f[WAVI_H_QUAD_F2];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            double uh = 0.5 * U[UI(i, j)] + 0.5 * U[UI(i + 1, j)];
            double vh = 0.5 * V[VI(i, j)] + 0.5 * V[VI(i, j + 1)];
            m->f[WAVI_H_U][k] = uh;
            m->f[WAVI_H_V][k] = vh;
            double ub = uh / (1.0 + beta[k] * f2[k]);
            double vb = vh / (1.0 + beta[k] * f2[k]);
            m->f[WAVI_H_UB][k] = ub;
            m->f[WAVI_H_VB][k] = vb;
            m->f[WAVI_H_US][k] = ub * (1.0 + beta[k] * f1[k]);
            m->f[WAVI_H_VS][k] = vb * (1.0 + beta[k] * f1[k]);
        }
    /* update_dhdt! (:184-189) */
    const double *hu = m->f[WAVI_U_H], *hv = m->f[WAVI_V_H];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            if (!m->h_mask[k]) continue;
            double q1 = m->u_mask[UI(i + 1, j)] ? hu[UI(i + 1, j)] * U[UI(i + 1, j)] : 0.0;
            double q0 = m->u_mask[UI(i, j)] ? hu[UI(i, j)] * U[UI(i, j)] : 0.0;
            double p1 = m->v_mask[VI(i, j + 1)] ? hv[VI(i, j + 1)] * V[VI(i, j + 1)] : 0.0;
            double p0 = m->v_mask[VI(i, j)] ? hv[VI(i, j)] * V[VI(i, j)] : 0.0;
            double div = (m->rdx * q1 - m->rdx * q0) + (m->rdy * p1 - m->rdy * p0);
            m->f[WAVI_H_DHDT][k] = (m->f[WAVI_H_ACCUMULATION][k] - m->f[WAVI_H_BASAL_MELT][k]) - div;
        }
    om_update_wavelets(m);
}

/* update_thickness! (src/Simulations/run_simulation.jl:24-42), evolveShelves. */
EXPORT void om_update_thickness(om_model *m, double dt)
{
    double *h = m->f[WAVI_H_H];
    const double *dhdt = m->f[WAVI_H_DHDT];
    for (int k = 0; k < m->nh; ++k) {
        if (!m->h_mask[k]) continue;
        double a = m->P.minimum_thickness - h[k], b = dt * dhdt[k];
        double upd = a > b ? a : b;
        if (m->h_isfixed[k]) upd = 0.0;
        h[k] = h[k] + upd;
    }
}

/* nsteps of timestep! (run_simulation.jl:6-17) */
EXPORT void om_run_steps(om_model *m, int nsteps, wavi_solve_report *rep)
{
    for (int n = 0; n < nsteps; ++n) {
        om_update_state(m, rep);
        om_update_thickness(m, m->P.dt);
    }
}

/* ------------------------------------------------------------------ */
/* Adjoint basal-drag gradient (no reference counterpart: WAVI.jl has no
 * adjoint).  The map and its gradient are specified beside
 * wavi_misfit_gradient in include/wavi_hip.h.  The forward steps are
 * inner_update! (update_velocities.jl:32-45, restated above) and linear
 * solves by the V-cycle of update_preconditioners.jl:54-83 iterated on the
 * residual; the reverse sweep is the chain rule of those steps, written
 * out by hand.  Sums here follow no fixed order: the GPU gradient is
 * compared with this one to a tolerance, and both with central finite
 * differences of the map.                                             */

/* x = A^{-1} b by V-cycles on the residual, from x = 0.  Each V-cycle's
 * coarse CG starts from zero (the Picard loop's warm start would be the
 * previous, much larger correction). */
static int solve_linear(om_model *m, const double *bu, const double *bv, double *xu, double *xv,
                        double tol, int maxit, int *vcycles, double *rel_out)
{
    int nu = m->nu, nv = m->nv;
    double *ru = (double *)malloc(sizeof(double) * nu), *rv = (double *)malloc(sizeof(double) * nv);
    double *eu = (double *)malloc(sizeof(double) * nu), *ev = (double *)malloc(sizeof(double) * nv);
    for (int k = 0; k < nu; ++k) xu[k] = 0.0;
    for (int k = 0; k < nv; ++k) xv[k] = 0.0;
    double bn = canonical_norm(m, bu, bv), rel = 0.0;
    int it = 0;
    for (;;) {
        resid(m, xu, xv, bu, bv, ru, rv);
        rel = bn > 0.0 ? canonical_norm(m, ru, rv) / bn : 0.0;
        if (rel <= tol || it >= maxit) break;
        memset(m->f[WAVI_WU_CORRECTION_COARSE], 0, sizeof(double) * nu);
        memset(m->f[WAVI_WV_CORRECTION_COARSE], 0, sizeof(double) * nv);
        om_vcycle(m, ru, rv, eu, ev);
        for (int k = 0; k < nu; ++k) if (m->u_inner[k]) xu[k] = xu[k] + eu[k];
        for (int k = 0; k < nv; ++k) if (m->v_inner[k]) xv[k] = xv[k] + ev[k];
        ++it;
    }
    free(ru); free(rv); free(eu); free(ev);
    *vcycles += it;
    *rel_out = rel;
    return rel <= tol;
}

typedef struct {
    double *u, *v, *eta, *beta, *f2;
} om_ckpt;

static void ckpt_alloc(const om_model *m, om_ckpt *c)
{
    c->u = (double *)malloc(sizeof(double) * m->nu);
    c->v = (double *)malloc(sizeof(double) * m->nv);
    c->eta = (double *)malloc(sizeof(double) * m->nh * m->ns);
    c->beta = (double *)malloc(sizeof(double) * m->nh);
    c->f2 = (double *)malloc(sizeof(double) * m->nh);
}

static void ckpt_free(om_ckpt *c)
{
    free(c->u); free(c->v); free(c->eta); free(c->beta); free(c->f2);
}

static void ckpt_save(const om_model *m, om_ckpt *c)
{
    memcpy(c->u, m->f[WAVI_U_U], sizeof(double) * m->nu);
    memcpy(c->v, m->f[WAVI_V_V], sizeof(double) * m->nv);
    memcpy(c->eta, m->f[WAVI_S_ETA], sizeof(double) * m->nh * m->ns);
    memcpy(c->beta, m->f[WAVI_H_BETA], sizeof(double) * m->nh);
    memcpy(c->f2, m->f[WAVI_H_QUAD_F2], sizeof(double) * m->nh);
}

static void ckpt_load(om_model *m, const om_ckpt *c)
{
    memcpy(m->f[WAVI_U_U], c->u, sizeof(double) * m->nu);
    memcpy(m->f[WAVI_V_V], c->v, sizeof(double) * m->nv);
    memcpy(m->f[WAVI_S_ETA], c->eta, sizeof(double) * m->nh * m->ns);
    memcpy(m->f[WAVI_H_BETA], c->beta, sizeof(double) * m->nh);
    memcpy(m->f[WAVI_H_QUAD_F2], c->f2, sizeof(double) * m->nh);
}

/* Reverse of Picard iteration k.  Fields on entry: inner_update! of
 * iteration k recomputed from checkpoint k-1 (ck_old: u_{k-1} and the
 * eta, beta, f2 it started from).  uk, vk: u_k; lu, lv: lambda_k =
 * A_k^{-1} ubar_k.  etab, betab, f2b: in, the adjoint of the eta, beta, f2
 * leaving iteration k; out, of those entering it.  ubp, vbp: out, the
 * adjoint of u_{k-1} at inner points.  wcb: += dJ/d(weertman_c on h). */
static void picard_reverse(om_model *m, const om_ckpt *ck_old, const double *uk, const double *vk,
                           const double *lu, const double *lv, double *etab, double *betab, double *f2b,
                           double *ubp, double *vbp, double *wcb)
{
    const int nx = m->nx, ny = m->ny, ns = m->ns, nh = m->nh, nc = m->nc;
    const double rdx = m->rdx, rdy = m->rdy;
    const double *up = ck_old->u, *vp = ck_old->v;
    const double *h = m->f[WAVI_H_H];
    const double *sr = m->f[WAVI_H_SHELF_STRAIN_RATE], *avs = m->f[WAVI_H_AV_SPEED];
    const double *bs = m->f[WAVI_H_BED_SPEED], *beta = m->f[WAVI_H_BETA], *tau = m->f[WAVI_H_TAU_BED];
    const double *f2 = m->f[WAVI_H_QUAD_F2], *eta = m->f[WAVI_S_ETA], *glen_b = m->f[WAVI_S_GLEN_B];
    const double *gfh = m->f[WAVI_H_GROUNDED_FRACTION];
    const double *gfu = m->f[WAVI_U_GROUNDED_FRACTION], *gfv = m->f[WAVI_V_GROUNDED_FRACTION];
    double *hhb = (double *)calloc(nh, sizeof(double));
    double *hcb = (double *)calloc(nc > 0 ? nc : 1, sizeof(double));
    double *scb = (double *)calloc(nc > 0 ? nc : 1, sizeof(double));
    double *sc = (double *)calloc(nc > 0 ? nc : 1, sizeof(double));
    double *bub = (double *)calloc(m->nu, sizeof(double)), *bvb = (double *)calloc(m->nv, sizeof(double));
    double *cub = (double *)calloc(nh, sizeof(double)), *cvb = (double *)calloc(nh, sizeof(double));
    double *uxb = (double *)calloc(nh, sizeof(double)), *vyb = (double *)calloc(nh, sizeof(double));
    double *shb = (double *)calloc(nh, sizeof(double));
    /* 1. coefficient adjoints cbar = -d(lambda^T A(c) u_k)/dc of op_fun!:
     *    h eta on h (extensional stresses), h eta on c (shear), beta_eff on
     *    u and v (drag) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            if (!m->h_mask[k]) continue;
            double dudx = rdx * uk[UI(i + 1, j)] - rdx * uk[UI(i, j)];
            double dvdy = rdy * vk[VI(i, j + 1)] - rdy * vk[VI(i, j)];
            double dlx = rdx * lu[UI(i + 1, j)] - rdx * lu[UI(i, j)];
            double dly = rdy * lv[VI(i, j + 1)] - rdy * lv[VI(i, j)];
            hhb[k] = 2.0 * (2.0 * dudx + dvdy) * dlx + 2.0 * (2.0 * dvdy + dudx) * dly;
        }
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            int k = CI(i, j);
            if (!m->c_mask[k]) continue;
            double s = (rdy * uk[UI(i + 1, j + 1)] - rdy * uk[UI(i + 1, j)]) +
                       (rdx * vk[VI(i + 1, j + 1)] - rdx * vk[VI(i, j + 1)]);
            double l = (rdy * lu[UI(i + 1, j + 1)] - rdy * lu[UI(i + 1, j)]) +
                       (rdx * lv[VI(i + 1, j + 1)] - rdx * lv[VI(i, j + 1)]);
            hcb[k] = s * l;
        }
    for (int k = 0; k < m->nu; ++k) bub[k] = m->u_mask[k] ? lu[k] * uk[k] : 0.0;
    for (int k = 0; k < m->nv; ++k) bvb[k] = m->v_mask[k] ? lv[k] * vk[k] : 0.0;
    /* shear strain of u_{k-1} on c (update_shelf_strain_rate!) */
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            double u1 = m->u_mask[UI(i + 1, j + 1)] ? up[UI(i + 1, j + 1)] : 0.0;
            double u0 = m->u_mask[UI(i + 1, j)] ? up[UI(i + 1, j)] : 0.0;
            double v1 = m->v_mask[VI(i + 1, j + 1)] ? vp[VI(i + 1, j + 1)] : 0.0;
            double v0 = m->v_mask[VI(i, j + 1)] ? vp[VI(i, j + 1)] : 0.0;
            sc[CI(i, j)] = m->c_mask[CI(i, j)] ? (rdy * u1 - rdy * u0) + (rdx * v1 - rdx * v0) : 0.0;
        }
    /* 2. per h cell, in reverse order of inner_update! */
    const double gexp = 1.0 / m->P.glen_n - 1.0, wexp = 1.0 / m->P.weertman_m - 1.0;
    const double greg2 = m->P.glen_reg_strain_rate * m->P.glen_reg_strain_rate;
    const double wreg2 = m->P.weertman_reg_speed * m->P.weertman_reg_speed;
    const int npass = m->S.n_iter_viscosity;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = HI(i, j);
            /* update_rheological_operators!: h eta_av on h and (cent) on c */
            double hqb = m->h_mask[k] ? hhb[k] : 0.0;
            for (int b = j - 1; b <= j; ++b)
                for (int a = i - 1; a <= i; ++a)
                    if (a >= 0 && a < nx - 1 && b >= 0 && b < ny - 1 && m->c_mask[CI(a, b)])
                        hqb += 0.25 * hcb[CI(a, b)];
            double etaavb = h[k] * hqb;
            /* update_betaeff_on_uv_grids! */
            double beffb = 0.0;
            if (m->h_mask[k]) {
                for (int a = i; a <= i + 1; ++a) {
                    int ku = UI(a, j);
                    if (!m->u_mask[ku]) continue;
                    double den = mask_avg_u(m, a, j), ip = avgm_h_to_u(m, gfh, a, j) / den;
                    beffb += bub[ku] * (0.5 / den) * (ip > 0.0 ? gfu[ku] / ip : 1.0);
                }
                for (int b = j; b <= j + 1; ++b) {
                    int kv = VI(i, b);
                    if (!m->v_mask[kv]) continue;
                    double den = mask_avg_v(m, i, b), ip = avgm_h_to_v(m, gfh, i, b) / den;
                    beffb += bvb[kv] * (0.5 / den) * (ip > 0.0 ? gfv[kv] / ip : 1.0);
                }
            }
            /* update_betaeff!: beta_eff = beta / (1 + f2 beta) */
            double D = 1.0 + f2[k] * beta[k];
            double bb = betab[k] + beffb / (D * D);
            double fb = f2b[k] - beffb * (beta[k] * beta[k]) / (D * D);
            double taub = 0.0, srb = 0.0;
            if (m->h_mask[k]) {
                for (int s = 0; s < ns; ++s) {
                    int k3 = k + nh * s;
                    double z = m->zeta[s];
                    /* update_av_viscosity!, update_quadrature_falpha! */
                    double eb = etab[k3] + m->qw[s] * etaavb -
                                fb * ((m->qw[s] * h[k]) * (z * z)) / (eta[k3] * eta[k3]);
                    /* inner_update_viscosity!: npass fixed-point passes */
                    double e[9];
                    e[0] = ck_old->eta[k3];
                    for (int p = 0; p < npass; ++p) {
                        double q = (tau[k] * z) / e[p];
                        double arg = sqrt((sr[k] * sr[k] + 0.25 * (q * q)) + greg2);
                        e[p + 1] = (0.5 * glen_b[k3]) * om_det_pow(arg, gexp);
                    }
                    for (int p = npass - 1; p >= 0; --p) {
                        double q = (tau[k] * z) / e[p];
                        double arg = sqrt((sr[k] * sr[k] + 0.25 * (q * q)) + greg2);
                        double g = eb * e[p + 1] * gexp / arg;
                        srb += g * sr[k] / arg;
                        double qb = g * 0.25 * q / arg;
                        taub += qb * z / e[p];
                        eb = -qb * q / e[p];
                    }
                    etab[k3] = eb;
                }
            }
            /* update_basal_drag!: tau = beta bs; update_beta!: beta =
             * wc (bs^2 + reg^2)^(wexp/2) */
            double bsb = taub * beta[k];
            bb += taub * bs[k];
            double S = bs[k] * bs[k] + wreg2;
            wcb[k] += bb * om_det_pow(sqrt(S), wexp);
            bsb += bb * beta[k] * wexp * bs[k] / S;
            /* update_bed_speed!: bs = avs / (1 + f2_old beta_old) */
            double D0 = 1.0 + ck_old->f2[k] * ck_old->beta[k];
            double avsb = bsb / D0;
            f2b[k] = -bsb * avs[k] * ck_old->beta[k] / (D0 * D0);
            betab[k] = -bsb * avs[k] * ck_old->f2[k] / (D0 * D0);
            if (!m->h_mask[k]) continue;
            /* update_av_speed! of u_{k-1} */
            double u0 = m->u_mask[UI(i, j)] ? up[UI(i, j)] : 0.0, u1 = m->u_mask[UI(i + 1, j)] ? up[UI(i + 1, j)] : 0.0;
            double v0 = m->v_mask[VI(i, j)] ? vp[VI(i, j)] : 0.0, v1 = m->v_mask[VI(i, j + 1)] ? vp[VI(i, j + 1)] : 0.0;
            double cu = 0.5 * u0 + 0.5 * u1, cv = 0.5 * v0 + 0.5 * v1;
            if (avs[k] > 0.0) {
                cub[k] = avsb * cu / avs[k];
                cvb[k] = avsb * cv / avs[k];
            }
            /* update_shelf_strain_rate! of u_{k-1} */
            if (sr[k] > 0.0) {
                double ux = rdx * u1 - rdx * u0, vy = rdy * v1 - rdy * v0, sh = cent_c_to_h(m, sc, i, j);
                uxb[k] = srb * (2.0 * ux + vy) / (2.0 * sr[k]);
                vyb[k] = srb * (2.0 * vy + ux) / (2.0 * sr[k]);
                shb[k] = srb * 0.25 * sh / sr[k];
            }
        }
    /* 3. shear on c: sh = cent(sc) on h */
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            int k = CI(i, j);
            if (!m->c_mask[k]) continue;
            scb[k] = 0.25 * (shb[HI(i, j)] + shb[HI(i + 1, j)] + shb[HI(i, j + 1)] + shb[HI(i + 1, j + 1)]);
        }
    /* 4. adjoint of u_{k-1} at inner points */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = UI(i, j);
            double a = 0.0;
            if (m->u_inner[k]) {
                if (i >= 1) a += 0.5 * cub[HI(i - 1, j)] + rdx * uxb[HI(i - 1, j)];
                if (i <= nx - 1) a += 0.5 * cub[HI(i, j)] - rdx * uxb[HI(i, j)];
                if (i >= 1 && i <= nx - 1) {
                    if (j >= 1) a += rdy * scb[CI(i - 1, j - 1)];
                    if (j <= ny - 2) a -= rdy * scb[CI(i - 1, j)];
                }
            }
            ubp[k] = a;
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = VI(i, j);
            double a = 0.0;
            if (m->v_inner[k]) {
                if (j >= 1) a += 0.5 * cvb[HI(i, j - 1)] + rdy * vyb[HI(i, j - 1)];
                if (j <= ny - 1) a += 0.5 * cvb[HI(i, j)] - rdy * vyb[HI(i, j)];
                if (j >= 1 && j <= ny - 1) {
                    if (i >= 1) a += rdx * scb[CI(i - 1, j - 1)];
                    if (i <= nx - 2) a -= rdx * scb[CI(i, j - 1)];
                }
            }
            vbp[k] = a;
        }
    free(hhb); free(hcb); free(scb); free(sc); free(bub); free(bvb);
    free(cub); free(cvb); free(uxb); free(vyb); free(shb);
}

EXPORT int om_misfit_gradient(om_model *m, int npic, double tol, int maxit, const double *uobs,
                              const double *vobs, double *J, double *grad, wavi_adjoint_report *rep)
{
    const int nu = m->nu, nv = m->nv, nh = m->nh;
    for (int k = 0; k < nu; ++k) if (m->u_isfixed[k]) return WAVI_EARG;
    for (int k = 0; k < nv; ++k) if (m->v_isfixed[k]) return WAVI_EARG;
    if (npic < 1 || m->S.n_iter_viscosity > 8) return WAVI_EARG;
    wavi_adjoint_report lrep;
    if (!rep) rep = &lrep;
    memset(rep, 0, sizeof(*rep));
    rep->n_picard = npic;
    om_ckpt *ck = (om_ckpt *)malloc(sizeof(om_ckpt) * (npic + 1));
    for (int k = 0; k <= npic; ++k) ckpt_alloc(m, &ck[k]);
    double *bu = (double *)malloc(sizeof(double) * nu), *bv = (double *)malloc(sizeof(double) * nv);
    double *xu = (double *)malloc(sizeof(double) * nu), *xv = (double *)malloc(sizeof(double) * nv);
    /* forward */
    for (int k = 0; k < npic; ++k) {
        ckpt_save(m, &ck[k]);
        om_inner_update(m);
        om_rhs(m, bu, bv);
        double rel;
        if (!solve_linear(m, bu, bv, xu, xv, tol, maxit, &rep->fwd_vcycles, &rel)) rep->unconverged++;
        if (rel > rep->max_fwd_rel) rep->max_fwd_rel = rel;
        for (int q = 0; q < nu; ++q) if (m->u_inner[q]) m->f[WAVI_U_U][q] = xu[q];
        for (int q = 0; q < nv; ++q) if (m->v_inner[q]) m->f[WAVI_V_V][q] = xv[q];
    }
    ckpt_save(m, &ck[npic]);
    double js = 0.0;
    for (int q = 0; q < nu; ++q) if (m->u_inner[q]) { double d = m->f[WAVI_U_U][q] - uobs[q]; js += d * d; }
    for (int q = 0; q < nv; ++q) if (m->v_inner[q]) { double d = m->f[WAVI_V_V][q] - vobs[q]; js += d * d; }
    *J = 0.5 * js;
    if (grad) {
        double *ubar = (double *)calloc(nu, sizeof(double)), *vbar = (double *)calloc(nv, sizeof(double));
        double *etab = (double *)calloc((size_t)nh * m->ns, sizeof(double));
        double *betab = (double *)calloc(nh, sizeof(double)), *f2b = (double *)calloc(nh, sizeof(double));
        double *wcb = (double *)calloc(nh, sizeof(double));
        for (int q = 0; q < nu; ++q) ubar[q] = m->u_inner[q] ? ck[npic].u[q] - uobs[q] : 0.0;
        for (int q = 0; q < nv; ++q) vbar[q] = m->v_inner[q] ? ck[npic].v[q] - vobs[q] : 0.0;
        for (int k = npic; k >= 1; --k) {
            ckpt_load(m, &ck[k - 1]);
            om_inner_update(m);
            double rel;
            if (!solve_linear(m, ubar, vbar, xu, xv, tol, maxit, &rep->adj_vcycles, &rel)) rep->unconverged++;
            if (rel > rep->max_adj_rel) rep->max_adj_rel = rel;
            picard_reverse(m, &ck[k - 1], ck[k].u, ck[k].v, xu, xv, etab, betab, f2b, ubar, vbar, wcb);
        }
        const double *wc = m->f[WAVI_H_WEERTMAN_C];
        for (int q = 0; q < nh; ++q) grad[q] = wc[q] * wcb[q];
        /* leave the state of the last forward iteration */
        ckpt_load(m, &ck[npic - 1]);
        om_inner_update(m);
        memcpy(m->f[WAVI_U_U], ck[npic].u, sizeof(double) * nu);
        memcpy(m->f[WAVI_V_V], ck[npic].v, sizeof(double) * nv);
        free(ubar); free(vbar); free(etab); free(betab); free(f2b); free(wcb);
    }
    for (int k = 0; k <= npic; ++k) ckpt_free(&ck[k]);
    free(ck); free(bu); free(bv); free(xu); free(xv);
    return 0;
}
