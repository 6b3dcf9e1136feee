// wavi_device.h - device-side building blocks of the gfx950 velocity solve.
//
// Everything here is written for 64-wide wavefronts and 256-thread
// workgroups that own one 16 x 16 block of the u grid and the v grid.  All
// FP64 arithmetic is written in the order in which WAVI.jl's sparse and
// Kronecker products accumulate it (src/KroneckerProduct.jl:57-73,
// src/utilities.jl:61-145); the library is compiled with -ffp-contract=off
// so every a*b+c keeps two roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wavi {

constexpr int TILE = 16;              // reduction / transform tile edge
constexpr int TPB = TILE * TILE;      // threads per block-tile workgroup
constexpr int MAX_HAAR_NNZ = 16;      // a row of the forward or transposed inverse touches a whole block

// ---------------------------------------------------------------------------
// Geometry of the staggered grids.
struct GridDims {
    int nx, ny;           // h grid
    int nxu, nyu;         // u grid (nx+1, ny)
    int nxv, nyv;         // v grid (nx, ny+1)
    int nbxu, nbyu;       // 16x16 blocks of the u grid
    int nbxv, nbyv;       // 16x16 blocks of the v grid
    int nbx, nby;         // launch grid of block-tile kernels: max of u and v
    int npart;            // nbxu*nbyu + nbxv*nbyv partials of a [u; v] reduction
};

// ---------------------------------------------------------------------------
// Deterministic x^y (double-double log and exp; see the oracle's det_pow for
// the derivation).  Uses only IEEE +,-,*,/, fma, frexp, ldexp and rint, so it
// returns the same bits on the CPU and the GPU.
__device__ __forceinline__ double det_pow(double x, double y)
{
    const double LN2_HI = 0x1.62e42fefa39efp-1;
    const double LN2_LO = 0x1.abc9e3b39803fp-56;
    const double INV_LN2 = 0x1.71547652b82fep+0;
    if (y == 0.0) return 1.0;
    if (__builtin_isnan(x) || __builtin_isnan(y)) return __builtin_nan("");
    if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_inf();
    if (x < 0.0) return __builtin_nan("");
    if (__builtin_isinf(x)) return y > 0.0 ? __builtin_inf() : 0.0;
    int k;
    double m = frexp(x, &k);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        k = k - 1;
    }
    double f = m - 1.0;
    double d_hi = 2.0 + f;
    double d_lo = f - (d_hi - 2.0);
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
    p = p * t;
    double lm_hi = 2.0 * s_hi;
    double lm_lo = 2.0 * s_lo + 2.0 * (s_hi * p);
    double dk = (double)k;
    double kh = dk * LN2_HI;
    double kl = fma(dk, LN2_HI, -kh) + dk * LN2_LO;
    double sh = kh + lm_hi;
    double bb = sh - kh;
    double sl = (kh - (sh - bb)) + (lm_hi - bb);
    sl = sl + (kl + lm_lo);
    double lh = sh + sl;
    double ll = sl - (lh - sh);
    double zh = y * lh;
    double zl = fma(y, lh, -zh) + y * ll;
    if (zh > 709.78) return __builtin_inf();
    if (zh < -745.2) return 0.0;
    double n = rint(zh * INV_LN2);
    double r = fma(-n, LN2_HI, zh);
    r = r - n * LN2_LO;
    r = r + zl;
    double q = 1.0 / 6227020800.0;
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
    double em1 = r + (r * r) * q;
    double ee = 1.0 + em1;
    return ldexp(ee, (int)n);
}

// ---------------------------------------------------------------------------
// Canonical reduction trees.  A 256-thread workgroup reduces one value per
// thread: each 64-wide wavefront sums its lanes with t[e] += t[e+s],
// s = 32..1 (shuffles, lane e reads lane e+s), and the four wavefront sums
// combine as (w0 + w2) + (w1 + w3).  The oracle sums in the same order.
template <typename T>
__device__ __forceinline__ T tree256(T v, T *lds)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    if (l == 0) lds[w] = v;
    __syncthreads();
    const T r = (lds[0] + lds[2]) + (lds[1] + lds[3]);
    __syncthreads();
    return r;
}

// Two independent trees in one pass.
__device__ __forceinline__ void tree256x2(double &a, double &b, double *lds)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        a = a + __shfl_down(a, s, 64);
        b = b + __shfl_down(b, s, 64);
    }
    if (l == 0) {
        lds[w] = a;
        lds[4 + w] = b;
    }
    __syncthreads();
    a = (lds[0] + lds[2]) + (lds[1] + lds[3]);
    b = (lds[4] + lds[6]) + (lds[5] + lds[7]);
    __syncthreads();
}

// Fold `np` block partials into 256 lanes (lane L = thread L sums P[L],
// P[L+256], ... in order; missing terms add +0) and sum the lanes with
// tree256.  Loads are issued eight at a time.  Every workgroup that calls
// this obtains the same bits, in all threads.
__device__ __forceinline__ double lanes_sum(const double *P, int np, double *lds)
{
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int q0 = 0; q0 < np; q0 += 8 * 256) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = q0 + t + 256 * k;
            v[k] = q < np ? P[q] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = acc + v[k];
    }
    return tree256(acc, lds);
}

// lanes_sum of two partial arrays of the same length in one pass.
__device__ __forceinline__ void lanes_sum2(const double *P, const double *Q, int np, double *lds, double &a,
                                           double &b)
{
    const int t = threadIdx.x;
    double x = 0.0, y = 0.0;
    for (int q0 = 0; q0 < np; q0 += 8 * 256) {
        double v[8], w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = q0 + t + 256 * k;
            v[k] = q < np ? P[q] : 0.0;
            w[k] = q < np ? Q[q] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x = x + v[k];
            y = y + w[k];
        }
    }
    tree256x2(x, y, lds);
    a = x;
    b = y;
}

// ---------------------------------------------------------------------------
// Non-temporal store (streamed past the caches).
template <typename T>
__device__ __forceinline__ void st_nt(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// ---------------------------------------------------------------------------
// Lane exchange within the 16-lane rows of a wavefront by DPP (every lane of
// the wavefront must be active).
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]: lane ^ 1
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]: lane ^ 2
constexpr int DPP_MIRROR = 0x140;      // row_mirror: lane ^ 15
constexpr int DPP_HALF_MIRROR = 0x141; // row_half_mirror: lane ^ 7

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over a 16-lane row, left in every lane of the row.  Each stage adds a
// lane and its partner, so both get the same bits and the result does not
// depend on the lane.
__device__ __forceinline__ double row_sum16(double v)
{
    v = v + dpp_d<DPP_XOR1>(v);
    v = v + dpp_d<DPP_XOR2>(v);
    v = v + dpp_d<DPP_HALF_MIRROR>(v);
    v = v + dpp_d<DPP_MIRROR>(v);
    return v;
}

// DPP with a row mask: rows not in ROWMASK read 0.
constexpr int DPP_ROW_BCAST15 = 0x142; // lane 15 of each row to the next row
constexpr int DPP_ROW_BCAST31 = 0x143; // lane 31 to rows 2 and 3

template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_rows_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_rows_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wavefront total from row sums (every lane of a row holding its row's sum):
// rows 1 and 3 add rows 0 and 2, then rows 2 and 3 add rows 0 + 1; lane 63
// ends with (r3 + r2) + (r1 + r0).
__device__ __forceinline__ double wave_total_rows(double v)
{
    v = v + dpp_rows_d<DPP_ROW_BCAST15, 0xA>(v);
    return v + dpp_rows_d<DPP_ROW_BCAST31, 0xC>(v);
}

__device__ __forceinline__ float wave_total_rows_f(float v)
{
    v = v + dpp_rows_f<DPP_ROW_BCAST15, 0xA>(v);
    return v + dpp_rows_f<DPP_ROW_BCAST31, 0xC>(v);
}

// Workgroup sums of a and b (256 threads): wavefront totals by DPP, the four
// totals through LDS (red holds 8 doubles and must not be written again
// before the next barrier).  Every thread gets the same totals.
__device__ __forceinline__ void wg_sum2(double &a, double &b, double *red)
{
    a = wave_total_rows(row_sum16(a));
    b = wave_total_rows(row_sum16(b));
    if ((threadIdx.x & 63) == 63) {
        red[threadIdx.x >> 6] = a;
        red[4 + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    a = (red[0] + red[1]) + (red[2] + red[3]);
    b = (red[4] + red[5]) + (red[6] + red[7]);
}

// ---------------------------------------------------------------------------
// The momentum operator op_fun! (src/utilities.jl:61-145), evaluated at one
// u point or one v point from an input accessor.  Intermediates (r_xx, r_yy,
// r_xy, extra) are recomputed by each output that needs them with the same
// expression, so a point's value does not depend on which kernel computes it.
template <typename T>
struct OpCoef {
    const T *heta_h;    // h*eta_av on the h mask, 0 off it          (gh.dneghηav)
    const T *heta_c;    // corner average of h*eta_av on the c mask   (gc.dneghηav)
    const T *dimp;      // -rho_i g theta dt dsdh on the h mask       (gh.dimplicit)
    const T *hu, *hv;   // thickness on u and v points                (gu.h, gv.h)
    const T *bu, *bv;   // beta_eff on u and v points                 (gu.βeff, gv.βeff)
    const uint8_t *hm, *um, *vm, *cm, *ui, *vi;
    int nx, ny;
    T rdx, rdy;
};

// Unit vector at one u (or v) point: the probe vector of get_op_diag seen
// from inside the stencil's reach.
template <typename T>
struct DeltaIn {
    int i0, j0, comp;  // comp 0: u point, 1: v point
    __device__ __forceinline__ T u(int i, int j) const { return (comp == 0 && i == i0 && j == j0) ? T(1) : T(0); }
    __device__ __forceinline__ T v(int i, int j) const { return (comp == 1 && i == i0 && j == j0) ? T(1) : T(0); }
};

template <typename T, class In>
__device__ __forceinline__ T op_rxx(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int k = i + c.nx * j;
    if (!c.hm[k]) return T(0);
    T dudx = c.rdx * in.u(i + 1, j) - c.rdx * in.u(i, j);
    T dvdy = c.rdy * in.v(i, j + 1) - c.rdy * in.v(i, j);
    return T(2) * (c.heta_h[k] * (T(2) * dudx + dvdy));
}

template <typename T, class In>
__device__ __forceinline__ T op_ryy(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int k = i + c.nx * j;
    if (!c.hm[k]) return T(0);
    T dudx = c.rdx * in.u(i + 1, j) - c.rdx * in.u(i, j);
    T dvdy = c.rdy * in.v(i, j + 1) - c.rdy * in.v(i, j);
    return T(2) * (c.heta_h[k] * (T(2) * dvdy + dudx));
}

// corner (i, j) of the c grid, i < nx-1, j < ny-1
template <typename T, class In>
__device__ __forceinline__ T op_rxy(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int k = i + (c.nx - 1) * j;
    if (!c.cm[k]) return T(0);
    T dudy = c.rdy * in.u(i + 1, j + 1) - c.rdy * in.u(i + 1, j);
    T dvdx = c.rdx * in.v(i + 1, j + 1) - c.rdx * in.v(i, j + 1);
    return c.heta_c[k] * (dudy + dvdx);
}

template <typename T, class In>
__device__ __forceinline__ T op_extra(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int nxu = c.nx + 1;
    const int k = i + c.nx * j;
    T qx1 = c.um[(i + 1) + nxu * j] ? c.hu[(i + 1) + nxu * j] * in.u(i + 1, j) : T(0);
    T qx0 = c.um[i + nxu * j] ? c.hu[i + nxu * j] * in.u(i, j) : T(0);
    T qy1 = c.vm[i + c.nx * (j + 1)] ? c.hv[i + c.nx * (j + 1)] * in.v(i, j + 1) : T(0);
    T qy0 = c.vm[i + c.nx * j] ? c.hv[i + c.nx * j] * in.v(i, j) : T(0);
    T divq = (c.rdx * qx1 - c.rdx * qx0) + (c.rdy * qy1 - c.rdy * qy0);
    return c.dimp[k] * divq;
}

// fx at u point (i, j), 0 <= i <= nx, 0 <= j < ny (caller checks inner mask)
template <typename T, class In>
__device__ __forceinline__ T op_u(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int nx = c.nx, ny = c.ny;
    const int k = i + (nx + 1) * j;
    T a1 = i < nx ? op_rxx(c, in, i, j) : T(0);
    T a0 = i > 0 ? op_rxx(c, in, i - 1, j) : T(0);
    T d_rxx_dx = c.rdx * a1 - c.rdx * a0;
    T c1 = (i >= 1 && i <= nx - 1 && j <= ny - 2) ? op_rxy(c, in, i - 1, j) : T(0);
    T c0 = (i >= 1 && i <= nx - 1 && j >= 1) ? op_rxy(c, in, i - 1, j - 1) : T(0);
    T d_rxy_dy = c.rdy * c1 - c.rdy * c0;
    T taubx = c.um[k] ? c.bu[k] * in.u(i, j) : T(0);
    T e1 = i < nx ? op_extra(c, in, i, j) : T(0);
    T e0 = i > 0 ? op_extra(c, in, i - 1, j) : T(0);
    T d_extra_dx = c.rdx * e1 - c.rdx * e0;
    return ((d_rxx_dx + d_rxy_dy) - taubx) - c.hu[k] * d_extra_dx;
}

// fy at v point (i, j), 0 <= i < nx, 0 <= j <= ny
template <typename T, class In>
__device__ __forceinline__ T op_v(const OpCoef<T> &c, const In &in, int i, int j)
{
    const int nx = c.nx, ny = c.ny;
    const int k = i + nx * j;
    T a1 = j < ny ? op_ryy(c, in, i, j) : T(0);
    T a0 = j > 0 ? op_ryy(c, in, i, j - 1) : T(0);
    T d_ryy_dy = c.rdy * a1 - c.rdy * a0;
    T c1 = (j >= 1 && j <= ny - 1 && i <= nx - 2) ? op_rxy(c, in, i, j - 1) : T(0);
    T c0 = (j >= 1 && j <= ny - 1 && i >= 1) ? op_rxy(c, in, i - 1, j - 1) : T(0);
    T d_rxy_dx = c.rdx * c1 - c.rdx * c0;
    T tauby = c.vm[k] ? c.bv[k] * in.v(i, j) : T(0);
    T e1 = j < ny ? op_extra(c, in, i, j) : T(0);
    T e0 = j > 0 ? op_extra(c, in, i, j - 1) : T(0);
    T d_extra_dy = c.rdy * e1 - c.rdy * e0;
    return ((d_ryy_dy + d_rxy_dx) - tauby) - c.hv[k] * d_extra_dy;
}

// ---------------------------------------------------------------------------
// Haar tables: the composed 1-D matrix of wavelet_matrix(m, levels, dir)
// (src/Wavelets/update_wavelets.jl:38-89) for every block length m <= B,
// row r: nnz[m][r] terms (col, coef), columns increasing.  dir 0 forward,
// 1 inverse, 2 transposed inverse.  Coefficients are +-2^-k (k <= 4), exact in
// float.  Filled on the host at context creation.
struct HaarTable {
    int B;
    int8_t nnz[3][17][16];
    int8_t col[3][17][16][MAX_HAAR_NNZ];
    float coef[3][17][16][MAX_HAAR_NNZ];
};

}  // namespace wavi
