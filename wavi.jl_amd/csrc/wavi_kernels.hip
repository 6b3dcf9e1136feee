// wavi_kernels.hip - gfx950 kernels of the WAVI.jl velocity solve.
//
// Layout: every field is a full grid in HBM, x fastest (u: (nx+1) x ny,
// v: nx x (ny+1), h: nx x ny, c: (nx-1) x (ny-1), sigma: nx x ny x nsigma
// with sigma slowest).  Vectors of the reference's "inner" unknowns are full
// u/v grids that are zero off the inner masks, so no gather/scatter index
// maps exist on the device.
//
// Kernels named k_tile_* run one 256-thread workgroup per 16 x 16 block of
// the u grid and of the v grid (grid (nbx, nby)); they own the Haar blocks
// (levels <= 3) and write one reduction partial per block.  Other kernels
// are flat (256 threads, one point each).
#include "wavi_device.h"

namespace wavi {

// ===========================================================================
// small utilities

// p[k] where ok, else dflt, with the load issued unconditionally (at index 0
// where !ok; every array here holds at least one element, the grid being at
// least 2 x 2): a guarded load becomes a branch with a wait of its own, so
// that a launch's loads would go out one round trip after another.
template <typename V>
__device__ __forceinline__ V ld_or(bool ok, const V *p, int k, V dflt)
{
    const V v = p[ok ? k : 0];
    return ok ? v : dflt;
}

template <typename T>
__global__ void k_fill(T *a, T v, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

// u / v pairs of the elementwise kernels above in one launch: blockIdx.y
// selects the u (0) or v (1) arrays.
template <typename Tin, typename Tout>
__global__ void k_select_uv(const Tin *iu, const uint8_t *mu, Tout *ou, int nu, const Tin *iv, const uint8_t *mv,
                            Tout *ov, int nv)
{
    const bool v = blockIdx.y != 0;
    const Tin *in = v ? iv : iu;
    const uint8_t *mask = v ? mv : mu;
    Tout *out = v ? ov : ou;
    const int n = v ? nv : nu;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = mask[k] ? (Tout)in[k] : Tout(0);
}

template <typename T>
__global__ void k_copy_masked_uv(T *du, const T *su, const uint8_t *mu, int nu, T *dv, const T *sv,
                                 const uint8_t *mv, int nv)
{
    const bool v = blockIdx.y != 0;
    T *dst = v ? dv : du;
    const T *src = v ? sv : su;
    const uint8_t *mask = v ? mv : mu;
    const int n = v ? nv : nu;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && mask[k]) dst[k] = src[k];
}

__global__ void k_set_velocities_uv(double *U, const double *xu, const double *eu, const uint8_t *iu, int nu,
                                    double *V, const double *xv, const double *ev, const uint8_t *iv, int nv)
{
    const bool v = blockIdx.y != 0;
    double *W = v ? V : U;
    const double *x = v ? xv : xu, *e = v ? ev : eu;
    const uint8_t *inner = v ? iv : iu;
    const int n = v ? nv : nu;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && inner[k]) W[k] = x[k] + e[k];
}

template <typename T>
__global__ void k_fill_uv(T *au, int nu, T *av, int nv, T val)
{
    const bool v = blockIdx.y != 0;
    T *a = v ? av : au;
    const int n = v ? nv : nu;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = val;
}

// Up to CONV_MAX conversions in one launch: blockIdx.y selects the array.
constexpr int CONV_MAX = 12;
template <typename Tin, typename Tout>
struct ConvList {
    const Tin *in[CONV_MAX];
    Tout *out[CONV_MAX];
    int n[CONV_MAX];
};

template <typename Tin, typename Tout>
__global__ void k_convert_list(ConvList<Tin, Tout> L)
{
    const int a = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L.n[a]) L.out[a][k] = (Tout)L.in[a][k];
}

// out = mask ? in : 0
template <typename Tin, typename Tout>
__global__ void k_select(const Tin *in, const uint8_t *mask, Tout *out, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = mask[k] ? (Tout)in[k] : Tout(0);
}

// dst = src where mask
template <typename T>
__global__ void k_copy_masked(T *dst, const T *src, const uint8_t *mask, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && mask[k]) dst[k] = src[k];
}

// mask as 0.0/1.0 <-> uint8
__global__ void k_mask_from_double(const double *m, uint8_t *out, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = m[k] > 0.5 ? 1 : 0;
}

__global__ void k_mask_to_double(const uint8_t *m, double *out, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = m[k] ? 1.0 : 0.0;
}

// ===========================================================================
// Canonical [u; v] dot-product partials: one per 16x16 block (u blocks then
// v blocks).  b == nullptr gives sum of squares.
template <typename T>
__global__ void k_tile_dot(GridDims g, const T *au, const T *av, const T *bu, const T *bv, double *P)
{
    __shared__ double lds[256];
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int i = bx * TILE + lx, j = by * TILE + ly;
    if (bx < g.nbxu && by < g.nbyu) {
        double v = 0.0;
        if (i < g.nxu && j < g.nyu) {
            int k = i + g.nxu * j;
            v = bu ? (double)au[k] * (double)bu[k] : (double)au[k] * (double)au[k];
        }
        double s = tree256(v, lds);
        if (threadIdx.x == 0) P[bx + g.nbxu * by] = s;
    }
    if (bx < g.nbxv && by < g.nbyv) {
        double v = 0.0;
        if (i < g.nxv && j < g.nyv) {
            int k = i + g.nxv * j;
            v = bv ? (double)av[k] * (double)bv[k] : (double)av[k] * (double)av[k];
        }
        double s = tree256(v, lds);
        if (threadIdx.x == 0) P[g.nbxu * g.nbyu + bx + g.nbxv * by] = s;
    }
}

// out[slot] = sqrt(sum of partials) for each of `count` partial arrays.
__global__ void k_finish_norms(const double *P0, const double *P1, int np, double *out)
{
    __shared__ double lds[256];
    double a = lanes_sum(P0, np, lds);
    double b = P1 ? lanes_sum(P1, np, lds) : 0.0;
    if (threadIdx.x == 0) {
        out[0] = sqrt(a);
        out[1] = sqrt(b);
    }
}

// out[c] = sum of partial array c (P + c * np), c < count (no square root)
__global__ void k_finish_sums(const double *P, int np, int count, double *out)
{
    __shared__ double lds[256];
    for (int c = 0; c < count; ++c) {
        double a = lanes_sum(P + (size_t)c * np, np, lds);
        if (threadIdx.x == 0) out[c] = a;
        __syncthreads();
    }
}

// flexible CG: x += alpha p, r -= alpha q on the u then the v grid
__global__ void k_fcg_step(int nu, int nv, double alpha, double *xu, double *xv, double *ru, double *rv,
                           const double *pu, const double *pv, const double *qu, const double *qv)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nu) {
        xu[t] = xu[t] + alpha * pu[t];
        ru[t] = ru[t] - alpha * qu[t];
    } else if (t < nu + nv) {
        t -= nu;
        xv[t] = xv[t] + alpha * pv[t];
        rv[t] = rv[t] - alpha * qv[t];
    }
}

// flexible CG: p = z + beta p at inner points, 0 elsewhere
__global__ void k_fcg_dir(int nu, int nv, double beta, const uint8_t *ui, const uint8_t *vi, const double *zu,
                          const double *zv, double *pu, double *pv)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nu) {
        pu[t] = ui[t] ? zu[t] + beta * pu[t] : 0.0;
    } else if (t < nu + nv) {
        t -= nu;
        pv[t] = vi[t] ? zv[t] + beta * pv[t] : 0.0;
    }
}

// ===========================================================================
// update_state! steps before the solve (src/Models/update_state.jl:51-147).

struct GeomArgs {
    const uint8_t *hm, *um, *vm;
    const double *b, *h;
    double *s, *haf, *q;  // q: 4 quadrant fractions, 4*nh
    double *gfh, *gfu, *gfv;
    double *hu, *su, *hv, *sv;
    double *wc, *dsdh, *acc, *mel;
    const double *wc_param, *acc_param;
    double sl, ri, rw, melt;
};

// update_surface_elevation! (:51-57), update_height_above_floatation! (:80-85)
__global__ void k_surface_haf(GridDims g, GeomArgs a)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nx * g.ny) return;
    double h = a.h[k], b = a.b[k];
    if (a.hm[k]) {
        double s1 = b + h, s2 = a.sl + h * (1.0 - a.ri / a.rw);
        double s = s1 > s2 ? s1 : s2;
        if (__builtin_isnan(s1) || __builtin_isnan(s2)) s = __builtin_nan("");
        a.s[k] = s;
    }
    a.haf[k] = h - (a.rw / a.ri) * (a.sl - b);
}

// pos_fraction quadrant area (src/utilities.jl:263-360) of one cell from the
// values at the cell (v1, its mask m1) and at the quadrant's two neighbours
// (z2, z3 with masks m2, m3; an unmasked neighbour takes the cell's value).
__device__ __forceinline__ double pos_fraction_area(double v1, bool m1, double z2, bool m2, double z3, bool m3)
{
    double v2 = m2 ? z2 : v1;
    double v3 = m3 ? z3 : v1;
    double s1 = (v1 > 0) - (v1 < 0), s2 = (v2 > 0) - (v2 < 0), s3 = (v3 > 0) - (v3 < 0);
    double area = 0.0;
    if (!((s1 == s2) && (s1 == s3))) {
        bool flip = fabs(v3 - v1) < fabs(v2 - v1);
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
    if (!m1) area = 0.0;
    return area;
}


// update_grounded_fraction (:92-99) on h, then accumulation, melt,
// weertman_c and dsdh (:106-147)
__global__ void k_gf_h(GridDims g, GeomArgs a)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nh = g.nx * g.ny;
    if (k >= nh) return;
    int i = k % g.nx, j = k / g.nx;
    // the cell and its four neighbours (clamped), every load before the
    // first store
    const int M = g.nx, N = g.ny;
    const int kE = (i + 1 < M ? i + 1 : M - 1) + M * j, kW = (i > 0 ? i - 1 : 0) + M * j;
    const int kN = i + M * (j + 1 < N ? j + 1 : N - 1), kS = i + M * (j > 0 ? j - 1 : 0);
    const double zc = a.haf[k], zE = a.haf[kE], zN = a.haf[kN], zW = a.haf[kW], zS = a.haf[kS];
    const bool mc = a.hm[k], mE = a.hm[kE], mN = a.hm[kN], mW = a.hm[kW], mS = a.hm[kS];
    const double accp = a.acc_param[k], wcp = a.wc_param[k];
    // quadrants 1..4 take points (2, 3) = (E, N), (N, W), (W, S), (S, E),
    // each clamped to the cell at the grid's edge (utilities.jl:263-360)
    const double q[4] = {pos_fraction_area(zc, mc, zE, mE, zN, mN), pos_fraction_area(zc, mc, zN, mN, zW, mW),
                         pos_fraction_area(zc, mc, zW, mW, zS, mS), pos_fraction_area(zc, mc, zS, mS, zE, mE)};
    double af = 0.0;
#pragma unroll
    for (int quad = 1; quad <= 4; ++quad) {
        a.q[(quad - 1) * nh + k] = q[quad - 1];
        af = af + 0.25 * q[quad - 1];
    }
    a.gfh[k] = af;
    a.acc[k] = accp;
    a.mel[k] = a.melt;
    a.wc[k] = wcp * af;
    a.dsdh[k] = (1.0 - a.ri / a.rw) + (a.ri / a.rw) * af;
}

// masked h -> face averages of update_geometry_on_uv_grids! (:65-73) and the
// u/v grounded fractions of pos_fraction (utilities.jl:367-390)
__global__ void k_geom_uv(GridDims g, GeomArgs a)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = g.nxu * g.nyu, nv = g.nxv * g.nyv, nh = g.nx * g.ny;
    const int nx = g.nx, ny = g.ny;
    if (t >= nu + nv) return;
    // the face between h cells ka (left / below) and kb; every load issued
    // before the first is used (ld_or), the sums in the same order as before
    const bool v = t >= nu;
    const int tt = v ? t - nu : t;
    const int i = v ? tt % g.nxv : tt % g.nxu, j = v ? tt / g.nxv : tt / g.nxu;
    const bool oka = v ? j >= 1 : i >= 1, okb = v ? j < ny : i < nx;
    const int ka = v ? i + nx * (j - 1) : (i - 1) + nx * j, kb = i + nx * j;
    // grounded fraction: on u quadrants 1, 4 from cell i - 1 and 2, 3 from
    // cell i; on v quadrants 1, 2 from cell j - 1 and 3, 4 from cell j
    const double q1 = ld_or(oka, a.q, ka, 0.0);
    const double q2 = v ? ld_or(oka, a.q, nh + ka, 0.0) : ld_or(okb, a.q, nh + kb, 0.0);
    const double q3 = ld_or(okb, a.q, 2 * nh + kb, 0.0);
    const double q4 = v ? ld_or(okb, a.q, 3 * nh + kb, 0.0) : ld_or(oka, a.q, 3 * nh + ka, 0.0);
    const bool fm = (v ? a.vm : a.um)[tt];
    const uint8_t ha = ld_or(oka, a.hm, ka, (uint8_t)0), hb = ld_or(okb, a.hm, kb, (uint8_t)0);
    const double hha = ld_or(oka, a.h, ka, 0.0), hhb = ld_or(okb, a.h, kb, 0.0);
    const double sa = ld_or(oka, a.s, ka, 0.0), sb = ld_or(okb, a.s, kb, 0.0);
    double gf = 0.0;
    gf = gf + 0.25 * q1;
    gf = gf + 0.25 * q2;
    gf = gf + 0.25 * q3;
    gf = gf + 0.25 * q4;
    (v ? a.gfv : a.gfu)[tt] = gf;
    if (fm) {
        double den = 0.0, nh_ = 0.0, ns_ = 0.0;
        if (oka) {
            den = den + 0.5 * (double)ha;
            nh_ = nh_ + 0.5 * (ha ? hha : 0.0);
            ns_ = ns_ + 0.5 * (ha ? sa : 0.0);
        }
        if (okb) {
            den = den + 0.5 * (double)hb;
            nh_ = nh_ + 0.5 * (hb ? hhb : 0.0);
            ns_ = ns_ + 0.5 * (hb ? sb : 0.0);
        }
        (v ? a.hv : a.hu)[tt] = nh_ / den;
        (v ? a.sv : a.su)[tt] = ns_ / den;
    }
}

// ===========================================================================
// inner_update! (src/Models/update_velocities.jl:32-45, 202-383)

struct InnerArgs {
    const uint8_t *hm, *um, *vm, *cm;
    const double *U, *V;
    const double *h, *wc, *gfh, *gfu, *gfv, *dsdh, *glen_b;
    const double *zeta, *qw;       // nsigma
    double *sr, *avs, *bs, *beta, *tau, *eta, *etaav, *f1, *f2, *beff;
    double *bu, *bv;
    double *heta_h, *heta_c, *dimp;
    double rdx, rdy;
    double wexp, wreg2, gexp, greg2, c0;
    int ns, niter_visc, xfirst_c2h, yfirst_h2c;
};

// shear strain (uy + vx) on the c mask at corner (a, b), 0 outside the c grid
__device__ __forceinline__ double corner_shear(const GridDims &g, const InnerArgs &p, int a, int b)
{
    const int nxc = g.nx - 1, nyc = g.ny - 1;
    const bool in = a >= 0 && a < nxc && b >= 0 && b < nyc;
    const int nxu = g.nxu, nx = g.nx;
    // every load issued at once (ld_or), values selected
    const int ku1 = (a + 1) + nxu * (b + 1), ku0 = (a + 1) + nxu * b;
    const int kv1 = (a + 1) + nx * (b + 1), kv0 = a + nx * (b + 1);
    const bool cm = ld_or(in, p.cm, a + nxc * b, (uint8_t)0) != 0;
    const bool mu1 = ld_or(in, p.um, ku1, (uint8_t)0) != 0, mu0 = ld_or(in, p.um, ku0, (uint8_t)0) != 0;
    const bool mv1 = ld_or(in, p.vm, kv1, (uint8_t)0) != 0, mv0 = ld_or(in, p.vm, kv0, (uint8_t)0) != 0;
    const double U1 = ld_or(in, p.U, ku1, 0.0), U0 = ld_or(in, p.U, ku0, 0.0);
    const double V1 = ld_or(in, p.V, kv1, 0.0), V0 = ld_or(in, p.V, kv0, 0.0);
    const double u1 = mu1 ? U1 : 0.0, u0 = mu0 ? U0 : 0.0;
    const double v1 = mv1 ? V1 : 0.0, v0 = mv0 ? V0 : 0.0;
    const double dudy = p.rdy * u1 - p.rdy * u0;
    const double dvdx = p.rdx * v1 - p.rdx * v0;
    return (in && cm) ? dudy + dvdx : 0.0;
}

// gc.cent (CGrid.jl:45) at h cell (i, j) in the reference's factor order
__device__ __forceinline__ double cent_c_to_h(const GridDims &g, const InnerArgs &p, int i, int j)
{
    const int nxc = g.nx - 1, nyc = g.ny - 1;
    // the four corners' loads first, then the sums in the reference's order
    const double s00 = corner_shear(g, p, i - 1, j - 1), s10 = corner_shear(g, p, i, j - 1);
    const double s01 = corner_shear(g, p, i - 1, j), s11 = corner_shear(g, p, i, j);
    double out = 0.0;
    if (p.xfirst_c2h) {
        if (j - 1 >= 0 && j - 1 < nyc) {
            double t0 = 0.0;
            if (i - 1 >= 0 && i - 1 < nxc) t0 = t0 + 0.5 * s00;
            if (i >= 0 && i < nxc) t0 = t0 + 0.5 * s10;
            out = out + 0.5 * t0;
        }
        if (j >= 0 && j < nyc) {
            double t1 = 0.0;
            if (i - 1 >= 0 && i - 1 < nxc) t1 = t1 + 0.5 * s01;
            if (i >= 0 && i < nxc) t1 = t1 + 0.5 * s11;
            out = out + 0.5 * t1;
        }
    } else {
        if (i - 1 >= 0 && i - 1 < nxc) {
            double t0 = 0.0;
            if (j - 1 >= 0 && j - 1 < nyc) t0 = t0 + 0.5 * s00;
            if (j >= 0 && j < nyc) t0 = t0 + 0.5 * s01;
            out = out + 0.5 * t0;
        }
        if (i >= 0 && i < nxc) {
            double t1 = 0.0;
            if (j - 1 >= 0 && j - 1 < nyc) t1 = t1 + 0.5 * s10;
            if (j >= 0 && j < nyc) t1 = t1 + 0.5 * s11;
            out = out + 0.5 * t1;
        }
    }
    return out;
}

// All h-cell updates of inner_update!: strain rate, speeds, beta, basal drag,
// viscosity on the sigma levels, quadratures, beta_eff and the h-grid
// rheological coefficients.
constexpr int INNER_NS_PREFETCH = 8;  // sigma levels whose eta and glen_b k_inner_h loads up front

__global__ void k_inner_h(GridDims g, InnerArgs p)
{
    const int nh = g.nx * g.ny;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh) return;
    const int i = k % g.nx, j = k / g.nx;
    const int nxu = g.nxu, nx = g.nx;
    // every load of the cell issued before the first is used: masks with
    // their fields (the values selected), the shear's corners, and eta and
    // glen_b on up to INNER_NS_PREFETCH levels
    const bool hm = p.hm[k];
    const double Ua = p.U[i + nxu * j], Ub = p.U[(i + 1) + nxu * j];
    const double Va = p.V[i + nx * j], Vb = p.V[i + nx * (j + 1)];
    const bool ma = p.um[i + nxu * j], mb = p.um[(i + 1) + nxu * j];
    const bool na = p.vm[i + nx * j], nb = p.vm[i + nx * (j + 1)];
    const double shc = cent_c_to_h(g, p, i, j);
    double eta_k[INNER_NS_PREFETCH], gb_k[INNER_NS_PREFETCH], zeta_k[INNER_NS_PREFETCH], qw_k[INNER_NS_PREFETCH];
#pragma unroll
    for (int ks = 0; ks < INNER_NS_PREFETCH; ++ks) {
        eta_k[ks] = ld_or(ks < p.ns, p.eta, k + nh * ks, 0.0);
        gb_k[ks] = ld_or(ks < p.ns, p.glen_b, k + nh * ks, 0.0);
        zeta_k[ks] = ld_or(ks < p.ns, p.zeta, ks, 0.0);
        qw_k[ks] = ld_or(ks < p.ns, p.qw, ks, 0.0);
    }
    // (and before the first store, which the compiler cannot move them past)
    const double f2_old = p.f2[k], beta_old = p.beta[k], wck = p.wc[k], hk = p.h[k], dsdhk = p.dsdh[k];
    // update_shelf_strain_rate! (:202-209)
    double ux = 0.0, vy = 0.0, sh = 0.0;
    double ub = 0.0, vb = 0.0;
    {
        double u0 = ma ? Ua : 0.0;
        double u1 = mb ? Ub : 0.0;
        double v0 = na ? Va : 0.0;
        double v1 = nb ? Vb : 0.0;
        if (hm) {
            ux = p.rdx * u1 - p.rdx * u0;
            vy = p.rdy * v1 - p.rdy * v0;
            sh = shc;
            ub = 0.5 * u0 + 0.5 * u1;
            vb = 0.5 * v0 + 0.5 * v1;
        }
    }
    double sr = sqrt(((ux * ux + vy * vy) + ux * vy) + 0.25 * (sh * sh));
    p.sr[k] = sr;
    // update_av_speed! (:216-220)
    double avs = sqrt(ub * ub + vb * vb);
    p.avs[k] = avs;
    // update_bed_speed!, update_β!, update_basal_drag! (:227-255)
    double bs = avs / (1.0 + f2_old * beta_old);
    double beta = wck * det_pow(sqrt(bs * bs + p.wreg2), p.wexp);
    double tau = beta * bs;
    p.bs[k] = bs;
    p.beta[k] = beta;
    p.tau[k] = tau;
    // inner_update_viscosity!, update_av_viscosity!, update_quadrature_falpha! (:264-328)
    double av = 0.0, q1 = 0.0, q2 = 0.0;
    if (hm) {
        // level ks from its eta, glen_b, zeta and weight (the levels in
        // increasing order: the prefetched ones unrolled, then any others)
        auto level = [&](int ks, double eta, double gb, double zeta, double w) {
            for (int it = 0; it < p.niter_visc; ++it) {
                double q = (tau * zeta) / eta;
                double arg = sqrt((sr * sr + 0.25 * (q * q)) + p.greg2);
                eta = (0.5 * gb) * det_pow(arg, p.gexp);
            }
            p.eta[k + nh * ks] = eta;
            av = av + w * eta;
            q1 = q1 + ((w * hk) * zeta) / eta;
            q2 = q2 + ((w * hk) * (zeta * zeta)) / eta;
        };
#pragma unroll
        for (int ks = 0; ks < INNER_NS_PREFETCH; ++ks)
            if (ks < p.ns) level(ks, eta_k[ks], gb_k[ks], zeta_k[ks], qw_k[ks]);
        for (int ks = INNER_NS_PREFETCH; ks < p.ns; ++ks)
            level(ks, p.eta[k + nh * ks], p.glen_b[k + nh * ks], p.zeta[ks], p.qw[ks]);
    }
    p.etaav[k] = av;
    p.f1[k] = q1;
    p.f2[k] = q2;
    // update_βeff! (:335-339)
    p.beff[k] = beta / (1.0 + q2 * beta);
    // update_rheological_operators! (:374-383), h-grid parts
    p.heta_h[k] = hm ? hk * av : 0.0;
    p.dimp[k] = hm ? p.c0 * dsdhk : 0.0;
}

// update_βeff_on_uv_grids! (:348-366) and the c-grid coefficient
// gc.dneghηav (:378)
__global__ void k_inner_uvc(GridDims g, InnerArgs p)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = g.nxu * g.nyu, nv = g.nxv * g.nyv, nc = (g.nx - 1) * (g.ny - 1);
    const int nx = g.nx;
    if (t < nu + nv) {
        // the face between h cells ka (left / below) and kb; every load
        // issued before the first is used (ld_or), the sums as before
        const bool v = t >= nu;
        const int tt = v ? t - nu : t;
        const int i = v ? tt % g.nxv : tt % g.nxu, j = v ? tt / g.nxv : tt / g.nxu;
        const bool oka = v ? j >= 1 : i >= 1, okb = v ? j <= g.ny - 1 : i <= nx - 1;
        const int ka = v ? i + nx * (j - 1) : (i - 1) + nx * j, kb = i + nx * j;
        const bool fm = (v ? p.vm : p.um)[tt];
        const double gff = (v ? p.gfv : p.gfu)[tt];
        const uint8_t ha = ld_or(oka, p.hm, ka, (uint8_t)0), hb = ld_or(okb, p.hm, kb, (uint8_t)0);
        const double ba = ld_or(oka, p.beff, ka, 0.0), bb = ld_or(okb, p.beff, kb, 0.0);
        const double ga = ld_or(oka, p.gfh, ka, 0.0), gb = ld_or(okb, p.gfh, kb, 0.0);
        if (!fm) return;
        double den = 0.0, nb = 0.0, ng = 0.0;
        if (oka) {
            den = den + 0.5 * (double)ha;
            nb = nb + 0.5 * (ha ? ba : 0.0);
            ng = ng + 0.5 * (ha ? ga : 0.0);
        }
        if (okb) {
            den = den + 0.5 * (double)hb;
            nb = nb + 0.5 * (hb ? bb : 0.0);
            ng = ng + 0.5 * (hb ? gb : 0.0);
        }
        double b = nb / den;
        double ip = ng / den;
        if (ip > 0.0) b = (b * gff) / ip;
        (v ? p.bv : p.bu)[tt] = b;
    } else if (t < nu + nv + nc) {
        int tc = t - nu - nv;
        int i = tc % (nx - 1), j = tc / (nx - 1);
        // the four cells' h and eta_av loaded with the corner's mask
        const int k00 = i + nx * j, k10 = k00 + 1, k01 = k00 + nx, k11 = k01 + 1;
        const double h00 = p.h[k00], h10 = p.h[k10], h01 = p.h[k01], h11 = p.h[k11];
        const double e00 = p.etaav[k00], e10 = p.etaav[k10], e01 = p.etaav[k01], e11 = p.etaav[k11];
        if (!p.cm[tc]) {
            p.heta_c[tc] = 0.0;
            return;
        }
        // gh.cent_xy (HGrid.jl:203) of h*eta_av in the reference's factor order
        double a00 = h00 * e00;
        double a10 = h10 * e10;
        double a01 = h01 * e01;
        double a11 = h11 * e11;
        double out;
        if (p.yfirst_h2c) {
            double t0 = 0.5 * a00 + 0.5 * a01;
            double t1 = 0.5 * a10 + 0.5 * a11;
            out = 0.5 * t0 + 0.5 * t1;
        } else {
            double t0 = 0.5 * a00 + 0.5 * a10;
            double t1 = 0.5 * a01 + 0.5 * a11;
            out = 0.5 * t0 + 0.5 * t1;
        }
        p.heta_c[tc] = out;
    }
}

// ===========================================================================
// Reverse of one Picard iteration, for the adjoint basal-drag gradient (no
// reference counterpart).  Iteration k is inner_update! from the state
// (u_{k-1}, eta, beta, f2 of checkpoint k-1), then u_k = A_k^{-1} b.  With
// lambda_k = A_k^{-1} ubar_k, the coefficient adjoints are
// cbar = -d(lambda^T A(c) u_k)/dc, and the chain rule runs back through
// update_rheological_operators!, update_βeff_on_uv_grids!, update_βeff!,
// the quadratures, the viscosity passes, update_basal_drag!, update_β!,
// update_bed_speed!, update_av_speed! and update_shelf_strain_rate!
// (update_velocities.jl:202-383).  k_adj_h does the per-cell part, k_adj_uv
// the gathers onto u_{k-1} and v_{k-1}.

struct AdjArgs {
    InnerArgs p;                               // iteration k recomputed; p.U, p.V = u_{k-1}
    const double *uk, *vk;                     // u_k
    const double *lu, *lv;                     // lambda_k
    const double *eta_old, *beta_old, *f2_old; // checkpoint k-1
    double *etab, *betab, *f2b;                // state adjoints: leaving k on entry, entering k on exit
    double *wcb;                               // += dJ/d(weertman_c on h)
    double *cub, *cvb, *uxb, *vyb, *shb;       // per h cell, for k_adj_uv
    double *ubar, *vbar;                       // out: adjoint of u_{k-1} at inner points
};

// (du/dy + dv/dx) of (U, V) at corner (a, b), unmasked as op_fun! reads it
__device__ __forceinline__ double corner_rate(const GridDims &g, const double *U, const double *V, double rdx,
                                              double rdy, int a, int b)
{
    const int nxu = g.nxu, nx = g.nx;
    return (rdy * U[(a + 1) + nxu * (b + 1)] - rdy * U[(a + 1) + nxu * b]) +
           (rdx * V[(a + 1) + nx * (b + 1)] - rdx * V[a + nx * (b + 1)]);
}

__global__ void k_adj_h(GridDims g, AdjArgs a)
{
    const int nh = g.nx * g.ny;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh) return;
    const InnerArgs &p = a.p;
    const int nx = g.nx, ny = g.ny, nxu = g.nxu, nxc = g.nx - 1, nyc = g.ny - 1;
    const int i = k % nx, j = k / nx;
    const bool hm = p.hm[k];
    const double rdx = p.rdx, rdy = p.rdy;
    // update_rheological_operators!: h eta_av on h, and cent onto c
    double hqb = 0.0;
    if (hm) {
        const double dudx = rdx * a.uk[(i + 1) + nxu * j] - rdx * a.uk[i + nxu * j];
        const double dvdy = rdy * a.vk[i + nx * (j + 1)] - rdy * a.vk[i + nx * j];
        const double dlx = rdx * a.lu[(i + 1) + nxu * j] - rdx * a.lu[i + nxu * j];
        const double dly = rdy * a.lv[i + nx * (j + 1)] - rdy * a.lv[i + nx * j];
        hqb = 2.0 * (2.0 * dudx + dvdy) * dlx + 2.0 * (2.0 * dvdy + dudx) * dly;
    }
    for (int b = j - 1; b <= j; ++b)
        for (int c = i - 1; c <= i; ++c)
            if (c >= 0 && c < nxc && b >= 0 && b < nyc && p.cm[c + nxc * b])
                hqb += 0.25 * (corner_rate(g, a.uk, a.vk, rdx, rdy, c, b) *
                               corner_rate(g, a.lu, a.lv, rdx, rdy, c, b));
    const double etaavb = p.h[k] * hqb;
    // update_βeff_on_uv_grids!
    double beffb = 0.0;
    if (hm) {
        for (int c = i; c <= i + 1; ++c) {
            const int ku = c + nxu * j;
            if (!p.um[ku]) continue;
            double den = 0.0, ng = 0.0;
            if (c >= 1) {
                den = den + 0.5 * (double)p.hm[(c - 1) + nx * j];
                ng = ng + 0.5 * (p.hm[(c - 1) + nx * j] ? p.gfh[(c - 1) + nx * j] : 0.0);
            }
            if (c <= nx - 1) {
                den = den + 0.5 * (double)p.hm[c + nx * j];
                ng = ng + 0.5 * (p.hm[c + nx * j] ? p.gfh[c + nx * j] : 0.0);
            }
            const double ip = ng / den;
            beffb += (a.lu[ku] * a.uk[ku]) * (0.5 / den) * (ip > 0.0 ? p.gfu[ku] / ip : 1.0);
        }
        for (int b = j; b <= j + 1; ++b) {
            const int kv = i + nx * b;
            if (!p.vm[kv]) continue;
            double den = 0.0, ng = 0.0;
            if (b >= 1) {
                den = den + 0.5 * (double)p.hm[i + nx * (b - 1)];
                ng = ng + 0.5 * (p.hm[i + nx * (b - 1)] ? p.gfh[i + nx * (b - 1)] : 0.0);
            }
            if (b <= ny - 1) {
                den = den + 0.5 * (double)p.hm[i + nx * b];
                ng = ng + 0.5 * (p.hm[i + nx * b] ? p.gfh[i + nx * b] : 0.0);
            }
            const double ip = ng / den;
            beffb += (a.lv[kv] * a.vk[kv]) * (0.5 / den) * (ip > 0.0 ? p.gfv[kv] / ip : 1.0);
        }
    }
    // update_βeff!: beta_eff = beta / (1 + f2 beta)
    const double beta = p.beta[k], f2 = p.f2[k];
    const double D = 1.0 + f2 * beta;
    double bb = a.betab[k] + beffb / (D * D);
    const double fb = a.f2b[k] - beffb * (beta * beta) / (D * D);
    const double sr = p.sr[k], tau = p.tau[k], bs = p.bs[k];
    double taub = 0.0, srb = 0.0;
    if (hm) {
        for (int s = 0; s < p.ns; ++s) {
            const int k3 = k + nh * s;
            const double z = p.zeta[s], w = p.qw[s], eta = p.eta[k3], gb = p.glen_b[k3];
            // update_av_viscosity!, update_quadrature_falpha!
            double eb = a.etab[k3] + w * etaavb - fb * ((w * p.h[k]) * (z * z)) / (eta * eta);
            // inner_update_viscosity!: the passes from eta_old, then back
            double e[9];
            e[0] = a.eta_old[k3];
            for (int q = 0; q < p.niter_visc; ++q) {
                const double qq = (tau * z) / e[q];
                const double arg = sqrt((sr * sr + 0.25 * (qq * qq)) + p.greg2);
                e[q + 1] = (0.5 * gb) * det_pow(arg, p.gexp);
            }
            for (int q = p.niter_visc - 1; q >= 0; --q) {
                const double qq = (tau * z) / e[q];
                const double arg = sqrt((sr * sr + 0.25 * (qq * qq)) + p.greg2);
                const double gg = eb * e[q + 1] * p.gexp / arg;
                srb += gg * sr / arg;
                const double qb = gg * 0.25 * qq / arg;
                taub += qb * z / e[q];
                eb = -qb * qq / e[q];
            }
            a.etab[k3] = eb;
        }
    }
    // update_basal_drag!: tau = beta bs; update_β!: beta = wc (bs^2 + reg^2)^(wexp/2)
    double bsb = taub * beta;
    bb += taub * bs;
    const double S = bs * bs + p.wreg2;
    a.wcb[k] += bb * det_pow(sqrt(S), p.wexp);
    bsb += bb * beta * p.wexp * bs / S;
    // update_bed_speed!: bs = avs / (1 + f2_old beta_old)
    const double D0 = 1.0 + a.f2_old[k] * a.beta_old[k];
    const double avsb = bsb / D0, avs = p.avs[k];
    a.f2b[k] = -bsb * avs * a.beta_old[k] / (D0 * D0);
    a.betab[k] = -bsb * avs * a.f2_old[k] / (D0 * D0);
    // update_av_speed! and update_shelf_strain_rate! of u_{k-1}
    double cub = 0.0, cvb = 0.0, uxb = 0.0, vyb = 0.0, shb = 0.0;
    if (hm) {
        const double u0 = p.um[i + nxu * j] ? p.U[i + nxu * j] : 0.0;
        const double u1 = p.um[(i + 1) + nxu * j] ? p.U[(i + 1) + nxu * j] : 0.0;
        const double v0 = p.vm[i + nx * j] ? p.V[i + nx * j] : 0.0;
        const double v1 = p.vm[i + nx * (j + 1)] ? p.V[i + nx * (j + 1)] : 0.0;
        if (avs > 0.0) {
            cub = avsb * (0.5 * u0 + 0.5 * u1) / avs;
            cvb = avsb * (0.5 * v0 + 0.5 * v1) / avs;
        }
        if (sr > 0.0) {
            const double ux = rdx * u1 - rdx * u0, vy = rdy * v1 - rdy * v0;
            const double sh = cent_c_to_h(g, p, i, j);
            uxb = srb * (2.0 * ux + vy) / (2.0 * sr);
            vyb = srb * (2.0 * vy + ux) / (2.0 * sr);
            shb = srb * 0.25 * sh / sr;
        }
    }
    a.cub[k] = cub;
    a.cvb[k] = cvb;
    a.uxb[k] = uxb;
    a.vyb[k] = vyb;
    a.shb[k] = shb;
}

// adjoint of the corner shear at c (a, b): cent (c -> h) transposed
__device__ __forceinline__ double adj_corner(const GridDims &g, const AdjArgs &a, int c, int b)
{
    const int nx = g.nx;
    if (!a.p.cm[c + (nx - 1) * b]) return 0.0;
    return 0.25 * (a.shb[c + nx * b] + a.shb[(c + 1) + nx * b] + a.shb[c + nx * (b + 1)] +
                   a.shb[(c + 1) + nx * (b + 1)]);
}

__global__ void k_adj_uv(GridDims g, AdjArgs a)
{
    const int nu = g.nxu * g.nyu, nv = g.nxv * g.nyv;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.nx, ny = g.ny;
    const double rdx = a.p.rdx, rdy = a.p.rdy;
    if (t < nu) {
        const int i = t % g.nxu, j = t / g.nxu;
        double s = 0.0;
        if (a.p.um[t]) {
            if (i >= 1) s += 0.5 * a.cub[(i - 1) + nx * j] + rdx * a.uxb[(i - 1) + nx * j];
            if (i <= nx - 1) s += 0.5 * a.cub[i + nx * j] - rdx * a.uxb[i + nx * j];
            if (i >= 1 && i <= nx - 1) {
                if (j >= 1) s += rdy * adj_corner(g, a, i - 1, j - 1);
                if (j <= ny - 2) s -= rdy * adj_corner(g, a, i - 1, j);
            }
        }
        a.ubar[t] = s;
    } else if (t < nu + nv) {
        const int tv = t - nu;
        const int i = tv % g.nxv, j = tv / g.nxv;
        double s = 0.0;
        if (a.p.vm[tv]) {
            if (j >= 1) s += 0.5 * a.cvb[i + nx * (j - 1)] + rdy * a.vyb[i + nx * (j - 1)];
            if (j <= ny - 1) s += 0.5 * a.cvb[i + nx * j] - rdy * a.vyb[i + nx * j];
            if (j >= 1 && j <= ny - 1) {
                if (i >= 1) s += rdx * adj_corner(g, a, i - 1, j - 1);
                if (i <= nx - 2) s -= rdx * adj_corner(g, a, i, j - 1);
            }
        }
        a.vbar[tv] = s;
    }
}

// grad = wc * wcb (dJ/d log C through weertman_c = C * grounded fraction)
__global__ void k_adj_grad(int nh, const double *wc, const double *wcb, double *grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nh) grad[k] = wc[k] * wcb[k];
}

// ===========================================================================
// get_rhs (update_velocities.jl:106-182): f1 + f2 + f3 with f3 = -A(u_fixed)
// passed in as the operator applied to the fixed values.
struct RhsArgs {
    const uint8_t *hm, *ui, *vi;
    const double *s, *dsdh, *acc, *mel;
    const double *hu, *su, *hv, *sv;
    const double *f3u, *f3v;   // A applied to the fixed velocities
    double *bu, *bv;
    double rdx, rdy, g, ri, rw, sl, tdt;
};

__global__ void k_rhs(GridDims g, RhsArgs a)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = g.nxu * g.nyu, nv = g.nxv * g.nyv;
    const int nx = g.nx, ny = g.ny;
    const double rg = a.ri * a.g, hri = 0.5 * a.ri, hrw = 0.5 * a.rw;
    if (t >= nu + nv) return;
    // a face between h cells ka (left / below) and kb: every load issued
    // before the first is used (ld_or), then the reference's f1, f2, f3 terms
    const bool v = t >= nu;
    const int tt = v ? t - nu : t;
    const int i = v ? tt % g.nxv : tt % g.nxu, j = v ? tt / g.nxv : tt / g.nxu;
    const bool oka = v ? j > 0 : i > 0, okb = v ? j < ny : i < nx;
    const int ka = v ? i + nx * (j - 1) : (i - 1) + nx * j, kb = i + nx * j;
    const bool inner = (v ? a.vi : a.ui)[tt];
    const uint8_t ha = ld_or(oka, a.hm, ka, (uint8_t)0), hb = ld_or(okb, a.hm, kb, (uint8_t)0);
    const double sa = ld_or(oka, a.s, ka, 0.0), sb = ld_or(okb, a.s, kb, 0.0);
    const double da = ld_or(oka, a.dsdh, ka, 0.0), db = ld_or(okb, a.dsdh, kb, 0.0);
    const double aa = ld_or(oka, a.acc, ka, 0.0), ab = ld_or(okb, a.acc, kb, 0.0);
    const double ma = ld_or(oka, a.mel, ka, 0.0), mb = ld_or(okb, a.mel, kb, 0.0);
    const double h = (v ? a.hv : a.hu)[tt], s = (v ? a.sv : a.su)[tt], f3 = (v ? a.f3v : a.f3u)[tt];
    double *out = v ? a.bv : a.bu;
    if (!inner) {
        out[tt] = 0.0;
        return;
    }
    const double rd = v ? a.rdy : a.rdx;
    // the surface adjusted for the implicit step, s + theta dt dsdh (a - m),
    // on the two cells (0 off the mask)
    double a1 = okb ? (hb ? sb + (a.tdt * db) * (ab - mb) : 0.0) : 0.0;
    double a0 = oka ? (ha ? sa + (a.tdt * da) * (aa - ma) : 0.0) : 0.0;
    double ts = rd * a1 - rd * a0;
    double f1 = (rg * h) * ts;
    double m1 = okb ? (hb ? 1.0 : 0.0) : 0.0;
    double m0 = oka ? (ha ? 1.0 : 0.0) : 0.0;
    double tm = rd * m1 - rd * m0;
    double d = h - (s - a.sl);
    if (!(d > 0.0)) d = 0.0;
    double X = (hri * (h * h) - hrw * (d * d)) - (a.ri * h) * s;
    double f2 = (tm * a.g) * X;
    out[tt] = (f1 + f2) + (-f3);
}

// ===========================================================================
// Operator application, residual, diagonal, colour sweeps.

// diag(A) at inner points: the stencil evaluated on a unit vector, which is
// what get_op_diag's colour probes (update_preconditioners.jl:92-109) see.
template <typename T>
__global__ void k_diag(GridDims g, OpCoef<T> c, T *du, T *dv)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = g.nxu * g.nyu, nv = g.nxv * g.nyv;
    if (t < nu) {
        int i = t % g.nxu, j = t / g.nxu;
        DeltaIn<T> in{i, j, 0};
        du[t] = c.ui[t] ? op_u(c, in, i, j) : T(0);
    } else if (t < nu + nv) {
        int tv = t - nu;
        int i = tv % g.nxv, j = tv / g.nxv;
        DeltaIn<T> in{i, j, 1};
        dv[tv] = c.vi[tv] ? op_v(c, in, i, j) : T(0);
    }
}

// ===========================================================================
// Haar transforms on 16x16 tiles in LDS.  dir: 0 forward, 1 inverse,
// 2 transposed inverse.  y factor first, then x (the order KroneckerProduct
// mul! takes for square factors); each output sums its terms in increasing
// source index.

// One composed row of a table-driven transform: sum over the row's n terms
// of coef * src[col * stride], in the table's order.  The row's whole table
// entries are loaded first (MAX_HAAR_NNZ of each; those past n are not used),
// so that a term waits on its LDS read only, not on two table loads.
template <typename T>
__device__ __forceinline__ T haar_table_row(const HaarTable *tab, int dir, int m, int r, const T *src, int stride)
{
    const int n = tab->nnz[dir][m][r];
    float cf[MAX_HAAR_NNZ];
    int cl[MAX_HAAR_NNZ];
#pragma unroll
    for (int q = 0; q < MAX_HAAR_NNZ; ++q) {
        cf[q] = tab->coef[dir][m][r][q];
        cl[q] = tab->col[dir][m][r][q];
    }
    T acc = T(0);
#pragma unroll
    for (int q = 0; q < MAX_HAAR_NNZ; ++q)
        if (q < n) acc = acc + (T)cf[q] * src[cl[q] * stride];
    return acc;
}

template <typename T>
__device__ __forceinline__ void haar_tile(const HaarTable *tab, int dir, T (*tile)[TILE + 1], T (*tmp)[TILE + 1],
                                          int i0, int j0, int nxw, int nyw)
{
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int B = tab->B;
    const int i = i0 + lx, j = j0 + ly;
    // y pass
    if (i < nxw && j < nyw) {
        int jb = ly - (ly % B);
        int m = nyw - (j0 + jb);
        m = m < B ? m : B;
        int r = ly - jb;
        tmp[ly][lx] = haar_table_row<T>(tab, dir, m, r, &tile[jb][lx], TILE + 1);
    }
    __syncthreads();
    // x pass, with the transposed element mapping of haar_pm_tile2m (the
    // lanes of a wave share the column, so rows of one length)
    const int lyx = threadIdx.x & 15, lxx = threadIdx.x >> 4;
    if (i0 + lxx < nxw && j0 + lyx < nyw) {
        int ib = lxx - (lxx % B);
        int m = nxw - (i0 + ib);
        m = m < B ? m : B;
        int r = lxx - ib;
        tile[lyx][lxx] = haar_table_row<T>(tab, dir, m, r, &tmp[lyx][ib], 1);
    }
    __syncthreads();
}

// update_wavelets! (src/Wavelets/update_wavelets.jl:6-32): dwt of the masked
// velocities, threshold mask, zero warm start.
__global__ void k_tile_wavelets(GridDims g, const HaarTable *tab, const double *U, const double *V,
                                const uint8_t *um, const uint8_t *vm, double *Wu, double *Wv,
                                uint8_t *mwu, uint8_t *mwv, double *cu, double *cv, double thr)
{
    __shared__ double tile[TILE][TILE + 1];
    __shared__ double tmp[TILE][TILE + 1];
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int i = i0 + lx, j = j0 + ly;
    if (bx < g.nbxu && by < g.nbyu) {
        int k = i + g.nxu * j;
        bool ok = i < g.nxu && j < g.nyu;
        const double x = ld_or(ok, U, k, 0.0);
        tile[ly][lx] = (ld_or(ok, um, k, (uint8_t)0) != 0) ? x : 0.0;
        __syncthreads();
        haar_tile<double>(tab, 0, tile, tmp, i0, j0, g.nxu, g.nyu);
        if (ok) {
            double w = tile[ly][lx];
            Wu[k] = w;
            mwu[k] = fabs(w) >= thr ? 1 : 0;
            cu[k] = 0.0;
        }
        __syncthreads();
    }
    if (bx < g.nbxv && by < g.nbyv) {
        int k = i + g.nxv * j;
        bool ok = i < g.nxv && j < g.nyv;
        const double x = ld_or(ok, V, k, 0.0);
        tile[ly][lx] = (ld_or(ok, vm, k, (uint8_t)0) != 0) ? x : 0.0;
        __syncthreads();
        haar_tile<double>(tab, 0, tile, tmp, i0, j0, g.nxv, g.nyv);
        if (ok) {
            double w = tile[ly][lx];
            Wv[k] = w;
            mwv[k] = fabs(w) >= thr ? 1 : 0;
            cv[k] = 0.0;
        }
        __syncthreads();
    }
}

// ===========================================================================
// Haar transforms with +-1 coefficients (the inverse and its transpose): each
// composed row is a pair of 16-bit masks, bit c of `nz` marking a term in
// column c and bit c of `neg` its sign, one table per `levels` (0..3).
// Terms are summed in increasing column order, as the oracle sums them.
struct HaarMasks {
    uint16_t nz[2][17][16];
    uint16_t neg[2][17][16];
};
__constant__ HaarMasks c_haar[4];

template <typename T>
__device__ __forceinline__ T haar_pm_row(uint32_t nz, uint32_t neg, const T *src, int stride)
{
    T acc = T(0);
    while (nz) {
        const int c = __builtin_ctz(nz);
        const T v = src[c * stride];
        acc = ((neg >> c) & 1u) ? acc - v : acc + v;
        nz &= nz - 1u;
    }
    return acc;
}

// haar_pm_row with the B = 2 << levels entries of the row's block loaded
// first, then added in increasing column order through selects: the same
// additions as haar_pm_row, so the same bits, with one LDS wait instead of
// one per term.  The block lies inside the 16 x 16 tile, so every load is in
// bounds; entries outside the row's terms are loaded and discarded.  Shorter
// where a tile's latency chain is the launch's cost (grids of few tiles);
// the extra VALU work costs more than it saves when many tiles share a CU.
template <typename T, int B>
__device__ __forceinline__ T haar_pm_row_ld(uint32_t nz, uint32_t neg, const T *src, int stride)
{
    T v[B];
#pragma unroll
    for (int c = 0; c < B; ++c) v[c] = src[c * stride];
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < B; ++c) {
        const T t = ((neg >> c) & 1u) ? acc - v[c] : acc + v[c];
        acc = ((nz >> c) & 1u) ? t : acc;
    }
    return acc;
}

template <typename T, bool LD>
__device__ __forceinline__ T haar_pm_row_sel(int levels, uint32_t nz, uint32_t neg, const T *src, int stride)
{
    if constexpr (!LD) {
        return haar_pm_row<T>(nz, neg, src, stride);
    } else {
        switch (levels) {
        case 0:
            return haar_pm_row_ld<T, 2>(nz, neg, src, stride);
        case 1:
            return haar_pm_row_ld<T, 4>(nz, neg, src, stride);
        case 2:
            return haar_pm_row_ld<T, 8>(nz, neg, src, stride);
        default:
            return haar_pm_row_ld<T, 16>(nz, neg, src, stride);
        }
    }
}

// The row masks one thread reads in the two passes of a transform of the u
// and the v tile of block (i0, j0), all loaded at once (so that they can be
// loaded with a kernel's other inputs instead of one round trip per pass).
// The y pass computes element (ly, lx) = (tid >> 4, tid & 15), the x pass
// element (tid & 15, tid >> 4): the lanes of a wave then share the row of
// their y pass and the column of their x pass, and with it the number of
// terms, so the long rows (16 terms in the transposed inverse at 3 levels)
// fall in one wave of four in both passes.
struct HaarRowMasks {
    uint32_t yz_u, yn_u, yz_v, yn_v;  // y pass: terms and signs, u and v tile
    uint32_t xz_u, xn_u, xz_v, xn_v;  // x pass
};

// p[k] where ok, else 0, with the load issued unconditionally (index 0
// where !ok).
__device__ __forceinline__ uint32_t haar_mask_or0(bool ok, const uint16_t *p, int k)
{
    const uint32_t v = p[ok ? k : 0];
    return ok ? v : 0u;
}

__device__ __forceinline__ HaarRowMasks haar_row_masks(int levels, int dir, int i0, int j0, const GridDims &g,
                                                       bool has_u, bool has_v)
{
    // the y pass's element (ly, lx)
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int B = 2 << levels;
    const int i = i0 + lx, j = j0 + ly;
    const int jb = ly & ~(B - 1), ry = ly - jb;
    const bool ou = has_u && i < g.nxu && j < g.nyu, ov = has_v && i < g.nxv && j < g.nyv;
    const uint16_t *nz = &c_haar[levels].nz[dir][0][0], *ng = &c_haar[levels].neg[dir][0][0];
    const int yu = min(g.nyu - (j0 + jb), B) * 16 + ry, yv = min(g.nyv - (j0 + jb), B) * 16 + ry;
    // the x pass's element (lyx, lxx)
    const int lyx = threadIdx.x & 15, lxx = threadIdx.x >> 4;
    const int ix = i0 + lxx, jx = j0 + lyx, ibx = lxx & ~(B - 1), rxx = lxx - ibx;
    const bool oux = has_u && ix < g.nxu && jx < g.nyu, ovx = has_v && ix < g.nxv && jx < g.nyv;
    const int xux = min(g.nxu - (i0 + ibx), B) * 16 + rxx, xvx = min(g.nxv - (i0 + ibx), B) * 16 + rxx;
    HaarRowMasks m;
    m.yz_u = haar_mask_or0(ou, nz, yu);
    m.yn_u = haar_mask_or0(ou, ng, yu);
    m.yz_v = haar_mask_or0(ov, nz, yv);
    m.yn_v = haar_mask_or0(ov, ng, yv);
    m.xz_u = haar_mask_or0(oux, nz, xux);
    m.xn_u = haar_mask_or0(oux, ng, xux);
    m.xz_v = haar_mask_or0(ovx, nz, xvx);
    m.xn_v = haar_mask_or0(ovx, ng, xvx);
    return m;
}

// The same transform on the u tile and the v tile of one block at once, with
// the row masks given (haar_row_masks; LD: the rows as haar_pm_row_ld).
template <typename T, bool LD = false>
__device__ __forceinline__ void haar_pm_tile2m(int levels, const HaarRowMasks &hm, T (*tu)[TILE + 1],
                                               T (*mu)[TILE + 1], T (*tv)[TILE + 1], T (*mv)[TILE + 1], int i0,
                                               int j0, const GridDims &g, bool has_u, bool has_v)
{
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int B = 2 << levels;
    const int i = i0 + lx, j = j0 + ly;
    const int jb = ly & ~(B - 1);
    if (has_u && i < g.nxu && j < g.nyu)
        mu[ly][lx] = haar_pm_row_sel<T, LD>(levels, hm.yz_u, hm.yn_u, &tu[jb][lx], TILE + 1);
    if (has_v && i < g.nxv && j < g.nyv)
        mv[ly][lx] = haar_pm_row_sel<T, LD>(levels, hm.yz_v, hm.yn_v, &tv[jb][lx], TILE + 1);
    __syncthreads();
    // the x pass with the transposed element mapping (haar_row_masks)
    const int lyx = threadIdx.x & 15, lxx = threadIdx.x >> 4;
    const int ix = i0 + lxx, jx = j0 + lyx, ibx = lxx & ~(B - 1);
    if (has_u && ix < g.nxu && jx < g.nyu)
        tu[lyx][lxx] = haar_pm_row_sel<T, LD>(levels, hm.xz_u, hm.xn_u, &mu[lyx][ibx], 1);
    if (has_v && ix < g.nxv && jx < g.nyv)
        tv[lyx][lxx] = haar_pm_row_sel<T, LD>(levels, hm.xz_v, hm.xn_v, &mv[lyx][ibx], 1);
    __syncthreads();
}

// haar_pm_tile2m with the masks loaded here.
template <typename T, bool LD = false>
__device__ __forceinline__ void haar_pm_tile2(int levels, int dir, T (*tu)[TILE + 1], T (*mu)[TILE + 1],
                                              T (*tv)[TILE + 1], T (*mv)[TILE + 1], int i0, int j0, const GridDims &g,
                                              bool has_u, bool has_v)
{
    const HaarRowMasks hm = haar_row_masks(levels, dir, i0, j0, g, has_u, has_v);
    haar_pm_tile2m<T, LD>(levels, hm, tu, mu, tv, mv, i0, j0, g, has_u, has_v);
}

// ===========================================================================
// The operator on one 16x16 tile of u and v outputs, every intermediate
// computed once per tile in LDS.  U and V are staged with a one-point halo
// (18 x 18, s.U[y][x] = U(i0-1+x, j0-1+y)); r_xx, r_yy and the implicit
// term on the 17 x 17 h cells and r_xy on the 17 x 17 c corners starting one
// point before the tile.  Arithmetic is that of op_u / op_v (same expressions,
// same order, absent terms zero), so results are bit-identical.
template <typename T>
struct OpTile {
    T U[18][19];
    T V[18][19];
    T rxx[17][18];
    T ryy[17][18];
    T ext[17][18];
    T rxy[17][18];
};

// Coefficients of one h cell and its c corner (the cell (i0-1+x, j0-1+y) of
// the 17 x 17 window, q = x + 17 y), loaded before they are needed so that a
// tile costs one round trip to memory.
template <typename T>
struct CellCoef {
    T heta, dimp, hu0, hu1, hv0, hv1, hetac;
    uint8_t hm, um0, um1, vm0, vm1, cm;
    bool cell, corner;
};

// Coefficients of h cell (a, b) and its c corner; in_window = false leaves
// both absent.
template <typename T>
__device__ __forceinline__ void cell_prefetch_at(const GridDims &g, const OpCoef<T> &c, int a, int b,
                                                 bool in_window, CellCoef<T> &cc)
{
    const int nx = g.nx, ny = g.ny, nxu = g.nxu;
    cc.cell = in_window && a >= 0 && a < nx && b >= 0 && b < ny;
    cc.corner = in_window && a >= 0 && a < nx - 1 && b >= 0 && b < ny - 1;
    // loads issued unconditionally (ld_or); absent entries are never read
    {
        const bool ok = cc.cell;
        const int k = a + nx * b, ku0 = a + nxu * b, kv0 = a + nx * b, kv1 = kv0 + nx;
        cc.hm = ld_or(ok, c.hm, k, (uint8_t)0);
        cc.heta = ld_or(ok, c.heta_h, k, T(0));
        cc.dimp = ld_or(ok, c.dimp, k, T(0));
        cc.um0 = ld_or(ok, c.um, ku0, (uint8_t)0);
        cc.um1 = ld_or(ok, c.um, ku0 + 1, (uint8_t)0);
        cc.hu0 = ld_or(ok, c.hu, ku0, T(0));
        cc.hu1 = ld_or(ok, c.hu, ku0 + 1, T(0));
        cc.vm0 = ld_or(ok, c.vm, kv0, (uint8_t)0);
        cc.vm1 = ld_or(ok, c.vm, kv1, (uint8_t)0);
        cc.hv0 = ld_or(ok, c.hv, kv0, T(0));
        cc.hv1 = ld_or(ok, c.hv, kv1, T(0));
    }
    {
        const int kc = a + (nx - 1) * b;
        cc.cm = ld_or(cc.corner, c.cm, kc, (uint8_t)0);
        cc.hetac = ld_or(cc.corner, c.heta_c, kc, T(0));
    }
}

template <typename T>
__device__ __forceinline__ void cell_prefetch(const GridDims &g, const OpCoef<T> &c, int i0, int j0, int q,
                                              CellCoef<T> &cc)
{
    cell_prefetch_at(g, c, i0 - 1 + q % 17, j0 - 1 + q / 17, q < 289, cc);
}

// Intermediates of the cell / corner at (x, y) of a staged window whose
// points U[y][x] and cells share one origin (same expressions as op_rxx,
// op_ryy, op_extra, op_rxy).
template <typename T, class S>
__device__ __forceinline__ void cell_eval(const OpCoef<T> &c, const CellCoef<T> &cc, int x, int y, S &s)
{
    T rxx = T(0), ryy = T(0), ext = T(0), rxy = T(0);
    if (cc.cell) {
        const T u0 = s.U[y][x], u1 = s.U[y][x + 1], v0 = s.V[y][x], v1 = s.V[y + 1][x];
        if (cc.hm) {
            const T dudx = c.rdx * u1 - c.rdx * u0;
            const T dvdy = c.rdy * v1 - c.rdy * v0;
            rxx = T(2) * (cc.heta * (T(2) * dudx + dvdy));
            ryy = T(2) * (cc.heta * (T(2) * dvdy + dudx));
        }
        const T qx1 = cc.um1 ? cc.hu1 * u1 : T(0);
        const T qx0 = cc.um0 ? cc.hu0 * u0 : T(0);
        const T qy1 = cc.vm1 ? cc.hv1 * v1 : T(0);
        const T qy0 = cc.vm0 ? cc.hv0 * v0 : T(0);
        const T divq = (c.rdx * qx1 - c.rdx * qx0) + (c.rdy * qy1 - c.rdy * qy0);
        ext = cc.dimp * divq;
    }
    if (cc.corner && cc.cm) {
        const T dudy = c.rdy * s.U[y + 1][x + 1] - c.rdy * s.U[y][x + 1];
        const T dvdx = c.rdx * s.V[y + 1][x + 1] - c.rdx * s.V[y + 1][x];
        rxy = cc.hetac * (dudy + dvdx);
    }
    s.rxx[y][x] = rxx;
    s.ryy[y][x] = ryy;
    s.ext[y][x] = ext;
    s.rxy[y][x] = rxy;
}

template <typename T>
__device__ __forceinline__ void cell_compute(const OpCoef<T> &c, const CellCoef<T> &cc, int q, OpTile<T> &s)
{
    if (q >= 289) return;
    cell_eval(c, cc, q % 17, q / 17, s);
}

// Coefficients of the thread's own u and v output points.
template <typename T>
struct PointCoef {
    T bu, hu, bv, hv;
    bool oku, okv, ui, um, vi, vm, mwu, mwv;
    int ku, kv;
};

template <typename T>
__device__ __forceinline__ void point_prefetch(const GridDims &g, const OpCoef<T> &c, const uint8_t *mwu,
                                               const uint8_t *mwv, int i, int j, bool has_u, bool has_v,
                                               PointCoef<T> &pc)
{
    pc.oku = has_u && i < g.nxu && j < g.nyu;
    pc.okv = has_v && i < g.nxv && j < g.nyv;
    pc.ku = i + g.nxu * j;
    pc.kv = i + g.nxv * j;
    // loads issued unconditionally (ld_or)
    pc.ui = ld_or(pc.oku, c.ui, pc.ku, (uint8_t)0) != 0;
    pc.um = ld_or(pc.oku, c.um, pc.ku, (uint8_t)0) != 0;
    pc.bu = ld_or(pc.oku, c.bu, pc.ku, T(0));
    pc.hu = ld_or(pc.oku, c.hu, pc.ku, T(0));
    pc.mwu = ld_or(pc.oku, mwu, pc.ku, (uint8_t)0) != 0;
    pc.vi = ld_or(pc.okv, c.vi, pc.kv, (uint8_t)0) != 0;
    pc.vm = ld_or(pc.okv, c.vm, pc.kv, (uint8_t)0) != 0;
    pc.bv = ld_or(pc.okv, c.bv, pc.kv, T(0));
    pc.hv = ld_or(pc.okv, c.hv, pc.kv, T(0));
    pc.mwv = ld_or(pc.okv, mwv, pc.kv, (uint8_t)0) != 0;
}

// fx at the u point (i0 + x, j0 + y), whose value is s.U[y + 1][x + 1] and
// whose cell (i, j) is s.rxx[y + 1][x + 1]
template <typename T, class S>
__device__ __forceinline__ T optile_u(const OpCoef<T> &c, const S &s, const PointCoef<T> &pc, int x, int y)
{
    const T d_rxx_dx = c.rdx * s.rxx[y + 1][x + 1] - c.rdx * s.rxx[y + 1][x];
    const T d_rxy_dy = c.rdy * s.rxy[y + 1][x] - c.rdy * s.rxy[y][x];
    const T taubx = pc.um ? pc.bu * s.U[y + 1][x + 1] : T(0);
    const T d_extra_dx = c.rdx * s.ext[y + 1][x + 1] - c.rdx * s.ext[y + 1][x];
    return ((d_rxx_dx + d_rxy_dy) - taubx) - pc.hu * d_extra_dx;
}

// fy at the v point (i0 + x, j0 + y), with the same window convention
template <typename T, class S>
__device__ __forceinline__ T optile_v(const OpCoef<T> &c, const S &s, const PointCoef<T> &pc, int x, int y)
{
    const T d_ryy_dy = c.rdy * s.ryy[y + 1][x + 1] - c.rdy * s.ryy[y][x + 1];
    const T d_rxy_dx = c.rdx * s.rxy[y][x + 1] - c.rdx * s.rxy[y][x];
    const T tauby = pc.vm ? pc.bv * s.V[y + 1][x + 1] : T(0);
    const T d_extra_dy = c.rdy * s.ext[y + 1][x + 1] - c.rdy * s.ext[y][x + 1];
    return ((d_ryy_dy + d_rxy_dx) - tauby) - pc.hv * d_extra_dy;
}

// Stage one tile of a fine vector: each thread loads its (at most two)
// staging entries into registers; optile_store puts them in LDS.
template <typename T>
struct StageRegs {
    T u0, v0, u1, v1;
};

template <typename T>
__device__ __forceinline__ void optile_fetch(const GridDims &g, const T *U, const T *V, int i0, int j0,
                                             StageRegs<T> &r)
{
    const int q0 = threadIdx.x, q1 = threadIdx.x + 256;
    {
        const int x = q0 % 18, y = q0 / 18, i = i0 - 1 + x, j = j0 - 1 + y;
        r.u0 = ld_or(i >= 0 && i < g.nxu && j >= 0 && j < g.nyu, U, i + g.nxu * j, T(0));
        r.v0 = ld_or(i >= 0 && i < g.nxv && j >= 0 && j < g.nyv, V, i + g.nxv * j, T(0));
    }
    {
        const int x = q1 % 18, y = q1 / 18, i = i0 - 1 + x, j = j0 - 1 + y;
        const bool in = q1 < 324;
        r.u1 = ld_or(in && i >= 0 && i < g.nxu && j >= 0 && j < g.nyu, U, i + g.nxu * j, T(0));
        r.v1 = ld_or(in && i >= 0 && i < g.nxv && j >= 0 && j < g.nyv, V, i + g.nxv * j, T(0));
    }
}

template <typename T>
__device__ __forceinline__ void optile_store(const StageRegs<T> &r, OpTile<T> &s)
{
    const int q0 = threadIdx.x, q1 = threadIdx.x + 256;
    s.U[q0 / 18][q0 % 18] = r.u0;
    s.V[q0 / 18][q0 % 18] = r.v0;
    if (q1 < 324) {
        s.U[q1 / 18][q1 % 18] = r.u1;
        s.V[q1 / 18][q1 % 18] = r.v1;
    }
}

// The whole tile operator with prefetch: stage t, compute intermediates,
// outputs into tu / tv (zero off the inner masks).  Ends with a barrier.
template <typename T>
__device__ __forceinline__ void optile_apply(const GridDims &g, const OpCoef<T> &c, const T *tU, const T *tV,
                                             int i0, int j0, const PointCoef<T> &pc, OpTile<T> &s,
                                             T (*tu)[TILE + 1], T (*tv)[TILE + 1])
{
    StageRegs<T> sr;
    CellCoef<T> c0, c1;
    optile_fetch(g, tU, tV, i0, j0, sr);
    cell_prefetch(g, c, i0, j0, threadIdx.x, c0);
    cell_prefetch(g, c, i0, j0, threadIdx.x + 256, c1);
    optile_store(sr, s);
    __syncthreads();
    cell_compute(c, c0, threadIdx.x, s);
    cell_compute(c, c1, threadIdx.x + 256, s);
    __syncthreads();
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    tu[ly][lx] = (pc.oku && pc.ui) ? optile_u(c, s, pc, lx, ly) : T(0);
    tv[ly][lx] = (pc.okv && pc.vi) ? optile_v(c, s, pc, lx, ly) : T(0);
    __syncthreads();
}

// ===========================================================================
// Operator application and residual on 16 x 16 tiles of u and v points: the
// tile stages its inputs with a one-point halo and its coefficients once
// (OpTile), computes every intermediate once, and each thread evaluates its
// u and its v point.  The arithmetic is that of op_u / op_v, so the results
// equal the per-point forms bit for bit.  Launched as nblk workgroups
// (xcd_band_tile).

// Tile (bx, by) of this workgroup in a one-dimensional launch of nblk (a
// multiple of 8) workgroups over an nbx x nby tile grid.  Workgroups are
// dispatched round-robin over the 8 XCDs; workgroup w takes tile
// L = (w % 8) (nblk / 8) + w / 8 of a column-major walk (cols = true: each
// XCD takes a band of whole tile columns, so the halos that vertical and
// horizontal neighbours share mostly stay in its L2) or a row-major one (a
// band of whole tile rows).  Returns false for the padding workgroups.
__device__ __forceinline__ bool xcd_band_tile(int nbx, int nby, int nblk, bool cols, int &bx, int &by)
{
    const int w = blockIdx.x;
    const int L = (w & 7) * (nblk >> 3) + (w >> 3);
    bx = cols ? L / nby : L % nbx;
    by = cols ? L % nby : L / nbx;
    return L < nbx * nby;
}

// The coefficients a tile's own u and v points read (no wavelet masks).
template <typename T>
__device__ __forceinline__ void point_prefetch_op(const GridDims &g, const OpCoef<T> &c, int i, int j, bool has_u,
                                                  bool has_v, PointCoef<T> &pc)
{
    pc.oku = has_u && i < g.nxu && j < g.nyu;
    pc.okv = has_v && i < g.nxv && j < g.nyv;
    pc.ku = i + g.nxu * j;
    pc.kv = i + g.nxv * j;
    pc.mwu = pc.mwv = false;
    // loads issued unconditionally (ld_or)
    pc.ui = ld_or(pc.oku, c.ui, pc.ku, (uint8_t)0) != 0;
    pc.um = ld_or(pc.oku, c.um, pc.ku, (uint8_t)0) != 0;
    pc.bu = ld_or(pc.oku, c.bu, pc.ku, T(0));
    pc.hu = ld_or(pc.oku, c.hu, pc.ku, T(0));
    pc.vi = ld_or(pc.okv, c.vi, pc.kv, (uint8_t)0) != 0;
    pc.vm = ld_or(pc.okv, c.vm, pc.kv, (uint8_t)0) != 0;
    pc.bv = ld_or(pc.okv, c.bv, pc.kv, T(0));
    pc.hv = ld_or(pc.okv, c.hv, pc.kv, T(0));
}

// Stage the tile's window of x and its intermediates (ends with a barrier).
template <typename T>
__device__ __forceinline__ void optile_prepare(const GridDims &g, const OpCoef<T> &c, const T *xu, const T *xv,
                                               int i0, int j0, OpTile<T> &s)
{
    StageRegs<T> sr;
    CellCoef<T> c0, c1;
    optile_fetch(g, xu, xv, i0, j0, sr);
    cell_prefetch(g, c, i0, j0, threadIdx.x, c0);
    cell_prefetch(g, c, i0, j0, threadIdx.x + 256, c1);
    optile_store(sr, s);
    __syncthreads();
    cell_compute(c, c0, threadIdx.x, s);
    cell_compute(c, c1, threadIdx.x + 256, s);
    __syncthreads();
}

// y = A x at inner points, 0 elsewhere
template <typename T>
__global__ void __launch_bounds__(256) k_op(GridDims g, OpCoef<T> c, const T *xu, const T *xv, T *yu, T *yv,
                                            int nblk)
{
    __shared__ OpTile<T> s;
    int bx, by;
    if (!xcd_band_tile(g.nbx, g.nby, nblk, true, bx, by)) return;
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int i0 = bx * TILE, j0 = by * TILE, lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    PointCoef<T> pc;
    point_prefetch_op(g, c, i0 + lx, j0 + ly, has_u, has_v, pc);
    optile_prepare(g, c, xu, xv, i0, j0, s);
    if (pc.oku) yu[pc.ku] = pc.ui ? optile_u(c, s, pc, lx, ly) : T(0);
    if (pc.okv) yv[pc.kv] = pc.vi ? optile_v(c, s, pc, lx, ly) : T(0);
}

// r = b - A x at inner points (get_resid!, utilities.jl:549-552); per-tile
// partials of sum r^2 into Pr and, if Pb != nullptr, of sum b^2 into Pb (u
// tiles first, then v tiles); if fu != nullptr, r also into fu, fv at the
// inner points (the residual fields of set_residual!,
// update_velocities.jl:428-433; their other points are left as they are).
template <typename T>
__global__ void __launch_bounds__(256) k_tile_resid(GridDims g, OpCoef<T> c, const T *xu, const T *xv, const T *bu,
                                                    const T *bv, T *ru, T *rv, double *Pr, double *Pb, int nblk,
                                                    T *fu = nullptr, T *fv = nullptr)
{
    __shared__ OpTile<T> s;
    __shared__ double lds[8];
    int bx, by;
    if (!xcd_band_tile(g.nbx, g.nby, nblk, true, bx, by)) return;
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int i0 = bx * TILE, j0 = by * TILE, lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    PointCoef<T> pc;
    point_prefetch_op(g, c, i0 + lx, j0 + ly, has_u, has_v, pc);
    const T b_u = ld_or(pc.oku, bu, pc.ku, T(0)), b_v = ld_or(pc.okv, bv, pc.kv, T(0));
    optile_prepare(g, c, xu, xv, i0, j0, s);
    if (has_u) {
        double r2 = 0.0, b2 = 0.0;
        if (pc.oku) {
            const T r = pc.ui ? b_u - optile_u(c, s, pc, lx, ly) : T(0);
            ru[pc.ku] = r;
            if (fu && pc.ui) fu[pc.ku] = r;
            r2 = (double)r * (double)r;
            b2 = (double)b_u * (double)b_u;
        }
        tree256x2(r2, b2, lds);
        if (threadIdx.x == 0) {
            Pr[bx + g.nbxu * by] = r2;
            if (Pb) Pb[bx + g.nbxu * by] = b2;
        }
    }
    if (has_v) {
        double r2 = 0.0, b2 = 0.0;
        if (pc.okv) {
            const T r = pc.vi ? b_v - optile_v(c, s, pc, lx, ly) : T(0);
            rv[pc.kv] = r;
            if (fv && pc.vi) fv[pc.kv] = r;
            r2 = (double)r * (double)r;
            b2 = (double)b_v * (double)b_v;
        }
        tree256x2(r2, b2, lds);
        if (threadIdx.x == 0) {
            const int o = g.nbxu * g.nbyu + bx + g.nbxv * by;
            Pr[o] = r2;
            if (Pb) Pb[o] = b2;
        }
    }
}

// ===========================================================================
// One iteration of gauss_seidel_smoother! (update_preconditioners.jl:133-150)
// in one launch: the four colour sweeps x[c] += omega (b - A x)[c] / diag[c]
// in the order given, on the u and v points of one tile.  Colours 1
// and 2 are the u points with 1 + mod(i - j, 2) == colour, 3 and 4 the v
// points with 3 + mod(i - j, 2) (get_preconditioner,
// update_preconditioners.jl:7-43).  A point's update reads its neighbours
// within one point, so the tile stages a window with a halo of SM_H = 4
// points; sweep s updates its colour inside the window shrunk by s points,
// whose neighbours hold the values of sweep s - 1, and the fourth sweep
// covers the tile.  Halo points are updated by the neighbouring tiles with
// the same inputs and arithmetic, and only the tile's own points are
// written, into a second buffer (xu_out, xv_out): the other tiles read the
// inputs of the iteration from the first.  Every update has the arithmetic of
// op_u / op_v and of the one-colour sweep, so the result is bit-identical to
// four sweeps of one colour each.
constexpr int SM_H = 4;  // halo points

// A TX x TY tile: window of PX x PY points (the tile and its halo), of which
// the CX x CY cells starting one point in are computed.  Points at
// [py][px] = (i0 - SM_H - 1 + px, j0 - SM_H - 1 + py), px in 1..PX (index 0
// unused); cells and corners with the same origin.  Window position (x, y)
// of a point is px = x + 1, py = y + 1, as optile_u / optile_v expect.
template <int TX, int TY>
struct SmoothShape {
    static constexpr int PX = TX + 2 * SM_H, PY = TY + 2 * SM_H;
    static constexpr int CX = PX - 1, CY = PY - 1;
    static constexpr int SLOTS = (PX * PY + 255) / 256;   // window points per thread
    static constexpr int CSLOTS = (CX * CY + 255) / 256;  // cells per thread
};

template <typename T, int TX, int TY>
struct SmoothWin {
    using S = SmoothShape<TX, TY>;
    T U[S::PY + 1][S::PX + 2];
    T V[S::PY + 1][S::PX + 2];
    T rxx[S::CY + 1][S::CX + 2];
    T ryy[S::CY + 1][S::CX + 2];
    T ext[S::CY + 1][S::CX + 2];
    T rxy[S::CY + 1][S::CX + 2];
};

// The coefficients one window point's update reads.
template <typename T>
struct SmoothPoint {
    PointCoef<T> pc;
    T bu, du, bv, dv;  // right-hand side and diagonal at the u and the v point
};

// Tile shapes of the smoother (the autotuner's candidates): 16 x 16, 32 x 16
// and 16 x 32 points.  Any shape gives the same bits: a point's update does
// not depend on the tile it is computed in.
template <typename T, int TX, int TY>
__global__ void __launch_bounds__(256) k_smooth4(GridDims g, OpCoef<T> c, const T *xu, const T *xv, T *xu_out,
                                                 T *xv_out, const T *bu, const T *bv, const T *du, const T *dv,
                                                 T omega, int order, int ntx, int nty, int nblk, int cols)
{
    using S = SmoothShape<TX, TY>;
    __shared__ SmoothWin<T, TX, TY> s;
    int bx, by;
    if (!xcd_band_tile(ntx, nty, nblk, cols != 0, bx, by)) return;
    const int i0 = bx * TX, j0 = by * TY;
    const int ox = i0 - SM_H, oy = j0 - SM_H;  // window position 0
    // stage the window and load the coefficients the four sweeps read
    SmoothPoint<T> sp[S::SLOTS];
#pragma unroll
    for (int k = 0; k < S::SLOTS; ++k) {
        const int q = threadIdx.x + 256 * k;
        const int x = q % S::PX, y = q / S::PX, i = ox + x, j = oy + y;
        PointCoef<T> &pc = sp[k].pc;
        pc.oku = q < S::PX * S::PY && i >= 0 && j >= 0 && i < g.nxu && j < g.nyu;
        pc.okv = q < S::PX * S::PY && i >= 0 && j >= 0 && i < g.nxv && j < g.nyv;
        pc.ku = i + g.nxu * j;
        pc.kv = i + g.nxv * j;
        pc.ui = pc.um = pc.vi = pc.vm = pc.mwu = pc.mwv = false;
        pc.bu = pc.hu = pc.bv = pc.hv = T(0);
        sp[k].bu = sp[k].du = sp[k].bv = sp[k].dv = T(0);
        T u = T(0), v = T(0);
        if (pc.oku) {
            u = xu[pc.ku];
            pc.ui = c.ui[pc.ku];
            pc.um = c.um[pc.ku];
            pc.bu = c.bu[pc.ku];
            pc.hu = c.hu[pc.ku];
            sp[k].bu = bu[pc.ku];
            sp[k].du = du[pc.ku];
        }
        if (pc.okv) {
            v = xv[pc.kv];
            pc.vi = c.vi[pc.kv];
            pc.vm = c.vm[pc.kv];
            pc.bv = c.bv[pc.kv];
            pc.hv = c.hv[pc.kv];
            sp[k].bv = bv[pc.kv];
            sp[k].dv = dv[pc.kv];
        }
        if (q < S::PX * S::PY) {
            s.U[y + 1][x + 1] = u;
            s.V[y + 1][x + 1] = v;
        }
    }
    // cell (a, b) = (ox + q % CX, oy + q / CX) at [1 + q / CX][1 + q % CX]
    CellCoef<T> cc[S::CSLOTS];
#pragma unroll
    for (int k = 0; k < S::CSLOTS; ++k) {
        const int q = threadIdx.x + 256 * k;
        cell_prefetch_at(g, c, ox + q % S::CX, oy + q / S::CX, q < S::CX * S::CY, cc[k]);
    }
    for (int sw = 1; sw <= 4; ++sw) {
        const int colour = (order >> (4 * (sw - 1))) & 15;
        __syncthreads();  // the window holds the values after sweep sw - 1
        // sweep sw updates window positions [sw, P - sw), which read the
        // cells at indices [sw, P - sw]
#pragma unroll
        for (int k = 0; k < S::CSLOTS; ++k) {
            const int q = threadIdx.x + 256 * k;
            const int cx = 1 + q % S::CX, cy = 1 + q / S::CX;
            if (q < S::CX * S::CY && cx >= sw && cy >= sw && cx <= S::PX - sw && cy <= S::PY - sw)
                cell_eval(c, cc[k], cx, cy, s);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < S::SLOTS; ++k) {
            const int q = threadIdx.x + 256 * k;
            const int x = q % S::PX, y = q / S::PX;
            if (q >= S::PX * S::PY || x < sw || y < sw || x >= S::PX - sw || y >= S::PY - sw) continue;
            const PointCoef<T> &pc = sp[k].pc;
            const int i = ox + x, j = oy + y;
            if (colour <= 2) {
                if (pc.oku && pc.ui && ((i - j) & 1) == colour - 1) {
                    const T r = sp[k].bu - optile_u(c, s, pc, x, y);
                    s.U[y + 1][x + 1] = s.U[y + 1][x + 1] + (omega * r) / sp[k].du;
                }
            } else if (pc.okv && pc.vi && ((i - j) & 1) == colour - 3) {
                const T r = sp[k].bv - optile_v(c, s, pc, x, y);
                s.V[y + 1][x + 1] = s.V[y + 1][x + 1] + (omega * r) / sp[k].dv;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S::SLOTS; ++k) {
        const int q = threadIdx.x + 256 * k;
        const int x = q % S::PX, y = q / S::PX;
        if (q >= S::PX * S::PY || x < SM_H || y < SM_H || x >= SM_H + TX || y >= SM_H + TY) continue;
        if (sp[k].pc.oku) xu_out[sp[k].pc.ku] = s.U[y + 1][x + 1];
        if (sp[k].pc.okv) xv_out[sp[k].pc.kv] = s.V[y + 1][x + 1];
    }
}

// ===========================================================================
// The whole smoothing of a small grid in one workgroup: nsmooth iterations of
// gauss_seidel_smoother! (update_preconditioners.jl:133-150), each the four
// colour sweeps in `order`, with the grid staged in LDS once and block
// barriers between the sweeps, so that a V-cycle's smoothing is one launch
// instead of one per iteration.  The window is the grid with a border of one
// absent point: window position (x, y) is the point (x - 1, y - 1), its value
// at [y + 1][x + 1] and its cell at [y + 1][x + 1] as optile_u / optile_v
// expect.  A point's update is that of k_smooth4 (the arithmetic of op_u /
// op_v and of the one-colour sweep), so the bits are the same.  The result
// goes to xu_out, xv_out, which may be xu, xv: every read of x precedes the
// first barrier and every write follows the last.
constexpr int SMG_THREADS = 1024;
constexpr int SMG_MAX_SLOTS = 2;  // window points per thread

// Rows of a 2-D array in dynamic LDS.
template <typename T>
struct LdsRows {
    T *p;
    int ld;
    __device__ __forceinline__ T *operator[](int y) const { return p + y * ld; }
};

template <typename T>
struct SmoothGridWin {
    LdsRows<T> U, V, rxx, ryy, ext, rxy;
};

// Extents of the whole-grid window: PX x PY points and CX x CY cells, the
// cell q = (q % CX - 1, q / CX - 1) at [1 + q / CX][1 + q % CX]; U and V have
// PY + 1 rows of PX + 2, the cell arrays CY + 1 rows of CX + 2.
struct SmoothGridDims {
    int PX, PY, CX, CY;
    __host__ __device__ explicit SmoothGridDims(const GridDims &g)
        : PX((g.nxu > g.nxv ? g.nxu : g.nxv) + 2), PY((g.nyu > g.nyv ? g.nyu : g.nyv) + 2), CX(PX - 1), CY(PY - 1)
    {
    }
    __host__ __device__ int points() const { return PX * PY; }
    __host__ __device__ int cells() const { return CX * CY; }
    // LDS elements of the six arrays
    __host__ __device__ size_t elems() const
    {
        return 2 * (size_t)(PY + 1) * (PX + 2) + 4 * (size_t)(CY + 1) * (CX + 2);
    }
};

template <typename T, int SLOTS>
__global__ void __launch_bounds__(SMG_THREADS) k_smooth_grid(GridDims g, OpCoef<T> c, const T *xu, const T *xv,
                                                            T *xu_out, T *xv_out, const T *bu, const T *bv,
                                                            const T *du, const T *dv, T omega, int order, int nsmooth)
{
    extern __shared__ __align__(16) unsigned char smg_lds[];
    const SmoothGridDims d(g);
    const int PX = d.PX, PY = d.PY, CX = d.CX, CY = d.CY;
    SmoothGridWin<T> s;
    {
        T *p = reinterpret_cast<T *>(smg_lds);
        const int lp = PX + 2, lc = CX + 2;
        s.U = {p, lp};
        p += (size_t)(PY + 1) * lp;
        s.V = {p, lp};
        p += (size_t)(PY + 1) * lp;
        s.rxx = {p, lc};
        p += (size_t)(CY + 1) * lc;
        s.ryy = {p, lc};
        p += (size_t)(CY + 1) * lc;
        s.ext = {p, lc};
        p += (size_t)(CY + 1) * lc;
        s.rxy = {p, lc};
    }
    // stage the grid and load the coefficients every sweep reads
    SmoothPoint<T> sp[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int q = threadIdx.x + SMG_THREADS * k;
        const int x = q % PX, y = q / PX, i = x - 1, j = y - 1;
        PointCoef<T> &pc = sp[k].pc;
        pc.oku = q < PX * PY && i >= 0 && j >= 0 && i < g.nxu && j < g.nyu;
        pc.okv = q < PX * PY && i >= 0 && j >= 0 && i < g.nxv && j < g.nyv;
        pc.ku = i + g.nxu * j;
        pc.kv = i + g.nxv * j;
        pc.ui = pc.um = pc.vi = pc.vm = pc.mwu = pc.mwv = false;
        pc.bu = pc.hu = pc.bv = pc.hv = T(0);
        sp[k].bu = sp[k].du = sp[k].bv = sp[k].dv = T(0);
        T u = T(0), v = T(0);
        if (pc.oku) {
            u = xu[pc.ku];
            pc.ui = c.ui[pc.ku];
            pc.um = c.um[pc.ku];
            pc.bu = c.bu[pc.ku];
            pc.hu = c.hu[pc.ku];
            sp[k].bu = bu[pc.ku];
            sp[k].du = du[pc.ku];
        }
        if (pc.okv) {
            v = xv[pc.kv];
            pc.vi = c.vi[pc.kv];
            pc.vm = c.vm[pc.kv];
            pc.bv = c.bv[pc.kv];
            pc.hv = c.hv[pc.kv];
            sp[k].bv = bv[pc.kv];
            sp[k].dv = dv[pc.kv];
        }
        if (q < PX * PY) {
            s.U[y + 1][x + 1] = u;
            s.V[y + 1][x + 1] = v;
        }
    }
    CellCoef<T> cc[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int q = threadIdx.x + SMG_THREADS * k;
        cell_prefetch_at(g, c, q % CX - 1, q / CX - 1, q < CX * CY, cc[k]);
    }
    for (int it = 0; it < nsmooth; ++it) {
        for (int sw = 0; sw < 4; ++sw) {
            const int colour = (order >> (4 * sw)) & 15;
            __syncthreads();  // the window holds the values after the previous sweep
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                const int q = threadIdx.x + SMG_THREADS * k;
                if (q < CX * CY) cell_eval(c, cc[k], 1 + q % CX, 1 + q / CX, s);
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                const int q = threadIdx.x + SMG_THREADS * k;
                if (q >= PX * PY) continue;
                const int x = q % PX, y = q / PX, i = x - 1, j = y - 1;
                const PointCoef<T> &pc = sp[k].pc;
                if (colour <= 2) {
                    if (pc.oku && pc.ui && ((i - j) & 1) == colour - 1) {
                        const T r = sp[k].bu - optile_u(c, s, pc, x, y);
                        s.U[y + 1][x + 1] = s.U[y + 1][x + 1] + (omega * r) / sp[k].du;
                    }
                } else if (pc.okv && pc.vi && ((i - j) & 1) == colour - 3) {
                    const T r = sp[k].bv - optile_v(c, s, pc, x, y);
                    s.V[y + 1][x + 1] = s.V[y + 1][x + 1] + (omega * r) / sp[k].dv;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int q = threadIdx.x + SMG_THREADS * k;
        if (q >= PX * PY) continue;
        const int x = q % PX, y = q / PX;
        if (sp[k].pc.oku) xu_out[sp[k].pc.ku] = s.U[y + 1][x + 1];
        if (sp[k].pc.okv) xv_out[sp[k].pc.kv] = s.V[y + 1][x + 1];
    }
}

// ===========================================================================
// Restricted additive Schwarz (SharedMemorySpec, src/SharedMemorySpec.jl):
// moving velocities between the full grid and a subdomain context on the
// same device.  A subdomain covers the cells [i0, i0 + nxg) x [j0, j0 + nyg)
// of the full grid (its u points i0 .. i0 + nxg, its v points j0 .. j0 + nyg).
struct RasPart {
    int i0, j0, nxg, nyg;
    int leave_i1, leave_im, leave_j1, leave_jn;  // edges where the partition of unity stays 1
    int overlap;
};

// schwarzRestrictVelocities! (SharedMemorySpec.jl:306-331): the subdomain's
// u and v are the full model's on its points (blockIdx.y 0: u, 1: v).
__global__ void k_ras_restrict(RasPart p, int nxuf, int nxvf, const double *Uf, const double *Vf, double *Ug,
                               double *Vg)
{
    const bool v = blockIdx.y != 0;
    const int mx = v ? p.nxg : p.nxg + 1, my = v ? p.nyg + 1 : p.nyg;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx;
    if (v)
        Vg[k] = Vf[(p.i0 + i) + nxvf * (p.j0 + j)];
    else
        Ug[k] = Uf[(p.i0 + i) + nxuf * (p.j0 + j)];
}

// One factor of schwarzPartitionOfUnity (SharedMemorySpec.jl:395-407) at the
// 1-based index i of m: min(1, (i - 1) / ov, (m - i) / ov), an edge left out
// where it is kept at 1.
__device__ __forceinline__ double ras_pou(int i, int m, bool leave_lo, bool leave_hi, int ov)
{
    const double lo = leave_lo ? INFINITY : (double)(i - 1) / (double)ov;
    const double hi = leave_hi ? INFINITY : (double)(m - i) / (double)ov;
    return fmin(1.0, fmin(lo, hi));
}

// schwarzProlongVelocities! (SharedMemorySpec.jl:351-382) for one
// subdomain: u += ((1 - damping) pou) u_g on its points, with the partition
// of unity ramping over 2 overlap points across u's x edges and 2 overlap - 1
// across its y edges (the other way round for v).
__global__ void k_ras_prolong(RasPart p, int nxuf, int nxvf, double *Uf, double *Vf, const double *Ug,
                              const double *Vg, double one_minus_damping)
{
    const bool v = blockIdx.y != 0;
    const int mx = v ? p.nxg : p.nxg + 1, my = v ? p.nyg + 1 : p.nyg;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx;
    const int ovx = v ? 2 * p.overlap - 1 : 2 * p.overlap, ovy = v ? 2 * p.overlap : 2 * p.overlap - 1;
    const double pou = ras_pou(i + 1, mx, p.leave_i1, p.leave_im, ovx) * ras_pou(j + 1, my, p.leave_j1, p.leave_jn, ovy);
    const double t = one_minus_damping * pou;
    if (v) {
        double *a = &Vf[(p.i0 + i) + nxvf * (p.j0 + j)];
        *a = *a + t * Vg[k];
    } else {
        double *a = &Uf[(p.i0 + i) + nxuf * (p.j0 + j)];
        *a = *a + t * Ug[k];
    }
}

// A subdomain model as schwarzModel builds it (SharedMemorySpec.jl:181-287)
// from the full model's fields, with the initial fields of setup_fields
// (src/Fields/Fields.jl:26-115): bed, thickness, grounded fraction, u, v,
// viscosity, glen_b and the Weertman coefficient and accumulation arrays
// sliced from the full grid; eta_av the bottom viscosity level; quad_f2 =
// h / (3 eta_av) on the mask; dsdh and the u/v grounded fractions 1; the
// wavelet masks all set; every other field 0.  One thread per entry of the
// largest of the subdomain's h, u, v and sigma grids.
struct RasRefresh {
    RasPart p;
    int nxf, nyf, ns;  // full h grid, sigma levels
    const double *b, *h, *gf, *wc, *acc, *u, *v, *eta, *glenb;  // full-grid sources
    double *f[WAVI_NFIELDS];                                   // subdomain fields (nullptr: not a double field)
    uint8_t *mwu, *mwv;
    const uint8_t *hm;                                         // subdomain h mask
};

__global__ void k_ras_refresh(RasRefresh r)
{
    const RasPart &p = r.p;
    const int nxg = p.nxg, nyg = p.nyg, nh = nxg * nyg;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nh) {
        const int i = k % nxg, j = k / nxg;
        const int sidx = (p.i0 + i) + r.nxf * (p.j0 + j);
        for (int id = WAVI_H_B; id <= WAVI_P_ACCUMULATION_RATE; ++id) r.f[id][k] = 0.0;
        r.f[WAVI_H_B][k] = r.b[sidx];
        r.f[WAVI_H_H][k] = r.h[sidx];
        r.f[WAVI_H_GROUNDED_FRACTION][k] = r.gf[sidx];
        r.f[WAVI_P_WEERTMAN_C][k] = r.wc[sidx];
        r.f[WAVI_P_ACCUMULATION_RATE][k] = r.acc[sidx];
        r.f[WAVI_H_DSDH][k] = 1.0;
        const double eta0 = r.eta[sidx];
        r.f[WAVI_H_ETA_AV][k] = eta0;
        r.f[WAVI_H_QUAD_F2][k] = r.hm[k] ? r.h[sidx] / (3.0 * eta0) : 0.0;
    }
    const int nu = (nxg + 1) * nyg;
    if (k < nu) {
        const int i = k % (nxg + 1), j = k / (nxg + 1);
        for (int id = WAVI_U_S; id <= WAVI_U_RESIDUAL; ++id) r.f[id][k] = 0.0;
        r.f[WAVI_U_U][k] = r.u[(p.i0 + i) + (r.nxf + 1) * (p.j0 + j)];
        r.f[WAVI_U_GROUNDED_FRACTION][k] = 1.0;
        r.f[WAVI_WU_WAVELETS][k] = 0.0;
        r.f[WAVI_WU_CORRECTION_COARSE][k] = 0.0;
        r.mwu[k] = 1;
    }
    const int nv = nxg * (nyg + 1);
    if (k < nv) {
        const int i = k % nxg, j = k / nxg;
        for (int id = WAVI_V_S; id <= WAVI_V_RESIDUAL; ++id) r.f[id][k] = 0.0;
        r.f[WAVI_V_V][k] = r.v[(p.i0 + i) + r.nxf * (p.j0 + j)];
        r.f[WAVI_V_GROUNDED_FRACTION][k] = 1.0;
        r.f[WAVI_WV_WAVELETS][k] = 0.0;
        r.f[WAVI_WV_CORRECTION_COARSE][k] = 0.0;
        r.mwv[k] = 1;
    }
    if (k < nh * r.ns) {
        const int l = k / nh, q = k % nh;
        const int i = q % nxg, j = q / nxg;
        const int sidx = (p.i0 + i) + r.nxf * ((p.j0 + j) + r.nyf * l);
        r.f[WAVI_S_ETA][k] = r.eta[sidx];
        r.f[WAVI_S_GLEN_B][k] = r.glenb[sidx];
    }
}

// u = damping u, v = damping v on the full grid (SharedMemorySpec.jl:134-135).
__global__ void k_scale_uv(double *u, int nu, double *v, int nv, double a)
{
    const bool isv = blockIdx.y != 0;
    double *x = isv ? v : u;
    const int n = isv ? nv : nu;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = a * x[k];
}

// ===========================================================================
// Coarse CG: IterativeSolvers.cg! on A_c = R A P (update_preconditioners.jl:73)
// with x0 the cached warm start: three kernels per iteration and no host
// round trip.  Kernels that end in a reduction write one partial per 16x16
// block; the next kernel folds them with the canonical tree in every
// workgroup (all get the same bits), so no workgroup waits for another inside
// a kernel.  Scalars a kernel publishes are double-buffered by iteration
// parity and every workgroup writes the same value; a kernel never reads a
// scalar that another workgroup of the same kernel writes.
// PRE = true (inside the flexible CG's linear solves only) runs
// preconditioned CG with M = D^-1, D = |diag A_c| (k_coarse_diag): the
// direction takes D^-1 r, alpha and beta take (r, D^-1 r), and the stop test
// stays on ||r||.  PRE = false is cg! as the reference runs it.

// Stop report of a coarse CG in host-mapped pinned memory: written once, by
// the launch that stops, so the host learns of the stop without a copy on
// the stream.
struct CGMHostFlag {
    int done;
    int iters;
};

struct CGScalars {
    double res[2];    // residual norm by iteration parity (written by k_cg_dir)
    double tol;       // max(reltol ||r0||, abstol)
    int it[2];        // iterations done, by parity
    int done;         // sticky stop flag, written only by k_cg_dir
    int iters;        // iterations at termination
    double rz[2];     // Jacobi-preconditioned CG: (r, D^-1 r) by parity (written by k_cg_dir)
};

template <typename T>
struct CGArgs {
    GridDims g;
    OpCoef<T> c;
    int levels;
    const uint8_t *mwu, *mwv, *ui, *vi;
    T *xu, *xv;       // solution (warm start in, result out)
    T *ru, *rv;       // residual (right-hand side on entry to setup)
    T *pu, *pv;       // search direction u
    T *qu, *qv;       // c = A_c u
    T *tu, *tv;       // P u on the fine grid
    double *Pr, *Pd, *Pb;
    double *Pz;                       // Jacobi: partials of (r, D^-1 r)
    const double *dinvu, *dinvv;      // Jacobi: 1 / |diag(A_c)|, 0 off the wavelet masks
    CGScalars *st;
    CGMHostFlag *hflag;   // host-mapped stop report
    double tol_coarse, reltol;
    int maxiter;
    int nblk;         // launched workgroups, a multiple of 8
};

// XCD-aware tile order: workgroups are dispatched round-robin over the 8
// XCDs, so workgroup w takes tile (w % 8) * (nblk / 8) + w / 8 and each XCD
// walks one contiguous band of tiles in row-major order (whole grid rows of
// the arrays, which the coarse-CG kernels read fastest: a column-major walk
// was measured 4.5% slower for k_cgm_iter at 1280 x 160).
__device__ __forceinline__ int xcd_tile(int nblk)
{
    const int w = blockIdx.x;
    return (w & 7) * (nblk >> 3) + (w >> 3);
}

__device__ __forceinline__ int band_tile(const GridDims &g, int L, int &bx, int &by)
{
    bx = L % g.nbx;
    by = L / g.nbx;
    return L;
}

// c = R A t for the tile: A into tu / tv, then the transposed inverse Haar,
// zero off the wavelet masks.
template <typename T, bool LD = false>
__device__ __forceinline__ void tile_apply_restrict(const CGArgs<T> &a, int i0, int j0, bool has_u, bool has_v,
                                                    const PointCoef<T> &pc, OpTile<T> &s, T (*tu)[TILE + 1],
                                                    T (*mu)[TILE + 1], T (*tv)[TILE + 1], T (*mv)[TILE + 1])
{
    const HaarRowMasks hm = haar_row_masks(a.levels, 1, i0, j0, a.g, has_u, has_v);
    optile_apply(a.g, a.c, a.tu, a.tv, i0, j0, pc, s, tu, tv);
    haar_pm_tile2m<T, LD>(a.levels, hm, tu, mu, tv, mv, i0, j0, a.g, has_u, has_v);
}

// setup (t = P x0 computed beforehand): r = b - R A t (b arrives in r),
// u = 0, partials of ||r0||^2 and ||b||^2 (PRE: also of (r0, D^-1 r0));
// resets the scalars.
template <typename T, bool PRE>
__global__ void __launch_bounds__(256) k_cg_setup(CGArgs<T> a)
{
    __shared__ OpTile<T> s;
    __shared__ T tu[TILE][TILE + 1], mu[TILE][TILE + 1], tv[TILE][TILE + 1], mv[TILE][TILE + 1];
    __shared__ double lds[8];
    const GridDims &g = a.g;
    const int L = xcd_tile(a.nblk);
    if (L >= g.nbx * g.nby) return;
    if (L == 0 && threadIdx.x == 0) {
        a.st->done = 0;
        a.st->iters = 0;
        a.st->it[0] = 0;
        a.st->it[1] = 0;
        // also reset on the stream: launches queued past the previous
        // CG's stop, which rewrite the report, may still have been running
        // when the host reset it
        a.hflag->done = 0;
        a.hflag->iters = 0;
    }
    int bx, by;
    band_tile(g, L, bx, by);
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    PointCoef<T> pc;
    point_prefetch(g, a.c, a.mwu, a.mwv, i0 + lx, j0 + ly, has_u, has_v, pc);
    const T bu = pc.oku ? a.ru[pc.ku] : T(0), bv = pc.okv ? a.rv[pc.kv] : T(0);
    tile_apply_restrict<T>(a, i0, j0, has_u, has_v, pc, s, tu, mu, tv, mv);
    double r2u = 0.0, b2u = 0.0, r2v = 0.0, b2v = 0.0, z2u = 0.0, z2v = 0.0;
    if (pc.oku) {
        const T r = bu - (pc.mwu ? tu[ly][lx] : T(0));
        a.ru[pc.ku] = r;
        a.pu[pc.ku] = T(0);
        r2u = (double)r * (double)r;
        b2u = (double)bu * (double)bu;
        if constexpr (PRE) z2u = (double)r * (a.dinvu[pc.ku] * (double)r);
    }
    if (pc.okv) {
        const T r = bv - (pc.mwv ? tv[ly][lx] : T(0));
        a.rv[pc.kv] = r;
        a.pv[pc.kv] = T(0);
        r2v = (double)r * (double)r;
        b2v = (double)bv * (double)bv;
        if constexpr (PRE) z2v = (double)r * (a.dinvv[pc.kv] * (double)r);
    }
    tree256x2(r2u, r2v, lds);
    tree256x2(b2u, b2v, lds);
    if constexpr (PRE) tree256x2(z2u, z2v, lds);
    if (threadIdx.x == 0) {
        if (has_u) {
            a.Pr[bx + g.nbxu * by] = r2u;
            a.Pb[bx + g.nbxu * by] = b2u;
            if constexpr (PRE) a.Pz[bx + g.nbxu * by] = z2u;
        }
        if (has_v) {
            const int o = g.nbxu * g.nbyu + bx + g.nbxv * by;
            a.Pr[o] = r2v;
            a.Pb[o] = b2v;
            if constexpr (PRE) a.Pz[o] = z2v;
        }
    }
}

// A CG scalar written by an earlier kernel, read as a vector load: a scalar
// load of it comes back under the same wait as the kernel's arguments, which
// the launch's later vector loads (the partials) need first, so its round
// trip would come before theirs.  The index is a zero the compiler cannot
// see through (a register move, not a load), which makes the load a vector
// load, waited for where its value is used.
template <typename V>
__device__ __forceinline__ V cg_scalar(const V *p)
{
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return p[z];
}

// part 1 (iteration parity par): residual norm and stop test, then
// u = r + beta u with beta = res^2 / prev^2 (PRE: u = D^-1 r + beta u with
// beta = rz / prev rz), t = P u.  The stop decision is
// recomputed from the partials each time, so it needs no flag: after a stop
// the partials are frozen and it[1 - par] repeats the count.  LD: the Haar
// rows loads first (haar_pm_row_ld), the same bits.
template <typename T, bool PRE, bool LD = false>
__device__ __forceinline__ bool cg_dir_tile(const CGArgs<T> &a, int par, const int L)
{
    __shared__ T tu[TILE][TILE + 1], mu[TILE][TILE + 1], tv[TILE][TILE + 1], mv[TILE][TILE + 1];
    __shared__ double lds[8];
    const GridDims &g = a.g;
    if (L >= g.nbx * g.nby) return false;
    int bx, by;
    band_tile(g, L, bx, by);
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int i = i0 + lx, j = j0 + ly;
    const bool oku = has_u && i < g.nxu && j < g.nyu, okv = has_v && i < g.nxv && j < g.nyv;
    const int ku = i + g.nxu * j, kv = i + g.nxv * j;
    // independent loads first
    const int it = cg_scalar(&a.st->it[par]);
    const double told = cg_scalar(&a.st->tol), prev_res = cg_scalar(&a.st->res[1 - par]);
    const T ru = ld_or(oku, a.ru, ku, T(0)), pu = ld_or(oku, a.pu, ku, T(0));
    const T rv = ld_or(okv, a.rv, kv, T(0)), pv = ld_or(okv, a.pv, kv, T(0));
    const bool mwu = ld_or(oku, a.mwu, ku, (uint8_t)0) != 0, mwv = ld_or(okv, a.mwv, kv, (uint8_t)0) != 0;
    const bool iu = ld_or(oku, a.ui, ku, (uint8_t)0) != 0, iv = ld_or(okv, a.vi, kv, (uint8_t)0) != 0;
    const HaarRowMasks hm = haar_row_masks(a.levels, 0, i0, j0, g, has_u, has_v);
    double dzu = 0.0, dzv = 0.0, rz = 0.0, prev_rz = 0.0;
    if constexpr (PRE) {
        dzu = ld_or(oku, a.dinvu, ku, 0.0);
        dzv = ld_or(okv, a.dinvv, kv, 0.0);
        prev_rz = cg_scalar(&a.st->rz[1 - par]);
    }
    double res;
    if constexpr (PRE) {
        double r2;
        lanes_sum2(a.Pr, a.Pz, g.npart, lds, r2, rz);
        res = sqrt(r2);
    } else {
        res = sqrt(lanes_sum(a.Pr, g.npart, lds));
    }
    double tol;
    if (it == 0) {
        const double abstol = a.tol_coarse * sqrt(lanes_sum(a.Pb, g.npart, lds));
        tol = a.reltol * res;
        if (abstol > tol) tol = abstol;
    } else {
        tol = told;
    }
    const bool done = (it >= a.maxiter) || (res <= tol);
    if (threadIdx.x == 0) {
        if (it == 0) a.st->tol = tol;
        a.st->res[par] = res;
        if constexpr (PRE) a.st->rz[par] = rz;
        if (done) {
            a.st->done = 1;
            a.st->iters = it;
            a.st->it[1 - par] = it;
            if (L == 0) {
                a.hflag->iters = it;
                a.hflag->done = 1;
            }
        }
    }
    if (done) return true;
    T beta;
    if constexpr (PRE) {
        beta = it == 0 ? T(0) : (T)(rz / prev_rz);
    } else {
        const double prev = it == 0 ? 1.0 : prev_res;
        beta = (T)((res * res) / (prev * prev));
    }
    T unu = T(0), unv = T(0);
    if (oku) {
        if constexpr (PRE)
            unu = (T)(dzu * (double)ru) + beta * pu;
        else
            unu = ru + beta * pu;
        a.pu[ku] = unu;
    }
    if (okv) {
        if constexpr (PRE)
            unv = (T)(dzv * (double)rv) + beta * pv;
        else
            unv = rv + beta * pv;
        a.pv[kv] = unv;
    }
    tu[ly][lx] = mwu ? unu : T(0);
    tv[ly][lx] = mwv ? unv : T(0);
    __syncthreads();
    haar_pm_tile2m<T, LD>(a.levels, hm, tu, mu, tv, mv, i0, j0, g, has_u, has_v);
    if (oku) a.tu[ku] = iu ? tu[ly][lx] : T(0);
    if (okv) a.tv[kv] = iv ? tv[ly][lx] : T(0);
    return false;
}

// part 2: c = R A t and partials of dot(u, c).  EARLY: return at once after
// a stop (written only by earlier kernels).  Otherwise the launch runs and
// stores in full: after a stop only k_cg_update reads c and the partials, and
// it changes nothing then; a return on the flag would put its round trip
// before every load of the launch.
template <typename T, bool LD = false, bool EARLY = true>
__device__ __forceinline__ void cg_apply_tile(const CGArgs<T> &a, const int L)
{
    __shared__ OpTile<T> s;
    __shared__ T tu[TILE][TILE + 1], mu[TILE][TILE + 1], tv[TILE][TILE + 1], mv[TILE][TILE + 1];
    __shared__ double lds[8];
    const GridDims &g = a.g;
    if (L >= g.nbx * g.nby) return;
    if (EARLY && a.st->done) return;
    int bx, by;
    band_tile(g, L, bx, by);
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    PointCoef<T> pc;
    point_prefetch(g, a.c, a.mwu, a.mwv, i0 + lx, j0 + ly, has_u, has_v, pc);
    const T pu = ld_or(pc.oku, a.pu, pc.ku, T(0)), pv = ld_or(pc.okv, a.pv, pc.kv, T(0));
    tile_apply_restrict<T, LD>(a, i0, j0, has_u, has_v, pc, s, tu, mu, tv, mv);
    double du = 0.0, dv = 0.0;
    if (pc.oku) {
        const T cv = pc.mwu ? tu[ly][lx] : T(0);
        a.qu[pc.ku] = cv;
        du = (double)(pu * cv);
    }
    if (pc.okv) {
        const T cv = pc.mwv ? tv[ly][lx] : T(0);
        a.qv[pc.kv] = cv;
        dv = (double)(pv * cv);
    }
    tree256x2(du, dv, lds);
    if (threadIdx.x == 0) {
        if (has_u) a.Pd[bx + g.nbxu * by] = du;
        if (has_v) a.Pd[g.nbxu * g.nbyu + bx + g.nbxv * by] = dv;
    }
}

// part 3: alpha = res^2 / dot(u, c) (PRE: rz / dot(u, c)), x += alpha u,
// r -= alpha c, partials of ||r||^2 (PRE: and of (r, D^-1 r)) and the
// iteration count.  EARLY: return at once after a stop.  Otherwise a launch
// after a stop stores x, r, their partials and the count unchanged (selected
// by the flag, which is then needed only at the stores), so that the flag's
// round trip does not precede the loads.
template <typename T, bool PRE, bool EARLY = true>
__device__ __forceinline__ void cg_update_tile(const CGArgs<T> &a, int par, const int L)
{
    __shared__ double lds[8];
    const GridDims &g = a.g;
    if (L >= g.nbx * g.nby) return;
    if (EARLY && a.st->done) return;
    const bool stopped = !EARLY && cg_scalar(&a.st->done) != 0;
    int bx, by;
    band_tile(g, L, bx, by);
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int i = bx * TILE + lx, j = by * TILE + ly;
    const bool oku = has_u && i < g.nxu && j < g.nyu, okv = has_v && i < g.nxv && j < g.nyv;
    const int ku = i + g.nxu * j, kv = i + g.nxv * j;
    const int it = cg_scalar(&a.st->it[par]);
    const double res = cg_scalar(&a.st->res[par]);
    const T xu = ld_or(oku, a.xu, ku, T(0)), pu = ld_or(oku, a.pu, ku, T(0)), ru = ld_or(oku, a.ru, ku, T(0)),
            qu = ld_or(oku, a.qu, ku, T(0));
    const T xv = ld_or(okv, a.xv, kv, T(0)), pv = ld_or(okv, a.pv, kv, T(0)), rv = ld_or(okv, a.rv, kv, T(0)),
            qv = ld_or(okv, a.qv, kv, T(0));
    double dzu = 0.0, dzv = 0.0, rzc = 0.0;
    if constexpr (PRE) {
        dzu = ld_or(oku, a.dinvu, ku, 0.0);
        dzv = ld_or(okv, a.dinvv, kv, 0.0);
        rzc = cg_scalar(&a.st->rz[par]);
    }
    const double dot = lanes_sum(a.Pd, g.npart, lds);
    T alpha;
    if constexpr (PRE)
        alpha = (T)(rzc / dot);
    else
        alpha = (T)((res * res) / dot);
    if (threadIdx.x == 0) a.st->it[1 - par] = stopped ? it : it + 1;
    double r2u = 0.0, r2v = 0.0, z2u = 0.0, z2v = 0.0;
    if (oku) {
        const T x = xu + alpha * pu, rn = ru - alpha * qu;
        a.xu[ku] = stopped ? xu : x;
        const T r = stopped ? ru : rn;
        a.ru[ku] = r;
        r2u = (double)r * (double)r;
        if constexpr (PRE) z2u = (double)r * (dzu * (double)r);
    }
    if (okv) {
        const T x = xv + alpha * pv, rn = rv - alpha * qv;
        a.xv[kv] = stopped ? xv : x;
        const T r = stopped ? rv : rn;
        a.rv[kv] = r;
        r2v = (double)r * (double)r;
        if constexpr (PRE) z2v = (double)r * (dzv * (double)r);
    }
    tree256x2(r2u, r2v, lds);
    if constexpr (PRE) tree256x2(z2u, z2v, lds);
    if (threadIdx.x == 0) {
        if (has_u) a.Pr[bx + g.nbxu * by] = r2u;
        if (has_v) a.Pr[g.nbxu * g.nbyu + bx + g.nbxv * by] = r2v;
        if constexpr (PRE) {
            if (has_u) a.Pz[bx + g.nbxu * by] = z2u;
            if (has_v) a.Pz[g.nbxu * g.nbyu + bx + g.nbxv * by] = z2v;
        }
    }
}

template <typename T, bool PRE, bool LD>
__global__ void __launch_bounds__(256) k_cg_dir(CGArgs<T> a, int par)
{
    (void)cg_dir_tile<T, PRE, LD>(a, par, xcd_tile(a.nblk));
}

template <typename T, bool LD, bool EARLY>
__global__ void __launch_bounds__(256) k_cg_apply(CGArgs<T> a)
{
    cg_apply_tile<T, LD, EARLY>(a, xcd_tile(a.nblk));
}

template <typename T, bool PRE, bool EARLY>
__global__ void __launch_bounds__(256) k_cg_update(CGArgs<T> a, int par)
{
    cg_update_tile<T, PRE, EARLY>(a, par, xcd_tile(a.nblk));
}

// Diagonal of the coarse-CG operator A_c = M_w R M_i A M_i P M_w
// (update_preconditioners.jl:73, with the masks the CG applies) at the
// points k of the wavelet masks: e_k' A_c e_k = phi' M_i A M_i phi with
// phi = P e_k, the Haar basis function of k, which is separable
// (Py (x) Px, from the inverse-transform rows of c_haar) and lives in k's
// Haar block, inside k's tile.  Workgroup (tile, component, coarse row ky)
// stages phi for each coarse column kx with a zero halo, evaluates the
// operator at the tile's points and sums phi (A phi) (FP64, canonical
// tree).  Writes 1 / |diagonal| (0 where the diagonal is 0 or off the
// masks) and, if diag is given, the diagonal.  Grid: (tiles, 32).
__global__ void __launch_bounds__(256) k_coarse_diag(GridDims g, OpCoef<double> c, const uint8_t *mwu,
                                                     const uint8_t *mwv, int levels, double *dinvu, double *dinvv,
                                                     double *diagu, double *diagv)
{
    __shared__ OpTile<double> s;
    __shared__ double lds[8];
    const int L = blockIdx.x;
    const int comp = blockIdx.y & 1, ky = blockIdx.y >> 1;
    const int bx = L % g.nbx, by = L / g.nbx;
    const bool has = comp == 0 ? (bx < g.nbxu && by < g.nbyu) : (bx < g.nbxv && by < g.nbyv);
    const int nxw = comp == 0 ? g.nxu : g.nxv, nyw = comp == 0 ? g.nyu : g.nyv;
    const int i0 = bx * TILE, j0 = by * TILE;
    if (!has || j0 + ky >= nyw) return;  // the whole workgroup
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int B = 2 << levels;
    PointCoef<double> pc;
    point_prefetch_op(g, c, i0 + lx, j0 + ly, comp == 0, comp == 1, pc);
    CellCoef<double> c0, c1;
    cell_prefetch(g, c, i0, j0, threadIdx.x, c0);
    cell_prefetch(g, c, i0, j0, threadIdx.x + 256, c1);
    const bool live = comp == 0 ? (pc.oku && pc.ui) : (pc.okv && pc.vi);
    // y factor of phi at this thread's row
    const int jb = ly & ~(B - 1), kjb = ky & ~(B - 1);
    double fy = 0.0;
    if (jb == kjb && j0 + ly < nyw) {
        const int my = min(nyw - (j0 + jb), B), col = ky - kjb;
        const uint32_t nz = c_haar[levels].nz[0][my][ly - jb], ng = c_haar[levels].neg[0][my][ly - jb];
        if ((nz >> col) & 1u) fy = ((ng >> col) & 1u) ? -1.0 : 1.0;
    }
    const int ib = lx & ~(B - 1);
    const int mx = i0 + lx < nxw ? min(nxw - (i0 + ib), B) : 0;
    const uint32_t nzx = mx > 0 ? c_haar[levels].nz[0][mx][lx - ib] : 0u;
    const uint32_t ngx = mx > 0 ? c_haar[levels].neg[0][mx][lx - ib] : 0u;
    // zero windows (the halo and the other component stay zero)
    for (int q = threadIdx.x; q < 18 * 19; q += 256) {
        (&s.U[0][0])[q] = 0.0;
        (&s.V[0][0])[q] = 0.0;
    }
    const uint8_t *mw = comp == 0 ? mwu : mwv;
    for (int kx = 0; kx < TILE && i0 + kx < nxw; ++kx) {
        const int kib = kx & ~(B - 1), col = kx - kib;
        double fx = 0.0;
        if (ib == kib && ((nzx >> col) & 1u)) fx = ((ngx >> col) & 1u) ? -1.0 : 1.0;
        const double phi = live ? fx * fy : 0.0;
        __syncthreads();  // previous intermediates read
        if (comp == 0)
            s.U[ly + 1][lx + 1] = phi;
        else
            s.V[ly + 1][lx + 1] = phi;
        __syncthreads();
        cell_compute(c, c0, threadIdx.x, s);
        cell_compute(c, c1, threadIdx.x + 256, s);
        __syncthreads();
        double v = 0.0;
        if (phi != 0.0) v = phi * (comp == 0 ? optile_u(c, s, pc, lx, ly) : optile_v(c, s, pc, lx, ly));
        v = tree256(v, lds);
        if (threadIdx.x == 0) {
            const int k = (i0 + kx) + nxw * (j0 + ky);
            const double d = mw[k] ? v : 0.0;
            (comp == 0 ? dinvu : dinvv)[k] = d != 0.0 ? 1.0 / fabs(d) : 0.0;
            if (diagu) (comp == 0 ? diagu : diagv)[k] = d;
        }
    }
}

// The whole coarse CG of a small grid in one workgroup (after k_cg_setup):
// the workgroup runs each part of an iteration on every tile in turn, so
// the barrier between two tiles stands in for the kernel boundary.  Tiles
// read only what earlier parts wrote and write only their own points,
// partials and parity slots, so the results are those of the three
// launches per iteration, bit for bit.  Ends at the stop, which every
// thread of every tile decides alike from the same partials and scalars,
// after maxiter + 1 iterations at most.
template <typename T, bool PRE>
__global__ void __launch_bounds__(256) k_cg_solve1(CGArgs<T> a)
{
    const int ntile = a.g.nbx * a.g.nby;
    for (int q = 0; q <= a.maxiter; ++q) {
        const int par = q & 1;
        bool stop = false;
        for (int L = 0; L < ntile; ++L) {
            stop = cg_dir_tile<T, PRE>(a, par, L);
            __syncthreads();
        }
        if (stop) break;
        for (int L = 0; L < ntile; ++L) {
            cg_apply_tile<T>(a, L);
            __syncthreads();
        }
        for (int L = 0; L < ntile; ++L) {
            cg_update_tile<T, PRE>(a, par, L);
            __syncthreads();
        }
    }
}

// ===========================================================================
// Mixed-precision coarse CG: the Chronopoulos-Gear form of CG (one combined
// reduction per iteration) on FP32 vectors, with FP32 sums inside a
// workgroup and FP64 sums across workgroups and FP64 scalars, one launch per
// iteration.  In exact arithmetic it produces the iterates of the CG above;
// launch k computes
//     p = r + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,
//     w = A_c r,  gamma = (r, r),  delta = (w, r),
//     beta = gamma / gamma_old,  alpha = gamma / (delta - beta gamma / alpha_old).
// w = R A P r needs P r on the tile and on a one-point halo around it.  The
// tile's own part is the block transform of its own r (the Haar blocks never
// cross a tile).  The halo belongs to the neighbours' rings (their outermost
// rows and columns), where every workgroup stores the exact images P r, P w
// and P s of its own vectors at the end of each launch; the next launch forms
// the halo of P r_new = P r - alpha (P w + beta P s) from them, one linear
// step from exact values.  The rings, the partials and the scalars are
// double-buffered: the launch of parity PAR reads set PAR and writes set
// 1 - PAR, so no workgroup reads what another workgroup of the same launch
// writes.
//
// Layout: every array of the CG lives in one FP32 arena in a zero-padded 2-D
// layout shared by the u grid, the v grid and the h cells: point (i, j) at
// (i + 1) + px (j + 1), with a zero border of one point and room for whole
// tiles, so that windows, halos and tile stores need no bounds tests (what a
// tile stores past the end of a grid is zero).

struct CGMState {
    double igamma;  // 1 / ||r||^2 at the start of the launch that wrote this state
    double alpha;   // that launch's step length
    double ialpha;  // 1 / alpha
    double tol2;    // square of the stopping tolerance, fixed at iteration 0
    int it;         // iterations done
    int done;       // stop decided
};

struct CGMScalars {
    CGMState s[2];  // by launch parity
    int done;       // sticky stop flag, read by the host only
    int iters;      // iterations at the stop
};

// Arena arrays (component u; the v component is the next array).
enum CGMArray {
    CGM_R = 0,       // residual r
    CGM_W = 2,       // w = A_c r
    CGM_P = 4,       // search direction p
    CGM_S = 6,       // s = A_c p
    CGM_X = 8,       // solution
    CGM_HU = 10,     // thickness on u / v points, 0 off the masks
    CGM_BU = 12,     // beta_eff on u / v points, 0 off the masks
    CGM_CELL = 14,   // per h cell: (h*eta_av on the h mask, dimp) pairs (two arrays' room)
    CGM_CORNER = 16, // per h cell: corner h*eta_av (0 off the c grid and mask)
    CGM_DINV = 17,   // Jacobi-preconditioned CG: 1 / |diag(A_c)| on u / v points, 0 off the masks
    CGM_NARRAY = 19
};

// Ring images, compact per tile: for each parity, tile (ntile + 1 tiles, the
// last one all zero and never written, standing in for tiles past the edge),
// kind (0 P r, 1 P s, 2 P w) and component, 64 slots: the tile's bottom row
// (y = 0, slots 0-15 by x), top row (y = 15, 16-31), left column (x = 0,
// 32-47 by y) and right column (x = 15, 48-63).  A side of a neighbour's halo
// is then 16 consecutive floats.
__device__ __forceinline__ unsigned ring_at(int par, int ntile, int tile, int kind, int comp)
{
    return ((unsigned)(par * (ntile + 1) + tile) * 6u + (unsigned)(2 * kind + comp)) * 64u;
}

// Store v at the ring slots of tile position (x, y) (a corner has two).
__device__ __forceinline__ void ring_put(float *R, unsigned base, int x, int y, float v)
{
    if (y == 0) R[base + x] = v;
    if (y == 15) R[base + 16 + x] = v;
    if (x == 0) R[base + 32 + y] = v;
    if (x == 15) R[base + 48 + y] = v;
}

// The ring slots of tile position (x, y): a row slot and a column slot, -1
// where there is none.
struct RingSlots {
    int row, col;
    __device__ __forceinline__ RingSlots(int x, int y)
    {
        row = y == 0 ? x : (y == 15 ? 16 + x : -1);
        col = x == 0 ? 32 + y : (x == 15 ? 48 + y : -1);
    }
    __device__ __forceinline__ void put(float *R, unsigned base, float v) const
    {
        if (row >= 0) R[base + (unsigned)row] = v;
        if (col >= 0) R[base + (unsigned)col] = v;
    }
};

struct CGMLayout {
    float *F;       // FP32 arena, array A at F + A * ns
    uint8_t *M;     // mask bits (bit 0 inner point, bit 1 wavelet mask): u at M, v at M + ns
    unsigned px;    // row pitch
    unsigned ns;    // array stride (px * rows)
    __device__ __forceinline__ unsigned at(int i, int j) const { return (unsigned)(i + 1) + px * (unsigned)(j + 1); }
};

// Coefficients and masks into the arena (once per V-cycle); 0 off the masks.
__global__ void k_cgm_pack(GridDims g, OpCoef<float> c, const uint8_t *mwu, const uint8_t *mwv, CGMLayout Lo,
                           const double *dinvu, const double *dinvv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nh = g.nx * g.ny, nu = g.nxu * g.nyu, nv = g.nxv * g.nyv;
    float *F = Lo.F;
    const unsigned ns = Lo.ns;
    // every load first (ld_or): the stores into the arena come after, since
    // the compiler cannot move a load past them
    const bool kh = k < nh, ku = k < nu, kv = k < nv;
    const int a = k % g.nx, b = k / g.nx;
    const bool okc = kh && a < g.nx - 1 && b < g.ny - 1;
    const int kc = a + (g.nx - 1) * b;
    const bool cmc = ld_or(okc, c.cm, kc, (uint8_t)0) != 0, hmk = ld_or(kh, c.hm, k, (uint8_t)0) != 0;
    const float hcc = ld_or(okc, c.heta_c, kc, 0.f), hhk = ld_or(kh, c.heta_h, k, 0.f), dpk = ld_or(kh, c.dimp, k, 0.f);
    const bool umk = ld_or(ku, c.um, k, (uint8_t)0) != 0, uik = ld_or(ku, c.ui, k, (uint8_t)0) != 0;
    const bool mwuk = ld_or(ku, mwu, k, (uint8_t)0) != 0;
    const float huk = ld_or(ku, c.hu, k, 0.f), buk = ld_or(ku, c.bu, k, 0.f);
    const bool vmk = ld_or(kv, c.vm, k, (uint8_t)0) != 0, vik = ld_or(kv, c.vi, k, (uint8_t)0) != 0;
    const bool mwvk = ld_or(kv, mwv, k, (uint8_t)0) != 0;
    const float hvk = ld_or(kv, c.hv, k, 0.f), bvk = ld_or(kv, c.bv, k, 0.f);
    const double dzu = dinvu ? ld_or(ku, dinvu, k, 0.0) : 0.0, dzv = dinvv ? ld_or(kv, dinvv, k, 0.0) : 0.0;
    if (kh) {
        const float hc = okc ? (cmc ? hcc : 0.f) : 0.f;
        const unsigned o = Lo.at(a, b);
        reinterpret_cast<float2 *>(F + CGM_CELL * ns)[o] = make_float2(hmk ? hhk : 0.f, dpk);
        F[CGM_CORNER * ns + o] = hc;
    }
    if (ku) {
        const unsigned o = Lo.at(k % g.nxu, k / g.nxu);
        F[CGM_HU * ns + o] = umk ? huk : 0.f;
        F[CGM_BU * ns + o] = umk ? buk : 0.f;
        Lo.M[o] = (uint8_t)((uik ? 1 : 0) | (mwuk ? 2 : 0));
        if (dinvu) F[CGM_DINV * ns + o] = (float)fmin(dzu, 1e30);  // finite in FP32
    }
    if (kv) {
        const unsigned o = Lo.at(k % g.nxv, k / g.nxv);
        F[(CGM_HU + 1) * ns + o] = vmk ? hvk : 0.f;
        F[(CGM_BU + 1) * ns + o] = vmk ? bvk : 0.f;
        Lo.M[ns + o] = (uint8_t)((vik ? 1 : 0) | (mwvk ? 2 : 0));
        if (dinvv) F[(CGM_DINV + 1) * ns + o] = (float)fmin(dzv, 1e30);  // finite in FP32
    }
}

// The solution back into the context's compact vectors.
__global__ void k_cgm_unpack(GridDims g, CGMLayout Lo, float *xu, float *xv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < g.nxu * g.nyu) xu[k] = Lo.F[CGM_X * Lo.ns + Lo.at(k % g.nxu, k / g.nxu)];
    if (k < g.nxv * g.nyv) xv[k] = Lo.F[(CGM_X + 1) * Lo.ns + Lo.at(k % g.nxv, k / g.nxv)];
}

struct CGMArgs {
    GridDims g;
    OpCoef<float> c;         // compact coefficients (setup)
    CGMLayout Lo;
    int levels;
    const uint8_t *mwu, *mwv;
    const float *bu, *bv;    // right-hand side (compact)
    const float *x0u, *x0v;  // warm start (compact)
    const float *t0u, *t0v;  // P x0 (compact)
    float *R;                // ring images
    double *Pb;              // partials of ||b||^2 per u / v block
    CGMScalars *st;
    CGMHostFlag *hflag;      // host-mapped stop report, reset here
    int nblk;
};

// setup (P x0 computed beforehand): r0 = b - R A P x0, p = s = 0, x = x0 in
// the arena, P r0 into the P r rings of set 0, partials of ||b||^2; resets
// the scalars.
template <bool PRE>
__global__ void __launch_bounds__(256) k_cgm_setup(CGMArgs a)
{
    __shared__ OpTile<float> s;
    __shared__ float tu[TILE][TILE + 1], mu[TILE][TILE + 1], tv[TILE][TILE + 1], mv[TILE][TILE + 1];
    __shared__ double lds[8];
    const GridDims &g = a.g;
    const int L = xcd_tile(a.nblk);
    if (L >= g.nbx * g.nby) return;
    if (L == 0 && threadIdx.x == 0) {
        for (int p = 0; p < 2; ++p) {
            a.st->s[p].igamma = 0.0;
            a.st->s[p].alpha = 0.0;
            a.st->s[p].ialpha = 0.0;
            a.st->s[p].tol2 = 0.0;
            a.st->s[p].it = 0;
            a.st->s[p].done = 0;
        }
        a.st->done = 0;
        a.st->iters = 0;
        a.hflag->done = 0;
        a.hflag->iters = 0;
    }
    int bx, by;
    const int tile_id = band_tile(g, L, bx, by);
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    float *F = a.Lo.F;
    const unsigned ns = a.Lo.ns, o = a.Lo.at(i0 + lx, j0 + ly);
    PointCoef<float> pc;
    point_prefetch(g, a.c, a.mwu, a.mwv, i0 + lx, j0 + ly, has_u, has_v, pc);
    const float bu = pc.oku ? a.bu[pc.ku] : 0.f, bv = pc.okv ? a.bv[pc.kv] : 0.f;
    const float xu = pc.oku ? a.x0u[pc.ku] : 0.f, xv = pc.okv ? a.x0v[pc.kv] : 0.f;
    optile_apply(g, a.c, a.t0u, a.t0v, i0, j0, pc, s, tu, tv);
    haar_pm_tile2<float>(a.levels, 1, tu, mu, tv, mv, i0, j0, g, has_u, has_v);
    float ru = 0.f, rv = 0.f;
    double b2u = 0.0, b2v = 0.0;
    // Jacobi: the rings hold P D^-1 r and ||b||^2 is the D^-1 norm
    const float dzu = PRE && pc.oku ? F[CGM_DINV * ns + o] : 1.f, dzv = PRE && pc.okv ? F[(CGM_DINV + 1) * ns + o] : 1.f;
    if (pc.oku) {
        ru = bu - (pc.mwu ? tu[ly][lx] : 0.f);
        F[CGM_R * ns + o] = ru;
        F[CGM_P * ns + o] = 0.f;
        F[CGM_S * ns + o] = 0.f;
        F[CGM_X * ns + o] = xu;
        b2u = PRE ? (double)bu * ((double)dzu * (double)bu) : (double)bu * (double)bu;
    }
    if (pc.okv) {
        rv = bv - (pc.mwv ? tv[ly][lx] : 0.f);
        F[(CGM_R + 1) * ns + o] = rv;
        F[(CGM_P + 1) * ns + o] = 0.f;
        F[(CGM_S + 1) * ns + o] = 0.f;
        F[(CGM_X + 1) * ns + o] = xv;
        b2v = PRE ? (double)bv * ((double)dzv * (double)bv) : (double)bv * (double)bv;
    }
    // each thread rewrites only its own entry; the transform ended with a barrier
    tu[ly][lx] = pc.mwu ? (PRE ? dzu * ru : ru) : 0.f;
    tv[ly][lx] = pc.mwv ? (PRE ? dzv * rv : rv) : 0.f;
    __syncthreads();
    haar_pm_tile2<float>(a.levels, 0, tu, mu, tv, mv, i0, j0, g, has_u, has_v);
    if (pc.oku) ring_put(a.R, ring_at(0, g.nbx * g.nby, tile_id, 0, 0), lx, ly, pc.ui ? tu[ly][lx] : 0.f);
    if (pc.okv) ring_put(a.R, ring_at(0, g.nbx * g.nby, tile_id, 0, 1), lx, ly, pc.vi ? tv[ly][lx] : 0.f);
    tree256x2(b2u, b2v, lds);
    if (threadIdx.x == 0) {
        if (has_u) a.Pb[bx + g.nbxu * by] = b2u;
        if (has_v) a.Pb[g.nbxu * g.nbyu + bx + g.nbxv * by] = b2v;
    }
}

// Block Haar transforms of 16-point lines held one point per lane (pos = the
// point's position in the line, blocks of B = 2^(LEV+1) points) as the
// butterflies that compose wavelet_matrix (src/Wavelets/update_wavelets.jl:
// 38-89).  A pair is (o, e = o + h) with o a multiple of 2h.
//   P (inverse), h = B/2 .. 1:           y_o = x_o - x_e,  y_e = x_o + x_e
//   R (transposed inverse), h = 1 .. B/2: y_o = x_o + x_e,  y_e = x_e - x_o
// Points past the end of the grid must hold zero: a pair whose e lies past
// the end then leaves x_o unchanged, as wavelet_matrix does for a partial
// block, and the garbage left in those points never reaches a point in range.
// Two lines at once (the u and the v tile), each stage one v_fmac_f32 with a
// DPP source per line, v += sign * partner(v) with sign in {-1, 0, +1}: the
// product is exact, so the fused multiply-add rounds exactly as the addition
// would.  Every block starts with two wait states: a DPP read of a register
// that the previous VALU instruction wrote needs them, and the compiler
// does not see the VALU writes inside these blocks.  The H = 4 partner is
// row_half_mirror of quad_perm [0,0,0,0], two DPP steps.
#define WAVI_DPP_TAIL " row_mask:0xf bank_mask:0xf bound_ctrl:1"

template <int H>
__device__ __forceinline__ void haar_stage_dpp(float &v0, float &v1, float sg)
{
    if constexpr (H == 1)
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %2 quad_perm:[1,0,3,2]" WAVI_DPP_TAIL
            "\n\tv_fmac_f32_dpp %1, %1, %2 quad_perm:[1,0,3,2]" WAVI_DPP_TAIL
            : "+v"(v0), "+v"(v1) : "v"(sg));
    else if constexpr (H == 2)
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %2 quad_perm:[2,3,0,1]" WAVI_DPP_TAIL
            "\n\tv_fmac_f32_dpp %1, %1, %2 quad_perm:[2,3,0,1]" WAVI_DPP_TAIL
            : "+v"(v0), "+v"(v1) : "v"(sg));
    else if constexpr (H == 8)
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %2 row_ror:8" WAVI_DPP_TAIL
            "\n\tv_fmac_f32_dpp %1, %1, %2 row_ror:8" WAVI_DPP_TAIL
            : "+v"(v0), "+v"(v1) : "v"(sg));
    else {
        float t0, t1;
        asm("s_nop 1\n\tv_mov_b32_dpp %2, %0 quad_perm:[0,0,0,0]" WAVI_DPP_TAIL
            "\n\tv_mov_b32_dpp %3, %1 quad_perm:[0,0,0,0]" WAVI_DPP_TAIL
            "\n\ts_nop 1\n\tv_fmac_f32_dpp %0, %2, %4 row_half_mirror" WAVI_DPP_TAIL
            "\n\tv_fmac_f32_dpp %1, %3, %4 row_half_mirror" WAVI_DPP_TAIL
            : "+v"(v0), "+v"(v1), "=&v"(t0), "=&v"(t1) : "v"(sg));
    }
}

template <int H>
__device__ __forceinline__ float haar_sign_P(int pos)
{
    return ((pos & (H - 1)) == 0) ? ((pos & H) ? 1.f : -1.f) : 0.f;
}

template <int H>
__device__ __forceinline__ float haar_sign_R(int pos)
{
    return ((pos & (H - 1)) == 0) ? ((pos & H) ? -1.f : 1.f) : 0.f;
}

template <int LEV>
__device__ __forceinline__ void haar_lines_P(float &v0, float &v1, int pos)
{
    if constexpr (LEV >= 3) haar_stage_dpp<8>(v0, v1, haar_sign_P<8>(pos));
    if constexpr (LEV >= 2) haar_stage_dpp<4>(v0, v1, haar_sign_P<4>(pos));
    if constexpr (LEV >= 1) haar_stage_dpp<2>(v0, v1, haar_sign_P<2>(pos));
    haar_stage_dpp<1>(v0, v1, haar_sign_P<1>(pos));
}

template <int LEV>
__device__ __forceinline__ void haar_lines_R(float &v0, float &v1, int pos)
{
    haar_stage_dpp<1>(v0, v1, haar_sign_R<1>(pos));
    if constexpr (LEV >= 1) haar_stage_dpp<2>(v0, v1, haar_sign_R<2>(pos));
    if constexpr (LEV >= 2) haar_stage_dpp<4>(v0, v1, haar_sign_R<4>(pos));
    if constexpr (LEV >= 3) haar_stage_dpp<8>(v0, v1, haar_sign_R<8>(pos));
}

// Halo point h (0..67) of the 18 x 18 window: window coordinates.
__device__ __forceinline__ void halo_xy(int h, int &x, int &y)
{
    if (h < 18) {
        x = h;
        y = 0;
    } else if (h < 36) {
        x = h - 18;
        y = 17;
    } else if (h < 52) {
        x = 0;
        y = h - 35;
    } else {
        x = 17;
        y = h - 51;
    }
}

// Sum over a 16-lane row in FP32, left in every lane of the row.
__device__ __forceinline__ float row_sum16f(float v)
{
    v = v + dpp_f<DPP_XOR1>(v);
    v = v + dpp_f<DPP_XOR2>(v);
    v = v + dpp_f<DPP_HALF_MIRROR>(v);
    v = v + dpp_f<DPP_MIRROR>(v);
    return v;
}

struct CGMIterArgs {
    CGMLayout Lo;
    float *R;               // ring images
    float2 *Pgd;            // [2][npad]: per tile (r, r) and (w, r) partials by parity (0 past the tiles)
    CGMScalars *st;
    CGMHostFlag *hflag;     // host-mapped stop report
    const double *Pb;       // partials of ||b||^2 per u / v block
    GridDims g;
    int npad;               // partials per parity, a multiple of 1024
    float rdx, rdy;
    double tol_coarse, reltol;
    int maxiter, nblk;
};

// One iteration (INIT: the launch before the first iteration, which only
// computes w0 = A_c r0 and the rings of set 1: P r0, P s = 0, P w0, and the
// partial (r0, r0), (w0, r0)).  All loads are issued up front, unconditional
// in the padded layout, so that the memory round trip is paid once.  Seven
// barriers: partial sums; block transforms of r and s (x pass, transpose,
// y pass straight into the window); intermediates; operator and R x pass;
// R y pass and P y pass; P x pass and the new partials.
// PRE (inside the flexible CG's linear solves only): the Jacobi-
// preconditioned form with u = D^-1 r, w = A_c u, gamma = (r, u),
// delta = (w, u); D^-1 is pointwise, so u, D^-1 s and D^-1 w are formed
// where they are used and the rings hold P D^-1 r, P D^-1 s, P D^-1 w (the
// halo of P u_new is then the same linear step).
template <int PAR, bool INIT, int LEV, bool PRE>
__device__ __forceinline__ bool cgm_iter_tile(const CGMIterArgs &a, const int L)
{
    static_assert(!INIT || PAR == 0, "the initial launch reads set 0");
    struct Smem {
        float U[18][19], V[18][19];      // window of P r (u and v points)
        float HU[18][19], HV[18][19];    // thickness windows, 0 off the masks
        float rxx[17][18], ryy[17][18], ext[17][18], rxy[17][18];
        float T[4][TILE][TILE + 1];      // transposes of the block transforms
        uint8_t mb[2][TILE][TILE + 1];   // mask bits of the tile
        float dz[PRE ? 2 : 1][TILE][TILE + 1];  // Jacobi: 1 / |diag(A_c)| of the tile
        double red0[8];
        float red1[8];
        double red2[8];
    };
    __shared__ Smem sm;
    const GridDims &g = a.g;
    const int ntile = g.nbx * g.nby;
    if (L >= ntile) return false;
    const int tid = threadIdx.x;
    int bx, by;
    const int tile_id = band_tile(g, L, bx, by);
    const int i0 = bx * TILE, j0 = by * TILE;
    const int lx = tid & 15, ly = tid >> 4;    // row-major position
    const int cx = tid >> 4, cy = tid & 15;    // column-major position (transposes)
    constexpr int NP = 1 - PAR;
    // Launch q >= 1 performs iteration k = q - 1, so PAR == 0 iff k is odd.
    // x is advanced every other iteration: an even iteration leaves it, the
    // odd one after it applies both steps, x_{k+1} = x_{k-1} +
    // (alpha_{k-1} + alpha_k beta_k) p_{k-1} + alpha_k r_k, and a stop after
    // an even iteration applies the pending alpha_{k-1} p_{k-1}.
    constexpr bool XUPD = PAR == 0;
    const unsigned px = a.Lo.px, ns = a.Lo.ns;
    float *F = a.Lo.F;
    const uint8_t *M = a.Lo.M;
    const float rdx = a.rdx, rdy = a.rdy;
    // Array A is the uniform base F + A ns; a point is a 32-bit byte offset
    // into it (the scalar-base addressing of global loads and stores).
    auto arr = [&](int A) { return F + (size_t)A * ns; };
    auto el = [](float *base, unsigned boff) -> float & { return *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + boff); };
    // window entry (x, y) is arena offset wb + x + px y
    // (24-bit products: a 32-bit multiply-add is emitted as a 64-bit one whose
    // unused high addend can be a register with a load in flight)
    const unsigned wb = (unsigned)i0 + __umul24(px, (unsigned)j0);
    const unsigned o = wb + (unsigned)(lx + 1) + __umul24(px, (unsigned)(ly + 1));   // own point
    const unsigned ob = 4u * o;
    // ---- loads, all issued before the first use, in the order they are
    // needed: the wait for a load also waits for every load issued before
    // it, so the partials come first and the coefficient windows last
    const CGMState S = a.st->s[PAR];
    // partials of the previous launch
    float2 pv[4];
    if (!INIT) {
#pragma unroll
        for (int e = 0; e < 4; ++e) pv[e] = a.Pgd[PAR * a.npad + tid + 256 * e];
    }
    const uint8_t mbu = M[o], mbv = (M + ns)[o];
    float r_u = el(arr(CGM_R), ob), r_v = el(arr(CGM_R + 1), ob);
    float w_u0 = 0.f, p_u0 = 0.f, s_u0 = 0.f, x_u0 = 0.f, w_v0 = 0.f, p_v0 = 0.f, s_v0 = 0.f, x_v0 = 0.f;
    if (!INIT) {
        w_u0 = el(arr(CGM_W), ob);
        w_v0 = el(arr(CGM_W + 1), ob);
        p_u0 = el(arr(CGM_P), ob);
        p_v0 = el(arr(CGM_P + 1), ob);
        s_u0 = el(arr(CGM_S), ob);
        s_v0 = el(arr(CGM_S + 1), ob);
        if (XUPD) {
            x_u0 = el(arr(CGM_X), ob);
            x_v0 = el(arr(CGM_X + 1), ob);
        }
    }
    // halo of the window from the neighbours' rings: threads 0..67 the u
    // grid, 68..135 the v grid
    int hx, hy;
    const int hcomp = tid >= 68 ? 1 : 0;
    const int hh = tid < 136 ? tid - 68 * hcomp : 0;
    halo_xy(hh, hx, hy);
    // neighbour tile (dx, dy) and the slot in its ring
    const int hdx = hx == 0 ? -1 : (hx == 17 ? 1 : 0), hdy = hy == 0 ? -1 : (hy == 17 ? 1 : 0);
    const int hslot = hdy == -1 ? 16 + (hdx == -1 ? 15 : (hdx == 1 ? 0 : hx - 1))
                    : hdy == 1 ? (hdx == -1 ? 15 : (hdx == 1 ? 0 : hx - 1))
                    : hdx == -1 ? 48 + (hy - 1) : 32 + (hy - 1);
    const int nbxx = bx + hdx, nbyy = by + hdy;
    const int ntl = (nbxx >= 0 && nbxx < g.nbx && nbyy >= 0 && nbyy < g.nby) ? nbxx + g.nbx * nbyy : ntile;
    const float *Rr = a.R;
    const float hr = Rr[ring_at(PAR, ntile, ntl, 0, hcomp) + hslot];
    float hs = 0.f, hw = 0.f;
    if (!INIT) {
        hs = Rr[ring_at(PAR, ntile, ntl, 1, hcomp) + hslot];
        hw = Rr[ring_at(PAR, ntile, ntl, 2, hcomp) + hslot];
    }
    // thickness windows, entries q = tid and tid + 256 (< 324) of the 18 x 18 window
    const int q1 = tid + 256 < 324 ? tid + 256 : tid;
    const unsigned t0 = 4u * (wb + (unsigned)(tid % 18) + __umul24(px, (unsigned)(tid / 18)));
    const unsigned t1 = 4u * (wb + (unsigned)(q1 % 18) + __umul24(px, (unsigned)(q1 / 18)));
    const float thu0 = el(arr(CGM_HU), t0), thv0 = el(arr(CGM_HU + 1), t0);
    const float thu1 = el(arr(CGM_HU), t1), thv1 = el(arr(CGM_HU + 1), t1);
    // cell coefficients, cells q = tid and tid + 256 (< 289) of the 17 x 17 window
    const int c1q = tid + 256 < 289 ? tid + 256 : tid;
    const unsigned c0o = wb + (unsigned)(tid % 17) + __umul24(px, (unsigned)(tid / 17));
    const unsigned c1o = wb + (unsigned)(c1q % 17) + __umul24(px, (unsigned)(c1q / 17));
    const char *cell2 = reinterpret_cast<const char *>(arr(CGM_CELL));
    const float2 cc0 = *reinterpret_cast<const float2 *>(cell2 + 8u * c0o);
    const float2 cc1 = *reinterpret_cast<const float2 *>(cell2 + 8u * c1o);
    const float ch0 = el(arr(CGM_CORNER), 4u * c0o), ch1 = el(arr(CGM_CORNER), 4u * c1o);
    const float bu0 = el(arr(CGM_BU), ob), bv0 = el(arr(CGM_BU + 1), ob);
    float dz_u = 1.f, dz_v = 1.f;
    if constexpr (PRE) {
        dz_u = el(arr(CGM_DINV), ob);
        dz_v = el(arr(CGM_DINV + 1), ob);
    }
    const bool inu = mbu & 1, inv = mbv & 1, mwu = mbu & 2, mwv = mbv & 2;
    // ---- scalars from the previous launch's partials (FP64 across workgroups)
    float fa = 0.f, fb = 0.f, fxp = 0.f;
    if (!INIT) {
        double gamma = 0.0, delta = 0.0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            gamma = gamma + (double)pv[e].x;
            delta = delta + (double)pv[e].y;
        }
        for (int q = tid + 1024; q < a.npad; q += 256) {
            const float2 v = a.Pgd[PAR * a.npad + q];
            gamma = gamma + (double)v.x;
            delta = delta + (double)v.y;
        }
        wg_sum2(gamma, delta, sm.red0);
        if (S.done) {
            if (L == 0 && tid == 0) a.st->s[NP] = S;
            return true;
        }
        // stop when ||r|| <= tol, tested as ||r||^2 <= tol^2
        double tol2 = S.tol2;
        if (S.it == 0) {
            const double abstol = a.tol_coarse * sqrt(lanes_sum(a.Pb, g.npart, sm.red2));
            double tol = a.reltol * sqrt(gamma);
            if (abstol > tol) tol = abstol;
            tol2 = tol * tol;
        }
        if (S.it >= a.maxiter || gamma <= tol2) {
            if (L == 0 && tid == 0) {
                CGMState n = S;
                n.tol2 = tol2;
                n.done = 1;
                a.st->s[NP] = n;
                a.st->done = 1;
                a.st->iters = S.it;
                a.hflag->iters = S.it;
                a.hflag->done = 1;
            }
            if (XUPD && S.it > 0) {
                const float fp = (float)S.alpha;
                el(arr(CGM_X), ob) = x_u0 + fp * p_u0;
                el(arr(CGM_X + 1), ob) = x_v0 + fp * p_v0;
            }
            return true;
        }
        const double beta = S.it == 0 ? 0.0 : gamma * S.igamma;
        const double alpha = gamma / (S.it == 0 ? delta : delta - beta * gamma * S.ialpha);
        if (L == 0 && tid == 0) {
            CGMState n;
            n.igamma = 1.0 / gamma;
            n.alpha = alpha;
            n.ialpha = 1.0 / alpha;
            n.tol2 = tol2;
            n.it = S.it + 1;
            n.done = 0;
            a.st->s[NP] = n;
        }
        fa = (float)alpha;
        fb = (float)beta;
        fxp = XUPD ? (float)(S.alpha + alpha * beta) : 0.f;
    }
    // ---- vector updates on the tile (zero past the ends of the grids); with
    // Jacobi the direction and x take u = D^-1 r where they take r
    float s_u = 0.f, s_v = 0.f;
    if (!INIT) {
        const float z_u = PRE ? dz_u * r_u : r_u, z_v = PRE ? dz_v * r_v : r_v;
        const float p_u = z_u + fb * p_u0, p_v = z_v + fb * p_v0;
        if (XUPD) {
            st_nt(&el(arr(CGM_X), ob), (x_u0 + fxp * p_u0) + fa * z_u);
            st_nt(&el(arr(CGM_X + 1), ob), (x_v0 + fxp * p_v0) + fa * z_v);
        }
        s_u = w_u0 + fb * s_u0;
        s_v = w_v0 + fb * s_v0;
        r_u = r_u - fa * s_u;
        r_v = r_v - fa * s_v;
        st_nt(&el(arr(CGM_P), ob), p_u);
        st_nt(&el(arr(CGM_P + 1), ob), p_v);
        st_nt(&el(arr(CGM_S), ob), s_u);
        st_nt(&el(arr(CGM_S + 1), ob), s_v);
        st_nt(&el(arr(CGM_R), ob), r_u);
        st_nt(&el(arr(CGM_R + 1), ob), r_v);
    }
    // ---- x pass of P r and P s (Jacobi: of D^-1 r and D^-1 s); halo and
    // thickness windows into LDS
    {
        float a0 = mwu ? (PRE ? dz_u * r_u : r_u) : 0.f, a1 = mwv ? (PRE ? dz_v * r_v : r_v) : 0.f;
        haar_lines_P<LEV>(a0, a1, lx);
        sm.T[0][ly][lx] = a0;
        sm.T[1][ly][lx] = a1;
        if (!INIT) {
            float a2 = mwu ? (PRE ? dz_u * s_u : s_u) : 0.f, a3 = mwv ? (PRE ? dz_v * s_v : s_v) : 0.f;
            haar_lines_P<LEV>(a2, a3, lx);
            sm.T[2][ly][lx] = a2;
            sm.T[3][ly][lx] = a3;
        }
    }
    sm.mb[0][ly][lx] = mbu;
    sm.mb[1][ly][lx] = mbv;
    if constexpr (PRE) {
        sm.dz[0][ly][lx] = dz_u;
        sm.dz[1][ly][lx] = dz_v;
    }
    if (tid < 136) {
        const float t = INIT ? hr : hr - fa * (hw + fb * hs);
        if (!hcomp)
            sm.U[hy][hx] = t;
        else
            sm.V[hy][hx] = t;
    }
    __syncthreads();
    // ---- y pass in the column-major lanes: P r into the window, rings of P r
    // and P s; thickness windows into LDS (read after the next barrier)
    sm.HU[tid / 18][tid % 18] = thu0;
    sm.HV[tid / 18][tid % 18] = thv0;
    if (tid + 256 < 324) {
        sm.HU[q1 / 18][q1 % 18] = thu1;
        sm.HV[q1 / 18][q1 % 18] = thv1;
    }
    {
        const uint8_t m0 = sm.mb[0][cy][cx], m1 = sm.mb[1][cy][cx];
        float y0 = sm.T[0][cy][cx], y1 = sm.T[1][cy][cx];
        haar_lines_P<LEV>(y0, y1, cy);
        const float pr_u = (m0 & 1) ? y0 : 0.f, pr_v = (m1 & 1) ? y1 : 0.f;
        sm.U[cy + 1][cx + 1] = pr_u;
        sm.V[cy + 1][cx + 1] = pr_v;
        float ps_u = 0.f, ps_v = 0.f;
        if (!INIT) {
            float y2 = sm.T[2][cy][cx], y3 = sm.T[3][cy][cx];
            haar_lines_P<LEV>(y2, y3, cy);
            ps_u = (m0 & 1) ? y2 : 0.f;
            ps_v = (m1 & 1) ? y3 : 0.f;
        }
        const RingSlots rs(cx, cy);
        const unsigned rb = ring_at(NP, ntile, tile_id, 0, 0);
        rs.put(a.R, rb, pr_u);
        rs.put(a.R, rb + 64, pr_v);
        rs.put(a.R, rb + 128, ps_u);
        rs.put(a.R, rb + 192, ps_v);
    }
    __syncthreads();
    // ---- intermediates on the 17 x 17 cells and corners
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int q = tid + 256 * e;
        if (q < 289) {
            const float hh = e ? cc1.x : cc0.x, dimp = e ? cc1.y : cc0.y, hc = e ? ch1 : ch0;
            const int x = q % 17, y = q / 17;
            const float u0 = sm.U[y][x], u1 = sm.U[y][x + 1], v0 = sm.V[y][x], v1 = sm.V[y + 1][x];
            const float dudx = rdx * (u1 - u0), dvdy = rdy * (v1 - v0);
            sm.rxx[y][x] = 2.f * hh * (2.f * dudx + dvdy);
            sm.ryy[y][x] = 2.f * hh * (2.f * dvdy + dudx);
            const float divq = rdx * (sm.HU[y][x + 1] * u1 - sm.HU[y][x] * u0) +
                               rdy * (sm.HV[y + 1][x] * v1 - sm.HV[y][x] * v0);
            sm.ext[y][x] = dimp * divq;
            sm.rxy[y][x] = hc * (rdy * (sm.U[y + 1][x + 1] - sm.U[y][x + 1]) +
                                 rdx * (sm.V[y + 1][x + 1] - sm.V[y + 1][x]));
        }
    }
    __syncthreads();
    // ---- A at the tile's points (zero off the inner masks), R x pass
    {
        const int x = lx, y = ly;
        const float d_rxx = rdx * (sm.rxx[y + 1][x + 1] - sm.rxx[y + 1][x]);
        const float d_rxy_u = rdy * (sm.rxy[y + 1][x] - sm.rxy[y][x]);
        const float d_ext_u = rdx * (sm.ext[y + 1][x + 1] - sm.ext[y + 1][x]);
        const float Au = ((d_rxx + d_rxy_u) - bu0 * sm.U[y + 1][x + 1]) - sm.HU[y + 1][x + 1] * d_ext_u;
        const float d_ryy = rdy * (sm.ryy[y + 1][x + 1] - sm.ryy[y][x + 1]);
        const float d_rxy_v = rdx * (sm.rxy[y][x + 1] - sm.rxy[y][x]);
        const float d_ext_v = rdy * (sm.ext[y + 1][x + 1] - sm.ext[y][x + 1]);
        const float Av = ((d_ryy + d_rxy_v) - bv0 * sm.V[y + 1][x + 1]) - sm.HV[y + 1][x + 1] * d_ext_v;
        float a0 = inu ? Au : 0.f, a1 = inv ? Av : 0.f;
        haar_lines_R<LEV>(a0, a1, lx);
        sm.T[0][ly][lx] = a0;
        sm.T[1][ly][lx] = a1;
    }
    __syncthreads();
    // ---- R y pass, wavelet mask (w), P y pass
    {
        float c0 = sm.T[0][cy][cx], c1 = sm.T[1][cy][cx];
        haar_lines_R<LEV>(c0, c1, cy);
        const float w0 = (sm.mb[0][cy][cx] & 2) ? c0 : 0.f;
        const float w1 = (sm.mb[1][cy][cx] & 2) ? c1 : 0.f;
        float q0 = w0, q1v = w1;
        if constexpr (PRE) {
            q0 = sm.dz[0][cy][cx] * w0;
            q1v = sm.dz[1][cy][cx] * w1;
        }
        haar_lines_P<LEV>(q0, q1v, cy);
        sm.T[0][cy][cx] = w0;
        sm.T[1][cy][cx] = w1;
        sm.T[2][cy][cx] = q0;
        sm.T[3][cy][cx] = q1v;
    }
    __syncthreads();
    // ---- P x pass, stores, partials of (r, r) and (w, r)
    const float wn_u = sm.T[0][ly][lx], wn_v = sm.T[1][ly][lx];
    float tw_u = sm.T[2][ly][lx], tw_v = sm.T[3][ly][lx];
    haar_lines_P<LEV>(tw_u, tw_v, lx);
    st_nt(&el(arr(CGM_W), ob), wn_u);
    st_nt(&el(arr(CGM_W + 1), ob), wn_v);
    {
        const RingSlots rs(lx, ly);
        const unsigned rb = ring_at(NP, ntile, tile_id, 2, 0);
        rs.put(a.R, rb, inu ? tw_u : 0.f);
        rs.put(a.R, rb + 64, inv ? tw_v : 0.f);
    }
    // (r, u) and (w, u) with u = D^-1 r under Jacobi, (r, r) and (w, r) without
    const float zr_u = PRE ? dz_u * r_u : r_u, zr_v = PRE ? dz_v * r_v : r_v;
    float gg = r_u * zr_u + r_v * zr_v;
    float dd = wn_u * zr_u + wn_v * zr_v;
    gg = wave_total_rows_f(row_sum16f(gg));
    dd = wave_total_rows_f(row_sum16f(dd));
    if ((tid & 63) == 63) {
        sm.red1[tid >> 6] = gg;
        sm.red1[4 + (tid >> 6)] = dd;
    }
    __syncthreads();
    if (tid == 0)
        a.Pgd[NP * a.npad + tile_id] = make_float2((sm.red1[0] + sm.red1[1]) + (sm.red1[2] + sm.red1[3]),
                                             (sm.red1[4] + sm.red1[5]) + (sm.red1[6] + sm.red1[7]));
    return false;
}

// One coarse-CG iteration of one CG problem (the tile of workgroup
// blockIdx.x, XCD-aware).
template <int PAR, bool INIT, int LEV, bool PRE>
__global__ void __launch_bounds__(256) k_cgm_iter(CGMIterArgs a)
{
    (void)cgm_iter_tile<PAR, INIT, LEV, PRE>(a, xcd_tile(a.nblk));
}

// The whole mixed coarse CG of a small grid in one workgroup (after
// k_cgm_setup): the initial launch's work and then each iteration on every
// tile in turn, the barrier between two tiles standing in for the kernel
// boundary.  A tile reads the other tiles' rings and partials of the
// previous iteration only, so the results are those of one launch per
// iteration, bit for bit.  Every thread decides the stop alike (same
// partials, same scalars); at most maxiter + 1 iterations.
template <int LEV, bool PRE>
__global__ void __launch_bounds__(256) k_cgm_solve1(CGMIterArgs a)
{
    const int ntile = a.g.nbx * a.g.nby;
    for (int L = 0; L < ntile; ++L) {
        (void)cgm_iter_tile<0, true, LEV, PRE>(a, L);
        __syncthreads();
    }
    for (int q = 1; q <= a.maxiter + 1; ++q) {
        bool stop = false;
        for (int L = 0; L < ntile; ++L) {
            stop = (q & 1) ? cgm_iter_tile<1, false, LEV, PRE>(a, L) : cgm_iter_tile<0, false, LEV, PRE>(a, L);
            __syncthreads();
        }
        if (stop) break;
    }
}

// The same iteration of several independent CG problems (an ensemble's
// members, each with its own arena, rings, partials and scalars) in one
// launch.  Member m owns the workgroups first[m] .. first[m] + nblk_m - 1;
// every first[m] and nblk_m is a multiple of 8, so group_member, indexed by
// blockIdx.x / 8, names the member, and a member's workgroups map to its
// tiles exactly as k_cgm_iter's do (blockIdx.x mod 8 is preserved).
struct CGMBatch {
    const CGMIterArgs *members;
    const int *group_member;
    const int *first;
};

template <int PAR, bool INIT, int LEV, bool PRE>
__global__ void __launch_bounds__(256) k_cgm_iter_batch(CGMBatch B)
{
    const int m = B.group_member[blockIdx.x >> 3];
    const CGMIterArgs a = B.members[m];
    const int w = (int)blockIdx.x - B.first[m];
    (void)cgm_iter_tile<PAR, INIT, LEV, PRE>(a, (w & 7) * (a.nblk >> 3) + (w >> 3));
}

// out = mask_out ? H(mask_in ? in : 0) : 0 on the u grid (comp 0) and v grid
// mask[k] != 0 at a point ok of a tile, true there without a mask (mask ==
// nullptr), false where !ok.  The load is unconditional (a branch on the
// pointer would give it a wait of its own): without a mask it reads byte 0
// of `any`, a valid buffer, and ignores it.
__device__ __forceinline__ bool mask_or_all(const uint8_t *mask, const void *any, bool ok, int k)
{
    const bool has = mask != nullptr;
    const uint8_t *p = has ? mask : static_cast<const uint8_t *>(any);
    const uint8_t b = p[(ok & has) ? k : 0];
    return ok & (!has | (b != 0));  // no short circuit, which would branch around the load
}

// Haar transform (dir 0 forward, 1 inverse, 2 transposed inverse) of the
// masked u and v inputs into the masked outputs.  With accu != nullptr (a
// +-1 transform only) the result is added into accu, accv at the points of
// the output masks instead (x += P y, as an add of the masked output would
// do), and outu, outv are not written.
template <typename Tin, typename Tout, bool LD = false>
__global__ void k_tile_haar(GridDims g, const HaarTable *tab, int levels, int dir,
                            const Tin *inu, const Tin *inv, const uint8_t *min_u, const uint8_t *min_v,
                            Tout *outu, Tout *outv, const uint8_t *mout_u, const uint8_t *mout_v,
                            Tout *accu = nullptr, Tout *accv = nullptr)
{
    __shared__ Tout tile[TILE][TILE + 1];
    __shared__ Tout tmp[TILE][TILE + 1];
    __shared__ Tout tile_v[TILE][TILE + 1];
    __shared__ Tout tmp_v[TILE][TILE + 1];
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int i0 = bx * TILE, j0 = by * TILE;
    const int i = i0 + lx, j = j0 + ly;
    const bool has_u = bx < g.nbxu && by < g.nbyu, has_v = bx < g.nbxv && by < g.nbyv;
    if (dir != 0) {
        // the +-1 transforms: the u and the v tile together, every input,
        // mask and row mask loaded first (ld_or, haar_row_masks)
        const int ku = i + g.nxu * j, kv = i + g.nxv * j;
        const bool oku = has_u && i < g.nxu && j < g.nyu, okv = has_v && i < g.nxv && j < g.nyv;
        const Tin xu = ld_or(oku, inu, ku, Tin(0)), xv = ld_or(okv, inv, kv, Tin(0));
        const bool iu = mask_or_all(min_u, inu, oku, ku), iv = mask_or_all(min_v, inv, okv, kv);
        const bool ou = mask_or_all(mout_u, inu, oku, ku), ov = mask_or_all(mout_v, inv, okv, kv);
        const HaarRowMasks hm = haar_row_masks(levels, dir - 1, i0, j0, g, has_u, has_v);
        // the accumulation target's values, loaded with the inputs (a uniform
        // branch: their first use is after the transform)
        const bool acc = accu != nullptr;
        Tout au = Tout(0), av = Tout(0);
        if (acc) {
            au = ld_or(oku, accu, ku, Tout(0));
            av = ld_or(okv, accv, kv, Tout(0));
        }
        tile[ly][lx] = (oku && iu) ? (Tout)xu : Tout(0);
        tile_v[ly][lx] = (okv && iv) ? (Tout)xv : Tout(0);
        __syncthreads();
        haar_pm_tile2m<Tout, LD>(levels, hm, tile, tmp, tile_v, tmp_v, i0, j0, g, has_u, has_v);
        if (acc) {
            if (oku && ou) accu[ku] = au + tile[ly][lx];
            if (okv && ov) accv[kv] = av + tile_v[ly][lx];
        } else {
            if (oku) outu[ku] = ou ? tile[ly][lx] : Tout(0);
            if (okv) outv[kv] = ov ? tile_v[ly][lx] : Tout(0);
        }
        return;
    }
    if (has_u) {
        int k = i + g.nxu * j;
        bool ok = i < g.nxu && j < g.nyu;
        tile[ly][lx] = (ok && (!min_u || min_u[k])) ? (Tout)inu[k] : Tout(0);
        __syncthreads();
        haar_tile<Tout>(tab, 0, tile, tmp, i0, j0, g.nxu, g.nyu);
        if (ok) outu[k] = (!mout_u || mout_u[k]) ? tile[ly][lx] : Tout(0);
        __syncthreads();
    }
    if (has_v) {
        int k = i + g.nxv * j;
        bool ok = i < g.nxv && j < g.nyv;
        tile[ly][lx] = (ok && (!min_v || min_v[k])) ? (Tout)inv[k] : Tout(0);
        __syncthreads();
        haar_tile<Tout>(tab, 0, tile, tmp, i0, j0, g.nxv, g.nyv);
        if (ok) outv[k] = (!mout_v || mout_v[k]) ? tile[ly][lx] : Tout(0);
        __syncthreads();
    }
}

// update_velocities_on_h_grid! (update_state.jl:164-178) and update_dhdt!
// (:184-189)
__global__ void k_post_solve(GridDims g, const uint8_t *hm, const uint8_t *um, const uint8_t *vm,
                             const double *U, const double *V, const double *beta, const double *f1,
                             const double *f2, const double *hu, const double *hv, const double *acc,
                             const double *mel, double *uh, double *vh, double *ubo, double *vbo,
                             double *uso, double *vso, double *dhdt, double rdx, double rdy)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nx * g.ny) return;
    const int i = k % g.nx, j = k / g.nx;
    const int nxu = g.nxu, nx = g.nx;
    // every load issued before the first store (which the compiler cannot
    // move them past) and before the mask is tested
    const int u1 = (i + 1) + nxu * j, u0 = i + nxu * j, v1 = i + nx * (j + 1), v0 = i + nx * j;
    const double U0 = U[u0], U1 = U[u1], V0 = V[v0], V1 = V[v1];
    const double bk = beta[k], f1k = f1[k], f2k = f2[k], acck = acc[k], melk = mel[k];
    const double hu0 = hu[u0], hu1 = hu[u1], hv0 = hv[v0], hv1 = hv[v1];
    const bool hk = hm[k], mu0 = um[u0], mu1 = um[u1], mv0 = vm[v0], mv1 = vm[v1];
    double u_ = 0.5 * U0 + 0.5 * U1;
    double v_ = 0.5 * V0 + 0.5 * V1;
    uh[k] = u_;
    vh[k] = v_;
    double ub = u_ / (1.0 + bk * f2k);
    double vb = v_ / (1.0 + bk * f2k);
    ubo[k] = ub;
    vbo[k] = vb;
    uso[k] = ub * (1.0 + bk * f1k);
    vso[k] = vb * (1.0 + bk * f1k);
    if (hk) {
        double q1 = mu1 ? hu1 * U1 : 0.0;
        double q0 = mu0 ? hu0 * U0 : 0.0;
        double p1 = mv1 ? hv1 * V1 : 0.0;
        double p0 = mv0 ? hv0 * V0 : 0.0;
        double div = (rdx * q1 - rdx * q0) + (rdy * p1 - rdy * p0);
        dhdt[k] = (acck - melk) - div;
    }
}

// update_thickness! (run_simulation.jl:24-42), evolveShelves = true
__global__ void k_update_thickness(int nh, const uint8_t *hm, const uint8_t *hfix, double *h, const double *dhdt,
                                   double hmin, double dt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh || !hm[k]) return;
    double a = hmin - h[k], b = dt * dhdt[k];
    double upd = a > b ? a : b;
    if (hfix[k]) upd = 0.0;
    h[k] = h[k] + upd;
}

// quad_f2 = h/(3 eta_av) on the h mask (src/Fields/HGrid.jl:229)
__global__ void k_init_quad_f2(int nh, const uint8_t *hm, const double *h, const double *etaav, double *f2)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nh && hm[k]) f2[k] = h[k] / (3.0 * etaav[k]);
}

// device math probe: fn 0 det_pow(x, y), 1 sqrt(x), 2 x / y
__global__ void k_math_probe(int fn, const double *x, const double *y, double *out, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (fn == 0) out[k] = det_pow(x[k], y[k]);
    else if (fn == 1) out[k] = sqrt(x[k]);
    else out[k] = x[k] / y[k];
}

}  // namespace wavi
