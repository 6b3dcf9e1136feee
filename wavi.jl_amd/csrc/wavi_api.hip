// wavi_api.hip - C ABI of libwavi_hip.so (include/wavi_hip.h): context,
// device-resident fields and the orchestration of update_velocities!.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/wavi_hip.h"
// all kernels of the library live in this translation unit
#include "wavi_kernels.hip"

using namespace wavi;

static thread_local std::string g_create_error;

// Serialises HIP stream captures with the library's device allocations,
// frees and synchronous copies, which other threads may make while a context
// captures (a batched call, or a caller creating or filling contexts on
// another thread): such a call can end another thread's capture.
static std::mutex g_graph_mu;

// The scratch of wavi_update_velocities_lockstep_noalloc: the device tables
// of its shared coarse-CG launches, in a fixed buffer that every GPU the
// library runs on holds until the process ends (each GPU has its own copy at
// its own address).  In host code `scratch` is only a stand-in for the
// symbol: its device address on each GPU is looked up once, at the first
// context creation on that GPU, and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[WAVI_NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
// LDS the whole-grid smoother may take (at most the device's per-workgroup limit)
constexpr size_t SMG_LDS_MAX = 65536;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[MAX_DEVICES] = {};

struct DevVec {
    double *u = nullptr, *v = nullptr;
};

template <typename T>
struct DevVecT {
    T *u = nullptr, *v = nullptr;
};

// Buffers of one V-cycle precision: operator coefficients, smoother and CG
// vectors, coarse warm start.  The FP64 set aliases the context's fields;
// the FP32 set (mixed precision) has its own copies.
template <typename T>
struct VSet {
    T *heta_h = nullptr, *heta_c = nullptr, *dimp = nullptr;
    DevVecT<T> h, beff, e, r, d, rho, t, cr, cp, cq, ct, xc;
};

struct wavi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    GridDims g{};
    wavi_phys_params P{};
    wavi_solver_params S{};
    int precision = 0;
    int nh = 0, nu = 0, nv = 0, nc = 0, ns3 = 0;
    double rdx = 0, rdy = 0;
    // masks (uint8, device)
    uint8_t *hm = nullptr, *hfix = nullptr, *um = nullptr, *vm = nullptr, *cm = nullptr;
    uint8_t *ui = nullptr, *vi = nullptr, *ufix = nullptr, *vfix = nullptr, *mwu = nullptr, *mwv = nullptr;
    bool has_fixed = false;     // any u_isfixed / v_isfixed point
    std::vector<uint8_t> h_ui, h_vi;  // host copies of the inner masks
    double *adj_base = nullptr;       // adjoint workspace, kept between gradient calls
    size_t adj_cap = 0;               // its length in doubles
    struct CgRendezvous *rdv = nullptr;  // set while a batch runs this context's coarse CGs with others'
    int rdv_index = -1;               // the context's index in that rendezvous
    bool rdv_done = false;            // its current coarse CG has been run
    int rdv_rc = 0, rdv_iters = 0;    // that CG's return code and iteration count
    // fields (double, device); mask fields use mwu/mwv instead
    double *f[WAVI_NFIELDS] = {};
    // sigma-level constants
    double *zeta = nullptr, *qw = nullptr;
    // rheological coefficients
    double *heta_h = nullptr, *heta_c = nullptr, *dimp = nullptr;
    // Picard / V-cycle vectors (full u/v grids)
    DevVec x, b, r, e, d, rho, t, f3, fixedv;
    // coarse CG vectors
    DevVec cr, cp, cq, ct;
    VSet<float> v32;
    // mixed-precision coarse CG: s, w, the parity-buffered images P r, P s,
    // P w and partials of (r, r) and (w, r)
    // mixed coarse CG: one FP32 arena (enum CGMArray) and a mask-byte arena
    // in the padded layout of pitch px and array stride ns
    float *arena = nullptr;
    uint8_t *marena = nullptr;
    unsigned px = 0, ns = 0;
    float2 *Pgd = nullptr;  // [2][npad] (r, r) and (w, r) partials by parity
    float *rings = nullptr; // ring images [2][tiles + 1][3][2][64]
    int npad = 0;
    CGMScalars *stm = nullptr;
    CGMHostFlag *h_flag = nullptr;  // pinned, host-mapped stop report of the mixed coarse CG
    CGMHostFlag *d_flag = nullptr;  // its device address
    hipEvent_t cg_ev[3] = {};       // chunk completion events of the mixed coarse CG
    hipEvent_t ev_ls = nullptr, ev_ls_done = nullptr;  // set-up and batch end in a lockstep batch without allocation
    hipEvent_t ev_tune[2] = {};     // timing events of the smoother's tile-shape tuning
    int smooth_shape[2] = {-1, -1}; // smoother variant by V-cycle precision (0 FP64, 1 FP32); -1: tune
    size_t lds_max = 0;             // dynamic LDS a workgroup may take (the whole-grid smoother)
    int cg_mode = WAVI_CG_AUTO;      // coarse CG: one launch per iteration or one workgroup (wavi_set_coarse_cg)
    int cg_prev = 0;                 // iterations of the context's previous coarse CG (0: none yet)
    hipGraphExec_t cgm_chunk = nullptr;  // 32 k_cgm_iter launches (parities 1, 0, ...) as one graph
    double *Pr0 = nullptr, *Pr1 = nullptr, *Pd = nullptr, *Pb = nullptr, *Pn = nullptr;
    double *Pz = nullptr;           // partials of (r, D^-1 r) of the Jacobi-preconditioned coarse CG
    DevVec cdinv;                   // 1 / |diag(A_c)| for the Jacobi-preconditioned coarse CG
    bool cg_pre = false;            // the V-cycles of do_linear_solve run the Jacobi-preconditioned coarse CG
    bool lin_jacobi = true;         // ... when set (wavi_set_linear_coarse_jacobi)
    double *quad = nullptr;  // 4 * nh pos_fraction quadrants
    double *h_norms = nullptr; // pinned, host-mapped: 2 doubles the finishing kernels write
    double *d_norms = nullptr; // its device address
    CGScalars *st = nullptr;
    CGScalars *h_st = nullptr;  // pinned host copy
    int nblk = 0;               // workgroups of the XCD-ordered tile kernels
    HaarTable *tab = nullptr;
    // saved state
    std::vector<double *> saved;
    bool has_saved = false;
    uint8_t *saved_mwu = nullptr, *saved_mwv = nullptr;
    // kernel timing (wavi_profile_*)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;  // (kernel id, first event index)
    double prof_ms[WAVI_KT_COUNT] = {};
    int64_t prof_n[WAVI_KT_COUNT] = {};
    int64_t prof_seen[WAVI_KT_COUNT] = {};  // launches seen; every 16th is timed
    std::vector<hipEvent_t> burst_ev;        // event pairs around CG chunks
    struct Burst {
        int ev, last, n;  // first event index, index of the chunk's last launch, launches
    };
    std::vector<Burst> bursts;
    std::string err;
};

// Timed launch: when profiling is on, record events around the launch.
struct ProfScope {
    wavi_ctx *ctx;
    int kid;
    int idx = -1;
    ProfScope(wavi_ctx *c, int k, bool on = true) : ctx(c), kid(k)
    {
        if (!ctx->prof || !on) return;
        if ((ctx->prof_seen[kid]++ & 15) != 0) return;
        size_t need = (ctx->ev_pending.size() + 1) * 2;
        while (ctx->ev_pool.size() < need) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            ctx->ev_pool.push_back(e);
        }
        idx = (int)ctx->ev_pending.size() * 2;
        (void)hipEventRecord(ctx->ev_pool[idx], ctx->stream);
    }
    ~ProfScope()
    {
        if (idx < 0) return;
        (void)hipEventRecord(ctx->ev_pool[idx + 1], ctx->stream);
        ctx->ev_pending.push_back({kid, idx});
    }
};

// After a stream synchronisation: fold the recorded event pairs.
static void prof_collect(wavi_ctx *ctx)
{
    if (ctx->ev_pending.empty()) return;
    for (auto &p : ctx->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev_pool[p.second], ctx->ev_pool[p.second + 1]) == hipSuccess) {
            ctx->prof_ms[p.first] += ms;
            ctx->prof_n[p.first] += 1;
        }
    }
    ctx->ev_pending.clear();
}

// Burst timing of a chunk of back-to-back CG launches: burst_begin records
// the opening event (returns its index, or -1 when not profiling or out of
// events), burst_end the closing one.
static int burst_begin(wavi_ctx *ctx)
{
    if (!ctx->prof) return -1;
    const size_t need = (ctx->bursts.size() + 1) * 2;
    while (ctx->burst_ev.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->burst_ev.push_back(e);
    }
    const int idx = (int)ctx->bursts.size() * 2;
    if (hipEventRecord(ctx->burst_ev[idx], ctx->stream) != hipSuccess) return -1;
    return idx;
}

static void burst_end(wavi_ctx *ctx, int idx, int last, int n)
{
    if (idx < 0) return;
    if (hipEventRecord(ctx->burst_ev[idx + 1], ctx->stream) != hipSuccess) return;
    ctx->bursts.push_back({idx, last, n});
}

// After the stream synchronisation that ends a coarse CG: fold the bursts
// whose launches all iterated (last <= last_active) into kernel id kid.
static void burst_collect(wavi_ctx *ctx, int kid, int last_active)
{
    for (const wavi_ctx::Burst &b : ctx->bursts) {
        if (b.last > last_active) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->burst_ev[b.ev], ctx->burst_ev[b.ev + 1]) == hipSuccess) {
            ctx->prof_ms[kid] += ms;
            ctx->prof_n[kid] += b.n;
        }
    }
    ctx->bursts.clear();
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return WAVI_EHIP;                                                               \
        }                                                                                   \
    } while (0)

#define CHK(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != WAVI_OK) return rc_; \
    } while (0)

static inline dim3 flat(int n) { return dim3((unsigned)((n + 255) / 256)); }
// workgroups of a launch over an nbx x nby tile grid with xcd_band_tile: a multiple of 8
static inline int band_blocks(int nbx, int nby) { return (nbx * nby + 7) / 8 * 8; }
// grid of a u / v pair launch: blockIdx.y 0 for the u arrays, 1 for the v arrays
static inline dim3 flat_uv(int nu, int nv) { return dim3((unsigned)(((nu > nv ? nu : nv) + 255) / 256), 2); }

static int field_len(const wavi_ctx *ctx, int id)
{
    if (id < 0 || id >= WAVI_NFIELDS) return -1;
    if (id <= WAVI_P_ACCUMULATION_RATE) return ctx->nh;
    if (id <= WAVI_U_RESIDUAL) return ctx->nu;
    if (id <= WAVI_V_RESIDUAL) return ctx->nv;
    if (id <= WAVI_S_GLEN_B) return ctx->ns3;
    if (id <= WAVI_WU_CORRECTION_COARSE) return ctx->nu;
    return ctx->nv;
}

static OpCoef<double> coef64(const wavi_ctx *ctx)
{
    OpCoef<double> c;
    c.heta_h = ctx->heta_h;
    c.heta_c = ctx->heta_c;
    c.dimp = ctx->dimp;
    c.hu = ctx->f[WAVI_U_H];
    c.hv = ctx->f[WAVI_V_H];
    c.bu = ctx->f[WAVI_U_BETA_EFF];
    c.bv = ctx->f[WAVI_V_BETA_EFF];
    c.hm = ctx->hm;
    c.um = ctx->um;
    c.vm = ctx->vm;
    c.cm = ctx->cm;
    c.ui = ctx->ui;
    c.vi = ctx->vi;
    c.nx = ctx->g.nx;
    c.ny = ctx->g.ny;
    c.rdx = ctx->rdx;
    c.rdy = ctx->rdy;
    return c;
}

// ---------------------------------------------------------------------------
// Haar tables from wavelet_matrix (src/Wavelets/update_wavelets.jl:38-89).
static void wavelet_matrix(int n, int levels, bool forward, std::vector<double> &W)
{
    W.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) W[(size_t)i * n + i] = 1.0;
    int step = 1;
    std::vector<double> tmp(n);
    for (int jl = levels; jl >= 0; --jl) {
        step *= 2;
        int h = step / 2;
        for (int o = 0; o + 1 <= n - h; o += step) {
            int e = o + h;
            if (forward) {
                for (int c = 0; c < n; ++c) {
                    double ro = W[(size_t)o * n + c], re = W[(size_t)e * n + c];
                    tmp[c] = 0.5 * (ro + re);
                    W[(size_t)e * n + c] = 0.5 * (re - ro);
                }
                for (int c = 0; c < n; ++c) W[(size_t)o * n + c] = tmp[c];
            } else {
                for (int r = 0; r < n; ++r) {
                    double co = W[(size_t)r * n + o], ce = W[(size_t)r * n + e];
                    W[(size_t)r * n + o] = ce + co;
                    W[(size_t)r * n + e] = ce - co;
                }
            }
        }
    }
}

static int build_haar_table(int levels, HaarTable &t, std::string &err)
{
    memset(&t, 0, sizeof(t));
    t.B = 1 << (levels + 1);
    for (int m = 1; m <= t.B; ++m) {
        std::vector<double> Wf, Wi;
        wavelet_matrix(m, levels, true, Wf);
        wavelet_matrix(m, levels, false, Wi);
        for (int dir = 0; dir < 3; ++dir)
            for (int r = 0; r < m; ++r) {
                int n = 0;
                for (int c = 0; c < m; ++c) {
                    double w = dir == 0 ? Wf[(size_t)r * m + c] : dir == 1 ? Wi[(size_t)r * m + c] : Wi[(size_t)c * m + r];
                    if (w != 0.0) {
                        if (n >= MAX_HAAR_NNZ) {
                            err = "Haar table overflow";
                            return WAVI_EARG;
                        }
                        t.col[dir][m][r][n] = (int8_t)c;
                        t.coef[dir][m][r][n] = (float)w;
                        if ((double)t.coef[dir][m][r][n] != w) {
                            err = "Haar coefficient not exact in float";
                            return WAVI_EARG;
                        }
                        ++n;
                    }
                }
                t.nnz[dir][m][r] = (int8_t)n;
            }
    }
    return WAVI_OK;
}

// +-1 Haar rows as bit masks for levels 0..3 (inverse, transposed inverse)
static void build_haar_masks(HaarMasks tabs[4])
{
    memset(tabs, 0, 4 * sizeof(HaarMasks));
    for (int lev = 0; lev < 4; ++lev) {
        const int B = 2 << lev;
        for (int m = 1; m <= B; ++m) {
            std::vector<double> Wi;
            wavelet_matrix(m, lev, false, Wi);
            for (int dir = 0; dir < 2; ++dir)
                for (int r = 0; r < m; ++r) {
                    uint16_t nz = 0, neg = 0;
                    for (int c = 0; c < m; ++c) {
                        double w = dir == 0 ? Wi[(size_t)r * m + c] : Wi[(size_t)c * m + r];
                        if (w != 0.0) {
                            nz |= (uint16_t)(1u << c);
                            if (w < 0.0) neg |= (uint16_t)(1u << c);
                        }
                    }
                    tabs[lev].nz[dir][m][r] = nz;
                    tabs[lev].neg[dir][m][r] = neg;
                }
        }
    }
}

// Run fn(ctx, k) for members k < n on a pool of nthreads host threads (the
// calling thread included), each first making the member's device current.
// codes (NULL or n entries) receives every member's return code; the
// result is the first nonzero one in member order.
// A context may appear only once in a batch: its members run concurrently.
static bool batch_has_duplicates(wavi_ctx **ctxs, int n)
{
    std::vector<wavi_ctx *> v(ctxs, ctxs + n);
    std::sort(v.begin(), v.end());
    for (int k = 1; k < n; ++k)
        if (v[k] == v[k - 1]) {
            v[k]->err = g_create_error = "the same context appears twice in the batch";
            return true;
        }
    return false;
}

template <typename F>
static int run_members(wavi_ctx **ctxs, int n, int nthreads, int *codes, F &&fn)
{
    std::vector<int> rcs(n, WAVI_OK);
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (int k = next.fetch_add(1); k < n; k = next.fetch_add(1)) {
            wavi_ctx *ctx = ctxs[k];
            if (hipSetDevice(ctx->device) != hipSuccess) {
                ctx->err = "hipSetDevice failed";
                rcs[k] = WAVI_EHIP;
                continue;
            }
            try {
                rcs[k] = fn(ctx, k);
            } catch (const std::exception &ex) {
                ctx->err = std::string("host exception: ") + ex.what();
                rcs[k] = WAVI_ESTATE;
            }
        }
    };
    const int nt = std::min(nthreads, std::max(n, 1));
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    } catch (const std::system_error &) {
        // fewer threads: the members left are taken by those running
    }
    worker();
    for (std::thread &t : pool) t.join();
    if (codes)
        for (int k = 0; k < n; ++k) codes[k] = rcs[k];
    for (int k = 0; k < n; ++k)
        if (rcs[k] != WAVI_OK) return rcs[k];
    return WAVI_OK;
}

// ---------------------------------------------------------------------------
extern "C" {
#pragma GCC visibility push(default)

const char *wavi_last_error(const wavi_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

static void ctx_free(wavi_ctx *ctx);

int wavi_ctx_create(const wavi_grid_desc *grid, const wavi_phys_params *phys, const wavi_solver_params *solver,
                    int precision_mode, wavi_ctx **out)
{
    g_create_error.clear();
    if (!grid || !phys || !solver || !out) {
        g_create_error = "null argument";
        return WAVI_EARG;
    }
    *out = nullptr;
    if (grid->nx <= 1 || grid->ny <= 1 || grid->nsigma <= 1) {
        g_create_error = "nx, ny and nsigma must be integers larger than one";
        return WAVI_EARG;
    }
    if (!grid->h_mask || !grid->sigma || !grid->quadrature_weights) {
        g_create_error = "h_mask, sigma and quadrature_weights are required";
        return WAVI_EARG;
    }
    if (!(grid->dx > 0) || !(grid->dy > 0)) {
        g_create_error = "dx and dy must be positive";
        return WAVI_EARG;
    }
    if (solver->n_iter_viscosity != 2) {
        g_create_error = "n_iter_viscosity must be 2";
        return WAVI_EARG;
    }
    if (solver->levels < 0 || solver->levels > 3) {
        g_create_error = "levels must be in 0..3 (Haar blocks of at most 16)";
        return WAVI_EARG;
    }
    if (solver->nsmooth < 0 || solver->maxiter_picard < 0 || solver->maxiter_coarse < 0) {
        g_create_error = "iteration counts must be non-negative";
        return WAVI_EARG;
    }
    if (solver->maxiter_coarse > (1 << 30)) {
        g_create_error = "maxiter_coarse must be at most 2^30";
        return WAVI_EARG;
    }
    if (precision_mode != WAVI_PRECISION_FP64 && precision_mode != WAVI_PRECISION_MIXED) {
        g_create_error = "unknown precision mode";
        return WAVI_EARG;
    }
    if (!phys->evolve_shelves) {
        g_create_error = "evolve_shelves = false is not supported";
        return WAVI_EARG;
    }
    {
        // The mixed coarse CG addresses its arena with 32-bit byte offsets
        // and 24-bit row products: (nx + 1) and (ny + 1) padded to whole
        // 16-point tiles must give fewer than 2^28 points per array.
        const size_t tx = ((size_t)grid->nx + 1 + TILE - 1) / TILE, ty = ((size_t)grid->ny + 1 + TILE - 1) / TILE;
        const size_t pitch = (TILE * tx + 2 + 31) / 32 * 32;
        if (pitch * (TILE * ty + 2) >= ((size_t)1 << 28)) {
            g_create_error = "grid too large: at most 2^28 padded points per array";
            return WAVI_EDIM;
        }
    }
    std::lock_guard<std::mutex> lock(g_graph_mu);
    wavi_ctx *ctx = new wavi_ctx();
    ctx->P = *phys;
    ctx->S = *solver;
    ctx->precision = precision_mode;
    const int nx = grid->nx, ny = grid->ny;
    GridDims &g = ctx->g;
    g.nx = nx;
    g.ny = ny;
    g.nxu = nx + 1;
    g.nyu = ny;
    g.nxv = nx;
    g.nyv = ny + 1;
    g.nbxu = (g.nxu + TILE - 1) / TILE;
    g.nbyu = (g.nyu + TILE - 1) / TILE;
    g.nbxv = (g.nxv + TILE - 1) / TILE;
    g.nbyv = (g.nyv + TILE - 1) / TILE;
    g.nbx = g.nbxu > g.nbxv ? g.nbxu : g.nbxv;
    g.nby = g.nbyu > g.nbyv ? g.nbyu : g.nbyv;
    g.npart = g.nbxu * g.nbyu + g.nbxv * g.nbyv;
    ctx->nh = nx * ny;
    ctx->nu = (nx + 1) * ny;
    ctx->nv = nx * (ny + 1);
    ctx->nc = (nx - 1) * (ny - 1);
    ctx->ns3 = nx * ny * grid->nsigma;
    ctx->rdx = 1.0 / grid->dx;
    ctx->rdy = 1.0 / grid->dy;
    int ns = grid->nsigma;

    auto fail = [&](int rc) {
        g_create_error = ctx->err;
        ctx_free(ctx);
        return rc;
    };
#define CREATECHK(expr)                                                           \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return fail(WAVI_EHIP);                                               \
        }                                                                         \
    } while (0)
    CREATECHK(hipGetDevice(&ctx->device));
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && lds > 0)
            ctx->lds_max = std::min((size_t)lds, SMG_LDS_MAX);
    }
    if (ctx->device >= 0 && ctx->device < MAX_DEVICES) {
        // the scratch address of the entry point without allocation, once per device
        std::lock_guard<std::mutex> slock(g_scratch_mu);
        if (!g_scratch_ptr[ctx->device]) {
            void *p = nullptr;
            CREATECHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_ptr[ctx->device] = (unsigned char *)p;
        }
    }
    CREATECHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));

    // host masks (setup_fields, src/Fields/Fields.jl:28-35)
    std::vector<uint8_t> hm(ctx->nh), hfix(ctx->nh, 0), um(ctx->nu), vm(ctx->nv), cm(ctx->nc > 0 ? ctx->nc : 1, 0);
    std::vector<uint8_t> ui(ctx->nu), vi(ctx->nv), ufix(ctx->nu, 0), vfix(ctx->nv, 0), ones_u(ctx->nu, 1), ones_v(ctx->nv, 1);
    for (int k = 0; k < ctx->nh; ++k) {
        hm[k] = grid->h_mask[k] ? 1 : 0;
        hfix[k] = grid->h_isfixed && grid->h_isfixed[k] ? 1 : 0;
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int k = i + (nx + 1) * j;
            int m = (i < nx && hm[i + nx * j]) || (i > 0 && hm[(i - 1) + nx * j]);
            if (grid->u_iszero && grid->u_iszero[k]) m = 0;
            um[k] = (uint8_t)m;
            ufix[k] = grid->u_isfixed && grid->u_isfixed[k] ? 1 : 0;
            ui[k] = (uint8_t)(m && !ufix[k]);
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int k = i + nx * j;
            int m = (j < ny && hm[i + nx * j]) || (j > 0 && hm[i + nx * (j - 1)]);
            if (grid->v_iszero && grid->v_iszero[k]) m = 0;
            vm[k] = (uint8_t)m;
            vfix[k] = grid->v_isfixed && grid->v_isfixed[k] ? 1 : 0;
            vi[k] = (uint8_t)(m && !vfix[k]);
        }
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i)
            cm[i + (nx - 1) * j] = hm[i + nx * j] && hm[i + nx * (j + 1)] && hm[(i + 1) + nx * j] && hm[(i + 1) + nx * (j + 1)];

    auto alloc8 = [&](uint8_t **p, const std::vector<uint8_t> &src) -> hipError_t {
        hipError_t e = hipMalloc(p, src.size());
        if (e != hipSuccess) return e;
        return hipMemcpy(*p, src.data(), src.size(), hipMemcpyHostToDevice);
    };
    ctx->h_ui = ui;
    ctx->h_vi = vi;
    for (uint8_t f : ufix) ctx->has_fixed = ctx->has_fixed || f;
    for (uint8_t f : vfix) ctx->has_fixed = ctx->has_fixed || f;
    CREATECHK(alloc8(&ctx->hm, hm));
    CREATECHK(alloc8(&ctx->hfix, hfix));
    CREATECHK(alloc8(&ctx->um, um));
    CREATECHK(alloc8(&ctx->vm, vm));
    CREATECHK(alloc8(&ctx->cm, cm));
    CREATECHK(alloc8(&ctx->ui, ui));
    CREATECHK(alloc8(&ctx->vi, vi));
    CREATECHK(alloc8(&ctx->ufix, ufix));
    CREATECHK(alloc8(&ctx->vfix, vfix));
    CREATECHK(alloc8(&ctx->mwu, ones_u));  // UWavelets mask = trues (UWavelets.jl:38)
    CREATECHK(alloc8(&ctx->mwv, ones_v));

    for (int id = 0; id < WAVI_NFIELDS; ++id) {
        if (id == WAVI_WU_MASK || id == WAVI_WV_MASK) continue;
        size_t n = (size_t)field_len(ctx, id);
        CREATECHK(hipMalloc(&ctx->f[id], n * sizeof(double)));
        CREATECHK(hipMemset(ctx->f[id], 0, n * sizeof(double)));
    }
    // constructor defaults: dsdh = 1, gu/gv grounded_fraction = 1 (HGrid.jl:213, UGrid.jl:94)
    {
        std::vector<double> one(std::max(ctx->nh, std::max(ctx->nu, ctx->nv)), 1.0);
        CREATECHK(hipMemcpy(ctx->f[WAVI_H_DSDH], one.data(), ctx->nh * sizeof(double), hipMemcpyHostToDevice));
        CREATECHK(hipMemcpy(ctx->f[WAVI_U_GROUNDED_FRACTION], one.data(), ctx->nu * sizeof(double), hipMemcpyHostToDevice));
        CREATECHK(hipMemcpy(ctx->f[WAVI_V_GROUNDED_FRACTION], one.data(), ctx->nv * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<double> zeta(ns), qw(ns);
    for (int k = 0; k < ns; ++k) {
        zeta[k] = 1.0 - grid->sigma[k];  // Grid.jl:170
        qw[k] = grid->quadrature_weights[k];
    }
    CREATECHK(hipMalloc(&ctx->zeta, ns * sizeof(double)));
    CREATECHK(hipMalloc(&ctx->qw, ns * sizeof(double)));
    CREATECHK(hipMemcpy(ctx->zeta, zeta.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(ctx->qw, qw.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    auto zalloc = [&](double **p, size_t n) -> hipError_t {
        hipError_t e = hipMalloc(p, n * sizeof(double));
        if (e != hipSuccess) return e;
        return hipMemset(*p, 0, n * sizeof(double));
    };
    CREATECHK(zalloc(&ctx->heta_h, ctx->nh));
    CREATECHK(zalloc(&ctx->heta_c, ctx->nc > 0 ? ctx->nc : 1));
    CREATECHK(zalloc(&ctx->dimp, ctx->nh));
    for (DevVec *dv : {&ctx->x, &ctx->b, &ctx->r, &ctx->e, &ctx->d, &ctx->rho, &ctx->t, &ctx->f3, &ctx->fixedv,
                       &ctx->cr, &ctx->cp, &ctx->cq, &ctx->ct}) {
        CREATECHK(zalloc(&dv->u, ctx->nu));
        CREATECHK(zalloc(&dv->v, ctx->nv));
    }
    CREATECHK(zalloc(&ctx->Pr0, g.npart));
    CREATECHK(zalloc(&ctx->Pr1, g.npart));
    CREATECHK(zalloc(&ctx->Pd, g.npart));
    CREATECHK(zalloc(&ctx->Pb, g.npart));
    CREATECHK(zalloc(&ctx->Pn, g.npart));
    CREATECHK(zalloc(&ctx->Pz, g.npart));
    CREATECHK(zalloc(&ctx->cdinv.u, ctx->nu));
    CREATECHK(zalloc(&ctx->cdinv.v, ctx->nv));
    CREATECHK(zalloc(&ctx->quad, 4 * (size_t)ctx->nh));
    CREATECHK(hipHostMalloc((void **)&ctx->h_norms, 2 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    CREATECHK(hipHostGetDevicePointer((void **)&ctx->d_norms, ctx->h_norms, 0));
    CREATECHK(hipMalloc(&ctx->st, sizeof(CGScalars)));
    CREATECHK(hipMemset(ctx->st, 0, sizeof(CGScalars)));
    CREATECHK(hipHostMalloc((void **)&ctx->h_st, sizeof(CGScalars), hipHostMallocDefault));
    ctx->nblk = ((g.nbx * g.nby + 7) / 8) * 8;
    CREATECHK(hipHostMalloc((void **)&ctx->h_flag, sizeof(CGMHostFlag), hipHostMallocMapped | hipHostMallocCoherent));
    CREATECHK(hipHostGetDevicePointer((void **)&ctx->d_flag, ctx->h_flag, 0));
    for (hipEvent_t &e : ctx->cg_ev) CREATECHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    CREATECHK(hipEventCreateWithFlags(&ctx->ev_ls, hipEventDisableTiming));
    for (hipEvent_t &e : ctx->ev_tune) CREATECHK(hipEventCreate(&e));
    CREATECHK(hipEventCreateWithFlags(&ctx->ev_ls_done, hipEventDisableTiming));
    if (precision_mode == WAVI_PRECISION_MIXED) {
        VSet<float> &w = ctx->v32;
        auto falloc = [&](float **p, size_t n) -> hipError_t {
            hipError_t e = hipMalloc(p, n * sizeof(float));
            if (e != hipSuccess) return e;
            return hipMemset(*p, 0, n * sizeof(float));
        };
        CREATECHK(falloc(&w.heta_h, ctx->nh));
        CREATECHK(falloc(&w.heta_c, ctx->nc > 0 ? ctx->nc : 1));
        CREATECHK(falloc(&w.dimp, ctx->nh));
        for (DevVecT<float> *dv : {&w.h, &w.beff, &w.e, &w.r, &w.d, &w.rho, &w.t, &w.cr, &w.cp, &w.cq, &w.ct, &w.xc}) {
            CREATECHK(falloc(&dv->u, ctx->nu));
            CREATECHK(falloc(&dv->v, ctx->nv));
        }
        ctx->px = (unsigned)((TILE * g.nbx + 2 + 31) / 32 * 32);
        ctx->ns = ctx->px * (unsigned)(TILE * g.nby + 2);
        CREATECHK(falloc(&ctx->arena, (size_t)CGM_NARRAY * ctx->ns));
        CREATECHK(hipMalloc(&ctx->marena, 2 * (size_t)ctx->ns));
        CREATECHK(hipMemset(ctx->marena, 0, 2 * (size_t)ctx->ns));
        ctx->npad = (g.nbx * g.nby + 1023) / 1024 * 1024;
        CREATECHK(falloc(&ctx->rings, 2 * (size_t)(g.nbx * g.nby + 1) * 6 * 64));
        CREATECHK(hipMalloc(&ctx->Pgd, 2 * (size_t)ctx->npad * sizeof(float2)));
        CREATECHK(hipMemset(ctx->Pgd, 0, 2 * (size_t)ctx->npad * sizeof(float2)));
        CREATECHK(hipMalloc(&ctx->stm, sizeof(CGMScalars)));
        CREATECHK(hipMemset(ctx->stm, 0, sizeof(CGMScalars)));
    }
    {
        HaarMasks hm_tab[4];
        build_haar_masks(hm_tab);
        CREATECHK(hipMemcpyToSymbol(HIP_SYMBOL(c_haar), hm_tab, sizeof(hm_tab), 0, hipMemcpyHostToDevice));
    }
    HaarTable tab;
    if (build_haar_table(solver->levels, tab, ctx->err) != WAVI_OK) return fail(WAVI_EARG);
    CREATECHK(hipMalloc(&ctx->tab, sizeof(HaarTable)));
    CREATECHK(hipMemcpy(ctx->tab, &tab, sizeof(HaarTable), hipMemcpyHostToDevice));
#undef CREATECHK
    *out = ctx;
    return WAVI_OK;
}

int wavi_ctx_destroy(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_OK;
    std::lock_guard<std::mutex> lock(g_graph_mu);
    ctx_free(ctx);
    return WAVI_OK;
}

static void ctx_free(wavi_ctx *ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (uint8_t *p : {ctx->hm, ctx->hfix, ctx->um, ctx->vm, ctx->cm, ctx->ui, ctx->vi, ctx->ufix, ctx->vfix,
                       ctx->mwu, ctx->mwv, ctx->saved_mwu, ctx->saved_mwv})
        if (p) (void)hipFree(p);
    for (int id = 0; id < WAVI_NFIELDS; ++id)
        if (ctx->f[id]) (void)hipFree(ctx->f[id]);
    for (double *p : ctx->saved)
        if (p) (void)hipFree(p);
    for (double *p : {ctx->zeta, ctx->qw, ctx->heta_h, ctx->heta_c, ctx->dimp, ctx->Pr0, ctx->Pr1, ctx->Pd, ctx->Pb,
                      ctx->Pn, ctx->Pz, ctx->cdinv.u, ctx->cdinv.v, ctx->quad})
        if (p) (void)hipFree(p);
    for (DevVec *dv : {&ctx->x, &ctx->b, &ctx->r, &ctx->e, &ctx->d, &ctx->rho, &ctx->t, &ctx->f3, &ctx->fixedv,
                       &ctx->cr, &ctx->cp, &ctx->cq, &ctx->ct}) {
        if (dv->u) (void)hipFree(dv->u);
        if (dv->v) (void)hipFree(dv->v);
    }
    {
        VSet<float> &w = ctx->v32;
        for (float *p : {w.heta_h, w.heta_c, w.dimp})
            if (p) (void)hipFree(p);
        for (DevVecT<float> *dv : {&w.h, &w.beff, &w.e, &w.r, &w.d, &w.rho, &w.t, &w.cr, &w.cp, &w.cq, &w.ct, &w.xc}) {
            if (dv->u) (void)hipFree(dv->u);
            if (dv->v) (void)hipFree(dv->v);
        }
    }
    for (void *p : {(void *)ctx->arena, (void *)ctx->marena, (void *)ctx->Pgd, (void *)ctx->rings})
        if (p) (void)hipFree(p);
    if (ctx->stm) (void)hipFree(ctx->stm);
    if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
    for (hipEvent_t e : ctx->cg_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ctx->ev_ls, ctx->ev_ls_done, ctx->ev_tune[0], ctx->ev_tune[1]})
        if (e) (void)hipEventDestroy(e);
    if (ctx->st) (void)hipFree(ctx->st);
    if (ctx->tab) (void)hipFree(ctx->tab);
    if (ctx->h_norms) (void)hipHostFree(ctx->h_norms);
    if (ctx->h_st) (void)hipHostFree(ctx->h_st);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->burst_ev) (void)hipEventDestroy(e);
    if (ctx->cgm_chunk) (void)hipGraphExecDestroy(ctx->cgm_chunk);
    if (ctx->adj_base) (void)hipFree(ctx->adj_base);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int wavi_set_dt(wavi_ctx *ctx, double dt)
{
    if (!ctx) return WAVI_EARG;
    if (!(dt > 0)) {
        ctx->err = "timestep dt must be a positive number";
        return WAVI_EARG;
    }
    ctx->P.dt = dt;
    return WAVI_OK;
}

int wavi_field_size(const wavi_ctx *ctx, int field, size_t *n)
{
    if (!ctx || !n) return WAVI_EARG;
    int len = field_len(ctx, field);
    if (len < 0) return WAVI_EARG;
    *n = (size_t)len;
    return WAVI_OK;
}

int wavi_set_field(wavi_ctx *ctx, int field, const double *host, size_t n)
{
    if (!ctx) return WAVI_EARG;
    int len = field_len(ctx, field);
    if (len < 0 || !host) {
        ctx->err = "unknown field or null pointer";
        return WAVI_EARG;
    }
    if ((size_t)len != n) {
        ctx->err = "field size mismatch: expected " + std::to_string(len) + ", got " + std::to_string(n);
        return WAVI_EDIM;
    }
    std::lock_guard<std::mutex> lock(g_graph_mu);
    if (field == WAVI_WU_MASK || field == WAVI_WV_MASK) {
        double *tmp = nullptr;
        HIPCHK(hipMalloc(&tmp, n * sizeof(double)));
        HIPCHK(hipMemcpy(tmp, host, n * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mask_from_double, flat(len), dim3(256), 0, ctx->stream, tmp,
                           field == WAVI_WU_MASK ? ctx->mwu : ctx->mwv, len);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(tmp));
        return WAVI_OK;
    }
    HIPCHK(hipMemcpy(ctx->f[field], host, n * sizeof(double), hipMemcpyHostToDevice));
    return WAVI_OK;
}

int wavi_get_field(wavi_ctx *ctx, int field, double *host, size_t n)
{
    if (!ctx) return WAVI_EARG;
    int len = field_len(ctx, field);
    if (len < 0 || !host) {
        ctx->err = "unknown field or null pointer";
        return WAVI_EARG;
    }
    if ((size_t)len != n) {
        ctx->err = "field size mismatch: expected " + std::to_string(len) + ", got " + std::to_string(n);
        return WAVI_EDIM;
    }
    std::lock_guard<std::mutex> lock(g_graph_mu);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (field == WAVI_WU_MASK || field == WAVI_WV_MASK) {
        double *tmp = nullptr;
        HIPCHK(hipMalloc(&tmp, n * sizeof(double)));
        hipLaunchKernelGGL(k_mask_to_double, flat(len), dim3(256), 0, ctx->stream,
                           field == WAVI_WU_MASK ? ctx->mwu : ctx->mwv, tmp, len);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipFree(tmp));
        return WAVI_OK;
    }
    HIPCHK(hipMemcpy(host, ctx->f[field], n * sizeof(double), hipMemcpyDeviceToHost));
    return WAVI_OK;
}

int wavi_init_quad_f2(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    hipLaunchKernelGGL(k_init_quad_f2, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->nh, ctx->hm,
                       ctx->f[WAVI_H_H], ctx->f[WAVI_H_ETA_AV], ctx->f[WAVI_H_QUAD_F2]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_state_save(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    if (!ctx->has_saved) {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        ctx->saved.assign(WAVI_NFIELDS, nullptr);
        for (int id = 0; id < WAVI_NFIELDS; ++id) {
            if (!ctx->f[id]) continue;
            HIPCHK(hipMalloc(&ctx->saved[id], field_len(ctx, id) * sizeof(double)));
        }
        HIPCHK(hipMalloc(&ctx->saved_mwu, ctx->nu));
        HIPCHK(hipMalloc(&ctx->saved_mwv, ctx->nv));
        ctx->has_saved = true;
    }
    for (int id = 0; id < WAVI_NFIELDS; ++id)
        if (ctx->f[id])
            HIPCHK(hipMemcpyAsync(ctx->saved[id], ctx->f[id], field_len(ctx, id) * sizeof(double),
                                  hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->saved_mwu, ctx->mwu, ctx->nu, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->saved_mwv, ctx->mwv, ctx->nv, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_state_restore(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    if (!ctx->has_saved) {
        ctx->err = "no saved state";
        return WAVI_ESTATE;
    }
    for (int id = 0; id < WAVI_NFIELDS; ++id)
        if (ctx->f[id])
            HIPCHK(hipMemcpyAsync(ctx->f[id], ctx->saved[id], field_len(ctx, id) * sizeof(double),
                                  hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->mwu, ctx->saved_mwu, ctx->nu, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->mwv, ctx->saved_mwv, ctx->nv, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

#pragma GCC visibility pop
}  // extern "C"

// ---------------------------------------------------------------------------
// pieces of the solve (device resident)

static InnerArgs inner_args(wavi_ctx *ctx)
{
    InnerArgs p;
    p.hm = ctx->hm;
    p.um = ctx->um;
    p.vm = ctx->vm;
    p.cm = ctx->cm;
    p.U = ctx->f[WAVI_U_U];
    p.V = ctx->f[WAVI_V_V];
    p.h = ctx->f[WAVI_H_H];
    p.wc = ctx->f[WAVI_H_WEERTMAN_C];
    p.gfh = ctx->f[WAVI_H_GROUNDED_FRACTION];
    p.gfu = ctx->f[WAVI_U_GROUNDED_FRACTION];
    p.gfv = ctx->f[WAVI_V_GROUNDED_FRACTION];
    p.dsdh = ctx->f[WAVI_H_DSDH];
    p.glen_b = ctx->f[WAVI_S_GLEN_B];
    p.zeta = ctx->zeta;
    p.qw = ctx->qw;
    p.sr = ctx->f[WAVI_H_SHELF_STRAIN_RATE];
    p.avs = ctx->f[WAVI_H_AV_SPEED];
    p.bs = ctx->f[WAVI_H_BED_SPEED];
    p.beta = ctx->f[WAVI_H_BETA];
    p.tau = ctx->f[WAVI_H_TAU_BED];
    p.eta = ctx->f[WAVI_S_ETA];
    p.etaav = ctx->f[WAVI_H_ETA_AV];
    p.f1 = ctx->f[WAVI_H_QUAD_F1];
    p.f2 = ctx->f[WAVI_H_QUAD_F2];
    p.beff = ctx->f[WAVI_H_BETA_EFF];
    p.bu = ctx->f[WAVI_U_BETA_EFF];
    p.bv = ctx->f[WAVI_V_BETA_EFF];
    p.heta_h = ctx->heta_h;
    p.heta_c = ctx->heta_c;
    p.dimp = ctx->dimp;
    p.rdx = ctx->rdx;
    p.rdy = ctx->rdy;
    p.wexp = 1.0 / ctx->P.weertman_m - 1.0;
    p.wreg2 = ctx->P.weertman_reg_speed * ctx->P.weertman_reg_speed;
    p.gexp = 1.0 / ctx->P.glen_n - 1.0;
    p.greg2 = ctx->P.glen_reg_strain_rate * ctx->P.glen_reg_strain_rate;
    p.c0 = (((-ctx->P.density_ice) * ctx->P.g) * ctx->S.super_implicitness) * ctx->P.dt;
    p.ns = ctx->ns3 / ctx->nh;
    p.niter_visc = ctx->S.n_iter_viscosity;
    p.xfirst_c2h = ctx->g.nx > ctx->g.ny ? 1 : 0;  // KroneckerProduct branch for gc.cent
    p.yfirst_h2c = ctx->g.ny <= ctx->g.nx ? 1 : 0; // and for gh.cent_xy
    return p;
}

static int do_inner_update(wavi_ctx *ctx)
{
    InnerArgs p = inner_args(ctx);
    hipLaunchKernelGGL(k_inner_h, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->g, p);
    hipLaunchKernelGGL(k_inner_uvc, flat(ctx->nu + ctx->nv + ctx->nc), dim3(256), 0, ctx->stream, ctx->g, p);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int do_op(wavi_ctx *ctx, const double *xu, const double *xv, double *yu, double *yv)
{
    const int nb = band_blocks(ctx->g.nbx, ctx->g.nby);
    hipLaunchKernelGGL(k_op<double>, dim3(nb), dim3(256), 0, ctx->stream, ctx->g, coef64(ctx), xu, xv, yu, yv, nb);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// get_rhs into ctx->b
static int do_rhs(wavi_ctx *ctx)
{
    hipLaunchKernelGGL((k_select_uv<double, double>), flat_uv(ctx->nu, ctx->nv), dim3(256), 0, ctx->stream,
                       ctx->f[WAVI_U_U], ctx->ufix, ctx->fixedv.u, ctx->nu, ctx->f[WAVI_V_V], ctx->vfix, ctx->fixedv.v,
                       ctx->nv);
    CHK(do_op(ctx, ctx->fixedv.u, ctx->fixedv.v, ctx->f3.u, ctx->f3.v));
    RhsArgs a;
    a.hm = ctx->hm;
    a.ui = ctx->ui;
    a.vi = ctx->vi;
    a.s = ctx->f[WAVI_H_S];
    a.dsdh = ctx->f[WAVI_H_DSDH];
    a.acc = ctx->f[WAVI_H_ACCUMULATION];
    a.mel = ctx->f[WAVI_H_BASAL_MELT];
    a.hu = ctx->f[WAVI_U_H];
    a.su = ctx->f[WAVI_U_S];
    a.hv = ctx->f[WAVI_V_H];
    a.sv = ctx->f[WAVI_V_S];
    a.f3u = ctx->f3.u;
    a.f3v = ctx->f3.v;
    a.bu = ctx->b.u;
    a.bv = ctx->b.v;
    a.rdx = ctx->rdx;
    a.rdy = ctx->rdy;
    a.g = ctx->P.g;
    a.ri = ctx->P.density_ice;
    a.rw = ctx->P.density_ocean;
    a.sl = ctx->P.sea_level_wrt_geoid;
    a.tdt = ctx->S.super_implicitness * ctx->P.dt;
    hipLaunchKernelGGL(k_rhs, flat(ctx->nu + ctx->nv), dim3(256), 0, ctx->stream, ctx->g, a);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

template <typename T>
static OpCoef<T> coef_of(const wavi_ctx *ctx, const VSet<T> &w)
{
    OpCoef<T> c;
    c.heta_h = w.heta_h;
    c.heta_c = w.heta_c;
    c.dimp = w.dimp;
    c.hu = w.h.u;
    c.hv = w.h.v;
    c.bu = w.beff.u;
    c.bv = w.beff.v;
    c.hm = ctx->hm;
    c.um = ctx->um;
    c.vm = ctx->vm;
    c.cm = ctx->cm;
    c.ui = ctx->ui;
    c.vi = ctx->vi;
    c.nx = ctx->g.nx;
    c.ny = ctx->g.ny;
    c.rdx = (T)ctx->rdx;
    c.rdy = (T)ctx->rdy;
    return c;
}

// the FP64 buffer set: views of the context's fields and vectors
static VSet<double> vset64(wavi_ctx *ctx)
{
    VSet<double> w;
    w.heta_h = ctx->heta_h;
    w.heta_c = ctx->heta_c;
    w.dimp = ctx->dimp;
    w.h = {ctx->f[WAVI_U_H], ctx->f[WAVI_V_H]};
    w.beff = {ctx->f[WAVI_U_BETA_EFF], ctx->f[WAVI_V_BETA_EFF]};
    w.e = {ctx->e.u, ctx->e.v};
    w.r = {ctx->r.u, ctx->r.v};
    w.d = {ctx->d.u, ctx->d.v};
    w.rho = {ctx->rho.u, ctx->rho.v};
    w.t = {ctx->t.u, ctx->t.v};
    w.cr = {ctx->cr.u, ctx->cr.v};
    w.cp = {ctx->cp.u, ctx->cp.v};
    w.cq = {ctx->cq.u, ctx->cq.v};
    w.ct = {ctx->ct.u, ctx->ct.v};
    w.xc = {ctx->f[WAVI_WU_CORRECTION_COARSE], ctx->f[WAVI_WV_CORRECTION_COARSE]};
    return w;
}

template <typename T>
static int do_diag(wavi_ctx *ctx, const VSet<T> &w)
{
    hipLaunchKernelGGL(k_diag<T>, flat(ctx->nu + ctx->nv), dim3(256), 0, ctx->stream, ctx->g, coef_of(ctx, w),
                       w.d.u, w.d.v);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// Tile shapes of the smoother kernel k_smooth4<T, TX, TY>, each walked in
// XCD bands of tile columns or of tile rows: variant v is shape v % 3,
// column bands when v < 3.
constexpr int SMOOTH_SHAPES = 3, SMOOTH_VARIANTS = 6;
// Variant SMOOTH_GRID: the whole grid in one workgroup (k_smooth_grid), all
// smoothing iterations in one launch, for grids whose window fits
// SMG_MAX_SLOTS points per thread and the context's LDS limit.
constexpr int SMOOTH_GRID = SMOOTH_VARIANTS;
static const int SMOOTH_TX[SMOOTH_SHAPES] = {16, 32, 16};
static const int SMOOTH_TY[SMOOTH_SHAPES] = {16, 16, 32};

template <typename T>
static void launch_smooth(wavi_ctx *ctx, int variant, const OpCoef<T> &c, const VSet<T> &w, const T *xu, const T *xv,
                          T *yu, T *yv, int packed)
{
    const GridDims &g = ctx->g;
    const int shape = variant % SMOOTH_SHAPES, cols = variant < SMOOTH_SHAPES ? 1 : 0;
    const int tx = SMOOTH_TX[shape], ty = SMOOTH_TY[shape];
    const int ntx = (std::max(g.nxu, g.nxv) + tx - 1) / tx, nty = (std::max(g.nyu, g.nyv) + ty - 1) / ty;
    const int nb = band_blocks(ntx, nty);
    const dim3 grid((unsigned)nb);
    const T om = (T)ctx->S.smoother_omega;
    switch (shape) {
    case 1:
        hipLaunchKernelGGL((k_smooth4<T, 32, 16>), grid, dim3(256), 0, ctx->stream, g, c, xu, xv, yu, yv, w.r.u,
                           w.r.v, w.d.u, w.d.v, om, packed, ntx, nty, nb, cols);
        break;
    case 2:
        hipLaunchKernelGGL((k_smooth4<T, 16, 32>), grid, dim3(256), 0, ctx->stream, g, c, xu, xv, yu, yv, w.r.u,
                           w.r.v, w.d.u, w.d.v, om, packed, ntx, nty, nb, cols);
        break;
    default:
        hipLaunchKernelGGL((k_smooth4<T, 16, 16>), grid, dim3(256), 0, ctx->stream, g, c, xu, xv, yu, yv, w.r.u,
                           w.r.v, w.d.u, w.d.v, om, packed, ntx, nty, nb, cols);
        break;
    }
}

static bool smooth_grid_fits(const wavi_ctx *ctx, size_t elem_bytes)
{
    const SmoothGridDims d(ctx->g);
    return d.points() <= SMG_THREADS * SMG_MAX_SLOTS && d.elems() * elem_bytes <= ctx->lds_max;
}

// nsmooth smoothing iterations of the whole grid in one launch, from x into y
// (which may be x); the grid must fit (smooth_grid_fits).
template <typename T>
static void launch_smooth_grid(wavi_ctx *ctx, const OpCoef<T> &c, const VSet<T> &w, const T *xu, const T *xv, T *yu,
                               T *yv, int packed, int nsmooth)
{
    const GridDims &g = ctx->g;
    const SmoothGridDims d(g);
    const size_t lds = d.elems() * sizeof(T);
    const T om = (T)ctx->S.smoother_omega;
    if (d.points() <= SMG_THREADS)
        hipLaunchKernelGGL((k_smooth_grid<T, 1>), dim3(1), dim3(SMG_THREADS), lds, ctx->stream, g, c, xu, xv, yu, yv,
                           w.r.u, w.r.v, w.d.u, w.d.v, om, packed, nsmooth);
    else
        hipLaunchKernelGGL((k_smooth_grid<T, 2>), dim3(1), dim3(SMG_THREADS), lds, ctx->stream, g, c, xu, xv, yu, yv,
                           w.r.u, w.r.v, w.d.u, w.d.v, om, packed, nsmooth);
}

// The autotuner's results, shared by the contexts of one process: the
// fastest smoother variant (tile shape and band order) per (nx, ny, V-cycle
// precision).
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, int> g_smooth_tuned;

// The smoother variant of the context, chosen on first use: the one set by
// wavi_set_smoother_tile, else one tuned before for a grid of this size and
// precision, else measured now: after a warm-up of each, three timed runs of
// one smoothing phase (nsmooth iterations) of every candidate from the
// current iterate w.e into w.t (free while smoothing; the results are not
// used), each run between two events.  The candidates are the three tile
// shapes in column and in row bands (one launch per iteration) and, where
// the grid fits, the whole grid in one workgroup (one launch per phase).  A
// variant's time is its fastest run (launches of small grids take a few
// microseconds, and the fastest of several runs filters out the jitter of the
// others), and the fastest variant is kept.  Every variant gives the same
// bits, so the choice never changes a result.
template <typename T>
static int smooth_shape(wavi_ctx *ctx, const VSet<T> &w, const OpCoef<T> &c, int packed)
{
    const int pi = sizeof(T) == sizeof(double) ? 0 : 1;
    if (ctx->smooth_shape[pi] >= 0) return ctx->smooth_shape[pi];
    const auto key = std::make_tuple(ctx->g.nx, ctx->g.ny, pi);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_smooth_tuned.find(key);
        if (it != g_smooth_tuned.end()) return ctx->smooth_shape[pi] = it->second;
    }
    const int n = ctx->S.nsmooth;
    if (n <= 0) return 0;  // nothing to smooth, nothing to tune
    const int candidates = SMOOTH_VARIANTS + (smooth_grid_fits(ctx, sizeof(T)) ? 1 : 0);
    auto phase = [&](int sh) {
        if (sh == SMOOTH_GRID)
            launch_smooth_grid(ctx, c, w, w.e.u, w.e.v, w.t.u, w.t.v, packed, n);
        else
            for (int r = 0; r < n; ++r) launch_smooth(ctx, sh, c, w, w.e.u, w.e.v, w.t.u, w.t.v, packed);
    };
    int best = 0;
    float best_ms = 0.f;
    for (int sh = 0; sh < candidates; ++sh) {
        phase(sh);
        float ms = 0.f;
        for (int run = 0; run < 3; ++run) {
            if (hipEventRecord(ctx->ev_tune[0], ctx->stream) != hipSuccess) return 0;
            phase(sh);
            if (hipEventRecord(ctx->ev_tune[1], ctx->stream) != hipSuccess) return 0;
            if (hipEventSynchronize(ctx->ev_tune[1]) != hipSuccess) return 0;
            float t = 0.f;
            if (hipEventElapsedTime(&t, ctx->ev_tune[0], ctx->ev_tune[1]) != hipSuccess) return 0;
            ms = run == 0 ? t : std::min(ms, t);
        }
        if (sh == 0 || ms < best_ms) {
            best = sh;
            best_ms = ms;
        }
    }
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_smooth_tuned.emplace(key, best);
    }
    return ctx->smooth_shape[pi] = best;
}

// gauss_seidel_smoother! (update_preconditioners.jl:133-150) on w.e with
// right-hand side w.r and diagonal w.d: nsmooth iterations of the four
// colours in `order`, one k_smooth4 launch per iteration.  A launch reads
// one buffer and writes the other, so the iterations alternate between w.e
// and w.t (free during smoothing); after an odd count the result is copied
// back into w.e.
template <typename T>
static int do_smoother(wavi_ctx *ctx, const VSet<T> &w, const int order[4])
{
    OpCoef<T> c = coef_of(ctx, w);
    const int packed = order[0] | order[1] << 4 | order[2] << 8 | order[3] << 12;
    int shape = smooth_shape(ctx, w, c, packed);
    if (shape == SMOOTH_GRID) {
        if (smooth_grid_fits(ctx, sizeof(T))) {
            ProfScope ps(ctx, WAVI_KT_COLOUR);
            if (ctx->S.nsmooth > 0)
                launch_smooth_grid(ctx, c, w, w.e.u, w.e.v, w.e.u, w.e.v, packed, ctx->S.nsmooth);
            HIPCHK(hipGetLastError());
            return WAVI_OK;
        }
        shape = 0;  // set for a V-cycle of the other precision, whose window does not fit
    }
    T *xu = w.e.u, *xv = w.e.v, *yu = w.t.u, *yv = w.t.v;
    for (int it = 0; it < ctx->S.nsmooth; ++it) {
        ProfScope ps(ctx, WAVI_KT_COLOUR);
        launch_smooth(ctx, shape, c, w, xu, xv, yu, yv, packed);
        std::swap(xu, yu);
        std::swap(xv, yv);
    }
    if (ctx->S.nsmooth & 1)
        hipLaunchKernelGGL(k_copy_masked_uv<T>, flat_uv(ctx->nu, ctx->nv), dim3(256), 0, ctx->stream, w.e.u, w.t.u,
                           ctx->ui, ctx->nu, w.e.v, w.t.v, ctx->vi, ctx->nv);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// Longest chunk of coarse-CG launches queued at once (cg_chunk_len).
constexpr int CG_CHUNK = 32;

template <typename T>
static CGArgs<T> cg_args(wavi_ctx *ctx, const VSet<T> &w)
{
    CGArgs<T> a;
    a.g = ctx->g;
    a.c = coef_of(ctx, w);
    a.levels = ctx->S.levels;
    a.mwu = ctx->mwu;
    a.mwv = ctx->mwv;
    a.ui = ctx->ui;
    a.vi = ctx->vi;
    a.xu = w.xc.u;
    a.xv = w.xc.v;
    a.ru = w.cr.u;
    a.rv = w.cr.v;
    a.pu = w.cp.u;
    a.pv = w.cp.v;
    a.qu = w.cq.u;
    a.qv = w.cq.v;
    a.tu = w.ct.u;
    a.tv = w.ct.v;
    a.Pr = ctx->Pr0;
    a.Pd = ctx->Pd;
    a.Pb = ctx->Pb;
    a.Pz = ctx->Pz;
    a.dinvu = ctx->cdinv.u;
    a.dinvv = ctx->cdinv.v;
    a.st = ctx->st;
    a.hflag = ctx->d_flag;
    a.tol_coarse = ctx->S.tol_coarse;
    a.reltol = std::sqrt(2.220446049250313e-16);  // IterativeSolvers' default sqrt(eps)
    a.maxiter = ctx->S.maxiter_coarse;
    a.nblk = ctx->nblk;
    return a;
}

// Full chunks of CG iterations replay one captured HIP graph per context
// (about 1% less time per launch on MI355X than separate launches);
// WAVI_CG_GRAPH=0 launches them one by one instead.
static bool use_cg_graph()
{
    static const bool on = !(getenv("WAVI_CG_GRAPH") && atoi(getenv("WAVI_CG_GRAPH")) == 0);
    return on;
}

// Iterations (launches, or triplets of launches in FP64) of coarse-CG chunk
// k.  The host reads the stop report of chunk k - 1 once chunk k is queued,
// so one chunk keeps the GPU busy while the host waits, and every launch
// queued past the stop runs as a no-op (the kernels read the stop from the
// partials).  Without a prediction (prev = 0) chunk k holds min(32,
// 2^(k+1)), so a CG that stops after a few iterations leaves few launches
// past its stop while a long CG runs in chunks of 32 after its first 30.
// With prev = the iterations of the context's previous coarse CG (a
// transient's CGs change length slowly from step to step), chunk 0 holds
// prev + 1, at most 32 (the launches up to and including the one that
// stops, if this CG is as long), and chunk k >= 1 holds min(32, 2^k).
static int cg_chunk_len(int k, int prev)
{
    if (prev > 0) return k == 0 ? std::min(prev + 1, CG_CHUNK) : (k >= 5 ? CG_CHUNK : 1 << k);
    return k >= 4 ? CG_CHUNK : 2 << k;
}

// Relative stop of the mixed-precision coarse CG: 2^-19 = 16 FP32 ulps.
constexpr double CGM_FP32_RELTOL = 0x1p-19;
// ... and of its Jacobi-preconditioned form in a linear solve's V-cycles,
// in the D^-1 norm: its FP32 recurrences stall above 2^-19 (the scaling
// spans the Haar levels' orders of magnitude), and the flexible CG around
// the V-cycle needs no more.  Measured on the 1280 x 160 gradient: 48,178
// coarse iterations at 2^-19, 22,134 at 2^-14, the same V-cycle count
// within 7%.
constexpr double CGM_FP32_JACOBI_RELTOL = 0x1p-14;

// The coarse CG runs in one workgroup (k_cg_solve1, k_cgm_solve1) only when
// wavi_set_coarse_cg asks for it: on MI355X one workgroup spends about 3 us
// per tile and iteration (a tile's chain of loads, barriers and stores),
// so even a 6-tile grid (80 x 10) takes 3.4 (mixed) to 3.6 (FP64) times as
// long per CG iteration as the launches over all tiles, whose tiles run
// side by side.
static bool cg_one_workgroup(const wavi_ctx *ctx)
{
    return ctx->cg_mode == WAVI_CG_WORKGROUP;
}

// The three-kernel coarse CG takes its Haar rows loads first
// (haar_pm_row_ld; the same bits) on grids of at most this many tiles,
// where a launch costs one tile's chain of LDS waits.
constexpr int HAAR_LD_MAX_TILES = 256;

static bool haar_loads_first(const wavi_ctx *ctx)
{
    return ctx->g.nbx * ctx->g.nby <= HAAR_LD_MAX_TILES;
}

// A +-1 Haar transform (dir 1: P, 2: R) of the u and v vectors on the tiles,
// with the Haar rows loads first where haar_loads_first says so.
// (accu != nullptr: the result added into accu, accv at the output masks'
// points instead, outu, outv untouched)
template <typename T>
static void launch_tile_haar(wavi_ctx *ctx, int dir, const T *inu, const T *inv, const uint8_t *min_u,
                             const uint8_t *min_v, T *outu, T *outv, const uint8_t *mout_u, const uint8_t *mout_v,
                             T *accu = nullptr, T *accv = nullptr)
{
    const GridDims &g = ctx->g;
    const dim3 tg(g.nbx, g.nby);
    if (haar_loads_first(ctx))
        hipLaunchKernelGGL((k_tile_haar<T, T, true>), tg, dim3(256), 0, ctx->stream, g, ctx->tab, ctx->S.levels, dir,
                           inu, inv, min_u, min_v, outu, outv, mout_u, mout_v, accu, accv);
    else
        hipLaunchKernelGGL((k_tile_haar<T, T, false>), tg, dim3(256), 0, ctx->stream, g, ctx->tab, ctx->S.levels, dir,
                           inu, inv, min_u, min_v, outu, outv, mout_u, mout_v, accu, accv);
}

// coarse CG on b_c held in w.cr; solution in the warm start w.xc
template <typename T>
static int do_coarse_cg(wavi_ctx *ctx, const VSet<T> &w, int *iters)
{
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby), tb(ctx->nblk);
    CGArgs<T> a = cg_args(ctx, w);
    // t = P x0
    launch_tile_haar<T>(ctx, 1, a.xu, a.xv, ctx->mwu, ctx->mwv, a.tu, a.tv, ctx->ui, ctx->vi);
    ctx->h_flag->done = 0;
    ctx->h_flag->iters = 0;
    const bool pre = ctx->cg_pre;
    const bool ld = haar_loads_first(ctx);
    // k_cg_apply / k_cg_update return on the stop flag at once after a short
    // previous CG (one that fitted its first chunk): such CGs stop within a
    // chunk, so the launches queued past the stop are many.  Long CGs have
    // about one, and their launches do without the flag's round trip.
    const bool early = ctx->cg_prev < CG_CHUNK;
    if (pre)
        hipLaunchKernelGGL((k_cg_setup<T, true>), tb, dim3(256), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_cg_setup<T, false>), tb, dim3(256), 0, ctx->stream, a);
    HIPCHK(hipGetLastError());
    volatile CGMHostFlag *hf = ctx->h_flag;
    if (cg_one_workgroup(ctx)) {
        if (pre)
            hipLaunchKernelGGL((k_cg_solve1<T, true>), dim3(1), dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL((k_cg_solve1<T, false>), dim3(1), dim3(256), 0, ctx->stream, a);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (!hf->done) {
            ctx->err = "coarse CG did not terminate";
            return WAVI_ESTATE;
        }
        *iters = hf->iters;
        return WAVI_OK;
    }
    // Triplet q performs iteration q unless the CG has stopped; triplet
    // maxiter always stops, so no more are needed.  Triplets are queued in
    // chunks (cg_chunk_len); the host reads the stop report between chunks
    // and queues no more after a stop.
    const int nq = ctx->S.maxiter_coarse + 1;
    int q = 0, k = 0;
    while (q < nq) {
        const int n = std::min(cg_chunk_len(k, ctx->cg_prev), nq - q);
        // when profiling, odd chunks time sampled launches and even chunks
        // of full length time the whole chunk
        const bool sampled = ctx->prof && (k & 1) != 0;
        const int bev = sampled || n != CG_CHUNK ? -1 : burst_begin(ctx);
        for (int s = 0; s < n; ++s, ++q) {
            const int par = q & 1;
            {
                ProfScope ps(ctx, WAVI_KT_CG_DIR, sampled);
                if (pre && ld)
                    hipLaunchKernelGGL((k_cg_dir<T, true, true>), tb, dim3(256), 0, ctx->stream, a, par);
                else if (pre)
                    hipLaunchKernelGGL((k_cg_dir<T, true, false>), tb, dim3(256), 0, ctx->stream, a, par);
                else if (ld)
                    hipLaunchKernelGGL((k_cg_dir<T, false, true>), tb, dim3(256), 0, ctx->stream, a, par);
                else
                    hipLaunchKernelGGL((k_cg_dir<T, false, false>), tb, dim3(256), 0, ctx->stream, a, par);
            }
            {
                ProfScope ps(ctx, WAVI_KT_CG_APPLY, sampled);
                if (ld && early)
                    hipLaunchKernelGGL((k_cg_apply<T, true, true>), tb, dim3(256), 0, ctx->stream, a);
                else if (ld)
                    hipLaunchKernelGGL((k_cg_apply<T, true, false>), tb, dim3(256), 0, ctx->stream, a);
                else if (early)
                    hipLaunchKernelGGL((k_cg_apply<T, false, true>), tb, dim3(256), 0, ctx->stream, a);
                else
                    hipLaunchKernelGGL((k_cg_apply<T, false, false>), tb, dim3(256), 0, ctx->stream, a);
            }
            {
                ProfScope ps(ctx, WAVI_KT_CG_UPDATE, sampled);
                if (pre && early)
                    hipLaunchKernelGGL((k_cg_update<T, true, true>), tb, dim3(256), 0, ctx->stream, a, par);
                else if (pre)
                    hipLaunchKernelGGL((k_cg_update<T, true, false>), tb, dim3(256), 0, ctx->stream, a, par);
                else if (early)
                    hipLaunchKernelGGL((k_cg_update<T, false, true>), tb, dim3(256), 0, ctx->stream, a, par);
                else
                    hipLaunchKernelGGL((k_cg_update<T, false, false>), tb, dim3(256), 0, ctx->stream, a, par);
            }
        }
        burst_end(ctx, bev, q - 1, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->cg_ev[k % 3], ctx->stream));
        if (k >= 1) {
            HIPCHK(hipEventSynchronize(ctx->cg_ev[(k - 1) % 3]));
            if (hf->done) break;
        }
        ++k;
    }
    // A stop seen between chunks needs no wait for the launches queued past
    // it (they change nothing); the timing events do.
    if (!hf->done || ctx->prof) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        prof_collect(ctx);
    }
    // triplet q iterates when q < iters
    burst_collect(ctx, WAVI_KT_CG_ITER_FP64, hf->done ? hf->iters - 1 : -1);
    if (!hf->done) {
        ctx->err = "coarse CG did not terminate";
        return WAVI_ESTATE;
    }
    *iters = hf->iters;
    ctx->cg_prev = hf->iters;
    return WAVI_OK;
}

static CGMLayout cgm_layout(const wavi_ctx *ctx)
{
    CGMLayout Lo;
    Lo.F = ctx->arena;
    Lo.M = ctx->marena;
    Lo.px = ctx->px;
    Lo.ns = ctx->ns;
    return Lo;
}

template <int PAR, bool INIT, bool PRE>
static void launch_cgm_iter_p(wavi_ctx *ctx, dim3 tb, const CGMIterArgs &a)
{
    switch (ctx->S.levels) {
    case 0: hipLaunchKernelGGL((k_cgm_iter<PAR, INIT, 0, PRE>), tb, dim3(256), 0, ctx->stream, a); break;
    case 1: hipLaunchKernelGGL((k_cgm_iter<PAR, INIT, 1, PRE>), tb, dim3(256), 0, ctx->stream, a); break;
    case 2: hipLaunchKernelGGL((k_cgm_iter<PAR, INIT, 2, PRE>), tb, dim3(256), 0, ctx->stream, a); break;
    default: hipLaunchKernelGGL((k_cgm_iter<PAR, INIT, 3, PRE>), tb, dim3(256), 0, ctx->stream, a); break;
    }
}

// k_cgm_iter at the context's Haar level, Jacobi-preconditioned inside a
// linear solve (ctx->cg_pre).
template <int PAR, bool INIT>
static void launch_cgm_iter(wavi_ctx *ctx, dim3 tb, const CGMIterArgs &a)
{
    if (ctx->cg_pre)
        launch_cgm_iter_p<PAR, INIT, true>(ctx, tb, a);
    else
        launch_cgm_iter_p<PAR, INIT, false>(ctx, tb, a);
}

// The arguments of k_cgm_iter, fixed for the context.
static CGMIterArgs cgm_iter_args(wavi_ctx *ctx)
{
    CGMIterArgs ia;
    ia.Lo = cgm_layout(ctx);
    ia.R = ctx->rings;
    ia.Pgd = ctx->Pgd;
    ia.st = ctx->stm;
    ia.hflag = ctx->d_flag;
    ia.Pb = ctx->Pb;
    ia.g = ctx->g;
    ia.npad = ctx->npad;
    ia.rdx = (float)ctx->rdx;
    ia.rdy = (float)ctx->rdy;
    ia.tol_coarse = ctx->S.tol_coarse;
    // IterativeSolvers' default reltol sqrt(eps), raised to FP32's attainable
    // accuracy: the FP32 recurrences track the true residual down to about
    // 16 ulps of ||r0||, and iterating past that point lets the recursive
    // residual drift from the true one and the iterate grow (the outer
    // Picard loop then stalls near 1e-8 instead of converging).
    ia.reltol = ctx->cg_pre ? CGM_FP32_JACOBI_RELTOL : std::max(std::sqrt(2.220446049250313e-16), CGM_FP32_RELTOL);
    ia.maxiter = ctx->S.maxiter_coarse;
    ia.nblk = ctx->nblk;
    return ia;
}

// Set-up of the mixed coarse CG on b_c in w.cr with the warm start w.xc:
// coefficients and masks into the arena, P x0, r0 = b - A_c x0, the
// scalars and the host stop report reset.
static int cgm_prepare(wavi_ctx *ctx, const VSet<float> &w)
{
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby), tb(ctx->nblk);
    const CGMLayout Lo = cgm_layout(ctx);
    const OpCoef<float> c = coef_of(ctx, w);
    hipLaunchKernelGGL(k_cgm_pack, flat(std::max(ctx->nh, std::max(ctx->nu, ctx->nv))), dim3(256), 0, ctx->stream, g,
                       c, ctx->mwu, ctx->mwv, Lo, ctx->cg_pre ? (const double *)ctx->cdinv.u : nullptr,
                       ctx->cg_pre ? (const double *)ctx->cdinv.v : nullptr);
    // P x0
    launch_tile_haar<float>(ctx, 1, w.xc.u, w.xc.v, ctx->mwu, ctx->mwv, w.ct.u, w.ct.v, ctx->ui, ctx->vi);
    CGMArgs sa;
    sa.g = g;
    sa.c = c;
    sa.Lo = Lo;
    sa.levels = ctx->S.levels;
    sa.mwu = ctx->mwu;
    sa.mwv = ctx->mwv;
    sa.bu = w.cr.u;
    sa.bv = w.cr.v;
    sa.x0u = w.xc.u;
    sa.x0v = w.xc.v;
    sa.t0u = w.ct.u;
    sa.t0v = w.ct.v;
    sa.R = ctx->rings;
    sa.Pb = ctx->Pb;
    sa.st = ctx->stm;
    sa.hflag = ctx->d_flag;
    sa.nblk = ctx->nblk;
    ctx->h_flag->done = 0;
    ctx->h_flag->iters = 0;
    if (ctx->cg_pre)
        hipLaunchKernelGGL(k_cgm_setup<true>, tb, dim3(256), 0, ctx->stream, sa);
    else
        hipLaunchKernelGGL(k_cgm_setup<false>, tb, dim3(256), 0, ctx->stream, sa);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// Mixed-precision coarse CG (Chronopoulos-Gear form, one launch per
// iteration) on b_c held in v32.cr; solution in the warm start v32.xc.  The
// CG works in the padded arena: k_cgm_pack fills the coefficients and masks,
// k_cgm_setup the vectors, k_cgm_unpack returns the solution.
static int rdv_coarse_cg(wavi_ctx *ctx, const VSet<float> &w, int *iters);

template <>
int do_coarse_cg<float>(wavi_ctx *ctx, const VSet<float> &w, int *iters)
{
    if (ctx->rdv) return rdv_coarse_cg(ctx, w, iters);
    const GridDims &g = ctx->g;
    dim3 tb(ctx->nblk);
    const CGMLayout Lo = cgm_layout(ctx);
    CHK(cgm_prepare(ctx, w));
    const CGMIterArgs ia = cgm_iter_args(ctx);
    volatile CGMHostFlag *hf = ctx->h_flag;
    if (cg_one_workgroup(ctx)) {
        const int sel = ctx->S.levels + (ctx->cg_pre ? 4 : 0);
        switch (sel) {
        case 0: hipLaunchKernelGGL((k_cgm_solve1<0, false>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 1: hipLaunchKernelGGL((k_cgm_solve1<1, false>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 2: hipLaunchKernelGGL((k_cgm_solve1<2, false>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 3: hipLaunchKernelGGL((k_cgm_solve1<3, false>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 4: hipLaunchKernelGGL((k_cgm_solve1<0, true>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 5: hipLaunchKernelGGL((k_cgm_solve1<1, true>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        case 6: hipLaunchKernelGGL((k_cgm_solve1<2, true>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        default: hipLaunchKernelGGL((k_cgm_solve1<3, true>), dim3(1), dim3(256), 0, ctx->stream, ia); break;
        }
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_cgm_unpack, flat(std::max(ctx->nu, ctx->nv)), dim3(256), 0, ctx->stream, g, Lo, w.xc.u,
                           w.xc.v);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (!hf->done) {
            ctx->err = "coarse CG did not terminate";
            return WAVI_ESTATE;
        }
        *iters = hf->iters;
        return WAVI_OK;
    }
    launch_cgm_iter<0, true>(ctx, tb, ia);
    HIPCHK(hipGetLastError());
    // Launch q (q >= 1) performs iteration q - 1 unless the CG has stopped;
    // launch maxiter + 1 always stops, so no more are needed.  Launches are
    // queued in chunks (cg_chunk_len); the host reads the stop report
    // between chunks and queues no more after a stop.  Full chunks are
    // timed (bursts) and replay one captured graph.
    const int qmax = ctx->S.maxiter_coarse + 1;
    const bool use_graph = use_cg_graph();
    int q = 1, k = 0;
    while (q <= qmax) {
        const int n = std::min(cg_chunk_len(k, ctx->cg_prev), qmax - q + 1);
        const int bev = n == CG_CHUNK ? burst_begin(ctx) : -1;
        if (n == CG_CHUNK && use_graph && (q & 1) && !ctx->cg_pre) {
            // the graph's launches have parities 1, 0, 1, ...: it replays a chunk that starts at an odd q
            if (!ctx->cgm_chunk) {
                std::lock_guard<std::mutex> lock(g_graph_mu);
                hipGraph_t gr = nullptr;
                HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < CG_CHUNK; ++s) {
                    if (s & 1)
                        launch_cgm_iter<0, false>(ctx, tb, ia);
                    else
                        launch_cgm_iter<1, false>(ctx, tb, ia);
                }
                HIPCHK(hipStreamEndCapture(ctx->stream, &gr));
                const hipError_t e = hipGraphInstantiate(&ctx->cgm_chunk, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                HIPCHK(e);
            }
            HIPCHK(hipGraphLaunch(ctx->cgm_chunk, ctx->stream));
            q += CG_CHUNK;
        } else {
            for (int s = 0; s < n; ++s, ++q) {
                if (q & 1)
                    launch_cgm_iter<1, false>(ctx, tb, ia);
                else
                    launch_cgm_iter<0, false>(ctx, tb, ia);
            }
        }
        burst_end(ctx, bev, q - 1, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->cg_ev[k % 3], ctx->stream));
        if (k >= 1) {
            HIPCHK(hipEventSynchronize(ctx->cg_ev[(k - 1) % 3]));
            if (hf->done) break;
        }
        ++k;
    }
    hipLaunchKernelGGL(k_cgm_unpack, flat(std::max(ctx->nu, ctx->nv)), dim3(256), 0, ctx->stream, g, Lo, w.xc.u,
                       w.xc.v);
    HIPCHK(hipGetLastError());
    // A stop seen between chunks needs no wait for the launches queued past
    // it (they change nothing); the timing events do.
    if (!hf->done || ctx->prof) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        prof_collect(ctx);
    }
    // launch q iterates when q <= iters
    burst_collect(ctx, WAVI_KT_CG_ITER_MIXED, hf->done ? hf->iters : -1);
    if (!hf->done) {
        ctx->err = "coarse CG did not terminate";
        return WAVI_ESTATE;
    }
    *iters = hf->iters;
    ctx->cg_prev = hf->iters;
    return WAVI_OK;
}

// precondition!(x, p, b) with x = 0, in the arithmetic of T: b in w.r,
// correction into w.e (update_preconditioners.jl:54-83).  vcycle_pre runs it
// up to the coarse right-hand side b_c = R (r - A e) in w.cr: diagonal,
// pre-smoothing, residual, restriction.
template <typename T>
static int vcycle_pre(wavi_ctx *ctx, const VSet<T> &w)
{
    static const int fwd[4] = {1, 3, 2, 4};
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby);
    CHK(do_diag(ctx, w));
    hipLaunchKernelGGL(k_fill_uv<T>, flat_uv(ctx->nu, ctx->nv), dim3(256), 0, ctx->stream, w.e.u, ctx->nu, w.e.v,
                       ctx->nv, T(0));
    CHK(do_smoother(ctx, w, fwd));
    // rho = r - A e
    hipLaunchKernelGGL(k_tile_resid<T>, dim3(band_blocks(g.nbx, g.nby)), dim3(256), 0, ctx->stream, g, coef_of(ctx, w),
                       w.e.u, w.e.v, w.r.u, w.r.v, w.rho.u, w.rho.v, ctx->Pn, (double *)nullptr,
                       band_blocks(g.nbx, g.nby));
    // b_c = R rho, into the CG residual buffer
    launch_tile_haar<T>(ctx, 2, w.rho.u, w.rho.v, ctx->ui, ctx->vi, w.cr.u, w.cr.v, ctx->mwu, ctx->mwv);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// The V-cycle after its coarse solve: e += P x_c, post-smoothing.
template <typename T>
static int vcycle_post(wavi_ctx *ctx, const VSet<T> &w)
{
    static const int rev[4] = {4, 2, 3, 1};
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby);
    // e += P x_c
    launch_tile_haar<T>(ctx, 1, w.xc.u, w.xc.v, ctx->mwu, ctx->mwv, w.t.u, w.t.v, ctx->ui, ctx->vi, w.e.u, w.e.v);
    CHK(do_smoother(ctx, w, rev));
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

template <typename T>
static int do_vcycle_t(wavi_ctx *ctx, const VSet<T> &w, int *cg_iters)
{
    CHK(vcycle_pre(ctx, w));
    CHK(do_coarse_cg(ctx, w, cg_iters));
    return vcycle_post(ctx, w);
}

// Conversions of several arrays, one launch.
template <typename Tin, typename Tout>
struct Conversions {
    ConvList<Tin, Tout> L{};
    int count = 0, nmax = 0;
    void add(const Tin *in, Tout *out, int n)
    {
        L.in[count] = in;
        L.out[count] = out;
        L.n[count] = n;
        ++count;
        nmax = n > nmax ? n : nmax;
    }
    void launch(wavi_ctx *ctx) const
    {
        hipLaunchKernelGGL((k_convert_list<Tin, Tout>), dim3((unsigned)((nmax + 255) / 256), (unsigned)count),
                           dim3(256), 0, ctx->stream, L);
    }
};

// V-cycle on the FP64 Picard residual ctx->r; correction into ctx->e.  In
// mixed precision the operator coefficients, the residual and the warm start
// are rounded to FP32, the V-cycle runs in FP32 (reductions accumulated in
// FP64) and the correction and warm start are widened back.
static int vcycle_narrow(wavi_ctx *ctx)
{
    VSet<float> &w = ctx->v32;
    Conversions<double, float> down;
    down.add(ctx->heta_h, w.heta_h, ctx->nh);
    down.add(ctx->heta_c, w.heta_c, ctx->nc > 0 ? ctx->nc : 1);
    down.add(ctx->dimp, w.dimp, ctx->nh);
    down.add(ctx->f[WAVI_U_H], w.h.u, ctx->nu);
    down.add(ctx->f[WAVI_V_H], w.h.v, ctx->nv);
    down.add(ctx->f[WAVI_U_BETA_EFF], w.beff.u, ctx->nu);
    down.add(ctx->f[WAVI_V_BETA_EFF], w.beff.v, ctx->nv);
    down.add(ctx->r.u, w.r.u, ctx->nu);
    down.add(ctx->r.v, w.r.v, ctx->nv);
    down.add(ctx->f[WAVI_WU_CORRECTION_COARSE], w.xc.u, ctx->nu);
    down.add(ctx->f[WAVI_WV_CORRECTION_COARSE], w.xc.v, ctx->nv);
    down.launch(ctx);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int vcycle_widen(wavi_ctx *ctx)
{
    VSet<float> &w = ctx->v32;
    Conversions<float, double> up;
    up.add(w.e.u, ctx->e.u, ctx->nu);
    up.add(w.e.v, ctx->e.v, ctx->nv);
    up.add(w.xc.u, ctx->f[WAVI_WU_CORRECTION_COARSE], ctx->nu);
    up.add(w.xc.v, ctx->f[WAVI_WV_CORRECTION_COARSE], ctx->nv);
    up.launch(ctx);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int do_vcycle(wavi_ctx *ctx, int *cg_iters)
{
    if (ctx->precision == WAVI_PRECISION_FP64) return do_vcycle_t<double>(ctx, vset64(ctx), cg_iters);
    CHK(vcycle_narrow(ctx));
    VSet<float> &w = ctx->v32;
    CHK(do_vcycle_t<float>(ctx, w, cg_iters));
    return vcycle_widen(ctx);
}

// precondition!(model, ::BasicParallelSpec) (update_velocities.jl:50-83),
// in three parts: precondition_resid queues the residual r = b - A x, the
// residual fields and the copy of ||r||, ||b|| to the host;
// precondition_decide waits for them; precondition_update queues
// u = x + e after the V-cycle.
static int precondition_resid(wavi_ctx *ctx)
{
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby);
    hipLaunchKernelGGL((k_select_uv<double, double>), flat_uv(ctx->nu, ctx->nv), dim3(256), 0, ctx->stream,
                       ctx->f[WAVI_U_U], ctx->ui, ctx->x.u, ctx->nu, ctx->f[WAVI_V_V], ctx->vi, ctx->x.v, ctx->nv);
    CHK(do_rhs(ctx));
    // r = b - A x, with set_residual! (update_velocities.jl:428-433): the
    // residual fields at inner points
    hipLaunchKernelGGL(k_tile_resid<double>, dim3(band_blocks(g.nbx, g.nby)), dim3(256), 0, ctx->stream, g,
                       coef64(ctx), ctx->x.u, ctx->x.v, ctx->b.u, ctx->b.v, ctx->r.u, ctx->r.v, ctx->Pr0, ctx->Pb,
                       band_blocks(g.nbx, g.nby), ctx->f[WAVI_U_RESIDUAL], ctx->f[WAVI_V_RESIDUAL]);
    hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(256), 0, ctx->stream, ctx->Pr0, ctx->Pb, g.npart,
                       ctx->d_norms);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int precondition_decide(wavi_ctx *ctx, int *converged, double *rel)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    *rel = ctx->h_norms[0] / ctx->h_norms[1];
    *converged = *rel < ctx->S.tol_picard;
    return WAVI_OK;
}

static int precondition_update(wavi_ctx *ctx)
{
    hipLaunchKernelGGL(k_set_velocities_uv, flat_uv(ctx->nu, ctx->nv), dim3(256), 0, ctx->stream,
                       ctx->f[WAVI_U_U], ctx->x.u, ctx->e.u, ctx->ui, ctx->nu, ctx->f[WAVI_V_V], ctx->x.v,
                       ctx->e.v, ctx->vi, ctx->nv);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// The coarse warm start of a mixed-precision V-cycle in the Picard loop
// (set/get_correction_coarse, update_preconditioners.jl:195-216): carried
// over from the previous Picard iteration while the relative residual is at
// least CGM_ZERO_START_REL, zero below it.  The coarse operator is singular
// (the Haar space spans points the fine operator crops); carried in FP32,
// the warm start gathers components in that null space, whose rounding
// feeds back into r0 and holds the Picard residual near 2e-9 on irregular
// domains.  A zero start has none.  Far from convergence the carried start
// keeps the mixed solve on the FP64 one's trajectory, so it stays there.
// FP64 solves always carry it, as the reference does.
constexpr double CGM_ZERO_START_REL = 1e-7;

static int coarse_start(wavi_ctx *ctx, double rel)
{
    if (ctx->precision != WAVI_PRECISION_MIXED || !(rel < CGM_ZERO_START_REL)) return WAVI_OK;
    hipLaunchKernelGGL(k_fill<double>, flat(ctx->nu), dim3(256), 0, ctx->stream, ctx->f[WAVI_WU_CORRECTION_COARSE],
                       0.0, ctx->nu);
    hipLaunchKernelGGL(k_fill<double>, flat(ctx->nv), dim3(256), 0, ctx->stream, ctx->f[WAVI_WV_CORRECTION_COARSE],
                       0.0, ctx->nv);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int do_precondition(wavi_ctx *ctx, int *converged, double *rel, int *cg_iters)
{
    CHK(precondition_resid(ctx));
    CHK(precondition_decide(ctx, converged, rel));
    *cg_iters = 0;
    if (!*converged) {
        CHK(coarse_start(ctx, *rel));
        CHK(do_vcycle(ctx, cg_iters));
        CHK(precondition_update(ctx));
    }
    return WAVI_OK;
}

// update_velocities!; wait = false leaves the work queued on the stream
// (wavi_run_steps, whose later steps follow on the same stream).
static int do_update_velocities(wavi_ctx *ctx, wavi_solve_report *rep, bool wait = true)
{
    wavi_solve_report local;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof(*rep));
    int converged = 0, it = 0, total = 0;
    double rel = INFINITY;
    while (!converged && it < ctx->S.maxiter_picard) {
        it = it + 1;
        CHK(do_inner_update(ctx));
        int cg = 0;
        CHK(do_precondition(ctx, &converged, &rel, &cg));
        total += cg;
        if (it <= WAVI_REPORT_MAX_PICARD) {
            rep->coarse_cg_iters[it - 1] = cg;
            rep->rel_resid_hist[it - 1] = rel;
        }
    }
    if (wait) HIPCHK(hipStreamSynchronize(ctx->stream));
    rep->picard_iters = it;
    rep->converged = converged;
    rep->rel_resid = rel;
    rep->coarse_cg_total = total;
    // mixed precision: every Picard iteration is one FP64 residual refinement
    rep->refinements = ctx->precision == WAVI_PRECISION_MIXED ? it : 0;
    return WAVI_OK;
}

// Work vectors of do_linear_solve: direction p, q = A p, and dot partials.
struct LinScratch {
    DevVec p, q;
    double *part = nullptr;  // 2 * npart
};

// Canonical dots of [u; v] pairs (b == nullptr: sums of squares), summed on
// the device and returned to the host.
static int dots(wavi_ctx *ctx, const LinScratch &s, int count, const double *const *au, const double *const *av,
                const double *const *bu, const double *const *bv, double *out)
{
    const GridDims &g = ctx->g;
    dim3 tg(g.nbx, g.nby);
    for (int c = 0; c < count; ++c)
        hipLaunchKernelGGL(k_tile_dot<double>, tg, dim3(256), 0, ctx->stream, g, au[c], av[c], bu[c], bv[c],
                           s.part + (size_t)c * g.npart);
    if (count < 1 || count > 2) return WAVI_EARG;  // h_norms holds 2
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(256), 0, ctx->stream, s.part, g.npart, count, ctx->d_norms);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < count; ++c) out[c] = ctx->h_norms[c];
    return WAVI_OK;
}

// x = A^{-1} b at inner points (b zero elsewhere) by flexible CG in FP64
// (Polak-Ribiere direction updates, so that the preconditioner may vary)
// preconditioned by the V-cycle of update_velocities! in the context's
// precision, from x = 0, until ||b - A x|| <= tol ||b|| (recomputed from
// x at the end) or maxit V-cycles.  Each V-cycle starts its coarse CG from
// zero: the warm start of the Picard loop (the previous correction) would
// be orders of magnitude larger than the shrinking residual's correction,
// and cancelling it costs the digits the solve is after.
// 1 / |diag(A_c)| of the current coefficients into ctx->cdinv (and the
// diagonal into diag when given).
static int do_coarse_diag(wavi_ctx *ctx, const DevVec *diag)
{
    const GridDims &g = ctx->g;
    hipLaunchKernelGGL(k_coarse_diag, dim3(g.nbx * g.nby, 2 * TILE), dim3(256), 0, ctx->stream, g, coef64(ctx),
                       ctx->mwu, ctx->mwv, ctx->S.levels, ctx->cdinv.u, ctx->cdinv.v, diag ? diag->u : nullptr,
                       diag ? diag->v : nullptr);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// The V-cycles of one linear solve run the Jacobi-preconditioned coarse CG.
struct CgPreScope {
    wavi_ctx *ctx;
    CgPreScope(wavi_ctx *c, bool on) : ctx(c) { ctx->cg_pre = on; }
    ~CgPreScope() { ctx->cg_pre = false; }
};

static int do_linear_solve(wavi_ctx *ctx, const DevVec &b, const DevVec &x, const LinScratch &s, double tol,
                           int maxit, int *vcycles, int64_t *cg_iters, double *rel_out, bool *converged)
{
    const GridDims &g = ctx->g;
    const int nu = ctx->nu, nv = ctx->nv;
    dim3 tg(g.nbx, g.nby);
    // The solve is the flexible CG's, so its preconditioner is free: the
    // V-cycle's coarse CG is scaled by the coarse operator's diagonal, which
    // the coarse problem (Haar levels of very different scale) needs far
    // more than the reference's plain CG can give it.
    const bool jacobi = ctx->lin_jacobi;
    if (jacobi) CHK(do_coarse_diag(ctx, nullptr));
    CgPreScope pre(ctx, jacobi);
    int64_t cg_total = 0;
    auto vcycle = [&](int *cg) -> int {
        hipLaunchKernelGGL(k_fill<double>, flat(nu), dim3(256), 0, ctx->stream, ctx->f[WAVI_WU_CORRECTION_COARSE],
                           0.0, nu);
        hipLaunchKernelGGL(k_fill<double>, flat(nv), dim3(256), 0, ctx->stream, ctx->f[WAVI_WV_CORRECTION_COARSE],
                           0.0, nv);
        CHK(do_vcycle(ctx, cg));
        cg_total += *cg;
        return WAVI_OK;
    };
    hipLaunchKernelGGL(k_fill<double>, flat(nu), dim3(256), 0, ctx->stream, x.u, 0.0, nu);
    hipLaunchKernelGGL(k_fill<double>, flat(nv), dim3(256), 0, ctx->stream, x.v, 0.0, nv);
    auto true_rel = [&](double *rel) -> int {
        hipLaunchKernelGGL(k_tile_resid<double>, dim3(band_blocks(g.nbx, g.nby)), dim3(256), 0, ctx->stream, g,
                           coef64(ctx), x.u, x.v, b.u, b.v, ctx->r.u, ctx->r.v, ctx->Pr0, ctx->Pb,
                           band_blocks(g.nbx, g.nby));
        hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(256), 0, ctx->stream, ctx->Pr0, ctx->Pb, g.npart,
                           ctx->d_norms);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *rel = ctx->h_norms[1] > 0.0 ? ctx->h_norms[0] / ctx->h_norms[1] : 0.0;
        return WAVI_OK;
    };
    // FCG from the true residual of x; restarted (at most 3 times) when the
    // recursively updated residual met the tolerance but the true one does not
    double rel = 0.0, bn = 0.0;
    int it = 0;
    for (int restart = 0; restart < 4; ++restart) {
        CHK(true_rel(&rel));
        if (restart == 0) bn = ctx->h_norms[1];
        if (rel <= tol || it >= maxit) break;
        int cg = 0;
        CHK(vcycle(&cg));  // z = M r into ctx->e
        ++it;
        hipLaunchKernelGGL(k_fcg_dir, flat(nu + nv), dim3(256), 0, ctx->stream, nu, nv, 0.0, ctx->ui, ctx->vi,
                           ctx->e.u, ctx->e.v, s.p.u, s.p.v);
        double rz;
        {
            const double *au[1] = {ctx->r.u}, *av[1] = {ctx->r.v}, *bu[1] = {ctx->e.u}, *bv[1] = {ctx->e.v};
            CHK(dots(ctx, s, 1, au, av, bu, bv, &rz));
        }
        for (;;) {
            hipLaunchKernelGGL(k_op<double>, dim3(band_blocks(g.nbx, g.nby)), dim3(256), 0, ctx->stream, g,
                               coef64(ctx), s.p.u, s.p.v, s.q.u, s.q.v, band_blocks(g.nbx, g.nby));
            double pq;
            {
                const double *au[1] = {s.p.u}, *av[1] = {s.p.v}, *bu[1] = {s.q.u}, *bv[1] = {s.q.v};
                CHK(dots(ctx, s, 1, au, av, bu, bv, &pq));
            }
            const double alpha = rz / pq;
            hipLaunchKernelGGL(k_fcg_step, flat(nu + nv), dim3(256), 0, ctx->stream, nu, nv, alpha, x.u, x.v,
                               ctx->r.u, ctx->r.v, s.p.u, s.p.v, s.q.u, s.q.v);
            double rr;
            {
                const double *au[1] = {ctx->r.u}, *av[1] = {ctx->r.v}, *bu[1] = {nullptr}, *bv[1] = {nullptr};
                CHK(dots(ctx, s, 1, au, av, bu, bv, &rr));
            }
            if (sqrt(rr) <= tol * bn || it >= maxit) break;
            CHK(vcycle(&cg));
            ++it;
            double zz[2];
            {
                const double *au[2] = {ctx->e.u, ctx->e.u}, *av[2] = {ctx->e.v, ctx->e.v};
                const double *bu[2] = {s.q.u, ctx->r.u}, *bv[2] = {s.q.v, ctx->r.v};
                CHK(dots(ctx, s, 2, au, av, bu, bv, zz));
            }
            const double beta = -alpha * zz[0] / rz;
            rz = zz[1];
            hipLaunchKernelGGL(k_fcg_dir, flat(nu + nv), dim3(256), 0, ctx->stream, nu, nv, beta, ctx->ui, ctx->vi,
                               ctx->e.u, ctx->e.v, s.p.u, s.p.v);
            HIPCHK(hipGetLastError());
        }
        if (restart == 3) CHK(true_rel(&rel));
    }
    *vcycles += it;
    *cg_iters += cg_total;
    *rel_out = rel;
    *converged = rel <= tol;
    return WAVI_OK;
}

// Device storage of the adjoint: checkpoints of (u, v, eta, beta, f2) for
// npic + 1 states, the state adjoints and the per-cell work arrays.
// Views into the context's adjoint workspace (ctx->adj_base).
struct AdjStore {
    std::vector<double *> u, v, eta, beta, f2;
    double *ubar = nullptr, *vbar = nullptr, *lu = nullptr, *lv = nullptr;
    double *etab = nullptr, *betab = nullptr, *f2b = nullptr, *wcb = nullptr, *grad = nullptr;
    double *cub = nullptr, *cvb = nullptr, *uxb = nullptr, *vyb = nullptr, *shb = nullptr;
    LinScratch ls;
};

static int adj_alloc(wavi_ctx *ctx, int npic, AdjStore &A)
{
    const size_t nu = ctx->nu, nv = ctx->nv, nh = ctx->nh, ns3 = (size_t)ctx->ns3;
    const size_t per = nu + nv + ns3 + 2 * nh;
    const size_t npart = ctx->g.npart;
    const size_t total = per * (npic + 1) + 4 * (nu + nv) + ns3 + 4 * nh + 5 * nh + 2 * npart;
    // The workspace is allocated on the first gradient call and kept with the
    // context (grown when a call needs more): a free, which can wait for the
    // whole device, would otherwise stall the other members of a batch.
    if (ctx->adj_cap < total) {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        if (ctx->adj_base) HIPCHK(hipFree(ctx->adj_base));
        ctx->adj_base = nullptr;
        ctx->adj_cap = 0;
        HIPCHK(hipMalloc(&ctx->adj_base, total * sizeof(double)));
        ctx->adj_cap = total;
    }
    double *q = ctx->adj_base;
    auto take = [&](size_t n) {
        double *r = q;
        q += n;
        return r;
    };
    for (int k = 0; k <= npic; ++k) {
        A.u.push_back(take(nu));
        A.v.push_back(take(nv));
        A.eta.push_back(take(ns3));
        A.beta.push_back(take(nh));
        A.f2.push_back(take(nh));
    }
    A.ubar = take(nu);
    A.vbar = take(nv);
    A.lu = take(nu);
    A.lv = take(nv);
    A.etab = take(ns3);
    A.betab = take(nh);
    A.f2b = take(nh);
    A.wcb = take(nh);
    A.grad = take(nh);
    A.cub = take(nh);
    A.cvb = take(nh);
    A.uxb = take(nh);
    A.vyb = take(nh);
    A.shb = take(nh);
    A.ls.p.u = take(nu);
    A.ls.p.v = take(nv);
    A.ls.q.u = take(nu);
    A.ls.q.v = take(nv);
    A.ls.part = take(2 * npart);
    return WAVI_OK;
}

static int adj_ckpt(wavi_ctx *ctx, AdjStore &A, int k, bool save)
{
    const size_t nh = ctx->nh, ns3 = (size_t)ctx->ns3;
    struct {
        double *field, *slot;
        size_t n;
    } items[5] = {{ctx->f[WAVI_U_U], A.u[k], (size_t)ctx->nu},
                  {ctx->f[WAVI_V_V], A.v[k], (size_t)ctx->nv},
                  {ctx->f[WAVI_S_ETA], A.eta[k], ns3},
                  {ctx->f[WAVI_H_BETA], A.beta[k], nh},
                  {ctx->f[WAVI_H_QUAD_F2], A.f2[k], nh}};
    for (auto &it : items)
        HIPCHK(hipMemcpyAsync(save ? it.slot : it.field, save ? it.field : it.slot, it.n * sizeof(double),
                              hipMemcpyDeviceToDevice, ctx->stream));
    return WAVI_OK;
}

static int do_misfit_gradient(wavi_ctx *ctx, int npic, double tol, int maxit, const double *uobs,
                              const double *vobs, double *J, double *grad, wavi_adjoint_report *rep)
{
    const int nu = ctx->nu, nv = ctx->nv, nh = ctx->nh;
    memset(rep, 0, sizeof(*rep));
    rep->n_picard = npic;
    AdjStore A;
    CHK(adj_alloc(ctx, npic, A));
    const std::vector<uint8_t> &ui = ctx->h_ui, &vi = ctx->h_vi;
    // forward: npic Picard iterations with converged linear solves
    for (int k = 0; k < npic; ++k) {
        CHK(adj_ckpt(ctx, A, k, true));
        CHK(do_inner_update(ctx));
        CHK(do_rhs(ctx));
        double rel = 0.0;
        bool ok = false;
        CHK(do_linear_solve(ctx, ctx->b, ctx->x, A.ls, tol, maxit, &rep->fwd_vcycles, &rep->coarse_cg_iters, &rel, &ok));
        if (!ok) rep->unconverged++;
        rep->max_fwd_rel = std::max(rep->max_fwd_rel, rel);
        hipLaunchKernelGGL(k_copy_masked<double>, flat(nu), dim3(256), 0, ctx->stream, ctx->f[WAVI_U_U], ctx->x.u,
                           ctx->ui, nu);
        hipLaunchKernelGGL(k_copy_masked<double>, flat(nv), dim3(256), 0, ctx->stream, ctx->f[WAVI_V_V], ctx->x.v,
                           ctx->vi, nv);
        HIPCHK(hipGetLastError());
    }
    CHK(adj_ckpt(ctx, A, npic, true));
    // J and ubar on the host
    std::vector<double> U(nu), V(nv);
    HIPCHK(hipMemcpyAsync(U.data(), A.u[npic], nu * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(V.data(), A.v[npic], nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double js = 0.0;
    for (int q = 0; q < nu; ++q) {
        U[q] = ui[q] ? U[q] - uobs[q] : 0.0;
        js += U[q] * U[q];
    }
    for (int q = 0; q < nv; ++q) {
        V[q] = vi[q] ? V[q] - vobs[q] : 0.0;
        js += V[q] * V[q];
    }
    *J = 0.5 * js;
    if (!grad) return WAVI_OK;
    // reverse sweep
    HIPCHK(hipMemcpyAsync(A.ubar, U.data(), nu * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(A.vbar, V.data(), nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const size_t ns3 = (size_t)ctx->ns3;
    hipLaunchKernelGGL(k_fill<double>, flat((int)ns3), dim3(256), 0, ctx->stream, A.etab, 0.0, (int)ns3);
    hipLaunchKernelGGL(k_fill<double>, flat(3 * nh), dim3(256), 0, ctx->stream, A.betab, 0.0, 3 * nh);
    HIPCHK(hipGetLastError());
    const DevVec bar{A.ubar, A.vbar}, lam{A.lu, A.lv};
    for (int k = npic; k >= 1; --k) {
        CHK(adj_ckpt(ctx, A, k - 1, false));
        CHK(do_inner_update(ctx));
        double rel = 0.0;
        bool ok = false;
        CHK(do_linear_solve(ctx, bar, lam, A.ls, tol, maxit, &rep->adj_vcycles, &rep->coarse_cg_iters, &rel, &ok));
        if (!ok) rep->unconverged++;
        rep->max_adj_rel = std::max(rep->max_adj_rel, rel);
        AdjArgs a;
        a.p = inner_args(ctx);
        a.uk = A.u[k];
        a.vk = A.v[k];
        a.lu = A.lu;
        a.lv = A.lv;
        a.eta_old = A.eta[k - 1];
        a.beta_old = A.beta[k - 1];
        a.f2_old = A.f2[k - 1];
        a.etab = A.etab;
        a.betab = A.betab;
        a.f2b = A.f2b;
        a.wcb = A.wcb;
        a.cub = A.cub;
        a.cvb = A.cvb;
        a.uxb = A.uxb;
        a.vyb = A.vyb;
        a.shb = A.shb;
        a.ubar = A.ubar;
        a.vbar = A.vbar;
        hipLaunchKernelGGL(k_adj_h, flat(nh), dim3(256), 0, ctx->stream, ctx->g, a);
        hipLaunchKernelGGL(k_adj_uv, flat(nu + nv), dim3(256), 0, ctx->stream, ctx->g, a);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_adj_grad, flat(nh), dim3(256), 0, ctx->stream, nh, ctx->f[WAVI_H_WEERTMAN_C], A.wcb,
                       A.grad);
    HIPCHK(hipMemcpyAsync(grad, A.grad, nh * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    // leave the state of the last forward iteration
    CHK(adj_ckpt(ctx, A, npic - 1, false));
    CHK(do_inner_update(ctx));
    HIPCHK(hipMemcpyAsync(ctx->f[WAVI_U_U], A.u[npic], nu * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->f[WAVI_V_V], A.v[npic], nv * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

static GeomArgs geom_args(wavi_ctx *ctx)
{
    GeomArgs a;
    a.hm = ctx->hm;
    a.um = ctx->um;
    a.vm = ctx->vm;
    a.b = ctx->f[WAVI_H_B];
    a.h = ctx->f[WAVI_H_H];
    a.s = ctx->f[WAVI_H_S];
    a.haf = ctx->f[WAVI_H_HAF];
    a.q = ctx->quad;
    a.gfh = ctx->f[WAVI_H_GROUNDED_FRACTION];
    a.gfu = ctx->f[WAVI_U_GROUNDED_FRACTION];
    a.gfv = ctx->f[WAVI_V_GROUNDED_FRACTION];
    a.hu = ctx->f[WAVI_U_H];
    a.su = ctx->f[WAVI_U_S];
    a.hv = ctx->f[WAVI_V_H];
    a.sv = ctx->f[WAVI_V_S];
    a.wc = ctx->f[WAVI_H_WEERTMAN_C];
    a.dsdh = ctx->f[WAVI_H_DSDH];
    a.acc = ctx->f[WAVI_H_ACCUMULATION];
    a.mel = ctx->f[WAVI_H_BASAL_MELT];
    a.wc_param = ctx->f[WAVI_P_WEERTMAN_C];
    a.acc_param = ctx->f[WAVI_P_ACCUMULATION_RATE];
    a.sl = ctx->P.sea_level_wrt_geoid;
    a.ri = ctx->P.density_ice;
    a.rw = ctx->P.density_ocean;
    a.melt = ctx->P.uniform_melt_rate;
    return a;
}

static int do_update_geometry(wavi_ctx *ctx)
{
    GeomArgs a = geom_args(ctx);
    hipLaunchKernelGGL(k_surface_haf, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->g, a);
    hipLaunchKernelGGL(k_gf_h, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->g, a);
    hipLaunchKernelGGL(k_geom_uv, flat(ctx->nu + ctx->nv), dim3(256), 0, ctx->stream, ctx->g, a);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int do_update_wavelets(wavi_ctx *ctx)
{
    dim3 tg(ctx->g.nbx, ctx->g.nby);
    hipLaunchKernelGGL(k_tile_wavelets, tg, dim3(256), 0, ctx->stream, ctx->g, ctx->tab, ctx->f[WAVI_U_U],
                       ctx->f[WAVI_V_V], ctx->um, ctx->vm, ctx->f[WAVI_WU_WAVELETS], ctx->f[WAVI_WV_WAVELETS],
                       ctx->mwu, ctx->mwv, ctx->f[WAVI_WU_CORRECTION_COARSE], ctx->f[WAVI_WV_CORRECTION_COARSE],
                       ctx->S.wavelet_threshold);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

static int do_after_solve(wavi_ctx *ctx)
{
    hipLaunchKernelGGL(k_post_solve, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->g, ctx->hm, ctx->um, ctx->vm,
                       ctx->f[WAVI_U_U], ctx->f[WAVI_V_V], ctx->f[WAVI_H_BETA], ctx->f[WAVI_H_QUAD_F1],
                       ctx->f[WAVI_H_QUAD_F2], ctx->f[WAVI_U_H], ctx->f[WAVI_V_H], ctx->f[WAVI_H_ACCUMULATION],
                       ctx->f[WAVI_H_BASAL_MELT], ctx->f[WAVI_H_U], ctx->f[WAVI_H_V], ctx->f[WAVI_H_UB],
                       ctx->f[WAVI_H_VB], ctx->f[WAVI_H_US], ctx->f[WAVI_H_VS], ctx->f[WAVI_H_DHDT], ctx->rdx,
                       ctx->rdy);
    HIPCHK(hipGetLastError());
    CHK(do_update_wavelets(ctx));
    return WAVI_OK;
}

static int do_update_state(wavi_ctx *ctx, wavi_solve_report *rep, bool wait = true)
{
    CHK(do_update_geometry(ctx));
    CHK(do_update_velocities(ctx, rep, wait));
    return do_after_solve(ctx);
}

static int do_update_thickness(wavi_ctx *ctx, double dt)
{
    hipLaunchKernelGGL(k_update_thickness, flat(ctx->nh), dim3(256), 0, ctx->stream, ctx->nh, ctx->hm, ctx->hfix,
                       ctx->f[WAVI_H_H], ctx->f[WAVI_H_DHDT], ctx->P.minimum_thickness, dt);
    HIPCHK(hipGetLastError());
    return WAVI_OK;
}

// ---------------------------------------------------------------------------
// k_cgm_iter_batch at the Haar level of the members.
template <int PAR, bool INIT, bool PRE>
static void launch_cgm_batch_p(hipStream_t s, int levels, int nwg, const CGMBatch &B)
{
    switch (levels) {
    case 0: hipLaunchKernelGGL((k_cgm_iter_batch<PAR, INIT, 0, PRE>), dim3(nwg), dim3(256), 0, s, B); break;
    case 1: hipLaunchKernelGGL((k_cgm_iter_batch<PAR, INIT, 1, PRE>), dim3(nwg), dim3(256), 0, s, B); break;
    case 2: hipLaunchKernelGGL((k_cgm_iter_batch<PAR, INIT, 2, PRE>), dim3(nwg), dim3(256), 0, s, B); break;
    default: hipLaunchKernelGGL((k_cgm_iter_batch<PAR, INIT, 3, PRE>), dim3(nwg), dim3(256), 0, s, B); break;
    }
}

// k_cgm_iter_batch, Jacobi-preconditioned when the members' CGs are (pre).
template <int PAR, bool INIT>
static void launch_cgm_batch(hipStream_t s, int levels, int nwg, const CGMBatch &B, bool pre)
{
    if (pre)
        launch_cgm_batch_p<PAR, INIT, true>(s, levels, nwg, B);
    else
        launch_cgm_batch_p<PAR, INIT, false>(s, levels, nwg, B);
}

// Device tables of a lockstep batch and the events that order the members'
// streams around its shared coarse CG (ev[k]: member k's set-up, ev[n]: a
// batch's end).  LockTables only refers to them; LockstepRes owns them.
struct LockTables {
    CGMIterArgs *members = nullptr;
    int *group = nullptr, *first = nullptr;
    std::vector<hipEvent_t> ev;
};

struct LockstepRes : LockTables {
    ~LockstepRes()
    {
        {
            std::lock_guard<std::mutex> lock(g_graph_mu);
            if (members) (void)hipFree(members);
            if (group) (void)hipFree(group);
            if (first) (void)hipFree(first);
        }
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Tables and events of a batch of the n members ctxs (tables sized for all
// of them at once; n + 1 events).
static int lockstep_res_alloc(LockstepRes &R, wavi_ctx *const *ctxs, int n)
{
    size_t ngroups = 0;
    for (int k = 0; k < n; ++k) ngroups += (size_t)(ctxs[k]->nblk / 8);
    wavi_ctx *ctx = ctxs[0];
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        HIPCHK(hipMalloc(&R.members, n * sizeof(CGMIterArgs)));
        HIPCHK(hipMalloc(&R.group, ngroups * sizeof(int)));
        HIPCHK(hipMalloc(&R.first, n * sizeof(int)));
    }
    R.ev.assign(n + 1, nullptr);
    for (hipEvent_t &e : R.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return WAVI_OK;
}

struct LockMember {
    wavi_ctx *ctx;
    wavi_solve_report rep;
    int it = 0, total = 0, converged = 0, cg = 0;
    double rel = INFINITY;
};

// Coarse CGs of the members mem (each set up by cgm_prepare on its own
// stream and marked there by the event pre[j]) run together on the stream of
// mem[0]: one k_cgm_iter_batch launch per CG iteration over the tiles of all
// of them, in chunks as for one member (cg_chunk_len), until every member's
// CG has stopped.  On return the CGs are complete, iters[j] holds member j's
// iteration count, and every member's stream waits for the event done
// before its next work.  On an error *bad names the member concerned.
static int run_cg_batch(const std::vector<wavi_ctx *> &mem, const std::vector<hipEvent_t> &pre, hipEvent_t done,
                        LockTables &R, std::vector<int> &iters, int *bad)
{
    wavi_ctx *lead = mem[0];
    hipStream_t cs = lead->stream;
    std::vector<CGMIterArgs> hm(mem.size());
    std::vector<int> hf(mem.size()), hg;
    int nwg = 0, qmax = 0;
    for (size_t j = 0; j < mem.size(); ++j) {
        *bad = (int)j;
        wavi_ctx *ctx = mem[j];
        if (j > 0) HIPCHK(hipStreamWaitEvent(cs, pre[j], 0));
        hm[j] = cgm_iter_args(ctx);
        hf[j] = nwg;
        for (int q = 0; q < ctx->nblk / 8; ++q) hg.push_back((int)j);
        nwg += ctx->nblk;
        qmax = std::max(qmax, ctx->S.maxiter_coarse + 1);
        if (ctx->cg_pre != lead->cg_pre) {
            ctx->err = "members of one coarse-CG batch disagree on the Jacobi-preconditioned coarse CG";
            return WAVI_ESTATE;
        }
    }
    *bad = 0;
    wavi_ctx *ctx = lead;
    const bool jac = lead->cg_pre;
    HIPCHK(hipMemcpyAsync(R.members, hm.data(), hm.size() * sizeof(CGMIterArgs), hipMemcpyHostToDevice, cs));
    HIPCHK(hipMemcpyAsync(R.group, hg.data(), hg.size() * sizeof(int), hipMemcpyHostToDevice, cs));
    HIPCHK(hipMemcpyAsync(R.first, hf.data(), hf.size() * sizeof(int), hipMemcpyHostToDevice, cs));
    const CGMBatch B{R.members, R.group, R.first};
    const int levels = lead->S.levels;
    launch_cgm_batch<0, true>(cs, levels, nwg, B, jac);
    HIPCHK(hipGetLastError());
    // launch q performs iteration q - 1 of every member whose CG has not stopped
    int q = 1, c = 0;
    while (q <= qmax) {
        const int n = std::min(cg_chunk_len(c, 0), qmax - q + 1);
        for (int s = 0; s < n; ++s, ++q) {
            if (q & 1)
                launch_cgm_batch<1, false>(cs, levels, nwg, B, jac);
            else
                launch_cgm_batch<0, false>(cs, levels, nwg, B, jac);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(lead->cg_ev[c % 3], cs));
        if (c >= 1) {
            HIPCHK(hipEventSynchronize(lead->cg_ev[(c - 1) % 3]));
            bool all = true;
            for (wavi_ctx *m : mem) all = all && ((volatile CGMHostFlag *)m->h_flag)->done;
            if (all) break;
        }
        ++c;
    }
    HIPCHK(hipEventRecord(done, cs));
    HIPCHK(hipStreamSynchronize(cs));
    iters.assign(mem.size(), 0);
    for (size_t j = 0; j < mem.size(); ++j) {
        *bad = (int)j;
        wavi_ctx *ctx = mem[j];
        volatile CGMHostFlag *f = ctx->h_flag;
        if (!f->done) {
            ctx->err = "coarse CG did not terminate";
            return WAVI_ESTATE;
        }
        iters[j] = f->iters;
        if (j > 0) HIPCHK(hipStreamWaitEvent(ctx->stream, done, 0));
    }
    return WAVI_OK;
}

// A rendezvous of the coarse CGs of ensemble members that run concurrently,
// each driven by its own thread.  A member whose V-cycle reaches its coarse
// CG sets it up on its own stream and waits.  When every registered member
// waits, the thread that completed the set runs all of their CGs as one
// batch (run_cg_batch) and releases them.  Members register when a thread
// starts them and leave when it is done with them, so the set always
// completes: the last member to arrive, or to leave, runs the batch.
struct CgRendezvous {
    std::mutex mu;
    std::condition_variable cv;
    int active = 0;
    std::vector<wavi_ctx *> pending;
    LockstepRes res;  // tables for all members; ev[j]: member j's set-up, ev.back(): a batch's end
};

// Runs the waiting members' CGs; called with lk held, returns with it held.
static void rdv_run(CgRendezvous &R, std::unique_lock<std::mutex> &lk)
{
    std::vector<wavi_ctx *> batch;
    batch.swap(R.pending);
    lk.unlock();
    std::vector<hipEvent_t> pre;
    for (wavi_ctx *c : batch) pre.push_back(R.res.ev[c->rdv_index]);
    std::vector<int> iters;
    int bad = 0, rc;
    try {
        rc = run_cg_batch(batch, pre, R.res.ev.back(), R.res, iters, &bad);
    } catch (const std::exception &ex) {
        batch[0]->err = std::string("host exception: ") + ex.what();
        rc = WAVI_ESTATE;
        bad = 0;
    }
    lk.lock();
    for (size_t j = 0; j < batch.size(); ++j) {
        wavi_ctx *c = batch[j];
        c->rdv_iters = rc == WAVI_OK ? iters[j] : 0;
        c->rdv_rc = rc == WAVI_OK ? WAVI_OK : ((int)j == bad ? rc : WAVI_ESTATE);
        if (rc != WAVI_OK && (int)j != bad) c->err = "a batched coarse CG failed for another member";
        c->rdv_done = true;
    }
    R.cv.notify_all();
}

static void rdv_join(wavi_ctx *ctx, CgRendezvous *R, int index)
{
    std::lock_guard<std::mutex> lk(R->mu);
    ctx->rdv = R;
    ctx->rdv_index = index;
    ++R->active;
}

static void rdv_leave(wavi_ctx *ctx)
{
    CgRendezvous &R = *ctx->rdv;
    std::unique_lock<std::mutex> lk(R.mu);
    ctx->rdv = nullptr;
    --R.active;
    if (R.active > 0 && (int)R.pending.size() == R.active) rdv_run(R, lk);
}

static int rdv_coarse_cg(wavi_ctx *ctx, const VSet<float> &w, int *iters)
{
    CHK(cgm_prepare(ctx, w));
    CgRendezvous &R = *ctx->rdv;
    HIPCHK(hipEventRecord(R.res.ev[ctx->rdv_index], ctx->stream));
    {
        std::unique_lock<std::mutex> lk(R.mu);
        ctx->rdv_done = false;
        R.pending.push_back(ctx);
        if ((int)R.pending.size() == R.active) rdv_run(R, lk);
        R.cv.wait(lk, [&] { return ctx->rdv_done; });
    }
    CHK(ctx->rdv_rc);
    hipLaunchKernelGGL(k_cgm_unpack, flat(std::max(ctx->nu, ctx->nv)), dim3(256), 0, ctx->stream, ctx->g,
                       cgm_layout(ctx), w.xc.u, w.xc.v);
    HIPCHK(hipGetLastError());
    *iters = ctx->rdv_iters;
    return WAVI_OK;
}

// One Picard iteration of every member that still iterates: each member's
// inner update and residual on its own stream; then, for the members whose
// residual asks for a V-cycle, the V-cycle up to the coarse right-hand side
// on their streams, the coarse CGs of all of them together on the stream of
// member 0 (one k_cgm_iter_batch launch per iteration over all their tiles,
// until every member's CG has stopped), and the rest of each V-cycle and the
// velocity update back on the members' streams.  Returns the number of
// members that iterated, or a negative error code (*bad = the member).
static int lockstep_picard(std::vector<LockMember> &M, LockTables &R, int it, int *bad)
{
    std::vector<int> act, cgm;
    for (int k = 0; k < (int)M.size(); ++k)
        if (!M[k].converged && M[k].it < M[k].ctx->S.maxiter_picard) act.push_back(k);
    for (int k : act) {
        *bad = k;
        wavi_ctx *ctx = M[k].ctx;
        M[k].it = it;
        M[k].cg = 0;
        CHK(do_inner_update(ctx));
        CHK(precondition_resid(ctx));
    }
    for (int k : act) {
        *bad = k;
        CHK(precondition_decide(M[k].ctx, &M[k].converged, &M[k].rel));
        if (!M[k].converged) cgm.push_back(k);
    }
    if (!cgm.empty()) {
        std::vector<wavi_ctx *> mem;
        std::vector<hipEvent_t> pre;
        for (int k : cgm) {
            *bad = k;
            wavi_ctx *ctx = M[k].ctx;
            CHK(coarse_start(ctx, M[k].rel));
            CHK(vcycle_narrow(ctx));
            CHK(vcycle_pre(ctx, ctx->v32));
            CHK(cgm_prepare(ctx, ctx->v32));
            HIPCHK(hipEventRecord(R.ev[k], ctx->stream));
            mem.push_back(ctx);
            pre.push_back(R.ev[k]);
        }
        std::vector<int> iters;
        int j = 0;
        const int rc = run_cg_batch(mem, pre, R.ev[M.size()], R, iters, &j);
        if (rc != WAVI_OK) {
            *bad = cgm[j];
            return rc;
        }
        for (size_t jj = 0; jj < cgm.size(); ++jj) {
            const int k = cgm[jj];
            *bad = k;
            wavi_ctx *ctx = M[k].ctx;
            M[k].cg = iters[jj];
            const VSet<float> &w = ctx->v32;
            hipLaunchKernelGGL(k_cgm_unpack, flat(std::max(ctx->nu, ctx->nv)), dim3(256), 0, ctx->stream, ctx->g,
                               cgm_layout(ctx), w.xc.u, w.xc.v);
            HIPCHK(hipGetLastError());
            CHK(vcycle_post(ctx, w));
            CHK(vcycle_widen(ctx));
            CHK(precondition_update(ctx));
        }
    }
    for (int k : act) {
        LockMember &m = M[k];
        m.total += m.cg;
        if (it <= WAVI_REPORT_MAX_PICARD) {
            m.rep.coarse_cg_iters[it - 1] = m.cg;
            m.rep.rel_resid_hist[it - 1] = m.rel;
        }
    }
    return (int)act.size();
}

extern "C" {
#pragma GCC visibility push(default)

int wavi_update_velocities(wavi_ctx *ctx, wavi_solve_report *report)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_update_velocities(ctx, report));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_misfit_gradient(wavi_ctx *ctx, int n_picard, double tol_linear, int maxit_linear, const double *u_obs,
                         const double *v_obs, double *J, double *grad_logc, wavi_adjoint_report *report)
{
    if (!ctx) return WAVI_EARG;
    if (n_picard < 1 || maxit_linear < 0 || !(tol_linear >= 0.0) || !u_obs || !v_obs || !J) {
        ctx->err = "wavi_misfit_gradient: n_picard >= 1, maxit_linear >= 0, tol_linear >= 0, u_obs, v_obs and J "
                   "are required";
        return WAVI_EARG;
    }
    if (ctx->has_fixed) {
        ctx->err = "wavi_misfit_gradient: fixed velocities (u_isfixed / v_isfixed) are not supported";
        return WAVI_EARG;
    }
    wavi_adjoint_report local;
    return do_misfit_gradient(ctx, n_picard, tol_linear, maxit_linear, u_obs, v_obs, J, grad_logc,
                              report ? report : &local);
}

// Argument checks shared by the lockstep entry points (0 or WAVI_EARG, the
// message in g_create_error and, where a context is at fault, its error).
static int lockstep_check(wavi_ctx **ctxs, int n)
{
    if (n < 0 || (n > 0 && !ctxs)) return WAVI_EARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return WAVI_EARG;
    if (batch_has_duplicates(ctxs, n)) return WAVI_EARG;
    for (int k = 0; k < n; ++k)
        if (ctxs[k]->precision != WAVI_PRECISION_MIXED || ctxs[k]->device != ctxs[0]->device ||
            ctxs[k]->S.levels != ctxs[0]->S.levels) {
            ctxs[k]->err = g_create_error =
                "lockstep batch: every context must be mixed precision, on one device, with one Haar level count";
            return WAVI_EARG;
        }
    return WAVI_OK;
}

// The Picard iterations of a lockstep batch of n >= 1 checked members with
// the tables and events T; fills reports and codes.
static int lockstep_run(wavi_ctx **ctxs, int n, LockTables &T, int rc, wavi_solve_report *reports, int *codes)
{
    std::vector<LockMember> M(n);
    for (int k = 0; k < n; ++k) {
        M[k].ctx = ctxs[k];
        memset(&M[k].rep, 0, sizeof(M[k].rep));
    }
    int bad = 0;
    for (int it = 1; rc == WAVI_OK; ++it) {
        const int r = lockstep_picard(M, T, it, &bad);
        if (r < 0) rc = r;
        if (r <= 0) break;
    }
    for (int k = 0; k < n; ++k) {
        LockMember &m = M[k];
        const int code = rc == WAVI_OK ? WAVI_OK : (k == bad ? rc : WAVI_ESTATE);
        if (rc != WAVI_OK && k != bad) m.ctx->err = "lockstep batch stopped by another member's error";
        if (hipStreamSynchronize(m.ctx->stream) != hipSuccess && rc == WAVI_OK) rc = WAVI_EHIP;
        m.rep.picard_iters = m.it;
        m.rep.converged = m.converged;
        m.rep.rel_resid = m.rel;
        m.rep.coarse_cg_total = m.total;
        m.rep.refinements = m.it;
        if (reports) reports[k] = m.rep;
        if (codes) codes[k] = code;
    }
    if (rc != WAVI_OK && g_create_error.empty()) g_create_error = ctxs[bad]->err;
    return rc;
}

int wavi_update_velocities_lockstep(wavi_ctx **ctxs, int n, wavi_solve_report *reports, int *codes)
{
    g_create_error.clear();
    CHK(lockstep_check(ctxs, n));
    if (n == 0) return WAVI_OK;
    if (hipSetDevice(ctxs[0]->device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return WAVI_EHIP;
    }
    LockstepRes R;
    const int rc = lockstep_res_alloc(R, ctxs, n);
    return lockstep_run(ctxs, n, R, rc, reports, codes);
}

int wavi_update_velocities_lockstep_noalloc(wavi_ctx **ctxs, int n, wavi_solve_report *reports, int *codes)
{
    g_create_error.clear();
    if (n > WAVI_NOALLOC_MAX_MEMBERS) {
        g_create_error = "wavi_update_velocities_lockstep_noalloc: at most " +
                         std::to_string(WAVI_NOALLOC_MAX_MEMBERS) + " members (got " + std::to_string(n) + ")";
        return WAVI_EARG;
    }
    CHK(lockstep_check(ctxs, n));
    if (n == 0) return WAVI_OK;
    const int dev = ctxs[0]->device;
    // the tables in the static scratch: member arguments, workgroup group ->
    // member, first workgroup of each member, each section 256-byte aligned
    size_t ngroups = 0;
    for (int k = 0; k < n; ++k) ngroups += (size_t)(ctxs[k]->nblk / 8);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t off_group = up((size_t)n * sizeof(CGMIterArgs));
    const size_t off_first = up(off_group + ngroups * sizeof(int));
    const size_t need = off_first + (size_t)n * sizeof(int);
    if (need > WAVI_NOALLOC_SCRATCH_BYTES) {
        g_create_error = "wavi_update_velocities_lockstep_noalloc: the batch's coarse-CG tables need " +
                         std::to_string(need) + " bytes, more than the " +
                         std::to_string((size_t)WAVI_NOALLOC_SCRATCH_BYTES) + "-byte static scratch";
        return WAVI_EARG;
    }
    unsigned char *base = dev >= 0 && dev < MAX_DEVICES ? g_scratch_ptr[dev] : nullptr;
    if (!base) {
        g_create_error = "wavi_update_velocities_lockstep_noalloc: no scratch address for this device "
                         "(the contexts' creation did not set it up)";
        return WAVI_ESTATE;
    }
    if (hipSetDevice(dev) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return WAVI_EHIP;
    }
    LockTables T;
    T.members = (CGMIterArgs *)base;
    T.group = (int *)(base + off_group);
    T.first = (int *)(base + off_first);
    T.ev.resize(n + 1);
    for (int k = 0; k < n; ++k) T.ev[k] = ctxs[k]->ev_ls;
    T.ev[n] = ctxs[0]->ev_ls_done;
    return lockstep_run(ctxs, n, T, WAVI_OK, reports, codes);
}

// Members that join a batch's lockstep group: mixed precision, at most
// LOCK_TILES tiles, on the device and Haar level of the group's first
// member, while the group's arenas stay within LOCK_ARENA_BYTES.  The
// group's coarse CGs touch all of its arenas in every launch; kept within
// half the 256 MB Infinity Cache they are served from it, and small members
// gain most from sharing launches.  Larger members run one per thread.
constexpr int LOCK_TILES = 256;
constexpr size_t LOCK_ARENA_BYTES = (size_t)128 << 20;

static size_t arena_bytes(const wavi_ctx *ctx)
{
    return (size_t)CGM_NARRAY * ctx->ns * sizeof(float) + 2 * (size_t)ctx->ns;
}

// The members of a batch that share coarse-CG launches, in member order.
// For gradients (linear = true) they also share the setting of
// wavi_set_linear_coarse_jacobi, since every coarse CG of a gradient is one
// of a linear solve and one launch serves one form of the CG.
static std::vector<int> lock_group(wavi_ctx *const *ctxs, int n, bool linear = false)
{
    std::vector<int> small;
    size_t bytes = 0;
    for (int k = 0; k < n; ++k) {
        const wavi_ctx *c = ctxs[k], *ref = small.empty() ? c : ctxs[small[0]];
        if (c->precision == WAVI_PRECISION_MIXED && c->g.nbx * c->g.nby <= LOCK_TILES && c->device == ref->device &&
            c->S.levels == ref->S.levels && (!linear || c->lin_jacobi == ref->lin_jacobi) &&
            bytes + arena_bytes(c) <= LOCK_ARENA_BYTES) {
            small.push_back(k);
            bytes += arena_bytes(c);
        }
    }
    return small;
}

int wavi_update_velocities_batch(wavi_ctx **ctxs, int n, int nthreads,
                                 wavi_solve_report *reports, int *codes)
{
    if (n < 0 || nthreads < 1 || (n > 0 && !ctxs)) return WAVI_EARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return WAVI_EARG;
    if (batch_has_duplicates(ctxs, n)) return WAVI_EARG;
    std::vector<int> small = lock_group(ctxs, n), large;
    for (int k = 0; k < n; ++k)
        if (std::find(small.begin(), small.end(), k) == small.end()) large.push_back(k);
    if (small.size() < 2) {
        return run_members(ctxs, n, nthreads, codes, [&](wavi_ctx *ctx, int k) {
            return wavi_update_velocities(ctx, reports ? &reports[k] : nullptr);
        });
    }
    // the lockstep group on one thread, the other members on the others
    std::vector<wavi_ctx *> sc, lc;
    for (int k : small) sc.push_back(ctxs[k]);
    for (int k : large) lc.push_back(ctxs[k]);
    std::vector<wavi_solve_report> srep(sc.size()), lrep(lc.size());
    std::vector<int> scode(sc.size(), WAVI_OK), lcode(lc.size(), WAVI_OK);
    int src = WAVI_OK, lrc = WAVI_OK;
    auto run_small = [&]() {
        try {
            src = wavi_update_velocities_lockstep(sc.data(), (int)sc.size(), srep.data(), scode.data());
        } catch (const std::exception &ex) {
            sc[0]->err = std::string("host exception: ") + ex.what();
            src = scode[0] = WAVI_ESTATE;
        }
    };
    auto run_large = [&](int threads) {
        lrc = run_members(lc.data(), (int)lc.size(), threads, lcode.data(), [&](wavi_ctx *ctx, int k) {
            return wavi_update_velocities(ctx, &lrep[k]);
        });
    };
    std::thread t;
    if (nthreads > 1 && !lc.empty()) {
        try {
            t = std::thread(run_small);
        } catch (const std::system_error &) {
        }
    }
    if (t.joinable()) {
        run_large(nthreads - 1);
        t.join();
    } else {
        run_small();
        run_large(nthreads);
    }
    for (size_t j = 0; j < small.size(); ++j) {
        if (reports) reports[small[j]] = srep[j];
        if (codes) codes[small[j]] = scode[j];
    }
    for (size_t j = 0; j < large.size(); ++j) {
        if (reports) reports[large[j]] = lrep[j];
        if (codes) codes[large[j]] = lcode[j];
    }
    if (src == WAVI_OK && lrc == WAVI_OK) return WAVI_OK;
    for (int k = 0; k < n; ++k) {
        const bool is_small = std::find(small.begin(), small.end(), k) != small.end();
        const int c = is_small ? scode[std::find(small.begin(), small.end(), k) - small.begin()]
                               : lcode[std::find(large.begin(), large.end(), k) - large.begin()];
        if (c != WAVI_OK) return c;
    }
    return src != WAVI_OK ? src : lrc;
}

int wavi_update_state_batch(wavi_ctx **ctxs, int n, int nthreads, wavi_solve_report *reports, int *codes)
{
    if (n < 0 || nthreads < 1 || (n > 0 && !ctxs)) return WAVI_EARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return WAVI_EARG;
    if (batch_has_duplicates(ctxs, n)) return WAVI_EARG;
    return run_members(ctxs, n, nthreads, codes, [&](wavi_ctx *ctx, int k) {
        return wavi_update_state(ctx, reports ? &reports[k] : nullptr);
    });
}

int wavi_misfit_gradient_batch(wavi_ctx **ctxs, int n, int nthreads, int n_picard, double tol_linear,
                               int maxit_linear, const double *const *u_obs, const double *const *v_obs, double *J,
                               double *const *grad_logc, wavi_adjoint_report *reports, int *codes)
{
    if (n < 0 || nthreads < 1 || (n > 0 && (!ctxs || !u_obs || !v_obs || !J))) return WAVI_EARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return WAVI_EARG;
    if (batch_has_duplicates(ctxs, n)) return WAVI_EARG;
    // the small mixed-precision members meet at every coarse CG, so the
    // members that run at the same time share the CG launches
    const std::vector<int> small = lock_group(ctxs, n, true);
    std::unique_ptr<CgRendezvous> R;
    std::vector<int> slot(n, -1);
    if (small.size() >= 2) {
        R.reset(new CgRendezvous);
        std::vector<wavi_ctx *> sc;
        for (size_t j = 0; j < small.size(); ++j) {
            slot[small[j]] = (int)j;
            sc.push_back(ctxs[small[j]]);
        }
        if (lockstep_res_alloc(R->res, sc.data(), (int)sc.size()) != WAVI_OK) R.reset();  // then alone
    }
    auto gradient = [&](wavi_ctx *ctx, int k) {
        const bool meet = R && slot[k] >= 0;
        if (meet) rdv_join(ctx, R.get(), slot[k]);
        int rc;
        try {
            rc = wavi_misfit_gradient(ctx, n_picard, tol_linear, maxit_linear, u_obs[k], v_obs[k], &J[k],
                                      grad_logc ? grad_logc[k] : nullptr, reports ? &reports[k] : nullptr);
        } catch (const std::exception &ex) {
            ctx->err = std::string("host exception: ") + ex.what();
            rc = WAVI_ESTATE;
        }
        if (meet) rdv_leave(ctx);
        return rc;
    };
    if (!R) return run_members(ctxs, n, nthreads, codes, gradient);
    // The small members get threads of their own, up to RDV_THREADS, beside
    // the nthreads that take the others: more of them then meet at each
    // coarse CG (a member waiting there holds its thread but no GPU time).
    constexpr int RDV_THREADS = 32;
    std::vector<int> large;
    for (int k = 0; k < n; ++k)
        if (slot[k] < 0) large.push_back(k);
    std::vector<wavi_ctx *> sc, lc;
    for (int k : small) sc.push_back(ctxs[k]);
    for (int k : large) lc.push_back(ctxs[k]);
    std::vector<int> scode(sc.size(), WAVI_OK), lcode(lc.size(), WAVI_OK);
    int src = WAVI_OK, lrc = WAVI_OK;
    auto run_small = [&]() {
        src = run_members(sc.data(), (int)sc.size(), std::min(RDV_THREADS, (int)sc.size()), scode.data(),
                          [&](wavi_ctx *ctx, int j) { return gradient(ctx, small[j]); });
    };
    auto run_large = [&]() {
        lrc = run_members(lc.data(), (int)lc.size(), nthreads, lcode.data(),
                          [&](wavi_ctx *ctx, int j) { return gradient(ctx, large[j]); });
    };
    std::thread t;
    if (!lc.empty()) {
        try {
            t = std::thread(run_large);
        } catch (const std::system_error &) {
        }
    }
    run_small();
    if (t.joinable())
        t.join();
    else
        run_large();
    for (size_t j = 0; j < small.size(); ++j)
        if (codes) codes[small[j]] = scode[j];
    for (size_t j = 0; j < large.size(); ++j)
        if (codes) codes[large[j]] = lcode[j];
    for (int k = 0; k < n; ++k) {
        const int c = slot[k] >= 0 ? scode[slot[k]] : lcode[std::find(large.begin(), large.end(), k) - large.begin()];
        if (c != WAVI_OK) return c;
    }
    return src != WAVI_OK ? src : lrc;
}

int wavi_inner_update(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_inner_update(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_update_geometry(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_update_geometry(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_update_state(wavi_ctx *ctx, wavi_solve_report *report)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_update_state(ctx, report));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

// Subdomain (ig, jg) (1-based) of an ngx x ngy restricted additive Schwarz
// partition with `overlap` cells: the bounds of schwarzModel
// (SharedMemorySpec.jl:183-199).  False when the grid does not split into
// whole subdomains.
static bool ras_part(const GridDims &g, int ig, int jg, int ngx, int ngy, int overlap, RasPart &p)
{
    if (g.nx % ngx != 0 || g.ny % ngy != 0) return false;
    const int nxd = g.nx / ngx, nyd = g.ny / ngy;
    const int i0 = std::max((ig - 1) * nxd + 1 - overlap, 1), i1 = std::min(ig * nxd + overlap, g.nx);
    const int j0 = std::max((jg - 1) * nyd + 1 - overlap, 1), j1 = std::min(jg * nyd + overlap, g.ny);
    p.i0 = i0 - 1;
    p.j0 = j0 - 1;
    p.nxg = i1 - i0 + 1;
    p.nyg = j1 - j0 + 1;
    p.leave_i1 = ig == 1;
    p.leave_im = ig == ngx;
    p.leave_j1 = jg == 1;
    p.leave_jn = jg == ngy;
    p.overlap = overlap;
    return true;
}

// Checks of a restricted additive Schwarz call: the subdomain contexts
// (subdomain (igrid, jgrid) at (igrid - 1) ngridsy + jgrid - 1) exist, are
// distinct, on the full model's device and on schwarzModel's grids; their
// parts into `parts`.
static int ras_check(wavi_ctx *full, wavi_ctx **subs, int ngridsx, int ngridsy, int overlap,
                     std::vector<RasPart> &parts)
{
    wavi_ctx *ctx = full;
    const int n = ngridsx * ngridsy;
    parts.assign(n, RasPart{});
    for (int ig = 1; ig <= ngridsx; ++ig)
        for (int jg = 1; jg <= ngridsy; ++jg) {
            const int k = (ig - 1) * ngridsy + (jg - 1);
            if (!subs[k] || subs[k] == full || subs[k]->device != full->device) {
                ctx->err = "restricted additive Schwarz: every subdomain needs a context of its own on the full "
                           "model's device";
                return WAVI_EARG;
            }
            if (!ras_part(full->g, ig, jg, ngridsx, ngridsy, overlap, parts[k])) {
                ctx->err = "Model domain is not an integer number of subdomains";
                return WAVI_EARG;
            }
            if (subs[k]->g.nx != parts[k].nxg || subs[k]->g.ny != parts[k].nyg || subs[k]->ns3 / subs[k]->nh != full->ns3 / full->nh) {
                ctx->err = "restricted additive Schwarz: subdomain " + std::to_string(k) + " has the wrong grid size";
                return WAVI_EDIM;
            }
        }
    if (batch_has_duplicates(subs, n)) {
        ctx->err = g_create_error;
        return WAVI_EARG;
    }
    return WAVI_OK;
}

int wavi_schwarz_refresh(wavi_ctx *full, wavi_ctx **subs, int ngridsx, int ngridsy, int overlap)
{
    g_create_error.clear();
    if (!full) return WAVI_EARG;
    wavi_ctx *ctx = full;
    if (!subs || ngridsx < 1 || ngridsy < 1 || overlap < 1) {
        ctx->err = "wavi_schwarz_refresh: subs is required; ngridsx, ngridsy and overlap must be at least 1";
        return WAVI_EARG;
    }
    std::vector<RasPart> parts;
    CHK(ras_check(full, subs, ngridsx, ngridsy, overlap, parts));
    HIPCHK(hipSetDevice(full->device));
    const int n = ngridsx * ngridsy;
    for (int k = 0; k < n; ++k) {
        wavi_ctx *sc = subs[k];
        RasRefresh r{};
        r.p = parts[k];
        r.nxf = full->g.nx;
        r.nyf = full->g.ny;
        r.ns = full->ns3 / full->nh;
        r.b = full->f[WAVI_H_B];
        r.h = full->f[WAVI_H_H];
        r.gf = full->f[WAVI_H_GROUNDED_FRACTION];
        r.wc = full->f[WAVI_P_WEERTMAN_C];
        r.acc = full->f[WAVI_P_ACCUMULATION_RATE];
        r.u = full->f[WAVI_U_U];
        r.v = full->f[WAVI_V_V];
        r.eta = full->f[WAVI_S_ETA];
        r.glenb = full->f[WAVI_S_GLEN_B];
        for (int id = 0; id < WAVI_NFIELDS; ++id) r.f[id] = sc->f[id];
        r.mwu = sc->mwu;
        r.mwv = sc->mwv;
        r.hm = sc->hm;
        const int m = std::max(std::max(sc->nu, sc->nv), sc->ns3);
        hipLaunchKernelGGL(k_ras_refresh, flat(m), dim3(256), 0, full->stream, r);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(full->stream));
    return WAVI_OK;
}

int wavi_schwarz_precondition(wavi_ctx *full, wavi_ctx **subs, int ngridsx, int ngridsy, int overlap,
                              int niterations, double damping, int nthreads, int *converged, double *rel_resid)
{
    g_create_error.clear();
    if (!full) return WAVI_EARG;
    wavi_ctx *ctx = full;
    if (!subs || !converged || !rel_resid || ngridsx < 1 || ngridsy < 1 || overlap < 1 || niterations < 0 ||
        nthreads < 1) {
        ctx->err = "wavi_schwarz_precondition: subs, converged and rel_resid are required; ngridsx, ngridsy, "
                   "overlap and nthreads must be at least 1 and niterations at least 0";
        return WAVI_EARG;
    }
    const int n = ngridsx * ngridsy;
    std::vector<RasPart> parts;
    CHK(ras_check(full, subs, ngridsx, ngridsy, overlap, parts));
    HIPCHK(hipSetDevice(full->device));
    int conv = 0;
    double rel = 0.0;
    CHK(precondition_resid(full));
    CHK(precondition_decide(full, &conv, &rel));
    const int nxuf = full->g.nxu, nxvf = full->g.nxv;
    for (int it = 0; !conv && it < niterations; ++it) {
        // schwarzRestrictVelocities! into every subdomain
        for (int k = 0; k < n; ++k) {
            const RasPart &p = parts[k];
            const int m = std::max((p.nxg + 1) * p.nyg, p.nxg * (p.nyg + 1));
            hipLaunchKernelGGL(k_ras_restrict, dim3((unsigned)((m + 255) / 256), 2), dim3(256), 0, full->stream, p,
                               nxuf, nxvf, full->f[WAVI_U_U], full->f[WAVI_V_V], subs[k]->f[WAVI_U_U],
                               subs[k]->f[WAVI_V_V]);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(full->stream));
        // update_state! of every subdomain, one stream and host thread each
        std::vector<int> codes(n, WAVI_OK);
        const int rc = run_members(subs, n, nthreads, codes.data(),
                                   [](wavi_ctx *c, int) { return wavi_update_state(c, nullptr); });
        if (rc != WAVI_OK) {
            for (int k = 0; k < n; ++k)
                if (codes[k] != WAVI_OK) {
                    ctx->err = "subdomain " + std::to_string(k) + ": " + subs[k]->err;
                    break;
                }
            return rc;
        }
        HIPCHK(hipSetDevice(full->device));
        // damping, then schwarzProlongVelocities! of the subdomains in order
        hipLaunchKernelGGL(k_scale_uv, flat_uv(full->nu, full->nv), dim3(256), 0, full->stream, full->f[WAVI_U_U],
                           full->nu, full->f[WAVI_V_V], full->nv, damping);
        for (int k = 0; k < n; ++k) {
            const RasPart &p = parts[k];
            const int m = std::max((p.nxg + 1) * p.nyg, p.nxg * (p.nyg + 1));
            hipLaunchKernelGGL(k_ras_prolong, dim3((unsigned)((m + 255) / 256), 2), dim3(256), 0, full->stream, p,
                               nxuf, nxvf, full->f[WAVI_U_U], full->f[WAVI_V_V], subs[k]->f[WAVI_U_U],
                               subs[k]->f[WAVI_V_V], 1.0 - damping);
        }
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(full->stream));
    *converged = conv;
    *rel_resid = rel;
    return WAVI_OK;
}

int wavi_after_solve(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_after_solve(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_update_thickness(wavi_ctx *ctx, double dt)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_update_thickness(ctx, dt));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_run_steps(wavi_ctx *ctx, int nsteps, wavi_solve_report *last_report)
{
    if (!ctx) return WAVI_EARG;
    if (nsteps < 0) {
        ctx->err = "nsteps must be non-negative";
        return WAVI_EARG;
    }
    for (int n = 0; n < nsteps; ++n) {
        CHK(do_update_state(ctx, last_report, false));
        CHK(do_update_thickness(ctx, ctx->P.dt));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

int wavi_update_wavelets(wavi_ctx *ctx)
{
    if (!ctx) return WAVI_EARG;
    CHK(do_update_wavelets(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAVI_OK;
}

static int upload2(wavi_ctx *ctx, const double *hu, const double *hv, DevVec &d)
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    HIPCHK(hipMemcpy(d.u, hu, ctx->nu * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.v, hv, ctx->nv * sizeof(double), hipMemcpyHostToDevice));
    return WAVI_OK;
}

static int download2(wavi_ctx *ctx, const DevVec &d, double *hu, double *hv)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lock(g_graph_mu);
    HIPCHK(hipMemcpy(hu, d.u, ctx->nu * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv, d.v, ctx->nv * sizeof(double), hipMemcpyDeviceToHost));
    return WAVI_OK;
}

int wavi_op_apply(wavi_ctx *ctx, const double *xu, const double *xv, double *yu, double *yv, int sampled)
{
    if (!ctx) return WAVI_EARG;
    if (!xu || !xv || !yu || !yv) {
        ctx->err = "null pointer";
        return WAVI_EARG;
    }
    CHK(upload2(ctx, xu, xv, ctx->fixedv));
    if (sampled) {
        hipLaunchKernelGGL((k_select<double, double>), flat(ctx->nu), dim3(256), 0, ctx->stream, ctx->fixedv.u,
                           ctx->ui, ctx->fixedv.u, ctx->nu);
        hipLaunchKernelGGL((k_select<double, double>), flat(ctx->nv), dim3(256), 0, ctx->stream, ctx->fixedv.v,
                           ctx->vi, ctx->fixedv.v, ctx->nv);
    }
    CHK(do_op(ctx, ctx->fixedv.u, ctx->fixedv.v, ctx->t.u, ctx->t.v));
    return download2(ctx, ctx->t, yu, yv);
}

int wavi_rhs(wavi_ctx *ctx, double *bu, double *bv)
{
    if (!ctx) return WAVI_EARG;
    if (!bu || !bv) {
        ctx->err = "null pointer";
        return WAVI_EARG;
    }
    CHK(do_rhs(ctx));
    return download2(ctx, ctx->b, bu, bv);
}

int wavi_op_diag(wavi_ctx *ctx, double *du, double *dv)
{
    if (!ctx) return WAVI_EARG;
    if (!du || !dv) {
        ctx->err = "null pointer";
        return WAVI_EARG;
    }
    CHK(do_diag(ctx, vset64(ctx)));
    return download2(ctx, ctx->d, du, dv);
}

int wavi_coarse_diag(wavi_ctx *ctx, double *du, double *dv)
{
    if (!ctx) return WAVI_EARG;
    if (!du || !dv) {
        ctx->err = "null pointer";
        return WAVI_EARG;
    }
    CHK(do_coarse_diag(ctx, &ctx->t));
    return download2(ctx, ctx->t, du, dv);
}

int wavi_set_linear_coarse_jacobi(wavi_ctx *ctx, int on)
{
    if (!ctx) return WAVI_EARG;
    ctx->lin_jacobi = on != 0;
    return WAVI_OK;
}

int wavi_vcycle(wavi_ctx *ctx, const double *ru, const double *rv, double *eu, double *ev, int32_t *coarse_iters)
{
    if (!ctx) return WAVI_EARG;
    if (!ru || !rv || !eu || !ev) {
        ctx->err = "null pointer";
        return WAVI_EARG;
    }
    CHK(upload2(ctx, ru, rv, ctx->r));
    int it = 0;
    CHK(do_vcycle(ctx, &it));
    if (coarse_iters) *coarse_iters = it;
    return download2(ctx, ctx->e, eu, ev);
}

int wavi_haar(wavi_ctx *ctx, int comp, int dir, const double *in, double *out)
{
    if (!ctx) return WAVI_EARG;
    if (!in || !out || comp < 0 || comp > 1 || dir < 0 || dir > 2) {
        ctx->err = "invalid argument";
        return WAVI_EARG;
    }
    const int n = comp == 0 ? ctx->nu : ctx->nv;
    double *din = comp == 0 ? ctx->fixedv.u : ctx->fixedv.v;
    double *dout = comp == 0 ? ctx->t.u : ctx->t.v;
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        HIPCHK(hipMemcpy(din, in, n * sizeof(double), hipMemcpyHostToDevice));
    }
    dim3 tg(ctx->g.nbx, ctx->g.nby);
    hipLaunchKernelGGL((k_tile_haar<double, double>), tg, dim3(256), 0, ctx->stream, ctx->g, ctx->tab, ctx->S.levels, dir,
                       ctx->fixedv.u, ctx->fixedv.v, (const uint8_t *)nullptr, (const uint8_t *)nullptr, ctx->t.u,
                       ctx->t.v, (const uint8_t *)nullptr, (const uint8_t *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lock(g_graph_mu);
    HIPCHK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost));
    return WAVI_OK;
}

int wavi_set_smoother_tile(wavi_ctx *ctx, int tx, int ty)
{
    if (!ctx) return WAVI_EARG;
    int shape = -1;
    for (int k = 0; k < SMOOTH_SHAPES; ++k)
        if (SMOOTH_TX[k] == tx && SMOOTH_TY[k] == ty) shape = k;
    if (tx == -1 && ty == -1) {
        const size_t w = ctx->precision == WAVI_PRECISION_MIXED ? sizeof(float) : sizeof(double);
        if (!smooth_grid_fits(ctx, w)) {
            ctx->err = "wavi_set_smoother_tile: the grid does not fit one workgroup (-1 x -1)";
            return WAVI_EARG;
        }
        shape = SMOOTH_GRID;
    } else if (shape < 0 && !(tx == 0 && ty == 0)) {
        ctx->err = "wavi_set_smoother_tile: the tile shapes are 16 x 16, 32 x 16 and 16 x 32 (0 x 0: tune; -1 x -1: "
                   "the whole grid in one workgroup)";
        return WAVI_EARG;
    }
    // a fixed shape is walked in column bands
    ctx->smooth_shape[0] = ctx->smooth_shape[1] = shape;
    return WAVI_OK;
}

int wavi_smoother_tile(const wavi_ctx *ctx, int *tx, int *ty)
{
    if (!ctx || !tx || !ty) return WAVI_EARG;
    const int pi = ctx->precision == WAVI_PRECISION_MIXED ? 1 : 0;
    const int sh = ctx->smooth_shape[pi];
    if (sh == SMOOTH_GRID) {
        *tx = *ty = -1;
        return WAVI_OK;
    }
    *tx = sh >= 0 ? SMOOTH_TX[sh % SMOOTH_SHAPES] : 0;
    *ty = sh >= 0 ? SMOOTH_TY[sh % SMOOTH_SHAPES] : 0;
    return WAVI_OK;
}

int wavi_set_coarse_cg(wavi_ctx *ctx, int mode)
{
    if (!ctx) return WAVI_EARG;
    if (mode != WAVI_CG_AUTO && mode != WAVI_CG_LAUNCHES && mode != WAVI_CG_WORKGROUP) {
        ctx->err = "wavi_set_coarse_cg: the modes are WAVI_CG_AUTO, WAVI_CG_LAUNCHES and WAVI_CG_WORKGROUP";
        return WAVI_EARG;
    }
    ctx->cg_mode = mode;
    return WAVI_OK;
}

int wavi_coarse_cg(const wavi_ctx *ctx, int *mode)
{
    if (!ctx || !mode) return WAVI_EARG;
    *mode = cg_one_workgroup(ctx) ? WAVI_CG_WORKGROUP : WAVI_CG_LAUNCHES;
    return WAVI_OK;
}

int wavi_profile_enable(wavi_ctx *ctx, int enable)
{
    if (!ctx) return WAVI_EARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    ctx->bursts.clear();
    ctx->prof = enable != 0;
    if (enable)
        for (int k = 0; k < WAVI_KT_COUNT; ++k) {
            ctx->prof_ms[k] = 0.0;
            ctx->prof_n[k] = 0;
            ctx->prof_seen[k] = 0;
        }
    return WAVI_OK;
}

int wavi_profile_read(wavi_ctx *ctx, int kernel, double *total_ms, int64_t *launches)
{
    if (!ctx || kernel < 0 || kernel >= WAVI_KT_COUNT || !total_ms || !launches) return WAVI_EARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    *total_ms = ctx->prof_ms[kernel];
    *launches = ctx->prof_n[kernel];
    return WAVI_OK;
}

int wavi_math_probe(int fn, const double *x, const double *y, double *out, size_t n)
{
    if (fn < 0 || fn > 2 || !x || !out || (fn != 1 && !y) || n > (size_t)(1 << 28)) {
        g_create_error = "invalid argument";
        return WAVI_EARG;
    }
    if (n == 0) return WAVI_OK;
    std::lock_guard<std::mutex> lock(g_graph_mu);
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, 3 * n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && y) e = hipMemcpy(d + n, y, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math_probe, flat((int)n), dim3(256), 0, 0, fn, d, d + n, d + 2 * n, (int)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    if (e != hipSuccess) {
        g_create_error = std::string("wavi_math_probe: ") + hipGetErrorString(e);
        return WAVI_EHIP;
    }
    return WAVI_OK;
}

#pragma GCC visibility pop
}  // extern "C"
