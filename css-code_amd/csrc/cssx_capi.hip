// C-ABI entry points (include/cssx.h): context, plan cache, derivative launches.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/cssx.h"
#include "cssx_internal.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"

namespace cssx {

// ---------------------------------------------------------------------------------------------
// Plan cache: one device buffer per (device, plan key, precision), built on first use and kept
// until the process ends (plans are a few KB each).
// ---------------------------------------------------------------------------------------------
namespace {
std::mutex g_plan_mu;
std::map<std::tuple<int, PlanKey, int>, void *> g_plans;
}  // namespace

int host_plan(const PlanKey &key, std::vector<double> &plan) {
    if (key.variant == 1) {
        const int MR = key.seg;
        if (MR != 4 && MR != 8 && MR != 16) return CSSX_EINVAL;
        if ((key.n + MR - 1) / MR > kMaxLanes) return CSSX_EINVAL;
        build_plan_contig(key, MR, plan);
    } else if (key.variant == 2) {
        if (key.seg <= 0 || (key.n + key.seg - 1) / key.seg > kMaxLanes) return CSSX_EINVAL;
        build_plan_contig(key, key.seg, plan);
    } else {
        const int M = key.seg;
        if (M <= 0 || seg_count(key.n, M) > kMaxSeg) return CSSX_EINVAL;
        build_plan(key, seg_count(key.n, M), plan);
    }
    return CSSX_OK;
}

int get_plan(const PlanKey &key, int prec, const void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CSSX_EHIP;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto tk = std::make_tuple(dev, key, prec);
    auto it = g_plans.find(tk);
    if (it != g_plans.end()) {
        *out = it->second;
        return CSSX_OK;
    }
    std::vector<double> plan;
    int rc = host_plan(key, plan);
    if (rc) return rc;
    void *d = nullptr;
    size_t bytes = plan.size() * (prec == CSSX_PREC_F64 ? sizeof(double) : sizeof(float));
    if (hipMalloc(&d, bytes) != hipSuccess) return CSSX_ENOMEM;
    hipError_t e;
    if (prec == CSSX_PREC_F64) {
        e = hipMemcpy(d, plan.data(), bytes, hipMemcpyHostToDevice);
    } else {
        std::vector<float> pf(plan.begin(), plan.end());
        e = hipMemcpy(d, pf.data(), bytes, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        (void)hipFree(d);
        return CSSX_EHIP;
    }
    g_plans[tk] = d;
    *out = d;
    return CSSX_OK;
}

// Pencil geometry of an axis on the reference layout f(nth, nphi, nr).
void axis_geometry(int axis, int nth, int nphi, int nr, AxisGeom &g) {
    const long sth = 1, sph = nth, sr = (long)nth * nphi;
    g.phase = axis == CSSX_AXIS_R ? CSSX_PHASE_DERIV_R : axis == CSSX_AXIS_THETA ? CSSX_PHASE_DERIV_THETA
                                                                                : CSSX_PHASE_DERIV_PHI;
    if (axis == CSSX_AXIS_R) {          // pencils (it, ip), lanes over theta
        g.n = nr; g.na = nth; g.nb = nphi; g.sa = sth; g.sb = sph; g.srow = sr;
        g.g_k = (long)nth * nphi; g.g_a = 1; g.g_b = nth;
    } else if (axis == CSSX_AXIS_THETA) {   // pencils (ip, ir), lanes over phi
        g.n = nth; g.na = nphi; g.nb = nr; g.sa = sph; g.sb = sr; g.srow = sth;
        g.g_k = 1; g.g_a = 4; g.g_b = 4L * nphi;
    } else {                            // pencils (it, ir), lanes over theta
        g.n = nphi; g.na = nth; g.nb = nr; g.sa = sth; g.sb = sr; g.srow = sph;
        g.g_k = nth; g.g_a = 1; g.g_b = 4L * nth;
    }
}

// The contiguous kernel serves pencils with unit row stride, and strided pencils longer than
// 512 rows, whose strided segments (64 rows) would not fit in registers: there each lane reads
// its MR rows with the row stride, and neighbouring pencils of the same workgroup share the
// cache lines.  Its D2 upper-wall closure needs five rows below row n-1 inside one lane's
// window, i.e. the last lane must own >= 2 rows.
static bool use_contig(const AxisGeom &g, int order, int hi_kind, bool f64) {
    const int MR = choose_MR(g.n, f64);
    if (MR < 0 || (g.srow != 1 && g.n <= 512)) return false;
    if (order == 2 && hi_kind == END_WALL && g.n % MR == 1) return false;
    return true;
}

static int lanes_per_pencil(int n, int MR) {
    const int nl = (n + MR - 1) / MR;
    int Lw = 1;
    while (Lw < nl) Lw *= 2;
    return Lw;
}

template <typename T>
static void fill_launch_geom(DerivLaunch<T> &L, const AxisGeom &g) {
    L.n = g.n;
    L.M = strided_M(g.n, sizeof(T) == 8);
    L.S = seg_count(g.n, L.M);
    L.na = g.na; L.nb = g.nb; L.sa = g.sa; L.sb = g.sb; L.srow = g.srow;
    L.njobs = 0;
}

// Kernel family of a batched launch: 2 = the wide kernel (long strided pencils, several
// segments per wavefront), 1 = the contiguous kernel, 0 = the strided kernel; seg = its rows
// per segment (or per lane).  The wide and contiguous D2 upper-wall closures need the last
// segment to own >= 2 rows.
static void launch_family(const DerivSpec *sp, int nspec, const AxisGeom &g, bool f64, int &variant, int &seg) {
    int WM = 0, WG = 0;
    if (g.srow != 1) wide_shape(g.n, f64, WM, WG);
    bool wide = WM > 0;
    for (int i = 0; i < nspec && wide; i++)
        if (sp[i].order == 2 && sp[i].hi_kind == END_WALL && g.n % WM == 1) wide = false;
    bool contig = !wide;
    for (int i = 0; i < nspec; i++) contig = contig && use_contig(g, sp[i].order, sp[i].hi_kind, f64);
    variant = wide ? 2 : contig ? 1 : 0;
    seg = wide ? WM : contig ? choose_MR(g.n, f64) : strided_M(g.n, f64);
}

PlanKey launch_plan_key(const DerivSpec *specs, int nspec, int i, const AxisGeom &g, bool f64) {
    int variant = 0, seg = 0;
    launch_family(specs, nspec, g, f64, variant, seg);
    const DerivSpec &s = specs[i];
    PlanKey key{g.n, s.order, s.lo_kind == END_WALL ? END_WALL : END_HALO,
                s.hi_kind == END_WALL ? END_WALL : END_HALO, (s.lo_kind == END_WALL || s.hi_kind == END_WALL) ? s.bdry : 0};
    key.variant = variant;
    key.seg = seg;
    return key;
}

// Appends spec i of a launch of `nspec` specs along g to L, with its plan (the spec's own or the
// cache's).
template <typename T>
static int add_job(DerivLaunch<T> &L, const DerivSpec *specs, int nspec, int i, const AxisGeom &g) {
    const bool f64 = sizeof(T) == 8;
    const DerivSpec &s = specs[i];
    const PlanKey key = launch_plan_key(specs, nspec, i, g, f64);
    const void *plan = s.plan;
    if (!plan) {
        int rc = get_plan(key, f64 ? CSSX_PREC_F64 : CSSX_PREC_F32, &plan);
        if (rc) return rc;
    }
    DerivJob<T> &J = L.jobs[L.njobs++];
    J.in = (const T *)s.in; J.out = (T *)s.out; J.rhs_out = (T *)s.rhs_out;
    J.glo = (const T *)s.glo; J.ghi = (const T *)s.ghi;
    J.g_k = s.g_k >= 0 ? s.g_k : g.g_k; J.g_a = s.g_k >= 0 ? s.g_a : g.g_a; J.g_b = s.g_k >= 0 ? s.g_b : g.g_b;
    J.plan = (const T *)plan;
    J.pre = (const T *)s.pre; J.post = (const T *)s.post;
    J.mul_a = (const T *)s.mul_a; J.mul_b = (const T *)s.mul_b;
    J.tap_in = (T *)s.tap_in; J.tap_out = (T *)s.tap_out; J.tap_side = s.tap_side;
    J.scale = (T)s.scale; J.neu_mul = (T)s.neu_mul; J.halo_s = (T)s.halo_s;
    J.rhs_scale = (T)s.rhs_scale; J.beta = (T)s.beta;
    J.bdry = s.bdry; J.lo_kind = s.lo_kind; J.hi_kind = s.hi_kind; J.lo_r0_val = s.lo_r0_val;
    return CSSX_OK;
}

template <typename T>
static int deriv_impl(const DerivSpec &sp, const AxisGeom &g, int nspec, hipStream_t stream) {
    DerivLaunch<T> L;
    fill_launch_geom(L, g);
    const bool f64 = sizeof(T) == 8;
    int variant = 0, seg = 0;
    launch_family(&sp, nspec, g, f64, variant, seg);
    const bool wide = variant == 2, contig = variant == 1;
    if (contig) { L.M = seg; L.S = lanes_per_pencil(g.n, L.M); }
    for (int i = 0; i < nspec && !contig; i++)   // only the contiguous kernel writes end-row taps
        if ((&sp)[i].tap_in || (&sp)[i].tap_out) return CSSX_EINVAL;
    if (wide) { L.M = seg; L.S = (g.n + seg - 1) / seg; }
    double passes = 0.0;
    for (int i = 0; i < nspec; i++) {
        const DerivSpec &s = (&sp)[i];
        int rc = add_job(L, &sp, nspec, i, g);
        if (rc) return rc;
        // algorithmic passes over the field: read f, write f' (and read the old f' when
        // accumulating into it)
        passes += s.beta != 0.0 ? 3.0 : 2.0;
        if (L.njobs == kMaxJobs || i == nspec - 1) {
            const int phase = g.phase;
            const int tok = timing_begin(phase, stream);
            hipError_t e = wide ? launch_deriv_wide<T>(L, s.order, stream)
                                : contig ? launch_deriv_contig<T>(L, s.order, stream) : launch_deriv<T>(L, s.order, stream);
            timing_end(tok, stream, passes * g.n * (double)g.na * g.nb * sizeof(T));
            if (e != hipSuccess) return CSSX_EHIP;
            L.njobs = 0;
            passes = 0.0;
        }
    }
    return CSSX_OK;
}

// Batched derivative launch; all specs share the axis geometry and the order.
int run_derivs(const DerivSpec *specs, int nspec, const AxisGeom &g, int prec, hipStream_t stream) {
    if (nspec <= 0) return CSSX_OK;
    for (int i = 1; i < nspec; i++)
        if (specs[i].order != specs[0].order) return CSSX_EINVAL;
    if (choose_M(g.n) < 0 || g.n < 8) return CSSX_EINVAL;
    return prec == CSSX_PREC_F64 ? deriv_impl<double>(specs[0], g, nspec, stream)
                                 : deriv_impl<float>(specs[0], g, nspec, stream);
}

}  // namespace cssx

using namespace cssx;

struct cssx_ctx {
    cssx_grid grid;
    int device;
};

extern "C" int cssx_version(void) { return 100; }

extern "C" const char *cssx_strerror(int code) {
    switch (code) {
    case CSSX_OK: return "ok";
    case CSSX_ENOMEM: return "allocation failed";
    case CSSX_EINVAL: return "invalid argument (shape, boundary code, axis or pointer)";
    case CSSX_ENAN: return "non-finite value detected";
    case CSSX_EHIP: return "HIP runtime error";
    case CSSX_ESIZE: return "size beyond the fixed scratch of a no-allocation entry point";
    default: return "unknown error";
    }
}

extern "C" int cssx_ctx_create(const cssx_grid *grid, cssx_ctx **out) {
    if (!grid || !out) return CSSX_EINVAL;
    *out = nullptr;
    if (grid->nr < 8 || grid->nth < 8 || grid->nphi < 8) return CSSX_EINVAL;
    if (choose_M(grid->nr) < 0 || choose_M(grid->nth) < 0 || choose_M(grid->nphi) < 0) return CSSX_EINVAL;
    cssx_ctx *c = new (std::nothrow) cssx_ctx;
    if (!c) return CSSX_ENOMEM;
    c->grid = *grid;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return CSSX_EHIP;
    }
    // build every plan this grid can ask for, so that later calls do not allocate
    int ns[3] = {grid->nr, grid->nth, grid->nphi};
    for (int ax = 0; ax < 3; ax++)
        for (int order = 1; order <= 2; order++)
            for (int lo = 0; lo < 2; lo++)
                for (int hi = 0; hi < 2; hi++)
                    for (int bd = 0; bd < 5; bd++) {
                        if (ax == CSSX_AXIS_PHI && (lo == 0 || hi == 0)) continue;
                        if (lo == 1 && hi == 1 && bd > 0) continue;
                        PlanKey key{ns[ax], order, lo ? END_HALO : END_WALL, hi ? END_HALO : END_WALL, bd};
                        const void *p;
                        for (int var = 0; var < 2; var++) {
                            if (var == 1 && (choose_MR(ns[ax]) < 0 || (ax != CSSX_AXIS_THETA && ns[ax] <= 512)))
                                continue;
                            key.variant = var;
                            for (int prec = 0; prec < 2; prec++) {
                                key.seg = var == 0 ? strided_M(ns[ax], prec == CSSX_PREC_F64)
                                                   : choose_MR(ns[ax], prec == CSSX_PREC_F64);
                                int rc = get_plan(key, prec, &p);
                                if (rc) {
                                    delete c;
                                    return rc;
                                }
                            }
                        }
                        if (ax != CSSX_AXIS_THETA)
                            for (int prec = 0; prec < 2; prec++) {
                                int WM, WG;
                                wide_shape(ns[ax], prec == CSSX_PREC_F64, WM, WG);
                                if (WM == 0) continue;
                                key.variant = 2;
                                key.seg = WM;
                                int rc = get_plan(key, prec, &p);
                                if (rc) {
                                    delete c;
                                    return rc;
                                }
                            }
                    }
    *out = c;
    return CSSX_OK;
}

extern "C" void cssx_ctx_destroy(cssx_ctx *ctx) { delete ctx; }

extern "C" int cssx_deriv(cssx_ctx *ctx, int axis, int order, int bdry, int prec, const void *in, void *out,
                          const void *ghost_lo, const void *ghost_hi, void *stream) {
    if (!ctx || !in || !out || axis < 0 || axis > 2 || (order != 1 && order != 2) || bdry < 0 || bdry > 4 ||
        (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    const cssx_grid &gr = ctx->grid;
    AxisGeom g;
    axis_geometry(axis, gr.nth, gr.nphi, gr.nr, g);
    DerivSpec s;
    s.order = order;
    s.in = in; s.out = out;
    s.glo = ghost_lo; s.ghi = ghost_hi;
    s.bdry = bdry;
    double inv;
    if (axis == CSSX_AXIS_R) {
        s.lo_kind = gr.r_lo_wall ? END_WALL : END_HALO;
        s.hi_kind = gr.r_hi_wall ? END_WALL : END_HALO;
        inv = gr.dri;
    } else if (axis == CSSX_AXIS_THETA) {
        s.lo_kind = gr.th_lo_wall ? END_WALL : END_HALO;
        s.hi_kind = gr.th_hi_wall ? END_WALL : END_HALO;
        inv = gr.dti;
    } else {
        s.lo_kind = ghost_lo ? END_HALO : END_PERIODIC;
        s.hi_kind = ghost_hi ? END_HALO : END_PERIODIC;
        inv = gr.dpi;
        s.bdry = 0;
    }
    if ((s.lo_kind == END_HALO && !ghost_lo) || (s.hi_kind == END_HALO && !ghost_hi)) return CSSX_EINVAL;
    s.scale = order == 1 ? inv : inv * inv;
    // a prescribed wall derivative enters the rows divided by dri (D2bydt2 also uses dri, T6)
    s.neu_mul = order == 1 ? 1.0 / inv : 1.0 / gr.dri;
    return run_derivs(&s, 1, g, prec, (hipStream_t)stream);
}

// Pencils of 5..7 points (too short for the segmented solve, whose end closures need 8 rows):
// one thread per pencil assembles the rows of derivatives.py:37-160 and runs the Thomas
// recurrence of linalg.tri_solver (its=0, linalg.py:93-113) in registers.  A singular system
// (n = 5 with Dirichlet-type closures at both walls) divides by a zero pivot and gives +-inf,
// as the reference does.
template <int N>
__global__ __launch_bounds__(256) void k_compact_small(long npen, double dri, int ibc, int dtype,
                                                       const double *__restrict__ arr, const double *__restrict__ b1,
                                                       const double *__restrict__ b2, double *darr,
                                                       double *__restrict__ res) {
    // no fused multiply-adds: the rows and the recurrence round as the reference's numpy does,
    // which decides the sign of a zero pivot in the singular case
#pragma clang fp contract(off)
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npen) return;
    double a[N], d[N], lo[N], up[N];
    for (int i = 0; i < N; i++) {
        a[i] = arr[i * npen + p];
        d[i] = darr ? darr[i * npen + p] : 0.0;
        lo[i] = up[i] = 1.0 / 3.0;
    }
    double g1[4] = {0, 0, 0, 0}, g2[4] = {0, 0, 0, 0};
    const bool lo_halo = dtype == 1 || dtype == 3, hi_halo = dtype == 2 || dtype == 3;
    for (int k = 0; k < 4; k++) {
        if (lo_halo) g1[k] = b1[k * npen + p];
        if (hi_halo) g2[k] = b2[k * npen + p];
    }
    const double dli = ibc == 5 ? (double)N : (double)(N - 1);
    const double alpha = 1.0 / 3.0, alpha2 = 3.0 / 4.0, gamma2 = 1.0 / 8.0, alpha1 = 6.0, beta1 = 3.0;
    const double cst = 1.0 / (alpha2 - beta1 * gamma2), up1 = cst * (alpha1 * alpha2 - beta1);
    double c1 = dli * 7.0 / 9.0, c2 = dli / 36.0;
    if (lo_halo) {
        d[0] = c1 * (a[1] - g1[3]) + c2 * (a[2] - g1[2]);
        d[1] = c1 * (a[2] - a[0]) + c2 * (a[3] - g1[3]);
    }
    if (hi_halo) {
        d[N - 2] = c1 * (a[N - 1] - a[N - 3]) + c2 * (g2[0] - a[N - 4]);
        d[N - 1] = c1 * (g2[0] - a[N - 2]) + c2 * (g2[1] - a[N - 3]);
    }
    for (int i = 2; i <= N - 3; i++) d[i] = c1 * (a[i + 1] - a[i - 1]) + c2 * (a[i + 2] - a[i - 2]);
    // the 6th-order halo row, times the Python's extra dli/60 (derivatives.py:88-91, :137-141)
    auto halo_lo = [&]() {
        d[0] += -alpha * (-g1[0] + 9.0 * g1[1] - 45.0 * g1[2] + 45.0 * a[0] - 9.0 * a[1] + a[2]);
        d[0] *= dli / 60.0;
    };
    auto halo_hi = [&]() {
        d[N - 1] += -alpha * (-a[N - 3] + 9.0 * a[N - 2] - 45.0 * a[N - 1] + 45.0 * g2[1] - 9.0 * g2[2] + g2[3]);
        d[N - 1] *= dli / 60.0;
    };
    if (dtype < 3) {
        const double e1 = -43.0 / 96 * dli, e2 = -5.0 / 6.0 * dli, e3 = 9.0 / 8.0 * dli, e4 = 1.0 / 6.0 * dli,
                     e5 = -1.0 / 96.0 * dli;
        const double f1 = -10.0 / 3.0 * dli, f2 = -3.0 * dli, f3 = 6.0 * dli, f4 = 1.0 / 3.0 * dli;
        if (dtype == 1 || dtype == 0) {   // upper wall
            d[N - 2] = -(e1 * a[N - 1] + e2 * a[N - 2] + e3 * a[N - 3] + e4 * a[N - 4] + e5 * a[N - 5]);
            if (ibc == 0) d[N - 1] = 0.0;
            if (ibc == 1 || ibc == 2) d[N - 1] = -(f1 * a[N - 1] + f2 * a[N - 2] + f3 * a[N - 3] + f4 * a[N - 4]);
            if (dtype == 1) halo_lo();
            if (ibc == 1 || ibc == 2) d[N - 1] = cst * (alpha2 * d[N - 1] - beta1 * d[N - 2]);
            lo[N - 2] = alpha2;
            up[N - 2] = gamma2;
            lo[N - 1] = (ibc == 1 || ibc == 2) ? up1 : 0.0;
        }
        if (dtype == 2 || dtype == 0) {   // lower wall
            d[1] = e1 * a[0] + e2 * a[1] + e3 * a[2] + e4 * a[3] + e5 * a[4];
            if (ibc == 0) d[0] = 0.0;
            if (ibc == 1 || ibc == 3) d[0] = f1 * a[0] + f2 * a[1] + f3 * a[2] + f4 * a[3];
            if (dtype == 2) halo_hi();
            const bool odd = (ibc % 2) != 0;
            if (odd) d[0] = cst * (alpha2 * d[0] - beta1 * d[1]);
            up[0] = odd ? up1 : 0.0;
            up[1] = alpha2;
            lo[1] = gamma2;
        }
    } else {
        halo_lo();
        halo_hi();
    }
    lo[0] = 0.0;
    up[N - 1] = 0.0;
    if (darr)
        for (int i = 0; i < N; i++) darr[i * npen + p] = d[i];
    // Thomas (unit diagonal)
    double cp[N], dp[N];
    cp[0] = up[0];
    dp[0] = d[0];
    for (int i = 1; i <= N - 2; i++) {
        const double piv = 1.0 - cp[i - 1] * lo[i];
        cp[i] = up[i] / piv;
        dp[i] = (d[i] - dp[i - 1] * lo[i]) / piv;
    }
    dp[N - 1] = (d[N - 1] - dp[N - 2] * lo[N - 1]) / (1.0 - cp[N - 2] * lo[N - 1]);
    double x = dp[N - 1];
    res[(N - 1) * npen + p] = x * dri;
    for (int i = N - 2; i >= 0; i--) {
        x = dp[i] - cp[i] * x;
        res[i * npen + p] = x * dri;
    }
}

// Argument checks, the register-resident kernel for n < 8, and the derivative spec for n >= 8
// shared by cssx_compact_fd6 and cssx_compact_fd6_noalloc.  Returns CSSX_OK with *done set
// when the short-pencil kernel already ran.
static int fd6_prepare(int n, long npencils, double dri, int ibc, int dtype, const double *arr, const double *b1,
                       const double *b2, double *darr, double *result, hipStream_t st, bool *done, DerivSpec &s,
                       AxisGeom &g) {
    *done = false;
    if (n < 5 || npencils < 1 || !arr || !result || ibc < 0 || ibc > 5 || dtype < 0 || dtype > 3 ||
        npencils > (long)INT32_MAX)
        return CSSX_EINVAL;
    const bool lo_halo = (dtype == 1 || dtype == 3), hi_halo = (dtype == 2 || dtype == 3);
    if ((lo_halo && !b1) || (hi_halo && !b2)) return CSSX_EINVAL;
    if (n < 8) {
        const dim3 grid((unsigned)((npencils + 255) / 256));
        if (n == 5)
            hipLaunchKernelGGL(k_compact_small<5>, grid, dim3(256), 0, st, npencils, dri, ibc, dtype, arr, b1, b2, darr, result);
        else if (n == 6)
            hipLaunchKernelGGL(k_compact_small<6>, grid, dim3(256), 0, st, npencils, dri, ibc, dtype, arr, b1, b2, darr, result);
        else
            hipLaunchKernelGGL(k_compact_small<7>, grid, dim3(256), 0, st, npencils, dri, ibc, dtype, arr, b1, b2, darr, result);
        *done = true;
        return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
    }
    if (choose_M(n) < 0) return CSSX_EINVAL;
    const double dli = (ibc == 5) ? (double)n : (double)(n - 1);
    // ibc 5: dli = n; at an upper wall it is Neumann with the preloaded value (not 1/2,
    // derivatives.py:82-115); at a lower wall it keeps the Dirichlet-type tridiagonal row
    // (5 is odd, :154-160) but its R0 is the preloaded darr[0] (no formula fills it, :134-135).
    const int bd = (ibc == 5) ? 3 : ibc;
    g.phase = CSSX_PHASE_OTHER;
    g.n = n; g.na = (int)npencils; g.nb = 1; g.sa = 1; g.sb = 0; g.srow = npencils;
    g.g_k = npencils; g.g_a = 1; g.g_b = 0;
    s.order = 1;
    s.in = arr; s.out = result; s.rhs_out = darr; s.rhs_scale = dli;
    s.lo_kind = lo_halo ? END_HALO : END_WALL;
    s.hi_kind = hi_halo ? END_HALO : END_WALL;
    // wall ends read the preloaded Neumann values (f'/dri) from darr rows 0 / n-1
    s.glo = lo_halo ? (const void *)b1 : (const void *)darr;
    s.ghi = hi_halo ? (const void *)b2 : (const void *)(darr ? darr + (long)(n - 1) * npencils : nullptr);
    s.bdry = bd;
    s.lo_r0_val = (ibc == 5);
    s.scale = dli * dri;
    s.neu_mul = 1.0 / dli;
    s.halo_s = dli / 60.0;   // derivatives.py:91,141,177,181 multiply the whole halo row by dli/60
    return CSSX_OK;
}

extern "C" int cssx_compact_fd6(int n, long npencils, double dri, int ibc, int dtype, const double *arr,
                                const double *b1, const double *b2, double *darr, double *result,
                                void *stream) {
    DerivSpec s;
    AxisGeom g;
    bool done = false;
    const int rc = fd6_prepare(n, npencils, dri, ibc, dtype, arr, b1, b2, darr, result, (hipStream_t)stream, &done, s, g);
    if (rc || done) return rc;
    return run_derivs(&s, 1, g, CSSX_PREC_F64, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------------------------
// No-allocation variant of cssx_compact_fd6.  Its only scratch, the FP64 plan table of the
// launch (5n + 80 doubles for the strided kernel, 5n + 960 for the wide one: at most 6080 for
// n <= 1024), lives in `scratch`, one fixed array per device.  The plan is built on the host
// into a host-side cache (kept for the process, so the asynchronous copy's source stays valid)
// and copied into the array on the caller's stream ahead of the launch.  All calls on one
// device share the array: calls that overlap on different streams of one device race.
// ---------------------------------------------------------------------------------------------
constexpr int kScratchDoubles = 8192;   // 64 KiB per device
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;                 // device -> address of its `scratch`
std::map<PlanKey, std::vector<double>> g_host_plans;   // FP64 plans, host memory only
}  // namespace

// the current device's address of `scratch`, looked up once per device
static int scratch_address(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CSSX_EHIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        double *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CSSX_EHIP;
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *out = it->second;
    return CSSX_OK;
}

extern "C" int cssx_compact_fd6_noalloc(int n, long npencils, double dri, int ibc, int dtype, const double *arr,
                                        const double *b1, const double *b2, double *darr, double *result,
                                        void *stream) {
    DerivSpec s;
    AxisGeom g;
    bool done = false;
    const hipStream_t st = (hipStream_t)stream;
    int rc = fd6_prepare(n, npencils, dri, ibc, dtype, arr, b1, b2, darr, result, st, &done, s, g);
    if (rc || done) return rc;
    double *ptr = nullptr;
    if ((rc = scratch_address(&ptr))) return rc;
    const PlanKey key = launch_plan_key(&s, 1, 0, g, true);
    const std::vector<double> *plan = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_host_plans.find(key);
        if (it == g_host_plans.end()) {
            std::vector<double> h;
            if ((rc = host_plan(key, h))) return rc;
            it = g_host_plans.emplace(key, std::move(h)).first;
        }
        plan = &it->second;
    }
    if (plan->size() > (size_t)kScratchDoubles) return CSSX_ESIZE;
    if (hipMemcpyAsync(ptr, plan->data(), plan->size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return CSSX_EHIP;
    s.plan = ptr;
    return run_derivs(&s, 1, g, CSSX_PREC_F64, st);
}

extern "C" int cssx_plan_tables(int n, int order, int lo_kind, int hi_kind, int bdry, int variant,
                                double *plan_out, int *nseg) {
    if (!plan_out || n < 8 || (order != 1 && order != 2) || bdry < 0 || bdry > 4 || lo_kind < 0 || lo_kind > 1 ||
        hi_kind < 0 || hi_kind > 1 || variant < 0 || variant > 1)
        return CSSX_EINVAL;
    std::vector<double> plan;
    PlanKey key{n, order, lo_kind, hi_kind, bdry};
    int S;
    if (variant == 0) {
        if (choose_M(n) < 0) return CSSX_EINVAL;
        S = seg_count(n, strided_M(n));
        build_plan(key, S, plan);
    } else {
        if (choose_MR(n) < 0) return CSSX_EINVAL;
        S = (n + choose_MR(n) - 1) / choose_MR(n);
        build_plan_contig(key, choose_MR(n), plan);
    }
    for (size_t i = 0; i < plan.size(); i++) plan_out[i] = plan[i];
    if (nseg) *nseg = S;
    return CSSX_OK;
}
