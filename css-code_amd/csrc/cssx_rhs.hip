// MHD right-hand side, Bound and RK4 stages on gfx950.
//
// RHS(call_num) of rhs_mpi.f90:279-783 (stratified, magnetic, do_radheat=F; with or without the
// Laplacian / SLaplacian explicit diffusion) is evaluated as:
//   1. curl A: six batched derivative sweeps whose epilogues apply the metric factors and add
//      the second term of each component onto the first, + k_curl_walls on the wall points
//      (Compute_B, derivatives.f90:1830-1892)
//   2. k_walls_state: wall values of the closures that depend on the state (rhs_mpi.f90:410-455,
//      :578-637)
//   3. the 19 first-derivative sweeps of Source_Theta_Terms / Phi_Terms / Radial_Terms, batched
//      per axis (one launch per axis, one job per field); with explicit diffusion also the 12
//      second derivatives, the 6 cross-term sweeps of first-derivative slots and the 6 sweeps
//      of curl B (Laplacian_Cross_Terms_*, VectorCurl)
//   4. main_terms / viscous_terms: every pointwise term of those three routines in the
//      reference's order, and the explicit-diffusion branches
//   5. k_walls_density: the density wall derivatives that zero the normal momentum tendency
//      (rhs_mpi.f90:1553-1564, :1750-1773); they need the terms of step 4 at the wall points,
//      which it evaluates there
//   6. two density sweeps + k_rhs_final: the terms of step 4 at every point, the density terms,
//      top relaxation, u x B (and -etar curl B with its Ohmic heating), xi_cs and the fused RK4
//      stage; no tendency field is written and re-read in between
// Temperature T = rho^(gamma-1) exp(s/Cv)/((gamma-1)Cv) and 1/rho are recomputed where used
// instead of being stored (rhs_mpi.f90:377-387).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cssx.h"
#include "cssx_internal.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"

// bracket one launch with the phase timer (no-op unless cssx_timing_enable(1))
#define CSSX_TIMED(phase, stream, bytes, ...)            \
    do {                                                 \
        const int tok_ = timing_begin((phase), (stream)); \
        __VA_ARGS__;                                     \
        timing_end(tok_, (stream), (double)(bytes));     \
    } while (0)

namespace cssx {

constexpr double kOneOver4Pi = 1.0 / 3.1415926535897932384626433832795028841972 / 4.0;
// 4 pi / c_light^2 of the Ohmic heating (c_light = 2.99792458d10, constants.f90:7)
constexpr double kJconst = 4.0 * 3.1415926535897932384626433832795028841972 / (2.99792458e10 * 2.99792458e10);

template <typename T>
struct Prof {
    T *r = nullptr, *rinv = nullptr, *dxdr = nullptr, *grav = nullptr, *eps = nullptr;
    T *drr = nullptr;   // dx/dr / r: the r sweeps of the curl

    T *dr = nullptr;    // cell widths dr_ind (Compute_Step_Size)
    T *sn = nullptr, *csc = nullptr, *cs = nullptr, *cot = nullptr;
    T *dsdr1 = nullptr, *dsdr2 = nullptr, *sr1 = nullptr;
    // hydrostatic background (nr): rho_bar, s_bar, T_bar and the physical radial derivatives
    // of rho_bar and s_bar by the compact operator with zero wall values
    T *rbar = nullptr, *sbar = nullptr, *tbar = nullptr, *drbar = nullptr, *dsbar = nullptr;
    // well-balanced radial momentum of the fluctuation form: hres = the background's radial
    // momentum residual -g - pconst (T_bar/rho_bar) drbar - (gamma-1) T_bar dsbar (zero on
    // Neumann density walls, truncation-small inside), tr = T_bar/rho_bar
    T *hres = nullptr, *tr = nullptr;
    // SLD face geometry (diffusion.f90:805-838, 1219-1250, 1316-1348): r face radii rfc (nr+1,
    // 0 on the walls so their flux vanishes), 3/(r_h^3-r_l^3) rv (nr), sin of the theta faces
    // sph (nth+1, 0 on the walls), 1.5(r_h^2-r_l^2)/(r_h^3-r_l^3) rfac (nr),
    // 1/(cos th_l - cos th_h) ct (nth), (th_h-th_l)/(cos th_l - cos th_h) gt (nth), and the
    // SLD speed^2 of each radius vsp (nr)
    T *rfc = nullptr, *rv = nullptr, *sph = nullptr, *rfac = nullptr, *ct = nullptr, *gt = nullptr, *vsp = nullptr;
    // explicit diffusion (nr): viscosity mur, d ln mur/dr, thermal diffusivity ks, d ln ks/dr,
    // resistivity etar, d2x/dr2 / dx/dr (Radial_Terms' work9), and the x-space second
    // derivative of s_bar (the background part of d2s/dx2 in the fluctuation form)
    T *mur = nullptr, *dlnmu = nullptr, *ks = nullptr, *dlnks = nullptr, *etar = nullptr, *w9 = nullptr,
      *d2sbar = nullptr;
    T *buf = nullptr;
};

// derivative scratch slots
enum Slot {
    D_TW = 0, D_TV, D_TU, D_TS, D_TBP, D_TBR,                 // theta: w v u s bp br
    D_PRHO, D_PV, D_PW, D_PU, D_PS, D_PBT, D_PBR,             // phi:   rho v w u s bt br
    D_RW, D_RV, D_RU, D_RS, D_RBT, D_RBP,                     // r:     w v u s bt bp
    D_THRHO, D_RRHO,                                          // density: theta, r
    D_NSLOT
};

// Where the theta first derivatives of one point are: w v u s bp br at p[k*stride], k = D_TW ..
// D_TBR (the derivative scratch: p = D + i, stride N).  Each value is read where the terms use
// it, so that no register holds it longer.
template <typename T>
struct ThetaD {
    const T *p;
    long stride;
    __device__ __forceinline__ T operator[](int k) const { return p[k * stride]; }
};
template <typename T>
__device__ __forceinline__ ThetaD<T> theta_from(const T *D, long N, long i) {
    return ThetaD<T>{D + i, N};
}

// explicit-diffusion scratch slots (Laplacian / SLaplacian)
enum LapSlot {
    L_TW2 = 0, L_TV2, L_TU2, L_TS2,          // theta second derivatives: w v u s
    L_PV2, L_PW2, L_PU2, L_PS2,              // phi second derivatives:   v w u s
    L_RW2, L_RV2, L_RU2, L_RS2,              // r second derivatives (x-space): w v u s
    L_XR1, L_XR2, L_XR3, L_XR4,              // cross terms: d/dr of dw/dth, du/dph, du/dth, dv/dph
    L_XT1, L_XT2,                            // cross terms: d/dth of dv/dph, dw/dph
    L_JT1, L_JT2, L_JP1, L_JP2, L_JRT, L_JRP,  // curl B: d/dth (sin B_phi), d/dth B_r, d/dph B_r,
                                             // d/dph B_theta, d/dr (r B_theta), d/dr (r B_phi)
    L_NSLOT
};

template <typename T>
struct Scratch {
    T *D = nullptr;      // D_NSLOT fields
    T *B = nullptr;      // 3 fields: B_theta, B_phi, B_r
    T *wall = nullptr;   // wall value buffers
    T *y1 = nullptr;     // RK stage state (nv fields)
    T *k = nullptr;      // RK stage tendency (nv fields)
    T *dif = nullptr;    // SLD diffusion tendency of stage 1 (nv fields, SLDiffusion runs)
    T *vcc = nullptr;    // SLD speed^2 per cell (1 field, SLDiffusion runs)
    // the theta-wall block: the state (5), B (3) and every derivative slot (D_NSLOT) at the
    // theta-wall points, [field][side][ir][ip] (side 0: theta row 0, 1: row nth-1), so that the
    // density wall closure reads them contiguously
    T *tw = nullptr;
    T *acc = nullptr;    // FP32 runs: RK4 tendency accumulator (nv fields)
    T *y0f = nullptr;    // FP32 runs: the FP64 state y0 in stage precision, fluctuation form
    T *L = nullptr;      // explicit diffusion: L_NSLOT fields
    T *lwall = nullptr;  // explicit diffusion: cross-term wall values (2 r, 2 theta buffers)
    bool ready = false;
};

}  // namespace cssx

struct cssx_mhd {
    cssx_grid grid;
    cssx_mhd_params p;
    int nr, nth, nphi, nv;
    long N;
    double Cv, gamm1, pconst, invpconst, invtconst, invCv;
    double dth, dphi;
    std::vector<double> h_r, h_dr, h_dxdr, h_theta, h_grav, h_eps, h_dsdr1, h_dsdr2, h_sr1, h_sr2, h_rho_bar,
        h_s_bar, h_diff_speed;
    // background: T_bar, compact radial derivatives of rho_bar / s_bar, radial momentum residual
    std::vector<double> h_tbar, h_drbar, h_dsbar, h_hres;
    // explicit diffusion: profiles and the compact second derivative of s_bar (x-space)
    std::vector<double> h_mur, h_dlnmu, h_ks, h_dlnks, h_etar, h_d2xdr2, h_d2sbar;
    double wbar_lo = 0.0, wbar_hi = 0.0;   // d rho_bar/dr imposed at Neumann density walls
    cssx::Prof<double> p64;
    cssx::Prof<float> p32;
    cssx::Scratch<double> s64;
    cssx::Scratch<float> s32;
    double *save = nullptr;   // FP64 RK accumulator (nv fields)
    double *red = nullptr;    // reduction partials
    int nred = 0;
    cssx_mlp *closure = nullptr;   // attached learned closure (not owned) or null
    float *ext = nullptr;          // its tendency (8 fields, FP32)
    // the closure tendency runs on a side stream (non-blocking, default priority), forked from
    // the RHS's stream at its start and joined before the final kernel, so the network's
    // matrix-core work overlaps the memory-bound derivative sweeps
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // RK4 stage 1's SLD diffusion runs on a stream of its own (non-blocking, default priority),
    // forked after Compute_B and joined before the final kernel: its VALU-bound column march
    // overlaps the memory-bound derivative sweeps
    hipStream_t sld_side = nullptr;
    hipEvent_t ev_sld_fork = nullptr, ev_sld_join = nullptr;
};

namespace cssx {

// ---------------------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Consts {
    int nth, nphi, nr;
    long N;
    T gamm1, invtconst, invCv, pconst, invpconst, Cv, omega0, om2, xi2, gamma, T0, damp;
    int magnetic, top_relax, xi_on;
    int fluct;   // state slots 1 and 5 hold rho - rho_bar(r) and s - s_bar(r) (FP32 path)
    int lap, slap;        // Laplacian; Laplacian or SLaplacian (thermal diffusion)
    T s_lap_lo, s_lap_hi; // constant-flux entropy wall values under slap: -lumr/(4 pi r^2 ks dxdr)
    T wbar_lo, wbar_hi;   // d rho_bar/dr prescribed at the Neumann density walls (0 elsewhere)
    T dT0top;             // T0 - T_bar at the top radius (FP64 difference)
};

template <typename T>
__device__ __forceinline__ T temperature(T rho, T s, const Consts<T> &c) {
    T t = c.invtconst * pow(rho, c.gamm1);
    return t * exp(s * c.invCv);
}
template <>
__device__ __forceinline__ float temperature<float>(float rho, float s, const Consts<float> &c) {
    float t = c.invtconst * powf(rho, c.gamm1);
    return t * expf(s * c.invCv);
}

// full density and temperature of point (ir) from the stored slots.  In the fluctuation form
// T = T_bar (1 + rho'/rho_bar)^(gamma-1) exp(s'/Cv), which equals
// invtconst rho^(gamma-1) exp(s/Cv) exactly but keeps FP32 relative precision.
template <typename T>
struct Thermo {
    T rho, t;
    T a, b;   // fluctuation form: T/rho = (T_bar/rho_bar)(1+a), T = T_bar (1+b)
};
template <typename T>
__device__ __forceinline__ Thermo<T> thermo(T rho_st, T s_st, int ir, const Consts<T> &c, const Prof<T> &P) {
    Thermo<T> o;
    if (c.fluct) {
        const T rb = P.rbar[ir];
        o.rho = rb + rho_st;
        // log1p/expm1 keep the relative precision of the small departures a, b themselves
        const T lq = log1p(rho_st / rb), e1 = s_st * c.invCv;
        o.b = expm1(c.gamm1 * lq + e1);
        o.a = expm1((c.gamm1 - T(1)) * lq + e1);
        o.t = P.tbar[ir] * (T(1) + o.b);
    } else {
        o.rho = rho_st;
        o.t = temperature(rho_st, s_st, c);
        o.a = o.b = T(0);
    }
    return o;
}

template <typename T>
__device__ __forceinline__ void coords(long i, const Consts<T> &c, int &it, int &ip, int &ir) {
    it = (int)(i % c.nth);
    long q = i / c.nth;
    ip = (int)(q % c.nphi);
    ir = (int)(q / c.nphi);
}

// ---------------------------------------------------------------------------------------------
// Bound (rhs_mpi.f90:2395-2491) as a per-point predicate, fused into the stage update
// ---------------------------------------------------------------------------------------------
struct BoundFlags {
    int theta_stress_free, bot_sf, top_sf, bot_ctemp, magnetic, perf_cond, bot_radial, top_radial;
};

template <typename T>
__device__ __forceinline__ T bound_value(int iv, int it, int ip, int ir, T val, const BoundFlags &f, int nth,
                                         int nr, const T *sr1) {
    const bool thw = (it == 0 || it == nth - 1);
    if (iv == 1 && thw) return T(0);                                     // w at theta walls
    if (iv == 3 && thw && !f.theta_stress_free) return T(0);
    if (ir == 0) {
        if ((iv == 1 || iv == 2) && !f.bot_sf) return T(0);
        if (iv == 3) return T(0);                                        // u_r at r1
        if (iv == 4 && f.bot_ctemp) return sr1[it];
    }
    if (ir == nr - 1) {
        if ((iv == 1 || iv == 2) && !f.top_sf) return T(0);
        if (iv == 3) return T(0);                                        // u_r at r2
    }
    if (f.magnetic && iv >= 5) {
        if (thw) {
            if (f.perf_cond) { if (iv == 6 || iv == 7) return T(0); }
            else if (iv == 5) return T(0);                               // A_theta at theta walls
        }
        if (iv == 7 && ((ir == 0 && f.bot_radial) || (ir == nr - 1 && f.top_radial))) return T(0);
    }
    return val;
}

// ---------------------------------------------------------------------------------------------
// 1. Compute_B wall conditions (derivatives.f90:1840-1883) on the wall points of B = curl A:
//    at the theta walls B_theta = 0 (perfectly conducting) or B_phi = B_r = 0 (normal field);
//    at a radial-field r wall B_theta = B_phi = 0.  Thread q < nr*nphi takes the theta-wall
//    pair of pencil (ip, ir); thread q < nth*nphi the r-wall pair of column (it, ip).
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_curl_walls(Consts<T> c, T *__restrict__ B, int perf_cond, int bot_radial,
                                                    int top_radial) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = c.N, plane = (long)c.nth * c.nphi;
    if (q < (long)c.nr * c.nphi) {
        const long row = q * c.nth;   // q = ir * nphi + ip
        for (int side = 0; side < 2; side++) {
            const long i = row + (side ? c.nth - 1 : 0);
            if (perf_cond) {
                B[i] = T(0);
            } else {
                B[N + i] = T(0);
                B[2 * N + i] = T(0);
            }
        }
    }
    if (q < plane) {
        if (bot_radial) {
            B[q] = T(0);
            B[N + q] = T(0);
        }
        if (top_radial) {
            const long i = (long)(c.nr - 1) * plane + q;
            B[i] = T(0);
            B[N + i] = T(0);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// 2. state-dependent wall values.  theta buffers are (nr, nphi) = pencils of the theta axis,
//    r buffers (nphi, nth) = pencils of the r axis.
// ---------------------------------------------------------------------------------------------
struct WallPtrs {
    long off_tv_lo, off_tv_hi, off_tbp_lo, off_tbp_hi, off_trho_lo, off_trho_hi;   // theta (nr*nphi)
    long off_rw_lo, off_rw_hi, off_rv_lo, off_rv_hi, off_rs_lo, off_rs_hi, off_rrho_lo, off_rrho_hi;  // r
    long total;
};

inline WallPtrs wall_layout(int nr, int nth, int nphi) {
    WallPtrs w;
    long nt = (long)nr * nphi, nrr = (long)nth * nphi, o = 0;
    w.off_tv_lo = o; o += nt; w.off_tv_hi = o; o += nt;
    w.off_tbp_lo = o; o += nt; w.off_tbp_hi = o; o += nt;
    w.off_trho_lo = o; o += nt; w.off_trho_hi = o; o += nt;
    w.off_rw_lo = o; o += nrr; w.off_rw_hi = o; o += nrr;
    w.off_rv_lo = o; o += nrr; w.off_rv_hi = o; o += nrr;
    w.off_rs_lo = o; o += nrr; w.off_rs_hi = o; o += nrr;
    w.off_rrho_lo = o; o += nrr; w.off_rrho_hi = o; o += nrr;
    w.total = o;
    return w;
}

template <typename T>
__global__ __launch_bounds__(256) void k_walls_state(Consts<T> c, Prof<T> P, WallPtrs W, T *__restrict__ wall,
                                                     const T *__restrict__ st, const T *__restrict__ B,
                                                     int bot_sf, int top_sf, int bot_flux, int top_flux,
                                                     int theta_normal) {
    const long nt = (long)c.nr * c.nphi, nrr = (long)c.nth * c.nphi;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const T *w = st + c.N, *v = st + 2 * c.N;
    if (i < nt) {   // theta walls (rhs_mpi.f90:410-455): pencil (ip, ir)
        const int ip = (int)(i % c.nphi), ir = (int)(i / c.nphi);
        const long q0 = ((long)ir * c.nphi + ip) * c.nth, q1 = q0 + c.nth - 1;
        wall[W.off_tv_lo + i] = P.cs[0] / P.sn[0] * v[q0];
        wall[W.off_tv_hi + i] = P.cs[c.nth - 1] / P.sn[c.nth - 1] * v[q1];
        if (c.magnetic) {
            const T *bp = B + c.N;
            wall[W.off_tbp_lo + i] = theta_normal ? -P.cs[0] / P.sn[0] * bp[q0] : T(0);
            wall[W.off_tbp_hi + i] = theta_normal ? -P.cs[c.nth - 1] / P.sn[c.nth - 1] * bp[q1] : T(0);
        }
    }
    if (i < nrr) {  // radial walls (rhs_mpi.f90:578-637): pencil (it, ip)
        const int it = (int)(i % c.nth);
        const long q0 = i, q1 = (long)(c.nr - 1) * nrr + i;
        const T r1 = P.r[0], r2 = P.r[c.nr - 1], x1 = P.dxdr[0], x2 = P.dxdr[c.nr - 1];
        wall[W.off_rw_lo + i] = bot_sf ? w[q0] / r1 / x1 : T(0);
        wall[W.off_rw_hi + i] = top_sf ? w[q1] / r2 / x2 : w[q1];
        wall[W.off_rv_lo + i] = bot_sf ? v[q0] / r1 / x1 : T(0);
        wall[W.off_rv_hi + i] = top_sf ? v[q1] / r2 / x2 : v[q1];
        wall[W.off_rs_lo + i] = bot_flux ? (c.slap ? c.s_lap_lo : P.dsdr1[it] / x1) : T(0);
        wall[W.off_rs_hi + i] = top_flux ? (c.slap ? c.s_lap_hi : P.dsdr2[it] / x2) : T(0);
    }
}

// ---------------------------------------------------------------------------------------------
// 4. pointwise terms of Source_Theta_Terms, Phi_Terms, Radial_Terms (reference order)
// ---------------------------------------------------------------------------------------------
// The terms of point i (theta index it, r index ir) in out[0..4] = drho, dw, dv, du, ds.
template <typename T>
__device__ __forceinline__ void main_terms(const Consts<T> &c, const Prof<T> &P, const T *__restrict__ st,
                                           const T *__restrict__ B, const T *__restrict__ D, const ThetaD<T> &td,
                                           long i, int it, int ir, T (&out)[5]) {
    const long N = c.N;
    const T ri = P.rinv[ir], rv = P.r[ir], epsr = P.eps[ir];
    const T omsq = rv * c.omega0 * c.omega0, om2 = c.om2;
    const T sn = P.sn[it], cs = P.cs[it], cot = P.cot[it];
    const T W = st[N + i], V = st[2 * N + i], U = st[3 * N + i];
    const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
    const T Rho = th.rho, t = th.t;
    const T invrho = T(1) / Rho;
    const bool mag = c.magnetic != 0;
    T Bt = T(0), Bp = T(0), Br = T(0);
    if (mag) { Bt = B[i]; Bp = B[N + i]; Br = B[2 * N + i]; }
    T xdw, xdv, xdu, xdrho, xds, wk, drho_out, dw_out = T(0), dv_out = T(0), du_out = T(0);
    // ---- Source_Theta_Terms (rhs_mpi.f90:812-1027)
    xdw = omsq * sn * cs;
    // fluctuation form: gravity enters through the background residual hres, the pressure
    // gradients below through their departures from the background's
    xdu = (c.fluct ? P.hres[ir] : -P.grav[ir]) + omsq * sn * sn;
    xdrho = -Rho * W * cot * ri;
    xdv = -om2 * cs * W;
    xdu = xdu + ri * W * W;
    wk = td[D_TW] * ri;
    xdrho = xdrho - Rho * wk;
    xdw = xdw - W * wk;
    xdw = xdw + V * (ri * cot * V + om2 * cs);
    xdv = xdv - ri * V * W * cot;
    xdu = xdu + V * (ri * V + om2 * sn);
    wk = td[D_TV] * ri;
    xdv = xdv - W * wk;
    drho_out = xdrho - T(2) * ri * Rho * U;
    xdv = xdv - U * (ri * V + om2 * sn);
    xdw = xdw - ri * U * W;
    if (!mag) dv_out = xdv;
    wk = td[D_TU] * ri;
    xdu = xdu - W * wk;
    if (!mag) du_out = xdu;
    xds = (fabs(epsr) > T(0)) ? epsr * invrho / t : T(0);
    wk = td[D_TS] * ri;
    if (mag) xdw = xdw - c.gamm1 * t * wk;
    else dw_out = xdw - c.gamm1 * t * wk;
    xds = xds - W * wk;
    if (mag) {
        const T oorho = invrho * T(kOneOver4Pi) * ri;
        xdu = xdu - Bt * Bt * oorho;
        xdw = xdw - Bp * Bp * cot * oorho;
        xdv = xdv + Bp * Bt * cot * oorho;
        xdu = xdu - Bp * Bp * oorho;
        wk = td[D_TBP] * oorho;
        xdv = xdv + Bt * wk;
        xdw = xdw - Bp * wk;
        wk = td[D_TBR] * oorho;
        dw_out = xdw + Br * (Bt * oorho - wk);
        dv_out = xdv + Br * Bp * oorho;
        du_out = xdu + Bt * wk;
    }
    // ---- Phi_Terms (rhs_mpi.f90:1064-1260)
    {
        const T oors = ri * P.csc[it];
        T oorho = invrho;
        wk = D[D_PRHO * N + i] * oors;
        xdrho = drho_out;
        xdv = dv_out;
        xdrho = xdrho - V * wk;
        xdv = xdv - c.pconst * wk * t * oorho;
        // ds carried in xds
        wk = D[D_PV * N + i] * oors;
        xdv = xdv - V * wk;
        drho_out = xdrho - Rho * wk;
        xdw = dw_out;
        wk = D[D_PW * N + i] * oors;
        xdw = xdw - V * wk;
        xdu = du_out;
        if (!mag) dw_out = xdw;
        wk = D[D_PU * N + i] * oors;
        xdu = xdu - V * wk;
        if (!mag) du_out = xdu;
        wk = D[D_PS * N + i] * oors;
        xds = xds - V * wk;
        xdv = xdv - c.gamm1 * t * wk;
        if (!mag) dv_out = xdv;
        if (mag) {
            oorho = T(kOneOver4Pi) * oorho * oors;
            wk = D[D_PBT * N + i] * oorho;
            xdv = xdv - Bt * wk;
            dw_out = xdw + Bp * wk;
            wk = D[D_PBR * N + i] * oorho;
            dv_out = xdv - Br * wk;
            du_out = xdu + Bp * wk;
        }
    }
    // ---- Radial_Terms (rhs_mpi.f90:1297-1518); r derivatives already carry dx/dr
    {
        wk = D[D_RW * N + i];
        xdw = dw_out - U * wk;
        if (!mag) dw_out = xdw;
        wk = D[D_RV * N + i];
        xdv = dv_out - U * wk;
        if (!mag) dv_out = xdv;
        xdu = du_out;
        wk = D[D_RU * N + i];
        xdu = xdu - U * wk;
        drho_out = drho_out - Rho * wk;
        wk = D[D_RS * N + i];
        if (c.fluct) {
            xdu = xdu - c.gamm1 * (t * wk + P.tbar[ir] * th.b * P.dsbar[ir]);
            wk = wk + P.dsbar[ir];
        } else {
            xdu = xdu - c.gamm1 * t * wk;
        }
        xds = xds - U * wk;
        if (!mag) du_out = xdu;
        if (mag) {
            const T oorho = invrho * T(kOneOver4Pi);
            wk = D[D_RBT * N + i] * oorho;
            xdu = xdu - Bt * wk;
            dw_out = xdw + Br * wk;
            wk = D[D_RBP * N + i] * oorho;
            dv_out = xdv + Br * wk;
            du_out = xdu - Bp * wk;
        }
    }
    out[0] = drho_out;
    out[1] = dw_out;
    out[2] = dv_out;
    out[3] = du_out;
    out[4] = xds;
}

// ---------------------------------------------------------------------------------------------
// 4b. explicit diffusion (Laplacian / SLaplacian branches of Source_Theta_Terms, Phi_Terms,
//     Radial_Terms and Laplacian_Cross_Terms_Radial/_Theta, rhs_mpi.f90:844-1474, :2070-2296).
//     Every branch adds to the tendencies without reading them, so viscous_terms adds the lot
//     to main_terms' sum; the density wall closure sees the viscous momentum through it.
// ---------------------------------------------------------------------------------------------
// Wall values of the cross-term sweeps.  The reference differentiates r*vars2(k) (r sweeps) and
// sin*vars2(k) (theta sweeps), vars2 holding the first derivatives over r or r sin; here the
// first-derivative slots themselves are swept (r * dw/dth / r = dw/dth; the 1/sin and 1/r
// factors are constant along the pencil and applied afterwards), so the wall values are the
// reference's divided by the same factors:
//   r, dw/dth at r1:    r1 * vars2(1)/r1/dxdr(1)  -> D_TW/(r1 dxdr(1))      (Bottom_Stress_Free)
//   r, dv/dph at r2:    r2 * vars2(4)/r2/dxdr(nr) -> D_PV/(r2 dxdr(nr))     (Bottom_Stress_Free,
//                       as written at rhs_mpi.f90:2120 for the top wall)
//   theta, dv/dph:      sin * vars2(4) cot        -> cot D_PV on both theta walls (:2243-2250)
// every other wall value is zero (the reference's dbinfo walls are never written).
// Layouts: r buffers (nphi, nth), theta buffers (nr, nphi).
template <typename T>
__global__ __launch_bounds__(256) void k_lap_walls(Consts<T> c, Prof<T> P, const T *__restrict__ D,
                                                   T *__restrict__ lw, int bot_sf) {
    const long nt = (long)c.nr * c.nphi, nrr = (long)c.nth * c.nphi, N = c.N;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrr) {
        const long q0 = i, q1 = (long)(c.nr - 1) * nrr + i;
        lw[i] = bot_sf ? D[D_TW * N + q0] * P.rinv[0] / P.dxdr[0] : T(0);
        lw[nrr + i] = bot_sf ? D[D_PV * N + q1] * P.rinv[c.nr - 1] / P.dxdr[c.nr - 1] : T(0);
    }
    if (i < nt) {
        const int ip = (int)(i % c.nphi), ir = (int)(i / c.nphi);
        const long q0 = ((long)ir * c.nphi + ip) * c.nth, q1 = q0 + c.nth - 1;
        lw[2 * nrr + i] = P.cot[0] * D[D_PV * N + q0];
        lw[2 * nrr + nt + i] = P.cot[c.nth - 1] * D[D_PV * N + q1];
    }
}

// adds the explicit-diffusion terms of point i to out[1..4] (dw, dv, du, ds)
template <typename T>
__device__ __forceinline__ void viscous_terms(const Consts<T> &c, const Prof<T> &P, const T *__restrict__ st,
                                              const T *__restrict__ D, const ThetaD<T> &td, const T *__restrict__ L,
                                              long i, int it, int ir, T (&out)[5]) {
    const long N = c.N;
    const T ri = P.rinv[ir], csc = P.csc[it], cot = P.cot[it];
    const T W = st[N + i], V = st[2 * N + i], U = st[3 * N + i];
    const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
    const T invrho = T(1) / th.rho, t = th.t, invt = T(1) / th.t;
    const T x = P.dxdr[ir], w9 = P.w9[ir], oors = ri * csc;
    T ddw = T(0), ddv = T(0), ddu = T(0), dds = T(0);
    if (c.lap) {
        const T nu = P.mur[ir] * invrho, dlm = P.dlnmu[ir], w5 = T(2) * ri + dlm, csc2 = csc * csc;
        const T tw = td[D_TW] * ri, tv = td[D_TV] * ri, tu = td[D_TU] * ri;
        const T pv = D[D_PV * N + i] * oors, pw = D[D_PW * N + i] * oors, pu = D[D_PU * N + i] * oors;
        const T rw = D[D_RW * N + i], rv = D[D_RV * N + i], ru = D[D_RU * N + i];
        const T c43 = T(4) / T(3), c13 = T(1) / T(3), c73 = T(7) / T(3), c23 = T(2) / T(3);
        // Source_Theta_Terms (rhs_mpi.f90:844-926)
        ddw += (-c43 * ri * ri * csc2 - ri * dlm) * nu * W;
        ddu += -c13 * (T(7) * ri + T(2) * dlm) * (ri * nu * cot * W + tw * nu);
        ddw += c43 * ri * cot * tw * nu;
        ddw += c43 * ri * ri * nu * L[L_TW2 * N + i];
        ddv += ri * nu * (cot * tv - (ri * csc2 + dlm) * V);
        ddv += ri * ri * nu * L[L_TV2 * N + i];
        ddw += (T(2) * c43 * ri + dlm) * tu * nu;
        ddu += ri * cot * tu * nu;
        ddu += ri * ri * nu * L[L_TU2 * N + i];
        // Phi_Terms (:1123-1198); heating theta-theta / phi-phi and theta-phi
        dds += c43 * nu * (tw + U * ri) * (tw + T(2) * U * ri - pv - ri * cot * W) * invt;
        ddw += -c73 * ri * cot * pv * nu;
        ddu += (-c73 * ri - c23 * dlm) * pv * nu;
        ddv += c43 * L[L_PV2 * N + i] * oors * oors * nu;
        T g = tv + pw - ri * cot * V;
        dds += nu * g * g * invt;
        ddv += c73 * ri * cot * pw * nu;
        ddw += L[L_PW2 * N + i] * oors * oors * nu;
        ddv += (T(2) * c43 * ri + dlm) * pu * nu;
        ddu += L[L_PU2 * N + i] * oors * oors * nu;
        // Radial_Terms (:1317-1431); the cot/r dw/dr term has no mu/rho factor (:1322) and the
        // r-phi heating term is added a second time with du/dr in place of dv/dr (:1418)
        ddu += c13 * cot * ri * rw;
        ddw += nu * w5 * rw + w9 * rw * nu + L[L_RW2 * N + i] * x * x * nu;
        g = rw + tu - W * ri;
        dds += nu * g * g * invt;
        g = rv + pu - V * ri;
        dds += nu * g * g * invt;
        ddv += rv * w5 * nu + w9 * rv * nu + L[L_RV2 * N + i] * x * x * nu;
        ddu += c43 * nu * w5 * (ru - U * ri) + c43 * w9 * ru * nu;
        dds += c43 * nu * ru * (ru - tw - U * ri) * invt;
        g = ru + pu - V * ri;
        dds += nu * g * g * invt;
        const T pp = pv + (cot * W + U) * ri;
        dds += c43 * nu * pp * (pp - ru) * invt;
        ddu += c43 * L[L_RU2 * N + i] * x * x * nu;
        // Laplacian_Cross_Terms_Radial / _Theta (:2126-2295): mur/(3 r rho) times d/dr of
        // r*vars2 (swept slots carry dx/dr), and mur/(3 r rho sin) times d/dth of sin*vars2
        const T c6 = c13 * nu * ri;
        ddu += c6 * L[L_XR1 * N + i];
        ddv += c6 * L[L_XR2 * N + i] * csc;
        ddw += c6 * L[L_XR3 * N + i];
        ddu += c6 * L[L_XR4 * N + i] * csc;
        ddw += c6 * csc * ri * L[L_XT1 * N + i];
        ddv += c6 * csc * ri * L[L_XT2 * N + i];
    }
    if (c.slap) {
        // thermal diffusion (:959-975, :1221-1228, :1456-1472); the theta second-derivative term
        // carries 1/r^3 as in the reference (:970-973)
        const T ks = P.ks[ir], ort = invrho * invt;
        const T ts = td[D_TS] * ri;
        T rs = D[D_RS * N + i];
        T d2r = L[L_RS2 * N + i];
        if (c.fluct) {
            rs += P.dsbar[ir];
            d2r += P.d2sbar[ir];
        }
        dds += cot * ri * ks * ts * ort;
        dds += ks * ri * ri * ri * L[L_TS2 * N + i] * ort;
        dds += ks * L[L_PS2 * N + i] * oors * oors * ort;
        dds += rs * ort * (T(2) * ri + P.dlnks[ir]) * ks + w9 * rs * ort * ks + ks * d2r * x * x * ort;
    }
    out[1] += ddw;
    out[2] += ddv;
    out[3] += ddu;
    out[4] += dds;
}

// every RHS term before the density terms at point i: main_terms plus, with explicit
// diffusion, viscous_terms (rhs_mpi.f90:786-1519, :2070-2296)
template <typename T>
__device__ __forceinline__ void pre_density_terms(const Consts<T> &c, const Prof<T> &P, const T *__restrict__ st,
                                                  const T *__restrict__ B, const T *__restrict__ D,
                                                  const ThetaD<T> &td, const T *__restrict__ L, long i, int it,
                                                  int ir, T (&out)[5]) {
    main_terms(c, P, st, B, D, td, i, it, ir, out);
    if (c.slap) viscous_terms(c, P, st, D, td, L, i, it, ir, out);
}

// ---------------------------------------------------------------------------------------------
// 5a. the theta-wall block's fields that no theta sweep reads or writes: rho, B_theta and the
//     phi / r derivative slots of the momentum tendency's theta component (D_PW, D_PBT, D_RW,
//     D_RBT), copied from the theta-wall points (one thread per (ir, ip), both walls)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_walls_gather(Consts<T> c, const T *__restrict__ st, const T *__restrict__ B,
                                                      const T *__restrict__ D, T *__restrict__ tw) {
    const long nt = (long)c.nr * c.nphi, N = c.N, M = 2 * nt;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const long q0 = i * c.nth;   // (ir, ip) row of theta points, theta index 0
    for (int side = 0; side < 2; side++) {
        const long q = q0 + (side ? c.nth - 1 : 0), w = side * nt + i;
        tw[w] = st[q];
        tw[(8 + D_PW) * M + w] = D[D_PW * N + q];
        tw[(8 + D_RW) * M + w] = D[D_RW * N + q];
        if (c.magnetic) {
            tw[5 * M + w] = B[q];
            tw[(8 + D_PBT) * M + w] = D[D_PBT * N + q];
            tw[(8 + D_RBT) * M + w] = D[D_RBT * N + q];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// 5. density wall derivatives from the momentum tendency (rhs_mpi.f90:1553-1564, :1750-1773)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_walls_density(Consts<T> c, Prof<T> P, WallPtrs W, T *__restrict__ wall,
                                                       const T *__restrict__ st, const T *__restrict__ B,
                                                       const T *__restrict__ D, const T *__restrict__ L,
                                                       int lbcr_rho, const T *__restrict__ tw) {
    const long nt = (long)c.nr * c.nphi, nrr = (long)c.nth * c.nphi, N = c.N;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // the momentum tendency before the density terms, evaluated at the wall points themselves
    T o0[5], o1[5];
    if (tw && i < nt) {
        // the theta-wall block (k_walls_gather and the theta sweeps' end-row taps): the same
        // values, one field after another with stride 2*nt
        const int ir = (int)(i / c.nphi);
        const T rv = P.r[ir] * c.invpconst;
        Consts<T> cw = c;
        cw.N = 2 * nt;
        const T *stw = tw, *Bw = tw + 5 * cw.N, *Dw = tw + 8 * cw.N;
        const long q0 = i, q1 = nt + i;
        const Thermo<T> a0 = thermo(stw[q0], stw[4 * cw.N + q0], ir, cw, P);
        const Thermo<T> a1 = thermo(stw[q1], stw[4 * cw.N + q1], ir, cw, P);
        pre_density_terms(cw, P, stw, Bw, Dw, theta_from(Dw, cw.N, q0), L, q0, 0, ir, o0);
        pre_density_terms(cw, P, stw, Bw, Dw, theta_from(Dw, cw.N, q1), L, q1, c.nth - 1, ir, o1);
        wall[W.off_trho_lo + i] = rv * o0[1] * a0.rho / a0.t;
        wall[W.off_trho_hi + i] = rv * o1[1] * a1.rho / a1.t;
    } else if (i < nt) {
        const int ir = (int)(i / c.nphi), ip = (int)(i % c.nphi);
        const T rv = P.r[ir] * c.invpconst;
        const long q0 = ((long)ir * c.nphi + ip) * c.nth, q1 = q0 + c.nth - 1;
        const Thermo<T> a0 = thermo(st[q0], st[4 * N + q0], ir, c, P), a1 = thermo(st[q1], st[4 * N + q1], ir, c, P);
        const T rho0 = a0.rho, rho1 = a1.rho, t0 = a0.t, t1 = a1.t;
        pre_density_terms(c, P, st, B, D, theta_from(D, N, q0), L, q0, 0, ir, o0);
        pre_density_terms(c, P, st, B, D, theta_from(D, N, q1), L, q1, c.nth - 1, ir, o1);
        wall[W.off_trho_lo + i] = rv * o0[1] * rho0 / t0;
        wall[W.off_trho_hi + i] = rv * o1[1] * rho1 / t1;
    }
    if (i < nrr) {
        const long q0 = i, q1 = (long)(c.nr - 1) * nrr + i;
        const int it = (int)(i % c.nth);
        const Thermo<T> a0 = thermo(st[q0], st[4 * N + q0], 0, c, P);
        const Thermo<T> a1 = thermo(st[q1], st[4 * N + q1], c.nr - 1, c, P);
        const T rho0 = a0.rho, rho1 = a1.rho, t0 = a0.t, t1 = a1.t;
        const bool nlo = lbcr_rho == 2 || lbcr_rho == 4, nhi = lbcr_rho == 3 || lbcr_rho == 4;
        T du0 = T(0), du1 = T(0);
        if (nlo) {
            pre_density_terms(c, P, st, B, D, theta_from(D, N, q0), L, q0, it, 0, o0);
            du0 = o0[3];
        }
        if (nhi) {
            pre_density_terms(c, P, st, B, D, theta_from(D, N, q1), L, q1, it, c.nr - 1, o1);
            du1 = o1[3];
        }
        T lo = nlo ? du0 * rho0 * c.invpconst / t0 : rho0;
        T hi = nhi ? du1 * rho1 * c.invpconst / t1 : rho1;
        if (c.fluct) {
            // the sweep differentiates rho - rho_bar, whose background derivative drbar already
            // carries wbar at these walls: prescribe the departure
            //   (du/dt without density terms + pconst (T_bar/rho_bar) wbar) rho/(pconst T) - wbar
            //   = du_partial (rho/T)/pconst - wbar a/(1+a)      (du_partial: hres-based, small)
            if (nlo) lo = du0 * (rho0 / t0) * c.invpconst - c.wbar_lo * a0.a / (T(1) + a0.a);
            if (nhi) hi = du1 * (rho1 / t1) * c.invpconst - c.wbar_hi * a1.a / (T(1) + a1.a);
        }
        wall[W.off_rrho_lo + i] = lo / P.dxdr[0];
        wall[W.off_rrho_hi + i] = hi / P.dxdr[c.nr - 1];
    }
}

// ---------------------------------------------------------------------------------------------
// 6. density terms, top relaxation, induction, xi_cs (rhs_mpi.f90:1580-1596, :1781-1818,
//    :2342-2346, :769)
// ---------------------------------------------------------------------------------------------
// RK4 stage update fused into the last RHS kernel (step_rk4.f90:392, 418-419, 444-445, 470-471):
//   stage 1: acc = Dif + k/6   y1 = y0 + (dt/2) k    (SLDiffusion; without it acc = k, :387)
//   stage 2: acc += k/3        y1 = y0 + (dt/2) k
//   stage 3: acc += k/3        y1 = y0 + dt k
//   stage 4: acc += k/6        y0 = y0 + dt acc
// The tendency k of every field at a point is consumed where it is produced instead of being
// stored and re-read.
// stage 0 = plain RHS (write dv).  FP64 path: acc64 (FP64 accumulator), y1 = stage state.
// Mixed path (T = float): acc32 (FP32 accumulator), y0f = FP32 copy of y0 (fluctuation form),
// y1 = FP32 stage state; stage 4 forms y0 + dt acc in FP64.  Bound is applied to every new
// value.  dif = the SLD tendency (stage 1 only) or null.
template <typename T>
struct StageFuse {
    int stage;
    double dt;
    BoundFlags f;
    double *y0;
    double *acc64;
    float *acc32;
    T *y0f;
    const T *dif;
    // stage 1 with SLD Ohmic heating: the curl of the SLD A tendency (3 fields), whose heating
    // term (diffusion.f90:430-461) the stage subtracts from the entropy's SLD tendency here
    // instead of a separate pass over ds
    const T *ohm;
    // stage 1 with SLD: the viscous-heating / mass-diffusion corrections of the SLD tendency
    // (diffusion.f90:375-428) are applied here, at the point, instead of in the SLD march
    int corr, visc, mass;
    T gcv;
    T *y1;
    const double *sr1_64, *rbar64, *sbar64;
    const T *sr1_t;   // constant-temperature wall entropy in stage precision/form
};

template <typename T>
__device__ __forceinline__ void stage_value(const StageFuse<T> &F, const Consts<T> &c, int iv, long i, int it,
                                            int ip, int ir, T kk_t, bool own_dif = false, T difv = T(0)) {
    const long j = (long)iv * c.N + i;
    const int st = F.stage;
    if constexpr (sizeof(T) == sizeof(double)) {
        const double kk = (double)kk_t;
        double s;
        // stage 1: vars_save = D + k1/6 with SLDiffusion, vars_save = k1 without it
        // (step_rk4.f90:369-389: k1 then carries weight 1, as in the reference)
        if (st == 1) s = F.dif ? (double)(own_dif ? difv : F.dif[j]) + kk / 6.0 : kk;
        else s = F.acc64[j] + kk / (st == 4 ? 6.0 : 3.0);
        const double y = st == 4 ? F.y0[j] + F.dt * s : F.y0[j] + (st == 3 ? F.dt : F.dt / 2.0) * kk;
        const double yb = bound_value<double>(iv, it, ip, ir, y, F.f, c.nth, c.nr, F.sr1_64);
        if (st < 4) {
            F.acc64[j] = s;
            F.y1[j] = (T)yb;
        } else {
            F.y0[j] = yb;
        }
    } else {
        const float kk = (float)kk_t;
        if (st < 4) {
            const float a = st == 1 ? (F.dif ? (float)(own_dif ? difv : F.dif[j]) + kk / 6.0f : kk) : F.acc32[j] + kk / 3.0f;
            const float h = (float)(st == 3 ? F.dt : 0.5 * F.dt);
            const float y = (float)F.y0f[j] + h * kk;
            F.acc32[j] = a;
            F.y1[j] = (T)bound_value<float>(iv, it, ip, ir, y, F.f, c.nth, c.nr, (const float *)F.sr1_t);
        } else {
            const float a = F.acc32[j] + kk / 6.0f;
            const double y = F.y0[j] + F.dt * (double)a;
            const double yb = bound_value<double>(iv, it, ip, ir, y, F.f, c.nth, c.nr, F.sr1_64);
            const double bg = iv == 0 ? F.rbar64[ir] : iv == 4 ? F.sbar64[ir] : 0.0;
            F.y0[j] = yb;
            F.y0f[j] = (T)(yb - bg);
        }
    }
}

// Every term of point i from the theta derivatives td, the density derivatives *dthrho (theta)
// and *drrho (r) and the phi / r derivative slots of D, then the fused stage (or dv).
template <typename T>
__device__ __forceinline__ void final_point(const Consts<T> &c, const Prof<T> &P, const T *st,
                                            const T *__restrict__ B, const ThetaD<T> &td, const T *dthrho,
                                            const T *drrho,
                                            T *dv, const StageFuse<T> &F, const float *__restrict__ ext,
                                            const T *__restrict__ D, const T *__restrict__ L, long i, int it,
                                            int ip, int ir) {
    const long N = c.N;
    const T W = st[N + i], V = st[2 * N + i], U = st[3 * N + i];
    const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
    const T Rho = th.rho, t = th.t;
    const T invrho = T(1) / Rho;
    T o[5];
    pre_density_terms(c, P, st, B, D, td, L, i, it, ir, o);
    T drho = o[0], dw = o[1], du = o[3], ds = o[4];
    // Density_Terms_Theta
    T wk = P.rinv[ir] * *dthrho;
    drho = drho - W * wk;
    wk = wk * invrho;
    dw = dw - c.pconst * wk * t;
    // Density_Terms_Radial
    wk = *drrho;
    if (c.fluct) {
        // pconst (T/rho) (drho'/dr + drbar) minus the background part already in hres
        du = du - c.pconst * ((t * invrho) * wk + P.tr[ir] * th.a * P.drbar[ir]);
        wk = wk + P.drbar[ir];
        drho = drho - U * wk;
    } else {
        drho = drho - U * wk;
        wk = wk * invrho;
        du = du - c.pconst * t * wk;
    }
    if (c.top_relax && ir == c.nr - 1) {
        const T dt_ = c.damp * (c.gamma - T(2));
        const T dT = c.fluct ? c.dT0top - P.tbar[ir] * th.b : c.T0 - t;   // T0 - T
        drho = drho + Rho * dT / (dt_ * t);
        ds = ds - c.Cv * dT / (dt_ * t);
    }
    if (c.xi_on) drho = c.xi2 * drho;
    T da0 = T(0), da1 = T(0), da2 = T(0);
    if (c.magnetic) {
        const T Bt = B[i], Bp = B[N + i], Br = B[2 * N + i];
        da0 = V * Br - U * Bp;
        da1 = U * Bt - W * Br;
        da2 = W * Bp - V * Bt;
        if (c.lap) {
            // Induction_Terms with Laplacian (rhs_mpi.f90:2303-2322, VectorCurl derivatives.f90
            // :1894-2104): dA/dt -= etar curl B, ds/dt += (4 pi/c^2) etar |curl B|^2 / (rho T)
            const T ri = P.rinv[ir], csc = P.csc[it], eta = P.etar[ir];
            const T jt = ri * L[L_JP1 * N + i] * csc - L[L_JRP * N + i] * ri;
            const T jp = -L[L_JT2 * N + i] * ri + L[L_JRT * N + i] * ri;
            const T jr = L[L_JT1 * N + i] * csc * ri - ri * L[L_JP2 * N + i] * csc;
            da0 -= eta * jt;
            da1 -= eta * jp;
            da2 -= eta * jr;
            ds += (T(kJconst) * eta) * (jt * jt + jp * jp + jr * jr) * invrho / t;
        }
    }
    T dvv = o[2];
    if (ext) {   // learned closure tendency
        drho += (T)ext[i];
        dw += (T)ext[N + i];
        dvv += (T)ext[2 * N + i];
        du += (T)ext[3 * N + i];
        ds += (T)ext[4 * N + i];
        if (c.magnetic) {
            da0 += (T)ext[5 * N + i];
            da1 += (T)ext[6 * N + i];
            da2 += (T)ext[7 * N + i];
        }
    }
    if (F.stage == 0) {
        dv[i] = drho;
        dv[2 * N + i] = dvv;
        dv[N + i] = dw;
        dv[3 * N + i] = du;
        dv[4 * N + i] = ds;
        if (c.magnetic) {
            dv[5 * N + i] = da0;
            dv[6 * N + i] = da1;
            dv[7 * N + i] = da2;
        }
    } else {
        // st may alias F.y1 (the stage state is updated in place): every read of st at this
        // point happened above
        const bool sld1 = F.stage == 1 && F.dif;
        if (sld1 && F.corr) {
            // the SLD march left rho..s uncorrected: viscous heating, then mass diffusion, the
            // operations of k_sld_col's point corrections in their order (the same state values
            // and thermodynamics at the point)
            // (d4 takes the three viscous terms in field order before the mass term; each velocity
            // tendency enters d4 before its own mass correction; one velocity field at a time)
            const T d0 = F.dif[i];
            const T itemp = T(1) / t;
            T d4 = F.dif[4 * N + i];
            stage_value(F, c, 0, i, it, ip, ir, drho, true, d0);
            {
                T d1 = F.dif[N + i];
                if (F.visc) d4 = d4 - W * d1 * itemp;
                if (F.mass) d1 = d1 - invrho * W * d0;
                stage_value(F, c, 1, i, it, ip, ir, dw, true, d1);
            }
            {
                T d2 = F.dif[2 * N + i];
                if (F.visc) d4 = d4 - V * d2 * itemp;
                if (F.mass) d2 = d2 - invrho * V * d0;
                stage_value(F, c, 2, i, it, ip, ir, dvv, true, d2);
            }
            {
                T d3 = F.dif[3 * N + i];
                if (F.visc) d4 = d4 - U * d3 * itemp;
                if (F.mass) d3 = d3 - invrho * U * d0;
                stage_value(F, c, 3, i, it, ip, ir, du, true, d3);
            }
            if (F.mass) d4 = d4 - F.gcv * invrho * d0;
            if (F.ohm) {
                const T tt = T(0.5 * kOneOver4Pi) * invrho * (T(1) / t);
                d4 = d4 - B[i] * F.ohm[i] * tt;
                d4 = d4 - B[N + i] * F.ohm[N + i] * tt;
                d4 = d4 - B[2 * N + i] * F.ohm[2 * N + i] * tt;
            }
            stage_value(F, c, 4, i, it, ip, ir, ds, true, d4);
        } else {
            stage_value(F, c, 0, i, it, ip, ir, drho);
            stage_value(F, c, 1, i, it, ip, ir, dw);
            stage_value(F, c, 2, i, it, ip, ir, dvv);
            stage_value(F, c, 3, i, it, ip, ir, du);
            if (F.ohm && sld1) {
                // SLD Ohmic heating: ds_sld -= B . curl(dA_sld/dt) / (8 pi rho T), the operations
                // of k_sld_finish in its order
                const T tt = T(0.5 * kOneOver4Pi) * invrho * (T(1) / t);
                T difs = F.dif[4 * N + i];
                difs = difs - B[i] * F.ohm[i] * tt;
                difs = difs - B[N + i] * F.ohm[N + i] * tt;
                difs = difs - B[2 * N + i] * F.ohm[2 * N + i] * tt;
                stage_value(F, c, 4, i, it, ip, ir, ds, true, difs);
            } else {
                stage_value(F, c, 4, i, it, ip, ir, ds);
            }
        }
        if (c.magnetic) {
            stage_value(F, c, 5, i, it, ip, ir, da0);
            stage_value(F, c, 6, i, it, ip, ir, da1);
            stage_value(F, c, 7, i, it, ip, ir, da2);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 3 : 7) void k_rhs_final(Consts<T> c, Prof<T> P, const T *st,
                                                   const T *__restrict__ B, const T *__restrict__ dth_rho,
                                                   const T *__restrict__ dr_rho, T *dv, StageFuse<T> F,
                                                   const float *__restrict__ ext, const T *__restrict__ D,
                                                   const T *__restrict__ L) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    int it, ip, ir;
    coords(i, c, it, ip, ir);
    final_point(c, P, st, B, theta_from(D, c.N, i), dth_rho + i, dr_rho + i, dv, F, ext, D, L, i, it, ip, ir);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bound(Consts<T> c, BoundFlags f, const T *__restrict__ sr1, T *st, int nv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N * nv) return;
    const int iv = (int)(i / c.N);
    const long j = i - (long)iv * c.N;
    int it, ip, ir;
    coords(j, c, it, ip, ir);
    st[i] = bound_value(iv, it, ip, ir, st[i], f, c.nth, c.nr, sr1);
}

// ---------------------------------------------------------------------------------------------
// y1 = (TS)(y0 - background) (first step of an FP32 run; rho and s as fluctuations)
template <typename TS>
__global__ __launch_bounds__(256) void k_cast_state(Consts<TS> c, int nv, const double *__restrict__ rbar64,
                                                    const double *__restrict__ sbar64, const double *__restrict__ y0,
                                                    TS *__restrict__ y1) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N * nv) return;
    const int iv = (int)(i / c.N);
    const int ir = (int)((i - (long)iv * c.N) / ((long)c.nth * c.nphi));
    const double bg = iv == 0 ? rbar64[ir] : iv == 4 ? sbar64[ir] : 0.0;
    y1[i] = (TS)(y0[i] - bg);
}


// Stage of Step_RK3_TVD / Step_RK2_TVD (step_rk4.f90:166-335, :31-162) with Bound:
//   mode 0 (stage 1)  y = y0 + dtc k
//   mode 1            y = a0 y0 + a1 ys + dtc k      (ys = the previous stage state, FP64)
//   last stage        y = y + dt dif                  (SLDiffusion's Euler diffusion)
// Intermediate stages keep y in ys (FP64) and its stage-precision copy in y1; the last writes y0.
template <typename TS>
__global__ __launch_bounds__(256) void k_tvd_stage(Consts<TS> c, BoundFlags f, const double *__restrict__ sr1_64,
                                                   const double *__restrict__ rbar64,
                                                   const double *__restrict__ sbar64, int mode, double a0,
                                                   double a1, double dtc, double dt, int last,
                                                   double *__restrict__ y0, double *__restrict__ ys,
                                                   const TS *__restrict__ k, const TS *__restrict__ dif,
                                                   TS *__restrict__ y1, int nv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N * nv) return;
    const int iv = (int)(i / c.N);
    const long j = i - (long)iv * c.N;
    int it, ip, ir;
    coords(j, c, it, ip, ir);
    const double kk = (double)k[i];
    double y = (mode == 0) ? y0[i] + dtc * kk : a0 * y0[i] + a1 * ys[i] + dtc * kk;
    if (last && dif) y = y + dt * (double)dif[i];
    const double yb = bound_value<double>(iv, it, ip, ir, y, f, c.nth, c.nr, sr1_64);
    const double bg = (sizeof(TS) == sizeof(double)) ? 0.0 : (iv == 0 ? rbar64[ir] : iv == 4 ? sbar64[ir] : 0.0);
    if (!last) {
        ys[i] = yb;
        y1[i] = (TS)(yb - bg);
    } else {
        y0[i] = yb;
        if (sizeof(TS) != sizeof(double)) y1[i] = (TS)(yb - bg);
    }
}

// ---------------------------------------------------------------------------------------------
// Compute_Step_Size (rhs_mpi.f90:145-185, stratified branch): per-block minima of the sound/flow
// crossing time and of the Alfven-included crossing time.
// ---------------------------------------------------------------------------------------------
// Compute_Step_Size (rhs_mpi.f90:92-275) per point: the sound/flow and Alfven crossing times
// (block minima) and the conservation sums E, L, m, m r sin(theta) (block sums) of
// rhs_mpi.f90:156-183: with dV = r^2 sin(th) dr dth dphi,
//   m += rho dV,  mrsint += r sin^3(th) rho r^2 dr dth dphi,
//   L += r rho sqrt(w^2 + sin^2(th) (v + r sin(th) Omega0)^2) dV,
//   E += rho (Cv T + |u|^2/2) dV + |B|^2/(8 pi) dV.
// The grid strides over the points (at most kCflMaxBlocks blocks); red holds 6 values per
// block: min, min, E, L, m, mrsint.
constexpr int kCflMaxBlocks = 4096;

__device__ __forceinline__ void cfl_block_reduce(double v[6], double *red) {
    __shared__ double sh[6][256];
    for (int q = 0; q < 6; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = fmin(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            for (int q = 2; q < 6; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; q++) red[q] = sh[q][0];
}

__global__ __launch_bounds__(256) void k_cfl(Consts<double> c, Prof<double> P, double dth, double dphi, double Cp,
                                             const double *__restrict__ st, const double *__restrict__ B,
                                             double *__restrict__ red) {
    double v[6] = {1e4, 1e4, 0.0, 0.0, 0.0, 0.0};
    const long N = c.N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        int it, ip, ir;
        coords(i, c, it, ip, ir);
        const double radius = P.r[ir], sinth = P.sn[it], drad = P.dr[ir];
        const double dxmin = fmin(drad, radius * fmin(dth, sinth * dphi));
        const double rho = st[i], w = st[N + i], vv = st[2 * N + i], u = st[3 * N + i], s = st[4 * N + i];
        const double t = temperature(rho, s, c);
        const double dvol = radius * radius * sinth * drad * dth * dphi;
        const double ms = rho * dvol;
        v[4] += ms;
        v[5] += radius * sinth * sinth * ms;
        const double vp = vv + radius * sinth * c.omega0;
        v[3] += radius * rho * sqrt(w * w + sinth * sinth * vp * vp) * dvol;
        double e = rho * (c.Cv * t + 0.5 * (u * u + w * w + vv * vv)) * dvol;
        const double ut = u * u + vv * vv + w * w + (c.gamm1 * Cp * c.xi2) * t;
        v[0] = fmin(v[0], dxmin / sqrt(ut));
        if (c.magnetic) {
            const double b2 = B[i] * B[i] + B[N + i] * B[N + i] + B[2 * N + i] * B[2 * N + i];
            v[1] = fmin(v[1], dxmin / sqrt(ut + b2 * kOneOver4Pi / rho));
            e += 0.5 * b2 * dvol * kOneOver4Pi;
        }
        v[2] += e;
    }
    cfl_block_reduce(v, red + 6 * blockIdx.x);
}

// one block: the nb block partials of k_cfl -> 6 values at out
__global__ __launch_bounds__(256) void k_cfl_final(const double *__restrict__ red, int nb, double *__restrict__ out) {
    double v[6] = {1e4, 1e4, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        v[0] = fmin(v[0], red[6 * b]);
        v[1] = fmin(v[1], red[6 * b + 1]);
        for (int q = 2; q < 6; q++) v[q] += red[6 * b + q];
    }
    cfl_block_reduce(v, out);
}

}  // namespace cssx

#include "cssx_sld.hpp"

// =============================================================================================
// Host orchestration
// =============================================================================================
namespace cssx {

template <typename T>
static Consts<T> make_consts(const cssx_mhd *m) {
    Consts<T> c;
    c.nth = m->nth; c.nphi = m->nphi; c.nr = m->nr; c.N = m->N;
    c.gamm1 = (T)m->gamm1; c.invtconst = (T)m->invtconst; c.invCv = (T)m->invCv; c.pconst = (T)m->pconst;
    c.invpconst = (T)m->invpconst; c.Cv = (T)m->Cv; c.omega0 = (T)m->p.omega0; c.om2 = (T)(2.0 * m->p.omega0);
    c.xi2 = (T)(m->p.xi_cs * m->p.xi_cs); c.gamma = (T)m->p.gamma; c.T0 = (T)m->p.T0; c.damp = (T)m->p.damping_time;
    c.magnetic = m->p.magnetic; c.top_relax = m->p.top_relax; c.xi_on = m->p.xi_cs < 1.0;
    c.fluct = sizeof(T) == 4;
    c.wbar_lo = (T)m->wbar_lo;
    c.wbar_hi = (T)m->wbar_hi;
    c.dT0top = m->h_tbar.empty() ? T(0) : (T)(m->p.T0 - m->h_tbar[m->nr - 1]);
    c.lap = m->p.laplacian != 0;
    c.slap = m->p.laplacian || m->p.slaplacian;
    c.s_lap_lo = c.s_lap_hi = T(0);
    if (c.slap) {
        const int nr = m->nr;
        const double r1 = m->h_r[0], r2 = m->h_r[nr - 1];
        c.s_lap_lo = (T)(-m->p.lumr1 * kOneOver4Pi / (r1 * r1 * m->h_ks[0] * m->h_dxdr[0]));
        c.s_lap_hi = (T)(-m->p.lumr2 * kOneOver4Pi / (r2 * r2 * m->h_ks[nr - 1] * m->h_dxdr[nr - 1]));
    }
    return c;
}

static BoundFlags make_bound_flags(const cssx_mhd *m) {
    BoundFlags f;
    f.theta_stress_free = m->p.theta_stress_free; f.bot_sf = m->p.bottom_stress_free; f.top_sf = m->p.top_stress_free;
    f.bot_ctemp = m->p.bottom_const_temp; f.magnetic = m->p.magnetic; f.perf_cond = m->p.theta_perf_cond;
    f.bot_radial = m->p.bottom_radial_field; f.top_radial = m->p.top_radial_field;
    return f;
}

template <typename T>
static int upload_profiles(cssx_mhd *m, Prof<T> &P) {
    const int nr = m->nr, nth = m->nth;
    const std::vector<double> &drbar = m->h_drbar, &dsbar = m->h_dsbar, &hres = m->h_hres;
    // the profiles back to back in one host vector that grows with every table put into it;
    // each table starts on a 16-element boundary, so that a kernel can read a lane's run of a
    // theta profile as 16-byte vectors
    std::vector<T> h;
    h.reserve(32 * (size_t)nr + 12 * (size_t)nth + 1024);
    size_t o = 0;
    auto put = [&](size_t n, auto gen) {
        o = (o + 15) & ~(size_t)15;
        const size_t at = o;
        h.resize(o + n);
        for (size_t i = 0; i < n; i++) h[o + i] = (T)gen(i);
        o += n;
        return at;
    };
    const bool f32 = sizeof(T) == 4;
    const size_t o_r = put(nr, [&](size_t i) { return m->h_r[i]; });
    const size_t o_ri = put(nr, [&](size_t i) { return 1.0 / m->h_r[i]; });
    const size_t o_x = put(nr, [&](size_t i) { return m->h_dxdr[i]; });
    const size_t o_drr = put(nr, [&](size_t i) { return m->h_dxdr[i] / m->h_r[i]; });
    const size_t o_g = put(nr, [&](size_t i) { return m->h_grav[i]; });
    const size_t o_e = put(nr, [&](size_t i) { return m->h_eps[i]; });
    const size_t o_dr = put(nr, [&](size_t i) { return m->h_dr[i]; });
    const size_t o_rb = put(nr, [&](size_t i) { return m->h_rho_bar[i]; });
    const size_t o_sb = put(nr, [&](size_t i) { return m->h_s_bar[i]; });
    const size_t o_tb = put(nr, [&](size_t i) { return m->h_tbar[i]; });
    const size_t o_tr = put(nr, [&](size_t i) { return m->h_tbar[i] / m->h_rho_bar[i]; });
    const size_t o_hr = put(nr, [&](size_t i) { return hres.empty() ? 0.0 : hres[i]; });
    const size_t o_drb = put(nr, [&](size_t i) { return drbar.empty() ? 0.0 : drbar[i]; });
    const size_t o_dsb = put(nr, [&](size_t i) { return dsbar.empty() ? 0.0 : dsbar[i]; });
    const size_t o_sn = put(nth, [&](size_t i) { return std::sin(m->h_theta[i]); });
    const size_t o_csc = put(nth, [&](size_t i) { return 1.0 / std::sin(m->h_theta[i]); });
    const size_t o_cs = put(nth, [&](size_t i) { return std::cos(m->h_theta[i]); });
    const size_t o_cot = put(nth, [&](size_t i) { return std::cos(m->h_theta[i]) / std::sin(m->h_theta[i]); });
    const size_t o_d1 = put(nth, [&](size_t i) { return m->h_dsdr1[i]; });
    const size_t o_d2 = put(nth, [&](size_t i) { return m->h_dsdr2[i]; });
    // FP32 states hold s - s_bar: the constant-temperature wall value is stored the same way
    const size_t o_s1 = put(nth, [&](size_t i) { return m->h_sr1[i] - (f32 ? m->h_s_bar[0] : 0.0); });
    // SLD geometry: cell faces at the midpoints, the wall faces half a cell out
    const std::vector<double> &r = m->h_r, &dr = m->h_dr, &th = m->h_theta;
    auto r_lo = [&](size_t i) { return i == 0 ? r[0] - 0.5 * dr[0] : 0.5 * (r[i] + r[i - 1]); };
    auto r_hi = [&](size_t i) { return i + 1 == (size_t)nr ? r[nr - 1] + 0.5 * dr[nr - 1] : 0.5 * (r[i + 1] + r[i]); };
    auto t_lo = [&](size_t i) { return i == 0 ? th[0] - 0.5 * m->dth : 0.5 * (th[i] + th[i - 1]); };
    auto t_hi = [&](size_t i) { return i + 1 == (size_t)nth ? th[nth - 1] + 0.5 * m->dth : 0.5 * (th[i + 1] + th[i]); };
    const size_t o_rfc = put(nr + 1, [&](size_t k) { return (k == 0 || k == (size_t)nr) ? 0.0 : 0.5 * (r[k] + r[k - 1]); });
    const size_t o_rv = put(nr, [&](size_t i) {
        const double a = r_lo(i), b = r_hi(i);
        return 3.0 / (b * b * b - a * a * a);
    });
    const size_t o_sph = put(nth + 1, [&](size_t k) {
        return (k == 0 || k == (size_t)nth) ? 0.0 : std::sin(0.5 * (th[k] + th[k - 1]));
    });
    const size_t o_rfac = put(nr, [&](size_t i) {
        const double a = r_lo(i), b = r_hi(i);
        return 1.5 * (b * b - a * a) / (b * b * b - a * a * a);
    });
    const size_t o_ct = put(nth, [&](size_t i) { return 1.0 / (std::cos(t_lo(i)) - std::cos(t_hi(i))); });
    const size_t o_gt = put(nth, [&](size_t i) {
        return (t_hi(i) - t_lo(i)) / (std::cos(t_lo(i)) - std::cos(t_hi(i)));
    });
    const size_t o_vsp = put(nr, [&](size_t i) { return m->h_diff_speed.empty() ? 0.0 : m->h_diff_speed[i + 1]; });
    auto opt = [](const std::vector<double> &v, size_t i) { return v.empty() ? 0.0 : v[i]; };
    const size_t o_mur = put(nr, [&](size_t i) { return opt(m->h_mur, i); });
    const size_t o_dlm = put(nr, [&](size_t i) { return opt(m->h_dlnmu, i); });
    const size_t o_ks = put(nr, [&](size_t i) { return opt(m->h_ks, i); });
    const size_t o_dlk = put(nr, [&](size_t i) { return opt(m->h_dlnks, i); });
    const size_t o_eta = put(nr, [&](size_t i) { return opt(m->h_etar, i); });
    const size_t o_w9 = put(nr, [&](size_t i) { return opt(m->h_d2xdr2, i) / m->h_dxdr[i]; });
    const size_t o_d2s = put(nr, [&](size_t i) { return opt(m->h_d2sbar, i); });
    const size_t tot = h.size();
    T *d = nullptr;
    if (P.buf) {
        d = P.buf;   // the same tables as the first upload: the same size
    } else if (hipMalloc(&d, tot * sizeof(T)) != hipSuccess) {
        return CSSX_ENOMEM;
    }
    if (hipMemcpy(d, h.data(), tot * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return CSSX_EHIP;
    P.buf = d;
    P.r = d + o_r; P.rinv = d + o_ri; P.dxdr = d + o_x; P.grav = d + o_g; P.eps = d + o_e; P.dr = d + o_dr;
    P.drr = d + o_drr;
    P.rbar = d + o_rb; P.sbar = d + o_sb; P.tbar = d + o_tb; P.tr = d + o_tr; P.hres = d + o_hr; P.drbar = d + o_drb; P.dsbar = d + o_dsb;
    P.sn = d + o_sn; P.csc = d + o_csc; P.cs = d + o_cs; P.cot = d + o_cot;
    P.dsdr1 = d + o_d1; P.dsdr2 = d + o_d2; P.sr1 = d + o_s1;
    P.rfc = d + o_rfc; P.rv = d + o_rv; P.sph = d + o_sph; P.rfac = d + o_rfac; P.ct = d + o_ct; P.gt = d + o_gt;
    P.vsp = d + o_vsp;
    P.mur = d + o_mur; P.dlnmu = d + o_dlm; P.ks = d + o_ks; P.dlnks = d + o_dlk; P.etar = d + o_eta;
    P.w9 = d + o_w9; P.d2sbar = d + o_d2s;
    return CSSX_OK;
}

// Physical radial derivative (times dx/dr) of a radial profile by the FP64 compact operator
// with zero wall values: the background part of d/dr in the FP32 fluctuation form.
static int background_derivative(cssx_mhd *m, const double *prof_dev, int code, double wlo, double whi,
                                 std::vector<double> &out, int order = 1) {
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * (m->nr + 2)) != hipSuccess) return CSSX_ENOMEM;
    // prescribed physical wall derivatives, in x units (the sweep multiplies by dx/dr)
    const double wv[2] = {wlo / m->h_dxdr[0], whi / m->h_dxdr[m->nr - 1]};
    if (hipMemcpy(d + m->nr, wv, sizeof(wv), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return CSSX_EHIP;
    }
    AxisGeom g;
    g.n = m->nr; g.na = 1; g.nb = 1; g.sa = 1; g.sb = 0; g.srow = 1;
    g.g_k = 0; g.g_a = 0; g.g_b = 0; g.phase = CSSX_PHASE_OTHER;
    DerivSpec s;
    s.order = order;
    s.in = prof_dev; s.out = d;
    s.glo = d + m->nr; s.ghi = d + m->nr + 1;
    s.lo_kind = s.hi_kind = END_WALL;
    s.bdry = code;
    // order 2 stays in x-space (the caller applies dx/dr^2)
    s.scale = order == 1 ? m->grid.dri : m->grid.dri * m->grid.dri;
    s.neu_mul = 1.0 / m->grid.dri;
    s.post = order == 1 ? m->p64.dxdr : nullptr;
    int rc = run_derivs(&s, 1, g, CSSX_PREC_F64, nullptr);
    out.assign(m->nr, 0.0);
    if (!rc && (hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy(out.data(), d, sizeof(double) * m->nr, hipMemcpyDeviceToHost) != hipSuccess))
        rc = CSSX_EHIP;
    (void)hipFree(d);
    return rc;
}

template <typename T>
static Scratch<T> &work(cssx_mhd *m);
template <>
Scratch<double> &work<double>(cssx_mhd *m) { return m->s64; }
template <>
Scratch<float> &work<float>(cssx_mhd *m) { return m->s32; }
template <typename T>
static Prof<T> &prof(cssx_mhd *m);
template <>
Prof<double> &prof<double>(cssx_mhd *m) { return m->p64; }
template <>
Prof<float> &prof<float>(cssx_mhd *m) { return m->p32; }

template <typename T>
static int ensure_scratch(cssx_mhd *m) {
    Scratch<T> &S = work<T>(m);
    if (S.ready) return CSSX_OK;
    const long N = m->N;
    WallPtrs W = wall_layout(m->nr, m->nth, m->nphi);
    if (hipMalloc(&S.D, sizeof(T) * N * D_NSLOT) != hipSuccess) return CSSX_ENOMEM;
    if (hipMalloc(&S.B, sizeof(T) * N * 3) != hipSuccess) return CSSX_ENOMEM;
    if (hipMalloc(&S.wall, sizeof(T) * W.total) != hipSuccess) return CSSX_ENOMEM;
    if (hipMalloc(&S.y1, sizeof(T) * N * m->nv) != hipSuccess) return CSSX_ENOMEM;
    if (hipMalloc(&S.k, sizeof(T) * N * m->nv) != hipSuccess) return CSSX_ENOMEM;
    if (m->p.sld_on && hipMalloc(&S.dif, sizeof(T) * N * m->nv) != hipSuccess) return CSSX_ENOMEM;
    if (m->p.sld_on && hipMalloc(&S.vcc, sizeof(T) * N) != hipSuccess) return CSSX_ENOMEM;
    {
        const size_t tw = (size_t)(8 + D_NSLOT) * 2 * m->nr * m->nphi;
        if (hipMalloc(&S.tw, sizeof(T) * tw) != hipSuccess) return CSSX_ENOMEM;
        if (hipMemset(S.tw, 0, sizeof(T) * tw) != hipSuccess) return CSSX_EHIP;   // slots never written stay 0
    }
    if (sizeof(T) == 4 && (hipMalloc(&S.acc, sizeof(T) * N * m->nv) != hipSuccess ||
                           hipMalloc(&S.y0f, sizeof(T) * N * m->nv) != hipSuccess))
        return CSSX_ENOMEM;
    if (hipMemset(S.wall, 0, sizeof(T) * W.total) != hipSuccess) return CSSX_EHIP;
    if (m->p.laplacian || m->p.slaplacian) {
        const long nw = 2 * ((long)m->nth * m->nphi + (long)m->nr * m->nphi);
        if (hipMalloc(&S.L, sizeof(T) * N * L_NSLOT) != hipSuccess) return CSSX_ENOMEM;
        if (hipMalloc(&S.lwall, sizeof(T) * nw) != hipSuccess) return CSSX_ENOMEM;
    }
    S.ready = true;
    return CSSX_OK;
}

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Whether the SLD diffusion of an RHS may run on its own stream: not while the caller's stream
// is captured into a graph (the fork is kept out of captured graphs), not while phase timing is
// on (its phases are timed one after another), and not with CSSX_SLD_FORK=0 (read on every
// call); creates the stream and its events on first use.
static bool sld_fork_ok(cssx_mhd *m, hipStream_t stream) {
    const char *e = std::getenv("CSSX_SLD_FORK");
    if (e && std::atoi(e) == 0) return false;
    if (timing_on()) return false;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return false;
    if (!m->sld_side && hipStreamCreateWithFlags(&m->sld_side, hipStreamNonBlocking) != hipSuccess) {
        m->sld_side = nullptr;
        return false;
    }
    if (!m->ev_sld_fork && hipEventCreateWithFlags(&m->ev_sld_fork, hipEventDisableTiming) != hipSuccess) {
        m->ev_sld_fork = nullptr;
        return false;
    }
    if (!m->ev_sld_join && hipEventCreateWithFlags(&m->ev_sld_join, hipEventDisableTiming) != hipSuccess) {
        m->ev_sld_join = nullptr;
        return false;
    }
    return true;
}

// one batched derivative launch along an axis
template <typename T>
static int derivs_axis(cssx_mhd *m, int axis, DerivSpec *specs, int ns, hipStream_t st) {
    AxisGeom g;
    axis_geometry(axis, m->nth, m->nphi, m->nr, g);
    return run_derivs(specs, ns, g, sizeof(T) == 8 ? CSSX_PREC_F64 : CSSX_PREC_F32, st);
}

// DerivSpec for one field on the single-subdomain walls
template <typename T>
static DerivSpec spec(cssx_mhd *m, int axis, const T *in, T *out, int code, const T *wlo, const T *whi,
                      const T *pre = nullptr, const T *post = nullptr) {
    DerivSpec s;
    s.order = 1;
    s.in = in; s.out = out;
    s.glo = wlo; s.ghi = whi;
    s.pre = pre; s.post = post;
    s.bdry = code;
    if (axis == CSSX_AXIS_PHI) {
        s.lo_kind = s.hi_kind = END_PERIODIC;
        s.scale = m->grid.dpi;
        s.bdry = 0;
    } else {
        s.lo_kind = s.hi_kind = END_WALL;
        const double inv = axis == CSSX_AXIS_R ? m->grid.dri : m->grid.dti;
        s.scale = inv;
        s.neu_mul = 1.0 / inv;
        // wall-value buffers are dense per pencil: (nr, nphi) on theta, (nphi, nth) on r
        if (axis == CSSX_AXIS_THETA) { s.g_k = 0; s.g_a = 1; s.g_b = m->nphi; }
        else { s.g_k = 0; s.g_a = 1; s.g_b = m->nth; }
    }
    return s;
}

// second-derivative DerivSpec (D2bydr2 / D2bydt2 / D2bydp2): scale the inverse spacing squared,
// wall derivatives divided by the radial dri on theta too (derivatives.f90:1322, :1338); r
// results stay in x-space
template <typename T>
static DerivSpec spec2(cssx_mhd *m, int axis, const T *in, T *out, int code, const T *wlo, const T *whi) {
    DerivSpec s = spec<T>(m, axis, in, out, code, wlo, whi);
    s.order = 2;
    s.scale = s.scale * s.scale;
    if (axis != CSSX_AXIS_PHI) s.neu_mul = 1.0 / m->grid.dri;
    return s;
}

template <typename T>
static int curl_impl(cssx_mhd *m, const T *A, T *B, hipStream_t stream, bool zero_walls = true) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    Prof<T> &P = prof<T>(m);
    const Consts<T> c = make_consts<T>(m);
    const long N = m->N;
    const int *ibcr = m->p.ibcr, *ibct = m->p.ibct;
    const T *at = A, *ap = A + N, *ar = A + 2 * N;
    T *bt = B, *bp = B + N, *br = B + 2 * N;
    // B = curl A with every factor applied in the sweeps' epilogues, the second sweep of each
    // component adding to the first (no derivative fields are written):
    //   B_r     =  csc/r d/dth (sin A_phi)  -  csc/r d/dph A_theta
    //   B_phi   = -1/r   d/dth A_r          +  (dx/dr)/r d/dx (r A_theta)
    //   B_theta =  csc/r d/dph A_r          -  (dx/dr)/r d/dx (r A_phi)
    // theta pencils: post = csc (row theta), mul_b = 1/r (slow index r)
    DerivSpec th[2] = {spec<T>(m, CSSX_AXIS_THETA, ap, br, ibct[11], nullptr, nullptr, P.sn, P.csc),
                       spec<T>(m, CSSX_AXIS_THETA, ar, bp, ibct[12], nullptr, nullptr)};
    th[0].mul_b = P.rinv;
    th[1].mul_b = P.rinv;
    th[1].scale = -th[1].scale;
    if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, th, 2, stream))) return rc;
    // phi pencils: mul_a = csc (fast index theta), mul_b = 1/r (slow index r)
    DerivSpec ph[2] = {spec<T>(m, CSSX_AXIS_PHI, ar, bt, 0, nullptr, nullptr),
                       spec<T>(m, CSSX_AXIS_PHI, at, br, 0, nullptr, nullptr)};
    for (DerivSpec &s : ph) {
        s.mul_a = P.csc;
        s.mul_b = P.rinv;
    }
    ph[1].scale = -ph[1].scale;
    ph[1].beta = 1.0;
    if ((rc = derivs_axis<T>(m, CSSX_AXIS_PHI, ph, 2, stream))) return rc;
    // r pencils: pre = r, post = (dx/dr)/r (row r)
    DerivSpec rr[2] = {spec<T>(m, CSSX_AXIS_R, at, bp, ibcr[10], nullptr, nullptr, P.r, P.drr),
                       spec<T>(m, CSSX_AXIS_R, ap, bt, ibcr[11], nullptr, nullptr, P.r, P.drr)};
    rr[0].beta = rr[1].beta = 1.0;
    rr[1].scale = -rr[1].scale;
    if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, rr, 2, stream))) return rc;
    if (zero_walls) {
        const long nw = std::max((long)m->nr * m->nphi, (long)m->nth * m->nphi);
        CSSX_TIMED(CSSX_PHASE_POINTWISE, stream, 0.0,
                   hipLaunchKernelGGL(k_curl_walls<T>, dim3(nblocks(nw, 256)), dim3(256), 0, stream, c, B,
                                      m->p.theta_perf_cond, m->p.bottom_radial_field, m->p.top_radial_field));
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

template <typename T>
static int sld_impl(cssx_mhd *m, const T *st, const T *B, T *dif, hipStream_t stream, T *ohm = nullptr,
                    bool defer_corr = false);

// RHS(call_num) of state st into dv.  With a stage descriptor F (F.stage > 0) the last kernel
// applies that RK4 stage instead of writing the full tendency; with sld_dif the SLD diffusion
// of st is computed into sld_dif once B is known (Compute_Diffusion(1) after RHS(1)).
template <typename T>
static int rhs_body(cssx_mhd *m, const T *st, T *dv, T *Bout, hipStream_t stream, const StageFuse<T> *F,
                    T *sld_dif, bool fork_closure);

template <typename T>
static int rhs_impl(cssx_mhd *m, const T *st, T *dv, T *Bout, hipStream_t stream, const StageFuse<T> *F = nullptr,
                    T *sld_dif = nullptr) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    // 0. the learned closure's tendency of the stage state, on the side stream (it reads only
    //    the state, which nothing here writes before the final kernel).  While the caller's
    //    stream is being captured into a graph the closure stays on that stream: the fork has
    //    not been exercised under capture.
    bool fork_closure = false;
    if (m->closure) {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cap) != hipSuccess) return CSSX_EHIP;
        fork_closure = cap == hipStreamCaptureStatusNone;
    }
    if (fork_closure) {
        const Prof<T> &P = prof<T>(m);
        if (hipEventRecord(m->ev_fork, stream) != hipSuccess || hipStreamWaitEvent(m->side, m->ev_fork, 0) != hipSuccess)
            return CSSX_EHIP;
        rc = closure_tendency(m->closure, sizeof(T) == 8 ? CSSX_PREC_F64 : CSSX_PREC_F32, st, make_consts<T>(m).fluct,
                              P.rbar, P.sbar, m->nr, m->nphi, m->nth, m->ext, m->side);
        // the join event is recorded whatever the closure returned, so that every exit below
        // leaves the caller's stream ordered after the side stream's work
        const hipError_t e = hipEventRecord(m->ev_join, m->side);
        if (rc || e != hipSuccess) {
            if (e == hipSuccess) (void)hipStreamWaitEvent(stream, m->ev_join, 0);
            else (void)hipStreamSynchronize(m->side);
            return rc ? rc : CSSX_EHIP;
        }
    }
    rc = rhs_body<T>(m, st, dv, Bout, stream, F, sld_dif, fork_closure);
    if (rc && fork_closure) (void)hipStreamWaitEvent(stream, m->ev_join, 0);   // failed before its own join
    return rc;
}

template <typename T>
static int rhs_body(cssx_mhd *m, const T *st, T *dv, T *Bout, hipStream_t stream, const StageFuse<T> *F,
                    T *sld_dif, bool fork_closure) {
    int rc = CSSX_OK;
    Scratch<T> &S = work<T>(m);
    Prof<T> &P = prof<T>(m);
    const Consts<T> c = make_consts<T>(m);
    const long N = m->N;
    const int *ibcr = m->p.ibcr, *ibct = m->p.ibct;
    T *D = S.D;
    T *B = Bout ? Bout : S.B;
    const WallPtrs W = wall_layout(m->nr, m->nth, m->nphi);
    T *wall = S.wall;
    const T *rho = st, *w = st + N, *v = st + 2 * N, *u = st + 3 * N, *s = st + 4 * N;
    const unsigned nb = nblocks(N, 256);
    // 1. B = curl A
    if (m->p.magnetic && (rc = curl_impl<T>(m, st + 5 * N, B, stream))) return rc;
    const T *bt = B, *bp = B + N, *br = B + 2 * N;
    // 2. wall values
    {
        long nw = std::max((long)m->nr * m->nphi, (long)m->nth * m->nphi);
        hipLaunchKernelGGL(k_walls_state<T>, dim3(nblocks(nw, 256)), dim3(256), 0, stream, c, P, W, wall, st, B,
                           m->p.bottom_stress_free, m->p.top_stress_free, m->p.bottom_const_flux,
                           m->p.top_const_flux, m->p.theta_normal_field);
    }
    // SLD diffusion of the stage-1 state: it needs only B and the state, and writes the SLD
    // tendency, its speed buffer and (RK4) the Ohmic curl, which nothing before the final kernel
    // reads; off a captured stream and with phase timing off it runs on its own stream
    bool fork_sld = false;
    // the SLD point corrections are left to this RHS's fused stage-1 kernel
    const bool defer_corr = F && F->stage == 1 && F->corr;
    if (sld_dif) {
        fork_sld = sld_fork_ok(m, stream);
        if (fork_sld) {
            if (hipEventRecord(m->ev_sld_fork, stream) != hipSuccess ||
                hipStreamWaitEvent(m->sld_side, m->ev_sld_fork, 0) != hipSuccess)
                return CSSX_EHIP;
            rc = sld_impl<T>(m, st, B, sld_dif, m->sld_side, F ? const_cast<T *>(F->ohm) : nullptr, defer_corr);
            // recorded whatever sld_impl returned: every exit below joins it
            const hipError_t e = hipEventRecord(m->ev_sld_join, m->sld_side);
            if (rc || e != hipSuccess) {
                if (e == hipSuccess) (void)hipStreamWaitEvent(stream, m->ev_sld_join, 0);
                else (void)hipStreamSynchronize(m->sld_side);
                return rc ? rc : CSSX_EHIP;
            }
        } else if ((rc = sld_impl<T>(m, st, B, sld_dif, stream, F ? const_cast<T *>(F->ohm) : nullptr, defer_corr))) {
            return rc;
        }
    }
    // from here every error return first joins the SLD stream
    auto fail = [&](int code) {
        if (fork_sld) (void)hipStreamWaitEvent(stream, m->ev_sld_join, 0);
        return code;
    };
    // the theta-wall block for the density wall closure: without explicit diffusion (whose
    // viscous terms the closure would read too)
    T *tw = (!c.slap && !c.lap) ? S.tw : nullptr;
    const long twM = 2L * m->nr * m->nphi;
    // 3. derivative sweeps
    {
        const int mag = m->p.magnetic;
        DerivSpec th[6] = {spec<T>(m, CSSX_AXIS_THETA, w, D + D_TW * N, ibct[2], nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_THETA, v, D + D_TV * N, ibct[3], wall + W.off_tv_lo, wall + W.off_tv_hi),
                           spec<T>(m, CSSX_AXIS_THETA, u, D + D_TU * N, ibct[4], nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_THETA, s, D + D_TS * N, ibct[5], nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_THETA, bp, D + D_TBP * N, ibct[8], wall + W.off_tbp_lo, wall + W.off_tbp_hi),
                           spec<T>(m, CSSX_AXIS_THETA, br, D + D_TBR * N, ibct[9], nullptr, nullptr)};
        if (tw) {
            // the theta sweeps copy their inputs and results at the theta-wall rows into the
            // theta-wall block (slots: w v u s -> 1..4, bp br -> B 1..2, results -> D slots)
            const int in_slot[6] = {1, 2, 3, 4, 6, 7};
            const int out_slot[6] = {D_TW, D_TV, D_TU, D_TS, D_TBP, D_TBR};
            for (int k = 0; k < 6; k++) {
                th[k].tap_in = tw + in_slot[k] * twM;
                th[k].tap_out = tw + (8 + out_slot[k]) * twM;
                th[k].tap_side = twM / 2;
            }
        }
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, th, mag ? 6 : 4, stream))) return fail(rc);
        DerivSpec ph[7] = {spec<T>(m, CSSX_AXIS_PHI, rho, D + D_PRHO * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, v, D + D_PV * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, w, D + D_PW * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, u, D + D_PU * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, s, D + D_PS * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, bt, D + D_PBT * N, 0, nullptr, nullptr),
                           spec<T>(m, CSSX_AXIS_PHI, br, D + D_PBR * N, 0, nullptr, nullptr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_PHI, ph, mag ? 7 : 5, stream))) return fail(rc);
        DerivSpec rr[6] = {spec<T>(m, CSSX_AXIS_R, w, D + D_RW * N, ibcr[2], wall + W.off_rw_lo, wall + W.off_rw_hi, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, v, D + D_RV * N, ibcr[3], wall + W.off_rv_lo, wall + W.off_rv_hi, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, u, D + D_RU * N, ibcr[4], nullptr, nullptr, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, s, D + D_RS * N, ibcr[5], wall + W.off_rs_lo, wall + W.off_rs_hi, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, bt, D + D_RBT * N, ibcr[7], nullptr, nullptr, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, bp, D + D_RBP * N, ibcr[8], nullptr, nullptr, nullptr, P.dxdr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, rr, mag ? 6 : 4, stream))) return fail(rc);
    }
    // 3b. explicit diffusion: second derivatives (same wall values as the first), the cross-term
    //     sweeps of first-derivative slots, and curl B
    T *L = S.L;
    if (c.slap) {
        const int nd = c.lap ? 4 : 1;
        DerivSpec th[4] = {spec2<T>(m, CSSX_AXIS_THETA, s, L + L_TS2 * N, ibct[5], nullptr, nullptr),
                           spec2<T>(m, CSSX_AXIS_THETA, w, L + L_TW2 * N, ibct[2], nullptr, nullptr),
                           spec2<T>(m, CSSX_AXIS_THETA, v, L + L_TV2 * N, ibct[3], wall + W.off_tv_lo, wall + W.off_tv_hi),
                           spec2<T>(m, CSSX_AXIS_THETA, u, L + L_TU2 * N, ibct[4], nullptr, nullptr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, th, nd, stream))) return fail(rc);
        DerivSpec ph[4] = {spec2<T>(m, CSSX_AXIS_PHI, s, L + L_PS2 * N, 0, nullptr, nullptr),
                           spec2<T>(m, CSSX_AXIS_PHI, v, L + L_PV2 * N, 0, nullptr, nullptr),
                           spec2<T>(m, CSSX_AXIS_PHI, w, L + L_PW2 * N, 0, nullptr, nullptr),
                           spec2<T>(m, CSSX_AXIS_PHI, u, L + L_PU2 * N, 0, nullptr, nullptr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_PHI, ph, nd, stream))) return fail(rc);
        DerivSpec rr[4] = {spec2<T>(m, CSSX_AXIS_R, s, L + L_RS2 * N, ibcr[5], wall + W.off_rs_lo, wall + W.off_rs_hi),
                           spec2<T>(m, CSSX_AXIS_R, w, L + L_RW2 * N, ibcr[2], wall + W.off_rw_lo, wall + W.off_rw_hi),
                           spec2<T>(m, CSSX_AXIS_R, v, L + L_RV2 * N, ibcr[3], wall + W.off_rv_lo, wall + W.off_rv_hi),
                           spec2<T>(m, CSSX_AXIS_R, u, L + L_RU2 * N, ibcr[4], nullptr, nullptr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, rr, nd, stream))) return fail(rc);
    }
    if (c.lap) {
        const long nrr = (long)m->nth * m->nphi, nt = (long)m->nr * m->nphi;
        T *lw = S.lwall;
        hipLaunchKernelGGL(k_lap_walls<T>, dim3(nblocks(std::max(nrr, nt), 256)), dim3(256), 0, stream, c, P,
                           (const T *)D, lw, m->p.bottom_stress_free);
        DerivSpec xr[4] = {spec<T>(m, CSSX_AXIS_R, D + D_TW * N, L + L_XR1 * N, ibcr[2], lw, nullptr, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, D + D_PU * N, L + L_XR2 * N, ibcr[4], nullptr, nullptr, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, D + D_TU * N, L + L_XR3 * N, ibcr[4], nullptr, nullptr, nullptr, P.dxdr),
                           spec<T>(m, CSSX_AXIS_R, D + D_PV * N, L + L_XR4 * N, ibcr[3], nullptr, lw + nrr, nullptr, P.dxdr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, xr, 4, stream))) return fail(rc);
        DerivSpec xt[2] = {spec<T>(m, CSSX_AXIS_THETA, D + D_PV * N, L + L_XT1 * N, ibct[3], lw + 2 * nrr, lw + 2 * nrr + nt),
                           spec<T>(m, CSSX_AXIS_THETA, D + D_PW * N, L + L_XT2 * N, ibct[2], nullptr, nullptr)};
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, xt, 2, stream))) return fail(rc);
        if (m->p.magnetic) {
            // VectorCurl: theta sweeps with code 1 (Theta_Normal_Field) or 0, whose wall values go
            // unused; r sweeps with zero wall values (the radial-field walls zero those halos)
            const int tcode = m->p.theta_normal_field ? 1 : 0;
            DerivSpec jt[2] = {spec<T>(m, CSSX_AXIS_THETA, bp, L + L_JT1 * N, tcode, nullptr, nullptr, P.sn),
                               spec<T>(m, CSSX_AXIS_THETA, br, L + L_JT2 * N, tcode, nullptr, nullptr)};
            if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, jt, 2, stream))) return fail(rc);
            DerivSpec jp[2] = {spec<T>(m, CSSX_AXIS_PHI, br, L + L_JP1 * N, 0, nullptr, nullptr),
                               spec<T>(m, CSSX_AXIS_PHI, bt, L + L_JP2 * N, 0, nullptr, nullptr)};
            if ((rc = derivs_axis<T>(m, CSSX_AXIS_PHI, jp, 2, stream))) return fail(rc);
            DerivSpec jr[2] = {spec<T>(m, CSSX_AXIS_R, bt, L + L_JRT * N, ibcr[7], nullptr, nullptr, P.r, P.dxdr),
                               spec<T>(m, CSSX_AXIS_R, bp, L + L_JRP * N, ibcr[8], nullptr, nullptr, P.r, P.dxdr)};
            if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, jr, 2, stream))) return fail(rc);
        }
    }
    // 4-5. density wall derivatives from the momentum tendency, which k_walls_density evaluates
    //      (every term before the density terms) at the wall points themselves
    {
        long nw = std::max((long)m->nr * m->nphi, (long)m->nth * m->nphi);
        if (tw)
            hipLaunchKernelGGL(k_walls_gather<T>, dim3(nblocks((long)m->nr * m->nphi, 256)), dim3(256), 0, stream, c,
                               st, (const T *)B, (const T *)D, tw);
        hipLaunchKernelGGL(k_walls_density<T>, dim3(nblocks(nw, 256)), dim3(256), 0, stream, c, P, W, wall, st,
                           (const T *)B, (const T *)D, (const T *)L, ibcr[1], (const T *)tw);
    }
    // 6. density sweeps + every pointwise term, the density terms and the stage in one kernel
    {
        DerivSpec th = spec<T>(m, CSSX_AXIS_THETA, rho, D + D_THRHO * N, ibct[1], wall + W.off_trho_lo,
                               wall + W.off_trho_hi);
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_THETA, &th, 1, stream))) return fail(rc);
        DerivSpec rr = spec<T>(m, CSSX_AXIS_R, rho, D + D_RRHO * N, ibcr[1], wall + W.off_rrho_lo,
                               wall + W.off_rrho_hi, nullptr, P.dxdr);
        if ((rc = derivs_axis<T>(m, CSSX_AXIS_R, &rr, 1, stream))) return fail(rc);
        StageFuse<T> F0{};
        const StageFuse<T> &FF = F ? *F : F0;
        const float *ext = nullptr;
        if (fork_closure) {   // join the closure's side stream
            if (hipStreamWaitEvent(stream, m->ev_join, 0) != hipSuccess) return fail(CSSX_EHIP);
            ext = m->ext;
        } else if (m->closure) {
            rc = closure_tendency(m->closure, sizeof(T) == 8 ? CSSX_PREC_F64 : CSSX_PREC_F32, st, c.fluct, P.rbar,
                                  P.sbar, m->nr, m->nphi, m->nth, m->ext, stream);
            if (rc) return fail(rc);
            ext = m->ext;
        }
        // reads: 5 state, 3 B, 19 (13) first-derivative and 2 density-derivative fields, 18 (3)
        // explicit-diffusion fields with Laplacian (SLaplacian); the plain RHS writes 8 (5)
        // tendencies, a fused stage moves the stage traffic of every field instead
        const bool mag = m->p.magnetic;
        const double reads = (mag ? 29.0 : 20.0) + (c.lap ? 18.0 : c.slap ? 3.0 : 0.0);
        const double fb = !F ? (mag ? 8.0 : 5.0) * sizeof(T)
                          : sizeof(T) == 8 ? (FF.stage == 1 ? 24.0 + (FF.dif ? 8.0 : 0.0) : FF.stage < 4 ? 32.0 : 24.0) * m->nv
                                           : (FF.stage == 1 ? 12.0 + (FF.dif ? 4.0 : 0.0) : FF.stage < 4 ? 16.0 : 24.0) * m->nv;
        const double fo = (F && FF.ohm && FF.stage == 1) ? 3.0 * sizeof(T) : 0.0;   // the folded Ohmic curl
        if (fork_sld && hipStreamWaitEvent(stream, m->ev_sld_join, 0) != hipSuccess) {
            (void)hipStreamSynchronize(m->sld_side);
            return CSSX_EHIP;
        }
        CSSX_TIMED(CSSX_PHASE_RHS_FINAL, stream, reads * N * sizeof(T) + (fb + fo) * N + (ext ? 32.0 * N : 0.0),
                   hipLaunchKernelGGL(k_rhs_final<T>, dim3(nb), dim3(256), 0, stream, c, P, st, B,
                                      (const T *)(D + D_THRHO * N), (const T *)(D + D_RRHO * N), dv, FF, ext,
                                      (const T *)D, (const T *)L));
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

template <typename T>
static SldParams<T> make_sld(const cssx_mhd *m) {
    SldParams<T> sp;
    const cssx_mhd_params &p = m->p;
    sp.coef = (T)p.sld_coef;
    const double tv[8] = {1e-8 * m->h_rho_bar[m->nr - 1], 1e-8, 1e-8, 1e-8, 1e-3, 1.0, 1.0, 1.0};
    for (int f = 0; f < 8; f++) {
        sp.tiny[f] = (T)tv[f];
        sp.itiny2[f] = (T)(1.0 / (tv[f] * tv[f]));
        const int iv = f + 1;
        // Diffusion_R scales rho,w,v by sld_pr and s,A by sld_prm (:846-854); Diffusion_Th/Phi
        // scale rho,w,v,u by sld_pr and A by sld_prm (:1252-1258, :1476-1482)
        sp.mr[f] = (T)(iv <= 3 ? p.sld_pr : (iv >= 5 ? p.sld_prm : 1.0));
        sp.mt[f] = (T)(iv < 5 ? p.sld_pr : (iv > 5 ? p.sld_prm : 1.0));
    }
    sp.invdphi = (T)(1.0 / m->dphi);
    sp.gcv = (T)(m->gamm1 * m->Cv);
    sp.first = p.mass_diffusion ? 0 : 1;
    sp.last = (p.magnetic && p.magnetic_diffusion) ? 7 : 4;
    sp.dstype = p.dstype;
    sp.magnetic = p.magnetic;
    sp.visc = p.viscous_heating;
    sp.mass = p.mass_diffusion;
    sp.ohmic = p.magnetic && p.ohmic_heating && p.magnetic_diffusion;
    sp.bot_pc = p.bottom_perf_cond;
    sp.top_pc = p.top_perf_cond;
    sp.th_normal = p.theta_normal_field;
    sp.th_pc = p.theta_perf_cond;
    const bool slap = p.laplacian || p.slaplacian;
    sp.sflux_lo = slap && p.bottom_const_flux;
    sp.sflux_hi = slap && p.top_const_flux;
    sp.soff_lo = sp.soff_hi = T(0);
    if (slap) {
        const int nr = m->nr;
        const double r1 = m->h_r[0], r2 = m->h_r[nr - 1], d1 = m->h_dr[0];
        sp.soff_lo = (T)(d1 * p.lumr1 * kOneOver4Pi / (r1 * r1 * m->h_ks[0] * m->h_dxdr[0]));
        sp.soff_hi = (T)(d1 * p.lumr2 * kOneOver4Pi / (r2 * r2 * m->h_ks[nr - 1] * m->h_dxdr[nr - 1]));
    }
    return sp;
}

// Compute_Diffusion(1): dif (nv fields) = the SLD tendency of state st, B = curl A with the
// Compute_B wall conditions (magnetic runs).  dif must not alias the derivative scratch (the
// Ohmic curl uses slots 0-8, the face speeds 15-17, the SLD speed slot 18).
// The FP32 SLD runs in two passes (rho..s, then A), as FP64 does: each pass keeps more
// wavefronts per SIMD than the single eight-field pass (126 and 93 VGPRs against 160), which
// measured 1 ms per 256x512x1024 step faster.  CSSX_SLD_SPLIT=0 selects the single pass (a
// tuning hook, read once).
static bool sld_split() {
    static const int v = [] {
        const char *e = std::getenv("CSSX_SLD_SPLIT");
        return (e && std::atoi(e) == 0) ? 0 : 1;
    }();
    return v != 0;
}

template <typename T>
// ohm: with Ohmic heating, a 3-field buffer that receives the curl of the A tendency and the
// caller applies the heating term (the fused RK4 stage 1); null: k_sld_finish applies it to dif
static int sld_impl(cssx_mhd *m, const T *st, const T *B, T *dif, hipStream_t stream, T *ohm, bool defer_corr) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    Scratch<T> &S = work<T>(m);
    Prof<T> &P = prof<T>(m);
    const Consts<T> c = make_consts<T>(m);
    const SldParams<T> sp = make_sld<T>(m);
    const long N = m->N;
    T *vcc = S.vcc;
    const unsigned nb = nblocks(N, 256);
    const int nv = m->nv;
    const double ts = (double)sizeof(T);
    const double b_speed = ts * N * (1.0 + (sp.dstype == 1 ? 3.0 + (sp.magnetic ? 4.0 : 0.0) : 0.0));
    CSSX_TIMED(CSSX_PHASE_SLD, stream, b_speed,
               hipLaunchKernelGGL(k_sld_speed<T>, dim3(nb), dim3(256), 0, stream, c, P, sp, st, B, vcc));
    // tiles of 64 theta x kSldRows phi columns (one thread each); the r range is cut into segments of
    // at least 16 levels so that the launch holds ~16K wavefronts where the grid allows
    const int ntt = (m->nth + 63) / 64, ntp = (m->nphi + kSldRows - 1) / kSldRows;
    const long tiles = (long)ntt * ntp;
    int nseg = (int)std::min<long>((4096 + tiles - 1) / tiles, std::max(1, m->nr / 16));
    nseg = std::max(nseg, 1);
    const int rseg = (m->nr + nseg - 1) / nseg;
    nseg = (m->nr + rseg - 1) / rseg;
    const double nd = (double)(sp.last - sp.first + 1);
    const double b_march = ts * N * (nd + 1.0 + (double)nv + ((sp.visc || sp.mass) ? 2.0 : 0.0));
    // rho..s and the A components in two passes (FP32 may take all eight in one, see
    // sld_split).  The kernel addresses a field with 32-bit buffer offsets.
    if ((double)N * sizeof(T) >= 4294967295.0) return CSSX_EINVAL;
    const dim3 grid((unsigned)(tiles * nseg));
    const double na = sp.last >= 5 ? (double)(sp.last - 4) : 0.0;   // diffused A components
    auto two_pass = [&]() {
        CSSX_TIMED(CSSX_PHASE_SLD, stream, b_march - ts * N * 2.0 * na,
                   hipLaunchKernelGGL((k_sld_col<T, 0, 5>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st, (const T *)vcc,
                                      dif, ntt, ntp, rseg));
        if (na > 0.0)
            CSSX_TIMED(CSSX_PHASE_SLD, stream, ts * N * (2.0 * na + 1.0),
                       hipLaunchKernelGGL((k_sld_col<T, 5, 3>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st,
                                          (const T *)vcc, dif, ntt, ntp, rseg));
    };
    if (defer_corr) {
        // the point corrections are the caller's: rho..s in two passes of at most three fields,
        // which fit three (FP32) / two (FP64) workgroups per CU, then A
        CSSX_TIMED(CSSX_PHASE_SLD, stream, ts * N * 7.0,
                   hipLaunchKernelGGL((k_sld_col<T, 0, 3>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st,
                                      (const T *)vcc, dif, ntt, ntp, rseg));
        CSSX_TIMED(CSSX_PHASE_SLD, stream, ts * N * 5.0,
                   hipLaunchKernelGGL((k_sld_col<T, 3, 2>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st,
                                      (const T *)vcc, dif, ntt, ntp, rseg));
        if (na > 0.0)
            CSSX_TIMED(CSSX_PHASE_SLD, stream, ts * N * (2.0 * na + 1.0),
                       hipLaunchKernelGGL((k_sld_col<T, 5, 3>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st,
                                          (const T *)vcc, dif, ntt, ntp, rseg));
    } else if constexpr (sizeof(T) == 4) {
        if (sp.last >= 5 && !sld_split())
            CSSX_TIMED(CSSX_PHASE_SLD, stream, b_march,
                       hipLaunchKernelGGL((k_sld_col<T, 0, 8>), grid, dim3(64 * kSldRows), 0, stream, c, P, sp, st,
                                          (const T *)vcc, dif, ntt, ntp, rseg));
        else
            two_pass();
    } else {
        two_pass();
    }
    if (sp.ohmic && ohm) {
        if ((rc = curl_impl<T>(m, dif + 5 * N, ohm, stream, false))) return rc;
    } else if (sp.ohmic) {
        T *C = S.D + 6 * N;
        if ((rc = curl_impl<T>(m, dif + 5 * N, C, stream, false))) return rc;
        CSSX_TIMED(CSSX_PHASE_SLD, stream, ts * N * (3.0 + 3.0 + 2.0 + 2.0),
                   hipLaunchKernelGGL(k_sld_finish<T>, dim3(nb), dim3(256), 0, stream, c, P, st, B, (const T *)C,
                                      dif));
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

template <typename T>
static int rk4_impl(cssx_mhd *m, double *y0, double dt, int nsteps, hipStream_t stream) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    Scratch<T> &S = work<T>(m);
    const Consts<T> c = make_consts<T>(m);
    const long NV = m->N * m->nv;
    const unsigned nb = nblocks(NV, 256);
    T *dif = m->p.sld_on ? S.dif : nullptr;
    // SLD Ohmic heating folded into stage 1: the curl of the A tendency goes to S.k, which the
    // fused stages do not use (they write no tendency field)
    const SldParams<T> sp = make_sld<T>(m);
    const bool ohm = dif && sp.ohmic;
    constexpr bool mixed = sizeof(T) != sizeof(double);
    StageFuse<T> F{};
    // stage 1's final kernel applies the SLD point corrections (the march then runs rho..s in
    // passes of fewer registers)
    F.corr = dif && (sp.visc || sp.mass);
    F.visc = sp.visc;
    F.mass = sp.mass;
    F.gcv = sp.gcv;
    F.dt = dt;
    F.f = make_bound_flags(m);
    F.y0 = y0;
    F.acc64 = m->save;
    F.acc32 = mixed ? (float *)S.acc : nullptr;
    F.y0f = mixed ? S.y0f : nullptr;
    F.y1 = S.y1;
    F.sr1_64 = m->p64.sr1;
    F.rbar64 = m->p64.rbar;
    F.sbar64 = m->p64.sbar;
    F.sr1_t = prof<T>(m).sr1;
    if (mixed)
        hipLaunchKernelGGL(k_cast_state<T>, dim3(nb), dim3(256), 0, stream, c, m->nv, (const double *)m->p64.rbar,
                           (const double *)m->p64.sbar, (const double *)y0, S.y0f);
    // stage 1 reads vars1 = vars0 (its FP32 copy on the mixed path); each stage's update is
    // fused into the RHS's last kernel; Compute_Diffusion(1) runs inside stage 1's RHS
    const T *in1 = mixed ? (const T *)S.y0f : (const T *)y0;
    for (int step = 0; step < nsteps; step++) {
        for (int stage = 1; stage <= 4; stage++) {
            F.stage = stage;
            F.dif = stage == 1 ? dif : nullptr;
            F.ohm = (stage == 1 && ohm) ? S.k : nullptr;
            if ((rc = rhs_impl<T>(m, stage == 1 ? in1 : S.y1, S.k, nullptr, stream, &F, stage == 1 ? dif : nullptr)))
                return rc;
        }
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

// nsteps of Step_RK3_TVD (order 3) or Step_RK2_TVD (order 2)
template <typename T>
static int tvd_impl(cssx_mhd *m, double *y0, double dt, int nsteps, int order, hipStream_t stream) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    Scratch<T> &S = work<T>(m);
    const Consts<T> c = make_consts<T>(m);
    const BoundFlags f = make_bound_flags(m);
    const long NV = m->N * m->nv;
    const unsigned nb = nblocks(NV, 256);
    const bool f32 = sizeof(T) != sizeof(double);
    T *dif = m->p.sld_on ? S.dif : nullptr;
    const double ts = (double)sizeof(T);
    auto stage = [&](int mode, double a0, double a1, double dtc, int last) {
        const double by = (ts + 8.0 + (mode ? 8.0 : 0.0) + (last && dif ? ts : 0.0) +
                           (last ? 8.0 + (f32 ? ts : 0.0) : 8.0 + ts)) * (double)NV;
        CSSX_TIMED(CSSX_PHASE_RK_UPDATE, stream, by,
                   hipLaunchKernelGGL(k_tvd_stage<T>, dim3(nb), dim3(256), 0, stream, c, f,
                                      (const double *)m->p64.sr1, (const double *)m->p64.rbar,
                                      (const double *)m->p64.sbar, mode, a0, a1, dtc, dt, last, y0, m->save,
                                      (const T *)S.k, (const T *)dif, S.y1, m->nv));
    };
    if (f32)
        hipLaunchKernelGGL(k_cast_state<T>, dim3(nb), dim3(256), 0, stream, c, m->nv, (const double *)m->p64.rbar,
                           (const double *)m->p64.sbar, (const double *)y0, S.y1);
    for (int step = 0; step < nsteps; step++) {
        const T *in1 = f32 ? S.y1 : (const T *)y0;
        if ((rc = rhs_impl<T>(m, in1, S.k, nullptr, stream))) return rc;
        if (dif && (rc = sld_impl<T>(m, in1, S.B, dif, stream))) return rc;
        stage(0, 1.0, 0.0, dt, 0);
        if ((rc = rhs_impl<T>(m, S.y1, S.k, nullptr, stream))) return rc;
        if (order == 2) {
            stage(1, 0.5, 0.5, 0.5 * dt, 1);
        } else {
            stage(1, 0.75, 0.25, dt / 4.0, 0);
            if ((rc = rhs_impl<T>(m, S.y1, S.k, nullptr, stream))) return rc;
            stage(1, 1.0 / 3.0, 2.0 / 3.0, 2.0 * dt / 3.0, 1);
        }
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

}  // namespace cssx

using namespace cssx;

extern "C" int cssx_mhd_create(const cssx_grid *grid, const cssx_mhd_params *params, const cssx_mhd_profiles *pr,
                               cssx_mhd **out) {
    if (!grid || !params || !pr || !out) return CSSX_EINVAL;
    *out = nullptr;
    if (!pr->r || !pr->dr || !pr->dxdr || !pr->theta || !pr->gravity || !pr->eps || !pr->dsdr1 || !pr->dsdr2 ||
        !pr->sr1 || !pr->sr2 || !pr->rho_bar || !pr->s_bar)
        return CSSX_EINVAL;
    if (grid->nr < 8 || grid->nth < 8 || grid->nphi < 8 || choose_M(grid->nr) < 0 || choose_M(grid->nth) < 0 ||
        choose_M(grid->nphi) < 0)
        return CSSX_EINVAL;
    if (!grid->r_lo_wall || !grid->r_hi_wall || !grid->th_lo_wall || !grid->th_hi_wall) return CSSX_EINVAL;
    if (params->gamma <= 1.0 || params->Cp <= 0.0) return CSSX_EINVAL;
    for (int i = 1; i <= 12; i++)
        if (params->ibcr[i] < 0 || params->ibcr[i] > 4 || params->ibct[i] < 0 || params->ibct[i] > 4)
            return CSSX_EINVAL;
    // explicit diffusion needs its profiles
    const bool slap = params->laplacian || params->slaplacian;
    if (slap && (!pr->ks || !pr->dlnks || !pr->d2xdr2)) return CSSX_EINVAL;
    if (params->laplacian && (!pr->mur || !pr->dlnmu || !pr->etar)) return CSSX_EINVAL;
    cssx_mhd *m = new (std::nothrow) cssx_mhd;
    if (!m) return CSSX_ENOMEM;
    m->grid = *grid;
    m->p = *params;
    m->nr = grid->nr; m->nth = grid->nth; m->nphi = grid->nphi;
    m->nv = params->magnetic ? 8 : 5;
    m->N = (long)grid->nr * grid->nth * grid->nphi;
    m->Cv = params->Cp / params->gamma;
    m->invCv = 1.0 / m->Cv;
    m->gamm1 = params->gamma - 1.0;
    m->pconst = m->gamm1 * params->Cp;
    m->invpconst = (1.0 / params->Cp) / m->gamm1;
    m->invtconst = m->invCv / m->gamm1;
    const int nr = grid->nr, nth = grid->nth;
    m->h_r.assign(pr->r, pr->r + nr);
    m->h_dr.assign(pr->dr, pr->dr + nr);
    m->h_dxdr.assign(pr->dxdr, pr->dxdr + nr);
    m->h_theta.assign(pr->theta, pr->theta + nth);
    m->h_grav.assign(pr->gravity, pr->gravity + nr);
    m->h_eps.assign(pr->eps, pr->eps + nr);
    m->h_dsdr1.assign(pr->dsdr1, pr->dsdr1 + nth);
    m->h_dsdr2.assign(pr->dsdr2, pr->dsdr2 + nth);
    m->h_sr1.assign(pr->sr1, pr->sr1 + nth);
    m->h_sr2.assign(pr->sr2, pr->sr2 + nth);
    m->h_rho_bar.assign(pr->rho_bar, pr->rho_bar + nr);
    m->h_s_bar.assign(pr->s_bar, pr->s_bar + nr);
    if (pr->diffusion_speed) m->h_diff_speed.assign(pr->diffusion_speed, pr->diffusion_speed + nr + 2);
    if (slap) {
        m->h_ks.assign(pr->ks, pr->ks + nr);
        m->h_dlnks.assign(pr->dlnks, pr->dlnks + nr);
        m->h_d2xdr2.assign(pr->d2xdr2, pr->d2xdr2 + nr);
    }
    if (params->laplacian) {
        m->h_mur.assign(pr->mur, pr->mur + nr);
        m->h_dlnmu.assign(pr->dlnmu, pr->dlnmu + nr);
        m->h_etar.assign(pr->etar, pr->etar + nr);
    }
    m->dth = (m->h_theta[nth - 1] - m->h_theta[0]) / (nth - 1);
    m->dphi = 1.0 / grid->dpi;
    m->h_tbar.resize(nr);
    for (int i = 0; i < nr; i++)
        m->h_tbar[i] = m->invtconst * std::pow(m->h_rho_bar[i], m->gamm1) * std::exp(m->h_s_bar[i] * m->invCv);
    // background derivatives need the FP64 profiles and the plans on the device first
    int rc = upload_profiles<double>(m, m->p64);
    if (!rc) {
        cssx_ctx *ctx = nullptr;
        rc = cssx_ctx_create(grid, &ctx);
        cssx_ctx_destroy(ctx);
    }
    // s_bar with zero wall derivatives; rho_bar with the wall derivatives that make the
    // background's radial momentum balance exactly on Neumann density walls (the density wall
    // rule of Density_Terms_Radial, rhs_mpi.f90:1750-1773, applied to the background)
    if (!rc) rc = background_derivative(m, m->p64.sbar, params->ibcr[5], 0.0, 0.0, m->h_dsbar);
    if (!rc && slap) rc = background_derivative(m, m->p64.sbar, params->ibcr[5], 0.0, 0.0, m->h_d2sbar, 2);
    if (!rc) {
        const int code = params->ibcr[1];
        auto du_bar = [&](int i) { return -m->h_grav[i] - m->gamm1 * m->h_tbar[i] * m->h_dsbar[i]; };
        auto wall = [&](int i) { return du_bar(i) * m->h_rho_bar[i] / (m->pconst * m->h_tbar[i]); };
        m->wbar_lo = (code == 2 || code == 4) ? wall(0) : 0.0;
        m->wbar_hi = (code == 3 || code == 4) ? wall(nr - 1) : 0.0;
        rc = background_derivative(m, m->p64.rbar, code, m->wbar_lo, m->wbar_hi, m->h_drbar);
        m->h_hres.resize(nr);
        for (int i = 0; i < nr && !rc; i++)
            m->h_hres[i] = du_bar(i) - m->pconst * (m->h_tbar[i] / m->h_rho_bar[i]) * m->h_drbar[i];
    }
    if (!rc) rc = upload_profiles<double>(m, m->p64);
    if (!rc) rc = upload_profiles<float>(m, m->p32);
    if (!rc && hipMalloc(&m->save, sizeof(double) * m->N * m->nv) != hipSuccess) rc = CSSX_ENOMEM;
    if (rc) {
        cssx_mhd_destroy(m);
        return rc;
    }
    *out = m;
    return CSSX_OK;
}

extern "C" void cssx_mhd_destroy(cssx_mhd *m) {
    if (!m) return;
    // the closure's side stream may still read ext and the state: drain it before any free
    if (m->side) {
        (void)hipStreamSynchronize(m->side);
        (void)hipStreamDestroy(m->side);
        m->side = nullptr;
    }
    if (m->sld_side) {
        (void)hipStreamSynchronize(m->sld_side);
        (void)hipStreamDestroy(m->sld_side);
        m->sld_side = nullptr;
    }
    for (void *p : {(void *)m->p64.buf, (void *)m->p32.buf, (void *)m->s64.D, (void *)m->s64.B, (void *)m->s64.wall,
                    (void *)m->s64.y1, (void *)m->s64.k, (void *)m->s32.D, (void *)m->s32.B, (void *)m->s32.wall,
                    (void *)m->s32.y1, (void *)m->s32.k, (void *)m->save, (void *)m->red, (void *)m->s64.dif,
                    (void *)m->s32.dif, (void *)m->s32.acc, (void *)m->s32.y0f, (void *)m->ext, (void *)m->s64.L,
                    (void *)m->s64.lwall, (void *)m->s32.L, (void *)m->s32.lwall, (void *)m->s64.vcc,
                    (void *)m->s32.vcc, (void *)m->s64.tw, (void *)m->s32.tw})
        if (p) (void)hipFree(p);
    if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
    if (m->ev_join) (void)hipEventDestroy(m->ev_join);
    if (m->ev_sld_fork) (void)hipEventDestroy(m->ev_sld_fork);
    if (m->ev_sld_join) (void)hipEventDestroy(m->ev_sld_join);
    delete m;
}

extern "C" int cssx_mhd_set_closure(cssx_mhd *m, cssx_mlp *net) {
    if (!m) return CSSX_EINVAL;
    if (net && m->nv != 8) return CSSX_EINVAL;
    if (net && !m->ext && hipMalloc((void **)&m->ext, (size_t)8 * m->N * sizeof(float)) != hipSuccess)
        return CSSX_ENOMEM;
    // each handle is created only if missing, so a call that failed part-way can be retried
    if (net && !m->side && hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking) != hipSuccess) {
        m->side = nullptr;
        return CSSX_EHIP;
    }
    if (net && !m->ev_fork && hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess) {
        m->ev_fork = nullptr;
        return CSSX_EHIP;
    }
    if (net && !m->ev_join && hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming) != hipSuccess) {
        m->ev_join = nullptr;
        return CSSX_EHIP;
    }
    m->closure = net;
    return CSSX_OK;
}

extern "C" int cssx_rhs(cssx_mhd *m, int prec, const void *state, void *rhs, void *b_out, void *stream) {
    if (!m || !state || !rhs) return CSSX_EINVAL;
    if (prec == CSSX_PREC_F64)
        return rhs_impl<double>(m, (const double *)state, (double *)rhs, (double *)b_out, (hipStream_t)stream);
    if (prec == CSSX_PREC_F32)
        return rhs_impl<float>(m, (const float *)state, (float *)rhs, (float *)b_out, (hipStream_t)stream);
    return CSSX_EINVAL;
}

extern "C" int cssx_bound(cssx_mhd *m, int prec, void *state, void *stream) {
    if (!m || !state) return CSSX_EINVAL;
    const long NV = m->N * m->nv;
    const BoundFlags f = make_bound_flags(m);
    if (prec == CSSX_PREC_F64)
        hipLaunchKernelGGL(k_bound<double>, dim3(nblocks(NV, 256)), dim3(256), 0, (hipStream_t)stream,
                           make_consts<double>(m), f, (const double *)m->p64.sr1, (double *)state, m->nv);
    else if (prec == CSSX_PREC_F32)
        hipLaunchKernelGGL(k_bound<float>, dim3(nblocks(NV, 256)), dim3(256), 0, (hipStream_t)stream,
                           make_consts<float>(m), f, (const float *)m->p32.sr1, (float *)state, m->nv);
    else
        return CSSX_EINVAL;
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_rk4(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream) {
    if (!m || !y0 || nsteps < 0 || !(dt > 0.0)) return CSSX_EINVAL;
    if (prec == CSSX_PREC_F64) return rk4_impl<double>(m, (double *)y0, dt, nsteps, (hipStream_t)stream);
    if (prec == CSSX_PREC_F32) return rk4_impl<float>(m, (double *)y0, dt, nsteps, (hipStream_t)stream);
    return CSSX_EINVAL;
}

template <typename T>
static int diffusion_entry(cssx_mhd *m, const T *st, const T *b, T *dif, hipStream_t stream) {
    int rc = ensure_scratch<T>(m);
    if (rc) return rc;
    Scratch<T> &S = work<T>(m);
    const T *B = b;
    if (m->p.magnetic && !B) {
        if ((rc = curl_impl<T>(m, st + 5 * m->N, S.B, stream))) return rc;
        B = S.B;
    }
    return sld_impl<T>(m, st, B, dif, stream);
}

extern "C" int cssx_diffusion(cssx_mhd *m, int prec, const void *state, const void *b, void *dif, void *stream) {
    if (!m || !state || !dif) return CSSX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (prec == CSSX_PREC_F64) return diffusion_entry<double>(m, (const double *)state, (const double *)b, (double *)dif, st);
    if (prec == CSSX_PREC_F32) return diffusion_entry<float>(m, (const float *)state, (const float *)b, (float *)dif, st);
    return CSSX_EINVAL;
}

extern "C" int cssx_rk3_tvd(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream) {
    if (!m || !y0 || nsteps < 0 || !(dt > 0.0)) return CSSX_EINVAL;
    if (prec == CSSX_PREC_F64) return tvd_impl<double>(m, (double *)y0, dt, nsteps, 3, (hipStream_t)stream);
    if (prec == CSSX_PREC_F32) return tvd_impl<float>(m, (double *)y0, dt, nsteps, 3, (hipStream_t)stream);
    return CSSX_EINVAL;
}

extern "C" int cssx_rk2_tvd(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream) {
    if (!m || !y0 || nsteps < 0 || !(dt > 0.0)) return CSSX_EINVAL;
    if (prec == CSSX_PREC_F64) return tvd_impl<double>(m, (double *)y0, dt, nsteps, 2, (hipStream_t)stream);
    if (prec == CSSX_PREC_F32) return tvd_impl<float>(m, (double *)y0, dt, nsteps, 2, (hipStream_t)stream);
    return CSSX_EINVAL;
}

// the block partials of k_cfl for state y (6 per block, host vector)
static int cfl_partials(cssx_mhd *m, const double *y, hipStream_t st, std::vector<double> &h) {
    int rc = ensure_scratch<double>(m);
    if (rc) return rc;
    const long N = m->N;
    if (m->p.magnetic && (rc = curl_impl<double>(m, y + 5 * N, m->s64.B, st))) return rc;
    const unsigned nb = std::min(nblocks(N, 256), (unsigned)kCflMaxBlocks);
    if (m->nred < (int)nb) {
        if (m->red) (void)hipFree(m->red);
        m->red = nullptr;
        if (hipMalloc(&m->red, sizeof(double) * 6 * kCflMaxBlocks) != hipSuccess) return CSSX_ENOMEM;
        m->nred = kCflMaxBlocks;
    }
    hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, st, make_consts<double>(m), m->p64, m->dth, m->dphi, m->p.Cp,
                       y, (const double *)m->s64.B, m->red);
    h.assign(6 * (size_t)nb, 0.0);
    hipError_t e = hipMemcpyAsync(h.data(), m->red, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

// dt from the crossing-time minima, with the explicit diffusion limit (rhs_mpi.f90:194-198,
// :222-229)
static int cfl_dt(const cssx_mhd *m, double dtc, double dtb, double cfl_safety, double *dt_out) {
    double dtd = 1e10;
    if (m->p.laplacian) {
        const int nr = m->nr, nth = m->nth;
        const double drmin = *std::min_element(m->h_dr.begin(), m->h_dr.end());
        const double sinmin = std::min(std::sin(m->h_theta[0]), std::sin(m->h_theta[nth - 1]));
        const double dx = std::min(drmin, m->h_r[0] * std::min(sinmin * m->dphi, m->dth));
        double dmax = 0.0;
        for (int i = 0; i < nr; i++)
            dmax = std::max({dmax, m->h_ks[i], m->h_mur[i] / m->h_rho_bar[i], m->h_etar[i]});
        dtd = dx * dx / dmax;
    }
    *dt_out = std::min(std::min(dtb, dtc), dtd) * cfl_safety;
    return std::isfinite(*dt_out) ? CSSX_OK : CSSX_ENAN;
}

extern "C" int cssx_cfl(cssx_mhd *m, const void *y0_f64, double cfl_safety, double *dt_out, void *stream) {
    if (!m || !y0_f64 || !dt_out || !(cfl_safety > 0.0)) return CSSX_EINVAL;
    std::vector<double> h;
    int rc = cfl_partials(m, (const double *)y0_f64, (hipStream_t)stream, h);
    if (rc) return rc;
    double dtc = 1e4, dtb = 1e4;
    for (size_t b = 0; b < h.size() / 6; b++) {
        dtc = std::min(dtc, h[6 * b]);
        dtb = std::min(dtb, h[6 * b + 1]);
    }
    return cfl_dt(m, dtc, dtb, cfl_safety, dt_out);
}

extern "C" int cssx_conservation(cssx_mhd *m, const void *y0_f64, double *out4, void *stream) {
    if (!m || !y0_f64 || !out4) return CSSX_EINVAL;
    std::vector<double> h;
    int rc = cfl_partials(m, (const double *)y0_f64, (hipStream_t)stream, h);
    if (rc) return rc;
    double acc[4] = {0, 0, 0, 0};
    for (size_t b = 0; b < h.size() / 6; b++)
        for (int q = 0; q < 4; q++) acc[q] += h[6 * b + 2 + q];
    for (int q = 0; q < 4; q++) out4[q] = acc[q];
    return (std::isfinite(acc[0]) && std::isfinite(acc[1])) ? CSSX_OK : CSSX_ENAN;
}
