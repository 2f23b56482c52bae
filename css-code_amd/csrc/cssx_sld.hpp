// SLD diffusion of Compute_Diffusion(1) (diffusion.f90:70-474, Diffusion_R/Th/Phi :476-1495)
// on gfx950.  Included by cssx_rhs.hip after the state/profile types.
//
// Two launches plus, with Ohmic heating, the curl of the A tendency and a third:
//   k_sld_speed   vcc = speed^2(r) [+ |u|^2 + B^2/(4 pi rho) for dstype 1]   (:177-302)
//   k_sld_col     every field's tendency: one thread per (theta, phi) column marching up r,
//                 with a four-level window of every diffused field in registers and the upper
//                 r face's flux carried to the next level (every r face limited once); the
//                 theta and phi faces of each cell from the neighbouring columns at the level
//                 (cache hits).  Then the viscous heating and mass diffusion corrections
//                 (:375-428), which need only the point itself (under RK4 the stage-1 final
//                 kernel applies them, and rho..s runs as two smaller passes).
//   (curl of the A tendency, the same batched derivative sweeps as Compute_B without its wall
//    zeroing, VectorCurlA of :437)
//   k_sld_finish  Ohmic heating (:430-461)
// The background meanV removed from rho and s (:168-175) is the hydrostatic profile
// (update_background = .False.); FP32 states already store exactly those fluctuations.
// Ghost values at the four walls follow the per-variable rules of Diffusion_R (:586-711) and
// Diffusion_Th (:1008-1120); the ghost-cell speeds are copies of the nearest cell
// (:290-302) and phi is periodic.  Face values enter only through the face-flux weights rfc
// and sph, which are zero on the walls.
#pragma once

#include "cssx_bufio.hpp"

namespace cssx {

template <typename T>
struct SldParams {
    T coef;                 // sld_coef
    T tiny[8];              // tinyvar by field (diffusion.f90:76-80)
    T itiny2[8];            // 1 / tinyvar^2
    T mr[8], mt[8];         // sld_pr / sld_prm multipliers of the r and the theta/phi fluxes
    T invdphi, gcv;         // 1/dphi, (gamma-1) Cv
    int first, last;        // diffused fields, 0-based (varlist of :86-112)
    int dstype, magnetic, visc, mass, ohmic;
    int bot_pc, top_pc, th_normal, th_pc;
    // entropy ghost of a constant-flux radial wall under explicit thermal diffusion
    // (diffusion.f90:614-618, :688-692): wall value +- dr_ind(1) lumr/(4 pi r^2 ks dxdr)
    int sflux_lo, sflux_hi;
    T soff_lo, soff_hi;
};

// Van Albada slope of diffusion.f90:744-762, q * max(0, min(1, q qph / max(|q|, tiny)^2)).
// Where |q| >= tiny this is minmod(q, qph) exactly, which needs no reciprocal; below tiny the
// closed form q min(1, q qph / tiny^2) is kept.  In FP32 both are medians of three: minmod(q, qph)
// = med3(q, qph, 0) (the one nearer zero when the signs agree, else 0) and the small-q form
// q * med3(q qph / tiny^2, 0, 1) (v_med3_f32; the same values, a zero may carry q's sign).
template <typename T>
__device__ __forceinline__ T van_albada(T q, T qph, T tiny, T itiny2) {
    const T s = q * qph;
    const T t = s * itiny2;
    if constexpr (sizeof(T) == 4) {
        const T slow = q * __builtin_amdgcn_fmed3f(t, 0.0f, 1.0f);
        const T fast = __builtin_amdgcn_fmed3f(q, qph, 0.0f);
        return fabs(q) >= tiny ? fast : slow;
    } else {
        // the same medians from min and max (FP64 has no three-operand median)
        const T slow = q * fmin(fmax(t, T(0)), T(1));
        const T fast = fmax(fmin(q, qph), fmin(fmax(q, qph), T(0)));
        return fabs(q) >= tiny ? fast : slow;
    }
}

// the SLD face speed sld_coef sqrt(x): FP32 takes the hardware square root (1 ulp), FP64 the
// correctly rounded one
__device__ __forceinline__ float sld_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sld_sqrt(double x) { return sqrt(x); }

// limited face difference delue = qph - (phiR + phiL)/2 of the face between f0 and fm1 (:766)
template <typename T>
__device__ __forceinline__ T sld_delue(T fm2, T fm1, T f0, T fp1, T tiny, T itiny2) {
    const T qmh = fm1 - fm2, qph = f0 - fm1, qp3h = fp1 - f0;
    return qph - T(0.5) * (van_albada(qp3h, qph, tiny, itiny2) + van_albada(qmh, qph, tiny, itiny2));
}

template <typename T>
__global__ __launch_bounds__(256) void k_sld_speed(Consts<T> c, Prof<T> P, SldParams<T> sp,
                                                   const T *__restrict__ st, const T *__restrict__ B,
                                                   T *__restrict__ vcc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    const int ir = (int)(i / ((long)c.nth * c.nphi));
    T v = P.vsp[ir];
    if (sp.dstype == 1) {
        const T w = st[c.N + i], vv = st[2 * c.N + i], u = st[3 * c.N + i];
        v += w * w;
        v += vv * vv;
        v += u * u;
        if (sp.magnetic) {
            const T rho = c.fluct ? P.rbar[ir] + st[i] : st[i];
            const T invrho = T(1) / rho;
            const T k = T(kOneOver4Pi);
            const T bt = B[i], bp = B[c.N + i], br = B[2 * c.N + i];
            v += k * invrho * bt * bt;
            v += k * invrho * bp * bp;
            v += k * invrho * br * br;
        }
    }
    vcc[i] = v;
}

// value of field f at a point of radius index ir: rho and s relative to the background
template <typename T>
__device__ __forceinline__ T sld_val(const Consts<T> &c, const Prof<T> &P, const T *__restrict__ fp, int f, long idx,
                                     int ir) {
    T x = fp[idx];
    if (!c.fluct) {
        if (f == 0) x -= P.rbar[ir];
        else if (f == 4) x -= P.sbar[ir];
    }
    return x;
}

// the ghost just outside a radial wall: b1(:,4) (lo) or b2(:,1) (hi) of diffusion.f90:586-711
template <typename T>
__device__ T sld_ghost_r(const Consts<T> &c, const Prof<T> &P, const SldParams<T> &sp, const T *__restrict__ fp,
                         int f, long col, long plane, bool hi) {
    const int j0 = hi ? c.nr - 1 : 0, dj = hi ? -1 : 1;
    switch (f) {
    case 0: return T(0);
    case 4:
        if (hi ? sp.sflux_hi : sp.sflux_lo) {
            const T t0 = sld_val(c, P, fp, f, col + (long)j0 * plane, j0);
            return hi ? t0 - sp.soff_hi : t0 + sp.soff_lo;
        }
        return T(0);
    case 1: case 2: {
        const T t0 = sld_val(c, P, fp, f, col + (long)j0 * plane, j0);
        const T t1 = sld_val(c, P, fp, f, col + (long)(j0 + dj) * plane, j0 + dj);
        const T t2 = sld_val(c, P, fp, f, col + (long)(j0 + 2 * dj) * plane, j0 + 2 * dj);
        const T t3 = sld_val(c, P, fp, f, col + (long)(j0 + 3 * dj) * plane, j0 + 3 * dj);
        return T(-10.0 / 3.0) * t0 / P.r[j0] + T(6) * t1 / P.r[j0 + dj] - T(2) * t2 / P.r[j0 + 2 * dj] +
               T(1.0 / 3.0) * t3 / P.r[j0 + 3 * dj];
    }
    case 5: case 6:
        if (hi ? sp.top_pc : sp.bot_pc) return T(0);
        return sld_val(c, P, fp, f, col + (long)j0 * plane, j0);
    default: return sld_val(c, P, fp, f, col + (long)j0 * plane, j0);
    }
}

// the ghost just outside a theta wall (diffusion.f90:1008-1120)
template <typename T>
__device__ T sld_ghost_t(const Consts<T> &c, const Prof<T> &P, const SldParams<T> &sp, const T *__restrict__ fp,
                         int f, long row, int ir, bool hi) {
    const int j0 = hi ? c.nth - 1 : 0, dj = hi ? -1 : 1;
    switch (f) {
    case 1: return T(0);
    case 2: {
        const T t0 = fp[row + j0], t1 = fp[row + j0 + dj], t2 = fp[row + j0 + 2 * dj], t3 = fp[row + j0 + 3 * dj];
        return T(-10.0 / 3.0) * t0 / P.sn[j0] + T(6) * t1 / P.sn[j0 + dj] - T(2) * t2 / P.sn[j0 + 2 * dj] +
               T(1.0 / 3.0) * t3 / P.sn[j0 + 3 * dj];
    }
    case 5: return sp.th_normal ? T(0) : fp[row + j0];
    case 6: case 7: return sp.th_pc ? T(0) : fp[row + j0];
    default: return sld_val(c, P, fp, f, row + j0, ir);
    }
}

// phi rows (wavefronts) per workgroup of k_sld_col
constexpr int kSldRows = 8;

// logical block index with each XCD (8 per chip, workgroups dealt round-robin) walking a
// contiguous range of tiles, so that neighbouring tiles share their halo rows in one L2
__device__ __forceinline__ long xcd_block(long b, long nb) {
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Compute_Diffusion's fluxes (Diffusion_R :476-891, Diffusion_Th :893-1263, Diffusion_Phi
// :1265-1495) and the viscous-heating / mass-diffusion corrections (:375-428) of every field.
// One thread per (theta, phi) column: a workgroup is kSldRows wavefronts, 64 consecutive theta
// columns each, on kSldRows consecutive phi rows (the phi neighbours two rows out of the tile
// come from other workgroups: 8 rows keep that overhang to half the tile), and it marches up the r levels [r0, r0 + rseg).  Each thread
// keeps a four-level window of its column of every diffused field in registers and carries the
// upper r face's flux to the next level as that level's lower face, so every r face is
// limited once.  The theta and phi faces of a cell are evaluated from the neighbouring values
// at the level, which neighbouring lanes and wavefronts load as their own columns (cache
// hits).  No LDS and no barriers: a thread whose column lies outside the grid returns at once.
// The tile index comes from the flat block index, XCD-contiguous, so that phi-neighbouring
// rows share one L2.  Register targets: passes of at most three fields at 6 (FP32) / 4 (FP64)
// wavefronts per SIMD fit three / two 8-wavefront workgroups per CU (80 / 128 VGPRs; their few
// spills are in the lower theta wall's ghost branch), the FP32 two-field pass at 8 fits four
// (64 VGPRs; spills outside the main loop); the rho..s pass with the point corrections does not
// fit the next step without spilling in its main loop.
template <typename T, int F0, int NF>
__global__ __launch_bounds__(64 * kSldRows) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (NF >= 5 ? 3 : NF <= 2 ? 8 : 6) : (NF >= 5 ? 2 : 4)))) void k_sld_col(Consts<T> c, Prof<T> P, SldParams<T> sp,
                                                  const T *__restrict__ st, const T *__restrict__ vcc,
                                                  T *__restrict__ D, int ntt, int ntp, int rseg) {
    // fields [F0, F0 + NF): rho..s with the point corrections (which need all five), parts of
    // rho..s without them (the caller applies them: RK4 stage 1's final kernel), and A
    static_assert((F0 == 0 && NF >= 5 && NF <= 8) || (F0 + NF <= 5 && NF <= 3) || (F0 >= 5 && F0 + NF <= 8),
                  "field group");
    constexpr bool CORR = F0 == 0 && NF >= 5;
    // FP32 states hold rho and s as fluctuations already; FP64 states are absolute
    constexpr bool FLUCT = sizeof(T) == 4;
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const int tt = (int)(lb % ntt), tp = (int)((lb / ntt) % ntp), seg = (int)(lb / ((long)ntt * ntp));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nth = c.nth, nphi = c.nphi, nr = c.nr;
    const int it = tt * 64 + lane, ip = tp * kSldRows + wv;
    if (it >= nth || ip >= nphi) return;
    const long N = c.N, plane = (long)nth * nphi;
    const int r0 = seg * rseg, r1 = min(nr, r0 + rseg);
    // the wavefront holds a cell whose theta stencil reaches a ghost: it < 2 (lower wall) or
    // it + 2 >= nth (upper wall; the tile's last column is tt*64 + 63, so the tile before the
    // last holds column nth - 2 when nth = 1 mod 64)
    const bool wall_lo = tt == 0, wall_hi = tt * 64 + 65 >= nth;
    // Loads are raw buffer loads: a scalar resource per field, the level's byte offset in a
    // scalar register and one of nine 32-bit lane byte offsets within the level plane, shared
    // by every field, so that no per-lane address is live per field across the march (a field
    // stays below 4 GiB).  The few stores per level use plain global stores (no resources).
    constexpr unsigned E = sizeof(T);
    const unsigned oc = (unsigned)(ip * nth + it) * E;
    const unsigned orow = (unsigned)(ip * nth);
    const unsigned otm2 = (orow + (unsigned)max(it - 2, 0)) * E, otm1 = (orow + (unsigned)max(it - 1, 0)) * E;
    const unsigned otp1 = (orow + (unsigned)min(it + 1, nth - 1)) * E, otp2 = (orow + (unsigned)min(it + 2, nth - 1)) * E;
    const unsigned opm2 = (unsigned)(((ip + nphi - 2) % nphi) * nth + it) * E;
    const unsigned opm1 = (unsigned)(((ip + nphi - 1) % nphi) * nth + it) * E;
    const unsigned opp1 = (unsigned)(((ip + 1) % nphi) * nth + it) * E;
    const unsigned opp2 = (unsigned)(((ip + 2) % nphi) * nth + it) * E;
    const unsigned pb = (unsigned)plane * E;   // bytes per level
    const __amdgpu_buffer_rsrc_t rv = field_rsrc(vcc);
    // field f (level j, lane offset ofs) and the speeds; the tendency of field f at level j
    auto ldf = [&](int f, int j, unsigned ofs) -> T { return bload<T>(field_rsrc(st + (long)f * N), ofs, (unsigned)j * pb); };
    auto ldv = [&](int j, unsigned ofs) -> T { return bload<T>(rv, ofs, (unsigned)j * pb); };
    auto dl = [&](int f, int j) -> T * { return D + (long)f * N + (long)j * plane; };
    // the background removed from rho (f = 0) and s (f = 4) of an absolute state at level j
    auto bg = [&](int f, int j) -> T {
        if constexpr (FLUCT) return T(0);
        return f == 0 ? P.rbar[j] : f == 4 ? P.sbar[j] : T(0);
    };
    // the column of field f at level j, with the r-wall ghosts below and above the grid
    auto colval = [&](int f, int j) -> T {
        const T *fp = st + (long)f * N;
        if (j < 0) return sld_ghost_r(c, P, sp, fp, f, (long)(ip * nth + it), plane, false);
        if (j >= nr) return sld_ghost_r(c, P, sp, fp, f, (long)(ip * nth + it), plane, true);
        return ldf(f, j, oc) - bg(f, j);
    };
    auto colvcc = [&](int j) -> T { return ldv(min(max(j, 0), nr - 1), oc); };
    // four-level windows w[q][k] = level ir - 1 + k of field F0 + q; frl = the flux of level ir's
    // lower r face.  Every field of the group is evaluated; rho, the only field a configuration
    // can leave undiffused (no mass diffusion), has its tendency zeroed at the store.
    T w[NF][4], frl[NF];
    T vc_cur = colvcc(r0), vc_next = colvcc(r0 + 1);
    {
        const T vlow = colvcc(r0 - 1);
        const T rl = P.rfc[r0];
        const T spd = sp.coef * sld_sqrt(T(0.5) * (vc_cur + vlow));
        // levels r0-2 .. r0+2 clamped into the grid, then the wall ghosts where a level lies
        // outside it (uniform branches: only the first and the last segment take them)
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const int f = F0 + q;
            const int jm = max(r0 - 2, 0);
            T wm2 = ldf(f, jm, oc) - bg(f, jm);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int j = min(max(r0 - 1 + k, 0), nr - 1);
                w[q][k] = ldf(f, j, oc) - bg(f, j);
            }
            if (r0 < 2) {
                const T g = colval(f, -1);
                wm2 = g;
                if (r0 == 0) w[q][0] = g;
            }
            if (r0 + 2 >= nr) {
                const T g = colval(f, nr);
                if (r0 + 1 >= nr) w[q][2] = g;
                w[q][3] = g;
            }
            frl[q] = spd * (sld_delue(wm2, w[q][0], w[q][1], w[q][2], sp.tiny[f], sp.itiny2[f]) * rl * rl);
        }
    }
    const T wtl = P.sph[it], wth = P.sph[it + 1];
    T gtop[NF];   // the top wall's ghosts, set before the tail levels
    // one r level; TAIL: within three levels of the top wall, where the next window level may
    // be the wall's ghost (the levels below keep that branch out of the main loop)
    auto level = [&](int ir, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        // next level of every window (consumed at the end of the step); past the top wall the
        // ghost (levels beyond nr meet only the wall face, whose weight is zero)
        const bool past = TAIL && ir + 3 >= nr;
        const int jn = min(ir + 3, nr - 1);
        const T bgn0 = bg(0, jn), bgn4 = bg(4, jn), bg0 = bg(0, ir), bg4 = bg(4, ir);
        T nx[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const int f = F0 + q;
            const T x = ldf(f, jn, oc) - (f == 0 ? bgn0 : f == 4 ? bgn4 : T(0));
            nx[q] = past ? gtop[q] : x;
        }
        const T vc_nx = ldv(min(ir + 2, nr - 1), oc);
        // face speeds sld_coef sqrt(mean of the two cells' speed^2); ghost cells copy the nearest
        // cell (diffusion.f90:290-302)
        const T vtm = ldv(ir, otm1), vtp = ldv(ir, otp1), vpm = ldv(ir, opm1), vpp = ldv(ir, opp1);
        const T s_tl = sp.coef * sld_sqrt(T(0.5) * (vc_cur + vtm)), s_th = sp.coef * sld_sqrt(T(0.5) * (vtp + vc_cur));
        const T s_pl = sp.coef * sld_sqrt(T(0.5) * (vpm + vc_cur)), s_ph = sp.coef * sld_sqrt(T(0.5) * (vc_cur + vpp));
        const T spd = sp.coef * sld_sqrt(T(0.5) * (vc_next + vc_cur));
        const T rh = P.rfc[ir + 1];
        const T gr = P.rv[ir], gtt = P.rfac[ir] * P.ct[it], gp = P.gt[it] * (P.rfac[ir] * sp.invdphi);
        T out[5];
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const int f = F0 + q;
            if (f < sp.first) {   // rho without mass diffusion
                if constexpr (CORR) out[0] = T(0);
                else at_row(dl(0, ir), oc) = T(0);
                continue;
            }
            const T b = f == 0 ? bg0 : f == 4 ? bg4 : T(0);
            const T tiny = sp.tiny[f], it2 = sp.itiny2[f];
            // theta neighbours at this level; past a wall, that wall's ghost
            T a2 = ldf(f, ir, otm2) - b, a1 = ldf(f, ir, otm1) - b, b1 = ldf(f, ir, otp1) - b, b2 = ldf(f, ir, otp2) - b;
            if (wall_lo) {
                const T g = sld_ghost_t(c, P, sp, st + (long)f * N, f, (long)ir * plane + orow, ir, false);
                a2 = it < 2 ? g : a2;
                a1 = it < 1 ? g : a1;
            }
            if (wall_hi) {
                const T g = sld_ghost_t(c, P, sp, st + (long)f * N, f, (long)ir * plane + orow, ir, true);
                b1 = it + 1 >= nth ? g : b1;
                b2 = it + 2 >= nth ? g : b2;
            }
            const T pm2 = ldf(f, ir, opm2) - b, pm1 = ldf(f, ir, opm1) - b;
            const T pp1 = ldf(f, ir, opp1) - b, pp2 = ldf(f, ir, opp2) - b;
            const T f0 = w[q][1];
            const T ft_lo = (s_tl * wtl) * sld_delue(a2, a1, f0, b1, tiny, it2);
            const T ft_hi = (s_th * wth) * sld_delue(a1, f0, b1, b2, tiny, it2);
            const T fp_lo = s_pl * sld_delue(pm2, pm1, f0, pp1, tiny, it2);
            const T fp_hi = s_ph * sld_delue(pm1, f0, pp1, pp2, tiny, it2);
            const T fh = spd * (sld_delue(w[q][0], w[q][1], w[q][2], w[q][3], tiny, it2) * rh * rh);
            const T Dr = sp.mr[f] * (gr * (fh - frl[q]));
            frl[q] = fh;
            const T Dt = sp.mt[f] * (gtt * (ft_hi - ft_lo));
            const T Dp = gp * (sp.mt[f] * (fp_hi - fp_lo));
            const T o = (Dr + Dt) + Dp;
            if (CORR && f < 5) out[f] = o;
            else if (f < 5 || c.magnetic) at_row(dl(f, ir), oc) = o;       // no correction here
        }
        if constexpr (CORR) {
            // viscous heating and mass diffusion (:375-428), at the point
            if (sp.visc || sp.mass) {
                // the stored slots (the windows hold rho and s with the FP64 background removed)
                const Thermo<T> th = thermo(ldf(0, ir, oc), ldf(4, ir, oc), ir, c, P);
                const T invrho = T(1) / th.rho, itemp = T(1) / th.t;
                const T u1 = w[1][1], u2 = w[2][1], u3 = w[3][1];
                if (sp.visc) {
                    out[4] = out[4] - u1 * out[1] * itemp;
                    out[4] = out[4] - u2 * out[2] * itemp;
                    out[4] = out[4] - u3 * out[3] * itemp;
                }
                if (sp.mass) {
                    const T drho = out[0];
                    out[1] = out[1] - invrho * u1 * drho;
                    out[2] = out[2] - invrho * u2 * drho;
                    out[3] = out[3] - invrho * u3 * drho;
                    out[4] = out[4] - sp.gcv * invrho * drho;
                }
            }
#pragma unroll
            for (int f = 0; f < 5; f++) at_row(dl(f, ir), oc) = out[f];
        }
        // shift the windows
#pragma unroll
        for (int q = 0; q < NF; q++) {
#pragma unroll
            for (int k = 0; k < 3; k++) w[q][k] = w[q][k + 1];
            w[q][3] = nx[q];
        }
        vc_cur = vc_next;
        vc_next = vc_nx;
    };
    const int rm = max(r0, min(r1, nr - 3));
    for (int ir = r0; ir < rm; ir++) level(ir, std::false_type{});
    if (rm < r1) {
#pragma unroll
        for (int q = 0; q < NF; q++) gtop[q] = colval(F0 + q, nr);
        for (int ir = rm; ir < r1; ir++) level(ir, std::true_type{});
    }
}

// Ohmic heating (diffusion.f90:430-461) at one point: ds -= B . curl(dA/dt) / (8 pi rho T),
// C = curl of the A tendency (3 fields)
template <typename T>
__global__ __launch_bounds__(256) void k_sld_finish(Consts<T> c, Prof<T> P, const T *__restrict__ st,
                                                    const T *__restrict__ B, const T *__restrict__ C,
                                                    T *__restrict__ D) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    const long N = c.N;
    const int ir = (int)(i / ((long)c.nth * c.nphi));
    const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
    const T invrho = T(1) / th.rho, itemp = T(1) / th.t;
    T ds = D[4 * N + i];
    const T t = T(0.5 * kOneOver4Pi) * invrho * itemp;
    ds = ds - B[i] * C[i] * t;
    ds = ds - B[N + i] * C[N + i] * t;
    ds = ds - B[2 * N + i] * C[2 * N + i] * t;
    D[4 * N + i] = ds;
}

}  // namespace cssx
