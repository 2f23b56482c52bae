// SLD diffusion of Compute_Diffusion(1) (diffusion.f90:70-474, Diffusion_R/Th/Phi :476-1495)
// on gfx950.  Included by cssx_rhs.hip after the state/profile types.
//
// Three launches plus the Ohmic curl:
//   k_sld_speed   vcc = speed^2(r) [+ |u|^2 + B^2/(4 pi rho) for dstype 1]   (:177-302)
//   k_sld_faces   the face speeds sld_coef*sqrt(mean vcc) of every cell's lower faces
//   k_sld_flux    one field per blockIdx.y: the six face fluxes of every cell in the three
//                 directions, Van Albada limited, summed in the reference's order r, theta, phi
//   (curl of the A tendency, the same batched derivative sweeps as Compute_B without its wall
//    zeroing, VectorCurlA of :437)
//   k_sld_finish  viscous heating, mass diffusion and Ohmic heating corrections (:375-461)
// The background meanV removed from rho and s (:168-175) is the hydrostatic profile
// (update_background = .False.); FP32 states already store exactly those fluctuations.
// Ghost values at the four walls follow the per-variable rules of Diffusion_R (:586-711) and
// Diffusion_Th (:1008-1120); the ghost-cell speeds are copies of the nearest cell
// (:290-302) and phi is periodic.  Face values enter only through the face-flux weights rfc
// and sph, which are zero on the walls.
#pragma once

namespace cssx {

template <typename T>
struct SldParams {
    T coef;                 // sld_coef
    T tiny[8];              // tinyvar by field (diffusion.f90:76-80)
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

// 1/x: IEEE division in FP64 (the reference's arithmetic), the hardware reciprocal (1 ulp) in
// the FP32 path
template <typename T>
__device__ __forceinline__ T recip(T x) { return T(1) / x; }
template <>
__device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Van Albada slope of diffusion.f90:744-762, operation for operation
template <typename T>
__device__ __forceinline__ T van_albada(T q, T qph, T tiny) {
    T rt = fabs(q);
    rt = recip(fmax(rt, tiny));
    rt = q * rt * rt;
    rt = rt * qph;
    rt = fmin(T(1), rt);
    return q * fmax(T(0), rt);
}

// limited face difference delue = qph - (phiR + phiL)/2 of the face between f0 and fm1 (:766)
template <typename T>
__device__ __forceinline__ T sld_delue(T fm2, T fm1, T f0, T fp1, T tiny) {
    const T qmh = fm1 - fm2, qph = f0 - fm1, qp3h = fp1 - f0;
    return qph - T(0.5) * (van_albada(qp3h, qph, tiny) + van_albada(qmh, qph, tiny));
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

// face speeds sld_coef*sqrt((vcc_a + vcc_b)/2) of the lower r, theta and phi face of every
// cell (diffusion.f90:836, :1246, :1466); ghost cells copy the nearest cell (:290-302), phi
// wraps.  The wall faces get a finite placeholder: their flux weight is zero.
template <typename T>
__global__ __launch_bounds__(256) void k_sld_faces(Consts<T> c, SldParams<T> sp, const T *__restrict__ vcc,
                                                   T *__restrict__ vt) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    int it, ip, ir;
    coords(i, c, it, ip, ir);
    const long plane = (long)c.nth * c.nphi;
    const T h = T(0.5), vc = vcc[i];
    const T vrm = vcc[ir > 0 ? i - plane : i];
    const T vtm = vcc[it > 0 ? i - 1 : i];
    const int ipm = ip > 0 ? ip - 1 : c.nphi - 1;
    const T vpm = vcc[i + (long)(ipm - ip) * c.nth];
    vt[i] = sp.coef * sqrt(h * (vc + vrm));
    vt[c.N + i] = sp.coef * sqrt(h * (vc + vtm));
    vt[2 * c.N + i] = sp.coef * sqrt(h * (vpm + vc));
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

// logical block index with each XCD (8 per chip, workgroups dealt round-robin) walking a
// contiguous range of rows, so the phi and r neighbours of a row are mostly in its own L2
__device__ __forceinline__ long xcd_block(long b, long nb) {
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sld_flux(Consts<T> c, Prof<T> P, SldParams<T> sp,
                                                  const T *__restrict__ st, const T *__restrict__ vt,
                                                  T *__restrict__ D) {
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const long i = lb * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= c.N) return;
    T *out = D + (long)f * c.N;
    if (f < sp.first || f > sp.last) {
        out[i] = T(0);
        return;
    }
    int it, ip, ir;
    coords(i, c, it, ip, ir);
    const int nth = c.nth, nphi = c.nphi, nr = c.nr;
    const long plane = (long)nth * nphi;
    const long col = i - (long)ir * plane;                 // (it, ip) column of the r pencil
    const long row = i - it;                               // (ip, ir) row of the theta pencil
    const T *fp = st + (long)f * c.N;
    const T tiny = sp.tiny[f];
    // ---- face speeds: lower face of this cell, lower face of the next cell ----
    const long N = c.N;
    const int ipp = ip < nphi - 1 ? ip + 1 : 0;
    const T sr_l = vt[i], sr_h = vt[ir < nr - 1 ? i + plane : i];
    const T st_l = vt[N + i], st_h = vt[N + (it < nth - 1 ? i + 1 : i)];
    const T sp_l = vt[2 * N + i], sp_h = vt[2 * N + i + (long)(ipp - ip) * nth];
    // ---- r (Diffusion_R :745-858) ----
    T Dr;
    {
        T w[5];
#pragma unroll
        for (int d = -2; d <= 2; d++) {
            const int j = min(max(ir + d, 0), nr - 1);
            w[d + 2] = sld_val(c, P, fp, f, col + (long)j * plane, j);
        }
        if (ir <= 1) {
            const T g = sld_ghost_r(c, P, sp, fp, f, col, plane, false);
            w[1 - ir] = g;
            if (ir == 0) w[0] = g;
        }
        if (ir >= nr - 2) {
            const T g = sld_ghost_r(c, P, sp, fp, f, col, plane, true);
            w[nr + 2 - ir] = g;     // position of index nr
            if (ir == nr - 1) w[4] = g;
        }
        const T rl = P.rfc[ir], rh = P.rfc[ir + 1];
        const T Fl = sr_l * (sld_delue(w[0], w[1], w[2], w[3], tiny) * rl * rl);
        const T Fh = sr_h * (sld_delue(w[1], w[2], w[3], w[4], tiny) * rh * rh);
        Dr = sp.mr[f] * (P.rv[ir] * (Fh - Fl));
    }
    // ---- theta (Diffusion_Th :1122-1262) ----
    T Dt;
    {
        T w[5];
#pragma unroll
        for (int d = -2; d <= 2; d++) {
            const int j = min(max(it + d, 0), nth - 1);
            w[d + 2] = sld_val(c, P, fp, f, row + j, ir);
        }
        if (it <= 1) {
            const T g = sld_ghost_t(c, P, sp, fp, f, row, ir, false);
            w[1 - it] = g;
            if (it == 0) w[0] = g;
        }
        if (it >= nth - 2) {
            const T g = sld_ghost_t(c, P, sp, fp, f, row, ir, true);
            w[nth + 2 - it] = g;
            if (it == nth - 1) w[4] = g;
        }
        const T Fl = (st_l * P.sph[it]) * sld_delue(w[0], w[1], w[2], w[3], tiny);
        const T Fh = (st_h * P.sph[it + 1]) * sld_delue(w[1], w[2], w[3], w[4], tiny);
        Dt = sp.mt[f] * ((P.rfac[ir] * P.ct[it]) * (Fh - Fl));
    }
    // ---- phi, periodic (Diffusion_Phi :1362-1494) ----
    T Dp;
    {
        T w[5];
#pragma unroll
        for (int d = -2; d <= 2; d++) {
            int j = ip + d;
            j = j < 0 ? j + nphi : (j >= nphi ? j - nphi : j);
            w[d + 2] = sld_val(c, P, fp, f, i + (long)(j - ip) * nth, ir);
        }
        const T Fl = sp_l * sld_delue(w[0], w[1], w[2], w[3], tiny);
        const T Fh = sp_h * sld_delue(w[1], w[2], w[3], w[4], tiny);
        const T geom = P.gt[it] * (P.rfac[ir] * sp.invdphi);
        Dp = geom * (sp.mt[f] * (Fh - Fl));
    }
    out[i] = (Dr + Dt) + Dp;
}

// viscous heating, mass diffusion, Ohmic heating (diffusion.f90:375-461) at one point.
// C = curl of the A tendency (3 fields) or null.
template <typename T>
__global__ __launch_bounds__(256) void k_sld_finish(Consts<T> c, Prof<T> P, SldParams<T> sp,
                                                    const T *__restrict__ st, const T *__restrict__ B,
                                                    const T *__restrict__ C, T *__restrict__ D) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    const long N = c.N;
    const int ir = (int)(i / ((long)c.nth * c.nphi));
    const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
    const T invrho = T(1) / th.rho, itemp = T(1) / th.t;
    const T u1 = st[N + i], u2 = st[2 * N + i], u3 = st[3 * N + i];
    T d1 = D[N + i], d2 = D[2 * N + i], d3 = D[3 * N + i], ds = D[4 * N + i];
    if (sp.visc) {
        ds = ds - u1 * d1 * itemp;
        ds = ds - u2 * d2 * itemp;
        ds = ds - u3 * d3 * itemp;
    }
    if (sp.mass) {
        const T drho = D[i];
        d1 = d1 - invrho * u1 * drho;
        d2 = d2 - invrho * u2 * drho;
        d3 = d3 - invrho * u3 * drho;
        ds = ds - sp.gcv * invrho * drho;
        D[N + i] = d1;
        D[2 * N + i] = d2;
        D[3 * N + i] = d3;
    }
    if (C) {
        const T t = T(0.5 * kOneOver4Pi) * invrho * itemp;
        ds = ds - B[i] * C[i] * t;
        ds = ds - B[N + i] * C[N + i] * t;
        ds = ds - B[2 * N + i] * C[2 * N + i] * t;
    }
    D[4 * N + i] = ds;
}

}  // namespace cssx
