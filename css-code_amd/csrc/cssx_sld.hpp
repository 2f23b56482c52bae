// SLD diffusion of Compute_Diffusion(1) (diffusion.f90:70-474, Diffusion_R/Th/Phi :476-1495)
// on gfx950.  Included by cssx_rhs.hip after the state/profile types.
//
// Two launches plus, with Ohmic heating, the curl of the A tendency and a third:
//   k_sld_speed   vcc = speed^2(r) [+ |u|^2 + B^2/(4 pi rho) for dstype 1]   (:177-302)
//   k_sld_march   every field's tendency: a workgroup owns a tile of 64 theta x PH phi columns
//                 and marches up r.  Each thread keeps a five-level window of its column of
//                 every diffused field in registers and carries the upper r face's flux to the
//                 next level as that level's lower face; the current level of the tile (with
//                 two cells of halo) sits in LDS, where each cell's lower theta and phi face
//                 fluxes are left for the neighbouring cells to take as their upper faces, so
//                 every face of the grid is limited once.  Then the viscous heating and mass
//                 diffusion corrections (:375-428), which need only the point itself.
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
// closed form q min(1, q qph / tiny^2) is kept.
template <typename T>
__device__ __forceinline__ T van_albada(T q, T qph, T tiny, T itiny2) {
    const T s = q * qph;
    const T a = fabs(q);
    const T r = a >= tiny ? (fabs(qph) < a ? qph : q) : q * fmin(T(1), s * itiny2);
    return s > T(0) ? r : T(0);
}

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

// logical block index with each XCD (8 per chip, workgroups dealt round-robin) walking a
// contiguous range of tiles, so that neighbouring tiles share their halo rows in one L2
__device__ __forceinline__ long xcd_block(long b, long nb) {
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// value at (it, row) of field f at one r level for the theta stencils: inside the grid the
// stored value (background removed), one or two cells past a theta wall that wall's ghost, and
// zero further out (beyond a tile that overhangs the grid; finite, never weighted)
template <typename T>
__device__ T sld_tile_value(const Consts<T> &c, const Prof<T> &P, const SldParams<T> &sp, const T *__restrict__ fp,
                            int f, long row, int it, int ir) {
    if (it >= 0 && it < c.nth) return sld_val(c, P, fp, f, row + it, ir);
    if (it >= -2 && it < 0) return sld_ghost_t(c, P, sp, fp, f, row, ir, false);
    if (it >= c.nth && it <= c.nth + 1) return sld_ghost_t(c, P, sp, fp, f, row, ir, true);
    return T(0);
}

// speed^2 of (it, row) with the ghost cells copying the nearest cell (diffusion.f90:290-302)
template <typename T>
__device__ __forceinline__ T sld_tile_vcc(const Consts<T> &c, const T *__restrict__ vcc, long row, int it) {
    return vcc[row + min(max(it, 0), c.nth - 1)];
}

// Compute_Diffusion's fluxes (Diffusion_R :476-891, Diffusion_Th :893-1263, Diffusion_Phi
// :1265-1495) and the viscous-heating / mass-diffusion corrections (:375-428) of every field.
// Workgroup = 64 theta lanes x PH phi rows (one wavefront per row), one tile, r levels
// [r0, r0 + rseg).  The tile index comes from the flat block index, XCD-contiguous.
template <typename T, int PH>
__global__ __launch_bounds__(64 * PH) void k_sld_march(Consts<T> c, Prof<T> P, SldParams<T> sp,
                                                        const T *__restrict__ st, const T *__restrict__ vcc,
                                                        T *__restrict__ D, int ntt, int ntp, int rseg) {
    constexpr int NF = 8, TW = 68, TR = PH + 4;
    __shared__ T sv[NF + 1][TR][TW];          // level values (fields, then vcc) with 2-cell halo
    __shared__ T fth[NF][PH][65];             // lower theta face fluxes (65th: the tile's upper edge)
    __shared__ T fph[NF][PH + 1][64];         // lower phi face fluxes (row PH: the tile's upper edge)
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const int tt = (int)(lb % ntt), tp = (int)((lb / ntt) % ntp), seg = (int)(lb / ((long)ntt * ntp));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nth = c.nth, nphi = c.nphi, nr = c.nr;
    const long N = c.N, plane = (long)nth * nphi;
    const int th0 = tt * 64, ph0 = tp * PH;
    const int r0 = seg * rseg, r1 = min(nr, r0 + rseg);
    const int it = th0 + lane, ip = (ph0 + wv) % nphi;
    const bool own = it < nth;                 // this lane's column is a grid column
    const long col = (long)ip * nth + (own ? it : 0);
    auto active = [&](int f) { return f >= sp.first && f <= sp.last; };
    // the column of field f at level j, with the r-wall ghosts below and above the grid
    auto colval = [&](int f, int j) -> T {
        const T *fp = st + (long)f * N;
        if (j < 0) return sld_ghost_r(c, P, sp, fp, f, col, plane, false);
        if (j >= nr) return sld_ghost_r(c, P, sp, fp, f, col, plane, true);
        return sld_val(c, P, fp, f, col + (long)j * plane, j);
    };
    auto colvcc = [&](int j) -> T { return vcc[col + (long)min(max(j, 0), nr - 1) * plane]; };
    // four-level windows w[f][k] = level ir - 1 + k; frl = the flux of level ir's lower r face
    T w[NF][4], frl[NF];
    T vc_cur = T(0), vc_next = T(0);
    if (own) {
        vc_cur = colvcc(r0);
        vc_next = colvcc(r0 + 1);
        const T vlow = colvcc(r0 - 1);
        const T rl = P.rfc[r0];
        const T spd = sp.coef * sqrt(T(0.5) * (vc_cur + vlow));
#pragma unroll
        for (int f = 0; f < NF; f++) {
            frl[f] = T(0);
            if (!active(f)) continue;
            const T wm2 = colval(f, r0 - 2);
#pragma unroll
            for (int k = 0; k < 4; k++) w[f][k] = colval(f, r0 - 1 + k);
            frl[f] = spd * (sld_delue(wm2, w[f][0], w[f][1], w[f][2], sp.tiny[f], sp.itiny2[f]) * rl * rl);
        }
    }
    const int hc_rows = 4 * TW, hc_cols = 4 * PH, nhalo = hc_rows + hc_cols;
    for (int ir = r0; ir < r1; ir++) {
        // next level of every window (consumed at the end of the step)
        T nx[NF];
        T vc_nx = T(0);
        if (own) {
#pragma unroll
            for (int f = 0; f < NF; f++) nx[f] = active(f) ? colval(f, ir + 3) : T(0);
            vc_nx = colvcc(ir + 2);
        }
        // ---- the tile's level ir in LDS: own cells, then the halo ----
#pragma unroll
        for (int f = 0; f < NF; f++)
            if (active(f)) sv[f][wv + 2][lane + 2] = own ? w[f][1] : T(0);
        sv[NF][wv + 2][lane + 2] = own ? vc_cur : T(0);
        if (!own) {
            // a tile overhanging the grid: the cells just past the upper theta wall are ghosts
            const long row = (long)ir * plane + (long)ip * nth;
#pragma unroll
            for (int f = 0; f < NF; f++)
                if (active(f)) sv[f][wv + 2][lane + 2] = sld_tile_value(c, P, sp, st + (long)f * N, f, row, it, ir);
            sv[NF][wv + 2][lane + 2] = sld_tile_vcc(c, vcc, row, it);
        }
        for (int h = threadIdx.x; h < nhalo; h += 64 * PH) {
            int hr, hcol;
            if (h < hc_rows) {                      // rows -2, -1, PH, PH+1, all 68 columns
                const int q = h / TW;
                hr = q < 2 ? q : PH + q;
                hcol = h % TW;
            } else {                                // columns -2, -1, 64, 65 of rows 0..PH-1
                const int q = h - hc_rows;
                hr = 2 + q / 4;
                const int k = q % 4;
                hcol = k < 2 ? k : 64 + k;
            }
            const int hit = th0 + hcol - 2;
            const int hip = ((ph0 + hr - 2) % nphi + nphi) % nphi;
            const long row = (long)ir * plane + (long)hip * nth;
#pragma unroll
            for (int f = 0; f < NF; f++)
                if (active(f)) sv[f][hr][hcol] = sld_tile_value(c, P, sp, st + (long)f * N, f, row, hit, ir);
            sv[NF][hr][hcol] = sld_tile_vcc(c, vcc, row, hit);
        }
        __syncthreads();
        // ---- lower theta and phi face fluxes of every cell; the tile's upper edges ----
        {
            const int tc = lane + 2, tr = wv + 2;
            const T vt = sp.coef * sqrt(T(0.5) * (sv[NF][tr][tc] + sv[NF][tr][tc - 1]));
            const T vp = sp.coef * sqrt(T(0.5) * (sv[NF][tr - 1][tc] + sv[NF][tr][tc]));
            const T wt = it <= nth ? P.sph[it] : T(0);
#pragma unroll
            for (int f = 0; f < NF; f++) {
                if (!active(f)) continue;
                const T tiny = sp.tiny[f], it2 = sp.itiny2[f];
                fth[f][wv][lane] = (vt * wt) * sld_delue(sv[f][tr][tc - 2], sv[f][tr][tc - 1], sv[f][tr][tc],
                                                         sv[f][tr][tc + 1], tiny, it2);
                fph[f][wv][lane] = vp * sld_delue(sv[f][tr - 2][tc], sv[f][tr - 1][tc], sv[f][tr][tc],
                                                  sv[f][tr + 1][tc], tiny, it2);
            }
            if (threadIdx.x < 64) {                 // phi faces of row PH (the next tile's row 0)
                const int r = PH + 2;
                const T v2 = sp.coef * sqrt(T(0.5) * (sv[NF][r - 1][tc] + sv[NF][r][tc]));
#pragma unroll
                for (int f = 0; f < NF; f++)
                    if (active(f))
                        fph[f][PH][lane] = v2 * sld_delue(sv[f][r - 2][tc], sv[f][r - 1][tc], sv[f][r][tc],
                                                          sv[f][r + 1][tc], sp.tiny[f], sp.itiny2[f]);
            } else if (threadIdx.x < 64 + PH) {     // theta faces of column 64 (the next tile's 0)
                const int rr = threadIdx.x - 64 + 2, cc = 66, ite = th0 + 64;
                const T v2 = sp.coef * sqrt(T(0.5) * (sv[NF][rr][cc] + sv[NF][rr][cc - 1]));
                const T w2 = ite <= nth ? P.sph[ite] : T(0);
#pragma unroll
                for (int f = 0; f < NF; f++)
                    if (active(f))
                        fth[f][rr - 2][64] = (v2 * w2) * sld_delue(sv[f][rr][cc - 2], sv[f][rr][cc - 1], sv[f][rr][cc],
                                                                   sv[f][rr][cc + 1], sp.tiny[f], sp.itiny2[f]);
            }
        }
        __syncthreads();
        // ---- the tendency of this level ----
        if (own) {
            const long i = (long)ir * plane + col;
            const T rh = P.rfc[ir + 1];
            const T spd = sp.coef * sqrt(T(0.5) * (vc_next + vc_cur));
            const T gr = P.rv[ir], gtt = P.rfac[ir] * P.ct[it], gp = P.gt[it] * (P.rfac[ir] * sp.invdphi);
            T out[5];
#pragma unroll
            for (int f = 0; f < NF; f++) {
                T o = T(0);
                if (active(f)) {
                    const T fh = spd * (sld_delue(w[f][0], w[f][1], w[f][2], w[f][3], sp.tiny[f], sp.itiny2[f]) * rh * rh);
                    const T Dr = sp.mr[f] * (gr * (fh - frl[f]));
                    frl[f] = fh;
                    const T Dt = sp.mt[f] * (gtt * (fth[f][wv][lane + 1] - fth[f][wv][lane]));
                    const T Dp = gp * (sp.mt[f] * (fph[f][wv + 1][lane] - fph[f][wv][lane]));
                    o = (Dr + Dt) + Dp;
                }
                if (f < 5) out[f] = o;
                else if (c.magnetic) D[(long)f * N + i] = o;       // A: no correction
            }
            // viscous heating and mass diffusion (:375-428), at the point
            if (sp.visc || sp.mass) {
                // the stored slots (the windows hold rho and s with the FP64 background removed)
                const Thermo<T> th = thermo(st[i], st[4 * N + i], ir, c, P);
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
            for (int f = 0; f < 5; f++) D[(long)f * N + i] = out[f];
            // shift the windows
#pragma unroll
            for (int f = 0; f < NF; f++) {
                if (!active(f)) continue;
#pragma unroll
                for (int k = 0; k < 3; k++) w[f][k] = w[f][k + 1];
                w[f][3] = nx[f];
            }
            vc_cur = vc_next;
            vc_next = vc_nx;
        }
        __syncthreads();
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
