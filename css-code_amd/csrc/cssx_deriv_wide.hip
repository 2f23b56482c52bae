// Compact derivative along a long strided axis (r or phi pencils of more than 16 segments'
// worth of rows) for gfx950: the operator of derivatives.f90 Dbydr/Dbydp (:547-796,
// :1480-1591) and D2bydr2/D2bydp2 (:799-1007, :1594-1684), as k_deriv computes it.
//
// k_deriv gives each wavefront one segment of 64 pencils, so a workgroup of 16 wavefronts cuts
// a pencil into at most 16 segments; a 1024-row pencil then needs 64-row segments, whose rows
// no longer fit in the 128 registers a lane has at that occupancy.  Here a wavefront holds G
// segments of P = 64/G pencils instead (lane = g*P + p), so a workgroup covers P pencils cut
// into 16*G segments of M rows:
//   * every lane loads its segment's rows plus 4 on each side with the row stride; the P lanes
//     of a group read P consecutive values of one row (P*sizeof(T) = 128 or 256 bytes), so the
//     loads stay whole-cache-line coalesced;
//   * the right-hand side rows and the segment-local LU solve are those of k_deriv, with the
//     plan coefficients read from LDS (one padding slot per segment keeps the G segments of a
//     wavefront on different banks);
//   * the 16*G-1 interface rows of a pencil are solved by parallel cyclic reduction over the
//     lanes of one wavefront (thread t takes interface row t mod 16G of pencil t / 16G), with
//     data-independent multipliers from the plan (the contiguous kernel's plan, cut at M);
//   * every lane corrects its rows with the two interface values around its segment and
//     stores them.
// One read and one write of every point reach HBM; the 8 halo rows per segment are re-read
// from L2.
#include <hip/hip_runtime.h>

#include "cssx_bufio.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"
#include "cssx_rows.hpp"

namespace cssx {

// capacity of each interface exchange array: 1024 slots plus one padding slot per row
constexpr int kWideShared = 1024 + 64;

// FP32: at most 64 VGPRs (8 waves per SIMD), so that two 16-wave workgroups share a CU and one
// workgroup's loads overlap the other's solve and barriers; at 93 VGPRs a CU held one
// workgroup and the sweep ran 17% slower.
template <typename T, int ORDER, int M, int G>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 1)))
void k_deriv_wide(const DerivLaunch<T> L) {
    constexpr int P = 64 / G;    // pencils per workgroup
    constexpr int SM = 16 * G;   // segment slots per pencil
    constexpr int W = M + 8;     // window w[i] = row st-4+i
    extern __shared__ unsigned char smem_raw[];
    T *sp = reinterpret_cast<T *>(smem_raw);
    const DerivJob<T> &J = L.jobs[blockIdx.z];
    const int n = L.n;
    const int nl = (n + M - 1) / M;   // segments in use (<= SM)
    const int nI = nl - 1;
    const int pn = n + nl;            // padded row count: row r at r + r/M
    const int tid = threadIdx.x;
    {
        // the five row tables (5n <= 5120 values): every thread's loads issued before the first
        // wait
        T pv[5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int i = tid + k * 1024;
            pv[k] = i < 5 * n ? J.plan[i] : T(0);
        }
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int i = tid + k * 1024;
            if (i < 5 * n) {
                const int c = i / n, r = i - c * n;
                sp[c * pn + r + r / M] = pv[k];
            }
        }
    }
    // exchange arrays, one padding slot per row so that the lanes of a wavefront, which walk
    // the other index, hit different banks: [segment][pencil] with stride PS, [pencil][segment]
    // with stride XS
    constexpr int PS = P + 1, XS = SM + 1, SHN = kWideShared;
    static_assert(SM * PS <= SHN && P * XS <= SHN, "exchange arrays");
    T *sh_xf = sp + 5 * pn, *sh_xl = sh_xf + SHN, *sh_dI = sh_xl + SHN, *sh_xI = sh_dI + SHN;
    const T *cf = sp, *cl = cf + pn, *cu = cl + pn, *vL = cu + pn, *vR = vL + pn;

    const int lane = tid & 63;
    const int g = lane / P, p = lane & (P - 1);
    const int s = (tid >> 6) * G + g;          // this lane's segment (== tid / P)
    const long a = (long)blockIdx.x * P + p;
    const int b = blockIdx.y;
    const bool pen_ok = a < L.na;
    const bool seg_ok = s < nl;
    const bool active = pen_ok && seg_ok;
    const int st = s * M;
    const int m = seg_ok ? min(M, n - st) : 0;
    const int lo_kind = J.lo_kind, hi_kind = J.hi_kind;
    const long gbase = a * J.g_a + (long)b * J.g_b;
    const long bofs = (long)b * L.sb;
    const unsigned lofs = (unsigned)((pen_ok ? a : 0) * L.sa) * (unsigned)sizeof(T);
    const unsigned rowb = (unsigned)L.srow * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rin = field_rsrc(J.in + bofs);
    const bool periodic = lo_kind == END_PERIODIC;

    // ---- 1. window: own rows and 4 on each side (wrapped on a periodic pencil, clamped
    //         otherwise; rows outside a non-periodic pencil are replaced below) ---------------
    T w[W];
    // FP64, wave-uniform: every window row of the wavefront's G segments lies inside the pencil
    // (all wavefronts but the first and the last ones), so the rows are a plain strided run with
    // no per-row wrap or clamp.  (In FP32 this path and the store path below measured slower:
    // 0.30 against 0.28 ms per 1024-row field, with 25% fewer VALU instructions.)
    constexpr bool kFast = sizeof(T) == 8;
    const int ws = (tid >> 6) * G;                  // the wavefront's first segment
    if (kFast && ws * M >= 4 && (ws + G) * M + 4 <= n) {
        const unsigned vb = lofs + (unsigned)(st - 4) * rowb;
#pragma unroll
        for (int i = 0; i < W; i++) {
            w[i] = bload<T>(rin, vb + (unsigned)i * rowb, 0u);
            if (J.pre) w[i] *= J.pre[st - 4 + i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < W; i++) {
            int q = st - 4 + i;
            if (periodic) q = q < 0 ? q + n : (q >= n ? q - n : q);
            else q = q < 0 ? 0 : (q >= n ? n - 1 : q);
            w[i] = bload<T>(rin, lofs + (unsigned)q * rowb, 0u);
            if (J.pre) w[i] *= J.pre[q];
        }
    }
    // only the end segments' windows leave the pencil
    if (!periodic && (st < 4 || st + M + 4 > n)) {
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int q = st - 4 + i;
            if (q < 0) w[i] = (lo_kind == END_HALO && pen_ok) ? J.glo[gbase + (long)(q + 4) * J.g_k] : T(0);
            if (q >= n) w[i] = (hi_kind == END_HALO && pen_ok && q < n + 4) ? J.ghi[gbase + (long)(q - n) * J.g_k] : T(0);
        }
    }
    T lo_val = T(0), hi_val = T(0);
    if (active && s == 0 && lo_kind == END_WALL && J.glo) lo_val = J.glo[gbase];
    if (active && s == nl - 1 && hi_kind == END_WALL && J.ghi) hi_val = J.ghi[gbase];

    // ---- 2. right-hand side rows -----------------------------------------------------------------
    const int bd = J.bdry;
    T d[M];
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        T val;
        if constexpr (ORDER == 1)
            val = Rows<T, 1>::interior(w[jj + 2], w[jj + 3], w[jj + 5], w[jj + 6]);
        else
            val = Rows<T, 2>::interior(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6]);
        const int row = st + jj;
        if (row == 0) {
            if (lo_kind == END_WALL) {
                if constexpr (ORDER == 1) {
                    T r1 = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
                    T vv = (bd == 0) ? T(0) : lo_val * J.neu_mul;
                    val = Rows<T, 1>::wall0(bd == 1 || bd == 3, vv, w[4], w[5], w[6], w[7], r1, J.lo_r0_val != 0);
                } else {
                    T vv = (bd == 0) ? T(0) : -lo_val * J.neu_mul;
                    val = Rows<T, 2>::wall0(bd == 1 || bd == 3, vv, w[4], w[5], w[6], w[7], w[8], w[9]);
                }
            } else {   // halo, or the periodic seam with the wrapped rows as ghosts
                if constexpr (ORDER == 1)
                    val = Rows<T, 1>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6], J.halo_s);
                else
                    val = Rows<T, 2>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
            }
        } else if (row == 1 && lo_kind == END_WALL) {
            if constexpr (ORDER == 1) val = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
            else val = Rows<T, 2>::wall1(w[4], w[5], w[6], w[7], w[8]);
        }
        if (row == n - 1) {
            if (hi_kind == END_WALL) {
                if constexpr (ORDER == 1) {
                    T rn2 = -Rows<T, 1>::wall1(w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj]);
                    T vv = (bd == 3 || bd == 4) ? hi_val * J.neu_mul : T(0);
                    T rn1 = Rows<T, 1>::wall0(bd == 1 || bd == 2, vv, w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], -rn2);
                    val = (bd == 1 || bd == 2) ? -rn1 : rn1;
                } else {
                    T vv = (bd == 0) ? T(0) : hi_val * J.neu_mul;
                    val = Rows<T, 2>::wall0(bd == 1 || bd == 2, vv, w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1],
                                            w[jj], w[jj - 1 < 0 ? 0 : jj - 1]);
                }
            } else {
                if constexpr (ORDER == 1)
                    val = Rows<T, 1>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6], w[jj + 7],
                                               w[jj + 8], J.halo_s);
                else
                    val = Rows<T, 2>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6], w[jj + 7],
                                               w[jj + 8]);
            }
        } else if (row == n - 2 && hi_kind == END_WALL) {
            // row n-2's window sits one row lower: f[n-1] = w[jj+5]
            if constexpr (ORDER == 1) val = -Rows<T, 1>::wall1(w[jj + 5], w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1]);
            else val = Rows<T, 2>::wall1(w[jj + 5], w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1]);
        }
        d[jj] = val;
    }
    if (J.rhs_out && active) {
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj < m) J.rhs_out[bofs + a * L.sa + (long)(st + jj) * L.srow] = d[jj] * J.rhs_scale;
    }
    __syncthreads();   // plan staged

    // ---- 3. segment-local solve (a non-final segment's last row is its interface row) ----------
    const bool has_iface = seg_ok && s < nl - 1;
    const int mi = has_iface ? M - 1 : m;
    const int pst = st + s;   // padded index of row st
    const T dI = d[M - 1];
    T y = T(0);
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        if (jj < mi) {
            y = cf[pst + jj] * d[jj] - cl[pst + jj] * y;
            d[jj] = y;
        }
    }
    T xn = T(0);
#pragma unroll
    for (int jj = M - 1; jj >= 0; jj--) {
        if (jj < mi) {
            xn = d[jj] - cu[pst + jj] * xn;
            d[jj] = xn;
        }
    }
    sh_xf[s * PS + p] = d[0];
    sh_xl[s * PS + p] = d[M - 2];   // last interior row of a full segment
    sh_dI[s * PS + p] = dI;
    __syncthreads();

    // ---- 4. interface system: thread t solves interface row q = t mod SM of pencil t / SM by
    //         parallel cyclic reduction over the SM lanes of its group --------------------------
    {
        const int q = tid & (SM - 1);
        const int pp = tid / SM;
        const bool iq = q < nI && (long)blockIdx.x * P + pp < L.na;
        const T *rl = J.plan + 5 * n, *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes, *k2 = k1 + kPcrSteps * kMaxLanes,
                *invb = k2 + kPcrSteps * kMaxLanes;
        const int qc = q < kMaxLanes ? q : kMaxLanes - 1;
        T R = T(0);
        if (iq) R = sh_dI[q * PS + pp] - rl[qc] * sh_xl[q * PS + pp] - ru[qc] * sh_xf[(q + 1) * PS + pp];
        int steps = 0;
        while ((1 << steps) < nI) steps++;
        for (int k = 0; k < steps; k++) {
            const int dd = 1 << k;
            const T Rm = __shfl(R, lane - dd, 64);
            const T Rp = __shfl(R, lane + dd, 64);
            // the multipliers vanish where q-dd or q+dd leaves the system, so a neighbouring
            // group's (finite) value never enters
            if (iq) R = R - k1[k * kMaxLanes + qc] * Rm - k2[k * kMaxLanes + qc] * Rp;
        }
        sh_xI[pp * XS + q] = iq ? R * invb[qc] : T(0);
    }
    __syncthreads();
    const T xl = (s > 0 && seg_ok) ? sh_xI[p * XS + s - 1] : T(0);
    const T xr = has_iface ? sh_xI[p * XS + s] : T(0);

    // ---- 5. correct and store ------------------------------------------------------------------
    if (!active) return;
    const T *__restrict__ post = J.post;
    const T scale = J.scale, beta = J.beta;
    const bool has_pm = J.mul_a || J.mul_b;
    const T pm = (J.mul_a ? J.mul_a[a] : T(1)) * (J.mul_b ? J.mul_b[b] : T(1));
    const __amdgpu_buffer_rsrc_t rout = field_rsrc(J.out + bofs);
    // with beta != 0 the old values of four rows are loaded ahead of their stores (a store may
    // alias the next row's load, so a load per store would wait a memory round trip each)
    constexpr int C = 4;
    if (kFast && ws + G < nl && !post && beta == T(0)) {
        // wave-uniform: every segment of the wavefront is full and ends in an interface row
        const unsigned vo = lofs + (unsigned)st * rowb;
#pragma unroll
        for (int jj = 0; jj < M; jj++) {
            const T x = jj < M - 1 ? d[jj] + vL[pst + jj] * xl + vR[pst + jj] * xr : xr;
            T o = scale * x;
            if (has_pm) o *= pm;
            bstore(o, rout, vo + (unsigned)jj * rowb, 0u);
        }
        return;
    }
#pragma unroll
    for (int jc = 0; jc < M; jc += C) {
        T old[C];
#pragma unroll
        for (int t = 0; t < C; t++)
            old[t] = (beta != T(0) && jc + t < m) ? bload<T>(rout, lofs + (unsigned)(st + jc + t) * rowb, 0u) : T(0);
#pragma unroll
        for (int t = 0; t < C; t++) {
            const int jj = jc + t;
            if (jj < m) {
                const int row = st + jj;
                const T x = (jj < mi) ? d[jj] + vL[pst + jj] * xl + vR[pst + jj] * xr : xr;
                T o = scale * x;
                if (post) o *= post[row];
                if (has_pm) o *= pm;
                if (beta != T(0)) o += beta * old[t];
                bstore(o, rout, lofs + (unsigned)row * rowb, 0u);
            }
        }
    }
}

// (M, G) of the wide kernel for a strided pencil of n rows, or M = 0 when k_deriv serves it
// (FP32 up to 512 rows, FP64 up to 256: segments of at most 32 / 16 rows).
void wide_shape(int n, bool f64, int &M, int &G) {
    M = G = 0;
    if (!f64) {
        if (n > 16 * 32 && n <= 2 * 16 * 32) { M = 32; G = 2; }
    } else {
        if (n > 16 * 16 && n <= 2 * 16 * 16) { M = 16; G = 2; }
        else if (n > 2 * 16 * 16 && n <= 4 * 16 * 16) { M = 16; G = 4; }
    }
}

size_t wide_lds_bytes(int n, int M, size_t elem) {
    const int nl = (n + M - 1) / M;
    return ((size_t)5 * (n + nl) + 4 * kWideShared) * elem;
}

template <typename T>
hipError_t launch_deriv_wide(const DerivLaunch<T> &L, int order, hipStream_t stream) {
    int M, G;
    wide_shape(L.n, sizeof(T) == 8, M, G);
    if (M == 0 || L.M != M || L.njobs < 1 || L.njobs > kMaxJobs || (L.n + M - 1) / M > 16 * G)
        return hipErrorInvalidValue;
    if (((double)(L.na - 1) * (double)L.sa + (double)(L.n - 1) * (double)L.srow + 1.0) * sizeof(T) >= 4294967295.0)
        return hipErrorInvalidValue;
    const int P = 64 / G;
    dim3 grid((unsigned)((L.na + P - 1) / P), L.nb, L.njobs), block(1024);
    const size_t shm = wide_lds_bytes(L.n, M, sizeof(T));
    // more than 64 KiB of dynamic LDS must be allowed per kernel
#define CSSX_LAUNCH_W(MM, GG)                                                                            \
    do {                                                                                                 \
        const void *fn = order == 1 ? (const void *)k_deriv_wide<T, 1, MM, GG>                           \
                                    : (const void *)k_deriv_wide<T, 2, MM, GG>;                          \
        if (shm > 65536) {                                                                               \
            hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);  \
            if (e != hipSuccess) return e;                                                               \
        }                                                                                                \
        if (order == 1) hipLaunchKernelGGL((k_deriv_wide<T, 1, MM, GG>), grid, block, shm, stream, L);   \
        else hipLaunchKernelGGL((k_deriv_wide<T, 2, MM, GG>), grid, block, shm, stream, L);              \
    } while (0)
    if constexpr (sizeof(T) == 4) {
        CSSX_LAUNCH_W(32, 2);
    } else {
        if (G == 2) CSSX_LAUNCH_W(16, 2);
        else CSSX_LAUNCH_W(16, 4);
    }
#undef CSSX_LAUNCH_W
    return hipGetLastError();
}

template hipError_t launch_deriv_wide<float>(const DerivLaunch<float> &, int, hipStream_t);
template hipError_t launch_deriv_wide<double>(const DerivLaunch<double> &, int, hipStream_t);

}  // namespace cssx
