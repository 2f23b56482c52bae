// Compact sixth-order derivative kernels for gfx950 (first derivative alpha=1/3, second
// derivative alpha=2/11), the operators of derivatives.f90 (Dbydr :547-796, Dbydt
// :1010-1262, Dbydp :1480-1591, D2byd*2 :799-1007, :1265-1477, :1594-1684).
//
// Work decomposition.  A workgroup owns 64 pencils (one per lane) and cuts each pencil into
// S segments of about M rows (one wavefront per segment, S <= 16).  Every lane:
//   1. loads its segment's rows plus 4 rows on either side (halo rows come from the ghost
//      buffers, the periodic wrap, or are not needed at a wall) - one coalesced HBM read per
//      value on the strided axes, the lanes walking the contiguous theta index;
//   2. builds the right-hand side rows d_j (interior Lele stencil, wall or halo closures);
//   3. solves its segment's rows with the precomputed local LU, in registers;
//   4. publishes three numbers per segment to LDS; every wave then solves the S-1 interface
//      system (the reduced tridiagonal system of the segmented solve);
//   5. corrects its rows with the interface values and streams f' out.
// No intermediate (the Thomas forward sweep) ever touches memory: one read + one write per
// point, the algorithmic minimum.
#include <hip/hip_runtime.h>

#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"
#include "cssx_rows.hpp"

namespace cssx {

template <typename T, int ORDER, int M>
__global__ __launch_bounds__(1024) void k_deriv(const DerivLaunch<T> L) {
    constexpr int W = M + 8;
    const DerivJob<T> &J = L.jobs[blockIdx.z];
    const int lane = threadIdx.x & 63;
    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // segment (wavefront) index
    const int S = L.S, n = L.n;
    const int a = blockIdx.x * 64 + lane;
    const int b = blockIdx.y;
    const bool active = a < L.na;
    const int st = seg_start(n, S, k);
    const int m = seg_start(n, S, k + 1) - st;  // rows in this segment
    const int mi = (k < S - 1) ? m - 1 : m;     // interior rows (last row = interface)
    const long pbase = (long)a * L.sa + (long)b * L.sb;
    const long gbase = (long)a * J.g_a + (long)b * J.g_b;

    // ---- 1. load rows st-4 .. st+M+3 ------------------------------------------------------
    // Every load is issued unconditionally from a clamped (or wrapped) row of a valid pencil, so
    // the whole window is in flight at once; rows outside the pencil are fixed up afterwards.
    const T *__restrict__ in = J.in;
    const T *__restrict__ pre = J.pre;
    const int lo_kind = J.lo_kind, hi_kind = J.hi_kind;
    const long pb_ld = (long)(active ? a : 0) * L.sa + (long)b * L.sb;
    const long srow = L.srow;
    // Interior rows read 2 neighbours on each side; only the closure rows of the first and
    // last segment (halo rows, derivatives.f90:684-720) reach 4 rows out.
    const bool edge = (k == 0) || (k == S - 1);
    T w[W];
#pragma unroll
    for (int i = 0; i < W; i++) {
        const int q = st - 4 + i;
        int qc;
        if (lo_kind == END_PERIODIC) qc = q < 0 ? q + n : (q >= n ? q - n : q);
        else qc = q < 0 ? 0 : (q >= n ? n - 1 : q);
        if (i >= 2 && i < W - 2) {
            w[i] = in[pb_ld + (long)qc * srow];
            if (pre) w[i] *= pre[qc];
        } else {
            w[i] = T(0);
            if (edge) {
                w[i] = in[pb_ld + (long)qc * srow];
                if (pre) w[i] *= pre[qc];
            }
        }
    }
    if (k == 0 && lo_kind != END_PERIODIC) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            w[i] = (lo_kind == END_HALO && active) ? J.glo[gbase + (long)i * J.g_k] : T(0);
    }
    if (k == S - 1 && hi_kind != END_PERIODIC) {
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int q = st - 4 + i;
            if (q >= n) w[i] = (hi_kind == END_HALO && active && q - n < 4) ? J.ghi[gbase + (long)(q - n) * J.g_k] : T(0);
        }
    }
    // wall values (physical first derivative prescribed at a Neumann end)
    T lo_val = T(0), hi_val = T(0);
    if (active && k == 0 && lo_kind == END_WALL && J.glo) lo_val = J.glo[gbase];
    if (active && k == S - 1 && hi_kind == END_WALL && J.ghi) hi_val = J.ghi[gbase];

    // ---- 2. closure rows (pencil ends) -------------------------------------------------------
    const int bd = J.bdry;
    T sp_lo0 = T(0), sp_lo1 = T(0), sp_hi1 = T(0), sp_hi0 = T(0);
    const bool lo_w = (k == 0) && J.lo_kind == END_WALL;
    const bool hi_w = (k == S - 1) && J.hi_kind == END_WALL;
    if (k == 0) {
        if (J.lo_kind == END_WALL) {
            if constexpr (ORDER == 1) {
                sp_lo1 = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
                T val = (bd == 0) ? T(0) : lo_val * J.neu_mul;
                sp_lo0 = Rows<T, 1>::wall0(bd == 1 || bd == 3, val, w[4], w[5], w[6], w[7], sp_lo1,
                                           J.lo_r0_val != 0);
            } else {
                T val = (bd == 0) ? T(0) : -lo_val * J.neu_mul;
                sp_lo0 = Rows<T, 2>::wall0(bd == 1 || bd == 3, val, w[4], w[5], w[6], w[7], w[8], w[9]);
                sp_lo1 = Rows<T, 2>::wall1(w[4], w[5], w[6], w[7], w[8]);
            }
        } else {
            if constexpr (ORDER == 1)
                sp_lo0 = Rows<T, 1>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6], J.halo_s);
            else
                sp_lo0 = Rows<T, 2>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
        }
    }
    if (k == S - 1) {
        // rows n-2 and n-1 sit at local rows m-2, m-1 of the last segment
#pragma unroll
        for (int jj = 1; jj < M; jj++) {
            if (jj == m - 1) {
                // window: row n-1 -> w[jj+4], n-2 -> w[jj+3], ..., ghosts n.. -> w[jj+5..]
                if (J.hi_kind == END_WALL) {
                    if constexpr (ORDER == 1) {
                        T rn2 = -Rows<T, 1>::wall1(w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj]);
                        T val = (bd == 3 || bd == 4) ? hi_val * J.neu_mul : T(0);
                        T rn1 = Rows<T, 1>::wall0(bd == 1 || bd == 2, val, w[jj + 4], w[jj + 3], w[jj + 2],
                                                  w[jj + 1], -rn2);
                        sp_hi0 = (bd == 1 || bd == 2) ? -rn1 : rn1;
                        sp_hi1 = rn2;
                    } else {
                        T val = (bd == 0) ? T(0) : hi_val * J.neu_mul;
                        sp_hi0 = Rows<T, 2>::wall0(bd == 1 || bd == 2, val, w[jj + 4], w[jj + 3], w[jj + 2],
                                                   w[jj + 1], w[jj], w[jj - 1]);
                        sp_hi1 = Rows<T, 2>::wall1(w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj]);
                    }
                } else {
                    if constexpr (ORDER == 1)
                        sp_hi0 = Rows<T, 1>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6],
                                                      w[jj + 7], w[jj + 8], J.halo_s);
                    else
                        sp_hi0 = Rows<T, 2>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6],
                                                      w[jj + 7], w[jj + 8]);
                }
            }
        }
    }

    // ---- 3. right-hand side rows fused with the segment-local forward sweep ---------------------
    const T *cf = J.plan, *cl = cf + n, *cu = cl + n, *vL = cu + n, *vR = vL + n;
    const T *rl = J.plan + 5 * n, *ru = rl + kMaxSeg, *rf = ru + kMaxSeg, *rcl = rf + kMaxSeg,
            *rcu = rcl + kMaxSeg;
    T d[M];
    T dI = T(0), y = T(0);
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        // window index of row st+jj is jj+4
        T v;
        if constexpr (ORDER == 1)
            v = Rows<T, 1>::interior(w[jj + 2], w[jj + 3], w[jj + 5], w[jj + 6]);
        else
            v = Rows<T, 2>::interior(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6]);
        if (jj == 0 && k == 0) v = sp_lo0;
        if (jj == 1 && lo_w) v = sp_lo1;
        if (k == S - 1) {
            if (jj == m - 1) v = sp_hi0;
            if (jj == m - 2 && hi_w) v = sp_hi1;
        }
        if (J.rhs_out && active && jj < m) J.rhs_out[pbase + (long)(st + jj) * L.srow] = v * J.rhs_scale;
        if (jj == m - 1) dI = v;
        if (jj < mi) {
            y = cf[st + jj] * v - cl[st + jj] * y;
            d[jj] = y;
        } else {
            d[jj] = T(0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    T xn = T(0);
#pragma unroll
    for (int jj = M - 1; jj >= 0; jj--) {
        if (jj < mi) {
            xn = d[jj] - cu[st + jj] * xn;
            d[jj] = xn;
        }
    }
    T xfirst = d[0], xlast = T(0);
#pragma unroll
    for (int jj = 0; jj < M; jj++)
        if (jj == mi - 1) xlast = d[jj];

    // ---- 4. interface system ----------------------------------------------------------------
    __shared__ T sh[3][kMaxSeg][64];
    T xl = T(0), xr = T(0);
    if (S > 1) {
        sh[0][k][lane] = xfirst;
        sh[1][k][lane] = xlast;
        sh[2][k][lane] = dI;
        __syncthreads();
        T z[kMaxSeg];
        T zp = T(0);
#pragma unroll
        for (int q = 0; q < kMaxSeg - 1; q++) {
            if (q < S - 1) {
                T R = sh[2][q][lane] - rl[q] * sh[1][q][lane] - ru[q] * sh[0][q + 1][lane];
                zp = rf[q] * R - rcl[q] * zp;
                z[q] = zp;
            }
        }
        T xp = T(0);
#pragma unroll
        for (int q = kMaxSeg - 2; q >= 0; q--) {
            if (q < S - 1) {
                xp = z[q] - rcu[q] * xp;
                z[q] = xp;
                if (q == k - 1) xl = xp;
                if (q == k) xr = xp;
            }
        }
    }

    __builtin_amdgcn_sched_barrier(0);
    // ---- 5. correct and store -----------------------------------------------------------------
    if (!active) return;
    T *__restrict__ out = J.out;
    const T *__restrict__ post = J.post;
    const T scale = J.scale, beta = J.beta;
    T o[M];
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        const int j = st + jj < n ? st + jj : n - 1;
        T x = (jj < mi) ? d[jj] + vL[j] * xl + vR[j] * xr : xr;
        o[jj] = scale * x;
        if (post) o[jj] *= post[j];
    }
    if (beta != T(0)) {
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj < m) o[jj] += beta * out[pbase + (long)(st + jj) * srow];
    }
#pragma unroll
    for (int jj = 0; jj < M; jj++)
        if (jj < m) out[pbase + (long)(st + jj) * srow] = o[jj];
}

template <typename T>
hipError_t launch_deriv(const DerivLaunch<T> &L, int order, hipStream_t stream) {
    const int M = L.M;
    if ((M != 8 && M != 16 && M != 32 && M != 64) || L.S != seg_count(L.n, M) || L.S > kMaxSeg || L.njobs < 1 ||
        L.njobs > kMaxJobs)
        return hipErrorInvalidValue;
    dim3 grid((L.na + 63) / 64, L.nb, L.njobs), block(64 * L.S);
#define CSSX_LAUNCH(MM)                                                                              \
    do {                                                                                             \
        if (order == 1) hipLaunchKernelGGL((k_deriv<T, 1, MM>), grid, block, 0, stream, L);          \
        else hipLaunchKernelGGL((k_deriv<T, 2, MM>), grid, block, 0, stream, L);                     \
    } while (0)
    switch (M) {
    case 8: CSSX_LAUNCH(8); break;
    case 16: CSSX_LAUNCH(16); break;
    case 32: CSSX_LAUNCH(32); break;
    default: CSSX_LAUNCH(64); break;
    }
#undef CSSX_LAUNCH
    return hipGetLastError();
}

template hipError_t launch_deriv<float>(const DerivLaunch<float> &, int, hipStream_t);
template hipError_t launch_deriv<double>(const DerivLaunch<double> &, int, hipStream_t);

}  // namespace cssx
