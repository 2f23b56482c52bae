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

#include "cssx_bufio.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"
#include "cssx_rows.hpp"

namespace cssx {

// closure rows of the pencil's last two rows; jj = local index of row n-1 in the window
template <typename T, int ORDER>
__device__ __forceinline__ void hi_closure(const T *w, int jj, const DerivJob<T> &J, int bd, T hi_val, T &sp_hi0,
                                           T &sp_hi1) {
    // window: row n-1 -> w[jj+4], n-2 -> w[jj+3], ..., ghosts n.. -> w[jj+5..]
    if (J.hi_kind == END_WALL) {
        if constexpr (ORDER == 1) {
            T rn2 = -Rows<T, 1>::wall1(w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj]);
            T val = (bd == 3 || bd == 4) ? hi_val * J.neu_mul : T(0);
            T rn1 = Rows<T, 1>::wall0(bd == 1 || bd == 2, val, w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], -rn2);
            sp_hi0 = (bd == 1 || bd == 2) ? -rn1 : rn1;
            sp_hi1 = rn2;
        } else {
            T val = (bd == 0) ? T(0) : hi_val * J.neu_mul;
            sp_hi0 = Rows<T, 2>::wall0(bd == 1 || bd == 2, val, w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj],
                                       w[jj - 1]);
            sp_hi1 = Rows<T, 2>::wall1(w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1], w[jj]);
        }
    } else {
        if constexpr (ORDER == 1)
            sp_hi0 = Rows<T, 1>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6], w[jj + 7], w[jj + 8],
                                          J.halo_s);
        else
            sp_hi0 = Rows<T, 2>::halo_hi0(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6], w[jj + 7], w[jj + 8]);
    }
}

template <typename T, int ORDER, int M>
// FP32 segments of up to 16 rows fit 8 waves per SIMD (two 1024-thread workgroups per CU) once
// the compiler is held to that occupancy; the default register split stops at 7 (SGPRs)
__global__ __launch_bounds__(1024, (sizeof(T) == 4 && M <= 16) ? 8 : 1) void k_deriv(const DerivLaunch<T> L) {
    constexpr int W = M + 8;   // window w[i] = row st-4+i
    const DerivJob<T> &J = L.jobs[blockIdx.z];
    const int lane = threadIdx.x & 63;
    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // segment (wavefront) index
    const int S = L.S, n = L.n;
    const int a = blockIdx.x * 64 + lane;
    const int b = blockIdx.y;
    const bool active = a < L.na;
    const int st = seg_start(n, S, k);
    const int m = seg_start(n, S, k + 1) - st;   // rows in this segment
    const bool first = k == 0, last = k == S - 1;
    const bool full = m == M;                     // every segment is full when S divides n
    const int lo_kind = J.lo_kind, hi_kind = J.hi_kind;
    const long gbase = (long)a * J.g_a + (long)b * J.g_b;
    const unsigned lofs = (unsigned)((active ? a : 0) * L.sa) * (unsigned)sizeof(T);   // lane offset, bytes
    const long bofs = (long)b * L.sb;                                                   // pencil-row offset
    const long srow = L.srow;
    const unsigned rowb = (unsigned)srow * (unsigned)sizeof(T);                         // row stride, bytes
    const __amdgpu_buffer_rsrc_t rin = field_rsrc(J.in + bofs);
    // Plan coefficients: lane j holds those of segment row j (and of interface row j), so the
    // sweeps read them with v_readlane instead of keeping ~5M of them in scalar registers.
    const T *cf = J.plan, *cl = cf + n, *cu = cl + n, *vL = cu + n, *vR = vL + n;
    const T *rl = J.plan + 5 * n, *ru = rl + kMaxSeg, *rf = ru + kMaxSeg, *rcl = rf + kMaxSeg,
            *rcu = rcl + kMaxSeg;
    const int crow = min(st + (lane < M ? lane : M - 1), n - 1);
    const T p_cf = cf[crow], p_cl = cl[crow], p_cu = cu[crow], p_vL = vL[crow], p_vR = vR[crow];
    const T p_post = J.post ? J.post[crow] : T(1);
    const int cq = lane < kMaxSeg ? lane : kMaxSeg - 1;
    const T p_rl = rl[cq], p_ru = ru[cq], p_rf = rf[cq], p_rcl = rcl[cq], p_rcu = rcu[cq];
    // the row a window slot reads: wrapped on a periodic pencil, clamped otherwise (rows
    // outside the pencil are replaced below); wave-uniform, so scalar arithmetic
    auto row_of = [&](int i) {
        const int q = st - 4 + i;
        if (lo_kind == END_PERIODIC) return q < 0 ? q + n : (q >= n ? q - n : q);
        return q < 0 ? 0 : (q >= n ? n - 1 : q);
    };

    // ---- 1. window ---------------------------------------------------------------------------
    // Interior rows read 2 neighbours on each side (slots 2 .. M+5); the 4-row reach of the
    // closure rows (slots 0, 1, M+6, M+7) is loaded for the first and last segment only.
    T w[W];
#pragma unroll
    for (int i = 2; i < W - 2; i++) w[i] = bload<T>(rin, lofs, (unsigned)row_of(i) * rowb);
    w[0] = w[1] = w[W - 2] = w[W - 1] = T(0);
    if (first || last) {
        w[0] = bload<T>(rin, lofs, (unsigned)row_of(0) * rowb);
        w[1] = bload<T>(rin, lofs, (unsigned)row_of(1) * rowb);
        w[W - 2] = bload<T>(rin, lofs, (unsigned)row_of(W - 2) * rowb);
        w[W - 1] = bload<T>(rin, lofs, (unsigned)row_of(W - 1) * rowb);
    }
    if (J.pre) {
#pragma unroll
        for (int i = 0; i < W; i++) w[i] *= J.pre[row_of(i)];
    }
    if (first && lo_kind != END_PERIODIC) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            w[i] = (lo_kind == END_HALO && active) ? J.glo[gbase + (long)i * J.g_k] : T(0);
    }
    if (last && hi_kind != END_PERIODIC) {
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int q = st - 4 + i;
            if (q >= n) w[i] = (hi_kind == END_HALO && active && q - n < 4) ? J.ghi[gbase + (long)(q - n) * J.g_k] : T(0);
        }
    }

    __builtin_amdgcn_sched_barrier(0);
    // ---- 2. right-hand side rows: the Lele interior stencil, then the closure rows ----------------
    T d[M];
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        if constexpr (ORDER == 1)
            d[jj] = Rows<T, 1>::interior(w[jj + 2], w[jj + 3], w[jj + 5], w[jj + 6]);
        else
            d[jj] = Rows<T, 2>::interior(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6]);
    }
    const int bd = J.bdry;
    if (first) {
        if (lo_kind == END_WALL) {
            const T lo_val = (active && J.glo) ? J.glo[gbase] : T(0);   // prescribed wall derivative
            if constexpr (ORDER == 1) {
                const T r1 = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
                const T val = (bd == 0) ? T(0) : lo_val * J.neu_mul;
                d[0] = Rows<T, 1>::wall0(bd == 1 || bd == 3, val, w[4], w[5], w[6], w[7], r1, J.lo_r0_val != 0);
                d[1] = r1;
            } else {
                const T val = (bd == 0) ? T(0) : -lo_val * J.neu_mul;
                d[0] = Rows<T, 2>::wall0(bd == 1 || bd == 3, val, w[4], w[5], w[6], w[7], w[8], w[9]);
                d[1] = Rows<T, 2>::wall1(w[4], w[5], w[6], w[7], w[8]);
            }
        } else if (lo_kind == END_HALO) {
            if constexpr (ORDER == 1)
                d[0] = Rows<T, 1>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6], J.halo_s);
            else
                d[0] = Rows<T, 2>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
        } else {   // periodic: the seam row is a halo row with the wrapped rows as ghosts
            if constexpr (ORDER == 1)
                d[0] = Rows<T, 1>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6], J.halo_s);
            else
                d[0] = Rows<T, 2>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
        }
    }
    if (last) {
        const T hi_val = (active && hi_kind == END_WALL && J.ghi) ? J.ghi[gbase] : T(0);
        const bool hw = hi_kind == END_WALL;
        T sp0 = T(0), sp1 = T(0);
        if (full) {
            hi_closure<T, ORDER>(w, M - 1, J, bd, hi_val, sp0, sp1);
            d[M - 1] = sp0;
            if (hw) d[M - 2] = sp1;
        } else {
#pragma unroll
            for (int jj = 1; jj < M; jj++)
                if (jj == m - 1) hi_closure<T, ORDER>(w, jj, J, bd, hi_val, sp0, sp1);
#pragma unroll
            for (int jj = 0; jj < M; jj++) {
                if (jj == m - 1) d[jj] = sp0;
                if (hw && jj == m - 2) d[jj] = sp1;
            }
        }
    }
    if (J.rhs_out && active) {   // CSS-Python compact_fd6 leaves the assembled rows in darr
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj < m) at_row(J.rhs_out + bofs + (long)(st + jj) * srow, lofs) = d[jj] * J.rhs_scale;
    }

    // ---- 3. segment-local solve (the last row of every segment but the pencil's last is an
    //         interface row, solved in step 4) -----------------------------------------------------
    T dI = T(0), xfirst, xlast = T(0);
    // scheduling fences keep each phase's coefficient reads (readlanes) next to their use
    __builtin_amdgcn_sched_barrier(0);
    if (full && !last) {
        dI = d[M - 1];
        T y = T(0);
#pragma unroll
        for (int jj = 0; jj < M - 1; jj++) {
            y = lane_val(p_cf, jj) * d[jj] - lane_val(p_cl, jj) * y;
            d[jj] = y;
        }
        __builtin_amdgcn_sched_barrier(0);
        T xn = T(0);
#pragma unroll
        for (int jj = M - 2; jj >= 0; jj--) {
            xn = d[jj] - lane_val(p_cu, jj) * xn;
            d[jj] = xn;
        }
        xlast = d[M - 2];
    } else if (full) {
        T y = T(0);
#pragma unroll
        for (int jj = 0; jj < M; jj++) {
            y = lane_val(p_cf, jj) * d[jj] - lane_val(p_cl, jj) * y;
            d[jj] = y;
        }
        __builtin_amdgcn_sched_barrier(0);
        T xn = T(0);
#pragma unroll
        for (int jj = M - 1; jj >= 0; jj--) {
            xn = d[jj] - lane_val(p_cu, jj) * xn;
            d[jj] = xn;
        }
        xlast = d[M - 1];
    } else {
        const int mi = last ? m : m - 1;   // rows solved locally
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj == m - 1) dI = d[jj];
        T y = T(0);
#pragma unroll
        for (int jj = 0; jj < M; jj++) {
            if (jj < mi) {
                y = cf[st + jj] * d[jj] - cl[st + jj] * y;
                d[jj] = y;
            }
        }
        T xn = T(0);
#pragma unroll
        for (int jj = M - 1; jj >= 0; jj--) {
            if (jj < mi) {
                xn = d[jj] - cu[st + jj] * xn;
                d[jj] = xn;
            }
        }
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj == mi - 1) xlast = d[jj];
    }
    xfirst = d[0];

    __builtin_amdgcn_sched_barrier(0);
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
                T R = sh[2][q][lane] - lane_val(p_rl, q) * sh[1][q][lane] - lane_val(p_ru, q) * sh[0][q + 1][lane];
                zp = lane_val(p_rf, q) * R - lane_val(p_rcl, q) * zp;
                z[q] = zp;
            }
        }
        T xp = T(0);
#pragma unroll
        for (int q = kMaxSeg - 2; q >= 0; q--) {
            if (q < S - 1) {
                xp = z[q] - lane_val(p_rcu, q) * xp;
                z[q] = xp;
                if (q == k - 1) xl = xp;
                if (q == k) xr = xp;
            }
        }
    }

    __builtin_amdgcn_sched_barrier(0);
    // ---- 5. correct and store -----------------------------------------------------------------
    if (!active) return;
    const T *__restrict__ post = J.post;
    const T scale = J.scale, beta = J.beta;
    const bool has_pm = J.mul_a || J.mul_b;
    const T pm = (J.mul_a ? J.mul_a[a] : T(1)) * (J.mul_b ? J.mul_b[b] : T(1));
    const __amdgpu_buffer_rsrc_t rout = field_rsrc(J.out + bofs);
    // with beta != 0 the old values of four rows are loaded ahead of their stores (a store may
    // alias the next row's load, so a load per store would wait a memory round trip each)
    constexpr int C = 4;
    if (full) {
#pragma unroll
        for (int jc = 0; jc < M; jc += C) {
            T old[C];
#pragma unroll
            for (int t = 0; t < C; t++)
                old[t] = (beta != T(0) && jc + t < M) ? bload<T>(rout, lofs, (unsigned)(st + jc + t) * rowb) : T(0);
#pragma unroll
            for (int t = 0; t < C; t++) {
                const int jj = jc + t;
                if (jj < M) {
                    const int j = st + jj;
                    // the interface row (last row of a non-final segment) is the interface value
                    const T x = (jj < M - 1 || last) ? d[jj] + lane_val(p_vL, jj) * xl + lane_val(p_vR, jj) * xr : xr;
                    T o = scale * x;
                    if (post) o *= lane_val(p_post, jj);
                    if (has_pm) o *= pm;
                    if (beta != T(0)) o += beta * old[t];
                    bstore(o, rout, lofs, (unsigned)j * rowb);
                }
            }
        }
    } else {
        const int mi = last ? m : m - 1;
#pragma unroll
        for (int jc = 0; jc < M; jc += C) {
            T old[C];
#pragma unroll
            for (int t = 0; t < C; t++)
                old[t] = (beta != T(0) && jc + t < m) ? bload<T>(rout, lofs, (unsigned)(st + jc + t) * rowb) : T(0);
#pragma unroll
            for (int t = 0; t < C; t++) {
                const int jj = jc + t;
                if (jj < m) {
                    const int j = st + jj;
                    const T x = (jj < mi) ? d[jj] + vL[j] * xl + vR[j] * xr : xr;
                    T o = scale * x;
                    if (post) o *= post[j];
                    if (has_pm) o *= pm;
                    if (beta != T(0)) o += beta * old[t];
                    bstore(o, rout, lofs, (unsigned)j * rowb);
                }
            }
        }
    }
}

template <typename T>
hipError_t launch_deriv(const DerivLaunch<T> &L, int order, hipStream_t stream) {
    const int M = L.M;
    if ((M != 8 && M != 16 && M != 32 && M != 64) || L.S != seg_count(L.n, M) || L.S > kMaxSeg || L.njobs < 1 ||
        L.njobs > kMaxJobs)
        return hipErrorInvalidValue;
    // the kernel addresses a pencil row with 32-bit byte offsets (lane part + row part)
    if (((double)(L.na - 1) * (double)L.sa + (double)(L.n - 1) * (double)L.srow + 1.0) * sizeof(T) >= 4294967295.0)
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
