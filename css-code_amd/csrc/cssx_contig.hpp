// Device pieces of the contiguous-axis compact derivative (theta, stride 1) for gfx950, used by
// k_deriv_contig (cssx_deriv_contig.hip).  The operator is that of derivatives.f90 Dbydt :1010-1262 and D2bydt2
// :1265-1477: lane j of a pencil group owns rows [j*MR, j*MR+MR), the 4-row halos of a lane's
// stencil come from its neighbours by wavefront shifts, and the nl-1 interface rows are solved
// by parallel cyclic reduction over shuffles with precomputed multipliers.
#pragma once
#include <hip/hip_runtime.h>

#include "cssx_bufio.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"
#include "cssx_rows.hpp"

namespace cssx {

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) {
    return __shfl(v, src, 64);
}

// MR consecutive values as 16-byte vector accesses (p 16-byte aligned)
template <typename T, int MR>
__device__ __forceinline__ void load_run(const T *__restrict__ p, T (&v)[MR]) {
    static_assert((MR * sizeof(T)) % 16 == 0, "run must be whole 16-byte vectors");
    constexpr int NC = MR * (int)sizeof(T) / 16;
    uint4 t[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) t[c] = reinterpret_cast<const uint4 *>(p)[c];
    __builtin_memcpy(v, t, sizeof(v));
}
template <typename T, int MR>
__device__ __forceinline__ void store_run(T *__restrict__ p, const T (&v)[MR]) {
    constexpr int NC = MR * (int)sizeof(T) / 16;
    uint4 t[NC];
    __builtin_memcpy(t, v, sizeof(v));
#pragma unroll
    for (int c = 0; c < NC; c++) reinterpret_cast<uint4 *>(p)[c] = t[c];
}

// the MR coefficients of lane j from a row table staged in 16-byte chunks (see k_deriv_contig)
template <typename T, int MR>
__device__ __forceinline__ void lds_rows(const T *tab, int nlT, int j, T (&v)[MR]) {
    constexpr int E = 16 / (int)sizeof(T);
#pragma unroll
    for (int h = 0; h < MR / E; h++) {
        const uint4 t = *reinterpret_cast<const uint4 *>(tab + ((long)h * nlT + j) * E);
        __builtin_memcpy(&v[h * E], &t, 16);
    }
}

// Plan rows of one lane: the five row tables staged in LDS in 16-byte chunks (see
// k_deriv_contig).  (Read straight from the L2-resident plan buffer instead, the 512-row FP32
// sweep ran 13% slower.)
template <typename T, int MR>
struct LdsRows {
    const T *sp;
    int nT, nlT;
    __device__ __forceinline__ void rows(int c, int j, T (&v)[MR]) const { lds_rows<T, MR>(sp + c * nT, nlT, j, v); }
    __device__ __forceinline__ const T *tail() const { return sp + 5 * nT; }
};

// The raw rows [j*MR, j*MR+MR) of lane j of pencil a (contig_lane's input v): issued by the
// caller ahead of its plan staging, so that the loads are in flight meanwhile.
template <typename T, int MR, bool FULL>
__device__ __forceinline__ void contig_load(const DerivJob<T> &J, int n, int nl, long sa, long sb, long srow, long a,
                                            int b, int j, bool pen_ok, T (&v)[MR]) {
    const T *__restrict__ in = J.in;
    const long pb_ld = (pen_ok ? a : 0) * sa + (long)b * sb;
    const int st = j * MR;
    if constexpr (FULL) {
        if (j < nl) load_run<T, MR>(in + pb_ld + st, v);
    } else {
#pragma unroll
        for (int k = 0; k < MR; k++) {
            const int q = min(st + k, n - 1);
            v[k] = in[pb_ld + (long)q * srow];
        }
    }
}

// One lane's share of a contiguous compact derivative: rows [st, st+MR) of pencil a (v: their
// raw values from contig_load), solved with the rest of the pencil's lanes (halos by wave
// shifts, interfaces by PCR); x = the solution before the job's scale.  lo_val / hi_val: the prescribed wall derivative at a wall end (used by
// lane 0 / lane nl-1).
// FULL: unit-stride pencils whose length is a multiple of MR.  Every lane then holds MR rows,
// lane 0 alone touches the lower end and lane nl-1 alone the upper one, so the wall, halo and
// interface special cases become per-lane tests at compile-time row positions instead of
// per-row comparisons.
template <typename T, int ORDER, int MR, bool FULL, typename PlanRows>
__device__ __forceinline__ void contig_lane(const DerivJob<T> &J, const PlanRows &PR, int n, int nl, long sa, long sb,
                                            long srow, long a, int b, int j, int lane, bool pen_ok, T (&v)[MR],
                                            T lo_val, T hi_val, T (&x)[MR]) {
    constexpr bool vec = FULL;
    constexpr int W = MR + 8;
    const int nI = nl - 1;                    // interface equations
    const bool active = pen_ok && j < nl;
    const int st = j * MR;
    const int m = active ? min(MR, n - st) : 0;
    const long pbase = a * sa + (long)b * sb;
    const long gbase = a * J.g_a + (long)b * J.g_b;
    const T *rl = PR.tail(), *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes, *k2 = k1 + kPcrSteps * kMaxLanes,
            *invb = k2 + kPcrSteps * kMaxLanes;

    // ---- own rows (loaded by contig_load), then the halo rows from the neighbouring lanes ------
    const T *__restrict__ in = J.in;
    if (J.pre) {
        // a row profile: the lane's run as 16-byte vectors where the pencil splits into whole
        // runs and the table is 16-byte aligned (lanes past the pencil read the last run)
        T pv[MR];
        if (FULL && (reinterpret_cast<uintptr_t>(J.pre) & 15) == 0) {
            load_run<T, MR>(J.pre + min(st, n - MR), pv);
        } else {
#pragma unroll
            for (int k = 0; k < MR; k++) pv[k] = J.pre[min(st + k, n - 1)];
        }
#pragma unroll
        for (int k = 0; k < MR; k++) v[k] *= pv[k];
    }
    if constexpr (FULL) {
        if (!active) {
#pragma unroll
            for (int k = 0; k < MR; k++) v[k] = T(0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < MR; k++)
            if (k >= m) v[k] = T(0);
    }
    T w[W];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        w[k] = from_prev(v[MR - 4 + k]);     // rows st-4 .. st-1
        w[MR + 4 + k] = from_next(v[k]);      // rows st+MR .. st+MR+3
    }
#pragma unroll
    for (int k = 0; k < MR; k++) w[4 + k] = v[k];
    // rows outside the pencil: ghosts (halo end), the wrapped rows (periodic pencil), or zero
    // (wall end; not read by wall rows)
    auto ghost = [&](int q) {
        T g = T(0);
        if (pen_ok && q >= -4 && q < n + 4) {
            if (q < 0) {
                if (J.lo_kind == END_HALO) g = J.glo[gbase + (long)(q + 4) * J.g_k];
                else if (J.lo_kind == END_PERIODIC) g = in[pbase + (long)(q + n) * srow];
            } else {
                if (J.hi_kind == END_HALO) g = J.ghi[gbase + (long)(q - n) * J.g_k];
                else if (J.hi_kind == END_PERIODIC) g = in[pbase + (long)(q - n) * srow];
            }
            if (J.pre && (q < 0 ? J.lo_kind : J.hi_kind) == END_PERIODIC) g *= J.pre[q < 0 ? q + n : q - n];
        }
        return g;
    };
    if constexpr (FULL) {
        // a wall end's closure rows read no row beyond the pencil: only halo and periodic ends
        // need the ghosts
        if (j == 0 && J.lo_kind != END_WALL) {
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = ghost(i - 4);
        }
        if (j == nl - 1 && J.hi_kind != END_WALL) {
#pragma unroll
            for (int i = MR + 4; i < W; i++) w[i] = ghost(n + i - MR - 4);
        }
    } else {
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int q = st - 4 + i;
            if (q < 0 || q >= n) w[i] = ghost(q);
        }
    }
    // ---- right-hand side rows ---------------------------------------------------------------
    const int bd = J.bdry;
    T d[MR];
#pragma unroll
    for (int jj = 0; jj < MR; jj++) {
        T val;
        if constexpr (ORDER == 1)
            val = Rows<T, 1>::interior(w[jj + 2], w[jj + 3], w[jj + 5], w[jj + 6]);
        else
            val = Rows<T, 2>::interior(w[jj + 2], w[jj + 3], w[jj + 4], w[jj + 5], w[jj + 6]);
        const int row = st + jj;
        const bool r0 = FULL ? (jj == 0 && j == 0) : row == 0;
        const bool r1 = FULL ? (jj == 1 && j == 0) : row == 1;
        const bool rn1 = FULL ? (jj == MR - 1 && j == nl - 1) : row == n - 1;
        const bool rn2 = FULL ? (jj == MR - 2 && j == nl - 1) : row == n - 2;
        if (r0) {
            if (J.lo_kind == END_WALL) {
                if constexpr (ORDER == 1) {
                    T r1 = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
                    T vv = (bd == 0) ? T(0) : lo_val * J.neu_mul;
                    val = Rows<T, 1>::wall0(bd == 1 || bd == 3, vv, w[4], w[5], w[6], w[7], r1, J.lo_r0_val != 0);
                } else {
                    T vv = (bd == 0) ? T(0) : -lo_val * J.neu_mul;
                    val = Rows<T, 2>::wall0(bd == 1 || bd == 3, vv, w[4], w[5], w[6], w[7], w[8], w[9]);
                }
            } else {
                if constexpr (ORDER == 1)
                    val = Rows<T, 1>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6], J.halo_s);
                else
                    val = Rows<T, 2>::halo_lo0(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
            }
        } else if (r1 && J.lo_kind == END_WALL) {
            if constexpr (ORDER == 1) val = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
            else val = Rows<T, 2>::wall1(w[4], w[5], w[6], w[7], w[8]);
        }
        if (rn1) {
            if (J.hi_kind == END_WALL) {
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
        } else if (rn2 && J.hi_kind == END_WALL) {
            // row n-2's window sits one row lower: f[n-1] = w[jj+5]
            if constexpr (ORDER == 1) val = -Rows<T, 1>::wall1(w[jj + 5], w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1]);
            else val = Rows<T, 2>::wall1(w[jj + 5], w[jj + 4], w[jj + 3], w[jj + 2], w[jj + 1]);
        }
        d[jj] = val;
    }
    if (J.rhs_out && active) {
#pragma unroll
        for (int jj = 0; jj < MR; jj++)
            if (jj < m) J.rhs_out[pbase + (long)(st + jj) * srow] = d[jj] * J.rhs_scale;
    }

    // ---- segment-local solve --------------------------------------------------------------------
    const bool has_iface = active && j < nl - 1;
    const int mi = has_iface ? MR - 1 : m;
    const T dI = d[MR - 1];
    T tf[MR], tl[MR], tu[MR];
    PR.rows(0, j, tf);
    PR.rows(1, j, tl);
    PR.rows(2, j, tu);
    // rows solved locally: all but an interface row (FULL: the last row of every lane but the
    // pencil's last)
    auto local = [&](int jj) { return FULL ? (jj < MR - 1 || !has_iface) : jj < mi; };
    T y = T(0);
#pragma unroll
    for (int jj = 0; jj < MR; jj++) {
        if (local(jj)) {
            y = tf[jj] * d[jj] - tl[jj] * y;
            d[jj] = y;
        }
    }
    T xn = T(0);
#pragma unroll
    for (int jj = MR - 1; jj >= 0; jj--) {
        if (local(jj)) {
            xn = d[jj] - tu[jj] * xn;
            d[jj] = xn;
        }
    }
    const T xfirst = d[0];
    const T xlast = d[MR - 2];   // last interior row of a full segment

    // ---- interface system by parallel cyclic reduction -----------------------------------------
    const T xf_next = from_next(xfirst);
    T R = T(0);
    if (has_iface) R = dI - rl[j] * xlast - ru[j] * xf_next;
    int steps = 0;
    while ((1 << steps) < nI) steps++;
    if (steps > 0) {
        const T Rm = from_prev(R), Rp = from_next(R);
        if (has_iface) R = R - k1[j] * Rm - k2[j] * Rp;
    }
#pragma unroll
    for (int s = 1; s < kPcrSteps; s++) {
        if (s < steps) {   // uniform
            const int dd = 1 << s;
            const T Rm = shfl(R, lane - dd);
            const T Rp = shfl(R, lane + dd);
            if (has_iface) R = R - k1[s * kMaxLanes + j] * Rm - k2[s * kMaxLanes + j] * Rp;
        }
    }
    const T xI = has_iface ? R * invb[j] : T(0);
    T xl = from_prev(xI);
    if (j == 0) xl = T(0);
    const T xr = xI;

    T tvl[MR], tvr[MR];
    PR.rows(3, j, tvl);
    PR.rows(4, j, tvr);
#pragma unroll
    for (int jj = 0; jj < MR; jj++) x[jj] = local(jj) ? d[jj] + tvl[jj] * xl + tvr[jj] * xr : xr;
}

}  // namespace cssx
