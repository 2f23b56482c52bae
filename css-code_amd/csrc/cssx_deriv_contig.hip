// Compact derivative along a contiguous axis (theta, stride 1) for gfx950.
//
// Same operator as k_deriv (derivatives.f90 Dbydt :1010-1262, D2bydt2 :1265-1477), but the
// lanes walk the pencil instead of across pencils: lane j of a pencil group owns rows
// [j*MR, j*MR+MR), so a wavefront reads and writes whole pencils as contiguous runs (fully
// coalesced).  The 4-row halos of a lane's stencil come from its neighbours by wavefront
// shuffles, the nl-1 interface rows are solved by parallel cyclic reduction over shuffles with
// precomputed multipliers, and no LDS traffic or barrier touches the data path (the plan
// tables are staged in LDS once per workgroup).
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

// FULL: unit-stride pencils whose length is a multiple of MR.  Every lane then holds MR rows,
// lane 0 alone touches the lower end and lane nl-1 alone the upper one, so the wall, halo and
// interface special cases become per-lane tests at compile-time row positions instead of
// per-row comparisons.
template <typename T, int ORDER, int MR, bool FULL>
__global__ __launch_bounds__(1024) void k_deriv_contig(const DerivLaunch<T> L) {
    extern __shared__ unsigned char smem_raw[];
    T *sp = reinterpret_cast<T *>(smem_raw);
    const DerivJob<T> &J = L.jobs[blockIdx.z];
    const int n = L.n;
    // The five row tables are staged in 16-byte chunks of E = 16/sizeof(T) rows: rows
    // jj = h*E..h*E+E-1 of lane j sit at [(h*nl + j)*E], so a lane reads E coefficients with one
    // 16-byte LDS load and the 16 lanes of a load phase read consecutive chunks (no bank
    // conflicts); the interface tables follow them unchanged.
    constexpr int E = 16 / (int)sizeof(T);
    const int nlT = (n + MR - 1) / MR, nT = nlT * MR;
    // (one pass per table: MR and E are powers of two, so no division by n)
#pragma unroll
    for (int c = 0; c < 5; c++)
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            const int jl = r / MR, jr = r % MR;
            sp[c * nT + ((jr / E) * nlT + jl) * E + jr % E] = J.plan[c * n + r];
        }
    constexpr int tail = 3 * kMaxLanes + 2 * kPcrSteps * kMaxLanes;
    for (int i = threadIdx.x; i < tail; i += blockDim.x) sp[5 * nT + i] = J.plan[5 * n + i];
    __syncthreads();
    // unit-stride pencils whose length is a multiple of MR move as 16-byte vectors
    constexpr bool vec = FULL;
    const int b = blockIdx.y;
    const T *cf = sp, *cl = cf + nT, *cu = cl + nT, *vL = cu + nT, *vR = vL + nT;
    const T *rl = sp + 5 * nT, *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes, *k2 = k1 + kPcrSteps * kMaxLanes,
            *invb = k2 + kPcrSteps * kMaxLanes;

    constexpr int W = MR + 8;
    const int Lw = L.S;                       // lanes per pencil group (power of two)
    const int nl = (n + MR - 1) / MR;         // lanes holding rows
    const int nI = nl - 1;                    // interface equations
    const int lane = threadIdx.x & 63;
    const int j = lane & (Lw - 1);
    const int ppw = 64 / Lw;
    const long a = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * ppw + lane / Lw;
    const bool pen_ok = a < L.na;
    const bool active = pen_ok && j < nl;
    const int st = j * MR;
    const int m = active ? min(MR, n - st) : 0;
    const long pbase = a * L.sa + (long)b * L.sb;
    const long gbase = a * J.g_a + (long)b * J.g_b;

    // ---- own rows, then the halo rows from the neighbouring lanes ----------------------------
    const T *__restrict__ in = J.in;
    const long pb_ld = (pen_ok ? a : 0) * L.sa + (long)b * L.sb;
    T v[MR];
    if (vec) {
        if (j < nl) load_run<T, MR>(in + pb_ld + st, v);
    } else {
#pragma unroll
        for (int k = 0; k < MR; k++) {
            const int q = min(st + k, n - 1);
            v[k] = in[pb_ld + (long)q * L.srow];
        }
    }
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
                else if (J.lo_kind == END_PERIODIC) g = in[pbase + (long)(q + n) * L.srow];
            } else {
                if (J.hi_kind == END_HALO) g = J.ghi[gbase + (long)(q - n) * J.g_k];
                else if (J.hi_kind == END_PERIODIC) g = in[pbase + (long)(q - n) * L.srow];
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
    T lo_val = T(0), hi_val = T(0);
    if (active && j == 0 && J.lo_kind == END_WALL && J.glo) lo_val = J.glo[gbase];
    if (active && j == nl - 1 && J.hi_kind == END_WALL && J.ghi) hi_val = J.ghi[gbase];

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
            if (jj < m) J.rhs_out[pbase + (long)(st + jj) * L.srow] = d[jj] * J.rhs_scale;
    }

    // ---- segment-local solve --------------------------------------------------------------------
    const bool has_iface = active && j < nl - 1;
    const int mi = has_iface ? MR - 1 : m;
    const T dI = d[MR - 1];
    T tf[MR], tl[MR], tu[MR];
    lds_rows<T, MR>(cf, nlT, j, tf);
    lds_rows<T, MR>(cl, nlT, j, tl);
    lds_rows<T, MR>(cu, nlT, j, tu);
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

    // ---- correct and store ------------------------------------------------------------------------
    if (active) {
    T *__restrict__ out = J.out;
    const T *__restrict__ post = J.post;
    const T scale = J.scale, beta = J.beta;
    const bool has_pm = J.mul_a || J.mul_b;
    const T pm = (J.mul_a ? J.mul_a[a] : T(1)) * (J.mul_b ? J.mul_b[b] : T(1));
    T o[MR], tvl[MR], tvr[MR], pw[MR];
    lds_rows<T, MR>(vL, nlT, j, tvl);
    lds_rows<T, MR>(vR, nlT, j, tvr);
    if (post) {   // as pre above
        if (FULL && (reinterpret_cast<uintptr_t>(post) & 15) == 0) {
            load_run<T, MR>(post + min(st, n - MR), pw);
        } else {
#pragma unroll
            for (int jj = 0; jj < MR; jj++) pw[jj] = post[min(st + jj, n - 1)];
        }
    }
#pragma unroll
    for (int jj = 0; jj < MR; jj++) {
        T x = local(jj) ? d[jj] + tvl[jj] * xl + tvr[jj] * xr : xr;
        o[jj] = scale * x;
        if (post) o[jj] *= pw[jj];
        if (has_pm) o[jj] *= pm;
    }
    if (beta != T(0)) {
#pragma unroll
        for (int jj = 0; jj < MR; jj++)
            if (jj < m) o[jj] += beta * out[pbase + (long)(st + jj) * L.srow];
    }
    if (vec) {
        store_run<T, MR>(out + pbase + st, o);
    } else {
#pragma unroll
        for (int jj = 0; jj < MR; jj++)
            if (jj < m) out[pbase + (long)(st + jj) * L.srow] = o[jj];
    }
    }   // active
}

template <typename T>
hipError_t launch_deriv_contig(const DerivLaunch<T> &L, int order, hipStream_t stream) {
    const int MR = L.M;
    if ((MR != 4 && MR != 8 && MR != 16) || (L.n + MR - 1) / MR > kMaxLanes || L.njobs < 1 || L.njobs > kMaxJobs)
        return hipErrorInvalidValue;
    const int Lw = L.S;
    const int ppw = 64 / Lw;
    // 16 wavefronts per workgroup share one staging of the plan (5n + 960 values)
    const int wpb = 16;
    const unsigned gx = (unsigned)((L.na + (long)ppw * wpb - 1) / ((long)ppw * wpb));
    dim3 grid(gx, L.nb, L.njobs), block(64 * wpb);
    const size_t shm = (size_t)(5 * MR * ((L.n + MR - 1) / MR) + 3 * kMaxLanes + 2 * kPcrSteps * kMaxLanes) * sizeof(T);
    const bool full = L.srow == 1 && L.n % MR == 0;
#define CSSX_LAUNCH_C(MM)                                                                                     \
    do {                                                                                                      \
        if (full) {                                                                                           \
            if (order == 1) hipLaunchKernelGGL((k_deriv_contig<T, 1, MM, true>), grid, block, shm, stream, L);  \
            else hipLaunchKernelGGL((k_deriv_contig<T, 2, MM, true>), grid, block, shm, stream, L);             \
        } else {                                                                                              \
            if (order == 1) hipLaunchKernelGGL((k_deriv_contig<T, 1, MM, false>), grid, block, shm, stream, L); \
            else hipLaunchKernelGGL((k_deriv_contig<T, 2, MM, false>), grid, block, shm, stream, L);            \
        }                                                                                                     \
    } while (0)
    switch (MR) {
    case 4: CSSX_LAUNCH_C(4); break;
    case 8: CSSX_LAUNCH_C(8); break;
    default: CSSX_LAUNCH_C(16); break;
    }
#undef CSSX_LAUNCH_C
    return hipGetLastError();
}

template hipError_t launch_deriv_contig<float>(const DerivLaunch<float> &, int, hipStream_t);
template hipError_t launch_deriv_contig<double>(const DerivLaunch<double> &, int, hipStream_t);

}  // namespace cssx
