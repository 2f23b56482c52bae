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

#include "cssx_contig.hpp"

namespace cssx {

template <typename T, int ORDER, int MR, bool FULL>
__global__ __launch_bounds__(1024) void k_deriv_contig(const DerivLaunch<T> L) {
    extern __shared__ unsigned char smem_raw[];
    T *sp = reinterpret_cast<T *>(smem_raw);
    const DerivJob<T> &J = L.jobs[blockIdx.z];
    const int n = L.n;
    // unit-stride pencils whose length is a multiple of MR move as 16-byte vectors
    constexpr bool vec = FULL;
    const int b = blockIdx.y;
    const int Lw = L.S;                       // lanes per pencil group (power of two)
    const int nl = (n + MR - 1) / MR;         // lanes holding rows
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
    T lo_val = T(0), hi_val = T(0);
    if (active && j == 0 && J.lo_kind == END_WALL && J.glo) lo_val = J.glo[gbase];
    if (active && j == nl - 1 && J.hi_kind == END_WALL && J.ghi) hi_val = J.ghi[gbase];
    // own rows first: in flight while the plan is staged
    T v[MR];
    contig_load<T, MR, FULL>(J, n, nl, L.sa, L.sb, L.srow, a, b, j, pen_ok, v);
    // The five row tables are staged in 16-byte chunks of E = 16/sizeof(T) rows: rows
    // jj = h*E..h*E+E-1 of lane j sit at [(h*nl + j)*E], so a lane reads E coefficients with one
    // 16-byte LDS load and the 16 lanes of a load phase read consecutive chunks (no bank
    // conflicts); the interface tables follow them unchanged.
    constexpr int E = 16 / (int)sizeof(T);
    const int nlT = (n + MR - 1) / MR, nT = nlT * MR;
    constexpr int tail = 3 * kMaxLanes + 2 * kPcrSteps * kMaxLanes;
    if (n <= (int)blockDim.x && tail <= (int)blockDim.x) {
        // one row of every table and one tail value per thread: all six loads issued before
        // the first wait
        const int r = threadIdx.x;
        T pv[5], tv = T(0);
        if (r < n) {
#pragma unroll
            for (int c = 0; c < 5; c++) pv[c] = J.plan[c * n + r];
        }
        if (r < tail) tv = J.plan[5 * n + r];
        if (r < n) {
            const int jl = r / MR, jr = r % MR;
#pragma unroll
            for (int c = 0; c < 5; c++) sp[c * nT + ((jr / E) * nlT + jl) * E + jr % E] = pv[c];
        }
        if (r < tail) sp[5 * nT + r] = tv;
    } else {
        // (one pass per table: MR and E are powers of two, so no division by n)
#pragma unroll
        for (int c = 0; c < 5; c++)
            for (int r = threadIdx.x; r < n; r += blockDim.x) {
                const int jl = r / MR, jr = r % MR;
                sp[c * nT + ((jr / E) * nlT + jl) * E + jr % E] = J.plan[c * n + r];
            }
        for (int i = threadIdx.x; i < tail; i += blockDim.x) sp[5 * nT + i] = J.plan[5 * n + i];
    }
    __syncthreads();
    // end-row taps of the raw input (before contig_lane applies a row profile to v); the lane's
    // last row is m - 1
    const long tap_at = (long)b * L.na + a;
    if (J.tap_in && active) {
        if (j == 0) J.tap_in[tap_at] = v[0];
        if (j == nl - 1) {
            T e = v[MR - 1];
            if constexpr (!FULL) {
#pragma unroll
                for (int k = 0; k < MR - 1; k++)
                    if (k == m - 1) e = v[k];
            }
            J.tap_in[J.tap_side + tap_at] = e;
        }
    }
    T x[MR];
    contig_lane<T, ORDER, MR, FULL>(J, LdsRows<T, MR>{sp, nT, nlT}, n, nl, L.sa, L.sb, L.srow, a, b, j, lane, pen_ok,
                                    v, lo_val, hi_val, x);

    // ---- correct and store ------------------------------------------------------------------------
    if (active) {
    T *__restrict__ out = J.out;
    const T *__restrict__ post = J.post;
    const T scale = J.scale, beta = J.beta;
    const bool has_pm = J.mul_a || J.mul_b;
    const T pm = (J.mul_a ? J.mul_a[a] : T(1)) * (J.mul_b ? J.mul_b[b] : T(1));
    T o[MR], pw[MR];
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
        o[jj] = scale * x[jj];
        if (post) o[jj] *= pw[jj];
        if (has_pm) o[jj] *= pm;
    }
    if (beta != T(0)) {
#pragma unroll
        for (int jj = 0; jj < MR; jj++)
            if (jj < m) o[jj] += beta * out[pbase + (long)(st + jj) * L.srow];
    }
    if (J.tap_out) {
        if (j == 0) J.tap_out[tap_at] = o[0];
        if (j == nl - 1) {
            T e = o[MR - 1];
            if constexpr (!FULL) {
#pragma unroll
                for (int k = 0; k < MR - 1; k++)
                    if (k == m - 1) e = o[k];
            }
            J.tap_out[J.tap_side + tap_at] = e;
        }
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
