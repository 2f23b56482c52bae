// Batched compact derivatives of many independent patches of uneven shape in one persistent
// launch (gfx950).  Each patch is a small subdomain f(nth, nphi, nr) with walls on all six
// faces (the Dbydr/Dbydt/Dbydp closures of derivatives.f90:547-796, :1010-1262 applied on
// every axis, as Determine_Dtype's dtype 0 would ask); all three first derivatives of every
// patch are work items of one kernel:
//   * r and phi (strided) items: a workgroup of NT lanes covers P = NT/SM pencils cut
//     into SM >= ceil(n/M) segments of M rows (lane = segment * P + pencil), loads stay
//     coalesced along theta, the segment solves run in registers and the interface system by
//     parallel cyclic reduction over SM lanes (the contiguous-kernel plan cut at M); each
//     strided job takes M or M2 rows per segment, whichever wastes fewer lanes;
//   * theta (contiguous) items: lanes walk the pencil in runs of MR rows (the contiguous
//     kernel's scheme) with the halo rows exchanged by DPP wave shifts.
// Workgroups loop over the item table (persistent, as many as can be resident), staging each
// item's plan in LDS.  (M, M2, MR, NT) are template choices; cssx_patchset_autotune times the
// instantiated variants on the actual patch set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/cssx.h"
#include "cssx_bufio.hpp"
#include "cssx_internal.hpp"
#include "cssx_kernels.hpp"
#include "cssx_plan.hpp"
#include "cssx_rows.hpp"

namespace cssx {

template <typename T>
struct PatchJob {
    const T *in;
    T *out;
    const T *plan;
    int n, na, nb, lanes;   // pencil length, pencils (a fast, b slow), SM or lanes per pencil
    long sa, sb, srow;
    T scale;
    int bdry, kind;         // kind 0 strided (M rows per segment), 2 strided (M2), 1 contiguous
};

constexpr int kPatchTail = 3 * kMaxLanes + 2 * kPcrSteps * kMaxLanes;
// capacity of each padded interface exchange array: SM*(P+1) or P*(SM+1) = NT + max(SM, P)
// slots with P = NT/SM, at most 2*NT
__host__ __device__ constexpr int patch_shared(int nt) { return 2 * nt; }

// A raw buffer resource over bytes [0, bytes) of base: loads outside it read zero and stores
// outside it are dropped (the range check of raw buffers), which replaces per-row bounds tests.
// A lane with no pencil takes the offset kNoPencil; a row below zero wraps to >= 2^32 - 4*rowb.
// The host keeps (n + 20) * rowb <= 2^31, so that neither can come back into range.
constexpr unsigned kNoPencil = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t span_rsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}

// First-derivative right-hand side rows st .. st+MS-1 of a pencil of n rows with walls at
// both ends (prescribed wall derivatives zero): w[i] holds row st-4+i (zero outside the
// pencil) and end(k) returns row n-1-k.  Interior rows everywhere, then the wall closures in
// the segment that starts the pencil (rows 0, 1) and in the one or two that end it (rows
// n-2, n-1).
template <typename T, int MS, class End>
__device__ __forceinline__ void patch_rows(T (&d)[MS], const T *w, int st, int n, int bd, End end) {
#pragma unroll
    for (int jj = 0; jj < MS; jj++) d[jj] = Rows<T, 1>::interior(w[jj + 2], w[jj + 3], w[jj + 5], w[jj + 6]);
    if (st == 0) {
        const T r1 = Rows<T, 1>::wall1(w[4], w[5], w[6], w[7], w[8]);
        d[1] = r1;
        d[0] = Rows<T, 1>::wall0(bd == 1 || bd == 3, T(0), w[4], w[5], w[6], w[7], r1);
    }
    const int jl = n - 1 - st;   // row n-1 at jj = jl
    if (jl >= 0 && jl <= MS) {
        const T e0 = end(0), e1 = end(1), e2 = end(2), e3 = end(3), e4 = end(4);
        const T rn2 = -Rows<T, 1>::wall1(e0, e1, e2, e3, e4);
        const T rn1 = Rows<T, 1>::wall0(bd == 1 || bd == 2, T(0), e0, e1, e2, e3, -rn2);
        const T vn1 = (bd == 1 || bd == 2) ? -rn1 : rn1;
#pragma unroll
        for (int jj = 0; jj < MS; jj++) {
            if (jj == jl) d[jj] = vn1;
            if (jj == jl - 1) d[jj] = rn2;
        }
    }
}

template <typename T, int M, int NT>
__device__ __forceinline__ void strided_item(const PatchJob<T> &J, int ablk, int b, T *sp) {
    constexpr int W = M + 8;
    const int n = J.n, nl = (n + M - 1) / M, nI = nl - 1, SM = J.lanes, P = NT / SM;
    const int pn = n + nl;
    const int tid = threadIdx.x, lane = tid & 63;
    // exchange arrays with one padding slot per row ([segment][pencil] stride PS, [pencil]
    // [segment] stride XS), so that the lanes of a wavefront, which walk the other index, hit
    // different banks
    const int PS = P + 1, XS = SM + 1;
    constexpr int XC = patch_shared(NT);
    T *sh_xf = sp + 5 * pn, *sh_xl = sh_xf + XC, *sh_dI = sh_xl + XC, *sh_xI = sh_dI + XC;
    T *sh_tail = sh_xI + XC;
    const T *cf = sp, *cl = cf + pn, *cu = cl + pn, *vL = cu + pn, *vR = vL + pn;
    const int s = tid / P, p = tid - s * P;
    const long a = (long)ablk * P + p;
    const bool pen_ok = a < J.na, seg_ok = s < nl;
    const int st = s * M;
    const int m = seg_ok ? min(M, n - st) : 0;
    const long bofs = (long)b * J.sb;
    const unsigned rowb = (unsigned)J.srow * (unsigned)sizeof(T);
    // the resource spans rows [0, n) of the item's pencils: halo rows past the walls read zero
    // and stores of rows >= n, of lanes without a pencil or a segment, are dropped
    const unsigned lofs = pen_ok ? (unsigned)(a * J.sa) * (unsigned)sizeof(T) : kNoPencil;
    const unsigned span = (unsigned)n * rowb;
    const int sr = seg_ok ? st : n;
    const __amdgpu_buffer_rsrc_t rin = span_rsrc(J.in + bofs, span);
    const unsigned v0 = lofs + (unsigned)(sr - 4) * rowb;
    T w[W];
#pragma unroll
    for (int i = 0; i < W; i++) w[i] = bload<T>(rin, v0 + (unsigned)i * rowb, 0u);
    // plan staged in LDS while the rows are in flight: segment tables with one slot of padding
    // per segment, then the interface-solve tables
    for (int i = tid; i < 5 * n; i += NT) {
        const int c = i / n, r = i - c * n;
        sp[c * pn + r + r / M] = J.plan[i];
    }
    for (int i = tid; i < kPatchTail; i += NT) sh_tail[i] = J.plan[5 * n + i];
    T d[M];
    patch_rows<T, M>(d, w, sr, n, J.bdry, [&](int k) { return bload<T>(rin, lofs + (unsigned)(n - 1 - k) * rowb, 0u); });
    __syncthreads();   // plan staged
    const bool has_iface = seg_ok && s < nI;
    const int mi = has_iface ? M - 1 : m;
    const int pst = st + s;
    const T dI = d[M - 1];
    T y = T(0);
#pragma unroll
    for (int jj = 0; jj < M; jj++)
        if (jj < mi) {
            y = cf[pst + jj] * d[jj] - cl[pst + jj] * y;
            d[jj] = y;
        }
    T xn = T(0);
#pragma unroll
    for (int jj = M - 1; jj >= 0; jj--)
        if (jj < mi) {
            xn = d[jj] - cu[pst + jj] * xn;
            d[jj] = xn;
        }
    sh_xf[s * PS + p] = d[0];
    sh_xl[s * PS + p] = d[M - 2];
    sh_dI[s * PS + p] = dI;
    __syncthreads();
    {
        const int q = tid & (SM - 1), pp = tid / SM;
        const bool iq = q < nI && (long)ablk * P + pp < J.na;
        const T *rl = sh_tail, *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes, *k2 = k1 + kPcrSteps * kMaxLanes,
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
            if (iq) R = R - k1[k * kMaxLanes + qc] * Rm - k2[k * kMaxLanes + qc] * Rp;
        }
        sh_xI[pp * XS + q] = iq ? R * invb[qc] : T(0);
    }
    __syncthreads();
    const T xl = (s > 0 && seg_ok) ? sh_xI[p * XS + s - 1] : T(0);
    const T xr = has_iface ? sh_xI[p * XS + s] : T(0);
    const __amdgpu_buffer_rsrc_t rout = span_rsrc(J.out + bofs, span);
    const unsigned o0 = lofs + (unsigned)sr * rowb;
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        const T x = (jj < mi) ? d[jj] + vL[pst + jj] * xl + vR[pst + jj] * xr : xr;
        bstore(J.scale * x, rout, o0 + (unsigned)jj * rowb, 0u);
    }
    __syncthreads();   // LDS reused by the next item
}

template <typename T, int MR, int NT>
__device__ __forceinline__ void contig_item(const PatchJob<T> &J, int ablk, int b, T *sp) {
    constexpr int W = MR + 8;
    const int n = J.n, nl = (n + MR - 1) / MR, nI = nl - 1, Lw = J.lanes, nT = nl * MR;
    const int tid = threadIdx.x, lane = tid & 63;
    const T *cf = sp, *cl = cf + nT, *cu = cl + nT, *vL = cu + nT, *vR = vL + nT;
    const T *rl = sp + 5 * nT, *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes, *k2 = k1 + kPcrSteps * kMaxLanes,
            *invb = k2 + kPcrSteps * kMaxLanes;
    const int j = lane & (Lw - 1);
    const int ppw = 64 / Lw;
    const long a = ((long)ablk * (NT / 64) + (tid >> 6)) * ppw + lane / Lw;
    const bool pen_ok = a < J.na, active = pen_ok && j < nl;
    const int st = j * MR;
    const int m = active ? min(MR, n - st) : 0;
    // the resource spans the item's plane of pencils (contiguous rows, srow 1): a run's rows
    // past n belong to the next pencil and are zeroed below, lanes without a pencil read zero
    // and their stores are dropped
    const unsigned pofs = pen_ok ? (unsigned)(a * J.sa) * (unsigned)sizeof(T) : kNoPencil;
    const unsigned span = (unsigned)(J.na * J.sa) * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rin = span_rsrc(J.in + (long)b * J.sb, span);
    const unsigned o0 = pofs + (unsigned)st * (unsigned)sizeof(T);
    T v[MR];
#pragma unroll
    for (int k = 0; k < MR; k++) v[k] = bload<T>(rin, o0 + (unsigned)(k * sizeof(T)), 0u);
    // plan staged in LDS (row-major over runs, so lanes read consecutive slots) while the rows
    // are in flight
    for (int i = tid; i < 5 * n; i += NT) {
        const int c = i / n, r = i - c * n;
        const int jl = r / MR, jr = r - jl * MR;
        sp[c * nT + jr * nl + jl] = J.plan[i];
    }
    for (int i = tid; i < kPatchTail; i += NT) sp[5 * nT + i] = J.plan[5 * n + i];
#pragma unroll
    for (int k = 0; k < MR; k++)
        if (k >= m) v[k] = T(0);
    // halo rows from the neighbouring lanes by DPP wave shifts.  The first run's left halo
    // (rows < 0) and the last run's right halo (rows >= n) must read zero: the lanes that feed
    // them across a pencil boundary (the last lane of the previous pencil, the first lane of the
    // next) send zero, and inactive lanes hold zero runs.  The masking sits on the sending side
    // so that no select follows a shift: the compiler may turn such a select into an exec-masked
    // branch around the shift, whose source lanes then read as disabled (zero).
    T w[W];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        w[k] = from_prev(j == Lw - 1 ? T(0) : v[MR - 4 + k]);
        w[MR + 4 + k] = from_next(j == 0 ? T(0) : v[k]);
    }
#pragma unroll
    for (int k = 0; k < MR; k++) w[4 + k] = v[k];
    T d[MR];
    patch_rows<T, MR>(d, w, st, n, J.bdry,
                      [&](int k) { return bload<T>(rin, pofs + (unsigned)((n - 1 - k) * sizeof(T)), 0u); });
    __syncthreads();   // plan staged
    const bool has_iface = active && j < nI;
    const int mi = has_iface ? MR - 1 : m;
    const T dI = d[MR - 1];
    T y = T(0);
#pragma unroll
    for (int jj = 0; jj < MR; jj++)
        if (jj < mi) {
            y = cf[jj * nl + j] * d[jj] - cl[jj * nl + j] * y;
            d[jj] = y;
        }
    T xn = T(0);
#pragma unroll
    for (int jj = MR - 1; jj >= 0; jj--)
        if (jj < mi) {
            xn = d[jj] - cu[jj * nl + j] * xn;
            d[jj] = xn;
        }
    const T xf_next = __shfl(d[0], lane + 1, 64);
    T R = T(0);
    if (has_iface) R = dI - rl[j] * d[MR - 2] - ru[j] * xf_next;
    int steps = 0;
    while ((1 << steps) < nI) steps++;
    for (int s = 0; s < steps; s++) {
        const int dd = 1 << s;
        const T Rm = __shfl(R, lane - dd, 64);
        const T Rp = __shfl(R, lane + dd, 64);
        if (has_iface) R = R - k1[s * kMaxLanes + j] * Rm - k2[s * kMaxLanes + j] * Rp;
    }
    const T xI = has_iface ? R * invb[j] : T(0);
    T xl = __shfl(xI, lane - 1, 64);
    if (j == 0) xl = T(0);
    // rows past n (the next pencil's) go to an offset outside the resource
    const __amdgpu_buffer_rsrc_t rout = span_rsrc(J.out + (long)b * J.sb, span);
#pragma unroll
    for (int jj = 0; jj < MR; jj++) d[jj] = J.scale * ((jj < mi) ? d[jj] + vL[jj * nl + j] * xl + vR[jj * nl + j] * xI : xI);
    if (m == MR) {
        // a full run as 16-byte stores (buffer stores need 4-byte alignment only)
        constexpr int NC = MR * (int)sizeof(T) / 16;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            cssx_u32x4 q;
            __builtin_memcpy(&q, &d[c * (16 / (int)sizeof(T))], 16);
            __builtin_amdgcn_raw_buffer_store_b128(q, rout, o0 + (unsigned)(c * 16), 0u, 0);
        }
    } else {
#pragma unroll
        for (int jj = 0; jj < MR; jj++) bstore(d[jj], rout, jj < m ? o0 + (unsigned)(jj * sizeof(T)) : kNoPencil, 0u);
    }
    __syncthreads();   // LDS reused by the next item
}

// FP32 with runs of 8 rows fits 64 VGPRs: eight waves per SIMD (two 1024-lane or four 512-lane
// workgroups per CU); the other variants would spill there
template <typename T, int M, int M2, int MR, int NT>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 && MR == 8) ? 8 : 1) void k_patch_derivs(const PatchJob<T> *__restrict__ jobs,
                                                       const int4 *__restrict__ items, int nitems) {
    extern __shared__ unsigned char smem_raw[];
    T *sp = reinterpret_cast<T *>(smem_raw);
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 w = items[it];
        const PatchJob<T> J = jobs[w.x];
        if (J.kind == 0) strided_item<T, M, NT>(J, w.y, w.z, sp);
        else if (M2 != M && J.kind == 2) strided_item<T, M2, NT>(J, w.y, w.z, sp);
        else contig_item<T, MR, NT>(J, w.y, w.z, sp);
    }
}

// LDS bytes an item of pencil length n needs (the larger of the two schemes)
inline size_t patch_lds(int nmax, int M, int MR, int NT, size_t elem) {
    const int nl = (nmax + M - 1) / M;   // M: the shorter segment length in use
    const size_t a = (size_t)5 * (nmax + nl) + 4 * (size_t)patch_shared(NT) + kPatchTail;
    const size_t b = (size_t)5 * MR * ((nmax + MR - 1) / MR) + kPatchTail;
    return std::max(a, b) * elem;
}

// the variants the autotuner chooses from: strided segment lengths (M, M2: each strided job
// takes the one that wastes fewer lanes to the power-of-two segment count), contiguous run
// length MR and workgroup size
constexpr int kPatchVariants = 4;
constexpr int kVarM[kPatchVariants] = {16, 16, 16, 16};
constexpr int kVarM2[kPatchVariants] = {12, 16, 12, 12};
constexpr int kVarMR[kPatchVariants] = {8, 8, 16, 8};
constexpr int kVarNT[kPatchVariants] = {1024, 1024, 1024, 512};

// ---- synthetic patch data: N(0,1) (SplitMix64 + Box-Muller, the generator of cssx/model.py's
//      normal()) smoothed by [1,2,1]/4 along each axis with the edge value repeated --------------
__device__ __forceinline__ double splitmix_u01(unsigned long long seed, unsigned long long idx) {
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + idx * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E5B5ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_patch_noise(T *out, long n, unsigned long long seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long m = (n + 1) / 2, k = i >> 1;
    const double u1 = splitmix_u01(seed, (unsigned long long)k), u2 = splitmix_u01(seed, (unsigned long long)(m + k));
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    out[i] = (T)((i & 1) ? rad * sin(ang) : rad * cos(ang));
}

// out = [1,2,1]/4 of in along one axis (element stride `stride`, length len), edges repeated
template <typename T>
__global__ __launch_bounds__(256) void k_patch_smooth(const T *__restrict__ in, T *__restrict__ out, long n, long stride,
                                                      int len) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = (int)((i / stride) % len);
    const T c = in[i];
    const T lo = q > 0 ? in[i - stride] : c;
    const T hi = q < len - 1 ? in[i + stride] : c;
    out[i] = T(0.25) * (lo + hi) + T(0.5) * c;
}

}  // namespace cssx

using namespace cssx;

extern "C" int cssx_normal_fill(void *out, int prec, long n, unsigned long long seed, void *stream) {
    if (!out || n < 1 || (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32)) return CSSX_EINVAL;
    const unsigned g = (unsigned)((n + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
    if (prec == CSSX_PREC_F64)
        hipLaunchKernelGGL(k_patch_noise<double>, dim3(g), dim3(256), 0, s, (double *)out, n, seed);
    else
        hipLaunchKernelGGL(k_patch_noise<float>, dim3(g), dim3(256), 0, s, (float *)out, n, seed);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_patch_fill(void *out, void *tmp, int prec, int nr, int nth, int nphi, unsigned long long seed,
                               void *stream) {
    if (!out || !tmp || nr < 1 || nth < 1 || nphi < 1 || (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    const long n = (long)nr * nth * nphi;
    const unsigned g = (unsigned)((n + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
    if (prec == CSSX_PREC_F64) {
        double *o = (double *)out, *t = (double *)tmp;
        hipLaunchKernelGGL(k_patch_noise<double>, dim3(g), dim3(256), 0, s, t, n, seed);
        hipLaunchKernelGGL(k_patch_smooth<double>, dim3(g), dim3(256), 0, s, t, o, n, 1L, nth);
        hipLaunchKernelGGL(k_patch_smooth<double>, dim3(g), dim3(256), 0, s, o, t, n, (long)nth, nphi);
        hipLaunchKernelGGL(k_patch_smooth<double>, dim3(g), dim3(256), 0, s, t, o, n, (long)nth * nphi, nr);
    } else {
        float *o = (float *)out, *t = (float *)tmp;
        hipLaunchKernelGGL(k_patch_noise<float>, dim3(g), dim3(256), 0, s, t, n, seed);
        hipLaunchKernelGGL(k_patch_smooth<float>, dim3(g), dim3(256), 0, s, t, o, n, 1L, nth);
        hipLaunchKernelGGL(k_patch_smooth<float>, dim3(g), dim3(256), 0, s, o, t, n, (long)nth, nphi);
        hipLaunchKernelGGL(k_patch_smooth<float>, dim3(g), dim3(256), 0, s, t, o, n, (long)nth * nphi, nr);
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

struct cssx_patchset {
    int prec = 0, npatch = 0, nmax = 0, best = 0;
    long nitems[kPatchVariants] = {};
    void *jobs[kPatchVariants] = {};
    int4 *items[kPatchVariants] = {};
};

namespace {

unsigned next_pow2(unsigned v) {
    unsigned p = 1;
    while (p < v) p <<= 1;
    return p;
}

template <typename T>
int build_variant(cssx_patchset *ps, const cssx_patch *pt, int var) {
    const int M = kVarM[var], M2 = kVarM2[var], MR = kVarMR[var], NT = kVarNT[var];
    // lanes a strided job of n rows and na pencils per plane occupies with segments of m rows
    auto strided_cost = [&](int n, long na, int m) {
        const long sm = (long)next_pow2((unsigned)((n + m - 1) / m)), p = NT / sm;
        return (na + p - 1) / p * (long)NT * m;
    };
    std::vector<PatchJob<T>> jobs;
    std::vector<int4> items;
    const int prec = sizeof(T) == 8 ? CSSX_PREC_F64 : CSSX_PREC_F32;
    for (int i = 0; i < ps->npatch; i++) {
        const cssx_patch &q = pt[i];
        const long sth = 1, sph = q.nth, sr = (long)q.nth * q.nphi;
        void *outs[3] = {q.d_r, q.d_th, q.d_ph};
        for (int ax = 0; ax < 3; ax++) {
            if (!outs[ax]) continue;
            PatchJob<T> J{};
            J.in = (const T *)q.in;
            J.out = (T *)outs[ax];
            J.bdry = q.bdry;
            if (ax == 0) { J.n = q.nr; J.na = q.nth; J.nb = q.nphi; J.sa = sth; J.sb = sph; J.srow = sr; J.scale = (T)q.dri; }
            else if (ax == 1) { J.n = q.nth; J.na = q.nphi; J.nb = q.nr; J.sa = sph; J.sb = sr; J.srow = sth; J.scale = (T)q.dti; }
            else { J.n = q.nphi; J.na = q.nth; J.nb = q.nr; J.sa = sth; J.sb = sr; J.srow = sph; J.scale = (T)q.dpi; }
            J.kind = ax == 1 ? 1 : 0;
            const int Mj = (J.kind == 0 && strided_cost(J.n, J.na, M2) < strided_cost(J.n, J.na, M)) ? M2 : M;
            if (Mj != M) J.kind = 2;
            // the kernels' buffer ranges (span_rsrc): offsets up to kNoPencil + (n + 20) rows
            // stay below 2^32, and contiguous items walk unit-stride rows
            if (J.kind != 1 && (double)(J.n + 20) * (double)J.srow * sizeof(T) > 2147483648.0) return CSSX_EINVAL;
            if (J.kind == 1 && (J.srow != 1 || (double)J.na * (double)J.sa * sizeof(T) > 2147483648.0))
                return CSSX_EINVAL;
            PlanKey key{J.n, 1, END_WALL, END_WALL, q.bdry};
            key.variant = J.kind == 1 ? 1 : 2;
            key.seg = J.kind == 1 ? MR : Mj;
            const void *plan = nullptr;
            int rc = get_plan(key, prec, &plan);
            if (rc) return rc;
            J.plan = (const T *)plan;
            long per_b;
            if (J.kind != 1) {
                J.lanes = (int)next_pow2((unsigned)((J.n + Mj - 1) / Mj));
                per_b = (J.na + (NT / J.lanes) - 1) / (NT / J.lanes);
            } else {
                J.lanes = (int)next_pow2((unsigned)((J.n + MR - 1) / MR));
                const int ppw = 64 / J.lanes;
                const long pw = (long)(NT / 64) * ppw;
                per_b = (J.na + pw - 1) / pw;
            }
            const int jid = (int)jobs.size();
            jobs.push_back(J);
            for (int b = 0; b < J.nb; b++)
                for (long x = 0; x < per_b; x++) items.push_back(make_int4(jid, (int)x, b, 0));
        }
    }
    if (items.size() > (size_t)INT32_MAX) return CSSX_EINVAL;
    ps->nitems[var] = (long)items.size();
    if (hipMalloc(&ps->jobs[var], jobs.size() * sizeof(PatchJob<T>)) != hipSuccess ||
        hipMalloc((void **)&ps->items[var], std::max<size_t>(1, items.size()) * sizeof(int4)) != hipSuccess)
        return CSSX_ENOMEM;
    if (hipMemcpy(ps->jobs[var], jobs.data(), jobs.size() * sizeof(PatchJob<T>), hipMemcpyHostToDevice) != hipSuccess ||
        (items.size() && hipMemcpy(ps->items[var], items.data(), items.size() * sizeof(int4),
                                   hipMemcpyHostToDevice) != hipSuccess))
        return CSSX_EHIP;
    return CSSX_OK;
}

template <typename T, int M, int M2, int MR, int NT>
hipError_t launch_variant(const cssx_patchset *ps, int var, hipStream_t s) {
    const size_t shm = patch_lds(ps->nmax, std::min(M, M2), MR, NT, sizeof(T));
    const void *fn = (const void *)k_patch_derivs<T, M, M2, MR, NT>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    int dev = 0, ncu = 256, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent grid: as many workgroups as can be resident at once, so none waits for another
    // to drain the item list
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, shm);
    if (e != hipSuccess) return e;
    const long g = std::min<long>(ps->nitems[var], (long)ncu * std::max(1, per_cu));
    if (g < 1) return hipSuccess;
    hipLaunchKernelGGL((k_patch_derivs<T, M, M2, MR, NT>), dim3((unsigned)g), dim3(NT), shm, s,
                       (const PatchJob<T> *)ps->jobs[var], (const int4 *)ps->items[var], (int)ps->nitems[var]);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_any(const cssx_patchset *ps, int var, hipStream_t s) {
    switch (var) {
    case 0: return launch_variant<T, 16, 12, 8, 1024>(ps, var, s);
    case 1: return launch_variant<T, 16, 16, 8, 1024>(ps, var, s);
    case 2: return launch_variant<T, 16, 12, 16, 1024>(ps, var, s);
    default: return launch_variant<T, 16, 12, 8, 512>(ps, var, s);
    }
}

}  // namespace

extern "C" void cssx_patchset_destroy(cssx_patchset *ps) {
    if (!ps) return;
    (void)hipDeviceSynchronize();
    for (int v = 0; v < kPatchVariants; v++) {
        if (ps->jobs[v]) (void)hipFree(ps->jobs[v]);
        if (ps->items[v]) (void)hipFree(ps->items[v]);
    }
    delete ps;
}

extern "C" int cssx_patchset_create(const cssx_patch *patches, int npatch, int prec, cssx_patchset **out) {
    if (!patches || npatch < 1 || !out || (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32)) return CSSX_EINVAL;
    *out = nullptr;
    int nmax = 0;
    for (int i = 0; i < npatch; i++) {
        const cssx_patch &q = patches[i];
        if (!q.in || q.nr < 8 || q.nth < 8 || q.nphi < 8 || q.nr > 512 || q.nth > 512 || q.nphi > 512 || q.bdry < 0 ||
            q.bdry > 4)
            return CSSX_EINVAL;
        const double bytes = (double)q.nr * q.nth * q.nphi * (prec == CSSX_PREC_F64 ? 8.0 : 4.0);
        if (bytes >= 4294967295.0) return CSSX_EINVAL;   // 32-bit buffer offsets
        nmax = std::max({nmax, q.nr, q.nth, q.nphi});
    }
    cssx_patchset *ps = new (std::nothrow) cssx_patchset;
    if (!ps) return CSSX_ENOMEM;
    ps->prec = prec;
    ps->npatch = npatch;
    ps->nmax = nmax;
    for (int v = 0; v < kPatchVariants; v++) {
        int rc = prec == CSSX_PREC_F64 ? build_variant<double>(ps, patches, v) : build_variant<float>(ps, patches, v);
        if (rc) {
            cssx_patchset_destroy(ps);
            return rc;
        }
    }
    *out = ps;
    return CSSX_OK;
}

extern "C" int cssx_patchset_run(cssx_patchset *ps, int variant, void *stream) {
    if (!ps || variant < -1 || variant >= kPatchVariants) return CSSX_EINVAL;
    const int v = variant < 0 ? ps->best : variant;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ps->prec == CSSX_PREC_F64 ? launch_any<double>(ps, v, s) : launch_any<float>(ps, v, s);
    return e == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_patchset_autotune(cssx_patchset *ps, int reps, void *stream, int *best, float *ms_out) {
    if (!ps || reps < 1) return CSSX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return CSSX_EHIP;
    float bestms = 0.0f;
    int bv = 0;
    int rc = CSSX_OK;
    for (int v = 0; v < kPatchVariants && !rc; v++) {
        if ((rc = cssx_patchset_run(ps, v, stream))) break;   // warm
        if (hipEventRecord(e0, s) != hipSuccess) { rc = CSSX_EHIP; break; }
        for (int r = 0; r < reps && !rc; r++) rc = cssx_patchset_run(ps, v, stream);
        float ms = 0.0f;
        if (rc || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = rc ? rc : CSSX_EHIP;
            break;
        }
        ms /= reps;
        if (ms_out) ms_out[v] = ms;
        if (v == 0 || ms < bestms) {
            bestms = ms;
            bv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    ps->best = bv;
    if (best) *best = bv;
    return CSSX_OK;
}
