// Learned subgrid closure for a coarse grid (no counterpart in the reference; it stands beside
// diffusion.f90's SLD as the alternative subgrid model): a 216-128-128-8 GELU MLP on the
// 3x3x3 stencils of the 8 state fields, trained and evaluated by the kernels in this file.
//
//   * k_box_filter      fine state / tendency -> coarse, mean over f^3 blocks (FP32 out)
//   * k_features        coarse state -> 216 standardized BF16 features per point (the 27
//                       neighbours of each field; rho and s as fluctuations about the
//                       hydrostatic background; walls clamp, phi wraps)
//   * k_gemm_bf16       C = A B on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
//                       accumulation) with the layer epilogues fused: bias + GELU (forward),
//                       GELU' (backward), split-K atomic accumulation (weight gradients)
//   * k_mse, k_colsum, k_adam   loss gradient, bias gradients, Adam (FP32 master weights,
//                       BF16 copies for the matrix cores)
//   * k_add_closure     un-standardized network output -> the FP32 tendency that the RHS adds
//
// Every matrix here is row-major.  A layer is Z = X W + b, H = GELU(Z), with W stored as
// (fan_in, fan_out).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <new>
#include <set>
#include <utility>
#include <vector>

#include "../../include/cssx.h"
#include "cssx_bufio.hpp"
#include "cssx_internal.hpp"

typedef __bf16 cssx_bf16;
typedef __bf16 cssx_bf16x8 __attribute__((ext_vector_type(8)));
typedef float cssx_f32x4 __attribute__((ext_vector_type(4)));

namespace cssx {

constexpr int kFeat = 216;   // 8 fields x 27 neighbours
constexpr int kOut = 8;

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, formula 7.1.26
// (|error| <= 1.5e-7, below what the BF16 rounding of every activation keeps):
// erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2), t = 1 / (1 + p x), x >= 0,
// odd in x.  About a dozen instructions (one v_rcp, one v_exp) against ~30 for erff; training
// and inference share it, so the network is trained with the activation it runs with.
__device__ __forceinline__ float exp_neg_sq(float x) { return __builtin_amdgcn_exp2f(-1.4426950408889634f * x * x); }
__device__ __forceinline__ float erf_as(float x) {
    const float ax = __builtin_fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    return __builtin_copysignf(1.0f - poly * exp_neg_sq(ax), x);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) + z * 0.39894228040143268f * exp_neg_sq(0.70710678118654752f * z);
}

// ---- box filter -------------------------------------------------------------------------------
template <typename Tin>
__global__ __launch_bounds__(256) void k_box_filter(const Tin *__restrict__ in, float *__restrict__ out, int nv,
                                                    int nr, int nphi, int nth, int f) {
    const int cr = nr / f, cp = nphi / f, ct = nth / f;
    const long Nc = (long)cr * cp * ct;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc * nv) return;
    const int iv = (int)(i / Nc);
    long j = i - (long)iv * Nc;
    const int it = (int)(j % ct);
    j /= ct;
    const int ip = (int)(j % cp);
    const int ir = (int)(j / cp);
    const Tin *base = in + (long)iv * nr * nphi * nth;
    double s = 0.0;
    for (int a = 0; a < f; a++)
        for (int b = 0; b < f; b++) {
            const Tin *row = base + ((long)(ir * f + a) * nphi + (ip * f + b)) * nth + (long)it * f;
            for (int c = 0; c < f; c++) s += (double)row[c];
        }
    out[i] = (float)(s / ((double)f * f * f));
}

// ---- features ---------------------------------------------------------------------------------
// thread per (point, feature): feature q = field*27 + (dr+1)*9 + (dp+1)*3 + (dt+1)
template <typename T>
__global__ __launch_bounds__(256) void k_features(const T *__restrict__ st, long N, int nr, int nphi, int nth,
                                                  int fluct, const T *__restrict__ rbar, const T *__restrict__ sbar,
                                                  long row0, long rows, const float *__restrict__ mean,
                                                  const float *__restrict__ istd, cssx_bf16 *__restrict__ xb,
                                                  float *__restrict__ xf) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * kFeat) return;
    const long pt = row0 + e / kFeat;
    const int q = (int)(e - (e / kFeat) * kFeat);
    const int iv = q / 27, nb = q - iv * 27;
    const int dr = nb / 9 - 1, dp = (nb / 3) % 3 - 1, dt = nb % 3 - 1;
    const int it = (int)(pt % nth);
    const long rest = pt / nth;
    const int ip = (int)(rest % nphi);
    const int ir = (int)(rest / nphi);
    const int jr = min(max(ir + dr, 0), nr - 1);
    const int jt = min(max(it + dt, 0), nth - 1);
    int jp = ip + dp;
    jp = jp < 0 ? jp + nphi : (jp >= nphi ? jp - nphi : jp);
    float v = (float)st[(long)iv * N + ((long)jr * nphi + jp) * nth + jt];
    if (!fluct) {
        if (iv == 0) v = (float)((double)st[(long)iv * N + ((long)jr * nphi + jp) * nth + jt] - (double)rbar[jr]);
        if (iv == 4) v = (float)((double)st[(long)iv * N + ((long)jr * nphi + jp) * nth + jt] - (double)sbar[jr]);
    }
    if (xf) xf[e] = v;
    if (xb) xb[e] = (cssx_bf16)((v - mean[q]) * istd[q]);
}

// ---- GEMM on the BF16 matrix cores -------------------------------------------------------------
enum Epi : int {
    EPI_STORE = 0,       // C = acc (+ bias)
    EPI_ATOMIC = 1,      // C += acc (split-K partial, FP32 atomics)
    EPI_BIAS_GELU = 2,   // C = Z = acc + bias (FP32, may be null), H = bf16(GELU(Z))
    EPI_GELU_BWD = 3,    // C = dZ = acc * GELU'(Zin) (FP32), H = bf16(dZ)
};

struct GemmArgs {
    const cssx_bf16 *A;
    long sam, sak;       // A(m, k) = A[m*sam + k*sak]
    const cssx_bf16 *B;
    long sbk, sbn;       // B(k, n) = B[k*sbk + n*sbn]
    int M, N, K;
    int kchunk;          // K range of one blockIdx.z
    int epi;
    const float *bias;
    float *C;
    long ldc;
    cssx_bf16 *H;
    long ldh;
    const float *Zin;
    long ldz;
};

// A workgroup of 4 wavefronts computes a 32x32 tile of C, each wavefront one 16x16 block with
// v_mfma_f32_16x16x32_bf16: lane l supplies row (l mod 16) of A and column (l mod 16) of B at
// k = 8*(l/16) .. 8*(l/16)+7 of every 32-wide K step, and receives C rows 4*(l/16) .. +3 of
// column l mod 16.
__global__ __launch_bounds__(256) void k_gemm_bf16(const GemmArgs g) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.x * 32 + (w & 1) * 16, n0 = blockIdx.y * 32 + (w >> 1) * 16;
    const int li = lane & 15, kb = (lane >> 4) * 8;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const int am = m0 + li, bn = n0 + li;
    const bool aok = am < g.M, bok = bn < g.N;
    const cssx_bf16 *Ar = g.A + (long)(aok ? am : 0) * g.sam;
    const cssx_bf16 *Bc = g.B + (long)(bok ? bn : 0) * g.sbn;
    const cssx_bf16 zero = (cssx_bf16)0.0f;
    cssx_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        cssx_bf16x8 a, b;
        const int kk = k0 + kb;
        if (g.sak == 1 && kk + 8 <= kend && aok && ((reinterpret_cast<uintptr_t>(Ar + kk) & 15) == 0)) {
            a = *reinterpret_cast<const cssx_bf16x8 *>(Ar + kk);
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) a[e] = (aok && kk + e < kend) ? Ar[(long)(kk + e) * g.sak] : zero;
        }
        if (g.sbk == 1 && kk + 8 <= kend && bok && ((reinterpret_cast<uintptr_t>(Bc + kk) & 15) == 0)) {
            b = *reinterpret_cast<const cssx_bf16x8 *>(Bc + kk);
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) b[e] = (bok && kk + e < kend) ? Bc[(long)(kk + e) * g.sbk] : zero;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + li;
    if (n >= g.N) return;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = m0 + (lane >> 4) * 4 + r;
        if (m >= g.M) continue;
        float v = acc[r];
        switch (g.epi) {
        case EPI_STORE:
            if (g.bias) v += g.bias[n];
            g.C[(long)m * g.ldc + n] = v;
            break;
        case EPI_ATOMIC:
            atomicAdd(&g.C[(long)m * g.ldc + n], v);
            break;
        case EPI_BIAS_GELU:
            if (g.bias) v += g.bias[n];
            if (g.C) g.C[(long)m * g.ldc + n] = v;
            g.H[(long)m * g.ldh + n] = (cssx_bf16)gelu(v);
            break;
        default:   // EPI_GELU_BWD
            v *= gelu_grad(g.Zin[(long)m * g.ldz + n]);
            g.C[(long)m * g.ldc + n] = v;
            g.H[(long)m * g.ldh + n] = (cssx_bf16)v;
            break;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill(float *p, long n, float v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// loss = mean (Y-T)^2 over B x nout; dY = 2 (Y-T) / (B nout) as FP32 and BF16
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ Tg, long n,
                                             float *__restrict__ dY, cssx_bf16 *__restrict__ dYb, float *loss) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.f;
    if (i < n) {
        const float d = Y[i] - Tg[i];
        const float g = 2.0f * d / (float)n;
        dY[i] = g;
        dYb[i] = (cssx_bf16)g;
        d2 = d * d;
    }
    for (int o = 32; o > 0; o >>= 1) d2 += __shfl_xor(d2, o, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) / (float)n);
}

// db[c] += sum over a chunk of rows of X[r][c]
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ X, int rows, int cols, int chunk,
                                                float *__restrict__ db) {
    const int r0 = blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        float s = 0.f;
        for (int r = r0; r < r1; r++) s += X[(long)r * cols + c];
        atomicAdd(&db[c], s);
    }
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, cssx_bf16 *__restrict__ pb, long n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = w;
    pb[i] = (cssx_bf16)w;
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ p, cssx_bf16 *__restrict__ pb, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = (cssx_bf16)p[i];
}

// ext[iv*N + row0 + r] = Y[r][iv] * ystd[iv] + ymean[iv]
__global__ __launch_bounds__(256) void k_add_closure(const float *__restrict__ Y, long rows, long row0, long N,
                                                     const float *__restrict__ ymean, const float *__restrict__ ystd,
                                                     float *__restrict__ ext) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * kOut) return;
    const long r = e / kOut;
    const int iv = (int)(e - r * kOut);
    ext[(long)iv * N + row0 + r] = Y[e] * ystd[iv] + ymean[iv];
}

// standardized row-major targets: out[r][iv] = (fine[iv*N + r] - coarse[iv*N + r] - ymean) / ystd
template <typename T>
__global__ __launch_bounds__(256) void k_targets(const float *__restrict__ fine, const T *__restrict__ coarse, long N,
                                                 const float *__restrict__ ymean, const float *__restrict__ ystd,
                                                 float *__restrict__ out, float *__restrict__ raw) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * kOut) return;
    const long r = e / kOut;
    const int iv = (int)(e - r * kOut);
    const float d = (float)((double)fine[(long)iv * N + r] - (double)coarse[(long)iv * N + r]);
    if (raw) raw[e] = d;
    if (out) out[e] = (d - ymean[iv]) / ystd[iv];
}

// column sums and sums of squares of a row-major (rows x cols) matrix, FP64 accumulation
__global__ __launch_bounds__(256) void k_colstats(const float *__restrict__ X, long rows, int cols, int chunk,
                                                  double *__restrict__ acc) {
    const long r0 = (long)blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        double s = 0.0, q = 0.0;
        for (long r = r0; r < r1; r++) {
            const double v = X[r * cols + c];
            s += v;
            q += v * v;
        }
        atomicAdd(&acc[c], s);
        atomicAdd(&acc[cols + c], q);
    }
}

// per-field sums and sums of squares of a field-major state, rho and s about the background
// when fluct == 0
template <typename T>
__global__ __launch_bounds__(256) void k_field_stats(const T *__restrict__ st, int nv, long N, int plane, int fluct,
                                                     const T *__restrict__ rbar, const T *__restrict__ sbar,
                                                     double *__restrict__ acc) {
    const int iv = blockIdx.y;
    double s = 0.0, q = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double v = (double)st[(long)iv * N + i];
        if (!fluct && iv == 0) v -= (double)rbar[i / plane];
        if (!fluct && iv == 4) v -= (double)sbar[i / plane];
        s += v;
        q += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        q += __shfl_xor(q, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[iv], s);
        atomicAdd(&acc[nv + iv], q);
    }
}

// dst[i] = src[idx[i]] for rows of rowbytes bytes (a multiple of 4)
__global__ __launch_bounds__(256) void k_gather_rows(const unsigned *__restrict__ src, const long *__restrict__ idx,
                                                     long n, int words, unsigned *__restrict__ dst) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * words) return;
    const long r = e / words;
    const int w = (int)(e - r * words);
    dst[e] = src[idx[r] * words + w];
}

// (x - mean[c]) / std[c] of a row-major (rows x cols) FP32 matrix, to BF16 and/or FP32
__global__ __launch_bounds__(256) void k_standardize(const float *__restrict__ x, long rows, int cols,
                                                     const float *__restrict__ mean, const float *__restrict__ sd,
                                                     cssx_bf16 *__restrict__ ob, float *__restrict__ of) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int c = (int)(e % cols);
    const float v = (x[e] - mean[c]) / sd[c];
    if (ob) ob[e] = (cssx_bf16)v;
    if (of) of[e] = v;
}

// ---- fused inference: features -> 3 layers -> tendency, activations never leave registers ------
// Computed transposed (Z^T = W^T X^T): the samples are the MFMA's B columns, so a layer's
// accumulators (lane l: units 16*nb + 4*(l/16) + r of sample l mod 16) are, after bias+GELU
// and a BF16 cast, already the next layer's B operand, provided the next weights' K slots are
// permuted to match (slot 8g+e of K step ks <-> unit 16*(2ks + e/4) + 4g + e mod 4).  The
// weights are packed once per call into that layout and held in LDS by every workgroup.
constexpr int kH = 128;                 // hidden width of the fused kernel
constexpr int kKp1 = 224;               // 216 features padded to 7 K steps of 32
constexpr int kS1 = kKp1 + 8, kS2 = kH + 8;   // LDS row strides (BF16), padded against bank conflicts
constexpr int kPackBf = kH * kS1 + kH * kS2 + 16 * kS2;
constexpr int kPackF = 2 * kH + 16;     // biases b1, b2, b3 (padded to 16), FP32
constexpr size_t kPackBytes = (size_t)kPackBf * 2 + (size_t)kPackF * 4;

__device__ __forceinline__ int perm_unit(int ks, int slot) {   // slot = 8g + e
    const int g = slot >> 3, e = slot & 7;
    return 16 * (2 * ks + (e >> 2)) + 4 * g + (e & 3);
}

__global__ __launch_bounds__(256) void k_pack_infer(const float *__restrict__ W, long oW1, long ob1, long oW2, long ob2,
                                                    long oW3, long ob3, cssx_bf16 *__restrict__ pk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const cssx_bf16 zero = (cssx_bf16)0.0f;
    if (i < (long)kH * kS1) {
        // layer-1 K slots are neighbour-major: slot 8*nb + field <-> feature field*27 + nb, so a
        // lane's 8 slots of a K step are one neighbour of all 8 fields (one address computation)
        const int n = (int)(i / kS1), k = (int)(i % kS1);
        const int nbh = k >> 3, iv = k & 7;
        pk[i] = (k < kKp1 && nbh < 27) ? (cssx_bf16)W[oW1 + (long)(iv * 27 + nbh) * kH + n] : zero;
    } else if (i < (long)kH * kS1 + (long)kH * kS2) {
        const long j = i - (long)kH * kS1;
        const int o = (int)(j / kS2), k = (int)(j % kS2);
        pk[i] = k < kH ? (cssx_bf16)W[oW2 + (long)perm_unit(k >> 5, k & 31) * kH + o] : zero;
    } else if (i < kPackBf) {
        const long j = i - (long)kH * kS1 - (long)kH * kS2;
        const int o = (int)(j / kS2), k = (int)(j % kS2);
        pk[i] = (k < kH && o < kOut) ? (cssx_bf16)W[oW3 + (long)perm_unit(k >> 5, k & 31) * kOut + o] : zero;
    } else if (i < kPackBf + kPackF) {
        const int j = (int)(i - kPackBf);
        float *bf = reinterpret_cast<float *>(pk + kPackBf);
        bf[j] = j < kH ? W[ob1 + j] : j < 2 * kH ? W[ob2 + j - kH] : (j - 2 * kH < kOut ? W[ob3 + j - 2 * kH] : 0.0f);
    }
}

template <typename T>
__global__ __launch_bounds__(512) void k_closure_infer(const T *__restrict__ st, long N, int nr, int nphi, int nth,
                                                        int fluct, const T *__restrict__ rbar,
                                                        const T *__restrict__ sbar, const float *__restrict__ xmean,
                                                        const float *__restrict__ xistd, const float *__restrict__ ymean,
                                                        const float *__restrict__ ystd,
                                                        const cssx_bf16 *__restrict__ pk, float *__restrict__ ext) {
    extern __shared__ uint4 lds4[];
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(pk);
        for (int i = threadIdx.x; i < (int)(kPackBytes / 16); i += blockDim.x) lds4[i] = src[i];
    }
    __syncthreads();
    const cssx_bf16 *W1T = reinterpret_cast<const cssx_bf16 *>(lds4);
    const cssx_bf16 *W2P = W1T + kH * kS1, *W3P = W2P + kH * kS2;
    const float *b1 = reinterpret_cast<const float *>(W1T + kPackBf), *b2 = b1 + kH, *b3 = b2 + kH;
    const int lane = threadIdx.x & 63, g = lane >> 4, li = lane & 15;
    const long ntiles = (N + 15) / 16;
    const int wpb = blockDim.x >> 6;
    const cssx_bf16 zero = (cssx_bf16)0.0f;
    for (long t = (long)blockIdx.x * wpb + (threadIdx.x >> 6); t < ntiles; t += (long)gridDim.x * wpb) {
        const long pt = t * 16 + li;
        const bool ok = pt < N;
        const long pc = ok ? pt : N - 1;
        const int it = (int)(pc % nth);
        const long rest = pc / nth;
        const int ip = (int)(rest % nphi), ir = (int)(rest / nphi);
        cssx_f32x4 acc[8];
#pragma unroll
        for (int nb = 0; nb < 8; nb++) acc[nb] = cssx_f32x4{0.f, 0.f, 0.f, 0.f};
        // layer 1: the lane's 8 features of each K step, gathered and standardized in place
#pragma unroll 1
        for (int ks = 0; ks < kKp1 / 32; ks++) {
            cssx_bf16x8 b;
            const int nbh = 4 * ks + g;   // this lane's neighbour; its 8 slots are the 8 fields
            if (nbh < 27) {
                const int jr = min(max(ir + nbh / 9 - 1, 0), nr - 1);
                const int jt = min(max(it + nbh % 3 - 1, 0), nth - 1);
                int jp = ip + (nbh / 3) % 3 - 1;
                jp = jp < 0 ? jp + nphi : (jp >= nphi ? jp - nphi : jp);
                const T *p0 = st + ((long)jr * nphi + jp) * nth + jt;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const T raw = p0[(long)e * N];
                    float x;
                    if (!fluct && e == 0) x = (float)((double)raw - (double)rbar[jr]);
                    else if (!fluct && e == 4) x = (float)((double)raw - (double)sbar[jr]);
                    else x = (float)raw;
                    const int q = e * 27 + nbh;
                    b[e] = (cssx_bf16)((x - xmean[q]) * xistd[q]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) b[e] = zero;
            }
#pragma unroll
            for (int nb = 0; nb < 8; nb++) {
                const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W1T + (nb * 16 + li) * kS1 + ks * 32 + 8 * g);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nb], 0, 0, 0);
            }
        }
        cssx_bf16x8 h[4];
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) h[nb >> 1][(nb & 1) * 4 + r] = (cssx_bf16)gelu(acc[nb][r] + b1[nb * 16 + 4 * g + r]);
        // layer 2
#pragma unroll
        for (int nb = 0; nb < 8; nb++) acc[nb] = cssx_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nb = 0; nb < 8; nb++) {
                const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W2P + (nb * 16 + li) * kS2 + ks * 32 + 8 * g);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[ks], acc[nb], 0, 0, 0);
            }
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) h[nb >> 1][(nb & 1) * 4 + r] = (cssx_bf16)gelu(acc[nb][r] + b2[nb * 16 + 4 * g + r]);
        // layer 3 (8 outputs padded to one 16-row block)
        cssx_f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W3P + li * kS2 + ks * 32 + 8 * g);
            y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[ks], y, 0, 0, 0);
        }
        if (ok && g < 2) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = 4 * g + r;
                ext[(long)o * N + pt] = (y[r] + b3[o]) * ystd[o] + ymean[o];
            }
        }
    }
}

// The same network with the tile's stencil footprint staged per wavefront: when nth is a
// multiple of 16, a tile's 16 points share (ir, ip) and cover theta rows it0 .. it0+15, so its
// 27-point stencils read 3 (r) x 3 (phi) x 18 (theta) points of each of the 8 fields.  The wave
// loads those 1296 values with one batch of independent loads (one memory latency per tile
// instead of one per K step), converts them to FP32 (background subtracted as above) into its
// slot of LDS, and the K steps gather their B operands from there; the standardization tables
// sit in LDS beside the packed weights.
constexpr int kFoot = 8 * 3 * 3 * 18;   // footprint values per tile
constexpr int kInferWaves = 8;
constexpr size_t kInferRowsBytes = kPackBytes + 2 * kFeat * sizeof(float) + (size_t)kInferWaves * kFoot * sizeof(float);

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__global__ __launch_bounds__(64 * kInferWaves) void k_closure_infer_rows(
    const T *__restrict__ st, long N, int nr, int nphi, int nth, int fluct, const T *__restrict__ rbar,
    const T *__restrict__ sbar, const float *__restrict__ xmean, const float *__restrict__ xistd,
    const float *__restrict__ ymean, const float *__restrict__ ystd, const cssx_bf16 *__restrict__ pk,
    float *__restrict__ ext) {
    extern __shared__ uint4 lds4[];
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(pk);
        for (int i = threadIdx.x; i < (int)(kPackBytes / 16); i += blockDim.x) lds4[i] = src[i];
    }
    float *xm = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(lds4) + kPackBytes);
    float *xs = xm + kFeat;
    for (int i = threadIdx.x; i < kFeat; i += blockDim.x) {
        xm[i] = xmean[i];
        xs[i] = xistd[i];
    }
    __syncthreads();
    const cssx_bf16 *W1T = reinterpret_cast<const cssx_bf16 *>(lds4);
    const cssx_bf16 *W2P = W1T + kH * kS1, *W3P = W2P + kH * kS2;
    const float *b1 = reinterpret_cast<const float *>(W1T + kPackBf), *b2 = b1 + kH, *b3 = b2 + kH;
    const int lane = threadIdx.x & 63, g = lane >> 4, li = lane & 15, wv = threadIdx.x >> 6;
    float *fp = xs + kFeat + wv * kFoot;
    // footprint slot of the lane; lanes >= 54 are idle (their loads stay in bounds, unused)
    const int dp = min(lane / 18, 2), tt = lane - 18 * (lane / 18);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(st);
    const long ntiles = N / 16;
    const cssx_bf16 zero = (cssx_bf16)0.0f;
    for (long t = (long)blockIdx.x * kInferWaves + wv; t < ntiles; t += (long)gridDim.x * kInferWaves) {
        const long p0 = t * 16;
        const int it0 = (int)(p0 % nth);
        const long rest = p0 / nth;
        const int ip = (int)(rest % nphi), ir = (int)(rest / nphi);
        // footprint: value f = ((e*3 + dr)*3 + dp)*18 + tt of field e at r row ir+dr-1 (clamped),
        // phi ip+dp-1 (periodic), theta it0+tt-1 (clamped).  Lane l < 54 owns dp = l/18 and
        // tt = l%18, and pass k = 3e + dr fills slot 54k + l: the lane's offset is fixed per tile
        // and each pass adds a uniform (scalar) offset, so a load costs no address arithmetic.
        const int jt = min(max(it0 + tt - 1, 0), nth - 1);
        int jp = ip + dp - 1;
        jp = jp < 0 ? jp + nphi : (jp >= nphi ? jp - nphi : jp);
        const unsigned voff = (unsigned)(((long)jp * nth + jt) * (long)sizeof(T));
        float v[24];
#pragma unroll
        for (int k = 0; k < 24; k++) {
            const int e = k / 3, dr = k % 3;
            const int jr = min(max(ir + dr - 1, 0), nr - 1);
            const T raw = bload<T>(rs, voff, (unsigned)(((long)e * N + (long)jr * nphi * nth) * (long)sizeof(T)));
            if (!fluct && e == 0) v[k] = (float)((double)raw - (double)rbar[jr]);
            else if (!fluct && e == 4) v[k] = (float)((double)raw - (double)sbar[jr]);
            else v[k] = (float)raw;
        }
        wave_sync();   // the previous tile's reads of fp are done
        if (lane < 54) {
#pragma unroll
            for (int k = 0; k < 24; k++) fp[54 * k + lane] = v[k];
        }
        wave_sync();
        cssx_f32x4 acc[8];
#pragma unroll
        for (int nb = 0; nb < 8; nb++) acc[nb] = cssx_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kKp1 / 32; ks++) {
            cssx_bf16x8 b;
            const int nbh = 4 * ks + g;   // this lane's neighbour; its 8 slots are the 8 fields
            if (nbh < 27) {
                const int dr = nbh / 9, dp = (nbh / 3) % 3, dt = nbh % 3;
                const float *f0 = fp + (dr * 3 + dp) * 18 + li + dt;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int q = e * 27 + nbh;
                    b[e] = (cssx_bf16)((f0[e * 162] - xm[q]) * xs[q]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) b[e] = zero;
            }
#pragma unroll
            for (int nb = 0; nb < 8; nb++) {
                const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W1T + (nb * 16 + li) * kS1 + ks * 32 + 8 * g);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nb], 0, 0, 0);
            }
        }
        cssx_bf16x8 h[4];
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) h[nb >> 1][(nb & 1) * 4 + r] = (cssx_bf16)gelu(acc[nb][r] + b1[nb * 16 + 4 * g + r]);
#pragma unroll
        for (int nb = 0; nb < 8; nb++) acc[nb] = cssx_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nb = 0; nb < 8; nb++) {
                const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W2P + (nb * 16 + li) * kS2 + ks * 32 + 8 * g);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[ks], acc[nb], 0, 0, 0);
            }
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) h[nb >> 1][(nb & 1) * 4 + r] = (cssx_bf16)gelu(acc[nb][r] + b2[nb * 16 + 4 * g + r]);
        cssx_f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const cssx_bf16x8 a = *reinterpret_cast<const cssx_bf16x8 *>(W3P + li * kS2 + ks * 32 + 8 * g);
            y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[ks], y, 0, 0, 0);
        }
        if (g < 2) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = 4 * g + r;
                ext[(long)o * N + p0 + li] = (y[r] + b3[o]) * ystd[o] + ymean[o];
            }
        }
    }
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace cssx

using namespace cssx;

// ------------------------------------------------------------------------------------------------
// the network object
// ------------------------------------------------------------------------------------------------
struct cssx_mlp {
    int nin, h1, h2, nout, max_rows;
    long np;                        // parameter count
    long oW1, ob1, oW2, ob2, oW3, ob3;
    float *W = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;   // FP32 master, grad, Adam moments
    cssx_bf16 *Wb = nullptr;        // BF16 copy
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *loss = nullptr;
    cssx_bf16 *H1 = nullptr, *H2 = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr, *X = nullptr;
    float *xmean = nullptr, *xistd = nullptr, *ymean = nullptr, *ystd = nullptr;
    cssx_bf16 *pk = nullptr;        // weights packed for the fused inference kernel (h1 = h2 = 128)
    long step = 0;
    void *blk = nullptr;
};

namespace {

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int gemm(const GemmArgs &g0, int ksplit, hipStream_t s) {
    GemmArgs g = g0;
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return CSSX_OK;
    if (ksplit < 1) ksplit = 1;
    g.kchunk = ((g.K + ksplit - 1) / ksplit + 31) / 32 * 32;
    const int kz = (g.K + g.kchunk - 1) / g.kchunk;
    dim3 grid((g.M + 31) / 32, (g.N + 31) / 32, kz);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, g);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

GemmArgs ga(const cssx_bf16 *A, long sam, long sak, const cssx_bf16 *B, long sbk, long sbn, int M, int N, int K,
            int epi) {
    GemmArgs g{};
    g.A = A; g.sam = sam; g.sak = sak; g.B = B; g.sbk = sbk; g.sbn = sbn; g.M = M; g.N = N; g.K = K; g.epi = epi;
    return g;
}

// forward of `rows` standardized rows X (BF16) -> Y (FP32, standardized); keeps Z1/H1/Z2/H2
int mlp_forward(cssx_mlp *n, const cssx_bf16 *X, int rows, bool keep_z, hipStream_t s) {
    GemmArgs g = ga(X, n->nin, 1, n->Wb + n->oW1, n->h1, 1, rows, n->h1, n->nin, EPI_BIAS_GELU);
    g.bias = n->W + n->ob1; g.C = keep_z ? n->Z1 : nullptr; g.ldc = n->h1; g.H = n->H1; g.ldh = n->h1;
    int rc = gemm(g, 1, s);
    if (rc) return rc;
    g = ga(n->H1, n->h1, 1, n->Wb + n->oW2, n->h2, 1, rows, n->h2, n->h1, EPI_BIAS_GELU);
    g.bias = n->W + n->ob2; g.C = keep_z ? n->Z2 : nullptr; g.ldc = n->h2; g.H = n->H2; g.ldh = n->h2;
    if ((rc = gemm(g, 1, s))) return rc;
    g = ga(n->H2, n->h2, 1, n->Wb + n->oW3, n->nout, 1, rows, n->nout, n->h2, EPI_STORE);
    g.bias = n->W + n->ob3; g.C = n->Y; g.ldc = n->nout;
    return gemm(g, 1, s);
}

// gradients of the mean squared error of the last forward into n->G (zeroed first)
int mlp_backward(cssx_mlp *n, const cssx_bf16 *X, const float *Tg, int rows, hipStream_t s) {
    const long ny = (long)rows * n->nout;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n->np)), dim3(256), 0, s, n->G, n->np, 0.0f);
    hipLaunchKernelGGL(k_mse, dim3(nblk(ny)), dim3(256), 0, s, n->Y, Tg, ny, n->dY, n->dYb, n->loss);
    const int ks = max(1, rows / 512);   // split the batch reduction of the weight gradients
    int rc;
    // dW3 = H2^T dY, db3 = colsum dY
    GemmArgs g = ga(n->H2, 1, n->h2, n->dYb, n->nout, 1, n->h2, n->nout, rows, EPI_ATOMIC);
    g.C = n->G + n->oW3; g.ldc = n->nout;
    if ((rc = gemm(g, ks, s))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(rows, 512)), dim3(256), 0, s, n->dY, rows, n->nout, 512, n->G + n->ob3);
    // dZ2 = (dY W3^T) * GELU'(Z2)
    g = ga(n->dYb, n->nout, 1, n->Wb + n->oW3, 1, n->nout, rows, n->h2, n->nout, EPI_GELU_BWD);
    g.C = n->dZ2; g.ldc = n->h2; g.H = n->dZ2b; g.ldh = n->h2; g.Zin = n->Z2; g.ldz = n->h2;
    if ((rc = gemm(g, 1, s))) return rc;
    // dW2 = H1^T dZ2, db2
    g = ga(n->H1, 1, n->h1, n->dZ2b, n->h2, 1, n->h1, n->h2, rows, EPI_ATOMIC);
    g.C = n->G + n->oW2; g.ldc = n->h2;
    if ((rc = gemm(g, ks, s))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(rows, 512)), dim3(256), 0, s, n->dZ2, rows, n->h2, 512, n->G + n->ob2);
    // dZ1 = (dZ2 W2^T) * GELU'(Z1)
    g = ga(n->dZ2b, n->h2, 1, n->Wb + n->oW2, 1, n->h2, rows, n->h1, n->h2, EPI_GELU_BWD);
    g.C = n->dZ1; g.ldc = n->h1; g.H = n->dZ1b; g.ldh = n->h1; g.Zin = n->Z1; g.ldz = n->h1;
    if ((rc = gemm(g, 1, s))) return rc;
    // dW1 = X^T dZ1, db1
    g = ga(X, 1, n->nin, n->dZ1b, n->h1, 1, n->nin, n->h1, rows, EPI_ATOMIC);
    g.C = n->G + n->oW1; g.ldc = n->h1;
    if ((rc = gemm(g, ks, s))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(rows, 512)), dim3(256), 0, s, n->dZ1, rows, n->h1, 512, n->G + n->ob1);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

template <typename T>
int features_impl(cssx_mlp *n, const T *st, long N, int nr, int nphi, int nth, int fluct, const T *rbar,
                  const T *sbar, long row0, long rows, cssx_bf16 *xb, float *xf, hipStream_t s) {
    hipLaunchKernelGGL(k_features<T>, dim3(nblk(rows * kFeat)), dim3(256), 0, s, st, N, nr, nphi, nth, fluct, rbar,
                       sbar, row0, rows, n ? n->xmean : nullptr, n ? n->xistd : nullptr, xb, xf);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

}  // namespace

namespace cssx {

// The LDS limit of a fused inference kernel is raised once per (kernel, device), and the CU
// count read once per device: a closure tendency runs once per RHS evaluation, and these host
// calls would otherwise leave the GPU idle between the weight packing and the inference.
int fused_setup(const void *fn, size_t lds, int *ncu) {
    static std::mutex mu;
    static std::set<std::pair<const void *, int>> raised;
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CSSX_EHIP;
    std::lock_guard<std::mutex> lk(mu);
    if (!raised.count({fn, dev})) {
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return CSSX_EHIP;
        raised.insert({fn, dev});
    }
    int c = dev >= 0 && dev < 64 ? cus[dev] : 0;
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        if (dev >= 0 && dev < 64) cus[dev] = c;
    }
    *ncu = c;
    return CSSX_OK;
}

// closure tendency of a whole state: ext (8 fields, FP32) = network(features(st)), in chunks
// of the network's row capacity
int closure_tendency(cssx_mlp *n, int prec, const void *st, int fluct, const void *rbar, const void *sbar, int nr,
                     int nphi, int nth, float *ext, hipStream_t s) {
    const long N = (long)nr * nphi * nth;
    if (n->h1 == kH && n->h2 == kH) {
        const long tot = kPackBf + kPackF;
        hipLaunchKernelGGL(k_pack_infer, dim3(nblk(tot)), dim3(256), 0, s, n->W, n->oW1, n->ob1, n->oW2, n->ob2,
                           n->oW3, n->ob3, n->pk);
        int ncu = 256;
        const long tiles = (N + 15) / 16;
        if (nth % 16 == 0 && 8.0 * (double)N * (prec == CSSX_PREC_F64 ? 8.0 : 4.0) < 4294967295.0) {
            // tiles of 16 points inside one theta row (and 32-bit byte offsets over the 8 fields): the
            // footprint-staged kernel
            const void *fr = prec == CSSX_PREC_F64 ? (const void *)k_closure_infer_rows<double>
                                                   : (const void *)k_closure_infer_rows<float>;
            if (int rc = fused_setup(fr, kInferRowsBytes, &ncu)) return rc;
            const unsigned grid = (unsigned)std::min<long>((long)ncu, (tiles + kInferWaves - 1) / kInferWaves);
            if (prec == CSSX_PREC_F64)
                hipLaunchKernelGGL(k_closure_infer_rows<double>, dim3(grid), dim3(64 * kInferWaves), kInferRowsBytes, s,
                                   (const double *)st, N, nr, nphi, nth, fluct, (const double *)rbar,
                                   (const double *)sbar, n->xmean, n->xistd, n->ymean, n->ystd, n->pk, ext);
            else
                hipLaunchKernelGGL(k_closure_infer_rows<float>, dim3(grid), dim3(64 * kInferWaves), kInferRowsBytes, s,
                                   (const float *)st, N, nr, nphi, nth, fluct, (const float *)rbar,
                                   (const float *)sbar, n->xmean, n->xistd, n->ymean, n->ystd, n->pk, ext);
            return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
        }
        const void *fn = prec == CSSX_PREC_F64 ? (const void *)k_closure_infer<double> : (const void *)k_closure_infer<float>;
        if (int rc = fused_setup(fn, kPackBytes, &ncu)) return rc;
        const unsigned grid = (unsigned)std::min<long>((long)ncu, (tiles + 7) / 8);
        if (prec == CSSX_PREC_F64)
            hipLaunchKernelGGL(k_closure_infer<double>, dim3(grid), dim3(512), kPackBytes, s, (const double *)st, N, nr,
                               nphi, nth, fluct, (const double *)rbar, (const double *)sbar, n->xmean, n->xistd,
                               n->ymean, n->ystd, n->pk, ext);
        else
            hipLaunchKernelGGL(k_closure_infer<float>, dim3(grid), dim3(512), kPackBytes, s, (const float *)st, N, nr,
                               nphi, nth, fluct, (const float *)rbar, (const float *)sbar, n->xmean, n->xistd,
                               n->ymean, n->ystd, n->pk, ext);
        return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
    }
    for (long r0 = 0; r0 < N; r0 += n->max_rows) {
        const long rows = min((long)n->max_rows, N - r0);
        int rc = prec == CSSX_PREC_F64
                     ? features_impl<double>(n, (const double *)st, N, nr, nphi, nth, fluct, (const double *)rbar,
                                             (const double *)sbar, r0, rows, n->X, nullptr, s)
                     : features_impl<float>(n, (const float *)st, N, nr, nphi, nth, fluct, (const float *)rbar,
                                            (const float *)sbar, r0, rows, n->X, nullptr, s);
        if (rc) return rc;
        if ((rc = mlp_forward(n, n->X, (int)rows, false, s))) return rc;
        hipLaunchKernelGGL(k_add_closure, dim3(nblk(rows * kOut)), dim3(256), 0, s, n->Y, rows, r0, N, n->ymean,
                           n->ystd, ext);
    }
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

}  // namespace cssx

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
extern "C" int cssx_mlp_create(int nin, int h1, int h2, int nout, int max_rows, unsigned long long seed,
                               cssx_mlp **out) {
    if (!out || nin != kFeat || nout != kOut || h1 < 16 || h2 < 16 || h1 > 4096 || h2 > 4096 || max_rows < 1)
        return CSSX_EINVAL;
    *out = nullptr;
    cssx_mlp *n = new (std::nothrow) cssx_mlp;
    if (!n) return CSSX_ENOMEM;
    n->nin = nin; n->h1 = h1; n->h2 = h2; n->nout = nout; n->max_rows = max_rows;
    n->oW1 = 0; n->ob1 = n->oW1 + (long)nin * h1; n->oW2 = n->ob1 + h1; n->ob2 = n->oW2 + (long)h1 * h2;
    n->oW3 = n->ob2 + h2; n->ob3 = n->oW3 + (long)h2 * nout; n->np = n->ob3 + nout;
    const long R = max_rows;
    const long fl = 4 * n->np + R * (2L * h1 + 2L * h2 + 2L * nout) + 1 + 2L * nin + 2L * nout;
    const long bf = n->np + R * (h1 + h2 + nout + h2 + h1 + nin);
    char *p = nullptr;
    const size_t bytes = (size_t)fl * 4 + (size_t)bf * 2 + kPackBytes + 64 * 16;
    if (hipMalloc((void **)&p, bytes) != hipSuccess) {
        delete n;
        return CSSX_ENOMEM;
    }
    n->blk = p;
    auto takef = [&](long cnt) { float *q = (float *)p; p += ((cnt * 4 + 255) / 256) * 256; return q; };
    auto takeb = [&](long cnt) { cssx_bf16 *q = (cssx_bf16 *)p; p += ((cnt * 2 + 255) / 256) * 256; return q; };
    n->W = takef(n->np); n->G = takef(n->np); n->Mo = takef(n->np); n->Vo = takef(n->np);
    n->Z1 = takef(R * h1); n->dZ1 = takef(R * h1); n->Z2 = takef(R * h2); n->dZ2 = takef(R * h2);
    n->Y = takef(R * nout); n->dY = takef(R * nout); n->loss = takef(1);
    n->xmean = takef(nin); n->xistd = takef(nin); n->ymean = takef(nout); n->ystd = takef(nout);
    n->Wb = takeb(n->np); n->H1 = takeb(R * h1); n->H2 = takeb(R * h2); n->dYb = takeb(R * nout);
    n->dZ2b = takeb(R * h2); n->dZ1b = takeb(R * h1); n->X = takeb(R * nin);
    n->pk = takeb((long)(kPackBytes / 2));
    // PyTorch-Linear-style initialization: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for W and b
    std::vector<float> w(n->np), zeros(n->np, 0.0f), ones(nin, 1.0f);
    uint64_t s = seed;
    auto fill = [&](long off, long cnt, int fan_in) {
        const double b = 1.0 / std::sqrt((double)fan_in);
        for (long i = 0; i < cnt; i++) w[off + i] = (float)((2.0 * ((splitmix(s) >> 11) * 0x1.0p-53) - 1.0) * b);
    };
    fill(n->oW1, (long)nin * h1, nin); fill(n->ob1, h1, nin);
    fill(n->oW2, (long)h1 * h2, h1); fill(n->ob2, h2, h1);
    fill(n->oW3, (long)h2 * nout, h2); fill(n->ob3, nout, h2);
    std::vector<float> zo(nout, 0.0f), oo(nout, 1.0f), zi(nin, 0.0f);
    bool ok = hipMemcpy(n->W, w.data(), n->np * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->Mo, zeros.data(), n->np * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->Vo, zeros.data(), n->np * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->xmean, zi.data(), nin * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->xistd, ones.data(), nin * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->ymean, zo.data(), nout * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->ystd, oo.data(), nout * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n->np)), dim3(256), 0, 0, n->W, n->Wb, n->np);
        ok = hipDeviceSynchronize() == hipSuccess;
    }
    if (!ok) {
        (void)hipFree(n->blk);
        delete n;
        return CSSX_EHIP;
    }
    *out = n;
    return CSSX_OK;
}

extern "C" void cssx_mlp_destroy(cssx_mlp *n) {
    if (!n) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(n->blk);
    delete n;
}

extern "C" long cssx_mlp_num_params(const cssx_mlp *n) { return n ? n->np : -1; }

extern "C" int cssx_mlp_set_params(cssx_mlp *n, const float *host_params) {
    if (!n || !host_params) return CSSX_EINVAL;
    if (hipMemcpy(n->W, host_params, n->np * 4, hipMemcpyHostToDevice) != hipSuccess) return CSSX_EHIP;
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n->np)), dim3(256), 0, 0, n->W, n->Wb, n->np);
    return hipDeviceSynchronize() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_mlp_get_params(cssx_mlp *n, float *host_params, float *host_grads) {
    if (!n || (!host_params && !host_grads)) return CSSX_EINVAL;
    if (hipDeviceSynchronize() != hipSuccess) return CSSX_EHIP;
    if (host_params && hipMemcpy(host_params, n->W, n->np * 4, hipMemcpyDeviceToHost) != hipSuccess) return CSSX_EHIP;
    if (host_grads && hipMemcpy(host_grads, n->G, n->np * 4, hipMemcpyDeviceToHost) != hipSuccess) return CSSX_EHIP;
    return CSSX_OK;
}

extern "C" int cssx_mlp_set_norm(cssx_mlp *n, const float *x_mean, const float *x_std, const float *y_mean,
                                 const float *y_std) {
    if (!n || !x_mean || !x_std || !y_mean || !y_std) return CSSX_EINVAL;
    std::vector<float> xi(n->nin);
    for (int i = 0; i < n->nin; i++) {
        if (!(x_std[i] > 0.0f)) return CSSX_EINVAL;
        xi[i] = 1.0f / x_std[i];
    }
    for (int i = 0; i < n->nout; i++)   // y_std 0 switches that output's tendency off
        if (!(y_std[i] >= 0.0f)) return CSSX_EINVAL;
    bool ok = hipMemcpy(n->xmean, x_mean, n->nin * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->xistd, xi.data(), n->nin * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->ymean, y_mean, n->nout * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n->ystd, y_std, n->nout * 4, hipMemcpyHostToDevice) == hipSuccess;
    return ok ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_mlp_forward(cssx_mlp *n, const void *x_bf16, int rows, float *y, void *stream) {
    if (!n || !x_bf16 || !y || rows < 1 || rows > n->max_rows) return CSSX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    int rc = mlp_forward(n, (const cssx_bf16 *)x_bf16, rows, true, s);
    if (rc) return rc;
    return hipMemcpyAsync(y, n->Y, (size_t)rows * n->nout * 4, hipMemcpyDeviceToDevice, s) == hipSuccess ? CSSX_OK
                                                                                                         : CSSX_EHIP;
}

extern "C" int cssx_mlp_grad(cssx_mlp *n, const void *x_bf16, const float *target, int rows, float *loss,
                             void *stream) {
    if (!n || !x_bf16 || !target || rows < 1 || rows > n->max_rows) return CSSX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, n->loss, 1L, 0.0f);
    int rc = mlp_forward(n, (const cssx_bf16 *)x_bf16, rows, true, s);
    if (!rc) rc = mlp_backward(n, (const cssx_bf16 *)x_bf16, target, rows, s);
    if (rc) return rc;
    if (loss && hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return CSSX_EHIP;
    return CSSX_OK;
}

extern "C" int cssx_mlp_train_step(cssx_mlp *n, const void *x_bf16, const float *target, int rows, float lr,
                                   float *loss, void *stream) {
    int rc = cssx_mlp_grad(n, x_bf16, target, rows, loss, stream);
    if (rc) return rc;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)n->step)), c2 = (float)(1.0 - std::pow(0.999, (double)n->step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->np)), dim3(256), 0, (hipStream_t)stream, n->W, n->G, n->Mo, n->Vo, n->Wb,
                       n->np, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_gemm_bf16(const void *A, long sam, long sak, const void *B, long sbk, long sbn, int M, int N,
                              int K, int ksplit, float *C, long ldc, void *stream) {
    if (!A || !B || !C || M < 1 || N < 1 || K < 1 || ldc < N) return CSSX_EINVAL;
    GemmArgs g = ga((const cssx_bf16 *)A, sam, sak, (const cssx_bf16 *)B, sbk, sbn, M, N, K,
                    ksplit > 1 ? EPI_ATOMIC : EPI_STORE);
    g.C = C; g.ldc = ldc;
    hipStream_t s = (hipStream_t)stream;
    if (ksplit > 1) {
        if (ldc != N) return CSSX_EINVAL;
        hipLaunchKernelGGL(k_fill, dim3(nblk((long)M * N)), dim3(256), 0, s, C, (long)M * N, 0.0f);
    }
    return gemm(g, ksplit, s);
}

extern "C" int cssx_box_filter(const void *in, int prec, int nv, int nr, int nphi, int nth, int factor, float *out,
                               void *stream) {
    if (!in || !out || nv < 1 || factor < 1 || nr % factor || nphi % factor || nth % factor ||
        (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    const long Nc = (long)(nr / factor) * (nphi / factor) * (nth / factor) * nv;
    hipStream_t s = (hipStream_t)stream;
    if (prec == CSSX_PREC_F64)
        hipLaunchKernelGGL(k_box_filter<double>, dim3(nblk(Nc)), dim3(256), 0, s, (const double *)in, out, nv, nr, nphi,
                           nth, factor);
    else
        hipLaunchKernelGGL(k_box_filter<float>, dim3(nblk(Nc)), dim3(256), 0, s, (const float *)in, out, nv, nr, nphi,
                           nth, factor);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_closure_features(cssx_mlp *n, const void *state, int prec, int fluct, const void *rho_bar,
                                     const void *s_bar, int nr, int nphi, int nth, long row0, long rows, void *x_bf16,
                                     float *x_raw, void *stream) {
    const long N = (long)nr * nphi * nth;
    if (!state || nr < 1 || nphi < 1 || nth < 1 || row0 < 0 || rows < 1 || row0 + rows > N || (!x_bf16 && !x_raw) ||
        (x_bf16 && !n) || (!fluct && (!rho_bar || !s_bar)) || (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    return prec == CSSX_PREC_F64
               ? features_impl<double>(n, (const double *)state, N, nr, nphi, nth, fluct, (const double *)rho_bar,
                                       (const double *)s_bar, row0, rows, (cssx_bf16 *)x_bf16, x_raw, s)
               : features_impl<float>(n, (const float *)state, N, nr, nphi, nth, fluct, (const float *)rho_bar,
                                      (const float *)s_bar, row0, rows, (cssx_bf16 *)x_bf16, x_raw, s);
}

extern "C" int cssx_closure_targets(const float *fine_filtered, const void *coarse_rhs, int prec, long npts,
                                    const float *y_mean, const float *y_std, float *target, float *raw, void *stream) {
    if (!fine_filtered || !coarse_rhs || npts < 1 || (!target && !raw) || (target && (!y_mean || !y_std)) ||
        (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    if (prec == CSSX_PREC_F64)
        hipLaunchKernelGGL(k_targets<double>, dim3(nblk(npts * kOut)), dim3(256), 0, s, fine_filtered,
                           (const double *)coarse_rhs, npts, y_mean, y_std, target, raw);
    else
        hipLaunchKernelGGL(k_targets<float>, dim3(nblk(npts * kOut)), dim3(256), 0, s, fine_filtered,
                           (const float *)coarse_rhs, npts, y_mean, y_std, target, raw);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_colstats(const float *X, long rows, int cols, double *mean, double *stddev, void *stream) {
    if (!X || rows < 2 || cols < 1 || !mean || !stddev) return CSSX_EINVAL;
    double *acc = nullptr;
    if (hipMalloc((void **)&acc, 2 * cols * sizeof(double)) != hipSuccess) return CSSX_ENOMEM;
    hipStream_t s = (hipStream_t)stream;
    std::vector<double> h(2 * cols, 0.0);
    bool ok = hipMemcpyAsync(acc, h.data(), 2 * cols * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    const int chunk = 4096;
    if (ok) {
        hipLaunchKernelGGL(k_colstats, dim3(nblk(rows, chunk)), dim3(256), 0, s, X, rows, cols, chunk, acc);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(h.data(), acc, 2 * cols * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    (void)hipFree(acc);
    if (!ok) return CSSX_EHIP;
    for (int c = 0; c < cols; c++) {
        const double m = h[c] / rows;
        mean[c] = m;
        stddev[c] = std::sqrt(std::max(0.0, h[cols + c] / rows - m * m));
    }
    return CSSX_OK;
}

extern "C" int cssx_field_stats(const void *state, int prec, int nv, int nr, int nphi, int nth, int fluct,
                                const void *rho_bar, const void *s_bar, double *mean, double *rms, void *stream) {
    if (!state || nv < 1 || nv > 8 || nr < 1 || nphi < 1 || nth < 1 || !mean || !rms ||
        (!fluct && (!rho_bar || !s_bar)) || (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    const long N = (long)nr * nphi * nth;
    double *acc = nullptr;
    if (hipMalloc((void **)&acc, 16 * sizeof(double)) != hipSuccess) return CSSX_ENOMEM;
    hipStream_t s = (hipStream_t)stream;
    std::vector<double> h(16, 0.0);
    bool ok = hipMemcpyAsync(acc, h.data(), 16 * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) {
        dim3 grid((unsigned)std::min<long>(nblk(N), 2048), nv);
        if (prec == CSSX_PREC_F64)
            hipLaunchKernelGGL(k_field_stats<double>, grid, dim3(256), 0, s, (const double *)state, nv, N, nphi * nth,
                               fluct, (const double *)rho_bar, (const double *)s_bar, acc);
        else
            hipLaunchKernelGGL(k_field_stats<float>, grid, dim3(256), 0, s, (const float *)state, nv, N, nphi * nth,
                               fluct, (const float *)rho_bar, (const float *)s_bar, acc);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(h.data(), acc, 16 * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    (void)hipFree(acc);
    if (!ok) return CSSX_EHIP;
    for (int iv = 0; iv < nv; iv++) {
        mean[iv] = h[iv] / N;
        rms[iv] = std::sqrt(std::max(0.0, h[nv + iv] / N));
    }
    return CSSX_OK;
}

extern "C" int cssx_gather_rows(const void *src, long row_bytes, const long *idx, long n, void *dst, void *stream) {
    if (!src || !idx || !dst || n < 1 || row_bytes < 4 || row_bytes % 4) return CSSX_EINVAL;
    const int words = (int)(row_bytes / 4);
    hipLaunchKernelGGL(k_gather_rows, dim3(nblk(n * words)), dim3(256), 0, (hipStream_t)stream,
                       (const unsigned *)src, idx, n, words, (unsigned *)dst);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_standardize_rows(const float *x, long rows, int cols, const float *mean, const float *stddev,
                                     void *out_bf16, float *out_f32, void *stream) {
    if (!x || rows < 1 || cols < 1 || !mean || !stddev || (!out_bf16 && !out_f32)) return CSSX_EINVAL;
    hipLaunchKernelGGL(k_standardize, dim3(nblk(rows * cols)), dim3(256), 0, (hipStream_t)stream, x, rows, cols, mean,
                       stddev, (cssx_bf16 *)out_bf16, out_f32);
    return hipGetLastError() == hipSuccess ? CSSX_OK : CSSX_EHIP;
}

extern "C" int cssx_closure_tendency(cssx_mlp *n, const void *state, int prec, int fluct, const void *rho_bar,
                                     const void *s_bar, int nr, int nphi, int nth, float *ext, void *stream) {
    if (!n || !state || !ext || nr < 1 || nphi < 1 || nth < 1 || (!fluct && (!rho_bar || !s_bar)) ||
        (prec != CSSX_PREC_F64 && prec != CSSX_PREC_F32))
        return CSSX_EINVAL;
    return closure_tendency(n, prec, state, fluct, rho_bar, s_bar, nr, nphi, nth, ext, (hipStream_t)stream);
}
