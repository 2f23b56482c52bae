// Kernel launch descriptors shared by the HIP sources and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

namespace cssx {

constexpr int kMaxJobs = 8;

// One field's derivative inside a batched launch.  All jobs of a launch share the pencil
// geometry (length n, pencil strides) but each has its own data, ghosts, closure and plan.
template <typename T>
struct DerivJob {
    const T *in;        // field (reference layout f(nth,nphi,nr), any axis)
    T *out;             // f' (may alias nothing that the job reads)
    T *rhs_out;         // optional: the assembled right-hand side rows (times rhs_scale)
    const T *glo;       // lower ghosts (halo: 4 values, slot 3 adjacent) or wall value (slot 0)
    const T *ghi;       // upper ghosts (slot 0 adjacent) or wall value (slot 0)
    long g_k, g_a, g_b; // ghost strides: slot, fast pencil index, slow pencil index
    const T *plan;      // device plan buffer (cssx_plan.hpp layout)
    const T *pre;       // optional per-row multiplier applied to the loaded field
    const T *post;      // optional per-row multiplier applied to f' (e.g. dx/dr)
    const T *mul_a;     // optional multiplier of f' by the pencil's fast index a
    const T *mul_b;     // optional multiplier of f' by the pencil's slow index b
    T *tap_in, *tap_out;  // contiguous kernel only, optional: the input and f' of the pencil's end
                          // rows copied to [side][b][a] (side 0: row 0, 1: row n-1; stride tap_side)
    long tap_side;
    T scale;            // 1/dx (order 1) or 1/dx^2 (order 2)
    T neu_mul;          // converts a prescribed wall derivative to index units
    T halo_s;           // multiplier of the stencil part of the halo rows (1 = Fortran)
    T rhs_scale;
    T beta;             // out = f' + beta*out
    int bdry;           // boundary code 0..4
    int lo_kind, hi_kind;
    int lo_r0_val;      // lower Dirichlet-type wall row takes R0 from the wall value (CSS-Python
                        // compact_fd6 with ibc=5, derivatives.py:134-155)
};

template <typename T>
struct DerivLaunch {
    int n, S;           // pencil length, segments per pencil (contiguous kernel: lanes per pencil)
    int M;              // strided kernels: rows-per-segment template size (S = ceil(n/M))
    int na, nb;         // pencils: fast index a (lanes) x slow index b (grid.y)
    long sa, sb, srow;  // element strides of a, b and along the pencil
    int njobs;
    DerivJob<T> jobs[kMaxJobs];
};

template <typename T>
hipError_t launch_deriv(const DerivLaunch<T> &L, int order, hipStream_t stream);
template <typename T>
hipError_t launch_deriv_contig(const DerivLaunch<T> &L, int order, hipStream_t stream);
template <typename T>
hipError_t launch_deriv_wide(const DerivLaunch<T> &L, int order, hipStream_t stream);
// (M, G) of the wide strided kernel for a pencil of n rows (M = 0: k_deriv serves it)
void wide_shape(int n, bool f64, int &M, int &G);

}  // namespace cssx
