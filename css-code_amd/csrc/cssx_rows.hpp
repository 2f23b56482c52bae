// Right-hand-side rows of the compact systems (shared by the derivative kernels).
#pragma once
#include <hip/hip_runtime.h>

namespace cssx {

template <typename T, int ORDER>
struct Rows;

// First-derivative rows in index units (the scale 1/dx is applied after the solve).
template <typename T>
struct Rows<T, 1> {
    // interior, derivatives.f90:725-728
    static __device__ __forceinline__ T interior(T xm2, T xm1, T xp1, T xp2) {
        return T(7.0 / 9.0) * (xp1 - xm1) + T(1.0 / 36.0) * (xp2 - xm2);
    }
    // lower halo row 0, derivatives.f90:565-580 written as stencil + explicit ghost derivative:
    // S0*hs - alpha*E0/60 (hs = 1 reproduces the Fortran; hs = dli/60 the CSS-Python rows).
    static __device__ __forceinline__ T halo_lo0(T g0, T g1, T g2, T g3, T x0, T x1, T x2, T hs) {
        T s = T(7.0 / 9.0) * (x1 - g3) + T(1.0 / 36.0) * (x2 - g2);
        T e = -g0 + T(9.0) * g1 - T(45.0) * g2 + T(45.0) * x0 - T(9.0) * x1 + x2;
        return s * hs - T(1.0 / 180.0) * e;
    }
    // upper halo row n-1, derivatives.f90:669-680
    static __device__ __forceinline__ T halo_hi0(T x3, T x2, T x1, T b0, T b1, T b2, T b3, T hs) {
        // x1 = f[n-1], x2 = f[n-2], x3 = f[n-3]; b0.. = ghosts n, n+1, ...
        T s = T(7.0 / 9.0) * (b0 - x2) + T(1.0 / 36.0) * (b1 - x3);
        T e = -x3 + T(9.0) * x2 - T(45.0) * x1 + T(45.0) * b1 - T(9.0) * b2 + b3;
        return s * hs - T(1.0 / 180.0) * e;
    }
    // wall row 1, derivatives.f90:633-640 (and its mirror :585-592)
    static __device__ __forceinline__ T wall1(T x0, T x1, T x2, T x3, T x4) {
        return T(-43.0 / 96.0) * x0 + T(-5.0 / 6.0) * x1 + T(9.0 / 8.0) * x2 + T(1.0 / 6.0) * x3 +
               T(-1.0 / 96.0) * x4;
    }
    // wall row 0: Neumann value or the 5th-order closure preconditioned to tridiagonal form
    // (derivatives.f90:623-662): 2*R0 - 8*R1.
    static __device__ __forceinline__ T wall0(int dirichlet, T val, T x0, T x1, T x2, T x3, T r1,
                                              bool r0_from_val = false) {
        if (dirichlet) {
            T r0 = r0_from_val ? val : T(-10.0 / 3.0) * x0 + T(-3.0) * x1 + T(6.0) * x2 + T(1.0 / 3.0) * x3;
            return T(2.0) * r0 + T(-8.0) * r1;
        }
        return val;
    }
};

// Second-derivative rows, derivatives.f90:816-938.
template <typename T>
struct Rows<T, 2> {
    static __device__ __forceinline__ T interior(T xm2, T xm1, T x0, T xp1, T xp2) {
        return T(12.0 / 11.0) * (xp1 - T(2.0) * x0 + xm1) + T(3.0 / 44.0) * (xm2 - T(2.0) * x0 + xp2);
    }
    static __device__ __forceinline__ T halo_lo0(T g0, T g1, T g2, T g3, T x0, T x1, T x2) {
        T r = T(-1.0 / 990.0) * (T(2.0) * (g0 + x2) + T(-27.0) * (g1 + x1) + T(270.0) * (g2 + x0) +
                                 T(-490.0) * g3);
        return r + T(12.0 / 11.0) * (x1 - T(2.0) * x0 + g3) + T(3.0 / 44.0) * (g2 - T(2.0) * x0 + x2);
    }
    static __device__ __forceinline__ T halo_hi0(T x3, T x2, T x1, T b0, T b1, T b2, T b3) {
        T r = T(12.0 / 11.0) * (b0 - T(2.0) * x1 + x2) + T(3.0 / 44.0) * (x3 - T(2.0) * x1 + b1);
        return r + T(-1.0 / 990.0) * (T(2.0) * (x3 + b3) + T(-27.0) * (x2 + b2) + T(270.0) * (x1 + b1) +
                                      T(-490.0) * b0);
    }
    static __device__ __forceinline__ T wall1(T x0, T x1, T x2, T x3, T x4) {
        return T(1.2375) * x0 + T(-3.0) * x1 + T(2.325) * x2 + T(-0.6) * x3 + T(0.0375) * x4;
    }
    // row 0: Dirichlet-type 5th-order closure, or the closure with the first derivative given
    // (sgn = -1 at the lower end, +1 at the upper end; val already includes 1/dri).
    static __device__ __forceinline__ T wall0(int dirichlet, T sgn_val, T x0, T x1, T x2, T x3, T x4, T x5) {
        if (dirichlet)
            return T(1955.0 / 156.0) * x0 + T(-4057.0 / 156.0) * x1 + T(1117.0 / 78.0) * x2 +
                   T(-55.0 / 78.0) * x3 + T(-29.0 / 156.0) * x4 + T(7.0 / 156.0) * x5;
        return T(575.0 / 193.0) * sgn_val + T(-5827.0 / 2316.0) * x1 + T(1987.0 / 1158.0) * x2 +
               T(465.0 / 386.0) * x3 + T(-547.0 / 1158.0) * x4 + T(157.0 / 2316.0) * x5;
    }
};

}  // namespace cssx
