// Host-side plans for the segmented compact solve.
//
// A compact derivative along one axis solves, for every pencil, the tridiagonal system
//   l_j f'_{j-1} + f'_j + u_j f'_{j+1} = d_j,   j = 0..n-1
// whose coefficients depend only on the pencil length, the derivative order, the closure at
// each end (physical wall or 4-deep halo) and the boundary code.  The reference precomputes
// Thomas-LU tables for that system once (derivatives.f90:106-518) and sweeps every pencil
// serially.  Here the pencil is cut into S segments, one wavefront each: every segment solves
// its own rows with zero coupling to its neighbours, the S-1 interface rows form a small
// tridiagonal system, and the segment solutions are corrected by the interface values.
// All of that is linear algebra on the coefficients alone, so it is precomputed here in FP64
// (per row: local LU factors and the two response vectors; per interface: the LU factors of
// the reduced system) and the GPU only streams the data.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cssx {

enum EndKind : int { END_WALL = 0, END_HALO = 1, END_PERIODIC = 2 };

struct PlanKey {
    int n, order, lo, hi, bdry;
    int variant = 0;   // 0: strided (waves = segments), 1: contiguous (lanes = segments),
                       // 2: wide strided (contiguous-kernel plan cut at M rows)
    int seg = 0;       // strided / wide: rows-per-segment size M the plan is cut for
    bool operator<(const PlanKey &o) const {
        if (variant != o.variant) return variant < o.variant;
        if (seg != o.seg) return seg < o.seg;
        if (n != o.n) return n < o.n;
        if (order != o.order) return order < o.order;
        if (lo != o.lo) return lo < o.lo;
        if (hi != o.hi) return hi < o.hi;
        return bdry < o.bdry;
    }
};

// Layout of one plan buffer (T = float or double), n rows, S segments (S <= kMaxSeg):
//   [0n) cf  : 1/pivot of the segment-local LU
//   [1n) cl  : l_j / pivot_j        (forward:  y_j = cf_j d_j - cl_j y_{j-1})
//   [2n) cu  : u_j / pivot_j        (backward: x_j = y_j - cu_j x_{j+1})
//   [3n) vL  : response of row j to the left interface value
//   [4n) vR  : response of row j to the right interface value
//   [5n + 0*kMaxSeg) rl  : interface row lower coefficient l_{I_k}
//   [5n + 1*kMaxSeg) ru  : interface row upper coefficient u_{I_k}
//   [5n + 2*kMaxSeg) rf  : reduced-system 1/pivot
//   [5n + 3*kMaxSeg) rcl : reduced-system lower/pivot
//   [5n + 4*kMaxSeg) rcu : reduced-system upper/pivot
constexpr int kMaxSeg = 16;
__host__ __device__ inline int plan_len(int n) { return 5 * n + 5 * kMaxSeg; }

// Segment split used by both the plan and the kernels: S segments, the first n%S get one
// extra row.
__host__ __device__ inline int seg_count(int n, int M) { return (n + M - 1) / M; }
__host__ __device__ inline int seg_start(int n, int S, int k) {
    int base = n / S, rem = n % S;
    return k * base + (k < rem ? k : rem);
}

// Rows-per-segment template size for a pencil of n rows (S = ceil(n/M) <= kMaxSeg).
inline int choose_M(int n) {
    if (n <= 16 * 8 && n <= 64) return 8;
    if (n <= 16 * 16) return 16;
    if (n <= 16 * 32) return 32;
    if (n <= 16 * 64) return 64;
    return -1;
}

// Rows-per-segment size the strided kernel uses for a pencil of n rows: choose_M, unless the
// environment variable CSSX_DERIV_M (8, 16, 32 or 64) names a size that fits (S <= kMaxSeg);
// read once per process.  A tuning hook for the segment length, which trades the serial
// sweep length against the halo rows re-read per segment.
// FP32 pencils of 129-256 rows take 32-row segments (8 per pencil: fewer halo rows re-read per
// row; measured 9% faster on 256-row r pencils than 16-row segments); FP64 keeps choose_M,
// whose 32-row instantiation would spill.
int strided_M(int n, bool f64 = true);

// Tridiagonal coefficients l_j, u_j (diagonal 1) of the compact system.
//   order 1: alpha = 1/3, wall rows derivatives.f90:108-125 (row 0 upper = 4 for codes 1,3
//            else 0; row 1 = (1/8, 3/4); mirrored at the upper wall, row n-1 lower = 4 for
//            codes 1,2 else 0).
//   order 2: alpha = 2/11, wall rows derivatives.f90:196-216, :256-261 (row 0 upper =
//            137/13 for Dirichlet-type ends else 707/193; row 1 = (0.1, -0.35)).
void lhs_coefficients(int order, int n, int lo, int hi, int bdry, std::vector<double> &l,
                      std::vector<double> &u);

// Fills a plan buffer (FP64) for the given key and segment count S.
void build_plan(const PlanKey &key, int S, std::vector<double> &plan);

// ---- contiguous-axis variant: lanes walk the pencil --------------------------------------------
// Segment j of a pencil is rows [j*MR, min((j+1)*MR, n)), one lane each, nl = ceil(n/MR) <= 64
// lanes.  The nl-1 interface rows form a tridiagonal system that the lanes solve together by
// parallel cyclic reduction over wavefront shuffles; its multipliers are data-independent.
// Layout (n rows, K = kMaxLanes):
//   [0n..5n)       cf, cl, cu, vL, vR            (as in the strided plan)
//   [5n)           rl[K], ru[K]                  interface row coefficients
//   [5n+2K)        k1[kPcrSteps][K], k2[kPcrSteps][K]   PCR multipliers
//   [5n+2K+2SK)    inv_b[K]                      1/diagonal after the last PCR step
constexpr int kMaxLanes = 64;
constexpr int kPcrSteps = 6;
__host__ __device__ inline int plan_contig_len(int n) { return 5 * n + 3 * kMaxLanes + 2 * kPcrSteps * kMaxLanes; }
// rows per lane for a pencil of n rows (4, 8 or 16), -1 if n > 1024.  FP32 takes 8 rows per
// lane from 256 rows down (two pencils per wavefront, half the shuffles and PCR steps per
// row); FP64 keeps 4 there, where 8 would cost it occupancy.
inline int choose_MR(int n, bool f64 = true) {
    if (n <= 64 * 4 && f64) return 4;
    if (n <= 64 * 8) return 8;
    if (n <= 64 * 16) return 16;
    return -1;
}
void build_plan_contig(const PlanKey &key, int MR, std::vector<double> &plan);

}  // namespace cssx
