// Internal host-side interfaces shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cssx_plan.hpp"

struct cssx_mlp;

namespace cssx {

// Pencil geometry of one axis on the reference layout (see axis_geometry).
struct AxisGeom {
    int n = 0, na = 0, nb = 0;
    long sa = 0, sb = 0, srow = 0;
    long g_k = 0, g_a = 0, g_b = 0;   // default ghost strides (reference halo layout)
    int phase = 0;                    // timing bucket of launches along this axis
};

// One derivative request (precision-neutral; pointers are device pointers).
struct DerivSpec {
    int order = 1;
    const void *in = nullptr;
    void *out = nullptr;
    void *rhs_out = nullptr;
    const void *glo = nullptr, *ghi = nullptr;
    long g_k = -1, g_a = 0, g_b = 0;      // -1: use the axis default ghost layout
    const void *pre = nullptr, *post = nullptr;
    const void *mul_a = nullptr, *mul_b = nullptr;   // per-pencil multipliers (fast / slow index)
    void *tap_in = nullptr, *tap_out = nullptr;      // contiguous axis only: end-row copies
    long tap_side = 0;
    double scale = 1.0, neu_mul = 1.0, halo_s = 1.0, rhs_scale = 1.0, beta = 0.0;
    int bdry = 0;
    int lo_kind = END_WALL, hi_kind = END_WALL;
    int lo_r0_val = 0;
    const void *plan = nullptr;           // device plan to use instead of the plan cache's
};

// per-phase event timing (cssx_timing.hip); token < 0 = disabled
int timing_begin(int phase, hipStream_t stream);
bool timing_on();
void timing_end(int token, hipStream_t stream, double bytes);

// learned closure tendency of a whole state (cssx_closure.hip)
int closure_tendency(cssx_mlp *n, int prec, const void *st, int fluct, const void *rbar, const void *sbar, int nr,
                     int nphi, int nth, float *ext, hipStream_t s);

void axis_geometry(int axis, int nth, int nphi, int nr, AxisGeom &g);
int get_plan(const PlanKey &key, int prec, const void **out);
// the plan key (variant and segment size included) that run_derivs uses for spec s of a
// launch of `nspec` specs along geometry g
PlanKey launch_plan_key(const DerivSpec *specs, int nspec, int i, const AxisGeom &g, bool f64);
// the FP64 plan of a key, as get_plan would upload it
int host_plan(const PlanKey &key, std::vector<double> &plan);
int run_derivs(const DerivSpec *specs, int nspec, const AxisGeom &g, int prec, hipStream_t stream);

}  // namespace cssx
