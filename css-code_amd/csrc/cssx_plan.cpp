// Host-side plan builder for the segmented compact solve (see cssx_plan.hpp).
#include "cssx_plan.hpp"

#include <cmath>
#include <cstdlib>

namespace cssx {

int strided_M(int n, bool f64) {
    static const int forced = [] {
        const char *e = std::getenv("CSSX_DERIV_M");
        const int v = e ? std::atoi(e) : 0;
        return (v == 8 || v == 16 || v == 32 || v == 64) ? v : 0;
    }();
    if (forced && seg_count(n, forced) <= kMaxSeg && n >= forced) return forced;
    if (!f64 && n > 128 && n <= 256) return 32;
    return choose_M(n);
}

void lhs_coefficients(int order, int n, int lo, int hi, int bdry, std::vector<double> &l,
                      std::vector<double> &u) {
    const double a = (order == 1) ? 1.0 / 3.0 : 2.0 / 11.0;
    l.assign(n, a);
    u.assign(n, a);
    l[0] = 0.0;
    u[n - 1] = 0.0;
    double up1, up2, r_lo, r_up;
    if (order == 1) {
        up1 = (bdry == 1 || bdry == 3) ? 4.0 : 0.0;   // (a1*a2-b1)/(a2-b1*g2) = 4
        up2 = (bdry == 1 || bdry == 2) ? 4.0 : 0.0;
        r_lo = 0.125;
        r_up = 0.75;
    } else {
        up1 = (bdry == 1 || bdry == 3) ? 137.0 / 13.0 : 707.0 / 193.0;
        up2 = (bdry == 1 || bdry == 2) ? 137.0 / 13.0 : 707.0 / 193.0;
        r_lo = 0.1;
        r_up = -0.35;
    }
    if (lo == END_WALL) {
        u[0] = up1;
        l[1] = r_lo;
        u[1] = r_up;
    }
    if (hi == END_WALL) {
        l[n - 1] = up2;
        u[n - 2] = r_lo;
        l[n - 2] = r_up;
    }
}

// Segment-local LU factors and interface response vectors for an arbitrary segmentation
// (segment k = rows [start(k), start(k+1)), the last row of segments 0..S-2 is an interface).
template <typename Start>
static void local_factors(int n, int S, const std::vector<double> &l, const std::vector<double> &u, Start start,
                          double *cf, double *cl, double *cu, double *vL, double *vR) {
    for (int k = 0; k < S; k++) {
        int st = start(k), en = start(k + 1);   // [st, en)
        int last = (k < S - 1) ? en - 2 : en - 1;
        double piv_prev = 0.0, up_prev = 0.0;
        for (int j = st; j <= last; j++) {
            double piv = (j == st) ? 1.0 : 1.0 - l[j] * up_prev / piv_prev;
            cf[j] = 1.0 / piv;
            cl[j] = (j == st) ? 0.0 : l[j] / piv;
            cu[j] = (j == last) ? 0.0 : u[j] / piv;
            piv_prev = piv;
            up_prev = u[j];
        }
        if (last < st) continue;
        std::vector<double> y(last - st + 1);
        auto solve = [&](int which, double *v) {
            for (int j = st; j <= last; j++) {
                double rhs = 0.0;
                if (which == 0 && j == st) rhs = -l[st];
                if (which == 1 && j == last) rhs = -u[last];
                y[j - st] = cf[j] * rhs - ((j == st) ? 0.0 : cl[j] * y[j - st - 1]);
            }
            for (int j = last; j >= st; j--) v[j] = y[j - st] - ((j == last) ? 0.0 : cu[j] * v[j + 1]);
        };
        if (k > 0) solve(0, vL);
        if (k < S - 1) solve(1, vR);
    }
}

// Reduced (interface) system coefficients A_k x_{k-1} + B_k x_k + C_k x_{k+1} = R_k.
template <typename Start>
static void reduced_system(int S, const std::vector<double> &l, const std::vector<double> &u, Start start,
                           const double *vL, const double *vR, std::vector<double> &A, std::vector<double> &B,
                           std::vector<double> &C) {
    const int nI = S - 1;
    A.assign(nI > 0 ? nI : 0, 0.0);
    B.assign(nI > 0 ? nI : 0, 1.0);
    C.assign(nI > 0 ? nI : 0, 0.0);
    for (int k = 0; k < nI; k++) {
        int I = start(k + 1) - 1;
        A[k] = (k > 0) ? l[I] * vL[I - 1] : 0.0;
        B[k] = 1.0 + l[I] * vR[I - 1] + u[I] * vL[I + 1];
        C[k] = (k + 1 < nI) ? u[I] * vR[I + 1] : 0.0;
    }
}

void build_plan(const PlanKey &key, int S, std::vector<double> &plan) {
    const int n = key.n;
    std::vector<double> l, u;
    lhs_coefficients(key.order, n, key.lo, key.hi, key.bdry, l, u);
    plan.assign(plan_len(n), 0.0);
    double *cf = plan.data(), *cl = cf + n, *cu = cl + n, *vL = cu + n, *vR = vL + n;
    double *rl = plan.data() + 5 * n, *ru = rl + kMaxSeg, *rf = ru + kMaxSeg, *rcl = rf + kMaxSeg,
           *rcu = rcl + kMaxSeg;
    auto start = [&](int k) { return seg_start(n, S, k); };
    local_factors(n, S, l, u, start, cf, cl, cu, vL, vR);
    std::vector<double> A, B, C;
    reduced_system(S, l, u, start, vL, vR, A, B, C);
    // Thomas factors of the reduced system
    double piv_prev = 0.0, C_prev = 0.0;
    for (int k = 0; k < S - 1; k++) {
        int I = start(k + 1) - 1;
        rl[k] = l[I];
        ru[k] = u[I];
        double piv = (k == 0) ? B[k] : B[k] - A[k] * C_prev / piv_prev;
        rf[k] = 1.0 / piv;
        rcl[k] = (k == 0) ? 0.0 : A[k] / piv;
        rcu[k] = C[k] / piv;
        piv_prev = piv;
        C_prev = C[k];
    }
}

void build_plan_contig(const PlanKey &key, int MR, std::vector<double> &plan) {
    const int n = key.n;
    const int nl = (n + MR - 1) / MR;
    std::vector<double> l, u;
    lhs_coefficients(key.order, n, key.lo, key.hi, key.bdry, l, u);
    plan.assign(plan_contig_len(n), 0.0);
    double *cf = plan.data(), *cl = cf + n, *cu = cl + n, *vL = cu + n, *vR = vL + n;
    double *rl = plan.data() + 5 * n, *ru = rl + kMaxLanes, *k1 = ru + kMaxLanes,
           *k2 = k1 + kPcrSteps * kMaxLanes, *invb = k2 + kPcrSteps * kMaxLanes;
    auto start = [&](int k) { return k >= nl ? n : k * MR; };
    local_factors(n, nl, l, u, start, cf, cl, cu, vL, vR);
    std::vector<double> A, B, C;
    reduced_system(nl, l, u, start, vL, vR, A, B, C);
    const int nI = nl - 1;
    for (int k = 0; k < nI; k++) {
        int I = start(k + 1) - 1;
        rl[k] = l[I];
        ru[k] = u[I];
    }
    // parallel cyclic reduction: after step s (distance d = 2^s) equation k couples x_{k-2d}, x_{k+2d}
    for (int s = 0, d = 1; s < kPcrSteps; s++, d *= 2) {
        std::vector<double> A2(nI), B2(nI), C2(nI);
        for (int k = 0; k < nI; k++) {
            double a = (k - d >= 0) ? A[k] / B[k - d] : 0.0;
            double c = (k + d < nI) ? C[k] / B[k + d] : 0.0;
            k1[s * kMaxLanes + k] = a;
            k2[s * kMaxLanes + k] = c;
            A2[k] = (k - d >= 0) ? -A[k - d] * a : 0.0;
            C2[k] = (k + d < nI) ? -C[k + d] * c : 0.0;
            B2[k] = B[k] - ((k - d >= 0) ? C[k - d] * a : 0.0) - ((k + d < nI) ? A[k + d] * c : 0.0);
        }
        A = A2; B = B2; C = C2;
    }
    for (int k = 0; k < nI; k++) invb[k] = 1.0 / B[k];
}

}  // namespace cssx
