// Host-side plan builder for the segmented compact solve (see cssx_plan.hpp).
#include "cssx_plan.hpp"

#include <cmath>

namespace cssx {

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

void build_plan(const PlanKey &key, int S, std::vector<double> &plan) {
    const int n = key.n;
    std::vector<double> l, u;
    lhs_coefficients(key.order, n, key.lo, key.hi, key.bdry, l, u);
    plan.assign(plan_len(n), 0.0);
    double *cf = plan.data(), *cl = cf + n, *cu = cl + n, *vL = cu + n, *vR = vL + n;
    double *rl = plan.data() + 5 * n, *ru = rl + kMaxSeg, *rf = ru + kMaxSeg, *rcl = rf + kMaxSeg,
           *rcu = rcl + kMaxSeg;

    // segment-local LU of the interior rows; interface row = last row of segments 0..S-2
    for (int k = 0; k < S; k++) {
        int st = seg_start(n, S, k), en = seg_start(n, S, k + 1);   // [st, en)
        int last = (k < S - 1) ? en - 2 : en - 1;                  // last interior row
        double piv_prev = 0.0, up_prev = 0.0;
        for (int j = st; j <= last; j++) {
            double piv = (j == st) ? 1.0 : 1.0 - l[j] * up_prev / piv_prev;
            cf[j] = 1.0 / piv;
            cl[j] = (j == st) ? 0.0 : l[j] / piv;
            cu[j] = (j == last) ? 0.0 : u[j] / piv;
            piv_prev = piv;
            up_prev = u[j];
        }
        // responses to the left / right interface values (solve T v = -l_st e_st, -u_last e_last)
        std::vector<double> y(last - st + 1);
        auto solve = [&](int which, double *v) {
            for (int j = st; j <= last; j++) {
                double rhs = 0.0;
                if (which == 0 && j == st) rhs = -l[st];
                if (which == 1 && j == last) rhs = -u[last];
                y[j - st] = cf[j] * rhs - ((j == st) ? 0.0 : cl[j] * y[j - st - 1]);
            }
            for (int j = last; j >= st; j--)
                v[j] = y[j - st] - ((j == last) ? 0.0 : cu[j] * v[j + 1]);
        };
        if (k > 0) solve(0, vL);
        if (k < S - 1) solve(1, vR);
    }
    // reduced system over interfaces I_k = seg_start(k+1)-1, k = 0..S-2:
    //   A_k x_{I_{k-1}} + B_k x_{I_k} + C_k x_{I_{k+1}} = d_{I_k} - l x'_{I_k-1} - u x'_{I_k+1}
    int nI = S - 1;
    double piv_prev = 0.0, C_prev = 0.0;
    for (int k = 0; k < nI; k++) {
        int I = seg_start(n, S, k + 1) - 1;
        double A = (k > 0) ? l[I] * vL[I - 1] : 0.0;
        double B = 1.0 + l[I] * vR[I - 1] + u[I] * vL[I + 1];
        double C = (k + 1 < nI) ? u[I] * vR[I + 1] : 0.0;
        rl[k] = l[I];
        ru[k] = u[I];
        double piv = (k == 0) ? B : B - A * C_prev / piv_prev;
        rf[k] = 1.0 / piv;
        rcl[k] = (k == 0) ? 0.0 : A / piv;
        rcu[k] = C / piv;
        piv_prev = piv;
        C_prev = C;
    }
}

}  // namespace cssx
