/*
 * CPU oracle: FP64 restatement of the CSS compact sixth-order derivative operators.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library (libcssx.so) never links or calls it.
 *
 * What it restates (reference file:line):
 *   - Thomas-LU factor tables gam1/gam2/gam3 for the first derivative (alpha=1/3) and the
 *     second derivative (alpha=2/11), per boundary code 0..4:
 *       derivatives.f90:106-191 (r, d1), :194-287 (r, d2), :290-373 (theta, d1),
 *       :376-469 (theta, d2), :472-493 (phi, d1), :497-518 (phi, d2).
 *   - One pencil of Dbydr/Dbydt (derivatives.f90:547-796, :1010-1262): boundary rows for a
 *     lower/upper physical wall (Select Case 2 / Case 1) or a 4-deep halo (Case 1 / 2 / 3),
 *     interior rows, forward sweep, back substitution, scale.  A pencil with two walls is
 *     the composition of the Case 2 lower rows and the Case 1 upper rows (the Fortran has
 *     no Case 0 at all; see the header of derivatives.py:66-166 for the same composition).
 *   - Dbydp (derivatives.f90:1480-1591): halo rows at both ends, its own sweep form.
 *   - D2bydr2/D2bydt2/D2bydp2 (derivatives.f90:799-1007, :1265-1477, :1594-1684).
 *   - CSS-Python compact_fd6 (CSS-Python/derivatives.py:32-185) with its Thomas solver
 *     tri_solver(its=0) (CSS-Python/linalg.py:94-111), operation for operation, including
 *     the halo-row scaling of dtype 1/2/3 (the stencil part of those rows is multiplied by
 *     dli/60 a second time, derivatives.py:49,91,138-141,175-181).
 *
 * Parity pins (tests/test_oracle_*.py): dense solve of the assembled banded system,
 * polynomial exactness of every row to its order, the factor-table fixed point,
 * analytic derivatives of trigonometric fields, Python/Fortran agreement at n=61 where
 * the Python halo scaling is exact.
 *
 * Row indices below are 0-based (Fortran row J is j=J-1 here).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ALPHA1 (1.0 / 3.0)
#define ALPHA2 (2.0 / 11.0)

/* ---------------------------------------------------------------------------------------
 * Factor tables.  lo_wall/hi_wall say whether row 0 / row n-1 is a physical wall (the
 * Fortran's myr1==1 / myr2==nr).  iv is the boundary code 0..4.  Tables have n entries.
 * derivatives.f90:108-191 (first derivative).
 * ------------------------------------------------------------------------------------- */
void orc_factors_d1(int n, int lo_wall, int hi_wall, int iv, double *g1, double *g2, double *g3)
{
    const double alpha = 1.0 / 3.0, a2 = 0.75, g2c = 0.125, a1 = 6.0, b1 = 3.0;
    double cn = 1.0 / (a2 - b1 * g2c);
    double up1 = (a1 * a2 - b1) * cn, up2;
    double c1 = 1.0, c2 = 1.0, c3;
    int j;
    up2 = (iv != 1 && iv != 2) ? 0.0 : up1;
    if (iv % 2 == 0) up1 = 0.0;
    for (j = 0; j < n; j++) { g1[j] = 0; g2[j] = 0; g3[j] = 0; }
    g1[0] = c2; g2[0] = c2; g3[0] = c2;
    if (lo_wall) {
        c3 = up1 * c2; g1[1] = c3; c2 = c1 / (c1 - g2c * c3); g2[1] = c2;
        c3 = a2 * c2; g1[2] = c3; c2 = c1 / (c1 - alpha * c3); g2[2] = c2;
        g3[1] = -g2c * g2[1];
    } else {
        c3 = alpha * c2; g1[1] = c3; c2 = c1 / (c1 - alpha * c3); g2[1] = c2;
        g3[1] = -alpha * g2[1];
        c3 = alpha * c2; g1[2] = c3; c2 = c1 / (c1 - alpha * c3); g2[2] = c2;
    }
    for (j = 3; j <= n - 3; j++) {          /* Fortran ir=4..mynr-2 */
        c3 = alpha * c2; g1[j] = c3; c2 = c1 / (c1 - alpha * c3); g2[j] = c2;
    }
    for (j = 2; j <= n - 3; j++) g3[j] = -alpha * g2[j];
    if (hi_wall) {
        c3 = alpha * c2; g1[n - 2] = c3; c2 = c1 / (c1 - a2 * c3); g2[n - 2] = c2;
        c3 = g2c * c2; g1[n - 1] = c3; c2 = c1 / (c1 - up2 * c3); g2[n - 1] = c2;
        g3[n - 2] = -a2 * g2[n - 2];
        g3[n - 1] = -up2 * g2[n - 1];
    } else {
        c3 = alpha * c2; g1[n - 2] = c3; c2 = c1 / (c1 - alpha * c3); g2[n - 2] = c2;
        c3 = alpha * c2; g1[n - 1] = c3; c2 = c1 / (c1 - alpha * c3); g2[n - 1] = c2;
        g3[n - 2] = -alpha * g2[n - 2];
        g3[n - 1] = -alpha * g2[n - 1];
    }
    for (j = 0; j < n; j++) g1[j] = -g1[j];
}

/* derivatives.f90:194-287 (second derivative, alpha=2/11). */
void orc_factors_d2(int n, int lo_wall, int hi_wall, int iv, double *g1, double *g2, double *g3)
{
    const double alpha = 2.0 / 11.0;
    double up1 = 0, up2 = 0, a2, g2c, c1 = 1.0, c2 = 1.0, c3;
    int j;
    switch (iv) {
    case 0: up1 = 707.0 / 193.0; up2 = 707.0 / 193.0; break;
    case 1: up1 = 137.0 / 13.0; up2 = 137.0 / 13.0; break;
    case 2: up1 = 707.0 / 193.0; up2 = 137.0 / 13.0; break;
    case 3: up1 = 137.0 / 13.0; up2 = 707.0 / 193.0; break;
    default: up1 = 707.0 / 193.0; up2 = 707.0 / 193.0; break;
    }
    a2 = -0.35; g2c = 0.1;
    for (j = 0; j < n; j++) { g1[j] = 0; g2[j] = 0; g3[j] = 0; }
    g1[0] = c2; g2[0] = c2; g3[0] = c2;
    if (lo_wall) {
        c3 = up1; g1[1] = c3; c2 = c1 / (c1 - g2c * c3); g2[1] = c2;
        c3 = a2 * c2; g1[2] = c3; c2 = c1 / (c1 - alpha * c3); g2[2] = c2;
        g3[1] = -g2c * g2[1];
    } else {
        c3 = alpha * c2; g1[1] = c3; c2 = c1 / (c1 - alpha * c3); g2[1] = c2;
        g3[1] = -alpha * g2[1];
        c3 = alpha * c2; g1[2] = c3; c2 = c1 / (c1 - alpha * c3); g2[2] = c2;
    }
    for (j = 3; j <= n - 3; j++) {
        c3 = alpha * c2; g1[j] = c3; c2 = c1 / (c1 - alpha * c3); g2[j] = c2;
    }
    for (j = 2; j <= n - 3; j++) g3[j] = -alpha * g2[j];
    g2c = a2;   /* -0.35: lower coefficient of row n-2 */
    a2 = 0.1;   /* upper coefficient of row n-2 */
    if (hi_wall) {
        c3 = alpha * c2; g1[n - 2] = c3; c2 = c1 / (c1 - g2c * c3); g2[n - 2] = c2;
        c3 = a2 * c2; g1[n - 1] = c3; c2 = c1 / (c1 - up2 * c3); g2[n - 1] = c2;
        g3[n - 2] = -g2c * g2[n - 2];
        g3[n - 1] = -up2 * g2[n - 1];
    } else {
        c3 = alpha * c2; g1[n - 2] = c3; c2 = c1 / (c1 - alpha * c3); g2[n - 2] = c2;
        c3 = alpha * c2; g1[n - 1] = c3; c2 = c1 / (c1 - alpha * c3); g2[n - 1] = c2;
        g3[n - 2] = -alpha * g2[n - 2];
        g3[n - 1] = -alpha * g2[n - 1];
    }
    for (j = 0; j < n; j++) g1[j] = -g1[j];
}

/* ---------------------------------------------------------------------------------------
 * Right-hand-side rows of the first-derivative system, in index units (scaled by dri at the
 * end of the solve).  x: contiguous pencil of n values.  b1[0..3], b2[0..3]: halo ghosts
 * (b1[3] and b2[0] adjacent to the pencil) or, at a wall, b1[0]/b2[0] = the prescribed
 * physical first derivative (used when the code asks for Neumann at that end).
 * derivatives.f90:564-728.
 * ------------------------------------------------------------------------------------- */
static void rows_d1(int n, const double *x, int lo_wall, int hi_wall, int bdry,
                    const double *b1, const double *b2, double dri, double *d)
{
    int j;
    double c1, c2, c3, c4, c5, cst, beta1, alpha1, alpha2, gamma2;
    if (!lo_wall) {                              /* Case (1)/(3): lower halo, :565-583 */
        c1 = -1.0 / 180.0; c2 = 9.0; c3 = 45.0; c4 = 7.0 / 9.0; c5 = 1.0 / 36.0;
        c2 = c1 * c2; c3 = c1 * c3;
        cst = -(c3 + c5); beta1 = c4 - c2; alpha1 = c1 + c5;
        c1 = -c1; c4 = -c4;
        d[0] = c1 * b1[0] + c2 * b1[1] + cst * b1[2] + c4 * b1[3] + c3 * x[0] + beta1 * x[1] + alpha1 * x[2];
        d[1] = (7.0 / 9.0) * (x[2] - x[0]) + (1.0 / 36.0) * (x[3] - b1[3]);
    } else {                                     /* Case (2): lower wall, :623-662 */
        d[0] = 0.0;
        if (bdry % 2 == 0) d[0] = b1[0] / dri;
        if (bdry == 0) d[0] = 0.0;
        c1 = -43.0 / 96.0; c2 = -5.0 / 6.0; c3 = 9.0 / 8.0; c4 = 1.0 / 6.0; c5 = -1.0 / 96.0;
        d[1] = c1 * x[0] + c2 * x[1] + c3 * x[2] + c4 * x[3] + c5 * x[4];
        if (bdry == 1 || bdry == 3) {
            alpha2 = 0.75; gamma2 = 0.125; beta1 = 3.0;
            cst = 1.0 / (alpha2 - beta1 * gamma2);
            alpha2 = alpha2 * cst; beta1 = -beta1 * cst;
            c1 = -10.0 / 3.0; c2 = -3.0; c3 = 6.0; c4 = 1.0 / 3.0;
            d[0] = c1 * x[0] + c2 * x[1] + c3 * x[2] + c4 * x[3];
            d[0] = alpha2 * d[0] + beta1 * d[1];
        }
    }
    if (!hi_wall) {                              /* Case (2)/(3): upper halo, :664-680 */
        c1 = 7.0 / 9.0; c2 = 1.0 / 36.0; c3 = -1.0 / 180.0; c4 = 9.0; c5 = 45.0;
        d[n - 2] = c1 * (x[n - 1] - x[n - 3]) + c2 * (b2[0] - x[n - 4]);
        cst = c2 + c3 * c5; beta1 = c3 * c4 - c1; alpha1 = -(c2 + c3);
        c4 = -c3 * c4; c5 = -c3 * c5;
        d[n - 1] = c1 * b2[0] + cst * b2[1] + c4 * b2[2] + c3 * b2[3] + beta1 * x[n - 2] + alpha1 * x[n - 3] + c5 * x[n - 1];
    } else {                                     /* Case (1): upper wall, :584-620 */
        c1 = 43.0 / 96.0; c2 = 5.0 / 6.0; c3 = -9.0 / 8.0; c4 = -1.0 / 6.0; c5 = 1.0 / 96.0;
        d[n - 2] = c1 * x[n - 1] + c2 * x[n - 2] + c3 * x[n - 3] + c4 * x[n - 4] + c5 * x[n - 5];
        if (bdry == 3 || bdry == 4) d[n - 1] = b2[0] / dri;
        else d[n - 1] = 0.0;
        if (bdry == 1 || bdry == 2) {
            alpha2 = 0.75; gamma2 = 0.125; beta1 = 3.0;
            cst = 1.0 / (alpha2 - beta1 * gamma2);
            alpha2 = alpha2 * cst; beta1 = -beta1 * cst;
            c1 = 10.0 / 3.0; c2 = 3.0; c3 = -6.0; c4 = -1.0 / 3.0;
            d[n - 1] = c1 * x[n - 1] + c2 * x[n - 2] + c3 * x[n - 3] + c4 * x[n - 4];
            d[n - 1] = alpha2 * d[n - 1] + beta1 * d[n - 2];
        }
    }
    c4 = 7.0 / 9.0; c5 = 1.0 / 36.0;             /* interior, :725-728 */
    for (j = 2; j <= n - 3; j++) d[j] = c4 * (x[j + 1] - x[j - 1]) + c5 * (x[j + 2] - x[j - 2]);
}

/* Second-derivative rows, derivatives.f90:816-938.  hneu = the inverse spacing the Fortran
 * divides a prescribed first derivative by (dri for r; also dri for theta, T6). */
static void rows_d2(int n, const double *x, int lo_wall, int hi_wall, int bdry,
                    const double *b1, const double *b2, double hneu, double *d)
{
    int j;
    double c1, c2, c3, c4, c5, c6, c7;
    if (!lo_wall) {                              /* halo rows, :817-826 */
        c1 = 12.0 / 11.0; c2 = 3.0 / 44.0; c3 = -1.0 / 990.0; c4 = 2.0; c5 = -27.0; c6 = 270.0; c7 = -490.0;
        d[0] = c3 * (c4 * (b1[0] + x[2]) + c5 * (b1[1] + x[1]) + c6 * (b1[2] + x[0]) + c7 * b1[3]);
        d[0] = d[0] + c1 * (x[1] - 2.0 * x[0] + b1[3]) + c2 * (b1[2] - 2.0 * x[0] + x[2]);
        d[1] = c1 * (x[2] - 2.0 * x[1] + x[0]) + c2 * (b1[3] - 2.0 * x[1] + x[3]);
    } else {                                     /* lower wall, :867-900 */
        if (bdry % 2 == 0) {
            c1 = -575.0 / 193.0 / hneu; c2 = -5827.0 / 2316.0; c3 = 1987.0 / 1158.0;
            c4 = 465.0 / 386.0; c5 = -547.0 / 1158.0; c6 = 157.0 / 2316.0;
            if (bdry == 0) c1 = 0.0;
            d[0] = c1 * b1[0] + c2 * x[1] + c3 * x[2] + c4 * x[3] + c5 * x[4] + c6 * x[5];
        } else {
            c1 = 1955.0 / 156.0; c2 = -4057.0 / 156.0; c3 = 1117.0 / 78.0; c4 = -55.0 / 78.0;
            c5 = -29.0 / 156.0; c6 = 7.0 / 156.0;
            d[0] = c1 * x[0] + c2 * x[1] + c3 * x[2] + c4 * x[3] + c5 * x[4] + c6 * x[5];
        }
        c1 = 1.2375; c2 = -3.0; c3 = 2.325; c4 = -0.6; c5 = 0.0375;
        d[1] = c1 * x[0] + c2 * x[1] + c3 * x[2] + c4 * x[3] + c5 * x[4];
    }
    if (!hi_wall) {                              /* halo rows, :910-917 */
        c1 = 12.0 / 11.0; c2 = 3.0 / 44.0; c3 = -1.0 / 990.0; c4 = 2.0; c5 = -27.0; c6 = 270.0; c7 = -490.0;
        d[n - 2] = c1 * (x[n - 1] - 2.0 * x[n - 2] + x[n - 3]) + c2 * (x[n - 4] - 2.0 * x[n - 2] + b2[0]);
        d[n - 1] = c1 * (b2[0] - 2.0 * x[n - 1] + x[n - 2]) + c2 * (x[n - 3] - 2.0 * x[n - 1] + b2[1]);
        d[n - 1] = d[n - 1] + c3 * (c4 * (x[n - 3] + b2[3]) + c5 * (x[n - 2] + b2[2]) + c6 * (x[n - 1] + b2[1]) + c7 * b2[0]);
    } else {                                     /* upper wall, :828-864 */
        c1 = 1.2375; c2 = -3.0; c3 = 2.325; c4 = -0.6; c5 = 0.0375;
        d[n - 2] = c1 * x[n - 1] + c2 * x[n - 2] + c3 * x[n - 3] + c4 * x[n - 4] + c5 * x[n - 5];
        if (bdry == 1 || bdry == 2) {
            c1 = 1955.0 / 156.0; c2 = -4057.0 / 156.0; c3 = 1117.0 / 78.0; c4 = -55.0 / 78.0;
            c5 = -29.0 / 156.0; c6 = 7.0 / 156.0;
            d[n - 1] = c1 * x[n - 1] + c2 * x[n - 2] + c3 * x[n - 3] + c4 * x[n - 4] + c5 * x[n - 5] + c6 * x[n - 6];
        } else {
            c1 = 575.0 / 193.0 / hneu; c2 = -5827.0 / 2316.0; c3 = 1987.0 / 1158.0;
            c4 = 465.0 / 386.0; c5 = -547.0 / 1158.0; c6 = 157.0 / 2316.0;
            if (bdry == 0) c1 = 0.0;
            d[n - 1] = c1 * b2[0] + c2 * x[n - 2] + c3 * x[n - 3] + c4 * x[n - 4] + c5 * x[n - 5] + c6 * x[n - 6];
        }
    }
    c1 = 12.0 / 11.0; c2 = 3.0 / 44.0;          /* interior, :935-938 */
    for (j = 2; j <= n - 3; j++)
        d[j] = c1 * (x[j + 1] - 2.0 * x[j] + x[j - 1]) + c2 * (x[j - 2] - 2.0 * x[j] + x[j + 2]);
}

/* Forward sweep + back substitution of Dbydr/Dbydt (derivatives.f90:730-795): rows 1, n-2,
 * n-1 use the gam3 form, interior rows the gam2*(d - alpha*y) form. */
static void sweep_rt(int n, double alpha, const double *g1, const double *g2, const double *g3,
                     double *d, double *y, double scale, double *out, long os)
{
    int j;
    double am = -alpha;
    y[0] = g2[0] * d[0];
    y[1] = g2[1] * d[1] + g3[1] * y[0];
    for (j = 2; j <= n - 3; j++) y[j] = g2[j] * (d[j] + am * y[j - 1]);
    y[n - 2] = g2[n - 2] * d[n - 2] + g3[n - 2] * y[n - 3];
    y[n - 1] = g2[n - 1] * d[n - 1] + g3[n - 1] * y[n - 2];
    d[n - 1] = y[n - 1];
    for (j = n - 2; j >= 0; j--) d[j] = y[j] + g1[j + 1] * d[j + 1];
    for (j = 0; j < n; j++) out[(long)j * os] = scale * d[j];
}

/* Dbydp's sweep (derivatives.f90:1535-1590): gam2*(d - alpha*y) on every row after the first. */
static void sweep_p(int n, double alpha, const double *g1, const double *g2,
                    double *d, double *y, double scale, double *out, long os)
{
    int j;
    double am = -alpha;
    y[0] = g2[0] * d[0];
    for (j = 1; j < n; j++) y[j] = g2[j] * (d[j] + am * y[j - 1]);
    d[n - 1] = y[n - 1];
    for (j = n - 2; j >= 0; j--) d[j] = y[j] + g1[j + 1] * d[j + 1];
    for (j = 0; j < n; j++) out[(long)j * os] = scale * d[j];
}

/* ---------------------------------------------------------------------------------------
 * Public pencil operator.
 *   order 1/2; x[j*xs] for j<n; out[j*os].
 *   lo_wall/hi_wall: 1 = physical wall, 0 = 4-deep halo from b1/b2.
 *   phi_form: use Dbydp's sweep (both ends must be halos).
 *   scale: the Fortran dri/dti/dpi (order 1) or its square (order 2).
 *   hneu: inverse spacing used to convert a prescribed wall derivative (order 1: equals scale;
 *         order 2: dri).
 * ------------------------------------------------------------------------------------- */
int orc_deriv_pencil(int order, int n, const double *x, long xs, double *out, long os,
                     int lo_wall, int hi_wall, int bdry, const double *b1, const double *b2,
                     double scale, double hneu, int phi_form)
{
    double *xc, *d, *y, *g1, *g2, *g3;
    int j;
    if (n < 8 || (order != 1 && order != 2) || bdry < 0 || bdry > 4) return 22;
    xc = (double *)malloc(sizeof(double) * 6 * n);
    if (!xc) return 12;
    d = xc + n; y = d + n; g1 = y + n; g2 = g1 + n; g3 = g2 + n;
    for (j = 0; j < n; j++) xc[j] = x[(long)j * xs];
    if (order == 1) {
        orc_factors_d1(n, lo_wall, hi_wall, bdry, g1, g2, g3);
        rows_d1(n, xc, lo_wall, hi_wall, bdry, b1, b2, hneu, d);
        if (phi_form) sweep_p(n, ALPHA1, g1, g2, d, y, scale, out, os);
        else sweep_rt(n, ALPHA1, g1, g2, g3, d, y, scale, out, os);
    } else {
        orc_factors_d2(n, lo_wall, hi_wall, bdry, g1, g2, g3);
        rows_d2(n, xc, lo_wall, hi_wall, bdry, b1, b2, hneu, d);
        if (phi_form) sweep_p(n, ALPHA2, g1, g2, d, y, scale, out, os);
        else sweep_rt(n, ALPHA2, g1, g2, g3, d, y, scale, out, os);
    }
    free(xc);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * CSS-Python compact_fd6 (derivatives.py:32-185) + tri_solver its=0 (linalg.py:94-111).
 * darr (length n) holds the caller's preloaded values on entry (Neumann values f'/dri at
 * darr[0]/darr[n-1]) and, as in the Python, the assembled right-hand side on exit.
 * b1/b2 point to 4 values or are NULL (the Python's scalar case).  Result = solution*dri.
 * ------------------------------------------------------------------------------------- */
int orc_compact_fd6_py(double dri, int n1, const double *arr, const double *b1, const double *b2,
                       int ibc, int dtype, double *darr, double *result)
{
    double *uppr, *lowr, *diag, *cp, *dp;
    double dli, c1, c2, c3, c4, c5, alpha, alpha2, gamma2, alpha1, beta1, cst, up1;
    int i;
    if (n1 < 5) return 22;   /* the rows reach arr[4] and arr[n1-5] (derivatives.py:65-66, :118) */
    if ((dtype == 1 || dtype == 3) && !b1) return 22;
    if ((dtype == 2 || dtype == 3) && !b2) return 22;
    uppr = (double *)malloc(sizeof(double) * 5 * n1);
    if (!uppr) return 12;
    lowr = uppr + n1; diag = lowr + n1; cp = diag + n1; dp = cp + n1;

    dli = (ibc == 5) ? (double)n1 : (double)(n1 - 1);
    c1 = dli * 7. / 9.;
    c2 = dli / 36.;
    if (dtype == 1 || dtype == 3) {
        darr[0] = c1 * (arr[1] - b1[3]) + c2 * (arr[2] - b1[2]);
        darr[1] = c1 * (arr[2] - arr[0]) + c2 * (arr[3] - b1[3]);
    }
    if (dtype == 2 || dtype == 3) {
        darr[n1 - 2] = c1 * (arr[n1 - 1] - arr[n1 - 3]) + c2 * (b2[0] - arr[n1 - 4]);
        darr[n1 - 1] = c1 * (b2[0] - arr[n1 - 2]) + c2 * (b2[1] - arr[n1 - 3]);
    }
    for (i = 2; i <= n1 - 3; i++) darr[i] = c1 * (arr[i + 1] - arr[i - 1]) + c2 * (arr[i + 2] - arr[i - 2]);

    alpha = 1. / 3.;
    for (i = 0; i < n1; i++) { diag[i] = 1.0; uppr[i] = alpha; lowr[i] = alpha; }

    if (dtype < 3) {
        if (dtype == 1 || dtype == 0) {
            c1 = -43. / 96 * dli; c2 = -5. / 6. * dli; c3 = 9. / 8. * dli; c4 = 1. / 6. * dli; c5 = -1. / 96. * dli;
            darr[n1 - 2] = -(c1 * arr[n1 - 1] + c2 * arr[n1 - 2] + c3 * arr[n1 - 3] + c4 * arr[n1 - 4] + c5 * arr[n1 - 5]);
            c1 = -10. / 3. * dli; c2 = -3. * dli; c3 = 6. * dli; c4 = 1. / 3. * dli;
            if (ibc == 0) darr[n1 - 1] = 0.;
            if (ibc == 1 || ibc == 2)
                darr[n1 - 1] = -(c1 * arr[n1 - 1] + c2 * arr[n1 - 2] + c3 * arr[n1 - 3] + c4 * arr[n1 - 4]);
            if (dtype == 1) {
                darr[0] += -alpha * (-b1[0] + 9. * b1[1] - 45. * b1[2] + 45. * arr[0] - 9. * arr[1] + arr[2]);
                darr[0] *= dli / 60.;
            }
            alpha2 = 3. / 4.; gamma2 = 1. / 8.; alpha1 = 6.; beta1 = 3.;
            cst = 1. / (alpha2 - beta1 * gamma2);
            up1 = cst * (alpha1 * alpha2 - beta1);
            if (ibc == 1 || ibc == 2) darr[n1 - 1] = cst * (alpha2 * darr[n1 - 1] - beta1 * darr[n1 - 2]);
            lowr[n1 - 2] = alpha2;
            uppr[n1 - 2] = gamma2;
            if (ibc >= 0) {
                if (ibc != 1 && ibc != 2) lowr[n1 - 1] = 0.;
                else lowr[n1 - 1] = up1;
            }
        }
        if (dtype == 2 || dtype == 0) {
            c1 = -43. / 96 * dli; c2 = -5. / 6. * dli; c3 = 9. / 8. * dli; c4 = 1. / 6. * dli; c5 = -1. / 96. * dli;
            darr[1] = c1 * arr[0] + c2 * arr[1] + c3 * arr[2] + c4 * arr[3] + c5 * arr[4];
            c1 = -10. / 3. * dli; c2 = -3. * dli; c3 = 6. * dli; c4 = 1. / 3. * dli;
            if (ibc == 0) darr[0] = 0.;
            if (ibc == 1 || ibc == 3) darr[0] = c1 * arr[0] + c2 * arr[1] + c3 * arr[2] + c4 * arr[3];
            if (dtype == 2) {
                darr[n1 - 1] += -alpha * (-arr[n1 - 3] + 9. * arr[n1 - 2] - 45. * arr[n1 - 1] + 45. * b2[1] - 9. * b2[2] + b2[3]);
                darr[n1 - 1] *= dli / 60.;
            }
            alpha2 = 3. / 4.; gamma2 = 1. / 8.; alpha1 = 6.; beta1 = 3.;
            cst = 1. / (alpha2 - beta1 * gamma2);
            up1 = cst * (alpha1 * alpha2 - beta1);
            /* Python's ibc % 2 for ibc >= 0 */
            if (((ibc % 2) + 2) % 2 != 0) darr[0] = cst * (alpha2 * darr[0] - beta1 * darr[1]);
            if (((ibc % 2) + 2) % 2 == 0) uppr[0] = 0.;
            else uppr[0] = up1;
            uppr[1] = alpha2;
            lowr[1] = gamma2;
        }
        lowr[0] = 0.;
        uppr[n1 - 1] = 0.;
    } else {
        lowr[0] = 0.;
        uppr[n1 - 1] = 0.;
        darr[0] += -alpha * (-b1[0] + 9. * b1[1] - 45. * b1[2] + 45. * arr[0] - 9. * arr[1] + arr[2]);
        darr[0] *= dli / 60.;
        darr[n1 - 1] += -alpha * (-arr[n1 - 3] + 9. * arr[n1 - 2] - 45. * arr[n1 - 1] + 45. * b2[1] - 9. * b2[2] + b2[3]);
        darr[n1 - 1] *= dli / 60.;
    }

    /* tri_solver its=0 */
    cp[0] = uppr[0] / diag[0];
    dp[0] = darr[0] / diag[0];
    for (i = 1; i <= n1 - 2; i++) {
        cp[i] = uppr[i] / (diag[i] - cp[i - 1] * lowr[i]);
        dp[i] = (darr[i] - dp[i - 1] * lowr[i]) / (diag[i] - cp[i - 1] * lowr[i]);
    }
    dp[n1 - 1] = (darr[n1 - 1] - dp[n1 - 2] * lowr[n1 - 1]) / (diag[n1 - 1] - cp[n1 - 2] * lowr[n1 - 1]);
    result[n1 - 1] = dp[n1 - 1];
    for (i = n1 - 2; i >= 0; i--) result[i] = dp[i] - cp[i] * result[i + 1];
    for (i = 0; i < n1; i++) result[i] *= dri;
    free(uppr);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * Whole-field operator on the reference layout f(nth, nphi, nr) (theta fastest), one call =
 * one Dbydr / Dbydt / Dbydp over every pencil of one rank's subdomain.
 *   axis 0 = r, 1 = theta, 2 = phi.
 *   lo_wall/hi_wall as in orc_deriv_pencil (phi: both 0).
 *   ghost_lo/ghost_hi: the reference halo buffers for that axis, or the wall values:
 *     r:     g[(k*nphi + ip)*nth + it]   (binfo(1,side)%bdata(nth,nphi,4))
 *     theta: g[(ir*nphi + ip)*4 + k]     (binfo(2,side)%bdata(4,nphi,nr))
 *     phi:   g[(ir*4 + k)*nth + it]      (binfo(3,side)%bdata(nth,4,nr))
 *   NULL ghosts on a wall end mean a zero wall value; NULL ghosts on the phi axis mean the
 *   single-subdomain periodic wrap b1 = f(n-4..n-1), b2 = f(0..3).
 *   metric: optional per-row multiplier along the axis applied after the scale (dxdr for r).
 * ------------------------------------------------------------------------------------- */
int orc_deriv_field(int axis, int order, int nth, int nphi, int nr, const double *f, double *out,
                    int lo_wall, int hi_wall, int bdry, const double *ghost_lo, const double *ghost_hi,
                    double scale, double hneu, const double *metric)
{
    long sth = 1, sph = nth, sr = (long)nth * nphi;
    int n, npen, p, k, rc = 0;
    if (axis == 0) { n = nr; npen = nth * nphi; }
    else if (axis == 1) { n = nth; npen = nphi * nr; }
    else { n = nphi; npen = nth * nr; }
#pragma omp parallel for schedule(static) reduction(| : rc)
    for (p = 0; p < npen; p++) {
        long base, xs;
        int it, ip, ir, j;
        double b1[4] = {0, 0, 0, 0}, b2[4] = {0, 0, 0, 0};
        if (axis == 0) {
            it = p % nth; ip = p / nth;
            base = it * sth + ip * sph; xs = sr;
            for (k = 0; k < 4; k++) {
                if (ghost_lo) b1[k] = ghost_lo[((long)k * nphi + ip) * nth + it];
                if (ghost_hi) b2[k] = ghost_hi[((long)k * nphi + ip) * nth + it];
            }
        } else if (axis == 1) {
            ip = p % nphi; ir = p / nphi;
            base = ip * sph + ir * sr; xs = sth;
            for (k = 0; k < 4; k++) {
                if (ghost_lo) b1[k] = ghost_lo[((long)ir * nphi + ip) * 4 + k];
                if (ghost_hi) b2[k] = ghost_hi[((long)ir * nphi + ip) * 4 + k];
            }
        } else {
            it = p % nth; ir = p / nth;
            base = it * sth + ir * sr; xs = sph;
            for (k = 0; k < 4; k++) {
                if (ghost_lo) b1[k] = ghost_lo[((long)ir * 4 + k) * nth + it];
                else b1[k] = f[base + (long)(n - 4 + k) * xs];
                if (ghost_hi) b2[k] = ghost_hi[((long)ir * 4 + k) * nth + it];
                else b2[k] = f[base + (long)k * xs];
            }
        }
        rc |= orc_deriv_pencil(order, n, f + base, xs, out + base, xs,
                               axis == 2 ? 0 : lo_wall, axis == 2 ? 0 : hi_wall, bdry, b1, b2,
                               scale, hneu, axis == 2);
        if (metric)
            for (j = 0; j < n; j++) out[base + (long)j * xs] *= metric[j];
    }
    return rc;
}
