/*
 * CPU oracle: FP64 restatement of the CSS MHD right-hand side, Bound, Compute_B and the
 * classical RK4 step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_deriv.c).  The parameter structs come from the public
 * header include/cssx.h (data layout only).
 *
 * Restated (reference file:line):
 *   RHS                    rhs_mpi.f90:279-783  (stratified, do_radheat=F; Laplacian and
 *                          SLaplacian branches included)
 *   Laplacian_Cross_Terms  rhs_mpi.f90:2070-2296
 *   VectorCurl             derivatives.f90:1894-2104
 *   Compute_B/VectorCurlA  derivatives.f90:1830-1892, :2106-2294
 *   Source_Theta_Terms     rhs_mpi.f90:786-1029
 *   Phi_Terms              rhs_mpi.f90:1032-1261
 *   Radial_Terms           rhs_mpi.f90:1264-1519
 *   Density_Terms_Theta    rhs_mpi.f90:1522-1605
 *   Density_Terms_Radial   rhs_mpi.f90:1710-1855
 *   Induction_Terms        rhs_mpi.f90:2298-2366 (u x B; the div B diagnostic is separate)
 *   wall ghost values      rhs_mpi.f90:410-455, :578-637
 *   Bound                  rhs_mpi.f90:2395-2491
 *   Step_RK4               step_rk4.f90:339-493 (SLDiffusion branch: Euler diffusion in stage 1)
 *   Step_RK3_TVD           step_rk4.f90:166-335
 *   Step_RK2_TVD           step_rk4.f90:31-162
 *   Compute_Diffusion      diffusion.f90:70-474, Diffusion_R/Th/Phi :476-1495 (SLD)
 *   Compute_Step_Size      rhs_mpi.f90:92-275 (stratified branch: CFL and conservation totals)
 * Parity status: the RHS is pinned by a manufactured solution, hydrostatic balance and wall
 * invariants, its Laplacian cross-term wall values by stress-free manufactured states; the
 * diffusion by conservation, the minmod identity, continuum geometry, the corrections'
 * energy/momentum identities and, for its wall ghost rules, the closed-form limiter response of
 * fields linear up to a wall; Compute_Step_Size by closed forms on uniform grids; the steppers by
 * their order of accuracy.
 * The subdomain has physical walls at both r ends and both theta ends; each wall is handled as
 * the rank touching it would (the Fortran's If (myr1==1) ... ElseIf (myr2==nr) chains need
 * two ranks per direction to visit both walls; here both are applied).
 *
 * Layout: field f[(ir*nphi + ip)*nth + it]; state = nv such fields, variable iv at offset
 * (iv-1)*N.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/cssx.h"

int orc_deriv_field(int axis, int order, int nth, int nphi, int nr, const double *f, double *out,
                    int lo_wall, int hi_wall, int bdry, const double *ghost_lo, const double *ghost_hi,
                    double scale, double hneu, const double *metric);

#define PI_ 3.1415926535897932384626433832795028841972
#define ONE_OVER_4PI (1.0 / PI_ / 4.0)

typedef struct {
    int nr, nth, nphi;
    long N;
    double dri, dti, dpi;
    cssx_mhd_params p;
    double gamm1, Cv, pconst, invpconst, invtconst, invCv;
    const double *r, *dr, *dxdr, *theta, *gravity, *eps, *dsdr1, *dsdr2, *sr1, *sr2;
    const double *mur, *dlnmu, *ks, *dlnks, *etar, *d2xdr2;
    int lap, slap;            /* Laplacian; Laplacian or SLaplacian (the thermal diffusion) */
    double *sines, *cosines;
} orc_mhd;

static void setup(orc_mhd *m, const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr)
{
    int it;
    m->nr = g->nr; m->nth = g->nth; m->nphi = g->nphi;
    m->N = (long)g->nr * g->nphi * g->nth;
    m->dri = g->dri; m->dti = g->dti; m->dpi = g->dpi;
    m->p = *p;
    /* rhs_mpi.f90:59-73, stratified branch */
    m->Cv = p->Cp / p->gamma;
    m->invCv = 1.0 / m->Cv;
    m->gamm1 = p->gamma - 1.0;
    m->pconst = m->gamm1 * p->Cp;
    m->invpconst = (1.0 / p->Cp) / m->gamm1;
    m->invtconst = m->invCv / m->gamm1;
    m->r = pr->r; m->dr = pr->dr; m->dxdr = pr->dxdr; m->theta = pr->theta;
    m->gravity = pr->gravity; m->eps = pr->eps;
    m->dsdr1 = pr->dsdr1; m->dsdr2 = pr->dsdr2; m->sr1 = pr->sr1; m->sr2 = pr->sr2;
    m->mur = pr->mur; m->dlnmu = pr->dlnmu; m->ks = pr->ks; m->dlnks = pr->dlnks; m->etar = pr->etar;
    m->d2xdr2 = pr->d2xdr2;
    m->lap = p->laplacian != 0;
    m->slap = p->laplacian || p->slaplacian;
    m->sines = (double *)malloc(sizeof(double) * 2 * g->nth);
    m->cosines = m->sines + g->nth;
    for (it = 0; it < g->nth; it++) { m->sines[it] = sin(pr->theta[it]); m->cosines[it] = cos(pr->theta[it]); }
}

#define IDX(m, it, ip, ir) ((((long)(ir) * (m)->nphi + (ip)) * (m)->nth) + (it))

/* derivative along theta with wall values per (ip, ir) pencil (NULL = 0) */
static void dth(const orc_mhd *m, const double *x, double *out, int code, const double *wlo, const double *whi,
                double *glo, double *ghi)
{
    long k, np = (long)m->nphi * m->nr;
    for (k = 0; k < np; k++) {
        glo[k * 4] = wlo ? wlo[k] : 0.0;
        ghi[k * 4] = whi ? whi[k] : 0.0;
    }
    orc_deriv_field(1, 1, m->nth, m->nphi, m->nr, x, out, 1, 1, code, glo, ghi, m->dti, m->dti, NULL);
}

static void dph(const orc_mhd *m, const double *x, double *out)
{
    orc_deriv_field(2, 1, m->nth, m->nphi, m->nr, x, out, 0, 0, 0, NULL, NULL, m->dpi, m->dpi, NULL);
}

/* derivative along r (x-space) times dx/dr; wall values per (it, ip) in x-space units */
static void dr_(const orc_mhd *m, const double *x, double *out, int code, const double *wlo, const double *whi,
                double *glo, double *ghi)
{
    long k, np = (long)m->nphi * m->nth;
    for (k = 0; k < np; k++) {
        glo[k] = wlo ? wlo[k] : 0.0;
        ghi[k] = whi ? whi[k] : 0.0;
    }
    orc_deriv_field(0, 1, m->nth, m->nphi, m->nr, x, out, 1, 1, code, glo, ghi, m->dri, m->dri, m->dxdr);
}

/* second derivatives (D2bydt2 / D2bydp2 / D2bydr2): scale the inverse spacing squared; a
 * prescribed wall derivative enters divided by the radial dri on theta as well (derivatives.f90
 * :1322, :1338).  d2r_ returns the x-space derivative (callers multiply by dxdr^2). */
static void d2th(const orc_mhd *m, const double *x, double *out, int code, const double *wlo, const double *whi,
                 double *glo, double *ghi)
{
    long k, np = (long)m->nphi * m->nr;
    for (k = 0; k < np; k++) {
        glo[k * 4] = wlo ? wlo[k] : 0.0;
        ghi[k * 4] = whi ? whi[k] : 0.0;
    }
    orc_deriv_field(1, 2, m->nth, m->nphi, m->nr, x, out, 1, 1, code, glo, ghi, m->dti * m->dti, m->dri, NULL);
}

static void d2ph(const orc_mhd *m, const double *x, double *out)
{
    orc_deriv_field(2, 2, m->nth, m->nphi, m->nr, x, out, 0, 0, 0, NULL, NULL, m->dpi * m->dpi, m->dri, NULL);
}

static void d2r_(const orc_mhd *m, const double *x, double *out, int code, const double *wlo, const double *whi,
                 double *glo, double *ghi)
{
    long k, np = (long)m->nphi * m->nth;
    for (k = 0; k < np; k++) {
        glo[k] = wlo ? wlo[k] : 0.0;
        ghi[k] = whi ? whi[k] : 0.0;
    }
    orc_deriv_field(0, 2, m->nth, m->nphi, m->nr, x, out, 1, 1, code, glo, ghi, m->dri * m->dri, m->dri, NULL);
}

/* Compute_B: B = curl A with wall zeroing; a, b point to 3 fields (theta, phi, r). */
static void compute_b(const orc_mhd *m, const double *a, double *b, double *work, double *work2, double *glo,
                      double *ghi, int zero_walls)
{
    const int nr = m->nr, nth = m->nth, nphi = m->nphi;
    const long N = m->N;
    const double *at = a, *ap = a + N, *ar = a + 2 * N;
    double *bt = b, *bp = b + N, *br = b + 2 * N;
    int it, ip, ir;
    long i;
    /* theta sweeps (derivatives.f90:2163-2178): B_r <- d(sin A_phi)/dth, B_phi <- -dA_r/dth */
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) work[IDX(m, it, ip, ir)] = m->sines[it] * ap[IDX(m, it, ip, ir)];
    dth(m, work, work2, m->p.ibct[11], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                br[i] = work2[i] * (1.0 / m->sines[it]) * (1.0 / m->r[ir]);
            }
    dth(m, ar, work2, m->p.ibct[12], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) bp[i] = -work2[i] * (1.0 / m->r[ir]);
    /* phi sweeps (:2186-2207) */
    dph(m, ar, work2);
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                bt[i] = (1.0 / m->r[ir]) * work2[i] * (1.0 / m->sines[it]);
            }
    dph(m, at, work2);
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                br[i] = br[i] - (1.0 / m->r[ir]) * work2[i] * (1.0 / m->sines[it]);
            }
    /* r sweeps of r*A_theta, r*A_phi (:2250-2293); wall A halos are zero, codes at/ap */
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) work[i] = m->r[ir] * at[i];
    dr_(m, work, work2, m->p.ibcr[10], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) bp[i] = bp[i] + work2[i] * (1.0 / m->r[ir]);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) work[i] = m->r[ir] * ap[i];
    dr_(m, work, work2, m->p.ibcr[11], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) bt[i] = bt[i] - work2[i] * (1.0 / m->r[ir]);
    /* wall conditions (derivatives.f90:1840-1883); VectorCurlA alone (Ohmic heating,
     * diffusion.f90:437) leaves them out */
    if (!zero_walls) return;
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++) {
            if (m->p.theta_perf_cond) {
                bt[IDX(m, 0, ip, ir)] = 0.0;
                bt[IDX(m, nth - 1, ip, ir)] = 0.0;
            } else {
                bp[IDX(m, 0, ip, ir)] = 0.0; br[IDX(m, 0, ip, ir)] = 0.0;
                bp[IDX(m, nth - 1, ip, ir)] = 0.0; br[IDX(m, nth - 1, ip, ir)] = 0.0;
            }
        }
    for (ip = 0; ip < nphi; ip++)
        for (it = 0; it < nth; it++) {
            if (m->p.bottom_radial_field) { bt[IDX(m, it, ip, 0)] = 0.0; bp[IDX(m, it, ip, 0)] = 0.0; }
            if (m->p.top_radial_field) { bt[IDX(m, it, ip, nr - 1)] = 0.0; bp[IDX(m, it, ip, nr - 1)] = 0.0; }
        }
}

#define ORC_C_LIGHT 2.99792458e10   /* c_light, constants.f90:7 */

/* VectorCurl (derivatives.f90:1894-2104): J = curl B into jt, jp, jr.  theta sweeps of sin*B_phi
 * and B_r with code 1 under Theta_Normal_Field, else 0 (their wall values then go unused);
 * phi sweeps periodic; r sweeps of r*B_theta, r*B_phi with codes ibcr(7), ibcr(8) and zero wall
 * values (the radial-field walls zero those halos, rhs_mpi.f90:596-601). */
static void vector_curl(const orc_mhd *m, const double *bt, const double *bp, const double *br, double *jt,
                        double *jp, double *jr, double *work, double *work1, double *work2, double *wlo,
                        double *whi, double *glo, double *ghi)
{
    const int nr = m->nr, nth = m->nth, nphi = m->nphi;
    const int code = m->p.theta_normal_field ? 1 : 0;
    int it, ip, ir;
    long i, kk;
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++) {
            kk = (long)ir * nphi + ip;
            /* sin_wall * binfo(2,*)(7) = sin * (-cot B_phi) with Theta_Normal_Field */
            wlo[kk] = m->p.theta_normal_field ? -m->cosines[0] * bp[IDX(m, 0, ip, ir)] : 0.0;
            whi[kk] = m->p.theta_normal_field ? -m->cosines[nth - 1] * bp[IDX(m, nth - 1, ip, ir)] : 0.0;
            for (it = 0; it < nth; it++) work[IDX(m, it, ip, ir)] = m->sines[it] * bp[IDX(m, it, ip, ir)];
        }
    dth(m, work, work1, code, wlo, whi, glo, ghi);
    dth(m, br, work2, code, NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                jr[i] = work1[i] * (1.0 / m->sines[it]) * (1.0 / m->r[ir]);
                jp[i] = -work2[i] * (1.0 / m->r[ir]);
            }
    dph(m, br, work1);
    dph(m, bt, work2);
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                jt[i] = (1.0 / m->r[ir]) * work1[i] * (1.0 / m->sines[it]);
                jr[i] = jr[i] - (1.0 / m->r[ir]) * work2[i] * (1.0 / m->sines[it]);
            }
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) work[i] = m->r[ir] * bt[i];
    dr_(m, work, work1, m->p.ibcr[7], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) jp[i] = jp[i] + work1[i] * (1.0 / m->r[ir]);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) work[i] = m->r[ir] * bp[i];
    dr_(m, work, work1, m->p.ibcr[8], NULL, NULL, glo, ghi);
    for (ir = 0; ir < nr; ir++)
        for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) jt[i] = jt[i] - work1[i] * (1.0 / m->r[ir]);
}

/* When set, orc_rhs copies the six Laplacian cross-term sweeps here (6 fields): the r sweeps
 * d(r vars2(k))/dr, k = 1..4, then the theta sweeps d(sin vars2(4))/dth, d(sin vars2(5))/dth. */
static double *g_cross_dump = NULL;

/* RHS on state v (nv fields with A in 6-8) -> dv (nv fields); bout (3 fields) optional. */
int orc_rhs(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, const double *v,
            double *dv, double *bout)
{
    orc_mhd M, *m = &M;
    long N, i, np;
    int nr, nth, nphi, it, ip, ir, nv;
    double *B, *T, *ir_, *work, *work1, *work2, *wlo, *whi, *glo, *ghi;
    const double *rho, *w, *vv, *u, *s, *bt, *bp, *br;
    double *drho, *dw, *dvv, *du, *ds;
    double *v2[5] = {NULL, NULL, NULL, NULL, NULL};   /* the Laplacian branches' vars2 */
    setup(m, g, p, pr);
    if (m->slap && (!pr->ks || !pr->dlnks || !pr->d2xdr2)) return 22;
    if (m->lap && (!pr->mur || !pr->dlnmu || !pr->etar)) return 22;
    N = m->N; nr = m->nr; nth = m->nth; nphi = m->nphi;
    nv = p->magnetic ? 8 : 5;
    np = (long)nphi * (nr > nth ? nr : nth) * 4 + 16;
    B = (double *)calloc(3 * N, sizeof(double));
    T = (double *)malloc(sizeof(double) * N);
    ir_ = (double *)malloc(sizeof(double) * N);
    work = (double *)malloc(sizeof(double) * N);
    work1 = (double *)malloc(sizeof(double) * N);
    work2 = (double *)malloc(sizeof(double) * N);
    wlo = (double *)calloc((size_t)nphi * (nr > nth ? nr : nth) + 16, sizeof(double));
    whi = (double *)calloc((size_t)nphi * (nr > nth ? nr : nth) + 16, sizeof(double));
    glo = (double *)calloc(np * 4 > 4L * nth * nphi ? np * 4 : 4L * nth * nphi, sizeof(double));
    ghi = (double *)calloc(np * 4 > 4L * nth * nphi ? np * 4 : 4L * nth * nphi, sizeof(double));
    if (!B || !T || !ir_ || !work || !work1 || !work2 || !wlo || !whi || !glo || !ghi) return 12;
    if (m->lap)
        for (i = 0; i < 5; i++)
            if (!(v2[i] = (double *)calloc(N, sizeof(double)))) return 12;
    rho = v; w = v + N; vv = v + 2 * N; u = v + 3 * N; s = v + 4 * N;
    drho = dv; dw = dv + N; dvv = dv + 2 * N; du = dv + 3 * N; ds = dv + 4 * N;
    memset(dv, 0, sizeof(double) * nv * N);

    /* Compute_B(1): B replaces A inside the RHS */
    if (p->magnetic) {
        compute_b(m, v + 5 * N, B, work, work2, glo, ghi, 1);
        if (bout) memcpy(bout, B, sizeof(double) * 3 * N);
    }
    bt = B; bp = B + N; br = B + 2 * N;

    /* temperature, 1/rho (rhs_mpi.f90:377-387) */
    for (i = 0; i < N; i++) {
        double tmp = pow(rho[i], m->gamm1);
        T[i] = m->invtconst * tmp;
        tmp = s[i] * m->invCv;
        tmp = exp(tmp);
        T[i] = T[i] * tmp;
        ir_[i] = 1.0 / rho[i];
    }

    /* ---- Source_Theta_Terms (rhs_mpi.f90:786-1029) ---- */
    {
        const double om2 = 2.0 * p->omega0;
        long kk;
        /* theta wall values of v: cot(theta)*v (rhs_mpi.f90:414, :441) */
        for (ir = 0; ir < nr; ir++)
            for (ip = 0; ip < nphi; ip++) {
                kk = (long)ir * nphi + ip;
                wlo[kk] = m->cosines[0] / m->sines[0] * vv[IDX(m, 0, ip, ir)];
                whi[kk] = m->cosines[nth - 1] / m->sines[nth - 1] * vv[IDX(m, nth - 1, ip, ir)];
            }
        /* the theta derivatives first (each is a whole-field call here); the second
         * derivatives of the Laplacian branches use the same wall values */
        {
            double *d_w = (double *)malloc(sizeof(double) * N), *d_v = (double *)malloc(sizeof(double) * N);
            double *d_u = (double *)malloc(sizeof(double) * N), *d_s = (double *)malloc(sizeof(double) * N);
            double *d_bp = NULL, *d_br = NULL;
            double *e_w = NULL, *e_v = NULL, *e_u = NULL, *e_s = NULL;
            dth(m, w, d_w, p->ibct[2], NULL, NULL, glo, ghi);
            dth(m, vv, d_v, p->ibct[3], wlo, whi, glo, ghi);
            dth(m, u, d_u, p->ibct[4], NULL, NULL, glo, ghi);
            dth(m, s, d_s, p->ibct[5], NULL, NULL, glo, ghi);
            if (m->lap) {
                e_w = (double *)malloc(sizeof(double) * N);
                e_v = (double *)malloc(sizeof(double) * N);
                e_u = (double *)malloc(sizeof(double) * N);
                d2th(m, w, e_w, p->ibct[2], NULL, NULL, glo, ghi);
                d2th(m, vv, e_v, p->ibct[3], wlo, whi, glo, ghi);
                d2th(m, u, e_u, p->ibct[4], NULL, NULL, glo, ghi);
            }
            if (m->slap) {
                e_s = (double *)malloc(sizeof(double) * N);
                d2th(m, s, e_s, p->ibct[5], NULL, NULL, glo, ghi);
            }
            if (p->magnetic) {
                /* Theta_Normal_Field wall values: bp -> -cot*bp, br -> 0 (rhs_mpi.f90:418-422) */
                for (ir = 0; ir < nr; ir++)
                    for (ip = 0; ip < nphi; ip++) {
                        kk = (long)ir * nphi + ip;
                        wlo[kk] = p->theta_normal_field ? -m->cosines[0] / m->sines[0] * bp[IDX(m, 0, ip, ir)] : 0.0;
                        whi[kk] = p->theta_normal_field ? -m->cosines[nth - 1] / m->sines[nth - 1] * bp[IDX(m, nth - 1, ip, ir)] : 0.0;
                    }
                d_bp = (double *)malloc(sizeof(double) * N);
                d_br = (double *)malloc(sizeof(double) * N);
                dth(m, bp, d_bp, p->ibct[8], wlo, whi, glo, ghi);
                dth(m, br, d_br, p->ibct[9], NULL, NULL, glo, ghi);
            }
            for (ir = 0; ir < nr; ir++) {
                const double ri = 1.0 / m->r[ir], rv = m->r[ir], epsr = m->eps ? m->eps[ir] : 0.0;
                const double omsq = rv * p->omega0 * p->omega0;
                const double dlm = m->lap ? m->dlnmu[ir] : 0.0;
                for (ip = 0; ip < nphi; ip++)
                    for (it = 0; it < nth; it++) {
                        const long q = IDX(m, it, ip, ir);
                        const double sn = m->sines[it], cs = m->cosines[it], cot = cs / sn;
                        double Rho = rho[q], W = w[q], V = vv[q], U = u[q], t = T[q];
                        double xdw, xdv, xdu, xds, wk, work6 = 0.0, ddiff;
                        double xdrho;
                        xdw = omsq * sn * cs;
                        xdu = -m->gravity[ir] + omsq * sn * sn;
                        xdrho = -Rho * W * cot * ri;
                        xdv = -om2 * cs * W;
                        xdu = xdu + ri * W * W;
                        wk = d_w[q] * ri;
                        xdrho = xdrho - Rho * wk;
                        xdw = xdw - W * wk;
                        if (m->lap) {   /* :844-864 */
                            work6 = m->mur[ir] * ir_[q];
                            ddiff = -(4.0 / 3.0) * ri * ri;
                            xdw = xdw + (ddiff * (1.0 / (sn * sn)) - ri * dlm) * work6 * W;
                            v2[0][q] = wk;
                            wk = wk * work6;
                            ddiff = -(1.0 / 3.0) * (7.0 * ri + 2.0 * dlm);
                            xdu = xdu + ddiff * (ri * work6 * cot * W + wk);
                            ddiff = (4.0 / 3.0) * ri;
                            xdw = xdw + ddiff * cot * wk;
                            xdw = xdw + ddiff * ri * work6 * e_w[q];
                        }
                        xdw = xdw + V * (ri * cot * V + om2 * cs);
                        xdv = xdv - ri * V * W * cot;
                        xdu = xdu + V * (ri * V + om2 * sn);
                        wk = d_v[q] * ri;
                        xdv = xdv - W * wk;
                        if (m->lap) {   /* :883-891 */
                            xdv = xdv + ri * work6 * (cot * wk - (ri * (1.0 / (sn * sn)) + dlm) * V);
                            v2[1][q] = wk;
                            xdv = xdv + ri * ri * work6 * e_v[q];
                        }
                        drho[q] = xdrho - 2.0 * ri * Rho * U;
                        xdv = xdv - U * (ri * V + om2 * sn);
                        xdw = xdw - ri * U * W;
                        if (!p->magnetic) dvv[q] = xdv;
                        wk = d_u[q] * ri;
                        xdu = xdu - W * wk;
                        if (m->lap) {   /* :914-926 */
                            v2[2][q] = wk;
                            wk = wk * work6;
                            xdw = xdw + (2.0 * (4.0 / 3.0) * ri + dlm) * wk;
                            xdu = xdu + ri * cot * wk;
                            xdu = xdu + ri * ri * work6 * e_u[q];
                        }
                        if (!p->magnetic) du[q] = xdu;
                        xds = (fabs(epsr) > 0.0) ? epsr * ir_[q] / t : 0.0;
                        wk = d_s[q] * ri;
                        if (p->magnetic) xdw = xdw - m->gamm1 * t * wk;
                        else dw[q] = xdw - m->gamm1 * t * wk;
                        xds = xds - W * wk;
                        if (m->slap) {   /* :959-975: the second-derivative term carries ri^3 */
                            const double oorho = ir_[q] / t;
                            ddiff = ri * m->ks[ir];
                            wk = ddiff * wk * oorho;
                            xds = xds + cot * wk;
                            ddiff = ddiff * ri * ri;
                            xds = xds + ddiff * e_s[q] * oorho;
                        }
                        ds[q] = xds;
                        if (p->magnetic) {
                            double oorho = ir_[q] * ONE_OVER_4PI * ri;
                            double Bt = bt[q], Bp = bp[q], Br = br[q];
                            xdu = xdu - Bt * Bt * oorho;
                            xdw = xdw - Bp * Bp * cot * oorho;
                            xdv = xdv + Bp * Bt * cot * oorho;
                            xdu = xdu - Bp * Bp * oorho;
                            wk = d_bp[q] * oorho;
                            xdv = xdv + Bt * wk;
                            xdw = xdw - Bp * wk;
                            wk = d_br[q] * oorho;
                            dw[q] = xdw + Br * (Bt * oorho - wk);
                            dvv[q] = xdv + Br * Bp * oorho;
                            du[q] = xdu + Bt * wk;
                        }
                    }
            }
            free(d_w); free(d_v); free(d_u); free(d_s); free(d_bp); free(d_br);
            free(e_w); free(e_v); free(e_u); free(e_s);
        }
    }

    /* ---- Phi_Terms (rhs_mpi.f90:1032-1261) ---- */
    {
        double *d_rho = (double *)malloc(sizeof(double) * N), *d_v = (double *)malloc(sizeof(double) * N);
        double *d_w = (double *)malloc(sizeof(double) * N), *d_u = (double *)malloc(sizeof(double) * N);
        double *d_s = (double *)malloc(sizeof(double) * N), *d_bt = NULL, *d_br = NULL;
        double *e_v = NULL, *e_w = NULL, *e_u = NULL, *e_s = NULL;
        dph(m, rho, d_rho); dph(m, vv, d_v); dph(m, w, d_w); dph(m, u, d_u); dph(m, s, d_s);
        if (m->lap) {
            e_v = (double *)malloc(sizeof(double) * N);
            e_w = (double *)malloc(sizeof(double) * N);
            e_u = (double *)malloc(sizeof(double) * N);
            d2ph(m, vv, e_v); d2ph(m, w, e_w); d2ph(m, u, e_u);
        }
        if (m->slap) {
            e_s = (double *)malloc(sizeof(double) * N);
            d2ph(m, s, e_s);
        }
        if (p->magnetic) {
            d_bt = (double *)malloc(sizeof(double) * N);
            d_br = (double *)malloc(sizeof(double) * N);
            dph(m, bt, d_bt); dph(m, br, d_br);
        }
        for (ir = 0; ir < nr; ir++) {
            const double ri = 1.0 / m->r[ir];
            const double dlm = m->lap ? m->dlnmu[ir] : 0.0;
            for (ip = 0; ip < nphi; ip++)
                for (it = 0; it < nth; it++) {
                    const long q = IDX(m, it, ip, ir);
                    const double oors = ri * (1.0 / m->sines[it]);
                    const double cot = m->cosines[it] / m->sines[it];
                    double Rho = rho[q], V = vv[q], W = w[q], U = u[q], t = T[q], oorho = ir_[q];
                    double xdrho, xdv, xdw, xdu, xds, wk, work2 = 0.0, ddiff, ddiff2;
                    wk = d_rho[q] * oors;
                    xdrho = drho[q];
                    xdv = dvv[q];
                    xdrho = xdrho - V * wk;
                    xdv = xdv - m->pconst * wk * t * oorho;
                    xds = ds[q];
                    wk = d_v[q] * oors;
                    xdv = xdv - V * wk;
                    drho[q] = xdrho - Rho * wk;
                    xdw = dw[q];
                    xdu = du[q];
                    if (m->lap) {   /* :1123-1141 */
                        work2 = oorho * m->mur[ir];
                        v2[3][q] = wk;
                        /* viscous heating theta-theta, phi-phi */
                        xds = xds + (4.0 / 3.0) * work2 * (v2[0][q] + U * ri) *
                                        (v2[0][q] + 2.0 * U * ri - wk - ri * cot * W) / t;
                        wk = wk * work2;
                        ddiff = -(7.0 / 3.0) * ri;
                        ddiff2 = ddiff - (2.0 / 3.0) * dlm;
                        xdw = xdw + ddiff * cot * wk;
                        xdu = xdu + ddiff2 * wk;
                        xdv = xdv + (4.0 / 3.0) * (e_v[q] * oors * oors) * work2;
                    }
                    wk = d_w[q] * oors;
                    xdw = xdw - V * wk;
                    if (m->lap) {   /* :1156-1169: viscous heating theta-phi */
                        xds = xds + work2 * (v2[1][q] + wk - ri * cot * V) * (v2[1][q] + wk - ri * cot * V) / t;
                        v2[1][q] = wk;
                        ddiff = (7.0 / 3.0) * ri;
                        wk = wk * work2;
                        xdv = xdv + ddiff * cot * wk;
                        xdw = xdw + (e_w[q] * oors * oors) * work2;
                    }
                    if (!p->magnetic) dw[q] = xdw;
                    wk = d_u[q] * oors;
                    xdu = xdu - V * wk;
                    if (m->lap) {   /* :1189-1198 */
                        v2[4][q] = wk;
                        ddiff = 2.0 * (4.0 / 3.0) * ri + dlm;
                        xdv = xdv + ddiff * wk * work2;
                        xdu = xdu + (e_u[q] * oors * oors) * work2;
                    }
                    if (!p->magnetic) du[q] = xdu;
                    wk = d_s[q] * oors;
                    xds = xds - V * wk;
                    xdv = xdv - m->gamm1 * t * wk;
                    if (!p->magnetic) dvv[q] = xdv;
                    if (m->slap)   /* :1221-1228 */
                        xds = xds + m->ks[ir] * (e_s[q] * (oors * oors)) * (oorho / t);
                    ds[q] = xds;
                    if (p->magnetic) {
                        double Bt = bt[q], Bp = bp[q], Br = br[q];
                        oorho = ONE_OVER_4PI * oorho * oors;
                        wk = d_bt[q] * oorho;
                        xdv = xdv - Bt * wk;
                        dw[q] = xdw + Bp * wk;
                        wk = d_br[q] * oorho;
                        dvv[q] = xdv - Br * wk;
                        du[q] = xdu + Bp * wk;
                    }
                }
        }
        free(d_rho); free(d_v); free(d_w); free(d_u); free(d_s); free(d_bt); free(d_br);
        free(e_v); free(e_w); free(e_u); free(e_s);
    }

    /* ---- radial wall values (rhs_mpi.f90:578-637) and Radial_Terms (:1264-1519) ---- */
    {
        double *d_w = (double *)malloc(sizeof(double) * N), *d_v = (double *)malloc(sizeof(double) * N);
        double *d_u = (double *)malloc(sizeof(double) * N), *d_s = (double *)malloc(sizeof(double) * N);
        double *d_bt = NULL, *d_bp = NULL;
        double *e_w = NULL, *e_v = NULL, *e_u = NULL, *e_s = NULL;
        const double r1 = m->r[0], r2 = m->r[nr - 1], x1 = m->dxdr[0], x2 = m->dxdr[nr - 1];
        long kk;
        if (m->lap) {
            e_w = (double *)malloc(sizeof(double) * N);
            e_v = (double *)malloc(sizeof(double) * N);
            e_u = (double *)malloc(sizeof(double) * N);
        }
        if (m->slap) e_s = (double *)malloc(sizeof(double) * N);
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                kk = (long)ip * nth + it;
                wlo[kk] = p->bottom_stress_free ? w[IDX(m, it, ip, 0)] / r1 / x1 : 0.0;
                whi[kk] = p->top_stress_free ? w[IDX(m, it, ip, nr - 1)] / r2 / x2 : w[IDX(m, it, ip, nr - 1)];
            }
        dr_(m, w, d_w, p->ibcr[2], wlo, whi, glo, ghi);
        if (m->lap) d2r_(m, w, e_w, p->ibcr[2], wlo, whi, glo, ghi);
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                kk = (long)ip * nth + it;
                wlo[kk] = p->bottom_stress_free ? vv[IDX(m, it, ip, 0)] / r1 / x1 : 0.0;
                whi[kk] = p->top_stress_free ? vv[IDX(m, it, ip, nr - 1)] / r2 / x2 : vv[IDX(m, it, ip, nr - 1)];
            }
        dr_(m, vv, d_v, p->ibcr[3], wlo, whi, glo, ghi);
        if (m->lap) d2r_(m, vv, e_v, p->ibcr[3], wlo, whi, glo, ghi);
        dr_(m, u, d_u, p->ibcr[4], NULL, NULL, glo, ghi);
        if (m->lap) d2r_(m, u, e_u, p->ibcr[4], NULL, NULL, glo, ghi);
        /* constant-flux entropy walls: dsdr/dxdr, or -lumr/(4 pi r^2 ks dxdr) under the explicit
         * thermal diffusion (:584-590, :616-622) */
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                kk = (long)ip * nth + it;
                wlo[kk] = !p->bottom_const_flux ? 0.0
                          : m->slap ? -p->lumr1 * ONE_OVER_4PI / (r1 * r1 * m->ks[0] * x1) : m->dsdr1[it] / x1;
                whi[kk] = !p->top_const_flux ? 0.0
                          : m->slap ? -p->lumr2 * ONE_OVER_4PI / (r2 * r2 * m->ks[nr - 1] * x2) : m->dsdr2[it] / x2;
            }
        dr_(m, s, d_s, p->ibcr[5], wlo, whi, glo, ghi);
        if (m->slap) d2r_(m, s, e_s, p->ibcr[5], wlo, whi, glo, ghi);
        if (p->magnetic) {
            d_bt = (double *)malloc(sizeof(double) * N);
            d_bp = (double *)malloc(sizeof(double) * N);
            dr_(m, bt, d_bt, p->ibcr[7], NULL, NULL, glo, ghi);
            dr_(m, bp, d_bp, p->ibcr[8], NULL, NULL, glo, ghi);
        }
        for (ir = 0; ir < nr; ir++) {
            const double x = m->dxdr[ir], ri = 1.0 / m->r[ir];
            /* Radial_Terms' work9 = d2xdr2/dxdr, work5 = 2/r + dlnmu, work6 = (2/r + dlnks) ks */
            const double w9 = m->slap ? m->d2xdr2[ir] / x : 0.0;
            const double w5 = m->lap ? 2.0 * ri + m->dlnmu[ir] : 0.0;
            const double w6 = m->slap ? (2.0 * ri + m->dlnks[ir]) * m->ks[ir] : 0.0;
            for (ip = 0; ip < nphi; ip++)
                for (it = 0; it < nth; it++) {
                    const long q = IDX(m, it, ip, ir);
                    const double cot = m->cosines[it] / m->sines[it];
                    double U = u[q], W = w[q], V = vv[q], t = T[q], xdw, xdv, xdu, xds, wk, work2 = 0.0, pw;
                    wk = d_w[q];
                    xdw = dw[q] - U * wk;
                    xdu = du[q];
                    xds = ds[q];
                    if (m->lap) {   /* :1317-1343 */
                        work2 = m->mur[ir] * ir_[q];
                        xdu = xdu + (1.0 / 3.0) * cot * ri * wk;   /* no mur/rho factor here (:1322) */
                        xdw = xdw + work2 * w5 * wk;
                        /* viscous heating r-theta */
                        xds = xds + work2 * (wk + v2[2][q] - W * ri) * (wk + v2[2][q] - W * ri) / t;
                        xdw = xdw + w9 * wk * work2;
                        xdw = xdw + (e_w[q] * x * x) * work2;
                    }
                    if (!p->magnetic) dw[q] = xdw;
                    wk = d_v[q];
                    xdv = dvv[q] - U * wk;
                    if (m->lap) {   /* :1364-1380 */
                        /* viscous heating r-phi (listed as r-theta in the reference) */
                        xds = xds + work2 * (wk + v2[4][q] - V * ri) * (wk + v2[4][q] - V * ri) / t;
                        xdv = xdv + wk * w5 * work2;
                        xdv = xdv + w9 * wk * work2;
                        xdv = xdv + (e_v[q] * x * x) * work2;
                    }
                    if (!p->magnetic) dvv[q] = xdv;
                    wk = d_u[q];
                    xdu = xdu - U * wk;
                    drho[q] = drho[q] - rho[q] * wk;
                    if (m->lap) {   /* :1409-1431; the r-phi heating term appears again with du/dr in
                                       place of dv/dr (:1418) */
                        xdu = xdu + (4.0 / 3.0) * work2 * w5 * (wk - U * ri);
                        xdu = xdu + (4.0 / 3.0) * w9 * wk * work2;
                        xds = xds + (4.0 / 3.0) * work2 * wk * (wk - v2[0][q] - U * ri) / t;
                        xds = xds + work2 * (wk + v2[4][q] - V * ri) * (wk + v2[4][q] - V * ri) / t;
                        pw = v2[3][q] + (cot * W + U) * ri;
                        xds = xds + (4.0 / 3.0) * work2 * pw * (pw - wk) / t;
                        xdu = xdu + (4.0 / 3.0) * (e_u[q] * x * x) * work2;
                    }
                    wk = d_s[q];
                    xdu = xdu - m->gamm1 * t * wk;
                    xds = xds - U * wk;
                    if (m->slap) {   /* :1456-1472 */
                        const double oorho = ir_[q] / t;
                        wk = wk * oorho;
                        xds = xds + wk * w6;
                        xds = xds + w9 * wk * m->ks[ir];
                        xds = xds + m->ks[ir] * (e_s[q] * x * x) * oorho;
                    }
                    if (!p->magnetic) du[q] = xdu;
                    ds[q] = xds;
                    if (p->magnetic) {
                        double oorho = ir_[q] * ONE_OVER_4PI;
                        wk = d_bt[q] * oorho;
                        xdu = xdu - bt[q] * wk;
                        dw[q] = xdw + br[q] * wk;
                        wk = d_bp[q] * oorho;
                        dvv[q] = xdv + br[q] * wk;
                        du[q] = xdu - bp[q] * wk;
                    }
                }
        }
        free(d_w); free(d_v); free(d_u); free(d_s); free(d_bt); free(d_bp);
        free(e_w); free(e_v); free(e_u); free(e_s);
    }

    /* ---- Laplacian_Cross_Terms_Radial / _Theta (rhs_mpi.f90:2070-2296) ---- */
    if (m->lap) {
        double *X = (double *)malloc(sizeof(double) * N), *dX = (double *)malloc(sizeof(double) * N);
        double *tmp;
        long kk;
        const double r1 = m->r[0], r2 = m->r[nr - 1];
        /* vars2 slots 2 and 5 trade places (:2103-2105): 1 dw/dth/r, 2 du/dph/(r sin),
         * 3 du/dth/r, 4 dv/dph/(r sin), 5 dw/dph/(r sin) */
        tmp = v2[1]; v2[1] = v2[4]; v2[4] = tmp;
        /* radial: d/dr of r*vars2(k) with wall values r_wall*dbinfo; only dbinfo(1,1)(1) and
         * dbinfo(1,2)(4) are set, both under Bottom_Stress_Free (:2113-2125), the rest stay 0.
         * The mur/r/rho factor takes invrho at the point itself (the reference indexes
         * invrho(it+ir-1,...) with it unset, :2133). */
        {
            static const int slot[4] = {0, 1, 2, 3};
            const int code[4] = {p->ibcr[2], p->ibcr[4], p->ibcr[4], p->ibcr[3]};
            int k;
            for (k = 0; k < 4; k++) {
                for (ir = 0; ir < nr; ir++)
                    for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) X[i] = m->r[ir] * v2[slot[k]][i];
                for (ip = 0; ip < nphi; ip++)
                    for (it = 0; it < nth; it++) {
                        kk = (long)ip * nth + it;
                        wlo[kk] = (k == 0 && p->bottom_stress_free)
                                      ? r1 * (v2[0][IDX(m, it, ip, 0)] / r1 / m->dxdr[0]) : 0.0;
                        whi[kk] = (k == 3 && p->bottom_stress_free)
                                      ? r2 * (v2[3][IDX(m, it, ip, nr - 1)] / r2 / m->dxdr[nr - 1]) : 0.0;
                    }
                dr_(m, X, dX, code[k], wlo, whi, glo, ghi);
                if (g_cross_dump) memcpy(g_cross_dump + k * N, dX, sizeof(double) * N);
                for (ir = 0; ir < nr; ir++) {
                    const double c6 = (1.0 / 3.0) * m->mur[ir] / m->r[ir];
                    for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) {
                        const double add = c6 * ir_[i] * dX[i];
                        if (k == 0 || k == 3) du[i] += add;
                        else if (k == 1) dvv[i] += add;
                        else dw[i] += add;
                    }
                }
            }
        }
        /* theta: d/dth of sin*vars2(4) and sin*vars2(5), wall values sin_wall*dbinfo with
         * dbinfo(2,*)(4) = vars2(4) cot at the walls (:2243-2250) and (5) = 0 */
        {
            int k;
            for (k = 0; k < 2; k++) {
                const double *src = v2[3 + k];
                for (ir = 0; ir < nr; ir++)
                    for (ip = 0; ip < nphi; ip++)
                        for (it = 0; it < nth; it++) {
                            const long q = IDX(m, it, ip, ir);
                            X[q] = m->sines[it] * src[q];
                        }
                for (ir = 0; ir < nr; ir++)
                    for (ip = 0; ip < nphi; ip++) {
                        kk = (long)ir * nphi + ip;
                        wlo[kk] = k == 0 ? m->sines[0] * (src[IDX(m, 0, ip, ir)] * m->cosines[0] / m->sines[0]) : 0.0;
                        whi[kk] = k == 0 ? m->sines[nth - 1] * (src[IDX(m, nth - 1, ip, ir)] * m->cosines[nth - 1] /
                                                                 m->sines[nth - 1])
                                         : 0.0;
                    }
                dth(m, X, dX, k == 0 ? p->ibct[3] : p->ibct[2], wlo, whi, glo, ghi);
                if (g_cross_dump) memcpy(g_cross_dump + (4 + k) * N, dX, sizeof(double) * N);
                for (ir = 0; ir < nr; ir++) {
                    const double rm = m->mur[ir] / m->r[ir];
                    for (ip = 0; ip < nphi; ip++)
                        for (it = 0; it < nth; it++) {
                            const long q = IDX(m, it, ip, ir);
                            const double add = (1.0 / 3.0) * rm * ir_[q] / m->sines[it] * dX[q];
                            if (k == 0) dw[q] += add;
                            else dvv[q] += add;
                        }
                }
            }
        }
        free(X); free(dX);
    }

    /* ---- Density_Terms_Theta (rhs_mpi.f90:1522-1605) ---- */
    {
        long kk;
        for (ir = 0; ir < nr; ir++) {
            const double rv = m->r[ir] * m->invpconst;
            for (ip = 0; ip < nphi; ip++) {
                long q0 = IDX(m, 0, ip, ir), q1 = IDX(m, nth - 1, ip, ir);
                kk = (long)ir * nphi + ip;
                wlo[kk] = rv * dw[q0] * rho[q0] / T[q0];
                whi[kk] = rv * dw[q1] * rho[q1] / T[q1];
            }
        }
        dth(m, rho, work1, p->ibct[1], wlo, whi, glo, ghi);
        for (ir = 0; ir < nr; ir++) {
            const double ri = 1.0 / m->r[ir];
            for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) {
                double wk = ri * work1[i];
                drho[i] = drho[i] - w[i] * wk;
                wk = wk * ir_[i];
                dw[i] = dw[i] - m->pconst * wk * T[i];
            }
        }
    }

    /* ---- Density_Terms_Radial (rhs_mpi.f90:1710-1855) ---- */
    {
        long kk;
        const int lb = p->ibcr[1];
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                long q0 = IDX(m, it, ip, 0), q1 = IDX(m, it, ip, nr - 1);
                kk = (long)ip * nth + it;
                wlo[kk] = (lb == 2 || lb == 4) ? du[q0] * rho[q0] * m->invpconst / T[q0] : rho[q0];
                wlo[kk] = wlo[kk] / m->dxdr[0];
                whi[kk] = (lb == 3 || lb == 4) ? du[q1] * rho[q1] * m->invpconst / T[q1] : rho[q1];
                whi[kk] = whi[kk] / m->dxdr[nr - 1];
            }
        dr_(m, rho, work1, lb, wlo, whi, glo, ghi);
        for (i = 0; i < N; i++) {
            double wk = work1[i];
            drho[i] = drho[i] - u[i] * wk;
            wk = wk * ir_[i];
            du[i] = du[i] - m->pconst * T[i] * wk;
        }
        if (p->top_relax) {
            const double dt_ = p->damping_time * (p->gamma - 2.0);
            for (ip = 0; ip < nphi; ip++)
                for (it = 0; it < nth; it++) {
                    long q = IDX(m, it, ip, nr - 1);
                    double t = T[q];
                    drho[q] = drho[q] + rho[q] * (p->T0 - t) / (dt_ * t);
                    ds[q] = ds[q] - m->Cv * (p->T0 - t) / (dt_ * t);
                }
        }
    }

    /* ---- Induction_Terms: u x B (rhs_mpi.f90:2342-2346); with Laplacian the resistive term
     * -etar curl B and the Ohmic heating 4 pi etar |curl B|^2 / c^2 / (rho T) (:2303-2322) ---- */
    if (p->magnetic) {
        double *dat = dv + 5 * N, *dap = dv + 6 * N, *dar = dv + 7 * N;
        if (m->lap) {
            const double jconst = 4.0 * PI_ / (ORC_C_LIGHT * ORC_C_LIGHT);
            vector_curl(m, bt, bp, br, dat, dap, dar, work, work1, work2, wlo, whi, glo, ghi);
            for (ir = 0; ir < nr; ir++) {
                const double eta = m->etar[ir];
                for (i = IDX(m, 0, 0, ir); i < IDX(m, 0, 0, ir + 1); i++) {
                    v2[0][i] = -eta * dat[i];
                    v2[1][i] = -eta * dap[i];
                    v2[2][i] = -eta * dar[i];
                    ds[i] = ds[i] + (jconst * eta) * (dat[i] * dat[i] + dap[i] * dap[i] + dar[i] * dar[i]) * ir_[i] / T[i];
                }
            }
        }
        for (i = 0; i < N; i++) {
            dat[i] = vv[i] * br[i] - u[i] * bp[i];
            dap[i] = u[i] * bt[i] - w[i] * br[i];
            dar[i] = w[i] * bp[i] - vv[i] * bt[i];
        }
        if (m->lap)
            for (i = 0; i < N; i++) {
                dat[i] += v2[0][i];
                dap[i] += v2[1][i];
                dar[i] += v2[2][i];
            }
    }
    if (p->xi_cs < 1.0)
        for (i = 0; i < N; i++) drho[i] = (p->xi_cs * p->xi_cs) * drho[i];

    free(B); free(T); free(ir_); free(work); free(work1); free(work2);
    free(wlo); free(whi); free(glo); free(ghi); free(m->sines);
    for (i = 0; i < 5; i++) free(v2[i]);
    return 0;
}

/* Bound (rhs_mpi.f90:2395-2491) on a state in place (A in slots 6-8). */
int orc_bound(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, double *v)
{
    const int nr = g->nr, nth = g->nth, nphi = g->nphi;
    const long N = (long)nr * nth * nphi;
    double *w = v + N, *vv = v + 2 * N, *u = v + 3 * N, *s = v + 4 * N;
    int it, ip, ir;
#define Q(it, ip, ir) ((((long)(ir) * nphi + (ip)) * nth) + (it))
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++) {
            w[Q(0, ip, ir)] = 0.0;
            w[Q(nth - 1, ip, ir)] = 0.0;
            if (!p->theta_stress_free) { u[Q(0, ip, ir)] = 0.0; u[Q(nth - 1, ip, ir)] = 0.0; }
        }
    for (ip = 0; ip < nphi; ip++)
        for (it = 0; it < nth; it++) {
            if (!p->bottom_stress_free) { w[Q(it, ip, 0)] = 0.0; vv[Q(it, ip, 0)] = 0.0; }
            u[Q(it, ip, 0)] = 0.0;
            if (p->bottom_const_temp) s[Q(it, ip, 0)] = pr->sr1[it];
            if (!p->top_stress_free) { w[Q(it, ip, nr - 1)] = 0.0; vv[Q(it, ip, nr - 1)] = 0.0; }
            u[Q(it, ip, nr - 1)] = 0.0;
            /* Top_Constant_Temp fixes s only with Set_Stop, which this build does not model */
        }
    if (p->magnetic) {
        double *at = v + 5 * N, *ap = v + 6 * N, *ar = v + 7 * N;
        for (ir = 0; ir < nr; ir++)
            for (ip = 0; ip < nphi; ip++) {
                if (p->theta_perf_cond) {
                    ap[Q(0, ip, ir)] = 0.0; ar[Q(0, ip, ir)] = 0.0;
                    ap[Q(nth - 1, ip, ir)] = 0.0; ar[Q(nth - 1, ip, ir)] = 0.0;
                } else {
                    at[Q(0, ip, ir)] = 0.0;
                    at[Q(nth - 1, ip, ir)] = 0.0;
                }
            }
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                if (p->bottom_radial_field) ar[Q(it, ip, 0)] = 0.0;
                if (p->top_radial_field) ar[Q(it, ip, nr - 1)] = 0.0;
            }
    }
#undef Q
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Compute_Diffusion (diffusion.f90:70-474) with Diffusion_R (:476-891), Diffusion_Th
 * (:893-1301), Diffusion_Phi (:1303-1495): slope-limited (Van Albada) second-order fluxes,
 * speed sld_coef*sqrt(vcc), zero flux through the r and theta walls, then viscous heating,
 * mass diffusion and Ohmic heating corrections.  v = vars1 as the RHS(1) call leaves it
 * (B in slots 6-8), a = the vector potential A (maga), dv = the diffusion tendency.
 * The background meanV is the hydrostatic profile (update_background = .False.) and the
 * optional flux limiters are off (their defaults, diffusion.f90:8).
 * ------------------------------------------------------------------------------------------- */
static double van_albada(double q, double qph, double tiny)
{
    double rt = fabs(q);
    rt = 1.0 / fmax(rt, tiny);
    rt = q * rt * rt;
    rt = rt * qph;
    rt = fmin(1.0, rt);
    return q * fmax(0.0, rt);
}

/* face fluxes of one pencil: tmp[0..n-1], ghosts gm (the value before tmp[0], "b1(:,4)") and
 * gp (after tmp[n-1], "b2(:,1)"), gm2/gp2 the next ones; delue[0..n] = qph - (phiL+phiR)/2 */
static void sld_faces(int n, const double *tmp, double gm2, double gm, double gp, double gp2, double tiny,
                      double *delue)
{
    int k;
    for (k = 0; k <= n; k++) {
        /* face k sits between tmp[k-1] and tmp[k] (Fortran face k+1) */
        double fm2 = (k - 2 >= 0) ? tmp[k - 2] : (k - 2 == -1 ? gm : gm2);
        double fm1 = (k - 1 >= 0) ? tmp[k - 1] : gm;
        double f0 = (k < n) ? tmp[k] : gp;
        double fp1 = (k + 1 < n) ? tmp[k + 1] : (k + 1 == n ? gp : gp2);
        double qmh = fm1 - fm2, qph = f0 - fm1, qp3h = fp1 - f0;
        double phiL = van_albada(qmh, qph, tiny);
        double phiR = van_albada(qp3h, qph, tiny);
        delue[k] = qph - 0.5 * (phiR + phiL);
    }
}

int orc_diffusion(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, const double *v,
                  const double *a, double *dv)
{
    orc_mhd M, *m = &M;
    const int nr = g->nr, nth = g->nth, nphi = g->nphi, nv = p->magnetic ? 8 : 5;
    const long N = (long)nr * nth * nphi;
    const int vth = nth + 2, vph = nphi + 2;
    double *vcc, *f, *invrho, *itemp, *tmp, *delue, *work, *work2, *glo, *ghi, *curl;
    double tinyvar[9], dth, dphi;
    int varlist[8], nsld = 0, isl, iv, it, ip, ir, k;
    const int slap = p->laplacian || p->slaplacian;
    long i;
    setup(m, g, p, pr);
    memset(dv, 0, sizeof(double) * nv * N);
    if (!p->sld_on) { free(m->sines); return 0; }
    /* tinyvar (diffusion.f90:76-80), 1-based by variable */
    tinyvar[1] = 1e-8 * pr->rho_bar[nr - 1];
    tinyvar[2] = tinyvar[3] = tinyvar[4] = 1e-8;
    tinyvar[5] = 1e-3;
    tinyvar[6] = tinyvar[7] = tinyvar[8] = 1.0;
    /* varlist (:86-112) */
    for (iv = (p->mass_diffusion ? 1 : 2); iv <= ((p->magnetic && p->magnetic_diffusion) ? 8 : 5); iv++)
        varlist[nsld++] = iv;
    dth = (pr->theta[nth - 1] - pr->theta[0]) / (nth - 1);
    dphi = 1.0 / g->dpi;
    vcc = (double *)calloc((size_t)vth * vph * (nr + 2), sizeof(double));
    f = (double *)malloc(sizeof(double) * 8 * N);
    invrho = (double *)malloc(sizeof(double) * N);
    itemp = (double *)malloc(sizeof(double) * N);
    tmp = (double *)malloc(sizeof(double) * (nr + nth + nphi + 8));
    delue = (double *)malloc(sizeof(double) * (nr + nth + nphi + 8));
    work = (double *)malloc(sizeof(double) * N);
    work2 = (double *)malloc(sizeof(double) * N);
    curl = (double *)malloc(sizeof(double) * 3 * N);
    glo = (double *)calloc(4 * (size_t)N / (nr < nth ? nr : nth) + 64, sizeof(double));
    ghi = (double *)calloc(4 * (size_t)N / (nr < nth ? nr : nth) + 64, sizeof(double));
    if (!vcc || !f || !invrho || !itemp || !tmp || !delue || !work || !work2 || !curl || !glo || !ghi) return 12;
#define VC(it, ip, ir) vcc[((long)(ir) * vph + (ip)) * vth + (it)]
    /* 1/rho and 1/T (:114-121), fields with the background removed (:168-175) */
    for (ir = 0; ir < nr; ir++)
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                i = IDX(m, it, ip, ir);
                invrho[i] = 1.0 / v[i];
                itemp[i] = pow(v[i], -m->gamm1) * exp(-v[4 * N + i] * m->invCv) / m->invtconst;
                f[i] = v[i] - pr->rho_bar[ir];
                f[N + i] = v[N + i]; f[2 * N + i] = v[2 * N + i]; f[3 * N + i] = v[3 * N + i];
                f[4 * N + i] = v[4 * N + i] - pr->s_bar[ir];
                if (p->magnetic) { f[5 * N + i] = a[i]; f[6 * N + i] = a[N + i]; f[7 * N + i] = a[2 * N + i]; }
            }
    /* vcc: speed^2 with theta/phi/r ghosts (:177-287); phi ghosts are the periodic wrap */
    for (ir = 0; ir <= nr + 1; ir++)
        for (ip = 0; ip < vph; ip++)
            for (it = 0; it < vth; it++) VC(it, ip, ir) = pr->diffusion_speed[ir];
    if (p->dstype == 1) {
        for (ir = 0; ir < nr; ir++)
            for (ip = -1; ip <= nphi; ip++) {
                int ipw = (ip + nphi) % nphi;
                for (it = 0; it < nth; it++) {
                    i = IDX(m, it, ipw, ir);
                    for (iv = 2; iv <= 4; iv++) VC(it + 1, ip + 1, ir + 1) += v[(iv - 1) * N + i] * v[(iv - 1) * N + i];
                }
            }
        if (p->magnetic)
            for (ir = 0; ir < nr; ir++)
                for (ip = -1; ip <= nphi; ip++) {
                    int ipw = (ip + nphi) % nphi;
                    for (it = 0; it < nth; it++) {
                        i = IDX(m, it, ipw, ir);
                        for (iv = 6; iv <= 8; iv++)
                            VC(it + 1, ip + 1, ir + 1) += ONE_OVER_4PI * invrho[i] * v[(iv - 1) * N + i] * v[(iv - 1) * N + i];
                    }
                }
    }
    for (ip = 0; ip < vph; ip++)
        for (it = 0; it < vth; it++) { VC(it, ip, 0) = VC(it, ip, 1); VC(it, ip, nr + 1) = VC(it, ip, nr); }
    for (ir = 0; ir <= nr + 1; ir++)
        for (ip = 0; ip < vph; ip++) { VC(0, ip, ir) = VC(1, ip, ir); VC(nth + 1, ip, ir) = VC(nth, ip, ir); }

    for (isl = 0; isl < nsld; isl++) {
        const double tiny = tinyvar[varlist[isl]];
        const double *fv;
        double *dvv;
        iv = varlist[isl];
        fv = f + (long)(iv - 1) * N;
        dvv = dv + (long)(iv - 1) * N;
        /* ---- Diffusion_R ---- */
        for (ip = 0; ip < nphi; ip++)
            for (it = 0; it < nth; it++) {
                double gm, gp;
                for (ir = 0; ir < nr; ir++) tmp[ir] = fv[IDX(m, it, ip, ir)];
                switch (iv) {           /* b1(:,4) and b2(:,1) at the walls (:586-711) */
                case 1: gm = 0.0; gp = 0.0; break;
                case 5:
                    /* constant-flux walls under explicit thermal diffusion (:614-618, :688-692;
                     * dr_ind(1) at both walls) */
                    gm = (slap && p->bottom_const_flux)
                             ? tmp[0] + pr->dr[0] * p->lumr1 * ONE_OVER_4PI /
                                            (pr->r[0] * pr->r[0] * pr->ks[0] * pr->dxdr[0])
                             : 0.0;
                    gp = (slap && p->top_const_flux)
                             ? tmp[nr - 1] - pr->dr[0] * p->lumr2 * ONE_OVER_4PI /
                                                 (pr->r[nr - 1] * pr->r[nr - 1] * pr->ks[nr - 1] * pr->dxdr[nr - 1])
                             : 0.0;
                    break;
                case 2: case 3:
                    gm = -10.0 / 3.0 * tmp[0] / pr->r[0] + 6.0 * tmp[1] / pr->r[1] - 2.0 * tmp[2] / pr->r[2] +
                         1.0 / 3.0 * tmp[3] / pr->r[3];
                    gp = -10.0 / 3.0 * tmp[nr - 1] / pr->r[nr - 1] + 6.0 * tmp[nr - 2] / pr->r[nr - 2] -
                         2.0 * tmp[nr - 3] / pr->r[nr - 3] + 1.0 / 3.0 * tmp[nr - 4] / pr->r[nr - 4];
                    break;
                case 6: case 7:
                    gm = p->bottom_perf_cond ? 0.0 : tmp[0];
                    gp = p->top_perf_cond ? 0.0 : tmp[nr - 1];
                    break;
                default: gm = tmp[0]; gp = tmp[nr - 1]; break;   /* 4 and 8 */
                }
                sld_faces(nr, tmp, gm, gm, gp, gp, tiny, delue);
                delue[0] = 0.0;
                delue[nr] = 0.0;
                for (k = 1; k < nr; k++) {
                    double rph = 0.5 * (pr->r[k] + pr->r[k - 1]);
                    delue[k] = delue[k] * rph * rph;
                }
                for (k = 0; k <= nr; k++) {
                    double vt = p->sld_coef * sqrt(0.5 * (VC(it + 1, ip + 1, k + 1) + VC(it + 1, ip + 1, k)));
                    delue[k] = vt * delue[k];
                }
                for (ir = 0; ir < nr; ir++) {
                    double rl = (ir == 0) ? pr->r[0] - 0.5 * pr->dr[0] : 0.5 * (pr->r[ir] + pr->r[ir - 1]);
                    double rh = (ir == nr - 1) ? pr->r[nr - 1] + 0.5 * pr->dr[nr - 1] : 0.5 * (pr->r[ir + 1] + pr->r[ir]);
                    double fl = 3.0 / (rh * rh * rh - rl * rl * rl) * (delue[ir + 1] - delue[ir]);
                    if (p->sld_pr != 1.0 || p->sld_prm != 1.0) {
                        if (iv <= 3) fl = p->sld_pr * fl;
                        else if (iv >= 5) fl = p->sld_prm * fl;
                    }
                    dvv[IDX(m, it, ip, ir)] += fl;
                }
            }
        /* ---- Diffusion_Th ---- */
        for (ir = 0; ir < nr; ir++) {
            double rm = (ir == 0) ? pr->r[0] - 0.5 * pr->dr[0] : 0.5 * (pr->r[ir] + pr->r[ir - 1]);
            double rp = (ir == nr - 1) ? pr->r[nr - 1] + 0.5 * pr->dr[nr - 1] : 0.5 * (pr->r[ir + 1] + pr->r[ir]);
            double rfac = 1.5 * (rp * rp - rm * rm) / (rp * rp * rp - rm * rm * rm);
            for (ip = 0; ip < nphi; ip++) {
                double gm, gp;
                for (it = 0; it < nth; it++) tmp[it] = fv[IDX(m, it, ip, ir)];
                switch (iv) {           /* b1(:,4), b2(:,1) at the theta walls (:1008-1120) */
                case 2: gm = 0.0; gp = 0.0; break;
                case 3:
                    gm = -10.0 / 3.0 * tmp[0] / m->sines[0] + 6.0 * tmp[1] / m->sines[1] -
                         2.0 * tmp[2] / m->sines[2] + 1.0 / 3.0 * tmp[3] / m->sines[3];
                    gp = -10.0 / 3.0 * tmp[nth - 1] / m->sines[nth - 1] + 6.0 * tmp[nth - 2] / m->sines[nth - 2] -
                         2.0 * tmp[nth - 3] / m->sines[nth - 3] + 1.0 / 3.0 * tmp[nth - 4] / m->sines[nth - 4];
                    break;
                case 6:
                    gm = p->theta_normal_field ? 0.0 : tmp[0];
                    gp = p->theta_normal_field ? 0.0 : tmp[nth - 1];
                    break;
                case 7: case 8:
                    gm = p->theta_perf_cond ? 0.0 : tmp[0];
                    gp = p->theta_perf_cond ? 0.0 : tmp[nth - 1];
                    break;
                default: gm = tmp[0]; gp = tmp[nth - 1]; break;
                }
                sld_faces(nth, tmp, gm, gm, gp, gp, tiny, delue);
                delue[0] = 0.0;
                delue[nth] = 0.0;
                for (k = 0; k <= nth; k++) {
                    double th = (k == 0) ? pr->theta[0] - 0.5 * dth
                              : (k == nth) ? pr->theta[nth - 1] + 0.5 * dth : 0.5 * (pr->theta[k] + pr->theta[k - 1]);
                    double sph = (k == 0) ? fabs(sin(th)) : sin(th);
                    double vt = p->sld_coef * sqrt(0.5 * (VC(k + 1, ip + 1, ir + 1) + VC(k, ip + 1, ir + 1)));
                    delue[k] = vt * sph * delue[k];
                }
                for (it = 0; it < nth; it++) {
                    double thl = (it == 0) ? pr->theta[0] - 0.5 * dth : 0.5 * (pr->theta[it] + pr->theta[it - 1]);
                    double thh = (it == nth - 1) ? pr->theta[nth - 1] + 0.5 * dth : 0.5 * (pr->theta[it + 1] + pr->theta[it]);
                    double t = rfac / (cos(thl) - cos(thh));
                    double fl = t * (delue[it + 1] - delue[it]);
                    if (p->sld_pr != 1.0 || p->sld_prm != 1.0) {
                        if (iv < 5) fl = p->sld_pr * fl;
                        else if (iv > 5) fl = p->sld_prm * fl;
                    }
                    dvv[IDX(m, it, ip, ir)] += fl;
                }
            }
        }
        /* ---- Diffusion_Phi ---- */
        for (ir = 0; ir < nr; ir++) {
            double rm = (ir == 0) ? pr->r[0] - 0.5 * pr->dr[0] : 0.5 * (pr->r[ir] + pr->r[ir - 1]);
            double rp = (ir == nr - 1) ? pr->r[nr - 1] + 0.5 * pr->dr[nr - 1] : 0.5 * (pr->r[ir + 1] + pr->r[ir]);
            double rfac = 1.5 * (rp * rp - rm * rm) / (rp * rp * rp - rm * rm * rm) / dphi;
            for (it = 0; it < nth; it++) {
                double thl = (it == 0) ? pr->theta[0] - 0.5 * dth : 0.5 * (pr->theta[it] + pr->theta[it - 1]);
                double thh = (it == nth - 1) ? pr->theta[nth - 1] + 0.5 * dth : 0.5 * (pr->theta[it + 1] + pr->theta[it]);
                double geom = ((thh - thl) / (cos(thl) - cos(thh))) * rfac;
                for (ip = 0; ip < nphi; ip++) tmp[ip] = fv[IDX(m, it, ip, ir)];
                sld_faces(nphi, tmp, tmp[nphi - 2], tmp[nphi - 1], tmp[0], tmp[1], tiny, delue);
                for (k = 0; k <= nphi; k++) {
                    double vt = p->sld_coef * sqrt(0.5 * (VC(it + 1, k, ir + 1) + VC(it + 1, k + 1, ir + 1)));
                    delue[k] = vt * delue[k];
                }
                for (ip = 0; ip < nphi; ip++) {
                    double fl = delue[ip + 1] - delue[ip];
                    if (p->sld_pr != 1.0 || p->sld_prm != 1.0) {
                        if (iv < 5) fl = p->sld_pr * fl;
                        else if (iv > 5) fl = p->sld_prm * fl;
                    }
                    dvv[IDX(m, it, ip, ir)] += geom * fl;
                }
            }
        }
    }
    /* viscous heating (:375-393) with temperature = 1/T, then mass diffusion (:395-428) */
    if (p->viscous_heating)
        for (iv = 2; iv <= 4; iv++)
            for (i = 0; i < N; i++) dv[4 * N + i] = dv[4 * N + i] - v[(iv - 1) * N + i] * dv[(iv - 1) * N + i] * itemp[i];
    if (p->mass_diffusion) {
        for (iv = 2; iv <= 4; iv++)
            for (i = 0; i < N; i++) dv[(iv - 1) * N + i] = dv[(iv - 1) * N + i] - invrho[i] * v[(iv - 1) * N + i] * dv[i];
        for (i = 0; i < N; i++) dv[4 * N + i] = dv[4 * N + i] - (m->gamm1 * m->Cv) * invrho[i] * dv[i];
    }
    /* Ohmic heating (:430-461): curl of the A tendency against B */
    if (p->magnetic && p->ohmic_heating) {
        compute_b(m, dv + 5 * N, curl, work, work2, glo, ghi, 0);
        for (iv = 6; iv <= 8; iv++)
            for (i = 0; i < N; i++) {
                double t = (0.5 * ONE_OVER_4PI) * invrho[i] * itemp[i];
                dv[4 * N + i] = dv[4 * N + i] - v[(iv - 1) * N + i] * curl[(iv - 6) * N + i] * t;
            }
    }
#undef VC
    free(vcc); free(f); free(invrho); free(itemp); free(tmp); free(delue); free(work); free(work2); free(curl);
    free(glo); free(ghi); free(m->sines);
    return 0;
}

/* nsteps of Step_RK4 (step_rk4.f90:339-493), SLDiffusion branch: k1 with Euler diffusion. */
int orc_rk4(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, double *y0, double dt,
            int nsteps)
{
    const int nv = p->magnetic ? 8 : 5;
    const long N = (long)g->nr * g->nth * g->nphi, NV = nv * N;
    double *y1 = (double *)malloc(sizeof(double) * NV), *k = (double *)malloc(sizeof(double) * NV);
    double *save = (double *)malloc(sizeof(double) * NV), *dd = (double *)malloc(sizeof(double) * NV);
    double *B = (double *)malloc(sizeof(double) * 3 * N);
    long i;
    int step, rc = 0;
    if (!y1 || !k || !save || !dd || !B) return 12;
    for (step = 0; step < nsteps && !rc; step++) {
        double dtcur = dt / 2.0;
        memcpy(y1, y0, sizeof(double) * NV);
        rc |= orc_rhs(g, p, pr, y1, k, B);
        if (p->sld_on) {
            /* Compute_Diffusion(1) sees vars1 with B in slots 6-8 (RHS(1) leaves it there) */
            double *yb = dd;
            memcpy(yb, y1, sizeof(double) * NV);
            if (p->magnetic) memcpy(yb + 5 * N, B, sizeof(double) * 3 * N);
            {
                double *D = (double *)calloc(NV, sizeof(double));
                rc |= orc_diffusion(g, p, pr, yb, y1 + 5 * N, D);   /* A = slots 6-8 of vars0 */
                for (i = 0; i < NV; i++) save[i] = D[i] + k[i] / 6.0;
                free(D);
            }
        } else {
            for (i = 0; i < NV; i++) save[i] = k[i];
        }
        for (i = 0; i < NV; i++) y1[i] = y0[i] + dtcur * k[i];
        orc_bound(g, p, pr, y1);
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) save[i] = save[i] + k[i] / 3.0;
        for (i = 0; i < NV; i++) y1[i] = y0[i] + dtcur * k[i];
        orc_bound(g, p, pr, y1);
        dtcur = dt;
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) save[i] = save[i] + k[i] / 3.0;
        for (i = 0; i < NV; i++) y1[i] = y0[i] + dtcur * k[i];
        orc_bound(g, p, pr, y1);
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) save[i] = save[i] + k[i] / 6.0;
        for (i = 0; i < NV; i++) y1[i] = y0[i] + dtcur * save[i];
        orc_bound(g, p, pr, y1);
        memcpy(y0, y1, sizeof(double) * NV);
    }
    free(y1); free(k); free(save); free(dd); free(B);
    return rc;
}

/* Stage 1 shared by the TVD schemes: k1 = RHS(1)(y0), the Euler diffusion D of y0 (with B as
 * RHS(1) leaves it) when SLDiffusion, and y1 = Bound(y0 + dt k1). */
static int tvd_first_stage(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr,
                           const double *y0, double dt, long NV, long N, double *y1, double *k, double *D,
                           double *B, double *tmp)
{
    long i;
    int rc = 0;
    memcpy(y1, y0, sizeof(double) * NV);
    rc |= orc_rhs(g, p, pr, y1, k, B);
    if (p->sld_on) {
        memcpy(tmp, y1, sizeof(double) * NV);
        if (p->magnetic) memcpy(tmp + 5 * N, B, sizeof(double) * 3 * N);
        rc |= orc_diffusion(g, p, pr, tmp, y1 + 5 * N, D);
    }
    for (i = 0; i < NV; i++) y1[i] = y0[i] + dt * k[i];
    orc_bound(g, p, pr, y1);
    return rc;
}

/* nsteps of Step_RK2_TVD (step_rk4.f90:31-162): y1 = y0 + dt k1;
 * y = y0/2 + y1/2 + (dt/2) k2 (+ dt D). */
int orc_rk2_tvd(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, double *y0, double dt,
                int nsteps)
{
    const int nv = p->magnetic ? 8 : 5;
    const long N = (long)g->nr * g->nth * g->nphi, NV = nv * N;
    double *y1 = (double *)malloc(sizeof(double) * NV), *k = (double *)malloc(sizeof(double) * NV);
    double *D = (double *)calloc(NV, sizeof(double)), *tmp = (double *)malloc(sizeof(double) * NV);
    double *B = (double *)malloc(sizeof(double) * 3 * N);
    long i;
    int step, rc = 0;
    if (!y1 || !k || !D || !tmp || !B) return 12;
    for (step = 0; step < nsteps && !rc; step++) {
        double dtcur;
        rc |= tvd_first_stage(g, p, pr, y0, dt, NV, N, y1, k, D, B, tmp);
        dtcur = 0.5 * dt;
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) y1[i] = 0.5 * y0[i] + 0.5 * y1[i] + dtcur * k[i];
        if (p->sld_on)
            for (i = 0; i < NV; i++) y1[i] = y1[i] + dt * D[i];
        orc_bound(g, p, pr, y1);
        memcpy(y0, y1, sizeof(double) * NV);
    }
    free(y1); free(k); free(D); free(tmp); free(B);
    return rc;
}

/* nsteps of Step_RK3_TVD (step_rk4.f90:166-335): y1 = y0 + dt k1;
 * y2 = 3/4 y0 + 1/4 y1 + (dt/4) k2; y = 1/3 y0 + 2/3 y2 + (2dt/3) k3 (+ dt D). */
int orc_rk3_tvd(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, double *y0, double dt,
                int nsteps)
{
    const int nv = p->magnetic ? 8 : 5;
    const long N = (long)g->nr * g->nth * g->nphi, NV = nv * N;
    const double onethird = 1.0 / 3.0, twothirds = 2.0 / 3.0;
    double *y1 = (double *)malloc(sizeof(double) * NV), *k = (double *)malloc(sizeof(double) * NV);
    double *D = (double *)calloc(NV, sizeof(double)), *tmp = (double *)malloc(sizeof(double) * NV);
    double *B = (double *)malloc(sizeof(double) * 3 * N);
    long i;
    int step, rc = 0;
    if (!y1 || !k || !D || !tmp || !B) return 12;
    for (step = 0; step < nsteps && !rc; step++) {
        double dtcur;
        rc |= tvd_first_stage(g, p, pr, y0, dt, NV, N, y1, k, D, B, tmp);
        dtcur = dt / 4.0;
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) y1[i] = 0.75 * y0[i] + 0.25 * y1[i] + dtcur * k[i];
        orc_bound(g, p, pr, y1);
        dtcur = 2.0 * dt / 3.0;
        rc |= orc_rhs(g, p, pr, y1, k, NULL);
        for (i = 0; i < NV; i++) y1[i] = onethird * y0[i] + twothirds * y1[i] + dtcur * k[i];
        if (p->sld_on)
            for (i = 0; i < NV; i++) y1[i] = y1[i] + dt * D[i];
        orc_bound(g, p, pr, y1);
        memcpy(y0, y1, sizeof(double) * NV);
    }
    free(y1); free(k); free(D); free(tmp); free(B);
    return rc;
}

/* The six cross-term sweeps of Laplacian_Cross_Terms_Radial / _Theta (rhs_mpi.f90:2070-2296)
 * that orc_rhs forms for state v (Laplacian on), into out (6 fields, order as g_cross_dump). */
int orc_lap_cross(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, const double *v,
                  double *out)
{
    const int nv = p->magnetic ? 8 : 5;
    const long N = (long)g->nr * g->nth * g->nphi;
    double *dv = (double *)malloc(sizeof(double) * nv * N);
    int rc;
    if (!dv) return 12;
    if (!p->laplacian) { free(dv); return 22; }
    g_cross_dump = out;
    rc = orc_rhs(g, p, pr, v, dv, NULL);
    g_cross_dump = NULL;
    free(dv);
    return rc;
}

/* ---------------------------------------------------------------------------------------------
 * Compute_Step_Size (rhs_mpi.f90:92-275), stratified branch, on an FP64 state with A in slots
 * 6-8.  B = curl A with the Compute_B wall conditions (Compute_B(0), :110).  Loops in the
 * reference's order: ir, then it, with the phi sums of each (it, ir) row (:145-185).
 *   dxmin = min(dr, r min(dth, sin(th) dphi)); dth = (th_N - th_1)/(nth-1), dphi = 1/dpi
 *   T = invtconst rho^(gamma-1) exp(s/Cv) (:107)
 *   ut = u^2 + v^2 + w^2 + (gamma-1) Cp xi_cs^2 T; dtmaxc = min dxmin/sqrt(ut) (:169-173)
 *   dtmaxb = min dxmin/sqrt(ut + |B|^2/(4 pi rho))          (:174-181, magnetic)
 *   dtmaxd = min(min dr, r1 min(min(sin th1, sin thN) dphi, dth))^2
 *            / max(max ks, max mur/rho_bar, max etar)        (:194-198, Laplacian; else 1e10)
 *   dt = min(dtmaxb, dtmaxc, dtmaxd) cfl_safety              (:229); dtmaxb/c start at 1e4
 * out[0] = dt, out[1..4] = E_tot, L_tot, m_tot, mrsint (:156-183), out[5..7] = dtmaxc, dtmaxb,
 * dtmaxd.
 * ------------------------------------------------------------------------------------------- */
int orc_step_size(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, const double *v,
                  double cfl_safety, double *out)
{
    orc_mhd M, *m = &M;
    long N, np;
    int nr, nth, nphi, it, ip, ir;
    double *B = NULL, *work, *work2, *glo, *ghi;
    double Etot = 0.0, Ltot = 0.0, mtot = 0.0, mrsint = 0.0, dtmaxc = 1e4, dtmaxb = 1e4, dtmaxd = 1e10;
    double dth, dphi;
    setup(m, g, p, pr);
    N = m->N; nr = m->nr; nth = m->nth; nphi = m->nphi;
    dth = (pr->theta[nth - 1] - pr->theta[0]) / (double)(nth - 1);
    dphi = 1.0 / g->dpi;
    np = (long)nphi * (nr > nth ? nr : nth) * 4 + 16;
    work = (double *)malloc(sizeof(double) * N);
    work2 = (double *)malloc(sizeof(double) * N);
    glo = (double *)calloc(np * 4 > 4L * nth * nphi ? np * 4 : 4L * nth * nphi, sizeof(double));
    ghi = (double *)calloc(np * 4 > 4L * nth * nphi ? np * 4 : 4L * nth * nphi, sizeof(double));
    if (p->magnetic) B = (double *)calloc(3 * N, sizeof(double));
    if (!work || !work2 || !glo || !ghi || (p->magnetic && !B)) return 12;
    if (p->magnetic) compute_b(m, v + 5 * N, B, work, work2, glo, ghi, 1);
    for (ir = 0; ir < nr; ir++) {
        const double radius = pr->r[ir], drad = pr->dr[ir];
        for (it = 0; it < nth; it++) {
            const double sinth = m->sines[it];
            const double dxmin = fmin(drad, radius * fmin(dth, sinth * dphi));
            double srho = 0.0, sl = 0.0, se = 0.0, sb = 0.0, vmax = 1e300, vamax = 1e300;
            for (ip = 0; ip < nphi; ip++) {
                const long i = IDX(m, it, ip, ir);
                const double rho = v[i], w = v[N + i], vv = v[2 * N + i], u = v[3 * N + i], s = v[4 * N + i];
                const double tt = m->invtconst * pow(rho, m->gamm1) * exp(s * m->invCv);
                const double wt = w * w, ut = u * u, vp = vv + radius * sinth * p->omega0;
                srho += rho;
                sl += rho * sqrt(wt + sinth * sinth * vp * vp);
                se += rho * (m->Cv * tt + 0.5 * (ut + wt + vv * vv));
                {
                    const double ct = m->gamm1 * p->Cp * p->xi_cs * p->xi_cs * tt;
                    const double us = ut + vv * vv + wt + ct;
                    vmax = fmin(vmax, dxmin / sqrt(us));
                    if (p->magnetic) {
                        const double b2 = B[i] * B[i] + B[N + i] * B[N + i] + B[2 * N + i] * B[2 * N + i];
                        sb += b2;
                        vamax = fmin(vamax, dxmin / sqrt(us + b2 * ONE_OVER_4PI / rho));
                    }
                }
            }
            mtot += radius * radius * sinth * srho * drad * dth * dphi;
            mrsint += radius * radius * radius * sinth * sinth * sinth * srho * drad * dth * dphi;
            Ltot += radius * radius * radius * sinth * sl * drad * dth * dphi;
            Etot += radius * radius * sinth * se * drad * dth * dphi;
            dtmaxc = fmin(dtmaxc, vmax);
            if (p->magnetic) {
                Etot += 0.5 * radius * radius * sinth * sb * drad * dth * dphi * ONE_OVER_4PI;
                dtmaxb = fmin(dtmaxb, vamax);
            }
        }
    }
    if (p->laplacian) {
        double drmin = pr->dr[0], dmax = 0.0, sinmin = fmin(m->sines[0], m->sines[nth - 1]), x;
        for (ir = 0; ir < nr; ir++) {
            drmin = fmin(drmin, pr->dr[ir]);
            dmax = fmax(dmax, pr->ks[ir]);
            dmax = fmax(dmax, pr->mur[ir] / pr->rho_bar[ir]);
            dmax = fmax(dmax, pr->etar[ir]);
        }
        x = fmin(drmin, pr->r[0] * fmin(sinmin * dphi, dth));
        dtmaxd = x * x / dmax;
    }
    out[0] = fmin(fmin(dtmaxb, dtmaxc), dtmaxd) * cfl_safety;
    out[1] = Etot; out[2] = Ltot; out[3] = mtot; out[4] = mrsint;
    out[5] = dtmaxc; out[6] = dtmaxb; out[7] = dtmaxd;
    free(work); free(work2); free(glo); free(ghi); free(B); free(m->sines);
    return 0;
}
