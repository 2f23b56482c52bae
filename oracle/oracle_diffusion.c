/*
 * CPU oracle: Compute_Diffusion (diffusion.f90:70-474), the slope-limited diffusion the
 * reference folds into RK4 stage 1 (step_rk4.f90:369-389).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_deriv.c).
 *
 * With sld_coef = 0 every flux of Diffusion_R/_Th/_Phi is sld_coef*sqrt(...)*delue = 0
 * (diffusion.f90:840, :1245, :1453), so the viscous, mass and Ohmic corrections built from
 * those fluxes (:375-461) vanish too and the tendency is exactly zero.
 */
#include <string.h>

#include "../include/cssx.h"

int orc_diffusion(const cssx_grid *g, const cssx_mhd_params *p, const cssx_mhd_profiles *pr, const double *v,
                  const double *a, double *dv)
{
    const int nv = p->magnetic ? 8 : 5;
    const long N = (long)g->nr * g->nth * g->nphi;
    (void)pr; (void)v; (void)a;
    if (p->sld_coef != 0.0) return 22;    /* the limited fluxes are not restated yet */
    memset(dv, 0, sizeof(double) * nv * N);
    return 0;
}
