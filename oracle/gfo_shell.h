/*
 * ORACLE (test infrastructure only) -- Shell_1 restatement interface.
 */
#ifndef GFO_SHELL_H
#define GFO_SHELL_H
#include "gfo_internal.h"

gfo_shell_const *gfo_shell_alloc(int n);
gfo_shell_const *gfo_shell_at(gfo_shell_const *base, int e);
void gfo_shell_precalc(const double *x /*[18]*/, const double *prop, gfo_shell_const *c, double *gp);
/* res (or NULL): Gauss-point means of n_r1 m_r1 n_r2 m_r2 eta_r1 kappa_r1 eta_r2 kappa_r2. */
void gfo_shell_mount(const gfo_shell_const *c, const double *gp, const double *d /*[36]*/, double *K, double *P,
                     double *res /*[24]*/);
/* utot: total translation (converged - reference) of the 6 nodes [18] */
void gfo_shell_pressure(const gfo_shell_const *c, const double *utot, const double *d /*[36]*/,
                        double p, int area_update, double *K, double *P);
void gfo_shell_gravity(const gfo_shell_const *c, double lf, const double *G, double *P);
void gfo_shell_save(const gfo_shell_const *c, double *gp, const double *d);
#endif
