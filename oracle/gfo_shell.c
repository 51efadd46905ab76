/*
 * ORACLE (test infrastructure only) -- Shell_1 restatement (placeholder).
 */
#include <stdlib.h>
#include "gfo_shell.h"

struct gfo_shell_const { double unused; };
gfo_shell_const *gfo_shell_alloc(int n) { return (gfo_shell_const *)calloc(n ? n : 1, sizeof(gfo_shell_const)); }
gfo_shell_const *gfo_shell_at(gfo_shell_const *base, int e) { return base + e; }
void gfo_shell_precalc(const double *x, const double *prop, gfo_shell_const *c, double *gp) { (void)x; (void)prop; (void)c; (void)gp; }
void gfo_shell_mount(const gfo_shell_const *c, const double *gp, const double *d, double *K, double *P) { (void)c; (void)gp; (void)d; (void)K; (void)P; }
void gfo_shell_pressure(const gfo_shell_const *c, const double *xc, const double *d, double p, int au, double *K, double *P) { (void)c; (void)xc; (void)d; (void)p; (void)au; (void)K; (void)P; }
void gfo_shell_save(const gfo_shell_const *c, double *gp, const double *d) { (void)c; (void)gp; (void)d; }
