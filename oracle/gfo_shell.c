/*
 * ORACLE (test infrastructure only) -- Shell_1 restated in plain C, FP64.
 *
 * Follows (homogeneous section, static part):
 *   Shell_1 active DOFs (corners u, mid-sides u + alpha) src/Shell_1.cpp:51-68
 *   Shell_1::PreCalc          src/Shell_1.cpp:1666-1697, 1975-2365
 *     (element frame :1985-1997, Lame constants and drilling stiffness E t^3
 *      :2016-2024, shape functions at the mid-side points :2029-2181, Cowper
 *      6-point rule for loads :2185-2354, initial Gauss-point state :2356-2364)
 *   Shell_1::TransformMatrix  src/Shell_1.cpp:1470-1511
 *   Shell_1::Mount            src/Shell_1.cpp:899-1332
 *   Shell_1::MountFieldLoads  src/Shell_1.cpp:1335-1381 (self-weight, applied twice
 *                             as the reference does)
 *   Shell_1::MountShellSpecialLoads src/Shell_1.cpp:1392-1467 (follower pressure)
 *   Shell_1::SaveLagrange     src/Shell_1.cpp:1650-1664
 *
 * Element products are evaluated densely in the reference's order
 * (K_c = dN^T (Psi^T D Psi) dN, K_g = dN^T G dN, f = dN^T Psi^T sigma).
 *
 * Pinned by tests/test_oracle_shell.py: central finite differences of f_int and
 * of the pressure load vector against the tangents, rigid-body motions giving
 * zero internal force, a membrane patch test (uniform stretch of a mesh gives
 * uniform force), and the small-deflection clamped plate (Timoshenko &
 * Woinowsky-Krieger w_c = 0.00126 q a^4 / D).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "gfo_shell.h"

struct gfo_shell_const {
    double T3[9];                 /* rows e1r, e2r, e3r (global components) */
    double e1r[3], e2r[3];
    double area, alpha1;
    double mu, lambda, thick, stiff_drill, rho;
    double Nu[3][6], Nu1[3][6], Nu2[3][6];   /* quadratic u shapes at the 3 mid-side points */
    double Na[3][3], Na1[3][3], Na2[3][3];   /* linear alpha shapes (nodes 4,5,6) */
    double N4u[6][6], N4u1[6][6], N4u2[6][6], w4[6];  /* Cowper points */
};

gfo_shell_const *gfo_shell_alloc(int n) { return (gfo_shell_const *)calloc(n ? n : 1, sizeof(gfo_shell_const)); }
gfo_shell_const *gfo_shell_at(gfo_shell_const *base, int e) { return base + e; }

static void v_sub(const double *a, const double *b, double *c) { for (int i = 0; i < 3; i++) c[i] = a[i] - b[i]; }
static double v_norm(const double *a) { return sqrt(v3_dot(a, a)); }

static void shapes(double L1, double L2, double L3, const double *dL1, const double *dL2,
                   double *Nu, double *Nu1, double *Nu2, double *Na, double *Na1, double *Na2)
{
    /* dL[k] = (dLk/dx1, dLk/dx2) */
    double L1x = dL1[0], L2x = dL1[1], L3x = dL1[2];
    double L1y = dL2[0], L2y = dL2[1], L3y = dL2[2];
    Nu[0] = (2 * L1 - 1) * L1; Nu[1] = (2 * L2 - 1) * L2; Nu[2] = (2 * L3 - 1) * L3;
    Nu[3] = 4 * L1 * L2; Nu[4] = 4 * L2 * L3; Nu[5] = 4 * L3 * L1;
    Nu1[0] = 4 * L1x * L1 - L1x; Nu1[1] = 4 * L2x * L2 - L2x; Nu1[2] = 4 * L3x * L3 - L3x;
    Nu1[3] = 4 * L1x * L2 + 4 * L1 * L2x; Nu1[4] = 4 * L2x * L3 + 4 * L2 * L3x; Nu1[5] = 4 * L3x * L1 + 4 * L3 * L1x;
    Nu2[0] = 4 * L1y * L1 - L1y; Nu2[1] = 4 * L2y * L2 - L2y; Nu2[2] = 4 * L3y * L3 - L3y;
    Nu2[3] = 4 * L1y * L2 + 4 * L1 * L2y; Nu2[4] = 4 * L2y * L3 + 4 * L2 * L3y; Nu2[5] = 4 * L3y * L1 + 4 * L3 * L1y;
    if (Na) {
        Na[0] = 1 - 2 * L3; Na[1] = 1 - 2 * L1; Na[2] = 1 - 2 * L2;
        Na1[0] = -2 * L3x; Na1[1] = -2 * L1x; Na1[2] = -2 * L2x;
        Na2[0] = -2 * L3y; Na2[1] = -2 * L1y; Na2[2] = -2 * L2y;
    }
}

void gfo_shell_precalc(const double *x, const double *prop, gfo_shell_const *c, double *gp)
{
    memset(c, 0, sizeof(*c));
    const double *x1 = x, *x2 = x + 3, *x3 = x + 6;
    double d21[3], d31[3], nrm[3];
    v_sub(x2, x1, d21);
    v_sub(x3, x1, d31);
    v3_cross(d21, d31, nrm);
    double A = 0.5 * v_norm(nrm);
    c->area = A;
    double e3[3], e1[3], e2[3];
    double nn = v_norm(nrm);
    for (int i = 0; i < 3; i++) e3[i] = nrm[i] / nn;
    double eg[3] = {1.0, 0.0, 0.0};
    if (fabs(v3_dot(eg, e3)) >= 1.0 - 1e-4) { eg[0] = 0.0; eg[1] = 1.0; }
    double p = v3_dot(eg, e3);
    for (int i = 0; i < 3; i++) e1[i] = eg[i] - p * e3[i];
    double n1 = v_norm(e1);
    for (int i = 0; i < 3; i++) e1[i] = (1.0 / n1) * e1[i];
    v3_cross(e3, e1, e2);
    for (int i = 0; i < 3; i++) { c->T3[i] = e1[i]; c->T3[3 + i] = e2[i]; c->T3[6 + i] = e3[i]; c->e1r[i] = e1[i]; c->e2r[i] = e2[i]; }
    double E = prop[0], nu = prop[1], rho = prop[2], t = prop[3];
    c->mu = E / (2.0 * (1 + nu));
    c->lambda = 2.0 * c->mu * nu / (1 - 2.0 * nu);
    c->stiff_drill = E * t * t * t;
    c->thick = t;
    c->rho = rho;
    /* derivatives of the area coordinates in the (e1r, e2r) plane */
    double t23[3], t31[3], t12[3], t32[3], t13[3], t21[3];
    v_sub(x2, x3, t23); v_sub(x3, x1, t31); v_sub(x1, x2, t12);
    v_sub(x3, x2, t32); v_sub(x1, x3, t13); v_sub(x2, x1, t21);
    double dLx[3] = {0.5 * v3_dot(t23, e2) / A, 0.5 * v3_dot(t31, e2) / A, 0.5 * v3_dot(t12, e2) / A};
    double dLy[3] = {0.5 * v3_dot(t32, e1) / A, 0.5 * v3_dot(t13, e1) / A, 0.5 * v3_dot(t21, e1) / A};
    for (int g = 0; g < 3; g++) {
        const double *xp = x + 3 * (3 + g); /* mid-side nodes 4, 5, 6 */
        double a2[3], a3[3], cr[3];
        double L[3];
        v_sub(x2, xp, a2); v_sub(x3, xp, a3); v3_cross(a2, a3, cr); L[0] = 0.5 * v_norm(cr) / A;
        v_sub(x3, xp, a2); v_sub(x1, xp, a3); v3_cross(a2, a3, cr); L[1] = 0.5 * v_norm(cr) / A;
        v_sub(x1, xp, a2); v_sub(x2, xp, a3); v3_cross(a2, a3, cr); L[2] = 0.5 * v_norm(cr) / A;
        shapes(L[0], L[1], L[2], dLx, dLy, c->Nu[g], c->Nu1[g], c->Nu2[g], c->Na[g], c->Na1[g], c->Na2[g]);
    }
    /* Cowper (1973) six-point rule, constants as src/Shell_1.cpp:2190-2233 */
    static const double cw[6][4] = {
        {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.109951743655322},
        {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.109951743655322},
        {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.109951743655322},
        {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.223381589678011},
        {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.223381589678011},
        {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.223381589678011}};
    for (int g = 0; g < 6; g++) {
        c->w4[g] = A * cw[g][3];
        shapes(cw[g][0], cw[g][1], cw[g][2], dLx, dLy, c->N4u[g], c->N4u1[g], c->N4u2[g], NULL, NULL, NULL);
    }
    c->alpha1 = A / 3.0;
    for (int g = 0; g < 3; g++) {
        double *st = gp + g * GFO_SHELL_GPS;
        memset(st, 0, sizeof(double) * GFO_SHELL_GPS);
        m3_eye(st);            /* Q_i */
        st[21] = 1.0;          /* z_x1_i = e1 (local) */
        st[25] = 1.0;          /* z_x2_i = e2 (local) */
    }
}

/* Per-point kinematics shared by Mount and SaveLagrange.  GP state layout:
 * Q_i 0-8, alpha_i 9-11, u_i 12-14, kr1 15-17, kr2 18-20, zx1 21-23, zx2 24-26. */
typedef struct shell_kin {
    double u[3], u1[3], u2[3], a[3], a1[3], a2[3];
    double Q[9], Xi[9], Xi1[9], Xi2[9], g;
    double z1[3], z2[3], QQ[9];
    double eta1[3], eta2[3], k1[3], k2[3];
} shell_kin;

static void shell_kinematics(const gfo_shell_const *c, const double *st, const double *d, int g, shell_kin *k)
{
    for (int i = 0; i < 3; i++) {
        k->u[i] = k->u1[i] = k->u2[i] = 0.0;
        for (int a = 0; a < 6; a++) {
            double v = d[6 * a + i];
            k->u[i] += v * c->Nu[g][a];
            k->u1[i] += v * c->Nu1[g][a];
            k->u2[i] += v * c->Nu2[g][a];
        }
        k->a[i] = k->a1[i] = k->a2[i] = 0.0;
        for (int b = 0; b < 3; b++) {
            double v = d[6 * (3 + b) + 3 + i];
            k->a[i] += v * c->Na[g][b];
            k->a1[i] += v * c->Na1[g][b];
            k->a2[i] += v * c->Na2[g][b];
        }
    }
    m3_vec(c->T3, k->u, k->u);
    m3_vec(c->T3, k->u1, k->u1);
    m3_vec(c->T3, k->u2, k->u2);
    m3_vec(c->T3, k->a, k->a);
    m3_vec(c->T3, k->a1, k->a1);
    m3_vec(c->T3, k->a2, k->a2);
    gfo_rodrigues(k->a, k->Q, k->Xi, &k->g);
    double A1[9], A2[9];
    m3_skew(k->a1, A1);
    m3_skew(k->a2, A2);
    double p1 = v3_dot(k->a, k->a1), p2 = v3_dot(k->a, k->a2);
    for (int i = 0; i < 9; i++) {
        k->Xi1[i] = -0.5 * k->g * (p1 * k->Xi[i] - A1[i]);
        k->Xi2[i] = -0.5 * k->g * (p2 * k->Xi[i] - A2[i]);
    }
    const double *Qi = st, *kr1 = st + 15, *kr2 = st + 18, *zx1 = st + 21, *zx2 = st + 24;
    for (int i = 0; i < 3; i++) { k->z1[i] = k->u1[i] + zx1[i]; k->z2[i] = k->u2[i] + zx2[i]; }
    m3_mul(k->Q, Qi, k->QQ);
    m3_trvec(k->QQ, k->z1, k->eta1);
    m3_trvec(k->QQ, k->z2, k->eta2);
    k->eta1[0] -= 1.0;
    k->eta2[1] -= 1.0;
    double t[3];
    m3_trvec(k->Xi, k->a1, t);
    m3_trvec(Qi, t, k->k1);
    m3_trvec(k->Xi, k->a2, t);
    m3_trvec(Qi, t, k->k2);
    for (int i = 0; i < 3; i++) { k->k1[i] += kr1[i]; k->k2[i] += kr2[i]; }
}

static void mm_(int r, int kk, int cc, const double *A, const double *B, double *C)
{
    for (int i = 0; i < r; i++)
        for (int j = 0; j < cc; j++) {
            double s = 0.0;
            for (int l = 0; l < kk; l++) s += A[i * kk + l] * B[l * cc + j];
            C[i * cc + j] = s;
        }
}
static void mtm_(int r, int kk, int cc, const double *A, const double *B, double *C)
{
    for (int i = 0; i < r; i++)
        for (int j = 0; j < cc; j++) {
            double s = 0.0;
            for (int l = 0; l < kk; l++) s += A[l * r + i] * B[l * cc + j];
            C[i * cc + j] = s;
        }
}
static void put3(double *M, int ld, int r0, int c0, const double *a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[(r0 + i) * ld + c0 + j] = a[3 * i + j];
}
static void acc3(double *a, double s, const double *b) { for (int i = 0; i < 9; i++) a[i] += s * b[i]; }

/* Through-thickness integration of the hyperelastic law (src/Shell_1.cpp:1064-1162). */
static void thickness_integration(const gfo_shell_const *c, const shell_kin *k, double *D, double *sig)
{
    static const double xi[3] = {-0.77459666924148337703585307995648, 0.0, +0.77459666924148337703585307995648};
    static const double wt[3] = {0.55555555555555555555555555555556, 0.88888888888888888888888888888889,
                                 0.55555555555555555555555555555556};
    double mu = c->mu, la = c->lambda, th = c->thick, J = th / 2.0;
    double dn1e1[9] = {0}, dn1e2[9] = {0}, dn1k1[9] = {0}, dn1k2[9] = {0};
    double dn2e1[9] = {0}, dn2e2[9] = {0}, dn2k1[9] = {0}, dn2k2[9] = {0};
    double dm1e1[9] = {0}, dm1e2[9] = {0}, dm1k1[9] = {0}, dm1k2[9] = {0};
    double dm2e1[9] = {0}, dm2e2[9] = {0}, dm2k1[9] = {0}, dm2k2[9] = {0};
    double n1[3] = {0}, n2[3] = {0}, m1[3] = {0}, m2[3] = {0};
    const double e3[3] = {0, 0, 1};
    double E3[9];
    m3_skew(e3, E3);
    for (int q = 0; q < 3; q++) {
        double zeta = th * xi[q] / 2.0;
        double c1[3], c2[3], g1[3], g2[3];
        v3_cross(k->k1, e3, c1);
        v3_cross(k->k2, e3, c2);
        for (int i = 0; i < 3; i++) { g1[i] = k->eta1[i] + zeta * c1[i]; g2[i] = k->eta2[i] + zeta * c2[i]; }
        double g11 = g1[0], g12 = g1[1], g13 = g1[2], g21 = g2[0], g22 = g2[1], g23 = g2[2];
        double jb = (1.0 + g11) * (1.0 + g22) - g12 * g21;
        double v = (la * (jb * jb * jb - 1.0) + 2.0 * mu * (jb - 1.0)) / (la * jb * jb * jb + 2.0 * mu * jb);
        double dv = ((la + 2.0 * mu) * (3.0 * la * jb * jb + 2.0 * mu)) /
                    (jb * jb * (la * jb * jb + 2.0 * mu) * (la * jb * jb + 2.0 * mu));
        double tau1[3] = {mu * v * (1.0 + g22) + mu * (g11 - g22), mu * v * (-g21) + mu * (g12 + g21), 0 + mu * g13};
        double tau2[3] = {mu * v * (-g12) + mu * (g12 + g21), mu * v * (1.0 + g11) + mu * (g22 - g11), 0 + mu * g23};
        double C11[9] = {0}, C22[9] = {0}, C12[9] = {0}, C21[9];
        C11[0] = mu * ((1.0 + g22) * (1.0 + g22) * dv + 1.0);
        C11[1] = -mu * (1.0 + g22) * g21 * dv;
        C11[3] = -mu * (1.0 + g22) * g21 * dv;
        C11[4] = mu * (g21 * g21 * dv + 1.0);
        C11[8] = mu;
        C22[0] = mu * (g12 * g12 * dv + 1.0);
        C22[1] = -mu * ((1.0 + g11) * g12 * dv);
        C22[3] = -mu * ((1.0 + g11) * g12 * dv);
        C22[4] = mu * ((1.0 + g11) * (1.0 + g11) * dv + 1.0);
        C22[8] = mu;
        C12[0] = -mu * ((1.0 + g22) * g12 * dv);
        C12[1] = mu * (v - 1.0 + (1.0 + g11) * (1.0 + g22) * dv);
        C12[3] = mu * (1.0 - v + g12 * g21 * dv);
        C12[4] = -mu * ((1.0 + g11) * g21 * dv);
        m3_tr(C12, C21);
        double w = wt[q] * J;
        double T[9], T2[9];
        acc3(dn1e1, w, C11);
        acc3(dn1e2, w, C12);
        m3_mul(C11, E3, T); acc3(dn1k1, -w * zeta, T);
        m3_mul(C12, E3, T); acc3(dn1k2, -w * zeta, T);
        acc3(dn2e1, w, C21);
        acc3(dn2e2, w, C22);
        m3_mul(C21, E3, T); acc3(dn2k1, -w * zeta, T);
        m3_mul(C22, E3, T); acc3(dn2k2, -w * zeta, T);
        m3_mul(E3, C11, T); acc3(dm1e1, w * zeta, T);
        m3_mul(E3, C12, T); acc3(dm1e2, w * zeta, T);
        m3_mul(E3, C11, T); m3_mul(T, E3, T2); acc3(dm1k1, -w * zeta * zeta, T2);
        m3_mul(E3, C12, T); m3_mul(T, E3, T2); acc3(dm1k2, -w * zeta * zeta, T2);
        m3_mul(E3, C21, T); acc3(dm2e1, w * zeta, T);
        m3_mul(E3, C22, T); acc3(dm2e2, w * zeta, T);
        m3_mul(E3, C21, T); m3_mul(T, E3, T2); acc3(dm2k1, -w * zeta * zeta, T2);
        m3_mul(E3, C22, T); m3_mul(T, E3, T2); acc3(dm2k2, -w * zeta * zeta, T2);
        double x1[3], x2[3];
        v3_cross(e3, tau1, x1);
        v3_cross(e3, tau2, x2);
        for (int i = 0; i < 3; i++) {
            n1[i] += w * tau1[i];
            n2[i] += w * tau2[i];
            m1[i] += w * zeta * x1[i];
            m2[i] += w * zeta * x2[i];
        }
    }
    memset(D, 0, sizeof(double) * 144);
    put3(D, 12, 0, 0, dn1e1); put3(D, 12, 0, 3, dn1k1); put3(D, 12, 0, 6, dn1e2); put3(D, 12, 0, 9, dn1k2);
    put3(D, 12, 3, 0, dm1e1); put3(D, 12, 3, 3, dm1k1); put3(D, 12, 3, 6, dm1e2); put3(D, 12, 3, 9, dm1k2);
    put3(D, 12, 6, 0, dn2e1); put3(D, 12, 6, 3, dn2k1); put3(D, 12, 6, 6, dn2e2); put3(D, 12, 6, 9, dn2k2);
    put3(D, 12, 9, 0, dm2e1); put3(D, 12, 9, 3, dm2k1); put3(D, 12, 9, 6, dm2e2); put3(D, 12, 9, 9, dm2k2);
    /* drilling stiffness (src/Shell_1.cpp:1212-1217) */
    D[5 * 12 + 5] = c->stiff_drill;
    D[11 * 12 + 11] = c->stiff_drill;
    m1[2] = c->stiff_drill * k->k1[2];
    m2[2] = c->stiff_drill * k->k2[2];
    for (int i = 0; i < 3; i++) { sig[i] = n1[i]; sig[3 + i] = m1[i]; sig[6 + i] = n2[i]; sig[9 + i] = m2[i]; }
}

void gfo_shell_mount(const gfo_shell_const *c, const double *gp, const double *d, double *K, double *P, double *res)
{
    double Kl[27 * 27], fl[27];
    memset(Kl, 0, sizeof(Kl));
    memset(fl, 0, sizeof(fl));
    if (res) memset(res, 0, 24 * sizeof(double));
    for (int g = 0; g < 3; g++) {
        const double *st = gp + g * GFO_SHELL_GPS;
        shell_kin k;
        shell_kinematics(c, st, d, g, &k);
        double D[144], sig[12];
        thickness_integration(c, &k, D, sig);
        /* Shell_1::WriteResults src/Shell_1.cpp:636-646: x_avg = x_avg + (1/3) x_gp */
        if (res) {
            const double *src[8] = {sig, sig + 3, sig + 6, sig + 9, k.eta1, k.k1, k.eta2, k.k2};
            for (int q = 0; q < 8; q++)
                for (int i = 0; i < 3; i++) res[3 * q + i] = res[3 * q + i] + (1.0 / 3.0) * src[q][i];
        }
        /* deltaN (15 x 27), src/Shell_1.cpp:2114-2161 */
        double dN[15 * 27];
        memset(dN, 0, sizeof(dN));
        for (int a = 0; a < 6; a++)
            for (int i = 0; i < 3; i++) {
                dN[(0 + i) * 27 + 3 * a + i] = c->Nu1[g][a];
                dN[(6 + i) * 27 + 3 * a + i] = c->Nu2[g][a];
            }
        for (int b = 0; b < 3; b++)
            for (int i = 0; i < 3; i++) {
                dN[(3 + i) * 27 + 18 + 3 * b + i] = c->Na1[g][b];
                dN[(9 + i) * 27 + 18 + 3 * b + i] = c->Na2[g][b];
                dN[(12 + i) * 27 + 18 + 3 * b + i] = c->Na[g][b];
            }
        /* Psi (12 x 15), src/Shell_1.cpp:1239-1268 */
        double Qt[9], tM[9], Z1[9], Z2[9], t1[9];
        m3_tr(k.QQ, Qt);
        m3_skew(k.z1, Z1);
        m3_skew(k.z2, Z2);
        double Psi[12 * 15];
        memset(Psi, 0, sizeof(Psi));
        put3(Psi, 15, 0, 0, Qt);
        put3(Psi, 15, 6, 6, Qt);
        m3_mul(Qt, k.Xi, tM);
        put3(Psi, 15, 3, 3, tM);
        put3(Psi, 15, 9, 9, tM);
        m3_mul(Qt, Z1, t1); m3_mul(t1, k.Xi, tM); put3(Psi, 15, 0, 12, tM);
        m3_mul(Qt, k.Xi1, tM); put3(Psi, 15, 3, 12, tM);
        m3_mul(Qt, Z2, t1); m3_mul(t1, k.Xi, tM); put3(Psi, 15, 6, 12, tM);
        m3_mul(Qt, k.Xi2, tM); put3(Psi, 15, 9, 12, tM);
        /* K_c = alpha1 dN^T (((Psi^T D) Psi) dN) */
        double PtD[15 * 12], PtDP[15 * 15], M1[15 * 27], Kc[27 * 27];
        mtm_(15, 12, 12, Psi, D, PtD);
        mm_(15, 12, 15, PtD, Psi, PtDP);
        mm_(15, 15, 27, PtDP, dN, M1);
        mtm_(27, 15, 27, dN, M1, Kc);
        for (int i = 0; i < 27 * 27; i++) Kl[i] += c->alpha1 * Kc[i];
        /* forces and moments, G (15 x 15), src/Shell_1.cpp:1277-1319 */
        double n_1[3], n_2[3], m_1[3], m_2[3];
        m3_vec(k.QQ, sig + 0, n_1);
        m3_vec(k.QQ, sig + 6, n_2);
        m3_vec(k.QQ, sig + 3, m_1);
        m3_vec(k.QQ, sig + 9, m_2);
        double Zn1[3], Zn2[3], VZ1[9], VZ2[9], Vm1[9], Vm2[9], dV1[9], dV2[9];
        m3_vec(Z1, n_1, Zn1);
        m3_vec(Z2, n_2, Zn2);
        gfo_V(k.a, Zn1, VZ1);
        gfo_V(k.a, Zn2, VZ2);
        gfo_V(k.a, m_1, Vm1);
        gfo_V(k.a, m_2, Vm2);
        gfo_dV(k.a, k.a1, m_1, dV1);
        gfo_dV(k.a, k.a2, m_2, dV2);
        double Sn1[9], Sn2[9], Sm1[9], Sm2[9];
        m3_skew(n_1, Sn1); m3_skew(n_2, Sn2); m3_skew(m_1, Sm1); m3_skew(m_2, Sm2);
        double Gua1[9], Gua2[9], Gau1[9], Gau2[9], Vm1t[9], Vm2t[9];
        m3_mul(Sn1, k.Xi, Gua1); for (int i = 0; i < 9; i++) Gua1[i] = -1.0 * Gua1[i];
        m3_mul(Sn2, k.Xi, Gua2); for (int i = 0; i < 9; i++) Gua2[i] = -1.0 * Gua2[i];
        m3_tr(Gua1, Gau1);
        m3_tr(Gua2, Gau2);
        m3_tr(Vm1, Vm1t);
        m3_tr(Vm2, Vm2t);
        double XiT[9], Gaa[9], ta[9], tb[9], Xi1T[9], Xi2T[9];
        m3_tr(k.Xi, XiT);
        m3_tr(k.Xi1, Xi1T);
        m3_tr(k.Xi2, Xi2T);
        /* Xi^T (Z_x1 skew(n_1)) Xi - V(a, Z_x1 n_1) + dV(a, a_x1, m_1) - Xi_x1^T (skew(m_1) Xi), + same for 2 */
        m3_mul(Z1, Sn1, ta); m3_mul(XiT, ta, tb); m3_mul(tb, k.Xi, ta);
        for (int i = 0; i < 9; i++) Gaa[i] = ta[i] - VZ1[i] + dV1[i];
        m3_mul(Sm1, k.Xi, ta); m3_mul(Xi1T, ta, tb);
        for (int i = 0; i < 9; i++) Gaa[i] -= tb[i];
        double Gaa2[9];
        m3_mul(Z2, Sn2, ta); m3_mul(XiT, ta, tb); m3_mul(tb, k.Xi, ta);
        for (int i = 0; i < 9; i++) Gaa2[i] = ta[i] - VZ2[i] + dV2[i];
        m3_mul(Sm2, k.Xi, ta); m3_mul(Xi2T, ta, tb);
        for (int i = 0; i < 9; i++) Gaa2[i] -= tb[i];
        for (int i = 0; i < 9; i++) Gaa[i] += Gaa2[i];
        double G[15 * 15];
        memset(G, 0, sizeof(G));
        put3(G, 15, 0, 12, Gua1);
        put3(G, 15, 3, 12, Vm1t);
        put3(G, 15, 6, 12, Gua2);
        put3(G, 15, 9, 12, Vm2t);
        put3(G, 15, 12, 0, Gau1);
        put3(G, 15, 12, 3, Vm1);
        put3(G, 15, 12, 6, Gau2);
        put3(G, 15, 12, 9, Vm2);
        put3(G, 15, 12, 12, Gaa);
        double Kg[27 * 27];
        mm_(15, 15, 27, G, dN, M1);
        mtm_(27, 15, 27, dN, M1, Kg);
        for (int i = 0; i < 27 * 27; i++) Kl[i] += c->alpha1 * Kg[i];
        /* f += alpha1 dN^T Psi^T sigma */
        double Pts[15], fg[27];
        mtm_(15, 12, 1, Psi, sig, Pts);
        mtm_(27, 15, 1, dN, Pts, fg);
        for (int i = 0; i < 27; i++) fl[i] += c->alpha1 * fg[i];
    }
    /* K = T^T K T, f = T^T f, T = blockdiag(T3 x 9) */
    double T[27 * 27], TK[27 * 27];
    memset(T, 0, sizeof(T));
    for (int b = 0; b < 9; b++) put3(T, 27, 3 * b, 3 * b, c->T3);
    mtm_(27, 27, 27, T, Kl, TK);
    mm_(27, 27, 27, TK, T, K);
    mtm_(27, 27, 1, T, fl, P);
}

/* MountFieldLoads: self-weight, P -= e_loading; the reference accumulates the
 * same weight twice (src/Shell_1.cpp:1340-1354 and :1357-1375). */
void gfo_shell_gravity(const gfo_shell_const *c, double lf, const double *G, double *P)
{
    double mult = lf * 1.0 * c->rho * c->thick;
    for (int rep = 0; rep < 2; rep++)
        for (int g = 0; g < 6; g++)
            for (int a = 0; a < 6; a++)
                for (int k = 0; k < 3; k++) P[3 * a + k] -= c->w4[g] * c->N4u[g][a] * mult * G[k];
}

void gfo_shell_pressure(const gfo_shell_const *c, const double *xc, const double *d, double p, int area_update,
                        double *K, double *P)
{
    (void)xc;
    /* pu_ip = (copy - ref) + incremental: passed in as xc (total displacement of the 6 nodes, [18]) */
    double pu[18];
    for (int a = 0; a < 6; a++)
        for (int k = 0; k < 3; k++) pu[3 * a + k] = xc[3 * a + k] + d[6 * a + k];
    double mult = 1.0;
    for (int g = 0; g < 6; g++) {
        double t1[3], t2[3];
        for (int k = 0; k < 3; k++) {
            t1[k] = c->e1r[k];
            t2[k] = c->e2r[k];
            for (int a = 0; a < 6; a++) {
                t1[k] += c->N4u1[g][a] * pu[3 * a + k];
                t2[k] += c->N4u2[g][a] * pu[3 * a + k];
            }
        }
        double cr[3];
        v3_cross(t1, t2, cr);
        double ncr = v_norm(cr);
        double n[3] = {cr[0] / ncr, cr[1] / ncr, cr[2] / ncr};
        double q[3] = {-mult * p * n[0], -mult * p * n[1], -mult * p * n[2]};
        double S1[9], S2[9];
        m3_skew(t1, S1);
        m3_skew(t2, S2);
        /* Kpaa (3 x 18) = s (I - n n^T)(skew(t1) N_x2 - skew(t2) N_x1) */
        double Kpaa[3 * 18];
        double B[3 * 18];
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 6; a++)
                for (int j = 0; j < 3; j++)
                    B[i * 18 + 3 * a + j] = S1[3 * i + j] * c->N4u2[g][a] - S2[3 * i + j] * c->N4u1[g][a];
        double jac = ncr;
        if (!area_update) {
            double Pm[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) Pm[3 * i + j] = (i == j ? 1.0 : 0.0) - n[i] * n[j];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 18; j++) {
                    double s = 0.0;
                    for (int l = 0; l < 3; l++) s += Pm[3 * i + l] * B[l * 18 + j];
                    Kpaa[i * 18 + j] = (1.0 / ncr) * s;
                }
            jac = 1.0;
        } else {
            memcpy(Kpaa, B, sizeof(B));
        }
        /* P -= w N4^T q (jac); K += w p N4^T Kp (only translational rows/cols) */
        for (int a = 0; a < 6; a++)
            for (int k = 0; k < 3; k++) P[3 * a + k] -= c->w4[g] * jac * c->N4u[g][a] * q[k];
        for (int a = 0; a < 6; a++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 18; j++)
                    K[(3 * a + i) * 27 + j] += c->w4[g] * mult * p * c->N4u[g][a] * Kpaa[i * 18 + j];
    }
}

void gfo_shell_save(const gfo_shell_const *c, double *gp, const double *d)
{
    for (int g = 0; g < 3; g++) {
        double *st = gp + g * GFO_SHELL_GPS;
        shell_kin k;
        shell_kinematics(c, st, d, g, &k);
        /* kappa_i <- kappa_r (uses the old Q_i), then Q_i <- Q_delta Q_i */
        for (int i = 0; i < 3; i++) { st[15 + i] = k.k1[i]; st[18 + i] = k.k2[i]; }
        double Qn[9];
        m3_mul(k.Q, st, Qn);
        for (int i = 0; i < 9; i++) st[i] = Qn[i];
        for (int i = 0; i < 3; i++) st[12 + i] += k.u[i];
        double a3[3];
        gfo_compose(k.a, st + 9, a3);
        for (int i = 0; i < 3; i++) st[9 + i] = a3[i];
        for (int i = 0; i < 3; i++) { st[21 + i] += k.u1[i]; st[24 + i] += k.u2[i]; }
    }
}

/* Single-element handles for the element-level tests. */
void *gfo_shell_new(const double *x, const double *prop, double *gp)
{
    gfo_shell_const *c = gfo_shell_alloc(1);
    gfo_shell_precalc(x, prop, c, gp);
    return c;
}
void gfo_shell_free(void *c) { free(c); }
