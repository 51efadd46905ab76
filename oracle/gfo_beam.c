/*
 * ORACLE (test infrastructure only) -- Beam_1 restated in plain C, FP64.
 *
 * Follows:
 *   Beam_1::PreCalc            src/Beam_1.cpp:501-692 (Hooke / Tube branch)
 *   Beam_1::TransformMatrix    src/Beam_1.cpp:1385-1427
 *   Beam_1::Mount              src/Beam_1.cpp:695-835
 *   Beam_1::SaveLagrange       src/Beam_1.cpp:1494-1506
 *   Pipe_1::PreCalc / Mount    src/Pipe_1.cpp:1106-1135, 836-973 (the Beam_1 kinematics
 *                              with the pipe section's D)
 *   LagrangeSave initial state src/LagrangeSave.cpp:41-49
 *   V, d_V, skew               src/Matrix.cpp:1812-1833, 1999-2033
 *   Node Rodrigues composition src/Node.cpp:342
 *
 * The element products are evaluated densely, in the order the reference
 * writes them (B = B1 B2, K_c = dN^T (B^T D B) dN, K_g = dN^T G dN).
 *
 * Pinned by (tests/test_oracle_beam.py): central finite differences of f_int
 * against K (consistent tangent), rigid-body translation/rotation giving zero
 * strain and force, Q(a) orthogonality and the rotation-angle identity
 * theta = 2 atan(|a|/2), the composition formula against Q(a2)Q(a1), and the
 * cantilever closed forms (Timoshenko tip deflection, the Simo-Vu-Quoc
 * roll-up under a follower end moment).
 */
#include <math.h>
#include <string.h>
#include "gfo_internal.h"

void gfo_rodrigues(const double *a, double *Q, double *Xi, double *g_out)
{
    double aa = sqrt(v3_dot(a, a));
    double g = 4.0 / (4.0 + aa * aa);
    double A[9], A2[9];
    m3_skew(a, A);
    m3_mul(A, A, A2);
    if (Q) {
        for (int i = 0; i < 9; i++) Q[i] = g * (A[i] + 0.5 * A2[i]);
        Q[0] += 1.0; Q[4] += 1.0; Q[8] += 1.0;
    }
    if (Xi) {
        for (int i = 0; i < 9; i++) Xi[i] = g * 0.5 * A[i];
        Xi[0] += g; Xi[4] += g; Xi[8] += g;
    }
    if (g_out) *g_out = g;
}

/* dyadic(a, b)(i,j) = a_i b_j */
static void add_dyad(double *M, double s, const double *a, const double *b)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[3 * i + j] += s * a[i] * b[j];
}

void gfo_V(const double *x, const double *t, double *out)
{
    double ax = sqrt(v3_dot(x, x));
    double h = 4.0 / (4.0 + ax * ax);
    double h2 = 0.5 * h, h4 = -0.25 * h * h, h8 = -0.5 * h * h;
    double xt[3], w[3], S[9];
    v3_cross(x, t, xt);
    for (int i = 0; i < 3; i++) w[i] = h8 * t[i] - h4 * xt[i];
    m3_skew(t, S);
    for (int i = 0; i < 9; i++) out[i] = h2 * S[i];
    add_dyad(out, 1.0, w, x);
}

void gfo_dV(const double *x, const double *dx, const double *t, double *out)
{
    double ax = sqrt(v3_dot(x, x));
    double h = 4.0 / (4.0 + ax * ax);
    double h4 = -0.25 * h * h, h6 = 0.25 * h * h * h, h8 = -0.5 * h * h, h9 = 0.5 * h * h * h;
    double xdx = v3_dot(x, dx);
    double xt[3], dxt[3], w1[3], w2[3], w3[3], S[9];
    v3_cross(x, t, xt);
    v3_cross(dx, t, dxt);
    for (int i = 0; i < 3; i++) {
        w1[i] = h9 * t[i] - h6 * xt[i];
        w2[i] = h8 * t[i] - h4 * xt[i];
        w3[i] = -h4 * dxt[i];
    }
    m3_skew(t, S);
    for (int i = 0; i < 9; i++) out[i] = h4 * xdx * S[i];
    add_dyad(out, xdx, w1, x);
    add_dyad(out, 1.0, w2, dx);
    add_dyad(out, 1.0, w3, x);
}

void gfo_compose(const double *a2, const double *a1, double *a3)
{
    double c[3];
    v3_cross(a2, a1, c);
    double s = 4.0 / (4.0 - v3_dot(a2, a1));
    for (int i = 0; i < 3; i++) a3[i] = s * (a2[i] + a1[i] + 0.5 * c[i]);
}

/* Gauss abscissa exactly as written in src/Beam_1.cpp:629-632. */
#define BEAM_CSI 0.577350269189626

void gfo_beam_precalc(const double *x, const double *cs, const double *prop, int kind,
                      gfo_beam_const *c, double *gp)
{
    double T0 = 0.0;
    memset(c, 0, sizeof(*c));
    if (kind == GF_BEAM_KIND_PIPE1) {
        /* Pipe_1::PreCalc src/Pipe_1.cpp:1108-1135: D = diag(GA, GA, EA, EI, EI, GJ),
         * mass per length Rho, no pretension */
        c->D[0 * 6 + 0] = prop[3];
        c->D[1 * 6 + 1] = prop[3];
        c->D[2 * 6 + 2] = prop[0];
        c->D[3 * 6 + 3] = prop[1];
        c->D[4 * 6 + 4] = prop[1];
        c->D[5 * 6 + 5] = prop[2];
        c->rhoA = prop[4];
    } else {
        double E = prop[0], nu = prop[1], rho = prop[2], A = prop[3];
        double I1 = prop[4], I2 = prop[5], I12 = prop[6], It = prop[7];
        T0 = prop[8];
        double G = E / (2 * (1 + nu));
        double sf = 1.0;
        c->D[0 * 6 + 0] = sf * G * A;
        c->D[1 * 6 + 1] = sf * G * A;
        c->D[2 * 6 + 2] = E * A;
        c->D[3 * 6 + 3] = E * I1;
        c->D[4 * 6 + 4] = E * I2;
        c->D[3 * 6 + 4] = E * I12;
        c->D[4 * 6 + 3] = E * I12;
        c->D[5 * 6 + 5] = G * It;
        c->rhoA = rho * A;
    }
    /* transform3 = [E1; E2; E3] (dot products with the global basis) */
    for (int i = 0; i < 9; i++) c->T3[i] = cs[i];
    double e3[3] = {x[6] - x[0], x[7] - x[1], x[8] - x[2]};
    double L0 = sqrt(v3_dot(e3, e3));
    double inv = 1.0 / L0;
    for (int i = 0; i < 3; i++) e3[i] *= inv;
    m3_vec(c->T3, e3, c->e3r);
    double du0 = T0 / c->D[2 * 6 + 2];
    double length = L0 / (1.0 + du0);
    c->jac = length / 2.0;
    for (int g = 0; g < 2; g++) {
        double s = (g == 0) ? -BEAM_CSI : BEAM_CSI;
        c->N[g][0] = 0.5 * s * (s - 1.0);
        c->N[g][1] = 1.0 - s * s;
        c->N[g][2] = 0.5 * s * (1.0 + s);
        c->dN[g][0] = (1.0 / c->jac) * (s - 0.5);
        c->dN[g][1] = (1.0 / c->jac) * (-2.0 * s);
        c->dN[g][2] = (1.0 / c->jac) * (0.5 + s);
        double *st = gp + g * GFO_BEAM_GPS;
        memset(st, 0, sizeof(double) * GFO_BEAM_GPS);
        m3_eye(st);              /* Q_i = I      */
        st[9 + 3 + 2] = 1.0 + du0; /* z'_i = (0,0,1+du0) */
    }
}

/* Per-Gauss-point kinematics shared by Mount and SaveLagrange. */
typedef struct beam_kin {
    double a[3], da[3], u[3], du[3];
    double Q[9], Xi[9], dXi[9], g;
    double dz[3], QQ[9];
    double eta[3], kap[3];
} beam_kin;

static void beam_kinematics(const gfo_beam_const *c, const double *st, const double *d, int gpi, beam_kin *k)
{
    const double *N = c->N[gpi], *dN = c->dN[gpi];
    for (int i = 0; i < 3; i++) {
        k->a[i] = d[3 + i] * N[0] + d[9 + i] * N[1] + d[15 + i] * N[2];
        k->da[i] = d[3 + i] * dN[0] + d[9 + i] * dN[1] + d[15 + i] * dN[2];
        k->u[i] = d[0 + i] * N[0] + d[6 + i] * N[1] + d[12 + i] * N[2];
        k->du[i] = d[0 + i] * dN[0] + d[6 + i] * dN[1] + d[12 + i] * dN[2];
    }
    m3_vec(c->T3, k->a, k->a);
    m3_vec(c->T3, k->da, k->da);
    m3_vec(c->T3, k->u, k->u);
    m3_vec(c->T3, k->du, k->du);
    gfo_rodrigues(k->a, k->Q, k->Xi, &k->g);
    double dA[9];
    m3_skew(k->da, dA);
    double ada = v3_dot(k->a, k->da);
    for (int i = 0; i < 9; i++) k->dXi[i] = -0.5 * k->g * (ada * k->Xi[i] - dA[i]);
    const double *Qi = st, *kap_i = st + 9, *dz_i = st + 12;
    for (int i = 0; i < 3; i++) k->dz[i] = k->du[i] + dz_i[i];
    m3_mul(k->Q, Qi, k->QQ);
    /* eta_r = (Q Q_i)^T z' - e3r ; kappa_r = Q_i^T Xi^T a' + kappa_i */
    m3_trvec(k->QQ, k->dz, k->eta);
    for (int i = 0; i < 3; i++) k->eta[i] -= c->e3r[i];
    double t[3];
    m3_trvec(k->Xi, k->da, t);
    m3_trvec(Qi, t, k->kap);
    for (int i = 0; i < 3; i++) k->kap[i] += kap_i[i];
}

/* C (r x c) = A (r x k) * B (k x c), row-major */
static void mm(int r, int kk, int cc, const double *A, const double *B, double *C)
{
    for (int i = 0; i < r; i++)
        for (int j = 0; j < cc; j++) {
            double s = 0.0;
            for (int l = 0; l < kk; l++) s += A[i * kk + l] * B[l * cc + j];
            C[i * cc + j] = s;
        }
}
/* C (r x c) = A^T B with A (k x r), B (k x c) */
static void mtm(int r, int kk, int cc, const double *A, const double *B, double *C)
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

void gfo_beam_mount(const gfo_beam_const *c, const double *gp, const double *d, double *K, double *f, double *res)
{
    double Kl[18 * 18], fl[18], sg[2][6], ep[2][6];
    memset(Kl, 0, sizeof(Kl));
    memset(fl, 0, sizeof(fl));
    for (int g = 0; g < 2; g++) {
        const double *st = gp + g * GFO_BEAM_GPS;
        beam_kin k;
        beam_kinematics(c, st, d, g, &k);
        /* deltaN (9 x 18), src/Beam_1.cpp:641-666 */
        double dNm[9 * 18];
        memset(dNm, 0, sizeof(dNm));
        for (int a = 0; a < 3; a++) {
            for (int i = 0; i < 6; i++) dNm[i * 18 + 6 * a + i] = c->dN[g][a];
            for (int i = 0; i < 3; i++) dNm[(6 + i) * 18 + 6 * a + 3 + i] = c->N[g][a];
        }
        /* B = B1 B2 (6 x 9), src/Beam_1.cpp:756-775 */
        double Qt[9], dZ[9], dZXi[9];
        m3_tr(k.QQ, Qt);
        m3_skew(k.dz, dZ);
        m3_mul(dZ, k.Xi, dZXi);
        double B1[36], B2[54], B[54];
        memset(B1, 0, sizeof(B1));
        memset(B2, 0, sizeof(B2));
        put3(B1, 6, 0, 0, Qt);
        put3(B1, 6, 3, 3, Qt);
        B2[0 * 9 + 0] = B2[1 * 9 + 1] = B2[2 * 9 + 2] = 1.0;
        put3(B2, 9, 0, 6, dZXi);
        put3(B2, 9, 3, 3, k.Xi);
        put3(B2, 9, 3, 6, k.dXi);
        mm(6, 6, 9, B1, B2, B);
        /* K_c = dN^T ((B^T D) B) dN */
        double BtD[54], BtDB[81], M1[81 * 2], Kc[18 * 18];
        mtm(9, 6, 6, B, c->D, BtD);
        mm(9, 6, 9, BtD, B, BtDB);
        mm(9, 9, 18, BtDB, dNm, M1);
        mtm(18, 9, 18, dNm, M1, Kc);
        /* strains and stresses */
        double eps[6], sig[6];
        for (int i = 0; i < 3; i++) { eps[i] = k.eta[i]; eps[3 + i] = k.kap[i]; }
        mm(6, 6, 1, c->D, eps, sig);
        for (int i = 0; i < 6; i++) { sg[g][i] = sig[i]; ep[g][i] = eps[i]; }
        double n[3], m[3];
        m3_vec(k.QQ, sig, n);
        m3_vec(k.QQ, sig + 3, m);
        /* G (9x9), src/Beam_1.cpp:799-822 */
        double dZn[3], V1[9], V2[9], dV[9], Sn[9], Sm[9], t1[9], t2[9];
        m3_vec(dZ, n, dZn);
        gfo_V(k.a, dZn, V1);
        gfo_V(k.a, m, V2);
        gfo_dV(k.a, k.da, m, dV);
        m3_skew(n, Sn);
        m3_skew(m, Sm);
        double Gm[81];
        memset(Gm, 0, sizeof(Gm));
        double Gua[9], Gau[9], Gaa[9], Gada[9], XiT[9], dXiT[9];
        m3_mul(Sn, k.Xi, t1);
        for (int i = 0; i < 9; i++) Gua[i] = -1.0 * t1[i];
        m3_tr(k.Xi, XiT);
        m3_mul(XiT, Sn, Gau);
        /* Xi^T (dZ Sn) Xi - V1 + dV - dXi^T (Sm Xi) */
        m3_mul(dZ, Sn, t1);
        m3_mul(XiT, t1, t2);
        m3_mul(t2, k.Xi, t1);
        m3_tr(k.dXi, dXiT);
        double SmXi[9];
        m3_mul(Sm, k.Xi, SmXi);
        m3_mul(dXiT, SmXi, t2);
        for (int i = 0; i < 9; i++) Gaa[i] = t1[i] - V1[i] + dV[i] - t2[i];
        m3_mul(XiT, SmXi, t1);
        for (int i = 0; i < 9; i++) Gada[i] = V2[i] - t1[i];
        put3(Gm, 9, 0, 6, Gua);
        put3(Gm, 9, 6, 0, Gau);
        put3(Gm, 9, 6, 6, Gaa);
        put3(Gm, 9, 3, 6, Gada);
        put3(Gm, 9, 6, 3, V2);
        double Kg[18 * 18];
        mm(9, 9, 18, Gm, dNm, M1);
        mtm(18, 9, 18, dNm, M1, Kg);
        double w = 1.0 * c->jac;
        for (int i = 0; i < 18 * 18; i++) Kl[i] += w * (Kc[i] + Kg[i]);
        /* f += w dN^T B^T sigma */
        double Bts[9], fg[18];
        mtm(9, 6, 1, B, sig, Bts);
        mtm(18, 9, 1, dNm, Bts, fg);
        for (int i = 0; i < 18; i++) fl[i] += w * fg[i];
    }
    /* K = T^T K T, f = T^T f with T = blockdiag(T3 x 6) */
    double T[18 * 18], TK[18 * 18];
    memset(T, 0, sizeof(T));
    for (int b = 0; b < 6; b++) put3(T, 18, 3 * b, 3 * b, c->T3);
    mtm(18, 18, 18, T, Kl, TK);
    mm(18, 18, 18, TK, T, K);
    mtm(18, 18, 1, T, fl, f);
    /* Gauss-point means of sigma_r and epsilon_r (Beam_1::WriteResults src/Beam_1.cpp:444-458) */
    if (res)
        for (int i = 0; i < 6; i++) {
            res[i] = (sg[0][i] + sg[1][i]) / 2;
            res[6 + i] = (ep[0][i] + ep[1][i]) / 2;
        }
}

void gfo_beam_save(const gfo_beam_const *c, double *gp, const double *d)
{
    for (int g = 0; g < 2; g++) {
        double *st = gp + g * GFO_BEAM_GPS;
        beam_kin k;
        beam_kinematics(c, st, d, g, &k);
        double Qn[9];
        m3_mul(k.Q, st, Qn);
        for (int i = 0; i < 9; i++) st[i] = Qn[i];
        for (int i = 0; i < 3; i++) {
            st[9 + i] = k.kap[i];
            st[12 + i] = k.dz[i];
            st[15 + i] += k.u[i];
        }
        double a3[3];
        gfo_compose(k.a, st + 18, a3);
        for (int i = 0; i < 3; i++) st[18 + i] = a3[i];
    }
}
