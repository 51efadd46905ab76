// Shell_1 on gfx950: element tangent + internal force fused with the
// conflict-free scatter, follower pressure, and the Gauss-point commit.
//
// Reference: Shell_1::Mount src/Shell_1.cpp:899-1332 (through-thickness
// hyperelastic integration :1064-1162, drilling stiffness :1212-1217,
// Psi :1239-1268, G :1277-1319), MountFieldLoads (self-weight, twice)
// :1335-1381, MountShellSpecialLoads :1392-1467, SaveLagrange :1650-1664,
// PreCalc :1666-2365, MountGlobal DOF order :1513-1648.
//
// Notation: the 15 rows of dN are five 3-row blocks R = 0 u_x1, 1 alpha_x1,
// 2 u_x2, 3 alpha_x2, 4 alpha; the 27 element DOFs are nine 3-column blocks
// I = 0..5 (u of nodes 1..6) and 6..8 (alpha of nodes 4..6).  dN block (R, I)
// is c_I^R times the identity, so K_IJ = sum_gp w sum_RS c_I^R c_J^S M_RS with
// M = Psi^T D Psi + G assembled block-wise (Psi has two non-zero blocks per
// block row: (R, R) and (R, 4)).
//
// Workgroup: 256 lanes, SEPB = 4 elements; LDS holds per Gauss point the
// 16 blocks of D, Psi, G, M and Psi^T sigma (48 KB).  In the Newton loop phase 1
// runs for all elements in one launch with every lane busy (shell_record_kernel),
// the follower pressure's terms likewise (shell_pressure_record_kernel), then one
// launch of phases 2-4 for all elements writes each element's K and f to HBM,
// and a light scatter launch per colour (shell_scatter_kernel) adds them, with the
// pressure terms and the self-weight, into the tangent and residual.
//   phase 1  lane per (element, gp): kinematics, 3-point thickness integration
//   phase 2  lane per (element, gp, block RS): the 25 blocks of M
//   phase 3  lane per (element, I, J): K_IJ, rotated to global
//   phase 4  lane per (element, I): internal force
#include "gf_internal.h"
#include "gf_device.h"

namespace {

constexpr int SEPB = 4;
constexpr int SWG = 256;
constexpr int L_D = 0;     // 16 blocks D(i,j), rows [n1 m1 n2 m2] x cols [eta1 k1 eta2 k2]
constexpr int L_P = 144;   // Psi(R,R), R = 0..3
constexpr int L_W = 180;   // Psi(R,4), R = 0..3
constexpr int L_G = 216;   // Gua1, Gua2, Vm1, Vm2, Gaa
constexpr int L_S = 261;   // Psi^T sigma (15)
constexpr int L_M = 276;   // M blocks (R,S), 25 x 9
constexpr int L_GP = 501;
constexpr int SPR = 57;    // pressure record of (element, u-node a): P_a (3), then K_p rows of a (3 x 18)

__device__ __forceinline__ m3<double> ld3(const double *p)
{
    m3<double> r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.a[i] = p[i];
    return r;
}
__device__ __forceinline__ void st3(double *p, const m3<double> &m)
{
#pragma unroll
    for (int i = 0; i < 9; i++) p[i] = m.a[i];
}
__device__ __forceinline__ void acc3(double *p, double s, const m3<double> &m)
{
#pragma unroll
    for (int i = 0; i < 9; i++) p[i] += s * m.a[i];
}
// X E3 and E3 X with E3 = skew(e3), e3 = (0,0,1)
__device__ __forceinline__ m3<double> mE3(const m3<double> &X)
{
    m3<double> r;
#pragma unroll
    for (int i = 0; i < 3; i++) { r.a[3 * i] = X.a[3 * i + 1]; r.a[3 * i + 1] = -X.a[3 * i]; r.a[3 * i + 2] = 0.0; }
    return r;
}
__device__ __forceinline__ m3<double> E3m(const m3<double> &X)
{
    m3<double> r;
#pragma unroll
    for (int j = 0; j < 3; j++) { r.a[j] = -X.a[3 + j]; r.a[3 + j] = X.a[j]; r.a[6 + j] = 0.0; }
    return r;
}

// Geometric block G(idx, 4) (row4 = false) or G(4, idx) (row4 = true), idx < 4:
// G(R,4) = Gua1, Vm1^T, Gua2, Vm2^T and G(4,S) = Gua1^T, Vm1, Gua2^T, Vm2, read
// from the Gua1 | Gua2 | Vm1 | Vm2 blocks at L_G with one indexed load per entry.
__device__ __forceinline__ m3<double> g_block(const double *o, int idx, bool row4)
{
    const int off = L_G + ((idx & 1) ? 18 + 9 * (idx >> 1) : 9 * (idx >> 1));
    const bool transpose = ((idx & 1) != 0) != row4;
    m3<double> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.a[3 * i + j] = o[off + (transpose ? 3 * j + i : 3 * i + j)];
    return r;
}

struct ShellKin {
    v3<double> u, u1, u2, a, a1, a2, z1, z2, eta1, eta2, k1, k2;
    m3<double> Q, Xi, Xi1, Xi2, QQ;
    double g;
};

// ZERO: the incremental displacements are taken as zero (disp and conn unused).
template <bool ZERO = false>
__device__ void shell_kin(const double *__restrict__ cst, const double *__restrict__ gps, size_t ld, int e, int gp,
                          const double *__restrict__ disp, const int *__restrict__ conn, ShellKin &k)
{
    double u[3] = {0, 0, 0}, u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
    for (int a = 0; a < (ZERO ? 0 : 6); a++) {
        const double *p = disp + 6 * (size_t)conn[a * ld + e];
        double c0 = cst[(SC_NU + 6 * gp + a) * ld + e], c1 = cst[(SC_NU1 + 6 * gp + a) * ld + e],
               c2 = cst[(SC_NU2 + 6 * gp + a) * ld + e];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double v = p[i];
            u[i] += v * c0;
            u1[i] += v * c1;
            u2[i] += v * c2;
        }
    }
    double al[3] = {0, 0, 0}, al1[3] = {0, 0, 0}, al2[3] = {0, 0, 0};
    for (int b = 0; b < (ZERO ? 0 : 3); b++) {
        const double *p = disp + 6 * (size_t)conn[(3 + b) * ld + e] + 3;
        double c0 = cst[(SC_NA + 3 * gp + b) * ld + e], c1 = cst[(SC_NA1 + 3 * gp + b) * ld + e],
               c2 = cst[(SC_NA2 + 3 * gp + b) * ld + e];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double v = p[i];
            al[i] += v * c0;
            al1[i] += v * c1;
            al2[i] += v * c2;
        }
    }
    m3<double> T3;
#pragma unroll
    for (int i = 0; i < 9; i++) T3.a[i] = cst[(SC_T3 + i) * ld + e];
    k.u = mv(T3, mk3(u[0], u[1], u[2]));
    k.u1 = mv(T3, mk3(u1[0], u1[1], u1[2]));
    k.u2 = mv(T3, mk3(u2[0], u2[1], u2[2]));
    k.a = mv(T3, mk3(al[0], al[1], al[2]));
    k.a1 = mv(T3, mk3(al1[0], al1[1], al1[2]));
    k.a2 = mv(T3, mk3(al2[0], al2[1], al2[2]));
    rodrigues(k.a, k.Q, k.Xi, k.g);
    m3<double> A1 = skew3(k.a1), A2 = skew3(k.a2);
    double p1 = dot3(k.a, k.a1), p2 = dot3(k.a, k.a2);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        k.Xi1.a[i] = -0.5 * k.g * (p1 * k.Xi.a[i] - A1.a[i]);
        k.Xi2.a[i] = -0.5 * k.g * (p2 * k.Xi.a[i] - A2.a[i]);
    }
    const size_t o = (size_t)SG_PER_GP * gp;
    m3<double> Qi;
#pragma unroll
    for (int i = 0; i < 9; i++) Qi.a[i] = gps[(o + i) * ld + e];
    v3<double> kr1 = mk3(gps[(o + 9) * ld + e], gps[(o + 10) * ld + e], gps[(o + 11) * ld + e]);
    v3<double> kr2 = mk3(gps[(o + 12) * ld + e], gps[(o + 13) * ld + e], gps[(o + 14) * ld + e]);
    v3<double> zx1 = mk3(gps[(o + 15) * ld + e], gps[(o + 16) * ld + e], gps[(o + 17) * ld + e]);
    v3<double> zx2 = mk3(gps[(o + 18) * ld + e], gps[(o + 19) * ld + e], gps[(o + 20) * ld + e]);
    k.z1 = add3(k.u1, zx1);
    k.z2 = add3(k.u2, zx2);
    k.QQ = mm3(k.Q, Qi);
    k.eta1 = mtv(k.QQ, k.z1);
    k.eta2 = mtv(k.QQ, k.z2);
    k.eta1.x -= 1.0;
    k.eta2.y -= 1.0;
    k.k1 = add3(mtv(Qi, mtv(k.Xi, k.a1)), kr1);
    k.k2 = add3(mtv(Qi, mtv(k.Xi, k.a2)), kr2);
}

// One point of the through-thickness rule (src/Shell_1.cpp:1083-1145): the
// strains gamma_a = eta_a + zeta kappa_a x e3 (e3 = (0,0,1)), v and dv of the
// hyperelastic law at jbar = det of the in-plane stretch, and the stress vectors
// tau_1, tau_2.
struct ShellPt {
    double g11, g12, g13, g21, g22, g23, v, dv;
    v3<double> tau1, tau2;
};
__device__ __forceinline__ ShellPt shell_point(const ShellKin &k, double mu, double la, double zeta)
{
    ShellPt p;
    p.g11 = k.eta1.x + zeta * k.k1.y;
    p.g12 = k.eta1.y - zeta * k.k1.x;
    p.g13 = k.eta1.z;
    p.g21 = k.eta2.x + zeta * k.k2.y;
    p.g22 = k.eta2.y - zeta * k.k2.x;
    p.g23 = k.eta2.z;
    const double jb = (1.0 + p.g11) * (1.0 + p.g22) - p.g12 * p.g21;
    p.v = (la * (jb * jb * jb - 1.0) + 2.0 * mu * (jb - 1.0)) / (la * jb * jb * jb + 2.0 * mu * jb);
    p.dv = ((la + 2.0 * mu) * (3.0 * la * jb * jb + 2.0 * mu)) /
           (jb * jb * (la * jb * jb + 2.0 * mu) * (la * jb * jb + 2.0 * mu));
    p.tau1 = mk3(mu * p.v * (1.0 + p.g22) + mu * (p.g11 - p.g22), mu * p.v * (-p.g21) + mu * (p.g12 + p.g21),
                 mu * p.g13);
    p.tau2 = mk3(mu * p.v * (-p.g12) + mu * (p.g12 + p.g21), mu * p.v * (1.0 + p.g11) + mu * (p.g22 - p.g11),
                 mu * p.g23);
    return p;
}

// Column block I of dN: rows (R) and their coefficients at Gauss point gp.
__device__ __forceinline__ int col_rows(const double *__restrict__ cst, size_t ld, int e, int gp, int I, int *R,
                                         double *c)
{
    if (I < 6) {
        R[0] = 0; c[0] = cst[(SC_NU1 + 6 * gp + I) * ld + e];
        R[1] = 2; c[1] = cst[(SC_NU2 + 6 * gp + I) * ld + e];
        return 2;
    }
    int b = I - 6;
    R[0] = 1; c[0] = cst[(SC_NA1 + 3 * gp + b) * ld + e];
    R[1] = 3; c[1] = cst[(SC_NA2 + 3 * gp + b) * ld + e];
    R[2] = 4; c[2] = cst[(SC_NA + 3 * gp + b) * ld + e];
    return 3;
}

// DOF block I of an element -> (node, first component)
__device__ __forceinline__ void dof_block(const int *__restrict__ conn, size_t ld, int e, int I, int &node, int &c0)
{
    if (I < 6) { node = conn[I * ld + e]; c0 = 0; }
    else { node = conn[(3 + I - 6) * ld + e]; c0 = 3; }
}

// Phase 1 of Shell_1::Mount for one (element, Gauss point): kinematics, the
// 3-point thickness integration, and the record o[0 .. L_M) (D, Psi, G blocks
// and Psi^T sigma) that the later phases read.  The D blocks, accumulated over
// the thickness points, go to oD[0 .. 144) (o + L_D, or scratch of their own
// that the caller copies into the record); the rest is written once, into o.
__device__ __forceinline__ void shell_gp_record(const double *__restrict__ cst, const double *__restrict__ gps,
                                                size_t ld, int e, int gp, const double *__restrict__ disp,
                                                const int *__restrict__ conn, double *oD, double *o)
{
    ShellKin k;
    shell_kin(cst, gps, ld, e, gp, disp, conn, k);
    const double mu = cst[SC_MU * ld + e], la = cst[SC_LAMBDA * ld + e], th = cst[SC_THICK * ld + e];
    const double drill = cst[SC_DRILL * ld + e];
    for (int i = 0; i < 144; i++) oD[i] = 0.0;
    v3<double> n1 = mk3(0.0, 0.0, 0.0), n2 = n1, m1 = n1, m2 = n1;
    const double xi[3] = {-0.77459666924148337703585307995648, 0.0, 0.77459666924148337703585307995648};
    const double wt[3] = {0.55555555555555555555555555555556, 0.88888888888888888888888888888889,
                          0.55555555555555555555555555555556};
    const double J = th / 2.0;
    for (int q = 0; q < 3; q++) {
        double zeta = th * xi[q] / 2.0;
        const ShellPt pt = shell_point(k, mu, la, zeta);
        const double g11 = pt.g11, g12 = pt.g12, g21 = pt.g21, g22 = pt.g22, v = pt.v, dv = pt.dv;
        const v3<double> tau1 = pt.tau1, tau2 = pt.tau2;
        m3<double> C11, C22, C12, C21;
#pragma unroll
        for (int i = 0; i < 9; i++) C11.a[i] = C22.a[i] = C12.a[i] = 0.0;
        C11.a[0] = mu * ((1.0 + g22) * (1.0 + g22) * dv + 1.0);
        C11.a[1] = -mu * (1.0 + g22) * g21 * dv;
        C11.a[3] = -mu * (1.0 + g22) * g21 * dv;
        C11.a[4] = mu * (g21 * g21 * dv + 1.0);
        C11.a[8] = mu;
        C22.a[0] = mu * (g12 * g12 * dv + 1.0);
        C22.a[1] = -mu * ((1.0 + g11) * g12 * dv);
        C22.a[3] = -mu * ((1.0 + g11) * g12 * dv);
        C22.a[4] = mu * ((1.0 + g11) * (1.0 + g11) * dv + 1.0);
        C22.a[8] = mu;
        C12.a[0] = -mu * ((1.0 + g22) * g12 * dv);
        C12.a[1] = mu * (v - 1.0 + (1.0 + g11) * (1.0 + g22) * dv);
        C12.a[3] = mu * (1.0 - v + g12 * g21 * dv);
        C12.a[4] = -mu * ((1.0 + g11) * g21 * dv);
        C21 = tr3(C12);
        double w = wt[q] * J;
        acc3(oD + 9 * 0, w, C11);            // dn1/deta1
        acc3(oD + 9 * 1, -w * zeta, mE3(C11)); // dn1/dk1
        acc3(oD + 9 * 2, w, C12);            // dn1/deta2
        acc3(oD + 9 * 3, -w * zeta, mE3(C12)); // dn1/dk2
        acc3(oD + 9 * 4, w * zeta, E3m(C11));  // dm1/deta1
        acc3(oD + 9 * 5, -w * zeta * zeta, mE3(E3m(C11)));
        acc3(oD + 9 * 6, w * zeta, E3m(C12));
        acc3(oD + 9 * 7, -w * zeta * zeta, mE3(E3m(C12)));
        acc3(oD + 9 * 8, w, C21);            // dn2/deta1
        acc3(oD + 9 * 9, -w * zeta, mE3(C21));
        acc3(oD + 9 * 10, w, C22);
        acc3(oD + 9 * 11, -w * zeta, mE3(C22));
        acc3(oD + 9 * 12, w * zeta, E3m(C21)); // dm2/deta1
        acc3(oD + 9 * 13, -w * zeta * zeta, mE3(E3m(C21)));
        acc3(oD + 9 * 14, w * zeta, E3m(C22));
        acc3(oD + 9 * 15, -w * zeta * zeta, mE3(E3m(C22)));
        n1 = add3(n1, scl3(w, tau1));
        n2 = add3(n2, scl3(w, tau2));
        // e3 x tau = (-tau.y, tau.x, 0)
        m1 = add3(m1, scl3(w * zeta, mk3(-tau1.y, tau1.x, 0.0)));
        m2 = add3(m2, scl3(w * zeta, mk3(-tau2.y, tau2.x, 0.0)));
    }
    // drilling (overwrites, as the reference)
    oD[9 * 5 + 8] = drill;   // D(5,5): block (1,1) entry (2,2)
    oD[9 * 15 + 8] = drill;  // D(11,11): block (3,3) entry (2,2)
    m1.z = drill * k.k1.z;
    m2.z = drill * k.k2.z;
    m3<double> Qt = tr3(k.QQ);
    m3<double> QtXi = mm3(Qt, k.Xi);
    m3<double> Z1 = skew3(k.z1), Z2 = skew3(k.z2);
    m3<double> W0 = mm3(mm3(Qt, Z1), k.Xi), W1 = mm3(Qt, k.Xi1);
    m3<double> W2 = mm3(mm3(Qt, Z2), k.Xi), W3 = mm3(Qt, k.Xi2);
    st3(o + L_P + 0, Qt);
    st3(o + L_P + 9, QtXi);
    st3(o + L_P + 18, Qt);
    st3(o + L_P + 27, QtXi);
    st3(o + L_W + 0, W0);
    st3(o + L_W + 9, W1);
    st3(o + L_W + 18, W2);
    st3(o + L_W + 27, W3);
    // Psi^T sigma
    v3<double> s0 = mtv(Qt, n1), s1 = mtv(QtXi, m1), s2 = mtv(Qt, n2), s3 = mtv(QtXi, m2);
    v3<double> s4 = add3(add3(mtv(W0, n1), mtv(W1, m1)), add3(mtv(W2, n2), mtv(W3, m2)));
    double *S = o + L_S;
    S[0] = s0.x; S[1] = s0.y; S[2] = s0.z; S[3] = s1.x; S[4] = s1.y; S[5] = s1.z;
    S[6] = s2.x; S[7] = s2.y; S[8] = s2.z; S[9] = s3.x; S[10] = s3.y; S[11] = s3.z;
    S[12] = s4.x; S[13] = s4.y; S[14] = s4.z;
    // geometric blocks
    v3<double> n_1 = mv(k.QQ, n1), n_2 = mv(k.QQ, n2), m_1 = mv(k.QQ, m1), m_2 = mv(k.QQ, m2);
    m3<double> Gua1 = mscl(-1.0, mm3(skew3(n_1), k.Xi));
    m3<double> Gua2 = mscl(-1.0, mm3(skew3(n_2), k.Xi));
    m3<double> Vm1 = opV(k.a, m_1), Vm2 = opV(k.a, m_2);
    m3<double> XiT = tr3(k.Xi);
    m3<double> Ga1 = msub(mm3(XiT, mm3(mm3(Z1, skew3(n_1)), k.Xi)), opV(k.a, mv(Z1, n_1)));
    Ga1 = madd(Ga1, msub(opdV(k.a, k.a1, m_1), mtm3(k.Xi1, mm3(skew3(m_1), k.Xi))));
    m3<double> Ga2 = msub(mm3(XiT, mm3(mm3(Z2, skew3(n_2)), k.Xi)), opV(k.a, mv(Z2, n_2)));
    Ga2 = madd(Ga2, msub(opdV(k.a, k.a2, m_2), mtm3(k.Xi2, mm3(skew3(m_2), k.Xi))));
    st3(o + L_G + 0, Gua1);
    st3(o + L_G + 9, Gua2);
    st3(o + L_G + 18, Vm1);
    st3(o + L_G + 27, Vm2);
    st3(o + L_G + 36, madd(Ga1, Ga2));
}

template <bool PRE>
__global__ __launch_bounds__(SWG, PRE ? 3 : 1) void shell_element_kernel(
    int e0, int e1, size_t ld, const int *__restrict__ conn, const double *__restrict__ cst,
    const double *__restrict__ gps, const double *__restrict__ disp, double *__restrict__ Kout,
    double *__restrict__ fout, const double *__restrict__ pre)
{
    __shared__ double sm[SEPB * 3 * L_GP];
    const int tid = threadIdx.x;
    const int ebase = e0 + blockIdx.x * SEPB;

    // ---- phase 1 (or its records from shell_record_kernel) ----------------------
    if constexpr (PRE) {
        const int nv = min(SEPB, e1 - ebase);
        const double *src = pre + (size_t)ebase * 3 * L_M;
        for (int t = tid; t < nv * 3 * L_M; t += SWG) sm[(t / L_M) * L_GP + t % L_M] = src[t];
    } else if (tid < 3 * SEPB) {
        int le = tid / 3, gp = tid - 3 * (tid / 3);
        int e = ebase + le;
        double *o = sm + (le * 3 + gp) * L_GP;
        if (e < e1) shell_gp_record(cst, gps, ld, e, gp, disp, conn, o + L_D, o);
    }
    __syncthreads();

    // ---- phase 2: M blocks ---------------------------------------------------
    for (int t = tid; t < SEPB * 3 * 25; t += SWG) {
        int le = t / 75, rem = t - 75 * le, gp = rem / 25, rs = rem - 25 * gp;
        int R = rs / 5, S = rs - 5 * R;
        int e = ebase + le;
        if (e >= e1) continue;
        double *o = sm + (le * 3 + gp) * L_GP;
        m3<double> Mb;
        if (R < 4 && S < 4) {
            Mb = mtm3(ld3(o + L_P + 9 * R), mm3(ld3(o + L_D + 9 * (4 * R + S)), ld3(o + L_P + 9 * S)));
        } else if (R < 4) {
            m3<double> V = mm3(ld3(o + L_D + 9 * (4 * R)), ld3(o + L_W));
#pragma unroll
            for (int j = 1; j < 4; j++) V = madd(V, mm3(ld3(o + L_D + 9 * (4 * R + j)), ld3(o + L_W + 9 * j)));
            Mb = mtm3(ld3(o + L_P + 9 * R), V);
            Mb = madd(Mb, g_block(o, R, false));
        } else if (S < 4) {
            m3<double> V = mtm3(ld3(o + L_W), ld3(o + L_D + 9 * S));
#pragma unroll
            for (int i = 1; i < 4; i++) V = madd(V, mtm3(ld3(o + L_W + 9 * i), ld3(o + L_D + 9 * (4 * i + S))));
            Mb = mm3(V, ld3(o + L_P + 9 * S));
            Mb = madd(Mb, g_block(o, S, true));
        } else {
            m3<double> acc;
#pragma unroll
            for (int q = 0; q < 9; q++) acc.a[q] = 0.0;
            for (int i = 0; i < 4; i++) {
                m3<double> V = mm3(ld3(o + L_D + 9 * (4 * i)), ld3(o + L_W));
                for (int j = 1; j < 4; j++) V = madd(V, mm3(ld3(o + L_D + 9 * (4 * i + j)), ld3(o + L_W + 9 * j)));
                acc = madd(acc, mtm3(ld3(o + L_W + 9 * i), V));
            }
            Mb = madd(acc, ld3(o + L_G + 36));
        }
        st3(o + L_M + 9 * rs, Mb);
    }
    __syncthreads();

    // ---- phase 3: K_IJ -------------------------------------------------------
    for (int t = tid; t < SEPB * 81; t += SWG) {
        int le = t / 81, ij = t - 81 * le;
        int I = ij / 9, J = ij - 9 * (ij / 9);
        int e = ebase + le;
        if (e >= e1) continue;
        double w = cst[SC_ALPHA1 * ld + e];
        m3<double> K;
#pragma unroll
        for (int q = 0; q < 9; q++) K.a[q] = 0.0;
        for (int gp = 0; gp < 3; gp++) {
            const double *o = sm + (le * 3 + gp) * L_GP;
            int RI[3], RJ[3];
            double cI[3], cJ[3];
            int nI = col_rows(cst, ld, e, gp, I, RI, cI);
            int nJ = col_rows(cst, ld, e, gp, J, RJ, cJ);
            for (int a = 0; a < nI; a++)
                for (int b = 0; b < nJ; b++) acc3(K.a, w * cI[a] * cJ[b], ld3(o + L_M + 9 * (5 * RI[a] + RJ[b])));
        }
        m3<double> T3;
#pragma unroll
        for (int q = 0; q < 9; q++) T3.a[q] = cst[(SC_T3 + q) * ld + e];
        K = mtm3(T3, mm3(K, T3));
        double *Ke = Kout + (size_t)e * 729;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) Ke[(3 * I + i) * 27 + 3 * J + j] = K.a[3 * i + j];
    }

    // ---- phase 4: internal force ---------------------------------------------
    for (int t = tid; t < SEPB * 9; t += SWG) {
        int le = t / 9, I = t - 9 * le;
        int e = ebase + le;
        if (e >= e1) continue;
        double w = cst[SC_ALPHA1 * ld + e];
        v3<double> f = mk3(0.0, 0.0, 0.0);
        for (int gp = 0; gp < 3; gp++) {
            const double *o = sm + (le * 3 + gp) * L_GP;
            int RI[3];
            double cI[3];
            int nI = col_rows(cst, ld, e, gp, I, RI, cI);
            for (int a = 0; a < nI; a++) {
                const double *S = o + L_S + 3 * RI[a];
                f = add3(f, scl3(w * cI[a], mk3(S[0], S[1], S[2])));
            }
        }
        m3<double> T3;
#pragma unroll
        for (int q = 0; q < 9; q++) T3.a[q] = cst[(SC_T3 + q) * ld + e];
        f = mtv(T3, f);
        fout[(size_t)e * 27 + 3 * I + 0] = f.x;
        fout[(size_t)e * 27 + 3 * I + 1] = f.y;
        fout[(size_t)e * 27 + 3 * I + 2] = f.z;
    }
}

// Follower pressure (MountShellSpecialLoads), lane per (element, u-node a) over all
// elements: the record out[(6 e + a) SPR ...] holds the unbalanced-force change P_a
// (3) and the load stiffness rows of node a (3 x 18, global frame); zeros for an
// element without pressure.  The colour launches of shell_scatter_kernel add them
// to the element's own scatter.
__global__ __launch_bounds__(256) void shell_pressure_record_kernel(
    int ns, size_t ld, const int *__restrict__ conn, const double *__restrict__ cst,
    const double *__restrict__ pc, const double *__restrict__ disp, const double *__restrict__ copy,
    const double *__restrict__ copy0, const int *__restrict__ pload, const int *__restrict__ pflag,
    const double *__restrict__ lrec, const double *__restrict__ tab_t, const double *__restrict__ tab_v, const double *__restrict__ tp,
    double *__restrict__ out)
{
    const double time = *tp;  // load time, device-resident (a captured Newton step reads the current one)
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int e = t / 6, a = t - 6 * (t / 6);
    if (e >= ns) return;
    double *o = out + (size_t)t * SPR;
    int L = pload[e];
    if (L < 0) {
        for (int q = 0; q < SPR; q++) o[q] = 0.0;
        return;
    }
    const double *rec = lrec + (size_t)L * NL_REC;
    double p = gf_table_value(tab_t, tab_v, (int)rec[NL_TAB0], (int)rec[NL_TABN], time, 0);
    int au = pflag[e];
    double pu[18];
    for (int b = 0; b < 6; b++) {
        int nd = conn[b * ld + e];
#pragma unroll
        for (int k = 0; k < 3; k++) pu[3 * b + k] = (copy[6 * (size_t)nd + k] - copy0[6 * (size_t)nd + k]) + disp[6 * (size_t)nd + k];
    }
    v3<double> e1r = mk3(cst[(SC_T3 + 0) * ld + e], cst[(SC_T3 + 1) * ld + e], cst[(SC_T3 + 2) * ld + e]);
    v3<double> e2r = mk3(cst[(SC_T3 + 3) * ld + e], cst[(SC_T3 + 4) * ld + e], cst[(SC_T3 + 5) * ld + e]);
    double Kr[3][18];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 18; j++) Kr[i][j] = 0.0;
    v3<double> Pa = mk3(0.0, 0.0, 0.0);
    for (int g = 0; g < 6; g++) {
        v3<double> t1 = e1r, t2 = e2r;
        for (int b = 0; b < 6; b++) {
            double c1 = pc[(SP_N4U1 + 6 * g + b) * ld + e], c2 = pc[(SP_N4U2 + 6 * g + b) * ld + e];
            t1 = add3(t1, scl3(c1, mk3(pu[3 * b], pu[3 * b + 1], pu[3 * b + 2])));
            t2 = add3(t2, scl3(c2, mk3(pu[3 * b], pu[3 * b + 1], pu[3 * b + 2])));
        }
        v3<double> cr = cross3(t1, t2);
        double ncr = sqrt(dot3(cr, cr));
        v3<double> n = scl3(1.0 / ncr, cr);
        double w4 = pc[(SP_W4 + g) * ld + e];
        double Na = pc[(SP_N4U + 6 * g + a) * ld + e];
        double jac = au ? ncr : 1.0;
        // unbalanced force change: P -= w4 jac N4^T q with q = -p n
        Pa = add3(Pa, scl3(w4 * jac * Na * p, n));
        m3<double> S1 = skew3(t1), S2 = skew3(t2);
        m3<double> Pm;
        if (!au) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    double ni = i == 0 ? n.x : (i == 1 ? n.y : n.z), nj = j == 0 ? n.x : (j == 1 ? n.y : n.z);
                    Pm.a[3 * i + j] = ((i == j ? 1.0 : 0.0) - ni * nj) / ncr;
                }
        }
        double s = w4 * 1.0 * p * Na;
        for (int b = 0; b < 6; b++) {
            double c1 = pc[(SP_N4U1 + 6 * g + b) * ld + e], c2 = pc[(SP_N4U2 + 6 * g + b) * ld + e];
            m3<double> B = msub(mscl(c2, S1), mscl(c1, S2));
            if (!au) B = mm3(Pm, B);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) Kr[i][3 * b + j] += s * B.a[3 * i + j];
        }
    }
    // the element's unbalanced force grows by Pa (the residual takes its negative)
    o[0] = Pa.x;
    o[1] = Pa.y;
    o[2] = Pa.z;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 18; j++) o[3 + 18 * i + j] = Kr[i][j];
}

// Phase 1 of every element's Gauss points in one launch: lane per (element, gp),
// its D blocks accumulated in LDS (74 KB per 64 lanes, two workgroups per CU;
// a lane's stride of 145 entries keeps a wavefront's accesses off shared banks)
// and copied out record by record, the write-once rest stored directly.  The
// element kernel reads the records instead of recomputing phase 1 on 12 lanes
// per workgroup.
constexpr int REC_LANES = 64, REC_DS = 145;
__global__ __launch_bounds__(REC_LANES) void shell_record_kernel(int ns, size_t ld, const int *__restrict__ conn,
                                                                  const double *__restrict__ cst,
                                                                  const double *__restrict__ gps,
                                                                  const double *__restrict__ disp, double *__restrict__ rec)
{
    __shared__ double sD[REC_LANES * REC_DS];
    const int t0 = blockIdx.x * REC_LANES, t = t0 + threadIdx.x;
    const int nt = min(REC_LANES, 3 * ns - t0);
    if ((int)threadIdx.x < nt) {
        const int e = t / 3, gp = t - 3 * (t / 3);
        shell_gp_record(cst, gps, ld, e, gp, disp, conn, sD + threadIdx.x * REC_DS, rec + (size_t)t * L_M);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nt * 144; q += REC_LANES) {
        const int r = q / 144, i = q - 144 * r;
        rec[(size_t)(t0 + r) * L_M + L_D + i] = sD[r * REC_DS + i];
    }
}

__global__ __launch_bounds__(256) void shell_commit_kernel(int ns, size_t ld, const int *__restrict__ conn,
                                                           const double *__restrict__ cst, double *__restrict__ gps,
                                                           const double *__restrict__ disp)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * ns) return;
    int e = t / 3, gp = t - 3 * (t / 3);
    ShellKin k;
    shell_kin(cst, gps, ld, e, gp, disp, conn, k);
    const size_t o = (size_t)SG_PER_GP * gp;
    m3<double> Qi;
#pragma unroll
    for (int i = 0; i < 9; i++) Qi.a[i] = gps[(o + i) * ld + e];
    m3<double> Qn = mm3(k.Q, Qi);
#pragma unroll
    for (int i = 0; i < 9; i++) gps[(o + i) * ld + e] = Qn.a[i];
    gps[(o + 9) * ld + e] = k.k1.x;
    gps[(o + 10) * ld + e] = k.k1.y;
    gps[(o + 11) * ld + e] = k.k1.z;
    gps[(o + 12) * ld + e] = k.k2.x;
    gps[(o + 13) * ld + e] = k.k2.y;
    gps[(o + 14) * ld + e] = k.k2.z;
    gps[(o + 15) * ld + e] += k.u1.x;
    gps[(o + 16) * ld + e] += k.u1.y;
    gps[(o + 17) * ld + e] += k.u1.z;
    gps[(o + 18) * ld + e] += k.u2.x;
    gps[(o + 19) * ld + e] += k.u2.y;
    gps[(o + 20) * ld + e] += k.u2.z;
}

// Results of the committed state (Shell_1::WriteResults src/Shell_1.cpp:624-666):
// per element the means over the three Gauss points of n_r1, m_r1, n_r2, m_r2,
// eta_r1, kappa_r1, eta_r2, kappa_r2.  With zero incremental displacements the
// kinematics reduce to the committed Q_i, kappa_i, z_x1, z_x2, which hold the
// converged configuration of the last Mount.  out: [ns][24], colour-sorted.
__global__ __launch_bounds__(256) void shell_results_kernel(int ns, size_t ld, const double *__restrict__ cst,
                                                            const double *__restrict__ gps, double *__restrict__ out)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ns) return;
    const double mu = cst[SC_MU * ld + e], la = cst[SC_LAMBDA * ld + e], th = cst[SC_THICK * ld + e];
    const double drill = cst[SC_DRILL * ld + e];
    const double xi[3] = {-0.77459666924148337703585307995648, 0.0, 0.77459666924148337703585307995648};
    const double wt[3] = {0.55555555555555555555555555555556, 0.88888888888888888888888888888889,
                          0.55555555555555555555555555555556};
    const double J = th / 2.0;
    v3<double> avg[8];
#pragma unroll
    for (int i = 0; i < 8; i++) avg[i] = mk3(0.0, 0.0, 0.0);
    for (int gp = 0; gp < 3; gp++) {
        ShellKin k;
        shell_kin<true>(cst, gps, ld, e, gp, nullptr, nullptr, k);
        v3<double> n1 = mk3(0.0, 0.0, 0.0), n2 = n1, m1 = n1, m2 = n1;
        for (int q = 0; q < 3; q++) {
            const double zeta = th * xi[q] / 2.0;
            const ShellPt pt = shell_point(k, mu, la, zeta);
            const double w = wt[q] * J;
            n1 = add3(n1, scl3(w, pt.tau1));
            n2 = add3(n2, scl3(w, pt.tau2));
            m1 = add3(m1, scl3(w * zeta, mk3(-pt.tau1.y, pt.tau1.x, 0.0)));
            m2 = add3(m2, scl3(w * zeta, mk3(-pt.tau2.y, pt.tau2.x, 0.0)));
        }
        m1.z = drill * k.k1.z;
        m2.z = drill * k.k2.z;
        const v3<double> r[8] = {n1, m1, n2, m2, k.eta1, k.k1, k.eta2, k.k2};
#pragma unroll
        for (int i = 0; i < 8; i++) avg[i] = add3(avg[i], scl3(1.0 / 3.0, r[i]));
    }
    double *o = out + 24 * (size_t)e;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        o[3 * i] = avg[i].x;
        o[3 * i + 1] = avg[i].y;
        o[3 * i + 2] = avg[i].z;
    }
}

}  // namespace

// ---- host: constants (Shell_1::PreCalc restated for the device layout) ---------
static void vsub(const double *a, const double *b, double *c) { for (int i = 0; i < 3; i++) c[i] = a[i] - b[i]; }
static void vcross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double vdot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double vnorm(const double *a) { return std::sqrt(vdot(a, a)); }

static void tri_shapes(double L1, double L2, double L3, const double *dx, const double *dy, double *Nu, double *Nu1,
                       double *Nu2, double *Na, double *Na1, double *Na2)
{
    double L1x = dx[0], L2x = dx[1], L3x = dx[2], L1y = dy[0], L2y = dy[1], L3y = dy[2];
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

void gf_shell_host_const(const double *x, const double *prop, double *rec, double *prec, double *gp)
{
    for (int i = 0; i < SC_COUNT; i++) rec[i] = 0.0;
    for (int i = 0; i < SP_COUNT; i++) prec[i] = 0.0;
    const double *x1 = x, *x2 = x + 3, *x3 = x + 6;
    double d21[3], d31[3], nrm[3];
    vsub(x2, x1, d21);
    vsub(x3, x1, d31);
    vcross(d21, d31, nrm);
    double nn = vnorm(nrm);
    double A = 0.5 * nn;
    double e3[3] = {nrm[0] / nn, nrm[1] / nn, nrm[2] / nn};
    double eg[3] = {1.0, 0.0, 0.0};
    if (std::fabs(vdot(eg, e3)) >= 1.0 - 1e-4) { eg[0] = 0.0; eg[1] = 1.0; }
    double pr = vdot(eg, e3), e1[3], e2[3];
    for (int i = 0; i < 3; i++) e1[i] = eg[i] - pr * e3[i];
    double n1 = vnorm(e1);
    for (int i = 0; i < 3; i++) e1[i] = (1.0 / n1) * e1[i];
    vcross(e3, e1, e2);
    for (int i = 0; i < 3; i++) { rec[SC_T3 + i] = e1[i]; rec[SC_T3 + 3 + i] = e2[i]; rec[SC_T3 + 6 + i] = e3[i]; }
    double E = prop[0], nu = prop[1], rho = prop[2], t = prop[3];
    double mu = E / (2.0 * (1 + nu));
    rec[SC_MU] = mu;
    rec[SC_LAMBDA] = 2.0 * mu * nu / (1 - 2.0 * nu);
    rec[SC_THICK] = t;
    rec[SC_DRILL] = E * t * t * t;
    rec[SC_RHO] = rho;
    rec[SC_ALPHA1] = A / 3.0;
    double t23[3], t31[3], t12[3], t32[3], t13[3], t21[3];
    vsub(x2, x3, t23); vsub(x3, x1, t31); vsub(x1, x2, t12);
    vsub(x3, x2, t32); vsub(x1, x3, t13); vsub(x2, x1, t21);
    double dx[3] = {0.5 * vdot(t23, e2) / A, 0.5 * vdot(t31, e2) / A, 0.5 * vdot(t12, e2) / A};
    double dy[3] = {0.5 * vdot(t32, e1) / A, 0.5 * vdot(t13, e1) / A, 0.5 * vdot(t21, e1) / A};
    for (int g = 0; g < 3; g++) {
        const double *xp = x + 3 * (3 + g);
        double a2[3], a3[3], cr[3], L[3];
        vsub(x2, xp, a2); vsub(x3, xp, a3); vcross(a2, a3, cr); L[0] = 0.5 * vnorm(cr) / A;
        vsub(x3, xp, a2); vsub(x1, xp, a3); vcross(a2, a3, cr); L[1] = 0.5 * vnorm(cr) / A;
        vsub(x1, xp, a2); vsub(x2, xp, a3); vcross(a2, a3, cr); L[2] = 0.5 * vnorm(cr) / A;
        tri_shapes(L[0], L[1], L[2], dx, dy, rec + SC_NU + 6 * g, rec + SC_NU1 + 6 * g, rec + SC_NU2 + 6 * g,
                   rec + SC_NA + 3 * g, rec + SC_NA1 + 3 * g, rec + SC_NA2 + 3 * g);
    }
    // Cowper (1973) six-point rule, constants as src/Shell_1.cpp:2190-2233
    static const double cw[6][4] = {
        {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.109951743655322},
        {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.109951743655322},
        {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.109951743655322},
        {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.223381589678011},
        {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.223381589678011},
        {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.223381589678011}};
    for (int g = 0; g < 6; g++) {
        prec[SP_W4 + g] = A * cw[g][3];
        tri_shapes(cw[g][0], cw[g][1], cw[g][2], dx, dy, prec + SP_N4U + 6 * g, prec + SP_N4U1 + 6 * g,
                   prec + SP_N4U2 + 6 * g, nullptr, nullptr, nullptr);
    }
    for (int a = 0; a < 6; a++) {
        double s = 0.0;
        for (int g = 0; g < 6; g++) s += prec[SP_W4 + g] * prec[SP_N4U + 6 * g + a];
        rec[SC_GW + a] = s;
    }
    for (int g = 0; g < 3; g++) {
        double *st = gp + SG_PER_GP * g;
        for (int i = 0; i < SG_PER_GP; i++) st[i] = 0.0;
        st[0] = st[4] = st[8] = 1.0;
        st[15] = 1.0;  // z_x1_i = e1 (local)
        st[19] = 1.0;  // z_x2_i = e2 (local)
    }
}

// Scatter of one colour's element matrices and forces (written for all elements by
// shell_element_kernel): lane per (element, I, J) adding K_IJ, plus the follower
// pressure stiffness of u-node pair (I, J), into the tangent; lane per (element, I)
// adding -(f_I - self-weight + P_I) into r.  Elements of a colour share no node, and
// an element's lanes touch distinct entries, so no two lanes write the same place.
__global__ __launch_bounds__(256) void shell_scatter_kernel(int e0, int e1, size_t ld, const int *__restrict__ conn,
                                                            const double *__restrict__ cst, const int *__restrict__ dof,
                                                            gf_gmat gm, double *__restrict__ r,
                                                            const double *__restrict__ Ke, const double *__restrict__ fe,
                                                            const double *__restrict__ ppre, const double *__restrict__ lfp,
                                                            double gx, double gy, double gz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ne = e1 - e0;
    if (t < 81 * ne) {
        const int e = e0 + t / 81, ij = t - 81 * (t / 81);
        const int I = ij / 9, J = ij - 9 * (ij / 9);
        const double *k = Ke + (size_t)e * 729 + (3 * I) * 27 + 3 * J;
        double V9[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) V9[3 * i + j] = k[27 * i + j];
        if (ppre && I < 6 && J < 6) {
            const double *kp = ppre + ((size_t)e * 6 + I) * SPR + 3 + 3 * J;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) V9[3 * i + j] += kp[18 * i + j];
        }
        int na, ca, nb_, cb;
        dof_block(conn, ld, e, I, na, ca);
        dof_block(conn, ld, e, J, nb_, cb);
        int R9[9], C9[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                R9[3 * i + j] = dof[6 * na + ca + i];
                C9[3 * i + j] = dof[6 * nb_ + cb + j];
            }
        gm_add_n<9>(gm, R9, C9, V9);
        return;
    }
    const int u = t - 81 * ne;
    if (u >= 9 * ne) return;
    const int e = e0 + u / 9, I = u - 9 * (u / 9);
    const double lf = lfp ? *lfp : 0.0;
    const double *fp = fe + (size_t)e * 27 + 3 * I;
    v3<double> f = mk3(fp[0], fp[1], fp[2]);
    if (I < 6 && lf != 0.0) {
        // self-weight, accumulated twice as src/Shell_1.cpp:1340-1375 does
        double mult = lf * cst[SC_RHO * ld + e] * cst[SC_THICK * ld + e];
        double gw = cst[(SC_GW + I) * ld + e];
        for (int rep = 0; rep < 2; rep++) {
            f.x -= gw * mult * gx;
            f.y -= gw * mult * gy;
            f.z -= gw * mult * gz;
        }
    }
    int node, c0;
    dof_block(conn, ld, e, I, node, c0);
    int R3[3];
    double V3[3] = {-f.x, -f.y, -f.z};
    if (ppre && I < 6) {
        const double *pa = ppre + ((size_t)e * 6 + I) * SPR;
#pragma unroll
        for (int i = 0; i < 3; i++) V3[i] -= pa[i];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) R3[i] = dof[6 * node + c0 + i];
    vec_add_n<3>(r, R3, V3);
}

// ---- launchers --------------------------------------------------------------
int gf_shell_assemble(gf_model *m, const double *lfp, const double *tp)
{
    const size_t ld = (size_t)m->ns;
    int ncol = (int)m->shell_color_off.size() - 1;
    int launches = 0;
    if (m->ns == 0) return 0;
    // phase 1 of every element and the pressure terms, then one scatter launch per colour
    hipLaunchKernelGGL(shell_record_kernel, dim3((3 * m->ns + REC_LANES - 1) / REC_LANES), dim3(REC_LANES), 0,
                       m->stream, m->ns, ld, m->d_sconn, m->d_sconst, m->d_sgp, m->d_disp, m->d_shell_rec);
    launches++;
    if (m->any_pressure) {
        hipLaunchKernelGGL(shell_pressure_record_kernel, dim3((6 * m->ns + 255) / 256), dim3(256), 0, m->stream,
                           m->ns, ld, m->d_sconn, m->d_sconst, m->d_spconst, m->d_disp, m->d_copy, m->d_copy0,
                           m->d_spload, m->d_spflag, m->d_lrec, m->d_tab_t, m->d_tab_v, tp, m->d_shell_prec);
        launches++;
    }
    // every element's K and f (phases 2-4 from the records), then one scatter launch per colour
    {
        const int pr = gf_probe_begin(m->stream, GF_KC_SHELL_ASSEMBLE);
        hipLaunchKernelGGL(shell_element_kernel<true>, dim3((m->ns + SEPB - 1) / SEPB), dim3(SWG), 0, m->stream, 0,
                           m->ns, ld, m->d_sconn, m->d_sconst, m->d_sgp, m->d_disp, m->d_shell_K, m->d_shell_f,
                           (const double *)m->d_shell_rec);
        gf_probe_end(m->stream, pr, m->ns);
        launches++;
    }
    const double *ppre = m->any_pressure ? m->d_shell_prec : nullptr;
    for (int c = 0; c < ncol; c++) {
        int e0 = m->shell_color_off[c], e1 = m->shell_color_off[c + 1];
        if (e1 <= e0) continue;
        launches++;
        const int threads = 90 * (e1 - e0);
        hipLaunchKernelGGL(shell_scatter_kernel, dim3((threads + 255) / 256), dim3(256), 0, m->stream, e0, e1, ld,
                           m->d_sconn, m->d_sconst, m->d_dof, gf_model_gmat(m), m->d_r, (const double *)m->d_shell_K,
                           (const double *)m->d_shell_f, ppre, lfp, m->gravity[0], m->gravity[1], m->gravity[2]);
    }
    GF_HIP(hipGetLastError());
    return launches;
}

int gf_shell_commit(gf_model *m)
{
    if (m->ns == 0) return GF_OK;
    hipLaunchKernelGGL(shell_commit_kernel, dim3((3 * m->ns + 255) / 256), dim3(256), 0, m->stream, m->ns,
                       (size_t)m->ns, m->d_sconn, m->d_sconst, m->d_sgp, m->d_disp);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_shell_results(gf_model *m, double *d_out)
{
    if (m->ns == 0) return GF_OK;
    hipLaunchKernelGGL(shell_results_kernel, dim3((m->ns + 255) / 256), dim3(256), 0, m->stream, m->ns, (size_t)m->ns,
                       m->d_sconst, m->d_sgp, d_out);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_shell_eval_device(int count, const double *x, const double *prop, const double *d, const double *gp,
                         double *K_out, double *f_out)
{
    if (count <= 0) return GF_OK;
    size_t ld = (size_t)count;
    std::vector<int> conn(6 * ld);
    std::vector<double> cst(SC_COUNT * ld), gps(SG_COUNT * ld), disp(36 * ld);
    std::vector<double> rec(SC_COUNT), prec(SP_COUNT), g0(SG_COUNT);
    for (int e = 0; e < count; e++) {
        gf_shell_host_const(x + 18 * (size_t)e, prop + GF_SHELL_NPROP * (size_t)e, rec.data(), prec.data(), g0.data());
        for (int k = 0; k < SC_COUNT; k++) cst[k * ld + e] = rec[k];
        for (int k = 0; k < SG_COUNT; k++) gps[k * ld + e] = gp[(size_t)e * SG_COUNT + k];
        for (int a = 0; a < 6; a++) conn[a * ld + e] = 6 * e + a;
        for (int k = 0; k < 36; k++) disp[36 * (size_t)e + k] = d[36 * (size_t)e + k];
    }
    int *d_conn = nullptr;
    double *d_cst = nullptr, *d_gps = nullptr, *d_disp = nullptr, *d_K = nullptr, *d_f = nullptr;
    hipStream_t s;
    GF_HIP(hipStreamCreate(&s));
    GF_HIP(hipMalloc(&d_conn, sizeof(int) * conn.size()));
    GF_HIP(hipMalloc(&d_cst, sizeof(double) * cst.size()));
    GF_HIP(hipMalloc(&d_gps, sizeof(double) * gps.size()));
    GF_HIP(hipMalloc(&d_disp, sizeof(double) * disp.size()));
    GF_HIP(hipMalloc(&d_K, sizeof(double) * 729 * ld));
    GF_HIP(hipMalloc(&d_f, sizeof(double) * 27 * ld));
    GF_HIP(hipMemcpyAsync(d_conn, conn.data(), sizeof(int) * conn.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_cst, cst.data(), sizeof(double) * cst.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_gps, gps.data(), sizeof(double) * gps.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_disp, disp.data(), sizeof(double) * disp.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(shell_element_kernel<false>, dim3((count + SEPB - 1) / SEPB), dim3(SWG), 0, s, 0, count, ld,
                       d_conn, d_cst, d_gps, d_disp, d_K, d_f, (const double *)nullptr);
    GF_HIP(hipGetLastError());
    GF_HIP(hipMemcpyAsync(K_out, d_K, sizeof(double) * 729 * ld, hipMemcpyDeviceToHost, s));
    GF_HIP(hipMemcpyAsync(f_out, d_f, sizeof(double) * 27 * ld, hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_conn); (void)hipFree(d_cst); (void)hipFree(d_gps);
    (void)hipFree(d_disp); (void)hipFree(d_K); (void)hipFree(d_f);
    (void)hipStreamDestroy(s);
    return GF_OK;
}
