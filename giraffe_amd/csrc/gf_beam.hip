// Beam_1 on gfx950: element tangent + internal force fused with the
// conflict-free scatter into the block-tridiagonal global matrix.
//
// Reference: Beam_1::Mount src/Beam_1.cpp:695-835, MountElementLoads /
// MountFieldLoading (gravity) :838-878, MountGlobal :1429-1492,
// SaveLagrange :1494-1506, PreCalc :501-692.
//
// Work split per workgroup of one wavefront (64 lanes) and EPB = 7 elements
// (a colour of 5,000 elements is 715 workgroups, so every CU holds several and
// their latencies overlap):
//   phase 1 -- one lane per (element, Gauss point): Rodrigues kinematics,
//              back-rotated strains, stresses and the 9x9 operator
//              M = B^T D B + G (block-sparse: 7 non-zero 3x3 blocks) plus
//              B^T sigma, staged in LDS;
//   phase 2 -- one lane per (element, node pair a,b): the 6x6 block
//              K_ab = sum_gp w dN_a^T M dN_b using the sparsity of dN
//              (dN_a = [dN_a I, 0; 0, dN_a I; 0, N_a I]), rotated to the
//              global frame (T3^T S T3) and added into the matrix;
//   phase 3 -- one lane per (element, node): internal force minus self-weight
//              subtracted from the residual r = f_ext - f_int.
// Elements are sorted by colour (no two elements of a colour share a node),
// so one launch per colour scatters without atomics and deterministically.
#include "gf_internal.h"
#include "gf_device.h"

namespace {

constexpr int EPB = 7;       // elements per workgroup
constexpr int WG = 64;
constexpr int LDS_GP = 73;   // 7 blocks * 9 + 9 (B^T sigma) + the Gauss point's mass per length

struct BeamGP {
    v3<double> a, da, dz, eta, kap;
    m3<double> Q, Xi, dXi, QQ;
    double g;
};

__device__ __forceinline__ void beam_kin(const double *__restrict__ cst, const double *__restrict__ gps,
                                         size_t ld, int e, int gp, const double d[18], BeamGP &k)
{
    m3<double> T3;
#pragma unroll
    for (int i = 0; i < 9; i++) T3.a[i] = cst[(BC_T3 + i) * ld + e];
    double N[3], dN[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        N[i] = cst[(BC_N + 3 * gp + i) * ld + e];
        dN[i] = cst[(BC_DN + 3 * gp + i) * ld + e];
    }
    v3<double> a, da, du;
    a = mk3(d[3] * N[0] + d[9] * N[1] + d[15] * N[2], d[4] * N[0] + d[10] * N[1] + d[16] * N[2],
            d[5] * N[0] + d[11] * N[1] + d[17] * N[2]);
    da = mk3(d[3] * dN[0] + d[9] * dN[1] + d[15] * dN[2], d[4] * dN[0] + d[10] * dN[1] + d[16] * dN[2],
             d[5] * dN[0] + d[11] * dN[1] + d[17] * dN[2]);
    du = mk3(d[0] * dN[0] + d[6] * dN[1] + d[12] * dN[2], d[1] * dN[0] + d[7] * dN[1] + d[13] * dN[2],
             d[2] * dN[0] + d[8] * dN[1] + d[14] * dN[2]);
    k.a = mv(T3, a);
    k.da = mv(T3, da);
    du = mv(T3, du);
    rodrigues(k.a, k.Q, k.Xi, k.g);
    m3<double> dA = skew3(k.da);
    double ada = dot3(k.a, k.da);
#pragma unroll
    for (int i = 0; i < 9; i++) k.dXi.a[i] = -0.5 * k.g * (ada * k.Xi.a[i] - dA.a[i]);
    const size_t o = (size_t)BG_PER_GP * gp;
    m3<double> Qi;
#pragma unroll
    for (int i = 0; i < 9; i++) Qi.a[i] = gps[(o + i) * ld + e];
    v3<double> kap_i = mk3(gps[(o + 9) * ld + e], gps[(o + 10) * ld + e], gps[(o + 11) * ld + e]);
    v3<double> dz_i = mk3(gps[(o + 12) * ld + e], gps[(o + 13) * ld + e], gps[(o + 14) * ld + e]);
    k.dz = add3(du, dz_i);
    k.QQ = mm3(k.Q, Qi);
    v3<double> e3r = mk3(cst[(BC_E3R) * ld + e], cst[(BC_E3R + 1) * ld + e], cst[(BC_E3R + 2) * ld + e]);
    k.eta = sub3(mtv(k.QQ, k.dz), e3r);
    k.kap = add3(mtv(Qi, mtv(k.Xi, k.da)), kap_i);
}

__device__ __forceinline__ void gather_disp(const double *__restrict__ disp, const int *__restrict__ conn,
                                            size_t ld, int e, double d[18])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double *p = disp + 6 * (size_t)conn[a * ld + e];
#pragma unroll
        for (int k = 0; k < 6; k++) d[6 * a + k] = p[k];
    }
}

// MODE 0: scatter into (A,B,C) and r.  MODE 1: write dense element outputs.
// buoy (null: none): Pipe_1 buoyant mass per length, applied at Gauss points
// below the ocean surface through `surf` (a point of it), with the converged
// coordinates `copy`.
template <int MODE>
__global__ __launch_bounds__(WG) void beam_assemble_kernel(
    int e0, int e1, size_t ld, const int *__restrict__ conn, const double *__restrict__ cst,
    const double *__restrict__ gps, const double *__restrict__ disp, const int *__restrict__ dof,
    gf_gmat gm, double *__restrict__ r, const double *__restrict__ lfp, double gx, double gy, double gz,
    double *__restrict__ Kout, double *__restrict__ fout, const double *__restrict__ buoy,
    const double *__restrict__ copy, double sx, double sy, double sz)
{
    __shared__ double sm[EPB * 2 * LDS_GP];
    const int tid = threadIdx.x;
    const int ebase = e0 + blockIdx.x * EPB;
    // gravity ramp factor, device-resident so a captured Newton step reads the current one
    const double lf = lfp ? *lfp : 0.0;

    // ---- phase 1: per (element, Gauss point) -------------------------------
    if (tid < 2 * EPB) {
        int le = tid >> 1, gp = tid & 1;
        int e = ebase + le;
        if (e < e1) {
            double d[18];
            gather_disp(disp, conn, ld, e, d);
            BeamGP k;
            beam_kin(cst, gps, ld, e, gp, d, k);
            double GA1 = cst[(BC_D + 0) * ld + e], GA2 = cst[(BC_D + 1) * ld + e], EA = cst[(BC_D + 2) * ld + e];
            double EI1 = cst[(BC_D + 3) * ld + e], EI2 = cst[(BC_D + 4) * ld + e], EI12 = cst[(BC_D + 5) * ld + e];
            double GJ = cst[(BC_D + 6) * ld + e];
            v3<double> sn = mk3(GA1 * k.eta.x, GA2 * k.eta.y, EA * k.eta.z);
            v3<double> smv = mk3(EI1 * k.kap.x + EI12 * k.kap.y, EI12 * k.kap.x + EI2 * k.kap.y, GJ * k.kap.z);
            v3<double> n = mv(k.QQ, sn), mm = mv(k.QQ, smv);
            m3<double> QQt = tr3(k.QQ);
            m3<double> dZ = skew3(k.dz);
            m3<double> B00 = QQt;
            m3<double> B02 = mm3(QQt, mm3(dZ, k.Xi));
            m3<double> B11 = mm3(QQt, k.Xi);
            m3<double> B12 = mm3(QQt, k.dXi);
            // D-weighted copies: (Dn B)_ij = Dn_i B_ij, (Dm B) = Dm B
            m3<double> DB00, DB02, DB11, DB12;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                DB00.a[j] = GA1 * B00.a[j]; DB00.a[3 + j] = GA2 * B00.a[3 + j]; DB00.a[6 + j] = EA * B00.a[6 + j];
                DB02.a[j] = GA1 * B02.a[j]; DB02.a[3 + j] = GA2 * B02.a[3 + j]; DB02.a[6 + j] = EA * B02.a[6 + j];
                DB11.a[j] = EI1 * B11.a[j] + EI12 * B11.a[3 + j];
                DB11.a[3 + j] = EI12 * B11.a[j] + EI2 * B11.a[3 + j];
                DB11.a[6 + j] = GJ * B11.a[6 + j];
                DB12.a[j] = EI1 * B12.a[j] + EI12 * B12.a[3 + j];
                DB12.a[3 + j] = EI12 * B12.a[j] + EI2 * B12.a[3 + j];
                DB12.a[6 + j] = GJ * B12.a[6 + j];
            }
            m3<double> M00 = mtm3(B00, DB00);
            m3<double> M02 = mtm3(B00, DB02);
            m3<double> M11 = mtm3(B11, DB11);
            m3<double> M12 = mtm3(B11, DB12);
            m3<double> M22 = madd(mtm3(B02, DB02), mtm3(B12, DB12));
            // geometric part G
            m3<double> Sn = skew3(n), Sm = skew3(mm);
            m3<double> XiT = tr3(k.Xi);
            m3<double> G02 = mscl(-1.0, mm3(Sn, k.Xi));
            m3<double> G20 = mm3(XiT, Sn);
            m3<double> V1 = opV(k.a, mv(dZ, n));
            m3<double> V2 = opV(k.a, mm);
            m3<double> dV = opdV(k.a, k.da, mm);
            m3<double> SmXi = mm3(Sm, k.Xi);
            m3<double> G22 = madd(msub(mm3(XiT, mm3(mm3(dZ, Sn), k.Xi)), V1), msub(dV, mtm3(k.dXi, SmXi)));
            m3<double> G12 = msub(V2, mm3(XiT, SmXi));
            double *o = sm + (le * 2 + gp) * LDS_GP;
            m3<double> P[7];
            P[0] = M00;                    // (0,0)
            P[1] = madd(M02, G02);         // (0,2)
            P[2] = M11;                    // (1,1)
            P[3] = madd(M12, G12);         // (1,2)
            P[4] = madd(tr3(M02), G20);    // (2,0)
            P[5] = madd(tr3(M12), V2);     // (2,1)
            P[6] = madd(M22, G22);         // (2,2)
#pragma unroll
            for (int b = 0; b < 7; b++)
#pragma unroll
                for (int i = 0; i < 9; i++) o[9 * b + i] = P[b].a[i];
            v3<double> s0 = mtv(B00, sn), s1 = mtv(B11, smv);
            v3<double> s2 = add3(mtv(B02, sn), mtv(B12, smv));
            o[63] = s0.x; o[64] = s0.y; o[65] = s0.z;
            o[66] = s1.x; o[67] = s1.y; o[68] = s1.z;
            o[69] = s2.x; o[70] = s2.y; o[71] = s2.z;
            if (buoy) {
                // Pipe_1::MountFieldLoading src/Pipe_1.cpp:1289-1306: Gauss-point position
                // (converged + incremental), depth along G, Rho - rho_f Aext below the surface
                double z[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    double c = 0.0, u = 0.0;
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const double Na = cst[(BC_N + 3 * gp + a) * ld + e];
                        c += copy[6 * (size_t)conn[a * ld + e] + k] * Na;
                        u += d[6 * a + k] * Na;
                    }
                    z[k] = c + u;
                }
                const double depth =
                    ((z[0] - sx) * gx + (z[1] - sy) * gy + (z[2] - sz) * gz) * (1.0 / sqrt(gx * gx + gy * gy + gz * gz));
                const double rhoA = cst[BC_RHOA * ld + e];
                o[72] = depth > 0 ? rhoA - buoy[e] : rhoA;
            }
        }
    }
    __syncthreads();

    // ---- phase 2: per (element, node pair) --------------------------------
    if (tid < 9 * EPB) {
        int le = tid / 9, ab = tid - 9 * (tid / 9);
        int a = ab / 3, b = ab - 3 * (ab / 3);
        int e = ebase + le;
        if (e < e1) {
            double w = cst[BC_JAC * ld + e];
            m3<double> uu, ua, au, aa;
#pragma unroll
            for (int i = 0; i < 9; i++) uu.a[i] = ua.a[i] = au.a[i] = aa.a[i] = 0.0;
#pragma unroll
            for (int gp = 0; gp < 2; gp++) {
                const double *o = sm + (le * 2 + gp) * LDS_GP;
                double Na = cst[(BC_N + 3 * gp + a) * ld + e], Nb = cst[(BC_N + 3 * gp + b) * ld + e];
                double dNa = cst[(BC_DN + 3 * gp + a) * ld + e], dNb = cst[(BC_DN + 3 * gp + b) * ld + e];
                double c00 = w * dNa * dNb, c02 = w * dNa * Nb, c20 = w * Na * dNb, c22 = w * Na * Nb;
#pragma unroll
                for (int i = 0; i < 9; i++) {
                    uu.a[i] += c00 * o[i];
                    ua.a[i] += c02 * o[9 + i];
                    au.a[i] += c20 * o[36 + i];
                    aa.a[i] += c00 * o[18 + i] + c02 * o[27 + i] + c20 * o[45 + i] + c22 * o[54 + i];
                }
            }
            m3<double> T3;
#pragma unroll
            for (int i = 0; i < 9; i++) T3.a[i] = cst[(BC_T3 + i) * ld + e];
            m3<double> blk[4];
            blk[0] = mtm3(T3, mm3(uu, T3));
            blk[1] = mtm3(T3, mm3(ua, T3));
            blk[2] = mtm3(T3, mm3(au, T3));
            blk[3] = mtm3(T3, mm3(aa, T3));
            if (MODE == 0) {
                int na = conn[a * ld + e], nbn = conn[b * ld + e];
                int rr[6], cc[6];
#pragma unroll
                for (int k = 0; k < 6; k++) { rr[k] = dof[6 * na + k]; cc[k] = dof[6 * nbn + k]; }
                int R36[36], C36[36];
                double V36[36];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    int ro = (q >> 1) * 3, co = (q & 1) * 3;
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            const int k = 9 * q + 3 * i + j;
                            R36[k] = rr[ro + i];
                            C36[k] = cc[co + j];
                            V36[k] = blk[q].a[3 * i + j];
                        }
                }
                gm_add_n<36>(gm, R36, C36, V36);
            } else {
                double *Ke = Kout + (size_t)e * 324;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    int ro = (q >> 1) * 3, co = (q & 1) * 3;
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) Ke[(6 * a + ro + i) * 18 + 6 * b + co + j] = blk[q].a[3 * i + j];
                }
            }
        }
    }

    // ---- phase 3: per (element, node): internal force - self-weight --------
    if (tid < 3 * EPB) {
        int le = tid / 3, a = tid - 3 * (tid / 3);
        int e = ebase + le;
        if (e < e1) {
            double w = cst[BC_JAC * ld + e];
            v3<double> fu = mk3(0.0, 0.0, 0.0), fa = mk3(0.0, 0.0, 0.0);
            double Nsum = 0.0;
#pragma unroll
            for (int gp = 0; gp < 2; gp++) {
                const double *o = sm + (le * 2 + gp) * LDS_GP;
                double Na = cst[(BC_N + 3 * gp + a) * ld + e], dNa = cst[(BC_DN + 3 * gp + a) * ld + e];
                Nsum += Na;
                fu = add3(fu, scl3(w * dNa, mk3(o[63], o[64], o[65])));
                fa = add3(fa, add3(scl3(w * dNa, mk3(o[66], o[67], o[68])), scl3(w * Na, mk3(o[69], o[70], o[71]))));
            }
            m3<double> T3;
#pragma unroll
            for (int i = 0; i < 9; i++) T3.a[i] = cst[(BC_T3 + i) * ld + e];
            fu = mtv(T3, fu);
            fa = mtv(T3, fa);
            double P[6] = {fu.x, fu.y, fu.z, fa.x, fa.y, fa.z};
            if (MODE == 0) {
                double mult = lf * w * cst[BC_RHOA * ld + e] * Nsum;
                if (buoy) {
                    double acc = 0.0;
#pragma unroll
                    for (int gp = 0; gp < 2; gp++)
                        acc += cst[(BC_N + 3 * gp + a) * ld + e] * sm[(le * 2 + gp) * LDS_GP + 72];
                    mult = lf * w * acc;
                }
                P[0] -= mult * gx; P[1] -= mult * gy; P[2] -= mult * gz;
                int na = conn[a * ld + e];
                int R6[6];
                double V6[6];
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    R6[k] = dof[6 * na + k];
                    V6[k] = -P[k];
                }
                vec_add_n<6>(r, R6, V6);
            } else {
#pragma unroll
                for (int k = 0; k < 6; k++) fout[(size_t)e * 18 + 6 * a + k] = P[k];
            }
        }
    }
}

// SaveLagrange: Q_i <- Q_delta Q_i, kappa_i <- kappa_r, z'_i <- z'.
__global__ __launch_bounds__(256) void beam_commit_kernel(int nb, size_t ld, const int *__restrict__ conn,
                                                          const double *__restrict__ cst, double *__restrict__ gps,
                                                          const double *__restrict__ disp)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nb) return;
    int e = t >> 1, gp = t & 1;
    double d[18];
    gather_disp(disp, conn, ld, e, d);
    BeamGP k;
    beam_kin(cst, gps, ld, e, gp, d, k);
    const size_t o = (size_t)BG_PER_GP * gp;
    m3<double> Qi;
#pragma unroll
    for (int i = 0; i < 9; i++) Qi.a[i] = gps[(o + i) * ld + e];
    m3<double> Qn = mm3(k.Q, Qi);
#pragma unroll
    for (int i = 0; i < 9; i++) gps[(o + i) * ld + e] = Qn.a[i];
    gps[(o + 9) * ld + e] = k.kap.x;
    gps[(o + 10) * ld + e] = k.kap.y;
    gps[(o + 11) * ld + e] = k.kap.z;
    gps[(o + 12) * ld + e] = k.dz.x;
    gps[(o + 13) * ld + e] = k.dz.y;
    gps[(o + 14) * ld + e] = k.dz.z;
}

// Results of the committed state (Beam_1::WriteResults src/Beam_1.cpp:444-458):
// per element the means over both Gauss points of sigma_r = D epsilon_r and of
// epsilon_r = (eta_r, kappa_r).  With zero incremental displacements the
// kinematics reduce to the committed Q_i, kappa_i, z'_i, so eta_r and kappa_r
// equal the values the last Mount computed at the converged configuration.
// out: [nb][12] in colour-sorted order.
__global__ __launch_bounds__(256) void beam_results_kernel(int nb, size_t ld, const double *__restrict__ cst,
                                                           const double *__restrict__ gps, double *__restrict__ out)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb) return;
    const double GA1 = cst[(BC_D + 0) * ld + e], GA2 = cst[(BC_D + 1) * ld + e], EA = cst[(BC_D + 2) * ld + e];
    const double EI1 = cst[(BC_D + 3) * ld + e], EI2 = cst[(BC_D + 4) * ld + e], EI12 = cst[(BC_D + 5) * ld + e];
    const double GJ = cst[(BC_D + 6) * ld + e];
    double d[18];
#pragma unroll
    for (int i = 0; i < 18; i++) d[i] = 0.0;
    double sg[2][6], ep[2][6];
#pragma unroll
    for (int gp = 0; gp < 2; gp++) {
        BeamGP k;
        beam_kin(cst, gps, ld, e, gp, d, k);
        ep[gp][0] = k.eta.x; ep[gp][1] = k.eta.y; ep[gp][2] = k.eta.z;
        ep[gp][3] = k.kap.x; ep[gp][4] = k.kap.y; ep[gp][5] = k.kap.z;
        sg[gp][0] = GA1 * k.eta.x;
        sg[gp][1] = GA2 * k.eta.y;
        sg[gp][2] = EA * k.eta.z;
        sg[gp][3] = EI1 * k.kap.x + EI12 * k.kap.y;
        sg[gp][4] = EI12 * k.kap.x + EI2 * k.kap.y;
        sg[gp][5] = GJ * k.kap.z;
    }
    double *o = out + 12 * (size_t)e;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        o[i] = (sg[0][i] + sg[1][i]) / 2;
        o[6 + i] = (ep[0][i] + ep[1][i]) / 2;
    }
}

}  // namespace

// Gauss abscissa as written in src/Beam_1.cpp:629-632.
static const double kBeamCsi = 0.577350269189626;

void gf_beam_host_const(const double *x, const double *cs, const double *prop, int kind, double *rec, double *gp)
{
    double T0 = 0.0;
    for (int i = 0; i < 9; i++) rec[BC_T3 + i] = cs[i];
    if (kind == GF_BEAM_KIND_PIPE1) {
        // Pipe_1::PreCalc src/Pipe_1.cpp:1108-1135: D from the pipe section, Rho per length
        const double EA = prop[0], EI = prop[1], GJ = prop[2], GA = prop[3], Rho = prop[4];
        rec[BC_D + 0] = GA;
        rec[BC_D + 1] = GA;
        rec[BC_D + 2] = EA;
        rec[BC_D + 3] = EI;
        rec[BC_D + 4] = EI;
        rec[BC_D + 5] = 0.0;
        rec[BC_D + 6] = GJ;
        rec[BC_RHOA] = Rho;
    } else {
        double E = prop[0], nu = prop[1], rho = prop[2], A = prop[3];
        double I1 = prop[4], I2 = prop[5], I12 = prop[6], It = prop[7];
        T0 = prop[8];
        double G = E / (2 * (1 + nu));
        double sf = 1.0;
        rec[BC_D + 0] = sf * G * A;
        rec[BC_D + 1] = sf * G * A;
        rec[BC_D + 2] = E * A;
        rec[BC_D + 3] = E * I1;
        rec[BC_D + 4] = E * I2;
        rec[BC_D + 5] = E * I12;
        rec[BC_D + 6] = G * It;
        rec[BC_RHOA] = rho * A;
    }
    double e3[3] = {x[6] - x[0], x[7] - x[1], x[8] - x[2]};
    double L0 = std::sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    for (int i = 0; i < 3; i++) e3[i] *= 1.0 / L0;
    for (int i = 0; i < 3; i++) rec[BC_E3R + i] = cs[3 * i] * e3[0] + cs[3 * i + 1] * e3[1] + cs[3 * i + 2] * e3[2];
    double du0 = T0 / rec[BC_D + 2];
    double jac = (L0 / (1.0 + du0)) / 2.0;
    rec[BC_JAC] = jac;
    for (int g = 0; g < 2; g++) {
        double s = g == 0 ? -kBeamCsi : kBeamCsi;
        rec[BC_N + 3 * g + 0] = 0.5 * s * (s - 1.0);
        rec[BC_N + 3 * g + 1] = 1.0 - s * s;
        rec[BC_N + 3 * g + 2] = 0.5 * s * (1.0 + s);
        rec[BC_DN + 3 * g + 0] = (1.0 / jac) * (s - 0.5);
        rec[BC_DN + 3 * g + 1] = (1.0 / jac) * (-2.0 * s);
        rec[BC_DN + 3 * g + 2] = (1.0 / jac) * (0.5 + s);
        double *st = gp + BG_PER_GP * g;
        for (int i = 0; i < BG_PER_GP; i++) st[i] = 0.0;
        st[0] = st[4] = st[8] = 1.0;
        st[14] = 1.0 + du0;
    }
}

int gf_beam_assemble(gf_model *m, const double *lfp)
{
    const size_t ld = (size_t)m->nb;
    int ncol = (int)m->beam_color_off.size() - 1;
    int launches = 0;
    for (int c = 0; c < ncol; c++) {
        int e0 = m->beam_color_off[c], e1 = m->beam_color_off[c + 1];
        if (e1 <= e0) continue;
        launches++;
        int grid = (e1 - e0 + EPB - 1) / EPB;
        const int pr = gf_probe_begin(m->stream, GF_KC_BEAM_ASSEMBLE);
        hipLaunchKernelGGL(beam_assemble_kernel<0>, dim3(grid), dim3(WG), 0, m->stream, e0, e1, ld, m->d_bconn,
                           m->d_bconst, m->d_bgp, m->d_disp, m->d_dof, gf_model_gmat(m), m->d_r, lfp, m->gravity[0],
                           m->gravity[1], m->gravity[2], (double *)nullptr, (double *)nullptr,
                           (const double *)m->d_bbuoy, (const double *)m->d_copy, m->ocean_surface[0],
                           m->ocean_surface[1], m->ocean_surface[2]);
        gf_probe_end(m->stream, pr, e1 - e0);
    }
    GF_HIP(hipGetLastError());
    return launches;
}

int gf_beam_commit(gf_model *m)
{
    if (m->nb == 0) return GF_OK;
    int grid = (2 * m->nb + 255) / 256;
    hipLaunchKernelGGL(beam_commit_kernel, dim3(grid), dim3(256), 0, m->stream, m->nb, (size_t)m->nb, m->d_bconn,
                       m->d_bconst, m->d_bgp, m->d_disp);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_beam_results(gf_model *m, double *d_out)
{
    if (m->nb == 0) return GF_OK;
    hipLaunchKernelGGL(beam_results_kernel, dim3((m->nb + 255) / 256), dim3(256), 0, m->stream, m->nb, (size_t)m->nb,
                       m->d_bconst, m->d_bgp, d_out);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

// Element-level entry: independent elements, each with its own three nodes.
int gf_beam_eval_device(int count, const double *x, const double *cs, const double *prop, const double *d,
                        const double *gp, double *K_out, double *f_out)
{
    if (count <= 0) return GF_OK;
    size_t ld = (size_t)count;
    std::vector<int> conn(3 * ld);
    std::vector<double> cst(BC_COUNT * ld), gps(BG_COUNT * ld), disp(18 * ld);
    std::vector<double> rec(BC_COUNT), g0(BG_COUNT);
    for (int e = 0; e < count; e++) {
        gf_beam_host_const(x + 9 * (size_t)e, cs + 9 * (size_t)e, prop + GF_BEAM_NPROP * (size_t)e, GF_BEAM_KIND_BEAM1,
                           rec.data(), g0.data());
        for (int k = 0; k < BC_COUNT; k++) cst[k * ld + e] = rec[k];
        for (int k = 0; k < BG_COUNT; k++) gps[k * ld + e] = gp[(size_t)e * BG_COUNT + k];
        for (int a = 0; a < 3; a++) conn[a * ld + e] = 3 * e + a;
        for (int k = 0; k < 18; k++) disp[18 * (size_t)e + k] = d[18 * (size_t)e + k];
    }
    int *d_conn = nullptr;
    double *d_cst = nullptr, *d_gps = nullptr, *d_disp = nullptr, *d_K = nullptr, *d_f = nullptr;
    hipStream_t s;
    GF_HIP(hipStreamCreate(&s));
    GF_HIP(hipMalloc(&d_conn, sizeof(int) * conn.size()));
    GF_HIP(hipMalloc(&d_cst, sizeof(double) * cst.size()));
    GF_HIP(hipMalloc(&d_gps, sizeof(double) * gps.size()));
    GF_HIP(hipMalloc(&d_disp, sizeof(double) * disp.size()));
    GF_HIP(hipMalloc(&d_K, sizeof(double) * 324 * ld));
    GF_HIP(hipMalloc(&d_f, sizeof(double) * 18 * ld));
    GF_HIP(hipMemcpyAsync(d_conn, conn.data(), sizeof(int) * conn.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_cst, cst.data(), sizeof(double) * cst.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_gps, gps.data(), sizeof(double) * gps.size(), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(d_disp, disp.data(), sizeof(double) * disp.size(), hipMemcpyHostToDevice, s));
    int grid = (count + EPB - 1) / EPB;
    gf_gmat none = {};
    hipLaunchKernelGGL(beam_assemble_kernel<1>, dim3(grid), dim3(WG), 0, s, 0, count, ld, d_conn, d_cst, d_gps,
                       d_disp, (const int *)nullptr, none, (double *)nullptr, (const double *)nullptr, 0.0, 0.0, 0.0, d_K,
                       d_f,
                       (const double *)nullptr, (const double *)nullptr, 0.0, 0.0, 0.0);
    GF_HIP(hipGetLastError());
    GF_HIP(hipMemcpyAsync(K_out, d_K, sizeof(double) * 324 * ld, hipMemcpyDeviceToHost, s));
    GF_HIP(hipMemcpyAsync(f_out, d_f, sizeof(double) * 18 * ld, hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_conn); (void)hipFree(d_cst); (void)hipFree(d_gps);
    (void)hipFree(d_disp); (void)hipFree(d_K); (void)hipFree(d_f);
    (void)hipStreamDestroy(s);
    return GF_OK;
}
