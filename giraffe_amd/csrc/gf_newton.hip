// Newton-step support kernels: nodal loads, FP64 max-norm scans, displacement
// update and the Rodrigues commit of the nodes.
//
// Reference: NodalLoad::Mount src/NodalLoad.cpp:322-401,
// NodalFollowerLoad::Mount src/NodalFollowerLoad.cpp:243-326,
// Table::GetValueAt src/Table.cpp:73-101,
// ConvergenceCriteria scans src/ConvergenceCriteria.cpp:187-288, 289-459, 460-598,
// Solution::UpdateDisps src/Solution.cpp:390-423,
// Node::SaveConfiguration src/Node.cpp:325-384.
#include "gf_internal.h"

#include <algorithm>
#include "gf_device.h"

namespace {

__global__ void fill_kernel(double *p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Zeroes three ranges; in the second, the entries dfirst + k dstride (k < dcount)
// get 1 instead (the unit diagonal of the padding rows of the global matrix).
__global__ __launch_bounds__(256) void clear3_kernel(double *a, size_t na, double *b, size_t nb, double *c, size_t nc,
                                                     size_t dfirst, size_t dstride, size_t dcount)
{
    const size_t n = na + nb + nc;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double *p = i < na ? a + i : i < na + nb ? b + (i - na) : c + (i - na - nb);
        double v = 0.0;
        if (i >= na + dfirst && i < na + nb) {
            const size_t d = i - na - dfirst;
            if (d % dstride == 0 && d / dstride < dcount) v = 1.0;
        }
        *p = v;
    }
}

// Zeroes a and c, and the band entries (row, cols[k]) for k in [ptr[row],
// ptr[row + 1]) of rows < nrows (the positions the element scatter writes): one
// wavefront per band row, the short ranges by every thread.
__global__ __launch_bounds__(256) void clear_csr_kernel(double *a, size_t na, double *c, size_t nc, double *ab,
                                                         int ld, int hb, const int *ptr, const int *cols, int nrows)
{
    const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (gid < na) a[gid] = 0.0;
    if (gid < nc) c[gid] = 0.0;
    const int row = (int)(gid >> 6), lane = threadIdx.x & 63;
    if (row >= nrows) return;
    double *base = ab + (size_t)row * (size_t)ld + (size_t)hb;
    for (int k = ptr[row] + lane; k < ptr[row + 1]; k += 64) base[cols[k]] = 0.0;
}

__global__ void fill_u32_kernel(unsigned *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}


// One thread per loaded node: all loads on that node, in load order.
__global__ __launch_bounds__(128) void nodal_loads_kernel(int n_ln, const int *__restrict__ ln_node, const int *__restrict__ ln_off,
                                   const int *__restrict__ ln_ids, const double *__restrict__ rec,
                                   const double *__restrict__ tab_t, const double *__restrict__ tab_v,
                                   const double *__restrict__ disp, const double *__restrict__ copy,
                                   const int *__restrict__ dof, gf_gmat gm, double *r, const double *__restrict__ tp)
{
    const double t = *tp;  // load time, device-resident (a captured Newton step reads the current one)
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_ln) return;
    int node = ln_node[k];
    int R[6];
#pragma unroll
    for (int i = 0; i < 6; i++) R[i] = dof[6 * node + i];
    v3<double> alpha = mk3(disp[6 * node + 3], disp[6 * node + 4], disp[6 * node + 5]);
    for (int q = ln_off[k]; q < ln_off[k + 1]; q++) {
        const double *L = rec + (size_t)ln_ids[q] * NL_REC;
        int kind = (int)L[NL_KIND];
        m3<double> CS;
#pragma unroll
        for (int i = 0; i < 9; i++) CS.a[i] = L[NL_CS + i];
        int r0 = (int)L[NL_TAB0], nt = (int)L[NL_TABN];
        double tv6[6];
        gf_table_values6(tab_t, tab_v, r0, nt, t, tv6);
        v3<double> f = mk3(L[NL_MULT + 0] * tv6[0], L[NL_MULT + 1] * tv6[1], L[NL_MULT + 2] * tv6[2]);
        v3<double> mo = mk3(L[NL_MULT + 3] * tv6[3], L[NL_MULT + 4] * tv6[4], L[NL_MULT + 5] * tv6[5]);
        if (kind == GF_LOAD_NODAL) {
            // dead force + pseudo-moment Xi^T m with stiffness -V(alpha, Xi^T m)
            f = mtv(CS, f);
            mo = mtv(CS, mo);
            m3<double> Q, Xi;
            double g;
            rodrigues(alpha, Q, Xi, g);
            mo = mtv(Xi, mo);
            m3<double> S = opV(alpha, mo);
            const double q6[6] = {f.x, f.y, f.z, mo.x, mo.y, mo.z};
            vec_add_n<6>(r, R, q6);
            int R9[9], C9[9];
            double V9[9];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    R9[3 * i + j] = R[3 + i];
                    C9[3 * i + j] = R[3 + j];
                    V9[3 * i + j] = -S.a[3 * i + j];
                }
            gm_add_n<9>(gm, R9, C9, V9);
        } else if (kind == GF_LOAD_NODAL_FOLLOWER) {
            v3<double> ai = mk3(copy[6 * node + 3], copy[6 * node + 4], copy[6 * node + 5]);
            m3<double> Qi, Xii, Q, Xi;
            double gi, g;
            rodrigues(ai, Qi, Xii, gi);
            Qi = mm3(Qi, tr3(CS));
            rodrigues(alpha, Q, Xi, g);
            v3<double> fip = mv(mm3(Q, Qi), f);
            v3<double> Qim = mv(Qi, mo);
            v3<double> mip = mv(Xi, Qim);
            m3<double> K12 = mscl(-1.0, mm3(skew3(fip), Xi));
            m3<double> dy;
#pragma unroll
            for (int i = 0; i < 9; i++) dy.a[i] = 0.0;
            add_outer(dy, 1.0, Qim, alpha);
            m3<double> K22 = mscl(-0.5 * g, madd(skew3(Qim), mm3(Xi, dy)));
            const double q6[6] = {fip.x, fip.y, fip.z, mip.x, mip.y, mip.z};
            vec_add_n<6>(r, R, q6);
            int R18[18], C18[18];
            double V18[18];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    R18[3 * i + j] = R[i];
                    C18[3 * i + j] = R[3 + j];
                    V18[3 * i + j] = -K12.a[3 * i + j];
                    R18[9 + 3 * i + j] = R[3 + i];
                    C18[9 + 3 * i + j] = R[3 + j];
                    V18[9 + 3 * i + j] = -K22.a[3 * i + j];
                }
            gm_add_n<18>(gm, R18, C18, V18);
        }
    }
}

// Max-norms of a vector over band rows split by DOF kind (translation /
// rotation) plus a non-finite flag.  out (as unsigned long long bits of
// non-negative doubles): [0] trans, [1] rot, [2] flag.
__device__ __forceinline__ void amax(unsigned long long *p, double v)
{
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Workgroup max of NV non-negative values (256 lanes): wave shuffles, then LDS
// across the 4 waves, then one atomic per value and workgroup.
template <int NV> __device__ __forceinline__ void block_amax(double (&v)[NV], unsigned long long *out)
{
    __shared__ double sred[4][NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        for (int off = 32; off > 0; off >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], off));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sred[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double mx = sred[0][threadIdx.x];
        for (int w = 1; w < 4; w++) mx = fmax(mx, sred[w][threadIdx.x]);
        amax(out + threadIdx.x, mx);
    }
}

// max |r| over translational and rotational free DOFs, and a non-finite flag.
__global__ __launch_bounds__(256) void scan_kernel(int n, const double *__restrict__ v,
                                                   const int *__restrict__ row_nc, unsigned long long *out)
{
    double m[3] = {0.0, 0.0, 0.0};  // translations, rotations, non-finite
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = fabs(v[i]);
        if (!(a == a) || a >= 1e300) m[2] = 1.0;
        else if (row_nc[i] % 6 < 3) m[0] = fmax(m[0], a);
        else m[1] = fmax(m[1], a);
    }
    block_amax<3>(m, out);
}

// d += x on free DOFs; scans |x| and the updated |d| (CheckGLConvergence).
__global__ __launch_bounds__(256) void update_kernel(int n, const double *__restrict__ x,
                                                     const int *__restrict__ row_nc, double *__restrict__ disp,
                                                     unsigned long long *out)
{
    double m[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |x| t/r, |d| t/r, non-finite
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int nc = row_nc[i];
        double xi = x[i];
        double dn = disp[nc] + xi;
        disp[nc] = dn;
        double ax = fabs(xi), ad = fabs(dn);
        if (!(ax == ax) || ax >= 1e300) m[4] = 1.0;
        if (nc % 6 < 3) { m[0] = fmax(m[0], ax); m[2] = fmax(m[2], ad); }
        else { m[1] = fmax(m[1], ax); m[3] = fmax(m[3], ad); }
    }
    block_amax<5>(m, out);
}

// Node::SaveConfiguration: x += u, alpha <- compose(alpha_delta, alpha), Euler vector.
__global__ void node_commit_kernel(int nn, double *__restrict__ copy, const double *__restrict__ disp,
                                   double *__restrict__ euler)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double *c = copy + 6 * (size_t)i;
    const double *d = disp + 6 * (size_t)i;
    c[0] += d[0]; c[1] += d[1]; c[2] += d[2];
    v3<double> a3 = compose(mk3(d[3], d[4], d[5]), mk3(c[3], c[4], c[5]));
    c[3] = a3.x; c[4] = a3.y; c[5] = a3.z;
    double na = sqrt(dot3(a3, a3));
    double th = 2.0 * atan(na / 2.0);
    double s = na != 0.0 ? th / na : 0.0;
    euler[3 * (size_t)i + 0] = s * a3.x;
    euler[3 * (size_t)i + 1] = s * a3.y;
    euler[3 * (size_t)i + 2] = s * a3.z;
}

// Unit diagonal for the padding rows past n_free.

}  // namespace

static inline int grid_for(size_t n, int block, int cap = 4096)
{
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

int gf_launch_fill(hipStream_t s, double *p, size_t n, double v)
{
    if (n == 0) return GF_OK;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_launch_clear_csr(hipStream_t s, double *a, size_t na, double *c, size_t nc, double *ab, int ld, int hb,
                        const int *ptr, const int *cols, int nrows)
{
    const size_t n = std::max((size_t)nrows * 64, std::max(na, nc));
    hipLaunchKernelGGL(clear_csr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, na, c, nc, ab, ld, hb,
                       ptr, cols, nrows);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_launch_clear3(hipStream_t s, double *a, size_t na, double *b, size_t nb, double *c, size_t nc, size_t dfirst,
                     size_t dstride, size_t dcount)
{
    const size_t n = na + nb + nc;
    if (n == 0) return GF_OK;
    if (dstride == 0) dcount = 0, dstride = 1;
    hipLaunchKernelGGL(clear3_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, a, na, b, nb, c, nc, dfirst, dstride,
                       dcount);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

namespace {
__global__ void set2_kernel(double *p, double a, double b)
{
    if (threadIdx.x == 0) {
        p[0] = a;
        p[1] = b;
    }
}
}  // namespace

int gf_launch_set2(hipStream_t s, double *p, double a, double b)
{
    hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(64), 0, s, p, a, b);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_launch_fill_bytes(hipStream_t s, void *p, size_t bytes)
{
    size_t n = bytes / 4;
    if (n == 0) return GF_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, (unsigned *)p, n);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_nodal_loads(gf_model *m, const double *tp)
{
    if (m->n_load_nodes == 0) return GF_OK;
    hipLaunchKernelGGL(nodal_loads_kernel, dim3((m->n_load_nodes + 127) / 128), dim3(128), 0, m->stream,
                       m->n_load_nodes, m->d_ln_node, m->d_ln_off, m->d_ln_ids, m->d_lrec, m->d_tab_t, m->d_tab_v,
                       m->d_disp, m->d_copy, m->d_dof, gf_model_gmat(m), m->d_r, tp);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

// The norms land in d_scal[GF_SCAL_RES..+2], which the assembly's clear zeroed.
int gf_scan_residual(gf_model *m)
{
    hipLaunchKernelGGL(scan_kernel, dim3(grid_for(m->n_free, 256, 128)), dim3(256), 0, m->stream, m->n_free, m->d_r,
                       m->d_row_nc, (unsigned long long *)(m->d_scal + GF_SCAL_RES));
    GF_HIP(hipGetLastError());
    return GF_OK;
}

// zero_norms: clear d_scal[0..7] first (callers that cleared it pass false).
int gf_update_disp(gf_model *m, bool zero_norms)
{
    if (zero_norms) {
        int st = gf_launch_fill_bytes(m->stream, m->d_scal, GF_SCAL_FLAG * sizeof(double));
        if (st) return st;
    }
    hipLaunchKernelGGL(update_kernel, dim3(grid_for(m->n_free, 256, 128)), dim3(256), 0, m->stream, m->n_free,
                       m->d_x, m->d_row_nc, m->d_disp, (unsigned long long *)m->d_scal);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int gf_node_commit(gf_model *m)
{
    hipLaunchKernelGGL(node_commit_kernel, dim3((m->nn + 255) / 256), dim3(256), 0, m->stream, m->nn, m->d_copy,
                       m->d_disp, m->d_euler);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

