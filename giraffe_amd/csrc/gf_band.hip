// Blocked band LU on gfx950 for wide bands (shell meshes): the tangent solve
// that replaces sparsesystem -> MKL PARDISO (src/SparseMatrix.cpp:74-166)
// when the half-bandwidth does not fit the block-tridiagonal solver.
//
// Storage (GM_BAND): entry (i, c), |i - c| <= hb, at ab[i*2hb + c + hb], so any
// sub-matrix inside the band is a strided matrix with leading dimension 2hb.
// Entries outside the band alias neighbouring rows and are never touched:
// every load masks them to zero and every store skips them.
//
// Right-looking, no row interchanges (PARDISO's mtype 11 default is static
// pivoting), panel width NB = 64 over the padded order n (multiple of NB):
//   diag   LU of the NB x NB diagonal block            (one workgroup)
//   trsm   L21 = A21 U11^-1 and U12 = L11^-1 A12         (64-row / 64-column tiles)
//   update A22 -= L21 U12 on the band window, MFMA 16x16x4 tiles of 64x64;
//          the workgroup owning the next diagonal block factors it in place.
// Zero or non-finite pivots raise the failure flag (info_fail).
// Triangular solves: one launch per panel (block of NB unknowns solved by one
// workgroup, the coupled rows below / above updated by the rest).
#include "gf_internal.h"

namespace {

constexpr int NB = 64;
constexpr int KT = 32;  // k-chunk of the update tiles

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct mfma;
// Operands of both: A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k.  The
// 16 x 16 result differs: item v of lane l holds row 4 v + l / 16 for f64 and
// row 4 (l / 16) + v for f32, column l % 16 in both (checked on gfx950).
template <> struct mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ size_t bidx(int i, int c, int hb) { return (size_t)i * (size_t)(2 * hb) + (size_t)(c + hb); }
__device__ __forceinline__ bool inband(int i, int c, int hb) { return i - c <= hb && c - i <= hb; }

template <typename T> __device__ __forceinline__ bool bad_pivot(T v)
{
    return !(v == v) || v == T(0) || fabs((double)v) >= 1e300;
}

// In-LDS LU (no pivoting) of a 64 x 64 tile with leading dimension ld.
template <typename T> __device__ void lds_lu64(T *s, int ld, int *flag)
{
    const int tid = threadIdx.x;
    for (int k = 0; k < NB; k++) {
        T piv = s[k * ld + k];
        if (bad_pivot(piv)) {
            if (tid == 0) atomicExch(flag, 1);
            piv = T(1);
        }
        if (tid > k && tid < NB) s[tid * ld + k] /= piv;
        __syncthreads();
        int w = NB - 1 - k;
        for (int t = tid; t < w * w; t += 256) {
            int i = k + 1 + t / w, j = k + 1 + t % w;
            s[i * ld + j] -= s[i * ld + k] * s[k * ld + j];
        }
        __syncthreads();
    }
}

template <typename T> __global__ __launch_bounds__(256) void band_diag_kernel(T *ab, int hb, int k0, int *flag)
{
    __shared__ T s[NB * NB];
    for (int t = threadIdx.x; t < NB * NB; t += 256) {
        int i = k0 + t / NB, c = k0 + t % NB;
        s[t] = inband(i, c, hb) ? ab[bidx(i, c, hb)] : T(0);
    }
    __syncthreads();
    lds_lu64(s, NB, flag);
    for (int t = threadIdx.x; t < NB * NB; t += 256) {
        int i = k0 + t / NB, c = k0 + t % NB;
        if (inband(i, c, hb)) ab[bidx(i, c, hb)] = s[t];
    }
}

// blockIdx.x < nrt: a 64-row tile of L21; otherwise a 64-column tile of U12.
template <typename T> __global__ __launch_bounds__(256) void band_trsm_kernel(T *ab, int hb, int k0, int n, int nrt)
{
    __shared__ T sd[NB * NB];  // diagonal block (L11 \ U11)
    __shared__ T sx[NB * NB];  // tile
    const int tid = threadIdx.x;
    const int lim = min(n, k0 + NB + hb);
    for (int t = tid; t < NB * NB; t += 256) {
        int i = k0 + t / NB, c = k0 + t % NB;
        sd[t] = inband(i, c, hb) ? ab[bidx(i, c, hb)] : T(0);
    }
    if ((int)blockIdx.x < nrt) {
        const int r0 = k0 + NB + blockIdx.x * NB;
        for (int t = tid; t < NB * NB; t += 256) {
            int i = r0 + t / NB, c = k0 + t % NB;
            sx[t] = (i < lim && inband(i, c, hb)) ? ab[bidx(i, c, hb)] : T(0);
        }
        __syncthreads();
        // X U11 = A21, column by column
        for (int c = 0; c < NB; c++) {
            if (tid < NB) sx[tid * NB + c] /= sd[c * NB + c];
            __syncthreads();
            int w = NB - 1 - c;
            for (int t = tid; t < NB * w; t += 256) {
                int i = t / w, j = c + 1 + t % w;
                sx[i * NB + j] -= sx[i * NB + c] * sd[c * NB + j];
            }
            __syncthreads();
        }
        for (int t = tid; t < NB * NB; t += 256) {
            int i = r0 + t / NB, c = k0 + t % NB;
            if (i < lim && inband(i, c, hb)) ab[bidx(i, c, hb)] = sx[t];
        }
    } else {
        const int c0 = k0 + NB + (blockIdx.x - nrt) * NB;
        for (int t = tid; t < NB * NB; t += 256) {
            int i = k0 + t / NB, c = c0 + t % NB;
            sx[t] = (c < lim && inband(i, c, hb)) ? ab[bidx(i, c, hb)] : T(0);
        }
        __syncthreads();
        // L11 Y = A12 (unit lower), row by row
        for (int r = 0; r < NB - 1; r++) {
            int h = NB - 1 - r;
            for (int t = tid; t < h * NB; t += 256) {
                int i = r + 1 + t / NB, j = t % NB;
                sx[i * NB + j] -= sd[i * NB + r] * sx[r * NB + j];
            }
            __syncthreads();
        }
        for (int t = tid; t < NB * NB; t += 256) {
            int i = k0 + t / NB, c = c0 + t % NB;
            if (c < lim && inband(i, c, hb)) ab[bidx(i, c, hb)] = sx[t];
        }
    }
}

// A22 -= L21 U12 on 64 x 64 tiles of the window [k0+NB, lim)^2, band entries only.
template <typename T> __global__ __launch_bounds__(256) void band_update_kernel(T *ab, int hb, int k0, int n, int nt,
                                                                                 int *flag)
{
    typedef typename mfma<T>::acc_t acc_t;
    // staging for the A ([row][k], padded) and B ([k][col]) chunks; after the
    // k loop the same storage holds the tile of the fused diagonal LU
    __shared__ T smem[NB * (KT + 1) + KT * NB];
    static_assert(NB * (KT + 1) + KT * NB >= NB * (NB + 1), "fused tile must fit the staging buffers");
    T *sa = smem, *sb = smem + NB * (KT + 1), *sc = smem;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
    const int base = k0 + NB;
    const int lim = min(n, k0 + NB + hb);
    const int R0 = base + ti * NB, C0 = base + tj * NB;
    // skip tiles that do not meet the band
    if (R0 - (C0 + NB - 1) > hb || C0 - (R0 + NB - 1) > hb) return;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{};
    for (int kc = 0; kc < NB; kc += KT) {
        for (int t = tid; t < NB * KT; t += 256) {
            int i = t / KT, kk = t % KT;
            int gi = R0 + i, gk = k0 + kc + kk;
            sa[i * (KT + 1) + kk] = (gi < lim && inband(gi, gk, hb)) ? ab[bidx(gi, gk, hb)] : T(0);
        }
        for (int t = tid; t < KT * NB; t += 256) {
            int kk = t / NB, j = t % NB;
            int gk = k0 + kc + kk, gj = C0 + j;
            sb[kk * NB + j] = (gj < lim && inband(gk, gj, hb)) ? ab[bidx(gk, gj, hb)] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KT; kk += 4) {
            T av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; a++) av[a] = sa[(wr + 16 * a + (lane & 15)) * (KT + 1) + kk + (lane >> 4)];
#pragma unroll
            for (int b = 0; b < 2; b++) bv[b] = sb[(kk + (lane >> 4)) * NB + wc + 16 * b + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = mfma<T>::run(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    const bool fuse = (ti == 0 && tj == 0 && base + NB <= n);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = wr + 16 * a + mfma<T>::row(lane, v), j = wc + 16 * b + (lane & 15);
                int gi = R0 + i, gj = C0 + j;
                // rows/columns past the window carry a zero update (masked operands)
                if (gi < n && gj < n && inband(gi, gj, hb)) {
                    T nv = ab[bidx(gi, gj, hb)] - acc[a][b][v];
                    if (fuse) sc[i * (NB + 1) + j] = nv;
                    else ab[bidx(gi, gj, hb)] = nv;
                } else if (fuse) {
                    sc[i * (NB + 1) + j] = T(0);
                }
            }
    if (!fuse) return;
    __syncthreads();
    lds_lu64(sc, NB + 1, flag);
    for (int t = tid; t < NB * NB; t += 256) {
        int i = t / NB, j = t % NB;
        int gi = R0 + i, gj = C0 + j;
        if (inband(gi, gj, hb)) ab[bidx(gi, gj, hb)] = sc[i * (NB + 1) + j];
    }
}

// ---- triangular solves ------------------------------------------------------------
// Forward, panel k0: every workgroup stages L11 in LDS and wave 0 solves
// L11 y_p = b_p (unit lower) with register shuffles; workgroup 0 stores y_p in
// z; then each wave takes rows below (one row per wave, coalesced along the
// row) and subtracts L21(row, :) y_p from b.  b_p is only read and the rows
// below are only written, so the workgroups of one launch never race.
// Launches follow panel order; z holds the forward solution at the end.
template <typename T, typename V>
__global__ __launch_bounds__(256) void band_fwd_kernel(const T *ab, int hb, int k0, int n, V *b, V *z,
                                                      int rows_per_wg)
{
    __shared__ T sl[NB * (NB + 1)];
    __shared__ V sy[NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < NB * NB; t += 256) {
        int i = t / NB, c = t % NB;
        sl[i * (NB + 1) + c] = (c < i && inband(k0 + i, k0 + c, hb)) ? ab[bidx(k0 + i, k0 + c, hb)] : T(0);
    }
    __syncthreads();
    if (wave == 0) {
        V yi = b[k0 + lane];
        for (int c = 0; c < NB - 1; c++) {
            V yc = __shfl(yi, c);
            if (lane > c) yi -= (V)sl[lane * (NB + 1) + c] * yc;
        }
        sy[lane] = yi;
        if (blockIdx.x == 0) z[k0 + lane] = yi;
    }
    __syncthreads();
    const int lim = min(n, k0 + NB + hb);
    const int rbeg = k0 + NB + blockIdx.x * rows_per_wg;
    const int rend = min(lim, rbeg + rows_per_wg);
    const V yl = sy[lane];
    for (int r = rbeg + wave; r < rend; r += 4) {
        V s = V(0);
        int c = k0 + lane;
        if (r - c <= hb) s = (V)ab[bidx(r, c, hb)] * yl;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) b[r] -= s;
    }
}

// Backward, panel k0 (from the last): wave 0 solves U11 x_p = z_p and
// workgroup 0 stores x_p in x; rows above that couple to the panel subtract
// U(row, panel) x_p from z.  Same read/write split as the forward sweep.
template <typename T, typename V>
__global__ __launch_bounds__(256) void band_bwd_kernel(const T *ab, int hb, int k0, V *z, V *x, int rows_per_wg)
{
    __shared__ T su[NB * (NB + 1)];
    __shared__ V sx[NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < NB * NB; t += 256) {
        int i = t / NB, c = t % NB;
        su[i * (NB + 1) + c] = (c >= i && inband(k0 + i, k0 + c, hb)) ? ab[bidx(k0 + i, k0 + c, hb)] : T(0);
    }
    __syncthreads();
    if (wave == 0) {
        V yi = z[k0 + lane];
        for (int c = NB - 1; c >= 0; c--) {
            V xc = __shfl(yi, c) / (V)su[c * (NB + 1) + c];
            if (lane == c) yi = xc;
            else if (lane < c) yi -= (V)su[lane * (NB + 1) + c] * xc;
        }
        sx[lane] = yi;
        if (blockIdx.x == 0) x[k0 + lane] = yi;
    }
    __syncthreads();
    const int top = max(0, k0 - hb);
    const int rbeg = top + blockIdx.x * rows_per_wg;
    const int rend = min(k0, rbeg + rows_per_wg);
    const V xl = sx[lane];
    for (int r = rbeg + wave; r < rend; r += 4) {
        V s = V(0);
        int c = k0 + lane;
        if (c - r <= hb) s = (V)ab[bidx(r, c, hb)] * xl;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) z[r] -= s;
    }
}

}  // namespace

// ---- host -----------------------------------------------------------------------
template <typename T> static int band_factor_t(T *ab, int n, int hb, hipStream_t s, int *flag)
{
    int launches = 0;
    hipLaunchKernelGGL(band_diag_kernel<T>, dim3(1), dim3(256), 0, s, ab, hb, 0, flag);
    launches++;
    for (int k0 = 0; k0 < n; k0 += NB) {
        int lim = std::min(n, k0 + NB + hb);
        int rem = lim - (k0 + NB);
        if (rem <= 0) break;
        int nt = (rem + NB - 1) / NB;
        hipLaunchKernelGGL(band_trsm_kernel<T>, dim3(2 * nt), dim3(256), 0, s, ab, hb, k0, n, nt);
        hipLaunchKernelGGL(band_update_kernel<T>, dim3(nt * nt), dim3(256), 0, s, ab, hb, k0, n, nt, flag);
        launches += 2;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

// x: right-hand side in, solution out; z: n entries of work space.
template <typename T, typename V> static int band_solve_t(const T *ab, int n, int hb, hipStream_t s, V *x, V *z)
{
    const int RPW = 64;  // coupled rows per workgroup
    int launches = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        int lim = std::min(n, k0 + NB + hb);
        int rows = std::max(0, lim - (k0 + NB));
        int grid = std::max(1, (rows + RPW - 1) / RPW);
        hipLaunchKernelGGL((band_fwd_kernel<T, V>), dim3(grid), dim3(256), 0, s, ab, hb, k0, n, x, z, RPW);
        launches++;
    }
    for (int k0 = n - NB; k0 >= 0; k0 -= NB) {
        int rows = k0 - std::max(0, k0 - hb);
        int grid = std::max(1, (rows + RPW - 1) / RPW);
        hipLaunchKernelGGL((band_bwd_kernel<T, V>), dim3(grid), dim3(256), 0, s, ab, hb, k0, z, x, RPW);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

int gf_band_factor(void *ab, int elem_size, int n, int hb, hipStream_t s, int *flag)
{
    if (n % NB != 0) {
        gf_set_error("band factor: order must be a multiple of 64");
        return GF_E_ARG;
    }
    if (elem_size == 8) return band_factor_t<double>((double *)ab, n, hb, s, flag);
    return band_factor_t<float>((float *)ab, n, hb, s, flag);
}

int gf_band_solve(const void *ab, int elem_size, int n, int hb, hipStream_t s, double *x, double *work)
{
    if (elem_size == 8) return band_solve_t<double, double>((const double *)ab, n, hb, s, x, work);
    return band_solve_t<float, double>((const float *)ab, n, hb, s, x, work);
}

int gf_band_pad(int n) { return (n + NB - 1) / NB * NB; }
