// Block-tridiagonal cyclic reduction on gfx950: the tangent solve for narrow
// bands (replaces sparsesystem -> MKL PARDISO, src/SparseMatrix.cpp:74-166).
//
// A band matrix of half-bandwidth b, cut into diagonal blocks of size m >= b,
// is block-tridiagonal: rows of block i couple only to blocks i-1 (A_i), i (B_i)
// and i+1 (C_i).  Cyclic reduction removes the even positions of the active
// block rows at each level (all independent), keeping the odd ones:
//   elimination row e:  B_e = P L U,  X_e = B_e^-1 A_e,  Y_e = B_e^-1 C_e
//   kept row k (neighbours l = p-1, r = p+1):
//       B_k -= A_k Y_l + C_k X_r,   A_k <- -A_k X_l,   C_k <- -C_k Y_r
// until one block row is left.  The right-hand side follows the same levels
// (z_e = B_e^-1 f_e, f_k -= A_k z_l + C_k z_r with the pre-update A_k, C_k)
// and back-substitution runs the levels in reverse (x_e = z_e - X_e x_l - Y_e x_r).
// Block LU uses partial pivoting inside each m x m diagonal block.
// Active row p at level L is global block row (p+1)*2^L - 1.
//
// One 64-lane wavefront (one workgroup) per block row, blocks staged in LDS,
// m <= 32.
#include "gf_internal.h"
#include <cmath>

namespace {

constexpr int MMAX = 32;

__device__ __forceinline__ int lvl_row(int p, int L) { return ((p + 1) << L) - 1; }

template <typename T> __device__ __forceinline__ bool bad(T v) { return !(v == v) || fabs((double)v) >= 1e300 || v == T(0); }

// LU with partial pivoting of sB (m x m, row-major) in LDS; piv -> global.
template <typename T>
__device__ void wave_lu(T *sB, int m, int *piv_out, int *flag, int *spiv)
{
    const int lane = threadIdx.x;
    for (int k = 0; k < m; k++) {
        // pivot search in column k (m <= 32 <= 64 lanes)
        T v = (lane >= k && lane < m) ? (T)fabs((double)sB[lane * m + k]) : T(-1);
        int idx = lane;
        for (int off = 32; off > 0; off >>= 1) {
            T ov = __shfl_xor(v, off);
            int oi = __shfl_xor(idx, off);
            if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
        }
        int p = idx;
        if (lane == 0) spiv[k] = p;
        if (p != k && lane < m) {
            T t = sB[k * m + lane];
            sB[k * m + lane] = sB[p * m + lane];
            sB[p * m + lane] = t;
        }
        __syncthreads();
        T pivot = sB[k * m + k];
        if (bad(pivot)) {
            if (lane == 0) atomicExch(flag, 1);
            pivot = T(1);
        }
        if (lane > k && lane < m) sB[lane * m + k] /= pivot;
        __syncthreads();
        int w = m - k - 1;
        for (int t = lane; t < w * w; t += 64) {
            int i = k + 1 + t / w, j = k + 1 + t % w;
            sB[i * m + j] -= sB[i * m + k] * sB[k * m + j];
        }
        __syncthreads();
    }
    if (lane < m) piv_out[lane] = spiv[lane];
}

// Solve (P L U) y = rhs for the column `col` held by this lane (strided by ld in LDS).
template <typename T>
__device__ __forceinline__ void lane_lu_solve(const T *sB, const int *spiv, int m, T *c, int ld)
{
    for (int k = 0; k < m; k++) {
        int p = spiv[k];
        if (p != k) { T t = c[k * ld]; c[k * ld] = c[p * ld]; c[p * ld] = t; }
    }
    for (int i = 1; i < m; i++) {
        T s = c[i * ld];
        for (int k = 0; k < i; k++) s -= sB[i * m + k] * c[k * ld];
        c[i * ld] = s;
    }
    for (int i = m - 1; i >= 0; i--) {
        T s = c[i * ld];
        for (int k = i + 1; k < m; k++) s -= sB[i * m + k] * c[k * ld];
        c[i * ld] = s / sB[i * m + i];
    }
}

template <typename T>
__global__ __launch_bounds__(64) void btd_elim_kernel(int L, int nL, int m, T *A, T *B, T *C, int *piv, int *flag)
{
    __shared__ T sB[MMAX * MMAX];
    __shared__ T sR[MMAX * 2 * MMAX];  // [m][2m]: X | Y right-hand sides
    __shared__ int spiv[MMAX];
    const int p = 2 * blockIdx.x;      // even positions are eliminated
    if (p >= nL) return;
    const int e = lvl_row(p, L);
    const size_t bs = (size_t)m * m;
    const int lane = threadIdx.x;
    const bool hasL = p >= 1, hasR = p + 1 < nL;
    for (int t = lane; t < m * m; t += 64) {
        int i = t / m, j = t % m;
        sB[t] = B[e * bs + t];
        sR[i * 2 * m + j] = hasL ? A[e * bs + t] : T(0);
        sR[i * 2 * m + m + j] = hasR ? C[e * bs + t] : T(0);
    }
    __syncthreads();
    wave_lu(sB, m, piv + (size_t)e * m, flag, spiv);
    if (lane < 2 * m) lane_lu_solve(sB, spiv, m, sR + lane, 2 * m);
    __syncthreads();
    for (int t = lane; t < m * m; t += 64) {
        int i = t / m, j = t % m;
        B[e * bs + t] = sB[t];
        A[e * bs + t] = sR[i * 2 * m + j];        // X_e
        C[e * bs + t] = sR[i * 2 * m + m + j];    // Y_e
    }
}

template <typename T>
__global__ __launch_bounds__(64) void btd_kept_kernel(int L, int nL, int m, T *A, T *B, T *C, T *Lsv, T *Rsv)
{
    __shared__ T sAk[MMAX * MMAX], sCk[MMAX * MMAX], sXl[MMAX * MMAX], sYl[MMAX * MMAX], sXr[MMAX * MMAX],
        sYr[MMAX * MMAX];
    const int q = blockIdx.x;
    const int p = 2 * q + 1;
    if (p >= nL) return;
    const int k = lvl_row(p, L), l = lvl_row(p - 1, L);
    const bool hasR = p + 1 < nL;
    const int r = hasR ? lvl_row(p + 1, L) : 0;
    const size_t bs = (size_t)m * m;
    const int lane = threadIdx.x;
    for (int t = lane; t < m * m; t += 64) {
        T ak = A[k * bs + t], ck = C[k * bs + t];
        sAk[t] = ak;
        sCk[t] = ck;
        Lsv[(size_t)q * bs + t] = ak;
        Rsv[(size_t)q * bs + t] = ck;
        sXl[t] = A[l * bs + t];
        sYl[t] = C[l * bs + t];
        sXr[t] = hasR ? A[r * bs + t] : T(0);
        sYr[t] = hasR ? C[r * bs + t] : T(0);
    }
    __syncthreads();
    for (int t = lane; t < m * m; t += 64) {
        int i = t / m, j = t % m;
        T b = T(0), a = T(0), c = T(0);
        for (int s = 0; s < m; s++) {
            T aks = sAk[i * m + s], cks = sCk[i * m + s];
            b += aks * sYl[s * m + j] + cks * sXr[s * m + j];
            a += aks * sXl[s * m + j];
            c += cks * sYr[s * m + j];
        }
        B[k * bs + t] -= b;
        A[k * bs + t] = -a;
        C[k * bs + t] = -c;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void btd_top_kernel(int L, int m, T *B, int *piv, int *flag)
{
    __shared__ T sB[MMAX * MMAX];
    __shared__ int spiv[MMAX];
    const int e = lvl_row(0, L);
    const size_t bs = (size_t)m * m;
    for (int t = threadIdx.x; t < m * m; t += 64) sB[t] = B[e * bs + t];
    __syncthreads();
    wave_lu(sB, m, piv + (size_t)e * m, flag, spiv);
    for (int t = threadIdx.x; t < m * m; t += 64) B[e * bs + t] = sB[t];
}

// ---- solve ------------------------------------------------------------------
// z_e = B_e^-1 f_e for the eliminated rows of level L (in place in x).
template <typename T, typename V>
__global__ __launch_bounds__(64) void btd_fwd_elim_kernel(int L, int nL, int m, const T *B, const int *piv, V *x, int stride2)
{
    __shared__ T sB[MMAX * MMAX];
    __shared__ T sc[MMAX];
    __shared__ int spiv[MMAX];
    const int p = stride2 * blockIdx.x;
    if (p >= nL) return;
    const int e = lvl_row(p, L);
    const size_t bs = (size_t)m * m;
    const int lane = threadIdx.x;
    for (int t = lane; t < m * m; t += 64) sB[t] = B[e * bs + t];
    if (lane < m) { sc[lane] = (T)x[(size_t)e * m + lane]; spiv[lane] = piv[(size_t)e * m + lane]; }
    __syncthreads();
    if (lane == 0) lane_lu_solve(sB, spiv, m, sc, 1);
    __syncthreads();
    if (lane < m) x[(size_t)e * m + lane] = (V)sc[lane];
}

// f_k -= Lsv z_l + Rsv z_r for the kept rows of level L.
template <typename T, typename V>
__global__ __launch_bounds__(64) void btd_fwd_kept_kernel(int L, int nL, int m, const T *Lsv, const T *Rsv, V *x)
{
    const int q = blockIdx.x;
    const int p = 2 * q + 1;
    if (p >= nL) return;
    const int k = lvl_row(p, L), l = lvl_row(p - 1, L);
    const bool hasR = p + 1 < nL;
    const int r = hasR ? lvl_row(p + 1, L) : 0;
    const size_t bs = (size_t)m * m;
    const int i = threadIdx.x;
    if (i >= m) return;
    V s = V(0);
    for (int j = 0; j < m; j++) {
        s += (V)Lsv[(size_t)q * bs + i * m + j] * x[(size_t)l * m + j];
        if (hasR) s += (V)Rsv[(size_t)q * bs + i * m + j] * x[(size_t)r * m + j];
    }
    x[(size_t)k * m + i] -= s;
}

// x_e = z_e - X_e x_l - Y_e x_r for the eliminated rows of level L.
template <typename T, typename V>
__global__ __launch_bounds__(64) void btd_bwd_kernel(int L, int nL, int m, const T *A, const T *C, V *x)
{
    __shared__ V sl[MMAX], sr[MMAX];
    const int p = 2 * blockIdx.x;
    if (p >= nL) return;
    const int e = lvl_row(p, L);
    const bool hasL = p >= 1, hasR = p + 1 < nL;
    const int l = hasL ? lvl_row(p - 1, L) : 0, r = hasR ? lvl_row(p + 1, L) : 0;
    const size_t bs = (size_t)m * m;
    const int i = threadIdx.x;
    if (i < m) {
        sl[i] = hasL ? x[(size_t)l * m + i] : V(0);
        sr[i] = hasR ? x[(size_t)r * m + i] : V(0);
    }
    __syncthreads();
    if (i >= m) return;
    V s = x[(size_t)e * m + i];
    for (int j = 0; j < m; j++) s -= (V)A[e * bs + i * m + j] * sl[j] + (V)C[e * bs + i * m + j] * sr[j];
    x[(size_t)e * m + i] = s;
}

}  // namespace

int gf_btd_alloc(gf_btd &b, int m, int nblk, int elem_size)
{
    if (m < 1 || m > MMAX) {
        gf_set_error("block-tridiagonal solver: block size out of range");
        return GF_E_UNSUPPORTED;
    }
    b.m = m;
    b.nblk = nblk;
    b.bsz = (size_t)m * m;
    b.elem_size = elem_size;
    b.lvl_n.clear();
    b.lvl_save.clear();
    int n = nblk, L = 0;
    size_t save = 0;
    while (n > 1) {
        b.lvl_n.push_back(n);
        b.lvl_save.push_back(save);
        save += n / 2;
        n /= 2;
        L++;
    }
    b.levels = L;
    size_t blocks = (size_t)nblk * b.bsz * elem_size;
    size_t sv = (save ? save : 1) * b.bsz * elem_size;
    GF_HIP(hipMalloc(&b.A, blocks));
    GF_HIP(hipMalloc(&b.B, blocks));
    GF_HIP(hipMalloc(&b.C, blocks));
    GF_HIP(hipMalloc(&b.Lsv, sv));
    GF_HIP(hipMalloc(&b.Rsv, sv));
    GF_HIP(hipMalloc(&b.piv, sizeof(int) * (size_t)nblk * m));
    return GF_OK;
}

void gf_btd_free(gf_btd &b)
{
    (void)hipFree(b.A); (void)hipFree(b.B); (void)hipFree(b.C);
    (void)hipFree(b.Lsv); (void)hipFree(b.Rsv); (void)hipFree(b.piv);
    b.A = b.B = b.C = b.Lsv = b.Rsv = nullptr;
    b.piv = nullptr;
}

template <typename T>
static int btd_factor_t(gf_btd &b, hipStream_t s, int *d_flag)
{
    T *A = (T *)b.A, *B = (T *)b.B, *C = (T *)b.C;
    int launches = 0;
    for (int L = 0; L < b.levels; L++) {
        int nL = b.lvl_n[L];
        T *Lsv = (T *)b.Lsv + b.lvl_save[L] * b.bsz, *Rsv = (T *)b.Rsv + b.lvl_save[L] * b.bsz;
        hipLaunchKernelGGL(btd_elim_kernel<T>, dim3((nL + 1) / 2), dim3(64), 0, s, L, nL, b.m, A, B, C, b.piv, d_flag);
        launches++;
        if (nL / 2 > 0) {
            hipLaunchKernelGGL(btd_kept_kernel<T>, dim3(nL / 2), dim3(64), 0, s, L, nL, b.m, A, B, C, Lsv, Rsv);
            launches++;
        }
    }
    hipLaunchKernelGGL(btd_top_kernel<T>, dim3(1), dim3(64), 0, s, b.levels, b.m, B, b.piv, d_flag);
    launches++;
    GF_HIP(hipGetLastError());
    return launches;
}

int gf_btd_factor(gf_btd &b, hipStream_t s, int *d_flag)
{
    if (b.elem_size == 8) return btd_factor_t<double>(b, s, d_flag);
    return btd_factor_t<float>(b, s, d_flag);
}

template <typename T>
static int btd_solve_t(gf_btd &b, hipStream_t s, double *x)
{
    const T *A = (const T *)b.A, *B = (const T *)b.B, *C = (const T *)b.C;
    int launches = 0;
    for (int L = 0; L < b.levels; L++) {
        int nL = b.lvl_n[L];
        const T *Lsv = (const T *)b.Lsv + b.lvl_save[L] * b.bsz, *Rsv = (const T *)b.Rsv + b.lvl_save[L] * b.bsz;
        hipLaunchKernelGGL((btd_fwd_elim_kernel<T, double>), dim3((nL + 1) / 2), dim3(64), 0, s, L, nL, b.m, B, b.piv, x, 2);
        launches++;
        if (nL / 2 > 0) {
            hipLaunchKernelGGL((btd_fwd_kept_kernel<T, double>), dim3(nL / 2), dim3(64), 0, s, L, nL, b.m, Lsv, Rsv, x);
            launches++;
        }
    }
    hipLaunchKernelGGL((btd_fwd_elim_kernel<T, double>), dim3(1), dim3(64), 0, s, b.levels, 1, b.m, B, b.piv, x, 2);
    launches++;
    for (int L = b.levels - 1; L >= 0; L--) {
        int nL = b.lvl_n[L];
        hipLaunchKernelGGL((btd_bwd_kernel<T, double>), dim3((nL + 1) / 2), dim3(64), 0, s, L, nL, b.m, A, C, x);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

// Solves in place: x holds the right-hand side (nblk*m entries) on entry.
int gf_btd_solve(gf_btd &b, hipStream_t s, const double *rhs, double *x, int n)
{
    if (rhs != x) GF_HIP(hipMemcpyAsync(x, rhs, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    if (b.elem_size == 8) return btd_solve_t<double>(b, s, x);
    return btd_solve_t<float>(b, s, x);
}
