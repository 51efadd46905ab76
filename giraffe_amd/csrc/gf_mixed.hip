// Mixed-precision tangent solve: FP32 factor of the equilibrated tangent,
// FP64 iterative refinement against the FP64 tangent.  The refinement is a
// sequence of enqueued sweeps; the host reads each sweep's norms when its
// event has completed and enqueues the next sweep, the update, or the FP64
// fallback (gf_refine_step), so a server's host thread never waits on it.
//
//   D = diag(|K_ii|)^-1/2,  K32 = fl32(D K D)  (factored in FP32)
//   x = 0, s = r
//   repeat: w = K32^-1 (D s) ;  x += D w ;  s = r - K x   (FP64)
// until the correction |D w| is below tol |x| or stops shrinking; a stagnating
// refinement falls back to the FP64 factorisation, so the Newton increment is
// always the FP64 solution of the FP64 tangent.  The residual r, the tangent K
// and every update stay FP64; only the factor is FP32.
#include "gf_internal.h"

#include <algorithm>
#include <cstring>

namespace {

__global__ void diag_scale_kernel(gf_gmat g, int n, int npad, double *D)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double v;
    if (g.layout == GM_BTD) {
        int bi = i / g.m, lr = i - bi * g.m;
        v = g.B[((size_t)bi * g.m + lr) * g.m + lr];
    } else {
        v = g.ab[(size_t)i * (size_t)g.ld + (size_t)(i + g.hb)];
    }
    double a = fabs(v);
    D[i] = (i < n && a > 0.0 && a < 1e300) ? 1.0 / sqrt(a) : 1.0;
}

// BTD blocks (A, B, C) -> scaled FP32 copies.
__global__ void convert_btd_kernel(const double *A, const double *B, const double *C, float *A32, float *B32,
                                   float *C32, int m, int nblk, const double *D)
{
    size_t bs = (size_t)m * m;
    size_t total = (size_t)nblk * bs;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t blk = t / bs;
        int lr = (int)((t % bs) / m), lc = (int)(t % m);
        int R = (int)blk * m + lr;
        double dr = D[R];
        double db = D[blk * m + lc];
        B32[t] = (float)(dr * B[t] * db);
        A32[t] = blk > 0 ? (float)(dr * A[t] * D[(blk - 1) * m + lc]) : 0.0f;
        C32[t] = blk + 1 < (size_t)nblk ? (float)(dr * C[t] * D[(blk + 1) * m + lc]) : 0.0f;
    }
}

// Band storage -> scaled FP32 copy (same skewed layout, row stride ld).
__global__ void convert_band_kernel(const double *ab, float *ab32, int npad, int hb, int ld, const double *D)
{
    size_t w = (size_t)(ld + 1);
    size_t total = (size_t)npad * w;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(t / w);
        int c = i - hb + (int)(t % w);
        double v = 0.0;
        if (c >= 0 && c < npad) v = D[i] * ab[t] * D[c];
        ab32[t] = (float)v;
    }
}

// s = r - K x, one wavefront per row.
__global__ void residual_kernel(gf_gmat g, int npad, const double *r, const double *x, double *s)
{
    const int lane = threadIdx.x & 63;
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (row >= npad) return;
    double acc = 0.0;
    if (g.layout == GM_BTD) {
        int m = g.m, bi = row / m, lr = row - bi * m;
        int nblk = npad / m;
        for (int t = lane; t < 3 * m; t += 64) {
            int which = t / m, c = t - which * m;
            int bj = bi + which - 1;
            if (bj < 0 || bj >= nblk) continue;
            const double *base = which == 0 ? g.A : (which == 1 ? g.B : g.C);
            acc += base[((size_t)bi * m + lr) * m + c] * x[(size_t)bj * m + c];
        }
    } else {
        int hb = g.hb;
        for (int k = lane; k <= 2 * hb; k += 64) {
            int c = row - hb + k;
            if (c < 0 || c >= npad) continue;
            acc += g.ab[(size_t)row * (size_t)g.ld + (size_t)(c + hb)] * x[c];
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) s[row] = r[row] - acc;
}

// s = r - K x over the tangent's structure (ptr, cols), one wavefront per row;
// the padding rows carry the unit diagonal.
__global__ void residual_csr_kernel(const double *ab, int ld, int hb, const int *ptr, const int *cols, int nrows,
                                    int npad, const double *r, const double *x, double *s)
{
    const int lane = threadIdx.x & 63;
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (row >= npad) return;
    if (row >= nrows) {
        if (lane == 0) s[row] = r[row] - x[row];
        return;
    }
    const double *base = ab + (size_t)row * (size_t)ld + (size_t)hb;
    double acc = 0.0;
    for (int k = ptr[row] + lane; k < ptr[row + 1]; k += 64) {
        const int c = cols[k];
        acc += base[c] * x[c];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) s[row] = r[row] - acc;
}
__global__ void scale_kernel(int n, const double *D, const double *v, double *w)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = D[i] * v[i];
}

// x += D w; out[0] = max|D w|, out[1] = max|x| (as bit patterns of non-negative doubles)
__global__ void correct_kernel(int n, const double *D, const double *w, double *x, unsigned long long *out)
{
    double dmax = 0.0, xmax = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double dx = D[i] * w[i];
        double xn = x[i] + dx;
        x[i] = xn;
        double a = fabs(dx);
        dmax = (a == a) ? fmax(dmax, a) : 1e308;
        xmax = fmax(xmax, fabs(xn));
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, off));
        xmax = fmax(xmax, __shfl_xor(xmax, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out + 0, (unsigned long long)__double_as_longlong(dmax));
        atomicMax(out + 1, (unsigned long long)__double_as_longlong(xmax));
    }
}

inline int grid_for(size_t n, int block, int cap = 4096)
{
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

}  // namespace

int gf_mixed_alloc(gf_model *m)
{
    size_t n = (size_t)m->npad;
    GF_TRY(gf_pool_alloc((void **)&m->d_scale, n * 8));
    GF_TRY(gf_pool_alloc((void **)&m->d_res, n * 8));
    GF_TRY(gf_pool_alloc((void **)&m->d_w, n * 8));
    if (m->layout == GM_BTD) {
        int st = gf_btd_alloc(m->btd32, m->btd.m, m->btd.nblk, 4);
        if (st) return st;
    } else if (m->fronts_on) {
        GF_TRY(gf_pool_alloc((void **)&m->d_front32, (size_t)std::max(1LL, m->fronts.f_elems) * 4));
        GF_TRY(gf_pool_alloc((void **)&m->d_front32_e, (size_t)std::max(1LL, m->fronts.e_elems) * 4));
    } else {
        GF_TRY(gf_pool_alloc((void **)&m->d_band32, m->band_elems * 4));
        GF_TRY(gf_pool_alloc((void **)&m->d_band32_inv, gf_band_inv_elems(m->npad) * 4));
    }
    return GF_OK;
}

void gf_mixed_free(gf_model *m)
{
    void *p[] = {m->d_scale, m->d_res, m->d_w, m->d_band32, m->d_band32_inv, m->d_front32, m->d_front32_e};
    for (void *q : p) gf_pool_free(q);
    if (m->btd32.A) gf_btd_free(m->btd32);
}

// Factor the equilibrated FP32 copy; returns launches or a negative status.
int gf_mixed_factor(gf_model *m)
{
    gf_gmat g = gf_model_gmat(m);
    int npad = m->npad;
    hipLaunchKernelGGL(diag_scale_kernel, dim3((npad + 255) / 256), dim3(256), 0, m->stream, g, m->n_free, npad,
                       m->d_scale);
    if (m->layout == GM_BTD) {
        size_t total = (size_t)m->btd.nblk * m->btd.bsz;
        hipLaunchKernelGGL(convert_btd_kernel, dim3(grid_for(total, 256)), dim3(256), 0, m->stream,
                           (const double *)m->btd.A, (const double *)m->btd.B, (const double *)m->btd.C,
                           (float *)m->btd32.A, (float *)m->btd32.B, (float *)m->btd32.C, m->btd.m, m->btd.nblk,
                           m->d_scale);
        GF_HIP(hipGetLastError());
        int l = gf_btd_factor(m->btd32, m->stream, m->d_flag);
        return l < 0 ? l : l + 2;
    }
    if (m->fronts_on) {
        // the fronts take the scaled entries straight from the FP64 band
        GF_HIP(hipGetLastError());
        const int l = gf_front_factor(m->fronts, 4, m->d_band, m->band_ld, m->half_band, m->d_scale, m->d_front32,
                                      m->d_front32_e, m->stream, m->d_flag, nullptr, nullptr);
        return l < 0 ? l : l + 1;
    }
    hipLaunchKernelGGL(convert_band_kernel, dim3(grid_for(m->band_elems, 256)), dim3(256), 0, m->stream, m->d_band,
                       m->d_band32, npad, m->half_band, m->band_ld, m->d_scale);
    GF_HIP(hipGetLastError());
    int l = gf_band_factor(m->d_band32, 4, npad, m->half_band, m->stream, m->d_flag, m->d_band32_inv);
    return l < 0 ? l : l + 2;
}

// Refinement of K x = r (r = d_r, x = d_x), enqueue-only pieces (the caller
// decides between sweeps from the norms each sweep copies back, without
// blocking the host: gf_refine_step).

// x = 0, s = r.
int gf_mixed_refine_begin(gf_model *m)
{
    GF_HIP(hipMemcpyAsync(m->d_res, m->d_r, sizeof(double) * m->npad, hipMemcpyDeviceToDevice, m->stream));
    return gf_launch_fill(m->stream, m->d_x, (size_t)m->npad, 0.0);
}

// One sweep: w = K32^-1 (D s), x += D w, s = r - K x; max|D w| and max|x| to
// d_scal[GF_SCAL_REF..+1] and, with the factorisation flag, to the host mirror.
// Returns launches or a negative status.
int gf_mixed_sweep(gf_model *m)
{
    const int npad = m->npad;
    gf_gmat g = gf_model_gmat(m);
    hipLaunchKernelGGL(scale_kernel, dim3((npad + 255) / 256), dim3(256), 0, m->stream, npad, m->d_scale, m->d_res,
                       m->d_w);
    // BTD: solution in d_swork (d_w is reduced in place); band: in place in d_w
    const int l = m->layout == GM_BTD ? gf_btd_solve(m->btd32, m->stream, m->d_w, m->d_swork, npad, m->d_w)
                  : m->fronts_on    ? gf_front_solve(m->fronts, 4, m->d_front32, m->d_front32_e, m->d_w, m->d_w,
                                                     m->d_front_v, m->stream, false)
                                    : gf_band_solve(m->d_band32, 4, npad, m->half_band, m->stream, m->d_w,
                                                    m->d_swork, m->d_band32_inv);
    if (l < 0) return l;
    const double *sol = m->layout == GM_BTD ? m->d_swork : m->d_w;
    GF_TRY(gf_launch_fill_bytes(m->stream, m->d_scal + GF_SCAL_REF, 2 * sizeof(double)));
    hipLaunchKernelGGL(correct_kernel, dim3(grid_for(npad, 256, 512)), dim3(256), 0, m->stream, npad, m->d_scale, sol,
                       m->d_x, (unsigned long long *)(m->d_scal + GF_SCAL_REF));
    if (m->fronts_on)
        hipLaunchKernelGGL(residual_csr_kernel, dim3((npad * 64 + 255) / 256), dim3(256), 0, m->stream, m->d_band,
                           m->band_ld, m->half_band, m->d_csr_ptr, m->d_csr_col, m->n_free, npad, m->d_r, m->d_x,
                           m->d_res);
    else
        hipLaunchKernelGGL(residual_kernel, dim3((npad * 64 + 255) / 256), dim3(256), 0, m->stream, g, npad, m->d_r,
                           m->d_x, m->d_res);
    GF_HIP(hipGetLastError());
    GF_HIP(hipMemcpyAsync(m->h_scal + GF_SCAL_FLAG, m->d_scal + GF_SCAL_FLAG, sizeof(double), hipMemcpyDeviceToHost,
                          m->stream));
    GF_HIP(hipMemcpyAsync(m->h_scal + GF_SCAL_REF, m->d_scal + GF_SCAL_REF, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, m->stream));
    return l + 3;
}

// The FP64 factorisation and solve of the FP64 tangent (x = d_x).
int gf_mixed_fallback(gf_model *m)
{
    const int npad = m->npad;
    GF_TRY(gf_launch_fill_bytes(m->stream, m->d_flag, 4));
    if (m->layout == GM_BTD) return gf_btd_factor_solve(m->btd, m->stream, m->d_flag, m->d_r, m->d_x, npad, m->d_res);
    if (m->fronts_on) {
        const int l = gf_front_factor(m->fronts, 8, m->d_band, m->band_ld, m->half_band, nullptr, m->d_front,
                                      m->d_front_e, m->stream, m->d_flag, m->d_r, m->d_front_v);
        if (l < 0) return l;
        const int l2 =
            gf_front_solve(m->fronts, 8, m->d_front, m->d_front_e, m->d_r, m->d_x, m->d_front_v, m->stream, true);
        return l2 < 0 ? l2 : l + l2;
    }
    const int l = gf_band_factor(m->d_band, 8, npad, m->half_band, m->stream, m->d_flag, m->d_band_inv);
    if (l < 0) return l;
    GF_HIP(hipMemcpyAsync(m->d_x, m->d_r, sizeof(double) * npad, hipMemcpyDeviceToDevice, m->stream));
    const int l2 = gf_band_solve(m->d_band, 8, npad, m->half_band, m->stream, m->d_x, m->d_swork, m->d_band_inv);
    return l2 < 0 ? l2 : l + l2;
}
