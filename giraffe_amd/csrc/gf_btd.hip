// Block-tridiagonal cyclic reduction on gfx950: the tangent solve for narrow
// bands (replaces sparsesystem -> MKL PARDISO, src/SparseMatrix.cpp:74-166).
//
// A band matrix of half-bandwidth b, cut into diagonal blocks of size m >= b,
// is block-tridiagonal: rows of block i couple only to blocks i-1 (A_i), i (B_i)
// and i+1 (C_i).  Cyclic reduction removes the even positions of the active
// block rows at each level (all independent), keeping the odd ones:
//   eliminated row e:  Binv_e = B_e^-1,  X_e = Binv_e A_e,  Y_e = Binv_e C_e
//   kept row k (neighbours l = p-1, r = p+1):
//       B_k -= A_k Y_l + C_k X_r,   A_k <- -A_k X_l,   C_k <- -C_k Y_r
// until one block row (the top) is left.  The right-hand side follows the same
// levels (z_e = Binv_e f_e, f_k -= A_k z_l + C_k z_r with the pre-update A_k,
// C_k) and back-substitution runs the levels in reverse
// (x_e = z_e - X_e x_l - Y_e x_r).  Active row p at level L is global block
// row (p+1)*2^L - 1.
//
// gfx950 mapping:
//  * one workgroup of two wavefronts per kept row and level: each wavefront
//    eliminates one neighbour with an in-register Gauss-Jordan sweep over the
//    augmented [B | I | A | C] (lane = column, 4 MT columns over 64 lanes,
//    threshold partial pivoting with wave-uniform pivot choice), then the two
//    waves form the kept-row products with 16x16x4 MFMA from LDS.  Both kept
//    neighbours of an eliminated row recompute it (one writes it), so a level
//    is one launch and no workgroup waits on another;
//  * the eliminated-row results go to separate arrays (Bi, X, Y), so the
//    redundant readers never see a half-written block and the assembled
//    blocks of eliminated rows stay untouched;
//  * the last levels (few rows) run in one workgroup with barriers between
//    levels, together with the top and their back-substitution;
//  * the right-hand side can ride along with the factorisation (factor_solve),
//    so a Newton step pays one launch per level plus one per back level.
// Blocks are m x m row-major with m <= 32; kernels work on MT = 16 or 32
// (padding rows/columns get identity on B and zeros elsewhere).
#include "gf_internal.h"

#include <algorithm>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA.  A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k; result item v
// of lane l is row 4 v + l / 16 (f64) or 4 (l / 16) + v (f32), column l % 16.
template <typename T> struct mma;
template <> struct mma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct mma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ double rdl(double v, int l)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float rdl(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ int lvl_row(int p, int L) { return ((p + 1) << L) - 1; }

template <typename T> __device__ __forceinline__ bool bad(T v)
{
    double a = fabs((double)v);
    return !(a >= (sizeof(T) == 4 ? 1e-37 : 1e-300)) || a >= 1e300 || !(v == v) || (sizeof(T) == 4 && a >= 3.0e38);
}

template <int MT> struct cfg {
    static constexpr int S = MT / 16;        // augmented columns per lane (4 MT / 64)
    static constexpr int LD = MT + 1;        // LDS leading dimension (odd: no bank conflicts)
    static constexpr int BLK = MT * LD;      // one padded block in LDS
    static constexpr int PAIRS = MT == 16 ? 4 : 1;  // wave pairs of the tail workgroup
};

// Per wave pair LDS: eliminated neighbours [l, r] x [Binv, X, Y], kept row A_k, C_k, z_l, z_r.
template <typename T, int MT> struct PairSmem {
    T e[2][3][cfg<MT>::BLK];
    T k[2][cfg<MT>::BLK];
    T z[2][MT];
};

// Loads the augmented [B | I | A | C] of block row e into the wave's registers
// (column c = lane + 64 s).  hasA / hasC: couplings that exist at this level.
// Every lane loads from a clamped in-block address and selects afterwards, so
// the MT loads of a lane are in flight together (no branch around a load).
template <typename T, int MT>
__device__ __forceinline__ void load_aug(T (&col)[cfg<MT>::S][MT], const T *A, const T *B, const T *C, int e, int m,
                                         bool hasA, bool hasC)
{
    const int lane = threadIdx.x & 63;
    const size_t bs = (size_t)m * m, o = (size_t)e * bs;
#pragma unroll
    for (int s = 0; s < cfg<MT>::S; s++) {
        const int c = lane + 64 * s, part = c / MT, cc = c % MT;
        const T *src = (part == 2 ? A : part == 3 ? C : B) + o + (cc < m ? cc : m - 1);
        const bool live = part == 0 || (part == 2 && hasA) || (part == 3 && hasC);
        T raw[MT];
#pragma unroll
        for (int i = 0; i < MT; i++) raw[i] = src[(size_t)(i < m ? i : m - 1) * m];
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const T diag = i == cc ? T(1) : T(0);
            col[s][i] = part == 1 ? diag : (i < m && cc < m) ? (live ? raw[i] : T(0)) : part == 0 ? diag : T(0);
        }
    }
}

// Gauss-Jordan on the wave-held augmented matrix: [B | I | A | C] -> [I | Binv | X | Y].
// Pivot: the largest |entry| of column k in rows >= k unless |B_kk| >= half of it.
// Columns k >= m are the identity padding (zero elsewhere in rows and columns
// >= m), so the sweep ends after m pivots.
template <typename T, int MT> __device__ void wave_gj(T (&col)[cfg<MT>::S][MT], int *flag, int m)
{
    constexpr int S = cfg<MT>::S;
#pragma unroll
    for (int k = 0; k < MT; k++) {
        if (k >= m) continue;  // wave-uniform; keeps the loop unrolled
        T mk[MT];
#pragma unroll
        for (int i = 0; i < MT; i++) mk[i] = rdl(col[0][i], k);  // column k sits in lane k, slot 0
        T amax = T(0);
#pragma unroll
        for (int i = k; i < MT; i++) amax = fmax(amax, fabs(mk[i]));
        if (fabs(mk[k]) < T(0.5) * amax) {  // wave-uniform; rarely taken
            int p = k;
            T best = T(0);
#pragma unroll
            for (int i = k; i < MT; i++) {
                const T a = fabs(mk[i]);
                if (a > best) { best = a; p = i; }
            }
#pragma unroll
            for (int s = 0; s < S; s++) {
                T ck = col[s][k], cp = ck;
#pragma unroll
                for (int i = k + 1; i < MT; i++) cp = i == p ? col[s][i] : cp;
#pragma unroll
                for (int i = k + 1; i < MT; i++) col[s][i] = i == p ? ck : col[s][i];
                col[s][k] = cp;
            }
            T tk = mk[k], tp = tk;
#pragma unroll
            for (int i = k + 1; i < MT; i++) tp = i == p ? mk[i] : tp;
#pragma unroll
            for (int i = k + 1; i < MT; i++) mk[i] = i == p ? tk : mk[i];
            mk[k] = tp;
        }
        T piv = mk[k];
        if (bad(piv)) {
            if ((threadIdx.x & 63) == 0) atomicExch(flag, 1);
            piv = T(1);
        }
        const T inv = pivot_rcp(piv);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const T rk = col[s][k] * inv;
#pragma unroll
            for (int i = 0; i < MT; i++)
                if (i != k) col[s][i] -= mk[i] * rk;
            col[s][k] = rk;
        }
    }
}

// Stores Binv, X, Y of the wave's augmented matrix into LDS (rows padded to LD)
// and, when `to_global`, into Bi/X/Y of block row e.
template <typename T, int MT>
__device__ __forceinline__ void store_elim(const T (&col)[cfg<MT>::S][MT], T *sE, T *Bi, T *X, T *Y, int e, int m,
                                           bool to_global, bool keep_bi)
{
    const int lane = threadIdx.x & 63;
    const size_t bs = (size_t)m * m, o = (size_t)e * bs;
#pragma unroll
    for (int s = 0; s < cfg<MT>::S; s++) {
        const int c = lane + 64 * s, part = c / MT, cc = c % MT;
        if (part == 0) continue;
        T *dst = sE + (part - 1) * cfg<MT>::BLK;
        T *g = part == 1 ? Bi : part == 2 ? X : Y;
#pragma unroll
        for (int i = 0; i < MT; i++) {
            dst[i * cfg<MT>::LD + cc] = col[s][i];
            if (to_global && (part != 1 || keep_bi) && i < m && cc < m) g[o + (size_t)i * m + cc] = col[s][i];
        }
    }
}

// acc (MT/16 x MT/16 tiles) += P (LDS, LD) * Q (LDS, LD), MT x MT x MT, one wavefront.
template <typename T, int MT>
__device__ __forceinline__ void mma_acc(typename mma<T>::acc_t (&acc)[MT / 16][MT / 16], const T *P, const T *Q)
{
    const int lane = threadIdx.x & 63;
    constexpr int LD = cfg<MT>::LD;
#pragma unroll
    for (int k0 = 0; k0 < MT; k0 += 4) {
        T a[MT / 16], b[MT / 16];
#pragma unroll
        for (int t = 0; t < MT / 16; t++) {
            a[t] = P[(16 * t + (lane & 15)) * LD + k0 + (lane >> 4)];
            b[t] = Q[(k0 + (lane >> 4)) * LD + 16 * t + (lane & 15)];
        }
#pragma unroll
        for (int ti = 0; ti < MT / 16; ti++)
#pragma unroll
            for (int tj = 0; tj < MT / 16; tj++) acc[ti][tj] = mma<T>::run(a[ti], b[tj], acc[ti][tj]);
    }
}

// dst (global m x m block) = acc.
template <typename T, int MT>
__device__ __forceinline__ void store_acc(const typename mma<T>::acc_t (&acc)[MT / 16][MT / 16], T *dst, int m)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int ti = 0; ti < MT / 16; ti++)
#pragma unroll
        for (int tj = 0; tj < MT / 16; tj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = 16 * ti + mma<T>::row(lane, v), j = 16 * tj + (lane & 15);
                if (i < m && j < m) dst[(size_t)i * m + j] = acc[ti][tj][v];
            }
}

// dst (global m x m block) = -acc.
template <typename T, int MT>
__device__ __forceinline__ void store_neg(const typename mma<T>::acc_t (&acc)[MT / 16][MT / 16], T *dst, int m)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int ti = 0; ti < MT / 16; ti++)
#pragma unroll
        for (int tj = 0; tj < MT / 16; tj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = 16 * ti + mma<T>::row(lane, v), j = 16 * tj + (lane & 15);
                if (i < m && j < m) dst[(size_t)i * m + j] = -acc[ti][tj][v];
            }
}

struct Fac {
    void *A, *B, *C, *Bi, *X, *Y;
    int m;
    int *flag;
};

// One kept row q of level L (active rows nL), processed by the wave pair `wp`
// (threads 128 wp .. 128 wp + 127 of the workgroup).  Every thread of the
// workgroup must call it the same number of times (it holds barriers); pairs
// with nothing to do pass active = false.
// f: the right-hand side being reduced (null: factor only); zout: z of the
// eliminated rows.
// Every global read of the step (the eliminated neighbour's augmented row, the
// kept row's couplings and diagonal block, both right-hand-side pieces) is
// issued before the Gauss-Jordan sweep, so a step pays one memory round trip.
template <typename T, int MT>
__device__ void fac_step(const Fac &F, PairSmem<T, MT> &sm, bool active, int L, int nL, int q, T *Lsv, T *Rsv,
                         double *f, double *zout)
{
    typedef typename mma<T>::acc_t acc_t;
    constexpr int NT = MT / 16, KV = MT * MT / 64;
    T *A = (T *)F.A, *B = (T *)F.B, *C = (T *)F.C;
    const int m = F.m;
    const size_t bs = (size_t)m * m;
    const int lane = threadIdx.x & 63, wv = (threadIdx.x >> 6) & 1;
    const int p = 2 * q + 1;
    const bool hasR = active && p + 1 < nL;
    const int k = lvl_row(p, L), l = lvl_row(p - 1, L), r = hasR ? lvl_row(p + 1, L) : 0;
    const bool doE = active && (wv == 0 || hasR);
    const int e = wv == 0 ? l : r, pe = wv == 0 ? p - 1 : p + 1;
    const bool own = wv == 0 || p + 2 == nL;  // the right neighbour is written here only if it is the last row
    // B^-1 and the kept-row couplings are stored only for later solve-only sweeps;
    // a factorisation that reduces its right-hand side (f) needs X and Y alone.
    const bool keep = f == nullptr;

    // ---- reads
    // (unconditional loads from clamped in-block addresses, selected afterwards;
    // k, e and q are valid rows also for inactive pairs)
    T kv[KV];  // kept-row coupling (wave 0: A_k, wave 1: C_k), element lane + 64 j
    {
        const T *src = (wv == 0 ? A : C) + (size_t)k * bs;
        T raw[KV];
#pragma unroll
        for (int jv = 0; jv < KV; jv++) {
            const int t = lane + 64 * jv, i = t / MT, jj = t % MT;
            raw[jv] = src[(size_t)(i < m ? i : m - 1) * m + (jj < m ? jj : m - 1)];
        }
#pragma unroll
        for (int jv = 0; jv < KV; jv++) {
            const int t = lane + 64 * jv, i = t / MT, jj = t % MT;
            kv[jv] = active && i < m && jj < m ? raw[jv] : T(0);
        }
    }
    T bk[NT][NT][4];  // wave 0: B_k at its MFMA result positions
#pragma unroll
    for (int ti = 0; ti < NT; ti++)
#pragma unroll
        for (int tj = 0; tj < NT; tj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * ti + mma<T>::row(lane, v), jj = 16 * tj + (lane & 15);
                bk[ti][tj][v] = B[(size_t)k * bs + (size_t)(i < m ? i : m - 1) * m + (jj < m ? jj : m - 1)];
            }
#pragma unroll
    for (int ti = 0; ti < NT; ti++)
#pragma unroll
        for (int tj = 0; tj < NT; tj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * ti + mma<T>::row(lane, v), jj = 16 * tj + (lane & 15);
                bk[ti][tj][v] = active && wv == 0 && i < m && jj < m ? bk[ti][tj][v] : T(0);
            }
    // right-hand side pieces: f_e (staged for z_e, read for j < m only) and f_k
    // (used by wave 0's lanes < m); lanes >= m hold a duplicate entry
    double fe = 0.0, fk = 0.0;
    if (f) {
        const int li = lane < m ? lane : m - 1;
        fe = f[(size_t)e * m + li];
        fk = f[(size_t)k * m + li];
    }
    T col[cfg<MT>::S][MT];  // (loaded by every wave: e is a valid row also when it is not eliminated here)
    load_aug<T, MT>(col, A, B, C, e, m, pe >= 1, pe + 1 < nL);

    // ---- kept-row couplings to LDS, saved for later solves
    if (active) {
        T *sv = (wv == 0 ? Lsv : Rsv) + (size_t)q * bs;
        T *dst = sm.k[wv];
#pragma unroll
        for (int jv = 0; jv < KV; jv++) {
            const int t = lane + 64 * jv, i = t / MT, jj = t % MT;
            dst[i * cfg<MT>::LD + jj] = kv[jv];
            if (keep && i < m && jj < m) sv[(size_t)i * m + jj] = kv[jv];
        }
        if (f && lane < MT) sm.z[wv][lane] = T(fe);
    }
    // ---- eliminated neighbour
    if (doE) {
        wave_gj<T, MT>(col, F.flag, m);
        store_elim<T, MT>(col, &sm.e[wv][0][0], (T *)F.Bi, (T *)F.X, (T *)F.Y, e, m, own, keep);
    } else if (active) {
        // no right neighbour: zero X_r, Y_r so the products below need no special case
        for (int t = lane; t < 3 * cfg<MT>::BLK; t += 64) (&sm.e[1][0][0])[t] = T(0);
    }
    __syncthreads();
    T zl = T(0);
    if (active && f && lane < MT) {
        // z of the eliminated neighbour: Binv_e f_e (f_e staged in sm.z)
        if (doE && lane < m) {
            // (the padding entries of B^-1 and z are zero: all MT terms are exact)
#pragma unroll
            for (int j = 0; j < MT; j++) zl += sm.e[wv][0][lane * cfg<MT>::LD + j] * sm.z[wv][j];
            if (own) zout[(size_t)e * m + lane] = (double)zl;
        }
    }
    if (active) {
        if (wv == 0) {
            acc_t acc[NT][NT];
#pragma unroll
            for (int a = 0; a < NT; a++)
#pragma unroll
                for (int b = 0; b < NT; b++) acc[a][b] = acc_t{};
            mma_acc<T, MT>(acc, sm.k[0], sm.e[0][2]);  // A_k Y_l
            mma_acc<T, MT>(acc, sm.k[1], sm.e[1][1]);  // C_k X_r
#pragma unroll
            for (int ti = 0; ti < NT; ti++)
#pragma unroll
                for (int tj = 0; tj < NT; tj++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[ti][tj][v] = bk[ti][tj][v] - acc[ti][tj][v];
            store_acc<T, MT>(acc, B + (size_t)k * bs, m);
        } else {
            acc_t acc[NT][NT];
#pragma unroll
            for (int a = 0; a < NT; a++)
#pragma unroll
                for (int b = 0; b < NT; b++) acc[a][b] = acc_t{};
            mma_acc<T, MT>(acc, sm.k[0], sm.e[0][1]);  // A_k X_l
            store_neg<T, MT>(acc, A + (size_t)k * bs, m);
#pragma unroll
            for (int a = 0; a < NT; a++)
#pragma unroll
                for (int b = 0; b < NT; b++) acc[a][b] = acc_t{};
            mma_acc<T, MT>(acc, sm.k[1], sm.e[1][2]);  // C_k Y_r
            store_neg<T, MT>(acc, C + (size_t)k * bs, m);
        }
    }
    __syncthreads();  // sm.z is overwritten with z below (all readers of f_e are done)
    if (active && f && lane < MT) sm.z[wv][lane] = zl;
    __syncthreads();
    if (active && f && wv == 0 && lane < m) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < MT; j++)
            s += sm.k[0][lane * cfg<MT>::LD + j] * sm.z[0][j] + sm.k[1][lane * cfg<MT>::LD + j] * sm.z[1][j];
        f[(size_t)k * m + lane] = fk - (double)s;
    }
    __syncthreads();
}

// Top row (the single active row after the last level): Binv_top, and
// zout_top = Binv_top f_top when f is given.  One wavefront (wave 0 of pair 0).
template <typename T, int MT>
__device__ void top_step(const Fac &F, PairSmem<T, MT> &sm, int levels, const double *f, double *zout)
{
    const int m = F.m, lane = threadIdx.x & 63;
    const int e = lvl_row(0, levels);
    if (threadIdx.x < 64) {
        T col[cfg<MT>::S][MT];
        load_aug<T, MT>(col, (const T *)F.A, (const T *)F.B, (const T *)F.C, e, m, false, false);
        wave_gj<T, MT>(col, F.flag, m);
        store_elim<T, MT>(col, &sm.e[0][0][0], (T *)F.Bi, (T *)F.X, (T *)F.Y, e, m, true, f == nullptr);
        if (f && lane < m) {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < MT; j++)
                if (j < m) s += sm.e[0][0][lane * cfg<MT>::LD + j] * (T)f[(size_t)e * m + (j < m ? j : 0)];
            zout[(size_t)e * m + lane] = (double)s;
        }
    }
    __syncthreads();
}

// Back-substitution of level L for eliminated rows [tid0 .. ) strided by nthr:
// x_e = z_e - X_e x_l - Y_e x_r (x holds z_e on entry for eliminated rows).
// The dot products run over MT (the padded block size) with clamped indices and
// masked terms, in chunks of 16, so each chunk's loads are in flight together;
// the terms are subtracted in the order j = 0 .. m-1, left block first.
template <typename T, int MT>
__device__ __forceinline__ void back_rows(const T *X, const T *Y, int m, int L, int nL, double *x, int tid0, int nthr)
{
    const int ne = (nL + 1) / 2;
    const size_t bs = (size_t)m * m;
    for (int t = tid0; t < ne * m; t += nthr) {
        const int pe = 2 * (t / m), i = t % m;
        const int e = lvl_row(pe, L);
        const bool hasL = pe >= 1, hasR = pe + 1 < nL;
        const int l = hasL ? lvl_row(pe - 1, L) : e, r = hasR ? lvl_row(pe + 1, L) : e;
        double s = x[(size_t)e * m + i];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const T *Mr = (side == 0 ? X : Y) + (size_t)e * bs + (size_t)i * m;
            const double *xn = x + (size_t)(side == 0 ? l : r) * m;
            const bool has = side == 0 ? hasL : hasR;
#pragma unroll
            for (int c0 = 0; c0 < MT; c0 += 16) {
                T mv[16];
                double xv[16];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int jj = c0 + j < m ? c0 + j : m - 1;
                    mv[j] = Mr[jj];
                    xv[j] = xn[jj];
                }
#pragma unroll
                for (int j = 0; j < 16; j++)
                    if (has && c0 + j < m) s -= (double)mv[j] * xv[j];
            }
        }
        x[(size_t)e * m + i] = s;
    }
}

// ---- factor kernels -----------------------------------------------------------
template <typename T, int MT>
__global__ __launch_bounds__(128) void btd_level_kernel(Fac F, int L, int nL, T *Lsv, T *Rsv, double *f,
                                                        double *zout)
{
    __shared__ PairSmem<T, MT> sm;
    fac_step<T, MT>(F, sm, true, L, nL, blockIdx.x, Lsv, Rsv, f, zout);
}

// Levels L0 .. levels-1, the top and (with f) the back-substitution of levels
// levels-1 .. Lb (Lb <= L0: the upper levels below the tail have so few
// eliminated rows that their back-substitution costs less here than a launch),
// in one workgroup of PAIRS wave pairs.  nblk: block rows of level 0 (level L
// has nblk >> L active rows).
template <typename T, int MT>
__global__ __launch_bounds__(128 * cfg<MT>::PAIRS) void btd_tail_kernel(Fac F, int L0, int levels, int nL0,
                                                                        T *Lsv, T *Rsv, const long long *save,
                                                                        double *f, double *x, int Lb, int nblk)
{
    constexpr int PAIRS = cfg<MT>::PAIRS;
    __shared__ PairSmem<T, MT> sm[PAIRS];
    const int wp = threadIdx.x >> 7;
    int nL = nL0;
    for (int L = L0; L < levels; L++) {
        const int nk = nL / 2;
        for (int base = 0; base < nk; base += PAIRS) {
            const int q = base + wp;
            fac_step<T, MT>(F, sm[wp], q < nk, L, nL, q < nk ? q : 0, Lsv + save[L] * (size_t)F.m * F.m,
                            Rsv + save[L] * (size_t)F.m * F.m, f, x);
        }
        nL /= 2;
    }
    top_step<T, MT>(F, sm[0], levels, f, x);
    if (!f) return;
    // back-substitution of levels levels-1 .. Lb
    for (int L = levels - 1; L >= Lb; L--) {
        back_rows<T, MT>((const T *)F.X, (const T *)F.Y, F.m, L, nblk >> L, x, threadIdx.x, blockDim.x);
        __syncthreads();
    }
}

// ---- solve-only kernels (factors already in place) ------------------------------
// Forward of level L: one wavefront per kept row q; lanes 0..31 handle z_l,
// lanes 32..63 z_r, then f_k -= Lsv z_l + Rsv z_r.
template <typename T>
__device__ __forceinline__ void fwd_row(const T *Bi, const T *Lsv, const T *Rsv, int m, int L, int nL, int q,
                                        double *f, double *zout, double *sz)
{
    const int lane = threadIdx.x & 63;
    const int p = 2 * q + 1;
    const bool hasR = p + 1 < nL;
    const int k = lvl_row(p, L), l = lvl_row(p - 1, L), r = hasR ? lvl_row(p + 1, L) : 0;
    const size_t bs = (size_t)m * m;
    const int side = lane >> 5, i = lane & 31;
    double z = 0.0;
    if (i < m && (side == 0 || hasR)) {
        const int e = side == 0 ? l : r;
        const T *Br = Bi + (size_t)e * bs + (size_t)i * m;
        for (int j = 0; j < m; j++) z += (double)Br[j] * f[(size_t)e * m + j];
        if (side == 0 || p + 2 == nL) zout[(size_t)e * m + i] = z;
    }
    sz[lane] = z;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < m) {
        double s = 0.0;
        const T *Lr = Lsv + (size_t)q * bs + (size_t)lane * m, *Rr = Rsv + (size_t)q * bs + (size_t)lane * m;
        for (int j = 0; j < m; j++) s += (double)Lr[j] * sz[j] + (hasR ? (double)Rr[j] * sz[32 + j] : 0.0);
        f[(size_t)k * m + lane] -= s;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void btd_fwd_kernel(const T *Bi, const T *Lsv, const T *Rsv, int m, int L, int nL,
                                                      double *f, double *zout)
{
    __shared__ double sz[4][64];
    const int wave = threadIdx.x >> 6;
    const int q = blockIdx.x * 4 + wave;
    if (q >= nL / 2) return;
    fwd_row<T>(Bi, Lsv, Rsv, m, L, nL, q, f, zout, sz[wave]);
}

template <typename T, int MT>
__global__ __launch_bounds__(1024) void btd_fwd_tail_kernel(const T *Bi, const T *X, const T *Y, const T *Lsv,
                                                            const T *Rsv, const long long *save, int m, int L0,
                                                            int levels, int nL0, double *f, double *x)
{
    __shared__ double sz[16][64];
    const int wave = threadIdx.x >> 6;
    const size_t bs = (size_t)m * m;
    int nL = nL0;
    for (int L = L0; L < levels; L++) {
        for (int q = wave; q < nL / 2; q += 16)
            fwd_row<T>(Bi, Lsv + save[L] * bs, Rsv + save[L] * bs, m, L, nL, q, f, x, sz[wave]);
        __syncthreads();
        nL /= 2;
    }
    const int top = lvl_row(0, levels);
    if (threadIdx.x < (unsigned)m) {
        double s = 0.0;
        const T *Br = Bi + (size_t)top * bs + (size_t)threadIdx.x * m;
        for (int j = 0; j < m; j++) s += (double)Br[j] * f[(size_t)top * m + j];
        x[(size_t)top * m + threadIdx.x] = s;
    }
    __syncthreads();
    for (int L = levels - 1; L >= L0; L--) {
        int n = nL0;
        for (int j = L0; j < L; j++) n /= 2;
        back_rows<T, MT>(X, Y, m, L, n, x, threadIdx.x, blockDim.x);
        __syncthreads();
    }
}

template <typename T, int MT>
__global__ __launch_bounds__(256) void btd_back_kernel(const T *X, const T *Y, int m, int L, int nL, double *x)
{
    back_rows<T, MT>(X, Y, m, L, nL, x, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

}  // namespace

// Level structure of the cyclic reduction (no device call).  Returns the number
// of saved coupling blocks (the Lsv / Rsv extent).
int gf_btd_plan(gf_btd &b, int m, int nblk, int elem_size)
{
    if (m < 1 || m > 32) {
        gf_set_error("block-tridiagonal solver: block size out of range");
        return GF_E_UNSUPPORTED;
    }
    b.m = m;
    b.mt = m <= 16 ? 16 : 32;
    b.nblk = nblk;
    b.bsz = (size_t)m * m;
    b.elem_size = elem_size;
    b.lvl_n.clear();
    b.lvl_save.clear();
    const int tail_rows = b.mt == 16 ? 16 : 4;  // 4 x the tail workgroup's wave pairs
    int n = nblk, L = 0;
    size_t save = 0;
    b.tail_level = -1;
    while (n > 1) {
        if (b.tail_level < 0 && n <= tail_rows) b.tail_level = L;
        b.lvl_n.push_back(n);
        b.lvl_save.push_back((long long)save);
        save += n / 2;
        n /= 2;
        L++;
    }
    b.levels = L;
    if (b.tail_level < 0) b.tail_level = L;
    // the back-substitution of the levels above the tail runs as one wide launch
    // per level: inside the single tail workgroup it took longer than the launches
    b.back_level = b.tail_level;
    b.n_save = (long long)save;
    return GF_OK;
}

int gf_btd_alloc(gf_btd &b, int m, int nblk, int elem_size)
{
    GF_TRY(gf_btd_plan(b, m, nblk, elem_size));
    const size_t save = (size_t)b.n_save;
    size_t blocks = (size_t)nblk * b.bsz * elem_size;
    size_t sv = (save ? save : 1) * b.bsz * elem_size;
    // A, B, C contiguous: the assembly clears them with one fill
    GF_TRY(gf_pool_alloc(&b.A, 3 * blocks));
    b.B = (char *)b.A + blocks;
    b.C = (char *)b.A + 2 * blocks;
    void **arrs[] = {&b.Bi, &b.X, &b.Y};
    for (void **a : arrs) GF_TRY(gf_pool_alloc(a, blocks));
    GF_TRY(gf_pool_alloc(&b.Lsv, sv));
    GF_TRY(gf_pool_alloc(&b.Rsv, sv));
    GF_TRY(gf_pool_alloc((void **)&b.d_save, sizeof(long long) * (b.lvl_save.size() + 1)));
    std::vector<long long> sv_h(b.lvl_save);
    sv_h.push_back((long long)save);
    GF_HIP(hipMemcpy(b.d_save, sv_h.data(), sizeof(long long) * sv_h.size(), hipMemcpyHostToDevice));
    // X and Y of rows that are never eliminated with a coupling stay zero
    GF_HIP(hipMemset(b.X, 0, blocks));
    GF_HIP(hipMemset(b.Y, 0, blocks));
    return GF_OK;
}

void gf_btd_free(gf_btd &b)
{
    void *arrs[] = {b.A, b.Bi, b.X, b.Y, b.Lsv, b.Rsv, b.d_save};  // B and C live inside A's allocation
    for (void *a : arrs) gf_pool_free(a);
    b.A = b.B = b.C = b.Bi = b.X = b.Y = b.Lsv = b.Rsv = nullptr;
    b.d_save = nullptr;
}

template <typename T, int MT>
static int btd_factor_t(gf_btd &b, hipStream_t s, int *d_flag, double *f, double *x)
{
    Fac F{b.A, b.B, b.C, b.Bi, b.X, b.Y, b.m, d_flag};
    T *Lsv = (T *)b.Lsv, *Rsv = (T *)b.Rsv;
    const size_t bs = b.bsz;
    int launches = 0;
    for (int L = 0; L < b.tail_level; L++) {
        const int nL = b.lvl_n[L];
        const int pr = gf_probe_begin(s, GF_KC_BTD_LEVEL);
        hipLaunchKernelGGL((btd_level_kernel<T, MT>), dim3(nL / 2), dim3(128), 0, s, F, L, nL, Lsv + b.lvl_save[L] * bs,
                           Rsv + b.lvl_save[L] * bs, f, x);
        gf_probe_end(s, pr, nL / 2);
        launches++;
    }
    const int nL0 = b.tail_level < b.levels ? b.lvl_n[b.tail_level] : 1;
    const int pr = gf_probe_begin(s, GF_KC_BTD_TAIL);
    hipLaunchKernelGGL((btd_tail_kernel<T, MT>), dim3(1), dim3(128 * cfg<MT>::PAIRS), 0, s, F, b.tail_level, b.levels,
                       nL0, Lsv, Rsv, (const long long *)b.d_save, f, x, b.back_level, b.nblk);
    gf_probe_end(s, pr, nL0);
    launches++;
    if (f) {
        for (int L = b.back_level - 1; L >= 0; L--) {
            const int nL = b.lvl_n[L];
            const int work = (nL + 1) / 2 * b.m;
            hipLaunchKernelGGL((btd_back_kernel<T, MT>), dim3((work + 255) / 256), dim3(256), 0, s, (const T *)b.X,
                               (const T *)b.Y, b.m, L, nL, x);
            launches++;
        }
    }
    GF_HIP(hipGetLastError());
    return launches;
}

template <typename T> static int btd_factor_dispatch(gf_btd &b, hipStream_t s, int *d_flag, double *f, double *x)
{
    return b.mt == 16 ? btd_factor_t<T, 16>(b, s, d_flag, f, x) : btd_factor_t<T, 32>(b, s, d_flag, f, x);
}

int gf_btd_factor(gf_btd &b, hipStream_t s, int *d_flag)
{
    if (b.elem_size == 8) return btd_factor_dispatch<double>(b, s, d_flag, nullptr, nullptr);
    return btd_factor_dispatch<float>(b, s, d_flag, nullptr, nullptr);
}

int gf_btd_factor_solve(gf_btd &b, hipStream_t s, int *d_flag, const double *rhs, double *x, int n, double *work)
{
    if (rhs == x || work == x) {
        gf_set_error("block-tridiagonal factor_solve: x must differ from rhs and work");
        return GF_E_ARG;
    }
    if (rhs != work) GF_HIP(hipMemcpyAsync(work, rhs, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    if (b.elem_size == 8) return btd_factor_dispatch<double>(b, s, d_flag, work, x);
    return btd_factor_dispatch<float>(b, s, d_flag, work, x);
}

template <typename T, int MT> static int btd_solve_t(gf_btd &b, hipStream_t s, double *f, double *x)
{
    const T *Bi = (const T *)b.Bi, *X = (const T *)b.X, *Y = (const T *)b.Y;
    const T *Lsv = (const T *)b.Lsv, *Rsv = (const T *)b.Rsv;
    const size_t bs = b.bsz;
    int launches = 0;
    for (int L = 0; L < b.tail_level; L++) {
        const int nL = b.lvl_n[L];
        hipLaunchKernelGGL((btd_fwd_kernel<T>), dim3((nL / 2 + 3) / 4), dim3(256), 0, s, Bi, Lsv + b.lvl_save[L] * bs,
                           Rsv + b.lvl_save[L] * bs, b.m, L, nL, f, x);
        launches++;
    }
    const int nL0 = b.tail_level < b.levels ? b.lvl_n[b.tail_level] : 1;
    hipLaunchKernelGGL((btd_fwd_tail_kernel<T, MT>), dim3(1), dim3(1024), 0, s, Bi, X, Y, Lsv, Rsv,
                       (const long long *)b.d_save, b.m, b.tail_level, b.levels, nL0, f, x);
    launches++;
    for (int L = b.tail_level - 1; L >= 0; L--) {
        const int nL = b.lvl_n[L];
        const int work = (nL + 1) / 2 * b.m;
        hipLaunchKernelGGL((btd_back_kernel<T, MT>), dim3((work + 255) / 256), dim3(256), 0, s, X, Y, b.m, L, nL, x);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

// x = K^-1 rhs with the factors in place; work (n entries) holds the reduced
// right-hand side; rhs may alias work but not x.
int gf_btd_solve(gf_btd &b, hipStream_t s, const double *rhs, double *x, int n, double *work)
{
    if (rhs == x || work == x) {
        gf_set_error("block-tridiagonal solve: x must differ from rhs and work");
        return GF_E_ARG;
    }
    if (rhs != work) GF_HIP(hipMemcpyAsync(work, rhs, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    if (b.elem_size == 8)
        return b.mt == 16 ? btd_solve_t<double, 16>(b, s, work, x) : btd_solve_t<double, 32>(b, s, work, x);
    return b.mt == 16 ? btd_solve_t<float, 16>(b, s, work, x) : btd_solve_t<float, 32>(b, s, work, x);
}
