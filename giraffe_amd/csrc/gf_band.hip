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
// pivoting), panel width NB = 64 over the padded order n (multiple of NB).
// Per panel p (columns k0 = 64 p ..):
//   panel  L21 = A21 U11^-1 and U12 = L11^-1 A12 as MFMA products with the
//          panel's stored inverses (no sequential substitution);
//   update A22 -= L21 U12 on the band window, MFMA 16x16x4 tiles of 64 x 64;
//          the workgroup owning the next diagonal block factors it in LDS
//          (register-blocked, one barrier per pivot) and writes its triangular
//          inverses, which the next panel and both triangular solves use.
// Zero or non-finite pivots raise the failure flag (info_fail).
//
// Triangular solves: one single-workgroup launch per sweep over 64-row blocks,
// each block finished by its stored inverse (no other workgroup takes part).
#include "gf_internal.h"

namespace {

constexpr int NB = 64;
constexpr int KT = 32;  // k-chunk of the MFMA tiles

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Operands of both: A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k.  The
// 16 x 16 result differs: item v of lane l holds row 4 v + l / 16 for f64 and
// row 4 (l / 16) + v for f32, column l % 16 in both (checked on gfx950).
template <typename T> struct mfma;
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

// Masked band load without a branch around it: the address is clamped into the
// stored band of row min(i, n - 1) (always inside the n (2hb+1) allocation) and
// the value is selected afterwards, so a thread's loads are all in flight
// together instead of each waiting in its own branch.  i, c >= 0.
template <typename T> __device__ __forceinline__ T band_ld(const T *ab, int i, int c, int hb, int n, bool ok)
{
    const int ii = min(i, n - 1);
    const int cc = min(max(c, ii - hb), ii + hb);
    const T v = ab[bidx(ii, cc, hb)];
    return ok && i < n && inband(i, c, hb) ? v : T(0);
}

template <typename T> __device__ __forceinline__ bool bad_pivot(T v)
{
    const double a = fabs((double)v);
    return !(v == v) || a >= 1e300 || a < (sizeof(T) == 4 ? 1e-37 : 1e-300);
}

// Per-panel inverses: inv + p * 2 * NB * NB holds Uinv (row-major), then Linv (unit lower).
template <typename T> __device__ __forceinline__ T *uinv_of(T *inv, int p) { return inv + (size_t)p * 2 * NB * NB; }
template <typename T> __device__ __forceinline__ T *linv_of(T *inv, int p) { return inv + (size_t)p * 2 * NB * NB + NB * NB; }

// ---- factorisation of one 64 x 64 diagonal block --------------------------------
// Blocked over 16 x 16 blocks D(a, b) of the LDS tile s (leading dimension ld),
// by a workgroup of 4 wavefronts:
//   for kb = 0..3:  wave 0 factors D(kb, kb) in registers (wave_lu16: L\U in
//                   place, its triangular inverses into li[kb] / ui[kb]);
//                   L(i, kb) = D(i, kb) Uinv(kb), U(kb, j) = Linv(kb) D(kb, j);
//                   D(i, j) -= L(i, kb) U(kb, j)            (MFMA, all waves)
//   then wave j assembles block column j of the 64 x 64 Linv and Uinv by block
//   substitution (MFMA) and writes both to global memory.
// Barriers: 3 per kb; the sequential pivot steps run inside one wavefront.
constexpr int LB = 16, LLD = LB + 1;  // 16 x 16 blocks in LDS with leading dimension 17

template <typename T> __device__ __forceinline__ T rdl(T v, int l);
template <> __device__ __forceinline__ double rdl<double>(double v, int l)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
template <> __device__ __forceinline__ float rdl<float>(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// One wavefront: LU (no pivoting) of the 16 x 16 block at s (leading dimension
// ld), with Linv (unit lower) and Uinv written row-major to li / ui (leading
// dimension LLD).  Lane c of the first 16 holds column c of the block, lanes
// 16..31 the identity that the forward elimination turns into Linv, lanes
// 32..47 the identity that the backward substitution turns into Uinv.  On
// return s holds U on and above the diagonal; below it the values are not the
// multipliers (nothing reads them).
template <typename T> __device__ __forceinline__ void wave_lu16(T *s, int ld, T *li, T *ui, int *flag)
{
    // Branch-free: the masks are applied arithmetically and a bad pivot is only
    // recorded, so the unrolled sweep is one basic block and the scheduler can
    // start the next pivot's reciprocal while the current updates issue.
    const int lane = threadIdx.x & 63, part = lane >> 4, c = lane & 15;
    T col[LB];
#pragma unroll
    for (int i = 0; i < LB; i++) {
        const T d = s[i * ld + c];
        col[i] = part == 0 ? d : (part == 1 && i == c) ? T(1) : T(0);
    }
    T rp[LB];
    bool any_bad = false;
#pragma unroll
    for (int k = 0; k < LB; k++) {
        T p = rdl<T>(col[k], k);
        const bool bad = bad_pivot(p);
        any_bad |= bad;
        p = bad ? T(1) : p;
        rp[k] = pivot_rcp(p);
        T a[LB];
#pragma unroll
        for (int i = k + 1; i < LB; i++) a[i] = rdl<T>(col[i], k);
        // row i -= (a_ik / a_kk) row k, applied to the columns right of k and to
        // every column of the identity part (elsewhere the scale is zero); the
        // multipliers themselves are not kept: only the inverses of the
        // diagonal blocks are used downstream.
        const bool act = (part == 0 && c > k) || part == 1;
        const T ck = act ? col[k] * rp[k] : T(0);
#pragma unroll
        for (int i = k + 1; i < LB; i++) col[i] -= a[i] * ck;
    }
    const bool inv_u = part == 2;
#pragma unroll
    for (int i = 0; i < LB; i++) col[i] = inv_u ? (i == c ? T(1) : T(0)) : col[i];
#pragma unroll
    for (int k = LB - 1; k >= 0; k--) {
        T u[LB];
#pragma unroll
        for (int i = 0; i < k; i++) u[i] = rdl<T>(col[i], k);
        const T xk = col[k] * rp[k];
        col[k] = inv_u ? xk : col[k];
        const T xm = inv_u ? xk : T(0);
#pragma unroll
        for (int i = 0; i < k; i++) col[i] -= u[i] * xm;
    }
    if (any_bad && lane == 0) atomicExch(flag, 1);
    if (part < 3) {
        T *dst = part == 0 ? s : part == 1 ? li : ui;
        const int ldd = part == 0 ? ld : LLD;
#pragma unroll
        for (int i = 0; i < LB; i++) dst[i * ldd + c] = col[i];
    }
}

// acc += P Q for 16 x 16 blocks in LDS (one wavefront, 4 MFMA 16x16x4).
template <typename T>
__device__ __forceinline__ void mm16(typename mfma<T>::acc_t &acc, const T *P, int ldp, const T *Q, int ldq)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k0 = 0; k0 < LB; k0 += 4) {
        const T a = P[(lane & 15) * ldp + k0 + (lane >> 4)];
        const T b = Q[(k0 + (lane >> 4)) * ldq + (lane & 15)];
        acc = mfma<T>::run(a, b, acc);
    }
}

// Stores sign * acc (result layout) as a row-major 16 x 16 block.
template <typename T>
__device__ __forceinline__ void st16(const typename mfma<T>::acc_t &acc, T *dst, int ld, T sign)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 4; v++) dst[mfma<T>::row(lane, v) * ld + (lane & 15)] = sign * acc[v];
}

template <typename T> __device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

template <typename T> struct LuSmem {
    T li[4][LB * LLD], ui[4][LB * LLD];  // inverses of the 16 x 16 diagonal blocks
    T ws[4][LB * LLD];                   // per-wave staging block
};

// s: the 64 x 64 tile (ld) in LDS; uinv / linv: 64 x 64 row-major outputs.
// All 256 threads call it.
template <typename T>
__device__ __forceinline__ void tile_lu_inv(T *s, int ld, LuSmem<T> &w, T *uinv, T *linv, int *flag)
{
    typedef typename mfma<T>::acc_t acc_t;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto D = [&](int a, int b) { return s + (LB * a) * ld + LB * b; };
    for (int kb = 0; kb < 4; kb++) {
        if (wave == 0) wave_lu16<T>(D(kb, kb), ld, w.li[kb], w.ui[kb], flag);
        __syncthreads();
        // L(i, kb) = D(i, kb) Uinv(kb) for i > kb; U(kb, j) = Linv(kb) D(kb, j) for j > kb
        const int nr = 3 - kb;
        for (int t = wave; t < 2 * nr; t += 4) {
            acc_t acc = acc_t{};
            T *dst;
            if (t < nr) {
                dst = D(kb + 1 + t, kb);
                mm16<T>(acc, dst, ld, w.ui[kb], LLD);
            } else {
                dst = D(kb, kb + 1 + (t - nr));
                mm16<T>(acc, w.li[kb], LLD, dst, ld);
            }
            wave_sync_lds<T>();
            st16<T>(acc, dst, ld, T(1));
        }
        __syncthreads();
        // D(i, j) -= L(i, kb) U(kb, j)
        for (int t = wave; t < nr * nr; t += 4) {
            const int i = kb + 1 + t / nr, j = kb + 1 + t % nr;
            acc_t acc = acc_t{};
            mm16<T>(acc, D(i, kb), ld, D(kb, j), ld);
            T *dst = D(i, j);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T &e = dst[mfma<T>::row(lane, v) * ld + (lane & 15)];
                e -= acc[v];
            }
        }
        __syncthreads();
    }
    // Block column j = wave of Linv (forward) and Uinv (backward) substitution.
    const int j = wave;
    T *ws = w.ws[j];
    auto put = [&](T *out, int bi, const acc_t &acc, T sign) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            out[(LB * bi + mfma<T>::row(lane, v)) * NB + LB * j + (lane & 15)] = sign * acc[v];
    };
    auto put_lds = [&](T *out, int bi, const T *src) {  // copy a 16 x 16 LDS block (LLD) to the output
        for (int t = lane; t < LB * LB; t += 64)
            out[(LB * bi + t / LB) * NB + LB * j + t % LB] = src[(t / LB) * LLD + t % LB];
    };
    auto zero = [&](T *out, int bi) {
        for (int t = lane; t < LB * LB; t += 64) out[(LB * bi + t / LB) * NB + LB * j + t % LB] = T(0);
    };
    // Linv(:, j): X(j) = Linv(j); X(i) = -Linv(i) sum_{k=j}^{i-1} L(i, k) X(k), i > j
    // (loops unrolled over the block indices; j is wave-uniform)
    {
        for (int i = 0; i < j; i++) zero(linv, i);
        put_lds(linv, j, w.li[j]);
        acc_t acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = acc_t{};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k < j) continue;
            const T *xk = k == j ? w.li[j] : ws;  // X(k) as a product operand
#pragma unroll
            for (int i = k + 1; i < 4; i++) mm16<T>(acc[i], D(i, k), ld, xk, LLD);
            // X(k+1) = -Linv(k+1) acc[k+1]
            wave_sync_lds<T>();
            st16<T>(acc[k + 1], ws, LLD, T(1));
            wave_sync_lds<T>();
            acc_t x = acc_t{};
            mm16<T>(x, w.li[k + 1], LLD, ws, LLD);
            put(linv, k + 1, x, T(-1));
            wave_sync_lds<T>();
            st16<T>(x, ws, LLD, T(-1));
            wave_sync_lds<T>();
        }
    }
    wave_sync_lds<T>();
    // Uinv(:, j): X(j) = Uinv(j); X(i) = -Uinv(i) sum_{k=i+1}^{j} U(i, k) X(k), i < j
    {
        for (int i = j + 1; i < 4; i++) zero(uinv, i);
        put_lds(uinv, j, w.ui[j]);
        acc_t acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = acc_t{};
#pragma unroll
        for (int k = 3; k > 0; k--) {
            if (k > j) continue;
            const T *xk = k == j ? w.ui[j] : ws;
#pragma unroll
            for (int i = 0; i < k; i++) mm16<T>(acc[i], D(i, k), ld, xk, LLD);
            wave_sync_lds<T>();
            st16<T>(acc[k - 1], ws, LLD, T(1));
            wave_sync_lds<T>();
            acc_t x = acc_t{};
            mm16<T>(x, w.ui[k - 1], LLD, ws, LLD);
            put(uinv, k - 1, x, T(-1));
            wave_sync_lds<T>();
            st16<T>(x, ws, LLD, T(-1));
            wave_sync_lds<T>();
        }
    }
    __syncthreads();
}

// 64 x 64 x 64 product on 256 threads (4 waves, 2 x 2 of 32 x 32):
// acc = P Q with P and Q staged chunk by chunk through sa [64][KT+1], sb [KT][64].
template <typename T, typename LP, typename LQ>
__device__ __forceinline__ void tile_mm(typename mfma<T>::acc_t (&acc)[2][2], T *sa, T *sb, LP loadP, LQ loadQ)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    for (int kc = 0; kc < NB; kc += KT) {
        // all loads of the chunk first (in flight together), then the LDS stores
        constexpr int PER = NB * KT / 256;
        T pv[PER], qv[PER];
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int t = tid + 256 * r;
            pv[r] = loadP(t / KT, kc + t % KT);
            qv[r] = loadQ(kc + t / NB, t % NB);
        }
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int t = tid + 256 * r;
            sa[(t / KT) * (KT + 1) + t % KT] = pv[r];
            sb[(t / NB) * NB + t % NB] = qv[r];
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
}

template <typename T> struct TileSmem {
    T ab[NB * (KT + 1) + KT * NB];  // staging of tile_mm; after it, the LU tile [64][65]
};
template <typename T> struct DiagSmem {
    T ab[NB * (KT + 1) + KT * NB];
    LuSmem<T> lu;
};

// First diagonal block: LU in place and its inverses (panel 0).
template <typename T>
__global__ __launch_bounds__(256) void band_diag_kernel(T *ab, int hb, int n, T *inv, int *flag)
{
    __shared__ DiagSmem<T> sm;
    T *s = sm.ab;
    T v[NB * NB / 256];
#pragma unroll
    for (int r = 0; r < NB * NB / 256; r++) {
        const int t = threadIdx.x + 256 * r;
        v[r] = band_ld(ab, t / NB, t % NB, hb, n, true);
    }
#pragma unroll
    for (int r = 0; r < NB * NB / 256; r++) {
        const int t = threadIdx.x + 256 * r;
        s[(t / NB) * (NB + 1) + t % NB] = v[r];
    }
    __syncthreads();
    tile_lu_inv<T>(s, NB + 1, sm.lu, uinv_of(inv, 0), linv_of(inv, 0), flag);
    __syncthreads();
    for (int t = threadIdx.x; t < NB * NB; t += 256) {
        int i = t / NB, c = t % NB;
        if (inband(i, c, hb)) ab[bidx(i, c, hb)] = s[i * (NB + 1) + c];
    }
}

// Panel p: blockIdx < nt: L21 tile = A21 Uinv; otherwise U12 tile = Linv A12.
template <typename T> __global__ __launch_bounds__(256) void band_panel_kernel(T *ab, int hb, int p, int n, int nt,
                                                                               const T *inv)
{
    typedef typename mfma<T>::acc_t acc_t;
    __shared__ TileSmem<T> sm;
    T *sa = sm.ab, *sb = sm.ab + NB * (KT + 1);
    const int k0 = p * NB, lim = min(n, k0 + NB + hb);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{};
    const T *U = uinv_of(const_cast<T *>(inv), p), *L = linv_of(const_cast<T *>(inv), p);
    const bool lower = (int)blockIdx.x < nt;
    const int o0 = k0 + NB + (lower ? blockIdx.x : blockIdx.x - nt) * NB;  // first row (lower) or column (upper)
    if (lower) {
        tile_mm<T>(acc, sa, sb,
                   [&](int i, int k) { int gi = o0 + i; return band_ld(ab, gi, k0 + k, hb, n, gi < lim); },
                   [&](int k, int j) { return U[k * NB + j]; });
    } else {
        tile_mm<T>(acc, sa, sb, [&](int i, int k) { return L[i * NB + k]; },
                   [&](int k, int j) { int gj = o0 + j; return band_ld(ab, k0 + k, gj, hb, n, gj < lim); });
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = wr + 16 * a + mfma<T>::row(lane, v), j = wc + 16 * b + (lane & 15);
                int gi = lower ? o0 + i : k0 + i, gj = lower ? k0 + j : o0 + j;
                bool ok = lower ? gi < lim : gj < lim;
                if (ok && inband(gi, gj, hb)) ab[bidx(gi, gj, hb)] = acc[a][b][v];
            }
}

// A22 -= L21 U12 on 64 x 64 tiles of the window [k0+NB, lim)^2; the tile that
// holds the next diagonal block then factors it and writes panel p+1's inverses.
template <typename T> __global__ __launch_bounds__(256) void band_update_kernel(T *ab, int hb, int p, int n, int nt,
                                                                                T *inv, int *flag)
{
    typedef typename mfma<T>::acc_t acc_t;
    __shared__ DiagSmem<T> sm;
    T *sa = sm.ab, *sb = sm.ab + NB * (KT + 1), *sc = sm.ab;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k0 = p * NB, base = k0 + NB, lim = min(n, k0 + NB + hb);
    const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
    const int R0 = base + ti * NB, C0 = base + tj * NB;
    if (R0 - (C0 + NB - 1) > hb || C0 - (R0 + NB - 1) > hb) return;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{};
    // the old values of this thread's result entries, read before the product
    // so their latency hides behind it
    T old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = wr + 16 * a + mfma<T>::row(lane, v), j = wc + 16 * b + (lane & 15);
                old[a][b][v] = band_ld(ab, R0 + i, C0 + j, hb, n, C0 + j < n);
            }
    tile_mm<T>(acc, sa, sb,
               [&](int i, int k) { int gi = R0 + i; return band_ld(ab, gi, k0 + k, hb, n, gi < lim); },
               [&](int k, int j) { int gj = C0 + j; return band_ld(ab, k0 + k, gj, hb, n, gj < lim); });
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
                const bool in = gi < n && gj < n && inband(gi, gj, hb);
                const T nv = old[a][b][v] - acc[a][b][v];
                if (fuse) sc[i * (NB + 1) + j] = in ? nv : T(0);
                else if (in) ab[bidx(gi, gj, hb)] = nv;
            }
    if (!fuse) return;
    __syncthreads();
    tile_lu_inv<T>(sc, NB + 1, sm.lu, uinv_of(inv, p + 1), linv_of(inv, p + 1), flag);
    __syncthreads();
    for (int t = tid; t < NB * NB; t += 256) {
        int i = t / NB, j = t % NB;
        int gi = R0 + i, gj = C0 + j;
        if (inband(gi, gj, hb)) ab[bidx(gi, gj, hb)] = sc[i * (NB + 1) + j];
    }
}

// ---- triangular solves ------------------------------------------------------------
// Right-looking block sweeps, one launch per batch of K consecutive 64-row
// blocks (forward: L y = b with unit L, blocks ascending; backward: U x = y,
// blocks descending).  The right-hand side f is updated in place: when a batch
// starts, every contribution of the blocks before it has been subtracted.
// Every workgroup of the launch solves the batch itself from f and the stored
// inverses (K inverse and K(K-1)/2 coupling products, the same data for all
// workgroups, so it comes from L2); workgroup 0 writes the batch's solution and
// workgroup g >= 1 subtracts the batch's couplings from block g past it.  No
// workgroup reads what another workgroup of the same launch writes, so the
// stream order of the launches is the only synchronisation, and the coupling
// strips are streamed by as many CUs as the band reaches blocks.
template <typename T, typename V>
__device__ __forceinline__ V mv_part(const T *row, int c0, const V *v)  // 16 products of one row quarter
{
    T a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = row[c0 + j];
    V s = V(0);
#pragma unroll
    for (int j = 0; j < 16; j++) s += (V)a[j] * v[c0 + j];
    return s;
}

template <typename V> __device__ __forceinline__ V quad_sum(V s)
{
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    return s;
}

// Row i = tid / 4 of the 64 x 64 block coupling band rows R0.. to band columns
// C0.., times v (LDS): the sum over the quad (masked to the band).
template <typename T, typename V>
__device__ __forceinline__ V coupling_row(const T *ab, int hb, int R0, int C0, const V *v)
{
    const int tid = threadIdx.x, i = tid >> 2, c0 = 16 * (tid & 3);
    const T *rowp = ab + (size_t)(R0 + i) * (size_t)(2 * hb) + hb + C0;  // entry (R0+i, C0+c) at rowp[c]
    T a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = rowp[c0 + j];
    V s = V(0);
#pragma unroll
    for (int j = 0; j < 16; j++) s += inband(R0 + i, C0 + c0 + j, hb) ? (V)a[j] * v[c0 + j] : V(0);
    return quad_sum(s);
}

template <typename T, typename V, bool FWD, int K>
__global__ __launch_bounds__(256) void band_sweep_kernel(const T *ab, int hb, int n, const T *inv, int r0, int kc,
                                                         V *f, V *out)
{
    __shared__ V fb[K][NB], ys[K][NB];
    const int tid = threadIdx.x, i = tid >> 2, c0 = 16 * (tid & 3);
    auto blk = [&](int k) { return FWD ? r0 + k : r0 - k; };
    for (int t = tid; t < kc * NB; t += 256) fb[t / NB][t % NB] = f[(size_t)blk(t / NB) * NB + t % NB];
    __syncthreads();
    for (int k = 0; k < kc; k++) {
        const int b = blk(k);
        const T *Ti = FWD ? linv_of(const_cast<T *>(inv), b) : uinv_of(const_cast<T *>(inv), b);
        const V yv = quad_sum(mv_part<T, V>(Ti + i * NB, c0, fb[k]));
        if ((tid & 3) == 0) ys[k][i] = yv;
        __syncthreads();
        for (int k2 = k + 1; k2 < kc; k2++) {
            const V cv = coupling_row<T, V>(ab, hb, blk(k2) * NB, b * NB, ys[k]);
            if ((tid & 3) == 0) fb[k2][i] -= cv;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        for (int t = tid; t < kc * NB; t += 256) out[(size_t)blk(t / NB) * NB + t % NB] = ys[t / NB][t % NB];
        return;
    }
    const int q = FWD ? r0 + kc - 1 + (int)blockIdx.x : r0 - kc + 1 - (int)blockIdx.x;  // target block
    V acc = V(0);
    for (int k = 0; k < kc; k++) acc += coupling_row<T, V>(ab, hb, q * NB, blk(k) * NB, ys[k]);
    if ((tid & 3) == 0) f[(size_t)q * NB + i] -= acc;
}

}  // namespace

// ---- host -----------------------------------------------------------------------
template <typename T> static int band_factor_t(T *ab, int n, int hb, hipStream_t s, int *flag, T *inv)
{
    int launches = 0;
    hipLaunchKernelGGL(band_diag_kernel<T>, dim3(1), dim3(256), 0, s, ab, hb, n, inv, flag);
    launches++;
    for (int p = 0; p * NB < n; p++) {
        int k0 = p * NB;
        int lim = std::min(n, k0 + NB + hb);
        int rem = lim - (k0 + NB);
        if (rem <= 0) break;
        int nt = (rem + NB - 1) / NB;
        int pr = gf_probe_begin(s, GF_KC_BAND_PANEL);
        hipLaunchKernelGGL(band_panel_kernel<T>, dim3(2 * nt), dim3(256), 0, s, ab, hb, p, n, nt, (const T *)inv);
        gf_probe_end(s, pr, NB);
        pr = gf_probe_begin(s, GF_KC_BAND_UPDATE);
        hipLaunchKernelGGL(band_update_kernel<T>, dim3(nt * nt), dim3(256), 0, s, ab, hb, p, n, nt, inv, flag);
        gf_probe_end(s, pr, NB);
        launches += 2;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

constexpr int SWEEP_K = 2;  // blocks per sweep launch (4 measured slower: longer in-batch chain)

template <typename T, typename V>
static int band_solve_t(const T *ab, int n, int hb, hipStream_t s, V *x, V *z, const T *inv)
{
    const int nbk = n / NB, reach = (hb + NB - 1) / NB;
    int launches = 0;
    // forward: L z = x, x updated in place as the right-hand side
    for (int r0 = 0; r0 < nbk; r0 += SWEEP_K) {
        const int kc = std::min(SWEEP_K, nbk - r0), last = r0 + kc - 1;
        const int g = 1 + std::min(reach, nbk - 1 - last);
        const int pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
        hipLaunchKernelGGL((band_sweep_kernel<T, V, true, SWEEP_K>), dim3(g), dim3(256), 0, s, ab, hb, n, inv, r0, kc,
                           x, z);
        gf_probe_end(s, pr, kc * NB);
        launches++;
    }
    // backward: U x = z, z updated in place
    for (int r0 = nbk - 1; r0 >= 0; r0 -= SWEEP_K) {
        const int kc = std::min(SWEEP_K, r0 + 1), last = r0 - kc + 1;
        const int g = 1 + std::min(reach, last);
        const int pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
        hipLaunchKernelGGL((band_sweep_kernel<T, V, false, SWEEP_K>), dim3(g), dim3(256), 0, s, ab, hb, n, inv, r0,
                           kc, z, x);
        gf_probe_end(s, pr, kc * NB);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

size_t gf_band_inv_elems(int n) { return (size_t)(n / NB) * 2 * NB * NB; }

int gf_band_factor(void *ab, int elem_size, int n, int hb, hipStream_t s, int *flag, void *inv)
{
    if (n % NB != 0) {
        gf_set_error("band factor: order must be a multiple of 64");
        return GF_E_ARG;
    }
    if (elem_size == 8) return band_factor_t<double>((double *)ab, n, hb, s, flag, (double *)inv);
    return band_factor_t<float>((float *)ab, n, hb, s, flag, (float *)inv);
}

// x: right-hand side in, solution out; work: n entries.
int gf_band_solve(const void *ab, int elem_size, int n, int hb, hipStream_t s, double *x, double *work, const void *inv)
{
    if (elem_size == 8)
        return band_solve_t<double, double>((const double *)ab, n, hb, s, x, work, (const double *)inv);
    return band_solve_t<float, double>((const float *)ab, n, hb, s, x, work, (const float *)inv);
}

int gf_band_pad(int n) { return (n + NB - 1) / NB * NB; }
