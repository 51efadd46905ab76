// 64 x 64 tile machinery shared by the wide-band and the multifrontal
// factorisations on gfx950: FP64 / FP32 MFMA 16x16x4 products of 64 x 64
// tiles staged through LDS, the in-register 16 x 16 Gauss-Jordan leaf and the
// 64 x 64 block inverse with pivot perturbation, row-major tile transfers.
// Device code only; every definition has internal linkage.
#pragma once
#include "gf_internal.h"

namespace {

constexpr int NB = 64;   // block size
constexpr int KT = 32;   // k-chunk of the staged MFMA products
// LDS strides (elements) chosen for the MFMA operand patterns in FP64 (64
// banks of 4 bytes, 32 lanes of 8-byte reads per cycle): an A operand reads
// rows i = 0..15 at column offsets c = 0..1, a B operand rows c at columns
// j = 0..15.  Row stride 66 (2 mod 32) and 34 spread A reads over all banks,
// 80 (16 mod 32) does the same for B reads, 18 for the saved 16 x 16 blocks.
constexpr int LDT = NB + 2, LDQ = NB + 16;
constexpr int SA_LD = KT + 2, SB_LD = NB + 16, TMP_LD = 18;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA.  Operands: A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k.
// Result item v of lane l: row 4 v + l / 16 (f64) or 4 (l / 16) + v (f32),
// column l % 16 (checked on gfx950).
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

// Per-wave part of a 64 x 64 result: rows wr + 16a + row(lane, v), columns wc + 16b + lane % 16.
template <typename T> struct Acc {
    typename mfma<T>::acc_t a[2][2];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) a[x][y] = typename mfma<T>::acc_t{};
    }
};
__device__ __forceinline__ int acc_row(int wave) { return (wave >> 1) * 32; }
__device__ __forceinline__ int acc_col(int wave) { return (wave & 1) * 32; }
template <typename T, typename F> __device__ __forceinline__ void acc_each(F f)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                f(x, y, v, acc_row(wave) + 16 * x + mfma<T>::row(lane, v), acc_col(wave) + 16 * y + (lane & 15));
}

// acc += P Q (64 x 64 x 64, 256 threads).  P and Q are read through loaders,
// every load of both K-chunks issued before the first LDS store (one memory
// round trip per product), then per chunk of KT the four waves run their
// 2 x 2 MFMA tiles from LDS.
template <typename T, typename LP, typename LQ>
__device__ __forceinline__ void tile_mm(Acc<T> &acc, T *sa, T *sb, LP loadP, LQ loadQ)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = acc_row(wave), wc = acc_col(wave);
    constexpr int PER = NB * KT / 256, NCH = NB / KT;
    T pv[NCH][PER], qv[NCH][PER];
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int t = tid + 256 * r;
            pv[c][r] = loadP(t / KT, c * KT + t % KT);
            qv[c][r] = loadQ(c * KT + t / NB, t % NB);
        }
#pragma unroll
    for (int c = 0; c < NCH; c++) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int t = tid + 256 * r;
            sa[(t / KT) * SA_LD + t % KT] = pv[c][r];
            sb[(t / NB) * SB_LD + t % NB] = qv[c][r];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KT; kk += 4) {
            T av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; x++) av[x] = sa[(wr + 16 * x + (lane & 15)) * SA_LD + kk + (lane >> 4)];
#pragma unroll
            for (int y = 0; y < 2; y++) bv[y] = sb[(kk + (lane >> 4)) * SB_LD + wc + 16 * y + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) acc.a[x][y] = mfma<T>::run(av[x], bv[y], acc.a[x][y]);
        }
        __syncthreads();
    }
}

// acc += P Q with both operands in LDS (64 x 64, row strides lp and lq); no
// barrier inside.
template <typename T> __device__ __forceinline__ void lds_mm(Acc<T> &acc, const T *P, int lp, const T *Q, int lq)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = acc_row(wave), wc = acc_col(wave);
#pragma unroll 4
    for (int k = 0; k < NB; k += 4) {
        T av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; x++) av[x] = P[(wr + 16 * x + (lane & 15)) * lp + k + (lane >> 4)];
#pragma unroll
        for (int y = 0; y < 2; y++) bv[y] = Q[(k + (lane >> 4)) * lq + wc + 16 * y + (lane & 15)];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) acc.a[x][y] = mfma<T>::run(av[x], bv[y], acc.a[x][y]);
    }
}

template <typename T> __device__ __forceinline__ T rcp_of(T p);
template <> __device__ __forceinline__ double rcp_of<double>(double p) { return pivot_rcp(p); }
template <> __device__ __forceinline__ float rcp_of<float>(float p) { return pivot_rcp(p); }

template <typename T> struct PivotGuard {
    T rel, huge;
    __device__ __forceinline__ PivotGuard() : rel(sizeof(T) == 8 ? T(1e-13) : T(1e-6)), huge(sizeof(T) == 8 ? T(1e300) : T(1e37)) {}
};

// Cross-lane moves of the 16 x 16 inverse.  A wavefront is four DPP rows of
// 16 lanes.  row_newbcast:K (DPP) gives every lane the value of lane K of its
// own row; rowcast<R> gives every lane the value at the same position of row R
// (v_permlane16_swap: each pair of rows swaps its odd row with the even one,
// so one of the two results holds the pair's even row twice, the other its
// odd row; v_permlane32_swap does the same for the two halves).  Register
// moves only: no LDS round trip on the pivot chain.
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v)
{
    // one 64-bit move where the control allows it (row_newbcast), else two 32-bit ones
    return __builtin_amdgcn_update_dpp(0.0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int R> __device__ __forceinline__ unsigned rowcast_u32(unsigned v)
{
    const auto h = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    const unsigned t = (R & 1) ? h[1] : h[0];
    const auto w = __builtin_amdgcn_permlane32_swap(t, t, false, false);
    return (R & 2) ? w[1] : w[0];
}
template <int R> __device__ __forceinline__ double rowcast(double v)
{
    const unsigned lo = rowcast_u32<R>((unsigned)__double2loint(v)), hi = rowcast_u32<R>((unsigned)__double2hiint(v));
    return __hiloint2double((int)hi, (int)lo);
}
template <int R> __device__ __forceinline__ float rowcast(float v)
{
    return __uint_as_float(rowcast_u32<R>(__float_as_uint(v)));
}
// Pivot K of the in-place Gauss-Jordan inverse of a 16 x 16 block held by one
// wavefront: lane 16 q + i holds row i, columns 4 q .. 4 q + 3 (a[0..3]).  Rows
// are kept scaled: the true row i is s_i times the stored one, s_i = 1 until
// row i has been the pivot row and the pivot's reciprocal after, so the pivot
// row itself is left as it is (its diagonal entry set to 1) and every other
// row takes one FMA per entry, a(i, j) -= (a(i, K) / p) a(K, j); the scales
// are applied once at the end.  Row K is unscaled when it is the pivot row, so
// its entries and the pivot are true values.  The pivot reaches every lane by
// two register moves (column K of the lane's row from row block K / 4, then
// lane K of each DPP row): no readlane, no LDS access on the pivot chain.  A
// pivot below a threshold eps_K (|p| / eps_K < 1; ieps[K] = 1 / eps_K in LDS,
// read per pivot off the dependent chain) takes the reciprocal of +-eps_K, a
// non-finite one the reciprocal 1.
template <typename T, int K>
__device__ __forceinline__ void inv16_pivot(T (&a)[4], T &sc, int i, int q, const T *ieps, T huge, bool &bad,
                                            bool &pert)
{
    constexpr int KQ = K >> 2, KE = K & 3;
    const T ie = ieps[K];
    T rk[4];
#pragma unroll
    for (int e = 0; e < 4; e++) rk[e] = dpp_mov<0x150 + K>(a[e]);  // row K, this lane's columns
    const T f = rowcast<KQ>(a[KE]);                                  // row i, column K
    const T p = dpp_mov<0x150 + K>(f);                               // the pivot
    const T r0 = rcp_of<T>(p);  // replaced below when p is tiny or not finite
    const T ap = fabs(p);
    const bool nonfinite = !(ap < huge);
    const bool small = !nonfinite && ap * ie < T(1);
    bad |= nonfinite;
    pert |= small;
    // rcp(-x) = -rcp(x) exactly, rcp(1) = 1
    const T ip = nonfinite ? T(1) : small ? (p >= T(0) ? ie : -ie) : r0;
    const bool isk = i == K;
    const T h = isk ? T(0) : f * ip;  // row K keeps its entries
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const T upd = fma(-h, rk[e], a[e]);
        if (e == KE)
            a[e] = q == KQ ? (isk ? T(1) : -h) : upd;  // column K: -a(i, K) / p; s_K / s_K on the diagonal
        else
            a[e] = upd;
    }
    sc = isk ? ip : sc;
    if constexpr (K + 1 < 16) inv16_pivot<T, K + 1>(a, sc, i, q, ieps, huge, bad, pert);
}

// One wavefront: in-place Gauss-Jordan inverse (no interchanges) of the 16 x 16
// block at b (stride LDT); the 16 pivots run in registers with DPP and
// permlane moves, no barrier and no LDS access on the pivot chain.  ieps: the
// reciprocals of the sixteen pivot thresholds (LDS).
template <typename T>
__device__ __forceinline__ void inv16(T *b, const T *ieps, T huge, bool &bad, bool &pert)
{
    const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
    T a[4], sc = T(1);
#pragma unroll
    for (int e = 0; e < 4; e++) a[e] = b[i * LDT + 4 * q + e];
    inv16_pivot<T, 0>(a, sc, i, q, ieps, huge, bad, pert);
#pragma unroll
    for (int e = 0; e < 4; e++) b[i * LDT + 4 * q + e] = sc * a[e];
}

// acc = P Q for 16 x 16 blocks in LDS (strides lp, lq); one wavefront.
template <typename T>
__device__ __forceinline__ typename mfma<T>::acc_t mm16(const T *P, int lp, const T *Q, int lq)
{
    const int lane = threadIdx.x & 63;
    typename mfma<T>::acc_t acc{};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma<T>::run(P[(lane & 15) * lp + k0 + (lane >> 4)], Q[(k0 + (lane >> 4)) * lq + (lane & 15)], acc);
    return acc;
}

// Scratch of inverse64: two sets of three saved 16 x 16 column blocks.
constexpr int INV_TMP = 2 * 3 * 16 * TMP_LD;

// In-place inverse of the 64 x 64 matrix M ([64][LDT]) by 256 threads: block
// Gauss-Jordan over 16 x 16 blocks.  Step kb, with the pivot block P already
// inverted in place and the column blocks C_i = M(i, kb) (i != kb) saved: waves
// 0-2 form the row blocks R_j = P^-1 M(kb, j); then the four waves update
// M(i, j) -= C_i R_j and set M(i, kb) = -C_i P^-1 (MFMA 16x16x4).  The next
// pivot block is updated first, by wave 0, which then inverts it (inv16) while
// waves 1-3 finish the step's other blocks and save the next step's column
// blocks: two barriers per step.  No interchanges: a pivot smaller than 1e-13
// (FP32: 1e-6) of its row's original diagonal entry (of the block's largest
// entry where that is zero) -- a threshold that does not depend on how rows and
// columns are scaled -- is replaced by that size with its sign (bit 2 of
// *flag); a non-finite pivot or an all-zero block sets bit 1.  (An FP32 factor
// with perturbed pivots is still refined against the FP64 tangent, which
// decides whether it was good enough.)  tmp: INV_TMP entries; aux: 4 + 64
// entries.  Runs on the first four wavefronts; those of a larger workgroup only
// take part in the barriers.
template <typename T> __device__ __forceinline__ void inverse64(T *M, T *tmp, T *aux, int *flag)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool act = tid < 256;
    const PivotGuard<T> pg;
    // largest magnitude in the block, and the original diagonal
    if (act) {
        const int r = tid >> 2, c0 = 16 * (tid & 3);
        T mx = T(0);
#pragma unroll
        for (int j = 0; j < 16; j++) mx = fmax(mx, fabs(M[r * LDT + c0 + j]));
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));
        if (lane == 0) aux[wave] = mx;
        if (tid < NB) aux[4 + tid] = fabs(M[tid * LDT + tid]);
    }
    __syncthreads();
    const T mx = fmax(fmax(aux[0], aux[1]), fmax(aux[2], aux[3]));
    bool bad = !(mx > T(0)) || !(mx < pg.huge), pert = false;
    // wave 0: pivot thresholds of block kb (their reciprocals replace the
    // original diagonal in aux; written and read by this wave only), then the
    // inverse of the block in place
    auto pivot_block = [&](int kb) {
        T *ie = aux + 4 + 16 * kb;
        if (lane < 16) {
            const T dk = ie[lane];
            ie[lane] = rcp_of<T>(pg.rel * (dk > T(0) ? dk : mx));
        }
        inv16<T>(M + 16 * kb * LDT + 16 * kb, ie, pg.huge, bad, pert);
    };
    constexpr int CS = 16 * TMP_LD, CB = 3 * CS;  // one saved block, one set
    // step 0's pivot block and column blocks
    if (wave == 0) {
        pivot_block(0);
    } else if (act) {
        const T *src = M + 16 * wave * LDT;
        for (int t = lane; t < 256; t += 64) tmp[(wave - 1) * CS + (t >> 4) * TMP_LD + (t & 15)] = src[(t >> 4) * LDT + (t & 15)];
    }
    __syncthreads();
    for (int kb = 0; kb < 4; kb++) {
        const T *cur = tmp + (kb & 1) * CB;
        T *nxt = tmp + ((kb + 1) & 1) * CB;
        T *P = M + 16 * kb * LDT + 16 * kb;
        if (wave < 3) {
            T *R = M + 16 * kb * LDT + 16 * (wave < kb ? wave : wave + 1);
            const typename mfma<T>::acc_t acc = mm16<T>(P, LDT, R, LDT);
#pragma unroll
            for (int v = 0; v < 4; v++) R[mfma<T>::row(lane, v) * LDT + (lane & 15)] = acc[v];
        }
        __syncthreads();
        // the twelve block updates u = 4 ri + j (row block i = ri, skipping kb);
        // with a next step, wave 0 takes the next pivot block (u0) and the other
        // eleven go to waves 1-3 in turn, else three to each wave
        const int nb = kb + 1, u0 = 5 * kb + 1;
        auto update = [&](int u) {
            const int ri = u >> 2, j = u & 3, i = ri < kb ? ri : ri + 1;
            T *Mij = M + 16 * i * LDT + 16 * j;
            const T *Ci = cur + ri * CS;
            const bool piv = j == kb;
            const typename mfma<T>::acc_t acc = mm16<T>(Ci, TMP_LD, piv ? P : M + 16 * kb * LDT + 16 * j, LDT);
            // the next step's column block of row block i (i != nb)
            T *sv = (nb < 4 && j == nb && i != nb) ? nxt + (i < nb ? i : i - 1) * CS : nullptr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = mfma<T>::row(lane, v), c = lane & 15;
                T &e = Mij[r * LDT + c];
                const T x = piv ? -acc[v] : e - acc[v];
                e = x;
                if (sv) sv[r * TMP_LD + c] = x;
            }
        };
        if (nb < 4) {
            if (wave == 0) {
                update(u0);
                pivot_block(nb);
            } else if (act) {
                for (int k = wave - 1; k < 11; k += 3) update(k < u0 ? k : k + 1);
                if (wave == 3) {
                    // row block kb of the next column blocks: R_nb, final since the products above
                    const T *src = M + 16 * kb * LDT + 16 * nb;
                    T *dst = nxt + kb * CS;
                    for (int t = lane; t < 256; t += 64) dst[(t >> 4) * TMP_LD + (t & 15)] = src[(t >> 4) * LDT + (t & 15)];
                }
            }
        } else if (act) {
            for (int u = wave; u < 12; u += 4) update(u);
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int bits = (bad ? 1 : 0) | (pert ? 2 : 0);
        if (bits) atomicOr(flag, bits);
    }
}

// Row-major 64 x 64 tile through registers: thread t holds elements t + 256 r
// (each row of 64 read or written by one 64-lane wave, coalesced).
template <typename T, typename F> __device__ __forceinline__ void tile_load(T (&v)[16], F load)
{
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int t = threadIdx.x + 256 * r;
        v[r] = load(t >> 6, t & 63);
    }
}
template <typename T> __device__ __forceinline__ void tile_put(T *dst, int ld, const T (&v)[16])
{
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int t = threadIdx.x + 256 * r;
        dst[(t >> 6) * ld + (t & 63)] = v[r];
    }
}

// LDS of the step kernel: P holds A operands (E_p, then L(i, p); the diagonal
// block D for the inverse), Q B operands (A'(p, j), then U(p, j)); the
// inverse's scratch reuses Q.  75 KB in FP64, so two workgroups share a CU and
// a launch's tiles are all resident at once.
template <typename T> struct StepSmem {
    T P[NB * LDT];
    union {
        T Q[NB * LDQ];
        T tmp[INV_TMP];
    };
    T aux[4 + NB];
};

// Wavefront sum.
template <typename V> __device__ __forceinline__ V sum64(V s)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o);
    return s;
}

}  // namespace
