// Wide-band tangent solve on gfx950 (shell meshes): the solve that replaces
// sparsesystem -> MKL PARDISO (src/SparseMatrix.cpp:74-166) when the
// half-bandwidth does not fit the block-tridiagonal solver.
//
// Storage (GM_BAND): entry (r, c) with -kl <= c - r <= ku at ab[r*ld + c + kl],
// ld = kl + ku, so any sub-matrix inside the band is a strided matrix with
// leading dimension ld.  The assembled tangent has kl = ku = hb; the stored
// band has ku = hb + 63, the fill of the block factorisation's upper strips.
// Entries outside the band alias neighbouring rows: every load masks them to
// zero and every store skips them.
//
// Factorisation: block LU over 64 x 64 blocks, no interchanges between blocks
// (PARDISO's mtype 11 default is static pivoting with perturbation),
//   A = L U,  L block lower with L(p,p) = D_p,  U block unit upper,
//   E_p = D_p^-1 (Gauss-Jordan in one workgroup),  U(p,j) = E_p A'(p,j),
//   A(i,j) -= A(i,p) U(p,j)                      (MFMA 64 x 64 tiles),
// so L(i,p) is the updated A(i,p) itself (no triangular solve on the panel).
// The band stays closed under the update: U(p, .) reaches at most 63 columns
// past the band of its first row (hence ku), and L U adds nothing outside
// -hb..hb.  The upper strip of row p is kept as A'(p,j) (updated, before the
// multiplication by E_p) while the factorisation needs it, and every update
// tile forms U(p,j) = E_p A'(p,j) itself, so one launch per block column p
// holds all of step p:
//   update    A(i,j) -= A(i,p) (E_p A'(p,j))  for i, j in p+1 .. p+R (two
//             64^3 MFMA products per tile, operands staged through LDS);
//   diagonal  the same for (p+1, p+1), then E_{p+1} = D^-1 (one workgroup:
//             the launch's critical path is one memory round trip, two
//             products and the inverse);
//   convert   A'(p-1,j) := E_{p-1} A'(p-1,j), the strip that step p-1 was
//             the last to read.
// No workgroup waits on another: the groups touch disjoint blocks and read
// only what earlier launches wrote, so stream order is the only
// synchronisation.  A pivot below 1e-13 of its row's original diagonal entry
// is perturbed to that size (flag bit 2); a non-finite one fails (bit 1).
//
// Triangular solves: the diagonal blocks are grouped in batches of SB = 8;
// after the factorisation each batch's diagonal part of L is inverted
// (G = L_bb^-1, lower, E_p on its diagonal) and of U (H = U_bb^-1, unit
// upper), one launch per block distance.  A solve then takes, per batch, one
// launch that applies the batch inverse (all rows in parallel) and one that
// subtracts the batch's couplings from the rows within reach.
#include "gf_internal.h"

#include <algorithm>

namespace {

constexpr int NB = 64;   // block size
constexpr int SB = 8;    // blocks per solve batch
constexpr int BW = NB * SB;
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

struct Geo {
    int n;       // padded order (multiple of NB)
    int kl, ku;  // stored band
    int ld;      // kl + ku
    int nbk;     // n / NB
    int R;       // reach in blocks of L below and U right of the diagonal block
};

__device__ __forceinline__ bool inband(int r, int c, const Geo &g) { return r - c <= g.kl && c - r <= g.ku; }

// Masked band load without a branch around it: the address is clamped into
// the stored band of row min(max(r, 0), n - 1) (inside the n (ld + 1)
// allocation) and the value selected afterwards, so a thread's loads are all
// in flight together.
template <typename T> __device__ __forceinline__ T band_ld(const T *ab, int r, int c, const Geo &g)
{
    const int rr = min(max(r, 0), g.n - 1);
    const int cc = min(max(c, rr - g.kl), rr + g.ku);
    const T v = ab[(size_t)rr * (size_t)g.ld + (size_t)(cc + g.kl)];
    return r >= 0 && r < g.n && c >= 0 && c < g.n && inband(r, c, g) ? v : T(0);
}
template <typename T> __device__ __forceinline__ void band_st(T *ab, int r, int c, const Geo &g, T v)
{
    if (r >= 0 && r < g.n && c >= 0 && c < g.n && inband(r, c, g)) ab[(size_t)r * (size_t)g.ld + (size_t)(c + g.kl)] = v;
}

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
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
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
// wavefront: lane 16 q + i holds row i, columns 4 q .. 4 q + 3 (a[0..3]).  The
// pivot reaches every lane by two register moves (column K of the lane's row
// from row block K / 4, then lane K of each DPP row), so the chain from one
// pivot to the next is VALU only: no readlane, no scalar selects.  A pivot
// below eps[K] takes the reciprocal of +-eps[K] (ieps, computed before the
// first pivot), a non-finite one the reciprocal 1.
template <typename T, int K>
__device__ __forceinline__ void inv16_pivot(T (&a)[4], int i, int q, const T (&eps)[16], const T (&ieps)[16],
                                            T huge, bool &bad, bool &pert)
{
    constexpr int KQ = K >> 2, KE = K & 3;
    T rk[4];
#pragma unroll
    for (int e = 0; e < 4; e++) rk[e] = dpp_mov<0x150 + K>(a[e]);  // row K, this lane's columns
    const T f = rowcast<KQ>(a[KE]);                                  // row i, column K
    const T p = dpp_mov<0x150 + K>(f);                               // the pivot
    const T ap = fabs(p);
    const bool nonfinite = !(ap < huge);
    const bool small = !nonfinite && ap < eps[K];
    bad |= nonfinite;
    pert |= small;
    const T ip = nonfinite ? T(1) : small ? (p >= T(0) ? ieps[K] : -ieps[K]) : rcp_of<T>(p);
    const bool isk = i == K;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const T r = rk[e] * ip;
        const T upd = isk ? r : a[e] - f * r;
        a[e] = (4 * q + e == K) ? (isk ? ip : -f * ip) : upd;
    }
    if constexpr (K + 1 < 16) inv16_pivot<T, K + 1>(a, i, q, eps, ieps, huge, bad, pert);
}

// One wavefront: in-place Gauss-Jordan inverse (no interchanges) of the 16 x 16
// block at b (stride LDT); the 16 pivots run in registers with DPP and
// permlane moves, no barrier and no LDS access between the load and the store.
// eps[k]: the perturbation size of pivot k.
template <typename T>
__device__ __forceinline__ void inv16(T *b, const T (&eps)[16], T huge, bool &bad, bool &pert)
{
    const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
    T a[4], ieps[16];
#pragma unroll
    for (int e = 0; e < 4; e++) a[e] = b[i * LDT + 4 * q + e];
#pragma unroll
    for (int k = 0; k < 16; k++) ieps[k] = rcp_of<T>(eps[k]);
    inv16_pivot<T, 0>(a, i, q, eps, ieps, huge, bad, pert);
#pragma unroll
    for (int e = 0; e < 4; e++) b[i * LDT + 4 * q + e] = a[e];
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

// In-place inverse of the 64 x 64 matrix M ([64][LDT]) by 256 threads: block
// Gauss-Jordan over 16 x 16 blocks.  For each diagonal block kb: wave 0 inverts
// the pivot block P (inv16) while waves 1-3 save the column blocks C_i; waves
// 0-2 form the row blocks R_j = P^-1 M(kb, j); then the four waves update
// M(i, j) -= C_i R_j and set M(i, kb) = -C_i P^-1 (MFMA 16x16x4), three
// barriers per block.  No interchanges: a pivot smaller than 1e-13 (FP32:
// 1e-6) of its row's original diagonal entry (of the block's largest entry
// where that is zero) -- a threshold that does not depend on how rows and
// columns are scaled -- is replaced by that size with its sign (bit 2 of
// *flag); a non-finite pivot or an all-zero block sets bit 1.  (An FP32 factor
// with perturbed pivots is still refined against the FP64 tangent, which
// decides whether it was good enough.)  tmp: 3 x 16 x TMP_LD entries; aux:
// 4 + 64 entries.
template <typename T> __device__ __forceinline__ void inverse64(T *M, T *tmp, T *aux, int *flag)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const PivotGuard<T> pg;
    // largest magnitude in the block, and the original diagonal
    {
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
    for (int kb = 0; kb < 4; kb++) {
        T *P = M + 16 * kb * LDT + 16 * kb;
        if (wave == 0) {
            T eps[16];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const T dk = aux[4 + 16 * kb + k];
                eps[k] = pg.rel * (dk > T(0) ? dk : mx);
            }
            inv16<T>(P, eps, pg.huge, bad, pert);
        } else {
            const int ri = wave - 1, i = ri < kb ? ri : ri + 1;
            const T *src = M + 16 * i * LDT + 16 * kb;
            for (int t = lane; t < 256; t += 64) tmp[ri * 16 * TMP_LD + (t >> 4) * TMP_LD + (t & 15)] = src[(t >> 4) * LDT + (t & 15)];
        }
        __syncthreads();
        if (wave < 3) {
            T *R = M + 16 * kb * LDT + 16 * (wave < kb ? wave : wave + 1);
            const typename mfma<T>::acc_t acc = mm16<T>(P, LDT, R, LDT);
#pragma unroll
            for (int v = 0; v < 4; v++) R[mfma<T>::row(lane, v) * LDT + (lane & 15)] = acc[v];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 3; s++) {
            const int u = wave + 4 * s, ri = u >> 2, cj = u & 3;
            const int i = ri < kb ? ri : ri + 1;
            T *Mij = M + 16 * i * LDT + 16 * cj;
            const T *Ci = tmp + ri * 16 * TMP_LD;
            const bool piv = cj == kb;
            const typename mfma<T>::acc_t acc = mm16<T>(Ci, TMP_LD, piv ? P : M + 16 * kb * LDT + 16 * cj, LDT);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T &e = Mij[mfma<T>::row(lane, v) * LDT + (lane & 15)];
                e = piv ? -acc[v] : e - acc[v];
            }
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
        T tmp[3 * 16 * TMP_LD];
    };
    T aux[4 + NB];
};

// LDS of the batch-inverse kernel: the staging of tile_mm, reused as X (a
// [64][LDT] operand) once the products that stage through it are done.
template <typename T> struct BatchSmem {
    union {
        T stage[NB * SA_LD + KT * SB_LD];
        T X[NB * LDT];
    };
    T D[NB * LDT];
    T tmp[3 * 16 * TMP_LD];
    T aux[4 + NB];
};
static_assert(NB * SA_LD + KT * SB_LD >= NB * LDT, "X fits in the staging area");

// Batch store: per batch b, a BW x BW row-major matrix; block (r, c) holds G(r, c)
// for c <= r (G = L_bb^-1, E on the diagonal) and H(r, c) for c > r (H = U_bb^-1).
template <typename T> __device__ __forceinline__ T *batch_blk(T *gh, int b, int r, int c)
{
    return gh + (size_t)b * BW * BW + (size_t)(r * NB) * BW + c * NB;
}

// Tile of workgroup b in launch p: the update tiles (i, j), i, j in
// p+1 .. p+nr (for p = -1 only the diagonal tile (0, 0)), then the nv
// conversion tiles of row p-1 (columns p ..).  Workgroups are dispatched
// round-robin over the 8 XCDs (b mod 8), and row i is always given to XCD
// i mod 8, so a band row's tiles are updated by the same XCD at every step
// (a row of the update window is 22 tiles, ~0.7 MB in FP64); within an XCD the
// rows are taken in order, so the diagonal tile (p+1, p+1) -- the only one
// with a dependent inverse -- is the first workgroup of its XCD.  kind: 0
// diagonal, 1 update, 2 conversion.  Returns false past the XCD's share.
__device__ __forceinline__ bool step_tile(int b, int p, int nr, int nv, int &kind, int &bi, int &bj)
{
    const int x = b & 7;
    int k = b >> 3;
    const int ncol = p < 0 ? 1 : nr;
    for (int i = p + 1 + (((x - (p + 1)) % 8) + 8) % 8; i <= p + nr; i += 8) {
        if (k < ncol) {
            bi = i;
            bj = p + 1 + k;
            kind = bi == p + 1 && bj == p + 1 ? 0 : 1;
            return true;
        }
        k -= ncol;
    }
    if (nv > 0 && ((p - 1) & 7) == x && k < nv) {
        bi = p - 1;
        bj = p + k;
        kind = 2;
        return true;
    }
    return false;
}

// Launch p of the factorisation (see the file header):
//   update    A(i, j) -= A(i, p) (E_p A'(p, j))            i, j in p+1 .. p+R
//   diagonal  the same for (p+1, p+1), then E_{p+1} = D^-1 into the batch store
//   convert   A'(p-1, j) := U(p-1, j) = E_{p-1} A'(p-1, j)  (no later launch
//             reads A'(p-1, .), and no workgroup of this one touches row p-1)
// Every operand is final before the launch starts, so stream order is the only
// synchronisation.  Each workgroup issues all of its loads before the first
// LDS store (one memory round trip), then runs the products from LDS.
template <typename T>
__global__ __launch_bounds__(256) void band_step_kernel(T *ab, Geo g, int p, int nr, int nv, T *gh, int *flag)
{
    __shared__ StepSmem<T> sm;
    int kind, ti, tj;
    if (!step_tile(blockIdx.x, p, nr, nv, kind, ti, tj)) return;
    // the diagonal tile is the launch's critical path: its waves go first when
    // they share a CU with an update tile
    if (kind == 0) __builtin_amdgcn_s_setprio(3);
    const int tid = threadIdx.x;
    const int R0 = ti * NB, C0 = tj * NB;
    Acc<T> acc;
    acc.zero();
    if (kind == 2) {
        T e[16], a[16];
        const T *E = batch_blk(gh, ti / SB, ti % SB, ti % SB);
        tile_load(e, [&](int i, int k) { return E[(size_t)i * BW + k]; });
        tile_load(a, [&](int k, int j) { return band_ld(ab, R0 + k, C0 + j, g); });
        tile_put(sm.P, LDT, e);
        tile_put(sm.Q, LDQ, a);
        __syncthreads();
        lds_mm<T>(acc, sm.P, LDT, sm.Q, LDQ);
        acc_each<T>([&](int x, int y, int v, int i, int j) { band_st(ab, R0 + i, C0 + j, g, acc.a[x][y][v]); });
        return;
    }
    T old[2][2][4];
    if (p >= 0) {
        // E_p and A'(p, j) first (32 loads per thread, one round trip); L(i, p)
        // and the old tile are issued behind the first product, which hides
        // their latency (a wave keeps at most 63 loads in flight)
        const int K0 = p * NB;
        T e[16], a[16], l[16];
        const T *E = batch_blk(gh, p / SB, p % SB, p % SB);
        tile_load(e, [&](int i, int k) { return E[(size_t)i * BW + k]; });
        tile_load(a, [&](int k, int j) { return band_ld(ab, K0 + k, C0 + j, g); });
        tile_put(sm.P, LDT, e);
        tile_put(sm.Q, LDQ, a);
        __syncthreads();
        tile_load(l, [&](int i, int k) { return band_ld(ab, R0 + i, K0 + k, g); });
        acc_each<T>([&](int x, int y, int v, int i, int j) { old[x][y][v] = band_ld(ab, R0 + i, C0 + j, g); });
        Acc<T> w;
        w.zero();
        lds_mm<T>(w, sm.P, LDT, sm.Q, LDQ);  // U(p, j)
        __syncthreads();
        acc_each<T>([&](int x, int y, int v, int i, int j) { sm.Q[i * LDQ + j] = w.a[x][y][v]; });
        tile_put(sm.P, LDT, l);
        __syncthreads();
        lds_mm<T>(acc, sm.P, LDT, sm.Q, LDQ);  // L(i, p) U(p, j)
    } else {
        acc_each<T>([&](int x, int y, int v, int i, int j) { old[x][y][v] = band_ld(ab, R0 + i, C0 + j, g); });
    }
    if (kind == 1) {
        acc_each<T>([&](int x, int y, int v, int i, int j) {
            band_st(ab, R0 + i, C0 + j, g, old[x][y][v] - acc.a[x][y][v]);
        });
        return;
    }
    __syncthreads();  // the products' LDS reads are done
    acc_each<T>([&](int x, int y, int v, int i, int j) { sm.P[i * LDT + j] = old[x][y][v] - acc.a[x][y][v]; });
    __syncthreads();
    inverse64<T>(sm.P, sm.tmp, sm.aux, flag);  // E_{p+1}
    T *E = batch_blk(gh, ti / SB, ti % SB, ti % SB);
    for (int t = tid; t < NB * NB; t += 256) E[(t / NB) * BW + t % NB] = sm.P[(t / NB) * LDT + t % NB];
}

// Batch inverses for block distance d (1 <= d < SB).  Workgroup (b, c, kind):
//   kind 0:  G(c+d, c) = -E_{c+d} sum_{q=c}^{c+d-1} L(c+d, q) G(q, c)
//   kind 1:  H(c, c+d) = -(U(c, c+d) + sum_{q=c+1}^{c+d-1} U(c, q) H(q, c+d))
// (block indices inside batch b; L and U from the factored band).
template <typename T>
__global__ __launch_bounds__(256) void band_batch_inv_kernel(const T *ab, Geo g, int d, T *gh)
{
    __shared__ BatchSmem<T> sm;
    T *sa = sm.stage, *sb = sm.stage + NB * SA_LD;
    const int per = SB - d;
    const int w = blockIdx.x % (2 * per), b = blockIdx.x / (2 * per);
    const int kind = w / per, c = w % per;
    const int kc = min(SB, g.nbk - b * SB);
    if (c + d >= kc) return;
    const int B0 = b * SB * NB;  // first row / column of the batch
    Acc<T> acc;
    acc.zero();
    if (kind == 0) {
        const int r = c + d;
        for (int q = c; q < r; q++) {
            const T *Gq = batch_blk(gh, b, q, c);
            tile_mm<T>(acc, sa, sb, [&](int i, int k) { return band_ld(ab, B0 + r * NB + i, B0 + q * NB + k, g); },
                       [&](int k, int j) { return Gq[(size_t)k * BW + j]; });
        }
        acc_each<T>([&](int x, int y, int v, int i, int j) { sm.X[i * LDT + j] = acc.a[x][y][v]; });
        const T *Er = batch_blk(gh, b, r, r);
        for (int t = threadIdx.x; t < NB * NB; t += 256) sm.D[(t / NB) * LDT + t % NB] = Er[(size_t)(t / NB) * BW + t % NB];
        __syncthreads();
        Acc<T> out;
        out.zero();
        lds_mm<T>(out, sm.D, LDT, sm.X, LDT);
        T *Gr = batch_blk(gh, b, r, c);
        acc_each<T>([&](int x, int y, int v, int i, int j) { Gr[(size_t)i * BW + j] = -out.a[x][y][v]; });
    } else {
        const int r = c, cc = c + d;
        for (int q = r + 1; q < cc; q++) {
            const T *Hq = batch_blk(gh, b, q, cc);
            tile_mm<T>(acc, sa, sb, [&](int i, int k) { return band_ld(ab, B0 + r * NB + i, B0 + q * NB + k, g); },
                       [&](int k, int j) { return Hq[(size_t)k * BW + j]; });
        }
        T *Hr = batch_blk(gh, b, r, cc);
        acc_each<T>([&](int x, int y, int v, int i, int j) {
            Hr[(size_t)i * BW + j] = -(acc.a[x][y][v] + band_ld(ab, B0 + r * NB + i, B0 + cc * NB + j, g));
        });
    }
}

// ---- triangular solves -----------------------------------------------------------
// One wavefront per row: lane l takes columns l, l + 64, ... of the batch's 512,
// so each load instruction of a wave reads 64 consecutive entries of the row
// (coalesced) and every lane's loads are independent; the row sum is a
// wavefront reduction.
template <typename V> __device__ __forceinline__ V sum64(V s)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o);
    return s;
}

// FWD: y = G f over batch b (out = z); BWD: x = H g, H unit upper (out = x).
// One row per workgroup of one wavefront; rows of block row rb use block
// columns [0, rb] (FWD) or (rb, kc) (BWD) of the batch store.
template <typename T, typename V, bool FWD>
__global__ __launch_bounds__(64) void band_batch_apply_kernel(const T *gh, Geo g, int b, const V *rhs, V *out)
{
    const int kc = min(SB, g.nbk - b * SB), B0 = b * BW;
    const int row = blockIdx.x, l = threadIdx.x;  // grid = kc * NB
    const int rb = row / NB;
    const T *M = gh + (size_t)b * BW * BW + (size_t)row * BW;
    const int c_lo = FWD ? 0 : rb + 1, c_hi = FWD ? rb + 1 : kc;  // block columns
    T m[SB];
    V y[SB];
#pragma unroll
    for (int cb = 0; cb < SB; cb++) {
        const bool use = cb >= c_lo && cb < c_hi;
        m[cb] = use ? M[cb * NB + l] : T(0);
        y[cb] = use ? rhs[B0 + cb * NB + l] : V(0);
    }
    V s = V(0);
#pragma unroll
    for (int cb = 0; cb < SB; cb++) s += (V)m[cb] * y[cb];
    s = sum64(s);
    if (l == 0) out[B0 + row] = FWD ? s : rhs[B0 + row] + s;
}

// FWD: f_q -= L(q, batch) y_batch for the rows q after batch b within reach;
// BWD: g_q -= U(q, batch) x_batch for the rows before it.  CPL_ROWS rows per
// workgroup, one wavefront each.
constexpr int CPL_ROWS = 4;
template <typename T, typename V, bool FWD>
__global__ __launch_bounds__(64 * CPL_ROWS) void band_couple_kernel(const T *ab, Geo g, int b, const V *sol, V *f, int q0)
{
    const int kc = min(SB, g.nbk - b * SB), B0 = b * BW;
    const int row = q0 * NB + blockIdx.x * CPL_ROWS + (threadIdx.x >> 6), l = threadIdx.x & 63;
    T m[SB];
    V y[SB];
#pragma unroll
    for (int cb = 0; cb < SB; cb++) {
        const bool use = cb < kc;
        m[cb] = use ? band_ld(ab, row, B0 + cb * NB + l, g) : T(0);
        y[cb] = use ? sol[B0 + cb * NB + l] : V(0);
    }
    V s = V(0);
#pragma unroll
    for (int cb = 0; cb < SB; cb++) s += (V)m[cb] * y[cb];
    s = sum64(s);
    if (l == 0 && row < g.n) f[row] -= s;
}

}  // namespace

// ---- host -----------------------------------------------------------------------
int gf_band_ku(int hb) { return hb + NB - 1; }

static Geo make_geo(int n, int hb)
{
    Geo g;
    g.n = n;
    g.kl = hb;
    g.ku = gf_band_ku(hb);
    g.ld = g.kl + g.ku;
    g.nbk = n / NB;
    g.R = (NB - 1 + hb) / NB;
    return g;
}

template <typename T> static int band_factor_t(T *ab, int n, int hb, hipStream_t s, int *flag, T *gh)
{
    const Geo g = make_geo(n, hb);
    int launches = 0;
    const int last = g.nbk - 1;
    // launch p = -1 .. last: E_{p+1} and the update of step p (p < last), the
    // conversion of row p-1 to U (p >= 1)
    for (int p = -1; p <= last; p++) {
        const int nr = p < 0 ? 1 : p < last ? std::min(last, p + g.R) - p : 0;
        const int nv = p >= 1 ? std::min(last, p - 1 + g.R) - p + 1 : 0;
        if (nr == 0 && nv == 0) continue;
        // tiles per XCD (row i on XCD i mod 8): the grid covers the largest share
        int share[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = p + 1; i <= p + nr; i++) share[i & 7] += p < 0 ? 1 : nr;
        if (nv > 0) share[(p - 1) & 7] += nv;
        const int grid = 8 * *std::max_element(share, share + 8);
        const int pr = gf_probe_begin(s, GF_KC_BAND_UPDATE);
        hipLaunchKernelGGL(band_step_kernel<T>, dim3(grid), dim3(256), 0, s, ab, g, p, nr, nv, gh, flag);
        gf_probe_end(s, pr, nr > 0 ? NB : 0);  // block column p+1 eliminated
        launches++;
    }
    const int nbatch = (g.nbk + SB - 1) / SB;
    for (int d = 1; d < std::min(SB, g.nbk); d++) {
        const int pr = gf_probe_begin(s, GF_KC_BAND_PANEL);
        hipLaunchKernelGGL(band_batch_inv_kernel<T>, dim3(nbatch * 2 * (SB - d)), dim3(256), 0, s, (const T *)ab, g,
                           d, gh);
        // flops of the launch: a G block at distance d takes d + 1 products of
        // 64 x 64 x 64, an H block d of them
        double blocks = 0.0;
        for (int bb = 0; bb < nbatch; bb++) blocks += std::max(0, std::min(SB, g.nbk - bb * SB) - d);
        gf_probe_end(s, pr, blocks * (2.0 * d + 1.0) * 2.0 * NB * NB * NB);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

template <typename T, typename V>
static int band_solve_t(const T *ab, int n, int hb, hipStream_t s, V *x, V *z, const T *gh)
{
    const Geo g = make_geo(n, hb);
    const int nbatch = (g.nbk + SB - 1) / SB;
    int launches = 0;
    // forward: L z = x (x updated in place as the right-hand side)
    for (int b = 0; b < nbatch; b++) {
        const int kc = std::min(SB, g.nbk - b * SB), q0 = b * SB + kc;
        const int q1 = std::min(g.nbk - 1, b * SB + kc - 1 + g.R);
        int pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
        hipLaunchKernelGGL((band_batch_apply_kernel<T, V, true>), dim3(kc * NB), dim3(64), 0, s, gh, g, b,
                           (const V *)x, z);
        gf_probe_end(s, pr, kc * NB);
        launches++;
        if (q1 >= q0) {
            pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
            hipLaunchKernelGGL((band_couple_kernel<T, V, true>), dim3((q1 - q0 + 1) * NB / CPL_ROWS), dim3(64 * CPL_ROWS), 0, s, ab, g,
                               b, (const V *)z, x, q0);
            gf_probe_end(s, pr, 0.0);  // rows are counted on the apply launches
            launches++;
        }
    }
    // backward: U x = z (z updated in place)
    for (int b = nbatch - 1; b >= 0; b--) {
        const int kc = std::min(SB, g.nbk - b * SB);
        const int q1 = b * SB - 1, q0 = std::max(0, b * SB - g.R);
        int pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
        hipLaunchKernelGGL((band_batch_apply_kernel<T, V, false>), dim3(kc * NB), dim3(64), 0, s, gh, g, b,
                           (const V *)z, x);
        gf_probe_end(s, pr, kc * NB);
        launches++;
        if (q1 >= q0) {
            pr = gf_probe_begin(s, GF_KC_BAND_SWEEP);
            hipLaunchKernelGGL((band_couple_kernel<T, V, false>), dim3((q1 - q0 + 1) * NB / CPL_ROWS), dim3(64 * CPL_ROWS), 0, s, ab,
                               g, b, (const V *)x, z, q0);
            gf_probe_end(s, pr, 0.0);  // rows are counted on the apply launches
            launches++;
        }
    }
    GF_HIP(hipGetLastError());
    return launches;
}

size_t gf_band_inv_elems(int n)
{
    const int nbk = n / NB;
    return (size_t)((nbk + SB - 1) / SB) * BW * BW;
}

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
