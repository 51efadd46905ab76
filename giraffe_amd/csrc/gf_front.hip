// Multifrontal tangent solve on gfx950: the nested-dissection counterpart of
// sparsesystem -> MKL PARDISO (src/SparseMatrix.cpp:74-166; PARDISO orders by
// nested dissection, iparm[1] = 3, and factors supernode by supernode).
//
// Fronts (plan: gf_nd.cpp) are dense row-major matrices [mp][mp] over a
// supernode's own DOFs (kp = 64 kb rows, the last kp - k of them identity
// padding) followed by its update DOFs.  Per level of the supernode tree:
//
//   assemble  F(i, j) = sum over children c of U_c(a_c(i), a_c(j)) (the child's
//             Schur complement rows that map to i and j, in child order), plus
//             the tangent entry K(i, j) read from the band storage when row i or
//             column j is one of the front's own DOFs (the front of the first of
//             the two in elimination order holds it).  Every entry is written by
//             one thread, in a fixed order: no atomics, bit-reproducible.
//             The workgroup of each front's first diagonal block also inverts it
//             (E_0).
//   factor    the block LU of the band solver on each front, restricted to its
//             kb own block columns, over all fronts of the level at once:
//             launch p = 0 .. kb - 1 updates A(i, j) -= A(i, p) (E_p A'(p, j))
//             for i, j > p (the update blocks included: after the last own step
//             they hold the Schur complement the parent assembles), inverts the
//             next diagonal block (E_{p+1}) and converts row p - 1 to U.  The
//             last own block row stays A'(kb - 1, .) = D U(kb - 1, .): the
//             backward solve applies E_{kb-1} to it, which saves a launch per
//             level.  Pivots: as the band solver (perturbation below 1e-13 of the
//             block's diagonal, flag bit 2; non-finite, bit 1).
//
// Solve, with A = L U (L(p, p) = D_p = E_p^-1, U block unit upper):
//   forward   per level: a front gathers its own right-hand side and its
//             children's update vectors, solves with its own L blocks (one
//             workgroup, rows kept in LDS), then every update row tile forms
//             its update vector  c = sum_children c_child - L(t, own) y.
//   backward  per level from the root: x_own = y - U(own, update) x_update
//             (row tiles in parallel, x_update read from the ancestors' rows of
//             x), then the own-block back substitution and the scatter of x.
#include "gf_tile.h"

#include <algorithm>

namespace {

constexpr int KMAX = 1024;  // largest own block (rows in LDS of the solve); the plan splits larger supernodes

struct FrontTab {
    const gf_front_dev *fr;
    const int *rows, *child, *inv, *items;
    const long long *child_inv;
    const int *orig, *asm_orig;
};

// Update vector entry of front position i gathered from the children.
__device__ __forceinline__ double child_sum(const FrontTab &t, const gf_front_dev &F, int i, const double *Vb)
{
    double w = 0.0;
    for (int c = F.ch0; c < F.ch1; c++) {
        const int a = t.inv[t.child_inv[c] + i];
        if (a >= 0) {
            const gf_front_dev C = t.fr[t.child[c]];
            w += Vb[C.V + NB * C.kb + a];
        }
    }
    return w;
}

// LDS of the assembly kernel: the first diagonal block of a front and the
// inverse's scratch.
template <typename T> struct AsmSmem {
    T P[NB * LDT];
    T tmp[INV_TMP];
    T aux[4 + NB];
};

// 512 threads (eight wavefronts) per tile: the inverse's 48 KB of LDS then still
// leaves 24 wavefronts per CU to overlap the gathers.
constexpr int ASM_T = 512, ASM_R = NB * NB / ASM_T;

template <typename T>
__global__ __launch_bounds__(ASM_T) void front_assemble_kernel(FrontTab t, int first, const double *ab, int ld, int hb,
                                                               const double *D, T *Fb, T *Eb, int *flag,
                                                               const double *rhs, double *Vb)
{
    __shared__ AsmSmem<T> sm;
    const int *it = t.items + 4 * (size_t)(first + blockIdx.x);
    const int f = it[0];
    const gf_front_dev F = t.fr[f];
    const int r0 = NB * it[1], c0 = NB * it[2];
    const int tid = threadIdx.x, j = tid & 63, i0 = tid >> 6;
    constexpr int RS = ASM_T / NB;  // row stride of a thread's rows
    const int cj = c0 + j;
    const int *rows = t.rows + F.rows;
    T v[ASM_R];
#pragma unroll
    for (int r = 0; r < ASM_R; r++) v[r] = T(0);
    for (int c = F.ch0; c < F.ch1; c++) {
        const gf_front_dev C = t.fr[t.child[c]];
        const int *im = t.inv + t.child_inv[c];
        const int b = im[cj];
        if (b < 0) continue;
        const T *U = Fb + C.F + (size_t)(NB * C.kb) * C.mp + NB * C.kb + b;
#pragma unroll
        for (int r = 0; r < ASM_R; r++) {
            const int a = im[r0 + i0 + RS * r];
            if (a >= 0) v[r] += U[(size_t)a * C.mp];
        }
    }
    // the children's sums into the tile in LDS, the tile's tangent entries (its
    // precomputed list) added there, then the tile written out row by row
#pragma unroll
    for (int r = 0; r < ASM_R; r++) {
        const int ri = r0 + i0 + RS * r;
        if (ri == cj && ri >= F.k && ri < NB * F.kb) v[r] = T(1);  // padding of the own block: identity
        sm.P[(i0 + RS * r) * LDT + j] = v[r];
    }
    __syncthreads();
    const int *ol = t.orig + 3 * (size_t)t.asm_orig[2 * (size_t)(first + blockIdx.x)];
    const int no = t.asm_orig[2 * (size_t)(first + blockIdx.x) + 1];
    for (int q = tid; q < no; q += ASM_T) {
        const int pos = ol[3 * q], oi = ol[3 * q + 1], oc = ol[3 * q + 2];
        double a = ab[(size_t)oi * (size_t)ld + (size_t)(oc + hb)];
        if (D) a = D[oi] * a * D[oc];
        sm.P[(pos >> 6) * LDT + (pos & 63)] += (T)a;
    }
    __syncthreads();
    T *out = Fb + F.F + cj;
#pragma unroll
    for (int r = 0; r < ASM_R; r++) {
        v[r] = sm.P[(i0 + RS * r) * LDT + j];
        out[(size_t)(r0 + i0 + RS * r) * F.mp] = v[r];
    }
    if (rhs && it[2] == 0 && tid < NB) {
        // the right-hand side rides along the factorisation: the front vector starts
        // as the own rows of b plus the children's update vectors
        const int i = r0 + tid, o = rows[i];
        Vb[F.V + i] = (i < F.k && o >= 0 ? rhs[o] : 0.0) + child_sum(t, F, i, Vb);
    }
    if (it[3] == 0) return;
    // the front's first diagonal block (already in LDS): E_0
    inverse64<T>(sm.P, sm.tmp, sm.aux, flag);
    T *E = Eb + F.E;
    for (int q = threadIdx.x; q < NB * NB; q += ASM_T) E[q] = sm.P[(q / NB) * LDT + q % NB];
}

// LDS of the step kernel, 52.6 KB in FP64 so that three workgroups share a CU: X
// holds one full 64 x 64 operand (E_p, then U(p, j) transposed) and C a 32-deep
// chunk of the other (A'(p, j) transposed, then L(i, p)), both at row strides
// for which the MFMA operand reads (16 rows x 4 consecutive entries) hit distinct
// banks; a B operand is kept transposed so that it is read the same way.  The
// inverse's scratch reuses C.
template <typename T> struct FStepSmem {
    T X[NB * LDT];
    union {
        T C[NB * SA_LD];
        T tmp[INV_TMP];
    };
    T aux[4 + NB];
};

// acc += X(:, k0 .. k0+31) * CT^T, X [i][k] (stride LDT), CT [j][k - k0] (stride SA_LD)
template <typename T> __device__ __forceinline__ void mm_xc(Acc<T> &acc, const T *X, const T *CT, int k0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = acc_row(wave), wc = acc_col(wave);
#pragma unroll 2
    for (int kk = 0; kk < KT; kk += 4) {
        T av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; x++) av[x] = X[(wr + 16 * x + (lane & 15)) * LDT + k0 + kk + (lane >> 4)];
#pragma unroll
        for (int y = 0; y < 2; y++) bv[y] = CT[(wc + 16 * y + (lane & 15)) * SA_LD + kk + (lane >> 4)];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) acc.a[x][y] = mfma<T>::run(av[x], bv[y], acc.a[x][y]);
    }
}
// acc += C * XT(:, k0 .. k0+31)^T, C [i][k - k0] (stride SA_LD), XT [j][k] (stride LDT)
template <typename T> __device__ __forceinline__ void mm_cx(Acc<T> &acc, const T *C, const T *XT, int k0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = acc_row(wave), wc = acc_col(wave);
#pragma unroll 2
    for (int kk = 0; kk < KT; kk += 4) {
        T av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; x++) av[x] = C[(wr + 16 * x + (lane & 15)) * SA_LD + kk + (lane >> 4)];
#pragma unroll
        for (int y = 0; y < 2; y++) bv[y] = XT[(wc + 16 * y + (lane & 15)) * LDT + k0 + kk + (lane >> 4)];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) acc.a[x][y] = mfma<T>::run(av[x], bv[y], acc.a[x][y]);
    }
}
// A 32 x 64 row chunk (rows r0 .. r0+31 of a row-major tile at src, stride ld) into
// registers (thread t holds t + 256 r: row (t + 256 r) / 64, column t % 64), and
// into C transposed (C [column][row]).
template <typename T> __device__ __forceinline__ void rowchunk_load(T (&v)[8], const T *src, size_t ld, int r0)
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int t = threadIdx.x + 256 * r;
        v[r] = src[(size_t)(r0 + (t >> 6)) * ld + (t & 63)];
    }
}
template <typename T> __device__ __forceinline__ void rowchunk_put_t(T *C, const T (&v)[8])
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int t = threadIdx.x + 256 * r;
        C[(t & 63) * SA_LD + (t >> 6)] = v[r];
    }
}
// A 64 x 32 column chunk (columns c0 .. c0+31) into registers (thread t holds
// t + 256 r: row (t + 256 r) / 32, column t % 32), and into C as it is.
template <typename T> __device__ __forceinline__ void colchunk_load(T (&v)[8], const T *src, size_t ld, int c0)
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int t = threadIdx.x + 256 * r;
        v[r] = src[(size_t)(t >> 5) * ld + c0 + (t & 31)];
    }
}
template <typename T> __device__ __forceinline__ void colchunk_put(T *C, const T (&v)[8])
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int t = threadIdx.x + 256 * r;
        C[(t >> 5) * SA_LD + (t & 31)] = v[r];
    }
}

// Launch p >= 0 of a level's factorisation (the band step on every front of the
// level; see the file header).  Item kinds: 0 diagonal, 1 update, 2 conversion.
// Each product runs in two 32-deep chunks; all of a workgroup's loads are issued
// before the first product.
template <typename T>
__global__ __launch_bounds__(256, 3) void front_step_kernel(FrontTab t, int first, int p, T *Fb, T *Eb, int *flag,
                                                         double *Vb)
{
    __shared__ FStepSmem<T> sm;
    __shared__ double rv[2][NB];  // the right-hand side block of row p (or p - 1) and E times it
    const int *it = t.items + 4 * (size_t)(first + blockIdx.x);
    const int f = it[0];
    if (f < 0) return;
    const int ti = it[1], tj = it[2], kind = it[3];
    if (kind == 0) __builtin_amdgcn_s_setprio(3);
    const gf_front_dev F = t.fr[f];
    T *A = Fb + F.F;
    const size_t ld = (size_t)F.mp;
    T *Ef = Eb + F.E;
    const int tid = threadIdx.x;
    const int R0 = ti * NB, C0 = tj * NB;
    // with the right-hand side in the front vector (Vb): the first conversion tile of
    // row p - 1 converts its block, y = E_{p-1} b_{p-1}; the first update tile of row
    // i subtracts L(i, p) y_p, y_p = E_p b_p, from block i (each block written by one
    // workgroup; b_p itself is final and unconverted during launch p)
    const bool rhs_row = Vb && tj == (kind == 2 ? p : p + 1);
    double *vrow = Vb ? Vb + F.V + (size_t)R0 : nullptr;
    // product 1: W = E_q A'(q, j), q = p (update, diagonal) or p - 1 (conversion)
    const int q = kind == 2 ? ti : p;
    const T *E = Ef + (size_t)q * NB * NB;
    const T *Aq = A + (size_t)(q * NB) * ld + C0;  // A'(q, j)
    T e[16], a0[8], a1[8];
    tile_load(e, [&](int i, int k) { return E[i * NB + k]; });
    rowchunk_load(a0, Aq, ld, 0);
    rowchunk_load(a1, Aq, ld, KT);
    if (rhs_row && tid < NB) rv[0][tid] = kind == 2 ? vrow[tid] : Vb[F.V + (size_t)(p * NB) + tid];
    tile_put(sm.X, LDT, e);
    rowchunk_put_t(sm.C, a0);
    __syncthreads();
    T l0[8], l1[8], old[2][2][4];
    const T *Lp = A + (size_t)R0 * ld + p * NB;  // L(i, p)
    if (kind != 2) colchunk_load(l0, Lp, ld, 0);
    if (rhs_row && tid < NB) {
        double yv = 0.0;
        for (int c = 0; c < NB; c++) yv += (double)sm.X[tid * LDT + c] * rv[0][c];
        if (kind == 2) vrow[tid] = yv;
        else rv[1][tid] = yv;  // y_p
    }
    Acc<T> w;
    w.zero();
    mm_xc<T>(w, sm.X, sm.C, 0);
    __syncthreads();
    rowchunk_put_t(sm.C, a1);
    __syncthreads();
    if (kind != 2) {
        // behind the second half of product 1
        colchunk_load(l1, Lp, ld, KT);
        acc_each<T>([&](int x, int y, int v, int i, int j) { old[x][y][v] = A[(R0 + i) * ld + C0 + j]; });
    }
    mm_xc<T>(w, sm.X, sm.C, KT);
    if (kind == 2) {
        acc_each<T>([&](int x, int y, int v, int i, int j) { A[(R0 + i) * ld + C0 + j] = w.a[x][y][v]; });
        return;
    }
    __syncthreads();
    // product 2: old - L(i, p) W, W kept transposed in X
    acc_each<T>([&](int x, int y, int v, int i, int j) { sm.X[j * LDT + i] = w.a[x][y][v]; });
    colchunk_put(sm.C, l0);
    __syncthreads();
    double s = 0.0;
    if (rhs_row && tid < NB)
        for (int c = 0; c < KT; c++) s += (double)sm.C[tid * SA_LD + c] * rv[1][c];
    Acc<T> acc;
    acc.zero();
    mm_cx<T>(acc, sm.C, sm.X, 0);
    __syncthreads();
    colchunk_put(sm.C, l1);
    __syncthreads();
    if (rhs_row && tid < NB) {
        for (int c = 0; c < KT; c++) s += (double)sm.C[tid * SA_LD + c] * rv[1][KT + c];
        vrow[tid] -= s;  // b_i -= L(i, p) y_p
    }
    mm_cx<T>(acc, sm.C, sm.X, KT);
    if (kind == 1) {
        acc_each<T>([&](int x, int y, int v, int i, int j) { A[(R0 + i) * ld + C0 + j] = old[x][y][v] - acc.a[x][y][v]; });
        return;
    }
    __syncthreads();
    acc_each<T>([&](int x, int y, int v, int i, int j) { sm.X[i * LDT + j] = old[x][y][v] - acc.a[x][y][v]; });
    __syncthreads();
    inverse64<T>(sm.X, sm.tmp, sm.aux, flag);  // E_{p+1}
    T *En = Ef + (size_t)ti * NB * NB;
    for (int q2 = tid; q2 < NB * NB; q2 += 256) En[q2] = sm.X[(q2 / NB) * LDT + q2 % NB];
}

// ---- solve -----------------------------------------------------------------------
// Tile rows are summed by quads of lanes: thread t takes row t / 4 of a 64-row
// tile and its 16 columns 16 (t % 4) .. 16 (t % 4) + 15 (one 128-byte run of an
// FP64 row, read with 16-byte vector loads), so a row's partial sums sit in four
// neighbouring lanes and two DPP moves reduce them.  Front rows are 512-byte
// aligned (fronts and diagonal inverses start at multiples of 64 x 64 entries,
// rows are multiples of 64 entries long).

template <typename T> __device__ __forceinline__ void load16(const T *p, T (&v)[16]);
template <> __device__ __forceinline__ void load16<double>(const double *p, double (&v)[16])
{
    const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const double2 w = q[e];
        v[2 * e] = w.x;
        v[2 * e + 1] = w.y;
    }
}
template <> __device__ __forceinline__ void load16<float>(const float *p, float (&v)[16])
{
    const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const float4 w = q[e];
        v[4 * e] = w.x;
        v[4 * e + 1] = w.y;
        v[4 * e + 2] = w.z;
        v[4 * e + 3] = w.w;
    }
}
template <typename T> __device__ __forceinline__ double dot16(const T (&v)[16], const double *y)
{
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 16; e++) s += (double)v[e] * y[e];
    return s;
}
// Sum over the four lanes of a quad (DPP quad_perm [1,0,3,2] then [2,3,0,1]).
__device__ __forceinline__ double quad_sum(double s)
{
    s += dpp_mov<0xB1>(s);
    return s + dpp_mov<0x4E>(s);
}

// Own rows of a front: w = b_own + children's update vectors, then
// y_p = E_p (w_p - sum_{q<p} L(p, q) y_q) block by block (one workgroup, y in LDS).
template <typename T>
__global__ __launch_bounds__(256) void front_fwd_own_kernel(FrontTab t, int f0, const T *Fb, const T *Eb,
                                                            const double *rhs, double *Vb)
{
    __shared__ double y[KMAX];
    __shared__ double tmp[NB];
    const int f = f0 + blockIdx.x;
    const gf_front_dev F = t.fr[f];
    const int kp = NB * F.kb, tid = threadIdx.x, r = tid >> 2, sg = 16 * (tid & 3);
    const int *rows = t.rows + F.rows;
    const size_t ld = (size_t)F.mp;
    for (int i = tid; i < kp; i += 256) {
        const int o = rows[i];
        y[i] = (o >= 0 ? rhs[o] : 0.0) + child_sum(t, F, i, Vb);
    }
    __syncthreads();
    for (int p = 0; p < F.kb; p++) {
        T e[16];
        load16(Eb + F.E + (size_t)p * NB * NB + r * NB + sg, e);
        const T *L = Fb + F.F + (size_t)(NB * p + r) * ld + sg;
        double part = 0.0;
        for (int q = 0; q < p; q++) {
            T v[16];
            load16(L + NB * q, v);
            part += dot16(v, y + NB * q + sg);
        }
        part = quad_sum(part);
        if ((tid & 3) == 0) tmp[r] = y[NB * p + r] - part;
        __syncthreads();
        part = quad_sum(dot16(e, tmp + sg));
        if ((tid & 3) == 0) y[NB * p + r] = part;
        __syncthreads();
    }
    for (int i = tid; i < kp; i += 256) Vb[F.V + i] = y[i];
}

// Update row tile t of a front: c = sum_children c_child - L(t, own) y_own.
template <typename T>
__global__ __launch_bounds__(256) void front_fwd_upd_kernel(FrontTab t, int first, const T *Fb, double *Vb)
{
    __shared__ double y[KMAX];
    const int *it = t.items + 4 * (size_t)(first + blockIdx.x);
    const gf_front_dev F = t.fr[it[0]];
    const int tr = it[1], tid = threadIdx.x, r = tid >> 2, sg = 16 * (tid & 3);
    const int kp = NB * F.kb, i = NB * tr + r;
    for (int j = tid; j < kp; j += 256) y[j] = Vb[F.V + j];
    const double c = (tid & 3) == 0 ? child_sum(t, F, i, Vb) : 0.0;
    __syncthreads();
    const T *L = Fb + F.F + (size_t)i * F.mp + sg;
    double part = 0.0;
    for (int q = 0; q < F.kb; q++) {
        T v[16];
        load16(L + NB * q, v);
        part += dot16(v, y + NB * q + sg);
    }
    part = quad_sum(part);
    if ((tid & 3) == 0) Vb[F.V + i] = c - part;
}

// Own row tile p of a front with an update part: z_p = y_p - U(p, update) x_update
// (x_update gathered from the ancestors' rows of x into LDS, 1024 at a time).
template <typename T>
__global__ __launch_bounds__(256) void front_bwd_upd_kernel(FrontTab t, int first, const T *Fb, const T *Eb,
                                                            const double *x, double *Vb, int raw_last)
{
    __shared__ double xu[KMAX];
    const int *it = t.items + 4 * (size_t)(first + blockIdx.x);
    const gf_front_dev F = t.fr[it[0]];
    const int p = it[1], tid = threadIdx.x, r = tid >> 2, sg = 16 * (tid & 3);
    const int *rows = t.rows + F.rows;
    const T *U = Fb + F.F + (size_t)(NB * p + r) * F.mp + sg;
    double part = 0.0;
    for (int c0 = NB * F.kb; c0 < F.mp; c0 += KMAX) {
        const int cn = min(KMAX, F.mp - c0);
        __syncthreads();
        for (int j = tid; j < cn; j += 256) {
            const int o = rows[c0 + j];
            xu[j] = o >= 0 ? x[o] : 0.0;
        }
        __syncthreads();
        for (int q = 0; q < cn / NB; q++) {
            T v[16];
            load16(U + c0 + NB * q, v);
            part += dot16(v, xu + NB * q + sg);
        }
    }
    part = quad_sum(part);
    if (p < F.kb - 1) {
        if ((tid & 3) == 0) Vb[F.V + NB * p + r] -= part;
        return;
    }
    // the last own block row holds A'(p, .) = D_p U(p, .): subtract E_p (A' x); when
    // the factorisation carried the right-hand side, its block is unconverted too:
    // z_p = E_p (b_p - A' x)
    T e[16];
    load16(Eb + F.E + (size_t)p * NB * NB + r * NB + sg, e);
    __syncthreads();
    if ((tid & 3) == 0) xu[r] = raw_last ? Vb[F.V + NB * p + r] - part : part;
    __syncthreads();
    part = quad_sum(dot16(e, xu + sg));
    if ((tid & 3) == 0) {
        if (raw_last) Vb[F.V + NB * p + r] = part;
        else Vb[F.V + NB * p + r] -= part;
    }
}

// Own-block back substitution x_p = z_p - sum_{p < j < kb} U(p, j) x_j, then the
// scatter of the own rows of x.
template <typename T>
__global__ __launch_bounds__(256) void front_bwd_own_kernel(FrontTab t, int f0, const T *Fb, const T *Eb,
                                                            const double *Vb, double *x, int raw_last)
{
    __shared__ double y[KMAX];
    const int f = f0 + blockIdx.x;
    const gf_front_dev F = t.fr[f];
    const int kp = NB * F.kb, tid = threadIdx.x, r = tid >> 2, sg = 16 * (tid & 3);
    const int *rows = t.rows + F.rows;
    const size_t ld = (size_t)F.mp;
    for (int i = tid; i < kp; i += 256) y[i] = Vb[F.V + i];
    __syncthreads();
    if (raw_last && F.mb == F.kb) {
        // no update part, so no front_bwd_upd_kernel converted the last block: y = E b
        const int p = F.kb - 1;
        T e[16];
        load16(Eb + F.E + (size_t)p * NB * NB + r * NB + sg, e);
        const double v = quad_sum(dot16(e, y + NB * p + sg));
        __syncthreads();
        if ((tid & 3) == 0) y[NB * p + r] = v;
        __syncthreads();
    }
    for (int p = F.kb - 2; p >= 0; p--) {
        const T *U = Fb + F.F + (size_t)(NB * p + r) * ld + sg;
        double part = 0.0;
        for (int j = p + 1; j < F.kb; j++) {
            T v[16];
            load16(U + NB * j, v);
            part += dot16(v, y + NB * j + sg);
        }
        part = quad_sum(part);
        if ((tid & 3) == 0) y[NB * p + r] -= part;
        __syncthreads();
    }
    for (int i = tid; i < F.k; i += 256) x[rows[i]] = y[i];
}

FrontTab tab_of(const gf_fronts &fp)
{
    FrontTab t;
    t.fr = fp.d_fr;
    t.rows = fp.d_rows;
    t.child = fp.d_child;
    t.inv = fp.d_inv;
    t.items = fp.d_items;
    t.child_inv = fp.d_child_inv;
    t.orig = fp.d_orig;
    t.asm_orig = fp.d_asm_orig;
    return t;
}

template <typename T>
int front_factor_t(const gf_fronts &fp, const double *ab, int ld, int hb, const double *D, T *F, T *E, hipStream_t s,
                   int *flag, const double *rhs, double *V)
{
    const FrontTab t = tab_of(fp);
    int launches = 0;
    for (int l = 0; l < fp.n_level; l++) {
        int pr = gf_probe_begin(s, GF_KC_FRONT_ASSEMBLE);
        hipLaunchKernelGGL(front_assemble_kernel<T>, dim3(fp.asm_count[l]), dim3(ASM_T), 0, s, t, fp.asm_first[l], ab, ld,
                           hb, D, F, E, flag, rhs, V);
        gf_probe_end(s, pr, (double)fp.asm_count[l] * NB * NB);
        launches++;
        for (int k = fp.lev_step[l]; k < fp.lev_step[l + 1]; k++) {
            if (fp.step_count[k] == 0) continue;
            const int p = k - fp.lev_step[l];
            pr = gf_probe_begin(s, GF_KC_FRONT_STEP);
            hipLaunchKernelGGL(front_step_kernel<T>, dim3(fp.step_count[k]), dim3(256), 0, s, t, fp.step_first[k], p,
                               F, E, flag, rhs ? V : nullptr);
            gf_probe_end(s, pr, fp.step_flops[k]);
            launches++;
        }
    }
    GF_HIP(hipGetLastError());
    return launches;
}

template <typename T>
int front_solve_t(const gf_fronts &fp, const T *F, const T *E, const double *rhs, double *x, double *V, hipStream_t s,
                  bool carried)
{
    const FrontTab t = tab_of(fp);
    int launches = 0;
    // probe units: factor bytes each launch reads (the own triangle with the
    // diagonal inverses, or the update panel)
    auto own_bytes = [&](int l, bool fwd) {
        double b = 0.0;
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++) {
            const double kp = NB * fp.fr[f].kb;
            b += fwd ? kp * (kp + NB) / 2.0 : kp * (kp - NB) / 2.0;
        }
        return b * sizeof(T);
    };
    auto panel_bytes = [&](int l) {
        double b = 0.0;
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++)
            b += (double)NB * fp.fr[f].kb * (fp.fr[f].mp - NB * fp.fr[f].kb);
        return b * sizeof(T);
    };
    for (int l = 0; l < fp.n_level && !carried; l++) {
        const int nf = fp.lev_off[l + 1] - fp.lev_off[l];
        int pr = gf_probe_begin(s, GF_KC_FRONT_SOLVE);
        hipLaunchKernelGGL(front_fwd_own_kernel<T>, dim3(nf), dim3(256), 0, s, t, fp.lev_off[l], F, E, rhs, V);
        gf_probe_end(s, pr, pr >= 0 ? own_bytes(l, true) : 0.0);
        launches++;
        if (fp.fupd_count[l] > 0) {
            pr = gf_probe_begin(s, GF_KC_FRONT_SOLVE);
            hipLaunchKernelGGL(front_fwd_upd_kernel<T>, dim3(fp.fupd_count[l]), dim3(256), 0, s, t, fp.fupd_first[l], F, V);
            gf_probe_end(s, pr, pr >= 0 ? panel_bytes(l) : 0.0);
            launches++;
        }
    }
    for (int l = fp.n_level - 1; l >= 0; l--) {
        const int nf = fp.lev_off[l + 1] - fp.lev_off[l];
        int pr;
        if (fp.bupd_count[l] > 0) {
            pr = gf_probe_begin(s, GF_KC_FRONT_SOLVE);
            hipLaunchKernelGGL(front_bwd_upd_kernel<T>, dim3(fp.bupd_count[l]), dim3(256), 0, s, t, fp.bupd_first[l], F,
                               E, (const double *)x, V, carried ? 1 : 0);
            gf_probe_end(s, pr, pr >= 0 ? panel_bytes(l) : 0.0);
            launches++;
        }
        pr = gf_probe_begin(s, GF_KC_FRONT_SOLVE);
        hipLaunchKernelGGL(front_bwd_own_kernel<T>, dim3(nf), dim3(256), 0, s, t, fp.lev_off[l], F, E, (const double *)V,
                           x, carried ? 1 : 0);
        gf_probe_end(s, pr, pr >= 0 ? own_bytes(l, false) : 0.0);
        launches++;
    }
    GF_HIP(hipGetLastError());
    return launches;
}

template <typename T> int put(T **dst, const std::vector<T> &src, hipStream_t s)
{
    const size_t n = std::max<size_t>(src.size(), 1);
    GF_TRY(gf_pool_alloc((void **)dst, n * sizeof(T)));
    if (!src.empty()) GF_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return GF_OK;
}

}  // namespace

int gf_front_upload(gf_fronts &fp, hipStream_t s)
{
    if (fp.max_k > KMAX) {
        gf_set_error("multifrontal plan: supernode larger than the solve supports");
        return GF_E_STATE;
    }
    GF_TRY(put(&fp.d_fr, fp.fr, s));
    GF_TRY(put(&fp.d_rows, fp.rows, s));
    GF_TRY(put(&fp.d_child, fp.child, s));
    GF_TRY(put(&fp.d_child_inv, fp.child_inv, s));
    GF_TRY(put(&fp.d_inv, fp.inv, s));
    GF_TRY(put(&fp.d_items, fp.items, s));
    GF_TRY(put(&fp.d_orig, fp.orig, s));
    GF_TRY(put(&fp.d_asm_orig, fp.asm_orig, s));
    return GF_OK;
}

void gf_front_free(gf_fronts &fp)
{
    void *p[] = {fp.d_fr, fp.d_rows, fp.d_child, fp.d_child_inv, fp.d_inv, fp.d_items, fp.d_orig, fp.d_asm_orig};
    for (void *q : p) gf_pool_free(q);
    fp.d_fr = nullptr;
    fp.d_rows = fp.d_child = fp.d_inv = fp.d_items = fp.d_orig = fp.d_asm_orig = nullptr;
    fp.d_child_inv = nullptr;
}

int gf_front_factor(const gf_fronts &fp, int elem_size, const double *ab, int ld, int hb, const double *D, void *F,
                    void *E, hipStream_t s, int *flag, const double *rhs, double *V)
{
    if (elem_size == 8) return front_factor_t<double>(fp, ab, ld, hb, D, (double *)F, (double *)E, s, flag, rhs, V);
    return front_factor_t<float>(fp, ab, ld, hb, D, (float *)F, (float *)E, s, flag, rhs, V);
}

int gf_front_solve(const gf_fronts &fp, int elem_size, const void *F, const void *E, const double *rhs, double *x,
                   double *V, hipStream_t s, bool carried)
{
    if (elem_size == 8) return front_solve_t<double>(fp, (const double *)F, (const double *)E, rhs, x, V, s, carried);
    return front_solve_t<float>(fp, (const float *)F, (const float *)E, rhs, x, V, s, carried);
}
