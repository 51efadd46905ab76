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
#include "gf_tile.h"

#include <algorithm>

namespace {

constexpr int SB = 8;    // blocks per solve batch
constexpr int BW = NB * SB;

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
