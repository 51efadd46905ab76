// Internal definitions of libgiraffe_amd (host side + kernel launch prototypes).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <string>
#include <thread>
#include <vector>
#include "../../include/giraffe_amd.h"

// Mixed precision: FP64 factorisations in a row after two FP32 fallbacks in a
// row (the first back-off; later ones double while the retries keep failing).
constexpr int GF_MIXED_BACKOFF = 16;
// Multifrontal path, mixed precision: the cost of an FP64 factorisation (with
// the right-hand side carried) plus the backward solve, and of an FP32
// factorisation, in FP32 refinement sweeps (tree solve, residual, correction),
// as measured on MI355X for the 84,867-DOF Shell_1 plate: 2.0 ms, 1.1 ms and
// 0.39 ms.  An FP32-refined solve that costs more than the FP64 one starts a
// back-off as a fallback does.
constexpr double GF_MIXED_F64_SWEEPS = 5.0, GF_MIXED_F32_SWEEPS = 3.0;

// Bits of a factorisation's flag word: a zero or non-finite pivot (the solve
// fails, sparsesystem's info_fail), and a tiny pivot replaced by a perturbation
// of the block's scale (PARDISO's static-pivoting default; the solve goes on).
constexpr int GF_FLAG_FAIL = 1, GF_FLAG_PERTURBED = 2;

// ---- error handling --------------------------------------------------------
void gf_set_error(const std::string &msg);
#define GF_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            gf_set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            return GF_E_DEVICE;                                                       \
        }                                                                             \
    } while (0)

#define GF_TRY(x)                   \
    do {                            \
        int s_ = (x);               \
        if (s_ != GF_OK) return s_; \
    } while (0)

// ---- host planning threads -------------------------------------------------
// fn(worker, item) for item = 0 .. n - 1, items handed out in order to at most
// 16 threads (worker = 0 .. workers - 1), or run inline when `work` (the
// caller's cost estimate) is too small to pay for threads. The caller keeps the
// result independent of which worker takes which item.
int gf_plan_workers(long long work);
template <class Fn> void gf_parallel_for(int n, int workers, Fn fn)
{
    if (workers <= 1 || n <= 1) {
        for (int i = 0; i < n; i++) fn(0, i);
        return;
    }
    std::atomic<int> next{0};
    auto run = [&](int w) {
        for (int i; (i = next.fetch_add(1)) < n;) fn(w, i);
    };
    std::vector<std::thread> th;
    for (int w = 1; w < workers; w++) th.emplace_back(run, w);
    run(0);
    for (auto &t : th) t.join();
}

// ---- device memory cache (gf_pool.cpp) ---------------------------------------
int gf_pool_alloc(void **out, size_t bytes);
void gf_pool_free(void *p);      // only after the last device use has completed
int gf_pinned_alloc(void **out, size_t bytes);
void gf_pinned_free(void *p);

// ---- per-element constant record of Beam_1 (structure of arrays, [BC_N][nb]) ----
enum {
    BC_T3 = 0,        // 9: transform3 rows E1,E2,E3
    BC_D = 9,         // 7: GA1 GA2 EA EI1 EI2 EI12 GJ
    BC_E3R = 16,      // 3: reference axis in the local frame
    BC_JAC = 19,      // 1: L/2
    BC_RHOA = 20,     // 1: rho*A
    BC_N = 21,        // 6: N[gp][3]
    BC_DN = 27,       // 6: dN[gp][3]
    BC_COUNT = 33
};
// Gauss-point state of Beam_1, [BG_COUNT][nb]: per gp Q_i(9) kappa_i(3) z'_i(3)
enum { BG_PER_GP = 15, BG_COUNT = 30 };

// ---- global tangent storage -------------------------------------------------
// GM_BTD : block-tridiagonal blocks A (sub), B (diag), C (super), [nblk][m][m]
//          row-major, consecutive in one allocation (B = A + nblk m m,
//          C = B + nblk m m); used when the half-bandwidth fits a block of m <= 32.
// GM_BAND: skewed dense band, entry (i, c) with -hb <= c - i <= ku at
//          ab[i * ld + c + hb], ld = hb + ku (each row a contiguous window of
//          ld + 1); ku = gf_band_ku(hb) leaves room for the factorisation's
//          fill.  Sub-matrices inside the band are plain strided matrices with
//          leading dimension ld, which is what the band LU tiles use.
enum { GM_BTD = 0, GM_BAND = 1 };
struct gf_gmat {
    int layout;
    int m;
    double *A, *B, *C;
    double *ab;
    int hb;
    int ld;
};
int gf_band_ku(int hb);

// Reciprocal of a pivot that is finite, non-zero and not subnormal (the solvers
// check this first): the hardware estimate refined by Newton-Raphson steps
// (each doubles the correct bits), within an ulp of 1/p at a fraction of the
// latency of the IEEE division sequence, which sits on the pivot chain.
__device__ __forceinline__ double pivot_rcp(double p)
{
    double r = __builtin_amdgcn_rcp(p);
    double e = fma(-p, r, 1.0);
    r = fma(r, e, r);
    e = fma(-p, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float pivot_rcp(float p)
{
    float r = __builtin_amdgcn_rcpf(p);
    const float e = fmaf(-p, r, 1.0f);
    return fmaf(r, e, r);
}

// Block index q = floor(x / m) for 0 <= x < 2^24 and m <= 32, through a float
// product (one rounding, then one correction step each way): no integer divide.
__device__ __forceinline__ int gm_block_of(int x, int m, float inv_m)
{
    int q = (int)(((float)x + 0.5f) * inv_m);
    q -= q * m > x ? 1 : 0;
    q += (q + 1) * m <= x ? 1 : 0;
    return q;
}

// Addresses of N entries (R, Cc); a constrained entry (R or Cc negative) gets a
// valid stand-in address whose value is never stored back.  The layout test is
// one uniform branch around the whole batch, so loads through the addresses are
// straight-line code and all in flight together.
template <int N>
__device__ __forceinline__ void gm_addr_n(const gf_gmat &g, const int (&R)[N], const int (&Cc)[N], double *(&p)[N])
{
    if (g.layout == GM_BTD) {
        const int m = g.m;
        const float inv_m = 1.0f / (float)m;
        const ptrdiff_t stride = g.B - g.A;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const bool ok = R[k] >= 0 && Cc[k] >= 0;
            const int r = ok ? R[k] : 0, c = ok ? Cc[k] : 0;
            const int bi = gm_block_of(r, m, inv_m), bj = gm_block_of(c, m, inv_m);
            const size_t off = ((size_t)bi * m + (r - bi * m)) * m + (c - bj * m);
            // A, B and C are one allocation in that order (gf_btd_alloc): the block
            // kind selects an offset; a selected pointer would be re-read from the
            // kernel arguments, one dependent load per entry
            const ptrdiff_t kind = bj < bi ? 0 : (bj == bi ? 1 : 2);
            p[k] = g.A + kind * stride + off;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const bool ok = R[k] >= 0 && Cc[k] >= 0;
            const int r = ok ? R[k] : 0, c = ok ? Cc[k] : 0;
            p[k] = g.ab + (size_t)r * (size_t)g.ld + (size_t)(c + g.hb);
        }
    }
}

// N entries added at once: every address, all loads, then the stores of the
// unconstrained ones.
template <int N>
__device__ __forceinline__ void gm_add_n(const gf_gmat &g, const int (&R)[N], const int (&Cc)[N], const double (&v)[N])
{
    double *p[N];
    double old[N];
    gm_addr_n<N>(g, R, Cc, p);
#pragma unroll
    for (int k = 0; k < N; k++) old[k] = *p[k];
#pragma unroll
    for (int k = 0; k < N; k++)
        if (R[k] >= 0 && Cc[k] >= 0) *p[k] = old[k] + v[k];
}

// r[R[k]] += v[k] for the unconstrained rows, loads first.
template <int N>
__device__ __forceinline__ void vec_add_n(double *r, const int (&R)[N], const double (&v)[N])
{
    double old[N];
#pragma unroll
    for (int k = 0; k < N; k++) old[k] = r[R[k] >= 0 ? R[k] : 0];
#pragma unroll
    for (int k = 0; k < N; k++)
        if (R[k] >= 0) r[R[k]] = old[k] + v[k];
}

// Table::GetValueAt (src/Table.cpp:73-101): piecewise-linear, clamped at both ends.
__device__ __forceinline__ double gf_table_value(const double *tt, const double *tv, int r0, int nt, double t, int col)
{
    if (nt == 0) return 0.0;
    const double *T = tt + r0;
    const double *Vv = tv + 6 * (size_t)r0;
    if (t >= T[nt - 1]) return Vv[6 * (nt - 1) + col];
    if (t <= T[0]) return Vv[col];
    int idx = 0;
    while (t >= T[idx]) idx++;
    double fac = (t - T[idx - 1]) / (T[idx] - T[idx - 1]);
    return Vv[6 * (idx - 1) + col] + fac * (Vv[6 * idx + col] - Vv[6 * (idx - 1) + col]);
}

// All six columns of the same table at time t: one segment search, the same
// per-column arithmetic as gf_table_value.
__device__ __forceinline__ void gf_table_values6(const double *tt, const double *tv, int r0, int nt, double t,
                                                 double (&out)[6])
{
    if (nt == 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) out[c] = 0.0;
        return;
    }
    const double *T = tt + r0;
    const double *Vv = tv + 6 * (size_t)r0;
    int lo, hi;
    double fac;
    if (t >= T[nt - 1]) { lo = hi = nt - 1; fac = -1.0; }
    else if (t <= T[0]) { lo = hi = 0; fac = -1.0; }
    else {
        int idx = 0;
        while (t >= T[idx]) idx++;
        lo = idx - 1;
        hi = idx;
        fac = (t - T[idx - 1]) / (T[idx] - T[idx - 1]);
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const double a = Vv[6 * lo + c], b = Vv[6 * hi + c];
        out[c] = fac < 0.0 ? a : a + fac * (b - a);
    }
}

// ---- Shell_1 constant record ([SC_COUNT][ns]) and Gauss-point state ([SG_COUNT][ns]) ----
enum {
    SC_T3 = 0,        // 9: rows e1r, e2r, e3r
    SC_ALPHA1 = 9,    // area / 3
    SC_MU = 10, SC_LAMBDA = 11, SC_THICK = 12, SC_DRILL = 13, SC_RHO = 14,
    SC_NU = 15,       // [3][6] quadratic u shapes at the mid-side points
    SC_NU1 = 33, SC_NU2 = 51,
    SC_NA = 69,       // [3][3] linear alpha shapes
    SC_NA1 = 78, SC_NA2 = 87,
    SC_GW = 96,       // [6] sum_g w4_g N4u_g[a] (self-weight)
    SC_COUNT = 102
};
enum { SP_N4U = 0, SP_N4U1 = 36, SP_N4U2 = 72, SP_W4 = 108, SP_COUNT = 114 };  // Cowper points [6][6]
enum { SG_PER_GP = 21, SG_COUNT = 63 };  // per gp: Q_i 9, kr1 3, kr2 3, zx1 3, zx2 3

// Nodal loads grouped by node (one device thread per loaded node).
enum { NL_KIND = 0, NL_CS = 1, NL_MULT = 10, NL_TAB0 = 16, NL_TABN = 17, NL_REC = 18 };

// Block-tridiagonal factor workspace (cyclic reduction), precision T.
struct gf_btd {
    int m = 0, mt = 16, nblk = 0, levels = 0, tail_level = 0;
    int back_level = 0;             // lowest level whose back-substitution the tail kernel does
    size_t bsz = 0;                 // m*m
    void *A = nullptr, *B = nullptr, *C = nullptr;     // [nblk][m][m] assembled blocks (kept rows reduced in place)
    void *Bi = nullptr, *X = nullptr, *Y = nullptr;    // eliminated rows: B^-1, B^-1 A, B^-1 C
    void *Lsv = nullptr, *Rsv = nullptr;               // pre-update A_k, C_k per (level, kept row)
    long long *d_save = nullptr;    // lvl_save on the device
    std::vector<int> lvl_n;         // rows active at each level
    std::vector<long long> lvl_save; // offsets (in blocks) into Lsv/Rsv per level
    long long n_save = 0;           // saved coupling blocks over all levels
    int elem_size = 8;
};
int gf_btd_plan(gf_btd &b, int m, int nblk, int elem_size);  // level structure only, no device call

// Nested-dissection multifrontal factorisation (gf_nd.cpp plans it on the host,
// gf_front.hip runs it).  Supernodes are the parts and separators of a recursive
// coordinate bisection of the node graph; the front of a supernode is a dense
// matrix over its own DOFs (k, padded to kp = 64 kb) followed by its update DOFs
// (u: the DOFs of ancestors it couples to, padded to up), mp = kp + up.  Fronts
// are numbered level by level (leaves first, a parent's level above all of its
// children's); fronts of one level are independent.
struct gf_front_dev {
    long long F, E, V;  // offsets (elements) of the front matrix [mp][mp], its kb diagonal inverses [64][64], its vector [mp]
    int mp, kb, mb, k;  // padded order, own blocks, blocks, own DOFs
    int rows;           // offset of the row map: mp entries, band row of each front row or -1 (padding)
    int ch0, ch1;       // children: entries [ch0, ch1) of the child lists
    int pad;
};
struct gf_fronts {
    int n_front = 0, n_level = 0, max_k = 0;
    std::vector<gf_front_dev> fr;
    std::vector<int> lev_off;          // fronts of level l: [lev_off[l], lev_off[l+1])
    std::vector<int> lev_kb;           // largest kb of level l
    std::vector<int> rows;             // row maps
    std::vector<int> child;            // child front of each child entry
    std::vector<long long> child_inv;  // offset of each child entry's map in inv (parent position -> child update row or -1)
    std::vector<int> inv;
    std::vector<int> items;            // work items, 4 ints each: front, tile row, tile column, kind (front -1: none)
    std::vector<int> step_first, step_count;  // factor launch p of level l: index lev_step[l] + p, p = 0 .. lev_kb[l] - 1
    std::vector<int> lev_step;
    std::vector<int> asm_first, asm_count;    // per level: assembly tiles
    std::vector<int> fupd_first, fupd_count;  // per level: forward update row tiles (front, tile row)
    std::vector<int> bupd_first, bupd_count;  // per level: backward own row tiles with an update part
    std::vector<double> step_flops;    // algorithmic flops of each factor launch
    std::vector<int> orig;             // tangent entries of the assembly tiles: position in the tile, band row, band column
    std::vector<int> asm_orig;         // per work item (2 ints): first entry of its tile in orig, count
    long long f_elems = 0, e_elems = 0, v_elems = 0;
    double flops = 0.0;                // algorithmic flops of one factorisation
    long long n_sep_dofs = 0;          // DOFs owned by separators (non-leaf fronts)
    // device copies
    gf_front_dev *d_fr = nullptr;
    int *d_rows = nullptr, *d_child = nullptr, *d_inv = nullptr, *d_items = nullptr;
    int *d_orig = nullptr, *d_asm_orig = nullptr;
    long long *d_child_inv = nullptr;
};
// Plans the fronts of the free DOFs dof[6 node + comp] (band rows, -1 when not
// free) over the node graph of the description's elements; host only.
int gf_nd_plan(gf_fronts &fp, const gf_model_desc *d, const std::vector<int> &dof, int nn, int npad);
// The assembly tiles' lists of tangent entries from the tangent's structure (CSR
// over band rows); after gf_nd_plan.
int gf_nd_orig_lists(gf_fronts &fp, const std::vector<int> &ptr, const std::vector<int> &col, int npad);

// gf_model_create in two halves, for callers that plan on one thread and touch
// the device on another. gf_model_preplan: the host plan without the
// multifrontal path's per-entry lists, and the device bytes the model will take
// (what gf_model_footprint reports); no device call. gf_model_plan_lists: those
// lists (host only). gf_model_create_planned: the device half (allocations,
// uploads; the lists first if still missing); on failure the caller destroys
// the model with gf_model_destroy, as it does a planned model it never creates.
int gf_model_preplan(const gf_model_desc *d, const gf_options *opt, gf_model **out, unsigned long long *bytes);
int gf_model_plan_lists(const gf_model_desc *d, gf_model *m);
int gf_model_create_planned(const gf_model_desc *d, const gf_options *opt, gf_model *m);
// Launch count of one factorisation and of one solve of a plan.
int gf_front_factor_launches(const gf_fronts &fp);
int gf_front_solve_launches(const gf_fronts &fp);
int gf_front_upload(gf_fronts &fp, hipStream_t s);  // device copies of the plan's tables
void gf_front_free(gf_fronts &fp);
// Factorisation of the band-stored tangent ab (row stride ld, half-bandwidth hb,
// FP64) into fronts of precision elem_size (8, or 4 with the equilibration
// scaling D applied on the way, D = null for none); returns launches.  rhs
// non-null: the right-hand side rides along (forward substitution done in the
// factorisation's launches, its state in the vector workspace V).
int gf_front_factor(const gf_fronts &fp, int elem_size, const double *ab, int ld, int hb, const double *D, void *F,
                    void *E, hipStream_t s, int *flag, const double *rhs, double *V);
// Solve with a factor: rhs (band-row vector) in, x out (may alias rhs); V: the
// plan's vector workspace (v_elems doubles); carried: the factorisation carried
// this right-hand side (only the backward substitution is left).  Returns launches.
int gf_front_solve(const gf_fronts &fp, int elem_size, const void *F, const void *E, const double *rhs, double *x,
                   double *V, hipStream_t s, bool carried);

// ---- per-kernel timing --------------------------------------------------------
// While a model's timing is on and its probe mask selects a kernel class, every
// launch of that class is bracketed by a pair of HIP events on its stream; the
// pairs are read back when the phase that issued them is accumulated.  `units`
// counts the work of each launch in the class's unit (block rows eliminated,
// band columns eliminated, elements assembled, rows swept).
struct gf_kprobe {
    unsigned mask = 0;
    std::vector<hipEvent_t> ev;    // 2 per slot
    std::vector<int> cls;
    std::vector<double> units;
    int used = 0;
    double ms[GF_NKC] = {}, units_done[GF_NKC] = {};
    long long launches[GF_NKC] = {};
    int open_slot = -1;

    bool wants(int c) const { return (mask >> c) & 1u; }
    int begin(hipStream_t s, int c);           // before the launch; -1 when not probed
    void end(hipStream_t s, int slot, double units);
    int flush();                               // waits for the recorded pairs, accumulates
    ~gf_kprobe();
};
// Probe of the model whose device work the calling thread is enqueuing (null when off).
extern thread_local gf_kprobe *gf_probe;
struct gf_probe_scope {
    gf_kprobe *prev;
    explicit gf_probe_scope(gf_kprobe *p) : prev(gf_probe) { gf_probe = p; }
    ~gf_probe_scope() { gf_probe = prev; }
};
inline int gf_probe_begin(hipStream_t s, int c) { return gf_probe && gf_probe->wants(c) ? gf_probe->begin(s, c) : -1; }
inline void gf_probe_end(hipStream_t s, int slot, double units)
{
    if (slot >= 0) gf_probe->end(s, slot, units);
}

// scalar block of a model (doubles): increment norms, failure flag, residual
// norms, refinement norms
// refinement norms, and the load time and gravity ramp factor of the current
// step (read by the assembly kernels, so a captured Newton step replays with
// the step's values)
enum { GF_SCAL_INC = 0, GF_SCAL_FLAG = 8, GF_SCAL_RES = 16, GF_SCAL_REF = 24, GF_SCAL_TIME = 28, GF_SCAL_LF = 29,
       GF_SCAL_N = 32 };

struct gf_model {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int precision = GF_PREC_FP64;
    int refine_max = 12;
    double refine_tol = 1e-11;
    bool refine_reuse = true;  // keep the FP32 factor across tangents while it converges
    int solver = GF_SOLVER_AUTO;

    int nn = 0, nb = 0, ns = 0, nl = 0;
    int n_free = 0;

    // host mirrors
    std::vector<int> dof;           // [nn*6] band row or -1
    std::vector<int> row_nodecomp;  // [n_free] node*6+comp
    std::vector<unsigned char> active, fix;
    std::vector<int> beam_order;    // colour-sorted position -> element id
    std::vector<int> beam_color_off;
    std::vector<int> beam_number, shell_number;  // deck element numbers (results writer)
    std::vector<double> h_nodes_ref;
    std::vector<int> h_beam_conn, h_shell_conn;  // description order
    double sol_start = 0, sol_end = 1;
    double t_last = 0, dt = 0;
    double gravity[3] = {0, 0, 0};
    int gravity_on = 0;
    int gravity_mode = 0;  // GF_GRAVITY_*: the BoolTable factor across solution steps
    int half_band = 0;
    int band_ld = 0;                // GM_BAND row stride: half_band + gf_band_ku(half_band)
    bool node_order_rcm = false;    // DOFs numbered in reverse Cuthill-McKee node order

    // device: node state
    double *d_disp = nullptr;       // [nn][6] incremental displacements
    double *d_copy = nullptr;       // [nn][6] converged X Y Z, alpha
    double *d_euler = nullptr;      // [nn][3]
    int *d_dof = nullptr;           // [nn][6]
    int *d_row_nc = nullptr;        // [n_free]

    // device: beams (colour-sorted)
    int *d_bconn = nullptr;         // [3][nb]
    double *d_bconst = nullptr;     // [BC_COUNT][nb]
    double *d_bgp = nullptr;        // [BG_COUNT][nb]
    double *d_bbuoy = nullptr;      // [nb] Pipe_1 buoyant mass per length (null: no buoyancy)
    std::vector<int> beam_kind;     // description order, GF_BEAM_KIND_*
    double ocean_surface[3] = {0, 0, 0};

    // device: shells (colour-sorted)
    int *d_sconn = nullptr;         // [6][ns]
    double *d_sconst = nullptr;     // [SC_COUNT][ns]
    double *d_spconst = nullptr;    // [SP_COUNT][ns]
    double *d_sgp = nullptr;        // [SG_COUNT][ns]
    double *d_sgp0 = nullptr;
    std::vector<int> shell_order, shell_color_off;
    // shell pressure loads: per shell (colour-sorted) the load record index or -1
    int *d_spload = nullptr;        // [ns]
    int *d_spflag = nullptr;        // [ns] AreaUpdate
    bool any_pressure = false;

    // device: nodal loads
    int n_load_nodes = 0;
    int *d_ln_node = nullptr, *d_ln_off = nullptr, *d_ln_ids = nullptr;
    double *d_lrec = nullptr;       // [nl][NL_REC]
    double *d_tab_t = nullptr, *d_tab_v = nullptr;

    // device: global system
    double *d_r = nullptr;          // residual r = f_ext - f_int [n_free]
    double *d_x = nullptr;          // increment [n_free]
    double *d_swork = nullptr;      // band triangular-solve work vector [npad]
    int layout = GM_BTD;
    gf_btd btd;
    double *d_band = nullptr;       // GM_BAND storage, FP64
    size_t band_elems = 0;
    double *d_band_inv = nullptr;   // per-panel triangular inverses (FP64 factor)
    float *d_band32_inv = nullptr;  // per-panel triangular inverses (FP32 factor)
    // mixed precision: equilibrated FP32 factor + FP64 refinement buffers
    gf_btd btd32;
    float *d_band32 = nullptr;
    // GM_BAND assembly with the multifrontal solve: the band holds the tangent
    // only (no fill), the factor lives in the fronts
    bool fronts_on = false;
    gf_fronts fronts;
    // host plan made ahead of the device part (gf_model_preplan): the band
    // layout's block size and count, and whether the per-entry lists exist
    bool planned = false, lists_done = false;
    int plan_bm = -1, plan_nblk = 0;
    double *d_front = nullptr, *d_front_e = nullptr, *d_front_v = nullptr;  // FP64 factor, vector workspace
    float *d_front32 = nullptr, *d_front32_e = nullptr;                      // FP32 factor (mixed)
    double *d_shell_rec = nullptr;   // per (Shell_1, Gauss point) phase-1 record [3 ns][276]
    double *d_shell_prec = nullptr;  // per (Shell_1, u-node) follower-pressure terms [6 ns][57]
    double *d_shell_K = nullptr, *d_shell_f = nullptr;  // per Shell_1 K [729] and f [27] of the current iteration
    // structure of the assembled tangent (the band entries of element node pairs,
    // row by row): what the assembly clears and the mixed residual multiplies
    std::vector<int> csr_ptr, csr_col;
    int *d_csr_ptr = nullptr, *d_csr_col = nullptr;
    double *d_scale = nullptr, *d_res = nullptr, *d_w = nullptr;
    long long n_solves = 0, n_sweeps = 0, n_fallbacks = 0, n_direct = 0, n_f32_factor = 0;
    int fallback_streak = 0, fp64_direct_left = 0;  // mixed-precision back-off
    int solve_sweeps = 0, solve_f32 = 0;            // sweeps and FP32 factorisations of the current solve
    bool f32_have = false;     // the FP32 factor holds an earlier tangent's factorisation
    bool f32_reused = false;   // the current refinement runs on such a factor
    int backoff_len = GF_MIXED_BACKOFF;
    // enqueue/result pairs (asynchronous drivers)
    bool ev_pending[GF_NPHASE] = {};
    long long pending_launches[GF_NPHASE] = {};
    int pending_info_fail = 0;
    // asynchronous refinement of the current tangent solve (mixed precision)
    bool refining = false, refine_then_assemble = false;
    int refine_it = 0;
    double refine_prev = 0.0;
    int npad = 0;                   // padded system size
    double *d_scal = nullptr;       // small reduction results
    double *h_scal = nullptr;       // pinned mirror
    int *d_flag = nullptr;          // solver failure flag
    bool factored = false;

    // reset support: initial node copy and Gauss-point state, kept on the device
    double *d_copy0 = nullptr, *d_bgp0 = nullptr;

    // phase timing (HIP events on `stream`)
    bool timing = false;
    hipEvent_t ev[2 * GF_NPHASE] = {};
    double t_ms[GF_NPHASE] = {0, 0, 0};
    long long launches[GF_NPHASE] = {0, 0, 0};
    gf_kprobe kprobe;
    // FP64 solve sequence replayed from a graph (GF_NO_GRAPH in the environment
    // at model creation switches this off)
    bool use_graph = true;
    hipGraphExec_t solve_graph = nullptr;
    // whole FP64 Newton iteration (solve, update, next assembly, scalar copy)
    hipGraphExec_t iter_graph = nullptr;
    hipGraphExec_t f32_factor_graph = nullptr, sweep_graph = nullptr;  // mixed precision
    int graph_launches = 0;
};

// ---- kernel launchers (defined in the .hip files) ----------------------------
int gf_launch_fill(hipStream_t s, double *p, size_t n, double v);
// zeroes three double ranges in one launch; entries dfirst + k dstride (k < dcount)
// of the second range are set to 1
int gf_launch_clear3(hipStream_t s, double *a, size_t na, double *b, size_t nb, double *c, size_t nc,
                     size_t dfirst = 0, size_t dstride = 0, size_t dcount = 0);
int gf_launch_fill_bytes(hipStream_t s, void *p, size_t bytes);
// Zeroes a[0..na), c[0..nc) and the band entries of the structure (ptr, cols)
// (rows < nrows), stored as the band ab (row stride ld, half-bandwidth hb).
int gf_launch_clear_csr(hipStream_t s, double *a, size_t na, double *c, size_t nc, double *ab, int ld, int hb,
                        const int *ptr, const int *cols, int nrows);
// p[0] = a, p[1] = b (a kernel, so it can sit in a captured graph)
int gf_launch_set2(hipStream_t s, double *p, double a, double b);
// Returns the number of kernel launches issued (>= 0) or a negative status.
int gf_beam_assemble(gf_model *m, const double *lfp);
int gf_beam_commit(gf_model *m);
int gf_shell_assemble(gf_model *m, const double *lfp, const double *tp);  // elements + follower pressure
int gf_shell_eval_device(int count, const double *x, const double *prop, const double *d, const double *gp,
                         double *K_out, double *f_out);
int gf_shell_commit(gf_model *m);
// Gauss-point means of the committed state for the results writer, colour-sorted:
// Beam_1 [nb][12] (sigma_r, epsilon_r), Shell_1 [ns][24] (n_r1 m_r1 n_r2 m_r2
// eta_r1 kappa_r1 eta_r2 kappa_r2).
int gf_beam_results(gf_model *m, double *d_out);
int gf_shell_results(gf_model *m, double *d_out);
void gf_shell_host_const(const double *x18, const double *prop, double *rec /*[SC_COUNT]*/,
                         double *prec /*[SP_COUNT]*/, double *gp /*[SG_COUNT]*/);
gf_gmat gf_model_gmat(const gf_model *m);
int gf_nodal_loads(gf_model *m, const double *tp);
int gf_scan_residual(gf_model *m);
int gf_update_disp(gf_model *m, bool zero_norms);
int gf_node_commit(gf_model *m);

int gf_btd_alloc(gf_btd &b, int m, int nblk, int elem_size);
void gf_btd_free(gf_btd &b);
// Both return the number of kernel launches issued (>= 0) or a negative status.
int gf_btd_factor(gf_btd &b, hipStream_t s, int *d_flag);
// x = K^-1 rhs (factors in place); work: n entries, may alias rhs, not x
int gf_btd_solve(gf_btd &b, hipStream_t s, const double *rhs, double *x, int n, double *work);
// factorisation with the right-hand side reduced alongside: x = K^-1 rhs
int gf_btd_factor_solve(gf_btd &b, hipStream_t s, int *d_flag, const double *rhs, double *x, int n, double *work);

// Asynchronous halves of gf_assemble and gf_solve + gf_update: *_enqueue puts the
// work on the model's stream, *_result reads the norms once the stream passed it.
int gf_assemble_enqueue(gf_model *m, bool copy_norms);
int gf_assemble_result(gf_model *m, double *res_out);
int gf_solve_update_enqueue(gf_model *m, bool then_assemble);
int gf_solve_update_result(gf_model *m, int *info_fail, double *inc_out);

int gf_commit_enqueue(gf_model *m);

// The Static loop as a resumable state machine (gf_static.cpp).
struct gf_static_run {
    enum Phase { PH_STEP, PH_ASSEMBLED, PH_REFINE, PH_SOLVED, PH_ITER_END, PH_DONE };
    gf_model *m = nullptr;
    gf_static_params p{};
    gf_static_report r{};
    Phase phase = PH_DONE;
    hipEvent_t ev = nullptr;
    bool has_f = false, has_m = false, aborted = false, diverged = false, res_conv = false, gl_conv = false;
    double force_crit = 0, moment_crit = 0, time_step = 0, time = 0, last_converged = 0, last_rel = 0;
    int n_conv_evaluated = 0, convergence_counter = 0, it = 0, info_fail = 0;
    std::chrono::steady_clock::time_point t0, t_mark;
    int start(gf_model *model, const gf_static_params &params, int *step_iters, int cap);
    int advance();   // 1: device work pending (event ev), 0: finished, < 0: error
    int wait();      // blocks on ev
    ~gf_static_run();

  private:
    int enqueue_assemble();
    int record();
};

int gf_mixed_alloc(gf_model *m);
void gf_mixed_free(gf_model *m);
int gf_mixed_factor(gf_model *m);
int gf_mixed_refine_begin(gf_model *m);
int gf_mixed_sweep(gf_model *m);      // enqueue one refinement sweep (norms copied to h_scal)
int gf_mixed_fallback(gf_model *m);   // enqueue the FP64 factorisation + solve
// Mixed-precision refinement in progress (gf_solve_update_enqueue started it):
// after the last enqueued sweep's event, decides and enqueues the next step.
// *more = 1: another sweep is enqueued; 0: the update (and assembly) are.
int gf_refine_step(gf_model *m, int *more);

int gf_band_pad(int n);
size_t gf_band_inv_elems(int n);   // per-panel triangular inverses of a factor of order n
int gf_band_factor(void *ab, int elem_size, int n, int hb, hipStream_t s, int *flag, void *inv);
int gf_band_solve(const void *ab, int elem_size, int n, int hb, hipStream_t s, double *x, double *work,
                  const void *inv);

int gf_beam_eval_device(int count, const double *x, const double *cs, const double *prop,
                        const double *d, const double *gp, double *K_out, double *f_out);
void gf_beam_host_const(const double *x9, const double *cs, const double *prop, int kind /*GF_BEAM_KIND_**/,
                        double *rec /*[BC_COUNT]*/, double *gp /*[BG_COUNT]*/);
