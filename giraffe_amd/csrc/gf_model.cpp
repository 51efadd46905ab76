// Host side of libgiraffe_amd: model set-up and the per-phase C ABI.
//
// Reference: Database::PreCalc src/Database.cpp:704-757; Solution::DOFsActive /
// SetGlobalDOFs src/Solution.cpp:121-224, 40-118; NodalLoad::UpdateforSolutionStep
// src/NodalLoad.cpp:236-266; Solution::Zeros / Clear / MountLocal /
// MountElementLoads / MountLoads / MountGlobal / UpdateDisps / SaveConfiguration /
// RestoreConfiguration src/Solution.cpp:227-574, 798-863; sparsesystem
// src/SparseMatrix.cpp:74-166.
//
// Device layout (HBM):
//   node state    disp[nn][6], copy[nn][6] (X Y Z alpha), euler[nn][3]   FP64
//   DOF map       dof[nn][6] -> band row or -1; row_nc[n_free]           int32
//   Beam_1        conn[3][nb], const[33][nb], gp state[30][nb]           SoA, colour-sorted
//   system        r, x [nblk*m] FP64; A,B,C [nblk][m][m] block-tridiagonal
#include "gf_internal.h"

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

static thread_local std::string g_err;
void gf_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *gf_last_error(void) { return g_err.c_str(); }
extern "C" int gf_version(void) { return 100; }


extern "C" int gf_default_options(gf_options *o)
{
    if (!o) return GF_E_ARG;
    o->precision = GF_PREC_FP64;
    o->device = 0;
    o->stream = nullptr;
    o->refine_max = 12;
    o->refine_tol = 1e-11;
    o->refine_reuse = 1;
    o->solver = GF_SOLVER_AUTO;
    return GF_OK;
}

template <typename T> static int dalloc(T **p, size_t n)
{
    return gf_pool_alloc((void **)p, sizeof(T) * (n ? n : 1));
}
template <typename T> static int upload(gf_model *m, T *dst, const T *src, size_t n)
{
    if (n == 0) return GF_OK;
    GF_HIP(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, m->stream));
    return GF_OK;
}

// Launchers that return a launch count (>= 0) or a negative status.
#define GF_TRYN(var, x)              \
    int var = (x);                   \
    do {                             \
        if (var < 0) return var;     \
    } while (0)

static void tick(gf_model *m, int ph, bool start)
{
    if (!m->timing) return;
    (void)hipEventRecord(m->ev[2 * ph + (start ? 0 : 1)], m->stream);
    if (!start) m->ev_pending[ph] = true;
}
// Accumulates the phases whose end events the stream has passed (enqueue/result pairs).
static int flush_timing(gf_model *m)
{
    for (int ph = 0; ph < GF_NPHASE; ph++) {
        if (m->timing && m->ev_pending[ph]) {
            float ms = 0.f;
            GF_HIP(hipEventSynchronize(m->ev[2 * ph + 1]));
            GF_HIP(hipEventElapsedTime(&ms, m->ev[2 * ph], m->ev[2 * ph + 1]));
            m->t_ms[ph] += ms;
            m->launches[ph] += m->pending_launches[ph];
        }
        m->ev_pending[ph] = false;
        m->pending_launches[ph] = 0;
    }
    return m->kprobe.flush();
}

thread_local gf_kprobe *gf_probe = nullptr;

int gf_kprobe::begin(hipStream_t s, int c)
{
    const int cap = 8192;
    if (used == cap && flush() != GF_OK) return -1;
    if (2 * used + 2 > (int)ev.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess) return -1;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return -1;
        }
        ev.push_back(a);
        ev.push_back(b);
        cls.push_back(0);
        units.push_back(0.0);
    }
    const int slot = used++;
    cls[slot] = c;
    (void)hipEventRecord(ev[2 * slot], s);
    return slot;
}

void gf_kprobe::end(hipStream_t s, int slot, double u)
{
    units[slot] = u;
    (void)hipEventRecord(ev[2 * slot + 1], s);
}

int gf_kprobe::flush()
{
    if (!used) return GF_OK;
    GF_HIP(hipEventSynchronize(ev[2 * used - 1]));
    for (int i = 0; i < used; i++) {
        float t = 0.f;
        GF_HIP(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
        ms[cls[i]] += t;
        units_done[cls[i]] += units[i];
        launches[cls[i]]++;
    }
    used = 0;
    return GF_OK;
}

gf_kprobe::~gf_kprobe()
{
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

static gf_kprobe *probe_of(gf_model *m) { return m->timing && m->kprobe.mask ? &m->kprobe : nullptr; }

static int solve_fp64_enqueue(gf_model *m);

gf_gmat gf_model_gmat(const gf_model *m)
{
    gf_gmat g = {};
    g.layout = m->layout;
    g.m = m->btd.m;
    g.A = (double *)m->btd.A;
    g.B = (double *)m->btd.B;
    g.C = (double *)m->btd.C;
    g.ab = m->d_band;
    g.hb = m->half_band;
    g.ld = m->band_ld;
    return g;
}

static void release(gf_model *m)
{
    if (!m) return;
    // the cache hands freed blocks to the next model: nothing may still run on them
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->solve_graph) (void)hipGraphExecDestroy(m->solve_graph);
    if (m->iter_graph) (void)hipGraphExecDestroy(m->iter_graph);
    if (m->f32_factor_graph) (void)hipGraphExecDestroy(m->f32_factor_graph);
    if (m->sweep_graph) (void)hipGraphExecDestroy(m->sweep_graph);
    m->f32_factor_graph = m->sweep_graph = nullptr;
    m->solve_graph = nullptr;
    for (int i = 0; i < 2 * GF_NPHASE; i++)
        if (m->ev[i]) (void)hipEventDestroy(m->ev[i]);
    void *bufs[] = {m->d_disp, m->d_copy, m->d_euler, m->d_dof, m->d_row_nc, m->d_bconn, m->d_bconst, m->d_bgp,
                    m->d_ln_node, m->d_ln_off, m->d_ln_ids, m->d_lrec, m->d_tab_t, m->d_tab_v, m->d_r, m->d_x,
                    m->d_scal, m->d_copy0, m->d_bgp0, m->d_sconn, m->d_sconst, m->d_spconst, m->d_sgp,
                    m->d_sgp0, m->d_spload, m->d_spflag, m->d_band, m->d_swork, m->d_band_inv, m->d_bbuoy,
                    m->d_front, m->d_front_e, m->d_front_v, m->d_csr_ptr, m->d_csr_col, m->d_shell_rec,
                    m->d_shell_prec, m->d_shell_K, m->d_shell_f};
    for (void *b : bufs) gf_pool_free(b);
    gf_front_free(m->fronts);
    if (m->h_scal) gf_pinned_free(m->h_scal);
    gf_btd_free(m->btd);
    gf_mixed_free(m);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

// Greedy colouring: no two elements of a colour share a node.  Returns the
// colour-sorted element order and the colour offsets.
static int colour_elements(const int *conn, int ne, int nper, int nn, std::vector<int> &order, std::vector<int> &off)
{
    std::vector<unsigned long long> used(nn, 0ull);
    std::vector<int> color(ne);
    int ncol = 0;
    for (int e = 0; e < ne; e++) {
        unsigned long long u = 0;
        for (int a = 0; a < nper; a++) u |= used[conn[nper * (size_t)e + a]];
        int c = 0;
        while (c < 64 && (u >> c & 1ull)) c++;
        if (c == 64) {
            gf_set_error("colouring needs more than 64 colours");
            return GF_E_UNSUPPORTED;
        }
        color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < nper; a++) used[conn[nper * (size_t)e + a]] |= 1ull << c;
    }
    order.resize(ne);
    off.assign(ncol + 1, 0);
    for (int e = 0; e < ne; e++) off[color[e] + 1]++;
    for (int c = 0; c < ncol; c++) off[c + 1] += off[c];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int e = 0; e < ne; e++) order[fill[color[e]]++] = e;
    return GF_OK;
}

// Reverse Cuthill-McKee order of the nodes (Cuthill & McKee 1969; George & Liu,
// "Computer Solution of Large Sparse Positive Definite Systems", 1981): nodes are
// adjacent when an element holds both.  Each connected component starts at a
// pseudo-peripheral node (repeated breadth-first searches from a node of least
// degree in the last level, while the eccentricity grows), is numbered level by
// level with the neighbours of each node in increasing degree, and the whole
// order is reversed.  Nodes no element holds come last.
static std::vector<int> rcm_order(const gf_model_desc *d, int nn)
{
    std::vector<std::vector<int>> adj(nn);
    auto link = [&](const int *cn, int nper) {
        for (int a = 0; a < nper; a++)
            for (int b = 0; b < nper; b++)
                if (a != b) adj[cn[a]].push_back(cn[b]);
    };
    for (int e = 0; e < d->n_beams; e++) link(d->beam_conn + 3 * (size_t)e, 3);
    for (int e = 0; e < d->n_shells; e++) link(d->shell_conn + 6 * (size_t)e, 6);
    for (auto &v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<int> deg(nn);
    for (int i = 0; i < nn; i++) deg[i] = (int)adj[i].size();
    for (auto &v : adj)
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> level(nn, -1), order, stamp(nn, -1);
    order.reserve(nn);
    // breadth-first levels from s over the component (stamp marks the visit round)
    int round = 0;
    auto bfs = [&](int s, std::vector<int> &seq) {
        seq.clear();
        seq.push_back(s);
        stamp[s] = round;
        level[s] = 0;
        for (size_t h = 0; h < seq.size(); h++)
            for (int v : adj[seq[h]])
                if (stamp[v] != round) {
                    stamp[v] = round;
                    level[v] = level[seq[h]] + 1;
                    seq.push_back(v);
                }
        round++;
        return level[seq.back()];
    };
    std::vector<char> done(nn, 0);
    std::vector<int> seq;
    std::vector<int> by_degree(nn);
    for (int i = 0; i < nn; i++) by_degree[i] = i;
    std::stable_sort(by_degree.begin(), by_degree.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    for (int s0 : by_degree) {
        if (done[s0] || deg[s0] == 0) continue;
        int s = s0, ecc = bfs(s, seq);
        for (int tries = 0; tries < 16; tries++) {
            int best = -1;
            for (size_t k = seq.size(); k-- > 0 && level[seq[k]] == ecc;)
                if (best < 0 || deg[seq[k]] < deg[best]) best = seq[k];
            std::vector<int> seq2;
            const int e2 = bfs(best, seq2);
            if (e2 <= ecc) break;
            s = best;
            ecc = e2;
            seq.swap(seq2);
        }
        // Cuthill-McKee from s: neighbours in increasing degree (adj is sorted so)
        const size_t first = order.size();
        order.push_back(s);
        done[s] = 1;
        for (size_t h = first; h < order.size(); h++)
            for (int v : adj[order[h]])
                if (!done[v]) {
                    done[v] = 1;
                    order.push_back(v);
                }
    }
    std::reverse(order.begin(), order.end());
    for (int i = 0; i < nn; i++)
        if (!done[i]) order.push_back(i);
    return order;
}

// Band entries of the tangent's structure, row by row: for each free DOF row,
// the free DOF columns of every node that shares an element with the row's node
// (itself included), in increasing column order. With count_only, ptr alone.
static void tangent_structure(const gf_model_desc *d, const gf_model *m, std::vector<int> &ptr, std::vector<int> &col,
                              bool count_only = false)
{
    const int nn = m->nn, nb = m->nb, ne = m->nb + m->ns;
    auto conn = [&](int e, int &nper) {
        nper = e < nb ? 3 : 6;
        return e < nb ? d->beam_conn + 3 * (size_t)e : d->shell_conn + 6 * (size_t)(e - nb);
    };
    std::vector<int> eoff(nn + 1, 0);
    for (int e = 0, np; e < ne; e++) {
        const int *cn = conn(e, np);
        for (int a = 0; a < np; a++) eoff[cn[a] + 1]++;
    }
    for (int v = 0; v < nn; v++) eoff[v + 1] += eoff[v];
    std::vector<int> elist(eoff[nn]), at(eoff.begin(), eoff.end() - 1);
    for (int e = 0, np; e < ne; e++) {
        const int *cn = conn(e, np);
        for (int a = 0; a < np; a++) elist[at[cn[a]]++] = e;
    }
    // columns of each node that holds a free DOF
    const int W = gf_plan_workers(36LL * eoff[nn]);
    std::vector<std::vector<int>> ncols(nn);
    const int chunk = 256, nchunk = (nn + chunk - 1) / chunk;
    gf_parallel_for(nchunk, W, [&](int, int x) {
        std::vector<int> nbr;
        for (int a = x * chunk; a < std::min(nn, (x + 1) * chunk); a++) {
            bool has = false;
            for (int c = 0; c < 6; c++) has |= m->dof[6 * (size_t)a + c] >= 0;
            if (!has) continue;
            nbr.clear();
            for (int q = eoff[a], np; q < eoff[a + 1]; q++) {
                const int *cn = conn(elist[q], np);
                nbr.insert(nbr.end(), cn, cn + np);
            }
            std::sort(nbr.begin(), nbr.end());
            nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
            std::vector<int> &cols = ncols[a];
            for (int b : nbr)
                for (int c = 0; c < 6; c++)
                    if (m->dof[6 * (size_t)b + c] >= 0) cols.push_back(m->dof[6 * (size_t)b + c]);
            std::sort(cols.begin(), cols.end());
        }
    });
    const int nf = m->n_free;
    ptr.assign(nf + 1, 0);
    for (int r = 0; r < nf; r++) ptr[r + 1] = ptr[r] + (int)ncols[m->row_nodecomp[r] / 6].size();
    col.clear();
    if (count_only) return;
    col.resize(ptr[nf]);
    const int rchunk = 1024;
    gf_parallel_for((nf + rchunk - 1) / rchunk, W, [&](int, int x) {
        for (int r = x * rchunk; r < std::min(nf, (x + 1) * rchunk); r++) {
            const std::vector<int> &cols = ncols[m->row_nodecomp[r] / 6];
            std::copy(cols.begin(), cols.end(), col.begin() + ptr[r]);
        }
    });
}

// Host-side planning shared by create() and gf_model_footprint: validation,
// DOFsActive + SetGlobalDOFs, half-bandwidth, storage layout, colourings. Without
// `lists` (sizes only), the multifrontal path counts the tangent's entries but
// builds neither the column array nor the per-tile entry lists.
static int plan(const gf_model_desc *d, gf_model *m, int &bm, int &nblk, int solver, bool lists = true)
{
    if (solver < GF_SOLVER_AUTO || solver > GF_SOLVER_FRONTS) { gf_set_error("unknown solver"); return GF_E_ARG; }
    if (d->n_nodes <= 0 || !d->node_xyz) { gf_set_error("model has no nodes"); return GF_E_ARG; }
    m->nn = d->n_nodes;
    m->nb = d->n_beams;
    m->ns = d->n_shells;
    m->nl = d->n_loads;
    const int nn = m->nn, nb = m->nb, ns = m->ns;
    for (int e = 0; e < nb; e++)
        for (int a = 0; a < 3; a++) {
            int n = d->beam_conn[3 * e + a];
            if (n < 0 || n >= nn) { gf_set_error("beam connectivity out of range"); return GF_E_ARG; }
        }
    for (int e = 0; e < ns; e++)
        for (int a = 0; a < 6; a++) {
            int n = d->shell_conn[6 * e + a];
            if (n < 0 || n >= nn) { gf_set_error("shell connectivity out of range"); return GF_E_ARG; }
        }
    // DOFsActive (Beam_1: 6 per node; Shell_1: u at corners, u + alpha at mid-sides,
    // src/Shell_1.cpp:51-68) + NodalConstraint + SetGlobalDOFs (node-major numbering)
    m->active.assign(6 * (size_t)nn, 0);
    m->fix.assign(6 * (size_t)nn, 0);
    for (int e = 0; e < nb; e++)
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 6; k++) m->active[6 * (size_t)d->beam_conn[3 * e + a] + k] = 1;
    for (int e = 0; e < ns; e++)
        for (int a = 0; a < 6; a++)
            for (int k = 0; k < (a < 3 ? 3 : 6); k++) m->active[6 * (size_t)d->shell_conn[6 * e + a] + k] = 1;
    if (d->node_fix)
        for (size_t i = 0; i < 6 * (size_t)nn; i++) m->fix[i] = d->node_fix[i] ? 1 : 0;
    // Node order of the DOF numbering: the caller's, or reverse Cuthill-McKee
    // when that gives the narrower band (the reference leaves the ordering to
    // PARDISO's nested dissection, src/SparseMatrix.cpp:115-117; a band solver
    // needs a bandwidth-reducing order of its own).
    std::vector<int> lo(nb + ns), hi(nb + ns);
    auto number = [&](const std::vector<int> &order) {
        m->dof.assign(6 * (size_t)nn, -1);
        m->row_nodecomp.clear();
        int nf = 0;
        for (int i : order)
            for (int k = 0; k < 6; k++) {
                size_t id = 6 * (size_t)i + k;
                if (m->active[id] && !m->fix[id]) {
                    m->dof[id] = nf++;
                    m->row_nodecomp.push_back((int)id);
                }
            }
        m->n_free = nf;
        // half-bandwidth; row span of every element
        int hb = 0;
        for (int e = 0; e < nb + ns; e++) {
            const int *cn = e < nb ? d->beam_conn + 3 * (size_t)e : d->shell_conn + 6 * (size_t)(e - nb);
            int nper = e < nb ? 3 : 6;
            int l = 1 << 30, h = -1;
            for (int a = 0; a < nper; a++)
                for (int k = 0; k < 6; k++) {
                    int r = m->dof[6 * (size_t)cn[a] + k];
                    if (r >= 0) { l = std::min(l, r); h = std::max(h, r); }
                }
            lo[e] = l; hi[e] = h;
            if (h >= 0) hb = std::max(hb, h - l);
        }
        return hb;
    };
    std::vector<int> natural(nn);
    for (int i = 0; i < nn; i++) natural[i] = i;
    int hb = number(natural);
    if (m->n_free == 0) { gf_set_error("model has no free DOFs"); return GF_E_ARG; }
    m->node_order_rcm = false;
    if (hb > 0 && getenv("GF_NO_RCM") == nullptr) {
        const std::vector<int> rcm = rcm_order(d, nn);
        const int hb_rcm = number(rcm);
        if (hb_rcm < hb) {
            hb = hb_rcm;
            m->node_order_rcm = true;
        } else {
            number(natural);
        }
    }
    const int nf = m->n_free;
    m->half_band = hb;
    // narrow bands: block-tridiagonal form with the smallest block size m <= 32
    bm = -1;
    for (int cand = std::max(1, (hb + 2) / 2); cand <= 32 && bm < 0; cand++) {
        bool ok = true;
        for (int e = 0; e < nb + ns && ok; e++)
            if (hi[e] >= 0 && hi[e] / cand - lo[e] / cand > 1) ok = false;
        if (ok) bm = cand;
    }
    // the block-tridiagonal scatter finds a row's block through a float product
    // (gm_block_of), exact for rows below 2^24: larger orders take the band layout
    if (bm > 0 && (long long)(nf + bm - 1) / bm * bm >= (1LL << 24)) bm = -1;
    m->layout = bm > 0 ? GM_BTD : GM_BAND;
    nblk = 0;
    if (m->layout == GM_BTD) {
        nblk = (nf + bm - 1) / bm;
        m->npad = nblk * bm;
    } else {
        m->npad = gf_band_pad(nf);
    }
    // Multifrontal solve over a nested-dissection order (PARDISO's strategy,
    // src/SparseMatrix.cpp:115-117): when forced, or for a band too wide for the
    // block-tridiagonal form when it takes less than half the band LU's flops
    // (2 n hb^2) in fewer launches (the band LU's critical path is one launch
    // per 64 columns).
    m->fronts_on = false;
    m->fronts = gf_fronts();
    if (solver == GF_SOLVER_FRONTS || (solver == GF_SOLVER_AUTO && m->layout == GM_BAND)) {
        const int npad = gf_band_pad(nf);
        GF_TRY(gf_nd_plan(m->fronts, d, m->dof, nn, npad));
        const double band_flops = 2.0 * (double)nf * (double)hb * (double)hb;
        const bool take = solver == GF_SOLVER_FRONTS ||
                          (m->fronts.flops < 0.5 * band_flops && gf_front_factor_launches(m->fronts) < npad / 64);
        if (take) {
            m->fronts_on = true;
            tangent_structure(d, m, m->csr_ptr, m->csr_col, !lists);
            if (lists) GF_TRY(gf_nd_orig_lists(m->fronts, m->csr_ptr, m->csr_col, npad));
            m->lists_done = lists;
            m->layout = GM_BAND;
            m->npad = npad;
            bm = -1;
            nblk = 0;
        } else {
            m->fronts = gf_fronts();
        }
    }
    GF_TRY(colour_elements(d->beam_conn, nb, 3, nn, m->beam_order, m->beam_color_off));
    GF_TRY(colour_elements(d->shell_conn, ns, 6, nn, m->shell_order, m->shell_color_off));
    return GF_OK;
}

// Device bytes create() allocates for a planned model.
static size_t footprint(const gf_model_desc *d, const gf_model *m, int precision, int bm, int nblk)
{
    const size_t nn = (size_t)m->nn, nb = (size_t)m->nb, ns = (size_t)m->ns, nf = (size_t)m->n_free;
    const size_t npad = (size_t)m->npad;
    size_t b = 0;
    b += nn * (6 + 6 + 6 + 3) * 8 + nn * 6 * 4 + nf * 4;                       // node state, DOF map
    b += nb * (3 * 4 + BC_COUNT * 8 + 2 * BG_COUNT * 8 + 8);                  // Beam_1 / Pipe_1 records
    b += ns * (6 * 4 + SC_COUNT * 8 + SP_COUNT * 8 + 2 * SG_COUNT * 8 + 8);   // Shell_1 records
    b += ns * (3 * 276 + 6 * 57 + 729 + 27) * 8;                               // Shell_1 assembly records
    size_t items = d->n_loads ? (size_t)d->load_set_off[d->n_loads] : 0;
    size_t rows = d->n_loads ? (size_t)d->load_tab_off[d->n_loads] : 0;
    b += (size_t)d->n_loads * NL_REC * 8 + items * 12 + rows * 7 * 8;         // loads and tables
    b += 3 * npad * 8 + 256;                                                   // r, x, work, scalars
    size_t tangent;
    if (m->layout == GM_BTD) {
        size_t blocks = 6 * (size_t)nblk + 2 * (size_t)nblk;                  // A B C Bi X Y + saved couplings
        tangent = blocks * (size_t)bm * bm;
    } else if (m->fronts_on) {
        const gf_fronts &fp = m->fronts;
        const size_t factor = (size_t)(fp.f_elems + fp.e_elems);
        b += (npad * (size_t)(2 * m->half_band + 1) + factor + (size_t)fp.v_elems) * 8;
        b += (fp.rows.size() + fp.child.size() + fp.inv.size() + fp.items.size()) * 4 + fp.child_inv.size() * 8 +
             fp.fr.size() * sizeof(gf_front_dev);
        // CSR structure and per-tile entry lists (3 ints per entry, 2 per assembly item)
        const size_t nnz = m->csr_ptr.empty() ? 0 : (size_t)m->csr_ptr.back();
        b += (m->csr_ptr.size() + nnz + 3 * nnz + fp.items.size() / 2) * 4;
        if (precision == GF_PREC_MIXED) b += factor * 4 + 3 * npad * 8;
        return b;
    } else {
        const int hb = m->half_band;
        tangent = npad * (size_t)(hb + gf_band_ku(hb) + 1) + gf_band_inv_elems(m->npad);
    }
    b += tangent * 8;
    if (precision == GF_PREC_MIXED) b += tangent * 4 + 3 * npad * 8;
    return b;
}

static int create(const gf_model_desc *d, const gf_options *opt, gf_model *m)
{
    int bm = -1, nblk = 0;
    if (m->planned) {
        bm = m->plan_bm;
        nblk = m->plan_nblk;
        GF_TRY(gf_model_plan_lists(d, m));
    } else {
        GF_TRY(plan(d, m, bm, nblk, m->solver));
    }
    const int nn = m->nn, nb = m->nb, ns = m->ns, nf = m->n_free;
    const int hb = m->half_band;
    // device + stream
    m->device = opt ? opt->device : 0;
    GF_HIP(hipSetDevice(m->device));
    if (opt && opt->stream) {
        m->stream = (hipStream_t)opt->stream;
    } else {
        GF_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        m->own_stream = true;
    }
    m->precision = opt ? opt->precision : GF_PREC_FP64;
    m->use_graph = getenv("GF_NO_GRAPH") == nullptr;
    if (m->precision != GF_PREC_FP64 && m->precision != GF_PREC_MIXED) {
        gf_set_error("unknown precision");
        return GF_E_ARG;
    }
    // node state
    std::vector<double> copy(6 * (size_t)nn, 0.0);
    for (int i = 0; i < nn; i++)
        for (int k = 0; k < 3; k++) copy[6 * (size_t)i + k] = d->node_xyz[3 * (size_t)i + k];
    m->h_nodes_ref.assign(d->node_xyz, d->node_xyz + 3 * (size_t)nn);
    m->h_beam_conn.assign(d->beam_conn, d->beam_conn + 3 * (size_t)nb);
    m->beam_kind.assign(nb, GF_BEAM_KIND_BEAM1);
    for (int e = 0; e < nb && d->beam_kind; e++) {
        if (d->beam_kind[e] != GF_BEAM_KIND_BEAM1 && d->beam_kind[e] != GF_BEAM_KIND_PIPE1) {
            gf_set_error("unknown beam kind");
            return GF_E_ARG;
        }
        m->beam_kind[e] = d->beam_kind[e];
    }
    m->h_shell_conn.assign(d->shell_conn, d->shell_conn + 6 * (size_t)ns);
    {
        // element numbers: a permutation of 1..nb+ns (the reference stores element i as number i+1)
        m->beam_number.resize(nb);
        m->shell_number.resize(ns);
        std::vector<char> seen((size_t)nb + ns + 1, 0);
        for (int e = 0; e < nb + ns; e++) {
            const int num = e < nb ? (d->beam_number ? d->beam_number[e] : e + 1)
                                   : (d->shell_number ? d->shell_number[e - nb] : e + 1);
            if (num < 1 || num > nb + ns || seen[num]) {
                gf_set_error("element numbers must be a permutation of 1..n_beams+n_shells");
                return GF_E_ARG;
            }
            seen[num] = 1;
            (e < nb ? m->beam_number[e] : m->shell_number[e - nb]) = num;
        }
    }
    GF_TRY(dalloc(&m->d_disp, 6 * (size_t)nn));
    GF_TRY(dalloc(&m->d_copy, 6 * (size_t)nn));
    GF_TRY(dalloc(&m->d_euler, 3 * (size_t)nn));
    GF_TRY(dalloc(&m->d_dof, 6 * (size_t)nn));
    GF_TRY(dalloc(&m->d_row_nc, (size_t)nf));
    GF_TRY(upload(m, m->d_copy, copy.data(), copy.size()));
    GF_TRY(dalloc(&m->d_copy0, copy.size()));
    GF_TRY(upload(m, m->d_copy0, copy.data(), copy.size()));
    GF_TRY(upload(m, m->d_dof, m->dof.data(), m->dof.size()));
    GF_TRY(upload(m, m->d_row_nc, m->row_nodecomp.data(), (size_t)nf));
    GF_TRY(gf_launch_fill(m->stream, m->d_disp, 6 * (size_t)nn, 0.0));
    GF_TRY(gf_launch_fill(m->stream, m->d_euler, 3 * (size_t)nn, 0.0));
    // beams, colour-sorted SoA
    if (nb > 0) {
        size_t ld = (size_t)nb;
        std::vector<int> conn(3 * ld);
        std::vector<double> cst(BC_COUNT * ld), gps(BG_COUNT * ld), rec(BC_COUNT), g0(BG_COUNT);
        for (int s = 0; s < nb; s++) {
            int e = m->beam_order[s];
            double x[9];
            for (int a = 0; a < 3; a++)
                for (int k = 0; k < 3; k++) x[3 * a + k] = d->node_xyz[3 * (size_t)d->beam_conn[3 * e + a] + k];
            gf_beam_host_const(x, d->beam_cs + 9 * (size_t)e, d->beam_prop + GF_BEAM_NPROP * (size_t)e,
                               m->beam_kind[e], rec.data(), g0.data());
            for (int k = 0; k < BC_COUNT; k++) cst[k * ld + s] = rec[k];
            for (int k = 0; k < BG_COUNT; k++) gps[k * ld + s] = g0[k];
            for (int a = 0; a < 3; a++) conn[a * ld + s] = d->beam_conn[3 * e + a];
        }
        GF_TRY(dalloc(&m->d_bconn, conn.size()));
        GF_TRY(dalloc(&m->d_bconst, cst.size()));
        GF_TRY(dalloc(&m->d_bgp, gps.size()));
        GF_TRY(upload(m, m->d_bconn, conn.data(), conn.size()));
        GF_TRY(upload(m, m->d_bconst, cst.data(), cst.size()));
        GF_TRY(upload(m, m->d_bgp, gps.data(), gps.size()));
        GF_TRY(dalloc(&m->d_bgp0, gps.size()));
        GF_TRY(upload(m, m->d_bgp0, gps.data(), gps.size()));
        // Pipe_1 buoyancy below the ocean surface (colour-sorted like the records)
        bool buoyant = false;
        std::vector<double> buoy(ld, 0.0);
        if (d->ocean_on && d->beam_buoy)
            for (int s = 0; s < nb; s++) {
                buoy[s] = d->beam_buoy[m->beam_order[s]];
                buoyant = buoyant || buoy[s] != 0.0;
            }
        if (buoyant) {
            GF_TRY(dalloc(&m->d_bbuoy, buoy.size()));
            GF_TRY(upload(m, m->d_bbuoy, buoy.data(), buoy.size()));
            for (int k = 0; k < 3; k++) m->ocean_surface[k] = d->ocean_surface[k];
        }
    }
    // shells, colour-sorted SoA
    if (ns > 0) {
        size_t ld = (size_t)ns;
        std::vector<int> conn(6 * ld);
        std::vector<double> cst(SC_COUNT * ld), pcs(SP_COUNT * ld), gps(SG_COUNT * ld);
        std::vector<double> rec(SC_COUNT), prec(SP_COUNT), g0(SG_COUNT);
        for (int s2 = 0; s2 < ns; s2++) {
            int e = m->shell_order[s2];
            double x[18];
            for (int a = 0; a < 6; a++)
                for (int k = 0; k < 3; k++) x[3 * a + k] = d->node_xyz[3 * (size_t)d->shell_conn[6 * e + a] + k];
            gf_shell_host_const(x, d->shell_prop + GF_SHELL_NPROP * (size_t)e, rec.data(), prec.data(), g0.data());
            for (int k = 0; k < SC_COUNT; k++) cst[k * ld + s2] = rec[k];
            for (int k = 0; k < SP_COUNT; k++) pcs[k * ld + s2] = prec[k];
            for (int k = 0; k < SG_COUNT; k++) gps[k * ld + s2] = g0[k];
            for (int a = 0; a < 6; a++) conn[a * ld + s2] = d->shell_conn[6 * e + a];
        }
        GF_TRY(dalloc(&m->d_sconn, conn.size()));
        GF_TRY(dalloc(&m->d_sconst, cst.size()));
        GF_TRY(dalloc(&m->d_spconst, pcs.size()));
        GF_TRY(dalloc(&m->d_sgp, gps.size()));
        GF_TRY(dalloc(&m->d_sgp0, gps.size()));
        GF_TRY(upload(m, m->d_sconn, conn.data(), conn.size()));
        GF_TRY(upload(m, m->d_sconst, cst.data(), cst.size()));
        GF_TRY(upload(m, m->d_spconst, pcs.data(), pcs.size()));
        GF_TRY(upload(m, m->d_sgp, gps.data(), gps.size()));
        GF_TRY(upload(m, m->d_sgp0, gps.data(), gps.size()));
        // ShellLoad -> per element (colour-sorted position) load index and AreaUpdate flag
        std::vector<int> pos(ns);
        for (int s2 = 0; s2 < ns; s2++) pos[m->shell_order[s2]] = s2;
        std::vector<int> pl(ns, -1), pf(ns, 0);
        for (int l = 0; l < m->nl; l++) {
            if (d->load_kind[l] != GF_LOAD_SHELL_PRESSURE) continue;
            for (int q = d->load_set_off[l]; q < d->load_set_off[l + 1]; q++) {
                int e = d->load_set_items[q];
                if (e < 0 || e >= ns) { gf_set_error("pressure load element out of range"); return GF_E_ARG; }
                if (pl[pos[e]] >= 0) { gf_set_error("more than one pressure load on a shell element"); return GF_E_UNSUPPORTED; }
                pl[pos[e]] = l;
                pf[pos[e]] = d->load_flags ? d->load_flags[l] : 0;
                m->any_pressure = true;
            }
        }
        GF_TRY(dalloc(&m->d_spload, pl.size()));
        GF_TRY(dalloc(&m->d_spflag, pf.size()));
        GF_TRY(upload(m, m->d_spload, pl.data(), pl.size()));
        GF_TRY(upload(m, m->d_spflag, pf.data(), pf.size()));
        // phase-1 records and pressure terms of the assembly (gf_shell_assemble)
        GF_TRY(dalloc(&m->d_shell_rec, 3 * (size_t)ns * 276));
        if (m->any_pressure) GF_TRY(dalloc(&m->d_shell_prec, 6 * (size_t)ns * 57));
        GF_TRY(dalloc(&m->d_shell_K, 729 * (size_t)ns));
        GF_TRY(dalloc(&m->d_shell_f, 27 * (size_t)ns));
    }
    // nodal loads grouped by node; multipliers as UpdateforSolutionStep
    {
        std::vector<std::pair<int, int>> pairs;  // (node, load)
        std::vector<double> lrec((size_t)std::max(1, m->nl) * NL_REC, 0.0);
        for (int l = 0; l < m->nl; l++) {
            int kind = d->load_kind[l];
            if (kind == GF_LOAD_SHELL_PRESSURE) {
                double *R = lrec.data() + (size_t)l * NL_REC;
                R[NL_KIND] = kind;
                R[NL_TAB0] = d->load_tab_off[l];
                R[NL_TABN] = d->load_tab_off[l + 1] - d->load_tab_off[l];
                continue;
            }
            if (kind != GF_LOAD_NODAL && kind != GF_LOAD_NODAL_FOLLOWER) { gf_set_error("unknown load kind"); return GF_E_ARG; }
            int cf[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
            for (int s = d->load_set_off[l]; s < d->load_set_off[l + 1]; s++) {
                int node = d->load_set_items[s];
                if (node < 0 || node >= nn) { gf_set_error("load node out of range"); return GF_E_ARG; }
                for (int j = 0; j < 3; j++) {
                    if (m->active[6 * (size_t)node + j]) cf[j]++;
                    if (m->active[6 * (size_t)node + 3 + j]) cm[j]++;
                }
                pairs.emplace_back(node, l);
            }
            double *R = lrec.data() + (size_t)l * NL_REC;
            R[NL_KIND] = kind;
            for (int i = 0; i < 9; i++) R[NL_CS + i] = d->load_cs[9 * (size_t)l + i];
            for (int j = 0; j < 3; j++) {
                R[NL_MULT + j] = 1.0 / cf[j];
                R[NL_MULT + 3 + j] = 1.0 / cm[j];
            }
            R[NL_TAB0] = d->load_tab_off[l];
            R[NL_TABN] = d->load_tab_off[l + 1] - d->load_tab_off[l];
        }
        std::stable_sort(pairs.begin(), pairs.end(),
                         [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
        std::vector<int> ln_node, ln_off(1, 0), ln_ids;
        for (size_t i = 0; i < pairs.size(); i++) {
            if (i == 0 || pairs[i].first != pairs[i - 1].first) {
                if (i) ln_off.push_back((int)ln_ids.size());
                ln_node.push_back(pairs[i].first);
            }
            ln_ids.push_back(pairs[i].second);
        }
        if (!pairs.empty()) ln_off.push_back((int)ln_ids.size());
        m->n_load_nodes = (int)ln_node.size();
        int rows = m->nl ? d->load_tab_off[m->nl] : 0;
        GF_TRY(dalloc(&m->d_ln_node, ln_node.size()));
        GF_TRY(dalloc(&m->d_ln_off, ln_off.size()));
        GF_TRY(dalloc(&m->d_ln_ids, ln_ids.size()));
        GF_TRY(dalloc(&m->d_lrec, lrec.size()));
        GF_TRY(dalloc(&m->d_tab_t, (size_t)rows));
        GF_TRY(dalloc(&m->d_tab_v, 6 * (size_t)rows));
        GF_TRY(upload(m, m->d_ln_node, ln_node.data(), ln_node.size()));
        GF_TRY(upload(m, m->d_ln_off, ln_off.data(), ln_off.size()));
        GF_TRY(upload(m, m->d_ln_ids, ln_ids.data(), ln_ids.size()));
        GF_TRY(upload(m, m->d_lrec, lrec.data(), lrec.size()));
        GF_TRY(upload(m, m->d_tab_t, d->tab_time, (size_t)rows));
        GF_TRY(upload(m, m->d_tab_v, d->tab_val, 6 * (size_t)rows));
    }
    for (int k = 0; k < 3; k++) m->gravity[k] = d->gravity[k];
    m->gravity_on = d->gravity_on;
    m->gravity_mode = d->gravity_on ? GF_GRAVITY_RAMP_UP : GF_GRAVITY_OFF;
    // global system
    size_t npad = (size_t)m->npad;
    GF_TRY(dalloc(&m->d_r, npad));
    GF_TRY(dalloc(&m->d_x, npad));
    GF_TRY(dalloc(&m->d_swork, npad));
    GF_TRY(gf_launch_fill(m->stream, m->d_x, npad, 0.0));  // padding rows stay zero (the fronts solve writes free rows only)
    if (m->layout == GM_BTD) {
        GF_TRY(gf_btd_alloc(m->btd, bm, nblk, 8));
    } else if (m->fronts_on) {
        // the band holds the assembled tangent only; the factor lives in the fronts
        m->band_ld = 2 * hb;
        m->band_elems = npad * (size_t)(m->band_ld + 1);
        GF_TRY(dalloc(&m->d_band, m->band_elems));
        GF_TRY(dalloc(&m->d_front, (size_t)m->fronts.f_elems));
        GF_TRY(dalloc(&m->d_front_e, (size_t)m->fronts.e_elems));
        GF_TRY(dalloc(&m->d_front_v, (size_t)m->fronts.v_elems));
        GF_TRY(gf_front_upload(m->fronts, m->stream));
        GF_TRY(dalloc(&m->d_csr_ptr, m->csr_ptr.size()));
        GF_TRY(dalloc(&m->d_csr_col, m->csr_col.size()));
        GF_TRY(upload(m, m->d_csr_ptr, m->csr_ptr.data(), m->csr_ptr.size()));
        GF_TRY(upload(m, m->d_csr_col, m->csr_col.data(), m->csr_col.size()));
        // the assembly clears only the structure: everything else of the band is
        // zero from here on, the padding rows' unit diagonal included
        const size_t nfr = (size_t)m->n_free;
        GF_TRY(gf_launch_clear3(m->stream, nullptr, 0, m->d_band, m->band_elems, nullptr, 0,
                                nfr * (size_t)(m->band_ld + 1) + (size_t)hb, (size_t)(m->band_ld + 1), npad - nfr));
    } else {
        m->band_ld = hb + gf_band_ku(hb);
        m->band_elems = npad * (size_t)(m->band_ld + 1);
        GF_TRY(dalloc(&m->d_band, m->band_elems));
        GF_TRY(dalloc(&m->d_band_inv, gf_band_inv_elems(m->npad)));
    }
    if (m->precision == GF_PREC_MIXED) GF_TRY(gf_mixed_alloc(m));
    // scalar block (doubles): [0..4] increment norms, [8] solver failure flag
    // (int), [16..18] residual norms, [24..25] refinement norms; one copy of
    // [0..18] returns a whole Newton iteration's results
    GF_TRY(dalloc(&m->d_scal, GF_SCAL_N));
    m->d_flag = (int *)(m->d_scal + GF_SCAL_FLAG);
    GF_TRY(gf_pinned_alloc((void **)&m->h_scal, GF_SCAL_N * sizeof(double)));
    for (int i = 0; i < 2 * GF_NPHASE; i++) GF_HIP(hipEventCreate(&m->ev[i]));
    GF_HIP(hipStreamSynchronize(m->stream));
    return GF_OK;
}

extern "C" int gf_reset_state(gf_model *m)
{
    if (!m) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    GF_HIP(hipMemcpyAsync(m->d_copy, m->d_copy0, sizeof(double) * 6 * (size_t)m->nn, hipMemcpyDeviceToDevice,
                          m->stream));
    if (m->nb)
        GF_HIP(hipMemcpyAsync(m->d_bgp, m->d_bgp0, sizeof(double) * BG_COUNT * (size_t)m->nb,
                              hipMemcpyDeviceToDevice, m->stream));
    if (m->ns)
        GF_HIP(hipMemcpyAsync(m->d_sgp, m->d_sgp0, sizeof(double) * SG_COUNT * (size_t)m->ns,
                              hipMemcpyDeviceToDevice, m->stream));
    GF_TRY(gf_launch_fill(m->stream, m->d_disp, 6 * (size_t)m->nn, 0.0));
    GF_TRY(gf_launch_fill(m->stream, m->d_euler, 3 * (size_t)m->nn, 0.0));
    GF_HIP(hipStreamSynchronize(m->stream));
    m->factored = false;
    // the mixed-precision back-off starts over: a solve after a reset is
    // independent of the ones before it
    m->fallback_streak = 0;
    m->fp64_direct_left = 0;
    m->backoff_len = GF_MIXED_BACKOFF;
    m->f32_have = false;
    return GF_OK;
}

extern "C" int gf_timing_enable(gf_model *m, int on)
{
    if (!m) return GF_E_ARG;
    m->timing = on != 0;
    return GF_OK;
}

extern "C" int gf_kernel_probe(gf_model *m, unsigned mask)
{
    if (!m || mask >= (1u << GF_NKC)) return GF_E_ARG;
    GF_TRY(m->kprobe.flush());
    m->kprobe.mask = mask;
    return GF_OK;
}

extern "C" int gf_kernel_probe_read(gf_model *m, double *ms_out, long long *launches_out, double *units_out, int reset)
{
    if (!m) return GF_E_ARG;
    GF_TRY(m->kprobe.flush());
    gf_kprobe &k = m->kprobe;
    for (int c = 0; c < GF_NKC; c++) {
        if (ms_out) ms_out[c] = k.ms[c];
        if (launches_out) launches_out[c] = k.launches[c];
        if (units_out) units_out[c] = k.units_done[c];
        if (reset) { k.ms[c] = 0.0; k.launches[c] = 0; k.units_done[c] = 0.0; }
    }
    return GF_OK;
}

extern "C" int gf_timing_read(gf_model *m, double *ms_out, long long *launches_out, int reset)
{
    if (!m) return GF_E_ARG;
    for (int i = 0; i < GF_NPHASE; i++) {
        if (ms_out) ms_out[i] = m->t_ms[i];
        if (launches_out) launches_out[i] = m->launches[i];
        if (reset) { m->t_ms[i] = 0.0; m->launches[i] = 0; }
    }
    return GF_OK;
}

extern "C" int gf_model_plan(const gf_model_desc *desc, long long *o)
{
    if (!desc || !o) return GF_E_ARG;
    gf_model tmp;
    int bm = -1, nblk = 0;
    GF_TRY(plan(desc, &tmp, bm, nblk, GF_SOLVER_AUTO, false));
    o[0] = tmp.n_free;
    o[1] = tmp.half_band;
    o[2] = tmp.layout == GM_BTD ? bm : 0;
    o[3] = tmp.node_order_rcm ? 1 : 0;
    return GF_OK;
}

extern "C" int gf_front_plan(const gf_model_desc *desc, long long *o, int *front_out, int *rows_out, int *dof_out)
{
    if (!desc || !o) return GF_E_ARG;
    gf_model tmp;
    int bm = -1, nblk = 0;
    GF_TRY(plan(desc, &tmp, bm, nblk, GF_SOLVER_FRONTS));
    const gf_fronts &fp = tmp.fronts;
    o[0] = fp.n_front;
    o[1] = fp.n_level;
    o[2] = (long long)fp.rows.size();
    o[3] = fp.f_elems;
    o[4] = fp.e_elems;
    o[5] = (long long)fp.flops;
    o[6] = gf_front_factor_launches(fp);
    o[7] = gf_front_solve_launches(fp);
    if (front_out) {
        for (int f = 0; f < fp.n_front; f++) {
            int *q = front_out + 6 * (size_t)f;
            int lev = 0;
            while (fp.lev_off[lev + 1] <= f) lev++;
            q[0] = fp.fr[f].k;
            q[1] = 64 * fp.fr[f].kb;
            q[2] = fp.fr[f].mp;
            q[3] = lev;
            q[4] = -1;
            q[5] = fp.fr[f].rows;
        }
        for (int f = 0; f < fp.n_front; f++)
            for (int c = fp.fr[f].ch0; c < fp.fr[f].ch1; c++) front_out[6 * (size_t)fp.child[c] + 4] = f;
    }
    if (rows_out) std::copy(fp.rows.begin(), fp.rows.end(), rows_out);
    if (dof_out) std::copy(tmp.dof.begin(), tmp.dof.end(), dof_out);
    return GF_OK;
}

extern "C" int gf_model_footprint(const gf_model_desc *desc, const gf_options *opt, unsigned long long *bytes)
{
    if (!desc || !bytes) return GF_E_ARG;
    gf_model tmp;
    int bm = -1, nblk = 0;
    GF_TRY(plan(desc, &tmp, bm, nblk, opt ? opt->solver : GF_SOLVER_AUTO, false));
    *bytes = footprint(desc, &tmp, opt ? opt->precision : GF_PREC_FP64, bm, nblk);
    return GF_OK;
}

static gf_model *new_model(const gf_options *opt)
{
    gf_model *m = new gf_model();
    if (opt) {
        m->device = opt->device;
        m->refine_max = opt->refine_max;
        m->refine_tol = opt->refine_tol;
        m->refine_reuse = opt->refine_reuse != 0;
        m->solver = opt->solver;
    }
    return m;
}

int gf_model_preplan(const gf_model_desc *d, const gf_options *opt, gf_model **out, unsigned long long *bytes)
{
    if (!d || !out || !bytes) { gf_set_error("null argument"); return GF_E_ARG; }
    *out = nullptr;
    gf_model *m = new_model(opt);
    int st = plan(d, m, m->plan_bm, m->plan_nblk, m->solver, false);
    if (st != GF_OK) {
        delete m;
        return st;
    }
    m->planned = true;
    *bytes = footprint(d, m, opt ? opt->precision : GF_PREC_FP64, m->plan_bm, m->plan_nblk);
    *out = m;
    return GF_OK;
}

int gf_model_plan_lists(const gf_model_desc *d, gf_model *m)
{
    if (!m->fronts_on || m->lists_done) return GF_OK;
    tangent_structure(d, m, m->csr_ptr, m->csr_col);
    GF_TRY(gf_nd_orig_lists(m->fronts, m->csr_ptr, m->csr_col, m->npad));
    m->lists_done = true;
    return GF_OK;
}

int gf_model_create_planned(const gf_model_desc *d, const gf_options *opt, gf_model *m)
{
    if (!d || !m || !m->planned) { gf_set_error("model not planned"); return GF_E_ARG; }
    return create(d, opt, m);
}

extern "C" int gf_model_create(const gf_model_desc *desc, const gf_options *opt, gf_model **out)
{
    if (!desc || !out) { gf_set_error("null argument"); return GF_E_ARG; }
    *out = nullptr;
    gf_model *m = new_model(opt);
    int st = create(desc, opt, m);
    if (st != GF_OK) {
        release(m);
        return st;
    }
    *out = m;
    return GF_OK;
}

extern "C" int gf_model_destroy(gf_model *m)
{
    if (m) {
        (void)hipSetDevice(m->device);
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        release(m);
    }
    return GF_OK;
}

extern "C" int gf_model_info(const gf_model *m, long long *o)
{
    if (!m || !o) return GF_E_ARG;
    o[0] = m->n_free;
    o[1] = m->half_band;
    o[2] = m->layout == GM_BTD ? m->btd.m : 0;
    o[3] = m->layout == GM_BTD ? m->btd.nblk : m->npad;
    o[4] = (long long)std::max(m->beam_color_off.size(), m->shell_color_off.size()) - 1;
    o[5] = m->layout == GM_BTD ? (long long)(3 * (size_t)m->btd.nblk * m->btd.bsz * m->btd.elem_size)
                               : (long long)((m->band_elems + m->fronts.f_elems + m->fronts.e_elems) * 8);
    o[6] = m->fronts_on ? 2 : m->layout;
    o[7] = m->precision;
    return GF_OK;
}

extern "C" int gf_model_stats(const gf_model *m, long long *out)
{
    if (!m || !out) return GF_E_ARG;
    out[0] = m->n_solves;
    out[1] = m->n_sweeps;
    out[2] = m->n_fallbacks;
    out[3] = m->n_direct;
    out[4] = m->n_f32_factor;
    return GF_OK;
}

// The step's load time t = t_last + dt and gravity factor into the scalar
// block, where the assembly kernels read them.  The factor follows
// BoolTable::GetLinearFactorAtCurrentTime (src/BoolTable.cpp:84-109): a ramp
// 0 -> 1 across the solution window where gravity turns on, 1 while it stays
// on, 1 -> 0 where it turns off.
static int set_step_scalars(gf_model *m)
{
    const double t = m->t_last + m->dt;
    const double r = (t - m->sol_start) / (m->sol_end - m->sol_start);
    double lf = 0.0;
    if (m->gravity_on) {
        switch (m->gravity_mode) {
        case GF_GRAVITY_RAMP_UP: lf = r; break;
        case GF_GRAVITY_ON: lf = 1.0; break;
        case GF_GRAVITY_RAMP_DOWN: lf = 1.0 - r; break;
        default: lf = 0.0;
        }
    }
    return gf_launch_set2(m->stream, m->d_scal + GF_SCAL_TIME, t, lf);
}

extern "C" int gf_set_gravity_mode(gf_model *m, int mode)
{
    if (!m || mode < GF_GRAVITY_OFF || mode > GF_GRAVITY_RAMP_DOWN) return GF_E_ARG;
    m->gravity_mode = mode;
    GF_HIP(hipSetDevice(m->device));
    return set_step_scalars(m);
}

extern "C" int gf_begin_step(gf_model *m, double t_last, double dt)
{
    if (!m) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    m->t_last = t_last;
    m->dt = dt;
    m->factored = false;
    GF_TRY(set_step_scalars(m));
    return gf_launch_fill(m->stream, m->d_disp, 6 * (size_t)m->nn, 0.0);
}

// Sets the solution-time window used by the gravity ramp (BoolTable factor).
extern "C" int gf_set_solution_window(gf_model *m, double start, double end)
{
    if (!m) return GF_E_ARG;
    m->sol_start = start;
    m->sol_end = end;
    GF_HIP(hipSetDevice(m->device));
    return set_step_scalars(m);
}

// Enqueues Clear + MountLocal/ElementLoads/Loads + MountGlobal and the residual
// max-norm scan (norms at d_scal[GF_SCAL_RES..+2]); copy_norms: also copy them
// to h_scal (a caller that copies the scalar block itself passes false).
int gf_assemble_enqueue(gf_model *m, bool copy_norms)
{
    gf_probe_scope ps(probe_of(m));
    GF_HIP(hipSetDevice(m->device));
    const size_t npad = (size_t)m->npad;
    double *tan = m->layout == GM_BTD ? (double *)m->btd.A : m->d_band;
    const size_t ntan = m->layout == GM_BTD ? 3 * (size_t)m->btd.nblk * m->btd.bsz : m->band_elems;
    // the padding rows [n_free, npad) carry a unit diagonal (all in the last block
    // row for the block-tridiagonal layout)
    size_t dfirst, dstride;
    const size_t dcount = npad - (size_t)m->n_free, nf = (size_t)m->n_free;
    if (m->layout == GM_BTD) {
        const size_t mb = (size_t)m->btd.m, last = (size_t)m->btd.nblk - 1, i0 = nf - last * mb;
        dfirst = (size_t)m->btd.nblk * m->btd.bsz + last * m->btd.bsz + i0 * mb + i0;  // inside B
        dstride = mb + 1;
    } else {
        dfirst = nf * (size_t)(m->band_ld + 1) + (size_t)m->half_band;
        dstride = (size_t)(m->band_ld + 1);
    }
    if (m->fronts_on)
        GF_TRY(gf_launch_clear_csr(m->stream, m->d_r, npad, m->d_scal + GF_SCAL_RES, 8, m->d_band, m->band_ld,
                                   m->half_band, m->d_csr_ptr, m->d_csr_col, m->n_free));
    else
        GF_TRY(gf_launch_clear3(m->stream, m->d_r, npad, tan, ntan, m->d_scal + GF_SCAL_RES, 8, dfirst, dstride, dcount));
    // load time and gravity ramp factor from the scalar block (set by gf_begin_step)
    const double *tp = m->d_scal + GF_SCAL_TIME, *lfp = m->d_scal + GF_SCAL_LF;
    GF_TRY(gf_nodal_loads(m, tp));
    tick(m, GF_PH_ASSEMBLE, true);
    GF_TRYN(nl_beam, gf_beam_assemble(m, lfp));
    GF_TRYN(nl_shell, gf_shell_assemble(m, lfp, tp));
    tick(m, GF_PH_ASSEMBLE, false);
    m->pending_launches[GF_PH_ASSEMBLE] += nl_beam + nl_shell;
    m->factored = false;
    GF_TRY(gf_scan_residual(m));
    if (copy_norms)
        GF_HIP(hipMemcpyAsync(m->h_scal + GF_SCAL_RES, m->d_scal + GF_SCAL_RES, 3 * sizeof(double),
                              hipMemcpyDeviceToHost, m->stream));
    return GF_OK;
}

// Reads what gf_assemble_enqueue left in h_scal (the stream must have passed it).
int gf_assemble_result(gf_model *m, double *res_out)
{
    GF_TRY(flush_timing(m));
    unsigned long long *u = (unsigned long long *)(m->h_scal + GF_SCAL_RES);
    double v0, v1;
    memcpy(&v0, &u[0], 8);
    memcpy(&v1, &u[1], 8);
    res_out[0] = v0;
    res_out[1] = v1;
    res_out[2] = u[2] ? 1.0 : 0.0;
    return GF_OK;
}

extern "C" int gf_assemble(gf_model *m, double *res_out)
{
    if (!m) return GF_E_ARG;
    GF_TRY(gf_assemble_enqueue(m, true));
    GF_HIP(hipStreamSynchronize(m->stream));
    double res[3];
    GF_TRY(gf_assemble_result(m, res));
    if (res_out) memcpy(res_out, res, sizeof(res));
    return GF_OK;
}

// UpdateDisps, the next iteration's assembly (then_assemble) and one copy of
// the scalar block, after a solve whose increment is in d_x.
static int update_tail_enqueue(gf_model *m, bool then_assemble)
{
    GF_TRY(gf_update_disp(m, false));
    if (then_assemble) GF_TRY(gf_assemble_enqueue(m, false));
    const int nscal = then_assemble ? GF_SCAL_RES + 3 : GF_SCAL_FLAG + 1;
    GF_HIP(hipMemcpyAsync(m->h_scal, m->d_scal, nscal * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    return GF_OK;
}

static void note_fallback(gf_model *m, bool fell_back, bool divergent = false, bool slow = false);

// Outside timing, a launch sequence that depends on the model only is
// captured into *exec on first use (kernel and copy nodes) and replayed: a
// dependent launch costs about 1.6 us from a graph instead of 2.4 us from the
// stream.  enqueue() returns its launch count or a negative status.
template <typename F> static int replay_or_launch(gf_model *m, hipGraphExec_t &exec, F enqueue)
{
    if (m->timing || !m->use_graph) return enqueue();
    if (!exec) {
        GF_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->stream, &g);
        if (st < 0 || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (st < 0) return st;
            GF_HIP(e);
        }
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) exec = nullptr;
        GF_HIP(ei);
    }
    GF_HIP(hipGraphLaunch(exec, m->stream));
    return 1;
}

static int mixed_sweep_enqueue(gf_model *m)
{
    return replay_or_launch(m, m->sweep_graph, [m] { return gf_mixed_sweep(m); });
}

// The FP32 factorisation of the equilibrated current tangent.
static int mixed_factor_enqueue(gf_model *m)
{
    tick(m, GF_PH_FACTOR, true);
    GF_TRYN(nl_f32, replay_or_launch(m, m->f32_factor_graph, [m] { return gf_mixed_factor(m); }));
    tick(m, GF_PH_FACTOR, false);
    m->pending_launches[GF_PH_FACTOR] += nl_f32;
    m->n_f32_factor++;
    m->solve_f32++;
    m->f32_have = true;
    m->f32_reused = false;
    return GF_OK;
}

// Mixed precision: the FP32 factorisation of the equilibrated tangent and the
// first refinement sweep (the model is then `refining` until gf_refine_step
// has enqueued the update).  With refine_reuse the FP32 factor of an earlier
// tangent (and its equilibration) is kept as long as the refinement against
// the current FP64 tangent keeps converging: the refinement's residual is
// always the current tangent's, so only the number of sweeps changes; a
// refinement that stagnates on the kept factor refactors and continues.
static int mixed_start_enqueue(gf_model *m)
{
    m->solve_sweeps = 1;
    m->solve_f32 = 0;
    if (m->refine_reuse && m->f32_have) m->f32_reused = true;
    else GF_TRY(mixed_factor_enqueue(m));
    GF_TRY(gf_mixed_refine_begin(m));
    tick(m, GF_PH_SOLVE, true);
    GF_TRYN(nl_sw, mixed_sweep_enqueue(m));
    tick(m, GF_PH_SOLVE, false);
    m->pending_launches[GF_PH_SOLVE] += nl_sw;
    m->n_sweeps++;
    m->refining = true;
    m->refine_it = 1;
    m->refine_prev = 1e308;
    return GF_OK;
}

// After the last sweep's copy has reached the host: converged (the correction is
// below refine_tol of x, or stagnates at the FP64 rounding floor), another sweep,
// or the FP64 factorisation (an unusable FP32 pivot, a non-finite or stagnating
// correction, refine_max sweeps).  *next: 0 converged, 1 another sweep enqueued,
// 2 FP64 fallback enqueued.
static int refine_decide(gf_model *m, int *next)
{
    GF_TRY(flush_timing(m));
    int flag = 0;
    double dn, xn;
    memcpy(&flag, &m->h_scal[GF_SCAL_FLAG], 4);
    memcpy(&dn, m->h_scal + GF_SCAL_REF, 8);
    memcpy(&xn, m->h_scal + GF_SCAL_REF + 1, 8);
    const bool finite = dn == dn && xn == xn;
    bool ok = false, fall = (flag & GF_FLAG_FAIL) != 0 || !finite, divergent = false;
    if (!fall) {
        if (dn <= m->refine_tol * xn) ok = true;
        else if (m->refine_it > 1 && dn > 0.5 * m->refine_prev) {
            if (dn <= 1e-10 * xn) ok = true;
            else {
                fall = true;
                // no contraction at all: the FP32 factor does not approximate
                // this model's tangent (spectral radius of the sweep >= 1)
                divergent = m->refine_it == 2 && dn >= m->refine_prev;
            }
        } else if (m->refine_it >= std::max(1, m->refine_max)) fall = true;
    }
    // A kept FP32 factor (an earlier tangent's) is replaced by the current
    // tangent's when its refinement would fall back, or when it contracts by
    // less than 10x per sweep: then the sweeps still needed cost more than a
    // factorisation and a fresh refinement (at the plate's sizes one FP32
    // factorisation costs about eight sweeps, and a fresh factor contracts
    // about 300x per sweep).
    bool refactor = false;
    if (m->f32_reused && (flag & GF_FLAG_FAIL) == 0) {
        if (fall) refactor = true;
        else if (!ok && m->refine_it > 1 && dn > 0.1 * m->refine_prev) refactor = true;
    }
    if ((!ok && !fall) || refactor) {
        if (refactor) {
            GF_TRY(mixed_factor_enqueue(m));
            // continue from the current x (already accurate to the kept
            // factor's sweeps), unless it is not finite
            if (!finite) GF_TRY(gf_mixed_refine_begin(m));
            m->refine_prev = 1e308;
            m->refine_it = 1;
        } else {
            m->refine_prev = dn;
            m->refine_it++;
        }
        m->n_sweeps++;
        m->solve_sweeps++;
        tick(m, GF_PH_SOLVE, true);
        GF_TRYN(l, mixed_sweep_enqueue(m));
        tick(m, GF_PH_SOLVE, false);
        m->pending_launches[GF_PH_SOLVE] += l;
        *next = 1;
        return GF_OK;
    }
    m->refining = false;
    m->factored = true;
    // a converged FP32-refined solve that cost more than the FP64 one would have
    const bool slow = ok && !fall && m->fronts_on &&
                      GF_MIXED_F32_SWEEPS * m->solve_f32 + m->solve_sweeps > GF_MIXED_F64_SWEEPS;
    note_fallback(m, fall, divergent, slow);
    *next = ok ? 0 : 2;
    if (fall) {
        // a fresh FP32 factor that could not drive the refinement is not kept
        m->f32_have = false;
        m->n_fallbacks++;
        tick(m, GF_PH_FACTOR, true);
        GF_TRYN(l, gf_mixed_fallback(m));
        tick(m, GF_PH_FACTOR, false);
        m->pending_launches[GF_PH_FACTOR] += l;
    }
    return GF_OK;
}

int gf_refine_step(gf_model *m, int *more)
{
    gf_probe_scope ps(probe_of(m));
    GF_HIP(hipSetDevice(m->device));
    int next = 0;
    GF_TRY(refine_decide(m, &next));
    *more = next == 1 ? 1 : 0;
    return next == 1 ? GF_OK : update_tail_enqueue(m, m->refine_then_assemble);
}

static int solve_fp64_launches(gf_model *m, int &nl_fac, int &nl_sol);

// The whole FP64 Newton iteration -- clear of the increment norms and flag, the
// factorisation and solve, UpdateDisps, the next iteration's assembly and
// residual scan, the copy of the scalar block -- captured once per model into a
// graph and replayed on every iteration (the launch sequence depends on the
// model only; the load time and gravity factor are read from the scalar block).
// The graph holds kernel and copy nodes, no memset.
static int iteration_graph_launch(gf_model *m)
{
    if (!m->iter_graph) {
        GF_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        int nf = 0, ns = 0;
        int st = gf_launch_fill_bytes(m->stream, m->d_scal, (GF_SCAL_FLAG + 1) * sizeof(double));
        if (!st) st = solve_fp64_launches(m, nf, ns);
        if (!st) st = update_tail_enqueue(m, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->stream, &g);
        if (st || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (st) return st;
            GF_HIP(e);
        }
        const hipError_t ei = hipGraphInstantiate(&m->iter_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) m->iter_graph = nullptr;
        GF_HIP(ei);
    }
    GF_HIP(hipGraphLaunch(m->iter_graph, m->stream));
    m->factored = false;  // the iteration ends with the next assembly
    return GF_OK;
}

// Enqueues the tangent solve, UpdateDisps and the increment norms (d_scal[0..4],
// failure flag at d_scal[GF_SCAL_FLAG]).  then_assemble: the next iteration's
// assembly and residual norms follow in the same batch (the Newton loop
// assembles after every solve that does not end the load step; assembly changes
// no state, so a batch whose step ends on the increments simply does not use
// it).  One copy brings the scalar block back.  Mixed precision (not backed
// off): only the FP32 factorisation and the first refinement sweep are enqueued
// here; gf_refine_step continues once their event has completed, and enqueues
// the update when the refinement has ended.
int gf_solve_update_enqueue(gf_model *m, bool then_assemble)
{
    gf_probe_scope ps(probe_of(m));
    GF_HIP(hipSetDevice(m->device));
    m->n_solves++;
    m->pending_info_fail = -1;
    const bool mixed = m->precision == GF_PREC_MIXED && m->fp64_direct_left == 0;
    // FP64, or a mixed model backed off to FP64
    if (!mixed && m->precision == GF_PREC_MIXED) {
        m->fp64_direct_left--;
        m->n_direct++;
    }
    if (!mixed && then_assemble && !m->timing && m->use_graph) return iteration_graph_launch(m);
    GF_TRY(gf_launch_fill_bytes(m->stream, m->d_scal, (GF_SCAL_FLAG + 1) * sizeof(double)));
    if (mixed) {
        m->refine_then_assemble = then_assemble;
        return mixed_start_enqueue(m);
    }
    GF_TRY(solve_fp64_enqueue(m));
    return update_tail_enqueue(m, then_assemble);
}

int gf_solve_update_result(gf_model *m, int *info_fail, double *inc_out)
{
    GF_TRY(flush_timing(m));
    if (m->pending_info_fail >= 0) {
        *info_fail = m->pending_info_fail;
    } else {
        int flag = 0;
        memcpy(&flag, &m->h_scal[GF_SCAL_FLAG], 4);
        *info_fail = (flag & GF_FLAG_FAIL) ? 1 : 0;
    }
    unsigned long long *u = (unsigned long long *)m->h_scal;
    for (int i = 0; i < 4; i++) memcpy(&inc_out[i], &u[i], 8);
    inc_out[4] = u[4] ? 1.0 : 0.0;
    return GF_OK;
}

// The launches of one FP64 factorisation + solve (no events); launch counts out.
static int solve_fp64_launches(gf_model *m, int &nl_fac, int &nl_sol)
{
    const size_t npad = (size_t)m->npad;
    nl_sol = 0;
    if (m->fronts_on) {
        nl_fac = gf_front_factor(m->fronts, 8, m->d_band, m->band_ld, m->half_band, nullptr, m->d_front, m->d_front_e,
                                 m->stream, m->d_flag, m->d_r, m->d_front_v);
        if (nl_fac < 0) return nl_fac;
        nl_sol = gf_front_solve(m->fronts, 8, m->d_front, m->d_front_e, m->d_r, m->d_x, m->d_front_v, m->stream, true);
        return nl_sol < 0 ? nl_sol : GF_OK;
    }
    if (m->layout == GM_BTD)
        nl_fac = gf_btd_factor_solve(m->btd, m->stream, m->d_flag, m->d_r, m->d_x, (int)npad, m->d_swork);
    else nl_fac = gf_band_factor(m->d_band, 8, m->npad, m->half_band, m->stream, m->d_flag, m->d_band_inv);
    if (nl_fac < 0) return nl_fac;
    if (m->layout != GM_BTD) {
        GF_HIP(hipMemcpyAsync(m->d_x, m->d_r, sizeof(double) * npad, hipMemcpyDeviceToDevice, m->stream));
        nl_sol = gf_band_solve(m->d_band, 8, m->npad, m->half_band, m->stream, m->d_x, m->d_swork, m->d_band_inv);
    }
    return nl_sol < 0 ? nl_sol : GF_OK;
}

// FP64 factorisation + solve of K x = r on the model's stream (x in d_x).
// Block-tridiagonal: the right-hand side is reduced alongside the factorisation
// (the factor phase covers the forward sweep too; the solve phase is empty).
// The launch sequence depends on the model only, so outside timing it is
// captured once into a graph and replayed (a dependent launch costs about
// 1.6 us from a graph instead of 2.4 us from the stream on MI355X); the graph
// holds kernel nodes and one device-to-device copy, no memset.
static int solve_fp64_enqueue(gf_model *m)
{
    if (!m->timing && m->use_graph) {
        if (!m->solve_graph) {
            GF_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
            int nf = 0, ns = 0;
            const int st = solve_fp64_launches(m, nf, ns);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(m->stream, &g);
            if (st || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (st) return st;
                GF_HIP(e);
            }
            const hipError_t ei = hipGraphInstantiate(&m->solve_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) m->solve_graph = nullptr;
            GF_HIP(ei);
            m->graph_launches = nf + ns;
        }
        GF_HIP(hipGraphLaunch(m->solve_graph, m->stream));
        m->factored = true;
        return GF_OK;
    }
    tick(m, GF_PH_FACTOR, true);
    int nl_fac = 0, nl_sol = 0;
    if (m->fronts_on) {
        nl_fac = gf_front_factor(m->fronts, 8, m->d_band, m->band_ld, m->half_band, nullptr, m->d_front, m->d_front_e,
                                 m->stream, m->d_flag, m->d_r, m->d_front_v);
        if (nl_fac < 0) return nl_fac;
        tick(m, GF_PH_FACTOR, false);
        tick(m, GF_PH_SOLVE, true);
        nl_sol = gf_front_solve(m->fronts, 8, m->d_front, m->d_front_e, m->d_r, m->d_x, m->d_front_v, m->stream, true);
        if (nl_sol < 0) return nl_sol;
    } else if (m->layout == GM_BTD) {
        GF_TRY(solve_fp64_launches(m, nl_fac, nl_sol));
        tick(m, GF_PH_FACTOR, false);
        tick(m, GF_PH_SOLVE, true);
    } else {
        const size_t npad = (size_t)m->npad;
        nl_fac = gf_band_factor(m->d_band, 8, m->npad, m->half_band, m->stream, m->d_flag, m->d_band_inv);
        if (nl_fac < 0) return nl_fac;
        tick(m, GF_PH_FACTOR, false);
        tick(m, GF_PH_SOLVE, true);
        GF_HIP(hipMemcpyAsync(m->d_x, m->d_r, sizeof(double) * npad, hipMemcpyDeviceToDevice, m->stream));
        nl_sol = gf_band_solve(m->d_band, 8, m->npad, m->half_band, m->stream, m->d_x, m->d_swork, m->d_band_inv);
        if (nl_sol < 0) return nl_sol;
    }
    m->factored = true;
    tick(m, GF_PH_SOLVE, false);
    m->pending_launches[GF_PH_FACTOR] += nl_fac;
    m->pending_launches[GF_PH_SOLVE] += nl_sol;
    return GF_OK;
}

// Consecutive fallbacks of the mixed path; two in a row start a back-off of
// backoff_len FP64 solves.  The first FP32 attempt after a back-off counts as
// the second of a row, and a back-off that ends in another fallback doubles the
// next one (up to 64 times the first); an FP32 solve that converges resets both.
// A refinement whose second correction is no smaller than its first (the FP32
// factor of this tangent does not contract the error at all, as for slender
// beam chains whose equilibrated tangent is beyond FP32's range) starts the
// longest back-off at once.  On the multifrontal path a solve that converged
// but whose sweeps and FP32 factorisations cost more than an FP64 factorisation
// and solve (GF_MIXED_*_SWEEPS) starts a back-off too, without discarding the
// counts of fallbacks; the FP32 factor is then dropped, since the next attempt
// comes a back-off's tangents later.
static void note_fallback(gf_model *m, bool fell_back, bool divergent, bool slow)
{
    if (!fell_back && slow) {
        m->fp64_direct_left = m->backoff_len;
        m->backoff_len = std::min(2 * m->backoff_len, 64 * GF_MIXED_BACKOFF);
        m->fallback_streak = 0;
        m->f32_have = false;
        return;
    }
    if (!fell_back) {
        m->fallback_streak = 0;
        m->backoff_len = GF_MIXED_BACKOFF;
        return;
    }
    if (divergent) {
        m->fp64_direct_left = 64 * GF_MIXED_BACKOFF;
        m->backoff_len = 64 * GF_MIXED_BACKOFF;
        m->fallback_streak = 1;
        return;
    }
    if (++m->fallback_streak >= 2) {
        m->fp64_direct_left = m->backoff_len;
        m->backoff_len = std::min(2 * m->backoff_len, 64 * GF_MIXED_BACKOFF);
        m->fallback_streak = 1;
    }
}

extern "C" int gf_solve(gf_model *m, int *info_fail)
{
    if (!m || !info_fail) return GF_E_ARG;
    gf_probe_scope ps(probe_of(m));
    GF_HIP(hipSetDevice(m->device));
    *info_fail = 0;
    GF_TRY(gf_launch_fill_bytes(m->stream, m->d_flag, 4));
    m->n_solves++;
    // Mixed precision backs off after two fallbacks in a row: the next
    // GF_MIXED_BACKOFF solves factor in FP64 directly, then FP32 is tried again.
    const bool direct = m->precision == GF_PREC_MIXED && m->fp64_direct_left > 0;
    if (direct) {
        m->fp64_direct_left--;
        m->n_direct++;
    }
    if (m->precision == GF_PREC_MIXED && !direct) {
        // the refinement of gf_solve_update_enqueue, waiting for each sweep
        GF_TRY(mixed_start_enqueue(m));
        int next = 1;
        while (next == 1) {
            GF_HIP(hipStreamSynchronize(m->stream));
            GF_TRY(refine_decide(m, &next));
        }
    } else {
        GF_TRY(solve_fp64_enqueue(m));
    }
    GF_TRY(flush_timing(m));
    int flag = 0;
    GF_HIP(hipMemcpyAsync(&m->h_scal[GF_SCAL_FLAG], m->d_flag, 4, hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    GF_TRY(flush_timing(m));
    memcpy(&flag, &m->h_scal[GF_SCAL_FLAG], 4);
    if (flag & GF_FLAG_FAIL) *info_fail = 1;
    return GF_OK;
}

extern "C" int gf_update(gf_model *m, double *inc_out)
{
    if (!m) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    GF_TRY(gf_update_disp(m, true));
    if (inc_out) {
        GF_HIP(hipMemcpyAsync(m->h_scal, m->d_scal, 5 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        GF_HIP(hipStreamSynchronize(m->stream));
        unsigned long long *u = (unsigned long long *)m->h_scal;
        for (int i = 0; i < 4; i++) memcpy(&inc_out[i], &u[i], 8);
        inc_out[4] = u[4] ? 1.0 : 0.0;
    }
    return GF_OK;
}

int gf_commit_enqueue(gf_model *m)
{
    GF_HIP(hipSetDevice(m->device));
    GF_TRY(gf_node_commit(m));
    GF_TRY(gf_beam_commit(m));
    GF_TRY(gf_shell_commit(m));
    return GF_OK;
}

extern "C" int gf_commit(gf_model *m)
{
    if (!m) return GF_E_ARG;
    GF_TRY(gf_commit_enqueue(m));
    GF_HIP(hipStreamSynchronize(m->stream));
    return GF_OK;
}

extern "C" int gf_rollback(gf_model *m)
{
    if (!m) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    GF_TRY(gf_launch_fill(m->stream, m->d_disp, 6 * (size_t)m->nn, 0.0));
    return GF_OK;
}

extern "C" int gf_read_nodes(gf_model *m, double *out)
{
    if (!m || !out) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    std::vector<double> c(6 * (size_t)m->nn), eu(3 * (size_t)m->nn);
    GF_HIP(hipMemcpyAsync(c.data(), m->d_copy, sizeof(double) * c.size(), hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipMemcpyAsync(eu.data(), m->d_euler, sizeof(double) * eu.size(), hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    for (int i = 0; i < m->nn; i++) {
        for (int k = 0; k < 3; k++) out[6 * (size_t)i + k] = c[6 * (size_t)i + k];
        for (int k = 0; k < 3; k++) out[6 * (size_t)i + 3 + k] = eu[3 * (size_t)i + k];
    }
    return GF_OK;
}

extern "C" int gf_read_disp(gf_model *m, double *out)
{
    if (!m || !out) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    GF_HIP(hipMemcpyAsync(out, m->d_disp, sizeof(double) * 6 * (size_t)m->nn, hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    return GF_OK;
}

extern "C" int gf_write_disp(gf_model *m, const double *in)
{
    if (!m || !in) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    GF_HIP(hipMemcpyAsync(m->d_disp, in, sizeof(double) * 6 * (size_t)m->nn, hipMemcpyHostToDevice, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    return GF_OK;
}

static int read_rows(gf_model *m, const double *dev, double *out)
{
    GF_HIP(hipSetDevice(m->device));
    std::vector<double> v((size_t)m->n_free);
    GF_HIP(hipMemcpyAsync(v.data(), dev, sizeof(double) * v.size(), hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    std::fill(out, out + 6 * (size_t)m->nn, 0.0);
    for (int r = 0; r < m->n_free; r++) out[m->row_nodecomp[r]] = v[r];
    return GF_OK;
}

extern "C" int gf_read_residual(gf_model *m, double *out)
{
    if (!m || !out) return GF_E_ARG;
    return read_rows(m, m->d_r, out);
}

extern "C" int gf_read_increment(gf_model *m, double *out)
{
    if (!m || !out) return GF_E_ARG;
    return read_rows(m, m->d_x, out);
}

extern "C" int gf_read_tangent_dense(gf_model *m, double *out, long long cap)
{
    if (!m || !out) return GF_E_ARG;
    long long n6 = 6LL * m->nn;
    if (cap < n6 * n6) { gf_set_error("output buffer too small"); return GF_E_CAPACITY; }
    if (m->factored) { gf_set_error("tangent already factored"); return GF_E_STATE; }
    GF_HIP(hipSetDevice(m->device));
    if (m->layout == GM_BAND) {
        std::vector<double> ab(m->band_elems);
        GF_HIP(hipMemcpyAsync(ab.data(), m->d_band, m->band_elems * 8, hipMemcpyDeviceToHost, m->stream));
        GF_HIP(hipStreamSynchronize(m->stream));
        std::fill(out, out + n6 * n6, 0.0);
        int hb = m->half_band;
        for (int R = 0; R < m->n_free; R++)
            for (int Cc = std::max(0, R - hb); Cc <= std::min(m->n_free - 1, R + hb); Cc++)
                out[(size_t)m->row_nodecomp[R] * n6 + m->row_nodecomp[Cc]] =
                    ab[(size_t)R * (size_t)m->band_ld + (size_t)(Cc + hb)];
        return GF_OK;
    }
    size_t nbl = (size_t)m->btd.nblk * m->btd.bsz;
    std::vector<double> A(nbl), B(nbl), C(nbl);
    GF_HIP(hipMemcpyAsync(A.data(), m->btd.A, nbl * 8, hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipMemcpyAsync(B.data(), m->btd.B, nbl * 8, hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipMemcpyAsync(C.data(), m->btd.C, nbl * 8, hipMemcpyDeviceToHost, m->stream));
    GF_HIP(hipStreamSynchronize(m->stream));
    std::fill(out, out + n6 * n6, 0.0);
    int bm = m->btd.m;
    for (int R = 0; R < m->n_free; R++)
        for (int Cc = std::max(0, (R / bm - 1) * bm); Cc < std::min(m->n_free, (R / bm + 2) * bm); Cc++) {
            int bi = R / bm, bj = Cc / bm;
            size_t off = ((size_t)bi * bm + (R - bi * bm)) * bm + (Cc - bj * bm);
            double v = bj == bi ? B[off] : (bj == bi - 1 ? A[off] : C[off]);
            out[(size_t)m->row_nodecomp[R] * n6 + m->row_nodecomp[Cc]] = v;
        }
    return GF_OK;
}

extern "C" int gf_bandsystem(int n, int hb, const double *ab_rows, const double *b, double *x, int *info_fail)
{
    if (n <= 0 || hb < 0 || !ab_rows || !b || !x || !info_fail) return GF_E_ARG;
    *info_fail = 0;
    int npad = gf_band_pad(n);
    const size_t w = (size_t)(2 * hb + 1);                  // caller's row length
    const size_t ws = (size_t)(hb + gf_band_ku(hb) + 1);    // stored row length (room for the fill)
    size_t elems = (size_t)npad * ws;
    // rows of the band: row i holds columns i-hb .. i+hb; columns outside [0, n) must be zero
    std::vector<double> rows((size_t)npad * ws, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < (int)w; k++) {
            int c = i - hb + k;
            rows[(size_t)i * ws + k] = (c >= 0 && c < n) ? ab_rows[(size_t)i * w + k] : 0.0;
        }
    for (int i = n; i < npad; i++) rows[(size_t)i * ws + hb] = 1.0;
    double *d_ab = nullptr, *d_x = nullptr, *d_z = nullptr, *d_inv = nullptr;
    int *d_flag = nullptr;
    hipStream_t s = nullptr;
    int flag = 0;
    int st = GF_OK;
    auto run = [&]() -> int {
        GF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        GF_HIP(hipMalloc(&d_ab, elems * 8));
        GF_HIP(hipMalloc(&d_x, (size_t)npad * 8));
        GF_HIP(hipMalloc(&d_z, (size_t)npad * 8));
        GF_HIP(hipMalloc(&d_inv, gf_band_inv_elems(npad) * 8));
        GF_HIP(hipMalloc(&d_flag, 4));
        GF_TRY(gf_launch_fill(s, d_ab, elems, 0.0));
        GF_TRY(gf_launch_fill(s, d_x, (size_t)npad, 0.0));
        GF_TRY(gf_launch_fill_bytes(s, d_flag, 4));
        GF_HIP(hipMemcpyAsync(d_ab, rows.data(), rows.size() * 8, hipMemcpyHostToDevice, s));
        GF_HIP(hipMemcpyAsync(d_x, b, (size_t)n * 8, hipMemcpyHostToDevice, s));
        int l = gf_band_factor(d_ab, 8, npad, hb, s, d_flag, d_inv);
        if (l < 0) return l;
        l = gf_band_solve(d_ab, 8, npad, hb, s, d_x, d_z, d_inv);
        if (l < 0) return l;
        GF_HIP(hipMemcpyAsync(x, d_x, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        GF_HIP(hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        return GF_OK;
    };
    st = run();
    void *bufs[] = {d_ab, d_x, d_z, d_inv, d_flag};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
    if (st) return st;
    if (flag & GF_FLAG_FAIL) *info_fail = 1;
    return GF_OK;
}

extern "C" int gf_btdsystem(int m, int nblk, const double *A, const double *B, const double *C, const double *b,
                            double *x, int *info_fail)
{
    if (m < 1 || m > 32 || nblk < 1 || !A || !B || !C || !b || !x || !info_fail) return GF_E_ARG;
    *info_fail = 0;
    const size_t bs = (size_t)m * m, blocks = (size_t)nblk * bs, n = (size_t)nblk * m;
    gf_btd t;
    double *d_r = nullptr, *d_x = nullptr, *d_w = nullptr;
    int *d_flag = nullptr;
    hipStream_t s = nullptr;
    int flag = 0;
    auto run = [&]() -> int {
        GF_TRY(gf_btd_alloc(t, m, nblk, 8));
        GF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        GF_HIP(hipMalloc(&d_r, n * 8));
        GF_HIP(hipMalloc(&d_x, n * 8));
        GF_HIP(hipMalloc(&d_w, n * 8));
        GF_HIP(hipMalloc(&d_flag, 4));
        GF_TRY(gf_launch_fill_bytes(s, d_flag, 4));
        GF_HIP(hipMemcpyAsync(t.A, A, blocks * 8, hipMemcpyHostToDevice, s));
        GF_HIP(hipMemcpyAsync(t.B, B, blocks * 8, hipMemcpyHostToDevice, s));
        GF_HIP(hipMemcpyAsync(t.C, C, blocks * 8, hipMemcpyHostToDevice, s));
        // the couplings past either end do not exist
        GF_HIP(hipMemsetAsync(t.A, 0, bs * 8, s));
        GF_HIP(hipMemsetAsync((double *)t.C + (blocks - bs), 0, bs * 8, s));
        GF_HIP(hipMemcpyAsync(d_r, b, n * 8, hipMemcpyHostToDevice, s));
        int l = gf_btd_factor_solve(t, s, d_flag, d_r, d_x, (int)n, d_w);
        if (l < 0) return l;
        GF_HIP(hipMemcpyAsync(x, d_x, n * 8, hipMemcpyDeviceToHost, s));
        GF_HIP(hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        return GF_OK;
    };
    const int st = run();
    if (s) (void)hipStreamSynchronize(s);
    void *bufs[] = {d_r, d_x, d_w, d_flag};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    gf_btd_free(t);
    if (s) (void)hipStreamDestroy(s);
    if (st) return st;
    if (flag) *info_fail = 1;
    return GF_OK;
}

extern "C" int gf_shell_eval(int count, const double *x, const double *prop, const double *d, const double *gp,
                             double *K_out, double *f_out)
{
    if (count < 0 || (count > 0 && (!x || !prop || !d || !gp || !K_out || !f_out))) return GF_E_ARG;
    return gf_shell_eval_device(count, x, prop, d, gp, K_out, f_out);
}

extern "C" int gf_beam_eval(int count, const double *x, const double *cs, const double *prop, const double *d,
                            const double *gp, double *K_out, double *f_out)
{
    if (count < 0 || (count > 0 && (!x || !cs || !prop || !d || !gp || !K_out || !f_out))) return GF_E_ARG;
    return gf_beam_eval_device(count, x, cs, prop, d, gp, K_out, f_out);
}
