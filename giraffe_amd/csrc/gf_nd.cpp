// Nested-dissection plan of the multifrontal tangent solve (host only).
//
// The reference hands the tangent to MKL PARDISO, which orders it by nested
// dissection (iparm[1] = 3, src/SparseMatrix.cpp:115-117) and factors it
// supernode by supernode.  The plan here does the same on the node graph of the
// mesh (nodes adjacent when an element holds both, every free DOF of a node in
// the node's supernode):
//
//   ordering   recursive coordinate bisection: a node set splits across the axis
//              of its largest extent at the position, among those between two
//              distinct coordinates with 25-75 % of the set's DOFs on either
//              side, whose one-sided vertex separator (the nodes of one side
//              adjacent to the other side) has the fewest DOFs; the parts recurse
//              until they hold at most LEAF_DOFS DOFs.  Separators above KMAX
//              DOFs become chains of supernodes.
//   structure  the update set of a supernode: the ancestor nodes its own nodes
//              are adjacent to, and those its children pass on (the classical
//              supernodal symbolic factorisation; George & Liu 1981, Liu 1992).
//   schedule   fronts by level (height in the tree); per level one assembly
//              launch, one launch per 64-column block step, and the work items
//              (tiles) of every launch, so the device needs no index arithmetic
//              beyond a table read.
#include "gf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int TB = 64;          // tile size of the fronts
constexpr int LEAF_DOFS = 256;  // largest leaf part
constexpr int KMAX = 1024;      // largest supernode (the solve keeps its own rows in LDS)

struct Nd {
    const double *xyz = nullptr;
    std::vector<std::vector<int>> adj;  // free nodes only
    std::vector<int> nd;                // free DOFs of each node
    std::vector<int> stamp, pos;
    int round = 0;
    std::vector<std::vector<int>> own;  // supernodes in creation order (children first)
    std::vector<std::vector<int>> kids;

    long long dofs(const std::vector<int> &s) const
    {
        long long t = 0;
        for (int v : s) t += nd[v];
        return t;
    }

    int make(std::vector<int> nodes, std::vector<int> children)
    {
        // chain of supernodes of at most KMAX DOFs each (a separator larger than
        // the solve's own-row capacity)
        size_t a = 0;
        int last = -1;
        while (a < nodes.size()) {
            std::vector<int> part;
            long long t = 0;
            while (a < nodes.size() && (part.empty() || t + nd[nodes[a]] <= KMAX)) {
                t += nd[nodes[a]];
                part.push_back(nodes[a++]);
            }
            own.push_back(std::move(part));
            kids.push_back(last < 0 ? children : std::vector<int>{last});
            last = (int)own.size() - 1;
        }
        return last;
    }

    std::vector<int> rec(std::vector<int> S)
    {
        if (S.empty()) return {};
        const long long D = dofs(S);
        const int N = (int)S.size();
        if (D <= LEAF_DOFS || N < 4) return {make(std::move(S), {})};
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int v : S)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], xyz[3 * (size_t)v + a]);
                hi[a] = std::max(hi[a], xyz[3 * (size_t)v + a]);
            }
        int ax = 0;
        for (int a = 1; a < 3; a++)
            if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        if (!(hi[ax] > lo[ax])) return {make(std::move(S), {})};
        std::sort(S.begin(), S.end(), [&](int a, int b) {
            const double xa = xyz[3 * (size_t)a + ax], xb = xyz[3 * (size_t)b + ax];
            return xa < xb || (xa == xb && a < b);
        });
        round++;
        for (int i = 0; i < N; i++) {
            stamp[S[i]] = round;
            pos[S[i]] = i;
        }
        // one-sided separator sizes of every split position s (A = [0, s), B = [s, N))
        std::vector<long long> dA(N + 2, 0), dB(N + 2, 0), cum(N + 1, 0);
        std::vector<int> mx(N), mn(N);
        for (int i = 0; i < N; i++) {
            const int v = S[i];
            int hiw = -1, low = N;
            for (int w : adj[v])
                if (stamp[w] == round) {
                    hiw = std::max(hiw, pos[w]);
                    low = std::min(low, pos[w]);
                }
            mx[i] = hiw;
            mn[i] = low;
            if (hiw > i) { dA[i + 1] += nd[v]; dA[hiw + 1] -= nd[v]; }
            if (low < i) { dB[low + 1] += nd[v]; dB[i + 1] -= nd[v]; }
            cum[i + 1] = cum[i] + nd[v];
        }
        long long sa = 0, sb = 0;
        int best = -1;
        bool best_a = true;
        double best_cost = 1e300;
        for (int pass = 0; pass < 2 && best < 0; pass++) {
            const double f0 = pass == 0 ? 0.25 : 0.05;
            sa = sb = 0;
            for (int s = 1; s < N; s++) {
                sa += dA[s];
                sb += dB[s];
                const double xa = xyz[3 * (size_t)S[s - 1] + ax], xb = xyz[3 * (size_t)S[s] + ax];
                if (!(xa < xb)) continue;
                const double frac = (double)cum[s] / (double)D;
                if (frac < f0 || frac > 1.0 - f0) continue;
                const long long sep = std::min(sa, sb);
                const double cost = (double)sep * (1.0 + std::fabs(frac - 0.5));
                if (cost < best_cost) {
                    best_cost = cost;
                    best = s;
                    best_a = sa <= sb;
                }
            }
        }
        if (best < 0) return {make(std::move(S), {})};
        std::vector<int> A, B, sep;
        for (int i = 0; i < N; i++) {
            const bool inA = i < best;
            const bool on_sep = best_a ? (inA && mx[i] >= best) : (!inA && mn[i] < best);
            (on_sep ? sep : inA ? A : B).push_back(S[i]);
        }
        if (dofs(sep) * 10 >= D * 9) return {make(std::move(S), {})};
        std::vector<int> roots = rec(std::move(A));
        std::vector<int> rb = rec(std::move(B));
        roots.insert(roots.end(), rb.begin(), rb.end());
        if (sep.empty()) return roots;
        return {make(std::move(sep), std::move(roots))};
    }
};

inline int pad64(int x) { return (x + TB - 1) / TB * TB; }

}  // namespace

int gf_nd_plan(gf_fronts &fp, const gf_model_desc *d, const std::vector<int> &dof, int nn, int npad)
{
    fp = gf_fronts();
    Nd g;
    g.xyz = d->node_xyz;
    g.adj.assign(nn, {});
    g.nd.assign(nn, 0);
    g.stamp.assign(nn, 0);
    g.pos.assign(nn, 0);
    for (int v = 0; v < nn; v++)
        for (int c = 0; c < 6; c++) g.nd[v] += dof[6 * (size_t)v + c] >= 0;
    auto link = [&](const int *cn, int nper) {
        for (int a = 0; a < nper; a++)
            for (int b = 0; b < nper; b++)
                if (a != b && g.nd[cn[a]] && g.nd[cn[b]]) g.adj[cn[a]].push_back(cn[b]);
    };
    for (int e = 0; e < d->n_beams; e++) link(d->beam_conn + 3 * (size_t)e, 3);
    for (int e = 0; e < d->n_shells; e++) link(d->shell_conn + 6 * (size_t)e, 6);
    for (auto &v : g.adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<int> all;
    for (int v = 0; v < nn; v++)
        if (g.nd[v]) all.push_back(v);
    if (all.empty()) { gf_set_error("nested dissection: no free DOFs"); return GF_E_ARG; }
    g.rec(all);
    const int S = (int)g.own.size();
    // supernode of each node; own nodes in band order
    std::vector<int> sn(nn, -1);
    auto first_dof = [&](int v) {
        for (int c = 0; c < 6; c++)
            if (dof[6 * (size_t)v + c] >= 0) return dof[6 * (size_t)v + c];
        return -1;
    };
    for (int s = 0; s < S; s++) {
        std::sort(g.own[s].begin(), g.own[s].end(), [&](int a, int b) { return first_dof(a) < first_dof(b); });
        for (int v : g.own[s]) sn[v] = s;
    }
    // update sets (ancestor nodes), children before parents
    std::vector<std::vector<int>> upd(S);
    std::vector<int> mark(nn, -1);
    for (int s = 0; s < S; s++) {
        std::vector<int> &u = upd[s];
        auto take = [&](int w) {
            if (sn[w] > s && mark[w] != s) {
                mark[w] = s;
                u.push_back(w);
            }
        };
        for (int v : g.own[s])
            for (int w : g.adj[v]) take(w);
        for (int c : g.kids[s])
            for (int w : upd[c]) take(w);
        std::sort(u.begin(), u.end(), [&](int a, int b) {
            return sn[a] < sn[b] || (sn[a] == sn[b] && first_dof(a) < first_dof(b));
        });
    }
    // levels (height) and front numbering level by level
    std::vector<int> height(S, 0);
    int nlev = 0;
    for (int s = 0; s < S; s++) {
        for (int c : g.kids[s]) height[s] = std::max(height[s], height[c] + 1);
        nlev = std::max(nlev, height[s] + 1);
    }
    std::vector<int> order(S), fid(S);
    for (int s = 0; s < S; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return height[a] < height[b]; });
    for (int f = 0; f < S; f++) fid[order[f]] = f;
    fp.n_front = S;
    fp.n_level = nlev;
    fp.lev_off.assign(nlev + 1, 0);
    for (int s = 0; s < S; s++) fp.lev_off[height[s] + 1]++;
    for (int l = 0; l < nlev; l++) fp.lev_off[l + 1] += fp.lev_off[l];
    fp.fr.resize(S);
    fp.lev_kb.assign(nlev, 0);
    // fronts: row maps, storage offsets
    std::vector<int> kvec(S), uvec(S);
    for (int f = 0; f < S; f++) {
        const int s = order[f];
        gf_front_dev &F = fp.fr[f];
        std::vector<int> rows;
        for (int v : g.own[s])
            for (int c = 0; c < 6; c++)
                if (dof[6 * (size_t)v + c] >= 0) rows.push_back(dof[6 * (size_t)v + c]);
        const int k = (int)rows.size(), kp = pad64(k);
        rows.resize(kp, -1);
        for (int w : upd[s])
            for (int c = 0; c < 6; c++)
                if (dof[6 * (size_t)w + c] >= 0) rows.push_back(dof[6 * (size_t)w + c]);
        const int u = (int)rows.size() - kp, up = pad64(u);
        rows.resize(kp + up, -1);
        kvec[f] = k;
        uvec[f] = u;
        F.k = k;
        F.kb = kp / TB;
        F.mp = kp + up;
        F.mb = F.mp / TB;
        F.rows = (int)fp.rows.size();
        fp.rows.insert(fp.rows.end(), rows.begin(), rows.end());
        F.F = fp.f_elems;
        fp.f_elems += (long long)F.mp * F.mp;
        F.E = fp.e_elems;
        fp.e_elems += (long long)F.kb * TB * TB;
        F.V = fp.v_elems;
        fp.v_elems += F.mp;
        F.pad = 0;
        fp.max_k = std::max(fp.max_k, kp);
        fp.lev_kb[height[s]] = std::max(fp.lev_kb[height[s]], F.kb);
        if (!g.kids[s].empty()) fp.n_sep_dofs += k;
    }
    // children and their maps (parent front position -> child update row)
    std::vector<int> ppos(npad, -1);
    for (int f = 0; f < S; f++) {
        const int s = order[f];
        gf_front_dev &F = fp.fr[f];
        const int *pr = fp.rows.data() + F.rows;
        for (int i = 0; i < F.mp; i++)
            if (pr[i] >= 0) ppos[pr[i]] = i;
        F.ch0 = (int)fp.child.size();
        for (int cs : g.kids[s]) {
            const int c = fid[cs];
            const gf_front_dev &C = fp.fr[c];
            fp.child.push_back(c);
            fp.child_inv.push_back((long long)fp.inv.size());
            const size_t base = fp.inv.size();
            fp.inv.resize(base + F.mp, -1);
            const int *cr = fp.rows.data() + C.rows + C.kb * TB;
            for (int a = 0; a < uvec[c]; a++) {
                const int i = ppos[cr[a]];
                if (i < 0) { gf_set_error("nested dissection: update row outside the parent front"); return GF_E_STATE; }
                fp.inv[base + i] = a;
            }
        }
        F.ch1 = (int)fp.child.size();
        for (int i = 0; i < F.mp; i++)
            if (pr[i] >= 0) ppos[pr[i]] = -1;
    }
    // factor launches: per level, launch s = step p = 0 .. lev_kb - 1 (the first
    // diagonal block is inverted by the assembly launch, and the last own block
    // row is left unconverted: the backward solve applies its inverse); items go
    // to XCD (front + tile row) mod 8 (workgroup b runs on XCD b mod 8), the
    // diagonal tiles first
    auto emit = [&](std::vector<std::vector<int>> &q) {
        size_t len = 0;
        for (auto &v : q) len = std::max(len, v.size() / 4);
        const int first = (int)fp.items.size() / 4;
        for (size_t r = 0; r < len; r++)
            for (int x = 0; x < 8; x++) {
                if (r * 4 < q[x].size()) fp.items.insert(fp.items.end(), q[x].begin() + 4 * r, q[x].begin() + 4 * r + 4);
                else fp.items.insert(fp.items.end(), {-1, 0, 0, 0});
            }
        return first;
    };
    fp.lev_step.assign(nlev + 1, 0);
    for (int l = 0; l < nlev; l++) {
        fp.lev_step[l] = (int)fp.step_first.size();
        for (int p = 0; p < fp.lev_kb[l]; p++) {
            std::vector<std::vector<int>> diag(8), rest(8);
            double fl = 0.0;
            for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++) {
                const gf_front_dev &F = fp.fr[f];
                const int x0 = f;
                if (p < F.kb) {
                    for (int i = p + 1; i < F.mb; i++)
                        for (int j = p + 1; j < F.mb; j++) {
                            const bool dg = i == p + 1 && j == p + 1 && p + 1 < F.kb;
                            auto &q = dg ? diag[(x0 + i) & 7] : rest[(x0 + i) & 7];
                            q.insert(q.end(), {f, i, j, dg ? 0 : 1});
                        }
                    const int m = kvec[f] + uvec[f];
                    for (int c = TB * p; c < std::min(TB * (p + 1), kvec[f]); c++) fl += 2.0 * (double)(m - c - 1) * (m - c - 1);
                }
                if (p >= 1 && p <= F.kb - 1)
                    for (int j = p; j < F.mb; j++) rest[(x0 + p - 1) & 7].insert(rest[(x0 + p - 1) & 7].end(), {f, p - 1, j, 2});
            }
            for (int x = 0; x < 8; x++) diag[x].insert(diag[x].end(), rest[x].begin(), rest[x].end());
            const int first = emit(diag);
            fp.step_first.push_back(first);
            fp.step_count.push_back((int)fp.items.size() / 4 - first);
            fp.step_flops.push_back(fl);
            fp.flops += fl;
        }
    }
    fp.lev_step[nlev] = (int)fp.step_first.size();
    // assembly tiles and solve row tiles per level
    for (int l = 0; l < nlev; l++) {
        // assembly tiles, the first diagonal block of every front first (kind 1:
        // assembled and inverted)
        const int a0 = (int)fp.items.size() / 4;
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++) fp.items.insert(fp.items.end(), {f, 0, 0, 1});
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++)
            for (int i = 0; i < fp.fr[f].mb; i++)
                for (int j = 0; j < fp.fr[f].mb; j++)
                    if (i || j) fp.items.insert(fp.items.end(), {f, i, j, 0});
        fp.asm_first.push_back(a0);
        fp.asm_count.push_back((int)fp.items.size() / 4 - a0);
        const int u0 = (int)fp.items.size() / 4;
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++)
            for (int t = fp.fr[f].kb; t < fp.fr[f].mb; t++) fp.items.insert(fp.items.end(), {f, t, 0, 0});
        fp.fupd_first.push_back(u0);
        fp.fupd_count.push_back((int)fp.items.size() / 4 - u0);
        const int b0 = (int)fp.items.size() / 4;
        for (int f = fp.lev_off[l]; f < fp.lev_off[l + 1]; f++)
            if (fp.fr[f].mb > fp.fr[f].kb)
                for (int p = 0; p < fp.fr[f].kb; p++) fp.items.insert(fp.items.end(), {f, p, 0, 0});
        fp.bupd_first.push_back(b0);
        fp.bupd_count.push_back((int)fp.items.size() / 4 - b0);
    }
    return GF_OK;
}

int gf_front_factor_launches(const gf_fronts &fp)
{
    int n = 0;
    for (int l = 0; l < fp.n_level; l++) n += 1 + fp.lev_step[l + 1] - fp.lev_step[l];
    return n;
}

int gf_front_solve_launches(const gf_fronts &fp)
{
    int n = 0;
    for (int l = 0; l < fp.n_level; l++) n += 2 + (fp.fupd_count[l] > 0) + (fp.bupd_count[l] > 0);
    return n;
}

// Original tangent entries of every assembly tile: for each front's own DOF a and
// each structural neighbour b of a (the tangent's structure, ptr / col, row by row)
// the entry (a, b) and, when b is an update DOF of the front, (b, a); entries whose
// other DOF is in no row of the front belong to a descendant's front.  Per
// assembly item: the first entry of its tile and their count.
int gf_plan_workers(long long work)
{
    if (work < (1LL << 20)) return 1;
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc ? hc : 1u, 16u));
}

int gf_nd_orig_lists(gf_fronts &fp, const std::vector<int> &ptr, const std::vector<int> &col, int npad)
{
    const int S = fp.n_front;
    fp.orig.clear();
    fp.asm_orig.assign(fp.items.size() / 2, 0);  // 2 ints per item
    // largest fronts first (dynamic hand-out); every front's entries land at
    // offsets fixed by the count pass, so the lists do not depend on the threads
    std::vector<int> by_size(S);
    for (int f = 0; f < S; f++) by_size[f] = f;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) {
        return (long long)fp.fr[a].k * fp.fr[a].mp > (long long)fp.fr[b].k * fp.fr[b].mp;
    });
    const int W = gf_plan_workers((long long)col.size());
    std::vector<std::vector<int>> ppos(W);
    std::vector<std::vector<long long>> toff(S);
    // the (front position, front position, band row, band column) of every
    // tangent entry whose earlier-eliminated index is one of the front's own rows
    auto each = [&](int w, int f, auto emit) {
        const gf_front_dev &F = fp.fr[f];
        const int *rows = fp.rows.data() + F.rows;
        std::vector<int> &pp = ppos[w];
        if (pp.empty()) pp.assign(npad, -1);
        for (int i = 0; i < F.mp; i++)
            if (rows[i] >= 0) pp[rows[i]] = i;
        for (int i = 0; i < F.k; i++) {
            const int oa = rows[i];
            for (int q = ptr[oa]; q < ptr[oa + 1]; q++) {
                const int ob = col[q], j = pp[ob];
                if (j < 0) continue;
                emit(i, j, oa, ob);
                if (j >= F.k) emit(j, i, ob, oa);
            }
        }
        for (int i = 0; i < F.mp; i++)
            if (rows[i] >= 0) pp[rows[i]] = -1;
    };
    gf_parallel_for(S, W, [&](int w, int x) {
        const int f = by_size[x];
        const int mb = fp.fr[f].mb;
        std::vector<long long> &to = toff[f];
        to.assign((size_t)mb * mb + 1, 0);
        each(w, f, [&](int i, int j, int, int) { to[(i / TB) * mb + j / TB + 1]++; });
        for (int t = 0; t < mb * mb; t++) to[t + 1] += to[t];
    });
    std::vector<long long> front_base(S + 1, 0);
    for (int f = 0; f < S; f++) front_base[f + 1] = front_base[f] + toff[f].back();
    // every structural entry is counted once, by the front that eliminates the
    // earlier of its row and column
    if (front_base[S] != (long long)col.size()) { gf_set_error("multifrontal plan: tangent entries outside the fronts"); return GF_E_STATE; }
    if (front_base[S] > (long long)0x7fffffff) { gf_set_error("multifrontal plan: too many tangent entries"); return GF_E_UNSUPPORTED; }
    fp.orig.resize(3 * (size_t)front_base[S]);
    gf_parallel_for(S, W, [&](int w, int x) {
        const int f = by_size[x];
        const int mb = fp.fr[f].mb;
        std::vector<long long> fill(toff[f].begin(), toff[f].end() - 1);
        int *o = fp.orig.data() + 3 * (size_t)front_base[f];
        each(w, f, [&](int i, int j, int oa, int ob) {
            int *e = o + 3 * (size_t)fill[(i / TB) * mb + j / TB]++;
            e[0] = (i % TB) * TB + j % TB;
            e[1] = oa;
            e[2] = ob;
        });
    });
    for (int l = 0; l < fp.n_level; l++)
        for (int it = fp.asm_first[l]; it < fp.asm_first[l] + fp.asm_count[l]; it++) {
            const int f = fp.items[4 * (size_t)it], ti = fp.items[4 * (size_t)it + 1], tj = fp.items[4 * (size_t)it + 2];
            const int t = ti * fp.fr[f].mb + tj;
            fp.asm_orig[2 * (size_t)it] = (int)(front_base[f] + toff[f][t]);
            fp.asm_orig[2 * (size_t)it + 1] = (int)(toff[f][t + 1] - toff[f][t]);
        }
    return GF_OK;
}

