/*
 * ORACLE (test infrastructure only) -- model set-up, global assembly, Newton
 * update / commit / rollback and the Static load-step loop, restated in C.
 *
 * Follows:
 *   Solution::DOFsActive / SetGlobalDOFs  src/Solution.cpp:121-224, 40-118
 *   NodalConstraint::Mount                src/NodalConstraint.cpp:152-173
 *   NodalLoad::UpdateforSolutionStep      src/NodalLoad.cpp:236-266
 *   NodalLoad::Mount                      src/NodalLoad.cpp:322-401
 *   NodalFollowerLoad::Mount              src/NodalFollowerLoad.cpp:243-326
 *   Table::GetValueAt                     src/Table.cpp:73-101
 *   BoolTable::GetLinearFactorAtCurrentTime src/BoolTable.cpp:84-109 (first solution on)
 *   Beam_1::MountElementLoads / MountFieldLoading (gravity) src/Beam_1.cpp:838-878
 *   Solution::MountLocal / MountElementLoads / MountLoads / MountGlobal
 *                                         src/Solution.cpp:227-387
 *   Beam_1::MountGlobal                   src/Beam_1.cpp:1429-1492
 *   sparsesystem                          src/SparseMatrix.cpp:74-166
 *   Solution::UpdateDisps                 src/Solution.cpp:390-423
 *   Solution::SaveConfiguration, Node::SaveConfiguration
 *                                         src/Solution.cpp:426-529, src/Node.cpp:325-384
 *   Solution::RestoreConfiguration        src/Solution.cpp:532-574
 *   ConvergenceCriteria                   src/ConvergenceCriteria.cpp:15-58, 187-617
 *   Static::Solve                         src/Static.cpp:140-311
 *
 * Linear solver: 0 = the band LU of gfo_band.c; 1 = MKL PARDISO called
 * exactly as sparsesystem does (mtype 11, fresh handle, phase 13 then -1,
 * iparm[1] = 3), loaded with dlopen when libmkl_rt is present.
 *
 * Pinned by tests/test_oracle_static.py: the config-1 roll-up closed form,
 * the small-load Timoshenko cantilever, band-vs-PARDISO agreement, hand-built
 * step-control sequences (bisection, abort, step growth) and the criteria
 * history average.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "gfo_internal.h"
#include "gfo_shell.h"

typedef struct gfo_model {
    int nn;
    double *xyz_ref;          /* [nn][3] */
    double *copy;             /* [nn][6] converged X Y Z and Rodrigues alpha */
    double *euler;            /* [nn][3] */
    double *d;                /* [nn][6] incremental displacements           */
    double *xinc;             /* [nn][6] last increment                      */
    int *GL;                  /* [nn][6] >0 free (1-based), <0 fixed, 0 inactive */
    unsigned char *active, *fix;
    int n_free, n_fixed;
    /* beams */
    int nb;
    int *bconn;
    gfo_beam_const *bc;
    double *bgp;              /* [nb][2][GFO_BEAM_GPS] */
    double *bK, *bP;          /* [nb][324], [nb][18] */
    double *bres;             /* [nb][12] last Mount's sigma_r, epsilon_r means */
    /* shells */
    int ns;
    int *sconn;
    gfo_shell_const *sc;
    double *sgp;              /* [ns][3][GFO_SHELL_GPS] */
    double *sK, *sP;          /* [ns][729], [ns][27] */
    double *sres;             /* [ns][24] last Mount's resultant and strain means */
    /* loads */
    int nl;
    int *lkind, *lset_off, *lset, *lflags, *ltab_off;
    double *lcs, *tab_t, *tab_v;
    double *mult;             /* [nl][6] */
    double grav[3];
    int grav_on;
    int grav_mode; /* 0 off, 1 ramp up, 2 on, 3 ramp down (BoolTable across solutions) */
    double *bbuoy;            /* [nb] Pipe_1 buoyant mass per length (with OceanData) */
    int ocean_on;
    double surf[3];
    /* global system */
    int kl, ku, ldab;
    double *ab;
    int *ipiv;
    double *P;                /* residual / solution, n_free */
    int factored;
    int solver;
    /* CSR for PARDISO */
    int *ia, *ja;
    double *a;
    long long nnz;
    /* time */
    double t_last, dt, sol_start, sol_end;
    int threads;
} gfo_model;

/* ---------------------------------------------------------------- tables */
static double table_value(const gfo_model *m, int load, double t, int col)
{
    int r0 = m->ltab_off[load], r1 = m->ltab_off[load + 1];
    int nt = r1 - r0;
    if (nt == 0) return 0.0;
    const double *tt = m->tab_t + r0;
    const double *vv = m->tab_v + 6 * (size_t)r0;
    if (t >= tt[nt - 1]) return vv[6 * (nt - 1) + col];
    if (t <= tt[0]) return vv[col];
    int idx = 0;
    while (t >= tt[idx]) idx++;
    double fac = (t - tt[idx - 1]) / (tt[idx] - tt[idx - 1]);
    return vv[6 * (idx - 1) + col] + fac * (vv[6 * idx + col] - vv[6 * (idx - 1) + col]);
}

static void *dupmem(const void *src, size_t bytes)
{
    void *p = malloc(bytes ? bytes : 1);
    if (src && bytes) memcpy(p, src, bytes);
    else memset(p, 0, bytes ? bytes : 1);
    return p;
}

/* ---------------------------------------------------------------- create */
static int csr_build(gfo_model *m);

gfo_model *gfo_create(const gf_model_desc *desc, int solver)
{
    gfo_model *m = (gfo_model *)calloc(1, sizeof(gfo_model));
    int nn = m->nn = desc->n_nodes;
    m->xyz_ref = (double *)dupmem(desc->node_xyz, sizeof(double) * 3 * nn);
    m->copy = (double *)calloc(6 * (size_t)nn, sizeof(double));
    for (int i = 0; i < nn; i++)
        for (int k = 0; k < 3; k++) m->copy[6 * i + k] = m->xyz_ref[3 * i + k];
    m->euler = (double *)calloc(3 * (size_t)nn, sizeof(double));
    m->d = (double *)calloc(6 * (size_t)nn, sizeof(double));
    m->xinc = (double *)calloc(6 * (size_t)nn, sizeof(double));
    m->GL = (int *)calloc(6 * (size_t)nn, sizeof(int));
    m->active = (unsigned char *)calloc(6 * (size_t)nn, 1);
    m->fix = (unsigned char *)dupmem(desc->node_fix, 6 * (size_t)nn);
    if (!desc->node_fix) memset(m->fix, 0, 6 * (size_t)nn);
    /* beams */
    m->nb = desc->n_beams;
    m->bconn = (int *)dupmem(desc->beam_conn, sizeof(int) * 3 * m->nb);
    m->bc = (gfo_beam_const *)calloc(m->nb ? m->nb : 1, sizeof(gfo_beam_const));
    m->bgp = (double *)calloc((size_t)m->nb * 2 * GFO_BEAM_GPS + 1, sizeof(double));
    m->bK = (double *)calloc((size_t)m->nb * 324 + 1, sizeof(double));
    m->bP = (double *)calloc((size_t)m->nb * 18 + 1, sizeof(double));
    m->bres = (double *)calloc((size_t)m->nb * 12 + 1, sizeof(double));
    for (int e = 0; e < m->nb; e++) {
        double x[9];
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 3; k++) x[3 * a + k] = m->xyz_ref[3 * m->bconn[3 * e + a] + k];
        gfo_beam_precalc(x, desc->beam_cs + 9 * e, desc->beam_prop + GF_BEAM_NPROP * e,
                         desc->beam_kind ? desc->beam_kind[e] : GF_BEAM_KIND_BEAM1, &m->bc[e],
                         m->bgp + (size_t)e * 2 * GFO_BEAM_GPS);
    }
    /* shells */
    m->ns = desc->n_shells;
    m->sconn = (int *)dupmem(desc->shell_conn, sizeof(int) * 6 * m->ns);
    m->sc = gfo_shell_alloc(m->ns);
    m->sgp = (double *)calloc((size_t)m->ns * 3 * GFO_SHELL_GPS + 1, sizeof(double));
    m->sK = (double *)calloc((size_t)m->ns * 729 + 1, sizeof(double));
    m->sP = (double *)calloc((size_t)m->ns * 27 + 1, sizeof(double));
    m->sres = (double *)calloc((size_t)m->ns * 24 + 1, sizeof(double));
    for (int e = 0; e < m->ns; e++) {
        double x[18];
        for (int a = 0; a < 6; a++)
            for (int k = 0; k < 3; k++) x[3 * a + k] = m->xyz_ref[3 * m->sconn[6 * e + a] + k];
        gfo_shell_precalc(x, desc->shell_prop + GF_SHELL_NPROP * e, gfo_shell_at(m->sc, e),
                          m->sgp + (size_t)e * 3 * GFO_SHELL_GPS);
    }
    /* loads */
    m->nl = desc->n_loads;
    int nl = m->nl;
    m->lkind = (int *)dupmem(desc->load_kind, sizeof(int) * nl);
    m->lset_off = (int *)dupmem(desc->load_set_off, sizeof(int) * (nl + 1));
    int nitems = nl ? desc->load_set_off[nl] : 0;
    m->lset = (int *)dupmem(desc->load_set_items, sizeof(int) * nitems);
    m->lcs = (double *)dupmem(desc->load_cs, sizeof(double) * 9 * nl);
    m->lflags = (int *)dupmem(desc->load_flags, sizeof(int) * nl);
    m->ltab_off = (int *)dupmem(desc->load_tab_off, sizeof(int) * (nl + 1));
    int rows = nl ? desc->load_tab_off[nl] : 0;
    m->tab_t = (double *)dupmem(desc->tab_time, sizeof(double) * rows);
    m->tab_v = (double *)dupmem(desc->tab_val, sizeof(double) * 6 * rows);
    m->mult = (double *)calloc(6 * (size_t)(nl ? nl : 1), sizeof(double));
    for (int k = 0; k < 3; k++) m->grav[k] = desc->gravity[k];
    m->grav_on = desc->gravity_on;
    m->grav_mode = m->grav_on ? 1 : 0;
    m->ocean_on = desc->ocean_on;
    for (int k = 0; k < 3; k++) m->surf[k] = desc->ocean_surface[k];
    m->bbuoy = (double *)calloc((size_t)m->nb + 1, sizeof(double));
    if (desc->ocean_on && desc->beam_buoy) memcpy(m->bbuoy, desc->beam_buoy, sizeof(double) * (size_t)m->nb);
    m->solver = solver;
#ifdef _OPENMP
    m->threads = omp_get_max_threads();
#else
    m->threads = 1;
#endif
    return m;
}

void gfo_destroy(gfo_model *m)
{
    if (!m) return;
    free(m->xyz_ref); free(m->copy); free(m->euler); free(m->d); free(m->xinc);
    free(m->GL); free(m->active); free(m->fix);
    free(m->bconn); free(m->bc); free(m->bgp); free(m->bK); free(m->bP); free(m->bres); free(m->bbuoy);
    free(m->sconn); free(m->sc); free(m->sgp); free(m->sK); free(m->sP); free(m->sres);
    free(m->lkind); free(m->lset_off); free(m->lset); free(m->lcs); free(m->lflags);
    free(m->ltab_off); free(m->tab_t); free(m->tab_v); free(m->mult);
    free(m->ab); free(m->ipiv); free(m->P);
    free(m->ia); free(m->ja); free(m->a);
    free(m);
}

/* DOFsActive + constraints + SetGlobalDOFs + load multipliers + band sizes. */
int gfo_begin_solution(gfo_model *m, double start_time, double end_time)
{
    int nn = m->nn;
    memset(m->active, 0, 6 * (size_t)nn);
    for (int e = 0; e < m->nb; e++)
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 6; k++) m->active[6 * m->bconn[3 * e + a] + k] = 1;
    for (int e = 0; e < m->ns; e++)
        for (int a = 0; a < 6; a++)
            for (int k = 0; k < 6; k++)
                if (k < 3 || a > 2) m->active[6 * m->sconn[6 * e + a] + k] = 1;
    int nf = 0, nx = 0;
    for (int i = 0; i < nn; i++)
        for (int k = 0; k < 6; k++) {
            int c = m->fix[6 * i + k] ? 1 : 0;
            m->GL[6 * i + k] = 0;
            if (c == 0 && m->active[6 * i + k]) m->GL[6 * i + k] = ++nf;
            if (c == 1 && m->active[6 * i + k]) m->GL[6 * i + k] = -(++nx);
        }
    m->n_free = nf;
    m->n_fixed = nx;
    /* NodalLoad / NodalFollowerLoad::UpdateforSolutionStep */
    for (int l = 0; l < m->nl; l++) {
        if (m->lkind[l] == GF_LOAD_SHELL_PRESSURE) continue;
        int cf[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
        for (int s = m->lset_off[l]; s < m->lset_off[l + 1]; s++) {
            int node = m->lset[s];
            for (int j = 0; j < 3; j++) {
                if (m->active[6 * node + j]) cf[j]++;
                if (m->active[6 * node + j + 3]) cm[j]++;
            }
        }
        for (int j = 0; j < 3; j++) {
            m->mult[6 * l + j] = 1.0 / cf[j];
            m->mult[6 * l + 3 + j] = 1.0 / cm[j];
        }
    }
    /* half-bandwidth of K_AA */
    int b = 0;
    for (int e = 0; e < m->nb; e++) {
        int lo = 1 << 30, hi = -1;
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 6; k++) {
                int g = m->GL[6 * m->bconn[3 * e + a] + k];
                if (g > 0) { if (g < lo) lo = g; if (g > hi) hi = g; }
            }
        if (hi >= 0 && hi - lo > b) b = hi - lo;
    }
    for (int e = 0; e < m->ns; e++) {
        int lo = 1 << 30, hi = -1;
        for (int a = 0; a < 6; a++)
            for (int k = 0; k < 6; k++) {
                int g = m->GL[6 * m->sconn[6 * e + a] + k];
                if (g > 0) { if (g < lo) lo = g; if (g > hi) hi = g; }
            }
        if (hi >= 0 && hi - lo > b) b = hi - lo;
    }
    m->kl = m->ku = b;
    m->ldab = 2 * b + b + 1;
    free(m->ab); free(m->ipiv); free(m->P);
    if (m->solver == 0) {
        m->ab = (double *)calloc((size_t)m->ldab * (size_t)(nf ? nf : 1), sizeof(double));
        if (!m->ab) return GF_E_NOMEM;
    } else {
        m->ab = NULL;
    }
    m->ipiv = (int *)calloc(nf ? nf : 1, sizeof(int));
    m->P = (double *)calloc(nf ? nf : 1, sizeof(double));
    m->sol_start = start_time;
    m->sol_end = end_time;
    if (m->solver == 1) return csr_build(m);
    return GF_OK;
}

int gfo_n_free(const gfo_model *m) { return m->n_free; }
int gfo_bandwidth(const gfo_model *m) { return m->kl; }
void gfo_dof_map(const gfo_model *m, int *out) { memcpy(out, m->GL, sizeof(int) * 6 * (size_t)m->nn); }

/* ------------------------------------------------------------ step control */
void gfo_begin_step(gfo_model *m, double t_last, double dt)
{
    m->t_last = t_last;
    m->dt = dt;
    memset(m->d, 0, sizeof(double) * 6 * (size_t)m->nn); /* Solution::Zeros */
}

static double gravity_factor(const gfo_model *m)
{
    /* BoolTable::GetLinearFactorAtCurrentTime src/BoolTable.cpp:84-109: on in this solution and
       off in the previous one ramps up across the solution, on in both keeps 1, off now and on
       before ramps down, off in both is 0. */
    const double r = (m->t_last + m->dt - m->sol_start) / (m->sol_end - m->sol_start);
    switch (m->grav_mode) {
    case 1: return r;
    case 2: return 1.0;
    case 3: return 1.0 - r;
    default: return 0.0;
    }
}

void gfo_set_gravity_mode(gfo_model *m, int mode) { m->grav_mode = mode; }

/* ------------------------------------------------------------ global adds */
static inline void kadd(gfo_model *m, int gi, int gj, double v);

static int csr_find(const gfo_model *m, int r, int c)
{
    int lo = m->ia[r], hi = m->ia[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (m->ja[mid] == c) return mid;
        if (m->ja[mid] < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

static inline void kadd(gfo_model *m, int gi, int gj, double v)
{
    /* gi, gj: 0-based free indices */
    if (m->solver == 0) {
        int kv = m->kl + m->ku;
        m->ab[(size_t)(kv + gi - gj) + (size_t)gj * m->ldab] += v;
    } else {
        int p = csr_find(m, gi, gj);
        if (p >= 0) m->a[p] += v;
    }
}

static void gather_d(const gfo_model *m, const int *conn, int nnodes, double *d)
{
    for (int a = 0; a < nnodes; a++)
        for (int k = 0; k < 6; k++) d[6 * a + k] = m->d[6 * conn[a] + k];
}

/* NodalLoad::Mount (pseudo-moment) and NodalFollowerLoad::Mount. */
static void mount_nodal_loads(gfo_model *m)
{
    double t = m->t_last + m->dt;
    for (int l = 0; l < m->nl; l++) {
        int kind = m->lkind[l];
        if (kind == GF_LOAD_SHELL_PRESSURE) continue;
        const double *CS = m->lcs + 9 * l;
        for (int s = m->lset_off[l]; s < m->lset_off[l + 1]; s++) {
            int node = m->lset[s];
            double f[3], mo[3];
            for (int i = 0; i < 3; i++) {
                f[i] = m->mult[6 * l + i] * table_value(m, l, t, i);
                mo[i] = m->mult[6 * l + 3 + i] * table_value(m, l, t, i + 3);
            }
            const int *GL = m->GL + 6 * node;
            const unsigned char *act = m->active + 6 * node;
            double *alpha = m->d + 6 * node + 3;
            if (kind == GF_LOAD_NODAL) {
                m3_trvec(CS, f, f);
                m3_trvec(CS, mo, mo);
                double Xi[9];
                gfo_rodrigues(alpha, NULL, Xi, NULL);
                m3_trvec(Xi, mo, mo);
                double S[9];
                gfo_V(alpha, mo, S);
                for (int i = 0; i < 3; i++)
                    if (act[i] && GL[i] > 0) m->P[GL[i] - 1] += -1.0 * f[i];
                for (int i = 0; i < 3; i++) {
                    int gi = GL[3 + i];
                    if (act[3 + i] && gi > 0) m->P[gi - 1] += -1.0 * mo[i];
                    for (int j = 0; j < 3; j++) {
                        int gj = GL[3 + j];
                        if (act[3 + j] && gi > 0 && gj > 0) kadd(m, gi - 1, gj - 1, -1.0 * S[3 * i + j]);
                    }
                }
            } else {
                double Qi[9], CSt[9], Q[9], Xi[9], g;
                gfo_rodrigues(m->copy + 6 * node + 3, Qi, NULL, NULL);
                m3_tr(CS, CSt);
                m3_mul(Qi, CSt, Qi);
                gfo_rodrigues(alpha, Q, Xi, &g);
                double QQi[9], fip[3], mip[3], Qim[3];
                m3_mul(Q, Qi, QQi);
                m3_vec(QQi, f, fip);
                m3_vec(Qi, mo, Qim);
                m3_vec(Xi, Qim, mip);
                double Sf[9], K12[9], K22[9], SQm[9], dy[9], t1[9];
                m3_skew(fip, Sf);
                m3_mul(Sf, Xi, K12);
                for (int i = 0; i < 9; i++) K12[i] = -1.0 * K12[i];
                m3_skew(Qim, SQm);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) dy[3 * i + j] = Qim[i] * alpha[j];
                m3_mul(Xi, dy, t1);
                for (int i = 0; i < 9; i++) K22[i] = -0.5 * g * (SQm[i] + t1[i]);
                double q[6], dq[36];
                memset(dq, 0, sizeof(dq));
                for (int i = 0; i < 3; i++) {
                    q[i] = fip[i];
                    q[3 + i] = mip[i];
                    for (int j = 0; j < 3; j++) {
                        dq[6 * i + 3 + j] = K12[3 * i + j];
                        dq[6 * (3 + i) + 3 + j] = K22[3 * i + j];
                    }
                }
                for (int i = 0; i < 6; i++) {
                    int gi = GL[i];
                    if (act[i] && gi > 0) m->P[gi - 1] += -1.0 * q[i];
                    for (int j = 0; j < 6; j++) {
                        int gj = GL[j];
                        if (act[j] && gi > 0 && gj > 0) kadd(m, gi - 1, gj - 1, -1.0 * dq[6 * i + j]);
                    }
                }
            }
        }
    }
}

static void scatter_element(gfo_model *m, const int *conn, int nnodes, const int (*dofs)[2], int ndof,
                            const double *K, const double *P)
{
    int g[27];
    for (int i = 0; i < ndof; i++) g[i] = m->GL[6 * conn[dofs[i][0]] + dofs[i][1]];
    for (int i = 0; i < ndof; i++) {
        if (g[i] > 0) m->P[g[i] - 1] += P[i];
        for (int j = 0; j < ndof; j++)
            if (g[i] > 0 && g[j] > 0) kadd(m, g[i] - 1, g[j] - 1, K[ndof * i + j]);
    }
    (void)nnodes;
}

static int beam_dofs[18][2];
static int shell_dofs[27][2];
static void init_dof_tables(void)
{
    for (int i = 0; i < 18; i++) { beam_dofs[i][0] = i / 6; beam_dofs[i][1] = i % 6; }
    /* Shell_1::MountGlobal src/Shell_1.cpp:1523-1557: 6 nodes x u(3), then mid-side nodes 4..6 x alpha(3) */
    for (int i = 0; i < 18; i++) { shell_dofs[i][0] = i / 3; shell_dofs[i][1] = i % 3; }
    for (int i = 0; i < 9; i++) { shell_dofs[18 + i][0] = 3 + i / 3; shell_dofs[18 + i][1] = 3 + i % 3; }
}

static void scan_residual(const gfo_model *m, double *out)
{
    double mf = 0, mm_ = 0;
    int bad = 0;
    for (int i = 0; i < m->nn; i++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * i + k];
            if (g > 0) {
                double v = fabs(m->P[g - 1]);
                if (!(v == v) || v >= 1e300) bad = 1;
                if (k < 3) { if (v > mf) mf = v; }
                else { if (v > mm_) mm_ = v; }
            }
        }
    out[0] = mf; out[1] = mm_; out[2] = bad ? 1.0 : 0.0;
}

/* Clear + MountLocal + MountElementLoads + MountLoads + MountGlobal + sign flip. */
int gfo_assemble(gfo_model *m, double *res)
{
    static int init = 0;
    if (!init) { init_dof_tables(); init = 1; }
    int nf = m->n_free;
    memset(m->P, 0, sizeof(double) * (size_t)(nf ? nf : 1));
    if (m->solver == 0) memset(m->ab, 0, sizeof(double) * (size_t)m->ldab * (size_t)(nf ? nf : 1));
    else memset(m->a, 0, sizeof(double) * (size_t)m->nnz);
    m->factored = 0;
    double lf = m->grav_on ? gravity_factor(m) : 0.0;
    /* MountLocal + MountElementLoads (element-owned buffers, OpenMP like the reference) */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->nb; e++) {
        double d[18];
        gather_d(m, m->bconn + 3 * e, 3, d);
        double *K = m->bK + 324 * (size_t)e, *P = m->bP + 18 * (size_t)e;
        gfo_beam_mount(&m->bc[e], m->bgp + (size_t)e * 2 * GFO_BEAM_GPS, d, K, P, m->bres + 12 * (size_t)e);
        if (m->grav_on) {
            const gfo_beam_const *c = &m->bc[e];
            double el[18];
            memset(el, 0, sizeof(el));
            for (int g = 0; g < 2; g++) {
                double mult = lf * 1.0 * 1.0 * c->jac * c->rhoA;
                if (m->ocean_on && m->bbuoy[e] != 0.0) {
                    /* Pipe_1::MountFieldLoading src/Pipe_1.cpp:1289-1306: below the surface
                     * (depth along G > 0) the weight per length is Rho - rho_f Aext */
                    double z[3], depth = 0.0, gn = sqrt(v3_dot(m->grav, m->grav));
                    for (int k = 0; k < 3; k++) {
                        z[k] = 0.0;
                        for (int a = 0; a < 3; a++) z[k] += m->copy[6 * (size_t)m->bconn[3 * e + a] + k] * c->N[g][a];
                        for (int a = 0; a < 3; a++) z[k] += d[6 * a + k] * c->N[g][a];
                        depth += (z[k] - m->surf[k]) * m->grav[k];
                    }
                    depth = depth * (1.0 / gn);
                    if (depth > 0) mult = lf * 1.0 * 1.0 * c->jac * (c->rhoA - m->bbuoy[e]);
                }
                for (int a = 0; a < 3; a++)
                    for (int k = 0; k < 3; k++) el[6 * a + k] += mult * c->N[g][a] * m->grav[k];
            }
            for (int i = 0; i < 18; i++) P[i] = P[i] - el[i];
        }
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->ns; e++) {
        double d[36];
        gather_d(m, m->sconn + 6 * e, 6, d);
        gfo_shell_mount(gfo_shell_at(m->sc, e), m->sgp + (size_t)e * 3 * GFO_SHELL_GPS, d,
                        m->sK + 729 * (size_t)e, m->sP + 27 * (size_t)e, m->sres + 24 * (size_t)e);
        if (m->grav_on) gfo_shell_gravity(gfo_shell_at(m->sc, e), lf, m->grav, m->sP + 27 * (size_t)e);
    }
    /* MountLoads: nodal loads write the global system; ShellLoad edits element buffers */
    mount_nodal_loads(m);
    double t = m->t_last + m->dt;
    for (int l = 0; l < m->nl; l++) {
        if (m->lkind[l] != GF_LOAD_SHELL_PRESSURE) continue;
        double p = table_value(m, l, t, 0);
        for (int s = m->lset_off[l]; s < m->lset_off[l + 1]; s++) {
            int e = m->lset[s];
            double d[36];
            gather_d(m, m->sconn + 6 * e, 6, d);
            double xc[18];
            for (int a = 0; a < 6; a++)
                for (int k = 0; k < 3; k++) {
                    int nd = m->sconn[6 * e + a];
                    xc[3 * a + k] = m->copy[6 * nd + k] - m->xyz_ref[3 * nd + k];
                }
            gfo_shell_pressure(gfo_shell_at(m->sc, e), xc, d, p, m->lflags[l],
                               m->sK + 729 * (size_t)e, m->sP + 27 * (size_t)e);
        }
    }
    /* MountGlobal (serial, element order) */
    for (int e = 0; e < m->nb; e++)
        scatter_element(m, m->bconn + 3 * e, 3, (const int(*)[2])beam_dofs, 18, m->bK + 324 * (size_t)e,
                        m->bP + 18 * (size_t)e);
    for (int e = 0; e < m->ns; e++)
        scatter_element(m, m->sconn + 6 * e, 6, (const int(*)[2])shell_dofs, 27, m->sK + 729 * (size_t)e,
                        m->sP + 27 * (size_t)e);
    for (int i = 0; i < nf; i++) m->P[i] = -1.0 * m->P[i];
    if (res) scan_residual(m, res);
    return GF_OK;
}

/* ------------------------------------------------------------ PARDISO */
typedef void (*pardisoinit_t)(void *, const int *, int *);
typedef void (*pardiso_t)(void *, const int *, const int *, const int *, const int *, const int *,
                          const double *, const int *, const int *, int *, const int *, int *,
                          const int *, double *, double *, int *);
typedef void (*mkl_threads_t)(int);
static pardisoinit_t p_init;
static pardiso_t p_solve;
static mkl_threads_t p_threads;
static int mkl_state = 0; /* 0 unknown, 1 ok, -1 absent */

/* SolverOptions::PreCalc src/SolverOptions.cpp:63-70: MKL threads = P-1 */
static int default_mkl_threads(void)
{
#ifdef _OPENMP
    int P = omp_get_max_threads();
#else
    int P = 1;
#endif
    return P != 1 ? P - 1 : P;
}

int gfo_pardiso_available(void)
{
    if (mkl_state) return mkl_state > 0;
    /* one OpenMP runtime in the process: MKL threads through libgomp like the
       element loop (two runtimes spinning on the same cores stall each other) */
    setenv("MKL_THREADING_LAYER", "GNU", 0);
    const char *names[] = {"libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", NULL};
    void *h = NULL;
    for (int i = 0; names[i] && !h; i++) h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
    if (!h) { mkl_state = -1; return 0; }
    p_init = (pardisoinit_t)dlsym(h, "pardisoinit");
    p_solve = (pardiso_t)dlsym(h, "pardiso");
    p_threads = (mkl_threads_t)dlsym(h, "MKL_Set_Num_Threads");
    if (!p_init || !p_solve) { mkl_state = -1; return 0; }
    if (p_threads) p_threads(default_mkl_threads());
    mkl_state = 1;
    return 1;
}

/* PARDISO's thread count: n > 0 sets it, n <= 0 restores the reference's P-1.
   PARDISO's rounding depends on its thread count (and on its dynamic schedule
   when threaded); one thread makes a run reproducible.  Returns 0 when MKL is
   absent or cannot set it. */
int gfo_pardiso_threads(int n)
{
    if (!gfo_pardiso_available() || !p_threads) return 0;
    p_threads(n > 0 ? n : default_mkl_threads());
    return 1;
}

static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

/* CSR pattern of K_AA from the element / load couplings (what setFromTriplets sees). */
static int csr_build(gfo_model *m)
{
    if (!gfo_pardiso_available()) return GF_E_UNSUPPORTED;
    int nf = m->n_free;
    int *cnt = (int *)calloc(nf + 1, sizeof(int));
    /* collect per-row column lists via element node sets */
    int maxdeg = 0;
    /* node -> elements adjacency */
    int nn = m->nn;
    int *nodeg = (int *)calloc(nn + 1, sizeof(int));
    for (int e = 0; e < m->nb; e++) for (int a = 0; a < 3; a++) nodeg[m->bconn[3 * e + a] + 1]++;
    for (int e = 0; e < m->ns; e++) for (int a = 0; a < 6; a++) nodeg[m->sconn[6 * e + a] + 1]++;
    for (int i = 0; i < nn; i++) nodeg[i + 1] += nodeg[i];
    int *nodel = (int *)malloc(sizeof(int) * (nodeg[nn] + 1));
    int *fill = (int *)calloc(nn, sizeof(int));
    for (int e = 0; e < m->nb; e++) for (int a = 0; a < 3; a++) { int n = m->bconn[3 * e + a]; nodel[nodeg[n] + fill[n]++] = e; }
    for (int e = 0; e < m->ns; e++) for (int a = 0; a < 6; a++) { int n = m->sconn[6 * e + a]; nodel[nodeg[n] + fill[n]++] = m->nb + e; }
    /* neighbouring nodes of each node */
    int cap = 256;
    int *nb = (int *)malloc(sizeof(int) * cap);
    long long total = 0;
    int **rows = (int **)calloc(nn, sizeof(int *));
    int *rown = (int *)calloc(nn, sizeof(int));
    for (int n = 0; n < nn; n++) {
        int k = 0;
        for (int p = nodeg[n]; p < nodeg[n + 1]; p++) {
            int e = nodel[p];
            const int *conn = e < m->nb ? m->bconn + 3 * e : m->sconn + 6 * (e - m->nb);
            int nc = e < m->nb ? 3 : 6;
            for (int a = 0; a < nc; a++) {
                if (k + 1 >= cap) { cap *= 2; nb = (int *)realloc(nb, sizeof(int) * cap); }
                nb[k++] = conn[a];
            }
        }
        qsort(nb, k, sizeof(int), cmp_int);
        int u = 0;
        for (int i = 0; i < k; i++) if (u == 0 || nb[u - 1] != nb[i]) nb[u++] = nb[i];
        rows[n] = (int *)malloc(sizeof(int) * (u ? u : 1));
        memcpy(rows[n], nb, sizeof(int) * u);
        rown[n] = u;
    }
    free(nb);
    /* count entries row by row (free DOFs, node-major) */
    m->ia = (int *)calloc(nf + 1, sizeof(int));
    for (int n = 0; n < nn; n++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * n + k];
            if (g <= 0) continue;
            int c = 0;
            for (int q = 0; q < rown[n]; q++)
                for (int kk = 0; kk < 6; kk++) if (m->GL[6 * rows[n][q] + kk] > 0) c++;
            m->ia[g] = c;
            if (c > maxdeg) maxdeg = c;
        }
    for (int i = 0; i < nf; i++) m->ia[i + 1] += m->ia[i];
    total = m->ia[nf];
    m->nnz = total;
    m->ja = (int *)malloc(sizeof(int) * (total ? total : 1));
    m->a = (double *)calloc(total ? total : 1, sizeof(double));
    for (int n = 0; n < nn; n++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * n + k];
            if (g <= 0) continue;
            int p = m->ia[g - 1];
            for (int q = 0; q < rown[n]; q++)
                for (int kk = 0; kk < 6; kk++) {
                    int gc = m->GL[6 * rows[n][q] + kk];
                    if (gc > 0) m->ja[p++] = gc - 1;
                }
            qsort(m->ja + m->ia[g - 1], m->ia[g] - m->ia[g - 1], sizeof(int), cmp_int);
        }
    for (int n = 0; n < nn; n++) free(rows[n]);
    free(rows); free(rown); free(nodeg); free(nodel); free(fill); free(cnt);
    return GF_OK;
}

/* ------------------------------------------------------------ solve */
int gfo_solve(gfo_model *m, int *info_fail)
{
    int nf = m->n_free;
    *info_fail = 0;
    if (nf == 0) return GF_OK;
    if (m->solver == 0) {
        int info = gfo_gbtrf(nf, m->kl, m->ku, m->ab, m->ldab, m->ipiv);
        m->factored = 1;
        if (info != 0) { *info_fail = 1; return GF_OK; }
        gfo_gbtrs(nf, m->kl, m->ku, m->ab, m->ldab, m->ipiv, m->P);
        for (int i = 0; i < nf; i++)
            if (!(m->P[i] == m->P[i]) || fabs(m->P[i]) >= 1e300) { *info_fail = 1; break; }
        return GF_OK;
    }
    /* sparsesystem: mtype 11, fresh handle, phase 13 then -1 */
    void *pt[64];
    int iparm[64];
    int mtype = 11, nrhs = 1, maxfct = 1, mnum = 1, phase, msglvl = 0, err = 0, idum = 0;
    double ddum = 0.0;
    memset(pt, 0, sizeof(pt));
    p_init(pt, &mtype, iparm);
    iparm[1] = 3;
    for (int i = 0; i <= nf; i++) m->ia[i] += 1;
    for (long long i = 0; i < m->nnz; i++) m->ja[i] += 1;
    double *x = (double *)calloc(nf, sizeof(double));
    phase = 13;
    p_solve(pt, &maxfct, &mnum, &mtype, &phase, &nf, m->a, m->ia, m->ja, &idum, &nrhs, iparm, &msglvl, m->P, x, &err);
    for (int i = 0; i <= nf; i++) m->ia[i] -= 1;
    for (long long i = 0; i < m->nnz; i++) m->ja[i] -= 1;
    int e1 = err;
    phase = -1;
    p_solve(pt, &maxfct, &mnum, &mtype, &phase, &nf, &ddum, m->ia, m->ja, &idum, &nrhs, iparm, &msglvl, &ddum, &ddum, &err);
    if (e1 != 0 || err != 0) *info_fail = 1;
    memcpy(m->P, x, sizeof(double) * nf);
    free(x);
    m->factored = 1;
    return GF_OK;
}

/* ------------------------------------------------------------ update */
int gfo_update(gfo_model *m, double *out)
{
    double mu = 0, mr = 0, vu = 0, vr = 0;
    int bad = 0;
    memset(m->xinc, 0, sizeof(double) * 6 * (size_t)m->nn);
    for (int i = 0; i < m->nn; i++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * i + k];
            if (g > 0 && m->active[6 * i + k] && !m->fix[6 * i + k]) {
                m->d[6 * i + k] += m->P[g - 1];
                m->xinc[6 * i + k] = m->P[g - 1];
            }
        }
    /* CheckGLConvergence scans (after the update) */
    for (int i = 0; i < m->nn; i++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * i + k];
            if (g <= 0) continue;
            double inc = fabs(m->P[g - 1]);
            if (!(inc == inc) || inc >= 1e300) bad = 1;
            double val = fabs(m->d[6 * i + k]);
            if (k < 3) { if (inc > mu) mu = inc; if (val > vu) vu = val; }
            else { if (inc > mr) mr = inc; if (val > vr) vr = val; }
        }
    if (out) { out[0] = mu; out[1] = mr; out[2] = vu; out[3] = vr; out[4] = bad ? 1.0 : 0.0; }
    return GF_OK;
}

int gfo_commit(gfo_model *m)
{
    for (int i = 0; i < m->nn; i++) {
        double *c = m->copy + 6 * i, *d = m->d + 6 * i;
        for (int k = 0; k < 3; k++) c[k] += d[k];
        double a3[3];
        gfo_compose(d + 3, c + 3, a3);
        for (int k = 0; k < 3; k++) c[3 + k] = a3[k];
        double na = sqrt(v3_dot(a3, a3));
        double th = 2.0 * atan(na / 2.0);
        for (int k = 0; k < 3; k++) m->euler[3 * i + k] = na != 0.0 ? (th / na) * a3[k] : 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->nb; e++) {
        double d[18];
        gather_d(m, m->bconn + 3 * e, 3, d);
        gfo_beam_save(&m->bc[e], m->bgp + (size_t)e * 2 * GFO_BEAM_GPS, d);
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->ns; e++) {
        double d[36];
        gather_d(m, m->sconn + 6 * e, 6, d);
        gfo_shell_save(gfo_shell_at(m->sc, e), m->sgp + (size_t)e * 3 * GFO_SHELL_GPS, d);
    }
    return GF_OK;
}

int gfo_rollback(gfo_model *m)
{
    memset(m->d, 0, sizeof(double) * 6 * (size_t)m->nn);
    return GF_OK;
}

/* ------------------------------------------------------------ reads */
void gfo_read_nodes(const gfo_model *m, double *out)
{
    for (int i = 0; i < m->nn; i++) {
        for (int k = 0; k < 3; k++) out[6 * i + k] = m->copy[6 * i + k];
        for (int k = 0; k < 3; k++) out[6 * i + 3 + k] = m->euler[3 * i + k];
    }
}
void gfo_read_copy(const gfo_model *m, double *out) { memcpy(out, m->copy, sizeof(double) * 6 * (size_t)m->nn); }
void gfo_read_disp(const gfo_model *m, double *out) { memcpy(out, m->d, sizeof(double) * 6 * (size_t)m->nn); }
void gfo_write_disp(gfo_model *m, const double *in) { memcpy(m->d, in, sizeof(double) * 6 * (size_t)m->nn); }
void gfo_read_increment(const gfo_model *m, double *out) { memcpy(out, m->xinc, sizeof(double) * 6 * (size_t)m->nn); }
void gfo_read_residual(const gfo_model *m, double *out)
{
    for (int i = 0; i < m->nn; i++)
        for (int k = 0; k < 6; k++) {
            int g = m->GL[6 * i + k];
            out[6 * i + k] = g > 0 ? m->P[g - 1] : 0.0;
        }
}
/* dense K_AA in node-DOF space (before the solve) */
int gfo_read_tangent_dense(const gfo_model *m, double *out)
{
    if (m->factored) return GF_E_STATE;
    int N6 = 6 * m->nn;
    memset(out, 0, sizeof(double) * (size_t)N6 * N6);
    int kv = m->kl + m->ku;
    for (int i = 0; i < N6; i++) {
        int gi = m->GL[i];
        if (gi <= 0) continue;
        for (int j = 0; j < N6; j++) {
            int gj = m->GL[j];
            if (gj <= 0) continue;
            double v = 0.0;
            if (m->solver == 0) {
                if (gi - gj <= m->kl && gj - gi <= m->ku)
                    v = m->ab[(size_t)(kv + (gi - 1) - (gj - 1)) + (size_t)(gj - 1) * m->ldab];
            } else {
                int p = csr_find(m, gi - 1, gj - 1);
                if (p >= 0) v = m->a[p];
            }
            out[(size_t)i * N6 + j] = v;
        }
    }
    return GF_OK;
}
void gfo_read_gp_beam(const gfo_model *m, double *out) { memcpy(out, m->bgp, sizeof(double) * (size_t)m->nb * 2 * GFO_BEAM_GPS); }
void gfo_read_gp_shell(const gfo_model *m, double *out) { memcpy(out, m->sgp, sizeof(double) * (size_t)m->ns * 3 * GFO_SHELL_GPS); }
/* The element results of the last Mount (the values Beam_1 / Shell_1 ::WriteResults print). */
void gfo_read_element_results(const gfo_model *m, double *beam_out, double *shell_out)
{
    if (beam_out) memcpy(beam_out, m->bres, sizeof(double) * 12 * (size_t)m->nb);
    if (shell_out) memcpy(shell_out, m->sres, sizeof(double) * 24 * (size_t)m->ns);
}
void gfo_read_element_beam(const gfo_model *m, int e, double *K, double *P)
{
    memcpy(K, m->bK + 324 * (size_t)e, sizeof(double) * 324);
    memcpy(P, m->bP + 18 * (size_t)e, sizeof(double) * 18);
}
void gfo_read_element_shell(const gfo_model *m, int e, double *K, double *P)
{
    memcpy(K, m->sK + 729 * (size_t)e, sizeof(double) * 729);
    memcpy(P, m->sP + 27 * (size_t)e, sizeof(double) * 27);
}

/* ------------------------------------------------------------ Static::Solve */
typedef struct conv_state {
    double force_crit, moment_crit;
    int n_conv_evaluated;
} conv_state;

/* EstablishResidualCriteria src/ConvergenceCriteria.cpp:254-287: the criteria of
 * a load step's first iteration (tol x max(max |r|, min_ref)), averaged with the
 * criteria of all earlier steps of the solution (n_conv_evaluated of them).
 * crit[0..1] = force / moment criterion in and out; *n counts the evaluations. */
void gfo_residual_criteria(double *crit, int *n, double max_force, double max_moment, const gf_static_params *p)
{
    double pf = crit[0], pm = crit[1];
    double fc = (max_force * p->force_tol > p->force_min * p->force_tol) ? max_force * p->force_tol : p->force_min * p->force_tol;
    double mc = (max_moment * p->moment_tol > p->moment_min * p->moment_tol) ? max_moment * p->moment_tol : p->moment_min * p->moment_tol;
    crit[0] = (*n * pf + fc) / (*n + 1);
    crit[1] = (*n * pm + mc) / (*n + 1);
    (*n)++;
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int gfo_static_solve(gfo_model *m, const gf_static_params *p, gf_static_report *rep)
{
    double t0 = now_ms();
    gf_static_report r;
    memset(&r, 0, sizeof(r));
    r.step_iters = rep ? rep->step_iters : NULL;
    r.step_iters_cap = rep ? rep->step_iters_cap : 0;
    int st = gfo_begin_solution(m, p->start_time, p->end_time);
    if (st != GF_OK) return st;
    conv_state cs = {0.0, 0.0, 0};
    int aborted = 0, convergence_counter = 0, converged_number = 0, diverged = 0;
    int verbose = getenv("GFO_VERBOSE") ? atoi(getenv("GFO_VERBOSE")) : 0;
    double time_step = p->time_step;
    double time = p->start_time;
    double last_converged = p->start_time;
    while ((time < p->end_time || diverged) && !aborted) {
        last_converged = time;
        time += time_step;
        if (time > 0.99999999 * p->end_time) {
            time = p->end_time;
            time_step = p->end_time - last_converged;
        }
        gfo_begin_step(m, last_converged, time_step);
        int it = 1;
        int res_conv = 0, gl_conv = 0;
        diverged = 0;
        double last_rel = 0.0;
        while ((!res_conv || !gl_conv) && !diverged && !aborted) {
            double res[3], ta = now_ms();
            gfo_assemble(m, res);
            r.t_assemble_ms += now_ms() - ta;
            r.total_iterations++;
            if (it == 1) {
                /* EstablishResidualCriteria src/ConvergenceCriteria.cpp:187-288 */
                double crit[2] = {cs.force_crit, cs.moment_crit};
                gfo_residual_criteria(crit, &cs.n_conv_evaluated, res[0], res[1], p);
                cs.force_crit = crit[0];
                cs.moment_crit = crit[1];
            }
            /* CheckResidualConvergence src/ConvergenceCriteria.cpp:460-598 */
            int has_f = 0, has_m = 0;
            for (int i = 0; i < 6 * m->nn && !(has_f && has_m); i++)
                if (m->GL[i] > 0) { if (i % 6 < 3) has_f = 1; else has_m = 1; }
            if (res[2] != 0.0) diverged = 1;
            if (res[0] > p->divergence_ref || res[1] > p->divergence_ref) { diverged = 1; res_conv = 0; }
            else {
                res_conv = 1;
                if (has_f && res[0] > cs.force_crit) res_conv = 0;
                if (has_m && res[1] > cs.moment_crit) res_conv = 0;
            }
            last_rel = 0.0;
            if (has_f && cs.force_crit > 0) last_rel = fmax(last_rel, res[0] / (cs.force_crit / p->force_tol));
            if (has_m && cs.moment_crit > 0) last_rel = fmax(last_rel, res[1] / (cs.moment_crit / p->moment_tol));
            int info_fail = 0;
            if (!(res_conv && gl_conv)) {
                double ts = now_ms();
                gfo_solve(m, &info_fail);
                r.t_solve_ms += now_ms() - ts;
                if (verbose > 1) {
                    fprintf(stderr, "oracle: time %.6f iteration %d solved (%.1f s)\n", time, it, (now_ms() - t0) * 1e-3);
                    fflush(stderr);
                }
                r.total_solves++;
                res_conv = 0;
                double inc[5];
                gfo_update(m, inc);
                /* CheckGLConvergence src/ConvergenceCriteria.cpp:289-459 */
                if (inc[4] != 0.0) diverged = 1;
                double dc = (inc[2] * p->disp_tol > p->disp_min * p->disp_tol) ? p->disp_tol * inc[2] : p->disp_min * p->disp_tol;
                double rc = (inc[3] * p->rot_tol > p->rot_min * p->rot_tol) ? p->rot_tol * inc[3] : p->rot_min * p->rot_tol;
                if (dc > p->divergence_ref || rc > p->divergence_ref) { diverged = 1; gl_conv = 0; }
                else {
                    gl_conv = 1;
                    if (has_f && inc[0] > dc) gl_conv = 0;
                    if (has_m && inc[1] > rc) gl_conv = 0;
                }
            }
            if ((it == p->max_it && !(res_conv && gl_conv)) || info_fail == 1) diverged = 1;
            if (time_step < p->min_time_step && time != p->end_time) aborted = 1;
            it++;
        }
        if (!diverged) {
            converged_number++;
            gfo_commit(m);
            convergence_counter++;
            if (r.n_step_iters < r.step_iters_cap && r.step_iters) r.step_iters[r.n_step_iters] = it - 1;
            r.n_step_iters++;
            r.converged_steps++;
            if (last_rel > r.max_rel_residual) r.max_rel_residual = last_rel;
            if (verbose) {
                fprintf(stderr, "oracle: time %.6f step %.6f converged in %d iterations (%.1f s)\n", time, time_step,
                        it - 1, (now_ms() - t0) * 1e-3);
                fflush(stderr);
            }
            if (it <= p->min_it || convergence_counter >= p->conv_increase) {
                time_step = time_step * p->inc_factor;
                if (time_step > p->max_time_step) time_step = p->max_time_step;
                convergence_counter = 0;
            }
        } else {
            gfo_rollback(m);
            if (r.n_step_iters < r.step_iters_cap && r.step_iters) r.step_iters[r.n_step_iters] = -(it - 1);
            r.n_step_iters++;
            r.diverged_steps++;
            time -= time_step;
            time_step = time_step / 2.0;
            diverged = 0;
            convergence_counter = 0;
        }
    }
    r.aborted = aborted;
    r.final_time = time;
    r.t_total_ms = now_ms() - t0;
    if (rep) *rep = r;
    return GF_OK;
}
