// The Static load-step loop (drop-in for bool Static::Solve(), src/Static.cpp:140-311)
// driving the device Newton step through the C ABI.
//
// Control flow kept from the reference:
//   * load-step loop: t += dt, clamp at EndTime (src/Static.cpp:176-186);
//   * Newton loop while !(res_converged && GL_converged) (:200-250), with the
//     solve skipped when the residual converged on already-converged increments;
//   * residual criteria fixed at iteration 1 and averaged over the history of
//     the solution (ConvergenceCriteria::EstablishResidualCriteria
//     src/ConvergenceCriteria.cpp:187-288); increment criteria from the step's
//     accumulated displacements (CheckGLConvergence :289-459);
//   * divergence at MaxIt, on a solver failure or a non-finite / > DivergenceReference
//     norm; rollback and bisection (src/Static.cpp:238-242, 286-294); abort below
//     MinTimeStep (:244-248); step growth by IncFactor (:262-268).
#include "gf_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>

extern "C" int gf_set_solution_window(gf_model *m, double start, double end);

namespace {
double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

extern "C" int gf_static_solve(gf_model *m, const gf_static_params *p, gf_static_report *rep)
{
    if (!m || !p) return GF_E_ARG;
    auto t0 = std::chrono::steady_clock::now();
    gf_static_report r;
    memset(&r, 0, sizeof(r));
    if (rep) {
        r.step_iters = rep->step_iters;
        r.step_iters_cap = rep->step_iters_cap;
    }
    int st = gf_set_solution_window(m, p->start_time, p->end_time);
    if (st) return st;
    bool has_f = false, has_m = false;
    for (int nc : m->row_nodecomp) {
        if (nc % 6 < 3) has_f = true;
        else has_m = true;
    }
    double force_crit = 0.0, moment_crit = 0.0;
    int n_conv_evaluated = 0;
    bool aborted = false, diverged = false;
    int convergence_counter = 0;
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
        if ((st = gf_begin_step(m, last_converged, time_step))) return st;
        int it = 1;
        bool res_conv = false, gl_conv = false;
        diverged = false;
        double last_rel = 0.0;
        while ((!res_conv || !gl_conv) && !diverged && !aborted) {
            double res[3];
            auto ta = std::chrono::steady_clock::now();
            if ((st = gf_assemble(m, res))) return st;
            r.t_assemble_ms += ms_since(ta);
            r.total_iterations++;
            if (it == 1) {
                double pf = force_crit, pm = moment_crit;
                double fc = (res[0] * p->force_tol > p->force_min * p->force_tol) ? res[0] * p->force_tol
                                                                                  : p->force_min * p->force_tol;
                double mc = (res[1] * p->moment_tol > p->moment_min * p->moment_tol) ? res[1] * p->moment_tol
                                                                                     : p->moment_min * p->moment_tol;
                force_crit = (n_conv_evaluated * pf + fc) / (n_conv_evaluated + 1);
                moment_crit = (n_conv_evaluated * pm + mc) / (n_conv_evaluated + 1);
                n_conv_evaluated++;
            }
            if (res[2] != 0.0) diverged = true;
            if (res[0] > p->divergence_ref || res[1] > p->divergence_ref) {
                diverged = true;
                res_conv = false;
            } else {
                res_conv = true;
                if (has_f && res[0] > force_crit) res_conv = false;
                if (has_m && res[1] > moment_crit) res_conv = false;
            }
            last_rel = 0.0;
            if (has_f && force_crit > 0) last_rel = std::fmax(last_rel, res[0] / (force_crit / p->force_tol));
            if (has_m && moment_crit > 0) last_rel = std::fmax(last_rel, res[1] / (moment_crit / p->moment_tol));
            int info_fail = 0;
            if (!(res_conv && gl_conv)) {
                auto ts = std::chrono::steady_clock::now();
                if ((st = gf_solve(m, &info_fail))) return st;
                r.t_solve_ms += ms_since(ts);
                r.total_solves++;
                res_conv = false;
                double inc[5];
                if ((st = gf_update(m, inc))) return st;
                if (inc[4] != 0.0) diverged = true;
                double dc = (inc[2] * p->disp_tol > p->disp_min * p->disp_tol) ? p->disp_tol * inc[2]
                                                                               : p->disp_min * p->disp_tol;
                double rc = (inc[3] * p->rot_tol > p->rot_min * p->rot_tol) ? p->rot_tol * inc[3]
                                                                            : p->rot_min * p->rot_tol;
                if (dc > p->divergence_ref || rc > p->divergence_ref) {
                    diverged = true;
                    gl_conv = false;
                } else {
                    gl_conv = true;
                    if (has_f && inc[0] > dc) gl_conv = false;
                    if (has_m && inc[1] > rc) gl_conv = false;
                }
            }
            if ((it == p->max_it && !(res_conv && gl_conv)) || info_fail == 1) diverged = true;
            if (time_step < p->min_time_step && time != p->end_time) aborted = true;
            it++;
        }
        if (!diverged) {
            if ((st = gf_commit(m))) return st;
            convergence_counter++;
            if (r.step_iters && r.n_step_iters < r.step_iters_cap) r.step_iters[r.n_step_iters] = it - 1;
            r.n_step_iters++;
            r.converged_steps++;
            if (last_rel > r.max_rel_residual) r.max_rel_residual = last_rel;
            if (it <= p->min_it || convergence_counter >= p->conv_increase) {
                time_step = time_step * p->inc_factor;
                if (time_step > p->max_time_step) time_step = p->max_time_step;
                convergence_counter = 0;
            }
        } else {
            if ((st = gf_rollback(m))) return st;
            if (r.step_iters && r.n_step_iters < r.step_iters_cap) r.step_iters[r.n_step_iters] = -(it - 1);
            r.n_step_iters++;
            r.diverged_steps++;
            time -= time_step;
            time_step = time_step / 2.0;
            diverged = false;
            convergence_counter = 0;
        }
    }
    r.aborted = aborted ? 1 : 0;
    r.final_time = time;
    r.t_total_ms = ms_since(t0);
    if (rep) *rep = r;
    return GF_OK;
}
