// The Static load-step loop (drop-in for bool Static::Solve(), src/Static.cpp:140-311)
// driving the device Newton step.
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
//
// The loop is a resumable state machine (gf_static_run): every device step is
// enqueued on the model's stream and followed by an event, and the host logic
// resumes once the event has completed.  A solve batch also assembles the next
// iteration (the loop always assembles after a solve unless the step ends, and
// assembly changes no state), so a Newton iteration costs one host round trip.  gf_static_solve drives one run to the
// end, waiting on each event; the request server drives many runs at once,
// polling their events, so the device interleaves the requests' work.
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

int gf_static_run::start(gf_model *model, const gf_static_params &params, int *step_iters, int cap)
{
    m = model;
    p = params;
    memset(&r, 0, sizeof(r));
    r.step_iters = step_iters;
    r.step_iters_cap = cap;
    int st = gf_set_solution_window(m, p.start_time, p.end_time);
    if (st) return st;
    has_f = has_m = false;
    for (int nc : m->row_nodecomp) {
        if (nc % 6 < 3) has_f = true;
        else has_m = true;
    }
    force_crit = moment_crit = 0.0;
    n_conv_evaluated = 0;
    aborted = diverged = false;
    convergence_counter = 0;
    time_step = p.time_step;
    time = p.start_time;
    last_converged = p.start_time;
    phase = PH_STEP;
    if (!ev) GF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    t0 = std::chrono::steady_clock::now();
    return GF_OK;
}

gf_static_run::~gf_static_run()
{
    if (ev) (void)hipEventDestroy(ev);
}

// Advances the run as far as the host can without device results.
// Returns 1 when device work is pending (wait for `ev`), 0 when the run ended.
int gf_static_run::advance()
{
    int st;
    for (;;) {
        switch (phase) {
        case PH_STEP: {
            if (!((time < p.end_time || diverged) && !aborted)) {
                r.aborted = aborted ? 1 : 0;
                r.final_time = time;
                r.t_total_ms = ms_since(t0);
                phase = PH_DONE;
                return 0;
            }
            last_converged = time;
            time += time_step;
            if (time > 0.99999999 * p.end_time) {
                time = p.end_time;
                time_step = p.end_time - last_converged;
            }
            if ((st = gf_begin_step(m, last_converged, time_step))) return st;
            it = 1;
            res_conv = gl_conv = false;
            diverged = false;
            last_rel = 0.0;
            if ((st = enqueue_assemble())) return st;
            return 1;
        }
        case PH_ASSEMBLED: {
            double res[3];
            if ((st = gf_assemble_result(m, res))) return st;
            r.total_iterations++;
            if (it == 1) {
                double pf = force_crit, pm = moment_crit;
                double fc = (res[0] * p.force_tol > p.force_min * p.force_tol) ? res[0] * p.force_tol
                                                                              : p.force_min * p.force_tol;
                double mc = (res[1] * p.moment_tol > p.moment_min * p.moment_tol) ? res[1] * p.moment_tol
                                                                                 : p.moment_min * p.moment_tol;
                force_crit = (n_conv_evaluated * pf + fc) / (n_conv_evaluated + 1);
                moment_crit = (n_conv_evaluated * pm + mc) / (n_conv_evaluated + 1);
                n_conv_evaluated++;
            }
            if (res[2] != 0.0) diverged = true;
            if (res[0] > p.divergence_ref || res[1] > p.divergence_ref) {
                diverged = true;
                res_conv = false;
            } else {
                res_conv = true;
                if (has_f && res[0] > force_crit) res_conv = false;
                if (has_m && res[1] > moment_crit) res_conv = false;
            }
            last_rel = 0.0;
            if (has_f && force_crit > 0) last_rel = std::fmax(last_rel, res[0] / (force_crit / p.force_tol));
            if (has_m && moment_crit > 0) last_rel = std::fmax(last_rel, res[1] / (moment_crit / p.moment_tol));
            info_fail = 0;
            if (!(res_conv && gl_conv)) {
                t_mark = std::chrono::steady_clock::now();
                // the next iteration's assembly rides in the same batch
                if ((st = gf_solve_update_enqueue(m, true))) return st;
                if ((st = record())) return st;
                r.total_solves++;
                res_conv = false;
                phase = m->refining ? PH_REFINE : PH_SOLVED;
                return 1;
            }
            phase = PH_ITER_END;
            break;
        }
        case PH_REFINE: {
            // mixed precision: the last refinement sweep's norms are on the host
            int more = 0;
            if ((st = gf_refine_step(m, &more))) return st;
            if ((st = record())) return st;
            if (!more) phase = PH_SOLVED;
            return 1;
        }
        case PH_SOLVED: {
            r.t_solve_ms += ms_since(t_mark);
            double inc[5];
            if ((st = gf_solve_update_result(m, &info_fail, inc))) return st;
            if (inc[4] != 0.0) diverged = true;
            double dc = (inc[2] * p.disp_tol > p.disp_min * p.disp_tol) ? p.disp_tol * inc[2]
                                                                       : p.disp_min * p.disp_tol;
            double rc = (inc[3] * p.rot_tol > p.rot_min * p.rot_tol) ? p.rot_tol * inc[3] : p.rot_min * p.rot_tol;
            if (dc > p.divergence_ref || rc > p.divergence_ref) {
                diverged = true;
                gl_conv = false;
            } else {
                gl_conv = true;
                if (has_f && inc[0] > dc) gl_conv = false;
                if (has_m && inc[1] > rc) gl_conv = false;
            }
            phase = PH_ITER_END;
            break;
        }
        case PH_ITER_END: {
            if ((it == p.max_it && !(res_conv && gl_conv)) || info_fail == 1) diverged = true;
            if (time_step < p.min_time_step && time != p.end_time) aborted = true;
            it++;
            if ((!res_conv || !gl_conv) && !diverged && !aborted) {
                // every iteration that continues follows a solve (res_conv is reset
                // there), whose batch already assembled and scanned the residual
                phase = PH_ASSEMBLED;
                break;
            }
            // end of the load step
            if (!diverged) {
                if ((st = gf_commit_enqueue(m))) return st;
                convergence_counter++;
                if (r.step_iters && r.n_step_iters < r.step_iters_cap) r.step_iters[r.n_step_iters] = it - 1;
                r.n_step_iters++;
                r.converged_steps++;
                if (last_rel > r.max_rel_residual) r.max_rel_residual = last_rel;
                if (it <= p.min_it || convergence_counter >= p.conv_increase) {
                    time_step = time_step * p.inc_factor;
                    if (time_step > p.max_time_step) time_step = p.max_time_step;
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
            phase = PH_STEP;
            break;
        }
        case PH_DONE:
            return 0;
        }
    }
}

int gf_static_run::enqueue_assemble()
{
    t_mark = std::chrono::steady_clock::now();
    int st = gf_assemble_enqueue(m, true);
    if (st) return st;
    if ((st = record())) return st;
    phase = PH_ASSEMBLED;
    return GF_OK;
}

int gf_static_run::record()
{
    GF_HIP(hipEventRecord(ev, m->stream));
    return GF_OK;
}

// Waits for the pending device work of the run (blocking driver).
int gf_static_run::wait()
{
    GF_HIP(hipEventSynchronize(ev));
    if (phase == PH_ASSEMBLED) r.t_assemble_ms += ms_since(t_mark);
    return GF_OK;
}

extern "C" int gf_static_solve(gf_model *m, const gf_static_params *p, gf_static_report *rep)
{
    if (!m || !p) return GF_E_ARG;
    GF_HIP(hipSetDevice(m->device));
    gf_static_run run;
    int st = run.start(m, *p, rep ? rep->step_iters : nullptr, rep ? rep->step_iters_cap : 0);
    if (st) return st;
    while ((st = run.advance()) == 1)
        if ((st = run.wait())) return st;
    if (st < 0) return st;
    if (rep) *rep = run.r;
    return GF_OK;
}

// The sequence of solution steps as the reference's main loop runs them
// (src/main.cpp:53-62): each Static step in turn, its gravity factor mode set
// first, stopping after a step that aborts (Static::Solve returns false).  The
// state carries over; each step starts its own criteria history, as
// Static::Solve does (src/Static.cpp:146).  reps[i] is filled for the steps
// run; *n_run counts them.
extern "C" int gf_static_solve_steps(gf_model *m, int n, const gf_static_params *steps, const int *gravity_modes,
                                     gf_static_report *reps, int *n_run)
{
    if (!m || n < 0 || (n && (!steps || !gravity_modes || !reps)) || !n_run) return GF_E_ARG;
    *n_run = 0;
    for (int i = 0; i < n; i++) {
        int st = gf_set_gravity_mode(m, gravity_modes[i]);
        if (st) return st;
        if ((st = gf_static_solve(m, &steps[i], &reps[i]))) return st;
        *n_run = i + 1;
        if (reps[i].aborted) break;
    }
    return GF_OK;
}
