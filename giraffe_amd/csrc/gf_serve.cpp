// Request server: many nonlinear static solves resident on one GPU at once.
//
// A request is a model description plus Static parameters.  Submitted
// requests wait in FIFO order until the device can hold them: the server
// admits the head of the queue when fewer than max_inflight requests are
// resident and its device footprint (gf_model_footprint) fits under mem_cap.
// Every resident request owns its Newton state, its tangent storage and a
// stream from the server's pool; its Static loop is a gf_static_run whose
// device steps are enqueued on that stream.  gf_server_pump polls the
// requests' events and advances each one whose device work has finished, so
// the device always holds the queued steps of every other request while the
// host decides the next step of one (continuous batching at Newton-step
// granularity).  Finished requests release their memory and stream at once
// and the next queued request is admitted in the same pump.
//
// Host planning stays off the thread that drives the device: submit plans the
// request's sizes (gf_model_preplan: the device footprint that decides
// admission) and keeps that plan while fewer than KEEP_PLANS requests are
// queued (beyond that it is dropped, so a long queue holds only the copied
// descriptions), and a planner thread completes the plans of the first
// PLAN_AHEAD queued requests (the sizes again where they were dropped, then the
// multifrontal per-entry lists) while the driving thread keeps the resident
// requests going; the queue head is admitted once its plan is complete, and
// admission then only allocates and uploads.
//
// Streams are created at the default priority.  One host thread drives a server
// (submit, pump, take, counts); the planner thread is the server's own.
#include "gf_internal.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

// Owned copy of a gf_model_desc.
struct DescCopy {
    std::vector<double> xyz, beam_cs, beam_prop, shell_prop, load_cs, tab_time, tab_val, beam_buoy;
    std::vector<unsigned char> fix;
    std::vector<int> beam_conn, shell_conn, load_kind, load_set_off, load_set_items, load_flags, load_tab_off;
    std::vector<int> beam_kind, beam_number, shell_number;
    gf_model_desc d{};

    // Every array the description points at is copied: the model is created
    // later, when the request is admitted, after the caller's arrays may be gone.
    void copy_from(const gf_model_desc &s)
    {
        const size_t nn = (size_t)s.n_nodes, nb = (size_t)s.n_beams, ns = (size_t)s.n_shells, nl = (size_t)s.n_loads;
        xyz.assign(s.node_xyz, s.node_xyz + 3 * nn);
        if (s.node_fix) fix.assign(s.node_fix, s.node_fix + 6 * nn);
        if (nb) {
            beam_conn.assign(s.beam_conn, s.beam_conn + 3 * nb);
            beam_cs.assign(s.beam_cs, s.beam_cs + 9 * nb);
            beam_prop.assign(s.beam_prop, s.beam_prop + GF_BEAM_NPROP * nb);
            if (s.beam_kind) beam_kind.assign(s.beam_kind, s.beam_kind + nb);
            if (s.beam_buoy) beam_buoy.assign(s.beam_buoy, s.beam_buoy + nb);
            if (s.beam_number) beam_number.assign(s.beam_number, s.beam_number + nb);
        }
        if (ns) {
            shell_conn.assign(s.shell_conn, s.shell_conn + 6 * ns);
            shell_prop.assign(s.shell_prop, s.shell_prop + GF_SHELL_NPROP * ns);
            if (s.shell_number) shell_number.assign(s.shell_number, s.shell_number + ns);
        }
        if (nl) {
            load_kind.assign(s.load_kind, s.load_kind + nl);
            load_set_off.assign(s.load_set_off, s.load_set_off + nl + 1);
            load_set_items.assign(s.load_set_items, s.load_set_items + load_set_off[nl]);
            load_cs.assign(s.load_cs, s.load_cs + 9 * nl);
            if (s.load_flags) load_flags.assign(s.load_flags, s.load_flags + nl);
            load_tab_off.assign(s.load_tab_off, s.load_tab_off + nl + 1);
            size_t rows = (size_t)load_tab_off[nl];
            tab_time.assign(s.tab_time, s.tab_time + rows);
            tab_val.assign(s.tab_val, s.tab_val + 6 * rows);
        }
        d = s;
        d.node_xyz = xyz.data();
        d.node_fix = fix.empty() ? nullptr : fix.data();
        d.beam_conn = beam_conn.data();
        d.beam_cs = beam_cs.data();
        d.beam_prop = beam_prop.data();
        d.shell_conn = shell_conn.data();
        d.shell_prop = shell_prop.data();
        d.load_kind = load_kind.data();
        d.load_set_off = load_set_off.empty() ? nullptr : load_set_off.data();
        d.load_set_items = load_set_items.data();
        d.load_cs = load_cs.data();
        d.load_flags = load_flags.empty() ? nullptr : load_flags.data();
        d.load_tab_off = load_tab_off.empty() ? nullptr : load_tab_off.data();
        d.tab_time = tab_time.data();
        d.tab_val = tab_val.data();
        d.beam_kind = beam_kind.empty() ? nullptr : beam_kind.data();
        d.beam_buoy = beam_buoy.empty() ? nullptr : beam_buoy.data();
        d.beam_number = beam_number.empty() ? nullptr : beam_number.data();
        d.shell_number = shell_number.empty() ? nullptr : shell_number.data();
    }
};

// Lists of at most this many queued requests are built ahead of admission; the
// size plans of at most KEEP_PLANS queued requests are kept from submit.
constexpr int PLAN_AHEAD = 4, KEEP_PLANS = 32;

enum { PLAN_SIZES = 0, PLAN_BUSY = 1, PLAN_DONE = 2 };

struct Request {
    long long id = 0;
    DescCopy desc;
    gf_static_params params{};
    unsigned long long bytes = 0;
    gf_model *model = nullptr;  // planned (submit, planner thread), created at admission
    std::atomic<int> plan_state{PLAN_SIZES};
    int plan_status = GF_OK;    // of the planner's list build
    std::string plan_error;
    int stream_slot = -1;
    gf_static_run run;
    std::vector<int> step_iters;
    std::vector<double> nodes;
    gf_static_report report{};
    int status = GF_OK;
    clk::time_point t_submit, t_admit, t_done;
};

}  // namespace

struct gf_server {
    gf_server_options opt{};
    std::vector<hipStream_t> streams;
    std::vector<int> free_slots;
    std::mutex mu;               // guards queue (the planner reads it)
    gf_options plan_opt{};       // what the planner plans requests with
    std::condition_variable cv;  // planner wake-up: a new request, a queue move, stop
    bool stop = false;
    std::thread planner;
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> inflight;
    std::map<long long, std::unique_ptr<Request>> done;
    unsigned long long mem_in_use = 0;
    long long next_id = 1;
};

extern "C" int gf_default_server_options(gf_server_options *o)
{
    if (!o) return GF_E_ARG;
    memset(o, 0, sizeof(*o));
    o->precision = GF_PREC_FP64;
    o->device = 0;
    o->max_inflight = 16;
    o->mem_cap = 200ull << 30;
    return GF_OK;
}

// Planner thread: builds the lists of the first queued requests, in queue order.
static void planner_main(gf_server *s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        Request *next = nullptr;
        s->cv.wait(lk, [&] {
            if (s->stop) return true;
            const size_t n = std::min(s->queue.size(), (size_t)PLAN_AHEAD);
            for (size_t i = 0; i < n; i++)
                if (s->queue[i]->plan_state.load() == PLAN_SIZES) {
                    next = s->queue[i].get();
                    return true;
                }
            return false;
        });
        if (s->stop) return;
        // the driving thread admits a request only once it is PLAN_DONE, so
        // `next` stays in the queue (and alive) until then
        next->plan_state.store(PLAN_BUSY);
        lk.unlock();
        int st;
        try {
            unsigned long long bytes = 0;
            st = next->model ? GF_OK : gf_model_preplan(&next->desc.d, &s->plan_opt, &next->model, &bytes);
            if (st == GF_OK) st = gf_model_plan_lists(&next->desc.d, next->model);
            if (st != GF_OK) next->plan_error = gf_last_error();
        } catch (const std::exception &e) {
            st = GF_E_UNSUPPORTED;
            next->plan_error = std::string("request planning: ") + e.what();
        }
        next->plan_status = st;
        lk.lock();
        next->plan_state.store(PLAN_DONE);
    }
}

extern "C" int gf_server_create(const gf_server_options *opt, gf_server **out)
{
    if (!opt || !out || opt->max_inflight < 1 || opt->max_inflight > 256) return GF_E_ARG;
    auto s = std::make_unique<gf_server>();
    s->opt = *opt;
    GF_HIP(hipSetDevice(opt->device));
    s->streams.resize(opt->max_inflight, nullptr);
    for (int i = 0; i < opt->max_inflight; i++) {
        GF_HIP(hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking));
        s->free_slots.push_back(opt->max_inflight - 1 - i);
    }
    gf_default_options(&s->plan_opt);
    s->plan_opt.precision = opt->precision;
    s->plan_opt.device = opt->device;
    s->planner = std::thread(planner_main, s.get());
    *out = s.release();
    return GF_OK;
}

extern "C" int gf_server_destroy(gf_server *s)
{
    if (!s) return GF_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->planner.join();
    (void)hipSetDevice(s->opt.device);
    for (auto &r : s->queue) gf_model_destroy(r->model);
    for (auto &r : s->inflight) {
        if (r->model) {
            (void)hipStreamSynchronize(s->streams[r->stream_slot]);
            gf_model_destroy(r->model);
        }
    }
    for (hipStream_t st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    delete s;
    return GF_OK;
}

extern "C" int gf_server_submit(gf_server *s, const gf_model_desc *d, const gf_static_params *p, long long *id_out)
{
    if (!s || !d || !p || !id_out) return GF_E_ARG;
    auto r = std::make_unique<Request>();
    r->desc.copy_from(*d);
    r->params = *p;
    gf_options o;
    gf_default_options(&o);
    o.precision = s->opt.precision;
    o.device = s->opt.device;
    GF_TRY(gf_model_preplan(&r->desc.d, &o, &r->model, &r->bytes));
    if (r->bytes > s->opt.mem_cap || (int)s->queue.size() >= KEEP_PLANS) {
        gf_model_destroy(r->model);
        r->model = nullptr;
    }
    if (r->bytes > s->opt.mem_cap) {
        gf_set_error("request needs more device memory than the server's cap");
        return GF_E_UNSUPPORTED;
    }
    r->id = s->next_id++;
    r->step_iters.assign(4096, 0);
    r->t_submit = clk::now();
    *id_out = r->id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queue.push_back(std::move(r));
    }
    s->cv.notify_one();
    return GF_OK;
}

// Creates the head request's model on a pooled stream and starts its run.
static int admit(gf_server *s, Request &r)
{
    gf_options o;
    gf_default_options(&o);
    o.precision = s->opt.precision;
    o.device = s->opt.device;
    r.stream_slot = s->free_slots.back();
    s->free_slots.pop_back();
    o.stream = (void *)s->streams[r.stream_slot];
    r.t_admit = clk::now();
    if (r.plan_status != GF_OK) {
        gf_model_destroy(r.model);  // a partial plan, if any
        r.model = nullptr;
        gf_set_error(r.plan_error);
        return r.plan_status;
    }
    s->mem_in_use += r.bytes;  // finish() gives it back, created or not
    int st = gf_model_create_planned(&r.desc.d, &o, r.model);
    if (st) return st;
    st = r.run.start(r.model, r.params, r.step_iters.data(), (int)r.step_iters.size());
    if (st) return st;
    st = r.run.advance();
    return st < 0 ? st : GF_OK;
}

static void finish(gf_server *s, std::unique_ptr<Request> &r, int status)
{
    r->status = status;
    r->t_done = clk::now();
    if (r->model) {
        if (status == GF_OK) {
            r->report = r->run.r;
            r->nodes.resize(6 * (size_t)r->desc.d.n_nodes);
            if (gf_read_nodes(r->model, r->nodes.data()) != GF_OK) r->nodes.clear();
        }
        (void)hipStreamSynchronize(s->streams[r->stream_slot]);
        gf_model_destroy(r->model);
        r->model = nullptr;
        s->mem_in_use -= r->bytes;
    }
    if (r->stream_slot >= 0) s->free_slots.push_back(r->stream_slot);
    r->stream_slot = -1;
    s->done[r->id] = std::move(r);
}

extern "C" int gf_server_pump(gf_server *s, int wait_us)
{
    if (!s) return GF_E_ARG;
    GF_HIP(hipSetDevice(s->opt.device));
    int finished = 0;
    auto deadline = clk::now() + std::chrono::microseconds(wait_us > 0 ? wait_us : 0);
    for (;;) {
        // admission: FIFO head while slots and memory allow and its plan is complete
        for (;;) {
            std::unique_ptr<Request> r;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->queue.empty() || (int)s->inflight.size() >= s->opt.max_inflight ||
                    s->mem_in_use + s->queue.front()->bytes > s->opt.mem_cap ||
                    s->queue.front()->plan_state.load() != PLAN_DONE)
                    break;
                r = std::move(s->queue.front());
                s->queue.pop_front();
            }
            s->cv.notify_one();
            int st = admit(s, *r);
            if (st) {
                finish(s, r, st);
                finished++;
                continue;
            }
            if (r->run.phase == gf_static_run::PH_DONE) {
                finish(s, r, GF_OK);
                finished++;
            } else {
                s->inflight.push_back(std::move(r));
            }
        }
        // progress: advance every request whose device work has completed
        bool progressed = false;
        for (size_t i = 0; i < s->inflight.size();) {
            Request &r = *s->inflight[i];
            hipError_t q = hipEventQuery(r.run.ev);
            if (q == hipErrorNotReady) { i++; continue; }
            if (q != hipSuccess) {
                gf_set_error(std::string("request event: ") + hipGetErrorString(q));
                return GF_E_DEVICE;
            }
            progressed = true;
            int st = r.run.advance();
            if (st == 1) { i++; continue; }
            finish(s, s->inflight[i], st < 0 ? st : GF_OK);
            s->inflight.erase(s->inflight.begin() + (long)i);
            finished++;
        }
        if (s->inflight.empty() && s->queue.empty()) break;
        if (clk::now() >= deadline) break;
        if (!progressed) std::this_thread::yield();
    }
    return finished;
}

extern "C" int gf_server_counts(const gf_server *s, long long *out)
{
    if (!s || !out) return GF_E_ARG;
    out[0] = (long long)s->queue.size();
    out[1] = (long long)s->inflight.size();
    out[2] = (long long)s->done.size();
    out[3] = (long long)s->mem_in_use;
    return GF_OK;
}

extern "C" int gf_server_take(gf_server *s, long long id, gf_static_report *rep, double *nodes_out, double *times_ms)
{
    if (!s) return GF_E_ARG;
    auto it = s->done.find(id);
    if (it == s->done.end()) return GF_E_ARG;
    Request &r = *it->second;
    int status = r.status;
    if (rep) {
        int *iters = rep->step_iters;
        int cap = rep->step_iters_cap;
        *rep = r.report;
        rep->step_iters = iters;
        rep->step_iters_cap = cap;
        if (iters)
            for (int k = 0; k < r.report.n_step_iters && k < cap && k < (int)r.step_iters.size(); k++)
                iters[k] = r.step_iters[k];
    }
    if (nodes_out && !r.nodes.empty()) memcpy(nodes_out, r.nodes.data(), sizeof(double) * r.nodes.size());
    if (times_ms) {
        times_ms[0] = ms_between(r.t_submit, r.t_admit);
        times_ms[1] = ms_between(r.t_submit, r.t_done);
        times_ms[2] = ms_between(r.t_admit, r.t_done);
    }
    s->done.erase(it);
    return status;
}
