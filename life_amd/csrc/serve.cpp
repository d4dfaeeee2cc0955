// Serving engine (life_server_*): continuous batching of independent LIFE
// cases on one device.
//
// Requests (params.h values, geometry.config text, a step count) queue in
// arrival order and are admitted while a slot is free and their device
// footprint fits the HBM budget.  Every engine step advances each admitted
// case by one timestep:
//  - one table of per-case descriptors (CaseDev) goes to the device;
//  - the lattice phase of all cases is two hipGraph launches, captured
//    again only when the set of cases changes: the bulk stream-collide over
//    a packed (case, block) index, then the tail (outlet speed, support-site
//    re-push, boundaries);
//  - when the next step will advance the same cases (none finishes, none
//    can be admitted), the next step's bulk stream-collide is launched on a
//    second stream as soon as this step's lattice phase is done, so it runs
//    beside this step's coupling; the next step then runs only the tail;
//  - the coupling loop (ObjectsClass::objectKernel, Objects.cpp:26-60) runs
//    in lockstep: the device work of every case still subiterating is one
//    batched launch per kernel, marker data crosses the bus in one packed
//    copy each way, and the FP64 FEM runs on host threads, one case each;
//  - one batched spread.
// A finished case frees its slot and memory at the end of the step.  Each
// case goes through the same kernels' arithmetic as life_case_step, so its
// results do not depend on the cases beside it.
//
// Compiled with -ffp-contract=off.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "case_impl.h"

using namespace life;

namespace {

using Clock = std::chrono::steady_clock;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return 2;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

struct Request {
    int id = 0;
    life_params p{};
    std::string geometry;
    int nsteps = 0;
    bool keep_fields = false;
    life_request_info info{};
    std::string error;
    std::vector<double> rho, u, mpos;
};

struct Slot {
    Request* req = nullptr;
    life_case* c = nullptr;
    int moff = 0;     // first marker in the packed staging arrays
    int fem_off = 0;  // first flexible body in the engine's device-FEM table
};

// Device footprint of a case: two population buffers, force_ibm, the
// initial rho/u, the spread's per-site counters, plus a margin for the
// boundary and marker arrays.
double case_bytes(const life_params& p, int prec) {
    const double N = (double)p.Nx * p.Ny;
    const double S = std::ceil(N / 256.0) * 256.0;
    const double T = prec == LIFE_FP64 ? 8.0 : 4.0;
    return S * (18.0 * T + 16.0 + 24.0) + N * 8.0 + (1 << 20);
}

template <typename T>
T* host_pinned(size_t n) {
    void* p = nullptr;
    LIFE_HIP(hipHostMalloc(&p, (n ? n : 1) * sizeof(T), hipHostMallocDefault));
    return static_cast<T*>(p);
}

// A device array with a pinned host mirror, grown on demand.
template <typename T>
struct Mirror {
    T* d = nullptr;
    T* h = nullptr;
    size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        release();
        cap = std::max<size_t>(n, 2 * cap);
        d = dalloc<T>(cap);
        h = host_pinned<T>(cap);
    }
    void release() {
        dfree(d);
        if (h) (void)hipHostFree(h);
        d = nullptr;
        h = nullptr;
        cap = 0;
    }
    void upload(size_t n, hipStream_t s) {
        if (n) LIFE_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(size_t n, hipStream_t s) {
        if (n) LIFE_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Optional phase timing (environment LIFE_TIMING=1): synchronises the
// stream at each phase boundary, so it is for diagnosis only.
struct EngineTimes {
    bool on = std::getenv("LIFE_TIMING") != nullptr;
    // admit+rebuild, lattice, host recompute, device recompute+interp+copies,
    // host FEM, spread, retire
    double t[7] = {0, 0, 0, 0, 0, 0, 0};
    double steps = 0, subits = 0;
    Clock::time_point last;
    void start(hipStream_t st) {
        if (!on) return;
        LIFE_HIP(hipStreamSynchronize(st));
        last = Clock::now();
    }
    void lap(hipStream_t st, int k) {
        if (!on) return;
        LIFE_HIP(hipStreamSynchronize(st));
        const auto now = Clock::now();
        t[k] += std::chrono::duration<double, std::milli>(now - last).count();
        last = now;
    }
    void report() const {
        if (!on || steps == 0) return;
        std::fprintf(stderr, "life server timing over %.0f engine steps (%.0f subiterations), ms/step: admit %.3f "
                     "lattice %.3f recompute-host %.3f device %.3f fem %.3f spread %.3f retire %.3f\n", steps, subits,
                     t[0] / steps, t[1] / steps, t[2] / steps, t[3] / steps, t[4] / steps, t[5] / steps, t[6] / steps);
    }
};

// Case lists of the lattice phase, fixed while the set of cases is.
struct LatticePlan {
    std::vector<CaseBlock> blocks[2];  // BGK, central moments
    std::vector<int> conv, ibm[2], bc[2];
    int max_sites[2] = {0, 0}, max_nbc[2] = {0, 0};
    int max_ny = 0;  // largest Ny of the cases with a convective outlet
};

}  // namespace

struct life_server {
    life_server_config cfg{};
    hipStream_t stream = nullptr;
    Clock::time_point t0;
    int next_id = 0;
    std::map<int, std::unique_ptr<Request>> requests;
    std::deque<Request*> queue;
    std::vector<Slot> slots;
    double used_bytes = 0.0;
    bool dirty = true;
    // device tables
    Mirror<CaseDev> cases;
    Mirror<CaseBlock> blocks[2];
    Mirror<int> lists;  // static lists of the lattice phase, back to back
    Mirror<int> dyn;    // per-subiteration list
    Mirror<CaseSys> sys;
    Mirror<int> outside;
    Mirror<double> stage_in, stage_out;
    LatticePlan plan;
    int off_conv = 0, off_ibm[2] = {0, 0}, off_bc[2] = {0, 0}, off_ibm_all = 0, n_ibm_all = 0;
    int total_markers = 0, max_markers = 0, max_bodies = 0;
    hipGraph_t graph = nullptr;  // lattice tail
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_bulk = nullptr;  // bulk stream-collide
    hipGraphExec_t exec_bulk = nullptr;
    // early bulk stream-collide of the next step (environment
    // LIFE_SERVE_EARLY): 0 off, 1 launched when the lattice phase is
    // queued (beside the whole coupling), 2 launched when the FEM of the
    // first subiteration starts (which paid while the FEM ran on host
    // threads and left the device idle; with the device FEM, capacity run of
    // 256 cases x 2000 steps: 207 000-208 100 cases/hour with 1, 181 500-
    // 182 800 with 2)
    int early = 1;
    bool prepushed = false;         // this step's bulk stream-collide already ran on stream2
    hipStream_t stream2 = nullptr;  // shared per device (masked_stream), not owned
    hipEvent_t ev_lattice = nullptr, ev_bulk = nullptr;
    // the last upload of the bulk case table (cases_bulk) from its
    // page-locked host copy; the host rewrites that copy only after it
    hipEvent_t ev_bulk_up = nullptr;
    bool bulk_up_pending = false;
    Mirror<CaseDev> cases_bulk;     // descriptors the bulk stream-collide reads
    double early_steps = 0.0;
    // FEM on the device (FP32-shifted populations): every case's flexible
    // bodies back to back, and per subiteration the list of the bodies of
    // the cases still subiterating, their step or relaxation factor and
    // their residual terms
    bool fem_device = false;
    size_t fem_lds = 0;
    Mirror<FemDevBody> fem_bodies;
    Mirror<int> fem_list, fem_list2;
    Mirror<double> fem_param, fem_out;
    EngineTimes times;
    // statistics
    double site_updates = 0.0, steps = 0.0, busy = 0.0, captures = 0.0, peak_running = 0.0, peak_bytes = 0.0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void drop_graph(life_server* s) {
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->exec_bulk) (void)hipGraphExecDestroy(s->exec_bulk);
    if (s->graph_bulk) (void)hipGraphDestroy(s->graph_bulk);
    s->exec = nullptr;
    s->graph = nullptr;
    s->exec_bulk = nullptr;
    s->graph_bulk = nullptr;
}

void fail(life_server* s, Slot& sl, const std::string& msg) {
    sl.req->info.status = LIFE_REQ_FAILED;
    sl.req->error = msg;
    (void)s;
}

// The head of the queue fits a free slot and the budget.
bool can_admit(const life_server* s) {
    if (s->queue.empty() || (int)s->slots.size() >= s->cfg.max_cases) return false;
    const double bytes = case_bytes(s->queue.front()->p, s->cfg.precision);
    return s->slots.empty() || s->used_bytes + bytes <= s->cfg.hbm_budget;
}

// Admit queued requests in arrival order while a slot and budget remain.
void admit(life_server* s) {
    while (can_admit(s)) {
        Request* r = s->queue.front();
        const double bytes = case_bytes(r->p, s->cfg.precision);
        s->queue.pop_front();
        r->info.t_admit = s->now();
        r->info.bytes = bytes;
        life_case* c = nullptr;
        if (r->p.womersley != 0.0 && (r->p.gravityX != 0.0 || r->p.gravityY != 0.0)) {
            // force_xy per site: the batched kernels take a uniform force
            r->info.status = LIFE_REQ_FAILED;
            r->error = "WOMERSLEY forcing with gravity runs on the single-case path (life_case) only";
            r->info.t_finish = s->now();
            continue;
        }
        if (life_case_create(&r->p, r->geometry.c_str(), s->cfg.precision, s->cfg.device, &c)) {
            r->info.status = LIFE_REQ_FAILED;
            r->error = life_last_error();
            r->info.t_finish = s->now();
            continue;
        }
        // force_ibm is zero from creation and no spread has run, so the
        // first spread needs no full clear
        c->g->dense_clear_next = false;
        c->g->check_every = s->cfg.check_every == 0 ? 100 : (s->cfg.check_every > 0 ? s->cfg.check_every : 0);
        r->info.status = LIFE_REQ_RUNNING;
        r->info.markers = c->nm;
        s->slots.push_back(Slot{r, c, 0});
        s->used_bytes += bytes;
        s->dirty = true;
    }
    s->peak_running = std::max(s->peak_running, (double)s->slots.size());
    s->peak_bytes = std::max(s->peak_bytes, s->used_bytes);
}

// Case lists and packed indices for the current set of cases.
void rebuild(life_server* s) {
    LatticePlan& P = s->plan;
    P = LatticePlan{};
    s->total_markers = s->max_markers = s->max_bodies = 0;
    std::vector<int> ibm_all;
    for (int q = 0; q < (int)s->slots.size(); q++) {
        Slot& sl = s->slots[q];
        life_grid* g = sl.c->g;
        const int cm = g->g.cm ? 1 : 0;
        const int bs = s->cfg.precision == LIFE_FP64 ? kServeBlock64 : kServeBlock32;
        const long long nb = (g->g.N + bs - 1) / bs;
        for (long long b = 0; b < nb; b++) P.blocks[cm].push_back(CaseBlock{q, (int)b});
        if (g->convective) {
            P.conv.push_back(q);
            P.max_ny = std::max(P.max_ny, g->g.Ny);
        }
        if (g->nbc > 0) {
            P.bc[cm].push_back(q);
            P.max_nbc[cm] = std::max(P.max_nbc[cm], g->nbc);
        }
        if (g->md.n > 0) {
            P.ibm[cm].push_back(q);
            P.max_sites[cm] = std::max(P.max_sites[cm], g->sc.cap);
        }
        if (sl.c->hasIBM) ibm_all.push_back(q);
        sl.moff = s->total_markers;
        s->total_markers += sl.c->nm;
        s->max_markers = std::max(s->max_markers, sl.c->nm);
        s->max_bodies = std::max(s->max_bodies, g->nbodies);
    }
    for (int k = 0; k < 2; k++) {
        s->blocks[k].reserve(P.blocks[k].size());
        std::copy(P.blocks[k].begin(), P.blocks[k].end(), s->blocks[k].h);
        s->blocks[k].upload(P.blocks[k].size(), s->stream);
    }
    std::vector<int> all;
    auto put = [&](const std::vector<int>& v) {
        const int off = (int)all.size();
        all.insert(all.end(), v.begin(), v.end());
        return off;
    };
    s->off_conv = put(P.conv);
    for (int k = 0; k < 2; k++) {
        s->off_ibm[k] = put(P.ibm[k]);
        s->off_bc[k] = put(P.bc[k]);
    }
    s->off_ibm_all = put(ibm_all);
    s->n_ibm_all = (int)ibm_all.size();
    s->lists.reserve(all.size());
    std::copy(all.begin(), all.end(), s->lists.h);
    s->lists.upload(all.size(), s->stream);
    s->cases.reserve(s->slots.size());
    s->cases_bulk.reserve(s->slots.size());
    s->stage_in.reserve(4 * (size_t)s->total_markers);
    s->stage_out.reserve(3 * (size_t)s->total_markers);
    s->dyn.reserve(s->slots.size());
    // device FEM when every flexible case has it (FP32-shifted populations)
    std::vector<FemDevBody> fb;
    bool all_dev = true, any_flex = false;
    s->fem_lds = 0;
    for (Slot& sl : s->slots) {
        life_case* c = sl.c;
        if (!c->hasFlex) continue;
        any_flex = true;
        if (!c->fdev.ok()) {
            all_dev = false;
            continue;
        }
        sl.fem_off = (int)fb.size();
        fb.insert(fb.end(), c->fdev.host.begin(), c->fdev.host.end());
        s->fem_lds = std::max(s->fem_lds, c->fdev.lds);
    }
    s->fem_device = any_flex && all_dev;
    if (s->fem_device) {
        s->fem_bodies.reserve(fb.size());
        std::copy(fb.begin(), fb.end(), s->fem_bodies.h);
        s->fem_bodies.upload(fb.size(), s->stream);
        s->fem_list.reserve(fb.size());
        s->fem_list2.reserve(fb.size());
        s->fem_param.reserve(fb.size());
        s->fem_out.reserve(3 * fb.size());
    }
    LIFE_HIP(hipStreamSynchronize(s->stream));  // host list buffers are reused next time
    drop_graph(s);
    s->dirty = false;
}

// Descriptor of every case for this step (ramp of step t+1, current buffers).
void fill_table(life_server* s) {
    for (int q = 0; q < (int)s->slots.size(); q++) {
        const Slot& sl = s->slots[q];
        life_grid* g = sl.c->g;
        CaseDev& d = s->cases.h[q];
        g->g = step_scalars(*g, sl.c->t + 1);  // this step's force_xy (WOMERSLEY) for later host-side use
        d.g = g->g;
        d.src = g->f[g->cur];
        d.dst = g->f[1 - g->cur];
        d.rho0 = g->rho0;
        d.u0 = g->u0;
        d.fibm = g->fibm;
        d.explicit_state = g->explicit_state ? 1 : 0;
        d.convective = g->convective ? 1 : 0;
        d.bc = BCArgs{g->bc, g->nbc, g->u_in, g->rho_in, g->delU, ramp_coefficient(g->p, g->cs.Dt, sl.c->t + 1)};
        d.ucol = g->ucol;
        d.m = g->md;
        d.sc = g->sc;
        d.nbodies = g->nbodies;
        d.moff = sl.moff;
        d.eps = EpsDev{g->eps_nsys, g->eps_first, g->eps_n, g->eps_off, g->eps_outside, g->eps_band, g->eps_y};
    }
    s->cases.upload(s->slots.size(), s->stream);
}

// GridClass::lbmKernel of every case (Grid.cpp:36-100), same stage order
// as lbm_step: outlet speed, stream-collide, support sites, boundaries.
// Bulk stream-collide of every case (descriptors: cases_bulk).
void launch_bulk(life_server* s, hipStream_t st) {
    const LatticePlan& P = s->plan;
    const bool f64 = s->cfg.precision == LIFE_FP64;
    for (int k = 0; k < 2; k++) {
        const int nb = (int)P.blocks[k].size();
        if (f64) launch_push64_batch(k == 1, s->cases_bulk.d, s->blocks[k].d, nb, st);
        else launch_push32_batch(k == 1, s->cases_bulk.d, s->blocks[k].d, nb, st);
    }
    LIFE_HIP(hipGetLastError());
}

// Lattice tail: outlet speed (reads the incoming populations only, so it
// may follow the bulk kernel), support-site re-push, boundaries.
void launch_tail(life_server* s, hipStream_t st) {
    const LatticePlan& P = s->plan;
    const bool f64 = s->cfg.precision == LIFE_FP64;
    const int* L = s->lists.d;
    if (!P.conv.empty()) {
        if (f64)
            launch_convective_batch<double>(s->cases.d, L + s->off_conv, (int)P.conv.size(), P.max_ny, true, st);
        else
            launch_convective_batch<float>(s->cases.d, L + s->off_conv, (int)P.conv.size(), P.max_ny, false, st);
    }
    for (int k = 0; k < 2; k++) {
        const int n = (int)P.ibm[k].size();
        if (f64) launch_support_push64_batch(k == 1, s->cases.d, L + s->off_ibm[k], n, P.max_sites[k], st);
        else launch_support_push32_batch(k == 1, s->cases.d, L + s->off_ibm[k], n, P.max_sites[k], st);
    }
    for (int k = 0; k < 2; k++) {
        const int n = (int)P.bc[k].size();
        if (f64) launch_bc_batch<double>(k == 1, s->cases.d, L + s->off_bc[k], n, P.max_nbc[k], st);
        else launch_bc_batch<float>(k == 1, s->cases.d, L + s->off_bc[k], n, P.max_nbc[k], st);
    }
    LIFE_HIP(hipGetLastError());
}

// Run body(stream) on `on`: directly, or as a graph captured on the
// engine stream the first time and launched on `on`.
template <typename Body>
void run_phase(life_server* s, hipGraph_t& graph, hipGraphExec_t& exec, hipStream_t on, Body&& body) {
    if (!s->cfg.use_graph) {
        body(on);
        return;
    }
    if (!exec) {
        LIFE_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        try {
            body(s->stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        LIFE_HIP(hipStreamEndCapture(s->stream, &graph));
        LIFE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        s->captures += 1.0;
    }
    LIFE_HIP(hipGraphLaunch(exec, on));
}

// Wait until the last upload of the bulk case table has read its host copy.
void bulk_table_free(life_server* s) {
    if (!s->bulk_up_pending) return;
    LIFE_HIP(hipEventSynchronize(s->ev_bulk_up));
    s->bulk_up_pending = false;
}

void lattice_phase(life_server* s) {
    if (s->prepushed) {
        // the bulk stream-collide of this step ran on stream2 beside the
        // previous step's coupling
        LIFE_HIP(hipStreamWaitEvent(s->stream, s->ev_bulk, 0));
        s->prepushed = false;
    } else {
        bulk_table_free(s);
        std::copy(s->cases.h, s->cases.h + s->slots.size(), s->cases_bulk.h);
        s->cases_bulk.upload(s->slots.size(), s->stream);
        LIFE_HIP(hipEventRecord(s->ev_bulk_up, s->stream));
        s->bulk_up_pending = true;
        run_phase(s, s->graph_bulk, s->exec_bulk, s->stream, [&](hipStream_t st) { launch_bulk(s, st); });
    }
    run_phase(s, s->graph, s->exec, s->stream, [&](hipStream_t st) { launch_tail(s, st); });
}

// Next step's bulk stream-collide on stream2, beside this step's coupling,
// when the next step advances exactly the cases of this one: none finishes
// now and none can be admitted.  It reads the populations this step's
// lattice phase wrote and writes the other buffer, which the coupling does
// not touch; the support sites and boundaries are redone by the next
// step's tail after the spread, as in the plain order.
void launch_bulk_early(life_server* s) {
    if (s->prepushed || s->slots.empty() || can_admit(s)) return;
    for (const Slot& sl : s->slots)
        if (sl.c->t >= sl.req->nsteps || sl.req->info.status == LIFE_REQ_FAILED) return;
    bulk_table_free(s);
    for (int q = 0; q < (int)s->slots.size(); q++) {
        const life_case* c = s->slots[q].c;
        const life_grid* g = c->g;
        CaseDev& d = s->cases_bulk.h[q];
        d = s->cases.h[q];
        d.g = step_scalars(*g, c->t + 1);
        d.src = g->f[g->cur];
        d.dst = g->f[1 - g->cur];
        d.explicit_state = 0;
    }
    // The table goes up on stream2 behind the previous bulk kernel (the
    // last reader of the old table) and ahead of the wait for this step's
    // lattice phase: it is in place by the time that phase ends, and the
    // bulk kernel starts on the event alone.  (A profiled run showed about
    // 40 us between the boundaries and the bulk kernel with the copy after
    // the wait; unprofiled capacity is the same within the spread.)
    s->cases_bulk.upload(s->slots.size(), s->stream2);
    LIFE_HIP(hipEventRecord(s->ev_bulk_up, s->stream2));
    s->bulk_up_pending = true;
    LIFE_HIP(hipEventRecord(s->ev_lattice, s->stream));
    LIFE_HIP(hipStreamWaitEvent(s->stream2, s->ev_lattice, 0));
    run_phase(s, s->graph_bulk, s->exec_bulk, s->stream2, [&](hipStream_t st) { launch_bulk(s, st); });
    LIFE_HIP(hipEventRecord(s->ev_bulk, s->stream2));
    s->prepushed = true;
    s->early_steps += 1.0;
}

// Upload the host marker positions/velocities of the listed cases.
void markers_in(life_server* s, const std::vector<int>& list) {
    if (list.empty()) return;
    for (int q : list) {
        const Slot& sl = s->slots[q];
        const life_case* c = sl.c;
        double* o = s->stage_in.h + 4 * (size_t)sl.moff;
        for (int k = 0; k < c->nm; k++) {
            o[4 * k] = c->pos[2 * k];
            o[4 * k + 1] = c->pos[2 * k + 1];
            o[4 * k + 2] = c->vel[2 * k];
            o[4 * k + 3] = c->vel[2 * k + 1];
        }
    }
    s->stage_in.upload(4 * (size_t)s->total_markers, s->stream);
    std::copy(list.begin(), list.end(), s->dyn.h);
    s->dyn.upload(list.size(), s->stream);
    launch_markers_in_batch(s->cases.d, s->dyn.d, (int)list.size(), s->max_markers, s->stage_in.d, s->stream);
    LIFE_HIP(hipGetLastError());
}

void launch_interpolate_batch_any(bool f64, life_server* s, const int* list, int n) {
    if (f64) launch_interpolate_batch<double>(s->cases.d, list, n, s->max_markers, s->stream);
    else launch_interpolate_batch<float>(s->cases.d, list, n, s->max_markers, s->stream);
    LIFE_HIP(hipGetLastError());
}

// fn(slot) for the listed cases on host threads.  Cases are independent
// (disjoint host state); a case whose host work throws is marked failed
// and leaves the lockstep loop.
template <typename Fn>
void for_cases(life_server* s, std::vector<int>& list, Fn&& fn) {
    const int n = (int)list.size();
    std::vector<std::string> err(n);
#pragma omp parallel for schedule(dynamic, 1) if (n > 1)
    for (int k = 0; k < n; k++) {
        try {
            fn(s->slots[list[k]]);
        } catch (const std::exception& e) {
            err[k] = e.what();
            if (err[k].empty()) err[k] = "host error";
        }
    }
    std::vector<int> keep;
    for (int k = 0; k < n; k++) {
        if (err[k].empty()) keep.push_back(list[k]);
        else fail(s, s->slots[list[k]], err[k]);
    }
    list.swap(keep);
}

// The FEM half of recomputeObjectVals on the device for the active cases:
// predictor at subiteration 0, else the Aitken relaxation (factor per case,
// Objects.cpp:179-210).  Returns the number of listed bodies (the list
// stays in fem_list for the solve).
int fem_recompute_device(life_server* s, const std::vector<int>& active) {
    int n = 0;
    bool predict = false;
    for (int q : active) {
        life_case* c = s->slots[q].c;
        predict = c->subIt == 0;  // the same for every case still subiterating
        if (!predict) case_aitken(c);
        for (int b = 0; b < c->fdev.nb; b++) {
            s->fem_list.h[n] = s->slots[q].fem_off + b;
            s->fem_param.h[n] = predict ? (double)c->t : c->relax;
            n++;
        }
    }
    s->fem_list.upload(n, s->stream);
    s->fem_param.upload(n, s->stream);
    if (predict)
        launch_fem_predict(s->fem_bodies.d, s->fem_list.d, n, 0, s->fem_param.d, s->fem_lds, s->stream);
    else
        launch_fem_relax(s->fem_bodies.d, s->fem_list.d, n, 0.0, s->fem_param.d, s->fem_lds, s->stream);
    LIFE_HIP(hipGetLastError());
    return n;
}

// femKernel of the active cases on the device, then each case's
// subiteration residual (case_fem's sums, in body order).  Epsilon systems
// found not to be banded get the dense solve and their cases' FEM runs
// again with it.
void fem_solve_device(life_server* s, const std::vector<int>& active, const std::vector<CaseSys>& sysl, int nfb) {
    hipStream_t st = s->stream;
    launch_fem_dynamic(s->fem_bodies.d, s->fem_list.d, nfb, s->fem_lds, false, st);
    // the device has room beside the FEM's few workgroups: the next step's
    // bulk stream-collide starts now
    if (s->early == 2) launch_bulk_early(s);
    // each body's terms go straight into the page-locked host array
    launch_fem_outs(s->fem_bodies.d, s->fem_list.d, nfb, s->fem_out.h, st);
    LIFE_HIP(hipGetLastError());
    LIFE_HIP(hipStreamSynchronize(st));
    std::vector<int> redo;  // slots whose epsilon changed
    for (size_t k = 0; k < sysl.size(); k++)
        if (s->outside.h[k]) {
            life_case* c = s->slots[sysl[k].c].c;
            case_solve_eps_dense(c, c->g->eps_first_h[sysl[k].k], c->g->eps_n_h[sysl[k].k]);
            LIFE_HIP(hipStreamSynchronize(c->g->stream));
            if (redo.empty() || redo.back() != sysl[k].c) redo.push_back(sysl[k].c);
        }
    if (!redo.empty()) {
        int n = 0;
        for (int q : redo)
            for (int b = 0; b < s->slots[q].c->fdev.nb; b++) s->fem_list2.h[n++] = s->slots[q].fem_off + b;
        s->fem_list2.upload(n, st);
        launch_fem_dynamic(s->fem_bodies.d, s->fem_list2.d, n, s->fem_lds, true, st);
        launch_fem_outs(s->fem_bodies.d, s->fem_list.d, nfb, s->fem_out.h, st);
        LIFE_HIP(hipGetLastError());
        LIFE_HIP(hipStreamSynchronize(st));
    }
    s->times.lap(st, 4);
    s->times.subits += 1;
    int k = 0;
    for (int q : active) {
        life_case* c = s->slots[q].c;
        double res = 0.0, num = 0.0, den = 0.0;
        for (int b = 0; b < c->fdev.nb; b++, k++) {
            res += s->fem_out.h[3 * k];
            num += s->fem_out.h[3 * k + 1];
            den += s->fem_out.h[3 * k + 2];
        }
        c->subRes = std::sqrt(res) / (c->p.ref_L * std::sqrt((double)c->simDOFs));
        c->subNum = num;
        c->subDen = den;
        c->subIt++;
    }
}

// ObjectsClass::objectKernel (Objects.cpp:26-60) of every case with bodies.
void coupling_phase(life_server* s) {
    std::vector<int> flex;
    for (int q = 0; q < (int)s->slots.size(); q++) {
        life_case* c = s->slots[q].c;
        if (!c->hasIBM) continue;
        c->subIt = 0;
        c->forces_on_host = false;
        if (c->hasFlex) flex.push_back(q);
    }
    if (s->n_ibm_all == 0) return;
    const bool f64 = s->cfg.precision == LIFE_FP64;
    hipStream_t st = s->stream;
    std::vector<int> active = flex;
    bool first = true;
    std::vector<CaseSys> sysl;
    std::vector<std::pair<int, int>> dense;  // (slot, first marker of a system) needing the dense path
    std::vector<int> dense_n;
    const bool devfem = s->fem_device;
    int nfb = 0;  // bodies in this subiteration's device-FEM list
    while (true) {
        sysl.clear();
        dense.clear();
        dense_n.clear();
        if (!active.empty()) {
            // recomputeObjectVals: predictor / relaxation (host or device
            // FEM), then supports, ds and epsilon on the device
            if (devfem) {
                nfb = fem_recompute_device(s, active);
                std::copy(active.begin(), active.end(), s->dyn.h);
                s->dyn.upload(active.size(), st);
            } else {
                for_cases(s, active, [&](Slot& sl) { case_recompute_host(sl.c); });
                s->times.lap(st, 2);
                markers_in(s, active);
            }
            const int na = (int)active.size();
            int max_n = 0;
            bool any_uni = false;
            for (int q : active) {
                life_case* c = s->slots[q].c;
                life_grid* g = c->g;
                if (c->p.uni_epsilon) {
                    any_uni = true;
                    dense.push_back({q, 0});
                    dense_n.push_back(c->nm);
                    continue;
                }
                for (int k = 0; k < g->eps_nsys; k++) {
                    sysl.push_back(CaseSys{q, k});
                    max_n = std::max(max_n, g->eps_n_h[k]);
                }
            }
            if (!sysl.empty()) {
                s->sys.reserve(sysl.size());
                s->outside.reserve(sysl.size());
                std::copy(sysl.begin(), sysl.end(), s->sys.h);
                s->sys.upload(sysl.size(), st);
            }
            // one workgroup per flexible body for supports, spacing and the
            // banded epsilon system when every case can take that path
            // FP32-shifted cases solve their banded systems by cyclic
            // reduction, as life_case does
            const bool cr = s->cfg.precision == LIFE_FP32_SHIFTED;
            // the out-of-band flags are gathered straight into the
            // page-locked host array, read after the next synchronisation
            const bool fused = !any_uni && !sysl.empty() &&
                               launch_recompute_fused_batch(s->cases.d, s->sys.d, (int)sysl.size(), max_n,
                                                            s->outside.h, cr, st);
            if (!fused) {
                launch_find_support_batch(s->cases.d, s->dyn.d, na, s->max_markers, 1, st);
                launch_compute_ds_batch(s->cases.d, s->dyn.d, na, s->max_bodies, 1, st);
                if (!sysl.empty())
                    launch_eps_band_batch(s->cases.d, s->sys.d, (int)sysl.size(), max_n, s->outside.h, cr, st);
            }
        }
        // ibmKernelInterp: every case with bodies on the first pass, then
        // the cases still subiterating
        if (first) {
            launch_interpolate_batch_any(f64, s, s->lists.d + s->off_ibm_all, s->n_ibm_all);
        } else {
            launch_interpolate_batch_any(f64, s, s->dyn.d, (int)active.size());
        }
        first = false;
        if (active.empty()) break;
        if (devfem) {
            fem_solve_device(s, active, sysl, nfb);
            std::vector<int> next;
            for (int q : active)
                if (case_subiterate(s->slots[q].c)) next.push_back(q);
            active.swap(next);
            if (active.empty()) break;
            continue;
        }
        launch_markers_out_batch(s->cases.d, s->dyn.d, (int)active.size(), s->max_markers, s->stage_out.d, st);
        LIFE_HIP(hipGetLastError());
        s->stage_out.download(3 * (size_t)s->total_markers, st);
        LIFE_HIP(hipStreamSynchronize(st));
        for (size_t k = 0; k < sysl.size(); k++)
            if (s->outside.h[k]) {
                life_grid* g = s->slots[sysl[k].c].c->g;
                dense.push_back({sysl[k].c, g->eps_first_h[sysl[k].k]});
                dense_n.push_back(g->eps_n_h[sysl[k].k]);
            }
        for (int q : active) {
            const Slot& sl = s->slots[q];
            life_case* c = sl.c;
            const double* o = s->stage_out.h + 3 * (size_t)sl.moff;
            for (int k = 0; k < c->nm; k++) {
                c->force[2 * k] = o[3 * k];
                c->force[2 * k + 1] = o[3 * k + 1];
                c->eps[k] = o[3 * k + 2];
            }
        }
        // systems that are not banded: computeEpsilon's dense solve
        for (size_t k = 0; k < dense.size(); k++) {
            life_case* c = s->slots[dense[k].first].c;
            case_solve_eps_dense(c, dense[k].second, dense_n[k]);
            LIFE_HIP(hipStreamSynchronize(c->g->stream));
        }
        // femKernel on host threads, then the convergence test per case; the
        // device is idle meanwhile, so the next step's bulk kernel starts now
        if (s->early == 2) launch_bulk_early(s);
        s->times.lap(st, 3);
        s->times.subits += 1;
        for_cases(s, active, [&](Slot& sl) { case_fem(sl.c); });
        s->times.lap(st, 4);
        std::vector<int> next;
        for (int q : active)
            if (case_subiterate(s->slots[q].c)) next.push_back(q);
        active.swap(next);
        if (active.empty()) break;
    }
    // ibmKernelSpread of every case with bodies
    launch_spread_batch(s->cases.d, s->lists.d + s->off_ibm_all, s->n_ibm_all, s->max_markers, st);
    LIFE_HIP(hipGetLastError());
    // the last FEM solve moved the markers: the device takes the final
    // positions (as life_case_step does after its spread); with the FEM on
    // the device they are there already, and the host copies lag
    if (!devfem) markers_in(s, flex);
    for (int q : flex) {
        life_case* c = s->slots[q].c;
        c->forces_on_host = !devfem;
        c->host_stale = devfem;
        if (c->subIt == kMaxSubIt) fail(s, s->slots[q], "Subiteration scheme hit 20 iterations");
    }
}

// Results of a finished case, then its slot and memory go back.
void retire(life_server* s, Slot& sl) {
    Request* r = sl.req;
    life_case* c = sl.c;
    if (r->info.status != LIFE_REQ_FAILED) {
        try {
            case_sync_host(c);
            if (c->hasIBM) check_flags(c->g);
            if (life_case_total_force(c, r->info.total_force)) throw std::runtime_error(life_last_error());
            if (c->nm > 0) {
                const int k = c->nm - 1;
                r->info.tip[0] = c->pos[2 * k];
                r->info.tip[1] = c->pos[2 * k + 1];
                r->info.tip[2] = c->vel[2 * k];
                r->info.tip[3] = c->vel[2 * k + 1];
            }
            if (r->keep_fields) {
                const size_t N = (size_t)c->g->g.N;
                r->rho.resize(N);
                r->u.resize(2 * N);
                if (life_grid_download(c->g, nullptr, r->rho.data(), r->u.data(), nullptr))
                    throw std::runtime_error(life_last_error());
                r->mpos.assign(c->pos.begin(), c->pos.end());
            }
            r->info.status = LIFE_REQ_DONE;
        } catch (const HipError& e) {
            fail(s, sl, std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        } catch (const std::exception& e) {
            fail(s, sl, e.what());
        }
    }
    r->info.t_finish = s->now();
    life_case_destroy(c);
    sl.c = nullptr;
    s->used_bytes -= r->info.bytes;
}

void engine_step(life_server* s) {
    s->times.start(s->stream);
    // after an early bulk launch the set of cases stays as it was: only a
    // failure in the last coupling can have freed a slot, and a case
    // admitted now would miss its bulk stream-collide, so admission waits
    // one step
    if (s->prepushed && s->slots.empty()) s->prepushed = false;
    if (!s->prepushed) admit(s);
    if (s->slots.empty()) return;
    const double t_begin = s->now();
    if (s->dirty) rebuild(s);
    fill_table(s);
    s->times.lap(s->stream, 0);
    lattice_phase(s);
    s->times.lap(s->stream, 1);
    for (Slot& sl : s->slots) {
        life_case* c = sl.c;
        c->t++;
        c->g->cur = 1 - c->g->cur;
        c->g->explicit_state = false;
        s->site_updates += (double)c->g->g.N;
    }
    if (s->early == 1) launch_bulk_early(s);
    coupling_phase(s);
    // the blow-up test of the cases due (GridClass::writeInfo's NaN scan
    // at the tinfo cadence); it only reads the populations, so a case's
    // neighbours are unaffected whatever it finds
    for (Slot& sl : s->slots) {
        life_grid* g = sl.c->g;
        if (g->check_every > 0 && sl.c->t % g->check_every == 0 && sl.req->info.status != LIFE_REQ_FAILED)
            blowup_scan(g, sl.c->t, s->stream);
    }
    // no FEM this step (or it failed): the next bulk kernel still runs
    // beside the spread and the step's end
    if (s->early == 2) launch_bulk_early(s);
    LIFE_HIP(hipStreamSynchronize(s->stream));
    for (Slot& sl : s->slots) {
        if (sl.c->g->blow_t < 0) continue;
        try {
            blowup_check(sl.c->g, true);
        } catch (const std::exception& e) {
            fail(s, sl, e.what());  // retired below, its slot and memory freed
        }
    }
    // A case that retires below may still be read by the early bulk kernel
    // (launched only while no case finishes this step, so after a failure):
    // the second stream is drained first.  Otherwise the host goes on to
    // the next step while that kernel runs, and the next step's lattice
    // phase waits for it on the device (draining it every step held the
    // host's launches of the next tail until the bulk kernel had ended).
    // The bulk table's host copy has its own event (bulk_table_free).
    bool retiring = false;
    for (const Slot& sl : s->slots)
        retiring = retiring || sl.c->t >= sl.req->nsteps || sl.req->info.status == LIFE_REQ_FAILED;
    if (retiring && s->prepushed) LIFE_HIP(hipStreamSynchronize(s->stream2));
    s->times.lap(s->stream, 5);
    s->times.steps += 1;
    s->steps += 1.0;
    for (int q = (int)s->slots.size() - 1; q >= 0; q--) {
        Slot& sl = s->slots[q];
        sl.req->info.steps_done = sl.c->t;
        sl.req->info.subits = sl.c->subIt;
        if (sl.c->t >= sl.req->nsteps || sl.req->info.status == LIFE_REQ_FAILED) {
            retire(s, sl);
            s->slots.erase(s->slots.begin() + q);
            s->dirty = true;
        }
    }
    s->times.lap(s->stream, 6);
    s->busy += s->now() - t_begin;
}

int submit(life_server* s, const life_params* p, const char* geometry, int nsteps, int keep_fields,
           double t_arrival) {
    if (!p) throw std::runtime_error("null params");
    if (nsteps < 1) throw std::runtime_error("nsteps must be positive");
    std::unique_ptr<Request> r(new Request());
    r->id = s->next_id++;
    r->p = *p;
    r->geometry = geometry ? geometry : "";
    r->nsteps = nsteps;
    r->keep_fields = keep_fields != 0;
    r->info.status = LIFE_REQ_QUEUED;
    r->info.nsteps = nsteps;
    r->info.t_arrival = t_arrival;
    r->info.t_admit = r->info.t_finish = -1.0;
    r->info.sites = (double)p->Nx * p->Ny;
    s->queue.push_back(r.get());
    const int id = r->id;
    s->requests[id] = std::move(r);
    return id;
}

const Request* find(const life_server* s, int id) {
    auto it = s->requests.find(id);
    if (it == s->requests.end()) throw std::runtime_error("unknown request id");
    return it->second.get();
}

}  // namespace

extern "C" {

int life_server_create(const life_server_config* cfg, life_server** out) {
    *out = nullptr;
    life_server* s = new life_server();
    handle_acquire();
    const int rc = guarded([&] {
        if (!cfg) throw std::runtime_error("null config");
        if (cfg->precision != LIFE_FP64 && cfg->precision != LIFE_FP32_SHIFTED)
            throw std::runtime_error("unknown precision");
        if (cfg->max_cases < 1) throw std::runtime_error("max_cases must be positive");
        if (!(cfg->hbm_budget > 0.0)) throw std::runtime_error("hbm_budget must be positive");
        s->cfg = *cfg;
        LIFE_HIP(hipSetDevice(cfg->device));
        LIFE_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        // CUs the early bulk kernel leaves to the coupling kernels beside it
        // (environment LIFE_SERVE_COUPLING_CUS, default 32 of 256; capacity
        // run with the early launch at the lattice phase: 207 500-208 100
        // cases/hour with 32, 206 600-207 200 with 16, 190 700-191 800 with 0)
        const char* rc_env = std::getenv("LIFE_SERVE_COUPLING_CUS");
        s->stream2 = masked_stream(cfg->device, rc_env ? std::atoi(rc_env) : 32);
        LIFE_HIP(hipEventCreateWithFlags(&s->ev_lattice, hipEventDisableTiming));
        LIFE_HIP(hipEventCreateWithFlags(&s->ev_bulk, hipEventDisableTiming));
        LIFE_HIP(hipEventCreateWithFlags(&s->ev_bulk_up, hipEventDisableTiming));
        if (const char* e = std::getenv("LIFE_SERVE_EARLY")) s->early = std::atoi(e);
        s->t0 = Clock::now();
    });
    if (rc) {
        life_server_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int life_server_destroy(life_server* s) {
    if (!s) return 0;
    const int rc = guarded([&] {
        (void)hipSetDevice(s->cfg.device);
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        if (s->stream2) (void)hipStreamSynchronize(s->stream2);
        s->times.report();
        for (Slot& sl : s->slots) life_case_destroy(sl.c);
        s->slots.clear();
        drop_graph(s);
        s->cases.release();
        s->cases_bulk.release();
        for (auto& b : s->blocks) b.release();
        s->lists.release();
        s->dyn.release();
        s->sys.release();
        s->outside.release();
        s->stage_in.release();
        s->stage_out.release();
        s->fem_bodies.release();
        s->fem_list.release();
        s->fem_list2.release();
        s->fem_param.release();
        s->fem_out.release();
        if (s->ev_lattice) (void)hipEventDestroy(s->ev_lattice);
        if (s->ev_bulk) (void)hipEventDestroy(s->ev_bulk);
        if (s->ev_bulk_up) (void)hipEventDestroy(s->ev_bulk_up);
        if (s->stream) (void)hipStreamDestroy(s->stream);
    });
    delete s;
    handle_release();
    return rc;
}

int life_server_submit(life_server* s, const life_params* p, const char* geometry, int nsteps, int keep_fields,
                       int* id) {
    return guarded([&] { *id = submit(s, p, geometry, nsteps, keep_fields, s->now()); });
}

int life_server_step(life_server* s, int* running, int* queued) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(s->cfg.device));
        engine_step(s);
        if (running) *running = (int)s->slots.size();
        if (queued) *queued = (int)s->queue.size();
    });
}

int life_server_run(life_server* s, int n, const life_params* p, const char* const* geometry, const int* nsteps,
                    const double* arrival, int* ids) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(s->cfg.device));
        for (int k = 1; k < n; k++)
            if (arrival[k] < arrival[k - 1]) throw std::runtime_error("arrival times must be ascending");
        const double base = s->now();
        int next = 0;
        while (true) {
            const double t = s->now() - base;
            while (next < n && arrival[next] <= t) {
                ids[next] = submit(s, &p[next], geometry[next], nsteps[next], 0, base + arrival[next]);
                next++;
            }
            if (s->slots.empty() && s->queue.empty()) {
                if (next >= n) break;
                std::this_thread::sleep_for(std::chrono::duration<double>(arrival[next] - t));
                continue;
            }
            engine_step(s);
        }
    });
}

int life_server_info(const life_server* s, int id, life_request_info* out) {
    return guarded([&] { *out = find(s, id)->info; });
}

const char* life_server_request_error(const life_server* s, int id) {
    auto it = s->requests.find(id);
    return it == s->requests.end() ? "unknown request id" : it->second->error.c_str();
}

int life_server_fields(const life_server* s, int id, double* rho, double* u, double* marker_pos) {
    return guarded([&] {
        const Request* r = find(s, id);
        if (r->info.status != LIFE_REQ_DONE) throw std::runtime_error("request has not finished");
        if (!r->keep_fields) throw std::runtime_error("request was submitted without keep_fields");
        if (rho) std::copy(r->rho.begin(), r->rho.end(), rho);
        if (u) std::copy(r->u.begin(), r->u.end(), u);
        if (marker_pos) std::copy(r->mpos.begin(), r->mpos.end(), marker_pos);
    });
}

int life_server_release(life_server* s, int id) {
    return guarded([&] {
        auto it = s->requests.find(id);
        if (it == s->requests.end()) throw std::runtime_error("unknown request id");
        const int st = it->second->info.status;
        if (st != LIFE_REQ_DONE && st != LIFE_REQ_FAILED) throw std::runtime_error("request is still queued or running");
        s->requests.erase(it);
    });
}

int life_server_stats(const life_server* s, double out[7]) {
    out[0] = s->site_updates;
    out[1] = s->steps;
    out[2] = s->busy;
    out[3] = s->captures;
    out[4] = s->peak_running;
    out[5] = s->peak_bytes;
    out[6] = s->early_steps;
    return 0;
}

}  // extern "C"
