// Device lattice runtime and the layer-1 C ABI (life_grid_*, life_lbm_*,
// life_ibm_*).  Compiled with -ffp-contract=off: the derived scalars, the
// inlet profile and the ramp follow the reference's rounding.
#include "runtime.h"

#include "lu.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <vector>

namespace life {

namespace {
// Freed pinned blocks by size (the cache object is never destroyed, so no
// HIP call runs from static destructors; its blocks are freed when the
// last grid or server goes).
struct PinnedCache {
    std::mutex mu;
    std::map<size_t, std::vector<void*>> blocks;
    size_t bytes = 0;
};
PinnedCache& pinned_cache() {
    static PinnedCache* c = new PinnedCache();
    return *c;
}
constexpr size_t kPinnedCacheMax = size_t(256) << 20;
}  // namespace

void* pinned_take(size_t bytes) {
    {
        PinnedCache& pc = pinned_cache();
        std::lock_guard<std::mutex> lock(pc.mu);
        auto it = pc.blocks.find(bytes);
        if (it != pc.blocks.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            pc.bytes -= bytes;
            return p;
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    return p;
}

// Free every kept pinned block.
void pinned_release_all() {
    PinnedCache& pc = pinned_cache();
    std::lock_guard<std::mutex> lock(pc.mu);
    for (auto& kv : pc.blocks)
        for (void* p : kv.second) (void)hipHostFree(p);
    pc.blocks.clear();
    pc.bytes = 0;
}

void pinned_give(void* p, size_t bytes) {
    if (!p) return;
    PinnedCache& pc = pinned_cache();
    {
        std::lock_guard<std::mutex> lock(pc.mu);
        if (pc.bytes + bytes <= kPinnedCacheMax) {
            pc.blocks[bytes].push_back(p);
            pc.bytes += bytes;
            return;
        }
    }
    (void)hipHostFree(p);
}

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

namespace {
struct DeviceCache {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void*>> blocks;  // (device, bytes) -> free blocks
    std::map<void*, std::pair<int, size_t>> live;                // blocks handed out
    size_t bytes = 0;
};
DeviceCache& device_cache() {
    static DeviceCache* c = new DeviceCache();  // never destroyed, like the blocks it keeps
    return *c;
}
constexpr size_t kDeviceCacheMax = size_t(24) << 30;

// Return every kept block of `device` (caller holds the lock).
void release_device_blocks(DeviceCache& dc, int device) {
    for (auto it = dc.blocks.begin(); it != dc.blocks.end(); ++it) {
        if (it->first.first != device) continue;
        for (void* p : it->second) {
            (void)hipFree(p);
            dc.bytes -= it->first.second;
        }
        it->second.clear();
    }
}
}  // namespace

// Grid streams of destroyed grids, per device, handed to the next grid
// (creating a stream costs more than a small case's step).
std::mutex g_stream_mu;
std::map<int, std::vector<hipStream_t>> g_free_streams;

hipStream_t take_grid_stream(int device) {
    {
        std::lock_guard<std::mutex> lock(g_stream_mu);
        std::vector<hipStream_t>& v = g_free_streams[device];
        if (!v.empty()) {
            hipStream_t st = v.back();
            v.pop_back();
            return st;
        }
    }
    hipStream_t st = nullptr;
    LIFE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
}

// The stream has no queued work (synchronised by the caller).
void give_grid_stream(int device, hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_stream_mu);
    g_free_streams[device].push_back(st);
}

void* dalloc_cached_bytes(size_t bytes) {
    int dev = 0;
    LIFE_HIP(hipGetDevice(&dev));
    DeviceCache& dc = device_cache();
    std::lock_guard<std::mutex> lock(dc.mu);
    auto it = dc.blocks.find({dev, bytes});
    if (it != dc.blocks.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        dc.bytes -= bytes;
        dc.live[p] = {dev, bytes};
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        release_device_blocks(dc, dev);
        LIFE_HIP(hipMalloc(&p, bytes));
    }
    dc.live[p] = {dev, bytes};
    return p;
}

void dfree_cached(void* p) {
    if (!p) return;
    DeviceCache& dc = device_cache();
    std::lock_guard<std::mutex> lock(dc.mu);
    auto it = dc.live.find(p);
    if (it == dc.live.end()) {
        (void)hipFree(p);
        return;
    }
    const std::pair<int, size_t> key = it->second;
    dc.live.erase(it);
    if (dc.bytes + key.second <= kDeviceCacheMax) {
        dc.blocks[key].push_back(p);
        dc.bytes += key.second;
    } else {
        (void)hipFree(p);
    }
}

CaseScalars derive_scalars(const life_params& p) {
    CaseScalars s{};
    const double rho0 = 1.0;
    s.cs = 1.0 / std::sqrt(3.0);
    s.cs2 = s.cs * s.cs;
    s.cs4 = s.cs * s.cs * s.cs * s.cs;
    if (p.omega > 0.0) {
        s.omega = p.omega;
    } else {
        const double c = 1.0 / std::sqrt(3.0);
        const double h = p.height_p / (p.Ny - 1);
        s.omega = 1.0 / (p.nu_p * p.tStep / ((c * c) * (h * h)) + 0.5);
    }
    s.tau = 1.0 / s.omega;
    s.nu = (s.tau - 0.5) * s.cs2;
    s.Dx = p.height_p / (p.Ny - 1);
    s.Dt = (s.Dx * s.Dx) * s.nu / p.nu_p;
    s.Dm = (p.rho_p / rho0) * (s.Dx * s.Dx * s.Dx);
    s.Drho = p.rho_p / rho0;
    return s;
}

double ramp_coefficient(const life_params& p, double Dt, int t) {
    if (p.inlet_ramp > 0.0 && Dt * t <= p.inlet_ramp) return (1.0 - std::cos(M_PI * Dt * t / p.inlet_ramp)) / 2.0;
    return 1.0;
}

namespace {

// Site type of the reference's type map (Grid.cpp:931-957): bottom/top
// override left/right at the corners.
int site_type(const life_params& p, int i, int j) {
    int t = LIFE_FLUID;
    if (i == 0) t = p.wall_left;
    else if (i == p.Nx - 1) t = p.wall_right;
    if (j == 0) t = p.wall_bottom;
    else if (j == p.Ny - 1) t = p.wall_top;
    return t;
}

// GridClass::getNormalVector, Grid.cpp:498-545.
void normal_vector(const life_params& p, int i, int j, int& nx, int& ny, int& nd) {
    nx = ny = 0;
    nd = 0;
    if (i == 0) {
        nx = 1;
        nd = 0;
    } else if (i == p.Nx - 1) {
        nx = -1;
        nd = 0;
    }
    if (j == 0) {
        ny = 1;
        nd = 1;
    } else if (j == p.Ny - 1) {
        ny = -1;
        nd = 1;
    }
    if (nx != 0 && ny != 0) {
        const int tx = site_type(p, i + nx, j), ty = site_type(p, i, j + ny);
        if (tx == LIFE_FLUID && ty == LIFE_FLUID)
            throw std::runtime_error("Corner node is surrounded by fluid lattice sites");
        else if (tx == LIFE_FLUID) {
            nd = 0;
            ny = 0;
        } else if (ty == LIFE_FLUID) {
            nd = 1;
            nx = 0;
        }
    }
}

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

}  // namespace

void body_force(const life_grid& g, int t, double F[2]) {
    const life_params& p = g.p;
    const CaseScalars& cs = g.cs;
    const double DxDt = cs.Dx * cs.Dt;
    const double rho_init = 1.0;  // the initial density the reference multiplies by (Grid.cpp:1042)
    double wave = 1.0;
    if (p.womersley != 0.0 && t > 0) {
        // Grid.cpp:59-60; the density factor only meets gravity, which is 0 here
        const double period = ((p.height_p * p.height_p) * M_PI) / (2.0 * (p.womersley * p.womersley) * p.nu_p);
        wave = std::cos(2.0 * M_PI * t * cs.Dt / period);
        F[0] = (rho_init * cs.Drho * p.gravityX + p.dpdx * wave) * (DxDt * DxDt) / cs.Dm;
        F[1] = (rho_init * cs.Drho * p.gravityY + p.dpdy * wave) * (DxDt * DxDt) / cs.Dm;
        return;
    }
    // Grid.cpp:1042-1043
    F[0] = (rho_init * cs.Drho * p.gravityX + p.dpdx) * (DxDt * DxDt) / cs.Dm;
    F[1] = (rho_init * cs.Drho * p.gravityY + p.dpdy) * (DxDt * DxDt) / cs.Dm;
}

GridScalars step_scalars(const life_grid& g, int t) {
    GridScalars s = g.g;
    if (g.p.womersley == 0.0) return s;
    double F[2], Fm[2];
    body_force(g, t, F);
    body_force(g, t - 1, Fm);
    s.Fx = F[0];
    s.Fy = F[1];
    s.Fx0 = F[0] + 0.0;
    s.Fy0 = F[1] + 0.0;
    s.Fxm = Fm[0];
    s.Fym = Fm[1];
    if (g.fsite[0]) {
        // gravity with WOMERSLEY: force_xy meets each site's density
        // (Grid.cpp:59-60); this step writes plane t & 1 and reads the other
        const life_params& p = g.p;
        const CaseScalars& cs = g.cs;
        double wave = 1.0;
        if (t > 0) {
            const double period = ((p.height_p * p.height_p) * M_PI) / (2.0 * (p.womersley * p.womersley) * p.nu_p);
            wave = std::cos(2.0 * M_PI * t * cs.Dt / period);
        }
        const double DxDt = cs.Dx * cs.Dt;
        s.fcur = g.fsite[t & 1];
        s.fprev = g.fsite[(t + 1) & 1];
        s.fDrho = cs.Drho;
        s.fgx = p.gravityX;
        s.fgy = p.gravityY;
        s.fdx = p.dpdx * wave;
        s.fdy = p.dpdy * wave;
        s.fsc = DxDt * DxDt;
        s.fDm = cs.Dm;
    }
    return s;
}

void lbm_join(life_grid* g) {
    if (!g->tail_pending) return;
    LIFE_HIP(hipStreamWaitEvent(g->stream, g->ev_tail, 0));
    g->tail_pending = false;
}

void lbm_bulk_early(life_grid* g, hipEvent_t ev_begin, hipEvent_t ev_end) {
    if (g->explicit_state) throw std::runtime_error("early stream-collide needs the populations' own moments");
    if (g->fsite[0]) throw std::runtime_error("early stream-collide does not take a per-site force");
    const int src = g->cur, dst = 1 - g->cur;
    // populations final: the step's tail is already queued on stream2, or
    // the step ran on the grid's stream
    if (!g->tail_pending) {
        LIFE_HIP(hipEventRecord(g->ev_ready, g->stream));
        LIFE_HIP(hipStreamWaitEvent(g->stream2, g->ev_ready, 0));
    } else if (g->tail_stream != g->stream2) {
        LIFE_HIP(hipStreamWaitEvent(g->stream2, g->ev_tail, 0));
    }
    hipStream_t s = g->stream2;
    if (ev_begin) LIFE_HIP(hipEventRecord(ev_begin, s));
    const GridScalars next = step_scalars(*g, g->t_step + 1);  // the step this kernel belongs to
    if (g->prec == kFP64)
        launch_push64(static_cast<const double*>(g->f[src]), static_cast<double*>(g->f[dst]), next, g->rho0, g->u0,
                      g->fibm, false, s);
    else
        launch_push32(static_cast<const float*>(g->f[src]), static_cast<float*>(g->f[dst]), next, g->rho0, g->u0,
                      g->fibm, false, s);
    if (ev_end) LIFE_HIP(hipEventRecord(ev_end, s));
    LIFE_HIP(hipGetLastError());
    LIFE_HIP(hipEventRecord(g->ev_bulk, s));
    g->bulk_pending = true;
}

void lbm_step(life_grid* g, int t, hipEvent_t ev_begin, hipEvent_t ev_end) {
    g->g = step_scalars(*g, t);
    g->t_step = t;
    const double ramp = ramp_coefficient(g->p, g->cs.Dt, t);
    const int src = g->cur, dst = 1 - g->cur;
    BCArgs a{g->bc, g->nbc, g->u_in, g->rho_in, g->delU, ramp};
    const bool ex = g->explicit_state;
    hipStream_t s = g->stream;
    const bool early = g->bulk_pending;
    bool off = early;  // the step runs on a stream other than the grid's
    if (early) {
        // the rest of the step follows the bulk kernel on stream2, after
        // the grid stream's work so far (the spread it reads)
        LIFE_HIP(hipEventRecord(g->ev_ready, g->stream));
        LIFE_HIP(hipStreamWaitEvent(g->stream2, g->ev_ready, 0));
        s = g->stream2;
        g->bulk_pending = false;
    } else {
        lbm_join(g);
        if (g->stream_pool) {
            // the grid's stream holds part of the device: the whole step
            // runs on the grid's own stream, after that stream's work so far
            LIFE_HIP(hipEventRecord(g->ev_ready, g->stream));
            LIFE_HIP(hipStreamWaitEvent(g->stream_pool, g->ev_ready, 0));
            s = g->stream_pool;
            off = true;
        }
    }
    if (g->prec == kFP64) {
        const double* fs = static_cast<const double*>(g->f[src]);
        double* fd = static_cast<double*>(g->f[dst]);
        if (g->convective) launch_convective_speed<double>(fs, g->g, g->fibm, g->u0, ex, g->ucol, g->delU, true, s);
        if (!early) {
            if (ev_begin) LIFE_HIP(hipEventRecord(ev_begin, s));
            launch_push64(fs, fd, g->g, g->rho0, g->u0, g->fibm, ex, s);
            if (ev_end) LIFE_HIP(hipEventRecord(ev_end, s));
        }
        if (!ex && g->md.n > 0) launch_support_push64(fs, fd, g->g, g->fibm, g->sc.uniq, g->sc.uniq_n, g->sc.cap, s);
        launch_bc<double>(fs, fd, g->g, a, g->fibm, g->rho0, g->u0, ex, s);
    } else {
        const float* fs = static_cast<const float*>(g->f[src]);
        float* fd = static_cast<float*>(g->f[dst]);
        if (g->convective) launch_convective_speed<float>(fs, g->g, g->fibm, g->u0, ex, g->ucol, g->delU, false, s);
        if (!early) {
            if (ev_begin) LIFE_HIP(hipEventRecord(ev_begin, s));
            launch_push32(fs, fd, g->g, g->rho0, g->u0, g->fibm, ex, s);
            if (ev_end) LIFE_HIP(hipEventRecord(ev_end, s));
        }
        if (!ex && g->md.n > 0) launch_support_push32(fs, fd, g->g, g->fibm, g->sc.uniq, g->sc.uniq_n, g->sc.cap, s);
        launch_bc<float>(fs, fd, g->g, a, g->fibm, g->rho0, g->u0, ex, s);
    }
    LIFE_HIP(hipGetLastError());
    if (off) {
        LIFE_HIP(hipEventRecord(g->ev_tail, s));
        g->tail_pending = true;
        g->tail_stream = s;
    }
    g->cur = dst;
    g->explicit_state = false;
}

void blowup_scan(life_grid* g, int t, hipStream_t s) {
    if (!g->blow_d) {
        g->blow_d = dalloc<unsigned long long>(1);
        LIFE_HIP(hipHostMalloc((void**)&g->blow_h, sizeof(unsigned long long), hipHostMallocDefault));
        LIFE_HIP(hipEventCreateWithFlags(&g->ev_blow, hipEventDisableTiming));
    }
    if (g->prec == kFP64) launch_blowup_scan<double>(static_cast<const double*>(g->f[g->cur]), g->g, g->blow_d, s);
    else launch_blowup_scan<float>(static_cast<const float*>(g->f[g->cur]), g->g, g->blow_d, s);
    LIFE_HIP(hipGetLastError());
    LIFE_HIP(hipMemcpyAsync(g->blow_h, g->blow_d, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    LIFE_HIP(hipEventRecord(g->ev_blow, s));
    g->blow_t = t;
}

void blowup_check(life_grid* g, bool wait) {
    if (g->blow_t < 0) return;
    if (!wait) {
        const hipError_t q = hipEventQuery(g->ev_blow);
        if (q == hipErrorNotReady) return;
        LIFE_HIP(q);
    } else {
        LIFE_HIP(hipEventSynchronize(g->ev_blow));
    }
    const int t = g->blow_t;
    g->blow_t = -1;
    const unsigned long long id = *g->blow_h;
    if (id == ~0ull) return;
    const long long Ny = g->g.Ny;
    // GridClass::writeInfo, Grid.cpp:573-579
    throw std::runtime_error("Simulation blew up (t = " + std::to_string(t) + ") at i = " +
                             std::to_string((long long)id / Ny) + ", j = " + std::to_string((long long)id % Ny) +
                             "...exiting");
}

void blowup_cadence(life_grid* g, int t, hipStream_t s) {
    if (g->check_every <= 0 || t % g->check_every != 0) return;
    // the previous scan was queued check_every steps ago: usually done, else
    // the host waits for it (so it never runs more than that far ahead)
    blowup_check(g, true);
    blowup_scan(g, t, s);
}

void ibm_interpolate(life_grid* g) {
    if (g->md.n <= 0) return;
    lbm_join(g);
    if (g->prec == kFP64) launch_interpolate<double>(static_cast<const double*>(g->f[g->cur]), g->md, g->g, g->stream);
    else launch_interpolate<float>(static_cast<const float*>(g->f[g->cur]), g->md, g->g, g->stream);
    LIFE_HIP(hipGetLastError());
}

void ibm_spread(life_grid* g) {
    launch_spread(g->md, g->g, g->fibm, g->sc, g->dense_clear_next, g->stream);
    g->dense_clear_next = false;
    LIFE_HIP(hipGetLastError());
}

void ibm_interp_spread(life_grid* g) {
    if (g->md.n <= 0) return;
    lbm_join(g);
    if (launch_interp_spread_small(g->f[g->cur], g->prec == kFP64, g->md, g->g, g->fibm, g->sc, g->dense_clear_next,
                                   g->stream)) {
        g->dense_clear_next = false;
        LIFE_HIP(hipGetLastError());
        return;
    }
    ibm_interpolate(g);
    ibm_spread(g);
}

void ibm_find_support(life_grid* g, int flexible_only) {
    if (g->md.n <= 0) return;
    launch_support_ds(g->md, g->g, g->nbodies, flexible_only, g->stream);
    LIFE_HIP(hipGetLastError());
}

void ibm_epsilon_matrix(life_grid* g, int first, int n, double* A_host) {
    const size_t need = (size_t)n * n;
    if (need > g->epsA_cap) {
        dfree(g->epsA);
        g->epsA = dalloc<double>(need);
        g->epsA_cap = need;
    }
    launch_epsilon_matrix(g->md, g->g, first, n, g->epsA, g->stream);
    LIFE_HIP(hipGetLastError());
    LIFE_HIP(hipMemcpyAsync(A_host, g->epsA, need * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream));
}

void ibm_epsilon_band_setup(life_grid* g, const std::vector<int>& first, const std::vector<int>& n, int KB) {
    for (void* p : {(void*)g->eps_first, (void*)g->eps_n, (void*)g->eps_off, (void*)g->eps_outside,
                    (void*)g->eps_band, (void*)g->eps_y})
        dfree(p);
    if (KB != 8) throw std::runtime_error("banded epsilon solve is built for a half-bandwidth of 8");
    g->eps_nsys = (int)first.size();
    g->eps_KB = KB;
    g->eps_first_h = first;
    g->eps_n_h = n;
    std::vector<long long> off(first.size());
    long long tot = 0;
    for (size_t k = 0; k < first.size(); k++) {
        off[k] = tot;
        tot += (long long)n[k] * (3 * KB + 1);
    }
    g->eps_first = dalloc<int>(first.size());
    g->eps_n = dalloc<int>(first.size());
    g->eps_off = dalloc<long long>(first.size());
    g->eps_outside = dalloc<int>(first.size());
    if (g->eps_outside_h) (void)hipHostFree(g->eps_outside_h);
    g->eps_outside_h = nullptr;
    LIFE_HIP(hipHostMalloc((void**)&g->eps_outside_h, sizeof(int) * (first.empty() ? 1 : first.size()),
                           hipHostMallocDefault));
    g->eps_band = dalloc<double>((size_t)tot);
    g->eps_y = dalloc<double>((size_t)(g->md.n > 0 ? g->md.n : 1));
    if (!first.empty()) {
        LIFE_HIP(hipMemcpyAsync(g->eps_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice,
                                g->stream));
        LIFE_HIP(hipMemcpyAsync(g->eps_n, n.data(), sizeof(int) * n.size(), hipMemcpyHostToDevice, g->stream));
        LIFE_HIP(hipMemcpyAsync(g->eps_off, off.data(), sizeof(long long) * off.size(), hipMemcpyHostToDevice,
                                g->stream));
    }
    LIFE_HIP(hipStreamSynchronize(g->stream));
}

void ibm_epsilon_band_launch(life_grid* g, bool copy_flags) {
    if (g->eps_nsys <= 0) return;
    // (the band kernels write every system's out-of-band flag)
    // FP32-shifted cases: cyclic reduction; the plain method: the pivoted
    // elimination in the reference's order
    launch_eps_band(g->md, g->g, g->eps_first, g->eps_n, g->eps_n_h.data(), g->eps_nsys, g->eps_KB, g->eps_band, g->eps_off,
                    g->eps_outside, g->eps_y, g->prec == LIFE_FP32_SHIFTED, g->stream);
    LIFE_HIP(hipGetLastError());
    if (copy_flags)
        LIFE_HIP(hipMemcpyAsync(g->eps_outside_h, g->eps_outside, sizeof(int) * g->eps_nsys, hipMemcpyDeviceToHost,
                                g->stream));
}

std::vector<int> ibm_epsilon_band_dense(const life_grid* g) {
    std::vector<int> dense;
    for (int k = 0; k < g->eps_nsys; k++)
        if (g->eps_outside_h[k]) dense.push_back(k);
    return dense;
}

void check_flags(life_grid* g) {
    if (!g->flags) return;
    int h[2] = {0, 0};
    LIFE_HIP(hipMemcpyAsync(h, g->flags, sizeof h, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream));
    if (h[0]) throw std::runtime_error("Support buffer size is not big enough for number of support points");
    if (h[1]) throw std::runtime_error("more than 64 markers share one support site");
}

static void free_markers(life_grid* g) {
    MarkerDev& m = g->md;
    for (void* p : {(void*)m.pos, (void*)m.vel, (void*)m.force, (void*)m.eps, (void*)m.ds, (void*)m.body,
                    (void*)m.supp_cnt, (void*)m.supp_site, (void*)m.supp_ij, (void*)m.supp_delta,
                    (void*)g->body_first, (void*)g->body_n, (void*)g->body_flex, (void*)g->sc.site_count,
                    (void*)g->sc.site_uid, (void*)g->sc.uniq, (void*)g->sc.uniq_n, (void*)g->sc.pair_slot,
                    (void*)g->sc.list, (void*)g->flags})
        dfree(p);
    m = MarkerDev{};
    g->sc = SpreadScratch{};
    g->body_first = g->body_n = g->body_flex = nullptr;
    g->flags = nullptr;
}

namespace {
// CU-masked streams, one per (device, reserve), shared by every grid and
// server on the device, created on first use and kept until the host's
// final teardown (life_release_cached).
struct MaskedStreams {
    std::mutex mu;
    std::map<std::pair<int, int>, hipStream_t> streams;
};
MaskedStreams& masked_streams() {
    static MaskedStreams* m = new MaskedStreams();
    return *m;
}

// Live grids and servers.  When the last one goes, the pooled memory is
// released; the streams stay for the next grid or server.
std::mutex g_handles_mu;
int g_handles = 0;
}  // namespace

// The stream is masked to all CUs but the reserve.  The reserved CUs are
// spread over the 8 XCDs whether the mask bits number the CUs XCD by XCD or
// round-robin over the XCDs: CU x*G + ((x + 8*k) mod G) for XCD x and
// k = 0, 1, ... (G = CUs per XCD) falls on XCD x either way.  A CU-masked
// stream has a hardware queue of its own, so there is one per device and
// reserve for the process, shared by every grid or server on it.  Sharing
// cannot deadlock: a user records the event its kernel waits for before it
// queues that wait.
hipStream_t masked_stream(int device, int reserve_cus) {
    MaskedStreams& ms = masked_streams();
    std::lock_guard<std::mutex> lock(ms.mu);
    int ncu = 0;
    LIFE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    const int reserve = reserve_cus < 0 ? 0 : (reserve_cus < ncu / 8 ? reserve_cus : ncu / 8);  // at most G/8 per XCD
    auto it = ms.streams.find({device, reserve});
    if (it != ms.streams.end()) return it->second;
    hipStream_t st = nullptr;
    if (reserve == 0 || ncu % 8 != 0) {
        LIFE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    } else {
        const int G = ncu / 8;
        std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
        for (int c = 0; c < ncu; c++) mask[(size_t)c / 32] |= 1u << (c % 32);
        for (int r = 0; r < reserve; r++) {
            const int x = r % 8, k = r / 8;
            const int c = x * G + (x + 8 * k) % G;
            mask[(size_t)c / 32] &= ~(1u << (c % 32));
        }
        LIFE_HIP(hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()));
    }
    ms.streams[{device, reserve}] = st;
    return st;
}

// Frees the pooled memory (kept device and page-locked blocks) and, with
// streams, the idle grid streams and the CU-masked streams too (only when
// no grid or server is live: the caller holds g_handles_mu and checked).
// Streams are released only at the host's final teardown
// (life_release_cached), never when the last handle goes: destroying a
// CU-masked stream and creating the next one within a process (every time
// a test or a host closed its last case and opened another) stalled the
// next life_server_create, inside stream creation, for minutes; a stream
// created once per (device, reserve) for the process does not.
static void release_pools(bool streams) {
    {
        DeviceCache& dc = device_cache();
        std::lock_guard<std::mutex> lock(dc.mu);
        for (auto& kv : dc.blocks)
            for (void* p : kv.second) (void)hipFree(p);
        dc.blocks.clear();
        dc.bytes = 0;
    }
    pinned_release_all();
    if (!streams) return;
    {
        std::lock_guard<std::mutex> lock(g_stream_mu);
        for (auto& kv : g_free_streams)
            for (hipStream_t st : kv.second) (void)hipStreamDestroy(st);
        g_free_streams.clear();
    }
    MaskedStreams& ms = masked_streams();
    std::lock_guard<std::mutex> lock(ms.mu);
    for (auto& kv : ms.streams) {
        (void)hipStreamSynchronize(kv.second);
        (void)hipStreamDestroy(kv.second);
    }
    ms.streams.clear();
}

void handle_acquire() {
    std::lock_guard<std::mutex> lock(g_handles_mu);
    ++g_handles;
}

void handle_release() {
    std::lock_guard<std::mutex> lock(g_handles_mu);
    if (--g_handles == 0) release_pools(false);
}

void grid_reserve_coupling_cus(life_grid* g, int cus) {
    if (std::getenv("LIFE_COUPLING_CUS")) return;  // the environment's choice stands
    hipStream_t s = masked_stream(g->device, cus);
    if (s == g->stream2) return;
    LIFE_HIP(hipStreamSynchronize(g->stream2));
    g->stream2 = s;
}

// A stream on the device's first `xcds` XCDs (inside) or on the others,
// made once per device and kept with the CU-masked streams (keys below
// zero).  The CUs of XCD x are x*G .. x*G+G-1 (G = CUs per XCD).  Only
// whole XCDs: the workgroups of a launch are dealt to the XCDs in turn, so
// an XCD with part of its CUs enabled holds up every launch (loaded canopy,
// coupling on 48 or 56 CUs: 0.66 ms/step; on 64, two whole XCDs: 0.544).
// Null where the CUs do not split into eight equal XCDs.
static hipStream_t xcd_group_stream(int device, int xcds, bool inside) {
    MaskedStreams& ms = masked_streams();
    std::lock_guard<std::mutex> lock(ms.mu);
    const std::pair<int, int> key{device, inside ? -1 - xcds : -101 - xcds};
    auto it = ms.streams.find(key);
    if (it != ms.streams.end()) return it->second;
    int ncu = 0;
    LIFE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    if (ncu % 8 != 0 || xcds <= 0 || xcds >= 8) return nullptr;
    const int G = ncu / 8;
    std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
    for (int c = 0; c < ncu; c++)
        if ((c < xcds * G) == inside) mask[(size_t)c / 32] |= 1u << (c % 32);
    hipStream_t st = nullptr;
    LIFE_HIP(hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()));
    ms.streams[key] = st;
    return st;
}

void grid_split_coupling(life_grid* g, int xcds, int fallback_cus) {
    if (std::getenv("LIFE_COUPLING_CUS")) return;  // the environment's choice stands
    if (const char* e = std::getenv("LIFE_COUPLING_XCDS")) xcds = std::atoi(e);
    hipStream_t cs = xcds > 0 ? xcd_group_stream(g->device, xcds, true) : nullptr;
    hipStream_t bs = cs ? xcd_group_stream(g->device, xcds, false) : nullptr;
    if (!cs || !bs) {
        grid_reserve_coupling_cus(g, fallback_cus);
        return;
    }
    LIFE_HIP(hipStreamSynchronize(g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream2));
    if (!g->stream_pool) g->stream_pool = g->stream;
    g->stream = cs;
    g->stream2 = bs;
}

}  // namespace life

using namespace life;

extern "C" {

const char* life_last_error(void) { return life::g_error.c_str(); }

const char* life_version(void) { return "life_hip 0.1 (gfx950)"; }

int life_release_cached(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_handles_mu);
        release_pools(g_handles == 0);
    });
}

int life_device_count(int* count) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

namespace {
// CUs kept free of the early bulk stream-collide (second stream) for the
// coupling kernels that run beside it (environment LIFE_COUPLING_CUS,
// default 0: no reserve).  A reserve lets the small coupling workgroups
// start without waiting for stream-collide workgroups to retire; it paid
// while the coupling chain was the critical path, and costs stream-collide
// time once the coupling fits beside it.
int coupling_cus() {
    static const int n = [] {
        const char* e = std::getenv("LIFE_COUPLING_CUS");
        const int v = e ? std::atoi(e) : 0;
        return v < 0 ? 0 : v;
    }();
    return n;
}

// Wave priority of the coupling kernels (LIFE_COUPLING_PRIO, 0..3, default
// 3), set once per device.
void init_coupling_priority(int device) {
    static std::mutex mu;
    static std::map<int, bool> done;
    std::lock_guard<std::mutex> lock(mu);
    if (done[device]) return;
    const char* e = std::getenv("LIFE_COUPLING_PRIO");
    int v = e ? std::atoi(e) : 3;
    v = v < 0 ? 0 : (v > 3 ? 3 : v);
    LIFE_HIP(set_coupling_priority(v));
    done[device] = true;
}

// The second stream of a grid: the early bulk stream-collide.
hipStream_t bulk_stream(int device) { return masked_stream(device, coupling_cus()); }
}  // namespace



int life_grid_create(const life_params* p, int precision, int device, life_grid** out) {
    *out = nullptr;
    life_grid* g = new life_grid();
    handle_acquire();
    int rc = guarded([&] {
        if (p->Nx < 4 || p->Ny < 4) throw std::runtime_error("grid must be at least 4x4");
        if (precision != LIFE_FP64 && precision != LIFE_FP32_SHIFTED) throw std::runtime_error("unknown precision");
        if (p->wall_left == LIFE_CONVECTIVE || p->wall_bottom == LIFE_CONVECTIVE || p->wall_top == LIFE_CONVECTIVE)
            throw std::runtime_error("Currently convective BC only supported for right boundary");
        if ((long long)p->Nx * p->Ny > (1LL << 31) - 1) throw std::runtime_error("lattice too large for 32-bit site ids");
        g->device = device;
        LIFE_HIP(hipSetDevice(device));
        g->stream = take_grid_stream(device);
        LIFE_HIP(hipEventCreate(&g->ev_a));
        LIFE_HIP(hipEventCreate(&g->ev_b));
        LIFE_HIP(hipEventCreate(&g->ev_k0));
        LIFE_HIP(hipEventCreate(&g->ev_k1));
        g->stream2 = bulk_stream(device);
        init_coupling_priority(device);
        LIFE_HIP(hipEventCreateWithFlags(&g->ev_ready, hipEventDisableTiming));
        LIFE_HIP(hipEventCreateWithFlags(&g->ev_bulk, hipEventDisableTiming));
        LIFE_HIP(hipEventCreateWithFlags(&g->ev_tail, hipEventDisableTiming));
        g->p = *p;
        g->prec = precision;
        g->cs = derive_scalars(*p);
        const int Nx = p->Nx, Ny = p->Ny;
        GridScalars& s = g->g;
        s.Nx = Nx;
        s.Ny = Ny;
        s.N = (long long)Nx * Ny;
        s.S = (s.N + 255) / 256 * 256;
        s.cm = p->central_moments ? 1 : 0;
        s.omega = g->cs.omega;
        s.cs2 = g->cs.cs2;
        s.cs4 = g->cs.cs4;
        const bool site_force = p->womersley != 0.0 && (p->gravityX != 0.0 || p->gravityY != 0.0);
        if (site_force && precision != LIFE_FP64)
            throw std::runtime_error("WOMERSLEY forcing with gravity needs FP64 populations (force_xy varies "
                                     "from site to site)");
        s.fcur = nullptr;
        s.fprev = nullptr;
        double F[2];
        body_force(*g, 0, F);
        s.Fx = F[0];
        s.Fy = F[1];
        s.Fx0 = s.Fx + 0.0;
        s.Fy0 = s.Fy + 0.0;
        s.Fxm = s.Fx;
        s.Fym = s.Fy;
        s.Dx = g->cs.Dx;
        s.velScale = g->cs.Dt / g->cs.Dx;
        if (site_force) {
            // force_xy of the previous and the current step, per site
            for (int k = 0; k < 2; k++) {
                g->fsite[k] = dalloc<double>(2 * (size_t)s.S);
                launch_fill_f64(g->fsite[k], 0.0, 2 * s.S, g->stream);
            }
        }

        // boundary sites in the reference's i-major order (Grid.cpp:931-957)
        for (int i = 0; i < Nx; i++) {
            const bool edge_i = (i == 0 || i == Nx - 1);
            for (int j = 0; j < Ny; j++) {
                if (!edge_i && j != 0 && j != Ny - 1) continue;
                const int t = site_type(*p, i, j);
                if (t == LIFE_FLUID) continue;
                int nx, ny, nd;
                normal_vector(*p, i, j, nx, ny, nd);
                BCSite b{};
                b.id = i * Ny + j;
                b.type = (int8_t)t;
                b.nx = (int8_t)nx;
                b.ny = (int8_t)ny;
                b.nd = (int8_t)nd;
                b.i = i;
                b.j = j;
                g->bc_host.push_back(b);
                if (t == LIFE_CONVECTIVE) g->convective = true;
            }
        }
        g->nbc = (int)g->bc_host.size();
        // inlet profile (Grid.cpp:959-1001)
        const double Dx = g->cs.Dx, Dt = g->cs.Dt;
        g->u_in_host.assign((size_t)Ny * 2, 0.0);
        g->rho_in_host.assign(Ny, 1.0);
        for (int j = 0; j < Ny; j++) {
            double* ui = &g->u_in_host[2 * j];
            if (p->profile == LIFE_UNIFORM) {
                ui[0] = p->uxInlet_p * Dt / Dx;
                ui[1] = p->uyInlet_p * Dt / Dx;
            } else if (p->profile == LIFE_PARABOLIC) {
                const double R = p->height_p / 2.0, Y = j * Dx - R;
                ui[0] = 1.5 * (p->uxInlet_p * Dt / Dx) * (1.0 - (Y / R) * (Y / R));
                ui[1] = 1.5 * (p->uyInlet_p * Dt / Dx) * (1.0 - (Y / R) * (Y / R));
            } else if (p->profile == LIFE_SHEAR) {
                const double H = p->height_p, Y = j * Dx;
                ui[0] = (p->uxInlet_p * Dt / Dx) * (Y / H);
                ui[1] = (p->uyInlet_p * Dt / Dx) * (Y / H);
            } else if (p->profile == LIFE_BOUNDARY_LAYER) {
                const double H = p->height_p, Y = j * Dx;
                ui[0] = ((1.5 * p->uxInlet_p * Dt / Dx) / (H * H)) * Y * (2.0 * H - Y);
                ui[1] = ((1.5 * p->uyInlet_p * Dt / Dx) / (H * H)) * Y * (2.0 * H - Y);
            } else {
                throw std::runtime_error("unknown inlet profile");
            }
        }
        // device buffers
        const size_t S = (size_t)s.S;
        for (int b = 0; b < 2; b++)
            g->f[b] = precision == LIFE_FP64 ? (void*)dalloc_cached<double>(9 * S)
                                             : (void*)dalloc_cached<float>(9 * S);
        g->rho0 = dalloc_cached<double>(S);
        g->u0 = dalloc_cached<double>(2 * S);
        g->fibm = dalloc_cached<double>(2 * S);
        g->bc = dalloc<BCSite>(g->nbc);
        g->u_in = dalloc<double>(2 * Ny);
        g->rho_in = dalloc<double>(Ny);
        g->delU = dalloc<double>(2 * Ny);
        g->ucol = dalloc<double>(6 * Ny);
        hipStream_t st = g->stream;
        if (g->nbc)
            LIFE_HIP(hipMemcpyAsync(g->bc, g->bc_host.data(), sizeof(BCSite) * g->nbc, hipMemcpyHostToDevice, st));
        LIFE_HIP(hipMemcpyAsync(g->u_in, g->u_in_host.data(), sizeof(double) * 2 * Ny, hipMemcpyHostToDevice, st));
        LIFE_HIP(hipMemcpyAsync(g->rho_in, g->rho_in_host.data(), sizeof(double) * Ny, hipMemcpyHostToDevice, st));
        launch_fill_f64(g->delU, 0.0, 2 * Ny, st);
        launch_fill_f64(g->fibm, 0.0, 2 * (long long)S, st);
        // initial moments (Grid.cpp:1004-1033) and f = f_eq (Grid.cpp:1048-1061)
        int mode = 0;
        double ux0 = 0.0, uy0 = 0.0;
        if (p->inlet_ramp > 0.0) {
            mode = 0;
        } else if (p->profile != LIFE_UNIFORM) {
            mode = 1;
        } else {
            ux0 = p->ux0_p * Dt / Dx;
            uy0 = p->uy0_p * Dt / Dx;
        }
        launch_init_moments(g->rho0, g->u0, s, mode, ux0, uy0, g->u_in, g->bc, g->nbc, st);
        if (precision == LIFE_FP64) launch_init_equilibrium<double>(static_cast<double*>(g->f[0]), s, g->rho0, g->u0, st);
        else launch_init_equilibrium<float>(static_cast<float*>(g->f[0]), s, g->rho0, g->u0, st);
        LIFE_HIP(hipGetLastError());
        LIFE_HIP(hipStreamSynchronize(st));
        g->cur = 0;
        g->explicit_state = true;
    });
    if (rc) {
        life_grid_destroy(g);
        return rc;
    }
    *out = g;
    return 0;
}

int life_grid_destroy(life_grid* g) {
    if (!g) return 0;
    const int rc = guarded([&] {
        (void)hipSetDevice(g->device);
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        if (g->stream_pool) (void)hipStreamSynchronize(g->stream_pool);
        if (g->ev_bulk) (void)hipEventSynchronize(g->ev_bulk);  // this grid's last bulk kernel
        if (g->ev_tail) (void)hipEventSynchronize(g->ev_tail);  // and step tail on stream2
        free_markers(g);
        // the grid's work is done (synchronised above): its lattice blocks can be handed out again
        for (void* p : {g->f[0], g->f[1], (void*)g->rho0, (void*)g->u0, (void*)g->fibm}) dfree_cached(p);
        for (void* p : {(void*)g->bc, (void*)g->u_in, (void*)g->rho_in, (void*)g->delU, (void*)g->ucol, (void*)g->epsA,
                        (void*)g->eps_first, (void*)g->eps_n, (void*)g->eps_off, (void*)g->eps_outside,
                        (void*)g->eps_band, (void*)g->eps_y, (void*)g->fsite[0], (void*)g->fsite[1]})
            dfree(p);
        if (g->eps_outside_h) (void)hipHostFree(g->eps_outside_h);
        if (g->blow_h) (void)hipHostFree(g->blow_h);
        dfree(g->blow_d);
        if (g->ev_blow) (void)hipEventDestroy(g->ev_blow);
        for (hipEvent_t e : {g->ev_a, g->ev_b, g->ev_k0, g->ev_k1, g->ev_ready, g->ev_bulk, g->ev_tail})
            if (e) (void)hipEventDestroy(e);
        // the coupling stream of a split grid is the device's, not the grid's
        if (g->stream_pool) give_grid_stream(g->device, g->stream_pool);
        else if (g->stream) give_grid_stream(g->device, g->stream);
        delete g;
    });
    handle_release();
    return rc;
}

int life_grid_scalars(const life_grid* g, double out[10]) {
    const CaseScalars& s = g->cs;
    const double v[10] = {s.omega, s.tau, s.nu, s.Dx, s.Dt, s.Dm, s.Drho, s.cs, s.cs2, s.cs4};
    std::memcpy(out, v, sizeof v);
    return 0;
}

int life_lbm_kernel(life_grid* g, int t) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        lbm_join(g);
        lbm_step(g, t, nullptr, nullptr);
    });
}

int life_grid_blowup_check(life_grid* g, int* i, int* j) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        lbm_join(g);
        *i = *j = -1;
        blowup_check(g, true);  // a scan still in flight from life_case_step is read first
        blowup_scan(g, g->t_step, g->stream);
        LIFE_HIP(hipEventSynchronize(g->ev_blow));
        g->blow_t = -1;
        const unsigned long long id = *g->blow_h;
        if (id != ~0ull) {
            *i = (int)((long long)id / g->g.Ny);
            *j = (int)((long long)id % g->g.Ny);
        }
    });
}

int life_grid_sync(life_grid* g) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        lbm_join(g);
        LIFE_HIP(hipStreamSynchronize(g->stream));
    });
}

int life_grid_download(life_grid* g, double* f_aos, double* rho, double* u, double* force_ibm) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        lbm_join(g);
        const size_t N = (size_t)g->g.N;
        double *df = nullptr, *dr = nullptr, *du = nullptr, *db = nullptr;
        try {
            if (f_aos) df = dalloc<double>(9 * N);
            if (rho) dr = dalloc<double>(N);
            if (u) du = dalloc<double>(2 * N);
            if (force_ibm) db = dalloc<double>(2 * N);
            if (g->prec == kFP64)
                launch_materialise<double>(static_cast<const double*>(g->f[g->cur]), g->g, g->fibm, g->rho0, g->u0,
                                           g->explicit_state, df, dr, du, db, g->stream);
            else
                launch_materialise<float>(static_cast<const float*>(g->f[g->cur]), g->g, g->fibm, g->rho0, g->u0,
                                          g->explicit_state, df, dr, du, db, g->stream);
            LIFE_HIP(hipGetLastError());
            if (df) LIFE_HIP(hipMemcpyAsync(f_aos, df, 9 * N * sizeof(double), hipMemcpyDeviceToHost, g->stream));
            if (dr) LIFE_HIP(hipMemcpyAsync(rho, dr, N * sizeof(double), hipMemcpyDeviceToHost, g->stream));
            if (du) LIFE_HIP(hipMemcpyAsync(u, du, 2 * N * sizeof(double), hipMemcpyDeviceToHost, g->stream));
            if (db) LIFE_HIP(hipMemcpyAsync(force_ibm, db, 2 * N * sizeof(double), hipMemcpyDeviceToHost, g->stream));
            LIFE_HIP(hipStreamSynchronize(g->stream));
        } catch (...) {
            dfree(df), dfree(dr), dfree(du), dfree(db);
            throw;
        }
        dfree(df), dfree(dr), dfree(du), dfree(db);
    });
}

int life_grid_upload(life_grid* g, const double* f_aos, const double* rho, const double* u, const double* force_ibm) {
    return guarded([&] {
        if (!f_aos || !rho || !u) throw std::runtime_error("life_grid_upload needs f, rho and u");
        LIFE_HIP(hipSetDevice(g->device));
        lbm_join(g);
        const size_t N = (size_t)g->g.N;
        double* df = dalloc<double>(9 * N);
        double* db = nullptr;
        try {
            LIFE_HIP(hipMemcpyAsync(df, f_aos, 9 * N * sizeof(double), hipMemcpyHostToDevice, g->stream));
            LIFE_HIP(hipMemcpyAsync(g->rho0, rho, N * sizeof(double), hipMemcpyHostToDevice, g->stream));
            LIFE_HIP(hipMemcpyAsync(g->u0, u, 2 * N * sizeof(double), hipMemcpyHostToDevice, g->stream));
            if (g->prec == kFP64) launch_upload<double>(static_cast<double*>(g->f[g->cur]), g->g, df, g->stream);
            else launch_upload<float>(static_cast<float*>(g->f[g->cur]), g->g, df, g->stream);
            if (force_ibm) {
                db = dalloc<double>(2 * N);
                LIFE_HIP(hipMemcpyAsync(db, force_ibm, 2 * N * sizeof(double), hipMemcpyHostToDevice, g->stream));
                launch_fibm_from_aos(g->fibm, g->g, db, g->stream);
            } else {
                launch_fill_f64(g->fibm, 0.0, 2 * g->g.S, g->stream);
            }
            LIFE_HIP(hipGetLastError());
            LIFE_HIP(hipStreamSynchronize(g->stream));
        } catch (...) {
            dfree(df), dfree(db);
            throw;
        }
        dfree(df), dfree(db);
        g->explicit_state = true;
        g->dense_clear_next = true;
    });
}

int life_solve_lapack(int n, int bc, const double* A, const double* b, double* x) {
    return guarded([&] {
        if (n < 0 || bc < 0 || bc > n) throw std::runtime_error("bad system size");
        std::vector<double> w;
        std::vector<int> iw;
        const int info = lu_solve_sparse(n, bc, A, b, x, w, iw);
        if (info) throw std::runtime_error("singular matrix (zero pivot column " + std::to_string(info - 1) + ")");
    });
}

int life_ibm_set_markers(life_grid* g, int n, const double* pos, const double* vel, const int* body, int nbodies,
                         const int* body_flex) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        LIFE_HIP(hipStreamSynchronize(g->stream));
        free_markers(g);
        if (n < 0 || nbodies < 0) throw std::runtime_error("negative marker or body count");
        std::vector<int> first(nbodies, -1), cnt(nbodies, 0);
        for (int k = 0; k < n; k++) {
            const int b = body[k];
            if (b < 0 || b >= nbodies) throw std::runtime_error("marker body index out of range");
            if (first[b] < 0) first[b] = k;
            else if (first[b] + cnt[b] != k) throw std::runtime_error("markers of a body must be contiguous");
            cnt[b]++;
        }
        for (int b = 0; b < nbodies; b++)
            if (first[b] < 0) first[b] = 0;
        g->body_first_h = first;
        g->body_n_h = cnt;
        g->body_flex_h.assign(body_flex, body_flex + nbodies);
        g->nbodies = nbodies;
        MarkerDev& m = g->md;
        m.n = n;
        m.pos = dalloc<double>(2 * (size_t)n);
        m.vel = dalloc<double>(2 * (size_t)n);
        m.force = dalloc<double>(2 * (size_t)n);
        m.eps = dalloc<double>(n);
        m.ds = dalloc<double>(n);
        m.body = dalloc<int>(n);
        m.supp_cnt = dalloc<int>(n);
        m.supp_site = dalloc<int>((size_t)n * kSupp);
        m.supp_ij = dalloc<int>((size_t)n * kSupp * 2);
        m.supp_delta = dalloc<double>((size_t)n * kSupp);
        g->body_first = dalloc<int>(nbodies);
        g->body_n = dalloc<int>(nbodies);
        g->body_flex = dalloc<int>(nbodies);
        g->flags = dalloc<int>(2);
        m.body_first = g->body_first;
        m.body_n = g->body_n;
        m.body_flex = g->body_flex;
        m.overflow = g->flags;
        SpreadScratch& sc = g->sc;
        sc.cap = n * kSupp;
        sc.site_count = dalloc<int>((size_t)g->g.N);
        sc.site_uid = dalloc<int>((size_t)g->g.N);
        sc.uniq = dalloc<int>(sc.cap);
        sc.uniq_n = dalloc<int>(1);
        sc.pair_slot = dalloc<int>(sc.cap);
        sc.list = dalloc<int>((size_t)sc.cap * kMaxPerSite);
        sc.overflow = g->flags + 1;
        hipStream_t s = g->stream;
        if (n) {
            LIFE_HIP(hipMemcpyAsync(m.pos, pos, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
            if (vel) LIFE_HIP(hipMemcpyAsync(m.vel, vel, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
            else launch_fill_f64(m.vel, 0.0, 2 * (long long)n, s);
            LIFE_HIP(hipMemcpyAsync(m.body, body, sizeof(int) * n, hipMemcpyHostToDevice, s));
        }
        if (nbodies) {
            LIFE_HIP(hipMemcpyAsync(g->body_first, first.data(), sizeof(int) * nbodies, hipMemcpyHostToDevice, s));
            LIFE_HIP(hipMemcpyAsync(g->body_n, cnt.data(), sizeof(int) * nbodies, hipMemcpyHostToDevice, s));
            LIFE_HIP(hipMemcpyAsync(g->body_flex, body_flex, sizeof(int) * nbodies, hipMemcpyHostToDevice, s));
        }
        launch_fill_f64(m.force, 0.0, 2 * (long long)n, s);
        launch_fill_f64(m.eps, 0.0, n, s);
        launch_fill_f64(m.ds, 0.0, n, s);
        launch_fill_i32(m.supp_cnt, 0, n, s);
        launch_fill_i32(g->flags, 0, 2, s);
        launch_fill_i32(sc.site_count, 0, g->g.N, s);
        launch_fill_i32(sc.uniq_n, 0, 1, s);
        LIFE_HIP(hipGetLastError());
        LIFE_HIP(hipStreamSynchronize(s));
        g->dense_clear_next = true;
    });
}

int life_ibm_update_markers(life_grid* g, const double* pos, const double* vel) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        const size_t n = (size_t)g->md.n;
        if (pos) LIFE_HIP(hipMemcpyAsync(g->md.pos, pos, sizeof(double) * 2 * n, hipMemcpyHostToDevice, g->stream));
        if (vel) LIFE_HIP(hipMemcpyAsync(g->md.vel, vel, sizeof(double) * 2 * n, hipMemcpyHostToDevice, g->stream));
    });
}

int life_ibm_find_support(life_grid* g, int flexible_only) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        ibm_find_support(g, flexible_only);
        check_flags(g);
    });
}

int life_ibm_epsilon_matrix(life_grid* g, int first, int n, double* A) {
    return guarded([&] {
        if (first < 0 || n < 0 || first + n > g->md.n) throw std::runtime_error("epsilon system out of range");
        LIFE_HIP(hipSetDevice(g->device));
        ibm_epsilon_matrix(g, first, n, A);
    });
}

int life_ibm_set_epsilon(life_grid* g, const double* eps) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        LIFE_HIP(hipMemcpyAsync(g->md.eps, eps, sizeof(double) * g->md.n, hipMemcpyHostToDevice, g->stream));
        LIFE_HIP(hipStreamSynchronize(g->stream));
    });
}

int life_ibm_interpolate(life_grid* g, double* force) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        ibm_interpolate(g);
        if (force) {
            LIFE_HIP(hipMemcpyAsync(force, g->md.force, sizeof(double) * 2 * g->md.n, hipMemcpyDeviceToHost,
                                    g->stream));
            LIFE_HIP(hipStreamSynchronize(g->stream));
        }
    });
}

int life_ibm_spread(life_grid* g) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        ibm_spread(g);
    });
}

int life_ibm_download(life_grid* g, double* pos, double* vel, double* force, double* eps, double* ds, int* supp_count,
                      int* supp_ij, double* supp_delta) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(g->device));
        const size_t n = (size_t)g->md.n;
        hipStream_t s = g->stream;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) LIFE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        };
        cp(pos, g->md.pos, 16 * n);
        cp(vel, g->md.vel, 16 * n);
        cp(force, g->md.force, 16 * n);
        cp(eps, g->md.eps, 8 * n);
        cp(ds, g->md.ds, 8 * n);
        cp(supp_count, g->md.supp_cnt, 4 * n);
        cp(supp_ij, g->md.supp_ij, 4 * n * kSupp * 2);
        cp(supp_delta, g->md.supp_delta, 8 * n * kSupp);
        LIFE_HIP(hipStreamSynchronize(s));
        check_flags(g);
    });
}

}  // extern "C"
