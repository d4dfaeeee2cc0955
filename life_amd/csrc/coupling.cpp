// Layer 2: the library's own host driver for one LIFE case (life_case_*).
//
// Owns a life_grid, the bodies read from geometry.config text, their FP64
// FEM and the strongly coupled subiteration loop, and runs the reference's
// main-loop body (GridClass::solver, Grid.cpp:25-33) on top of the layer-1
// entry points.  FEM, the epsilon solve and the Aitken relaxation stay FP64
// on the host; everything on the lattice runs on the device.
//
// Compiled with -ffp-contract=off.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "case_impl.h"
#include "lu.h"

using namespace life;

namespace {

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

// IBMBodyClass circle constructor, IBMBody.cpp:45-76.
void build_circle(life_case* c, int ID, const double pos[2], double radius) {
    Body b;
    b.ID = ID;
    b.flex = 1;
    b.kind = 0;
    const int n = (int)std::floor(2.0 * M_PI * radius / c->cs.Dx);
    for (int i = 0; i < n; i++) {
        b.mpos.push_back(pos[0] + radius * std::cos(i * 2.0 * M_PI / n));
        b.mpos.push_back(pos[1] + radius * std::sin(i * 2.0 * M_PI / n));
    }
    c->bodies.push_back(std::move(b));
}

// IBMBodyClass filament constructor, IBMBody.cpp:79-135.
void build_filament(life_case* c, int ID, const double pos[2], const double geom[2], double angleDeg,
                    const std::string& flex, const std::string& nEl, const std::string& bc, double rho, double E) {
    Body b;
    b.ID = ID;
    if (flex == "FLEXIBLE") b.flex = 0;
    else if (flex == "RIGID") b.flex = 1;
    else throw std::runtime_error("Not a proper flexible type set");
    b.kind = 1;
    const double L = geom[0];
    const double ang = angleDeg * M_PI / 180.0;
    const double ca = std::cos(ang), sa = std::sin(ang);
    const int n = (int)std::floor(L / c->cs.Dx) + 1;
    for (int i = 0; i < n; i++) {
        const double px = i * L / (n - 1), py = 0.0;
        const double rx = 0.0 + ca * px + (-sa) * py, ry = 0.0 + sa * px + ca * py;
        b.mpos.push_back(pos[0] + rx);
        b.mpos.push_back(pos[1] + ry);
    }
    b.fpos[0] = pos[0];
    b.fpos[1] = pos[1];
    b.fgeom[0] = geom[0];
    b.fgeom[1] = geom[1];
    b.fangle = ang;
    b.fnel = nEl;
    b.fbc = bc;
    b.frho = rho;
    b.fE = E;
    c->bodies.push_back(std::move(b));
}

// ObjectsClass::geometryReadIn, Objects.cpp:324-488, on text.
void parse_geometry(life_case* c, const char* text) {
    if (!text) return;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        size_t k = line.find_first_not_of(" \t\r");
        if (k == std::string::npos || line[k] == '#') continue;
        std::istringstream ls(line.substr(k));
        std::string kw;
        int number, ID;
        double st[2], sp[2];
        if (!(ls >> kw >> number >> ID >> st[0] >> st[1] >> sp[0] >> sp[1]))
            throw std::runtime_error("malformed geometry line: " + line);
        if (kw == "CIRCLE") {
            double r;
            if (!(ls >> r)) throw std::runtime_error("malformed CIRCLE line: " + line);
            for (int i = 0; i < number; i++) {
                const double pos[2] = {st[0] + i * sp[0], st[1] + i * sp[1]};
                build_circle(c, ID++, pos, r);
            }
        } else if (kw == "FILAMENT") {
            double g[2], ang, rho, E;
            std::string flex, nel, bc;
            if (!(ls >> g[0] >> g[1] >> ang >> flex >> nel >> bc >> rho >> E))
                throw std::runtime_error("malformed FILAMENT line: " + line);
            for (int i = 0; i < number; i++) {
                const double pos[2] = {st[0] + i * sp[0], st[1] + i * sp[1]};
                build_filament(c, ID++, pos, g, ang, flex, nel, bc, rho, E);
            }
        } else {
            throw std::runtime_error("unknown body keyword '" + kw + "'");
        }
    }
    for (size_t i = 0; i < c->bodies.size(); i++)
        for (size_t j = i + 1; j < c->bodies.size(); j++)
            if (c->bodies[i].ID == c->bodies[j].ID) throw std::runtime_error("Duplicate body IDs in geometry.config");
}

// Lay the markers out in global order and drop overlapping ones
// (ObjectsClass::removeOverlapMarkers, Objects.cpp:867-938).
void finalise_markers(life_case* c) {
    std::vector<double> P;
    std::vector<int> B;
    for (size_t b = 0; b < c->bodies.size(); b++)
        for (size_t k = 0; k < c->bodies[b].mpos.size() / 2; k++) {
            P.push_back(c->bodies[b].mpos[2 * k]);
            P.push_back(c->bodies[b].mpos[2 * k + 1]);
            B.push_back((int)b);
        }
    const int nm = (int)B.size();
    std::vector<char> erase(nm, 0);
    for (int a = 0; a < nm; a++)
        for (int q = a + 1; q < nm; q++) {
            const double d[2] = {P[2 * a] - P[2 * q], P[2 * a + 1] - P[2 * q + 1]};
            const double mag = std::sqrt(0.0 + d[0] * d[0] + d[1] * d[1]) / c->cs.Dx;
            if (mag < 0.5) {
                const int fa = c->bodies[B[a]].flex, fb = c->bodies[B[q]].flex;
                if (fa == 0 && fb == 0) throw std::runtime_error("Cannot place two flexible bodies so close to each other");
                erase[(fa == 0 && fb == 1) ? q : a] = 1;
            }
        }
    for (auto& b : c->bodies) b.n = 0;
    c->pos.reserve(2 * (size_t)nm);  // one pinned allocation, not one per growth
    for (int a = 0; a < nm; a++) {
        if (erase[a]) continue;
        Body& b = c->bodies[B[a]];
        if (b.n == 0) b.first = (int)c->mbody.size();
        b.n++;
        c->pos.push_back(P[2 * a]);
        c->pos.push_back(P[2 * a + 1]);
        c->mbody.push_back(B[a]);
    }
    c->nm = (int)c->mbody.size();
    c->vel.assign(2 * (size_t)c->nm, 0.0);
    c->force.assign(2 * (size_t)c->nm, 0.0);
    c->eps.assign(c->nm, 0.0);
    c->ds.assign(c->nm, 0.0);
    for (auto& b : c->bodies) b.mpos.clear();
}

// ObjectsClass::computeEpsilon (Objects.cpp:235-321).  UNI_EPSILON and the
// initial call: dense device matrix, host canonical solve of A eps = 1.
// Later calls (flexible bodies only): banded systems assembled and solved
// on the device, with the dense host path for any system that is not banded.
// copy_flags: queue the copy of the banded systems' out-of-band flags to the
// host (the device FEM's last workgroup brings them over otherwise).
void compute_epsilon(life_case* c, bool initial, bool copy_flags = true) {
    auto solve = [&](int first, int n) { case_solve_eps_dense(c, first, n); };
    if (c->p.uni_epsilon) {
        solve(0, c->nm);
    } else if (initial) {
        for (auto& b : c->bodies)
            if (b.n > 0) solve(b.first, b.n);
    } else {
        // systems found not to be banded are solved when the forces are
        // downloaded (finish_epsilon); the interpolation in between does
        // not read epsilon
        ibm_epsilon_band_launch(c->g, copy_flags);
        c->eps_check_pending = true;
    }
}

// Dense solves of the systems that the banded path left (after a
// synchronisation of the grid stream).
bool finish_epsilon(life_case* c) {
    if (!c->eps_check_pending) return false;
    c->eps_check_pending = false;
    const std::vector<int> dense = ibm_epsilon_band_dense(c->g);
    for (int k : dense) case_solve_eps_dense(c, c->g->eps_first_h[k], c->g->eps_n_h[k]);
    return !dense.empty();
}

// Download the marker forces and epsilon; returns once they are on the
// host.  With spread set, the spread is queued behind the copies and runs
// on the device while the host goes on (the FEM does not change what it
// reads: supports, forces, epsilon); the return value says whether that
// spread used the final epsilon (false after a dense fallback solve, which
// rewrites epsilon).
bool download_forces(life_case* c, bool spread) {
    life_grid* g = c->g;
    LIFE_HIP(hipMemcpyAsync(c->force.data(), g->md.force, sizeof(double) * 2 * c->nm, hipMemcpyDeviceToHost,
                            g->stream));
    LIFE_HIP(hipMemcpyAsync(c->eps.data(), g->md.eps, sizeof(double) * c->nm, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipEventRecord(c->ev_forces, g->stream));
    if (spread) ibm_spread(g);
    LIFE_HIP(hipEventSynchronize(c->ev_forces));
    return !finish_epsilon(c) && spread;
}

// Run fn(body) for every flexible body.  Bodies are independent (disjoint
// FEM state and marker ranges), so they run on host threads; anything
// order-dependent is reduced afterwards in body order.
template <typename Fn>
void for_flexible(life_case* c, Fn&& fn) {
    const int nf = (int)c->flex_bodies.size();
#pragma omp parallel for schedule(dynamic, 1) if (nf > 1)
    for (int q = 0; q < nf; q++) fn(*c->flex_bodies[q]);
}

}  // namespace

void life::case_aitken(life_case* c) {
    if (c->subIt == 1) {
        const double a = std::fabs(c->relax), r = a < c->p.relaxMax ? a : c->p.relaxMax;
        c->relax = (double)(((0.0 < c->relax) - (c->relax < 0.0)) * r);
    } else {
        c->relax = -c->relax * c->subNum / c->subDen;
    }
}

// Host part of ObjectsClass::recomputeObjectVals, Objects.cpp:152-232.
void life::case_recompute_host(life_case* c) {
    if (c->subIt == 0) {
        for_flexible(c, [&](Body& b) {
            b.fem->reset_values();
            b.fem->predictor(c->t, c->view(b));
        });
    } else {
        case_aitken(c);
        const double w = c->relax;
        for_flexible(c, [&](Body& b) { b.fem->relax(w, c->view(b)); });
    }
}

void life::case_sync_host(life_case* c) {
    if (!c->host_stale) return;
    life_grid* g = c->g;
    LIFE_HIP(hipSetDevice(g->device));
    lbm_join(g);
    std::vector<FemBody*> fb;
    for (Body* b : c->flex_bodies) fb.push_back(b->fem.get());
    fem_dev_download(c->fdev, fb, g->stream);
    const size_t b2 = sizeof(double) * 2 * c->nm, b1 = sizeof(double) * c->nm;
    LIFE_HIP(hipMemcpyAsync(c->pos.data(), g->md.pos, b2, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->vel.data(), g->md.vel, b2, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->force.data(), g->md.force, b2, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->eps.data(), g->md.eps, b1, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->ds.data(), g->md.ds, b1, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream));
    c->host_stale = false;
}

void life::case_push_fem(life_case* c) {
    if (!c->fdev.ok()) return;
    std::vector<const FemBody*> fb;
    for (Body* b : c->flex_bodies) fb.push_back(b->fem.get());
    fem_dev_upload(c->fdev, fb, c->g->stream);
    c->host_stale = false;
}

// ObjectsClass::femKernel, Objects.cpp:63-99 (ORDERED: sums in body order).
void life::case_fem(life_case* c) {
    for_flexible(c, [&](Body& b) { b.fem->dynamic(c->view(b)); });
    double res = 0.0, num = 0.0, den = 0.0;
    for (Body* b : c->flex_bodies) {
        res += b->fem->sub_res();
        num += b->fem->sub_num();
        den += b->fem->sub_den();
    }
    c->subRes = std::sqrt(res) / (c->p.ref_L * std::sqrt((double)c->simDOFs));
    c->subNum = num;
    c->subDen = den;
    c->subIt++;
}

// computeEpsilon's solve of one system by the dense path.
void life::case_solve_eps_dense(life_case* c, int first, int n) {
    c->A.resize((size_t)n * n);
    ibm_epsilon_matrix(c->g, first, n, c->A.data());
    c->rhs.assign(n, 1.0);
    c->sol.assign(n, 0.0);
    lu_solve_sparse(n, 0, c->A.data(), c->rhs.data(), c->sol.data(), c->luw, c->luiw);
    std::memcpy(&c->eps[first], c->sol.data(), sizeof(double) * n);
    LIFE_HIP(hipMemcpyAsync(c->g->md.eps + first, &c->eps[first], sizeof(double) * n, hipMemcpyHostToDevice,
                            c->g->stream));
}

void life::case_restart_objects(life_case* c) {
    life_grid* g = c->g;
    LIFE_HIP(hipSetDevice(g->device));
    LIFE_HIP(hipEventSynchronize(c->ev_upload));
    const size_t b2 = sizeof(double) * 2 * c->nm;
    LIFE_HIP(hipMemcpyAsync(g->md.pos, c->pos.data(), b2, hipMemcpyHostToDevice, g->stream));
    LIFE_HIP(hipMemcpyAsync(g->md.vel, c->vel.data(), b2, hipMemcpyHostToDevice, g->stream));
    LIFE_HIP(hipMemcpyAsync(g->md.force, c->force.data(), b2, hipMemcpyHostToDevice, g->stream));
    ibm_find_support(g, 0);
    compute_epsilon(c, false);
    LIFE_HIP(hipStreamSynchronize(g->stream));
    finish_epsilon(c);
    LIFE_HIP(hipMemcpyAsync(c->eps.data(), g->md.eps, sizeof(double) * c->nm, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->ds.data(), g->md.ds, sizeof(double) * c->nm, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream));
    check_flags(g);
}

namespace {

// ObjectsClass::recomputeObjectVals, Objects.cpp:152-232.
void recompute_object_vals(life_case* c) {
    // the host rewrites pos/vel: the previous step's upload must be done
    LIFE_HIP(hipEventSynchronize(c->ev_upload));
    case_recompute_host(c);
    life_grid* g = c->g;
    LIFE_HIP(hipMemcpyAsync(g->md.pos, c->pos.data(), sizeof(double) * 2 * c->nm, hipMemcpyHostToDevice, g->stream));
    LIFE_HIP(hipMemcpyAsync(g->md.vel, c->vel.data(), sizeof(double) * 2 * c->nm, hipMemcpyHostToDevice, g->stream));
    ibm_find_support(g, 1);
    compute_epsilon(c, false);
}

// ObjectsClass::objectKernel, Objects.cpp:26-60.
// Pending early stream-collide of the next step (launch point, events).
struct EarlyBulk {
    bool want = false;
    hipEvent_t begin = nullptr, end = nullptr;
};

void launch_pending(life_case* c, EarlyBulk& eb) {
    if (!eb.want) return;
    lbm_bulk_early(c->g, eb.begin, eb.end);
    eb.want = false;
}

// How long the host polls for the device FEM's sums before it waits on the
// launch's event instead (microseconds; environment LIFE_SUMS_SPIN_US, 0:
// the event wait alone).
double sums_spin_us() {
    static const double v = [] {
        const char* e = std::getenv("LIFE_SUMS_SPIN_US");
        return e && *e ? std::atof(e) : 20000.0;
    }();
    return v;
}

// objectKernel with the FEM on the device: per subiteration the host only
// reads the three residual sums (and the out-of-band flags of epsilon).
void object_kernel_dev(life_case* c, EarlyBulk& eb, int mode) {
    life_grid* g = c->g;
    FemDevCase& fd = c->fdev;
    PhaseTimes& T = *c->timing;
    c->subIt = 0;
    c->forces_on_host = false;
    bool more = false, spread_done = false;
    do {
        // recomputeObjectVals: predictor or relaxation moves the markers
        if (c->subIt == 0) {
            launch_fem_predict(fd.d_bodies, nullptr, fd.nb, c->t, nullptr, fd.lds, g->stream);
        } else {
            case_aitken(c);
            launch_fem_relax(fd.d_bodies, nullptr, fd.nb, c->relax, nullptr, fd.lds, g->stream);
        }
        LIFE_HIP(hipGetLastError());
        ibm_find_support(g, 1);
        // the out-of-band flags reach the host with the FEM's sums
        compute_epsilon(c, false, false);
        T.lap(g, 1);
        ibm_interpolate(g);
        if (mode == 2) launch_pending(c, eb);
        T.lap(g, 2);
        // femKernel; again with epsilon from the dense path if a system was
        // found not to be banded
        bool redo = false;
        FemSumsOut so;
        so.done = fd.done;
        so.sums_h = fd.sums_h;
        if (c->eps_check_pending) {
            so.flags = g->eps_outside;
            so.flags_h = g->eps_outside_h;
            so.nflags = g->eps_nsys;
        }
        // The final spread (ObjectsClass::ibmKernelSpread) takes the forces
        // of the last interpolation and the supports found before the last
        // FEM solve, so it does not wait for that solve's residual: when
        // this subiteration is predicted to be the last, it is queued
        // behind the FEM and runs while the host reads the residual.  If
        // another subiteration follows after all, its spread clears every
        // site of this one and writes its own, so the result is the same.
        const bool spec = c->spec_spread && c->subIt + 1 >= c->prev_subits;
        for (;;) {
            so.seq = fd.seq += 1.0;
            launch_fem_dynamic(fd.d_bodies, nullptr, fd.nb, fd.lds, redo, g->stream, so);
            LIFE_HIP(hipEventRecord(c->ev_forces, g->stream));
            if (spec) ibm_spread(g);
            spread_done = spec;
            // the host polls the sequence word the launch writes last (an
            // event wait wakes the host tens of microseconds later)
            fem_wait_sums(so, c->ev_forces, sums_spin_us());
            // a redo changes epsilon: the spread is queued again behind it
            if (!finish_epsilon(c)) break;
            redo = true;
        }
        T.subits++;
        T.lap(g, 3);
        // FEMBodyClass::subResidual summed over the bodies (case_fem)
        c->subRes = std::sqrt(fd.sums_h[0]) / (c->p.ref_L * std::sqrt((double)c->simDOFs));
        c->subNum = fd.sums_h[1];
        c->subDen = fd.sums_h[2];
        c->subIt++;
        more = case_subiterate(c);
    } while (more);
    if (!spread_done) ibm_spread(g);
    c->prev_subits = c->subIt;
    T.lap(g, 4);
    c->host_stale = true;
    if (c->subIt == kMaxSubIt) throw std::runtime_error("Subiteration scheme hit 20 iterations");
}

void object_kernel(life_case* c, EarlyBulk& eb, int mode) {
    if (c->fdev.ok()) {
        object_kernel_dev(c, eb, mode);
        return;
    }
    c->subIt = 0;
    c->forces_on_host = false;
    PhaseTimes& T = *c->timing;
    if (!c->hasFlex && mode != 2 && !T.on) {
        // rigid bodies only: the spread follows the interpolation with
        // nothing between them (one launch for a small case)
        ibm_interp_spread(c->g);
        T.lap(c->g, 4);
        return;
    }
    bool spread_final = false, more = false;
    do {
        if (c->hasFlex) recompute_object_vals(c);
        T.lap(c->g, 1);
        ibm_interpolate(c->g);
        // mode 2: the next step's stream-collide starts once the first
        // pass of the coupling kernels is queued, beside the host FEM
        if (mode == 2) launch_pending(c, eb);
        if (!c->hasFlex) break;
        // the spread with these forces starts beside the host FEM: if the
        // FEM ends the subiterations they are the final forces; otherwise
        // the spread is done again below, and a spread overwrites every
        // site of the previous one, so the result is the same
        spread_final = download_forces(c, true);
        T.lap(c->g, 2);
        T.subits++;
        case_fem(c);
        T.lap(c->g, 3);
        more = case_subiterate(c);
    } while (more);
    if (!spread_final) ibm_spread(c->g);
    T.lap(c->g, 4);
    if (c->hasFlex) {
        // the last FEM solve moved the markers (FEMBody.cpp:63-64); the
        // spread above used the supports found before it, as the reference
        // does, and the device copy now takes the final positions for output
        LIFE_HIP(hipMemcpyAsync(c->g->md.pos, c->pos.data(), sizeof(double) * 2 * c->nm, hipMemcpyHostToDevice,
                                c->g->stream));
        LIFE_HIP(hipMemcpyAsync(c->g->md.vel, c->vel.data(), sizeof(double) * 2 * c->nm, hipMemcpyHostToDevice,
                                c->g->stream));
        LIFE_HIP(hipEventRecord(c->ev_upload, c->g->stream));
        c->forces_on_host = true;
    }
    if (c->subIt == kMaxSubIt) throw std::runtime_error("Subiteration scheme hit 20 iterations");
}

// The bulk stream-collide of step t+1 reads only the populations after
// step t and the uniform force, so it is launched on the grid's second
// stream before the coupling of step t and overlaps it (disabled with
// LIFE_OVERLAP=0).  Support sites and boundaries of step t+1 still follow
// the spread of step t, as in the reference order.
// LIFE_OVERLAP = 0 (off), 1 (launch right after this step's lattice
// kernels) or 2 (after the first interpolation of the coupling); unset:
// 1 for cases with flexible bodies on lattices of at least 2^22 sites, else
// 0.  On smaller lattices, or with rigid bodies only, the stream-collide is
// too short, or the coupling too light, to pay for the cross-stream waits
// (measured: rigid cylinder on 1024x256 -24%, cylinder + plate on 2048x410
// unchanged, canopy on 8192x4096 +35%).  Read at every step, so a caller can
// switch it between calls.
constexpr long long kOverlapMinSites = 1LL << 22;
int overlap_mode(const life_case* c) {
    // a per-site force (WOMERSLEY with gravity) lives in two planes that
    // alternate by step: an early stream-collide would overwrite the one
    // this step's coupling reads
    if (c->g->fsite[0]) return 0;
    const char* e = std::getenv("LIFE_OVERLAP");
    if (e) {
        const int v = std::atoi(e);
        return v >= 0 && v <= 2 ? v : 1;
    }
    return c->hasFlex && (long long)c->p.Nx * c->p.Ny >= kOverlapMinSites ? 1 : 0;
}

// One main-loop step; next_begin/next_end (may be null) are recorded
// around the next step's stream-collide when it is launched early
// (more = a next step follows in this call).
void case_step(life_case* c, hipEvent_t ev_begin, hipEvent_t ev_end, bool more, hipEvent_t next_begin,
               hipEvent_t next_end) {
    PhaseTimes& T = *c->timing;
    T.start(c->g);
    c->t++;
    lbm_step(c->g, c->t, ev_begin, ev_end);
    if (T.on) lbm_join(c->g);  // the lattice phase ends with the step's tail, wherever it ran
    T.lap(c->g, 0);
    // without immersed bodies there is nothing to overlap the early
    // stream-collide with: the plain order avoids the cross-stream waits
    const int mode = T.on || !c->hasIBM ? 0 : overlap_mode(c);
    EarlyBulk eb{more && mode != 0, next_begin, next_end};
    if (mode == 1) launch_pending(c, eb);
    if (c->hasIBM) object_kernel(c, eb, mode);
    launch_pending(c, eb);
    T.t[5] += 1;
}

// The blow-up test of the reference's writeInfo (Grid.cpp:562-581) at the
// grid's cadence, after step c->t.
void case_blowup_cadence(life_case* c) {
    life_grid* g = c->g;
    if (g->check_every <= 0 || c->t % g->check_every != 0) return;
    lbm_join(g);
    blowup_cadence(g, c->t, g->stream);
}

// The FEM runs on the device for FP32-shifted populations unless the
// environment sets LIFE_DEVICE_FEM = 0 (or a body does not fit the device
// FEM's LDS).  The plain method keeps the host FEM, whose transcendental
// functions match the oracle's bit for bit, unless LIFE_DEVICE_FEM = 1 asks
// for the device FEM (which then agrees with the oracle to rounding).
bool device_fem_wanted(int precision) {
    const char* e = std::getenv("LIFE_DEVICE_FEM");
    const bool set = e && *e;
    if (precision == LIFE_FP32_SHIFTED) return !(set && std::atoi(e) == 0);
    return set && std::atoi(e) != 0;
}

// CUs kept free of the bulk stream-collide for a device-FEM coupling chain
// (loaded canopy, 8192x4096, two subiterations per step: 0.69 ms/step with
// no reserve, 0.60 with 16 or 32 CUs), where the XCD split below is not
// available.
constexpr int kDeviceFemCus = 32;
// The device-FEM coupling chain on two whole XCDs and the early bulk
// stream-collide on the other six (loaded canopy: 0.543-0.545 ms/step
// against 0.559-0.560 with the 32-CU reserve, which lets the chain's
// kernels share CUs with the bulk kernel; one or three XCDs for the chain:
// 0.677 and 0.571).
constexpr int kDeviceFemXcds = 2;

void setup_device_fem(life_case* c) {
    std::vector<const FemBody*> fb;
    std::vector<int> first, nm;
    for (Body* b : c->flex_bodies) {
        if (fem_dev_lds_bytes(b->fem->ndof(), b->n) > kFemMaxLds) return;  // too large: host FEM
        fb.push_back(b->fem.get());
        first.push_back(b->first);
        nm.push_back(b->n);
    }
    const life_params* p = &c->p;
    const FemDevScalars s{c->cs.Dx, c->cs.Dt, c->cs.Dm, p->alpha, p->delta, p->gravityX, p->gravityY, p->ref_L};
    life_grid* g = c->g;
    fem_dev_create(c->fdev, fb, first, nm, g->md.pos, g->md.vel, g->md.force, g->md.eps, s, g->stream);
    // With the FEM on the device the whole coupling chain competes with the
    // early bulk stream-collide for CUs: the two get XCDs of their own.
    grid_split_coupling(g, kDeviceFemXcds, kDeviceFemCus);
}

}  // namespace

extern "C" {

int life_case_create(const life_params* p, const char* geometry, int precision, int device, life_case** out) {
    *out = nullptr;
    life_case* c = new life_case();
    int rc = guarded([&] {
        c->p = *p;
        c->cs = derive_scalars(*p);
        c->relax = p->relaxMax;
        parse_geometry(c, geometry);
        finalise_markers(c);
        c->hasIBM = !c->bodies.empty();
        FemScalars fs{c->cs.Dx, c->cs.Dt, c->cs.Dm, p->alpha, p->delta, p->gravityX, p->gravityY, p->ref_L};
        for (auto& b : c->bodies) {
            if (b.flex == 0) {
                c->hasFlex = true;
                b.fem.reset(new FemBody(fs, b.fpos, b.fgeom, b.fangle, b.fnel, b.fbc, b.frho, b.fE, b.n));
                c->simDOFs += b.fem->ndof();
                c->flex_bodies.push_back(&b);
            }
        }
        // ObjectsClass::initialDeflect, Objects.cpp:943-947, 964-976
        if (p->initial_deflect != 0.0)
            for (Body* b : c->flex_bodies) b->fem->set_initial_deflection(p->initial_deflect, c->view(*b));
        if (life_grid_create(p, precision, device, &c->g)) throw std::runtime_error(life_last_error());
        LIFE_HIP(hipEventCreateWithFlags(&c->ev_upload, hipEventDisableTiming));
        LIFE_HIP(hipEventCreateWithFlags(&c->ev_forces, hipEventDisableTiming));
        if (c->hasIBM) {
            std::vector<int> flex(c->bodies.size());
            for (size_t b = 0; b < c->bodies.size(); b++) flex[b] = c->bodies[b].flex;
            if (life_ibm_set_markers(c->g, c->nm, c->pos.data(), c->vel.data(), c->mbody.data(),
                                     (int)c->bodies.size(), flex.data()))
                throw std::runtime_error(life_last_error());
            // ObjectsClass::initialiseObjects, Objects.cpp:941-961
            ibm_find_support(c->g, 0);
            compute_epsilon(c, true);
            if (c->hasFlex && !p->uni_epsilon) {
                std::vector<int> first, n;
                for (auto& b : c->bodies)
                    if (b.flex == 0 && b.n > 0) {
                        first.push_back(b.first);
                        n.push_back(b.n);
                    }
                ibm_epsilon_band_setup(c->g, first, n, 8);
            }
            LIFE_HIP(hipMemcpyAsync(c->ds.data(), c->g->md.ds, sizeof(double) * c->nm, hipMemcpyDeviceToHost,
                                    c->g->stream));
            LIFE_HIP(hipStreamSynchronize(c->g->stream));
            check_flags(c->g);
            if (c->hasFlex && device_fem_wanted(precision)) setup_device_fem(c);
            if (const char* e = std::getenv("LIFE_SPEC_SPREAD")) c->spec_spread = std::atoi(e) != 0;
        }
    });
    if (rc) {
        life_case_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int life_case_destroy(life_case* c) {
    if (!c) return 0;
    int rc = 0;
    c->timing->report();
    if (c->g) {
        (void)hipSetDevice(c->g->device);
        for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
        if (c->ev_upload) (void)hipEventDestroy(c->ev_upload);
        if (c->ev_forces) (void)hipEventDestroy(c->ev_forces);
        (void)hipStreamSynchronize(c->g->stream);
        fem_dev_destroy(c->fdev);
        rc = life_grid_destroy(c->g);
    }
    delete c;
    return rc;
}

int life_case_step(life_case* c, int nsteps) {
    return guarded([&] {
        LIFE_HIP(hipSetDevice(c->g->device));
        for (int s = 0; s < nsteps; s++) {
            case_step(c, nullptr, nullptr, s + 1 < nsteps, nullptr, nullptr);
            case_blowup_cadence(c);
        }
        lbm_join(c->g);
        LIFE_HIP(hipStreamSynchronize(c->g->stream));
        blowup_check(c->g, true);
        if (c->hasIBM) check_flags(c->g);
    });
}

int life_case_set_check_interval(life_case* c, int every) {
    c->g->check_every = every > 0 ? every : 0;
    return 0;
}

int life_case_step_timed(life_case* c, int nsteps, double* ms_total, double* ms_kernel) {
    return guarded([&] {
        life_grid* g = c->g;
        LIFE_HIP(hipSetDevice(g->device));
        LIFE_HIP(hipStreamSynchronize(g->stream));
        // one event pair per step around the stream-collide launch, read
        // after the last step so the host never waits inside the loop.
        // They order nothing, so they are recorded without the system-scope
        // fence (cache writeback and invalidate) that a default event adds
        // and that the untimed step does not have
        while ((int)c->kev.size() < 2 * nsteps) {
            hipEvent_t e;
            LIFE_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            c->kev.push_back(e);
        }
        LIFE_HIP(hipEventRecord(g->ev_a, g->stream));
        // a step's stream-collide events are recorded where it is launched:
        // in the step itself, or during the previous step when launched early
        for (int s = 0; s < nsteps; s++) {
            const bool early = g->bulk_pending;
            hipEvent_t nb = s + 1 < nsteps ? c->kev[2 * (s + 1)] : nullptr;
            hipEvent_t ne = s + 1 < nsteps ? c->kev[2 * (s + 1) + 1] : nullptr;
            case_step(c, early ? nullptr : c->kev[2 * s], early ? nullptr : c->kev[2 * s + 1], s + 1 < nsteps, nb,
                      ne);
            case_blowup_cadence(c);
        }
        lbm_join(g);
        LIFE_HIP(hipEventRecord(g->ev_b, g->stream));
        LIFE_HIP(hipEventSynchronize(g->ev_b));
        blowup_check(g, true);
        float ms = 0.0f;
        LIFE_HIP(hipEventElapsedTime(&ms, g->ev_a, g->ev_b));
        double kms = 0.0;
        for (int s = 0; s < nsteps; s++) {
            float k = 0.0f;
            LIFE_HIP(hipEventElapsedTime(&k, c->kev[2 * s], c->kev[2 * s + 1]));
            kms += k;
        }
        if (ms_total) *ms_total = ms;
        if (ms_kernel) *ms_kernel = kms;
        if (c->hasIBM) check_flags(g);
    });
}

int life_case_time(const life_case* c, int* t) {
    *t = c->t;
    return 0;
}

life_grid* life_case_grid(life_case* c) { return c->g; }

int life_case_num_markers(const life_case* c, int* n) {
    *n = c->nm;
    return 0;
}

int life_case_num_bodies(const life_case* c, int* n) {
    *n = (int)c->bodies.size();
    return 0;
}

int life_case_body_info(const life_case* c, int b, int info[5]) {
    return guarded([&] {
        if (b < 0 || b >= (int)c->bodies.size()) throw std::runtime_error("body index out of range");
        const Body& B = c->bodies[b];
        info[0] = B.ID;
        info[1] = B.flex;
        info[2] = B.n;
        info[3] = B.n ? B.first : -1;
        info[4] = B.fem ? B.fem->ndof() : 0;
    });
}

int life_case_get_fem(const life_case* c, int b, double* U, double* Udot, double* Udotdot) {
    return guarded([&] {
        if (b < 0 || b >= (int)c->bodies.size() || !c->bodies[b].fem) throw std::runtime_error("not a flexible body");
        case_sync_host(const_cast<life_case*>(c));
        const FemBody& F = *c->bodies[b].fem;
        const size_t n = (size_t)F.ndof() * sizeof(double);
        if (U) std::memcpy(U, F.U.data(), n);
        if (Udot) std::memcpy(Udot, F.Ud.data(), n);
        if (Udotdot) std::memcpy(Udotdot, F.Udd.data(), n);
    });
}

int life_case_coupling_info(const life_case* c, double out[3]) {
    out[0] = c->subIt;
    out[1] = c->subRes;
    out[2] = c->relax;
    return 0;
}

int life_case_total_force(life_case* c, double out[4]) {
    return guarded([&] {
        life_grid* g = c->g;
        LIFE_HIP(hipSetDevice(g->device));
        case_sync_host(c);
        if (c->nm) {
            LIFE_HIP(hipMemcpyAsync(c->force.data(), g->md.force, sizeof(double) * 2 * c->nm, hipMemcpyDeviceToHost,
                                    g->stream));
            LIFE_HIP(hipMemcpyAsync(c->ds.data(), g->md.ds, sizeof(double) * c->nm, hipMemcpyDeviceToHost, g->stream));
            LIFE_HIP(hipStreamSynchronize(g->stream));
        }
        // ObjectsClass::writeTotalForces, Objects.cpp:517-534
        const CaseScalars& s = c->cs;
        const double scale = s.Dm * s.Dx / (s.Dt * s.Dt) * 1.0 / s.Dx;
        double F[2] = {0.0, 0.0};
        for (int n = 0; n < c->nm; n++)
            for (int d = 0; d < 2; d++) F[d] -= c->force[2 * n + d] * 1.0 * c->eps[n] * c->ds[n] * scale;
        const double ND = 0.5 * c->p.ref_rho * (c->p.ref_U * c->p.ref_U) * c->p.ref_L;
        out[0] = F[0];
        out[1] = F[1];
        out[2] = F[0] / ND;
        out[3] = F[1] / ND;
    });
}

}  // extern "C"
