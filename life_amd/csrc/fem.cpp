// FP64 corotational beam; see fem.h.  Compiled with -ffp-contract=off.
#include "fem.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace life {

namespace {

// Utils::shiftAngle, Utils.h:235-247: wrap into [-pi, pi).  fmod(x, y)
// is x itself for |x| < y (exactly, -0.0 included), so the library call is
// only made outside that range.
double wrap_angle(double a) {
    const double x = a + M_PI, y = 2.0 * M_PI;
    a = (x > -y && x < y) ? x : std::fmod(x, y);
    if (a < 0.0) a += 2.0 * M_PI;
    return a - M_PI;
}

double dot(const double* a, const double* b, int n) {
    double r = 0.0;
    for (int i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

// The element transformation T (FEMElement.cpp:255-262) is block diagonal:
// rows/columns {0,1} and {3,4} hold the rotation, 2 and 5 the identity.
// The reference multiplies the full 6x6 matrices, accumulating k = 0..5
// from 0.0; the terms with a structural zero of T add an exact zero, so
// summing only the structurally non-zero terms in increasing k gives the
// same values.  kNZ[i] lists the structurally non-zero k of row/column i.
constexpr int kNZ[6][2] = {{0, 1}, {0, 1}, {2, -1}, {3, 4}, {3, 4}, {5, -1}};

// Transpose(T) * A * T (loops fully unrolled, so the sparsity pattern is
// resolved at compile time)
void rotate6(const double T[6][6], const double A[6][6], double out[6][6]) {
    double tmp[6][6];
#pragma GCC unroll 6
    for (int i = 0; i < 6; i++)
#pragma GCC unroll 6
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
#pragma GCC unroll 2
            for (int q = 0; q < 2; q++)
                if (kNZ[i][q] >= 0) s += T[kNZ[i][q]][i] * A[kNZ[i][q]][j];
            tmp[i][j] = s;
        }
#pragma GCC unroll 6
    for (int i = 0; i < 6; i++)
#pragma GCC unroll 6
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
#pragma GCC unroll 2
            for (int q = 0; q < 2; q++)
                if (kNZ[j][q] >= 0) s += tmp[i][kNZ[j][q]] * T[kNZ[j][q]][j];
            out[i][j] = s;
        }
}

void mul_t6(const double T[6][6], const double* v, double* out) {  // Transpose(T) * v
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int q = 0; q < 2 && kNZ[i][q] >= 0; q++) s += T[kNZ[i][q]][i] * v[kNZ[i][q]];
        out[i] = s;
    }
}

void mul6(const double T[6][6], const double* v, double* out) {  // T * v
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int q = 0; q < 2 && kNZ[i][q] >= 0; q++) s += T[i][kNZ[i][q]] * v[kNZ[i][q]];
        out[i] = s;
    }
}

}  // namespace

FemBody::FemBody(const FemScalars& s, const double pos[2], const double geom[2], double angle,
                 const std::string& nEl, const std::string& bc, double rho, double E, int n_markers)
    : s_(s) {
    if (bc == "CLAMPED") nbc_ = 3;
    else if (bc == "SUPPORTED") nbc_ = 2;
    else throw std::runtime_error("unknown FEM boundary condition '" + bc + "'");
    angle0_ = angle;
    L0_ = geom[0];
    int nel;
    if (nEl == "CONFORMING") nel = (int)std::floor(L0_ / s.Dx);
    else nel = (int)std::strtod(nEl.c_str(), nullptr);
    if (nel < 1) throw std::runtime_error("filament needs at least one element");
    const int nnode = nel + 1;
    ndof_ = nnode * 3;
    tCrit_ = (L0_ / nel) / std::sqrt(E / rho);
    const double ca = std::cos(angle), sa = std::sin(angle);
    node_.resize(nnode);
    for (int i = 0; i < nnode; i++) {
        const double px = i * L0_ / nel, py = 0.0;
        const double rx = 0.0 + ca * px + (-sa) * py, ry = 0.0 + sa * px + ca * py;
        Node& n = node_[i];
        n.pos0[0] = n.pos[0] = pos[0] + rx;
        n.pos0[1] = n.pos[1] = pos[1] + ry;
        n.angle0 = n.angle = angle;
    }
    el_.resize(nel);
    for (int i = 0; i < nel; i++) {
        Elem& e = el_[i];
        e.L0 = e.L = L0_ / nel;
        e.angle = angle;
        e.A = s.Dx * geom[1];
        e.I = s.Dx * (geom[1] * geom[1] * geom[1]) / 12.0;
        e.E = E;
        e.rho = rho;
        e.n0 = i;
        e.n1 = i + 1;
        std::memset(e.T, 0, sizeof e.T);
        transform(e);
        local_matrices(e);
    }
    // FEMBodyClass::computeNodeMapping, FEMBody.cpp:292-338
    const int nI = n_markers, nF = nnode;
    pmap_.resize(nI);
    pmap_[0] = {0, -1.0};
    for (int i = 1; i < nI; i++) {
        const double r = i * ((nF - 1.0) / (nI - 1.0));
        if (std::fmod(r, 1.0) == 0.0) pmap_[i] = {(int)std::floor(r) - 1, 1.0};
        else pmap_[i] = {(int)std::floor(r), std::fmod(r, 1.0) * 2.0 - 1.0};
    }
    for (int el = 0; el < nel; el++) {
        for (int nd = 0; nd < nI; nd++) {
            double n1 = -1 + 2.0 * (((double)nd - 0.5) * ((nF - 1.0) / (nI - 1.0)) - (double)el) / 1.0;
            double n2 = -1 + 2.0 * (((double)nd + 0.5) * ((nF - 1.0) / (nI - 1.0)) - (double)el) / 1.0;
            if ((n1 > -1.0 && n1 < 1.0) || (n2 > -1.0 && n2 < 1.0)) {
                if (n1 < -1.0) n1 = -1.0;
                if (n2 > 1.0) n2 = 1.0;
                el_[el].map.push_back({nd, n1, n2});
            }
        }
    }
    // an element couples DOFs at most 5 apart
    M_.resize(ndof_, 5, 5);
    K_.resize(ndof_, 5, 5);
    for (auto* v : {&U, &Ud, &Udd, &Un, &Unm1, &Unm2, &Udn, &Uddn, &Ukm1, &Rk, &Rkm1, &R, &F, &delU})
        v->assign(ndof_, 0.0);
    vec_.assign(ndof_, 0.0);
}

void FemBody::node_state(int n, double out[6]) const {
    const Node& d = node_[n];
    out[0] = d.pos0[0];
    out[1] = d.pos0[1];
    out[2] = d.pos[0];
    out[3] = d.pos[1];
    out[4] = d.angle0;
    out[5] = d.angle;
}

void FemBody::set_element_state(int e, double L, double angle) {
    Elem& el = el_[e];
    el.L = L;
    el.angle = angle;
    transform(el);
}

void FemBody::set_node_state(int n, const double in[6]) {
    Node& d = node_[n];
    d.pos0[0] = in[0];
    d.pos0[1] = in[1];
    d.pos[0] = in[2];
    d.pos[1] = in[3];
    d.angle0 = in[4];
    d.angle = in[5];
}

void FemBody::layout(Layout& out) const {
    out = Layout{};
    for (const Node& n : node_) {
        for (double v : {n.pos0[0], n.pos0[1], n.pos[0], n.pos[1], n.angle0, n.angle}) out.node.push_back(v);
    }
    out.map_start.push_back(0);
    for (const Elem& e : el_) {
        for (double v : {e.L0, e.L, e.angle, e.A, e.I, e.E, e.rho}) out.el.push_back(v);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                out.Mloc.push_back(e.M[i][j]);
                out.KLloc.push_back(e.KL[i][j]);
            }
        for (const Map& m : e.map) {
            out.map_node.push_back(m.node);
            out.map_z.push_back(m.z1);
            out.map_z.push_back(m.z2);
        }
        out.map_start.push_back((int)out.map_node.size());
    }
    for (const PosMap& pm : pmap_) {
        out.pm_el.push_back(pm.el);
        out.pm_zeta.push_back(pm.zeta);
    }
}

void FemBody::import_state(const double* vecs, const double* node, const double* el, double subRes, double subNum,
                           double subDen, int itNR, double resNR) {
    std::vector<double>* v[] = {&U, &Ud, &Udd, &Un, &Unm1, &Unm2, &Udn, &Uddn, &Ukm1, &Rk, &Rkm1};
    for (int k = 0; k < 11; k++) std::memcpy(v[k]->data(), vecs + (size_t)k * ndof_, sizeof(double) * ndof_);
    for (size_t n = 0; n < node_.size(); n++) {
        Node& d = node_[n];
        const double* q = node + 6 * n;
        d.pos0[0] = q[0];
        d.pos0[1] = q[1];
        d.pos[0] = q[2];
        d.pos[1] = q[3];
        d.angle0 = q[4];
        d.angle = q[5];
    }
    for (size_t e = 0; e < el_.size(); e++) {
        el_[e].L = el[7 * e + 1];
        el_[e].angle = el[7 * e + 2];
        transform(el_[e]);
    }
    subRes_ = subRes;
    subNum_ = subNum;
    subDen_ = subDen;
    itNR_ = itNR;
    resNR_ = resNR;
}

void FemBody::transform(Elem& e) {
    const double ca = std::cos(e.angle), sa = std::sin(e.angle);
    e.T[0][0] = e.T[1][1] = e.T[3][3] = e.T[4][4] = ca;
    e.T[0][1] = e.T[3][4] = sa;
    e.T[1][0] = e.T[4][3] = -sa;
    e.T[2][2] = e.T[5][5] = 1.0;
}

void FemBody::local_matrices(Elem& e) {
    std::memset(e.M, 0, sizeof e.M);
    std::memset(e.KL, 0, sizeof e.KL);
    std::memset(e.KNL, 0, sizeof e.KNL);
    std::memset(e.R, 0, sizeof e.R);
    std::memset(e.F, 0, sizeof e.F);
    const double L0 = e.L0, E = e.E, A = e.A, I = e.I;
    const double C1 = e.rho * A * L0 / 420.0;
    const double L2 = L0 * L0, L3 = L0 * L0 * L0;
    e.M[0][0] = C1 * 140.0;
    e.M[0][3] = C1 * 70.0;
    e.M[1][1] = C1 * 156.0;
    e.M[1][2] = C1 * 22.0 * L0;
    e.M[1][4] = C1 * 54;
    e.M[1][5] = C1 * (-13.0 * L0);
    e.M[2][2] = C1 * 4.0 * L2;
    e.M[2][4] = C1 * 13.0 * L0;
    e.M[2][5] = C1 * (-3.0 * L2);
    e.M[3][3] = C1 * 140.0;
    e.M[4][4] = C1 * 156.0;
    e.M[4][5] = C1 * (-22.0 * L0);
    e.M[5][5] = C1 * 4.0 * L2;
    e.KL[0][0] = E * A / L0;
    e.KL[0][3] = -E * A / L0;
    e.KL[1][1] = 12.0 * E * I / L3;
    e.KL[1][2] = 6.0 * E * I / L2;
    e.KL[1][4] = -12.0 * E * I / L3;
    e.KL[1][5] = 6.0 * E * I / L2;
    e.KL[2][2] = 4.0 * E * I / L0;
    e.KL[2][4] = -6.0 * E * I / L2;
    e.KL[2][5] = 2.0 * E * I / L0;
    e.KL[3][3] = E * A / L0;
    e.KL[4][4] = 12.0 * E * I / L3;
    e.KL[4][5] = -6.0 * E * I / L2;
    e.KL[5][5] = 4.0 * E * I / L0;
    for (int i = 1; i < 6; i++)
        for (int j = 0; j < i; j++) {
            e.M[i][j] = e.M[j][i];
            e.KL[i][j] = e.KL[j][i];
        }
}

void FemBody::shape(const Elem& e, const double v[6], double zeta, double out[2]) {
    const double z = (zeta + 1.0) / 2.0;
    const double N0 = 1.0 - (zeta + 1.0) / 2.0;
    const double N1 = 1.0 - 3.0 * (z * z) + 2.0 * (z * z * z);
    const double N2 = (z - 2.0 * (z * z) + (z * z * z)) * e.L;
    const double N3 = z;
    const double N4 = 3.0 * (z * z) - 2.0 * (z * z * z);
    const double N5 = (-(z * z) + (z * z * z)) * e.L;
    out[0] = v[0] * N0 + v[3] * N3;
    out[1] = v[1] * N1 + v[2] * N2 + v[4] * N4 + v[5] * N5;
}

void FemBody::update_fem() {
    for (size_t n = 0; n < node_.size(); n++) {
        Node& nd = node_[n];
        for (int d = 0; d < 2; d++) nd.pos[d] = nd.pos0[d] + U[3 * n + d];
        nd.angle = nd.angle0 + U[3 * n + 2];
    }
    for (Elem& e : el_) {
        const double ev[2] = {node_[e.n1].pos[0] - node_[e.n0].pos[0], node_[e.n1].pos[1] - node_[e.n0].pos[1]};
        e.angle = std::atan2(ev[1], ev[0]);
        e.L = std::sqrt(dot(ev, ev, 2));
        transform(e);
    }
}

void FemBody::update_ibm(MarkerView mv) {
    // An element's displacement and velocity in its local frame do not
    // depend on the marker: computed once per element, then interpolated
    // to each of its markers (the same operations as per marker).
    const int nel = (int)el_.size();
    elu_.resize((size_t)nel * 12);
    for (int q = 0; q < nel; q++) {
        const Elem& e = el_[q];
        const int d0 = 3 * e.n0;
        double dU[6], dUd[6];
        for (int k = 0; k < 6; k++) {
            dU[k] = U[d0 + k];
            dUd[k] = Ud[d0 + k];
        }
        const Node* nn[2] = {&node_[e.n0], &node_[e.n1]};
        for (int n = 0; n < 2; n++) {
            dU[n * 3 + 0] += nn[n]->pos0[0] - nn[0]->pos[0];
            dU[n * 3 + 1] += nn[n]->pos0[1] - nn[0]->pos[1];
            dU[n * 3 + 2] += nn[n]->angle0 - e.angle;
            dU[n * 3 + 2] = wrap_angle(dU[n * 3 + 2]);
        }
        mul6(e.T, dU, &elu_[(size_t)q * 12]);
        mul6(e.T, dUd, &elu_[(size_t)q * 12 + 6]);
    }
    for (int i = 0; i < mv.n; i++) {
        const int q = pmap_[i].el;
        const Elem& e = el_[q];
        const double* lU = &elu_[(size_t)q * 12];
        double s[2], sd[2];
        shape(e, lU, pmap_[i].zeta, s);
        shape(e, lU + 6, pmap_[i].zeta, sd);
        const double* p0 = node_[e.n0].pos;
        double* p = mv.pos + 2 * (mv.first + i);
        double* v = mv.vel + 2 * (mv.first + i);
        for (int r = 0; r < 2; r++) {
            const double gs = 0.0 + e.T[0][r] * s[0] + e.T[1][r] * s[1];
            const double gsd = 0.0 + e.T[0][r] * sd[0] + e.T[1][r] * sd[1];
            p[r] = p0[r] + gs;
            v[r] = gsd;
        }
    }
}

void FemBody::load_vector(Elem& e, MarkerView mv) {
    const double wt[2] = {e.rho * e.A * s_.gravityX, e.rho * e.A * s_.gravityY};
    const double fs = s_.Dm / (s_.Dt * s_.Dt);
    const double L = e.L;
    for (const Map& mp : e.map) {
        const int k = mv.first + mp.node;
        const double a = mp.z1, b = mp.z2;
        const double sc = -mv.eps[k] * 1.0 * fs;
        const double vv[2] = {sc * mv.force[2 * k] + wt[0], sc * mv.force[2 * k + 1] + wt[1]};
        double Fl[2];
        for (int r = 0; r < 2; r++) Fl[r] = 0.0 + e.T[r][0] * vv[0] + e.T[r][1] * vv[1];
        const double a2 = a * a, b2 = b * b, a3 = a * a * a, b3 = b * b * b;
        e.R[0] = Fl[0] * 0.5 * L * (0.5 * b - 0.5 * a + 0.25 * a2 - 0.25 * b2);
        e.R[1] = Fl[1] * 0.5 * L * (0.5 * b - 0.5 * a - a2 * a2 / 16.0 + b2 * b2 / 16.0 + 3.0 * a2 / 8.0 - 3.0 * b2 / 8.0);
        e.R[2] = Fl[1] * 0.5 * L *
                 (L * (-a2 * a2 + b2 * b2) / 32.0 - L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 + L * (b - a) / 8.0);
        e.R[3] = Fl[0] * 0.5 * L * (-0.25 * a2 + 0.25 * b2 + 0.5 * b - 0.5 * a);
        e.R[4] = Fl[1] * 0.5 * L * (0.5 * b - 0.5 * a + a2 * a2 / 16.0 - b2 * b2 / 16.0 - 3.0 * a2 / 8.0 + 3.0 * b2 / 8.0);
        e.R[5] = Fl[1] * 0.5 * L *
                 (L * (-a2 * a2 + b2 * b2) / 32.0 + L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 - L * (b - a) / 8.0);
        double RG[6];
        mul_t6(e.T, e.R, RG);
        const int d0 = 3 * e.n0;
        for (int q = 0; q < 6; q++) R[d0 + q] += RG[q];
    }
}

void FemBody::force_vector(Elem& e) {
    const double u = (e.L * e.L - e.L0 * e.L0) / (e.L + e.L0);
    const double th1 = wrap_angle(node_[e.n0].angle - e.angle);
    const double th2 = wrap_angle(node_[e.n1].angle - e.angle);
    const double F0 = (e.E * e.A / e.L0) * u;
    const double M1 = (2 * e.E * e.I / e.L0) * (2.0 * th1 + th2);
    const double M2 = (2 * e.E * e.I / e.L0) * (th1 + 2.0 * th2);
    e.F[0] = -F0;
    e.F[1] = (1.0 / e.L0) * (M1 + M2);
    e.F[2] = M1;
    e.F[3] = F0;
    e.F[4] = -(1.0 / e.L0) * (M1 + M2);
    e.F[5] = M2;
    double FG[6];
    mul_t6(e.T, e.F, FG);
    const int d0 = 3 * e.n0;
    for (int q = 0; q < 6; q++) F[d0 + q] += FG[q];
}

void FemBody::mass_stiff(Elem& e) {
    double MG[6][6], KG[6][6], KN[6][6];
    const int d0 = 3 * e.n0;
    rotate6(e.T, e.M, MG);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) M_.at(d0 + i, d0 + j) += MG[i][j];
    rotate6(e.T, e.KL, KG);
    const double F0 = -e.F[0], V0 = e.F[4], L0 = e.L0;
    e.KNL[0][1] = -V0 / L0;
    e.KNL[0][4] = V0 / L0;
    e.KNL[1][0] = -V0 / L0;
    e.KNL[1][1] = F0 / L0;
    e.KNL[1][3] = V0 / L0;
    e.KNL[1][4] = -F0 / L0;
    e.KNL[3][1] = V0 / L0;
    e.KNL[3][4] = -V0 / L0;
    e.KNL[4][0] = V0 / L0;
    e.KNL[4][1] = -F0 / L0;
    e.KNL[4][3] = -V0 / L0;
    e.KNL[4][4] = F0 / L0;
    rotate6(e.T, e.KNL, KN);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) K_.at(d0 + i, d0 + j) += KG[i][j] + KN[i][j];
}

void FemBody::build_global() {
    M_.zero();
    K_.zero();
    std::fill(F.begin(), F.end(), 0.0);
    for (Elem& e : el_) {
        force_vector(e);
        mass_stiff(e);
    }
}

void FemBody::finish_newmark() {
    const double Dt = s_.Dt, al = s_.alpha, de = s_.delta;
    const double a6 = 1.0 / (al * (Dt * Dt));
    const double a7 = -1.0 / (al * Dt);
    const double a8 = -(1.0 / (2.0 * al) - 1.0);
    const double a9 = Dt * (1.0 - de);
    const double a10 = de * Dt;
    for (int k = 0; k < ndof_; k++) {
        Udd[k] = a6 * (U[k] - Un[k]) + a7 * Udn[k] + a8 * Uddn[k];
        Ud[k] = Udn[k] + a9 * Uddn[k] + a10 * Udd[k];
    }
}

void FemBody::newton_dynamic() {
    build_global();
    const double Dt = s_.Dt, al = s_.alpha;
    const double a0 = 1.0 / (al * (Dt * Dt));
    const double a2 = 1.0 / (al * Dt);
    const double a3 = 1.0 / (2.0 * al) - 1.0;
    for (int k = 0; k < ndof_; k++) vec_[k] = a0 * (Un[k] - U[k]) + a2 * Ud[k] + a3 * Udd[k];
    const int kb = M_.kl;
    for (int i = 0; i < ndof_; i++) {
        // MatMultiply row i over the band, increasing column order from 0.0
        double s = 0.0;
        const int lo = i - kb < 0 ? 0 : i - kb, hi = i + kb > ndof_ - 1 ? ndof_ - 1 : i + kb;
        for (int j = lo; j <= hi; j++) s += M_.at(i, j) * vec_[j];
        F[i] = R[i] - F[i] + s;
    }
    for (int i = 0; i < ndof_; i++) {
        const int lo = i - kb < 0 ? 0 : i - kb, hi = i + kb > ndof_ - 1 ? ndof_ - 1 : i + kb;
        for (int j = lo; j <= hi; j++) K_.at(i, j) = K_.at(i, j) + a0 * M_.at(i, j);
    }
    lu_solve_band(K_, nbc_, F.data(), delU.data(), W_, y_);
    for (int k = 0; k < ndof_; k++) U[k] = U[k] + delU[k];
    update_fem();
}

void FemBody::dynamic(MarkerView mv) {
    U = Un;
    Ud = Udn;
    Udd = Uddn;
    update_fem();
    std::fill(R.begin(), R.end(), 0.0);
    for (Elem& e : el_) load_vector(e, mv);
    itNR_ = 0;
    do {
        newton_dynamic();
        resNR_ = std::sqrt(dot(delU.data(), delU.data(), ndof_)) / (s_.ref_L * std::sqrt((double)ndof_));
        itNR_++;
    } while (resNR_ > 1e-10 && itNR_ < 20);
    finish_newmark();
    update_ibm(mv);
    // FEMBodyClass::subResidual, FEMBody.cpp:244-256
    Rkm1.swap(Rk);
    for (int k = 0; k < ndof_; k++) Rk[k] = U[k] - Ukm1[k];
    subRes_ = dot(Rk.data(), Rk.data(), ndof_);
    double num = 0.0, den = 0.0;
    for (int k = 0; k < ndof_; k++) num += Rkm1[k] * (Rk[k] - Rkm1[k]);
    for (int k = 0; k < ndof_; k++) den += (Rk[k] - Rkm1[k]) * (Rk[k] - Rkm1[k]);
    subNum_ = num;
    subDen_ = den;
}

// FEMBodyClass::staticFEM + newtonRaphsonStatic, FEMBody.cpp:448-484:
// Newton on K dU = R - F (stiffness only), then the markers follow.
void FemBody::static_fem(MarkerView mv) {
    itNR_ = 0;
    do {
        build_global();
        for (int k = 0; k < ndof_; k++) vec_[k] = R[k] - F[k];
        lu_solve_band(K_, nbc_, vec_.data(), delU.data(), W_, y_);
        for (int k = 0; k < ndof_; k++) U[k] = U[k] + delU[k];
        update_fem();
        resNR_ = std::sqrt(dot(delU.data(), delU.data(), ndof_)) / (s_.ref_L * std::sqrt((double)ndof_));
        itNR_++;
    } while (resNR_ > 1e-10 && itNR_ < 20);
    update_ibm(mv);
}

namespace {
// 2x2 matrix times vector, accumulating from 0.0 (Utils.h:375-389).
void mat2_vec(const double T[2][2], const double v[2], double out[2]) {
    for (int i = 0; i < 2; i++) {
        double r = 0.0;
        for (int j = 0; j < 2; j++) r += T[i][j] * v[j];
        out[i] = r;
    }
}
}  // namespace

// FEMBodyClass::setInitialDeflection, FEMBody.cpp:352-445: a tip load in
// the body frame, corrected by a bounded proportional feedback until the
// static tip deflection is -fraction * L0.
void FemBody::set_initial_deflection(double fraction, MarkerView mv) {
    const int nd = ndof_;
    const double deflect = -fraction;
    const double E = el_[0].E, I = el_[0].I;
    const double linearForce = 3.0 * E * I * deflect / (L0_ * L0_);
    const int nSteps = 100;
    const double dFmax = linearForce / (double)nSteps;
    const double T[2][2] = {{std::cos(angle0_), -std::sin(angle0_)}, {std::sin(angle0_), std::cos(angle0_)}};
    const double Tt[2][2] = {{T[0][0], T[1][0]}, {T[0][1], T[1][1]}};
    double local[2] = {0.0, dFmax}, global[2];
    mat2_vec(T, local, global);
    R[nd - 3] = global[0];
    R[nd - 2] = global[1];
    const double gain = 1e-3, TOL = 1e-10;
    const int MAXIT = 10000;
    int it = 0;
    double error;
    do {
        static_fem(mv);
        const double dv[2] = {U[nd - 3], U[nd - 2]};
        double dl[2];
        mat2_vec(Tt, dv, dl);
        error = deflect - dl[1] / L0_;
        double fc = error * gain;
        if (std::fabs(fc) > std::fabs(dFmax)) fc = (double)((0.0 < fc) - (fc < 0.0)) * std::fabs(dFmax);
        local[1] += fc;
        mat2_vec(T, local, global);
        R[nd - 3] = global[0];
        R[nd - 2] = global[1];
        it++;
        if (it >= MAXIT) throw std::runtime_error("max iterations hit! Try changing the gain...exiting");
    } while (std::fabs(error) > TOL);
    itNR_ = 0;
    resNR_ = 0.0;
    std::fill(R.begin(), R.end(), 0.0);
    std::fill(delU.begin(), delU.end(), 0.0);
}

void FemBody::predictor(int t, MarkerView mv) {
    if (t > 2)
        for (int k = 0; k < ndof_; k++) U[k] = 2.5 * Un[k] - 2.0 * Unm1[k] + 0.5 * Unm2[k];
    else if (t == 2)
        for (int k = 0; k < ndof_; k++) U[k] = 2.0 * Un[k] - Unm1[k];
    else if (t == 1)
        U = Un;
    update_fem();
    finish_newmark();
    update_ibm(mv);
    Ukm1.swap(U);
}

void FemBody::relax(double omega, MarkerView mv) {
    for (int k = 0; k < ndof_; k++) U[k] = Ukm1[k] + omega * (U[k] - Ukm1[k]);
    update_fem();
    finish_newmark();
    update_ibm(mv);
    Ukm1.swap(U);
}

void FemBody::reset_values() {
    Unm2.swap(Unm1);
    Unm1.swap(Un);
    Un.swap(U);
    Udn.swap(Ud);
    Uddn.swap(Udd);
}

}  // namespace life
