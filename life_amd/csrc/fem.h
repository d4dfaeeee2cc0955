// FP64 corotational Euler-Bernoulli beam (host side of the coupling).
//
// Restates FEMBodyClass / FEMElementClass / FEMNodeClass (FEMBody.cpp,
// FEMElement.cpp, FEMNode.cpp): 3 DOF per node, consistent mass, linear +
// geometric stiffness in a corotating frame, Newmark-beta time integration
// and Newton-Raphson iterations.  The global matrices are assembled in
// banded storage (an element couples DOFs at most 5 apart) and solved with
// the banded form of the canonical elimination, which performs the same
// floating-point operations as the reference's dense solve on the same
// matrix; assembly and products keep the reference's accumulation order.
#pragma once

#include <string>
#include <vector>

#include "lu.h"

namespace life {

// Marker arrays of the owning body (global marker order, physical units).
struct MarkerView {
    double* pos;          // 2 per marker
    double* vel;          // 2 per marker
    const double* force;  // 2 per marker (lattice units)
    const double* eps;    // 1 per marker
    int first, n;
};

struct FemScalars {
    double Dx, Dt, Dm;
    double alpha, delta;  // Newmark (params.h:76-77)
    double gravityX, gravityY;
    double ref_L;
};

class FemBody {
  public:
    // FEMBodyClass constructor (FEMBody.cpp:487-560); angle in radians.
    FemBody(const FemScalars& s, const double pos[2], const double geom[2], double angle, const std::string& nEl,
            const std::string& bc, double rho, double E, int n_markers);

    void reset_values();                          // FEMBody.cpp:341-349
    void predictor(int t, MarkerView mv);         // FEMBody.cpp:259-289
    void relax(double omega, MarkerView mv);      // Objects.cpp:198-210
    void dynamic(MarkerView mv);                  // FEMBody.cpp:26-68
    // INITIAL_DEFLECT: static tip deflection of -fraction * L0 (FEMBody.cpp:352-484)
    void set_initial_deflection(double fraction, MarkerView mv);

    int ndof() const { return ndof_; }
    int bc_dofs() const { return nbc_; }
    double sub_res() const { return subRes_; }
    double sub_num() const { return subNum_; }
    double sub_den() const { return subDen_; }
    int it_nr() const { return itNR_; }
    double res_nr() const { return resNR_; }
    double L0() const { return L0_; }
    double angle0() const { return angle0_; }
    double t_crit() const { return tCrit_; }

    std::vector<double> U, Ud, Udd, Un, Unm1, Unm2, Udn, Uddn, Ukm1, Rk, Rkm1, R, F, delU;

    // Element and node state of the restart file (FEM.restart, Objects.cpp:1206-1386):
    // per element L and angle; per node pos0[2], pos[2], angle0, angle.
    int num_elements() const { return (int)el_.size(); }
    int num_nodes() const { return (int)node_.size(); }
    void element_state(int e, double& L, double& angle) const {
        L = el_[e].L;
        angle = el_[e].angle;
    }
    void node_state(int n, double out[6]) const;
    // Restore them as ObjectsClass::readRestart does (Objects.cpp:1164-1201):
    // the element transforms follow the angles; the local matrices depend on
    // L0 only and are unchanged.
    void set_element_state(int e, double L, double angle);
    void set_node_state(int n, const double in[6]);

    // Flat copy of the body for the device FEM (fem_dev.hip): per node
    // pos0[2], pos[2], angle0, angle; per element L0, L, angle, A, I, E,
    // rho; the local mass and linear stiffness matrices (constant); the
    // element-to-marker load map (CSR: per element its entries' marker,
    // z1, z2) and the marker-to-element map (element, zeta).
    struct Layout {
        std::vector<double> node, el, Mloc, KLloc, map_z, pm_zeta;
        std::vector<int> map_start, map_node, pm_el;
    };
    void layout(Layout& out) const;
    // Back from the device: the vectors (U, Ud, Udd, Un, Unm1, Unm2, Udn,
    // Uddn, Ukm1, Rk, Rkm1 in that order, ndof each), node and element
    // state in the Layout's order, and the last solve's scalars.
    void import_state(const double* vecs, const double* node, const double* el, double subRes, double subNum,
                      double subDen, int itNR, double resNR);

  private:
    struct Node {
        double pos0[2], pos[2], angle0, angle;
    };
    struct Map {
        int node;
        double z1, z2;
    };
    struct Elem {
        double L0, L, angle, A, I, E, rho;
        double T[6][6], M[6][6], KL[6][6], KNL[6][6], R[6], F[6];
        int n0, n1;
        std::vector<Map> map;
    };
    struct PosMap {
        int el;
        double zeta;
    };

    void static_fem(MarkerView mv);    // FEMBody.cpp:448-484
    void update_fem();                 // FEMBody.cpp:198-223
    void update_ibm(MarkerView mv);    // FEMBody.cpp:147-195
    void finish_newmark();             // FEMBody.cpp:129-144
    void build_global();               // FEMBody.cpp:90-109
    void newton_dynamic();             // FEMBody.cpp:71-87, 112-126
    void load_vector(Elem& e, MarkerView mv);  // FEMElement.cpp:27-69
    void force_vector(Elem& e);        // FEMElement.cpp:72-97
    void mass_stiff(Elem& e);          // FEMElement.cpp:100-138
    static void transform(Elem& e);    // FEMElement.cpp:255-262
    static void local_matrices(Elem& e);  // FEMElement.cpp:203-252
    static void shape(const Elem& e, const double v[6], double zeta, double out[2]);  // FEMElement.cpp:141-160

    FemScalars s_;
    double L0_, angle0_, tCrit_;
    int ndof_, nbc_, itNR_ = 0;
    double resNR_ = 0.0, subRes_ = 0.0, subNum_ = 0.0, subDen_ = 0.0;
    std::vector<Node> node_;
    std::vector<Elem> el_;
    std::vector<PosMap> pmap_;
    std::vector<double> elu_;  // per element: local displacement, velocity (update_ibm)
    BandMatrix M_, K_, W_;
    std::vector<double> vec_, y_;
};

}  // namespace life
