// Device form of the FP64 corotational beam FEM (fem_dev.hip): the
// femKernel and the FEM half of recomputeObjectVals (FEMBody.cpp:26-289,
// FEMElement.cpp:27-262, Objects.cpp:152-232) for flexible bodies, one
// workgroup per body, with the operations of FemBody (fem.cpp) in the same
// order.  The transcendental functions (atan2, sin, cos) are the device
// library's, so results agree with the host FEM to rounding, not bit for
// bit; the plain method (FP64 populations) keeps the host FEM.
#pragma once

#include <hip/hip_runtime_api.h>

#include <vector>

namespace life {

class FemBody;

// Vectors of one body, ndof each, in this order in FemDevBody::vec: the
// state a restart or the host copy needs.
enum FemVec { kFU, kFUd, kFUdd, kFUn, kFUnm1, kFUnm2, kFUdn, kFUddn, kFUkm1, kFRk, kFRkm1, kFemState };
constexpr int kFemElStride = 16;  // per element: L0 L angle A I E rho | ca sa | F[6] | pad
constexpr int kFemKL = 5;         // an element couples DOFs at most 5 apart
constexpr int kFemBandW = 2 * kFemKL + 1;  // band row: columns r-5 .. r+5
constexpr int kFemMatNZ = 13;  // upper-triangle non-zeros of an element's mass or linear stiffness matrix
constexpr int kFemThreads = 256;

struct FemDevScalars {
    double Dx, Dt, Dm, alpha, delta, gx, gy, ref_L;
};

// One flexible body: its FEM state, its constants and its markers.
struct FemDevBody {
    int ndof, nbc, nel, nnode, nm;
    FemDevScalars s;
    // vec, node, el and mats lie in this order in one device block
    double* node;          // nnode x 6: pos0 x, y, pos x, y, angle0, angle
    double* el;            // nel x kFemElStride
    const double* mats;    // nel x 2 x kFemMatNZ: M's then KL's upper-triangle non-zeros (constant)
    const int* map_start;  // nel + 1
    const int* map_node;   // marker (body-local) of each load-map entry
    const double* map_z;   // 2 per entry
    const int* pm_el;      // nm: element of each marker
    const double* pm_zeta; // nm
    double* vec;           // kFemState x ndof
    double* out;           // subRes, subNum, subDen, itNR, resNR
    // the body's markers (physical positions and velocities, lattice
    // forces, epsilon), already offset to its first marker
    double* mpos;
    double* mvel;
    const double* mforce;
    const double* meps;
};

// Device FEM of one case's flexible bodies (in the case's body order).
struct FemDevCase {
    int nb = 0;
    std::vector<FemDevBody> host;  // descriptors (device pointers)
    FemDevBody* d_bodies = nullptr;
    std::vector<void*> blocks;     // device allocations
    int* done = nullptr;           // device: bodies finished in the running launch
    // page-locked: res, num, den over the bodies, in order; [3] the sequence
    // number of the launch that wrote them (written last)
    double* sums_h = nullptr;
    double seq = 0.0;              // sequence number of the latest launch
    size_t lds = 0;                // dynamic LDS of the solve kernels
    bool ok() const { return nb > 0 && d_bodies != nullptr; }
};

// LDS a launch needs for a body of ndof unknowns and nm markers (the
// solve's global matrices, vectors and marker loads, and the body's state
// and element data mirrored from device memory); bodies beyond the limit
// (gfx950: 160 KiB of LDS per workgroup, less the kernels' static 6 KiB)
// stay on the host FEM.
size_t fem_dev_lds_bytes(int ndof, int nm);
constexpr size_t kFemMaxLds = 150 * 1024;

// Device state from the host bodies (their present state).  mk: the case's
// marker arrays; first[b], nm[b]: each body's marker range.
void fem_dev_create(FemDevCase& fc, const std::vector<const FemBody*>& bodies, const std::vector<int>& first,
                    const std::vector<int>& nm, double* pos, double* vel, const double* force, const double* eps,
                    const FemDevScalars& s, hipStream_t st);
void fem_dev_destroy(FemDevCase& fc);
// Host bodies take the device state (synchronises the stream).
void fem_dev_download(FemDevCase& fc, const std::vector<FemBody*>& bodies, hipStream_t st);
// Device state takes the host bodies' state.
void fem_dev_upload(FemDevCase& fc, const std::vector<const FemBody*>& bodies, hipStream_t st);

// Launchers over nb bodies: list[k] (k < nb) indexes the descriptors
// (device memory; bodies of several cases may be mixed), or with list null
// the bodies 0 .. nb-1.
// reset_values + predictor (subiteration 0 of a step): step t, or per
// listed body t_d[k].
void launch_fem_predict(const FemDevBody* b, const int* list, int nb, int t, const double* t_d, size_t lds,
                        hipStream_t st);
// relax (later subiterations): factor omega, or per listed body omega_d[k].
void launch_fem_relax(const FemDevBody* b, const int* list, int nb, double omega, const double* omega_d, size_t lds,
                      hipStream_t st);
// Where the dynamic launch over bodies 0 .. nb-1 of one case leaves the
// case's subiteration sums (done null: nowhere).  done: a device int, zero
// between launches; sums_h: page-locked, 4 doubles: the sums, then seq,
// written after them and the flags with a system-scope fence between, so a
// host that sees seq there sees the rest; flags (device) -> flags_h
// (page-locked): nflags ints copied alongside.
struct FemSumsOut {
    int* done = nullptr;
    double* sums_h = nullptr;
    double seq = 0.0;
    const int* flags = nullptr;
    int* flags_h = nullptr;
    int nflags = 0;
};

// Wait until the launch with sequence number seq has written its sums
// (polling the page-locked word, for at most spin_us microseconds), else
// until ev (recorded after that launch) completes.  Returns with the sums
// and flags readable.
void fem_wait_sums(const FemSumsOut& so, hipEvent_t ev, double spin_us);

// femKernel: the dynamic Newmark/Newton solve of every listed body, its new
// marker positions and velocities, and its subiteration residual terms.
// redo: again after epsilon changed, with the previous residual kept.  With
// so.done set (bodies 0 .. nb-1 only), the last workgroup also writes the
// sums over the bodies and the flags to the host (synchronise the stream
// before reading them).
void launch_fem_dynamic(const FemDevBody* b, const int* list, int nb, size_t lds, bool redo, hipStream_t st,
                        const FemSumsOut& so = FemSumsOut());
// out[0..2] of the listed bodies packed into outs (3 per body; outs may be
// page-locked host memory, readable after the stream is synchronised).
void launch_fem_outs(const FemDevBody* b, const int* list, int nb, double* outs, hipStream_t st);

}  // namespace life
