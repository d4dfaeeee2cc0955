/*
 * liblife_hip.so -- MI355X (gfx950) implementation of the LIFE fluid and
 * immersed-boundary coupling hot path, behind a plain C ABI.
 *
 * Two layers are exported:
 *
 *  1. life_grid_* / life_lbm_* / life_ibm_*: device lattice and marker
 *     state with one entry point per reference method on the per-step path.
 *     A maintainer of the reference keeps GridClass/ObjectsClass/FEM on the
 *     host and replaces the bodies of those methods with these calls.
 *
 *  2. life_case_*: the library's own C++ host driver, which owns a lattice,
 *     its bodies and their FP64 FEM, and runs the reference's main loop body
 *     (GridClass::solver, Grid.cpp:25-33) on top of layer 1.
 *
 * Every function returns 0 on success and a non-zero code on failure;
 * life_last_error() then returns the message (thread-local).  No C++
 * exception crosses this boundary.  A handle is bound to one HIP device and
 * one HIP stream and is not thread-safe; distinct handles are independent.
 *
 * Units and layouts follow the reference: lattice id = i*Ny + j (j fastest,
 * Grid.cpp:70), populations in D2Q9 order v0..v8 (Grid.cpp:1249), host-side
 * population arrays in the reference's AoS order f[id*9 + v], vectors as
 * [id*2 + d], marker positions/velocities in physical units (IBMNode.h:52-53).
 */
#ifndef LIFE_HIP_H
#define LIFE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* eLatType, inc/defs.h:52 */
enum { LIFE_FLUID = 0, LIFE_WALL = 1, LIFE_VELOCITY = 2, LIFE_FREESLIP = 3, LIFE_PRESSURE = 4, LIFE_CONVECTIVE = 5 };
/* eProfileType, inc/defs.h:53; LIFE_UNIFORM = PROFILE left undefined */
enum { LIFE_UNIFORM = -1, LIFE_PARABOLIC = 0, LIFE_SHEAR = 1, LIFE_BOUNDARY_LAYER = 2 };
/* Population storage. */
enum {
    LIFE_FP64 = 0,          /* plain method: FP64 populations, reference operation order */
    LIFE_FP32_SHIFTED = 1   /* f_i - w_i*rho0 stored in FP32, FP32 collision, FP64 reductions */
};

/* Runtime form of inc/params.h:26-110 (same names and meanings). */
typedef struct life_params {
    int Nx, Ny;
    double height_p, rho_p, nu_p;
    double ux0_p, uy0_p;
    double gravityX, gravityY, dpdx, dpdy;
    int wall_left, wall_right, wall_bottom, wall_top;
    int profile;
    double uxInlet_p, uyInlet_p;
    double alpha, delta;
    double relaxMax, subTol;
    double omega;            /* > 0: used directly; else derived from tStep (params.h:89-90) */
    double tStep;
    int central_moments;     /* CENTRAL_MOMENTS (params.h:32) */
    double inlet_ramp;       /* INLET_RAMP seconds; <= 0: undefined (params.h:33) */
    int uni_epsilon;         /* UNI_EPSILON (params.h:35) */
    double initial_deflect;  /* INITIAL_DEFLECT; 0: undefined (params.h:37) */
    double ref_nu, ref_rho, ref_P, ref_L, ref_U;
    double womersley;        /* WOMERSLEY number of the oscillating pressure gradient dpdx/dpdy
                                (params.h:34, Grid.cpp:52-61); 0: undefined.  Needs
                                gravityX = gravityY = 0 (the force is then uniform). */
} life_params;

typedef struct life_grid life_grid;
typedef struct life_case life_case;

const char *life_last_error(void);
/* Library version string. */
const char *life_version(void);
/* Number of visible HIP devices (0 without a GPU). */
int life_device_count(int *count);
/* Free the library's pooled HIP resources: the device and page-locked
 * blocks kept for reuse by later grids and -- when no grid or server is
 * live -- the pooled and CU-masked streams, which otherwise live until the
 * process ends.  The last destroy of a grid or server frees the blocks by
 * itself; a host calls this once before it exits (the Python package does
 * so at interpreter exit).  No reference counterpart (the reference
 * allocates no device resources). */
int life_release_cached(void);

/* ---------------------------------------------------------------- layer 1 */

/* Replaces GridClass::GridClass + initialiseGrid (Grid.cpp:1232-1289,
 * 916-1062): allocate the device lattice for a case and set f = f_eq. */
int life_grid_create(const life_params *p, int precision, int device, life_grid **out);
int life_grid_destroy(life_grid *g);

/* Derived scalars omega, tau, nu, Dx, Dt, Dm, Drho, c_s, c_s^2, c_s^4
 * (Grid.cpp:1246-1260). */
int life_grid_scalars(const life_grid *g, double out[10]);

/* Replaces GridClass::lbmKernel (Grid.cpp:36-100) for time step t:
 * convective speed, fused stream-collide, boundary conditions.  Enqueued on
 * the grid's stream; returns without waiting. */
int life_lbm_kernel(life_grid *g, int t);

/* Materialise the post-step state in the reference layout (the arrays that
 * GridClass::writeVTK/writeRestart read, Grid.cpp:790-913, 1163-1229).
 * Any pointer may be NULL.  Blocks until the data is on the host. */
int life_grid_download(life_grid *g, double *f_aos, double *rho, double *u, double *force_ibm);
/* Load a full state (GridClass::readRestart, Grid.cpp:1072-1160).  The next
 * step uses rho/u as given, as the reference does after a restart. */
int life_grid_upload(life_grid *g, const double *f_aos, const double *rho, const double *u,
                     const double *force_ibm);
/* Wait for all work queued on the grid's stream. */
int life_grid_sync(life_grid *g);
/* The NaN test of GridClass::writeInfo (Grid.cpp:562-581) on the current
 * state: *i, *j = the first site in the reference's i-major order whose
 * velocity magnitude is NaN, or -1, -1 when there is none.  Blocks. */
int life_grid_blowup_check(life_grid *g, int *i, int *j);

/* Markers: replaces the IBMNodeClass arrays (IBMNode.h:45-70).  n markers in
 * the reference's global order; body[k] is the index of marker k's body
 * (bodies hold contiguous marker ranges); body_flex[b] is 0 for eFlexible,
 * 1 for eRigid (defs.h:48). */
int life_ibm_set_markers(life_grid *g, int n, const double *pos, const double *vel, const int *body,
                         int nbodies, const int *body_flex);
/* Update marker positions/velocities (after FEMBodyClass::updateIBMValues,
 * FEMBody.cpp:147-195).  pos/vel may be NULL. */
int life_ibm_update_markers(life_grid *g, const double *pos, const double *vel);
/* IBMNodeClass::findSupport + computeDs (IBMNode.cpp:139-204) for all
 * markers (flexible_only = 0, Objects.cpp:949-955) or only the markers of
 * flexible bodies (flexible_only = 1, Objects.cpp:217-227). */
int life_ibm_find_support(life_grid *g, int flexible_only);
/* Epsilon matrix of ObjectsClass::computeEpsilon (Objects.cpp:258-302) for
 * the system made of markers [first, first+n): A row-major n x n.  */
int life_ibm_epsilon_matrix(life_grid *g, int first, int n, double *A);
/* life_ibm_epsilon_matrix without any device allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory (stream
 * capture, tight serving loops).  The matrix is assembled in a fixed 8 MiB
 * device array per GPU (n <= 1024 markers; larger n returns an error) that
 * every call on that GPU shares: overlapping calls on one GPU race.  */
int life_ibm_epsilon_matrix_noalloc(life_grid *g, int first, int n, double *A);
/* Set epsilon for all markers (Objects.cpp:309-310). */
int life_ibm_set_epsilon(life_grid *g, const double *eps);
/* ObjectsClass::ibmKernelInterp (Objects.cpp:102-117): interpolate rho and
 * rho*u to every marker and compute its direct-forcing force; force (n*2,
 * may be NULL) receives the marker forces.  Blocks when force != NULL. */
int life_ibm_interpolate(life_grid *g, double *force);
/* ObjectsClass::ibmKernelSpread (Objects.cpp:120-149): spread marker forces
 * onto the lattice in marker order and refresh the support-site moments. */
int life_ibm_spread(life_grid *g);
/* Marker data for output and tests.  Any pointer may be NULL. */
int life_ibm_download(life_grid *g, double *pos, double *vel, double *force, double *eps, double *ds,
                      int *supp_count, int *supp_ij, double *supp_delta);

/* Replaces Utils::solveLAPACK (Utils.cpp:288-311): solve A x = b for the
 * trailing (n-bc) block of the row-major n x n matrix A, x[0..bc) = 0, with
 * the library's canonical partial-pivot elimination (host, FP64; no GPU). */
int life_solve_lapack(int n, int bc, const double *A, const double *b, double *x);

/* ---------------------------------------------------------------- layer 2 */

/* Build a case from params and geometry.config text (Objects.cpp:324-488),
 * including body construction, overlap removal and initialiseObjects
 * (Objects.cpp:867-961). */
int life_case_create(const life_params *p, const char *geometry, int precision, int device, life_case **out);
int life_case_destroy(life_case *c);
/* nsteps iterations of the main loop body (main.cpp:70-73):
 * t++, lbmKernel, objectKernel. */
int life_case_step(life_case *c, int nsteps);
/* As life_case_step, also returning the elapsed device time of the steps
 * (ms, HIP events on the case stream) and of the fused stream-collide kernel
 * launches alone (ms, summed). */
int life_case_step_timed(life_case *c, int nsteps, double *ms_total, double *ms_kernel);
int life_case_time(const life_case *c, int *t);
/* Cadence of the blow-up test (the NaN velocity scan of GridClass::writeInfo,
 * Grid.cpp:562-581, which the reference runs every tinfo steps): every
 * `every` steps (default 100; 0 turns it off) life_case_step scans the
 * populations on the device.  A NaN makes life_case_step return non-zero
 * with "Simulation blew up (t = ..) at i = .., j = ..", t being the step
 * scanned and (i, j) the first such site in i-major order; the steps after
 * it that the call had already queued have run. */
int life_case_set_check_interval(life_case *c, int every);
life_grid *life_case_grid(life_case *c);
int life_case_num_markers(const life_case *c, int *n);
int life_case_num_bodies(const life_case *c, int *n);
/* body info: ID, flex, nmarkers, first marker, FEM dofs */
int life_case_body_info(const life_case *c, int b, int info[5]);
int life_case_get_fem(const life_case *c, int b, double *U, double *Udot, double *Udotdot);
/* subIt, subRes, relax of the last objectKernel (Objects.cpp:26-60) */
int life_case_coupling_info(const life_case *c, double out[3]);
/* ObjectsClass::writeTotalForces quantities (Objects.cpp:517-534): Fx, Fy, Cx, Cy */
int life_case_total_force(life_case *c, double out[4]);

/* Result files in the reference's formats, under a results directory laid
 * out as the reference's Results/ folder.  Binary data is in the host's byte
 * order (little-endian here), as the reference writes it on such a host. */
/* GridClass::writeRestart + ObjectsClass::writeRestart (Grid.cpp:1163-1229,
 * Objects.cpp:1206-1386): <results>/Restart/Fluid.restart, IBM.restart (bodies)
 * and FEM.restart (flexible bodies), each written to a .temp file first. */
int life_case_write_restart(life_case *c, const char *results);
/* GridClass::readRestart + ObjectsClass::readRestart + restartCleanup
 * (Grid.cpp:1072-1160, Objects.cpp:979-1204, 1387-1586): restores the lattice
 * (the next step uses the stored rho/u), the time step, the markers and the
 * FEM state, recomputes supports, spacing and epsilon, and trims the force and
 * tip logs to the steps before the restart. */
int life_case_read_restart(life_case *c, const char *results);
/* GridClass::writeVTK + ObjectsClass::writeVTK (Grid.cpp:790-913,
 * Objects.cpp:711-864): <results>/VTK/Fluid.<t>.vti, IBM.<t>.vtp and, with
 * fem = 1 (the reference's VTK_FEM), FEM.<t>.vtp. */
int life_case_write_vtk(life_case *c, const char *results, int fem);
/* ObjectsClass::writeTotalForces (Objects.cpp:491-539): appends a line to
 * <results>/TotalForces.out (header first when the file is new). */
int life_case_write_forces(life_case *c, const char *results);
/* ObjectsClass::writeTips (Objects.cpp:542-594): appends to
 * <results>/TipPositions.out and TipVelocities.out. */
int life_case_write_tips(life_case *c, const char *results);

/* ---------------------------------------------------------------- serving
 *
 * Continuous batching of independent cases on one device.  The reference
 * runs one case per process (main.cpp:33-95); it has no counterpart for
 * this layer, which drives many layer-2 cases at once.  Requests queue in
 * arrival order and are admitted while a slot is free and their device
 * footprint fits the HBM budget; every engine step advances each admitted
 * case by one timestep (lattice phase of all cases as one hipGraph launch
 * over a packed (case, block) index, batched coupling kernels, host FP64
 * FEM on host threads), and a finished case frees its slot for the next
 * request.  Each case's results are those of life_case_step on the same
 * inputs, whatever else runs beside it. */
typedef struct life_server life_server;

typedef struct life_server_config {
    int device;
    int precision;      /* LIFE_FP32_SHIFTED or LIFE_FP64, for every case */
    int max_cases;      /* admission slots */
    double hbm_budget;  /* bytes of device memory the admitted cases may hold */
    int use_graph;      /* 1: lattice phase as one hipGraph launch per step */
    int check_every;    /* blow-up test cadence in steps (as life_case_set_check_interval);
                           0: every 100 steps, < 0: off */
} life_server_config;

enum { LIFE_REQ_QUEUED = 0, LIFE_REQ_RUNNING = 1, LIFE_REQ_DONE = 2, LIFE_REQ_FAILED = 3 };

typedef struct life_request_info {
    int status;             /* LIFE_REQ_* */
    int steps_done, nsteps;
    int subits;             /* subiterations of the last step */
    int markers;            /* Lagrangian markers of the case (after admission) */
    double t_arrival;       /* seconds since the server was created */
    double t_admit, t_finish;
    double sites;           /* Nx * Ny */
    double bytes;           /* device footprint charged to the budget */
    double total_force[4];  /* Fx Fy Cx Cy after the last step, as life_case_total_force */
    double tip[4];          /* last marker after the last step: x y vx vy */
} life_request_info;

int life_server_create(const life_server_config *cfg, life_server **out);
int life_server_destroy(life_server *s);
/* Queue a request arriving now; keep_fields = 1 keeps rho, u and the marker
 * positions of the last step for life_server_fields. */
int life_server_submit(life_server *s, const life_params *p, const char *geometry, int nsteps, int keep_fields,
                       int *id);
/* One engine step: admit, advance every running case, retire finished ones. */
int life_server_step(life_server *s, int *running, int *queued);
/* Persistent request loop over a timed stream: request k arrives arrival[k]
 * seconds after the call (ascending) and is submitted then; returns when
 * all are finished.  ids[k] receives request k's id. */
int life_server_run(life_server *s, int n, const life_params *p, const char *const *geometry, const int *nsteps,
                    const double *arrival, int *ids);
int life_server_info(const life_server *s, int id, life_request_info *out);
/* Message of a failed request. */
const char *life_server_request_error(const life_server *s, int id);
/* rho[N], u[2N] (reference layout) and marker positions [2*markers] of a
 * finished request submitted with keep_fields = 1. */
int life_server_fields(const life_server *s, int id, double *rho, double *u, double *marker_pos);
int life_server_release(life_server *s, int id);
/* site updates, engine steps, busy seconds, graph captures, peak running
 * cases, peak admitted bytes, engine steps whose next bulk stream-collide
 * was launched early (beside the coupling) */
int life_server_stats(const life_server *s, double out[7]);

#ifdef __cplusplus
}
#endif
#endif
