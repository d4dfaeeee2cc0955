/*
 * Oracle for the LIFE fluid + immersed-boundary + FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, CPU restatement of the
 * reference algorithm (joconnor22/LIFE v1.0.3), kept in the reference's own
 * array-of-structures layout f[(i*Ny+j)*9+v], push streaming and operation
 * order.  It exists to check the HIP product path (tests/, the smoke check
 * and the CPU-baseline leg of bench.py); the product never links or calls it.
 */
#ifndef LIFE_ORACLE_H
#define LIFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice site types, Grid values of eLatType (inc/defs.h:52). */
enum { ORC_FLUID = 0, ORC_WALL, ORC_VELOCITY, ORC_FREESLIP, ORC_PRESSURE, ORC_CONVECTIVE };
/* Inlet profile (inc/defs.h:53); ORC_UNIFORM = PROFILE left undefined. */
enum { ORC_UNIFORM = -1, ORC_PARABOLIC = 0, ORC_SHEAR = 1, ORC_BOUNDARY_LAYER = 2 };

/* Runtime form of the compile-time case header inc/params.h:26-110. */
typedef struct orc_params {
    int Nx, Ny;
    double height_p, rho_p, nu_p;
    double ux0_p, uy0_p;
    double gravityX, gravityY, dpdx, dpdy;
    int wall_left, wall_right, wall_bottom, wall_top;
    int profile;
    double uxInlet_p, uyInlet_p;
    double alpha, delta;            /* Newmark */
    double relaxMax, subTol;        /* coupling */
    double omega;                   /* if > 0 used directly, else from tStep */
    double tStep;
    int central_moments;            /* CENTRAL_MOMENTS */
    double inlet_ramp;              /* INLET_RAMP seconds, <= 0 = undefined */
    int uni_epsilon;                /* UNI_EPSILON */
    double initial_deflect;         /* INITIAL_DEFLECT, 0 = undefined */
    double ref_nu, ref_rho, ref_P, ref_L, ref_U;
    double womersley;               /* WOMERSLEY, 0 = undefined */
} orc_params;

typedef struct orc_case orc_case;

const char *orc_last_error(void);
void orc_set_threads(int n);
int orc_get_threads(void);

orc_case *orc_create(const orc_params *p, const char *geometry);
void orc_destroy(orc_case *c);

/* Derived scalars: omega, tau, nu, Dx, Dt, Dm, Drho, c_s, c_s^2, c_s^4. */
void orc_scalars(const orc_case *c, double out[10]);
int orc_time(const orc_case *c);

/* One main-loop iteration per step (main.cpp:70-86 without output): t++, solver(). */
int orc_step(orc_case *c, int nsteps);
/* lbmKernel only (Grid.cpp:36-100), for kernel-level comparisons. */
int orc_lbm_kernel(orc_case *c);

/* Grid state.  f in AoS [(i*Ny+j)*9+v]; rho [N]; u [N*2]; fibm [N*2]. */
void orc_get_grid(const orc_case *c, double *f, double *rho, double *u, double *fibm);
void orc_set_grid(orc_case *c, const double *f, const double *rho, const double *u, const double *fibm);
void orc_get_types(const orc_case *c, int *type);

/* Markers in global order. */
int orc_num_markers(const orc_case *c);
int orc_num_bodies(const orc_case *c);
void orc_get_markers(const orc_case *c, double *pos, double *vel, double *force,
                     double *eps, double *ds, int *body);
void orc_get_supports(const orc_case *c, int *count, int *ij, double *delta);
void orc_body_info(const orc_case *c, int b, int *info /* ID, flex, nmarkers, first, ndof */);
int orc_body_nodes(const orc_case *c, int b, int *idx /* n */);
int orc_get_fem(const orc_case *c, int b, double *U, double *Udot, double *Udotdot);
int orc_get_fem_state(const orc_case *c, int b, int *sizes /* nel, nnode */, double *Un, double *Unm1,
                      double *Rk, double *el /* 2 nel */, double *nodes /* 6 nnode */);
void orc_coupling_info(const orc_case *c, double *out /* subIt, subRes, relax */);

/* writeTotalForces quantities (Objects.cpp:491-539): Fx, Fy, Cx, Cy. */
void orc_total_force(const orc_case *c, double out[4]);

/* Unit-level restatements. */
double orc_dirac_delta(double dist);
int orc_find_support(double x, double y, double Dx, int Nx, int Ny, int *ij, double *delta);
/* Canonical solve of A x = b on the trailing (n-bc) block (Utils.cpp:288-311). */
int orc_lu_solve(int n, int bc, const double *A, const double *b, double *x);
/* Single-site collision, for invariant tests: rho/u/F at the site -> 9 post-collision values. */
void orc_collide_site(int cm, double omega, const double *fsite, double rho, double ux, double uy,
                      double Fx, double Fy, double *out);
void orc_equilibrium(int cm, double rho, double ux, double uy, double *feq);
/* Clamped horizontal beam alone: static tip deflection under tip load P
 * (nsteps == 0, tip_out[0]) or tip displacement history after release. */
int orc_fem_cantilever(double L, double h, double E, double rho, int nel, double Dt, double P, int nsteps,
                       double *tip_out);

/* Unit-level access, for checking each restated step on its own. */
void orc_ibm_interpolate(orc_case *c);
void orc_ibm_spread(orc_case *c);
void orc_set_markers(orc_case *c, const double *pos, const double *vel, const double *force, const double *eps,
                     const double *ds);
int orc_ibm_supports(orc_case *c);
int orc_epsilon_matrix(const orc_case *c, int sys, double *A);
void orc_get_interp(const orc_case *c, double *rho, double *mom);
int orc_normal_vector(const orc_case *c, int i, int j, int *n, int *nd);
void orc_get_delu(const orc_case *c, double *delU);
void orc_get_inlet(const orc_case *c, double *u_in);
int orc_fem_load(orc_case *c, int b, double *R);
double orc_aitken_relax(int subIt, double relax, double relaxMax, double num, double den);
void orc_sub_residual(int nd, const double *U, const double *Ukm1, double *Rk, const double *Rkm1, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
