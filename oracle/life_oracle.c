/*
 * Oracle for the LIFE fluid + immersed-boundary + FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C FP64 CPU restatement of the reference
 * (joconnor22/LIFE v1.0.3).  Only tests, the smoke check and the CPU-baseline
 * leg of bench.py load it; the HIP product never does.
 *
 * It keeps the reference's AoS layout f[(i*Ny+j)*9+v], push streaming with
 * periodic wrap, and the reference's per-site floating-point operation order,
 * so that it is a bit-level specification of the path.  Every function names
 * the reference file:line it restates.  Built with -ffp-contract=off.
 *
 * One deliberate difference: LAPACK dgetrf/dgetrs (Utils.cpp:288-311, an
 * unpinned external library) is replaced by one canonical unblocked
 * partial-pivot Gaussian elimination, orc_lu_solve(), which the product
 * also implements.
 */
#include "life_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV 9
#define SUPP 9

/* D2Q9 tables (Grid.cpp:1247-1250). */
static const int CX[NV] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
static const int CY[NV] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
static const int OPP[NV] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
static double W[NV];

static char g_err[512];
static int g_threads = 1;

const char *orc_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

static void init_tables(void) {
    W[0] = 4.0 / 9.0;
    for (int v = 1; v < 5; v++) W[v] = 1.0 / 9.0;
    for (int v = 5; v < 9; v++) W[v] = 1.0 / 36.0;
}

/* ------------------------------------------------------------------------- */
/* Structures                                                                 */
/* ------------------------------------------------------------------------- */

typedef struct { int i, j; double d; } supp_t;

typedef struct {
    int body, ID;
    double pos[2], vel[2], force[2];
    double ds, eps, interpMom[2], interpRho;
    int nsupp;
    supp_t supp[SUPP];
} marker_t;

typedef struct { double pos0[2], pos[2], angle0, angle; int dofs[3]; } fnode_t;
typedef struct { int node; double z1, z2; } fmap_t;
typedef struct {
    double L0, L, angle, A, I, E, rho;
    double T[6][6], M[6][6], KL[6][6], KNL[6][6], R[6], F[6];
    int dofs[6], n0, n1, nmap;
    fmap_t *map;
} elem_t;
typedef struct { int el; double zeta; } pmap_t;

typedef struct {
    int body;
    double L0, angle0, tCrit, resNR, subRes, subNum, subDen;
    int ndof, nbc, itNR, nnode, nel;
    double *M, *K, *R, *F, *U, *delU, *Ud, *Udd, *Un, *Unm1, *Unm2, *Udn, *Uddn, *Ukm1, *Rk, *Rkm1;
    fnode_t *node;
    elem_t *el;
    pmap_t *pmap;
} fem_t;

typedef struct {
    int ID, flex /* 0 flexible, 1 rigid (eFlexibleType) */, kind /* 0 circle 1 filament */;
    int n, cap;
    int *nodes; /* indices into the global marker array */
    fem_t *fem;
} body_t;

struct orc_case {
    orc_params p;
    int Nx, Ny;
    long N;
    double cs, cs2, cs4, omega, tau, nu, Dx, Dt, Dm, Drho;
    int t;
    double *f, *f_n, *u, *u_n, *rho, *rho_n, *fxy, *fibm, *delU, *u_in, *rho_in;
    int *type, *bc, nbc;
    /* objects */
    int nm, nb, hasIBM, hasFlex, subIt, simDOFs;
    marker_t *m;
    body_t *b;
    double relax, subRes, subNum, subDen;
};

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static void swapp(double **a, double **b) { double *t = *a; *a = *b; *b = t; }

/* ------------------------------------------------------------------------- */
/* Small helpers                                                              */
/* ------------------------------------------------------------------------- */

/* Utils::diracDelta, Utils.h:220-232 (Roma et al. 1999, 3-point). */
double orc_dirac_delta(double dist) {
    double a = fabs(dist);
    if (a > 1.5) return 0.0;
    if (a > 0.5) return (5.0 - 3.0 * a - sqrt(-3.0 * ((1.0 - a) * (1.0 - a)) + 1.0)) / 6.0;
    return (1.0 + sqrt(1.0 - 3.0 * (a * a))) / 3.0;
}

/* Utils::shiftAngle, Utils.h:235-247. */
static double shift_angle(double a) {
    a = fmod(a + M_PI, 2.0 * M_PI);
    if (a < 0.0) a += 2.0 * M_PI;
    return a - M_PI;
}

/* Utils::sgn, Utils.h:250-255. */
static int sgn(double v) { return (0.0 < v) - (v < 0.0); }

/* Dense dot product with the reference's accumulation (Utils.h:478-494). */
static double vdot(const double *a, const double *b, int n) {
    double r = 0.0;
    for (int i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

/* Canonical solve replacing Utils::solveLAPACK (Utils.cpp:288-311): solve
 * B x = b for the trailing block B = A[bc:, bc:] of the row-major n x n A,
 * leading bc entries of x set to zero.  Right-looking unblocked Gaussian
 * elimination with partial pivoting (first maximum |a_ik|), row updates in
 * increasing column order, back substitution in increasing column order. */
int orc_lu_solve(int n, int bc, const double *A, const double *b, double *x) {
    int m = n - bc;
    double *B = (double *)xcalloc((size_t)m * m, sizeof(double));
    double *y = (double *)xcalloc((size_t)m, sizeof(double));
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < m; j++) B[(size_t)i * m + j] = A[(size_t)(i + bc) * n + (j + bc)];
        y[i] = b[i + bc];
    }
    int info = 0;
    for (int k = 0; k < m; k++) {
        int p = k;
        double best = fabs(B[(size_t)k * m + k]);
        for (int i = k + 1; i < m; i++) {
            double a = fabs(B[(size_t)i * m + k]);
            if (a > best) { best = a; p = i; }
        }
        if (best == 0.0) { info = k + 1; continue; }
        if (p != k) {
            for (int j = 0; j < m; j++) {
                double t = B[(size_t)k * m + j];
                B[(size_t)k * m + j] = B[(size_t)p * m + j];
                B[(size_t)p * m + j] = t;
            }
            double t = y[k]; y[k] = y[p]; y[p] = t;
        }
        double piv = B[(size_t)k * m + k];
        for (int i = k + 1; i < m; i++) {
            double l = B[(size_t)i * m + k] / piv;
            B[(size_t)i * m + k] = l;
            for (int j = k + 1; j < m; j++) B[(size_t)i * m + j] = B[(size_t)i * m + j] - l * B[(size_t)k * m + j];
            y[i] = y[i] - l * y[k];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        double s = y[i];
        for (int j = i + 1; j < m; j++) s = s - B[(size_t)i * m + j] * x[j + bc];
        x[i + bc] = s / B[(size_t)i * m + i];
    }
    for (int i = 0; i < bc; i++) x[i] = 0.0;
    free(B);
    free(y);
    return info;
}

/* ------------------------------------------------------------------------- */
/* LBM: equilibria, forcing, collision (Grid.cpp:103-299)                     */
/* ------------------------------------------------------------------------- */

/* GridClass::equilibrium, Grid.cpp:249-264. */
static inline double feq_one(int cm, double rho, double ux, double uy, int v) {
    int cx = CX[v], cy = CY[v];
    if (cm) {
        double ax = 9.0 * (cx * cx) * (ux * ux) + 6.0 * cx * ux - 3.0 * (ux * ux) + 2.0;
        double ay = 9.0 * (cy * cy) * (uy * uy) + 6.0 * cy * uy - 3.0 * (uy * uy) + 2.0;
        return 0.25 * rho * W[v] * ax * ay;
    }
    double cu = cx * ux + cy * uy;
    double q = ux * ux * (cx * cx - 1.0 / 3.0) + uy * uy * (cy * cy - 1.0 / 3.0);
    return rho * W[v] * (1.0 + 3.0 * cu + 4.5 * q + 9.0 * cx * cy * ux * uy);
}

void orc_equilibrium(int cm, double rho, double ux, double uy, double *feq) {
    init_tables();
    for (int v = 0; v < NV; v++) feq[v] = feq_one(cm, rho, ux, uy, v);
}

/* GridClass::latticeForce (Guo forcing), Grid.cpp:267-279. */
static inline double guo_term(double ux, double uy, double Fx, double Fy, int v) {
    int cx = CX[v], cy = CY[v];
    return 3.0 * W[v] * (Fx * (cx - ux + cx * 3.0 * (cx * ux + cy * uy)) +
                         (Fy * (cy - uy + cy * 3.0 * (cx * ux + cy * uy))));
}

/* Central-moments collision, Grid.cpp:106-223 (De Rosis 2019): returns the
 * nine post-collision populations for one site. */
static void collide_cm(double omega, double cs2, double cs4, const double *fs, double rho,
                       double ux, double uy, double Fx, double Fy, double *fo) {
    double k4p = 0.0, k5p = 0.0;
    for (int v = 0; v < NV; v++) {
        double dx = CX[v] - ux, dy = CY[v] - uy;
        k4p += fs[v] * (dx * dx - dy * dy);
        k5p += fs[v] * dx * dy;
    }
    const double k0 = rho;
    const double k1 = 0.5 * Fx;
    const double k2 = 0.5 * Fy;
    const double k3 = 2.0 * rho * cs2;
    const double k4 = (1.0 - omega) * k4p;
    const double k5 = (1.0 - omega) * k5p;
    const double k6 = 0.5 * Fy * cs2;
    const double k7 = 0.5 * Fx * cs2;
    const double k8 = rho * cs4;
    const double X = ux, Y = uy, X2 = ux * ux, Y2 = uy * uy;

    /* Back-transformation f* = M^-1(u) k; each row is a sum of nine terms
     * taken left to right, coefficient polynomials as in Grid.cpp:143-223. */
    fo[0] = (X2 * Y2 - X2 - Y2 + 1.0) * k0
          + (2.0 * X * Y2 - 2.0 * X) * k1
          + (2.0 * Y * X2 - 2.0 * Y) * k2
          + (0.5 * X2 + 0.5 * Y2 - 1.0) * k3
          + (0.5 * Y2 - 0.5 * X2) * k4
          + 4.0 * X * Y * k5
          + 2.0 * Y * k6
          + 2.0 * X * k7
          + k8;
    fo[1] = (0.5 * X2 - 0.5 * (X2 * Y2) - 0.5 * (X * Y2) + 0.5 * X) * k0
          + (X - X * Y2 - 0.5 * Y2 + 0.5) * k1
          + (-Y * X2 - Y * X) * k2
          + (-0.25 * X2 - 0.25 * X - 0.25 * Y2 + 0.25) * k3
          + (0.25 * X2 + 0.25 * X - 0.25 * Y2 + 0.25) * k4
          + (-Y - 2.0 * X * Y) * k5
          + (-Y) * k6
          + (-X - 0.5) * k7
          - 0.5 * k8;
    fo[2] = (-0.5 * (X2 * Y2) + 0.5 * X2 + 0.5 * (X * Y2) - 0.5 * X) * k0
          + (X - X * Y2 + 0.5 * Y2 - 0.5) * k1
          + (-Y * X2 + Y * X) * k2
          + (-0.25 * X2 + 0.25 * X - 0.25 * Y2 + 0.25) * k3
          + (0.25 * X2 - 0.25 * X - 0.25 * Y2 + 0.25) * k4
          + (Y - 2.0 * X * Y) * k5
          + (-Y) * k6
          + (0.5 - X) * k7
          - 0.5 * k8;
    fo[3] = (0.5 * Y2 - 0.5 * (X2 * Y) - 0.5 * (X2 * Y2) + 0.5 * Y) * k0
          + (-X * Y2 - X * Y) * k1
          + (Y - X2 * Y - 0.5 * X2 + 0.5) * k2
          + (-0.25 * X2 - 0.25 * Y2 - 0.25 * Y + 0.25) * k3
          + (0.25 * X2 - 0.25 * Y2 - 0.25 * Y - 0.25) * k4
          + (-X - 2.0 * X * Y) * k5
          + (-Y - 0.5) * k6
          + (-X) * k7
          - 0.5 * k8;
    fo[4] = (-0.5 * (X2 * Y2) + 0.5 * (X2 * Y) + 0.5 * Y2 - 0.5 * Y) * k0
          + (-X * Y2 + X * Y) * k1
          + (Y - X2 * Y + 0.5 * X2 - 0.5) * k2
          + (-0.25 * X2 - 0.25 * Y2 + 0.25 * Y + 0.25) * k3
          + (0.25 * X2 - 0.25 * Y2 + 0.25 * Y - 0.25) * k4
          + (X - 2.0 * X * Y) * k5
          + (0.5 - Y) * k6
          + (-X) * k7
          - 0.5 * k8;
    fo[5] = (0.25 * (X2 * Y2) + 0.25 * (X2 * Y) + 0.25 * (X * Y2) + 0.25 * (X * Y)) * k0
          + (0.25 * Y + 0.5 * (X * Y) + 0.5 * (X * Y2) + 0.25 * Y2) * k1
          + (0.25 * X + 0.5 * (X * Y) + 0.5 * (X2 * Y) + 0.25 * X2) * k2
          + (0.125 * X2 + 0.125 * X + 0.125 * Y2 + 0.125 * Y) * k3
          + (-0.125 * X2 - 0.125 * X + 0.125 * Y2 + 0.125 * Y) * k4
          + (0.5 * X + 0.5 * Y + X * Y + 0.25) * k5
          + (0.5 * Y + 0.25) * k6
          + (0.5 * X + 0.25) * k7
          + 0.25 * k8;
    fo[6] = (0.25 * (X2 * Y2) - 0.25 * (X2 * Y) - 0.25 * (X * Y2) + 0.25 * (X * Y)) * k0
          + (0.25 * Y - 0.5 * (X * Y) + 0.5 * (X * Y2) - 0.25 * Y2) * k1
          + (0.25 * X - 0.5 * (X * Y) + 0.5 * (X2 * Y) - 0.25 * X2) * k2
          + (0.125 * X2 - 0.125 * X + 0.125 * Y2 - 0.125 * Y) * k3
          + (-0.125 * X2 + 0.125 * X + 0.125 * Y2 - 0.125 * Y) * k4
          + (X * Y - 0.5 * Y - 0.5 * X + 0.25) * k5
          + (0.5 * Y - 0.25) * k6
          + (0.5 * X - 0.25) * k7
          + 0.25 * k8;
    fo[7] = (0.25 * (X2 * Y2) - 0.25 * (X2 * Y) + 0.25 * (X * Y2) - 0.25 * (X * Y)) * k0
          + (0.5 * (X * Y2) - 0.5 * (X * Y) - 0.25 * Y + 0.25 * Y2) * k1
          + (0.5 * (X * Y) - 0.25 * X + 0.5 * (X2 * Y) - 0.25 * X2) * k2
          + (0.125 * X2 + 0.125 * X + 0.125 * Y2 - 0.125 * Y) * k3
          + (-0.125 * X2 - 0.125 * X + 0.125 * Y2 - 0.125 * Y) * k4
          + (0.5 * Y - 0.5 * X + X * Y - 0.25) * k5
          + (0.5 * Y - 0.25) * k6
          + (0.5 * X + 0.25) * k7
          + 0.25 * k8;
    fo[8] = (0.25 * (X2 * Y2) + 0.25 * (X2 * Y) - 0.25 * (X * Y2) - 0.25 * (X * Y)) * k0
          + (0.5 * (X * Y) - 0.25 * Y + 0.5 * (X * Y2) - 0.25 * Y2) * k1
          + (0.5 * (X2 * Y) - 0.5 * (X * Y) - 0.25 * X + 0.25 * X2) * k2
          + (0.125 * X2 - 0.125 * X + 0.125 * Y2 + 0.125 * Y) * k3
          + (-0.125 * X2 + 0.125 * X + 0.125 * Y2 + 0.125 * Y) * k4
          + (0.5 * X - 0.5 * Y + X * Y - 0.25) * k5
          + (0.5 * Y + 0.25) * k6
          + (0.5 * X - 0.25) * k7
          + 0.25 * k8;
}

/* BGK collision with Guo forcing, Grid.cpp:236-244. */
static void collide_bgk(double omega, const double *fs, double rho, double ux, double uy,
                        double Fx, double Fy, double *fo) {
    for (int v = 0; v < NV; v++)
        fo[v] = fs[v] + omega * (feq_one(0, rho, ux, uy, v) - fs[v]) +
                (1.0 - 0.5 * omega) * guo_term(ux, uy, Fx, Fy, v);
}

void orc_collide_site(int cm, double omega, const double *fsite, double rho, double ux, double uy,
                      double Fx, double Fy, double *out) {
    init_tables();
    double cs = 1.0 / sqrt(3.0);
    double cs2 = cs * cs, cs4 = cs * cs * cs * cs;
    if (cm) collide_cm(omega, cs2, cs4, fsite, rho, ux, uy, Fx, Fy, out);
    else collide_bgk(omega, fsite, rho, ux, uy, Fx, Fy, out);
}

/* GridClass::macroscopic, Grid.cpp:282-299 (F = force_xy). */
static inline void macro_site(const double *fs, double Fx, double Fy, double *rho, double *ux, double *uy) {
    double r = 0.0, mx = 0.0, my = 0.0;
    for (int v = 0; v < NV; v++) {
        r += fs[v];
        mx += CX[v] * fs[v];
        my += CY[v] * fs[v];
    }
    *rho = r;
    *ux = (mx + 0.5 * Fx) / r;
    *uy = (my + 0.5 * Fy) / r;
}

/* ------------------------------------------------------------------------- */
/* Grid set-up (Grid.cpp:916-1062, 1232-1289)                                 */
/* ------------------------------------------------------------------------- */

static int grid_init(orc_case *c) {
    const orc_params *p = &c->p;
    int Nx = c->Nx = p->Nx, Ny = c->Ny = p->Ny;
    long N = c->N = (long)Nx * Ny;
    if (Nx < 4 || Ny < 4) { set_err("grid must be at least 4x4"); return -1; }
    const double rho0 = 1.0;
    c->cs = 1.0 / sqrt(3.0);
    c->cs2 = c->cs * c->cs;
    c->cs4 = c->cs * c->cs * c->cs * c->cs;
    if (p->omega > 0.0) {
        c->omega = p->omega;
    } else {
        /* params.h:89-90, pow(x, 2.0) == x*x correctly rounded */
        double s = 1.0 / sqrt(3.0);
        double h = p->height_p / (Ny - 1);
        c->omega = 1.0 / (p->nu_p * p->tStep / ((s * s) * (h * h)) + 0.5);
    }
    c->t = 0;
    c->tau = 1.0 / c->omega;
    c->nu = (c->tau - 0.5) * c->cs2;
    c->Dx = p->height_p / (Ny - 1);
    c->Dt = (c->Dx * c->Dx) * c->nu / p->nu_p;
    c->Dm = (p->rho_p / rho0) * (c->Dx * c->Dx * c->Dx);
    c->Drho = p->rho_p / rho0;

    c->u = xcalloc((size_t)N * 2, sizeof(double));
    c->u_n = xcalloc((size_t)N * 2, sizeof(double));
    c->rho = xcalloc((size_t)N, sizeof(double));
    c->rho_n = xcalloc((size_t)N, sizeof(double));
    for (long k = 0; k < N; k++) c->rho[k] = c->rho_n[k] = rho0;
    c->fxy = xcalloc((size_t)N * 2, sizeof(double));
    c->fibm = xcalloc((size_t)N * 2, sizeof(double));
    c->type = xcalloc((size_t)N, sizeof(int));
    c->f = xcalloc((size_t)N * NV, sizeof(double));
    c->f_n = xcalloc((size_t)N * NV, sizeof(double));
    c->u_in = xcalloc((size_t)Ny * 2, sizeof(double));
    c->rho_in = xcalloc((size_t)Ny, sizeof(double));
    for (int j = 0; j < Ny; j++) c->rho_in[j] = rho0;
    c->delU = xcalloc((size_t)Ny * 2, sizeof(double));
    c->bc = xcalloc((size_t)2 * (Nx + Ny), sizeof(int));

    if (p->wall_left == ORC_CONVECTIVE || p->wall_bottom == ORC_CONVECTIVE || p->wall_top == ORC_CONVECTIVE) {
        set_err("Currently convective BC only supported for right boundary");
        return -1;
    }
    /* type map, bottom/top override left/right at corners (Grid.cpp:931-957) */
    c->nbc = 0;
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) {
            long id = (long)i * Ny + j;
            if (i == 0) c->type[id] = p->wall_left;
            else if (i == Nx - 1) c->type[id] = p->wall_right;
            if (j == 0) c->type[id] = p->wall_bottom;
            else if (j == Ny - 1) c->type[id] = p->wall_top;
            if (c->type[id] != ORC_FLUID) c->bc[c->nbc++] = (int)id;
        }
    /* inlet profile (Grid.cpp:959-1001) */
    for (int j = 0; j < Ny; j++) {
        double *ui = &c->u_in[j * 2];
        if (p->profile == ORC_UNIFORM) {
            ui[0] = p->uxInlet_p * c->Dt / c->Dx;
            ui[1] = p->uyInlet_p * c->Dt / c->Dx;
        } else if (p->profile == ORC_PARABOLIC) {
            double R = p->height_p / 2.0;
            double Y = j * c->Dx - R;
            ui[0] = 1.5 * (p->uxInlet_p * c->Dt / c->Dx) * (1.0 - (Y / R) * (Y / R));
            ui[1] = 1.5 * (p->uyInlet_p * c->Dt / c->Dx) * (1.0 - (Y / R) * (Y / R));
        } else if (p->profile == ORC_SHEAR) {
            double H = p->height_p, Y = j * c->Dx;
            ui[0] = (p->uxInlet_p * c->Dt / c->Dx) * (Y / H);
            ui[1] = (p->uyInlet_p * c->Dt / c->Dx) * (Y / H);
        } else {
            double H = p->height_p, Y = j * c->Dx;
            ui[0] = ((1.5 * p->uxInlet_p * c->Dt / c->Dx) / (H * H)) * Y * (2.0 * H - Y);
            ui[1] = ((1.5 * p->uyInlet_p * c->Dt / c->Dx) / (H * H)) * Y * (2.0 * H - Y);
        }
    }
    /* initial velocity (Grid.cpp:1004-1033) */
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) {
            long id = (long)i * Ny + j;
            if (p->inlet_ramp > 0.0) {
                c->u[id * 2] = 0.0;
                c->u[id * 2 + 1] = 0.0;
            } else if (p->profile != ORC_UNIFORM) {
                c->u[id * 2] = c->u_in[j * 2];
                c->u[id * 2 + 1] = c->u_in[j * 2 + 1];
            } else {
                c->u[id * 2] = p->ux0_p * c->Dt / c->Dx;
                c->u[id * 2 + 1] = p->uy0_p * c->Dt / c->Dx;
            }
            if (c->type[id] == ORC_WALL) c->u[id * 2] = c->u[id * 2 + 1] = 0.0;
        }
    memcpy(c->u_n, c->u, sizeof(double) * N * 2);
    memcpy(c->rho_n, c->rho, sizeof(double) * N);
    /* body force (Grid.cpp:1035-1045) */
    for (long id = 0; id < N; id++) {
        c->fxy[id * 2] = (c->rho[id] * c->Drho * p->gravityX + p->dpdx) * ((c->Dx * c->Dt) * (c->Dx * c->Dt)) / c->Dm;
        c->fxy[id * 2 + 1] = (c->rho[id] * c->Drho * p->gravityY + p->dpdy) * ((c->Dx * c->Dt) * (c->Dx * c->Dt)) / c->Dm;
    }
    /* equilibrium populations (Grid.cpp:1048-1061) */
    for (long id = 0; id < N; id++)
        for (int v = 0; v < NV; v++)
            c->f[id * NV + v] = feq_one(p->central_moments, c->rho_n[id], c->u_n[id * 2], c->u_n[id * 2 + 1], v);
    memcpy(c->f_n, c->f, sizeof(double) * N * NV);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* LBM step (Grid.cpp:36-100) and boundary conditions (Grid.cpp:302-556)      */
/* ------------------------------------------------------------------------- */

/* GridClass::getRampCoefficient, Grid.cpp:548-556. */
static double ramp_coeff(const orc_case *c) {
    if (c->p.inlet_ramp > 0.0 && c->Dt * c->t <= c->p.inlet_ramp)
        return (1.0 - cos(M_PI * c->Dt * c->t / c->p.inlet_ramp)) / 2.0;
    return 1.0;
}

/* GridClass::getNormalVector, Grid.cpp:498-545.  Returns 0 or -1 on error. */
static int normal_vector(const orc_case *c, int i, int j, int n[2], int *nd) {
    int Nx = c->Nx, Ny = c->Ny;
    n[0] = n[1] = 0;
    *nd = 0;
    if (i == 0) { n[0] = 1; *nd = 0; }
    else if (i == Nx - 1) { n[0] = -1; *nd = 0; }
    if (j == 0) { n[1] = 1; *nd = 1; }
    else if (j == Ny - 1) { n[1] = -1; *nd = 1; }
    if (n[0] != 0 && n[1] != 0) {
        int tx = c->type[(long)(i + n[0]) * Ny + j];
        int ty = c->type[(long)i * Ny + j + n[1]];
        if (tx == ORC_FLUID && ty == ORC_FLUID) { set_err("Corner node is surrounded by fluid lattice sites"); return -1; }
        else if (tx == ORC_FLUID) { *nd = 0; n[1] = 0; }
        else if (ty == ORC_FLUID) { *nd = 1; n[0] = 0; }
    }
    return 0;
}

static inline long sid(const orc_case *c, int i, int j) { return (long)i * c->Ny + j; }

/* GridClass::regularisedBC, Grid.cpp:387-465 (Latt et al. 2008). */
static void regularised_bc(orc_case *c, int i, int j, long id, const int n[2], int nd) {
    int cm = c->p.central_moments;
    int ty = c->type[id];
    double *fs = &c->f[id * NV];
    int velocity_like = (ty == ORC_VELOCITY || ty == ORC_WALL || ty == ORC_FREESLIP);
    int corner = (n[0] != 0 && n[1] != 0);
    if (corner) {
        if (velocity_like)  /* Utils::extrapolate order 1, Utils.h:150-162 */
            c->rho_n[id] = 2.0 * c->rho[sid(c, i + n[0], j + n[1])] - c->rho[sid(c, i + 2 * n[0], j + 2 * n[1])];
    } else {
        double fplus = 0.0, fzero = 0.0;
        int nn = n[nd];
        for (int v = 0; v < NV; v++) {
            int cn = nd == 0 ? CX[v] : CY[v];
            if (cn == -nn) fplus += fs[v];
            else if (cn == 0) fzero += fs[v];
        }
        if (velocity_like)
            c->rho_n[id] = (2.0 * fplus + fzero) / (1.0 - nn * c->u_n[id * 2 + nd]);
        else if (ty == ORC_PRESSURE)
            c->u_n[id * 2 + nd] = nn * (1.0 - (2.0 * fplus + fzero) / c->rho_n[id]);
    }
    double rho = c->rho_n[id], ux = c->u_n[id * 2], uy = c->u_n[id * 2 + 1];
    double Sxx = 0.0, Syy = 0.0, Sxy = 0.0;
    for (int v = 0; v < NV; v++) {
        double feq = feq_one(cm, rho, ux, uy, v);
        int unknown;
        if (corner) unknown = (CX[v] == n[0] || CY[v] == n[1]);
        else unknown = ((nd == 0 ? CX[v] : CY[v]) == n[nd]);
        if (unknown) {
            if (corner && n[0] * CX[v] + n[1] * CY[v] == 0) fs[v] = feq;
            else fs[v] = feq + (fs[OPP[v]] - feq_one(cm, rho, ux, uy, OPP[v]));
        }
        double fneq = fs[v] - feq;
        Sxx += CX[v] * CX[v] * fneq;
        Syy += CY[v] * CY[v] * fneq;
        Sxy += CX[v] * CY[v] * fneq;
    }
    for (int v = 0; v < NV; v++)
        fs[v] = feq_one(cm, rho, ux, uy, v) +
                (W[v] / (2.0 * c->cs4)) * (((CX[v] * CX[v] - c->cs2) * Sxx) + ((CY[v] * CY[v] - c->cs2) * Syy) +
                                           (2.0 * CX[v] * CY[v] * Sxy));
}

/* GridClass::applyBCs, Grid.cpp:302-384. */
static int apply_bc(orc_case *c, int i, int j, long id, double ramp) {
    int n[2], nd;
    if (normal_vector(c, i, j, n, &nd)) return -1;
    switch (c->type[id]) {
    case ORC_WALL:
        c->u_n[id * 2] = 0.0;
        c->u_n[id * 2 + 1] = 0.0;
        regularised_bc(c, i, j, id, n, nd);
        break;
    case ORC_VELOCITY:
        c->u_n[id * 2] = c->u_in[j * 2] * ramp;
        c->u_n[id * 2 + 1] = c->u_in[j * 2 + 1] * ramp;
        regularised_bc(c, i, j, id, n, nd);
        break;
    case ORC_FREESLIP:
    case ORC_PRESSURE:
        if (c->type[id] == ORC_FREESLIP) c->u_n[id * 2 + nd] = 0.0;
        else c->rho_n[id] = c->rho_in[j];
        for (int d = 0; d < 2; d++)
            if (d != nd) /* Utils::zeroGradient order 2, Utils.h:200-210 */
                c->u_n[id * 2 + d] = (4.0 / 3.0) * c->u[sid(c, i + n[0], j + n[1]) * 2 + d] -
                                     (1.0 / 3.0) * c->u[sid(c, i + 2 * n[0], j + 2 * n[1]) * 2 + d];
        regularised_bc(c, i, j, id, n, nd);
        break;
    case ORC_CONVECTIVE: {
        /* GridClass::convectiveBC, Grid.cpp:468-474 */
        static const int vs[3] = {2, 6, 8};
        for (int q = 0; q < 3; q++) {
            int v = vs[q];
            c->f[id * NV + v] = c->f_n[id * NV + v] +
                                3.0 * W[v] * (c->delU[j * 2] * CX[v] + c->delU[j * 2 + 1] * CY[v]);
        }
        break;
    }
    default:
        break;
    }
    return 0;
}

/* GridClass::convectiveSpeed, Grid.cpp:477-495 (serial mean). */
static void convective_speed(orc_case *c) {
    int Nx = c->Nx, Ny = c->Ny;
    double uOut = 0.0;
    for (int j = 0; j < Ny; j++) uOut += c->u[sid(c, Nx - 1, j) * 2];
    uOut /= (double)Ny;
    for (int j = 0; j < Ny; j++)
        for (int d = 0; d < 2; d++)
            c->delU[j * 2 + d] = (-uOut / 2.0) * (3.0 * c->u[sid(c, Nx - 1, j) * 2 + d] -
                                                  4.0 * c->u[sid(c, Nx - 2, j) * 2 + d] +
                                                  c->u[sid(c, Nx - 3, j) * 2 + d]);
}

/* GridClass::lbmKernel, Grid.cpp:36-100. */
int orc_lbm_kernel(orc_case *c) {
    int Nx = c->Nx, Ny = c->Ny;
    long N = c->N;
    int cm = c->p.central_moments;
    if (c->p.wall_right == ORC_CONVECTIVE) convective_speed(c);
    swapp(&c->f_n, &c->f);
    swapp(&c->u_n, &c->u);
    swapp(&c->rho_n, &c->rho);
    /* WOMERSLEY: oscillating pressure-gradient forcing, Grid.cpp:52-61 */
    if (c->p.womersley != 0.0) {
        const orc_params *p = &c->p;
        const double period = ((p->height_p * p->height_p) * M_PI) / (2.0 * (p->womersley * p->womersley) * p->nu_p);
        const double wave = cos(2.0 * M_PI * c->t * c->Dt / period);
        const double sc = (c->Dx * c->Dt) * (c->Dx * c->Dt);
        for (long id = 0; id < N; id++) {
            c->fxy[id * 2] = (c->rho_n[id] * c->Drho * p->gravityX + p->dpdx * wave) * sc / c->Dm;
            c->fxy[id * 2 + 1] = (c->rho_n[id] * c->Drho * p->gravityY + p->dpdy * wave) * sc / c->Dm;
        }
    }
    const double omega = c->omega, cs2 = c->cs2, cs4 = c->cs4;
    double *f = c->f;
    const double *fn = c->f_n, *un = c->u_n, *rn = c->rho_n, *fxy = c->fxy, *fib = c->fibm;

#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long id = 0; id < N; id++) {
        int i = (int)(id / Ny), j = (int)(id - (long)i * Ny);
        double fo[NV];
        double Fx = fxy[id * 2] + fib[id * 2], Fy = fxy[id * 2 + 1] + fib[id * 2 + 1];
        if (cm) collide_cm(omega, cs2, cs4, &fn[id * NV], rn[id], un[id * 2], un[id * 2 + 1], Fx, Fy, fo);
        else collide_bgk(omega, &fn[id * NV], rn[id], un[id * 2], un[id * 2 + 1], Fx, Fy, fo);
        for (int v = 0; v < NV; v++) {
            long r = (long)((i + CX[v] + Nx) % Nx) * Ny + ((j + CY[v] + Ny) % Ny);
            f[r * NV + v] = fo[v];
        }
    }
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long id = 0; id < N; id++)
        if (c->type[id] == ORC_FLUID)
            macro_site(&f[id * NV], fxy[id * 2], fxy[id * 2 + 1], &c->rho[id], &c->u[id * 2], &c->u[id * 2 + 1]);

    double ramp = ramp_coeff(c);
    int err = 0;
    for (int b = 0; b < c->nbc; b++) {
        long id = c->bc[b];
        int i = (int)(id / Ny), j = (int)(id - (long)i * Ny);
        if (apply_bc(c, i, j, id, ramp)) err = -1;
        macro_site(&f[id * NV], fxy[id * 2], fxy[id * 2 + 1], &c->rho[id], &c->u[id * 2], &c->u[id * 2 + 1]);
    }
    return err;
}

/* ------------------------------------------------------------------------- */
/* IBM markers (IBMNode.cpp)                                                  */
/* ------------------------------------------------------------------------- */

/* IBMNodeClass::findSupport, IBMNode.cpp:139-179. */
int orc_find_support(double x, double y, double Dx, int Nx, int Ny, int *ij, double *delta) {
    int inear = (int)round(x / Dx), jnear = (int)round(y / Dx);
    int n = 0;
    for (int i = inear - 2; i <= inear + 2; i++) {
        double dx = fabs(x / Dx - i);
        for (int j = jnear - 2; j <= jnear + 2; j++) {
            double dy = fabs(y / Dx - j);
            if (dx < 1.5 && dy < 1.5 && i >= 0 && i <= Nx - 1 && j >= 0 && j <= Ny - 1) {
                if (n == SUPP) return -1;
                ij[2 * n] = i;
                ij[2 * n + 1] = j;
                delta[n] = orc_dirac_delta(dx) * orc_dirac_delta(dy);
                n++;
            }
        }
    }
    return n;
}

static int marker_support(orc_case *c, marker_t *m) {
    int ij[2 * SUPP];
    double d[SUPP];
    int n = orc_find_support(m->pos[0], m->pos[1], c->Dx, c->Nx, c->Ny, ij, d);
    if (n < 0) { set_err("Support buffer size is not big enough"); return -1; }
    for (int s = 0; s < SUPP; s++) { m->supp[s].i = 0; m->supp[s].j = 0; m->supp[s].d = 0.0; }
    for (int s = 0; s < n; s++) { m->supp[s].i = ij[2 * s]; m->supp[s].j = ij[2 * s + 1]; m->supp[s].d = d[s]; }
    m->nsupp = n;
    return 0;
}

/* IBMNodeClass::computeDs, IBMNode.cpp:182-204. */
static void marker_ds(orc_case *c, marker_t *m) {
    const body_t *b = &c->b[m->body];
    double ds = 10.0;
    for (int k = 0; k < b->n; k++) {
        const marker_t *o = &c->m[b->nodes[k]];
        if (o->ID == m->ID) continue;
        double d[2] = {m->pos[0] - o->pos[0], m->pos[1] - o->pos[1]};
        double mag = sqrt(vdot(d, d, 2)) / c->Dx;
        if (mag < ds) ds = mag;
    }
    m->ds = ds;
}

/* IBMNodeClass::interpolate + forceCalc, IBMNode.cpp:26-58. */
static void marker_interp_force(orc_case *c, marker_t *m) {
    int Ny = c->Ny;
    m->interpRho = 0.0;
    m->interpMom[0] = m->interpMom[1] = 0.0;
    for (int s = 0; s < m->nsupp; s++) {
        long id = (long)m->supp[s].i * Ny + m->supp[s].j;
        m->interpRho += c->rho[id] * m->supp[s].d * 1.0 * 1.0;
        for (int d = 0; d < 2; d++) m->interpMom[d] += c->rho[id] * c->u[id * 2 + d] * m->supp[s].d * 1.0 * 1.0;
    }
    double vs = c->Dt / c->Dx;
    for (int d = 0; d < 2; d++) m->force[d] = 2.0 * (vs * m->interpRho * m->vel[d] - m->interpMom[d]);
}

/* IBMNodeClass::spread, IBMNode.cpp:61-94 (ORDERED). */
static void marker_spread(orc_case *c, const marker_t *m) {
    for (int s = 0; s < m->nsupp; s++) {
        long id = (long)m->supp[s].i * c->Ny + m->supp[s].j;
        double Fx = m->force[0] * m->eps * m->ds * 1.0 * m->supp[s].d;
        double Fy = m->force[1] * m->eps * m->ds * 1.0 * m->supp[s].d;
        c->fibm[id * 2] += Fx;
        c->fibm[id * 2 + 1] += Fy;
    }
}

/* IBMNodeClass::updateMacroscopic, IBMNode.cpp:97-136. */
static void marker_update_macro(orc_case *c, const marker_t *m) {
    for (int s = 0; s < m->nsupp; s++) {
        long id = (long)m->supp[s].i * c->Ny + m->supp[s].j;
        const double *fs = &c->f[id * NV];
        double r = 0.0, mx = 0.0, my = 0.0;
        for (int v = 0; v < NV; v++) {
            r += fs[v];
            mx += CX[v] * fs[v];
            my += CY[v] * fs[v];
        }
        mx = (mx + 0.5 * (c->fxy[id * 2] + c->fibm[id * 2])) / r;
        my = (my + 0.5 * (c->fxy[id * 2 + 1] + c->fibm[id * 2 + 1])) / r;
        c->rho[id] = r;
        c->u[id * 2] = mx;
        c->u[id * 2 + 1] = my;
    }
}

/* ------------------------------------------------------------------------- */
/* FEM (FEMBody.cpp, FEMElement.cpp, FEMNode.cpp)                              */
/* ------------------------------------------------------------------------- */

static void mat6_tmul(double T[6][6], double A[6][6], double out[6][6]) {
    /* Transpose(T) * A * T, evaluated left to right (FEMElement.cpp:102, 114) */
    double tmp[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += T[k][i] * A[k][j];
            tmp[i][j] = s;
        }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += tmp[i][k] * T[k][j];
            out[i][j] = s;
        }
}

static void tvec6(double T[6][6], const double *v, double *out) {
    /* Transpose(T) * v (FEMElement.cpp:64, 92) */
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int j = 0; j < 6; j++) s += T[j][i] * v[j];
        out[i] = s;
    }
}

static void vec6(double T[6][6], const double *v, double *out) {
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int j = 0; j < 6; j++) s += T[i][j] * v[j];
        out[i] = s;
    }
}

/* FEMElementClass::setElementTransform, FEMElement.cpp:255-262. */
static void el_transform(elem_t *e) {
    double ca = cos(e->angle), sa = sin(e->angle);
    e->T[0][0] = e->T[1][1] = e->T[3][3] = e->T[4][4] = ca;
    e->T[0][1] = e->T[3][4] = sa;
    e->T[1][0] = e->T[4][3] = -sa;
    e->T[2][2] = e->T[5][5] = 1.0;
}

/* FEMElementClass::setLocalMatrices, FEMElement.cpp:203-252. */
static void el_local(elem_t *e) {
    memset(e->M, 0, sizeof e->M);
    memset(e->KL, 0, sizeof e->KL);
    memset(e->KNL, 0, sizeof e->KNL);
    memset(e->R, 0, sizeof e->R);
    memset(e->F, 0, sizeof e->F);
    double L0 = e->L0, E = e->E, A = e->A, I = e->I;
    double C1 = e->rho * A * L0 / 420.0;
    e->M[0][0] = C1 * 140.0;
    e->M[0][3] = C1 * 70.0;
    e->M[1][1] = C1 * 156.0;
    e->M[1][2] = C1 * 22.0 * L0;
    e->M[1][4] = C1 * 54;
    e->M[1][5] = C1 * (-13.0 * L0);
    e->M[2][2] = C1 * 4.0 * (L0 * L0);
    e->M[2][4] = C1 * 13.0 * L0;
    e->M[2][5] = C1 * (-3.0 * (L0 * L0));
    e->M[3][3] = C1 * 140.0;
    e->M[4][4] = C1 * 156.0;
    e->M[4][5] = C1 * (-22.0 * L0);
    e->M[5][5] = C1 * 4.0 * (L0 * L0);
    double L2 = L0 * L0, L3 = L0 * L0 * L0;
    e->KL[0][0] = E * A / L0;
    e->KL[0][3] = -E * A / L0;
    e->KL[1][1] = 12.0 * E * I / L3;
    e->KL[1][2] = 6.0 * E * I / L2;
    e->KL[1][4] = -12.0 * E * I / L3;
    e->KL[1][5] = 6.0 * E * I / L2;
    e->KL[2][2] = 4.0 * E * I / L0;
    e->KL[2][4] = -6.0 * E * I / L2;
    e->KL[2][5] = 2.0 * E * I / L0;
    e->KL[3][3] = E * A / L0;
    e->KL[4][4] = 12.0 * E * I / L3;
    e->KL[4][5] = -6.0 * E * I / L2;
    e->KL[5][5] = 4.0 * E * I / L0;
    for (int i = 1; i < 6; i++)
        for (int j = 0; j < i; j++) {
            e->M[i][j] = e->M[j][i];
            e->KL[i][j] = e->KL[j][i];
        }
}

/* FEMElementClass::shapeFuns, FEMElement.cpp:141-160. */
static void el_shape(const elem_t *e, const double *v, double zeta, double out[2]) {
    double z = (zeta + 1.0) / 2.0;
    double N0 = 1.0 - (zeta + 1.0) / 2.0;
    double N1 = 1.0 - 3.0 * (z * z) + 2.0 * (z * z * z);
    double N2 = (z - 2.0 * (z * z) + (z * z * z)) * e->L;
    double N3 = z;
    double N4 = 3.0 * (z * z) - 2.0 * (z * z * z);
    double N5 = (-(z * z) + (z * z * z)) * e->L;
    out[0] = v[0] * N0 + v[3] * N3;
    out[1] = v[1] * N1 + v[2] * N2 + v[4] * N4 + v[5] * N5;
}

static fem_t *fem_create(orc_case *c, body_t *b, const double pos[2], const double geom[2], double angle,
                         const char *nElStr, const char *bcStr, double rho, double E) {
    fem_t *F = xcalloc(1, sizeof(fem_t));
    F->nbc = strcmp(bcStr, "CLAMPED") == 0 ? 3 : (strcmp(bcStr, "SUPPORTED") == 0 ? 2 : 0);
    F->angle0 = angle;
    F->L0 = geom[0];
    int nel = strcmp(nElStr, "CONFORMING") == 0 ? (int)floor(F->L0 / c->Dx) : (int)strtod(nElStr, NULL);
    F->nel = nel;
    F->nnode = nel + 1;
    F->ndof = F->nnode * 3;
    F->tCrit = (F->L0 / nel) / sqrt(E / rho);
    double ca = cos(angle), sa = sin(angle);
    F->node = xcalloc(F->nnode, sizeof(fnode_t));
    for (int i = 0; i < F->nnode; i++) {
        double px = i * F->L0 / nel, py = 0.0;
        double rx = 0.0 + ca * px + (-sa) * py, ry = 0.0 + sa * px + ca * py;
        fnode_t *n = &F->node[i];
        n->pos0[0] = n->pos[0] = pos[0] + rx;
        n->pos0[1] = n->pos[1] = pos[1] + ry;
        n->angle0 = n->angle = angle;
        for (int d = 0; d < 3; d++) n->dofs[d] = i * 3 + d;
    }
    F->el = xcalloc(nel, sizeof(elem_t));
    for (int i = 0; i < nel; i++) {
        elem_t *e = &F->el[i];
        e->L0 = e->L = F->L0 / nel;
        e->angle = angle;
        e->A = c->Dx * geom[1];
        e->I = c->Dx * (geom[1] * geom[1] * geom[1]) / 12.0;
        e->E = E;
        e->rho = rho;
        e->n0 = i;
        e->n1 = i + 1;
        for (int k = 0; k < 2; k++)
            for (int d = 0; d < 3; d++) e->dofs[k * 3 + d] = (i + k) * 3 + d;
        memset(e->T, 0, sizeof e->T);
        el_transform(e);
        el_local(e);
    }
    /* FEMBodyClass::computeNodeMapping, FEMBody.cpp:292-338 */
    int nI = b->n, nF = F->nnode;
    F->pmap = xcalloc(nI, sizeof(pmap_t));
    F->pmap[0].el = 0;
    F->pmap[0].zeta = -1.0;
    for (int i = 1; i < nI; i++) {
        double r = i * ((nF - 1.0) / (nI - 1.0));
        if (fmod(r, 1.0) == 0.0) {
            F->pmap[i].el = (int)floor(r) - 1;
            F->pmap[i].zeta = 1.0;
        } else {
            F->pmap[i].el = (int)floor(r);
            F->pmap[i].zeta = fmod(r, 1.0) * 2.0 - 1.0;
        }
    }
    for (int el = 0; el < nel; el++) {
        elem_t *e = &F->el[el];
        e->map = xcalloc(nI, sizeof(fmap_t));
        for (int nd = 0; nd < nI; nd++) {
            double n1 = -1 + 2.0 * (((double)nd - 0.5) * ((nF - 1.0) / (nI - 1.0)) - (double)el) / 1.0;
            double n2 = -1 + 2.0 * (((double)nd + 0.5) * ((nF - 1.0) / (nI - 1.0)) - (double)el) / 1.0;
            if ((n1 > -1.0 && n1 < 1.0) || (n2 > -1.0 && n2 < 1.0)) {
                if (n1 < -1.0) n1 = -1.0;
                if (n2 > 1.0) n2 = 1.0;
                e->map[e->nmap].node = nd;
                e->map[e->nmap].z1 = n1;
                e->map[e->nmap].z2 = n2;
                e->nmap++;
            }
        }
    }
    int nd = F->ndof;
    F->M = xcalloc((size_t)nd * nd, sizeof(double));
    F->K = xcalloc((size_t)nd * nd, sizeof(double));
    double **vecs[] = {&F->R, &F->F, &F->U, &F->delU, &F->Ud, &F->Udd, &F->Un, &F->Unm1, &F->Unm2,
                       &F->Udn, &F->Uddn, &F->Ukm1, &F->Rk, &F->Rkm1};
    for (size_t k = 0; k < sizeof vecs / sizeof vecs[0]; k++) *vecs[k] = xcalloc(nd, sizeof(double));
    return F;
}

/* FEMBodyClass::updateFEMValues, FEMBody.cpp:198-223. */
static void fem_update_fem(fem_t *F) {
    for (int n = 0; n < F->nnode; n++) {
        fnode_t *nd = &F->node[n];
        for (int d = 0; d < 2; d++) nd->pos[d] = nd->pos0[d] + F->U[nd->dofs[d]];
        nd->angle = nd->angle0 + F->U[nd->dofs[2]];
    }
    for (int el = 0; el < F->nel; el++) {
        elem_t *e = &F->el[el];
        double ev[2] = {F->node[e->n1].pos[0] - F->node[e->n0].pos[0], F->node[e->n1].pos[1] - F->node[e->n0].pos[1]};
        e->angle = atan2(ev[1], ev[0]);
        e->L = sqrt(vdot(ev, ev, 2));
        el_transform(e);
    }
}

/* FEMBodyClass::updateIBMValues, FEMBody.cpp:147-195. */
static void fem_update_ibm(orc_case *c, fem_t *F) {
    body_t *b = &c->b[F->body];
    for (int i = 0; i < b->n; i++) {
        elem_t *e = &F->el[F->pmap[i].el];
        double zeta = F->pmap[i].zeta;
        double dU[6], dUd[6], lU[6], lUd[6];
        for (int k = 0; k < 6; k++) { dU[k] = F->U[e->dofs[k]]; dUd[k] = F->Ud[e->dofs[k]]; }
        const fnode_t *nn[2] = {&F->node[e->n0], &F->node[e->n1]};
        for (int n = 0; n < 2; n++) {
            dU[n * 3 + 0] += nn[n]->pos0[0] - nn[0]->pos[0];
            dU[n * 3 + 1] += nn[n]->pos0[1] - nn[0]->pos[1];
            dU[n * 3 + 2] += nn[n]->angle0 - e->angle;
            dU[n * 3 + 2] = shift_angle(dU[n * 3 + 2]);
        }
        vec6(e->T, dU, lU);
        vec6(e->T, dUd, lUd);
        double s[2], sd[2];
        el_shape(e, lU, zeta, s);
        el_shape(e, lUd, zeta, sd);
        double gs[2], gsd[2];
        for (int r = 0; r < 2; r++) {
            gs[r] = 0.0 + e->T[0][r] * s[0] + e->T[1][r] * s[1];
            gsd[r] = 0.0 + e->T[0][r] * sd[0] + e->T[1][r] * sd[1];
        }
        marker_t *m = &c->m[b->nodes[i]];
        m->pos[0] = nn[0]->pos[0] + gs[0];
        m->pos[1] = nn[0]->pos[1] + gs[1];
        m->vel[0] = gsd[0];
        m->vel[1] = gsd[1];
    }
}

/* FEMElementClass::loadVector, FEMElement.cpp:27-69. */
static void el_load(orc_case *c, fem_t *F, elem_t *e) {
    body_t *b = &c->b[F->body];
    double wt[2] = {e->rho * e->A * c->p.gravityX, e->rho * e->A * c->p.gravityY};
    double fs = c->Dm / (c->Dt * c->Dt);
    double L = e->L;
    for (int n = 0; n < e->nmap; n++) {
        const marker_t *m = &c->m[b->nodes[e->map[n].node]];
        double a = e->map[n].z1, bb = e->map[n].z2;
        double sc = -m->eps * 1.0 * fs;
        double vv[2] = {sc * m->force[0] + wt[0], sc * m->force[1] + wt[1]};
        double Fl[2];
        for (int r = 0; r < 2; r++) Fl[r] = 0.0 + e->T[r][0] * vv[0] + e->T[r][1] * vv[1];
        double a2 = a * a, b2 = bb * bb;
        e->R[0] = Fl[0] * 0.5 * L * (0.5 * bb - 0.5 * a + 0.25 * a2 - 0.25 * b2);
        e->R[1] = Fl[1] * 0.5 * L * (0.5 * bb - 0.5 * a - a2 * a2 / 16.0 + b2 * b2 / 16.0 + 3.0 * a2 / 8.0 - 3.0 * b2 / 8.0);
        e->R[2] = Fl[1] * 0.5 * L * (L * (-a2 * a2 + b2 * b2) / 32.0 - L * (-(a * a * a) + (bb * bb * bb)) / 24.0 -
                                     L * (-a2 + b2) / 16.0 + L * (bb - a) / 8.0);
        e->R[3] = Fl[0] * 0.5 * L * (-0.25 * a2 + 0.25 * b2 + 0.5 * bb - 0.5 * a);
        e->R[4] = Fl[1] * 0.5 * L * (0.5 * bb - 0.5 * a + a2 * a2 / 16.0 - b2 * b2 / 16.0 - 3.0 * a2 / 8.0 + 3.0 * b2 / 8.0);
        e->R[5] = Fl[1] * 0.5 * L * (L * (-a2 * a2 + b2 * b2) / 32.0 + L * (-(a * a * a) + (bb * bb * bb)) / 24.0 -
                                     L * (-a2 + b2) / 16.0 - L * (bb - a) / 8.0);
        double RG[6];
        tvec6(e->T, e->R, RG);
        for (int k = 0; k < 6; k++) F->R[e->dofs[k]] += RG[k];
    }
}

/* FEMElementClass::forceVector, FEMElement.cpp:72-97. */
static void el_force(fem_t *F, elem_t *e) {
    double u = (e->L * e->L - e->L0 * e->L0) / (e->L + e->L0);
    double th1 = shift_angle(F->node[e->n0].angle - e->angle);
    double th2 = shift_angle(F->node[e->n1].angle - e->angle);
    double F0 = (e->E * e->A / e->L0) * u;
    double M1 = (2 * e->E * e->I / e->L0) * (2.0 * th1 + th2);
    double M2 = (2 * e->E * e->I / e->L0) * (th1 + 2.0 * th2);
    e->F[0] = -F0;
    e->F[1] = (1.0 / e->L0) * (M1 + M2);
    e->F[2] = M1;
    e->F[3] = F0;
    e->F[4] = -(1.0 / e->L0) * (M1 + M2);
    e->F[5] = M2;
    double FG[6];
    tvec6(e->T, e->F, FG);
    for (int k = 0; k < 6; k++) F->F[e->dofs[k]] += FG[k];
}

static void assemble6(fem_t *F, const elem_t *e, double G[6][6], double *glob) {
    int nd = F->ndof;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) glob[(size_t)e->dofs[i] * nd + e->dofs[j]] += G[i][j];
}

/* FEMElementClass::massMatrix + stiffMatrix, FEMElement.cpp:100-138. */
static void el_mass_stiff(fem_t *F, elem_t *e) {
    double MG[6][6], KG[6][6], KN[6][6];
    mat6_tmul(e->T, e->M, MG);
    assemble6(F, e, MG, F->M);
    mat6_tmul(e->T, e->KL, KG);
    double F0 = -e->F[0], V0 = e->F[4], L0 = e->L0;
    e->KNL[0][1] = -V0 / L0;
    e->KNL[0][4] = V0 / L0;
    e->KNL[1][0] = -V0 / L0;
    e->KNL[1][1] = F0 / L0;
    e->KNL[1][3] = V0 / L0;
    e->KNL[1][4] = -F0 / L0;
    e->KNL[3][1] = V0 / L0;
    e->KNL[3][4] = -V0 / L0;
    e->KNL[4][0] = V0 / L0;
    e->KNL[4][1] = -F0 / L0;
    e->KNL[4][3] = -V0 / L0;
    e->KNL[4][4] = F0 / L0;
    mat6_tmul(e->T, e->KNL, KN);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) KG[i][j] = KG[i][j] + KN[i][j];
    assemble6(F, e, KG, F->K);
}

/* FEMBodyClass::buildGlobalMatrices, FEMBody.cpp:90-109. */
static void fem_build(fem_t *F) {
    size_t nd = F->ndof;
    memset(F->M, 0, sizeof(double) * nd * nd);
    memset(F->K, 0, sizeof(double) * nd * nd);
    memset(F->F, 0, sizeof(double) * nd);
    for (int el = 0; el < F->nel; el++) {
        el_force(F, &F->el[el]);
        el_mass_stiff(F, &F->el[el]);
    }
}

/* FEMBodyClass::finishNewmark, FEMBody.cpp:129-144. */
static void fem_finish_newmark(orc_case *c, fem_t *F) {
    double Dt = c->Dt, al = c->p.alpha, de = c->p.delta;
    double a6 = 1.0 / (al * (Dt * Dt));
    double a7 = -1.0 / (al * Dt);
    double a8 = -(1.0 / (2.0 * al) - 1.0);
    double a9 = Dt * (1.0 - de);
    double a10 = de * Dt;
    for (int k = 0; k < F->ndof; k++) {
        F->Udd[k] = a6 * (F->U[k] - F->Un[k]) + a7 * F->Udn[k] + a8 * F->Uddn[k];
        F->Ud[k] = F->Udn[k] + a9 * F->Uddn[k] + a10 * F->Udd[k];
    }
}

/* FEMBodyClass::newtonRaphsonDynamic + setNewmark, FEMBody.cpp:71-126. */
static void fem_newton_dynamic(orc_case *c, fem_t *F) {
    int nd = F->ndof;
    fem_build(F);
    double Dt = c->Dt, al = c->p.alpha;
    double a0 = 1.0 / (al * (Dt * Dt));
    double a2 = 1.0 / (al * Dt);
    double a3 = 1.0 / (2.0 * al) - 1.0;
    double *vec = xcalloc(nd, sizeof(double));
    for (int k = 0; k < nd; k++) vec[k] = a0 * (F->Un[k] - F->U[k]) + a2 * F->Ud[k] + a3 * F->Udd[k];
    for (int i = 0; i < nd; i++) {
        double s = 0.0;
        for (int j = 0; j < nd; j++) s += F->M[(size_t)i * nd + j] * vec[j];
        F->F[i] = F->R[i] - F->F[i] + s;
    }
    for (size_t k = 0; k < (size_t)nd * nd; k++) F->K[k] = F->K[k] + a0 * F->M[k];
    orc_lu_solve(nd, F->nbc, F->K, F->F, F->delU);
    for (int k = 0; k < nd; k++) F->U[k] = F->U[k] + F->delU[k];
    fem_update_fem(F);
    free(vec);
}

/* FEMBodyClass::subResidual, FEMBody.cpp:244-256: R_k = U - U_km1 (R_km1
 * holds the previous subiteration's R_k), then |R_k|^2 and the numerator
 * R_km1.(R_k - R_km1) and denominator |R_k - R_km1|^2 of the Aitken factor. */
void orc_sub_residual(int nd, const double *U, const double *Ukm1, double *Rk, const double *Rkm1, double out[3]) {
    for (int k = 0; k < nd; k++) Rk[k] = U[k] - Ukm1[k];
    out[0] = vdot(Rk, Rk, nd);
    double num = 0.0, den = 0.0;
    for (int k = 0; k < nd; k++) num += Rkm1[k] * (Rk[k] - Rkm1[k]);
    for (int k = 0; k < nd; k++) den += (Rk[k] - Rkm1[k]) * (Rk[k] - Rkm1[k]);
    out[1] = num;
    out[2] = den;
}

/* FEMBodyClass::dynamicFEM, FEMBody.cpp:26-68. */
static void fem_dynamic(orc_case *c, fem_t *F) {
    int nd = F->ndof;
    memcpy(F->U, F->Un, sizeof(double) * nd);
    memcpy(F->Ud, F->Udn, sizeof(double) * nd);
    memcpy(F->Udd, F->Uddn, sizeof(double) * nd);
    fem_update_fem(F);
    memset(F->R, 0, sizeof(double) * nd);
    for (int el = 0; el < F->nel; el++) el_load(c, F, &F->el[el]);
    F->itNR = 0;
    do {
        fem_newton_dynamic(c, F);
        F->resNR = sqrt(vdot(F->delU, F->delU, nd)) / (c->p.ref_L * sqrt((double)nd));
        F->itNR++;
    } while (F->resNR > 1e-10 && F->itNR < 20);
    fem_finish_newmark(c, F);
    fem_update_ibm(c, F);
    double *t = F->Rkm1; F->Rkm1 = F->Rk; F->Rk = t;  /* R_km1 = R_k (old R_km1 content discarded) */
    double out[3];
    orc_sub_residual(nd, F->U, F->Ukm1, F->Rk, F->Rkm1, out);
    F->subRes = out[0];
    F->subNum = out[1];
    F->subDen = out[2];
}

/* FEMBodyClass::staticFEM + newtonRaphsonStatic, FEMBody.cpp:448-484:
 * Newton on K dU = R - F (stiffness only), then the markers follow. */
static void fem_static(orc_case *c, fem_t *F) {
    int nd = F->ndof;
    double *rhs = xcalloc(nd, sizeof(double));
    F->itNR = 0;
    do {
        fem_build(F);
        for (int k = 0; k < nd; k++) rhs[k] = F->R[k] - F->F[k];
        orc_lu_solve(nd, F->nbc, F->K, rhs, F->delU);
        for (int k = 0; k < nd; k++) F->U[k] = F->U[k] + F->delU[k];
        fem_update_fem(F);
        F->resNR = sqrt(vdot(F->delU, F->delU, nd)) / (c->p.ref_L * sqrt((double)nd));
        F->itNR++;
    } while (F->resNR > 1e-10 && F->itNR < 20);
    fem_update_ibm(c, F);
    free(rhs);
}

/* 2x2 matrix times vector with the reference's accumulation from 0.0
 * (Utils.h:375-389). */
static void mat2_vec(const double T[2][2], const double v[2], double out[2]) {
    for (int i = 0; i < 2; i++) {
        double r = 0.0;
        for (int j = 0; j < 2; j++) r += T[i][j] * v[j];
        out[i] = r;
    }
}

/* FEMBodyClass::setInitialDeflection, FEMBody.cpp:352-445: a tip load in
 * the body frame, corrected by a bounded proportional feedback until the
 * static tip deflection is -INITIAL_DEFLECT * L0. */
static int fem_initial_deflection(orc_case *c, fem_t *F, double fraction) {
    int nd = F->ndof;
    double deflect = -fraction;
    double E = F->el[0].E, I = F->el[0].I;
    double linearForce = 3.0 * E * I * deflect / (F->L0 * F->L0);
    int nSteps = 100;
    double dFmax = linearForce / (double)nSteps;
    double T[2][2] = {{cos(F->angle0), -sin(F->angle0)}, {sin(F->angle0), cos(F->angle0)}};
    double Tt[2][2] = {{T[0][0], T[1][0]}, {T[0][1], T[1][1]}};
    double local[2] = {0.0, dFmax}, global[2];
    mat2_vec(T, local, global);
    F->R[nd - 3] = global[0];
    F->R[nd - 2] = global[1];
    double gain = 1e-3, TOL = 1e-10, error;
    int MAXIT = 10000, it = 0;
    do {
        fem_static(c, F);
        double dv[2] = {F->U[nd - 3], F->U[nd - 2]}, dl[2];
        mat2_vec(Tt, dv, dl);
        error = deflect - dl[1] / F->L0;
        double fc = error * gain;
        if (fabs(fc) > fabs(dFmax)) fc = (double)((0.0 < fc) - (fc < 0.0)) * fabs(dFmax);
        local[1] += fc;
        mat2_vec(T, local, global);
        F->R[nd - 3] = global[0];
        F->R[nd - 2] = global[1];
        it++;
        if (it >= MAXIT) { set_err("max iterations hit! Try changing the gain...exiting"); return -1; }
    } while (fabs(error) > TOL);
    F->itNR = 0;
    F->resNR = 0.0;
    memset(F->R, 0, sizeof(double) * nd);
    memset(F->delU, 0, sizeof(double) * nd);
    return 0;
}

/* FEMBodyClass::predictor, FEMBody.cpp:259-289. */
static void fem_predictor(orc_case *c, fem_t *F) {
    int nd = F->ndof;
    if (c->t > 2)
        for (int k = 0; k < nd; k++) F->U[k] = 2.5 * F->Un[k] - 2.0 * F->Unm1[k] + 0.5 * F->Unm2[k];
    else if (c->t == 2)
        for (int k = 0; k < nd; k++) F->U[k] = 2.0 * F->Un[k] - F->Unm1[k];
    else if (c->t == 1)
        memcpy(F->U, F->Un, sizeof(double) * nd);
    fem_update_fem(F);
    fem_finish_newmark(c, F);
    fem_update_ibm(c, F);
    swapp(&F->Ukm1, &F->U);
}

/* FEMBodyClass::resetValues, FEMBody.cpp:341-349. */
static void fem_reset(fem_t *F) {
    swapp(&F->Unm2, &F->Unm1);
    swapp(&F->Unm1, &F->Un);
    swapp(&F->Un, &F->U);
    swapp(&F->Udn, &F->Ud);
    swapp(&F->Uddn, &F->Udd);
}

/* ------------------------------------------------------------------------- */
/* Bodies, geometry, epsilon, coupling (Objects.cpp, IBMBody.cpp)             */
/* ------------------------------------------------------------------------- */

static marker_t *add_marker(orc_case *c, int body, double x, double y) {
    c->m = realloc(c->m, sizeof(marker_t) * (c->nm + 1));
    marker_t *m = &c->m[c->nm];
    memset(m, 0, sizeof *m);
    m->body = body;
    m->ID = c->nm;
    m->pos[0] = x;
    m->pos[1] = y;
    body_t *b = &c->b[body];
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64;
        b->nodes = realloc(b->nodes, sizeof(int) * b->cap);
    }
    b->nodes[b->n++] = c->nm;
    c->nm++;
    return m;
}

static body_t *add_body(orc_case *c, int ID, int flex, int kind) {
    c->b = realloc(c->b, sizeof(body_t) * (c->nb + 1));
    body_t *b = &c->b[c->nb++];
    memset(b, 0, sizeof *b);
    b->ID = ID;
    b->flex = flex;
    b->kind = kind;
    return b;
}

/* IBMBodyClass circle constructor, IBMBody.cpp:45-76. */
static void build_circle(orc_case *c, int ID, const double pos[2], double radius) {
    add_body(c, ID, 1, 0);
    int bi = c->nb - 1;
    int n = (int)floor(2.0 * M_PI * radius / c->Dx);
    for (int i = 0; i < n; i++)
        add_marker(c, bi, pos[0] + radius * cos(i * 2.0 * M_PI / n), pos[1] + radius * sin(i * 2.0 * M_PI / n));
}

/* IBMBodyClass filament constructor, IBMBody.cpp:79-135. */
static int build_filament(orc_case *c, int ID, const double pos[2], const double geom[2], double angleDeg,
                          const char *flex, const char *nEl, const char *bc, double rho, double E) {
    int fl;
    if (strcmp(flex, "FLEXIBLE") == 0) fl = 0;
    else if (strcmp(flex, "RIGID") == 0) fl = 1;
    else { set_err("Not a proper flexible type set"); return -1; }
    add_body(c, ID, fl, 1);
    int bi = c->nb - 1;
    double L = geom[0];
    double ang = angleDeg * M_PI / 180.0;
    double ca = cos(ang), sa = sin(ang);
    int n = (int)floor(L / c->Dx) + 1;
    for (int i = 0; i < n; i++) {
        double px = i * L / (n - 1), py = 0.0;
        double rx = 0.0 + ca * px + (-sa) * py, ry = 0.0 + sa * px + ca * py;
        add_marker(c, bi, pos[0] + rx, pos[1] + ry);
    }
    if (fl == 0) {
        c->b[bi].fem = fem_create(c, &c->b[bi], pos, geom, ang, nEl, bc, rho, E);
        c->b[bi].fem->body = bi;
    }
    return 0;
}

/* ObjectsClass::geometryReadIn, Objects.cpp:324-488 (text instead of a file). */
static int parse_geometry(orc_case *c, const char *text) {
    if (!text) return 0;
    const char *s = text;
    char line[1024];
    while (*s) {
        size_t k = 0;
        while (*s && *s != '\n' && k + 1 < sizeof line) line[k++] = *s++;
        line[k] = 0;
        if (*s == '\n') s++;
        char *p = line;
        while (*p == ' ' || *p == '\t' || *p == '\r') p++;
        if (*p == '#' || *p == 0) continue;
        char kw[64];
        int number, ID;
        double st[2], sp[2];
        int used = 0;
        if (sscanf(p, "%63s %d %d %lf %lf %lf %lf%n", kw, &number, &ID, &st[0], &st[1], &sp[0], &sp[1], &used) != 7) {
            set_err("malformed geometry line");
            return -1;
        }
        p += used;
        if (strcmp(kw, "CIRCLE") == 0) {
            double r;
            if (sscanf(p, "%lf", &r) != 1) { set_err("malformed CIRCLE line"); return -1; }
            for (int i = 0; i < number; i++) {
                double pos[2] = {st[0] + i * sp[0], st[1] + i * sp[1]};
                build_circle(c, ID, pos, r);
                ID++;
            }
        } else if (strcmp(kw, "FILAMENT") == 0) {
            double g[2], ang, rho, E;
            char flex[32], nel[32], bc[32];
            if (sscanf(p, "%lf %lf %lf %31s %31s %31s %lf %lf", &g[0], &g[1], &ang, flex, nel, bc, &rho, &E) != 8) {
                set_err("malformed FILAMENT line");
                return -1;
            }
            for (int i = 0; i < number; i++) {
                double pos[2] = {st[0] + i * sp[0], st[1] + i * sp[1]};
                if (build_filament(c, ID, pos, g, ang, flex, nel, bc, rho, E)) return -1;
                ID++;
            }
        } else {
            set_err("unknown body keyword");
            return -1;
        }
    }
    for (int i = 0; i < c->nb; i++)
        for (int j = i + 1; j < c->nb; j++)
            if (c->b[i].ID == c->b[j].ID) { set_err("Duplicate body IDs in geometry.config"); return -1; }
    return 0;
}

/* ObjectsClass::removeOverlapMarkers, Objects.cpp:867-938. */
static int remove_overlap(orc_case *c) {
    int nm = c->nm;
    char *erase = xcalloc(nm, 1);
    int *order = xcalloc(nm + 1, sizeof(int));
    int ne = 0;
    for (int a = 0; a < nm; a++)
        for (int b = a + 1; b < nm; b++) {
            double d[2] = {c->m[a].pos[0] - c->m[b].pos[0], c->m[a].pos[1] - c->m[b].pos[1]};
            double mag = sqrt(vdot(d, d, 2)) / c->Dx;
            if (mag < 0.5) {
                int fa = c->b[c->m[a].body].flex, fb = c->b[c->m[b].body].flex;
                if (fa == 0 && fb == 0) { set_err("Cannot place two flexible bodies so close to each other"); free(erase); free(order); return -1; }
                int victim = (fa == 0 && fb == 1) ? b : a;
                if (!erase[victim]) { erase[victim] = 1; ne++; }
            }
        }
    if (ne) {
        int *newidx = xcalloc(nm, sizeof(int));
        int k = 0;
        for (int a = 0; a < nm; a++) {
            newidx[a] = erase[a] ? -1 : k;
            if (!erase[a]) c->m[k++] = c->m[a];
        }
        c->nm = k;
        for (int bi = 0; bi < c->nb; bi++) {
            body_t *b = &c->b[bi];
            int q = 0;
            for (int t = 0; t < b->n; t++)
                if (newidx[b->nodes[t]] >= 0) b->nodes[q++] = newidx[b->nodes[t]];
            b->n = q;
        }
        free(newidx);
    }
    for (int a = 0; a < c->nm; a++) c->m[a].ID = a;
    free(erase);
    free(order);
    return 0;
}

/* The matrix of ObjectsClass::computeEpsilon, Objects.cpp:258-302, for the
 * markers nodes[0..dim): A_ij = ds_j * sum over the supports s of marker i
 * of delta_i(s) * delta(x_j - s). */
static void eps_matrix(const orc_case *c, const int *nodes, int dim, double *A) {
    for (int i = 0; i < dim; i++) {
        const marker_t *mi = &c->m[nodes[i]];
        for (int j = 0; j < dim; j++) {
            const marker_t *mj = &c->m[nodes[j]];
            double a = 0.0;
            for (int s = 0; s < mi->nsupp; s++) {
                double dx = fabs(mj->pos[0] / c->Dx - mi->supp[s].i);
                double dy = fabs(mj->pos[1] / c->Dx - mi->supp[s].j);
                a += mi->supp[s].d * (orc_dirac_delta(dx) * orc_dirac_delta(dy));
            }
            a *= 1.0 * 1.0 * mj->ds;
            A[(size_t)i * dim + j] = a;
        }
    }
}

/* ObjectsClass::computeEpsilon, Objects.cpp:235-321. */
static int compute_epsilon(orc_case *c) {
    int nsys = c->p.uni_epsilon ? 1 : c->nb;
    for (int sys = 0; sys < nsys; sys++) {
        int dim;
        const int *nodes;
        int *all = NULL;
        if (c->p.uni_epsilon) {
            dim = c->nm;
            all = xcalloc(dim, sizeof(int));
            for (int k = 0; k < dim; k++) all[k] = k;
            nodes = all;
        } else {
            if (!(c->t == 0 || c->b[sys].flex == 0)) continue;
            dim = c->b[sys].n;
            nodes = c->b[sys].nodes;
        }
        double *A = xcalloc((size_t)dim * dim, sizeof(double));
        double *rhs = xcalloc(dim, sizeof(double));
        double *x = xcalloc(dim, sizeof(double));
        eps_matrix(c, nodes, dim, A);
        for (int i = 0; i < dim; i++) rhs[i] = 1.0;
        orc_lu_solve(dim, 0, A, rhs, x);
        for (int i = 0; i < dim; i++) c->m[nodes[i]].eps = x[i];
        free(A);
        free(rhs);
        free(x);
        free(all);
    }
    return 0;
}

/* The relaxation factor of ObjectsClass::recomputeObjectVals,
 * Objects.cpp:179-190: at the first relaxed subiteration the previous
 * factor's sign with its magnitude capped at relaxMax, afterwards the Aitken
 * update -relax * num / den (num, den from orc_sub_residual). */
double orc_aitken_relax(int subIt, double relax, double relaxMax, double num, double den) {
    if (subIt == 1) {
        double r = fabs(relax) < relaxMax ? fabs(relax) : relaxMax;
        return (double)(sgn(relax) * r);
    }
    return -relax * num / den;
}

/* ObjectsClass::recomputeObjectVals, Objects.cpp:152-232. */
static int recompute_object_vals(orc_case *c) {
    if (c->subIt == 0) {
        for (int b = 0; b < c->nb; b++)
            if (c->b[b].flex == 0) {
                fem_reset(c->b[b].fem);
                fem_predictor(c, c->b[b].fem);
            }
    } else {
        c->relax = orc_aitken_relax(c->subIt, c->relax, c->p.relaxMax, c->subNum, c->subDen);
        for (int b = 0; b < c->nb; b++)
            if (c->b[b].flex == 0) {
                fem_t *F = c->b[b].fem;
                for (int k = 0; k < F->ndof; k++) F->U[k] = F->Ukm1[k] + c->relax * (F->U[k] - F->Ukm1[k]);
                fem_update_fem(F);
                fem_finish_newmark(c, F);
                fem_update_ibm(c, F);
                swapp(&F->Ukm1, &F->U);
            }
    }
    for (int n = 0; n < c->nm; n++)
        if (c->b[c->m[n].body].flex == 0) {
            if (marker_support(c, &c->m[n])) return -1;
            marker_ds(c, &c->m[n]);
        }
    return compute_epsilon(c);
}

/* ObjectsClass::objectKernel, Objects.cpp:26-60. */
static int object_kernel(orc_case *c) {
    c->subIt = 0;
    const int MAXIT = 20;
    do {
        if (c->hasFlex && recompute_object_vals(c)) return -1;
        /* ibmKernelInterp, Objects.cpp:102-117 */
        memset(c->fibm, 0, sizeof(double) * c->N * 2);
        for (int n = 0; n < c->nm; n++) marker_interp_force(c, &c->m[n]);
        if (!c->hasFlex) break;
        /* femKernel, Objects.cpp:63-99 (ORDERED) */
        double res = 0.0, num = 0.0, den = 0.0;
        for (int b = 0; b < c->nb; b++)
            if (c->b[b].flex == 0) {
                fem_dynamic(c, c->b[b].fem);
                res += c->b[b].fem->subRes;
                num += c->b[b].fem->subNum;
                den += c->b[b].fem->subDen;
            }
        c->subRes = sqrt(res) / (c->p.ref_L * sqrt((double)c->simDOFs));
        c->subNum = num;
        c->subDen = den;
        c->subIt++;
    } while (c->subIt < MAXIT && c->subRes > c->p.subTol);
    /* ibmKernelSpread, Objects.cpp:120-149 */
    memset(c->fibm, 0, sizeof(double) * c->N * 2);
    for (int n = 0; n < c->nm; n++) marker_spread(c, &c->m[n]);
    for (int n = 0; n < c->nm; n++) marker_update_macro(c, &c->m[n]);
    if (c->subIt == MAXIT) { set_err("Subiteration scheme hit 20 iterations"); return -1; }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Public API                                                                 */
/* ------------------------------------------------------------------------- */

orc_case *orc_create(const orc_params *p, const char *geometry) {
    init_tables();
    orc_case *c = xcalloc(1, sizeof(orc_case));
    c->p = *p;
    if (grid_init(c)) { orc_destroy(c); return NULL; }
    c->relax = p->relaxMax;
    if (parse_geometry(c, geometry) || remove_overlap(c)) { orc_destroy(c); return NULL; }
    c->hasIBM = c->nb > 0;
    for (int b = 0; b < c->nb; b++) if (c->b[b].flex == 0) c->hasFlex = 1;
    if (c->hasIBM) {
        /* ObjectsClass::initialiseObjects, Objects.cpp:941-961 */
        /* ObjectsClass::initialDeflect, Objects.cpp:964-976 */
        if (p->initial_deflect != 0.0 && c->hasFlex)
            for (int b = 0; b < c->nb; b++)
                if (c->b[b].flex == 0 && fem_initial_deflection(c, c->b[b].fem, p->initial_deflect)) {
                    orc_destroy(c);
                    return NULL;
                }
        for (int n = 0; n < c->nm; n++) {
            if (marker_support(c, &c->m[n])) { orc_destroy(c); return NULL; }
            marker_ds(c, &c->m[n]);
        }
        compute_epsilon(c);
    }
    c->simDOFs = 0;
    for (int b = 0; b < c->nb; b++) if (c->b[b].flex == 0) c->simDOFs += c->b[b].fem->ndof;
    return c;
}

static void fem_free(fem_t *F) {
    if (!F) return;
    for (int e = 0; e < F->nel; e++) free(F->el[e].map);
    free(F->el); free(F->node); free(F->pmap);
    double *v[] = {F->M, F->K, F->R, F->F, F->U, F->delU, F->Ud, F->Udd, F->Un, F->Unm1, F->Unm2,
                   F->Udn, F->Uddn, F->Ukm1, F->Rk, F->Rkm1};
    for (size_t k = 0; k < sizeof v / sizeof v[0]; k++) free(v[k]);
    free(F);
}

void orc_destroy(orc_case *c) {
    if (!c) return;
    double *v[] = {c->f, c->f_n, c->u, c->u_n, c->rho, c->rho_n, c->fxy, c->fibm, c->delU, c->u_in, c->rho_in};
    for (size_t k = 0; k < sizeof v / sizeof v[0]; k++) free(v[k]);
    free(c->type); free(c->bc);
    for (int b = 0; b < c->nb; b++) { free(c->b[b].nodes); fem_free(c->b[b].fem); }
    free(c->b); free(c->m);
    free(c);
}

void orc_scalars(const orc_case *c, double out[10]) {
    out[0] = c->omega; out[1] = c->tau; out[2] = c->nu; out[3] = c->Dx; out[4] = c->Dt;
    out[5] = c->Dm; out[6] = c->Drho; out[7] = c->cs; out[8] = c->cs2; out[9] = c->cs4;
}
int orc_time(const orc_case *c) { return c->t; }

int orc_step(orc_case *c, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        c->t++;
        if (orc_lbm_kernel(c)) return -1;
        if (c->hasIBM && object_kernel(c)) return -1;
    }
    return 0;
}

void orc_get_grid(const orc_case *c, double *f, double *rho, double *u, double *fibm) {
    if (f) memcpy(f, c->f, sizeof(double) * c->N * NV);
    if (rho) memcpy(rho, c->rho, sizeof(double) * c->N);
    if (u) memcpy(u, c->u, sizeof(double) * c->N * 2);
    if (fibm) memcpy(fibm, c->fibm, sizeof(double) * c->N * 2);
}

void orc_set_grid(orc_case *c, const double *f, const double *rho, const double *u, const double *fibm) {
    if (f) memcpy(c->f, f, sizeof(double) * c->N * NV);
    if (rho) memcpy(c->rho, rho, sizeof(double) * c->N);
    if (u) memcpy(c->u, u, sizeof(double) * c->N * 2);
    if (fibm) memcpy(c->fibm, fibm, sizeof(double) * c->N * 2);
}

void orc_get_types(const orc_case *c, int *type) { memcpy(type, c->type, sizeof(int) * c->N); }

int orc_num_markers(const orc_case *c) { return c->nm; }
int orc_num_bodies(const orc_case *c) { return c->nb; }

void orc_get_markers(const orc_case *c, double *pos, double *vel, double *force, double *eps, double *ds, int *body) {
    for (int n = 0; n < c->nm; n++) {
        const marker_t *m = &c->m[n];
        if (pos) { pos[2 * n] = m->pos[0]; pos[2 * n + 1] = m->pos[1]; }
        if (vel) { vel[2 * n] = m->vel[0]; vel[2 * n + 1] = m->vel[1]; }
        if (force) { force[2 * n] = m->force[0]; force[2 * n + 1] = m->force[1]; }
        if (eps) eps[n] = m->eps;
        if (ds) ds[n] = m->ds;
        if (body) body[n] = m->body;
    }
}

void orc_get_supports(const orc_case *c, int *count, int *ij, double *delta) {
    for (int n = 0; n < c->nm; n++) {
        const marker_t *m = &c->m[n];
        count[n] = m->nsupp;
        for (int s = 0; s < SUPP; s++) {
            ij[(n * SUPP + s) * 2] = s < m->nsupp ? m->supp[s].i : -1;
            ij[(n * SUPP + s) * 2 + 1] = s < m->nsupp ? m->supp[s].j : -1;
            delta[n * SUPP + s] = s < m->nsupp ? m->supp[s].d : 0.0;
        }
    }
}

void orc_body_info(const orc_case *c, int b, int *info) {
    const body_t *B = &c->b[b];
    info[0] = B->ID;
    info[1] = B->flex;
    info[2] = B->n;
    info[3] = B->n ? B->nodes[0] : -1;
    info[4] = B->fem ? B->fem->ndof : 0;
}

/* Body b's shape (0 circle, 1 filament) and its markers' global indices in
   the body's own node order (iBody[ib].node, IBMBody.h). */
int orc_body_nodes(const orc_case *c, int b, int *idx) {
    const body_t *B = &c->b[b];
    if (idx) memcpy(idx, B->nodes, sizeof(int) * B->n);
    return B->kind;
}

int orc_get_fem(const orc_case *c, int b, double *U, double *Udot, double *Udotdot) {
    const fem_t *F = c->b[b].fem;
    if (!F) return -1;
    if (U) memcpy(U, F->U, sizeof(double) * F->ndof);
    if (Udot) memcpy(Udot, F->Ud, sizeof(double) * F->ndof);
    if (Udotdot) memcpy(Udotdot, F->Udd, sizeof(double) * F->ndof);
    return F->ndof;
}

/* The rest of a flexible body's restart state (FEM.restart as
   ObjectsClass::writeRestart lays it out, Objects.cpp:1290-1352): U_n,
   U_nm1, R_k (ndof each), (L, angle) per element and (pos0, pos, angle0,
   angle) per node.  sizes: nel, nnode. */
int orc_get_fem_state(const orc_case *c, int b, int *sizes, double *Un, double *Unm1, double *Rk, double *el,
                      double *nodes) {
    const fem_t *F = c->b[b].fem;
    if (!F) return -1;
    sizes[0] = F->nel;
    sizes[1] = F->nnode;
    if (Un) memcpy(Un, F->Un, sizeof(double) * F->ndof);
    if (Unm1) memcpy(Unm1, F->Unm1, sizeof(double) * F->ndof);
    if (Rk) memcpy(Rk, F->Rk, sizeof(double) * F->ndof);
    if (el)
        for (int e = 0; e < F->nel; e++) {
            el[2 * e] = F->el[e].L;
            el[2 * e + 1] = F->el[e].angle;
        }
    if (nodes)
        for (int n = 0; n < F->nnode; n++) {
            double *q = nodes + 6 * n;
            q[0] = F->node[n].pos0[0];
            q[1] = F->node[n].pos0[1];
            q[2] = F->node[n].pos[0];
            q[3] = F->node[n].pos[1];
            q[4] = F->node[n].angle0;
            q[5] = F->node[n].angle;
        }
    return F->ndof;
}

void orc_coupling_info(const orc_case *c, double *out) {
    out[0] = c->subIt;
    out[1] = c->subRes;
    out[2] = c->relax;
}

/* A clamped horizontal filament FEM body on its own (no fluid, no
 * markers), for checking the beam against closed forms: with nsteps == 0,
 * the static tip deflection under a tip load P through the static Newton
 * iteration of FEMBodyClass::staticFEM (FEMBody.cpp:437-484); with
 * nsteps > 0, the tip y displacement after releasing that static state,
 * one Newmark step (FEMBody.cpp:26-144 without the IBM load) per entry. */
int orc_fem_cantilever(double L, double h, double E, double rho, int nel, double Dt, double P, int nsteps,
                       double *tip_out) {
    init_tables();
    orc_case c;
    memset(&c, 0, sizeof c);
    c.Dx = 1.0;
    c.Dt = Dt;
    c.Dm = 1.0;
    c.p.alpha = 0.25;
    c.p.delta = 0.5;
    c.p.ref_L = L;
    body_t b;
    memset(&b, 0, sizeof b);
    b.n = nel + 1;
    char nels[32];
    snprintf(nels, sizeof nels, "%d", nel);
    double pos[2] = {0.0, 0.0}, geom[2] = {L, h};
    fem_t *F = fem_create(&c, &b, pos, geom, 0.0, nels, "CLAMPED", rho, E);
    int nd = F->ndof;
    /* static Newton under the tip load */
    F->R[nd - 3 + 1] = P;
    double *rhs = xcalloc(nd, sizeof(double));
    int it = 0;
    double res;
    do {
        fem_build(F);
        for (int k = 0; k < nd; k++) rhs[k] = F->R[k] - F->F[k];
        orc_lu_solve(nd, F->nbc, F->K, rhs, F->delU);
        for (int k = 0; k < nd; k++) F->U[k] = F->U[k] + F->delU[k];
        fem_update_fem(F);
        res = sqrt(vdot(F->delU, F->delU, nd)) / (L * sqrt((double)nd));
        it++;
    } while (res > 1e-12 && it < 100);
    if (nsteps == 0) {
        tip_out[0] = F->U[nd - 3 + 1];
        free(rhs);
        fem_free(F);
        return it;
    }
    /* release: R = 0, consistent initial acceleration M a = -F_int */
    memset(F->R, 0, sizeof(double) * nd);
    fem_build(F);
    for (int k = 0; k < nd; k++) rhs[k] = -F->F[k];
    orc_lu_solve(nd, F->nbc, F->M, rhs, F->Uddn);
    memcpy(F->Un, F->U, sizeof(double) * nd);
    memset(F->Udn, 0, sizeof(double) * nd);
    for (int s = 0; s < nsteps; s++) {
        memcpy(F->U, F->Un, sizeof(double) * nd);
        memcpy(F->Ud, F->Udn, sizeof(double) * nd);
        memcpy(F->Udd, F->Uddn, sizeof(double) * nd);
        fem_update_fem(F);
        int k = 0;
        do {
            fem_newton_dynamic(&c, F);
            res = sqrt(vdot(F->delU, F->delU, nd)) / (L * sqrt((double)nd));
            k++;
        } while (res > 1e-12 && k < 20);
        fem_finish_newmark(&c, F);
        tip_out[s] = F->U[nd - 3 + 1];
        memcpy(F->Un, F->U, sizeof(double) * nd);
        memcpy(F->Udn, F->Ud, sizeof(double) * nd);
        memcpy(F->Uddn, F->Udd, sizeof(double) * nd);
    }
    free(rhs);
    fem_free(F);
    return 0;
}

/* ObjectsClass::writeTotalForces, Objects.cpp:517-534. */
void orc_total_force(const orc_case *c, double out[4]) {
    double scale = c->Dm * c->Dx / (c->Dt * c->Dt) * 1.0 / c->Dx;
    double F[2] = {0.0, 0.0};
    for (int n = 0; n < c->nm; n++)
        for (int d = 0; d < 2; d++) F[d] -= c->m[n].force[d] * 1.0 * c->m[n].eps * c->m[n].ds * scale;
    double ND = 0.5 * c->p.ref_rho * (c->p.ref_U * c->p.ref_U) * c->p.ref_L;
    out[0] = F[0];
    out[1] = F[1];
    out[2] = F[0] / ND;
    out[3] = F[1] / ND;
}

/* ------------------------------------------------------------------------- */
/* Unit-level access to the restated steps, so that each can be checked on   */
/* its own against what the mathematics fixes.                               */
/* ------------------------------------------------------------------------- */

/* ObjectsClass::ibmKernelInterp, Objects.cpp:102-117, on the stored rho/u. */
void orc_ibm_interpolate(orc_case *c) {
    memset(c->fibm, 0, sizeof(double) * c->N * 2);
    for (int n = 0; n < c->nm; n++) marker_interp_force(c, &c->m[n]);
}

/* ObjectsClass::ibmKernelSpread, Objects.cpp:120-149. */
void orc_ibm_spread(orc_case *c) {
    memset(c->fibm, 0, sizeof(double) * c->N * 2);
    for (int n = 0; n < c->nm; n++) marker_spread(c, &c->m[n]);
    for (int n = 0; n < c->nm; n++) marker_update_macro(c, &c->m[n]);
}

/* Overwrite marker values (NULL: keep). */
void orc_set_markers(orc_case *c, const double *pos, const double *vel, const double *force, const double *eps,
                     const double *ds) {
    for (int n = 0; n < c->nm; n++) {
        marker_t *m = &c->m[n];
        if (pos) { m->pos[0] = pos[2 * n]; m->pos[1] = pos[2 * n + 1]; }
        if (vel) { m->vel[0] = vel[2 * n]; m->vel[1] = vel[2 * n + 1]; }
        if (force) { m->force[0] = force[2 * n]; m->force[1] = force[2 * n + 1]; }
        if (eps) m->eps = eps[n];
        if (ds) m->ds = ds[n];
    }
}

/* findSupport + computeDs for every marker (Objects.cpp:949-955). */
int orc_ibm_supports(orc_case *c) {
    for (int n = 0; n < c->nm; n++) {
        if (marker_support(c, &c->m[n])) return -1;
        marker_ds(c, &c->m[n]);
    }
    return 0;
}

/* The epsilon system `sys` (UNI_EPSILON: the one system of all markers;
 * else body sys): its matrix into A (dim x dim, row-major, may be NULL).
 * Returns dim. */
int orc_epsilon_matrix(const orc_case *c, int sys, double *A) {
    if (c->p.uni_epsilon) {
        int dim = c->nm;
        if (A) {
            int *all = xcalloc(dim, sizeof(int));
            for (int k = 0; k < dim; k++) all[k] = k;
            eps_matrix(c, all, dim, A);
            free(all);
        }
        return dim;
    }
    if (sys < 0 || sys >= c->nb) return -1;
    if (A) eps_matrix(c, c->b[sys].nodes, c->b[sys].n, A);
    return c->b[sys].n;
}

/* Interpolated density and momentum of every marker from the last
 * interpolation (IBMNode.h:56-57). */
void orc_get_interp(const orc_case *c, double *rho, double *mom) {
    for (int n = 0; n < c->nm; n++) {
        rho[n] = c->m[n].interpRho;
        mom[2 * n] = c->m[n].interpMom[0];
        mom[2 * n + 1] = c->m[n].interpMom[1];
    }
}

/* GridClass::getNormalVector at one site: n[2] and the normal direction. */
int orc_normal_vector(const orc_case *c, int i, int j, int *n, int *nd) { return normal_vector(c, i, j, n, nd); }

/* Convective outlet increments of the last lattice step (Ny x 2). */
void orc_get_delu(const orc_case *c, double *delU) { memcpy(delU, c->delU, sizeof(double) * c->Ny * 2); }

/* Inlet velocity profile in lattice units (Ny x 2). */
void orc_get_inlet(const orc_case *c, double *u_in) { memcpy(u_in, c->u_in, sizeof(double) * c->Ny * 2); }

/* FEMBodyClass::constructRVector, FEMBody.cpp:226-233: the load vector of
 * flexible body b from its markers' present forces and epsilon. */
int orc_fem_load(orc_case *c, int b, double *R) {
    fem_t *F = c->b[b].fem;
    if (!F) return -1;
    memset(F->R, 0, sizeof(double) * F->ndof);
    for (int el = 0; el < F->nel; el++) el_load(c, F, &F->el[el]);
    memcpy(R, F->R, sizeof(double) * F->ndof);
    return F->ndof;
}
