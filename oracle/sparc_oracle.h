/*
 * FP64 CPU oracle for the CheFSI + AAR-electrostatics path of SPARC.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithms and is used as the checker by tests/, by the smoke()
 * entry point and by the cpu_baseline leg of bench.py.  The product library
 * (sparc_amd/, libsparcgpu.so) never links, loads or calls it.
 *
 * Every function names the reference routine it restates (file:line in
 * SPARC-X/SPARC).  Data layout follows the reference: a block of ncol vectors
 * on an Nx x Ny x Nz grid is column-major with leading dimension ld >= Nd and
 * X(n,i,j,k) = x[n*ld + k*Nx*Ny + j*Nx + i] (src/lapVecRoutines.c:313).
 */
#ifndef SPARC_ORACLE_H
#define SPARC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_FDN 8

/* Orthogonal grid + FD stencil description.  bc* = 0 periodic, 1 Dirichlet
 * (same convention as pSPARC->BCx, src/lapVecRoutines.c:330-332). */
typedef struct {
    int nx, ny, nz;
    double hx, hy, hz;
    int bcx, bcy, bcz;
    int fdn;                      /* order / 2 */
    double d2x[ORC_MAX_FDN + 1];  /* D2_stencil_coeffs_x (weights / hx^2) */
    double d2y[ORC_MAX_FDN + 1];
    double d2z[ORC_MAX_FDN + 1];
} orc_grid;

/* Flattened nonlocal projector set (ATOM_NLOC_INFLUENCE_OBJ + NLOC_PROJ_OBJ,
 * src/include/isddft.h:202-216,247-265).  One entry per influencing (image)
 * atom; entry a owns ndc[a] grid points grid_pos[pos_off[a] ...] and a
 * column-major ndc[a] x nproj[a] block of Chi at chi[chi_off[a]].  Its
 * projector coefficients live at alpha rows alpha_off[a] ... (IP_displ of
 * the parent atom, so periodic images of one atom share rows).  gamma[] holds
 * Gamma_{l,p} expanded per projector row. */
typedef struct {
    int n_influence;
    const int *ndc;
    const int *nproj;
    const long *pos_off;
    const int *grid_pos;
    const long *chi_off;
    const double *chi;
    const int *alpha_off;
    int n_alpha;
    const double *gamma;
    double dV;
} orc_nloc;

int  orc_fd_d2_weights(int fdn, double *w);          /* unscaled FDweights_D2[0..fdn] */
int  orc_grid_init(orc_grid *g, int nx, int ny, int nz, double hx, double hy, double hz,
                   int bcx, int bcy, int bcz, int order);
double orc_max_eig_mhalf_lap(const orc_grid *g);
int  orc_mesh2chebdegree(double h);

/* y = (a*Lap + b*diag(v) + c*I) x, per column (src/lapVecRoutines.c:306-611). */
void orc_lap_plus_diag(const orc_grid *g, int ncol, double a, double b, double c,
                       const double *v, const double *x, int ldi, double *y, int ldo);
/* Lapx = (Lap + c) x  (src/lapVecRoutines.c:37-58). */
void orc_lap_vec_mult(const orc_grid *g, int ncol, double c, const double *x, int ldi,
                      double *y, int ldo);
/* Hx += Vnl x (src/nlocVecRoutines.c:798-883). */
void orc_vnl_vec_mult(const orc_nloc *nl, int ncol, const double *x, int ldi, double *Hx, int ldo);
/* Hx = (-0.5 Lap + Veff + c) x + Vnl x (src/hamiltonianVecRoutines.c:45-121). */
void orc_hamiltonian(const orc_grid *g, const double *veff, const orc_nloc *nl, int ncol,
                     double c, const double *x, int ldi, double *Hx, int ldo);
/* Chebyshev filter (src/eigenSolver.c:722-798); X is clobbered (ends as Y_{m-1}). */
void orc_chebyshev_filter(const orc_grid *g, const double *veff, const orc_nloc *nl,
                          double *X, int ldi, double *Y, int ldo, int ncol, int m,
                          double a, double b, double a0);
/* Lanczos bounds (src/eigenSolver.c:1920-2129).  Returns iterations done. */
int  orc_lanczos(const orc_grid *g, const double *veff, const orc_nloc *nl, const double *x0,
                 double tol_min, double tol_max, int maxit, double *eigmin, double *eigmax);
/* Min/max eigenvalues of the symmetric tridiagonal (d, e) of size n. */
void orc_tridiag_minmax(int n, const double *d, const double *e, double *lmin, double *lmax);
/* Mp = Y'Y, Hp = Y'HY, HY overwritten (src/eigenSolver.c:939-1082). */
void orc_project_hamiltonian(const orc_grid *g, const double *veff, const orc_nloc *nl,
                             const double *Y, int ldi, double *HY, int ldo, int ncol,
                             double *Hp, double *Mp);
/* C = A' B for column-major A (nd x na, lda), B (nd x nb, ldb); C na x nb. */
void orc_gemm_tn(int nd, int na, int nb, const double *A, int lda, const double *B, int ldb,
                 double *C);
/* out = Y Q (src/eigenSolver.c:1405-1415). */
void orc_subspace_rotation(int nd, int ns, const double *Y, int ldy, const double *Q,
                           double *out, int ldo);
/* Generalized symmetric eigenproblem Hp Q = Mp Q L, eigenvalues ascending,
 * Q M-orthonormal (the contract of LAPACKE_dsygvd itype 1, src/eigenSolver.c:1318).
 * Hp, Mp are overwritten.  Returns 0 on success. */
int  orc_sygv(int n, double *Hp, double *Mp, double *lambda, double *Q);
/* Symmetric eigendecomposition by cyclic Jacobi; A overwritten, V = eigenvectors. */
int  orc_syev_jacobi(int n, double *A, double *w, double *V);

/* Anderson coefficient Gamma = pinv(F'F) F'f (src/mixing.c:115-139). */
void orc_anderson_coeff(int N, int m, const double *f, const double *F, double *Gamma);
/* Anderson extrapolation x_kp1 = x_k - X G + beta (f - F G) (src/mixing.c:48-106). */
void orc_anderson_extrapolation(int N, int m, double *x_kp1, const double *x_k, const double *f,
                                const double *X, const double *F, double beta);

typedef void (*orc_res_fn)(void *ctx, int N, double c, const double *x, const double *b, double *r);
typedef void (*orc_precond_fn)(void *ctx, int N, double c, const double *r, double *f);
/* AAR (src/linearSolver.c:37-159).  Returns the iteration count; *relres gets
 * ||r||/||b|| at the last norm evaluation. */
int  orc_aar(void *ctx, orc_res_fn res, orc_precond_fn precond, double c, int N, double *x,
             const double *b, double omega, double beta, int m, int p, double tol, int max_iter,
             double *relres);
/* Poisson residual r = b + (Lap + c) x (src/lapVecRoutines.c:69-79); ctx = orc_grid*. */
void orc_poisson_residual(void *ctx, int N, double c, const double *x, const double *b, double *r);
/* Jacobi f = -r / (D2x0 + D2y0 + D2z0 + c) (src/electrostatics.c:1682-1699); ctx = orc_grid*. */
void orc_jacobi_precond(void *ctx, int N, double c, const double *r, double *f);
/* AAR Poisson solve with the reference's defaults, then mean shift for
 * periodic BC (src/electrostatics.c:1593-1679). */
int  orc_poisson_solve(const orc_grid *g, double *phi, const double *rhs, double tol, int maxit,
                       double *relres);

/* Smearing, Fermi level and occupations (src/occupation.c:31-465). */
double orc_smearing(double beta, double lambda, double ef, int type);
double orc_fermi_level(int ns, const double *lambda, double occfac, double nelec, double beta,
                       int type, double x1, double x2, double tol, int max_iter, int *status);
/* rho_i = sum_n occfac g_n psi_{i,n}^2 / dV (src/electronDensity.c:104-202). */
void orc_density(int nd, int ns, const double *psi, int ld, const double *occ, double occfac,
                 double dV, double *rho);

/* LDA: Slater exchange + Perdew-Zunger correlation (src/exchangeCorrelation.c:459-529). */
void orc_lda_pz(int n, const double *rho, double *exc, double *vxc);
/* LDA: Slater exchange + Perdew-Wang 92 correlation (src/exchangeCorrelation.c:475-499). */
void orc_lda_pw(int n, const double *rho, double *exc, double *vxc);

/* Set the number of OpenMP threads used by the column-parallel loops. */
void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
