/*
 * sparc_gpu.h -- C ABI of libsparcgpu.so, the MI355X (gfx950) implementation
 * of SPARC's Chebyshev-filtered subspace iteration (CheFSI) and AAR
 * electrostatics hot path.
 *
 * One context owns one orthogonal grid on one GPU (the "one MPI rank owns the
 * whole domain" case of SPARC).  Operator data (Veff, nonlocal projectors) is
 * uploaded once and stays in HBM.  Two families of entry points exist:
 *
 *  - Host-pointer drop-ins with the argument meaning of the reference C entry
 *    points (column-major blocks, leading dimensions ldi/ldo, FP64 host data).
 *    Each one cites the reference function it replaces.  They copy their
 *    inputs to the device, run, and copy results back.
 *  - Device-resident variants (`*_dev`) working on sgpu_block handles, so
 *    that Psi stays in HBM across filter -> projection -> rotation.
 *
 * Every function returns an int status: SGPU_OK (0) on success, a negative
 * SGPU_ERR_* on failure (message via sgpu_last_error()), or SGPU_NOT_CONVERGED
 * (1) where an iterative method stopped at its iteration limit.  The reference
 * functions are void and exit() on failure; the status is this library's
 * addition (see each entry).
 *
 * Precision: SGPU_F64 runs everything in FP64.  SGPU_F32 stores the vector
 * block in FP32 and runs the stencil/Veff/Chebyshev/nonlocal arithmetic in
 * FP32; projections return FP64 matrices (FP32 products, FP64 accumulation
 * across tiles), and the dense eigensolve is always FP64.
 */
#ifndef SPARC_GPU_H
#define SPARC_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SGPU_OK              0
#define SGPU_NOT_CONVERGED   1
#define SGPU_ERR_ARG        -1
#define SGPU_ERR_HIP        -2
#define SGPU_ERR_NOMEM      -3
#define SGPU_ERR_STATE      -4
#define SGPU_ERR_SOLVER     -5

#define SGPU_F64 0
#define SGPU_F32 1

typedef struct sgpu_context sgpu_context;
typedef struct sgpu_block sgpu_block;

/* ---- context -------------------------------------------------------- */

/* Create a context for an nx*ny*nz orthogonal grid with mesh (hx,hy,hz),
 * boundary conditions bc* (0 periodic, 1 Dirichlet; pSPARC->BCx convention,
 * src/lapVecRoutines.c:330-332) and FD order (even, <= 16).  Computes the
 * second-derivative weights of src/initialization.c:2113-2147. */
int sgpu_create(sgpu_context **ctx, int device, int nx, int ny, int nz,
                double hx, double hy, double hz, int bcx, int bcy, int bcz, int order);
/* Destroying a context releases the storage of its remaining blocks and
 * U-Nets; their handles stay valid only for sgpu_block_destroy /
 * sgpu_unet_destroy (any other call on them returns SGPU_ERR_STATE).  Any
 * teardown order is safe. */
int sgpu_destroy(sgpu_context *ctx);
const char *sgpu_last_error(void);
/* Scaled D2 stencil coefficients D2_stencil_coeffs_{x,y,z}[0..order/2]. */
int sgpu_get_stencil_coeffs(const sgpu_context *ctx, double *d2x, double *d2y, double *d2z);
/* Synchronise the context's stream. */
int sgpu_sync(sgpu_context *ctx);
/* Diagnostic, no counterpart in the reference: the x-y tile (tx x ty points,
 * ry rows per thread) the radius-6 stencil uses for blocks of ncol columns of
 * precision dtype; 0 with *tx = 0 if no launch of that shape has chosen one. */
int sgpu_stencil_tile(const sgpu_context *ctx, int dtype, int ncol, int *tx, int *ty, int *ry);
/* Diagnostic, no counterpart in the reference: the name of that stencil
 * kernel ("stencil_tiled<float,32,16,2>", "stencil_rows<96,16,2>", ...) into
 * name[len]; an empty string if no launch of that shape has chosen one. */
int sgpu_stencil_kernel(const sgpu_context *ctx, int dtype, int ncol, char *name, int len);

/* Veff_loc_dmcomm (one spin channel, Nd values).  Replaces the field read by
 * Hamiltonian_vectors_mult (src/hamiltonianVecRoutines.c:45-49). */
int sgpu_set_veff(sgpu_context *ctx, const double *veff);

/* Nonlocal projectors, flattened from ATOM_NLOC_INFLUENCE_OBJ / NLOC_PROJ_OBJ
 * (src/include/isddft.h:202-216,247-265): influence entry a has ndc[a] grid
 * points grid_pos[pos_off[a]...], Chi as a column-major ndc[a] x nproj[a]
 * block at chi[chi_off[a]], and projector rows starting at alpha_off[a]
 * (IP_displ of its parent atom; images of one atom share rows).  gamma[r] is
 * Gamma_{l,p} of projector row r (n_alpha rows).  dV is the volume element. */
int sgpu_set_nonlocal(sgpu_context *ctx, int n_influence, const int *ndc, const int *nproj,
                      const long *pos_off, const int *grid_pos, const long *chi_off,
                      const double *chi, const int *alpha_off, int n_alpha,
                      const double *gamma, double dV);

/* Diagnostic, no counterpart in the reference: the projector layout that
 * sgpu_set_nonlocal chose (0 bricks of 16 x 8 x 4 points, 1 per-atom spheres)
 * and the projector values it stores per copy, zero padding included; one
 * projection or one apply issues 2 x stored multiply-adds per column. */
int sgpu_nonlocal_layout(const sgpu_context *ctx, int *mode, long *stored);

/* ---- device blocks -------------------------------------------------- */

/* A block of ncol grid vectors in HBM, column-major, dtype SGPU_F64/F32. */
int sgpu_block_create(sgpu_context *ctx, int ncol, int dtype, sgpu_block **blk);
int sgpu_block_destroy(sgpu_block *blk);
int sgpu_block_ncol(const sgpu_block *blk);
int sgpu_block_dtype(const sgpu_block *blk);
long sgpu_block_ld(const sgpu_block *blk);
/* Copy ncol host columns (leading dimension ld) in / out, converting type. */
int sgpu_block_upload(sgpu_block *blk, const double *host, long ld, int ncol);
int sgpu_block_download(const sgpu_block *blk, double *host, long ld, int ncol);
/* Fill with uniform [lo,hi) values from a counter-based hash of (seed, index). */
int sgpu_block_fill_uniform(sgpu_block *blk, unsigned long long seed, double lo, double hi);
/* Copy one block into another of the same shape (dtype may differ). */
int sgpu_block_copy(const sgpu_block *src, sgpu_block *dst);

/* ---- reference entry points: host-pointer drop-ins ------------------- */

/* Lapx = (Lap + c I) x.  Replaces Lap_vec_mult (src/lapVecRoutines.c:37-58). */
int sgpu_Lap_vec_mult(sgpu_context *ctx, int ncol, double c, const double *x, int ldi,
                      double *Lapx, int ldo);
/* y = (a Lap + b diag(v) + c I) x.  Replaces Lap_plus_diag_vec_mult_orth
 * (src/lapVecRoutines.c:306-611); v == NULL or |b| < 1e-14 drops the diagonal. */
int sgpu_Lap_plus_diag_vec_mult(sgpu_context *ctx, int ncol, double a, double b, double c,
                                const double *v, const double *x, int ldi, double *y, int ldo);
/* Hx += Vnl x.  Replaces Vnl_vec_mult (src/nlocVecRoutines.c:798-883). */
int sgpu_Vnl_vec_mult(sgpu_context *ctx, int ncol, const double *x, int ldi, double *Hx, int ldo);
/* Hx = (-1/2 Lap + Veff + c) x + Vnl x.  Replaces Hamiltonian_vectors_mult
 * (src/hamiltonianVecRoutines.c:45-121); Hx is overwritten. */
int sgpu_Hamiltonian_vectors_mult(sgpu_context *ctx, int ncol, double c, const double *x, int ldi,
                                  double *Hx, int ldo);
/* Chebyshev filter of degree m with cutoff a, upper bound b, lower bound a0.
 * Replaces ChebyshevFiltering (src/eigenSolver.c:722-798): Y receives the
 * filtered block and X is clobbered (ends as Y_{m-1}).  dtype selects the
 * arithmetic (SGPU_F64 or SGPU_F32). */
int sgpu_ChebyshevFiltering(sgpu_context *ctx, double *X, int ldi, double *Y, int ldo, int ncol,
                            int m, double a, double b, double a0, int dtype);
/* HY = H Y, Mp = Y'Y, Hp = Y'HY (Ns x Ns, column-major).  Replaces
 * DP_Project_Hamiltonian (src/eigenSolver.c:939-1082) on one rank. */
int sgpu_DP_Project_Hamiltonian(sgpu_context *ctx, const double *Y, int ldi, double *HY, int ldo,
                                int ncol, double *Hp, double *Mp, int dtype);
/* Solve Hp Q = Mp Q Lambda (itype 1, upper triangles read).  Replaces the
 * LAPACKE_dsygvd call of DP_Solve_Generalized_EigenProblem
 * (src/eigenSolver.c:1262-1375); Hp is overwritten by Q, lambda ascending.
 * Runs rocSOLVER dsygvd in FP64 on the device.  Returns SGPU_ERR_SOLVER with
 * the LAPACK info in the message when the factorisation fails (the reference
 * ignores info). */
int sgpu_DP_Solve_Generalized_EigenProblem(sgpu_context *ctx, int n, double *Hp, double *Mp,
                                           double *lambda);
/* Psi_rot = Y Q.  Replaces DP_Subspace_Rotation (src/eigenSolver.c:1386-1439). */
int sgpu_DP_Subspace_Rotation(sgpu_context *ctx, const double *Y, int ldy, const double *Q,
                              int ns, double *Psi_rot, int ldo, int dtype);

/* Residual / preconditioner callbacks of AAR with the argument meaning of
 * the reference's res_fun / precond_fun (src/linearSolver.c:37-43), acting on
 * FP64 device vectors of length N on the context's stream. */
typedef void (*sgpu_res_fn)(void *user, int N, double c, const double *x_dev,
                            const double *b_dev, double *r_dev);
typedef void (*sgpu_precond_fn)(void *user, int N, double c, const double *r_dev,
                                double *f_dev);
/* Alternating Anderson-Richardson.  Replaces AAR (src/linearSolver.c:37-159).
 * res == NULL selects the built-in poisson_residual (src/lapVecRoutines.c:69)
 * and precond == NULL the built-in Jacobi_preconditioner
 * (src/electrostatics.c:1682).  x and b are host arrays of length N = Nd.
 * iters / relres (may be NULL) receive the iteration count and ||r||/||b||
 * at the last norm evaluation.  Returns SGPU_NOT_CONVERGED if max_iter was
 * reached (the reference returns silently). */
int sgpu_AAR(sgpu_context *ctx, sgpu_res_fn res, sgpu_precond_fn precond, void *user, double c,
             int N, double *x, const double *b, double omega, double beta, int m, int p,
             double tol, int max_iter, int *iters, double *relres);

/* ---- device-resident variants ---------------------------------------- */

/* y = (a Lap + b diag(Veff) + c) x on blocks (diagonal from sgpu_set_veff). */
int sgpu_lap_plus_diag_dev(sgpu_context *ctx, double a, double b, double c,
                           const sgpu_block *x, sgpu_block *y);
/* Hx = (H + c) x, Hx overwritten. */
int sgpu_hamiltonian_dev(sgpu_context *ctx, double c, const sgpu_block *x, sgpu_block *Hx);
/* Chebyshev filter on blocks; same contract as sgpu_ChebyshevFiltering
 * (Y result, X = Y_{m-1}); the two handles may exchange their storage. */
int sgpu_chebyshev_filter_dev(sgpu_context *ctx, sgpu_block *X, sgpu_block *Y, int m,
                              double a, double b, double a0);
/* HY = H Y and Mp = Y'Y, Hp = Y'HY into host arrays (FP64). */
int sgpu_project_dev(sgpu_context *ctx, const sgpu_block *Y, sgpu_block *HY, double *Hp,
                     double *Mp);
/* out = Y Q with host Q (ns x ns FP64). */
int sgpu_rotate_dev(sgpu_context *ctx, const sgpu_block *Y, const double *Q, sgpu_block *out);
/* Lanczos bounds of H (src/eigenSolver.c:1920-2129) from host start vector
 * x0 (FP64); returns the iteration count in *iters. */
int sgpu_lanczos(sgpu_context *ctx, const double *x0, double tol_min, double tol_max, int maxit,
                 double *eigmin, double *eigmax, int *iters);
/* One CheFSI pass (src/eigenSolver.c:269-445) on resident blocks: filter X
 * into Y with bounds (cutoff a, eigmax b, eigmin a0), project, solve the
 * FP64 generalized problem, rotate into X.  lambda receives Ns eigenvalues. */
int sgpu_chefsi_dev(sgpu_context *ctx, sgpu_block *X, sgpu_block *Y, int m, double a, double b,
                    double a0, double *lambda);

/* Bytes of the device scratch array sgpu_hamiltonian_dev_noalloc uses (one
 * copy per GPU, held by the library until the process ends). */
#define SGPU_NOALLOC_SCRATCH_BYTES (64u << 20)

/* sgpu_hamiltonian_dev that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * stream capture, tight serving loops).  Same result as sgpu_hamiltonian_dev,
 * bit for bit.  The nonlocal partials live in a file-scope device array of
 * SGPU_NOALLOC_SCRATCH_BYTES per GPU instead of the context's scratch; they
 * take (projector rows of the layout) x ncol x element size bytes, and a call
 * that needs more is refused with SGPU_ERR_ARG (apply fewer columns per
 * call).  The stencil uses the tile shape already chosen for this
 * (precision, column count) by an earlier allocating call, else 32 x 16 x 2
 * (every shape gives the same bits); event timers are not recorded.  All
 * calls on one GPU share its array: calls that overlap on one GPU race. */
int sgpu_hamiltonian_dev_noalloc(sgpu_context *ctx, double c, const sgpu_block *x,
                                 sgpu_block *Hx);

/* ---- self-consistent field loop ------------------------------------- */

/* Exchange-correlation functionals (EXCHANGE_CORRELATION: LDA_PZ / LDA_PW) */
#define SGPU_XC_LDA_PZ 0
#define SGPU_XC_LDA_PW 1

/* Parameters of the density-mixing SCF (defaults of SPARC in brackets,
 * src/initialization.c:677-989,2233-2608). */
typedef struct {
    int nstates;            /* Nstates */
    int cheb_degree;        /* ChebDegree (Mesh2ChebDegree(h)) */
    int rho_trigger;        /* CheFSI passes in the first SCF [4] */
    int max_scf;            /* MAXIT_SCF [100] */
    int min_scf;            /* MINIT_SCF [2] */
    double tol_scf;         /* TOL_SCF */
    double tol_poisson;     /* TOL_POISSON (relative to ||rhs||) */
    int maxit_poisson;      /* MAXIT_POISSON [3000] */
    double tol_lanczos;     /* TOL_LANCZOS [1e-2] */
    double beta;            /* 1/(k_B T) [1/Ha] */
    int smear_type;         /* elec_T_type: 0 Fermi-Dirac, 1 Gaussian [1] */
    double nelec;           /* electrons (-NegCharge) */
    double occfac;          /* 2 for spin-unpolarized */
    int mix_history;        /* MixingHistory [7], <= 8 */
    int pulay_freq;         /* PulayFrequency [1] */
    double mix_beta;        /* MixingParameter [0.3] */
    double mix_omega;       /* MixingParameterSimple [= MixingParameter] */
    double kerker_ktf;      /* precond_kerker_kTF [1.0] */
    double kerker_thresh;   /* precond_kerker_thresh [0.1] */
    double tol_precond;     /* TOL_PRECOND [h^2 1e-3] */
    double Esc;             /* self + correction energy of the pseudocharges */
    int xc;                 /* SGPU_XC_LDA_PZ or SGPU_XC_LDA_PW */
    int dtype;              /* SGPU_F64 or SGPU_F32 vector blocks (filter, Gram, rotation) */
    unsigned long long psi_seed;      /* initial orbitals uniform [-0.5, 0.5) */
    unsigned long long lanczos_seed;  /* Lanczos start vector uniform [0, 1) */
    /* precond_fun of the Poisson AAR solves; NULL selects the reference's
     * Jacobi_preconditioner.  sgpu_unet_precond with a trained network
     * (poisson_precond_user = the sgpu_unet) gives the learned solve. */
    sgpu_precond_fn poisson_precond;
    void *poisson_precond_user;
    /* precond_fun of the Kerker AAR solves (Kerker_precond,
     * src/mixing.c:462-527, shift c = -kTF^2); NULL selects Jacobi.  A solve
     * with it that does not converge is redone with Jacobi-AAR. */
    sgpu_precond_fn kerker_precond;
    void *kerker_precond_user;
    /* optional outputs for the forces after the SCF: the final orbitals
     * (Xorb; a block of this context with nstates columns and the SCF's
     * dtype) and Vxc of the final density (XCPotential, Nd host doubles);
     * NULL skips either */
    sgpu_block *psi_out;
    double *vxc_out;
} sgpu_scf_params;

typedef struct {
    int scf_iters;          /* SCF iterations done */
    int converged;
    double Etot, Eband, Exc, E1, E2, E3, Entropy, Efermi;  /* last iteration (src/energy.c:165) */
    double scf_error;       /* last |rho_out - rho_in| / |rho_out| */
    double mean_iter_ms;    /* mean wall time of iterations 2..N */
    double total_ms;
    long poisson_iters, kerker_iters;  /* AAR iterations summed over the run */
    int poisson_fallbacks;  /* Poisson solves redone with Jacobi-AAR because the solve with
                               poisson_precond did not converge */
    int kerker_fallbacks;   /* the same for the Kerker solves and kerker_precond */
} sgpu_scf_result;

/* The integrals of the nonlocal forces for every orbital of psi:
 * out[0] = dV <Chi, psi> (Compute_Integral_psi_Chi, src/forces.c:849-888) and
 * out[1..3] = <Chi, d psi / dx>, <Chi, d psi / dy>, <Chi, d psi / dz>
 * (Compute_Integral_Chi_Dpsi, :890-924, with the order-2n first-derivative
 * stencil of Gradient_vectors_dir), periodic images summed into their atom's
 * rows; out is [4][n_alpha][ncol] (row-major) host doubles.  The forces
 * follow as -2 occfac sum_n g_n sum_p Gamma_p out[0] out[1+d]
 * (Compute_force_nloc_by_integrals, :926-993). */
int sgpu_nonlocal_integrals(sgpu_context *ctx, const sgpu_block *psi, double *out);

/* Radial tables of one pseudopotential on its uniform grid r[0..size): r V_loc
 * with its spline derivatives, the NLCC core density / 4 pi with its spline
 * derivatives (NULL without NLCC), the valence charge and V_loc(0)
 * (PSD_OBJ fields RadialGrid, rVloc, SplinerVlocD, rho_c_table, SplineRhocD,
 * Znucl, Vloc_0). */
typedef struct {
    int size;
    const double *r, *rvloc, *rvloc_d, *rhoc, *rhoc_d;
    double zion, vloc0;
} sgpu_psp_radial;

/* Local and NLCC force terms (Calculate_local_forces_linear,
 * src/forces.c:157-457; Calculate_forces_xc_linear, :473-707) for n_img
 * image atoms: img_box[6k..6k+5] = xs, ys, zs, xe, ye, ze of image k's
 * pseudocharge box (grid indices inside the grid), img_centre its position,
 * img_type / img_atom its pseudopotential and atom; phi, Vc, b + b_ref and
 * Vxc on the grid (host, Nd each).  f_loc / f_xc receive [n_atom][3]
 * (not symmetrized).  Periodic orthogonal cells. */
int sgpu_local_forces(sgpu_context *ctx, int n_types, const sgpu_psp_radial *types, int n_img,
                      const int *img_type, const int *img_atom, const int *img_box,
                      const double *img_centre, double ref_cutoff, const double *phi,
                      const double *vc, const double *bsum, const double *vxc, int n_atom,
                      double *f_loc, double *f_xc);

/* Density-mixing SCF with CheFSI: replaces scf / scf_loop
 * (src/electronicGroundState.c:607-1147) for one spin-unpolarized Gamma-point
 * system whose grid, Veff storage and nonlocal projectors live in ctx.  b is
 * the pseudocharge density, rho0 the initial electron density and rho_core
 * the NLCC model core density added to rho for exchange-correlation
 * (add_rho_core, src/exchangeCorrelation.c:1254-1270; may be NULL), Nd each,
 * host.  Outputs: lambda (nstates eigenvalues of the last pass), optional
 * final mixed density / electrostatic potential (Nd each) and per-iteration
 * wall time, SCF error and total energy (max_scf entries each).  Returns
 * SGPU_NOT_CONVERGED if max_scf was reached. */
int sgpu_scf_run(sgpu_context *ctx, const sgpu_scf_params *par, const double *b,
                 const double *rho0, const double *rho_core, double *lambda, double *rho_out,
                 double *phi_out,
                 double *iter_ms, double *iter_err, double *iter_etot, sgpu_scf_result *res);

/* ---- learned AAR preconditioner --------------------------------------- */

/* A linear 3-level 3-D convolutional U-Net (16 channels, BF16 activations,
 * FP32 accumulation and master weights) bound to a context's periodic grid
 * (at least 8 points per axis; on a grid whose dimensions are not multiples
 * of 4 the network runs on the grid rounded up to multiples of 4, input and
 * output interpolated periodically-trilinearly).  Applied as
 * f = (r + U(r)) / |D0| with
 * D0 = D2x[0] + D2y[0] + D2z[0]: the reference's Jacobi_preconditioner
 * (src/electrostatics.c:1682-1699) is the U = 0 case, which is also the
 * network's initial state.  No counterpart in the reference beyond that. */
typedef struct sgpu_unet sgpu_unet;
int  sgpu_unet_create(sgpu_context *ctx, int channels, unsigned long long seed, sgpu_unet **net);
void sgpu_unet_destroy(sgpu_unet *net);
int  sgpu_unet_nparams(const sgpu_unet *net);
int  sgpu_unet_get_weights(const sgpu_unet *net, float *w);
int  sgpu_unet_set_weights(sgpu_unet *net, const float *w);
/* A precond_fun for sgpu_AAR (user = the sgpu_unet*), the signature of the
 * reference's precond_fun (src/linearSolver.c:37-43): FP64 device vectors. */
void sgpu_unet_precond(void *user, int N, double c, const double *r_dev, double *f_dev);
/* f = P(r) on host vectors of length Nd. */
int  sgpu_unet_apply(sgpu_unet *net, const double *r, double *f);
/* Loss mean_i |P(r_i) - e_i|^2 over npair host pairs (Nd each, stacked) and
 * its gradient with respect to the weights (sgpu_unet_nparams floats). */
int  sgpu_unet_loss_grad(sgpu_unet *net, const double *r, const double *e, int npair,
                         double *loss, float *grad);
/* Training pairs from a Jacobi-AAR solve of the context's Poisson problem
 * (rhs b, start x0, host vectors): every stride-th iterate k gives
 * (r_k / |e_k|, (e_k - mean e_k) / |e_k|) with e_k = x* - x_k. */
int  sgpu_unet_harvest(sgpu_unet *net, const double *b, const double *x0, double tol,
                       int max_iter, int stride, int max_pairs, int *added);
int  sgpu_unet_npairs(const sgpu_unet *net);
/* The shift c (<= 0) of the operator -(Lap + c) that sgpu_unet_harvest solves
 * and the training output scale 1/|D0 + c| use: 0 (the default) for the
 * Poisson solve, -kTF^2 for the Kerker solve (Kerker_precond,
 * src/mixing.c:462-527).  No counterpart in the reference. */
int  sgpu_unet_set_shift(sgpu_unet *net, double c);
/* Adam (lr, 0.9, 0.999) over the stored pairs, batch pairs per step. */
int  sgpu_unet_train(sgpu_unet *net, int steps, int batch, double lr, unsigned long long seed,
                     double *loss_hist);

/* ---- timing ---------------------------------------------------------- */

/* Average device time (ms) of the last timed launches of the named kernel
 * group ("filter_stencil", "nonlocal", "filter", ...), measured with HIP
 * events on the context's stream; *count gets the number of samples. */
int sgpu_timer_get(sgpu_context *ctx, const char *name, double *total_ms, int *count);
int sgpu_timer_reset(sgpu_context *ctx);
/* Enable (1) or disable (0) per-kernel event timing (default off). */
int sgpu_timer_enable(sgpu_context *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
