// Internal definitions shared by the libsparcgpu.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <map>
#include <utility>
#include <string>
#include <vector>

#include "sparc_gpu.h"

#define SGPU_MAX_FDN 8

// Thread-local last error message (sgpu_last_error).
void sgpu_set_error(const char *fmt, ...);

#define SGPU_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            sgpu_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,               \
                           hipGetErrorString(e_));                                  \
            return SGPU_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

#define SGPU_BLAS(call)                                                             \
    do {                                                                            \
        rocblas_status s_ = (call);                                                 \
        if (s_ != rocblas_status_success) {                                         \
            sgpu_set_error("%s:%d %s: rocblas status %d", __FILE__, __LINE__, #call, \
                           (int)s_);                                                \
            return SGPU_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

#define SGPU_CHECK(expr)                 \
    do {                                 \
        int st_ = (expr);                \
        if (st_ < 0) return st_;         \
    } while (0)

struct sgpu_block {
    sgpu_context *ctx;
    int ncol;
    int dtype;   // SGPU_F64 / SGPU_F32
    long ld;     // elements between columns (>= Nd, padded to 64)
    void *ptr;   // device storage, ncol * ld elements
};

// Nonlocal projectors re-laid out in bricks of kBX x kBY x kBZ grid points.
// For every brick touched by at least one projector sphere, Phi_b is a dense
// kBP x np_b row-major block (np_b = projector columns of all influence
// entries touching the brick, padded to a multiple of 4 with zero columns);
// brick row j maps to projector (alpha) row row_alpha[brick_row[b] + j]
// (-1 for padding).  Periodic images of one atom are separate entries whose
// columns map to the same alpha rows.
constexpr int kBX = 16, kBY = 8, kBZ = 4, kBP = kBX * kBY * kBZ;

struct sgpu_nonlocal {
    int n_infl = 0;
    int n_alpha = 0;
    int n_bricks = 0;
    long n_rows = 0;            // sum of padded brick rows
    int max_np = 0;
    int *brick_org = nullptr;   // x0, y0, z0 per brick
    int *brick_np = nullptr;    // padded rows per brick
    long *brick_row = nullptr;  // first brick row of each brick
    int *row_alpha = nullptr;   // alpha row of each brick row (-1: padding)
    double *phi64 = nullptr;    // Phi_b row-major (kBP x np_b) at brick_row[b] * kBP
    float *phi32 = nullptr;
    double *phic64 = nullptr;   // the same blocks column-major (np_b x kBP)
    float *phic32 = nullptr;
    int *arow_ptr = nullptr;    // alpha row r sums brick rows arow_src[arow_ptr[r]..]
    int *arow_src = nullptr;
    double *gamma = nullptr;    // Gamma per alpha row
    double dV = 0.0;
    // sphere layout (small projector spheres): per influence entry its point
    // list and Chi, entries ordered by colour (no two entries of a colour
    // share a grid point), images of one atom found through atom_img
    int mode = 0;               // 0 brick layout, 1 sphere layout
    long stored = 0;            // projector values per copy, zero padding included
    int n_ent = 0, n_atoms = 0, n_colors = 0, max_qpad = 0, max_npad = 0;
    std::vector<int> color_start;   // host: first entry of each colour (+ end)
    int *ent_info = nullptr;    // per entry: qpad, pos offset, np, npad, atom, chi offset (6 ints)
    long *ent_chi = nullptr;    // per entry: element offset of its Chi blocks
    int *gpos = nullptr;        // grid index per padded point (-1: padding)
    double *chir64 = nullptr;   // Chi [q][npad] per entry
    float *chir32 = nullptr;
    double *chit64 = nullptr;   // Chi [p][qpad] per entry
    float *chit32 = nullptr;
    int *atom_img_ptr = nullptr;  // atom a: entries atom_img[ptr[a]..ptr[a+1])
    int *atom_img = nullptr;
    int *atom_row = nullptr;      // alpha row of the atom's first projector
    // partials carried from one application of a chain to the next: two
    // sphere partial buffers (chain_cur: the current one)
    void *chain = nullptr;
    size_t chain_bytes = 0;
    int chain_cur = 0;
    const void *chain_x = nullptr;  // the block whose partials are held
};

struct sgpu_timer_acc {
    double total_ms = 0.0;
    int count = 0;
};

struct sgpu_context {
    int device = 0;
    int nx = 0, ny = 0, nz = 0;
    long nd = 0;
    long ld = 0;  // padded column stride used by device blocks
    double hx = 0, hy = 0, hz = 0;
    int bcx = 0, bcy = 0, bcz = 0;
    int fdn = 0;
    double d2x[SGPU_MAX_FDN + 1] = {0}, d2y[SGPU_MAX_FDN + 1] = {0}, d2z[SGPU_MAX_FDN + 1] = {0};
    // radius-6 tile shape (TX x TY, RY rows per thread): 0 = 32x16x2,
    // 1 = 32x32x4, 2 = 64x16x4, 3 = 32x32x8, 4 = 64x32x8, 5 / 6 = 32x16x2 /
    // 32x8x2 with an 8-plane LDS ring (FP32; FP64 uses 0), 7 = 32x32x2,
    // 8 = 64x16x2 (512 threads), 9 / 10 = 32x16x2 with z-chunks of at least
    // 8 / 4 planes (default 16), 11 / 12 = 96x16x2 / 80x16x2 (whole x-rows
    // on grids of 81-96 / 65-80 points), 13 / 14 = the same shapes as
    // stencil_rows when nx is exactly 96 / 80 (FP32); -1 (default) = the
    // fastest of 0 / 7 / 8 (9 / 10 below 16 columns, 11-14 where they fit),
    // timed once per (precision, column count) and kept in tile_choice.  SPARC_AMD_STENCIL_TILE fixes one at
    // creation.
    int stencil_cfg = -1;
    std::map<std::pair<int, int>, int> tile_choice;
    // fuse each filter step's nonlocal projection into the previous step's
    // apply (SPARC_AMD_NL_CHAIN=0 turns it off; results are bitwise equal)
    bool nl_chain = true;
    // the brick layout's fused apply + next projection (nlb_apply_rs with the
    // projection; SPARC_AMD_NL_CHAIN_BRICK=1 selects it): bitwise equal, but
    // slower on the 96^3 x 512 operator case (2.23 ms per application against
    // 1.19 + 0.74 ms for nlb_apply_rs and nlb_project_rs), so off by default
    bool nl_chain_brick = false;
    // eigensolves of up to 180 states reduce to tridiagonal form in one
    // workgroup's LDS instead of rocsolver's dsytrd (SPARC_AMD_LDS_TRIDIAG=0
    // keeps dsyevd everywhere)
    bool lds_tridiag = true;
    // larger eigensolves reduce to tridiagonal form in one cooperative launch
    // of this many workgroups, each holding a share of the rows in LDS
    // (SPARC_AMD_COOP_TRIDIAG=<groups>; 0 keeps dsyevd)
    int coop_tridiag = 32;
    int coop_threads = 256;
    // the eigensolve's overlap factor by the blocked Cholesky kernels below,
    // with C = L^-1 Hp L^-T and Q = L^-T Z as products with L^-1
    // (SPARC_AMD_CHOL=rocsolver keeps dpotrf and the triangular solves)
    bool own_chol = true;
    // the back-transformation Q1 Z by ormtr_lower_wave (SPARC_AMD_ORMTR=rocsolver
    // keeps rocSOLVER's dormtr)
    bool own_ormtr = true;
    // the tridiagonal eigenvectors by sgpu_tdc's divide and conquer
    // (SPARC_AMD_STEDC=rocsolver keeps rocSOLVER's dstedc)
    bool own_stedc = true;
    // its lowest levels (merges of up to this many poles) in one launch, one
    // wave per leaf block (SPARC_AMD_TDC_LEAF=1|8|16|32; 1: none)
    int tdc_leaf = 8;
    // Mp and Hp of a CheFSI projection in one BF16-piece launch
    // (SPARC_AMD_GRAM3_PAIR=0: two)
    bool gram3_pair = true;
    int n_cu = 256;   // compute units of the device (set by sgpu_create)
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;

    double *veff64 = nullptr;
    float *veff32 = nullptr;
    bool has_veff = false;

    sgpu_nonlocal nl;

    // scratch (grown on demand, owned by the context)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    double *red_buf = nullptr;      // reduction partials
    size_t red_elems = 0;
    double *host_red = nullptr;     // pinned host mirror of small results
    double *aar_mem = nullptr;      // AAR work vectors (r, x_old, f, f_old, X, F)
    size_t aar_elems = 0;

    bool timing = false;
    std::map<std::string, sgpu_timer_acc> timers;
    struct pending_pair {
        hipEvent_t a, b;
        std::string name;
    };
    std::vector<pending_pair> pending;
    std::vector<hipEvent_t> ev_pool;
    // live blocks; sgpu_destroy releases their storage and orphans them
    std::vector<sgpu_block *> blocks;
    // live U-Nets; sgpu_destroy releases their device memory and orphans them
    std::vector<sgpu_unet *> unets;
};


// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (linear
// workgroup id modulo 8), and each XCD has its own L2.  xcd_remap turns the
// linear id into a logical id such that consecutive logical ids run on the
// same XCD, so blocks that share data (the column blocks of one brick, the
// tiles of one column) share an L2.  Bijective for any grid size.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    const int xcd = bid % kNumXcd, slot = bid / kNumXcd;
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + slot;
}

// ---- helpers implemented in sgpu_core.hip ----
int sgpu_scratch(sgpu_context *ctx, size_t bytes, void **p);
int sgpu_reduce_buf(sgpu_context *ctx, size_t elems, double **p);
size_t sgpu_dtype_size(int dtype);

// RAII-less event timer: begin returns a token, end accumulates.
struct sgpu_tmark {
    hipEvent_t a = nullptr, b = nullptr;
};
int sgpu_tbegin(sgpu_context *ctx, sgpu_tmark *m);
int sgpu_tend(sgpu_context *ctx, sgpu_tmark *m, const char *name);
int sgpu_timer_resolve(sgpu_context *ctx);

// ---- tridiagonal divide and conquer (tdc.hip) ----
// Largest n sgpu_tdc takes, and its workspace in doubles for n.
constexpr int kTdcMaxN = 1536;
size_t sgpu_tdc_work_doubles(int n);
// D (n) <- eigenvalues ascending, E (n - 1) the off-diagonal (read only),
// Z (n x n, leading dimension n) <- the eigenvectors, on ctx->stream.
int sgpu_tdc(sgpu_context *ctx, int n, double *dD, const double *dE, double *dZ, double *work);

// ---- kernel launchers (stencil.hip) ----
// out = s1 * ((a Lap + bdiag diag(v) + c) x) + s2 * z, per column (z may be null).
// may_tune = false never times tile shapes (the timing allocates a temporary
// output): a shape already chosen for (dtype, ncol) is used, else 32x16x2.
int sgpu_launch_stencil(sgpu_context *ctx, int dtype, int ncol, double a, double bdiag, double c,
                        const void *v, const void *x, long ldx, void *out, long ldo,
                        const void *z, long ldz, double s1, double s2, bool allow_tiled,
                        bool may_tune = true);
// ---- nonlocal.hip ----
void sgpu_free_nonlocal(sgpu_nonlocal &nl);
// Scratch bytes sgpu_launch_vnl takes for ncol columns of dtype.
size_t sgpu_vnl_scratch_bytes(const sgpu_context *ctx, int dtype, int ncol);
// out += scale * Vnl x  (alpha computed from x); dtype of x/out.  The
// partials go to ws (ws_bytes >= sgpu_vnl_scratch_bytes) when ws is given,
// else to the context's scratch.
int sgpu_launch_vnl(sgpu_context *ctx, int dtype, int ncol, const void *x, long ldx, void *out,
                    long ldo, double scale, void *ws = nullptr, size_t ws_bytes = 0);
// The same for one application of a chain in which each out is the next x
// (the Chebyshev recurrence).  have_part: the partials of x were left by the
// previous call (its *projected was true, and x is that call's out,
// unchanged since); want_next: also form the partials of out after the
// apply, fused into it, where the layout allows (the sphere layout with one
// colour; *projected tells).
int sgpu_launch_vnl_chain(sgpu_context *ctx, int dtype, int ncol, const void *x, long ldx,
                          void *out, long ldo, double scale, bool have_part, bool want_next,
                          bool *projected);
// ---- noalloc.hip ----
// Look up (once per device, with that device current) the address of the
// file-scope scratch array that sgpu_hamiltonian_dev_noalloc uses.
int sgpu_noalloc_setup(int device);
// ---- unet.hip ----
// Release a network's device memory and detach it from its context (called
// by sgpu_destroy; the handle then only accepts sgpu_unet_destroy).
void sgpu_unet_orphan(sgpu_unet *u);
// ---- aar.hip ----
void anderson_coeff(int m, const double *FtF, const double *Ftf, double *G);
int sgpu_anderson_coeff_dev(sgpu_context *ctx, const double *F, long ldf, int m, const double *f,
                            long n, double *G);
int sgpu_anderson_update_dev(sgpu_context *ctx, double *x, const double *xk, const double *f,
                             const double *Xh, const double *Fh, long ldh, int m, const double *G,
                             double beta, long n);
// AAR on device vectors (x in/out, b), returns SGPU_OK / SGPU_NOT_CONVERGED / error.
// Optional observer of the AAR iterates: fn(user, x, r) after every residual
// evaluation (r = b + (Lap + c) x), on the context's stream.
struct sgpu_aar_recorder {
    void *user = nullptr;
    void (*fn)(void *user, const double *x_dev, const double *r_dev) = nullptr;
};
int sgpu_aar_device(sgpu_context *ctx, sgpu_res_fn res, sgpu_precond_fn precond, void *user,
                    double c, int N, double *x, const double *b, double omega, double beta, int m,
                    int p, double tol, int max_iter, int *iters, double *relres,
                    const sgpu_aar_recorder *rec = nullptr);
// ---- blas.hip ----
int sgpu_dot_many(sgpu_context *ctx, int n, int npairs, const double *const *a,
                  const double *const *b, double *host_out);
int sgpu_convert(sgpu_context *ctx, const void *src, int st, void *dst, int dt, long n);
