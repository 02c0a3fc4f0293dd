/*
 * FP64 CPU oracle for the CheFSI + AAR-electrostatics path of SPARC.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference
 * algorithms, used as the checker by tests/, smoke() and bench.py's
 * cpu_baseline leg.  The product library never links or calls it.
 * Each function cites the reference routine (SPARC-X/SPARC file:line) it
 * follows.  Pinning of each function is listed in oracle/README.md and
 * exercised by tests/test_oracle_*.py.
 */
#include "sparc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

/* fract(n,k) = n!^2 / ((n-k)! (n+k)!) (src/tools.c:725-733). */
static double fract(int n, int k)
{
    double nr = 1.0, dr = 1.0;
    for (int i = n - k + 1; i <= n; i++) nr *= i;
    for (int i = n + 1; i <= n + k; i++) dr *= i;
    return nr / dr;
}

/* FDweights_D2 (src/initialization.c:2118-2122). */
int orc_fd_d2_weights(int fdn, double *w)
{
    if (fdn < 1 || fdn > ORC_MAX_FDN) return -1;
    w[0] = 0.0;
    for (int p = 1; p <= fdn; p++) {
        w[0] -= 2.0 / (p * p);
        w[p] = (2 * (p % 2) - 1) * 2 * fract(fdn, p) / (p * p);
    }
    return 0;
}

/* D2_stencil_coeffs_{x,y,z} for the orthogonal cell (src/initialization.c:2134-2147). */
int orc_grid_init(orc_grid *g, int nx, int ny, int nz, double hx, double hy, double hz,
                  int bcx, int bcy, int bcz, int order)
{
    int fdn = order / 2;
    double w[ORC_MAX_FDN + 1];
    if (order % 2 || orc_fd_d2_weights(fdn, w)) return -1;
    /* the reference's single-rank periodic halo copies fdn planes from the
     * opposite face, which needs at least fdn points per direction */
    if (nx < fdn || ny < fdn || nz < fdn) return -2;
    memset(g, 0, sizeof(*g));
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->hx = hx; g->hy = hy; g->hz = hz;
    g->bcx = bcx; g->bcy = bcy; g->bcz = bcz;
    g->fdn = fdn;
    for (int p = 0; p <= fdn; p++) {
        g->d2x[p] = w[p] / (hx * hx);
        g->d2y[p] = w[p] / (hy * hy);
        g->d2z[p] = w[p] / (hz * hz);
    }
    return 0;
}

/* MaxEigVal_mhalfLap closed form (src/initialization.c:2184-2197). */
double orc_max_eig_mhalf_lap(const orc_grid *g)
{
    double s = g->d2x[0] + g->d2y[0] + g->d2z[0];
    double sx = (g->nx - g->nx % 2) / (double)g->nx;
    double sy = (g->ny - g->ny % 2) / (double)g->ny;
    double sz = (g->nz - g->nz % 2) / (double)g->nz;
    for (int p = 1; p <= g->fdn; p++)
        s += 2.0 * (g->d2x[p] * cos(M_PI * p * sx) + g->d2y[p] * cos(M_PI * p * sy) +
                    g->d2z[p] * cos(M_PI * p * sz));
    return -0.5 * s;
}

/* Mesh2ChebDegree (src/eigenSolver.c:2347-2362). */
int orc_mesh2chebdegree(double h)
{
    double p3 = -700. / 3., p2 = 1240. / 3., p1 = -773. / 3., p0 = 1078. / 15.;
    double npl = h > 0.7 ? 14 : ((p3 * h + p2) * h + p1) * h + p0;
    return (int)round(npl);
}

/* Neighbour value along one axis with the reference's halo semantics:
 * periodic wraps (equivalent to the opposite-face copy of
 * src/lapVecRoutines.c:536-577 when n >= fdn), Dirichlet reads zero. */
static inline int wrap(int i, int n, int bc, int *valid)
{
    if (i >= 0 && i < n) { *valid = 1; return i; }
    if (bc) { *valid = 0; return 0; }
    *valid = 1;
    return i < 0 ? i + n : i - n;
}

/* One column of y = (a Lap + b diag(v) + c) x, following
 * Lap_plus_diag_vec_mult_orth + stencil_3axis_thread_radius6
 * (src/lapVecRoutines.c:185-227,306-611): centre weight
 * w2_diag = a(D2x0+D2y0+D2z0) + c, pairs summed per radius, and
 * b*(v*x) added last.  b==0 or v==NULL drops the diagonal (:321-322). */
static void lap_plus_diag_col(const orc_grid *g, double a, double b, double c, const double *v,
                              const double *x, double *y)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, fdn = g->fdn;
    const long nxny = (long)nx * ny;
    double wx[ORC_MAX_FDN + 1], wy[ORC_MAX_FDN + 1], wz[ORC_MAX_FDN + 1];
    for (int p = 0; p <= fdn; p++) {
        wx[p] = g->d2x[p] * a;
        wy[p] = g->d2y[p] * a;
        wz[p] = g->d2z[p] * a;
    }
    const double w0 = wx[0] + wy[0] + wz[0] + c;
    double bb = b;
    if (fabs(b) < 1e-14 || v == NULL) bb = 0.0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long idx = k * nxny + (long)j * nx + i;
                double res = w0 * x[idx];
                for (int r = 1; r <= fdn; r++) {
                    int v1, v2;
                    int im = wrap(i - r, nx, g->bcx, &v1), ip = wrap(i + r, nx, g->bcx, &v2);
                    double xm = v1 ? x[k * nxny + (long)j * nx + im] : 0.0;
                    double xp = v2 ? x[k * nxny + (long)j * nx + ip] : 0.0;
                    double rx = (xm + xp) * wx[r];
                    int jm = wrap(j - r, ny, g->bcy, &v1), jp = wrap(j + r, ny, g->bcy, &v2);
                    xm = v1 ? x[k * nxny + (long)jm * nx + i] : 0.0;
                    xp = v2 ? x[k * nxny + (long)jp * nx + i] : 0.0;
                    double ry = (xm + xp) * wy[r];
                    int km = wrap(k - r, nz, g->bcz, &v1), kp = wrap(k + r, nz, g->bcz, &v2);
                    xm = v1 ? x[km * nxny + (long)j * nx + i] : 0.0;
                    xp = v2 ? x[kp * nxny + (long)j * nx + i] : 0.0;
                    double rz = (xm + xp) * wz[r];
                    res += rx + ry + rz;
                }
                y[idx] = bb != 0.0 ? res + bb * (v[idx] * x[idx]) : res;
            }
}

void orc_lap_plus_diag(const orc_grid *g, int ncol, double a, double b, double c,
                       const double *v, const double *x, int ldi, double *y, int ldo)
{
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int n = 0; n < ncol; n++)
        lap_plus_diag_col(g, a, b, c, v, x + (long)n * ldi, y + (long)n * ldo);
}

void orc_lap_vec_mult(const orc_grid *g, int ncol, double c, const double *x, int ldi,
                      double *y, int ldo)
{
    orc_lap_plus_diag(g, ncol, 1.0, 0.0, c, NULL, x, ldi, y, ldo);
}

/* Vnl_vec_mult (src/nlocVecRoutines.c:798-883): alpha = dV Chi' x_rc per
 * influencing atom (accumulated over images), alpha *= Gamma, then
 * Hx[grid_pos] += Chi alpha. */
void orc_vnl_vec_mult(const orc_nloc *nl, int ncol, const double *x, int ldi, double *Hx, int ldo)
{
    if (!nl || nl->n_influence == 0 || nl->n_alpha == 0) return;
    const int na = nl->n_alpha;
    double *alpha = (double *)calloc((size_t)na * ncol, sizeof(double));
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int n = 0; n < ncol; n++) {
        const double *xn = x + (long)n * ldi;
        for (int a = 0; a < nl->n_influence; a++) {
            const int ndc = nl->ndc[a], np = nl->nproj[a];
            const int *pos = nl->grid_pos + nl->pos_off[a];
            const double *chi = nl->chi + nl->chi_off[a];
            for (int p = 0; p < np; p++) {
                double s = 0.0;
                for (int i = 0; i < ndc; i++) s += chi[(long)p * ndc + i] * xn[pos[i]];
                alpha[(long)n * na + nl->alpha_off[a] + p] += nl->dV * s;
            }
        }
        for (int r = 0; r < na; r++) alpha[(long)n * na + r] *= nl->gamma[r];
        double *hn = Hx + (long)n * ldo;
        for (int a = 0; a < nl->n_influence; a++) {
            const int ndc = nl->ndc[a], np = nl->nproj[a];
            const int *pos = nl->grid_pos + nl->pos_off[a];
            const double *chi = nl->chi + nl->chi_off[a];
            const double *al = alpha + (long)n * na + nl->alpha_off[a];
            for (int i = 0; i < ndc; i++) {
                double s = 0.0;
                for (int p = 0; p < np; p++) s += chi[(long)p * ndc + i] * al[p];
                hn[pos[i]] += s;
            }
        }
    }
    free(alpha);
}

void orc_hamiltonian(const orc_grid *g, const double *veff, const orc_nloc *nl, int ncol,
                     double c, const double *x, int ldi, double *Hx, int ldo)
{
    orc_lap_plus_diag(g, ncol, -0.5, 1.0, c, veff, x, ldi, Hx, ldo);
    orc_vnl_vec_mult(nl, ncol, x, ldi, Hx, ldo);
}

/* ChebyshevFiltering (src/eigenSolver.c:722-798). */
void orc_chebyshev_filter(const orc_grid *g, const double *veff, const orc_nloc *nl,
                          double *X, int ldi, double *Y, int ldo, int ncol, int m,
                          double a, double b, double a0)
{
    const long nd = (long)g->nx * g->ny * g->nz;
    double e = 0.5 * (b - a), c = 0.5 * (b + a);
    double sigma = e / (a0 - c), sigma1 = sigma, gamma = 2.0 / sigma1;
    orc_hamiltonian(g, veff, nl, ncol, -c, X, ldi, Y, ldo);
    double vscal = sigma1 / e;
    for (int n = 0; n < ncol; n++)
        for (long i = 0; i < nd; i++) Y[i + n * (long)ldo] *= vscal;
    double *Ynew = (double *)malloc(sizeof(double) * nd * ncol);
    for (int j = 1; j < m; j++) {
        double sigma2 = 1.0 / (gamma - sigma);
        orc_hamiltonian(g, veff, nl, ncol, -c, Y, ldo, Ynew, (int)nd);
        double vs = 2.0 * sigma2 / e, vs2 = sigma * sigma2;
#pragma omp parallel for num_threads(g_threads)
        for (int n = 0; n < ncol; n++)
            for (long i = 0; i < nd; i++) {
                double yn = Ynew[i + n * nd] * vs;
                yn -= vs2 * X[i + n * (long)ldi];
                X[i + n * (long)ldi] = Y[i + n * (long)ldo];
                Y[i + n * (long)ldo] = yn;
            }
        sigma = sigma2;
    }
    free(Ynew);
}

/* Sturm count: number of eigenvalues of tridiag(e, d, e) below x. */
static int sturm_count(int n, const double *d, const double *e, double x)
{
    int cnt = 0;
    double q = d[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < n; i++) {
        if (q == 0.0) q = 1e-300;
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
        if (q < 0) cnt++;
    }
    return cnt;
}

/* Extreme eigenvalues by Sturm bisection; stands in for the min/max of
 * LAPACKE_dsterf's spectrum (src/eigenSolver.c:2072-2081). */
void orc_tridiag_minmax(int n, const double *d, const double *e, double *lmin, double *lmax)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        if (d[i] - r < lo) lo = d[i] - r;
        if (d[i] + r > hi) hi = d[i] + r;
    }
    double span = hi - lo + 1e-300;
    for (int which = 0; which < 2; which++) {
        double l = lo - 1e-12 * span, h = hi + 1e-12 * span;
        int target = which == 0 ? 1 : n;  /* count(x) >= target */
        for (int it = 0; it < 200 && h - l > 4e-16 * (fabs(l) + fabs(h)) + 1e-300; it++) {
            double mid = 0.5 * (l + h);
            if (sturm_count(n, d, e, mid) >= target) h = mid; else l = mid;
        }
        if (which == 0) *lmin = 0.5 * (l + h); else *lmax = 0.5 * (l + h);
    }
}

static double dot(long n, const double *a, const double *b)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Lanczos (src/eigenSolver.c:1920-2129), x0 as the start vector. */
int orc_lanczos(const orc_grid *g, const double *veff, const orc_nloc *nl, const double *x0,
                double tol_min, double tol_max, int maxit, double *eigmin, double *eigmax)
{
    const long nd = (long)g->nx * g->ny * g->nz;
    double *Vj = malloc(sizeof(double) * nd), *Vjm1 = malloc(sizeof(double) * nd),
           *Vjp1 = malloc(sizeof(double) * nd);
    double *a = malloc(sizeof(double) * (maxit + 2)), *b = malloc(sizeof(double) * (maxit + 2));
    memcpy(Vjm1, x0, sizeof(double) * nd);
    double vs = 1.0 / sqrt(dot(nd, Vjm1, Vjm1));
    for (long i = 0; i < nd; i++) Vjm1[i] *= vs;
    orc_hamiltonian(g, veff, nl, 1, 0.0, Vjm1, (int)nd, Vj, (int)nd);
    a[0] = dot(nd, Vjm1, Vj);
    for (long i = 0; i < nd; i++) Vj[i] -= a[0] * Vjm1[i];
    b[0] = sqrt(dot(nd, Vj, Vj));
    vs = b[0] == 0.0 ? 1.0 : 1.0 / b[0];
    for (long i = 0; i < nd; i++) Vj[i] *= vs;
    double emin_pre = 0.0, emax_pre = 0.0, err_min = tol_min + 1.0, err_max = tol_max + 1.0;
    *eigmin = *eigmax = 0.0;
    int j = 0;
    while ((err_min > tol_min || err_max > tol_max) && j < maxit) {
        orc_hamiltonian(g, veff, nl, 1, 0.0, Vj, (int)nd, Vjp1, (int)nd);
        a[j + 1] = dot(nd, Vj, Vjp1);
        for (long i = 0; i < nd; i++) {
            Vjp1[i] -= (a[j + 1] * Vj[i] + b[j] * Vjm1[i]);
            Vjm1[i] = Vj[i];
        }
        b[j + 1] = sqrt(dot(nd, Vjp1, Vjp1));
        if (b[j + 1] == 0.0) break;
        vs = 1.0 / b[j + 1];
        for (long i = 0; i < nd; i++) Vj[i] = Vjp1[i] * vs;
        orc_tridiag_minmax(j + 2, a, b, eigmin, eigmax);
        err_min = fabs(*eigmin - emin_pre);
        err_max = fabs(*eigmax - emax_pre);
        emin_pre = *eigmin;
        emax_pre = *eigmax;
        j++;
    }
    free(Vj); free(Vjm1); free(Vjp1); free(a); free(b);
    return j;
}

void orc_gemm_tn(int nd, int na, int nb, const double *A, int lda, const double *B, int ldb,
                 double *C)
{
#pragma omp parallel for num_threads(g_threads) collapse(2)
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++)
            C[i + (long)j * na] = dot(nd, A + (long)i * lda, B + (long)j * ldb);
}

/* DP_Project_Hamiltonian (src/eigenSolver.c:939-1082) on one rank. */
void orc_project_hamiltonian(const orc_grid *g, const double *veff, const orc_nloc *nl,
                             const double *Y, int ldi, double *HY, int ldo, int ncol,
                             double *Hp, double *Mp)
{
    const int nd = g->nx * g->ny * g->nz;
    orc_hamiltonian(g, veff, nl, ncol, 0.0, Y, ldi, HY, ldo);
    orc_gemm_tn(nd, ncol, ncol, Y, ldi, Y, ldi, Mp);
    orc_gemm_tn(nd, ncol, ncol, Y, ldi, HY, ldo, Hp);
}

/* DP_Subspace_Rotation (src/eigenSolver.c:1386-1439): out = Y Q. */
void orc_subspace_rotation(int nd, int ns, const double *Y, int ldy, const double *Q,
                           double *out, int ldo)
{
#pragma omp parallel for num_threads(g_threads)
    for (int j = 0; j < ns; j++) {
        double *o = out + (long)j * ldo;
        for (int i = 0; i < nd; i++) o[i] = 0.0;
        for (int k = 0; k < ns; k++) {
            const double q = Q[k + (long)j * ns];
            const double *y = Y + (long)k * ldy;
            for (int i = 0; i < nd; i++) o[i] += y[i] * q;
        }
    }
}

/* Cyclic Jacobi eigensolver for a symmetric matrix (column-major). */
int orc_syev_jacobi(int n, double *A, double *w, double *V)
{
#define AA(i, j) A[(i) + (long)(j) * n]
#define VV(i, j) V[(i) + (long)(j) * n]
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) VV(i, j) = i == j;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                tot += AA(i, j) * AA(i, j);
                if (i != j) off += AA(i, j) * AA(i, j);
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = AA(p, q);
                if (fabs(apq) < 1e-300) continue;
                double theta = (AA(q, q) - AA(p, p)) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    double akp = AA(k, p), akq = AA(k, q);
                    AA(k, p) = cs * akp - sn * akq;
                    AA(k, q) = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = AA(p, k), aqk = AA(q, k);
                    AA(p, k) = cs * apk - sn * aqk;
                    AA(q, k) = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = VV(k, p), vkq = VV(k, q);
                    VV(k, p) = cs * vkp - sn * vkq;
                    VV(k, q) = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = AA(i, i);
    /* sort ascending */
    for (int i = 0; i < n - 1; i++) {
        int kmin = i;
        for (int k = i + 1; k < n; k++) if (w[k] < w[kmin]) kmin = k;
        if (kmin != i) {
            double t = w[i]; w[i] = w[kmin]; w[kmin] = t;
            for (int r = 0; r < n; r++) { t = VV(r, i); VV(r, i) = VV(r, kmin); VV(r, kmin) = t; }
        }
    }
    return 0;
#undef AA
#undef VV
}

/* Generalized problem by Cholesky reduction + Jacobi (the mathematical
 * contract of dsygvd itype 1, src/eigenSolver.c:1318). */
int orc_sygv(int n, double *Hp, double *Mp, double *lambda, double *Q)
{
#define M(i, j) Mp[(i) + (long)(j) * n]
#define H(i, j) Hp[(i) + (long)(j) * n]
    /* lower Cholesky M = L L' into the lower triangle of Mp */
    for (int j = 0; j < n; j++) {
        double s = M(j, j);
        for (int k = 0; k < j; k++) s -= M(j, k) * M(j, k);
        if (s <= 0.0) return j + 1;
        M(j, j) = sqrt(s);
        for (int i = j + 1; i < n; i++) {
            double t = M(i, j);
            for (int k = 0; k < j; k++) t -= M(i, k) * M(j, k);
            M(i, j) = t / M(j, j);
        }
    }
    double *C = malloc(sizeof(double) * n * n), *V = malloc(sizeof(double) * n * n);
    /* symmetrise H from its upper triangle (uplo 'U') */
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) H(i, j) = H(j, i);
    /* W = L^{-1} H (forward substitution per column) */
    for (int c = 0; c < n; c++)
        for (int i = 0; i < n; i++) {
            double t = H(i, c);
            for (int k = 0; k < i; k++) t -= M(i, k) * H(k, c);
            H(i, c) = t / M(i, i);
        }
    /* C = W L^{-T}: solve C L' = W row-wise */
    for (int r = 0; r < n; r++)
        for (int i = 0; i < n; i++) {
            double t = H(r, i);
            for (int k = 0; k < i; k++) t -= C[r + (long)k * n] * M(i, k);
            C[r + (long)i * n] = t / M(i, i);
        }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) {
            double s = 0.5 * (C[i + (long)j * n] + C[j + (long)i * n]);
            C[i + (long)j * n] = C[j + (long)i * n] = s;
        }
    orc_syev_jacobi(n, C, lambda, V);
    /* Q = L^{-T} V (back substitution) */
    for (int c = 0; c < n; c++)
        for (int i = n - 1; i >= 0; i--) {
            double t = V[i + (long)c * n];
            for (int k = i + 1; k < n; k++) t -= M(k, i) * Q[k + (long)c * n];
            Q[i + (long)c * n] = t / M(i, i);
        }
    free(C); free(V);
    return 0;
#undef M
#undef H
}

/* AndersonExtrapCoeff (src/mixing.c:115-139): LAPACKE_dgelsd with rcond = -1
 * on the symmetric F'F is the minimum-norm least-squares solution, i.e. the
 * eigen-pseudo-inverse with singular values below eps * s_max dropped. */
void orc_anderson_coeff(int N, int m, const double *f, const double *F, double *Gamma)
{
    double *FtF = malloc(sizeof(double) * m * m), *V = malloc(sizeof(double) * m * m);
    double *w = malloc(sizeof(double) * m), *Ftf = malloc(sizeof(double) * m);
    for (int j = 0; j < m; j++) {
        Ftf[j] = dot(N, F + (long)j * N, f);
        for (int i = 0; i <= j; i++) {
            double s = dot(N, F + (long)i * N, F + (long)j * N);
            FtF[i + j * m] = FtF[j + i * m] = s;
        }
    }
    orc_syev_jacobi(m, FtF, w, V);
    double smax = 0.0;
    for (int i = 0; i < m; i++) if (fabs(w[i]) > smax) smax = fabs(w[i]);
    const double thr = 2.220446049250313e-16 * smax;
    for (int i = 0; i < m; i++) Gamma[i] = 0.0;
    for (int k = 0; k < m; k++) {
        if (fabs(w[k]) <= thr) continue;
        double proj = 0.0;
        for (int i = 0; i < m; i++) proj += V[i + k * m] * Ftf[i];
        proj /= w[k];
        for (int i = 0; i < m; i++) Gamma[i] += V[i + k * m] * proj;
    }
    free(FtF); free(V); free(w); free(Ftf);
}

void orc_anderson_extrapolation(int N, int m, double *x_kp1, const double *x_k, const double *f,
                                const double *X, const double *F, double beta)
{
    double *G = malloc(sizeof(double) * m);
    orc_anderson_coeff(N, m, f, F, G);
    for (int i = 0; i < N; i++) {
        double xw = x_k[i], fw = f[i];
        for (int j = 0; j < m; j++) {
            xw -= X[i + (long)j * N] * G[j];
            fw -= F[i + (long)j * N] * G[j];
        }
        x_kp1[i] = xw + beta * fw;
    }
    free(G);
}

int orc_aar(void *ctx, orc_res_fn res, orc_precond_fn precond, double c, int N, double *x,
            const double *b, double omega, double beta, int m, int p, double tol, int max_iter,
            double *relres)
{
    double *r = malloc(sizeof(double) * N), *x_old = malloc(sizeof(double) * N);
    double *f = malloc(sizeof(double) * N), *f_old = malloc(sizeof(double) * N);
    double *X = calloc((size_t)N * m, sizeof(double)), *F = calloc((size_t)N * m, sizeof(double));
    memcpy(x_old, x, sizeof(double) * N);
    double bnorm = sqrt(dot(N, b, b));
    res(ctx, N, c, x, b, r);
    tol *= bnorm;
    double rnorm = tol + 1.0;
    int it = 0;
    while (rnorm > tol && it < max_iter) {
        precond(ctx, N, c, r, f);
        if (it > 0) {
            int ih = (it - 1) % m;
            for (int i = 0; i < N; i++) {
                X[i + (long)ih * N] = x[i] - x_old[i];
                F[i + (long)ih * N] = f[i] - f_old[i];
            }
        }
        memcpy(x_old, x, sizeof(double) * N);
        memcpy(f_old, f, sizeof(double) * N);
        if ((it + 1) % p == 0 && it > 0) {
            orc_anderson_extrapolation(N, m, x, x_old, f, X, F, beta);
            res(ctx, N, c, x, b, r);
            rnorm = sqrt(dot(N, r, r));
        } else {
            for (int i = 0; i < N; i++) x[i] = x_old[i] + omega * f[i];
            res(ctx, N, c, x, b, r);
        }
        it++;
    }
    if (relres) *relres = bnorm > 0 ? sqrt(dot(N, r, r)) / bnorm : 0.0;
    free(r); free(x_old); free(f); free(f_old); free(X); free(F);
    return it;
}

void orc_poisson_residual(void *ctx, int N, double c, const double *x, const double *b, double *r)
{
    const orc_grid *g = (const orc_grid *)ctx;
    orc_lap_vec_mult(g, 1, c, x, N, r, N);
    for (int i = 0; i < N; i++) r[i] += b[i];
}

void orc_jacobi_precond(void *ctx, int N, double c, const double *r, double *f)
{
    const orc_grid *g = (const orc_grid *)ctx;
    double minv = g->d2x[0] + g->d2y[0] + g->d2z[0] + c;
    if (fabs(minv) < 1e-14) minv = 1.0;
    minv = -1.0 / minv;
    for (int i = 0; i < N; i++) f[i] = minv * r[i];
}

int orc_poisson_solve(const orc_grid *g, double *phi, const double *rhs, double tol, int maxit,
                      double *relres)
{
    const int N = g->nx * g->ny * g->nz;
    int it = orc_aar((void *)g, orc_poisson_residual, orc_jacobi_precond, 0.0, N, phi, rhs, 0.6,
                     0.6, 7, 6, tol, maxit, relres);
    if (g->bcx == 0 && g->bcy == 0 && g->bcz == 0) {
        double s = 0.0;
        for (int i = 0; i < N; i++) s += phi[i];
        s /= N;
        for (int i = 0; i < N; i++) phi[i] -= s;
    }
    return it;
}

/* smearing_function (src/occupation.c:31-36): type 1 Gaussian, 0 Fermi-Dirac. */
double orc_smearing(double beta, double lambda, double ef, int type)
{
    return type ? 0.5 * (1.0 - erf(beta * (lambda - ef))) : 1.0 / (1.0 + exp(beta * (lambda - ef)));
}

static double occ_constraint(int ns, const double *lambda, double occfac, double nelec,
                             double beta, int type, double ef)
{
    double g = 0.0;
    for (int n = 0; n < ns; n++) g += occfac * orc_smearing(beta, lambda[n], ef, type);
    return g - nelec;   /* g + NegCharge with NegCharge = -Nelectron */
}

/* local_Calculate_FermiLevel: Brent's method with bracket expansion
 * (src/occupation.c:291-408), Gamma point, one spin channel. */
double orc_fermi_level(int ns, const double *lambda, double occfac, double nelec, double beta,
                       int type, double x1, double x2, double tol, int max_iter, int *status)
{
    const double EPS = 1e-16;
    double a = x1 < x2 ? x1 : x2, b = x1 < x2 ? x2 : x1, c = b, d = 0, e = 0;
    double fa = occ_constraint(ns, lambda, occfac, nelec, beta, type, a);
    double fb = occ_constraint(ns, lambda, occfac, nelec, beta, type, b), fc;
    int ext = 0;
    *status = 0;
    while (fa * fb > 0.0 && ext++ < 10) {
        double w = b - a;
        a -= w / 2.0; b += w / 2.0; c = b;
        fa = occ_constraint(ns, lambda, occfac, nelec, beta, type, a);
        fb = occ_constraint(ns, lambda, occfac, nelec, beta, type, b);
    }
    if (fa * fb > 0.0) { *status = 1; return 0.0; }
    fc = fb;
    for (int it = 1; it <= max_iter; it++) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; e = d = b - a; }
        if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        double tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol, xm = 0.5 * (c - b);
        if (fabs(xm) <= tol1 || fabs(fb) < EPS) return b;
        if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
            double s = fb / fa, p, q, r;
            if (a == c) { p = 2.0 * xm * s; q = 1.0 - s; }
            else {
                q = fa / fc; r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = fabs(p);
            double min1 = 3.0 * xm * q - fabs(tol1 * q), min2 = fabs(e * q);
            if (2.0 * p < (min1 < min2 ? min1 : min2)) { e = d; d = p / q; }
            else { d = xm; e = d; }
        } else { d = xm; e = d; }
        a = b; fa = fb;
        if (fabs(d) > tol1) b += d; else b += xm > 0.0 ? fabs(tol1) : -fabs(tol1);
        fb = occ_constraint(ns, lambda, occfac, nelec, beta, type, b);
    }
    *status = 2;
    return b;
}

void orc_density(int nd, int ns, const double *psi, int ld, const double *occ, double occfac,
                 double dV, double *rho)
{
    for (int i = 0; i < nd; i++) rho[i] = 0.0;
    for (int n = 0; n < ns; n++) {
        const double g = occfac * occ[n];
        const double *p = psi + (long)n * ld;
        for (int i = 0; i < nd; i++) rho[i] += g * p[i] * p[i];
    }
    const double s = 1.0 / dV;
    for (int i = 0; i < nd; i++) rho[i] *= s;
}

/* slater + pz (src/exchangeCorrelation.c:459-529); exc = ex + ec, vxc = vx + vc. */
void orc_lda_pz(int n, const double *rho, double *exc, double *vxc)
{
    const double C2 = 0.75 * cbrt(3.0 / M_PI), C3 = cbrt(3.0 / M_PI);
    const double C31 = cbrt(3.0 / (4.0 * M_PI));
    const double A = 0.0311, B = -0.048, C = 0.002, D = -0.0116;
    const double g1 = -0.1423, b1 = 1.0529, b2 = 0.3334;
    for (int i = 0; i < n; i++) {
        double rc = cbrt(rho[i]);
        double ex = -C2 * rc, vx = -C3 * rc, ec, vc;
        double rs = C31 / rc;
        if (rs < 1.0) {
            double lr = log(rs);
            ec = A * lr + B + C * rs * lr + D * rs;
            vc = lr * (A + (2.0 / 3.0) * C * rs) + (B - A / 3.0) + (2.0 * D - C) * rs / 3.0;
        } else {
            double sq = sqrt(rs), den = 1.0 + b1 * sq + b2 * rs;
            ec = g1 / den;
            vc = (g1 + (7.0 / 6.0) * g1 * b1 * sq + (4.0 / 3.0) * g1 * b2 * rs) / (den * den);
        }
        exc[i] = ex + ec;
        vxc[i] = vx + vc;
    }
}

/* Slater exchange + Perdew-Wang 1992 correlation, p = 1 parameter set
 * (J. P. Perdew and Y. Wang, PRB 45, 13244 (1992), eq. 10; SPARC pw,
 * src/exchangeCorrelation.c:475-499).  v_c = e_c - (rs/3) de_c/drs. */
void orc_lda_pw(int n, const double *rho, double *exc, double *vxc)
{
    const double C2 = 0.75 * cbrt(3.0 / M_PI), C3 = cbrt(3.0 / M_PI);
    const double C31 = cbrt(3.0 / (4.0 * M_PI));
    const double A = 0.031091, a1 = 0.21370;
    const double b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
    for (int i = 0; i < n; i++) {
        double rc = cbrt(rho[i]);
        double ex = -C2 * rc, vx = -C3 * rc;
        double rs = C31 / rc, sq = sqrt(rs);
        double Q = b1 * sq + b2 * rs + b3 * rs * sq + b4 * rs * rs;
        double dQ = 0.5 * b1 / sq + b2 + 1.5 * b3 * sq + 2.0 * b4 * rs;
        double L = log(1.0 + 1.0 / (2.0 * A * Q));
        double ec = -2.0 * A * (1.0 + a1 * rs) * L;
        /* d/drs ln(1 + 1/(2AQ)) = -dQ / (Q (1 + 2AQ)) */
        double dec = -2.0 * A * a1 * L + 2.0 * A * (1.0 + a1 * rs) * dQ / (Q * (1.0 + 2.0 * A * Q));
        exc[i] = ex + ec;
        vxc[i] = vx + ec - rs / 3.0 * dec;
    }
}
