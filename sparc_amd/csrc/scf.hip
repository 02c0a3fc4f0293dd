// Self-consistent field loop on the device: the reference's scf_loop with
// density mixing (SPARC src/electronicGroundState.c:607-1147) around the
// CheFSI eigensolver.  Per SCF iteration:
//
//   CheFSI pass(es)            eigSolve_CheFSI        src/eigenSolver.c:150-263
//   Fermi level, occupations   Calculate_occupation   src/occupation.c:43-465
//   rho_out = occfac sum g psi^2 / dV                 src/electronDensity.c:104-202
//   free energy with rho_in    Calculate_Free_Energy  src/energy.c:41-366
//   error |rho_out-rho_in|/|rho_out|                  src/electronicGroundState.c:1160-1193
//   Pulay + Kerker mixing      Mixing_periodic_pulay  src/mixing.c:213-527
//   phi = Poisson(rho) (mean zero), Vxc = LDA-PZ(rho), Veff = phi + Vxc
//
// Psi stays in HBM for the whole run (FP64 or FP32 blocks); every Nd-vector of
// the loop (densities, potentials, mixing history) is FP64 on the device; the
// dense eigensolve, occupations and energies are FP64.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "sgpu_internal.h"

namespace {

// rho[i] = sum_n w_n psi[n][i]^2, FP64 accumulation, one thread per point.
template <typename T>
__global__ void density_kernel(const T *__restrict__ psi, long ld, int ns,
                               const double *__restrict__ w, long nd, double *__restrict__ rho)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    double s = 0.0;
    for (int n = 0; n < ns; n++) {
        const double v = (double)psi[(long)n * ld + i];
        s += w[n] * v * v;
    }
    rho[i] = s;
}

// LDA exchange-correlation at rho_xc = max(rho + rho_core, 1e-14), as
// add_rho_core does (src/exchangeCorrelation.c:1254-1270): Slater exchange
// (slater, :459-473) plus Perdew-Zunger 1981 correlation (pz, :505-529;
// PRB 23, 5048, Ceperley-Alder fit) or Perdew-Wang 1992 correlation (pw,
// :475-499; PRB 45, 13244, p = 1 parameter set).  Writes the energy density
// rho_xc * e_xc (Calculate_Exc, :429-454) and v_xc.
__global__ void lda_kernel(const double *__restrict__ rho_in, const double *__restrict__ rho_core,
                           long n, int xc, double *__restrict__ exc_density,
                           double *__restrict__ vxc)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = fmax(rho_in[i] + (rho_core ? rho_core[i] : 0.0), 1e-14);
    const double C2 = 0.738558766382022406;   // 3/4 (3/pi)^(1/3)
    const double C3 = 0.984745021842697;      // (3/pi)^(1/3)
    const double C31 = 0.620350490899400087;  // (3/(4 pi))^(1/3)
    const double rc = cbrt(r);
    const double ex = -C2 * rc, vx = -C3 * rc;
    const double rs = C31 / rc;
    double ec, vc;
    if (xc == SGPU_XC_LDA_PW) {
        const double A = 0.031091, a1 = 0.21370;
        const double b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
        const double sq = sqrt(rs);
        const double q = b1 * sq + b2 * rs + b3 * rs * sq + b4 * rs * rs;   // p = 1
        const double G2 = 2.0 * A * q;
        const double G1 = log(1.0 + 1.0 / G2);
        ec = -2.0 * A * (1.0 + a1 * rs) * G1;
        // d ec / d rs, with d G2 / d rs = A (b1 / sqrt(rs) + 2 b2 + 3 b3 sqrt(rs) + 4 b4 rs)
        const double dG2 = A * (b1 / sq + 2.0 * b2 + 3.0 * b3 * sq + 4.0 * b4 * rs);
        const double dec = -2.0 * A * a1 * G1 + 2.0 * A * (1.0 + a1 * rs) * dG2 / (G2 * (G2 + 1.0));
        vc = ec - rs / 3.0 * dec;
    } else {
        const double A = 0.0311, B = -0.048, C = 0.002, D = -0.0116;
        const double g1 = -0.1423, b1 = 1.0529, b2 = 0.3334;
        if (rs < 1.0) {
            const double lr = log(rs);
            ec = A * lr + B + C * rs * lr + D * rs;
            vc = lr * (A + (2.0 / 3.0) * C * rs) + (B - A / 3.0) + (2.0 * D - C) * rs / 3.0;
        } else {
            const double sq = sqrt(rs), den = 1.0 + b1 * sq + b2 * rs;
            ec = g1 / den;
            vc = (g1 + (7.0 / 6.0) * g1 * b1 * sq + (4.0 / 3.0) * g1 * b2 * rs) / (den * den);
        }
    }
    exc_density[i] = r * (ex + ec);
    vxc[i] = vx + vc;
}

// veff64 = phi + vxc, veff32 = (float) veff64
__global__ void veff_kernel(const double *__restrict__ phi, const double *__restrict__ vxc, long n,
                            double *__restrict__ v64, float *__restrict__ v32)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = phi[i] + vxc[i];
    v64[i] = v;
    v32[i] = (float)v;
}

// y = a*x + b*z  (z may be null; y may alias x or z elementwise)
__global__ void axpbz(double *y, double a, const double *x, double b, const double *z, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = a * x[i] + (z ? b * z[i] : 0.0);
}

__global__ void fill_kernel(double *__restrict__ y, double a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a;
}

// y = max(y, 0); flag set when a negative entry was clipped
__global__ void clamp_nonneg(double *__restrict__ y, long n, int *__restrict__ flag)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (y[i] < 0.0) {
        y[i] = 0.0;
        *flag = 1;
    }
}

// Pulay history columns R(:,ih) = x_k - x_km1, F(:,ih) = f_k - f_km1, with
// f_km1 <- f_k, f_k <- g_k - x_k first (src/mixing.c:262-290).
__global__ void pulay_hist(const double *__restrict__ g, const double *__restrict__ xk,
                           const double *__restrict__ xkm1, double *__restrict__ fk,
                           double *__restrict__ fkm1, double *__restrict__ Rc,
                           double *__restrict__ Fc, int store, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fo = fk[i];
    const double fn = g[i] - xk[i];
    fkm1[i] = fo;
    fk[i] = fn;
    if (store) {
        Rc[i] = xk[i] - xkm1[i];
        Fc[i] = fn - fo;
    }
}

// x_wavg = x_k - R G, f_wavg = f_k - F G  (AndersonExtrapWtdAvg, src/mixing.c:77-106)
__global__ void wtd_avg(double *__restrict__ xw, double *__restrict__ fw,
                        const double *__restrict__ xk, const double *__restrict__ fk,
                        const double *__restrict__ R, const double *__restrict__ F, long ldh, int m,
                        double g0, double g1, double g2, double g3, double g4, double g5,
                        double g6, double g7, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g[8] = {g0, g1, g2, g3, g4, g5, g6, g7};
    double x = xk[i], f = fk[i];
#pragma unroll
    for (int a = 0; a < 8; a++)
        if (a < m) {
            x -= R[a * ldh + i] * g[a];
            f -= F[a * ldh + i] * g[a];
        }
    xw[i] = x;
    fw[i] = f;
}

// ---------------------------------------------------------------- host math

// smearing_function (src/occupation.c:31-36): type 1 Gaussian, 0 Fermi-Dirac
double smearing(double beta, double lam, double ef, int type)
{
    return type ? 0.5 * (1.0 - std::erf(beta * (lam - ef))) : 1.0 / (1.0 + std::exp(beta * (lam - ef)));
}

double occ_constraint(const std::vector<double> &lam, double occfac, double nelec, double beta,
                      int type, double ef)
{
    double g = 0.0;
    for (double l : lam) g += occfac * smearing(beta, l, ef, type);
    return g - nelec;
}

// Brent's method with bracket expansion (local_Calculate_FermiLevel,
// src/occupation.c:291-408).  Returns 0 on success.
int fermi_level(const std::vector<double> &lam, double occfac, double nelec, double beta, int type,
                double x1, double x2, double tol, int max_iter, double *ef_out)
{
    const double EPS = 1e-16;
    double a = std::min(x1, x2), b = std::max(x1, x2), c = b, d = 0, e = 0;
    double fa = occ_constraint(lam, occfac, nelec, beta, type, a);
    double fb = occ_constraint(lam, occfac, nelec, beta, type, b), fc;
    int ext = 0;
    while (fa * fb > 0.0 && ext++ < 10) {
        const double w = b - a;
        a -= w / 2.0;
        b += w / 2.0;
        c = b;
        fa = occ_constraint(lam, occfac, nelec, beta, type, a);
        fb = occ_constraint(lam, occfac, nelec, beta, type, b);
    }
    if (fa * fb > 0.0) return 1;
    fc = fb;
    for (int it = 1; it <= max_iter; it++) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; e = d = b - a; }
        if (std::fabs(fc) < std::fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        const double tol1 = 2.0 * EPS * std::fabs(b) + 0.5 * tol, xm = 0.5 * (c - b);
        if (std::fabs(xm) <= tol1 || std::fabs(fb) < EPS) { *ef_out = b; return 0; }
        if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
            double s = fb / fa, p, q, r;
            if (a == c) { p = 2.0 * xm * s; q = 1.0 - s; }
            else {
                q = fa / fc; r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = std::fabs(p);
            const double min1 = 3.0 * xm * q - std::fabs(tol1 * q), min2 = std::fabs(e * q);
            if (2.0 * p < std::min(min1, min2)) { e = d; d = p / q; }
            else { d = xm; e = d; }
        } else { d = xm; e = d; }
        a = b; fa = fb;
        b += std::fabs(d) > tol1 ? d : (xm > 0.0 ? std::fabs(tol1) : -std::fabs(tol1));
        fb = occ_constraint(lam, occfac, nelec, beta, type, b);
    }
    return 2;
}

// Calculate_electronicEntropy / Calculate_entropy_term (src/energy.c:241-330)
double entropy(const std::vector<double> &lam, const std::vector<double> &occ, double ef,
               double beta, int type, double occfac)
{
    double s = 0.0;
    const double c = 0.5 / std::sqrt(M_PI);
    for (size_t n = 0; n < lam.size(); n++) {
        if (type == 0) {
            const double g = occ[n];
            if (g > 1e-12 && 1.0 - g > 1e-12) s += -(g * std::log(g) + (1.0 - g) * std::log(1.0 - g));
        } else {
            const double x = beta * (lam[n] - ef);
            s += c * std::exp(-x * x);
        }
    }
    return -occfac / beta * s;
}

}  // namespace

extern "C" int sgpu_scf_run(sgpu_context *ctx, const sgpu_scf_params *par, const double *b_host,
                            const double *rho0_host, const double *rho_core_host,
                            double *lambda_out, double *rho_out_host,
                            double *phi_out_host, double *iter_ms, double *iter_err,
                            double *iter_etot, sgpu_scf_result *res)
{
    if (!ctx || !par || !b_host || !rho0_host || !lambda_out || !res) {
        sgpu_set_error("sgpu_scf_run: null argument");
        return SGPU_ERR_ARG;
    }
    const int ns = par->nstates, m = par->mix_history;
    if (ns <= 0 || par->cheb_degree < 1 || m < 1 || m > 8 || par->max_scf < 1 || par->beta <= 0 ||
        par->nelec <= 0 || (par->dtype != SGPU_F64 && par->dtype != SGPU_F32) ||
        par->nelec > par->occfac * ns || par->rho_trigger < 1 || par->pulay_freq < 1 ||
        (par->xc != SGPU_XC_LDA_PZ && par->xc != SGPU_XC_LDA_PW)) {
        sgpu_set_error("sgpu_scf_run: bad parameters");
        return SGPU_ERR_ARG;
    }
    if (par->psi_out && (par->psi_out->ctx != ctx || par->psi_out->ncol != ns)) {
        sgpu_set_error("sgpu_scf_run: psi_out must be a block of this context with nstates columns");
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    memset(res, 0, sizeof(*res));
    const long nd = ctx->nd, ld = ctx->ld;
    const double dV = ctx->hx * ctx->hy * ctx->hz;
    const unsigned nb = (unsigned)((nd + 255) / 256);
    hipStream_t s = ctx->stream;

    // FP64 device vectors of stride ld, then the Pulay history and small arrays
    enum { RHO_IN, RHO_OUT, XKM1, FK, FKM1, PF, XW, FW, PHI, VXC, EXC, BCH, RHS, LF, ONES, CORE,
           PHI0, PF0, NVEC };
    double *mem = nullptr;
    const size_t elems = (size_t)(NVEC + 2 * m) * ld + ns + 64;
    SGPU_HIP(hipMalloc((void **)&mem, elems * sizeof(double)));
    double *vec[NVEC];
    for (int k = 0; k < NVEC; k++) vec[k] = mem + (long)k * ld;
    double *R = mem + (long)NVEC * ld, *F = R + (long)m * ld;
    double *w_occ = F + (long)m * ld;
    int *flag = reinterpret_cast<int *>(w_occ + ns);
    double *rho_in = vec[RHO_IN], *rho_out = vec[RHO_OUT], *phi = vec[PHI];

    sgpu_block *X = nullptr, *Y = nullptr, *L0 = nullptr;
    std::vector<double> lam(ns), occ(ns), x0(nd), wh(ns);

    // phi = Poisson(4 pi (b + rho)) with zero mean, Vxc / exc = LDA-PZ(rho),
    // Veff = phi + Vxc  (Calculate_elecstPotential src/electrostatics.c:1593-1679,
    // Calculate_Vxc, Calculate_Veff_loc_dmcomm_phi)
    auto potential = [&](const double *rho) -> int {
        axpbz<<<nb, 256, 0, s>>>(vec[RHS], 4.0 * M_PI, vec[BCH], 4.0 * M_PI, rho, nd);
        SGPU_HIP(hipGetLastError());
        int pit = 0;
        double prr = 0.0;
        sgpu_tmark tp;
        SGPU_CHECK(sgpu_tbegin(ctx, &tp));
        if (par->poisson_precond)
            SGPU_HIP(hipMemcpyAsync(vec[PHI0], phi, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
        int st = sgpu_aar_device(ctx, nullptr, par->poisson_precond, par->poisson_precond_user,
                                 0.0, (int)nd, phi, vec[RHS], 0.6, 0.6, 7, 6, par->tol_poisson,
                                 par->maxit_poisson, &pit, &prr);
        if (st < 0) return st;
        if (st != SGPU_OK && par->poisson_precond) {
            // the solve with the caller's preconditioner did not converge:
            // Jacobi-AAR (the reference's solve) from the same start
            res->poisson_iters += pit;
            res->poisson_fallbacks++;
            SGPU_HIP(hipMemcpyAsync(phi, vec[PHI0], nd * sizeof(double), hipMemcpyDeviceToDevice, s));
            st = sgpu_aar_device(ctx, nullptr, nullptr, nullptr, 0.0, (int)nd, phi, vec[RHS], 0.6,
                                 0.6, 7, 6, par->tol_poisson, par->maxit_poisson, &pit, &prr);
            if (st < 0) return st;
        }
        SGPU_CHECK(sgpu_tend(ctx, &tp, "scf_poisson"));
        res->poisson_iters += pit;
        double sum = 0.0;
        const double *pa[1] = {phi}, *pb[1] = {vec[ONES]};
        SGPU_CHECK(sgpu_dot_many(ctx, (int)nd, 1, pa, pb, &sum));
        axpbz<<<nb, 256, 0, s>>>(phi, 1.0, phi, -sum / (double)nd, vec[ONES], nd);
        lda_kernel<<<nb, 256, 0, s>>>(rho, rho_core_host ? vec[CORE] : nullptr, nd, par->xc,
                                      vec[EXC], vec[VXC]);
        veff_kernel<<<nb, 256, 0, s>>>(phi, vec[VXC], nd, ctx->veff64, ctx->veff32);
        SGPU_HIP(hipGetLastError());
        ctx->has_veff = true;
        return SGPU_OK;
    };

    auto run = [&]() -> int {
        SGPU_HIP(hipMemsetAsync(mem, 0, elems * sizeof(double), s));
        fill_kernel<<<nb, 256, 0, s>>>(vec[ONES], 1.0, nd);
        SGPU_HIP(hipMemcpyAsync(vec[BCH], b_host, nd * sizeof(double), hipMemcpyHostToDevice, s));
        if (rho_core_host)
            SGPU_HIP(hipMemcpyAsync(vec[CORE], rho_core_host, nd * sizeof(double),
                                    hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(rho_in, rho0_host, nd * sizeof(double), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(vec[XKM1], rho0_host, nd * sizeof(double), hipMemcpyHostToDevice, s));
        SGPU_CHECK(sgpu_block_create(ctx, ns, par->dtype, &X));
        SGPU_CHECK(sgpu_block_create(ctx, ns, par->dtype, &Y));
        SGPU_CHECK(sgpu_block_fill_uniform(X, par->psi_seed, -0.5, 0.5));
        SGPU_CHECK(sgpu_block_create(ctx, 1, SGPU_F64, &L0));
        SGPU_CHECK(sgpu_block_fill_uniform(L0, par->lanczos_seed, 0.0, 1.0));
        SGPU_CHECK(sgpu_block_download(L0, x0.data(), nd, 1));
        if (!ctx->veff64) SGPU_HIP(hipMalloc((void **)&ctx->veff64, nd * sizeof(double)));
        if (!ctx->veff32) SGPU_HIP(hipMalloc((void **)&ctx->veff32, nd * sizeof(float)));
        SGPU_CHECK(potential(rho_in));
        SGPU_HIP(hipStreamSynchronize(s));

        double eigmin = 0.0, eigmax = 0.0, cutoff = 0.0, ef = 0.0;
        double sum_ms = 0.0;
        int scf = 0;
        for (; scf < par->max_scf; scf++) {
            const auto t0 = std::chrono::steady_clock::now();
            // ---- eigSolve_CheFSI (src/eigenSolver.c:150-263)
            const int npass = scf == 0 ? par->rho_trigger : 1;
            for (int pass = 0; pass < npass; pass++) {
                const int count = scf == 0 ? pass : par->rho_trigger + scf - 1;
                // Chebyshevfilter_constants (src/eigenSolver.c:589-715)
                if (count == 0) {
                    int lit = 0;
                    sgpu_tmark tl;
                    SGPU_CHECK(sgpu_tbegin(ctx, &tl));
                    SGPU_CHECK(sgpu_lanczos(ctx, x0.data(), par->tol_lanczos, par->tol_lanczos, 1000,
                                            &eigmin, &eigmax, &lit));
                    SGPU_CHECK(sgpu_tend(ctx, &tl, "scf_lanczos"));
                    eigmax *= 1.01;
                    eigmin -= 0.1;
                    cutoff = 0.5 * (eigmin + eigmax);
                } else {
                    if (count >= par->rho_trigger) eigmin = lam[0];
                    cutoff = lam[ns - 1] + 0.1;
                }
                sgpu_tmark tc;
                SGPU_CHECK(sgpu_tbegin(ctx, &tc));
                SGPU_CHECK(sgpu_chefsi_dev(ctx, X, Y, par->cheb_degree, cutoff, eigmax, eigmin,
                                           lam.data()));
                SGPU_CHECK(sgpu_tend(ctx, &tc, "scf_chefsi"));
                // Calculate_occupation(eigmin - 1, eigmax + 1, 1e-12, 100)
                if (fermi_level(lam, par->occfac, par->nelec, par->beta, par->smear_type, lam[0] - 1.0,
                                lam[ns - 1] + 1.0, 1e-12, 100, &ef) != 0) {
                    sgpu_set_error("sgpu_scf_run: Fermi level not bracketed (too few states?)");
                    return SGPU_ERR_SOLVER;
                }
                for (int n = 0; n < ns; n++) occ[n] = smearing(par->beta, lam[n], ef, par->smear_type);
            }
            // ---- rho_out = occfac sum_n g_n psi_n^2 / dV
            sgpu_tmark td;
            SGPU_CHECK(sgpu_tbegin(ctx, &td));
            for (int n = 0; n < ns; n++) wh[n] = par->occfac * occ[n] / dV;
            SGPU_HIP(hipMemcpyAsync(w_occ, wh.data(), ns * sizeof(double), hipMemcpyHostToDevice, s));
            if (X->dtype == SGPU_F64)
                density_kernel<double><<<nb, 256, 0, s>>>((const double *)X->ptr, X->ld, ns, w_occ,
                                                          nd, rho_out);
            else
                density_kernel<float><<<nb, 256, 0, s>>>((const float *)X->ptr, X->ld, ns, w_occ, nd,
                                                         rho_out);
            SGPU_HIP(hipGetLastError());
            // ---- free energy with rho_in (Calculate_Free_Energy, src/energy.c:41-199)
            double d[6];
            const double *pa[6] = {vec[EXC], vec[BCH], rho_in, rho_in, rho_out, rho_out};
            const double *pb[6] = {vec[ONES], phi, phi, vec[VXC], rho_out, rho_in};
            SGPU_CHECK(sgpu_dot_many(ctx, (int)nd, 6, pa, pb, d));
            double eband = 0.0;
            for (int n = 0; n < ns; n++) eband += par->occfac * occ[n] * lam[n];
            const double exc = d[0] * dV, e1 = 0.5 * d[1] * dV, e2 = 0.5 * d[2] * dV,
                         e3 = d[3] * dV;
            const double ent = entropy(lam, occ, ef, par->beta, par->smear_type, par->occfac);
            const double etot = eband + e1 - e2 - e3 + exc + par->Esc + ent;
            // ---- Evaluate_scf_error: |rho_out - rho_in| / |rho_out|
            double dd = 0.0;
            {
                axpbz<<<nb, 256, 0, s>>>(vec[LF], 1.0, rho_out, -1.0, rho_in, nd);
                const double *qa[1] = {vec[LF]}, *qb[1] = {vec[LF]};
                SGPU_CHECK(sgpu_dot_many(ctx, (int)nd, 1, qa, qb, &dd));
            }
            const double err = std::sqrt(dd / d[4]);
            SGPU_CHECK(sgpu_tend(ctx, &td, "scf_density_energy"));
            res->Etot = etot; res->Eband = eband; res->Exc = exc; res->E1 = e1; res->E2 = e2;
            res->E3 = e3; res->Entropy = ent; res->Efermi = ef; res->scf_error = err;
            if (!(etot == etot)) {
                sgpu_set_error("sgpu_scf_run: total energy is NaN");
                return SGPU_ERR_SOLVER;
            }
            // ---- Mixing_periodic_pulay (src/mixing.c:213-447), density, Kerker
            const bool pulay = ((scf + 1) % par->pulay_freq == 0) && scf > 0;
            const double amix = pulay ? par->mix_beta : par->mix_omega;
            const int ih = scf > 0 ? (scf - 1) % m : 0;
            pulay_hist<<<nb, 256, 0, s>>>(rho_out, rho_in, vec[XKM1], vec[FK], vec[FKM1],
                                          R + (long)ih * ld, F + (long)ih * ld, scf > 0 ? 1 : 0, nd);
            SGPU_HIP(hipGetLastError());
            if (pulay) {
                double G[8] = {0};
                SGPU_CHECK(sgpu_anderson_coeff_dev(ctx, F, ld, m, vec[FK], nd, G));
                wtd_avg<<<nb, 256, 0, s>>>(vec[XW], vec[FW], rho_in, vec[FK], R, F, ld, m, G[0], G[1],
                                           G[2], G[3], G[4], G[5], G[6], G[7], nd);
            } else {
                axpbz<<<nb, 256, 0, s>>>(vec[XW], 1.0, rho_in, 0.0, nullptr, nd);
                axpbz<<<nb, 256, 0, s>>>(vec[FW], 1.0, vec[FK], 0.0, nullptr, nd);
            }
            SGPU_HIP(hipGetLastError());
            // Kerker_precond (src/mixing.c:462-527): Lf = (L - kTF^2 idiemac) f,
            // solve -(L - kTF^2) Pf = Lf by AAR warm-started from the last Pf, Pf *= -amix
            const double k2 = par->kerker_ktf * par->kerker_ktf;
            sgpu_tmark tk;
            SGPU_CHECK(sgpu_tbegin(ctx, &tk));
            SGPU_CHECK(sgpu_launch_stencil(ctx, SGPU_F64, 1, 1.0, 0.0, -k2 * par->kerker_thresh,
                                           nullptr, vec[FW], ld, vec[LF], ld, nullptr, 0, 1.0, 0.0,
                                           true));
            int kit = 0;
            double krr = 0.0;
            if (par->kerker_precond)
                SGPU_HIP(hipMemcpyAsync(vec[PF0], vec[PF], nd * sizeof(double),
                                        hipMemcpyDeviceToDevice, s));
            int kst = sgpu_aar_device(ctx, nullptr, par->kerker_precond, par->kerker_precond_user,
                                      -k2, (int)nd, vec[PF], vec[LF], 0.6, 0.6, 7, 6,
                                      par->tol_precond, 1000, &kit, &krr);
            if (kst < 0) return kst;
            if (kst != SGPU_OK && par->kerker_precond) {
                // not converged with the caller's preconditioner: Jacobi-AAR
                // (the reference's solve) from the same start
                res->kerker_iters += kit;
                res->kerker_fallbacks++;
                SGPU_HIP(hipMemcpyAsync(vec[PF], vec[PF0], nd * sizeof(double),
                                        hipMemcpyDeviceToDevice, s));
                kst = sgpu_aar_device(ctx, nullptr, nullptr, nullptr, -k2, (int)nd, vec[PF],
                                      vec[LF], 0.6, 0.6, 7, 6, par->tol_precond, 1000, &kit, &krr);
                if (kst < 0) return kst;
            }
            res->kerker_iters += kit;
            if (std::fabs(par->kerker_ktf) < 1e-14) {
                double sum = 0.0;
                const double *qa[1] = {vec[PF]}, *qb[1] = {vec[ONES]};
                SGPU_CHECK(sgpu_dot_many(ctx, (int)nd, 1, qa, qb, &sum));
                axpbz<<<nb, 256, 0, s>>>(vec[PF], 1.0, vec[PF], -sum / (double)nd, vec[ONES], nd);
            }
            axpbz<<<nb, 256, 0, s>>>(vec[PF], -amix, vec[PF], 0.0, nullptr, nd);
            SGPU_CHECK(sgpu_tend(ctx, &tk, "scf_kerker"));
            // x_kp1 = x_wavg + Pf into the rho_out buffer, clipped at zero and
            // rescaled to the electron count
            axpbz<<<nb, 256, 0, s>>>(rho_out, 1.0, vec[XW], 1.0, vec[PF], nd);
            SGPU_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
            clamp_nonneg<<<nb, 256, 0, s>>>(rho_out, nd, flag);
            SGPU_HIP(hipGetLastError());
            double int_rho = 0.0;
            {
                const double *qa[1] = {rho_out}, *qb[1] = {vec[ONES]};
                SGPU_CHECK(sgpu_dot_many(ctx, (int)nd, 1, qa, qb, &int_rho));
            }
            int_rho *= dV;
            axpbz<<<nb, 256, 0, s>>>(rho_out, par->nelec / int_rho, rho_out, 0.0, nullptr, nd);
            // x_km1 = x_k, x_k = x_kp1
            SGPU_HIP(hipMemcpyAsync(vec[XKM1], rho_in, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
            SGPU_HIP(hipMemcpyAsync(rho_in, rho_out, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
            // ---- potential of the mixed density
            SGPU_CHECK(potential(rho_in));
            SGPU_HIP(hipStreamSynchronize(s));
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (iter_ms) iter_ms[scf] = ms;
            if (iter_err) iter_err[scf] = err;
            if (iter_etot) iter_etot[scf] = etot;
            res->total_ms += ms;
            if (scf >= 1) sum_ms += ms;
            if (err < par->tol_scf && scf + 1 >= par->min_scf) {
                res->converged = 1;
                scf++;
                break;
            }
        }
        res->scf_iters = scf;
        res->mean_iter_ms = scf > 1 ? sum_ms / (scf - 1) : res->total_ms;
        for (int n = 0; n < ns; n++) lambda_out[n] = lam[n];
        if (rho_out_host)
            SGPU_HIP(hipMemcpyAsync(rho_out_host, rho_in, nd * sizeof(double), hipMemcpyDeviceToHost, s));
        if (phi_out_host)
            SGPU_HIP(hipMemcpyAsync(phi_out_host, phi, nd * sizeof(double), hipMemcpyDeviceToHost, s));
        if (par->vxc_out)
            SGPU_HIP(hipMemcpyAsync(par->vxc_out, vec[VXC], nd * sizeof(double), hipMemcpyDeviceToHost,
                                    s));
        if (par->psi_out) SGPU_CHECK(sgpu_block_copy(X, par->psi_out));
        SGPU_HIP(hipStreamSynchronize(s));
        return SGPU_OK;
    };
    int st = run();
    (void)hipStreamSynchronize(s);
    sgpu_block_destroy(X);
    sgpu_block_destroy(Y);
    sgpu_block_destroy(L0);
    (void)hipFree(mem);
    if (st < 0) return st;
    return res->converged ? SGPU_OK : SGPU_NOT_CONVERGED;
}
