// Alternating Anderson-Richardson (AAR) linear solver on the device, with the
// reference's Poisson residual and Jacobi preconditioner built in
// (SPARC AAR src/linearSolver.c:37-159, AndersonExtrapolation
// src/mixing.c:48-139, poisson_residual src/lapVecRoutines.c:69-79,
// Jacobi_preconditioner src/electrostatics.c:1682-1699).
//
// Per Richardson iteration the built-in path runs two kernels: one fused
// pass (Jacobi f = -r/d, history columns X/F, x_old/f_old, x += omega f)
// and the residual stencil r = b + (Lap + c) x (the b add is fused into the
// stencil epilogue).  Anderson steps add one pass computing F'F and F'f
// together and one update pass; the m x m minimum-norm solve is done on the
// host in FP64.
#include <cmath>
#include <vector>

#include "sgpu_internal.h"

namespace {

constexpr int kAarMaxM = 8;

// f = jac * r (when jac != 0; otherwise f is given); history and Richardson.
__global__ void aar_fused(const double *__restrict__ r, double *__restrict__ f, double jac,
                          double *__restrict__ x, double *__restrict__ x_old,
                          double *__restrict__ f_old, double *__restrict__ Xh,
                          double *__restrict__ Fh, int store_hist, double omega, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fi = jac != 0.0 ? jac * r[i] : f[i];
    if (jac != 0.0) f[i] = fi;
    const double xi = x[i];
    if (store_hist) {
        Xh[i] = xi - x_old[i];
        Fh[i] = fi - f_old[i];
    }
    x_old[i] = xi;
    f_old[i] = fi;
    if (omega != 0.0) x[i] = xi + omega * fi;
}

// Partial sums of F'F (upper triangle, m(m+1)/2 values) and F'f (m values).
template <int M>
__global__ __launch_bounds__(256) void aar_gram(const double *__restrict__ F, long ldf,
                                                const double *__restrict__ f, int m, long n,
                                                double *__restrict__ partial)
{
    constexpr int NV = M * (M + 1) / 2 + M;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double v[M];
#pragma unroll
        for (int a = 0; a < M; a++) v[a] = a < m ? F[a * ldf + i] : 0.0;
        const double fi = f[i];
        int k = 0;
#pragma unroll
        for (int a = 0; a < M; a++)
#pragma unroll
            for (int b = a; b < M; b++) acc[k++] += v[a] * v[b];
#pragma unroll
        for (int a = 0; a < M; a++) acc[k++] += v[a] * fi;
    }
    __shared__ double sh[4][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = acc[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) sh[wave][k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += 256)
        partial[(long)blockIdx.x * NV + k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
}

// One wave per Gram value: lane l sums the partials l, l + 64, ... in order,
// then a fixed butterfly over the wave (deterministic).
__global__ __launch_bounds__(64) void aar_gram_final(const double *__restrict__ partial, int nblk,
                                                     int nv, double *__restrict__ out)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s += partial[(long)b * nv + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[k] = s;
}

// x = x_old - X G + beta (f - F G)
__global__ void aar_anderson_update(double *__restrict__ x, const double *__restrict__ x_old,
                                    const double *__restrict__ f, const double *__restrict__ Xh,
                                    const double *__restrict__ Fh, long ldh, int m, double g0,
                                    double g1, double g2, double g3, double g4, double g5,
                                    double g6, double g7, double beta, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g[kAarMaxM] = {g0, g1, g2, g3, g4, g5, g6, g7};
    double xw = x_old[i], fw = f[i];
#pragma unroll
    for (int a = 0; a < kAarMaxM; a++) {
        if (a < m) {
            xw -= Xh[a * ldh + i] * g[a];
            fw -= Fh[a * ldh + i] * g[a];
        }
    }
    x[i] = xw + beta * fw;
}

// Symmetric eigendecomposition by cyclic Jacobi rotations (m <= 8).
void jacobi_eig(int n, double *A, double *w, double *V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i + j * n] = i == j;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                tot += A[i + j * n] * A[i + j * n];
                if (i != j) off += A[i + j * n] * A[i + j * n];
            }
        if (off == 0.0 || off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p + q * n];
                if (std::fabs(apq) < 1e-300) continue;
                const double th = (A[q + q * n] - A[p + p * n]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = cs * akp - sn * akq;
                    A[k + q * n] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = cs * apk - sn * aqk;
                    A[q + k * n] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = cs * vkp - sn * vkq;
                    V[k + q * n] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i + i * n];
}

}  // namespace

// Minimum-norm least-squares solution of (F'F) G = F'f, singular values
// below eps * s_max dropped: the result LAPACKE_dgelsd(rcond = -1) returns in
// AndersonExtrapCoeff (src/mixing.c:135).
void anderson_coeff(int m, const double *FtF, const double *Ftf, double *G)
{
    double A[kAarMaxM * kAarMaxM], V[kAarMaxM * kAarMaxM], w[kAarMaxM];
    for (int i = 0; i < m * m; i++) A[i] = FtF[i];
    jacobi_eig(m, A, w, V);
    double smax = 0.0;
    for (int i = 0; i < m; i++) smax = std::max(smax, std::fabs(w[i]));
    const double thr = 2.220446049250313e-16 * smax;
    for (int i = 0; i < m; i++) G[i] = 0.0;
    for (int k = 0; k < m; k++) {
        if (std::fabs(w[k]) <= thr) continue;
        double proj = 0.0;
        for (int i = 0; i < m; i++) proj += V[i + k * m] * Ftf[i];
        proj /= w[k];
        for (int i = 0; i < m; i++) G[i] += V[i + k * m] * proj;
    }
}

// G = pinv(F'F) F'f for device F (n x m, column stride ldf) and f: one pass
// computes F'F and F'f together, the m x m solve runs on the host in FP64.
int sgpu_anderson_coeff_dev(sgpu_context *ctx, const double *F, long ldf, int m, const double *f,
                            long n, double *G)
{
    if (m < 1 || m > kAarMaxM) {
        sgpu_set_error("anderson: history %d outside 1..%d", m, kAarMaxM);
        return SGPU_ERR_ARG;
    }
    const int gram_blocks = 256;
    const int nv = kAarMaxM * (kAarMaxM + 1) / 2 + kAarMaxM;
    double *partial = nullptr;
    SGPU_CHECK(sgpu_reduce_buf(ctx, (size_t)gram_blocks * nv + nv, &partial));
    double *out = partial + (size_t)gram_blocks * nv;
    aar_gram<kAarMaxM><<<gram_blocks, 256, 0, ctx->stream>>>(F, ldf, f, m, n, partial);
    aar_gram_final<<<nv, 64, 0, ctx->stream>>>(partial, gram_blocks, nv, out);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipMemcpyAsync(ctx->host_red, out, nv * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    double FtF[kAarMaxM * kAarMaxM], Ftf[kAarMaxM];
    int k = 0;
    for (int a = 0; a < kAarMaxM; a++)
        for (int bb = a; bb < kAarMaxM; bb++, k++)
            if (a < m && bb < m) FtF[a + bb * m] = FtF[bb + a * m] = ctx->host_red[k];
    for (int a = 0; a < kAarMaxM; a++, k++)
        if (a < m) Ftf[a] = ctx->host_red[k];
    anderson_coeff(m, FtF, Ftf, G);
    return SGPU_OK;
}

// x = xk - X G + beta (f - F G) on device vectors (history stride ldh).
int sgpu_anderson_update_dev(sgpu_context *ctx, double *x, const double *xk, const double *f,
                             const double *Xh, const double *Fh, long ldh, int m, const double *G,
                             double beta, long n)
{
    double g[kAarMaxM] = {0};
    for (int a = 0; a < m && a < kAarMaxM; a++) g[a] = G[a];
    aar_anderson_update<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
        x, xk, f, Xh, Fh, ldh, m, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], beta, n);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// AAR on device vectors x (in/out) and b, both of length N = Nd.
int sgpu_aar_device(sgpu_context *ctx, sgpu_res_fn res, sgpu_precond_fn precond, void *user,
                    double c, int N, double *x, const double *b, double omega, double beta, int m,
                    int p, double tol, int max_iter, int *iters, double *relres,
                    const sgpu_aar_recorder *rec)
{
    if (!ctx || !x || !b || N != ctx->nd || m < 1 || m > kAarMaxM || p < 1 || max_iter < 0) {
        sgpu_set_error("AAR: bad arguments (N %d, Nd %ld, m %d, p %d)", N, ctx ? ctx->nd : 0L, m, p);
        return SGPU_ERR_ARG;
    }
    const long n = N;
    const long ld = ctx->ld;
    // r, x_old, f, f_old, X (m), F (m): cached on the context
    const size_t elems = (size_t)(4 + 2 * m) * ld;
    if (elems > ctx->aar_elems) {
        SGPU_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->aar_mem) SGPU_HIP(hipFree(ctx->aar_mem));
        ctx->aar_mem = nullptr;
        ctx->aar_elems = 0;
        SGPU_HIP(hipMalloc((void **)&ctx->aar_mem, elems * sizeof(double)));
        ctx->aar_elems = elems;
    }
    double *mem = ctx->aar_mem;
    double *r = mem, *x_old = mem + ld, *f = mem + 2 * ld, *f_old = mem + 3 * ld;
    double *Xh = mem + 4 * ld, *Fh = Xh + (long)m * ld;
    const unsigned nb = (unsigned)((n + 255) / 256);
    int it = 0;
    double rnorm = 0.0, bnorm = 0.0;
    const double dsum = ctx->d2x[0] + ctx->d2y[0] + ctx->d2z[0] + c;
    const double jac = -1.0 / (std::fabs(dsum) < 1e-14 ? 1.0 : dsum);

    auto residual = [&]() -> int {
        if (res) {
            res(user, N, c, x, b, r);
        } else {
            // r = b + (Lap + c) x
            sgpu_tmark t;
            SGPU_CHECK(sgpu_tbegin(ctx, &t));
            SGPU_CHECK(sgpu_launch_stencil(ctx, SGPU_F64, 1, 1.0, 0.0, c, nullptr, x, ld, r, ld, b,
                                           ld, 1.0, 1.0, true));
            SGPU_CHECK(sgpu_tend(ctx, &t, "aar_residual"));
        }
        if (rec && rec->fn) rec->fn(rec->user, x, r);
        return SGPU_OK;
    };
    SGPU_HIP(hipMemsetAsync(Xh, 0, 2 * (size_t)m * ld * sizeof(double), ctx->stream));
    SGPU_HIP(hipMemcpyAsync(x_old, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    const double *pa[1] = {b}, *pb[1] = {b};
    SGPU_CHECK(sgpu_dot_many(ctx, N, 1, pa, pb, &bnorm));
    bnorm = std::sqrt(bnorm);
    SGPU_CHECK(residual());
    const double atol = tol * bnorm;
    rnorm = atol + 1.0;
    while (rnorm > atol && it < max_iter) {
        const bool anderson = ((it + 1) % p == 0) && it > 0;
        const int ih = it > 0 ? (it - 1) % m : 0;
        if (precond) precond(user, N, c, r, f);
        aar_fused<<<nb, 256, 0, ctx->stream>>>(r, f, precond ? 0.0 : jac, x, x_old, f_old,
                                               Xh + (long)ih * ld, Fh + (long)ih * ld,
                                               it > 0 ? 1 : 0, anderson ? 0.0 : omega, n);
        SGPU_HIP(hipGetLastError());
        if (anderson) {
            double G[kAarMaxM] = {0};
            SGPU_CHECK(sgpu_anderson_coeff_dev(ctx, Fh, ld, m, f, n, G));
            SGPU_CHECK(sgpu_anderson_update_dev(ctx, x, x_old, f, Xh, Fh, ld, m, G, beta, n));
            SGPU_CHECK(residual());
            pa[0] = r; pb[0] = r;
            SGPU_CHECK(sgpu_dot_many(ctx, N, 1, pa, pb, &rnorm));
            rnorm = std::sqrt(rnorm);
            // diverged (not finite, or 1e6 x the right-hand side's norm, which
            // no converging solve reaches): reported as not converged at once
            // instead of after max_iter iterations
            if (!std::isfinite(rnorm) || (bnorm > 0 && rnorm > 1e6 * bnorm)) break;
        } else {
            SGPU_CHECK(residual());
        }
        it++;
    }
    double r2 = 0.0;
    pa[0] = r; pb[0] = r;
    SGPU_CHECK(sgpu_dot_many(ctx, N, 1, pa, pb, &r2));
    if (relres) *relres = bnorm > 0 ? std::sqrt(r2) / bnorm : 0.0;
    if (iters) *iters = it;
    return (!std::isfinite(rnorm) || rnorm > atol) ? SGPU_NOT_CONVERGED : SGPU_OK;
}

extern "C" int sgpu_AAR(sgpu_context *ctx, sgpu_res_fn res, sgpu_precond_fn precond, void *user,
                        double c, int N, double *x_host, const double *b_host, double omega,
                        double beta, int m, int p, double tol, int max_iter, int *iters,
                        double *relres)
{
    if (!ctx || !x_host || !b_host || N != ctx->nd) {
        sgpu_set_error("sgpu_AAR: bad arguments (N %d, Nd %ld)", N, ctx ? ctx->nd : 0L);
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    const long ld = ctx->ld;
    double *xb = nullptr;
    SGPU_HIP(hipMalloc((void **)&xb, 2 * ld * sizeof(double)));
    double *xd = xb, *bd = xb + ld;
    int st = SGPU_OK;
    if (hipMemcpyAsync(xd, x_host, (size_t)N * sizeof(double), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(bd, b_host, (size_t)N * sizeof(double), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess) {
        sgpu_set_error("sgpu_AAR: upload failed");
        st = SGPU_ERR_HIP;
    }
    if (st == SGPU_OK)
        st = sgpu_aar_device(ctx, res, precond, user, c, N, xd, bd, omega, beta, m, p, tol, max_iter,
                             iters, relres);
    if (st >= 0 && (hipMemcpyAsync(x_host, xd, (size_t)N * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream) != hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess)) {
        sgpu_set_error("sgpu_AAR: download failed");
        st = SGPU_ERR_HIP;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(xb);
    return st;
}
