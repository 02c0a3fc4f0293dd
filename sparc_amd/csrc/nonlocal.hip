// Kleinman-Bylander nonlocal apply  out += scale * sum_J Chi_J Gamma_J (dV Chi_J' x)
// (SPARC Vnl_vec_mult, src/nlocVecRoutines.c:798-883), in three launches:
//
//  1. nl_project: per influence entry and column group, alpha_part =
//     Chi_J' x[grid_pos] with one wavefront per CPW columns; lanes stride the
//     entry's points (coalesced gathers), each lane keeps CPW x NP partial sums
//     in registers, one cross-lane reduction at the end.  No atomics.
//  2. nl_reduce_scale: alpha_s[row] = scale * dV * Gamma[row] * sum of the
//     partial rows of every image of the row's atom (fixed order).
//  3. nl_apply: point-centric scatter-free apply over the covered grid points:
//     out[q] += sum_{(J,i) at q} sum_p Chi_J[i,p] alpha_s[J,p], CBP columns per
//     thread, so overlapping projector spheres need no atomics either.
#include "sgpu_internal.h"

namespace {

constexpr int kNlThreads = 256;
constexpr int kNlWaves = kNlThreads / 64;

template <typename T, int NP>
__global__ __launch_bounds__(kNlThreads) void nl_project(
    const T *__restrict__ x, long ldx, int ncol, const int *__restrict__ ndc,
    const int *__restrict__ nproj, const long *__restrict__ pos_off,
    const int *__restrict__ grid_pos, const long *__restrict__ chi_off, const T *__restrict__ chi,
    const int *__restrict__ part_row, int n_part, T *__restrict__ part)
{
    constexpr int CPW = NP >= 64 ? 1 : 64 / NP;
    const int a = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col0 = (blockIdx.y * kNlWaves + wave) * CPW;
    if (col0 >= ncol) return;
    const int nd_a = ndc[a], np = nproj[a];
    const int *pos = grid_pos + pos_off[a];
    const T *ch = chi + chi_off[a];
    T acc[CPW][NP];
#pragma unroll
    for (int c = 0; c < CPW; c++)
#pragma unroll
        for (int p = 0; p < NP; p++) acc[c][p] = T(0);
    for (int i = lane; i < nd_a; i += 64) {
        const long q = pos[i];
        T xv[CPW];
#pragma unroll
        for (int c = 0; c < CPW; c++)
            xv[c] = (col0 + c < ncol) ? x[(long)(col0 + c) * ldx + q] : T(0);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if (p < np) {
                const T cv = ch[(long)p * nd_a + i];
#pragma unroll
                for (int c = 0; c < CPW; c++) acc[c][p] += cv * xv[c];
            }
        }
    }
    const int prow = part_row[a];
#pragma unroll
    for (int c = 0; c < CPW; c++) {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if (p >= np) continue;
            T s = acc[c][p];
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (lane == 0 && col0 + c < ncol) part[(long)(col0 + c) * n_part + prow + p] = s;
        }
    }
}

template <typename T>
__global__ void nl_reduce_scale(const T *__restrict__ part, int n_part, int n_alpha, int ncol,
                                const int *__restrict__ arow_ptr, const int *__restrict__ arow_src,
                                const double *__restrict__ gamma, double scale,
                                T *__restrict__ alpha_s)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_alpha * ncol) return;
    const int col = (int)(t / n_alpha);
    const int r = (int)(t - (long)col * n_alpha);
    T s = T(0);
    for (int e = arow_ptr[r]; e < arow_ptr[r + 1]; e++) s += part[(long)col * n_part + arow_src[e]];
    alpha_s[(long)col * n_alpha + r] = (T)(scale * gamma[r]) * s;
}

constexpr int kCBP = 8;

template <typename T>
__global__ __launch_bounds__(256) void nl_apply(
    T *__restrict__ out, long ldo, int ncol, const int *__restrict__ cov_pts, int n_cov,
    const int *__restrict__ inv_ptr, const int *__restrict__ inv_ent,
    const int *__restrict__ inv_loc, const int *__restrict__ ndc, const int *__restrict__ nproj,
    const long *__restrict__ chi_off, const int *__restrict__ alpha_off,
    const T *__restrict__ chi, const T *__restrict__ alpha_s, int n_alpha)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_cov) return;
    const int col0 = blockIdx.y * kCBP;
    const int q = cov_pts[t];
    T acc[kCBP];
#pragma unroll
    for (int c = 0; c < kCBP; c++) acc[c] = T(0);
    for (int e = inv_ptr[q]; e < inv_ptr[q + 1]; e++) {
        const int a = inv_ent[e], i = inv_loc[e];
        const int np = nproj[a], nd_a = ndc[a];
        const T *ch = chi + chi_off[a] + i;
        const T *al = alpha_s + alpha_off[a];
        for (int p = 0; p < np; p++) {
            const T cv = ch[(long)p * nd_a];
#pragma unroll
            for (int c = 0; c < kCBP; c++)
                if (col0 + c < ncol) acc[c] += cv * al[(long)(col0 + c) * n_alpha + p];
        }
    }
#pragma unroll
    for (int c = 0; c < kCBP; c++)
        if (col0 + c < ncol) out[(long)(col0 + c) * ldo + q] += acc[c];
}

template <typename T, int NP>
void launch_project(sgpu_context *ctx, const T *x, long ldx, int ncol, const T *chi, T *part)
{
    constexpr int CPW = NP >= 64 ? 1 : 64 / NP;
    const sgpu_nonlocal &nl = ctx->nl;
    dim3 grid(nl.n_infl, (ncol + kNlWaves * CPW - 1) / (kNlWaves * CPW));
    nl_project<T, NP><<<grid, kNlThreads, 0, ctx->stream>>>(
        x, ldx, ncol, nl.ndc, nl.nproj, nl.pos_off, nl.grid_pos, nl.chi_off, chi, nl.part_row,
        nl.n_part, part);
}

template <typename T>
int launch_vnl_typed(sgpu_context *ctx, int ncol, const T *x, long ldx, T *out, long ldo,
                     double scale, const T *chi)
{
    const sgpu_nonlocal &nl = ctx->nl;
    if (nl.n_infl == 0 || nl.n_alpha == 0 || ncol <= 0) return SGPU_OK;
    if (ncol > 65535 * kCBP) {
        sgpu_set_error("nonlocal: ncol %d exceeds the launch limit", ncol);
        return SGPU_ERR_ARG;
    }
    void *buf = nullptr;
    const size_t part_elems = (size_t)nl.n_part * ncol;
    const size_t alpha_elems = (size_t)nl.n_alpha * ncol;
    SGPU_CHECK(sgpu_scratch(ctx, (part_elems + alpha_elems) * sizeof(T) + 256, &buf));
    T *part = (T *)buf;
    T *alpha_s = part + part_elems;
    if (nl.max_np <= 8) launch_project<T, 8>(ctx, x, ldx, ncol, chi, part);
    else if (nl.max_np <= 16) launch_project<T, 16>(ctx, x, ldx, ncol, chi, part);
    else if (nl.max_np <= 32) launch_project<T, 32>(ctx, x, ldx, ncol, chi, part);
    else launch_project<T, 64>(ctx, x, ldx, ncol, chi, part);
    SGPU_HIP(hipGetLastError());
    {
        long total = (long)nl.n_alpha * ncol;
        nl_reduce_scale<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
            part, nl.n_part, nl.n_alpha, ncol, nl.arow_ptr, nl.arow_src, nl.gamma,
            scale * nl.dV, alpha_s);
        SGPU_HIP(hipGetLastError());
    }
    if (nl.n_cov > 0) {
        dim3 grid((nl.n_cov + 255) / 256, (ncol + kCBP - 1) / kCBP);
        nl_apply<T><<<grid, 256, 0, ctx->stream>>>(out, ldo, ncol, nl.cov_pts, nl.n_cov,
                                                    nl.inv_ptr, nl.inv_ent, nl.inv_loc, nl.ndc,
                                                    nl.nproj, nl.chi_off, nl.alpha_off, chi,
                                                    alpha_s, nl.n_alpha);
        SGPU_HIP(hipGetLastError());
    }
    return SGPU_OK;
}

}  // namespace

int sgpu_launch_vnl(sgpu_context *ctx, int dtype, int ncol, const void *x, long ldx, void *out,
                    long ldo, double scale)
{
    if (dtype == SGPU_F64)
        return launch_vnl_typed<double>(ctx, ncol, (const double *)x, ldx, (double *)out, ldo,
                                        scale, ctx->nl.chi64);
    return launch_vnl_typed<float>(ctx, ncol, (const float *)x, ldx, (float *)out, ldo, scale,
                                   ctx->nl.chi32);
}
