// Elementwise, conversion, RNG fill and reduction kernels for libsparcgpu.so.
#include "sgpu_internal.h"

namespace {

template <typename S, typename D>
__global__ void convert_kernel(const S *__restrict__ src, D *__restrict__ dst, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = (D)src[i];
}

// splitmix64 finaliser: counter-based, so every element is independent of the
// launch shape.
__device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void fill_uniform_kernel(T *__restrict__ x, long nd, long ld, int ncol, uint64_t seed,
                                    double lo, double hi)
{
    long total = nd * ncol;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < total; i += stride) {
        long col = i / nd, row = i - col * nd;
        uint64_t r = mix64(seed ^ mix64((uint64_t)i));
        double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        x[col * ld + row] = (T)(lo + (hi - lo) * u);
    }
}

// Several dot products sum_i a_k[i] b_k[i] in one launch: block partials in
// FP64, second pass sums partials.  Deterministic (fixed reduction tree).
constexpr int kDotThreads = 256;
constexpr int kMaxPairs = 16;

struct DotArgs {
    const double *a[kMaxPairs];
    const double *b[kMaxPairs];
};

__global__ __launch_bounds__(kDotThreads) void dot_partial_kernel(DotArgs args, int npairs, long n,
                                                                  double *__restrict__ partial)
{
    __shared__ double sh[kDotThreads / 64];
    const int pair = blockIdx.y;
    const double *a = args.a[pair];
    const double *b = args.b[pair];
    double s = 0.0;
    for (long i = (long)blockIdx.x * kDotThreads + threadIdx.x; i < n;
         i += (long)gridDim.x * kDotThreads)
        s += a[i] * b[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kDotThreads / 64; w++) t += sh[w];
        partial[pair * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void dot_final_kernel(const double *__restrict__ partial, int nblk, int npairs,
                                 double *__restrict__ out)
{
    __shared__ double sh[256];
    const int pair = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += partial[pair * nblk + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[pair] = sh[0];
}

}  // namespace

int sgpu_convert(sgpu_context *ctx, const void *src, int st, void *dst, int dt, long n)
{
    if (n <= 0) return SGPU_OK;
    const int threads = 256;
    long blocks = std::min<long>((n + threads - 1) / threads, 65536);
    if (st == SGPU_F64 && dt == SGPU_F32)
        convert_kernel<<<blocks, threads, 0, ctx->stream>>>((const double *)src, (float *)dst, n);
    else if (st == SGPU_F32 && dt == SGPU_F64)
        convert_kernel<<<blocks, threads, 0, ctx->stream>>>((const float *)src, (double *)dst, n);
    else if (st == SGPU_F64 && dt == SGPU_F64)
        convert_kernel<<<blocks, threads, 0, ctx->stream>>>((const double *)src, (double *)dst, n);
    else
        convert_kernel<<<blocks, threads, 0, ctx->stream>>>((const float *)src, (float *)dst, n);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

int sgpu_dot_many(sgpu_context *ctx, int n, int npairs, const double *const *a,
                  const double *const *b, double *host_out)
{
    if (npairs <= 0 || npairs > kMaxPairs || n < 0) {
        sgpu_set_error("sgpu_dot_many: bad pair count %d", npairs);
        return SGPU_ERR_ARG;
    }
    DotArgs args;
    for (int k = 0; k < npairs; k++) { args.a[k] = a[k]; args.b[k] = b[k]; }
    int nblk = (int)std::min<long>(std::max<long>((n + kDotThreads - 1) / kDotThreads, 1), 1024);
    double *red = nullptr;
    SGPU_CHECK(sgpu_reduce_buf(ctx, (size_t)npairs * nblk + kMaxPairs, &red));
    double *out = red + (size_t)npairs * nblk;
    dot_partial_kernel<<<dim3(nblk, npairs), kDotThreads, 0, ctx->stream>>>(args, npairs, n, red);
    dot_final_kernel<<<npairs, 256, 0, ctx->stream>>>(red, nblk, npairs, out);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipMemcpyAsync(ctx->host_red, out, npairs * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < npairs; k++) host_out[k] = ctx->host_red[k];
    return SGPU_OK;
}

extern "C" int sgpu_block_fill_uniform(sgpu_block *blk, unsigned long long seed, double lo,
                                       double hi)
{
    if (!blk || !blk->ctx) {
        sgpu_set_error("sgpu_block_fill_uniform: null or orphaned block");
        return SGPU_ERR_ARG;
    }
    sgpu_context *ctx = blk->ctx;
    SGPU_HIP(hipSetDevice(ctx->device));
    long total = ctx->nd * blk->ncol;
    long blocks = std::min<long>((total + 255) / 256, 65536);
    if (blk->dtype == SGPU_F64)
        fill_uniform_kernel<double><<<blocks, 256, 0, ctx->stream>>>(
            (double *)blk->ptr, ctx->nd, blk->ld, blk->ncol, seed, lo, hi);
    else
        fill_uniform_kernel<float><<<blocks, 256, 0, ctx->stream>>>(
            (float *)blk->ptr, ctx->nd, blk->ld, blk->ncol, seed, lo, hi);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}
