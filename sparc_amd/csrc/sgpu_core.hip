// Context, device blocks, operator data upload and timers for libsparcgpu.so.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sgpu_internal.h"

static thread_local char g_err[1024] = "";

void sgpu_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *sgpu_last_error(void) { return g_err; }

size_t sgpu_dtype_size(int dtype) { return dtype == SGPU_F32 ? sizeof(float) : sizeof(double); }

// n!^2 / ((n-k)! (n+k)!) -- the factorial ratio of the central-difference
// second-derivative weights (SPARC computes it as fract(n,k), src/tools.c:725).
static double fd_ratio(int n, int k)
{
    double nr = 1.0, dr = 1.0;
    for (int i = n - k + 1; i <= n; i++) nr *= i;
    for (int i = n + 1; i <= n + k; i++) dr *= i;
    return nr / dr;
}

extern "C" int sgpu_create(sgpu_context **out, int device, int nx, int ny, int nz, double hx,
                           double hy, double hz, int bcx, int bcy, int bcz, int order)
{
    if (!out) { sgpu_set_error("sgpu_create: null output"); return SGPU_ERR_ARG; }
    *out = nullptr;
    const int fdn = order / 2;
    if (order % 2 || fdn < 1 || fdn > SGPU_MAX_FDN) {
        sgpu_set_error("sgpu_create: FD order %d unsupported (even, 2..16)", order);
        return SGPU_ERR_ARG;
    }
    if (nx < fdn || ny < fdn || nz < fdn) {
        // the reference's single-rank periodic halo copy needs >= fdn points per axis
        sgpu_set_error("sgpu_create: grid %dx%dx%d smaller than stencil radius %d", nx, ny, nz, fdn);
        return SGPU_ERR_ARG;
    }
    if (!(hx > 0 && hy > 0 && hz > 0) || (bcx | bcy | bcz) & ~1) {
        sgpu_set_error("sgpu_create: bad mesh or boundary condition");
        return SGPU_ERR_ARG;
    }
    if ((long)nx * ny * nz > (1L << 31) - 1) {
        sgpu_set_error("sgpu_create: grid too large for 32-bit point indices");
        return SGPU_ERR_ARG;
    }
    int ndev = 0;
    SGPU_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        sgpu_set_error("sgpu_create: device %d not present (%d devices)", device, ndev);
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(device));
    SGPU_CHECK(sgpu_noalloc_setup(device));
    sgpu_context *ctx = new sgpu_context();
    ctx->device = device;
    ctx->nx = nx; ctx->ny = ny; ctx->nz = nz;
    ctx->nd = (long)nx * ny * nz;
    ctx->ld = (ctx->nd + 63) / 64 * 64;
    ctx->hx = hx; ctx->hy = hy; ctx->hz = hz;
    ctx->bcx = bcx; ctx->bcy = bcy; ctx->bcz = bcz;
    ctx->fdn = fdn;
    if (const char *t = getenv("SPARC_AMD_STENCIL_TILE")) ctx->stencil_cfg = atoi(t);
    if (const char *t = getenv("SPARC_AMD_NL_CHAIN")) ctx->nl_chain = atoi(t) != 0;
    if (const char *t = getenv("SPARC_AMD_NL_CHAIN_BRICK")) ctx->nl_chain_brick = atoi(t) != 0;
    if (const char *t = getenv("SPARC_AMD_LDS_TRIDIAG")) ctx->lds_tridiag = atoi(t) != 0;
    if (const char *t = getenv("SPARC_AMD_COOP_TRIDIAG")) ctx->coop_tridiag = std::max(0, atoi(t));
    if (const char *t = getenv("SPARC_AMD_COOP_THREADS")) ctx->coop_threads = atoi(t) == 512 ? 512 : 256;
    if (const char *t = getenv("SPARC_AMD_CHOL")) ctx->own_chol = strcmp(t, "rocsolver") != 0;
    if (const char *t = getenv("SPARC_AMD_ORMTR")) ctx->own_ormtr = strcmp(t, "rocsolver") != 0;
    if (const char *t = getenv("SPARC_AMD_STEDC")) ctx->own_stedc = strcmp(t, "rocsolver") != 0;
    if (const char *t = getenv("SPARC_AMD_GRAM3_PAIR")) ctx->gram3_pair = atoi(t) != 0;
    if (const char *t = getenv("SPARC_AMD_TDC_LEAF")) {
        const int k = atoi(t);
        if (k == 1 || k == 8 || k == 16 || k == 32) ctx->tdc_leaf = k;
    }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
            ncu > 0)
            ctx->n_cu = ncu;
    }
    // second-derivative weights: w_p = 2 (-1)^{p+1} ratio(fdn,p) / p^2,
    // w_0 = -2 sum 1/p^2, divided by h^2 per axis
    double w[SGPU_MAX_FDN + 1];
    w[0] = 0.0;
    for (int p = 1; p <= fdn; p++) {
        w[0] -= 2.0 / (p * p);
        w[p] = (p % 2 ? 2.0 : -2.0) * fd_ratio(fdn, p) / (p * p);
    }
    for (int p = 0; p <= fdn; p++) {
        ctx->d2x[p] = w[p] / (hx * hx);
        ctx->d2y[p] = w[p] / (hy * hy);
        ctx->d2z[p] = w[p] / (hz * hz);
    }
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        sgpu_set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete ctx;
        return SGPU_ERR_HIP;
    }
    if (rocblas_create_handle(&ctx->blas) != rocblas_status_success ||
        rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success) {
        sgpu_set_error("rocblas handle creation failed");
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return SGPU_ERR_HIP;
    }
    if (hipHostMalloc((void **)&ctx->host_red, 4096 * sizeof(double), hipHostMallocDefault) !=
        hipSuccess) {
        sgpu_set_error("pinned host buffer allocation failed");
        rocblas_destroy_handle(ctx->blas);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return SGPU_ERR_NOMEM;
    }
    *out = ctx;
    return SGPU_OK;
}

extern "C" int sgpu_destroy(sgpu_context *ctx)
{
    if (!ctx) return SGPU_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (sgpu_block *b : ctx->blocks) {
        if (b->ptr) (void)hipFree(b->ptr);
        b->ptr = nullptr;
        b->ctx = nullptr;  // orphaned: sgpu_block_destroy only frees the handle
    }
    ctx->blocks.clear();
    for (sgpu_unet *u : ctx->unets) sgpu_unet_orphan(u);
    ctx->unets.clear();
    sgpu_free_nonlocal(ctx->nl);
    if (ctx->veff64) (void)hipFree(ctx->veff64);
    if (ctx->veff32) (void)hipFree(ctx->veff32);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->red_buf) (void)hipFree(ctx->red_buf);
    if (ctx->aar_mem) (void)hipFree(ctx->aar_mem);
    if (ctx->host_red) (void)hipHostFree(ctx->host_red);
    for (auto &pe : ctx->pending) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    for (hipEvent_t ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    (void)hipGetLastError();  // do not leave an ignored teardown error behind
    return SGPU_OK;
}

extern "C" int sgpu_get_stencil_coeffs(const sgpu_context *ctx, double *d2x, double *d2y,
                                       double *d2z)
{
    if (!ctx) return SGPU_ERR_ARG;
    for (int p = 0; p <= ctx->fdn; p++) {
        if (d2x) d2x[p] = ctx->d2x[p];
        if (d2y) d2y[p] = ctx->d2y[p];
        if (d2z) d2z[p] = ctx->d2z[p];
    }
    return SGPU_OK;
}

extern "C" int sgpu_sync(sgpu_context *ctx)
{
    if (!ctx) return SGPU_ERR_ARG;
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_scratch(sgpu_context *ctx, size_t bytes, void **p)
{
    if (bytes > ctx->scratch_bytes) {
        SGPU_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->scratch) SGPU_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        SGPU_HIP(hipMalloc(&ctx->scratch, bytes));
        ctx->scratch_bytes = bytes;
    }
    *p = ctx->scratch;
    return SGPU_OK;
}

int sgpu_reduce_buf(sgpu_context *ctx, size_t elems, double **p)
{
    if (elems > ctx->red_elems) {
        SGPU_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->red_buf) SGPU_HIP(hipFree(ctx->red_buf));
        ctx->red_buf = nullptr;
        ctx->red_elems = 0;
        SGPU_HIP(hipMalloc((void **)&ctx->red_buf, elems * sizeof(double)));
        ctx->red_elems = elems;
    }
    *p = ctx->red_buf;
    return SGPU_OK;
}

// ---- timers ----

static int get_event(sgpu_context *ctx, hipEvent_t *ev)
{
    if (!ctx->ev_pool.empty()) {
        *ev = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return SGPU_OK;
    }
    SGPU_HIP(hipEventCreate(ev));
    return SGPU_OK;
}

int sgpu_tbegin(sgpu_context *ctx, sgpu_tmark *m)
{
    m->a = m->b = nullptr;
    if (!ctx->timing) return SGPU_OK;
    SGPU_CHECK(get_event(ctx, &m->a));
    SGPU_CHECK(get_event(ctx, &m->b));
    SGPU_HIP(hipEventRecord(m->a, ctx->stream));
    return SGPU_OK;
}

// Records the end event without waiting; the pair is resolved lazily by
// sgpu_timer_get so timed regions are not serialised by the timers.
int sgpu_tend(sgpu_context *ctx, sgpu_tmark *m, const char *name)
{
    if (!ctx->timing || !m->a) return SGPU_OK;
    SGPU_HIP(hipEventRecord(m->b, ctx->stream));
    ctx->pending.push_back({m->a, m->b, std::string(name)});
    m->a = m->b = nullptr;
    if (ctx->pending.size() > 4096) return sgpu_timer_resolve(ctx);
    return SGPU_OK;
}

int sgpu_timer_resolve(sgpu_context *ctx)
{
    if (ctx->pending.empty()) return SGPU_OK;
    SGPU_HIP(hipEventSynchronize(ctx->pending.back().b));
    for (auto &pe : ctx->pending) {
        float ms = 0.f;
        SGPU_HIP(hipEventElapsedTime(&ms, pe.a, pe.b));
        sgpu_timer_acc &t = ctx->timers[pe.name];
        t.total_ms += ms;
        t.count += 1;
        ctx->ev_pool.push_back(pe.a);
        ctx->ev_pool.push_back(pe.b);
    }
    ctx->pending.clear();
    return SGPU_OK;
}

extern "C" int sgpu_timer_get(sgpu_context *ctx, const char *name, double *total_ms, int *count)
{
    if (!ctx || !name) return SGPU_ERR_ARG;
    SGPU_CHECK(sgpu_timer_resolve(ctx));
    auto it = ctx->timers.find(name);
    if (total_ms) *total_ms = it == ctx->timers.end() ? 0.0 : it->second.total_ms;
    if (count) *count = it == ctx->timers.end() ? 0 : it->second.count;
    return SGPU_OK;
}

extern "C" int sgpu_timer_reset(sgpu_context *ctx)
{
    if (!ctx) return SGPU_ERR_ARG;
    SGPU_CHECK(sgpu_timer_resolve(ctx));
    ctx->timers.clear();
    return SGPU_OK;
}

extern "C" int sgpu_timer_enable(sgpu_context *ctx, int on)
{
    if (!ctx) return SGPU_ERR_ARG;
    ctx->timing = on != 0;
    return SGPU_OK;
}

// ---- operator data ----

extern "C" int sgpu_set_veff(sgpu_context *ctx, const double *veff)
{
    if (!ctx || !veff) { sgpu_set_error("sgpu_set_veff: null argument"); return SGPU_ERR_ARG; }
    SGPU_HIP(hipSetDevice(ctx->device));
    if (!ctx->veff64) SGPU_HIP(hipMalloc((void **)&ctx->veff64, ctx->nd * sizeof(double)));
    if (!ctx->veff32) SGPU_HIP(hipMalloc((void **)&ctx->veff32, ctx->nd * sizeof(float)));
    SGPU_HIP(hipMemcpyAsync(ctx->veff64, veff, ctx->nd * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream));
    SGPU_CHECK(sgpu_convert(ctx, ctx->veff64, SGPU_F64, ctx->veff32, SGPU_F32, ctx->nd));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    ctx->has_veff = true;
    return SGPU_OK;
}

// ---- blocks ----

extern "C" int sgpu_block_create(sgpu_context *ctx, int ncol, int dtype, sgpu_block **blk)
{
    if (!ctx || !blk || ncol <= 0 || (dtype != SGPU_F64 && dtype != SGPU_F32)) {
        sgpu_set_error("sgpu_block_create: bad arguments");
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    sgpu_block *b = new sgpu_block();
    b->ctx = ctx;
    b->ncol = ncol;
    b->dtype = dtype;
    b->ld = ctx->ld;
    size_t bytes = (size_t)ncol * b->ld * sgpu_dtype_size(dtype);
    hipError_t e = hipMalloc(&b->ptr, bytes);
    if (e != hipSuccess) {
        sgpu_set_error("sgpu_block_create: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        delete b;
        return SGPU_ERR_NOMEM;
    }
    e = hipMemsetAsync(b->ptr, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        sgpu_set_error("sgpu_block_create: clear failed: %s", hipGetErrorString(e));
        (void)hipFree(b->ptr);
        delete b;
        return SGPU_ERR_HIP;
    }
    ctx->blocks.push_back(b);
    *blk = b;
    return SGPU_OK;
}

extern "C" int sgpu_block_destroy(sgpu_block *blk)
{
    if (!blk) return SGPU_OK;
    sgpu_context *ctx = blk->ctx;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (blk->ptr) (void)hipFree(blk->ptr);
        for (size_t i = 0; i < ctx->blocks.size(); i++)
            if (ctx->blocks[i] == blk) {
                ctx->blocks.erase(ctx->blocks.begin() + i);
                break;
            }
        (void)hipGetLastError();
    }
    delete blk;
    return SGPU_OK;
}

extern "C" int sgpu_block_ncol(const sgpu_block *blk) { return blk ? blk->ncol : 0; }
extern "C" int sgpu_block_dtype(const sgpu_block *blk) { return blk ? blk->dtype : -1; }
extern "C" long sgpu_block_ld(const sgpu_block *blk) { return blk ? blk->ld : 0; }

extern "C" int sgpu_block_upload(sgpu_block *blk, const double *host, long ld, int ncol)
{
    if (!blk || !blk->ctx || !host || ncol < 0 || ncol > blk->ncol || ld < blk->ctx->nd) {
        sgpu_set_error("sgpu_block_upload: bad arguments or orphaned block");
        return SGPU_ERR_ARG;
    }
    sgpu_context *ctx = blk->ctx;
    SGPU_HIP(hipSetDevice(ctx->device));
    if (ncol == 0) return SGPU_OK;
    if (blk->dtype == SGPU_F64) {
        SGPU_HIP(hipMemcpy2DAsync(blk->ptr, blk->ld * sizeof(double), host, ld * sizeof(double),
                                  ctx->nd * sizeof(double), ncol, hipMemcpyHostToDevice,
                                  ctx->stream));
    } else {
        // stage through FP64 scratch in column chunks, then convert on device
        const long chunk = std::max(1L, (256L << 20) / (ctx->ld * (long)sizeof(double)));
        void *stage = nullptr;
        SGPU_CHECK(sgpu_scratch(ctx, (size_t)std::min<long>(chunk, ncol) * ctx->ld * sizeof(double),
                                &stage));
        for (long c0 = 0; c0 < ncol; c0 += chunk) {
            long nc = std::min<long>(chunk, ncol - c0);
            SGPU_HIP(hipMemcpy2DAsync(stage, ctx->ld * sizeof(double), host + c0 * ld,
                                      ld * sizeof(double), ctx->nd * sizeof(double), nc,
                                      hipMemcpyHostToDevice, ctx->stream));
            SGPU_CHECK(sgpu_convert(ctx, stage, SGPU_F64, (float *)blk->ptr + c0 * blk->ld,
                                    SGPU_F32, nc * ctx->ld));
        }
    }
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

extern "C" int sgpu_block_download(const sgpu_block *blk, double *host, long ld, int ncol)
{
    if (!blk || !blk->ctx || !host || ncol < 0 || ncol > blk->ncol || ld < blk->ctx->nd) {
        sgpu_set_error("sgpu_block_download: bad arguments or orphaned block");
        return SGPU_ERR_ARG;
    }
    sgpu_context *ctx = blk->ctx;
    SGPU_HIP(hipSetDevice(ctx->device));
    if (ncol == 0) return SGPU_OK;
    if (blk->dtype == SGPU_F64) {
        SGPU_HIP(hipMemcpy2DAsync(host, ld * sizeof(double), blk->ptr, blk->ld * sizeof(double),
                                  ctx->nd * sizeof(double), ncol, hipMemcpyDeviceToHost,
                                  ctx->stream));
    } else {
        const long chunk = std::max(1L, (256L << 20) / (ctx->ld * (long)sizeof(double)));
        void *stage = nullptr;
        SGPU_CHECK(sgpu_scratch(ctx, (size_t)std::min<long>(chunk, ncol) * ctx->ld * sizeof(double),
                                &stage));
        for (long c0 = 0; c0 < ncol; c0 += chunk) {
            long nc = std::min<long>(chunk, ncol - c0);
            SGPU_CHECK(sgpu_convert(ctx, (const float *)blk->ptr + c0 * blk->ld, SGPU_F32, stage,
                                    SGPU_F64, nc * ctx->ld));
            SGPU_HIP(hipMemcpy2DAsync(host + c0 * ld, ld * sizeof(double), stage,
                                      ctx->ld * sizeof(double), ctx->nd * sizeof(double), nc,
                                      hipMemcpyDeviceToHost, ctx->stream));
            SGPU_HIP(hipStreamSynchronize(ctx->stream));
        }
    }
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

extern "C" int sgpu_block_copy(const sgpu_block *src, sgpu_block *dst)
{
    if (!src || !dst || !src->ctx || src->ctx != dst->ctx || src->ncol != dst->ncol) {
        sgpu_set_error("sgpu_block_copy: shape mismatch");
        return SGPU_ERR_ARG;
    }
    sgpu_context *ctx = src->ctx;
    SGPU_HIP(hipSetDevice(ctx->device));
    long n = (long)src->ncol * src->ld;
    if (src->dtype == dst->dtype)
        SGPU_HIP(hipMemcpyAsync(dst->ptr, src->ptr, n * sgpu_dtype_size(src->dtype),
                                hipMemcpyDeviceToDevice, ctx->stream));
    else
        SGPU_CHECK(sgpu_convert(ctx, src->ptr, src->dtype, dst->ptr, dst->dtype, n));
    return SGPU_OK;
}
