// Hamiltonian apply that allocates nothing: the nonlocal partials live in a
// file-scope device array instead of the context's grow-on-demand scratch,
// the stencil never times tile shapes (that allocates a temporary output) and
// the event timers are skipped.  For callers that cannot call hipMalloc (code
// under stream capture, tight serving loops).
//
// `scratch` below is a device symbol: in host code its name is a host-side
// stand-in, not a device address.  Its address on each GPU is looked up once
// per device with hipGetSymbolAddress (sgpu_noalloc_setup, called by
// sgpu_create with the context's device current) and only that address is
// passed to kernels.
#include <mutex>

#include "sgpu_internal.h"

namespace {

constexpr int kMaxDevices = 64;

// One copy per GPU the library runs on, held until the process ends.  Every
// call on a GPU shares that GPU's copy: calls that overlap on one GPU race.
__device__ __align__(256) unsigned char scratch[SGPU_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void *g_scratch[kMaxDevices];

int scratch_address(int device, void **p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        sgpu_set_error("sgpu_hamiltonian_dev_noalloc: device %d scratch not set up", device);
        return SGPU_ERR_STATE;
    }
    *p = g_scratch[device];
    return SGPU_OK;
}

}  // namespace

int sgpu_noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        sgpu_set_error("sgpu_create: device %d beyond the %d the library supports", device,
                       kMaxDevices);
        return SGPU_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return SGPU_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        sgpu_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device,
                       hipGetErrorString(e));
        return SGPU_ERR_HIP;
    }
    g_scratch[device] = p;
    return SGPU_OK;
}

extern "C" int sgpu_hamiltonian_dev_noalloc(sgpu_context *ctx, double c, const sgpu_block *x,
                                           sgpu_block *Hx)
{
    if (!ctx || !x || !Hx || x->ctx != ctx || Hx->ctx != ctx) {
        sgpu_set_error("sgpu_hamiltonian_dev_noalloc: block missing or from another context");
        return SGPU_ERR_ARG;
    }
    if (x->ncol != Hx->ncol || x->dtype != Hx->dtype) {
        sgpu_set_error("sgpu_hamiltonian_dev_noalloc: blocks differ in shape or type");
        return SGPU_ERR_ARG;
    }
    if (!ctx->has_veff) {
        sgpu_set_error("Hamiltonian apply before sgpu_set_veff");
        return SGPU_ERR_STATE;
    }
    const size_t need = ctx->nl.n_infl > 0 ? sgpu_vnl_scratch_bytes(ctx, x->dtype, x->ncol) : 0;
    if (need > (size_t)SGPU_NOALLOC_SCRATCH_BYTES) {
        sgpu_set_error("sgpu_hamiltonian_dev_noalloc: %d columns need %zu scratch bytes, more "
                       "than the %zu available; apply fewer columns per call",
                       x->ncol, need, (size_t)SGPU_NOALLOC_SCRATCH_BYTES);
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    void *ws = nullptr;
    SGPU_CHECK(scratch_address(ctx->device, &ws));
    const void *veff = x->dtype == SGPU_F64 ? (const void *)ctx->veff64 : (const void *)ctx->veff32;
    // the timers take events from a pool that grows with hipEventCreate
    const bool timing = ctx->timing;
    ctx->timing = false;
    int st = sgpu_launch_stencil(ctx, x->dtype, x->ncol, -0.5, 1.0, c, veff, x->ptr, x->ld,
                                 Hx->ptr, Hx->ld, nullptr, 0, 1.0, 0.0, true, false);
    if (st == SGPU_OK && ctx->nl.n_infl > 0)
        st = sgpu_launch_vnl(ctx, x->dtype, x->ncol, x->ptr, x->ld, Hx->ptr, Hx->ld, 1.0, ws,
                             (size_t)SGPU_NOALLOC_SCRATCH_BYTES);
    ctx->timing = timing;
    return st;
}
