// Learned preconditioner for the AAR Poisson solve: a linear 3-level 3-D
// convolutional U-Net with BF16 activations / gradients, FP32 accumulation
// and FP32 master weights, trained on the device on (residual, error) pairs
// that Jacobi-AAR solves produce, and applied as the reference's precond_fun
// (AAR, src/linearSolver.c:37-43; Jacobi_preconditioner,
// src/electrostatics.c:1682-1699, is the U = 0 case):
//
//     f = (r + U(r)) / |D0|,   D0 = D2x[0] + D2y[0] + D2z[0]
//
// Levels L0 = grid, L1 = grid / 2, L2 = grid / 4, C = 16 channels, periodic
// 3x3x3 kernels.  On a grid whose dimensions are not all multiples of 4 the
// network runs on the internal grid M (each dimension rounded up to a
// multiple of 4, same periodic cell): f = (r + I_{M->N} U(I_{N->M} r)) / |D0|
// with periodic trilinear interpolation I (interp_trilinear, prolong_out,
// and its transpose interp_adjoint in the backward pass).  Layers:
//   R1 conv/2 1->C (L0->L1), B1 conv C->C, R2 conv/2 (L1->L2), B2, M2 conv,
//   P2 transposed conv x2 (L2->L1) + skip from B1, C1 conv, P1 transposed
//   conv x2 C->1 (L1->L0).
// No activations: the preconditioner is a fixed linear operator, as AAR's
// Anderson mixing assumes.  Activations are stored channels-last
// ([voxel][C], 32 bytes per voxel) in BF16; products are BF16 x BF16 (weights
// rounded to BF16 when staged in LDS) accumulated in FP32.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sgpu_internal.h"

namespace {

constexpr int kC = 16;
constexpr int kLayers = 8;
// weights per layer: [cout][cin][27]
constexpr int kLayerCin[kLayers] = {1, kC, kC, kC, kC, kC, kC, kC};
constexpr int kLayerCout[kLayers] = {kC, kC, kC, kC, kC, kC, kC, 1};

__host__ __device__ inline int layer_size(int l) { return kLayerCin[l] * kLayerCout[l] * 27; }

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bfround(float f) { return bf2f(f2bf(f)); }

struct Dims {
    int nx, ny, nz;
    __host__ __device__ long vox(int x, int y, int z) const { return ((long)z * ny + y) * nx + x; }
    __host__ __device__ long n() const { return (long)nx * ny * nz; }
};

__device__ __forceinline__ int wrap(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Stage W[co][ci][27] (FP32) into LDS as BF16-rounded floats laid out
// [k][ci][co] so that the inner co loop reads contiguous words.
__device__ void stage_weights(const float *__restrict__ W, int cin, int cout, float *lds)
{
    const int total = cin * cout * 27;
#pragma unroll 8
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int k = t % 27, ci = (t / 27) % cin, co = t / (27 * cin);
        lds[(k * cin + ci) * cout + co] = bfround(W[t]);
    }
    __syncthreads();
}

__device__ __forceinline__ void load16(const uint16_t *__restrict__ p, float *v)
{
    const uint4 a = *reinterpret_cast<const uint4 *>(p);
    const uint4 b = *reinterpret_cast<const uint4 *>(p + 8);
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        v[2 * i] = __uint_as_float(w[i] << 16);
        v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
    }
}

__device__ __forceinline__ void store16(uint16_t *__restrict__ p, const float *v)
{
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = (uint32_t)f2bf(v[2 * i]) | ((uint32_t)f2bf(v[2 * i + 1]) << 16);
    *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    *reinterpret_cast<uint4 *>(p + 8) = make_uint4(w[4], w[5], w[6], w[7]);
}

// C -> C convolution, stride S (1 or 2): out[co][y] = sum W[co][ci][k] in[ci][S y + k - 1]
// (periodic), one thread per output voxel; optional skip added to the output.
template <int S>
__global__ __launch_bounds__(256) void conv_cc(const uint16_t *__restrict__ in, Dims di,
                                               uint16_t *__restrict__ out, Dims dout,
                                               const float *__restrict__ W,
                                               const uint16_t *__restrict__ skip)
{
    __shared__ float wl[27 * kC * kC];
    stage_weights(W, kC, kC, wl);
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dout.n()) return;
    const int vq = (int)v / dout.nx;  // 32-bit: a level holds < 2^31 voxels
    const int ox = (int)v - vq * dout.nx, oy = vq % dout.ny, oz = vq / dout.ny;
    float acc[kC];
#pragma unroll
    for (int c = 0; c < kC; c++) acc[c] = 0.f;
    for (int k = 0; k < 27; k++) {
        const int ix = wrap(S * ox + k % 3 - 1, di.nx);
        const int iy = wrap(S * oy + (k / 3) % 3 - 1, di.ny);
        const int iz = wrap(S * oz + k / 9 - 1, di.nz);
        float xv[kC];
        load16(in + di.vox(ix, iy, iz) * kC, xv);
        const float *wk = wl + k * kC * kC;
#pragma unroll
        for (int ci = 0; ci < kC; ci++)
#pragma unroll
            for (int co = 0; co < kC; co++) acc[co] += wk[ci * kC + co] * xv[ci];
    }
    if (skip) {
        float sv[kC];
        load16(skip + v * kC, sv);
#pragma unroll
        for (int c = 0; c < kC; c++) acc[c] += sv[c];
    }
    store16(out + v * kC, acc);
}

// Transposed stride-2 convolution C -> C: out[co][x] = sum over (y, k) with
// 2y + k - 1 = x of W[co][ci][k] in[ci][y]; per axis an even x takes k = 1
// (y = x/2) and an odd x takes k = 0 (y = (x+1)/2) and k = 2 (y = (x-1)/2).
__global__ __launch_bounds__(256) void convT_cc(const uint16_t *__restrict__ in, Dims di,
                                                uint16_t *__restrict__ out, Dims dout,
                                                const float *__restrict__ W,
                                                const uint16_t *__restrict__ skip)
{
    __shared__ float wl[27 * kC * kC];
    stage_weights(W, kC, kC, wl);
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dout.n()) return;
    const int vq = (int)v / dout.nx;  // 32-bit: a level holds < 2^31 voxels
    const int fx = (int)v - vq * dout.nx, fy = vq % dout.ny, fz = vq / dout.ny;
    int kxs[2], cxs[2], kys[2], cys[2], kzs[2], czs[2];
    const int nkx = (fx & 1) ? 2 : 1, nky = (fy & 1) ? 2 : 1, nkz = (fz & 1) ? 2 : 1;
    if (fx & 1) { kxs[0] = 0; cxs[0] = wrap((fx + 1) / 2, di.nx); kxs[1] = 2; cxs[1] = (fx - 1) / 2; }
    else { kxs[0] = 1; cxs[0] = fx / 2; }
    if (fy & 1) { kys[0] = 0; cys[0] = wrap((fy + 1) / 2, di.ny); kys[1] = 2; cys[1] = (fy - 1) / 2; }
    else { kys[0] = 1; cys[0] = fy / 2; }
    if (fz & 1) { kzs[0] = 0; czs[0] = wrap((fz + 1) / 2, di.nz); kzs[1] = 2; czs[1] = (fz - 1) / 2; }
    else { kzs[0] = 1; czs[0] = fz / 2; }
    float acc[kC];
#pragma unroll
    for (int c = 0; c < kC; c++) acc[c] = 0.f;
    for (int a = 0; a < nkz; a++)
        for (int b = 0; b < nky; b++)
            for (int c = 0; c < nkx; c++) {
                const int k = kzs[a] * 9 + kys[b] * 3 + kxs[c];
                float xv[kC];
                load16(in + di.vox(cxs[c], cys[b], czs[a]) * kC, xv);
                const float *wk = wl + k * kC * kC;
#pragma unroll
                for (int ci = 0; ci < kC; ci++)
#pragma unroll
                    for (int co = 0; co < kC; co++) acc[co] += wk[ci * kC + co] * xv[ci];
            }
    if (skip) {
        float sv[kC];
        load16(skip + v * kC, sv);
#pragma unroll
        for (int c = 0; c < kC; c++) acc[c] += sv[c];
    }
    store16(out + v * kC, acc);
}

// ---- the same two C -> C layers on matrix cores -------------------------
//
// v_mfma_f32_16x16x16bf16_1k per tap: A = 16 output voxels along x x 16 input
// channels (one 8-byte channels-last load per lane), B = the tap's 16 x 16
// weights from LDS, D = 16 voxels x 16 output channels in FP32.  Operand
// layout (checked on the device): A(i,k) in lane i + 16(k/4) element k%4,
// B(k,n) in lane n + 16(k/4) element k%4, D(i,n) in lane n + 16(i/4)
// register i%4.  One wave per 16-voxel tile, four tiles per block.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kTilesPerWave = 8;  // 16-voxel tiles per wave: amortises the weight staging

// C -> C weights packed once per pass as BF16 in the layout the MFMA B
// operand reads, [k][co][ci] (pack_bf16), so that a block stages them with
// 864 coalesced 16-byte copies.  (Converting W[co][ci][27] while staging
// scattered 2-byte LDS stores 512 bytes apart: ~26 bank-conflict cycles per
// LDS instruction, the largest cost of every layer.)
constexpr int kWcc = 27 * kC * kC;  // weights of one C -> C layer

__global__ void pack_bf16(const float *__restrict__ W, int nlayer, uint16_t *__restrict__ Wb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nlayer * kWcc) return;
    const int l = e / kWcc, u = e - l * kWcc;
    const int k = u / (kC * kC), co = (u / kC) % kC, ci = u % kC;
    Wb[e] = f2bf(W[(long)l * kWcc + (co * kC + ci) * 27 + k]);
}

// 256-thread blocks: 864 16-byte pieces, consecutive threads on consecutive
// pieces (coalesced reads, conflict-free LDS writes)
__device__ void stage_packed(const uint16_t *__restrict__ Wb, uint16_t *lds)
{
    constexpr int kPieces = kWcc / 8;  // 864 = 4 x 216
    if (threadIdx.x < kPieces / 4) {
        const uint4 *src = reinterpret_cast<const uint4 *>(Wb);
        uint4 *dst = reinterpret_cast<uint4 *>(lds);
        uint4 v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = src[threadIdx.x + j * (kPieces / 4)];
#pragma unroll
        for (int j = 0; j < 4; j++) dst[threadIdx.x + j * (kPieces / 4)] = v[j];
    }
    __syncthreads();
}

__device__ __forceinline__ void store_tile(uint16_t *__restrict__ out, const uint16_t *__restrict__ skip,
                                           long vox, int co, float val)
{
    if (skip) val += bf2f(skip[vox * kC + co]);
    out[vox * kC + co] = f2bf(val);
}

template <int S>
__global__ __launch_bounds__(256) void conv_cc_mfma(const uint16_t *__restrict__ in, Dims di,
                                                    uint16_t *__restrict__ out, Dims dout,
                                                    const uint16_t *__restrict__ Wb,
                                                    const uint16_t *__restrict__ skip, int tpw)
{
    __shared__ __align__(16) uint16_t wl[27 * kC * kC];
    stage_packed(Wb, wl);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int txn = (dout.nx + 15) / 16;
    const long ntiles = (long)txn * dout.ny * dout.nz;
    for (int t = 0; t < tpw; t++) {
        const long tile = ((long)blockIdx.x * tpw + t) * 4 + wave;
        if (tile >= ntiles) return;
        const int tx = (int)(tile % txn);
        const long row = tile / txn;
        const int oy = (int)(row % dout.ny), oz = (int)(row / dout.ny);
        const int ox = min(tx * 16 + li, dout.nx - 1);  // A row voxel of this lane (clamped)
        // all 27 tap loads issued before the dependent MFMA chain
        int xs[3], ys[3], zs[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            xs[d] = wrap(S * ox + d - 1, di.nx);
            ys[d] = wrap(S * oy + d - 1, di.ny);
            zs[d] = wrap(S * oz + d - 1, di.nz);
        }
        bf16x4 av[27];
#pragma unroll
        for (int k = 0; k < 27; k++)
            av[k] = *reinterpret_cast<const bf16x4 *>(in + di.vox(xs[k % 3], ys[(k / 3) % 3], zs[k / 9]) * kC + 4 * g);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 27; k++) {
            const bf16x4 b = *reinterpret_cast<const bf16x4 *>(wl + (k * kC + li) * kC + 4 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[k], b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int x = tx * 16 + 4 * g + j;
            if (x < dout.nx) store_tile(out, skip, dout.vox(x, oy, oz), li, acc[j]);
        }
    }
}

// Stride-1 C -> C layer with the block's input staged in LDS: the block
// computes 16 x 4 x 4 output voxels (16 tiles, 4 per wave) from an
// 18 x 6 x 6-voxel input box loaded once (coalesced 16-byte pieces of the
// channels-last rows), instead of each tile fetching its 27 taps through L2
// (27 x the layer's input in L2 traffic).  Per tile the same 27 MFMAs with
// the same operands as conv_cc_mfma, so the output is bitwise the same.
// Staged coordinates are taken modulo the grid (periodic); lanes and rows
// past the grid edge compute on wrapped data and are not stored.
constexpr int kLX = 16, kLY = 4, kLZ = 4;
constexpr int kLBX = kLX + 2, kLBY = kLY + 2, kLBZ = kLZ + 2;

__device__ __forceinline__ int pmod(int i, int n)
{
    const int r = i % n;
    return r < 0 ? r + n : r;
}

__global__ __launch_bounds__(256) void conv_cc_lds(const uint16_t *__restrict__ in, Dims d,
                                                   uint16_t *__restrict__ out,
                                                   const uint16_t *__restrict__ Wb,
                                                   const uint16_t *__restrict__ skip)
{
    __shared__ __align__(16) uint16_t wl[27 * kC * kC];
    // the input box, padded to a whole number of 16-byte pieces per thread
    constexpr int kPieces = kLBZ * kLBY * kLBX * 2, kPer = (kPieces + 255) / 256;
    __shared__ __align__(16) uint16_t xb[kPer * 256 * 8];
    const int bxn = (d.nx + kLX - 1) / kLX, byn = (d.ny + kLY - 1) / kLY;
    const int bx = blockIdx.x % bxn, by = (blockIdx.x / bxn) % byn, bz = blockIdx.x / (bxn * byn);
    const int x0 = bx * kLX, y0 = by * kLY, z0 = bz * kLZ;
    // input box: voxel v = (zz * kLBY + yy) * kLBX + xx, two 16-byte halves
    // each; every load of a thread is issued before its first LDS store.
    // Loads and stores are unconditional (clamped source, padded box): a
    // guarded store let the compiler sink each load into its branch and wait
    // for it there, one memory latency per piece.
    // Swizzle: the 8-byte channel group q of box voxel xx sits in slot
    // q ^ ((xx >> 2) & 3), so that 16 lanes reading group g of 16
    // consecutive voxels hit 16 different bank pairs (unswizzled, the 32-byte
    // voxel stride put them on 8 banks: a 4-way conflict on every A read).
    uint4 pv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int c = min((int)threadIdx.x + 256 * j, kPieces - 1);
        const int v = c >> 1, h = c & 1;
        const int xx = v % kLBX, yy = (v / kLBX) % kLBY, zz = v / (kLBX * kLBY);
        const long gv =
            d.vox(pmod(x0 + xx - 1, d.nx), pmod(y0 + yy - 1, d.ny), pmod(z0 + zz - 1, d.nz));
        pv[j] = *reinterpret_cast<const uint4 *>(in + gv * kC + 8 * h);
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int c = (int)threadIdx.x + 256 * j;  // < kPer * 256: the box is padded
        const int v = c >> 1, h = c & 1;
        const int sw = ((v % kLBX) >> 2) & 3;
        uint4 val = pv[j];
        if (sw & 1) val = uint4{val.z, val.w, val.x, val.y};  // swap the two 8-byte groups
        *reinterpret_cast<uint4 *>(xb + (2 * v + (h ^ (sw >> 1))) * 8) = val;
    }
    stage_packed(Wb, wl);  // ends with a barrier
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    // B operands (the 27 taps' weights) read once per wave; per tile the 27
    // A reads are issued before the MFMA chain
    bf16x4 bw[27];
#pragma unroll
    for (int k = 0; k < 27; k++) bw[k] = *reinterpret_cast<const bf16x4 *>(wl + (k * kC + li) * kC + 4 * g);
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
        const int row = wave * 4 + t;
        const int ty = row % kLY, tz = row / kLY;
        const int oy = y0 + ty, oz = z0 + tz;
        bf16x4 av[27];
#pragma unroll
        for (int k = 0; k < 27; k++) {
            const int dx = k % 3, dy = (k / 3) % 3, dz = k / 9;
            const int xx = li + dx;
            av[k] = *reinterpret_cast<const bf16x4 *>(
                xb + (((tz + dz) * kLBY + ty + dy) * kLBX + xx) * kC + 4 * (g ^ ((xx >> 2) & 3)));
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 27; k++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[k], bw[k], acc, 0, 0, 0);
        if (oy < d.ny && oz < d.nz) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = x0 + 4 * g + j;
                if (x < d.nx) store_tile(out, skip, d.vox(x, oy, oz), li, acc[j]);
            }
        }
    }
}

// Transposed stride-2 layer on matrix cores: a tile is 16 fine voxels of one
// x-parity (fine x = 2 (c0 + i) + px), so all of them use the same taps.
__global__ __launch_bounds__(256) void convT_cc_mfma(const uint16_t *__restrict__ in, Dims di,
                                                     uint16_t *__restrict__ out, Dims dout,
                                                     const uint16_t *__restrict__ Wb,
                                                     const uint16_t *__restrict__ skip, int tpw)
{
    __shared__ __align__(16) uint16_t wl[27 * kC * kC];
    stage_packed(Wb, wl);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int txn = (di.nx + 15) / 16;
    const long ntiles = (long)txn * 2 * dout.ny * dout.nz;
    for (int tt = 0; tt < tpw; tt++) {
    const long tile = ((long)blockIdx.x * tpw + tt) * 4 + wave;
    if (tile >= ntiles) return;
    const int tx = (int)(tile % txn);
    long rest = tile / txn;
    const int px = (int)(rest % 2);
    rest /= 2;
    const int fy = (int)(rest % dout.ny), fz = (int)(rest / dout.ny);
    const int c = min(tx * 16 + li, di.nx - 1);
    int kx[2], ox[2], ky[2], cy[2], kz[2], cz[2];
    const int nkx = px ? 2 : 1, nky = (fy & 1) ? 2 : 1, nkz = (fz & 1) ? 2 : 1;
    if (px) { kx[0] = 0; ox[0] = 1; kx[1] = 2; ox[1] = 0; } else { kx[0] = 1; ox[0] = 0; }
    if (fy & 1) { ky[0] = 0; cy[0] = wrap((fy + 1) / 2, di.ny); ky[1] = 2; cy[1] = (fy - 1) / 2; }
    else { ky[0] = 1; cy[0] = fy / 2; }
    if (fz & 1) { kz[0] = 0; cz[0] = wrap((fz + 1) / 2, di.nz); kz[1] = 2; cz[1] = (fz - 1) / 2; }
    else { kz[0] = 1; cz[0] = fz / 2; }
    // up to 8 taps: loads first, then the MFMA chain
    // unconditional loads (taps past nk* read the first tap again and are
    // skipped below): a guarded load became a branch with a wait after it
    bf16x4 av[8];
    int kk[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int a = t >> 2, b = (t >> 1) & 1, e = t & 1;
        const bool ok = a < nkz && b < nky && e < nkx;
        const int aa = ok ? a : 0, bb = ok ? b : 0, ee = ok ? e : 0;
        kk[t] = ok ? kz[aa] * 9 + ky[bb] * 3 + kx[ee] : -1;
        const int cx = wrap(c + ox[ee], di.nx);
        av[t] = *reinterpret_cast<const bf16x4 *>(in + di.vox(cx, cy[bb], cz[aa]) * kC + 4 * g);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if (kk[t] < 0) continue;
        const bf16x4 bv = *reinterpret_cast<const bf16x4 *>(wl + (kk[t] * kC + li) * kC + 4 * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[t], bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int cc = tx * 16 + 4 * g + j;
        if (cc < di.nx) store_tile(out, skip, dout.vox(2 * cc + px, fy, fz), li, acc[j]);
    }
    }
}

// 1 -> C stride-2 convolution from a scalar field at the fine level (FP64 or
// FP32 input, rounded to BF16): out[co][y] = sum_k W[co][0][k] in[2y + k - 1]
// times in_scale.
template <typename TI>
__global__ __launch_bounds__(256) void conv_1c(const TI *__restrict__ in, Dims di,
                                               uint16_t *__restrict__ out, Dims dout,
                                               const float *__restrict__ W, float in_scale)
{
    __shared__ float wl[27 * kC];
    stage_weights(W, 1, kC, wl);
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dout.n()) return;
    const int vq = (int)v / dout.nx;  // 32-bit: a level holds < 2^31 voxels
    const int ox = (int)v - vq * dout.nx, oy = vq % dout.ny, oz = vq / dout.ny;
    float acc[kC];
#pragma unroll
    for (int c = 0; c < kC; c++) acc[c] = 0.f;
    // all 27 tap loads issued before the accumulation (same order of sums)
    TI xin[27];
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int ix = wrap(2 * ox + k % 3 - 1, di.nx);
        const int iy = wrap(2 * oy + (k / 3) % 3 - 1, di.ny);
        const int iz = wrap(2 * oz + k / 9 - 1, di.nz);
        xin[k] = in[di.vox(ix, iy, iz)];
    }
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const float xv = bfround((float)xin[k] * in_scale);
#pragma unroll
        for (int co = 0; co < kC; co++) acc[co] += wl[k * kC + co] * xv;
    }
    store16(out + v * kC, acc);
}

// C -> 1 transposed stride-2 convolution to the fine level fused with the
// preconditioner output: U = P1(in); f = (r + U * out_scale) / |D0|.
// One thread per coarse voxel c writes the 2 x 2 x 2 fine voxels
// 2c + {0,1}^3: each takes its taps from the coarse voxels c + {0,1}^3 only
// (per axis, fine 2c: tap 1 of c; fine 2c+1: tap 0 of c+1, then tap 2 of c),
// so the eight 32-byte coarse voxels are loaded once for eight outputs
// instead of once per tap of every output.  Taps summed z, y, x outer to
// inner, channels in order.
template <typename TI, typename TO, bool HAS_R = true>
__global__ __launch_bounds__(256) void convT_c1_out_blk(const uint16_t *__restrict__ in, Dims di,
                                                        Dims dout, const float *__restrict__ W,
                                                        const TI *__restrict__ r, float out_scale,
                                                        double inv_d0, TO *__restrict__ f)
{
    __shared__ float wl[27 * kC];
    stage_weights(W, kC, 1, wl);
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= di.n()) return;
    const int vq = (int)v / di.nx;  // 32-bit: a level holds < 2^31 voxels
    const int cx = (int)v - vq * di.nx, cy = vq % di.ny, cz = vq / di.ny;
    // coarse voxel (cx + dx, cy + dy, cz + dz), d in {0,1}: 16 BF16 channels as 8 words
    uint32_t cw[8][8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int dx = t & 1, dy = (t >> 1) & 1, dz = t >> 2;
        const uint16_t *p = in + di.vox(wrap(cx + dx, di.nx), wrap(cy + dy, di.ny), wrap(cz + dz, di.nz)) * kC;
        const uint4 a = *reinterpret_cast<const uint4 *>(p);
        const uint4 b = *reinterpret_cast<const uint4 *>(p + 8);
        cw[t][0] = a.x; cw[t][1] = a.y; cw[t][2] = a.z; cw[t][3] = a.w;
        cw[t][4] = b.x; cw[t][5] = b.y; cw[t][6] = b.z; cw[t][7] = b.w;
    }
    TI rv[8];
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int px = o & 1, py = (o >> 1) & 1, pz = o >> 2;
        rv[o] = HAS_R ? r[dout.vox(2 * cx + px, 2 * cy + py, 2 * cz + pz)] : TI(0);
    }
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int px = o & 1, py = (o >> 1) & 1, pz = o >> 2;
        // per axis: (tap, coarse offset) pairs, odd parity tap 0 then tap 2
        const int nkx = px ? 2 : 1, nky = py ? 2 : 1, nkz = pz ? 2 : 1;
        float u = 0.f;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            if (a >= nkz) continue;
            const int kz = pz ? (a == 0 ? 0 : 2) : 1, dz = pz ? (a == 0 ? 1 : 0) : 0;
#pragma unroll
            for (int b = 0; b < 2; b++) {
                if (b >= nky) continue;
                const int ky = py ? (b == 0 ? 0 : 2) : 1, dy = py ? (b == 0 ? 1 : 0) : 0;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (c >= nkx) continue;
                    const int kx = px ? (c == 0 ? 0 : 2) : 1, dx = px ? (c == 0 ? 1 : 0) : 0;
                    const int k = kz * 9 + ky * 3 + kx;
                    const uint32_t *w = cw[dz * 4 + dy * 2 + dx];
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        u += wl[k * kC + 2 * i] * __uint_as_float(w[i] << 16);
                        u += wl[k * kC + 2 * i + 1] * __uint_as_float(w[i] & 0xffff0000u);
                    }
                }
            }
        }
        f[dout.vox(2 * cx + px, 2 * cy + py, 2 * cz + pz)] =
            (TO)(((double)rv[o] + (double)u * out_scale) * inv_d0);
    }
}

// ---- grids whose dimensions are not multiples of 4 -------------------------
// Periodic linear interpolation along one axis from n_src to n_dst points over
// the same period: target point j sits at s = j n_src / n_dst source points and
// takes (1 - t, t) of source points i0 = floor(s), i1 = i0 + 1 (mod n_src).
__device__ __forceinline__ void lin_taps(int j, int n_src, int n_dst, int &i0, int &i1, float &t)
{
    const long num = (long)j * n_src;
    i0 = (int)(num / n_dst);
    t = (float)(num - (long)i0 * n_dst) / (float)n_dst;
    i1 = i0 + 1 == n_src ? 0 : i0 + 1;
}

// out (FP32, grid dd) = trilinear periodic interpolation of in (grid ds).
template <typename TI>
__global__ __launch_bounds__(256) void interp_trilinear(const TI *__restrict__ in, Dims ds,
                                                       float *__restrict__ out, Dims dd)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dd.n()) return;
    const int vq = (int)v / dd.nx;
    const int x = (int)v - vq * dd.nx, y = vq % dd.ny, z = vq / dd.ny;
    int ix[2], iy[2], iz[2];
    float tx, ty, tz;
    lin_taps(x, ds.nx, dd.nx, ix[0], ix[1], tx);
    lin_taps(y, ds.ny, dd.ny, iy[0], iy[1], ty);
    lin_taps(z, ds.nz, dd.nz, iz[0], iz[1], tz);
    const float wx[2] = {1.f - tx, tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int a = c >> 2, b = (c >> 1) & 1, e = c & 1;
        acc += wz[a] * wy[b] * wx[e] * (float)in[ds.vox(ix[e], iy[b], iz[a])];
    }
    out[v] = acc;
}

// The preconditioner output on the grid from the network output um on the
// internal grid: f = (r + I_{M->N} um) * inv_d0.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void prolong_out(const float *__restrict__ um, Dims dm,
                                                  const TI *__restrict__ r, Dims d0, double inv_d0,
                                                  TO *__restrict__ f)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d0.n()) return;
    const int vq = (int)v / d0.nx;
    const int x = (int)v - vq * d0.nx, y = vq % d0.ny, z = vq / d0.ny;
    int ix[2], iy[2], iz[2];
    float tx, ty, tz;
    lin_taps(x, dm.nx, d0.nx, ix[0], ix[1], tx);
    lin_taps(y, dm.ny, d0.ny, iy[0], iy[1], ty);
    lin_taps(z, dm.nz, d0.nz, iz[0], iz[1], tz);
    const float wx[2] = {1.f - tx, tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
    float u = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int a = c >> 2, b = (c >> 1) & 1, e = c & 1;
        u += wz[a] * wy[b] * wx[e] * um[dm.vox(ix[e], iy[b], iz[a])];
    }
    f[v] = (TO)(((double)r[v] + (double)u) * inv_d0);
}

// Taps of the transpose of the interpolation from m to n points along one
// axis: the points i of the n-grid whose interpolation reads point j of the
// m-grid, with their weights (at most 5 candidates for n <= m; n >= 8 keeps
// the wrapped candidates distinct).  Returns the count.
__device__ __forceinline__ int adj_taps(int j, int n, int m, int *idx, float *w)
{
    int cnt = 0;
    const int lo = (int)floor((double)(j - 1) * n / m), hi = (int)ceil((double)(j + 1) * n / m);
    for (int ip = lo; ip <= hi && cnt < 6; ip++) {
        const int i = ((ip % n) + n) % n;
        int j0, j1;
        float t;
        lin_taps(i, m, n, j0, j1, t);
        const float ww = (j0 == j ? 1.f - t : 0.f) + (j1 == j ? t : 0.f);
        if (ww != 0.f) {
            idx[cnt] = i;
            w[cnt] = ww;
            cnt++;
        }
    }
    return cnt;
}

// out (internal grid dm) = transpose of the interpolation dm -> d0 applied to
// g (grid d0): the gradient of the network output from that of the
// preconditioner output.  Gather form, fixed order (no atomics).
__global__ __launch_bounds__(256) void interp_adjoint(const float *__restrict__ g, Dims d0,
                                                     float *__restrict__ out, Dims dm)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dm.n()) return;
    const int vq = (int)v / dm.nx;
    const int x = (int)v - vq * dm.nx, y = vq % dm.ny, z = vq / dm.ny;
    int ix[6], iy[6], iz[6];
    float wx[6], wy[6], wz[6];
    const int nx = adj_taps(x, d0.nx, dm.nx, ix, wx);
    const int ny = adj_taps(y, d0.ny, dm.ny, iy, wy);
    const int nz = adj_taps(z, d0.nz, dm.nz, iz, wz);
    float acc = 0.f;
    for (int a = 0; a < nz; a++)
        for (int b = 0; b < ny; b++)
            for (int e = 0; e < nx; e++) acc += wz[a] * wy[b] * wx[e] * g[d0.vox(ix[e], iy[b], iz[a])];
    out[v] = acc;
}

// Weight gradient G[a][b][k] = sum_y A[y][a] B[S y + k - 1][b] over the
// coarse voxels y of this block's chunk (partial per block; summed in a
// fixed order by wgrad_reduce).  A: BF16 channels-last (16 channels); B:
// BF16 channels-last (CB = 16) or a scalar FP32 field (CB = 1).  The chunk's
// A rows are staged in LDS once; each thread's B values are gathered kWB
// voxels at a time, all loads before the FMAs (one dependent load per voxel
// left the kernel latency-bound at ~190 us per 26^3 layer).  Every (a, b, k)
// still accumulates over y in order, so the sums are unchanged.
constexpr int kWChunk = 32;   // coarse voxels per block (26^3: 550 blocks)
constexpr int kWB = 16;       // B gathers in flight per thread
template <int S, int CB>
__global__ __launch_bounds__(256) void wgrad(const uint16_t *__restrict__ A, Dims da,
                                             const void *__restrict__ Bv, Dims db,
                                             float *__restrict__ partial)
{
    constexpr int NPK = 27 * CB;                      // (k, b) pairs
    constexpr int PPT = (NPK + 255) / 256;            // pairs per thread
    __shared__ float as[kWChunk][kC];
    float acc[PPT][kC];
#pragma unroll
    for (int p = 0; p < PPT; p++)
#pragma unroll
        for (int a = 0; a < kC; a++) acc[p][a] = 0.f;
    const long y0 = (long)blockIdx.x * kWChunk;
    const int ny = (int)min((long)kWChunk, da.n() - y0);
    for (int e = threadIdx.x; e < ny * kC; e += 256) as[e / kC][e % kC] = bf2f(A[y0 * kC + e]);
    __syncthreads();
    int pk_k[PPT], pk_b[PPT];
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        const int pk = min((int)threadIdx.x + p * 256, NPK - 1);
        pk_k[p] = pk / CB;
        pk_b[p] = pk % CB;
    }
    for (int yb = 0; yb < ny; yb += kWB) {
        float bv[kWB][PPT];
#pragma unroll
        for (int u = 0; u < kWB; u++) {
            // tail: a valid voxel, not accumulated; 32-bit index arithmetic
            // (a coarse level holds far fewer than 2^31 voxels)
            const int y = (int)y0 + min(yb + u, ny - 1);
            const int yq = y / da.nx;
            const int cx = y - yq * da.nx, cy = yq % da.ny, cz = yq / da.ny;
#pragma unroll
            for (int p = 0; p < PPT; p++) {
                const int k = pk_k[p], b = pk_b[p];
                const int ix = wrap(S * cx + k % 3 - 1, db.nx);
                const int iy = wrap(S * cy + (k / 3) % 3 - 1, db.ny);
                const int iz = wrap(S * cz + k / 9 - 1, db.nz);
                if (CB == 1)
                    bv[u][p] = bfround(reinterpret_cast<const float *>(Bv)[db.vox(ix, iy, iz)]);
                else
                    bv[u][p] = bf2f(reinterpret_cast<const uint16_t *>(Bv)[db.vox(ix, iy, iz) * kC + b]);
            }
        }
#pragma unroll
        for (int u = 0; u < kWB; u++) {
            if (yb + u >= ny) break;
#pragma unroll
            for (int p = 0; p < PPT; p++) {
                if ((int)threadIdx.x + p * 256 >= NPK) continue;
#pragma unroll
                for (int a = 0; a < kC; a++) acc[p][a] += as[yb + u][a] * bv[u][p];
            }
        }
    }
    // partial[block][a][b][k]
#pragma unroll
    for (int p = 0; p < PPT; p++) {
        const int pk = threadIdx.x + p * 256;
        if (pk >= NPK) continue;
        const int k = pk / CB, b = pk % CB;
#pragma unroll
        for (int a = 0; a < kC; a++)
            partial[(long)blockIdx.x * kC * CB * 27 + ((long)a * CB + b) * 27 + k] = acc[p][a];
    }
}

// out[i] (+)= sum over blocks of partial[block][i], fixed order; `transpose`
// maps G[a][b][k] (a < na, b < nb) to dW[b][a][k].
__global__ void wgrad_reduce(const float *__restrict__ partial, int nblocks, int na, int nb,
                             int transpose, float *__restrict__ dW)
{
    // eight lanes per output (g = lane % 8, all in one wave): lane g sums the
    // blocks b = g mod 8 in ascending order, then the eight sums are combined
    // pairwise, ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)), by an xor
    // butterfly (the same sum on every lane: addition is commutative)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = na * nb * 27;
    const int i = min(t >> 3, n - 1), g = t & 7;
    float s = 0.f;
#pragma unroll 8
    for (int b = g; b < nblocks; b += 8) s += partial[(long)b * n + i];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    if ((t >> 3) >= n || g != 0) return;
    const int k = i % 27, bb = (i / 27) % nb, a = i / (27 * nb);
    const int o = transpose ? (bb * na + a) * 27 + k : i;
    dW[o] += s;
}

// Weight transforms for the data-gradient passes: mode 0 W'[ci][co][k] =
// W[co][ci][26 - k] (stride-1 conv adjoint), mode 1 W'[ci][co][k] = W[co][ci][k]
// (stride-2 conv <-> transposed conv adjoints).
__global__ void wtransform(const float *__restrict__ W, int cout, int cin, int mode,
                           float *__restrict__ Wt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cout * cin * 27) return;
    const int k = i % 27, ci = (i / 27) % cin, co = i / (27 * cin);
    Wt[((long)ci * cout + co) * 27 + (mode == 0 ? 26 - k : k)] = W[i];
}

// loss partials: d = f - e; partial[block] = sum d^2; df = 2 d * gscale
__global__ __launch_bounds__(256) void loss_grad(const float *__restrict__ f,
                                                 const float *__restrict__ e, long n,
                                                 float gscale, float *__restrict__ df,
                                                 double *__restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float d = f[i] - e[i];
        s += (double)d * d;
        df[i] = 2.f * d * gscale;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void adam_step(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                          float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void fill_f32(float *p, float v, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// pair = (r / |e|, (e - mean e) / |e|) in FP32, from FP64 device vectors
__global__ void make_pair(const double *__restrict__ r, const double *__restrict__ xs,
                          const double *__restrict__ xk, double mean_e, double inv_norm, long n,
                          float *__restrict__ pr, float *__restrict__ pe)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pr[i] = (float)(r[i] * inv_norm);
    pe[i] = (float)((xs[i] - xk[i] - mean_e) * inv_norm);
}

// per-block partial sums of e = x - xk and e^2 (partial[0..B), partial[B..2B))
__global__ __launch_bounds__(256) void diff_stats(const double *__restrict__ x,
                                                  const double *__restrict__ xk, long n,
                                                  double *__restrict__ partial)
{
    __shared__ double r1[256], r2[256];
    double a = 0.0, b = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double e = x[i] - xk[i];
        a += e;
        b += e * e;
    }
    r1[threadIdx.x] = a;
    r2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r1[0];
        partial[gridDim.x + blockIdx.x] = r2[0];
    }
}

unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

bool g_unet_scalar = false;  // SPARC_AMD_UNET_SCALAR=1 selects the FP32-FMA layer kernels
bool g_unet_lds = true;      // SPARC_AMD_UNET_LDS=0: stride-1 layers without the LDS input box

// 16-voxel tiles per wave: enough to amortise the weight staging, few enough
// to keep ~2048 blocks in flight (the quarter-resolution layers are small)
int tiles_per_wave(long tiles)
{
    return (int)std::max(1L, std::min<long>(kTilesPerWave, tiles / (4L * 2048)));
}

// W: FP32 [co][ci][27] (FMA kernels); Wb: the same packed by pack_bf16
// (matrix-core kernels)
template <int S>
void launch_conv_cc(const uint16_t *in, Dims di, uint16_t *out, Dims dout, const float *W,
                    const uint16_t *Wb, const uint16_t *skip, hipStream_t s)
{
    if (g_unet_scalar) {
        conv_cc<S><<<nblk(dout.n()), 256, 0, s>>>(in, di, out, dout, W, skip);
    } else if (S == 1 && g_unet_lds) {
        const long nb = (long)((dout.nx + kLX - 1) / kLX) * ((dout.ny + kLY - 1) / kLY) *
                        ((dout.nz + kLZ - 1) / kLZ);
        conv_cc_lds<<<(unsigned)nb, 256, 0, s>>>(in, dout, out, Wb, skip);
    } else {
        const long tiles = (long)((dout.nx + 15) / 16) * dout.ny * dout.nz;
        const int tpw = tiles_per_wave(tiles);
        const long per = 4L * tpw;
        conv_cc_mfma<S><<<(unsigned)((tiles + per - 1) / per), 256, 0, s>>>(in, di, out, dout, Wb, skip,
                                                                         tpw);
    }
}

void launch_convT_cc(const uint16_t *in, Dims di, uint16_t *out, Dims dout, const float *W,
                     const uint16_t *Wb, const uint16_t *skip, hipStream_t s)
{
    if (g_unet_scalar) {
        convT_cc<<<nblk(dout.n()), 256, 0, s>>>(in, di, out, dout, W, skip);
    } else {
        const long tiles = (long)((di.nx + 15) / 16) * 2 * dout.ny * dout.nz;
        const int tpw = tiles_per_wave(tiles);
        const long per = 4L * tpw;
        convT_cc_mfma<<<(unsigned)((tiles + per - 1) / per), 256, 0, s>>>(in, di, out, dout, Wb, skip,
                                                                         tpw);
    }
}


}  // namespace

struct sgpu_unet {
    sgpu_context *ctx = nullptr;
    // d0: the context's grid; dm: the network's fine level (d0 with every
    // dimension rounded up to a multiple of 4); d1 = dm / 2, d2 = dm / 4
    Dims d0{}, dm{}, d1{}, d2{};
    bool resample = false;         // dm != d0: interpolate in and out
    float *rm = nullptr;           // the input interpolated onto dm
    float *um = nullptr;           // the network output on dm / its gradient
    double d0sum = -1.0;  // D2x[0] + D2y[0] + D2z[0] (the Laplacian's diagonal)
    double d0abs = 1.0;
    double shift = 0.0;   // c of the operator Lap + c harvested for and trained on
    int nparam = 0;
    int off[kLayers + 1] = {0};
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *wt = nullptr;
    uint16_t *wb = nullptr;   // layers 1..6 packed for the matrix cores (pack_bf16)
    bool wb_current = false;  // wb holds the packed current weights (cleared on every write of w)
    uint16_t *wtb = nullptr;  // wt packed likewise
    // wt / wtb hold the transforms of all eight layers (layer l at l * kWcc),
    // formed once per change of the weights, not once per pair
    bool wt_current = false;
    // activations (BF16 channels-last) and their gradients
    uint16_t *a1 = nullptr, *b1 = nullptr, *a2 = nullptr, *b2 = nullptr, *c2 = nullptr,
             *s1 = nullptr, *c1 = nullptr;
    uint16_t *ga = nullptr, *gb = nullptr, *gc = nullptr, *gd = nullptr;  // L1 grads x2, L2 x2
    float *f32buf = nullptr;       // fine-level f / df (training)
    float *partial = nullptr;      // weight-gradient partials
    size_t partial_elems = 0;
    double *lpart = nullptr;       // loss partials
    // training pairs
    float *pr = nullptr, *pe = nullptr;
    int npairs = 0, cap = 0;
    long adam_t = 0;
    std::vector<void *> owned;
};

namespace {

template <typename T>
int dalloc(sgpu_unet *u, T **p, size_t n)
{
    SGPU_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    u->owned.push_back((void *)*p);
    return SGPU_OK;
}

const float *layer_w(const sgpu_unet *u, const float *base, int l) { return base + u->off[l]; }

// 1 / |D0 + shift|: the output scale of training and of sgpu_unet_apply
double out_scale(const sgpu_unet *u) { return 1.0 / std::fabs(u->d0sum + u->shift); }

// forward pass for one input (fine field r of type TI); writes f (type TO)
template <typename TI, typename TO>
int forward(sgpu_unet *u, const TI *r, TO *f, double out_scale)
{
    hipStream_t s = u->ctx->stream;
    const float *W = u->w;
    // layers 1..6 (C -> C, contiguous in W) packed for the matrix cores, once
    // per change of the weights (stream order makes the pack visible to every
    // later pass)
    if (!u->wb_current) {
        pack_bf16<<<nblk(6L * kWcc), 256, 0, s>>>(layer_w(u, W, 1), 6, u->wb);
        u->wb_current = true;
    }
    auto wb = [&](int l) -> const uint16_t * { return u->wb + (long)(l - 1) * kWcc; };
    if (u->resample) {
        interp_trilinear<TI><<<nblk(u->dm.n()), 256, 0, s>>>(r, u->d0, u->rm, u->dm);
        conv_1c<float><<<nblk(u->d1.n()), 256, 0, s>>>(u->rm, u->dm, u->a1, u->d1, layer_w(u, W, 0),
                                                       1.f);
    } else {
        conv_1c<TI><<<nblk(u->d1.n()), 256, 0, s>>>(r, u->d0, u->a1, u->d1, layer_w(u, W, 0), 1.f);
    }
    launch_conv_cc<1>(u->a1, u->d1, u->b1, u->d1, layer_w(u, W, 1), wb(1), nullptr, s);
    launch_conv_cc<2>(u->b1, u->d1, u->a2, u->d2, layer_w(u, W, 2), wb(2), nullptr, s);
    launch_conv_cc<1>(u->a2, u->d2, u->b2, u->d2, layer_w(u, W, 3), wb(3), nullptr, s);
    launch_conv_cc<1>(u->b2, u->d2, u->c2, u->d2, layer_w(u, W, 4), wb(4), nullptr, s);
    launch_convT_cc(u->c2, u->d2, u->s1, u->d1, layer_w(u, W, 5), wb(5), u->b1, s);
    launch_conv_cc<1>(u->s1, u->d1, u->c1, u->d1, layer_w(u, W, 6), wb(6), nullptr, s);
    if (u->resample) {
        convT_c1_out_blk<float, float, false><<<nblk(u->d1.n()), 256, 0, s>>>(
            u->c1, u->d1, u->dm, layer_w(u, W, 7), nullptr, 1.f, 1.0, u->um);
        prolong_out<TI, TO><<<nblk(u->d0.n()), 256, 0, s>>>(u->um, u->dm, r, u->d0, out_scale, f);
    } else {
        convT_c1_out_blk<TI, TO><<<nblk(u->d1.n()), 256, 0, s>>>(u->c1, u->d1, u->d0, layer_w(u, W, 7),
                                                             r, 1.f, out_scale, f);
    }
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

int wgrad_launch(sgpu_unet *u, int stride, const uint16_t *A, Dims da, const void *B, int cb,
                 Dims db, int transpose, int layer)
{
    hipStream_t s = u->ctx->stream;
    const int nb = (int)((da.n() + kWChunk - 1) / kWChunk);
    const size_t need = (size_t)nb * kC * cb * 27;
    if (need > u->partial_elems) {
        sgpu_set_error("unet: weight-gradient partial buffer too small");
        return SGPU_ERR_STATE;
    }
    if (stride == 1 && cb == kC)
        wgrad<1, kC><<<nb, 256, 0, s>>>(A, da, B, db, u->partial);
    else if (stride == 2 && cb == kC)
        wgrad<2, kC><<<nb, 256, 0, s>>>(A, da, B, db, u->partial);
    else if (stride == 2 && cb == 1)
        wgrad<2, 1><<<nb, 256, 0, s>>>(A, da, B, db, u->partial);
    else {
        sgpu_set_error("unet: unsupported weight-gradient shape");
        return SGPU_ERR_ARG;
    }
    const int n = kC * cb * 27;
    wgrad_reduce<<<nblk(8L * n), 256, 0, s>>>(u->partial, nb, kC, cb, transpose, u->g + u->off[layer]);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// backward for one pair given df (dL/df, fine level, FP32) with the forward
// activations of that pair resident; accumulates into u->g
int backward(sgpu_unet *u, const float *r, const float *df)
{
    hipStream_t s = u->ctx->stream;
    const float *W = u->w;
    // the adjoint weights of layers 1..7 (mode 0: stride-1 conv adjoint, 1:
    // stride-2 conv <-> transposed conv), formed when the weights changed
    static const int kMode[kLayers] = {0, 0, 1, 0, 0, 1, 0, 1};
    if (!u->wt_current) {
        for (int l = 1; l < kLayers; l++) {
            wtransform<<<nblk(layer_size(l)), 256, 0, s>>>(layer_w(u, W, l), kLayerCout[l],
                                                            kLayerCin[l], kMode[l], u->wt + (long)l * kWcc);
            if (kLayerCin[l] == kC && kLayerCout[l] == kC)
                pack_bf16<<<nblk(kWcc), 256, 0, s>>>(u->wt + (long)l * kWcc, 1, u->wtb + (long)l * kWcc);
        }
        u->wt_current = true;
    }
    const float *wt = nullptr;
    const uint16_t *wtb = nullptr;
    auto transform = [&](int l, int mode) {
        (void)mode;  // kMode[l]
        wt = u->wt + (long)l * kWcc;
        wtb = u->wtb + (long)l * kWcc;
    };
    // f = (r + U) / |D0 + shift|  ->  dU = df / |D0 + shift|, applied by loss_grad;
    // on a resampled grid the network output's gradient is I_{M->N}' dU
    (void)df;
    const float *dU = u->f32buf;
    if (u->resample) {
        interp_adjoint<<<nblk(u->dm.n()), 256, 0, s>>>(u->f32buf, u->d0, u->um, u->dm);
        dU = u->um;
    }
    // P1 (convT C->1, L1->L0): dW7[0][ci][k] = sum_y c1[y][ci] dU[2y+k-1]
    //   -> G[a=ci][b=0][k], transposed to [co=0][ci][k]
    SGPU_CHECK(wgrad_launch(u, 2, u->c1, u->d1, dU, 1, u->dm, 1, 7));
    // dc1[ci][y] = sum_k W7[0][ci][k] dU[2y+k-1]: a 1->C stride-2 conv with W'[ci][0][k]
    transform(7, 1);
    conv_1c<float><<<nblk(u->d1.n()), 256, 0, s>>>(dU, u->dm, u->ga, u->d1, wt, 1.f);
    // C1 (conv C->C, L1): dW6 = sum_y dc1[y][co] s1[y+k-1][ci]
    SGPU_CHECK(wgrad_launch(u, 1, u->ga, u->d1, u->s1, kC, u->d1, 0, 6));
    transform(6, 0);
    launch_conv_cc<1>(u->ga, u->d1, u->gb, u->d1, wt, wtb, nullptr, s);  // ds1
    // P2 (convT L2->L1): dW5[co][ci][k] = sum_y c2[y][ci] ds1[2y+k-1][co] -> transpose
    SGPU_CHECK(wgrad_launch(u, 2, u->c2, u->d2, u->gb, kC, u->d1, 1, 5));
    transform(5, 1);
    launch_conv_cc<2>(u->gb, u->d1, u->gc, u->d2, wt, wtb, nullptr, s);  // dc2
    // M2 (conv L2): dW4 = sum dc2[y][co] b2[y+k-1][ci]
    SGPU_CHECK(wgrad_launch(u, 1, u->gc, u->d2, u->b2, kC, u->d2, 0, 4));
    transform(4, 0);
    launch_conv_cc<1>(u->gc, u->d2, u->gd, u->d2, wt, wtb, nullptr, s);  // db2
    // B2 (conv L2): dW3 = sum db2[y][co] a2[y+k-1][ci]
    SGPU_CHECK(wgrad_launch(u, 1, u->gd, u->d2, u->a2, kC, u->d2, 0, 3));
    transform(3, 0);
    launch_conv_cc<1>(u->gd, u->d2, u->gc, u->d2, wt, wtb, nullptr, s);  // da2
    // R2 (conv/2 L1->L2): dW2 = sum da2[y][co] b1[2y+k-1][ci]
    SGPU_CHECK(wgrad_launch(u, 2, u->gc, u->d2, u->b1, kC, u->d1, 0, 2));
    // db1 = ds1 (skip) + R2 adjoint(da2): transposed conv with W'[ci][co][k] = W2[co][ci][k]
    transform(2, 1);
    launch_convT_cc(u->gc, u->d2, u->ga, u->d1, wt, wtb, u->gb, s);  // db1
    // B1 (conv L1): dW1 = sum db1[y][co] a1[y+k-1][ci]
    SGPU_CHECK(wgrad_launch(u, 1, u->ga, u->d1, u->a1, kC, u->d1, 0, 1));
    transform(1, 0);
    launch_conv_cc<1>(u->ga, u->d1, u->gb, u->d1, wt, wtb, nullptr, s);  // da1
    // R1 (conv/2 1->C, L0->L1): dW0[co][0][k] = sum_y da1[y][co] r[2y+k-1]
    // (the interpolated input of this pair's forward pass on a resampled grid)
    SGPU_CHECK(wgrad_launch(u, 2, u->gb, u->d1, u->resample ? u->rm : r, 1, u->dm, 0, 0));
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// loss and gradient over a list of pairs (device FP32 r, e of length n0 each)
int loss_grad_pairs(sgpu_unet *u, const float *const *R, const float *const *E, int nb, double *loss)
{
    hipStream_t s = u->ctx->stream;
    const long n0 = u->d0.n();
    fill_f32<<<nblk(u->nparam), 256, 0, s>>>(u->g, 0.f, u->nparam);
    double total = 0.0;
    const int lblocks = 256;
    std::vector<double> hp(lblocks);
    for (int i = 0; i < nb; i++) {
        SGPU_CHECK((forward<float, float>(u, R[i], u->f32buf, out_scale(u))));
        // df = 2 (f - e) / nb, then dU = df / |D0 + shift|
        loss_grad<<<lblocks, 256, 0, s>>>(u->f32buf, E[i], n0, (float)(out_scale(u) / nb),
                                          u->f32buf, u->lpart);
        SGPU_HIP(hipGetLastError());
        SGPU_CHECK(backward(u, R[i], u->f32buf));
        SGPU_HIP(hipMemcpyAsync(hp.data(), u->lpart, lblocks * sizeof(double), hipMemcpyDeviceToHost, s));
        SGPU_HIP(hipStreamSynchronize(s));
        for (double v : hp) total += v;
    }
    if (loss) *loss = total / nb;
    return SGPU_OK;
}

}  // namespace

extern "C" int sgpu_unet_create(sgpu_context *ctx, int channels, unsigned long long seed,
                                sgpu_unet **out)
{
    if (!ctx || !out || channels != kC) {
        sgpu_set_error("sgpu_unet_create: bad arguments (channels must be %d)", kC);
        return SGPU_ERR_ARG;
    }
    if (ctx->bcx || ctx->bcy || ctx->bcz || ctx->nx < 8 || ctx->ny < 8 || ctx->nz < 8) {
        sgpu_set_error("sgpu_unet_create: needs a periodic grid of at least 8 points per axis "
                       "(got %d x %d x %d%s)", ctx->nx, ctx->ny, ctx->nz,
                       (ctx->bcx || ctx->bcy || ctx->bcz) ? ", Dirichlet" : "");
        return SGPU_ERR_ARG;
    }
    if (ctx->nd >= (1L << 31)) {  // the kernels index a level's voxels with int
        sgpu_set_error("sgpu_unet_create: grid of %ld points exceeds 2^31 - 1", ctx->nd);
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    {
        const char *env = getenv("SPARC_AMD_UNET_SCALAR");
        g_unet_scalar = env && env[0] == '1';
        const char *lds = getenv("SPARC_AMD_UNET_LDS");
        g_unet_lds = !(lds && lds[0] == '0');
    }
    sgpu_unet *u = new sgpu_unet();
    u->ctx = ctx;
    u->d0 = Dims{ctx->nx, ctx->ny, ctx->nz};
    u->dm = Dims{(ctx->nx + 3) / 4 * 4, (ctx->ny + 3) / 4 * 4, (ctx->nz + 3) / 4 * 4};
    u->resample = u->dm.nx != ctx->nx || u->dm.ny != ctx->ny || u->dm.nz != ctx->nz;
    u->d1 = Dims{u->dm.nx / 2, u->dm.ny / 2, u->dm.nz / 2};
    u->d2 = Dims{u->dm.nx / 4, u->dm.ny / 4, u->dm.nz / 4};
    u->d0sum = ctx->d2x[0] + ctx->d2y[0] + ctx->d2z[0];
    u->d0abs = std::fabs(u->d0sum);
    for (int l = 0; l < kLayers; l++) u->off[l + 1] = u->off[l] + layer_size(l);
    u->nparam = u->off[kLayers];
    auto fail = [&](int st) { sgpu_unet_destroy(u); return st; };
    int st;
    if ((st = dalloc(u, &u->w, u->nparam)) || (st = dalloc(u, &u->g, u->nparam)) ||
        (st = dalloc(u, &u->m, u->nparam)) || (st = dalloc(u, &u->v, u->nparam)) ||
        (st = dalloc(u, &u->wt, (size_t)kLayers * kWcc)) || (st = dalloc(u, &u->wb, 6 * (size_t)kWcc)) ||
        (st = dalloc(u, &u->wtb, (size_t)kLayers * kWcc)))
        return fail(st);
    const size_t n1 = (size_t)u->d1.n() * kC, n2 = (size_t)u->d2.n() * kC;
    uint16_t **l1[] = {&u->a1, &u->b1, &u->s1, &u->c1, &u->ga, &u->gb};
    uint16_t **l2[] = {&u->a2, &u->b2, &u->c2, &u->gc, &u->gd};
    for (auto p : l1)
        if ((st = dalloc(u, p, n1))) return fail(st);
    for (auto p : l2)
        if ((st = dalloc(u, p, n2))) return fail(st);
    if ((st = dalloc(u, &u->f32buf, (size_t)u->d0.n())) || (st = dalloc(u, &u->lpart, 256)))
        return fail(st);
    if (u->resample &&
        ((st = dalloc(u, &u->rm, (size_t)u->dm.n())) || (st = dalloc(u, &u->um, (size_t)u->dm.n()))))
        return fail(st);
    u->partial_elems = (size_t)((u->d1.n() + kWChunk - 1) / kWChunk) * kC * kC * 27;
    if ((st = dalloc(u, &u->partial, u->partial_elems))) return fail(st);
    // weights: layers 0..6 uniform(-a, a) with a = sqrt(3 / fan_in) * 0.5
    // (splitmix64 stream from seed), the last layer zero so that the initial
    // preconditioner is exactly Jacobi
    std::vector<float> hw(u->nparam, 0.f);
    unsigned long long z = seed;
    auto next = [&]() {
        z += 0x9e3779b97f4a7c15ull;
        unsigned long long t = z;
        t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
        t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
        t ^= t >> 31;
        return (double)(t >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < kLayers - 1; l++) {
        const double a = 0.5 * std::sqrt(3.0 / (27.0 * kLayerCin[l]));
        for (int i = u->off[l]; i < u->off[l + 1]; i++) hw[i] = (float)((2.0 * next() - 1.0) * a);
    }
    SGPU_HIP(hipMemcpyAsync(u->w, hw.data(), hw.size() * sizeof(float), hipMemcpyHostToDevice,
                            ctx->stream));
    u->wb_current = false;
    u->wt_current = false;
    SGPU_HIP(hipMemsetAsync(u->m, 0, u->nparam * sizeof(float), ctx->stream));
    SGPU_HIP(hipMemsetAsync(u->v, 0, u->nparam * sizeof(float), ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    ctx->unets.push_back(u);
    *out = u;
    return SGPU_OK;
}

namespace {

void unet_free_device(sgpu_unet *u)
{
    for (void *p : u->owned) (void)hipFree(p);
    u->owned.clear();
    if (u->pr) (void)hipFree(u->pr);
    if (u->pe) (void)hipFree(u->pe);
    u->pr = u->pe = nullptr;
    u->npairs = u->cap = 0;
}

// Entry points other than sgpu_unet_destroy need the network's context.
int unet_live(const sgpu_unet *u, const char *what)
{
    if (!u) {
        sgpu_set_error("%s: null network", what);
        return SGPU_ERR_ARG;
    }
    if (!u->ctx) {
        sgpu_set_error("%s: the network's context was destroyed", what);
        return SGPU_ERR_STATE;
    }
    return SGPU_OK;
}

}  // namespace

// Called by sgpu_destroy (the context's stream is already synchronised): the
// context's stream and device memory go away with it, so the network keeps
// only its host handle.
void sgpu_unet_orphan(sgpu_unet *u)
{
    if (!u) return;
    unet_free_device(u);
    u->ctx = nullptr;
}

extern "C" void sgpu_unet_destroy(sgpu_unet *u)
{
    if (!u) return;
    if (u->ctx) {
        (void)hipStreamSynchronize(u->ctx->stream);
        auto &v = u->ctx->unets;
        v.erase(std::remove(v.begin(), v.end(), u), v.end());
    }
    unet_free_device(u);
    delete u;
}

extern "C" int sgpu_unet_nparams(const sgpu_unet *u) { return u ? u->nparam : 0; }

extern "C" int sgpu_unet_get_weights(const sgpu_unet *u, float *w)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_get_weights"));
    if (!w) return SGPU_ERR_ARG;
    SGPU_HIP(hipMemcpyAsync(w, u->w, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
    SGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    return SGPU_OK;
}

extern "C" int sgpu_unet_set_weights(sgpu_unet *u, const float *w)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_set_weights"));
    if (!w) return SGPU_ERR_ARG;
    SGPU_HIP(hipMemcpyAsync(u->w, w, u->nparam * sizeof(float), hipMemcpyHostToDevice, u->ctx->stream));
    u->wb_current = false;
    u->wt_current = false;
    SGPU_HIP(hipMemsetAsync(u->m, 0, u->nparam * sizeof(float), u->ctx->stream));
    SGPU_HIP(hipMemsetAsync(u->v, 0, u->nparam * sizeof(float), u->ctx->stream));
    SGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    u->adam_t = 0;
    return SGPU_OK;
}

// precond_fun of AAR (src/linearSolver.c:37-43): f = (r + U(r)) / |D0 + c| on
// FP64 device vectors of the context's grid, with |D0 + c| the magnitude of
// the Jacobi diagonal of the operator Lap + c being solved (c = 0 for
// Poisson, -kTF^2 for the Kerker solve).
extern "C" void sgpu_unet_precond(void *user, int N, double c, const double *r_dev, double *f_dev)
{
    sgpu_unet *u = static_cast<sgpu_unet *>(user);
    if (!u || !u->ctx || N != u->d0.nx * u->d0.ny * u->d0.nz) return;
    const double diag = std::fabs(u->d0sum + c);
    (void)forward<double, double>(u, r_dev, f_dev, diag > 0.0 ? 1.0 / diag : 1.0 / u->d0abs);
}

extern "C" int sgpu_unet_apply(sgpu_unet *u, const double *r_host, double *f_host)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_apply"));
    if (!r_host || !f_host) return SGPU_ERR_ARG;
    const long n = u->d0.n();
    double *buf = nullptr;
    SGPU_HIP(hipMalloc((void **)&buf, 2 * n * sizeof(double)));
    hipStream_t s = u->ctx->stream;
    int st = SGPU_OK;
    if (hipMemcpyAsync(buf, r_host, n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        st = SGPU_ERR_HIP;
    if (st == SGPU_OK) st = forward<double, double>(u, buf, buf + n, out_scale(u));
    if (st == SGPU_OK &&
        (hipMemcpyAsync(f_host, buf + n, n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        st = SGPU_ERR_HIP;
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (st == SGPU_ERR_HIP) sgpu_set_error("sgpu_unet_apply: HIP error");
    return st;
}

extern "C" int sgpu_unet_loss_grad(sgpu_unet *u, const double *r_host, const double *e_host,
                                   int npair, double *loss, float *grad)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_loss_grad"));
    if (!r_host || !e_host || npair < 1) return SGPU_ERR_ARG;
    const long n = u->d0.n();
    float *buf = nullptr;
    SGPU_HIP(hipMalloc((void **)&buf, 2 * (size_t)npair * n * sizeof(float)));
    std::vector<float> tmp((size_t)npair * n);
    hipStream_t s = u->ctx->stream;
    for (size_t i = 0; i < tmp.size(); i++) tmp[i] = (float)r_host[i];
    SGPU_HIP(hipMemcpy(buf, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice));
    for (size_t i = 0; i < tmp.size(); i++) tmp[i] = (float)e_host[i];
    SGPU_HIP(hipMemcpy(buf + (size_t)npair * n, tmp.data(), tmp.size() * sizeof(float),
                       hipMemcpyHostToDevice));
    std::vector<const float *> R(npair), E(npair);
    for (int i = 0; i < npair; i++) {
        R[i] = buf + (size_t)i * n;
        E[i] = buf + (size_t)(npair + i) * n;
    }
    int st = loss_grad_pairs(u, R.data(), E.data(), npair, loss);
    if (st == SGPU_OK && grad) {
        if (hipMemcpyAsync(grad, u->g, u->nparam * sizeof(float), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = SGPU_ERR_HIP;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    return st;
}

// Jacobi-AAR solve of -(Lap + shift) x = b (rhs b, start x0, the reference's
// AAR parameters 0.6 / 0.6 / 7 / 6) that records every `stride`-th iterate;
// afterwards each recorded (r_k, x_k) becomes the pair (r_k / |e_k|,
// (e_k - mean e_k) / |e_k|) with e_k = x* - x_k (the mean is removed only for
// shift 0, where constants are the operator's null space).
extern "C" int sgpu_unet_harvest(sgpu_unet *u, const double *b_host, const double *x0_host,
                                 double tol, int max_iter, int stride, int max_pairs, int *added)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_harvest"));
    if (!b_host || !x0_host || stride < 1 || max_pairs < 1) return SGPU_ERR_ARG;
    sgpu_context *ctx = u->ctx;
    const long n = u->d0.n(), ld = ctx->ld;
    hipStream_t s = ctx->stream;
    // room for the new pairs
    if (u->npairs + max_pairs > u->cap) {
        const int cap = u->npairs + max_pairs;
        float *pr = nullptr, *pe = nullptr;
        SGPU_HIP(hipMalloc((void **)&pr, (size_t)cap * n * sizeof(float)));
        SGPU_HIP(hipMalloc((void **)&pe, (size_t)cap * n * sizeof(float)));
        if (u->npairs) {
            SGPU_HIP(hipMemcpyAsync(pr, u->pr, (size_t)u->npairs * n * sizeof(float), hipMemcpyDeviceToDevice, s));
            SGPU_HIP(hipMemcpyAsync(pe, u->pe, (size_t)u->npairs * n * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        SGPU_HIP(hipStreamSynchronize(s));
        if (u->pr) SGPU_HIP(hipFree(u->pr));
        if (u->pe) SGPU_HIP(hipFree(u->pe));
        u->pr = pr;
        u->pe = pe;
        u->cap = cap;
    }
    // x, b and the recorded iterates (x_k, r_k) on the device
    double *mem = nullptr;
    SGPU_HIP(hipMalloc((void **)&mem, (2 + 2 * (size_t)max_pairs) * ld * sizeof(double)));
    double *x = mem, *b = mem + ld, *rec = mem + 2 * ld;
    SGPU_HIP(hipMemcpyAsync(x, x0_host, n * sizeof(double), hipMemcpyHostToDevice, s));
    SGPU_HIP(hipMemcpyAsync(b, b_host, n * sizeof(double), hipMemcpyHostToDevice, s));
    struct Rec {
        double *rec;
        long ld, n;
        int stride, max, count, seen;
        hipStream_t s;
    } R{rec, ld, n, stride, max_pairs, 0, 0, s};
    sgpu_aar_recorder hook;
    hook.user = &R;
    hook.fn = [](void *p, const double *x_dev, const double *r_dev) {
        Rec *q = static_cast<Rec *>(p);
        if (q->count < q->max && (q->seen++ % q->stride) == 0) {
            double *dst = q->rec + 2 * (long)q->count * q->ld;
            (void)hipMemcpyAsync(dst, x_dev, q->n * sizeof(double), hipMemcpyDeviceToDevice, q->s);
            (void)hipMemcpyAsync(dst + q->ld, r_dev, q->n * sizeof(double), hipMemcpyDeviceToDevice, q->s);
            q->count++;
        }
    };
    int it = 0;
    double rr = 0.0;
    int st = sgpu_aar_device(ctx, nullptr, nullptr, nullptr, u->shift, (int)n, x, b, 0.6, 0.6, 7, 6,
                             tol, max_iter, &it, &rr, &hook);
    int made = 0;
    if (st >= 0) {
        // pairs from the recorded iterates (skipping any whose error vanished)
        double *part = nullptr;
        const int nbk = 256;
        SGPU_HIP(hipMalloc((void **)&part, 2 * nbk * sizeof(double)));
        std::vector<double> hp(2 * nbk);
        for (int k = 0; k < R.count && st >= 0; k++) {
            double *xk = rec + 2 * (long)k * ld, *rk = xk + ld;
            diff_stats<<<nbk, 256, 0, s>>>(x, xk, n, part);
            if (hipMemcpyAsync(hp.data(), part, 2 * nbk * sizeof(double), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                st = SGPU_ERR_HIP;
                break;
            }
            double se = 0.0, se2 = 0.0;
            for (int i = 0; i < nbk; i++) {
                se += hp[i];
                se2 += hp[nbk + i];
            }
            const double mean = u->shift == 0.0 ? se / (double)n : 0.0;
            const double en2 = se2 - (double)n * mean * mean;
            if (!(en2 > 0.0) || !std::isfinite(en2)) continue;
            make_pair<<<nblk(n), 256, 0, s>>>(rk, x, xk, mean, 1.0 / std::sqrt(en2), n,
                                              u->pr + (size_t)u->npairs * n,
                                              u->pe + (size_t)u->npairs * n);
            u->npairs++;
            made++;
        }
        (void)hipStreamSynchronize(s);
        (void)hipFree(part);
        if (st == SGPU_ERR_HIP) sgpu_set_error("sgpu_unet_harvest: HIP error");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(mem);
    if (added) *added = made;
    return st < 0 ? st : SGPU_OK;
}

extern "C" int sgpu_unet_npairs(const sgpu_unet *u) { return u ? u->npairs : 0; }

extern "C" int sgpu_unet_set_shift(sgpu_unet *u, double c)
{
    if (!u || !(c <= 0.0) || !std::isfinite(c)) {
        sgpu_set_error("sgpu_unet_set_shift: the shift must be finite and <= 0");
        return SGPU_ERR_ARG;
    }
    u->shift = c;
    return SGPU_OK;
}

// Adam over the stored pairs: `steps` steps of `batch` pairs drawn with a
// splitmix64 stream from `seed`; loss_hist (may be NULL) gets each step's loss.
extern "C" int sgpu_unet_train(sgpu_unet *u, int steps, int batch, double lr,
                               unsigned long long seed, double *loss_hist)
{
    SGPU_CHECK(unet_live(u, "sgpu_unet_train"));
    if (steps < 0 || batch < 1 || u->npairs < 1) {
        sgpu_set_error("sgpu_unet_train: bad arguments or no training pairs");
        return SGPU_ERR_ARG;
    }
    const long n = u->d0.n();
    unsigned long long z = seed;
    auto next = [&]() {
        z += 0x9e3779b97f4a7c15ull;
        unsigned long long t = z;
        t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
        t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
        return t ^ (t >> 31);
    };
    std::vector<const float *> R(batch), E(batch);
    for (int it = 0; it < steps; it++) {
        for (int i = 0; i < batch; i++) {
            const int k = (int)(next() % (unsigned long long)u->npairs);
            R[i] = u->pr + (size_t)k * n;
            E[i] = u->pe + (size_t)k * n;
        }
        double loss = 0.0;
        SGPU_CHECK(loss_grad_pairs(u, R.data(), E.data(), batch, &loss));
        if (loss_hist) loss_hist[it] = loss;
        u->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
        adam_step<<<nblk(u->nparam), 256, 0, u->ctx->stream>>>(u->w, u->g, u->m, u->v, u->nparam,
                                                               (float)lr, b1, b2, 1e-8f, c1, c2);
        u->wb_current = false;
        u->wt_current = false;
        SGPU_HIP(hipGetLastError());
    }
    SGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    return SGPU_OK;
}
