// Hamiltonian apply, Chebyshev filter, Rayleigh-Ritz projection, FP64
// generalized eigensolve, subspace rotation, Lanczos bounds and one CheFSI
// pass, plus the host-pointer drop-ins of the reference entry points.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sgpu_internal.h"

namespace {

int check_block(const sgpu_context *ctx, const sgpu_block *b, const char *what)
{
    if (!b || b->ctx != ctx) {
        sgpu_set_error("%s: block missing or from another context", what);
        return SGPU_ERR_ARG;
    }
    return SGPU_OK;
}

int check_pair(const sgpu_context *ctx, const sgpu_block *a, const sgpu_block *b, const char *what)
{
    SGPU_CHECK(check_block(ctx, a, what));
    SGPU_CHECK(check_block(ctx, b, what));
    if (a->ncol != b->ncol || a->dtype != b->dtype) {
        sgpu_set_error("%s: blocks differ in shape or type", what);
        return SGPU_ERR_ARG;
    }
    return SGPU_OK;
}

const void *veff_for(const sgpu_context *ctx, int dtype)
{
    return dtype == SGPU_F64 ? (const void *)ctx->veff64 : (const void *)ctx->veff32;
}

// out = s1 * (H + c) x + s2 * z   (z may alias out; x must not).
int apply_h(sgpu_context *ctx, int dtype, int ncol, double c, const void *x, long ldx, void *out,
            long ldo, const void *z, long ldz, double s1, double s2)
{
    if (!ctx->has_veff) {
        sgpu_set_error("Hamiltonian apply before sgpu_set_veff");
        return SGPU_ERR_STATE;
    }
    sgpu_tmark t;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_CHECK(sgpu_launch_stencil(ctx, dtype, ncol, -0.5, 1.0, c, veff_for(ctx, dtype), x, ldx, out,
                                   ldo, z, ldz, s1, s2, true));
    SGPU_CHECK(sgpu_tend(ctx, &t, "stencil"));
    if (ctx->nl.n_infl > 0) {
        SGPU_CHECK(sgpu_tbegin(ctx, &t));
        SGPU_CHECK(sgpu_launch_vnl(ctx, dtype, ncol, x, ldx, out, ldo, s1));
        SGPU_CHECK(sgpu_tend(ctx, &t, "nonlocal"));
    }
    return SGPU_OK;
}

// apply_h for one step of a chain in which each out is the next x (the
// Chebyshev recurrence): with want_next the nonlocal partials of out are
// formed inside the nonlocal apply and the next step reuses them instead of
// re-reading its input (*have_part carries that from step to step).
int apply_h_chain(sgpu_context *ctx, int dtype, int ncol, double c, const void *x, long ldx,
                  void *out, long ldo, const void *z, long ldz, double s1, double s2,
                  bool *have_part, bool want_next)
{
    if (!ctx->has_veff) {
        sgpu_set_error("Hamiltonian apply before sgpu_set_veff");
        return SGPU_ERR_STATE;
    }
    sgpu_tmark t;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_CHECK(sgpu_launch_stencil(ctx, dtype, ncol, -0.5, 1.0, c, veff_for(ctx, dtype), x, ldx, out,
                                   ldo, z, ldz, s1, s2, true));
    SGPU_CHECK(sgpu_tend(ctx, &t, "stencil"));
    bool projected = false;
    if (ctx->nl.n_infl > 0) {
        SGPU_CHECK(sgpu_tbegin(ctx, &t));
        SGPU_CHECK(sgpu_launch_vnl_chain(ctx, dtype, ncol, x, ldx, out, ldo, s1, *have_part,
                                         want_next && ctx->nl_chain, &projected));
        SGPU_CHECK(sgpu_tend(ctx, &t, "nonlocal"));
    }
    *have_part = projected;
    return SGPU_OK;
}

// Chebyshev filter on raw device storage (ChebyshevFiltering,
// src/eigenSolver.c:722-798).  On return *res points at the buffer holding
// Y_m and *prev at the one holding Y_{m-1}; bx starts as X, by as Y.
int filter_raw(sgpu_context *ctx, int dtype, int ncol, void *bx, void *by, long ld, int m,
               double a, double b, double a0, void **res, void **prev)
{
    const double e = 0.5 * (b - a), c = 0.5 * (b + a);
    double sigma = e / (a0 - c);
    const double sigma1 = sigma, gamma = 2.0 / sigma1;
    sgpu_tmark tall;
    SGPU_CHECK(sgpu_tbegin(ctx, &tall));
    // Y = (sigma1 / e) (H - c) X
    bool have_part = false;
    SGPU_CHECK(apply_h_chain(ctx, dtype, ncol, -c, bx, ld, by, ld, nullptr, 0, sigma1 / e, 0.0,
                             &have_part, m > 1));
    void *P = by, *Q = bx;
    for (int j = 1; j < m; j++) {
        const double sigma2 = 1.0 / (gamma - sigma);
        const double vs = 2.0 * sigma2 / e, vs2 = sigma * sigma2;
        // Q <- vs (H - c) P - vs2 Q   (in place over Y_{j-1})
        SGPU_CHECK(apply_h_chain(ctx, dtype, ncol, -c, P, ld, Q, ld, Q, ld, vs, -vs2, &have_part,
                                 j + 1 < m));
        std::swap(P, Q);
        sigma = sigma2;
    }
    SGPU_CHECK(sgpu_tend(ctx, &tall, "filter"));
    *res = P;
    *prev = Q;
    return SGPU_OK;
}

// Sum a [nslice][n*n] partial stack into an FP64 n*n matrix.
template <typename P>
__global__ void sum_partials(const P *__restrict__ part, int nslice, long nn,
                             double *__restrict__ out)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    // slices in order, the loads issued eight at a time
    double s = 0.0;
    int k = 0;
    for (; k + 8 <= nslice; k += 8) {
        P v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[(long)(k + u) * nn + t];
#pragma unroll
        for (int u = 0; u < 8; u++) s += (double)v[u];
    }
    for (; k < nslice; k++) s += (double)part[(long)k * nn + t];
    out[t] = s;
}

// Rows per K-slice of the split-K Gram product.  A TN product with n = Ns
// and K = Nd has only (n/64)^2 output tiles, far too few for 256 CUs when
// Ns is small (rocBLAS ran 140k x 160 in FP64 at under 1 TFLOP/s unsplit),
// so K is cut into slices until there are about 256 (FP64) or 2048 (FP32)
// tile-slices.  FP64 products with n >= 384 run unsplit: rocBLAS reached
// 45 TFLOP/s on 474k x 480 unsplit but 5-30x less with 4-32 slices.  FP32
// slices are capped at 8192 rows: FP32 accumulation stays inside a slice and
// the slices are summed in FP64.
long gram_kslice(int n, long nd, int dtype)
{
    const long tiles = std::max(1L, ((long)(n + 63) / 64) * ((n + 63) / 64));
    if (dtype == SGPU_F64 && n >= 384) return std::max(nd, 1L);
    const long target = dtype == SGPU_F32 ? 2048 : 256;
    long nslice = std::max(1L, (target + tiles - 1) / tiles);
    long ks = (nd + nslice - 1) / nslice;
    ks = std::max(1024L, (ks + 255) / 256 * 256);
    if (dtype == SGPU_F32) ks = std::min(ks, 8192L);
    return std::min(ks, std::max(nd, 1L));
}

// ---- FP32 Gram products from BF16 pieces on matrix cores ------------------
//
// C = A' B over one K-slice per block row.  Every FP32 operand is split into
// three BF16 pieces, a = a0 + a1 + a2 (a1 = bf16(a - a0), a2 = bf16(a - a0 -
// a1): 24 significant bits, the FP32 significand), and
//   C += A0'B0 + A0'B1 + A1'B0 + A0'B2 + A2'B0 + A1'B1
// on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation; the dropped products
// are below 2^-24 of |a||b|.  (Two pieces, three products, carry only ~17
// bits and stalled the SiC SCF above TOL_SCF = 1e-6.)  The slices are summed
// in FP64 by sum_partials.  Block: 64 x 64 output tile, 4 waves of 2 x 2 MFMA
// tiles; K staged through LDS 32 rows at a time, coalesced along the
// columns, stored k-contiguous so that each lane's operand is one 8-byte LDS
// read.  Operand layout as in the U-Net kernels: A(i,k) lane i + 16(k/4)
// element k%4, B(k,n) lane n + 16(k/4), D(i,n) lane n + 16(i/4) register i%4.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int kGT = 64;        // output tile
constexpr int kGK = 32;        // K rows per stage
constexpr int kGLD = kGK + 4;  // LDS stride (elements) of a staged column

__device__ __forceinline__ unsigned short bf16_bits(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// Two FP32 values to a BF16 pair (x in the low half), rounded to nearest
// even: one v_cvt_pk_bf16_f32 on gfx950, the same bits as bf16_bits.
typedef __bf16 bf16x2_pk_t __attribute__((ext_vector_type(2)));
typedef float f32x2_pk_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned bf16_pk(float x, float y)
{
    const f32x2_pk_t f = {x, y};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f, bf16x2_pk_t));
}

// split3 of two values at once, each piece as a packed pair (x low): three
// conversions, four unpacks and four subtractions for the two.
__device__ __forceinline__ void split3_pk(float x, float y, unsigned &h, unsigned &m, unsigned &l)
{
    h = bf16_pk(x, y);
    const float rx = x - __uint_as_float(h << 16), ry = y - __uint_as_float(h & 0xffff0000u);
    m = bf16_pk(rx, ry);
    l = bf16_pk(rx - __uint_as_float(m << 16), ry - __uint_as_float(m & 0xffff0000u));
}


__global__ __launch_bounds__(256) void gram_bf16x3(const float *__restrict__ A, long lda,
                                                   const float *__restrict__ B, long ldb, int n,
                                                   long nd, long kslice, float *__restrict__ part)
{
    __shared__ __align__(16) unsigned short sa[3][kGT * kGLD], sb[3][kGT * kGLD];  // [piece][col][k]
    const int tilesn = (n + kGT - 1) / kGT;
    const int ti = blockIdx.x % tilesn, tj = blockIdx.x / tilesn;
    const int slice = blockIdx.y;
    const long k0 = (long)slice * kslice, k1 = min(nd, k0 + kslice);
    const int i0 = ti * kGT, j0 = tj * kGT;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    f32x4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (long kb = k0; kb < k1; kb += kGK) {
        // stage: element e = (col, k) with k fastest across threads (coalesced)
#pragma unroll
        for (int r = 0; r < kGT * kGK / 256; r++) {
            const int e = threadIdx.x + r * 256;
            const int kk = e % kGK, c = e / kGK;
            const long k = kb + kk;
            float va = 0.f, vb = 0.f;
            if (k < k1) {
                if (i0 + c < n) va = A[(long)(i0 + c) * lda + k];
                if (j0 + c < n) vb = B[(long)(j0 + c) * ldb + k];
            }
            const int o = c * kGLD + kk;
            unsigned h, m, l;
            split3_pk(va, vb, h, m, l);
            sa[0][o] = (unsigned short)h;
            sa[1][o] = (unsigned short)m;
            sa[2][o] = (unsigned short)l;
            sb[0][o] = (unsigned short)(h >> 16);
            sb[1][o] = (unsigned short)(m >> 16);
            sb[2][o] = (unsigned short)(l >> 16);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kGK; ks += 16) {
            bf16x4_t av[3][2], bv[3][2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int ca = (wi + t * 16 + li) * kGLD + ks + 4 * g;
                const int cb = (wj + t * 16 + li) * kGLD + ks + 4 * g;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    av[q][t] = *reinterpret_cast<const bf16x4_t *>(&sa[q][ca]);
                    bv[q][t] = *reinterpret_cast<const bf16x4_t *>(&sb[q][cb]);
                }
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    // smallest products first
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[1][a], bv[1][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[0][a], bv[2][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[2][a], bv[0][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[0][a], bv[1][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[1][a], bv[0][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[0][a], bv[0][b], acc[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    float *pp = part + (long)slice * n * n;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wi + a * 16 + 4 * g + r;
                const int j = j0 + wj + b * 16 + li;
                if (i < n && j < n) pp[(long)j * n + i] = acc[a][b][r];
            }
}

// ---- FP32 subspace rotation from BF16 pieces on gfx950 matrix cores -------
//
// Psi = Y Q (DP_Subspace_Rotation, src/eigenSolver.c:1386-1439) for FP32
// blocks: every value is split into three BF16 pieces (as in gram_bf16x3) and
// the six products with i + j <= 2 are summed on v_mfma_f32_16x16x32_bf16
// (16 x 16 output tile, K = 32 per instruction, FP32 accumulation).  One block
// takes a PTS-point strip of Y and produces all n output columns, so Y is read
// and split once: the strip's 32-column K-chunks are staged in LDS as
// [piece][point][k] BF16 (the A operands, 16-byte reads), while Q comes as
// three BF16 planes prepared once per rotation (split_q_bf16x3), column-major
// with K padded to a multiple of 32 (the B operands, 16-byte loads from L2,
// issued with the chunk's Y loads).  The eight waves share the strip and take
// every eighth 16-column output tile.  Every block reads all of Q, so the
// strip sets the L2 traffic for Q: 128-point strips (253 VGPRs, still two
// waves per SIMD) took 1.45 ms at 474,552 x 480 on MI355X against 1.82 ms
// for 64-point strips.
// Operand layout of the 16x16x32 instruction: A(i, k) in lane i + 16 (k / 8),
// element k % 8; B(k, n) in lane n + 16 (k / 8), element k % 8; D(i, n) in
// lane n + 16 (i / 4), register i % 4, so each lane's four results are four
// consecutive points of one column (one 16-byte store).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
constexpr int kRotPts = 128;       // points per block
constexpr int kRotK = 32;          // K per chunk
constexpr int kRotLd = kRotK + 8;  // LDS row (bf16): 80 bytes
constexpr int kRotWaves = 8;       // waves per block
constexpr int kRotMaxCT = 4;       // output column tiles per wave (n <= 512)

__device__ __forceinline__ void split3(float v, unsigned short &p0, unsigned short &p1,
                                       unsigned short &p2)
{
    p0 = bf16_bits(v);
    const float r1 = v - bf16_val(p0);
    p1 = bf16_bits(r1);
    p2 = bf16_bits(r1 - bf16_val(p1));
}

// Q (n x n FP64, column-major) -> three BF16 planes qp[piece][c * nk + k],
// zero for k >= n (nk = n rounded up to kRotK).
__global__ void split_q_bf16x3(const double *__restrict__ q, int n, int nk,
                               unsigned short *__restrict__ qp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)n * nk;
    if (t >= tot) return;
    const int c = (int)(t / nk), k = (int)(t - (long)c * nk);
    unsigned short p0 = 0, p1 = 0, p2 = 0;
    if (k < n) split3((float)q[(long)c * n + k], p0, p1, p2);
    qp[t] = p0;
    qp[tot + t] = p1;
    qp[2 * tot + t] = p2;
}

template <int PTS>
__global__ __launch_bounds__(64 * kRotWaves) void rot_bf16x3(const float *__restrict__ Y, long ldy,
                                                  const unsigned short *__restrict__ qp, int n,
                                                  int nk, long nd, float *__restrict__ out,
                                                  long ldo)
{
    constexpr int NPT = PTS / 16;   // 16-point tiles per block
    __shared__ __align__(16) unsigned short sa[3][PTS * kRotLd];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const long p0 = (long)blockIdx.x * PTS;
    const int ntile = (n + 15) / 16;
    const long plane = (long)n * nk;
    f32x4_t acc[NPT][kRotMaxCT];
#pragma unroll
    for (int pt = 0; pt < NPT; pt++)
#pragma unroll
        for (int ct = 0; ct < kRotMaxCT; ct++) acc[pt][ct] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    // Y[p0 .. p0+PTS-1][k0 .. k0+31] into registers: consecutive threads on
    // consecutive points (256-byte column segments), all loads first; the
    // next chunk's loads are issued once this chunk is staged, so they are in
    // flight during its MFMAs
    constexpr int NT = 64 * kRotWaves;
    float v[PTS * kRotK / NT];
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < PTS * kRotK / NT; r++) {
            const int e = threadIdx.x + NT * r;
            const int kk = e / PTS, pp = e % PTS;
            const bool ok = k0 + kk < n && p0 + pp < nd;
            v[r] = Y[ok ? (long)(k0 + kk) * ldy + p0 + pp : 0L];
            if (!ok) v[r] = 0.f;
        }
    };
    load(0);
    for (int k0 = 0; k0 < nk; k0 += kRotK) {
        // this wave's Q pieces of the chunk, one 16-column tile per ct
        bf16x8_t bv[kRotMaxCT][3];
#pragma unroll
        for (int ct = 0; ct < kRotMaxCT; ct++) {
            const int col = min((wave + kRotWaves * ct) * 16 + li, n - 1);
#pragma unroll
            for (int q = 0; q < 3; q++)
                bv[ct][q] = *reinterpret_cast<const bf16x8_t *>(qp + q * plane + (long)col * nk + k0 + 8 * g);
        }
        __syncthreads();  // the previous chunk's operands have been read
        static_assert((PTS * kRotK / NT) % 2 == 0, "values split in pairs");
#pragma unroll
        for (int r = 0; r < PTS * kRotK / NT; r += 2) {
            const int e0 = threadIdx.x + NT * r, e1 = e0 + NT;
            const int o0 = (e0 % PTS) * kRotLd + e0 / PTS, o1 = (e1 % PTS) * kRotLd + e1 / PTS;
            unsigned h, m, l;
            split3_pk(v[r], v[r + 1], h, m, l);
            sa[0][o0] = (unsigned short)h;
            sa[1][o0] = (unsigned short)m;
            sa[2][o0] = (unsigned short)l;
            sa[0][o1] = (unsigned short)(h >> 16);
            sa[1][o1] = (unsigned short)(m >> 16);
            sa[2][o1] = (unsigned short)(l >> 16);
        }
        __syncthreads();
        if (k0 + kRotK < nk) load(k0 + kRotK);
#pragma unroll
        for (int pt = 0; pt < NPT; pt++) {
            bf16x8_t av[3];
#pragma unroll
            for (int q = 0; q < 3; q++)
                av[q] = *reinterpret_cast<const bf16x8_t *>(&sa[q][(pt * 16 + li) * kRotLd + 8 * g]);
#pragma unroll
            for (int ct = 0; ct < kRotMaxCT; ct++) {
                if (wave + kRotWaves * ct >= ntile) break;
                // smallest products first
                f32x4_t c = acc[pt][ct];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1], bv[ct][1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[ct][2], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[2], bv[ct][0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[ct][1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1], bv[ct][0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[ct][0], c, 0, 0, 0);
                acc[pt][ct] = c;
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < kRotMaxCT; ct++) {
        const int tile = wave + kRotWaves * ct;
        if (tile >= ntile) break;
        const int col = tile * 16 + li;
        if (col >= n) continue;
#pragma unroll
        for (int pt = 0; pt < NPT; pt++) {
            const long p = p0 + pt * 16 + 4 * g;
            float *o = out + (long)col * ldo + p;
            if (p + 3 < nd && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
                *reinterpret_cast<float4 *>(o) =
                    make_float4(acc[pt][ct][0], acc[pt][ct][1], acc[pt][ct][2], acc[pt][ct][3]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (p + r < nd) o[r] = acc[pt][ct][r];
            }
        }
    }
}

// FP32 Gram product C = A'B on v_mfma_f32_16x16x32_bf16 from three BF16
// pieces per value (six products, as gram_bf16x3), for the large state counts
// at which the per-block split of gram_bf16x3's 64 x 64 tiles lost to rocBLAS:
// 128 x 128 output tiles (each staged value is split once per 128 output
// columns), eight waves of 32 x 64, K staged 64 points at a time as
// [piece][column][k] BF16 in LDS (two MFMA K-steps per chunk, 16-byte operand
// reads).  Each thread stages four consecutive points of a column: one
// 16-byte load per operand and quad, one 8-byte LDS store per piece; the next
// chunk's loads are issued before this chunk's MFMAs.  SYM (A == B): only the
// tiles on and above the diagonal are computed and each off-diagonal one is
// also written transposed.  One block per (tile, K-slice); the slices are
// summed in FP64 by sum_partials.  Needs lda, ldb multiples of 4 and 16-byte
// aligned operands (the caller checks).
constexpr int kG3T = 128;          // output tile
constexpr int kG3K = 64;           // K per chunk
constexpr int kG3Ld = kG3K + 8;    // LDS row (bf16): 144 bytes
constexpr size_t kG3Shm = 2 * 3 * (size_t)kG3T * kG3Ld * sizeof(unsigned short);

template <bool SYM>
__global__ __launch_bounds__(512) void gram3_bf16(const float *__restrict__ A, long lda,
                                                  const float *__restrict__ B, long ldb, int n,
                                                  long nd, long kslice, float *__restrict__ part)
{
    extern __shared__ __align__(16) unsigned short g3sm[];
    unsigned short *sa = g3sm, *sb = g3sm + 3 * kG3T * kG3Ld;   // [piece][column][k]
    const int tiles = (n + kG3T - 1) / kG3T;
    // the tiles of one K-slice are consecutive logical ids -> one XCD, so
    // the slice's columns they share are L2 hits
    const int lid = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tid_blk = lid % gridDim.x;
    int ti, tj;
    if (SYM) {  // upper-triangle tile index -> (ti, tj), ti <= tj
        int t = tid_blk;
        ti = 0;
        while (t >= tiles - ti) {
            t -= tiles - ti;
            ti++;
        }
        tj = ti + t;
    } else {
        ti = tid_blk % tiles;
        tj = tid_blk / tiles;
    }
    const int slice = lid / gridDim.x;
    const long k0 = (long)slice * kslice, k1 = min(nd, k0 + kslice);
    const int i0 = ti * kG3T, j0 = tj * kG3T;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 64;  // wave's 32 x 64 sub-tile
    f32x4_t acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    // quad e = (column, k / 4), k fastest: 16 threads cover a column's 64
    // points (256 contiguous bytes)
    constexpr int PER = kG3T * kG3K / 4 / 512;
    float4 va[PER], vb[PER];
    auto load = [&](long kb) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + 512 * r;
            const int kq = 4 * (e % (kG3K / 4)), c = e / (kG3K / 4);
            const long k = kb + kq;
            // k is a multiple of 4 below nd, so k + 3 stays inside the column's
            // leading dimension; points at or past k1 are zeroed
            const bool oka = k < k1 && i0 + c < n, okb = k < k1 && j0 + c < n;
            va[r] = *reinterpret_cast<const float4 *>(A + (oka ? (long)(i0 + c) * lda + k : 0L));
            vb[r] = *reinterpret_cast<const float4 *>(B + (okb ? (long)(j0 + c) * ldb + k : 0L));
            const int nv = (int)min(4L, k1 - k);
            if (!oka) va[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!okb) vb[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (nv < 4) { va[r].w = 0.f; vb[r].w = 0.f; }
            if (nv < 3) { va[r].z = 0.f; vb[r].z = 0.f; }
            if (nv < 2) { va[r].y = 0.f; vb[r].y = 0.f; }
        }
    };
    auto put = [&](unsigned short *dst, int o, float4 v) {
        unsigned h0, m0, l0, h1, m1, l1;
        split3_pk(v.x, v.y, h0, m0, l0);
        split3_pk(v.z, v.w, h1, m1, l1);
        *reinterpret_cast<uint2 *>(dst + o) = make_uint2(h0, h1);
        *reinterpret_cast<uint2 *>(dst + kG3T * kG3Ld + o) = make_uint2(m0, m1);
        *reinterpret_cast<uint2 *>(dst + 2 * kG3T * kG3Ld + o) = make_uint2(l0, l1);
    };
    load(k0);
    for (long kb = k0; kb < k1; kb += kG3K) {
        __syncthreads();  // the previous chunk's operands have been read
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + 512 * r;
            const int kq = 4 * (e % (kG3K / 4)), c = e / (kG3K / 4);
            const int o = c * kG3Ld + kq;
            put(sa, o, va[r]);
            put(sb, o, vb[r]);
        }
        __syncthreads();
        if (kb + kG3K < k1) load(kb + kG3K);
#pragma unroll
        for (int ks = 0; ks < kG3K; ks += 32) {
            bf16x8_t av[2][3], bv[4][3];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int q = 0; q < 3; q++)
                    av[a][q] = *reinterpret_cast<const bf16x8_t *>(
                        &sa[q * kG3T * kG3Ld + (wi + a * 16 + li) * kG3Ld + ks + 8 * g]);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int q = 0; q < 3; q++)
                    bv[b][q] = *reinterpret_cast<const bf16x8_t *>(
                        &sb[q * kG3T * kG3Ld + (wj + b * 16 + li) * kG3Ld + ks + 8 * g]);
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    f32x4_t c = acc[a][b];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[b][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[b][2], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][2], bv[b][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[b][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[b][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[b][0], c, 0, 0, 0);
                    acc[a][b] = c;
                }
        }
    }
    float *pp = part + (long)slice * n * n;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wi + a * 16 + 4 * g + r;
                const int j = j0 + wj + b * 16 + li;
                if (i < n && j < n) {
                    pp[(long)j * n + i] = acc[a][b][r];
                    if (SYM && ti != tj) pp[(long)i * n + j] = acc[a][b][r];
                }
            }
}

// Mp = Y'Y and Hp = Y'(HY) in one launch (both symmetric: the same upper
// tiles): gram3_bf16<true>'s tiles, waves and products, with the block's
// column tile of Y split and staged once for both and the two B tiles (Y
// and HY) beside it.  Three staged tiles need 32-point chunks to fit the LDS
// (92 KB); a chunk is one MFMA K-step for each of the two products, so a
// barrier pair still covers 96 products per wave.  The K-steps, slices and
// the order of the six products per step are those of gram3_bf16 with the
// same slice length, so each of Mp and Hp is bitwise what the single-product
// kernel gives.
constexpr int kG3DK = 32;           // K per chunk
constexpr int kG3DLd = kG3DK + 8;   // LDS row (bf16): 80 bytes, conflict-free 16-byte reads
constexpr size_t kG3DShm = 3 * 3 * (size_t)kG3T * kG3DLd * sizeof(unsigned short);

__global__ __launch_bounds__(512) void gram3x2_bf16(const float *__restrict__ Y,
                                                    const float *__restrict__ HY, long ld, int n,
                                                    long nd, long kslice,
                                                    float *__restrict__ partM,
                                                    float *__restrict__ partH)
{
    extern __shared__ __align__(16) unsigned short g3dsm[];
    constexpr int PL = kG3T * kG3DLd;   // one piece of one tile
    unsigned short *sa = g3dsm, *sb = g3dsm + 3 * PL, *sc = g3dsm + 6 * PL;
    const int tiles = (n + kG3T - 1) / kG3T;
    const int lid = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    int t = lid % gridDim.x, ti = 0;
    while (t >= tiles - ti) {
        t -= tiles - ti;
        ti++;
    }
    const int tj = ti + t;
    const int slice = lid / gridDim.x;
    const long k0 = (long)slice * kslice, k1 = min(nd, k0 + kslice);
    const int i0 = ti * kG3T, j0 = tj * kG3T;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 64;
    f32x4_t am[2][4], ah[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            am[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
            ah[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
    constexpr int PER = kG3T * kG3DK / 4 / 512;
    float4 va[PER], vb[PER], vc[PER];
    auto load = [&](long kb) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + 512 * r;
            const int kq = 4 * (e % (kG3DK / 4)), c = e / (kG3DK / 4);
            const long k = kb + kq;
            const bool oka = k < k1 && i0 + c < n, okb = k < k1 && j0 + c < n;
            va[r] = *reinterpret_cast<const float4 *>(Y + (oka ? (long)(i0 + c) * ld + k : 0L));
            vb[r] = *reinterpret_cast<const float4 *>(Y + (okb ? (long)(j0 + c) * ld + k : 0L));
            vc[r] = *reinterpret_cast<const float4 *>(HY + (okb ? (long)(j0 + c) * ld + k : 0L));
            const int nv = (int)min(4L, k1 - k);
            if (!oka) va[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!okb) { vb[r] = make_float4(0.f, 0.f, 0.f, 0.f); vc[r] = vb[r]; }
            if (nv < 4) { va[r].w = 0.f; vb[r].w = 0.f; vc[r].w = 0.f; }
            if (nv < 3) { va[r].z = 0.f; vb[r].z = 0.f; vc[r].z = 0.f; }
            if (nv < 2) { va[r].y = 0.f; vb[r].y = 0.f; vc[r].y = 0.f; }
        }
    };
    auto put = [&](unsigned short *dst, int o, float4 v) {
        unsigned h0, m0, l0, h1, m1, l1;
        split3_pk(v.x, v.y, h0, m0, l0);
        split3_pk(v.z, v.w, h1, m1, l1);
        *reinterpret_cast<uint2 *>(dst + o) = make_uint2(h0, h1);
        *reinterpret_cast<uint2 *>(dst + PL + o) = make_uint2(m0, m1);
        *reinterpret_cast<uint2 *>(dst + 2 * PL + o) = make_uint2(l0, l1);
    };
    load(k0);
    for (long kb = k0; kb < k1; kb += kG3DK) {
        __syncthreads();  // the previous chunk's operands have been read
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + 512 * r;
            const int kq = 4 * (e % (kG3DK / 4)), c = e / (kG3DK / 4);
            const int o = c * kG3DLd + kq;
            put(sa, o, va[r]);
            put(sb, o, vb[r]);
            put(sc, o, vc[r]);
        }
        __syncthreads();
        if (kb + kG3DK < k1) load(kb + kG3DK);
        bf16x8_t av[2][3];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int q = 0; q < 3; q++)
                av[a][q] = *reinterpret_cast<const bf16x8_t *>(
                    &sa[q * PL + (wi + a * 16 + li) * kG3DLd + 8 * g]);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            bf16x8_t bv[3], cv[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                bv[q] = *reinterpret_cast<const bf16x8_t *>(&sb[q * PL + (wj + b * 16 + li) * kG3DLd + 8 * g]);
                cv[q] = *reinterpret_cast<const bf16x8_t *>(&sc[q * PL + (wj + b * 16 + li) * kG3DLd + 8 * g]);
            }
#pragma unroll
            for (int a = 0; a < 2; a++) {
                f32x4_t c = am[a][b];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[2], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][2], bv[0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[0], c, 0, 0, 0);
                am[a][b] = c;
                f32x4_t d = ah[a][b];
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], cv[1], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], cv[2], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][2], cv[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], cv[1], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], cv[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], cv[0], d, 0, 0, 0);
                ah[a][b] = d;
            }
        }
    }
    const long off = (long)slice * n * n;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wi + a * 16 + 4 * g + r;
                const int j = j0 + wj + b * 16 + li;
                if (i < n && j < n) {
                    partM[off + (long)j * n + i] = am[a][b][r];
                    partH[off + (long)j * n + i] = ah[a][b][r];
                    if (ti != tj) {
                        partM[off + (long)i * n + j] = am[a][b][r];
                        partH[off + (long)i * n + j] = ah[a][b][r];
                    }
                }
            }
}

template <typename T>
rocblas_status gemm_tn_batched(rocblas_handle h, int n, long k, const T *A, long lda, long sa,
                               const T *B, long ldb, long sb, T *C, long sc, int batch);
template <>
rocblas_status gemm_tn_batched<double>(rocblas_handle h, int n, long k, const double *A, long lda,
                                       long sa, const double *B, long ldb, long sb, double *C,
                                       long sc, int batch)
{
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, n,
                                         n, (rocblas_int)k, &one, A, (rocblas_int)lda, sa, B,
                                         (rocblas_int)ldb, sb, &zero, C, n, sc, batch);
}
template <>
rocblas_status gemm_tn_batched<float>(rocblas_handle h, int n, long k, const float *A, long lda,
                                      long sa, const float *B, long ldb, long sb, float *C,
                                      long sc, int batch)
{
    const float one = 1.f, zero = 0.f;
    return rocblas_sgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, n,
                                         n, (rocblas_int)k, &one, A, (rocblas_int)lda, sa, B,
                                         (rocblas_int)ldb, sb, &zero, C, n, sc, batch);
}

// SPARC_AMD_ROT=rocblas: rocBLAS sgemm for the FP32 rotation instead of rot_bf16x3.
int rot_mode()
{
    static const int v = [] {
        const char *e = getenv("SPARC_AMD_ROT");
        return (e && !strcmp(e, "rocblas")) ? 1 : 0;
    }();
    return v;
}

// SPARC_AMD_GRAM=rocblas: rocBLAS for every FP32 Gram; =bf16x3: the 64 x 64
// BF16-piece kernel for every FP32 Gram; unset: the 64 x 64 kernel below 256
// columns, the 128 x 128 one (gram3_bf16) from 256.
int gram_mode()
{
    static const int v = [] {
        const char *e = getenv("SPARC_AMD_GRAM");
        if (e && !strcmp(e, "rocblas")) return 1;
        if (e && !strcmp(e, "bf16x3")) return 2;
        return 0;
    }();
    return v;
}

// symmetric: C is symmetric (Y'Y, or Y'HY with H symmetric); the 128 x 128
// kernel then forms the tiles on and above the diagonal only and mirrors
// them, which is what the reference's dsygvd reads of Hp and Mp ('U').
// K-slice of gram3_bf16 (one 512-thread block per CU): the number of
// (tile, slice) blocks is chosen so that they fill whole rounds of the CUs,
// minimising rounds x slice rows with slices of at most 8192 rows (multiples
// of the 32-point chunk).  216-atom cell (474,552 rows, 10 tiles): 102
// slices of 4,672 rows in four full rounds instead of 58 of 8,192 in three
// rounds with the last one 27 % full (Mp + Hp 4.19 -> 3.30 ms on MI355X).
static long gram3_kslice(const sgpu_context *ctx, long nblk, long nd)
{
    const long cap = 8192, cus = std::max(1, ctx->n_cu);
    const long ns0 = std::max(1L, (nd + cap - 1) / cap);
    long best = std::min(cap, std::max(nd, 1L)), best_cost = -1;
    for (long ns = ns0; ns <= 2 * ns0 + 8; ns++) {
        long ks = (nd + ns - 1) / ns;
        ks = (ks + kG3K - 1) / kG3K * kG3K;
        if (ks > cap) continue;
        const long nsl = (nd + ks - 1) / ks;
        const long cost = (nblk * nsl + cus - 1) / cus * ks;
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = ks;
        }
    }
    return best;
}

template <typename T>
int gram_tn_typed(sgpu_context *ctx, int n, long nd, const T *A, long lda, const T *B, long ldb,
                  double *C, bool symmetric)
{
    long kslice = gram_kslice(n, nd, sizeof(T) == 4 ? SGPU_F32 : SGPU_F64);
    const bool g3 = sizeof(T) == 4 && n >= 256 && gram_mode() == 0 && lda % 4 == 0 &&
                    ldb % 4 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(B) & 15) == 0;
    const int tiles3 = (n + kG3T - 1) / kG3T;
    const bool sym3 = symmetric || (A == B && lda == ldb);
    const unsigned nblk3 = sym3 ? (unsigned)(tiles3 * (tiles3 + 1) / 2) : (unsigned)(tiles3 * tiles3);
    if (g3) kslice = gram3_kslice(ctx, nblk3, nd);
    const int nslice = (int)((nd + kslice - 1) / kslice);
    const long nn = (long)n * n;
    void *buf = nullptr;
    SGPU_CHECK(sgpu_scratch(ctx, (size_t)nslice * nn * sizeof(T), &buf));
    T *part = (T *)buf;
    // Below 256 states the 64 x 64 BF16-piece kernel; from 256 states the
    // 128 x 128 one (512k x 480: 4.8 ms for Mp + Hp, rocBLAS sgemm 4.7 ms;
    // the 64 x 64 kernel took 15 ms there, bound by splitting every staged
    // value once per 64 output columns)
    if (g3) {
        // 128 x 128 tiles of the gfx950 BF16-piece kernel; Mp = Y'Y computes
        // the upper triangle of tiles only
        const unsigned nblk = nblk3;
        SGPU_HIP(hipFuncSetAttribute((const void *)gram3_bf16<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kG3Shm));
        SGPU_HIP(hipFuncSetAttribute((const void *)gram3_bf16<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kG3Shm));
        if (sym3)
            gram3_bf16<true><<<dim3(nblk, nslice), 512, kG3Shm, ctx->stream>>>(
                (const float *)A, lda, (const float *)B, ldb, n, nd, kslice, (float *)part);
        else
            gram3_bf16<false><<<dim3(nblk, nslice), 512, kG3Shm, ctx->stream>>>(
                (const float *)A, lda, (const float *)B, ldb, n, nd, kslice, (float *)part);
        SGPU_HIP(hipGetLastError());
        sum_partials<T><<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(part, nslice, nn, C);
        SGPU_HIP(hipGetLastError());
        return SGPU_OK;
    }
    if (sizeof(T) == 4 && ((n < 256 && gram_mode() == 0) || gram_mode() == 2)) {
        const int tilesn = (n + kGT - 1) / kGT;
        gram_bf16x3<<<dim3(tilesn * tilesn, nslice), 256, 0, ctx->stream>>>(
            (const float *)A, lda, (const float *)B, ldb, n, nd, kslice, (float *)part);
        SGPU_HIP(hipGetLastError());
        sum_partials<T><<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(part, nslice, nn, C);
        SGPU_HIP(hipGetLastError());
        return SGPU_OK;
    }
    SGPU_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host));
    const long full = nd / kslice;
    if (full > 0)
        SGPU_BLAS(gemm_tn_batched<T>(ctx->blas, n, kslice, A, lda, kslice, B, ldb, kslice, part, nn,
                                     (int)full));
    const long rem = nd - full * kslice;
    if (rem > 0)
        SGPU_BLAS(gemm_tn_batched<T>(ctx->blas, n, rem, A + full * kslice, lda, 0,
                                     B + full * kslice, ldb, 0, part + full * nn, nn, 1));
    sum_partials<T><<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(part, nslice, nn, C);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// C (n x n, FP64, device) = A' B over nd rows, split-K (FP64 sums of slices).
int gram_tn(sgpu_context *ctx, int dtype, int n, long nd, const void *A, long lda, const void *B,
            long ldb, double *C, bool symmetric)
{
    if (dtype == SGPU_F64)
        return gram_tn_typed<double>(ctx, n, nd, (const double *)A, lda, (const double *)B, ldb, C,
                                     symmetric);
    return gram_tn_typed<float>(ctx, n, nd, (const float *)A, lda, (const float *)B, ldb, C,
                                symmetric);
}

// Mp = Y'Y and Hp = Y'HY by gram3x2_bf16 (FP32 from 128 states: at 160
// states on the 64-atom cell 0.32 ms against 0.54 for the 64 x 64 kernel's
// two launches); false when it does not apply (SPARC_AMD_GRAM3_PAIR=0: two
// gram_tn launches instead).
static bool gram3_pair_ok(const sgpu_context *ctx, int dtype, int n, const void *Y,
                          const void *HY, long ld)
{
    return ctx->gram3_pair && dtype == SGPU_F32 && n >= 128 && gram_mode() == 0 && ld % 4 == 0 &&
           (reinterpret_cast<uintptr_t>(Y) & 15) == 0 && (reinterpret_cast<uintptr_t>(HY) & 15) == 0;
}

static int gram3_pair(sgpu_context *ctx, int n, long nd, const float *Y, const float *HY, long ld,
                      double *Mp, double *Hp)
{
    const int tiles = (n + kG3T - 1) / kG3T;
    const unsigned nblk = (unsigned)(tiles * (tiles + 1) / 2);
    const long kslice = gram3_kslice(ctx, nblk, nd);   // the single-product kernel's slices
    const int nslice = (int)((nd + kslice - 1) / kslice);
    const long nn = (long)n * n;
    void *buf = nullptr;
    SGPU_CHECK(sgpu_scratch(ctx, 2 * (size_t)nslice * nn * sizeof(float), &buf));
    float *pm = (float *)buf, *ph = pm + (size_t)nslice * nn;
    SGPU_HIP(hipFuncSetAttribute((const void *)gram3x2_bf16,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kG3DShm));
    gram3x2_bf16<<<dim3(nblk, nslice), 512, kG3DShm, ctx->stream>>>(Y, HY, ld, n, nd, kslice, pm,
                                                                   ph);
    SGPU_HIP(hipGetLastError());
    sum_partials<float><<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(pm, nslice, nn, Mp);
    sum_partials<float><<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(ph, nslice, nn, Hp);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// HY = H Y, then Mp = Y'Y at *dmat and Hp = Y'HY at *dmat + ncol^2 (FP64,
// device, in the context's reduction buffer).
int project_device(sgpu_context *ctx, int dtype, int ncol, const void *Y, void *HY, long ld,
                   double **dmat_out)
{
    sgpu_tmark t;
    SGPU_CHECK(apply_h(ctx, dtype, ncol, 0.0, Y, ld, HY, ld, nullptr, 0, 1.0, 0.0));
    double *dmat = nullptr;
    const long nn = (long)ncol * ncol;
    SGPU_CHECK(sgpu_reduce_buf(ctx, (size_t)(2 * nn), &dmat));
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    if (gram3_pair_ok(ctx, dtype, ncol, Y, HY, ld)) {
        SGPU_CHECK(gram3_pair(ctx, ncol, ctx->nd, (const float *)Y, (const float *)HY, ld, dmat,
                              dmat + nn));
    } else {
        SGPU_CHECK(gram_tn(ctx, dtype, ncol, ctx->nd, Y, ld, Y, ld, dmat, true));
        SGPU_CHECK(gram_tn(ctx, dtype, ncol, ctx->nd, Y, ld, HY, ld, dmat + nn, true));
    }
    SGPU_CHECK(sgpu_tend(ctx, &t, "project_gemm"));
    *dmat_out = dmat;
    return SGPU_OK;
}

int project_raw(sgpu_context *ctx, int dtype, int ncol, const void *Y, void *HY, long ld,
                double *Hp_host, double *Mp_host)
{
    double *dmat = nullptr;
    const long nn = (long)ncol * ncol;
    SGPU_CHECK(project_device(ctx, dtype, ncol, Y, HY, ld, &dmat));
    SGPU_HIP(hipMemcpyAsync(Mp_host, dmat, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipMemcpyAsync(Hp_host, dmat + nn, nn * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

// out = Y Q with Q (ns x ns, FP64) on the device.
int rotate_device(sgpu_context *ctx, int dtype, int ns, const void *Y, long ldy, const double *q64,
                  void *out, long ldo)
{
    sgpu_tmark t;
    const long nn = (long)ns * ns;
    void *qd = nullptr;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host));
    if (dtype == SGPU_F64) {
        const double one = 1.0, zero = 0.0;
        SGPU_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                                (rocblas_int)ctx->nd, ns, ns, &one, (const double *)Y,
                                (rocblas_int)ldy, q64, ns, &zero, (double *)out, (rocblas_int)ldo));
    } else if (rot_mode() != 1 && ns <= 16 * kRotWaves * kRotMaxCT) {
        // BF16-piece rotation on matrix cores (SPARC_AMD_ROT=rocblas: rocBLAS sgemm)
        const int nk = (ns + kRotK - 1) / kRotK * kRotK;
        SGPU_CHECK(sgpu_scratch(ctx, 3 * (size_t)ns * nk * sizeof(unsigned short), &qd));
        const long tot = (long)ns * nk;
        split_q_bf16x3<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
            q64, ns, nk, (unsigned short *)qd);
        SGPU_HIP(hipGetLastError());
        rot_bf16x3<kRotPts><<<(unsigned)((ctx->nd + kRotPts - 1) / kRotPts), 64 * kRotWaves, 0,
                               ctx->stream>>>((const float *)Y, ldy, (const unsigned short *)qd, ns,
                                              nk, ctx->nd, (float *)out, ldo);
        SGPU_HIP(hipGetLastError());
    } else {
        SGPU_CHECK(sgpu_scratch(ctx, nn * sizeof(float), &qd));
        SGPU_CHECK(sgpu_convert(ctx, q64, SGPU_F64, qd, SGPU_F32, nn));
        const float one = 1.f, zero = 0.f;
        SGPU_BLAS(rocblas_sgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                                (rocblas_int)ctx->nd, ns, ns, &one, (const float *)Y,
                                (rocblas_int)ldy, (const float *)qd, ns, &zero, (float *)out,
                                (rocblas_int)ldo));
    }
    SGPU_CHECK(sgpu_tend(ctx, &t, "rotate_gemm"));
    return SGPU_OK;
}

int rotate_raw(sgpu_context *ctx, int dtype, int ns, const void *Y, long ldy, const double *Q_host,
               void *out, long ldo)
{
    const long nn = (long)ns * ns;
    double *q64 = nullptr;
    SGPU_CHECK(sgpu_reduce_buf(ctx, (size_t)nn, &q64));
    SGPU_HIP(hipMemcpyAsync(q64, Q_host, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SGPU_CHECK(rotate_device(ctx, dtype, ns, Y, ldy, q64, out, ldo));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

// dsygvd on the device (FP64), in place: dA (Hp, n x n) receives Q, dB (Mp)
// is overwritten (by L^-1, or by U with SPARC_AMD_CHOL=rocsolver); lambda
// (host) the eigenvalues in ascending order.  Hp Q = Mp Q Lambda with
// Q' Mp Q = I (LAPACKE_dsygvd itype 1 in the reference,
// src/eigenSolver.c:1262-1375), solved as Mp = L L' (chol_panel), L^-1
// (dtrtri), C = L^-1 Hp L^-T (two dgemm), C Z = Z Lambda (tridiagonal
// reduction, dstedc, ormtr_lower_wave) and Q = L^-T Z (dgemm).  Same reduction as
// dsygvd; at n = 480 on MI355X rocSOLVER's dsygst takes 16 ms of dsygvd's
// 29 ms, its dpotrf 1.06 ms and each dtrsm 0.16-0.18 ms, against 0.02 ms for
// a dgemm of that size.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum of v over the wave, the same in every lane: DPP row shifts by 1, 2, 4
// and 8 (an inclusive scan within each row of 16 lanes), row_bcast15 and
// row_bcast31 across the rows, then lane 63's total broadcast by readlane.
// Fixed order, VALU only (a __shfl_xor butterfly compiles to six dependent
// ds_bpermute round trips through the LDS crossbar).
__device__ __forceinline__ double dpp_f64(double v, int ctrl_sel)
{
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffff), hi = (int)(b >> 32);
    switch (ctrl_sel) {
        case 0: lo = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, true);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, true); break;
        case 1: lo = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xf, 0xf, true);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xf, 0xf, true); break;
        case 2: lo = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xf, 0xf, true);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xf, 0xf, true); break;
        case 3: lo = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xf, 0xf, true);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xf, 0xf, true); break;
        case 4: lo = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xa, 0xf, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xa, 0xf, false); break;
        default: lo = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xc, 0xf, false);
                 hi = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xc, 0xf, false); break;
    }
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int CW>
__device__ __forceinline__ void wave_sum_f64(double (&d)[CW])
{
#pragma unroll
    for (int s = 0; s < 6; s++)
#pragma unroll
        for (int j = 0; j < CW; j++) d[j] += dpp_f64(d[j], s);
#pragma unroll
    for (int j = 0; j < CW; j++) d[j] = readlane_f64(d[j], 63);
}

// Householder reduction of a symmetric matrix to tridiagonal form inside one
// workgroup, the matrix held in LDS: LAPACK dsytd2 with UPLO = 'L' (the
// storage rocsolver_dormtr reads back): for i = 0..n-2, H(i) = I - tau v v'
// with v(0:i) = 0, v(i+1) = 1 and v(i+2:n) stored in A(i+2:n, i); e(i) =
// beta, d(i) = A(i,i).  Each step is dlarfg (beta = -sign(alpha) |(alpha, x)|,
// tau = (beta - alpha) / beta, x /= alpha - beta), w = tau A22 v,
// w -= (tau / 2)(w'v) v and the rank-2 update A22 -= v w' + w v' on the lower
// triangle.  rocsolver's dsytrd takes ~1,800 dependent kernels for n = 480
// (3.7 us each on the device); here every step is a few workgroup barriers.
// The lower triangle is packed by columns: A(r, c), r >= c, at
// c (2n - c + 1) / 2 + (r - c).  Sums use a fixed order (deterministic).
constexpr int kTridMaxN = 180;   // packed triangle + work vectors fit in 160 KB of LDS
constexpr int kTridThreads = 1024;

__device__ __forceinline__ int tri_at(int n, int r, int c) { return c * (2 * n - c + 1) / 2 + (r - c); }

// Sum over the workgroup in a fixed order (DPP wave sums, then the 16 wave
// sums in order); the total is valid in thread 0.
__device__ __forceinline__ double tri_block_sum(double v, double *red)
{
    double a[1] = {v};
    wave_sum_f64<1>(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a[0];
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kTridThreads / 64; k++) s += red[k];
    return s;
}

__global__ __launch_bounds__(kTridThreads) void sytd2_lower_lds(double *__restrict__ A, int lda,
                                                                  int n, double *__restrict__ D,
                                                                  double *__restrict__ E,
                                                                  double *__restrict__ tau)
{
    extern __shared__ double sm[];
    const int npk = n * (n + 1) / 2;
    double *P = sm;                 // packed lower triangle
    double *v = P + npk;            // Householder vector of the current step (n)
    double *w = v + n;              // w (n)
    double *part = w + n;           // partial sums (kTridThreads)
    __shared__ double sc[4];        // tau, scale, beta, dot
    const int t = threadIdx.x;
    for (int e = t; e < n * n; e += kTridThreads) {
        const int c = e / n, r = e - c * n;
        if (r >= c) P[tri_at(n, r, c)] = A[(long)c * lda + r];
    }
    __syncthreads();
    for (int i = 0; i + 1 < n; i++) {
        const int m = n - i - 1;               // length of x = A(i+1:n, i)
        const int col = tri_at(n, i + 1, i);   // packed index of A(i+1, i)
        // xnorm^2 over A(i+2:n, i): per-thread partials, then thread 0 in order
        double ss = 0.0;
        for (int k = 1 + t; k < m; k += kTridThreads) ss += P[col + k] * P[col + k];
        const double xn2 = tri_block_sum(ss, part);
        if (t == 0) {
            const double alpha = P[col];
            if (xn2 == 0.0) {
                sc[0] = 0.0; sc[1] = 1.0; sc[2] = alpha;
            } else {
                const double beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                sc[0] = (beta - alpha) / beta;
                sc[1] = 1.0 / (alpha - beta);
                sc[2] = beta;
            }
        }
        __syncthreads();
        const double ti = sc[0];
        for (int k = t; k < m; k += kTridThreads)
            v[k] = k == 0 ? 1.0 : (ti != 0.0 ? P[col + k] * sc[1] : P[col + k]);
        __syncthreads();
        if (ti != 0.0)
            for (int k = 1 + t; k < m; k += kTridThreads) P[col + k] = v[k];
        if (ti != 0.0) {
            // w = tau A22 v over the symmetric trailing block (rows split into J chunks)
            const int J = max(1, kTridThreads / m);
            {
                const int r = t / J, j = t - r * J;
                double s2 = 0.0;
                if (r < m)
                    for (int c = j; c < m; c += J) {
                        const int rr = max(r, c), cc = min(r, c);
                        s2 += P[tri_at(n, i + 1 + rr, i + 1 + cc)] * v[c];
                    }
                part[t] = s2;
            }
            __syncthreads();
            if (t < m) {
                double s3 = 0.0;
                for (int j = 0; j < J; j++) s3 += part[t * J + j];
                w[t] = ti * s3;
            }
            __syncthreads();
            // alpha2 = -tau/2 (w'v), w += alpha2 v
            double dd = 0.0;
            for (int k = t; k < m; k += kTridThreads) dd += w[k] * v[k];
            const double s4 = tri_block_sum(dd, part);
            if (t == 0) sc[3] = -0.5 * ti * s4;
            __syncthreads();
            for (int k = t; k < m; k += kTridThreads) w[k] += sc[3] * v[k];
            __syncthreads();
            // A22 -= v w' + w v' (lower triangle)
            for (int e = t; e < m * m; e += kTridThreads) {
                const int c = e / m, r = e - c * m;
                if (r >= c) P[tri_at(n, i + 1 + r, i + 1 + c)] -= v[r] * w[c] + w[r] * v[c];
            }
        }
        __syncthreads();
        if (t == 0) {
            E[i] = sc[2];
            D[i] = P[tri_at(n, i, i)];
            tau[i] = ti;
            P[col] = sc[2];
        }
        __syncthreads();
    }
    if (t == 0) {
        D[n - 1] = P[tri_at(n, n - 1, n - 1)];
        tau[n - 1] = 0.0;
        E[n - 1] = 0.0;
    }
    __syncthreads();
    for (int e = t; e < n * n; e += kTridThreads) {
        const int c = e / n, r = e - c * n;
        if (r >= c) A[(long)c * lda + r] = P[tri_at(n, r, c)];
    }
}

// The same reduction (dsytd2, UPLO = 'L', the same D, E, tau and vectors)
// for matrices too large for one workgroup's LDS, in one cooperative launch
// of G workgroups.  Row r of the symmetric matrix lives in the LDS of
// workgroup r mod G (both triangles, so the product A22 v is row-local).
// Step i needs one grid barrier: before it every workgroup publishes its rows
// of p = A22 v with its share of p'v, and the owner of row i+1 publishes that
// row as it stands; after it every workgroup forms w = tau p - (tau^2 / 2)
// (p'v) v and the next column x = row(i+1) - v(0) w - w(0) v itself, in the
// same order (so every workgroup holds the same bits), computes the next
// Householder vector from x and, in one pass over its own rows, applies the
// rank-2 update and forms the next step's p.  Published values are written
// and read with agent-scope atomics (coherent across the XCDs' L2s).  The
// barrier: a release fence per wave, then each workgroup stores its step
// number in its own mark (no read-modify-write on a shared counter) and one
// wave polls all marks, then one acquire fence.  A poll that exceeds its time
// limit sets an abort flag that every other workgroup sees in its own polls,
// so the grid always drains.
constexpr int kCoopMaxN = 1536;
constexpr int kCoopMaxG = 128;

struct CoopSync {
    unsigned int abort;               // set by a poll that timed out
    unsigned int pad[63];
    unsigned int step[kCoopMaxG];     // last step each workgroup has published
};

__device__ __forceinline__ void coop_store(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double coop_load(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int coop_load_u(unsigned int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier: this workgroup marks step `target` published, then waits
// until every workgroup has; false if the grid aborted.  The first wave polls
// the G marks, one or two per lane.
__device__ __forceinline__ bool coop_barrier(CoopSync *sy, unsigned int target, long tick_limit,
                                             int *flag)
{
    const int G = gridDim.x, ln = threadIdx.x & 63;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // each wave's stores performed
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&sy->step[blockIdx.x], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 64) {
        const long t0 = wall_clock64();
        int ok = 1;
        for (;;) {
            const bool r0 = ln >= G || coop_load_u(&sy->step[ln]) >= target;
            const bool r1 = ln + 64 >= G || coop_load_u(&sy->step[ln + 64]) >= target;
            if (__all(r0 && r1)) break;
            if (coop_load_u(&sy->abort)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > tick_limit) {
                if (ln == 0)
                    __hip_atomic_store(&sy->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (ln == 0) *flag = ok;
    }
    __syncthreads();
    return *flag != 0;
}

// Sum over the workgroup in a fixed order, valid in every thread.
template <int NT>
__device__ __forceinline__ double coop_block_sum(double v, double *red)
{
    double a[1] = {v};
    wave_sum_f64<1>(a);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a[0];
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NT / 64; k++) s += red[k];
    return s;
}

// Householder vector of x(0:mm) (dlarfg): vv(0) = 1, vv(1:) = x(1:) /
// (alpha - beta); tau = 0 (and vv(1:) = x(1:)) when x(1:) = 0.
template <int NT>
__device__ __forceinline__ void coop_householder(const double *x, int mm, double *vv,
                                                 double *red, double &ti, double &beta)
{
    double ss = 0.0;
    for (int j = 1 + (int)threadIdx.x; j < mm; j += NT) ss += x[j] * x[j];
    const double xn2 = coop_block_sum<NT>(ss, red);
    const double alpha = x[0];
    double scal = 1.0;
    if (xn2 == 0.0) {
        ti = 0.0;
        beta = alpha;
    } else {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        ti = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    for (int j = threadIdx.x; j < mm; j += NT) vv[j] = j == 0 ? 1.0 : (ti != 0.0 ? x[j] * scal : x[j]);
    __syncthreads();
}

// pub: two buffers of (p[n], row[n], dots[kCoopMaxG]) doubles, by step parity.
template <int NT>
__global__ __launch_bounds__(NT) void sytrd_lower_coop(double *__restrict__ A, int lda, int n,
                                                       double *__restrict__ D,
                                                       double *__restrict__ E,
                                                       double *__restrict__ tau,
                                                       double *__restrict__ pub,
                                                       CoopSync *__restrict__ sy, long tick_limit)
{
    constexpr int NW = NT / 64;
    extern __shared__ double sm[];
    const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
    const int wv = t >> 6, ln = t & 63;
    const int R = (n - 1 - g) / G + 1;      // rows g, g + G, ... below n
    const int ldr = n + 1;
    double *Ar = sm;                        // R rows of n (+1 pad)
    double *x = Ar + (size_t)R * ldr;       // next column below the diagonal (n)
    double *v = x + n;                      // Householder vector of this step (n)
    double *vn = v + n;                     // ... of the next step (n)
    double *w = vn + n;                     // w (n)
    double *pown = w + n;                   // p of the own rows, by local index (R)
    double *red = pown + R;                 // block-sum slots (NW)
    double *wdot = red + NW;                // per-wave shares of p'v (NW)
    __shared__ int flag;
    const size_t pubsz = 2 * (size_t)n + kCoopMaxG;
    for (int k = 0; k < R; k++) {
        const int r = g + k * G;
        for (int c = t; c < n; c += NT)
            Ar[(size_t)k * ldr + c] = r >= c ? A[(long)c * lda + r] : A[(long)r * lda + c];
    }
    for (int j = t; j < n - 1; j += NT) x[j] = A[1 + j];
    double dcur = A[0];                     // A(i, i) of the current step
    double ti, beta;
    __syncthreads();
    coop_householder<NT>(x, n - 1, v, red, ti, beta);
    // p = A22 v for step 0 (rows 1..n-1), and this workgroup's share of p'v
    {
        double dsum = 0.0;
        for (int k = (g == 0 ? 1 : 0) + wv; k < R; k += NW) {
            const double *row = Ar + (size_t)k * ldr + 1;
            double s = 0.0;
            for (int c = ln; c < n - 1; c += 64) s += row[c] * v[c];
            double a[1] = {s};
            wave_sum_f64<1>(a);
            s = a[0];
            if (ln == 0) pown[k] = s;
            dsum += s * v[g + k * G - 1];
        }
        if (ln == 0) wdot[wv] = dsum;
    }
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int q = 0; q < NW; q++) s += wdot[q];
        coop_store(pub + 2 * n + g, s);
    }
    for (int i = 0; i + 1 < n; i++) {
        const int m = n - i - 1;
        double *pb = pub + (size_t)(i & 1) * pubsz;
        // publish p (own rows >= i + 1), p'v, and row i + 1 from its owner
        const int k0 = (i + 1 - g + G - 1) / G;   // first own row >= i + 1
        for (int k = k0 + t; k < R; k += NT) coop_store(pb + (g + k * G - i - 1), pown[k]);
        if (g == (i + 1) % G) {
            const double *row = Ar + (size_t)((i + 1) / G) * ldr + i + 1;
            for (int c = t; c < m; c += NT) coop_store(pb + n + c, row[c]);
        }
        if (!coop_barrier(sy, (unsigned)(i + 1), tick_limit, &flag)) return;
        if (g == 0) {
            for (int j = 1 + t; j < m; j += NT) A[(long)i * lda + i + 1 + j] = v[j];
            if (t == 0) {
                A[(long)i * lda + i] = dcur;
                A[(long)i * lda + i + 1] = beta;
                D[i] = dcur;
                E[i] = beta;
                tau[i] = ti;
            }
        }
        // p, row i + 1 and the shares of p'v, loaded in one batch; every
        // wave sums the shares in the same fixed order
        constexpr int QM = (kCoopMaxN + NT - 1) / NT;
        double pj[QM], rj[QM];
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int j = t + q * NT;
            pj[q] = j < m ? coop_load(pb + j) : 0.0;
            rj[q] = j < m ? coop_load(pb + n + j) : 0.0;
        }
        const double p0 = coop_load(pb), r0 = coop_load(pb + n);
        double pv = ln < G ? coop_load(pb + 2 * n + ln) : 0.0;
        if (ln + 64 < G) pv += coop_load(pb + 2 * n + 64 + ln);
        {
            double a[1] = {pv};
            wave_sum_f64<1>(a);
            pv = a[0];
        }
        const double a2 = -0.5 * ti * (ti * pv);
        // w = tau p + a2 v; the next column x = row(i+1) - v(0) w - w(0) v
        // below its diagonal (v(0) = 1), with the squares of x(1:) summed
        const double w0 = ti * p0 + a2 * 1.0;
        dcur = r0 - (w0 + w0);
        double ss = 0.0;
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int j = t + q * NT;
            if (j < m) {
                const double wj = ti * pj[q] + a2 * v[j];
                w[j] = wj;
                if (j >= 1) {
                    const double xv = rj[q] - (wj + w0 * v[j]);
                    x[j - 1] = xv;
                    if (j >= 2) ss += xv * xv;
                }
            }
        }
        if (m == 1) break;
        {
            double a[1] = {ss};
            wave_sum_f64<1>(a);
            ss = a[0];
        }
        if (ln == 0) red[wv] = ss;
        __syncthreads();
        // Householder vector of x (dlarfg), the same bits in every workgroup
        double xn2 = 0.0;
        for (int q = 0; q < NW; q++) xn2 += red[q];
        const double alpha = x[0];
        double ti1 = 0.0, beta1 = alpha, scal = 1.0;
        if (xn2 != 0.0) {
            beta1 = -copysign(sqrt(alpha * alpha + xn2), alpha);
            ti1 = (beta1 - alpha) / beta1;
            scal = 1.0 / (alpha - beta1);
        }
        for (int j = t; j < m - 1; j += NT) vn[j] = j == 0 ? 1.0 : (ti1 != 0.0 ? x[j] * scal : x[j]);
        __syncthreads();
        // own rows r >= i + 2: rank-2 update of columns i+2..n-1 and the next
        // step's p = A22 vn with this workgroup's share of p'vn; each wave
        // takes four rows at a time (the v, w, vn loads shared, the four
        // lane sums reduced together)
        {
            const int k1 = (i + 2 - g + G - 1) / G;
            double dsum = 0.0;
            for (int kb = k1 + wv; kb < R; kb += 4 * NW) {
                double acc[4] = {0.0, 0.0, 0.0, 0.0}, vj[4], wj[4];
                double *row[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = min(kb + u * NW, R - 1);
                    const int j = g + k * G - i - 1;
                    vj[u] = kb + u * NW < R ? v[j] : 0.0;
                    wj[u] = kb + u * NW < R ? w[j] : 0.0;
                    row[u] = Ar + (size_t)k * ldr + i + 1;
                }
                const int nu = min(4, (R - kb + NW - 1) / NW);
                // two columns per lane and pass, every load issued before the
                // stores (the LDS rows and vectors may alias for the compiler)
                for (int c0 = 1 + ln; c0 < m; c0 += 128) {
                    const int c1 = c0 + 64;
                    const bool h1 = c1 < m;
                    const double wc0 = w[c0], vc0 = v[c0], vnc0 = vn[c0 - 1];
                    const double wc1 = h1 ? w[c1] : 0.0, vc1 = h1 ? v[c1] : 0.0;
                    const double vnc1 = h1 ? vn[c1 - 1] : 0.0;
                    double a0[4], a1[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        a0[u] = u < nu ? row[u][c0] : 0.0;
                        a1[u] = u < nu && h1 ? row[u][c1] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        a0[u] -= vj[u] * wc0 + wj[u] * vc0;
                        a1[u] -= vj[u] * wc1 + wj[u] * vc1;
                        acc[u] += a0[u] * vnc0;
                        acc[u] += a1[u] * vnc1;
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        if (u < nu) {
                            row[u][c0] = a0[u];
                            if (h1) row[u][c1] = a1[u];
                        }
                }
                wave_sum_f64<4>(acc);
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (u < nu) {
                        const int k = kb + u * NW;
                        if (ln == 0) pown[k] = acc[u];
                        dsum += acc[u] * vn[g + k * G - i - 2];
                    }
            }
            if (ln == 0) wdot[wv] = dsum;
        }
        __syncthreads();
        if (t == 0) {
            double sd = 0.0;
            for (int q = 0; q < NW; q++) sd += wdot[q];
            coop_store(pub + (size_t)((i + 1) & 1) * pubsz + 2 * n + g, sd);
        }
        double *tmp = v; v = vn; vn = tmp;
        ti = ti1;
        beta = beta1;
    }
    if (g == 0 && t == 0) {
        A[(long)(n - 1) * lda + n - 1] = dcur;
        D[n - 1] = dcur;
        E[n - 1] = 0.0;
        tau[n - 1] = 0.0;
    }
}

// Cholesky factor of the overlap matrix, Mp = L L' (lower), left-looking in
// panels of 32 columns, one launch per panel: LAPACK dpotrf's factor (the
// reference's dsygvd, src/eigenSolver.c:1318, starts with it), where
// rocSOLVER's dpotrf spends ~1 ms in short launches at 480 states.  Block 0
// of a launch owns the diagonal block, block b >= 1 the 32 rows
// kb + nb + 32 (b - 1) ... below it.  Every block forms the diagonal block
// D = A11 - L10 L10' (the same bits in every block) and, for a row block,
// P = A21 - L20 L10' on FP64 matrix cores: the columns of L10 / L20 are
// staged four chunks of 32 at a time (one chunk per wave, the four partial
// sums added in a fixed order).  One wave factors D in registers (rows in lanes, the pivot column
// broadcast with readlane); the row blocks then solve L21 L11' = P row by
// row.  Block 0 keeps L11 in a staging buffer (chol_put_diag copies the
// diagonal blocks into place after the last panel) and zeroes the strict
// upper triangle above it (the factor is then used as a full matrix).  info gets the 1-based column
// of the first pivot that is not positive, and a launch after a failed one
// returns at once.
constexpr int kCholNB = 32;
constexpr int kCholThreads = 256;
constexpr int kCholLd = kCholNB + 1;
constexpr size_t kCholShm = (4 * 2 + 3) * (size_t)kCholNB * kCholLd * sizeof(double) +
                            kCholNB * sizeof(double);


__global__ __launch_bounds__(kCholThreads) void chol_panel(double *__restrict__ B, int n, int kb,
                                                           int *__restrict__ info,
                                                           double *__restrict__ l11buf)
{
    extern __shared__ double csm[];
    double(*St)[kCholLd] = (double(*)[kCholLd])csm;             // [4 chunks][Ld; Lr][32] rows
    double(*Dm)[kCholLd] = St + 8 * kCholNB;
    double(*Pm)[kCholLd] = Dm + kCholNB;
    double(*L11)[kCholLd] = Pm + kCholNB;
    double *rdiag = (double *)(L11 + kCholNB);
    __shared__ int fail;
    if (*info != 0) return;
    const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
    const int nb = min(kCholNB, n - kb);
    const int b = blockIdx.x;
    const int r0 = kb + nb + kCholNB * (b - 1);
    const int nr = b == 0 ? 0 : min(kCholNB, n - r0);
    // D and P of this wave's chunks on v_mfma_f64_16x16x4f64: 16 x 16 tiles
    // (a, c) of D = Ld Ld' and P = Lr Ld'; operand lane (li, lk) holds row li
    // of a tile at k = 4 s + lk, which for D's B operand is again Ld (row j).
    // D(i, j) of a tile ends in lane j + 16 (i % 4), register i / 4.
    const int li = ln & 15, lk = ln >> 4;
    typedef double f64x4_t __attribute__((ext_vector_type(4)));
    f64x4_t ad[2][2], ap[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            ad[a][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
            ap[a][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
        }
    for (int s0 = 0; s0 < kb; s0 += 4 * kCholNB) {
        const int nch = min(4, (kb - s0) / kCholNB);
        for (int e = t; e < nch * kCholNB * kCholNB; e += kCholThreads) {
            const int ii = e & 31, kk = (e >> 5) & 31, ch = e >> 10;
            const double *col = B + (long)(s0 + ch * kCholNB + kk) * n;
            St[(2 * ch) * kCholNB + ii][kk] = ii < nb ? col[kb + ii] : 0.0;
            St[(2 * ch + 1) * kCholNB + ii][kk] = ii < nr ? col[r0 + ii] : 0.0;
        }
        __syncthreads();
        if (wv < nch) {
            const double(*Ld)[kCholLd] = St + (2 * wv) * kCholNB;
            const double(*Lr)[kCholLd] = St + (2 * wv + 1) * kCholNB;
#pragma unroll
            for (int s = 0; s < kCholNB / 4; s++) {
                const int k = 4 * s + lk;
                const double d0 = Ld[li][k], d1 = Ld[16 + li][k];
                const double p0 = Lr[li][k], p1 = Lr[16 + li][k];
                ad[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(d0, d0, ad[0][0], 0, 0, 0);
                ad[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(d0, d1, ad[0][1], 0, 0, 0);
                ad[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(d1, d0, ad[1][0], 0, 0, 0);
                ad[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(d1, d1, ad[1][1], 0, 0, 0);
                ap[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(p0, d0, ap[0][0], 0, 0, 0);
                ap[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(p0, d1, ap[0][1], 0, 0, 0);
                ap[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(p1, d0, ap[1][0], 0, 0, 0);
                ap[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(p1, d1, ap[1][1], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // partial sums of the four waves, added in wave order
    if (kb > 0) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 16 * a + lk + 4 * r, j = 16 * c + li;
                    St[(2 * wv) * kCholNB + i][j] = ad[a][c][r];
                    St[(2 * wv + 1) * kCholNB + i][j] = ap[a][c][r];
                }
        __syncthreads();
    }
    for (int e = t; e < kCholNB * kCholNB; e += kCholThreads) {
        const int i = e >> 5, j = e & 31;
        double sd = 0.0, sp = 0.0;
        if (kb > 0)
            for (int w = 0; w < 4; w++) {
                sd += St[(2 * w) * kCholNB + i][j];
                sp += St[(2 * w + 1) * kCholNB + i][j];
            }
        double d;
        if (i < nb && j < nb)
            d = j <= i ? B[(long)(kb + j) * n + kb + i] - sd : 0.0;
        else
            d = i == j ? 1.0 : 0.0;
        Dm[i][j] = d;
        Pm[i][j] = i < nr && j < nb ? B[(long)(kb + j) * n + r0 + i] - sp : 0.0;
    }
    __syncthreads();
    // L11 by one wave: lane l < 32 holds row l of D; step j scales column j
    // by its pivot and subtracts L(l, j) L(k, j) from row l (k > j)
    if (wv == 0) {
        const int l = ln & 31;
        double row[kCholNB];
#pragma unroll
        for (int k = 0; k < kCholNB; k++) row[k] = Dm[l][k];
        int bad = 0;
#pragma unroll
        for (int j = 0; j < kCholNB; j++) {
            const double djj = readlane_f64(row[j], j);
            if (!(djj > 0.0)) {
                bad = j + 1;
                break;
            }
            const double piv = sqrt(djj), rp = 1.0 / piv;
            row[j] = l == j ? piv : row[j] * rp;
#pragma unroll
            for (int k = j + 1; k < kCholNB; k++) row[k] -= row[j] * readlane_f64(row[j], k);
        }
        if (ln < kCholNB) {
#pragma unroll
            for (int k = 0; k < kCholNB; k++) L11[l][k] = k <= l ? row[k] : 0.0;
            rdiag[l] = 1.0 / row[l];
        }
        if (ln == 0) fail = bad;
    }
    __syncthreads();
    if (fail) {
        if (b == 0 && t == 0) *info = kb + fail;
        return;
    }
    if (b == 0) {
        // L11 goes to l11buf, not into B: the row blocks of this launch may
        // still be reading A11 there (chol_put_diag copies it after the
        // last panel; later panels never read these rows)
        double *lb = l11buf + (size_t)(kb / kCholNB) * kCholNB * kCholNB;
        for (int e = t; e < kCholNB * kCholNB; e += kCholThreads) {
            const int ii = e & 31, jj = e >> 5;
            lb[e] = L11[ii][jj];
        }
        for (int jj = 0; jj < nb; jj++)
            for (int r = t; r < kb; r += kCholThreads) B[(long)(kb + jj) * n + r] = 0.0;
        return;
    }
    // L21 = P L11^-T: one thread per row, forward substitution
    if (t < nr) {
        double x[kCholNB];
#pragma unroll
        for (int j = 0; j < kCholNB; j++) {
            double s = Pm[t][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= x[k] * L11[j][k];
            x[j] = s * rdiag[j];
        }
#pragma unroll
        for (int j = 0; j < kCholNB; j++) Pm[t][j] = x[j];
    }
    __syncthreads();
    for (int e = t; e < kCholNB * kCholNB; e += kCholThreads) {
        const int ii = e & 31, jj = e >> 5;
        if (ii < nr && jj < nb) B[(long)(kb + jj) * n + r0 + ii] = Pm[ii][jj];
    }
}

// The diagonal blocks L11 of every panel from l11buf into B (zeros above the
// diagonal).
__global__ __launch_bounds__(256) void chol_put_diag(double *__restrict__ B, int n,
                                                     const double *__restrict__ l11buf)
{
    const int kb = blockIdx.x * kCholNB, nb = min(kCholNB, n - kb);
    const double *lb = l11buf + (size_t)blockIdx.x * kCholNB * kCholNB;
    for (int e = threadIdx.x; e < kCholNB * kCholNB; e += 256) {
        const int ii = e & 31, jj = e >> 5;
        if (ii < nb && jj < nb) B[(long)(kb + jj) * n + kb + ii] = lb[e];
    }
}

// l11buf: ceil(n / 32) * 32 * 32 doubles.
static int cholesky_device(sgpu_context *ctx, int n, double *dB, rocblas_int *dinfo,
                           double *l11buf)
{
    SGPU_HIP(hipFuncSetAttribute((const void *)chol_panel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCholShm));
    SGPU_HIP(hipMemsetAsync(dinfo, 0, sizeof(rocblas_int), ctx->stream));
    for (int kb = 0; kb < n; kb += kCholNB) {
        const int nb = std::min(kCholNB, n - kb);
        const int nblk = 1 + (n - kb - nb + kCholNB - 1) / kCholNB;
        chol_panel<<<nblk, kCholThreads, kCholShm, ctx->stream>>>(dB, n, kb, (int *)dinfo, l11buf);
        SGPU_HIP(hipGetLastError());
    }
    chol_put_diag<<<(unsigned)((n + kCholNB - 1) / kCholNB), 256, 0, ctx->stream>>>(dB, n, l11buf);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

// Q1 Z for the tridiagonal reduction's reflectors (LAPACK dormtr with
// SIDE = 'L', UPLO = 'L', TRANS = 'N', as rocSOLVER's, which takes 0.9 ms at
// n = 480 in short launches): Z <- H(0) H(1) ... H(n-2) Z with
// H(i) = I - tau(i) v v', v(i+1) = 1, v(i+2:n) = A(i+2:n, i).  Columns of Z
// are independent: each of the four waves of a block owns CW of them, its
// rows spread over the lanes (row lane + 64 q, q < NR) and held in registers,
// and applies the reflectors from the last to the first (v'z by the DPP
// wave sum).  The reflectors are staged in LDS KC at a time, double
// buffered: the next chunk's loads are in flight while the waves apply the
// current one, one barrier per chunk.
template <int NR, int CW, int KC>
__global__ __launch_bounds__(256) void ormtr_lower_blk(const double *__restrict__ A, int lda,
                                                       const double *__restrict__ tau,
                                                       double *__restrict__ Z, int ldz, int n)
{
    extern __shared__ double osm[];
    double *vs = osm;                       // [2][KC][n] reflector vectors
    double *ts = osm + 2 * (size_t)KC * n;  // [2][KC] tau
    const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
    const int c0 = (blockIdx.x * 4 + wv) * CW;
    constexpr int SPT = (KC * 64 * NR + 255) / 256;   // staged values per thread (n <= 64 NR)
    double z[NR][CW];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const int r = ln + 64 * q;
#pragma unroll
        for (int j = 0; j < CW; j++) z[q][j] = r < n && c0 + j < n ? Z[(long)(c0 + j) * ldz + r] : 0.0;
    }
    // chunk c holds reflectors n-2 - c KC - g, g < KC (zero past the first)
    double st[SPT], stt = 0.0;
    auto load_chunk = [&](int c) {
#pragma unroll
        for (int k = 0; k < SPT; k++) {
            const int e = t + 256 * k;
            const int g = e / n, r = e - g * n;
            const int i = n - 2 - c * KC - g;
            double v = 0.0;
            if (g < KC && i >= 0) v = r == i + 1 ? 1.0 : (r > i + 1 ? A[(long)i * lda + r] : 0.0);
            st[k] = v;
        }
        if (t < KC) {
            const int i = n - 2 - c * KC - t;
            stt = i >= 0 ? tau[i] : 0.0;
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int k = 0; k < SPT; k++) {
            const int e = t + 256 * k;
            if (e < KC * n) vs[(size_t)buf * KC * n + e] = st[k];
        }
        if (t < KC) ts[buf * KC + t] = stt;
    };
    const int nchunk = (n - 1 + KC - 1) / KC;
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        if (c + 1 < nchunk) load_chunk(c + 1);
        const double *vb = vs + (size_t)buf * KC * n;
        for (int g = 0; g < KC; g++) {
            if (n - 2 - c * KC - g < 0) break;
            const double ti = ts[buf * KC + g];
            double v[NR];
#pragma unroll
            for (int q = 0; q < NR; q++) {
                const int r = ln + 64 * q;
                v[q] = r < n ? vb[(size_t)g * n + r] : 0.0;
            }
            double d[CW];
#pragma unroll
            for (int j = 0; j < CW; j++) {
                d[j] = 0.0;
#pragma unroll
                for (int q = 0; q < NR; q++) d[j] += v[q] * z[q][j];
            }
            wave_sum_f64<CW>(d);
#pragma unroll
            for (int j = 0; j < CW; j++) {
                const double w = ti * d[j];
#pragma unroll
                for (int q = 0; q < NR; q++) z[q][j] -= v[q] * w;
            }
        }
        if (c + 1 < nchunk) store_chunk(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const int r = ln + 64 * q;
#pragma unroll
        for (int j = 0; j < CW; j++)
            if (r < n && c0 + j < n) Z[(long)(c0 + j) * ldz + r] = z[q][j];
    }
}

template <int NR, int CW, int KC>
static int launch_ormtr_blk(sgpu_context *ctx, int n, const double *dA, const double *dtau,
                            double *dZ)
{
    const size_t shm = (2 * (size_t)KC * n + 2 * KC) * sizeof(double);
    SGPU_HIP(hipFuncSetAttribute((const void *)ormtr_lower_blk<NR, CW, KC>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    const unsigned nblk = (unsigned)((n + 4 * CW - 1) / (4 * CW));
    ormtr_lower_blk<NR, CW, KC><<<nblk, 256, shm, ctx->stream>>>(dA, n, dtau, dZ, n, n);
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

static int ormtr_lower_device(sgpu_context *ctx, int n, const double *dA, const double *dtau,
                              double *dZ)
{
    // rows per lane NR, columns per wave CW, reflectors per chunk KC: z and
    // the staged chunk fit in registers, two chunks in LDS
    const int nr = (n + 63) / 64;
    if (nr <= 4) return launch_ormtr_blk<4, 4, 16>(ctx, n, dA, dtau, dZ);
    if (nr <= 8) return launch_ormtr_blk<8, 4, 16>(ctx, n, dA, dtau, dZ);
    if (nr <= 12) return launch_ormtr_blk<12, 4, 8>(ctx, n, dA, dtau, dZ);
    if (nr <= 16) return launch_ormtr_blk<16, 4, 8>(ctx, n, dA, dtau, dZ);
    if (nr <= 24) return launch_ormtr_blk<24, 2, 4>(ctx, n, dA, dtau, dZ);
    sgpu_set_error("eigensolve: back-transformation limited to %d states", 24 * 64);
    return SGPU_ERR_ARG;
}

// Workgroup count and dynamic LDS of sytrd_lower_coop for n (0 groups when
// it does not apply): at least the requested count, more where the rows
// would not fit in 150 KB of LDS per workgroup, at most one group per row.
static int coop_plan(const sgpu_context *ctx, int n, size_t *shm)
{
    if (ctx->coop_tridiag <= 0 || n < 2 || n > kCoopMaxN) return 0;
    const size_t lds_max = 150 * 1024;
    const size_t fixed = (4 * (size_t)n + 2 * (ctx->coop_threads / 64)) * sizeof(double);
    const size_t row = (size_t)(n + 2) * sizeof(double);   // the row and its p
    if (fixed + row > lds_max) return 0;
    const int rmax = (int)((lds_max - fixed) / row);
    int G = std::max(ctx->coop_tridiag, (n + rmax - 1) / rmax);
    G = std::min(G, n);
    if (G > kCoopMaxG) return 0;
    const int R = (n + G - 1) / G;
    *shm = fixed + (size_t)R * row;
    return G;
}

template <int NT>
static int launch_sytrd_coop(sgpu_context *ctx, int n, int G, size_t shm, double *dA, double *dD,
                             double *dE, double *dtau, double *pub, CoopSync *sy)
{
    int wc_khz = 0;
    SGPU_HIP(hipDeviceGetAttribute(&wc_khz, hipDeviceAttributeWallClockRate, ctx->device));
    // a barrier that waits 0.5 s is taken as a failed launch
    long tick_limit = (long)std::max(wc_khz, 1000) * 500;
    SGPU_HIP(hipFuncSetAttribute((const void *)sytrd_lower_coop<NT>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    SGPU_HIP(hipMemsetAsync(sy, 0, sizeof(CoopSync), ctx->stream));
    int lda = n;
    void *args[] = {&dA, &lda, &n, &dD, &dE, &dtau, &pub, &sy, &tick_limit};
    SGPU_HIP(hipLaunchCooperativeKernel((const void *)sytrd_lower_coop<NT>, dim3(G), dim3(NT), args,
                                        (unsigned)shm, ctx->stream));
    return SGPU_OK;
}

int eigensolve_device(sgpu_context *ctx, int n, double *dA, double *dB, double *lambda)
{
    void *buf = nullptr;
    // D, E, tau, the tridiagonal eigenvectors, the cooperative reduction's
    // exchange buffers, its barrier and info in the shared scratch (free again
    // once lambda is on the host)
    const size_t npub = 2 * (2 * (size_t)n + kCoopMaxG);
    const size_t nl11 = (size_t)((n + kCholNB - 1) / kCholNB) * kCholNB * kCholNB;
    const bool tdc = ctx->own_stedc && n <= kTdcMaxN;
    const size_t ntdc = tdc ? sgpu_tdc_work_doubles(n) : 0;
    SGPU_CHECK(sgpu_scratch(ctx, (3 * (size_t)n + (size_t)n * n + npub + nl11 + ntdc) * sizeof(double) +
                                     sizeof(CoopSync) + 320,
                            &buf));
    double *dD = (double *)buf, *dE = dD + n;
    double *pub = dE + 2 * n + (size_t)n * n;
    double *l11buf = pub + npub;
    CoopSync *sy = (CoopSync *)(l11buf + nl11);
    rocblas_int *dinfo = (rocblas_int *)(sy + 1);
    double *tdcw = (double *)(((uintptr_t)(dinfo + 1) + 255) & ~(uintptr_t)255);
    sgpu_tmark t, tp;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_CHECK(sgpu_tbegin(ctx, &tp));
    rocblas_int info = 0;
    SGPU_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host));
    sgpu_tmark tq;
    SGPU_CHECK(sgpu_tbegin(ctx, &tq));
    if (ctx->own_chol)
        SGPU_CHECK(cholesky_device(ctx, n, dB, dinfo, l11buf));
    else
        SGPU_BLAS(rocsolver_dpotrf(ctx->blas, rocblas_fill_upper, n, dB, n, dinfo));
    SGPU_CHECK(sgpu_tend(ctx, &tq, "eig_potrf"));
    SGPU_HIP(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    if (info != 0) {
        sgpu_set_error("eigensolve: overlap matrix not positive definite (Cholesky info = %d)",
                       (int)info);
        return SGPU_ERR_SOLVER;
    }
    const double one = 1.0, zero = 0.0;
    double *dW = dE + 2 * n;   // n x n work (the tridiagonal eigenvectors later)
    if (ctx->own_chol) {
        // dB <- L^-1 (lower, zero above), C = L^-1 Hp L^-T
        SGPU_BLAS(rocsolver_dtrtri(ctx->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, n, dB,
                                   n, dinfo));
        SGPU_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n,
                                &one, dB, n, dA, n, &zero, dW, n));
        SGPU_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose, n,
                                n, n, &one, dW, n, dB, n, &zero, dA, n));
    } else {
        SGPU_BLAS(rocblas_dtrsm(ctx->blas, rocblas_side_left, rocblas_fill_upper,
                                rocblas_operation_transpose, rocblas_diagonal_non_unit, n, n, &one,
                                dB, n, dA, n));
        SGPU_BLAS(rocblas_dtrsm(ctx->blas, rocblas_side_right, rocblas_fill_upper,
                                rocblas_operation_none, rocblas_diagonal_non_unit, n, n, &one, dB,
                                n, dA, n));
    }
    SGPU_CHECK(sgpu_tend(ctx, &tp, "eig_cholesky"));
    size_t coop_shm = 0;
    const bool lds = n <= kTridMaxN && ctx->lds_tridiag;
    const int G = lds ? 0 : coop_plan(ctx, n, &coop_shm);
    if (lds || G > 0) {
        // tridiagonal reduction (one workgroup, or one cooperative launch),
        // then the tridiagonal eigenvectors (dstedc) and the
        // back-transformation Q Z (dormtr): the three steps of dsyevd with its
        // dsytrd replaced
        double *dtau = dE + n, *dC = dtau + n;
        SGPU_CHECK(sgpu_tbegin(ctx, &tp));
        if (lds) {
            const size_t shm = ((size_t)n * (n + 1) / 2 + 2 * (size_t)n + kTridThreads) * sizeof(double);
            SGPU_HIP(hipFuncSetAttribute((const void *)sytd2_lower_lds,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            sytd2_lower_lds<<<1, kTridThreads, shm, ctx->stream>>>(dA, n, n, dD, dE, dtau);
            SGPU_HIP(hipGetLastError());
        } else {
            if (ctx->coop_threads == 512)
                SGPU_CHECK(launch_sytrd_coop<512>(ctx, n, G, coop_shm, dA, dD, dE, dtau, pub, sy));
            else
                SGPU_CHECK(launch_sytrd_coop<256>(ctx, n, G, coop_shm, dA, dD, dE, dtau, pub, sy));
            CoopSync hs{};
            SGPU_HIP(hipMemcpyAsync(&hs, sy, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
            SGPU_HIP(hipStreamSynchronize(ctx->stream));
            if (hs.abort) {
                sgpu_set_error("eigensolve: tridiagonal reduction's grid barrier timed out "
                               "(%d workgroups)", G);
                return SGPU_ERR_SOLVER;
            }
        }
        SGPU_CHECK(sgpu_tend(ctx, &tp, "eig_tridiag"));
        SGPU_CHECK(sgpu_tbegin(ctx, &tp));
        if (tdc)
            SGPU_CHECK(sgpu_tdc(ctx, n, dD, dE, dC, tdcw));
        else
            SGPU_BLAS(rocsolver_dstedc(ctx->blas, rocblas_evect_tridiagonal, n, dD, dE, dC, n,
                                       dinfo));
        SGPU_CHECK(sgpu_tend(ctx, &tp, "eig_stedc"));
        SGPU_CHECK(sgpu_tbegin(ctx, &tp));
        SGPU_CHECK(sgpu_tbegin(ctx, &tq));
        if (ctx->own_ormtr && n <= 24 * 64)
            SGPU_CHECK(ormtr_lower_device(ctx, n, dA, dtau, dC));
        else
            SGPU_BLAS(rocsolver_dormtr(ctx->blas, rocblas_side_left, rocblas_fill_lower,
                                       rocblas_operation_none, n, n, dA, n, dtau, dC, n));
        SGPU_CHECK(sgpu_tend(ctx, &tq, "eig_ormtr"));
        if (ctx->own_chol) {
            SGPU_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                                    n, n, n, &one, dB, n, dC, n, &zero, dA, n));
        } else {
            SGPU_BLAS(rocblas_dtrsm(ctx->blas, rocblas_side_left, rocblas_fill_upper,
                                    rocblas_operation_none, rocblas_diagonal_non_unit, n, n, &one,
                                    dB, n, dC, n));
            SGPU_HIP(hipMemcpyAsync(dA, dC, (size_t)n * n * sizeof(double),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        }
        SGPU_CHECK(sgpu_tend(ctx, &tp, "eig_back"));
    } else {
        SGPU_BLAS(rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, n, dA, n,
                                   dD, dE, dinfo));
        if (ctx->own_chol) {
            SGPU_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                                    n, n, n, &one, dB, n, dA, n, &zero, dW, n));
            SGPU_HIP(hipMemcpyAsync(dA, dW, (size_t)n * n * sizeof(double),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            SGPU_BLAS(rocblas_dtrsm(ctx->blas, rocblas_side_left, rocblas_fill_upper,
                                    rocblas_operation_none, rocblas_diagonal_non_unit, n, n, &one,
                                    dB, n, dA, n));
        }
    }
    SGPU_CHECK(sgpu_tend(ctx, &t, "eigensolve"));
    SGPU_HIP(hipMemcpyAsync(lambda, dD, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    if (info != 0) {
        sgpu_set_error("eigensolve: tridiagonal eigensolve did not converge (info = %d)", (int)info);
        return SGPU_ERR_SOLVER;
    }
    for (int i = 0; i < n; i++)
        if (!std::isfinite(lambda[i])) {
            sgpu_set_error("eigensolve: non-finite eigenvalue %d (non-finite Hp or Mp?)", i);
            return SGPU_ERR_SOLVER;
        }
    return SGPU_OK;
}

// The same from host matrices: Hp (host) is overwritten by Q.
int eigensolve(sgpu_context *ctx, int n, double *Hp, double *Mp, double *lambda)
{
    const long nn = (long)n * n;
    double *dmat = nullptr;
    SGPU_CHECK(sgpu_reduce_buf(ctx, (size_t)(2 * nn), &dmat));
    SGPU_HIP(hipMemcpyAsync(dmat, Hp, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SGPU_HIP(hipMemcpyAsync(dmat + nn, Mp, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SGPU_CHECK(eigensolve_device(ctx, n, dmat, dmat + nn, lambda));
    SGPU_HIP(hipMemcpyAsync(Hp, dmat, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

// ---- Lanczos vector kernels ----
__global__ void lz_orth(double *__restrict__ vjp1, const double *__restrict__ vj,
                        double *__restrict__ vjm1, double a, double b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vjp1[i] -= a * vj[i] + b * vjm1[i];
    vjm1[i] = vj[i];
}

__global__ void vscale(double *__restrict__ y, const double *__restrict__ x, double s, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = s * x[i];
}

__global__ void vaxpy(double *__restrict__ y, const double *__restrict__ x, double s, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += s * x[i];
}

int sturm_count(int n, const double *d, const double *e, double x)
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

// Extreme eigenvalues of the symmetric tridiagonal matrix (diagonal d,
// off-diagonal e) by Sturm-sequence bisection; the reference takes the ends
// of LAPACKE_dsterf's spectrum (src/eigenSolver.c:2072-2081).
void tridiag_extremes(int n, const double *d, const double *e, double *lmin, double *lmax)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double span = hi - lo + 1e-300;
    for (int which = 0; which < 2; which++) {
        double l = lo - 1e-12 * span, h = hi + 1e-12 * span;
        const int target = which == 0 ? 1 : n;
        for (int it = 0; it < 200 && h - l > 4e-16 * (std::fabs(l) + std::fabs(h)) + 1e-300; it++) {
            double mid = 0.5 * (l + h);
            if (sturm_count(n, d, e, mid) >= target) h = mid; else l = mid;
        }
        (which == 0 ? *lmin : *lmax) = 0.5 * (l + h);
    }
}

}  // namespace

// ======================= device-resident entry points ======================

extern "C" int sgpu_lap_plus_diag_dev(sgpu_context *ctx, double a, double b, double c,
                                      const sgpu_block *x, sgpu_block *y)
{
    if (!ctx) return SGPU_ERR_ARG;
    SGPU_CHECK(check_pair(ctx, x, y, "sgpu_lap_plus_diag_dev"));
    SGPU_HIP(hipSetDevice(ctx->device));
    const bool diag = std::fabs(b) >= 1e-14;
    if (diag && !ctx->has_veff) {
        sgpu_set_error("sgpu_lap_plus_diag_dev: diagonal requested before sgpu_set_veff");
        return SGPU_ERR_STATE;
    }
    SGPU_CHECK(sgpu_launch_stencil(ctx, x->dtype, x->ncol, a, b, c,
                                   diag ? veff_for(ctx, x->dtype) : nullptr, x->ptr, x->ld, y->ptr,
                                   y->ld, nullptr, 0, 1.0, 0.0, true));
    return SGPU_OK;
}

extern "C" int sgpu_hamiltonian_dev(sgpu_context *ctx, double c, const sgpu_block *x,
                                    sgpu_block *Hx)
{
    if (!ctx) return SGPU_ERR_ARG;
    SGPU_CHECK(check_pair(ctx, x, Hx, "sgpu_hamiltonian_dev"));
    SGPU_HIP(hipSetDevice(ctx->device));
    return apply_h(ctx, x->dtype, x->ncol, c, x->ptr, x->ld, Hx->ptr, Hx->ld, nullptr, 0, 1.0, 0.0);
}

extern "C" int sgpu_chebyshev_filter_dev(sgpu_context *ctx, sgpu_block *X, sgpu_block *Y, int m,
                                         double a, double b, double a0)
{
    if (!ctx) return SGPU_ERR_ARG;
    SGPU_CHECK(check_pair(ctx, X, Y, "sgpu_chebyshev_filter_dev"));
    if (m < 1 || !(b > a) || a0 == 0.5 * (a + b)) {
        sgpu_set_error("sgpu_chebyshev_filter_dev: bad degree or bounds");
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    void *res = nullptr, *prev = nullptr;
    SGPU_CHECK(filter_raw(ctx, X->dtype, X->ncol, X->ptr, Y->ptr, X->ld, m, a, b, a0, &res, &prev));
    Y->ptr = res;   // storage exchange instead of a copy when m is even
    X->ptr = prev;
    return SGPU_OK;
}

extern "C" int sgpu_project_dev(sgpu_context *ctx, const sgpu_block *Y, sgpu_block *HY,
                                double *Hp, double *Mp)
{
    if (!ctx || !Hp || !Mp) return SGPU_ERR_ARG;
    SGPU_CHECK(check_pair(ctx, Y, HY, "sgpu_project_dev"));
    SGPU_HIP(hipSetDevice(ctx->device));
    return project_raw(ctx, Y->dtype, Y->ncol, Y->ptr, HY->ptr, Y->ld, Hp, Mp);
}

extern "C" int sgpu_rotate_dev(sgpu_context *ctx, const sgpu_block *Y, const double *Q,
                               sgpu_block *out)
{
    if (!ctx || !Q) return SGPU_ERR_ARG;
    SGPU_CHECK(check_pair(ctx, Y, out, "sgpu_rotate_dev"));
    if (Y->ptr == out->ptr) {
        sgpu_set_error("sgpu_rotate_dev: output must not alias input");
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    return rotate_raw(ctx, Y->dtype, Y->ncol, Y->ptr, Y->ld, Q, out->ptr, out->ld);
}

extern "C" int sgpu_chefsi_dev(sgpu_context *ctx, sgpu_block *X, sgpu_block *Y, int m, double a,
                               double b, double a0, double *lambda)
{
    if (!ctx || !lambda) return SGPU_ERR_ARG;
    SGPU_CHECK(sgpu_chebyshev_filter_dev(ctx, X, Y, m, a, b, a0));
    // projection, eigensolve and rotation on device-resident matrices: Mp
    // and Hp stay in the reduction buffer, Hp becomes Q in place
    const int ns = Y->ncol;
    const long nn = (long)ns * ns;
    double *dmat = nullptr;
    SGPU_CHECK(project_device(ctx, Y->dtype, ns, Y->ptr, X->ptr, Y->ld, &dmat));
    SGPU_CHECK(eigensolve_device(ctx, ns, dmat + nn, dmat, lambda));
    SGPU_CHECK(rotate_device(ctx, Y->dtype, ns, Y->ptr, Y->ld, dmat + nn, X->ptr, X->ld));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

extern "C" int sgpu_lanczos(sgpu_context *ctx, const double *x0, double tol_min, double tol_max,
                            int maxit, double *eigmin, double *eigmax, int *iters)
{
    if (!ctx || !x0 || !eigmin || !eigmax || maxit < 1) return SGPU_ERR_ARG;
    SGPU_HIP(hipSetDevice(ctx->device));
    const long n = ctx->nd;
    void *buf = nullptr;
    // separate allocation: apply_h may use the shared scratch
    double *V = nullptr;
    SGPU_HIP(hipMalloc((void **)&V, 3 * ctx->ld * sizeof(double)));
    (void)buf;
    double *Vj = V, *Vjm1 = V + ctx->ld, *Vjp1 = V + 2 * ctx->ld;
    const unsigned nb = (unsigned)((n + 255) / 256);
    std::vector<double> a(maxit + 2), bb(maxit + 2);
    int status = SGPU_OK, j = 0;
    auto run = [&]() -> int {
        SGPU_HIP(hipMemcpyAsync(Vjm1, x0, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        double nrm2;
        const double *pa[1] = {Vjm1}, *pb[1] = {Vjm1};
        SGPU_CHECK(sgpu_dot_many(ctx, (int)n, 1, pa, pb, &nrm2));
        vscale<<<nb, 256, 0, ctx->stream>>>(Vjm1, Vjm1, 1.0 / std::sqrt(nrm2), n);
        SGPU_CHECK(apply_h(ctx, SGPU_F64, 1, 0.0, Vjm1, ctx->ld, Vj, ctx->ld, nullptr, 0, 1.0, 0.0));
        pa[0] = Vjm1; pb[0] = Vj;
        SGPU_CHECK(sgpu_dot_many(ctx, (int)n, 1, pa, pb, &a[0]));
        vaxpy<<<nb, 256, 0, ctx->stream>>>(Vj, Vjm1, -a[0], n);
        pa[0] = Vj; pb[0] = Vj;
        SGPU_CHECK(sgpu_dot_many(ctx, (int)n, 1, pa, pb, &nrm2));
        bb[0] = std::sqrt(nrm2);
        vscale<<<nb, 256, 0, ctx->stream>>>(Vj, Vj, bb[0] == 0.0 ? 1.0 : 1.0 / bb[0], n);
        double emin_pre = 0.0, emax_pre = 0.0, err_min = tol_min + 1.0, err_max = tol_max + 1.0;
        *eigmin = *eigmax = 0.0;
        while ((err_min > tol_min || err_max > tol_max) && j < maxit) {
            SGPU_CHECK(apply_h(ctx, SGPU_F64, 1, 0.0, Vj, ctx->ld, Vjp1, ctx->ld, nullptr, 0, 1.0, 0.0));
            pa[0] = Vj; pb[0] = Vjp1;
            SGPU_CHECK(sgpu_dot_many(ctx, (int)n, 1, pa, pb, &a[j + 1]));
            lz_orth<<<nb, 256, 0, ctx->stream>>>(Vjp1, Vj, Vjm1, a[j + 1], bb[j], n);
            pa[0] = Vjp1; pb[0] = Vjp1;
            SGPU_CHECK(sgpu_dot_many(ctx, (int)n, 1, pa, pb, &nrm2));
            bb[j + 1] = std::sqrt(nrm2);
            if (bb[j + 1] == 0.0) break;
            vscale<<<nb, 256, 0, ctx->stream>>>(Vj, Vjp1, 1.0 / bb[j + 1], n);
            tridiag_extremes(j + 2, a.data(), bb.data(), eigmin, eigmax);
            err_min = std::fabs(*eigmin - emin_pre);
            err_max = std::fabs(*eigmax - emax_pre);
            emin_pre = *eigmin;
            emax_pre = *eigmax;
            j++;
        }
        SGPU_HIP(hipGetLastError());
        return SGPU_OK;
    };
    status = run();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(V);
    if (iters) *iters = j;
    return status;
}

// ======================= host-pointer drop-ins ==============================

namespace {

// Temporary device blocks for the host-pointer entry points.
struct tmp_blocks {
    std::vector<sgpu_block *> b;
    ~tmp_blocks()
    {
        for (auto *p : b) sgpu_block_destroy(p);
    }
    int make(sgpu_context *ctx, int ncol, int dtype, sgpu_block **out)
    {
        SGPU_CHECK(sgpu_block_create(ctx, ncol, dtype, out));
        b.push_back(*out);
        return SGPU_OK;
    }
};

int check_host_args(const sgpu_context *ctx, int ncol, const void *x, int ldi, const void *y,
                    int ldo, const char *what)
{
    if (!ctx || ncol < 0 || (ncol > 0 && (!x || !y)) || ldi < ctx->nd || ldo < ctx->nd) {
        sgpu_set_error("%s: bad arguments (ncol %d, ldi %d, ldo %d, Nd %ld)", what, ncol, ldi, ldo,
                       ctx ? ctx->nd : 0L);
        return SGPU_ERR_ARG;
    }
    return SGPU_OK;
}

}  // namespace

extern "C" int sgpu_Lap_plus_diag_vec_mult(sgpu_context *ctx, int ncol, double a, double b,
                                           double c, const double *v, const double *x, int ldi,
                                           double *y, int ldo)
{
    SGPU_CHECK(check_host_args(ctx, ncol, x, ldi, y, ldo, "sgpu_Lap_plus_diag_vec_mult"));
    if (ncol == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *bx, *by;
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &bx));
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &by));
    SGPU_CHECK(sgpu_block_upload(bx, x, ldi, ncol));
    const bool diag = v != nullptr && std::fabs(b) >= 1e-14;
    double *dv = nullptr;
    if (diag) {
        void *s = nullptr;
        SGPU_CHECK(sgpu_scratch(ctx, ctx->nd * sizeof(double), &s));
        dv = (double *)s;
        SGPU_HIP(hipMemcpyAsync(dv, v, ctx->nd * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    SGPU_CHECK(sgpu_launch_stencil(ctx, SGPU_F64, ncol, a, diag ? b : 0.0, c, dv, bx->ptr, bx->ld,
                                   by->ptr, by->ld, nullptr, 0, 1.0, 0.0, true));
    return sgpu_block_download(by, y, ldo, ncol);
}

extern "C" int sgpu_Lap_vec_mult(sgpu_context *ctx, int ncol, double c, const double *x, int ldi,
                                 double *Lapx, int ldo)
{
    return sgpu_Lap_plus_diag_vec_mult(ctx, ncol, 1.0, 0.0, c, nullptr, x, ldi, Lapx, ldo);
}

extern "C" int sgpu_Vnl_vec_mult(sgpu_context *ctx, int ncol, const double *x, int ldi, double *Hx,
                                 int ldo)
{
    SGPU_CHECK(check_host_args(ctx, ncol, x, ldi, Hx, ldo, "sgpu_Vnl_vec_mult"));
    if (ncol == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *bx, *by;
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &bx));
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &by));
    SGPU_CHECK(sgpu_block_upload(bx, x, ldi, ncol));
    SGPU_CHECK(sgpu_block_upload(by, Hx, ldo, ncol));
    SGPU_CHECK(sgpu_launch_vnl(ctx, SGPU_F64, ncol, bx->ptr, bx->ld, by->ptr, by->ld, 1.0));
    return sgpu_block_download(by, Hx, ldo, ncol);
}

extern "C" int sgpu_Hamiltonian_vectors_mult(sgpu_context *ctx, int ncol, double c, const double *x,
                                             int ldi, double *Hx, int ldo)
{
    SGPU_CHECK(check_host_args(ctx, ncol, x, ldi, Hx, ldo, "sgpu_Hamiltonian_vectors_mult"));
    if (ncol == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *bx, *by;
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &bx));
    SGPU_CHECK(t.make(ctx, ncol, SGPU_F64, &by));
    SGPU_CHECK(sgpu_block_upload(bx, x, ldi, ncol));
    SGPU_CHECK(sgpu_hamiltonian_dev(ctx, c, bx, by));
    return sgpu_block_download(by, Hx, ldo, ncol);
}

extern "C" int sgpu_ChebyshevFiltering(sgpu_context *ctx, double *X, int ldi, double *Y, int ldo,
                                       int ncol, int m, double a, double b, double a0, int dtype)
{
    SGPU_CHECK(check_host_args(ctx, ncol, X, ldi, Y, ldo, "sgpu_ChebyshevFiltering"));
    if (dtype != SGPU_F64 && dtype != SGPU_F32) return SGPU_ERR_ARG;
    if (ncol == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *bx, *by;
    SGPU_CHECK(t.make(ctx, ncol, dtype, &bx));
    SGPU_CHECK(t.make(ctx, ncol, dtype, &by));
    SGPU_CHECK(sgpu_block_upload(bx, X, ldi, ncol));
    SGPU_CHECK(sgpu_chebyshev_filter_dev(ctx, bx, by, m, a, b, a0));
    SGPU_CHECK(sgpu_block_download(by, Y, ldo, ncol));
    return sgpu_block_download(bx, X, ldi, ncol);
}

extern "C" int sgpu_DP_Project_Hamiltonian(sgpu_context *ctx, const double *Y, int ldi, double *HY,
                                           int ldo, int ncol, double *Hp, double *Mp, int dtype)
{
    SGPU_CHECK(check_host_args(ctx, ncol, Y, ldi, HY, ldo, "sgpu_DP_Project_Hamiltonian"));
    if ((dtype != SGPU_F64 && dtype != SGPU_F32) || !Hp || !Mp) return SGPU_ERR_ARG;
    if (ncol == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *by, *bh;
    SGPU_CHECK(t.make(ctx, ncol, dtype, &by));
    SGPU_CHECK(t.make(ctx, ncol, dtype, &bh));
    SGPU_CHECK(sgpu_block_upload(by, Y, ldi, ncol));
    SGPU_CHECK(sgpu_project_dev(ctx, by, bh, Hp, Mp));
    return sgpu_block_download(bh, HY, ldo, ncol);
}

extern "C" int sgpu_DP_Solve_Generalized_EigenProblem(sgpu_context *ctx, int n, double *Hp,
                                                      double *Mp, double *lambda)
{
    if (!ctx || n <= 0 || !Hp || !Mp || !lambda) {
        sgpu_set_error("sgpu_DP_Solve_Generalized_EigenProblem: bad arguments");
        return SGPU_ERR_ARG;
    }
    // LAPACK would report non-finite input through info; the device path has
    // no such report, so it is refused here before anything is uploaded
    const long nn = (long)n * n;
    for (long i = 0; i < nn; i++)
        if (!std::isfinite(Hp[i]) || !std::isfinite(Mp[i])) {
            sgpu_set_error("sgpu_DP_Solve_Generalized_EigenProblem: non-finite entry in %s at %ld",
                           std::isfinite(Hp[i]) ? "Mp" : "Hp", i);
            return SGPU_ERR_ARG;
        }
    SGPU_HIP(hipSetDevice(ctx->device));
    return eigensolve(ctx, n, Hp, Mp, lambda);
}

extern "C" int sgpu_DP_Subspace_Rotation(sgpu_context *ctx, const double *Y, int ldy,
                                         const double *Q, int ns, double *Psi_rot, int ldo,
                                         int dtype)
{
    SGPU_CHECK(check_host_args(ctx, ns, Y, ldy, Psi_rot, ldo, "sgpu_DP_Subspace_Rotation"));
    if ((dtype != SGPU_F64 && dtype != SGPU_F32) || !Q) return SGPU_ERR_ARG;
    if (ns == 0) return SGPU_OK;
    SGPU_HIP(hipSetDevice(ctx->device));
    tmp_blocks t;
    sgpu_block *by, *bo;
    SGPU_CHECK(t.make(ctx, ns, dtype, &by));
    SGPU_CHECK(t.make(ctx, ns, dtype, &bo));
    SGPU_CHECK(sgpu_block_upload(by, Y, ldy, ns));
    SGPU_CHECK(sgpu_rotate_dev(ctx, by, Q, bo));
    return sgpu_block_download(bo, Psi_rot, ldo, ns);
}
