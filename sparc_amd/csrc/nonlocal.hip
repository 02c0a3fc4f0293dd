// Kleinman-Bylander nonlocal apply  out += scale * sum_J Chi_J Gamma_J (dV Chi_J' x)
// (SPARC Vnl_vec_mult, src/nlocVecRoutines.c:798-883).
//
// The reference gathers x at each atom's sphere points and calls two small
// dgemms per atom.  Here the projectors are re-laid out once per geometry in
// bricks of kBX x kBY x kBZ grid points (sgpu_set_nonlocal): for each brick a
// dense Phi_b (kBP points x np_b projector columns of every atom touching
// it).  The apply is then three launches over bricks x column blocks:
//
//  1. nlb_project_rs: part[brick rows][cols] = Phi_b' X_b on 16x16x4 MFMA,
//     64 (FP32) / 32 (FP64) columns per block, X streamed from memory
//     straight into the MFMA operands (16-byte loads of four consecutive
//     points per lane), each wavefront taking a quarter of the brick's
//     points, the four partial tiles summed in a fixed order through LDS.
//  2. nlb_reduce: alpha_s[r][c] = scale dV Gamma_r sum over the brick rows of
//     projector r (fixed order: deterministic, no atomics).
//  3. nlb_apply_rs: out_b += Phi_b beta_b on 16x16x4 MFMA with beta_b
//     (np_b x NC) gathered from alpha_s into LDS, computed with points as
//     MFMA rows so that each lane's results are four consecutive points of
//     one column (one 16-byte read-add-write).  Optionally it also forms the
//     next Chebyshev step's partials from the updated values (the brick
//     chain, SPARC_AMD_NL_CHAIN_BRICK=1).
//
// The LDS-staged forms (nlb_project, 16 columns per block; the transposed
// nlb_apply, 32 columns) stay selectable with SPARC_AMD_NLB=lds.  On the
// 96^3 x 512 FP32 operator case: 1.04 -> 0.74 ms (projection) and
// 1.23 -> 1.19 ms (apply) per application.
//
// Bricks touch each x value once per launch, so the gather redundancy of the
// reference (overlapping spheres read the same points once per atom) turns
// into dense zeros inside Phi_b; both are counted in the measured rates.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "sgpu_internal.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA per precision.  Operand layout (gfx950 matrix cores, checked
// on the device): A (16x4) element (i,k) in lane i + 16k, B (4x16) element
// (k,n) in lane n + 16k.  D (16x16) element (i,n) sits in lane n + 16(i/4),
// register i%4, for f32, and in lane n + 16(i%4), register i/4, for f64;
// drow(lk, r) gives the row i held in register r of a lane with lk = lane/16.
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = f32x4;
    using vec = float4;              // 16-byte staging vector
    static constexpr int NCP = 16;   // project: columns per block
    static constexpr int NCA = 32;   // sphere apply: columns per block
    static constexpr int NCB = 64;   // brick apply: columns per block
    __device__ static acc run(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int drow(int lk, int r) { return 4 * lk + r; }
};
template <>
struct Mfma<double> {
    using acc = f64x4;
    using vec = double2;
    static constexpr int NCP = 16;
    static constexpr int NCA = 16;
    static constexpr int NCB = 16;
    __device__ static acc run(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int drow(int lk, int r) { return lk + 4 * r; }
};

// 16-byte vectors for staging copies
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

constexpr int kWaves = 4;
constexpr int kRT = 4;            // 16-row tiles per projector chunk (64 rows)
constexpr int kLDQ = kBP + 4;     // LDS stride of a staged column (bank-conflict free)

__device__ __forceinline__ long brick_point(int q, int x0, int y0, int z0, int nx, int ny, int nz)
{
    const int qx = q % kBX, qy = (q / kBX) % kBY, qz = q / (kBX * kBY);
    const int gx = x0 + qx, gy = y0 + qy, gz = z0 + qz;
    if (gx >= nx || gy >= ny || gz >= nz) return -1;
    return ((long)gz * ny + gy) * nx + gx;
}

// part[brick row j][col] = sum_q Phi_b[q][j] X[col][q] on 16x16x4 MFMA, X
// streamed from global memory straight into the B operands (no LDS staging).
// The K index (points) is assigned so that a lane's 16-byte loads feed four
// MFMAs: a step covers one 16-point x-row of the brick, and in MFMA u of the
// step lane group lk supplies point 4 lk + u, so each lane loads the four
// consecutive points 4 lk .. 4 lk + 3 of its column li (one 16-byte access in
// FP32, two in FP64) and the 16 lanes of a group read 16 columns' segments of
// the row; the A operands (Phi_b rows of the same points) follow the same
// assignment, which is also the one nlb_apply_rs's fused projection uses, so
// both give the same bits.  NC columns per block share every Phi load (NC/16
// MFMAs per load); the next step's loads are issued before the current
// step's MFMAs.  The four waves split the 512 points and their partial tiles
// are summed in a fixed order (3 -> 2 -> 1 -> 0).  VEC: 16-byte loads (nx a
// multiple of 4, ldx of 16 / sizeof(T)); otherwise scalar loads with
// per-point checks.
template <typename T, int NC, bool VEC>
__global__ __launch_bounds__(256) void nlb_project_rs(const T *__restrict__ X, long ldx, int ncol,
                                                      int nx, int ny, int nz,
                                                      const int *__restrict__ brick_org,
                                                      const int *__restrict__ brick_np,
                                                      const long *__restrict__ brick_row,
                                                      const T *__restrict__ phi,
                                                      T *__restrict__ part)
{
    using M = Mfma<T>;
    using V = typename Vec16<T>::type;
    constexpr int VW = 16 / sizeof(T);         // elements per 16-byte access
    constexpr int NV = 4 / VW;                 // 16-byte accesses per 4 points
    constexpr int NCT = NC / 16;
    constexpr int NSTEP = kBP / kWaves / 16;   // 16-point steps per wave
    __shared__ typename M::acc red[kRT * NCT][64];
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int b = lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int x0 = brick_org[3 * b], y0 = brick_org[3 * b + 1], z0 = brick_org[3 * b + 2];
    const int np = brick_np[b];
    const long row0 = brick_row[b];
    const T *ph = phi + row0 * kBP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int qw = wave * (kBP / kWaves) + 4 * lk;  // first point of this lane, step 0
    // column offsets of my NCT columns (clamped to column 0 past ncol, zeroed after the load)
    long coff[NCT];
    bool cok[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const int col = col0 + ct * 16 + li;
        cok[ct] = col < ncol;
        coff[ct] = cok[ct] ? (long)col * ldx : 0L;
    }
    auto load_x = [&](int s, T (&xv)[NCT][4]) {
        const int q = qw + s * 16;
        if constexpr (VEC) {
            // the four points share y and z and lie inside the grid along x
            // together (x0 a multiple of 16, nx of VW)
            const long pt = brick_point(q, x0, y0, z0, nx, ny, nz);
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const V w = *reinterpret_cast<const V *>(X + coff[ct] + (pt >= 0 ? pt : 0L) + v * VW);
#pragma unroll
                    for (int e = 0; e < VW; e++)
                        xv[ct][v * VW + e] = (pt >= 0 && cok[ct]) ? w[e] : T(0);
                }
        } else {
            long pts[4];
#pragma unroll
            for (int u = 0; u < 4; u++) pts[u] = brick_point(q + u, x0, y0, z0, nx, ny, nz);
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    T v = X[coff[ct] + (pts[u] >= 0 ? pts[u] : 0L)];
                    xv[ct][u] = (pts[u] >= 0 && cok[ct]) ? v : T(0);
                }
        }
    };
    for (int j0 = 0; j0 < np; j0 += 16 * kRT) {
        typename M::acc acc[kRT][NCT];
#pragma unroll
        for (int rt = 0; rt < kRT; rt++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = typename M::acc{0, 0, 0, 0};
        // rows past np read row np - 1 (in range) and are never stored
        int jr[kRT];
#pragma unroll
        for (int rt = 0; rt < kRT; rt++) jr[rt] = min(j0 + rt * 16 + li, np - 1);
        auto load_phi = [&](int s, T (&av)[4][kRT]) {
            const int q = qw + s * 16;
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++) av[u][rt] = ph[(long)(q + u) * np + jr[rt]];
        };
        T xc[NCT][4], xn[NCT][4];
        T ac[4][kRT], an[4][kRT];
        load_x(0, xc);
        load_phi(0, ac);
        // not unrolled: one step's loads in flight while the previous step's
        // MFMAs run (a full unroll hoisted every load and spilled)
#pragma unroll 1
        for (int s = 0; s < NSTEP; s++) {
            const int sn = min(s + 1, NSTEP - 1);  // the last step reloads itself (cache hits)
            load_x(sn, xn);
            load_phi(sn, an);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++) {
                    if (j0 + rt * 16 < np) {
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++)
                            acc[rt][ct] = M::run(ac[u][rt], xc[ct][u], acc[rt][ct]);
                    }
                }
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int u = 0; u < 4; u++) xc[ct][u] = xn[ct][u];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++) ac[u][rt] = an[u][rt];
        }
        // fixed-order cross-wave sum: wave 3 -> 2 -> 1 -> 0
        for (int w = kWaves - 1; w >= 1; w--) {
            if (wave == w)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) red[rt * NCT + ct][lane] = acc[rt][ct];
            __syncthreads();
            if (wave == w - 1)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) acc[rt][ct] += red[rt * NCT + ct][lane];
            __syncthreads();
        }
        if (wave == 0) {
#pragma unroll
            for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = j0 + rt * 16 + M::drow(lk, r);
                        const int col = col0 + ct * 16 + li;
                        if (j < np && col < ncol) part[(row0 + j) * ncol + col] = acc[rt][ct][r];
                    }
        }
    }
}

// part[brick row j][col] = sum_q Phi_b[q][j] X[col][q]  on 16x16x4 MFMA.
// The brick's NC columns are staged in LDS column-major; the four waves split
// the 512 points and their partial tiles are summed in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void nlb_project(const T *__restrict__ X, long ldx, int ncol,
                                                   int nx, int ny, int nz,
                                                   const int *__restrict__ brick_org,
                                                   const int *__restrict__ brick_np,
                                                   const long *__restrict__ brick_row,
                                                   const T *__restrict__ phi, T *__restrict__ part)
{
    using M = Mfma<T>;
    constexpr int NC = M::NCP, NCT = NC / 16;
    __shared__ __align__(16) T xs[NC * kLDQ];
    __shared__ typename M::acc red[kRT * NCT][64];
    // column blocks of one brick are consecutive logical ids -> one XCD/L2
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int b = lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int x0 = brick_org[3 * b], y0 = brick_org[3 * b + 1], z0 = brick_org[3 * b + 2];
    constexpr int VW = sizeof(typename M::vec) / sizeof(T);
    // All loads of a thread are issued before its first LDS store, from
    // clamped addresses (X[0] for points outside the grid or columns past
    // ncol, zeroed after the load): a guarded load became a branch that
    // waited for each load in turn, one memory latency per element.
    if (nx % VW == 0 && ldx % VW == 0) {
        // 16-byte loads: a vector never straddles the grid edge because
        // x0 is a multiple of kBX and nx a multiple of VW
        constexpr int kPer = kBP * NC / VW / 256;
        typename M::vec v[kPer];
        bool ok[kPer];
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int e = threadIdx.x + 256 * j;
            const int c = e / (kBP / VW), q = (e - c * (kBP / VW)) * VW;
            const long pt = brick_point(q, x0, y0, z0, nx, ny, nz);
            ok[j] = pt >= 0 && col0 + c < ncol;
            v[j] = *reinterpret_cast<const typename M::vec *>(X + (ok[j] ? (long)(col0 + c) * ldx + pt : 0L));
        }
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int e = threadIdx.x + 256 * j;
            const int c = e / (kBP / VW), q = (e - c * (kBP / VW)) * VW;
            if (!ok[j]) v[j] = typename M::vec{};
            *reinterpret_cast<typename M::vec *>(xs + c * kLDQ + q) = v[j];
        }
    } else {
        constexpr int kPer = kBP * NC / 256;
        T v[kPer];
        bool ok[kPer];
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int e = threadIdx.x + 256 * j;
            const int c = e / kBP, q = e - c * kBP;
            const long pt = brick_point(q, x0, y0, z0, nx, ny, nz);
            ok[j] = pt >= 0 && col0 + c < ncol;
            v[j] = X[ok[j] ? (long)(col0 + c) * ldx + pt : 0L];
        }
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int e = threadIdx.x + 256 * j;
            const int c = e / kBP, q = e - c * kBP;
            xs[c * kLDQ + q] = ok[j] ? v[j] : T(0);
        }
    }
    __syncthreads();
    const int np = brick_np[b];
    const long row0 = brick_row[b];
    const T *ph = phi + row0 * kBP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int qs = wave * (kBP / kWaves);
    for (int j0 = 0; j0 < np; j0 += 16 * kRT) {
        typename M::acc acc[kRT][NCT];
#pragma unroll
        for (int rt = 0; rt < kRT; rt++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = typename M::acc{0, 0, 0, 0};
        // kU k-steps at a time: their Phi loads are all issued before the
        // MFMAs that consume them (one dependent global load per MFMA step
        // left the kernel latency-bound at ~1 TB/s)
        constexpr int kU = 8;
        for (int k = qs; k < qs + kBP / kWaves; k += 4 * kU) {
            T av[kU][kRT];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int q = k + 4 * u + lk;
#pragma unroll
                for (int rt = 0; rt < kRT; rt++) {
                    const int j = j0 + rt * 16 + li;
                    av[u][rt] = (j0 + rt * 16 < np && j < np) ? ph[(long)q * np + j] : T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int q = k + 4 * u + lk;
                T bv[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) bv[ct] = xs[(ct * 16 + li) * kLDQ + q];
#pragma unroll
                for (int rt = 0; rt < kRT; rt++) {
                    if (j0 + rt * 16 < np) {
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = M::run(av[u][rt], bv[ct], acc[rt][ct]);
                    }
                }
            }
        }
        // fixed-order cross-wave sum: wave 3 -> 2 -> 1 -> 0
        for (int w = kWaves - 1; w >= 1; w--) {
            if (wave == w)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) red[rt * NCT + ct][lane] = acc[rt][ct];
            __syncthreads();
            if (wave == w - 1)
#pragma unroll
                for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) acc[rt][ct] += red[rt * NCT + ct][lane];
            __syncthreads();
        }
        if (wave == 0) {
#pragma unroll
            for (int rt = 0; rt < kRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = j0 + rt * 16 + M::drow(lk, r);
                        const int col = col0 + ct * 16 + li;
                        if (j < np && col < ncol) part[(row0 + j) * ncol + col] = acc[rt][ct][r];
                    }
        }
    }
}

template <typename T>
__global__ void nlb_reduce(const T *__restrict__ part, int ncol, int n_alpha,
                           const int *__restrict__ arow_ptr, const int *__restrict__ arow_src,
                           const double *__restrict__ gamma, double scale, T *__restrict__ alpha_s)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_alpha * ncol) return;
    const int r = (int)(t / ncol);
    const int col = (int)(t - (long)r * ncol);
    // in the order of arow_src, the loads issued eight at a time (the row's
    // ~100 partials are otherwise one load latency each)
    T s = T(0);
    const int e1 = arow_ptr[r + 1];
    int e = arow_ptr[r];
    for (; e + 8 <= e1; e += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[(long)arow_src[e + u] * ncol + col];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; e < e1; e++) s += part[(long)arow_src[e] * ncol + col];
    alpha_s[(long)r * ncol + col] = (T)(scale * (gamma ? gamma[r] : 1.0)) * s;
}

// out_b[q][col] += sum_j Phi_b[q][j] beta_b[j][col]  on 16x16x4 MFMA, computed
// as the transposed tile (cols x points) = beta_b' Phi_b' so that the B
// operand (Phi_b column-major, points contiguous) is a coalesced load and
// each output register row is 16 consecutive points of one column.
template <typename T>
__global__ __launch_bounds__(256) void nlb_apply(T *__restrict__ out, long ldo, int ncol, int nx,
                                                 int ny, int nz, const int *__restrict__ brick_org,
                                                 const int *__restrict__ brick_np,
                                                 const long *__restrict__ brick_row,
                                                 const int *__restrict__ row_alpha,
                                                 const T *__restrict__ phic,
                                                 const T *__restrict__ alpha_s)
{
    using M = Mfma<T>;
    constexpr int NC = M::NCA, NCT = NC / 16, LDB = NC * 3 / 2;
    extern __shared__ unsigned char smem_raw[];
    T *beta = reinterpret_cast<T *>(smem_raw);  // np x LDB
    // column blocks of one brick are consecutive logical ids -> one XCD/L2
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int b = lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int x0 = brick_org[3 * b], y0 = brick_org[3 * b + 1], z0 = brick_org[3 * b + 2];
    const int np = brick_np[b];
    const long row0 = brick_row[b];
    // beta gathered four elements per thread at a time, every load issued
    // before the stores (clamped indices: the tail repeats the last element)
    for (int base = 0; base < np * NC; base += 4 * 256) {
        int jj[4], cc[4], rr[4];
        T vv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = min(base + (int)threadIdx.x + 256 * u, np * NC - 1);
            jj[u] = e / NC;
            cc[u] = e - jj[u] * NC;
            rr[u] = row_alpha[row0 + jj[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = rr[u] >= 0 && col0 + cc[u] < ncol;
            vv[u] = alpha_s[ok ? (long)rr[u] * ncol + col0 + cc[u] : 0L];
            if (!ok) vv[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) beta[jj[u] * LDB + cc[u]] = vv[u];
    }
    __syncthreads();
    const T *pc = phic + row0 * kBP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int pt_tile = wave; pt_tile < kBP / 16; pt_tile += kWaves) {
        const int q0 = pt_tile * 16;
        // the tile's current output values are loaded before the MFMA chain
        // so that their latency overlaps it
        const long pt = brick_point(q0 + li, x0, y0, z0, nx, ny, nz);
        typename M::acc acc[NCT], ov[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            acc[ct] = typename M::acc{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = col0 + ct * 16 + M::drow(lk, r);
                ov[ct][r] = (pt >= 0 && col < ncol) ? out[(long)col * ldo + pt] : T(0);
            }
        }
        // Phi loads of up to 8 j-steps issued before their MFMAs
        for (int jb = 0; jb < np; jb += 32) {
            T bqs[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = jb + 4 * u;
                bqs[u] = j < np ? pc[(long)(j + lk) * kBP + q0 + li] : T(0);  // B (j, point)
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = jb + 4 * u;
                if (j < np) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++)                               // A (col, j)
                        acc[ct] = M::run(beta[(j + lk) * LDB + ct * 16 + li], bqs[u], acc[ct]);
                }
            }
        }
        if (pt < 0) continue;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = col0 + ct * 16 + M::drow(lk, r);
                if (col < ncol) out[(long)col * ldo + pt] = ov[ct][r] + acc[ct][r];
            }
        }
    }
}

// out_b += Phi_b beta_b for NC columns, computed untransposed on 16x16x4 MFMA:
// D (16 points x 16 columns) = Phi_b(points, j) beta_b(j, columns) over one
// brick x-row, with the MFMA row of each point permuted (FP64: row
// 4 (p % 4) + p / 4 for point p; FP32: the identity) so that the four D
// registers of lane (li, lk) are the consecutive points 4 lk .. 4 lk + 3 of
// column li.  The read-add-write of out is then one 16-byte (FP32) or two
// 16-byte (FP64) accesses per lane and column tile, the 16 lanes of a group
// covering 16 columns' 64-byte (128-byte) segments of the row, and the A
// operand (Phi_b column-major, points contiguous) is a coalesced load shared
// by the NC / 16 column tiles.
// With PROJ the updated values, still in registers, are also the B operands
// of the projection Phi_b' out (point 4 lk + u in MFMA u, as in
// nlb_project_rs): the pass writes the partials of out for the next
// application of a Chebyshev chain into part_next, all RTM row tiles
// accumulated at once (np <= 16 RTM), the four waves' tiles summed in a fixed
// order.  Every lane's loads of a row step are issued before its MFMAs.
template <typename T, int NC, int RTM, bool PROJ>
__global__ __launch_bounds__(256) void nlb_apply_rs(T *__restrict__ out, long ldo, int ncol, int nx,
                                                    int ny, int nz, const int *__restrict__ brick_org,
                                                    const int *__restrict__ brick_np,
                                                    const long *__restrict__ brick_row,
                                                    const int *__restrict__ row_alpha,
                                                    const T *__restrict__ phi,
                                                    const T *__restrict__ phic,
                                                    const T *__restrict__ alpha_s,
                                                    T *__restrict__ part_next)
{
    using M = Mfma<T>;
    using V = typename Vec16<T>::type;
    constexpr int VW = 16 / sizeof(T);        // elements per 16-byte access
    constexpr int NV = 4 / VW;                // 16-byte accesses per 4 points
    constexpr int NCT = NC / 16;
    constexpr int LDB = NC + 64 / sizeof(T);  // LDS stride of beta rows (bank spread)
    constexpr int PW = kBP / kWaves;          // points per wave
    extern __shared__ unsigned char smem_raw[];
    T *beta = reinterpret_cast<T *>(smem_raw);  // np x LDB
    __shared__ typename M::acc red[PROJ ? RTM * NCT : 1][64];
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int b = lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int x0 = brick_org[3 * b], y0 = brick_org[3 * b + 1], z0 = brick_org[3 * b + 2];
    const int np = brick_np[b];
    const long row0 = brick_row[b];
    // beta gathered four elements per thread at a time, every load issued
    // before the stores (clamped indices: the tail repeats the last element)
    for (int base = 0; base < np * NC; base += 4 * 256) {
        int jj[4], cc[4], rr[4];
        T vv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = min(base + (int)threadIdx.x + 256 * u, np * NC - 1);
            jj[u] = e / NC;
            cc[u] = e - jj[u] * NC;
            rr[u] = row_alpha[row0 + jj[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = rr[u] >= 0 && col0 + cc[u] < ncol;
            vv[u] = alpha_s[ok ? (long)rr[u] * ncol + col0 + cc[u] : 0L];
            if (!ok) vv[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) beta[jj[u] * LDB + cc[u]] = vv[u];
    }
    __syncthreads();
    const T *pc = phic + row0 * kBP;
    const T *ph = phi + row0 * kBP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // MFMA row li of a tile holds point prow (the D layout puts row i of
    // lane group lk in register r at i = 4 lk + r (FP32) or lk + 4 r (FP64))
    const int prow = sizeof(T) == 4 ? li : 4 * (li & 3) + (li >> 2);
    long coff[NCT];
    bool cok[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const int col = col0 + ct * 16 + li;
        cok[ct] = col < ncol;
        coff[ct] = cok[ct] ? (long)col * ldo : 0L;
    }
    typename M::acc pacc[PROJ ? RTM : 1][NCT];
    if constexpr (PROJ) {
#pragma unroll
        for (int rt = 0; rt < RTM; rt++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) pacc[rt][ct] = typename M::acc{0, 0, 0, 0};
    }
    int jr[PROJ ? RTM : 1];
    if constexpr (PROJ) {
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) jr[rt] = min(rt * 16 + li, np - 1);
    }
#pragma unroll 1
    for (int t = 0; t < PW / 16; t++) {
        const int q0 = wave * PW + t * 16;
        // my four points 4 lk .. 4 lk + 3 of the row (all inside or all
        // outside the grid along y / z; along x when nx is a multiple of 4)
        long pts[4];
#pragma unroll
        for (int r = 0; r < 4; r++) pts[r] = brick_point(q0 + 4 * lk + r, x0, y0, z0, nx, ny, nz);
        const bool vec = pts[0] >= 0 && pts[3] == pts[0] + 3 && ((ldo & (VW - 1)) == 0) &&
                         (reinterpret_cast<uintptr_t>(out + pts[0]) & 15) == 0;
        // current values of out, loaded before the MFMA chain
        T ov[NCT][4];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            if (vec) {
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const V w = *reinterpret_cast<const V *>(out + coff[ct] + pts[0] + v * VW);
#pragma unroll
                    for (int e = 0; e < VW; e++) ov[ct][v * VW + e] = w[e];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++) ov[ct][r] = out[coff[ct] + (pts[r] >= 0 ? pts[r] : 0L)];
            }
        }
        typename M::acc acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[ct] = typename M::acc{0, 0, 0, 0};
        const T *pcq = pc + q0 + prow;
        for (int jb = 0; jb < np; jb += 32) {
            T aq[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = min(jb + 4 * u + lk, np - 1);
                aq[u] = pcq[(long)j * kBP];  // A (point, j)
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = jb + 4 * u;
                if (j < np) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++)  // B (j, col)
                        acc[ct] = M::run(aq[u], beta[(j + lk) * LDB + ct * 16 + li], acc[ct]);
                }
            }
        }
        // new values: D register r of lane (li, lk) is point 4 lk + r of column li
        T y[NCT][4];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int r = 0; r < 4; r++) y[ct][r] = ov[ct][r] + acc[ct][r];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            if (!cok[ct]) continue;
            if (vec) {
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    V w;
#pragma unroll
                    for (int e = 0; e < VW; e++) w[e] = y[ct][v * VW + e];
                    *reinterpret_cast<V *>(out + coff[ct] + pts[0] + v * VW) = w;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (pts[r] >= 0) out[coff[ct] + pts[r]] = y[ct][r];
            }
        }
        if constexpr (PROJ) {
            // points outside the grid or columns past ncol contribute zero
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (pts[r] < 0 || !cok[ct]) y[ct][r] = T(0);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                T aph[RTM];
#pragma unroll
                for (int rt = 0; rt < RTM; rt++) aph[rt] = ph[(long)(q0 + 4 * lk + u) * np + jr[rt]];
#pragma unroll
                for (int rt = 0; rt < RTM; rt++) {
                    if (rt * 16 < np) {
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++)
                            pacc[rt][ct] = M::run(aph[rt], y[ct][u], pacc[rt][ct]);
                    }
                }
            }
        }
    }
    if constexpr (PROJ) {
        for (int w = kWaves - 1; w >= 1; w--) {
            if (wave == w)
#pragma unroll
                for (int rt = 0; rt < RTM; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) red[rt * NCT + ct][lane] = pacc[rt][ct];
            __syncthreads();
            if (wave == w - 1)
#pragma unroll
                for (int rt = 0; rt < RTM; rt++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) pacc[rt][ct] += red[rt * NCT + ct][lane];
            __syncthreads();
        }
        if (wave == 0) {
#pragma unroll
            for (int rt = 0; rt < RTM; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = rt * 16 + M::drow(lk, r);
                        const int col = col0 + ct * 16 + li;
                        if (j < np && col < ncol) part_next[(row0 + j) * ncol + col] = pacc[rt][ct][r];
                    }
        }
    }
}

// ---- sphere layout: one block per (influence entry, column block) ----------
//
// For projectors whose spheres cover a small part of a brick (real
// pseudopotentials, rc ~ 2 Bohr: ~500 points against ~6000 brick points) the
// work follows the reference's gather / GEMM / scatter per atom instead:
//
//  1. nls_project: part[e][p][col] = sum_q Chi_e[q][p] X[col][pos_e[q]] on
//     16x16x4 MFMA; the entry's points are staged in LDS 256 at a time
//     (gathered, consecutive threads on consecutive points), the four waves
//     split each chunk and their tiles are summed in a fixed order.
//  2. nls_apply, one launch per colour: beta = scale dV Gamma sum over the
//     atom's images of part (fixed order), then out[col][pos_e[q]] +=
//     sum_p Chi_e[q][p] beta[p][col] as the transposed tile (cols x points)
//     so that every output register row is 16 consecutive points of one
//     column.  Entries of one colour share no grid point, so the
//     read-add-write needs no atomics and the result is deterministic.
constexpr int kQC = 256;           // points staged per chunk (multiple of 64)
constexpr int kQLD = kQC + 4;      // LDS stride of a staged column

template <typename T>
__global__ __launch_bounds__(256) void nls_project(const T *__restrict__ X, long ldx, int ncol,
                                                   const int *__restrict__ ent_info,
                                                   const long *__restrict__ ent_chi,
                                                   const int *__restrict__ gpos,
                                                   const T *__restrict__ chir, int max_np,
                                                   T *__restrict__ part)
{
    using M = Mfma<T>;
    constexpr int NC = M::NCP, NCT = NC / 16, MRT = 4;
    __shared__ __align__(16) T xs[NC * kQLD];
    __shared__ typename M::acc red[MRT * NCT][64];
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int e = lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int qpad = ent_info[6 * e], poff = ent_info[6 * e + 1];
    const int np = ent_info[6 * e + 2], npad = ent_info[6 * e + 3];
    const int nrt = npad / 16;
    const T *ch = chir + ent_chi[e];
    const int *gp = gpos + poff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    typename M::acc acc[MRT][NCT];
#pragma unroll
    for (int rt = 0; rt < MRT; rt++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = typename M::acc{0, 0, 0, 0};
    for (int qc = 0; qc < qpad; qc += kQC) {
        const int qn = min(kQC, qpad - qc);
        // eight elements per thread at a time: point indices, then values,
        // then stores (clamped: the tail repeats the last element)
        for (int base = 0; base < qn * NC; base += 8 * 256) {
            int cc[8], qq[8], gg[8];
            T vv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int t = min(base + (int)threadIdx.x + 256 * u, qn * NC - 1);
                cc[u] = t / qn;
                qq[u] = t - cc[u] * qn;
                gg[u] = gp[qc + qq[u]];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const bool ok = gg[u] >= 0 && col0 + cc[u] < ncol;
                vv[u] = X[ok ? (long)(col0 + cc[u]) * ldx + gg[u] : 0L];
                if (!ok) vv[u] = T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) xs[cc[u] * kQLD + qq[u]] = vv[u];
        }
        __syncthreads();
        const int qw = qn / 4;  // a multiple of 16
        for (int k = wave * qw; k < (wave + 1) * qw; k += 16) {
            // four k-steps: Chi loads first, then the MFMAs
            T av[4][MRT];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const T *arow = ch + (long)(qc + k + 4 * u + lk) * npad + li;
#pragma unroll
                for (int rt = 0; rt < MRT; rt++) av[u][rt] = rt < nrt ? arow[rt * 16] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = k + 4 * u + lk;
                T bv[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) bv[ct] = xs[(ct * 16 + li) * kQLD + q];
#pragma unroll
                for (int rt = 0; rt < MRT; rt++) {
                    if (rt < nrt) {
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = M::run(av[u][rt], bv[ct], acc[rt][ct]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // fixed-order cross-wave sum: wave 3 -> 2 -> 1 -> 0
    for (int w = kWaves - 1; w >= 1; w--) {
        if (wave == w)
#pragma unroll
            for (int rt = 0; rt < MRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) red[rt * NCT + ct][lane] = acc[rt][ct];
        __syncthreads();
        if (wave == w - 1)
#pragma unroll
            for (int rt = 0; rt < MRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) acc[rt][ct] += red[rt * NCT + ct][lane];
        __syncthreads();
    }
    if (wave == 0) {
#pragma unroll
        for (int rt = 0; rt < MRT; rt++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int p = rt * 16 + M::drow(lk, r);
                    const int col = col0 + ct * 16 + li;
                    if (p < np && col < ncol) part[((long)e * max_np + p) * ncol + col] = acc[rt][ct][r];
                }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void nls_apply(T *__restrict__ out, long ldo, int ncol, int e0,
                                                 const int *__restrict__ ent_info,
                                                 const long *__restrict__ ent_chi,
                                                 const int *__restrict__ gpos,
                                                 const T *__restrict__ chit,
                                                 const int *__restrict__ atom_img_ptr,
                                                 const int *__restrict__ atom_img,
                                                 const int *__restrict__ atom_row,
                                                 const double *__restrict__ gamma,
                                                 double scale_dv, const T *__restrict__ part,
                                                 int max_np)
{
    using M = Mfma<T>;
    constexpr int NC = M::NCA, NCT = NC / 16, LDB = NC * 3 / 2;
    __shared__ T beta[64 * LDB];
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int e = e0 + lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int qpad = ent_info[6 * e], poff = ent_info[6 * e + 1];
    const int np = ent_info[6 * e + 2], npad = ent_info[6 * e + 3], atom = ent_info[6 * e + 4];
    const int i0 = atom_img_ptr[atom], i1 = atom_img_ptr[atom + 1], row = atom_row[atom];
    // beta = scale dV Gamma sum over the atom's images of part, four
    // elements per thread at a time with every load before the stores
    // (clamped: the tail repeats the last element)
    {
        const int img0 = atom_img[i0];
        for (int base = 0; base < npad * NC; base += 4 * 256) {
            int pp[4], cc[4];
            bool ok[4];
            T vv[4];
            double gm[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = min(base + (int)threadIdx.x + 256 * u, npad * NC - 1);
                pp[u] = t / NC;
                cc[u] = t - pp[u] * NC;
                ok[u] = pp[u] < np && col0 + cc[u] < ncol;
                vv[u] = part[ok[u] ? ((long)img0 * max_np + pp[u]) * ncol + col0 + cc[u] : 0L];
                gm[u] = gamma[row + (ok[u] ? pp[u] : 0)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                T v = ok[u] ? vv[u] : T(0);
                if (ok[u]) {
                    for (int i = i0 + 1; i < i1; i++)
                        v += part[((long)atom_img[i] * max_np + pp[u]) * ncol + col0 + cc[u]];
                    v *= (T)(scale_dv * gm[u]);
                }
                beta[pp[u] * LDB + cc[u]] = v;
            }
        }
    }
    __syncthreads();
    const T *ct_ = chit + ent_chi[e];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int qt = wave; qt < qpad / 16; qt += kWaves) {
        const int q0 = qt * 16;
        // current output values loaded ahead of the MFMA chain
        const int g = gpos[poff + q0 + li];
        typename M::acc acc[NCT], ov[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            acc[ct] = typename M::acc{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = col0 + ct * 16 + M::drow(lk, r);
                ov[ct][r] = (g >= 0 && col < ncol) ? out[(long)col * ldo + g] : T(0);
            }
        }
        for (int jb = 0; jb < npad; jb += 16) {
            // npad is a multiple of 16: four p-steps, Chi loads first
            T bqs[4];
#pragma unroll
            for (int u = 0; u < 4; u++) bqs[u] = ct_[(long)(jb + 4 * u + lk) * qpad + q0 + li];  // B (p, point)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = jb + 4 * u;
#pragma unroll
                for (int ct = 0; ct < NCT; ct++)                    // A (col, p)
                    acc[ct] = M::run(beta[(j + lk) * LDB + ct * 16 + li], bqs[u], acc[ct]);
            }
        }
        if (g < 0) continue;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = col0 + ct * 16 + M::drow(lk, r);
                if (col < ncol) out[(long)col * ldo + g] = ov[ct][r] + acc[ct][r];
            }
        }
    }
}

// Orders a wavefront's LDS writes before its later LDS reads and the reads
// before later writes, for data exchanged between lanes of one wave only:
// LDS instructions of a wave execute in order, so it is enough that the
// compiler moves no LDS access across this point.
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Chained applies (the Chebyshev recurrence, where every Hamiltonian output
// is the next input), sphere layout with one colour (no two entries share a
// grid point): after an entry's apply its points hold final values, so the
// entry's partials for the next application are formed here from the tile
// still in registers instead of re-gathering the block in nls_project.  Each
// wave's 16-point x NC-column result is transposed through a wave-private
// LDS slice into the MFMA B layout (points along k); Chi is the A operand
// from the [p][qpad] copy.  The partials go to the other partials buffer
// (the atom's images still read this application's partials for beta).
// Wave w takes, chunk by chunk, the quarter of the points that nls_project
// gives it, in the same order, and the waves are summed in the same fixed
// order, so the partials are bitwise those of the unfused path.
// The waves run independently between the beta staging and the final sum:
// a tile's grid indices are read two tiles ahead and its current output
// values one tile ahead, and all of a tile's Chi operands are issued before
// its first MFMA.  MRT: 16-row projector tiles held (npad <= 16 MRT).  CHL:
// each kQC-point chunk of the entry's Chi is first copied to LDS by the
// whole block (coalesced 16-byte loads) and the operands are read there.
template <typename T, int MRT, bool CHL>
__global__ __launch_bounds__(256) void nls_apply_proj(T *__restrict__ out, long ldo, int ncol,
                                                      int e0, const int *__restrict__ ent_info,
                                                      const long *__restrict__ ent_chi,
                                                      const int *__restrict__ gpos,
                                                      const T *__restrict__ chit,
                                                      const int *__restrict__ atom_img_ptr,
                                                      const int *__restrict__ atom_img,
                                                      const int *__restrict__ atom_row,
                                                      const double *__restrict__ gamma,
                                                      double scale_dv, const T *__restrict__ part,
                                                      int max_np, T *__restrict__ part_next)
{
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NC = M::NCA, NCT = NC / 16, LDB = NC * 3 / 2, LDY = NC + 1;
    constexpr int LDQ = kQC + 16;  // LDS stride of a staged Chi row
    __shared__ T beta[16 * MRT * LDB];
    __shared__ T ytile[kWaves * 16 * LDY];
    __shared__ acc_t red[MRT * NCT][64];
    __shared__ __align__(16) T chs[CHL ? 16 * MRT * LDQ : 1];
    const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int e = e0 + lid / gridDim.x;
    const int col0 = (lid % gridDim.x) * NC;
    const int qpad = ent_info[6 * e], poff = ent_info[6 * e + 1];
    const int np = ent_info[6 * e + 2], npad = ent_info[6 * e + 3], atom = ent_info[6 * e + 4];
    const int nrt = npad / 16;
    const int i0 = atom_img_ptr[atom], i1 = atom_img_ptr[atom + 1], row = atom_row[atom];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // the wave's tiles: chunk t / 4 of kQC points, tile t % 4 of the wave's
    // quarter of the chunk (every chunk but the last holds four per wave)
    const int ntile = qpad / 64;
    auto tile_q0 = [&](int t) {
        const int qc = (t >> 2) * kQC;
        return qc + wave * (min(kQC, qpad - qc) >> 2) + (t & 3) * 16;
    };
    const int *gp = gpos + poff;
    auto load_out = [&](acc_t *ov, int g) {
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = col0 + ct * 16 + M::drow(lk, r);
                const bool ok = g >= 0 && col < ncol;
                const T v = out[ok ? (long)col * ldo + g : 0L];
                ov[ct][r] = ok ? v : T(0);
            }
    };
    // the first tile's outputs are requested before the beta staging
    int g_cur = gp[tile_q0(0) + li];
    int g_nxt = ntile > 1 ? gp[tile_q0(1) + li] : -1;
    acc_t ov[NCT];
    load_out(ov, g_cur);
    // beta = scale dV Gamma sum over the atom's images of part, four
    // elements per thread at a time with every load before the stores
    // (clamped: the tail repeats the last element)
    {
        const int img0 = atom_img[i0];
        for (int base = 0; base < npad * NC; base += 4 * 256) {
            int pp[4], cc[4];
            bool ok[4];
            T vv[4];
            double gm[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = min(base + (int)threadIdx.x + 256 * u, npad * NC - 1);
                pp[u] = t / NC;
                cc[u] = t - pp[u] * NC;
                ok[u] = pp[u] < np && col0 + cc[u] < ncol;
                vv[u] = part[ok[u] ? ((long)img0 * max_np + pp[u]) * ncol + col0 + cc[u] : 0L];
                gm[u] = gamma[row + (ok[u] ? pp[u] : 0)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                T v = ok[u] ? vv[u] : T(0);
                if (ok[u]) {
                    for (int i = i0 + 1; i < i1; i++)
                        v += part[((long)atom_img[i] * max_np + pp[u]) * ncol + col0 + cc[u]];
                    v *= (T)(scale_dv * gm[u]);
                }
                beta[pp[u] * LDB + cc[u]] = v;
            }
        }
    }
    __syncthreads();
    const T *ct_ = chit + ent_chi[e];
    T *yw = ytile + wave * 16 * LDY;
    acc_t pacc[MRT][NCT];
#pragma unroll
    for (int rt = 0; rt < MRT; rt++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) pacc[rt][ct] = acc_t{0, 0, 0, 0};
    for (int t = 0; t < ntile; t++) {
        const int q0 = tile_q0(t);
        const int qc = (t >> 2) * kQC;
        if (CHL && (t & 3) == 0) {
            // stage Chi rows 0..npad of points qc .. qc + qn (all loads first)
            using V = typename Vec16<T>::type;
            constexpr int VW = sizeof(V) / sizeof(T), kS = 16 * MRT * kQC / VW / 256;
            const int qv = min(kQC, qpad - qc) / VW, nv = npad * qv;
            V v[kS];
#pragma unroll
            for (int j = 0; j < kS; j++) {
                const int f = min((int)threadIdx.x + 256 * j, nv - 1);
                const int pr = f / qv, qq = (f - pr * qv) * VW;
                v[j] = *reinterpret_cast<const V *>(ct_ + (long)pr * qpad + qc + qq);
            }
            __syncthreads();  // the previous chunk's operands are read
#pragma unroll
            for (int j = 0; j < kS; j++) {
                const int f = threadIdx.x + 256 * j;
                if (f < nv) {
                    const int pr = f / qv, qq = (f - pr * qv) * VW;
                    *reinterpret_cast<V *>(chs + pr * LDQ + qq) = v[j];
                }
            }
            __syncthreads();
        }
        // Chi operands of the tile: B (p, point) of the apply, A (p, point)
        // of the projection (rows past npad are never read: nrt <= MRT)
        T bq[MRT][4], av[4][MRT];
#pragma unroll
        for (int rt = 0; rt < MRT; rt++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool ok = rt < nrt;
                if (CHL) {
                    const int ql = q0 - qc;
                    bq[rt][u] = ok ? chs[(rt * 16 + 4 * u + lk) * LDQ + ql + li] : T(0);
                    av[u][rt] = ok ? chs[(rt * 16 + li) * LDQ + ql + 4 * u + lk] : T(0);
                } else {
                    bq[rt][u] = ok ? ct_[(long)(rt * 16 + 4 * u + lk) * qpad + q0 + li] : T(0);
                    av[u][rt] = ok ? ct_[(long)(rt * 16 + li) * qpad + q0 + 4 * u + lk] : T(0);
                }
            }
        // the next tile's outputs, and the grid index of the one after
        acc_t ovn[NCT];
        load_out(ovn, t + 1 < ntile ? g_nxt : -1);
        const int g_nn = t + 2 < ntile ? gp[tile_q0(t + 2) + li] : -1;
        acc_t acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[ct] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int rt = 0; rt < MRT; rt++) {
            if (rt < nrt) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int j = rt * 16 + 4 * u;
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++)
                        acc[ct] = M::run(beta[(j + lk) * LDB + ct * 16 + li], bq[rt][u], acc[ct]);
                }
            }
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int cl = ct * 16 + M::drow(lk, r);
                const bool ok = g_cur >= 0 && col0 + cl < ncol;
                const T y = ov[ct][r] + acc[ct][r];
                if (ok) out[(long)(col0 + cl) * ldo + g_cur] = y;
                yw[li * LDY + cl] = ok ? y : T(0);
            }
        }
        wave_lds_order();
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) {
            T bv[NCT];
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) bv[ct] = yw[(4 * s4 + lk) * LDY + ct * 16 + li];
#pragma unroll
            for (int rt = 0; rt < MRT; rt++) {
                if (rt < nrt) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++)
                        pacc[rt][ct] = M::run(av[s4][rt], bv[ct], pacc[rt][ct]);
                }
            }
        }
        wave_lds_order();
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) ov[ct] = ovn[ct];
        g_cur = g_nxt;
        g_nxt = g_nn;
    }
    // fixed-order cross-wave sum: wave 3 -> 2 -> 1 -> 0
    for (int w = kWaves - 1; w >= 1; w--) {
        if (wave == w)
#pragma unroll
            for (int rt = 0; rt < MRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) red[rt * NCT + ct][lane] = pacc[rt][ct];
        __syncthreads();
        if (wave == w - 1)
#pragma unroll
            for (int rt = 0; rt < MRT; rt++)
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) pacc[rt][ct] += red[rt * NCT + ct][lane];
        __syncthreads();
    }
    if (wave == 0) {
#pragma unroll
        for (int rt = 0; rt < MRT; rt++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int p = rt * 16 + M::drow(lk, r);
                    const int col = col0 + ct * 16 + li;
                    if (p < np && col < ncol)
                        part_next[((long)e * max_np + p) * ncol + col] = pacc[rt][ct][r];
                }
    }
}

// SPARC_AMD_NLB=lds selects the LDS-staged nlb_project and the transposed
// nlb_apply instead of the register-streamed kernels (no brick chain).
bool nlb_lds_kernels()
{
    static const bool v = [] {
        const char *e = getenv("SPARC_AMD_NLB");
        return e && !strcmp(e, "lds");
    }();
    return v;
}

// Rows of Phi_b the fused apply + projection accumulates at once.
constexpr int kChainRTM = 7;

// out += Phi_b beta_b (beta_b gathered from alpha_s); with part_next also the
// partials Phi_b' out of the result.
template <typename T>
int launch_nlb_apply(sgpu_context *ctx, int ncol, T *out, long ldo, const T *phi, const T *phic,
                     const T *alpha_s, T *part_next)
{
    const sgpu_nonlocal &nl = ctx->nl;
    if (nlb_lds_kernels() && !part_next) {
        constexpr int NC = Mfma<T>::NCA;
        const size_t shmem = (size_t)nl.max_np * (NC * 3 / 2) * sizeof(T);
        if (shmem > 64 * 1024) {
            sgpu_set_error("nonlocal: %d projector rows in one brick exceed the LDS of the apply",
                           nl.max_np);
            return SGPU_ERR_ARG;
        }
        nlb_apply<T><<<dim3((ncol + NC - 1) / NC, nl.n_bricks), 256, shmem, ctx->stream>>>(
            out, ldo, ncol, ctx->nx, ctx->ny, ctx->nz, nl.brick_org, nl.brick_np, nl.brick_row,
            nl.row_alpha, phic, alpha_s);
        return SGPU_OK;
    }
    constexpr int NC = Mfma<T>::NCB, LDB = NC + 64 / (int)sizeof(T);
    const size_t shmem = (size_t)nl.max_np * LDB * sizeof(T);
    if (shmem > 96 * 1024) {
        sgpu_set_error("nonlocal: %d projector rows in one brick exceed the LDS of the apply",
                       nl.max_np);
        return SGPU_ERR_ARG;
    }
    const dim3 grid((ncol + NC - 1) / NC, nl.n_bricks);
    auto go = [&](auto kern) {
        if (shmem > 32 * 1024)
            SGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
        kern<<<grid, 256, shmem, ctx->stream>>>(out, ldo, ncol, ctx->nx, ctx->ny, ctx->nz,
                                                nl.brick_org, nl.brick_np, nl.brick_row,
                                                nl.row_alpha, phi, phic, alpha_s, part_next);
        return SGPU_OK;
    };
    if (part_next) {
        if (nl.max_np > 16 * kChainRTM) {
            sgpu_set_error("nonlocal: fused apply + projection limited to %d rows per brick",
                           16 * kChainRTM);
            return SGPU_ERR_ARG;
        }
        return go(nlb_apply_rs<T, NC, kChainRTM, true>);
    }
    return go(nlb_apply_rs<T, NC, 1, false>);
}

// part = Phi_b' X for every brick: nlb_project_rs with 64 (FP32) / 32 (FP64)
// columns per block, 16-byte loads where the grid and X allow them.
template <typename T>
void launch_nlb_project(sgpu_context *ctx, int ncol, const T *x, long ldx, const T *phi, T *part)
{
    const sgpu_nonlocal &nl = ctx->nl;
    const hipStream_t stream = ctx->stream;
    if (nlb_lds_kernels()) {
        constexpr int NCP = Mfma<T>::NCP;
        nlb_project<T><<<dim3((ncol + NCP - 1) / NCP, nl.n_bricks), 256, 0, stream>>>(
            x, ldx, ncol, ctx->nx, ctx->ny, ctx->nz, nl.brick_org, nl.brick_np, nl.brick_row, phi,
            part);
        return;
    }
    // 16-byte loads of four points: the points of a lane lie inside the grid
    // together when nx is a multiple of 4
    constexpr int NC = sizeof(T) == 4 ? 64 : 32, VW = 16 / (int)sizeof(T);
    const bool vec = ctx->nx % 4 == 0 && ldx % VW == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    const dim3 grid((ncol + NC - 1) / NC, nl.n_bricks);
    if (vec)
        nlb_project_rs<T, NC, true><<<grid, 256, 0, stream>>>(
            x, ldx, ncol, ctx->nx, ctx->ny, ctx->nz, nl.brick_org, nl.brick_np, nl.brick_row, phi,
            part);
    else
        nlb_project_rs<T, NC, false><<<grid, 256, 0, stream>>>(
            x, ldx, ncol, ctx->nx, ctx->ny, ctx->nz, nl.brick_org, nl.brick_np, nl.brick_row, phi,
            part);
}

// The partials buffer: the caller's workspace when one is given (it must
// hold sgpu_vnl_scratch_bytes), else the context's scratch.
int vnl_partials(sgpu_context *ctx, int dtype, int ncol, void *ws, size_t ws_bytes, void **buf)
{
    const size_t need = sgpu_vnl_scratch_bytes(ctx, dtype, ncol);
    if (!ws) return sgpu_scratch(ctx, need, buf);
    if (ws_bytes < need) {
        sgpu_set_error("nonlocal: %zu workspace bytes given, %zu needed", ws_bytes, need);
        return SGPU_ERR_ARG;
    }
    *buf = ws;
    return SGPU_OK;
}

template <typename T>
int launch_vnl_sphere(sgpu_context *ctx, int ncol, const T *x, long ldx, T *out, long ldo,
                      double scale, const T *chir, const T *chit, void *ws, size_t ws_bytes)
{
    const sgpu_nonlocal &nl = ctx->nl;
    constexpr int NCP = Mfma<T>::NCP, NC = Mfma<T>::NCA;
    const int max_np = 64;
    void *buf = nullptr;
    SGPU_CHECK(vnl_partials(ctx, sizeof(T) == 4 ? SGPU_F32 : SGPU_F64, ncol, ws, ws_bytes, &buf));
    T *part = (T *)buf;
    sgpu_tmark t;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    nls_project<T><<<dim3((ncol + NCP - 1) / NCP, nl.n_ent), 256, 0, ctx->stream>>>(
        x, ldx, ncol, nl.ent_info, nl.ent_chi, nl.gpos, chir, max_np, part);
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_project"));
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    for (int c = 0; c < nl.n_colors; c++) {
        const int e0 = nl.color_start[c], ne = nl.color_start[c + 1] - e0;
        if (ne <= 0) continue;
        nls_apply<T><<<dim3((ncol + NC - 1) / NC, ne), 256, 0, ctx->stream>>>(
            out, ldo, ncol, e0, nl.ent_info, nl.ent_chi, nl.gpos, chit, nl.atom_img_ptr,
            nl.atom_img, nl.atom_row, nl.gamma, scale * nl.dV, part, max_np);
        SGPU_HIP(hipGetLastError());
    }
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_apply"));
    return SGPU_OK;
}

template <typename T>
int launch_vnl_typed(sgpu_context *ctx, int ncol, const T *x, long ldx, T *out, long ldo,
                     double scale, const T *phi, const T *phic, void *ws, size_t ws_bytes)
{
    const sgpu_nonlocal &nl = ctx->nl;
    if (nl.n_bricks == 0 || nl.n_alpha == 0 || ncol <= 0) return SGPU_OK;
    if (nl.n_bricks > 65535) {
        sgpu_set_error("nonlocal: %d bricks exceed the launch limit", nl.n_bricks);
        return SGPU_ERR_ARG;
    }
    void *buf = nullptr;
    const size_t part_elems = (size_t)nl.n_rows * ncol;
    SGPU_CHECK(vnl_partials(ctx, sizeof(T) == 4 ? SGPU_F32 : SGPU_F64, ncol, ws, ws_bytes, &buf));
    T *part = (T *)buf;
    T *alpha_s = part + part_elems;
    sgpu_tmark t;
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    launch_nlb_project<T>(ctx, ncol, x, ldx, phi, part);
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_project"));
    const long total = (long)nl.n_alpha * ncol;
    nlb_reduce<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
        part, ncol, nl.n_alpha, nl.arow_ptr, nl.arow_src, nl.gamma, scale * nl.dV, alpha_s);
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_CHECK(launch_nlb_apply<T>(ctx, ncol, out, ldo, phi, phic, alpha_s, nullptr));
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_apply"));
    return SGPU_OK;
}

int chain_buffer(sgpu_context *ctx, size_t bytes, void **p)
{
    sgpu_nonlocal &nl = ctx->nl;
    if (bytes > nl.chain_bytes) {
        SGPU_HIP(hipStreamSynchronize(ctx->stream));
        if (nl.chain) SGPU_HIP(hipFree(nl.chain));
        nl.chain = nullptr;
        nl.chain_bytes = 0;
        SGPU_HIP(hipMalloc(&nl.chain, bytes));
        nl.chain_bytes = bytes;
    }
    *p = nl.chain;
    return SGPU_OK;
}

template <typename T>
int launch_vnl_chain_sphere(sgpu_context *ctx, int ncol, const T *x, long ldx, T *out, long ldo,
                            double scale, const T *chir, const T *chit, bool have_part,
                            bool want_next, bool *projected)
{
    sgpu_nonlocal &nl = ctx->nl;
    *projected = false;
    if (nl.n_ent == 0 || ncol <= 0) return SGPU_OK;
    constexpr int NCP = Mfma<T>::NCP, NC = Mfma<T>::NCA;
    const int max_np = 64;
    const size_t part_elems = (size_t)nl.n_ent * max_np * ncol;
    const size_t need = 2 * part_elems * sizeof(T) + 256;
    if (have_part && need > nl.chain_bytes) {
        sgpu_set_error("nonlocal chain: partials of a different shape");
        return SGPU_ERR_STATE;
    }
    void *buf = nullptr;
    SGPU_CHECK(chain_buffer(ctx, need, &buf));
    if (!have_part) nl.chain_cur = 0;
    T *part = (T *)buf + (size_t)nl.chain_cur * part_elems;
    T *part_next = (T *)buf + (size_t)(nl.chain_cur ^ 1) * part_elems;
    sgpu_tmark t;
    if (!have_part) {
        SGPU_CHECK(sgpu_tbegin(ctx, &t));
        nls_project<T><<<dim3((ncol + NCP - 1) / NCP, nl.n_ent), 256, 0, ctx->stream>>>(
            x, ldx, ncol, nl.ent_info, nl.ent_chi, nl.gpos, chir, max_np, part);
        SGPU_HIP(hipGetLastError());
        SGPU_CHECK(sgpu_tend(ctx, &t, "nl_project"));
    }
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    if (want_next && nl.n_colors == 1) {
        const dim3 grid((ncol + NC - 1) / NC, nl.n_ent);
        auto go = [&](auto kern) {
            kern<<<grid, 256, 0, ctx->stream>>>(out, ldo, ncol, 0, nl.ent_info, nl.ent_chi,
                                                nl.gpos, chit, nl.atom_img_ptr, nl.atom_img,
                                                nl.atom_row, nl.gamma, scale * nl.dV, part,
                                                max_np, part_next);
        };
        // Chi staged in LDS for FP32 with up to 32 projector rows (53 KB of
        // LDS per block, three blocks per CU); FP64 reads it from L2
        constexpr bool F32 = sizeof(T) == 4;
        if (nl.max_npad <= 32)
            go(nls_apply_proj<T, 2, F32>);
        else
            go(nls_apply_proj<T, 4, false>);
        SGPU_HIP(hipGetLastError());
        nl.chain_cur ^= 1;
        *projected = true;
    } else {
        for (int c = 0; c < nl.n_colors; c++) {
            const int e0 = nl.color_start[c], ne = nl.color_start[c + 1] - e0;
            if (ne <= 0) continue;
            nls_apply<T><<<dim3((ncol + NC - 1) / NC, ne), 256, 0, ctx->stream>>>(
                out, ldo, ncol, e0, nl.ent_info, nl.ent_chi, nl.gpos, chit, nl.atom_img_ptr,
                nl.atom_img, nl.atom_row, nl.gamma, scale * nl.dV, part, max_np);
            SGPU_HIP(hipGetLastError());
        }
    }
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_apply"));
    return SGPU_OK;
}

// One application of a Chebyshev chain with the brick layout: the partials of
// x are those the previous application left (have_part) or a projection of x;
// with want_next the apply also forms the partials of out (nlb_apply_rs with
// PROJ), so the next application skips its projection and the re-read of its
// input.  Partials ping-pong between two buffers in nl.chain, alpha_s after
// them.
template <typename T>
int launch_vnl_chain_brick(sgpu_context *ctx, int ncol, const T *x, long ldx, T *out, long ldo,
                           double scale, const T *phi, const T *phic, bool have_part,
                           bool want_next, bool *projected)
{
    sgpu_nonlocal &nl = ctx->nl;
    *projected = false;
    if (nl.n_bricks == 0 || nl.n_alpha == 0 || ncol <= 0) return SGPU_OK;
    if (nl.n_bricks > 65535) {
        sgpu_set_error("nonlocal: %d bricks exceed the launch limit", nl.n_bricks);
        return SGPU_ERR_ARG;
    }
    const bool fuse = want_next && ctx->nl_chain_brick && !nlb_lds_kernels() &&
                      nl.max_np <= 16 * kChainRTM;
    if (!fuse && !have_part)
        return launch_vnl_typed<T>(ctx, ncol, x, ldx, out, ldo, scale, phi, phic, nullptr, 0);
    const size_t part_elems = (size_t)nl.n_rows * ncol;
    const size_t need = (2 * part_elems + (size_t)nl.n_alpha * ncol) * sizeof(T) + 256;
    if (have_part && need > nl.chain_bytes) {
        sgpu_set_error("nonlocal chain: partials of a different shape");
        return SGPU_ERR_STATE;
    }
    void *buf = nullptr;
    SGPU_CHECK(chain_buffer(ctx, need, &buf));
    if (!have_part) nl.chain_cur = 0;
    T *part = (T *)buf + (size_t)nl.chain_cur * part_elems;
    T *part_next = (T *)buf + (size_t)(nl.chain_cur ^ 1) * part_elems;
    T *alpha_s = (T *)buf + 2 * part_elems;
    sgpu_tmark t;
    if (!have_part) {
        SGPU_CHECK(sgpu_tbegin(ctx, &t));
        launch_nlb_project<T>(ctx, ncol, x, ldx, phi, part);
        SGPU_HIP(hipGetLastError());
        SGPU_CHECK(sgpu_tend(ctx, &t, "nl_project"));
    }
    const long total = (long)nl.n_alpha * ncol;
    nlb_reduce<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
        part, ncol, nl.n_alpha, nl.arow_ptr, nl.arow_src, nl.gamma, scale * nl.dV, alpha_s);
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tbegin(ctx, &t));
    SGPU_CHECK(launch_nlb_apply<T>(ctx, ncol, out, ldo, phi, phic, alpha_s,
                                   fuse ? part_next : nullptr));
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tend(ctx, &t, "nl_apply"));
    if (fuse) {
        nl.chain_cur ^= 1;
        *projected = true;
    }
    return SGPU_OK;
}

template <typename T>
int upload(T **dst, const T *src, size_t n, hipStream_t s)
{
    SGPU_HIP(hipMalloc((void **)dst, (n ? n : 1) * sizeof(T)));
    if (n) SGPU_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return SGPU_OK;
}

// Sphere layout (see nls_project / nls_apply): entries greedily coloured so
// that no two entries of a colour share a grid point, each entry's point
// list padded to a multiple of 64 (-1 = padding) and its Chi stored twice,
// [q][npad] for the projection and [p][qpad] for the apply.
int build_sphere_layout(sgpu_context *ctx, int n_influence, const int *ndc, const int *nproj,
                        const long *pos_off, const int *grid_pos, const long *chi_off,
                        const double *chi, const int *alpha_off, int n_alpha,
                        const double *gamma, double dV)
{
    // atoms = distinct alpha offsets (images of one atom share them)
    std::vector<int> offs(alpha_off, alpha_off + n_influence);
    std::sort(offs.begin(), offs.end());
    offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
    const int n_atoms = (int)offs.size();
    auto atom_of = [&](int a) {
        return (int)(std::lower_bound(offs.begin(), offs.end(), alpha_off[a]) - offs.begin());
    };
    // greedy colouring over shared grid points
    std::vector<uint64_t> used((size_t)ctx->nd, 0);
    std::vector<int> color(n_influence);
    int n_colors = 0;
    for (int a = 0; a < n_influence; a++) {
        uint64_t m = 0;
        for (int i = 0; i < ndc[a]; i++) m |= used[grid_pos[pos_off[a] + i]];
        if (m == ~0ull) {
            sgpu_set_error("sgpu_set_nonlocal: more than 64 overlapping projector spheres");
            return SGPU_ERR_ARG;
        }
        int c = 0;
        while (m & (1ull << c)) c++;
        color[a] = c;
        n_colors = std::max(n_colors, c + 1);
        for (int i = 0; i < ndc[a]; i++) used[grid_pos[pos_off[a] + i]] |= 1ull << c;
    }
    std::vector<int> order(n_influence);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return color[x] < color[y]; });
    std::vector<int> color_start(n_colors + 1, 0);
    for (int a = 0; a < n_influence; a++) color_start[color[a] + 1]++;
    for (int c = 0; c < n_colors; c++) color_start[c + 1] += color_start[c];
    // per entry layout
    std::vector<int> info(6 * (size_t)n_influence), gp;
    std::vector<long> eoff(n_influence);
    std::vector<double> chir, chit;
    std::vector<std::vector<int>> imgs(n_atoms);
    long coff = 0;
    int max_qpad = 0, max_npad = 0;
    for (int e = 0; e < n_influence; e++) {
        const int a = order[e];
        const int n = ndc[a], np = nproj[a];
        const int qpad = std::max(64, (n + 63) / 64 * 64), npad = std::max(16, (np + 15) / 16 * 16);
        const int atom = atom_of(a);
        info[6 * e] = qpad;
        info[6 * e + 1] = (int)gp.size();
        info[6 * e + 2] = np;
        info[6 * e + 3] = npad;
        info[6 * e + 4] = atom;
        info[6 * e + 5] = 0;
        eoff[e] = coff;
        imgs[atom].push_back(e);
        max_qpad = std::max(max_qpad, qpad);
        max_npad = std::max(max_npad, npad);
        for (int q = 0; q < qpad; q++) gp.push_back(q < n ? grid_pos[pos_off[a] + q] : -1);
        chir.resize(coff + (size_t)qpad * npad, 0.0);
        chit.resize(coff + (size_t)qpad * npad, 0.0);
        for (int p = 0; p < np; p++)
            for (int q = 0; q < n; q++) {
                const double v = chi[chi_off[a] + (long)p * n + q];
                chir[coff + (long)q * npad + p] = v;
                chit[coff + (long)p * qpad + q] = v;
            }
        coff += (long)qpad * npad;
    }
    std::vector<int> img_ptr(n_atoms + 1, 0), img, arow(n_atoms);
    for (int t = 0; t < n_atoms; t++) {
        img_ptr[t + 1] = img_ptr[t] + (int)imgs[t].size();
        img.insert(img.end(), imgs[t].begin(), imgs[t].end());
        arow[t] = offs[t];
    }
    sgpu_nonlocal &nl = ctx->nl;
    hipStream_t s = ctx->stream;
    nl.mode = 1;
    nl.stored = (long)chir.size();
    nl.n_infl = n_influence;
    nl.n_alpha = n_alpha;
    nl.n_ent = n_influence;
    nl.n_atoms = n_atoms;
    nl.n_colors = n_colors;
    nl.max_qpad = max_qpad;
    nl.max_npad = max_npad;
    nl.color_start = color_start;
    nl.dV = dV;
    SGPU_CHECK(upload(&nl.ent_info, info.data(), info.size(), s));
    SGPU_CHECK(upload(&nl.ent_chi, eoff.data(), eoff.size(), s));
    SGPU_CHECK(upload(&nl.gpos, gp.data(), gp.size(), s));
    SGPU_CHECK(upload(&nl.chir64, chir.data(), chir.size(), s));
    SGPU_CHECK(upload(&nl.chit64, chit.data(), chit.size(), s));
    SGPU_HIP(hipMalloc((void **)&nl.chir32, std::max<size_t>(chir.size(), 1) * sizeof(float)));
    SGPU_HIP(hipMalloc((void **)&nl.chit32, std::max<size_t>(chit.size(), 1) * sizeof(float)));
    SGPU_CHECK(sgpu_convert(ctx, nl.chir64, SGPU_F64, nl.chir32, SGPU_F32, (long)chir.size()));
    SGPU_CHECK(sgpu_convert(ctx, nl.chit64, SGPU_F64, nl.chit32, SGPU_F32, (long)chit.size()));
    SGPU_CHECK(upload(&nl.atom_img_ptr, img_ptr.data(), img_ptr.size(), s));
    SGPU_CHECK(upload(&nl.atom_img, img.data(), img.size(), s));
    SGPU_CHECK(upload(&nl.atom_row, arow.data(), arow.size(), s));
    SGPU_CHECK(upload(&nl.gamma, gamma, (size_t)n_alpha, s));
    SGPU_HIP(hipStreamSynchronize(s));
    return SGPU_OK;
}

}  // namespace

void sgpu_free_nonlocal(sgpu_nonlocal &nl)
{
    void *ptrs[] = {nl.brick_org, nl.brick_np, nl.brick_row,  nl.row_alpha, nl.phi64,
                    nl.phi32,     nl.phic64,   nl.phic32,     nl.arow_ptr,  nl.arow_src,
                    nl.gamma,     nl.ent_info, nl.ent_chi,    nl.gpos,      nl.chir64,
                    nl.chir32,    nl.chit64,   nl.chit32,     nl.atom_img_ptr, nl.atom_img,
                    nl.atom_row,  nl.chain};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    nl = sgpu_nonlocal();
}

// ---- nonlocal force integrals -------------------------------------------
namespace {

// First derivative along one axis with the order-2F central stencil
// (Gradient_vectors_dir, src/gradVecRoutines.c), periodic wrap-around or
// zero outside a Dirichlet boundary; one thread per grid point and column.
struct D1W {
    double w[SGPU_MAX_FDN + 1];
};

template <typename T>
__global__ void grad_axis(const T *__restrict__ x, long ldx, T *__restrict__ y, long ldy, int nx,
                          int ny, int nz, int axis, int bc, int fdn, D1W w)
{
    const long nd = (long)nx * ny * nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const int col = blockIdx.y;
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((long)nx * ny));
    const int c = axis == 0 ? ix : axis == 1 ? iy : iz;
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const long stride = axis == 0 ? 1L : axis == 1 ? (long)nx : (long)nx * ny;
    const T *xc = x + (long)col * ldx + (i - (long)c * stride);
    double s = 0.0;
    for (int p = 1; p <= fdn; p++) {
        int a = c + p, b = c - p;
        double va = 0.0, vb = 0.0;
        if (a >= n) a = bc ? -1 : a - n;
        if (b < 0) b = bc ? -1 : b + n;
        if (a >= 0) va = (double)xc[(long)a * stride];
        if (b >= 0) vb = (double)xc[(long)b * stride];
        s += w.w[p] * (va - vb);
    }
    y[(long)col * ldy + i] = (T)s;
}

// alpha[atom_row + p][col] = scale * sum over the atom's images of part
// (sphere layout; images in the fixed order of atom_img)
template <typename T>
__global__ void nls_sum_images(const T *__restrict__ part, int ncol, int n_atoms, int max_np,
                               const int *__restrict__ atom_img_ptr,
                               const int *__restrict__ atom_img,
                               const int *__restrict__ atom_row,
                               const int *__restrict__ ent_info, double scale,
                               T *__restrict__ alpha)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_atoms * max_np * ncol) return;
    const int col = (int)(t % ncol);
    const long ap = t / ncol;
    const int p = (int)(ap % max_np), a = (int)(ap / max_np);
    const int i0 = atom_img_ptr[a], i1 = atom_img_ptr[a + 1];
    if (p >= ent_info[6 * atom_img[i0] + 2]) return;
    double v = 0.0;
    for (int i = i0; i < i1; i++) v += (double)part[((long)atom_img[i] * max_np + p) * ncol + col];
    alpha[(long)(atom_row[a] + p) * ncol + col] = (T)(scale * v);
}

// alpha (n_alpha x ncol, device, type T) = scale * sum over images Chi' x
template <typename T>
int project_alpha(sgpu_context *ctx, int ncol, const T *x, long ldx, double scale, T *alpha)
{
    const sgpu_nonlocal &nl = ctx->nl;
    constexpr int NCP = Mfma<T>::NCP;
    void *buf = nullptr;
    if (nl.mode == 1) {
        const int max_np = 64;
        SGPU_CHECK(sgpu_scratch(ctx, (size_t)nl.n_ent * max_np * ncol * sizeof(T) + 256, &buf));
        T *part = (T *)buf;
        nls_project<T><<<dim3((ncol + NCP - 1) / NCP, nl.n_ent), 256, 0, ctx->stream>>>(
            x, ldx, ncol, nl.ent_info, nl.ent_chi, nl.gpos,
            sizeof(T) == 8 ? (const T *)nl.chir64 : (const T *)nl.chir32, max_np, part);
        SGPU_HIP(hipGetLastError());
        const long total = (long)nl.n_atoms * max_np * ncol;
        nls_sum_images<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
            part, ncol, nl.n_atoms, max_np, nl.atom_img_ptr, nl.atom_img, nl.atom_row, nl.ent_info,
            scale, alpha);
        SGPU_HIP(hipGetLastError());
    } else {
        if (nl.n_bricks > 65535) {
            sgpu_set_error("nonlocal: %d bricks exceed the launch limit", nl.n_bricks);
            return SGPU_ERR_ARG;
        }
        SGPU_CHECK(sgpu_scratch(ctx, (size_t)nl.n_rows * ncol * sizeof(T) + 256, &buf));
        T *part = (T *)buf;
        launch_nlb_project<T>(ctx, ncol, x, ldx,
                              sizeof(T) == 8 ? (const T *)nl.phi64 : (const T *)nl.phi32, part);
        SGPU_HIP(hipGetLastError());
        const long total = (long)nl.n_alpha * ncol;
        nlb_reduce<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
            part, ncol, nl.n_alpha, nl.arow_ptr, nl.arow_src, nullptr, scale, alpha);
        SGPU_HIP(hipGetLastError());
    }
    return SGPU_OK;
}

template <typename T>
int nonlocal_integrals_typed(sgpu_context *ctx, const sgpu_block *psi, double *out)
{
    const sgpu_nonlocal &nl = ctx->nl;
    const int ncol = psi->ncol;
    const long na = (long)nl.n_alpha * ncol;
    hipStream_t s = ctx->stream;
    T *grad = nullptr, *alpha = nullptr;
    std::vector<T> host((size_t)na);
    D1W w{};
    int st = SGPU_OK;
    auto fail = [&](int code) { st = code; };
    if (hipMalloc((void **)&grad, (size_t)ncol * psi->ld * sizeof(T)) != hipSuccess ||
        hipMalloc((void **)&alpha, (size_t)std::max(na, 1L) * sizeof(T)) != hipSuccess) {
        sgpu_set_error("sgpu_nonlocal_integrals: device allocation failed");
        fail(SGPU_ERR_HIP);
    }
    const double h[3] = {ctx->hx, ctx->hy, ctx->hz};
    const int bc[3] = {ctx->bcx, ctx->bcy, ctx->bcz};
    for (int k = 0; k < 4 && st == SGPU_OK; k++) {
        const T *src = (const T *)psi->ptr;
        if (k > 0) {
            // w_p = (-1)^(p+1) F!^2 / ((F-p)! (F+p)!) / (p h)
            //      (src/initialization.c:2113-2132)
            for (int p = 1; p <= ctx->fdn; p++) {
                double fr = 1.0;
                for (int q = 1; q <= p; q++) fr *= (double)(ctx->fdn - p + q) / (double)(ctx->fdn + q);
                w.w[p] = ((p % 2) ? 1.0 : -1.0) * fr / p / h[k - 1];
            }
            const dim3 grid((unsigned)((ctx->nd + 255) / 256), (unsigned)ncol);
            grad_axis<T><<<grid, 256, 0, s>>>((const T *)psi->ptr, psi->ld, grad, psi->ld, ctx->nx,
                                              ctx->ny, ctx->nz, k - 1, bc[k - 1], ctx->fdn, w);
            if (hipGetLastError() != hipSuccess) {
                sgpu_set_error("sgpu_nonlocal_integrals: gradient launch failed");
                fail(SGPU_ERR_HIP);
                break;
            }
            src = grad;
        }
        const int pst = project_alpha<T>(ctx, ncol, src, psi->ld, k == 0 ? nl.dV : 1.0, alpha);
        if (pst < 0) {
            fail(pst);
            break;
        }
        if (hipMemcpyAsync(host.data(), alpha, (size_t)na * sizeof(T), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            sgpu_set_error("sgpu_nonlocal_integrals: copy failed");
            fail(SGPU_ERR_HIP);
            break;
        }
        for (long i = 0; i < na; i++) out[(long)k * na + i] = (double)host[i];
    }
    (void)hipStreamSynchronize(s);
    if (grad) (void)hipFree(grad);
    if (alpha) (void)hipFree(alpha);
    return st;
}

}  // namespace

// <Chi, psi> dV and <Chi, d psi / dx_d> for every orbital of psi (images
// summed into the atom's rows): Compute_Integral_psi_Chi and
// Compute_Integral_Chi_Dpsi (src/forces.c:849-924) with Gradient_vectors_dir;
// out[4][n_alpha][ncol] on the host.
extern "C" int sgpu_nonlocal_integrals(sgpu_context *ctx, const sgpu_block *psi, double *out)
{
    if (!ctx || !psi || !out || psi->ctx != ctx || !psi->ptr) return SGPU_ERR_ARG;
    if (ctx->nl.n_alpha == 0 || psi->ncol <= 0) return SGPU_OK;
    if (ctx->fdn < 1 || ctx->fdn > SGPU_MAX_FDN) return SGPU_ERR_ARG;
    if (psi->dtype == SGPU_F64) return nonlocal_integrals_typed<double>(ctx, psi, out);
    return nonlocal_integrals_typed<float>(ctx, psi, out);
}

extern "C" int sgpu_set_nonlocal(sgpu_context *ctx, int n_influence, const int *ndc,
                                 const int *nproj, const long *pos_off, const int *grid_pos,
                                 const long *chi_off, const double *chi, const int *alpha_off,
                                 int n_alpha, const double *gamma, double dV)
{
    if (!ctx || n_influence < 0 || n_alpha < 0) {
        sgpu_set_error("sgpu_set_nonlocal: bad arguments");
        return SGPU_ERR_ARG;
    }
    SGPU_HIP(hipSetDevice(ctx->device));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    sgpu_free_nonlocal(ctx->nl);
    if (n_influence == 0 || n_alpha == 0) return SGPU_OK;
    if (!ndc || !nproj || !pos_off || !grid_pos || !chi_off || !chi || !alpha_off || !gamma) {
        sgpu_set_error("sgpu_set_nonlocal: null array");
        return SGPU_ERR_ARG;
    }
    const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
    const int nbx = (nx + kBX - 1) / kBX, nby = (ny + kBY - 1) / kBY, nbz = (nz + kBZ - 1) / kBZ;
    // (brick, entry) pairs, validated
    std::vector<std::pair<int, int>> pairs;
    for (int a = 0; a < n_influence; a++) {
        if (ndc[a] < 0 || nproj[a] < 0 || alpha_off[a] < 0 || alpha_off[a] + nproj[a] > n_alpha ||
            pos_off[a] < 0 || chi_off[a] < 0) {
            sgpu_set_error("sgpu_set_nonlocal: entry %d out of range", a);
            return SGPU_ERR_ARG;
        }
        std::vector<int> bl;
        for (int i = 0; i < ndc[a]; i++) {
            const int q = grid_pos[pos_off[a] + i];
            if (q < 0 || q >= ctx->nd) {
                sgpu_set_error("sgpu_set_nonlocal: grid_pos %d of entry %d outside the grid", q, a);
                return SGPU_ERR_ARG;
            }
            const int ix = q % nx, iy = (q / nx) % ny, iz = q / (nx * ny);
            bl.push_back((iz / kBZ * nby + iy / kBY) * nbx + ix / kBX);
        }
        std::sort(bl.begin(), bl.end());
        bl.erase(std::unique(bl.begin(), bl.end()), bl.end());
        for (int b : bl) pairs.push_back({b, a});
    }
    std::sort(pairs.begin(), pairs.end());
    {
        // layout choice: the sphere layout when the spheres fill less than a
        // quarter of the bricks they touch (SPARC_AMD_NL_MODE=brick|sphere
        // overrides); it needs at most 64 projectors per entry
        long pts = 0;
        int np_max = 0;
        for (int a = 0; a < n_influence; a++) {
            pts += ndc[a];
            np_max = std::max(np_max, nproj[a]);
        }
        const double fill = pairs.empty() ? 1.0 : (double)pts / ((double)pairs.size() * kBP);
        int mode = fill < 0.25 ? 1 : 0;
        if (const char *env = getenv("SPARC_AMD_NL_MODE")) {
            if (!strcmp(env, "brick")) mode = 0;
            if (!strcmp(env, "sphere")) mode = 1;
        }
        if (np_max > 64) mode = 0;
        if (mode == 1)
            return build_sphere_layout(ctx, n_influence, ndc, nproj, pos_off, grid_pos, chi_off, chi,
                                       alpha_off, n_alpha, gamma, dV);
    }
    // bricks, their column layout and row offsets
    std::vector<int> bid, borg, bnp;
    std::vector<long> brow;
    std::vector<int> row_alpha;
    // column offset of entry a inside brick k: (k, a) -> col, looked up by pair index
    std::vector<int> pair_col(pairs.size());
    std::vector<int> pair_brick(pairs.size());
    long nrows = 0;
    int max_np = 0;
    for (size_t s = 0; s < pairs.size();) {
        size_t e = s;
        int cols = 0;
        while (e < pairs.size() && pairs[e].first == pairs[s].first) {
            pair_col[e] = cols;
            pair_brick[e] = (int)bid.size();
            cols += nproj[pairs[e].second];
            e++;
        }
        const int padded = (cols + 3) / 4 * 4;
        const int b = pairs[s].first;
        bid.push_back(b);
        borg.push_back((b % nbx) * kBX);
        borg.push_back(((b / nbx) % nby) * kBY);
        borg.push_back((b / (nbx * nby)) * kBZ);
        bnp.push_back(padded);
        brow.push_back(nrows);
        for (size_t k = s; k < e; k++)
            for (int p = 0; p < nproj[pairs[k].second]; p++)
                row_alpha.push_back(alpha_off[pairs[k].second] + p);
        for (int p = cols; p < padded; p++) row_alpha.push_back(-1);
        nrows += padded;
        max_np = std::max(max_np, padded);
        s = e;
    }
    const int nb = (int)bid.size();
    if ((size_t)max_np * 24 * sizeof(double) > 60 * 1024) {
        sgpu_set_error("sgpu_set_nonlocal: %d projector columns in one brick exceed LDS staging", max_np);
        return SGPU_ERR_ARG;
    }
    // dense Phi per brick
    std::vector<double> phi((size_t)nrows * kBP, 0.0);
    {
        // pair index lookup per entry: entries' bricks in ascending order
        std::vector<std::vector<std::pair<int, int>>> ent_pairs(n_influence);
        for (size_t k = 0; k < pairs.size(); k++)
            ent_pairs[pairs[k].second].push_back({pairs[k].first, (int)k});
        for (int a = 0; a < n_influence; a++) {
            auto &ep = ent_pairs[a];
            for (int i = 0; i < ndc[a]; i++) {
                const int q = grid_pos[pos_off[a] + i];
                const int ix = q % nx, iy = (q / nx) % ny, iz = q / (nx * ny);
                const int b = (iz / kBZ * nby + iy / kBY) * nbx + ix / kBX;
                auto it = std::lower_bound(ep.begin(), ep.end(), std::make_pair(b, -1));
                const int k = it->second;
                const int kb = pair_brick[k];
                const int ql = (ix % kBX) + kBX * ((iy % kBY) + kBY * (iz % kBZ));
                const long base = (brow[kb] * kBP) + (long)ql * bnp[kb] + pair_col[k];
                for (int p = 0; p < nproj[a]; p++)
                    phi[base + p] += chi[chi_off[a] + (long)p * ndc[a] + i];
            }
        }
    }
    // alpha row -> brick rows
    std::vector<int> arow_ptr(n_alpha + 1, 0), arow_src;
    for (long r = 0; r < nrows; r++)
        if (row_alpha[r] >= 0) arow_ptr[row_alpha[r] + 1]++;
    for (int r = 0; r < n_alpha; r++) arow_ptr[r + 1] += arow_ptr[r];
    arow_src.resize(std::max(1, arow_ptr[n_alpha]));
    {
        std::vector<int> fill(arow_ptr.begin(), arow_ptr.end() - 1);
        for (long r = 0; r < nrows; r++)
            if (row_alpha[r] >= 0) arow_src[fill[row_alpha[r]]++] = (int)r;
    }
    sgpu_nonlocal &nl = ctx->nl;
    hipStream_t s = ctx->stream;
    nl.n_infl = n_influence;
    nl.n_alpha = n_alpha;
    nl.n_bricks = nb;
    nl.n_rows = nrows;
    nl.stored = nrows * kBP;
    nl.max_np = max_np;
    nl.dV = dV;
    SGPU_CHECK(upload(&nl.brick_org, borg.data(), borg.size(), s));
    SGPU_CHECK(upload(&nl.brick_np, bnp.data(), bnp.size(), s));
    SGPU_CHECK(upload(&nl.brick_row, brow.data(), brow.size(), s));
    SGPU_CHECK(upload(&nl.row_alpha, row_alpha.data(), row_alpha.size(), s));
    SGPU_CHECK(upload(&nl.phi64, phi.data(), phi.size(), s));
    SGPU_HIP(hipMalloc((void **)&nl.phi32, std::max<size_t>(phi.size(), 1) * sizeof(float)));
    SGPU_CHECK(sgpu_convert(ctx, nl.phi64, SGPU_F64, nl.phi32, SGPU_F32, (long)phi.size()));
    {
        // column-major copy per brick (points contiguous) for the apply
        std::vector<double> phic(phi.size());
        for (int k = 0; k < nb; k++) {
            const double *src = phi.data() + brow[k] * kBP;
            double *dst = phic.data() + brow[k] * kBP;
            const int npk = bnp[k];
            for (int q = 0; q < kBP; q++)
                for (int j = 0; j < npk; j++) dst[(long)j * kBP + q] = src[(long)q * npk + j];
        }
        SGPU_CHECK(upload(&nl.phic64, phic.data(), phic.size(), s));
        SGPU_HIP(hipStreamSynchronize(s));
    }
    SGPU_HIP(hipMalloc((void **)&nl.phic32, std::max<size_t>(phi.size(), 1) * sizeof(float)));
    SGPU_CHECK(sgpu_convert(ctx, nl.phic64, SGPU_F64, nl.phic32, SGPU_F32, (long)phi.size()));
    SGPU_CHECK(upload(&nl.arow_ptr, arow_ptr.data(), arow_ptr.size(), s));
    SGPU_CHECK(upload(&nl.arow_src, arow_src.data(), arow_src.size(), s));
    SGPU_CHECK(upload(&nl.gamma, gamma, (size_t)n_alpha, s));
    SGPU_HIP(hipStreamSynchronize(s));
    return SGPU_OK;
}

size_t sgpu_vnl_scratch_bytes(const sgpu_context *ctx, int dtype, int ncol)
{
    const sgpu_nonlocal &nl = ctx->nl;
    const size_t s = sgpu_dtype_size(dtype);
    if (nl.n_alpha == 0 || ncol <= 0) return 0;
    if (nl.mode == 1) return (size_t)nl.n_ent * 64 * ncol * s + 256;  // partials, 64 rows per entry
    return ((size_t)nl.n_rows * ncol + (size_t)nl.n_alpha * ncol) * s + 256;
}

extern "C" int sgpu_nonlocal_layout(const sgpu_context *ctx, int *mode, long *stored)
{
    if (!ctx || !mode || !stored) {
        sgpu_set_error("sgpu_nonlocal_layout: bad arguments");
        return SGPU_ERR_ARG;
    }
    *mode = ctx->nl.mode;
    *stored = ctx->nl.n_alpha > 0 ? ctx->nl.stored : 0;
    return SGPU_OK;
}

int sgpu_launch_vnl(sgpu_context *ctx, int dtype, int ncol, const void *x, long ldx, void *out,
                    long ldo, double scale, void *ws, size_t ws_bytes)
{
    if (ctx->nl.mode == 1) {
        if (ctx->nl.n_ent == 0 || ncol <= 0) return SGPU_OK;
        if (dtype == SGPU_F64)
            return launch_vnl_sphere<double>(ctx, ncol, (const double *)x, ldx, (double *)out, ldo,
                                             scale, ctx->nl.chir64, ctx->nl.chit64, ws, ws_bytes);
        return launch_vnl_sphere<float>(ctx, ncol, (const float *)x, ldx, (float *)out, ldo, scale,
                                        ctx->nl.chir32, ctx->nl.chit32, ws, ws_bytes);
    }
    if (dtype == SGPU_F64)
        return launch_vnl_typed<double>(ctx, ncol, (const double *)x, ldx, (double *)out, ldo,
                                        scale, ctx->nl.phi64, ctx->nl.phic64, ws, ws_bytes);
    return launch_vnl_typed<float>(ctx, ncol, (const float *)x, ldx, (float *)out, ldo, scale,
                                   ctx->nl.phi32, ctx->nl.phic32, ws, ws_bytes);
}

int sgpu_launch_vnl_chain(sgpu_context *ctx, int dtype, int ncol, const void *x, long ldx,
                          void *out, long ldo, double scale, bool have_part, bool want_next,
                          bool *projected)
{
    sgpu_nonlocal &nl = ctx->nl;
    *projected = false;
    if (have_part && nl.chain_x != x) {
        sgpu_set_error("nonlocal chain: no partials held for this block");
        return SGPU_ERR_STATE;
    }
    nl.chain_x = nullptr;
    int st;
    if (nl.mode != 1) {
        st = dtype == SGPU_F64
                 ? launch_vnl_chain_brick<double>(ctx, ncol, (const double *)x, ldx, (double *)out,
                                                  ldo, scale, nl.phi64, nl.phic64, have_part,
                                                  want_next, projected)
                 : launch_vnl_chain_brick<float>(ctx, ncol, (const float *)x, ldx, (float *)out, ldo,
                                                 scale, nl.phi32, nl.phic32, have_part, want_next,
                                                 projected);
        if (st == SGPU_OK && *projected) nl.chain_x = out;
        return st;
    }
    st =
        dtype == SGPU_F64
            ? launch_vnl_chain_sphere<double>(ctx, ncol, (const double *)x, ldx, (double *)out, ldo,
                                              scale, nl.chir64, nl.chit64, have_part, want_next,
                                              projected)
            : launch_vnl_chain_sphere<float>(ctx, ncol, (const float *)x, ldx, (float *)out, ldo,
                                             scale, nl.chir32, nl.chit32, have_part, want_next,
                                             projected);
    if (st == SGPU_OK && *projected) nl.chain_x = out;
    return st;
}
