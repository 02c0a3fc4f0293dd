// Order-2R finite-difference stencil kernels with fused diagonal, shift and
// three-term (Chebyshev) update:
//
//   out = s1 * ((a Lap + bdiag diag(v) + c) x) + s2 * z          (per column)
//
// This is the arithmetic of Lap_plus_diag_vec_mult_orth +
// stencil_3axis_thread_radius6 (SPARC src/lapVecRoutines.c:185-227,306-611)
// with the scaling/recurrence of ChebyshevFiltering (src/eigenSolver.c:763-794)
// folded into the epilogue, so one pass reads x, v, z once and writes out once.
// Periodic halos wrap indices (the reference copies the opposite face,
// :536-577); Dirichlet halos read zero.
//
// Two kernels:
//  * stencil_tiled<T, TX, TY, RY>: radius 6.  A TX x TY tile of the x-y plane
//    (plus a 6-deep halo) is staged in LDS per z-plane while the block marches
//    along z; each thread keeps a 13-deep register queue of its own points'
//    z-neighbours, so every x value is read from HBM about once.
//  * stencil_rows<TX, TY, RY>: the FP32 whole-row form (TX == nx) of the
//    tiled kernel, with no x halo reads and the y halo read kR planes early.
//  * stencil_direct<T, R>: one thread per point, all 6R+1 loads from global;
//    used for other FD orders and as the independent cross-check.
#include "sgpu_internal.h"

namespace {

template <typename T>
struct StencilParams {
    const T *x;
    T *out;
    const T *z;
    const T *v;
    long ldx, ldo, ldz;
    int nx, ny, nz;
    int bcx, bcy, bcz;
    T wx[SGPU_MAX_FDN + 1], wy[SGPU_MAX_FDN + 1], wz[SGPU_MAX_FDN + 1];
    T w0;     // a (D2x0 + D2y0 + D2z0) + c
    T bdiag;  // coefficient of v (0 -> no diagonal term)
    T s1, s2;
};

// Index along one axis with the halo convention; returns -1 for a Dirichlet
// point outside the domain.
__device__ __forceinline__ int axis_index(int g, int n, int bc)
{
    if (g >= 0 && g < n) return g;
    if (bc) return -1;
    int w = g % n;
    return w < 0 ? w + n : w;
}

template <typename T, int R>
__global__ __launch_bounds__(256) void stencil_direct(StencilParams<T> p)
{
    const long nxny = (long)p.nx * p.ny;
    const long nd = nxny * p.nz;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nd) return;
    const int col = blockIdx.y;
    const int k = (int)(idx / nxny);
    const long rem = idx - (long)k * nxny;
    const int j = (int)(rem / p.nx);
    const int i = (int)(rem - (long)j * p.nx);
    const T *x = p.x + col * p.ldx;
    const T xc = x[idx];
    T res = p.w0 * xc;
#pragma unroll
    for (int r = 1; r <= R; r++) {
        int a, b;
        a = axis_index(i - r, p.nx, p.bcx);
        b = axis_index(i + r, p.nx, p.bcx);
        T xm = a >= 0 ? x[k * nxny + (long)j * p.nx + a] : T(0);
        T xp = b >= 0 ? x[k * nxny + (long)j * p.nx + b] : T(0);
        T rx = (xm + xp) * p.wx[r];
        a = axis_index(j - r, p.ny, p.bcy);
        b = axis_index(j + r, p.ny, p.bcy);
        xm = a >= 0 ? x[k * nxny + (long)a * p.nx + i] : T(0);
        xp = b >= 0 ? x[k * nxny + (long)b * p.nx + i] : T(0);
        T ry = (xm + xp) * p.wy[r];
        a = axis_index(k - r, p.nz, p.bcz);
        b = axis_index(k + r, p.nz, p.bcz);
        xm = a >= 0 ? x[a * nxny + (long)j * p.nx + i] : T(0);
        xp = b >= 0 ? x[b * nxny + (long)j * p.nx + i] : T(0);
        T rz = (xm + xp) * p.wz[r];
        res += rx + ry + rz;
    }
    if (p.bdiag != T(0)) res += p.bdiag * (p.v[idx] * xc);
    T o = p.s1 * res;
    if (p.z) o += p.s2 * p.z[col * p.ldz + idx];
    p.out[col * p.ldo + idx] = o;
}

constexpr int kR = 6;
// Radius-6 stencil: a TX x TY tile of the x-y plane plus its 6-deep halo is
// staged in LDS per z-plane while the block marches along z; each thread owns
// RY points of the tile and keeps their 13 z-neighbours in a register queue.
template <typename T, int TX, int TY, int RY>
__global__ __launch_bounds__(TX *(TY / RY)) void stencil_tiled(StencilParams<T> p, int zchunk)
{
    constexpr int NTY = TY / RY;
    constexpr int NT = TX * NTY;
    constexpr int LX = TX + 2 * kR;
    constexpr int LY = TY + 2 * kR;
    constexpr int NHALO = 2 * kR * TY + 2 * kR * TX;
    constexpr int HPT = (NHALO + NT - 1) / NT;
    __shared__ T tile[2][LY][LX];

    const int tilesx = (p.nx + TX - 1) / TX;
    // tiles of one column (and z-chunk) are consecutive logical ids -> one
    // XCD, so the halo rows neighbouring tiles re-read stay in that L2
    const int lid = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                              gridDim.x * gridDim.y * gridDim.z);
    const int tid_xy = lid % gridDim.x, rest = lid / gridDim.x;
    const int bx = tid_xy % tilesx;
    const int by = tid_xy / tilesx;
    const int col = rest % gridDim.y;
    const int zs = (rest / gridDim.y) * zchunk;
    const int ze = min(p.nz, zs + zchunk);
    const int tx = threadIdx.x % TX;
    const int tyb = threadIdx.x / TX;
    const int i0 = bx * TX, j0 = by * TY;
    const long nxny = (long)p.nx * p.ny;
    const T *x = p.x + col * p.ldx;

    // in-plane offsets of my interior points (-1: Dirichlet zero / outside)
    int coff[RY], ooff[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int ty = tyb + r * NTY;
        const int ig = i0 + tx, jg = j0 + ty;
        const int ia = axis_index(ig, p.nx, p.bcx), ja = axis_index(jg, p.ny, p.bcy);
        coff[r] = (ia >= 0 && ja >= 0) ? ja * p.nx + ia : -1;
        ooff[r] = (ig < p.nx && jg < p.ny) ? jg * p.nx + ig : -1;
    }
    // my share of the halo: LDS slot and in-plane offset
    int hslot[HPT], hoff[HPT];
#pragma unroll
    for (int h = 0; h < HPT; h++) {
        const int e = threadIdx.x + h * NT;
        hslot[h] = -1;
        hoff[h] = -1;
        if (e < NHALO) {
            int ly, lx;
            if (e < 2 * kR * TY) {  // x halo: left/right strips of the interior rows
                const int row = e / (2 * kR), c = e % (2 * kR);
                ly = kR + row;
                lx = c < kR ? c : TX + c;
            } else {  // y halo: top/bottom strips of the interior columns
                const int e2 = e - 2 * kR * TY;
                const int rr = e2 / TX, c = e2 % TX;
                ly = rr < kR ? rr : TY + rr;
                lx = kR + c;
            }
            hslot[h] = ly * LX + lx;
            const int ia = axis_index(i0 - kR + lx, p.nx, p.bcx);
            const int ja = axis_index(j0 - kR + ly, p.ny, p.bcy);
            hoff[h] = (ia >= 0 && ja >= 0) ? ja * p.nx + ia : -1;
        }
    }

    // q[r][s] holds plane z-kR+s of point r; nq / nh are the prefetched values
    // of the next step (plane z+1+kR of my points, halo of plane z+1), issued
    // before this step's barrier so their latency overlaps the compute.
    // (Fetching a plane's halo together with its interior, kR steps early,
    // cut L2 misses by a quarter on MI355X but cost more in registers than it
    // saved; the halo is fetched for the current plane.)
    T q[RY][2 * kR + 1];
#pragma unroll
    for (int s = 0; s < 2 * kR; s++) {
        const int za = axis_index(zs - kR + s, p.nz, p.bcz);
        const T *xa = x + (long)(za < 0 ? 0 : za) * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) q[r][s + 1] = (za >= 0 && coff[r] >= 0) ? xa[coff[r]] : T(0);
    }
    // nv / nz2: Veff and the recurrence's z term of the next output plane,
    // loaded one step ahead like the x values
    const bool has_v = p.bdiag != T(0), has_z = p.z != nullptr;
    const T *zc = has_z ? p.z + col * p.ldz : nullptr;
    T nq[RY], nh[HPT], nv[RY], nz2[RY];
    {
        const int zn = axis_index(zs + kR, p.nz, p.bcz);
        const T *xn = x + (long)(zn < 0 ? 0 : zn) * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) nq[r] = (zn >= 0 && coff[r] >= 0) ? xn[coff[r]] : T(0);
        const T *xz = x + (long)zs * nxny;
#pragma unroll
        for (int h = 0; h < HPT; h++) nh[h] = hoff[h] >= 0 ? xz[hoff[h]] : T(0);
        const long zb = (long)zs * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) {
            nv[r] = (has_v && ooff[r] >= 0) ? p.v[zb + ooff[r]] : T(0);
            nz2[r] = (has_z && ooff[r] >= 0) ? zc[zb + ooff[r]] : T(0);
        }
    }
    int buf = 0;
    for (int z = zs; z < ze; z++) {
        T cv[RY], cz[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) {
#pragma unroll
            for (int s = 0; s < 2 * kR; s++) q[r][s] = q[r][s + 1];
            q[r][2 * kR] = nq[r];
            cv[r] = nv[r];
            cz[r] = nz2[r];
        }
        T hv[HPT];
#pragma unroll
        for (int h = 0; h < HPT; h++) hv[h] = nh[h];
        if (z + 1 < ze) {
            const int zn = axis_index(z + 1 + kR, p.nz, p.bcz);
            const T *xn = x + (long)(zn < 0 ? 0 : zn) * nxny;
#pragma unroll
            for (int r = 0; r < RY; r++) nq[r] = (zn >= 0 && coff[r] >= 0) ? xn[coff[r]] : T(0);
            const T *xz = x + (long)(z + 1) * nxny;
#pragma unroll
            for (int h = 0; h < HPT; h++) nh[h] = hoff[h] >= 0 ? xz[hoff[h]] : T(0);
            const long zb1 = (long)(z + 1) * nxny;
#pragma unroll
            for (int r = 0; r < RY; r++) {
                nv[r] = (has_v && ooff[r] >= 0) ? p.v[zb1 + ooff[r]] : T(0);
                nz2[r] = (has_z && ooff[r] >= 0) ? zc[zb1 + ooff[r]] : T(0);
            }
        }
        // double-buffered plane: one barrier per step (a thread can only
        // overwrite this buffer two steps later, after everyone passed the
        // next step's barrier)
        T(*tl)[LX] = tile[buf];
#pragma unroll
        for (int r = 0; r < RY; r++) tl[kR + tyb + r * NTY][kR + tx] = q[r][kR];
#pragma unroll
        for (int h = 0; h < HPT; h++)
            if (hslot[h] >= 0) (&tl[0][0])[hslot[h]] = hv[h];
        __syncthreads();
        buf ^= 1;
        const long zb = (long)z * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) {
            if (ooff[r] < 0) continue;
            const int ly = kR + tyb + r * NTY, lx = kR + tx;
            const T xc = q[r][kR];
            T res = p.w0 * xc;
#pragma unroll
            for (int s = 1; s <= kR; s++) {
                const T rx = (tl[ly][lx - s] + tl[ly][lx + s]) * p.wx[s];
                const T ry = (tl[ly - s][lx] + tl[ly + s][lx]) * p.wy[s];
                const T rz = (q[r][kR - s] + q[r][kR + s]) * p.wz[s];
                res += rx + ry + rz;
            }
            const long o = zb + ooff[r];
            if (has_v) res += p.bdiag * (cv[r] * xc);
            T val = p.s1 * res;
            if (has_z) val += p.s2 * cz[r];
            p.out[col * p.ldo + o] = val;
        }
    }
}

// Whole-row variant (TX == nx, FP32): the x halo of a row is the row itself,
// so it is written into LDS from the block's own interior values (periodic x)
// or as zeros (Dirichlet x) and never fetched; only the y halo (2 kR rows of
// TX points) is read from global memory.  The y halo of plane z + kR is read
// at step z - 1, in the same step as the neighbouring tiles read those rows as
// the interior of their z-queue (one HBM read, the others L2 hits), and held
// for kR steps in a thread-private LDS queue hq (slot (z - zs) % kR: plane z
// is taken out before plane z + kR goes in).  Same arithmetic in the same
// order as stencil_tiled, so the results are bitwise those of the other tiles.
template <int TX, int TY, int RY>
__global__ __launch_bounds__(TX *(TY / RY)) void stencil_rows(StencilParams<float> p, int zchunk)
{
    using T = float;
    constexpr int NTY = TY / RY;
    constexpr int NT = TX * NTY;
    constexpr int LX = TX + 2 * kR;
    constexpr int LY = TY + 2 * kR;
    constexpr int NHALO = 2 * kR * TX;
    constexpr int HPT = (NHALO + NT - 1) / NT;
    __shared__ T tile[2][LY][LX];
    __shared__ T hq[kR][HPT][NT];
    static_assert(sizeof(tile) + sizeof(hq) <= 65536, "static LDS limit");

    const int lid = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                              gridDim.x * gridDim.y * gridDim.z);
    const int by = lid % gridDim.x, rest = lid / gridDim.x;
    const int col = rest % gridDim.y;
    const int zs = (rest / gridDim.y) * zchunk;
    const int ze = min(p.nz, zs + zchunk);
    const int tid = threadIdx.x;
    const int tx = tid % TX;
    const int tyb = tid / TX;
    const int j0 = by * TY;
    const long nxny = (long)p.nx * p.ny;
    const T *x = p.x + col * p.ldx;
    const bool xper = p.bcx == 0;

    int coff[RY], ooff[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int jg = j0 + tyb + r * NTY;
        const int ja = axis_index(jg, p.ny, p.bcy);
        coff[r] = ja >= 0 ? ja * p.nx + tx : -1;
        ooff[r] = jg < p.ny ? jg * p.nx + tx : -1;
    }
    int hslot[HPT], hoff[HPT];
#pragma unroll
    for (int h = 0; h < HPT; h++) {
        const int e = tid + h * NT;
        hslot[h] = -1;
        hoff[h] = -1;
        if (e < NHALO) {
            const int rr = e / TX, c = e % TX;
            const int ly = rr < kR ? rr : TY + rr;
            hslot[h] = ly * LX + kR + c;
            const int ja = axis_index(j0 - kR + ly, p.ny, p.bcy);
            hoff[h] = ja >= 0 ? ja * p.nx + c : -1;
        }
    }

    T q[RY][2 * kR + 1];
#pragma unroll
    for (int s = 0; s < 2 * kR; s++) {
        const int za = axis_index(zs - kR + s, p.nz, p.bcz);
        const T *xa = x + (long)(za < 0 ? 0 : za) * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) q[r][s + 1] = (za >= 0 && coff[r] >= 0) ? xa[coff[r]] : T(0);
    }
    // y halos of planes zs .. zs + kR - 1 straight into the queue
    for (int s = 0; s < kR && zs + s < ze; s++) {
        const T *xs = x + (long)(zs + s) * nxny;
#pragma unroll
        for (int h = 0; h < HPT; h++) hq[s][h][tid] = hoff[h] >= 0 ? xs[hoff[h]] : T(0);
    }
    const bool has_v = p.bdiag != T(0), has_z = p.z != nullptr;
    const T *zc = has_z ? p.z + col * p.ldz : nullptr;
    T nq[RY], nh[HPT], nv[RY], nz2[RY];
    {
        const int zn = axis_index(zs + kR, p.nz, p.bcz);
        const T *xn = x + (long)(zn < 0 ? 0 : zn) * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) nq[r] = (zn >= 0 && coff[r] >= 0) ? xn[coff[r]] : T(0);
        const T *xh = x + (long)min(zs + kR, p.nz - 1) * nxny;
#pragma unroll
        for (int h = 0; h < HPT; h++) nh[h] = (zs + kR < ze && hoff[h] >= 0) ? xh[hoff[h]] : T(0);
        const long zb = (long)zs * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) {
            nv[r] = (has_v && ooff[r] >= 0) ? p.v[zb + ooff[r]] : T(0);
            nz2[r] = (has_z && ooff[r] >= 0) ? zc[zb + ooff[r]] : T(0);
        }
    }
    int buf = 0, slot = 0;
    for (int z = zs; z < ze; z++) {
        T cv[RY], cz[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) {
#pragma unroll
            for (int s = 0; s < 2 * kR; s++) q[r][s] = q[r][s + 1];
            q[r][2 * kR] = nq[r];
            cv[r] = nv[r];
            cz[r] = nz2[r];
        }
        T hv[HPT];
#pragma unroll
        for (int h = 0; h < HPT; h++) hv[h] = nh[h];
        if (z + 1 < ze) {
            const int zn = axis_index(z + 1 + kR, p.nz, p.bcz);
            const T *xn = x + (long)(zn < 0 ? 0 : zn) * nxny;
#pragma unroll
            for (int r = 0; r < RY; r++) nq[r] = (zn >= 0 && coff[r] >= 0) ? xn[coff[r]] : T(0);
            const bool hn = z + 1 + kR < ze;
            const T *xh = x + (long)(hn ? z + 1 + kR : 0) * nxny;
#pragma unroll
            for (int h = 0; h < HPT; h++) nh[h] = (hn && hoff[h] >= 0) ? xh[hoff[h]] : T(0);
            const long zb1 = (long)(z + 1) * nxny;
#pragma unroll
            for (int r = 0; r < RY; r++) {
                nv[r] = (has_v && ooff[r] >= 0) ? p.v[zb1 + ooff[r]] : T(0);
                nz2[r] = (has_z && ooff[r] >= 0) ? zc[zb1 + ooff[r]] : T(0);
            }
        }
        T(*tl)[LX] = tile[buf];
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int ly = kR + tyb + r * NTY;
            const T v = q[r][kR];
            tl[ly][kR + tx] = v;
            if (tx < kR) tl[ly][TX + kR + tx] = xper ? v : T(0);
            if (tx >= TX - kR) tl[ly][tx - TX + kR] = xper ? v : T(0);
        }
#pragma unroll
        for (int h = 0; h < HPT; h++) {
            if (hslot[h] < 0) continue;
            (&tl[0][0])[hslot[h]] = hq[slot][h][tid];
            hq[slot][h][tid] = hv[h];
        }
        slot = slot + 1 == kR ? 0 : slot + 1;
        __syncthreads();
        buf ^= 1;
        const long zb = (long)z * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) {
            if (ooff[r] < 0) continue;
            const int ly = kR + tyb + r * NTY, lx = kR + tx;
            const T xc = q[r][kR];
            T res = p.w0 * xc;
#pragma unroll
            for (int s = 1; s <= kR; s++) {
                const T rx = (tl[ly][lx - s] + tl[ly][lx + s]) * p.wx[s];
                const T ry = (tl[ly - s][lx] + tl[ly + s][lx]) * p.wy[s];
                const T rz = (q[r][kR - s] + q[r][kR + s]) * p.wz[s];
                res += rx + ry + rz;
            }
            const long o = zb + ooff[r];
            if (has_v) res += p.bdiag * (cv[r] * xc);
            T val = p.s1 * res;
            if (has_z) val += p.s2 * cz[r];
            p.out[col * p.ldo + o] = val;
        }
    }
}

// Ring variant: an NS-slot LDS ring of whole planes (interior + halo).  Plane
// z + kR is fetched kR steps before its x/y use -- interior into the
// register z-queue and the ring, halo into the ring -- i.e. at the moment the
// neighbouring tiles fetch the same rows as their interior, so halo rows are
// served from L2 instead of HBM.  One barrier per step: the slot written at
// step z, (z + kR) % NS, was last read at step z + kR - NS <= z - 2, and
// every thread has passed step z - 1's barrier since then.
template <typename T, int TX, int TY, int RY, int NS>
__global__ __launch_bounds__(TX *(TY / RY)) void stencil_ring(StencilParams<T> p, int zchunk)
{
    static_assert(NS >= kR + 2, "ring too short for the one-barrier schedule");
    constexpr int NTY = TY / RY;
    constexpr int NT = TX * NTY;
    constexpr int LX = TX + 2 * kR;
    constexpr int LY = TY + 2 * kR;
    constexpr int NHALO = 2 * kR * TY + 2 * kR * TX;
    constexpr int HPT = (NHALO + NT - 1) / NT;
    __shared__ T ring[NS][LY][LX];

    const int tilesx = (p.nx + TX - 1) / TX;
    const int lid = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                              gridDim.x * gridDim.y * gridDim.z);
    const int tid_xy = lid % gridDim.x, rest = lid / gridDim.x;
    const int bx = tid_xy % tilesx;
    const int by = tid_xy / tilesx;
    const int col = rest % gridDim.y;
    const int zs = (rest / gridDim.y) * zchunk;
    const int ze = min(p.nz, zs + zchunk);
    const int tx = threadIdx.x % TX;
    const int tyb = threadIdx.x / TX;
    const int i0 = bx * TX, j0 = by * TY;
    const long nxny = (long)p.nx * p.ny;
    const T *x = p.x + col * p.ldx;

    int coff[RY], ooff[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int ty = tyb + r * NTY;
        const int ig = i0 + tx, jg = j0 + ty;
        const int ia = axis_index(ig, p.nx, p.bcx), ja = axis_index(jg, p.ny, p.bcy);
        coff[r] = (ia >= 0 && ja >= 0) ? ja * p.nx + ia : -1;
        ooff[r] = (ig < p.nx && jg < p.ny) ? jg * p.nx + ig : -1;
    }
    int hslot[HPT], hoff[HPT];
#pragma unroll
    for (int h = 0; h < HPT; h++) {
        const int e = threadIdx.x + h * NT;
        hslot[h] = -1;
        hoff[h] = -1;
        if (e < NHALO) {
            int ly, lx;
            if (e < 2 * kR * TY) {
                const int row = e / (2 * kR), c = e % (2 * kR);
                ly = kR + row;
                lx = c < kR ? c : TX + c;
            } else {
                const int e2 = e - 2 * kR * TY;
                const int rr = e2 / TX, c = e2 % TX;
                ly = rr < kR ? rr : TY + rr;
                lx = kR + c;
            }
            hslot[h] = ly * LX + lx;
            const int ia = axis_index(i0 - kR + lx, p.nx, p.bcx);
            const int ja = axis_index(j0 - kR + ly, p.ny, p.bcy);
            hoff[h] = (ia >= 0 && ja >= 0) ? ja * p.nx + ia : -1;
        }
    }
    auto plane = [&](int zz) -> const T * {
        const int za = axis_index(zz, p.nz, p.bcz);
        return za < 0 ? nullptr : x + (long)za * nxny;
    };

    // register queue: q[r][s + 1] = plane zs - kR + s; ring slots for planes
    // zs .. zs + kR - 1 (interior from the queue, halo fetched here)
    T q[RY][2 * kR + 1];
#pragma unroll
    for (int s = 0; s < 2 * kR; s++) {
        const T *xa = plane(zs - kR + s);
#pragma unroll
        for (int r = 0; r < RY; r++) q[r][s + 1] = (xa && coff[r] >= 0) ? xa[coff[r]] : T(0);
    }
#pragma unroll
    for (int t = 0; t < kR; t++) {
        T(*sl)[LX] = ring[(zs + t) % NS];
#pragma unroll
        for (int r = 0; r < RY; r++) sl[kR + tyb + r * NTY][kR + tx] = q[r][kR + t + 1];
        const T *xa = plane(zs + t);
#pragma unroll
        for (int h = 0; h < HPT; h++)
            if (hslot[h] >= 0) (&sl[0][0])[hslot[h]] = (xa && hoff[h] >= 0) ? xa[hoff[h]] : T(0);
    }
    // prefetch of plane zs + kR (interior and halo)
    T nq[RY], nh[HPT];
    {
        const T *xn = plane(zs + kR);
#pragma unroll
        for (int r = 0; r < RY; r++) nq[r] = (xn && coff[r] >= 0) ? xn[coff[r]] : T(0);
#pragma unroll
        for (int h = 0; h < HPT; h++) nh[h] = (xn && hoff[h] >= 0) ? xn[hoff[h]] : T(0);
    }
    for (int z = zs; z < ze; z++) {
#pragma unroll
        for (int r = 0; r < RY; r++) {
#pragma unroll
            for (int s = 0; s < 2 * kR; s++) q[r][s] = q[r][s + 1];
            q[r][2 * kR] = nq[r];
        }
        {
            T(*sl)[LX] = ring[(z + kR) % NS];
#pragma unroll
            for (int r = 0; r < RY; r++) sl[kR + tyb + r * NTY][kR + tx] = nq[r];
#pragma unroll
            for (int h = 0; h < HPT; h++)
                if (hslot[h] >= 0) (&sl[0][0])[hslot[h]] = nh[h];
        }
        if (z + 1 < ze) {
            const T *xn = plane(z + 1 + kR);
#pragma unroll
            for (int r = 0; r < RY; r++) nq[r] = (xn && coff[r] >= 0) ? xn[coff[r]] : T(0);
#pragma unroll
            for (int h = 0; h < HPT; h++) nh[h] = (xn && hoff[h] >= 0) ? xn[hoff[h]] : T(0);
        }
        __syncthreads();
        T(*tl)[LX] = ring[z % NS];
        const long zb = (long)z * nxny;
#pragma unroll
        for (int r = 0; r < RY; r++) {
            if (ooff[r] < 0) continue;
            const int ly = kR + tyb + r * NTY, lx = kR + tx;
            const T xc = q[r][kR];
            T res = p.w0 * xc;
#pragma unroll
            for (int s = 1; s <= kR; s++) {
                const T rx = (tl[ly][lx - s] + tl[ly][lx + s]) * p.wx[s];
                const T ry = (tl[ly - s][lx] + tl[ly + s][lx]) * p.wy[s];
                const T rz = (q[r][kR - s] + q[r][kR + s]) * p.wz[s];
                res += rx + ry + rz;
            }
            const long o = zb + ooff[r];
            if (p.bdiag != T(0)) res += p.bdiag * (p.v[o] * xc);
            T val = p.s1 * res;
            if (p.z) val += p.s2 * p.z[col * p.ldz + o];
            p.out[col * p.ldo + o] = val;
        }
    }
}

template <typename T, int TX, int TY, int RY, int NS>
void launch_ring(sgpu_context *ctx, const StencilParams<T> &p, int ncol)
{
    const int tiles = ((ctx->nx + TX - 1) / TX) * ((ctx->ny + TY - 1) / TY);
    long want = (2048 + (long)tiles * ncol - 1) / ((long)tiles * ncol);
    int nzc = (int)std::max(1L, std::min<long>(want, std::max(1, ctx->nz / 16)));
    int zchunk = (ctx->nz + nzc - 1) / nzc;
    nzc = (ctx->nz + zchunk - 1) / zchunk;
    dim3 grid(tiles, ncol, nzc);
    stencil_ring<T, TX, TY, RY, NS><<<grid, TX * (TY / RY), 0, ctx->stream>>>(p, zchunk);
}

// min_chunk: fewest z-planes per block (each chunk re-reads 2 x 6 planes)
template <typename T, int TX, int TY, int RY>
void launch_tiled(sgpu_context *ctx, const StencilParams<T> &p, int ncol, int min_chunk = 16)
{
    const int tiles = ((ctx->nx + TX - 1) / TX) * ((ctx->ny + TY - 1) / TY);
    // enough blocks to fill the chip: split z when there are few columns
    long want = (2048 + (long)tiles * ncol - 1) / ((long)tiles * ncol);
    int nzc = (int)std::max(1L, std::min<long>(want, std::max(1, ctx->nz / min_chunk)));
    int zchunk = (ctx->nz + nzc - 1) / nzc;
    nzc = (ctx->nz + zchunk - 1) / zchunk;
    dim3 grid(tiles, ncol, nzc);
    stencil_tiled<T, TX, TY, RY><<<grid, TX * (TY / RY), 0, ctx->stream>>>(p, zchunk);
}

// Whole-row tiles need TX == nx; other grids take the tiled kernel of that shape.
template <typename T, int TX, int TY, int RY>
void launch_rows(sgpu_context *ctx, const StencilParams<T> &p, int ncol)
{
    if constexpr (sizeof(T) == 4) {
        if (ctx->nx == TX) {
            const int tiles = (ctx->ny + TY - 1) / TY;
            long want = (2048 + (long)tiles * ncol - 1) / ((long)tiles * ncol);
            int nzc = (int)std::max(1L, std::min<long>(want, std::max(1, ctx->nz / 16)));
            int zchunk = (ctx->nz + nzc - 1) / nzc;
            nzc = (ctx->nz + zchunk - 1) / zchunk;
            dim3 grid(tiles, ncol, nzc);
            stencil_rows<TX, TY, RY><<<grid, TX * (TY / RY), 0, ctx->stream>>>(p, zchunk);
            return;
        }
    }
    launch_tiled<T, TX, TY, RY>(ctx, p, ncol);
}

template <typename T>
void launch_cfg(sgpu_context *ctx, const StencilParams<T> &p, int ncol, int cfg)
{
    switch (cfg) {
        case 13: launch_rows<T, 96, 16, 2>(ctx, p, ncol); break;
        case 14: launch_rows<T, 80, 16, 2>(ctx, p, ncol); break;
        case 5:
            if (sizeof(T) == 4) launch_ring<T, 32, 16, 2, 8>(ctx, p, ncol);
            else launch_tiled<T, 32, 16, 2>(ctx, p, ncol);
            break;
        case 6:
            if (sizeof(T) == 4) launch_ring<T, 32, 8, 2, 8>(ctx, p, ncol);
            else launch_tiled<T, 32, 16, 2>(ctx, p, ncol);
            break;
        case 1: launch_tiled<T, 32, 32, 4>(ctx, p, ncol); break;
        case 2: launch_tiled<T, 64, 16, 4>(ctx, p, ncol); break;
        case 3: launch_tiled<T, 32, 32, 8>(ctx, p, ncol); break;
        case 4: launch_tiled<T, 64, 32, 8>(ctx, p, ncol); break;
        case 7: launch_tiled<T, 32, 32, 2>(ctx, p, ncol); break;
        case 8: launch_tiled<T, 64, 16, 2>(ctx, p, ncol); break;
        case 9: launch_tiled<T, 32, 16, 2>(ctx, p, ncol, 8); break;
        case 10: launch_tiled<T, 32, 16, 2>(ctx, p, ncol, 4); break;
        // whole x-rows for grids of 65-96 / 65-80 points along x: the x halo
        // is the block's own row (L2 hits), so only y halos cost HBM traffic
        case 11: launch_tiled<T, 96, 16, 2>(ctx, p, ncol); break;
        case 12: launch_tiled<T, 80, 16, 2>(ctx, p, ncol); break;
        default: launch_tiled<T, 32, 16, 2>(ctx, p, ncol); break;
    }
}

// Tile shape for one (precision, column count): the fastest of the
// candidates, each timed over three launches after a first untimed one, into a
// temporary output (the real out may alias z).  Every shape computes each point with the same
// operations in the same order, so the choice changes no result bit.
// Measured on MI355X: 32x32x2 fastest on 96^3 x 512, 64x16x2 on 78^3 x 480
// FP32, 32x16x2 on 52^3 x 160 FP32 -- no single shape wins.
template <typename T>
int tune_tile(sgpu_context *ctx, StencilParams<T> p, int ncol, int *best_cfg)
{
    // few columns (the Poisson / Kerker residual, Lanczos): also thinner z-chunks
    static const int cands_many[] = {0, 7, 8}, cands_few[] = {0, 7, 8, 9, 10};
    const bool few = ncol < 16;
    int cands[8], ncand = 0;
    for (int i = 0; i < (few ? 5 : 3); i++) cands[ncand++] = few ? cands_few[i] : cands_many[i];
    if (ctx->nx > 64 && ctx->nx <= 80) cands[ncand++] = 12;
    else if (ctx->nx > 80 && ctx->nx <= 96) cands[ncand++] = 11;
    if (sizeof(T) == 4 && ctx->nx == 80) cands[ncand++] = 14;
    else if (sizeof(T) == 4 && ctx->nx == 96) cands[ncand++] = 13;
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    T *tmp = nullptr;
    SGPU_HIP(hipMalloc((void **)&tmp, (size_t)ncol * p.ldo * sizeof(T)));
    p.out = tmp;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = SGPU_OK;
    float best = 0.f;
    *best_cfg = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = SGPU_ERR_HIP;
    for (int rep = 0; rep < 2 && st == SGPU_OK; rep++) {
        for (int ci = 0; ci < ncand; ci++) {
            const int c = cands[ci];
            float ms = 0.f;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) { st = SGPU_ERR_HIP; break; }
            // timed pass: three launches back to back (one launch varies by
            // about 5 % run to run on MI355X)
            for (int k = 0; k < (rep == 0 ? 1 : 3); k++) launch_cfg<T>(ctx, p, ncol, c);
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, ctx->stream) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                st = SGPU_ERR_HIP;
                break;
            }
            // a shape must beat the current best by 3 %, so that timing noise
            // between near-equal shapes does not change the choice run to run
            if (rep == 1 && (c == cands[0] || ms < 0.97f * best)) {
                best = c == cands[0] ? ms : std::min(best, ms);
                *best_cfg = c;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(tmp);
    if (st != SGPU_OK) sgpu_set_error("stencil: tile tuning failed");
    return st;
}

template <typename T>
int launch_typed(sgpu_context *ctx, int ncol, double a, double bdiag, double c, const T *v,
                 const T *x, long ldx, T *out, long ldo, const T *z, long ldz, double s1,
                 double s2, bool allow_tiled, bool may_tune)
{
    StencilParams<T> p;
    p.x = x; p.out = out; p.z = z; p.v = v;
    p.ldx = ldx; p.ldo = ldo; p.ldz = ldz;
    p.nx = ctx->nx; p.ny = ctx->ny; p.nz = ctx->nz;
    p.bcx = ctx->bcx; p.bcy = ctx->bcy; p.bcz = ctx->bcz;
    for (int r = 0; r <= SGPU_MAX_FDN; r++) {
        p.wx[r] = (T)(r <= ctx->fdn ? ctx->d2x[r] * a : 0.0);
        p.wy[r] = (T)(r <= ctx->fdn ? ctx->d2y[r] * a : 0.0);
        p.wz[r] = (T)(r <= ctx->fdn ? ctx->d2z[r] * a : 0.0);
    }
    p.w0 = (T)(a * (ctx->d2x[0] + ctx->d2y[0] + ctx->d2z[0]) + c);
    p.bdiag = (v && std::fabs(bdiag) >= 1e-14) ? (T)bdiag : T(0);
    p.s1 = (T)s1;
    p.s2 = (T)s2;
    if (p.bdiag != T(0) && !v) {
        sgpu_set_error("stencil: diagonal requested without Veff");
        return SGPU_ERR_STATE;
    }
    if (ncol <= 0) return SGPU_OK;
    if (ncol > 65535) {
        sgpu_set_error("stencil: ncol %d exceeds the launch limit", ncol);
        return SGPU_ERR_ARG;
    }
    if (allow_tiled && ctx->fdn == kR) {
        int cfg = ctx->stencil_cfg;
        if (cfg < 0) {
            const auto key = std::make_pair((int)sizeof(T), ncol);
            auto it = ctx->tile_choice.find(key);
            if (it == ctx->tile_choice.end() && !may_tune) {
                cfg = 0;
            } else if (it == ctx->tile_choice.end()) {
                SGPU_CHECK(tune_tile<T>(ctx, p, ncol, &cfg));
                ctx->tile_choice[key] = cfg;
            } else {
                cfg = it->second;
            }
        }
        launch_cfg<T>(ctx, p, ncol, cfg);
    } else {
        dim3 grid((unsigned)((ctx->nd + 255) / 256), ncol);
        switch (ctx->fdn) {
            case 1: stencil_direct<T, 1><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 2: stencil_direct<T, 2><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 3: stencil_direct<T, 3><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 4: stencil_direct<T, 4><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 5: stencil_direct<T, 5><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 6: stencil_direct<T, 6><<<grid, 256, 0, ctx->stream>>>(p); break;
            case 7: stencil_direct<T, 7><<<grid, 256, 0, ctx->stream>>>(p); break;
            default: stencil_direct<T, 8><<<grid, 256, 0, ctx->stream>>>(p); break;
        }
    }
    SGPU_HIP(hipGetLastError());
    return SGPU_OK;
}

}  // namespace

int sgpu_launch_stencil(sgpu_context *ctx, int dtype, int ncol, double a, double bdiag, double c,
                        const void *v, const void *x, long ldx, void *out, long ldo, const void *z,
                        long ldz, double s1, double s2, bool allow_tiled, bool may_tune)
{
    if (dtype == SGPU_F64)
        return launch_typed<double>(ctx, ncol, a, bdiag, c, (const double *)v, (const double *)x,
                                    ldx, (double *)out, ldo, (const double *)z, ldz, s1, s2,
                                    allow_tiled, may_tune);
    return launch_typed<float>(ctx, ncol, a, bdiag, c, (const float *)v, (const float *)x, ldx,
                               (float *)out, ldo, (const float *)z, ldz, s1, s2, allow_tiled,
                               may_tune);
}

extern "C" int sgpu_stencil_tile(const sgpu_context *ctx, int dtype, int ncol, int *tx, int *ty,
                                 int *ry)
{
    if (!ctx || !tx || !ty || !ry) {
        sgpu_set_error("sgpu_stencil_tile: bad arguments");
        return SGPU_ERR_ARG;
    }
    int cfg = ctx->stencil_cfg;
    if (cfg < 0) {
        auto it = ctx->tile_choice.find(std::make_pair(dtype == SGPU_F64 ? 8 : 4, ncol));
        if (it == ctx->tile_choice.end()) {
            *tx = *ty = *ry = 0;
            return SGPU_OK;
        }
        cfg = it->second;
    }
    static const int shape[15][3] = {{32, 16, 2}, {32, 32, 4}, {64, 16, 4}, {32, 32, 8},
                                     {64, 32, 8}, {32, 16, 2}, {32, 8, 2},  {32, 32, 2},
                                     {64, 16, 2}, {32, 16, 2}, {32, 16, 2}, {96, 16, 2},
                                     {80, 16, 2}, {96, 16, 2}, {80, 16, 2}};
    if (cfg > 14 || ((cfg == 5 || cfg == 6) && dtype == SGPU_F64)) cfg = 0;
    *tx = shape[cfg][0];
    *ty = shape[cfg][1];
    *ry = shape[cfg][2];
    return SGPU_OK;
}

extern "C" int sgpu_stencil_kernel(const sgpu_context *ctx, int dtype, int ncol, char *name,
                                   int len)
{
    int tx = 0, ty = 0, ry = 0;
    if (!name || len <= 0) {
        sgpu_set_error("sgpu_stencil_kernel: bad arguments");
        return SGPU_ERR_ARG;
    }
    SGPU_CHECK(sgpu_stencil_tile(ctx, dtype, ncol, &tx, &ty, &ry));
    int cfg = ctx->stencil_cfg;
    if (cfg < 0) {
        auto it = ctx->tile_choice.find(std::make_pair(dtype == SGPU_F64 ? 8 : 4, ncol));
        cfg = it == ctx->tile_choice.end() ? 0 : it->second;
    }
    const char *t = dtype == SGPU_F64 ? "double" : "float";
    if (tx == 0)
        name[0] = 0;
    else if ((cfg == 13 || cfg == 14) && dtype != SGPU_F64 && ctx->nx == tx)
        snprintf(name, len, "stencil_rows<%d,%d,%d>", tx, ty, ry);
    else if ((cfg == 5 || cfg == 6) && dtype != SGPU_F64)
        snprintf(name, len, "stencil_ring<%s,%d,%d,%d,8>", t, tx, ty, ry);
    else
        snprintf(name, len, "stencil_tiled<%s,%d,%d,%d>", t, tx, ty, ry);
    return SGPU_OK;
}
