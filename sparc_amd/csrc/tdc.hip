// Eigenvalues and eigenvectors of a symmetric tridiagonal matrix by divide
// and conquer on the device (Cuppen's method, as LAPACK dstedc that the
// reference's dsyevd / dsygvd reach, src/eigenSolver.c:1318), in place of
// rocSOLVER's dstedc, which spends 2.4 ms at n = 480 mostly in one secular
// solver kernel per merge level.  The matrix is split at every off-diagonal
// down to 1 x 1 blocks (the diagonal entries on both sides of a split lose
// |e|), and blocks are merged level by level: a merge of [a, m) and [m, b)
// with coupling beta solves diag(lambda) + rho z z' with z the last row of
// the left block's eigenvectors and sign(beta) times the first row of the
// right block's, rho = |beta| |z|^2, z normalised (LAPACK dlaed2-dlaed4):
//  * tdc_leaf (one wave per block of 8 poles): the three lowest levels in
//    one launch, every step below done by lanes in LDS;
//  * tdc_prep (one block per merge): the two sorted runs merged by rank,
//    deflation in sorted order (rho |z_j| <= tol: eigenvalue d_j, vector
//    unchanged; two poles whose Givens rotation t c s <= tol: rotated,
//    the first deflated), tol = 8 eps max(|d|max, |z|max), evaluated in
//    parallel with the sequential scan's decisions; the deflated values
//    sorted; the kept columns grouped by the row halves they span;
//  * tdc_secular (one wave per root): the root in (d_i, d_i+1) (the last in
//    (d_K-1, d_K-1 + rho z'z]) as tau = lambda - o, o the pole nearer the
//    root, so that d_j - lambda = (d_j - o) - tau keeps its relative
//    accuracy, by Li's middle-way iteration (LAPACK dlaed4) inside a
//    bracket; the waves without a root apply the deflation's rotation
//    chains to the old eigenvectors;
//  * tdc_zhat (one wave per pole): Gu and Eisenstat's z from the computed
//    roots (Loewner's formula as a product of ratios), so that the vectors
//    are orthogonal to working accuracy, and the sorted positions of the
//    roots and the deflated values in the block;
//  * tdc_vec (one wave per root): u_i = zhat / (d - lambda_i), normalised;
//  * tdc_gemm: the block's old eigenvectors (the non-deflated columns) times
//    U into the other buffer at the roots' positions, the deflated columns
//    copied to theirs.
// Every sum and product runs in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "sgpu_internal.h"

namespace {

constexpr int kTdcMaxJ = 24;   // terms per lane of a wave (n <= 64 * 24)
constexpr double kEps = 2.220446049250313e-16;

__device__ __forceinline__ double tdc_readlane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double tdc_dpp(double v, int sel)
{
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffff), hi = (int)(b >> 32);
    switch (sel) {
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

// Wave sum, the same bits in every lane (all 64 lanes must take part).
__device__ __forceinline__ double tdc_wave_sum(double v)
{
#pragma unroll
    for (int s = 0; s < 6; s++) v += tdc_dpp(v, s);
    return tdc_readlane(v, 63);
}

// Wave product in lane order (lanes without a factor hold 1), the same bits
// in every lane.
__device__ __forceinline__ double tdc_wave_prod(double v)
{
    double p = 1.0;
    for (int l = 0; l < 64; l++) p *= tdc_readlane(v, l);
    return p;
}

// Leaves: lambda_i = d_i - |e_i-1| - |e_i|, Z = I, Z2 = 0 (a level writes
// only the diagonal blocks of its output, and these grow from level to
// level, so entries outside them stay zero in both buffers; tdc_leaf then
// rewrites lambda and Z's diagonal blocks of its leaf size).
__global__ void tdc_init(const double *__restrict__ d, const double *__restrict__ e, int n,
                         double *__restrict__ lam, double *__restrict__ Z,
                         double *__restrict__ Z2, int ld)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        double v = d[t];
        if (t > 0) v -= fabs(e[t - 1]);
        if (t + 1 < n) v -= fabs(e[t]);
        lam[t] = v;
    }
    if (t < (long)n * n) {
        const int c = (int)(t / n), r = (int)(t - (long)c * n);
        Z[(long)c * ld + r] = r == c ? 1.0 : 0.0;
        Z2[(long)c * ld + r] = 0.0;
    }
}

// Exclusive positions of the flagged items 0..cnt-1 (in order) across the
// block: out(i, pos) for each flagged i; returns the count.  256 threads.
template <class P, class O>
__device__ int tdc_compact(int cnt, P pred, O out, int *wsum)
{
    const int t = threadIdx.x, w = t >> 6, ln = t & 63;
    int base = 0;
    for (int r0 = 0; r0 < cnt; r0 += 256) {
        const int i = r0 + t;
        const bool f = i < cnt && pred(i);
        const unsigned long long bal = __ballot(f);
        const int pre = __popcll(bal & ((1ull << ln) - 1ull));
        if (ln == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int q = 0; q < w; q++) off += wsum[q];
        if (f) out(i, off + pre);
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    return base;
}

// A chain of deflating rotations from kept pair k (poles L[k], L[k+1] of the
// sorted list, dlaed2's test |t c s| <= tol with t the gap, c, s the Givens
// rotation that zeroes the first z): while a pair deflates, the second pole
// takes the rotated value and |z| and is tested against the next.  Returns
// the first pair that does not deflate (or M - 1).  With write, records
// each rotation (c, s, the deflated value), the second poles' new values and
// the row halves the last (kept) pole's column now spans.
__device__ int tdc_walk(int k, int M, const int *L, double *ds, double *zs, double tol, bool write,
                        int *rf, double *rcv, double *rsv, double *dvr, int *ctype)
{
    double dp = ds[L[k]], zp = zs[L[k]];
    int ty = write ? ctype[L[k]] : 0;
    while (k < M - 1) {
        const int j = L[k + 1];
        const double dn0 = ds[j], zn0 = zs[j];
        const double tau = hypot(zn0, zp);
        const double C = zn0 / tau, S = -zp / tau;
        if (!(fabs((dn0 - dp) * C * S) <= tol)) break;
        const double t2 = dp * C * C + dn0 * S * S;
        const double dn = dp * S * S + dn0 * C * C;
        if (write) {
            rf[k] = 1;
            rcv[k] = C;
            rsv[k] = S;
            dvr[k] = t2;
            ds[j] = dn;
            zs[j] = tau;
            ty |= ctype[j];
        }
        dp = dn;
        zp = tau;
        k++;
    }
    if (write) ctype[L[k]] = ty;
    return k;
}

// Per merge mg (block): a = 2 s mg, m = a + s, b = min(a + 2 s, n).
// Outputs at offset a: ndp / ndz / ndc (the K non-deflated poles, z values,
// Z columns, ascending), dfv / dfc (the deflated values, ascending, and
// their columns), and per merge K, rho.  Deflation follows LAPACK dlaed2's
// scan in sorted order (a pole with rho |z| <= tol deflates; of two
// consecutive kept poles whose rotation passes the test the first deflates
// and the second, rotated, is tested against the next), evaluated in
// parallel: the small-z poles by compaction, every consecutive kept pair
// tested on the original values, and one thread per pair that passes
// following its chain; a chain that starts inside an earlier one is dropped
// (the earlier chain's rotated values decide there).  The same decisions and
// values as the sequential scan.  The chains' rotations of Z's columns are
// left to tdc_secular (colk, rC, rS per kept pair, chs / che per chain,
// nch per merge).  For the product with the old eigenvectors the kept
// columns are ordered as in dlaed2: those nonzero only in rows [a, m), those
// nonzero in both halves (mixed by a rotation), those only in [m, b)
// (pinv: sorted index -> that order, zcol: Z column in that order, nlm /
// nmm: the first two counts).
struct TdcPrepOut {
    double *ndp, *ndz, *dfv, *rC, *rS;
    int *ndc, *dfc, *Km, *colk, *chs, *che, *nch, *pinv, *zcol, *nlm, *nmm;
    double *rhom;
};

__global__ __launch_bounds__(256) void tdc_prep(const double *__restrict__ lam,
                                                const double *__restrict__ Z, int ld, int n, int s,
                                                const double *__restrict__ e, TdcPrepOut o)
{
    extern __shared__ double tsm[];
    const int mg = blockIdx.x, t = threadIdx.x;
    const int a = 2 * s * mg, m = min(a + s, n), b = min(a + 2 * s, n);
    const int nl = b - a;
    double *ds = tsm;                 // sorted poles
    double *zs = ds + nl;             // their z
    double *rcv = zs + nl;            // per kept pair: rotation cosine
    double *rsv = rcv + nl;           //   sine
    double *dvr = rsv + nl;           //   deflated value
    double *dv = dvr + nl;            // deflated values (small z, then rotated)
    int *cols = (int *)(dv + nl);     // Z column of a sorted pole
    int *L = cols + nl;               // kept (not small z) sorted indices
    int *f0 = L + nl;                 // per kept pair: passes on original values
    int *rf = f0 + nl;                //   rotated
    int *cand = rf + nl;              // pairs that pass (chain starts)
    int *cend = cand + nl;            //   their chain ends
    int *ndl = cend + nl;             // non-deflated sorted indices
    int *dfl = ndl + nl;              // deflated sorted indices
    int *ctype = dfl + nl;            // row halves a sorted pole's column spans
    __shared__ double red[8];
    __shared__ int wsum[4];
    __shared__ int cnt3[3];
    if (m >= b) {
        // no partner at this level: every value deflated, in order
        for (int i = t; i < nl; i += 256) {
            o.dfv[a + i] = lam[a + i];
            o.dfc[a + i] = a + i;
        }
        if (t == 0) { o.Km[mg] = 0; o.rhom[mg] = 0.0; o.nch[mg] = 0; o.nlm[mg] = 0; o.nmm[mg] = 0; }
        return;
    }
    const double beta = e[m - 1];
    const double sg = beta >= 0.0 ? 1.0 : -1.0;
    // merge the two ascending runs by rank (left run first on ties)
    for (int i = a + t; i < b; i += 256) {
        const double v = lam[i];
        int pos;
        if (i < m) {
            int lo = m, hi = b;   // first right value >= v
            while (lo < hi) { const int md = (lo + hi) >> 1; if (lam[md] < v) lo = md + 1; else hi = md; }
            pos = (i - a) + (lo - m);
        } else {
            int lo = a, hi = m;   // first left value > v
            while (lo < hi) { const int md = (lo + hi) >> 1; if (lam[md] <= v) lo = md + 1; else hi = md; }
            pos = (i - m) + (lo - a);
        }
        ds[pos] = v;
        zs[pos] = i < m ? Z[(long)i * ld + (m - 1)] : sg * Z[(long)i * ld + m];
        cols[pos] = i;
        ctype[pos] = i < m ? 1 : 2;
    }
    __syncthreads();
    // |z|^2 and max |d| (fixed order: per-thread partials, then by wave)
    double z2 = 0.0, dmax = 0.0;
    for (int i = t; i < nl; i += 256) {
        z2 += zs[i] * zs[i];
        dmax = fmax(dmax, fabs(ds[i]));
    }
    {
        const double a2 = tdc_wave_sum(z2);
        for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
        if ((t & 63) == 0) { red[t >> 6] = a2; red[4 + (t >> 6)] = dmax; }
    }
    __syncthreads();
    const double zz = red[0] + red[1] + red[2] + red[3];
    const double dm = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    const double nz = sqrt(zz), rho = fabs(beta) * zz;
    __syncthreads();
    double zmax = 0.0;
    for (int i = t; i < nl; i += 256) {
        zs[i] = nz > 0.0 ? zs[i] / nz : 0.0;
        zmax = fmax(zmax, fabs(zs[i]));
    }
    for (int off = 32; off > 0; off >>= 1) zmax = fmax(zmax, __shfl_xor(zmax, off, 64));
    if ((t & 63) == 0) red[t >> 6] = zmax;
    __syncthreads();
    const double zm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const double tol = 8.0 * kEps * fmax(dm, zm);
    const bool all_small = rho * zm <= tol;
    auto small = [&](int j) { return all_small || rho * fabs(zs[j]) <= tol; };
    const int M = tdc_compact(nl, [&](int j) { return !small(j); }, [&](int j, int p) { L[p] = j; }, wsum);
    const int ns = tdc_compact(nl, small, [&](int j, int p) { dfl[p] = j; dv[p] = ds[j]; }, wsum);
    for (int k = t; k < M - 1; k += 256) {
        const double dp = ds[L[k]], zp = zs[L[k]], dn = ds[L[k + 1]], zn = zs[L[k + 1]];
        const double tau = hypot(zn, zp);
        f0[k] = fabs((dn - dp) * (zn / tau) * (-zp / tau)) <= tol ? 1 : 0;
        rf[k] = 0;
    }
    for (int k = t; k < M; k += 256) o.colk[a + k] = cols[L[k]];
    __syncthreads();
    const int nc = tdc_compact(M - 1, [&](int k) { return f0[k] != 0; }, [&](int k, int p) { cand[p] = k; }, wsum);
    int nr = 0, nch = 0;
    if (nc > 0) {
        for (int q = t; q < nc; q += 256)
            cend[q] = tdc_walk(cand[q], M, L, ds, zs, tol, false, rf, rcv, rsv, dvr, ctype);
        __syncthreads();
        if (t == 0) {
            int last = -1;
            for (int q = 0; q < nc; q++) {
                if (cand[q] > last) last = cend[q];
                else cand[q] = -1;
            }
        }
        __syncthreads();
        for (int q = t; q < nc; q += 256)
            if (cand[q] >= 0) tdc_walk(cand[q], M, L, ds, zs, tol, true, rf, rcv, rsv, dvr, ctype);
        __syncthreads();
        nr = tdc_compact(M - 1, [&](int k) { return rf[k] != 0; },
                         [&](int k, int p) {
                             dfl[ns + p] = L[k];
                             dv[ns + p] = dvr[k];
                             o.rC[a + k] = rcv[k];
                             o.rS[a + k] = rsv[k];
                         },
                         wsum);
        nch = tdc_compact(nc, [&](int q) { return cand[q] >= 0; },
                          [&](int q, int p) { o.chs[a + p] = cand[q]; o.che[a + p] = cend[q]; }, wsum);
    }
    const int K = tdc_compact(M, [&](int k) { return k == M - 1 || rf[k] == 0; },
                              [&](int k, int p) { ndl[p] = L[k]; }, wsum);
    const int nd = ns + nr;
    for (int k = t; k < K; k += 256) {
        const int j = ndl[k];
        o.ndp[a + k] = ds[j];
        o.ndz[a + k] = zs[j];
        o.ndc[a + k] = cols[j];
    }
    // kept columns grouped by the row halves they span (top only, both,
    // bottom only), each group in sorted order
    int base = 0;
    for (int ty = 1; ty <= 3; ty++) {
        const int want = ty == 1 ? 1 : (ty == 2 ? 3 : 2);
        const int c = tdc_compact(K, [&](int k) { return ctype[ndl[k]] == want; },
                                  [&](int k, int p) {
                                      o.pinv[a + k] = base + p;
                                      o.zcol[a + base + p] = cols[ndl[k]];
                                  },
                                  wsum);
        if (t == 0 && ty < 3) cnt3[ty - 1] = c;
        base += c;
    }
    // deflated values ascending (the small-z ones already are; with rotated
    // ones, by rank, ties in list order)
    for (int q = t; q < nd; q += 256) {
        const double v = dv[q];
        int rkq = q;
        if (nr > 0) {
            rkq = 0;
            for (int u = 0; u < nd; u++) {
                const double w = dv[u];
                rkq += (w < v || (w == v && u < q)) ? 1 : 0;
            }
        }
        o.dfv[a + rkq] = v;
        o.dfc[a + rkq] = cols[dfl[q]];
    }
    __syncthreads();
    if (t == 0) {
        o.Km[mg] = K;
        o.rhom[mg] = rho;
        o.nch[mg] = nch;
        o.nlm[mg] = cnt3[0];
        o.nmm[mg] = cnt3[1];
    }
}

// Merge and index of a wave (merges of 2 s values, one wave per value);
// true when the wave has a root (index below the merge's K).
__device__ __forceinline__ bool tdc_wave_item(int s, int n, const int *Km, int &mg, int &i, int &a,
                                              int &K)
{
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    mg = (int)(w / (2 * s));
    i = (int)(w - (long)mg * 2 * s);
    a = 2 * s * mg;
    K = 0;
    if (a >= n) return false;
    K = Km[mg];
    return i < K;
}

// Four wave sums at once (independent DPP chains), the same bits in every
// lane.
__device__ __forceinline__ void tdc_wave_sum4(double &p, double &q, double &u, double &v)
{
#pragma unroll
    for (int s = 0; s < 6; s++) {
        p += tdc_dpp(p, s);
        q += tdc_dpp(q, s);
        u += tdc_dpp(u, s);
        v += tdc_dpp(v, s);
    }
    p = tdc_readlane(p, 63);
    q = tdc_readlane(q, 63);
    u = tdc_readlane(u, 63);
    v = tdc_readlane(v, 63);
}

// Root i of 1/rho + sum_j z_j^2 / (d_j - lambda) = 0 (one wave; lane l holds
// poles l, l + 64, ...).  The root lies in (d_i, d_i+1), the last one in
// (d_K-1, d_K-1 + rho); it is found as tau = lambda - o from the pole o
// nearer to it (chosen by the sign of the function at the midpoint) by
// Li's "middle way" iteration, as in LAPACK dlaed4: at the current tau the
// sums over the poles left (psi) and right (phi) of the root are each
// replaced by a constant plus one pole term matching value and slope, and
// the resulting two-pole equation (a quadratic) is solved for the next tau.
// A bracket kept from the function's signs catches steps that leave it
// (bisection instead).  Stops when |f| is below its rounding-error bound.
__global__ __launch_bounds__(256) void tdc_secular(int n, int s, const int *__restrict__ Km,
                                                   const double *__restrict__ rhom,
                                                   const double *__restrict__ ndp,
                                                   const double *__restrict__ ndz,
                                                   double *__restrict__ orig,
                                                   double *__restrict__ tauv,
                                                   double *__restrict__ Z, int ld,
                                                   const int *__restrict__ nchm,
                                                   const int *__restrict__ chs,
                                                   const int *__restrict__ che,
                                                   const int *__restrict__ colk,
                                                   const double *__restrict__ rC,
                                                   const double *__restrict__ rS)
{
    int mg, i, a, K;
    const int ln = threadIdx.x & 63;
    if (!tdc_wave_item(s, n, Km, mg, i, a, K)) {
        // a wave without a root: rotation chain i - K of its merge (if any)
        // on the block's rows of Z, the carried column in registers
        if (a >= n || i < K || i - K >= nchm[mg]) return;
        const int q = i - K, b = min(a + 2 * s, n), nu = (b - a + 63) >> 6;
        const int c = chs[a + q], en = che[a + q];
        double x[kTdcMaxJ], y[kTdcMaxJ];
        const long c0 = (long)colk[a + c] * ld;
#pragma unroll
        for (int u = 0; u < kTdcMaxJ; u++) {
            const int r = a + ln + 64 * u;
            x[u] = u < nu && r < b ? Z[c0 + r] : 0.0;
        }
        for (int k = c; k < en; k++) {
            const double C = rC[a + k], S = rS[a + k];
            const long cp = (long)colk[a + k] * ld, cn = (long)colk[a + k + 1] * ld;
#pragma unroll
            for (int u = 0; u < kTdcMaxJ; u++) {
                const int r = a + ln + 64 * u;
                y[u] = u < nu && r < b ? Z[cn + r] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kTdcMaxJ; u++) {
                const int r = a + ln + 64 * u;
                if (u < nu && r < b) {
                    Z[cp + r] = C * x[u] + S * y[u];
                    x[u] = C * y[u] - S * x[u];
                }
            }
        }
        const long ce = (long)colk[a + en] * ld;
#pragma unroll
        for (int u = 0; u < kTdcMaxJ; u++) {
            const int r = a + ln + 64 * u;
            if (u < nu && r < b) Z[ce + r] = x[u];
        }
        return;
    }
    const int nq = (K + 63) >> 6;
    const double rho = rhom[mg], rinv = 1.0 / rho;
    const double *dd = ndp + a, *zz = ndz + a;
    const bool last = i == K - 1;
    double dl[kTdcMaxJ], w[kTdcMaxJ];
    double o = dd[i], lo = 0.0, hi, tt;
#pragma unroll
    for (int q = 0; q < kTdcMaxJ; q++) {
        const int j = ln + 64 * q;
        dl[q] = q < nq && j < K ? dd[j] - o : 1.0;
        w[q] = q < nq && j < K ? zz[j] * zz[j] : 0.0;
    }
    double psi, phi, dpsi, dphi;
    auto eval = [&](double x) {
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
        for (int q = 0; q < kTdcMaxJ; q++) {
            if (q < nq && ln + 64 * q < K) {
                const double r = 1.0 / (dl[q] - x);
                const double tw = w[q] * r;
                if (ln + 64 * q <= i) { p0 += tw; p2 += tw * r; }
                else { p1 += tw; p3 += tw * r; }
            }
        }
        tdc_wave_sum4(p0, p1, p2, p3);
        psi = p0; phi = p1; dpsi = p2; dphi = p3;
        return rinv + psi + phi;
    };
    double g;
    if (!last) {
        const double mid = 0.5 * (dd[i + 1] - dd[i]);
        g = eval(mid);
        if (g > 0.0) {
            hi = tt = mid;
        } else {
            o = dd[i + 1];
#pragma unroll
            for (int q = 0; q < kTdcMaxJ; q++) {
                const int j = ln + 64 * q;
                dl[q] = q < nq && j < K ? dd[j] - o : 1.0;
            }
            lo = tt = -mid;
            hi = 0.0;
            g = eval(tt);
        }
    } else {
        double ws = 0.0;
#pragma unroll
        for (int q = 0; q < kTdcMaxJ; q++) ws += w[q];
        hi = tt = rho * tdc_wave_sum(ws);
        g = eval(tt);
    }
    const double di = dd[i] - o, di1 = last ? 0.0 : dd[i + 1] - o;
    for (int it = 0; it < 160; it++) {
        const double err = 8.0 * (phi - psi + rinv) + fabs(tt) * (dpsi + dphi);
        if (fabs(g) <= kEps * err) break;
        if (g > 0.0) hi = tt;
        else lo = tt;
        if (!(hi > lo) || hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi))) break;
        const double a1 = di - tt;
        const double B = dpsi * a1 * a1;
        double tn = 0.5 * (lo + hi);
        if (!last) {
            const double a2 = di1 - tt;
            const double D = dphi * a2 * a2;
            const double c0 = rinv + (psi - B / a1) + (phi - D / a2);
            // c0 eta^2 - bq eta + a1 a2 g = 0
            const double bq = c0 * (a1 + a2) + B + D;
            const double r0 = a1 * a2 * g;
            const double disc = bq * bq - 4.0 * c0 * r0;
            if (disc >= 0.0) {
                const double qq = 0.5 * (bq + copysign(sqrt(disc), bq));
                const double e1 = qq != 0.0 ? r0 / qq : 0.0;
                const double e2 = c0 != 0.0 ? qq / c0 : 0.0;
                if (qq != 0.0 && lo < tt + e1 && tt + e1 < hi) tn = tt + e1;
                else if (c0 != 0.0 && lo < tt + e2 && tt + e2 < hi) tn = tt + e2;
            }
        } else {
            const double c0 = rinv + (psi - B / a1);
            if (c0 > 0.0) {
                const double x = tt + a1 + B / c0;
                if (lo < x && x < hi) tn = x;
            }
        }
        if (!(lo < tn && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == tt) break;
        tt = tn;
        g = eval(tt);
    }
    if (ln == 0) {
        orig[a + i] = o;
        tauv[a + i] = tt;
    }
}

// One wave per index i of a merge: for a root (i < K) Gu and Eisenstat's
// zhat_i and the root's sorted position within the block, i + #(deflated <
// lambda_i); for a deflated value (i < #deflated) its position, i + #(roots
// <= value).  The new eigenvalues go to lamn at those positions.
__global__ __launch_bounds__(256) void tdc_zhat(int n, int s, const int *__restrict__ Km,
                                                const double *__restrict__ rhom,
                                                const double *__restrict__ ndp,
                                                const double *__restrict__ ndz,
                                                const double *__restrict__ orig,
                                                const double *__restrict__ tauv,
                                                const double *__restrict__ dfv,
                                                double *__restrict__ zh, int *__restrict__ posr,
                                                int *__restrict__ posd, double *__restrict__ lamn)
{
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int mg = (int)(w / (2 * s)), j = (int)(w - (long)mg * 2 * s), a = 2 * s * mg;
    if (a >= n) return;
    const int K = Km[mg], nd = (min(a + 2 * s, n) - a) - K;
    const int ln = threadIdx.x & 63;
    if (j < K) {
        const double *dd = ndp + a;
        const double dj = dd[j];
        double p = 1.0;
        for (int i = ln; i < K; i += 64) {
            const double num = (orig[a + i] - dj) + tauv[a + i];
            p *= i == j ? num / rhom[mg] : num / (dd[i] - dj);
        }
        const double pr = tdc_wave_prod(p);
        if (ln == 0) {
            zh[a + j] = copysign(sqrt(fabs(pr)), ndz[a + j]);
            const double v = orig[a + j] + tauv[a + j];
            int lo = 0, hi = nd;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (dfv[a + md] < v) lo = md + 1;
                else hi = md;
            }
            posr[a + j] = j + lo;
            lamn[a + j + lo] = v;
        }
    }
    if (j < nd && ln == 0) {
        const double v = dfv[a + j];
        int lo = 0, hi = K;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (orig[a + md] + tauv[a + md] <= v) lo = md + 1;
            else hi = md;
        }
        posd[a + j] = j + lo;
        lamn[a + j + lo] = v;
    }
}

// U (column i of the merge's K x K block at U + a (ld + 1), column-major;
// row pinv[j] for pole j, the order of the old eigenvectors in the product).
__global__ __launch_bounds__(256) void tdc_vec(int n, int s, const int *__restrict__ Km,
                                               const double *__restrict__ ndp,
                                               const double *__restrict__ orig,
                                               const double *__restrict__ tauv,
                                               const double *__restrict__ zh,
                                               const int *__restrict__ pinv,
                                               double *__restrict__ U, int ld)
{
    int mg, i, a, K;
    if (!tdc_wave_item(s, n, Km, mg, i, a, K)) return;
    const int ln = threadIdx.x & 63;
    const double o = orig[a + i], tt = tauv[a + i];
    double uv[kTdcMaxJ], ss = 0.0;
#pragma unroll
    for (int q = 0; q < kTdcMaxJ; q++) {
        const int j = ln + 64 * q;
        uv[q] = j < K ? zh[a + j] / ((ndp[a + j] - o) - tt) : 0.0;
        ss += uv[q] * uv[q];
    }
    const double nrm = sqrt(tdc_wave_sum(ss));
    double *col = U + (long)(a + i) * ld + a;
#pragma unroll
    for (int q = 0; q < kTdcMaxJ; q++) {
        const int j = ln + 64 * q;
        if (j < K) col[pinv[a + j]] = uv[q] / nrm;
    }
}

// The block's new eigenvectors: column i < K of the merge (a root) is
// Z[rows a..b, zcol] U[:, i], column K + q a deflated column copied; both go
// to their sorted positions.  64 x 64 output tiles, 4 x 4 per thread (grid:
// row tiles, column tiles, merges), the sum in chunks of 16 old columns; a
// row tile inside [a, m) sums over the columns nonzero there (the first
// nlm + nmm), one inside [m, b) over the last K - nlm.
constexpr int kGT = 64, kGK = 32;
__global__ __launch_bounds__(256) void tdc_gemm(int n, int s, const int *__restrict__ Km,
                                                const int *__restrict__ nlm,
                                                const int *__restrict__ nmm,
                                                const double *__restrict__ Z, int ld,
                                                const int *__restrict__ zcol,
                                                const double *__restrict__ U,
                                                const int *__restrict__ posr,
                                                const int *__restrict__ dfc,
                                                const int *__restrict__ posd,
                                                double *__restrict__ Zn)
{
    const int mg = blockIdx.z, a = 2 * s * mg;
    if (a >= n) return;
    const int b = min(a + 2 * s, n), m = min(a + s, n), K = Km[mg];
    const int r0 = a + blockIdx.x * kGT, i0 = blockIdx.y * kGT;
    if (r0 >= b || i0 >= b - a) return;
    __shared__ double za[kGK][kGT + 1], ub[kGK][kGT + 1];
    const int t = threadIdx.x, tr = t & 15, tc = t >> 4;
    constexpr int kPer = kGK * kGT / 256;   // elements of each operand a thread stages
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
    if (i0 < K) {
        int jlo = 0, jhi = K;
        if (r0 + kGT <= m) jhi = nlm[mg] + nmm[mg];
        else if (r0 >= m) jlo = nlm[mg];
        double pa[kPer], pb[kPer];
        // chunk j0 of both operands into registers (zero outside)
        auto fetch = [&](int j0) {
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const int e = t + 256 * q;
                const int rr = e & (kGT - 1), jj = e / kGT;
                const int r = r0 + rr, j = j0 + jj;
                pa[q] = (r < b && j < jhi) ? Z[(long)zcol[a + j] * ld + r] : 0.0;
                const int ii = e / kGK, j2 = e & (kGK - 1);
                pb[q] = (i0 + ii < K && j0 + j2 < jhi) ? U[(long)(a + i0 + ii) * ld + a + j0 + j2]
                                                       : 0.0;
            }
        };
        if (jlo < jhi) fetch(jlo);
        for (int j0 = jlo; j0 < jhi; j0 += kGK) {
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const int e = t + 256 * q;
                za[e / kGT][e & (kGT - 1)] = pa[q];
                ub[e & (kGK - 1)][e / kGK] = pb[q];
            }
            __syncthreads();
            if (j0 + kGK < jhi) fetch(j0 + kGK);   // in flight during the products
#pragma unroll 8
            for (int jj = 0; jj < kGK; jj++) {
                double x[4], y[4];
#pragma unroll
                for (int u = 0; u < 4; u++) x[u] = za[jj][tr + 16 * u];
#pragma unroll
                for (int v = 0; v < 4; v++) y[v] = ub[jj][tc + 16 * v];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[u][v] += x[u] * y[v];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int i = i0 + tc + 16 * v;
        if (i >= b - a) continue;
        const bool root = i < K;
        const long dst = (long)(a + (root ? posr[a + i] : posd[a + i - K])) * ld;
        const long src = root ? 0 : (long)dfc[a + i - K] * ld;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = r0 + tr + 16 * u;
            if (r < b) Zn[dst + r] = root ? acc[u][v] : Z[src + r];
        }
    }
}

// ---- the lowest levels in one launch ----
// Merges of up to S poles run inside one wave per leaf block [a0, a0 + L)
// (L = S but for the last block): every level of the block in LDS, the same
// steps as the kernels above with lanes for elements: each merge's first
// lane runs dlaed2's deflation scan sequentially (at most S poles) and its
// rotations, one lane per root runs the middle-way iteration with the sums
// taken in order, one lane per pole Gu and Eisenstat's z, one lane per root
// its vector, and the lanes share the products.  The global levels then
// start at merges of 2 S.  With one wave per CU every step here is latency:
// S = 8 (three levels in one launch instead of fifteen) is the fastest on
// MI355X (n = 480: 0.39 ms against 0.47 without and 0.50 with S = 32).
constexpr int kLeafMax = 32;

// Root i of the merge's secular equation (poles dd[0..K), z values zz), as
// tdc_secular with sequential sums.
__device__ void tdc_root_seq(const double *dd, const double *zz, int K, int i, double rho,
                             double &o_out, double &t_out)
{
    const double rinv = 1.0 / rho;
    const bool last = i == K - 1;
    double o = dd[i], lo = 0.0, hi, tt;
    double psi = 0.0, phi = 0.0, dpsi = 0.0, dphi = 0.0;
    auto eval = [&](double x) {
        psi = phi = dpsi = dphi = 0.0;
        for (int j = 0; j < K; j++) {
            const double r = 1.0 / ((dd[j] - o) - x);
            const double tw = zz[j] * zz[j] * r;
            if (j <= i) { psi += tw; dpsi += tw * r; }
            else { phi += tw; dphi += tw * r; }
        }
        return rinv + psi + phi;
    };
    double g;
    if (!last) {
        const double mid = 0.5 * (dd[i + 1] - dd[i]);
        g = eval(mid);
        if (g > 0.0) {
            hi = tt = mid;
        } else {
            o = dd[i + 1];
            lo = tt = -mid;
            hi = 0.0;
            g = eval(tt);
        }
    } else {
        double ws = 0.0;
        for (int j = 0; j < K; j++) ws += zz[j] * zz[j];
        hi = tt = rho * ws;
        g = eval(tt);
    }
    const double di = dd[i] - o, di1 = last ? 0.0 : dd[i + 1] - o;
    for (int it = 0; it < 160; it++) {
        const double err = 8.0 * (phi - psi + rinv) + fabs(tt) * (dpsi + dphi);
        if (fabs(g) <= kEps * err) break;
        if (g > 0.0) hi = tt;
        else lo = tt;
        if (!(hi > lo) || hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi))) break;
        const double a1 = di - tt;
        const double B = dpsi * a1 * a1;
        double tn = 0.5 * (lo + hi);
        if (!last) {
            const double a2 = di1 - tt;
            const double D = dphi * a2 * a2;
            const double c0 = rinv + (psi - B / a1) + (phi - D / a2);
            const double bq = c0 * (a1 + a2) + B + D;
            const double r0 = a1 * a2 * g;
            const double disc = bq * bq - 4.0 * c0 * r0;
            if (disc >= 0.0) {
                const double qq = 0.5 * (bq + copysign(sqrt(disc), bq));
                const double e1 = qq != 0.0 ? r0 / qq : 0.0;
                const double e2 = c0 != 0.0 ? qq / c0 : 0.0;
                if (qq != 0.0 && lo < tt + e1 && tt + e1 < hi) tn = tt + e1;
                else if (c0 != 0.0 && lo < tt + e2 && tt + e2 < hi) tn = tt + e2;
            }
        } else {
            const double c0 = rinv + (psi - B / a1);
            if (c0 > 0.0) {
                const double x = tt + a1 + B / c0;
                if (lo < x && x < hi) tn = x;
            }
        }
        if (!(lo < tn && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == tt) break;
        tt = tn;
        g = eval(tt);
    }
    o_out = o;
    t_out = tt;
}

template <int S>
__global__ __launch_bounds__(64) void tdc_leaf(const double *__restrict__ d,
                                               const double *__restrict__ e, int n,
                                               double *__restrict__ lam_out,
                                               double *__restrict__ Zg, int ld)
{
    static_assert(S <= kLeafMax && S <= 64, "one lane per element");
    __shared__ double Zs[2][S * S];   // column-major, local rows and columns
    __shared__ double Us[S * S];
    __shared__ double lam[2][S];
    __shared__ double ds[S], zs[S], dfv[S], ndp[S], ndz[S], org[S], tau[S], zh[S];
    __shared__ int cols[S], ndc[S], dfc[S], posr[S], posd[S];
    __shared__ int Km[S];
    __shared__ double rhm[S];
    const int t = threadIdx.x, a0 = blockIdx.x * S, L = min(S, n - a0);
    if (t < L) {
        const int g = a0 + t;
        double v = d[g];
        if (g > 0) v -= fabs(e[g - 1]);
        if (g + 1 < n) v -= fabs(e[g]);
        lam[0][t] = v;
    }
    for (int f = t; f < S * S; f += 64) {
        Zs[0][f] = (f / S == f % S) ? 1.0 : 0.0;
        Zs[1][f] = 0.0;
    }
    __syncthreads();
    int cur = 0;
    for (int s = 1; s < L; s *= 2) {
        const double *Zc = Zs[cur];
        double *Zn = Zs[cur ^ 1];
        // merges [a, m) + [m, b) of the level; element t's merge
        const int mg = t / (2 * s), a = 2 * s * mg;
        const int m = min(a + s, L), b = min(a + 2 * s, L);
        const bool partner = t < L && m < b;
        // the two sorted runs merged by rank, z from the old eigenvectors
        if (partner) {
            const double sg = e[a0 + m - 1] >= 0.0 ? 1.0 : -1.0;
            const double v = lam[cur][t];
            int pos;
            if (t < m) {
                int lo = m, hi = b;
                while (lo < hi) { const int md = (lo + hi) >> 1; if (lam[cur][md] < v) lo = md + 1; else hi = md; }
                pos = (t - a) + (lo - m);
            } else {
                int lo = a, hi = m;
                while (lo < hi) { const int md = (lo + hi) >> 1; if (lam[cur][md] <= v) lo = md + 1; else hi = md; }
                pos = (t - m) + (lo - a);
            }
            ds[a + pos] = v;
            zs[a + pos] = t < m ? Zc[t * S + (m - 1)] : sg * Zc[t * S + m];
            cols[a + pos] = t;
        }
        __syncthreads();
        // deflation (dlaed2's scan) and rotations by each merge's first lane
        if (t < L && t == a) {
            const int nl = b - a;
            if (m >= b) {
                for (int q = 0; q < nl; q++) { dfv[a + q] = lam[cur][a + q]; dfc[a + q] = a + q; }
                Km[mg] = 0;
                rhm[mg] = 0.0;
            } else {
                const double beta = e[a0 + m - 1];
                double zz = 0.0, dm = 0.0, zm = 0.0;
                for (int j = a; j < b; j++) { zz += zs[j] * zs[j]; dm = fmax(dm, fabs(ds[j])); }
                const double nz = sqrt(zz), rho = fabs(beta) * zz;
                for (int j = a; j < b; j++) { zs[j] = nz > 0.0 ? zs[j] / nz : 0.0; zm = fmax(zm, fabs(zs[j])); }
                const double tol = 8.0 * kEps * fmax(dm, zm);
                int K = 0, nd = 0;
                // deflated values kept sorted (insertion, ties in scan order)
                auto push_defl = [&](double v, int c) {
                    int q = nd++;
                    while (q > 0 && dfv[a + q - 1] > v) {
                        dfv[a + q] = dfv[a + q - 1];
                        dfc[a + q] = dfc[a + q - 1];
                        q--;
                    }
                    dfv[a + q] = v;
                    dfc[a + q] = c;
                };
                if (rho * zm <= tol) {
                    for (int j = a; j < b; j++) push_defl(ds[j], cols[j]);
                } else {
                    int pj = -1;
                    for (int j = a; j < b; j++) {
                        if (rho * fabs(zs[j]) <= tol) { push_defl(ds[j], cols[j]); continue; }
                        if (pj < 0) { pj = j; continue; }
                        const double tau2 = hypot(zs[j], zs[pj]);
                        const double C = zs[j] / tau2, Sn = -zs[pj] / tau2;
                        if (fabs((ds[j] - ds[pj]) * C * Sn) <= tol) {
                            const int cp = cols[pj], cn = cols[j];
                            for (int r = a; r < b; r++) {
                                const double x = Zc[cp * S + r], y = Zc[cn * S + r];
                                Zs[cur][cp * S + r] = C * x + Sn * y;
                                Zs[cur][cn * S + r] = C * y - Sn * x;
                            }
                            const double t2 = ds[pj] * C * C + ds[j] * Sn * Sn;
                            ds[j] = ds[pj] * Sn * Sn + ds[j] * C * C;
                            zs[j] = tau2;
                            push_defl(t2, cp);
                        } else {
                            ndp[a + K] = ds[pj]; ndz[a + K] = zs[pj]; ndc[a + K] = cols[pj]; K++;
                        }
                        pj = j;
                    }
                    if (pj >= 0) { ndp[a + K] = ds[pj]; ndz[a + K] = zs[pj]; ndc[a + K] = cols[pj]; K++; }
                }
                Km[mg] = K;
                rhm[mg] = rho;
            }
        }
        __syncthreads();
        const int K = t < L ? Km[mg] : 0, nd = t < L ? (b - a) - K : 0, i = t - a;
        if (t < L && i < K) tdc_root_seq(ndp + a, ndz + a, K, i, rhm[mg], org[a + i], tau[a + i]);
        __syncthreads();
        // Gu-Eisenstat z, sorted positions, new eigenvalues
        if (t < L && i < K) {
            const double dj = ndp[a + i];
            double pr = ((org[a + i] - dj) + tau[a + i]) / rhm[mg];
            for (int q = 0; q < K; q++)
                if (q != i) pr *= ((org[a + q] - dj) + tau[a + q]) / (ndp[a + q] - dj);
            zh[a + i] = copysign(sqrt(fabs(pr)), ndz[a + i]);
            const double v = org[a + i] + tau[a + i];
            int lo = 0, hi = nd;
            while (lo < hi) { const int md = (lo + hi) >> 1; if (dfv[a + md] < v) lo = md + 1; else hi = md; }
            posr[a + i] = i + lo;
            lam[cur ^ 1][a + i + lo] = v;
        }
        if (t < L && i < nd) {
            const double v = dfv[a + i];
            int lo = 0, hi = K;
            while (lo < hi) { const int md = (lo + hi) >> 1; if (org[a + md] + tau[a + md] <= v) lo = md + 1; else hi = md; }
            posd[a + i] = i + lo;
            lam[cur ^ 1][a + i + lo] = v;
        }
        __syncthreads();
        // the vector of root i: zhat / (d - lambda_i), normalised
        if (t < L && i < K) {
            double ss = 0.0;
            for (int j = 0; j < K; j++) {
                const double u = zh[a + j] / ((ndp[a + j] - org[a + i]) - tau[a + i]);
                Us[(a + i) * S + a + j] = u;
                ss += u * u;
            }
            const double rn = 1.0 / sqrt(ss);
            for (int j = 0; j < K; j++) Us[(a + i) * S + a + j] *= rn;
        }
        __syncthreads();
        // new eigenvectors of each merge's block at their sorted positions
        for (int f = t; f < L * L; f += 64) {
            const int r = f % L, c = f / L;
            const int mc = c / (2 * s), ac = 2 * s * mc, bc = min(ac + 2 * s, L);
            if (r < ac || r >= bc) continue;
            const int Kc = Km[mc], ic = c - ac;
            if (ic < Kc) {
                double acc = 0.0;
                for (int j = 0; j < Kc; j++) acc += Zc[ndc[ac + j] * S + r] * Us[(ac + ic) * S + ac + j];
                Zn[(ac + posr[ac + ic]) * S + r] = acc;
            } else {
                const int q = ic - Kc;
                Zn[(ac + posd[ac + q]) * S + r] = Zc[dfc[ac + q] * S + r];
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    if (t < L) lam_out[a0 + t] = lam[cur][t];
    for (int f = t; f < L * L; f += 64) {
        const int r = f % L, c = f / L;
        Zg[(long)(a0 + c) * ld + a0 + r] = Zs[cur][c * S + r];
    }
}

}  // namespace

// Workspace doubles for n.
size_t sgpu_tdc_work_doubles(int n)
{
    const size_t nn = (size_t)n * n;
    return 2 * nn + 20 * (size_t)n + 64;
}

// D (n) -> eigenvalues ascending, E (n - 1) off-diagonal (read only), Z (n x n,
// leading dimension n) <- eigenvectors; work: sgpu_tdc_work_doubles(n).
int sgpu_tdc(sgpu_context *ctx, int n, double *dD, const double *dE, double *dZ, double *work)
{
    static_assert(kTdcMaxN <= 64 * kTdcMaxJ, "a wave holds every term of a secular equation");
    if (n < 1 || n > kTdcMaxN) {
        sgpu_set_error("tridiagonal divide and conquer: n = %d outside 1..%d", n, kTdcMaxN);
        return SGPU_ERR_ARG;
    }
    const hipStream_t st = ctx->stream;
    const size_t nn = (size_t)n * n;
    double *ZB = work, *U = ZB + nn, *lamA = U + nn, *lamB = lamA + n;
    double *orig = lamB + n, *tauv = orig + n, *zh = tauv + n;
    TdcPrepOut po;
    po.ndp = zh + n;
    po.ndz = po.ndp + n;
    po.dfv = po.ndz + n;
    po.rC = po.dfv + n;
    po.rS = po.rC + n;
    po.rhom = po.rS + n;
    po.ndc = (int *)(po.rhom + n);
    po.dfc = po.ndc + n;
    po.Km = po.dfc + n;
    po.colk = po.Km + n;
    po.chs = po.colk + n;
    po.che = po.chs + n;
    po.nch = po.che + n;
    po.pinv = po.nch + n;
    po.zcol = po.pinv + n;
    po.nlm = po.zcol + n;
    po.nmm = po.nlm + n;
    int *posr = po.nmm + n, *posd = posr + n;
    double *Za = dZ, *Zb = ZB, *la = lamA, *lb = lamB;
    // per-kernel timers only on request (their events cost about 5 us each)
    static const bool inner = getenv("SPARC_AMD_TDC_TIMERS") != nullptr;
    struct Restore {
        sgpu_context *c;
        bool v;
        ~Restore() { c->timing = v; }
    } restore{ctx, ctx->timing};
    ctx->timing = ctx->timing && inner;
    sgpu_tmark tm;
    SGPU_CHECK(sgpu_tbegin(ctx, &tm));
    tdc_init<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(dD, dE, n, la, Za, Zb, n);
    const int leaf = ctx->tdc_leaf;   // 1: no leaf kernel
    const unsigned nleaf = (unsigned)((n + leaf - 1) / leaf);
    if (leaf == 8)
        tdc_leaf<8><<<nleaf, 64, 0, st>>>(dD, dE, n, la, Za, n);
    else if (leaf == 16)
        tdc_leaf<16><<<nleaf, 64, 0, st>>>(dD, dE, n, la, Za, n);
    else if (leaf == 32)
        tdc_leaf<32><<<nleaf, 64, 0, st>>>(dD, dE, n, la, Za, n);
    SGPU_HIP(hipGetLastError());
    SGPU_CHECK(sgpu_tend(ctx, &tm, "tdc_init"));
    for (int s = leaf; s < n; s *= 2) {
        const int nm = (n + 2 * s - 1) / (2 * s);
        const int nl = std::min(2 * s, n);
        const size_t shm = (size_t)nl * (6 * sizeof(double) + 10 * sizeof(int));
        if (shm > 64 * 1024)
            SGPU_HIP(hipFuncSetAttribute((const void *)tdc_prep,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        SGPU_CHECK(sgpu_tbegin(ctx, &tm));
        tdc_prep<<<nm, 256, shm, st>>>(la, Za, n, n, s, dE, po);
        SGPU_CHECK(sgpu_tend(ctx, &tm, "tdc_prep"));
        const unsigned nwb = (unsigned)(((long)nm * 2 * s + 3) / 4);   // waves of 4 per block
        SGPU_CHECK(sgpu_tbegin(ctx, &tm));
        tdc_secular<<<nwb, 256, 0, st>>>(n, s, po.Km, po.rhom, po.ndp, po.ndz, orig, tauv, Za, n,
                                         po.nch, po.chs, po.che, po.colk, po.rC, po.rS);
        SGPU_CHECK(sgpu_tend(ctx, &tm, "tdc_secular"));
        SGPU_CHECK(sgpu_tbegin(ctx, &tm));
        tdc_zhat<<<nwb, 256, 0, st>>>(n, s, po.Km, po.rhom, po.ndp, po.ndz, orig, tauv, po.dfv, zh,
                                      posr, posd, lb);
        tdc_vec<<<nwb, 256, 0, st>>>(n, s, po.Km, po.ndp, orig, tauv, zh, po.pinv, U, n);
        SGPU_CHECK(sgpu_tend(ctx, &tm, "tdc_vec"));
        SGPU_CHECK(sgpu_tbegin(ctx, &tm));
        const unsigned nt = (unsigned)((nl + kGT - 1) / kGT);
        tdc_gemm<<<dim3(nt, nt, nm), 256, 0, st>>>(n, s, po.Km, po.nlm, po.nmm, Za, n, po.zcol, U,
                                                  posr, po.dfc, posd, Zb);
        SGPU_CHECK(sgpu_tend(ctx, &tm, "tdc_gemm"));
        SGPU_HIP(hipGetLastError());
        std::swap(Za, Zb);
        std::swap(la, lb);
    }
    if (Za != dZ)
        SGPU_HIP(hipMemcpyAsync(dZ, Za, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    SGPU_HIP(hipMemcpyAsync(dD, la, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return SGPU_OK;
}
