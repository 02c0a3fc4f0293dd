// Local (pseudocharge) and NLCC terms of the atomic forces on the device
// (SPARC Calculate_local_forces_linear, src/forces.c:157-457, and
// Calculate_forces_xc_linear, :473-707), orthogonal cells:
//
//   F_J^loc = dV sum over J's box [ -b_J grad(phi)
//             + 1/2 ( grad(Vc_J) (b + b_ref) - grad(Vc) (b_J + b_J^ref) ) ]
//   F_J^xc  = dV sum over J's box  Vxc grad(rho_c,J)
//
// per image atom J, with b_J = -Lap(V_J) / 4 pi, b_J^ref = -Lap(V_J^ref) / 4 pi
// and Vc_J = V_J^ref - V_J on the box extended by F = order / 2 points.
//
//  1. grad3: grad(phi) and grad(Vc) on the periodic grid (first-derivative
//     weights w_p = (-1)^(p+1) F!^2 / ((F-p)! (F+p)!) / (p h)).
//  2. box_eval: V_J (cubic Hermite spline of r V_loc on the uniform radial
//     grid divided by r, V_loc(0) at the nucleus, -Z / r beyond the table),
//     V_J^ref (Calculate_Pseudopot_Ref) and rho_c,J (spline, zero beyond the
//     table) at every point of every extended box, into one scratch array.
//  3. box_force: per inner box point the FD Laplacians and gradients of the
//     staged values times the global fields; per-block partial sums of the
//     six components, then per image and per atom in a fixed order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sgpu_internal.h"

namespace {

struct D1 {
    double w[3][SGPU_MAX_FDN + 1];   // first derivative, per axis (scaled)
    double l[3][SGPU_MAX_FDN + 1];   // second derivative, per axis (scaled)
};

// periodic first derivatives of f along x, y, z (one thread per point)
__global__ void grad3(const double *__restrict__ f, int nx, int ny, int nz, int fdn, D1 c,
                      double *__restrict__ gx, double *__restrict__ gy, double *__restrict__ gz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nd = (long)nx * ny * nz;
    if (i >= nd) return;
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((long)nx * ny));
    const long ro = i - ix, co = i - (long)iy * nx, po = i - (long)iz * nx * ny;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int p = 1; p <= fdn; p++) {
        int a = ix + p, b = ix - p;
        a -= a >= nx ? nx : 0; b += b < 0 ? nx : 0;
        sx += c.w[0][p] * (f[ro + a] - f[ro + b]);
        a = iy + p; b = iy - p;
        a -= a >= ny ? ny : 0; b += b < 0 ? ny : 0;
        sy += c.w[1][p] * (f[co + (long)a * nx] - f[co + (long)b * nx]);
        a = iz + p; b = iz - p;
        a -= a >= nz ? nz : 0; b += b < 0 ? nz : 0;
        sz += c.w[2][p] * (f[po + (long)a * nx * ny] - f[po + (long)b * nx * ny]);
    }
    gx[i] = sx; gy[i] = sy; gz[i] = sz;
}

// cubic Hermite spline on a uniform knot grid (SplineInterpUniform,
// src/tools.c:552-603): knots r, values y, derivatives yd
__device__ double spline_u(const double *__restrict__ r, const double *__restrict__ y,
                           const double *__restrict__ yd, int size, double q)
{
    int j = (int)floor((q - r[0]) / (r[1] - r[0]));
    j = max(0, min(j, size - 2));
    const double p1 = r[j], dxi = 1.0 / (r[j + 1] - r[j]);
    const double dy = (y[j + 1] - y[j]) * dxi;
    const double a0 = y[j], a1 = yd[j];
    const double a2 = dxi * (3.0 * dy - 2.0 * yd[j] - yd[j + 1]);
    const double a3 = dxi * dxi * (-2.0 * dy + yd[j] + yd[j + 1]);
    const double t = q - p1;
    return ((a3 * t + a2) * t + a1) * t + a0;
}

struct TypeDev {
    const double *r, *rvloc, *rvloc_d, *rhoc, *rhoc_d;
    int size;
    double rmax, zion, vloc0;
    int nlcc;
};

constexpr int kMaxTypes = 8;
struct Types {
    TypeDev t[kMaxTypes];
};

// per image: type, atom, extended-box origin (grid indices), extended-box
// sizes, scratch offset, centre
struct Img {
    int type, atom;
    int ex0, ey0, ez0;     // grid index of the extended box's first point
    int nex, ney, nez;     // extended box sizes (inner + 2F)
    long off;              // first value of this box in the scratch arrays
    double cx, cy, cz;
};

__global__ void box_eval(const Img *__restrict__ imgs, int n_img, Types T, double hx, double hy,
                         double hz, double ref_cutoff, double *__restrict__ VJ,
                         double *__restrict__ VR, double *__restrict__ RC)
{
    const int k = blockIdx.y;
    if (k >= n_img) return;
    const Img im = imgs[k];
    const long npt = (long)im.nex * im.ney * im.nez;
    const TypeDev &ty = T.t[im.type];
    const double rc3 = ref_cutoff * ref_cutoff * ref_cutoff, rc6 = rc3 * rc3;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npt; i += (long)gridDim.x * blockDim.x) {
        const int a = (int)(i % im.nex), b = (int)((i / im.nex) % im.ney), c = (int)(i / ((long)im.nex * im.ney));
        const double x = (im.ex0 + a) * hx - im.cx, y = (im.ey0 + b) * hy - im.cy,
                     z = (im.ez0 + c) * hz - im.cz;
        const double R = sqrt(x * x + y * y + z * z);
        double v;
        if (R > ty.rmax) v = -ty.zion / R;
        else if (fabs(R) < 1e-12) v = ty.vloc0;
        else v = spline_u(ty.r, ty.rvloc, ty.rvloc_d, ty.size, R) / R;
        double vr;
        if (R <= ref_cutoff) {
            const double R2 = R * R;
            vr = -ty.zion * (R2 * (R2 * R * (5.6 / rc6 + R * (1.8 * R / (rc6 * ref_cutoff * ref_cutoff) -
                                                               6.0 / (rc6 * ref_cutoff))) -
                                   2.8 / rc3) +
                             2.4 / ref_cutoff);
        } else {
            vr = -ty.zion / R;
        }
        const double rc = (ty.nlcc && R <= ty.rmax) ? spline_u(ty.r, ty.rhoc, ty.rhoc_d, ty.size, R) : 0.0;
        VJ[im.off + i] = v;
        VR[im.off + i] = vr;
        RC[im.off + i] = rc;
    }
}

// per inner point of image k: the six force integrands, block partials
__global__ __launch_bounds__(256) void box_force(const Img *__restrict__ imgs, int n_img, int fdn,
                                                 D1 c, int nx, int ny, const double *__restrict__ VJ,
                                                 const double *__restrict__ VR,
                                                 const double *__restrict__ RC,
                                                 const double *__restrict__ dphi,
                                                 const double *__restrict__ dvc,
                                                 const double *__restrict__ bsum,
                                                 const double *__restrict__ vxc, long nd,
                                                 int blocks_per_img, double *__restrict__ part)
{
    const int k = blockIdx.x / blocks_per_img, blk = blockIdx.x % blocks_per_img;
    __shared__ double red[6][256];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (k < n_img) {
        const Img im = imgs[k];
        const int F = fdn;
        const int ni = im.nex - 2 * F, nj = im.ney - 2 * F, nk = im.nez - 2 * F;
        const long npt = (long)ni * nj * nk;
        const long sy = im.nex, sz = (long)im.nex * im.ney;
        const double w0 = c.l[0][0] + c.l[1][0] + c.l[2][0];
        const double m4pi = -0.25 / M_PI;
        for (long i = (long)blk * 256 + threadIdx.x; i < npt; i += (long)blocks_per_img * 256) {
            const int a = (int)(i % ni), b = (int)((i / ni) % nj), cc = (int)(i / ((long)ni * nj));
            const long e = im.off + (long)(cc + F) * sz + (long)(b + F) * sy + (a + F);
            const long g = ((long)(im.ez0 + F + cc) * ny + (im.ey0 + F + b)) * nx + (im.ex0 + F + a);
            // b_J, b_J^ref: -(1 / 4 pi) Lap on the extended box (lap_box order)
            double lv = w0 * VJ[e], lr = w0 * VR[e];
            for (int p = 1; p <= F; p++) {
                lv += c.l[0][p] * (VJ[e + p] + VJ[e - p]);
                lr += c.l[0][p] * (VR[e + p] + VR[e - p]);
                lv += c.l[1][p] * (VJ[e + p * sy] + VJ[e - p * sy]);
                lr += c.l[1][p] * (VR[e + p * sy] + VR[e - p * sy]);
                lv += c.l[2][p] * (VJ[e + p * sz] + VJ[e - p * sz]);
                lr += c.l[2][p] * (VR[e + p * sz] + VR[e - p * sz]);
            }
            const double bJ = m4pi * lv, bJr = m4pi * lr;
            const long stride[3] = {1, sy, sz};
            const double bb = bsum[g], vx = vxc[g];
            for (int d = 0; d < 3; d++) {
                double gvc = 0.0, grc = 0.0;
                for (int p = 1; p <= F; p++) {
                    const long s = p * stride[d];
                    gvc += c.w[d][p] * ((VR[e + s] - VJ[e + s]) - (VR[e - s] - VJ[e - s]));
                    grc += c.w[d][p] * (RC[e + s] - RC[e - s]);
                }
                acc[d] += -bJ * dphi[d * nd + g] + 0.5 * (gvc * bb - dvc[d * nd + g] * (bJ + bJr));
                acc[3 + d] += vx * grc;
            }
        }
    }
    for (int q = 0; q < 6; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 6; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(long)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

extern "C" int sgpu_local_forces(sgpu_context *ctx, int n_types, const sgpu_psp_radial *types,
                                 int n_img, const int *img_type, const int *img_atom,
                                 const int *img_box, const double *img_centre, double ref_cutoff,
                                 const double *phi, const double *vc, const double *bsum,
                                 const double *vxc, int n_atom, double *f_loc, double *f_xc)
{
    if (!ctx || n_types < 1 || n_types > kMaxTypes || !types || n_img < 0 || n_atom < 1 || !phi ||
        !vc || !bsum || !vxc || !f_loc || !f_xc || (n_img > 0 && (!img_type || !img_atom ||
                                                                   !img_box || !img_centre))) {
        sgpu_set_error("sgpu_local_forces: bad arguments");
        return SGPU_ERR_ARG;
    }
    if (n_img > 65535) {
        sgpu_set_error("sgpu_local_forces: %d image boxes exceed the launch limit", n_img);
        return SGPU_ERR_ARG;
    }
    if (ctx->bcx || ctx->bcy || ctx->bcz) {
        sgpu_set_error("sgpu_local_forces: periodic cells only");
        return SGPU_ERR_ARG;
    }
    const int F = ctx->fdn;
    const long nd = ctx->nd;
    hipStream_t s = ctx->stream;
    D1 c{};
    const double h[3] = {ctx->hx, ctx->hy, ctx->hz};
    const double *d2[3] = {ctx->d2x, ctx->d2y, ctx->d2z};
    for (int d = 0; d < 3; d++) {
        for (int p = 1; p <= F; p++) {
            double fr = 1.0;
            for (int q = 1; q <= p; q++) fr *= (double)(F - p + q) / (double)(F + q);
            c.w[d][p] = ((p % 2) ? 1.0 : -1.0) * fr / p / h[d];
        }
        for (int p = 0; p <= F; p++) c.l[d][p] = d2[d][p];
    }
    // images: check the boxes and lay out the scratch
    std::vector<Img> im(n_img);
    long total = 0;
    const int n3[3] = {ctx->nx, ctx->ny, ctx->nz};
    for (int k = 0; k < n_img; k++) {
        const int *bx = img_box + 6 * k;   // xs, ys, zs, xe, ye, ze (inner, inside the grid)
        for (int d = 0; d < 3; d++)
            if (bx[d] < 0 || bx[3 + d] >= n3[d] || bx[d] > bx[3 + d]) {
                sgpu_set_error("sgpu_local_forces: image %d box outside the grid", k);
                return SGPU_ERR_ARG;
            }
        if (img_type[k] < 0 || img_type[k] >= n_types || img_atom[k] < 0 || img_atom[k] >= n_atom) {
            sgpu_set_error("sgpu_local_forces: image %d type or atom out of range", k);
            return SGPU_ERR_ARG;
        }
        Img &q = im[k];
        q.type = img_type[k];
        q.atom = img_atom[k];
        q.ex0 = bx[0] - F; q.ey0 = bx[1] - F; q.ez0 = bx[2] - F;
        q.nex = bx[3] - bx[0] + 1 + 2 * F;
        q.ney = bx[4] - bx[1] + 1 + 2 * F;
        q.nez = bx[5] - bx[2] + 1 + 2 * F;
        q.off = total;
        q.cx = img_centre[3 * k]; q.cy = img_centre[3 * k + 1]; q.cz = img_centre[3 * k + 2];
        total += (long)q.nex * q.ney * q.nez;
    }
    for (int i = 0; i < 3 * n_atom; i++) f_loc[i] = f_xc[i] = 0.0;
    if (n_img == 0) return SGPU_OK;
    Types T{};
    std::vector<double *> dev;
    int st = SGPU_OK;
    auto up = [&](const double *h_, size_t n) -> double * {
        double *p = nullptr;
        if (st != SGPU_OK) return nullptr;
        if (hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess ||
            (n && hipMemcpyAsync(p, h_, n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)) {
            st = SGPU_ERR_HIP;
            return nullptr;
        }
        dev.push_back(p);
        return p;
    };
    for (int t = 0; t < n_types; t++) {
        const sgpu_psp_radial &r = types[t];
        if (r.size < 2 || !r.r || !(r.r[1] > r.r[0]) || !r.rvloc || !r.rvloc_d) {
            sgpu_set_error("sgpu_local_forces: bad radial table of type %d", t);
            return SGPU_ERR_ARG;
        }
        TypeDev &d = T.t[t];
        d.size = r.size; d.rmax = r.r[r.size - 1];
        d.zion = r.zion; d.vloc0 = r.vloc0;
        d.nlcc = r.rhoc && r.rhoc_d ? 1 : 0;
        d.r = up(r.r, r.size);
        d.rvloc = up(r.rvloc, r.size);
        d.rvloc_d = up(r.rvloc_d, r.size);
        d.rhoc = d.nlcc ? up(r.rhoc, r.size) : nullptr;
        d.rhoc_d = d.nlcc ? up(r.rhoc_d, r.size) : nullptr;
    }
    double *dphi = nullptr, *dvc = nullptr, *tmp = nullptr, *vals = nullptr, *part = nullptr;
    Img *dimg = nullptr;
    const int bpi = 64;   // blocks per image
    if (st == SGPU_OK &&
        (hipMalloc((void **)&tmp, 4 * (size_t)nd * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&dphi, 3 * (size_t)nd * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&dvc, 3 * (size_t)nd * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&vals, 3 * (size_t)total * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&part, (size_t)n_img * bpi * 6 * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&dimg, (size_t)n_img * sizeof(Img)) != hipSuccess))
        st = SGPU_ERR_HIP;
    std::vector<double> hp;
    if (st == SGPU_OK) {
        double *dph = tmp, *dv = tmp + nd, *db = tmp + 2 * nd, *dx = tmp + 3 * nd;
        const unsigned nb = (unsigned)((nd + 255) / 256);
        if (hipMemcpyAsync(dph, phi, nd * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dv, vc, nd * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(db, bsum, nd * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dx, vxc, nd * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dimg, im.data(), n_img * sizeof(Img), hipMemcpyHostToDevice, s) != hipSuccess)
            st = SGPU_ERR_HIP;
        if (st == SGPU_OK) {
            grad3<<<nb, 256, 0, s>>>(dph, ctx->nx, ctx->ny, ctx->nz, F, c, dphi, dphi + nd, dphi + 2 * nd);
            grad3<<<nb, 256, 0, s>>>(dv, ctx->nx, ctx->ny, ctx->nz, F, c, dvc, dvc + nd, dvc + 2 * nd);
            box_eval<<<dim3(64, n_img), 256, 0, s>>>(dimg, n_img, T, ctx->hx, ctx->hy, ctx->hz,
                                                     ref_cutoff, vals, vals + total, vals + 2 * total);
            box_force<<<n_img * bpi, 256, 0, s>>>(dimg, n_img, F, c, ctx->nx, ctx->ny, vals,
                                                  vals + total, vals + 2 * total, dphi, dvc, db, dx,
                                                  nd, bpi, part);
            if (hipGetLastError() != hipSuccess) st = SGPU_ERR_HIP;
        }
        if (st == SGPU_OK) {
            hp.resize((size_t)n_img * bpi * 6);
            if (hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                st = SGPU_ERR_HIP;
        }
    }
    (void)hipStreamSynchronize(s);
    for (double *p : dev) (void)hipFree(p);
    for (void *p : {(void *)tmp, (void *)dphi, (void *)dvc, (void *)vals, (void *)part, (void *)dimg})
        if (p) (void)hipFree(p);
    if (st != SGPU_OK) {
        sgpu_set_error("sgpu_local_forces: HIP error");
        return st;
    }
    // per image in block order, per atom in image order
    const double dV = ctx->hx * ctx->hy * ctx->hz;
    for (int k = 0; k < n_img; k++) {
        double sum[6] = {0, 0, 0, 0, 0, 0};
        for (int b = 0; b < bpi; b++)
            for (int q = 0; q < 6; q++) sum[q] += hp[((size_t)k * bpi + b) * 6 + q];
        for (int d = 0; d < 3; d++) {
            f_loc[3 * im[k].atom + d] += sum[d] * dV;
            f_xc[3 * im[k].atom + d] += sum[3 + d] * dV;
        }
    }
    return SGPU_OK;
}
