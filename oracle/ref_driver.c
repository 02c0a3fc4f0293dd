/*
 * Thin driver around the reference's own stencil code, compiled from the
 * reference sources where they lie (see oracle/Makefile, target `ref`).
 * TEST INFRASTRUCTURE ONLY: it lets tests compare the oracle's
 * restatements with SPARC's Lap_plus_diag_vec_mult_orth
 * (src/lapVecRoutines.c:306-611) and Gradient_vec_dir
 * (src/gradVecRoutines.c:61-316) on the same inputs.  It is built only where
 * /root/reference is present and never linked into the product.
 */
#include <stdlib.h>
#include <string.h>
#include <mpi.h>
#include "isddft.h"
#include "lapVecRoutines.h"
#include "gradVecRoutines.h"

/* y = (a Lap + b diag(v) + c) x for ncol columns on one rank owning the whole
 * grid (dims = 1,1,1 selects the single-rank halo branch, :536-577). */
int ref_lap_plus_diag(int nx, int ny, int nz, int bcx, int bcy, int bcz, int order,
                      const double *d2x, const double *d2y, const double *d2z,
                      int ncol, double a, double b, double c, const double *v,
                      const double *x, int ldi, double *y, int ldo)
{
    SPARC_OBJ *s = (SPARC_OBJ *)calloc(1, sizeof(SPARC_OBJ));
    if (!s) return -1;
    int fdn = order / 2;
    s->order = order;
    s->BCx = bcx; s->BCy = bcy; s->BCz = bcz;
    s->D2_stencil_coeffs_x = (double *)malloc(sizeof(double) * (fdn + 1));
    s->D2_stencil_coeffs_y = (double *)malloc(sizeof(double) * (fdn + 1));
    s->D2_stencil_coeffs_z = (double *)malloc(sizeof(double) * (fdn + 1));
    memcpy(s->D2_stencil_coeffs_x, d2x, sizeof(double) * (fdn + 1));
    memcpy(s->D2_stencil_coeffs_y, d2y, sizeof(double) * (fdn + 1));
    memcpy(s->D2_stencil_coeffs_z, d2z, sizeof(double) * (fdn + 1));
    int DMVertices[6] = {0, nx - 1, 0, ny - 1, 0, nz - 1};
    int dims[3] = {1, 1, 1};
    int DMnd = nx * ny * nz;
    for (int n = 0; n < ncol; n++)
        Lap_plus_diag_vec_mult_orth(s, DMnd, DMVertices, 1, a, b, c, v,
                                    x + (long)n * ldi, ldi, y + (long)n * ldo, ldo,
                                    MPI_COMM_SELF, dims);
    free(s->D2_stencil_coeffs_x); free(s->D2_stencil_coeffs_y); free(s->D2_stencil_coeffs_z);
    free(s);
    return 0;
}

/* Dx = d/dx_dir x for ncol columns (Gradient_vec_dir with c = 0 on one rank
 * owning the whole grid, dims = 1,1,1). */
int ref_gradient(int nx, int ny, int nz, int bcx, int bcy, int bcz, int order,
                 const double *d1x, const double *d1y, const double *d1z, int ncol, int dir,
                 const double *x, int ldi, double *y, int ldo)
{
    SPARC_OBJ *s = (SPARC_OBJ *)calloc(1, sizeof(SPARC_OBJ));
    if (!s) return -1;
    int fdn = order / 2;
    s->order = order;
    s->BCx = bcx; s->BCy = bcy; s->BCz = bcz;
    s->D1_stencil_coeffs_x = (double *)malloc(sizeof(double) * (fdn + 1));
    s->D1_stencil_coeffs_y = (double *)malloc(sizeof(double) * (fdn + 1));
    s->D1_stencil_coeffs_z = (double *)malloc(sizeof(double) * (fdn + 1));
    memcpy(s->D1_stencil_coeffs_x, d1x, sizeof(double) * (fdn + 1));
    memcpy(s->D1_stencil_coeffs_y, d1y, sizeof(double) * (fdn + 1));
    memcpy(s->D1_stencil_coeffs_z, d1z, sizeof(double) * (fdn + 1));
    int DMVertices[6] = {0, nx - 1, 0, ny - 1, 0, nz - 1};
    int dims[3] = {1, 1, 1};
    int DMnd = nx * ny * nz;
    for (int n = 0; n < ncol; n++)
        Gradient_vec_dir(s, DMnd, DMVertices, 1, 0.0, x + (long)n * ldi, ldi, y + (long)n * ldo, ldo,
                         dir, MPI_COMM_SELF, dims);
    free(s->D1_stencil_coeffs_x); free(s->D1_stencil_coeffs_y); free(s->D1_stencil_coeffs_z);
    free(s);
    return 0;
}
