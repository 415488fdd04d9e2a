/*
 * adapt.c -- oracle: ghost-node synchronisation on non-uniform (2:1
 * balanced, leaf) forests, interpolating prediction and block refinement.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The ghost fill here works by
 * point location, independently of the device's patch tables: every ghost
 * point of every block (within the requested depth) looks up the leaf that
 * covers its position and takes
 *   - the leaf's value at that point (same level: copy; finer leaf: the
 *     coinciding fine point = restriction by decimation,
 *     restrict_predict_data.f90:90-98), in a first pass that reads interiors
 *     only (stages 1-2 of sync_ghosts_generic, synchronize_ghosts_generic.f90:
 *     235-239), or
 *   - for a coarser leaf, the value of that leaf's interpolated grid at the
 *     point (predict_data -> prediction, restrict_predict_data.f90:174-202,
 *     module_wavelets.f90:96-284), computed by the literal separable passes
 *     over the coarse block including the ghosts filled in the first pass
 *     (stage 3).
 * Pinning: tests/test_oracle_adapt.py -- the reference's --refine-everywhere
 * outputs (TESTING/wavelets/adaptive_CDF{20,40,60}/vor_00100.h5), which run
 * sync_ghosts_tree on a 112-block non-uniform grid, refineBlock, and the sync
 * before saving; plus exactness of prediction on polynomials of degree < order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wabbit_oracle.h"

static void pred_coefs(int order, double *c)
{
    /* module_wavelets.f90:121-145 */
    if (order == 2) {
        c[0] = 0.5; c[1] = 0.5;
    } else if (order == 4) {
        c[0] = -1.0 / 16.0; c[1] = 9.0 / 16.0; c[2] = 9.0 / 16.0; c[3] = -1.0 / 16.0;
    } else {
        c[0] = 3.0 / 256.0; c[1] = -25.0 / 256.0; c[2] = 150.0 / 256.0;
        c[3] = 150.0 / 256.0; c[4] = -25.0 / 256.0; c[5] = 3.0 / 256.0;
    }
}

/* `prediction` as whole-array passes: coarse (nx,ny,nz) -> fine
 * (2nx-1, 2ny-1, 2nz-1) (nz = 1 in 2D).  Index (i,j,k) at i + nx*(j + ny*k). */
void or_prediction(const double *coarse, int nx, int ny, int nz, double *fine, int order)
{
    or_prediction_zorder(coarse, nx, ny, nz, fine, order, order);
}

/* the same with a z pass of order zorder (zorder != order only as a planted
 * mistake in tests) */
void or_prediction_zorder(const double *coarse, int nx, int ny, int nz, double *fine, int order, int zorder)
{
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = nz > 1 ? 2 * nz - 1 : 1;
    int nc = order, N = order / 2;
    double c[6];
    pred_coefs(order, c);
#define F(i, j, k) fine[(i) + (size_t)fx * ((j) + (size_t)fy * (k))]
    for (size_t q = 0; q < (size_t)fx * fy * fz; ++q) fine[q] = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                F(2 * i, 2 * j, nz > 1 ? 2 * k : 0) = coarse[i + (size_t)nx * (j + (size_t)ny * k)];
    /* 1-based loops of the reference: ixfine = 2N, nxfine-(2N-1), 2 -> 0-based 2N-1 .. fx-2N */
    for (int k = 0; k < fz; k += (nz > 1 ? 2 : 1)) {
        for (int j = 0; j < fy; j += 2)
            for (int i = 2 * N - 1; i <= fx - 2 * N; i += 2) {
                double s = c[0] * F(i - 2 * N + 1, j, k);
                for (int t = 1; t < nc; ++t) s = s + c[t] * F(i - 2 * N + 1 + 2 * t, j, k);
                F(i, j, k) = s;
            }
        for (int j = 2 * N - 1; j <= fy - 2 * N; j += 2)
            for (int i = 0; i < fx; ++i) {
                double s = c[0] * F(i, j - 2 * N + 1, k);
                for (int t = 1; t < nc; ++t) s = s + c[t] * F(i, j - 2 * N + 1 + 2 * t, k);
                F(i, j, k) = s;
            }
    }
    pred_coefs(zorder, c);
    nc = zorder;
    N = zorder / 2;
    if (nz > 1)
        for (int k = 2 * N - 1; k <= fz - 2 * N; k += 2)
            for (int j = 0; j < fy; ++j)
                for (int i = 0; i < fx; ++i) {
                    double s = c[0] * F(i, j, k - 2 * N + 1);
                    for (int t = 1; t < nc; ++t) s = s + c[t] * F(i, j, k - 2 * N + 1 + 2 * t);
                    F(i, j, k) = s;
                }
#undef F
}

/* ---- forest lookup ------------------------------------------------------ */

typedef struct {
    const or_forest *f;
    int **by_level; /* per level: dense array over (2^l)^dim cells -> block or -1 */
} lookup_t;

static void ijk_of(const or_forest *f, int64_t tc, int level, int ijk[3])
{
    ijk[0] = ijk[1] = ijk[2] = 0;
    for (int l = 1; l <= level; ++l) {
        const int d = (int)((tc >> ((f->jmax - l) * f->dim)) & ((1 << f->dim) - 1));
        ijk[0] = 2 * ijk[0] + ((d >> 1) & 1);
        ijk[1] = 2 * ijk[1] + (d & 1);
        ijk[2] = 2 * ijk[2] + (f->dim == 3 ? ((d >> 2) & 1) : 0);
    }
}

static int lookup_build(lookup_t *L, const or_forest *f)
{
    L->f = f;
    L->by_level = (int **)calloc((size_t)f->jmax + 1, sizeof(int *));
    for (int l = 0; l <= f->jmax; ++l) {
        const size_t n = (size_t)1 << (l * f->dim);
        L->by_level[l] = (int *)malloc(n * sizeof(int));
        for (size_t q = 0; q < n; ++q) L->by_level[l][q] = -1;
    }
    for (int b = 0; b < f->n; ++b) {
        int ijk[3];
        ijk_of(f, f->tc[b], f->level[b], ijk);
        const size_t m = (size_t)1 << f->level[b];
        const size_t q = ijk[0] + m * (ijk[1] + (f->dim == 3 ? m * ijk[2] : 0));
        if (L->by_level[f->level[b]][q] != -1) return -1;
        L->by_level[f->level[b]][q] = b;
    }
    return 0;
}

static void lookup_free(lookup_t *L)
{
    for (int l = 0; l <= L->f->jmax; ++l) free(L->by_level[l]);
    free(L->by_level);
}

/* leaf covering the point X (integer coordinates at level jmax+1, periodic);
 * returns block index or -1; *lev gets its level */
static int cover(const lookup_t *L, const int64_t X[3], int *lev)
{
    const or_forest *f = L->f;
    for (int l = 0; l <= f->jmax; ++l) {
        const int64_t cell = (int64_t)1 << (f->jmax + 1 - l); /* point spacing at level l */
        const int64_t m = (int64_t)1 << l;
        size_t q = 0, mul = 1;
        for (int d = 0; d < f->dim; ++d) {
            const int64_t bi = X[d] / (cell * f->bs[d]);
            q += (size_t)(bi % m) * mul;
            mul *= (size_t)m;
        }
        const int b = L->by_level[l][q];
        if (b >= 0) {
            *lev = l;
            return b;
        }
    }
    return -1;
}

int or_sync_leaf(const or_forest *f, double *hvy, int neq, int gm, int gp, int order)
{
    const int dim = f->dim, g = f->g;
    const int nx = f->bs[0] + 2 * g, ny = f->bs[1] + 2 * g, nz = dim == 3 ? f->bs[2] + 2 * g : 1;
    const size_t cs = (size_t)nx * ny * nz, bsz = cs * neq;
    lookup_t L;
    if (lookup_build(&L, f)) return -1;
    int64_t ext[3];
    for (int d = 0; d < 3; ++d)
        ext[d] = d < dim ? ((int64_t)f->bs[d] << (f->jmax + 1)) : 1;
    double *snap = (double *)malloc(sizeof(double) * bsz * f->n);
    memcpy(snap, hvy, sizeof(double) * bsz * f->n);
    /* which coarse blocks are prediction senders: computed lazily */
    const int fxn = 2 * nx - 1, fyn = 2 * ny - 1, fzn = dim == 3 ? 2 * nz - 1 : 1;
    double **pred = (double **)calloc((size_t)f->n, sizeof(double *));
    int rc = 0;
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        if (pass == 1) memcpy(snap, hvy, sizeof(double) * bsz * f->n);
        for (int b = 0; b < f->n && !rc; ++b) {
            int ijk[3];
            ijk_of(f, f->tc[b], f->level[b], ijk);
            const int J = f->level[b];
            const int64_t sp = (int64_t)1 << (f->jmax + 1 - J);
            const int zl = dim == 3 ? g - gm : 0, zh = dim == 3 ? g + f->bs[2] + gp : 1;
            for (int iz = zl; iz < zh && !rc; ++iz)
                for (int iy = g - gm; iy < g + f->bs[1] + gp && !rc; ++iy)
                    for (int ix = g - gm; ix < g + f->bs[0] + gp && !rc; ++ix) {
                        const int ii[3] = {ix, iy, iz};
                        int inner = 1;
                        for (int d = 0; d < dim; ++d)
                            if (ii[d] < g || ii[d] >= g + f->bs[d]) inner = 0;
                        if (inner) continue;
                        int64_t X[3] = {0, 0, 0};
                        for (int d = 0; d < dim; ++d) {
                            X[d] = ((int64_t)ijk[d] * f->bs[d] + (ii[d] - g)) * sp;
                            X[d] = ((X[d] % ext[d]) + ext[d]) % ext[d];
                        }
                        int lc;
                        const int c = cover(&L, X, &lc);
                        if (c < 0) continue; /* outside a non-periodic domain */
                        if (lc < J - 1 || lc > J + 1) {
                            rc = -2; /* not 2:1 balanced */
                            break;
                        }
                        int cij[3];
                        ijk_of(f, f->tc[c], lc, cij);
                        const int64_t spc = (int64_t)1 << (f->jmax + 1 - lc);
                        const size_t me = (size_t)b * bsz +
                                          ((size_t)iz * ny + iy) * nx + ix;
                        if (lc >= J && pass == 0) {
                            int li[3] = {0, 0, 0};
                            for (int d = 0; d < dim; ++d)
                                li[d] = (int)((X[d] - (int64_t)cij[d] * f->bs[d] * spc) / spc) + g;
                            const size_t src = (size_t)c * bsz + ((size_t)li[2] * ny + li[1]) * nx + li[0];
                            for (int q = 0; q < neq; ++q) hvy[me + q * cs] = snap[src + q * cs];
                        } else if (lc == J - 1 && pass == 1) {
                            if (!pred[c]) {
                                pred[c] = (double *)malloc(sizeof(double) * (size_t)fxn * fyn * fzn * neq);
                                for (int q = 0; q < neq; ++q)
                                    or_prediction(snap + (size_t)c * bsz + q * cs, nx, ny, nz,
                                                  pred[c] + (size_t)q * fxn * fyn * fzn, order);
                            }
                            /* index in the coarse block's doubled ghost-inclusive space */
                            int fi[3] = {0, 0, 0};
                            for (int d = 0; d < dim; ++d) {
                                fi[d] = (int)((X[d] - (int64_t)cij[d] * f->bs[d] * spc) / sp) + 2 * g;
                                const int lim = d == 0 ? fxn : d == 1 ? fyn : fzn;
                                if (fi[d] < 0 || fi[d] >= lim) rc = -3;
                            }
                            if (rc) break;
                            const size_t src = fi[0] + (size_t)fxn * (fi[1] + (size_t)fyn * fi[2]);
                            for (int q = 0; q < neq; ++q)
                                hvy[me + q * cs] = pred[c][(size_t)q * fxn * fyn * fzn + src];
                        }
                    }
        }
    }
    for (int b = 0; b < f->n; ++b) free(pred[b]);
    free(pred);
    free(snap);
    lookup_free(&L);
    return rc;
}

void or_refine_block(int dim, const int bs[3], int g, int neq, const double *mother,
                     double *const *daughters, int order)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = dim == 3 ? 2 * nz - 1 : 1;
    const size_t cs = (size_t)nx * ny * nz;
    double *fine = (double *)malloc(sizeof(double) * (size_t)fx * fy * fz);
    for (int q = 0; q < neq; ++q) {
        or_prediction(mother + q * cs, nx, ny, nz, fine, order);
        /* refinementExecute.f90:90-104: daughter k takes fine[g + Bs*off ...] */
        for (int k = 0; k < (1 << dim); ++k) {
            const int off[3] = {(k >> 1) & 1, k & 1, (k >> 2) & 1};
            double *d = daughters[k] + q * cs;
            for (int iz = 0; iz < nz; ++iz)
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix) {
                        const int fi = g + bs[0] * off[0] + ix, fj = g + bs[1] * off[1] + iy;
                        const int fk = dim == 3 ? g + bs[2] * off[2] + iz : 0;
                        d[((size_t)iz * ny + iy) * nx + ix] =
                            fine[fi + (size_t)fx * (fj + (size_t)fy * fk)];
                    }
        }
    }
    free(fine);
}
