/*
 * acm.c -- oracle: ACM right-hand side, per-block dt, RK generic step and
 * same-level periodic sync for an equidistant grid.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see wabbit_oracle.h for the list of
 * checks that pin each function.  Compiled with -ffp-contract=off so every
 * product and sum rounds separately, in the reference's left-to-right order.
 */
#include "wabbit_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>

/* FD coefficient tables, rhs_ACM.f90:335-344 (2D) and :958-965 (3D). */
static const double A_FD4[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
static const double B_FD4[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
static const double A_TW4[7] = {-0.02651995, 0.18941314, -0.79926643, 0.0,
                                0.79926643, -0.18941314, 0.02651995};
static const double A_FD6[7] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0,
                                3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
static const double B_FD6[7] = {1.0 / 90.0, -3.0 / 20.0, 3.0 / 2.0, -49.0 / 18.0,
                                3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};

/* first derivative of f along stride s (rhs_ACM.f90:383-384 FD2, :540-542 FD4,
 * :460 TW4, :698 FD6); terms summed from the most negative offset up. */
static double d1(int scheme, const double *f, ptrdiff_t s, double inv, int plant)
{
    if ((plant & 2) && scheme == OR_FD6)
        return (A_FD6[0] * f[-3 * s] + A_FD6[1] * f[-2 * s] + A_FD6[2] * f[-s] +
                A_FD6[4] * f[s] + A_FD6[5] * f[2 * s]) * inv;
    if ((plant & 4) && scheme == OR_FD4_TW)
        return (A_TW4[0] * f[-3 * s] + A_TW4[1] * f[-2 * s] + A_TW4[2] * f[-s] +
                A_TW4[4] * f[s] + A_TW4[5] * f[2 * s] - A_TW4[6] * f[3 * s]) * inv;
    switch (scheme) {
    case OR_FD2:
        return (f[s] - f[-s]) * inv * 0.5;
    case OR_FD4:
        return (A_FD4[0] * f[-2 * s] + A_FD4[1] * f[-s] + A_FD4[3] * f[s] +
                A_FD4[4] * f[2 * s]) * inv;
    case OR_FD4_TW:
        return (A_TW4[0] * f[-3 * s] + A_TW4[1] * f[-2 * s] + A_TW4[2] * f[-s] +
                A_TW4[4] * f[s] + A_TW4[5] * f[2 * s] + A_TW4[6] * f[3 * s]) * inv;
    default: /* OR_FD6 */
        return (A_FD6[0] * f[-3 * s] + A_FD6[1] * f[-2 * s] + A_FD6[2] * f[-s] +
                A_FD6[4] * f[s] + A_FD6[5] * f[2 * s] + A_FD6[6] * f[3 * s]) * inv;
    }
}

/* first derivative of the product fa*fb (skew-symmetric terms uu_dx, uv_dy,
 * ...: rhs_ACM.f90:389-393, :545-551); each term is (coef*fa)*fb. */
static double d1p(int scheme, const double *fa, const double *fb, ptrdiff_t s, double inv)
{
    switch (scheme) {
    case OR_FD2:
        return (fa[s] * fb[s] - fa[-s] * fb[-s]) * inv * 0.5;
    case OR_FD4:
        return (A_FD4[0] * fa[-2 * s] * fb[-2 * s] + A_FD4[1] * fa[-s] * fb[-s] +
                A_FD4[3] * fa[s] * fb[s] + A_FD4[4] * fa[2 * s] * fb[2 * s]) * inv;
    case OR_FD4_TW:
        return (A_TW4[0] * fa[-3 * s] * fb[-3 * s] + A_TW4[1] * fa[-2 * s] * fb[-2 * s] +
                A_TW4[2] * fa[-s] * fb[-s] + A_TW4[4] * fa[s] * fb[s] +
                A_TW4[5] * fa[2 * s] * fb[2 * s] + A_TW4[6] * fa[3 * s] * fb[3 * s]) * inv;
    default:
        return (A_FD6[0] * fa[-3 * s] * fb[-3 * s] + A_FD6[1] * fa[-2 * s] * fb[-2 * s] +
                A_FD6[2] * fa[-s] * fb[-s] + A_FD6[4] * fa[s] * fb[s] +
                A_FD6[5] * fa[2 * s] * fb[2 * s] + A_FD6[6] * fa[3 * s] * fb[3 * s]) * inv;
    }
}

/* second derivative (rhs_ACM.f90:395-399 FD2, :554-558 FD4 -- also used by
 * the Tam-Webb scheme --, :712-716 FD6). */
static double d2(int scheme, const double *f, ptrdiff_t s, double inv2, int plant)
{
    switch (scheme) {
    case OR_FD2:
        if (plant & 8) return (f[-s] - 1.0 * f[0] + f[s]) * inv2;
        return (f[-s] - 2.0 * f[0] + f[s]) * inv2;
    case OR_FD4:
    case OR_FD4_TW:
        return (B_FD4[0] * f[-2 * s] + B_FD4[1] * f[-s] + B_FD4[2] * f[0] +
                B_FD4[3] * f[s] + B_FD4[4] * f[2 * s]) * inv2;
    default:
        return (B_FD6[0] * f[-3 * s] + B_FD6[1] * f[-2 * s] + B_FD6[2] * f[-s] +
                B_FD6[3] * f[0] + B_FD6[4] * f[s] + B_FD6[5] * f[2 * s] +
                B_FD6[6] * f[3 * s]) * inv2;
    }
}

static void dims_of(int dim, const int Bs[3], int g, int *nx, int *ny, int *nz)
{
    *nx = Bs[0] + 2 * g;
    *ny = Bs[1] + 2 * g;
    *nz = (dim == 3) ? Bs[2] + 2 * g : 1;
}

void or_rhs_acm_block(const or_acm_params *p, const int Bs[3], int g,
                      const double dx[3], const double *u, double *rhs,
                      const double *mask)
{
    const int dim = p->dim;
    int nx, ny, nz;
    dims_of(dim, Bs, g, &nx, &ny, &nz);
    const ptrdiff_t s[3] = {1, nx, (ptrdiff_t)nx * ny};
    const ptrdiff_t cs = (ptrdiff_t)nx * ny * nz; /* component stride */
    const int np = dim;                           /* index of p */
    double inv[3], inv2[3];
    for (int d = 0; d < 3; ++d) {
        inv[d] = 1.0 / dx[d];
        inv2[d] = 1.0 / (dx[d] * dx[d]);
    }
    const double c02 = p->c_0 * p->c_0;
    const double Cinv = 1.0 / p->C_eta; /* C_eta_apply_inv(color>0), rhs_ACM.f90:353-355 */
    const int z0 = (dim == 3) ? g : 0, z1 = (dim == 3) ? g + Bs[2] : 1;
    const int sc = p->discretization;

    for (int iz = z0; iz < z1; ++iz)
        for (int iy = g; iy < g + Bs[1]; ++iy)
            for (int ix = g; ix < g + Bs[0]; ++ix) {
                const ptrdiff_t i = ((ptrdiff_t)iz * ny + iy) * nx + ix;
                double du[4][3]; /* du[c][d] = d(u_c)/dx_d */
                double dd[3][3]; /* second derivatives of velocities */
                double uc[4];
                for (int c = 0; c <= dim; ++c) {
                    uc[c] = u[c * cs + i];
                    for (int d = 0; d < dim; ++d)
                        du[c][d] = d1(sc, u + c * cs + i, s[d], inv[d], p->plant);
                }
                for (int c = 0; c < dim; ++c)
                    for (int d = 0; d < dim; ++d)
                        dd[c][d] = d2(sc, u + c * cs + i, s[d], inv2[d], p->plant);

                for (int c = 0; c < dim; ++c) {
                    double penal = 0.0;
                    int has_penal = 0;
                    if (p->penalization && mask) {
                        const int color = (int)mask[4 * cs + i];
                        const double ci = (color == 0) ? 0.0 : Cinv;
                        const double chi = mask[0 * cs + i] * ci;
                        penal = -(chi * (uc[c] - mask[(1 + c) * cs + i]));
                        has_penal = 1;
                    }
                    double lap = dd[c][0];
                    for (int d = 1; d < dim; ++d) lap = lap + dd[c][d];
                    double t;
                    if (p->skew_symmetry) {
                        /* -0.5*(sum_d d(u_c u_d)/dx_d + sum_d u_d du_c/dx_d) */
                        double sum = d1p(sc, u + c * cs + i, u + 0 * cs + i, s[0], inv[0]);
                        for (int d = 1; d < dim; ++d)
                            sum = sum + d1p(sc, u + c * cs + i, u + d * cs + i, s[d], inv[d]);
                        for (int d = 0; d < dim; ++d) sum = sum + uc[d] * du[c][d];
                        t = -(0.5 * sum);
                    } else {
                        /* plant bit 0 swaps the first two derivative directions */
                        const int sw = (p->plant & 1) ? 1 : 0;
                        t = -(uc[0] * du[c][sw]);
                        for (int d = 1; d < dim; ++d) t = t - uc[d] * du[c][(sw && d == 1) ? 0 : d];
                    }
                    t = t - du[np][c];
                    t = t + p->nu * lap;
                    if (has_penal) t = t + penal;
                    rhs[c * cs + i] = t;
                }
                double div = du[0][0];
                for (int d = 1; d < dim; ++d) div = div + du[d][d];
                rhs[np * cs + i] = -(c02 * div) - p->gamma_p * uc[np];

                if (p->use_sponge && mask) {
                    /* rhs_ACM.f90:872-892 (2D), :1732-1750 (3D) */
                    const double spo = mask[5 * cs + i] * (1.0 / p->C_sponge);
                    for (int c = 0; c < dim; ++c)
                        rhs[c * cs + i] = rhs[c * cs + i] - (uc[c] - p->u_mean_set[c]) * spo;
                    rhs[np * cs + i] = rhs[np * cs + i] - uc[np] * spo;
                }
            }
}

double or_dt_block(const or_acm_params *p, const int Bs[3], int g,
                   const double dx[3], const double *u)
{
    const int dim = p->dim;
    int nx, ny, nz;
    dims_of(dim, Bs, g, &nx, &ny, &nz);
    const ptrdiff_t cs = (ptrdiff_t)nx * ny * nz;
    const int z0 = (dim == 3) ? g : 0, z1 = (dim == 3) ? g + Bs[2] : 1;
    double umag = -1.0;
    for (int iz = z0; iz < z1; ++iz)
        for (int iy = g; iy < g + Bs[1]; ++iy)
            for (int ix = g; ix < g + Bs[0]; ++ix) {
                const ptrdiff_t i = ((ptrdiff_t)iz * ny + iy) * nx + ix;
                double m = u[i] * u[i] + u[cs + i] * u[cs + i];
                if (dim == 3) m = m + u[2 * cs + i] * u[2 * cs + i];
                if (m > umag) umag = m;
            }
    double dxmin = dx[0];
    for (int d = 1; d < dim; ++d)
        if (dx[d] < dxmin) dxmin = dx[d];
    const double ueig = sqrt(umag) + sqrt(p->c_0 * p->c_0 + umag);
    double dt = (ueig >= 1.0e-6) ? p->CFL * dxmin / ueig : 1.0e-2;
    if (p->nu > 1.0e-13) {
        const double t = p->CFL_nu * (dxmin * dxmin) / p->nu;
        if (t < dt) dt = t;
    }
    if (p->gamma_p > 0.0) {
        const double t = p->CFL_eta * p->gamma_p;
        if (t < dt) dt = t;
    }
    if (p->penalization) {
        const double t = p->CFL_eta * p->C_eta;
        if (t < dt) dt = t;
    }
    if (p->use_sponge) {
        const double t = p->CFL_eta * p->C_sponge;
        if (t < dt) dt = t;
    }
    return dt;
}

double or_maxabs_block(const int Bs[3], int g, int neq, const double *u)
{
    const int dim = Bs[2] > 1 ? 3 : 2;
    int nx, ny, nz;
    dims_of(dim, Bs, g, &nx, &ny, &nz);
    const ptrdiff_t cs = (ptrdiff_t)nx * ny * nz;
    const int z0 = (dim == 3) ? g : 0, z1 = (dim == 3) ? g + Bs[2] : 1;
    double m = 0.0;
    for (int c = 0; c < neq; ++c)
        for (int iz = z0; iz < z1; ++iz)
            for (int iy = g; iy < g + Bs[1]; ++iy)
                for (int ix = g; ix < g + Bs[0]; ++ix) {
                    const double a = fabs(u[c * cs + ((ptrdiff_t)iz * ny + iy) * nx + ix]);
                    if (a > m) m = a;
                }
    return m;
}

static int wrap(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

void or_sync_equidistant(int dim, const int Bs[3], int g, int neq, int nb,
                         const int *block_at, int nblocks, double *hvy, int gs)
{
    int nx, ny, nz;
    dims_of(dim, Bs, g, &nx, &ny, &nz);
    const ptrdiff_t cs = (ptrdiff_t)nx * ny * nz;
    const ptrdiff_t bsz = cs * neq;
    const int nbz = (dim == 3) ? nb : 1;
    const int ntot[3] = {nb * Bs[0], nb * Bs[1], nb * Bs[2]};
    (void)nblocks;
#pragma omp parallel for collapse(2) schedule(static)
    for (int bk = 0; bk < nbz; ++bk)
        for (int bj = 0; bj < nb; ++bj)
            for (int bi = 0; bi < nb; ++bi) {
                const int me = block_at[((ptrdiff_t)bk * nb + bj) * nb + bi];
                const int lo = g - gs, hi = g + gs; /* box [lo, Bs+hi) */
                const int zl = (dim == 3) ? lo : 0, zh = (dim == 3) ? Bs[2] + hi : 1;
                for (int iz = zl; iz < zh; ++iz)
                    for (int iy = lo; iy < Bs[1] + hi; ++iy)
                        for (int ix = lo; ix < Bs[0] + hi; ++ix) {
                            const int inx = ix >= g && ix < g + Bs[0];
                            const int iny = iy >= g && iy < g + Bs[1];
                            const int inz = dim == 2 || (iz >= g && iz < g + Bs[2]);
                            if (inx && iny && inz) continue;
                            const int G0 = wrap(bi * Bs[0] + ix - g, ntot[0]);
                            const int G1 = wrap(bj * Bs[1] + iy - g, ntot[1]);
                            const int G2 = (dim == 3) ? wrap(bk * Bs[2] + iz - g, ntot[2]) : 0;
                            const int src = block_at[((ptrdiff_t)(G2 / (dim == 3 ? Bs[2] : 1)) * nb +
                                                      G1 / Bs[1]) * nb + G0 / Bs[0]];
                            const int sx = G0 % Bs[0] + g, sy = G1 % Bs[1] + g;
                            const int sz = (dim == 3) ? G2 % Bs[2] + g : 0;
                            for (int c = 0; c < neq; ++c)
                                hvy[me * bsz + c * cs + ((ptrdiff_t)iz * ny + iy) * nx + ix] =
                                    hvy[src * bsz + c * cs + ((ptrdiff_t)sz * ny + sy) * nx + sx];
                        }
            }
}

double or_dt_min(const or_acm_params *p, const int Bs[3], int g, int neq,
                 int nblocks, const double *dxs, const double *hvy)
{
    int nx, ny, nz;
    dims_of(p->dim, Bs, g, &nx, &ny, &nz);
    const ptrdiff_t bsz = (ptrdiff_t)nx * ny * nz * neq;
    double dt = 9.0e9;
    for (int b = 0; b < nblocks; ++b) {
        const double t = or_dt_block(p, Bs, g, dxs + 3 * b, hvy + b * bsz);
        if (t < dt) dt = t;
    }
    return dt;
}

/* interior-only axpy helpers of RungeKuttaGeneric */
static void interior_loop_bounds(int dim, const int Bs[3], int g, int *z0, int *z1)
{
    *z0 = (dim == 3) ? g : 0;
    *z1 = (dim == 3) ? g + Bs[2] : 1;
}

void or_rk_step_equidistant(const or_acm_params *p, const int Bs[3], int g,
                            int g_rhs, int neq, int nb, const int *block_at,
                            int nblocks, const double *dxs, double *hvy,
                            double *work, const double *mask,
                            const double *butcher, int nrows, double dt)
{
    const int dim = p->dim;
    int nx, ny, nz, z0, z1;
    dims_of(dim, Bs, g, &nx, &ny, &nz);
    interior_loop_bounds(dim, Bs, g, &z0, &z1);
    const ptrdiff_t cs = (ptrdiff_t)nx * ny * nz;
    const ptrdiff_t bsz = cs * neq;
    const ptrdiff_t slot = bsz * nblocks;
    const ptrdiff_t msz = cs * 6;
#define BT(r, c) butcher[(r) * nrows + (c)]
#define FOR_INTERIOR(body)                                                       \
    for (int c = 0; c < neq; ++c)                                                \
        for (int iz = z0; iz < z1; ++iz)                                         \
            for (int iy = g; iy < g + Bs[1]; ++iy)                               \
                for (int ix = g; ix < g + Bs[0]; ++ix) {                         \
                    const ptrdiff_t i = c * cs + ((ptrdiff_t)iz * ny + iy) * nx + ix; \
                    body;                                                        \
                }

    /* runge_kutta_generic.f90:52 sync, :63-67 u0 copy, :72 first RHS */
    or_sync_equidistant(dim, Bs, g, neq, nb, block_at, nblocks, hvy, g_rhs);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nblocks; ++b) {
        double *u = hvy + b * bsz, *w1 = work + b * bsz;
        FOR_INTERIOR(w1[i] = u[i]);
        or_rhs_acm_block(p, Bs, g, dxs + 3 * b, u, work + slot + b * bsz,
                         mask ? mask + b * msz : NULL);
    }
    /* :80-127 stages j = 2..nrows-1 */
    for (int j = 2; j <= nrows - 1; ++j) {
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nblocks; ++b) {
            double *u = hvy + b * bsz, *w1 = work + b * bsz;
            FOR_INTERIOR(u[i] = w1[i]);
            for (int l = 2; l <= j; ++l) {
                const double a = BT(j - 1, l - 1);
                if (fabs(a) < 1.0e-8) continue;
                const double *k = work + (ptrdiff_t)(l - 1) * slot + b * bsz;
                const double da = dt * a;
                FOR_INTERIOR(u[i] = u[i] + da * k[i]);
            }
        }
        or_sync_equidistant(dim, Bs, g, neq, nb, block_at, nblocks, hvy, g_rhs);
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nblocks; ++b)
            or_rhs_acm_block(p, Bs, g, dxs + 3 * b, hvy + b * bsz,
                             work + (ptrdiff_t)j * slot + b * bsz,
                             mask ? mask + b * msz : NULL);
    }
    /* :136-154 final stage */
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nblocks; ++b) {
        double *u = hvy + b * bsz, *w1 = work + b * bsz;
        FOR_INTERIOR(u[i] = w1[i]);
        for (int j = 2; j <= nrows; ++j) {
            const double bb = BT(nrows - 1, j - 1);
            if (fabs(bb) < 1.0e-8) continue;
            const double *k = work + (ptrdiff_t)(j - 1) * slot + b * bsz;
            const double db = dt * bb;
            FOR_INTERIOR(u[i] = u[i] + db * k[i]);
        }
    }
#undef BT
#undef FOR_INTERIOR
}
