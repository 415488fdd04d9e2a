/*
 * wavelet.c -- oracle: CDF wavelet filters, one-level FWT/IWT of a block,
 * the restriction filter and the Linfty detail of threshold_block.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see wavelet_oracle.h for pinning.
 *
 * Summation order.  The reference computes every filter output as a chain
 * of products added left to right (gfortran, no reassociation); the unrolled
 * cases of wavelet_decomposition_reconstruction.f90 (CDF20/22/40/42/44/60/62)
 * skip the zero taps and, in the reconstruction, start the chain with the
 * point's own coefficient; the generic case sums two separate SUM()s.  The
 * tables built by terms_fwt/terms_iwt restate those chains term by term so
 * the results agree bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wavelet_oracle.h"

/* Stencil_Int_* (module_wavelets.f90:14-22): interpolation stencils on
 * -(N-1)..N-1 */
static void stencil_int(int order, double *c /* 2*order-1 */)
{
    if (order == 2) {
        c[0] = 1.0 / 2.0; c[1] = 2.0 / 2.0; c[2] = 1.0 / 2.0;
    } else if (order == 4) {
        const double v[7] = {-1.0, 0.0, 9.0, 16.0, 9.0, 0.0, -1.0};
        for (int i = 0; i < 7; ++i) c[i] = v[i] / 16.0;
    } else {
        const double v[11] = {3.0, 0.0, -25.0, 0.0, 150.0, 256.0, 150.0, 0.0, -25.0, 0.0, 3.0};
        for (int i = 0; i < 11; ++i) c[i] = v[i] / 256.0;
    }
}

#define TAP(arr, lo, i) ((arr)[(i) - (lo) + 0])

int or_wavelet_setup(const char *name, int fd_max_size, or_wavelet *w)
{
    memset(w, 0, sizeof(*w));
    if (!name || strncmp(name, "CDF", 3) != 0 || strlen(name) != 5) return -1;
    const int X = name[3] - '0', Y = name[4] - '0';
    if (!(X == 2 || X == 4 || X == 6) || !(Y == 0 || Y == 2 || Y == 4 || Y == 6) || Y > X) return -1;
    w->X = X;
    w->Y = Y;
    w->lifted = Y != 0;
    /* HR = Stencil_Int_Xth (:1157-1177) */
    w->hr_lo = -(X - 1);
    w->hr_hi = X - 1;
    stencil_int(X, w->HR);
    if (Y == 0) {
        w->hd_lo = w->hd_hi = 0;
        w->HD[0] = 1.0;
    } else {
        /* HD(i) = delta_i0 + sum_j (-1)^j HR(j) h~(i-j) / 2 with h~ the
         * order-Y stencil whose centre is zeroed (:1182-1234) */
        double ht[11];
        stencil_int(Y, ht);
        const int ht_lo = -(Y - 1), ht_hi = Y - 1;
        ht[-ht_lo] = 0.0;
        w->hd_lo = w->hr_lo + ht_lo;
        w->hd_hi = w->hr_hi + ht_hi;
        for (int i = w->hd_lo; i <= w->hd_hi; ++i) {
            double v = i == 0 ? 1.0 : 0.0;
            for (int j = w->hr_lo; j <= w->hr_hi; ++j) {
                if (i - j < ht_lo || i - j > ht_hi) continue;
                const double s = (j % 2 == 0) ? 1.0 : -1.0;
                v = v + s * TAP(w->HR, w->hr_lo, j) * ht[i - j - ht_lo] / 2.0;
            }
            TAP(w->HD, w->hd_lo, i) = v;
        }
    }
    /* GD(i) = (-1)^i HR(i), GR(i) = (-1)^i HD(i) (:1275-1288) */
    w->gd_lo = w->hr_lo;
    w->gd_hi = w->hr_hi;
    for (int i = w->gd_lo; i <= w->gd_hi; ++i)
        TAP(w->GD, w->gd_lo, i) = ((i % 2 == 0) ? 1.0 : -1.0) * TAP(w->HR, w->hr_lo, i);
    w->gr_lo = w->hd_lo;
    w->gr_hi = w->hd_hi;
    for (int i = w->gr_lo; i <= w->gr_hi; ++i)
        TAP(w->GR, w->gr_lo, i) = ((i % 2 == 0) ? 1.0 : -1.0) * TAP(w->HD, w->hd_lo, i);
    w->order = X;
    /* g_min = largest |bound| of the four filters (:1325-1334) */
    int gm = 0;
    const int b[8] = {w->hd_lo, w->hd_hi, w->gd_lo, w->gd_hi, w->hr_lo, w->hr_hi, w->gr_lo, w->gr_hi};
    for (int i = 0; i < 8; ++i) gm = abs(b[i]) > gm ? abs(b[i]) : gm;
    w->g = gm;
    w->g_rhs = X / 2 > fd_max_size ? X / 2 : fd_max_size;
    /* coarse-extension widths (:1368-1420) */
    w->Nscl = abs(w->hd_lo) - 1 > 0 ? abs(w->hd_lo) - 1 : 0;
    w->Nwcl = w->Nscl + abs(w->gd_lo);
    w->Nreconl = w->Nwcl + abs(w->gr_lo);
    {
        const int t = (abs(w->hr_lo) + 1) / 2 - 1;
        w->Ndep2l = w->Nreconl + (t > 0 ? t : 0);
    }
    w->Nscr = w->hd_hi;
    w->Nwcr = w->Nscr + w->gd_hi;
    w->Nreconr = w->Nwcr + w->gr_hi;
    w->Ndep2r = w->Nreconr + (w->hr_hi + 1) / 2;
    /* lifted CDF24/26/28/46 widen by a; none of the supported names do */
    {
        const int m = 2 * fd_max_size;
        const int dl = m - w->Nwcl > 0 ? m - w->Nwcl : 0;
        w->Nwcl += dl;
        w->Nreconl += dl;
        w->Ndep2l += dl;
        const int dr = m - w->Nwcr > 0 ? m - w->Nwcr : 0;
        w->Nwcr += dr;
        w->Nreconr += dr;
        w->Ndep2r += dr;
    }
    w->specialised = (Y == 0) || (X == 2 && Y == 2) || (X == 4 && (Y == 2 || Y == 4)) ||
                     (X == 6 && Y == 2);
    return 0;
}

/* ---- term chains -------------------------------------------------------- */
typedef struct {
    int off;
    double c;
    int brk; /* 1: this term starts a new group added to the running total */
} term_t;

/* FWT output at an SC (sc=1) or WC position: nonzero taps of HD / GD in
 * index order, one chain (the unrolled forms; the generic SUM adds the zero
 * taps, which leaves the value unchanged). */
static int terms_fwt(const or_wavelet *w, int sc, term_t *t)
{
    int n = 0;
    const double *F = sc ? w->HD : w->GD;
    const int lo = sc ? w->hd_lo : w->gd_lo, hi = sc ? w->hd_hi : w->gd_hi;
    for (int j = lo; j <= hi; ++j)
        if (F[j - lo] != 0.0) t[n++] = (term_t){j, F[j - lo], 0};
    return n;
}

/* IWT output at an SC or WC position. */
static int terms_iwt(const or_wavelet *w, int sc, term_t *t)
{
    int n = 0;
    if (!w->lifted) {
        /* CDFX0: u_r = u; WC points: u_r(ix) + HR(-k) u(ix-k) + ... (odd k) */
        t[n++] = (term_t){0, 1.0, 0};
        if (!sc)
            for (int j = w->hr_lo; j <= w->hr_hi; ++j)
                if ((j & 1) && w->HR[j - w->hr_lo] != 0.0) t[n++] = (term_t){j, w->HR[j - w->hr_lo], 0};
        return n;
    }
    if (w->specialised) {
        if (sc) {
            /* u(ix) + GR(odd taps) in index order */
            t[n++] = (term_t){0, 1.0, 0};
            for (int j = w->gr_lo; j <= w->gr_hi; ++j)
                if ((j & 1) && w->GR[j - w->gr_lo] != 0.0) t[n++] = (term_t){j, w->GR[j - w->gr_lo], 0};
        } else {
            /* HR(odd taps), then GR(0), then GR(even taps != 0), one chain */
            for (int j = w->hr_lo; j <= w->hr_hi; ++j)
                if ((j & 1) && w->HR[j - w->hr_lo] != 0.0) t[n++] = (term_t){j, w->HR[j - w->hr_lo], 0};
            t[n++] = (term_t){0, w->GR[0 - w->gr_lo], 0};
            for (int j = w->gr_lo; j <= w->gr_hi; ++j)
                if (j != 0 && !(j & 1) && w->GR[j - w->gr_lo] != 0.0)
                    t[n++] = (term_t){j, w->GR[j - w->gr_lo], 0};
        }
        return n;
    }
    /* generic: sum(HR * SC-buffer) + sum(GR * WC-buffer) */
    if (sc) {
        t[n++] = (term_t){0, 1.0, 0};
        int first = 1;
        for (int j = w->gr_lo; j <= w->gr_hi; ++j)
            if ((j & 1) && w->GR[j - w->gr_lo] != 0.0) {
                t[n++] = (term_t){j, w->GR[j - w->gr_lo], first};
                first = 0;
            }
    } else {
        for (int j = w->hr_lo; j <= w->hr_hi; ++j)
            if ((j & 1) && w->HR[j - w->hr_lo] != 0.0) t[n++] = (term_t){j, w->HR[j - w->hr_lo], 0};
        int first = 1;
        for (int j = w->gr_lo; j <= w->gr_hi; ++j)
            if (!(j & 1) && w->GR[j - w->gr_lo] != 0.0) {
                t[n++] = (term_t){j, w->GR[j - w->gr_lo], first};
                first = 0;
            }
    }
    return n;
}

static inline double eval_chain(const term_t *t, int n, const double *p, long stride)
{
    double total = 0.0, cur = 0.0;
    int have = 0;
    for (int k = 0; k < n; ++k) {
        const double v = p[t[k].off * stride] * t[k].c;
        if (k == 0) {
            cur = v;
        } else if (t[k].brk) {
            total = have ? total + cur : cur;
            have = 1;
            cur = v;
        } else {
            cur = cur + v;
        }
    }
    return have ? total + cur : cur;
}

/* Three separable passes (x over an extended row range, then y, then z) of
 * a two-parity filter pair; the input of pass k+1 is the output of pass k.
 * ext: rows beyond the interior covered by the x pass (f_size_max). */
static void separable(int dim, const int bs[3], int g, int neq, const double *src, double *dst,
                      const term_t *tsc, int nsc, const term_t *twc, int nwc, int ext)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const long cs = (long)nx * ny * nz;
    const int zext = dim == 3 ? ext : 0;
    const int z0 = dim == 3 ? g : 0, z1 = dim == 3 ? g + bs[2] : 1;
    double *col = (double *)malloc(sizeof(double) * (size_t)(nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz)));
    for (int c = 0; c < neq; ++c) {
        const double *s = src + c * cs;
        double *d = dst + c * cs;
        /* x pass: rows iy in [g-ext, g+Bs+ext), iz likewise in 3D */
        for (int iz = z0 - zext; iz < z1 + zext; ++iz)
            for (int iy = g - ext; iy < g + bs[1] + ext; ++iy)
                for (int ix = g; ix < g + bs[0]; ++ix) {
                    const long q = ((long)iz * ny + iy) * nx + ix;
                    d[q] = ((ix - g) & 1) ? eval_chain(twc, nwc, s + q, 1) : eval_chain(tsc, nsc, s + q, 1);
                }
        /* y pass (in place through a column buffer) */
        for (int iz = z0 - zext; iz < z1 + zext; ++iz)
            for (int ix = g; ix < g + bs[0]; ++ix) {
                for (int iy = 0; iy < ny; ++iy) col[iy] = d[((long)iz * ny + iy) * nx + ix];
                for (int iy = g; iy < g + bs[1]; ++iy)
                    d[((long)iz * ny + iy) * nx + ix] =
                        ((iy - g) & 1) ? eval_chain(twc, nwc, col + iy, 1) : eval_chain(tsc, nsc, col + iy, 1);
            }
        if (dim == 3)
            for (int iy = g; iy < g + bs[1]; ++iy)
                for (int ix = g; ix < g + bs[0]; ++ix) {
                    for (int iz = 0; iz < nz; ++iz) col[iz] = d[((long)iz * ny + iy) * nx + ix];
                    for (int iz = g; iz < g + bs[2]; ++iz)
                        d[((long)iz * ny + iy) * nx + ix] = ((iz - g) & 1) ? eval_chain(twc, nwc, col + iz, 1)
                                                                          : eval_chain(tsc, nsc, col + iz, 1);
                }
    }
    free(col);
}

void or_fwt_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq, const double *src,
                  double *dst)
{
    term_t tsc[2 * OR_WMAX + 2], twc[2 * OR_WMAX + 2];
    const int nsc = terms_fwt(w, 1, tsc), nwc = terms_fwt(w, 0, twc);
    /* f_size_max = max(ub HD, ub GD) (wavelet_decomposition_reconstruction.f90:45-47) */
    const int ext = w->hd_hi > w->gd_hi ? w->hd_hi : w->gd_hi;
    separable(dim, bs, g, neq, src, dst, tsc, nsc, twc, nwc, ext);
}

void or_iwt_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq, const double *src,
                  double *dst)
{
    term_t tsc[2 * OR_WMAX + 4], twc[2 * OR_WMAX + 4];
    const int nsc = terms_iwt(w, 1, tsc), nwc = terms_iwt(w, 0, twc);
    const int ext = w->hr_hi > w->gr_hi ? w->hr_hi : w->gr_hi;
    if (!w->lifted) {
        /* CDFX0 copies the whole block first (:483, :509, :538) */
        const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
        memcpy(dst, src, sizeof(double) * (size_t)nx * ny * nz * neq);
    }
    separable(dim, bs, g, neq, src, dst, tsc, nsc, twc, nwc, ext);
}

void or_filter_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq, const double *src,
                     double *dst)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const long cs = (long)nx * ny * nz;
    memcpy(dst, src, sizeof(double) * (size_t)cs * neq);
    const int lo = w->hd_lo, hi = w->hd_hi;
    if (lo == 0 && hi == 0 && fabs(w->HD[0] - 1.0) <= 1.0e-10) return;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)cs);
    const int zf0 = dim == 3 ? g + lo : 0, zf1 = dim == 3 ? g + bs[2] + hi : 1;
    const int zs0 = dim == 3 ? g : 0, zs1 = dim == 3 ? g + bs[2] : 1, zstep = dim == 3 ? 2 : 1;
    for (int c = 0; c < neq; ++c) {
        const double *s = src + c * cs;
        double *d = dst + c * cs;
        /* x: rows g+lo .. g+Bs-1+hi, x at SC interior points */
        for (int iz = zf0; iz < zf1; ++iz)
            for (int iy = g + lo; iy < g + bs[1] + hi; ++iy)
                for (int ix = g; ix < g + bs[0]; ix += 2) {
                    const long q = ((long)iz * ny + iy) * nx + ix;
                    double v = 0.0;
                    for (int sh = lo; sh <= hi; ++sh) v = v + s[q + sh] * w->HD[sh - lo];
                    d[q] = v;
                }
        memcpy(tmp, d, sizeof(double) * (size_t)cs);
        for (int iz = zf0; iz < zf1; ++iz)
            for (int iy = g; iy < g + bs[1]; iy += 2)
                for (int ix = g; ix < g + bs[0]; ix += 2) {
                    const long q = ((long)iz * ny + iy) * nx + ix;
                    double v = 0.0;
                    for (int sh = lo; sh <= hi; ++sh) v = v + tmp[q + (long)sh * nx] * w->HD[sh - lo];
                    d[q] = v;
                }
        if (dim == 3) {
            memcpy(tmp, d, sizeof(double) * (size_t)cs);
            for (int iz = zs0; iz < zs1; iz += zstep)
                for (int iy = g; iy < g + bs[1]; iy += 2)
                    for (int ix = g; ix < g + bs[0]; ix += 2) {
                        const long q = ((long)iz * ny + iy) * nx + ix;
                        double v = 0.0;
                        for (int sh = lo; sh <= hi; ++sh) v = v + tmp[q + (long)sh * nx * ny] * w->HD[sh - lo];
                        d[q] = v;
                    }
        }
    }
    free(tmp);
}

/* planted mistakes of the eps_norm functions (or_set_norm_plant), for the
 * tests that show their pins have teeth; 0 in every real use */
static int norm_plant = 0;
void or_set_norm_plant(int bits) { norm_plant = bits; }

double or_renorm_value(int eps_norm, int dim, int level, int ix, int iy, int iz, int g, double v)
{
    /* the patch's copy of the block, then the level shift, then the SC
     * rows of x, of y, of z (module_wavelets.f90:1909-1953); Jref = 0 */
    const int Jref = 0;
    const int sc[3] = {((ix - g) & 1) == 0, ((iy - g) & 1) == 0, dim == 3 && ((iz - g) & 1) == 0};
    /* plant bit 2: the level enters the shift with the wrong sign */
    const int lv = (norm_plant & 4) ? -level - 2 : level;
    double r = v;
    if (eps_norm == OR_NORM_L1) {
        r = r * pow(2.0, (double)((Jref - lv - 1) * dim));
        for (int d = 0; d < dim; ++d)
            if (sc[d]) r = r / 4.0;
    } else if (eps_norm == OR_NORM_L2) {
        /* plant bit 0: the L2 shift without its square root (L1's) */
        r = r * pow(2.0, (double)((Jref - lv - 1) * dim) / ((norm_plant & 1) ? 1.0 : 2.0));
        for (int d = 0; d < dim; ++d)
            if (sc[d]) r = r / 2.0;
    } else if (eps_norm == OR_NORM_H1 && dim == 3) {
        r = r * pow(2.0, (double)(Jref - level) * (2.0 - (double)dim) / 2.0);
        for (int d = 0; d < dim; ++d)
            if (sc[d]) r = r * pow(2.0, 2.0 * ((double)dim - 2.0) / 3.0);
    }
    return r;
}

void or_detail_block(int dim, const int bs[3], int g, int neq, const int *groups, const double *u,
                     double *detail)
{
    or_detail_block_norm(dim, bs, g, neq, groups, u, 0, OR_NORM_LINFTY, detail);
}

void or_component_norm(int dim, const int bs[3], int g, int neq, const int *groups, int eps_norm, int nblk,
                       const double *const *u, const double *dx, double *norm)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const long cs = (long)nx * ny * nz;
    const int z0 = dim == 3 ? g : 0, z1 = dim == 3 ? g + bs[2] : 1;
    const int inf = eps_norm == OR_NORM_LINFTY;
    int maxgrp = 0;
    for (int c = 0; c < neq; ++c) maxgrp = groups[c] > maxgrp ? groups[c] : maxgrp;
    for (int c = 0; c < neq; ++c) norm[c] = inf ? -1.0 : 0.0;
    /* sum over the points of components cl[0..ncl) (Fortran order: x, y,
     * z, then component) of u^2, |u|, or the max of |u| */
    double nb;
#define NORM_SECTION(COND)                                                                  \
    do {                                                                                   \
        nb = 0.0;                                                                          \
        for (int c = 0; c < neq; ++c) {                                                    \
            if (!(COND)) continue;                                                         \
            for (int iz = z0; iz < z1; ++iz)                                               \
                for (int iy = g; iy < g + bs[1]; ++iy)                                     \
                    for (int ix = g; ix < g + bs[0]; ++ix) {                               \
                        const double v = ub[c * cs + ((long)iz * ny + iy) * nx + ix];      \
                        if (inf) nb = fabs(v) > nb ? fabs(v) : nb;                         \
                        else if (eps_norm == OR_NORM_L1) nb = nb + fabs(v);                \
                        else nb = nb + v * v;                                              \
                    }                                                                      \
        }                                                                                  \
    } while (0)
    for (int k = 0; k < nblk; ++k) {
        const double *ub = u[k];
        double vol = 1.0;
        for (int d = 0; d < dim; ++d) vol = vol * dx[k * 3 + d];
        if (norm_plant & 2) vol = 1.0; /* plant bit 1: the cell volume dropped */
        for (int p = 0; p < neq; ++p) {
            if (groups[p] != 1) continue;
            NORM_SECTION(c == p);
            norm[p] = inf ? (nb > norm[p] ? nb : norm[p]) : norm[p] + vol * nb;
        }
        for (int l = 2; l <= maxgrp; ++l) {
            NORM_SECTION(groups[c] == l);
            for (int c = 0; c < neq; ++c)
                if (groups[c] == l) norm[c] = inf ? (nb > norm[c] ? nb : norm[c]) : norm[c] + vol * nb;
        }
    }
#undef NORM_SECTION
    if (eps_norm == OR_NORM_L2 || eps_norm == OR_NORM_H1)
        for (int c = 0; c < neq; ++c) norm[c] = sqrt(norm[c]);
}

void or_detail_block_norm(int dim, const int bs[3], int g, int neq, const int *groups, const double *u,
                          int level, int eps_norm, double *detail)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const long cs = (long)nx * ny * nz;
    const int z0 = dim == 3 ? g : 0, z1 = dim == 3 ? g + bs[2] : 1;
    int maxgrp = 0;
    for (int c = 0; c < neq; ++c) maxgrp = groups[c] > maxgrp ? groups[c] : maxgrp;
    for (int c = 0; c < neq; ++c) detail[c] = -1.0;
    /* value with SCs zeroed, renormed (wavelet_renorm_block) */
#define WCV(c, iz, iy, ix)                                                                    \
    ((((ix - g) & 1) == 0 && ((iy - g) & 1) == 0 && (dim == 2 || ((iz - g) & 1) == 0))        \
         ? 0.0                                                                                \
         : or_renorm_value(eps_norm, dim, level, ix, iy, iz, g,                              \
                           u[(c) * cs + ((long)(iz) * ny + (iy)) * nx + (ix)]))
    for (int l = 2; l <= maxgrp; ++l) {
        double m = -1.0;
        int any = 0;
        for (int c = 0; c < neq; ++c) {
            if (groups[c] != l) continue;
            for (int iz = z0; iz < z1; ++iz)
                for (int iy = g; iy < g + bs[1]; ++iy)
                    for (int ix = g; ix < g + bs[0]; ++ix) {
                        const double v = WCV(c, iz, iy, ix);
                        const double a = sqrt(v * v);
                        if (!any || a > m) m = a;
                        any = 1;
                    }
        }
        for (int c = 0; c < neq; ++c)
            if (groups[c] == l) detail[c] = m;
    }
    for (int c = 0; c < neq; ++c) {
        if (groups[c] == 1) {
            double m = 0.0;
            int any = 0;
            for (int iz = z0; iz < z1; ++iz)
                for (int iy = g; iy < g + bs[1]; ++iy)
                    for (int ix = g; ix < g + bs[0]; ++ix) {
                        const double a = fabs(WCV(c, iz, iy, ix));
                        if (!any || a > m) m = a;
                        any = 1;
                    }
            detail[c] = m;
        } else if (groups[c] == 0) {
            detail[c] = 0.0;
        }
    }
#undef WCV
}
