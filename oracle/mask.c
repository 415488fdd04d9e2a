/*
 * mask.c -- oracle: the ACM penalisation mask of one block.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).
 *
 * CREATE_MASK_meta (module_physics_metamodule.f90:38-84) zeroes the mask and
 * sets the colour to 1, then create_mask_2D_ACM (create_mask.f90:183-320)
 * draws the geometry and the sponge:
 *   - cylinder: draw_circle (module_geometry.f90:315-381) over the block's
 *     interior plus its first upper ghost point, within a bounding box of
 *     the circle widened by the smoothing safety; chi = step_cosine4(dist),
 *     dist the signed distance to the circle (geometry_primitive_2D.f90:21-30);
 *   - sponge: sponge_2D (sponge.f90:4-58) on the interior, "rect" or
 *     "p-norm", with step_cosine3 (module_helpers.f90:436-456).
 * and in 3D create_mask_3D_acm (create_mask.f90:6-174) with "sphere-fixed"
 * (draw_sphere) and sponge_3D, see mask_block_3d.
 * Pinned by the reference's saved masks of TESTING/acm/acm_CDF44
 * (mask_0000000*.h5) on the same grids (tests/test_oracle_acm.py); the 3D
 * sphere and sponges by their slices through the centre plane, which equal
 * the pinned 2D circle and sponges bit for bit, and by radial symmetry
 * (tests/test_oracle_mask.py).
 */
#include <math.h>
#include <string.h>

#include "wavelet_oracle.h"

static double pi_val(void) { return 4.0 * atan(1.0); }

/* step_cosine4 / step_cosine3: 1 below -h, 0 above h, cosine in between */
static double step_cos(double x, double t, double h)
{
    const double xr = x - t;
    if (xr <= -h) return 1.0;
    if (xr >= h) return 0.0;
    return 0.5 * (1.0 + cos((xr + h) * pi_val() / (2.0 * h)));
}

/* create_mask_3D_acm (create_mask.f90:6-174) for geometry "sphere-fixed":
 * draw_sphere (module_geometry.f90:49-117) over the interior plus the first
 * upper ghost point in each direction within the sphere's bounding box
 * widened by the smoothing safety, dist = |x - c| - R
 * (signed_distance_sphere_3D, geometry_primitive_3D.f90:21-30, norm2_3d
 * module_globals.f90:253-258); sponge_3D (sponge.f90:61-233) on the
 * interior, "rect" (min of the six wall distances) or "p-norm". */
static void mask_block_3d(const or_mask_params *mp, const int bs[3], int g, const double x0[3],
                          const double dx[3], double *mask)
{
    const int n1[3] = {bs[0] + 2 * g, bs[1] + 2 * g, bs[2] + 2 * g};
    const long nx = n1[0], ny = n1[1], cs = (long)n1[0] * n1[1] * n1[2];
    if (mp->geometry == 2) {
        const double h = mp->smoothing_width, safety = mp->smoothing_safety, R = mp->R_cyl;
        double dmin = dx[0];
        for (int d = 1; d < 3; ++d) dmin = dx[d] < dmin ? dx[d] : dmin;
        const int nsafe = (int)ceil(safety / dmin);
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            const int bmin = (int)((mp->x_cntr[d] - R - x0[d]) / dx[d]) - nsafe + (g + 1);
            const int bmax = (int)ceil((mp->x_cntr[d] + R - x0[d]) / dx[d]) + nsafe + (g + 1);
            lo[d] = bmin > g + 1 ? bmin : g + 1;
            hi[d] = bmax < n1[d] - g + 1 ? bmax : n1[d] - g + 1;
        }
        for (int iz = lo[2]; iz <= hi[2]; ++iz) {
            const double z = (double)(iz - (g + 1)) * dx[2] + x0[2];
            for (int iy = lo[1]; iy <= hi[1]; ++iy) {
                const double y = (double)(iy - (g + 1)) * dx[1] + x0[1];
                for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                    const double x = (double)(ix - (g + 1)) * dx[0] + x0[0];
                    const double ax = x - mp->x_cntr[0], ay = y - mp->x_cntr[1], az = z - mp->x_cntr[2];
                    const double dist = sqrt(ax * ax + ay * ay + az * az) - R;
                    const double tmp = step_cos(dist, 0.0, h);
                    const long q = ((long)(iz - 1) * ny + (iy - 1)) * nx + (ix - 1);
                    if (tmp >= mask[q]) {
                        mask[q] = tmp;
                        mask[4 * cs + q] = 1.0; /* colour of the (only) geometry */
                    }
                }
            }
        }
    }
    if (mp->use_sponge) {
        const double Lh = 0.5 * mp->L_sponge;
        for (int iz = g; iz < g + bs[2]; ++iz)
            for (int iy = g; iy < g + bs[1]; ++iy)
                for (int ix = g; ix < g + bs[0]; ++ix) {
                    double tmp;
                    if (mp->sponge_type == 0) {
                        const double x = (double)(ix - g) * dx[0] + x0[0];
                        const double y = (double)(iy - g) * dx[1] + x0[1];
                        const double z = (double)(iz - g) * dx[2] + x0[2];
                        const double v[6] = {x, y, z, -(x - mp->domain[0]), -(y - mp->domain[1]),
                                             -(z - mp->domain[2])};
                        tmp = v[0];
                        for (int k = 1; k < 6; ++k) tmp = v[k] < tmp ? v[k] : tmp;
                    } else {
                        const double p = mp->p_sponge, off = 0.5 * mp->domain[0];
                        const double z = pow((double)(iz - g) * dx[2] + x0[2] - off, p);
                        const double y = pow((double)(iy - g) * dx[1] + x0[1] - off, p);
                        const double x = pow((double)(ix - g) * dx[0] + x0[0] - off, p);
                        tmp = -(pow(x + y + z, 1.0 / p) - off);
                    }
                    mask[5 * cs + ((long)iz * ny + iy) * nx + ix] = step_cos(tmp, Lh, Lh);
                }
    }
}

void or_mask_block(const or_mask_params *mp, int dim, const int bs[3], int g, const double x0[3],
                   const double dx[3], double *mask)
{
    const int nx = bs[0] + 2 * g, ny = bs[1] + 2 * g, nz = dim == 3 ? bs[2] + 2 * g : 1;
    const long cs = (long)nx * ny * nz;
    memset(mask, 0, sizeof(double) * (size_t)cs * 6);
    for (long q = 0; q < cs; ++q) mask[4 * cs + q] = 1.0;
    if (dim == 3) {
        mask_block_3d(mp, bs, g, x0, dx, mask);
        return;
    }
    if (mp->geometry == 1) {
        const double h = mp->smoothing_width, safety = mp->smoothing_safety, R = mp->R_cyl;
        const double dmin = dx[0] < dx[1] ? dx[0] : dx[1];
        const int nsafe = (int)ceil(safety / dmin);
        int bmin[2], bmax[2];
        const int n1[2] = {nx, ny};
        for (int d = 0; d < 2; ++d) {
            /* 1-based indices as in the reference; int() truncates toward zero */
            bmin[d] = (int)((mp->x_cntr[d] - R - x0[d]) / dx[d]) - nsafe + (g + 1);
            bmax[d] = (int)ceil((mp->x_cntr[d] + R - x0[d]) / dx[d]) + nsafe + (g + 1);
        }
        const int iy0 = bmin[1] > g + 1 ? bmin[1] : g + 1, iy1 = bmax[1] < n1[1] - g + 1 ? bmax[1] : n1[1] - g + 1;
        const int ix0 = bmin[0] > g + 1 ? bmin[0] : g + 1, ix1 = bmax[0] < n1[0] - g + 1 ? bmax[0] : n1[0] - g + 1;
        for (int iy = iy0; iy <= iy1; ++iy) {
            const double y = (double)(iy - (g + 1)) * dx[1] + x0[1];
            for (int ix = ix0; ix <= ix1; ++ix) {
                const double x = (double)(ix - (g + 1)) * dx[0] + x0[0];
                const double ax = x - mp->x_cntr[0], ay = y - mp->x_cntr[1];
                const double dist = sqrt(ax * ax + ay * ay) - R;
                const double tmp = step_cos(dist, 0.0, h);
                const long q = (long)(iy - 1) * nx + (ix - 1);
                if (tmp >= mask[q]) {
                    mask[q] = tmp;
                    mask[4 * cs + q] = 1.0; /* colour of the (only) geometry */
                }
            }
        }
    }
    if (mp->use_sponge) {
        const double Lh = 0.5 * mp->L_sponge;
        for (int iy = g; iy < g + bs[1]; ++iy)
            for (int ix = g; ix < g + bs[0]; ++ix) {
                double tmp;
                if (mp->sponge_type == 0) {
                    const double x = (double)(ix - g) * dx[0] + x0[0];
                    const double y = (double)(iy - g) * dx[1] + x0[1];
                    const double v[4] = {x, y, -(x - mp->domain[0]), -(y - mp->domain[1])};
                    tmp = v[0];
                    for (int k = 1; k < 4; ++k) tmp = v[k] < tmp ? v[k] : tmp;
                } else {
                    const double p = mp->p_sponge, off = 0.5 * mp->domain[0];
                    const double y = (double)(iy - g) * dx[1] + x0[1] - off;
                    const double x = (double)(ix - g) * dx[0] + x0[0] - off;
                    tmp = -(pow(pow(x, p) + pow(y, p), 1.0 / p) - off);
                }
                mask[5 * cs + (long)iy * nx + ix] = step_cos(tmp, Lh, Lh);
            }
    }
}
