// wh_mask.hip -- the ACM penalisation mask of listed blocks on the device
// (CREATE_MASK_meta "all-parts", LIB/EQUATION/module_physics_metamodule.f90:
// 38-84, with create_mask_2D_ACM, LIB/EQUATION/ACMnew/create_mask.f90:183-320,
// and in 3D create_mask_3D_acm, create_mask.f90:6-174, see k_mask3d):
//   component 0: chi of the cylinder (draw_circle, LIB/EQUATION/insects/
//                module_geometry.f90:315-381), written over the interior and
//                the first upper ghost point within the circle's bounding box
//                widened by the smoothing safety, 0 elsewhere;
//   components 1-3: u_s = 0;  component 4: colour 1;
//   component 5: sponge on the interior (sponge_2D, sponge.f90:4-58), "rect"
//                or "p-norm", 0 on the ghosts.
// Smoothing: step_cosine (LIB/HELPER/module_helpers.f90:436-487).
// Also the per-block statistics of chi the coarsening mask criterion needs
// (coarseningIndicatorMask_tree, coarseningIndicator_tree.f90:338-422).
#include <cmath>

#include "wh_internal.h"

namespace wh {
namespace {

struct MaskDev {
    int geometry, use_sponge, sponge_type;
    double cx, cy, cz, R, h, safety, Lh, p, pinv, off, Lx, Ly, Lz, pi;
};

__device__ __forceinline__ double step_cos(double x, double t, double h, double pi)
{
    const double xr = x - t;
    if (xr <= -h) return 1.0;
    if (xr >= h) return 0.0;
    return 0.5 * (1.0 + cos((xr + h) * pi / (2.0 * h)));
}

__global__ __launch_bounds__(256) void k_mask(double *__restrict__ mask, size_t mblk, size_t cs,
                                              const int32_t *__restrict__ ids,
                                              const double *__restrict__ x0dx, int nb, int nx, int ny,
                                              int g, int bs0, int bs1, const MaskDev *__restrict__ geos,
                                              const int32_t *__restrict__ gidx)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nb * (long)cs) return;
    const int k = (int)(e / (long)cs);
    const int q = (int)(e - (long)k * cs);
    const MaskDev M = geos[gidx ? gidx[k] : 0];  // the listed block's forest's geometry
    const int ix = q % nx, iy = q / nx;  // 2D: nz = 1
    const double x0 = x0dx[6 * k + 0], y0 = x0dx[6 * k + 1], dx = x0dx[6 * k + 3], dy = x0dx[6 * k + 4];
    double *m = mask + (size_t)ids[k] * mblk + q;
    double chi = 0.0;
    if (M.geometry == 1) {
        // draw_circle: 1-based indices, bounds from int()/ceiling() of the
        // circle's box in grid units, widened by ceiling(safety/min dx)
        const int nsafe = (int)ceil(M.safety / fmin(dx, dy));
        const int bxmin = (int)((M.cx - M.R - x0) / dx) - nsafe + (g + 1);
        const int bymin = (int)((M.cy - M.R - y0) / dy) - nsafe + (g + 1);
        const int bxmax = (int)ceil((M.cx + M.R - x0) / dx) + nsafe + (g + 1);
        const int bymax = (int)ceil((M.cy + M.R - y0) / dy) + nsafe + (g + 1);
        const int i1 = ix + 1, j1 = iy + 1;
        if (i1 >= max(bxmin, g + 1) && i1 <= min(bxmax, nx - g + 1) && j1 >= max(bymin, g + 1) &&
            j1 <= min(bymax, ny - g + 1)) {
            const double y = (double)(j1 - (g + 1)) * dy + y0;
            const double x = (double)(i1 - (g + 1)) * dx + x0;
            const double ax = x - M.cx, ay = y - M.cy;
            const double dist = sqrt(ax * ax + ay * ay) - M.R;
            chi = step_cos(dist, 0.0, M.h, M.pi);
        }
    }
    double sp = 0.0;
    if (M.use_sponge && ix >= g && ix < g + bs0 && iy >= g && iy < g + bs1) {
        double tmp;
        if (M.sponge_type == 0) {
            const double x = (double)(ix - g) * dx + x0, y = (double)(iy - g) * dy + y0;
            tmp = fmin(fmin(fmin(x, y), -(x - M.Lx)), -(y - M.Ly));
        } else {
            const double y = (double)(iy - g) * dy + y0 - M.off;
            const double x = (double)(ix - g) * dx + x0 - M.off;
            tmp = -(pow(pow(x, M.p) + pow(y, M.p), M.pinv) - M.off);
        }
        sp = step_cos(tmp, M.Lh, M.Lh, M.pi);
    }
    m[0] = chi;
    m[cs] = 0.0;
    m[2 * cs] = 0.0;
    m[3 * cs] = 0.0;
    m[4 * cs] = 1.0;
    m[5 * cs] = sp;
}

// 3D (create_mask_3D_acm, create_mask.f90:6-174): "sphere-fixed" by
// draw_sphere (module_geometry.f90:49-117: interior plus first upper ghost
// point within the sphere's box widened by the smoothing safety, dist =
// |x - c| - R as norm2_3d sums it), sponge_3D (sponge.f90:61-233) "rect"
// (min of the six wall distances) or "p-norm" on the interior.
__global__ __launch_bounds__(256) void k_mask3d(double *__restrict__ mask, size_t mblk, size_t cs,
                                                const int32_t *__restrict__ ids,
                                                const double *__restrict__ x0dx, int nb, int nx, int ny,
                                                int nz, int g, int bs0, int bs1, int bs2,
                                                const MaskDev *__restrict__ geos, const int32_t *__restrict__ gidx)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nb * (long)cs) return;
    const int k = (int)(e / (long)cs);
    const int q = (int)(e - (long)k * cs);
    const MaskDev M = geos[gidx ? gidx[k] : 0];
    const int ix = q % nx, iy = (q / nx) % ny, iz = q / (nx * ny);
    const double x0 = x0dx[6 * k + 0], y0 = x0dx[6 * k + 1], z0 = x0dx[6 * k + 2];
    const double dx = x0dx[6 * k + 3], dy = x0dx[6 * k + 4], dz = x0dx[6 * k + 5];
    double *m = mask + (size_t)ids[k] * mblk + q;
    double chi = 0.0;
    if (M.geometry == 2) {
        const int nsafe = (int)ceil(M.safety / fmin(fmin(dx, dy), dz));
        const int bxmin = (int)((M.cx - M.R - x0) / dx) - nsafe + (g + 1);
        const int bymin = (int)((M.cy - M.R - y0) / dy) - nsafe + (g + 1);
        const int bzmin = (int)((M.cz - M.R - z0) / dz) - nsafe + (g + 1);
        const int bxmax = (int)ceil((M.cx + M.R - x0) / dx) + nsafe + (g + 1);
        const int bymax = (int)ceil((M.cy + M.R - y0) / dy) + nsafe + (g + 1);
        const int bzmax = (int)ceil((M.cz + M.R - z0) / dz) + nsafe + (g + 1);
        const int i1 = ix + 1, j1 = iy + 1, k1 = iz + 1;
        if (i1 >= max(bxmin, g + 1) && i1 <= min(bxmax, nx - g + 1) && j1 >= max(bymin, g + 1) &&
            j1 <= min(bymax, ny - g + 1) && k1 >= max(bzmin, g + 1) && k1 <= min(bzmax, nz - g + 1)) {
            const double z = (double)(k1 - (g + 1)) * dz + z0;
            const double y = (double)(j1 - (g + 1)) * dy + y0;
            const double x = (double)(i1 - (g + 1)) * dx + x0;
            const double ax = x - M.cx, ay = y - M.cy, az = z - M.cz;
            const double dist = sqrt(ax * ax + ay * ay + az * az) - M.R;
            chi = step_cos(dist, 0.0, M.h, M.pi);
        }
    }
    double sp = 0.0;
    if (M.use_sponge && ix >= g && ix < g + bs0 && iy >= g && iy < g + bs1 && iz >= g && iz < g + bs2) {
        double tmp;
        if (M.sponge_type == 0) {
            const double x = (double)(ix - g) * dx + x0, y = (double)(iy - g) * dy + y0,
                         z = (double)(iz - g) * dz + z0;
            tmp = fmin(fmin(fmin(fmin(fmin(x, y), z), -(x - M.Lx)), -(y - M.Ly)), -(z - M.Lz));
        } else {
            const double z = pow((double)(iz - g) * dz + z0 - M.off, M.p);
            const double y = pow((double)(iy - g) * dy + y0 - M.off, M.p);
            const double x = pow((double)(ix - g) * dx + x0 - M.off, M.p);
            tmp = -(pow(x + y + z, M.pinv) - M.off);
        }
        sp = step_cos(tmp, M.Lh, M.Lh, M.pi);
    }
    m[0] = chi;
    m[cs] = 0.0;
    m[2 * cs] = 0.0;
    m[3 * cs] = 0.0;
    m[4 * cs] = 1.0;
    m[5 * cs] = sp;
}

// chi statistics over the interior of listed blocks: [between, max, min]
__global__ __launch_bounds__(256) void k_mask_stats(const double *__restrict__ mask, size_t mblk, size_t cs,
                                                    const int32_t *__restrict__ ids, int nx, int ny, int g,
                                                    int bs0, int bs1, int bs2, double *__restrict__ out)
{
    __shared__ double smax[256], smin[256];
    __shared__ int sbetw[256];
    const double *m = mask + (size_t)ids[blockIdx.x] * mblk;
    double mx = -1.0e300, mn = 1.0e300;
    int betw = 0;
    for (int e = threadIdx.x; e < bs0 * bs1 * bs2; e += blockDim.x) {
        const int ix = g + e % bs0, iy = g + (e / bs0) % bs1;
        const int iz = bs2 > 1 ? g + e / (bs0 * bs1) : 0;
        const double v = m[((size_t)iz * ny + iy) * nx + ix];
        if (v > 1.0e-12 && v < 1.0 - 1.0e-12) betw = 1;
        mx = fmax(mx, v);
        mn = fmin(mn, v);
    }
    smax[threadIdx.x] = mx;
    smin[threadIdx.x] = mn;
    sbetw[threadIdx.x] = betw;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            sbetw[threadIdx.x] |= sbetw[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x + 0] = (double)sbetw[0];
        out[3 * blockIdx.x + 1] = smax[0];
        out[3 * blockIdx.x + 2] = smin[0];
    }
}

MaskDev make_mask_dev(const wh_mask_params &mp)
{
    MaskDev M{};
    M.geometry = mp.geometry;
    M.use_sponge = mp.use_sponge;
    M.sponge_type = mp.sponge_type;
    M.cx = mp.x_cntr[0];
    M.cy = mp.x_cntr[1];
    M.cz = mp.x_cntr[2];
    M.R = mp.R_cyl;
    M.h = mp.smoothing_width;
    M.safety = mp.smoothing_safety;
    M.Lh = 0.5 * mp.L_sponge;
    M.p = mp.p_sponge;
    M.pinv = 1.0 / mp.p_sponge;
    M.off = 0.5 * mp.domain[0];
    M.Lx = mp.domain[0];
    M.Ly = mp.domain[1];
    M.Lz = mp.domain[2];
    M.pi = 4.0 * std::atan(1.0);
    return M;
}

}  // namespace

int check_mask_pool(wh_pool *P)
{
    WH_CHECK(P->mask && P->d.n_mask >= 6 && P->d.mask_per_block, WH_E_ARG,
             "mask: the pool needs a per-block mask with 6 components");
    return WH_OK;
}

int check_mask_params(wh_pool *P, const wh_mask_params &mp)
{
    const int dim = P->d.dim;
    WH_CHECK(mp.geometry == 0 || (dim == 2 && mp.geometry == 1) || (dim == 3 && mp.geometry == 2), WH_E_ARG,
             "mask: unknown geometry (2D: 1 cylinder; 3D: 2 sphere-fixed)");
    WH_CHECK(mp.sponge_type == 0 || mp.sponge_type == 1, WH_E_ARG, "mask: unknown sponge type");
    WH_CHECK(!(mp.use_sponge && mp.sponge_type == 1) ||
                 (std::fabs(mp.domain[1] - mp.domain[0]) <= 1.0e-10 &&
                  (dim == 2 || std::fabs(mp.domain[2] - mp.domain[0]) <= 1.0e-10)),
             WH_E_ARG, "mask: the p-norm sponge needs a square (cubic) domain");
    WH_CHECK(mp.geometry == 0 || (mp.smoothing_width > 0.0), WH_E_ARG, "mask: smoothing width must be > 0");
    return WH_OK;
}

int run_masks(wh_pool *P, const std::vector<wh_mask_params> &mps, const std::vector<int32_t> &gidx,
              const std::vector<int32_t> &ids, const std::vector<double> &x0dx)
{
    if (int rc = check_mask_pool(P)) return rc;
    for (const wh_mask_params &mp : mps)
        if (int rc = check_mask_params(P, mp)) return rc;
    const int n = (int)ids.size(), ng = (int)mps.size();
    if (n == 0) return WH_OK;
    WH_CHECK(ng > 0 && (ng == 1 ? gidx.empty() : gidx.size() == (size_t)n), WH_E_ARG,
             "mask: one geometry index per block");
    for (int32_t q : gidx) WH_CHECK(q >= 0 && q < ng, WH_E_ARG, "mask: geometry index out of range");
    for (int32_t h : ids) WH_CHECK(h >= 0 && h < P->d.nb_max, WH_E_ARG, "mask: hvy id out of range");
    WH_CHECK(x0dx.size() == (size_t)n * 6, WH_E_ARG, "mask: six origin/spacing values per block");
    std::vector<MaskDev> geos(ng);
    for (int i = 0; i < ng; ++i) geos[i] = make_mask_dev(mps[i]);
    if (int rc = stage_begin(P, sizeof(int32_t) * (n + gidx.size()) + sizeof(double) * n * 6 +
                                    sizeof(MaskDev) * ng))
        return rc;
    const int32_t *d_ids = (const int32_t *)stage_put(P, ids.data(), sizeof(int32_t) * n, true);
    const double *d_x0dx = (const double *)stage_put(P, x0dx.data(), sizeof(double) * n * 6, true);
    const MaskDev *d_geos = (const MaskDev *)stage_put(P, geos.data(), sizeof(MaskDev) * ng, true);
    const int32_t *d_gidx =
        gidx.empty() ? nullptr : (const int32_t *)stage_put(P, gidx.data(), sizeof(int32_t) * gidx.size(), true);
    if (int rc = stage_flush(P)) return rc;
    if (!d_ids || !d_x0dx || !d_geos || (!gidx.empty() && !d_gidx)) return WH_E_HIP;
    const long tot = (long)n * (long)P->cs;
    if (P->d.dim == 3)
        hipLaunchKernelGGL(k_mask3d, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, P->stream, P->mask,
                           P->mblk, P->cs, d_ids, d_x0dx, n, P->nx, P->ny, P->nz, P->d.g, P->d.bs[0], P->d.bs[1],
                           P->d.bs[2], d_geos, d_gidx);
    else
        hipLaunchKernelGGL(k_mask, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, P->stream, P->mask,
                           P->mblk, P->cs, d_ids, d_x0dx, n, P->nx, P->ny, P->d.g, P->d.bs[0], P->d.bs[1], d_geos,
                           d_gidx);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int run_mask(wh_pool *P, const wh_mask_params &mp, const std::vector<int32_t> &ids,
             const std::vector<double> &x0dx)
{
    return run_masks(P, std::vector<wh_mask_params>{mp}, std::vector<int32_t>{}, ids, x0dx);
}

int launch_mask_stats(wh_pool *P, const std::vector<int32_t> &ids, double *d_out)
{
    const int n = (int)ids.size();
    if (n == 0) return WH_OK;
    if (int rc = check_mask_pool(P)) return rc;
    for (int32_t h : ids) WH_CHECK(h >= 0 && h < P->d.nb_max, WH_E_ARG, "mask stats: hvy id out of range");
    if (int rc = stage_begin(P, sizeof(int32_t) * n)) return rc;
    const int32_t *d_ids = (const int32_t *)stage_put(P, ids.data(), sizeof(int32_t) * n);
    if (!d_ids) return WH_E_HIP;
    hipLaunchKernelGGL(k_mask_stats, dim3(n), dim3(256), 0, P->stream, P->mask, P->mblk, P->cs, d_ids, P->nx,
                       P->ny, P->d.g, P->d.bs[0], P->d.bs[1], P->d.dim == 3 ? P->d.bs[2] : 1, d_out);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int run_mask_stats(wh_pool *P, const std::vector<int32_t> &ids, std::vector<double> &stats)
{
    if (int rc = check_mask_pool(P)) return rc;
    const int n = (int)ids.size();
    stats.assign((size_t)n * 3, 0.0);
    if (n == 0) return WH_OK;
    double *d_out = nullptr;
    if (int rc = results_reserve(P, (size_t)n * 3, &d_out)) return rc;
    if (int rc = launch_mask_stats(P, ids, d_out)) return rc;
    return results_fetch(P, (size_t)n * 3, stats);
}

}  // namespace wh

extern "C" int wh_create_mask(wh_pool *pool, const wh_mask_params *mp, const int32_t *ids, int n,
                              const double *x0dx)
{
    using namespace wh;
    WH_CHECK(pool && mp, WH_E_ARG, "wh_create_mask: null argument");
    if (n < 0) {
        std::vector<int32_t> act(pool->h_active.begin(), pool->h_active.begin() + pool->nact);
        std::vector<double> xd(pool->h_x0dx.begin(), pool->h_x0dx.begin() + 6 * (size_t)pool->nact);
        return run_mask(pool, *mp, act, xd);
    }
    WH_CHECK(n == 0 || (ids && x0dx), WH_E_ARG, "wh_create_mask: null list");
    return run_mask(pool, *mp, std::vector<int32_t>(ids, ids + n), std::vector<double>(x0dx, x0dx + 6 * (size_t)n));
}
