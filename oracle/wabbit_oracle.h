/*
 * wabbit_oracle.h -- CPU restatement of WABBIT's per-block time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in wabbit_amd/ links or calls it.
 *
 * Data layout is the reference's heavy-data layout, Fortran column-major
 * hvy_block(nx, ny, nz, neq, Nb) (LIB/MESH/allocate_forest.f90:228), which in
 * C order is u[b][c][iz][iy][ix] with nx = Bs(1)+2g, ny = Bs(2)+2g,
 * nz = Bs(3)+2g in 3D and nz = 1 in 2D.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   or_rhs_acm_block   FD2/FD4/FD6 skew-symmetric 2D: reference golden
 *                      outputs TESTING/acm/3vortices/3vorticesEqui*; 3D:
 *                      TESTING/acm/taylorGreen/taylorGreenEqui_*; non-skew
 *                      and all schemes: exactness on polynomials of the
 *                      scheme's order and the skew == non-skew identity for
 *                      divergence-free linear fields.  FD_4th_central_optimized
 *                      (Tam-Webb) is pinned only by its polynomial order
 *                      (exact on quadratics) -- partially unpinned.
 *   or_dt_block        golden outputs above (dt sequence drives the final
 *                      state) and the closed form for constant fields.
 *   or_rk_*            golden outputs above and the textbook RK4 order test.
 *   or_sync_equidistant  exact periodic copy; brute-force global indexing.
 */
#ifndef WABBIT_ORACLE_H
#define WABBIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* order_discretization codes (LIB/EQUATION/ACMnew/rhs_ACM.f90:374-870) */
enum {
    OR_FD2 = 2,           /* "FD_2nd_central"            */
    OR_FD4 = 4,           /* "FD_4th_central"            */
    OR_FD4_TW = 40,       /* "FD_4th_central_optimized"  */
    OR_FD6 = 6            /* "FD_6th_central"            */
};

/* subset of type_params_acm (LIB/EQUATION/ACMnew/module_ACM.f90:56-149)
 * that the per-block time step reads. */
typedef struct {
    int dim;
    int discretization;   /* OR_FD* */
    int skew_symmetry;
    int penalization;     /* mask components 1..5 are read when set */
    int use_sponge;       /* mask component 6 is read when set */
    double c_0, nu, gamma_p;
    double C_eta, C_sponge;
    double u_mean_set[3];
    double CFL, CFL_nu, CFL_eta;
    /* planted mistakes, for tests that show a check has teeth (0 in every
     * real use): bit 0 the convective term takes du_c/dy where it needs
     * du_c/dx and vice versa; bit 1 FD6's first derivative drops its +3
     * term; bit 2 the Tam-Webb +3 coefficient has the wrong sign; bit 3
     * the FD2 second derivative uses -1 for the centre weight */
    int plant;
} or_acm_params;

/* RHS of one block: RHS_2D_acm / RHS_3D_acm (rhs_ACM.f90:292-922, 927-1779).
 * u, rhs: [neq][nz][ny][nx]; mask: [6][nz][ny][nx] or NULL.
 * Only interior points of rhs are written. */
void or_rhs_acm_block(const or_acm_params *p, const int Bs[3], int g,
                      const double dx[3], const double *u, double *rhs,
                      const double *mask);

/* GET_DT_BLOCK_ACM (module_ACM.f90:617-691). */
double or_dt_block(const or_acm_params *p, const int Bs[3], int g,
                   const double dx[3], const double *u);

/* Integral-stage divergence guard (rhs_ACM.f90:133-146): returns the max
 * absolute interior value over all neq components. */
double or_maxabs_block(const int Bs[3], int g, int neq, const double *u);

/* Same-level periodic ghost fill on an equidistant grid of nb^dim blocks.
 * block_at[(k*nb + j)*nb + i] is the pool slot of block (i,j,k).
 * Fills every ghost point within gs of the interior (the g_minus = g_plus =
 * gs box of sync_ghosts_RHS_tree) with the periodic neighbour's value. */
void or_sync_equidistant(int dim, const int Bs[3], int g, int neq, int nb,
                         const int *block_at, int nblocks, double *hvy,
                         int gs);

/* One RungeKuttaGeneric step (LIB/TIME/runge_kutta_generic.f90:1-156) on an
 * equidistant periodic grid.  hvy: [nblocks][neq][nz][ny][nx];
 * work: [nslots][nblocks][neq][nz][ny][nx] with nslots = nstages+1.
 * butcher: (nstages+1) x (nstages+1) row-major, rows as in
 * ini_file_to_params.f90:640-646.  x0dx: per block (x0[3], dx[3]).
 * mask: [nblocks][6][nz][ny][nx] or NULL. */
void or_rk_step_equidistant(const or_acm_params *p, const int Bs[3], int g,
                            int g_rhs, int neq, int nb, const int *block_at,
                            int nblocks, const double *dxs, double *hvy,
                            double *work, const double *mask,
                            const double *butcher, int nrows, double dt);

/* min over blocks of or_dt_block (calculate_time_step.f90:28-48, before the
 * host-side clamps). */
double or_dt_min(const or_acm_params *p, const int Bs[3], int g, int neq,
                 int nblocks, const double *dxs, const double *hvy);

/* A leaf forest: block b (0..n-1) has treecode tc[b] at level level[b] and
 * lives in slot b of the heavy array (layout as above). */
typedef struct {
    int dim, jmax, g;
    int bs[3];
    int n;
    const int64_t *tc;
    const int *level;
} or_forest;

/* `prediction` (module_wavelets.f90:96-284) as whole-array passes:
 * coarse (nx,ny,nz) -> fine (2nx-1, 2ny-1, 2nz-1); nz = 1 in 2D. */
void or_prediction(const double *coarse, int nx, int ny, int nz, double *fine, int order);
void or_prediction_zorder(const double *coarse, int nx, int ny, int nz, double *fine, int order, int zorder);

/* Ghost fill of depth gm/gp on a 2:1-balanced periodic leaf forest by point
 * location (see adapt.c).  Returns 0, -1 duplicate block, -2 not balanced,
 * -3 prediction outside the coarse block's interpolated grid. */
int or_sync_leaf(const or_forest *f, double *hvy, int neq, int gm, int gp, int order);

/* refineBlock (refinementExecute.f90:4-116) of one mother into 2^dim
 * daughter arrays (ghost-inclusive, same layout). */
void or_refine_block(int dim, const int bs[3], int g, int neq, const double *mother,
                     double *const *daughters, int order);

#ifdef __cplusplus
}
#endif
#endif
