/*
 * wavelet_oracle.h -- CPU restatement of WABBIT's CDF wavelet machinery and
 * of adapt_tree (coarsening on the full tree).
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load liboracle.so.
 *
 * Block layout as in wabbit_oracle.h: u[c][iz][iy][ix], x fastest,
 * nx = Bs+2g (nz = 1 in 2D).  Index conventions below are 0-based; the
 * reference's 1-based index i is i-1 here.
 *
 * Pinning:
 *   or_wavelet_setup     closed-form tables (CDF40/CDF44 HD/GD/HR/GR and CE
 *                        widths restated from the formulas), sum rules
 *                        (sum HD = 1, sum HR = 2, vanishing moments of GD)
 *                        (tests/test_oracle_wavelet.py).
 *   or_fwt/or_iwt_block  perfect reconstruction IWT(FWT(u)) = u on the
 *                        interior for every wavelet; SC of polynomials.
 *   or_tree_adapt        coarsening: the reference's --coarsen-everywhere
 *                        outputs (TESTING/wavelets/adaptive_CDF{20,22,40,42,
 *                        44,60,62}/vor_00200.h5), bit for bit for CDFX0;
 *                        threshold indicator with normalised eps, mask
 *                        criterion and security zone: the adaptive cylinder
 *                        run TESTING/acm/acm_CDF44 (grids and iteration
 *                        counts identical at t = 0.05 and 0.1,
 *                        tests/test_oracle_acm.py).
 *   or_tree_refine_everywhere / or_tree_sync_full / or_tree_rk_step
 *                        the same acm_CDF44 run (every step refines, syncs
 *                        and steps the whole tree).
 *   or_tree_dt           same run (the dt sequence fixes the iteration count).
 *   or_renorm_value / or_detail_block_norm / or_component_norm (eps_norm
 *                        L1, L2, H1): no reference fixture uses them; Linfty
 *                        and 2D H1 leave details unchanged, the level
 *                        scaling of the renormed details equals that of the
 *                        norm of the same data, the norms of constants,
 *                        vector groups and a sine in closed form
 *                        (tests/test_oracle_norm.py).  Parity unpinned: the
 *                        per-direction SC-to-WC factors, the -1 of the level
 *                        shift and H1 in 3D follow the reference's lines only.
 *   or_mask_block        the run's saved masks (p-norm sponge, cylinder);
 *                        the rect sponge by symmetry/support properties on a
 *                        non-square domain (tests/test_oracle_mask.py).
 */
#ifndef WAVELET_ORACLE_H
#define WAVELET_ORACLE_H

#include <stdint.h>

#include "wabbit_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_WMAX 16 /* filter support bound: taps -OR_WMAX..OR_WMAX */

/* setup_wavelet (module_wavelets.f90:1031-1430) for "CDFXY", X in {2,4,6},
 * Y in {0,2,4,6}, Y <= X.  Filters are stored on [lo, hi] with coef[i - lo]. */
typedef struct {
    int X, Y, lifted;
    int hd_lo, hd_hi, gd_lo, gd_hi, hr_lo, hr_hi, gr_lo, gr_hi;
    double HD[2 * OR_WMAX + 1], GD[2 * OR_WMAX + 1], HR[2 * OR_WMAX + 1], GR[2 * OR_WMAX + 1];
    int g;           /* g_min: largest filter half-width */
    int g_rhs;       /* max(X/2, FD_max_size) */
    int order;       /* predictor order (multiresolution_<X>th) */
    int Nscl, Nscr, Nwcl, Nwcr, Nreconl, Nreconr, Ndep2l, Ndep2r;
    int specialised; /* 1: CDF20/22/40/42/44/60/62 unrolled forms, 0: generic */
} or_wavelet;

/* returns 0, or -1 for an unsupported name. fd_max_size: reach of the FD
 * scheme (0 when no discretization is set, as in wabbit-post). */
int or_wavelet_setup(const char *name, int fd_max_size, or_wavelet *w);

/* waveletDecomposition_optimized_block (wavelet_decomposition_reconstruction.f90:23-389):
 * src ghost-inclusive block -> dst spaghetti form.  Only the region the
 * reference computes is written; the caller gives dst its prior content. */
void or_fwt_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq,
                  const double *src, double *dst);

/* waveletReconstruction_optimized_block (:426-841), same conventions. */
void or_iwt_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq,
                  const double *src, double *dst);

/* blockFilterXYZ_vct with do_restriction (module_wavelets.f90:307-401) with
 * the filter HD: dst = src, then SC interior points replaced by the
 * separable filter. */
void or_filter_block(const or_wavelet *w, int dim, const int bs[3], int g, int neq,
                     const double *src, double *dst);

/* threshold_block's detail for eps_norm = Linfty (threshold_block.f90:1-130,
 * wavelet_renorm_block module_wavelets.f90:1848-1960): per component the
 * max |WC| over the interior (SCs zeroed); components with group id l >= 2
 * take max over the group of sqrt(wc^2); group id 0 gives 0. */
void or_detail_block(int dim, const int bs[3], int g, int neq, const int *groups,
                     const double *u, double *detail);

/* params%eps_norm (ini_file_to_params.f90:494) */
enum { OR_NORM_LINFTY = 0, OR_NORM_L1 = 1, OR_NORM_L2 = 2, OR_NORM_H1 = 3 };

/* wavelet_renorm_block (module_wavelets.f90:1848-1960) for one value v at a
 * point of a block on `level`, level_ref 0 (threshold_block passes none,
 * threshold_block.f90:72-74): the level shift, then per direction d (x, y,
 * z in that order) in which the point lies on a scaling-coefficient
 * position (ix - g even, 0-based) the SC-to-WC factor.  Linfty and H1 in
 * 2D leave v as it is. */
double or_renorm_value(int eps_norm, int dim, int level, int ix, int iy, int iz, int g, double v);

/* Planted mistakes of the eps_norm functions, for tests that show a pin has
 * teeth (0 in every real use): bit 0 the L2 level shift without its square
 * root, bit 1 componentWiseNorm without the cell volume, bit 2 the level
 * with the wrong sign in the shift. */
void or_set_norm_plant(int bits);

/* or_detail_block with the renormalisation of eps_norm for a block on
 * `level` (threshold_block.f90:69-124); eps_norm OR_NORM_LINFTY is
 * or_detail_block. */
void or_detail_block_norm(int dim, const int bs[3], int g, int neq, const int *groups, const double *u,
                          int level, int eps_norm, double *detail);

/* componentWiseNorm_tree (componentWiseNorm_tree.f90:1-200) over the
 * listed blocks (the leaves, in the order given), norm_case default,
 * threshold_state_vector on: per component group 1 the sum over the blocks
 * of dx^dim * sum(u^2) (L2, H1) or of dx^dim * sum(|u|) (L1) over the
 * interior, the components of a group l >= 2 together; L2/H1 take the
 * square root; Linfty the max of |u| (grouped as threshold_block does).
 * dx[k*3 + d] is block k's spacing.  Group 0 components stay 0. */
void or_component_norm(int dim, const int bs[3], int g, int neq, const int *groups, int eps_norm, int nblk,
                       const double *const *u, const double *dx, double *norm);

/* ---- penalisation mask (create_mask_2D_ACM, create_mask.f90:183-320) --- */
typedef struct {
    int geometry;                   /* 0 none, 1 cylinder (a circle in 2D) */
    double x_cntr[3], R_cyl;
    double smoothing_width, smoothing_safety;   /* cosine smoothing */
    int use_sponge, sponge_type;    /* sponge_type 0: rect, 1: p-norm */
    double L_sponge, p_sponge;
    double domain[3];
} or_mask_params;

/* CREATE_MASK_meta "all-parts" for one block: mask[6][nz][ny][nx] = chi,
 * u_s (3), colour, sponge (module_physics_metamodule.f90:38-84,
 * draw_circle module_geometry.f90:315-381, sponge_2D sponge.f90:4-58). */
void or_mask_block(const or_mask_params *mp, int dim, const int bs[3], int g, const double x0[3],
                   const double dx[3], double *mask);

/* ---- adapt_tree on a periodic leaf forest ----------------------------- */
enum { OR_IND_EVERYWHERE = 0, OR_IND_THRESHOLD = 1, OR_IND_EXTERNAL = 2 };

typedef struct {
    int dim, jmin, jmax, g, neq;
    int bs[3];
    int indicator;                  /* OR_IND_* */
    int use_ce;                     /* params%useCoarseExtension */
    int force_maxlevel_dealiasing;
    double eps;
    double norm[8];                 /* eps * norm per component */
    int groups[8];                  /* threshold_state_vector_component */
    int eps_normalized;             /* norm = Linfty of the leaves (componentWiseNorm_tree) */
    int threshold_mask;             /* coarseningIndicatorMask_tree */
    int security_zone;              /* addSecurityZone_CE_tree */
    double domain[3];
    or_mask_params mask;
    /* wavelet_decompose_full_tree's branch: 0 the reference's rule
     * (leaf-first iff every Bs >= 3 max|HD tap|, adapt_tree.f90:321-322),
     * 1 leaf-first, 2 level-wise.  Forcing a branch is a test device: both
     * must give the same decomposition. */
    int decompose_branch;
    /* planted mistakes, for tests that show a comparison has teeth (0 in
     * every real use): bit 0 level-wise branch: leaves below the finest
     * level are never decomposed; bit 1 level-wise branch: the ghost sync
     * of iterations after the first is skipped; bit 2: the z pass of 3D
     * prediction in ghost syncs is of order 2 */
    int plant;
    int eps_norm;                   /* OR_NORM_*: params%eps_norm */
} or_adapt_params;

/* A forest with heavy data that the oracle owns.  n blocks in list order. */
typedef struct or_tree or_tree;

or_tree *or_tree_create(const or_wavelet *w, const or_adapt_params *p, int n, const int64_t *tc,
                        const int *level, const double *hvy);
void or_tree_destroy(or_tree *T);
int or_tree_size(const or_tree *T);
/* copies out the leaf list (sorted by level, then treecode) and data */
void or_tree_get(const or_tree *T, int64_t *tc, int *level, double *hvy);
/* sync_ghosts_tree (full sync of all ghosts, filtered restriction for lifted
 * wavelets) */
int or_tree_sync_full(or_tree *T);
/* refine_tree "everywhere" execution (refineBlock of every leaf below jmax) */
int or_tree_refine_everywhere(or_tree *T);
/* adapt_tree (adapt_tree.f90:11-255) without the final load balancing */
int or_tree_adapt(or_tree *T);
/* adapt_tree with indicator "nothing (external)": leaf_flags (-1/0/1) in
 * or_tree_get's order replace coarseningIndicator_tree */
int or_tree_adapt_flags(or_tree *T, const int *leaf_flags);
/* the ACM time step on the leaves: min block dt (before the host clamps)
 * and one RungeKuttaGeneric step with sync_ghosts_RHS_tree of depth g_rhs
 * and the mask of every leaf rebuilt from mp (mp NULL: no mask) */
double or_tree_dt(or_tree *T, const or_acm_params *acm);
int or_tree_rk_step(or_tree *T, const or_acm_params *acm, const or_mask_params *mp,
                    const double *butcher, int nrows, double dt, int g_rhs);
/* the mask (chi) of the leaves, layout as the data: [n][6][...] */
void or_tree_get_mask(or_tree *T, const or_mask_params *mp, double *mask);

#ifdef __cplusplus
}
#endif
#endif
