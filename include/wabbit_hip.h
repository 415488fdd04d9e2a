/*
 * wabbit_hip.h -- C ABI of the MI355X (gfx950) block time-step library.
 *
 * Drop-in for the per-block time-step path of WABBIT.  The host keeps the
 * tree, treecodes, neighbour lists and refinement flags; the library owns
 * the heavy data (hvy_block, hvy_work, hvy_tmp, hvy_mask) in device memory
 * and runs the ACM right-hand side, ghost-node synchronisation, Runge-Kutta
 * stage updates and the dt reduction as HIP kernels in FP64.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative WH_E* code on
 *    failure; wh_last_error() returns the message of the last failure on the
 *    calling thread.  The reference aborts instead of returning codes
 *    (LIB/MODULE/module_globals.f90:131-162); a Fortran binding maps non-zero
 *    to abort().
 *  - Heavy data layout is the reference's, Fortran column-major
 *    hvy_block(nx, ny, nz, neq, Nb) (LIB/MESH/allocate_forest.f90:228), i.e.
 *    C order [Nb][neq][nz][ny][nx], nx = Bs1+2g, nz = 1 in 2D.  Work slots are
 *    hvy_work(nx, ny, nz, neq, Nb, nslots): slot is the outermost index.
 *  - Block ids ("hvy ids") are 0-based here (the reference's are 1-based).
 *  - One host thread per pool; calls are stream-ordered on the pool's stream.
 */
#ifndef WABBIT_HIP_H
#define WABBIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    WH_OK = 0,
    WH_E_ARG = -1,      /* invalid argument / size out of range            */
    WH_E_HIP = -2,      /* HIP runtime error                               */
    WH_E_STATE = -3,    /* call made before its prerequisite (e.g. no active list) */
    WH_E_DIVERGED = -4, /* |u| > 1e12 in the divergence guard              */
    WH_E_NOMEM = -5
};

/* heavy-data arrays of a pool */
enum { WH_BLOCK = 0, WH_TMP = 1, WH_MASK = 2, WH_WORK = 16 /* + slot */ };

/* order_discretization (rhs_ACM.f90:374-870) */
enum { WH_FD2 = 2, WH_FD4 = 4, WH_FD4_TW = 40, WH_FD6 = 6 };

/* ghost-patch operations */
enum { WH_PATCH_COPY = 0, WH_PATCH_RESTRICT = 1, WH_PATCH_PREDICT = 2 };

typedef struct wh_pool wh_pool;

typedef struct {
    int dim;            /* 2 or 3 */
    int bs[3];          /* interior points per direction (bs[2] ignored in 2D) */
    int g;              /* stored ghost width */
    int neq;            /* equations in hvy_block (n_eqn) */
    int nb_max;         /* blocks in the pool (params%number_blocks) */
    int n_work_slots;   /* hvy_work slots (nrhs_slots) */
    int n_tmp_eqn;      /* components of hvy_tmp (0: none) */
    int n_mask;         /* components of hvy_mask (6 for ACM, 0: none) */
    int mask_per_block; /* 1: hvy_mask(..., Nb); 0: a single shared mask block */
} wh_pool_desc;

/* type_params_acm subset (LIB/EQUATION/ACMnew/module_ACM.f90:56-149) */
typedef struct {
    int dim;
    int scheme;         /* WH_FD* */
    int skew_symmetry;
    int penalization;
    int use_sponge;
    double c_0, nu, gamma_p, C_eta, C_sponge;
    double u_mean_set[3];
    double CFL, CFL_nu, CFL_eta;
} wh_acm_params;

/* One ghost patch: copy / restrict / predict a box of the sender into a box
 * of the receiver.  Indices are 0-based and ghost-inclusive.  For COPY the
 * sender box has the receiver box's extent; for RESTRICT the sender is read
 * at stride 2 from send_lo; PREDICT is evaluated from the coarser sender
 * (see wh_sync_ghosts).  The table is the device form of the reference's
 * (sender, receiver, relation, level-diff) list and ijkPatches boxes
 * (LIB/MPI/synchronize_ghosts_generic.f90:352-694, module_mpi.f90:318-396). */
typedef struct {
    int32_t recv, send;        /* hvy ids */
    int32_t recv_lo[3];
    int32_t send_lo[3];
    int32_t n[3];
    int32_t op;                /* WH_PATCH_* */
    int32_t stage;             /* 1..3: sync stage the patch belongs to */
    int32_t aux;               /* op-specific (PREDICT: parity/offset code) */
} wh_patch;

const char *wh_last_error(void);
int wh_version(void);
/* Free and total memory of the current device, through the library's own
 * HIP runtime: what a host sizes its pools from, as allocate_forest turns
 * --memory into a block count (LIB/MESH/allocate_forest.f90:160-196). */
int wh_device_memory(size_t *free_bytes, size_t *total_bytes);

/* ---- pool: allocate_forest (LIB/MESH/allocate_forest.f90:1-381) -------- */
int wh_pool_create(wh_pool **out, const wh_pool_desc *desc);
int wh_pool_destroy(wh_pool *pool);
/* bytes of one block of array `array` */
int wh_pool_block_bytes(const wh_pool *pool, int array, size_t *bytes);
/* copy `count` consecutive hvy ids starting at `first` between host and device */
int wh_pool_upload(wh_pool *pool, int array, int first, int count, const double *host);
int wh_pool_download(wh_pool *pool, int array, int first, int count, double *host);
/* device address of an array (for zero-copy interop on the same device) */
int wh_pool_device_ptr(wh_pool *pool, int array, void **dptr);
int wh_pool_fill(wh_pool *pool, int array, double value);
int wh_synchronize(wh_pool *pool);

/* ---- topology ----------------------------------------------------------- */
/* Active list hvy_active(1:hvy_n) (0-based ids) and, per listed block, its
 * origin and spacing (get_block_spacing_origin, get_block_spacing_origin.f90:6-19):
 * x0dx[6*k + 0..2] = x0, x0dx[6*k + 3..5] = dx. */
int wh_set_active(wh_pool *pool, const int32_t *hvy_active, int n, const double *x0dx);
/* Ghost patch table of one sync case (0: sync_ghosts_RHS_tree "full_leaf",
 * 1: sync_ghosts_tree "full") for ghost width gs. */
int wh_set_patches(wh_pool *pool, int sync_case, const wh_patch *patches, int npatch);

/* ---- hot path ----------------------------------------------------------- */
/* sync_ghosts_RHS_tree / sync_ghosts_tree (synchronize_ghosts_generic.f90:121-174)
 * on array `array` (WH_BLOCK, WH_TMP or WH_WORK+slot) using the patch table
 * of `sync_case`. */
int wh_sync_ghosts(wh_pool *pool, int array, int sync_case);

/* RHS_wrapper local stage (RHS_wrapper.f90:124-142 -> RHS_ACM rhs_ACM.f90:236-277):
 * rhs of every active block of `in_array` into work slot `out_slot`. */
int wh_rhs_acm(wh_pool *pool, const wh_acm_params *p, double time, int in_array, int out_slot);

/* integral-stage divergence guard (rhs_ACM.f90:133-146): returns
 * WH_E_DIVERGED if any interior |u| exceeds 1e12; *maxabs gets the max. */
int wh_divergence_guard(wh_pool *pool, int in_array, double *maxabs);

/* GET_DT_BLOCK_ACM over all active blocks and its minimum
 * (module_ACM.f90:617-691, calculate_time_step.f90:28-48). */
int wh_dt_reduce(wh_pool *pool, const wh_acm_params *p, double *dt_out);

/* RungeKuttaGeneric building block (runge_kutta_generic.f90:63-67, 90-112,
 * 136-154): interior of `out_array` = work[base_slot] + sum_i coef[i]*work[slot[i]]
 * over all active blocks, terms added in the given order; with base_slot < 0
 * the base is `out_array` itself.  Also used for the u0 copy (nterms = 0). */
int wh_rk_combine(wh_pool *pool, int out_array, int base_array, int nterms,
                  const int *term_arrays, const double *coefs);

/* One full RungeKuttaGeneric step: sync + (dt given) + stages + final update.
 * `path` 0: reference-structured (one kernel per operation, every active
 * block in one launch); 1: one launch per block ("plain", as the reference
 * loops over blocks); 2: fused stage kernels (RHS + RK update in one launch
 * per stage, needs a tableau whose stage rows have only a sub-diagonal
 * entry, e.g. classic RK4).  butcher is (nrows x nrows) row-major. */
int wh_rk_step(wh_pool *pool, const wh_acm_params *p, const double *butcher, int nrows,
               double time, double dt, int path);

/* Time controls of calculate_time_step (calculate_time_step.f90:22-118) and
 * the main-loop stop time (main.f90:293), for the device-side time loop. */
typedef struct {
    double time_max, dt_fixed, dt_max;
    double write_time, write_time_first, tsave_stats, write_backup_time;
    int write_fixed_time; /* write_method == "fixed_time" */
    int write_backup_n;
} wh_time_params;

int wh_time_set(wh_pool *pool, double time, int iteration);
int wh_time_get(wh_pool *pool, double *time, int *iteration, double *dt_last);
/* wh_time_fetch queues a copy of the device time, iteration and last dt to
 * pinned host memory on the pool's stream and returns at once;
 * wh_time_fetched waits for that copy only (not for later work) and returns
 * the values: the host learns a device-side step's dt without a stream
 * synchronisation of its own (main.f90's time bookkeeping, kept on the
 * device). */
int wh_time_fetch(wh_pool *pool);
int wh_time_fetched(wh_pool *pool, double *time, int *iteration, double *dt_last);

/* `nsteps` main-loop time steps with dt, its clamps and the time advance kept
 * on the device: per step the block-dt reduction, the clamps, a sync and the
 * fused stage kernels (path 2 of wh_rk_step).  use_graph = 1 captures one
 * step into a hipGraph once and replays it (re-captured when the active list,
 * the patch tables or the parameters change).  Steps at or beyond time_max
 * take dt = 0 and leave the state and iteration unchanged. */
int wh_run_steps(wh_pool *pool, const wh_acm_params *p, const double *butcher, int nrows,
                 const wh_time_params *tp, int nsteps, int use_graph);

/* ---- block groups: independent forests sharing one pool ----------------
 * The reference runs one forest per process; a batch of independent solves
 * (an ensemble, or the requests of a serving loop) can instead share one pool
 * and one launch per operation.  Each active block belongs to a group (its
 * forest); groups never exchange data (their ghost patches stay inside the
 * group) and each has its own dt.
 *
 * wh_set_block_groups: the group of every active block, in active-list order
 * (call after wh_set_active, which clears the groups). */
int wh_set_block_groups(wh_pool *pool, const int32_t *group, int n, int ngroups);
/* calculate_time_step's block-dt reduction per group (calculate_time_step.f90:
 * 28-48, GET_DT_BLOCK_ACM module_ACM.f90:617-691): dt_out[ngroups], before
 * the host clamps. */
int wh_dt_reduce_groups(wh_pool *pool, const wh_acm_params *p, double *dt_out);
/* RungeKuttaGeneric (runge_kutta_generic.f90:1-156) for every group at once,
 * group i with its own dt_groups[i]; path 2 (fused stages) or 1 (one launch
 * per block).  Each block's arithmetic is that of wh_rk_step with its
 * group's dt, so a group's result does not depend on the other groups. */
int wh_rk_step_groups(wh_pool *pool, const wh_acm_params *p, const double *butcher, int nrows,
                      const double *dt_groups, int path);

/* Physics constants per block group (independent forests of one pool may
 * differ in c_0, nu, gamma_p, C_eta, C_sponge, u_mean_set, CFL numbers,
 * penalization and sponge; they share dim, scheme and skew_symmetry, which
 * select the kernels).  params[ngroups], ngroups = the count given to
 * wh_set_block_groups (which, like wh_set_active, clears them); ngroups = 0
 * clears.  Used by wh_dt_reduce_groups and wh_batch_rk_step in place of the
 * constants of their `p` argument (type_params_acm, module_ACM.f90:56-149,
 * one per forest in the reference). */
int wh_set_group_params(wh_pool *pool, const wh_acm_params *params, int ngroups);

/* ---- batches: RungeKuttaGeneric over several pools at once -------------
 * The reference steps one forest per process (runge_kutta_generic.f90:1-156,
 * main.f90:350).  A batch steps the block groups of several pools of one
 * dimension (any block sizes, e.g. one pool per Bs class and host thread)
 * together: per RK stage one ghost fill per sync stage over the patch tables
 * of all pools and ONE fused stage kernel over the active blocks of all
 * pools (persistent kernels that read every pool's own device tables --
 * active list, origins/spacings, block groups, patch table -- through one
 * small descriptor per pool; the fused path's stage sequence, wh_rk_step
 * path 2).  Each pool needs its active list, patch table (case 0) and block
 * groups set, three work slots, and a mask where a group is penalised.  Per
 * block the arithmetic is that of wh_rk_step_groups, so each forest's result
 * does not depend on the batch.  The batch has its own stream, ordered after
 * the work queued on every pool's stream and before their later work; a
 * pool's next table change waits until the step has read its tables. */
typedef struct wh_batch wh_batch;
int wh_batch_create(wh_batch **out, wh_pool *const *pools, int npools);
int wh_batch_destroy(wh_batch *batch);
/* One step of every group of every pool: dt_groups[] holds one dt per group,
 * the groups of pool 0 first, then those of pool 1, ...  The constants come
 * from a pool's wh_set_group_params where set, else from *p.  use_graph = 1
 * replays the step from a hipGraph captured once (re-captured only when the
 * per-group arrays grow or the tableau, scheme or block sizes change); the
 * pool descriptors, per-group dt and constants are uploaded before each
 * replay (no memset nodes). */
int wh_batch_rk_step(wh_batch *batch, const wh_acm_params *p, const double *butcher, int nrows,
                     const double *dt_groups, int use_graph);
/* Average duration of one launch of the batch's fused stage kernel over
 * `reps` RK stage sequences back to back on the last step's tables, one HIP
 * event pair around them; the last stage writes a work slot, not the state.
 * *blocks: the blocks per launch.  No counterpart in the reference
 * (measurement). */
int wh_batch_time_stage_kernel(wh_batch *batch, const wh_acm_params *p, const double *butcher, int nrows,
                               int reps, double *ms, int *blocks);
int wh_batch_synchronize(wh_batch *batch);
/* The batch's 2D stage kernel (no counterpart in the reference):
 * WH_TUNE_K2D_STAGE 0 = one block per workgroup at a time, the first
 * point's operands loaded before the tile barrier; 1 = software pipelined
 * (the next block's metadata and halo tile in flight while a block is
 * computed; used where every tile is at most 32 x 32, i.e. Bs <= 28 with
 * FD4); 2 = 0 without the early loads; 3 / 4 = 2 with its registers bounded
 * for 6 / 8 waves per SIMD.  All give the same results bit for bit. */
int wh_batch_set_tuning(wh_batch *batch, int key, int value);
int wh_batch_get_tuning(wh_batch *batch, int key, int *value);

/* refineBlock (LIB/MESH/refinementExecute.f90:4-116) for a list of
 * (mother, daughter, digit) triples: daughter `digit` of a mother takes its
 * quadrant/octant of the mother's interpolated grid (prediction of order
 * `order` = 2/4/6 over the whole ghost-inclusive block, so the mothers' ghost
 * nodes must be synchronised first).  Daughters and mothers must be distinct
 * pool slots.  Ordered on the pool's stream; may return before the device
 * has finished. */
int wh_refine_blocks(wh_pool *pool, int array, const int32_t *triples, int ntriples, int order);

/* ---- host forest (metadata of one tree) --------------------------------
 * The reference keeps this on the host (lgt_block, hvy_neighbor, ...); these
 * helpers build the device tables from a block list (treecode, level, hvy id).
 * Treecodes as in LIB/TREE/module_treelib.f90 (dim bits per level, level l's
 * digit at bit offset (jmax-l)*dim, digit bits x=2, y=1, z=4). */
typedef struct wh_forest wh_forest;
int wh_forest_create(wh_forest **out, int dim, int jmax, const int *bs, int g,
                     const double *domain, const int *periodic, int nb_max);
int wh_forest_destroy(wh_forest *forest);
/* The forest's pool slots become [first, first + nb_max) (default first = 0),
 * so that several forests can share one pool; only while it has no blocks. */
int wh_forest_set_slot_base(wh_forest *forest, int first);
/* Replaces the block list.  WH_E_ARG for a slot id outside the forest's
 * range or used twice, or a level outside [0, jmax] (the forest is left as it
 * was), and for two blocks with one (level, treecode) (the forest is left
 * empty). */
int wh_forest_set_blocks(wh_forest *forest, int n, const int64_t *treecodes,
                         const int32_t *levels, const int32_t *hvy_ids);
int wh_forest_size(const wh_forest *forest, int *n);
/* blocks in list order; x0dx (6 per block) as get_block_spacing_origin */
int wh_forest_get_blocks(const wh_forest *forest, int cap, int64_t *treecodes, int32_t *levels,
                         int32_t *hvy_ids, double *x0dx);
/* Ghost patches of a 2:1-balanced leaf forest for ghost depth gminus/gplus
 * (replaces prepare_ghost_synch_metadata + ghosts_setup_patches for the
 * "full_leaf"/"full" cases, synchronize_ghosts_generic.f90:352-694,
 * calc_data_bounds.f90:46-189): each ghost box is filled from the leaf that
 * covers it -- COPY (stage 1), RESTRICT by decimation (stage 2) or PREDICT of
 * `order` (stage 3).  out may be NULL to query *npatch. */
int wh_forest_patches(const wh_forest *forest, int gminus, int gplus, int order, wh_patch *out,
                      int cap, int *npatch);
/* Refine blocks (by hvy id): daughters on free pool slots, mothers leave the
 * list; writes (mother, daughter, digit) triples for wh_refine_blocks. */
int wh_forest_refine(wh_forest *forest, int n, const int32_t *hvy_ids, int32_t *triples, int cap,
                     int *ntriples);

/* ---- CDF wavelets (LIB/WAVELETS/module_wavelets.f90) ---------------------
 * Filters on [lo, hi], the coefficient of tap i at [i - lo]. */
#define WH_WAVELET_TAPS 33
typedef struct {
    int X, Y, lifted, specialised;
    int hd_lo, hd_hi, gd_lo, gd_hi, hr_lo, hr_hi, gr_lo, gr_hi;
    double HD[WH_WAVELET_TAPS], GD[WH_WAVELET_TAPS], HR[WH_WAVELET_TAPS], GR[WH_WAVELET_TAPS];
    int g;       /* ghost nodes the filters need (g_min) */
    int g_rhs;   /* max(X/2, FD reach) */
    int order;   /* predictor order (multiresolution_<X>th) */
    int Nscl, Nscr, Nwcl, Nwcr, Nreconl, Nreconr, Ndep2l, Ndep2r;  /* coarse extension */
} wh_wavelet;

/* setup_wavelet (module_wavelets.f90:1031-1430) for "CDFXY" with X in
 * {2,4,6}, Y in {0,2,4,6}, Y <= X; fd_max_size is the reach of the FD scheme
 * (0 when none is set).  Host only. */
int wh_wavelet_setup(const char *name, int fd_max_size, wh_wavelet *out);
/* The wavelet the pool's transform kernels use. */
int wh_pool_set_wavelet(wh_pool *pool, const wh_wavelet *w);

/* The transform, box, patch-list, mask, wh_sync_tree and wh_adapt_tree calls
 * below are ordered on the pool's stream and may return before the device
 * has finished (the host lists are staged through pinned memory, so the
 * caller's buffers are free on return; wh_adapt_tree's new leaf list is in
 * the forest on return); downloads and wh_detail_blocks return with the
 * stream synchronised.
 *
 * Per-block transforms of the listed hvy ids, src and dst distinct arrays;
 * only the region the reference computes is written (dst keeps the rest):
 * waveletDecomposition_optimized_block (wavelet_decomposition_reconstruction.f90:23-389),
 * waveletReconstruction_optimized_block (:426-841),
 * blockFilterXYZ_vct with HD and do_restriction (module_wavelets.f90:307-401).
 * Work slot 0 is used as scratch. */
int wh_fwt_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids, int n);
int wh_iwt_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids, int n);
int wh_filter_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids, int n);
/* whole-block copies (hvy_tmp = hvy_block and back) */
int wh_copy_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids, int n);
/* threshold_block's detail with eps_norm = Linfty (threshold_block.f90:1-130):
 * detail[k*neq + c] = max over the interior of |WC| (SCs zeroed); components
 * with groups[c] = l >= 2 take the max over the group, groups[c] = 0 gives 0. */
int wh_detail_blocks(wh_pool *pool, int array, const int32_t *hvy_ids, int n,
                     const int32_t *groups, double *detail);

/* params%eps_norm (ini_file_to_params.f90:494; "Linfty" by default) */
enum { WH_NORM_LINFTY = 0, WH_NORM_L1 = 1, WH_NORM_L2 = 2, WH_NORM_H1 = 3 };
/* threshold_block's detail (threshold_block.f90:69-124) with the
 * wavelet_renorm_block scaling of eps_norm (module_wavelets.f90:1848-1960)
 * for blocks on levels[k] (level_ref 0, as threshold_block calls it): L1
 * 2^(-(J+1)dim) and L2 2^(-(J+1)dim/2) on every WC, then /4 (L1) or /2 (L2)
 * per direction in which the point lies on an SC position; H1 in 3D
 * 2^(J/2) and 2^(2/3) per such direction, in 2D as Linfty.  WH_NORM_LINFTY
 * is wh_detail_blocks. */
int wh_detail_blocks_norm(wh_pool *pool, int array, const int32_t *hvy_ids, const int32_t *levels, int n,
                          const int32_t *groups, int eps_norm, double *detail);
/* componentWiseNorm_tree(params, hvy_block, tree_ID, eps_norm, norm)
 * (OPERATORS/componentWiseNorm_tree.f90:1-200, norm_case default) over the
 * listed blocks (the caller passes the leaves): per component of group 1
 * the sum over the blocks of dx^dim * sum(u^2) (L2, H1) or dx^dim *
 * sum(|u|) (L1) over the interior, the components of a group l >= 2
 * together, then the square root for L2/H1; Linfty: the max of |u|.  dx of
 * block k is 2^-levels[k] * domain[d] / Bs[d]; groups NULL: all 1.  No
 * clamp (coarseningIndicator_tree's norm <= 1e-9 -> 1 is the caller's). */
int wh_component_norm(wh_pool *pool, int array, const int32_t *hvy_ids, const int32_t *levels, int n,
                      const int32_t *groups, int eps_norm, const double *domain, double *norm);

/* Coarse-extension boxes (coarseExtensionManipulateWC/SC_block,
 * module_wavelets.f90:877-1027): inclusive ghost-inclusive index box of block
 * hvy; WH_BOX_ZERO_WC zeroes the points that are not SC in every direction,
 * WH_BOX_COPY_SC copies the SC points from src_array. */
enum { WH_BOX_ZERO_WC = 0, WH_BOX_COPY_SC = 1 };
typedef struct {
    int32_t hvy;
    int32_t lo[3], hi[3];
    int32_t op;
} wh_box;
int wh_apply_boxes(wh_pool *pool, int array, int src_array, const wh_box *boxes, int n);

/* Run an explicit patch list on `array` (stages in order).  Bits 8-9 of a
 * patch's op select where its sender is read: 0 `array`, 1 `alt1`, 2 `alt2`
 * (the source-array choice of xfer_block_data.f90:321-437). */
int wh_run_patches(wh_pool *pool, int array, int alt1, int alt2, const wh_patch *patches, int n);

/* ---- penalisation mask (create_mask_2D_ACM, LIB/EQUATION/ACMnew/create_mask.f90:183-320) */
typedef struct {
    int geometry;                    /* 0 none, 1 cylinder (2D: draw_circle), 2 sphere-fixed (3D: draw_sphere) */
    double x_cntr[3], R_cyl;
    double smoothing_width, smoothing_safety;  /* cosine smoothing (module_ACM.f90:467-471) */
    int use_sponge, sponge_type;     /* sponge_type 0: "rect", 1: "p-norm" (sponge.f90:4-233) */
    double L_sponge, p_sponge;
    double domain[3];
} wh_mask_params;
/* CREATE_MASK_meta "all-parts" (module_physics_metamodule.f90:38-84; 2D
 * create_mask_2D_ACM create_mask.f90:183-320, 3D create_mask_3D_acm :6-174) for the
 * listed blocks into WH_MASK (6 components: chi, u_s, colour, sponge); x0dx
 * per listed block as in wh_set_active.  n < 0: all active blocks with the
 * pool's own origins and spacings. */
int wh_create_mask(wh_pool *pool, const wh_mask_params *mp, const int32_t *hvy_ids, int n,
                   const double *x0dx);

/* adapt_tree (LIB/MESH/adapt_tree.f90:11-255) on the forest's leaf blocks in
 * WH_BLOCK: full tree down to jmin, level-wise decomposition, coarsening
 * indicator, gradedness, coarse extension, reconstruction, pruning and the
 * closing sync_ghosts_tree.  The forest's block list is replaced by the new
 * leaves.  WH_TMP must hold >= neq components; work slots 0 and 1 are used
 * as scratch.  The pool needs free slots for the mother blocks. */
enum { WH_IND_EVERYWHERE = 0, WH_IND_THRESHOLD = 1, WH_IND_EXTERNAL = 2 };
typedef struct {
    int jmin, jmax;
    int indicator;                  /* WH_IND_* */
    int use_ce;                     /* useCoarseExtension */
    int force_maxlevel_dealiasing;
    double eps;
    double norm[8];                 /* per component: threshold eps * norm */
    int32_t groups[8];              /* threshold_state_vector_component */
    int eps_normalized;             /* norm = eps_norm of the leaves (componentWiseNorm_tree) */
    int threshold_mask;             /* coarseningIndicatorMask_tree, with `mask` */
    int security_zone;              /* addSecurityZone_CE_tree */
    double domain[3];
    wh_mask_params mask;
    /* wavelet_decompose_full_tree's branch: 0 the reference's rule
     * (leaf-first iff every Bs >= 3 max|HD tap|, adapt_tree.f90:321-322),
     * 1 leaf-first, 2 level-wise; both give the same decomposition */
    int decompose_branch;
    int eps_norm;                   /* WH_NORM_*: the details' renormalisation and the norm */
} wh_adapt_params;
int wh_adapt_tree(wh_pool *pool, wh_forest *forest, const wh_adapt_params *p);

/* adapt_tree with indicator "nothing (external)" (adapt_tree.f90:11-255,
 * coarseningIndicator_tree.f90:55): the host keeps the refinement flags and
 * passes them in, one per leaf in the forest's list order
 * (wh_forest_get_blocks): -1 coarsen, 0 stay, 1 refine (a refine flag is
 * kept as a status; adapt_tree never refines).  p->indicator must be
 * WH_IND_EXTERNAL; the flags then go through respectJmaxJmin_tree and
 * ensureGradedness_tree as the reference's do (no mask criterion, no
 * security zone -- adapt_tree.f90:127 skips it for this indicator).  n must
 * equal the forest's block count. */
int wh_adapt_tree_flags(wh_pool *pool, wh_forest *forest, const wh_adapt_params *p,
                        const int32_t *leaf_flags, int n);

/* coarseningIndicator_block -> threshold_block on the listed blocks
 * (coarseningIndicator_block.f90:12-108, threshold_block.f90:1-130):
 * waveletDecomposition_optimized_block of src_array into dst_array, then per
 * block the detail (max over the interior of |WC| per component, grouped by
 * groups[] as in wh_detail_blocks) into detail_out[k*neq + c], and the flag
 * flags_out[k] = -1 if every detail <= thresh[c] (eps * norm, Linfty), else
 * 0.  The host decides what to do with the flags (wh_adapt_tree_flags).
 * Ghost nodes of src_array must be synchronised.  Returns with the stream
 * synchronised. */
int wh_fwt_threshold(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids, int n,
                     const double *thresh, const int32_t *groups, int32_t *flags_out,
                     double *detail_out);

/* wh_fwt_threshold without allocating anything.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and checks as wh_fwt_threshold; the
 * block ids and the per-block details pass through a fixed device array of
 * the library (one per GPU, held until the process ends) instead of the
 * pool's staging and result buffers, so n is limited to
 * WH_NOALLOC_MAX_BLOCKS and neq to 8.  The array's address on a GPU is set
 * up when the first pool on that GPU is created.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several host threads)
 * race.  Returns with the pool's stream synchronised. */
#define WH_NOALLOC_MAX_BLOCKS 16384
int wh_fwt_threshold_noalloc(wh_pool *pool, int src_array, int dst_array, const int32_t *hvy_ids,
                             int n, const double *thresh, const int32_t *groups,
                             int32_t *flags_out, double *detail_out);

/* sync_ghosts_tree (synchronize_ghosts_generic.f90:121-144) on a leaf forest
 * with all ghosts, restriction filtered with HD for lifted wavelets
 * (restrict_copy_at_CE, restrict_predict_data.f90:101-170). */
int wh_sync_tree(wh_pool *pool, wh_forest *forest, int array);

/* sync_ghosts_tree and adapt_tree for n forests that share the pool
 * (independent solves, each in its own slot range): per forest the same
 * sequence of steps as wh_sync_tree / wh_adapt_tree, with every device
 * operation launched once for all forests over their concatenated lists and
 * the coarsening indicator's reductions of all forests read back together.
 * No counterpart in the reference (one forest per process there).
 * wh_adapt_trees: params[i] and status[i] per forest; status[i] is forest
 * i's own error (its checks, its slots), which leaves its block list
 * unchanged while the others go on; the return value is an error only when
 * an operation shared by all forests failed. */
int wh_sync_trees(wh_pool *pool, wh_forest *const *forests, int n, int array);
int wh_adapt_trees(wh_pool *pool, wh_forest *const *forests, int n, const wh_adapt_params *params,
                   int *status);

/* The rest of a main-loop iteration's mesh work for n forests that share the
 * pool (independent solves; no counterpart in the reference, which runs one
 * forest per process), each call made once per pool instead of once per
 * forest.  status[i] is forest i's own error; the return value is an error
 * only when an operation shared by all forests failed.
 * wh_refine_trees: refine_tree with the "everywhere" indicator
 * (LIB/MESH/refine_tree.f90, refinementExecute.f90:4-116): every leaf of
 * forest i below level jmax[i] gets its 2^dim daughters (block lists on the
 * host, then one wh_refine_blocks launch for all forests); a forest without
 * the free slots for all its daughters at once is left unchanged with
 * WH_E_NOMEM.
 * wh_create_masks: createMask_tree (LIB/PHYSICS/create_mask.f90:6-320) on
 * the leaves of each forest with its own params[i].
 * wh_set_tables_from_forests: the tables of the RK step -- active list and
 * origins/spacings of every forest's leaves (forest after forest, in leaf
 * order), block group k for the k-th forest taken, and the RHS ghost-patch
 * table of depth g_rhs with predictors of `order`
 * (synchronize_ghosts_generic.f90:155 sync_ghosts_RHS_tree) -- as
 * wh_set_active + wh_set_block_groups + wh_set_patches(pool, 0, ...) would
 * set them from wh_forest_get_blocks / wh_forest_patches; a forest whose
 * patches cannot be generated is left out (status[i], nblocks[i] = 0);
 * nblocks[i]: forest i's leaves in the tables. */
int wh_refine_trees(wh_pool *pool, wh_forest *const *forests, int n, const int32_t *jmax, int order, int array,
                    int *status);
int wh_create_masks(wh_pool *pool, wh_forest *const *forests, int n, const wh_mask_params *params, int *status);
int wh_set_tables_from_forests(wh_pool *pool, wh_forest *const *forests, int n, int g_rhs, int order, int *status,
                               int *nblocks);

/* Kernel tuning knobs.  WH_TUNE_K3D_YSPLIT: workgroups per block of the 3D
 * stage kernel (each owns Bs1/ysplit rows of (x,y) columns and marches in z);
 * 0 selects the flat one-point-per-thread kernel, -1 automatic.
 * WH_TUNE_K3D_ZSPLIT (default 1): the z range of a block is cut into zsplit
 * pieces marched by separate workgroups (a split that does not divide Bs2
 * into pieces of at least 2R planes, R the stencil reach, acts as 1).
 * WH_TUNE_K3D_WAVES (default 1): 3 selects the build of the z-marching
 * kernel whose register allocation allows 3 waves per SIMD, 4 the build of
 * the flat kernel allowing 4 (both may spill a few registers).
 * WH_TUNE_K3D_XCD (default 0): 1 remaps the flat 3D kernel's workgroups so
 * that those of one block run on one XCD (one L2).
 * WH_TUNE_K2D_STAGE: 2D stage kernel variant: 0 one workgroup per block;
 * 1 (default) the same, loading the mask and stage operands of each
 * thread's first point before the halo-tile barrier; 2, 3, 4 a persistent
 * kernel with 2, 4, 8 workgroups per compute unit that loads the next
 * block's halo tile while computing the current one (where the tile fits).
 * Every variant computes the same values bit for bit.
 * WH_TUNE_GATHER (default 1): where the RHS ghost-patch table of a 2D pool
 * consists of same-level copies only (a uniform grid) that cover the halo
 * of every active block, the fused RK stages read each block's halo tile
 * straight from its neighbours' interiors and the step launches no ghost
 * fill (the values are the same copies; the state's own ghost nodes are
 * then not refreshed by the step); 0 fills the ghost nodes first. */
enum {
    WH_TUNE_K3D_YSPLIT = 1,
    WH_TUNE_K2D_STAGE = 2,
    WH_TUNE_K3D_ZSPLIT = 3,
    WH_TUNE_K3D_WAVES = 4,
    WH_TUNE_K3D_XCD = 5,
    WH_TUNE_GATHER = 6
};
int wh_set_tuning(wh_pool *pool, int key, int value);
int wh_get_tuning(wh_pool *pool, int key, int *value);

/* Autotuner: times every valid variant of the RHS/stage kernel for this
 * pool's (Bs, g, neq) and parameters on the current active blocks, keeps the
 * fastest and returns it (3D: the WH_TUNE_K3D_YSPLIT value, with the
 * WH_TUNE_K3D_ZSPLIT, WH_TUNE_K3D_WAVES and WH_TUNE_K3D_XCD values set as
 * well; 2D: the
 * WH_TUNE_K2D_STAGE value).  With three or more work slots the timed kernel
 * is the fused RK stage kernel and work slots 0 and 1 are overwritten,
 * otherwise the RHS kernel and work slot 1. */
int wh_autotune(wh_pool *pool, const wh_acm_params *p, int reps, int *best_value, double *best_ms);

/* Average duration of one launch of the fused RK stage kernel (the tuned
 * variant) over `reps` classic-RK4 stage sequences (first, two middle and
 * last stage, 4*reps launches) back to back on the pool's stream, bracketed
 * by one pair of HIP events, on the current active blocks: the stage
 * kernel's time without the dispatch gaps a per-launch event pair adds.
 * Reads the state, writes only work slots 0-2 (overwritten); needs three
 * work slots.  No counterpart in the reference (measurement). */
int wh_time_stage_kernel(wh_pool *pool, const wh_acm_params *p, int reps, double *ms);

/* HIP-event timing of kernel kinds (0: RHS or fused stage kernel, 1: ghost
 * sync) on the pool's stream: average ms per launch since the last reset.
 * Timing serialises host and device; enable it only for measurement. */
int wh_kernel_timing_enable(wh_pool *pool, int on);
int wh_kernel_timing(wh_pool *pool, int which, double *avg_ms, int *count);

#ifdef __cplusplus
}
#endif
#endif
