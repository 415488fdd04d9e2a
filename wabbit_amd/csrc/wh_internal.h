// wh_internal.h -- shared declarations of the gfx950 block time-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wabbit_hip.h"

namespace wh {

void set_error(const std::string &msg);

#define WH_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::wh::set_error(std::string(#call) + ": " + hipGetErrorString(e_));      \
            return WH_E_HIP;                                                         \
        }                                                                            \
    } while (0)

#define WH_CHECK(cond, code, msg)                                                    \
    do {                                                                             \
        if (!(cond)) {                                                               \
            ::wh::set_error(msg);                                                    \
            return code;                                                             \
        }                                                                            \
    } while (0)

// Scalars the kernels read/write in device memory (one array per pool).
enum DevScalar {
    DS_DTMIN_BITS = 0,   // uint64 bit pattern of min block dt (atomicMin)
    DS_MAXABS_BITS = 1,  // uint64 bit pattern of max |u| (atomicMax)
    DS_TIME = 2,         // current time (device-side time loop)
    DS_DT = 3,           // dt of the current step
    DS_ITER = 4,         // iteration counter (stored as double)
    DS_COUNT = 8
};

// Kernel-side copy of the ACM parameters.
struct AcmDev {
    double c02, nu, gamma_p, Cinv, Csp_inv;
    double ums[3];
    double CFL, CFL_nu, CFL_eta, C_eta, C_sponge, c_0;
    int dim, scheme, skew, penal, sponge;
};

// Block geometry of a pool as the kernels see it.
struct Geo {
    int bs0, bs1, bs2, g, nx, ny;
    size_t blk, cs, mblk;
    int mask_per_block;
};

// A 3D ghost patch holds up to a few thousand points; the ghost fill cuts
// every patch into chunks of SYNC_CHUNK points (its first point `first`),
// one wavefront per chunk (four points per lane), so that the work spreads
// evenly and a wavefront finds its patch with one load.
constexpr int SYNC_CHUNK = 256;
struct SyncChunk {
    int32_t patch, first;
};

// One pool of a multi-pool launch: its arrays (0 hvy_block, 1..3 work slots
// 0..2), its mask (or null) and geometry, and where the launch finds the
// pool's own device tables -- active list, origin/spacing and block group
// per active position, the RHS ghost-patch table with its slot prefix sum --
// and the pool's place in the launch: its first entry (entries are the
// pools' active lists one after the other), its first block group, and per
// sync stage s its patch range [p0[s], p1[s]) and chunk range [c0[s],
// c1[s]), and where these start in the launch: its first chunk s0[s] among
// the stage's chunks and its first patch q0[s] among the stage's patches
// (the pools' ranges one after the other).
struct MPoolDev {
    double *arr[4];
    const double *mask;
    Geo G;
    int neq, dim;
    const int32_t *act;
    const double *x0dx;
    const int32_t *grp;
    const wh_patch *pt;
    const SyncChunk *ch;
    int ent0, goff;
    int p0[3], p1[3], c0[3], c1[3];
    int64_t s0[3], q0[3];
};

// Entry k of a multi-pool launch, resolved once per step (launch_entries):
// its pool, hvy id, block group in the launch (goff + grp) and position in
// the pool's active list.  The stage kernels read one record per block
// instead of searching the pools and then reading the active list and the
// group table one after the other.
struct MEnt {
    int32_t pool, b, g, j;
};

// The RK stage of a multi-pool fused stage launch: input / output array
// selectors into MPoolDev::arr, the epilogue mode and tableau coefficients.
struct MStage {
    int in_sel, out_sel, sout_sel, mode, use_a, use_b;
    int full_rows;  // the next stage input written in whole rows (stage_full_rows)
    double a, b;
};

// Whether a fused RK stage over nblk blocks of blk_bytes each writes the
// next stage input in whole rows (zeros in the ghost columns of the interior
// rows): worth it where the stage's arrays (input, accumulator, state,
// output) exceed the 256 MB last-level cache, so that partially written
// cache lines would go to HBM and be merged there with their old contents;
// below that the lines stay in the cache and the extra bytes only cost.
inline bool stage_full_rows(double nblk, double blk_bytes) { return 4.0 * nblk * blk_bytes > 256.0 * (1 << 20); }

// Alternate source arrays of explicit patch lists (bits 8-9 of wh_patch::op).
struct SyncSrc {
    const double *alt[2] = {nullptr, nullptr};
    size_t blk_alt[2] = {0, 0};
};

// Device workspace of the wavelet and adapt operations (grown on demand).
struct WaveletWS {
    bool set = false;
    wh_wavelet w{};
    int32_t *d_ids = nullptr;
    size_t ids_cap = 0;
    wh_patch *d_patches = nullptr;
    size_t patch_cap = 0;
    wh_box *d_boxes = nullptr;
    size_t box_cap = 0;
    double *d_out = nullptr;
    size_t out_cap = 0;
    // staging of the host-built lists (block ids, boxes, patches): copied to
    // a pinned buffer and from there to the same offset of a device buffer,
    // so that the copies run asynchronously on the pool stream; the space is
    // reclaimed only after a stream synchronisation
    char *h_stage = nullptr, *d_stage = nullptr;
    size_t stage_cap = 0, stage_off = 0;
    size_t pend_lo = ~(size_t)0;  // start of pieces put with defer, not yet copied
    const int32_t *cur_ids = nullptr;  // device block list of the current wavelet op
};

// Host-side time controls (calculate_time_step.f90:58-118) evaluated on device.
struct TimeDev {
    double time_max, dt_fixed, dt_max, write_time, write_time_first, tsave_stats,
        write_backup_time;
    int write_fixed_time, write_backup_n, use_stats;
};

struct PatchSet {
    std::vector<wh_patch> host;     // sorted by (stage, receiver)
    wh_patch *d_patches = nullptr;
    int32_t *d_recv_off = nullptr;  // CSR over (stage, receiver) rows
    int32_t *d_recv_ids = nullptr;  // receiver hvy id per CSR row
    SyncChunk *d_chunks = nullptr;   // the patches cut into wavefront-sized chunks
    std::vector<int32_t> h_recv_off, h_recv_ids;
    std::vector<SyncChunk> h_chunks;
    // a 2D table of same-level copies only: per receiver hvy id the senders of
    // its 9 directions (x fastest, centre unused), all copying `gather_depth`
    // ghost layers; gather_ok false otherwise (set_patch_runs)
    bool gather_ok = false;
    int gather_depth = 0;
    std::vector<int32_t> h_gnb;
    int32_t *d_gnb = nullptr;
    size_t cap_gnb = 0;
    int stage_begin[5] = {0, 0, 0, 0, 0};  // stage s (1..3): patches [stage_begin[s], stage_begin[s+1])
    int chunk_begin[5] = {0, 0, 0, 0, 0};  // and chunks [chunk_begin[s], chunk_begin[s+1])
    int n = 0, nrecv = 0;
    int max_stage = 0;
    // device capacities (element counts); buffers are reused until they must grow
    size_t cap_patches = 0, cap_recv_off = 0, cap_recv_ids = 0, cap_chunks = 0;
};

}  // namespace wh

struct wh_pool {
    wh_pool_desc d{};
    int nx = 0, ny = 0, nz = 0;
    size_t cs = 0;        // points per component (nx*ny*nz)
    size_t blk = 0;       // doubles per block of hvy_block / hvy_work
    size_t tblk = 0;      // doubles per block of hvy_tmp
    size_t mblk = 0;      // doubles per block of hvy_mask
    double *block = nullptr, *work = nullptr, *tmp = nullptr, *mask = nullptr;
    double *scal = nullptr;       // wh::DS_COUNT doubles
    double *h_scal = nullptr;     // pinned mirror
    double *h_time = nullptr;     // pinned copy of the scalars by wh_time_fetch
    hipEvent_t ev_time = nullptr; // recorded after that copy
    bool time_fetch_pending = false;
    int32_t *d_active = nullptr;
    double *d_x0dx = nullptr;
    std::vector<int32_t> h_active;
    std::vector<double> h_x0dx;
    int nact = 0, act_cap = 0;
    wh::PatchSet ps[2];
    hipStream_t stream = nullptr;
    // timing
    bool timing = false;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // recorded after the last queued work that reads the active list, the
    // patch tables or the block groups: replacing them waits for this event
    // only, not for everything queued since
    hipEvent_t ev_tables = nullptr;
    bool tables_recorded = false;
    double t_ms[2] = {0, 0};
    int t_cnt[2] = {0, 0};
    // graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t topo_version = 0, graph_topo = ~0ull;
    wh_acm_params graph_p{};
    std::vector<double> graph_bt;
    wh::TimeDev graph_tc{};
    // tuning: y-split of the 3D stage kernel (-1 auto, 0 flat kernel)
    int k3d_ysplit = -1;
    int k3d_zsplit = 1;
    int k3d_waves = 1;  // 3: the z-marching kernel built for 3 waves per SIMD, 4: the flat one for 4
    int k3d_xcd = 0;    // 1: the flat 3D kernel's workgroups swizzled per XCD
    // tuning of the 2D stage kernel: 0 plain, 1 mask and epilogue operands
    // loaded before the tile barrier, 2..4 persistent pipelined kernel with
    // 2, 4, 8 workgroups per CU
    int k2d_stage = 1;
    // WH_TUNE_GATHER, and whether the current tables allow it (checked once
    // per topology version: every active block has its 8 senders)
    int gather = 1;
    uint64_t gather_topo = ~0ull;
    bool gather_all = false;
    int n_cu = 256;  // compute units of the pool's device
    int device = 0;  // the device the pool lives on
    // wavelet transforms / adapt_tree
    wh::WaveletWS ws;
    // block groups (independent forests in one pool): group of each active
    // block, per-group dt and its per-position copy for the stage kernels
    int ngrp = 0, grp_cap = 0, dtg_cap = 0;
    std::vector<int32_t> h_grp;
    int32_t *d_grp = nullptr;
    double *d_dtb = nullptr;
    unsigned long long *d_dtg = nullptr;
    // per-group physics constants (wh_set_group_params); ngacm = 0: none
    int ngacm = 0, gacm_cap = 0;
    std::vector<wh_acm_params> h_gparams;
    wh::AcmDev *d_gacm = nullptr;
};

namespace wh {

double *array_ptr(wh_pool *p, int array, size_t *blk_out, int *ncomp);

// launchers (wh_kernels.hip)
int launch_rhs(wh_pool *P, const AcmDev &a, const double *in, double *out, const int32_t *act,
               const double *x0dx, int nact);
int launch_sync(wh_pool *P, int sync_case, double *arr, int one_receiver);
int launch_combine(wh_pool *P, double *out, const double *base, int nterms,
                   const double *const *terms, const double *coefs, const int32_t *act,
                   int nact);
int launch_dt(wh_pool *P, const AcmDev &a);
int launch_maxabs(wh_pool *P, const double *in);
// dtb: per active-list position dt (block groups), or null for the pool's DS_DT
// gnb: the halo-gather table of a 2D pool (the ghost fill skipped), or null
int launch_fused_stage(wh_pool *P, const AcmDev &a, const double *tin, double *tout,
                       const double *u0, double *acc, double *sout, int mode, double acoef,
                       double bcoef, const double *dtb, const int32_t *gnb = nullptr);
// per-group minimum block dt into P->d_dtg (bit patterns) for the P->ngrp groups
int launch_dt_groups(wh_pool *P, const AcmDev &a);
int launch_dt_finalize(wh_pool *P, const TimeDev &tc);
int launch_time_advance(wh_pool *P);
int launch_set_scalar(wh_pool *P, int which, double v);

// the chunks of patch i of a list (its points cut into pieces of
// SYNC_CHUNK), appended in point order
inline void add_chunks(std::vector<SyncChunk> &out, int i, const wh_patch &q)
{
    const int n = q.n[0] * q.n[1] * q.n[2];
    for (int r = 0; r < n; r += SYNC_CHUNK) out.push_back(SyncChunk{i, r});
}
// ghost fill of one sync stage of a device patch list: 2D the patches
// [p0, p1), one wavefront per patch; 3D the chunks [c0, c1), one wavefront
// per chunk
int launch_sync_flat(wh_pool *P, double *arr, const wh_patch *d_patches, const SyncChunk *d_chunks, int p0, int p1,
                     int c0, int c1, const SyncSrc &S);

// wavelet workspace (wh_wavelet.hip)
void free_wavelet_ws(wh_pool *P);
// staging: stage_begin reserves `bytes` (aligned pieces included) for the
// following stage_put calls, synchronising the stream and reusing the
// buffer from its start when it is full; stage_put copies host data and
// returns its device address; stage_sync waits for the stream and frees all
// staged space
int stage_begin(wh_pool *P, size_t bytes);
// copies src into the reserved pinned space and returns the device address of
// the same offset; the host-to-device copy is issued at once, or with defer
// by the next stage_flush (one copy for consecutive deferred pieces of one
// reservation)
void *stage_put(wh_pool *P, const void *src, size_t bytes, bool defer = false);
int stage_flush(wh_pool *P);
// copies src into the reserved pinned space and returns that host address
// (for a direct host-to-device copy into a buffer of the caller's)
void *stage_host(wh_pool *P, const void *src, size_t bytes);
int stage_sync(wh_pool *P);
// device-side steps used by the adapt orchestration (wh_adapt.cpp); the
// host vectors are uploaded on the pool's stream before use
int run_patch_list(wh_pool *P, int array, int alt1, int alt2, const std::vector<wh_patch> &patches);
int run_blocks_op(wh_pool *P, int op, int src_array, int dst_array, const std::vector<int32_t> &ids);
int run_boxes(wh_pool *P, int array, int src_array, const std::vector<wh_box> &boxes);
// resolve the device address of the no-allocation entry point's scratch
// array on device dev (the current device); once per device
void noalloc_setup(int dev);
int run_detail(wh_pool *P, int array, const std::vector<int32_t> &ids, const int32_t *groups,
               std::vector<double> &detail);
enum { WOP_FWT = 0, WOP_IWT = 1, WOP_FILTER = 2, WOP_COPY = 3 };
// per-box component maxima (|v|, SCs zeroed with zero_sc, sqrt(v*v) for
// grouped components with sqrt_groups): out[k*neq + c]
// Reductions launched into a caller's device range d_out (no copy, no
// synchronisation), so that several of them come back in one round trip:
// results_reserve sizes the pool's result buffer for n doubles (before any
// of the launches), results_fetch copies them back and synchronises.
int launch_region_max(wh_pool *P, int array, const std::vector<wh_box> &boxes, bool zero_sc, bool sqrt_groups,
                      const int32_t *groups, double *d_out);
int launch_detail_list(wh_pool *P, int array, const std::vector<int32_t> &ids, const int32_t *groups,
                       double *d_out, bool typed = false);
// eps_norm (WH_NORM_*) reductions of boxes: Linfty as launch_region_max
// (zone: zero_sc and sqrt_groups); other norms: zone -- the per-type maxima
// of |v| (2^dim - 1 per box and component), else the per-box sums of |v|
// (L1) or v*v (L2, H1)
int launch_region_norm(wh_pool *P, int array, const std::vector<wh_box> &boxes, int kind, bool zone,
                       const int32_t *groups, double *d_out);
// wavelet_renorm_block (module_wavelets.f90:1848-1960, level_ref 0) of the
// max m of |v| over the points of type t (bit d: SC position in direction
// d) of a block on `level`, in the reference's order of operations
double renorm_max(double m, int t, int level, int dim, int kind);
// typed maxima of n blocks -> their renormed details out[k*neq + c]
// (grouped components as sqrt(r*r), threshold_block.f90:86-91)
void renorm_typed(const double *typed, int n, int neq, int dim, int kind, const int *levels,
                  const int32_t *groups, double *out);
// componentWiseNorm_tree (componentWiseNorm_tree.f90:63-200) from per-block
// values v[k*neq + c] (maxima for Linfty, sums otherwise), blocks in order
void accumulate_norm(const double *v, int n, int neq, const int32_t *groups, const int *levels, int dim,
                     const int bs[3], const double domain[3], int kind, double *norm);
int launch_mask_stats(wh_pool *P, const std::vector<int32_t> &ids, double *d_out);
int results_reserve(wh_pool *P, size_t n, double **d_out);
int results_fetch(wh_pool *P, size_t n, std::vector<double> &out);
// threshold_block's grouping of detail components (threshold_block.f90:86-91),
// in place on n blocks of neq values
void detail_groups(const wh_pool *P, const int32_t *groups, double *detail, int n);
int run_region_max(wh_pool *P, int array, const std::vector<wh_box> &boxes, bool zero_sc, bool sqrt_groups,
                   const int32_t *groups, std::vector<double> &out);
// penalisation mask (wh_mask.hip)
int run_mask(wh_pool *P, const wh_mask_params &mp, const std::vector<int32_t> &ids,
             const std::vector<double> &x0dx);
// the masks of several forests' blocks in one launch: mps[gidx[k]] is
// block k's geometry (gidx empty: one geometry for all); the parameters'
// checks of run_mask for every geometry
int check_mask_params(wh_pool *P, const wh_mask_params &mp);
int run_masks(wh_pool *P, const std::vector<wh_mask_params> &mps, const std::vector<int32_t> &gidx,
              const std::vector<int32_t> &ids, const std::vector<double> &x0dx);
int run_mask_stats(wh_pool *P, const std::vector<int32_t> &ids, std::vector<double> &stats);

AcmDev make_acm_dev(const wh_acm_params *p);
Geo make_geo(const wh_pool *P);
// multi-pool launches (independent forests of a batch, wh_batch.hip)
// d_tot: [0] entries of the launch, [s] chunks of sync stage s (1..3), [3 + s]
// patches of sync stage s; dtg /
// gacm: dt and constants per block group of the launch; pipe: the 2D
// software-pipelined kernel (tile widths and heights <= 32)
// a run of ghost patches (set_patch_table takes several, one after the other)
struct PatchRun {
    const wh_patch *p;
    int n;
};
// the patch table of sync case 0/1 from runs of patches, without the checks
// of wh_set_patches (for lists the library generated itself)
int set_patch_table(wh_pool *P, int sync_case, const PatchRun *runs, int nruns);
// d_ents: the launch's entry records (launch_entries, capacity >= d_tot[0])
int launch_entries(hipStream_t stream, const MPoolDev *d_pools, int npools, const int64_t *d_tot, MEnt *d_ents,
                   int grid);
int launch_stage_multi(hipStream_t stream, int dim, int scheme, int skew, int xcd, const MPoolDev *d_pools,
                       int npools, const MEnt *d_ents, const double *d_dtg, const AcmDev *d_gacm,
                       const int64_t *d_tot, const MStage &ms, size_t lds, int grid, int nchunk, int pipe = 0);
int launch_sync_multi(hipStream_t stream, const MPoolDev *d_pools, int npools, int sel, int stage,
                      const int64_t *d_tot, int grid, int dim);
int reach_of(int scheme);
bool zm_valid(int bs0, int bs1, int scheme, int ys);
bool zm_zvalid(int bs2, int scheme, int zs);

}  // namespace wh
