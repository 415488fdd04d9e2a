// wh_batch.hip -- RungeKuttaGeneric over the block groups of several pools
// in one launch per RK stage (independent forests of an ensemble or of a
// serving loop, any block size), optionally replayed from a hipGraph.
//
// The reference runs one forest per process (LIB/MAIN/main.f90) and loops
// over its blocks (runge_kutta_generic.f90:1-156).  Here every pool keeps its
// own active list, origin/spacing, block groups and RHS ghost-patch table on
// the device (set as for wh_rk_step_groups); the batch passes the kernels
// one small descriptor per pool (where those tables are, and where the pool's
// entries and ghost slots start in the launch) plus dt and constants per
// block group, and runs each RK stage as one ghost fill per sync stage plus
// one fused stage kernel over the blocks of all pools.  The kernels are
// persistent and read their counts and the descriptors from device memory,
// so the captured launch sequence is replayed unchanged while the grids
// change from step to step (re-captured only when the group arrays must grow
// or the tableau / scheme / block sizes change).  No memset is captured.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "wh_internal.h"
#include "wh_prof.h"

using wh::AcmDev;
using wh::MPoolDev;
using wh::MStage;

struct wh_batch {
    std::vector<wh_pool *> pools;
    int dim = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr, ev_up = nullptr;
    bool up_recorded = false;
    std::vector<hipEvent_t> ev_pool;
    int n_cu = 256;
    // device tables: one descriptor per pool, dt and constants per block
    // group (capacity cap_grp), counts ([0] entries, [s] chunks and [3 + s]
    // patches of sync stage s)
    MPoolDev *d_pools = nullptr;
    wh::MEnt *d_ents = nullptr;  // entry records, one per slot of the pools
    double *d_dtg = nullptr;
    AcmDev *d_gacm = nullptr;
    int64_t *d_tot = nullptr;
    size_t cap_grp = 0;
    // pinned host copies of the tables
    char *h_buf = nullptr;
    size_t h_cap = 0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t buf_version = 1, graph_version = 0;
    std::vector<double> graph_key;
    // the last step's launch geometry
    size_t lds = 0;
    int nchunk = 1, grid_stage = 0, grid_sync = 0;
    int n_ent_last = 0;
    int sout_sel = 0;  // the last stage's output array (0: the state; 3 while timing)
    // 2D stage kernel (WH_TUNE_K2D_STAGE): 0 one block per workgroup at a
    // time with the first point's operands loaded before the tile barrier;
    // 1 software-pipelined (used where every tile is at most 32 x 32); 2 as
    // 0 without the early loads; 3 / 4 as 2 with registers for 6 / 8 waves
    // per SIMD
    int k2d = 0;
    bool pipe_ok = false;  // the last step's tiles fit the pipelined kernel
    int full_rows = 0;     // the last step's stages write their outputs in whole rows
};

namespace {

template <class T>
int grow(T *&p, size_t &cap, size_t n, uint64_t &version)
{
    if (n <= cap) return WH_OK;
    if (p) WH_HIP(hipFree(p));
    p = nullptr;
    const size_t c = std::max<size_t>(n, cap + cap / 2 + 64);
    WH_HIP(hipMalloc(&p, sizeof(T) * c));
    cap = c;
    ++version;
    return WH_OK;
}

void free_graph(wh_batch *B)
{
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->gexec = nullptr;
    B->graph = nullptr;
}

bool pipe_on(const wh_batch *B) { return B->dim == 2 && B->k2d == 1 && B->pipe_ok; }

// the launcher's kernel choice (0: the default kernel)
int k2d_sel(const wh_batch *B) { return B->dim != 2 ? 0 : B->k2d == 1 ? (pipe_on(B) ? 1 : 0) : B->k2d; }

// persistent stage-kernel grids, multiples of 8 (one per XCD); the pipelined
// kernel holds more registers, so its grid stays resident at 4 per CU
int stage_grid(const wh_batch *B) { return B->dim == 2 ? (pipe_on(B) ? 4 : 8) * B->n_cu : 16 * B->n_cu; }

// the RK stage sequence of the fused path (runge_kutta_generic.f90:63-154 with
// the stage-j RHS and its accumulation fused, as wh_rk_step path 2)
int launch_sequence(wh_batch *B, const wh_acm_params &p, const double *bt, int nrows)
{
    const int s = nrows - 1;
    const int np = (int)B->pools.size();
    // the step's entry records, read by every stage
    if (int rc = wh::launch_entries(B->stream, B->d_pools, np, B->d_tot, B->d_ents, 2 * B->n_cu)) return rc;
    auto sync = [&](int sel) {
        for (int st = 1; st <= 3; ++st)
            if (int rc = wh::launch_sync_multi(B->stream, B->d_pools, np, sel, st, B->d_tot, B->grid_sync, B->dim))
                return rc;
        return (int)WH_OK;
    };
    auto stage = [&](int in_sel, int out_sel, int mode, double a, double b) {
        MStage ms{};
        ms.in_sel = in_sel;
        ms.out_sel = out_sel;
        ms.mode = mode;
        ms.a = a;
        ms.b = b;
        ms.use_a = std::abs(a) >= 1.0e-8;
        ms.use_b = std::abs(b) >= 1.0e-8;
        ms.sout_sel = mode == 2 ? B->sout_sel : 0;
        ms.full_rows = B->full_rows;
        return wh::launch_stage_multi(B->stream, B->dim, p.scheme, p.skew_symmetry ? 1 : 0, 1, B->d_pools, np,
                                      B->d_ents, B->d_dtg, B->d_gacm, B->d_tot, ms, B->lds, B->grid_stage,
                                      B->nchunk, k2d_sel(B));
    };
    // arrays: 0 state, 1 accumulator (work 0), 2 / 3 ping-pong stage inputs
    if (int rc = sync(0)) return rc;
    if (int rc = stage(0, 2, 0, bt[1 * nrows + 1], bt[s * nrows + 1])) return rc;
    int cur = 2;
    for (int j = 2; j <= s; ++j) {
        if (int rc = sync(cur)) return rc;
        if (j < s) {
            const int nxt = cur == 2 ? 3 : 2;
            if (int rc = stage(cur, nxt, 1, bt[j * nrows + j], bt[s * nrows + j])) return rc;
            cur = nxt;
        } else {
            if (int rc = stage(cur, 0, 2, 0.0, bt[s * nrows + j])) return rc;
        }
    }
    return WH_OK;
}

bool same_kind(const wh_acm_params &a, const wh_acm_params &b)
{
    return a.dim == b.dim && a.scheme == b.scheme && !a.skew_symmetry == !b.skew_symmetry;
}

}  // namespace

extern "C" {

int wh_set_group_params(wh_pool *P, const wh_acm_params *params, int ngroups)
{
    WH_CHECK(P && (ngroups == 0 || params), WH_E_ARG, "wh_set_group_params: null argument");
    WH_CHECK(ngroups == 0 || ngroups == P->ngrp, WH_E_ARG,
             "wh_set_group_params: one parameter set per block group (wh_set_block_groups first)");
    for (int i = 0; i < ngroups; ++i) {
        const wh_acm_params &q = params[i];
        WH_CHECK(same_kind(q, params[0]), WH_E_ARG,
                 "wh_set_group_params: groups of one pool share dim, scheme and skew_symmetry");
        WH_CHECK(q.dim == P->d.dim && q.C_eta > 0.0 && q.C_sponge > 0.0, WH_E_ARG,
                 "wh_set_group_params: dim mismatch or C_eta / C_sponge <= 0");
        WH_CHECK(!((q.penalization || q.use_sponge) && (P->d.n_mask < 6 || !P->mask)), WH_E_ARG,
                 "wh_set_group_params: penalization/sponge need a 6-component mask");
    }
    P->h_gparams.assign(params, params + ngroups);
    P->ngacm = 0;
    if (ngroups == 0) return WH_OK;
    if (ngroups > P->gacm_cap) {
        WH_HIP(hipStreamSynchronize(P->stream));
        if (P->d_gacm) WH_HIP(hipFree(P->d_gacm));
        P->d_gacm = nullptr;
        WH_HIP(hipMalloc(&P->d_gacm, sizeof(AcmDev) * ngroups));
        P->gacm_cap = ngroups;
    }
    std::vector<AcmDev> a(ngroups);
    for (int i = 0; i < ngroups; ++i) a[i] = wh::make_acm_dev(&params[i]);
    if (int rc = wh::stage_begin(P, sizeof(AcmDev) * ngroups)) return rc;
    void *h = wh::stage_host(P, a.data(), sizeof(AcmDev) * ngroups);
    WH_HIP(hipMemcpyAsync(P->d_gacm, h, sizeof(AcmDev) * ngroups, hipMemcpyHostToDevice, P->stream));
    P->ngacm = ngroups;
    return WH_OK;
}

int wh_batch_create(wh_batch **out, wh_pool *const *pools, int npools)
{
    WH_CHECK(out && pools && npools >= 1 && npools <= (1 << 16) && pools[0], WH_E_ARG,
             "wh_batch_create: bad argument");
    const int dim = pools[0]->d.dim;
    for (int i = 0; i < npools; ++i) {
        const wh_pool *P = pools[i];
        WH_CHECK(P, WH_E_ARG, "wh_batch_create: null pool");
        WH_CHECK(P->d.dim == dim, WH_E_ARG, "wh_batch_create: the pools of a batch share dim");
        WH_CHECK(P->d.n_work_slots >= 3, WH_E_ARG, "wh_batch_create: the fused stages need 3 work slots");
        WH_CHECK(P->device == pools[0]->device, WH_E_ARG, "wh_batch_create: the pools share one device");
        for (int j = 0; j < i; ++j) WH_CHECK(pools[j] != P, WH_E_ARG, "wh_batch_create: a pool listed twice");
    }
    wh_batch *B = new wh_batch();
    B->pools.assign(pools, pools + npools);
    B->dim = dim;
    B->n_cu = pools[0]->n_cu;
    auto fail = [&](hipError_t e) {
        wh::set_error(std::string("wh_batch_create: ") + hipGetErrorString(e));
        wh_batch_destroy(B);
        return WH_E_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipEventCreateWithFlags(&B->ev_done, hipEventDisableTiming)) != hipSuccess) return fail(e);
    if ((e = hipEventCreateWithFlags(&B->ev_up, hipEventDisableTiming)) != hipSuccess) return fail(e);
    B->ev_pool.resize(npools, nullptr);
    for (auto &ev : B->ev_pool)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&B->d_tot, sizeof(int64_t) * 7)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&B->d_pools, sizeof(MPoolDev) * npools)) != hipSuccess) return fail(e);
    size_t slots = 0;
    for (int i = 0; i < npools; ++i) slots += (size_t)pools[i]->d.nb_max;
    if ((e = hipMalloc(&B->d_ents, sizeof(wh::MEnt) * slots)) != hipSuccess) return fail(e);
    *out = B;
    return WH_OK;
}

int wh_batch_destroy(wh_batch *B)
{
    if (!B) return WH_OK;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    free_graph(B);
    for (void *p : {(void *)B->d_pools, (void *)B->d_ents, (void *)B->d_dtg, (void *)B->d_gacm, (void *)B->d_tot})
        if (p) (void)hipFree(p);
    if (B->h_buf) (void)hipHostFree(B->h_buf);
    for (hipEvent_t ev : B->ev_pool)
        if (ev) (void)hipEventDestroy(ev);
    if (B->ev_done) (void)hipEventDestroy(B->ev_done);
    if (B->ev_up) (void)hipEventDestroy(B->ev_up);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return WH_OK;
}

int wh_batch_rk_step(wh_batch *B, const wh_acm_params *p, const double *bt, int nrows, const double *dt_groups,
                     int use_graph)
{
    WH_PROF("api.batch_rk_step");
    WH_CHECK(B && p && bt && dt_groups, WH_E_ARG, "wh_batch_rk_step: null argument");
    WH_CHECK(nrows >= 3 && nrows <= 9, WH_E_ARG, "wh_batch_rk_step: butcher tableau size");
    for (int j = 1; j < nrows - 1; ++j)
        for (int l = 1; l < nrows; ++l)
            WH_CHECK(l == j || std::fabs(bt[j * nrows + l]) < 1.0e-8, WH_E_ARG,
                     "wh_batch_rk_step: the fused stages need a tableau with sub-diagonal stage rows");
    WH_CHECK(p->dim == B->dim, WH_E_ARG, "wh_batch_rk_step: params dim does not match the pools");
    const int npools = (int)B->pools.size();
    // per pool: where its device tables are and where its entries, groups
    // and ghost slots start in the launch; per group: dt and constants
    std::vector<AcmDev> gacm;
    std::vector<double> dtg;
    std::vector<MPoolDev> mp(npools);
    size_t lds = 0;
    int nchunk = 1;
    const int R = p->scheme == WH_FD2 ? 1 : p->scheme == WH_FD4 ? 2 : 3;
    int goff = 0, nent = 0;
    double step_bytes = 0.0;  // one block of every entry
    int64_t tot[7] = {0, 0, 0, 0, 0, 0, 0};
    bool pipe_ok = true;
    for (int i = 0; i < npools; ++i) {
        wh_pool *P = B->pools[i];
        WH_CHECK(P->nact == (int)P->h_active.size() && (P->nact == 0 || P->ngrp > 0), WH_E_STATE,
                 "wh_batch_rk_step: every pool with blocks needs its block groups (wh_set_block_groups)");
        WH_CHECK(P->d.g >= R, WH_E_ARG, "wh_batch_rk_step: ghost width smaller than the stencil reach");
        WH_CHECK(P->d.dim == 3 || P->d.neq == 3, WH_E_ARG, "wh_batch_rk_step: 2D ACM pools hold 3 equations");
        const bool own = P->ngacm > 0 && P->ngacm == P->ngrp;
        bool masked = false;
        for (int gi = 0; gi < P->ngrp; ++gi) {
            const wh_acm_params &q = own ? P->h_gparams[gi] : *p;
            WH_CHECK(same_kind(q, *p), WH_E_ARG, "wh_batch_rk_step: a group differs in dim / scheme / skew");
            WH_CHECK(q.C_eta > 0.0 && q.C_sponge > 0.0, WH_E_ARG, "wh_batch_rk_step: C_eta and C_sponge must be > 0");
            WH_CHECK(dt_groups[goff + gi] >= 0.0, WH_E_ARG, "wh_batch_rk_step: negative dt");
            if (q.penalization || q.use_sponge) {
                WH_CHECK(P->d.n_mask >= 6 && P->mask, WH_E_ARG,
                         "wh_batch_rk_step: penalization/sponge need a 6-component mask");
                masked = true;
            }
            gacm.push_back(wh::make_acm_dev(&q));
            dtg.push_back(dt_groups[goff + gi]);
        }
        const wh::PatchSet &ps = P->ps[0];
        MPoolDev &m = mp[i];
        m.arr[0] = P->block;
        for (int s = 0; s < 3; ++s) m.arr[1 + s] = wh::array_ptr(P, WH_WORK + s, nullptr, nullptr);
        m.mask = masked ? P->mask : nullptr;
        m.G = wh::make_geo(P);
        m.neq = P->d.neq;
        m.dim = P->d.dim;
        m.act = P->d_active;
        m.x0dx = P->d_x0dx;
        m.grp = P->d_grp;
        m.pt = ps.d_patches;
        m.ch = ps.d_chunks;
        m.ent0 = nent;
        m.goff = goff;
        for (int st = 1; st <= 3; ++st) {
            const int a = ps.n ? ps.stage_begin[st] : 0, b = ps.n ? ps.stage_begin[st + 1] : 0;
            const int ca = ps.n ? ps.chunk_begin[st] : 0, cb = ps.n ? ps.chunk_begin[st + 1] : 0;
            m.p0[st - 1] = a;
            m.p1[st - 1] = b;
            m.c0[st - 1] = ca;
            m.c1[st - 1] = cb;
            m.s0[st - 1] = tot[st];
            m.q0[st - 1] = tot[3 + st];
            if (P->nact > 0 && b > a) {
                tot[st] += cb - ca;
                tot[3 + st] += b - a;
            } else {  // a pool without blocks fills no ghosts
                m.p1[st - 1] = a;
                m.c1[st - 1] = ca;
            }
        }
        if (P->nact > 0) {
            if (std::max(P->d.bs[0], P->d.bs[1]) + 2 * R > 32) pipe_ok = false;
            if (B->dim == 2)
                lds = std::max(lds, sizeof(double) * 3 * (P->d.bs[0] + 2 * R) * (P->d.bs[1] + 2 * R));
            else
                nchunk = std::max(nchunk, (P->d.bs[0] * P->d.bs[1] * P->d.bs[2] + 255) / 256);
        }
        nent += P->nact;
        goff += P->ngrp;
        step_bytes += (double)P->nact * sizeof(double) * (double)m.G.blk;
    }
    tot[0] = nent;
    if (gacm.empty()) {
        gacm.push_back(wh::make_acm_dev(p));
        dtg.push_back(0.0);
    }
    // group arrays (grown, never shrunk; growing re-captures the graph)
    uint64_t &ver = B->buf_version;
    if (gacm.size() > B->cap_grp) {
        WH_HIP(hipStreamSynchronize(B->stream));  // queued work may still read the old buffers
        size_t c = B->cap_grp;
        int rc = grow(B->d_dtg, c, gacm.size(), ver);
        c = B->cap_grp;
        if (!rc) rc = grow(B->d_gacm, c, gacm.size(), ver);
        if (rc) return rc;
        B->cap_grp = c;
    }
    // pinned copies: the previous step's uploads must have left the buffer
    const size_t bytes[4] = {sizeof(MPoolDev) * npools, sizeof(double) * dtg.size(), sizeof(AcmDev) * gacm.size(),
                             sizeof(tot)};
    size_t total = 0;
    for (size_t b : bytes) total += (b + 255) & ~(size_t)255;
    if (B->up_recorded) WH_HIP(hipEventSynchronize(B->ev_up));
    if (total > B->h_cap) {
        if (B->h_buf) WH_HIP(hipHostFree(B->h_buf));
        B->h_buf = nullptr;
        const size_t c = std::max(total, B->h_cap * 2);
        WH_HIP(hipHostMalloc((void **)&B->h_buf, c, hipHostMallocDefault));
        B->h_cap = c;
    }
    // order after each pool's queued work (mesh operations, masks, tables)
    for (int i = 0; i < npools; ++i) {
        WH_HIP(hipEventRecord(B->ev_pool[i], B->pools[i]->stream));
        WH_HIP(hipStreamWaitEvent(B->stream, B->ev_pool[i], 0));
    }
    size_t off = 0;
    auto put = [&](void *dst, const void *src, size_t n) -> int {
        if (!n) return WH_OK;
        std::memcpy(B->h_buf + off, src, n);
        WH_HIP(hipMemcpyAsync(dst, B->h_buf + off, n, hipMemcpyHostToDevice, B->stream));
        off += (n + 255) & ~(size_t)255;
        return WH_OK;
    };
    int rc = WH_OK;
    if (!rc) rc = put(B->d_pools, mp.data(), bytes[0]);
    if (!rc) rc = put(B->d_dtg, dtg.data(), bytes[1]);
    if (!rc) rc = put(B->d_gacm, gacm.data(), bytes[2]);
    if (!rc) rc = put(B->d_tot, tot, bytes[3]);
    if (rc) return rc;
    WH_HIP(hipEventRecord(B->ev_up, B->stream));
    B->up_recorded = true;
    // launch geometry
    B->lds = lds;
    B->nchunk = nchunk;
    B->pipe_ok = pipe_ok;
    B->full_rows = wh::stage_full_rows(1.0, step_bytes) ? 1 : 0;
    B->grid_stage = stage_grid(B);
    B->grid_sync = 4 * B->n_cu;
    B->n_ent_last = nent;
    if (use_graph) {
        std::vector<double> key(bt, bt + nrows * nrows);
        key.push_back(p->scheme);
        key.push_back(p->skew_symmetry ? 1 : 0);
        key.push_back((double)lds);
        key.push_back(nchunk);
        key.push_back(npools);
        key.push_back(k2d_sel(B));
        key.push_back(B->full_rows);
        if (!B->gexec || B->graph_version != ver || key != B->graph_key) {
            free_graph(B);
            WH_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            const int lrc = launch_sequence(B, *p, bt, nrows);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(B->stream, &g);
            if (lrc) {
                if (g) (void)hipGraphDestroy(g);
                return lrc;
            }
            WH_HIP(ec);
            B->graph = g;
            WH_HIP(hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0));
            B->graph_version = ver;
            B->graph_key = key;
        }
        WH_HIP(hipGraphLaunch(B->gexec, B->stream));
    } else {
        if (int lrc = launch_sequence(B, *p, bt, nrows)) return lrc;
    }
    // the pools' later work (adapt_tree, downloads) follows the step, and a
    // pool replaces its tables only after the step has read them
    WH_HIP(hipEventRecord(B->ev_done, B->stream));
    for (wh_pool *P : B->pools) {
        WH_HIP(hipStreamWaitEvent(P->stream, B->ev_done, 0));
        WH_HIP(hipEventRecord(P->ev_tables, P->stream));
        P->tables_recorded = true;
    }
    return WH_OK;
}

int wh_batch_time_stage_kernel(wh_batch *B, const wh_acm_params *p, const double *bt, int nrows, int reps,
                               double *ms, int *blocks)
{
    WH_CHECK(B && p && bt && ms && reps >= 1 && reps <= 1000, WH_E_ARG, "wh_batch_time_stage_kernel: bad argument");
    WH_CHECK(B->d_tot && B->grid_stage > 0, WH_E_STATE, "wh_batch_time_stage_kernel: no step run yet");
    const int s = nrows - 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    WH_HIP(hipEventCreate(&e0));
    WH_HIP(hipEventCreate(&e1));
    auto stage = [&](int in_sel, int out_sel, int mode, double a, double b) {
        MStage m{};
        m.in_sel = in_sel;
        m.out_sel = out_sel;
        m.sout_sel = 3;  // never the state
        m.mode = mode;
        m.a = a;
        m.b = b;
        m.use_a = std::abs(a) >= 1.0e-8;
        m.use_b = std::abs(b) >= 1.0e-8;
        m.full_rows = B->full_rows;
        return wh::launch_stage_multi(B->stream, B->dim, p->scheme, p->skew_symmetry ? 1 : 0, 1, B->d_pools,
                                      (int)B->pools.size(), B->d_ents, B->d_dtg, B->d_gacm, B->d_tot, m, B->lds,
                                      stage_grid(B), B->nchunk, k2d_sel(B));
    };
    // the last step's entry records (rebuilt from its descriptors)
    int rc = wh::launch_entries(B->stream, B->d_pools, (int)B->pools.size(), B->d_tot, B->d_ents, 2 * B->n_cu);
    WH_HIP(hipEventRecord(e0, B->stream));
    for (int r = 0; r < reps && !rc; ++r) {
        // one RK4 stage sequence that reads the state and writes work slots only
        rc = stage(0, 2, 0, bt[1 * nrows + 1], bt[s * nrows + 1]);
        int cur = 2;
        for (int j = 2; j <= s && !rc; ++j) {
            const int nxt = cur == 2 ? 3 : 2;
            rc = j < s ? stage(cur, nxt, 1, bt[j * nrows + j], bt[s * nrows + j])
                       : stage(cur, 0, 2, 0.0, bt[s * nrows + j]);
            cur = nxt;
        }
    }
    WH_HIP(hipEventRecord(e1, B->stream));
    WH_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    WH_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms = (double)t / (reps * s);
    if (blocks) *blocks = B->n_ent_last;
    return WH_OK;
}

int wh_batch_set_tuning(wh_batch *B, int key, int value)
{
    WH_CHECK(B, WH_E_ARG, "wh_batch_set_tuning: null argument");
    WH_CHECK(key == WH_TUNE_K2D_STAGE && value >= 0 && value <= 4, WH_E_ARG,
             "wh_batch_set_tuning: WH_TUNE_K2D_STAGE takes 0 to 4");
    B->k2d = value;
    return WH_OK;
}

int wh_batch_get_tuning(wh_batch *B, int key, int *value)
{
    WH_CHECK(B && value && key == WH_TUNE_K2D_STAGE, WH_E_ARG, "wh_batch_get_tuning: bad argument");
    *value = B->k2d;
    return WH_OK;
}

int wh_batch_synchronize(wh_batch *B)
{
    WH_CHECK(B, WH_E_ARG, "wh_batch_synchronize: null argument");
    WH_HIP(hipStreamSynchronize(B->stream));
    return WH_OK;
}

}  // extern "C"
