// wh_api.hip -- extern "C" entry points of libwabbit_hip.so (see include/wabbit_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "wh_internal.h"
#include "wh_prof.h"

namespace wh {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

double *array_ptr(wh_pool *P, int array, size_t *blk_out, int *ncomp)
{
    if (array == WH_BLOCK) {
        if (blk_out) *blk_out = P->blk;
        if (ncomp) *ncomp = P->d.neq;
        return P->block;
    }
    if (array == WH_TMP) {
        if (blk_out) *blk_out = P->tblk;
        if (ncomp) *ncomp = P->d.n_tmp_eqn;
        return P->tmp;
    }
    if (array == WH_MASK) {
        if (blk_out) *blk_out = P->mblk;
        if (ncomp) *ncomp = P->d.n_mask;
        return P->mask;
    }
    if (array >= WH_WORK && array < WH_WORK + P->d.n_work_slots) {
        if (blk_out) *blk_out = P->blk;
        if (ncomp) *ncomp = P->d.neq;
        return P->work + (size_t)(array - WH_WORK) * P->d.nb_max * P->blk;
    }
    return nullptr;
}

// Time one kernel-kind launch sequence with events on the pool stream.
template <class F>
static int timed(wh_pool *P, int which, F &&f)
{
    if (!P->timing) return f();
    WH_HIP(hipEventRecord(P->ev_a, P->stream));
    int rc = f();
    if (rc) return rc;
    WH_HIP(hipEventRecord(P->ev_b, P->stream));
    WH_HIP(hipEventSynchronize(P->ev_b));
    float ms = 0.f;
    WH_HIP(hipEventElapsedTime(&ms, P->ev_a, P->ev_b));
    P->t_ms[which] += ms;
    P->t_cnt[which] += 1;
    return WH_OK;
}

static int read_scalar(wh_pool *P, int which, double *out)
{
    WH_HIP(hipMemcpyAsync(P->h_scal + which, P->scal + which, sizeof(double),
                          hipMemcpyDeviceToHost, P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    *out = P->h_scal[which];
    return WH_OK;
}

static TimeDev make_time_dev(const wh_time_params *tp)
{
    TimeDev t{};
    t.time_max = tp->time_max;
    t.dt_fixed = tp->dt_fixed;
    t.dt_max = tp->dt_max;
    t.write_time = tp->write_time;
    t.write_time_first = tp->write_time_first;
    t.tsave_stats = tp->tsave_stats;
    t.write_backup_time = tp->write_backup_time;
    t.write_fixed_time = tp->write_fixed_time;
    t.write_backup_n = tp->write_backup_n;
    t.use_stats = std::fabs(tp->tsave_stats - 9999999.9) > 1e-3;
    return t;
}

// The fused path needs every stage row j>=2 of the tableau to carry only its
// sub-diagonal coefficient a(j, j-1) (classic RK4, midpoint rule, ...).
static bool fused_ok(const double *bt, int nrows)
{
    if (nrows < 3) return false;
    for (int j = 1; j < nrows - 1; ++j)
        for (int l = 1; l < nrows; ++l) {
            if (l == j) continue;  // a(j+1, j) in 1-based terms
            if (std::fabs(bt[j * nrows + l]) >= 1.0e-8) return false;
        }
    return true;
}

// The halo-gather table of the fused 2D stages (WH_TUNE_GATHER), or null:
// the RHS patch table holds same-level copies only, at least as deep as the
// stencil reach, and every active block has the senders of all 8
// directions (checked once per topology version).
static const int32_t *gather_table(wh_pool *P, const AcmDev &a)
{
    const PatchSet &ps = P->ps[0];
    if (!P->gather || P->d.dim != 2 || !ps.gather_ok || !ps.d_gnb || ps.gather_depth < reach_of(a.scheme))
        return nullptr;
    if (P->gather_topo != P->topo_version) {
        bool all = P->nact > 0;
        for (int k = 0; k < P->nact && all; ++k) {
            const int32_t *row = &ps.h_gnb[(size_t)P->h_active[k] * 9];
            for (int d = 0; d < 9 && all; ++d)
                if (d != 4 && row[d] < 0) all = false;
        }
        P->gather_all = all;
        P->gather_topo = P->topo_version;
    }
    return P->gather_all ? ps.d_gnb : nullptr;
}

// one step of the fused path; dt is taken from the device scalar DS_DT, or
// per block from dtb (block groups).  With a halo-gather table no ghost
// fill is launched: every stage reads its blocks' halos from the
// neighbours' interiors of its input array.
static int fused_step(wh_pool *P, const AcmDev &a, const double *bt, int nrows,
                      const double *dtb = nullptr)
{
    const int s = nrows - 1;  // number of stages
    double *S = P->block;
    double *acc = array_ptr(P, WH_WORK + 0, nullptr, nullptr);
    double *T[2] = {array_ptr(P, WH_WORK + 1, nullptr, nullptr),
                    array_ptr(P, WH_WORK + 2, nullptr, nullptr)};
    const int32_t *gnb = gather_table(P, a);
    // the ghost fill (timed as a sync launch); nothing at all with gather
    auto sync = [&](double *arr) {
        return gnb ? (int)WH_OK : timed(P, 1, [&] { return wh::launch_sync(P, 0, arr, -1); });
    };
    auto launch_fused_stage = [&](wh_pool *Q, const AcmDev &aa, const double *tin, double *tout, const double *u0,
                                  double *ac, double *sout, int mode, double acoef, double bcoef, const double *db) {
        return wh::launch_fused_stage(Q, aa, tin, tout, u0, ac, sout, mode, acoef, bcoef, db, gnb);
    };
    int rc;
    if ((rc = sync(S))) return rc;
    // stage 1: input S, writes acc and T[0]
    if ((rc = timed(P, 0, [&] {
             return launch_fused_stage(P, a, S, T[0], nullptr, acc, nullptr, 0, bt[1 * nrows + 1],
                                       bt[s * nrows + 1], dtb);
         })))
        return rc;
    int cur = 0;
    for (int j = 2; j <= s; ++j) {
        double *tin = T[cur];
        if ((rc = sync(tin))) return rc;
        if (j < s) {
            double *tout = T[cur ^ 1];
            if ((rc = timed(P, 0, [&] {
                     return launch_fused_stage(P, a, tin, tout, S, acc, nullptr, 1,
                                               bt[j * nrows + j], bt[s * nrows + j], dtb);
                 })))
                return rc;
            cur ^= 1;
        } else {
            if ((rc = timed(P, 0, [&] {
                     return launch_fused_stage(P, a, tin, nullptr, nullptr, acc, S, 2, 0.0,
                                               bt[s * nrows + j], dtb);
                 })))
                return rc;
        }
    }
    return WH_OK;
}

// reference-structured step; per_block = 1 launches every operation once per
// block, then dt_pos (host, per active-list position) may give each block its
// own dt (block groups)
static int generic_step(wh_pool *P, const AcmDev &a, const double *bt, int nrows, double dt,
                        int per_block, const double *dt_pos = nullptr)
{
    const int ncol = nrows;
    double *S = P->block;
    auto W = [&](int slot) { return array_ptr(P, WH_WORK + slot, nullptr, nullptr); };
    auto sync = [&](double *arr) -> int {
        if (!per_block) return timed(P, 1, [&] { return launch_sync(P, 0, arr, -1); });
        for (int r = 0; r < P->ps[0].nrecv; ++r) {
            int rc = timed(P, 1, [&] { return launch_sync(P, 0, arr, r); });
            if (rc) return rc;
        }
        return WH_OK;
    };
    auto rhs = [&](double *in, double *out) -> int {
        if (!per_block)
            return timed(P, 0, [&] {
                return launch_rhs(P, a, in, out, P->d_active, P->d_x0dx, P->nact);
            });
        for (int k = 0; k < P->nact; ++k) {
            int rc = timed(P, 0, [&] {
                return launch_rhs(P, a, in, out, P->d_active + k, P->d_x0dx + 6 * k, 1);
            });
            if (rc) return rc;
        }
        return WH_OK;
    };
    // raw holds the tableau entries a; the kernels get dt * a, with dt_pos
    // block k's own dt_pos[k] * a
    auto combine = [&](double *out, const double *base, int nt, const double *const *terms,
                       const double *raw) -> int {
        std::vector<double> ck(nt > 0 ? nt : 1);
        for (int t = 0; t < nt; ++t) ck[t] = dt * raw[t];
        if (!per_block) return launch_combine(P, out, base, nt, terms, ck.data(), P->d_active, P->nact);
        for (int k = 0; k < P->nact; ++k) {
            if (dt_pos)
                for (int t = 0; t < nt; ++t) ck[t] = dt_pos[k] * raw[t];
            int rc = launch_combine(P, out, base, nt, terms, ck.data(), P->d_active + k, 1);
            if (rc) return rc;
        }
        return WH_OK;
    };
    int rc;
    // runge_kutta_generic.f90:52 sync, :63-67 u0 copy, :72 first RHS
    if ((rc = sync(S))) return rc;
    if ((rc = combine(W(0), S, 0, nullptr, nullptr))) return rc;
    if ((rc = rhs(S, W(1)))) return rc;
    for (int j = 2; j <= nrows - 1; ++j) {  // :80-127
        std::vector<const double *> terms;
        std::vector<double> coefs;
        for (int l = 2; l <= j; ++l) {
            const double c = bt[(j - 1) * ncol + (l - 1)];
            if (std::fabs(c) < 1.0e-8) continue;
            terms.push_back(W(l - 1));
            coefs.push_back(c);
        }
        if ((rc = combine(S, W(0), (int)terms.size(), terms.data(), coefs.data()))) return rc;
        if ((rc = sync(S))) return rc;
        if ((rc = rhs(S, W(j)))) return rc;
    }
    // :136-154 final stage
    std::vector<const double *> terms;
    std::vector<double> coefs;
    for (int j = 2; j <= ncol; ++j) {
        const double c = bt[(nrows - 1) * ncol + (j - 1)];
        if (std::fabs(c) < 1.0e-8) continue;
        terms.push_back(W(j - 1));
        coefs.push_back(c);
    }
    return combine(S, W(0), (int)terms.size(), terms.data(), coefs.data());
}

// Prediction reads, for the refined-grid indices [fmin, fmax] along one
// direction: even f copies coarse point f/2, odd f reads coarse points
// f/2-N+1 .. f/2+N.  True when all of them lie in [0, ext).
static bool pred_range_ok(int fmin, int fmax, int N, int ext)
{
    if (fmin < 0 || fmax < fmin) return false;
    if ((fmax >> 1) >= ext) return false;
    const int omin = (fmin & 1) ? fmin : fmin + 1;
    const int omax = (fmax & 1) ? fmax : fmax - 1;
    if (omin > omax) return true;  // only copies
    return (omin >> 1) - N + 1 >= 0 && (omax >> 1) + N < ext;
}

// Checks that the kernels' reads stay inside the pool for these parameters.
static int check_acm(const wh_pool *P, const wh_acm_params *p)
{
    WH_CHECK(p->dim == P->d.dim, WH_E_ARG, "params dim does not match the pool");
    WH_CHECK(P->d.neq >= p->dim + 1, WH_E_ARG, "pool has fewer equations than the ACM needs");
    WH_CHECK(P->d.dim == 3 || P->d.neq == 3, WH_E_ARG, "2D ACM pools hold exactly 3 equations");
    const int reach = p->scheme == WH_FD2 ? 1 : p->scheme == WH_FD4 ? 2
                      : (p->scheme == WH_FD4_TW || p->scheme == WH_FD6) ? 3 : -1;
    WH_CHECK(reach > 0, WH_E_ARG, "unknown discretization scheme");
    WH_CHECK(P->d.g >= reach, WH_E_ARG, "ghost width smaller than the stencil reach");
    WH_CHECK(!((p->penalization || p->use_sponge) && (P->d.n_mask < 6 || !P->mask)), WH_E_ARG,
             "penalization/sponge need a 6-component mask");
    WH_CHECK(p->C_eta > 0.0 && p->C_sponge > 0.0, WH_E_ARG, "C_eta and C_sponge must be > 0");
    return WH_OK;
}

// grow a device buffer to at least n elements (contents are not kept);
// callers synchronise the stream first
template <class T>
static int grow_dev(T *&ptr, size_t &cap, size_t n)
{
    if (n <= cap) return WH_OK;
    if (ptr) WH_HIP(hipFree(ptr));
    ptr = nullptr;
    const size_t c = std::max(n, 2 * cap);
    cap = 0;
    WH_HIP(hipMalloc(&ptr, sizeof(T) * c));
    cap = c;
    return WH_OK;
}

static void free_patchset(PatchSet &ps)
{
    if (ps.d_patches) (void)hipFree(ps.d_patches);
    if (ps.d_recv_off) (void)hipFree(ps.d_recv_off);
    if (ps.d_recv_ids) (void)hipFree(ps.d_recv_ids);
    if (ps.d_chunks) (void)hipFree(ps.d_chunks);
    if (ps.d_gnb) (void)hipFree(ps.d_gnb);
    ps = PatchSet{};
}

static void destroy_graph(wh_pool *P)
{
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->gexec = nullptr;
    P->graph = nullptr;
    P->graph_topo = ~0ull;
}

}  // namespace wh

using namespace wh;

extern "C" {

const char *wh_last_error(void) { return g_err.c_str(); }

// marks the end of queued work that reads the pool's tables (on every return path)
struct TablesUse {
    wh_pool *P;
    ~TablesUse()
    {
        if (P && P->ev_tables && hipEventRecord(P->ev_tables, P->stream) == hipSuccess) P->tables_recorded = true;
    }
};

static int wait_tables(wh_pool *P)
{
    if (P->tables_recorded) WH_HIP(hipEventSynchronize(P->ev_tables));
    return WH_OK;
}

// host -> pinned staging -> device buffer `dst` (one host-to-device copy),
// ordered on the pool stream; the pinned space is reused only after a stream
// synchronisation
static int upload_staged(wh_pool *P, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return WH_OK;
    if (int rc = stage_begin(P, bytes)) return rc;
    void *h = stage_host(P, src, bytes);
    if (!h) return WH_E_HIP;
    WH_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, P->stream));
    return WH_OK;
}

int wh_version(void) { return 100; }

int wh_device_memory(size_t *free_bytes, size_t *total_bytes)
{
    WH_CHECK(free_bytes && total_bytes, WH_E_ARG, "wh_device_memory: null argument");
    WH_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return WH_OK;
}

int wh_pool_create(wh_pool **out, const wh_pool_desc *desc)
{
    WH_CHECK(out && desc, WH_E_ARG, "null argument");
    const wh_pool_desc &d = *desc;
    WH_CHECK(d.dim == 2 || d.dim == 3, WH_E_ARG, "dim must be 2 or 3");
    WH_CHECK(d.bs[0] >= 4 && d.bs[1] >= 4 && (d.dim == 2 || d.bs[2] >= 4), WH_E_ARG,
             "Bs must be >= 4");
    WH_CHECK(d.g >= 1 && d.neq >= 1 && d.nb_max >= 1 && d.n_work_slots >= 0, WH_E_ARG,
             "bad g/neq/nb_max/n_work_slots");
    WH_CHECK(d.n_tmp_eqn >= 0 && d.n_mask >= 0, WH_E_ARG, "bad n_tmp_eqn/n_mask");
    wh_pool *P = new wh_pool();
    P->d = d;
    P->nx = d.bs[0] + 2 * d.g;
    P->ny = d.bs[1] + 2 * d.g;
    P->nz = d.dim == 3 ? d.bs[2] + 2 * d.g : 1;
    P->cs = (size_t)P->nx * P->ny * P->nz;
    P->blk = P->cs * d.neq;
    P->tblk = P->cs * d.n_tmp_eqn;
    P->mblk = P->cs * d.n_mask;
    auto fail = [&](const char *what, hipError_t e) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        wh_pool_destroy(P);
        return e == hipErrorOutOfMemory ? WH_E_NOMEM : WH_E_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail("hipStreamCreate", e);
    {
        int dev = 0, ncu = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return fail("hipGetDevice", e);
        if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
            return fail("hipDeviceGetAttribute", e);
        P->n_cu = ncu > 0 ? ncu : 256;
        P->device = dev;
        wh::noalloc_setup(dev);
    }
    if ((e = hipMalloc(&P->block, sizeof(double) * P->blk * d.nb_max)) != hipSuccess)
        return fail("hipMalloc(hvy_block)", e);
    if (d.n_work_slots > 0 &&
        (e = hipMalloc(&P->work, sizeof(double) * P->blk * d.nb_max * d.n_work_slots)) != hipSuccess)
        return fail("hipMalloc(hvy_work)", e);
    if (P->tblk && (e = hipMalloc(&P->tmp, sizeof(double) * P->tblk * d.nb_max)) != hipSuccess)
        return fail("hipMalloc(hvy_tmp)", e);
    if (P->mblk &&
        (e = hipMalloc(&P->mask, sizeof(double) * P->mblk * (d.mask_per_block ? d.nb_max : 1))) !=
            hipSuccess)
        return fail("hipMalloc(hvy_mask)", e);
    if ((e = hipMalloc(&P->scal, sizeof(double) * DS_COUNT)) != hipSuccess)
        return fail("hipMalloc(scalars)", e);
    if ((e = hipHostMalloc((void **)&P->h_scal, sizeof(double) * DS_COUNT, hipHostMallocDefault)) !=
        hipSuccess)
        return fail("hipHostMalloc", e);
    if ((e = hipHostMalloc((void **)&P->h_time, sizeof(double) * DS_COUNT, hipHostMallocDefault)) !=
        hipSuccess)
        return fail("hipHostMalloc", e);
    if ((e = hipMemsetAsync(P->scal, 0, sizeof(double) * DS_COUNT, P->stream)) != hipSuccess)
        return fail("hipMemset", e);
    if ((e = hipMemsetAsync(P->block, 0, sizeof(double) * P->blk * d.nb_max, P->stream)) != hipSuccess)
        return fail("hipMemset", e);
    if (P->work && (e = hipMemsetAsync(P->work, 0, sizeof(double) * P->blk * d.nb_max * d.n_work_slots,
                                       P->stream)) != hipSuccess)
        return fail("hipMemset", e);
    if ((e = hipEventCreate(&P->ev_a)) != hipSuccess) return fail("hipEventCreate", e);
    if ((e = hipEventCreateWithFlags(&P->ev_tables, hipEventDisableTiming)) != hipSuccess)
        return fail("hipEventCreate", e);
    if ((e = hipEventCreate(&P->ev_b)) != hipSuccess) return fail("hipEventCreate", e);
    if ((e = hipEventCreateWithFlags(&P->ev_time, hipEventDisableTiming)) != hipSuccess)
        return fail("hipEventCreate", e);
    if ((e = hipStreamSynchronize(P->stream)) != hipSuccess) return fail("sync", e);
    *out = P;
    return WH_OK;
}

int wh_pool_destroy(wh_pool *P)
{
    if (!P) return WH_OK;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    destroy_graph(P);
    free_patchset(P->ps[0]);
    free_patchset(P->ps[1]);
    wh::free_wavelet_ws(P);
    for (double *p : {P->block, P->work, P->tmp, P->mask, P->scal, P->d_x0dx})
        if (p) (void)hipFree(p);
    if (P->d_active) (void)hipFree(P->d_active);
    if (P->d_grp) (void)hipFree(P->d_grp);
    if (P->d_dtb) (void)hipFree(P->d_dtb);
    if (P->d_dtg) (void)hipFree(P->d_dtg);
    if (P->d_gacm) (void)hipFree(P->d_gacm);
    if (P->h_scal) (void)hipHostFree(P->h_scal);
    if (P->h_time) (void)hipHostFree(P->h_time);
    if (P->ev_time) (void)hipEventDestroy(P->ev_time);
    if (P->ev_a) (void)hipEventDestroy(P->ev_a);
    if (P->ev_tables) (void)hipEventDestroy(P->ev_tables);
    if (P->ev_b) (void)hipEventDestroy(P->ev_b);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return WH_OK;
}

int wh_pool_block_bytes(const wh_pool *P, int array, size_t *bytes)
{
    WH_CHECK(P && bytes, WH_E_ARG, "null argument");
    size_t blk = 0;
    double *ptr = array_ptr(const_cast<wh_pool *>(P), array, &blk, nullptr);
    WH_CHECK(ptr || blk == 0, WH_E_ARG, "unknown array");
    *bytes = blk * sizeof(double);
    return WH_OK;
}

static int xfer(wh_pool *P, int array, int first, int count, void *host, bool up)
{
    WH_CHECK(P && host, WH_E_ARG, "null argument");
    size_t blk = 0;
    double *base = array_ptr(P, array, &blk, nullptr);
    WH_CHECK(base, WH_E_ARG, "unknown or unallocated array");
    const int nb = (array == WH_MASK && !P->d.mask_per_block) ? 1 : P->d.nb_max;
    WH_CHECK(first >= 0 && count >= 0 && first + count <= nb, WH_E_ARG, "block range out of pool");
    double *dev = base + (size_t)first * blk;
    const size_t bytes = sizeof(double) * blk * count;
    if (up)
        WH_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, P->stream));
    else
        WH_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    return WH_OK;
}

int wh_pool_upload(wh_pool *P, int array, int first, int count, const double *host)
{
    WH_PROF("wh_pool_upload");
    return xfer(P, array, first, count, const_cast<double *>(host), true);
}

int wh_pool_download(wh_pool *P, int array, int first, int count, double *host)
{
    WH_PROF("wh_pool_download");
    return xfer(P, array, first, count, host, false);
}

int wh_pool_device_ptr(wh_pool *P, int array, void **dptr)
{
    WH_CHECK(P && dptr, WH_E_ARG, "null argument");
    double *p = array_ptr(P, array, nullptr, nullptr);
    WH_CHECK(p, WH_E_ARG, "unknown or unallocated array");
    *dptr = p;
    return WH_OK;
}

int wh_pool_fill(wh_pool *P, int array, double value)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    size_t blk = 0;
    double *p = array_ptr(P, array, &blk, nullptr);
    WH_CHECK(p, WH_E_ARG, "unknown or unallocated array");
    const int nb = (array == WH_MASK && !P->d.mask_per_block) ? 1 : P->d.nb_max;
    std::vector<double> h(blk * nb, value);
    WH_HIP(hipMemcpyAsync(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    return WH_OK;
}

int wh_synchronize(wh_pool *P)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_HIP(hipStreamSynchronize(P->stream));
    return WH_OK;
}

int wh_set_active(wh_pool *P, const int32_t *act, int n, const double *x0dx)
{
    WH_PROF("wh_set_active");
    WH_CHECK(P && (n == 0 || (act && x0dx)), WH_E_ARG, "null argument");
    WH_CHECK(n >= 0 && n <= P->d.nb_max, WH_E_ARG, "active list longer than the pool");
    std::vector<char> seen(P->d.nb_max, 0);
    for (int k = 0; k < n; ++k) {
        WH_CHECK(act[k] >= 0 && act[k] < P->d.nb_max, WH_E_ARG, "hvy id out of pool range");
        WH_CHECK(!seen[act[k]], WH_E_ARG, "duplicate hvy id in active list");
        seen[act[k]] = 1;
        for (int d = 0; d < P->d.dim; ++d)
            WH_CHECK(x0dx[6 * k + 3 + d] > 0.0, WH_E_ARG, "block spacing must be positive");
    }
    if (int rc = wait_tables(P)) return rc;
    if (n > P->act_cap) {
        WH_HIP(hipStreamSynchronize(P->stream));  // before freeing buffers queued work may read
        if (P->d_active) WH_HIP(hipFree(P->d_active));
        if (P->d_x0dx) WH_HIP(hipFree(P->d_x0dx));
        P->d_active = nullptr;
        P->d_x0dx = nullptr;
        WH_HIP(hipMalloc(&P->d_active, sizeof(int32_t) * n));
        WH_HIP(hipMalloc(&P->d_x0dx, sizeof(double) * 6 * n));
        P->act_cap = n;
    }
    P->h_active.assign(act, act + n);
    P->h_x0dx.assign(x0dx, x0dx + 6 * n);
    if (int rc = upload_staged(P, P->d_active, act, sizeof(int32_t) * n)) return rc;
    if (int rc = upload_staged(P, P->d_x0dx, x0dx, sizeof(double) * 6 * n)) return rc;
    P->nact = n;
    P->ngrp = 0;  // block groups refer to the previous active list
    P->ngacm = 0;
    P->h_gparams.clear();
    P->topo_version++;
    return WH_OK;
}

// The patch table of a sync case from runs of patches (taken one after the
// other), without checks: stable counting sort by stage (1..3; within a
// stage the receivers keep their order), per-receiver rows, the chunks of
// the 3D ghost fill, upload.
static int set_patch_runs(wh_pool *P, int sync_case, const PatchRun *runs, int nruns)
{
    int npatch = 0;
    for (int r = 0; r < nruns; ++r) npatch += runs[r].n;
    if (int rc = wait_tables(P)) return rc;  // nothing queued reads the old table any more
    PatchSet &ps = P->ps[sync_case];
    ps.max_stage = 0;
    // stable counting sort by stage (1..3); the receivers keep their order
    // within a stage (rows of the per-receiver table are runs of one receiver)
    {
        int cnt[5] = {0, 0, 0, 0, 0};
        for (int r = 0; r < nruns; ++r)
            for (int i = 0; i < runs[r].n; ++i) ++cnt[runs[r].p[i].stage + 1];
        for (int st = 1; st < 5; ++st) cnt[st] += cnt[st - 1];
        ps.host.resize(npatch);
        for (int r = 0; r < nruns; ++r)
            for (int i = 0; i < runs[r].n; ++i) ps.host[cnt[runs[r].p[i].stage]++] = runs[r].p[i];
    }
    ps.h_recv_off.clear();
    ps.h_recv_ids.clear();
    ps.h_chunks.clear();
    for (int s = 0; s < 5; ++s) ps.stage_begin[s] = 0;
    for (int i = 0; i < npatch; ++i) {
        const wh_patch &q = ps.host[i];
        if (i == 0 || q.recv != ps.host[i - 1].recv || q.stage != ps.host[i - 1].stage) {
            ps.h_recv_off.push_back(i);
            ps.h_recv_ids.push_back(q.recv);
        }
        ps.max_stage = std::max(ps.max_stage, (int)q.stage);
        if (P->d.dim == 3) add_chunks(ps.h_chunks, i, q);  // 2D: one wavefront per patch, no chunks
    }
    {  // stage ranges over the (stage, receiver)-sorted table, in patches and in chunks
        int i = 0, c = 0;
        for (int s = 1; s <= 3; ++s) {
            ps.stage_begin[s] = i;
            ps.chunk_begin[s] = c;
            while (i < npatch && ps.host[i].stage == s) {
                const wh_patch &q = ps.host[i++];
                if (P->d.dim == 3) c += (q.n[0] * q.n[1] * q.n[2] + SYNC_CHUNK - 1) / SYNC_CHUNK;
            }
        }
        ps.stage_begin[4] = npatch;
        ps.chunk_begin[4] = c;
    }
    ps.h_recv_off.push_back(npatch);
    ps.n = npatch;
    // 2D, same-level copies only: the senders of every receiver's 8
    // directions (for the fused stages' halo gather)
    ps.gather_ok = sync_case == 0 && P->d.dim == 2 && npatch > 0;
    ps.gather_depth = -1;
    if (ps.gather_ok) {
        ps.h_gnb.assign((size_t)P->d.nb_max * 9, -1);
        const int g = P->d.g, b0 = P->d.bs[0], b1 = P->d.bs[1];
        auto dir = [&](int lo, int n, int bs, int &w) {  // -1 / 0 / +1, or 2: not a face or corner box
            if (lo + n == g && lo >= 0) return w = n, -1;
            if (lo == g + bs) return w = n, 1;
            if (lo == g && n == bs) return 0;
            return 2;
        };
        for (int i = 0; i < npatch && ps.gather_ok; ++i) {
            const wh_patch &q = ps.host[i];
            int wx = -1, wy = -1;
            const int dx = dir(q.recv_lo[0], q.n[0], b0, wx), dy = dir(q.recv_lo[1], q.n[1], b1, wy);
            const int w = wx >= 0 ? wx : wy;
            if ((q.op & 0x3ff) != WH_PATCH_COPY || q.stage != 1 || dx == 2 || dy == 2 || (dx == 0 && dy == 0) ||
                (wx >= 0 && wy >= 0 && wx != wy) || (ps.gather_depth >= 0 && w != ps.gather_depth) ||
                q.send_lo[0] != q.recv_lo[0] - dx * b0 || q.send_lo[1] != q.recv_lo[1] - dy * b1) {
                ps.gather_ok = false;
                break;
            }
            ps.gather_depth = w;
            int32_t &slot = ps.h_gnb[(size_t)q.recv * 9 + (dy + 1) * 3 + (dx + 1)];
            if (slot >= 0) ps.gather_ok = false;  // a direction filled twice
            slot = q.send;
        }
    }
    ps.nrecv = (int)ps.h_recv_ids.size();
    if (npatch) {
        if ((size_t)npatch > ps.cap_patches || ps.h_recv_off.size() > ps.cap_recv_off ||
            ps.h_recv_ids.size() > ps.cap_recv_ids || ps.h_chunks.size() > ps.cap_chunks)
            WH_HIP(hipStreamSynchronize(P->stream));  // before freeing buffers queued work may read
        if (int rc = grow_dev(ps.d_patches, ps.cap_patches, (size_t)npatch)) return rc;
        if (int rc = grow_dev(ps.d_recv_off, ps.cap_recv_off, ps.h_recv_off.size())) return rc;
        if (int rc = grow_dev(ps.d_recv_ids, ps.cap_recv_ids, ps.h_recv_ids.size())) return rc;
        if (int rc = grow_dev(ps.d_chunks, ps.cap_chunks, ps.h_chunks.size())) return rc;
        if (int rc = upload_staged(P, ps.d_chunks, ps.h_chunks.data(), sizeof(SyncChunk) * ps.h_chunks.size()))
            return rc;
        if (ps.gather_ok) {
            if (ps.h_gnb.size() > ps.cap_gnb) WH_HIP(hipStreamSynchronize(P->stream));
            if (int rc = grow_dev(ps.d_gnb, ps.cap_gnb, ps.h_gnb.size())) return rc;
            if (int rc = upload_staged(P, ps.d_gnb, ps.h_gnb.data(), sizeof(int32_t) * ps.h_gnb.size())) return rc;
        }
        if (int rc = upload_staged(P, ps.d_patches, ps.host.data(), sizeof(wh_patch) * npatch)) return rc;
        if (int rc = upload_staged(P, ps.d_recv_off, ps.h_recv_off.data(), sizeof(int32_t) * ps.h_recv_off.size()))
            return rc;
        if (int rc = upload_staged(P, ps.d_recv_ids, ps.h_recv_ids.data(), sizeof(int32_t) * ps.h_recv_ids.size()))
            return rc;
    }
    P->topo_version++;
    return WH_OK;
}

int wh_set_patches(wh_pool *P, int sync_case, const wh_patch *patches, int npatch)
{
    WH_PROF("wh_set_patches");
    WH_CHECK(P && (npatch == 0 || patches), WH_E_ARG, "null argument");
    WH_CHECK(sync_case == 0 || sync_case == 1, WH_E_ARG, "sync_case must be 0 or 1");
    WH_CHECK(npatch >= 0, WH_E_ARG, "negative patch count");
    const int ext[3] = {P->nx, P->ny, P->nz};
    for (int i = 0; i < npatch; ++i) {
        const wh_patch &q = patches[i];
        WH_CHECK(q.recv >= 0 && q.recv < P->d.nb_max && q.send >= 0 && q.send < P->d.nb_max,
                 WH_E_ARG, "patch block id out of pool range");
        WH_CHECK(q.op == WH_PATCH_COPY || q.op == WH_PATCH_RESTRICT || q.op == WH_PATCH_PREDICT,
                 WH_E_ARG, "unknown patch op");
        WH_CHECK(q.stage >= 1 && q.stage <= 3, WH_E_ARG, "patch stage must be 1..3");
        WH_CHECK(q.op != WH_PATCH_PREDICT || q.aux == 2 || q.aux == 4 || q.aux == 6, WH_E_ARG,
                 "predict patch order must be 2, 4 or 6");
        const int st = q.op == WH_PATCH_RESTRICT ? 2 : 1;
        for (int d = 0; d < 3; ++d) {
            WH_CHECK(q.n[d] >= 1, WH_E_ARG, "empty patch");
            WH_CHECK(q.recv_lo[d] >= 0 && q.recv_lo[d] + q.n[d] <= ext[d], WH_E_ARG,
                     "receiver box outside the block");
            if (q.op == WH_PATCH_PREDICT) {
                // coarse stencil of every point must stay inside the sender block
                const int N = q.aux / 2;
                const int fmin = q.send_lo[d], fmax = q.send_lo[d] + q.n[d] - 1;
                const bool interp = (d < P->d.dim);
                WH_CHECK(interp ? pred_range_ok(fmin, fmax, N, ext[d]) : (fmin == 0 && fmax == 0),
                         WH_E_ARG, "prediction stencil outside the sender block");
            } else {
                WH_CHECK(q.send_lo[d] >= 0 && q.send_lo[d] + st * (q.n[d] - 1) < ext[d], WH_E_ARG,
                         "sender box outside the block");
            }
        }
    }
    const PatchRun run{patches, npatch};
    return set_patch_runs(P, sync_case, &run, 1);
}

int wh_sync_ghosts(wh_pool *P, int array, int sync_case)
{
    TablesUse tables_use{P};
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_CHECK(sync_case == 0 || sync_case == 1, WH_E_ARG, "sync_case must be 0 or 1");
    WH_CHECK(array == WH_BLOCK || array >= WH_WORK, WH_E_ARG, "sync works on hvy_block/hvy_work");
    double *arr = array_ptr(P, array, nullptr, nullptr);
    WH_CHECK(arr, WH_E_ARG, "unknown or unallocated array");
    return timed(P, 1, [&] { return launch_sync(P, sync_case, arr, -1); });
}

int wh_rhs_acm(wh_pool *P, const wh_acm_params *p, double /*time*/, int in_array, int out_slot)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p, WH_E_ARG, "null argument");
    if (int rc = check_acm(P, p)) return rc;
    double *in = array_ptr(P, in_array, nullptr, nullptr);
    double *out = array_ptr(P, WH_WORK + out_slot, nullptr, nullptr);
    WH_CHECK(in && out && (in_array == WH_BLOCK || in_array >= WH_WORK), WH_E_ARG, "bad array");
    WH_CHECK(P->nact > 0 || P->h_active.empty(), WH_E_STATE, "no active list");
    const AcmDev a = make_acm_dev(p);
    return timed(P, 0, [&] { return launch_rhs(P, a, in, out, P->d_active, P->d_x0dx, P->nact); });
}

int wh_divergence_guard(wh_pool *P, int in_array, double *maxabs)
{
    TablesUse tables_use{P};
    WH_CHECK(P, WH_E_ARG, "null argument");
    double *in = array_ptr(P, in_array, nullptr, nullptr);
    WH_CHECK(in, WH_E_ARG, "bad array");
    int rc = launch_maxabs(P, in);
    if (rc) return rc;
    double m = 0;
    if ((rc = read_scalar(P, DS_MAXABS_BITS, &m))) return rc;
    if (maxabs) *maxabs = m;
    if (!(m <= 1.0e12)) {  // LIM_DIVERGED, module_globals.f90:52
        set_error("ACM fail: very very large values in state vector");
        return WH_E_DIVERGED;
    }
    return WH_OK;
}

int wh_dt_reduce(wh_pool *P, const wh_acm_params *p, double *dt_out)
{
    TablesUse tables_use{P};
    WH_PROF("wh_dt_reduce");
    WH_CHECK(P && p && dt_out, WH_E_ARG, "null argument");
    if (int rc = check_acm(P, p)) return rc;
    const AcmDev a = make_acm_dev(p);
    int rc = launch_dt(P, a);
    if (rc) return rc;
    return read_scalar(P, DS_DTMIN_BITS, dt_out);
}

int wh_rk_combine(wh_pool *P, int out_array, int base_array, int nterms, const int *term_arrays,
                  const double *coefs)
{
    TablesUse tables_use{P};
    WH_CHECK(P && (nterms == 0 || (term_arrays && coefs)), WH_E_ARG, "null argument");
    double *out = array_ptr(P, out_array, nullptr, nullptr);
    const double *base = array_ptr(P, base_array < 0 ? out_array : base_array, nullptr, nullptr);
    WH_CHECK(out && base, WH_E_ARG, "bad array");
    std::vector<const double *> terms(nterms);
    for (int t = 0; t < nterms; ++t) {
        terms[t] = array_ptr(P, term_arrays[t], nullptr, nullptr);
        WH_CHECK(terms[t], WH_E_ARG, "bad term array");
    }
    return launch_combine(P, out, base, nterms, terms.data(), coefs, P->d_active, P->nact);
}

int wh_rk_step(wh_pool *P, const wh_acm_params *p, const double *bt, int nrows, double /*time*/,
               double dt, int path)
{
    TablesUse tables_use{P};
    WH_PROF("wh_rk_step");
    WH_CHECK(P && p && bt, WH_E_ARG, "null argument");
    WH_CHECK(nrows >= 2 && nrows <= 9, WH_E_ARG, "butcher tableau size");
    if (int rc = check_acm(P, p)) return rc;
    const AcmDev a = make_acm_dev(p);
    if (path == 2) {
        WH_CHECK(fused_ok(bt, nrows), WH_E_ARG, "tableau not supported by the fused path");
        WH_CHECK(P->d.n_work_slots >= 3, WH_E_ARG, "fused path needs 3 work slots");
        int rc = launch_set_scalar(P, DS_DT, dt);
        if (rc) return rc;
        return fused_step(P, a, bt, nrows);
    }
    WH_CHECK(path == 0 || path == 1, WH_E_ARG, "path must be 0, 1 or 2");
    WH_CHECK(P->d.n_work_slots >= nrows, WH_E_ARG, "need nrows work slots");
    return generic_step(P, a, bt, nrows, dt, path);
}

int wh_set_block_groups(wh_pool *P, const int32_t *group, int n, int ngroups)
{
    WH_CHECK(P && (n == 0 || group), WH_E_ARG, "null argument");
    WH_CHECK(n == P->nact, WH_E_ARG, "one group per active block (set the active list first)");
    WH_CHECK(ngroups >= 1 && ngroups <= (1 << 20), WH_E_ARG, "ngroups out of range");
    for (int k = 0; k < n; ++k)
        WH_CHECK(group[k] >= 0 && group[k] < ngroups, WH_E_ARG, "group index out of range");
    if (int rc = wait_tables(P)) return rc;
    if (n > P->grp_cap || ngroups > P->dtg_cap)
        WH_HIP(hipStreamSynchronize(P->stream));  // before freeing buffers queued work may read
    if (n > P->grp_cap) {
        if (P->d_grp) WH_HIP(hipFree(P->d_grp));
        if (P->d_dtb) WH_HIP(hipFree(P->d_dtb));
        P->d_grp = nullptr;
        P->d_dtb = nullptr;
        WH_HIP(hipMalloc(&P->d_grp, sizeof(int32_t) * n));
        WH_HIP(hipMalloc(&P->d_dtb, sizeof(double) * n));
        P->grp_cap = n;
    }
    if (ngroups > P->dtg_cap) {
        if (P->d_dtg) WH_HIP(hipFree(P->d_dtg));
        P->d_dtg = nullptr;
        WH_HIP(hipMalloc(&P->d_dtg, sizeof(unsigned long long) * ngroups));
        P->dtg_cap = ngroups;
    }
    P->h_grp.assign(group, group + n);
    if (int rc = upload_staged(P, P->d_grp, group, sizeof(int32_t) * n)) return rc;
    P->ngrp = ngroups;
    P->ngacm = 0;  // new groups: their constants are set again (wh_set_group_params)
    P->h_gparams.clear();
    return WH_OK;
}

int wh_dt_reduce_groups(wh_pool *P, const wh_acm_params *p, double *dt_out)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p && dt_out, WH_E_ARG, "null argument");
    WH_CHECK(P->ngrp > 0, WH_E_STATE, "no block groups set");
    if (int rc = check_acm(P, p)) return rc;
    const AcmDev a = make_acm_dev(p);
    if (int rc = launch_dt_groups(P, a)) return rc;
    std::vector<unsigned long long> bits(P->ngrp);
    WH_HIP(hipMemcpyAsync(bits.data(), P->d_dtg, sizeof(unsigned long long) * P->ngrp,
                          hipMemcpyDeviceToHost, P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    for (int i = 0; i < P->ngrp; ++i) std::memcpy(dt_out + i, &bits[i], sizeof(double));
    return WH_OK;
}

int wh_rk_step_groups(wh_pool *P, const wh_acm_params *p, const double *bt, int nrows,
                      const double *dt_groups, int path)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p && bt && dt_groups, WH_E_ARG, "null argument");
    WH_CHECK(nrows >= 2 && nrows <= 9, WH_E_ARG, "butcher tableau size");
    WH_CHECK(P->ngrp > 0, WH_E_STATE, "no block groups set");
    WH_CHECK(P->ngacm == 0, WH_E_STATE, "per-group constants are stepped by wh_batch_rk_step");
    for (int i = 0; i < P->ngrp; ++i)
        WH_CHECK(dt_groups[i] >= 0.0, WH_E_ARG, "negative dt");
    if (int rc = check_acm(P, p)) return rc;
    const AcmDev a = make_acm_dev(p);
    std::vector<double> dt_pos(P->nact);
    for (int k = 0; k < P->nact; ++k) dt_pos[k] = dt_groups[P->h_grp[k]];
    if (path == 2) {
        WH_CHECK(fused_ok(bt, nrows), WH_E_ARG, "tableau not supported by the fused path");
        WH_CHECK(P->d.n_work_slots >= 3, WH_E_ARG, "fused path needs 3 work slots");
        if (int rc = upload_staged(P, P->d_dtb, dt_pos.data(), sizeof(double) * P->nact)) return rc;
        return fused_step(P, a, bt, nrows, P->d_dtb);
    }
    WH_CHECK(path == 1, WH_E_ARG, "block groups run on path 1 (per block) or 2 (fused)");
    WH_CHECK(P->d.n_work_slots >= nrows, WH_E_ARG, "need nrows work slots");
    return generic_step(P, a, bt, nrows, 1.0, 1, dt_pos.data());
}

int wh_time_set(wh_pool *P, double time, int iteration)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    int rc = launch_set_scalar(P, DS_TIME, time);
    if (!rc) rc = launch_set_scalar(P, DS_ITER, (double)iteration);
    if (!rc) rc = launch_set_scalar(P, DS_DT, 0.0);
    if (rc) return rc;
    WH_HIP(hipStreamSynchronize(P->stream));
    return WH_OK;
}

int wh_time_get(wh_pool *P, double *time, int *iteration, double *dt_last)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_HIP(hipMemcpyAsync(P->h_scal, P->scal, sizeof(double) * DS_COUNT, hipMemcpyDeviceToHost,
                          P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    if (time) *time = P->h_scal[DS_TIME];
    if (iteration) *iteration = (int)P->h_scal[DS_ITER];
    if (dt_last) *dt_last = P->h_scal[DS_DT];
    return WH_OK;
}

int wh_time_fetch(wh_pool *P)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_HIP(hipMemcpyAsync(P->h_time, P->scal, sizeof(double) * DS_COUNT, hipMemcpyDeviceToHost, P->stream));
    WH_HIP(hipEventRecord(P->ev_time, P->stream));
    P->time_fetch_pending = true;
    return WH_OK;
}

int wh_time_fetched(wh_pool *P, double *time, int *iteration, double *dt_last)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_CHECK(P->time_fetch_pending, WH_E_STATE, "wh_time_fetched: no wh_time_fetch before");
    WH_HIP(hipEventSynchronize(P->ev_time));
    if (time) *time = P->h_time[DS_TIME];
    if (iteration) *iteration = (int)P->h_time[DS_ITER];
    if (dt_last) *dt_last = P->h_time[DS_DT];
    return WH_OK;
}

static int one_device_step(wh_pool *P, const AcmDev &a, const double *bt, int nrows,
                           const TimeDev &tc)
{
    int rc;
    if ((rc = launch_dt(P, a))) return rc;
    if ((rc = launch_dt_finalize(P, tc))) return rc;
    if ((rc = fused_step(P, a, bt, nrows))) return rc;
    return launch_time_advance(P);
}

int wh_run_steps(wh_pool *P, const wh_acm_params *p, const double *bt, int nrows,
                 const wh_time_params *tp, int nsteps, int use_graph)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p && bt && tp, WH_E_ARG, "null argument");
    WH_CHECK(nrows >= 3 && nrows <= 9 && fused_ok(bt, nrows), WH_E_ARG,
             "tableau not supported by the fused path");
    WH_CHECK(P->d.n_work_slots >= 3, WH_E_ARG, "device time loop needs 3 work slots");
    if (int rc = check_acm(P, p)) return rc;
    WH_CHECK(nsteps >= 0, WH_E_ARG, "negative step count");
    const AcmDev a = make_acm_dev(p);
    const TimeDev tc = make_time_dev(tp);
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) {
            int rc = one_device_step(P, a, bt, nrows, tc);
            if (rc) return rc;
        }
        return WH_OK;
    }
    WH_CHECK(!P->timing, WH_E_STATE, "kernel timing cannot be captured into a graph");
    const bool same = P->gexec && P->graph_topo == P->topo_version &&
                      std::memcmp(&P->graph_p, p, sizeof(*p)) == 0 &&
                      std::memcmp(&P->graph_tc, &tc, sizeof(tc)) == 0 &&
                      P->graph_bt == std::vector<double>(bt, bt + nrows * nrows);
    if (!same) {
        destroy_graph(P);
        WH_HIP(hipStreamSynchronize(P->stream));
        WH_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
        int rc = one_device_step(P, a, bt, nrows, tc);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(P->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        WH_HIP(e);
        P->graph = g;
        WH_HIP(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        P->graph_topo = P->topo_version;
        P->graph_p = *p;
        P->graph_tc = tc;
        P->graph_bt.assign(bt, bt + nrows * nrows);
    }
    for (int s = 0; s < nsteps; ++s) WH_HIP(hipGraphLaunch(P->gexec, P->stream));
    return WH_OK;
}

int wh_kernel_timing_enable(wh_pool *P, int on)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    P->timing = on != 0;
    return WH_OK;
}

int wh_kernel_timing(wh_pool *P, int which, double *avg_ms, int *count)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    if (which < 0) {
        P->t_ms[0] = P->t_ms[1] = 0.0;
        P->t_cnt[0] = P->t_cnt[1] = 0;
        return WH_OK;
    }
    WH_CHECK(which <= 1, WH_E_ARG, "which must be 0 or 1");
    if (avg_ms) *avg_ms = P->t_cnt[which] ? P->t_ms[which] / P->t_cnt[which] : 0.0;
    if (count) *count = P->t_cnt[which];
    return WH_OK;
}

}  // extern "C"

extern "C" {

int wh_set_tuning(wh_pool *P, int key, int value)
{
    WH_CHECK(P, WH_E_ARG, "null argument");
    WH_CHECK(key == WH_TUNE_K3D_YSPLIT || key == WH_TUNE_K2D_STAGE || key == WH_TUNE_K3D_ZSPLIT ||
                 key == WH_TUNE_K3D_WAVES || key == WH_TUNE_K3D_XCD || key == WH_TUNE_GATHER,
             WH_E_ARG, "unknown tuning key");
    if (key == WH_TUNE_GATHER) {
        WH_CHECK(value == 0 || value == 1, WH_E_ARG, "GATHER is 0 or 1");
        P->gather = value;
        P->topo_version++;
        return WH_OK;
    }
    if (key == WH_TUNE_K3D_XCD) {
        WH_CHECK(value == 0 || value == 1, WH_E_ARG, "K3D_XCD is 0 or 1");
        P->k3d_xcd = value;
        P->topo_version++;
        return WH_OK;
    }
    if (key == WH_TUNE_K3D_WAVES) {
        WH_CHECK(value == 1 || value == 3 || value == 4, WH_E_ARG, "K3D_WAVES is 1, 3 or 4");
        P->k3d_waves = value;
        P->topo_version++;
        return WH_OK;
    }
    if (key == WH_TUNE_K3D_ZSPLIT) {
        WH_CHECK(value >= 1 && value <= 64, WH_E_ARG, "zsplit out of range");
        // a split that does not divide Bs2 into pieces of at least 2R planes
        // falls back to 1 at the launch
        P->k3d_zsplit = value;
        P->topo_version++;
        return WH_OK;
    }
    if (key == WH_TUNE_K2D_STAGE) {
        WH_CHECK(value >= 0 && value <= 4, WH_E_ARG, "K2D_STAGE is 0..4");
        P->k2d_stage = value;
        P->topo_version++;
        return WH_OK;
    }
    WH_CHECK(value >= -1 && value <= 64, WH_E_ARG, "ysplit out of range");
    // a split that is not valid for a launch's scheme falls back to the
    // automatic choice at that launch (zm_split)
    P->k3d_ysplit = value;
    P->topo_version++;  // captured graphs hold the old launch shape
    return WH_OK;
}

int wh_get_tuning(wh_pool *P, int key, int *value)
{
    WH_CHECK(P && value, WH_E_ARG, "null argument");
    WH_CHECK(key == WH_TUNE_K3D_YSPLIT || key == WH_TUNE_K2D_STAGE || key == WH_TUNE_K3D_ZSPLIT ||
                 key == WH_TUNE_K3D_WAVES || key == WH_TUNE_K3D_XCD || key == WH_TUNE_GATHER,
             WH_E_ARG, "unknown tuning key");
    *value = key == WH_TUNE_K2D_STAGE   ? P->k2d_stage
             : key == WH_TUNE_GATHER    ? P->gather
             : key == WH_TUNE_K3D_ZSPLIT ? P->k3d_zsplit
             : key == WH_TUNE_K3D_WAVES  ? P->k3d_waves
             : key == WH_TUNE_K3D_XCD    ? P->k3d_xcd
                                         : P->k3d_ysplit;
    return WH_OK;
}

int wh_time_stage_kernel(wh_pool *P, const wh_acm_params *p, int reps, double *ms)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p && ms, WH_E_ARG, "null argument");
    if (int rc = check_acm(P, p)) return rc;
    WH_CHECK(P->d.n_work_slots >= 3, WH_E_ARG, "stage-kernel timing needs 3 work slots");
    WH_CHECK(P->nact > 0, WH_E_STATE, "stage-kernel timing needs an active list");
    WH_CHECK(!P->timing, WH_E_STATE, "disable kernel timing first");
    reps = std::max(1, std::min(reps, 1000));
    const AcmDev a = make_acm_dev(p);
    double *acc = array_ptr(P, WH_WORK + 0, nullptr, nullptr), *tout = array_ptr(P, WH_WORK + 1, nullptr, nullptr),
           *sout = array_ptr(P, WH_WORK + 2, nullptr, nullptr);
    // the four stages of classic RK4 (first, two middle, last), reading the
    // state and writing only work slots; with the halo-gather table the
    // steps use, the launches gather their halos the same way
    const int32_t *gnb = gather_table(P, a);
    auto rk4 = [&]() -> int {
        const double ca[4] = {0.5, 0.5, 1.0, 0.0}, cb[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        const int mode[4] = {0, 1, 1, 2};
        for (int j = 0; j < 4; ++j)
            if (int rc = launch_fused_stage(P, a, P->block, mode[j] == 2 ? nullptr : tout,
                                            mode[j] == 1 ? P->block : nullptr, acc, mode[j] == 2 ? sout : nullptr,
                                            mode[j], ca[j], cb[j], nullptr, gnb))
                return rc;
        return WH_OK;
    };
    if (int rc = rk4()) return rc;  // warm
    WH_HIP(hipEventRecord(P->ev_a, P->stream));
    for (int r = 0; r < reps; ++r)
        if (int rc = rk4()) return rc;
    WH_HIP(hipEventRecord(P->ev_b, P->stream));
    WH_HIP(hipEventSynchronize(P->ev_b));
    float t = 0.f;
    WH_HIP(hipEventElapsedTime(&t, P->ev_a, P->ev_b));
    *ms = (double)t / (4.0 * reps);
    return WH_OK;
}

int wh_autotune(wh_pool *P, const wh_acm_params *p, int reps, int *best_value, double *best_ms)
{
    TablesUse tables_use{P};
    WH_CHECK(P && p, WH_E_ARG, "null argument");
    if (int rc = check_acm(P, p)) return rc;
    WH_CHECK(P->d.n_work_slots >= 2, WH_E_ARG, "autotune needs work slot 1 as scratch");
    WH_CHECK(P->nact > 0, WH_E_STATE, "autotune needs an active list");
    WH_CHECK(!P->timing, WH_E_STATE, "disable kernel timing before autotuning");
    reps = std::max(1, std::min(reps, 100));
    // 3D: the flat kernel and the z-marching (y-split, z-split) pairs; 2D:
    // the load order of the mask and stage operands.  The variant value is
    // reported in best_value (3D: the y-split; the z-split is set too).
    struct Var {
        int v, zs, waves, xcd;
    };
    std::vector<Var> cand;
    const bool three = P->d.dim == 3;
    if (three) {
        for (int xcd : {0, 1})
            for (int wv : {1, 4}) cand.push_back({0, 1, wv, xcd});
        for (int ys = 1; ys <= 8; ++ys)
            for (int zs = 1; zs <= 8; ++zs)
                if (zm_valid(P->d.bs[0], P->d.bs[1], p->scheme, ys) && zm_zvalid(P->d.bs[2], p->scheme, zs))
                    for (int wv : {1, 3}) cand.push_back({ys, zs, wv, 0});
    } else {
        for (int v = 0; v <= 4; ++v) cand.push_back({v, 1, 1, 0});
    }
    int &knob = three ? P->k3d_ysplit : P->k2d_stage;
    const AcmDev a = make_acm_dev(p);
    double *out = array_ptr(P, WH_WORK + 1, nullptr, nullptr);
    // with three work slots the timed kernel is the fused RK stage (a middle
    // stage: reads the block array, the accumulator in work slot 0, writes
    // work slots 0 and 1), otherwise the RHS; nothing else is written
    const bool fused = P->d.n_work_slots >= 3;
    double *acc = array_ptr(P, WH_WORK + 0, nullptr, nullptr);
    auto launch = [&]() -> int {
        if (fused) return launch_fused_stage(P, a, P->block, out, P->block, acc, nullptr, 1, 0.5, 1.0 / 3.0, nullptr);
        return launch_rhs(P, a, P->block, out, P->d_active, P->d_x0dx, P->nact);
    };
    const Var saved{knob, P->k3d_zsplit, P->k3d_waves, P->k3d_xcd};
    Var best = saved;
    double bms = 1e300;
    auto set = [&](const Var &v) {
        knob = v.v;
        if (three) {
            P->k3d_zsplit = v.zs;
            P->k3d_waves = v.waves;
            P->k3d_xcd = v.xcd;
        }
    };
    for (const Var &v : cand) {
        set(v);
        int rc = launch();  // warm
        if (rc) {
            set(saved);
            return rc;
        }
        WH_HIP(hipEventRecord(P->ev_a, P->stream));
        for (int r = 0; r < reps; ++r)
            if ((rc = launch())) {
                set(saved);
                return rc;
            }
        WH_HIP(hipEventRecord(P->ev_b, P->stream));
        WH_HIP(hipEventSynchronize(P->ev_b));
        float ms = 0.f;
        WH_HIP(hipEventElapsedTime(&ms, P->ev_a, P->ev_b));
        if (ms / reps < bms) {
            bms = ms / reps;
            best = v;
        }
    }
    set(best);
    P->topo_version++;
    if (best_value) *best_value = best.v;
    if (best_ms) *best_ms = bms;
    return WH_OK;
}

}  // extern "C"

namespace wh {
int launch_refine(wh_pool *P, double *arr, const int32_t *d_pairs, int npairs, int order);
}

extern "C" int wh_refine_blocks(wh_pool *P, int array, const int32_t *triples, int ntriples,
                                int order)
{
    WH_PROF("wh_refine_blocks");
    WH_CHECK(P && (ntriples == 0 || triples), WH_E_ARG, "null argument");
    WH_CHECK(order == 2 || order == 4 || order == 6, WH_E_ARG, "order must be 2, 4 or 6");
    WH_CHECK(ntriples >= 0, WH_E_ARG, "negative count");
    double *arr = array_ptr(P, array, nullptr, nullptr);
    WH_CHECK(arr && (array == WH_BLOCK || array >= WH_WORK), WH_E_ARG, "bad array");
    const int N = order / 2;
    // the daughters' points reach the mother's interpolated grid at fine index
    // g .. g + 2*Bs + g - 1 + ...: every stencil must stay in [0, nx)
    for (int d = 0; d < P->d.dim; ++d) {
        const int ext = d == 0 ? P->nx : d == 1 ? P->ny : P->nz;
        const int fmin = P->d.g, fmax = P->d.g + P->d.bs[d] + ext - 1;
        WH_CHECK(pred_range_ok(fmin, fmax, N, ext), WH_E_ARG,
                 "too few ghost nodes for this predictor order");
    }
    std::vector<char> mother(P->d.nb_max, 0), daughter(P->d.nb_max, 0);
    for (int i = 0; i < ntriples; ++i) {
        const int m = triples[3 * i], dgh = triples[3 * i + 1], k = triples[3 * i + 2];
        WH_CHECK(m >= 0 && m < P->d.nb_max && dgh >= 0 && dgh < P->d.nb_max, WH_E_ARG,
                 "hvy id out of pool range");
        WH_CHECK(k >= 0 && k < (1 << P->d.dim), WH_E_ARG, "daughter digit out of range");
        WH_CHECK(!daughter[dgh], WH_E_ARG, "daughter slot written twice");
        mother[m] = 1;
        daughter[dgh] = 1;
    }
    for (int h = 0; h < P->d.nb_max; ++h)
        WH_CHECK(!(mother[h] && daughter[h]), WH_E_ARG, "a slot is both mother and daughter");
    if (ntriples == 0) return WH_OK;
    const size_t bytes = sizeof(int32_t) * 3 * ntriples;
    if (int rc = stage_begin(P, bytes)) return rc;
    const int32_t *d_tr = (const int32_t *)stage_put(P, triples, bytes);
    if (!d_tr) return WH_E_HIP;
    return launch_refine(P, arr, d_tr, ntriples, order);
}

namespace wh {
int set_patch_table(wh_pool *P, int sync_case, const PatchRun *runs, int nruns)
{
    return set_patch_runs(P, sync_case, runs, nruns);
}
}  // namespace wh
