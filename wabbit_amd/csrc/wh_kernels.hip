// wh_kernels.hip -- FP64 HIP kernels of the block time step for gfx950.
//
// Every kernel works on the device block pool in the reference's heavy-data
// layout [Nb][neq][nz][ny][nx] (allocate_forest.f90:228) and on the active
// list hvy_active.  Arithmetic follows the reference's term order and is
// compiled with -ffp-contract=off, so each product and sum rounds on its own
// exactly as the CPU code does.
#include "wh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace wh {

// ---------------------------------------------------------------------------
// FD stencils (rhs_ACM.f90:335-344 and :958-965).  Coefficients are compile-
// time constants; each stencil takes a value accessor f(o) -> field at offset
// o along one direction, so the same code reads global memory, an LDS tile or
// a register queue (constant offsets keep the queue in registers).
// ---------------------------------------------------------------------------
constexpr double A4_m2 = 1.0 / 12.0, A4_m1 = -2.0 / 3.0, A4_p1 = 2.0 / 3.0, A4_p2 = -1.0 / 12.0;
constexpr double B4_m2 = -1.0 / 12.0, B4_m1 = 4.0 / 3.0, B4_0 = -5.0 / 2.0, B4_p1 = 4.0 / 3.0,
                 B4_p2 = -1.0 / 12.0;
constexpr double TW_m3 = -0.02651995, TW_m2 = 0.18941314, TW_m1 = -0.79926643,
                 TW_p1 = 0.79926643, TW_p2 = -0.18941314, TW_p3 = 0.02651995;
constexpr double A6_m3 = -1.0 / 60.0, A6_m2 = 3.0 / 20.0, A6_m1 = -3.0 / 4.0, A6_p1 = 3.0 / 4.0,
                 A6_p2 = -3.0 / 20.0, A6_p3 = 1.0 / 60.0;
constexpr double B6_m3 = 1.0 / 90.0, B6_m2 = -3.0 / 20.0, B6_m1 = 3.0 / 2.0, B6_0 = -49.0 / 18.0,
                 B6_p1 = 3.0 / 2.0, B6_p2 = -3.0 / 20.0, B6_p3 = 1.0 / 90.0;

template <int S> struct Reach;
template <> struct Reach<WH_FD2> { static constexpr int R = 1; };
template <> struct Reach<WH_FD4> { static constexpr int R = 2; };
template <> struct Reach<WH_FD4_TW> { static constexpr int R = 3; };
template <> struct Reach<WH_FD6> { static constexpr int R = 3; };

// first derivative; terms summed from the most negative offset up
template <int S, class F>
__device__ __forceinline__ double d1(const F &f, double inv)
{
    if constexpr (S == WH_FD2) {
        return (f(1) - f(-1)) * inv * 0.5;
    } else if constexpr (S == WH_FD4) {
        return (A4_m2 * f(-2) + A4_m1 * f(-1) + A4_p1 * f(1) + A4_p2 * f(2)) * inv;
    } else if constexpr (S == WH_FD4_TW) {
        return (TW_m3 * f(-3) + TW_m2 * f(-2) + TW_m1 * f(-1) + TW_p1 * f(1) + TW_p2 * f(2) +
                TW_p3 * f(3)) * inv;
    } else {
        return (A6_m3 * f(-3) + A6_m2 * f(-2) + A6_m1 * f(-1) + A6_p1 * f(1) + A6_p2 * f(2) +
                A6_p3 * f(3)) * inv;
    }
}

// first derivative of the product fa*fb (skew-symmetric form); terms (c*fa)*fb
template <int S, class F, class G>
__device__ __forceinline__ double d1p(const F &fa, const G &fb, double inv)
{
    if constexpr (S == WH_FD2) {
        return (fa(1) * fb(1) - fa(-1) * fb(-1)) * inv * 0.5;
    } else if constexpr (S == WH_FD4) {
        return (A4_m2 * fa(-2) * fb(-2) + A4_m1 * fa(-1) * fb(-1) + A4_p1 * fa(1) * fb(1) +
                A4_p2 * fa(2) * fb(2)) * inv;
    } else if constexpr (S == WH_FD4_TW) {
        return (TW_m3 * fa(-3) * fb(-3) + TW_m2 * fa(-2) * fb(-2) + TW_m1 * fa(-1) * fb(-1) +
                TW_p1 * fa(1) * fb(1) + TW_p2 * fa(2) * fb(2) + TW_p3 * fa(3) * fb(3)) * inv;
    } else {
        return (A6_m3 * fa(-3) * fb(-3) + A6_m2 * fa(-2) * fb(-2) + A6_m1 * fa(-1) * fb(-1) +
                A6_p1 * fa(1) * fb(1) + A6_p2 * fa(2) * fb(2) + A6_p3 * fa(3) * fb(3)) * inv;
    }
}

template <int S, class F>
__device__ __forceinline__ double d2(const F &f, double inv2)
{
    if constexpr (S == WH_FD2) {
        return (f(-1) - 2.0 * f(0) + f(1)) * inv2;
    } else if constexpr (S == WH_FD4 || S == WH_FD4_TW) {
        return (B4_m2 * f(-2) + B4_m1 * f(-1) + B4_0 * f(0) + B4_p1 * f(1) + B4_p2 * f(2)) * inv2;
    } else {
        return (B6_m3 * f(-3) + B6_m2 * f(-2) + B6_m1 * f(-1) + B6_0 * f(0) + B6_p1 * f(1) +
                B6_p2 * f(2) + B6_p3 * f(3)) * inv2;
    }
}

// ACM right-hand side at one point (RHS_2D_acm / RHS_3D_acm).  A(c, d, o) is
// component c of the state at offset o along direction d from the point.
// mk: mask component 0 at this point (components mcs apart) or null.
template <int DIM, int S, bool SKEW, class ACC>
__device__ __forceinline__ void acm_point(const ACC &A, const double *inv, const double *inv2,
                                          const AcmDev &a, const double *__restrict__ mk,
                                          ptrdiff_t mcs, double *r)
{
    double uv[DIM + 1];
    double du[DIM + 1][DIM];
#pragma unroll
    for (int c = 0; c <= DIM; ++c) {
        uv[c] = A(c, 0, 0);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            du[c][d] = d1<S>([&](int o) { return A(c, d, o); }, inv[d]);
    }
    double chi = 0.0;
    if (a.penal && mk) {
        const int color = (int)mk[4 * mcs];
        chi = mk[0] * (color == 0 ? 0.0 : a.Cinv);
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        double lap = d2<S>([&](int o) { return A(c, 0, o); }, inv2[0]);
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            lap = lap + d2<S>([&](int o) { return A(c, d, o); }, inv2[d]);
        double t;
        if constexpr (SKEW) {
            double sum = d1p<S>([&](int o) { return A(c, 0, o); },
                                [&](int o) { return A(0, 0, o); }, inv[0]);
#pragma unroll
            for (int d = 1; d < DIM; ++d)
                sum = sum + d1p<S>([&](int o) { return A(c, d, o); },
                                   [&](int o) { return A(d, d, o); }, inv[d]);
#pragma unroll
            for (int d = 0; d < DIM; ++d) sum = sum + uv[d] * du[c][d];
            t = -(0.5 * sum);
        } else {
            t = -(uv[0] * du[c][0]);
#pragma unroll
            for (int d = 1; d < DIM; ++d) t = t - uv[d] * du[c][d];
        }
        t = t - du[DIM][c];
        t = t + a.nu * lap;
        if (a.penal && mk) t = t + (-(chi * (uv[c] - mk[(1 + c) * mcs])));
        r[c] = t;
    }
    double div = du[0][0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) div = div + du[d][d];
    r[DIM] = -(a.c02 * div) - a.gamma_p * uv[DIM];
    if (a.sponge && mk) {
        const double spo = mk[5 * mcs] * a.Csp_inv;
#pragma unroll
        for (int c = 0; c < DIM; ++c) r[c] = r[c] - (uv[c] - a.ums[c]) * spo;
        r[DIM] = r[DIM] - uv[DIM] * spo;
    }
}

// accessor over one strided storage: base = component 0 of the point,
// components cs apart, strides s[d] along direction d
template <int DIM>
struct StridedAcc {
    const double *base;
    ptrdiff_t cs;
    ptrdiff_t s[3];
    __device__ __forceinline__ double operator()(int c, int d, int o) const
    {
        return base[c * cs + o * s[d]];
    }
};

// Runge-Kutta epilogue of the fused stage kernel.
//   mode 0 (first stage): u0 = input state; acc = u0 + (dt*b)*k; tout = u0 + (dt*a)*k
//   mode 1 (middle):      acc = acc + (dt*b)*k;  tout = u0 + (dt*a)*k
//   mode 2 (last):        sout = acc + (dt*b)*k
// b == 0 skips the acc term as the reference skips |coef| < 1e-8
// (runge_kutta_generic.f90:99,144).
// The accumulator lives only through one RK step and holds block interiors
// only: point p (x fastest) of component c of block b at b*NC*npts + c*npts
// + p, so its reads and writes cover whole cache lines.  u0, tout and sout
// are in the ghost-including block layout.
struct StageArgs {
    const double *u0;
    double *acc, *tout, *sout;
    const double *scal;
    const double *dtb;  // per active-list position dt (independent forests), or null
    double a, b;
    int mode, use_b, use_a;
    int full_rows;      // tout written in whole rows (stage_ghost_cols)
    const int32_t *gnb;  // 2D halo gather: senders of each block's 9 directions (by hvy id), or null
};

struct NoEpilogue {};

// The arrays of one block of a fused stage: base pointers of block b (the
// same in every lane, so they stay in scalar registers); the epilogue
// addresses points by 32-bit offsets inside the block.
struct StageBlk {
    const double *u0;
    double *acc, *tout, *sout;
};

template <int NC>
__device__ __forceinline__ StageBlk stage_blk(const StageArgs &sa, int b, size_t blk, int npts)
{
    const size_t o = (size_t)b * blk;
    return StageBlk{sa.u0 ? sa.u0 + o : nullptr, sa.acc + (size_t)b * NC * npts, sa.tout ? sa.tout + o : nullptr,
                    sa.sout ? sa.sout + o : nullptr};
}

// point p of the interior (offset gi in the block, components cs apart)
template <int NC>
__device__ __forceinline__ void stage_epilogue(const StageArgs &sa, const StageBlk &B, int k, int gi, int cs, int p,
                                               int npts, const double *r, const double *ucen, ptrdiff_t ucs)
{
    const double dt = sa.dtb ? sa.dtb[k] : sa.scal[DS_DT];
    const double ca = dt * sa.a, cb = dt * sa.b;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int i = gi + c * cs, ia = p + c * npts;
        if (sa.mode == 0) {
            const double u0 = ucen[c * ucs];
            B.acc[ia] = sa.use_b ? u0 + cb * r[c] : u0;
            if (sa.use_a) B.tout[i] = u0 + ca * r[c];
        } else if (sa.mode == 1) {
            const double u0 = B.u0[i];
            if (sa.use_b) B.acc[ia] = B.acc[ia] + cb * r[c];
            if (sa.use_a) B.tout[i] = u0 + ca * r[c];
        } else {
            B.sout[i] = sa.use_b ? B.acc[ia] + cb * r[c] : B.acc[ia];
        }
    }
}

// the fused epilogue with the acc / u0 values of the point already loaded
// (accv, u0v: one value per component)
template <int NC>
__device__ __forceinline__ void stage_epilogue_pre(const StageArgs &sa, const StageBlk &B, int k, int gi, int cs,
                                                   int p, int npts, const double *r, const double *ucen,
                                                   ptrdiff_t ucs, const double *accv, const double *u0v)
{
    const double dt = sa.dtb ? sa.dtb[k] : sa.scal[DS_DT];
    const double ca = dt * sa.a, cb = dt * sa.b;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int i = gi + c * cs, ia = p + c * npts;
        if (sa.mode == 0) {
            const double u0 = ucen[c * ucs];
            B.acc[ia] = sa.use_b ? u0 + cb * r[c] : u0;
            if (sa.use_a) B.tout[i] = u0 + ca * r[c];
        } else if (sa.mode == 1) {
            if (sa.use_b) B.acc[ia] = accv[c] + cb * r[c];
            if (sa.use_a) B.tout[i] = u0v[c] + ca * r[c];
        } else {
            B.sout[i] = sa.use_b ? accv[c] + cb * r[c] : accv[c];
        }
    }
}

// Zeros in the ghost columns (x < g and x >= g + bs0) of one interior row of
// the next stage input, written by the threads of the row's first 2g points
// (where sa.full_rows: stages whose arrays exceed the last-level cache):
// with the interior written point by point, every cache line of the row is
// then written whole, so no line is merged with its old contents in memory.
// Before the next stage the ghost fill overwrites every ghost value its
// stencil reads (the stencils run along the axes, within g_rhs of the
// interior); the other ghost values of a stage input are not read.  row: the
// offset of the row's first (x = 0) element in the block.
template <int NC>
__device__ __forceinline__ void stage_ghost_cols(const StageArgs &sa, const StageBlk &B, int row, int cs, int ix,
                                                 int bs0, int g)
{
    if (!sa.full_rows || sa.mode == 2 || !sa.use_a) return;
    for (int x = ix; x < 2 * g; x += bs0) {
        const int i = row + (x < g ? x : bs0 + x);
#pragma unroll
        for (int c = 0; c < NC; ++c) B.tout[i + c * cs] = 0.0;
    }
}

// stage_ghost_cols for every interior row of a 2D block, by the whole
// workgroup after its points (no per-point registers live meanwhile)
template <int NC>
__device__ __forceinline__ void stage_ghost_cols_2d(const StageArgs &sa, const StageBlk &B, const Geo &G)
{
    if (!sa.full_rows || sa.mode == 2 || !sa.use_a) return;
    const int w = 2 * G.g, cs = (int)G.cs;
    for (int i = threadIdx.x; i < G.bs1 * w; i += blockDim.x) {
        const int r = i / w, x = i - r * w;
        const int o = (G.g + r) * G.nx + (x < G.g ? x : G.bs0 + x);
#pragma unroll
        for (int c = 0; c < NC; ++c) B.tout[o + c * cs] = 0.0;
    }
}

// 2D: one workgroup per block, the (Bs+2R)^2 x 3 halo tile staged in LDS.
// With PRE each thread loads the mask and, for the fused stages, the acc /
// u0 values of its first point into registers before the tile barrier, so
// that these reads overlap the tile load instead of following it.
// rhs2d_block is the work of one workgroup on block b (active position k):
// shared by the one-pool kernel and the multi-pool kernel.  mk: the mask of
// block b or null.
// LOADED: the caller has already put the tile in LDS (and synchronised).
template <int S, bool SKEW, bool FUSED, bool PRE, bool LOADED = false>
__device__ __forceinline__ void rhs2d_block(double *tile, const double *__restrict__ in, double *__restrict__ out,
                                            const double *__restrict__ mk, int b, int k,
                                            const double *__restrict__ x0dx, const Geo &G, const AcmDev &a,
                                            const StageArgs &sa)
{
    static_assert(!(PRE && LOADED), "the operand prefetch belongs to the tile load");
    constexpr int R = Reach<S>::R;
    const int tx = G.bs0 + 2 * R, ty = G.bs1 + 2 * R, tcs = tx * ty;
    const int npts = G.bs0 * G.bs1;
    const size_t base = (size_t)b * G.blk;
    const int p0 = threadIdx.x;
    const int cs = (int)G.cs;
    const StageBlk SB = FUSED ? stage_blk<3>(sa, b, G.blk, npts) : StageBlk{};
    double mpre[6] = {0, 0, 0, 0, 0, 0}, apre[3] = {0, 0, 0}, upre[3] = {0, 0, 0};
    if constexpr (PRE) {
        if (p0 < npts) {
            const int iy = p0 / G.bs0, ix = p0 - iy * G.bs0;
            const int gi = (G.g + iy) * G.nx + (G.g + ix);
            if (mk) {
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    if (c != 3) mpre[c] = mk[c * G.cs + gi];  // component 3 (u_s,z) is 3D only
            }
            if (FUSED && sa.mode >= 1) {
#pragma unroll
                for (int c = 0; c < 3; ++c) apre[c] = SB.acc[c * npts + p0];
                if (sa.mode == 1) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) upre[c] = SB.u0[c * cs + gi];
                }
            }
        }
    }
    if constexpr (!LOADED) {
        const double *u = in + (size_t)b * G.blk;
        if (FUSED && sa.gnb) {
            // the halo from the neighbours' interiors (same-level senders, the
            // values the ghost fill would copy): tile point (x, y) relative to
            // the interior lies in direction (dx, dy) of the block
            const int32_t *nb = sa.gnb + (size_t)b * 9;
            for (int i = threadIdx.x; i < 3 * tcs; i += blockDim.x) {
                const int c = i / tcs, rr = i - c * tcs;
                const int y = rr / tx - R, x = rr - (y + R) * tx - R;
                const int dx = x < 0 ? -1 : (x >= G.bs0 ? 1 : 0), dy = y < 0 ? -1 : (y >= G.bs1 ? 1 : 0);
                const double *src = (dx | dy) ? in + (size_t)nb[(dy + 1) * 3 + dx + 1] * G.blk : u;
                tile[i] = src[c * G.cs + (size_t)(G.g + y - dy * G.bs1) * G.nx + (G.g + x - dx * G.bs0)];
            }
        } else {
            for (int i = threadIdx.x; i < 3 * tcs; i += blockDim.x) {
                const int c = i / tcs, rr = i - c * tcs;
                const int iy = rr / tx, ix = rr - iy * tx;
                tile[i] = u[c * G.cs + (size_t)(G.g - R + iy) * G.nx + (G.g - R + ix)];
            }
        }
        __syncthreads();
    }
    (void)ty;
    const double dx0 = x0dx[6 * k + 3], dx1 = x0dx[6 * k + 4];
    const double inv[2] = {1.0 / dx0, 1.0 / dx1};
    const double inv2[2] = {1.0 / (dx0 * dx0), 1.0 / (dx1 * dx1)};
    for (int pnt = threadIdx.x; pnt < npts; pnt += blockDim.x) {
        const int iy = pnt / G.bs0, ix = pnt - iy * G.bs0;
        const double *uc = tile + (iy + R) * tx + (ix + R);
        const int gi = (G.g + iy) * G.nx + (G.g + ix);
        double r[3];
        const StridedAcc<2> A{uc, tcs, {1, tx, 0}};
        if (PRE && pnt == p0) {
            // separate calls (not a pointer select) keep mpre in registers
            if (mk) acm_point<2, S, SKEW>(A, inv, inv2, a, mpre, 1, r);
            else acm_point<2, S, SKEW>(A, inv, inv2, a, nullptr, 1, r);
            if constexpr (FUSED) stage_epilogue_pre<3>(sa, SB, k, gi, cs, pnt, npts, r, uc, tcs, apre, upre);
        } else {
            acm_point<2, S, SKEW>(A, inv, inv2, a, mk ? mk + gi : nullptr, G.cs, r);
            if constexpr (FUSED) stage_epilogue<3>(sa, SB, k, gi, cs, pnt, npts, r, uc, tcs);
        }
        if constexpr (!FUSED) {
#pragma unroll
            for (int c = 0; c < 3; ++c) out[base + gi + c * G.cs] = r[c];
        }
    }
    if constexpr (FUSED) stage_ghost_cols_2d<3>(sa, SB, G);
}

template <int S, bool SKEW, bool FUSED, bool PRE>
__global__ __launch_bounds__(256) void k_rhs2d(const double *__restrict__ in, double *__restrict__ out,
                                               const double *__restrict__ mask,
                                               const int32_t *__restrict__ act,
                                               const double *__restrict__ x0dx, Geo G, AcmDev a,
                                               StageArgs sa)
{
    extern __shared__ double tile[];
    const int k = blockIdx.x;
    const int b = act[k];
    const double *mk = mask ? mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) : nullptr;
    rhs2d_block<S, SKEW, FUSED, PRE>(tile, in, out, mk, b, k, x0dx, G, a, sa);
}

// ---------------------------------------------------------------------------
// Several pools in one launch (independent forests of a batch, any block
// size): entry k of the launch is position k - ent0 of the active list of
// the pool whose entry range holds k (pools[] in entry order), with that
// position's origin/spacing and block group g = goff + grp[..] (dt dtg[g],
// constants gacm[g]), all read from the pool's own device tables; once per
// step k_entries resolves every entry into a record (pool, hvy id, group,
// position) that the stage kernels read with one load per block.  The
// stage's input and output arrays are selected per pool from arr[]:
// 0 hvy_block, 1 work slot 0 (the RK accumulator), 2 / 3 work slots 1 / 2.
// The kernels are persistent (grid-stride over the entries, their count read
// from device memory), so one captured launch serves any number of blocks.
// ---------------------------------------------------------------------------
// the last pool whose first entry (s < 0), first chunk of sync stage s (0..2) or
// first patch of sync stage s - 3 (3..5) is <= x
__device__ __forceinline__ int pool_of(const MPoolDev *__restrict__ pools, int npools, int64_t x, int s)
{
    int lo = 0, hi = npools - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t first = s < 0 ? (int64_t)pools[mid].ent0 : s < 3 ? pools[mid].s0[s] : pools[mid].q0[s - 3];
        if (first <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ StageArgs multi_stage_args(const MPoolDev &P, const MStage &st, const double *dtp)
{
    StageArgs sa{};
    sa.u0 = P.arr[0];
    sa.acc = P.arr[1];
    sa.tout = P.arr[st.out_sel];
    sa.sout = P.arr[st.sout_sel];
    sa.dtb = dtp;  // the entry's dt at index 0
    sa.a = st.a;
    sa.b = st.b;
    sa.mode = st.mode;
    sa.use_a = st.use_a;
    sa.use_b = st.use_b;
    sa.full_rows = st.full_rows;
    return sa;
}

// the entry records of a launch: one thread per entry
__global__ __launch_bounds__(256) void k_entries(const MPoolDev *__restrict__ pools, int npools,
                                                 const int64_t *__restrict__ tot, MEnt *__restrict__ ents)
{
    const int n = (int)tot[0];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int p = pool_of(pools, npools, k, -1);
        const MPoolDev &P = pools[p];
        const int j = k - P.ent0;
        ents[k] = MEnt{p, P.act[j], P.goff + P.grp[j], j};
    }
}

// entry k's record with every field wave-uniform, so that the descriptor and
// the group's constants are read through scalar loads
__device__ __forceinline__ MEnt entry_of(const MEnt *__restrict__ ents, int k)
{
    const MEnt e = ents[k];
    return MEnt{__builtin_amdgcn_readfirstlane(e.pool), __builtin_amdgcn_readfirstlane(e.b),
                __builtin_amdgcn_readfirstlane(e.g), __builtin_amdgcn_readfirstlane(e.j)};
}

// PRE: the first point's mask and operands loaded before the tile barrier;
// WPE: minimum waves per SIMD the register allocation must allow.
template <int S, bool SKEW, bool PRE, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_rhs2d_multi(
    const MPoolDev *__restrict__ pools, const MEnt *__restrict__ ents, const double *__restrict__ dtg,
    const AcmDev *__restrict__ gacm, const int64_t *__restrict__ tot, MStage st)
{
    extern __shared__ double tile[];
    const int nent = (int)tot[0];
    for (int k = blockIdx.x; k < nent; k += gridDim.x) {
        const MEnt e = entry_of(ents, k);
        const MPoolDev &P = pools[e.pool];
        const int j = e.j, b = e.b, g = e.g;
        const Geo &G = P.G;
        const AcmDev &a = gacm[g];
        const StageArgs sa = multi_stage_args(P, st, dtg + g);
        const double *mk = P.mask ? P.mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) : nullptr;
        rhs2d_block<S, SKEW, true, PRE>(tile, P.arr[st.in_sel], nullptr, mk, b, 0, P.x0dx + 6 * (size_t)j, G, a,
                                        sa);
        __syncthreads();  // the tile is reused by the next entry
    }
}

// 2D, several pools, software-pipelined: while block k is computed from
// the LDS tile, the workgroup already holds the next entry's metadata and
// has its halo tile in flight into registers; after the compute the
// registers go to LDS.  The tile is loaded by rows: lane (col, row0) =
// (t % 32, t / 32) takes column col of rows row0, row0 + 8, ... of every
// component, so a wavefront reads two contiguous row segments and no index
// is divided.  Needs tile widths <= 32 and heights <= 8 * MP_ROWS (Bs up to
// 28 with FD4); per block the arithmetic is that of k_rhs2d_multi.
constexpr int MP_ROWS = 4;
template <int S, bool SKEW>
__global__ __launch_bounds__(256) void k_rhs2d_multi_pipe(const MPoolDev *__restrict__ pools,
                                                          const MEnt *__restrict__ ents,
                                                          const double *__restrict__ dtg,
                                                          const AcmDev *__restrict__ gacm,
                                                          const int64_t *__restrict__ tot, MStage st)
{
    constexpr int R = Reach<S>::R;
    extern __shared__ double tile[];
    const int nent = (int)tot[0];
    int k = blockIdx.x;
    if (k >= nent) return;  // uniform over the workgroup
    const int col = threadIdx.x & 31, row0 = threadIdx.x >> 5;
    double reg[3][MP_ROWS];
    int cp, cj, cb, cg;  // the current entry: pool, active position, hvy id, group
    auto fetch = [&](int kk, int &p, int &j, int &b, int &g) {
        const MEnt e = entry_of(ents, kk);
        p = e.pool;
        j = e.j;
        b = e.b;
        g = e.g;
        const MPoolDev &P = pools[p];
        const Geo &G = P.G;
        const int tx = G.bs0 + 2 * R, ty = G.bs1 + 2 * R;
        const double *u = P.arr[st.in_sel] + (size_t)b * G.blk + (size_t)(G.g - R) * G.nx + (G.g - R);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < MP_ROWS; ++r) {
                const int row = row0 + 8 * r;
                if (col < tx && row < ty) reg[c][r] = u[c * G.cs + (size_t)row * G.nx + col];
            }
    };
    fetch(k, cp, cj, cb, cg);
    while (true) {
        __syncthreads();  // the previous block is done with the tile
        {
            const Geo &G = pools[cp].G;
            const int tx = G.bs0 + 2 * R, ty = G.bs1 + 2 * R, tcs = tx * ty;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < MP_ROWS; ++r) {
                    const int row = row0 + 8 * r;
                    if (col < tx && row < ty) tile[c * tcs + row * tx + col] = reg[c][r];
                }
        }
        __syncthreads();
        const int kn = k + (int)gridDim.x;
        int np = 0, nj = 0, nb = 0, ng = 0;
        if (kn < nent) fetch(kn, np, nj, nb, ng);  // in flight while block k is computed
        const MPoolDev &P = pools[cp];
        const Geo G = P.G;
        const AcmDev a = gacm[cg];
        const StageArgs sa = multi_stage_args(P, st, dtg + cg);
        const double *mk = P.mask ? P.mask + (G.mask_per_block ? (size_t)cb * G.mblk : 0) : nullptr;
        rhs2d_block<S, SKEW, true, false, true>(tile, nullptr, nullptr, mk, cb, 0, P.x0dx + 6 * (size_t)cj, G, a, sa);
        if (kn >= nent) break;  // uniform
        k = kn;
        cp = np;
        cj = nj;
        cb = nb;
        cg = ng;
    }
}

// 2D, persistent and software-pipelined: workgroup w handles the blocks at
// active positions w, w + G, w + 2G, ... (G workgroups); the halo tile of its
// next block is loaded into registers while the current block is computed
// from LDS, then stored to the other of two LDS buffers.  Per point the
// arithmetic is that of k_rhs2d.
constexpr int PIPE_MAXL = 6;  // tile elements per thread held in registers
template <int S, bool SKEW, bool FUSED>
__global__ __launch_bounds__(256) void k_rhs2d_pipe(const double *__restrict__ in, double *__restrict__ out,
                                                    const double *__restrict__ mask,
                                                    const int32_t *__restrict__ act,
                                                    const double *__restrict__ x0dx, Geo G, AcmDev a,
                                                    StageArgs sa, int nact)
{
    constexpr int R = Reach<S>::R;
    extern __shared__ double sm[];
    const int tx = G.bs0 + 2 * R, ty = G.bs1 + 2 * R, tcs = tx * ty, ntile = 3 * tcs;
    const int npts = G.bs0 * G.bs1;
    double reg[PIPE_MAXL];
    int off[PIPE_MAXL];  // block-independent source offsets of this thread's tile elements
#pragma unroll
    for (int j = 0; j < PIPE_MAXL; ++j) {
        const int i = threadIdx.x + j * blockDim.x;
        const int c = i / tcs, rr = i - c * tcs;
        const int iy = rr / tx, ix = rr - iy * tx;
        off[j] = i < ntile ? (int)(c * G.cs + (size_t)(G.g - R + iy) * G.nx + (G.g - R + ix)) : 0;
    }
    int k = blockIdx.x;
    if (k >= nact) return;  // uniform over the workgroup
    auto load = [&](int kk) {
        const double *u = in + (size_t)act[kk] * G.blk;
#pragma unroll
        for (int j = 0; j < PIPE_MAXL; ++j)
            if (threadIdx.x + j * blockDim.x < (unsigned)ntile) reg[j] = u[off[j]];
    };
    auto store = [&](double *dst) {
#pragma unroll
        for (int j = 0; j < PIPE_MAXL; ++j) {
            const int i = threadIdx.x + j * blockDim.x;
            if (i < ntile) dst[i] = reg[j];
        }
    };
    load(k);
    store(sm);
    __syncthreads();
    int cur = 0;
    while (true) {
        const int kn = k + (int)gridDim.x;
        if (kn < nact) load(kn);  // in flight while block k is computed
        const double *tile = sm + cur * ntile;
        const int b = act[k];
        const double dx0 = x0dx[6 * k + 3], dx1 = x0dx[6 * k + 4];
        const double inv[2] = {1.0 / dx0, 1.0 / dx1};
        const double inv2[2] = {1.0 / (dx0 * dx0), 1.0 / (dx1 * dx1)};
        const double *mk = mask ? mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) : nullptr;
        for (int pnt = threadIdx.x; pnt < npts; pnt += blockDim.x) {
            const int iy = pnt / G.bs0, ix = pnt - iy * G.bs0;
            const double *uc = tile + (iy + R) * tx + (ix + R);
            const size_t gi = (size_t)(G.g + iy) * G.nx + (G.g + ix);
            double r[3];
            const StridedAcc<2> A{uc, tcs, {1, tx, 0}};
            acm_point<2, S, SKEW>(A, inv, inv2, a, mk ? mk + gi : nullptr, G.cs, r);
            const size_t o = (size_t)b * G.blk + gi;
            if constexpr (FUSED) {
                const StageBlk SB = stage_blk<3>(sa, b, G.blk, npts);
                stage_epilogue<3>(sa, SB, k, (int)gi, (int)G.cs, pnt, npts, r, uc, tcs);
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) out[o + c * G.cs] = r[c];
            }
        }
        if constexpr (FUSED) stage_ghost_cols_2d<3>(sa, stage_blk<3>(sa, b, G.blk, npts), G);
        if (kn >= nact) break;  // uniform
        store(sm + (cur ^ 1) * ntile);
        __syncthreads();
        cur ^= 1;
        k = kn;
    }
}

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin:
// workgroup w runs on XCD w % 8.  The swizzle gives XCD x the consecutive
// logical indices [start(x), start(x) + count(x)), so the workgroups of one
// block share an L2 instead of being spread over all eight.  A permutation
// of [0, nw) for any nw.
__device__ __forceinline__ int xcd_swizzle(int w, int nw)
{
    constexpr int NX = 8;
    const int q = nw / NX, r = nw - q * NX, x = w % NX;
    return x * q + min(x, r) + w / NX;
}

// 3D: 256 interior points per workgroup, stencil reads served by L1/L2;
// XCD: workgroup indices swizzled so that a block's workgroups share an L2;
// WPE: minimum waves per SIMD of the register allocation (4 spills a few
// registers to scratch).
template <int S, bool SKEW, bool FUSED, bool XCD, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_rhs3d(const double *__restrict__ in, double *__restrict__ out,
                                               const double *__restrict__ mask,
                                               const int32_t *__restrict__ act,
                                               const double *__restrict__ x0dx, Geo G, AcmDev a,
                                               StageArgs sa)
{
    const int npts = G.bs0 * G.bs1 * G.bs2;
    const int nchunk = (npts + 255) / 256;
    const int wg = XCD ? xcd_swizzle((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    const int k = wg / nchunk;
    const int pnt = (wg - k * nchunk) * 256 + threadIdx.x;
    if (pnt >= npts) return;
    const int b = act[k];
    const int ix = pnt % G.bs0, iy = (pnt / G.bs0) % G.bs1, iz = pnt / (G.bs0 * G.bs1);
    const double dx0 = x0dx[6 * k + 3], dx1 = x0dx[6 * k + 4], dx2 = x0dx[6 * k + 5];
    const double inv[3] = {1.0 / dx0, 1.0 / dx1, 1.0 / dx2};
    const double inv2[3] = {1.0 / (dx0 * dx0), 1.0 / (dx1 * dx1), 1.0 / (dx2 * dx2)};
    const size_t gi = ((size_t)(G.g + iz) * G.ny + (G.g + iy)) * G.nx + (G.g + ix);
    const double *uc = in + (size_t)b * G.blk + gi;
    const double *mk = mask ? mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) + gi : nullptr;
    double r[4];
    const StridedAcc<3> A{uc, (ptrdiff_t)G.cs, {1, G.nx, (ptrdiff_t)G.nx * G.ny}};
    acm_point<3, S, SKEW>(A, inv, inv2, a, mk, G.cs, r);
    const size_t o = (size_t)b * G.blk + gi;
    if constexpr (FUSED) {
        const StageBlk SB = stage_blk<4>(sa, b, G.blk, npts);
        stage_epilogue<4>(sa, SB, k, (int)gi, (int)G.cs, pnt, npts, r, uc, G.cs);
        stage_ghost_cols<4>(sa, SB, (int)gi - (G.g + ix), (int)G.cs, ix, G.bs0, G.g);
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) out[o + c * G.cs] = r[c];
    }
}

// 3D, several pools in one launch (same arithmetic per point as k_rhs3d):
// work item t = (entry, chunk of 256 points), nchunk chunks per entry (the
// largest block of the launch; points past a smaller block's end idle);
// persistent over the items, their count nent * nchunk from device memory.
// With XCD the item index is swizzled so that a block's chunks run on one
// XCD (the grid is a multiple of 8, so item t runs on XCD t % 8).
template <int S, bool SKEW, bool XCD>
__global__ __launch_bounds__(256) void k_rhs3d_multi(const MPoolDev *__restrict__ pools,
                                                     const MEnt *__restrict__ ents,
                                                     const double *__restrict__ dtg,
                                                     const AcmDev *__restrict__ gacm,
                                                     const int64_t *__restrict__ tot, MStage st, int nchunk)
{
    const int total = (int)tot[0] * nchunk;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int w = XCD ? xcd_swizzle(t, total) : t;
        const int k = w / nchunk;
        const int pnt = (w - k * nchunk) * 256 + threadIdx.x;
        const MEnt e = entry_of(ents, k);
        const MPoolDev &P = pools[e.pool];
        const Geo &G = P.G;
        const int npts = G.bs0 * G.bs1 * G.bs2;
        if (pnt >= npts) continue;
        const int j = e.j, b = e.b, g = e.g;
        const AcmDev &a = gacm[g];
        const StageArgs sa = multi_stage_args(P, st, dtg + g);
        const double *xd = P.x0dx + 6 * (size_t)j;
        const int ix = pnt % G.bs0, iy = (pnt / G.bs0) % G.bs1, iz = pnt / (G.bs0 * G.bs1);
        const double dx0 = xd[3], dx1 = xd[4], dx2 = xd[5];
        const double inv[3] = {1.0 / dx0, 1.0 / dx1, 1.0 / dx2};
        const double inv2[3] = {1.0 / (dx0 * dx0), 1.0 / (dx1 * dx1), 1.0 / (dx2 * dx2)};
        const size_t gi = ((size_t)(G.g + iz) * G.ny + (G.g + iy)) * G.nx + (G.g + ix);
        const double *uc = P.arr[st.in_sel] + (size_t)b * G.blk + gi;
        const double *mk = P.mask ? P.mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) + gi : nullptr;
        double r[4];
        const StridedAcc<3> A{uc, (ptrdiff_t)G.cs, {1, G.nx, (ptrdiff_t)G.nx * G.ny}};
        acm_point<3, S, SKEW>(A, inv, inv2, a, mk, G.cs, r);
        const StageBlk SB = stage_blk<4>(sa, b, G.blk, npts);
        stage_epilogue<4>(sa, SB, 0, (int)gi, (int)G.cs, pnt, npts, r, uc, G.cs);
        stage_ghost_cols<4>(sa, SB, (int)gi - (G.g + ix), (int)G.cs, ix, G.bs0, G.g);
    }
}

// 3D, z-marching: a workgroup owns a slab of ty0 rows of (x, y) columns of
// one block (ysplit workgroups per block, one thread per column) and marches
// through z.  The x/y stencils read an LDS copy of the current z-plane slab
// (interior from the threads' registers, the R-wide x/y ring from global
// memory); the z stencil reads a per-thread register queue of 2R+1 planes.
// The next plane's values are fetched before the current one is computed.
// Every input value is read from memory once (plus the slab ring), except
// that with zsplit > 1 the z range of a block is cut into zsplit pieces,
// each with its own workgroups, which read the 2R planes around their cut.
// WPE: the minimum waves per SIMD the register allocation must allow (1: no
// constraint; 3 may spill a few registers to scratch).
template <int S, bool SKEW, bool FUSED, int WPE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(WPE))) void k_rhs3d_zm(const double *__restrict__ in,
                                                  double *__restrict__ out,
                                                  const double *__restrict__ mask,
                                                  const int32_t *__restrict__ act,
                                                  const double *__restrict__ x0dx, Geo G, AcmDev a,
                                                  StageArgs sa, int ysplit, int zsplit)
{
    constexpr int R = Reach<S>::R;
    constexpr int W = 2 * R + 1;
    constexpr int MAXRING = 4;  // ring elements per thread held in registers
    extern __shared__ double plane[];
    const int per_block = ysplit * zsplit;
    const int k = blockIdx.x / per_block;
    const int rem = blockIdx.x - k * per_block;
    const int zpart = rem / ysplit, part = rem - zpart * ysplit;
    const int tz0 = G.bs2 / zsplit, z0 = zpart * tz0, z1 = z0 + tz0;
    const int b = act[k];
    const int ty0 = G.bs1 / ysplit;
    const int y0 = part * ty0;
    const int tid = threadIdx.x;
    const int ix = tid % G.bs0, iyl = tid / G.bs0, iy = y0 + iyl;
    const int tx = G.bs0 + 2 * R, ty = ty0 + 2 * R, tcs = tx * ty;
    const size_t zs = (size_t)G.nx * G.ny;
    const double *u = in + (size_t)b * G.blk;
    const size_t col = (size_t)(G.g + iy) * G.nx + (G.g + ix);
    const int nthr = G.bs0 * ty0;
    const int ring = 2 * R * (G.bs0 + ty0);  // x/y ring of the slab, without corners
    const double dx0 = x0dx[6 * k + 3], dx1 = x0dx[6 * k + 4], dx2 = x0dx[6 * k + 5];
    const double inv[3] = {1.0 / dx0, 1.0 / dx1, 1.0 / dx2};
    const double inv2[3] = {1.0 / (dx0 * dx0), 1.0 / (dx1 * dx1), 1.0 / (dx2 * dx2)};
    const double *mk = mask ? mask + (G.mask_per_block ? (size_t)b * G.mblk : 0) : nullptr;

    // ring element h -> (component, offset in LDS, offset in the block plane)
    int r_lds[MAXRING];
    size_t r_glb[MAXRING];
    int nring = 0;
    for (int h = tid; h < 4 * ring && nring < MAXRING; h += nthr, ++nring) {
        const int c = h / ring;
        const int e = h - c * ring;
        int px, py;
        if (e < 2 * R * G.bs0) {
            const int row = e / G.bs0;
            px = R + (e - row * G.bs0);
            py = row < R ? row : ty0 + row;
        } else {
            const int e2 = e - 2 * R * G.bs0;
            const int kk = e2 % (2 * R);
            py = R + e2 / (2 * R);
            px = kk < R ? kk : G.bs0 + kk;
        }
        r_lds[nring] = c * tcs + py * tx + px;
        r_glb[nring] = c * G.cs + (size_t)(G.g + y0 - R + py) * G.nx + (G.g - R + px);
    }

    double q[4][W];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) q[c][j] = u[c * G.cs + (size_t)(G.g + z0 - R + j) * zs + col];
    double nq[4], nr[MAXRING];
#pragma unroll
    for (int c = 0; c < 4; ++c) nq[c] = u[c * G.cs + (size_t)(G.g + z0 + R) * zs + col];
#pragma unroll
    for (int j = 0; j < MAXRING; ++j)
        if (j < nring) nr[j] = u[(size_t)(G.g + z0) * zs + r_glb[j]];

    const int lc = (iyl + R) * tx + (ix + R);
    for (int iz = z0; iz < z1; ++iz) {
        const size_t zoff = (size_t)(G.g + iz) * zs;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            q[c][2 * R] = nq[c];
            plane[c * tcs + lc] = q[c][R];
        }
#pragma unroll
        for (int j = 0; j < MAXRING; ++j)
            if (j < nring) plane[r_lds[j]] = nr[j];
        if (iz + 1 < z1) {  // prefetch the next plane
            const size_t zn = zoff + zs;
#pragma unroll
            for (int c = 0; c < 4; ++c) nq[c] = u[c * G.cs + zn + R * zs + col];
#pragma unroll
            for (int j = 0; j < MAXRING; ++j)
                if (j < nring) nr[j] = u[zn + r_glb[j]];
        }
        __syncthreads();
        // x/y from the LDS plane, z from the register queue
        const auto A = [&](int c, int d, int o) -> double {
            if (d == 2) return q[c][R + o];
            return plane[c * tcs + lc + o * (d == 0 ? 1 : tx)];
        };
        const size_t gi = zoff + col;
        double r[4];
        acm_point<3, S, SKEW>(A, inv, inv2, a, mk ? mk + gi : nullptr, G.cs, r);
        const size_t o = (size_t)b * G.blk + gi;
        if constexpr (FUSED) {
            const double u0c[4] = {q[0][R], q[1][R], q[2][R], q[3][R]};
            const int npts = G.bs0 * G.bs1 * G.bs2;
            const StageBlk SB = stage_blk<4>(sa, b, G.blk, npts);
            stage_epilogue<4>(sa, SB, k, (int)gi, (int)G.cs, (iz * G.bs1 + iy) * G.bs0 + ix, npts, r, u0c, 1);
            stage_ghost_cols<4>(sa, SB, (int)gi - (G.g + ix), (int)G.cs, ix, G.bs0, G.g);
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) out[o + c * G.cs] = r[c];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int j = 0; j < 2 * R; ++j) q[c][j] = q[c][j + 1];
    }
}

int reach_of(int scheme)
{
    return scheme == WH_FD2 ? 1 : scheme == WH_FD4 ? 2 : 3;
}

// Is the z-marching kernel usable with this z-split?  It divides Bs2, and
// each piece has at least 2R planes.
bool zm_zvalid(int bs2, int scheme, int zs)
{
    return zs >= 1 && bs2 % zs == 0 && bs2 / zs >= 2 * reach_of(scheme);
}

// Is the z-marching kernel usable with this y-split?  Threads per workgroup
// in [64, 512], the slab ring fits the per-thread ring registers, LDS <= 64 KB.
bool zm_valid(int bs0, int bs1, int scheme, int ys)
{
    if (ys < 1 || bs1 % ys) return false;
    const int R = reach_of(scheme), ty0 = bs1 / ys, nthr = bs0 * ty0;
    if (nthr < 64 || nthr > 512) return false;
    if (2 * R * (bs0 + ty0) > nthr) return false;
    return sizeof(double) * 4 * (bs0 + 2 * R) * (ty0 + 2 * R) <= 64 * 1024;
}

// y-split of the 3D stage kernel: the pool's tuned value, or the smallest
// valid split; 0 selects the flat kernel.
static int zm_split(const wh_pool *P, const Geo &G, int scheme)
{
    if (P->k3d_ysplit == 0) return 0;
    if (P->k3d_ysplit > 0 && zm_valid(G.bs0, G.bs1, scheme, P->k3d_ysplit)) return P->k3d_ysplit;
    for (int ys = 1; ys <= 8; ++ys)
        if (zm_valid(G.bs0, G.bs1, scheme, ys)) return ys;
    return 0;
}

Geo make_geo(const wh_pool *P)
{
    Geo G;
    G.bs0 = P->d.bs[0];
    G.bs1 = P->d.bs[1];
    G.bs2 = P->d.dim == 3 ? P->d.bs[2] : 1;
    G.g = P->d.g;
    G.nx = P->nx;
    G.ny = P->ny;
    G.blk = P->blk;
    G.cs = P->cs;
    G.mblk = P->mblk;
    G.mask_per_block = P->d.mask_per_block;
    return G;
}

template <bool FUSED>
static int dispatch_rhs(wh_pool *P, const AcmDev &a, const double *in, double *out,
                        const int32_t *act, const double *x0dx, int nact, const StageArgs &sa)
{
    if (nact <= 0) return WH_OK;
    const Geo G = make_geo(P);
    const double *mask = (P->mask && (a.penal || a.sponge)) ? P->mask : nullptr;
    if (P->d.dim == 2) {
        int R = a.scheme == WH_FD2 ? 1 : a.scheme == WH_FD4 ? 2 : 3;
        size_t lds = sizeof(double) * 3 * (G.bs0 + 2 * R) * (G.bs1 + 2 * R);
        WH_CHECK(lds <= 160 * 1024, WH_E_ARG, "2D block tile exceeds LDS");
        const int pre = P->k2d_stage == 1;
        // the persistent pipelined variant: two tiles in LDS, each thread
        // holds at most PIPE_MAXL tile elements of the next block
        // (the halo gather reads through the one-workgroup-per-block kernel)
        const int wpc = (FUSED && sa.gnb) ? 0 : P->k2d_stage == 2 ? 2 : P->k2d_stage == 3 ? 4 : P->k2d_stage == 4 ? 8 : 0;
        if (wpc && 3 * (G.bs0 + 2 * R) * (G.bs1 + 2 * R) <= PIPE_MAXL * 256 && 2 * lds <= 160 * 1024) {
            const int nwg = std::min(nact, wpc * P->n_cu);
            dim3 pgrid(nwg), pblock(256);
#define L2P(SC, SK)                                                                                 \
    hipLaunchKernelGGL((k_rhs2d_pipe<SC, SK, FUSED>), pgrid, pblock, 2 * lds, P->stream, in, out, mask, \
                       act, x0dx, G, a, sa, nact)
            switch (a.scheme * 2 + a.skew) {
            case WH_FD2 * 2: L2P(WH_FD2, false); break;
            case WH_FD2 * 2 + 1: L2P(WH_FD2, true); break;
            case WH_FD4 * 2: L2P(WH_FD4, false); break;
            case WH_FD4 * 2 + 1: L2P(WH_FD4, true); break;
            case WH_FD4_TW * 2: L2P(WH_FD4_TW, false); break;
            case WH_FD4_TW * 2 + 1: L2P(WH_FD4_TW, true); break;
            case WH_FD6 * 2: L2P(WH_FD6, false); break;
            case WH_FD6 * 2 + 1: L2P(WH_FD6, true); break;
            default: set_error("unknown scheme"); return WH_E_ARG;
            }
#undef L2P
            WH_HIP(hipGetLastError());
            return WH_OK;
        }
        dim3 grid(nact), blockd(256);
#define L2D(SC, SK)                                                                              \
    do {                                                                                         \
        if (pre)                                                                                 \
            hipLaunchKernelGGL((k_rhs2d<SC, SK, FUSED, true>), grid, blockd, lds, P->stream, in, out, \
                               mask, act, x0dx, G, a, sa);                                       \
        else                                                                                     \
            hipLaunchKernelGGL((k_rhs2d<SC, SK, FUSED, false>), grid, blockd, lds, P->stream, in,     \
                               out, mask, act, x0dx, G, a, sa);                                  \
    } while (0)
        switch (a.scheme * 2 + a.skew) {
        case WH_FD2 * 2: L2D(WH_FD2, false); break;
        case WH_FD2 * 2 + 1: L2D(WH_FD2, true); break;
        case WH_FD4 * 2: L2D(WH_FD4, false); break;
        case WH_FD4 * 2 + 1: L2D(WH_FD4, true); break;
        case WH_FD4_TW * 2: L2D(WH_FD4_TW, false); break;
        case WH_FD4_TW * 2 + 1: L2D(WH_FD4_TW, true); break;
        case WH_FD6 * 2: L2D(WH_FD6, false); break;
        case WH_FD6 * 2 + 1: L2D(WH_FD6, true); break;
        default: set_error("unknown scheme"); return WH_E_ARG;
        }
#undef L2D
    } else if (const int ys = zm_split(P, G, a.scheme); ys > 0) {
        const int R = reach_of(a.scheme);
        const int zs = zm_zvalid(G.bs2, a.scheme, P->k3d_zsplit) ? P->k3d_zsplit : 1;
        const size_t lds = sizeof(double) * 4 * (G.bs0 + 2 * R) * (G.bs1 / ys + 2 * R);
        dim3 grid(nact * ys * zs), blockd(G.bs0 * (G.bs1 / ys));
#define L3D(SC, SK)                                                                                    \
    do {                                                                                                \
        if (P->k3d_waves == 3)                                                                          \
            hipLaunchKernelGGL((k_rhs3d_zm<SC, SK, FUSED, 3>), grid, blockd, lds, P->stream, in, out, mask, \
                               act, x0dx, G, a, sa, ys, zs);                                            \
        else                                                                                            \
            hipLaunchKernelGGL((k_rhs3d_zm<SC, SK, FUSED, 1>), grid, blockd, lds, P->stream, in, out, mask, \
                               act, x0dx, G, a, sa, ys, zs);                                            \
    } while (0)
        switch (a.scheme * 2 + a.skew) {
        case WH_FD2 * 2: L3D(WH_FD2, false); break;
        case WH_FD2 * 2 + 1: L3D(WH_FD2, true); break;
        case WH_FD4 * 2: L3D(WH_FD4, false); break;
        case WH_FD4 * 2 + 1: L3D(WH_FD4, true); break;
        case WH_FD4_TW * 2: L3D(WH_FD4_TW, false); break;
        case WH_FD4_TW * 2 + 1: L3D(WH_FD4_TW, true); break;
        case WH_FD6 * 2: L3D(WH_FD6, false); break;
        case WH_FD6 * 2 + 1: L3D(WH_FD6, true); break;
        default: set_error("unknown scheme"); return WH_E_ARG;
        }
#undef L3D
    } else {
        const int npts = G.bs0 * G.bs1 * G.bs2;
        dim3 grid(nact * ((npts + 255) / 256)), blockd(256);
#define L3F(SC, SK, X, W)                                                                          \
    hipLaunchKernelGGL((k_rhs3d<SC, SK, FUSED, X, W>), grid, blockd, 0, P->stream, in, out, mask, act, \
                       x0dx, G, a, sa)
#define L3D(SC, SK)                                            \
    do {                                                        \
        const bool w4 = P->k3d_waves == 4;                      \
        if (P->k3d_xcd && w4) L3F(SC, SK, true, 4);             \
        else if (P->k3d_xcd) L3F(SC, SK, true, 1);              \
        else if (w4) L3F(SC, SK, false, 4);                     \
        else L3F(SC, SK, false, 1);                             \
    } while (0)
        switch (a.scheme * 2 + a.skew) {
        case WH_FD2 * 2: L3D(WH_FD2, false); break;
        case WH_FD2 * 2 + 1: L3D(WH_FD2, true); break;
        case WH_FD4 * 2: L3D(WH_FD4, false); break;
        case WH_FD4 * 2 + 1: L3D(WH_FD4, true); break;
        case WH_FD4_TW * 2: L3D(WH_FD4_TW, false); break;
        case WH_FD4_TW * 2 + 1: L3D(WH_FD4_TW, true); break;
        case WH_FD6 * 2: L3D(WH_FD6, false); break;
        case WH_FD6 * 2 + 1: L3D(WH_FD6, true); break;
        default: set_error("unknown scheme"); return WH_E_ARG;
        }
#undef L3F
#undef L3D
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_rhs(wh_pool *P, const AcmDev &a, const double *in, double *out, const int32_t *act,
               const double *x0dx, int nact)
{
    StageArgs sa{};
    return dispatch_rhs<false>(P, a, in, out, act, x0dx, nact, sa);
}

int launch_fused_stage(wh_pool *P, const AcmDev &a, const double *tin, double *tout,
                       const double *u0, double *acc, double *sout, int mode, double acoef,
                       double bcoef, const double *dtb, const int32_t *gnb)
{
    StageArgs sa{};
    sa.gnb = gnb;
    sa.dtb = dtb;
    sa.u0 = u0;
    sa.acc = acc;
    sa.tout = tout;
    sa.sout = sout;
    sa.scal = P->scal;
    sa.a = acoef;
    sa.b = bcoef;
    sa.mode = mode;
    sa.use_a = std::abs(acoef) >= 1.0e-8;
    sa.use_b = std::abs(bcoef) >= 1.0e-8;
    sa.full_rows = stage_full_rows(P->nact, sizeof(double) * (double)make_geo(P).blk);
    return dispatch_rhs<true>(P, a, tin, nullptr, P->d_active, P->d_x0dx, P->nact, sa);
}


// Interpolating prediction (module_wavelets.f90:96-284, `prediction`) at one
// point of the refined grid.  f[d] is the point's index in the doubled
// ghost-inclusive index space of the coarse block u (coarse point c sits at
// fine index 2c).  Even indices copy, odd ones interpolate with the centred
// stencil of `order` (2: (1,1)/2, 4: (-1,9,9,-1)/16, 6: (3,-25,150,150,-25,3)/256)
// separably, x innermost, then y, then z, each sum left to right -- the
// reference's separable passes evaluated at one point.
template <int N>
__device__ __forceinline__ double pred_point(const double *__restrict__ u, int nx, int ny,
                                             int dim, const int f[3])
{
    constexpr double c2[2] = {0.5, 0.5};
    constexpr double c4[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
    constexpr double c6[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0,
                              150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};
    const double *cf = (N == 1) ? c2 : (N == 2) ? c4 : c6;
    const int cx = f[0] >> 1, px = f[0] & 1;
    const int cy = f[1] >> 1, py = f[1] & 1;
    const int cz = dim == 3 ? f[2] >> 1 : 0, pz = dim == 3 ? (f[2] & 1) : 0;
    auto vx = [&](int jy, int kz) -> double {
        const double *row = u + ((size_t)kz * ny + jy) * nx;
        if (!px) return row[cx];
        double s = cf[0] * row[cx - N + 1];
#pragma unroll
        for (int t = 1; t < 2 * N; ++t) s = s + cf[t] * row[cx - N + 1 + t];
        return s;
    };
    auto vy = [&](int kz) -> double {
        if (!py) return vx(cy, kz);
        double s = cf[0] * vx(cy - N + 1, kz);
#pragma unroll
        for (int t = 1; t < 2 * N; ++t) s = s + cf[t] * vx(cy - N + 1 + t, kz);
        return s;
    };
    if (!pz) return vy(cz);
    double s = cf[0] * vy(cz - N + 1);
#pragma unroll
    for (int t = 1; t < 2 * N; ++t) s = s + cf[t] * vy(cz - N + 1 + t);
    return s;
}

__device__ __forceinline__ double pred_any(const double *u, int nx, int ny, int dim, const int f[3],
                                           int order)
{
    if (order == 4) return pred_point<2>(u, nx, ny, dim, f);
    if (order == 2) return pred_point<1>(u, nx, ny, dim, f);
    return pred_point<3>(u, nx, ny, dim, f);
}

// The points [r0, r_end) of one ghost patch (x fastest), taken by the 64
// lanes of a wavefront in steps of 64 (lane l starts at r0 + l): per point
// the copy of a same-level sender, the decimation of a finer one (every
// other point) or the prediction of a coarser one (pred_any), every
// component, into the receiver's ghost nodes.  P is the same in every lane
// (read through the scalar unit); sbase: block P.send of the sender array.
__device__ __forceinline__ void sync_patch(double *__restrict__ arr, const wh_patch &P,
                                           const double *__restrict__ sbase, size_t blk, size_t cs, int nx, int ny,
                                           int neq, int dim, int r0, int r_end = 1 << 30)
{
    const int n0 = P.n[0], n1 = P.n[1], np = min(n0 * n1 * P.n[2], r_end);
    double *dbase = arr + (size_t)P.recv * blk;
    const int op = P.op & 0xff;
    const int st = (op == WH_PATCH_RESTRICT) ? 2 : 1;
    for (int r = r0; r < np; r += 64) {
        const int i = r % n0, j = (r / n0) % n1, kk = r / (n0 * n1);
        double *dst = dbase + ((size_t)(P.recv_lo[2] + kk) * ny + (P.recv_lo[1] + j)) * nx + (P.recv_lo[0] + i);
        if (op == WH_PATCH_PREDICT) {
            const int f[3] = {P.send_lo[0] + i, P.send_lo[1] + j, P.send_lo[2] + kk};
            for (int c = 0; c < neq; ++c) dst[c * cs] = pred_any(sbase + c * cs, nx, ny, dim, f, P.aux);
        } else {
            const double *src = sbase + ((size_t)(P.send_lo[2] + st * kk) * ny + (P.send_lo[1] + st * j)) * nx +
                                (P.send_lo[0] + st * i);
            for (int c = 0; c < neq; ++c) dst[c * cs] = src[c * cs];
        }
    }
}

// 2D ghost fill of the patches [p0, p1): one wavefront per patch (patch
// p0 + w for wavefront w of the launch), its record read once through the
// scalar unit -- no search; a 2D ghost patch has a few dozen points, so the
// wavefront takes them in one or two steps (3D patches are far larger and
// are cut into chunks, k_sync_chunks).  Bits 8-9 of a patch's op select the
// array its sender is read from (0: the synced array itself, 1/2: the
// alternates in S).
__global__ __launch_bounds__(256) void k_sync_patches(double *__restrict__ arr, const wh_patch *__restrict__ pt,
                                                      int p0, int p1, size_t blk, size_t cs, int nx, int ny,
                                                      int neq, SyncSrc S)
{
    const int q = p0 + (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (q >= p1) return;
    const wh_patch P = pt[q];
    const int sel = (P.op >> 8) & 3;
    const double *sbase = sel == 0 ? arr + (size_t)P.send * blk
                                   : S.alt[sel - 1] + (size_t)P.send * S.blk_alt[sel - 1];
    sync_patch(arr, P, sbase, blk, cs, nx, ny, neq, 2, (int)(threadIdx.x & 63u));
}

// 3D ghost fill of the chunks [c0, c1): one wavefront per chunk (chunk c0 + w
// for wavefront w of the launch), chunk and patch records read once through
// the scalar unit, lane l takes points first + l, first + l + 64, ... of the
// chunk.
__global__ __launch_bounds__(256) void k_sync_chunks(double *__restrict__ arr, const wh_patch *__restrict__ pt,
                                                     const SyncChunk *__restrict__ ch, int c0, int c1, size_t blk,
                                                     size_t cs, int nx, int ny, int neq, SyncSrc S)
{
    const int q = c0 + (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (q >= c1) return;
    const SyncChunk c = ch[q];
    const wh_patch P = pt[c.patch];
    const int sel = (P.op >> 8) & 3;
    const double *sbase = sel == 0 ? arr + (size_t)P.send * blk
                                   : S.alt[sel - 1] + (size_t)P.send * S.blk_alt[sel - 1];
    sync_patch(arr, P, sbase, blk, cs, nx, ny, neq, 3, c.first + (int)(threadIdx.x & 63u), c.first + SYNC_CHUNK);
}

int launch_sync_flat(wh_pool *P, double *arr, const wh_patch *d_patches, const SyncChunk *d_chunks, int p0, int p1,
                     int c0, int c1, const SyncSrc &S)
{
    if (P->d.dim == 2) {
        if (p1 > p0)
            hipLaunchKernelGGL(k_sync_patches, dim3((unsigned)((p1 - p0 + 3) / 4)), dim3(256), 0, P->stream, arr,
                               d_patches, p0, p1, P->blk, P->cs, P->nx, P->ny, P->d.neq, S);
    } else if (c1 > c0) {
        hipLaunchKernelGGL(k_sync_chunks, dim3((unsigned)((c1 - c0 + 3) / 4)), dim3(256), 0, P->stream, arr,
                           d_patches, d_chunks, c0, c1, P->blk, P->cs, P->nx, P->ny, P->d.neq, S);
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// 3D ghost fill over the chunks of several pools (one sync stage s): the
// stage's chunks are the pools' chunk ranges [c0[s], c1[s]) one after the
// other (pool p's from s0[s] on); wavefront w of the launch fills chunk
// w - s0[s] + c0[s] of the pool whose range holds w, as k_sync_chunks.
// Persistent over the chunks, their count tot[stage] read from device
// memory; into array `sel` of the pool.
__global__ __launch_bounds__(256) void k_sync_multi(const MPoolDev *__restrict__ pools, int npools, int sel,
                                                    int stage, const int64_t *__restrict__ tot)
{
    const int s = stage - 1;
    const int64_t nc = tot[stage];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int64_t w = (int64_t)blockIdx.x * 4 + wv; w < nc; w += (int64_t)gridDim.x * 4) {
        const MPoolDev &M = pools[__builtin_amdgcn_readfirstlane(pool_of(pools, npools, w, s))];
        const SyncChunk c = M.ch[M.c0[s] + (int)(w - M.s0[s])];
        const wh_patch P = M.pt[c.patch];
        double *arr = M.arr[sel];
        const Geo &G = M.G;
        sync_patch(arr, P, arr + (size_t)P.send * G.blk, G.blk, G.cs, G.nx, G.ny, M.neq, 3,
                   c.first + (int)(threadIdx.x & 63u), c.first + SYNC_CHUNK);
    }
}

// 2D ghost fill over the patches of several pools (one sync stage s): the
// stage's patches are the pools' patch ranges [p0[s], p1[s]) one after the
// other (pool p's from q0[s] on); wavefront w of the launch fills patch
// w - q0[s] + p0[s] of the pool whose range holds w, as k_sync_patches.
// Persistent over the patches, their count tot[4 + s] read from device
// memory.
__global__ __launch_bounds__(256) void k_sync_multi_patches(const MPoolDev *__restrict__ pools, int npools,
                                                            int sel, int stage, const int64_t *__restrict__ tot)
{
    const int s = stage - 1;
    const int64_t nq = tot[4 + s];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int64_t w = (int64_t)blockIdx.x * 4 + wv; w < nq; w += (int64_t)gridDim.x * 4) {
        const MPoolDev &M = pools[__builtin_amdgcn_readfirstlane(pool_of(pools, npools, w, 3 + s))];
        const wh_patch P = M.pt[M.p0[s] + (int)(w - M.q0[s])];
        double *arr = M.arr[sel];
        const Geo &G = M.G;
        sync_patch(arr, P, arr + (size_t)P.send * G.blk, G.blk, G.cs, G.nx, G.ny, M.neq, 2,
                   (int)(threadIdx.x & 63u));
    }
}

int launch_sync_multi(hipStream_t stream, const MPoolDev *d_pools, int npools, int sel, int stage,
                      const int64_t *d_tot, int grid, int dim)
{
    if (dim == 2)
        hipLaunchKernelGGL(k_sync_multi_patches, dim3(grid), dim3(256), 0, stream, d_pools, npools, sel, stage,
                           d_tot);
    else
        hipLaunchKernelGGL(k_sync_multi, dim3(grid), dim3(256), 0, stream, d_pools, npools, sel, stage, d_tot);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_entries(hipStream_t stream, const MPoolDev *d_pools, int npools, const int64_t *d_tot, MEnt *d_ents,
                   int grid)
{
    hipLaunchKernelGGL(k_entries, dim3(grid), dim3(256), 0, stream, d_pools, npools, d_tot, d_ents);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_stage_multi(hipStream_t stream, int dim, int scheme, int skew, int xcd, const MPoolDev *d_pools,
                       int npools, const MEnt *d_ents, const double *d_dtg, const AcmDev *d_gacm,
                       const int64_t *d_tot, const MStage &ms, size_t lds, int grid, int nchunk, int pipe)
{
    (void)npools;
    if (dim == 2) {
        WH_CHECK(lds <= 160 * 1024, WH_E_ARG, "2D block tile exceeds LDS");
#define M2D(SC, SK)                                                                                              \
    do {                                                                                                         \
        if (pipe == 1)                                                                                           \
            hipLaunchKernelGGL((k_rhs2d_multi_pipe<SC, SK>), dim3(grid), dim3(256), lds, stream, d_pools, d_ents, \
                               d_dtg, d_gacm, d_tot, ms);                                                        \
        else if (pipe == 2)                                                                                      \
            hipLaunchKernelGGL((k_rhs2d_multi<SC, SK, false, 1>), dim3(grid), dim3(256), lds, stream, d_pools,    \
                               d_ents, d_dtg, d_gacm, d_tot, ms);                                                \
        else if (pipe == 3)                                                                                      \
            hipLaunchKernelGGL((k_rhs2d_multi<SC, SK, false, 6>), dim3(grid), dim3(256), lds, stream, d_pools,    \
                               d_ents, d_dtg, d_gacm, d_tot, ms);                                                \
        else if (pipe == 4)                                                                                      \
            hipLaunchKernelGGL((k_rhs2d_multi<SC, SK, false, 8>), dim3(grid), dim3(256), lds, stream, d_pools,    \
                               d_ents, d_dtg, d_gacm, d_tot, ms);                                                \
        else                                                                                                     \
            hipLaunchKernelGGL((k_rhs2d_multi<SC, SK, true, 1>), dim3(grid), dim3(256), lds, stream, d_pools,     \
                               d_ents, d_dtg, d_gacm, d_tot, ms);                                                \
    } while (0)
        switch (scheme * 2 + skew) {
        case WH_FD2 * 2: M2D(WH_FD2, false); break;
        case WH_FD2 * 2 + 1: M2D(WH_FD2, true); break;
        case WH_FD4 * 2: M2D(WH_FD4, false); break;
        case WH_FD4 * 2 + 1: M2D(WH_FD4, true); break;
        case WH_FD4_TW * 2: M2D(WH_FD4_TW, false); break;
        case WH_FD4_TW * 2 + 1: M2D(WH_FD4_TW, true); break;
        case WH_FD6 * 2: M2D(WH_FD6, false); break;
        case WH_FD6 * 2 + 1: M2D(WH_FD6, true); break;
        default: set_error("unknown scheme"); return WH_E_ARG;
        }
#undef M2D
    } else {
#define M3D(SC, SK)                                                                                           \
    do {                                                                                                      \
        if (xcd)                                                                                              \
            hipLaunchKernelGGL((k_rhs3d_multi<SC, SK, true>), dim3(grid), dim3(256), 0, stream, d_pools, d_ents, \
                               d_dtg, d_gacm, d_tot, ms, nchunk);                                              \
        else                                                                                                  \
            hipLaunchKernelGGL((k_rhs3d_multi<SC, SK, false>), dim3(grid), dim3(256), 0, stream, d_pools, d_ents, \
                               d_dtg, d_gacm, d_tot, ms, nchunk);                                              \
    } while (0)
        switch (scheme * 2 + skew) {
        case WH_FD2 * 2: M3D(WH_FD2, false); break;
        case WH_FD2 * 2 + 1: M3D(WH_FD2, true); break;
        case WH_FD4 * 2: M3D(WH_FD4, false); break;
        case WH_FD4 * 2 + 1: M3D(WH_FD4, true); break;
        case WH_FD4_TW * 2: M3D(WH_FD4_TW, false); break;
        case WH_FD4_TW * 2 + 1: M3D(WH_FD4_TW, true); break;
        case WH_FD6 * 2: M3D(WH_FD6, false); break;
        case WH_FD6 * 2 + 1: M3D(WH_FD6, true); break;
        default: set_error("unknown scheme"); return WH_E_ARG;
        }
#undef M3D
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// Per-receiver ghost fill (the plain one-launch-per-block path): one
// workgroup walks the patches of one (stage, receiver) row of the table.
__global__ __launch_bounds__(256) void k_sync(double *__restrict__ arr,
                                              const wh_patch *__restrict__ pt,
                                              const int32_t *__restrict__ off, int row0,
                                              int stage, size_t blk, size_t cs, int nx, int ny,
                                              int neq)
{
    const int row = row0 + blockIdx.x;
    const int dim = (cs == (size_t)nx * ny) ? 2 : 3;
    for (int q = off[row]; q < off[row + 1]; ++q) {
        const wh_patch P = pt[q];
        if (stage > 0 && P.stage != stage) continue;
        const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
        const int npt = n0 * n1 * n2;
        const int st = (P.op == WH_PATCH_RESTRICT) ? 2 : 1;
        double *dst = arr + (size_t)P.recv * blk;
        const double *src = arr + (size_t)P.send * blk;
        for (int e = threadIdx.x; e < npt * neq; e += blockDim.x) {
            const int c = e / npt, r = e - c * npt;
            const int i = r % n0, j = (r / n0) % n1, kk = r / (n0 * n1);
            const size_t ri = ((size_t)(P.recv_lo[2] + kk) * ny + (P.recv_lo[1] + j)) * nx +
                              (P.recv_lo[0] + i);
            if (P.op == WH_PATCH_PREDICT) {
                const int f[3] = {P.send_lo[0] + i, P.send_lo[1] + j, P.send_lo[2] + kk};
                dst[c * cs + ri] = pred_any(src + c * cs, nx, ny, dim, f, P.aux);
                continue;
            }
            const size_t si = ((size_t)(P.send_lo[2] + st * kk) * ny + (P.send_lo[1] + st * j)) * nx +
                              (P.send_lo[0] + st * i);
            dst[c * cs + ri] = src[c * cs + si];
        }
    }
}

// refineBlock (refinementExecute.f90:4-116): daughter k of a mother takes
// points g + Bs*off_k + i (i over its ghost-inclusive block) of the mother's
// interpolated grid; one thread per daughter point.
__global__ __launch_bounds__(256) void k_refine(double *__restrict__ arr,
                                                const int32_t *__restrict__ pairs, int npairs,
                                                int bs0, int bs1, int bs2, int g, int nx, int ny,
                                                int nz, size_t blk, size_t cs, int neq, int dim,
                                                int order)
{
    const size_t npt = (size_t)nx * ny * nz;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npt * npairs) return;
    const int q = (int)(e / npt);
    const int r = (int)(e - (size_t)q * npt);
    const int i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
    const int mother = pairs[3 * q], daughter = pairs[3 * q + 1], dg = pairs[3 * q + 2];
    const int off[3] = {(dg >> 1) & 1, dg & 1, (dg >> 2) & 1};
    const int f[3] = {g + bs0 * off[0] + i, g + bs1 * off[1] + j,
                      dim == 3 ? g + bs2 * off[2] + k : 0};
    const double *src = arr + (size_t)mother * blk;
    double *dst = arr + (size_t)daughter * blk + r;
    for (int c = 0; c < neq; ++c) dst[c * cs] = pred_any(src + c * cs, nx, ny, dim, f, order);
}

// 2D refineBlock with the mother's component plane staged in LDS: one
// workgroup per (pair, component), each daughter point pred_point of the
// staged plane -- k_refine's arithmetic, its stencil reads from LDS
// instead of up to (2N)^2 cached global loads per point.  The host's
// stencil-range check keeps every read inside the plane.
template <int N>
__global__ __launch_bounds__(256) void k_refine2d_lds(double *__restrict__ arr, const int32_t *__restrict__ pairs,
                                                      int bs0, int bs1, int g, int nx, int ny, size_t blk, size_t cs,
                                                      int neq)
{
    extern __shared__ double plane[];
    const int q = (int)blockIdx.x / neq, c = (int)blockIdx.x - q * neq;
    const int mother = pairs[3 * q], daughter = pairs[3 * q + 1], dg = pairs[3 * q + 2];
    const int npt = nx * ny;
    const double *src = arr + (size_t)mother * blk + (size_t)c * cs;
    for (int e = threadIdx.x; e < npt; e += blockDim.x) plane[e] = src[e];
    __syncthreads();
    double *dst = arr + (size_t)daughter * blk + (size_t)c * cs;
    const int o0 = g + bs0 * ((dg >> 1) & 1), o1 = g + bs1 * (dg & 1);
    for (int r = threadIdx.x; r < npt; r += blockDim.x) {
        const int j = r / nx, i = r - j * nx;
        const int f[3] = {o0 + i, o1 + j, 0};
        dst[r] = pred_point<N>(plane, nx, ny, 2, f);
    }
}

// 3D refineBlock, one thread per (x, y) column of a daughter component,
// marching in z: the x-then-y interpolated values of the mother's coarse
// planes (pred_point's vy) pass through a register window of 2N planes, and
// each fine z point is either the window's centre plane (even index) or the
// left-to-right z sum over the window (odd index) -- pred_point's arithmetic
// in pred_point's order, each vy computed once per column instead of once
// per output point.  Window planes outside the mother are set to zero; the
// host's stencil-range check guarantees no output reads them.
template <int N>
__global__ __launch_bounds__(256) void k_refine3d(double *__restrict__ arr,
                                                  const int32_t *__restrict__ pairs, int npairs,
                                                  int wpc, int bs0, int bs1, int bs2, int g, int nx,
                                                  int ny, int nz, size_t blk, size_t cs, int neq)
{
    constexpr double c2[2] = {0.5, 0.5};
    constexpr double c4[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
    constexpr double c6[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0,
                              150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};
    const double *cf = (N == 1) ? c2 : (N == 2) ? c4 : c6;
    // `wpc` workgroups per (pair, component): the pair and component are a
    // scalar division of the (XCD-swizzled: the daughters of a mother, listed
    // together, read it through one L2) workgroup index
    const int ncol = nx * ny;
    const int wg = xcd_swizzle((int)blockIdx.x, (int)gridDim.x);
    const int qc = wg / wpc;
    const int q = qc / neq, c = qc - q * neq;
    const int r = (wg - qc * wpc) * (int)blockDim.x + (int)threadIdx.x;
    if (r >= ncol) return;
    const int j = r / nx, i = r - j * nx;
    const int mother = pairs[3 * q], daughter = pairs[3 * q + 1], dg = pairs[3 * q + 2];
    const int f0 = g + bs0 * ((dg >> 1) & 1) + i, f1 = g + bs1 * (dg & 1) + j;
    const int fz0 = g + bs2 * ((dg >> 2) & 1);
    const int cx = f0 >> 1, px = f0 & 1, cy = f1 >> 1, py = f1 & 1;
    const double *u = arr + (size_t)mother * blk + (size_t)c * cs;
    double *dst = arr + (size_t)daughter * blk + (size_t)c * cs + (size_t)j * nx + i;
    auto vx = [&](int jy, int kz) -> double {
        const double *row = u + ((size_t)kz * ny + jy) * nx;
        if (!px) return row[cx];
        double s = cf[0] * row[cx - N + 1];
#pragma unroll
        for (int t = 1; t < 2 * N; ++t) s = s + cf[t] * row[cx - N + 1 + t];
        return s;
    };
    auto vy = [&](int kz) -> double {
        if (kz < 0 || kz >= nz) return 0.0;
        if (!py) return vx(cy, kz);
        double s = cf[0] * vx(cy - N + 1, kz);
#pragma unroll
        for (int t = 1; t < 2 * N; ++t) s = s + cf[t] * vx(cy - N + 1 + t, kz);
        return s;
    };
    int cz = fz0 >> 1;
    double w[2 * N];  // vy(cz - N + 1 .. cz + N)
#pragma unroll
    for (int t = 0; t < 2 * N; ++t) w[t] = vy(cz - N + 1 + t);
    const size_t pstride = (size_t)ncol;
    for (int k = 0; k < nz; ++k) {
        const int f2 = fz0 + k;
        if ((f2 >> 1) != cz) {
            ++cz;
#pragma unroll
            for (int t = 0; t < 2 * N - 1; ++t) w[t] = w[t + 1];
            w[2 * N - 1] = vy(cz + N);
        }
        double v;
        if (f2 & 1) {
            v = cf[0] * w[0];
#pragma unroll
            for (int t = 1; t < 2 * N; ++t) v = v + cf[t] * w[t];
        } else {
            v = w[N - 1];
        }
        dst[(size_t)k * pstride] = v;
    }
}

// 3D refineBlock, one workgroup per (pair, component), marching through the
// mother's coarse planes: per plane kz the coarse rows and columns the
// daughter reaches are staged in LDS (A), the x interpolation of every
// staged row at every fine x goes to LDS (B; pred_point's vx), and each
// thread takes the y interpolation (vy) of its own fine (x, y) points into a
// register window of the last 2N planes.  When the window ends at plane
// cz + N, the fine planes 2cz and 2cz + 1 are complete: the even one is the
// window's centre plane, the odd one the left-to-right z sum over the window
// -- pred_point's arithmetic in its order.  Coarse points outside the mother
// are staged as zero; the host's stencil-range check guarantees no output
// reads them.  PPT: fine (x, y) points per thread.
template <int N, int PPT>
__global__ __launch_bounds__(256) void k_refine3d_lds(double *__restrict__ arr,
                                                      const int32_t *__restrict__ pairs, int bs0, int bs1,
                                                      int bs2, int g, int nx, int ny, int nz, size_t blk,
                                                      size_t cs, int neq)
{
    constexpr double c2[2] = {0.5, 0.5};
    constexpr double c4[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
    constexpr double c6[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0,
                              150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};
    const double *cf = (N == 1) ? c2 : (N == 2) ? c4 : c6;
    extern __shared__ double lds[];
    const int wg = xcd_swizzle((int)blockIdx.x, (int)gridDim.x);
    const int q = wg / neq, c = wg - q * neq;
    const int mother = pairs[3 * q], daughter = pairs[3 * q + 1], dg = pairs[3 * q + 2];
    const int fx0 = g + bs0 * ((dg >> 1) & 1), fy0 = g + bs1 * (dg & 1), fz0 = g + bs2 * ((dg >> 2) & 1);
    // staged coarse window: columns [cxa, cxa + ncol), rows [cya, cya + nrow)
    const int cxa = (fx0 >> 1) - N + 1, ncol = ((fx0 + nx - 1) >> 1) + N - cxa + 1;
    const int cya = (fy0 >> 1) - N + 1, nrow = ((fy0 + ny - 1) >> 1) + N - cya + 1;
    double *A = lds, *B = lds + nrow * ncol;
    const double *u = arr + (size_t)mother * blk + (size_t)c * cs;
    double *dst = arr + (size_t)daughter * blk + (size_t)c * cs;
    const int nxy = nx * ny, tid = threadIdx.x;
    double win[PPT][2 * N];
    const int kz0 = (fz0 >> 1) - N + 1, kz1 = ((fz0 + nz - 1) >> 1) + N;
    for (int kz = kz0; kz <= kz1; ++kz) {
        const bool inz = kz >= 0 && kz < nz;
        __syncthreads();  // the previous plane's B reads are done
        for (int e = tid; e < nrow * ncol; e += 256) {
            const int r = e / ncol, cc = e - r * ncol;
            const int y = cya + r, x = cxa + cc;
            A[e] = (inz && y >= 0 && y < ny && x >= 0 && x < nx) ? u[((size_t)kz * ny + y) * nx + x] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < nrow * nx; e += 256) {
            const int r = e / nx, i = e - r * nx;
            const int f0 = fx0 + i, cx = (f0 >> 1) - cxa;
            const double *row = A + r * ncol;
            double v;
            if (f0 & 1) {
                v = cf[0] * row[cx - N + 1];
#pragma unroll
                for (int t = 1; t < 2 * N; ++t) v = v + cf[t] * row[cx - N + 1 + t];
            } else {
                v = row[cx];
            }
            B[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < PPT; ++pp) {
            const int pt = tid + 256 * pp;
            double v = 0.0;
            if (pt < nxy) {
                const int j = pt / nx, i = pt - j * nx;
                const int f1 = fy0 + j, cy = (f1 >> 1) - cya;
                if (f1 & 1) {
                    v = cf[0] * B[(cy - N + 1) * nx + i];
#pragma unroll
                    for (int t = 1; t < 2 * N; ++t) v = v + cf[t] * B[(cy - N + 1 + t) * nx + i];
                } else {
                    v = B[cy * nx + i];
                }
            }
#pragma unroll
            for (int t = 0; t < 2 * N - 1; ++t) win[pp][t] = win[pp][t + 1];
            win[pp][2 * N - 1] = v;
        }
        // fine planes whose window ends here
        const int cz = kz - N;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int f2 = 2 * cz + h, k = f2 - fz0;
            if (k < 0 || k >= nz) continue;
#pragma unroll
            for (int pp = 0; pp < PPT; ++pp) {
                const int pt = tid + 256 * pp;
                if (pt >= nxy) continue;
                double v;
                if (h) {
                    v = cf[0] * win[pp][0];
#pragma unroll
                    for (int t = 1; t < 2 * N; ++t) v = v + cf[t] * win[pp][t];
                } else {
                    v = win[pp][N - 1];
                }
                dst[(size_t)k * nxy + pt] = v;
            }
        }
    }
}

// LDS bytes of k_refine3d_lds for this pool and order, 0 if it does not apply
static size_t refine3d_lds_bytes(const wh_pool *P, int N)
{
    if (P->nx * P->ny > 8 * 256) return 0;
    const int ncol = P->nx / 2 + 2 * N + 2, nrow = P->ny / 2 + 2 * N + 2;  // upper bounds
    const size_t b = sizeof(double) * ((size_t)nrow * ncol + (size_t)nrow * P->nx);
    return b <= 64 * 1024 ? b : 0;
}

int launch_refine(wh_pool *P, double *arr, const int32_t *d_pairs, int npairs, int order)
{
    if (npairs <= 0) return WH_OK;
    if (P->d.dim == 3 && refine3d_lds_bytes(P, order / 2) > 0) {
        const int N = order / 2, ppt = (P->nx * P->ny + 255) / 256;
        const size_t lds = refine3d_lds_bytes(P, N);
        const dim3 grid((unsigned)((size_t)P->d.neq * npairs));
#define LRL(NN, PP)                                                                                    \
    hipLaunchKernelGGL((k_refine3d_lds<NN, PP>), grid, dim3(256), lds, P->stream, arr, d_pairs, P->d.bs[0], \
                       P->d.bs[1], P->d.bs[2], P->d.g, P->nx, P->ny, P->nz, P->blk, P->cs, P->d.neq)
#define LRN(NN)                        \
    do {                               \
        if (ppt <= 2) LRL(NN, 2);      \
        else if (ppt <= 4) LRL(NN, 4); \
        else if (ppt <= 6) LRL(NN, 6); \
        else LRL(NN, 8);               \
    } while (0)
        if (N == 2) LRN(2);
        else if (N == 1) LRN(1);
        else LRN(3);
#undef LRN
#undef LRL
        WH_HIP(hipGetLastError());
        return WH_OK;
    }
    if (P->d.dim == 3) {
        const int wpc = (P->nx * P->ny + 255) / 256;
        const dim3 grid((unsigned)((size_t)wpc * P->d.neq * npairs));
#define LR3(NN)                                                                                    \
    hipLaunchKernelGGL(k_refine3d<NN>, grid, dim3(256), 0, P->stream, arr, d_pairs, npairs, wpc,   \
                       P->d.bs[0], P->d.bs[1], P->d.bs[2], P->d.g, P->nx, P->ny, P->nz, P->blk,    \
                       P->cs, P->d.neq)
        if (order == 4) LR3(2);
        else if (order == 2) LR3(1);
        else LR3(3);
#undef LR3
        WH_HIP(hipGetLastError());
        return WH_OK;
    }
    const size_t plane = sizeof(double) * (size_t)P->nx * P->ny;
    if (P->d.dim == 2 && plane <= 64 * 1024) {
        const dim3 grid((unsigned)((size_t)P->d.neq * npairs));
#define LR2(NN)                                                                                              \
    hipLaunchKernelGGL(k_refine2d_lds<NN>, grid, dim3(256), plane, P->stream, arr, d_pairs, P->d.bs[0], P->d.bs[1], \
                       P->d.g, P->nx, P->ny, P->blk, P->cs, P->d.neq)
        if (order == 4) LR2(2);
        else if (order == 2) LR2(1);
        else LR2(3);
#undef LR2
        WH_HIP(hipGetLastError());
        return WH_OK;
    }
    const size_t npt = (size_t)P->nx * P->ny * P->nz * npairs;
    hipLaunchKernelGGL(k_refine, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, P->stream, arr,
                       d_pairs, npairs, P->d.bs[0], P->d.bs[1], P->d.dim == 3 ? P->d.bs[2] : 1,
                       P->d.g, P->nx, P->ny, P->nz, P->blk, P->cs, P->d.neq, P->d.dim, order);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_sync(wh_pool *P, int sync_case, double *arr, int one_receiver)
{
    PatchSet &ps = P->ps[sync_case];
    if (ps.n == 0) return WH_OK;
    if (one_receiver >= 0) {
        // one launch per (stage, receiver) row: the reference's per-block loop
        hipLaunchKernelGGL(k_sync, dim3(1), dim3(256), 0, P->stream, arr, ps.d_patches,
                           ps.d_recv_off, one_receiver, 0, P->blk, P->cs, P->nx, P->ny, P->d.neq);
        WH_HIP(hipGetLastError());
        return WH_OK;
    }
    for (int stage = 1; stage <= 3; ++stage) {
        const int p0 = ps.stage_begin[stage], p1 = ps.stage_begin[stage + 1];
        if (p1 <= p0) continue;
        if (int rc = launch_sync_flat(P, arr, ps.d_patches, ps.d_chunks, p0, p1, ps.chunk_begin[stage],
                                      ps.chunk_begin[stage + 1], SyncSrc{}))
            return rc;
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// ---------------------------------------------------------------------------
// RK combination over block interiors: out = base + sum coef_i * term_i
// ---------------------------------------------------------------------------
struct CombineArgs {
    const double *base;
    const double *terms[8];
    double coefs[8];
    int nterms;
};

__global__ __launch_bounds__(256) void k_combine(double *__restrict__ out, CombineArgs ca,
                                                 const int32_t *__restrict__ act, Geo G, int neq)
{
    const int npts = G.bs0 * G.bs1 * G.bs2;
    const int nchunk = (npts * neq + 255) / 256;
    const int k = blockIdx.x / nchunk;
    const int e = (blockIdx.x - k * nchunk) * 256 + threadIdx.x;
    if (e >= npts * neq) return;
    const int b = act[k];
    const int c = e / npts, pnt = e - c * npts;
    const int ix = pnt % G.bs0, iy = (pnt / G.bs0) % G.bs1, iz = pnt / (G.bs0 * G.bs1);
    const int zoff = (G.bs2 > 1) ? G.g : 0;
    const size_t i = (size_t)b * G.blk + c * G.cs +
                     ((size_t)(zoff + iz) * G.ny + (G.g + iy)) * G.nx + (G.g + ix);
    double v = ca.base[i];
    for (int t = 0; t < ca.nterms; ++t) v = v + ca.coefs[t] * ca.terms[t][i];
    out[i] = v;
}

int launch_combine(wh_pool *P, double *out, const double *base, int nterms,
                   const double *const *terms, const double *coefs, const int32_t *act, int nact)
{
    WH_CHECK(nterms >= 0 && nterms <= 8, WH_E_ARG, "rk_combine: at most 8 terms");
    if (nact <= 0) return WH_OK;
    CombineArgs ca{};
    ca.base = base;
    ca.nterms = nterms;
    for (int t = 0; t < nterms; ++t) {
        ca.terms[t] = terms[t];
        ca.coefs[t] = coefs[t];
    }
    const Geo G = make_geo(P);
    const int npts = G.bs0 * G.bs1 * G.bs2;
    const int nchunk = (npts * P->d.neq + 255) / 256;
    hipLaunchKernelGGL(k_combine, dim3(nact * nchunk), dim3(256), 0, P->stream, out, ca, act, G,
                       P->d.neq);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// ---------------------------------------------------------------------------
// reductions: per-block dt (GET_DT_BLOCK_ACM) and max |u| (divergence guard)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ double block_max(double v)
{
    __shared__ double red[16];
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    if (w == 0) {
        v = (l < (int)(blockDim.x >> 6)) ? red[l] : -1.0;
        v = wave_max(v);
    }
    return v;
}

// GET_DT_BLOCK_ACM of active block k.  WAVES 1: one wavefront per block
// (four blocks per workgroup; the 2D blocks' few hundred points), 0: the
// whole workgroup per block.  The max and min are exact whatever the order,
// so the per-group minimum may be combined in any grouping.
template <int WAVES>
__global__ __launch_bounds__(256) void k_dt(const double *__restrict__ in,
                                            const int32_t *__restrict__ act,
                                            const double *__restrict__ x0dx, Geo G, AcmDev a0,
                                            unsigned long long *dtmin, const int32_t *__restrict__ grp,
                                            const AcmDev *__restrict__ gacm, int nact)
{
    const int lane = WAVES ? (int)(threadIdx.x & 63) : (int)threadIdx.x, step = WAVES ? 64 : (int)blockDim.x;
    const int w = (int)(threadIdx.x >> 6);
    const int k = WAVES ? (int)blockIdx.x * 4 + w : (int)blockIdx.x;
    const bool live = !WAVES || k < nact;  // wave-uniform
    double dt = 0.0;
    if (live) {
        const int b = act[k];
        const int npts = G.bs0 * G.bs1 * G.bs2;
        const int zoff = (G.bs2 > 1) ? G.g : 0;
        const double *u = in + (size_t)b * G.blk;
        double m = -1.0;
        for (int p = lane; p < npts; p += step) {
            const int ix = p % G.bs0, iy = (p / G.bs0) % G.bs1, iz = p / (G.bs0 * G.bs1);
            const size_t i = ((size_t)(zoff + iz) * G.ny + (G.g + iy)) * G.nx + (G.g + ix);
            double v = u[i] * u[i] + u[G.cs + i] * u[G.cs + i];
            if (G.bs2 > 1) v = v + u[2 * G.cs + i] * u[2 * G.cs + i];
            m = fmax(m, v);
        }
        m = WAVES ? wave_max(m) : block_max(m);
        if (lane == 0) {
            const AcmDev &a = gacm ? gacm[grp[k]] : a0;  // per-group constants of independent forests
            double dxmin = x0dx[6 * k + 3];
            for (int d = 1; d < a.dim; ++d) dxmin = fmin(dxmin, x0dx[6 * k + 3 + d]);
            const double ueig = sqrt(m) + sqrt(a.c_0 * a.c_0 + m);
            dt = (ueig >= 1.0e-6) ? a.CFL * dxmin / ueig : 1.0e-2;
            if (a.nu > 1.0e-13) dt = fmin(dt, a.CFL_nu * (dxmin * dxmin) / a.nu);
            if (a.gamma_p > 0.0) dt = fmin(dt, a.CFL_eta * a.gamma_p);
            if (a.penal) dt = fmin(dt, a.CFL_eta * a.C_eta);
            if (a.sponge) dt = fmin(dt, a.CFL_eta * a.C_sponge);
            if (!WAVES) atomicMin(grp ? dtmin + grp[k] : dtmin, (unsigned long long)__double_as_longlong(dt));
        }
    }
    if (WAVES) {
        // the workgroup's four blocks (consecutive active entries: mostly one
        // group) combined first, one atomic per distinct group
        __shared__ double sdt[4];
        __shared__ int sgr[4];
        if (lane == 0) {
            sdt[w] = dt;
            sgr[w] = live ? (grp ? grp[k] : 0) : -1;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int x = 0; x < 4; ++x) {
                if (sgr[x] < 0) continue;
                bool first = true;
                for (int y = 0; y < x; ++y) first = first && sgr[y] != sgr[x];
                if (!first) continue;
                double v = sdt[x];
                for (int y = x + 1; y < 4; ++y)
                    if (sgr[y] == sgr[x]) v = fmin(v, sdt[y]);
                atomicMin(dtmin + sgr[x], (unsigned long long)__double_as_longlong(v));
            }
    }
}

__global__ __launch_bounds__(256) void k_maxabs(const double *__restrict__ in,
                                                const int32_t *__restrict__ act, Geo G, int neq,
                                                unsigned long long *mx)
{
    const int k = blockIdx.x, b = act[k];
    const int npts = G.bs0 * G.bs1 * G.bs2;
    const int zoff = (G.bs2 > 1) ? G.g : 0;
    const double *u = in + (size_t)b * G.blk;
    double m = 0.0;
    for (int e = threadIdx.x; e < npts * neq; e += blockDim.x) {
        const int c = e / npts, p = e - c * npts;
        const int ix = p % G.bs0, iy = (p / G.bs0) % G.bs1, iz = p / (G.bs0 * G.bs1);
        const size_t i = c * G.cs + ((size_t)(zoff + iz) * G.ny + (G.g + iy)) * G.nx + (G.g + ix);
        m = fmax(m, fabs(u[i]));
    }
    m = block_max(m);
    if (threadIdx.x == 0) atomicMax(mx, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_scalar_init(double *scal, int which, double v)
{
    if (threadIdx.x == 0) scal[which] = v;
}

int launch_dt(wh_pool *P, const AcmDev &a)
{
    hipLaunchKernelGGL(k_scalar_init, dim3(1), dim3(64), 0, P->stream, P->scal, (int)DS_DTMIN_BITS,
                       9.0e9);
    if (P->nact > 0) {
        const Geo G = make_geo(P);
        if (P->d.dim == 2)
            hipLaunchKernelGGL(k_dt<1>, dim3((P->nact + 3) / 4), dim3(256), 0, P->stream, P->block, P->d_active,
                               P->d_x0dx, G, a, (unsigned long long *)(P->scal + DS_DTMIN_BITS),
                               (const int32_t *)nullptr, (const AcmDev *)nullptr, P->nact);
        else
            hipLaunchKernelGGL(k_dt<0>, dim3(P->nact), dim3(256), 0, P->stream, P->block, P->d_active,
                               P->d_x0dx, G, a, (unsigned long long *)(P->scal + DS_DTMIN_BITS),
                               (const int32_t *)nullptr, (const AcmDev *)nullptr, P->nact);
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_dt_groups(wh_pool *P, const AcmDev &a)
{
    const int n = P->ngrp;
    const double init = 9.0e9;
    unsigned long long bits;
    std::memcpy(&bits, &init, sizeof bits);
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, P->stream, P->d_dtg, n,
                       bits);
    if (P->nact > 0) {
        const Geo G = make_geo(P);
        const AcmDev *gacm = P->ngacm == P->ngrp ? P->d_gacm : nullptr;
        if (P->d.dim == 2)
            hipLaunchKernelGGL(k_dt<1>, dim3((P->nact + 3) / 4), dim3(256), 0, P->stream, P->block, P->d_active,
                               P->d_x0dx, G, a, P->d_dtg, (const int32_t *)P->d_grp, gacm, P->nact);
        else
            hipLaunchKernelGGL(k_dt<0>, dim3(P->nact), dim3(256), 0, P->stream, P->block, P->d_active,
                               P->d_x0dx, G, a, P->d_dtg, (const int32_t *)P->d_grp, gacm, P->nact);
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_maxabs(wh_pool *P, const double *in)
{
    hipLaunchKernelGGL(k_scalar_init, dim3(1), dim3(64), 0, P->stream, P->scal,
                       (int)DS_MAXABS_BITS, 0.0);
    if (P->nact > 0) {
        const Geo G = make_geo(P);
        hipLaunchKernelGGL(k_maxabs, dim3(P->nact), dim3(256), 0, P->stream, in, P->d_active, G,
                           P->d.neq, (unsigned long long *)(P->scal + DS_MAXABS_BITS));
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// dt clamps of calculate_time_step.f90:58-118 on the device, so that a whole
// step can be replayed from a hipGraph without a host round trip.
__global__ void k_dt_finalize(double *scal, TimeDev tc)
{
    if (threadIdx.x != 0) return;
    const double time = scal[DS_TIME];
    double dt;
    if (tc.dt_fixed > 0.0) {
        dt = tc.dt_fixed;
    } else {
        dt = scal[DS_DTMIN_BITS];  // the min was taken on the bit pattern of positive doubles
        if (tc.dt_max > 0.0) dt = fmin(tc.dt_max, dt);
    }
    if (tc.write_fixed_time) {
        if (fmod(time + dt, tc.write_time) < fmod(time + 1e-12, tc.write_time) &&
            !(fabs(fmod(time, tc.write_time)) < 1e-12) && time + 1e-12 > tc.write_time_first)
            dt = tc.write_time - fmod(time, tc.write_time);
    }
    if (tc.write_backup_n > 0) {
        const double wb = tc.write_backup_time;
        if (fmod(time + dt, wb) < fmod(time + 1e-12, wb) && !(fabs(fmod(time, wb)) < 1e-12))
            dt = wb - fmod(time, wb);
    }
    if (tc.use_stats) {
        const double ts = tc.tsave_stats;
        if (fmod(time + dt, ts) < fmod(time + 1e-12, ts) && !(fabs(fmod(time, ts)) < 1e-12))
            dt = ts - fmod(time, ts);
    }
    if (time + dt > tc.time_max && time <= tc.time_max) dt = tc.time_max - time;
    if (!(time < tc.time_max)) dt = 0.0;  // main.f90:293 loop condition
    scal[DS_DT] = dt;
}

__global__ void k_time_advance(double *scal)
{
    if (threadIdx.x == 0 && scal[DS_DT] > 0.0) {
        scal[DS_TIME] = scal[DS_TIME] + scal[DS_DT];
        scal[DS_ITER] = scal[DS_ITER] + 1.0;
    }
}

int launch_dt_finalize(wh_pool *P, const TimeDev &tc)
{
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, P->stream, P->scal, tc);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_time_advance(wh_pool *P)
{
    hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(64), 0, P->stream, P->scal);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_set_scalar(wh_pool *P, int which, double v)
{
    hipLaunchKernelGGL(k_scalar_init, dim3(1), dim3(64), 0, P->stream, P->scal, which, v);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

AcmDev make_acm_dev(const wh_acm_params *p)
{
    AcmDev a{};
    a.c02 = p->c_0 * p->c_0;
    a.c_0 = p->c_0;
    a.nu = p->nu;
    a.gamma_p = p->gamma_p;
    a.C_eta = p->C_eta;
    a.Cinv = 1.0 / p->C_eta;
    a.C_sponge = p->C_sponge;
    a.Csp_inv = 1.0 / p->C_sponge;
    for (int i = 0; i < 3; ++i) a.ums[i] = p->u_mean_set[i];
    a.CFL = p->CFL;
    a.CFL_nu = p->CFL_nu;
    a.CFL_eta = p->CFL_eta;
    a.dim = p->dim;
    a.scheme = p->scheme;
    a.skew = p->skew_symmetry ? 1 : 0;
    a.penal = p->penalization ? 1 : 0;
    a.sponge = p->use_sponge ? 1 : 0;
    return a;
}

}  // namespace wh
