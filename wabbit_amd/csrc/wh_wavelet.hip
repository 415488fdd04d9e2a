// wh_wavelet.hip -- CDF wavelet machinery on the device pool: filter setup
// (host), one-level FWT / IWT and the restriction filter of listed blocks,
// Linfty details for the coarsening indicator, coarse-extension boxes and
// explicit ghost-patch lists with selectable source arrays.
//
// Layout: block b of an array at base + b*blk, component c at + c*cs, point
// (ix, iy, iz) at ((iz*ny + iy)*nx + ix), x fastest (Fortran hvy_block).
//
// Summation order.  Every filter output is a chain of products added left
// to right, the order of the reference's loops (gfortran, no reassociation):
// the unrolled cases of wavelet_decomposition_reconstruction.f90 (CDF20/22/
// 40/42/44/60/62) skip zero taps and, in the reconstruction, start with the
// point's own coefficient; the generic case adds two separate sums;
// blockFilterXYZ_vct starts from zero and adds every tap.  A TermTable holds
// that chain per output parity, so results match the CPU restatement bit for
// bit (compiled with -ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "wh_internal.h"
#include "wh_prof.h"

namespace wh {

// ---------------------------------------------------------------------------
// setup_wavelet (module_wavelets.f90:1031-1430)
// ---------------------------------------------------------------------------
namespace {

// Stencil_Int_<order>th (module_wavelets.f90:14-22) on -(order-1)..order-1
void interp_stencil(int order, double *c)
{
    if (order == 2) {
        const double v[3] = {1.0, 2.0, 1.0};
        for (int i = 0; i < 3; ++i) c[i] = v[i] / 2.0;
    } else if (order == 4) {
        const double v[7] = {-1.0, 0.0, 9.0, 16.0, 9.0, 0.0, -1.0};
        for (int i = 0; i < 7; ++i) c[i] = v[i] / 16.0;
    } else {
        const double v[11] = {3.0, 0.0, -25.0, 0.0, 150.0, 256.0, 150.0, 0.0, -25.0, 0.0, 3.0};
        for (int i = 0; i < 11; ++i) c[i] = v[i] / 256.0;
    }
}

inline double &tap(double *f, int lo, int i) { return f[i - lo]; }
inline double tapc(const double *f, int lo, int i) { return f[i - lo]; }

}  // namespace
}  // namespace wh

extern "C" int wh_wavelet_setup(const char *name, int fd_max_size, wh_wavelet *w)
{
    using namespace wh;
    WH_CHECK(w && name && std::strlen(name) == 5 && std::strncmp(name, "CDF", 3) == 0, WH_E_ARG,
             "wh_wavelet_setup: expected a name CDFXY");
    const int X = name[3] - '0', Y = name[4] - '0';
    WH_CHECK((X == 2 || X == 4 || X == 6) && (Y == 0 || Y == 2 || Y == 4 || Y == 6) && Y <= X,
             WH_E_ARG, "wh_wavelet_setup: supported are CDFXY with X in {2,4,6}, Y in {0,2,4,6}, Y <= X");
    WH_CHECK(fd_max_size >= 0 && fd_max_size <= 8, WH_E_ARG, "wh_wavelet_setup: bad fd_max_size");
    std::memset(w, 0, sizeof(*w));
    w->X = X;
    w->Y = Y;
    w->lifted = Y != 0;
    w->hr_lo = -(X - 1);
    w->hr_hi = X - 1;
    interp_stencil(X, w->HR);
    if (Y == 0) {
        w->hd_lo = w->hd_hi = 0;
        w->HD[0] = 1.0;
    } else {
        // HD(i) = delta_i0 + 1/2 sum_j (-1)^j HR(j) h(i-j), h = order-Y stencil
        // with its centre tap removed (:1182-1234)
        double h[11];
        interp_stencil(Y, h);
        const int h_lo = -(Y - 1), h_hi = Y - 1;
        h[-h_lo] = 0.0;
        w->hd_lo = w->hr_lo + h_lo;
        w->hd_hi = w->hr_hi + h_hi;
        for (int i = w->hd_lo; i <= w->hd_hi; ++i) {
            double v = (i == 0) ? 1.0 : 0.0;
            for (int j = w->hr_lo; j <= w->hr_hi; ++j) {
                const int k = i - j;
                if (k < h_lo || k > h_hi) continue;
                const double sgn = (j % 2 == 0) ? 1.0 : -1.0;
                v = v + sgn * tapc(w->HR, w->hr_lo, j) * h[k - h_lo] / 2.0;
            }
            tap(w->HD, w->hd_lo, i) = v;
        }
    }
    // GD(i) = (-1)^i HR(i), GR(i) = (-1)^i HD(i) (:1275-1288)
    w->gd_lo = w->hr_lo;
    w->gd_hi = w->hr_hi;
    for (int i = w->gd_lo; i <= w->gd_hi; ++i)
        tap(w->GD, w->gd_lo, i) = ((i % 2 == 0) ? 1.0 : -1.0) * tapc(w->HR, w->hr_lo, i);
    w->gr_lo = w->hd_lo;
    w->gr_hi = w->hd_hi;
    for (int i = w->gr_lo; i <= w->gr_hi; ++i)
        tap(w->GR, w->gr_lo, i) = ((i % 2 == 0) ? 1.0 : -1.0) * tapc(w->HD, w->hd_lo, i);
    w->order = X;
    const int bounds[8] = {w->hd_lo, w->hd_hi, w->gd_lo, w->gd_hi,
                           w->hr_lo, w->hr_hi, w->gr_lo, w->gr_hi};
    w->g = 0;
    for (int b : bounds) w->g = std::max(w->g, std::abs(b));
    w->g_rhs = std::max(X / 2, fd_max_size);
    // coarse-extension widths (:1368-1420) with the FD minimum 2*reach
    w->Nscl = std::max(std::abs(w->hd_lo) - 1, 0);
    w->Nwcl = w->Nscl + std::abs(w->gd_lo);
    w->Nreconl = w->Nwcl + std::abs(w->gr_lo);
    w->Ndep2l = w->Nreconl + std::max((std::abs(w->hr_lo) + 1) / 2 - 1, 0);
    w->Nscr = w->hd_hi;
    w->Nwcr = w->Nscr + w->gd_hi;
    w->Nreconr = w->Nwcr + w->gr_hi;
    w->Ndep2r = w->Nreconr + (w->hr_hi + 1) / 2;
    const int m = 2 * fd_max_size;
    const int dl = std::max(m - w->Nwcl, 0), dr = std::max(m - w->Nwcr, 0);
    w->Nwcl += dl;
    w->Nreconl += dl;
    w->Ndep2l += dl;
    w->Nwcr += dr;
    w->Nreconr += dr;
    w->Ndep2r += dr;
    w->specialised = (Y == 0) || (X == 2 && Y == 2) || (X == 4 && (Y == 2 || Y == 4)) ||
                     (X == 6 && Y == 2);
    return WH_OK;
}

namespace wh {
namespace {

// ---------------------------------------------------------------------------
// term chains
// ---------------------------------------------------------------------------
constexpr int TMAX = 24;
struct TermTable {
    int n[2];  // [0]: output at an SC point of the pass direction, [1]: WC point
    int off[2][TMAX];
    double c[2][TMAX];
    int brk[2][TMAX];  // 1: term starts a new sum added to the running total
    int from_zero;     // accumulate from 0.0 (blockFilterXYZ_vct)
};

void add_term(TermTable &T, int par, int off, double c, int brk = 0)
{
    const int k = T.n[par]++;
    T.off[par][k] = off;
    T.c[par][k] = c;
    T.brk[par][k] = brk;
}

TermTable fwt_terms(const wh_wavelet &w)
{
    TermTable T{};
    for (int j = w.hd_lo; j <= w.hd_hi; ++j)
        if (tapc(w.HD, w.hd_lo, j) != 0.0) add_term(T, 0, j, tapc(w.HD, w.hd_lo, j));
    for (int j = w.gd_lo; j <= w.gd_hi; ++j)
        if (tapc(w.GD, w.gd_lo, j) != 0.0) add_term(T, 1, j, tapc(w.GD, w.gd_lo, j));
    return T;
}

TermTable iwt_terms(const wh_wavelet &w)
{
    TermTable T{};
    auto hr = [&](int j) { return tapc(w.HR, w.hr_lo, j); };
    auto gr = [&](int j) { return tapc(w.GR, w.gr_lo, j); };
    if (!w.lifted) {
        // u_r = u; WC points: u_r(ix) + HR(odd taps) (:483-537)
        add_term(T, 0, 0, 1.0);
        add_term(T, 1, 0, 1.0);
        for (int j = w.hr_lo; j <= w.hr_hi; ++j)
            if ((j & 1) && hr(j) != 0.0) add_term(T, 1, j, hr(j));
    } else if (w.specialised) {
        // SC: u(ix) + GR(odd); WC: HR(odd), GR(0), GR(even != 0), one chain
        add_term(T, 0, 0, 1.0);
        for (int j = w.gr_lo; j <= w.gr_hi; ++j)
            if ((j & 1) && gr(j) != 0.0) add_term(T, 0, j, gr(j));
        for (int j = w.hr_lo; j <= w.hr_hi; ++j)
            if ((j & 1) && hr(j) != 0.0) add_term(T, 1, j, hr(j));
        add_term(T, 1, 0, gr(0));
        for (int j = w.gr_lo; j <= w.gr_hi; ++j)
            if (j != 0 && !(j & 1) && gr(j) != 0.0) add_term(T, 1, j, gr(j));
    } else {
        // sum(HR * SC-buffer) + sum(GR * WC-buffer) (:797-838)
        add_term(T, 0, 0, 1.0);
        int first = 1;
        for (int j = w.gr_lo; j <= w.gr_hi; ++j)
            if ((j & 1) && gr(j) != 0.0) {
                add_term(T, 0, j, gr(j), first);
                first = 0;
            }
        for (int j = w.hr_lo; j <= w.hr_hi; ++j)
            if ((j & 1) && hr(j) != 0.0) add_term(T, 1, j, hr(j));
        first = 1;
        for (int j = w.gr_lo; j <= w.gr_hi; ++j)
            if (!(j & 1) && gr(j) != 0.0) {
                add_term(T, 1, j, gr(j), first);
                first = 0;
            }
    }
    return T;
}

TermTable filter_terms(const wh_wavelet &w)
{
    TermTable T{};
    T.from_zero = 1;
    for (int j = w.hd_lo; j <= w.hd_hi; ++j) add_term(T, 0, j, tapc(w.HD, w.hd_lo, j));
    return T;
}

// Chain kinds: a plain left-to-right sum (no breaks), a sum from zero, or
// the general form with breaks.  The plain and from-zero kinds evaluate the
// same additions as the general form without its per-term selects.
enum { CHAIN_PLAIN = 0, CHAIN_ZERO = 1, CHAIN_GENERAL = 2 };

int chain_kind(const TermTable &T)
{
    if (T.from_zero) return CHAIN_ZERO;
    for (int par = 0; par < 2; ++par) {
        if (T.n[par] < 1) return CHAIN_GENERAL;
        for (int k = 0; k < T.n[par]; ++k)
            if (T.brk[par][k]) return CHAIN_GENERAL;
    }
    return CHAIN_PLAIN;
}

template <int KIND, class F>
__device__ __forceinline__ double chain_f(const TermTable &T, int par, F at)
{
    const int n = T.n[par];
    if constexpr (KIND == CHAIN_ZERO) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s = s + at(T.off[par][k]) * T.c[par][k];
        return s;
    } else if constexpr (KIND == CHAIN_PLAIN) {
        double s = at(T.off[par][0]) * T.c[par][0];
        for (int k = 1; k < n; ++k) s = s + at(T.off[par][k]) * T.c[par][k];
        return s;
    }
    double total = 0.0, cur = 0.0;
    bool have = false;
    for (int k = 0; k < n; ++k) {
        const double v = at(T.off[par][k]) * T.c[par][k];
        if (k == 0) {
            cur = v;
        } else if (T.brk[par][k]) {
            total = have ? total + cur : cur;
            have = true;
            cur = v;
        } else {
            cur = cur + v;
        }
    }
    return have ? total + cur : cur;
}

// The plain and from-zero kinds unrolled to NT terms (n <= NT): all NT
// points are read first (unused entries have offset 0, an in-range point),
// then the n used ones are added in order -- the same additions as chain_f.
template <int KIND, int NT, class F>
__device__ __forceinline__ double chain_n(const TermTable &T, int par, F at)
{
    static_assert(KIND == CHAIN_PLAIN || KIND == CHAIN_ZERO, "unrolled chains have no breaks");
    const int n = T.n[par];
    double v[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) v[k] = at(T.off[par][k]);
    double s = KIND == CHAIN_ZERO ? 0.0 : v[0] * T.c[par][0];
#pragma unroll
    for (int k = KIND == CHAIN_ZERO ? 0 : 1; k < NT; ++k)
        if (k < n) s = s + v[k] * T.c[par][k];
    return s;
}

// kernel variant of a chain table: kind and unroll length (0: loop)
struct ChainVar {
    int kind, nt;
};
ChainVar chain_var(const TermTable &T)
{
    ChainVar v{chain_kind(T), 0};
    if (v.kind == CHAIN_GENERAL) return v;
    const int n = std::max(T.n[0], T.n[1]);
    v.nt = n <= 8 ? 8 : n <= 12 ? 12 : n <= 16 ? 16 : 0;
    return v;
}

// KIND and NT select the evaluation (NT 0: the loop of chain_f)
template <int KIND, int NT, class F>
__device__ __forceinline__ double chain_v(const TermTable &T, int par, F at)
{
    if constexpr (NT == 0) return chain_f<KIND>(T, par, at);
    else return chain_n<KIND, NT>(T, par, at);
}

// launch L(KIND, NT) for a ChainVar
#define CHAIN_DISPATCH(VAR, L)                                \
    do {                                                      \
        const ChainVar cv_ = (VAR);                           \
        if (cv_.kind == CHAIN_PLAIN && cv_.nt == 8) L(CHAIN_PLAIN, 8);        \
        else if (cv_.kind == CHAIN_PLAIN && cv_.nt == 12) L(CHAIN_PLAIN, 12); \
        else if (cv_.kind == CHAIN_PLAIN && cv_.nt == 16) L(CHAIN_PLAIN, 16); \
        else if (cv_.kind == CHAIN_PLAIN) L(CHAIN_PLAIN, 0);                  \
        else if (cv_.kind == CHAIN_ZERO && cv_.nt == 8) L(CHAIN_ZERO, 8);     \
        else if (cv_.kind == CHAIN_ZERO && cv_.nt == 12) L(CHAIN_ZERO, 12);   \
        else if (cv_.kind == CHAIN_ZERO && cv_.nt == 16) L(CHAIN_ZERO, 16);   \
        else if (cv_.kind == CHAIN_ZERO) L(CHAIN_ZERO, 0);                    \
        else L(CHAIN_GENERAL, 0);                                             \
    } while (0)

// the chain at p, p[off*stride] being the point at offset off
template <int KIND>
__device__ __forceinline__ double chain_k(const TermTable &T, int par, const double *p, long stride)
{
    return chain_f<KIND>(T, par, [&](int off) { return p[off * stride]; });
}

// n / d for 0 <= n, d < 2^16 by a multiply-high with m = ceil(2^32 / d):
// n*m / 2^32 exceeds n/d by less than 2^-16 <= 1/d, so the floor is exact.
struct FastDiv {
    uint32_t m;
    int one;
};
FastDiv fast_div(int d)
{
    FastDiv f;
    f.one = d == 1;
    f.m = d <= 1 ? 0u : (uint32_t)((0x100000000ull + (uint64_t)d - 1) / (uint64_t)d);
    return f;
}
__device__ __forceinline__ int fdiv(int n, FastDiv f) { return f.one ? n : (int)__umulhi((unsigned)n, f.m); }

// Output points of a pass: lo + step*i, i < n per direction.
struct Region {
    int lo[3], n[3], step[3];
};

struct ArrGeo {
    size_t blk_in, blk_out, cs;
    int nx, ny, neq;
};

// Per-point launches over listed blocks: `wpb` workgroups per (block,
// component) pair, so the pair is a scalar division of the workgroup index
// and a lane's point index t stays in 32 bits.
struct PtIdx {
    int k, c, t;
};
__device__ __forceinline__ PtIdx pt_index(int wpb, int neq)
{
    const int bc = (int)blockIdx.x / wpb;
    PtIdx r;
    r.k = bc / neq;
    r.c = bc - r.k * neq;
    r.t = ((int)blockIdx.x - bc * wpb) * (int)blockDim.x + (int)threadIdx.x;
    return r;
}

// One separable pass over the listed blocks: out = chain_k(in) along `axis`;
// the parity of the output point (relative to g) picks the SC/WC chain.
template <int KIND>
__global__ __launch_bounds__(256) void k_wpass(const double *__restrict__ in,
                                               double *__restrict__ out,
                                               const int32_t *__restrict__ ids, int wpb,
                                               ArrGeo A, Region R, int axis, int astride, int g,
                                               TermTable T)
{
    const PtIdx I = pt_index(wpb, A.neq);
    if (I.t >= R.n[0] * R.n[1] * R.n[2]) return;
    const int t1 = I.t / R.n[0];
    const int i0 = I.t - t1 * R.n[0], i2 = t1 / R.n[1], i1 = t1 - i2 * R.n[1];
    const int p0 = R.lo[0] + R.step[0] * i0, p1 = R.lo[1] + R.step[1] * i1,
              p2 = R.lo[2] + R.step[2] * i2;
    const int q = (p2 * A.ny + p1) * A.nx + p0;
    const int pa = axis == 0 ? p0 : (axis == 1 ? p1 : p2);
    const int par = (pa - g) & 1;
    const size_t b = (size_t)ids[I.k];
    out[b * A.blk_out + (size_t)I.c * A.cs + q] =
        chain_k<KIND>(T, par, in + b * A.blk_in + (size_t)I.c * A.cs + q, astride);
}

// Region membership along one direction, for steps 1 and 2 (checked on the
// host before launch).
__device__ __forceinline__ bool in_dir(const Region &R, int d, int p)
{
    const int r = p - R.lo[d];
    return r >= 0 && (r & (R.step[d] - 1)) == 0 && (r >> (R.step[d] - 1)) < R.n[d];
}

// Output positions along a pass direction split by chain parity: class 0
// holds i = 0, s, 2s, ... (parity par0), class 1 (steps of 1 only) i = 1, 3,
// ...; one class-row of items per enumeration index, each class padded to a
// whole number of wavefronts, so a wavefront evaluates one chain.
struct ParSplit {
    int par0, s, nA, nB;
};
ParSplit par_split(const Region &R, int d, int g)
{
    ParSplit P;
    P.par0 = ((R.lo[d] - g) % 2 + 2) % 2;
    if (R.step[d] % 2 == 0) {
        P.s = 1;
        P.nA = R.n[d];
        P.nB = 0;
    } else {
        P.s = 2;
        P.nA = (R.n[d] + 1) / 2;
        P.nB = R.n[d] / 2;
    }
    return P;
}

// The x and y passes of one z-plane of one component of one block in one
// workgroup (the whole pass in 2D).  The source plane is staged in LDS (S);
// the x pass writes Rx into D; the y pass reads D along y and writes Ry into
// S (which the x pass no longer needs); the plane's Rx points are written
// out (Ry points from S, the rest of Rx from D), and with init_copy the
// points outside Rx take the source's values.  The y chains read only rows
// of Rx and Ry's x positions are Rx's (checked on the host), so every value
// is the one of the pass-by-pass version bit for bit.  Outputs are
// enumerated by parity class (ParSplit), so each wavefront's chain is
// uniform and its table is read through the scalar unit.
struct SepDiv {
    FastDiv xa, xb, ry, rx, nx;  // by Px.nA, Px.nB, Ry.n[0], Rx.n[0], A.nx
};
template <int KIND, int NT>
__global__ __launch_bounds__(256) void k_sepxy(const double *__restrict__ in, double *__restrict__ out,
                                               const int32_t *__restrict__ ids, ArrGeo A, Region Rx,
                                               Region Ry, ParSplit Px, ParSplit Py, SepDiv V, int nzp,
                                               int init_copy, TermTable T)
{
    extern __shared__ double sp[];
    const int nxy = A.nx * A.ny;
    double *S = sp, *D = sp + nxy;
    const int bc = (int)blockIdx.x / nzp, z = (int)blockIdx.x - bc * nzp;
    const int k = bc / A.neq, c = bc - k * A.neq;
    const size_t b = (size_t)ids[k];
    const double *src = in + b * A.blk_in + (size_t)c * A.cs + (size_t)z * nxy;
    double *dst = out + b * A.blk_out + (size_t)c * A.cs + (size_t)z * nxy;
    if (!in_dir(Rx, 2, z)) {  // a plane outside the passes (uniform)
        if (init_copy)
            for (int i = threadIdx.x; i < nxy; i += blockDim.x) dst[i] = src[i];
        return;
    }
    for (int i = threadIdx.x; i < nxy; i += blockDim.x) S[i] = src[i];
    __syncthreads();
    {   // x pass: class items are (position class index j, row)
        const int na = Px.nA * Rx.n[1], pa = (na + 63) & ~63, tot = pa + Px.nB * Rx.n[1];
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const bool cb = e >= pa;
            const int r = cb ? e - pa : e, nc = cb ? Px.nB : Px.nA;
            if (!cb && e >= na) continue;
            const int i1 = fdiv(r, cb ? V.xb : V.xa), j = r - i1 * nc;
            const int i0 = cb ? 1 + 2 * j : Px.s * j;
            const int q = (Rx.lo[1] + Rx.step[1] * i1) * A.nx + Rx.lo[0] + Rx.step[0] * i0;
            const int par = __builtin_amdgcn_readfirstlane(cb ? 1 - Px.par0 : Px.par0);
            const double *sq = S + q;
            D[q] = chain_v<KIND, NT>(T, par, [&](int off) { return sq[off]; });
        }
    }
    __syncthreads();
    {   // y pass: class items are (column, row class index j)
        const int na = Py.nA * Ry.n[0], pa = (na + 63) & ~63, tot = pa + Py.nB * Ry.n[0];
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const bool cb = e >= pa;
            const int r = cb ? e - pa : e;
            if (!cb && e >= na) continue;
            const int j = fdiv(r, V.ry), i0 = r - j * Ry.n[0];
            const int i1 = cb ? 1 + 2 * j : Py.s * j;
            const int q = (Ry.lo[1] + Ry.step[1] * i1) * A.nx + Ry.lo[0] + Ry.step[0] * i0;
            const int par = __builtin_amdgcn_readfirstlane(cb ? 1 - Py.par0 : Py.par0);
            const double *dq = D + q;
            const int nx = A.nx;
            S[q] = chain_v<KIND, NT>(T, par, [&](int off) { return dq[off * nx]; });
        }
    }
    __syncthreads();
    if (init_copy) {
        for (int q = threadIdx.x; q < nxy; q += blockDim.x) {
            const int p1 = fdiv(q, V.nx), p0 = q - p1 * A.nx;
            const bool x_only = in_dir(Rx, 0, p0) && in_dir(Rx, 1, p1) && !in_dir(Ry, 1, p1);
            dst[q] = x_only ? D[q] : S[q];
        }
    } else {
        const int nrx = Rx.n[0] * Rx.n[1];
        for (int e = threadIdx.x; e < nrx; e += blockDim.x) {
            const int i1 = fdiv(e, V.rx), i0 = e - i1 * Rx.n[0];
            const int p1 = Rx.lo[1] + Rx.step[1] * i1;
            const int q = p1 * A.nx + Rx.lo[0] + Rx.step[0] * i0;
            dst[q] = in_dir(Ry, 1, p1) ? S[q] : D[q];
        }
    }
}

// A pass along y or z done in place, one thread per column, marching along
// it: a ring of WRING LDS slots per thread holds the column's points from
// the lowest offset of the current output's chain to the highest; a point is
// loaded once, before any output at or above it is written, so every chain
// reads the values before the pass, as the pass-then-copy-back version does.
// Columns of a wavefront are consecutive x positions (coalesced); the output
// position, and with it the chain, is uniform.
// WRING slots per column (16, or 32 for the longest chains), 4096 / WRING
// columns per workgroup: 32 KB of LDS either way.
template <int KIND, int NT, int WRING>
__global__ __launch_bounds__(256) void k_wring(double *__restrict__ arr, const int32_t *__restrict__ ids,
                                               int wpb, ArrGeo A, Region R, int axis, int len, int astride,
                                               int g, int lo_off, int hi_off, TermTable T)
{
    constexpr int WRING_LANES = 4096 / WRING;
    __shared__ double ring[WRING * WRING_LANES];  // [slot][lane]
    // the column runs along `axis`; x and the remaining direction (z for a
    // y pass, y for a z pass) number the columns
    const bool ya = axis == 1;
    const int n1 = ya ? R.n[2] : R.n[1], lo1 = ya ? R.lo[2] : R.lo[1], st1 = ya ? R.step[2] : R.step[1];
    const int na = ya ? R.n[1] : R.n[2], loa = ya ? R.lo[1] : R.lo[2], sta = ya ? R.step[1] : R.step[2];
    const PtIdx I = pt_index(wpb, A.neq);
    if (I.t >= R.n[0] * n1) return;
    const int i1 = I.t / R.n[0], i0 = I.t - i1 * R.n[0];
    const int px = R.lo[0] + R.step[0] * i0, p1 = lo1 + st1 * i1;
    const int q0 = ya ? p1 * A.ny * A.nx + px : p1 * A.nx + px;
    double *base = arr + (size_t)ids[I.k] * A.blk_out + (size_t)I.c * A.cs + q0;
    double *mine = ring + threadIdx.x;
    // points next .. next+3 are in flight in f0..f3 (zero past the column's
    // end, never read); each is loaded before any output at or above it is
    // stored, and program order keeps a thread's load before its later store
    // to the same address
    auto ld = [&](int x) { return x < len ? base[x * astride] : 0.0; };
    int next = max(loa + lo_off, 0);
    double f0 = ld(next), f1 = ld(next + 1), f2 = ld(next + 2), f3 = ld(next + 3);
    for (int i = 0; i < na; ++i) {
        const int p = loa + sta * i;
        const int need = min(p + hi_off, len - 1);
        for (; next <= need; ++next) {
            mine[(next & (WRING - 1)) * WRING_LANES] = f0;
            f0 = f1;
            f1 = f2;
            f2 = f3;
            f3 = ld(next + 4);
        }
        base[p * astride] =
            chain_v<KIND, NT>(T, (p - g) & 1, [&](int off) { return mine[((p + off) & (WRING - 1)) * WRING_LANES]; });
    }
}

// Whole listed blocks (all components, contiguous per block) copied in
// 16-byte words where every block offset and the length are even, in
// doubles otherwise.
template <typename W>
__global__ __launch_bounds__(256) void k_copy_blocks(const double *__restrict__ in, size_t blk_in,
                                                     double *__restrict__ out, size_t blk_out,
                                                     const int32_t *__restrict__ ids, int wpb, int nw)
{
    const int k = (int)blockIdx.x / wpb;
    const int t = ((int)blockIdx.x - k * wpb) * 512 + (int)threadIdx.x;
    const size_t b = (size_t)ids[k];
    const W *src = reinterpret_cast<const W *>(in + b * blk_in);
    W *dst = reinterpret_cast<W *>(out + b * blk_out);
    if (t < nw) dst[t] = src[t];
    if (t + 256 < nw) dst[t + 256] = src[t + 256];
}

// Linfty detail: per listed block and component, max over the interior of
// |v| (sqrt(v*v) for grouped components), SC points counted as 0.
struct DetailArgs {
    int g, bs[3], dim, neq;
    int groups[8];
};
__global__ __launch_bounds__(256) void k_detail(const double *__restrict__ arr,
                                                const int32_t *__restrict__ ids, ArrGeo A,
                                                DetailArgs D, double *__restrict__ out)
{
    __shared__ unsigned long long smax;
    if (threadIdx.x == 0) smax = 0ull;
    __syncthreads();
    const int k = (int)blockIdx.x / D.neq, c = (int)blockIdx.x - k * D.neq;
    const int b = ids[k];
    const int ncol = D.bs[0] * D.bs[1], nz = D.bs[2];
    const int zoff = D.dim == 3 ? D.g : 0;
    const bool sq = D.groups[c] >= 2;
    const double *u = arr + (size_t)b * A.blk_in + (size_t)c * A.cs;
    double m = 0.0;
    for (int e = threadIdx.x; e < ncol; e += blockDim.x) {
        const int iy = e / D.bs[0], ix = e - iy * D.bs[0];
        const bool sc2 = ((ix & 1) == 0) && ((iy & 1) == 0);
        const double *col = u + (size_t)(zoff * A.ny + D.g + iy) * A.nx + (D.g + ix);
        const size_t zs = (size_t)A.nx * A.ny;
        for (int iz = 0; iz < nz; ++iz) {
            const bool sc = sc2 && (D.dim == 2 || (iz & 1) == 0);
            const double v = sc ? 0.0 : col[iz * zs];
            const double a = sq ? sqrt(v * v) : fabs(v);
            m = a > m ? a : m;
        }
    }
    // non-negative doubles order like their bit patterns
    atomicMax(&smax, (unsigned long long)__double_as_longlong(m));
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = __longlong_as_double((long long)smax);
}

// k_detail with one wavefront per (block, component), four per workgroup:
// the 2D blocks' few hundred points; the wavefront's lanes reduce by
// shuffles (the max is exact in any order)
__global__ __launch_bounds__(256) void k_detail_wave(const double *__restrict__ arr,
                                                     const int32_t *__restrict__ ids, ArrGeo A,
                                                     DetailArgs D, int nitems, double *__restrict__ out)
{
    const int item = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    if (item >= nitems) return;  // whole wavefronts; no barrier in this kernel
    const int k = item / D.neq, c = item - k * D.neq;
    const int b = ids[k];
    const int ncol = D.bs[0] * D.bs[1], nz = D.bs[2];
    const int zoff = D.dim == 3 ? D.g : 0;
    const bool sq = D.groups[c] >= 2;
    const double *u = arr + (size_t)b * A.blk_in + (size_t)c * A.cs;
    double m = 0.0;
    for (int e = lane; e < ncol; e += 64) {
        const int iy = e / D.bs[0], ix = e - iy * D.bs[0];
        const bool sc2 = ((ix & 1) == 0) && ((iy & 1) == 0);
        const double *col = u + (size_t)(zoff * A.ny + D.g + iy) * A.nx + (D.g + ix);
        const size_t zs = (size_t)A.nx * A.ny;
        for (int iz = 0; iz < nz; ++iz) {
            const bool sc = sc2 && (D.dim == 2 || (iz & 1) == 0);
            const double v = sc ? 0.0 : col[iz * zs];
            const double a = sq ? sqrt(v * v) : fabs(v);
            m = a > m ? a : m;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(m, o, 64);
        m = t > m ? t : m;
    }
    if (lane == 0) out[item] = m;
}

// Point types: bit d set where the point lies on a scaling-coefficient
// position in direction d (index - g even); the all-SC type (2^dim - 1) is
// the zeroed SC itself.  wavelet_renorm_block scales all points of one
// type of one block by one factor, and rounding is monotone, so the max of
// the renormed values is the renormed max of |v| per type: the typed
// kernels below return those maxima (2^dim - 1 per block and component)
// and the host applies the level and direction factors.
__device__ inline void typed_max(double (&m)[7], int t, double a)
{
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if (q == t) m[q] = a > m[q] ? a : m[q];
}

// Per listed block and component: max over the interior of |v| per point
// type, out[(k*neq + c)*(2^dim - 1) + t].
__global__ __launch_bounds__(256) void k_detail_typed(const double *__restrict__ arr,
                                                      const int32_t *__restrict__ ids, ArrGeo A,
                                                      DetailArgs D, double *__restrict__ out)
{
    __shared__ unsigned long long smax[7];
    if (threadIdx.x < 7) smax[threadIdx.x] = 0ull;
    __syncthreads();
    const int k = (int)blockIdx.x / D.neq, c = (int)blockIdx.x - k * D.neq;
    const int b = ids[k];
    const int ncol = D.bs[0] * D.bs[1], nz = D.bs[2];
    const int zoff = D.dim == 3 ? D.g : 0;
    const int nt = (1 << D.dim) - 1;
    const double *u = arr + (size_t)b * A.blk_in + (size_t)c * A.cs;
    double m[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = threadIdx.x; e < ncol; e += blockDim.x) {
        const int iy = e / D.bs[0], ix = e - iy * D.bs[0];
        const int t2 = ((ix & 1) == 0 ? 1 : 0) | ((iy & 1) == 0 ? 2 : 0);
        const double *col = u + (size_t)(zoff * A.ny + D.g + iy) * A.nx + (D.g + ix);
        const size_t zs = (size_t)A.nx * A.ny;
        for (int iz = 0; iz < nz; ++iz) {
            const int t = t2 | ((D.dim == 3 && (iz & 1) == 0) ? 4 : 0);
            if (t != nt) typed_max(m, t, fabs(col[iz * zs]));
        }
    }
    for (int q = 0; q < nt; ++q) atomicMax(&smax[q], (unsigned long long)__double_as_longlong(m[q]));
    __syncthreads();
    if (threadIdx.x < nt) out[(size_t)blockIdx.x * nt + threadIdx.x] = __longlong_as_double((long long)smax[threadIdx.x]);
}

// Per box and component: max over the box of |v| per point type (the
// security zone's threshold_block with indices), out[(k*neq + c)*nt + t].
__global__ __launch_bounds__(256) void k_region_max_typed(const double *__restrict__ arr,
                                                          const wh_box *__restrict__ boxes, ArrGeo A, int g,
                                                          int dim, int neq, double *__restrict__ out)
{
    __shared__ unsigned long long smax[8 * 7];
    if (threadIdx.x < 8 * 7) smax[threadIdx.x] = 0ull;
    __syncthreads();
    const wh_box B = boxes[blockIdx.x];
    const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1, n2 = B.hi[2] - B.lo[2] + 1;
    const int npts = n0 * n1 * n2;
    const int nt = (1 << dim) - 1;
    for (int c = 0; c < neq; ++c) {
        double m[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const double *u = arr + (size_t)B.hvy * A.blk_in + c * A.cs;
        for (int e = threadIdx.x; e < npts; e += blockDim.x) {
            const int ix = B.lo[0] + e % n0, iy = B.lo[1] + (e / n0) % n1, iz = B.lo[2] + e / (n0 * n1);
            const int t = (((ix - g) & 1) == 0 ? 1 : 0) | (((iy - g) & 1) == 0 ? 2 : 0) |
                          ((dim == 3 && ((iz - g) & 1) == 0) ? 4 : 0);
            if (t != nt) typed_max(m, t, fabs(u[((size_t)iz * A.ny + iy) * A.nx + ix]));
        }
        for (int q = 0; q < nt; ++q) atomicMax(&smax[c * 7 + q], (unsigned long long)__double_as_longlong(m[q]));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < neq * nt; i += blockDim.x) {
        const int c = i / nt, q = i - c * nt;
        out[(size_t)blockIdx.x * neq * nt + i] = __longlong_as_double((long long)smax[c * 7 + q]);
    }
}

// Per box and component: the sum over the box of |v| (l1) or v*v (else),
// componentWiseNorm_tree's per-block sums; a fixed reduction order (each
// thread's strided partial sum, then a tree over the workgroup).
__global__ __launch_bounds__(256) void k_region_sum(const double *__restrict__ arr, const wh_box *__restrict__ boxes,
                                                    ArrGeo A, int neq, int l1, double *__restrict__ out)
{
    __shared__ double part[256];
    const wh_box B = boxes[blockIdx.x];
    const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1, n2 = B.hi[2] - B.lo[2] + 1;
    const int npts = n0 * n1 * n2;
    for (int c = 0; c < neq; ++c) {
        const double *u = arr + (size_t)B.hvy * A.blk_in + c * A.cs;
        double s = 0.0;
        for (int e = threadIdx.x; e < npts; e += blockDim.x) {
            const int ix = B.lo[0] + e % n0, iy = B.lo[1] + (e / n0) % n1, iz = B.lo[2] + e / (n0 * n1);
            const double v = u[((size_t)iz * A.ny + iy) * A.nx + ix];
            s = s + (l1 ? fabs(v) : v * v);
        }
        part[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * neq + c] = part[0];
        __syncthreads();
    }
}

// Per-box maxima (one workgroup per box): per component the max over the box
// of |v| (zero_sc: SC points count as 0; sqrt_groups: components of group
// >= 2 take sqrt(v*v)), as threshold_block with `indices` and the Linfty
// norm of componentWiseNorm_tree.
struct RegionArgs {
    int g, dim, neq, zero_sc, sqrt_groups;
    int groups[8];
};
__global__ __launch_bounds__(256) void k_region_max(const double *__restrict__ arr, const wh_box *__restrict__ boxes,
                                                    ArrGeo A, RegionArgs R, double *__restrict__ out)
{
    __shared__ unsigned long long smax[8];
    if (threadIdx.x < 8) smax[threadIdx.x] = 0ull;
    __syncthreads();
    const wh_box B = boxes[blockIdx.x];
    const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1, n2 = B.hi[2] - B.lo[2] + 1;
    const int npts = n0 * n1 * n2;
    for (int c = 0; c < R.neq; ++c) {
        double m = 0.0;
        const double *u = arr + (size_t)B.hvy * A.blk_in + c * A.cs;
        for (int e = threadIdx.x; e < npts; e += blockDim.x) {
            const int ix = B.lo[0] + e % n0, iy = B.lo[1] + (e / n0) % n1, iz = B.lo[2] + e / (n0 * n1);
            const bool sc = (((ix - R.g) & 1) == 0) && (((iy - R.g) & 1) == 0) &&
                            (R.dim == 2 || ((iz - R.g) & 1) == 0);
            const double v = (R.zero_sc && sc) ? 0.0 : u[((size_t)iz * A.ny + iy) * A.nx + ix];
            const double a = (R.sqrt_groups && R.groups[c] >= 2) ? sqrt(v * v) : fabs(v);
            m = a > m ? a : m;
        }
        atomicMax(&smax[c], (unsigned long long)__double_as_longlong(m));
    }
    __syncthreads();
    if (threadIdx.x < R.neq)
        out[(size_t)blockIdx.x * R.neq + threadIdx.x] = __longlong_as_double((long long)smax[threadIdx.x]);
}

// Coarse-extension boxes: one workgroup per box.
__global__ __launch_bounds__(256) void k_boxes(double *__restrict__ arr, const double *__restrict__ src,
                                               size_t blk_src, const wh_box *__restrict__ boxes,
                                               ArrGeo A, int g, int dim)
{
    const wh_box B = boxes[blockIdx.x];
    const int n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1, n2 = B.hi[2] - B.lo[2] + 1;
    const int npts = n0 * n1 * n2;
    for (int e = threadIdx.x; e < npts * A.neq; e += blockDim.x) {
        const int c = e / npts, r = e - c * npts;
        const int ix = B.lo[0] + r % n0, iy = B.lo[1] + (r / n0) % n1, iz = B.lo[2] + r / (n0 * n1);
        const bool sc = (((ix - g) & 1) == 0) && (((iy - g) & 1) == 0) && (dim == 2 || ((iz - g) & 1) == 0);
        const size_t q = c * A.cs + ((size_t)iz * A.ny + iy) * A.nx + ix;
        if (B.op == WH_BOX_ZERO_WC) {
            if (!sc) arr[(size_t)B.hvy * A.blk_out + q] = 0.0;
        } else if (sc) {
            arr[(size_t)B.hvy * A.blk_out + q] = src[(size_t)B.hvy * blk_src + q];
        }
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
template <class T>
int ensure(T *&ptr, size_t &cap, size_t n)
{
    if (n <= cap) return WH_OK;
    if (ptr) WH_HIP(hipFree(ptr));
    ptr = nullptr;
    size_t c = std::max<size_t>(n, 2 * cap);
    WH_HIP(hipMalloc(&ptr, sizeof(T) * c));
    cap = c;
    return WH_OK;
}

int upload_ids(wh_pool *P, const int32_t *ids, int n)
{
    if (int rc = stage_begin(P, sizeof(int32_t) * n)) return rc;
    P->ws.cur_ids = (const int32_t *)stage_put(P, ids, sizeof(int32_t) * n);
    return P->ws.cur_ids ? WH_OK : WH_E_HIP;
}

int check_ids(wh_pool *P, const int32_t *ids, int n)
{
    WH_CHECK(n >= 0 && (n == 0 || ids), WH_E_ARG, "block list: bad arguments");
    for (int k = 0; k < n; ++k)
        WH_CHECK(ids[k] >= 0 && ids[k] < P->d.nb_max, WH_E_ARG, "block list: hvy id out of range");
    return WH_OK;
}

int launch_pass(wh_pool *P, const double *in, size_t blk_in, double *out, size_t blk_out, int nb,
                const Region &R, int axis, int g, const TermTable &T)
{
    const int npts = R.n[0] * R.n[1] * R.n[2];
    if (npts <= 0 || nb <= 0) return WH_OK;
    const int wpb = (npts + 255) / 256;
    WH_CHECK((size_t)wpb * P->d.neq * nb < (1u << 31), WH_E_ARG, "wavelet pass: too many blocks in one launch");
    const ArrGeo A{blk_in, blk_out, P->cs, P->nx, P->ny, P->d.neq};
    const int astride = axis == 0 ? 1 : (axis == 1 ? P->nx : P->nx * P->ny);
    const dim3 grid((unsigned)((size_t)wpb * P->d.neq * nb));
    switch (chain_kind(T)) {
    case CHAIN_PLAIN:
        hipLaunchKernelGGL(k_wpass<CHAIN_PLAIN>, grid, dim3(256), 0, P->stream, in, out, P->ws.cur_ids, wpb, A, R,
                           axis, astride, g, T);
        break;
    case CHAIN_ZERO:
        hipLaunchKernelGGL(k_wpass<CHAIN_ZERO>, grid, dim3(256), 0, P->stream, in, out, P->ws.cur_ids, wpb, A, R,
                           axis, astride, g, T);
        break;
    default:
        hipLaunchKernelGGL(k_wpass<CHAIN_GENERAL>, grid, dim3(256), 0, P->stream, in, out, P->ws.cur_ids, wpb, A,
                           R, axis, astride, g, T);
    }
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int launch_copy_blocks(wh_pool *P, const double *in, size_t blk_in, double *out, size_t blk_out, int nb)
{
    if (nb <= 0) return WH_OK;
    const size_t len = P->cs * P->d.neq;
    const bool wide = blk_in % 2 == 0 && blk_out % 2 == 0 && len % 2 == 0;
    const size_t nw = wide ? len / 2 : len;
    WH_CHECK(nw < (1u << 30), WH_E_ARG, "block copy: blocks too large");
    const int wpb = (int)((nw + 511) / 512);
    WH_CHECK((size_t)wpb * nb < (1u << 31), WH_E_ARG, "block copy: too many blocks in one launch");
    const dim3 grid((unsigned)(wpb * nb));
    if (wide)
        hipLaunchKernelGGL(k_copy_blocks<double2>, grid, dim3(256), 0, P->stream, in, blk_in, out, blk_out,
                           P->ws.cur_ids, wpb, (int)nw);
    else
        hipLaunchKernelGGL(k_copy_blocks<double>, grid, dim3(256), 0, P->stream, in, blk_in, out, blk_out,
                           P->ws.cur_ids, wpb, (int)nw);
    WH_HIP(hipGetLastError());
    return WH_OK;
}


// in-place pass along y (axis 1) or z (axis 2) on `arr` over region R
int launch_wcol(wh_pool *P, double *arr, size_t blk, int nb, const Region &R, int axis, const TermTable &T)
{
    const int ncol = R.n[0] * (axis == 1 ? R.n[2] : R.n[1]);
    if (ncol <= 0 || nb <= 0) return WH_OK;
    const int len = axis == 1 ? P->ny : P->nz;
    const int astride = axis == 1 ? P->nx : P->nx * P->ny;
    // every point a chain reads lies in the column, and the chains' offsets
    // span at most the ring
    int lo_off = 0, hi_off = 0;
    for (int i = 0; i < R.n[axis]; ++i) {
        const int p = R.lo[axis] + R.step[axis] * i, par = (p - P->d.g) & 1;
        for (int k = 0; k < T.n[par]; ++k) {
            WH_CHECK(p + T.off[par][k] >= 0 && p + T.off[par][k] < len, WH_E_ARG,
                     "wavelet pass: filter support leaves the block");
            lo_off = std::min(lo_off, T.off[par][k]);
            hi_off = std::max(hi_off, T.off[par][k]);
        }
    }
    const int span = hi_off - lo_off + 1;
    WH_CHECK(span <= 32, WH_E_ARG, "wavelet pass: filter longer than the column ring");
    const int ring = span <= 16 ? 16 : 32, lanes = 4096 / ring;
    const int wpb = (ncol + lanes - 1) / lanes;
    WH_CHECK((size_t)wpb * P->d.neq * nb < (1u << 31), WH_E_ARG, "wavelet pass: too many blocks in one launch");
    const ArrGeo A{blk, blk, P->cs, P->nx, P->ny, P->d.neq};
    const dim3 grid((unsigned)((size_t)wpb * P->d.neq * nb));
#define LWR(K, N)                                                                                               \
    do {                                                                                                         \
        if (ring == 16)                                                                                          \
            hipLaunchKernelGGL((k_wring<K, N, 16>), grid, dim3(256), 0, P->stream, arr, P->ws.cur_ids, wpb, A, R, \
                               axis, len, astride, P->d.g, lo_off, hi_off, T);                                  \
        else                                                                                                     \
            hipLaunchKernelGGL((k_wring<K, N, 32>), grid, dim3(128), 0, P->stream, arr, P->ws.cur_ids, wpb, A, R, \
                               axis, len, astride, P->d.g, lo_off, hi_off, T);                                  \
    } while (0)
    CHAIN_DISPATCH(chain_var(T), LWR);
#undef LWR
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// the fused x/y passes, where two planes of a block fit in 64 KB of LDS
bool sepxy_ok(const wh_pool *P) { return 2 * sizeof(double) * P->nx * P->ny <= 64 * 1024; }

// reach of the chains along a pass direction from output position p
bool chain_in(const TermTable &T, int p, int g, int lo, int hi)
{
    const int par = (p - g) & 1;
    for (int k = 0; k < T.n[par]; ++k)
        if (p + T.off[par][k] < lo || p + T.off[par][k] > hi) return false;
    return true;
}

bool region_in(const Region &R, const int ext[3])
{
    for (int d = 0; d < 3; ++d)
        if (R.n[d] <= 0 || R.lo[d] < 0 || R.lo[d] + R.step[d] * (R.n[d] - 1) >= ext[d] ||
            (R.step[d] != 1 && R.step[d] != 2))
            return false;
    return true;
}

int launch_sepxy(wh_pool *P, const double *in, size_t blk_in, double *out, size_t blk_out, int nb,
                 const Region &Rx, const Region &Ry, int init_copy, const TermTable &T)
{
    if (nb <= 0) return WH_OK;
    const int g = P->d.g;
    const int ext[3] = {P->nx, P->ny, P->d.dim == 3 ? P->nz : 1};
    // what the kernel relies on: both regions inside the block, the same
    // planes, Ry's x positions are Rx's, the x chains stay in the row, the
    // y chains read rows of Rx
    bool ok = region_in(Rx, ext) && region_in(Ry, ext) && Rx.lo[2] == Ry.lo[2] && Rx.n[2] == Ry.n[2] &&
              Rx.step[2] == Ry.step[2] && Rx.lo[0] == Ry.lo[0] && Rx.n[0] == Ry.n[0] && Rx.step[0] == Ry.step[0];
    for (int i = 0; ok && i < Rx.n[0]; ++i) ok = chain_in(T, Rx.lo[0] + Rx.step[0] * i, g, 0, P->nx - 1);
    for (int i = 0; ok && i < Ry.n[1]; ++i) {
        const int p = Ry.lo[1] + Ry.step[1] * i, par = (p - g) & 1;
        for (int k = 0; ok && k < T.n[par]; ++k) {
            const int r = p + T.off[par][k] - Rx.lo[1];
            ok = r >= 0 && r % Rx.step[1] == 0 && r / Rx.step[1] < Rx.n[1];
        }
    }
    WH_CHECK(ok, WH_E_ARG, "wavelet pass: regions do not fit the fused x/y kernel");
    const int nzp = P->d.dim == 3 ? P->nz : 1;
    WH_CHECK((size_t)nb * P->d.neq * nzp < (1u << 31), WH_E_ARG, "wavelet pass: too many blocks in one launch");
    const ArrGeo A{blk_in, blk_out, P->cs, P->nx, P->ny, P->d.neq};
    const size_t lds = 2 * sizeof(double) * P->nx * P->ny;
    const ParSplit Px = par_split(Rx, 0, g), Py = par_split(Ry, 1, g);
    WH_CHECK(P->nx < (1 << 16) && P->nx * P->ny + 64 < (1 << 16), WH_E_ARG, "wavelet pass: plane too large");
    const SepDiv V{fast_div(Px.nA), fast_div(std::max(Px.nB, 1)), fast_div(Ry.n[0]), fast_div(Rx.n[0]),
                   fast_div(P->nx)};
    const dim3 grid((unsigned)(nb * P->d.neq * nzp));
#define LSX(K, N)                                                                                                 \
    hipLaunchKernelGGL((k_sepxy<K, N>), grid, dim3(256), lds, P->stream, in, out, P->ws.cur_ids, A, Rx, Ry, Px, \
                       Py, V, nzp, init_copy, T)
    CHAIN_DISPATCH(chain_var(T), LSX);
#undef LSX
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// Three separable passes (x over an extended row range, then y, then z) of
// the chains T: x reads `in` and writes `out`; y and z work in place on
// `out`, a column at a time.
int separable(wh_pool *P, const double *in, size_t blk_in, double *out, size_t blk_out, int nb,
              const TermTable &T, int ext, bool init_copy = false)
{
    const int g = P->d.g, dim = P->d.dim;
    const int B0 = P->d.bs[0], B1 = P->d.bs[1], B2 = dim == 3 ? P->d.bs[2] : 1;
    const int zlo = dim == 3 ? g - ext : 0, zn = dim == 3 ? B2 + 2 * ext : 1;
    Region Rx{{g, g - ext, zlo}, {B0, B1 + 2 * ext, zn}, {1, 1, 1}};
    Region Ry{{g, g, zlo}, {B0, B1, zn}, {1, 1, 1}};
    Region Rz{{g, g, g}, {B0, B1, B2}, {1, 1, 1}};
    if (sepxy_ok(P)) {
        if (int rc = launch_sepxy(P, in, blk_in, out, blk_out, nb, Rx, Ry, init_copy ? 1 : 0, T)) return rc;
        return dim == 3 ? launch_wcol(P, out, blk_out, nb, Rz, 2, T) : WH_OK;
    }
    if (init_copy)
        if (int rc = launch_copy_blocks(P, in, blk_in, out, blk_out, nb)) return rc;
    if (int rc = launch_pass(P, in, blk_in, out, blk_out, nb, Rx, 0, g, T)) return rc;
    if (int rc = launch_wcol(P, out, blk_out, nb, Ry, 1, T)) return rc;
    return dim == 3 ? launch_wcol(P, out, blk_out, nb, Rz, 2, T) : WH_OK;
}

// blockFilterXYZ_vct(HD, do_restriction): out = in, then the SC interior
// points are filtered along x (rows g+lo..g+Bs-1+hi), y, z.
int filter_passes(wh_pool *P, const double *in, size_t blk_in, double *out, size_t blk_out, int nb)
{
    const wh_wavelet &w = P->ws.w;
    const int g = P->d.g, dim = P->d.dim;
    const int B0 = P->d.bs[0], B1 = P->d.bs[1], B2 = dim == 3 ? P->d.bs[2] : 1;
    const bool trivial = w.hd_lo == 0 && w.hd_hi == 0 && std::fabs(w.HD[0] - 1.0) <= 1.0e-10;
    const TermTable T = filter_terms(w);
    const int lo = w.hd_lo, hi = w.hd_hi;
    const int zlo = dim == 3 ? g + lo : 0, zn = dim == 3 ? B2 - lo + hi : 1;
    // the parity of every filtered point is even, so chain 0 applies
    Region Rx{{g, g + lo, zlo}, {B0 / 2, B1 - lo + hi, zn}, {2, 1, 1}};
    Region Ry{{g, g, zlo}, {B0 / 2, B1 / 2, zn}, {2, 2, 1}};
    Region Rz{{g, g, g}, {B0 / 2, B1 / 2, B2 / 2}, {2, 2, 2}};
    if (!trivial && sepxy_ok(P)) {
        if (int rc = launch_sepxy(P, in, blk_in, out, blk_out, nb, Rx, Ry, 1, T)) return rc;
        return dim == 3 ? launch_wcol(P, out, blk_out, nb, Rz, 2, T) : WH_OK;
    }
    if (int rc = launch_copy_blocks(P, in, blk_in, out, blk_out, nb)) return rc;
    if (trivial) return WH_OK;
    if (int rc = launch_pass(P, in, blk_in, out, blk_out, nb, Rx, 0, g, T)) return rc;
    if (int rc = launch_wcol(P, out, blk_out, nb, Ry, 1, T)) return rc;
    return dim == 3 ? launch_wcol(P, out, blk_out, nb, Rz, 2, T) : WH_OK;
}

}  // namespace

static constexpr size_t STAGE_ALIGN = 256;

static size_t stage_round(size_t b) { return (b + STAGE_ALIGN - 1) / STAGE_ALIGN * STAGE_ALIGN; }

int stage_sync(wh_pool *P)
{
    WH_PROF("stage_sync");
    WH_HIP(hipStreamSynchronize(P->stream));
    P->ws.stage_off = 0;
    return WH_OK;
}

int stage_begin(wh_pool *P, size_t bytes)
{
    WaveletWS &W = P->ws;
    WH_CHECK(W.pend_lo == ~(size_t)0, WH_E_STATE, "staging: deferred pieces not flushed");
    const size_t need = stage_round(bytes) + 4 * STAGE_ALIGN;  // room for a few aligned pieces
    if (W.stage_off + need <= W.stage_cap) return WH_OK;
    // full: wait for the queued copies and kernels, then reuse the space; a
    // buffer that fills between the orchestration's own synchronisations is
    // doubled (up to 64 MB), so that this wait stops recurring
    const bool filled = W.stage_cap > 0;
    if (int rc = stage_sync(P)) return rc;
    const size_t grow = filled && W.stage_cap < ((size_t)64 << 20) ? 2 * W.stage_cap : 0;
    if (need <= W.stage_cap && !grow) return WH_OK;
    const size_t old_cap = W.stage_cap;
    if (W.h_stage) WH_HIP(hipHostFree(W.h_stage));
    if (W.d_stage) WH_HIP(hipFree(W.d_stage));
    W.h_stage = W.d_stage = nullptr;
    W.stage_cap = 0;
    const size_t cap = std::max<size_t>(need, std::max<size_t>(std::max(grow, 2 * old_cap), (size_t)1 << 20));
    WH_HIP(hipHostMalloc((void **)&W.h_stage, cap, hipHostMallocDefault));
    WH_HIP(hipMalloc((void **)&W.d_stage, cap));
    W.stage_cap = cap;
    return WH_OK;
}

int stage_flush(wh_pool *P)
{
    WaveletWS &W = P->ws;
    if (W.pend_lo == ~(size_t)0) return WH_OK;
    const size_t lo = W.pend_lo;
    W.pend_lo = ~(size_t)0;
    if (W.stage_off > lo)
        WH_HIP(hipMemcpyAsync(W.d_stage + lo, W.h_stage + lo, W.stage_off - lo, hipMemcpyHostToDevice, P->stream));
    return WH_OK;
}

void *stage_put(wh_pool *P, const void *src, size_t bytes, bool defer)
{
    WaveletWS &W = P->ws;
    if (!defer && stage_flush(P) != WH_OK) return nullptr;
    const size_t r = stage_round(bytes ? bytes : 1);
    if (W.stage_off + r > W.stage_cap) {
        set_error("staging: stage_put beyond the reserved space");
        return nullptr;
    }
    char *h = W.h_stage + W.stage_off, *d = W.d_stage + W.stage_off;
    if (bytes) {
        std::memcpy(h, src, bytes);
        if (defer) {
            if (W.pend_lo == ~(size_t)0) W.pend_lo = W.stage_off;
        } else if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, P->stream) != hipSuccess) {
            set_error("staging: hipMemcpyAsync failed");
            return nullptr;
        }
    }
    W.stage_off += r;
    return d;
}

void *stage_host(wh_pool *P, const void *src, size_t bytes)
{
    WaveletWS &W = P->ws;
    if (stage_flush(P) != WH_OK) return nullptr;
    const size_t r = stage_round(bytes ? bytes : 1);
    if (W.stage_off + r > W.stage_cap) {
        set_error("staging: stage_host beyond the reserved space");
        return nullptr;
    }
    char *h = W.h_stage + W.stage_off;
    if (bytes) std::memcpy(h, src, bytes);
    W.stage_off += r;
    return h;
}

void free_wavelet_ws(wh_pool *P)
{
    WaveletWS &W = P->ws;
    if (W.h_stage) (void)hipHostFree(W.h_stage);
    if (W.d_stage) (void)hipFree(W.d_stage);
    if (W.d_ids) (void)hipFree(W.d_ids);
    if (W.d_patches) (void)hipFree(W.d_patches);
    if (W.d_boxes) (void)hipFree(W.d_boxes);
    if (W.d_out) (void)hipFree(W.d_out);
    W = WaveletWS{};
}

int run_blocks_op(wh_pool *P, int op, int src_array, int dst_array, const std::vector<int32_t> &ids)
{
    WH_PROF("run_blocks_op");
    const int n = (int)ids.size();
    if (n == 0) return WH_OK;
    size_t blk_in = 0, blk_out = 0;
    int nc_in = 0, nc_out = 0;
    const double *in = array_ptr(P, src_array, &blk_in, &nc_in);
    double *out = array_ptr(P, dst_array, &blk_out, &nc_out);
    WH_CHECK(in && out, WH_E_ARG, "wavelet op: unknown array");
    WH_CHECK(src_array != dst_array, WH_E_ARG, "wavelet op: source and destination must differ");
    WH_CHECK(nc_in >= P->d.neq && nc_out >= P->d.neq, WH_E_ARG,
             "wavelet op: arrays need neq components");
    WH_CHECK(op == WOP_COPY || P->ws.set, WH_E_STATE, "wavelet op: no wavelet set on the pool");
    WH_CHECK(op == WOP_COPY || (P->d.n_work_slots >= 1 && src_array != WH_WORK && dst_array != WH_WORK),
             WH_E_ARG, "wavelet op: work slot 0 is scratch and needed");
    if (int rc = check_ids(P, ids.data(), n)) return rc;
    if (int rc = upload_ids(P, ids.data(), n)) return rc;
    const wh_wavelet &w = P->ws.w;
    WH_CHECK(op == WOP_COPY || P->d.g >= w.g, WH_E_ARG, "wavelet op: too few ghost nodes for the wavelet");
    if (op == WOP_COPY) return launch_copy_blocks(P, in, blk_in, out, blk_out, n);
    if (op == WOP_FWT)
        return separable(P, in, blk_in, out, blk_out, n, fwt_terms(w), std::max(w.hd_hi, w.gd_hi));
    if (op == WOP_IWT)  // CDFX0 starts from a copy of the whole block
        return separable(P, in, blk_in, out, blk_out, n, iwt_terms(w), std::max(w.hr_hi, w.gr_hi), !w.lifted);
    if (op == WOP_FILTER) return filter_passes(P, in, blk_in, out, blk_out, n);
    set_error("wavelet op: unknown op");
    return WH_E_ARG;
}

int run_boxes(wh_pool *P, int array, int src_array, const std::vector<wh_box> &boxes)
{
    WH_PROF("run_boxes");
    const int n = (int)boxes.size();
    if (n == 0) return WH_OK;
    size_t blk = 0, blk_src = 0;
    int nc = 0, nc_src = 0;
    double *arr = array_ptr(P, array, &blk, &nc);
    const double *src = array_ptr(P, src_array, &blk_src, &nc_src);
    WH_CHECK(arr && src && nc >= P->d.neq && nc_src >= P->d.neq, WH_E_ARG, "boxes: bad arrays");
    const int lim[3] = {P->nx, P->ny, P->nz};
    for (const wh_box &b : boxes) {
        WH_CHECK(b.hvy >= 0 && b.hvy < P->d.nb_max && (b.op == WH_BOX_ZERO_WC || b.op == WH_BOX_COPY_SC),
                 WH_E_ARG, "boxes: bad hvy id or op");
        for (int d = 0; d < 3; ++d)
            WH_CHECK(b.lo[d] >= 0 && b.hi[d] < lim[d] && b.lo[d] <= b.hi[d], WH_E_ARG,
                     "boxes: box outside the block");
    }
    if (int rc = stage_begin(P, sizeof(wh_box) * n)) return rc;
    const wh_box *d_boxes = (const wh_box *)stage_put(P, boxes.data(), sizeof(wh_box) * n);
    if (!d_boxes) return WH_E_HIP;
    const ArrGeo A{blk_src, blk, P->cs, P->nx, P->ny, P->d.neq};
    hipLaunchKernelGGL(k_boxes, dim3(n), dim3(256), 0, P->stream, arr, src, blk_src, d_boxes, A,
                       P->d.g, P->d.dim);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

// k_detail over n listed blocks whose ids are at d_ids on the device, into
// d_out[k*neq + c] (every entry written: nothing needs clearing).
int launch_detail(wh_pool *P, const double *arr, size_t blk, const int32_t *d_ids, int n, const DetailArgs &D,
                  double *d_out)
{
    const ArrGeo A{blk, blk, P->cs, P->nx, P->ny, D.neq};
    const int nitems = n * D.neq;
    if (D.dim == 2)
        hipLaunchKernelGGL(k_detail_wave, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, P->stream, arr, d_ids, A,
                           D, nitems, d_out);
    else
        hipLaunchKernelGGL(k_detail, dim3((unsigned)nitems), dim3(256), 0, P->stream, arr, d_ids, A, D, d_out);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

DetailArgs detail_args(const wh_pool *P, const int32_t *groups)
{
    DetailArgs D{};
    D.g = P->d.g;
    D.dim = P->d.dim;
    D.neq = P->d.neq;
    for (int d = 0; d < 3; ++d) D.bs[d] = d < P->d.dim ? P->d.bs[d] : 1;
    for (int c = 0; c < 8; ++c) D.groups[c] = (groups && c < D.neq) ? groups[c] : 1;
    return D;
}

// groups: the max over the group's components (threshold_block.f90:86-91),
// group 0 gives 0; in place on the host
void apply_detail_groups(double *detail, int n, const DetailArgs &D)
{
    const int neq = D.neq;
    int maxg = 0;
    for (int c = 0; c < neq; ++c) maxg = std::max(maxg, D.groups[c]);
    for (int k = 0; k < n; ++k) {
        double *dk = detail + (size_t)k * neq;
        for (int l = 2; l <= maxg; ++l) {
            double m = -1.0;
            for (int c = 0; c < neq; ++c)
                if (D.groups[c] == l) m = std::max(m, dk[c]);
            for (int c = 0; c < neq; ++c)
                if (D.groups[c] == l) dk[c] = m;
        }
        for (int c = 0; c < neq; ++c)
            if (D.groups[c] == 0) dk[c] = 0.0;
    }
}

void detail_groups(const wh_pool *P, const int32_t *groups, double *detail, int n)
{
    apply_detail_groups(detail, n, detail_args(P, groups));
}

int run_detail(wh_pool *P, int array, const std::vector<int32_t> &ids, const int32_t *groups,
               std::vector<double> &detail)
{
    WH_PROF("run_detail");
    const int n = (int)ids.size(), neq = P->d.neq;
    detail.assign((size_t)n * neq, 0.0);
    if (n == 0) return WH_OK;
    WaveletWS &W = P->ws;
    if (int rc = ensure(W.d_out, W.out_cap, (size_t)n * neq)) return rc;
    if (int rc = launch_detail_list(P, array, ids, groups, W.d_out)) return rc;
    WH_HIP(hipMemcpyAsync(detail.data(), W.d_out, sizeof(double) * n * neq, hipMemcpyDeviceToHost,
                          P->stream));
    if (int rc = stage_sync(P)) return rc;
    apply_detail_groups(detail.data(), n, detail_args(P, groups));
    return WH_OK;
}

// checks the boxes against the pool and stages them on the device
static int stage_boxes(wh_pool *P, int array, const std::vector<wh_box> &boxes, const double **arr, size_t *blk,
                       const wh_box **d_boxes)
{
    const int n = (int)boxes.size(), neq = P->d.neq;
    WH_CHECK(neq <= 8, WH_E_ARG, "region reduction: at most 8 components");
    int nc = 0;
    *arr = array_ptr(P, array, blk, &nc);
    WH_CHECK(*arr && nc >= neq, WH_E_ARG, "region reduction: bad array");
    const int lim[3] = {P->nx, P->ny, P->nz};
    for (const wh_box &b : boxes) {
        WH_CHECK(b.hvy >= 0 && b.hvy < P->d.nb_max, WH_E_ARG, "region reduction: bad hvy id");
        for (int d = 0; d < 3; ++d)
            WH_CHECK(b.lo[d] >= 0 && b.hi[d] < lim[d] && b.lo[d] <= b.hi[d], WH_E_ARG,
                     "region reduction: bad box");
    }
    if (int rc = stage_begin(P, sizeof(wh_box) * n)) return rc;
    *d_boxes = (const wh_box *)stage_put(P, boxes.data(), sizeof(wh_box) * n);
    return *d_boxes ? (int)WH_OK : (int)WH_E_HIP;
}

int launch_region_norm(wh_pool *P, int array, const std::vector<wh_box> &boxes, int kind, bool zone,
                       const int32_t *groups, double *d_out)
{
    if (kind == WH_NORM_LINFTY) return launch_region_max(P, array, boxes, zone, zone, groups, d_out);
    const int n = (int)boxes.size(), neq = P->d.neq;
    if (n == 0) return WH_OK;
    const double *arr = nullptr;
    size_t blk = 0;
    const wh_box *d_boxes = nullptr;
    if (int rc = stage_boxes(P, array, boxes, &arr, &blk, &d_boxes)) return rc;
    const ArrGeo A{blk, blk, P->cs, P->nx, P->ny, neq};
    if (zone)
        hipLaunchKernelGGL(k_region_max_typed, dim3(n), dim3(256), 0, P->stream, arr, d_boxes, A, P->d.g, P->d.dim,
                           neq, d_out);
    else
        hipLaunchKernelGGL(k_region_sum, dim3(n), dim3(256), 0, P->stream, arr, d_boxes, A, neq,
                           kind == WH_NORM_L1 ? 1 : 0, d_out);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

double renorm_max(double m, int t, int level, int dim, int kind)
{
    const int jref = 0;
    double r = m;
    if (kind == WH_NORM_L1) {
        r = r * std::pow(2.0, (double)((jref - level - 1) * dim));
        for (int d = 0; d < dim; ++d)
            if ((t >> d) & 1) r = r / 4.0;
    } else if (kind == WH_NORM_L2) {
        r = r * std::pow(2.0, (double)((jref - level - 1) * dim) / 2.0);
        for (int d = 0; d < dim; ++d)
            if ((t >> d) & 1) r = r / 2.0;
    } else if (kind == WH_NORM_H1 && dim == 3) {
        r = r * std::pow(2.0, (double)(jref - level) * (2.0 - (double)dim) / 2.0);
        const double f = std::pow(2.0, 2.0 * ((double)dim - 2.0) / 3.0);
        for (int d = 0; d < dim; ++d)
            if ((t >> d) & 1) r = r * f;
    }
    return r;
}

void renorm_typed(const double *typed, int n, int neq, int dim, int kind, const int *levels,
                  const int32_t *groups, double *out)
{
    const int nt = (1 << dim) - 1;
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < neq; ++c) {
            double r = 0.0;
            for (int t = 0; t < nt; ++t) r = std::max(r, renorm_max(typed[((size_t)k * neq + c) * nt + t], t, levels[k], dim, kind));
            out[(size_t)k * neq + c] = (groups && groups[c] >= 2) ? std::sqrt(r * r) : r;
        }
}

void accumulate_norm(const double *v, int n, int neq, const int32_t *groups, const int *levels, int dim,
                     const int bs[3], const double domain[3], int kind, double *norm)
{
    const bool inf = kind == WH_NORM_LINFTY;
    int maxg = 0;
    for (int c = 0; c < neq; ++c) maxg = std::max(maxg, (int)groups[c]);
    for (int c = 0; c < neq; ++c) norm[c] = inf ? -1.0 : 0.0;
    for (int k = 0; k < n; ++k) {
        const double *vk = v + (size_t)k * neq;
        double vol = 1.0;
        for (int d = 0; d < dim; ++d) vol = vol * (std::ldexp(1.0, -levels[k]) * domain[d] / (double)bs[d]);
        for (int c = 0; c < neq; ++c)
            if (groups[c] == 1) norm[c] = inf ? std::max(norm[c], vk[c]) : norm[c] + vol * vk[c];
        for (int l = 2; l <= maxg; ++l) {
            double nb = 0.0;
            for (int c = 0; c < neq; ++c)
                if (groups[c] == l) nb = inf ? std::max(nb, vk[c]) : nb + vk[c];
            for (int c = 0; c < neq; ++c)
                if (groups[c] == l) norm[c] = inf ? std::max(norm[c], nb) : norm[c] + vol * nb;
        }
    }
    if (kind == WH_NORM_L2 || kind == WH_NORM_H1)
        for (int c = 0; c < neq; ++c) norm[c] = std::sqrt(norm[c]);
}

int launch_region_max(wh_pool *P, int array, const std::vector<wh_box> &boxes, bool zero_sc, bool sqrt_groups,
                      const int32_t *groups, double *d_out)
{
    const int n = (int)boxes.size(), neq = P->d.neq;
    if (n == 0) return WH_OK;
    const double *arr = nullptr;
    size_t blk = 0;
    const wh_box *d_boxes = nullptr;
    if (int rc = stage_boxes(P, array, boxes, &arr, &blk, &d_boxes)) return rc;
    RegionArgs R{};
    R.g = P->d.g;
    R.dim = P->d.dim;
    R.neq = neq;
    R.zero_sc = zero_sc;
    R.sqrt_groups = sqrt_groups;
    for (int c = 0; c < 8; ++c) R.groups[c] = (groups && c < neq) ? groups[c] : 1;
    const ArrGeo A{blk, blk, P->cs, P->nx, P->ny, neq};
    hipLaunchKernelGGL(k_region_max, dim3(n), dim3(256), 0, P->stream, arr, d_boxes, A, R, d_out);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int run_region_max(wh_pool *P, int array, const std::vector<wh_box> &boxes, bool zero_sc, bool sqrt_groups,
                   const int32_t *groups, std::vector<double> &out)
{
    const int n = (int)boxes.size(), neq = P->d.neq;
    out.assign((size_t)n * neq, 0.0);
    if (n == 0) return WH_OK;
    WaveletWS &W = P->ws;
    if (int rc = ensure(W.d_out, W.out_cap, (size_t)n * neq)) return rc;
    if (int rc = launch_region_max(P, array, boxes, zero_sc, sqrt_groups, groups, W.d_out)) return rc;
    WH_HIP(hipMemcpyAsync(out.data(), W.d_out, sizeof(double) * n * neq, hipMemcpyDeviceToHost, P->stream));
    return stage_sync(P);
}

int results_reserve(wh_pool *P, size_t n, double **d_out)
{
    WaveletWS &W = P->ws;
    if (int rc = ensure(W.d_out, W.out_cap, std::max<size_t>(n, 1))) return rc;
    *d_out = W.d_out;
    return WH_OK;
}

int results_fetch(wh_pool *P, size_t n, std::vector<double> &out)
{
    out.assign(n, 0.0);
    if (n) WH_HIP(hipMemcpyAsync(out.data(), P->ws.d_out, sizeof(double) * n, hipMemcpyDeviceToHost, P->stream));
    return stage_sync(P);
}

int launch_detail_list(wh_pool *P, int array, const std::vector<int32_t> &ids, const int32_t *groups,
                       double *d_out, bool typed)
{
    const int n = (int)ids.size(), neq = P->d.neq;
    if (n == 0) return WH_OK;
    WH_CHECK(neq <= 8, WH_E_ARG, "detail: at most 8 components");
    size_t blk = 0;
    int nc = 0;
    const double *arr = array_ptr(P, array, &blk, &nc);
    WH_CHECK(arr && nc >= neq, WH_E_ARG, "detail: bad array");
    if (int rc = check_ids(P, ids.data(), n)) return rc;
    if (int rc = upload_ids(P, ids.data(), n)) return rc;
    if (!typed) return launch_detail(P, arr, blk, P->ws.cur_ids, n, detail_args(P, groups), d_out);
    const ArrGeo A{blk, blk, P->cs, P->nx, P->ny, neq};
    hipLaunchKernelGGL(k_detail_typed, dim3((unsigned)(n * neq)), dim3(256), 0, P->stream, arr,
                       (const int32_t *)P->ws.cur_ids, A, detail_args(P, groups), d_out);
    WH_HIP(hipGetLastError());
    return WH_OK;
}

int run_patch_list(wh_pool *P, int array, int alt1, int alt2, const std::vector<wh_patch> &patches_in)
{
    WH_PROF("run_patch_list");
    const int n = (int)patches_in.size();
    if (n == 0) return WH_OK;
    size_t blk = 0;
    int nc = 0;
    double *arr = array_ptr(P, array, &blk, &nc);
    WH_CHECK(arr && nc >= P->d.neq && blk == P->blk, WH_E_ARG, "patch list: bad target array");
    SyncSrc S;
    const int alts[2] = {alt1, alt2};
    for (int a = 0; a < 2; ++a) {
        if (alts[a] < 0) continue;
        int nca = 0;
        S.alt[a] = array_ptr(P, alts[a], &S.blk_alt[a], &nca);
        WH_CHECK(S.alt[a] && nca >= P->d.neq, WH_E_ARG, "patch list: bad alternate array");
    }
    // validate and order by stage (lists that are already ordered are used as they are)
    const int lim[3] = {P->nx, P->ny, P->nz};
    const auto by_stage = [](const wh_patch &a, const wh_patch &b) { return a.stage < b.stage; };
    std::vector<wh_patch> sorted_copy;
    const bool ordered = std::is_sorted(patches_in.begin(), patches_in.end(), by_stage);
    if (!ordered) {
        sorted_copy = patches_in;
        std::stable_sort(sorted_copy.begin(), sorted_copy.end(), by_stage);
    }
    const std::vector<wh_patch> &p = ordered ? patches_in : sorted_copy;
    for (const wh_patch &q : p) {
        const int op = q.op & 0xff, sel = (q.op >> 8) & 3;
        WH_CHECK(q.recv >= 0 && q.recv < P->d.nb_max && q.send >= 0 && q.send < P->d.nb_max, WH_E_ARG,
                 "patch list: hvy id out of range");
        WH_CHECK(op <= WH_PATCH_PREDICT && q.stage >= 1 && q.stage <= 3 && sel <= 2 &&
                     (sel == 0 || S.alt[sel - 1]),
                 WH_E_ARG, "patch list: bad op/stage/source");
        for (int d = 0; d < 3; ++d) {
            WH_CHECK(q.n[d] >= 1 && q.recv_lo[d] >= 0 && q.recv_lo[d] + q.n[d] <= lim[d], WH_E_ARG,
                     "patch list: receiver box outside the block");
            if (op == WH_PATCH_PREDICT) continue;
            const int st = op == WH_PATCH_RESTRICT ? 2 : 1;
            WH_CHECK(q.send_lo[d] >= 0 && q.send_lo[d] + st * (q.n[d] - 1) < lim[d], WH_E_ARG,
                     "patch list: sender box outside the block");
        }
        if (op == WH_PATCH_PREDICT) {
            WH_CHECK(q.aux == 2 || q.aux == 4 || q.aux == 6, WH_E_ARG, "patch list: predictor order");
            const int N = q.aux / 2;
            for (int d = 0; d < P->d.dim; ++d) {
                const int fmin = q.send_lo[d], fmax = q.send_lo[d] + q.n[d] - 1;
                WH_CHECK(fmin >= 0 && (fmax >> 1) < lim[d], WH_E_ARG, "patch list: prediction box");
                const int omin = (fmin & 1) ? fmin : fmin + 1, omax = (fmax & 1) ? fmax : fmax - 1;
                WH_CHECK(omin > omax || ((omin >> 1) - N + 1 >= 0 && (omax >> 1) + N < lim[d]), WH_E_ARG,
                         "patch list: prediction stencil leaves the coarse block");
            }
        }
    }
    // 3D: the patches cut into wavefront-sized chunks (cstart[i]: patch i's first)
    std::vector<SyncChunk> chunks;
    std::vector<int> cstart(n + 1, 0);
    if (P->d.dim == 3)
        for (int i = 0; i < n; ++i) {
            add_chunks(chunks, i, p[i]);
            cstart[i + 1] = (int)chunks.size();
        }
    if (int rc = stage_begin(P, sizeof(wh_patch) * n + sizeof(SyncChunk) * chunks.size())) return rc;
    const wh_patch *d_patches = (const wh_patch *)stage_put(P, p.data(), sizeof(wh_patch) * n, true);
    const SyncChunk *d_chunks =
        chunks.empty() ? nullptr : (const SyncChunk *)stage_put(P, chunks.data(), sizeof(SyncChunk) * chunks.size(), true);
    if (int rc = stage_flush(P)) return rc;
    if (!d_patches || (!chunks.empty() && !d_chunks)) return WH_E_HIP;
    int s0 = 0;
    while (s0 < n) {
        int s1 = s0;
        while (s1 < n && p[s1].stage == p[s0].stage) ++s1;
        if (int rc = launch_sync_flat(P, arr, d_patches, d_chunks, s0, s1, cstart[s0], cstart[s1], S)) return rc;
        s0 = s1;
    }
    return WH_OK;
}

}  // namespace wh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Device scratch of wh_fwt_threshold_noalloc: the listed block ids, then the
// per-block details (k_detail writes every entry it later copies, so nothing
// needs clearing between calls).  One copy per GPU, held until the process
// ends.  In host code `scratch` only names the symbol; its device address on
// each GPU comes from hipGetSymbolAddress (noalloc_setup) and is the only
// form host code passes on.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_IDS_BYTES = ((size_t)WH_NOALLOC_MAX_BLOCKS * sizeof(int32_t) + 255) / 256 * 256;
constexpr size_t NOALLOC_DETAIL_BYTES = (size_t)WH_NOALLOC_MAX_BLOCKS * 8 * sizeof(double);
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_IDS_BYTES + NOALLOC_DETAIL_BYTES];

namespace wh {
namespace {
std::mutex noalloc_mutex;
bool noalloc_done[NOALLOC_MAX_DEVICES];
hipError_t noalloc_err[NOALLOC_MAX_DEVICES];
unsigned char *noalloc_ptr[NOALLOC_MAX_DEVICES];
}  // namespace

void noalloc_setup(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_done[dev]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    noalloc_err[dev] = e;
    noalloc_ptr[dev] = e == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
    noalloc_done[dev] = true;
}

namespace {
int noalloc_scratch(int dev, unsigned char **out)
{
    WH_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, WH_E_ARG, "no-alloc scratch: device index out of range");
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    WH_CHECK(noalloc_done[dev], WH_E_STATE, "no-alloc scratch: not set up on this device (create a pool on it)");
    if (noalloc_err[dev] != hipSuccess) {
        set_error(std::string("no-alloc scratch: hipGetSymbolAddress: ") + hipGetErrorString(noalloc_err[dev]));
        return WH_E_HIP;
    }
    *out = noalloc_ptr[dev];
    return WH_OK;
}
}  // namespace
}  // namespace wh

extern "C" {

int wh_pool_set_wavelet(wh_pool *pool, const wh_wavelet *w)
{
    WH_CHECK(pool && w, WH_E_ARG, "wh_pool_set_wavelet: null argument");
    WH_CHECK(w->hd_lo >= -16 && w->hd_hi <= 16 && w->gd_lo >= -16 && w->gd_hi <= 16 &&
                 w->hr_lo >= -16 && w->hr_hi <= 16 && w->gr_lo >= -16 && w->gr_hi <= 16 &&
                 w->hd_hi - w->hd_lo + 1 <= 23 && w->gr_hi - w->gr_lo + 1 <= 23,
             WH_E_ARG, "wh_pool_set_wavelet: filter support too large");
    WH_CHECK(pool->d.g >= w->g, WH_E_ARG, "wh_pool_set_wavelet: the pool has fewer ghost nodes than the wavelet needs");
    pool->ws.w = *w;
    pool->ws.set = true;
    return WH_OK;
}

int wh_fwt_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || ids), WH_E_ARG, "wh_fwt_blocks: bad argument");
    return wh::run_blocks_op(pool, wh::WOP_FWT, src_array, dst_array, std::vector<int32_t>(ids, ids + n));
}

int wh_iwt_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || ids), WH_E_ARG, "wh_iwt_blocks: bad argument");
    return wh::run_blocks_op(pool, wh::WOP_IWT, src_array, dst_array, std::vector<int32_t>(ids, ids + n));
}

int wh_filter_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || ids), WH_E_ARG, "wh_filter_blocks: bad argument");
    return wh::run_blocks_op(pool, wh::WOP_FILTER, src_array, dst_array,
                             std::vector<int32_t>(ids, ids + n));
}

int wh_copy_blocks(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || ids), WH_E_ARG, "wh_copy_blocks: bad argument");
    return wh::run_blocks_op(pool, wh::WOP_COPY, src_array, dst_array, std::vector<int32_t>(ids, ids + n));
}

int wh_detail_blocks_norm(wh_pool *pool, int array, const int32_t *ids, const int32_t *levels, int n,
                          const int32_t *groups, int eps_norm, double *detail)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || (ids && levels && detail)), WH_E_ARG,
             "wh_detail_blocks_norm: bad argument");
    WH_CHECK(eps_norm >= WH_NORM_LINFTY && eps_norm <= WH_NORM_H1, WH_E_ARG, "wh_detail_blocks_norm: unknown norm");
    wh_pool *P = pool;
    const int neq = P->d.neq, dim = P->d.dim, nt = (1 << dim) - 1;
    if (eps_norm == WH_NORM_LINFTY) return wh_detail_blocks(pool, array, ids, n, groups, detail);
    if (n == 0) return WH_OK;
    WH_CHECK(neq <= 8, WH_E_ARG, "wh_detail_blocks_norm: at most 8 components");
    std::vector<double> typed;
    double *d_out = nullptr;
    if (int rc = wh::results_reserve(P, (size_t)n * neq * nt, &d_out)) return rc;
    if (int rc = wh::launch_detail_list(P, array, std::vector<int32_t>(ids, ids + n), groups, d_out, true)) return rc;
    if (int rc = wh::results_fetch(P, (size_t)n * neq * nt, typed)) return rc;
    std::vector<int> lv(levels, levels + n);
    wh::renorm_typed(typed.data(), n, neq, dim, eps_norm, lv.data(), groups, detail);
    wh::detail_groups(P, groups, detail, n);
    return WH_OK;
}

int wh_component_norm(wh_pool *pool, int array, const int32_t *ids, const int32_t *levels, int n,
                      const int32_t *groups, int eps_norm, const double *domain, double *norm)
{
    WH_CHECK(pool && n >= 0 && norm && domain && (n == 0 || (ids && levels)), WH_E_ARG,
             "wh_component_norm: bad argument");
    WH_CHECK(eps_norm >= WH_NORM_LINFTY && eps_norm <= WH_NORM_H1, WH_E_ARG, "wh_component_norm: unknown norm");
    wh_pool *P = pool;
    const int neq = P->d.neq, dim = P->d.dim;
    WH_CHECK(neq <= 8, WH_E_ARG, "wh_component_norm: at most 8 components");
    int32_t gr[8];
    for (int c = 0; c < 8; ++c) gr[c] = (groups && c < neq) ? groups[c] : 1;
    std::vector<wh_box> boxes((size_t)n);
    for (int k = 0; k < n; ++k) {
        wh_box &b = boxes[k];
        b = wh_box{};
        b.hvy = ids[k];
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = d < dim ? P->d.g : 0;
            b.hi[d] = d < dim ? P->d.g + P->d.bs[d] - 1 : 0;
        }
    }
    std::vector<double> v;
    if (n > 0) {
        double *d_out = nullptr;
        if (int rc = wh::results_reserve(P, (size_t)n * neq, &d_out)) return rc;
        if (int rc = wh::launch_region_norm(P, array, boxes, eps_norm, false, gr, d_out)) return rc;
        if (int rc = wh::results_fetch(P, (size_t)n * neq, v)) return rc;
    }
    std::vector<int> lv(levels, levels + n);
    wh::accumulate_norm(v.data(), n, neq, gr, lv.data(), dim, P->d.bs, domain, eps_norm, norm);
    return WH_OK;
}

int wh_detail_blocks(wh_pool *pool, int array, const int32_t *ids, int n, const int32_t *groups,
                     double *detail)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || (ids && detail)), WH_E_ARG, "wh_detail_blocks: bad argument");
    std::vector<double> d;
    if (int rc = wh::run_detail(pool, array, std::vector<int32_t>(ids, ids + n), groups, d)) return rc;
    std::copy(d.begin(), d.end(), detail);
    return WH_OK;
}

int wh_fwt_threshold(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n, const double *thresh,
                     const int32_t *groups, int32_t *flags_out, double *detail_out)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || (ids && thresh && flags_out && detail_out)), WH_E_ARG,
             "wh_fwt_threshold: bad argument");
    const std::vector<int32_t> list(ids, ids + n);
    if (int rc = wh::run_blocks_op(pool, wh::WOP_FWT, src_array, dst_array, list)) return rc;
    std::vector<double> d;
    if (int rc = wh::run_detail(pool, dst_array, list, groups, d)) return rc;
    const int neq = pool->d.neq;
    for (int k = 0; k < n; ++k) {
        bool all = true;  // threshold_block.f90:119-124
        for (int c = 0; c < neq; ++c)
            if (!(d[(size_t)k * neq + c] <= thresh[c])) all = false;
        flags_out[k] = all ? -1 : 0;
    }
    std::copy(d.begin(), d.end(), detail_out);
    return WH_OK;
}

int wh_apply_boxes(wh_pool *pool, int array, int src_array, const wh_box *boxes, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || boxes), WH_E_ARG, "wh_apply_boxes: bad argument");
    return wh::run_boxes(pool, array, src_array, std::vector<wh_box>(boxes, boxes + n));
}

int wh_run_patches(wh_pool *pool, int array, int alt1, int alt2, const wh_patch *patches, int n)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || patches), WH_E_ARG, "wh_run_patches: bad argument");
    return wh::run_patch_list(pool, array, alt1, alt2, std::vector<wh_patch>(patches, patches + n));
}

int wh_fwt_threshold_noalloc(wh_pool *pool, int src_array, int dst_array, const int32_t *ids, int n,
                             const double *thresh, const int32_t *groups, int32_t *flags_out, double *detail_out)
{
    WH_CHECK(pool && n >= 0 && (n == 0 || (ids && thresh && flags_out && detail_out)), WH_E_ARG,
             "wh_fwt_threshold_noalloc: bad argument");
    WH_CHECK(n <= WH_NOALLOC_MAX_BLOCKS, WH_E_ARG,
             "wh_fwt_threshold_noalloc: more blocks than WH_NOALLOC_MAX_BLOCKS (16384) in one call");
    wh_pool *P = pool;
    const int neq = P->d.neq;
    WH_CHECK(neq <= 8, WH_E_ARG, "wh_fwt_threshold_noalloc: at most 8 components");
    unsigned char *base = nullptr;
    if (int rc = wh::noalloc_scratch(P->device, &base)) return rc;
    if (n == 0) return WH_OK;
    // the checks of run_blocks_op (FWT) and run_detail
    size_t blk_in = 0, blk_out = 0;
    int nc_in = 0, nc_out = 0;
    const double *in = wh::array_ptr(P, src_array, &blk_in, &nc_in);
    double *out = wh::array_ptr(P, dst_array, &blk_out, &nc_out);
    WH_CHECK(in && out, WH_E_ARG, "wh_fwt_threshold_noalloc: unknown array");
    WH_CHECK(src_array != dst_array, WH_E_ARG, "wh_fwt_threshold_noalloc: source and destination must differ");
    WH_CHECK(nc_in >= neq && nc_out >= neq, WH_E_ARG, "wh_fwt_threshold_noalloc: arrays need neq components");
    WH_CHECK(P->ws.set, WH_E_STATE, "wh_fwt_threshold_noalloc: no wavelet set on the pool");
    WH_CHECK(P->d.n_work_slots >= 1 && src_array != WH_WORK && dst_array != WH_WORK, WH_E_ARG,
             "wh_fwt_threshold_noalloc: work slot 0 is scratch and needed");
    const wh_wavelet &w = P->ws.w;
    WH_CHECK(P->d.g >= w.g, WH_E_ARG, "wh_fwt_threshold_noalloc: too few ghost nodes for the wavelet");
    if (int rc = wh::check_ids(P, ids, n)) return rc;
    if (int rc = wh::stage_flush(P)) return rc;
    int32_t *d_ids = reinterpret_cast<int32_t *>(base);
    double *d_det = reinterpret_cast<double *>(base + NOALLOC_IDS_BYTES);
    WH_HIP(hipMemcpyAsync(d_ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, P->stream));
    const int32_t *saved = P->ws.cur_ids;
    P->ws.cur_ids = d_ids;
    const wh::DetailArgs D = wh::detail_args(P, groups);
    int rc = wh::separable(P, in, blk_in, out, blk_out, n, wh::fwt_terms(w), std::max(w.hd_hi, w.gd_hi));
    if (!rc) rc = wh::launch_detail(P, out, blk_out, d_ids, n, D, d_det);
    P->ws.cur_ids = saved;
    if (rc) return rc;
    WH_HIP(hipMemcpyAsync(detail_out, d_det, sizeof(double) * n * neq, hipMemcpyDeviceToHost, P->stream));
    WH_HIP(hipStreamSynchronize(P->stream));
    wh::apply_detail_groups(detail_out, n, D);
    for (int k = 0; k < n; ++k) {
        bool all = true;  // threshold_block.f90:119-124
        for (int c = 0; c < neq; ++c)
            if (!(detail_out[(size_t)k * neq + c] <= thresh[c])) all = false;
        flags_out[k] = all ? -1 : 0;
    }
    return WH_OK;
}

}  // extern "C"
