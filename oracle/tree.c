/*
 * tree.c -- oracle: adapt_tree on the full tree (mothers created down to
 * Jmin), the ghost syncs it uses, coarse extension, daughter-to-mother
 * transfer and the wavelet reconstruction, plus refine-everywhere.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see wavelet_oracle.h for pinning.
 *
 * Ghost syncs work by point location (like or_sync_leaf): a ghost point of
 * a receiver takes its value from the same-level block at its position, or
 * else from the finer block there (restriction), or else from the coarser
 * leaf (prediction), subject to the sender/receiver rules of each sync case
 * (prepare_ghost_synch_metadata, synchronize_ghosts_generic.f90:352-694)
 * and the source-array rules of xfer_block_data.f90:321-437.  Stages are
 * evaluated as in sync_ghosts_generic: copies first, restriction next (for
 * lifted wavelets the sender block is filtered with HD after the copies and
 * SC points near level jumps are put back, restrict_copy_at_CE), prediction
 * last from the coarse block's state after the first two stages.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wabbit_oracle.h"
#include "wavelet_oracle.h"

enum { REF_TMP_EMPTY = 10, REF_UNSIGNIFICANT_STAY = 9 };

struct or_tree {
    or_wavelet w;
    or_adapt_params p;
    int dim, g, neq, nx, ny, nz;
    size_t cs, bsz;
    int n, cap;
    int64_t *tc;
    int *level, *status;
    double **u, **t;
    /* hash (level, tc) -> index */
    int hcap;
    int *hkey_idx;
};

/* ---- bookkeeping ---------------------------------------------------------- */
static uint64_t hkey(int level, int64_t tc) { return ((uint64_t)level << 58) ^ (uint64_t)tc; }
static uint64_t hmix(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
}
static void rehash(or_tree *T)
{
    int hc = 64;
    while (hc < 4 * T->n + 4) hc *= 2;
    free(T->hkey_idx);
    T->hcap = hc;
    T->hkey_idx = (int *)malloc(sizeof(int) * hc);
    for (int i = 0; i < hc; ++i) T->hkey_idx[i] = -1;
    for (int b = 0; b < T->n; ++b) {
        uint64_t h = hmix(hkey(T->level[b], T->tc[b])) & (uint64_t)(hc - 1);
        while (T->hkey_idx[h] >= 0) h = (h + 1) & (uint64_t)(hc - 1);
        T->hkey_idx[h] = b;
    }
}
static int lookup(const or_tree *T, int level, int64_t tc)
{
    uint64_t h = hmix(hkey(level, tc)) & (uint64_t)(T->hcap - 1);
    while (T->hkey_idx[h] >= 0) {
        const int b = T->hkey_idx[h];
        if (T->level[b] == level && T->tc[b] == tc) return b;
        h = (h + 1) & (uint64_t)(T->hcap - 1);
    }
    return -1;
}
static int dshift(const or_tree *T, int l) { return (T->p.jmax - l) * T->dim; }
static void ijk_of(const or_tree *T, int64_t tc, int level, int ijk[3])
{
    ijk[0] = ijk[1] = ijk[2] = 0;
    for (int l = 1; l <= level; ++l) {
        const int d = (int)((tc >> dshift(T, l)) & ((1 << T->dim) - 1));
        ijk[0] = 2 * ijk[0] + ((d >> 1) & 1);
        ijk[1] = 2 * ijk[1] + (d & 1);
        ijk[2] = 2 * ijk[2] + (T->dim == 3 ? ((d >> 2) & 1) : 0);
    }
}
static int64_t tc_of(const or_tree *T, const int ijk[3], int level)
{
    int64_t tc = 0;
    for (int l = 1; l <= level; ++l) {
        const int sh = level - l;
        const int bx = (ijk[0] >> sh) & 1, by = (ijk[1] >> sh) & 1;
        const int bz = T->dim == 3 ? ((ijk[2] >> sh) & 1) : 0;
        tc |= (int64_t)(by + 2 * bx + 4 * bz) << dshift(T, l);
    }
    return tc;
}
/* block at (level, ijk) with periodic wrap, -1 if absent */
static int find(const or_tree *T, int level, const int ijk[3])
{
    if (level < 0) return -1;
    const int m = 1 << level;
    int w[3] = {0, 0, 0};
    for (int d = 0; d < T->dim; ++d) w[d] = ((ijk[d] % m) + m) % m;
    return lookup(T, level, tc_of(T, w, level));
}
static int valid(const or_tree *T, int b) { return b >= 0 && T->status[b] != REF_TMP_EMPTY; }
static int mother_of(const or_tree *T, int b)
{
    if (T->level[b] == 0) return -1;
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    const int m[3] = {ijk[0] >> 1, ijk[1] >> 1, ijk[2] >> 1};
    return find(T, T->level[b] - 1, m);
}
static int daughter_of(const or_tree *T, int b, int k)
{
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    const int d[3] = {2 * ijk[0] + ((k >> 1) & 1), 2 * ijk[1] + (k & 1), 2 * ijk[2] + ((k >> 2) & 1)};
    return find(T, T->level[b] + 1, d);
}
/* block_is_leaf (block_relations.f90), check_empty treats REF_TMP_EMPTY
 * daughters as absent */
static int is_leaf(const or_tree *T, int b, int check_empty)
{
    for (int k = 0; k < (1 << T->dim); ++k) {
        const int d = daughter_of(T, b, k);
        if (d >= 0 && (!check_empty || T->status[d] != REF_TMP_EMPTY)) return 0;
    }
    return 1;
}

static int add_block(or_tree *T, int64_t tc, int level, int status)
{
    if (T->n == T->cap) {
        T->cap = T->cap ? 2 * T->cap : 64;
        T->tc = (int64_t *)realloc(T->tc, sizeof(int64_t) * T->cap);
        T->level = (int *)realloc(T->level, sizeof(int) * T->cap);
        T->status = (int *)realloc(T->status, sizeof(int) * T->cap);
        T->u = (double **)realloc(T->u, sizeof(double *) * T->cap);
        T->t = (double **)realloc(T->t, sizeof(double *) * T->cap);
    }
    const int b = T->n++;
    T->tc[b] = tc;
    T->level[b] = level;
    T->status[b] = status;
    /* fresh slots hold garbage in the reference; NaN makes any use visible */
    T->u[b] = (double *)malloc(sizeof(double) * T->bsz);
    T->t[b] = (double *)malloc(sizeof(double) * T->bsz);
    for (size_t q = 0; q < T->bsz; ++q) T->u[b][q] = T->t[b][q] = NAN;
    return b;
}
/* delete all blocks with del[b] set, keeping list order */
static void delete_marked(or_tree *T, const char *del)
{
    int m = 0;
    for (int b = 0; b < T->n; ++b) {
        if (del[b]) {
            free(T->u[b]);
            free(T->t[b]);
            continue;
        }
        T->tc[m] = T->tc[b];
        T->level[m] = T->level[b];
        T->status[m] = T->status[b];
        T->u[m] = T->u[b];
        T->t[m] = T->t[b];
        ++m;
    }
    T->n = m;
    rehash(T);
}

or_tree *or_tree_create(const or_wavelet *w, const or_adapt_params *p, int n, const int64_t *tc,
                        const int *level, const double *hvy)
{
    or_tree *T = (or_tree *)calloc(1, sizeof(or_tree));
    T->w = *w;
    T->p = *p;
    T->dim = p->dim;
    T->g = p->g;
    T->neq = p->neq;
    T->nx = p->bs[0] + 2 * p->g;
    T->ny = p->bs[1] + 2 * p->g;
    T->nz = p->dim == 3 ? p->bs[2] + 2 * p->g : 1;
    T->cs = (size_t)T->nx * T->ny * T->nz;
    T->bsz = T->cs * p->neq;
    for (int b = 0; b < n; ++b) {
        const int k = add_block(T, tc[b], level[b], 0);
        memcpy(T->u[k], hvy + (size_t)b * T->bsz, sizeof(double) * T->bsz);
        memcpy(T->t[k], T->u[k], sizeof(double) * T->bsz);
    }
    rehash(T);
    return T;
}

void or_tree_destroy(or_tree *T)
{
    if (!T) return;
    for (int b = 0; b < T->n; ++b) {
        free(T->u[b]);
        free(T->t[b]);
    }
    free(T->tc);
    free(T->level);
    free(T->status);
    free(T->u);
    free(T->t);
    free(T->hkey_idx);
    free(T);
}

int or_tree_size(const or_tree *T) { return T->n; }

static int cmp_idx_T_level_tc(const or_tree *T, int a, int b)
{
    if (T->level[a] != T->level[b]) return T->level[a] < T->level[b] ? -1 : 1;
    return T->tc[a] < T->tc[b] ? -1 : (T->tc[a] > T->tc[b]);
}
void or_tree_get(const or_tree *T, int64_t *tc, int *level, double *hvy)
{
    int *ord = (int *)malloc(sizeof(int) * (T->n ? T->n : 1));
    for (int b = 0; b < T->n; ++b) ord[b] = b;
    /* insertion sort, small n */
    for (int i = 1; i < T->n; ++i) {
        const int v = ord[i];
        int j = i - 1;
        while (j >= 0 && cmp_idx_T_level_tc(T, ord[j], v) > 0) {
            ord[j + 1] = ord[j];
            --j;
        }
        ord[j + 1] = v;
    }
    for (int i = 0; i < T->n; ++i) {
        const int b = ord[i];
        if (tc) tc[i] = T->tc[b];
        if (level) level[i] = T->level[b];
        if (hvy) memcpy(hvy + (size_t)i * T->bsz, T->u[b], sizeof(double) * T->bsz);
    }
    free(ord);
}

/* ---- geometry ------------------------------------------------------------- */
static int nd_dirs(const or_tree *T, int dirs[26][3])
{
    int nd = 0;
    for (int oz = (T->dim == 3 ? -1 : 0); oz <= (T->dim == 3 ? 1 : 0); ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                if (!ox && !oy && !oz) continue;
                dirs[nd][0] = ox;
                dirs[nd][1] = oy;
                dirs[nd][2] = oz;
                ++nd;
            }
    return nd;
}
/* same-level neighbour in direction o (any status), -1 if absent */
static int nb_same(const or_tree *T, int b, const int o[3])
{
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    const int q[3] = {ijk[0] + o[0], ijk[1] + o[1], ijk[2] + o[2]};
    return find(T, T->level[b], q);
}
/* finer neighbours in direction o: 1 if any exists with status filter */
static int nb_finer_any(const or_tree *T, int b, const int o[3], int need_valid)
{
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    const int nbv[3] = {ijk[0] + o[0], ijk[1] + o[1], ijk[2] + o[2]};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        if (d >= T->dim) {
            lo[d] = hi[d] = 0;
        } else if (o[d] < 0) {
            lo[d] = hi[d] = 1;
        } else if (o[d] > 0) {
            lo[d] = hi[d] = 0;
        } else {
            lo[d] = 0;
            hi[d] = 1;
        }
    }
    for (int cz = lo[2]; cz <= hi[2]; ++cz)
        for (int cy = lo[1]; cy <= hi[1]; ++cy)
            for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                const int c[3] = {2 * nbv[0] + cx, 2 * nbv[1] + cy, T->dim == 3 ? 2 * nbv[2] + cz : 0};
                const int f = find(T, T->level[b] + 1, c);
                if (f >= 0 && (!need_valid || valid(T, f))) return 1;
            }
    return 0;
}
/* coarse relation in direction o (find_neighbors.f90): only at the mother's
 * boundary; returns the coarse block (any status) or -1 */
static int nb_coarse(const or_tree *T, int b, const int o[3])
{
    if (T->level[b] == 0) return -1;
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    for (int d = 0; d < T->dim; ++d) {
        if (o[d] > 0 && (ijk[d] & 1) != 1) return -1;
        if (o[d] < 0 && (ijk[d] & 1) != 0) return -1;
    }
    int c[3] = {0, 0, 0};
    for (int d = 0; d < T->dim; ++d) {
        const int v = ijk[d] + o[d];
        c[d] = v >= 0 ? v / 2 : -((-v + 1) / 2);
    }
    return find(T, T->level[b] - 1, c);
}

/* ---- coarse extension ------------------------------------------------------ */
static int is_sc(const or_tree *T, int ix, int iy, int iz)
{
    const int g = T->g;
    return ((ix - g) & 1) == 0 && ((iy - g) & 1) == 0 && (T->dim == 2 || ((iz - g) & 1) == 0);
}
/* get_indices_of_modify_patch (neighborhood.f90) for direction o, widths
 * Nl (lower side) / Nr (upper side), lvl_diff = +1; 0-based inclusive */
static void modify_box(const or_tree *T, const int o[3], int Nl, int Nr, int lo[3], int hi[3])
{
    const int g = T->g;
    for (int d = 0; d < 3; ++d) {
        if (d >= T->dim) {
            lo[d] = hi[d] = 0;
            continue;
        }
        const int B = T->p.bs[d];
        lo[d] = g;
        hi[d] = g + B - 1;
        if (o[d] < 0) hi[d] = g + Nl - 1;
        if (o[d] > 0) lo[d] = g + B - Nr;
    }
}
/* get_indices_of_ghost_patch with lvl_diff = +1 for the coarse relation of
 * block b in direction o: ghost slab, extended along each free direction
 * toward the block's sister side. */
static void ce_ghost_box(const or_tree *T, int b, const int o[3], int lo[3], int hi[3])
{
    const int g = T->g;
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    for (int d = 0; d < 3; ++d) {
        if (d >= T->dim) {
            lo[d] = hi[d] = 0;
            continue;
        }
        const int B = T->p.bs[d];
        if (o[d] < 0) {
            lo[d] = 0;
            hi[d] = g - 1;
        } else if (o[d] > 0) {
            lo[d] = g + B;
            hi[d] = B + 2 * g - 1;
        } else if (ijk[d] & 1) {
            lo[d] = 0;
            hi[d] = g + B - 1;
        } else {
            lo[d] = g;
            hi[d] = B + 2 * g - 1;
        }
    }
}
static void zero_wc_box(or_tree *T, double *u, const int lo[3], const int hi[3])
{
    for (int c = 0; c < T->neq; ++c)
        for (int iz = lo[2]; iz <= hi[2]; ++iz)
            for (int iy = lo[1]; iy <= hi[1]; ++iy)
                for (int ix = lo[0]; ix <= hi[0]; ++ix)
                    if (!is_sc(T, ix, iy, iz)) u[c * T->cs + ((size_t)iz * T->ny + iy) * T->nx + ix] = 0.0;
}
static void copy_sc_box(or_tree *T, double *dst, const double *src, const int lo[3], const int hi[3])
{
    for (int c = 0; c < T->neq; ++c)
        for (int iz = lo[2]; iz <= hi[2]; ++iz)
            for (int iy = lo[1]; iy <= hi[1]; ++iy)
                for (int ix = lo[0]; ix <= hi[0]; ++ix)
                    if (is_sc(T, ix, iy, iz)) {
                        const size_t q = c * T->cs + ((size_t)iz * T->ny + iy) * T->nx + ix;
                        dst[q] = src[q];
                    }
}

enum { CE_TREE = 0, CE_LEVEL = 1, CE_REF = 2 };
/* coarse_extension_modify (reconstruction_step.f90:3-100) */
static void ce_modify(or_tree *T, int ce_case, int s_level, int s_ref, int clear_wc, int copy_sc)
{
    int dirs[26][3];
    const int nd = nd_dirs(T, dirs);
    for (int b = 0; b < T->n; ++b) {
        if (ce_case == CE_LEVEL && T->level[b] != s_level) continue;
        if (ce_case == CE_REF && T->status[b] != s_ref) continue;
        if (!is_leaf(T, b, 0)) continue;
        for (int q = 0; q < nd; ++q) {
            const int *o = dirs[q];
            const int c = nb_coarse(T, b, o);
            if (c < 0) continue;
            if (valid(T, nb_same(T, b, o)) || nb_finer_any(T, b, o, 1)) continue;
            int lo[3], hi[3];
            if (clear_wc) {
                modify_box(T, o, T->w.Nwcl, T->w.Nwcr, lo, hi);
                zero_wc_box(T, T->u[b], lo, hi);
                ce_ghost_box(T, b, o, lo, hi);
                zero_wc_box(T, T->u[b], lo, hi);
            }
            if (copy_sc) {
                modify_box(T, o, T->w.Nscl, T->w.Nscr, lo, hi);
                copy_sc_box(T, T->u[b], T->t[b], lo, hi);
            }
        }
    }
}

/* restrict_copy_at_CE (restrict_predict_data.f90:101-170): the HD-filtered
 * block with the original SC values put back near level jumps of a leaf */
static void restricted_block(or_tree *T, int b, const double *src, double *out)
{
    or_filter_block(&T->w, T->dim, T->p.bs, T->g, T->neq, src, out);
    if (!is_leaf(T, b, 0)) return;
    int dirs[26][3];
    const int nd = nd_dirs(T, dirs);
    for (int q = 0; q < nd; ++q) {
        const int *o = dirs[q];
        if (valid(T, nb_same(T, b, o))) continue;
        if (nb_coarse(T, b, o) < 0 && !nb_finer_any(T, b, o, 0)) continue;
        int lo[3], hi[3];
        modify_box(T, o, T->w.Nscl, T->w.Nscr, lo, hi);
        copy_sc_box(T, out, src, lo, hi);
    }
}

/* ---- ghost sync -------------------------------------------------------------- */
enum { SRC_U = 0, SRC_NOT_FLAG = 1, SRC_DECOMPOSED = 2 };
typedef struct {
    int same, restrict_, predict;
    int recv_ref_on, recv_ref;
    int recv_level_on, recv_level;
    int src_mode, src_ref;
    int filter;
    int gm, gp;
} sync_spec;

static const double *src_array(const or_tree *T, const sync_spec *S, int sender, int kind /*0 same,1 res/pred*/)
{
    if (S->src_mode == SRC_NOT_FLAG && T->status[sender] != S->src_ref) return T->t[sender];
    if (S->src_mode == SRC_DECOMPOSED && kind == 1) return T->t[sender];
    return T->u[sender];
}

typedef struct {
    double *dst;
    double v;
} wr_t;

static int tree_sync(or_tree *T, const sync_spec *S)
{
    const int dim = T->dim, g = T->g;
    const int nx = T->nx, ny = T->ny, nz = T->nz;
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = dim == 3 ? 2 * nz - 1 : 1;
    size_t wcap = 1024, nw = 0;
    wr_t *W = (wr_t *)malloc(sizeof(wr_t) * wcap);
    double **filt = (double **)calloc((size_t)T->n, sizeof(double *));
    double **pred = (double **)calloc((size_t)T->n, sizeof(double *));
    int rc = 0;
    for (int stage = 1; stage <= 3 && !rc; ++stage) {
        nw = 0;
        for (int b = 0; b < T->n && !rc; ++b) {
            if (T->status[b] == REF_TMP_EMPTY) continue;
            if (S->recv_ref_on && T->status[b] != S->recv_ref) continue;
            if (S->recv_level_on && T->level[b] != S->recv_level) continue;
            int ijk[3];
            ijk_of(T, T->tc[b], T->level[b], ijk);
            const int J = T->level[b];
            const int zl = dim == 3 ? g - S->gm : 0, zh = dim == 3 ? g + T->p.bs[2] + S->gp : 1;
            for (int iz = zl; iz < zh && !rc; ++iz)
                for (int iy = g - S->gm; iy < g + T->p.bs[1] + S->gp && !rc; ++iy)
                    for (int ix = g - S->gm; ix < g + T->p.bs[0] + S->gp && !rc; ++ix) {
                        const int ii[3] = {ix, iy, iz};
                        int o[3] = {0, 0, 0}, inner = 1;
                        for (int d = 0; d < dim; ++d) {
                            if (ii[d] < g) o[d] = -1;
                            else if (ii[d] >= g + T->p.bs[d]) o[d] = 1;
                            if (o[d]) inner = 0;
                        }
                        if (inner) continue;
                        const size_t me = ((size_t)iz * ny + iy) * nx + ix;
                        /* global point index at level J (periodic) */
                        const int ns = nb_same(T, b, o);
                        if (valid(T, ns)) {
                            if (stage != 1 || !S->same) continue;
                            int li[3] = {0, 0, 0};
                            for (int d = 0; d < dim; ++d) li[d] = ii[d] - o[d] * T->p.bs[d];
                            const double *src = src_array(T, S, ns, 0);
                            const size_t sq = ((size_t)li[2] * ny + li[1]) * nx + li[0];
                            for (int c = 0; c < T->neq; ++c) {
                                if (nw == wcap) W = (wr_t *)realloc(W, sizeof(wr_t) * (wcap *= 2));
                                W[nw++] = (wr_t){T->u[b] + c * T->cs + me, src[c * T->cs + sq]};
                            }
                            continue;
                        }
                        /* finer block containing the point (fine grid index 2*(global)) */
                        int fglob[3] = {0, 0, 0}, cij[3] = {0, 0, 0};
                        for (int d = 0; d < dim; ++d) {
                            fglob[d] = 2 * (ijk[d] * T->p.bs[d] + (ii[d] - g));
                            cij[d] = fglob[d] >= 0 ? fglob[d] / T->p.bs[d] : -((-fglob[d] + T->p.bs[d] - 1) / T->p.bs[d]);
                        }
                        const int fb = find(T, J + 1, cij);
                        if (valid(T, fb)) {
                            if (stage != 2 || !S->restrict_) continue;
                            int fij[3];
                            ijk_of(T, T->tc[fb], J + 1, fij);
                            /* wrap the child index to compute the local offset */
                            int li[3] = {0, 0, 0};
                            for (int d = 0; d < dim; ++d)
                                li[d] = fglob[d] - cij[d] * T->p.bs[d] + g;
                            const double *src;
                            if (S->filter && T->w.lifted) {
                                if (!filt[fb]) {
                                    filt[fb] = (double *)malloc(sizeof(double) * T->bsz);
                                    restricted_block(T, fb, src_array(T, S, fb, 1), filt[fb]);
                                }
                                src = filt[fb];
                            } else {
                                src = src_array(T, S, fb, 1);
                            }
                            const size_t sq = ((size_t)li[2] * ny + li[1]) * nx + li[0];
                            for (int c = 0; c < T->neq; ++c) {
                                if (nw == wcap) W = (wr_t *)realloc(W, sizeof(wr_t) * (wcap *= 2));
                                W[nw++] = (wr_t){T->u[b] + c * T->cs + me, src[c * T->cs + sq]};
                            }
                            continue;
                        }
                        /* coarser leaf covering the point: the face patches of a coarse
                         * relation reach into the corner on the sister side
                         * (get_indices_of_ghost_patch, lvl_diff=+1), so the covering
                         * block is the sender also where no corner relation exists */
                        int P[3] = {0, 0, 0}, cu[3] = {0, 0, 0};
                        for (int d = 0; d < dim; ++d) {
                            P[d] = ijk[d] * T->p.bs[d] + (ii[d] - g);
                            const int s2 = 2 * T->p.bs[d];
                            cu[d] = P[d] >= 0 ? P[d] / s2 : -((-P[d] + s2 - 1) / s2);
                        }
                        const int cb = find(T, J - 1, cu);
                        if (!valid(T, cb) || !is_leaf(T, cb, 1)) continue;
                        if (stage != 3 || !S->predict) continue;
                        if (!pred[cb]) {
                            pred[cb] = (double *)malloc(sizeof(double) * (size_t)fx * fy * fz * T->neq);
                            const double *src = src_array(T, S, cb, 1);
                            for (int c = 0; c < T->neq; ++c)
                                or_prediction_zorder(src + c * T->cs, nx, ny, nz,
                                                     pred[cb] + (size_t)c * fx * fy * fz, T->w.order,
                                                     (T->p.plant & 4) ? 2 : T->w.order);
                        }
                        int fi[3] = {0, 0, 0};
                        for (int d = 0; d < dim; ++d) {
                            /* fine position relative to the (unwrapped) coarse block origin,
                             * in the coarse block's doubled ghost-inclusive index space */
                            fi[d] = P[d] - 2 * cu[d] * T->p.bs[d] + 2 * g;
                            const int lim = d == 0 ? fx : d == 1 ? fy : fz;
                            if (fi[d] < 0 || fi[d] >= lim) rc = -3;
                        }
                        if (rc) break;
                        const size_t sq = fi[0] + (size_t)fx * (fi[1] + (size_t)fy * fi[2]);
                        for (int c = 0; c < T->neq; ++c) {
                            if (nw == wcap) W = (wr_t *)realloc(W, sizeof(wr_t) * (wcap *= 2));
                            W[nw++] = (wr_t){T->u[b] + c * T->cs + me, pred[cb][(size_t)c * fx * fy * fz + sq]};
                        }
                    }
        }
        for (size_t k = 0; k < nw; ++k) *W[k].dst = W[k].v;
    }
    for (int b = 0; b < T->n; ++b) {
        free(filt[b]);
        free(pred[b]);
    }
    free(filt);
    free(pred);
    free(W);
    return rc;
}

/* sync_ghosts_tree: "full" on all blocks, filtered restriction */
int or_tree_sync_full(or_tree *T)
{
    sync_spec S = {0};
    S.same = S.restrict_ = S.predict = 1;
    S.src_mode = SRC_U;
    S.filter = 1;
    S.gm = S.gp = T->g;
    return tree_sync(T, &S);
}

/* ---- refinement -------------------------------------------------------------- */
int or_tree_refine_everywhere(or_tree *T)
{
    const int n0 = T->n;
    char *del = (char *)calloc((size_t)n0 + 1, 1);
    for (int b = 0; b < n0; ++b) {
        if (T->level[b] >= T->p.jmax) continue;
        double *d[8];
        int kk[8];
        for (int k = 0; k < (1 << T->dim); ++k) {
            const int64_t tcd = T->tc[b] | ((int64_t)k << dshift(T, T->level[b] + 1));
            kk[k] = add_block(T, tcd, T->level[b] + 1, 0);
        }
        for (int k = 0; k < (1 << T->dim); ++k) d[k] = T->u[kk[k]];
        or_refine_block(T->dim, T->p.bs, T->g, T->neq, T->u[b], d, T->w.order);
        for (int k = 0; k < (1 << T->dim); ++k) memcpy(T->t[kk[k]], T->u[kk[k]], sizeof(double) * T->bsz);
        del[b] = 1;
    }
    char *del2 = (char *)calloc((size_t)T->n, 1);
    memcpy(del2, del, (size_t)n0);
    delete_marked(T, del2);
    free(del);
    free(del2);
    return 0;
}

/* ---- adapt_tree -------------------------------------------------------------- */
/* init_full_tree (adapt_tree.f90:995-1084): mothers down to jmin */
static void init_full_tree(or_tree *T, int jmin)
{
    int changed = 1;
    while (changed) {
        changed = 0;
        rehash(T);
        const int n0 = T->n;
        for (int b = 0; b < n0; ++b) {
            if (T->level[b] == jmin || T->level[b] == 0) continue;
            if (mother_of(T, b) >= 0) continue;
            const int64_t tcm = T->tc[b] & ~((((int64_t)1 << T->dim) - 1) << dshift(T, T->level[b]));
            if (lookup(T, T->level[b] - 1, tcm) >= 0) continue;
            add_block(T, tcm, T->level[b] - 1, REF_TMP_EMPTY);
            rehash(T);
            changed = 1;
        }
    }
    rehash(T);
}

/* sync_D2M (executeCoarsening_tree.f90:125-227): daughters on `level` give
 * their SC to the mother's quadrant */
static void d2m(or_tree *T, int level)
{
    const int g = T->g;
    for (int b = 0; b < T->n; ++b) {
        if (T->level[b] != level) continue;
        const int m = mother_of(T, b);
        if (m < 0) continue;
        int ijk[3];
        ijk_of(T, T->tc[b], level, ijk);
        const int off[3] = {ijk[0] & 1, ijk[1] & 1, T->dim == 3 ? (ijk[2] & 1) : 0};
        const int hx = T->p.bs[0] / 2, hy = T->p.bs[1] / 2, hz = T->dim == 3 ? T->p.bs[2] / 2 : 1;
        for (int c = 0; c < T->neq; ++c)
            for (int k = 0; k < hz; ++k)
                for (int j = 0; j < hy; ++j)
                    for (int i = 0; i < hx; ++i) {
                        const int mz = T->dim == 3 ? g + off[2] * hz + k : 0;
                        const int dz = T->dim == 3 ? g + 2 * k : 0;
                        T->u[m][c * T->cs + ((size_t)mz * T->ny + g + off[1] * hy + j) * T->nx + g + off[0] * hx + i] =
                            T->u[b][c * T->cs + ((size_t)dz * T->ny + g + 2 * j) * T->nx + g + 2 * i];
                    }
    }
}

static int max_level(const or_tree *T)
{
    int m = -1;
    for (int b = 0; b < T->n; ++b) m = T->level[b] > m ? T->level[b] : m;
    return m;
}

/* wavelet_decompose_full_tree (adapt_tree.f90:264-545) */
static int decompose_full_tree(or_tree *T)
{
    const or_wavelet *w = &T->w;
    const int jmin = T->p.jmin;
    int *backup = (int *)malloc(sizeof(int) * (size_t)T->n);
    const int n_leaf = T->n;
    for (int b = 0; b < T->n; ++b) backup[b] = T->status[b];
    int hmax = abs(w->hd_lo) > w->hd_hi ? abs(w->hd_lo) : w->hd_hi;
    int leaf_first = 1;
    for (int d = 0; d < T->dim; ++d)
        if (T->p.bs[d] < 3 * hmax) leaf_first = 0;
    if (T->p.decompose_branch == 1) leaf_first = 1;
    if (T->p.decompose_branch == 2) leaf_first = 0;
    int level = max_level(T);
    for (int b = 0; b < T->n; ++b) {
        if (T->level[b] == level || leaf_first) T->status[b] = -1;
        memcpy(T->t[b], T->u[b], sizeof(double) * T->bsz);
    }
    init_full_tree(T, jmin);
    const int g_this = abs(w->hd_lo) > w->gd_hi ? abs(w->hd_lo) : w->gd_hi;
    int it = 0;
    int rc = 0;
    while (level >= jmin && !rc) {
        sync_spec S = {0};
        S.same = 1;
        S.restrict_ = 1;
        S.predict = !(T->p.use_ce && (!leaf_first || it > 0));
        S.recv_ref_on = 1;
        S.recv_ref = -1;
        S.src_mode = SRC_NOT_FLAG;
        S.src_ref = -1;
        S.filter = 1;
        S.gm = S.gp = g_this;
        if (!((T->p.plant & 2) && !leaf_first && it > 0)) rc = tree_sync(T, &S);
        if (rc) break;
        for (int b = 0; b < T->n; ++b) {
            if (T->status[b] != -1) continue;
            memcpy(T->t[b], T->u[b], sizeof(double) * T->bsz);
            if (T->level[b] == T->p.jmax && T->p.force_maxlevel_dealiasing)
                or_filter_block(w, T->dim, T->p.bs, T->g, T->neq, T->t[b], T->u[b]);
            else
                or_fwt_block(w, T->dim, T->p.bs, T->g, T->neq, T->t[b], T->u[b]);
        }
        if (T->p.use_ce) ce_modify(T, CE_REF, 0, -1, 1, 1);
        for (int b = 0; b < T->n; ++b) {
            if (T->status[b] == -1) T->status[b] = 0;
            if ((T->status[b] == REF_TMP_EMPTY || !leaf_first) && T->level[b] == level - 1) T->status[b] = -1;
            if ((T->p.plant & 1) && !leaf_first && b < n_leaf && T->status[b] == -1) T->status[b] = 0;
        }
        d2m(T, level);
        --level;
        ++it;
    }
    for (int b = 0; b < T->n; ++b) T->status[b] = b < n_leaf ? backup[b] : 0;
    free(backup);
    return rc;
}

/* ensureGradedness_tree (ensureGradedness_tree.f90:13-219), with
 * ensure_completeness_block (check_daughters) */
static void ensure_gradedness(or_tree *T, int stay)
{
    int dirs[26][3];
    const int nd = nd_dirs(T, dirs);
    const int nsis = 1 << T->dim;
    int changed = 1, counter = 0;
    while (changed && counter < 10 * T->p.jmax + 10) {
        changed = 0;
        ++counter;
        for (int b = 0; b < T->n; ++b) {
            int me = T->status[b];
            if (me == -1) {
                /* completeness */
                const int m = mother_of(T, b);
                int sis[8], all = 1;
                for (int k = 0; k < nsis; ++k) {
                    sis[k] = m >= 0 ? daughter_of(T, m, k) : -1;
                    if (sis[k] < 0) all = 0;
                }
                if (all) {
                    int any_empty = 0, any_stay = 0, all_neg = 1;
                    for (int k = 0; k < nsis; ++k) {
                        const int s = T->status[sis[k]];
                        if (s == REF_TMP_EMPTY) any_empty = 1;
                        if (s == 0 || s == stay) any_stay = 1;
                        if (s >= 0) all_neg = 0;
                    }
                    if (any_empty || any_stay) {
                        for (int k = 0; k < nsis; ++k)
                            if (T->status[sis[k]] < 0) T->status[sis[k]] = stay;
                    } else if (all_neg) {
                        for (int k = 0; k < nsis; ++k) T->status[sis[k]] = -1;
                    }
                    if (!is_leaf(T, b, 0)) {
                        for (int k = 0; k < nsis; ++k) {
                            const int d = daughter_of(T, b, k);
                            if (d >= 0 && (T->status[d] == 0 || T->status[d] == stay)) {
                                for (int l = 0; l < nsis; ++l)
                                    if (T->status[sis[l]] < 0) T->status[sis[l]] = stay;
                            }
                        }
                    }
                } else {
                    for (int k = 0; k < nsis; ++k)
                        if (sis[k] >= 0 && T->status[sis[k]] < -1) T->status[sis[k]] = stay;
                }
                if (T->status[b] != me) {
                    me = T->status[b];
                    changed = 1;
                }
            }
            if (me == -1) {
                for (int q = 0; q < nd; ++q) {
                    const int *o = dirs[q];
                    const int s = nb_same(T, b, o);
                    if (s >= 0 && T->status[s] == 1) {
                        T->status[b] = stay;
                        changed = 1;
                    }
                    /* finer neighbours */
                    int ijk[3];
                    ijk_of(T, T->tc[b], T->level[b], ijk);
                    for (int k = 0; k < nsis; ++k) {
                        int c[3];
                        const int cc[3] = {(k >> 1) & 1, k & 1, (k >> 2) & 1};
                        int adj = 1;
                        for (int d = 0; d < 3; ++d) {
                            if (d >= T->dim) {
                                c[d] = 0;
                                continue;
                            }
                            c[d] = 2 * (ijk[d] + o[d]) + cc[d];
                            if (o[d] < 0 && cc[d] != 1) adj = 0;
                            if (o[d] > 0 && cc[d] != 0) adj = 0;
                        }
                        if (!adj) continue;
                        const int f = find(T, T->level[b] + 1, c);
                        if (f < 0) continue;
                        if (T->status[f] == 1) {
                            T->status[b] = 1;
                            changed = 1;
                        } else if (T->status[f] == 0 || T->status[f] == stay) {
                            if (T->status[b] != 1) T->status[b] = stay;
                            changed = 1;
                        }
                    }
                }
            } else if (me == 0 || me == stay) {
                for (int q = 0; q < nd; ++q) {
                    const int *o = dirs[q];
                    int ijk[3];
                    ijk_of(T, T->tc[b], T->level[b], ijk);
                    for (int k = 0; k < nsis; ++k) {
                        int c[3];
                        const int cc[3] = {(k >> 1) & 1, k & 1, (k >> 2) & 1};
                        int adj = 1;
                        for (int d = 0; d < 3; ++d) {
                            if (d >= T->dim) {
                                c[d] = 0;
                                continue;
                            }
                            c[d] = 2 * (ijk[d] + o[d]) + cc[d];
                            if (o[d] < 0 && cc[d] != 1) adj = 0;
                            if (o[d] > 0 && cc[d] != 0) adj = 0;
                        }
                        if (!adj) continue;
                        const int f = find(T, T->level[b] + 1, c);
                        if (f >= 0 && T->status[f] == 1) {
                            T->status[b] = 1;
                            changed = 1;
                        }
                    }
                }
            }
        }
    }
}

/* wavelet_reconstruct_full_tree_CEoptimized (adapt_tree.f90:686-987) */
static int reconstruct_tree(or_tree *T)
{
    const or_wavelet *w = &T->w;
    int dirs[26][3];
    const int nd = nd_dirs(T, dirs);
    int *backup = (int *)malloc(sizeof(int) * (size_t)T->n);
    for (int b = 0; b < T->n; ++b) {
        backup[b] = T->status[b];
        T->status[b] = 0;
    }
    for (int b = 0; b < T->n; ++b) {
        if (!is_leaf(T, b, 0)) continue;
        for (int q = 0; q < nd; ++q) {
            const int *o = dirs[q];
            if (nb_coarse(T, b, o) < 0) continue;
            if (valid(T, nb_same(T, b, o)) || nb_finer_any(T, b, o, 1)) continue;
            T->status[b] = -1;
            break;
        }
    }
    int reconstruct_neighbors = 0;
    for (int d = 0; d < T->dim; ++d)
        if (T->p.bs[d] < w->Nreconl || T->p.bs[d] < w->Nreconr) reconstruct_neighbors = 1;
    if (reconstruct_neighbors) {
        for (int b = 0; b < T->n; ++b) {
            if (T->status[b] == -1) continue;
            for (int q = 0; q < nd; ++q) {
                const int s = nb_same(T, b, dirs[q]);
                if (s >= 0 && T->status[s] == -1) {
                    T->status[b] = -2;
                    break;
                }
            }
        }
        for (int b = 0; b < T->n; ++b)
            if (T->status[b] == -2) T->status[b] = -1;
    }
    sync_spec S = {0};
    S.same = 1;
    S.predict = 1;
    S.recv_ref_on = 1;
    S.recv_ref = -1;
    S.src_mode = SRC_DECOMPOSED;
    S.filter = 1;
    S.gm = S.gp = T->g;
    int rc = 0;
    int leaf_only = 1;
    for (int d = 0; d < T->dim; ++d)
        if (T->p.bs[d] < w->Ndep2l || T->p.bs[d] < w->Ndep2r) leaf_only = 0;
    if (leaf_only) {
        rc = tree_sync(T, &S);
        if (T->p.use_ce) ce_modify(T, CE_TREE, 0, 0, 1, 0);
        for (int b = 0; b < T->n; ++b) {
            if (T->status[b] == -1) or_iwt_block(w, T->dim, T->p.bs, T->g, T->neq, T->u[b], T->t[b]);
            memcpy(T->u[b], T->t[b], sizeof(double) * T->bsz);
        }
    } else {
        /* level by level from Jmin (adapt_tree.f90:880-980); sync_M2D only
         * when reconstruct_neighbors, which the supported sizes exclude */
        if (reconstruct_neighbors) rc = -11;
        const int lmax = max_level(T);
        for (int L = T->p.jmin; L <= lmax && !rc; ++L) {
            S.recv_level_on = 1;
            S.recv_level = L;
            rc = tree_sync(T, &S);
            if (T->p.use_ce) ce_modify(T, CE_LEVEL, L, 0, 1, 0);
            for (int b = 0; b < T->n; ++b) {
                if (T->level[b] != L) continue;
                if (T->status[b] == -1) or_iwt_block(w, T->dim, T->p.bs, T->g, T->neq, T->u[b], T->t[b]);
                memcpy(T->u[b], T->t[b], sizeof(double) * T->bsz);
            }
        }
    }
    for (int b = 0; b < T->n; ++b) T->status[b] = backup[b];
    free(backup);
    return rc;
}

/* get_block_spacing_origin: dx = 2^-J L / Bs, x0 = ijk Bs dx */
static void block_x0dx(const or_tree *T, int b, double x0[3], double dx[3])
{
    int ijk[3];
    ijk_of(T, T->tc[b], T->level[b], ijk);
    for (int d = 0; d < 3; ++d) {
        x0[d] = dx[d] = 0.0;
        if (d >= T->dim) continue;
        dx[d] = ldexp(1.0, -T->level[b]) * T->p.domain[d] / (double)T->p.bs[d];
        x0[d] = (double)(ijk[d] * T->p.bs[d]) * dx[d];
    }
}

/* detail of a box (threshold_block with indices) of a block on `level`,
 * SCs zeroed by the global parity, renormed (wavelet_renorm_block), grouped
 * components as in or_detail_block */
static void detail_box(const or_tree *T, const double *u, int level, const int lo[3], const int hi[3],
                       double *det)
{
    const int g = T->g;
    int maxgrp = 0;
    for (int c = 0; c < T->neq; ++c) maxgrp = T->p.groups[c] > maxgrp ? T->p.groups[c] : maxgrp;
    double m[8];
    for (int c = 0; c < T->neq; ++c) {
        m[c] = 0.0;
        for (int iz = lo[2]; iz <= hi[2]; ++iz)
            for (int iy = lo[1]; iy <= hi[1]; ++iy)
                for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                    if (is_sc(T, ix, iy, iz)) continue;
                    const double v = or_renorm_value(T->p.eps_norm, T->dim, level, ix, iy, iz, T->g,
                                                     u[c * T->cs + ((size_t)iz * T->ny + iy) * T->nx + ix]);
                    const double a = T->p.groups[c] >= 2 ? sqrt(v * v) : fabs(v);
                    if (a > m[c]) m[c] = a;
                }
        (void)g;
    }
    for (int c = 0; c < T->neq; ++c) det[c] = T->p.groups[c] == 1 ? m[c] : 0.0;
    for (int l = 2; l <= maxgrp; ++l) {
        double mm = -1.0;
        for (int c = 0; c < T->neq; ++c)
            if (T->p.groups[c] == l && m[c] > mm) mm = m[c];
        for (int c = 0; c < T->neq; ++c)
            if (T->p.groups[c] == l) det[c] = mm;
    }
}

/* componentWiseNorm_tree(hvy_tmp, "Linfty") over the leaves
 * (componentWiseNorm_tree.f90), then norm <= 1e-9 -> 1
 * (coarseningIndicator_tree.f90:160-163) */
static void leaf_linfty(const or_tree *T, double *norm)
{
    const int g = T->g;
    const int z0 = T->dim == 3 ? g : 0, z1 = T->dim == 3 ? g + T->p.bs[2] : 1;
    for (int c = 0; c < T->neq; ++c) norm[c] = -1.0;
    int maxgrp = 0;
    for (int c = 0; c < T->neq; ++c) maxgrp = T->p.groups[c] > maxgrp ? T->p.groups[c] : maxgrp;
    for (int b = 0; b < T->n; ++b) {
        if (!is_leaf(T, b, 0)) continue;
        double m[8];
        for (int c = 0; c < T->neq; ++c) {
            m[c] = 0.0;
            for (int iz = z0; iz < z1; ++iz)
                for (int iy = g; iy < g + T->p.bs[1]; ++iy)
                    for (int ix = g; ix < g + T->p.bs[0]; ++ix) {
                        const double a = fabs(T->t[b][c * T->cs + ((size_t)iz * T->ny + iy) * T->nx + ix]);
                        if (a > m[c]) m[c] = a;
                    }
        }
        for (int c = 0; c < T->neq; ++c)
            if (T->p.groups[c] == 1 && m[c] > norm[c]) norm[c] = m[c];
        for (int l = 2; l <= maxgrp; ++l) {
            double mm = 0.0;
            for (int c = 0; c < T->neq; ++c)
                if (T->p.groups[c] == l && m[c] > mm) mm = m[c];
            for (int c = 0; c < T->neq; ++c)
                if (T->p.groups[c] == l && mm > norm[c]) norm[c] = mm;
        }
    }
    for (int c = 0; c < T->neq; ++c)
        if (norm[c] <= 1.0e-9) norm[c] = 1.0;
}

/* componentWiseNorm_tree(hvy_tmp, eps_norm) over the leaves in list order,
 * then norm <= 1e-9 -> 1 (coarseningIndicator_tree.f90:146-163) */
static void leaf_norm(const or_tree *T, double *norm)
{
    if (T->p.eps_norm == OR_NORM_LINFTY) {
        leaf_linfty(T, norm);
        return;
    }
    const double **u = (const double **)malloc(sizeof(double *) * (T->n ? T->n : 1));
    double *dx = (double *)malloc(sizeof(double) * 3 * (T->n ? T->n : 1));
    int n = 0;
    for (int b = 0; b < T->n; ++b) {
        if (!is_leaf(T, b, 0)) continue;
        double x0[3];
        block_x0dx(T, b, x0, dx + 3 * n);
        u[n++] = T->t[b];
    }
    or_component_norm(T->dim, T->p.bs, T->g, T->neq, T->p.groups, T->p.eps_norm, n, u, dx, norm);
    for (int c = 0; c < T->neq; ++c)
        if (norm[c] <= 1.0e-9) norm[c] = 1.0;
    free(u);
    free(dx);
}

/* coarseningIndicatorMask_tree (coarseningIndicator_tree.f90:338-422): a
 * block whose chi is partly in (0,1) or not constant stays */
static void mask_criterion(or_tree *T)
{
    const int g = T->g;
    double *mk = (double *)malloc(sizeof(double) * T->cs * 6);
    const int z0 = T->dim == 3 ? g : 0, z1 = T->dim == 3 ? g + T->p.bs[2] : 1;
    for (int b = 0; b < T->n; ++b) {
        if (T->p.force_maxlevel_dealiasing && T->level[b] == T->p.jmax) continue;
        double x0[3], dx[3];
        block_x0dx(T, b, x0, dx);
        or_mask_block(&T->p.mask, T->dim, T->p.bs, g, x0, dx, mk);
        int stay = 0;
        double mx = -1.0e300, mn = 1.0e300;
        for (int iz = z0; iz < z1; ++iz)
            for (int iy = g; iy < g + T->p.bs[1]; ++iy)
                for (int ix = g; ix < g + T->p.bs[0]; ++ix) {
                    const double v = mk[((size_t)iz * T->ny + iy) * T->nx + ix];
                    if (v > 1.0e-12 && v < 1.0 - 1.0e-12) stay = 1;
                    mx = v > mx ? v : mx;
                    mn = v < mn ? v : mn;
                }
        if (mx - mn > 1.0e-12) stay = 1;
        const int ms = stay ? 0 : -1;
        if (ms > T->status[b]) T->status[b] = ms;
    }
    free(mk);
}

/* addSecurityZone_CE_tree (securityZone_tree.f90): a block that would
 * coarsen stays if a same-level neighbour that stays has significant
 * details within Nwc of their common boundary */
static void security_zone(or_tree *T, const double *norm)
{
    int dirs[26][3];
    const int nd = nd_dirs(T, dirs);
    char *flag = (char *)calloc((size_t)T->n * 26, 1);
    for (int b = 0; b < T->n; ++b) {
        if (T->level[b] == T->p.jmax && T->p.force_maxlevel_dealiasing) continue;
        if (T->status[b] != 0) continue;
        for (int q = 0; q < nd; ++q) {
            const int s = nb_same(T, b, dirs[q]);
            if (s < 0 || T->status[s] != -1) continue;
            int lo[3], hi[3];
            modify_box(T, dirs[q], T->w.Nwcl, T->w.Nwcr, lo, hi);
            double det[8];
            detail_box(T, T->u[b], T->level[b], lo, hi, det);
            int all = 1;
            for (int c = 0; c < T->neq; ++c)
                if (!(det[c] <= T->p.eps * norm[c])) all = 0;
            if (!all) flag[(size_t)b * 26 + q] = 1;
        }
    }
    char *keep = (char *)calloc((size_t)T->n, 1);
    for (int b = 0; b < T->n; ++b) {
        if (T->status[b] != -1) continue;
        for (int q = 0; q < nd; ++q) {
            const int s = nb_same(T, b, dirs[q]);
            if (s < 0) continue;
            /* the neighbour's patch toward b is its direction -o */
            int inv = -1;
            for (int r = 0; r < nd; ++r)
                if (dirs[r][0] == -dirs[q][0] && dirs[r][1] == -dirs[q][1] && dirs[r][2] == -dirs[q][2]) inv = r;
            if (flag[(size_t)s * 26 + inv]) keep[b] = 1;
        }
    }
    for (int b = 0; b < T->n; ++b)
        if (keep[b]) T->status[b] = 0;
    free(keep);
    free(flag);
}

int or_tree_adapt(or_tree *T) { return or_tree_adapt_flags(T, NULL); }

int or_tree_adapt_flags(or_tree *T, const int *leaf_flags)
{
    const int stay = REF_UNSIGNIFICANT_STAY;
    if (T->p.indicator == OR_IND_EXTERNAL) {
        if (!leaf_flags) return -10;
        /* the flags come in or_tree_get's order: by level, then treecode */
        int *ord = (int *)malloc(sizeof(int) * (T->n ? T->n : 1));
        for (int b = 0; b < T->n; ++b) ord[b] = b;
        for (int i = 1; i < T->n; ++i) {
            const int v = ord[i];
            int j = i - 1;
            while (j >= 0 && cmp_idx_T_level_tc(T, ord[j], v) > 0) {
                ord[j + 1] = ord[j];
                --j;
            }
            ord[j + 1] = v;
        }
        for (int i = 0; i < T->n; ++i) T->status[ord[i]] = leaf_flags[i];
        free(ord);
    }
    /* the decomposition hands the leaves' statuses back (its backup) */
    int rc = decompose_full_tree(T);
    if (rc) return rc;
    /* coarseningIndicator_tree (coarseningIndicator_tree.f90:4-331); the
     * external indicator returns at once (:55) */
    if (T->p.indicator != OR_IND_EXTERNAL)
        for (int b = 0; b < T->n; ++b) T->status[b] = 0;
    double norm[8];
    for (int c = 0; c < 8; ++c) norm[c] = T->p.norm[c];
    if (T->p.indicator == OR_IND_EXTERNAL) {
    } else if (T->p.indicator == OR_IND_EVERYWHERE) {
        for (int b = 0; b < T->n; ++b)
            if (is_leaf(T, b, 0)) T->status[b] = -1;
    } else {
        if (T->p.eps_normalized) leaf_norm(T, norm);
        double det[8];
        for (int b = 0; b < T->n; ++b) {
            if (T->p.force_maxlevel_dealiasing && T->level[b] == T->p.jmax) {
                T->status[b] = -1;
                continue;
            }
            or_detail_block_norm(T->dim, T->p.bs, T->g, T->neq, T->p.groups, T->u[b], T->level[b], T->p.eps_norm,
                                 det);
            int all = 1;
            for (int c = 0; c < T->neq; ++c)
                if (!(det[c] <= T->p.eps * norm[c])) all = 0;
            T->status[b] = all ? -1 : 0;
        }
        if (T->p.threshold_mask) mask_criterion(T);
        if (T->p.security_zone) security_zone(T, norm);
    }
    /* respectJmaxJmin_tree */
    for (int b = 0; b < T->n; ++b) {
        if (T->status[b] == 1 && T->level[b] >= T->p.jmax) T->status[b] = stay;
        if (T->status[b] == -1 && T->level[b] <= T->p.jmin) T->status[b] = stay;
    }
    ensure_gradedness(T, stay);
    {
        char *del = (char *)calloc((size_t)(T->n > 0 ? T->n : 1), 1);
        for (int b = 0; b < T->n; ++b) del[b] = T->status[b] == -1;
        delete_marked(T, del);
        free(del);
    }
    if (T->p.use_ce) ce_modify(T, CE_TREE, 0, 0, 1, 0);
    if (!T->p.use_ce) {
        for (int b = 0; b < T->n; ++b) memcpy(T->u[b], T->t[b], sizeof(double) * T->bsz);
    } else {
        rc = reconstruct_tree(T);
        if (rc) return rc;
    }
    /* prune_fulltree2leafs */
    {
        char *del = (char *)calloc((size_t)(T->n > 0 ? T->n : 1), 1);
        for (int b = 0; b < T->n; ++b) del[b] = !is_leaf(T, b, 0);
        delete_marked(T, del);
        free(del);
    }
    for (int b = 0; b < T->n; ++b) T->status[b] = 0;
    return or_tree_sync_full(T);
}

/* ---- ACM time step on the leaves ------------------------------------------------ */
double or_tree_dt(or_tree *T, const or_acm_params *acm)
{
    double dt = 9.0e9;
    for (int b = 0; b < T->n; ++b) {
        double x0[3], dx[3];
        block_x0dx(T, b, x0, dx);
        const double d = or_dt_block(acm, T->p.bs, T->g, dx, T->u[b]);
        if (d < dt) dt = d;
    }
    return dt;
}

static void masks_in_list_order(or_tree *T, const or_mask_params *mp, double *mask)
{
    for (int b = 0; b < T->n; ++b) {
        double x0[3], dx[3];
        block_x0dx(T, b, x0, dx);
        or_mask_block(mp, T->dim, T->p.bs, T->g, x0, dx, mask + (size_t)b * T->cs * 6);
    }
}

/* in or_tree_get's order (level, treecode) */
void or_tree_get_mask(or_tree *T, const or_mask_params *mp, double *mask)
{
    int *ord = (int *)malloc(sizeof(int) * (T->n ? T->n : 1));
    for (int b = 0; b < T->n; ++b) ord[b] = b;
    for (int i = 1; i < T->n; ++i) {
        const int v = ord[i];
        int j = i - 1;
        while (j >= 0 && cmp_idx_T_level_tc(T, ord[j], v) > 0) {
            ord[j + 1] = ord[j];
            --j;
        }
        ord[j + 1] = v;
    }
    for (int i = 0; i < T->n; ++i) {
        double x0[3], dx[3];
        block_x0dx(T, ord[i], x0, dx);
        or_mask_block(mp, T->dim, T->p.bs, T->g, x0, dx, mask + (size_t)i * T->cs * 6);
    }
    free(ord);
}

/* RungeKuttaGeneric (runge_kutta_generic.f90:1-156) on the leaves, the RHS
 * ghosts filled by sync_ghosts_RHS_tree (plain restriction) of depth g_rhs */
int or_tree_rk_step(or_tree *T, const or_acm_params *acm, const or_mask_params *mp, const double *bt,
                    int nrows, double dt, int g_rhs)
{
    const int n = T->n, g = T->g;
    const size_t bsz = T->bsz, cs = T->cs;
    const int z0 = T->dim == 3 ? g : 0, z1 = T->dim == 3 ? g + T->p.bs[2] : 1;
    double *work = (double *)malloc(sizeof(double) * bsz * (size_t)n * (size_t)nrows);
    double *mask = mp ? (double *)malloc(sizeof(double) * cs * 6 * (size_t)n) : NULL;
    double *dxs = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    for (int b = 0; b < n; ++b) {
        double x0[3];
        block_x0dx(T, b, x0, dxs + 3 * b);
    }
    if (mp) masks_in_list_order(T, mp, mask);
    sync_spec S = {0};
    S.same = S.restrict_ = S.predict = 1;
    S.src_mode = SRC_U;
    S.filter = 0;
    S.gm = S.gp = g_rhs;
#define SLOT(j, b) (work + ((size_t)(j) * n + (b)) * bsz)
#define BT(r, c) bt[(r) * nrows + (c)]
#define FOR_INT(body)                                                              \
    for (int c = 0; c < T->neq; ++c)                                                \
        for (int iz = z0; iz < z1; ++iz)                                            \
            for (int iy = g; iy < g + T->p.bs[1]; ++iy)                             \
                for (int ix = g; ix < g + T->p.bs[0]; ++ix) {                       \
                    const size_t i = c * cs + ((size_t)iz * T->ny + iy) * T->nx + ix; \
                    body;                                                           \
                }
    int rc = tree_sync(T, &S);
    for (int b = 0; b < n && !rc; ++b) {
        double *u = T->u[b], *w1 = SLOT(0, b);
        FOR_INT(w1[i] = u[i]);
        or_rhs_acm_block(acm, T->p.bs, g, dxs + 3 * b, u, SLOT(1, b), mask ? mask + (size_t)b * cs * 6 : NULL);
    }
    for (int j = 2; j <= nrows - 1 && !rc; ++j) {
        for (int b = 0; b < n; ++b) {
            double *u = T->u[b], *w1 = SLOT(0, b);
            FOR_INT(u[i] = w1[i]);
            for (int l = 2; l <= j; ++l) {
                const double a = BT(j - 1, l - 1);
                if (fabs(a) < 1.0e-8) continue;
                const double *k = SLOT(l - 1, b);
                const double da = dt * a;
                FOR_INT(u[i] = u[i] + da * k[i]);
            }
        }
        rc = tree_sync(T, &S);
        for (int b = 0; b < n && !rc; ++b)
            or_rhs_acm_block(acm, T->p.bs, g, dxs + 3 * b, T->u[b], SLOT(j, b),
                             mask ? mask + (size_t)b * cs * 6 : NULL);
    }
    for (int b = 0; b < n && !rc; ++b) {
        double *u = T->u[b], *w1 = SLOT(0, b);
        FOR_INT(u[i] = w1[i]);
        for (int j = 2; j <= nrows; ++j) {
            const double bb = BT(nrows - 1, j - 1);
            if (fabs(bb) < 1.0e-8) continue;
            const double *k = SLOT(j - 1, b);
            const double db = dt * bb;
            FOR_INT(u[i] = u[i] + db * k[i]);
        }
    }
#undef SLOT
#undef BT
#undef FOR_INT
    free(work);
    free(mask);
    free(dxs);
    return rc;
}
