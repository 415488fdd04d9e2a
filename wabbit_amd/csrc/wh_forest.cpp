// wh_forest.cpp -- host-side forest metadata for the device pool: block
// list (treecode, level, hvy id), neighbour search and ghost-patch tables,
// refinement bookkeeping.  Host code only (no kernels).  The forest struct
// and treecode helpers are in wh_forest_impl.h.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "wh_forest_impl.h"
#include "wh_prof.h"

using Block = wh::FBlock;

namespace {

// Ghost box of receiver b in direction o (per dim -1/0/+1): ghost-inclusive
// index ranges [lo, lo+n) with depth gm (lower) / gp (upper).
void ghost_box(const wh_forest &F, const int o[3], int gm, int gp, int lo[3], int n[3])
{
    for (int d = 0; d < 3; ++d) {
        if (d >= F.dim) {
            lo[d] = 0;
            n[d] = 1;
        } else if (o[d] < 0) {
            lo[d] = F.g - gm;
            n[d] = gm;
        } else if (o[d] > 0) {
            lo[d] = F.g + F.bs[d];
            n[d] = gp;
        } else {
            lo[d] = F.g;
            n[d] = F.bs[d];
        }
    }
}

// order of a predictor "multiresolution_<2N>th": stencil half-width N
}  // namespace

extern "C" {

int wh_forest_create(wh_forest **out, int dim, int jmax, const int *bs, int g,
                     const double *domain, const int *periodic, int nb_max)
{
    if (!out || !bs || !domain || (dim != 2 && dim != 3) || jmax < 0 || jmax > 18 || g < 1 ||
        nb_max < 1) {
        wh::set_error("wh_forest_create: bad argument");
        return WH_E_ARG;
    }
    wh_forest *F = new wh_forest();
    F->dim = dim;
    F->jmax = jmax;
    F->g = g;
    F->nb_max = nb_max;
    for (int d = 0; d < 3; ++d) {
        F->bs[d] = d < dim ? bs[d] : 1;
        F->domain[d] = d < dim ? domain[d] : 1.0;
        F->periodic[d] = periodic ? periodic[d] : 1;
    }
    for (int h = nb_max - 1; h >= 0; --h) F->free_hvy.push_back(h);
    *out = F;
    return WH_OK;
}

int wh_forest_set_slot_base(wh_forest *F, int first)
{
    if (!F || first < 0 || !F->blocks.empty()) {
        wh::set_error("wh_forest_set_slot_base: bad argument or the forest has blocks");
        return WH_E_ARG;
    }
    F->slot0 = first;
    F->free_hvy.clear();
    for (int h = F->nb_max - 1; h >= 0; --h) F->free_hvy.push_back(first + h);
    return WH_OK;
}

int wh_forest_destroy(wh_forest *F)
{
    delete F;
    return WH_OK;
}

int wh_forest_set_blocks(wh_forest *F, int n, const int64_t *tc, const int32_t *level,
                         const int32_t *hvy)
{
    if (!F || n < 0 || (n && (!tc || !level || !hvy)) || n > F->nb_max) {
        wh::set_error("wh_forest_set_blocks: bad argument");
        return WH_E_ARG;
    }
    // validated before the forest changes: a rejected list leaves it as it was
    std::vector<char> used(F->nb_max, 0);
    std::vector<Block> nb;
    nb.reserve((size_t)n);
    for (int i = 0; i < n; ++i) {
        const int h = hvy[i] - F->slot0;
        if (h < 0 || h >= F->nb_max || used[h] || level[i] < 0 || level[i] > F->jmax) {
            wh::set_error("wh_forest_set_blocks: hvy id / level out of range or duplicate");
            return WH_E_ARG;
        }
        // the digits of the levels below the block's own must be zero
        // (module_treelib.f90: a block's treecode holds level[i] digits)
        const int low_bits = (F->jmax - level[i]) * F->dim;
        if (tc[i] < 0 || (low_bits > 0 && (tc[i] & ((int64_t(1) << low_bits) - 1)) != 0) ||
            (F->jmax * F->dim < 63 && (tc[i] >> (F->jmax * F->dim)) != 0)) {
            wh::set_error("wh_forest_set_blocks: treecode has digits beyond its level or beyond jmax");
            return WH_E_ARG;
        }
        used[h] = 1;
        nb.push_back(Block{tc[i], level[i], hvy[i], 0});
    }
    F->blocks.swap(nb);
    auto free_all_but = [&](const std::vector<char> &u) {
        F->free_hvy.clear();
        for (int h = F->nb_max - 1; h >= 0; --h)
            if (!u[h]) F->free_hvy.push_back(F->slot0 + h);
    };
    free_all_but(used);
    if (!F->rebuild_map()) {
        // two blocks on one (level, treecode): the forest is left empty
        F->blocks.clear();
        free_all_but(std::vector<char>(F->nb_max, 0));
        F->rebuild_map();
        wh::set_error("wh_forest_set_blocks: duplicate (level, treecode)");
        return WH_E_ARG;
    }
    return WH_OK;
}

int wh_forest_size(const wh_forest *F, int *n)
{
    if (!F || !n) return WH_E_ARG;
    *n = (int)F->blocks.size();
    return WH_OK;
}

int wh_forest_get_blocks(const wh_forest *F, int cap, int64_t *tc, int32_t *level, int32_t *hvy,
                         double *x0dx)
{
    if (!F || cap < (int)F->blocks.size()) {
        wh::set_error("wh_forest_get_blocks: capacity too small");
        return WH_E_ARG;
    }
    for (size_t i = 0; i < F->blocks.size(); ++i) {
        const Block &b = F->blocks[i];
        if (tc) tc[i] = b.tc;
        if (level) level[i] = b.level;
        if (hvy) hvy[i] = b.hvy;
        if (x0dx) F->x0dx(b, x0dx + 6 * i);
    }
    return WH_OK;
}

// Ghost patches of a leaf forest (sync_ghosts_RHS_tree "full_leaf" and
// sync_ghosts_tree "full" on leaf grids, synchronize_ghosts_generic.f90:352-694):
// for every block and each of the 3^dim-1 directions the ghost box of depth
// gminus/gplus is filled from the leaf that covers it: a same-level neighbour
// (COPY, stage 1), the finer neighbours adjacent to the box (RESTRICT, stage 2:
// decimation), or the coarser neighbour (PREDICT, stage 3: interpolation of
// order `order`, 2/4/6).  aux of a PREDICT patch carries the order.
int wh_forest_patches(const wh_forest *F, int gminus, int gplus, int order, wh_patch *out,
                      int cap, int *npatch)
{
    WH_PROF("forest_patches");
    if (!F || !npatch || gminus < 0 || gplus < 0 || gminus > F->g || gplus > F->g ||
        (order != 2 && order != 4 && order != 6)) {
        wh::set_error("wh_forest_patches: bad argument");
        return WH_E_ARG;
    }
    if (F->pc_version == F->version && F->pc_gm == gminus && F->pc_gp == gplus && F->pc_order == order) {
        const int n = (int)F->pcache.size();
        *npatch = n;
        if (out) {
            if (cap < n) {
                wh::set_error("wh_forest_patches: output too small");
                return WH_E_ARG;
            }
            std::copy(F->pcache.begin(), F->pcache.end(), out);
        }
        return WH_OK;
    }
    const int dim = F->dim, g = F->g;
    const int N = order / 2;
    // the lists name blocks by their index in the block list: a block list
    // equal in treecodes and levels to the last generated one reuses them
    std::vector<int64_t> key = {gminus, gplus, order, dim, g, F->bs[0], F->bs[1], F->bs[2], F->jmax,
                                F->periodic[0], F->periodic[1], F->periodic[2], (int64_t)F->blocks.size()};
    key.reserve(key.size() + 2 * F->blocks.size());
    for (const Block &b : F->blocks) {
        key.push_back(b.tc);
        key.push_back(b.level);
    }
    auto finish = [&](const std::vector<wh_patch> &idx) -> int {
        std::vector<wh_patch> P(idx);
        for (wh_patch &p : P) {
            p.recv = F->blocks[p.recv].hvy;
            p.send = F->blocks[p.send].hvy;
        }
        *npatch = (int)P.size();
        if (out) {
            if (cap < (int)P.size()) {
                wh::set_error("wh_forest_patches: capacity too small");
                return WH_E_ARG;
            }
            std::copy(P.begin(), P.end(), out);
        }
        F->pcache.swap(P);
        F->pc_gm = gminus;
        F->pc_gp = gplus;
        F->pc_order = order;
        F->pc_version = F->version;
        return WH_OK;
    };
    if (key == F->pc_key) return finish(F->pc_idx);
    F->pc_key.clear();
    std::vector<wh_patch> P;
    std::vector<wh_patch> fp;
    int dirs[27][3], nd = 0;
    for (int oz = (dim == 3 ? -1 : 0); oz <= (dim == 3 ? 1 : 0); ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                if (!ox && !oy && !oz) continue;
                dirs[nd][0] = ox;
                dirs[nd][1] = oy;
                dirs[nd][2] = oz;
                ++nd;
            }
    const int nxyz[3] = {F->bs[0] + 2 * g, F->bs[1] + 2 * g, dim == 3 ? F->bs[2] + 2 * g : 1};
    for (int bi = 0; bi < (int)F->blocks.size(); ++bi) {
        const Block &b = F->blocks[bi];
        int ijk[3];
        F->ijk_of(b.tc, b.level, ijk);
        for (int q = 0; q < nd; ++q) {
            const int *o = dirs[q];
            int rlo[3], rn[3];
            ghost_box(*F, o, gminus, gplus, rlo, rn);
            if (rn[0] <= 0 || rn[1] <= 0 || rn[2] <= 0) continue;
            int nb[3] = {ijk[0] + o[0], ijk[1] + o[1], ijk[2] + o[2]}, w[3];
            // 1) same level
            int s = F->find(b.level, nb, w);
            if (s >= 0) {
                wh_patch p{};
                p.recv = bi;  // block indices, mapped to hvy ids by finish
                p.send = s;
                for (int d = 0; d < 3; ++d) {
                    p.recv_lo[d] = rlo[d];
                    p.n[d] = rn[d];
                    // receiver index i <-> sender index i - o*Bs
                    p.send_lo[d] = d < dim ? rlo[d] - o[d] * F->bs[d] : 0;
                }
                p.op = WH_PATCH_COPY;
                p.stage = 1;
                P.push_back(p);
                continue;
            }
            // 2) finer neighbours: children of the virtual same-level neighbour
            //    that touch the receiver box
            bool finer = false;
            if (b.level < F->jmax) {
                int lo_c[3], hi_c[3];  // child index range (0/1) per dim
                for (int d = 0; d < 3; ++d) {
                    if (d >= dim) {
                        lo_c[d] = hi_c[d] = 0;
                    } else if (o[d] < 0) {
                        lo_c[d] = hi_c[d] = 1;  // the child adjacent to the receiver
                    } else if (o[d] > 0) {
                        lo_c[d] = hi_c[d] = 0;
                    } else {
                        lo_c[d] = 0;
                        hi_c[d] = 1;
                    }
                }
                fp.clear();
                bool all = true;
                for (int cz = lo_c[2]; cz <= hi_c[2] && all; ++cz)
                    for (int cy = lo_c[1]; cy <= hi_c[1] && all; ++cy)
                        for (int cx = lo_c[0]; cx <= hi_c[0] && all; ++cx) {
                            const int cc[3] = {cx, cy, cz};
                            int cij[3];
                            for (int d = 0; d < 3; ++d) cij[d] = d < dim ? 2 * nb[d] + cc[d] : 0;
                            int cw[3];
                            const int f = F->find(b.level + 1, cij, cw);
                            if (f < 0) {
                                all = false;
                                break;
                            }
                            wh_patch p{};
                            p.recv = bi;
                            p.send = f;
                            bool ok = true;
                            for (int d = 0; d < 3; ++d) {
                                if (d >= dim) {
                                    p.recv_lo[d] = 0;
                                    p.n[d] = 1;
                                    p.send_lo[d] = 0;
                                    continue;
                                }
                                int lo = rlo[d], n = rn[d];
                                if (o[d] == 0) {  // half of the face covered by this child
                                    const int h = F->bs[d] / 2;
                                    lo = g + cc[d] * h;
                                    n = h;
                                }
                                p.recv_lo[d] = lo;
                                p.n[d] = n;
                                // receiver ghost-incl index i -> position (i-g)*dxc relative to
                                // receiver origin; child origin relative to receiver origin in
                                // fine units: (2*nb - 2*ijk + c)*Bs ... expressed per dim
                                const int child_off_f = (2 * o[d] + cc[d]) * F->bs[d];  // fine pts
                                // fine index in child (ghost-incl) = 2*(i-g) - child_off_f + g
                                p.send_lo[d] = 2 * (lo - g) - child_off_f + g;
                                if (p.send_lo[d] < 0 || p.send_lo[d] + 2 * (n - 1) >= nxyz[d]) ok = false;
                            }
                            if (!ok) {
                                all = false;
                                break;
                            }
                            p.op = WH_PATCH_RESTRICT;
                            p.stage = 2;
                            fp.push_back(p);
                        }
                if (all && !fp.empty()) {
                    P.insert(P.end(), fp.begin(), fp.end());
                    finer = true;
                }
            }
            if (finer) continue;
            // 3) coarser neighbour
            if (b.level > 0) {
                int cij[3];
                for (int d = 0; d < 3; ++d) cij[d] = d < dim ? (int)std::floor(nb[d] / 2.0) : 0;
                int cw[3];
                const int c = F->find(b.level - 1, cij, cw);
                if (c >= 0) {
                    wh_patch p{};
                    p.recv = bi;
                    p.send = c;
                    bool ok = true;
                    for (int d = 0; d < 3; ++d) {
                        p.recv_lo[d] = rlo[d];
                        p.n[d] = rn[d];
                        if (d >= dim) {
                            p.send_lo[d] = 0;
                            continue;
                        }
                        // receiver origin relative to coarse origin, in fine points
                        const int off_f = ijk[d] * F->bs[d] - 2 * cij[d] * F->bs[d];
                        // fine index in the coarse block's doubled ghost-incl index space
                        p.send_lo[d] = (rlo[d] - g) + off_f + 2 * g;
                        const int fmin = p.send_lo[d], fmax = p.send_lo[d] + rn[d] - 1;
                        // even f copies coarse f/2, odd f reads f/2-N+1 .. f/2+N
                        const int omin = (fmin & 1) ? fmin : fmin + 1;
                        const int omax = (fmax & 1) ? fmax : fmax - 1;
                        if (fmin < 0 || (fmax >> 1) >= nxyz[d]) ok = false;
                        if (omin <= omax && ((omin >> 1) - N + 1 < 0 || (omax >> 1) + N >= nxyz[d]))
                            ok = false;
                    }
                    if (!ok) {
                        wh::set_error("wh_forest_patches: prediction stencil leaves the coarse block "
                                      "(too few ghost nodes for this predictor)");
                        return WH_E_ARG;
                    }
                    p.op = WH_PATCH_PREDICT;
                    p.stage = 3;
                    p.aux = order;
                    P.push_back(p);
                }
            }
            // no cover: non-periodic domain boundary, left to boundary conditions
        }
    }
    F->pc_idx.swap(P);
    F->pc_key.swap(key);
    return finish(F->pc_idx);
}

// Refine the listed blocks (refinementExecute.f90:4-116 bookkeeping): each
// gets 2^dim daughters on fresh hvy ids (daughter k has digit k at level+1);
// the mother leaves the active list and its slot returns to the free list.
// pairs_out[3*i] = mother hvy, [3*i+1] = daughter hvy, [3*i+2] = daughter digit.
int wh_forest_refine(wh_forest *F, int n, const int32_t *hvy_list, int32_t *pairs_out, int cap,
                     int *npairs)
{
    WH_PROF("forest_refine");
    if (!F || n < 0 || (n && !hvy_list) || !npairs) return WH_E_ARG;
    const int nd = 1 << F->dim;
    if ((int)F->free_hvy.size() < n * nd || cap < n * nd) {
        wh::set_error("wh_forest_refine: pool out of memory (or output too small)");
        return WH_E_NOMEM;
    }
    // block index by pool slot (the forest's slots are [slot0, slot0 + nb_max))
    std::vector<int> by_hvy((size_t)F->nb_max, -1);
    for (size_t i = 0; i < F->blocks.size(); ++i) {
        const int h = F->blocks[i].hvy - F->slot0;
        if (h >= 0 && h < F->nb_max) by_hvy[h] = (int)i;
    }
    // every entry is checked before a slot is taken: a rejected list leaves
    // the forest (blocks and free slots) as it was
    std::vector<char> remove(F->blocks.size(), 0);
    for (int i = 0; i < n; ++i) {
        const int h = hvy_list[i] - F->slot0;
        const int bi = (h >= 0 && h < F->nb_max) ? by_hvy[h] : -1;
        if (bi < 0 || remove[bi]) {
            wh::set_error("wh_forest_refine: block not active or listed twice");
            return WH_E_ARG;
        }
        if (F->blocks[bi].level >= F->jmax) {
            wh::set_error("wh_forest_refine: block already at Jmax");
            return WH_E_ARG;
        }
        remove[bi] = 1;
    }
    std::vector<Block> added;
    added.reserve((size_t)n * nd);
    int np = 0;
    for (int i = 0; i < n; ++i) {
        const Block m = F->blocks[by_hvy[hvy_list[i] - F->slot0]];
        for (int k = 0; k < nd; ++k) {
            const int h = F->free_hvy.back();
            F->free_hvy.pop_back();
            const int64_t tc = m.tc | ((int64_t)k << F->digit_shift(m.level + 1));
            added.push_back(Block{tc, m.level + 1, h, 34});  // REF_FRESHLY_REFINED
            pairs_out[3 * np] = m.hvy;
            pairs_out[3 * np + 1] = h;
            pairs_out[3 * np + 2] = k;
            ++np;
        }
    }
    std::vector<Block> keep;
    for (size_t i = 0; i < F->blocks.size(); ++i)
        if (!remove[i]) keep.push_back(F->blocks[i]);
        else F->free_hvy.push_back(F->blocks[i].hvy);
    keep.insert(keep.end(), added.begin(), added.end());
    F->blocks.swap(keep);
    F->rebuild_map();
    *npairs = np;
    return WH_OK;
}

}  // extern "C"
