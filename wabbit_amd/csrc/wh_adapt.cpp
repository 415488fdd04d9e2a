// wh_adapt.cpp -- adapt_tree (LIB/MESH/adapt_tree.f90:11-255) and
// sync_ghosts_tree with filtered restriction, orchestrated on the host over
// the pool's device operations (wavelet transforms, coarse-extension boxes,
// ghost-patch lists).  Host code only.
//
// The full tree (leaves plus their mothers down to Jmin) lives in a Tree
// built from the forest's leaf list; mothers take free pool slots.  Ghost
// fills are generated per receiver and direction from the block covering
// the ghost box: the same-level block (COPY), else the finer blocks touching
// the box (RESTRICT), else the coarser leaf (PREDICT), gated by the sender
// and receiver rules of each sync case (prepare_ghost_synch_metadata,
// synchronize_ghosts_generic.f90:352-694) and reading the source array the
// reference's xfer_block_data.f90:321-437 picks (hvy_block or hvy_tmp).
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wh_forest_impl.h"
#include "wh_internal.h"
#include "wh_prof.h"

namespace wh {
namespace {

constexpr int REF_TMP_EMPTY = 10;          // module_globals.f90:25
constexpr int REF_UNSIGNIFICANT_STAY = 9;  // module_globals.f90:28
constexpr int A_U = WH_BLOCK, A_T = WH_TMP, A_FILT = WH_WORK + 1;  // work slot 0: scratch of the transforms

struct TNode {
    int64_t tc;
    int level;
    int hvy;
    int status;
    int q[3];  // grid index of the block at its level (from tc)
};

// hash key of (level, grid index); indices < 2^19 per direction
inline uint64_t qkey(int level, const int q[3])
{
    return (uint64_t)level | ((uint64_t)q[0] << 5) | ((uint64_t)q[1] << 24) | ((uint64_t)q[2] << 43);
}

enum SrcMode { SRC_U = 0, SRC_NOT_FLAG = 1, SRC_DECOMPOSED = 2 };
struct SyncSpec {
    bool same = true, restrict_ = true, predict = true;
    bool recv_ref_on = false, recv_level_on = false;
    int recv_ref = 0, recv_level = 0;
    SrcMode src = SRC_U;
    int src_ref = -1;
    bool filter = true;
    int gm = 0, gp = 0;
};

int dirs_of(int dim, int dirs[26][3])
{
    int nd = 0;
    for (int oz = (dim == 3 ? -1 : 0); oz <= (dim == 3 ? 1 : 0); ++oz)
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

// The device operations of the orchestration: launched at once, or logged
// per tree (Tree::log) and replayed for several trees of one pool together
// (replay): round r launches the r-th operation of every tree, operations of
// one kind and arguments as one launch over their concatenated lists.  The
// trees' blocks are disjoint slots of the pool's arrays, so only each tree's
// own order matters.
enum OpKind { OP_BLOCKS = 0, OP_BOXES = 1, OP_PATCHES = 2 };
struct DevOp {
    int kind = 0, a = 0, b = 0, c = 0;  // BLOCKS: op, src, dst; BOXES: array, src; PATCHES: array, alt1, alt2
    std::vector<int32_t> ids;
    std::vector<wh_box> boxes;
    std::vector<wh_patch> patches;
};
struct OpLog {
    std::vector<DevOp> ops;
};

struct Tree {
    OpLog *log = nullptr;
    wh_forest *F;
    wh_pool *P;
    wh_wavelet w;
    int dim, g, bs[3], jmax;
    std::vector<TNode> b;
    // (level, grid index) -> node: a dense table per level while it has at
    // most 2^20 cells, the hash map beyond
    std::unordered_map<uint64_t, int> map;
    std::vector<std::vector<int>> dense;
    static constexpr int DENSE_BITS = 20;

    Tree(wh_forest *f, wh_pool *p) : F(f), P(p), w(p->ws.w), dim(f->dim), g(f->g), jmax(f->jmax)
    {
        WH_PROF("tree_build");
        for (int d = 0; d < 3; ++d) bs[d] = f->bs[d];
        nd = dirs_of(dim, dirs);
        b.reserve(2 * f->blocks.size());
        for (const FBlock &x : f->blocks) b.push_back(node(x.tc, x.level, x.hvy, 0));
        rehash();
    }
    TNode node(int64_t tc, int level, int hvy, int status) const
    {
        TNode n{tc, level, hvy, status, {0, 0, 0}};
        F->ijk_of(tc, level, n.q);
        return n;
    }
    size_t lin(int level, const int q[3]) const
    {
        return ((size_t)q[2] << (2 * level)) | ((size_t)q[1] << level) | (size_t)q[0];
    }
    void insert(int i)
    {
        const int L = b[i].level;
        if (L < (int)dense.size() && !dense[L].empty()) dense[L][lin(L, b[i].q)] = i;
        else map[qkey(L, b[i].q)] = i;
    }
    void rehash()
    {
        map.clear();
        int lmax = 0;
        for (const TNode &x : b) lmax = std::max(lmax, x.level);
        dense.resize(lmax + 1);
        for (int L = 0; L <= lmax; ++L) {
            if (dim * L <= DENSE_BITS) dense[L].assign((size_t)1 << (dim * L), -1);
            else dense[L].clear();
        }
        for (size_t i = 0; i < b.size(); ++i) insert((int)i);
        topo_ok = false;
    }
    void ijk(int i, int out[3]) const
    {
        out[0] = b[i].q[0];
        out[1] = b[i].q[1];
        out[2] = b[i].q[2];
    }
    int find(int level, const int q[3]) const
    {
        if (level < 0) return -1;
        const int m = 1 << level;
        int wq[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            if (!F->periodic[d] && (q[d] < 0 || q[d] >= m)) return -1;
            wq[d] = q[d] & (m - 1);  // periodic wrap; m is a power of two
        }
        if (level < (int)dense.size() && !dense[level].empty()) return dense[level][lin(level, wq)];
        auto it = map.find(qkey(level, wq));
        return it == map.end() ? -1 : it->second;
    }
    bool valid(int i) const { return i >= 0 && b[i].status != REF_TMP_EMPTY; }

    // ---- relations, uncached (used while the block set changes) ----
    int mother_raw(int i) const
    {
        if (b[i].level == 0) return -1;
        const int *q = b[i].q;
        const int m[3] = {q[0] >> 1, q[1] >> 1, q[2] >> 1};
        return find(b[i].level - 1, m);
    }
    int daughter_raw(int i, int k) const
    {
        const int *q = b[i].q;
        const int d[3] = {2 * q[0] + ((k >> 1) & 1), 2 * q[1] + (k & 1), dim == 3 ? 2 * q[2] + ((k >> 2) & 1) : 0};
        return find(b[i].level + 1, d);
    }
    // coarse relation in direction o (find_neighbors.f90): exists only on the
    // mother's boundary in that direction
    int nb_coarse_raw(int i, const int o[3]) const
    {
        if (b[i].level == 0) return -1;
        const int *q = b[i].q;
        int c[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            if (o[d] > 0 && (q[d] & 1) != 1) return -1;
            if (o[d] < 0 && (q[d] & 1) != 0) return -1;
            const int v = q[d] + o[d];
            c[d] = v >= 0 ? v / 2 : -((-v + 1) / 2);
        }
        return find(b[i].level - 1, c);
    }

    // ---- relations, cached per block set ----
    // Per block, built on its first use after the block set changed (add,
    // removal): per direction the same-level neighbour, the coarse relation,
    // the coarser block covering the ghost box (floor((q+o)/2) one level up)
    // and the finer blocks adjacent in that direction; mother and daughters.
    int nd = 0;
    int dirs[26][3];
    mutable bool topo_ok = false;
    mutable std::vector<char> t_done;
    mutable std::vector<int> t_same, t_coarse, t_cover, t_mom, t_dau, t_fcnt;
    mutable std::vector<std::pair<int, std::array<int, 3>>> t_fin;  // maxc slots per block and direction
    int dir_index(const int o[3]) const
    {
        const int t = dim == 3 ? (o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1) : (o[1] + 1) * 3 + (o[0] + 1);
        const int c = dim == 3 ? 13 : 4;
        return t > c ? t - 1 : t;
    }
    int maxc() const { return 1 << (dim - 1); }
    void topo() const
    {
        if (topo_ok) return;
        const size_t n = b.size();
        t_done.assign(n, 0);
        t_same.resize(n * nd);
        t_coarse.resize(n * nd);
        t_cover.resize(n * nd);
        t_mom.resize(n);
        t_dau.resize(n << dim);
        t_fcnt.resize(n * nd);
        t_fin.resize(n * nd * maxc());
        topo_ok = true;
    }
    void node_topo(int i) const
    {
        topo();
        if (t_done[i]) return;
        const int *q = b[i].q, L = b[i].level, ns = 1 << dim, mc = maxc();
        t_mom[i] = mother_raw(i);
        for (int k = 0; k < ns; ++k) t_dau[(size_t)i * ns + k] = daughter_raw(i, k);
        for (int k = 0; k < nd; ++k) {
            const int *o = dirs[k];
            const size_t e = (size_t)i * nd + k;
            const int nq[3] = {q[0] + o[0], q[1] + o[1], q[2] + o[2]};
            t_same[e] = find(L, nq);
            t_coarse[e] = nb_coarse_raw(i, o);
            t_cover[e] = -1;
            if (L > 0) {
                int cu[3] = {0, 0, 0};
                for (int d = 0; d < dim; ++d) {
                    const int v = q[d] + o[d];
                    cu[d] = v >= 0 ? v / 2 : -((-v + 1) / 2);
                }
                t_cover[e] = find(L - 1, cu);
            }
            int lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
                if (d >= dim) lo[d] = hi[d] = 0;
                else if (o[d] < 0) lo[d] = hi[d] = 1;
                else if (o[d] > 0) lo[d] = hi[d] = 0;
                else {
                    lo[d] = 0;
                    hi[d] = 1;
                }
            }
            int cnt = 0;
            for (int cz = lo[2]; cz <= hi[2]; ++cz)
                for (int cy = lo[1]; cy <= hi[1]; ++cy)
                    for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                        const int c[3] = {2 * (q[0] + o[0]) + cx, 2 * (q[1] + o[1]) + cy,
                                          dim == 3 ? 2 * (q[2] + o[2]) + cz : 0};
                        t_fin[e * mc + cnt++] = {find(L + 1, c), {cx, cy, cz}};
                    }
            t_fcnt[e] = cnt;
        }
        t_done[i] = 1;
    }
    int mother(int i) const
    {
        node_topo(i);
        return t_mom[i];
    }
    int daughter(int i, int k) const
    {
        node_topo(i);
        return t_dau[((size_t)i << dim) + k];
    }
    // block_is_leaf (block_relations.f90); check_empty: REF_TMP_EMPTY daughters don't count
    bool is_leaf(int i, bool check_empty) const
    {
        node_topo(i);
        const int ns = 1 << dim;
        for (int k = 0; k < ns; ++k) {
            const int d = t_dau[(size_t)i * ns + k];
            if (d >= 0 && (!check_empty || b[d].status != REF_TMP_EMPTY)) return false;
        }
        return true;
    }
    int nb_same(int i, const int o[3]) const
    {
        node_topo(i);
        return t_same[(size_t)i * nd + dir_index(o)];
    }
    // the coarser block covering the ghost box of block i in direction o
    int cover(int i, const int o[3]) const
    {
        node_topo(i);
        return t_cover[(size_t)i * nd + dir_index(o)];
    }
    // the finer blocks of block i in direction index k, as a range of the table
    std::pair<const std::pair<int, std::array<int, 3>> *, const std::pair<int, std::array<int, 3>> *>
    fin_range(int i, int k) const
    {
        node_topo(i);
        const size_t e = (size_t)i * nd + k;
        const auto *p = t_fin.data() + e * maxc();
        return {p, p + t_fcnt[e]};
    }
    // the finer blocks adjacent to block i in direction o (children of the
    // virtual same-level neighbour); child cc per dimension in out
    void finer_children(int i, const int o[3], std::vector<std::pair<int, std::array<int, 3>>> &out) const
    {
        const auto r = fin_range(i, dir_index(o));
        out.assign(r.first, r.second);
    }
    bool any_finer(int i, const int o[3], bool need_valid) const
    {
        const auto r = fin_range(i, dir_index(o));
        for (auto *p = r.first; p != r.second; ++p)
            if (p->first >= 0 && (!need_valid || valid(p->first))) return true;
        return false;
    }
    int nb_coarse(int i, const int o[3]) const
    {
        node_topo(i);
        return t_coarse[(size_t)i * nd + dir_index(o)];
    }
    int add(int64_t tc, int level, int status)
    {
        WH_CHECK(!F->free_hvy.empty(), -1, "adapt_tree: pool out of slots for the full tree");
        const int h = F->free_hvy.back();
        F->free_hvy.pop_back();
        b.push_back(node(tc, level, h, status));
        insert((int)b.size() - 1);
        topo_ok = false;
        return (int)b.size() - 1;
    }
    void remove_marked(const std::vector<char> &del)
    {
        std::vector<TNode> keep;
        for (size_t i = 0; i < b.size(); ++i) {
            if (del[i]) F->free_hvy.push_back(b[i].hvy);
            else keep.push_back(b[i]);
        }
        b.swap(keep);
        rehash();
    }
    int max_level() const
    {
        int m = -1;
        for (const TNode &x : b) m = std::max(m, x.level);
        return m;
    }
};


int dev_blocks(Tree &T, int op, int src, int dst, std::vector<int32_t> ids)
{
    if (!T.log) return run_blocks_op(T.P, op, src, dst, ids);
    if (ids.empty()) return WH_OK;
    DevOp o;
    o.kind = OP_BLOCKS;
    o.a = op;
    o.b = src;
    o.c = dst;
    o.ids = std::move(ids);
    T.log->ops.push_back(std::move(o));
    return WH_OK;
}

int dev_boxes(Tree &T, int array, int src, std::vector<wh_box> boxes)
{
    if (!T.log) return run_boxes(T.P, array, src, boxes);
    if (boxes.empty()) return WH_OK;
    DevOp o;
    o.kind = OP_BOXES;
    o.a = array;
    o.b = src;
    o.boxes = std::move(boxes);
    T.log->ops.push_back(std::move(o));
    return WH_OK;
}

int dev_patches(Tree &T, int array, int alt1, int alt2, std::vector<wh_patch> patches)
{
    if (!T.log) return run_patch_list(T.P, array, alt1, alt2, patches);
    if (patches.empty()) return WH_OK;
    DevOp o;
    o.kind = OP_PATCHES;
    o.a = array;
    o.b = alt1;
    o.c = alt2;
    o.patches = std::move(patches);
    T.log->ops.push_back(std::move(o));
    return WH_OK;
}

int replay(wh_pool *P, const std::vector<OpLog *> &logs)
{
    WH_PROF("replay");
    size_t rounds = 0;
    for (const OpLog *l : logs) rounds = std::max(rounds, l->ops.size());
    int rc = WH_OK;
    for (size_t r = 0; r < rounds && !rc; ++r) {
        std::vector<DevOp> merged;
        for (OpLog *l : logs) {
            if (r >= l->ops.size()) continue;
            DevOp &o = l->ops[r];
            DevOp *m = nullptr;
            for (DevOp &x : merged)
                if (x.kind == o.kind && x.a == o.a && x.b == o.b && x.c == o.c) {
                    m = &x;
                    break;
                }
            if (!m) {
                merged.push_back(std::move(o));
                continue;
            }
            m->ids.insert(m->ids.end(), o.ids.begin(), o.ids.end());
            m->boxes.insert(m->boxes.end(), o.boxes.begin(), o.boxes.end());
            m->patches.insert(m->patches.end(), o.patches.begin(), o.patches.end());
        }
        for (const DevOp &m : merged) {
            if (m.kind == OP_BLOCKS) rc = run_blocks_op(P, m.a, m.b, m.c, m.ids);
            else if (m.kind == OP_BOXES) rc = run_boxes(P, m.a, m.b, m.boxes);
            else rc = run_patch_list(P, m.a, m.b, m.c, m.patches);
            if (rc) break;
        }
    }
    for (OpLog *l : logs) l->ops.clear();
    return rc;
}

// ---- coarse extension boxes ------------------------------------------------
// get_indices_of_modify_patch (neighborhood.f90) with lvl_diff = +1: the
// first Nl / last Nr interior points toward direction o, full interior across
void modify_box(const Tree &T, const int o[3], int Nl, int Nr, wh_box &bx)
{
    for (int d = 0; d < 3; ++d) {
        if (d >= T.dim) {
            bx.lo[d] = bx.hi[d] = 0;
            continue;
        }
        bx.lo[d] = T.g;
        bx.hi[d] = T.g + T.bs[d] - 1;
        if (o[d] < 0) bx.hi[d] = T.g + Nl - 1;
        if (o[d] > 0) bx.lo[d] = T.g + T.bs[d] - Nr;
    }
}
// get_indices_of_ghost_patch with lvl_diff = +1 for the coarse relation of
// block i in direction o: the ghost slab, extended across each free direction
// toward the sister side of the block
void ce_ghost_box(const Tree &T, int i, const int o[3], wh_box &bx)
{
    int q[3];
    T.ijk(i, q);
    for (int d = 0; d < 3; ++d) {
        if (d >= T.dim) {
            bx.lo[d] = bx.hi[d] = 0;
            continue;
        }
        const int B = T.bs[d], g = T.g;
        if (o[d] < 0) {
            bx.lo[d] = 0;
            bx.hi[d] = g - 1;
        } else if (o[d] > 0) {
            bx.lo[d] = g + B;
            bx.hi[d] = B + 2 * g - 1;
        } else if (q[d] & 1) {
            bx.lo[d] = 0;
            bx.hi[d] = g + B - 1;
        } else {
            bx.lo[d] = g;
            bx.hi[d] = B + 2 * g - 1;
        }
    }
}

enum CeCase { CE_TREE = 0, CE_LEVEL = 1, CE_REF = 2 };
// coarse_extension_modify (reconstruction_step.f90:3-100) on WH_BLOCK
// The cache of generated lists (TreeSyncCache) of a tree's forest: the key
// is tag and parameters (spec) followed by every node's treecode, level and
// status in order; the lists name blocks by node index.
void tree_key(const Tree &T, std::initializer_list<int64_t> spec, std::vector<int64_t> &key)
{
    key.assign(spec.begin(), spec.end());
    const int64_t geo[] = {T.w.order, T.w.lifted, T.w.Nscl, T.w.Nscr, T.w.Nwcl, T.w.Nwcr, T.dim, T.g,
                           T.bs[0], T.bs[1], T.bs[2], T.P->nx, T.P->ny, T.P->nz, (int64_t)T.b.size()};
    key.insert(key.end(), geo, geo + sizeof(geo) / sizeof(geo[0]));
    key.reserve(key.size() + 2 * T.b.size());
    for (const TNode &x : T.b) {
        key.push_back(x.tc);
        key.push_back(((int64_t)x.level << 32) | (int64_t)(uint32_t)x.status);
    }
}

TreeSyncCache *cache_find(Tree &T, const std::vector<int64_t> &key)
{
    for (TreeSyncCache &c : T.F->tscache)
        if (c.key == key) {
            c.used = ++T.F->ts_clock;
            return &c;
        }
    return nullptr;
}

// stores an entry (the least recently used one dropped beyond TS_CACHE)
TreeSyncCache &cache_put(Tree &T, TreeSyncCache &&c)
{
    std::vector<TreeSyncCache> &cache = T.F->tscache;
    c.used = ++T.F->ts_clock;
    if ((int)cache.size() < wh_forest::TS_CACHE) {
        cache.push_back(std::move(c));
        return cache.back();
    }
    TreeSyncCache *lru = &cache[0];
    for (TreeSyncCache &e : cache)
        if (e.used < lru->used) lru = &e;
    *lru = std::move(c);
    return *lru;
}

// node indices of cached boxes -> the nodes' pool slots
std::vector<wh_box> boxes_of(const Tree &T, const std::vector<wh_box> &src)
{
    std::vector<wh_box> out(src);
    for (wh_box &bx : out) bx.hvy = T.b[bx.hvy].hvy;
    return out;
}

int ce_modify(Tree &T, CeCase ce, int s_level, int s_ref, bool clear_wc, bool copy_sc)
{
    WH_PROF("ce_modify");
    std::vector<int64_t> key;
    tree_key(T, {2, (int64_t)ce, s_level, s_ref, clear_wc, copy_sc}, key);
    if (const TreeSyncCache *c = cache_find(T, key)) return dev_boxes(T, A_U, A_T, boxes_of(T, c->fboxes[0]));
    int dirs[26][3];
    const int nd = dirs_of(T.dim, dirs);
    std::vector<wh_box> boxes;
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (ce == CE_LEVEL && T.b[i].level != s_level) continue;
        if (ce == CE_REF && T.b[i].status != s_ref) continue;
        if (!T.is_leaf(i, false)) continue;
        for (int q = 0; q < nd; ++q) {
            const int *o = dirs[q];
            if (T.nb_coarse(i, o) < 0) continue;
            if (T.valid(T.nb_same(i, o)) || T.any_finer(i, o, true)) continue;
            wh_box bx{};
            bx.hvy = i;  // node index, mapped to its slot on use
            if (clear_wc) {
                bx.op = WH_BOX_ZERO_WC;
                modify_box(T, o, T.w.Nwcl, T.w.Nwcr, bx);
                boxes.push_back(bx);
                ce_ghost_box(T, i, o, bx);
                boxes.push_back(bx);
            }
            if (copy_sc) {
                bx.op = WH_BOX_COPY_SC;
                modify_box(T, o, T.w.Nscl, T.w.Nscr, bx);
                boxes.push_back(bx);
            }
        }
    }
    TreeSyncCache e;
    e.key = std::move(key);
    e.fboxes[0] = std::move(boxes);
    return dev_boxes(T, A_U, A_T, boxes_of(T, cache_put(T, std::move(e)).fboxes[0]));
}

// ---- ghost syncs -------------------------------------------------------------
int src_sel(const Tree &T, const SyncSpec &S, int sender, bool res_pred)
{
    if (S.src == SRC_NOT_FLAG && T.b[sender].status != S.src_ref) return 1;
    if (S.src == SRC_DECOMPOSED && res_pred) return 1;
    return 0;
}

// launches a cached sync: its node indices mapped to the nodes' pool slots
int tree_sync_run(Tree &T, const TreeSyncCache &c)
{
    auto patches = [&](const std::vector<wh_patch> &src) {
        std::vector<wh_patch> out(src);
        for (wh_patch &p : out) {
            p.recv = T.b[p.recv].hvy;
            p.send = T.b[p.send].hvy;
        }
        return out;
    };
    if (int rc = dev_patches(T, A_U, A_T, A_FILT, patches(c.st1))) return rc;
    // restrict_copy_at_CE's filtered senders, where the sync has any
    if (!c.fids[0].empty() || !c.fids[1].empty()) {
        for (int k = 0; k < 2; ++k) {
            const int arr = k == 0 ? A_U : A_T;
            std::vector<int32_t> ids(c.fids[k]);
            for (int32_t &i : ids) i = T.b[i].hvy;
            if (int rc = dev_blocks(T, WOP_FILTER, arr, A_FILT, std::move(ids))) return rc;
            if (int rc = dev_boxes(T, A_FILT, arr, boxes_of(T, c.fboxes[k]))) return rc;
        }
    }
    return dev_patches(T, A_U, A_T, A_FILT, patches(c.st23));
}

int tree_sync(Tree &T, const SyncSpec &S)
{
    WH_PROF("tree_sync");
    std::vector<int64_t> key;
    tree_key(T, {1, S.same, S.restrict_, S.predict, S.recv_ref_on, S.recv_level_on, S.recv_ref, S.recv_level,
                 (int64_t)S.src, S.src_ref, S.filter, S.gm, S.gp},
             key);
    if (const TreeSyncCache *c = cache_find(T, key)) {
        WH_PROF("tree_sync.hit");
        return tree_sync_run(T, *c);
    }
    const int dim = T.dim, g = T.g, order = T.w.order;
    const bool filt = S.filter && T.w.lifted;
    int dirs[26][3];
    const int nd = dirs_of(dim, dirs);
    const int nxyz[3] = {T.P->nx, T.P->ny, T.P->nz};
    std::vector<wh_patch> st1, st23;
    std::vector<int> restrict_senders;  // tree indices
    std::unique_ptr<ProfScope> gen(new ProfScope("tree_sync.gen"));
    T.topo();
    st1.reserve(T.b.size() * nd);
    st23.reserve(T.b.size() * 2);
    for (int r = 0; r < (int)T.b.size(); ++r) {
        const TNode &R = T.b[r];
        if (R.status == REF_TMP_EMPTY) continue;
        if (S.recv_ref_on && R.status != S.recv_ref) continue;
        if (S.recv_level_on && R.level != S.recv_level) continue;
        int q[3];
        T.ijk(r, q);
        for (int k = 0; k < nd; ++k) {
            const int *o = dirs[k];
            int rlo[3], rn[3];
            for (int d = 0; d < 3; ++d) {
                if (d >= dim) {
                    rlo[d] = 0;
                    rn[d] = 1;
                } else if (o[d] < 0) {
                    rlo[d] = g - S.gm;
                    rn[d] = S.gm;
                } else if (o[d] > 0) {
                    rlo[d] = g + T.bs[d];
                    rn[d] = S.gp;
                } else {
                    rlo[d] = g;
                    rn[d] = T.bs[d];
                }
            }
            if (rn[0] <= 0 || rn[1] <= 0 || rn[2] <= 0) continue;
            // 1) same level
            const int ns = T.nb_same(r, o);
            if (T.valid(ns)) {
                if (!S.same) continue;
                wh_patch p{};
                p.recv = r;  // node indices; tree_sync_run maps them to slots
                p.send = ns;
                for (int d = 0; d < 3; ++d) {
                    p.recv_lo[d] = rlo[d];
                    p.n[d] = rn[d];
                    p.send_lo[d] = d < dim ? rlo[d] - o[d] * T.bs[d] : 0;
                }
                p.op = WH_PATCH_COPY | (src_sel(T, S, ns, false) << 8);
                p.stage = 1;
                st1.push_back(p);
                continue;
            }
            // 2) finer blocks touching the box
            const auto fr = T.fin_range(r, k);
            const auto *ch0 = fr.first, *ch1 = fr.second;
            bool finer = ch1 != ch0;
            for (auto *cp = ch0; cp != ch1; ++cp) finer = finer && T.valid(cp->first);
            if (finer) {
                if (!S.restrict_) continue;
                for (auto *cp = ch0; cp != ch1; ++cp) {
                    const auto &c = *cp;
                    const int f = c.first;
                    wh_patch p{};
                    p.recv = r;
                    p.send = f;
                    for (int d = 0; d < 3; ++d) {
                        if (d >= dim) {
                            p.recv_lo[d] = 0;
                            p.n[d] = 1;
                            p.send_lo[d] = 0;
                            continue;
                        }
                        int lo = rlo[d], n = rn[d];
                        if (o[d] == 0) {
                            const int h = T.bs[d] / 2;
                            lo = g + c.second[d] * h;
                            n = h;
                        }
                        p.recv_lo[d] = lo;
                        p.n[d] = n;
                        p.send_lo[d] = 2 * (lo - g) - (2 * o[d] + c.second[d]) * T.bs[d] + g;
                        WH_CHECK(p.send_lo[d] >= 0 && p.send_lo[d] + 2 * (n - 1) < nxyz[d], WH_E_ARG,
                                 "sync: restriction box outside the finer block (too few points per block)");
                    }
                    p.op = WH_PATCH_RESTRICT | ((filt ? 2 : src_sel(T, S, f, true)) << 8);
                    p.stage = 2;
                    st23.push_back(p);
                    restrict_senders.push_back(f);
                }
                continue;
            }
            // 3) the coarser leaf covering the box
            if (!S.predict || R.level == 0) continue;
            int cu[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) {
                const int v = q[d] + o[d];
                cu[d] = v >= 0 ? v / 2 : -((-v + 1) / 2);
            }
            const int c = T.cover(r, o);
            if (!T.valid(c) || !T.is_leaf(c, true)) continue;
            wh_patch p{};
            p.recv = r;
            p.send = c;
            for (int d = 0; d < 3; ++d) {
                p.recv_lo[d] = rlo[d];
                p.n[d] = rn[d];
                if (d >= dim) {
                    p.send_lo[d] = 0;
                    continue;
                }
                // fine index in the coarse block's doubled ghost-inclusive index space
                p.send_lo[d] = q[d] * T.bs[d] + (rlo[d] - g) - 2 * cu[d] * T.bs[d] + 2 * g;
            }
            p.op = WH_PATCH_PREDICT | (src_sel(T, S, c, true) << 8);
            p.stage = 3;
            p.aux = order;
            st23.push_back(p);
        }
    }
    TreeSyncCache c;
    c.key = std::move(key);
    c.st1 = std::move(st1);
    c.st23 = std::move(st23);
    if (filt && !restrict_senders.empty()) {
        // restrict_copy_at_CE: HD-filtered copy of every restriction sender,
        // original SC points put back near level jumps of leaf senders
        std::sort(restrict_senders.begin(), restrict_senders.end());
        restrict_senders.erase(std::unique(restrict_senders.begin(), restrict_senders.end()),
                               restrict_senders.end());
        for (int arr : {A_U, A_T}) {
            std::vector<int32_t> &ids = c.fids[arr == A_U ? 0 : 1];
            std::vector<wh_box> &boxes = c.fboxes[arr == A_U ? 0 : 1];
            for (int s : restrict_senders) {
                if ((src_sel(T, S, s, true) == 1 ? A_T : A_U) != arr) continue;
                ids.push_back(s);
                if (!T.is_leaf(s, false)) continue;
                for (int k = 0; k < nd; ++k) {
                    const int *o = dirs[k];
                    if (T.valid(T.nb_same(s, o))) continue;
                    if (T.nb_coarse(s, o) < 0 && !T.any_finer(s, o, false)) continue;
                    wh_box bx{};
                    bx.hvy = s;
                    bx.op = WH_BOX_COPY_SC;
                    modify_box(T, o, T.w.Nscl, T.w.Nscr, bx);
                    boxes.push_back(bx);
                }
            }
        }
    }
    gen.reset();
    return tree_sync_run(T, cache_put(T, std::move(c)));
}

// ---- adapt_tree pieces ----------------------------------------------------------
// init_full_tree (adapt_tree.f90:995-1084): mothers down to jmin, REF_TMP_EMPTY
int init_full_tree(Tree &T, int jmin)
{
    WH_PROF("init_full_tree");
    bool changed = true;
    while (changed) {
        changed = false;
        const int n0 = (int)T.b.size();
        for (int i = 0; i < n0; ++i) {
            const TNode x = T.b[i];
            if (x.level == jmin || x.level == 0 || T.mother_raw(i) >= 0) continue;
            const int64_t mask = (((int64_t)1 << T.dim) - 1) << T.F->digit_shift(x.level);
            if (T.add(x.tc & ~mask, x.level - 1, REF_TMP_EMPTY) < 0) return WH_E_NOMEM;
            changed = true;
        }
    }
    return WH_OK;
}

// sync_D2M (executeCoarsening_tree.f90:125-227): daughters on `level` hand
// their SC (interior points g, g+2, ...) to the mother's quadrant/octant
int d2m(Tree &T, int level)
{
    WH_PROF("d2m");
    std::vector<wh_patch> p;
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (T.b[i].level != level) continue;
        const int m = T.mother(i);
        if (m < 0) continue;
        int q[3];
        T.ijk(i, q);
        wh_patch x{};
        x.recv = T.b[m].hvy;
        x.send = T.b[i].hvy;
        for (int d = 0; d < 3; ++d) {
            if (d >= T.dim) {
                x.recv_lo[d] = x.send_lo[d] = 0;
                x.n[d] = 1;
                continue;
            }
            const int h = T.bs[d] / 2;
            x.recv_lo[d] = T.g + (q[d] & 1) * h;
            x.send_lo[d] = T.g;
            x.n[d] = h;
        }
        x.op = WH_PATCH_RESTRICT;
        x.stage = 1;
        p.push_back(x);
    }
    return dev_patches(T, A_U, -1, -1, std::move(p));
}

std::vector<int32_t> hvy_of(const Tree &T, const std::vector<int> &idx)
{
    std::vector<int32_t> h;
    h.reserve(idx.size());
    for (int i : idx) h.push_back(T.b[i].hvy);
    return h;
}

// wavelet_decompose_full_tree (adapt_tree.f90:264-545)
int decompose_full_tree(Tree &T, const wh_adapt_params &p)
{
    WH_PROF("decompose_full_tree");
    const wh_wavelet &w = T.w;
    const int n_leaf = (int)T.b.size();
    std::vector<int> backup(n_leaf);
    for (int i = 0; i < n_leaf; ++i) backup[i] = T.b[i].status;
    const int hmax = std::max(std::abs(w.hd_lo), w.hd_hi);
    bool leaf_first = true;
    for (int d = 0; d < T.dim; ++d)
        if (T.bs[d] < 3 * hmax) leaf_first = false;
    if (p.decompose_branch == 1) leaf_first = true;
    if (p.decompose_branch == 2) leaf_first = false;
    int level = T.max_level();
    std::vector<int> all(n_leaf);
    for (int i = 0; i < n_leaf; ++i) {
        all[i] = i;
        if (T.b[i].level == level || leaf_first) T.b[i].status = -1;
    }
    if (int rc = dev_blocks(T, WOP_COPY, A_U, A_T, hvy_of(T, all))) return rc;
    if (int rc = init_full_tree(T, p.jmin)) return rc;
    const int g_this = std::max(std::abs(w.hd_lo), w.gd_hi);
    int it = 0;
    while (level >= p.jmin) {
        SyncSpec S;
        S.predict = !(p.use_ce && (!leaf_first || it > 0));  // MoF_ref vs full_ref
        S.recv_ref_on = true;
        S.recv_ref = -1;
        S.src = SRC_NOT_FLAG;
        S.src_ref = -1;
        S.gm = S.gp = g_this;
        if (int rc = tree_sync(T, S)) return rc;
        std::vector<int> dec, dea;
        for (int i = 0; i < (int)T.b.size(); ++i) {
            if (T.b[i].status != -1) continue;
            if (T.b[i].level == p.jmax && p.force_maxlevel_dealiasing) dea.push_back(i);
            else dec.push_back(i);
        }
        std::vector<int> both(dec);
        both.insert(both.end(), dea.begin(), dea.end());
        if (int rc = dev_blocks(T, WOP_COPY, A_U, A_T, hvy_of(T, both))) return rc;
        if (int rc = dev_blocks(T, WOP_FWT, A_T, A_U, hvy_of(T, dec))) return rc;
        if (int rc = dev_blocks(T, WOP_FILTER, A_T, A_U, hvy_of(T, dea))) return rc;
        if (p.use_ce)
            if (int rc = ce_modify(T, CE_REF, 0, -1, true, true)) return rc;
        for (TNode &x : T.b) {
            if (x.status == -1) x.status = 0;
            if ((x.status == REF_TMP_EMPTY || !leaf_first) && x.level == level - 1) x.status = -1;
        }
        if (int rc = d2m(T, level)) return rc;
        --level;
        ++it;
    }
    for (int i = 0; i < (int)T.b.size(); ++i) T.b[i].status = i < n_leaf ? backup[i] : 0;
    return WH_OK;
}

// ensureGradedness_tree (ensureGradedness_tree.f90:13-219) with
// ensure_completeness_block(check_daughters) (ensure_completeness_block.f90)
void ensure_gradedness(Tree &T, int stay)
{
    WH_PROF("ensure_gradedness");
    int dirs[26][3];
    const int nd = dirs_of(T.dim, dirs);
    const int nsis = 1 << T.dim;
    bool changed = true;
    int counter = 0;
    while (changed && counter < 10 * T.jmax + 10) {
        changed = false;
        ++counter;
        for (int i = 0; i < (int)T.b.size(); ++i) {
            int me = T.b[i].status;
            if (me == -1) {
                const int m = T.mother(i);
                int sis[8];
                bool all = true;
                for (int k = 0; k < nsis; ++k) {
                    sis[k] = m >= 0 ? T.daughter(m, k) : -1;
                    if (sis[k] < 0) all = false;
                }
                if (all) {
                    bool any_empty = false, any_stay = false, all_neg = true;
                    for (int k = 0; k < nsis; ++k) {
                        const int s = T.b[sis[k]].status;
                        if (s == REF_TMP_EMPTY) any_empty = true;
                        if (s == 0 || s == stay) any_stay = true;
                        if (s >= 0) all_neg = false;
                    }
                    if (any_empty || any_stay) {
                        for (int k = 0; k < nsis; ++k)
                            if (T.b[sis[k]].status < 0) T.b[sis[k]].status = stay;
                    } else if (all_neg) {
                        for (int k = 0; k < nsis; ++k) T.b[sis[k]].status = -1;
                    }
                    if (!T.is_leaf(i, false))
                        for (int k = 0; k < nsis; ++k) {
                            const int d = T.daughter(i, k);
                            if (d >= 0 && (T.b[d].status == 0 || T.b[d].status == stay))
                                for (int l = 0; l < nsis; ++l)
                                    if (T.b[sis[l]].status < 0) T.b[sis[l]].status = stay;
                        }
                } else {
                    for (int k = 0; k < nsis; ++k)
                        if (sis[k] >= 0 && T.b[sis[k]].status < -1) T.b[sis[k]].status = stay;
                }
                if (T.b[i].status != me) {
                    me = T.b[i].status;
                    changed = true;
                }
            }
            if (me == -1) {
                for (int q = 0; q < nd; ++q) {
                    const int *o = dirs[q];
                    const int s = T.nb_same(i, o);
                    if (s >= 0 && T.b[s].status == 1) {
                        T.b[i].status = stay;
                        changed = true;
                    }
                    const auto fr = T.fin_range(i, q);
                    for (auto *cp = fr.first; cp != fr.second; ++cp) {
                        const int f = cp->first;
                        if (f < 0) continue;
                        if (T.b[f].status == 1) {
                            T.b[i].status = 1;
                            changed = true;
                        } else if (T.b[f].status == 0 || T.b[f].status == stay) {
                            if (T.b[i].status != 1) T.b[i].status = stay;
                            changed = true;
                        }
                    }
                }
            } else if (me == 0 || me == stay) {
                for (int q = 0; q < nd; ++q) {
                    const auto fr = T.fin_range(i, q);
                    for (auto *cp = fr.first; cp != fr.second; ++cp)
                        if (cp->first >= 0 && T.b[cp->first].status == 1) {
                            T.b[i].status = 1;
                            changed = true;
                        }
                }
            }
        }
    }
}

// wavelet_reconstruct_full_tree_CEoptimized (adapt_tree.f90:686-987)
int reconstruct_tree(Tree &T, const wh_adapt_params &p)
{
    WH_PROF("reconstruct_tree");
    const wh_wavelet &w = T.w;
    int dirs[26][3];
    const int nd = dirs_of(T.dim, dirs);
    std::vector<int> backup(T.b.size());
    for (size_t i = 0; i < T.b.size(); ++i) {
        backup[i] = T.b[i].status;
        T.b[i].status = 0;
    }
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (!T.is_leaf(i, false)) continue;
        for (int q = 0; q < nd; ++q) {
            const int *o = dirs[q];
            if (T.nb_coarse(i, o) < 0) continue;
            if (T.valid(T.nb_same(i, o)) || T.any_finer(i, o, true)) continue;
            T.b[i].status = -1;
            break;
        }
    }
    bool reconstruct_neighbors = false, leaf_only = true;
    for (int d = 0; d < T.dim; ++d) {
        if (T.bs[d] < w.Nreconl || T.bs[d] < w.Nreconr) reconstruct_neighbors = true;
        if (T.bs[d] < w.Ndep2l || T.bs[d] < w.Ndep2r) leaf_only = false;
    }
    WH_CHECK(!reconstruct_neighbors, WH_E_ARG,
             "adapt_tree: blocks smaller than the reconstruction reach are not supported");
    SyncSpec S;
    S.restrict_ = false;
    S.recv_ref_on = true;
    S.recv_ref = -1;
    S.src = SRC_DECOMPOSED;
    S.gm = S.gp = T.g;
    auto finish = [&](int level_only) -> int {
        std::vector<int> rec, all;
        for (int i = 0; i < (int)T.b.size(); ++i) {
            if (level_only >= 0 && T.b[i].level != level_only) continue;
            all.push_back(i);
            if (T.b[i].status == -1) rec.push_back(i);
        }
        if (int rc = dev_blocks(T, WOP_IWT, A_U, A_T, hvy_of(T, rec))) return rc;
        return dev_blocks(T, WOP_COPY, A_T, A_U, hvy_of(T, all));
    };
    if (leaf_only) {
        if (int rc = tree_sync(T, S)) return rc;
        if (p.use_ce)
            if (int rc = ce_modify(T, CE_TREE, 0, 0, true, false)) return rc;
        if (int rc = finish(-1)) return rc;
    } else {
        const int lmax = T.max_level();
        S.recv_level_on = true;
        for (int L = p.jmin; L <= lmax; ++L) {
            S.recv_level = L;
            if (int rc = tree_sync(T, S)) return rc;
            if (p.use_ce)
                if (int rc = ce_modify(T, CE_LEVEL, L, 0, true, false)) return rc;
            if (int rc = finish(L)) return rc;
        }
    }
    for (size_t i = 0; i < T.b.size(); ++i) T.b[i].status = backup[i];
    return WH_OK;
}

// the max over each group l >= 2 of the per-component values, 0 for group 0
void combine_groups(double *v, int neq, const int32_t *groups)
{
    int maxg = 0;
    for (int c = 0; c < neq; ++c) maxg = std::max(maxg, (int)groups[c]);
    for (int l = 2; l <= maxg; ++l) {
        double m = -1.0;
        for (int c = 0; c < neq; ++c)
            if (groups[c] == l) m = std::max(m, v[c]);
        for (int c = 0; c < neq; ++c)
            if (groups[c] == l) v[c] = m;
    }
    for (int c = 0; c < neq; ++c)
        if (groups[c] == 0) v[c] = 0.0;
}

wh_box interior_box(const Tree &T, int i)
{
    wh_box b{};
    b.hvy = T.b[i].hvy;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = d < T.dim ? T.g : 0;
        b.hi[d] = d < T.dim ? T.g + T.bs[d] - 1 : 0;
    }
    return b;
}

// componentWiseNorm_tree(hvy_tmp, eps_norm) over the leaves
// (OPERATORS/componentWiseNorm_tree.f90): the leaves' interior boxes (their
// per-box maxima or sums are reduced on the device, accumulate_norm makes
// the norm of them, then norm <= 1e-9 -> 1, coarseningIndicator_tree.f90:160-163)
std::vector<wh_box> leaf_boxes(const Tree &T)
{
    std::vector<wh_box> boxes;
    for (int i = 0; i < (int)T.b.size(); ++i)
        if (T.is_leaf(i, false)) boxes.push_back(interior_box(T, i));
    return boxes;
}

// the levels of the leaf boxes, in leaf_boxes' order
std::vector<int> leaf_levels(const Tree &T)
{
    std::vector<int> lv;
    for (int i = 0; i < (int)T.b.size(); ++i)
        if (T.is_leaf(i, false)) lv.push_back(T.b[i].level);
    return lv;
}

// coarseningIndicatorMask_tree (coarseningIndicator_tree.f90:338-422): the
// mask of every block of the full tree (mask_blocks lists them and launches
// the mask); a block whose chi is partly inside (0,1) or not constant stays
// (mask_statuses, from [between, max, min] per listed block)
int mask_blocks(Tree &T, const wh_adapt_params &p, std::vector<int> &idx, std::vector<int32_t> &ids)
{
    WH_PROF("mask_criterion");
    std::vector<double> x0dx;
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (p.force_maxlevel_dealiasing && T.b[i].level == p.jmax) continue;
        idx.push_back(i);
        ids.push_back(T.b[i].hvy);
        double xd[6];
        T.F->x0dx(FBlock{T.b[i].tc, T.b[i].level, T.b[i].hvy, 0}, xd);
        x0dx.insert(x0dx.end(), xd, xd + 6);
    }
    return run_mask(T.P, p.mask, ids, x0dx);
}

void mask_statuses(Tree &T, const std::vector<int> &idx, const double *st)
{
    for (size_t k = 0; k < idx.size(); ++k) {
        const bool stay = st[3 * k] != 0.0 || (st[3 * k + 1] - st[3 * k + 2]) > 1.0e-12;
        const int ms = stay ? 0 : -1;
        int &s = T.b[idx[k]].status;
        s = std::max(s, ms);
    }
}

// addSecurityZone_CE_tree (securityZone_tree.f90): a block that would coarsen
// stays if a same-level neighbour that stays has significant details within
// Nwc of their common boundary.  zone_boxes lists the boxes whose maxima
// decide it (a block that stays, toward a neighbour that would coarsen);
// zone_apply applies the rule given their per-box maxima m.
void zone_boxes(const Tree &T, const wh_adapt_params &p, std::vector<wh_box> &boxes,
                std::vector<std::pair<int, int>> &who)
{
    WH_PROF("security_zone");
    int dirs[26][3];
    const int nd = dirs_of(T.dim, dirs);
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (T.b[i].level == p.jmax && p.force_maxlevel_dealiasing) continue;
        if (T.b[i].status != 0) continue;
        for (int q = 0; q < nd; ++q) {
            const int s = T.nb_same(i, dirs[q]);
            if (s < 0 || T.b[s].status != -1) continue;
            wh_box bx{};
            bx.hvy = T.b[i].hvy;
            modify_box(T, dirs[q], T.w.Nwcl, T.w.Nwcr, bx);
            boxes.push_back(bx);
            who.push_back({i, q});
        }
    }
}

void zone_apply(Tree &T, const wh_adapt_params &p, const double *norm, const std::vector<std::pair<int, int>> &who,
                double *m)
{
    int dirs[26][3];
    const int nd = dirs_of(T.dim, dirs);
    const int neq = T.P->d.neq;
    std::vector<char> flag(T.b.size() * 26, 0);
    for (size_t k = 0; k < who.size(); ++k) {
        double *v = &m[k * neq];
        combine_groups(v, neq, p.groups);
        bool all = true;
        for (int c = 0; c < neq; ++c)
            if (!(v[c] <= p.eps * norm[c])) all = false;
        if (!all) flag[(size_t)who[k].first * 26 + who[k].second] = 1;
    }
    std::vector<char> keep(T.b.size(), 0);
    for (int i = 0; i < (int)T.b.size(); ++i) {
        if (T.b[i].status != -1) continue;
        for (int q = 0; q < nd; ++q) {
            const int s = T.nb_same(i, dirs[q]);
            if (s < 0) continue;
            int inv = -1;
            for (int r = 0; r < nd; ++r)
                if (dirs[r][0] == -dirs[q][0] && dirs[r][1] == -dirs[q][1] && dirs[r][2] == -dirs[q][2]) inv = r;
            if (flag[(size_t)s * 26 + inv]) keep[i] = 1;
        }
    }
    for (size_t i = 0; i < T.b.size(); ++i)
        if (keep[i]) T.b[i].status = 0;
}

// leaves of the tree back into the forest's block list
void store_leaves(Tree &T)
{
    T.F->blocks.clear();
    for (const TNode &x : T.b) T.F->blocks.push_back(FBlock{x.tc, x.level, x.hvy, 0});
    T.F->rebuild_map();
}

int check_pool(wh_pool *P, wh_forest *F)
{
    WH_CHECK(P && F, WH_E_ARG, "adapt: null argument");
    WH_CHECK(P->ws.set, WH_E_STATE, "adapt: set the pool's wavelet first (wh_pool_set_wavelet)");
    WH_CHECK(F->dim == P->d.dim && F->g == P->d.g && F->slot0 + F->nb_max <= P->d.nb_max, WH_E_ARG,
             "adapt: forest and pool disagree on dim / ghost nodes / capacity");
    for (int d = 0; d < F->dim; ++d)
        WH_CHECK(F->bs[d] == P->d.bs[d] && F->bs[d] % 2 == 0 && 2 * F->g <= F->bs[d], WH_E_ARG,
                 "adapt: block size must match the pool, be even and at least 2g");
    WH_CHECK(P->d.n_work_slots >= 2 && P->tmp && P->d.n_tmp_eqn >= P->d.neq, WH_E_ARG,
             "adapt: needs two work slots and hvy_tmp with neq components");
    return WH_OK;
}

}  // namespace
}  // namespace wh

namespace wh {
namespace {

// Host threads for the per-forest work of the multi-forest calls (the
// forests are independent until their device operations are merged):
// WH_HOST_THREADS, default 4.  fn(i) for i in [0, n); device calls stay on
// the calling thread.
int host_threads()
{
    static const int n = [] {
        const char *e = std::getenv("WH_HOST_THREADS");
        const int v = e ? std::atoi(e) : 4;
        return std::max(1, std::min(v, 64));
    }();
    return n;
}

template <class F>
void parallel_for(int n, F &&fn)
{
    const int nt = std::min(host_threads(), n);
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&] {
        for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) fn(i);
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int k = 1; k < nt; ++k) th.emplace_back(work);
    work();
    for (std::thread &x : th) x.join();
}

// adapt_tree's checks of one forest (adapt_tree.f90:11-255 preconditions)
int check_adapt(wh_pool *P, wh_forest *F, const wh_adapt_params *p, const int32_t *leaf_flags)
{
    if (int rc = check_pool(P, F)) return rc;
    WH_CHECK(p && p->jmin >= 0 && p->jmax <= F->jmax && p->jmin <= p->jmax &&
                 (p->indicator == WH_IND_EVERYWHERE || p->indicator == WH_IND_THRESHOLD ||
                  (p->indicator == WH_IND_EXTERNAL && leaf_flags)) &&
                 p->decompose_branch >= 0 && p->decompose_branch <= 2 && p->eps_norm >= WH_NORM_LINFTY &&
                 p->eps_norm <= WH_NORM_H1,
             WH_E_ARG, "wh_adapt_tree: bad parameters");
    for (const FBlock &b : F->blocks)
        WH_CHECK(b.level >= p->jmin, WH_E_ARG, "wh_adapt_tree: a block is below jmin");
    if (leaf_flags)
        for (size_t i = 0; i < F->blocks.size(); ++i)
            WH_CHECK(leaf_flags[i] == -1 || leaf_flags[i] == 0 || leaf_flags[i] == 1, WH_E_ARG,
                     "wh_adapt_tree_flags: a flag is not -1, 0 or 1");
    return WH_OK;
}

// forests whose reductions can share a launch: equal groups and norm
bool same_groups(const wh_adapt_params &a, const wh_adapt_params &b, int neq)
{
    if (a.eps_norm != b.eps_norm) return false;
    for (int c = 0; c < neq; ++c)
        if (a.groups[c] != b.groups[c]) return false;
    return true;
}

// One forest of a multi-forest adapt_tree.
struct TreeRun {
    wh_forest *F = nullptr;
    const wh_adapt_params *p = nullptr;
    std::unique_ptr<Tree> T;
    std::vector<int> free0;
    int rc = WH_OK;
    OpLog log;
    double norm[8];
    std::vector<int> idx, midx;
    std::vector<int32_t> mids;
    std::vector<wh_box> nboxes, zboxes;
    std::vector<int> nlev;
    std::vector<std::pair<int, int>> zwho;
    // results per detail or zone box: one, or one per point type where the
    // norm renormalises (the host scales the typed maxima)
    int ntyp(int dim) const { return p->eps_norm != WH_NORM_LINFTY ? (1 << dim) - 1 : 1; }
    std::string err;  // the message of rc (set on the thread that failed)
    bool alive() const { return rc == WH_OK && T; }
};

// Region maxima of the listed boxes of several forests: consecutive forests
// with equal groups in one launch; forest t's maxima at d_out + off(t), the
// forests' ranges one after the other in forest order.
int region_max_runs(wh_pool *P, std::vector<TreeRun *> &R, std::vector<wh_box> TreeRun::*boxes, int array,
                    bool zone, double *d_out)
{
    const int neq = P->d.neq;
    size_t off = 0;
    for (size_t a = 0; a < R.size();) {
        size_t b = a + 1;
        while (b < R.size() && same_groups(*R[a]->p, *R[b]->p, neq)) ++b;
        std::vector<wh_box> all;
        for (size_t t = a; t < b; ++t) all.insert(all.end(), (R[t]->*boxes).begin(), (R[t]->*boxes).end());
        if (int rc = launch_region_norm(P, array, all, R[a]->p->eps_norm, zone, R[a]->p->groups, d_out + off))
            return rc;
        off += all.size() * neq * (zone ? R[a]->ntyp(P->d.dim) : 1);
        a = b;
    }
    return WH_OK;
}

// adapt_tree (adapt_tree.f90:11-255) for n forests of one pool: per forest
// the reference's sequence of steps, with the device operations of all
// forests launched together (op logs, replay) and the indicator's reductions
// of all forests read back in one round trip (one more for the security
// zone).  flags[i]: forest i's external leaf statuses (indicator "nothing
// (external)") or null.  status[i]: forest i's result; an error of one
// forest (its checks, its slots) leaves its block list unchanged and the
// others go on; the return value is an error only when an operation shared
// by all forests failed.
int adapt_trees_impl(wh_pool *P, wh_forest *const *Fs, int n, const wh_adapt_params *ps,
                     const int32_t *const *flags, int *status)
{
    const int stay = REF_UNSIGNIFICANT_STAY;
    const int neq = P->d.neq, dim = P->d.dim;
    std::vector<TreeRun> R(n);
    auto fail = [&](TreeRun &t, int rc) {
        if (t.rc != WH_OK) return;
        if (t.T) t.F->free_hvy = t.free0;  // restore the forest's slots; device data of the leaves is not
        t.rc = rc;
        t.err = wh_last_error();
    };
    // fn(t) for every live forest, on the host threads
    auto each_live = [&](auto &&fn) {
        std::vector<TreeRun *> v;
        for (TreeRun &t : R)
            if (t.alive()) v.push_back(&t);
        parallel_for((int)v.size(), [&](int i) { fn(*v[i]); });
    };
    for (int i = 0; i < n; ++i) {
        TreeRun &t = R[i];
        t.F = Fs[i];
        t.p = &ps[i];
        const int32_t *lf = flags ? flags[i] : nullptr;
        if (int rc = check_adapt(P, t.F, t.p, lf)) {
            t.rc = rc;
            continue;
        }
        t.T.reset(new Tree(t.F, P));
        if (lf)  // the leaves are T.b[0, n) in forest list order
            for (size_t k = 0; k < t.F->blocks.size(); ++k) t.T->b[k].status = lf[k];
        t.free0 = t.F->free_hvy;
        // one forest launches as it goes (its host work overlaps the device);
        // several log their operations and launch them together
        if (n > 1) t.T->log = &t.log;
    }
    int shared = WH_OK;
    auto live = [&]() {
        std::vector<TreeRun *> v;
        for (TreeRun &t : R)
            if (t.alive()) v.push_back(&t);
        return v;
    };
    auto run_logs = [&]() {
        std::vector<OpLog *> logs;
        for (TreeRun *t : live()) logs.push_back(&t->log);
        if (int rc = replay(P, logs)) {
            for (TreeRun *t : live()) fail(*t, rc);
            shared = rc;
        }
    };
    // wavelet_decompose_full_tree; it hands the leaves' statuses back (its
    // backup), the new mothers get 0
    each_live([&](TreeRun &t) {
        if (int rc = decompose_full_tree(*t.T, *t.p)) fail(t, rc);
    });
    run_logs();
    // coarseningIndicator_tree (coarseningIndicator_tree.f90:4-331); with the
    // external indicator it returns at once (:55).  The reductions that need
    // no host decision (leaf norms, details, mask statistics) of all forests
    // go out together: results [norms of all][details of all][mask of all]
    std::vector<TreeRun *> thr;
    for (TreeRun *t : live()) {
        Tree &T = *t->T;
        const wh_adapt_params &p = *t->p;
        if (p.indicator != WH_IND_EXTERNAL)
            for (TNode &x : T.b) x.status = 0;
        if (p.indicator == WH_IND_EVERYWHERE) {
            for (int i = 0; i < (int)T.b.size(); ++i)
                if (T.is_leaf(i, false)) T.b[i].status = -1;
        } else if (p.indicator == WH_IND_THRESHOLD) {
            for (int c = 0; c < 8; ++c) t->norm[c] = p.norm[c];
            if (p.eps_normalized) {
                t->nboxes = leaf_boxes(T);
                t->nlev = leaf_levels(T);
            }
            for (int i = 0; i < (int)T.b.size(); ++i) {
                if (p.force_maxlevel_dealiasing && T.b[i].level == p.jmax) T.b[i].status = -1;
                else t->idx.push_back(i);
            }
            thr.push_back(t);
        }
    }
    if (!thr.empty() && !shared) {
        WH_PROF("indicator_reductions");
        size_t n_norm = 0, n_det = 0, n_mask = 0;
        for (TreeRun *t : thr) {
            n_norm += t->nboxes.size() * neq;
            n_det += t->idx.size() * neq * t->ntyp(dim);
            if (t->p->threshold_mask) n_mask += t->idx.size() * 3;
        }
        double *d_res = nullptr;
        std::vector<double> res;
        int rc = results_reserve(P, n_norm + n_det + n_mask, &d_res);
        if (!rc) rc = region_max_runs(P, thr, &TreeRun::nboxes, A_T, false, d_res);
        size_t off = n_norm;
        for (size_t a = 0; a < thr.size() && !rc;) {  // details: consecutive forests with equal groups
            size_t b = a + 1;
            while (b < thr.size() && same_groups(*thr[a]->p, *thr[b]->p, neq)) ++b;
            std::vector<int32_t> ids;
            for (size_t t = a; t < b; ++t) {
                const std::vector<int32_t> h = hvy_of(*thr[t]->T, thr[t]->idx);
                ids.insert(ids.end(), h.begin(), h.end());
            }
            const int nt = thr[a]->ntyp(dim);
            rc = launch_detail_list(P, A_U, ids, thr[a]->p->groups, d_res + off, nt > 1);
            off += ids.size() * neq * nt;
            a = b;
        }
        std::vector<int32_t> mids;
        for (TreeRun *t : thr) {  // the masks: one launch per forest (its own geometry)
            if (rc || !t->p->threshold_mask) continue;
            rc = mask_blocks(*t->T, *t->p, t->midx, t->mids);
            if (!rc && t->mids.size() != t->idx.size()) {
                set_error("adapt: mask and detail block lists differ");
                rc = WH_E_STATE;
            }
            mids.insert(mids.end(), t->mids.begin(), t->mids.end());
        }
        if (!rc && !mids.empty()) rc = launch_mask_stats(P, mids, d_res + n_norm + n_det);
        if (!rc) rc = results_fetch(P, n_norm + n_det + n_mask, res);
        if (rc) {
            for (TreeRun *t : thr) fail(*t, rc);
            shared = rc;
        } else {
            std::vector<size_t> o_norm(thr.size()), o_det(thr.size()), o_mask(thr.size());
            size_t a = 0, b = n_norm, c = n_norm + n_det;
            for (size_t i = 0; i < thr.size(); ++i) {
                o_norm[i] = a;
                o_det[i] = b;
                o_mask[i] = c;
                a += thr[i]->nboxes.size() * neq;
                b += thr[i]->idx.size() * neq * thr[i]->ntyp(dim);
                if (thr[i]->p->threshold_mask) c += thr[i]->idx.size() * 3;
            }
            parallel_for((int)thr.size(), [&](int i) {
                TreeRun *t = thr[i];
                Tree &T = *t->T;
                const wh_adapt_params &p = *t->p;
                if (p.eps_normalized) {
                    accumulate_norm(res.data() + o_norm[i], (int)t->nboxes.size(), neq, p.groups, t->nlev.data(), dim,
                                    T.bs, p.domain, p.eps_norm, t->norm);
                    for (int c = 0; c < neq; ++c)
                        if (t->norm[c] <= 1.0e-9) t->norm[c] = 1.0;
                }
                double *det = res.data() + o_det[i];
                std::vector<double> renormed;
                if (t->ntyp(dim) > 1) {  // wavelet_renorm_block of the typed maxima
                    std::vector<int> lv(t->idx.size());
                    for (size_t k = 0; k < t->idx.size(); ++k) lv[k] = T.b[t->idx[k]].level;
                    renormed.resize(t->idx.size() * neq);
                    renorm_typed(det, (int)t->idx.size(), neq, dim, p.eps_norm, lv.data(), p.groups, renormed.data());
                    det = renormed.data();
                }
                detail_groups(P, p.groups, det, (int)t->idx.size());
                for (size_t k = 0; k < t->idx.size(); ++k) {
                    bool all = true;
                    for (int c = 0; c < neq; ++c)
                        if (!(det[k * neq + c] <= p.eps * t->norm[c])) all = false;
                    T.b[t->idx[k]].status = all ? -1 : 0;
                }
                if (p.threshold_mask) mask_statuses(T, t->midx, res.data() + o_mask[i]);
            });
        }
    }
    // the security zone of all forests in one more round trip
    std::vector<TreeRun *> zone;
    for (TreeRun *t : thr)
        if (t->alive() && t->p->security_zone) zone.push_back(t);
    parallel_for((int)zone.size(), [&](int i) { zone_boxes(*zone[i]->T, *zone[i]->p, zone[i]->zboxes, zone[i]->zwho); });
    if (!zone.empty() && !shared) {
        size_t nz = 0;
        for (TreeRun *t : zone) nz += t->zboxes.size() * neq * t->ntyp(dim);
        double *d_res = nullptr;
        std::vector<double> res;
        int rc = results_reserve(P, nz, &d_res);
        if (!rc) rc = region_max_runs(P, zone, &TreeRun::zboxes, A_U, true, d_res);
        if (!rc) rc = results_fetch(P, nz, res);
        if (rc) {
            for (TreeRun *t : zone) fail(*t, rc);
            shared = rc;
        } else {
            std::vector<size_t> off(zone.size());
            size_t o = 0;
            for (size_t i = 0; i < zone.size(); ++i) {
                off[i] = o;
                o += zone[i]->zboxes.size() * neq * zone[i]->ntyp(dim);
            }
            parallel_for((int)zone.size(), [&](int i) {
                TreeRun *t = zone[i];
                double *m = res.data() + off[i];
                std::vector<double> renormed;
                if (t->ntyp(dim) > 1) {  // wavelet_renorm_block of the typed maxima, the zone block's level
                    const int nb = (int)t->zboxes.size();
                    std::vector<int> lv(nb);
                    for (int k = 0; k < nb; ++k) lv[k] = t->T->b[t->zwho[k].first].level;
                    renormed.resize((size_t)nb * neq);
                    renorm_typed(m, nb, neq, dim, t->p->eps_norm, lv.data(), t->p->groups, renormed.data());
                    m = renormed.data();
                }
                zone_apply(*t->T, *t->p, t->norm, t->zwho, m);
            });
        }
    }
    each_live([&](TreeRun &tr) {
        TreeRun *t = &tr;
        Tree &T = *t->T;
        const wh_adapt_params &p = *t->p;
        // respectJmaxJmin_tree (respectJmaxJmin_tree.f90:1-39)
        for (TNode &x : T.b) {
            if (x.status == 1 && x.level >= p.jmax) x.status = stay;
            if (x.status == -1 && x.level <= p.jmin) x.status = stay;
        }
        ensure_gradedness(T, stay);
        std::vector<char> del(T.b.size(), 0);
        for (size_t i = 0; i < T.b.size(); ++i) del[i] = T.b[i].status == -1;
        T.remove_marked(del);
        if (p.use_ce) {
            int rc = ce_modify(T, CE_TREE, 0, 0, true, false);
            if (!rc) rc = reconstruct_tree(T, p);
            if (rc) {
                fail(*t, rc);
                t->log.ops.clear();
            }
        } else {
            std::vector<int> all(T.b.size());
            for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
            if (int rc = dev_blocks(T, WOP_COPY, A_T, A_U, hvy_of(T, all))) fail(*t, rc);
        }
    });
    run_logs();
    each_live([&](TreeRun &tr) {
        TreeRun *t = &tr;
        Tree &T = *t->T;
        // prune_fulltree2leafs
        std::vector<char> del(T.b.size(), 0);
        for (size_t i = 0; i < T.b.size(); ++i) del[i] = !T.is_leaf((int)i, false);
        T.remove_marked(del);
        for (TNode &x : T.b) x.status = 0;
        store_leaves(T);
        SyncSpec S;
        S.gm = S.gp = T.g;
        if (int rc = tree_sync(T, S)) fail(*t, rc);
    });
    run_logs();
    bool any_failed = false;
    for (int i = 0; i < n; ++i) {
        status[i] = R[i].rc;
        if (R[i].rc && !any_failed && !R[i].err.empty()) set_error(R[i].err);  // the first forest's message
        if (R[i].rc) any_failed = true;
    }
    if (any_failed) (void)stage_sync(P);
    return shared;
}

// sync_ghosts_tree of n forests of one pool, their ghost patches in one
// launch per operation
int sync_trees_impl(wh_pool *P, wh_forest *const *Fs, int n)
{
    for (int i = 0; i < n; ++i)
        if (int rc = check_pool(P, Fs[i])) return rc;
    std::vector<OpLog> logs(n);
    std::vector<int> rcs(n, WH_OK);
    std::vector<std::string> errs(n);
    parallel_for(n > 1 ? n : 0, [&](int i) {  // several forests: patch lists on the host threads
        Tree T(Fs[i], P);
        T.log = &logs[i];
        SyncSpec S;
        S.gm = S.gp = T.g;
        if ((rcs[i] = tree_sync(T, S))) errs[i] = wh_last_error();
    });
    if (n == 1) {  // one forest launches as it goes
        Tree T(Fs[0], P);
        SyncSpec S;
        S.gm = S.gp = T.g;
        return tree_sync(T, S);
    }
    for (int i = 0; i < n; ++i)
        if (rcs[i]) {
            set_error(errs[i]);
            return rcs[i];
        }
    std::vector<OpLog *> lp;
    for (OpLog &l : logs) lp.push_back(&l);
    return replay(P, lp);
}

// the forest's blocks fit the pool (dimension, block size, ghost nodes,
// slot range)
int check_fits(wh_pool *P, const wh_forest *F)
{
    WH_CHECK(F->dim == P->d.dim && F->g == P->d.g && F->slot0 + F->nb_max <= P->d.nb_max, WH_E_ARG,
             "forest and pool disagree on dim / ghost nodes / capacity");
    for (int d = 0; d < F->dim; ++d) WH_CHECK(F->bs[d] == P->d.bs[d], WH_E_ARG, "forest and pool block sizes differ");
    return WH_OK;
}

// refine_tree with the "everywhere" indicator (refine_tree.f90: every leaf
// below jmax[i] gets 2^dim daughters, which keeps a graded grid graded) for
// n forests of one pool: the block lists on the host threads, then one
// refinement launch (refineBlock, refinementExecute.f90:4-116) for the new
// blocks of all of them.  A forest without the free slots for all its
// daughters at once is left unchanged and reports WH_E_NOMEM.
int refine_trees_impl(wh_pool *P, wh_forest *const *Fs, int n, const int *jmax, int order, int array, int *status)
{
    std::vector<std::vector<int32_t>> tri(n);
    std::vector<std::string> errs(n);
    parallel_for(n, [&](int i) {
        wh_forest *F = Fs[i];
        status[i] = check_fits(P, F);
        if (status[i]) {
            errs[i] = wh_last_error();
            return;
        }
        std::vector<int32_t> ids;
        for (const FBlock &b : F->blocks)
            if (b.level < jmax[i]) ids.push_back(b.hvy);
        if (ids.empty()) return;
        const int nd = 1 << F->dim, cap = (int)ids.size() * nd;
        if ((int)F->free_hvy.size() < cap) {
            status[i] = WH_E_NOMEM;
            errs[i] = "refine_trees: the forest has too few free slots to refine all its blocks at once";
            return;
        }
        tri[i].resize(3 * (size_t)cap);
        int np = 0;
        if ((status[i] = wh_forest_refine(F, (int)ids.size(), ids.data(), tri[i].data(), cap, &np))) {
            errs[i] = wh_last_error();
            tri[i].clear();
            return;
        }
        tri[i].resize(3 * (size_t)np);
    });
    std::vector<int32_t> all;
    for (int i = 0; i < n; ++i) all.insert(all.end(), tri[i].begin(), tri[i].end());
    for (int i = 0; i < n; ++i)
        if (status[i]) set_error(errs[i]);  // the last failing forest's message
    if (all.empty()) return WH_OK;
    return wh_refine_blocks(P, array, all.data(), (int)(all.size() / 3), order);
}

// The RK tables of a pool from n forests' leaves: the active list and
// origin/spacing (forest after forest, each in its leaf order), block group
// k for the k-th forest taken, and the RHS ghost-patch table (depth g_rhs,
// predictor order; sync_ghosts_RHS_tree) -- what wh_set_active,
// wh_set_block_groups and wh_set_patches(0) set from the forests' own lists,
// with the patch lists generated on the host threads and moved into the
// pool's table directly.  A forest whose patches cannot be generated is left
// out (status); nblocks[i]: its leaves in the tables (0 if left out).
int set_tables_impl(wh_pool *P, wh_forest *const *Fs, int n, int g_rhs, int order, int *status, int *nblocks)
{
    std::vector<std::string> errs(n);
    parallel_for(n, [&](int i) {
        int np = 0;
        status[i] = check_fits(P, Fs[i]);
        if (!status[i]) status[i] = wh_forest_patches(Fs[i], g_rhs, g_rhs, order, nullptr, 0, &np);
        if (status[i]) errs[i] = wh_last_error();
    });
    size_t nact = 0;
    for (int i = 0; i < n; ++i) {
        nblocks[i] = status[i] ? 0 : (int)Fs[i]->blocks.size();
        nact += nblocks[i];
        if (status[i]) set_error(errs[i]);
    }
    WH_CHECK(nact <= (size_t)P->d.nb_max, WH_E_ARG, "set_tables: more leaves than pool slots");
    std::vector<int32_t> act(nact), grp(nact);
    std::vector<double> x0dx(6 * nact);
    std::vector<PatchRun> runs;
    size_t k = 0;
    int ng = 0;
    for (int i = 0; i < n; ++i) {
        if (status[i]) continue;
        const wh_forest *F = Fs[i];
        for (const FBlock &b : F->blocks) {
            act[k] = b.hvy;
            grp[k] = ng;
            F->x0dx(b, &x0dx[6 * k]);
            ++k;
        }
        // the list wh_forest_patches just generated (the forest's cache)
        runs.push_back(PatchRun{F->pcache.data(), (int)F->pcache.size()});
        ++ng;
    }
    if (int rc = wh_set_active(P, act.data(), (int)nact, x0dx.data())) return rc;
    if (int rc = set_patch_table(P, 0, runs.data(), (int)runs.size())) return rc;
    if (ng > 0)
        if (int rc = wh_set_block_groups(P, grp.data(), (int)nact, ng)) return rc;
    return WH_OK;
}

// createMask_tree (create_mask.f90:6-320) on the leaves of n forests of one
// pool, each with its own parameters, one mask launch per forest
int create_masks_impl(wh_pool *P, wh_forest *const *Fs, int n, const wh_mask_params *mp, int *status)
{
    // every forest that fits and whose geometry is valid, in one launch
    // (its blocks take their forest's geometry through a per-block index)
    std::vector<wh_mask_params> geos;
    std::vector<int32_t> gidx, ids;
    std::vector<double> x0dx;
    std::vector<int> in;
    for (int i = 0; i < n; ++i) {
        const wh_forest *F = Fs[i];
        if ((status[i] = check_fits(P, F))) continue;
        if ((status[i] = check_mask_params(P, mp[i]))) continue;
        in.push_back(i);
        const int q = (int)geos.size();
        geos.push_back(mp[i]);
        for (const FBlock &b : F->blocks) {
            ids.push_back(b.hvy);
            gidx.push_back(q);
            x0dx.resize(x0dx.size() + 6);
            F->x0dx(b, &x0dx[x0dx.size() - 6]);
        }
    }
    if (geos.empty()) return WH_OK;
    if (geos.size() == 1) gidx.clear();
    const int rc = run_masks(P, geos, gidx, ids, x0dx);
    for (int i : in) status[i] = rc;
    return rc == WH_E_HIP ? rc : (int)WH_OK;  // the device failed: every forest fails
}

}  // namespace
}  // namespace wh

extern "C" {

int wh_sync_tree(wh_pool *P, wh_forest *F, int array)
{
    WH_PROF("api.sync_tree");
    WH_CHECK(P && F, WH_E_ARG, "wh_sync_tree: null argument");
    WH_CHECK(array == WH_BLOCK, WH_E_ARG, "wh_sync_tree: only WH_BLOCK is supported");
    return wh::sync_trees_impl(P, &F, 1);
}

int wh_sync_trees(wh_pool *P, wh_forest *const *forests, int n, int array)
{
    WH_PROF("api.sync_tree");
    WH_CHECK(P && n >= 0 && (n == 0 || forests), WH_E_ARG, "wh_sync_trees: null argument");
    WH_CHECK(array == WH_BLOCK, WH_E_ARG, "wh_sync_trees: only WH_BLOCK is supported");
    for (int i = 0; i < n; ++i) {
        WH_CHECK(forests[i], WH_E_ARG, "wh_sync_trees: null forest");
        for (int j = 0; j < i; ++j) WH_CHECK(forests[j] != forests[i], WH_E_ARG, "wh_sync_trees: a forest listed twice");
    }
    return wh::sync_trees_impl(P, forests, n);
}

int wh_adapt_tree(wh_pool *P, wh_forest *F, const wh_adapt_params *p)
{
    WH_PROF("api.adapt_tree");
    WH_CHECK(P && F && p && p->indicator != WH_IND_EXTERNAL, WH_E_ARG,
             "wh_adapt_tree: the external indicator needs wh_adapt_tree_flags");
    int st = WH_OK;
    const int rc = wh::adapt_trees_impl(P, &F, 1, p, nullptr, &st);
    return rc ? rc : st;
}

int wh_adapt_tree_flags(wh_pool *P, wh_forest *F, const wh_adapt_params *p, const int32_t *leaf_flags, int n)
{
    WH_PROF("api.adapt_tree");
    WH_CHECK(P && F && p && p->indicator == WH_IND_EXTERNAL && leaf_flags && n == (int)F->blocks.size(),
             WH_E_ARG, "wh_adapt_tree_flags: needs WH_IND_EXTERNAL and one flag per leaf");
    int st = WH_OK;
    const int rc = wh::adapt_trees_impl(P, &F, 1, p, &leaf_flags, &st);
    return rc ? rc : st;
}

int wh_adapt_trees(wh_pool *P, wh_forest *const *forests, int n, const wh_adapt_params *params, int *status)
{
    WH_PROF("api.adapt_trees");
    WH_CHECK(P && n >= 0 && (n == 0 || (forests && params && status)), WH_E_ARG, "wh_adapt_trees: null argument");
    for (int i = 0; i < n; ++i) {
        WH_CHECK(forests[i] && params[i].indicator != WH_IND_EXTERNAL, WH_E_ARG,
                 "wh_adapt_trees: null forest or external indicator");
        for (int j = 0; j < i; ++j) WH_CHECK(forests[j] != forests[i], WH_E_ARG, "wh_adapt_trees: a forest listed twice");
    }
    return wh::adapt_trees_impl(P, forests, n, params, nullptr, status);
}

int wh_refine_trees(wh_pool *P, wh_forest *const *forests, int n, const int32_t *jmax, int order, int array,
                    int *status)
{
    WH_PROF("api.refine_trees");
    WH_CHECK(P && n >= 0 && (n == 0 || (forests && jmax && status)), WH_E_ARG, "wh_refine_trees: null argument");
    WH_CHECK(order == 2 || order == 4 || order == 6, WH_E_ARG, "wh_refine_trees: order must be 2, 4 or 6");
    for (int i = 0; i < n; ++i) {
        WH_CHECK(forests[i], WH_E_ARG, "wh_refine_trees: null forest");
        for (int j = 0; j < i; ++j) WH_CHECK(forests[j] != forests[i], WH_E_ARG, "wh_refine_trees: a forest listed twice");
    }
    return wh::refine_trees_impl(P, forests, n, jmax, order, array, status);
}

int wh_set_tables_from_forests(wh_pool *P, wh_forest *const *forests, int n, int g_rhs, int order, int *status,
                               int *nblocks)
{
    WH_PROF("api.set_tables_from_forests");
    WH_CHECK(P && n >= 0 && (n == 0 || (forests && status && nblocks)), WH_E_ARG,
             "wh_set_tables_from_forests: null argument");
    WH_CHECK(g_rhs >= 0 && g_rhs <= P->d.g && (order == 2 || order == 4 || order == 6), WH_E_ARG,
             "wh_set_tables_from_forests: g_rhs beyond the ghost nodes or order not 2, 4, 6");
    for (int i = 0; i < n; ++i) {
        WH_CHECK(forests[i], WH_E_ARG, "wh_set_tables_from_forests: null forest");
        for (int j = 0; j < i; ++j)
            WH_CHECK(forests[j] != forests[i], WH_E_ARG, "wh_set_tables_from_forests: a forest listed twice");
    }
    return wh::set_tables_impl(P, forests, n, g_rhs, order, status, nblocks);
}

int wh_create_masks(wh_pool *P, wh_forest *const *forests, int n, const wh_mask_params *params, int *status)
{
    WH_PROF("api.create_masks");
    WH_CHECK(P && n >= 0 && (n == 0 || (forests && params && status)), WH_E_ARG, "wh_create_masks: null argument");
    for (int i = 0; i < n; ++i) WH_CHECK(forests[i], WH_E_ARG, "wh_create_masks: null forest");
    return wh::create_masks_impl(P, forests, n, params, status);
}

}  // extern "C"

