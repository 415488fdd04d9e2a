// wh_forest_impl.h -- the host forest (block list of one tree) shared by the
// forest helpers (wh_forest.cpp) and the adapt_tree orchestration
// (wh_adapt.cpp).  Host code only.
//
// Treecode layout (LIB/TREE/module_treelib.f90:107-283): int64 with dim bits
// per level; the digit of level l sits at bit offset (Jmax-l)*dim; digit bits
// (x, y, z) = (2, 1, 4) (refinementExecute.f90:98-105).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wabbit_hip.h"

namespace wh {
void set_error(const std::string &msg);

struct FBlock {
    int64_t tc;
    int level;
    int hvy;
    int status;  // refinement status (module_globals.f90:25-32)
};

inline uint64_t fkey(int level, int64_t tc) { return ((uint64_t)level << 58) ^ (uint64_t)tc; }

// The generated lists of one full-tree ghost sync of adapt_tree / sync_tree
// (or of coarse-extension boxes) with blocks
// named by their node index in the tree (recv, send, box and filter ids),
// keyed by everything the generation reads: the sync's
// parameters, the pool's geometry and wavelet, and the tree's nodes
// (treecode, level, status) in order.  A step whose tree equals an earlier
// one's reuses the lists (mapped to the nodes' current pool slots) instead
// of generating them again.
struct TreeSyncCache {
    std::vector<int64_t> key;
    std::vector<wh_patch> st1, st23;  // sync stage 1, stages 2-3
    std::vector<int32_t> fids[2];  // HD-filtered senders per source array
    std::vector<wh_box> fboxes[2];
    uint64_t used = 0;
};
}  // namespace wh

struct wh_forest {
    int dim = 2, jmax = 0, g = 0;
    int bs[3] = {1, 1, 1};
    double domain[3] = {1, 1, 1};
    int periodic[3] = {1, 1, 1};
    int nb_max = 0;
    int slot0 = 0;                          // the forest's pool slots: [slot0, slot0 + nb_max)
    std::vector<wh::FBlock> blocks;         // active blocks (leaf list)
    std::unordered_map<uint64_t, int> map;  // (level, tc) -> index in blocks, levels beyond the dense tables
    std::vector<int> free_hvy;              // free pool slots

    // bumped by every change of the block list (all changes end in rebuild_map)
    uint64_t version = 0;
    // the last generated ghost-patch list: a count query and the fill call
    // that follows it generate once
    mutable std::vector<wh_patch> pcache;
    mutable int pc_gm = -1, pc_gp = -1, pc_order = -1;
    mutable uint64_t pc_version = ~0ull;
    // the last generation with blocks named by block-list index, and its
    // key (ghost widths, order, geometry, every block's treecode and level):
    // a block list that repeats in new pool slots reuses it
    mutable std::vector<wh_patch> pc_idx;
    mutable std::vector<int64_t> pc_key;
    // generated lists of adapt_tree's full-tree ghost syncs and coarse
    // extension boxes (least recently used dropped beyond TS_CACHE; a step
    // of adapt_tree keys about two dozen)
    static constexpr int TS_CACHE = 32;
    std::vector<wh::TreeSyncCache> tscache;
    uint64_t ts_clock = 0;

    // (level, grid index) -> index in blocks as a dense table per level with
    // at most 2^DENSE_BITS cells (the hash map beyond); the entries set by the
    // last rebuild are reset by the next, so a rebuild costs O(blocks)
    static constexpr int DENSE_BITS = 22;
    std::vector<std::vector<int>> dense;
    std::vector<std::pair<int, size_t>> dense_set;
    size_t lin(int level, const int q[3]) const
    {
        return ((size_t)q[2] << (2 * level)) | ((size_t)q[1] << level) | (size_t)q[0];
    }
    // rebuilds the lookup; returns false if two blocks share (level, treecode)
    bool rebuild_map()
    {
        ++version;
        map.clear();
        for (const auto &e : dense_set) dense[e.first][e.second] = -1;
        dense_set.clear();
        bool unique = true;
        for (size_t i = 0; i < blocks.size(); ++i) {
            const wh::FBlock &b = blocks[i];
            if (dim * b.level > DENSE_BITS) {
                unique = map.emplace(wh::fkey(b.level, b.tc), (int)i).second && unique;
                continue;
            }
            if ((int)dense.size() <= b.level) dense.resize(b.level + 1);
            std::vector<int> &t = dense[b.level];
            if (t.empty()) t.assign((size_t)1 << (dim * b.level), -1);
            int q[3];
            ijk_of(b.tc, b.level, q);
            const size_t k = lin(b.level, q);
            if (t[k] >= 0) {
                unique = false;
                continue;
            }
            t[k] = (int)i;
            dense_set.push_back({b.level, k});
        }
        return unique;
    }
    int digit_shift(int level) const { return (jmax - level) * dim; }
    void ijk_of(int64_t tc, int level, int ijk[3]) const
    {
        ijk[0] = ijk[1] = ijk[2] = 0;
        for (int l = 1; l <= level; ++l) {
            const int dgt = (int)((tc >> digit_shift(l)) & ((1 << dim) - 1));
            ijk[0] = 2 * ijk[0] + ((dgt >> 1) & 1);
            ijk[1] = 2 * ijk[1] + (dgt & 1);
            ijk[2] = 2 * ijk[2] + (dim == 3 ? ((dgt >> 2) & 1) : 0);
        }
    }
    int64_t tc_of(const int ijk[3], int level) const
    {
        int64_t tc = 0;
        for (int l = 1; l <= level; ++l) {
            const int sh = level - l;
            const int bx = (ijk[0] >> sh) & 1, by = (ijk[1] >> sh) & 1;
            const int bz = dim == 3 ? ((ijk[2] >> sh) & 1) : 0;
            tc |= (int64_t)(by + 2 * bx + 4 * bz) << digit_shift(l);
        }
        return tc;
    }
    // block at grid index ijk on level (periodic wrap); -1 if not active.
    int find(int level, const int ijk_in[3], int wrapped[3]) const
    {
        if (level < 0) return -1;
        const int n = 1 << level;
        for (int d = 0; d < 3; ++d) {
            int v = ijk_in[d];
            if (d < dim) {
                if (v < 0 || v >= n) {
                    if (!periodic[d]) return -1;
                    v = ((v % n) + n) % n;
                }
            } else {
                v = 0;
            }
            wrapped[d] = v;
        }
        if (dim * level <= DENSE_BITS) {
            if (level >= (int)dense.size() || dense[level].empty()) return -1;  // no block on this level
            return dense[level][lin(level, wrapped)];
        }
        auto it = map.find(wh::fkey(level, tc_of(wrapped, level)));
        return it == map.end() ? -1 : it->second;
    }
    void x0dx(const wh::FBlock &b, double out[6]) const
    {
        int ijk[3];
        ijk_of(b.tc, b.level, ijk);
        for (int d = 0; d < 3; ++d) {
            if (d < dim) {
                // get_block_spacing_origin_b: dx = 2^-J * L / Bs, x0 = ((ijk)*Bs) * dx
                const double dx = std::ldexp(1.0, -b.level) * domain[d] / (double)bs[d];
                out[3 + d] = dx;
                out[d] = (double)(ijk[d] * bs[d]) * dx;
            } else {
                out[d] = 0.0;
                out[3 + d] = 0.0;
            }
        }
    }
};
