// grid.hpp -- host-side block hierarchy and ghost tables.
//
// A leaf-block octree with MRAG's semantics (16^3 blocks, unit cube,
// periodic, vertex-centred interpolating wavelets):
//   * topology, neighbourhood with periodic images  (MRAGcore/MRAGrid.inl:841-1149)
//   * refine planning: Refiner_SpaceExtension + Refiner::createPlan chain
//     reaction (MRAGscience/MRAGRefiner_SpaceExtension.h:25-67,
//     MRAGcore/MRAGRefiner.h:43-198)
//   * compress planning: Compressor::createPlan (MRAGcore/MRAGCompressor.h:68-146)
//   * ghost tables: the value MRAG_BBInfoCreator assigns to every ghost point
//     (MRAGcore/MRAG_BBInfoCreator.inl:20-1315), flattened to one row of
//     (source cell, weight) terms per ghost point.  Terms reaching the same
//     leaf point are merged and rows are sorted canonically (leaf level, then
//     global z, y, x at that level), so an FP64 evaluation of a row is
//     independent of block numbering.
//
// The device never sees this structure: it receives flat per-slot tables.
#pragma once

#include <array>
#include <cstdint>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace gq {

constexpr int BS = 16;          // cells per block edge
constexpr int BS3 = BS * BS * BS;
// device look-up entries: slot | level << LUT_LEVEL_SHIFT (slots stay below
// 2^19: the pool holds at most 2^31 cells; levels are at most 8)
constexpr int LUT_LEVEL_SHIFT = 24;
constexpr int LUT_SLOT_MASK = (1 << LUT_LEVEL_SHIFT) - 1;

struct Leaf {
    int level;
    int idx[3];
    int block_id;
    int slot;  // device pool slot
};

struct Term {
    int slot;
    int cell;
    double w;
    int64_t key[4];  // level, gz, gy, gx
};

// Flattened ghost rows.
struct Rows {
    std::vector<int> ptr;    // nrows + 1
    std::vector<int> src;    // slot*BS3 + cell
    std::vector<double> w;
    void clear() { ptr.assign(1, 0); src.clear(); w.clear(); }
    int nrows() const { return (int)ptr.size() - 1; }
};

class Grid {
public:
    Grid(int blocks_per_dim, int max_level, int order);

    int order() const { return order_; }
    int max_level() const { return max_level_; }
    int base_level() const { return base_level_; }
    int blocks_per_dim() const { return bpd_; }
    // finest level used for location look-ups (>= every leaf level)
    int lut_level() const { return lut_level_; }

    const std::vector<Leaf> &leaves() const { return leaves_; }
    const std::vector<int> &lut() const { return lut_; }  // finest block cells -> leaf index
    std::vector<Leaf> &leaves_mut() { return leaves_; }

    // Wide lab stencil [s0, s1) of this wavelet (grid max stencil,
    // MRAGrid.inl:762-789): Interp4 [-2,4), Interp2 [-1,2).
    int lab_s0() const { return order_ == 4 ? -2 : -1; }
    int lab_s1() const { return order_ == 4 ? 4 : 2; }

    int find(int level, int x, int y, int z) const;  // leaf index or -1
    // leaf whose region contains the level-`lam` point p (unwrapped coords)
    int lying_leaf(int lam, const int64_t p[3]) const;

    // neighbours of leaf i: pairs (leaf index, via periodic image)
    void neighbors(int i, std::vector<std::pair<int, bool>> &out) const;

    // Planning.  Return leaf indices (refine) or parent keys (compress).
    // Flags by block id: flag[id] != 0 (ids at or above flag.size() unflagged).
    // nbr26: optional slot-indexed same-level neighbour table of exactly these
    // leaves (as build_sample_tables writes it); with it the planners look
    // only at the directions where a leaf meets another level (same plans).
    std::vector<int> plan_refine(const std::vector<char> &flag_by_id, const int *nbr26 = nullptr) const;
    std::vector<std::array<int, 4>> plan_compress(const std::vector<char> &flag_by_id,
                                                  const int *nbr26 = nullptr) const;
    std::vector<int> plan_refine(const std::unordered_set<int> &flagged_ids) const;
    std::vector<std::array<int, 4>> plan_compress(const std::unordered_set<int> &flagged_ids) const;
    std::vector<char> id_flags(const std::unordered_set<int> &ids) const;
    int id_bound() const { return next_id_; }  // every block id is below this

    // Topology edits (slots are assigned by the caller).
    int new_block_id() { return next_id_++; }
    void replace_leaves(const std::vector<int> &remove, const std::vector<Leaf> &add);
    void check_jump() const;  // throws on a level jump > 1

    // Ghost point expansion (merged, canonical order).
    int expand_point(int leaf, int lx, int ly, int lz, std::vector<Term> &scratch, std::vector<Term> &out) const;

private:
    template <class F>
    bool any_ring(int i, F &&f) const;
    template <class F>
    bool any_ring_balanced(int i, F &&f) const;  // any_ring on a 2:1-balanced grid
    template <class F>
    bool any_other_level(int i, const int *nbr26, F &&f) const;  // via the same-level table
    void rebuild_index();
    void expand_rec(int lam, const int64_t p[3], double w, std::vector<Term> &T, int depth) const;

    int bpd_, base_level_, max_level_, order_, lut_level_;
    int next_id_ = 0;
    std::vector<Leaf> leaves_;
    std::vector<int> lut_;  // (2^lut_level)^3 -> leaf index
};

inline uint64_t leaf_key(int l, int x, int y, int z)
{
    return ((uint64_t)l << 60) | ((uint64_t)(uint32_t)x << 40) | ((uint64_t)(uint32_t)y << 20) | (uint64_t)(uint32_t)z;
}

// Per-sample device tables derived from a Grid.  Ghost values across levels
// are evaluated on the device (depth-first over the finest-level look-up);
// the host only provides neighbour slots and the look-ups.
struct SampleTables {
    std::vector<int> order;     // leaf indices in Morton order
    std::vector<int> faces;     // slot*6+f of every cross-level face, leaves in Morton order
    std::vector<int> lut;       // (2^lut_level)^3 finest block cells -> slot | level << LUT_LEVEL_SHIFT
    std::vector<int> lut_slot;  // bpd^3 base-level blocks -> slot (uniform view)
};

// Writes, for every leaf slot of g (other slots are left untouched):
//   fnbr[slot*6+f]   same-level face neighbour slot or -1
//                    (f = 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)
//   nbr26[slot*26+c] same-level neighbour per wide lab region or -1
//                    (c = (dx+1) + 3(dy+1) + 9(dz+1), the centre left out)
// and fills t.  Grids whose leaves own disjoint slots can be built in parallel.
void build_sample_tables(const Grid &g, int *fnbr, int *nbr26, SampleTables &t);

// Morton order of the leaves (for the launch work list).
std::vector<int> morton_order(const Grid &g);

}  // namespace gq
