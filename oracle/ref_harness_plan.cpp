// ref_harness_plan.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own refine/compress planners, compiled from the
// reference sources:
//   MRAG::Refiner::createPlan               MRAGcore/MRAGRefiner.h:43-198
//   MRAG::Refiner_SpaceExtension::createPlan MRAGscience/MRAGRefiner_SpaceExtension.h:25-67
//   MRAG::Compressor::createPlan            MRAGcore/MRAGCompressor.h:68-146
// on caller-described leaf hierarchies.  The two container types the planners
// are written against are the ones MRAGrid.h:33-34 declares.  The Grid itself
// (MRAGrid.h) needs the TBB headers and is not compiled; the harness builds
// the inputs the Grid would hand the planners:
//   * the hierarchy (Grid::_setup / _createChildren, MRAGridHelpers.h:20-55,
//     and Grid::refine step 4, MRAGrid.inl:288-313): every ancestor of a leaf
//     is an empty node, hierarchy[NULL] holds the level-0 root;
//   * the neighbourhood (Grid::_computeNeighborhood, MRAGrid.inl:841-1149,
//     adjacency FlattenedBlock::areAdjacent, MRAGridHelpers.h:76-84): for
//     every leaf, every other leaf and every periodic image of a leaf whose
//     finest-level box touches its box; images are separate GridNode objects
//     (the reference's ghost nodes) carrying the leaf's parent and blockID,
//     and only leaves are keys;
//   * the flags as Grid::refine / Grid::compress step 1 sets them
//     (MRAGrid.inl:258-266, :391-396).
#include "MRAGCommon.h"
#undef max
#undef min
#include <algorithm>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>
using namespace std;
#include "MRAGGridNode.h"
namespace MRAG
{
typedef map<GridNode *, vector<GridNode *> > HierarchyType;
typedef map<GridNode *, vector<GridNode *> > NeighborhoodType;
}
#include "MRAGRefinementPlan.h"
#include "MRAGRefiner.h"
#include "MRAGscience/MRAGRefiner_SpaceExtension.h"
#include "MRAGCompressor.h"
using namespace MRAG;

#define GQ_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

struct Key {
    int l, x, y, z;
    bool operator<(const Key &o) const
    {
        if (l != o.l) return l < o.l;
        if (z != o.z) return z < o.z;
        if (y != o.y) return y < o.y;
        return x < o.x;
    }
};

struct Forest {
    HierarchyType hierarchy;
    NeighborhoodType neighborhood;
    map<Key, GridNode *> nodes;  // every real node, leaves and ancestors
    vector<GridNode *> leaves;   // caller order
    vector<GridNode *> images;   // periodic-image (ghost) nodes
    bool ok;

    GridNode *ancestor(int l, int x, int y, int z)
    {
        Key k = {l, x, y, z};
        map<Key, GridNode *>::iterator it = nodes.find(k);
        if (it != nodes.end()) return it->second;
        GridNode *parent = l == 0 ? NULL : ancestor(l - 1, x >> 1, y >> 1, z >> 1);
        GridNode *n = new GridNode(true, parent, -1, x, y, z, l);
        nodes[k] = n;
        hierarchy[parent].push_back(n);
        hierarchy[n];
        return n;
    }

    Forest(int n_leaves, const int *level, const int *index) : ok(true)
    {
        for (int i = 0; i < n_leaves; i++) {
            const int l = level[i];
            const int *ix = index + 3 * i;
            GridNode *parent = l == 0 ? NULL : ancestor(l - 1, ix[0] >> 1, ix[1] >> 1, ix[2] >> 1);
            Key k = {l, ix[0], ix[1], ix[2]};
            if (nodes.count(k)) { ok = false; return; }
            GridNode *n = new GridNode(false, parent, i, ix[0], ix[1], ix[2], l);
            nodes[k] = n;
            hierarchy[parent].push_back(n);
            hierarchy[n];
            leaves.push_back(n);
        }
        // ancestors must be empty nodes: a leaf may not have leaf descendants
        for (size_t i = 0; i < leaves.size(); i++)
            if (!hierarchy[leaves[i]].empty()) ok = false;
        int maxl = 0;
        for (int i = 0; i < n_leaves; i++) maxl = std::max(maxl, level[i]);
        // periodic images of every leaf (all 26 shifts by one domain length)
        struct Box { GridNode *node; int s[3], e[3]; };
        vector<Box> boxes;
        for (size_t i = 0; i < leaves.size(); i++) {
            GridNode *b = leaves[i];
            const int d = maxl - b->level;
            const int n = 1 << b->level;
            for (int c = 0; c < 27; c++) {
                const int o[3] = {c % 3 - 1, (c / 3) % 3 - 1, (c / 9) % 3 - 1};
                GridNode *t = b;
                if (c != 13) {
                    t = new GridNode(b->isEmpty, b->parent, b->blockID, b->index[0] + o[0] * n, b->index[1] + o[1] * n,
                                     b->index[2] + o[2] * n, b->level);
                    images.push_back(t);
                }
                Box bx;
                bx.node = t;
                for (int a = 0; a < 3; a++) {
                    bx.s[a] = t->index[a] << d;
                    bx.e[a] = (t->index[a] + 1) << d;
                }
                boxes.push_back(bx);
            }
        }
        for (size_t i = 0; i < leaves.size(); i++) {
            const Box &a = boxes[27 * i + 13];
            vector<GridNode *> &nb = neighborhood[a.node];
            for (size_t j = 0; j < boxes.size(); j++) {
                const Box &b = boxes[j];
                if (b.node == a.node) continue;
                bool adj = true;
                for (int k = 0; k < 3; k++) {
                    const int gap = a.s[k] < b.s[k] ? (b.s[k] - a.e[k]) : (a.s[k] - b.e[k]);
                    adj = adj && gap <= 0;
                }
                if (adj) nb.push_back(b.node);
            }
        }
    }

    ~Forest()
    {
        for (map<Key, GridNode *>::iterator it = nodes.begin(); it != nodes.end(); ++it) delete it->second;
        for (size_t i = 0; i < images.size(); i++) delete images[i];
    }
};

}  // namespace

// Grid::refine planning with Refiner_SpaceExtension(resJump = 1, max_level)
// for leaves flagged[i] != 0.  Writes the caller indices of the leaves the
// plan splits, in increasing order; returns their number, -1 on overflow,
// -2 on an invalid hierarchy.
GQ_EXPORT int gqref_plan_refine(int n_leaves, const int *level, const int *index, const unsigned char *flagged,
                                int max_level, int cap, int *out)
{
    Forest f(n_leaves, level, index);
    if (!f.ok) return -2;
    for (int i = 0; i < n_leaves; i++) f.leaves[i]->shouldBeRefined = flagged[i] != 0;
    Refiner_SpaceExtension refiner(1, max_level);
    const bool dirs[3] = {true, true, true};
    vector<NodeToRefine> v;
    RefinementPlan *plan = refiner.createPlan(f.hierarchy, f.neighborhood, dirs, v);
    vector<int> ids;
    for (size_t i = 0; i < v.size(); i++) ids.push_back(v[i].node->blockID);
    delete plan;
    std::sort(ids.begin(), ids.end());
    if ((int)ids.size() > cap) return -1;
    for (size_t i = 0; i < ids.size(); i++) out[i] = ids[i];
    return (int)ids.size();
}

// Grid::compress planning with Compressor(resJump = 1) for leaves
// flagged[i] != 0.  Writes (level, ix, iy, iz) of every parent the plan
// collapses, sorted; returns their number, -1 on overflow, -2 on an invalid
// hierarchy.
GQ_EXPORT int gqref_plan_compress(int n_leaves, const int *level, const int *index, const unsigned char *flagged,
                                  int cap, int *out)
{
    Forest f(n_leaves, level, index);
    if (!f.ok) return -2;
    for (int i = 0; i < n_leaves; i++) f.leaves[i]->shouldBeCompressed = flagged[i] != 0;
    Compressor compressor(1);
    vector<GridNodeCollapseInfo> v;
    CompressionPlan *plan = compressor.createPlan(f.hierarchy, f.neighborhood, v);
    vector<Key> parents;
    for (size_t i = 0; i < v.size(); i++) {
        const GridNode *p = v[i].node;
        Key k = {p->level, p->index[0], p->index[1], p->index[2]};
        parents.push_back(k);
    }
    delete plan;
    std::sort(parents.begin(), parents.end());
    if ((int)parents.size() > cap) return -1;
    for (size_t i = 0; i < parents.size(); i++) {
        out[4 * i] = parents[i].l;
        out[4 * i + 1] = parents[i].x;
        out[4 * i + 2] = parents[i].y;
        out[4 * i + 3] = parents[i].z;
    }
    return (int)parents.size();
}
