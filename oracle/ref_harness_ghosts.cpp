// ref_harness_ghosts.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own ghost-table builder, MRAG::MRAG_BBInfoCreator
// (MRAGcore/MRAG_BBInfoCreator.{h,inl}, MRAG_SmartBlockFinder.h,
// MRAGBoundaryBlockInfo.{h,cpp}, MRAGWavelets_StaticData.cpp), compiled from
// the reference sources, on caller-described toy hierarchies of leaf blocks.
// The neighbour list handed to createBoundaryInfoBlock is formed the way
// Grid::_computeNeighborhood forms it (MRAGrid.inl:841-1149,
// MRAGridHelpers.h:63-108): every leaf, and every periodic image of a leaf,
// whose finest-level box touches the target's box.  The instruction lists are
// flattened to (leaf, cell, weight) terms with weight = w0*w1*w2 taken from
// the double weight pool; the caller merges and compares them.
#include "MRAGCommon.h"
#undef max
#undef min
#include <algorithm>
#include <cmath>
#include <vector>
using namespace std;
#include "MRAGBlock.h"
#include "MRAGWavelets_Interp2ndOrder.h"
#include "MRAGWavelets_Interp4thOrder.h"
#include "MRAGBoundaryBlockInfo.h"
#include "MRAG_BBInfoCreator.h"
using namespace MRAG;

#define GQ_EXPORT extern "C" __attribute__((visibility("default")))

namespace {
template <typename W, int B>
int run(int n_leaves, const int *level, const int *index, int target, int s0, int s1,
        int cap, int *out_off, int *out_leaf, int *out_cell, double *out_w)
{
    typedef Block<float, B, B, B> BT;
    int maxl = 0;
    for (int i = 0; i < n_leaves; i++) maxl = std::max(maxl, level[i]);
    vector<GridNode *> nodes;
    for (int i = 0; i < n_leaves; i++)
        nodes.push_back(new GridNode(false, NULL, i, index[3 * i], index[3 * i + 1], index[3 * i + 2], level[i]));
    vector<GridNode *> images;
    vector<GridNode *> neighbors;
    const GridNode &b = *nodes[target];
    const int bd = maxl - b.level;
    const int bs[3] = {b.index[0] << bd, b.index[1] << bd, b.index[2] << bd};
    const int be[3] = {(b.index[0] + 1) << bd, (b.index[1] + 1) << bd, (b.index[2] + 1) << bd};
    for (int i = 0; i < n_leaves; i++) {
        const int n = 1 << level[i];
        for (int c = 0; c < 27; c++) {
            const int k[3] = {c % 3 - 1, (c / 3) % 3 - 1, (c / 9) % 3 - 1};
            if (i == target && k[0] == 0 && k[1] == 0 && k[2] == 0) continue;
            const int idx[3] = {index[3 * i] + k[0] * n, index[3 * i + 1] + k[1] * n, index[3 * i + 2] + k[2] * n};
            const int d = maxl - level[i];
            bool adj = true;
            for (int a = 0; a < 3; a++) {
                const int s = idx[a] << d, e = (idx[a] + 1) << d;
                const int gap = s < bs[a] ? (bs[a] - s - (e - s)) : (s - bs[a] - (be[a] - bs[a]));
                adj = adj && gap <= 0;
            }
            if (!adj) continue;
            if (k[0] == 0 && k[1] == 0 && k[2] == 0) neighbors.push_back(nodes[i]);
            else {
                GridNode *g = new GridNode(false, NULL, i, idx[0], idx[1], idx[2], level[i]);
                images.push_back(g);
                neighbors.push_back(g);
            }
        }
    }
    const int ss[3] = {s0, s0, s0}, se[3] = {s1, s1, s1}, bsz[3] = {B, B, B};
    MRAG_BBInfoCreator<W, BT> creator(ss, se, bsz);
    BoundaryInfoBlock *bb = creator.createBoundaryInfoBlock(b, neighbors);
    bb->lock();
    const vector<vector<IndexWP> > &ghosts = bb->getGhosts();
    const vector<PointIndex> &pool = bb->getIndexPool();
    int nterm = 0, ng = 0, rc = 0;
    const int L = B + s1 - s0 - 1;
    // emit in lab z,y,x order over non-interior points
    for (int z = s0; z < B + s1 - 1; z++)
        for (int y = s0; y < B + s1 - 1; y++)
            for (int x = s0; x < B + s1 - 1; x++) {
                if (x >= 0 && x < B && y >= 0 && y < B && z >= 0 && z < B) continue;
                const int d[3] = {x < 0 ? -1 : (x < B ? 0 : 1), y < 0 ? -1 : (y < B ? 0 : 1), z < 0 ? -1 : (z < B ? 0 : 1)};
                const int code = (d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1);
                const int st[3] = {d[0] < 0 ? s0 : (d[0] == 0 ? 0 : B), d[1] < 0 ? s0 : (d[1] == 0 ? 0 : B), d[2] < 0 ? s0 : (d[2] == 0 ? 0 : B)};
                const int en[3] = {d[0] < 0 ? 0 : (d[0] == 0 ? B : B + s1 - 1), d[1] < 0 ? 0 : (d[1] == 0 ? B : B + s1 - 1), d[2] < 0 ? 0 : (d[2] == 0 ? B : B + s1 - 1)};
                const int sz[3] = {en[0] - st[0], en[1] - st[1], en[2] - st[2]};
                const int gi = bb->boundary[code].start + (x - st[0]) + sz[0] * ((y - st[1]) + sz[1] * (z - st[2]));
                out_off[ng] = nterm;
                const vector<IndexWP> &v = ghosts[gi];
                for (size_t k = 0; k < v.size(); k++) {
                    if (nterm >= cap) { rc = -1; break; }
                    const PointIndex &p = pool[v[k].point_index];
                    out_leaf[nterm] = p.blockID;
                    out_cell[nterm] = p.index;
                    out_w[nterm] = bb->weightsPool[v[k].weights_index[0]] * bb->weightsPool[v[k].weights_index[1]] *
                                   bb->weightsPool[v[k].weights_index[2]];
                    nterm++;
                }
                ng++;
            }
    out_off[ng] = nterm;
    (void)L;
    bb->release();
    delete bb;
    for (size_t i = 0; i < images.size(); i++) delete images[i];
    for (size_t i = 0; i < nodes.size(); i++) delete nodes[i];
    return rc < 0 ? rc : ng;
}
}  // namespace

// Returns the number of ghost points emitted (lab order), or -1 on overflow,
// -2 on unsupported arguments.
GQ_EXPORT int gqref_ghost_terms(int n_leaves, const int *level, const int *index, int target, int order, int B,
                                int s0, int s1, int cap, int *out_off, int *out_leaf, int *out_cell, double *out_w)
{
    if (order == 4 && B == 16) return run<Wavelets_Interp4thOrder, 16>(n_leaves, level, index, target, s0, s1, cap, out_off, out_leaf, out_cell, out_w);
    if (order == 2 && B == 16) return run<Wavelets_Interp2ndOrder, 16>(n_leaves, level, index, target, s0, s1, cap, out_off, out_leaf, out_cell, out_w);
    if (order == 4 && B == 8) return run<Wavelets_Interp4thOrder, 8>(n_leaves, level, index, target, s0, s1, cap, out_off, out_leaf, out_cell, out_w);
    if (order == 2 && B == 8) return run<Wavelets_Interp2ndOrder, 8>(n_leaves, level, index, target, s0, s1, cap, out_off, out_leaf, out_cell, out_w);
    return -2;
}
