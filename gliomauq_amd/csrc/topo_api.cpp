// topo_api.cpp -- host-only topology entry points (no device).
#include <memory>
#include <stdexcept>
#include <unordered_set>
#include <vector>

#include "../../include/gq_api.h"
#include "grid.hpp"

struct gq_topo {
    std::unique_ptr<gq::Grid> g;
    std::vector<int> order;  // leaf indices in Morton order
    void reorder() { order = gq::morton_order(*g); }
};

namespace {
template <typename F>
gq_status tguard(gq_topo *t, F &&f)
{
    if (!t) return GQ_ERR_ARG;
    try {
        f();
        return GQ_OK;
    } catch (const std::invalid_argument &) {
        return GQ_ERR_ARG;
    } catch (const std::bad_alloc &) {
        return GQ_ERR_NOMEM;
    } catch (...) {
        return GQ_ERR_TOPOLOGY;
    }
}
}  // namespace

extern "C" {

gq_status gq_topo_create(gq_topo **out, int bpd, int max_level, int order)
{
    if (!out) return GQ_ERR_ARG;
    *out = nullptr;
    try {
        auto t = std::make_unique<gq_topo>();
        t->g = std::make_unique<gq::Grid>(bpd, max_level, order);
        t->reorder();
        *out = t.release();
        return GQ_OK;
    } catch (const std::invalid_argument &) {
        return GQ_ERR_ARG;
    } catch (...) {
        return GQ_ERR_NOMEM;
    }
}

gq_status gq_topo_destroy(gq_topo *t)
{
    if (!t) return GQ_ERR_ARG;
    delete t;
    return GQ_OK;
}

gq_status gq_topo_leaves(gq_topo *t, int32_t *level, int32_t *index, int32_t *block_id, int capacity, int *n)
{
    return tguard(t, [&] {
        const auto &L = t->g->leaves();
        const int m = (int)L.size();
        if (n) *n = m;
        if (!level && !index && !block_id) return;
        if (capacity < m) throw std::invalid_argument("capacity");
        for (int k = 0; k < m; k++) {
            const gq::Leaf &lf = L[t->order[k]];
            if (level) level[k] = lf.level;
            if (block_id) block_id[k] = lf.block_id;
            if (index)
                for (int a = 0; a < 3; a++) index[3 * k + a] = lf.idx[a];
        }
    });
}

gq_status gq_topo_refine(gq_topo *t, const int32_t *ids, int n, int *n_out)
{
    return tguard(t, [&] {
        if (n < 0 || (n > 0 && !ids)) throw std::invalid_argument("ids");
        std::unordered_set<int> s(ids, ids + n);
        std::vector<int> parents = t->g->plan_refine(s);
        std::vector<gq::Leaf> add;
        for (int p : parents) {
            const gq::Leaf pl = t->g->leaves()[p];
            for (int c = 0; c < 8; c++) {
                gq::Leaf ch;
                ch.level = pl.level + 1;
                ch.idx[0] = 2 * pl.idx[0] + (c & 1);
                ch.idx[1] = 2 * pl.idx[1] + ((c >> 1) & 1);
                ch.idx[2] = 2 * pl.idx[2] + ((c >> 2) & 1);
                ch.block_id = t->g->new_block_id();
                ch.slot = -1;
                add.push_back(ch);
            }
        }
        if (!parents.empty()) t->g->replace_leaves(parents, add);
        t->reorder();
        if (n_out) *n_out = (int)parents.size();
    });
}

gq_status gq_topo_compress(gq_topo *t, const int32_t *ids, int n, int *n_out)
{
    return tguard(t, [&] {
        if (n < 0 || (n > 0 && !ids)) throw std::invalid_argument("ids");
        std::unordered_set<int> s(ids, ids + n);
        auto plan = t->g->plan_compress(s);
        std::vector<int> remove;
        std::vector<gq::Leaf> add;
        for (const auto &p : plan) {
            gq::Leaf pl;
            pl.level = p[0];
            pl.idx[0] = p[1];
            pl.idx[1] = p[2];
            pl.idx[2] = p[3];
            pl.block_id = t->g->new_block_id();
            pl.slot = -1;
            add.push_back(pl);
            for (int c = 0; c < 8; c++)
                remove.push_back(t->g->find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1),
                                            2 * p[3] + ((c >> 2) & 1)));
        }
        if (!plan.empty()) t->g->replace_leaves(remove, add);
        t->reorder();
        if (n_out) *n_out = (int)plan.size();
    });
}

gq_status gq_topo_point_terms(gq_topo *t, int leaf, const int32_t loc[3], int capacity, int32_t *out_leaf,
                              int32_t *out_cell, double *out_w, int *n)
{
    return tguard(t, [&] {
        const int m = (int)t->g->leaves().size();
        if (leaf < 0 || leaf >= m || !loc) throw std::invalid_argument("leaf");
        std::vector<int> inv(m);
        for (int k = 0; k < m; k++) inv[t->order[k]] = k;
        std::vector<gq::Term> scratch, terms;
        const int r = t->g->expand_point(t->order[leaf], loc[0], loc[1], loc[2], scratch, terms);
        if (n) *n = r;
        if (r > capacity) throw std::invalid_argument("capacity");
        for (int k = 0; k < r; k++) {
            out_leaf[k] = inv[terms[k].slot];
            out_cell[k] = terms[k].cell;
            out_w[k] = terms[k].w;
        }
    });
}

}  // extern "C"
