// sanitize_topology.cpp -- CPU test driver for the product's host topology
// code (grid.cpp, topo_api.cpp), built with AddressSanitizer, LeakSanitizer
// and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.
//
// Runs long seeded sequences of the work the library's host side does around
// every adaptation -- refine/compress planning and topology edits, slot
// assignment as gq_api.cpp does it (free-list reuse), per-sample table builds
// (build_sample_tables), Morton ordering and ghost-point expansion -- and
// checks the level-jump invariant and the slot tables after every edit, and
// that the planners give the same plans through the same-level table as
// through their ring walks.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <unordered_set>
#include <vector>

#include "../../include/gq_api.h"
#include "../../gliomauq_amd/csrc/grid.hpp"

static int fail(const char *m, int it)
{
    std::fprintf(stderr, "FAIL %s at iteration %d\n", m, it);
    return 1;
}

static int run_grid(int bpd, int max_level, int order, int iters, unsigned seed)
{
    std::mt19937 rng(seed);
    gq::Grid g(bpd, max_level, order);
    std::vector<int> free_slots;
    int cap = 0;
    auto alloc = [&]() {
        if (free_slots.empty()) {
            const int ncap = cap + cap / 2 + 8;
            for (int s = ncap - 1; s >= cap; s--) free_slots.push_back(s);
            cap = ncap;
        }
        const int s = free_slots.back();
        free_slots.pop_back();
        return s;
    };
    for (auto &lf : g.leaves_mut()) lf.slot = alloc();
    std::vector<gq::Term> scratch, terms;
    for (int it = 0; it < iters; it++) {
        const auto &L = g.leaves();
        std::unordered_set<int> flagged;
        const double dens = std::uniform_real_distribution<double>(0.0, 1.0)(rng);
        for (const auto &lf : L)
            if (std::uniform_real_distribution<double>(0.0, 1.0)(rng) < dens * 0.5) flagged.insert(lf.block_id);
        if (it % 2 == 0) {
            std::vector<int> parents = g.plan_refine(flagged);
            std::vector<gq::Leaf> add;
            for (int p : parents) {
                const gq::Leaf pl = g.leaves()[p];
                if (pl.level >= max_level) return fail("refining a max-level leaf", it);
                free_slots.push_back(pl.slot);
                for (int c = 0; c < 8; c++) {
                    gq::Leaf ch;
                    ch.level = pl.level + 1;
                    for (int a = 0; a < 3; a++) ch.idx[a] = 2 * pl.idx[a] + ((c >> a) & 1);
                    ch.block_id = g.new_block_id();
                    ch.slot = alloc();
                    add.push_back(ch);
                }
            }
            if (!parents.empty()) g.replace_leaves(parents, add);
        } else {
            std::unordered_set<int> all;
            for (const auto &lf : L)
                if (std::uniform_real_distribution<double>(0.0, 1.0)(rng) < 0.9) all.insert(lf.block_id);
            auto plan = g.plan_compress(all);
            std::vector<int> remove;
            std::vector<gq::Leaf> add;
            for (const auto &p : plan) {
                gq::Leaf pl;
                pl.level = p[0];
                pl.idx[0] = p[1];
                pl.idx[1] = p[2];
                pl.idx[2] = p[3];
                pl.block_id = g.new_block_id();
                for (int c = 0; c < 8; c++) {
                    const int li = g.find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1),
                                          2 * p[3] + ((c >> 2) & 1));
                    if (li < 0) return fail("collapse child missing", it);
                    remove.push_back(li);
                    free_slots.push_back(g.leaves()[li].slot);
                }
                pl.slot = alloc();
                add.push_back(pl);
            }
            if (!plan.empty()) g.replace_leaves(remove, add);
        }
        g.check_jump();
        // tables over the whole slot range, as the context builds them
        std::vector<int> fnbr((size_t)cap * 6, -7), nbr26((size_t)cap * 26, -7);
        gq::SampleTables t;
        gq::build_sample_tables(g, fnbr.data(), nbr26.data(), t);
        if (t.order.size() != g.leaves().size()) return fail("morton order size", it);
        for (const auto &lf : g.leaves()) {
            if (lf.slot < 0 || lf.slot >= cap) return fail("slot range", it);
            for (int f = 0; f < 6; f++) {
                const int nb = fnbr[(size_t)lf.slot * 6 + f];
                if (nb < -1 || nb >= cap || nb == -7) return fail("fnbr entry", it);
            }
        }
        std::vector<int> slot_level(cap, -1);
        for (const auto &lf : g.leaves()) slot_level[lf.slot] = lf.level;
        for (int e : t.lut) {
            const int s = e & gq::LUT_SLOT_MASK;
            if (e < 0 || s >= cap || slot_level[s] != (e >> gq::LUT_LEVEL_SHIFT)) return fail("lut entry", it);
        }
        // the planners through the same-level table plan exactly as the ring walks
        for (int k = 0; k < 4; k++) {
            std::vector<char> fl(g.id_bound(), 0);
            const unsigned pct = k == 0 ? 2 : k == 1 ? 20 : k == 2 ? 60 : 95;
            for (const auto &lf : g.leaves())
                if (rng() % 100 < pct) fl[lf.block_id] = 1;
            if (g.plan_refine(fl) != g.plan_refine(fl, nbr26.data())) return fail("refine plan via table", it);
            if (g.plan_compress(fl) != g.plan_compress(fl, nbr26.data())) return fail("compress plan via table", it);
        }
        // ghost expansion of a few halo points of random leaves
        for (int k = 0; k < 8; k++) {
            const int leaf = (int)(rng() % g.leaves().size());
            const int lo = g.lab_s0(), hi = gq::BS + g.lab_s1() - 1;
            const int lx = lo + (int)(rng() % (hi - lo)), ly = lo + (int)(rng() % (hi - lo)),
                      lz = lo + (int)(rng() % (hi - lo));
            const int n = g.expand_point(leaf, lx, ly, lz, scratch, terms);
            double sw = 0;
            for (int q = 0; q < n; q++) sw += terms[q].w;
            if (n <= 0 || std::abs(sw - 1.0) > 1e-12) return fail("ghost weights do not sum to one", it);
        }
    }
    return 0;
}

static int run_api(int iters, unsigned seed)
{
    std::mt19937 rng(seed);
    gq_topo *t = nullptr;
    if (gq_topo_create(&t, 4, 4, 4) != GQ_OK) return fail("gq_topo_create", 0);
    std::vector<int32_t> lv, ix, id, ol, oc;
    std::vector<double> ow;
    for (int it = 0; it < iters; it++) {
        int n = 0;
        if (gq_topo_leaves(t, nullptr, nullptr, nullptr, 0, &n) != GQ_OK) return fail("leaves count", it);
        lv.resize(n);
        ix.resize(3 * n);
        id.resize(n);
        if (gq_topo_leaves(t, lv.data(), ix.data(), id.data(), n, &n) != GQ_OK) return fail("leaves", it);
        std::vector<int32_t> pick;
        for (int k = 0; k < n; k++)
            if (rng() % 5 == 0) pick.push_back(id[k]);
        int nout = 0;
        const gq_status st = it % 2 ? gq_topo_compress(t, pick.data(), (int)pick.size(), &nout)
                                    : gq_topo_refine(t, pick.data(), (int)pick.size(), &nout);
        if (st != GQ_OK) return fail("refine/compress", it);
        const int32_t loc[3] = {-1, 7, 16};
        ol.resize(4096);
        oc.resize(4096);
        ow.resize(4096);
        int m = 0;
        if (gq_topo_point_terms(t, (int)(rng() % n), loc, 4096, ol.data(), oc.data(), ow.data(), &m) != GQ_OK)
            return fail("point terms", it);
        // capacity too small: an error, no overflow
        if (m > 1 && gq_topo_point_terms(t, 0, loc, 0, ol.data(), oc.data(), ow.data(), &m) == GQ_OK)
            return fail("capacity not checked", it);
    }
    gq_topo_destroy(t);
    return 0;
}

int main(int argc, char **argv)
{
    const int iters = argc > 1 ? std::atoi(argv[1]) : 60;
    int rc = 0;
    rc |= run_grid(2, 4, 4, iters, 1);
    rc |= run_grid(4, 5, 4, iters, 2);
    rc |= run_grid(1, 3, 2, iters, 3);
    rc |= run_grid(8, 4, 2, iters / 2, 4);
    rc |= run_api(iters, 5);
    if (rc == 0) std::printf("ok\n");
    return rc;
}
