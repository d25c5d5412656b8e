// grid.cpp -- host-side block hierarchy, planning and ghost tables.
#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_set>

namespace gq {

namespace {
inline int64_t floordiv(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
inline int64_t wrapi(int64_t i, int64_t n)
{
    i %= n;
    return i < 0 ? i + n : i;
}
// Hs of the interpolating wavelets (MRAGWavelets_StaticData.cpp:16,34).
inline double hs(int order, int i)
{
    if (order == 2) {
        if (i < -1 || i > 1) return 0.0;
        return i == 0 ? 1.0 : 0.5;
    }
    switch (i) {
    case -3: case 3: return -1.0 / 16.0;
    case -1: case 1: return 9.0 / 16.0;
    case 0: return 1.0;
    default: return 0.0;
    }
}
inline bool term_less(const Term &a, const Term &b)
{
    for (int k = 0; k < 4; k++) {
        if (a.key[k] != b.key[k]) return a.key[k] < b.key[k];
    }
    return false;
}
inline bool term_eq(const Term &a, const Term &b)
{
    return a.key[0] == b.key[0] && a.key[1] == b.key[1] && a.key[2] == b.key[2] && a.key[3] == b.key[3];
}
}  // namespace

Grid::Grid(int bpd, int max_level, int order) : bpd_(bpd), max_level_(max_level), order_(order)
{
    int l = 0;
    while ((1 << l) < bpd) l++;
    if ((1 << l) != bpd || bpd < 1) throw std::invalid_argument("blocks_per_dim must be a power of two");
    if (order != 2 && order != 4) throw std::invalid_argument("wavelet order must be 2 or 4");
    if (max_level < l) throw std::invalid_argument("max_level below the initial level");
    if (max_level > 8) throw std::invalid_argument("max_level above 8 is not supported");
    base_level_ = l;
    lut_level_ = max_level;
    // Grid::_setup: uniform blocks at the level log2(bpd) (MRAGrid.inl:93-111)
    for (int z = 0; z < bpd; z++)
        for (int y = 0; y < bpd; y++)
            for (int x = 0; x < bpd; x++) {
                Leaf lf;
                lf.level = l;
                lf.idx[0] = x; lf.idx[1] = y; lf.idx[2] = z;
                lf.block_id = next_id_++;
                lf.slot = -1;
                leaves_.push_back(lf);
            }
    rebuild_index();
}

void Grid::rebuild_index()
{
    key2leaf_.clear();
    key2leaf_.reserve(leaves_.size() * 2);
    const int64_t nF = 1LL << lut_level_;
    lut_.assign((size_t)(nF * nF * nF), -1);
    for (int i = 0; i < (int)leaves_.size(); i++) {
        const Leaf &lf = leaves_[i];
        key2leaf_[leaf_key(lf.level, lf.idx[0], lf.idx[1], lf.idx[2])] = i;
        const int d = lut_level_ - lf.level;
        const int64_t s = 1LL << d;
        for (int64_t z = (int64_t)lf.idx[2] << d; z < ((int64_t)lf.idx[2] << d) + s; z++)
            for (int64_t y = (int64_t)lf.idx[1] << d; y < ((int64_t)lf.idx[1] << d) + s; y++)
                for (int64_t x = (int64_t)lf.idx[0] << d; x < ((int64_t)lf.idx[0] << d) + s; x++)
                    lut_[(size_t)(x + nF * (y + nF * z))] = i;
    }
    for (int v : lut_)
        if (v < 0) throw std::logic_error("leaves do not cover the domain");
}

int Grid::find(int level, int x, int y, int z) const
{
    auto it = key2leaf_.find(leaf_key(level, x, y, z));
    return it == key2leaf_.end() ? -1 : it->second;
}

int Grid::lying_leaf(int lam, const int64_t p[3]) const
{
    const int64_t nF = 1LL << lut_level_;
    int64_t f[3];
    for (int d = 0; d < 3; d++) {
        int64_t q = p[d];
        if (lut_level_ >= lam) q = q * (1LL << (lut_level_ - lam));
        else q = floordiv(q, 1LL << (lam - lut_level_));
        f[d] = wrapi(floordiv(q, BS), nF);
    }
    return lut_[(size_t)(f[0] + nF * (f[1] + nF * f[2]))];
}

void Grid::neighbors(int i, std::vector<std::pair<int, bool>> &out) const
{
    // Ring of finest cells around the leaf's box; a leaf is a neighbour if
    // its box touches this box (face, edge or corner), directly or through a
    // periodic image (Grid::_computeNeighborhood with FlattenedBlock::
    // areAdjacent, MRAGridHelpers.h:63-72).
    out.clear();
    const Leaf &lf = leaves_[i];
    const int64_t nF = 1LL << lut_level_;
    const int d = lut_level_ - lf.level;
    int64_t bs[3], be[3];
    for (int a = 0; a < 3; a++) {
        bs[a] = (int64_t)lf.idx[a] << d;
        be[a] = ((int64_t)lf.idx[a] + 1) << d;
    }
    std::set<std::pair<int, bool>> seen;
    for (int64_t z = bs[2] - 1; z <= be[2]; z++)
        for (int64_t y = bs[1] - 1; y <= be[1]; y++)
            for (int64_t x = bs[0] - 1; x <= be[0]; x++) {
                const bool inside = x >= bs[0] && x < be[0] && y >= bs[1] && y < be[1] && z >= bs[2] && z < be[2];
                if (inside) {
                    // skip the interior quickly
                    if (x == bs[0]) x = be[0] - 1;
                    continue;
                }
                const bool img = !(x >= 0 && x < nF && y >= 0 && y < nF && z >= 0 && z < nF);
                const int k = lut_[(size_t)(wrapi(x, nF) + nF * (wrapi(y, nF) + nF * wrapi(z, nF)))];
                if (k == i && !img) continue;
                seen.insert(std::make_pair(k, img));
            }
    out.assign(seen.begin(), seen.end());
}

std::vector<int> Grid::plan_refine(const std::unordered_set<int> &flagged_ids) const
{
    const int n = (int)leaves_.size();
    std::vector<char> flag(n, 0);
    for (int i = 0; i < n; i++) flag[i] = flagged_ids.count(leaves_[i].block_id) ? 1 : 0;
    std::vector<std::pair<int, bool>> nb;
    // Refiner_SpaceExtension step 3: a leaf with a finer, non-image neighbour
    // is refined as well (MRAGRefiner_SpaceExtension.h:35-51)
    for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        neighbors(i, nb);
        for (auto &p : nb)
            if (!p.second && leaves_[p.first].level > leaves_[i].level) { flag[i] = 1; break; }
    }
    // step 4: never refine on the finest allowed level (:54-63)
    for (int i = 0; i < n; i++)
        if (leaves_[i].level == max_level_) flag[i] = 0;
    // Refiner::createPlan step 3: chain reaction on the level jump
    // (MRAGRefiner.h:66-152) with resJump = 1
    std::vector<char> victim(flag);
    std::vector<int> old_v, new_v;
    for (int i = 0; i < n; i++)
        if (victim[i]) old_v.push_back(i);
    while (!old_v.empty()) {
        new_v.clear();
        for (int v : old_v) {
            const int nl = leaves_[v].level + 1;
            neighbors(v, nb);
            for (auto &p : nb) {
                if (std::abs(leaves_[p.first].level - nl) > 1 && !victim[p.first]) {
                    victim[p.first] = 1;
                    new_v.push_back(p.first);
                }
            }
        }
        std::swap(old_v, new_v);
    }
    std::vector<int> out;
    for (int i = 0; i < n; i++)
        if (victim[i]) out.push_back(i);
    return out;
}

std::vector<std::array<int, 4>> Grid::plan_compress(const std::unordered_set<int> &flagged_ids) const
{
    // Compressor::createPlan (MRAGCompressor.h:68-118): collapse a parent
    // whose 8 children are all leaves, all flagged, and whose children's
    // neighbours are within one level of the parent.
    std::set<std::array<int, 4>> parents;
    for (const Leaf &lf : leaves_)
        if (lf.level > 0) parents.insert({lf.level - 1, lf.idx[0] >> 1, lf.idx[1] >> 1, lf.idx[2] >> 1});
    std::vector<std::array<int, 4>> out;
    std::vector<std::pair<int, bool>> nb;
    for (const auto &p : parents) {
        int ch[8];
        bool ok = true;
        for (int c = 0; c < 8 && ok; c++) {
            ch[c] = find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1), 2 * p[3] + ((c >> 2) & 1));
            ok = ch[c] >= 0 && flagged_ids.count(leaves_[ch[c]].block_id);
        }
        if (!ok) continue;
        for (int c = 0; c < 8 && ok; c++) {
            neighbors(ch[c], nb);
            for (auto &q : nb)
                if (std::abs(leaves_[q.first].level - p[0]) > 1) { ok = false; break; }
        }
        if (ok) out.push_back(p);
    }
    return out;
}

void Grid::replace_leaves(const std::vector<int> &remove, const std::vector<Leaf> &add)
{
    std::vector<char> rm(leaves_.size(), 0);
    for (int i : remove) rm[i] = 1;
    std::vector<Leaf> nl;
    nl.reserve(leaves_.size() - remove.size() + add.size());
    for (size_t i = 0; i < leaves_.size(); i++)
        if (!rm[i]) nl.push_back(leaves_[i]);
    for (const Leaf &a : add) nl.push_back(a);
    leaves_.swap(nl);
    rebuild_index();
}

void Grid::check_jump() const
{
    std::vector<std::pair<int, bool>> nb;
    for (int i = 0; i < (int)leaves_.size(); i++) {
        neighbors(i, nb);
        for (auto &p : nb)
            if (std::abs(leaves_[p.first].level - leaves_[i].level) > 1)
                throw std::runtime_error("level jump > 1 between neighbouring blocks");
    }
}

void Grid::expand_rec(int lam, const int64_t p[3], double w, std::vector<Term> &T, int depth) const
{
    if (depth > 40) throw std::runtime_error("ghost expansion does not terminate");
    const int li = lying_leaf(lam, p);
    const Leaf &lf = leaves_[li];
    if (lf.level == lam) {
        const int64_t n = (int64_t)BS << lam;
        Term t;
        int loc[3];
        for (int d = 0; d < 3; d++) {
            const int64_t q = wrapi(p[d], n);
            loc[d] = (int)(q - (int64_t)lf.idx[d] * BS);
            t.key[3 - d] = q;  // key: level, z, y, x
        }
        t.key[0] = lam;
        t.slot = li;  // leaf index for now; mapped to slots by the caller
        t.cell = loc[0] + BS * (loc[1] + BS * loc[2]);
        t.w = w;
        T.push_back(t);
        return;
    }
    if (lf.level > lam) {  // finer leaf: analysis with Ha = {1}
        const int64_t q[3] = {2 * p[0], 2 * p[1], 2 * p[2]};
        expand_rec(lam + 1, q, w, T, depth + 1);
        return;
    }
    // coarser leaf: synthesis from level lam-1 (BastardGhost::generateRequests,
    // MRAG_BBInfoCreator.h: supports ceil((p+HsS0)/2) .. floor((p+HsS1-1)/2))
    const int lo = order_ == 4 ? -3 : -1, hi = order_ == 4 ? 4 : 2;
    int64_t s[3], e[3];
    for (int d = 0; d < 3; d++) {
        s[d] = -floordiv(-(p[d] + lo), 2);
        e[d] = 1 + floordiv(p[d] + hi - 1, 2);
    }
    int64_t i[3];
    for (i[2] = s[2]; i[2] < e[2]; i[2]++)
        for (i[1] = s[1]; i[1] < e[1]; i[1]++)
            for (i[0] = s[0]; i[0] < e[0]; i[0]++) {
                const double ww = hs(order_, (int)(2 * i[0] - p[0])) * hs(order_, (int)(2 * i[1] - p[1])) *
                                  hs(order_, (int)(2 * i[2] - p[2]));
                if (ww == 0.0) continue;
                expand_rec(lam - 1, i, w * ww, T, depth + 1);
            }
}

int Grid::expand_point(int leaf, int lx, int ly, int lz, std::vector<Term> &scratch, std::vector<Term> &out) const
{
    scratch.clear();
    const Leaf &lf = leaves_[leaf];
    const int64_t p[3] = {(int64_t)lf.idx[0] * BS + lx, (int64_t)lf.idx[1] * BS + ly, (int64_t)lf.idx[2] * BS + lz};
    expand_rec(lf.level, p, 1.0, scratch, 0);
    std::sort(scratch.begin(), scratch.end(), term_less);
    out.clear();
    for (const Term &t : scratch) {
        if (!out.empty() && term_eq(out.back(), t)) out.back().w += t.w;
        else out.push_back(t);
    }
    size_t r = 0;
    for (size_t k = 0; k < out.size(); k++)
        if (out[k].w != 0.0) out[r++] = out[k];
    out.resize(r);
    return (int)r;
}

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
namespace {
struct Region {
    int s[3], e[3];
};
inline Region region_of(int c, int s0, int s1)
{
    Region r;
    const int d[3] = {c % 3 - 1, (c / 3) % 3 - 1, c / 9 - 1};
    for (int a = 0; a < 3; a++) {
        r.s[a] = d[a] < 0 ? s0 : (d[a] == 0 ? 0 : BS);
        r.e[a] = d[a] < 0 ? 0 : (d[a] == 0 ? BS : BS + s1 - 1);
    }
    return r;
}
// face f -> region code with a 1-cell halo
inline int face_code(int f)
{
    static const int codes[6] = {12, 14, 10, 16, 4, 22};
    return codes[f];
}

}  // namespace

void build_tables(const Grid &g, int nslots, Tables &t, int nthreads)
{
    const auto &L = g.leaves();
    const int n = (int)L.size();
    const int s0 = g.lab_s0(), s1 = g.lab_s1();
    t.nslots = nslots;
    t.fnbr.assign((size_t)nslots * 6, -1);
    t.fbase.assign((size_t)nslots * 6, -1);
    t.nbr26.assign((size_t)nslots * 26, -1);
    t.rbase26.assign((size_t)nslots * 26, -1);
    t.lab_row_begin.assign((size_t)nslots, 0);
    t.lab_row_end.assign((size_t)nslots, 0);

    // same-level neighbours per region (periodic wrap on the leaf's level)
    std::vector<char> reg_cross((size_t)n * 27, 0);
    for (int i = 0; i < n; i++) {
        const Leaf &lf = L[i];
        const int m = 1 << lf.level;
        for (int c = 0; c < 27; c++) {
            if (c == 13) continue;
            const int d[3] = {c % 3 - 1, (c / 3) % 3 - 1, c / 9 - 1};
            const int j = g.find(lf.level, (int)wrapi(lf.idx[0] + d[0], m), (int)wrapi(lf.idx[1] + d[1], m),
                                 (int)wrapi(lf.idx[2] + d[2], m));
            const int cc = c < 13 ? c : c - 1;
            if (j >= 0) t.nbr26[(size_t)lf.slot * 26 + cc] = L[j].slot;
            else reg_cross[(size_t)i * 27 + c] = 1;
        }
        for (int f = 0; f < 6; f++) {
            const int cc = face_code(f) < 13 ? face_code(f) : face_code(f) - 1;
            t.fnbr[(size_t)lf.slot * 6 + f] = t.nbr26[(size_t)lf.slot * 26 + cc];
        }
    }

    // neighbour signatures (block ids) decide which cached rows stay valid
    std::vector<std::vector<int>> sig(n);
    std::vector<char> need(n, 0);
    {
        std::vector<std::pair<int, bool>> nb;
        for (int i = 0; i < n; i++) {
            g.neighbors(i, nb);
            std::vector<int> &v = sig[i];
            for (auto &p : nb) v.push_back(L[p.first].block_id);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            auto it = t.cache.find(L[i].block_id);
            need[i] = (it == t.cache.end() || it->second.nbr_ids != v) ? 1 : 0;
        }
    }
    std::vector<int> todo;
    for (int i = 0; i < n; i++)
        if (need[i]) todo.push_back(i);
    std::vector<LeafRowCache> fresh(todo.size());
    auto work = [&](int tid, int nt) {
        std::vector<Term> scratch, terms;
        for (size_t q = tid; q < todo.size(); q += nt) {
            const int i = todo[q];
            LeafRowCache &out = fresh[q];
            for (int c = 0; c < 27; c++) {
                if (c == 13 || !reg_cross[(size_t)i * 27 + c]) continue;
                const Region r = region_of(c, s0, s1);
                for (int z = r.s[2]; z < r.e[2]; z++)
                    for (int y = r.s[1]; y < r.e[1]; y++)
                        for (int x = r.s[0]; x < r.e[0]; x++) {
                            g.expand_point(i, x, y, z, scratch, terms);
                            out.counts.push_back((int)terms.size());
                            for (const Term &tm : terms) {
                                out.src.push_back(L[tm.slot].slot * BS3 + tm.cell);
                                out.w.push_back(tm.w);
                            }
                        }
            }
        }
    };
    int nt = std::max(1, std::min(nthreads, (int)todo.size()));
    if (nt <= 1) work(0, 1);
    else {
        std::vector<std::thread> th;
        std::vector<std::exception_ptr> errs(nt);
        for (int k = 0; k < nt; k++)
            th.emplace_back([&, k]() {
                try { work(k, nt); } catch (...) { errs[k] = std::current_exception(); }
            });
        for (auto &x : th) x.join();
        for (auto &e : errs)
            if (e) std::rethrow_exception(e);
    }
    for (size_t q = 0; q < todo.size(); q++) {
        const int i = todo[q];
        fresh[q].nbr_ids = std::move(sig[i]);
        t.cache[L[i].block_id] = std::move(fresh[q]);
    }
    t.last_expanded = (int)todo.size();
    t.last_reused = n - (int)todo.size();
    // drop cache entries of leaves that no longer exist
    {
        std::unordered_set<int> live;
        for (const Leaf &lf : L) live.insert(lf.block_id);
        for (auto it = t.cache.begin(); it != t.cache.end();)
            it = live.count(it->first) ? std::next(it) : t.cache.erase(it);
    }

    // concatenate in leaf order; face rows are the lab rows of the 1-cell halo
    t.face_rows.clear();
    t.lab_rows.clear();
    size_t tot = 0;
    for (int i = 0; i < n; i++) tot += t.cache[L[i].block_id].src.size();
    t.lab_rows.src.reserve(tot);
    t.lab_rows.w.reserve(tot);
    for (int i = 0; i < n; i++) {
        const Leaf &lf = L[i];
        const LeafRowCache &lc = t.cache[lf.block_id];
        t.lab_row_begin[lf.slot] = t.lab_rows.nrows();
        int k = 0;
        size_t off = 0;
        for (int c = 0; c < 27; c++) {
            if (c == 13 || !reg_cross[(size_t)i * 27 + c]) continue;
            const int cc = c < 13 ? c : c - 1;
            const int row0 = t.lab_rows.nrows();
            t.rbase26[(size_t)lf.slot * 26 + cc] = row0;
            const Region r = region_of(c, s0, s1);
            const int npts = (r.e[0] - r.s[0]) * (r.e[1] - r.s[1]) * (r.e[2] - r.s[2]);
            for (int q0 = 0; q0 < npts; q0++, k++) {
                const int cnt = lc.counts[k];
                t.lab_rows.src.insert(t.lab_rows.src.end(), lc.src.begin() + off, lc.src.begin() + off + cnt);
                t.lab_rows.w.insert(t.lab_rows.w.end(), lc.w.begin() + off, lc.w.begin() + off + cnt);
                off += cnt;
                t.lab_rows.ptr.push_back((int)t.lab_rows.src.size());
            }
            // a cross-level face region: its 1-cell-halo plane becomes the step rows
            int f = -1;
            for (int ff = 0; ff < 6; ff++)
                if (face_code(ff) == c) f = ff;
            if (f >= 0) {
                t.fbase[(size_t)lf.slot * 6 + f] = t.face_rows.nrows();
                const int ex = r.e[0] - r.s[0], ey = r.e[1] - r.s[1];
                for (int v = 0; v < BS; v++)
                    for (int u = 0; u < BS; u++) {
                        int x, y, z;
                        switch (f) {
                        case 0: x = -1; y = u; z = v; break;
                        case 1: x = BS; y = u; z = v; break;
                        case 2: x = u; y = -1; z = v; break;
                        case 3: x = u; y = BS; z = v; break;
                        case 4: x = u; y = v; z = -1; break;
                        default: x = u; y = v; z = BS; break;
                        }
                        const int lr = row0 + (x - r.s[0]) + ex * ((y - r.s[1]) + ey * (z - r.s[2]));
                        for (int q = t.lab_rows.ptr[lr]; q < t.lab_rows.ptr[lr + 1]; q++) {
                            t.face_rows.src.push_back(t.lab_rows.src[q]);
                            t.face_rows.w.push_back(t.lab_rows.w[q]);
                        }
                        t.face_rows.ptr.push_back((int)t.face_rows.src.size());
                    }
            }
        }
        t.lab_row_end[lf.slot] = t.lab_rows.nrows();
    }
    if (t.face_rows.src.size() > (size_t)INT32_MAX || t.lab_rows.src.size() > (size_t)INT32_MAX)
        throw std::runtime_error("ghost tables exceed 2^31 terms");

    // base-level look-up for the uniform dump / likelihood
    const int bpd = g.blocks_per_dim();
    const int bl = g.base_level();
    t.lut_slot.assign((size_t)bpd * bpd * bpd, -1);
    for (int z = 0; z < bpd; z++)
        for (int y = 0; y < bpd; y++)
            for (int x = 0; x < bpd; x++) {
                const int64_t p[3] = {(int64_t)x * BS, (int64_t)y * BS, (int64_t)z * BS};
                const int li = g.lying_leaf(bl, p);
                t.lut_slot[(size_t)x + bpd * ((size_t)y + bpd * z)] = L[li].slot;
            }
}

std::vector<int> morton_order(const Grid &g)
{
    const auto &L = g.leaves();
    const int lv = g.lut_level();
    std::vector<std::pair<uint64_t, int>> keys;
    keys.reserve(L.size());
    for (int i = 0; i < (int)L.size(); i++) {
        const int d = lv - L[i].level;
        uint64_t x = (uint64_t)L[i].idx[0] << d, y = (uint64_t)L[i].idx[1] << d, z = (uint64_t)L[i].idx[2] << d;
        uint64_t m = 0;
        for (int b = 0; b < 20; b++)
            m |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2);
        keys.push_back({m, i});
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int> out;
    out.reserve(keys.size());
    for (auto &k : keys) out.push_back(k.second);
    return out;
}

}  // namespace gq
