// grid.cpp -- host-side block hierarchy, planning and ghost tables.
#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <set>
#include <stdexcept>
#include <string>
#include <mutex>
#include <utility>
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
    const int64_t nF = 1LL << lut_level_;
    lut_.assign((size_t)(nF * nF * nF), -1);
    for (int i = 0; i < (int)leaves_.size(); i++) {
        const Leaf &lf = leaves_[i];
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
    // the finest-level look-up cell at the box's corner holds the leaf if it exists
    if (level < 0 || level > lut_level_) return -1;
    const int64_t n = 1LL << level;
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return -1;
    const int d = lut_level_ - level;
    const int64_t nF = 1LL << lut_level_;
    const int i = lut_[(size_t)(((int64_t)x << d) + nF * (((int64_t)y << d) + nF * ((int64_t)z << d)))];
    const Leaf &lf = leaves_[i];
    return (lf.level == level && lf.idx[0] == x && lf.idx[1] == y && lf.idx[2] == z) ? i : -1;
}

int Grid::lying_leaf(int lam, const int64_t p[3]) const
{
    // every extent is a power of two: floor division and periodic wrap are
    // arithmetic shifts and masks
    const int64_t mask = (1LL << lut_level_) - 1;
    int64_t f[3];
    for (int d = 0; d < 3; d++) {
        int64_t q = p[d];
        if (lut_level_ >= lam) q = q * (1LL << (lut_level_ - lam));
        else q = q >> (lam - lut_level_);
        f[d] = (q >> 4) & mask;  // BS = 16
    }
    return lut_[(size_t)(f[0] + (mask + 1) * (f[1] + (mask + 1) * f[2]))];
}

template <class F>
bool Grid::any_ring(int i, F &&f) const
{
    // Ring of finest look-up cells around the leaf's box; a leaf is a
    // neighbour if its box touches this box (face, edge or corner), directly
    // or through a periodic image (Grid::_computeNeighborhood with
    // FlattenedBlock::areAdjacent, MRAGridHelpers.h:63-72).  f(k, image)
    // returning true stops the walk; a neighbour may be visited repeatedly.
    const Leaf &lf = leaves_[i];
    const int64_t nF = 1LL << lut_level_, mask = nF - 1;
    const int d = lut_level_ - lf.level;
    int64_t bs[3], be[3];
    for (int a = 0; a < 3; a++) {
        bs[a] = (int64_t)lf.idx[a] << d;
        be[a] = ((int64_t)lf.idx[a] + 1) << d;
    }
    for (int64_t z = bs[2] - 1; z <= be[2]; z++) {
        const bool zin = z >= bs[2] && z < be[2];
        for (int64_t y = bs[1] - 1; y <= be[1]; y++) {
            const bool yzin = zin && y >= bs[1] && y < be[1];
            const int64_t row = nF * ((y & mask) + nF * (z & mask));
            const bool yzimg = y < 0 || y >= nF || z < 0 || z >= nF;
            // inside the box in y and z: only the two x ends of the row
            const int64_t step = yzin ? be[0] - bs[0] + 1 : 1;
            for (int64_t x = bs[0] - 1; x <= be[0]; x += step) {
                const bool img = yzimg || x < 0 || x >= nF;
                const int k = lut_[(size_t)((x & mask) + row)];
                if (k == i && !img) continue;
                if (f(k, img)) return true;
            }
        }
    }
    return false;
}

template <class F>
bool Grid::any_ring_balanced(int i, F &&f) const
{
    // The ring walk of any_ring on a grid whose neighbouring leaves differ by
    // at most one level (the invariant every plan keeps): a neighbour's box
    // is at least half the leaf's box s and aligned to its own size, so per
    // axis the look-up cells bs - 1, bs, bs + s/2 and be meet every
    // neighbour (the full ring has (s+2)^3 - s^3 cells; this walk at most 56).
    const Leaf &lf = leaves_[i];
    const int d = lut_level_ - lf.level;
    if (d <= 1) return any_ring(i, std::forward<F>(f));  // the full ring has at most 56 cells
    const int64_t nF = 1LL << lut_level_, mask = nF - 1;
    const int64_t sz = 1LL << d;
    int64_t c[3][4];
    bool in[3][4];
    int n = 0;
    for (int a = 0; a < 3; a++) {
        const int64_t bs = (int64_t)lf.idx[a] << d;
        int k = 0;
        c[a][k] = bs - 1, in[a][k++] = false;
        c[a][k] = bs, in[a][k++] = true;
        c[a][k] = bs + sz / 2, in[a][k++] = true;
        c[a][k] = bs + sz, in[a][k++] = false;
        n = k;
    }
    for (int kz = 0; kz < n; kz++) {
        const int64_t z = c[2][kz];
        for (int ky = 0; ky < n; ky++) {
            const int64_t y = c[1][ky];
            const int64_t row = nF * ((y & mask) + nF * (z & mask));
            const bool yzimg = y < 0 || y >= nF || z < 0 || z >= nF;
            for (int kx = 0; kx < n; kx++) {
                if (in[0][kx] && in[1][ky] && in[2][kz]) continue;
                const int64_t x = c[0][kx];
                const bool img = yzimg || x < 0 || x >= nF;
                const int k = lut_[(size_t)((x & mask) + row)];
                if (k == i && !img) continue;
                if (f(k, img)) return true;
            }
        }
    }
    return false;
}

template <class F>
bool Grid::any_other_level(int i, const int *nbr26, F &&f) const
{
    // The 26 same-level boxes around leaf i (periodic); where the table has
    // no same-level leaf, the box lies inside one coarser leaf or is split
    // into finer leaves, and the leaf holding its corner look-up cell has that
    // coarser or finer level.  Every neighbour of another level touching the
    // leaf is found this way: a coarser one contains the box on its side, a
    // finer one splits it.  f(k, image) per such box, image = the box is
    // reached through the periodic boundary; true stops the walk.
    const Leaf &lf = leaves_[i];
    const int *nb = nbr26 + (size_t)lf.slot * 26;
    const int m = 1 << lf.level, sh = lut_level_ - lf.level;
    const int64_t nF = 1LL << lut_level_;
    for (int c = 0; c < 26; c++) {
        if (nb[c] >= 0) continue;
        const int cc = c < 13 ? c : c + 1;
        const int d[3] = {cc % 3 - 1, cc / 3 % 3 - 1, cc / 9 - 1};
        bool img = false;
        int64_t q[3];
        for (int a = 0; a < 3; a++) {
            const int v = lf.idx[a] + d[a];
            img = img || v < 0 || v >= m;
            q[a] = (int64_t)(v & (m - 1)) << sh;
        }
        if (f(lut_[(size_t)(q[0] + nF * (q[1] + nF * q[2]))], img)) return true;
    }
    return false;
}

void Grid::neighbors(int i, std::vector<std::pair<int, bool>> &out) const
{
    out.clear();
    out.reserve(64);
    any_ring(i, [&](int k, bool img) {
        out.emplace_back(k, img);
        return false;
    });
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

std::vector<char> Grid::id_flags(const std::unordered_set<int> &ids) const
{
    std::vector<char> f(next_id_, 0);
    for (int id : ids)
        if (id >= 0 && id < next_id_) f[id] = 1;
    return f;
}

std::vector<int> Grid::plan_refine(const std::unordered_set<int> &flagged_ids) const
{
    return plan_refine(id_flags(flagged_ids));
}

std::vector<std::array<int, 4>> Grid::plan_compress(const std::unordered_set<int> &flagged_ids) const
{
    return plan_compress(id_flags(flagged_ids));
}

std::vector<int> Grid::plan_refine(const std::vector<char> &flag_by_id, const int *nbr26) const
{
    const int n = (int)leaves_.size();
    const int nf = (int)flag_by_id.size();
    std::vector<char> flag(n, 0);
    int top = 0;  // finest leaf level
    for (int i = 0; i < n; i++) {
        const int id = leaves_[i].block_id;
        flag[i] = id < nf && flag_by_id[id] ? 1 : 0;
        top = std::max(top, leaves_[i].level);
    }
    // Refiner_SpaceExtension step 3: a leaf with a finer, non-image neighbour
    // is refined as well (MRAGRefiner_SpaceExtension.h:35-51); leaves on the
    // finest level present have none, and step 4 unflags the finest allowed
    // level anyway
    for (int i = 0; i < n; i++) {
        const int li = leaves_[i].level;
        if (flag[i] || li >= top || li >= max_level_) continue;
        auto finer = [&](int k, bool img) { return !img && leaves_[k].level > li; };
        if (nbr26 ? any_other_level(i, nbr26, finer) : any_ring_balanced(i, finer)) flag[i] = 1;
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
            auto jump = [&](int k, bool) {
                if (std::abs(leaves_[k].level - nl) > 1 && !victim[k]) {
                    victim[k] = 1;
                    new_v.push_back(k);
                }
                return false;
            };
            if (nbr26) any_other_level(v, nbr26, jump);
            else any_ring_balanced(v, jump);
        }
        std::swap(old_v, new_v);
    }
    std::vector<int> out;
    for (int i = 0; i < n; i++)
        if (victim[i]) out.push_back(i);
    return out;
}

std::vector<std::array<int, 4>> Grid::plan_compress(const std::vector<char> &flag_by_id, const int *nbr26) const
{
    // Compressor::createPlan (MRAGCompressor.h:68-118): collapse a parent
    // whose 8 children are all leaves, all flagged, and whose children's
    // neighbours are within one level of the parent.  Flagged leaves grouped
    // by parent: eight distinct leaves with one parent are its eight
    // children.  Parents come out in (level, x, y, z) order.
    const int nf = (int)flag_by_id.size();
    std::vector<uint64_t> keys;
    for (const Leaf &lf : leaves_)
        if (lf.level > 0 && lf.block_id < nf && flag_by_id[lf.block_id])
            keys.push_back(leaf_key(lf.level - 1, lf.idx[0] >> 1, lf.idx[1] >> 1, lf.idx[2] >> 1));
    std::sort(keys.begin(), keys.end());
    std::vector<std::array<int, 4>> out;
    for (size_t a = 0; a + 8 <= keys.size();) {
        if (keys[a] != keys[a + 7]) {
            a++;
            continue;
        }
        const uint64_t k = keys[a];
        const std::array<int, 4> p = {(int)(k >> 60), (int)((k >> 40) & 0xfffff), (int)((k >> 20) & 0xfffff),
                                      (int)(k & 0xfffff)};
        bool ok = true;
        for (int c = 0; c < 8 && ok; c++) {
            const int ch = find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1), 2 * p[3] + ((c >> 2) & 1));
            auto jump = [&](int q, bool) { return std::abs(leaves_[q].level - p[0]) > 1; };
            ok = ch >= 0 && !(nbr26 ? any_other_level(ch, nbr26, jump) : any_ring_balanced(ch, jump));
        }
        if (ok) out.push_back(p);
        a += 8;
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
    for (int i = 0; i < (int)leaves_.size(); i++) {
        const int li = leaves_[i].level;
        if (any_ring(i, [&](int k, bool) { return std::abs(leaves_[k].level - li) > 1; }))
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
void build_sample_tables(const Grid &g, int *fnbr, int *nbr26, SampleTables &t)
{
    static const int face_codes[6] = {12, 14, 10, 16, 4, 22};
    const auto &L = g.leaves();
    t.order = morton_order(g);
    t.faces.clear();
    const std::vector<int> &gl = g.lut();
    const int lutl = g.lut_level();
    const size_t nF = (size_t)1 << lutl;
    // device look-up entries first: the neighbour search below reads slot and
    // level of a look-up cell in one load
    t.lut.resize(gl.size());
    for (size_t k = 0; k < gl.size(); k++) t.lut[k] = L[gl[k]].slot | L[gl[k]].level << LUT_LEVEL_SHIFT;
    const int *lut = t.lut.data();
    for (int i : t.order) {
        const Leaf &lf = L[i];
        const int m = 1 << lf.level, sh = lutl - lf.level;
        int *nb26 = nbr26 + (size_t)lf.slot * 26;
        // a same-level neighbour box is a leaf iff the leaf holding its corner
        // look-up cell has that level (boxes of one level are aligned)
        size_t xs[3], ys[3], zs[3];
        for (int d = 0; d < 3; d++) {
            xs[d] = (size_t)((lf.idx[0] + d - 1) & (m - 1)) << sh;
            ys[d] = nF * ((size_t)((lf.idx[1] + d - 1) & (m - 1)) << sh);
            zs[d] = nF * nF * ((size_t)((lf.idx[2] + d - 1) & (m - 1)) << sh);
        }
        const int want = lf.level << LUT_LEVEL_SHIFT;
        int c = 0;
        for (int dz = 0; dz < 3; dz++)
            for (int dy = 0; dy < 3; dy++) {
                const int *row = lut + ys[dy] + zs[dz];
                for (int dx = 0; dx < 3; dx++, c++) {
                    if (c == 13) continue;
                    const int e = row[xs[dx]];
                    nb26[c < 13 ? c : c - 1] = (e & ~LUT_SLOT_MASK) == want ? (e & LUT_SLOT_MASK) : -1;
                }
            }
        for (int f = 0; f < 6; f++) {
            const int cc = face_codes[f];
            const int nb = nb26[cc < 13 ? cc : cc - 1];
            fnbr[(size_t)lf.slot * 6 + f] = nb;
            if (nb < 0) t.faces.push_back(lf.slot * 6 + f);
        }
    }
    // base-level look-up for the uniform dump / likelihood: the leaf holding
    // base block (x, y, z) holds finest look-up cell (x, y, z) << (lut_level - base_level)
    const int bpd = g.blocks_per_dim();
    const int d = lutl - g.base_level();
    t.lut_slot.resize((size_t)bpd * bpd * bpd);
    for (int z = 0; z < bpd; z++)
        for (int y = 0; y < bpd; y++) {
            const int *row = t.lut.data() + nF * (((size_t)y << d) + nF * ((size_t)z << d));
            int *out = t.lut_slot.data() + (size_t)bpd * ((size_t)y + (size_t)bpd * z);
            for (int x = 0; x < bpd; x++) out[x] = row[(size_t)x << d] & LUT_SLOT_MASK;
        }
}

namespace {
// Interleave the low 21 bits of x with two zero bits between each.
inline uint64_t spread3(uint64_t x)
{
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}
}  // namespace

std::vector<int> morton_order(const Grid &g)
{
    const auto &L = g.leaves();
    const int lv = g.lut_level();
    if (lv <= 6) {
        // Leaves are aligned boxes of look-up cells, each a contiguous run of
        // the cells' Morton order starting at its corner: walking the look-up
        // in Morton order and taking each leaf at its first cell gives the
        // leaves in Morton order of their corners.
        static std::vector<int> walk[7];
        static std::once_flag once[7];
        std::call_once(once[lv], [&] {
            const int64_t n = 1LL << lv;
            std::vector<int> &w = walk[lv];
            w.resize((size_t)(n * n * n));
            for (int64_t m = 0; m < n * n * n; m++) {
                int64_t c[3] = {0, 0, 0};
                for (int b = 0; b < lv; b++)
                    for (int a = 0; a < 3; a++) c[a] |= ((m >> (3 * b + a)) & 1) << b;
                w[(size_t)m] = (int)(c[0] + n * (c[1] + n * c[2]));
            }
        });
        const std::vector<int> &lut = g.lut();
        std::vector<char> seen(L.size(), 0);
        std::vector<int> out;
        out.reserve(L.size());
        for (int cell : walk[lv]) {
            const int i = lut[(size_t)cell];
            if (!seen[i]) {
                seen[i] = 1;
                out.push_back(i);
            }
        }
        return out;
    }
    // codes use 3 * lut_level <= 24 bits: code << 32 | leaf index sorts as one integer
    std::vector<uint64_t> keys(L.size());
    for (size_t i = 0; i < L.size(); i++) {
        const int d = lv - L[i].level;
        const uint64_t m = spread3((uint64_t)L[i].idx[0] << d) | spread3((uint64_t)L[i].idx[1] << d) << 1 |
                           spread3((uint64_t)L[i].idx[2] << d) << 2;
        keys[i] = m << 32 | (uint64_t)i;
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int> out(keys.size());
    for (size_t k = 0; k < keys.size(); k++) out[k] = (int)(keys[k] & 0xffffffffu);
    return out;
}

}  // namespace gq
