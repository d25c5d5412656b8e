// gq_api.cpp -- extern "C" layer: context, block pool, adaptation driver.
//
// Host C++ that owns the grid topology (grid.hpp) and the device state, and
// launches the HIP kernels (kernels.hpp) on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/gq_api.h"
#include "grid.hpp"
#include "kernels.hpp"

namespace {

struct gq_error : std::runtime_error {
    gq_status code;
    gq_error(gq_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw gq_error(e_ == hipErrorOutOfMemory ? GQ_ERR_NOMEM : GQ_ERR_HIP,                   \
                           std::string(#x) + ": " + hipGetErrorString(e_));                         \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;  // capacity in elements
    void reserve(size_t m)
    {
        if (m <= n) return;
        release();
        size_t cap = std::max<size_t>(m, 16);
        HIPCHK(hipMalloc((void **)&p, cap * sizeof(T)));
        n = cap;
    }
    void upload(const std::vector<T> &v, hipStream_t st)
    {
        reserve(v.size());
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ScopeTimer {
    double &acc;
    double t0;
    explicit ScopeTimer(double &a) : acc(a), t0(now_s()) {}
    ~ScopeTimer() { acc += now_s() - t0; }
};

int hw_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    return (int)std::min(n, 16u);
}

}  // namespace

struct gq_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;

    std::unique_ptr<gq::Grid> grid;
    // block pool
    int cap = 0;
    float *phi[2] = {nullptr, nullptr};
    int cur = 0;
    float *pw = nullptr, *pg = nullptr;
    std::vector<int> free_slots;

    // tables
    bool tables_valid = false;
    gq::Tables tabs;
    std::vector<int> api_order;  // leaf indices in API (Morton) order
    DevBuf<int> d_work, d_level, d_index, d_fnbr, d_fbase, d_face_ptr, d_face_src, d_nbr26, d_lut, d_lutf,
        d_faces, d_counts;
    int face_nrows = 0;
    DevBuf<double> d_ghost;  // cross-level lab points (3 fields for a split)
    DevBuf<float> d_face_w, d_fik, d_gphi, d_ih2;
    bool fik_valid = false;
    float fik_Dw = 0, fik_Dg = 0;

    // scratch
    DevBuf<int> d_list, d_list2;
    DevBuf<float> d_dense;
    DevBuf<double> d_dbl, d_partial;

    // observations
    int obs_flip_gpd = 0;
    DevBuf<uint8_t> d_t1u8, d_t2u8;
    DevBuf<float> d_petf;
    int obs_ref_gpd = 0, obs_ref_npts = 0;
    DevBuf<double> d_pet_ref, d_t1_ref, d_t2_ref;
    DevBuf<int> d_pts;

    // step tile (slab depth) chosen per grid-size bucket
    bool autotune = true;
    std::map<int, int> zt_cache;
    int last_zt = 16;
    double tune_us[4] = {0, 0, 0, 0};

    // host-side time accounting (seconds) and counts
    double t_tables = 0, t_plan = 0, t_fwt = 0, t_split = 0, t_collapse = 0, t_tune = 0;
    int n_tables = 0, n_fwt = 0;

    // timing
    bool per_kernel_timing = false;
    std::vector<hipEvent_t> events;
    double last_ms = 0;
    int last_launches = 0;
    int64_t last_cells = 0;

    ~gq_ctx()
    {
        for (auto e : events) (void)hipEventDestroy(e);
        release_pool();
        if (st) (void)hipStreamDestroy(st);
    }

    void release_pool()
    {
        for (float *&q : phi) {
            if (q) (void)hipFree(q);
            q = nullptr;
        }
        if (pw) (void)hipFree(pw);
        if (pg) (void)hipFree(pg);
        pw = pg = nullptr;
        cap = 0;
    }

    void need_grid() const
    {
        if (!grid) throw gq_error(GQ_ERR_STATE, "no grid: call gq_grid_create first");
    }

    // Grow the pool to at least `want` slots, keeping slot contents.
    void grow(int want)
    {
        if (want <= cap) return;
        int ncap = std::max(want, cap + cap / 2);
        const size_t nb = (size_t)ncap * gq::BS3 * sizeof(float);
        const size_t ob = (size_t)cap * gq::BS3 * sizeof(float);
        float *np[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < 4; k++) {
            if (hipMalloc((void **)&np[k], nb) != hipSuccess) {
                for (int j = 0; j < k; j++) (void)hipFree(np[j]);
                throw gq_error(GQ_ERR_NOMEM, "block pool allocation failed");
            }
            HIPCHK(hipMemsetAsync(np[k], 0, nb, st));
        }
        float *old[4] = {phi[0], phi[1], pw, pg};
        for (int k = 0; k < 4; k++)
            if (old[k] && ob) HIPCHK(hipMemcpyAsync(np[k], old[k], ob, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int k = 0; k < 4; k++)
            if (old[k]) (void)hipFree(old[k]);
        phi[0] = np[0];
        phi[1] = np[1];
        pw = np[2];
        pg = np[3];
        for (int s = ncap - 1; s >= cap; s--) free_slots.push_back(s);
        cap = ncap;
        tables_valid = false;
    }

    int alloc_slot()
    {
        if (free_slots.empty()) grow(cap + 8);
        int s = free_slots.back();
        free_slots.pop_back();
        return s;
    }

    float *field_ptr(int f)
    {
        switch (f) {
        case GQ_FIELD_PHI: return phi[cur];
        case GQ_FIELD_PWM: return pw;
        case GQ_FIELD_PGM: return pg;
        default: throw gq_error(GQ_ERR_ARG, "unknown field");
        }
    }

    void ensure_tables()
    {
        need_grid();
        if (tables_valid) return;
        ScopeTimer tm(t_tables);
        n_tables++;
        gq::Tables &t = tabs;
        gq::build_tables(*grid, cap, t);
        api_order = gq::morton_order(*grid);
        const auto &L = grid->leaves();
        std::vector<int> work(api_order.size());
        for (size_t k = 0; k < api_order.size(); k++) work[k] = L[api_order[k]].slot;
        std::vector<int> lvl((size_t)cap, 0), idx((size_t)cap * 3, 0);
        for (const auto &lf : L) {
            lvl[lf.slot] = lf.level;
            for (int a = 0; a < 3; a++) idx[(size_t)lf.slot * 3 + a] = lf.idx[a];
        }
        std::vector<int> fbase((size_t)cap * 6, -1);
        for (size_t k = 0; k < t.faces.size(); k++) fbase[t.faces[k]] = (int)k * 256;
        d_work.upload(work, st);
        d_level.upload(lvl, st);
        d_index.upload(idx, st);
        d_fnbr.upload(t.fnbr, st);
        d_fbase.upload(fbase, st);
        d_nbr26.upload(t.nbr26, st);
        d_lut.upload(t.lut_slot, st);
        d_lutf.upload(t.lut, st);
        d_faces.upload(t.faces, st);
        // step rows of the cross-level faces, enumerated on the device
        const int nfaces = (int)t.faces.size();
        face_nrows = nfaces * 256;
        if (nfaces > 0) {
            const gq::LabArgs la = lab_args();
            d_counts.reserve((size_t)face_nrows);
            HIPCHK(gq::launch_face_rows_count(la, d_faces.p, nfaces, d_counts.p, st));
            std::vector<int> cnt((size_t)face_nrows), ptr((size_t)face_nrows + 1);
            HIPCHK(hipMemcpyAsync(cnt.data(), d_counts.p, cnt.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            ptr[0] = 0;
            for (int r = 0; r < face_nrows; r++) {
                if ((int64_t)ptr[r] + cnt[r] > INT32_MAX) throw gq_error(GQ_ERR_NOMEM, "face rows exceed 2^31 terms");
                ptr[r + 1] = ptr[r] + cnt[r];
            }
            d_face_ptr.upload(ptr, st);
            d_face_src.reserve((size_t)std::max(1, ptr[face_nrows]));
            d_face_w.reserve((size_t)std::max(1, ptr[face_nrows]));
            HIPCHK(gq::launch_face_rows_fill(la, d_faces.p, nfaces, d_face_ptr.p, d_face_src.p, d_face_w.p, st));
        }
        d_gphi.reserve((size_t)std::max(1, face_nrows));
        d_fik.reserve((size_t)cap * 6 * 256);
        std::vector<float> ih2(16);
        for (int l = 0; l < 16; l++) {
            const float h = (float)(std::ldexp(1.0, -l) / gq::BS);  // BlockInfo::h (Real)
            const double hd = h;
            ih2[l] = (float)(1. / (hd * hd));
        }
        d_ih2.upload(ih2, st);
        HIPCHK(hipStreamSynchronize(st));  // the host vectors above are released on return
        tables_valid = true;
        fik_valid = false;
    }

    gq::LabArgs lab_args() const
    {
        gq::LabArgs a;
        a.nbr26 = d_nbr26.p;
        a.lut = d_lutf.p;
        a.level = d_level.p;
        a.index = d_index.p;
        a.lut_level = grid->lut_level();
        a.order = grid->order();
        return a;
    }

    void ensure_fik(float Dw, float Dg)
    {
        ensure_tables();
        if (fik_valid && Dw == fik_Dw && Dg == fik_Dg) return;
        HIPCHK(gq::launch_face_prep(pw, pg, d_work.p, (int)api_order.size(), d_fnbr.p, d_fbase.p, d_face_ptr.p,
                                    d_face_src.p, d_face_w.p, d_fik.p, Dw, Dg, st));
        fik_valid = true;
        fik_Dw = Dw;
        fik_Dg = Dg;
    }

    // detail per API-order leaf subset (leaf indices)
    std::vector<double> fwt(const std::vector<int> &leaf_idx)
    {
        ensure_tables();
        const auto &L = grid->leaves();
        std::vector<int> slots(leaf_idx.size());
        for (size_t k = 0; k < leaf_idx.size(); k++) slots[k] = L[leaf_idx[k]].slot;
        std::vector<double> out(leaf_idx.size());
        if (slots.empty()) return out;
        ScopeTimer tm(t_fwt);
        n_fwt++;
        d_list.upload(slots, st);
        d_dbl.reserve(slots.size());
        const gq::LabArgs la = lab_args();
        d_ghost.reserve(slots.size() * (size_t)gq::lab_points(grid->order()));
        HIPCHK(gq::launch_lab_ghosts(la, phi[cur], d_list.p, (int)slots.size(), d_ghost.p, false, st));
        HIPCHK(gq::launch_fwt_detail(grid->order(), la, d_ghost.p, phi[cur], d_list.p, (int)slots.size(), d_dbl.p,
                                     st));
        HIPCHK(hipMemcpyAsync(out.data(), d_dbl.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return out;
    }

    int refine_ids(const std::unordered_set<int> &ids)
    {
        need_grid();
        std::vector<int> parents;
        {
            ScopeTimer tm(t_plan);
            parents = grid->plan_refine(ids);
        }
        if (parents.empty()) return 0;
        ensure_tables();  // parent labs use the pre-refine topology
        ScopeTimer tm(t_split);
        grow((int)(grid->leaves().size() + parents.size() * 8));
        const auto &L = grid->leaves();
        std::vector<int> pslots, cslots;
        std::vector<gq::Leaf> add;
        for (int p : parents) {
            const gq::Leaf &pl = L[p];
            pslots.push_back(pl.slot);
            for (int c = 0; c < 8; c++) {
                gq::Leaf ch;
                ch.level = pl.level + 1;
                ch.idx[0] = 2 * pl.idx[0] + (c & 1);
                ch.idx[1] = 2 * pl.idx[1] + ((c >> 1) & 1);
                ch.idx[2] = 2 * pl.idx[2] + ((c >> 2) & 1);
                ch.block_id = grid->new_block_id();
                ch.slot = alloc_slot();
                cslots.push_back(ch.slot);
                add.push_back(ch);
            }
        }
        d_list.upload(pslots, st);
        d_list2.upload(cslots, st);
        const gq::LabArgs la = lab_args();
        const size_t per = pslots.size() * (size_t)gq::lab_points(grid->order());
        d_ghost.reserve(3 * per);
        const float *fl[3] = {phi[cur], pw, pg};
        for (int f = 0; f < 3; f++)
            HIPCHK(gq::launch_lab_ghosts(la, fl[f], d_list.p, (int)pslots.size(), d_ghost.p + f * per, true, st));
        HIPCHK(gq::launch_split(grid->order(), la, d_ghost.p, phi[cur], pw, pg, d_list.p, d_list2.p,
                                (int)pslots.size(), st));
        HIPCHK(hipStreamSynchronize(st));
        for (int s : pslots) free_slots.push_back(s);
        grid->replace_leaves(parents, add);
        tables_valid = false;
        return (int)parents.size();
    }

    int compress_ids(const std::unordered_set<int> &ids)
    {
        need_grid();
        std::vector<std::array<int, 4>> plan;
        {
            ScopeTimer tm(t_plan);
            plan = grid->plan_compress(ids);
        }
        if (plan.empty()) return 0;
        ScopeTimer tm(t_collapse);
        grow((int)(grid->leaves().size() + plan.size()));
        std::vector<int> pslots, cslots, remove;
        std::vector<gq::Leaf> add;
        for (const auto &p : plan) {
            gq::Leaf pl;
            pl.level = p[0];
            pl.idx[0] = p[1];
            pl.idx[1] = p[2];
            pl.idx[2] = p[3];
            pl.block_id = grid->new_block_id();
            pl.slot = alloc_slot();
            pslots.push_back(pl.slot);
            add.push_back(pl);
            for (int c = 0; c < 8; c++) {
                const int li = grid->find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1),
                                          2 * p[3] + ((c >> 2) & 1));
                if (li < 0) throw gq_error(GQ_ERR_TOPOLOGY, "compress plan references a missing child");
                remove.push_back(li);
                cslots.push_back(grid->leaves()[li].slot);
            }
        }
        d_list.upload(pslots, st);
        d_list2.upload(cslots, st);
        HIPCHK(gq::launch_collapse(phi[cur], pw, pg, d_list.p, d_list2.p, (int)pslots.size(), st));
        HIPCHK(hipStreamSynchronize(st));
        for (int s : cslots) free_slots.push_back(s);
        grid->replace_leaves(remove, add);
        tables_valid = false;
        return (int)plan.size();
    }

    static int zt_heuristic(int nwork)
    {
        // enough workgroups to fill 256 CUs several times over
        if (nwork >= 2048) return 16;
        if (nwork >= 1024) return 8;
        if (nwork >= 512) return 4;
        return 2;
    }

    // Step tile autotuner: time every slab depth on the current grid (the
    // outputs go to the scratch ping-pong buffer, the state is unchanged) and
    // keep the fastest per power-of-two bucket of the block count.
    int choose_zt(gq::StepArgs a, int nrows)
    {
        const int nwork = a.nwork;
        int bucket = 0;
        while ((1 << bucket) < nwork) bucket++;
        auto it = zt_cache.find(bucket);
        if (it != zt_cache.end()) return it->second;
        if (!autotune) return zt_heuristic(nwork);
        ScopeTimer tm(t_tune);
        static const int cand[4] = {16, 8, 4, 2};
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        a.phi_in = phi[cur];
        a.phi_out = phi[cur ^ 1];
        if (nrows > 0)
            HIPCHK(gq::launch_face_ghosts(phi[cur], d_face_ptr.p, d_face_src.p, d_face_w.p, d_gphi.p, nrows, st));
        int best = 16;
        float best_ms = 1e30f;
        for (int k = 0; k < 4; k++) {
            for (int r = 0; r < 2; r++) HIPCHK(gq::launch_fk_step(a, cand[k], st));
            HIPCHK(hipEventRecord(e0, st));
            const int reps = 5;
            for (int r = 0; r < reps; r++) HIPCHK(gq::launch_fk_step(a, cand[k], st));
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            tune_us[k] = 1e3 * ms / reps;
            if (ms < best_ms) {
                best_ms = ms;
                best = cand[k];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        zt_cache[bucket] = best;
        return best;
    }

    gq::UniformArgs uniform_args()
    {
        ensure_tables();
        if (grid->max_level() != grid->base_level())
            throw gq_error(GQ_ERR_UNSUPPORTED,
                           "uniform view needs max_level == log2(blocks_per_dim) (no leaf finer than the dump grid)");
        gq::UniformArgs u;
        u.vol = nullptr;
        u.phi = phi[cur];
        u.lut_slot = d_lut.p;
        u.level = d_level.p;
        u.index = d_index.p;
        u.gpd = grid->blocks_per_dim() * gq::BS;
        u.hole = 0.0;
        return u;
    }
};

namespace {

gq_status fail(gq_ctx *ctx, gq_status s, const std::string &m)
{
    if (ctx) ctx->err = m;
    return s;
}

template <typename F>
gq_status guard(gq_ctx *ctx, F &&f)
{
    if (!ctx) return GQ_ERR_ARG;
    try {
        ctx->err.clear();
        f();
        return GQ_OK;
    } catch (const gq_error &e) {
        return fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(ctx, GQ_ERR_NOMEM, "host allocation failed");
    } catch (const std::invalid_argument &e) {
        return fail(ctx, GQ_ERR_ARG, e.what());
    } catch (const std::runtime_error &e) {
        return fail(ctx, GQ_ERR_TOPOLOGY, e.what());
    } catch (const std::exception &e) {
        return fail(ctx, GQ_ERR_HIP, e.what());
    }
}

}  // namespace

extern "C" {

gq_status gq_ctx_create(gq_ctx **out, int device)
{
    if (!out) return GQ_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GQ_ERR_HIP;
    if (device < 0 || device >= ndev) return GQ_ERR_ARG;
    gq_ctx *c = new (std::nothrow) gq_ctx();
    if (!c) return GQ_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GQ_ERR_HIP;
    }
    *out = c;
    return GQ_OK;
}

gq_status gq_ctx_destroy(gq_ctx *ctx)
{
    if (!ctx) return GQ_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    delete ctx;
    return GQ_OK;
}

const char *gq_status_string(gq_status s)
{
    switch (s) {
    case GQ_OK: return "ok";
    case GQ_ERR_ARG: return "invalid argument";
    case GQ_ERR_HIP: return "HIP runtime error";
    case GQ_ERR_NOMEM: return "out of memory";
    case GQ_ERR_STATE: return "invalid state";
    case GQ_ERR_TOPOLOGY: return "grid topology error";
    case GQ_ERR_UNSUPPORTED: return "unsupported configuration";
    case GQ_ERR_NUMERIC: return "non-finite parameter";
    default: return "unknown status";
    }
}

const char *gq_last_error(const gq_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

gq_status gq_grid_create(gq_ctx *ctx, int bpd, int max_level, int order)
{
    return guard(ctx, [&] {
        HIPCHK(hipSetDevice(ctx->device));
        auto g = std::make_unique<gq::Grid>(bpd, max_level, order);
        ctx->tabs = gq::Tables();  // block ids restart: drop the row cache of any previous grid
        ctx->tables_valid = false;
        ctx->fik_valid = false;
        ctx->release_pool();
        ctx->free_slots.clear();
        ctx->grid = std::move(g);
        const int n = (int)ctx->grid->leaves().size();
        ctx->grow(std::max(64, n + n / 2));
        // slots in Morton order of the initial leaves
        std::vector<int> mo = gq::morton_order(*ctx->grid);
        std::vector<char> used(ctx->cap, 0);
        auto &L = ctx->grid->leaves_mut();
        for (int k = 0; k < n; k++) {
            L[mo[k]].slot = k;
            used[k] = 1;
        }
        ctx->free_slots.clear();
        for (int s = ctx->cap - 1; s >= 0; s--)
            if (!used[s]) ctx->free_slots.push_back(s);
        ctx->cur = 0;
        ctx->ensure_tables();
    });
}

gq_status gq_grid_block_count(gq_ctx *ctx, int *n)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        if (!n) throw gq_error(GQ_ERR_ARG, "null output");
        *n = (int)ctx->grid->leaves().size();
    });
}

gq_status gq_grid_blocks_info(gq_ctx *ctx, gq_block_info *out, int capacity, int *n)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        const auto &L = ctx->grid->leaves();
        const int m = (int)L.size();
        if (n) *n = m;
        if (!out) return;
        if (capacity < m) throw gq_error(GQ_ERR_ARG, "capacity smaller than the block count");
        for (int k = 0; k < m; k++) {
            const gq::Leaf &lf = L[ctx->api_order[k]];
            gq_block_info &b = out[k];
            b.block_id = lf.block_id;
            b.level = (int16_t)lf.level;
            for (int a = 0; a < 3; a++) {
                b.index[a] = (int16_t)lf.idx[a];
                b.origin[a] = (float)std::ldexp((double)lf.idx[a], -lf.level);
            }
            b.h = (float)(std::ldexp(1.0, -lf.level) / gq::BS);
        }
    });
}

static void transfer(gq_ctx *ctx, int field, float *host, int n_blocks, bool write)
{
    ctx->ensure_tables();
    const auto &L = ctx->grid->leaves();
    const int m = (int)L.size();
    if (n_blocks != m) throw gq_error(GQ_ERR_ARG, "n_blocks does not match the grid");
    if (!host) throw gq_error(GQ_ERR_ARG, "null buffer");
    float *dst = ctx->field_ptr(field);
    std::vector<int> slots(m);
    for (int k = 0; k < m; k++) slots[k] = L[ctx->api_order[k]].slot;
    ctx->d_list.upload(slots, ctx->st);
    ctx->d_dense.reserve((size_t)m * gq::BS3);
    const size_t bytes = (size_t)m * gq::BS3 * sizeof(float);
    if (write) {
        HIPCHK(hipMemcpyAsync(ctx->d_dense.p, host, bytes, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(gq::launch_scatter_blocks(dst, ctx->d_dense.p, ctx->d_list.p, m, ctx->st));
        if (field != GQ_FIELD_PHI) ctx->fik_valid = false;
    } else {
        HIPCHK(gq::launch_gather_blocks(dst, ctx->d_dense.p, ctx->d_list.p, m, ctx->st));
        HIPCHK(hipMemcpyAsync(host, ctx->d_dense.p, bytes, hipMemcpyDeviceToHost, ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
}

gq_status gq_grid_write_field(gq_ctx *ctx, int field, const float *data, int n_blocks)
{
    return guard(ctx, [&] { transfer(ctx, field, const_cast<float *>(data), n_blocks, true); });
}

gq_status gq_grid_read_field(gq_ctx *ctx, int field, float *data, int n_blocks)
{
    return guard(ctx, [&] { transfer(ctx, field, data, n_blocks, false); });
}

gq_status gq_fk_steps(gq_ctx *ctx, const gq_fk_params *p, int n_steps)
{
    return guard(ctx, [&] {
        if (!p || n_steps < 0) throw gq_error(GQ_ERR_ARG, "bad arguments");
        if (!std::isfinite(p->Dw) || !std::isfinite(p->Dg) || !std::isfinite(p->rho) || !std::isfinite(p->dt) ||
            p->dt <= 0)
            throw gq_error(GQ_ERR_NUMERIC, "non-finite parameter or dt <= 0");
        // the operator stores Dw, Dg, rho as Real (float)
        const float Dw = (float)p->Dw, Dg = (float)p->Dg, rho = (float)p->rho;
        ctx->ensure_fik(Dw, Dg);
        const int nwork = (int)ctx->api_order.size();
        const int nrows = ctx->face_nrows;
        gq::StepArgs a;
        a.pw = ctx->pw;
        a.pg = ctx->pg;
        a.work = ctx->d_work.p;
        a.nwork = nwork;
        a.fnbr = ctx->d_fnbr.p;
        a.fbase = ctx->d_fbase.p;
        a.fik = ctx->d_fik.p;
        a.gphi = ctx->d_gphi.p;
        a.level = ctx->d_level.p;
        a.ih2_lvl = ctx->d_ih2.p;
        a.Dw = Dw;
        a.Dg = Dg;
        a.rho = rho;
        a.dt = (float)p->dt;
        const int zt = ctx->choose_zt(a, nrows);
        ctx->last_zt = zt;
        const int chunk = 1024;
        const int need_ev = ctx->per_kernel_timing ? 2 * std::min(n_steps, chunk) : 2;
        while ((int)ctx->events.size() < need_ev) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->events.push_back(e);
        }
        double total = 0.0;
        int done = 0;
        while (done < n_steps) {
            const int m = std::min(chunk, n_steps - done);
            if (!ctx->per_kernel_timing) HIPCHK(hipEventRecord(ctx->events[0], ctx->st));
            for (int s = 0; s < m; s++) {
                if (nrows > 0)
                    HIPCHK(gq::launch_face_ghosts(ctx->phi[ctx->cur], ctx->d_face_ptr.p, ctx->d_face_src.p,
                                                  ctx->d_face_w.p, ctx->d_gphi.p, nrows, ctx->st));
                a.phi_in = ctx->phi[ctx->cur];
                a.phi_out = ctx->phi[ctx->cur ^ 1];
                if (ctx->per_kernel_timing) HIPCHK(hipEventRecord(ctx->events[2 * s], ctx->st));
                HIPCHK(gq::launch_fk_step(a, zt, ctx->st));
                if (ctx->per_kernel_timing) HIPCHK(hipEventRecord(ctx->events[2 * s + 1], ctx->st));
                ctx->cur ^= 1;
            }
            if (!ctx->per_kernel_timing) HIPCHK(hipEventRecord(ctx->events[1], ctx->st));
            HIPCHK(hipStreamSynchronize(ctx->st));
            if (ctx->per_kernel_timing) {
                for (int s = 0; s < m; s++) {
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, ctx->events[2 * s], ctx->events[2 * s + 1]));
                    total += ms;
                }
            } else {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
                total += ms;
            }
            done += m;
        }
        ctx->last_ms = total;
        ctx->last_launches = n_steps;
        ctx->last_cells = (int64_t)n_steps * nwork * gq::BS3;
    });
}

gq_status gq_fwt_details(gq_ctx *ctx, double *out, int n_blocks)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        const int m = (int)ctx->grid->leaves().size();
        if (n_blocks != m || !out) throw gq_error(GQ_ERR_ARG, "n_blocks does not match the grid");
        std::vector<double> d = ctx->fwt(ctx->api_order);
        std::copy(d.begin(), d.end(), out);
    });
}

gq_status gq_refine_blocks(gq_ctx *ctx, const int32_t *ids, int n, int *n_out)
{
    return guard(ctx, [&] {
        if (n < 0 || (n > 0 && !ids)) throw gq_error(GQ_ERR_ARG, "bad id list");
        std::unordered_set<int> s(ids, ids + n);
        const int r = ctx->refine_ids(s);
        if (n_out) *n_out = r;
    });
}

gq_status gq_compress_blocks(gq_ctx *ctx, const int32_t *ids, int n, int *n_out)
{
    return guard(ctx, [&] {
        if (n < 0 || (n > 0 && !ids)) throw gq_error(GQ_ERR_ARG, "bad id list");
        std::unordered_set<int> s(ids, ids + n);
        const int r = ctx->compress_ids(s);
        if (n_out) *n_out = r;
    });
}

gq_status gq_auto_refine(gq_ctx *ctx, double tol, int max_level, int max_loops, int *n_out)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        std::unordered_set<int> nice;
        int loops = 0, total = 0;
        while (true) {
            ctx->ensure_tables();
            const auto &L = ctx->grid->leaves();
            std::vector<int> cand;
            for (int li : ctx->api_order) {
                if (max_level >= 0 && max_level <= L[li].level) continue;
                if (!nice.count(L[li].block_id)) cand.push_back(li);
            }
            if (cand.empty()) break;
            std::vector<double> det = ctx->fwt(cand);
            std::unordered_set<int> flagged;
            for (size_t k = 0; k < cand.size(); k++) {
                if (det[k] >= tol) flagged.insert(L[cand[k]].block_id);
                else nice.insert(L[cand[k]].block_id);
            }
            if (flagged.empty()) break;
            total += ctx->refine_ids(flagged);
            loops++;
            if (max_loops > 0 && loops >= max_loops) break;
        }
        if (n_out) *n_out = total;
    });
}

gq_status gq_auto_compress(gq_ctx *ctx, double tol, int max_loops, int *n_out)
{
    return guard(ctx, [&] {
        std::unordered_set<int> nice;
        int loops = 0, total = 0;
        while (true) {
            ctx->ensure_tables();
            const auto &L = ctx->grid->leaves();
            std::vector<int> cand;
            for (int li : ctx->api_order)
                if (!nice.count(L[li].block_id)) cand.push_back(li);
            std::vector<double> det = ctx->fwt(cand);
            std::unordered_set<int> flagged;
            for (size_t k = 0; k < cand.size(); k++) {
                if (det[k] < tol) flagged.insert(L[cand[k]].block_id);
                else nice.insert(L[cand[k]].block_id);
            }
            if (flagged.empty()) break;
            const int n = ctx->compress_ids(flagged);
            if (n == 0) break;
            total += n;
            loops++;
            if (max_loops > 0 && loops >= max_loops) break;
        }
        if (n_out) *n_out = total;
    });
}

gq_status gq_dump_uniform(gq_ctx *ctx, double *out, size_t n_out)
{
    return guard(ctx, [&] {
        gq::UniformArgs u = ctx->uniform_args();
        const size_t n = (size_t)u.gpd * u.gpd * u.gpd;
        if (!out || n_out != n) throw gq_error(GQ_ERR_ARG, "output size must be gpd^3");
        ctx->d_dbl.reserve(n);
        HIPCHK(hipMemcpyAsync(ctx->d_dbl.p, out, n * sizeof(double), hipMemcpyHostToDevice, ctx->st));
        HIPCHK(gq::launch_dump_uniform(u, ctx->d_dbl.p, ctx->st));
        HIPCHK(hipMemcpyAsync(out, ctx->d_dbl.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
    });
}

gq_status gq_obs_flip_upload(gq_ctx *ctx, int gpd, const uint8_t *t1, const uint8_t *t2, const float *pet)
{
    return guard(ctx, [&] {
        if (gpd <= 0 || !t1 || !t2 || !pet) throw gq_error(GQ_ERR_ARG, "bad observation arrays");
        const size_t n = (size_t)gpd * gpd * gpd;
        ctx->d_t1u8.reserve(n);
        ctx->d_t2u8.reserve(n);
        ctx->d_petf.reserve(n);
        HIPCHK(hipMemcpyAsync(ctx->d_t1u8.p, t1, n, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->d_t2u8.p, t2, n, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->d_petf.p, pet, n * sizeof(float), hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        ctx->obs_flip_gpd = gpd;
    });
}

gq_status gq_loglik_flip(gq_ctx *ctx, const double params[4], double out[4])
{
    return guard(ctx, [&] {
        if (!params || !out) throw gq_error(GQ_ERR_ARG, "null argument");
        gq::UniformArgs u = ctx->uniform_args();
        if (ctx->obs_flip_gpd != u.gpd) throw gq_error(GQ_ERR_STATE, "observations missing or of another size");
        const double th1 = params[0], th2 = params[1], eps = params[2], sigma = params[3];
        if (!(eps > 0 && eps < 1) || !(sigma > 0)) throw gq_error(GQ_ERR_NUMERIC, "flip rate / sigma out of range");
        const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
        const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
        const double is2 = 1. / (2. * sigma * sigma);
        const int64_t n = (int64_t)u.gpd * u.gpd * u.gpd;
        const int nb = gq::reduce_blocks(n);
        ctx->d_partial.reserve((size_t)nb * 3 + 8);
        double *res = ctx->d_partial.p + (size_t)nb * 3;
        HIPCHK(gq::launch_lik_flip(u, ctx->d_t1u8.p, ctx->d_t2u8.p, ctx->d_petf.p, th1, th2, lp1, lp0, c0, is2,
                                   ctx->d_partial.p, nb, res, ctx->st));
        double h[3];
        HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        out[1] = h[0];
        out[2] = h[1];
        out[3] = h[2];
        out[0] = h[0] + h[1] + h[2];
    });
}

gq_status gq_obs_ref_upload(gq_ctx *ctx, int gpd, const double *pet, const double *t1, const double *t2,
                            const int32_t *pts, int npts)
{
    return guard(ctx, [&] {
        if (gpd <= 0 || !pet || !t1 || !t2 || npts < 0 || (npts > 0 && !pts))
            throw gq_error(GQ_ERR_ARG, "bad observation arrays");
        const size_t n = (size_t)gpd * gpd * gpd;
        for (int k = 0; k < 3 * npts; k++)
            if (pts[k] < 0 || pts[k] >= gpd) throw gq_error(GQ_ERR_ARG, "point outside the grid");
        ctx->d_pet_ref.reserve(n);
        ctx->d_t1_ref.reserve(n);
        ctx->d_t2_ref.reserve(n);
        ctx->d_pts.reserve((size_t)std::max(1, 3 * npts));
        HIPCHK(hipMemcpyAsync(ctx->d_pet_ref.p, pet, n * 8, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->d_t1_ref.p, t1, n * 8, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->d_t2_ref.p, t2, n * 8, hipMemcpyHostToDevice, ctx->st));
        if (npts > 0) HIPCHK(hipMemcpyAsync(ctx->d_pts.p, pts, (size_t)npts * 12, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        ctx->obs_ref_gpd = gpd;
        ctx->obs_ref_npts = npts;
    });
}

gq_status gq_loglik_ref(gq_ctx *ctx, const double params[5], int selective, int step_pet, int step_ti, double out[4])
{
    return guard(ctx, [&] {
        if (!params || !out || step_pet <= 0 || step_ti <= 0) throw gq_error(GQ_ERR_ARG, "bad arguments");
        gq::UniformArgs u = ctx->uniform_args();
        if (ctx->obs_ref_gpd != u.gpd) throw gq_error(GQ_ERR_STATE, "observations missing or of another size");
        const double sigma2 = params[0], scale = params[1], t1uc = params[2], t2uc = params[3], slope = params[4];
        const int64_t n = (int64_t)u.gpd * u.gpd * u.gpd;
        const int nb = gq::reduce_blocks(n);
        ctx->d_partial.reserve((size_t)nb * 4 + 8);
        double *res = ctx->d_partial.p + (size_t)nb * 4;
        HIPCHK(gq::launch_lik_ref(u, ctx->d_pet_ref.p, ctx->d_t1_ref.p, ctx->d_t2_ref.p, ctx->d_pts.p,
                                  ctx->obs_ref_npts, selective ? 1 : 0, step_pet, step_ti, scale, t1uc, t2uc, slope,
                                  ctx->d_partial.p, nb, res, ctx->st));
        double h[4];
        HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        // HGG_Likelihood.cpp:70-71
        const int N = (int)h[1];
        const double p1 = -0.5 * N * std::log(2. * M_PI * sigma2);
        const double p2 = -0.5 * (1. / sigma2) * h[0];
        out[1] = p1 + p2;
        out[2] = h[2];
        out[3] = h[3];
        out[0] = out[1] + out[2] + out[3];
    });
}

gq_status gq_loglik_ref_volume(gq_ctx *ctx, const double *model, int gpd, const double params[5], int selective,
                               int step_pet, int step_ti, double out[4])
{
    return guard(ctx, [&] {
        if (!model || gpd <= 0 || !params || !out || step_pet <= 0 || step_ti <= 0)
            throw gq_error(GQ_ERR_ARG, "bad arguments");
        if (ctx->obs_ref_gpd != gpd) throw gq_error(GQ_ERR_STATE, "observations missing or of another size");
        const size_t n = (size_t)gpd * gpd * gpd;
        ctx->d_dbl.reserve(n);
        HIPCHK(hipMemcpyAsync(ctx->d_dbl.p, model, n * sizeof(double), hipMemcpyHostToDevice, ctx->st));
        gq::UniformArgs u;
        u.vol = ctx->d_dbl.p;
        u.phi = nullptr;
        u.lut_slot = nullptr;
        u.level = nullptr;
        u.index = nullptr;
        u.gpd = gpd;
        u.hole = 0.0;
        const double sigma2 = params[0], scale = params[1], t1uc = params[2], t2uc = params[3], slope = params[4];
        const int nb = gq::reduce_blocks((int64_t)n);
        ctx->d_partial.reserve((size_t)nb * 4 + 8);
        double *res = ctx->d_partial.p + (size_t)nb * 4;
        HIPCHK(gq::launch_lik_ref(u, ctx->d_pet_ref.p, ctx->d_t1_ref.p, ctx->d_t2_ref.p, ctx->d_pts.p,
                                  ctx->obs_ref_npts, selective ? 1 : 0, step_pet, step_ti, scale, t1uc, t2uc, slope,
                                  ctx->d_partial.p, nb, res, ctx->st));
        double h[4];
        HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        const int N = (int)h[1];
        const double p1 = -0.5 * N * std::log(2. * M_PI * sigma2);
        const double p2 = -0.5 * (1. / sigma2) * h[0];
        out[1] = p1 + p2;
        out[2] = h[2];
        out[3] = h[3];
        out[0] = out[1] + out[2] + out[3];
    });
}

gq_status gq_last_step_timing(gq_ctx *ctx, double *ms, int *launches, int64_t *cells)
{
    return guard(ctx, [&] {
        if (ms) *ms = ctx->last_ms;
        if (launches) *launches = ctx->last_launches;
        if (cells) *cells = ctx->last_cells;
    });
}

gq_status gq_set_autotune(gq_ctx *ctx, int mode, int fixed_zt)
{
    return guard(ctx, [&] {
        if (mode == 2 && fixed_zt != 16 && fixed_zt != 8 && fixed_zt != 4 && fixed_zt != 2)
            throw gq_error(GQ_ERR_ARG, "slab depth must be 16, 8, 4 or 2");
        ctx->zt_cache.clear();
        ctx->autotune = mode == 1;
        if (mode == 2)
            for (int b = 0; b < 32; b++) ctx->zt_cache[b] = fixed_zt;
    });
}

gq_status gq_autotune_info(gq_ctx *ctx, int *zt, double us[4])
{
    return guard(ctx, [&] {
        if (zt) *zt = ctx->last_zt;
        if (us)
            for (int k = 0; k < 4; k++) us[k] = ctx->tune_us[k];
    });
}

gq_status gq_host_stats(gq_ctx *ctx, double out[8])
{
    return guard(ctx, [&] {
        if (!out) throw gq_error(GQ_ERR_ARG, "null output");
        out[0] = ctx->t_tables;
        out[1] = ctx->t_plan;
        out[2] = ctx->t_fwt;
        out[3] = ctx->t_split;
        out[4] = ctx->t_collapse;
        out[5] = ctx->t_tune;
        out[6] = ctx->n_tables;
        out[7] = ctx->n_fwt;
    });
}

gq_status gq_set_step_timing(gq_ctx *ctx, int per_kernel)
{
    return guard(ctx, [&] { ctx->per_kernel_timing = per_kernel != 0; });
}

gq_status gq_memory_info(gq_ctx *ctx, int64_t *bytes, int *pool_blocks)
{
    return guard(ctx, [&] {
        int64_t b = (int64_t)ctx->cap * gq::BS3 * 4 * 4;
        b += (int64_t)(ctx->d_ghost.n * 8);
        b += (int64_t)(ctx->d_face_w.n * 4 + ctx->d_face_src.n * 4 + ctx->d_face_ptr.n * 4 + ctx->d_fik.n * 4);
        if (bytes) *bytes = b;
        if (pool_blocks) *pool_blocks = ctx->cap;
    });
}

}  // extern "C"
