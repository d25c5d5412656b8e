// gq_api.cpp -- extern "C" layer: context, block pool, adaptation driver.
//
// Host C++ that owns the grid topology (grid.hpp) and the device state, and
// launches the HIP kernels (kernels.hpp) on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <initializer_list>
#include <chrono>
#include <climits>
#include <cstdlib>
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
#include "pool.hpp"

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
        // regrowth takes 1.5x the old capacity: a grid that grows step by step
        // reallocates (hipFree synchronises the device) a logarithmic number of times
        const size_t grow = n ? n + n / 2 : 0;
        release();
        size_t cap = std::max<size_t>(std::max(m, grow), 16);
        HIPCHK(hipMalloc((void **)&p, cap * sizeof(T)));
        n = cap;
    }
    // Stream-ordered copy: v must stay alive until the stream is synchronised.
    void upload(const std::vector<T> &v, hipStream_t st)
    {
        reserve(v.size());
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void release()
    {
        if (p) {
            // kernels or copies still queued may read the old buffer
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Page-locked host staging buffer (hipHostMalloc): copies from and to it are
// plain DMA transfers, with no staging through a driver buffer.  Growing it
// synchronises the stream first, so a transfer in flight never loses its source.
template <typename T>
struct PinnedBuf {
    T *p = nullptr;
    size_t n = 0;
    void reserve(size_t m, hipStream_t st)
    {
        if (m <= n) return;
        const size_t grow = n + n / 2;
        if (p) {
            HIPCHK(hipStreamSynchronize(st));
            (void)hipHostFree(p);
            p = nullptr;
            n = 0;
        }
        const size_t cap = std::max<size_t>(std::max(m, grow), 4096);
        HIPCHK(hipHostMalloc((void **)&p, cap * sizeof(T), hipHostMallocDefault));
        n = cap;
    }
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

// Host int arrays copied to the device in one transfer: the parts lie back to
// back (each at a 64-int boundary) in one device buffer, staged in pinned
// memory.  The caller's vectors may be released as soon as upload() returns:
// the transfer reads the staging buffer, and the next upload through the same
// object waits for the previous transfer (an event) before rewriting it.  The
// device buffer is rewritten in stream order, after the kernels queued on the
// same stream that read its previous contents.
struct PackedInts {
    DevBuf<int> dev;
    PinnedBuf<int> host;
    std::vector<size_t> off;
    hipEvent_t done = nullptr;
    bool pending = false;
    size_t total = 0;
    void upload(std::initializer_list<const std::vector<int> *> parts, hipStream_t st)
    {
        std::vector<size_t> sizes;
        for (const std::vector<int> *v : parts) sizes.push_back(v->size());
        stage(sizes, st);
        size_t k = 0;
        for (const std::vector<int> *v : parts) std::copy(v->begin(), v->end(), host.p + off[k++]);
        send(st);
    }
    // Parts of the given sizes laid out in the staging buffer (once the
    // previous transfer through this object has completed); the caller fills
    // them through host_ptr() and then calls send().
    void stage(const std::vector<size_t> &sizes, hipStream_t st)
    {
        total = 0;
        off.clear();
        for (size_t n : sizes) {
            off.push_back(total);
            total += (n + 63) & ~(size_t)63;
        }
        if (pending) HIPCHK(hipEventSynchronize(done));
        pending = false;
        host.reserve(std::max<size_t>(total, 1), st);
        dev.reserve(std::max<size_t>(total, 1));
    }
    int *host_ptr(int i) const { return host.p + off[i]; }
    void send(hipStream_t st)
    {
        if (total) {
            HIPCHK(hipMemcpyAsync(dev.p, host.p, total * sizeof(int), hipMemcpyHostToDevice, st));
            if (!done) HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
            HIPCHK(hipEventRecord(done, st));
            pending = true;
        }
    }
    int *ptr(int i) const { return dev.p + off[i]; }
    ~PackedInts()
    {
        if (done) (void)hipEventDestroy(done);
    }
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

// One grid of an ensemble with its host-side tables.
struct Sample {
    std::unique_ptr<gq::Grid> grid;
    bool dirty = true;  // topology changed since the tables were built
    gq::SampleTables t;
};

struct gq_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;

    // samples: independent grids of one configuration sharing the block pool
    std::vector<Sample> samples;
    int cur_sample = 0;              // the sample single-grid calls address
    std::vector<int> sample_order;   // work-list order of the samples

    // block pool
    int cap = 0;
    float *phi[2] = {nullptr, nullptr};
    int cur = 0;
    float *pw = nullptr, *pg = nullptr;
    std::vector<int> free_slots;

    // merged tables (slot-indexed arrays cover the whole pool)
    bool tables_valid = false;
    std::vector<int> h_fnbr, h_nbr26;  // [cap*6], [cap*26]
    std::vector<int> work;             // live slots: samples in sample_order, leaves in Morton order
    std::vector<int> work_begin;       // [S+1] per position in sample_order
    std::vector<int> face_begin;       // [S+1] cross-level faces per position
    DevBuf<int> d_level, d_index, d_fnbr, d_fbase, d_face_src, d_nbr26, d_lutoff, d_slot_sample, d_counts;
    // per-topology lists in one transfer: work list, sample face-row bases, base-level and finest look-ups,
    // cross-level faces, changed slot rows
    PackedInts tabs;
    int *p_work = nullptr, *p_sfbase = nullptr, *p_lut = nullptr, *p_lutf = nullptr, *p_faces = nullptr;
    int *p_wrec = nullptr;  // per-work-item step records (gq::WREC ints each), in the same transfer
    PackedInts lists;  // slot lists of one adaptation kernel sequence
    PackedInts flist;  // slots whose face 1/K is recomputed
    PinnedBuf<double> h_det;   // FWT details read back
    PinnedBuf<int64_t> h_total;  // face-row term count read back
    DevBuf<int64_t> d_face_ptr;
    DevBuf<char> d_scan_tmp;
    int face_nrows = 0;
    bool face_rows_valid = false;  // device face-ghost rows (ptr/src/w) match the tables
    int face_rows_npos = 0;        // ... for the samples at the first face_rows_npos positions
    DevBuf<double> d_ghost;  // cross-level lab points (3 fields for a split)
    DevBuf<float> d_face_w, d_fik, d_gphi, d_ih2, d_sprm;
    DevBuf<float> d_ikc;          // per-cell 1/K codes of the step (pool layout)
    bool fuse2 = false;           // fuse step pairs without cross-level faces into one launch (opt-in)
    std::vector<int> buf_after;   // advance(): ping-pong buffer holding the state after each step
    int last_nlaunch = 0;         // advance(): kernel launches of the last call
    bool step_coeffs = true;      // false: the step recomputes 1/K from pw, pg every step
    bool fik_valid = false;       // false: every slot's face 1/K must be recomputed
    std::vector<char> fik_stale;  // per sample: recompute that sample's slots
    std::vector<float> fik_key;   // Dw, Dg (per sample when per-sample parameters are set)

    // scratch
    DevBuf<int> d_list;
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
    bool autotune_all = false;  // also time the column kernel and the in-step face-ghost evaluation
    // step configuration: slab depth zt (16/8/4/2, 1 = column streaming) and
    // whether the step evaluates the cross-level face-ghost rows itself
    // (fused) or reads them from a separate k_face_ghosts launch
    // and whether a chunk of steps runs as one persistent cooperative launch
    static int cfg_zt(int cfg) { return cfg & 31; }
    static bool cfg_fused(int cfg) { return (cfg >> 5) & 1; }
    static bool cfg_persist(int cfg) { return (cfg >> 6) & 1; }
    static int make_cfg(int zt, bool fused, bool persist = false)
    {
        return zt | (fused ? 32 : 0) | (persist ? 64 : 0);
    }
    // persistent chunks: 0 never (default: measured slower, the grid-wide
    // barriers cost more than a launch), 1 an autotuner candidate, 2 always
    int persist_mode = 0;
    double tune_persist_us = 0.0;
    DevBuf<float> d_tune;  // second state buffer of the persistent candidate's timing
    std::map<int, int> zt_cache;  // work-list size bucket -> cfg
    int last_cfg = 16;
    double tune_us[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // host-side time accounting (seconds) and counts
    double t_tables = 0, t_plan = 0, t_fwt = 0, t_split = 0, t_collapse = 0, t_tune = 0;
    int n_tables = 0, n_fwt = 0;

    // timing
    bool per_kernel_timing = false;
    int timing_stride = 1;       // time every timing_stride-th step launch
    uint64_t timing_count = 0;   // step launches seen by the timing, across calls
    int last_timed = 0;          // timed launches of the last advance
    int64_t last_timed_cells = 0;  // cell updates of those launches
    // Step timing without waiting: advance records an event pair per timed
    // launch (or one around a chunk) and queues it; the pairs are read when a
    // caller asks for timings (or when many are queued), so a chunk of steps
    // never ends in a stream synchronisation of its own.
    struct TimedPair {
        hipEvent_t e0, e1;
        int64_t cells;  // cell updates the pair covers
        uint64_t call;  // gq_fk_steps / gq_ensemble_run call that recorded it
    };
    struct TimedSpan {
        float ms;
        int64_t cells;
        uint64_t call;
    };
    std::vector<hipEvent_t> event_free;
    std::vector<TimedPair> timing_pending;
    std::vector<TimedSpan> timing_done;
    uint64_t timing_call = 0;
    static constexpr size_t TIMING_PENDING_MAX = 4096, TIMING_DONE_MAX = 1 << 20;

    hipEvent_t take_event()
    {
        if (event_free.empty()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = event_free.back();
        event_free.pop_back();
        return e;
    }
    void queue_pair(hipEvent_t e0, hipEvent_t e1, int64_t cells)
    {
        if (timing_pending.size() >= TIMING_PENDING_MAX) drain_timing();
        timing_pending.push_back({e0, e1, cells, timing_call});
    }
    // Wait for the queued pairs (stream order: the last one ends last) and
    // move their durations to timing_done; the events go back to the pool.
    void drain_timing()
    {
        if (timing_pending.empty()) return;
        HIPCHK(hipEventSynchronize(timing_pending.back().e1));
        if (timing_done.size() + timing_pending.size() > TIMING_DONE_MAX)
            timing_done.erase(timing_done.begin(), timing_done.begin() + timing_done.size() / 2);
        for (const TimedPair &q : timing_pending) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, q.e0, q.e1));
            timing_done.push_back({ms, q.cells, q.call});
            event_free.push_back(q.e0);
            event_free.push_back(q.e1);
        }
        timing_pending.clear();
    }
    // device ms of call `call`'s timed spans (drains first)
    double call_ms(uint64_t call)
    {
        drain_timing();
        double t = 0.0;
        for (auto it = timing_done.rbegin(); it != timing_done.rend() && it->call >= call; ++it)
            if (it->call == call) t += it->ms;
        return t;
    }
    std::vector<hipEvent_t> events;
    uint64_t last_call = 0;  // timing call id of the last gq_fk_steps
    int last_launches = 0;
    int64_t last_cells = 0;

    // ensemble run statistics (gq_ensemble_run)
    double ens_stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // discrete adjoint: checkpoints, recomputed states, adjoint fields, accumulators
    DevBuf<float> d_ckpt, d_stash, d_lam;
    DevBuf<double> d_gacc, d_icp, d_icg;
    DevBuf<float> d_adj_ik;  // per-cell 1/K of the adjoint

    // Captured step chunks, a few launch shapes at a time.  The key holds
    // everything a capture records (launch arguments, buffers, per-step
    // prefixes, configuration, start buffer): an identical key relaunches the
    // instantiated graph without a new capture.  A shape is captured and
    // instantiated on its second sighting (the first runs as plain launches);
    // the least recently used entry is destroyed when the cache is full
    // (callers alternating between shapes -- odd and even step counts,
    // stage-2 batches of different sizes -- no longer capture on every call).
    // Instantiated graphs are never updated in place.
    struct CachedGraph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        std::vector<int> buf_after;
        int final_c = 0, nlaunch = 0;
        uint64_t used = 0;
    };
    static constexpr int GRAPH_CACHE = 8;
    bool fk_graphs = true;  // gq_fk_steps replays its chunks as graphs (unless every launch is timed)
    std::vector<CachedGraph> graphs;
    uint64_t graph_clock = 0;
    int64_t n_captures = 0, n_relaunches = 0;
    // Hashes of recently launched keys that have no graph: a key is captured
    // on its second sighting only, so a launch shape that occurs once (an
    // adaptive chunk's prefixes, a sampler batch of an unusual size) runs as
    // plain launches and never costs a capture and an instantiation.
    static constexpr int SEEN_KEYS = 64;
    // Instantiations per context are bounded: past this many captures every
    // new shape runs as plain launches (cached graphs are still relaunched),
    // so a long-running caller never instantiates and destroys graphs without
    // end.
    static constexpr int64_t MAX_CAPTURES = 4096;
    uint64_t seen_keys[SEEN_KEYS] = {};
    int seen_next = 0;
    bool key_seen(uint64_t h)
    {
        for (uint64_t k : seen_keys)
            if (k == h) return true;
        seen_keys[seen_next] = h;
        seen_next = (seen_next + 1) % SEEN_KEYS;
        return false;
    }

    void clear_graphs()
    {
        if (graphs.empty()) return;
        if (st) (void)hipStreamSynchronize(st);  // a graph may still be running
        for (auto &c : graphs)
            if (c.exec) (void)hipGraphExecDestroy(c.exec);
        graphs.clear();
    }

    ~gq_ctx()
    {
        if (st) (void)hipStreamSynchronize(st);
        clear_graphs();
        for (auto e : events) (void)hipEventDestroy(e);
        for (const auto &q : timing_pending) {
            (void)hipEventDestroy(q.e0);
            (void)hipEventDestroy(q.e1);
        }
        for (auto e : event_free) (void)hipEventDestroy(e);
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

    int nsamples() const { return (int)samples.size(); }

    void need_grid() const
    {
        if (samples.empty()) throw gq_error(GQ_ERR_STATE, "no grid: call gq_grid_create first");
    }

    gq::Grid &G(int s) { return *samples[s].grid; }
    gq::Grid &G() { return *samples[cur_sample].grid; }
    const std::vector<int> &api_order() { return samples[cur_sample].t.order; }

    // Grow the pool to at least `want` slots, keeping slot contents.
    void grow(int want)
    {
        if (want <= cap) return;
        int ncap = std::max(want, cap + cap / 2);
        if ((int64_t)ncap * gq::BS3 > INT32_MAX) {
            ncap = INT32_MAX / gq::BS3;
            if (want > ncap) throw gq_error(GQ_ERR_NOMEM, "block pool beyond 2^31 cells");
        }
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
        tables_valid = false;  // slot-indexed device arrays are sized by the pool
    }

    void reserve_free(size_t k)
    {
        if (free_slots.size() < k) grow(cap + (int)(k - free_slots.size()));
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

    // Host worker threads (created on first use) for the data-parallel loops
    // of an adaptation: GQ_HOST_THREADS threads in all, the caller included
    // (default: the hardware threads, at most 16).
    std::unique_ptr<gq::WorkerPool> wpool;
    gq::WorkerPool &pool()
    {
        if (!wpool) {
            int nt = hw_threads();
            if (const char *e = std::getenv("GQ_HOST_THREADS")) {
                const int v = std::atoi(e);
                if (v >= 1) nt = std::min(v, 64);
            }
            wpool = std::make_unique<gq::WorkerPool>(nt - 1);
        }
        return *wpool;
    }
    // Loops over fewer items than this stay on the calling thread: at the
    // size of one adaptive grid's leaves, handing chunks to other threads
    // costs more than the loop (measured: the tables of a 256^3-equivalent
    // solve take 8.3 ms serial and 11-14 ms on 2-16 threads).
    static constexpr int PAR_ITEMS = 4096;

    // Run f(i) for i in [0, n) on the worker threads.
    template <typename F>
    void parallel_for(int n, F &&f)
    {
        if (n <= 1) {
            for (int i = 0; i < n; i++) f(i);
            return;
        }
        pool().run(n, 1, [&](int b, int e) {
            for (int i = b; i < e; i++) f(i);
        });
    }

    // Slot-indexed tables on the host (persistent, sized by the pool) and the
    // pool size the device copies were last sized for.
    std::vector<int> h_level, h_ssample, h_lutoff, h_fbase;
    int uploaded_cap = 0;
    static constexpr int ROW = 45;  // packed row: slot, level, index[3], sample, lutoff, fnbr[6], fbase[6], nbr26[26]

    void ensure_tables()
    {
        need_grid();
        if (tables_valid) return;
        ScopeTimer tm(t_tables);
        n_tables++;
        const int S = nsamples();
        if ((int)sample_order.size() != S) {
            sample_order.resize(S);
            for (int s = 0; s < S; s++) sample_order[s] = s;
        }
        const bool full = uploaded_cap != cap;
        bool pageable = false;  // copies straight from pageable host vectors are queued below
        h_fnbr.resize((size_t)cap * 6, -1);
        h_nbr26.resize((size_t)cap * 26, -1);
        h_level.resize((size_t)cap, 0);
        h_ssample.resize((size_t)cap, 0);
        h_lutoff.resize((size_t)cap, 0);
        h_fbase.resize((size_t)cap * 6, -1);
        std::vector<int> dirty;
        for (int s = 0; s < S; s++)
            if (samples[s].dirty) dirty.push_back(s);
        fik_stale.resize(S, 1);
        for (int s : dirty) fik_stale[s] = 1;
        const int lut_n = (int)samples[0].grid->lut().size();
        // per-sample tables and slot rows of the samples whose topology changed
        parallel_for((int)dirty.size(), [&](int k) {
            const int s = dirty[k];
            Sample &sm = samples[s];
            gq::build_sample_tables(*sm.grid, h_fnbr.data(), h_nbr26.data(), sm.t);
            const auto &L = sm.grid->leaves();
            for (const auto &lf : L) {
                h_level[lf.slot] = lf.level;
                h_ssample[lf.slot] = s;
                h_lutoff[lf.slot] = s * lut_n;
                for (int f = 0; f < 6; f++) h_fbase[(size_t)lf.slot * 6 + f] = -1;
            }
            // face rows relative to the sample's first row
            for (size_t j = 0; j < sm.t.faces.size(); j++) h_fbase[sm.t.faces[j]] = (int)j * 256;
            sm.dirty = false;
        });
        // merged lists in work order
        std::vector<int> faces, sfbase(S, 0);
        work.clear();
        work_begin.assign(1, 0);
        face_begin.assign(1, 0);
        for (int pos = 0; pos < S; pos++) {
            const int s = sample_order[pos];
            const Sample &sm = samples[s];
            const auto &L = sm.grid->leaves();
            for (int li : sm.t.order) work.push_back(L[li].slot);
            sfbase[s] = (int)faces.size() * 256;
            faces.insert(faces.end(), sm.t.faces.begin(), sm.t.faces.end());
            work_begin.push_back((int)work.size());
            face_begin.push_back((int)faces.size());
        }
        if ((int64_t)faces.size() * 256 > INT32_MAX) throw gq_error(GQ_ERR_NOMEM, "too many cross-level faces");
        // slot rows scattered on the device: those of the changed samples, or
        // after a pool growth (device arrays regrown, contents dropped) every
        // live slot's row -- no kernel reads the row of a free slot
        std::vector<int> row_samples;
        if (full) {
            d_level.reserve((size_t)cap);
            d_index.reserve((size_t)cap * 3);
            d_slot_sample.reserve((size_t)cap);
            d_lutoff.reserve((size_t)cap);
            d_fnbr.reserve((size_t)cap * 6);
            d_fbase.reserve((size_t)cap * 6);
            d_nbr26.reserve((size_t)cap * 26);
            uploaded_cap = cap;
            for (int s = 0; s < S; s++) row_samples.push_back(s);
        } else {
            row_samples = dirty;
        }
        size_t nrow = 0;
        for (int s : row_samples) nrow += samples[s].grid->leaves().size();
        const size_t lslot_n = samples[0].t.lut_slot.size();
        // one transfer, filled in place in the page-locked staging buffer: work
        // list, sample face-row bases, base-level look-ups, finest look-ups,
        // cross-level faces, changed slot rows, step records
        tabs.stage({work.size(), (size_t)S, (size_t)S * lslot_n, (size_t)S * lut_n, faces.size(), nrow * ROW,
                    std::max<size_t>(1, work.size()) * gq::WREC},
                   st);
        std::copy(work.begin(), work.end(), tabs.host_ptr(0));
        std::copy(sfbase.begin(), sfbase.end(), tabs.host_ptr(1));
        for (int s = 0; s < S; s++) {
            const Sample &sm = samples[s];
            std::copy(sm.t.lut_slot.begin(), sm.t.lut_slot.end(), tabs.host_ptr(2) + (size_t)s * lslot_n);
            std::copy(sm.t.lut.begin(), sm.t.lut.end(), tabs.host_ptr(3) + (size_t)s * lut_n);
        }
        std::copy(faces.begin(), faces.end(), tabs.host_ptr(4));
        if (nrow) {
            // rows of those samples in order; samples on the worker threads
            // when there are enough rows to pay for them
            std::vector<size_t> roff(row_samples.size() + 1, 0);
            for (size_t k = 0; k < row_samples.size(); k++)
                roff[k + 1] = roff[k] + samples[row_samples[k]].grid->leaves().size();
            auto pack = [&](int k) {
                const int s = row_samples[k];
                int *q = tabs.host_ptr(5) + roff[k] * ROW;
                for (const gq::Leaf &lf : samples[s].grid->leaves()) {
                    const size_t sl = (size_t)lf.slot;
                    q[0] = lf.slot;
                    q[1] = lf.level;
                    q[2] = lf.idx[0];
                    q[3] = lf.idx[1];
                    q[4] = lf.idx[2];
                    q[5] = s;
                    q[6] = h_lutoff[sl];
                    std::memcpy(q + 7, &h_fnbr[sl * 6], 6 * sizeof(int));
                    std::memcpy(q + 13, &h_fbase[sl * 6], 6 * sizeof(int));
                    std::memcpy(q + 19, &h_nbr26[sl * 26], 26 * sizeof(int));
                    q += ROW;
                }
            };
            if (nrow >= PAR_ITEMS) parallel_for((int)row_samples.size(), pack);
            else
                for (int k = 0; k < (int)row_samples.size(); k++) pack(k);
        }
        // per-work-item step records: slot, 1/h^2 of its level (float bits),
        // face neighbours, absolute first face-ghost rows, sample
        static const std::vector<float> ih2 = [] {
            std::vector<float> v(16);
            for (int l = 0; l < 16; l++) {
                const float h = (float)(std::ldexp(1.0, -l) / gq::BS);  // BlockInfo::h (Real)
                const double hd = h;
                v[l] = (float)(1. / (hd * hd));
            }
            return v;
        }();
        auto records = [&](int b, int e) {
            for (int i = b; i < e; i++) {
                const int sl = work[i];
                int *r = tabs.host_ptr(6) + (size_t)i * gq::WREC;
                r[0] = sl;
                std::memcpy(&r[1], &ih2[h_level[sl]], sizeof(float));
                for (int f = 0; f < 6; f++) {
                    r[2 + f] = h_fnbr[(size_t)sl * 6 + f];
                    const int fb = h_fbase[(size_t)sl * 6 + f];
                    r[8 + f] = fb >= 0 ? sfbase[h_ssample[sl]] + fb : -1;
                }
                r[14] = h_ssample[sl];
                r[15] = 0;
            }
        };
        if ((int)work.size() >= PAR_ITEMS) pool().run((int)work.size(), PAR_ITEMS, records);
        else records(0, (int)work.size());
        tabs.send(st);
        p_work = tabs.ptr(0);
        p_sfbase = tabs.ptr(1);
        p_lut = tabs.ptr(2);
        p_lutf = tabs.ptr(3);
        p_faces = tabs.ptr(4);
        p_wrec = tabs.ptr(6);
        if (nrow) {
            gq::SlotTables t{d_level.p, d_index.p, d_slot_sample.p, d_lutoff.p, d_fnbr.p, d_fbase.p, d_nbr26.p};
            HIPCHK(gq::launch_scatter_slot_rows(t, tabs.ptr(5), (int)nrow, st));
        }
        face_nrows = (int)faces.size() * 256;
        face_rows_valid = false;
        if (d_fik.n < (size_t)cap * 6 * 256 || d_ikc.n < (size_t)cap * gq::BS3) {
            d_fik.reserve((size_t)cap * 6 * 256);  // contents not kept
            d_ikc.reserve((size_t)cap * gq::BS3);
            fik_valid = false;
        }
        if (!d_ih2.p) {
            d_ih2.upload(ih2, st);
            pageable = true;
        }
        // the packed lists went through pinned staging; copies from pageable vectors finish first
        if (pageable) HIPCHK(hipStreamSynchronize(st));
        tables_valid = true;
    }

    // Face-ghost rows of the step (ptr/src/w of every cross-level face-halo
    // cell), enumerated on the device.  Only stepping needs them, so a table
    // rebuild between a refinement and the compression that follows does not.
    //
    // npos: only the samples at the first npos positions of the work order
    // (the ones still stepping: a prefix) need rows; the faces of the others
    // get none until a later call asks for them.
    void ensure_face_rows(int npos = INT_MAX)
    {
        ensure_tables();
        const int S = nsamples();
        npos = std::max(0, std::min(npos, S));
        if (face_rows_valid && face_rows_npos >= npos) return;
        ScopeTimer tm(t_tables);
        const int nfaces = face_begin[npos];
        const int nrows = nfaces * 256;
        d_face_ptr.reserve((size_t)nrows + 1);
        if (nfaces > 0) {
            const gq::LabArgs la = lab_args();
            d_counts.reserve((size_t)nrows);
            HIPCHK(gq::launch_face_rows_count(la, p_faces, nfaces, d_counts.p, st));
            size_t tb = 0;
            HIPCHK(gq::scan_counts(d_counts.p, d_face_ptr.p, nrows, nullptr, tb, st));
            d_scan_tmp.reserve(tb + 1);
            HIPCHK(gq::scan_counts(d_counts.p, d_face_ptr.p, nrows, d_scan_tmp.p, tb, st));
            h_total.reserve(1, st);
            HIPCHK(hipMemcpyAsync(h_total.p, d_face_ptr.p + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            const int64_t total = *h_total.p;
            if (total > INT32_MAX) throw gq_error(GQ_ERR_NOMEM, "face rows exceed 2^31 terms");
            d_face_src.reserve((size_t)std::max<int64_t>(1, total));
            d_face_w.reserve((size_t)std::max<int64_t>(1, total));
            HIPCHK(gq::launch_face_rows_fill(la, p_faces, nfaces, d_face_ptr.p, d_face_src.p, d_face_w.p, st));
        }
        d_gphi.reserve((size_t)std::max(1, face_nrows));
        face_rows_valid = true;
        face_rows_npos = npos;
    }

    gq::LabArgs lab_args() const
    {
        gq::LabArgs a;
        a.nbr26 = d_nbr26.p;
        a.lut = p_lutf;
        a.lutoff = d_lutoff.p;
        a.level = d_level.p;
        a.index = d_index.p;
        a.lut_level = samples[0].grid->lut_level();
        a.order = samples[0].grid->order();
        return a;
    }

    // 1/K of the face halos for (Dw, Dg), or for the per-sample parameters in d_sprm (per_sample).  Only the
    // slots of samples whose topology changed are recomputed unless the parameters or the anatomy changed.
    // npos: as for ensure_face_rows -- the samples at later positions keep
    // their stale mark until a call covers them.
    void ensure_fik(const std::vector<float> &key, bool per_sample, int npos = INT_MAX)
    {
        ensure_face_rows(npos);
        const int S = nsamples();
        const int P = face_rows_npos;  // samples whose face rows exist
        if (!fik_valid || key != fik_key) {
            fik_stale.assign(S, 1);
            fik_valid = true;
            fik_key = key;
        }
        bool all = P == S, any = false;
        for (int pos = 0; pos < P; pos++) {
            all = all && fik_stale[sample_order[pos]];
            any = any || fik_stale[sample_order[pos]];
        }
        if (!any) return;
        const int *list = p_work;
        int n = (int)work.size();
        if (!all) {
            std::vector<int> sl;
            for (int pos = 0; pos < P; pos++)
                if (fik_stale[sample_order[pos]])
                    sl.insert(sl.end(), work.begin() + work_begin[pos], work.begin() + work_begin[pos + 1]);
            flist.upload({&sl}, st);
            list = flist.ptr(0);
            n = (int)sl.size();
        }
        HIPCHK(gq::launch_face_prep(pw, pg, list, n, d_fnbr.p, d_fbase.p, p_sfbase, d_slot_sample.p, d_face_ptr.p,
                                    d_face_src.p, d_face_w.p, d_fik.p, d_ikc.p, key[0], key[1],
                                    per_sample ? d_sprm.p : nullptr, st));
        for (int pos = 0; pos < P; pos++) fik_stale[sample_order[pos]] = 0;
    }

    // FWT detail of each listed slot (any samples)
    std::vector<double> fwt_slots(const std::vector<int> &slots)
    {
        ensure_tables();
        std::vector<double> out(slots.size());
        if (slots.empty()) return out;
        ScopeTimer tm(t_fwt);
        n_fwt++;
        const gq::LabArgs la = lab_args();
        const int order = samples[0].grid->order();
        d_ghost.reserve(slots.size() * (size_t)gq::lab_points(order));
        const float *fl[1] = {phi[cur]};
        std::vector<int> items;
        gq::lab_items(slots, h_nbr26, order, items);
        lists.upload({&slots, &items}, st);
        HIPCHK(gq::launch_lab_ghosts(la, fl, 1, lists.ptr(0), lists.ptr(1), (int)items.size(), d_ghost.p, 0, false,
                                     st));
        // the details go straight to page-locked host memory (no copy launch);
        // they are visible to the host once the stream has synchronised
        h_det.reserve(out.size(), st);
        HIPCHK(gq::launch_fwt_detail(order, la, d_ghost.p, phi[cur], lists.ptr(0), (int)slots.size(), h_det.p, st));
        HIPCHK(hipStreamSynchronize(st));
        std::copy(h_det.p, h_det.p + out.size(), out.begin());
        return out;
    }

    // The same-level neighbour rows of sample s (slot-indexed) when they
    // describe its current leaves: every topology edit marks the sample
    // dirty, and only its own new leaves take its freed slots.
    const int *nbr26_of(int s) const { return samples[s].dirty ? nullptr : h_nbr26.data(); }

    // Refine the flagged block ids of each listed sample (Refiner_SpaceExtension plan + split).
    // flags[k]: per block id of sample which[k], nflag[k] of them set.
    std::vector<int> refine_batch(const std::vector<int> &which, const std::vector<std::vector<char>> &flags,
                                  const std::vector<int> &nflag)
    {
        need_grid();
        const int n = (int)which.size();
        std::vector<std::vector<int>> parents(n);
        {
            ScopeTimer tm(t_plan);
            parallel_for(n, [&](int k) {
                if (nflag[k] > 0) parents[k] = G(which[k]).plan_refine(flags[k], nbr26_of(which[k]));
            });
        }
        size_t np = 0;
        for (auto &p : parents) np += p.size();
        std::vector<int> counts(n, 0);
        if (np == 0) return counts;
        ensure_tables();  // parent labs use the pre-refine topology
        ScopeTimer tm(t_split);
        reserve_free(np * 8);
        std::vector<int> pslots, cslots;
        std::vector<std::vector<gq::Leaf>> add(n);
        for (int k = 0; k < n; k++) {
            gq::Grid &g = G(which[k]);
            const auto &L = g.leaves();
            for (int p : parents[k]) {
                const gq::Leaf &pl = L[p];
                pslots.push_back(pl.slot);
                for (int c = 0; c < 8; c++) {
                    gq::Leaf ch;
                    ch.level = pl.level + 1;
                    ch.idx[0] = 2 * pl.idx[0] + (c & 1);
                    ch.idx[1] = 2 * pl.idx[1] + ((c >> 1) & 1);
                    ch.idx[2] = 2 * pl.idx[2] + ((c >> 2) & 1);
                    ch.block_id = g.new_block_id();
                    ch.slot = alloc_slot();
                    cslots.push_back(ch.slot);
                    add[k].push_back(ch);
                }
            }
        }
        const gq::LabArgs la = lab_args();
        const int order = samples[0].grid->order();
        const size_t per = pslots.size() * (size_t)gq::lab_points(order);
        d_ghost.reserve(3 * per);
        const float *fl[3] = {phi[cur], pw, pg};
        std::vector<int> items;
        gq::lab_items(pslots, h_nbr26, order, items);
        lists.upload({&pslots, &cslots, &items}, st);
        HIPCHK(gq::launch_lab_ghosts(la, fl, 3, lists.ptr(0), lists.ptr(2), (int)items.size(), d_ghost.p, per, true,
                                     st));
        HIPCHK(gq::launch_split(order, la, d_ghost.p, phi[cur], pw, pg, lists.ptr(0), lists.ptr(1), (int)pslots.size(),
                                st));
        // no wait: the host updates the topology while the split runs; later work on
        // the parents' slots (reused by later allocations) is queued behind it
        for (int s : pslots) free_slots.push_back(s);
        parallel_for(n, [&](int k) {
            if (!parents[k].empty()) G(which[k]).replace_leaves(parents[k], add[k]);
        });
        for (int k = 0; k < n; k++) {
            counts[k] = (int)parents[k].size();
            if (counts[k]) samples[which[k]].dirty = true;
        }
        tables_valid = false;
        return counts;
    }

    std::vector<int> compress_batch(const std::vector<int> &which, const std::vector<std::vector<char>> &flags,
                                    const std::vector<int> &nflag)
    {
        need_grid();
        const int n = (int)which.size();
        std::vector<std::vector<std::array<int, 4>>> plan(n);
        {
            ScopeTimer tm(t_plan);
            parallel_for(n, [&](int k) {
                if (nflag[k] > 0) plan[k] = G(which[k]).plan_compress(flags[k], nbr26_of(which[k]));
            });
        }
        size_t np = 0;
        for (auto &p : plan) np += p.size();
        std::vector<int> counts(n, 0);
        if (np == 0) return counts;
        ScopeTimer tm(t_collapse);
        reserve_free(np);
        std::vector<int> pslots, cslots;
        std::vector<std::vector<int>> remove(n);
        std::vector<std::vector<gq::Leaf>> add(n);
        for (int k = 0; k < n; k++) {
            gq::Grid &g = G(which[k]);
            for (const auto &p : plan[k]) {
                gq::Leaf pl;
                pl.level = p[0];
                pl.idx[0] = p[1];
                pl.idx[1] = p[2];
                pl.idx[2] = p[3];
                pl.block_id = g.new_block_id();
                pl.slot = alloc_slot();
                pslots.push_back(pl.slot);
                add[k].push_back(pl);
                for (int c = 0; c < 8; c++) {
                    const int li =
                        g.find(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1), 2 * p[3] + ((c >> 2) & 1));
                    if (li < 0) throw gq_error(GQ_ERR_TOPOLOGY, "compress plan references a missing child");
                    remove[k].push_back(li);
                    cslots.push_back(g.leaves()[li].slot);
                }
            }
        }
        lists.upload({&pslots, &cslots}, st);
        HIPCHK(gq::launch_collapse(phi[cur], pw, pg, lists.ptr(0), lists.ptr(1), (int)pslots.size(), st));
        for (int s : cslots) free_slots.push_back(s);
        parallel_for(n, [&](int k) {
            if (!plan[k].empty()) G(which[k]).replace_leaves(remove[k], add[k]);
        });
        for (int k = 0; k < n; k++) {
            counts[k] = (int)plan[k].size();
            if (counts[k]) samples[which[k]].dirty = true;
        }
        tables_valid = false;
        return counts;
    }

    // Science::AutomaticRefinement<0,0> (MRAGscience/MRAGScienceCore.h:42-108) over several samples at once.
    std::vector<int> auto_refine(const std::vector<int> &which, double tol, int max_level, int max_loops)
    {
        const int n = (int)which.size();
        std::vector<std::vector<char>> nice(n), flagged(n);  // per block id
        std::vector<int> nflag(n), total(n, 0);
        std::vector<char> live(n, 1);
        int loops = 0;
        while (true) {
            ensure_tables();
            std::vector<int> cand_slots, cand_k, cand_li;
            for (int k = 0; k < n; k++) {
                if (!live[k]) continue;
                const Sample &sm = samples[which[k]];
                const auto &L = sm.grid->leaves();
                nice[k].resize(sm.grid->id_bound(), 0);
                for (int li : sm.t.order) {
                    if (max_level >= 0 && max_level <= L[li].level) continue;
                    if (nice[k][L[li].block_id]) continue;
                    cand_slots.push_back(L[li].slot);
                    cand_k.push_back(k);
                    cand_li.push_back(li);
                }
            }
            if (cand_slots.empty()) break;
            std::vector<double> det = fwt_slots(cand_slots);
            for (int k = 0; k < n; k++) {
                flagged[k].assign(G(which[k]).id_bound(), 0);
                nflag[k] = 0;
            }
            for (size_t c = 0; c < cand_slots.size(); c++) {
                const int k = cand_k[c];
                const int id = G(which[k]).leaves()[cand_li[c]].block_id;
                if (det[c] >= tol) flagged[k][id] = 1, nflag[k]++;
                else nice[k][id] = 1;
            }
            bool any = false;
            for (int k = 0; k < n; k++) {
                if (nflag[k] == 0) live[k] = 0;
                else any = true;
            }
            if (!any) break;
            std::vector<int> r = refine_batch(which, flagged, nflag);
            for (int k = 0; k < n; k++) total[k] += r[k];
            loops++;
            if (max_loops > 0 && loops >= max_loops) break;
        }
        return total;
    }

    // Science::AutomaticCompression<0,0> (MRAGscience/MRAGScienceCore.h:117-177) over several samples at once.
    std::vector<int> auto_compress(const std::vector<int> &which, double tol, int max_loops)
    {
        const int n = (int)which.size();
        std::vector<std::vector<char>> nice(n), flagged(n);  // per block id
        std::vector<int> nflag(n), total(n, 0);
        std::vector<char> live(n, 1), cand;
        int loops = 0;
        while (true) {
            ensure_tables();
            std::vector<int> cand_slots, cand_k, cand_li;
            for (int k = 0; k < n; k++) {
                if (!live[k]) continue;
                const Sample &sm = samples[which[k]];
                const auto &L = sm.grid->leaves();
                nice[k].resize(sm.grid->id_bound(), 0);
                if (loops > 0) {
                    // Only candidates (leaves not yet found nice) can be
                    // flagged, so the parents this pass could collapse are
                    // among those the plan collapses with every candidate
                    // flagged.  None: the pass collapses nothing whatever the
                    // details are and the loop ends here, as it would after
                    // the transform (MRAGScienceCore.h:161-162); the transform
                    // is skipped.
                    cand.assign(nice[k].size(), 0);
                    for (const auto &lf : L) cand[lf.block_id] = !nice[k][lf.block_id];
                    ScopeTimer tm(t_plan);
                    if (sm.grid->plan_compress(cand, nbr26_of(which[k])).empty()) {
                        live[k] = 0;
                        continue;
                    }
                }
                for (int li : sm.t.order) {
                    if (nice[k][L[li].block_id]) continue;
                    cand_slots.push_back(L[li].slot);
                    cand_k.push_back(k);
                    cand_li.push_back(li);
                }
            }
            if (cand_slots.empty()) break;
            std::vector<double> det = fwt_slots(cand_slots);
            for (int k = 0; k < n; k++) {
                flagged[k].assign(G(which[k]).id_bound(), 0);
                nflag[k] = 0;
            }
            for (size_t c = 0; c < cand_slots.size(); c++) {
                const int k = cand_k[c];
                const int id = G(which[k]).leaves()[cand_li[c]].block_id;
                if (det[c] < tol) flagged[k][id] = 1, nflag[k]++;
                else nice[k][id] = 1;
            }
            std::vector<int> r = compress_batch(which, flagged, nflag);
            bool any = false;
            for (int k = 0; k < n; k++) {
                total[k] += r[k];
                if (nflag[k] == 0 || r[k] == 0) live[k] = 0;
                else any = true;
            }
            if (!any) break;
            loops++;
            if (max_loops > 0 && loops >= max_loops) break;
        }
        return total;
    }

    static int zt_heuristic(int nwork)
    {
        // enough workgroups to fill 256 CUs several times over
        if (nwork >= 2048) return 16;
        if (nwork >= 1024) return 8;
        if (nwork >= 512) return 4;
        return 2;
    }

    // Step autotuner: time every slab depth, with the face-ghost rows from a
    // separate launch (timed together with the step) and evaluated inside the
    // step, on the current grid (outputs go to the scratch ping-pong buffer,
    // the state is unchanged); keep the fastest per power-of-two bucket of the
    // block count.  Every configuration gives the same bits.
    int choose_step(gq::StepArgs a, int nrows)
    {
        const int nwork = a.nwork;
        int bucket = 0;
        while ((1 << bucket) < nwork) bucket++;
        auto it = zt_cache.find(bucket);
        if (it != zt_cache.end()) return with_persist(it->second);
        if (!autotune) return with_persist(make_cfg(zt_heuristic(nwork), false));
        ScopeTimer tm(t_tune);
        static const int cand[5] = {16, 8, 4, 2, 1};
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        a.phi_in = phi[cur];
        a.phi_out = phi[cur ^ 1];
        int best = make_cfg(16, false);
        float best_ms = 1e30f;
        for (int k = 0; k < 10; k++) tune_us[k] = 0.0;
        const int ncand = autotune_all ? 5 : 4;
        for (int fused = 0; fused < (nrows > 0 && autotune_all ? 2 : 1); fused++) {
            gq::StepArgs b = a;
            if (fused) b.gphi = nullptr;
            for (int k = 0; k < ncand; k++) {
                auto rep = [&] {
                    if (nrows > 0 && !fused)
                        HIPCHK(gq::launch_face_ghosts(phi[cur], d_face_ptr.p, d_face_src.p, d_face_w.p, d_gphi.p,
                                                      nrows, st));
                    HIPCHK(gq::launch_fk_step(b, cand[k], st));
                };
                for (int r = 0; r < 2; r++) rep();
                HIPCHK(hipEventRecord(e0, st));
                const int reps = 5;
                for (int r = 0; r < reps; r++) rep();
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                tune_us[5 * fused + k] = 1e3 * ms / reps;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = make_cfg(cand[k], fused);
                }
            }
        }
        // the persistent chunk at the fastest depth, timed over 4-step launches
        // on a copy of the state (phi[cur^1] and a scratch buffer) so the
        // live state is never overwritten
        tune_persist_us = 0.0;
        const int bzt = cfg_zt(best);
        if (persist_mode == 1 && bzt >= 2 && nwork > 0) {
            gq::PersistArgs p;
            p.a = a;
            if (cfg_fused(best)) p.a.gphi = nullptr;
            d_tune.reserve((size_t)cap * gq::BS3);
            HIPCHK(hipMemcpyAsync(phi[cur ^ 1], phi[cur], (size_t)cap * gq::BS3 * sizeof(float),
                                  hipMemcpyDeviceToDevice, st));
            p.phi[0] = phi[cur ^ 1];
            p.phi[1] = d_tune.p;
            p.gout = d_gphi.p;
            p.c0 = 0;
            p.m = 4;
            for (int j = 0; j < p.m; j++) {
                p.nwork[j] = nwork;
                p.nrows[j] = nrows;
            }
            const hipError_t e = gq::launch_fk_persistent(p, bzt, st);
            if (e == hipSuccess) {
                HIPCHK(hipEventRecord(e0, st));
                for (int r = 0; r < 2; r++) HIPCHK(gq::launch_fk_persistent(p, bzt, st));
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                tune_persist_us = 1e3 * ms / (2 * p.m);
                if (ms < best_ms * (2 * p.m) / 5) best = make_cfg(bzt, cfg_fused(best), true);
            } else {
                (void)hipGetLastError();  // not launchable cooperatively here: not a candidate
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        zt_cache[bucket] = best;
        return with_persist(best);
    }

    // the persistent-chunk mode applied to a cached or chosen configuration
    // (mode 1 keeps what the autotuner measured faster)
    int with_persist(int cfg) const
    {
        if (persist_mode == 0 || cfg_zt(cfg) < 2) return cfg & ~64;
        if (persist_mode == 2) return cfg | 64;
        return cfg;
    }

    gq::StepArgs step_args(const gq_fk_params *p)
    {
        gq::StepArgs a;
        a.pw = pw;
        a.pg = pg;
        a.work = p_work;
        a.nwork = (int)work.size();
        a.fnbr = d_fnbr.p;
        a.fik = d_fik.p;
        a.ikc = step_coeffs ? d_ikc.p : nullptr;
        a.wrec = p_wrec;
        a.nbr26 = d_nbr26.p;
        a.gphi = d_gphi.p;
        a.fptr = d_face_ptr.p;
        a.fsrc = d_face_src.p;
        a.fw = d_face_w.p;
        a.level = d_level.p;
        a.ih2_lvl = d_ih2.p;
        a.slot_sample = d_slot_sample.p;
        if (p) {  // the operator stores Dw, Dg, rho as Real (float)
            a.Dw = (float)p->Dw;
            a.Dg = (float)p->Dg;
            a.rho = (float)p->rho;
            a.dt = (float)p->dt;
            a.sprm = nullptr;
        } else {
            a.Dw = a.Dg = a.rho = a.dt = 0.f;
            a.sprm = d_sprm.p;
        }
        return a;
    }

    void need_events(int n)
    {
        while ((int)events.size() < n) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            events.push_back(e);
        }
    }

    // Advance m steps; step j updates the first nwork[j] work slots and the
    // first nrows[j] face rows (a prefix of the samples in work order).
    // Queues the device timing (call_ms): the step kernels only with
    // per-kernel timing, else the whole chunk.  Does not wait for the steps.
    void advance(gq::StepArgs a, int cfg, const std::vector<int> &nwork, const std::vector<int> &nrows,
                 bool use_graph)
    {
        const int zt = cfg_zt(cfg);
        const bool fused = cfg_fused(cfg);
        if (fused) a.gphi = nullptr;
        const int m = (int)nwork.size();
        last_nlaunch = 0;
        last_timed = 0;
        last_timed_cells = 0;
        if (m == 0) return;
        const bool pk = per_kernel_timing && !use_graph;
        int64_t chunk_cells = 0;
        for (int j = 0; j < m; j++) chunk_cells += (int64_t)nwork[j] * gq::BS3;
        // steps j, j+1 over the same work prefix without cross-level faces run as one fused two-step launch
        const bool can_fuse = fuse2 && a.ikc != nullptr;
        buf_after.assign(m, 0);
        int nlaunch = 0, final_c = cur, ntimed = 0;
        int64_t timed_cells = 0;
        const bool persist = cfg_persist(cfg) && zt >= 2;
        auto enqueue = [&](int c0) {
            int c = c0;
            nlaunch = 0;
            if (persist) {  // launches of up to PMAX steps
                for (int j0 = 0; j0 < m; j0 += gq::PMAX) {
                    gq::PersistArgs p;
                    p.a = a;
                    p.phi[0] = phi[c];
                    p.phi[1] = phi[c ^ 1];
                    p.gout = d_gphi.p;
                    p.c0 = 0;
                    p.m = std::min(gq::PMAX, m - j0);
                    for (int j = 0; j < p.m; j++) {
                        p.nwork[j] = nwork[j0 + j];
                        p.nrows[j] = fused ? 0 : nrows[j0 + j];
                        buf_after[j0 + j] = c ^ ((j + 1) & 1);
                    }
                    hipEvent_t e0 = pk ? take_event() : nullptr, e1 = pk ? take_event() : nullptr;
                    if (pk) HIPCHK(hipEventRecord(e0, st));
                    HIPCHK(gq::launch_fk_persistent(p, zt, st));
                    if (pk) {
                        HIPCHK(hipEventRecord(e1, st));
                        int64_t cl = 0;
                        for (int j = 0; j < p.m; j++) cl += (int64_t)p.nwork[j] * gq::BS3;
                        queue_pair(e0, e1, cl);
                        ntimed++;
                        timed_cells += cl;
                    }
                    nlaunch++;
                    c ^= p.m & 1;
                }
                final_c = c;
                return;
            }
            for (int j = 0; j < m;) {
                const bool pair = can_fuse && j + 1 < m && nwork[j] > 0 && nwork[j + 1] == nwork[j] &&
                                  nrows[j] == 0 && nrows[j + 1] == 0;
                if (!pair && nrows[j] > 0 && !fused)
                    HIPCHK(gq::launch_face_ghosts(phi[c], d_face_ptr.p, d_face_src.p, d_face_w.p, d_gphi.p, nrows[j],
                                                  st));
                a.phi_in = phi[c];
                a.phi_out = phi[c ^ 1];
                a.nwork = nwork[j];
                // per-kernel timing: every timing_stride-th launch carries its events
                const bool tj = pk && timing_count++ % (uint64_t)timing_stride == 0;
                hipEvent_t e0 = tj ? take_event() : nullptr, e1 = tj ? take_event() : nullptr;
                if (pair) HIPCHK(gq::launch_fk_step2(a, st, e0, e1));
                else if (nwork[j] > 0) HIPCHK(gq::launch_fk_step(a, zt, st, e0, e1));
                else if (tj) {
                    HIPCHK(hipEventRecord(e0, st));
                    HIPCHK(hipEventRecord(e1, st));
                }
                if (tj) {
                    const int64_t cl = (int64_t)nwork[j] * (pair ? 2 : 1) * gq::BS3;
                    queue_pair(e0, e1, cl);
                    ntimed++;
                    timed_cells += cl;
                }
                c ^= 1;
                nlaunch++;
                buf_after[j] = c;  // a pair leaves no state after its first step (no sample stops there)
                if (pair) buf_after[j + 1] = c;
                j += pair ? 2 : 1;
            }
            final_c = c;
        };
        std::vector<uint64_t> key;
        if (use_graph && !persist) {
            auto u = [&](uint64_t v) { key.push_back(v); };
            auto pt = [&](const void *q) { key.push_back((uint64_t)(uintptr_t)q); };
            auto fb = [&](float f) {
                uint32_t b;
                std::memcpy(&b, &f, 4);
                key.push_back(b);
            };
            pt(a.pw), pt(a.pg), pt(a.work), pt(a.fnbr), pt(a.fik), pt(a.ikc), pt(a.gphi), pt(a.fptr);
            pt(a.fsrc), pt(a.fw), pt(a.level), pt(a.ih2_lvl), pt(a.slot_sample), pt(a.sprm), pt(a.wrec), pt(a.nbr26);
            fb(a.Dw), fb(a.Dg), fb(a.rho), fb(a.dt);
            pt(phi[0]), pt(phi[1]), pt(d_face_ptr.p), pt(d_face_src.p), pt(d_face_w.p), pt(d_gphi.p);
            u((uint64_t)cfg), u((uint64_t)cur), u((uint64_t)can_fuse), u((uint64_t)m);
            for (int j = 0; j < m; j++) u(((uint64_t)(uint32_t)nwork[j] << 32) | (uint32_t)nrows[j]);
        }
        CachedGraph *hit = nullptr;
        bool capture = false;
        if (use_graph && !persist) {
            for (auto &c : graphs)
                if (c.key == key) hit = &c;
            if (!hit) {
                uint64_t h = 1469598103934665603ULL;  // FNV-1a over the key words
                for (uint64_t v : key) h = (h ^ v) * 1099511628211ULL;
                capture = key_seen(h | 1) && n_captures < MAX_CAPTURES;  // | 1: never the empty slot value 0
            }
        }
        if (hit) {
            // same launches as an instantiated graph: relaunch it
            hit->used = ++graph_clock;
            n_relaunches++;
            buf_after = hit->buf_after;
            final_c = hit->final_c;
            nlaunch = hit->nlaunch;
            hipEvent_t e0 = take_event(), e1 = take_event();
            HIPCHK(hipEventRecord(e0, st));
            HIPCHK(hipGraphLaunch(hit->exec, st));
            HIPCHK(hipEventRecord(e1, st));
            queue_pair(e0, e1, chunk_cells);
        } else if (capture) {  // a cooperative launch is not captured: one launch per chunk anyway
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                enqueue(cur);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(st, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);  // the instantiated graph does not refer to its source
            HIPCHK(ie);
            n_captures++;
            if ((int)graphs.size() >= GRAPH_CACHE) {
                auto lru = std::min_element(graphs.begin(), graphs.end(),
                                            [](const CachedGraph &x, const CachedGraph &y) { return x.used < y.used; });
                HIPCHK(hipStreamSynchronize(st));  // the evicted graph may still be running
                (void)hipGraphExecDestroy(lru->exec);
                graphs.erase(lru);
            }
            CachedGraph c;
            c.key = key;
            c.exec = ex;
            c.buf_after = buf_after;
            c.final_c = final_c;
            c.nlaunch = nlaunch;
            c.used = ++graph_clock;
            graphs.push_back(std::move(c));
            hipEvent_t e0 = take_event(), e1 = take_event();
            HIPCHK(hipEventRecord(e0, st));
            HIPCHK(hipGraphLaunch(ex, st));
            HIPCHK(hipEventRecord(e1, st));
            queue_pair(e0, e1, chunk_cells);
        } else {
            hipEvent_t e0 = pk ? nullptr : take_event(), e1 = pk ? nullptr : take_event();
            if (!pk) HIPCHK(hipEventRecord(e0, st));
            enqueue(cur);
            if (!pk) {
                HIPCHK(hipEventRecord(e1, st));
                queue_pair(e0, e1, chunk_cells);
            }
        }
        cur = final_c;
        last_nlaunch = nlaunch;
        last_timed = ntimed;
        last_timed_cells = timed_cells;
    }

    gq::UniformArgs uniform_args(int sample)
    {
        ensure_tables();
        const gq::Grid &g = G(sample);
        if (g.max_level() != g.base_level())
            throw gq_error(GQ_ERR_UNSUPPORTED,
                           "uniform view needs max_level == log2(blocks_per_dim) (no leaf finer than the dump grid)");
        const int bpd = g.blocks_per_dim();
        gq::UniformArgs u;
        u.vol = nullptr;
        u.phi = phi[cur];
        u.lut_slot = p_lut + (size_t)sample * bpd * bpd * bpd;
        u.level = d_level.p;
        u.index = d_index.p;
        u.gpd = bpd * gq::BS;
        u.hole = 0.0;
        u.obs_stride = 1;
        return u;
    }
    gq::UniformArgs uniform_args() { return uniform_args(cur_sample); }
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

static void create_samples(gq_ctx *ctx, int nsamples, int bpd, int max_level, int order, bool keep_pool = false)
{
    if (nsamples <= 0) throw gq_error(GQ_ERR_ARG, "nsamples must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<Sample> smp(nsamples);
    for (auto &sm : smp) sm.grid = std::make_unique<gq::Grid>(bpd, max_level, order);
    ctx->clear_graphs();
    ctx->tables_valid = false;
    ctx->fik_valid = false;
    ctx->h_fnbr.clear();
    ctx->h_nbr26.clear();
    ctx->h_level.clear();
    ctx->h_ssample.clear();
    ctx->h_lutoff.clear();
    ctx->h_fbase.clear();
    ctx->uploaded_cap = 0;
    const int n = (int)smp[0].grid->leaves().size();
    const int64_t want = (int64_t)n * nsamples;
    if (want * gq::BS3 > INT32_MAX) throw gq_error(GQ_ERR_NOMEM, "block pool beyond 2^31 cells");
    if (!keep_pool || ctx->cap < want) ctx->release_pool();
    ctx->free_slots.clear();
    ctx->samples = std::move(smp);
    ctx->cur_sample = 0;
    ctx->sample_order.clear();
    ctx->grow((int)std::max<int64_t>(64, want + want / 2));
    // slots in Morton order of the initial leaves, sample after sample
    std::vector<int> mo = gq::morton_order(*ctx->samples[0].grid);
    std::vector<char> used(ctx->cap, 0);
    for (int s = 0; s < nsamples; s++) {
        auto &L = ctx->samples[s].grid->leaves_mut();
        for (int k = 0; k < n; k++) {
            L[mo[k]].slot = s * n + k;
            used[s * n + k] = 1;
        }
    }
    ctx->free_slots.clear();
    for (int s = ctx->cap - 1; s >= 0; s--)
        if (!used[s]) ctx->free_slots.push_back(s);
    ctx->cur = 0;
    ctx->ensure_tables();
}

gq_status gq_grid_create(gq_ctx *ctx, int bpd, int max_level, int order)
{
    return guard(ctx, [&] { create_samples(ctx, 1, bpd, max_level, order); });
}

gq_status gq_ensemble_create(gq_ctx *ctx, int nsamples, int bpd, int max_level, int order)
{
    return guard(ctx, [&] { create_samples(ctx, nsamples, bpd, max_level, order); });
}

gq_status gq_ensemble_reset(gq_ctx *ctx)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        const gq::Grid &g = ctx->G(0);
        create_samples(ctx, ctx->nsamples(), g.blocks_per_dim(), g.max_level(), g.order(), true);
    });
}

gq_status gq_ensemble_size(gq_ctx *ctx, int *nsamples)
{
    return guard(ctx, [&] {
        if (!nsamples) throw gq_error(GQ_ERR_ARG, "null output");
        *nsamples = ctx->nsamples();
    });
}

gq_status gq_sample_select(gq_ctx *ctx, int sample)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        if (sample < 0 || sample >= ctx->nsamples()) throw gq_error(GQ_ERR_ARG, "sample out of range");
        ctx->cur_sample = sample;
    });
}

gq_status gq_ensemble_block_counts(gq_ctx *ctx, int32_t *out)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        if (!out) throw gq_error(GQ_ERR_ARG, "null output");
        for (int s = 0; s < ctx->nsamples(); s++) out[s] = (int32_t)ctx->G(s).leaves().size();
    });
}

gq_status gq_grid_block_count(gq_ctx *ctx, int *n)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        if (!n) throw gq_error(GQ_ERR_ARG, "null output");
        *n = (int)ctx->G().leaves().size();
    });
}

gq_status gq_grid_blocks_info(gq_ctx *ctx, gq_block_info *out, int capacity, int *n)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        const auto &L = ctx->G().leaves();
        const auto &order = ctx->api_order();
        const int m = (int)L.size();
        if (n) *n = m;
        if (!out) return;
        if (capacity < m) throw gq_error(GQ_ERR_ARG, "capacity smaller than the block count");
        for (int k = 0; k < m; k++) {
            const gq::Leaf &lf = L[order[k]];
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
    const auto &L = ctx->G().leaves();
    const auto &order = ctx->api_order();
    const int m = (int)L.size();
    if (n_blocks != m) throw gq_error(GQ_ERR_ARG, "n_blocks does not match the grid");
    if (!host) throw gq_error(GQ_ERR_ARG, "null buffer");
    float *dst = ctx->field_ptr(field);
    std::vector<int> slots(m);
    for (int k = 0; k < m; k++) slots[k] = L[order[k]].slot;
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

gq_status gq_grid_write_volume(gq_ctx *ctx, int field, const float *vol, int gpd, int sample)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        if (!vol) throw gq_error(GQ_ERR_ARG, "null volume");
        const gq::Grid &g0 = ctx->G(0);
        if (gpd != g0.blocks_per_dim() * gq::BS) throw gq_error(GQ_ERR_ARG, "gpd must be 16 * blocks_per_dim");
        if (sample < -1 || sample >= ctx->nsamples()) throw gq_error(GQ_ERR_ARG, "sample out of range");
        for (int s = 0; s < ctx->nsamples(); s++)
            for (const auto &lf : ctx->G(s).leaves())
                if (lf.level > g0.base_level())
                    throw gq_error(GQ_ERR_UNSUPPORTED, "a leaf is finer than the volume");
        float *dst = ctx->field_ptr(field);
        std::vector<int> slots;
        for (int s = 0; s < ctx->nsamples(); s++)
            if (sample < 0 || s == sample)
                for (const auto &lf : ctx->G(s).leaves()) slots.push_back(lf.slot);
        const size_t n = (size_t)gpd * gpd * gpd;
        ctx->d_dense.reserve(n);
        ctx->d_list.upload(slots, ctx->st);
        HIPCHK(hipMemcpyAsync(ctx->d_dense.p, vol, n * sizeof(float), hipMemcpyHostToDevice, ctx->st));
        HIPCHK(gq::launch_sample_volume(dst, ctx->d_dense.p, gpd, g0.base_level(), ctx->d_list.p, (int)slots.size(),
                                        ctx->d_level.p, ctx->d_index.p, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        if (field != GQ_FIELD_PHI) ctx->fik_valid = false;
    });
}

static void check_params(const gq_fk_params &p)
{
    if (!std::isfinite(p.Dw) || !std::isfinite(p.Dg) || !std::isfinite(p.rho) || !std::isfinite(p.dt) || p.dt <= 0)
        throw gq_error(GQ_ERR_NUMERIC, "non-finite parameter or dt <= 0");
}

gq_status gq_fk_steps(gq_ctx *ctx, const gq_fk_params *p, int n_steps)
{
    return guard(ctx, [&] {
        if (!p || n_steps < 0) throw gq_error(GQ_ERR_ARG, "bad arguments");
        check_params(*p);
        ctx->ensure_fik({(float)p->Dw, (float)p->Dg}, false);
        gq::StepArgs a = ctx->step_args(p);
        const int nwork = (int)ctx->work.size();
        const int nrows = ctx->face_nrows;
        const int cfg = ctx->choose_step(a, nrows);
        ctx->last_cfg = cfg;
        const int chunk = 1024;
        int done = 0, launches = 0;
        int64_t timed_cells = 0;
        const bool graph = ctx->fk_graphs && !ctx->per_kernel_timing;
        ctx->timing_call++;
        while (done < n_steps) {
            const int m = std::min(chunk, n_steps - done);
            ctx->advance(a, cfg, std::vector<int>(m, nwork), std::vector<int>(m, nrows), graph);
            launches += ctx->per_kernel_timing ? ctx->last_timed : ctx->last_nlaunch;
            timed_cells += ctx->last_timed_cells;
            done += m;
        }
        ctx->last_call = ctx->timing_call;  // its device time is read on request (gq_last_step_timing)
        ctx->last_launches = launches;
        ctx->last_cells = ctx->per_kernel_timing ? timed_cells : (int64_t)n_steps * nwork * gq::BS3;
    });
}

gq_status gq_ensemble_run(gq_ctx *ctx, const gq_fk_params *p, const int32_t *n_steps, const gq_run_opts *o)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        const int S = ctx->nsamples();
        if (!p || !n_steps || !o) throw gq_error(GQ_ERR_ARG, "null argument");
        if (o->adapt_every < 0) throw gq_error(GQ_ERR_ARG, "adapt_every must be >= 0");
        for (int s = 0; s < S; s++) {
            check_params(p[s]);
            if (n_steps[s] < 0) throw gq_error(GQ_ERR_ARG, "negative step count");
        }
        const bool use_graph = (o->flags & GQ_RUN_GRAPH) != 0;
        const bool adaptive = o->adapt_every > 0;
        // samples in decreasing step count: the active samples of any step are a prefix of the work list
        std::vector<int> order(S);
        for (int s = 0; s < S; s++) order[s] = s;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n_steps[x] > n_steps[y]; });
        if (order != ctx->sample_order) {
            ctx->sample_order = order;
            ctx->tables_valid = false;
        }
        std::vector<float> prm((size_t)S * 4), key((size_t)S * 2);
        for (int s = 0; s < S; s++) {
            prm[4 * s + 0] = (float)p[s].Dw;  // Real, as the operator stores them
            prm[4 * s + 1] = (float)p[s].Dg;
            prm[4 * s + 2] = (float)p[s].rho;
            prm[4 * s + 3] = (float)p[s].dt;
            key[2 * s] = prm[4 * s];
            key[2 * s + 1] = prm[4 * s + 1];
        }
        ctx->d_sprm.upload(prm, ctx->st);
        std::vector<int> rem(n_steps, n_steps + S);  // steps left per sample
        const int K = adaptive ? o->adapt_every : std::max(1, std::min(1024, *std::max_element(rem.begin(), rem.end())));
        double t_adapt = 0.0, cells = 0.0;
        int chunks = 0, adapts = 0;
        const uint64_t call = ++ctx->timing_call;
        while (true) {
            const int mx = *std::max_element(rem.begin(), rem.end());
            if (mx <= 0) break;
            const int m = std::min(K, mx);
            int nrun = 0;  // samples still stepping: a prefix of the work order
            while (nrun < S && rem[ctx->sample_order[nrun]] > 0) nrun++;
            ctx->ensure_fik(key, true, nrun);
            // per-step prefixes
            std::vector<int> nw(m), nr(m);
            for (int j = 0; j < m; j++) {
                int npos = 0;
                while (npos < S && rem[ctx->sample_order[npos]] > j) npos++;
                nw[j] = ctx->work_begin[npos];
                nr[j] = ctx->face_begin[npos] * 256;
                cells += (double)nw[j] * gq::BS3;
            }
            gq::StepArgs a = ctx->step_args(nullptr);
            a.nwork = nw[0];
            const int cfg = ctx->choose_step(a, nr[0]);
            ctx->last_cfg = cfg;
            const int c0 = ctx->cur;
            ctx->advance(a, cfg, nw, nr, use_graph);
            chunks++;
            // a sample that stopped early may hold its state in the other ping-pong buffer
            for (int pos = 0; pos < S; pos++) {
                const int s = ctx->sample_order[pos];
                const int ms_s = std::min(rem[s], m);
                const int src = ms_s > 0 ? ctx->buf_after[ms_s - 1] : c0;
                if (src != ctx->cur) {
                    const int b0 = ctx->work_begin[pos], b1 = ctx->work_begin[pos + 1];
                    HIPCHK(gq::launch_copy_blocks(ctx->phi[ctx->cur], ctx->phi[src], ctx->p_work + b0, b1 - b0,
                                                  ctx->st));
                }
            }
            std::vector<int> adapt;
            for (int s = 0; s < S; s++) {
                const int ms_s = std::min(rem[s], m);
                rem[s] -= ms_s;
                if (adaptive && ms_s == K) adapt.push_back(s);
            }
            if (!adapt.empty()) {
                const double t0 = now_s();
                ctx->auto_refine(adapt, o->tol_refine, o->max_level, 1);
                ctx->auto_compress(adapt, o->tol_compress, -1);
                t_adapt += now_s() - t0;
                adapts++;
            }
        }
        if (adaptive) {  // final refinement before the dump (Glioma/Glioma_HG_UQ.cpp:574-575)
            std::vector<int> all(S);
            for (int s = 0; s < S; s++) all[s] = s;
            const double t0 = now_s();
            ctx->auto_refine(all, o->tol_refine, o->max_level, 1);
            t_adapt += now_s() - t0;
        }
        HIPCHK(hipStreamSynchronize(ctx->st));
        ctx->ens_stats[0] = cells;
        ctx->ens_stats[1] = ctx->call_ms(call);
        ctx->ens_stats[2] = chunks;
        ctx->ens_stats[3] = adapts;
        ctx->ens_stats[4] = t_adapt;
        ctx->ens_stats[5] = (double)ctx->work.size();
        ctx->ens_stats[6] = gq_ctx::cfg_zt(ctx->last_cfg);
        ctx->ens_stats[7] = use_graph ? 1 : 0;
    });
}

// observations on the model's grid or on a coarser one whose size divides it
static int flip_stride(gq_ctx *ctx, int gpd)
{
    const int og = ctx->obs_flip_gpd;
    if (og <= 0 || gpd % og != 0) throw gq_error(GQ_ERR_STATE, "observations missing or of another size");
    return gpd / og;
}

// every sample a uniform grid (all leaves at the base level): the adjoint and the Gaussian IC map cells to voxels 1:1
static void need_uniform(gq_ctx *ctx)
{
    ctx->ensure_tables();
    for (int s = 0; s < ctx->nsamples(); s++) {
        const gq::Grid &g = ctx->G(s);
        for (const auto &lf : g.leaves())
            if (lf.level != g.base_level()) throw gq_error(GQ_ERR_UNSUPPORTED, "needs uniform grids (no adaptation)");
    }
}

gq_status gq_ensemble_ic_gaussian(gq_ctx *ctx, const double *ic)
{
    return guard(ctx, [&] {
        need_uniform(ctx);
        if (!ic) throw gq_error(GQ_ERR_ARG, "null argument");
        const int S = ctx->nsamples();
        for (int k = 0; k < 5 * S; k++)
            if (!std::isfinite(ic[k])) throw gq_error(GQ_ERR_NUMERIC, "non-finite initial-condition parameter");
        for (int s = 0; s < S; s++)
            if (!(ic[5 * s + 4] > 0)) throw gq_error(GQ_ERR_NUMERIC, "sigma must be positive");
        const std::vector<double> icv(ic, ic + 5 * S);  // kept until the synchronisation below
        ctx->d_icp.upload(icv, ctx->st);
        const int gpd = ctx->G(0).blocks_per_dim() * gq::BS;
        HIPCHK(gq::launch_ic_gaussian(ctx->phi[ctx->cur], ctx->p_work, (int)ctx->work.size(), ctx->d_index.p,
                                      ctx->d_slot_sample.p, ctx->d_icp.p, gpd, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
    });
}

gq_status gq_adjoint_flip(gq_ctx *ctx, const gq_fk_params *p, int n_steps, int ckpt_every, const double lik[4],
                          const double *ic, double *out_ll, double *out_grad)
{
    return guard(ctx, [&] {
        need_uniform(ctx);
        const int S = ctx->nsamples();
        if (!p || !lik || !out_ll || !out_grad || n_steps < 1 || ckpt_every < 1)
            throw gq_error(GQ_ERR_ARG, "bad arguments");
        for (int s = 0; s < S; s++) {
            check_params(p[s]);
            if (!(p[s].Dw > 0) || !(p[s].Dg > 0)) throw gq_error(GQ_ERR_NUMERIC, "the adjoint needs Dw, Dg > 0");
        }
        const int gpd = ctx->G(0).blocks_per_dim() * gq::BS;
        const int ostride = flip_stride(ctx, gpd);
        const double sigma = lik[3], eps = lik[2];
        if (!(eps > 0 && eps < 1) || !(sigma > 0)) throw gq_error(GQ_ERR_NUMERIC, "flip rate / sigma out of range");
        if (ic) {
            const gq_status r = gq_ensemble_ic_gaussian(ctx, ic);
            if (r != GQ_OK) throw gq_error(r, ctx->err);
        }
        // all samples step together: identity order
        std::vector<int> ident(S);
        for (int s = 0; s < S; s++) ident[s] = s;
        if (ctx->sample_order != ident) {
            ctx->sample_order = ident;
            ctx->tables_valid = false;
        }
        ctx->ensure_tables();
        if (ctx->face_nrows != 0) throw gq_error(GQ_ERR_UNSUPPORTED, "needs uniform grids (no adaptation)");
        std::vector<float> prm((size_t)S * 4), key((size_t)S * 2);
        for (int s = 0; s < S; s++) {
            prm[4 * s + 0] = (float)p[s].Dw;
            prm[4 * s + 1] = (float)p[s].Dg;
            prm[4 * s + 2] = (float)p[s].rho;
            prm[4 * s + 3] = (float)p[s].dt;
            key[2 * s] = prm[4 * s];
            key[2 * s + 1] = prm[4 * s + 1];
        }
        ctx->d_sprm.upload(prm, ctx->st);
        ctx->ensure_fik(key, true);
        const int nw = (int)ctx->work.size();
        const size_t pool = (size_t)ctx->cap * gq::BS3;
        const int E = std::min(ckpt_every, n_steps);
        const int nseg = (n_steps + E - 1) / E;
        ctx->d_ckpt.reserve((size_t)nseg * pool);
        ctx->d_stash.reserve((size_t)E * pool);
        ctx->d_lam.reserve(2 * pool);
        gq::StepArgs a = ctx->step_args(nullptr);
        const int zt = gq_ctx::cfg_zt(ctx->choose_step(a, 0));  // uniform grid: no face rows
        ctx->need_events(4);
        hipStream_t st = ctx->st;
        HIPCHK(hipEventRecord(ctx->events[0], st));
        // forward, keeping the state at the start of every segment
        const size_t pbytes = pool * sizeof(float);
        for (int seg = 0; seg < nseg; seg++) {
            HIPCHK(hipMemcpyAsync(ctx->d_ckpt.p + seg * pool, ctx->phi[ctx->cur], pbytes, hipMemcpyDeviceToDevice, st));
            const int m = std::min(E, n_steps - seg * E);
            const bool pairs = ctx->fuse2 && a.ikc != nullptr;  // uniform grid: fused two-step launches
            for (int j = 0; j < m;) {
                a.phi_in = ctx->phi[ctx->cur];
                a.phi_out = ctx->phi[ctx->cur ^ 1];
                if (pairs && j + 1 < m) {
                    HIPCHK(gq::launch_fk_step2(a, st));
                    j += 2;
                } else {
                    HIPCHK(gq::launch_fk_step(a, zt, st));
                    j++;
                }
                ctx->cur ^= 1;
            }
        }
        HIPCHK(hipEventRecord(ctx->events[1], st));
        // log-likelihood at u_N and its gradient field
        {
            gq::UniformArgs u = ctx->uniform_args(0);
            u.obs_stride = ostride;
            const double th1 = lik[0], th2 = lik[1];
            const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
            const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
            const double is2 = 1. / (2. * sigma * sigma);
            const int64_t og = gpd / ostride;
            const int nb = gq::flip_blocks(og * og * og);
            ctx->d_partial.reserve((size_t)S * nb * 3 + (size_t)S * 3);
            double *res = ctx->d_partial.p + (size_t)S * nb * 3;
            HIPCHK(gq::launch_lik_flip(u, S, ctx->d_t1u8.p, ctx->d_t2u8.p, ctx->d_petf.p, th1, th2, lp1, lp0, c0, is2,
                                       ctx->d_partial.p, nb, res, st));
            std::vector<double> h((size_t)S * 3);
            HIPCHK(hipMemcpyAsync(h.data(), res, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int s = 0; s < S; s++) {
                out_ll[4 * s + 1] = h[3 * s];
                out_ll[4 * s + 2] = h[3 * s + 1];
                out_ll[4 * s + 3] = h[3 * s + 2];
                out_ll[4 * s] = h[3 * s] + h[3 * s + 1] + h[3 * s + 2];
            }
        }
        float *lam[2] = {ctx->d_lam.p, ctx->d_lam.p + pool};
        int lc = 0;
        HIPCHK(gq::launch_lik_flip_grad(ctx->phi[ctx->cur], ctx->d_petf.p, ctx->p_work, nw, ctx->d_index.p, gpd,
                                        ostride, 1.0 / (sigma * sigma), lam[lc], st));
        ctx->d_adj_ik.reserve(pool);
        HIPCHK(gq::launch_adj_prep(ctx->pw, ctx->pg, ctx->p_work, nw, ctx->d_slot_sample.p, ctx->d_sprm.p,
                                   ctx->d_adj_ik.p, st));
        const int NS = gq::adj_slabs();
        ctx->d_gacc.reserve((size_t)ctx->cap * NS * 2);
        HIPCHK(hipMemsetAsync(ctx->d_gacc.p, 0, (size_t)ctx->cap * NS * 2 * sizeof(double), st));
        HIPCHK(hipEventRecord(ctx->events[2], st));
        gq::AdjArgs g;
        g.ik = ctx->d_adj_ik.p;
        g.work = ctx->p_work;
        g.nwork = nw;
        g.fnbr = ctx->d_fnbr.p;
        g.level = ctx->d_level.p;
        g.ih2_lvl = ctx->d_ih2.p;
        g.slot_sample = ctx->d_slot_sample.p;
        g.sprm = ctx->d_sprm.p;
        g.gacc = ctx->d_gacc.p;
        const float *u_end = ctx->phi[ctx->cur];  // u_N
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int m = std::min(E, n_steps - seg * E);
            // states u_{s0} .. u_{s0+m}: checkpoint, m-1 recomputed, and the next checkpoint (or u_N)
            auto state = [&](int j) -> const float * {
                if (j == 0) return ctx->d_ckpt.p + seg * pool;
                if (j == m) return seg + 1 < nseg ? ctx->d_ckpt.p + (seg + 1) * pool : u_end;
                return ctx->d_stash.p + (size_t)(j - 1) * pool;
            };
            for (int j = 0; j + 1 < m; j++) {
                a.phi_in = state(j);
                a.phi_out = ctx->d_stash.p + (size_t)j * pool;
                HIPCHK(gq::launch_fk_step(a, zt, st));
            }
            for (int j = m - 1; j >= 0; j--) {
                g.u = state(j);
                g.mu_in = lam[lc];
                g.out = lam[lc ^ 1];
                g.out_lambda = seg == 0 && j == 0;  // lambda_0 for the initial-condition gradient
                HIPCHK(gq::launch_adj_step(g, st));
                lc ^= 1;
            }
        }
        HIPCHK(hipEventRecord(ctx->events[3], st));
        // gradients: per-slab accumulators and the initial-condition reduction, summed per sample in work order
        std::vector<double> gacc((size_t)ctx->cap * NS * 2);
        HIPCHK(hipMemcpyAsync(gacc.data(), ctx->d_gacc.p, gacc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        std::vector<double> icg;
        if (ic) {
            ctx->d_icg.reserve((size_t)nw * 4);
            HIPCHK(gq::launch_ic_grad(lam[lc], ctx->p_work, nw, ctx->d_index.p, ctx->d_slot_sample.p, ctx->d_icp.p,
                                      gpd, ctx->d_icg.p, st));
            icg.resize((size_t)nw * 4);
            HIPCHK(hipMemcpyAsync(icg.data(), ctx->d_icg.p, icg.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        for (int s = 0; s < S; s++) {
            double gr[6] = {0, 0, 0, 0, 0, 0};
            for (int k = ctx->work_begin[s]; k < ctx->work_begin[s + 1]; k++) {
                const int slot = ctx->work[k];
                for (int sl = 0; sl < NS; sl++) {
                    gr[0] += gacc[((size_t)slot * NS + sl) * 2];
                    gr[1] += gacc[((size_t)slot * NS + sl) * 2 + 1];
                }
                if (ic)
                    for (int q = 0; q < 4; q++) gr[2 + q] += icg[(size_t)k * 4 + q];
            }
            for (int q = 0; q < 6; q++) out_grad[6 * s + q] = gr[q];
        }
        float f_ms = 0, b_ms = 0;
        HIPCHK(hipEventElapsedTime(&f_ms, ctx->events[0], ctx->events[1]));
        HIPCHK(hipEventElapsedTime(&b_ms, ctx->events[2], ctx->events[3]));
        ctx->ens_stats[0] = (double)n_steps * nw * gq::BS3;
        ctx->ens_stats[1] = f_ms;
        ctx->ens_stats[2] = b_ms;
        ctx->ens_stats[3] = nseg;
        ctx->ens_stats[4] = 0;
        ctx->ens_stats[5] = nw;
        ctx->ens_stats[6] = zt;
        ctx->ens_stats[7] = 0;
    });
}

gq_status gq_ensemble_stats(gq_ctx *ctx, double out[8])
{
    return guard(ctx, [&] {
        if (!out) throw gq_error(GQ_ERR_ARG, "null output");
        for (int k = 0; k < 8; k++) out[k] = ctx->ens_stats[k];
    });
}

gq_status gq_fwt_details(gq_ctx *ctx, double *out, int n_blocks)
{
    return guard(ctx, [&] {
        ctx->ensure_tables();
        const auto &L = ctx->G().leaves();
        const int m = (int)L.size();
        if (n_blocks != m || !out) throw gq_error(GQ_ERR_ARG, "n_blocks does not match the grid");
        std::vector<int> slots;
        for (int li : ctx->api_order()) slots.push_back(L[li].slot);
        std::vector<double> d = ctx->fwt_slots(slots);
        std::copy(d.begin(), d.end(), out);
    });
}

gq_status gq_refine_blocks(gq_ctx *ctx, const int32_t *ids, int n, int *n_out)
{
    return guard(ctx, [&] {
        if (n < 0 || (n > 0 && !ids)) throw gq_error(GQ_ERR_ARG, "bad id list");
        const gq::Grid &g = ctx->G();
        std::vector<std::vector<char>> f(1, g.id_flags(std::unordered_set<int>(ids, ids + n)));
        int nf = 0;
        for (char c : f[0]) nf += c;
        const std::vector<int> r = ctx->refine_batch({ctx->cur_sample}, f, {nf});
        if (n_out) *n_out = r[0];
    });
}

gq_status gq_compress_blocks(gq_ctx *ctx, const int32_t *ids, int n, int *n_out)
{
    return guard(ctx, [&] {
        if (n < 0 || (n > 0 && !ids)) throw gq_error(GQ_ERR_ARG, "bad id list");
        const gq::Grid &g = ctx->G();
        std::vector<std::vector<char>> f(1, g.id_flags(std::unordered_set<int>(ids, ids + n)));
        int nf = 0;
        for (char c : f[0]) nf += c;
        const std::vector<int> r = ctx->compress_batch({ctx->cur_sample}, f, {nf});
        if (n_out) *n_out = r[0];
    });
}

gq_status gq_auto_refine(gq_ctx *ctx, double tol, int max_level, int max_loops, int *n_out)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        const std::vector<int> r = ctx->auto_refine({ctx->cur_sample}, tol, max_level, max_loops);
        if (n_out) *n_out = r[0];
    });
}

gq_status gq_auto_compress(gq_ctx *ctx, double tol, int max_loops, int *n_out)
{
    return guard(ctx, [&] {
        ctx->need_grid();
        const std::vector<int> r = ctx->auto_compress({ctx->cur_sample}, tol, max_loops);
        if (n_out) *n_out = r[0];
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
        u.obs_stride = flip_stride(ctx, u.gpd);
        const double th1 = params[0], th2 = params[1], eps = params[2], sigma = params[3];
        if (!(eps > 0 && eps < 1) || !(sigma > 0)) throw gq_error(GQ_ERR_NUMERIC, "flip rate / sigma out of range");
        const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
        const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
        const double is2 = 1. / (2. * sigma * sigma);
        const int64_t og = u.gpd / u.obs_stride;
        const int nb = gq::flip_blocks(og * og * og);
        ctx->d_partial.reserve((size_t)nb * 3 + 8);
        double *res = ctx->d_partial.p + (size_t)nb * 3;
        HIPCHK(gq::launch_lik_flip(u, 1, ctx->d_t1u8.p, ctx->d_t2u8.p, ctx->d_petf.p, th1, th2, lp1, lp0, c0, is2,
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

gq_status gq_ensemble_loglik_flip(gq_ctx *ctx, const double params[4], double *out)
{
    return guard(ctx, [&] {
        if (!params || !out) throw gq_error(GQ_ERR_ARG, "null argument");
        const int S = ctx->nsamples();
        for (int s = 1; s < S; s++) (void)ctx->uniform_args(s);  // every sample must map to the uniform view
        gq::UniformArgs u = ctx->uniform_args(0);
        u.obs_stride = flip_stride(ctx, u.gpd);
        const double th1 = params[0], th2 = params[1], eps = params[2], sigma = params[3];
        if (!(eps > 0 && eps < 1) || !(sigma > 0)) throw gq_error(GQ_ERR_NUMERIC, "flip rate / sigma out of range");
        const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
        const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
        const double is2 = 1. / (2. * sigma * sigma);
        const int64_t og = u.gpd / u.obs_stride;
        const int nb = gq::flip_blocks(og * og * og);
        ctx->d_partial.reserve((size_t)S * nb * 3 + (size_t)S * 3);
        double *res = ctx->d_partial.p + (size_t)S * nb * 3;
        HIPCHK(gq::launch_lik_flip(u, S, ctx->d_t1u8.p, ctx->d_t2u8.p, ctx->d_petf.p, th1, th2, lp1, lp0, c0, is2,
                                   ctx->d_partial.p, nb, res, ctx->st));
        std::vector<double> h((size_t)S * 3);
        HIPCHK(hipMemcpyAsync(h.data(), res, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        for (int s = 0; s < S; s++) {
            out[4 * s + 1] = h[3 * s];
            out[4 * s + 2] = h[3 * s + 1];
            out[4 * s + 3] = h[3 * s + 2];
            out[4 * s] = h[3 * s] + h[3 * s + 1] + h[3 * s + 2];
        }
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
        u.obs_stride = 1;
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

gq_status gq_step_timing_drain(gq_ctx *ctx, float *ms, int64_t *cells, int cap, int *n)
{
    return guard(ctx, [&] {
        if (!n) throw gq_error(GQ_ERR_ARG, "null count");
        ctx->drain_timing();
        const int cnt = (int)ctx->timing_done.size();
        if (!ms) {
            *n = cnt;
            return;
        }
        if (cap < cnt) throw gq_error(GQ_ERR_ARG, "capacity below the number of timed entries");
        for (int i = 0; i < cnt; i++) {
            ms[i] = ctx->timing_done[i].ms;
            if (cells) cells[i] = ctx->timing_done[i].cells;
        }
        ctx->timing_done.clear();
        *n = cnt;
    });
}

gq_status gq_last_step_timing(gq_ctx *ctx, double *ms, int *launches, int64_t *cells)
{
    return guard(ctx, [&] {
        if (ms) *ms = ctx->last_call ? ctx->call_ms(ctx->last_call) : 0.0;
        if (launches) *launches = ctx->last_launches;
        if (cells) *cells = ctx->last_cells;
    });
}

gq_status gq_set_autotune(gq_ctx *ctx, int mode, int fixed_zt)
{
    return guard(ctx, [&] {
        if (mode < 0 || mode > 4) throw gq_error(GQ_ERR_ARG, "autotune mode must be 0, 1, 2, 3 or 4");
        if ((mode == 2 || mode == 3) && fixed_zt != 16 && fixed_zt != 8 && fixed_zt != 4 && fixed_zt != 2 && fixed_zt != 1)
            throw gq_error(GQ_ERR_ARG, "slab depth must be 16, 8, 4, 2 or 1 (column streaming)");
        ctx->zt_cache.clear();
        ctx->autotune = mode == 1 || mode == 4;
        ctx->autotune_all = mode == 4;
        if (mode == 2 || mode == 3)
            for (int b = 0; b < 32; b++) ctx->zt_cache[b] = gq_ctx::make_cfg(fixed_zt, mode == 3);
    });
}

gq_status gq_set_step_fusion(gq_ctx *ctx, int on)
{
    return guard(ctx, [&] {
        if (on != 0 && on != 1) throw gq_error(GQ_ERR_ARG, "on must be 0 or 1");
        ctx->fuse2 = on == 1;
    });
}

gq_status gq_set_step_coeffs(gq_ctx *ctx, int precomputed)
{
    return guard(ctx, [&] {
        if (precomputed != 0 && precomputed != 1) throw gq_error(GQ_ERR_ARG, "precomputed must be 0 or 1");
        ctx->step_coeffs = precomputed == 1;
    });
}

gq_status gq_autotune_info(gq_ctx *ctx, int *zt, int *fused, double us[10])
{
    return guard(ctx, [&] {
        if (zt) *zt = gq_ctx::cfg_zt(ctx->last_cfg);
        if (fused) *fused = gq_ctx::cfg_fused(ctx->last_cfg) ? 1 : 0;
        if (us)
            for (int k = 0; k < 10; k++) us[k] = ctx->tune_us[k];
    });
}

gq_status gq_set_step_persistent(gq_ctx *ctx, int mode)
{
    return guard(ctx, [&] {
        if (mode < 0 || mode > 2) throw gq_error(GQ_ERR_ARG, "persistent mode must be 0, 1 or 2");
        ctx->persist_mode = mode;
        if (mode == 1 && ctx->autotune) ctx->zt_cache.clear();  // time the candidate per bucket
    });
}

gq_status gq_persistent_info(gq_ctx *ctx, int *used, double *tuned_us)
{
    return guard(ctx, [&] {
        if (used) *used = gq_ctx::cfg_persist(ctx->last_cfg) ? 1 : 0;
        if (tuned_us) *tuned_us = ctx->tune_persist_us;
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

gq_status gq_graph_info(gq_ctx *ctx, int64_t *captures, int64_t *relaunches, int *cached)
{
    return guard(ctx, [&] {
        if (captures) *captures = ctx->n_captures;
        if (relaunches) *relaunches = ctx->n_relaunches;
        if (cached) *cached = (int)ctx->graphs.size();
    });
}

gq_status gq_set_step_graphs(gq_ctx *ctx, int on)
{
    return guard(ctx, [&] { ctx->fk_graphs = on != 0; });
}

gq_status gq_set_step_timing(gq_ctx *ctx, int per_kernel)
{
    return guard(ctx, [&] {
        if (per_kernel < 0) throw gq_error(GQ_ERR_ARG, "per_kernel must be >= 0");
        ctx->per_kernel_timing = per_kernel != 0;
        ctx->timing_stride = std::max(1, per_kernel);
        ctx->timing_count = 0;
    });
}

gq_status gq_memory_info(gq_ctx *ctx, int64_t *bytes, int *pool_blocks)
{
    return guard(ctx, [&] {
        int64_t b = (int64_t)ctx->cap * gq::BS3 * 4 * 4;
        b += (int64_t)(ctx->d_ghost.n * 8);
        b += (int64_t)(ctx->d_face_w.n * 4 + ctx->d_face_src.n * 4 + ctx->d_face_ptr.n * 8 + ctx->d_fik.n * 4 +
                       ctx->d_ikc.n * 4);
        if (bytes) *bytes = b;
        if (pool_blocks) *pool_blocks = ctx->cap;
    });
}

}  // extern "C"
