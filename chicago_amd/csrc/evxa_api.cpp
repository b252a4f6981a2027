// evxa_api.cpp -- host runtime and C ABI of libevxa.so (see include/evxa.h).
//
// Pictures, record tables, coefficients and levels stay resident in HBM; the
// host enqueues frame jobs on the context's HIP stream.  Every analysis
// launch is: a reset of the work-list counters and MB predictor words (a fill
// kernel, so the sequence can be captured into a hipGraph), then one
// persistent analysis kernel over the MB wavefronts of all jobs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <set>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "evxa_internal.h"

using namespace evxa;

namespace {

thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw Failure{code, msg}; }

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        fail(EVXA_ERROR_EXECUTION_FAILURE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f)
{
    try {
        f();
        g_last_error.clear();
        return EVXA_SUCCESS;
    } catch (const Failure &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return EVXA_ERROR_OUTOFMEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return EVXA_ERROR_EXECUTION_FAILURE;
    }
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Plane geometry with apron; pitch rounded to 64 elements (128 bytes).
struct PlaneGeom {
    int32_t w, h, pitch, rows;
    size_t bytes() const { return (size_t)pitch * rows * sizeof(int16_t); }
    size_t origin() const { return (size_t)kApron * pitch + kApron; }
};

PlaneGeom plane_geom(int32_t w, int32_t h)
{
    PlaneGeom g;
    g.w = w;
    g.h = h;
    g.pitch = (int32_t)align_up((size_t)w + 2 * kApron, 64);
    g.rows = h + 2 * kApron;
    return g;
}

size_t picture_bytes(int32_t w, int32_t h)
{
    return plane_geom(w, h).bytes() + 2 * plane_geom(w / 2, h / 2).bytes();
}

size_t outputs_bytes(int32_t w, int32_t h)
{
    const size_t n = (size_t)(w / 16) * (h / 16);
    return 3 * n * sizeof(evxa_block_record) + 2 * n * EVXA_MB_SAMPLES * sizeof(int16_t) +
           n * EVXA_MAX_REFS * sizeof(SearchRec);
}

// Device buffer owned by RAII.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    void alloc(size_t bytes)
    {
        release();
        if (bytes == 0) return;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            fail(EVXA_ERROR_OUTOFMEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        n = bytes;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

struct PinnedBuf {
    void *p = nullptr;
    size_t n = 0;
    void ensure(size_t bytes)
    {
        if (bytes <= n) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            fail(EVXA_ERROR_OUTOFMEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        }
        n = bytes;
    }
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

// compute_lambda (common.h:264-266), 0.85 * powf(2, (quality - 12) / 3),
// as the fp32 bits the reference's own build computes for the frame
// qualities 0..63 (set_quality keeps QPs within them, evx3enc.cpp:64-82):
// a table, so no libm's powf rounding reaches the RD costs.
constexpr uint32_t kLambdaBits[64] = {
    0x3D59999Au, 0x3D891454u, 0x3DACB591u, 0x3DD9999Au, 0x3E091454u, 0x3E2CB591u,
    0x3E59999Au, 0x3E891455u, 0x3EACB591u, 0x3ED9999Au, 0x3F091455u, 0x3F2CB591u,
    0x3F59999Au, 0x3F891455u, 0x3FACB591u, 0x3FD9999Au, 0x40091455u, 0x402CB591u,
    0x4059999Au, 0x40891454u, 0x40ACB591u, 0x40D9999Au, 0x41091454u, 0x412CB591u,
    0x4159999Au, 0x41891456u, 0x41ACB590u, 0x41D9999Au, 0x42091456u, 0x422CB590u,
    0x4259999Au, 0x42891456u, 0x42ACB590u, 0x42D9999Au, 0x43091456u, 0x432CB590u,
    0x4359999Au, 0x43891452u, 0x43ACB593u, 0x43D9999Au, 0x44091452u, 0x442CB593u,
    0x4459999Au, 0x44891452u, 0x44ACB593u, 0x44D9999Au, 0x45091452u, 0x452CB593u,
    0x4559999Au, 0x45891452u, 0x45ACB593u, 0x45D9999Au, 0x46091452u, 0x462CB593u,
    0x4659999Au, 0x46891452u, 0x46ACB593u, 0x46D9999Au, 0x47091452u, 0x472CB593u,
    0x4759999Au, 0x47891458u, 0x47ACB58Bu, 0x47D9999Au,
};

float lambda_of(int32_t quality)
{
    float f;
    memcpy(&f, &kLambdaBits[quality & 63], sizeof(f));
    return f;
}

// Device-side resources of one launch configuration: job table, work list,
// predictor words and counters.  Graphs own theirs, so replays never see
// another launch's tables.
struct LaunchRes {
    DevBuf jobs, tickets, progress, counters;
    DevBuf state;                    // sweep mode: per-MB words, keys, predictors and decisions
    DevBuf lists;                    // sweep mode: the dynamic check and search lists (zero between launches)
    DevBuf fpred;                    // MB wavefront: the B jobs' forward predictions
    PinnedBuf h_jobs, h_tickets;
    hipEvent_t staged = nullptr;     // host staging may be reused after this
    hipEvent_t g_start = nullptr, g_stop = nullptr;   // captured timing (graphs)
    std::vector<uint64_t> key;       // composition of the cached work list
    int32_t n_tickets = 0;
    int32_t n_progress = 0;
    int32_t grid = 0;                // workgroups to launch for this composition
    int32_t waves = 1;               // waves per workgroup
    bool sweep = false;              // sweep mode (else the MB wavefront)
    int32_t n_reset = 0;             // progress words cleared before each launch
    int32_t cap_check = 0;           // sweep mode: entries per check sub-list (MBs of the launch)
    int32_t cap_search = 0;          // sweep mode: entries per search sub-list (tickets of the launch)
    int32_t max_sweeps = 0;          // sweep mode: sweep counter blocks
    int32_t n_final = 0;             // sweep mode: MBs of the final pass
    int32_t sctr_off = 0;            // sweep mode: progress word of sweep block 0
    int32_t reads = kSweepReadsGuess;  // sweep mode: how sweep 0 reads neighbours' words
    int32_t wait_ticks = 0;          // sweep mode: fresh reads' longest wait (10 ns ticks)
    int32_t batch = 4;               // sweep mode: tickets per claim
    SweepArgs sweep_args() const
    {
        SweepArgs a;
        a.check = lists.as<uint64_t>();
        a.search = a.check + (size_t)2 * kSubLists * cap_check;
        a.cap_check = cap_check;
        a.cap_search = cap_search;
        a.max_sweeps = max_sweeps;
        a.sctr = progress.as<int32_t>() + sctr_off;
        a.reads = reads;
        a.wait_ticks = wait_ticks;
        a.batch = batch;
        return a;
    }
    ~LaunchRes()
    {
        if (staged) (void)hipEventDestroy(staged);
        if (g_start) (void)hipEventDestroy(g_start);
        if (g_stop) (void)hipEventDestroy(g_stop);
    }
};

}  // namespace

struct evxa_stream;

struct evxa_context {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    // The allocation-free path's scratch on this device (its address from
    // hipGetSymbolAddress), and whether the last analysis used it.
    char *scratch = nullptr;
    bool scratch_last = false;
    // The counter block of the last in-loop encode, when that was the last
    // launch whose errors evxa_synchronize reports.
    const int32_t *enc_last = nullptr;
    // In-loop encode launches: job table, MB order (cached per composition),
    // counter block and per-MB words, coefficient candidates per workgroup.
    struct EncodeRes {
        DevBuf jobs, order, progress, coscratch;
        PinnedBuf h_jobs, h_order;
        hipEvent_t staged = nullptr;
        std::vector<uint64_t> key;
        int32_t n_tickets = 0, grid = 0;
        ~EncodeRes()
        {
            if (staged) (void)hipEventDestroy(staged);
        }
    } enc;
    int32_t enc_resident = 0;        // encode workgroups that fit the device
    uint64_t budget = 0;
    uint64_t used = 0;
    int32_t resident[2][3] = {};     // [sweep mode][waves]: workgroups of the analysis kernels that fit
    int32_t waves = 1;               // waves per search (1 or 2)
    int32_t mode = EVXA_MODE_WAVEFRONT;  // predictor dependencies: wavefront or sweeps
    int32_t sweep_reads = kSweepReadsFresh;  // sweep 0: guess table only, or fresh words too
    int32_t sweep_wait_us = 0;       // sweep 0, fresh reads: longest wait for a neighbour
    int32_t sweep_order = EVXA_SWEEP_ORDER_RASTER;  // sweep 0's ticket order
    int32_t claim_batch = 1;         // sweep mode: tickets per claim
    int32_t reach = 15;              // reach of the window each ticket stages (kernels' maximum)
    int32_t grid = 0;                // most analysis workgroups per launch (0: all that fit)
    int32_t lists = 0;               // work lists per launch (0: automatic)
    int32_t cap_pct = 60;            // tickets per list and virtual diagonal, % of its workgroups
    std::mutex mu;
    LaunchRes res;
    // Ring of (start, stop) event pairs around analysis kernels; launches
    // [ring_tail, ring_head) have not been reported by evxa_launch_times.
    static constexpr size_t kRing = 8192;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ring;
    uint64_t ring_head = 0, ring_tail = 0;
    hipEvent_t last_start = nullptr, last_stop = nullptr;
    std::map<int32_t, std::unique_ptr<LaunchRes>> graph_res;
    std::map<int32_t, hipGraphExec_t> graphs;
    std::map<int32_t, std::vector<evxa_stream *>> graph_streams;
    int32_t next_graph = 1;
    std::vector<evxa_stream *> streams;
};

struct evxa_stream {
    evxa_context *ctx = nullptr;
    int32_t w = 0, h = 0, wb = 0, hb = 0, n_slots = 0;
    PlaneGeom gy{}, gc{};
    // All picture slots in one allocation, slot k at k * slot_stride bytes:
    // one large contiguous buffer maps with the largest page fragments, so the
    // per-CU translation cache covers a stream's pictures with few entries.
    DevBuf pics;
    size_t slot_stride = 0;
    DevBuf tables;         // 3 record tables: P ping, P pong, B
    DevBuf coeffs, levels;
    DevBuf searches;       // EVXA_MAX_REFS SearchRec per MB
    DevBuf wide;           // per slot: the sample range words the analysis reads (2 words)
    DevBuf spread;         // per slot: the spread range record its producers grow
    int32_t p_cur = 0;
    bool temporal_valid = false;
    std::atomic<int32_t> inflight{0};   // frames queued on a server and not yet complete
    int32_t last_table = -1;
    DevJob last_job{};     // the last analysed frame (for reconstruction)
    bool last_encoded = false;   // the last frame came from the in-loop encode
    int32_t last_refs[EVXA_MAX_REFS] = {-1, -1, -1, -1};
    DevBuf recon_levels;   // levels handed in for reconstruction
    DevBuf deblock_src;    // unfiltered copy of the picture being deblocked (packed planes)
    DevBuf deblock_records;  // block descriptors handed in for deblocking
    DevBuf recon_flag;     // range record of a reconstruction headed for the filter (unused)
    DevBuf sao_words;      // per MB: published SAO decision; then the SAO counters
    DevBuf sao_table;      // evxa_sao_info per MB

    uint64_t bytes = 0;
    // Recorded after every asynchronous job that writes the stream's slots or
    // outputs, on the HIP stream that job ran on (evxa_stream_wait).
    hipEvent_t done = nullptr;

    int32_t *range_words(int32_t slot) const { return wide.as<int32_t>() + 2 * slot; }
    int32_t *range_spread(int32_t slot) const
    {
        return spread.as<int32_t>() + (size_t)kRangeWords * slot;
    }
    int16_t *slot_base(int32_t slot) const
    {
        return reinterpret_cast<int16_t *>(static_cast<char *>(pics.p) + (size_t)slot * slot_stride);
    }
    evxa_block_record *table(int k) const
    {
        return tables.as<evxa_block_record>() + (size_t)k * wb * hb;
    }
    DevPic pic(int32_t slot) const
    {
        const int16_t *base = slot_base(slot);
        DevPic p;
        p.y = base + gy.origin();
        const int16_t *ub = base + (size_t)gy.pitch * gy.rows;
        p.u = ub + gc.origin();
        p.v = ub + (size_t)gc.pitch * gc.rows + gc.origin();
        p.py = gy.pitch;
        p.pc = gc.pitch;
        p.wide = wide.as<int32_t>() + 2 * slot;
        return p;
    }
};

namespace {

void set_device(const evxa_context *ctx) { check(hipSetDevice(ctx->device), "hipSetDevice"); }

// The stream's work so far ends with what was just queued on `st`.
void mark_done(evxa_stream *s, hipStream_t st)
{
    check(hipEventRecord(s->done, st), "hipEventRecord");
}

void check_slot(const evxa_stream *s, int32_t slot)
{
    if (slot < 0 || slot >= s->n_slots) fail(EVXA_ERROR_INVALIDARG, "slot out of range");
}

// A slot's range record: cleared before the kernels that write the picture
// (they grow its spread record), folded into the two words the analysis
// reads after them.
void range_begin(evxa_stream *s, int32_t slot)
{
    check(launch_reset(s->range_spread(slot), kRangeWords, s->ctx->stream), "reset range record");
}

void range_end(evxa_stream *s, int32_t slot)
{
    check(launch_fold_range(s->range_spread(slot), s->range_words(slot), s->ctx->stream),
          "fold range record");
}

// Allocate a buffer a stream creates on first use and charge it to the
// context's HBM budget (the caller holds ctx->mu).
void lazy_alloc(evxa_stream *s, DevBuf &b, size_t bytes)
{
    if (b.p) return;
    evxa_context *ctx = s->ctx;
    if (ctx->used + bytes > ctx->budget) fail(EVXA_ERROR_CAPACITY_LIMIT, "buffer does not fit the HBM budget");
    b.alloc(bytes);
    s->bytes += bytes;
    ctx->used += bytes;
}

// Build the job descriptor of one frame and advance the stream's
// temporal-table state (P frames ping-pong, B frames use a third table).
DevJob make_job(evxa_stream *s, const evxa_frame_desc &d, bool advance)
{
    if (d.type != EVXA_FRAME_P && d.type != EVXA_FRAME_B) fail(EVXA_ERROR_INVALIDARG, "frame type");
    if (d.type == EVXA_FRAME_P && (d.n_refs < 1 || d.n_refs > EVXA_MAX_REFS))
        fail(EVXA_ERROR_INVALIDARG, "P frames take 1..4 references");
    if (d.type == EVXA_FRAME_B && d.n_refs != 2)
        fail(EVXA_ERROR_INVALIDARG, "B frames take 2 references");
    if (d.quality < 0 || d.quality > 63) fail(EVXA_ERROR_INVALIDARG, "quality must be 0..63");
    check_slot(s, d.src_slot);
    DevJob j;
    memset(&j, 0, sizeof(j));
    j.src = s->pic(d.src_slot);
    for (int k = 0; k < d.n_refs; k++) {
        check_slot(s, d.ref_slots[k]);
        j.refs[k] = s->pic(d.ref_slots[k]);
    }
    j.width = s->w;
    j.height = s->h;
    j.wb = s->wb;
    j.hb = s->hb;
    j.type = d.type;
    j.quality = d.quality;
    j.n_refs = d.n_refs;
    j.thresh = (d.quality >> 2) + 1;   // motion.cpp:589
    j.reach = s->ctx->reach;
    j.lambda = lambda_of(d.quality);
    j.target_ids = (uint32_t)d.target_ids[0] | ((uint32_t)d.target_ids[1] << 8) |
                   ((uint32_t)d.target_ids[2] << 16) | ((uint32_t)d.target_ids[3] << 24);
    int out;
    if (d.type == EVXA_FRAME_P) {
        out = s->p_cur;
        j.prev = (d.temporal_mvp && s->temporal_valid) ? s->table(s->p_cur ^ 1) : nullptr;
        // Sweep mode starts from the temporal table's motion (the information
        // the predictor itself uses), never from an earlier result of the
        // same frame.
        j.guess = j.prev;
    } else {
        out = 2;
        j.prev = nullptr;
    }
    j.records = s->table(out);
    j.coeffs = s->coeffs.as<int16_t>();
    j.levels = s->levels.as<int16_t>();
    j.searches = s->searches.as<SearchRec>();
    if (advance) {
        s->last_table = out;
        s->last_job = j;
        s->last_encoded = false;
        for (int k = 0; k < EVXA_MAX_REFS; k++) s->last_refs[k] = k < d.n_refs ? d.ref_slots[k] : -1;
        if (d.type == EVXA_FRAME_P && d.update_temporal) {
            s->p_cur ^= 1;
            s->temporal_valid = true;
        }
    }
    return j;
}

// Work list: anti-diagonal order t = bx + 2 by across all jobs, the
// reference searches of a MB adjacent, so every ticket's predecessors (left,
// top-left, top, top-right MBs) carry smaller tickets and are owned by waves
// already running.  Returns the largest number of items on one diagonal.
// Work buffer of a launch (header + tickets, layout at kWorkHeader).
//
// Lists: each job's MB rows are laid end to end, weighted by their work
// (MBs x refs), and cut into n_lists equal segments, so a large frame splits
// into row bands over several lists (XCDs) and small ones share a list.
//
// Order: ticket (job j, MB (bx, by), ref r) has virtual diagonal
// v = bx + 2 by + start[j]; each list runs by v, then job, row and ref.  The
// per-job start (the same in every list) staggers the jobs so that each list
// keeps about `per_list_cap` tickets per virtual diagonal: an XCD then works
// on few pictures at a time and its L2 holds their search fronts.  Within a
// job v follows the dependency order (left, top and top-right neighbours
// have smaller v), and across jobs there are no dependencies, so the item
// with the smallest unfinished v is always runnable: no list can deadlock.
// Returns the most tickets any list holds on one virtual diagonal.
//
// Sweep mode appends the final pass's lists (one ticket per MB, reference 0,
// in MB order, each MB in the list of its row) and fills the sweep fields
// of the header.  Its sweep 0 has no dependencies: `raster` lists each
// frame's rows in raster order instead (an XCD then works on few pictures
// at a time), while the anti-diagonal order lets sweep 0 read the MVs its
// neighbours decided before it.
int32_t build_work(const std::vector<DevJob> &jobs, int32_t n_lists, int32_t per_list_cap,
                   bool stagger, bool sweep, bool raster, std::vector<uint32_t> &out)
{
    int64_t total = 0;
    for (const auto &j : jobs) total += (int64_t)j.wb * j.hb * j.n_refs;
    std::vector<std::vector<uint8_t>> row_list(jobs.size());
    int64_t cum = 0;
    for (size_t k = 0; k < jobs.size(); k++) {
        const DevJob &j = jobs[k];
        row_list[k].resize(j.hb);
        for (int32_t by = 0; by < j.hb; by++) {
            const int64_t w = (int64_t)j.wb * j.n_refs;
            row_list[k][by] = (uint8_t)std::min<int64_t>(n_lists - 1, (cum + w / 2) * n_lists / total);
            cum += w;
        }
    }
    // Tickets per (list, diagonal) of each job.
    auto diagonals = [](const DevJob &j) { return j.wb + 2 * (j.hb - 1); };
    std::vector<std::vector<int32_t>> prof(jobs.size());
    for (size_t k = 0; k < jobs.size(); k++) {
        const DevJob &j = jobs[k];
        prof[k].assign((size_t)n_lists * diagonals(j), 0);
        for (int32_t by = 0; by < j.hb; by++)
            for (int32_t bx = 0; bx < j.wb; bx++)
                prof[k][(size_t)row_list[k][by] * diagonals(j) + bx + 2 * by] += j.n_refs;
    }
    // Starts: longest MB chains (most diagonals) first, then the most work,
    // each at the earliest start that keeps every list within the cap (or,
    // failing that, at the end of what is placed).  A frame cannot finish
    // sooner than its chain of diagonals, so long chains that start late
    // would set the end of the launch.
    std::vector<int32_t> order(jobs.size()), start(jobs.size(), 0);
    for (size_t k = 0; k < jobs.size(); k++) order[k] = (int32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        if (diagonals(jobs[a]) != diagonals(jobs[b]))
            return diagonals(jobs[a]) > diagonals(jobs[b]);
        return (int64_t)jobs[a].wb * jobs[a].hb * jobs[a].n_refs >
               (int64_t)jobs[b].wb * jobs[b].hb * jobs[b].n_refs;
    });
    std::vector<std::vector<int32_t>> load(n_lists);
    int32_t horizon = 0;
    for (int32_t k : order) {
        const int32_t nd = diagonals(jobs[k]);
        auto fits = [&](int32_t s0) {
            for (int32_t l = 0; l < n_lists; l++)
                for (int32_t t = 0; t < nd; t++) {
                    const int32_t add = prof[k][(size_t)l * nd + t];
                    if (!add) continue;
                    const int32_t have = s0 + t < (int32_t)load[l].size() ? load[l][s0 + t] : 0;
                    if (have > 0 && have + add > per_list_cap) return false;
                }
            return true;
        };
        int32_t s0 = 0;
        while (stagger && s0 < horizon && !fits(s0)) s0 += 2;
        start[k] = s0;
        for (int32_t l = 0; l < n_lists; l++) {
            if ((int32_t)load[l].size() < s0 + nd) load[l].resize(s0 + nd, 0);
            for (int32_t t = 0; t < nd; t++) load[l][s0 + t] += prof[k][(size_t)l * nd + t];
        }
        horizon = std::max(horizon, s0 + nd);
    }
    int32_t widest = 0;
    for (int32_t l = 0; l < n_lists; l++)
        for (int32_t v : load[l]) widest = std::max(widest, v);
    std::vector<std::vector<uint32_t>> lists(n_lists);
    // Raster order: each list takes its rows frame by frame (the references
    // of a MB adjacent, the last one designated).
    if (raster) horizon = 0;
    for (size_t k = 0; raster && k < jobs.size(); k++) {
        const DevJob &j = jobs[k];
        for (int32_t by = 0; by < j.hb; by++)
            for (int32_t bx = 0; bx < j.wb; bx++)
                for (int32_t r = 0; r < j.n_refs; r++)
                    lists[row_list[k][by]].push_back(((uint32_t)k << kJobShift) |
                                                     ((uint32_t)r << kRefShift) |
                                                     (uint32_t)(by * j.wb + bx));
    }
    for (int32_t v = 0; v < horizon; v++)
        for (size_t k = 0; k < jobs.size(); k++) {
            const DevJob &j = jobs[k];
            const int32_t t = v - start[k];
            if (t < 0 || t >= diagonals(j)) continue;
            for (int32_t by = 0; by < j.hb; by++) {
                const int32_t bx = t - 2 * by;
                if (bx < 0) break;
                if (bx >= j.wb) continue;
                auto &dst = lists[row_list[k][by]];
                for (int32_t r = 0; r < j.n_refs; r++)
                    dst.push_back(((uint32_t)k << kJobShift) | ((uint32_t)r << kRefShift) |
                                  (uint32_t)(by * j.wb + bx));
            }
        }
    out.assign(kWorkHeader, 0u);
    out[0] = (uint32_t)n_lists;
    uint32_t off = 0;
    for (int32_t l = 0; l < n_lists; l++) {
        out[1 + l] = off;
        off += (uint32_t)lists[l].size();
        out.insert(out.end(), lists[l].begin(), lists[l].end());
    }
    out[1 + n_lists] = off;
    if (sweep) {
        int32_t non_empty = 0, max_diag = 0;
        for (int32_t l = 0; l < n_lists; l++) non_empty += lists[l].empty() ? 0 : 1;
        for (const auto &j : jobs) max_diag = std::max(max_diag, diagonals(j));
        for (int32_t l = 0; l < n_lists; l++) {
            out[kWorkFinal + l] = off;
            for (size_t k = 0; k < jobs.size(); k++) {
                const DevJob &j = jobs[k];
                for (int32_t by = 0; by < j.hb; by++) {
                    if (row_list[k][by] != l) continue;
                    for (int32_t bx = 0; bx < j.wb; bx++)
                        out.push_back(((uint32_t)k << kJobShift) | (uint32_t)(by * j.wb + bx));
                    off += (uint32_t)j.wb;
                }
            }
        }
        out[kWorkFinal + n_lists] = off;
        out[kWorkNonEmpty] = (uint32_t)non_empty;
        // A MB at dependency depth d is final after sweep d, and depths stay
        // below the anti-diagonal count; sweep s + 2's block is addressed
        // during sweep s.
        out[kWorkMaxSweeps] = (uint32_t)(max_diag + 3);
    }
    return widest;
}

// Progress words (int32, cleared per launch) of a frame of n_mb MBs in
// sweep mode: the listed stamps (padded to 8 bytes), the predictor words of
// both parities and the published predictors (64-bit each).
int32_t sweep_words(int32_t n_mb) { return (int32_t)align_up((size_t)n_mb, 2) + 6 * n_mb; }

// The reset and analysis kernels of a staged launch (also captured into
// graphs).
hipError_t launch_staged(LaunchRes &r, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    hipError_t e = launch_reset(r.progress.as<int32_t>(), r.n_reset, st);
    if (e != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    e = r.sweep ? launch_sweep(r.jobs.as<DevJob>(), r.tickets.as<uint32_t>(), r.progress.as<int32_t>(),
                               r.sweep_args(), r.grid, r.waves, r.n_final, st)
                : launch_analyze(r.jobs.as<DevJob>(), r.tickets.as<uint32_t>(), r.progress.as<int32_t>(),
                                 r.grid, r.waves, st);
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    return hipSuccess;
}

// Stage jobs + work list into `r` and enqueue reset + analysis kernels on
// `st` (the context's queue unless given).  `grid_cap` bounds the waves of
// the launch (0: all that fit the device); `stagger` staggers the jobs'
// diagonals per list (batches of whole streams) or starts all at once
// (serving, where compositions change every step).  `ev0`/`ev1`, if given,
// are recorded around the analysis kernel instead of the context's ring.
// `launch` false stages only (graph capture launches it).
void enqueue(evxa_context *ctx, LaunchRes &r, std::vector<DevJob> &jobs, bool time_it,
             hipStream_t st = nullptr, int32_t grid_cap = 0, bool stagger = true,
             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, bool launch = true)
{
    if (!st) st = ctx->stream;
    if (st == ctx->stream) {
        ctx->scratch_last = false;
        ctx->enc_last = nullptr;
    }
    const bool sweep = ctx->mode == EVXA_MODE_SWEEP;
    const int32_t fit = ctx->resident[sweep ? 1 : 0][ctx->waves];
    const int32_t all = ctx->grid > 0 ? std::min(ctx->grid, fit) : fit;
    const int32_t max_grid = grid_cap > 0 ? std::min(grid_cap, all) : all;
    if (jobs.empty()) return;
    if ((int)jobs.size() > kMaxJobs) fail(EVXA_ERROR_CAPACITY_LIMIT, "too many jobs in one launch");
    std::vector<uint64_t> key;
    int32_t n_words = 0;           // per-MB words cleared per launch
    int64_t state_bytes = 0, n_mbs = 0, n_tickets = 0;
    for (const auto &j : jobs) {
        n_mbs += (int64_t)j.wb * j.hb;
        n_tickets += (int64_t)j.wb * j.hb * j.n_refs;
        if ((int64_t)j.wb * j.hb > (int64_t)kMbMask) fail(EVXA_ERROR_CAPACITY_LIMIT, "frame too large");
        key.push_back(((uint64_t)j.wb << 40) | ((uint64_t)j.hb << 8) | (uint32_t)j.n_refs);
        // Wavefront: one 64-bit predictor word per MB.  Sweep: the listed
        // stamp, the words of both parities and the predictor.
        n_words += sweep ? sweep_words(j.wb * j.hb) : j.wb * j.hb * 2;
        state_bytes += (int64_t)align_up((size_t)j.wb * j.hb * (sizeof(Decision) + 4), 256);
    }
    key.push_back(((uint64_t)max_grid << 32) | ((uint64_t)ctx->waves << 24) |
                  ((uint64_t)ctx->lists << 16) | ((uint64_t)(stagger ? 1 : 0) << 15) |
                  (uint64_t)(ctx->cap_pct & 0x7FFF));
    key.push_back(sweep ? 1 + ctx->sweep_order : 0);
    if (!r.staged) check(hipEventCreateWithFlags(&r.staged, hipEventDisableTiming), "hipEventCreate");
    // The previous launch's host->device copies must have consumed the
    // staging buffers before they are rewritten.
    check(hipEventSynchronize(r.staged), "hipEventSynchronize");

    if (key != r.key) {
        std::vector<uint32_t> t;
        // Wavefront: one list for a single picture (its rows all stay in
        // flight), one per XCD for batches.  Sweeps: one per XCD.
        const int32_t lists = ctx->lists > 0 ? ctx->lists : ((jobs.size() > 1 || sweep) ? kLists : 1);
        const int32_t n_lists = std::max(1, std::min(lists, max_grid));
        const int32_t cap = std::max(16, max_grid / n_lists * ctx->cap_pct / 100);
        const bool raster = sweep && ctx->sweep_order == EVXA_SWEEP_ORDER_RASTER;
        const int32_t widest = build_work(jobs, n_lists, cap, stagger && !raster, sweep, raster, t);
        if (sweep) {
            // Every workgroup that fits: sweep 0 is throughput work.
            r.grid = n_lists * std::max(1, max_grid / n_lists);
            r.max_sweeps = (int32_t)t[kWorkMaxSweeps];
            r.n_final = (int32_t)(t[kWorkFinal + t[0]] - t[kWorkFinal]);
        } else {
            // Per list, enough waves for its widest diagonal (plus slack for
            // the MBs still finishing the previous one), never more than fit
            // at once: waiting waves beyond that only add polling traffic.
            // The grid is a multiple of n_lists so that every list has the
            // same number of waves.
            const int32_t per_list = std::max(8, widest + widest / 2 + 4);
            r.grid = n_lists * std::max(1, std::min(max_grid / n_lists, per_list));
        }
        r.waves = ctx->waves;
        r.sweep = sweep;
        r.h_tickets.ensure(t.size() * sizeof(uint32_t));
        memcpy(r.h_tickets.p, t.data(), t.size() * sizeof(uint32_t));
        if (r.tickets.n < t.size() * sizeof(uint32_t)) r.tickets.alloc(t.size() * sizeof(uint32_t));
        check(hipMemcpyAsync(r.tickets.p, r.h_tickets.p, t.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, st),
              "copy work list");
        r.n_tickets = (int32_t)t.size() - kWorkHeader;
        r.key = key;
    }
    // Progress words: the counter block, the per-MB words, then (sweep mode)
    // the sweep counter blocks, of which blocks 0..2 are cleared here and
    // the kernel clears the others as it goes.
    r.n_progress = n_words;
    r.reads = ctx->sweep_reads;
    r.wait_ticks = ctx->sweep_wait_us * 100;
    r.batch = ctx->claim_batch;
    r.sctr_off = (int32_t)align_up((size_t)(kCounterWords + n_words), 64);
    const size_t prog_words = r.sweep ? (size_t)r.sctr_off + (size_t)r.max_sweeps * kSweepStride
                                      : (size_t)(kCounterWords + n_words);
    r.n_reset = r.sweep ? r.sctr_off + 3 * kSweepStride : kCounterWords + n_words;
    if (r.progress.n < prog_words * sizeof(int32_t)) r.progress.alloc(prog_words * sizeof(int32_t));
    if (r.sweep) {
        if (r.state.n < (size_t)state_bytes) r.state.alloc((size_t)state_bytes);
        // Each sub-list can receive every MB (check) or ticket (search) of
        // the launch: a workgroup fills only the sub-list of its index.
        r.cap_check = (int32_t)n_mbs;
        r.cap_search = (int32_t)n_tickets;
        const size_t list_bytes = 2 * kSubLists * (size_t)(n_mbs + n_tickets) * sizeof(uint64_t);
        if (r.lists.n < list_bytes) {
            r.lists.alloc(list_bytes);
            check(hipMemsetAsync(r.lists.p, 0, list_bytes, st), "clear ticket lists");
        }
    }
    int32_t *prog = r.progress.as<int32_t>();
    // MB wavefront, B frames: the forward search's wave stores its prediction
    // for the deciding wave, which then needs no interpolation of its own.
    size_t fpred_bytes = 0;
    if (!r.sweep)
        for (const auto &j : jobs)
            if (j.type == EVXA_FRAME_B) fpred_bytes += (size_t)j.wb * j.hb * EVXA_MB_SAMPLES * sizeof(int16_t);
    if (r.fpred.n < fpred_bytes) r.fpred.alloc(fpred_bytes);
    int16_t *fpred = r.fpred.as<int16_t>();
    int32_t off = kCounterWords;
    char *state = r.state.as<char>();
    for (auto &j : jobs) {
        const size_t n_mb = (size_t)j.wb * j.hb;
        j.fpred = nullptr;
        if (!r.sweep && j.type == EVXA_FRAME_B) {
            j.fpred = fpred;
            fpred += n_mb * EVXA_MB_SAMPLES;
        }
        if (r.sweep) {
            j.mvw = nullptr;
            const int32_t nl = (int32_t)align_up(n_mb, 2);
            j.sw_listed = reinterpret_cast<uint32_t *>(prog + off);
            j.sw_words = reinterpret_cast<uint64_t *>(prog + off + nl);
            j.sw_pred = reinterpret_cast<uint64_t *>(prog + off + nl + 4 * (int32_t)n_mb);
            off += sweep_words((int32_t)n_mb);
            j.sw_dec = reinterpret_cast<Decision *>(state);
            j.sw_key = reinterpret_cast<uint32_t *>(state + n_mb * sizeof(Decision));
            state += align_up(n_mb * (sizeof(Decision) + 4), 256);
            j.eager = 0;
        } else {
            j.mvw = reinterpret_cast<uint64_t *>(prog + off);
            off += (int32_t)n_mb * 2;
            // A lone picture is latency-bound: the wave of whichever
            // reference finishes last can select and publish at once,
            // without waiting for the designated wave to see its result.  In
            // batches the extra read per search costs throughput, so only
            // the designated wave selects.
            j.eager = jobs.size() == 1 && j.type == EVXA_FRAME_P && j.n_refs > 1;
        }
    }
    const size_t jb = jobs.size() * sizeof(DevJob);
    r.h_jobs.ensure(jb);
    memcpy(r.h_jobs.p, jobs.data(), jb);
    if (r.jobs.n < jb) r.jobs.alloc(jb);
    check(hipMemcpyAsync(r.jobs.p, r.h_jobs.p, jb, hipMemcpyHostToDevice, st), "copy jobs");
    check(hipEventRecord(r.staged, st), "hipEventRecord");
    if (!launch) return;
    std::pair<hipEvent_t, hipEvent_t> ev{ev0, ev1};
    if (time_it) {
        if (ctx->ring_head - ctx->ring_tail == evxa_context::kRing) ctx->ring_tail++;
        ev = ctx->ring[ctx->ring_head % evxa_context::kRing];
    }
    check(launch_staged(r, st, ev.first, ev.second), "analysis kernel");
    if (time_it) {
        ctx->ring_head++;
        ctx->last_start = ev.first;
        ctx->last_stop = ev.second;
    }
}

// Where the allocation-free path keeps a frame's launch in the scratch: the
// counter block and predictor words first (fixed offset, so a later call can
// read the error word), the job descriptor, then the work list; `end` is the
// bytes used.
struct NoallocLayout {
    size_t prog, job, work, end;
};

NoallocLayout noalloc_layout(size_t n_mb, int32_t n_refs)
{
    NoallocLayout l;
    l.prog = 0;
    l.job = align_up((kCounterWords + 2 * n_mb) * sizeof(int32_t), 256);
    l.work = l.job + align_up(sizeof(DevJob), 256);
    l.end = l.work + (kWorkHeader + n_mb * (size_t)n_refs) * sizeof(uint32_t);
    return l;
}

void check_counters(evxa_context *ctx, LaunchRes &r)
{
    if (!r.progress.p) return;
    int32_t c[kCounterError + 1];
    check(hipMemcpy(c, r.progress.p, sizeof(c), hipMemcpyDeviceToHost), "read counters");
    // A launch that stopped early may leave ticket entries behind.
    if (c[kCounterError] && r.lists.p) check(hipMemset(r.lists.p, 0, r.lists.n), "clear ticket lists");
    if (c[kCounterError] & kErrSweepLimit)
        fail(EVXA_ERROR_EXECUTION_FAILURE, "analysis kernel: sweep limit reached");
    if (c[kCounterError] & kErrWaitTimeout)
        fail(EVXA_ERROR_OPERATION_TIMEDOUT, "analysis kernel: dependency wait timed out");
    if (c[kCounterError] & kErrSearchFailed)
        fail(EVXA_ERROR_EXECUTION_FAILURE, "analysis kernel: motion search failed");
}

// The error word of the context's last analysis launch (its queue is idle).
void check_last_launch(evxa_context *ctx)
{
    if (ctx->enc_last) {
        int32_t c[kCounterError + 1];
        check(hipMemcpy(c, ctx->enc_last, sizeof(c), hipMemcpyDeviceToHost), "read counters");
        if (c[kCounterError] & kErrWaitTimeout)
            fail(EVXA_ERROR_OPERATION_TIMEDOUT, "encode kernel: dependency wait timed out");
        if (c[kCounterError] & kErrSearchFailed)
            fail(EVXA_ERROR_EXECUTION_FAILURE, "encode kernel: block search failed");
        return;
    }
    if (!ctx->scratch_last) {
        check_counters(ctx, ctx->res);
        return;
    }
    int32_t c[kCounterError + 1];
    check(hipMemcpy(c, ctx->scratch + noalloc_layout(0, 0).prog, sizeof(c), hipMemcpyDeviceToHost),
          "read counters");
    if (c[kCounterError] & kErrWaitTimeout)
        fail(EVXA_ERROR_OPERATION_TIMEDOUT, "analysis kernel: dependency wait timed out");
    if (c[kCounterError] & kErrSearchFailed)
        fail(EVXA_ERROR_EXECUTION_FAILURE, "analysis kernel: motion search failed");
}

// The frames' in-loop encode in one launch (the caller holds ctx->mu).
// Validates every descriptor before anything is queued.
void encode_launch(evxa_context *ctx, evxa_stream *const *streams, const evxa_encode_desc *descs,
                   int32_t n)
{
    if (n > (int32_t)(1u << (32 - kMbBits))) fail(EVXA_ERROR_CAPACITY_LIMIT, "too many frames in one launch");
    for (int32_t k = 0; k < n; k++) {
        evxa_stream *s = streams[k];
        const evxa_encode_desc *d = &descs[k];
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        if (d->type != EVXA_FRAME_I && d->type != EVXA_FRAME_P && d->type != EVXA_FRAME_B)
            fail(EVXA_ERROR_INVALIDARG, "frame type");
        if (d->quality < 0 || d->quality > 63) fail(EVXA_ERROR_INVALIDARG, "quality must be 0..63");
        if (s->w < 48 || s->h < 48) fail(EVXA_ERROR_INVALIDARG, "the encode needs frames of at least 48x48");
        if ((int64_t)s->wb * s->hb > (int64_t)kMbMask) fail(EVXA_ERROR_CAPACITY_LIMIT, "frame too large");
        check_slot(s, d->src_slot);
        check_slot(s, d->dst_slot);
        if (d->dst_slot == d->src_slot) fail(EVXA_ERROR_INVALIDARG, "dst_slot is the source");
        const int32_t n_refs = d->type == EVXA_FRAME_I ? 0 : d->type == EVXA_FRAME_P ? 1 : 2;
        for (int r = 0; r < n_refs; r++) {
            check_slot(s, d->ref_slots[r]);
            if (d->ref_slots[r] == d->dst_slot) fail(EVXA_ERROR_INVALIDARG, "dst_slot is a reference");
        }
        // A target id names the decoder's picture-buffer slot, so two
        // references share one only if they are the same picture (the
        // reference encoder's targets are its slots, encode.cpp:397, 418).
        if (n_refs == 2 && d->target_ids[0] == d->target_ids[1] && d->ref_slots[0] != d->ref_slots[1])
            fail(EVXA_ERROR_INVALIDARG, "two reference slots under one target id");
    }
    set_device(ctx);
    auto &r = ctx->enc;
    if (!r.staged) check(hipEventCreateWithFlags(&r.staged, hipEventDisableTiming), "hipEventCreate");
    check(hipEventSynchronize(r.staged), "hipEventSynchronize");   // staging free again
    if (ctx->enc_resident <= 0) ctx->enc_resident = encode_resident_groups(ctx->device);
    hipStream_t st = ctx->stream;

    // The MB order of the composition: diagonal v = bx + (lag + 1) by of
    // each frame, frames interleaved by v, a diagonal's MBs by row; every
    // MB's dependencies (same frame, smaller v) come first.
    const int32_t slope = kEncodeLag + 1;
    std::vector<uint64_t> key;
    size_t n_mb_total = 0;
    for (int32_t k = 0; k < n; k++) {
        key.push_back(((uint64_t)streams[k]->wb << 32) | (uint32_t)streams[k]->hb);
        n_mb_total += (size_t)streams[k]->wb * streams[k]->hb;
    }
    if (key != r.key) {
        int32_t max_v = 0;
        for (int32_t k = 0; k < n; k++)
            max_v = std::max(max_v, streams[k]->wb + slope * (streams[k]->hb - 1));
        std::vector<uint32_t> order;
        order.reserve(n_mb_total);
        int32_t widest = 1;
        for (int32_t v = 0; v < max_v; v++) {
            int32_t on_v = 0;
            for (int32_t k = 0; k < n; k++) {
                const int32_t wb = streams[k]->wb, hb = streams[k]->hb;
                for (int32_t by = 0; by < hb; by++) {
                    const int32_t bx = v - slope * by;
                    if (bx < 0) break;
                    if (bx < wb) {
                        order.push_back(((uint32_t)k << kMbBits) | (uint32_t)(by * wb + bx));
                        on_v++;
                    }
                }
            }
            widest = std::max(widest, on_v);
        }
        r.h_order.ensure(order.size() * sizeof(uint32_t));
        memcpy(r.h_order.p, order.data(), order.size() * sizeof(uint32_t));
        if (r.order.n < order.size() * sizeof(uint32_t)) r.order.alloc(order.size() * sizeof(uint32_t));
        check(hipMemcpyAsync(r.order.p, r.h_order.p, order.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, st),
              "copy encode order");
        r.n_tickets = (int32_t)order.size();
        // Enough workgroups for twice the widest diagonal (the MBs that can
        // run at once, and the next ones waiting), at most all that fit.
        r.grid = std::max(1, std::min(ctx->enc_resident, 2 * widest + 8));
        r.key = key;
    }
    // Counter block, then the per-MB words of every frame.
    const size_t prog_words = kCounterWords + 2 * n_mb_total;
    if (r.progress.n < prog_words * sizeof(int32_t)) r.progress.alloc(prog_words * sizeof(int32_t));
    const size_t cos_bytes = (size_t)ctx->enc_resident * kEncodeCands * EVXA_MB_SAMPLES * sizeof(int16_t);
    if (r.coscratch.n < cos_bytes) r.coscratch.alloc(cos_bytes);
    int32_t *counters = r.progress.as<int32_t>();
    uint64_t *words = reinterpret_cast<uint64_t *>(counters + kCounterWords);

    std::vector<EncJob> jobs(n);
    std::vector<int> outs(n);
    for (int32_t k = 0; k < n; k++) {
        evxa_stream *s = streams[k];
        const evxa_encode_desc *d = &descs[k];
        const int32_t n_refs = d->type == EVXA_FRAME_I ? 0 : d->type == EVXA_FRAME_P ? 1 : 2;
        DevJob &j = jobs[k].job;
        memset(&j, 0, sizeof(j));
        j.src = s->pic(d->src_slot);
        for (int r2 = 0; r2 < n_refs; r2++) j.refs[r2] = s->pic(d->ref_slots[r2]);
        j.width = s->w;
        j.height = s->h;
        j.wb = s->wb;
        j.hb = s->hb;
        j.type = d->type;
        j.quality = d->quality;
        j.n_refs = n_refs;
        j.thresh = (d->quality >> 2) + 1;   // motion.cpp:589
        j.lambda = lambda_of(d->quality);
        j.target_ids = (uint32_t)d->target_ids[0] | ((uint32_t)d->target_ids[1] << 8);
        // The reference keeps one temporal table, the block table of the
        // last non-B frame (encode.cpp:990-997); I and P frames write the P
        // table in use, B frames their own.
        outs[k] = d->type == EVXA_FRAME_B ? 2 : s->p_cur;
        j.prev = (d->type == EVXA_FRAME_P && d->temporal_mvp && s->temporal_valid)
                     ? s->table(s->p_cur ^ 1) : nullptr;
        j.records = s->table(outs[k]);
        j.coeffs = s->coeffs.as<int16_t>();
        j.levels = s->levels.as<int16_t>();
        j.mvw = words;
        words += (size_t)s->wb * s->hb;
        jobs[k].pic = s->pic(d->dst_slot);
        jobs[k].spread = s->range_spread(d->dst_slot);
    }
    const size_t jb = jobs.size() * sizeof(EncJob);
    r.h_jobs.ensure(jb);
    memcpy(r.h_jobs.p, jobs.data(), jb);
    if (r.jobs.n < jb) r.jobs.alloc(jb);
    check(hipMemcpyAsync(r.jobs.p, r.h_jobs.p, jb, hipMemcpyHostToDevice, st), "copy encode jobs");
    check(hipEventRecord(r.staged, st), "hipEventRecord");
    for (int32_t k = 0; k < n; k++) range_begin(streams[k], descs[k].dst_slot);
    check(launch_reset(counters, (int32_t)prog_words, st), "reset kernel");
    if (ctx->ring_head - ctx->ring_tail == evxa_context::kRing) ctx->ring_tail++;
    const auto ev = ctx->ring[ctx->ring_head % evxa_context::kRing];
    check(hipEventRecord(ev.first, st), "hipEventRecord");
    check(launch_encode(r.jobs.as<EncJob>(), r.order.as<uint32_t>(), r.n_tickets, counters,
                        r.coscratch.as<int16_t>(), r.grid, st),
          "encode kernel");
    check(hipEventRecord(ev.second, st), "hipEventRecord");
    ctx->ring_head++;
    ctx->last_start = ev.first;
    ctx->last_stop = ev.second;
    ctx->scratch_last = false;
    ctx->enc_last = counters;
    for (int32_t k = 0; k < n; k++) {
        evxa_stream *s = streams[k];
        const evxa_encode_desc *d = &descs[k];
        range_end(s, d->dst_slot);
        s->last_table = outs[k];
        s->last_job = jobs[k].job;
        s->last_encoded = true;
        const int32_t n_refs = d->type == EVXA_FRAME_I ? 0 : d->type == EVXA_FRAME_P ? 1 : 2;
        for (int r2 = 0; r2 < EVXA_MAX_REFS; r2++) s->last_refs[r2] = r2 < n_refs ? d->ref_slots[r2] : -1;
        if (d->type != EVXA_FRAME_B && d->update_temporal) {
            s->p_cur ^= 1;
            s->temporal_valid = true;
        }
        mark_done(s, st);
    }
}

}  // namespace

extern "C" {

int32_t evxa_version(void) { return 105; }

const char *evxa_last_error(void) { return g_last_error.c_str(); }

float evxa_lambda(int32_t quality) { return lambda_of(quality); }

uint64_t evxa_stream_bytes(int32_t width, int32_t height, int32_t n_slots)
{
    if (width <= 0 || height <= 0 || n_slots <= 0) return 0;
    return (uint64_t)n_slots * (align_up(picture_bytes(width, height), 4096) +
                                (2 + kRangeWords) * sizeof(int32_t)) +
           outputs_bytes(width, height);
}

uint64_t evxa_context_bytes(const evxa_context *ctx) { return ctx ? ctx->used : 0; }

int evxa_create(int32_t device_ordinal, uint64_t hbm_budget_bytes, evxa_context **out)
{
    return guarded([&] {
        if (!out) fail(EVXA_ERROR_INVALIDARG, "out is null");
        *out = nullptr;
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device_ordinal < 0 || device_ordinal >= n) fail(EVXA_ERROR_INVALIDARG, "no such device");
        auto ctx = std::make_unique<evxa_context>();
        ctx->device = device_ordinal;
        set_device(ctx.get());
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        ctx->ring.resize(evxa_context::kRing);
        for (auto &e : ctx->ring) {
            check(hipEventCreate(&e.first), "hipEventCreate");
            check(hipEventCreate(&e.second), "hipEventCreate");
        }
        if (hbm_budget_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            hbm_budget_bytes = (uint64_t)(free_b * 0.9);
        }
        ctx->budget = hbm_budget_bytes;
        void *sp = nullptr;
        check(noalloc_scratch(&sp), "hipGetSymbolAddress(scratch)");
        ctx->scratch = static_cast<char *>(sp);
        for (int m = 0; m < 2; m++)
            for (int w = 1; w <= 2; w++) ctx->resident[m][w] = analyze_resident_groups(device_ordinal, w, m != 0);
        if (const char *g = getenv("EVXA_MODE")) {
            const int v = atoi(g);
            if (v == EVXA_MODE_WAVEFRONT || v == EVXA_MODE_SWEEP) ctx->mode = v;
        }
        if (const char *g = getenv("EVXA_SWEEP_READS")) {
            const int v = atoi(g);
            if (v == kSweepReadsGuess || v == kSweepReadsFresh) ctx->sweep_reads = v;
        }
        if (const char *g = getenv("EVXA_SWEEP_WAIT_US")) {
            const int v = atoi(g);
            if (v >= 0 && v <= 100000) ctx->sweep_wait_us = v;
        }
        if (const char *g = getenv("EVXA_WAVES")) {
            const int v = atoi(g);
            if (v == 1 || v == 2) ctx->waves = v;
        }
        if (const char *g = getenv("EVXA_ANALYZE_GRID")) {
            const int v = atoi(g);
            if (v > 0) ctx->grid = v;
        }
        if (const char *g = getenv("EVXA_LISTS")) {
            const int v = atoi(g);
            if (v > 0 && v <= kLists) ctx->lists = v;
        }
        if (const char *g = getenv("EVXA_CAP_PCT")) {
            const int v = atoi(g);
            if (v > 0) ctx->cap_pct = v;
        }
        if (const char *g = getenv("EVXA_SWEEP_ORDER")) {
            const int v = atoi(g);
            if (v == EVXA_SWEEP_ORDER_RASTER || v == EVXA_SWEEP_ORDER_DIAGONAL) ctx->sweep_order = v;
        }
        if (const char *g = getenv("EVXA_CLAIM_BATCH")) {
            const int v = atoi(g);
            if (v >= 1 && v <= kMaxClaimBatch) ctx->claim_batch = v;
        }
        if (const char *g = getenv("EVXA_WINDOW_REACH")) {
            const int v = atoi(g);
            if (v >= 1 && v <= 15) ctx->reach = v;
        }
        *out = ctx.release();
    });
}

int evxa_destroy(evxa_context *ctx)
{
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
        ctx->graphs.clear();
        ctx->graph_res.clear();
        for (auto *s : ctx->streams) delete s;
        ctx->streams.clear();
        for (auto &e : ctx->ring) {
            if (e.first) (void)hipEventDestroy(e.first);
            if (e.second) (void)hipEventDestroy(e.second);
        }
        hipStream_t st = ctx->stream;
        delete ctx;
        if (st) (void)hipStreamDestroy(st);
    });
}

int evxa_stream_open(evxa_context *ctx, int32_t width, int32_t height, int32_t n_slots,
                     evxa_stream **out)
{
    return guarded([&] {
        if (!ctx || !out) fail(EVXA_ERROR_INVALIDARG, "null argument");
        *out = nullptr;
        if (width < 16 || height < 16 || width % 16 || height % 16 || width > 16384 ||
            height > 16384)
            fail(EVXA_ERROR_INVALIDARG, "width/height must be multiples of 16 in [16, 16384]");
        if (n_slots < 1 || n_slots > 64) fail(EVXA_ERROR_INVALIDARG, "n_slots must be in [1, 64]");
        std::lock_guard<std::mutex> lk(ctx->mu);
        const uint64_t need = evxa_stream_bytes(width, height, n_slots);
        if (ctx->used + need > ctx->budget)
            fail(EVXA_ERROR_CAPACITY_LIMIT, "stream does not fit the HBM budget");
        set_device(ctx);
        auto s = std::make_unique<evxa_stream>();
        s->ctx = ctx;
        s->w = width;
        s->h = height;
        s->wb = width / 16;
        s->hb = height / 16;
        s->n_slots = n_slots;
        s->gy = plane_geom(width, height);
        s->gc = plane_geom(width / 2, height / 2);
        const size_t pb = picture_bytes(width, height);
        s->slot_stride = align_up(pb, 4096);
        s->pics.alloc(s->slot_stride * n_slots);
        check(hipMemsetAsync(s->pics.p, 0, s->slot_stride * n_slots, ctx->stream), "clear slots");
        const size_t n = (size_t)s->wb * s->hb;
        s->tables.alloc(3 * n * sizeof(evxa_block_record));
        check(hipMemsetAsync(s->tables.p, 0, 3 * n * sizeof(evxa_block_record), ctx->stream),
              "clear tables");
        s->coeffs.alloc(n * EVXA_MB_SAMPLES * sizeof(int16_t));
        s->levels.alloc(n * EVXA_MB_SAMPLES * sizeof(int16_t));
        s->searches.alloc(n * EVXA_MAX_REFS * sizeof(SearchRec));
        check(hipMemsetAsync(s->searches.p, 0, n * EVXA_MAX_REFS * sizeof(SearchRec), ctx->stream),
              "clear search words");
        s->spread.alloc((size_t)n_slots * kRangeWords * sizeof(int32_t));
        check(hipMemsetAsync(s->spread.p, 0, (size_t)n_slots * kRangeWords * sizeof(int32_t),
                             ctx->stream),
              "clear range records");
        s->wide.alloc((size_t)n_slots * 2 * sizeof(int32_t));
        check(hipMemsetAsync(s->wide.p, 0, (size_t)n_slots * 2 * sizeof(int32_t), ctx->stream),
              "clear range words");
        check(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "hipEventCreate");
        mark_done(s.get(), ctx->stream);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        s->bytes = need;
        ctx->used += need;
        ctx->streams.push_back(s.get());
        *out = s.release();
    });
}

int evxa_stream_close(evxa_stream *s)
{
    return guarded([&] {
        if (!s) return;
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->used -= s->bytes;
        ctx->streams.erase(std::remove(ctx->streams.begin(), ctx->streams.end(), s),
                           ctx->streams.end());
        for (auto &g : ctx->graph_streams)
            g.second.erase(std::remove(g.second.begin(), g.second.end(), s), g.second.end());
        if (s->done) (void)hipEventDestroy(s->done);
        delete s;
    });
}

int evxa_stream_reset_temporal(evxa_stream *s)
{
    return guarded([&] {
        if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
        std::lock_guard<std::mutex> lk(s->ctx->mu);
        s->temporal_valid = false;
    });
}

int evxa_upload_picture(evxa_stream *s, int32_t slot, const int16_t *y, const int16_t *u,
                        const int16_t *v, int32_t stride_y, int32_t stride_c)
{
    return guarded([&] {
        if (!s || !y || !u || !v) fail(EVXA_ERROR_INVALIDARG, "null argument");
        check_slot(s, slot);
        if (stride_y < s->w || stride_c < s->w / 2) fail(EVXA_ERROR_INVALIDARG, "stride too small");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        int16_t *base = s->slot_base(slot);
        int16_t *py = base;
        int16_t *pu = base + (size_t)s->gy.pitch * s->gy.rows;
        int16_t *pv = pu + (size_t)s->gc.pitch * s->gc.rows;
        const PlaneGeom *g[3] = {&s->gy, &s->gc, &s->gc};
        int16_t *dst[3] = {py, pu, pv};
        const int16_t *src[3] = {y, u, v};
        const int32_t strides[3] = {stride_y, stride_c, stride_c};
        int32_t *wide = s->range_spread(slot);
        range_begin(s, slot);
        for (int k = 0; k < 3; k++) {
            check(hipMemcpy2DAsync(dst[k] + g[k]->origin(), (size_t)g[k]->pitch * 2, src[k],
                                   (size_t)strides[k] * 2, (size_t)g[k]->w * 2, g[k]->h,
                                   hipMemcpyHostToDevice, ctx->stream),
                  "upload plane");
            check(launch_range(dst[k] + g[k]->origin(), g[k]->pitch, g[k]->w, g[k]->h, wide,
                               ctx->stream),
                  "range kernel");
        }
        range_end(s, slot);
        check(launch_apron3(py + s->gy.origin(), pu + s->gc.origin(), pv + s->gc.origin(),
                            s->gy.pitch, s->gc.pitch, s->w, s->h, ctx->stream),
              "apron kernel");
        // Host buffers are the caller's: finish before returning.
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int evxa_download_picture(evxa_stream *s, int32_t slot, int16_t *y, int16_t *u, int16_t *v,
                          int32_t stride_y, int32_t stride_c)
{
    return guarded([&] {
        if (!s || !y || !u || !v) fail(EVXA_ERROR_INVALIDARG, "null argument");
        check_slot(s, slot);
        if (stride_y < s->w || stride_c < s->w / 2) fail(EVXA_ERROR_INVALIDARG, "stride too small");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        set_device(ctx);
        const DevPic p = s->pic(slot);
        const PlaneGeom *g[3] = {&s->gy, &s->gc, &s->gc};
        const int16_t *src[3] = {p.y, p.u, p.v};
        int16_t *dst[3] = {y, u, v};
        const int32_t strides[3] = {stride_y, stride_c, stride_c};
        for (int k = 0; k < 3; k++)
            check(hipMemcpy2DAsync(dst[k], (size_t)strides[k] * 2, src[k], (size_t)g[k]->pitch * 2,
                                   (size_t)g[k]->w * 2, g[k]->h, hipMemcpyDeviceToHost, ctx->stream),
                  "download plane");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int evxa_analyze_frame(evxa_stream *s, const evxa_frame_desc *desc)
{
    return guarded([&] {
        if (!s || !desc) fail(EVXA_ERROR_INVALIDARG, "null argument");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        set_device(ctx);
        std::vector<DevJob> jobs{make_job(s, *desc, true)};
        enqueue(ctx, ctx->res, jobs, true);
        mark_done(s, ctx->stream);
    });
}

int evxa_encode_frame(evxa_stream *s, const evxa_encode_desc *d)
{
    return guarded([&] {
        if (!s || !d) fail(EVXA_ERROR_INVALIDARG, "null argument");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        encode_launch(ctx, &s, d, 1);
    });
}

int evxa_encode_batch(evxa_context *ctx, evxa_stream *const *streams, const evxa_encode_desc *descs,
                      int32_t n)
{
    return guarded([&] {
        if (!ctx || n < 0 || (n > 0 && (!streams || !descs))) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (int32_t a = 0; a < n; a++) {
            if (!streams[a] || streams[a]->ctx != ctx) fail(EVXA_ERROR_INVALIDARG, "foreign stream");
            for (int32_t b = 0; b < a; b++)
                if (streams[a] == streams[b])
                    fail(EVXA_ERROR_INVALIDARG, "a stream appears twice in one batch");
        }
        if (n > 0) encode_launch(ctx, streams, descs, n);
    });
}

int evxa_analyze_frame_noalloc(evxa_stream *s, const evxa_frame_desc *desc)
{
    return guarded([&] {
        if (!s || !desc) fail(EVXA_ERROR_INVALIDARG, "null argument");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        if (!ctx->scratch) fail(EVXA_ERROR_INVALID_RESOURCE, "no scratch on this device");
        set_device(ctx);
        // Validate and size before the stream's temporal state advances.
        (void)make_job(s, *desc, false);
        const size_t n_mb = (size_t)s->wb * s->hb;
        const NoallocLayout lay = noalloc_layout(n_mb, desc->n_refs);
        if (lay.end > kNoAllocScratchBytes) {
            char msg[160];
            snprintf(msg, sizeof(msg),
                     "frame needs %zu bytes of scratch, the allocation-free path has %zu",
                     lay.end, kNoAllocScratchBytes);
            fail(EVXA_ERROR_CAPACITY_LIMIT, msg);
        }
        DevJob j = make_job(s, *desc, true);
        int32_t *prog = reinterpret_cast<int32_t *>(ctx->scratch + lay.prog);
        DevJob *jp = reinterpret_cast<DevJob *>(ctx->scratch + lay.job);
        uint32_t *work = reinterpret_cast<uint32_t *>(ctx->scratch + lay.work);
        j.mvw = reinterpret_cast<uint64_t *>(prog + kCounterWords);
        j.eager = j.type == EVXA_FRAME_P && j.n_refs > 1;
        // As the wavefront sizes one frame's launch: enough workgroups for
        // the widest anti-diagonal plus slack, at most all that fit.
        int32_t widest = 0;
        for (int32_t v = 0; v < s->wb + 2 * (s->hb - 1); v++) {
            int32_t c = 0;
            for (int32_t by = 0; by < s->hb; by++) c += (v - 2 * by >= 0 && v - 2 * by < s->wb) ? 1 : 0;
            widest = std::max(widest, c * j.n_refs);
        }
        const int32_t fit = ctx->resident[0][ctx->waves];
        const int32_t all = ctx->grid > 0 ? std::min(ctx->grid, fit) : fit;
        const int32_t grid = std::max(1, std::min(all, std::max(8, widest + widest / 2 + 4)));
        hipStream_t st = ctx->stream;
        check(launch_reset(prog, (int32_t)(kCounterWords + 2 * n_mb), st), "reset kernel");
        check(launch_set_job(j, jp, st), "job copy kernel");
        check(launch_wave_list(s->wb, s->hb, j.n_refs, work, st), "work list kernel");
        if (ctx->ring_head - ctx->ring_tail == evxa_context::kRing) ctx->ring_tail++;
        const auto ev = ctx->ring[ctx->ring_head % evxa_context::kRing];
        check(hipEventRecord(ev.first, st), "hipEventRecord");
        check(launch_analyze(jp, work, prog, grid, ctx->waves, st), "analysis kernel");
        check(hipEventRecord(ev.second, st), "hipEventRecord");
        ctx->ring_head++;
        ctx->last_start = ev.first;
        ctx->last_stop = ev.second;
        ctx->scratch_last = true;
        ctx->enc_last = nullptr;
        mark_done(s, st);
    });
}

int evxa_analyze_batch(evxa_context *ctx, evxa_stream *const *streams, const evxa_frame_desc *descs,
                       int32_t n)
{
    return guarded([&] {
        if (!ctx || (n > 0 && (!streams || !descs)) || n < 0) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        for (int32_t a = 0; a < n; a++) {
            if (!streams[a] || streams[a]->ctx != ctx) fail(EVXA_ERROR_INVALIDARG, "foreign stream");
            for (int32_t b = 0; b < a; b++)
                if (streams[a] == streams[b])
                    fail(EVXA_ERROR_INVALIDARG, "a stream appears twice in one batch");
        }
        for (int32_t a = 0; a < n; a++)
            if (streams[a]->inflight.load() > 0)
                fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        std::vector<DevJob> jobs;
        jobs.reserve(n);
        for (int32_t a = 0; a < n; a++) jobs.push_back(make_job(streams[a], descs[a], true));
        enqueue(ctx, ctx->res, jobs, true);
        for (int32_t a = 0; a < n; a++) mark_done(streams[a], ctx->stream);
    });
}

int evxa_synchronize(evxa_context *ctx)
{
    return guarded([&] {
        if (!ctx) fail(EVXA_ERROR_INVALIDARG, "null context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check_last_launch(ctx);
    });
}

int evxa_stream_wait(evxa_stream *s)
{
    return guarded([&] {
        if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
        hipEvent_t ev;
        {
            std::lock_guard<std::mutex> lk(s->ctx->mu);
            set_device(s->ctx);
            ev = s->done;
        }
        // Outside the lock: other threads keep queueing work meanwhile; the
        // wait covers what was queued for this stream before the call.
        check(hipEventSynchronize(ev), "hipEventSynchronize");
    });
}

int evxa_stream_query(evxa_stream *s, int32_t *done)
{
    return guarded([&] {
        if (!s || !done) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(s->ctx->mu);
        set_device(s->ctx);
        const hipError_t e = hipEventQuery(s->done);
        if (e == hipErrorNotReady) {
            *done = 0;
            return;
        }
        check(e, "hipEventQuery");
        *done = 1;
    });
}

int evxa_last_launch_ms(evxa_context *ctx, float *ms)
{
    return guarded([&] {
        if (!ctx || !ms) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (!ctx->last_stop) fail(EVXA_ERROR_INVALID_RESOURCE, "no timed launch yet");
        set_device(ctx);
        check(hipEventSynchronize(ctx->last_stop), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, ctx->last_start, ctx->last_stop), "hipEventElapsedTime");
    });
}

int evxa_launch_times(evxa_context *ctx, float *ms, int32_t max, int32_t *n_out)
{
    return guarded([&] {
        if (!ctx || !n_out || max < 0 || (max > 0 && !ms)) fail(EVXA_ERROR_INVALIDARG, "bad argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        int32_t n = 0;
        while (ctx->ring_tail < ctx->ring_head && n < max) {
            const auto &e = ctx->ring[ctx->ring_tail % evxa_context::kRing];
            check(hipEventElapsedTime(ms + n, e.first, e.second), "hipEventElapsedTime");
            ctx->ring_tail++;
            n++;
        }
        *n_out = n;
    });
}

int evxa_read_results(evxa_stream *s, evxa_block_record *records, int16_t *coeffs, int16_t *levels)
{
    return guarded([&] {
        if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->last_table < 0) fail(EVXA_ERROR_INVALID_RESOURCE, "no frame analysed yet");
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check_last_launch(ctx);
        const size_t n = (size_t)s->wb * s->hb;
        if (records)
            check(hipMemcpy(records, s->table(s->last_table), n * sizeof(evxa_block_record),
                            hipMemcpyDeviceToHost),
                  "read records");
        if (coeffs)
            check(hipMemcpy(coeffs, s->coeffs.p, n * EVXA_MB_SAMPLES * 2, hipMemcpyDeviceToHost),
                  "read coefficients");
        if (levels)
            check(hipMemcpy(levels, s->levels.p, n * EVXA_MB_SAMPLES * 2, hipMemcpyDeviceToHost),
                  "read levels");
    });
}

}  // extern "C"

namespace {

// Events around one kernel of the context's queue, for evxa_launch_times.
template <class F>
void timed(evxa_context *ctx, F &&launch)
{
    if (ctx->ring_head - ctx->ring_tail == evxa_context::kRing) ctx->ring_tail++;
    const auto ev = ctx->ring[ctx->ring_head % evxa_context::kRing];
    check(hipEventRecord(ev.first, ctx->stream), "hipEventRecord");
    launch();
    check(hipEventRecord(ev.second, ctx->stream), "hipEventRecord");
    ctx->ring_head++;
    ctx->last_start = ev.first;
    ctx->last_stop = ev.second;
}

// The packed unfiltered planes the deblocking kernel reads (luma pitch w,
// chroma pitch w / 2), allocated on first use.
void deblock_input(evxa_stream *s, int16_t **planes)
{
    const size_t ny = (size_t)s->w * s->h, nc = ny / 4;
    lazy_alloc(s, s->deblock_src, (ny + 2 * nc) * sizeof(int16_t));
    int16_t *scr = s->deblock_src.as<int16_t>();
    planes[0] = scr;
    planes[1] = scr + ny;
    planes[2] = scr + ny + nc;
}

// evxa_reconstruct, optionally followed by the deblocking filter: then the
// reconstruction goes to the filter's input planes, and the filter writes
// the slot (and its apron).
void reconstruct(evxa_stream *s, int32_t dst_slot, const int16_t *levels, bool filter)
{
    if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
    evxa_context *ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (s->last_table < 0) fail(EVXA_ERROR_INVALID_RESOURCE, "no frame analysed yet");
    if (s->last_encoded)
        fail(EVXA_ERROR_INVALID_RESOURCE,
             "the last frame came from the in-loop encode, which reconstructs it itself");
    if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
    check_slot(s, dst_slot);
    const DevJob &j = s->last_job;
    for (int k = 0; k < j.n_refs; k++) {
        if (s->last_refs[k] == dst_slot)
            fail(EVXA_ERROR_INVALIDARG, "destination slot is a reference of the frame");
        for (int m = 0; m < k; m++)
            if (((j.target_ids >> (8 * k)) & 0xFF) == ((j.target_ids >> (8 * m)) & 0xFF))
                fail(EVXA_ERROR_INVALIDARG, "reconstruction needs distinct target ids");
    }
    set_device(ctx);
    const size_t n = (size_t)s->wb * s->hb;
    const int16_t *lv = j.levels;
    if (levels) {
        lazy_alloc(s, s->recon_levels, n * EVXA_MB_SAMPLES * sizeof(int16_t));
        check(hipMemcpyAsync(s->recon_levels.p, levels, n * EVXA_MB_SAMPLES * sizeof(int16_t),
                             hipMemcpyHostToDevice, ctx->stream),
              "upload levels");
        lv = s->recon_levels.as<int16_t>();
    }
    const DevPic dst = s->pic(dst_slot);
    int32_t *wide = s->range_spread(dst_slot);
    range_begin(s, dst_slot);
    if (!filter) {
        timed(ctx, [&] { check(launch_recon(j, lv, dst, wide, ctx->stream), "reconstruction kernel"); });
        range_end(s, dst_slot);
        // The slot is a picture like an uploaded one: refresh its apron (the
        // range record came with the reconstruction).
        check(launch_apron3(const_cast<int16_t *>(dst.y), const_cast<int16_t *>(dst.u),
                            const_cast<int16_t *>(dst.v), dst.py, dst.pc, s->w, s->h, ctx->stream),
              "apron kernel");
    } else {
        int16_t *src[3];
        deblock_input(s, src);
        lazy_alloc(s, s->recon_flag, kRangeWords * sizeof(int32_t));
        DevPic mid{};
        mid.y = src[0];
        mid.u = src[1];
        mid.v = src[2];
        mid.py = s->w;
        mid.pc = s->w / 2;
        const int32_t sp[3] = {s->w, s->w / 2, s->w / 2}, dp[3] = {dst.py, dst.pc, dst.pc};
        int16_t *const out[3] = {const_cast<int16_t *>(dst.y), const_cast<int16_t *>(dst.u),
                                 const_cast<int16_t *>(dst.v)};
        timed(ctx, [&] {
            check(launch_recon(j, lv, mid, s->recon_flag.as<int32_t>(), ctx->stream),
                  "reconstruction kernel");
        });
        timed(ctx, [&] {
            check(launch_deblock(src, sp, out, dp, s->w, s->h, j.records, wide, ctx->stream),
                  "deblock kernel");
        });
        range_end(s, dst_slot);
    }
    mark_done(s, ctx->stream);
    // The caller's level buffer has been consumed once this returns.
    if (levels) check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" {

int evxa_reconstruct(evxa_stream *s, int32_t dst_slot, const int16_t *levels)
{
    return guarded([&] { reconstruct(s, dst_slot, levels, false); });
}

int evxa_reconstruct_deblocked(evxa_stream *s, int32_t dst_slot, const int16_t *levels)
{
    return guarded([&] { reconstruct(s, dst_slot, levels, true); });
}

int evxa_deblock(evxa_stream *s, int32_t slot, const evxa_block_record *records)
{
    return guarded([&] {
        if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        check_slot(s, slot);
        if (!records && s->last_table < 0) fail(EVXA_ERROR_INVALID_RESOURCE, "no frame analysed yet");
        set_device(ctx);
        const size_t n = (size_t)s->wb * s->hb;
        const evxa_block_record *rec = s->last_job.records;
        if (records) {
            lazy_alloc(s, s->deblock_records, n * sizeof(evxa_block_record));
            check(hipMemcpyAsync(s->deblock_records.p, records, n * sizeof(evxa_block_record),
                                 hipMemcpyHostToDevice, ctx->stream),
                  "upload records");
            rec = s->deblock_records.as<evxa_block_record>();
        }
        // The filter reads the unfiltered picture around each tile, so it
        // runs from a packed copy back into the slot (and its apron).
        int16_t *src[3];
        deblock_input(s, src);
        const DevPic p = s->pic(slot);
        int16_t *const dst[3] = {const_cast<int16_t *>(p.y), const_cast<int16_t *>(p.u),
                                 const_cast<int16_t *>(p.v)};
        const int32_t sp[3] = {s->w, s->w / 2, s->w / 2}, dp[3] = {p.py, p.pc, p.pc};
        int32_t *wide = s->range_spread(slot);
        check(launch_copy3(p.y, p.u, p.v, p.py, p.pc, src[0], s->w, s->h, wide, ctx->stream),
              "copy kernel");
        range_begin(s, slot);
        timed(ctx, [&] {
            check(launch_deblock(src, sp, dst, dp, s->w, s->h, rec, wide, ctx->stream),
                  "deblock kernel");
        });
        range_end(s, slot);
        mark_done(s, ctx->stream);
        // The caller's records have been consumed once this returns.
        if (records) check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

}  // extern "C"

namespace {

constexpr int kSaoCounters = 64;

// The SAO buffers of a stream, allocated on first use: sao_words holds the
// decided words (one per MB), then kSaoCounters counters, then 8 words per
// MB of per-stage scratch.
void sao_buffers(evxa_stream *s)
{
    const size_t n = (size_t)s->wb * s->hb;
    lazy_alloc(s, s->sao_words, n * sizeof(uint64_t) + kSaoCounters * sizeof(int32_t) +
                                    8 * n * sizeof(uint32_t));
    lazy_alloc(s, s->sao_table, n * sizeof(evxa_sao_info));
}

int32_t *sao_counters(evxa_stream *s)
{
    return reinterpret_cast<int32_t *>(s->sao_words.as<uint64_t>() + (size_t)s->wb * s->hb);
}

// The unfiltered copy of `slot` SAO reads, then SAO per the stream's SAO
// table into the slot and its apron; the error flags are read back.
void sao_apply_slot(evxa_stream *s, int32_t slot, int16_t *const rec[3])
{
    evxa_context *ctx = s->ctx;
    const DevPic p = s->pic(slot);
    int16_t *const dst[3] = {const_cast<int16_t *>(p.y), const_cast<int16_t *>(p.u),
                             const_cast<int16_t *>(p.v)};
    const int32_t dp[3] = {p.py, p.pc, p.pc};
    int32_t *counters = sao_counters(s);
    range_begin(s, slot);
    timed(ctx, [&] {
        check(launch_sao_apply(rec, dst, dp, s->w, s->h, s->sao_table.as<evxa_sao_info>(),
                               s->range_spread(slot), counters + 1, ctx->stream),
              "SAO apply kernel");
    });
    range_end(s, slot);
    check(launch_apron3(dst[0], dst[1], dst[2], p.py, p.pc, s->w, s->h, ctx->stream), "apron kernel");
    int32_t err = 0;
    check(hipMemcpyAsync(&err, counters + 1, sizeof(err), hipMemcpyDeviceToHost, ctx->stream),
          "read SAO flags");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (err & 1) fail(EVXA_ERROR_OPERATION_TIMEDOUT, "SAO dependency wait timed out");
    if (err) fail(EVXA_ERROR_EXECUTION_FAILURE, "SAO chain deeper than three blocks");
}

}  // namespace

extern "C" {

int evxa_sao(evxa_stream *s, int32_t slot, int32_t src_slot, int32_t quality, evxa_sao_info *table)
{
    return guarded([&] {
        if (!s) fail(EVXA_ERROR_INVALIDARG, "null stream");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        check_slot(s, slot);
        check_slot(s, src_slot);
        if (slot == src_slot) fail(EVXA_ERROR_INVALIDARG, "the source and the reconstruction share a slot");
        set_device(ctx);
        int16_t *rec[3];
        deblock_input(s, rec);
        sao_buffers(s);
        const size_t n = (size_t)s->wb * s->hb;
        const DevPic p = s->pic(slot);
        // The decisions and the filter read the unfiltered reconstruction.
        check(launch_copy3(p.y, p.u, p.v, p.py, p.pc, rec[0], s->w, s->h,
                           s->range_spread(slot), ctx->stream),
              "copy kernel");
        check(launch_reset(reinterpret_cast<int32_t *>(s->sao_words.p), (int32_t)(2 * n) + kSaoCounters,
                           ctx->stream),
              "reset SAO words");
        timed(ctx, [&] {
            check(launch_sao_decide(rec[0], rec[1], s->pic(src_slot), s->w, s->h, lambda_of(quality),
                                    reinterpret_cast<uint32_t *>(sao_counters(s) + kSaoCounters),
                                    s->sao_words.as<uint64_t>(), s->sao_table.as<evxa_sao_info>(),
                                    sao_counters(s), ctx->stream),
                  "SAO decision kernels");
        });
        sao_apply_slot(s, slot, rec);
        mark_done(s, ctx->stream);
        if (table)
            check(hipMemcpy(table, s->sao_table.p, n * sizeof(evxa_sao_info), hipMemcpyDeviceToHost),
                  "download SAO table");
    });
}

int evxa_sao_filter(evxa_stream *s, int32_t slot, const evxa_sao_info *table)
{
    return guarded([&] {
        if (!s || !table) fail(EVXA_ERROR_INVALIDARG, "null argument");
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        check_slot(s, slot);
        set_device(ctx);
        int16_t *rec[3];
        deblock_input(s, rec);
        sao_buffers(s);
        const size_t n = (size_t)s->wb * s->hb;
        check(hipMemcpyAsync(s->sao_table.p, table, n * sizeof(evxa_sao_info), hipMemcpyHostToDevice,
                             ctx->stream),
              "upload SAO table");
        check(launch_reset(sao_counters(s), kSaoCounters, ctx->stream), "reset SAO flags");
        const DevPic p = s->pic(slot);
        check(launch_copy3(p.y, p.u, p.v, p.py, p.pc, rec[0], s->w, s->h,
                           s->range_spread(slot), ctx->stream),
              "copy kernel");
        sao_apply_slot(s, slot, rec);
        mark_done(s, ctx->stream);
    });
}

int evxa_graph_create(evxa_context *ctx, evxa_stream *const *streams, const evxa_frame_desc *descs,
                      int32_t n, int32_t *graph_id)
{
    return guarded([&] {
        if (!ctx || !streams || !descs || !graph_id || n <= 0) fail(EVXA_ERROR_INVALIDARG, "bad argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        std::vector<DevJob> jobs;
        for (int32_t a = 0; a < n; a++) {
            if (!streams[a] || streams[a]->ctx != ctx) fail(EVXA_ERROR_INVALIDARG, "foreign stream");
            jobs.push_back(make_job(streams[a], descs[a], false));
            streams[a]->last_table = descs[a].type == EVXA_FRAME_P ? streams[a]->p_cur : 2;
        }
        auto res = std::make_unique<LaunchRes>();
        // Stage the work list and job table outside capture (plain copies),
        // then capture only the reset + analysis kernels.
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        enqueue(ctx, *res, jobs, false, nullptr, 0, true, nullptr, nullptr, false);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        hipGraph_t graph = nullptr;
        check(hipEventCreate(&res->g_start), "hipEventCreate");
        check(hipEventCreate(&res->g_stop), "hipEventCreate");
        check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        const hipError_t e1 = launch_staged(*res, ctx->stream, res->g_start, res->g_stop);
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
        check(e1, "capture analysis");
        check(e2, "end capture");
        hipGraphExec_t exec = nullptr;
        hipError_t e6 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e6, "hipGraphInstantiate");
        const int32_t id = ctx->next_graph++;
        ctx->graphs[id] = exec;
        ctx->graph_res[id] = std::move(res);
        ctx->graph_streams[id].assign(streams, streams + n);
        *graph_id = id;
    });
}

int evxa_graph_launch(evxa_context *ctx, int32_t graph_id)
{
    return guarded([&] {
        if (!ctx) fail(EVXA_ERROR_INVALIDARG, "null context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->graphs.find(graph_id);
        if (it == ctx->graphs.end()) fail(EVXA_ERROR_INVALIDARG, "no such graph");
        set_device(ctx);
        check(hipGraphLaunch(it->second, ctx->stream), "hipGraphLaunch");
        for (evxa_stream *gs : ctx->graph_streams[graph_id]) mark_done(gs, ctx->stream);
        const LaunchRes &r = *ctx->graph_res[graph_id];
        ctx->last_start = r.g_start;
        ctx->last_stop = r.g_stop;
    });
}

int evxa_graph_destroy(evxa_context *ctx, int32_t graph_id)
{
    return guarded([&] {
        if (!ctx) fail(EVXA_ERROR_INVALIDARG, "null context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->graphs.find(graph_id);
        if (it == ctx->graphs.end()) fail(EVXA_ERROR_INVALIDARG, "no such graph");
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        (void)hipGraphExecDestroy(it->second);
        ctx->graphs.erase(it);
        ctx->graph_res.erase(graph_id);
        ctx->graph_streams.erase(graph_id);
    });
}

int evxa_interp_blocks(evxa_stream *s, int32_t slot, int32_t n, const int32_t *xs, const int32_t *ys,
                       const int32_t *dir_x, const int32_t *dir_y, const int32_t *quarter, int16_t *out)
{
    return guarded([&] {
        if (!s || n < 0 || (n > 0 && (!xs || !ys || !dir_x || !dir_y || !quarter || !out)))
            fail(EVXA_ERROR_INVALIDARG, "bad argument");
        if (n == 0) return;
        check_slot(s, slot);
        std::vector<int32_t> p((size_t)n * 5);
        for (int32_t k = 0; k < n; k++) {
            if (xs[k] < 0 || xs[k] > s->w - 16 || ys[k] < 0 || ys[k] > s->h - 16)
                fail(EVXA_ERROR_INVALIDARG, "block base outside the picture");
            if (dir_x[k] < -1 || dir_x[k] > 1 || dir_y[k] < -1 || dir_y[k] > 1)
                fail(EVXA_ERROR_INVALIDARG, "direction components must be -1, 0 or 1");
            p[k * 5] = xs[k]; p[k * 5 + 1] = ys[k]; p[k * 5 + 2] = dir_x[k];
            p[k * 5 + 3] = dir_y[k]; p[k * 5 + 4] = quarter[k] ? 1 : 0;
        }
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        DevBuf dp(p.size() * 4), dout((size_t)n * EVXA_MB_SAMPLES * 2);
        check(hipMemcpy(dp.p, p.data(), p.size() * 4, hipMemcpyHostToDevice), "copy params");
        check(launch_interp_blocks(s->pic(slot), s->w, s->h, n, dp.as<int32_t>(), dout.as<int16_t>(),
                                   ctx->stream),
              "interp kernel");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check(hipMemcpy(out, dout.p, (size_t)n * EVXA_MB_SAMPLES * 2, hipMemcpyDeviceToHost), "copy out");
    });
}

int evxa_search_blocks(evxa_stream *s, int32_t src_slot, int32_t ref_slot, int32_t quality, int32_t n,
                       const int32_t *xs, const int32_t *ys, const int16_t *pred_x,
                       const int16_t *pred_y, evxa_search_result *out)
{
    return guarded([&] {
        if (!s || n < 0 || (n > 0 && (!xs || !ys || !pred_x || !pred_y || !out)))
            fail(EVXA_ERROR_INVALIDARG, "bad argument");
        if (quality < 0 || quality > 63) fail(EVXA_ERROR_INVALIDARG, "quality must be 0..63");
        if (n == 0) return;
        check_slot(s, src_slot);
        check_slot(s, ref_slot);
        std::vector<int32_t> p((size_t)n * 4);
        for (int32_t k = 0; k < n; k++) {
            if (xs[k] < 0 || xs[k] > s->w - 16 || ys[k] < 0 || ys[k] > s->h - 16)
                fail(EVXA_ERROR_INVALIDARG, "MB position outside the picture");
            p[k * 4] = xs[k]; p[k * 4 + 1] = ys[k]; p[k * 4 + 2] = pred_x[k]; p[k * 4 + 3] = pred_y[k];
        }
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        DevBuf dp(p.size() * 4), dout((size_t)n * sizeof(evxa_search_result));
        check(hipMemcpy(dp.p, p.data(), p.size() * 4, hipMemcpyHostToDevice), "copy params");
        check(launch_search_blocks(s->pic(src_slot), s->pic(ref_slot), s->w, s->h, quality,
                                   lambda_of(quality), n, dp.as<int32_t>(),
                                   dout.as<evxa_search_result>(), ctx->stream),
              "search kernel");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check(hipMemcpy(out, dout.p, (size_t)n * sizeof(evxa_search_result), hipMemcpyDeviceToHost),
              "copy out");
    });
}

int evxa_intra_search_blocks(evxa_stream *s, int32_t src_slot, int32_t pic_slot, int32_t quality,
                             int32_t n, const int32_t *xs, const int32_t *ys,
                             evxa_search_result *out)
{
    return guarded([&] {
        if (!s || n < 0 || (n > 0 && (!xs || !ys || !out))) fail(EVXA_ERROR_INVALIDARG, "bad argument");
        if (quality < 0 || quality > 63) fail(EVXA_ERROR_INVALIDARG, "quality must be 0..63");
        if (n == 0) return;
        check_slot(s, src_slot);
        check_slot(s, pic_slot);
        std::vector<int32_t> p((size_t)n * 2);
        for (int32_t k = 0; k < n; k++) {
            if (xs[k] < 0 || xs[k] > s->w - 16 || ys[k] < 0 || ys[k] > s->h - 16)
                fail(EVXA_ERROR_INVALIDARG, "MB position outside the picture");
            p[k * 2] = xs[k];
            p[k * 2 + 1] = ys[k];
        }
        evxa_context *ctx = s->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (s->inflight.load() > 0) fail(EVXA_ERROR_INVALID_RESOURCE, "stream has a frame in flight");
        set_device(ctx);
        DevBuf dp(p.size() * 4), dout((size_t)n * sizeof(evxa_search_result));
        check(hipMemcpy(dp.p, p.data(), p.size() * 4, hipMemcpyHostToDevice), "copy params");
        check(launch_intra_blocks(s->pic(src_slot), s->pic(pic_slot), s->w, s->h, quality,
                                  lambda_of(quality), n, dp.as<int32_t>(),
                                  dout.as<evxa_search_result>(), ctx->stream),
              "intra search kernel");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check(hipMemcpy(out, dout.p, (size_t)n * sizeof(evxa_search_result), hipMemcpyDeviceToHost),
              "copy out");
    });
}

int evxa_dct_quant_blocks(evxa_context *ctx, int32_t n, const int16_t *src, const int16_t *pred,
                          const int32_t *quality, int16_t *coeffs, int16_t *levels, int32_t *qp,
                          int32_t *variance)
{
    return guarded([&] {
        if (!ctx || n < 0 || (n > 0 && (!src || !pred || !quality || !coeffs || !levels || !qp || !variance)))
            fail(EVXA_ERROR_INVALIDARG, "bad argument");
        if (n == 0) return;
        for (int32_t k = 0; k < n; k++)
            if (quality[k] < 0 || quality[k] > 255) fail(EVXA_ERROR_INVALIDARG, "quality");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        const size_t sb = (size_t)n * EVXA_MB_SAMPLES * 2;
        DevBuf ds(sb), dp(sb), dq((size_t)n * 4), dc(sb), dl(sb), dqp((size_t)n * 4), dv((size_t)n * 4);
        check(hipMemcpy(ds.p, src, sb, hipMemcpyHostToDevice), "copy src");
        check(hipMemcpy(dp.p, pred, sb, hipMemcpyHostToDevice), "copy pred");
        check(hipMemcpy(dq.p, quality, (size_t)n * 4, hipMemcpyHostToDevice), "copy quality");
        check(launch_dct_quant_blocks(n, ds.as<int16_t>(), dp.as<int16_t>(), dq.as<int32_t>(),
                                      dc.as<int16_t>(), dl.as<int16_t>(), dqp.as<int32_t>(),
                                      dv.as<int32_t>(), ctx->stream),
              "dct/quant kernel");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check(hipMemcpy(coeffs, dc.p, sb, hipMemcpyDeviceToHost), "copy coefficients");
        check(hipMemcpy(levels, dl.p, sb, hipMemcpyDeviceToHost), "copy levels");
        check(hipMemcpy(qp, dqp.p, (size_t)n * 4, hipMemcpyDeviceToHost), "copy qp");
        check(hipMemcpy(variance, dv.p, (size_t)n * 4, hipMemcpyDeviceToHost), "copy variance");
    });
}

}  // extern "C"

// ---- serving: continuous batching over concurrent launches ----------------
//
// A server owns a dispatcher thread and `max_inflight` launch slots, each with
// its own HIP stream, job/work buffers and events.  Submitted frames queue in
// arrival order; whenever a slot is free, the dispatcher fuses the queued
// frames whose stream has no frame in flight and whose size equals the oldest
// such frame's (up to `max_frames`; a launch completes as a whole) into one
// persistent launch on that slot, sized to its share of the device, so
// launches overlap on the GPU and frames that arrive while others run start
// in the next free slot.  A stream has at most one frame in flight: the next
// P frame of a stream reads the previous one's MV table, and a stream's
// outputs are overwritten by its next frame.  Completion is detected by
// polling the slot's event; each frame then yields a completion record.

namespace {

double clock_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct evxa_server {
    evxa_context *ctx = nullptr;
    int32_t max_inflight = 2;
    int32_t max_frames = 64;
    std::thread thread;
    std::mutex mu;                       // guards everything below
    std::condition_variable cv;
    bool stop = false;
    struct Req {
        evxa_stream *s;
        evxa_frame_desc d;
        uint64_t tag;
        double t_arrival;
    };
    struct Slot {
        hipStream_t st = nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        hipEvent_t dep = nullptr;        // work queued on the context's stream before a launch
        LaunchRes res;
        PinnedBuf flags;                 // error word of the launch, copied back
        std::vector<Req> reqs;
        bool busy = false;
        double t_launch = 0.0;
    };
    std::deque<Req> pending;
    std::vector<std::unique_ptr<Slot>> slots;
    std::set<evxa_stream *> active;
    std::deque<evxa_completion> done;
    int32_t error = 0;
    std::string error_msg;
    uint64_t launches = 0, frames = 0;
};

namespace {

void server_fail(evxa_server *srv, int code, const std::string &msg)
{
    if (!srv->error) {
        srv->error = code;
        srv->error_msg = msg;
    }
}

// Launch `reqs` on `sl` (called with srv->mu held; takes ctx->mu).
void server_launch(evxa_server *srv, evxa_server::Slot &sl, std::vector<evxa_server::Req> &reqs)
{
    evxa_context *ctx = srv->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    set_device(ctx);
    std::vector<DevJob> jobs;
    jobs.reserve(reqs.size());
    for (auto &r : reqs) jobs.push_back(make_job(r.s, r.d, true));
    // Calls that return before their work is done (uploads aside: analysis,
    // reconstruction, deblocking, SAO) queue it on the context's stream; the
    // launch reads their pictures, range words and temporal tables, so it
    // waits for everything queued there so far.
    check(hipEventRecord(sl.dep, ctx->stream), "hipEventRecord");
    check(hipStreamWaitEvent(sl.st, sl.dep, 0), "hipStreamWaitEvent");
    // Each launch is sized by its own work (the widest diagonal of its
    // lists), capped at half the device: launches of chain-bound frames need
    // few workgroups, and several in flight may share the device.
    const int32_t fit = ctx->resident[ctx->mode == EVXA_MODE_SWEEP ? 1 : 0][ctx->waves];
    const int32_t all = ctx->grid > 0 ? std::min(ctx->grid, fit) : fit;
    const int32_t cap = std::max(1, all / std::min(srv->max_inflight, 2));
    enqueue(ctx, sl.res, jobs, false, sl.st, cap, false, sl.t0, nullptr);
    check(hipMemcpyAsync(sl.flags.p, sl.res.progress.p, (kCounterError + 1) * sizeof(int32_t),
                         hipMemcpyDeviceToHost, sl.st),
          "copy launch flags");
    check(hipEventRecord(sl.t1, sl.st), "hipEventRecord");
    for (auto &r : reqs) mark_done(r.s, sl.st);
    sl.reqs = std::move(reqs);
    sl.busy = true;
    sl.t_launch = clock_s();
    srv->launches++;
}

void server_complete(evxa_server *srv, evxa_server::Slot &sl)
{
    const double now = clock_s();
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, sl.t0, sl.t1);
    const int32_t flags = static_cast<const int32_t *>(sl.flags.p)[kCounterError];
    int32_t status = EVXA_SUCCESS;
    if (flags & kErrWaitTimeout) status = EVXA_ERROR_OPERATION_TIMEDOUT;
    else if (flags & kErrSearchFailed) status = EVXA_ERROR_EXECUTION_FAILURE;
    for (auto &r : sl.reqs) {
        evxa_completion c;
        memset(&c, 0, sizeof(c));
        c.tag = r.tag;
        c.status = status;
        c.batch_frames = (int32_t)sl.reqs.size();
        c.t_arrival = r.t_arrival;
        c.t_launch = sl.t_launch;
        c.t_done = now;
        c.device_ms = ms;
        srv->done.push_back(c);
        srv->active.erase(r.s);
        r.s->inflight--;
        srv->frames++;
    }
    sl.reqs.clear();
    sl.busy = false;
}

void server_run(evxa_server *srv)
{
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        bool any_busy = false;
        for (auto &sl : srv->slots) {
            if (!sl->busy) continue;
            const hipError_t q = hipEventQuery(sl->t1);
            if (q == hipSuccess) {
                server_complete(srv, *sl);
            } else if (q == hipErrorNotReady) {
                any_busy = true;
            } else {
                server_fail(srv, EVXA_ERROR_EXECUTION_FAILURE,
                            std::string("analysis launch: ") + hipGetErrorString(q));
                (void)hipStreamSynchronize(sl->st);
                server_complete(srv, *sl);
            }
        }
        evxa_server::Slot *free_slot = nullptr;
        for (auto &sl : srv->slots)
            if (!sl->busy) { free_slot = sl.get(); break; }
        if (free_slot && !srv->pending.empty()) {
            // A launch completes as a whole, so it fuses frames of one size
            // class (MBs per frame): the class of the oldest ready frame.
            std::vector<evxa_server::Req> pick;
            std::set<evxa_stream *> taken;
            int64_t klass = -1;
            for (auto it = srv->pending.begin();
                 it != srv->pending.end() && (int32_t)pick.size() < srv->max_frames;) {
                const int64_t k = (int64_t)it->s->wb * it->s->hb;
                if (srv->active.count(it->s) || taken.count(it->s) || (klass >= 0 && k != klass)) {
                    ++it;
                    continue;
                }
                klass = k;
                taken.insert(it->s);
                pick.push_back(*it);
                it = srv->pending.erase(it);
            }
            if (!pick.empty()) {
                for (auto &r : pick) srv->active.insert(r.s);
                try {
                    server_launch(srv, *free_slot, pick);
                } catch (const Failure &e) {
                    server_fail(srv, e.code, e.msg);
                    for (auto &r : pick) {
                        evxa_completion c;
                        memset(&c, 0, sizeof(c));
                        c.tag = r.tag;
                        c.status = e.code;
                        c.t_arrival = r.t_arrival;
                        c.t_launch = c.t_done = clock_s();
                        srv->done.push_back(c);
                        srv->active.erase(r.s);
                        r.s->inflight--;
                    }
                }
                continue;
            }
        }
        if (srv->stop && srv->pending.empty() && !any_busy) break;
        if (any_busy) srv->cv.wait_for(lk, std::chrono::microseconds(20));
        else srv->cv.wait(lk);
    }
}

}  // namespace

extern "C" {

double evxa_clock(void) { return clock_s(); }

int evxa_set_tuning(evxa_context *ctx, const evxa_tuning *t)
{
    return guarded([&] {
        if (!ctx || !t) fail(EVXA_ERROR_INVALIDARG, "null argument");
        if (t->waves_per_search < 0 || t->waves_per_search > 2 || t->max_groups < 0 ||
            t->lists < 0 || t->lists > kLists || t->cap_pct < 0 || t->cap_pct > 10000 ||
            (t->mode != 0 && t->mode != EVXA_MODE_WAVEFRONT && t->mode != EVXA_MODE_SWEEP) ||
            t->sweep_reads < 0 || t->sweep_reads > 2 || t->sweep_wait_us < -1 ||
            t->sweep_wait_us > 100000 || t->sweep_order < 0 || t->sweep_order > 2 ||
            t->claim_batch < 0 || t->claim_batch > kMaxClaimBatch || t->window_reach < 0 ||
            t->window_reach > 15)
            fail(EVXA_ERROR_INVALIDARG, "tuning out of range");
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (t->waves_per_search) ctx->waves = t->waves_per_search;
        if (t->mode) ctx->mode = t->mode;
        if (t->sweep_reads) ctx->sweep_reads = t->sweep_reads;
        if (t->sweep_wait_us >= 0) ctx->sweep_wait_us = t->sweep_wait_us;
        if (t->sweep_order) ctx->sweep_order = t->sweep_order;
        if (t->claim_batch) ctx->claim_batch = t->claim_batch;
        if (t->window_reach) ctx->reach = t->window_reach;
        ctx->grid = t->max_groups;
        ctx->lists = t->lists;
        ctx->cap_pct = t->cap_pct ? t->cap_pct : 60;
    });
}

int evxa_get_tuning(evxa_context *ctx, evxa_tuning *t, int32_t *resident_groups)
{
    return guarded([&] {
        if (!ctx || !t) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        t->waves_per_search = ctx->waves;
        t->max_groups = ctx->grid;
        t->lists = ctx->lists;
        t->cap_pct = ctx->cap_pct;
        t->mode = ctx->mode;
        t->sweep_reads = ctx->sweep_reads;
        t->sweep_wait_us = ctx->sweep_wait_us;
        t->sweep_order = ctx->sweep_order;
        t->claim_batch = ctx->claim_batch;
        t->window_reach = ctx->reach;
        if (resident_groups) *resident_groups = ctx->resident[ctx->mode == EVXA_MODE_SWEEP ? 1 : 0][ctx->waves];
    });
}

int evxa_sweep_stats(evxa_context *ctx, int32_t *listed, int32_t *searched, float *done_us,
                     int32_t max, int32_t *n_sweeps)
{
    return guarded([&] {
        if (!ctx || !n_sweeps || (max > 0 && (!listed || !searched)))
            fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        *n_sweeps = 0;
        const LaunchRes &r = ctx->res;
        if (!r.sweep || !r.progress.p) return;
        std::vector<int32_t> blk((size_t)r.max_sweeps * kSweepStride);
        check(hipMemcpy(blk.data(), r.progress.as<int32_t>() + r.sctr_off, blk.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost),
              "read sweep counters");
        // Sweeps 0 and 1 take every MB; a later sweep ran iff MBs were listed
        // for it (the block after the last sweep was cleared before that
        // sweep and stays empty).
        int32_t n = 0;
        uint64_t t0;
        memcpy(&t0, blk.data() + kSwStart, 8);
        for (int32_t s = 0; s < r.max_sweeps; s++) {
            const int32_t *b = blk.data() + (size_t)s * kSweepStride;
            int32_t l = 0;
            if (s <= 1) l = r.cap_check;
            else
                for (int k = 0; k < kSubLists; k++) l += b[k * kSubStride + kSubCheckTail];
            if (s > 1 && l == 0) break;
            if (n < max) {
                listed[n] = l;
                searched[n] = s == 0 ? r.cap_check : b[kSwSearched];   // sweep 0 searches every MB
                if (done_us) {
                    uint64_t t1;
                    memcpy(&t1, b + kSwTime, 8);
                    // A sweep whose checks changed nothing ends without searches
                    // (and without an end time of its own).
                    done_us[n] = b[kSwFlag] ? (float)((double)(int64_t)(t1 - t0) / 100.0) : -1.0f;
                }
            }
            n++;
        }
        *n_sweeps = n;
    });
}

int evxa_server_create(evxa_context *ctx, int32_t max_inflight, int32_t max_frames,
                       evxa_server **out)
{
    return guarded([&] {
        if (!ctx || !out) fail(EVXA_ERROR_INVALIDARG, "null argument");
        *out = nullptr;
        if (max_inflight < 1 || max_inflight > 16) fail(EVXA_ERROR_INVALIDARG, "max_inflight must be 1..16");
        if (max_frames < 1 || max_frames > kMaxJobs) fail(EVXA_ERROR_INVALIDARG, "max_frames out of range");
        auto srv = std::make_unique<evxa_server>();
        srv->ctx = ctx;
        srv->max_inflight = max_inflight;
        srv->max_frames = max_frames;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            set_device(ctx);
            for (int32_t k = 0; k < max_inflight; k++) {
                auto sl = std::make_unique<evxa_server::Slot>();
                check(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking), "hipStreamCreate");
                check(hipEventCreate(&sl->t0), "hipEventCreate");
                check(hipEventCreate(&sl->t1), "hipEventCreate");
                check(hipEventCreateWithFlags(&sl->dep, hipEventDisableTiming), "hipEventCreate");
                sl->flags.ensure(64);
                srv->slots.push_back(std::move(sl));
            }
        }
        evxa_server *raw = srv.get();
        raw->thread = std::thread(server_run, raw);
        *out = srv.release();
    });
}

int evxa_server_submit(evxa_server *srv, evxa_stream *s, const evxa_frame_desc *desc, uint64_t tag,
                       double t_arrival)
{
    return guarded([&] {
        if (!srv || !s || !desc) fail(EVXA_ERROR_INVALIDARG, "null argument");
        if (s->ctx != srv->ctx) fail(EVXA_ERROR_INVALIDARG, "foreign stream");
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            if (srv->stop) fail(EVXA_ERROR_INVALID_RESOURCE, "server is stopping");
            srv->pending.push_back({s, *desc, tag, t_arrival > 0.0 ? t_arrival : clock_s()});
            s->inflight++;
        }
        srv->cv.notify_one();
    });
}

int evxa_server_poll(evxa_server *srv, evxa_completion *out, int32_t max, int32_t *n_out)
{
    return guarded([&] {
        if (!srv || !n_out || (max > 0 && !out)) fail(EVXA_ERROR_INVALIDARG, "null argument");
        std::lock_guard<std::mutex> lk(srv->mu);
        int32_t n = 0;
        while (n < max && !srv->done.empty()) {
            out[n++] = srv->done.front();
            srv->done.pop_front();
        }
        *n_out = n;
        if (srv->error) fail(srv->error, srv->error_msg);
    });
}

int evxa_server_pending(evxa_server *srv, int32_t *queued, int32_t *inflight)
{
    return guarded([&] {
        if (!srv) fail(EVXA_ERROR_INVALIDARG, "null server");
        std::lock_guard<std::mutex> lk(srv->mu);
        if (queued) *queued = (int32_t)srv->pending.size();
        if (inflight) *inflight = (int32_t)srv->active.size();
    });
}

int evxa_server_destroy(evxa_server *srv)
{
    return guarded([&] {
        if (!srv) return;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv.notify_one();
        if (srv->thread.joinable()) srv->thread.join();
        for (auto &sl : srv->slots) {
            if (sl->st) (void)hipStreamSynchronize(sl->st);
            if (sl->t0) (void)hipEventDestroy(sl->t0);
            if (sl->t1) (void)hipEventDestroy(sl->t1);
            if (sl->dep) (void)hipEventDestroy(sl->dep);
            if (sl->st) (void)hipStreamDestroy(sl->st);
        }
        delete srv;
    });
}

}  // extern "C"
