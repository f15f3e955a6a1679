// Serving with continuous batching: spectroscopy jobs (LTS + process pairs)
// arrive at any time.  Builder threads (each with its own stream) take every
// queued job and build their games together, in one level loop over the
// disjoint union of the jobs' LTSs; each job's game comes out as its own
// contiguous, locally numbered range.  At a round boundary built jobs are
// placed into shared HBM pools and seeded into the shared device worklist, so
// every round evaluates the frontier of all resident jobs together.  A job
// retires as soon as none of its positions is on the worklist; its starting
// positions' budgets are gathered and its pool ranges are reused.  Jobs are
// disjoint subgraphs and every round is evaluated per node, so a job's result
// is the same alone and in the mix.
//
// There is no counterpart in the reference, which solves one game per
// TransitionSystem::compare / equivalences call (src/lib.rs:275-343).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "gamebuild.hpp"
#include "runtime.hpp"

using namespace geq;

namespace {

// per-node bytes of the pools (row, rrow, atk, node record, wl x2, chg, job_of, queued bits rounded up) and
// per-edge bytes (edges, rcol)
constexpr uint64_t NODE_BYTES = 4 + 4 + 1 + 16 * NODE_U4 + 4 + 4 + 1 + 4 + 1;
constexpr uint64_t EDGE_BYTES = 8 + 4;

// Place a built game at node base nb / edge base eb of the pools.  Its
// slots start empty, tagged so that slot 0 is current from round tag t0 on.
__global__ void k_place_nodes(const uint32_t* row, const uint8_t* atk, uint32_t V, uint32_t nb, uint32_t eb, uint32_t job,
                              uint32_t t0, uint32_t* prow, uint8_t* patk, uint4* st, uint32_t* job_of) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > V) return;
    prow[nb + v] = eb + row[v];             // v == V: the row end of the job (padding slot)
    patk[nb + v] = v < V ? atk[v] : (uint8_t)1;
    st[slot_index(nb + v, 0)] = make_uint4(0, 0, 0, ((t0 - 1u) & 0xffffu) << 16);
    st[slot_index(nb + v, 1)] = make_uint4(0, 0, 0, ((t0 - 2u) & 0xffffu) << 16);
    job_of[nb + v] = job;
}
// a released range holds no lists (the arena compaction walks every slot)
__global__ void k_clear_nodes(uint4* st, uint32_t nb, uint32_t n) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        st[slot_index(nb + v, 0)] = make_uint4(0, 0, 0, 0);
        st[slot_index(nb + v, 1)] = make_uint4(0, 0, 0, 0);
    }
}
// arena compaction (t: tag of the next round): arena entries each current
// list needs, then the move to the packed offsets.  The other slot of a node
// is emptied: it is never current again before it is rewritten, and holds no
// offsets into the old arena.
__global__ void k_arena_need(const uint4* st, uint32_t n, uint32_t t, uint32_t* need) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    const uint32_t len = v < n ? nl_len(load_cur(st, v, t)) : 0u;
    need[v] = len > INLINE_CAP ? len : 0u;
}
__global__ void k_arena_move(uint4* st, const uint16_t* src, uint16_t* dst, const uint32_t* offs, uint32_t n, uint32_t t) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t c = 0;
    uint4 d = load_cur(st, v, t, c);
    uint4 o = st[slot_index(v, c ^ 1u)];
    o.x &= 0xffff0000u;                     // length 0
    st[slot_index(v, c ^ 1u)] = o;
    const uint32_t len = nl_len(d);
    if (len <= INLINE_CAP) return;
    const uint32_t from = nl_off(d), to = offs[v];
    for (uint32_t j = 0; j < len; j++) dst[to + j] = src[from + j];
    d.y = to;
    st[slot_index(v, c)] = d;
}
__global__ void k_place_edges(const uint32_t* col, const uint32_t* wt, uint32_t E, uint32_t nb, uint32_t eb, uint2* edges) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) edges[eb + e] = make_uint2(nb + col[e], wt[e]);
}
__global__ void k_place_indeg(const uint32_t* col, uint32_t E, uint32_t* cnt) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) atomicAdd(&cnt[col[e]], 1u);
}
__global__ void k_place_rrow(const uint32_t* excl, uint32_t V, uint32_t eb, uint32_t nb, uint32_t* prrow, uint32_t* fill) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > V) return;
    prrow[nb + v] = eb + excl[v];
    if (v < V) fill[v] = excl[v];
}
__global__ void k_place_rcol(const uint32_t* row, const uint32_t* col, uint32_t V, uint32_t nb, uint32_t eb, uint32_t* fill,
                             uint32_t* prcol) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (uint32_t e = row[v]; e < row[v + 1]; e++) prcol[eb + atomicAdd(&fill[col[e]], 1u)] = nb + v;
}
// nodes to reach (defender dead ends): zero energy, and the seed list
__global__ void k_place_reach(const uint32_t* row, const uint8_t* atk, uint32_t V, uint32_t nb, uint32_t t0, uint4* st,
                              uint32_t* seed, uint32_t* n) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && !atk[v] && row[v] == row[v + 1]) {
        st[slot_index(nb + v, 0)] = make_uint4(1u, 0u, 0u, ((t0 - 1u) & 0xffffu) << 16);   // slot 0 stays current
        seed[atomicAdd(n, 1u)] = nb + v;
    }
}
// jobs that still have positions on the worklist of the current round
__global__ void k_mark_active(const Ctl* ctl, uint32_t* const* wl, const uint32_t* job_of, uint8_t* active) {
    const uint32_t r = ctl->round;
    const uint32_t n = ctl->wl_n[r % 3u];
    const uint32_t* list = wl[r & 1u];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) active[job_of[list[i]]] = 1;
}

enum JobState : uint32_t { QUEUED = 0, BUILDING = 1, BUILT = 2, ACTIVE = 3, DONE = 4, FAILED = 5 };

struct Job {
    uint64_t id = 0;
    uint32_t n_states = 0;
    std::vector<uint32_t> row, dst, pairs;
    std::vector<int32_t> lbl;
    std::shared_ptr<BuiltJobs> batch;   // the build this job's game sits in, until placed
    uint32_t bj = 0;                    // index of the job in `batch`
    geq_status status = GEQ_OK;
    std::string error;
    uint32_t V = 0, E = 0, n_start = 0, nb = 0, eb = 0, slot = 0;
    uint32_t positions = 0, edges = 0;   // game size once built
    uint32_t admit_round = 0, done_round = 0;
    double t_submit = 0, t_built = 0, t_admit = 0, t_done = 0;
    bool alone = false;                 // its merged build failed: it is built by itself
    bool exclusive = false;             // its own build ran out of device memory: built again once, with no other build running
    std::atomic<uint32_t> state{QUEUED};
    std::vector<uint32_t> offs, data;
};

#define GEQ_TRY(x)                          \
    do {                                    \
        const geq_status st_ = (x);         \
        if (st_ != GEQ_OK) return st_;      \
    } while (0)

// Scratch that holds at least n elements.  A buffer that has to grow is
// replaced only after the stream has drained: queued kernels may still use it.
template <class T>
geq_status ensure_buf(DevBuf<T>& b, size_t n, hipStream_t s) {
    if (b.p && b.n >= n) return GEQ_OK;
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    GEQ_HIP(b.alloc(std::max<size_t>(n + n / 4, 1024)), GEQ_ERR_OUT_OF_MEMORY);
    return GEQ_OK;
}

// first-fit allocator over an index range
struct RangeAlloc {
    std::map<uint32_t, uint32_t> free_;   // start -> length
    uint64_t total = 0, used = 0;
    void init(uint32_t n) {
        free_.clear();
        if (n) free_[0] = n;
        total = n;
        used = 0;
    }
    bool alloc(uint32_t len, uint32_t& start) {
        if (len == 0) {
            start = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                start = it->first;
                const uint32_t rest = it->second - len;
                free_.erase(it);
                if (rest) free_[start + len] = rest;
                used += len;
                return true;
            }
        }
        return false;
    }
    void release(uint32_t start, uint32_t len) {
        if (len == 0) return;
        used -= len;
        auto it = free_.emplace(start, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace

struct geq_server {
    geq_ctx* ctx = nullptr;
    std::vector<geq_ctx*> bctx;        // builders' contexts: a stream each
    geq_game* pool = nullptr;          // pools of positions / edges / energies
    DevBuf<uint32_t> job_of, seed_cnt, place_cnt, place_fill, place_excl;
    DevBuf<uint32_t> ret_nodes;        // start positions of the jobs retired in a step
    DevBuf<uint8_t> active;
    DevBuf<uint32_t> scan_tmp;   // block totals of the placement scans
    DevBuf<uint32_t*> wl_ptrs;
    std::vector<uint8_t> h_active;
    uint32_t max_jobs = 0;
    uint32_t rounds_per_step = 4;
    RangeAlloc nodes, edges;
    std::vector<uint32_t> free_slots;
    std::map<uint64_t, Job*> jobs;
    std::deque<Job*> build_q, ready_q;
    std::vector<Job*> resident;
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    uint32_t max_ready = 256;          // built jobs waiting for admission at most
    uint32_t max_batch = 64;           // jobs per merged build
    bool trace = false;                // GEQ_SERVER_TRACE: one stderr line per build and step
    std::vector<std::thread> builders;
    std::vector<uint64_t> finished;    // ids of jobs done or failed, not yet taken
    std::mutex spare_mu;
    std::shared_mutex build_rw;        // builds hold it shared; a job built again after running out of memory, exclusively
    bool exclusive_pending = false;    // such a job is taken: the other builders take no new job until it is built
    std::vector<BuiltJobs*> spare_outs;  // job-game buffers of placed builds, for reuse
    uint64_t next_id = 1;
    std::chrono::steady_clock::time_point t0;
    uint64_t steps = 0, rounds_total = 0, compactions = 0, jobs_done = 0, builds = 0;
    uint64_t split_builds = 0;          // merged builds that failed and were split into one-job builds
    uint64_t deferred_admissions = 0;   // admitted jobs sent back to wait: no device memory for their placement
    uint64_t force_defer = 0;           // GEQ_SERVER_DEFER_PLACEMENTS: placements treated as short of memory (tests)
    uint64_t arena_min = 0;            // arena entries the pools start with
    uint64_t compact_entries = 0;      // compaction threshold (0: 3/4 of the arena)
    uint64_t max_positions = 0;        // per-job game size limit (0: the node pool)
    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    ~geq_server();
};

namespace {

// A builder's scratch is replaced after a build that needed more positions
// than this, so that a rare very large job does not keep its device memory.
constexpr uint64_t BIG_BUILD_POSITIONS = 64ull << 20;

// Builder thread: takes every queued job (up to max_batch) and builds their
// games in one level loop over the union of their LTSs (gamebuild_jobs).
void builder_main(geq_server* S, geq_ctx* bctx) {
    (void)hipSetDevice(bctx->device);
    // reused by every build of this thread (replaced after a very large build,
    // so that its peak device memory is not kept for the server's lifetime)
    auto ws = std::make_unique<gb::BuildScratch>();
    auto store = std::make_unique<geq_built>();
    while (true) {
        std::vector<Job*> take;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // no new builds while a job waits to be built with no other build running
            S->cv.wait(lk, [&] {
                return S->stop || (!S->build_q.empty() && S->ready_q.size() < S->max_ready && !S->exclusive_pending);
            });
            if (S->stop) return;
            // a job whose merged build failed is built alone, so that it fails
            // (or succeeds) by itself
            while (!S->build_q.empty() && take.size() < S->max_batch) {
                Job* J = S->build_q.front();
                if (J->alone && !take.empty()) break;
                take.push_back(J);
                S->build_q.pop_front();
                J->state = BUILDING;
                if (J->exclusive) S->exclusive_pending = true;
                if (J->alone) break;
            }
        }
        // invalid jobs fail alone; jobs without pairs have nothing to solve
        std::vector<Job*> build;
        std::vector<geq_lts> ltss;
        std::vector<const uint32_t*> prs;
        std::vector<uint32_t> nps;
        for (Job* J : take) {
            geq_lts l{J->n_states, (uint32_t)J->dst.size(), J->row.data(), J->dst.data(), J->lbl.data()};
            const geq_status st = validate_job(&l, J->pairs.data(), J->n_start);
            std::lock_guard<std::mutex> lk(S->mu);
            if (st != GEQ_OK || J->n_start == 0) {
                J->t_built = J->t_done = S->now_ms();
                J->status = st;
                if (st != GEQ_OK) {
                    J->error = geq_last_error();
                    J->state = FAILED;
                } else {
                    J->offs.assign(1, 0);
                    J->state = DONE;
                    S->jobs_done++;
                }
                S->finished.push_back(J->id);
                continue;
            }
            build.push_back(J);
            ltss.push_back(l);
            prs.push_back(J->pairs.data());
            nps.push_back(J->n_start);
        }
        if (build.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Job* J : take)
                if (J->exclusive) {   // (it did not reach its build: the others may go on)
                    S->exclusive_pending = false;
                    S->cv.notify_all();
                }
            continue;
        }
        BuiltJobs* raw = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->spare_mu);
            if (!S->spare_outs.empty()) {
                raw = S->spare_outs.back();
                S->spare_outs.pop_back();
            }
        }
        if (!raw) raw = new BuiltJobs();
        // returned for reuse once its last job has been placed
        std::shared_ptr<BuiltJobs> B(raw, [S](BuiltJobs* b) {
            std::lock_guard<std::mutex> lk(S->spare_mu);
            S->spare_outs.push_back(b);
        });
        const double tb0 = S->now_ms();
        // A build that ran out of device memory by itself runs again alone:
        // the other builders finish their builds and wait meanwhile.
        const bool exclusive = build.size() == 1 && build[0]->exclusive;
        std::shared_lock<std::shared_mutex> shared(S->build_rw, std::defer_lock);
        std::unique_lock<std::shared_mutex> sole(S->build_rw, std::defer_lock);
        if (exclusive) sole.lock(); else shared.lock();
        const double tb1 = S->now_ms();
        const geq_status st = gamebuild_jobs(bctx, ltss, prs, nps, S->max_positions, *ws, *store, B.get());
        const double tb2 = S->now_ms();
        if (store->peak_nodes > BIG_BUILD_POSITIONS) {   // a very large build: its scratch goes
            ws = std::make_unique<gb::BuildScratch>();
            store = std::make_unique<geq_built>();
        }
        if (exclusive) sole.unlock(); else shared.unlock();
        const std::string err = st != GEQ_OK ? std::string(geq_last_error()) : std::string();
        if (S->trace) {
            uint64_t pos = 0, dead = 0, dead_pos = 0;
            for (uint32_t k = 0; st == GEQ_OK && k < build.size(); k++) {
                pos += B->n_pos[k];
                dead += B->dead[k];
                dead_pos += B->dead[k] ? B->n_pos[k] : 0;
            }
            std::fprintf(stderr,
                         "[geq server] t=%.1f ms build of %zu jobs: %llu positions, %u levels, %llu over the cap (%llu positions), "
                         "%.1f ms (lock %.1f, merge %.1f, level loop %.1f, extraction %.1f, teardown %.1f ms; status %d)\n",
                         tb0, build.size(), (unsigned long long)pos, B->levels, (unsigned long long)dead,
                         (unsigned long long)dead_pos, S->now_ms() - tb0, tb1 - tb0, B->prep_ms, B->build_ms, B->extract_ms,
                         tb2 - tb1 - B->prep_ms - B->build_ms - B->extract_ms, (int)st);
        }
        std::lock_guard<std::mutex> lk(S->mu);
        const double now = S->now_ms();
        S->builds++;
        if (st != GEQ_OK && build.size() > 1) {
            // the merged build failed (a level past the 32-bit totals, device
            // memory): its jobs go back to the front of the queue, each to be
            // built alone, so only the jobs that fail by themselves fail
            for (size_t k = build.size(); k-- > 0;) {
                build[k]->alone = true;
                build[k]->state = QUEUED;
                S->build_q.push_front(build[k]);
            }
            S->split_builds++;
            S->cv.notify_all();
            continue;
        }
        if (st == GEQ_ERR_OUT_OF_MEMORY && build.size() == 1 && !build[0]->exclusive &&
            err.find("2^3") == std::string::npos) {
            // device memory, not a 32-bit limit: once more with no other build running
            build[0]->exclusive = true;
            build[0]->alone = true;
            build[0]->state = QUEUED;
            S->build_q.push_front(build[0]);
            S->cv.notify_all();
            continue;
        }
        if (exclusive) {   // the other builders may take jobs again
            S->exclusive_pending = false;
            S->cv.notify_all();
        }
        for (uint32_t k = 0; k < build.size(); k++) {
            Job* J = build[k];
            J->t_built = now;
            if (st != GEQ_OK || B->dead[k]) {
                J->status = st != GEQ_OK ? st : GEQ_ERR_OUT_OF_MEMORY;
                J->error = st != GEQ_OK ? err : "the job's game grows past " + std::to_string(S->max_positions) + " positions";
                J->state = FAILED;
                J->t_done = now;
                S->finished.push_back(J->id);
                continue;
            }
            J->batch = B;
            J->bj = k;
            J->positions = B->n_pos[k];
            J->edges = B->n_edges[k];
            J->state = BUILT;
            S->ready_q.push_back(J);
        }
    }
}

// Retired jobs leave their long lists behind in the energy arena, whose top
// only grows while rounds run.  Between steps (when only the committed lists
// are live) the lists of every pool slot are packed into a fresh arena.
geq_status compact_arena(geq_server* S) {
    geq_game* g = S->pool;
    hipStream_t s = S->ctx->stream;
    const uint32_t n = g->V, tb = 256;
    GEQ_TRY(ensure_buf(S->place_cnt, n + 1, s));
    GEQ_TRY(ensure_buf(S->place_excl, n + 1, s));
    const uint32_t t = round_tag(g->h_ctl->round);
    k_arena_need<<<(n + 1 + tb - 1) / tb, tb, 0, s>>>(g->st.p, n, t, S->place_cnt.p);
    GEQ_TRY((exclusive_scan_dev<uint32_t, SumU32>(S->place_cnt.p, S->place_excl.p, n, S->scan_tmp, s, SumU32{})));
    uint32_t live = 0;
    GEQ_HIP(hipMemcpyAsync(&live, S->place_excl.p + n, 4, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(S->arena_min, 2ull * live + (1u << 20)), 0xffffffffull);
    uint16_t* p = nullptr;
    hipError_t e = device_malloc(reinterpret_cast<void**>(&p), 2ull * cap);
    if (e != hipSuccess) return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("compacting the energy arena: ") + hipGetErrorString(e));
    k_arena_move<<<(n + tb - 1) / tb, tb, 0, s>>>(g->st.p, g->arena.p, p, S->place_excl.p, n, t);
    e = hipMemcpyAsync(&g->ctl.p->arena_top, &live, 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return fail(GEQ_ERR_TRANSFER, std::string("compacting the energy arena: ") + hipGetErrorString(e));
    }
    (void)hipFree(g->arena.p);
    g->arena.p = p;
    g->arena.n = cap;
    g->arena_cap = (uint32_t)cap;
    g->h_ctl->arena_top = live;
    S->compactions++;
    return GEQ_OK;
}

}  // namespace

geq_server::~geq_server() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto& t : builders)
        if (t.joinable()) t.join();
    for (auto& kv : jobs) delete kv.second;
    for (BuiltJobs* b : spare_outs) delete b;
    if (pool) geq_game_free(pool);
    for (geq_ctx* c : bctx) geq_ctx_destroy(c);
}

extern "C" {

geq_status geq_server_create(geq_ctx* ctx, const geq_server_opts* opts, geq_server** out) {
    if (!ctx || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    const uint64_t budget = opts && opts->budget_bytes ? opts->budget_bytes : (4ull << 30);
    const double edge_ratio = opts && opts->edges_per_position > 0 ? opts->edges_per_position : 4.0;
    // 7/8 of the budget for the position / edge pools, 1/8 for the energy arena
    const uint64_t pool_bytes = budget - budget / 8;
    uint64_t V = (uint64_t)(pool_bytes / (NODE_BYTES + edge_ratio * EDGE_BYTES));
    V = std::min<uint64_t>(V, 0x7fff0000ull);
    uint64_t E = std::min<uint64_t>((uint64_t)(V * edge_ratio), 0xfff00000ull);
    if (V < 64) return fail(GEQ_ERR_INVALID_ARG, "serving budget too small");
    geq_server* S = new geq_server();
    S->ctx = ctx;
    S->t0 = std::chrono::steady_clock::now();
    S->max_jobs = opts && opts->max_jobs ? opts->max_jobs : 4096;
    S->rounds_per_step = opts && opts->rounds_per_step ? opts->rounds_per_step : 4;
    S->compact_entries = opts ? opts->compact_entries : 0;
    S->trace = std::getenv("GEQ_SERVER_TRACE") != nullptr;
    if (const char* env = std::getenv("GEQ_SERVER_DEFER_PLACEMENTS")) S->force_defer = (uint64_t)std::max(0, std::atoi(env));
    if (const char* env = std::getenv("GEQ_SERVER_MAX_BATCH"))   // jobs per merged build (the per-job counters hold 1024)
        S->max_batch = (uint32_t)std::min(std::max(std::atoi(env), 1), 1024);
    S->max_positions = opts && opts->max_positions ? std::min<uint64_t>(opts->max_positions, V) : V;
    const uint32_t n_builders = opts && opts->builders ? std::min(opts->builders, 16u) : 3u;
    geq_status st = GEQ_OK;
    for (uint32_t i = 0; i < n_builders; i++) {
        geq_ctx* c = nullptr;
        st = geq_ctx_create(ctx->device, &c);
        if (st != GEQ_OK) {
            delete S;
            return st;
        }
        S->bctx.push_back(c);
    }
    st = alloc_game_buffers(ctx, (uint32_t)V, (uint32_t)E, &S->pool, std::max<uint64_t>(budget / 16, 4096));
    if (st != GEQ_OK) {
        delete S;
        return st;
    }
    S->pool->tune.rounds_per_launch = S->rounds_per_step;
    S->pool->tune.mode = 1;
    // append-order worklists: the pool's bitmaps span every slot of the pool,
    // most of them free, so listing a round from its bitmap would read it whole
    S->pool->tune.ordered = 0;
    S->arena_min = S->pool->arena_cap;
    auto oom = [&](hipError_t e) {
        delete S;
        return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("serving pools: ") + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = S->job_of.alloc(V)) != hipSuccess) return oom(e);
    if ((e = S->active.alloc(S->max_jobs)) != hipSuccess) return oom(e);
    if ((e = S->seed_cnt.alloc(1)) != hipSuccess) return oom(e);
    if ((e = S->wl_ptrs.alloc(2)) != hipSuccess) return oom(e);
    uint32_t* w[2] = {S->pool->wl0.p, S->pool->wl1.p};
    hipStream_t s = ctx->stream;
    GEQ_HIP(hipMemcpyAsync(S->wl_ptrs.p, w, sizeof(w), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    Ctl c0{};
    *S->pool->h_ctl = c0;
    GEQ_HIP(hipMemcpyAsync(S->pool->ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemsetAsync(S->pool->stats.p, 0, (4ull * STAT_SLOTS + 4) * 8, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemsetAsync(S->active.p, 0, S->max_jobs, s), GEQ_ERR_TRANSFER);
    // no node is queued; a node's bits are clear again whenever it is on no list
    GEQ_HIP(hipMemsetAsync(S->pool->queued.p, 0, 8ull * queue_words(S->pool->V), s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    S->h_active.assign(S->max_jobs, 0);
    S->nodes.init((uint32_t)V);
    S->edges.init((uint32_t)E);
    for (uint32_t i = S->max_jobs; i-- > 0;) S->free_slots.push_back(i);
    for (geq_ctx* c : S->bctx) S->builders.emplace_back(builder_main, S, c);
    *out = S;
    return GEQ_OK;
}

void geq_server_destroy(geq_server* S) {
    if (!S) return;
    (void)hipSetDevice(S->ctx->device);
    delete S;
}

geq_status geq_server_submit(geq_server* S, const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs, uint64_t* job_id) {
    if (!S || !lts || !job_id || (n_pairs && !pairs) || !lts->row_offsets) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    Job* J = new Job();
    J->n_states = lts->n_states;
    J->row.assign(lts->row_offsets, lts->row_offsets + lts->n_states + 1);
    J->dst.assign(lts->targets, lts->targets + lts->n_transitions);
    J->lbl.assign(lts->labels, lts->labels + lts->n_transitions);
    J->pairs.assign(pairs, pairs + 2ull * n_pairs);
    J->n_start = n_pairs;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J->id = S->next_id++;
        J->t_submit = S->now_ms();
        S->jobs[J->id] = J;
        S->build_q.push_back(J);
    }
    S->cv.notify_all();
    *job_id = J->id;
    return GEQ_OK;
}

// One scheduling step: admit built jobs that fit, run rounds_per_step rounds
// over all resident jobs, retire the jobs whose worklist share is empty.
geq_status geq_server_step(geq_server* S, uint32_t* n_resident, uint32_t* n_waiting) {
    if (!S) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    (void)hipSetDevice(S->ctx->device);
    geq_game* g = S->pool;
    hipStream_t s = S->ctx->stream;
    const uint32_t tb = 256;
    // ---- admission (HBM budget = the pools' free ranges)
    std::vector<Job*> admit;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        while (!S->ready_q.empty()) {
            Job* J = S->ready_q.front();
            const uint32_t V = J->positions, E = J->edges;
            if (V + 1ull > S->nodes.total || E > S->edges.total) {
                // larger than the whole budget: it could never be admitted
                J->status = GEQ_ERR_OUT_OF_MEMORY;
                J->error = "job needs " + std::to_string(V) + " positions / " + std::to_string(E) +
                           " edges, more than the serving budget holds";
                J->state = FAILED;
                J->t_done = S->now_ms();
                S->finished.push_back(J->id);
                J->batch.reset();
                S->ready_q.pop_front();
                continue;
            }
            uint32_t nb = 0, eb = 0;
            if (S->free_slots.empty() || !S->nodes.alloc(V + 1, nb)) break;
            if (!S->edges.alloc(E, eb)) {
                S->nodes.release(nb, V + 1);
                break;
            }
            J->V = V;
            J->E = E;
            J->nb = nb;
            J->eb = eb;
            J->slot = S->free_slots.back();
            S->free_slots.pop_back();
            S->ready_q.pop_front();
            admit.push_back(J);
        }
    }
    S->cv.notify_all();
    // Placement of the admitted jobs: their kernels back to back, their seeds
    // in one list, then one wait and one seeding expansion for all of them.
    if (!admit.empty()) {
        uint64_t sum_v = 0;
        uint32_t max_v = 0;
        for (Job* J : admit) {
            sum_v += J->V;
            max_v = std::max(max_v, J->V);
        }
        // scratch for the largest job (and the scans' block totals); when device
        // memory runs short (the builders hold much of it) the jobs wait for a
        // later step instead of failing
        geq_status sst = ensure_buf(g->seed, sum_v, s);
        if (sst == GEQ_OK) sst = ensure_buf(S->place_cnt, max_v + 1ull, s);
        if (sst == GEQ_OK) sst = ensure_buf(S->place_excl, max_v + 1ull, s);
        if (sst == GEQ_OK) sst = ensure_buf(S->place_fill, max_v + 1ull, s);
        if (sst == GEQ_OK) sst = ensure_buf(S->scan_tmp, (max_v + SCAN_ITEMS - 1ull) / SCAN_ITEMS + 2ull, s);
        if (sst == GEQ_OK && S->force_defer) {
            S->force_defer--;
            sst = GEQ_ERR_OUT_OF_MEMORY;
        }
        if (sst == GEQ_ERR_OUT_OF_MEMORY) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto it = admit.rbegin(); it != admit.rend(); ++it) {
                Job* J = *it;
                S->nodes.release(J->nb, J->V + 1);
                S->edges.release(J->eb, J->E);
                S->free_slots.push_back(J->slot);
                S->ready_q.push_front(J);
            }
            S->deferred_admissions += admit.size();
            admit.clear();
        } else if (sst != GEQ_OK) {
            return sst;
        }
    }
    if (!admit.empty()) {
        GEQ_HIP(hipMemsetAsync(S->seed_cnt.p, 0, 4, s), GEQ_ERR_TRANSFER);
        // the next round to run; the seeds' slots are tagged relative to it
        const uint32_t t0 = round_tag(g->h_ctl->round);
        for (Job* J : admit) {
            const uint32_t V = J->V, E = J->E;
            const BuiltJobs& Bj = *J->batch;
            const uint32_t* brow = Bj.row.p + Bj.pos_start[J->bj] + J->bj;
            const uint32_t* bcol = Bj.col.p + Bj.edge_start[J->bj];
            const uint32_t* bwt = Bj.wt.p + Bj.edge_start[J->bj];
            const uint8_t* batk = Bj.atk.p + Bj.pos_start[J->bj];
            // the builder's stream finished this game before it was queued as BUILT
            k_place_nodes<<<(V + 1 + tb - 1) / tb, tb, 0, s>>>(brow, batk, V, J->nb, J->eb, J->slot, t0, g->row.p, g->atk.p,
                                                               g->st.p, S->job_of.p);
            if (E) k_place_edges<<<(E + tb - 1) / tb, tb, 0, s>>>(bcol, bwt, E, J->nb, J->eb, g->edges.p);
            // reverse CSR of the job's range (the scratch is reused by the next job, in stream order)
            GEQ_HIP(hipMemsetAsync(S->place_cnt.p, 0, 4ull * (V + 1), s), GEQ_ERR_TRANSFER);
            if (E) k_place_indeg<<<(E + tb - 1) / tb, tb, 0, s>>>(bcol, E, S->place_cnt.p);
            GEQ_TRY((exclusive_scan_dev<uint32_t, SumU32>(S->place_cnt.p, S->place_excl.p, V, S->scan_tmp, s, SumU32{})));
            k_place_rrow<<<(V + 1 + tb - 1) / tb, tb, 0, s>>>(S->place_excl.p, V, J->eb, J->nb, g->rrow.p, S->place_fill.p);
            k_place_rcol<<<(V + tb - 1) / tb, tb, 0, s>>>(brow, bcol, V, J->nb, J->eb, S->place_fill.p, g->rcol.p);
            k_node_hdr<<<(V + tb - 1) / tb, tb, 0, s>>>(g->row.p + J->nb, g->rrow.p + J->nb, g->atk.p + J->nb, V,
                                                        g->st.p + (size_t)NODE_U4 * J->nb);
            // nodes to reach: zero energy, and the seed list of this step
            k_place_reach<<<(V + tb - 1) / tb, tb, 0, s>>>(brow, batk, V, J->nb, t0, g->st.p, g->seed.p, S->seed_cnt.p);
        }
        uint32_t n_seed = 0;
        GEQ_HIP(hipMemcpyAsync(&n_seed, S->seed_cnt.p, 4, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        if (n_seed) {
            // the predecessors of every admitted job's seeds join the current round's list
            ExpandArgs sx = expand_args(g);
            sx.seed = g->seed.p;
            sx.n_seed = n_seed;
            k_expand<<<std::max(1u, std::min(1024u, (n_seed + 64u * WAVES - 1) / (64u * WAVES))), BLOCK, 0, s>>>(sx, 0u);
        }
        std::lock_guard<std::mutex> lk(S->mu);
        for (Job* J : admit) {
            J->batch.reset();   // the pools hold the game now (the build is freed with its last job)
            J->t_admit = S->now_ms();
            J->state = ACTIVE;
            S->resident.push_back(J);
        }
    }
    // ---- rounds over every resident job
    if (!S->resident.empty()) {
        const RoundShape shape = round_shape(g);
        EvalArgs ea = eval_args(g);
        ExpandArgs xa = expand_args(g);
        const uint32_t round0 = g->h_ctl->round;
        while (true) {
            if (!g->graph_exec || g->graph_arena != g->arena.p) {
                geq_status st = capture_rounds(g, ea, xa, shape);
                if (st != GEQ_OK) return st;
            }
            GEQ_HIP(hipGraphLaunch(g->graph_exec, s), GEQ_ERR_NO_DEVICE);
            GEQ_HIP(hipMemcpyAsync(g->h_ctl, g->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
            GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
            if (!g->h_ctl->overflow) {
                S->rounds_total += g->h_ctl->round - round0;
                GEQ_TRY(maybe_retag(g));
                break;
            }
            GEQ_TRY(recover_overflow(g));
            ea.arena = g->arena.p;
            ea.arena_cap = g->arena_cap;
        }
        S->steps++;
        if (S->trace)
            std::fprintf(stderr, "[geq server] t=%.1f ms step %llu: %zu resident, %zu admitted, round %u, worklist %u\n", S->now_ms(),
                         (unsigned long long)S->steps, S->resident.size(), admit.size(), g->h_ctl->round,
                         g->h_ctl->wl_n[g->h_ctl->round % 3u]);
        // ---- retirement
        GEQ_HIP(hipMemsetAsync(S->active.p, 0, S->max_jobs, s), GEQ_ERR_TRANSFER);
        k_mark_active<<<256, 256, 0, s>>>(g->ctl.p, S->wl_ptrs.p, S->job_of.p, S->active.p);
        GEQ_HIP(hipMemcpyAsync(S->h_active.data(), S->active.p, S->max_jobs, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        std::vector<Job*> still, done_jobs;
        for (Job* J : S->resident) (S->h_active[J->slot] ? still : done_jobs).push_back(J);
        const uint32_t retired = (uint32_t)done_jobs.size();
        if (retired) {
            // the start positions of every retired job, downloaded together
            std::vector<uint32_t> nodes;
            for (Job* J : done_jobs)
                for (uint32_t k = 0; k < J->n_start; k++) nodes.push_back(J->nb + k);
            geq_energies res{};
            geq_status st = ensure_buf(S->ret_nodes, nodes.size(), s);
            if (st == GEQ_OK && hipMemcpyAsync(S->ret_nodes.p, nodes.data(), 4 * nodes.size(), hipMemcpyHostToDevice, s) != hipSuccess)
                st = fail(GEQ_ERR_TRANSFER, "server: uploading the retired jobs' positions");
            if (st == GEQ_OK)
                st = download_mapped(g, g->st.p, S->ret_nodes.p, (uint32_t)nodes.size(), round_tag(g->h_ctl->round), &res);
            const std::string err = st != GEQ_OK ? std::string(geq_last_error()) : std::string();
            std::lock_guard<std::mutex> lk(S->mu);
            size_t at = 0;
            for (Job* J : done_jobs) {
                if (st != GEQ_OK) {
                    J->status = st;
                    J->error = err;
                    J->state = FAILED;
                } else {
                    const uint32_t b = res.offsets[at], e = res.offsets[at + J->n_start];
                    J->offs.resize(J->n_start + 1);
                    for (uint32_t k = 0; k <= J->n_start; k++) J->offs[k] = res.offsets[at + k] - b;
                    J->data.assign(res.data + b, res.data + e);
                    J->state = DONE;
                    S->jobs_done++;
                }
                at += J->n_start;
                J->t_done = S->now_ms();
                S->finished.push_back(J->id);
                k_clear_nodes<<<(J->V + 1 + tb - 1) / tb, tb, 0, s>>>(g->st.p, J->nb, J->V + 1);
                S->nodes.release(J->nb, J->V + 1);
                S->edges.release(J->eb, J->E);
                S->free_slots.push_back(J->slot);
            }
            if (st == GEQ_OK) geq_energies_free(&res);
        }
        S->resident.swap(still);
        const uint64_t top = g->h_ctl->arena_top;
        if (retired && (S->compact_entries ? top >= S->compact_entries : top * 4 > 3ull * g->arena_cap)) {
            geq_status st = compact_arena(S);
            if (st != GEQ_OK) return st;
        }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (n_resident) *n_resident = (uint32_t)S->resident.size();
    if (n_waiting) *n_waiting = (uint32_t)(S->build_q.size() + S->ready_q.size());
    return GEQ_OK;
}

geq_status geq_server_poll(geq_server* S, uint64_t job_id, geq_job_info* info, geq_energies* out) {
    if (!S || !info) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return fail(GEQ_ERR_INVALID_ARG, "unknown job id");
    Job* J = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = J->state;
    info->t_submit_ms = J->t_submit;
    info->t_built_ms = J->t_built;
    info->t_admit_ms = J->t_admit;
    info->t_done_ms = J->t_done;
    info->positions = J->positions;
    info->edges = J->edges;
    info->status = J->status;
    if (J->state == FAILED) {
        geq::set_error(J->error);
        return GEQ_OK;
    }
    if (J->state == DONE && out) {
        out->n_nodes = J->n_start;
        out->energy_size = 1;
        out->offsets = (uint32_t*)std::malloc(4ull * (J->n_start + 1));
        out->data = (uint32_t*)std::malloc(4ull * std::max<size_t>(J->data.size(), 1));
        std::memcpy(out->offsets, J->offs.data(), 4ull * (J->n_start + 1));
        if (!J->data.empty()) std::memcpy(out->data, J->data.data(), 4ull * J->data.size());
    }
    return GEQ_OK;
}

geq_status geq_server_get_stats(geq_server* S, geq_server_stats* out) {
    if (!S || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(out, 0, sizeof(*out));
    out->steps = S->steps;
    out->rounds = S->rounds_total;
    out->compactions = S->compactions;
    out->node_slots = S->nodes.total;
    out->node_slots_used = S->nodes.used;
    out->edge_slots = S->edges.total;
    out->edge_slots_used = S->edges.used;
    out->arena_entries = S->pool->arena_cap;
    out->arena_top = S->pool->h_ctl->arena_top;
    out->jobs_resident = S->resident.size();
    out->jobs_waiting = S->build_q.size() + S->ready_q.size();
    out->jobs_done = S->jobs_done;
    out->builds = S->builds;
    out->split_builds = S->split_builds;
    out->deferred_admissions = S->deferred_admissions;
    return GEQ_OK;
}

geq_status geq_server_finished(geq_server* S, uint64_t* ids, uint32_t cap, uint32_t* n) {
    if (!S || !n || (cap && !ids)) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    const uint32_t k = (uint32_t)std::min<size_t>(cap, S->finished.size());
    std::copy(S->finished.begin(), S->finished.begin() + k, ids);
    S->finished.erase(S->finished.begin(), S->finished.begin() + k);
    *n = k;
    return GEQ_OK;
}

geq_status geq_server_release(geq_server* S, uint64_t job_id) {
    if (!S) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return fail(GEQ_ERR_INVALID_ARG, "unknown job id");
    if (it->second->state != DONE && it->second->state != FAILED) return fail(GEQ_ERR_INVALID_ARG, "job still running");
    delete it->second;
    S->jobs.erase(it);
    return GEQ_OK;
}

}  // extern "C"
