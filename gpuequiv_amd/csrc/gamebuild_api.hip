// Host driver of the device gamebuild (level loop, capacities) and the C ABI
// entry points geq_gamebuild / geq_built_* of include/gpuequiv.h.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "gamebuild.hpp"
#include "runtime.hpp"

using namespace geq;
using namespace geq::gb;

namespace {

__global__ void k_edges_u2(const uint32_t* col, const uint32_t* wt, uint32_t E, uint2* edges) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) edges[e] = make_uint2(col[e], wt[e]);
}
// in-degrees, and each edge's rank among its target's predecessors (the count
// its atomic saw), so the placement below needs no atomics
__global__ void k_indeg(const uint32_t* col, uint32_t E, uint32_t* cnt, uint32_t* rank) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) rank[e] = atomicAdd(&cnt[col[e]], 1u);
}
// predecessor v of each edge's target at its rank (the edges of v are
// independent: no iteration waits on the one before)
__global__ void k_rev_scatter(const uint32_t* row, uint32_t V, const uint32_t* col, const uint32_t* rank, const uint32_t* rrow,
                              uint32_t* rcol) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (uint32_t e = row[v]; e < row[v + 1]; e++) rcol[rrow[col[e]] + rank[e]] = v;
}


template <class T>
geq_status ensure(DevBuf<T>& b, size_t need, hipStream_t s, bool keep, const char* what) {
    if (b.p && b.n >= need) return GEQ_OK;
    size_t cap = std::max<size_t>(need, b.n ? 2 * b.n : 1024);
    T* p = nullptr;
    size_t got = 0;
    if ((p = static_cast<T*>(spare_take(sizeof(T) * cap, &got))) != nullptr) {
        cap = got / sizeof(T);   // a spare buffer of an earlier build on this thread
    } else {
        hipError_t e = device_malloc(reinterpret_cast<void**>(&p), sizeof(T) * cap);
        if (e != hipSuccess)
            return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("gamebuild: allocating ") + what + ": " + hipGetErrorString(e));
    }
    hipError_t e = hipSuccess;
    if (keep && b.p && b.n) {
        // stream-ordered: the old buffer is retired below, not freed, when a
        // RetireScope is open; otherwise the copy must finish before the free
        e = hipMemcpyAsync(p, b.p, sizeof(T) * b.n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && !RetireScope::active()) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return fail(GEQ_ERR_TRANSFER, std::string("gamebuild: growing ") + what);
        }
    }
    b.retire();
    b.p = p;
    b.n = cap;
    return GEQ_OK;
}

// Exclusive scan of in[0..n) into out[0..n); out[n] = the total.
geq_status exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, DevBuf<uint32_t>& tmp, hipStream_t s) {
    return exclusive_scan_dev<uint32_t, SumU32>(in, out, n, tmp, s, SumU32{});
}

// The saturation point of the bound scans: 2^32 - 1, or GEQ_BUILD_LEVEL_LIMIT
// (lower values exercise the refusal in tests).
uint32_t level_limit() {
    if (const char* env = std::getenv("GEQ_BUILD_LEVEL_LIMIT")) {
        const long long v = std::atoll(env);
        if (v >= 1 && v < 0xffffffffll) return (uint32_t)v;
    }
    return 0xffffffffu;
}

// Exclusive scan of (candidates, set-pool words) pairs, one scan for both:
// out[0..n) and out[n] = the totals.
geq_status exclusive_scan_pairs(const uint2* in, uint2* out, size_t n, DevBuf<uint2>& tmp, hipStream_t s) {
    return exclusive_scan_dev<uint2, PairSum>(in, out, n, tmp, s, PairSum{level_limit()});
}

// Writes a chunk's candidate and set-pool totals (*a, *b) and its heavy
// positions (*h) to pinned host memory: they reach the host with one launch
// and one wait.
__global__ void k_readback(const uint32_t* a, const uint32_t* b, const uint32_t* h, uint32_t* out, uint32_t seq) {
    if (threadIdx.x != 0) return;
    out[RB_CAND] = *a;
    out[RB_WORDS] = *b;
    out[RB_HEAVY] = *h;
    rb_publish(out, seq);
}

}  // namespace


geq_status geq::alloc_game_buffers(geq_ctx* ctx, uint32_t V, uint32_t E, geq_game** out, uint64_t arena_entries) {
    geq_game* gm = new geq_game();
    gm->ctx = ctx;
    check_conf(geq_conf{6, 3}, gm->conf);
    gm->V = V;
    gm->E = E;
    gm->game_base = {0, V};
    auto oom = [&](hipError_t e, const char* what) -> geq_status {
        delete gm;
        return fail(e == hipErrorOutOfMemory ? GEQ_ERR_OUT_OF_MEMORY : GEQ_ERR_NO_DEVICE,
                    std::string("allocating ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = gm->row.alloc(V + 1)) != hipSuccess) return oom(e, "row_offsets");
    if ((e = gm->rrow.alloc(V + 1)) != hipSuccess) return oom(e, "rev_row_offsets");
    if ((e = gm->rcol.alloc(E)) != hipSuccess) return oom(e, "rev_column_indices");
    if ((e = gm->edges.alloc(E)) != hipSuccess) return oom(e, "edges");
    if ((e = gm->atk.alloc(V)) != hipSuccess) return oom(e, "attacker_pos");
    if ((e = gm->st.alloc((size_t)NODE_U4 * V)) != hipSuccess) return oom(e, "node records");
    if ((e = gm->wl0.alloc(V)) != hipSuccess) return oom(e, "worklist");
    if ((e = gm->wl1.alloc(V)) != hipSuccess) return oom(e, "worklist");
    if ((e = gm->chg.alloc(V)) != hipSuccess) return oom(e, "changed flags");
    if ((e = gm->queued.alloc(2ull * queue_words(V))) != hipSuccess) return oom(e, "queued bitmaps");
    if ((e = gm->ctl.alloc(1)) != hipSuccess) return oom(e, "control block");
    if ((e = gm->stats.alloc(4ull * STAT_SLOTS + 4)) != hipSuccess) return oom(e, "stats");
    gm->arena_cap = initial_arena_entries(E, arena_entries);
    if ((e = gm->arena.alloc(gm->arena_cap)) != hipSuccess) return oom(e, "energy arena");
    if ((e = alloc_host_ctl(gm)) != hipSuccess) return oom(e, "pinned control block");
    if (const char* env = std::getenv("GEQ_SOLVE_MODE")) gm->tune.mode = (uint32_t)std::atoi(env);
    gm->tune.ordered = worklist_order_default();
    *out = gm;
    return GEQ_OK;
}

namespace {
// Whether a cached game's buffers hold a game of V nodes and E edges.
bool game_fits(const geq_game* g, uint32_t V, uint32_t E) {
    return g->row.n >= V + 1ull && g->rrow.n >= V + 1ull && g->rcol.n >= E && g->edges.n >= E && g->atk.n >= V &&
           g->st.n >= (size_t)NODE_U4 * V && g->wl0.n >= V && g->wl1.n >= V && g->chg.n >= V &&
           g->queued.n >= 2ull * queue_words(V) && g->arena_cap >= initial_arena_entries(E, 0) && g->h_ctl && g->ctl.p &&
           g->stats.p;
}
}  // namespace

bool geq::drop_build_spares(gb::BuildScratch* ws) { return ws->spare.drop(); }

void geq::release_built_game(geq_game* g) {
    geq_ctx* ctx = g->ctx;
    geq_game* drop = g;
    {
        // keep the game with more node room: it fits more of the next games
        std::lock_guard<std::mutex> lk(ctx->game_mu);
        if (!ctx->spare_game || ctx->spare_game->st.n <= g->st.n) {
            drop = ctx->spare_game;
            ctx->spare_game = g;
        }
    }
    if (drop) geq_game_free(drop);
}

geq_status geq::game_from_device(geq_ctx* ctx, uint32_t V, uint32_t E, const uint32_t* d_row, const uint32_t* d_col,
                                 const uint32_t* d_wt, const uint8_t* d_atk, geq_game** out) {
    hipStream_t s = ctx->stream;
    geq_game* gm = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->game_mu);
        if (ctx->spare_game && game_fits(ctx->spare_game, V, E)) {
            gm = ctx->spare_game;
            ctx->spare_game = nullptr;
        }
    }
    if (gm) {
        // the buffers of a freed built game: everything of the old game is reset
        // (its lists and control block are reset by every solve)
        gm->V = V;
        gm->E = E;
        gm->game_base = {0, V};
        gm->tune = geq_game::Tune{};
        if (const char* env = std::getenv("GEQ_SOLVE_MODE")) gm->tune.mode = (uint32_t)std::atoi(env);
        gm->tune.ordered = worklist_order_default();
        gm->std_ready = false;
        gm->std_n = 0;
        gm->pg_ready = false;         // solve mode 3's game bounds and seed grouping
        gm->pg_std_grouped = false;
        gm->reach_limit = 0xffffffffu;
        gm->retag_round = 0;
        gm->solves = 0;
        if (gm->graph_exec) (void)hipGraphExecDestroy(gm->graph_exec);
        gm->graph_exec = nullptr;
        gm->graph_arena = nullptr;
        gm->h_atk.clear();
    } else {
        geq_status st0 = alloc_game_buffers(ctx, V, E, &gm);
        if (st0 != GEQ_OK) return st0;
    }
    // freed on every error path below (the cached buffers included)
    std::unique_ptr<geq_game> own(gm);
    auto oom = [&](hipError_t e, const char* what) -> geq_status {
        return fail(e == hipErrorOutOfMemory ? GEQ_ERR_OUT_OF_MEMORY : GEQ_ERR_NO_DEVICE,
                    std::string("allocating ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    DevBuf<uint32_t>& cnt = gm->rev_cnt;
    DevBuf<uint32_t>& rank = gm->rev_fill;   // one entry per edge
    DevBuf<uint32_t>& tmp = gm->rev_tmp;
    if (cnt.n < V + 1ull && (e = cnt.alloc(V + 1)) != hipSuccess) return oom(e, "reverse CSR scratch");
    if (rank.n < std::max<uint64_t>(E, 1) && (e = rank.alloc(std::max<uint64_t>(E, 1))) != hipSuccess)
        return oom(e, "reverse CSR scratch");
    const uint32_t tb = 256;
    GEQ_HIP(hipMemcpyAsync(gm->row.p, d_row, 4ull * (V + 1), hipMemcpyDeviceToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemcpyAsync(gm->atk.p, d_atk, V, hipMemcpyDeviceToDevice, s), GEQ_ERR_TRANSFER);
    if (E) k_edges_u2<<<(E + tb - 1) / tb, tb, 0, s>>>(d_col, d_wt, E, gm->edges.p);
    k_fill_u32<<<(V + 1 + tb - 1) / tb, tb, 0, s>>>(cnt.p, V + 1, 0u);
    if (E) k_indeg<<<(E + tb - 1) / tb, tb, 0, s>>>(d_col, E, cnt.p, rank.p);
    geq_status st = exclusive_scan(cnt.p, gm->rrow.p, V, tmp, s);
    if (st != GEQ_OK) return st;
    if (V) k_rev_scatter<<<(V + tb - 1) / tb, tb, 0, s>>>(gm->row.p, V, d_col, rank.p, gm->rrow.p, gm->rcol.p);
    if (V) k_node_hdr<<<(V + tb - 1) / tb, tb, 0, s>>>(gm->row.p, gm->rrow.p, gm->atk.p, V, gm->st.p);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    *out = own.release();
    return GEQ_OK;
}

// Checks of an LTS + start pairs: what geq_gamebuild rejects with INVALID_ARG.
geq_status geq::validate_job(const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs) {
    if (!lts || (n_pairs && !pairs)) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    const uint32_t n = lts->n_states, m = lts->n_transitions;
    if (n == 0 || !lts->row_offsets || (m && (!lts->targets || !lts->labels)))
        return fail(GEQ_ERR_INVALID_ARG, "empty or incomplete LTS");
    if (n >= (1u << 30)) return fail(GEQ_ERR_INVALID_ARG, "at most 2^30 states");
    if (lts->row_offsets[0] != 0 || lts->row_offsets[n] != m) return fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets malformed");
    for (uint32_t st = 0; st < n; st++) {
        const uint32_t b = lts->row_offsets[st], e = lts->row_offsets[st + 1];
        if (e < b) return fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets not monotone");
        for (uint32_t k = b; k < e; k++) {
            if (lts->targets[k] >= n) return fail(GEQ_ERR_INVALID_ARG, "LTS target out of range");
            if (k > b && lts->labels[k] < lts->labels[k - 1])
                return fail(GEQ_ERR_INVALID_ARG, "LTS transitions of a state must be sorted by label");
        }
    }
    for (uint32_t i = 0; i < 2 * n_pairs; i++)
        if (pairs[i] >= n) return fail(GEQ_ERR_INVALID_ARG, "start pair process out of range");
    return GEQ_OK;
}

namespace {

// Per-job bookkeeping of a merged build: the job of every state, each job's
// position count so far, its cap and whether it went past it.
struct JobTrack {
    std::vector<uint32_t> job_of_state;    // uploaded to BuildScratch::job_of_state by build_core
    std::vector<uint64_t> count;
    std::vector<uint8_t> dead;
    uint64_t cap = 0;
};

// New positions per job among a level's candidates: the winners (where the
// rank steps) counted by the job of their process (LDS histogram per block).
__global__ void k_job_count_cands(const Cand* cands, const uint32_t* rank, uint32_t n_cand, const uint32_t* job_of_state,
                                  uint32_t n_jobs, uint32_t* cnt) {
    __shared__ uint32_t h[1024];
    for (uint32_t j = threadIdx.x; j < n_jobs; j += blockDim.x) h[j] = 0;
    __syncthreads();
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cand; c += gridDim.x * blockDim.x)
        if (rank[c + 1] != rank[c]) atomicAdd(&h[job_of_state[cands[c].kind_p & 0x3fffffffu]], 1u);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n_jobs; j += blockDim.x)
        if (h[j]) atomicAdd(&cnt[j], h[j]);
}

// Positions per job among the ids [lo, hi) (LDS histogram per block).
__global__ void k_job_count(const uint32_t* p, uint32_t lo, uint32_t hi, const uint32_t* job_of_state, uint32_t n_jobs,
                            uint32_t* cnt) {
    __shared__ uint32_t h[1024];
    for (uint32_t j = threadIdx.x; j < n_jobs; j += blockDim.x) h[j] = 0;
    __syncthreads();
    for (uint32_t v = lo + blockIdx.x * blockDim.x + threadIdx.x; v < hi; v += gridDim.x * blockDim.x)
        atomicAdd(&h[job_of_state[p[v]]], 1u);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n_jobs; j += blockDim.x)
        if (h[j]) atomicAdd(&cnt[j], h[j]);
}

// The level-synchronous build of the game seeded with `pairs` over a
// label-sorted LTS (already validated).  With `jt`, the LTS is a disjoint
// union of jobs: a job whose position count passes jt->cap is marked dead and
// its positions get no further successors; `max_positions` is then unused.
geq_status build_core(geq_ctx* ctx, const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs, uint64_t max_positions,
                      JobTrack* jt, geq_built* B, BuildScratch& ws, geq_build_stats* stats) {
    const uint32_t n = lts->n_states, m = lts->n_transitions;
    std::vector<uint32_t> en_row(n + 1, 0);
    std::vector<int32_t> en_lbl;
    for (uint32_t st = 0; st < n; st++) {
        const uint32_t b = lts->row_offsets[st], e = lts->row_offsets[st + 1];
        for (uint32_t k = b; k < e; k++)
            if (k == b || lts->labels[k] != lts->labels[k - 1]) en_lbl.push_back(lts->labels[k]);
        en_row[st + 1] = (uint32_t)en_lbl.size();
    }
    // GameBuild::new_node keeps the last id of a repeated start position in its map
    std::vector<uint8_t> insert(n_pairs, 1);
    {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> last;
        for (uint32_t i = 0; i < n_pairs; i++) {
            auto key = std::make_pair(pairs[2 * i], pairs[2 * i + 1]);
            auto it = last.find(key);
            if (it != last.end()) insert[it->second] = 0;
            last[key] = i;
        }
    }

    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    RetireScope retire_scope(s, &ws.spare);   // outgrown buffers are kept for later builds
    auto t0 = std::chrono::steady_clock::now();
    B->ctx = ctx;
    B->n_start = n_pairs;
    B->V = B->E = 0;
    B->spool_top = 0;
    auto& ub = ws.ub; auto& off = ws.off; auto& nsucc = ws.nsucc;
    auto& rowl = ws.rowl; auto& flag = ws.flag; auto& rank = ws.rank; auto& keymin = ws.keymin;
    auto& cands = ws.cands; auto& table = ws.table;
    geq_status st = GEQ_OK;
    // a failed build waits for its queued work before it returns: the next
    // build on this scratch rewrites the pinned stage its upload reads
    auto bail = [&](geq_status e) {
        (void)hipStreamSynchronize(s);
        return e;
    };
#define GB_TRY(x)                      \
    do {                               \
        geq_status _s = (x);           \
        if (_s != GEQ_OK) return bail(_s); \
    } while (0)
#define GB_HIP(x, code)                                                                          \
    do {                                                                                         \
        hipError_t _e = (x);                                                                     \
        if (_e != hipSuccess) return bail(fail((code), std::string(#x) + ": " + hipGetErrorString(_e))); \
    } while (0)
    auto& d_jcnt = ws.d_jcnt;
    const uint32_t n_jobs = jt ? (uint32_t)jt->count.size() : 0u;
    // label bit sets when every label is in [0, 64) (label_range, compare_enabled)
    bool use_mask = !en_lbl.empty();
    for (const int32_t a : en_lbl) use_mask = use_mask && a >= 0 && a < 64;
    if (const char* env = std::getenv("GEQ_BUILD_LABEL_MASK")) use_mask = use_mask && std::atoi(env) != 0;
    std::vector<unsigned long long> mask;
    if (use_mask) {
        mask.assign(n, 0ull);
        for (uint32_t st = 0; st < n; st++)
            for (uint32_t k = en_row[st]; k < en_row[st + 1]; k++) mask[st] |= 1ull << en_lbl[k];
    }
    // The LTS, its label index and bit sets, the start pairs and (merged
    // builds) the job map and flags go up as one copy from pinned memory into
    // one device buffer (a copy from pageable memory waits for the runtime's
    // staging, each time).
    struct Part {
        const void* src;
        size_t bytes, off;
    };
    std::vector<Part> parts;
    size_t pack_bytes = 0;
    auto part = [&](const void* src, size_t bytes) {
        parts.push_back(Part{src, bytes, pack_bytes});
        pack_bytes = (pack_bytes + bytes + 15) & ~(size_t)15;
        return parts.back().off;
    };
    const size_t o_row = part(lts->row_offsets, 4ull * (n + 1));
    const size_t o_dst = part(lts->targets, 4ull * m);
    const size_t o_lbl = part(lts->labels, 4ull * m);
    const size_t o_enrow = part(en_row.data(), 4ull * (n + 1));
    const size_t o_enlbl = part(en_lbl.data(), 4ull * en_lbl.size());
    const size_t o_mask = part(mask.data(), 8ull * mask.size());
    const size_t o_pairs = part(pairs, 8ull * n_pairs);
    const size_t o_insert = part(insert.data(), n_pairs);
    const size_t o_job = part(jt ? jt->job_of_state.data() : nullptr, jt ? 4ull * n : 0);
    const size_t o_dead = part(jt ? jt->dead.data() : nullptr, jt ? (size_t)n_jobs : 0);
    if (ws.stage.bytes < pack_bytes) {
        // the last build's copy from the old stage ended before that build returned
        const size_t bytes = std::max<size_t>(pack_bytes, 2 * ws.stage.bytes);
        ws.stage.release();
        GB_HIP(hipHostMalloc(&ws.stage.p, bytes, hipHostMallocDefault), GEQ_ERR_OUT_OF_MEMORY);
        ws.stage.bytes = bytes;
    }
    for (const Part& pt : parts)
        if (pt.bytes) std::memcpy(static_cast<char*>(ws.stage.p) + pt.off, pt.src, pt.bytes);
    GB_TRY(ensure(ws.l_pack, std::max<size_t>(pack_bytes / 4, 1), s, false, "LTS"));
    if (pack_bytes) GB_HIP(hipMemcpyAsync(ws.l_pack.p, ws.stage.p, pack_bytes, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    char* const pack = reinterpret_cast<char*>(ws.l_pack.p);
    auto at32 = [&](size_t off) { return reinterpret_cast<uint32_t*>(pack + off); };
    uint32_t* const d_pairs = at32(o_pairs);
    uint8_t* const d_insert = reinterpret_cast<uint8_t*>(pack + o_insert);
    uint32_t* const d_job_of_state = at32(o_job);
    uint8_t* const d_dead = reinterpret_cast<uint8_t*>(pack + o_dead);
    ws.job_map = jt ? d_job_of_state : nullptr;
    if (jt) {
        GB_TRY(ensure(d_jcnt, n_jobs, s, false, "job counters"));
        GB_HIP(hipMemsetAsync(d_jcnt.p, 0, 4ull * n_jobs, s), GEQ_ERR_TRANSFER);
    }
    LtsArgs L{at32(o_row), at32(o_dst), reinterpret_cast<const int32_t*>(pack + o_lbl), at32(o_enrow),
              reinterpret_cast<const int32_t*>(pack + o_enlbl), jt ? d_job_of_state : nullptr, jt ? d_dead : nullptr,
              n, m, (uint32_t)en_lbl.size(), use_mask ? reinterpret_cast<const unsigned long long*>(pack + o_mask) : nullptr};
    // a small LTS is read from LDS by k_generate (its walks are chains of dependent reads)
    const uint32_t gen_lds = lts_words(n, m, (uint32_t)en_lbl.size()) <= LTS_LDS_WORDS
                                 ? 4u * lts_words(n, m, (uint32_t)en_lbl.size()) : 0u;
    // the level loop's counts come back through pinned words written by
    // k_readback and k_chunk_counts
    if (ws.rb.n < RB_JOBS + n_jobs) {
        ws.rb.release();
        const size_t words = std::max<size_t>(64, RB_JOBS + n_jobs);
        GB_HIP(hipHostMalloc(reinterpret_cast<void**>(&ws.rb.h), 4 * words, hipHostMallocMapped | hipHostMallocCoherent),
               GEQ_ERR_OUT_OF_MEMORY);
        if (hipHostGetDevicePointer(reinterpret_cast<void**>(&ws.rb.d), ws.rb.h, 0) != hipSuccess || !ws.rb.d) {
            ws.rb.release();   // usable only with both views
            return bail(fail(GEQ_ERR_INTERNAL, "gamebuild: no device view of the pinned count words"));
        }
        ws.rb.n = words;
        ws.rb.h[RB_SEQ] = ws.rb_seq;
    }
    const uint32_t* const h_rb = ws.rb.h;
    // The kernel that ends a wait writes the count words and then, last, the
    // wait's number into h_rb[RB_SEQ] (rb_publish).  The host spins on that
    // word, which is quicker than the runtime's completion wait; after 2 ms it
    // waits for the stream instead, so a failed kernel still ends the wait
    // (with the stream's error).
    auto next_seq = [&]() { return ++ws.rb_seq; };
    auto wait = [&](uint32_t seq) -> geq_status {
        const auto w0 = std::chrono::steady_clock::now();
        while (__atomic_load_n(h_rb + RB_SEQ, __ATOMIC_ACQUIRE) != seq) {
            if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(2)) {
                GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
                if (__atomic_load_n(h_rb + RB_SEQ, __ATOMIC_ACQUIRE) != seq)
                    return fail(GEQ_ERR_INTERNAL, "gamebuild: the level's count words did not arrive");
                break;
            }
            __builtin_ia32_pause();
        }
        return GEQ_OK;
    };

    // node, edge and set-pool capacities grow level by level
    auto node_cap = [&](size_t need) -> geq_status {
        B->peak_nodes = std::max<uint64_t>(B->peak_nodes, need);
        geq_status e;
        if ((e = ensure(B->kind, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->p, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->q_off, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->q_len, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->x_off, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->x_len, need, s, true, "positions")) != GEQ_OK) return e;
        if ((e = ensure(B->hash, need, s, true, "positions")) != GEQ_OK) return e;
        return ensure(B->row, need + 1, s, true, "row offsets");
    };
    uint32_t table_mask = 0;
    uint64_t table_slots = 0;     // power of two in use (the buffers may be larger)
    // The table grows eightfold (each growth refills it and rehashes every
    // position), to at most 2^31 slots.
    auto table_cap = [&](uint64_t entries) -> geq_status {
        if (table_slots >= 2 * entries) return GEQ_OK;
        uint64_t cap = 1024;
        while (cap < 2 * entries) cap <<= 1;
        cap = std::max<uint64_t>(cap, std::min<uint64_t>(8 * table_slots, 1ull << 31));
        if (cap > (1ull << 31)) return fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: position table exceeds 2^31 entries");
        table_slots = cap;
        geq_status e;
        if ((e = ensure(table, cap, s, false, "position table")) != GEQ_OK) return e;
        if ((e = ensure(keymin, cap, s, false, "position table")) != GEQ_OK) return e;
        table_mask = (uint32_t)(cap - 1);
        const uint32_t tb = 256;
        k_fill_u64<<<(uint32_t)((cap + tb - 1) / tb), tb, 0, s>>>(table.p, cap, EMPTY_ENTRY);
        k_fill_u32<<<(uint32_t)((cap + tb - 1) / tb), tb, 0, s>>>(keymin.p, cap, 0xffffffffu);
        if (B->V)
            k_rehash<<<(B->V + tb - 1) / tb, tb, 0, s>>>(B->hash.p, B->V, d_insert, std::min<uint32_t>(n_pairs, B->V), table.p,
                                                        table_mask);
        return GEQ_OK;
    };
    auto nodes = [&]() {
        return NodeArgs{B->kind.p, B->p.p, B->q_off.p, B->q_len.p, B->x_off.p, B->x_len.p, B->hash.p, B->spool.p};
    };

    const uint32_t tb = 256;
    // a standalone build's arrays start at the size of the context's last
    // build (repeated builds of similar games then allocate once)
    const uint64_t hint_cap = 1ull << 28;
    const uint64_t hv = jt ? 0 : std::min<uint64_t>(ctx->hint_positions.load(), hint_cap);
    const uint64_t he = jt ? 0 : std::min<uint64_t>(ctx->hint_edges.load(), hint_cap);
    const uint64_t hw = jt ? 0 : std::min<uint64_t>(ctx->hint_words.load(), hint_cap);
    GB_TRY(node_cap(std::max<uint64_t>(std::max<uint32_t>(n_pairs, 1), hv)));
    GB_TRY(ensure(B->spool, std::max<uint64_t>((uint64_t)std::max<uint32_t>(n_pairs, 1) + 1024, hw), s, false, "set pool"));
    if (he) {
        GB_TRY(ensure(B->col, he, s, false, "edges"));
        GB_TRY(ensure(B->wt, he, s, false, "edges"));
    }
    // a standalone build's table starts at the size of the context's last
    // build too (up to 2^24 positions: the fill costs as much as the table)
    GB_TRY(table_cap(std::max<uint64_t>(std::max<uint32_t>(n_pairs, 1), std::min<uint64_t>(hv, 1ull << 24))));
    if (n_pairs)   // the pairs and their insert flags arrived with the LTS
        k_starts<<<(n_pairs + tb - 1) / tb, tb, 0, s>>>(d_pairs, n_pairs, d_insert, nodes(), table.p, table_mask);
    B->V = n_pairs;
    B->spool_top = n_pairs;
    if (jt && n_pairs) {
        const uint32_t nb = std::min<uint32_t>(1024, (n_pairs + tb - 1) / tb);
        k_job_count<<<nb, tb, 0, s>>>(B->p.p, 0, n_pairs, d_job_of_state, n_jobs, d_jcnt.p);
    }
    uint64_t E = 0, levels = 0;
    uint32_t lo = 0, hi = n_pairs;   // [lo, hi): the rest of the current level's frontier
    uint32_t level_lo = 0;           // first position of the current level
    // GEQ_BUILD_TRACE=1: one stderr line per level (frontier, new positions, totals, time)
    const char* trace_env = std::getenv("GEQ_BUILD_TRACE");
    const bool trace = trace_env && std::atoi(trace_env) != 0;
    bool truncated = false;
    // Merged builds take a level in chunks of frontier positions and check the
    // per-job caps after each chunk, so a job that passes its cap stops within
    // the level instead of finishing it.  Chunks of one level, in parent
    // order, create exactly the positions, ids and edges the whole level would
    // (dedup across chunks goes through the position table).
    uint32_t chunk = (jt && jt->cap) ? (uint32_t)std::min<uint64_t>(std::max<uint64_t>(32768, jt->cap / 8), 0x7fffffffu)
                                     : 0xffffffffu;
    if (jt)
        if (const char* env = std::getenv("GEQ_BUILD_CHUNK")) {   // small chunks exercise the path in tests
            const long long v = std::atoll(env);
            if (v >= 1) chunk = (uint32_t)std::min<long long>(v, 0x7fffffffll);
        }
    // Attack positions whose serial walk weighs at least heavy_min are
    // generated by a block each (k_generate_heavy); 0 turns this off.
    uint32_t heavy_min = 1024;
    if (const char* env = std::getenv("GEQ_BUILD_HEAVY_MIN")) {   // small values exercise the path in tests
        const long long v = std::atoll(env);
        if (v >= 0) heavy_min = (uint32_t)std::min<long long>(v, 0xffffffffll);
    }
    // Chunks whose frontier, candidates and next bounds fit tail_max items
    // finish in one block (k_level_tail); GEQ_BUILD_TAIL_MAX lowers the limit
    // (0: never), so tests take both paths.
    // further small levels k_level_tail may build in its block (GEQ_BUILD_TAIL_LOOP=0: none)
    uint32_t tail_loop = 4096;
    if (const char* env = std::getenv("GEQ_BUILD_TAIL_LOOP")) {
        const long long v = std::atoll(env);
        if (v >= 0) tail_loop = (uint32_t)std::min<long long>(v, 1ll << 20);
    }
    // the largest level it generates itself (GEQ_BUILD_TAIL_GEN, at most TAIL_GEN positions)
    uint32_t tail_gen = 128;
    if (const char* env = std::getenv("GEQ_BUILD_TAIL_GEN")) {
        const long long v = std::atoll(env);
        if (v >= 0) tail_gen = (uint32_t)std::min<long long>(v, TAIL_GEN);
    }
    if (!tail_gen) tail_loop = 0;
    uint32_t tail_max = TAIL_MAX;
    if (const char* env = std::getenv("GEQ_BUILD_TAIL_MAX")) {
        const long long v = std::atoll(env);
        if (v >= 0) tail_max = (uint32_t)std::min<long long>(v, TAIL_MAX);
    }
    // Positions are generated by SUB_LANES lanes each (k_generate_sub, with
    // k_generate_wave and k_generate_heavy for the larger ones), or by one
    // thread each with GEQ_BUILD_GEN=thread (k_generate); GEQ_BUILD_GEN=wave /
    // block send every attack position that fits to the wave / block kernel.
    uint32_t gen_mode = GEN_BY_SUB;
    if (const char* env = std::getenv("GEQ_BUILD_GEN")) {
        if (!std::strcmp(env, "thread")) gen_mode = GEN_BY_THREAD;
        else if (!std::strcmp(env, "wave")) gen_mode = GEN_BY_WAVE;
        else if (!std::strcmp(env, "block")) gen_mode = GEN_BY_BLOCK;
    }
    GB_TRY(ensure(ws.hv_count, 1, s, false, "heavy positions"));
    auto heavy_args = [&]() { return HeavyArgs{ws.hv_flag.p, ws.hv_list.p, ws.hv_count.p, heavy_min, gen_mode}; };
    // The last chunk of a level also bounds and scans the next level before
    // its wait, so the next level's first chunk starts without one.
    bool have_next = false;
    uint32_t next_cand = 0, next_words = 0, next_heavy = 0;
    if (trace)
        std::fprintf(stderr, "gamebuild: setup %.3f ms\n",
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    while (lo < hi) {
        const uint32_t c_hi = hi - lo > chunk ? lo + chunk : hi;
        const uint32_t nf = c_hi - lo;
        uint32_t end = c_hi;   // where the next chunk starts (after k_level_tail's own levels: their last one's end)
        GB_TRY(ensure(nsucc, nf + 1, s, false, "frontier"));
        GB_TRY(ensure(rowl, nf + 1, s, false, "frontier"));
        uint32_t n_cand = next_cand, n_words = next_words, n_heavy = next_heavy;
        if (!have_next) {
            GB_TRY(ensure(ub, nf + 1, s, false, "frontier"));
            GB_TRY(ensure(off, nf + 1, s, false, "frontier"));
            GB_TRY(ensure(ws.hv_flag, std::max<uint32_t>(nf, 1), s, false, "heavy positions"));
            GB_TRY(ensure(ws.hv_list, std::max<uint32_t>(nf, 1), s, false, "heavy positions"));
            GB_HIP(hipMemsetAsync(ws.hv_count.p, 0, 4, s), GEQ_ERR_TRANSFER);
            k_bounds<<<(nf + 1 + tb - 1) / tb, tb, 0, s>>>(nodes(), lo, nf, L, ub.p, heavy_args());
            GB_TRY(exclusive_scan_pairs(ub.p, off.p, nf, ws.scan2, s));
            const uint32_t sq = next_seq();
            k_readback<<<1, 64, 0, s>>>(&off.p[nf].x, &off.p[nf].y, ws.hv_count.p, ws.rb.d, sq);
            GB_TRY(wait(sq));
            n_cand = h_rb[RB_CAND];
            n_words = h_rb[RB_WORDS];
            n_heavy = h_rb[RB_HEAVY];
        }
        have_next = false;
        if (n_cand >= level_limit() || n_words >= level_limit())
            return bail(fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: a level of " + std::to_string(nf) +
                                                        " positions needs 2^32 or more candidates or set-pool words"));
        if ((uint64_t)B->spool_top + n_words >= 0xffffffffull) return bail(fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: set pool exceeds 2^32 words"));
        B->peak_words = std::max<uint64_t>(B->peak_words, B->spool_top + n_words + 1);
        GB_TRY(ensure(B->spool, B->spool_top + n_words + 1, s, true, "set pool"));
        GB_TRY(ensure(cands, std::max<uint32_t>(n_cand, 1), s, false, "candidates"));
        GB_TRY(ensure(flag, n_cand + 1, s, false, "candidates"));
        GB_TRY(ensure(rank, n_cand + 1, s, false, "candidates"));
        GB_TRY(node_cap((size_t)B->V + n_cand + 1));
        GB_TRY(table_cap((uint64_t)B->V + n_cand));
        if (E + n_cand >= 0xffffffffull) return bail(fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: more than 2^32 edges"));
        B->peak_edges = std::max<uint64_t>(B->peak_edges, E + n_cand + 1);
        GB_TRY(ensure(B->col, E + n_cand + 1, s, true, "edges"));
        GB_TRY(ensure(B->wt, E + n_cand + 1, s, true, "edges"));
        if (gen_mode != GEN_BY_THREAD) {
            if (nf >= WAVE_TAG) return bail(fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: a chunk of 2^31 or more positions"));
            k_generate_sub<<<(uint32_t)(((uint64_t)nf + 1 + SUB_BLOCK / SUB_LANES - 1) / (SUB_BLOCK / SUB_LANES)), SUB_BLOCK, 0, s>>>(
                nodes(), lo, nf, L, off.p, ws.hv_flag.p, B->spool_top, cands.p, nsucc.p);
        } else {
            k_generate<<<(nf + 1 + tb - 1) / tb, tb, gen_lds, s>>>(nodes(), lo, nf, L, off.p, ws.hv_flag.p, B->spool_top,
                                                                   cands.p, nsucc.p);
        }
        if (n_heavy) {   // k_bounds listed them (their count came back with the totals)
            if (gen_mode != GEN_BY_THREAD)
                k_generate_wave<<<std::min<uint32_t>((n_heavy + 3) / 4, 2048), SUB_BLOCK, 0, s>>>(nodes(), lo, L, off.p,
                                                                                               B->spool_top, heavy_args(),
                                                                                               cands.p, nsucc.p);
            if (gen_mode != GEN_BY_WAVE)
                k_generate_heavy<<<std::min<uint32_t>(n_heavy, 1024), 256, 0, s>>>(nodes(), lo, L, off.p, B->spool_top,
                                                                              heavy_args(), cands.p, nsucc.p);
        }
        // After a level's last chunk, the next level is the positions the level
        // created: `add` by its earlier chunks and *(rank + n_cand) by this one.
        const bool last_chunk = c_hi == hi;
        const uint32_t add = B->V - hi;
        const uint64_t slots = last_chunk ? (uint64_t)add + n_cand + 1 : 0;
        if (slots > 0xffffffffull) return bail(fail(GEQ_ERR_OUT_OF_MEMORY, "gamebuild: more than 2^32 positions"));
        if (last_chunk) {
            GB_TRY(ensure(ub, slots, s, false, "frontier"));
            GB_TRY(ensure(off, slots, s, false, "frontier"));
            GB_TRY(ensure(ws.hv_flag, slots, s, false, "heavy positions"));
            GB_TRY(ensure(ws.hv_list, slots, s, false, "heavy positions"));
        }
        const bool mark = jt && jt->cap;
        if (nf <= tail_max && n_cand <= tail_max && (!last_chunk || slots - 1 <= tail_max)) {
            // a small chunk: the rest of it in one block
            const uint32_t sq = next_seq();
            // a one-chunk level no larger than the levels the block generates may go
            // on into the following small levels in the same block, within the
            // buffers as they are now (k_level_tail_loop spills registers, so a
            // larger level that will not go on takes k_level_tail)
            const bool loop = gen_mode == GEN_BY_SUB && last_chunk && add == 0 && tail_loop && nf <= tail_gen;
            auto cap3 = [](size_t a, size_t b, size_t c) { return std::min(std::min(a, b), c); };
            TailRoom room{};
            if (loop) {
                room.front = (uint32_t)std::min<size_t>(std::min(nsucc.n, rowl.n), 0xffffffffu);
                room.cand = (uint32_t)std::min<size_t>(cap3(cands.n, flag.n, rank.n), 0xffffffffu);
                room.slots = (uint32_t)std::min<size_t>(std::min(cap3(ub.n, off.n, ws.hv_flag.n), ws.hv_list.n), 0xffffffffu);
                room.nodes = std::min(std::min(cap3(B->kind.n, B->p.n, B->q_off.n), cap3(B->q_len.n, B->x_off.n, B->x_len.n)),
                                      std::min<size_t>(B->hash.n, B->row.n ? B->row.n - 1 : 0));
                room.edges = std::min<size_t>(std::min(B->col.n, B->wt.n), 0xffffffffull);
                room.words = std::min<size_t>(B->spool.n, 0xffffffffull);
                room.table = table_slots;
                room.max_positions = jt ? 0 : max_positions;   // merged builds cap each job instead
                room.spool_top = (uint32_t)B->spool_top;
                room.n_words = n_words;
                room.tail_max = tail_max;
                room.gen_max = tail_gen;
            }
            const TailArgs T{nodes(), L, heavy_args(), cands.p, nsucc.p, rowl.p, flag.p, rank.p, table.p, keymin.p, B->col.p,
                             B->wt.p, B->row.p, ub.p, off.p, jt ? d_jcnt.p : nullptr, mark ? d_dead : nullptr, n_jobs,
                             mark ? jt->cap : 0ull, ws.rb.d, lo, nf, n_cand, B->V, (uint32_t)E, table_mask, hi, add,
                             (uint32_t)slots, chunk, last_chunk ? 1u : 0u, level_limit(), sq, loop ? tail_loop : 0u, room};
            if (loop) k_level_tail_loop<<<1, TAIL_BLOCK, 0, s>>>(T);
            else k_level_tail<<<1, TAIL_BLOCK, 0, s>>>(T);
            GB_TRY(wait(sq));
            if (const uint32_t more = h_rb[RB_LOOP]) {
                // the block built `more` further levels: the launched one is done, and the
                // count words below describe the last of them, which started here
                if (trace)
                    std::fprintf(stderr, "gamebuild: levels %llu-%llu in one block: %u frontier -> ... -> %u frontier\n",
                                 (unsigned long long)levels, (unsigned long long)(levels + more), nf, h_rb[RB_LOOP_NF]);
                levels += more;
                level_lo = h_rb[RB_LOOP_LO];
                lo = level_lo;
                hi = h_rb[RB_LOOP_V];
                end = hi;
                B->V = hi;
                E = h_rb[RB_LOOP_E];
                B->spool_top = h_rb[RB_LOOP_SP];
                n_words = h_rb[RB_LOOP_WORDS];
                const uint64_t lc = h_rb[RB_LOOP_CAND];
                B->peak_words = std::max<uint64_t>(B->peak_words, B->spool_top + n_words + 1);
                B->peak_edges = std::max<uint64_t>(B->peak_edges, E + lc + 1);
                B->peak_nodes = std::max<uint64_t>(B->peak_nodes, B->V + lc + 1);
            }
        } else {
            GB_TRY(exclusive_scan(nsucc.p, rowl.p, nf, ws.scan32, s));
            if (n_cand)
                k_insert<<<(n_cand + tb - 1) / tb, tb, 0, s>>>(cands.p, n_cand, nodes(), table.p, keymin.p, table_mask);
            // the winners' ranks (their flags computed inside the scan)
            GB_TRY(scan_marks(cands.p, n_cand, keymin.p, rank.p, flag.p, ws.scan32, s));
            // without per-job counts, the edges, rows and next bounds are one launch
            const bool fused = !jt && n_cand;
            if (fused) {
                k_finalize<<<(n_cand + tb - 1) / tb, tb, 0, s>>>(cands.p, n_cand, rank.p, B->V, nodes(), table.p, keymin.p,
                                                                 last_chunk ? ws.hv_count.p : nullptr);
                const uint64_t nt = std::max<uint64_t>(std::max<uint64_t>(n_cand, nf), slots);
                k_level_out<<<(uint32_t)((nt + tb - 1) / tb), tb, 0, s>>>(cands.p, n_cand, rowl.p, (uint32_t)E, table.p, B->col.p,
                                                                          B->wt.p, nf, lo, B->row.p, nodes(), hi, rank.p + n_cand,
                                                                          add, (uint32_t)slots, chunk, L, ub.p, heavy_args());
                if (last_chunk) GB_TRY(exclusive_scan_pairs(ub.p, off.p, slots - 1, ws.scan2, s));
            } else if (n_cand) {
                k_finalize<<<(n_cand + tb - 1) / tb, tb, 0, s>>>(cands.p, n_cand, rank.p, B->V, nodes(), table.p, keymin.p,
                                                                 nullptr);
                k_edges<<<(n_cand + tb - 1) / tb, tb, 0, s>>>(cands.p, n_cand, rowl.p, (uint32_t)E, table.p, B->col.p, B->wt.p);
                if (jt) {
                    // this level's new positions per job (the winners among the candidates)
                    const uint32_t nb = std::min<uint32_t>(1024, (n_cand + tb - 1) / tb);
                    k_job_count_cands<<<nb, tb, 0, s>>>(cands.p, rank.p, n_cand, d_job_of_state, n_jobs, d_jcnt.p);
                }
            }
            // the rows; the heavy count cleared for k_bounds_next; jobs past the cap marked
            if (!fused)
                k_rows<<<(nf + tb - 1) / tb, tb, 0, s>>>(rowl.p, nf, lo, (uint32_t)E, B->row.p,
                                                         RowsTail{last_chunk ? ws.hv_count.p : nullptr, mark ? d_jcnt.p : nullptr,
                                                                  mark ? n_jobs : 0u, mark ? jt->cap : 0ull, mark ? d_dead : nullptr});
            if (last_chunk && !fused) {
                k_bounds_next<<<(uint32_t)((slots + tb - 1) / tb), tb, 0, s>>>(nodes(), hi, rank.p + n_cand, add, (uint32_t)slots,
                                                                               chunk, L, ub.p, heavy_args());
                GB_TRY(exclusive_scan_pairs(ub.p, off.p, slots - 1, ws.scan2, s));
            }
            // one synchronisation per chunk for the edge and position counts, the per-job
            // counts and the next level's first totals
            const uint32_t sq = next_seq();
            k_chunk_counts<<<1, 256, 0, s>>>(rowl.p + nf, rank.p + n_cand, add, chunk, last_chunk ? off.p : nullptr,
                                             ws.hv_count.p, jt ? d_jcnt.p : nullptr, n_jobs, ws.rb.d, sq);
            GB_TRY(wait(sq));
        }
        const uint32_t n_edges = h_rb[RB_EDGES], n_new = h_rb[RB_NEW];
        if (last_chunk) {
            have_next = true;
            next_cand = h_rb[RB_CAND];
            next_words = h_rb[RB_WORDS];
            next_heavy = h_rb[RB_HEAVY];
        }
        E += n_edges;
        B->spool_top += n_words;
        B->V += n_new;
        lo = end;
        const bool level_done = lo == hi;
        if (level_done) {        // the next level is every position this one created
            if (trace)
                std::fprintf(stderr, "gamebuild: level %llu: %u frontier -> %u new positions, %llu positions, %llu edges, %.3f ms\n",
                             (unsigned long long)levels, hi - level_lo, B->V - hi, (unsigned long long)B->V,
                             (unsigned long long)E,
                             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            level_lo = hi;
            hi = B->V;
            levels++;
        }
        if (jt) {
            // positions of every job so far (read with the counts above); jobs past
            // the cap stop growing (k_rows set the same flags on the device)
            for (uint32_t j = 0; j < n_jobs; j++) {
                jt->count[j] = h_rb[RB_JOBS + j];
                if (jt->cap && jt->count[j] > jt->cap) jt->dead[j] = 1;
            }
        } else if (level_done && max_positions && B->V > max_positions) {
            truncated = true;
            break;
        }
    }
    if (truncated) {
        // the last level's new positions get no successors: close their rows
        const uint64_t rest = (uint64_t)B->V - lo + 1;
        k_fill_u32<<<(uint32_t)((rest + tb - 1) / tb), tb, 0, s>>>(B->row.p + lo, rest, (uint32_t)E);
    } else {
        const uint32_t e32 = (uint32_t)E;
        GB_HIP(hipMemcpyAsync(B->row.p + B->V, &e32, 4, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    B->E = (uint32_t)E;
    B->truncated = truncated;
    B->frontier = truncated ? lo : B->V;
    GB_TRY(ensure(B->atk, std::max<uint32_t>(B->V, 1), s, false, "attacker flags"));
    if (B->V) k_attacker_flags<<<(B->V + tb - 1) / tb, tb, 0, s>>>(B->kind.p, B->V, B->atk.p);
    GB_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->positions = B->V;
        stats->edges = E;
        stats->levels = levels;
        stats->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->truncated = truncated ? 1u : 0u;
        stats->set_words = B->spool_top;
    }
    (void)st;
    return GEQ_OK;
#undef GB_TRY
#undef GB_HIP
}

__global__ void k_pos_job(const uint32_t* p, uint32_t V, const uint32_t* job_of_state, uint32_t* key, uint32_t* val) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        key[v] = job_of_state[p[v]];
        val[v] = v;
    }
}
__global__ void k_sorted_deg(const uint32_t* order, const uint32_t* row, uint32_t V, uint32_t* dsort) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) dsort[i] = row[order[i] + 1] - row[order[i]];
    else if (i == V) dsort[i] = 0;
}
__global__ void k_at(const uint32_t* a, const uint32_t* idx, uint32_t n, uint32_t* out) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = a[idx[j]];
}
__global__ void k_local_ids(const uint32_t* order, const uint32_t* key, const uint32_t* pos_start, uint32_t V, uint32_t* local) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) local[order[i]] = i - pos_start[key[i]];
}
// job-major rows: job j's row entries sit at row_out[pos_start[j] + j ...], relative to its first edge
__global__ void k_job_rows(const uint32_t* key, const uint32_t* escan, const uint32_t* pos_start, const uint32_t* edge_start,
                           uint32_t V, uint32_t* row_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const uint32_t j = key[i];
    row_out[i + j] = escan[i] - edge_start[j];
    if (i + 1 == pos_start[j + 1]) row_out[i + j + 1] = escan[i + 1] - edge_start[j];
}
__global__ void k_job_edges(const uint32_t* order, const uint32_t* row, const uint32_t* col, const uint32_t* wt,
                            const uint8_t* atk, const uint32_t* local, const uint32_t* escan, uint32_t V, uint32_t* col_out,
                            uint32_t* wt_out, uint8_t* atk_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const uint32_t v = order[i], r0 = row[v], deg = row[v + 1] - r0, b = escan[i];
    for (uint32_t k = 0; k < deg; k++) {
        col_out[b + k] = local[col[r0 + k]];
        wt_out[b + k] = wt[r0 + k];
    }
    atk_out[i] = atk[v];
}

}  // namespace

geq_status geq::gamebuild_jobs(geq_ctx* ctx, const std::vector<geq_lts>& ltss, const std::vector<const uint32_t*>& pairs,
                               const std::vector<uint32_t>& n_pairs, uint64_t cap, BuildScratch& ws, geq_built& store,
                               BuiltJobs* out) {
    const uint32_t J = (uint32_t)ltss.size();
    if (J == 0 || J > 1024) return fail(GEQ_ERR_INVALID_ARG, "gamebuild_jobs: 1..1024 jobs");
    const auto tp0 = std::chrono::steady_clock::now();
    // disjoint union of the jobs' LTSs; pairs shifted to the merged state ids
    uint64_t n = 0, m = 0, np = 0;
    for (uint32_t j = 0; j < J; j++) {
        n += ltss[j].n_states;
        m += ltss[j].n_transitions;
        np += n_pairs[j];
    }
    if (n >= (1u << 30) || m >= 0xffffffffull || np >= (1u << 30)) return fail(GEQ_ERR_INVALID_ARG, "gamebuild_jobs: batch too large");
    std::vector<uint32_t> row(n + 1), dst(std::max<uint64_t>(m, 1)), mp(std::max<uint64_t>(2 * np, 2));
    std::vector<int32_t> lbl(std::max<uint64_t>(m, 1));
    JobTrack jt;
    jt.job_of_state.resize(n);
    jt.count.assign(J, 0);
    jt.dead.assign(J, 0);
    jt.cap = cap;
    uint32_t sb = 0, tb_ = 0, pb = 0;
    row[0] = 0;
    for (uint32_t j = 0; j < J; j++) {
        const geq_lts& l = ltss[j];
        for (uint32_t st = 0; st < l.n_states; st++) {
            row[sb + st + 1] = tb_ + l.row_offsets[st + 1];
            jt.job_of_state[sb + st] = j;
        }
        for (uint32_t k = 0; k < l.n_transitions; k++) {
            dst[tb_ + k] = sb + l.targets[k];
            lbl[tb_ + k] = l.labels[k];
        }
        for (uint32_t i = 0; i < 2 * n_pairs[j]; i++) mp[2 * pb + i] = sb + pairs[j][i];
        sb += l.n_states;
        tb_ += l.n_transitions;
        pb += n_pairs[j];
    }
    geq_lts merged{(uint32_t)n, (uint32_t)m, row.data(), dst.data(), lbl.data()};
    RetireScope retire_scope(ctx->stream, &ws.spare);   // outgrown buffers are kept for later builds
    const auto tp1 = std::chrono::steady_clock::now();
    geq_build_stats bs{};
    geq_built* B = &store;
    geq_status st = build_core(ctx, &merged, mp.data(), (uint32_t)np, 0, &jt, B, ws, &bs);
    if (st != GEQ_OK) return st;
    const uint32_t V = B->V, E = B->E, tb = 256;
    uint64_t total = 0;
    for (uint64_t c : jt.count) total += c;
    if (total != V) return fail(GEQ_ERR_INTERNAL, "gamebuild_jobs: per-job position counts do not add up");
    hipStream_t s = ctx->stream;
    // stable sort of the positions by job: order[i] = merged id of the i-th position in job-major order
    auto& key = ws.key; auto& key_s = ws.key_s; auto& val = ws.val; auto& order = ws.order; auto& dsort = ws.dsort;
    auto& escan = ws.escan; auto& local = ws.local; auto& d_ps = ws.d_ps; auto& d_es = ws.d_es;
    for (auto* b : {&key, &key_s, &val, &order, &local})
        if ((st = ensure(*b, V, s, false, "job extraction")) != GEQ_OK) return st;
    for (auto* b : {&dsort, &escan})
        if ((st = ensure(*b, V + 1, s, false, "job extraction")) != GEQ_OK) return st;
    for (auto* b : {&d_ps, &d_es})
        if ((st = ensure(*b, J + 1, s, false, "job extraction")) != GEQ_OK) return st;
    k_pos_job<<<(V + tb - 1) / tb, tb, 0, s>>>(B->p.p, V, ws.job_map, key.p, val.p);
    int end_bit = 1;
    while ((1u << end_bit) < J) end_bit++;
    // dsort and local are free until after the sort: they hold its middle passes
    if ((st = radix_sort_pairs(key.p, val.p, key_s.p, order.p, dsort.p, local.p, V, (uint32_t)end_bit, ws.sort, s)) != GEQ_OK)
        return st;
    k_sorted_deg<<<(V + 1 + tb - 1) / tb, tb, 0, s>>>(order.p, B->row.p, V, dsort.p);
    st = exclusive_scan(dsort.p, escan.p, V, ws.scan32, s);
    if (st != GEQ_OK) return st;
    out->pos_start.assign(J + 1, 0);
    out->n_pos.assign(J, 0);
    for (uint32_t j = 0; j < J; j++) {
        out->n_pos[j] = (uint32_t)jt.count[j];
        out->pos_start[j + 1] = out->pos_start[j] + out->n_pos[j];
    }
    GEQ_HIP(hipMemcpyAsync(d_ps.p, out->pos_start.data(), 4ull * (J + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    k_at<<<(J + 1 + tb - 1) / tb, tb, 0, s>>>(escan.p, d_ps.p, J + 1, d_es.p);
    out->edge_start.assign(J + 1, 0);
    GEQ_HIP(hipMemcpyAsync(out->edge_start.data(), d_es.p, 4ull * (J + 1), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    k_local_ids<<<(V + tb - 1) / tb, tb, 0, s>>>(order.p, key_s.p, d_ps.p, V, local.p);
    if ((st = ensure(out->row, (size_t)V + J, s, false, "job games")) != GEQ_OK) return st;
    if ((st = ensure(out->col, std::max<uint32_t>(E, 1), s, false, "job games")) != GEQ_OK) return st;
    if ((st = ensure(out->wt, std::max<uint32_t>(E, 1), s, false, "job games")) != GEQ_OK) return st;
    if ((st = ensure(out->atk, std::max<uint32_t>(V, 1), s, false, "job games")) != GEQ_OK) return st;
    k_job_rows<<<(V + tb - 1) / tb, tb, 0, s>>>(key_s.p, escan.p, d_ps.p, d_es.p, V, out->row.p);
    k_job_edges<<<(V + tb - 1) / tb, tb, 0, s>>>(order.p, B->row.p, B->col.p, B->wt.p, B->atk.p, local.p, escan.p, V, out->col.p,
                                                 out->wt.p, out->atk.p);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    out->n_edges.assign(J, 0);
    for (uint32_t j = 0; j < J; j++) out->n_edges[j] = out->edge_start[j + 1] - out->edge_start[j];
    out->dead = jt.dead;
    out->levels = (uint32_t)bs.levels;
    out->build_ms = bs.build_ms;
    out->prep_ms = std::chrono::duration<double, std::milli>(tp1 - tp0).count();
    out->extract_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp1).count() - bs.build_ms;
    return GEQ_OK;
}

extern "C" {

geq_status geq_gamebuild(geq_ctx* ctx, const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs, uint64_t max_positions,
                         geq_built** out, geq_build_stats* stats) {
    if (!ctx || !lts || !out || (n_pairs && !pairs)) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    geq_status st = validate_job(lts, pairs, n_pairs);
    if (st != GEQ_OK) return st;
    // the context's scratch when no other build holds it
    std::unique_lock<std::mutex> lk(ctx->build_mu, std::try_to_lock);
    std::unique_ptr<BuildScratch> own;
    BuildScratch* ws = nullptr;
    if (lk.owns_lock()) {
        if (!ctx->build_ws) ctx->build_ws = std::make_shared<BuildScratch>();
        ws = ctx->build_ws.get();
    } else {
        own = std::make_unique<BuildScratch>();
        ws = own.get();
    }
    geq_built* B = new geq_built();
    st = build_core(ctx, lts, pairs, n_pairs, max_positions, nullptr, B, *ws, stats);
    if (st != GEQ_OK) {
        delete B;
        return st;
    }
    ctx->hint_positions = B->peak_nodes;
    ctx->hint_edges = B->peak_edges;
    ctx->hint_words = B->peak_words;
    *out = B;
    return GEQ_OK;
}

geq_status geq_built_download(geq_built* B, geq_built_game* o) {
    if (!B || !o) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    (void)hipSetDevice(B->ctx->device);
    hipStream_t s = B->ctx->stream;
    std::memset(o, 0, sizeof(*o));
    const uint32_t V = B->V, E = B->E;
    o->n_positions = V;
    o->n_edges = E;
    o->n_start = B->n_start;
    o->pool_len = B->spool_top;
    auto a32 = [](size_t n) { return (uint32_t*)std::malloc(4 * std::max<size_t>(n, 1)); };
    o->row_offsets = a32(V + 1);
    o->column_indices = a32(E);
    o->weights = a32(E);
    o->attacker_pos = (uint8_t*)std::malloc(std::max<size_t>(V, 1));
    o->kind = (uint8_t*)std::malloc(std::max<size_t>(V, 1));
    o->p = a32(V);
    o->q_off = a32(V);
    o->q_len = a32(V);
    o->x_off = a32(V);
    o->x_len = a32(V);
    o->pool = a32(B->spool_top);
    if (!o->row_offsets || !o->column_indices || !o->weights || !o->attacker_pos || !o->kind || !o->p || !o->q_off || !o->q_len ||
        !o->x_off || !o->x_len || !o->pool) {
        geq_built_game_free(o);
        return fail(GEQ_ERR_OUT_OF_MEMORY, "host allocation of the built game");
    }
    auto dn = [&](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = dn(o->row_offsets, B->row.p, 4ull * (V + 1));
    if (e == hipSuccess) e = dn(o->column_indices, B->col.p, 4ull * E);
    if (e == hipSuccess) e = dn(o->weights, B->wt.p, 4ull * E);
    if (e == hipSuccess) e = dn(o->attacker_pos, B->atk.p, V);
    if (e == hipSuccess) e = dn(o->kind, B->kind.p, V);
    if (e == hipSuccess) e = dn(o->p, B->p.p, 4ull * V);
    if (e == hipSuccess) e = dn(o->q_off, B->q_off.p, 4ull * V);
    if (e == hipSuccess) e = dn(o->q_len, B->q_len.p, 4ull * V);
    if (e == hipSuccess) e = dn(o->x_off, B->x_off.p, 4ull * V);
    if (e == hipSuccess) e = dn(o->x_len, B->x_len.p, 4ull * V);
    if (e == hipSuccess) e = dn(o->pool, B->spool.p, 4ull * B->spool_top);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        geq_built_game_free(o);
        return fail(GEQ_ERR_TRANSFER, std::string("downloading the built game: ") + hipGetErrorString(e));
    }
    return GEQ_OK;
}

void geq_built_game_free(geq_built_game* o) {
    if (!o) return;
    std::free(o->row_offsets); std::free(o->column_indices); std::free(o->weights); std::free(o->attacker_pos);
    std::free(o->kind); std::free(o->p); std::free(o->q_off); std::free(o->q_len); std::free(o->x_off); std::free(o->x_len);
    std::free(o->pool);
    std::memset(o, 0, sizeof(*o));
}

namespace {
geq_status built_solve(geq_built* B, int profile_kernels, geq_stats* stats) {
    (void)hipSetDevice(B->ctx->device);
    if (!B->game) {
        geq_game* g = nullptr;
        geq_status st = game_from_device(B->ctx, B->V, B->E, B->row.p, B->col.p, B->wt.p, B->atk.p, &g);
        if (st != GEQ_OK) return st;
        // the frontier of a capped build is never a position to reach
        g->reach_limit = B->frontier;
        st = standard_seed(g);
        if (st != GEQ_OK) {
            release_built_game(g);
            return st;
        }
        B->game = g;
        B->n_reach = g->std_n;
    }
    return solve_core(B->game, B->n_reach, nullptr, profile_kernels, stats, B->game->std_seed.p);
}
}  // namespace

geq_status geq_built_solve(geq_built* B, int profile_kernels, geq_stats* stats) {
    if (!B) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if (B->truncated)
        return fail(GEQ_ERR_INVALID_ARG, "the build stopped at max_positions: positions " + std::to_string(B->frontier) +
                                             ".. were never expanded, so this is not the complete game "
                                             "(geq_built_solve_capped solves it with that frontier left out)");
    return built_solve(B, profile_kernels, stats);
}

geq_status geq_built_solve_capped(geq_built* B, int profile_kernels, geq_stats* stats) {
    if (!B) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    return built_solve(B, profile_kernels, stats);
}

geq_status geq_built_frontier(geq_built* B, uint32_t* truncated, uint32_t* frontier) {
    if (!B) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if (truncated) *truncated = B->truncated ? 1u : 0u;
    if (frontier) *frontier = B->frontier;
    return GEQ_OK;
}

geq_status geq_built_energies(geq_built* B, uint32_t first, uint32_t count, geq_energies* out) {
    if (!B || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if (!B->game) return fail(GEQ_ERR_INVALID_ARG, "the built game was not solved");
    return download_energies(B->game, first, count, out);
}

geq_status geq_built_preorders(geq_built* B, uint32_t first, uint32_t count, const uint32_t* energies, uint32_t n_energies,
                               uint8_t* out) {
    if (!B) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if (!B->game) return fail(GEQ_ERR_INVALID_ARG, "the built game was not solved");
    return game_preorders(B->game, first, count, energies, n_energies, out);
}

void geq_built_free(geq_built* B) {
    if (!B) return;
    geq_ctx* ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    {
        // The game's position, edge and set arrays become spares of the
        // context's build scratch, which the next build takes instead of
        // allocating (and each hipFree would wait for the whole device).  Work
        // queued on the context's stream may still read them: it ends first.
        std::unique_lock<std::mutex> lk(ctx->build_mu, std::try_to_lock);
        if (lk.owns_lock() && ctx->build_ws && hipStreamSynchronize(ctx->stream) == hipSuccess) {
            SpareBufs& pool = ctx->build_ws->spare;
            auto give = [&](auto& b) {
                if (b.p) pool.bufs.push_back(SpareBufs::Buf{b.p, sizeof(*b.p) * std::max<size_t>(b.n, 1)});
                b.p = nullptr;
                b.n = 0;
            };
            give(B->kind); give(B->atk); give(B->p); give(B->q_off); give(B->q_len); give(B->x_off); give(B->x_len);
            give(B->hash); give(B->spool); give(B->row); give(B->col); give(B->wt);
            pool.trim();
        }
    }
    delete B;
}

}  // extern "C"
