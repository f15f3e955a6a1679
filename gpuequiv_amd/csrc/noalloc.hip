// geq_solve_noalloc: geq_solve (EnergyGame::run) without any device
// allocation, for callers that cannot allocate (code under graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  Every device buffer
// of the solve -- the graph in the solver's layout, the reverse CSR, node
// records, worklists, queued bitmaps, the energy arena, statistics, the
// control block and the output staging -- is a region of one file-scope
// device array, `scratch`, of NOALLOC_SCRATCH_BYTES.  Each device the library
// runs on holds its own copy of the array until the process ends; its device
// address is looked up once per device, when a context is created on it.
//
// Calls on one device share that device's array and would race; a per-device
// lock inside the library serialises them within the process.  Every region
// a solve needs cleared is cleared by a fill kernel at the start of the call
// (the array keeps the previous call's contents).
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "gamebuild.hpp"
#include "runtime.hpp"

namespace geq {

constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace geq

static __device__ __align__(256) unsigned char scratch[geq::NOALLOC_SCRATCH_BYTES];

namespace geq {

namespace {

struct DeviceScratch {
    unsigned char* ptr = nullptr;   // device address of `scratch` on this device
    std::mutex mu;                  // one call at a time uses it
};
std::array<DeviceScratch, NOALLOC_MAX_DEVICES> g_scratch;
std::mutex g_setup_mu;

// In-place exclusive scan of a[0..n) by one block of 1024 threads (each thread
// a contiguous share); a[n] receives the total.
__global__ __launch_bounds__(1024) void k_scan_one_block(uint32_t* a, uint32_t n) {
    __shared__ uint32_t s_tot[1024 / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = min(n, tid * per), e = min(n, b + per);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += a[i];
    const uint32_t incl = wave_incl_scan(sum, lane);
    if (lane == 63u) s_tot[wid] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < 1024u / 64u; w++) {
        base += w < wid ? s_tot[w] : 0u;
        total += s_tot[w];
    }
    uint32_t run = base + incl - sum;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
    if (tid == 0) a[n] = total;
}

__global__ void k_fill(uint32_t* a, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

// edges[e] = (column, update in the canonical nibbles), the in-degree counts
// of the reverse CSR, and the attacker flags as bytes.
__global__ void k_na_edges(const uint32_t* col, const uint32_t* wt, uint32_t E, uint32_t elements, uint32_t ubits,
                           uint2* edges, uint32_t* indeg) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t w = wt[e], mask = (1u << ubits) - 1u;
    uint32_t canon = 0;
    for (uint32_t i = 0; i < elements; i++) canon |= ((w >> (i * ubits)) & mask) << (4 * i);
    edges[e] = make_uint2(col[e], canon);
    atomicAdd(&indeg[col[e]], 1u);
}

__global__ void k_na_rev(const uint32_t* row, uint32_t V, const uint2* edges, uint32_t* fill, uint32_t* rcol) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (uint32_t e = row[v]; e < row[v + 1]; e++) rcol[atomicAdd(&fill[edges[e].x], 1u)] = v;
}

// A fresh solve: empty list slots, node v in `reach` at the zero energy
// (slot 0 current), no node queued, zero statistics, a fresh control block.
__global__ void k_na_init(uint4* st, uint32_t V, uint32_t* queued, uint32_t qwords2, unsigned long long* stats, uint32_t nstats,
                          Ctl* ctl) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (uint32_t v = tid; v < V; v += nt) {
        st[slot_index(v, 0)] = make_uint4(0, 0, 0, 0);
        st[slot_index(v, 1)] = make_uint4(0, 0, 0, 0);
        st[slot_index(v, 2)] = make_uint4(0, 0, 0, 0);
    }
    for (uint32_t i = tid; i < qwords2; i += nt) queued[i] = 0u;
    for (uint32_t i = tid; i < nstats; i += nt) stats[i] = 0ull;
    if (tid == 0) *ctl = Ctl{};
}
__global__ void k_na_seed(uint4* st, const uint32_t* ids, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st[slot_index(ids[i], 0)] = make_uint4(1u, 0u, 0u, 0u);   // one all-zero energy
}

// Regions of the scratch array for a game of V nodes and E edges.
struct Layout {
    size_t ctl, stats, st, offs, row, rrow, rcol, edges, atk, wl0, wl1, chg, queued, seeds, arena, end;
    uint32_t arena_cap;     // u16 entries
};
size_t up256(size_t x) { return (x + 255) & ~size_t(255); }
bool plan(uint32_t V, uint32_t E, uint32_t n_reach, Layout& L) {
    size_t p = 0;
    auto take = [&](size_t bytes) {
        const size_t at = p;
        p = up256(p + bytes);
        return at;
    };
    L.ctl = take(sizeof(Ctl));
    L.stats = take(8ull * (4ull * STAT_SLOTS + 4));
    L.st = take(16ull * NODE_U4 * V);
    L.offs = take(4ull * (V + 1));
    L.row = take(4ull * (V + 1));
    L.rrow = take(4ull * (V + 1));
    L.rcol = take(4ull * std::max<uint32_t>(E, 1));
    L.edges = take(8ull * std::max<uint32_t>(E, 1));
    L.atk = take(std::max<uint32_t>(V, 1));
    L.wl0 = take(4ull * (V + 1));
    L.wl1 = take(4ull * (V + 1));
    L.chg = take(std::max<uint32_t>(V, 1));
    L.queued = take(8ull * queue_words(V));
    L.seeds = take(4ull * std::max<uint32_t>(n_reach, 1));
    L.arena = p;
    // the arena takes the rest; it first stages the raw columns and updates
    if (p >= NOALLOC_SCRATCH_BYTES) return false;
    const size_t rest = NOALLOC_SCRATCH_BYTES - p;
    if (rest < std::max<size_t>(8ull * E, 1u << 17)) return false;
    L.arena_cap = (uint32_t)std::min<size_t>(rest / 2, 0x7fffffffull);
    L.end = NOALLOC_SCRATCH_BYTES;
    return true;
}

}  // namespace

// Set-up for a context on `device` (the current device): the device address of
// `scratch` on it, looked up once per device.
geq_status noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return GEQ_OK;   // the entry point then reports it
    std::lock_guard<std::mutex> lk(g_setup_mu);
    DeviceScratch& d = g_scratch[device];
    if (d.ptr) return GEQ_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(GEQ_ERR_NO_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = static_cast<unsigned char*>(p);
    return GEQ_OK;
}

}  // namespace geq

using namespace geq;

extern "C" geq_status geq_solve_noalloc(geq_ctx* ctx, const geq_graph* g, const uint32_t* to_reach, uint32_t n_reach,
                                        uint32_t* out_offsets, uint32_t* out_data, uint64_t out_cap, uint64_t* out_n,
                                        geq_stats* stats) {
    if (!ctx || !g || !out_offsets || (out_cap && !out_data) || (n_reach && !to_reach))
        return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    ConfInfo ci;
    geq_status st = check_conf(g->conf, ci);
    if (st != GEQ_OK) return st;
    if (ci.generic)
        return fail(GEQ_ERR_UNSUPPORTED, "geq_solve_noalloc takes energy configurations with elements <= 6 and max <= 3");
    const uint32_t V = g->n_vertices, E = g->n_edges;
    // the checks of geq_solve's upload, on the caller's arrays (nothing is copied)
    if (V == 0) return fail(GEQ_ERR_INVALID_ARG, "empty game graph");
    if (!g->row_offsets || !g->attacker_pos || (E && (!g->column_indices || !g->weights)))
        return fail(GEQ_ERR_INVALID_ARG, "missing graph array");
    if (g->row_offsets[0] != 0 || g->row_offsets[V] != E) return fail(GEQ_ERR_INVALID_ARG, "row_offsets must start at 0 and end at n_edges");
    for (uint32_t v = 0; v < V; v++)
        if (g->row_offsets[v + 1] < g->row_offsets[v]) return fail(GEQ_ERR_INVALID_ARG, "row_offsets not monotone");
    for (uint32_t k = 0; k < E; k++) {
        uint32_t canon;
        if (g->column_indices[k] >= V) return fail(GEQ_ERR_INVALID_ARG, "column index out of range at edge " + std::to_string(k));
        if (!update_to_canon(ci, g->weights[k], canon)) return fail(GEQ_ERR_INVALID_ARG, "invalid update at edge " + std::to_string(k));
    }
    for (uint32_t i = 0; i < n_reach; i++)
        if (to_reach[i] >= V) return fail(GEQ_ERR_INVALID_ARG, "to_reach id out of range");
    Layout L;
    if (!plan(V, E, n_reach, L))
        return fail(GEQ_ERR_OUT_OF_MEMORY, "game of " + std::to_string(V) + " positions and " + std::to_string(E) +
                                               " edges does not fit geq_solve_noalloc's static scratch of " +
                                               std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
    const int dev = ctx->device;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES || !g_scratch[dev].ptr)
        return fail(GEQ_ERR_NO_DEVICE, "no static scratch was set up on device " + std::to_string(dev));
    DeviceScratch& D = g_scratch[dev];
    std::lock_guard<std::mutex> lk(D.mu);
    (void)hipSetDevice(dev);
    hipStream_t s = ctx->stream;
    unsigned char* base = D.ptr;   // never the array's name in host code
    auto at = [&](size_t off) { return base + off; };
    Ctl* ctl = reinterpret_cast<Ctl*>(at(L.ctl));
    unsigned long long* sts = reinterpret_cast<unsigned long long*>(at(L.stats));
    uint4* nodes = reinterpret_cast<uint4*>(at(L.st));
    uint32_t* offs = reinterpret_cast<uint32_t*>(at(L.offs));
    uint32_t* row = reinterpret_cast<uint32_t*>(at(L.row));
    uint32_t* rrow = reinterpret_cast<uint32_t*>(at(L.rrow));
    uint32_t* rcol = reinterpret_cast<uint32_t*>(at(L.rcol));
    uint2* edges = reinterpret_cast<uint2*>(at(L.edges));
    uint8_t* atk = reinterpret_cast<uint8_t*>(at(L.atk));
    uint32_t* wl0 = reinterpret_cast<uint32_t*>(at(L.wl0));
    uint32_t* wl1 = reinterpret_cast<uint32_t*>(at(L.wl1));
    uint8_t* chg = reinterpret_cast<uint8_t*>(at(L.chg));
    uint32_t* queued = reinterpret_cast<uint32_t*>(at(L.queued));
    uint32_t* seeds = reinterpret_cast<uint32_t*>(at(L.seeds));
    uint16_t* arena = reinterpret_cast<uint16_t*>(at(L.arena));
    uint32_t* raw_col = reinterpret_cast<uint32_t*>(at(L.arena));   // staged in the arena before the solve
    uint32_t* raw_wt = raw_col + std::max<uint32_t>(E, 1);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t tb = 256;
    // ---- the graph in the solver's layout
    GEQ_HIP(hipMemcpyAsync(row, g->row_offsets, 4ull * (V + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemcpyAsync(atk, g->attacker_pos, V, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    if (E) {
        GEQ_HIP(hipMemcpyAsync(raw_col, g->column_indices, 4ull * E, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipMemcpyAsync(raw_wt, g->weights, 4ull * E, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    if (n_reach) GEQ_HIP(hipMemcpyAsync(seeds, to_reach, 4ull * n_reach, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    // in-degree counts into rrow, scanned in place; the fill counters of the scatter in wl0
    k_fill<<<std::min<uint32_t>((V + 1 + tb - 1) / tb, 4096), tb, 0, s>>>(rrow, V + 1ull, 0u);
    if (E) k_na_edges<<<(E + tb - 1) / tb, tb, 0, s>>>(raw_col, raw_wt, E, ci.elements, ci.ubits, edges, rrow);
    k_scan_one_block<<<1, 1024, 0, s>>>(rrow, V);
    GEQ_HIP(hipMemcpyAsync(wl0, rrow, 4ull * V, hipMemcpyDeviceToDevice, s), GEQ_ERR_TRANSFER);
    k_na_rev<<<(V + tb - 1) / tb, tb, 0, s>>>(row, V, edges, wl0, rcol);
    k_node_hdr<<<(V + tb - 1) / tb, tb, 0, s>>>(row, rrow, atk, V, nodes);
    // ---- a fresh solve (EnergyGame::with_reach, src/energygame.rs:355-365)
    const uint32_t q2 = 2u * queue_words(V), ns = 4u * STAT_SLOTS + 4u;
    k_na_init<<<std::min<uint32_t>((std::max(std::max(V, q2), ns) + tb - 1) / tb, 4096), tb, 0, s>>>(nodes, V, queued, q2, sts, ns,
                                                                                                    ctl);
    if (n_reach) k_na_seed<<<(n_reach + tb - 1) / tb, tb, 0, s>>>(nodes, seeds, n_reach);
    const uint32_t ncu = (uint32_t)std::max(1, ctx->n_cu);
    const uint32_t waves = MAX_WAVES, threads = 64u * waves;
    uint32_t blocks = std::min(ncu * 8u, ncu * eval_blocks_resident(true, waves));
    blocks = std::max(1u, std::min(std::min(blocks, (V + waves - 1) / waves), STAT_SLOTS));
    const EvalArgs ea{{wl0, wl1}, row, edges, atk, nodes, arena, ctl, L.arena_cap, chg, ci.max, sts, rrow, rcol, queued,
                      queue_words(V)};
    ExpandArgs xa{{wl0, wl1}, seeds, n_reach, chg, rrow, rcol, queued, queue_words(V), ctl};
    // the seeds' predecessors form round 0's worklist (initialize_visit_lists, :1461-1465)
    if (n_reach) k_expand<<<std::max(1u, std::min(ncu * 4u, (n_reach + BLOCK - 1) / BLOCK)), BLOCK, 0, s>>>(xa, 0u);
    // ---- rounds (GPURunner::execute_gpu, :1481-1523): batches of R rounds, one readback each
    const uint32_t R = 8;
    Ctl h{};
    h.wl_n[0] = n_reach ? 1u : 0u;
    uint32_t retag_round = 0;
    const uint32_t retag_every = retag_interval();
    while (h.wl_n[h.round % 3u] != 0) {
        for (uint32_t k = 0; k < R; k++)
            k_eval<true><<<blocks, threads, waves * EVAL_FUSED_LDS_PER_WAVE, s>>>(ea, k);
        k_advance<<<1, 64, 0, s>>>(ctl, R);
        GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
        GEQ_HIP(hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        if (h.overflow)
            return fail(GEQ_ERR_OUT_OF_MEMORY, "geq_solve_noalloc: the energy arena in the static scratch (" +
                                                   std::to_string(L.arena_cap) + " entries) is exhausted");
        if (h.round - retag_round >= retag_every) {
            k_retag<<<(V + tb - 1) / tb, tb, 0, s>>>(nodes, V, round_tag(h.round));
            retag_round = h.round;
        }
    }
    k_reduce_stats<<<1, 256, 0, s>>>(sts, blocks, ctl->totals);
    // ---- output: list lengths scanned into offsets, lists gathered into the
    // regions the solve no longer needs (from rcol on), ascending per node
    const uint32_t t = round_tag(h.round);
    k_list_lens<<<(V + tb - 1) / tb, tb, 0, s>>>(nodes, V, t, offs, nullptr);
    k_scan_one_block<<<1, 1024, 0, s>>>(offs, V);
    GEQ_HIP(hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemcpyAsync(out_offsets, offs, 4ull * (V + 1), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    const uint64_t tot = out_offsets[V];
    if (out_n) *out_n = tot;
    const size_t staging = (L.arena - L.rcol) / 4;   // words from rcol up to the arena
    if (tot > staging)
        return fail(GEQ_ERR_OUT_OF_MEMORY, "geq_solve_noalloc: " + std::to_string(tot) + " result energies exceed the static scratch");
    if (tot > out_cap)
        return fail(GEQ_ERR_INVALID_ARG, "out_data holds " + std::to_string(out_cap) + " energies, the result has " + std::to_string(tot) +
                                             " (*out_n)");
    uint32_t* stage = rcol;
    k_gather<<<(V + tb - 1) / tb, tb, 0, s>>>(nodes, arena, V, t, offs, stage, staging, nullptr);
    if (tot) GEQ_HIP(hipMemcpyAsync(out_data, stage, 4ull * tot, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    if (ci.ebits != 2)
        for (uint64_t i = 0; i < tot; i++) out_data[i] = energy_from_canon(ci, out_data[i]);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rounds = h.last_nonempty;
        stats->node_evals = h.totals[0];
        stats->edge_updates = h.totals[1];
        stats->alg_bytes = h.totals[2];
        stats->changed = h.totals[3];
        stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->mode = 3;
    }
    return GEQ_OK;
}
