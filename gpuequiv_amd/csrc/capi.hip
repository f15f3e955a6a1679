// C ABI of the solver (include/gpuequiv.h): context, game upload, the
// device-resident fixpoint loop and result download.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace geq {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
geq_status fail(geq_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

static uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
static uint32_t bitlen(uint32_t x) { return x == 0 ? 0 : 32u - (uint32_t)__builtin_clz(x); }

// EnergyConf arithmetic of src/energy.rs:48-78
geq_status check_conf(const geq_conf& c, ConfInfo& o) {
    if (c.elements == 0 || c.elements > 6 || c.max > 3)
        return fail(GEQ_ERR_UNSUPPORTED, "energy configuration outside the kernels (need 1 <= elements <= 6, max <= 3): elements=" +
                                             std::to_string(c.elements) + " max=" + std::to_string(c.max));
    o.elements = c.elements;
    o.max = c.max;
    o.ebits = next_pow2(bitlen(c.max));
    o.emask = 0xffffffffu >> (32 - o.ebits);
    o.ubits = next_pow2(bitlen(c.elements + 2));
    o.umask = 0xffffffffu >> (32 - o.ubits);
    return GEQ_OK;
}

bool energy_to_canon(const ConfInfo& c, uint32_t packed, uint32_t& canon) {
    // all fields beyond `elements` must be zero, every field <= max
    if (c.ebits * c.elements < 32 && (packed >> (c.ebits * c.elements)) != 0) return false;
    canon = 0;
    for (uint32_t i = 0; i < c.elements; i++) {
        uint32_t v = (packed >> (i * c.ebits)) & c.emask;
        if (v > c.max) return false;
        canon |= v << (2 * i);
    }
    return true;
}

uint32_t energy_from_canon(const ConfInfo& c, uint32_t canon) {
    uint32_t out = 0;
    for (uint32_t i = 0; i < c.elements; i++) out |= ((canon >> (2 * i)) & 3u) << (i * c.ebits);
    return out;
}

bool update_to_canon(const ConfInfo& c, uint32_t packed, uint32_t& canon) {
    if (c.ubits * c.elements < 32 && (packed >> (c.ubits * c.elements)) != 0) return false;
    canon = 0;
    for (uint32_t i = 0; i < c.elements; i++) {
        uint32_t v = (packed >> (i * c.ubits)) & c.umask;
        if (v > c.elements + 1) return false;   // Min(k) needs k <= elements (src/energy.rs:551-556)
        canon |= v << (4 * i);
    }
    return true;
}

}  // namespace geq

using namespace geq;

geq_game::~geq_game() {
    if (h_ctl) (void)hipHostFree(h_ctl);
}

extern "C" {

const char* geq_last_error(void) { return g_last_error.c_str(); }

geq_status geq_ctx_create(int device, geq_ctx** out) {
    if (!out) return fail(GEQ_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(GEQ_ERR_NO_ADAPTER, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(GEQ_ERR_NO_ADAPTER, "device ordinal " + std::to_string(device) + " out of range (" + std::to_string(n) + " devices)");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(GEQ_ERR_NO_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(GEQ_ERR_NO_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    geq_ctx* c = new geq_ctx();
    c->device = device;
    c->name = prop.gcnArchName;
    c->n_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(GEQ_ERR_NO_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return GEQ_OK;
}

void geq_ctx_destroy(geq_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

geq_status geq_ctx_info(geq_ctx* ctx, char* name, size_t name_len, int* n_cu) {
    if (!ctx) return fail(GEQ_ERR_INVALID_ARG, "ctx is NULL");
    if (name && name_len) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (n_cu) *n_cu = ctx->n_cu;
    return GEQ_OK;
}

geq_status geq_standard_reach(const geq_graph* g, uint32_t* out, uint32_t* n_out) {
    if (!g || !out || !n_out || !g->row_offsets || !g->attacker_pos) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    uint32_t k = 0;
    for (uint32_t v = 0; v < g->n_vertices; v++)
        if (!g->attacker_pos[v] && g->row_offsets[v + 1] == g->row_offsets[v]) out[k++] = v;
    *n_out = k;
    return GEQ_OK;
}

void geq_energies_free(geq_energies* e) {
    if (!e) return;
    std::free(e->offsets);
    std::free(e->data);
    e->offsets = nullptr;
    e->data = nullptr;
    e->n_nodes = 0;
}

geq_status geq_game_upload(geq_ctx* ctx, const geq_graph* g, geq_game** out) {
    if (!ctx || !g || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    ConfInfo ci;
    geq_status st = check_conf(g->conf, ci);
    if (st != GEQ_OK) return st;
    const uint32_t V = g->n_vertices, E = g->n_edges;
    if (V == 0) return fail(GEQ_ERR_INVALID_ARG, "empty game graph");
    if (!g->row_offsets || !g->attacker_pos || (E && (!g->column_indices || !g->weights)))
        return fail(GEQ_ERR_INVALID_ARG, "missing graph array");
    if (g->row_offsets[0] != 0 || g->row_offsets[V] != E) return fail(GEQ_ERR_INVALID_ARG, "row_offsets must start at 0 and end at n_edges");
    for (uint32_t v = 0; v < V; v++)
        if (g->row_offsets[v + 1] < g->row_offsets[v]) return fail(GEQ_ERR_INVALID_ARG, "row_offsets not monotone");
    std::vector<uint2> edges(E);
    for (uint32_t k = 0; k < E; k++) {
        uint32_t c = g->column_indices[k];
        if (c >= V) return fail(GEQ_ERR_INVALID_ARG, "column index out of range at edge " + std::to_string(k));
        uint32_t w;
        if (!update_to_canon(ci, g->weights[k], w)) return fail(GEQ_ERR_INVALID_ARG, "invalid update at edge " + std::to_string(k));
        edges[k] = make_uint2(c, w);
    }
    // reverse CSR: given, or derived as GameGraph::make_reverse does
    std::vector<uint32_t> rrow(V + 1, 0), rcol(E);
    if (g->rev_row_offsets && g->rev_column_indices) {
        if (g->rev_row_offsets[0] != 0 || g->rev_row_offsets[V] != E) return fail(GEQ_ERR_INVALID_ARG, "rev_row_offsets malformed");
        for (uint32_t v = 0; v <= V; v++) rrow[v] = g->rev_row_offsets[v];
        for (uint32_t v = 0; v < V; v++)
            if (rrow[v + 1] < rrow[v]) return fail(GEQ_ERR_INVALID_ARG, "rev_row_offsets not monotone");
        for (uint32_t k = 0; k < E; k++) {
            rcol[k] = g->rev_column_indices[k];
            if (rcol[k] >= V) return fail(GEQ_ERR_INVALID_ARG, "rev column index out of range");
        }
    } else {
        for (uint32_t k = 0; k < E; k++) rrow[g->column_indices[k] + 1]++;
        for (uint32_t v = 0; v < V; v++) rrow[v + 1] += rrow[v];
        std::vector<uint32_t> fill(rrow.begin(), rrow.end() - 1);
        for (uint32_t v = 0; v < V; v++)
            for (uint32_t k = g->row_offsets[v]; k < g->row_offsets[v + 1]; k++) rcol[fill[g->column_indices[k]]++] = v;
    }

    (void)hipSetDevice(ctx->device);
    geq_game* gm = new geq_game();
    gm->ctx = ctx;
    gm->conf = ci;
    gm->V = V;
    gm->E = E;
    gm->h_atk.assign(g->attacker_pos, g->attacker_pos + V);
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
    if ((e = gm->cur.alloc(V)) != hipSuccess) return oom(e, "energy lists");
    if ((e = gm->nxt.alloc(V)) != hipSuccess) return oom(e, "energy lists");
    if ((e = gm->wl0.alloc(V)) != hipSuccess) return oom(e, "worklist");
    if ((e = gm->wl1.alloc(V)) != hipSuccess) return oom(e, "worklist");
    if ((e = gm->chg.alloc(V)) != hipSuccess) return oom(e, "changed flags");
    if ((e = gm->stamp.alloc(V)) != hipSuccess) return oom(e, "stamps");
    if ((e = gm->ctl.alloc(1)) != hipSuccess) return oom(e, "control block");
    if ((e = gm->stats.alloc(4)) != hipSuccess) return oom(e, "stats");
    gm->arena_cap = std::max<uint32_t>(1u << 20, (uint32_t)std::min<uint64_t>(8ull * E + 1024, 0x7fffffffull));
    if ((e = gm->arena.alloc(gm->arena_cap)) != hipSuccess) return oom(e, "energy arena");
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&gm->h_ctl), sizeof(Ctl), hipHostMallocDefault)) != hipSuccess)
        return oom(e, "pinned control block");
    hipStream_t s = ctx->stream;
    auto up = [&](void* dst, const void* src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s); };
    if (up(gm->row.p, g->row_offsets, 4ull * (V + 1)) != hipSuccess || up(gm->rrow.p, rrow.data(), 4ull * (V + 1)) != hipSuccess ||
        (E && up(gm->rcol.p, rcol.data(), 4ull * E) != hipSuccess) || (E && up(gm->edges.p, edges.data(), 8ull * E) != hipSuccess) ||
        up(gm->atk.p, g->attacker_pos, V) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        delete gm;
        return fail(GEQ_ERR_TRANSFER, "uploading the game graph failed");
    }
    *out = gm;
    return GEQ_OK;
}

void geq_game_free(geq_game* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    delete g;
}

}  // extern "C"

namespace {

__global__ void k_seed_zero(uint4* cur, const uint32_t* ids, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cur[ids[i]] = make_uint4(1u, 0u, 0u, 0u);   // EnergyArray::zero(1): one all-zero energy
}

// Grow the arena keeping its content (lists of the committed state live in [0, top)).
geq_status grow_arena(geq_game* g, uint32_t needed_top) {
    uint64_t cap = std::max<uint64_t>(2ull * g->arena_cap, (uint64_t)needed_top + (1u << 20));
    cap = std::min<uint64_t>(cap, 0xffffffffull);
    if (cap <= g->arena_cap) return fail(GEQ_ERR_OUT_OF_MEMORY, "energy arena exceeds 2^32 entries");
    uint16_t* p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), 2ull * cap);
    if (e != hipSuccess) return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("growing the energy arena: ") + hipGetErrorString(e));
    e = hipMemcpyAsync(p, g->arena.p, 2ull * g->arena_cap, hipMemcpyDeviceToDevice, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return fail(GEQ_ERR_TRANSFER, std::string("copying the energy arena: ") + hipGetErrorString(e));
    }
    (void)hipFree(g->arena.p);
    g->arena.p = p;
    g->arena.n = cap;
    g->arena_cap = (uint32_t)cap;
    return GEQ_OK;
}

}  // namespace

extern "C" geq_status geq_game_solve(geq_game* g, const uint32_t* to_reach, uint32_t n_reach, const geq_energies* initial,
                                     int profile_kernels, geq_stats* stats) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "game is NULL");
    if (n_reach && !to_reach) return fail(GEQ_ERR_INVALID_ARG, "to_reach is NULL");
    const uint32_t V = g->V;
    for (uint32_t i = 0; i < n_reach; i++)
        if (to_reach[i] >= V) return fail(GEQ_ERR_INVALID_ARG, "to_reach id out of range");
    geq_ctx* ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;

    // ---- initial state: EnergyGame::with_reach (src/energygame.rs:355-365) or explicit lists
    std::vector<uint16_t> h_arena;
    if (initial) {
        if (initial->n_nodes != V || !initial->offsets || (initial->offsets[V] && !initial->data))
            return fail(GEQ_ERR_INVALID_ARG, "initial energies must cover every node");
        if (initial->energy_size != 1) return fail(GEQ_ERR_INVALID_ARG, "initial energies: energy_size must be 1 for this conf");
        std::vector<uint4> h_cur(V);
        for (uint32_t v = 0; v < V; v++) {
            uint32_t b = initial->offsets[v], e = initial->offsets[v + 1];
            if (e < b || e - b > LEN_MASK) return fail(GEQ_ERR_INVALID_ARG, "initial energies: bad offsets / list too long");
            uint32_t len = e - b;
            uint16_t h[8] = {0};
            h[0] = (uint16_t)(len | (len > 1 ? UNSORTED : 0u));
            std::vector<uint16_t> vals(len);
            for (uint32_t j = 0; j < len; j++) {
                uint32_t c;
                if (!energy_to_canon(g->conf, initial->data[b + j], c)) return fail(GEQ_ERR_INVALID_ARG, "initial energies: invalid energy");
                vals[j] = (uint16_t)c;
            }
            if (len <= INLINE_CAP) {
                for (uint32_t j = 0; j < len; j++) h[1 + j] = vals[j];
            } else {
                uint32_t off = (uint32_t)h_arena.size();
                h_arena.insert(h_arena.end(), vals.begin(), vals.end());
                h[1] = (uint16_t)(off & 0xffff);
                h[2] = (uint16_t)(off >> 16);
            }
            std::memcpy(&h_cur[v], h, 16);
        }
        if (h_arena.size() + (1u << 20) > g->arena_cap) {
            geq_status st = grow_arena(g, (uint32_t)h_arena.size());
            if (st != GEQ_OK) return st;
        }
        GEQ_HIP(hipMemcpyAsync(g->cur.p, h_cur.data(), 16ull * V, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
        if (!h_arena.empty())
            GEQ_HIP(hipMemcpyAsync(g->arena.p, h_arena.data(), 2ull * h_arena.size(), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    if (n_reach) {
        if (!g->seed.p || g->seed.n < n_reach) GEQ_HIP(g->seed.alloc(n_reach), GEQ_ERR_OUT_OF_MEMORY);
        GEQ_HIP(hipMemcpyAsync(g->seed.p, to_reach, 4ull * n_reach, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    if (!initial) {
        GEQ_HIP(hipMemsetAsync(g->cur.p, 0, 16ull * V, s), GEQ_ERR_TRANSFER);
        if (n_reach) k_seed_zero<<<(n_reach + 255) / 256, 256, 0, s>>>(g->cur.p, g->seed.p, n_reach);
    }
    GEQ_HIP(hipMemsetAsync(g->stamp.p, 0, 4ull * V, s), GEQ_ERR_TRANSFER);
    Ctl c0{};
    c0.arena_top = (uint32_t)h_arena.size();
    GEQ_HIP(hipMemcpyAsync(g->ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemsetAsync(g->stats.p, 0, 4 * sizeof(unsigned long long), s), GEQ_ERR_TRANSFER);

    hipEvent_t ev_start, ev_end;
    GEQ_HIP(hipEventCreate(&ev_start), GEQ_ERR_NO_DEVICE);
    GEQ_HIP(hipEventCreate(&ev_end), GEQ_ERR_NO_DEVICE);
    std::vector<hipEvent_t> prof;   // 3 events per round: before eval, after eval, after expand
    GEQ_HIP(hipEventRecord(ev_start, s), GEQ_ERR_NO_DEVICE);

    const uint32_t max_blocks = (uint32_t)std::max(1, ctx->n_cu) * 8u;
    auto expand_grid = [&](uint32_t n) { return std::max(1u, std::min(max_blocks, (n + 64u * WAVES - 1) / (64u * WAVES))); };
    auto eval_grid = [&](uint32_t n) { return std::max(1u, std::min(max_blocks, (n + WAVES - 1) / WAVES)); };
    uint32_t* wl[2] = {g->wl0.p, g->wl1.p};

    // seed: predecessors of the nodes to reach (initialize_visit_lists, :1461-1465)
    if (n_reach) {
        ExpandArgs x{g->seed.p, n_reach, nullptr, g->cur.p, nullptr, g->rrow.p, g->rcol.p, g->stamp.p, 1u, wl[0], g->ctl.p, 0u};
        k_expand<<<expand_grid(n_reach), BLOCK, 0, s>>>(x);
    }
    GEQ_HIP(hipMemcpyAsync(g->h_ctl, g->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);

    uint64_t rounds = 0, launches = 0;
    uint32_t r = 0;
    geq_status result = GEQ_OK;
    while (true) {
        const uint32_t n = g->h_ctl->wl_n[r & 1];
        if (n == 0) break;
        const uint32_t nx = (r + 1) & 1;
        GEQ_HIP(hipMemsetAsync(&g->ctl.p->wl_n[nx], 0, 4, s), GEQ_ERR_TRANSFER);
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (profile_kernels) {
            GEQ_HIP(hipEventCreate(&e0), GEQ_ERR_NO_DEVICE);
            GEQ_HIP(hipEventCreate(&e1), GEQ_ERR_NO_DEVICE);
            GEQ_HIP(hipEventCreate(&e2), GEQ_ERR_NO_DEVICE);
            GEQ_HIP(hipEventRecord(e0, s), GEQ_ERR_NO_DEVICE);
        }
        EvalArgs ea{wl[r & 1], n, g->row.p, g->edges.p, g->atk.p, g->cur.p, g->nxt.p, g->arena.p, g->ctl.p,
                    g->arena_cap, g->chg.p, g->conf.max, g->stats.p};
        k_eval<<<eval_grid(n), BLOCK, 0, s>>>(ea);
        launches++;
        if (profile_kernels) GEQ_HIP(hipEventRecord(e1, s), GEQ_ERR_NO_DEVICE);
        ExpandArgs x{wl[r & 1], n, g->chg.p, g->cur.p, g->nxt.p, g->rrow.p, g->rcol.p, g->stamp.p, r + 2u, wl[nx], g->ctl.p, nx};
        k_expand<<<expand_grid(n), BLOCK, 0, s>>>(x);
        if (profile_kernels) {
            GEQ_HIP(hipEventRecord(e2, s), GEQ_ERR_NO_DEVICE);
            prof.push_back(e0);
            prof.push_back(e1);
            prof.push_back(e2);
        }
        GEQ_HIP(hipMemcpyAsync(g->h_ctl, g->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        if (g->h_ctl->overflow) {
            // k_expand did nothing; grow the arena and evaluate the same round again
            result = grow_arena(g, g->h_ctl->arena_top);
            if (result != GEQ_OK) break;
            Ctl fix = *g->h_ctl;
            fix.overflow = 0;
            fix.wl_n[nx] = 0;
            GEQ_HIP(hipMemcpyAsync(g->ctl.p, &fix, sizeof(Ctl), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
            GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
            *g->h_ctl = fix;
            continue;
        }
        rounds++;
        r++;
    }
    GEQ_HIP(hipEventRecord(ev_end, s), GEQ_ERR_NO_DEVICE);
    GEQ_HIP(hipEventSynchronize(ev_end), GEQ_ERR_NO_DEVICE);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev_start, ev_end);
    double eval_ms = 0, expand_ms = 0;
    for (size_t i = 0; i + 2 < prof.size(); i += 3) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, prof[i], prof[i + 1]);
        (void)hipEventElapsedTime(&b, prof[i + 1], prof[i + 2]);
        eval_ms += a;
        expand_ms += b;
    }
    for (hipEvent_t ev : prof) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(ev_start);
    (void)hipEventDestroy(ev_end);
    if (result != GEQ_OK) return result;
    unsigned long long h_stats[4] = {0, 0, 0, 0};
    GEQ_HIP(hipMemcpy(h_stats, g->stats.p, sizeof(h_stats), hipMemcpyDeviceToHost), GEQ_ERR_TRANSFER);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rounds = rounds;
        stats->node_evals = h_stats[0];
        stats->edge_updates = h_stats[1];
        stats->alg_bytes = h_stats[2];
        stats->changed = h_stats[3];
        stats->solve_ms = ms;
        stats->eval_kernel_ms = eval_ms;
        stats->expand_kernel_ms = expand_ms;
        stats->eval_launches = launches;
    }
    return GEQ_OK;
}

extern "C" geq_status geq_game_energies(geq_game* g, geq_energies* out) {
    if (!g || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    (void)hipSetDevice(g->ctx->device);
    hipStream_t s = g->ctx->stream;
    const uint32_t V = g->V;
    if (!g->lens.p || g->lens.n < V) GEQ_HIP(g->lens.alloc(V), GEQ_ERR_OUT_OF_MEMORY);
    k_list_lens<<<(V + 255) / 256, 256, 0, s>>>(g->cur.p, V, g->lens.p);
    std::vector<uint32_t> lens(V), offs(V + 1);
    GEQ_HIP(hipMemcpyAsync(lens.data(), g->lens.p, 4ull * V, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    uint64_t tot = 0;
    for (uint32_t v = 0; v < V; v++) {
        offs[v] = (uint32_t)tot;
        tot += lens[v];
    }
    if (tot > 0xffffffffull) return fail(GEQ_ERR_OUT_OF_MEMORY, "result exceeds 2^32 energies");
    offs[V] = (uint32_t)tot;
    if (!g->offs.p || g->offs.n < V + 1) GEQ_HIP(g->offs.alloc(V + 1), GEQ_ERR_OUT_OF_MEMORY);
    if (!g->out.p || g->out.n < tot) GEQ_HIP(g->out.alloc(tot), GEQ_ERR_OUT_OF_MEMORY);
    GEQ_HIP(hipMemcpyAsync(g->offs.p, offs.data(), 4ull * (V + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    k_gather<<<(V + 255) / 256, 256, 0, s>>>(g->cur.p, g->arena.p, V, g->offs.p, g->out.p);
    uint32_t* data = (uint32_t*)std::malloc(4ull * (tot ? tot : 1));
    uint32_t* ooffs = (uint32_t*)std::malloc(4ull * (V + 1));
    if (!data || !ooffs) {
        std::free(data);
        std::free(ooffs);
        return fail(GEQ_ERR_OUT_OF_MEMORY, "host allocation of the result");
    }
    hipError_t e = hipMemcpyAsync(data, g->out.p, 4ull * tot, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        std::free(data);
        std::free(ooffs);
        return fail(GEQ_ERR_TRANSFER, std::string("downloading energies: ") + hipGetErrorString(e));
    }
    std::memcpy(ooffs, offs.data(), 4ull * (V + 1));
    const bool canon = (g->conf.ebits == 2);   // 2-bit fields pack exactly like the canonical layout
    if (!canon)
        for (uint64_t i = 0; i < tot; i++) data[i] = energy_from_canon(g->conf, data[i]);
    out->n_nodes = V;
    out->energy_size = 1;
    out->offsets = ooffs;
    out->data = data;
    return GEQ_OK;
}

extern "C" geq_status geq_solve(geq_ctx* ctx, const geq_graph* graph, const uint32_t* to_reach, uint32_t n_reach,
                                const geq_energies* initial, geq_energies* out, geq_stats* stats) {
    if (!out) return fail(GEQ_ERR_INVALID_ARG, "out is NULL");
    geq_game* g = nullptr;
    geq_status st = geq_game_upload(ctx, graph, &g);
    if (st != GEQ_OK) return st;
    st = geq_game_solve(g, to_reach, n_reach, initial, 0, stats);
    if (st == GEQ_OK) st = geq_game_energies(g, out);
    geq_game_free(g);
    return st;
}
