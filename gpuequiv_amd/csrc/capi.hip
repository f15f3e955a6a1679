// C ABI of the solver (include/gpuequiv.h): context, game upload, the
// device-resident fixpoint loop and result download.
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"
#include "scan.hpp"

namespace geq {

static thread_local std::string g_last_error;

// Every live context, so that an allocation that runs out of device memory
// can free the game buffers the contexts keep for reuse.
static std::mutex g_ctx_mu;
static std::vector<geq_ctx*> g_ctxs;

bool drop_spare_games(int device) {
    std::vector<geq_game*> drop;
    bool bufs = false;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        for (geq_ctx* c : g_ctxs) {
            if (c->device != device) continue;
            {
                std::lock_guard<std::mutex> gl(c->game_mu);
                if (c->spare_game) drop.push_back(c->spare_game);
                c->spare_game = nullptr;
            }
            // the spares of the context's build scratch, unless a build holds it
            // (that build's own allocation drops them through its RetireScope)
            std::unique_lock<std::mutex> bl(c->build_mu, std::try_to_lock);
            if (bl.owns_lock() && c->build_ws) bufs = drop_build_spares(c->build_ws.get()) || bufs;
        }
    }
    for (geq_game* g : drop) delete g;
    return !drop.empty() || bufs;
}

hipError_t alloc_host_ctl(geq_game* g) {
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, 128, hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    void* d = nullptr;
    if ((e = hipHostGetDevicePointer(&d, p, 0)) != hipSuccess || !d) {
        (void)hipHostFree(p);
        (void)hipGetLastError();
        return e != hipSuccess ? e : hipErrorInvalidValue;
    }
    std::memset(p, 0, 128);
    g->h_ctl = static_cast<Ctl*>(p);
    *g->h_ctl = Ctl{};
    g->d_ctl_host = static_cast<Ctl*>(d);
    g->sig_seq = 0;
    return hipSuccess;
}

geq_status wait_signal(geq_game* g, uint32_t seq) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(g->h_ctl) + 64);
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(w, __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            // a long batch, or a failed kernel: the stream's wait (and its error)
            GEQ_HIP(hipStreamSynchronize(g->ctx->stream), GEQ_ERR_TRANSFER);
            if (__atomic_load_n(w, __ATOMIC_ACQUIRE) != seq) return fail(GEQ_ERR_INTERNAL, "the device's signal did not arrive");
            return GEQ_OK;
        }
        __builtin_ia32_pause();
    }
    return GEQ_OK;
}

hipError_t device_malloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        int dev = 0;
        const bool games = hipGetDevice(&dev) == hipSuccess && drop_spare_games(dev);
        if (spare_drop() || games) {
            (void)hipGetLastError();   // the failed attempt's error is handled here
            e = hipMalloc(p, bytes);
        }
    }
    // the caller gets the error from the return value: it must not stay behind
    // for the next hipGetLastError (a launch check) to report
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

void* SpareBufs::take(size_t bytes, size_t* got) {
    size_t best = bufs.size();
    for (size_t i = 0; i < bufs.size(); i++)
        if (bufs[i].bytes >= bytes && bufs[i].bytes / 4 <= bytes && (best == bufs.size() || bufs[i].bytes < bufs[best].bytes))
            best = i;
    if (best == bufs.size()) return nullptr;
    void* p = bufs[best].p;
    *got = bufs[best].bytes;
    bufs.erase(bufs.begin() + best);
    return p;
}
bool SpareBufs::drop() {
    for (const Buf& b : bufs) (void)hipFree(b.p);
    const bool any = !bufs.empty();
    bufs.clear();
    return any;
}

namespace {
struct RetireList {
    int depth = 0;
    SpareBufs* pool = nullptr;   // the outermost open scope's
    std::vector<SpareBufs::Buf> bufs;
};
thread_local RetireList t_retired;
}  // namespace

RetireScope::RetireScope(hipStream_t s, SpareBufs* pool) : s_(s), outer_(t_retired.depth == 0) {
    if (outer_) t_retired.pool = pool;
    t_retired.depth++;
}
RetireScope::~RetireScope() {
    t_retired.depth--;
    if (!outer_) return;
    SpareBufs* pool = t_retired.pool;
    t_retired.pool = nullptr;
    if (t_retired.bufs.empty()) return;
    if (hipStreamSynchronize(s_) != hipSuccess || !pool) {   // their kernels' state is unknown: free them
        for (const SpareBufs::Buf& b : t_retired.bufs) (void)hipFree(b.p);
        t_retired.bufs.clear();
        return;
    }
    pool->bufs.insert(pool->bufs.end(), t_retired.bufs.begin(), t_retired.bufs.end());
    t_retired.bufs.clear();
    pool->trim();
}
// keep the largest buffers that fit SPARE_BUFS and SPARE_BYTES, and up to
// SPARE_SMALL_BUFS more of at most SPARE_SMALL bytes; free the rest
void SpareBufs::trim() {
    auto& v = bufs;
    std::sort(v.begin(), v.end(), [](const SpareBufs::Buf& a, const SpareBufs::Buf& b) { return a.bytes > b.bytes; });
    size_t bytes = 0, kept = 0, large = 0, small = 0;
    for (const SpareBufs::Buf& b : v) {
        if (large < SPARE_BUFS && bytes + b.bytes <= SPARE_BYTES) {
            v[kept++] = b;
            bytes += b.bytes;
            large++;
        } else if (b.bytes <= SPARE_SMALL && small < SPARE_SMALL_BUFS) {
            v[kept++] = b;
            small++;
        } else {
            (void)hipFree(b.p);
        }
    }
    v.resize(kept);
}
bool RetireScope::active() { return t_retired.depth > 0; }
bool RetireScope::take(void* p, size_t bytes) {
    if (t_retired.depth == 0) return false;
    t_retired.bufs.push_back(SpareBufs::Buf{p, bytes});
    return true;
}
void* geq::spare_take(size_t bytes, size_t* got) {
    return t_retired.depth && t_retired.pool ? t_retired.pool->take(bytes, got) : nullptr;
}
bool geq::spare_drop() { return t_retired.depth && t_retired.pool ? t_retired.pool->drop() : false; }

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
    if (c.elements == 0)
        return fail(GEQ_ERR_UNSUPPORTED, "energy configuration without elements");
    o = ConfInfo{};
    o.elements = c.elements;
    o.max = c.max;
    o.ebits = next_pow2(bitlen(c.max));
    o.emask = 0xffffffffu >> (32 - o.ebits);
    o.ubits = next_pow2(bitlen(c.elements + 2));
    o.umask = 0xffffffffu >> (32 - o.ubits);
    const uint64_t esz = ((uint64_t)o.ebits * c.elements + 31) / 32, usz = ((uint64_t)o.ubits * c.elements + 31) / 32;
    if (esz > GEN_MAX_WORDS)
        return fail(GEQ_ERR_UNSUPPORTED, "energy configuration needs " + std::to_string(esz) + " words per energy (at most " +
                                             std::to_string(GEN_MAX_WORDS) + "): elements=" + std::to_string(c.elements) +
                                             " max=" + std::to_string(c.max));
    o.esz = (uint32_t)esz;
    o.usz = (uint32_t)usz;
    o.generic = !(c.elements <= 6 && c.max <= 3);
    return GEQ_OK;
}

GConf gconf_of(const ConfInfo& c) {
    GConf g{};
    g.elements = c.elements;
    g.maxv = c.max;
    g.ebits = c.ebits;
    g.emask = c.emask;
    g.ubits = c.ubits;
    g.umask = c.umask;
    g.esz = c.esz;
    g.usz = c.usz;
    if (c.ebits < 32) {
        uint32_t E = 0, low = 0;
        for (uint32_t f = 0; f < 32; f += 2 * c.ebits) {   // the even fields of a word
            E |= c.emask << f;
            low |= 1u << f;
        }
        g.E = E;
        g.G = low << c.ebits;                              // lowest bit of every odd field
    }
    return g;
}

// Field k of a packed value of `bits`-wide fields (fields never straddle words).
static uint32_t field(const uint32_t* w, uint32_t k, uint32_t bits, uint32_t mask) {
    const uint64_t b = (uint64_t)k * bits;
    return (w[b >> 5] >> (b & 31)) & mask;
}
// The bits of the last word above the last field are zero.
static bool tail_clear(const uint32_t* w, uint32_t n_fields, uint32_t bits, uint32_t words) {
    const uint64_t used = (uint64_t)n_fields * bits - 32ull * (words - 1);   // bits used in the last word
    return used >= 32 || (w[words - 1] >> used) == 0;
}

bool energy_valid(const ConfInfo& c, const uint32_t* w) {
    if (!tail_clear(w, c.elements, c.ebits, c.esz)) return false;
    for (uint32_t k = 0; k < c.elements; k++)
        if (field(w, k, c.ebits, c.emask) > c.max) return false;
    return true;
}

bool update_valid(const ConfInfo& c, const uint32_t* w) {
    if (!tail_clear(w, c.elements, c.ubits, c.usz)) return false;
    for (uint32_t k = 0; k < c.elements; k++)
        if (field(w, k, c.ubits, c.umask) > c.elements + 1) return false;   // Min(k) needs k <= elements (src/energy.rs:551-556)
    return true;
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
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_stage) (void)hipHostFree(h_stage);   // (d_stage is its device view)
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
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
    // the static scratch of geq_solve_noalloc on this device (a failure here
    // only disables that entry point, which then reports it)
    (void)noalloc_setup(device);
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_ctxs.push_back(c);
    }
    *out = c;
    return GEQ_OK;
}

void geq_ctx_destroy(geq_ctx* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_ctxs.erase(std::remove(g_ctxs.begin(), g_ctxs.end(), ctx), g_ctxs.end());
    }
    (void)hipSetDevice(ctx->device);
    ctx->build_ws.reset();   // its device buffers, before the stream goes
    if (ctx->spare_game) geq_game_free(ctx->spare_game);
    ctx->spare_game = nullptr;
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

geq_status geq_ctx_memory(geq_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx) return fail(GEQ_ERR_INVALID_ARG, "ctx is NULL");
    (void)hipSetDevice(ctx->device);
    size_t f = 0, t = 0;
    GEQ_HIP(hipMemGetInfo(&f, &t), GEQ_ERR_NO_DEVICE);
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
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

}  // extern "C"

namespace {

// Validate one graph and append it to the merged host arrays at node offset vb / edge offset eb.
// gwt (generic confs only): the update words of every edge, usz per edge.
geq_status append_graph(const geq_graph* g, const ConfInfo& ci, uint32_t vb, uint64_t eb, std::vector<uint32_t>& row,
                        std::vector<uint2>& edges, std::vector<uint32_t>& rrow, std::vector<uint32_t>& rcol,
                        std::vector<uint8_t>& atk, std::vector<uint32_t>* gwt) {
    const uint32_t V = g->n_vertices, E = g->n_edges;
    if (V == 0) return fail(GEQ_ERR_INVALID_ARG, "empty game graph");
    if (!g->row_offsets || !g->attacker_pos || (E && (!g->column_indices || !g->weights)))
        return fail(GEQ_ERR_INVALID_ARG, "missing graph array");
    if (g->row_offsets[0] != 0 || g->row_offsets[V] != E) return fail(GEQ_ERR_INVALID_ARG, "row_offsets must start at 0 and end at n_edges");
    for (uint32_t v = 0; v < V; v++)
        if (g->row_offsets[v + 1] < g->row_offsets[v]) return fail(GEQ_ERR_INVALID_ARG, "row_offsets not monotone");
    for (uint32_t v = 0; v < V; v++) row[vb + v + 1] = (uint32_t)(eb + g->row_offsets[v + 1]);
    for (uint32_t k = 0; k < E; k++) {
        uint32_t c = g->column_indices[k];
        if (c >= V) return fail(GEQ_ERR_INVALID_ARG, "column index out of range at edge " + std::to_string(k));
        uint32_t w;
        if (ci.generic) {
            const uint32_t* uw = g->weights + (size_t)k * ci.usz;
            if (!update_valid(ci, uw)) return fail(GEQ_ERR_INVALID_ARG, "invalid update at edge " + std::to_string(k));
            std::copy(uw, uw + ci.usz, gwt->begin() + (eb + k) * ci.usz);
            w = uw[0];
        } else if (!update_to_canon(ci, g->weights[k], w)) {
            return fail(GEQ_ERR_INVALID_ARG, "invalid update at edge " + std::to_string(k));
        }
        edges[eb + k] = make_uint2(vb + c, w);
    }
    for (uint32_t v = 0; v < V; v++) atk[vb + v] = g->attacker_pos[v] ? 1 : 0;
    // reverse CSR: given, or derived as GameGraph::make_reverse does (:235-250)
    if (g->rev_row_offsets && g->rev_column_indices) {
        if (g->rev_row_offsets[0] != 0 || g->rev_row_offsets[V] != E) return fail(GEQ_ERR_INVALID_ARG, "rev_row_offsets malformed");
        for (uint32_t v = 0; v < V; v++)
            if (g->rev_row_offsets[v + 1] < g->rev_row_offsets[v]) return fail(GEQ_ERR_INVALID_ARG, "rev_row_offsets not monotone");
        for (uint32_t v = 0; v < V; v++) rrow[vb + v + 1] = (uint32_t)(eb + g->rev_row_offsets[v + 1]);
        for (uint32_t k = 0; k < E; k++) {
            uint32_t r = g->rev_column_indices[k];
            if (r >= V) return fail(GEQ_ERR_INVALID_ARG, "rev column index out of range");
            rcol[eb + k] = vb + r;
        }
    } else {
        std::vector<uint32_t> cnt(V + 1, 0);
        for (uint32_t k = 0; k < E; k++) cnt[g->column_indices[k] + 1]++;
        for (uint32_t v = 0; v < V; v++) cnt[v + 1] += cnt[v];
        for (uint32_t v = 0; v < V; v++) rrow[vb + v + 1] = (uint32_t)(eb + cnt[v + 1]);
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (uint32_t v = 0; v < V; v++)
            for (uint32_t k = g->row_offsets[v]; k < g->row_offsets[v + 1]; k++) rcol[eb + fill[g->column_indices[k]]++] = vb + v;
    }
    return GEQ_OK;
}

// Renumber the positions of each game (ids base[i] .. base[i+1]) in reverse
// breadth-first order from its defender dead ends (the standard reach): the
// positions a round of the fixpoint evaluates then sit in a few dense id
// ranges, so node records, edge runs, reverse runs and queued-bitmap words
// are read in whole cache lines instead of one random line each.  Positions
// the walk does not reach keep their order after the reached ones.  Rewrites
// every array in place; inv[old id] = new id.  Results are the same under any
// numbering (each game's fixpoint is unique).
void renumber(const std::vector<uint32_t>& base, uint32_t usz, std::vector<uint32_t>& row, std::vector<uint2>& edges,
              std::vector<uint32_t>& rrow, std::vector<uint32_t>& rcol, std::vector<uint8_t>& atk, std::vector<uint32_t>& gwt,
              std::vector<uint32_t>& inv) {
    const uint32_t V = (uint32_t)atk.size(), E = (uint32_t)edges.size();
    std::vector<uint32_t> perm(V);     // new id -> old id
    inv.assign(V, 0xffffffffu);
    for (size_t gi = 0; gi + 1 < base.size(); gi++) {
        const uint32_t b0 = base[gi], b1 = base[gi + 1];
        uint32_t head = b0, tail = b0;
        auto visit = [&](uint32_t v) {
            inv[v] = tail;
            perm[tail++] = v;
        };
        for (uint32_t v = b0; v < b1; v++)
            if (!atk[v] && row[v] == row[v + 1]) visit(v);
        while (head < tail) {
            const uint32_t v = perm[head++];
            for (uint32_t k = rrow[v]; k < rrow[v + 1]; k++)
                if (inv[rcol[k]] == 0xffffffffu) visit(rcol[k]);
        }
        for (uint32_t v = b0; v < b1; v++)
            if (inv[v] == 0xffffffffu) visit(v);
    }
    std::vector<uint32_t> row2(V + 1, 0), rrow2(V + 1, 0), rcol2(E);
    std::vector<uint2> edges2(E);
    std::vector<uint8_t> atk2(V);
    std::vector<uint32_t> gwt2(usz ? (size_t)E * usz : 0);
    for (uint32_t i = 0; i < V; i++) {
        const uint32_t o = perm[i];
        const uint32_t e0 = row[o], d = row[o + 1] - e0, r0 = rrow[o], rd = rrow[o + 1] - r0;
        row2[i + 1] = row2[i] + d;
        rrow2[i + 1] = rrow2[i] + rd;
        for (uint32_t k = 0; k < d; k++) edges2[row2[i] + k] = make_uint2(inv[edges[e0 + k].x], edges[e0 + k].y);
        if (usz) std::copy(gwt.begin() + (size_t)e0 * usz, gwt.begin() + (size_t)(e0 + d) * usz, gwt2.begin() + (size_t)row2[i] * usz);
        for (uint32_t k = 0; k < rd; k++) rcol2[rrow2[i] + k] = inv[rcol[r0 + k]];
        atk2[i] = atk[o];
    }
    row.swap(row2);
    rrow.swap(rrow2);
    rcol.swap(rcol2);
    edges.swap(edges2);
    atk.swap(atk2);
    gwt.swap(gwt2);
}

geq_status upload_games(geq_ctx* ctx, uint32_t n_games, const geq_graph* graphs, geq_game** out) {
    if (!ctx || !graphs || !out || n_games == 0) return fail(GEQ_ERR_INVALID_ARG, "NULL argument or no games");
    *out = nullptr;
    ConfInfo ci;
    geq_status st = check_conf(graphs[0].conf, ci);
    if (st != GEQ_OK) return st;
    uint64_t Vt = 0, Et = 0;
    std::vector<uint32_t> base(n_games + 1, 0);
    for (uint32_t i = 0; i < n_games; i++) {
        if (graphs[i].conf.elements != graphs[0].conf.elements || graphs[i].conf.max != graphs[0].conf.max)
            return fail(GEQ_ERR_INVALID_ARG, "batched games must share one energy configuration");
        base[i] = (uint32_t)Vt;
        Vt += graphs[i].n_vertices;
        Et += graphs[i].n_edges;
        if (Vt >= 0xffffffffull || Et >= 0xffffffffull) return fail(GEQ_ERR_INVALID_ARG, "more than 2^32 positions or edges");
    }
    base[n_games] = (uint32_t)Vt;
    const uint32_t V = (uint32_t)Vt, E = (uint32_t)Et;
    std::vector<uint32_t> row(V + 1, 0), rrow(V + 1, 0), rcol(E);
    std::vector<uint2> edges(E);
    std::vector<uint8_t> atkv(V);
    std::vector<uint32_t> gwt(ci.generic ? (size_t)E * ci.usz : 0);
    uint64_t eb = 0;
    for (uint32_t i = 0; i < n_games; i++) {
        st = append_graph(&graphs[i], ci, base[i], eb, row, edges, rrow, rcol, atkv, &gwt);
        if (st != GEQ_OK) {
            if (n_games > 1) set_error("game " + std::to_string(i) + ": " + geq_last_error());
            return st;
        }
        eb += graphs[i].n_edges;
    }
    // GEQ_RENUMBER=1: internal ids in reverse breadth-first order (off by
    // default: measured within noise on batches and 3% slower on single
    // games, whose consecutive predecessor ids then share queued-bitmap words)
    std::vector<uint32_t> inv;
    const char* renum_env = std::getenv("GEQ_RENUMBER");
    if (renum_env && std::atoi(renum_env) != 0 && V > 1)
        renumber(base, ci.generic ? ci.usz : 0, row, edges, rrow, rcol, atkv, gwt, inv);

    (void)hipSetDevice(ctx->device);
    geq_game* gm = new geq_game();
    gm->ctx = ctx;
    gm->conf = ci;
    gm->V = V;
    gm->E = E;
    gm->game_base = base;
    gm->h_atk = std::move(atkv);
    gm->inv = std::move(inv);
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
    if (const char* env = std::getenv("GEQ_SOLVE_MODE")) gm->tune.mode = (uint32_t)std::atoi(env);
    gm->tune.ordered = worklist_order_default();
    if (ci.generic) {
        // the word arena: the zero energy, then room for about one list per two edges
        gm->garena_cap = (uint32_t)std::min<uint64_t>(
            GEN_ZERO_WORDS + std::max<uint64_t>(64, (uint64_t)initial_arena_entries(E, 0) / 2) * ci.esz, 0x7fffffffull);
        if ((e = gm->garena.alloc(gm->garena_cap)) != hipSuccess) return oom(e, "energy arena");
        if ((e = gm->gwt.alloc(std::max<size_t>((size_t)E * ci.usz, 1))) != hipSuccess) return oom(e, "updates");
        gm->arena_cap = 0;
    } else {
        gm->arena_cap = initial_arena_entries(E, 0);
        if ((e = gm->arena.alloc(gm->arena_cap)) != hipSuccess) return oom(e, "energy arena");
    }
    if ((e = alloc_host_ctl(gm)) != hipSuccess) return oom(e, "pinned control block");
    hipStream_t s = ctx->stream;
    if (ci.generic) {
        if ((E && hipMemcpyAsync(gm->gwt.p, gwt.data(), 4ull * E * ci.usz, hipMemcpyHostToDevice, s) != hipSuccess) ||
            hipMemsetAsync(gm->garena.p, 0, 4ull * GEN_ZERO_WORDS, s) != hipSuccess) {
            delete gm;
            return fail(GEQ_ERR_TRANSFER, "uploading the game graph failed");
        }
    }
    auto up = [&](void* dst, const void* src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s); };
    if (up(gm->row.p, row.data(), 4ull * (V + 1)) != hipSuccess || up(gm->rrow.p, rrow.data(), 4ull * (V + 1)) != hipSuccess ||
        (E && up(gm->rcol.p, rcol.data(), 4ull * E) != hipSuccess) || (E && up(gm->edges.p, edges.data(), 8ull * E) != hipSuccess) ||
        up(gm->atk.p, gm->h_atk.data(), V) != hipSuccess) {
        delete gm;
        return fail(GEQ_ERR_TRANSFER, "uploading the game graph failed");
    }
    if (!gm->inv.empty()) {
        if ((e = gm->inv_map.alloc(V)) != hipSuccess) return oom(e, "id map");
        if (up(gm->inv_map.p, gm->inv.data(), 4ull * V) != hipSuccess) {
            delete gm;
            return fail(GEQ_ERR_TRANSFER, "uploading the game graph failed");
        }
    }
    if (V) k_node_hdr<<<(V + 255) / 256, 256, 0, s>>>(gm->row.p, gm->rrow.p, gm->atk.p, V, gm->st.p);
    if (hipStreamSynchronize(s) != hipSuccess) {
        delete gm;
        return fail(GEQ_ERR_TRANSFER, "uploading the game graph failed");
    }
    *out = gm;
    return GEQ_OK;
}

}  // namespace

extern "C" {

geq_status geq_game_upload(geq_ctx* ctx, const geq_graph* g, geq_game** out) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "graph is NULL");
    return upload_games(ctx, 1, g, out);
}

geq_status geq_batch_upload(geq_ctx* ctx, uint32_t n_games, const geq_graph* graphs, geq_game** out) {
    return upload_games(ctx, n_games, graphs, out);
}

void geq_game_free(geq_game* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    delete g;
}

}  // extern "C"

namespace {

// The per-solve reset in one launch: empty list slots (unless initial lists
// were uploaded), no node queued, zero statistics, a fresh control block.
__global__ void k_solve_init(uint4* st, uint32_t V, int clear_slots, uint32_t* queued, uint32_t qwords2,
                             unsigned long long* stats, uint32_t nstats, Ctl* ctl, uint32_t arena_top) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    if (clear_slots)
        for (uint32_t v = tid; v < V; v += nt) {   // the record after its header: both slots and the spare piece
            st[slot_index(v, 0)] = make_uint4(0, 0, 0, 0);
            st[slot_index(v, 1)] = make_uint4(0, 0, 0, 0);
            st[slot_index(v, 2)] = make_uint4(0, 0, 0, 0);
        }
    for (uint32_t i = tid; i < qwords2; i += nt) queued[i] = 0u;
    for (uint32_t i = tid; i < nstats; i += nt) stats[i] = 0ull;
    if (tid == 0) {
        Ctl c{};
        c.arena_top = arena_top;
        *ctl = c;
    }
}

// The download's offsets (n + 1) and lists (offs[n] words) into the staging.
__global__ void k_stage_out(const uint32_t* offs, uint32_t n, const uint32_t* data, uint64_t cap, uint32_t* h_offs,
                            uint32_t* h_data) {
    const uint64_t tot = min((uint64_t)offs[n], cap);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= n; i += stride) h_offs[i] = offs[i];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += stride) h_data[i] = data[i];
}

// The signal word after the control block: written last, at system scope.
__device__ __forceinline__ uint32_t* signal_word(Ctl* host_ctl) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(host_ctl) + 64);
}

// The control block into the host's pinned copy, then `seq` into the signal
// word (one wave: its system-scope fence orders its own stores).
__global__ void k_ctl_out(const Ctl* ctl, Ctl* host_ctl, uint32_t seq) {
    static_assert(sizeof(Ctl) % 4 == 0 && sizeof(Ctl) / 4 <= 64, "one word per thread");
    const uint32_t t = threadIdx.x;
    if (t < sizeof(Ctl) / 4) reinterpret_cast<uint32_t*>(host_ctl)[t] = reinterpret_cast<const uint32_t*>(ctl)[t];
    __threadfence_system();
    __syncthreads();
    if (t == 0) __hip_atomic_store(signal_word(host_ctl), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}


// slot 0 (tag 0, current while both slots are untouched) of every node to reach
__global__ void k_seed_zero(uint4* st, const uint32_t* ids, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st[slot_index(ids[i], 0)] = make_uint4(1u, 0u, 0u, 0u);   // EnergyArray::zero(1): one all-zero energy
}

}  // namespace

// Grow the arena keeping its content (lists of the committed state live in [0, top)).
uint32_t geq::initial_arena_entries(uint64_t E, uint64_t want) {
    // most antichains fit the 6 inline entries of a record; the arena grows on demand
    uint64_t n = want ? want : std::max<uint64_t>(1u << 20, E / 2);
    if (const char* env = std::getenv("GEQ_ARENA_ENTRIES")) {
        const long long v = std::atoll(env);
        if (v >= 64) n = (uint64_t)v;
    }
    return (uint32_t)std::min<uint64_t>(n, 0x7fffffffull);
}

geq_status geq::grow_arena(geq_game* g, uint32_t needed_top) {
    if (g->conf.generic) {   // the word arena of the list kernels
        uint64_t cap = std::max<uint64_t>(2ull * g->garena_cap, (uint64_t)needed_top + (1u << 18));
        cap = std::min<uint64_t>(cap, 0xffffffffull);
        if (cap <= g->garena_cap) return fail(GEQ_ERR_OUT_OF_MEMORY, "energy arena exceeds 2^32 words");
        uint32_t* p = nullptr;
        hipError_t e = device_malloc(reinterpret_cast<void**>(&p), 4ull * cap);
        if (e != hipSuccess) return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("growing the energy arena: ") + hipGetErrorString(e));
        e = hipMemcpyAsync(p, g->garena.p, 4ull * g->garena_cap, hipMemcpyDeviceToDevice, g->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return fail(GEQ_ERR_TRANSFER, std::string("copying the energy arena: ") + hipGetErrorString(e));
        }
        (void)hipFree(g->garena.p);
        g->garena.p = p;
        g->garena.n = cap;
        g->garena_cap = (uint32_t)cap;
        return GEQ_OK;
    }
    uint64_t cap = std::max<uint64_t>(2ull * g->arena_cap, (uint64_t)needed_top + (1u << 20));
    cap = std::min<uint64_t>(cap, 0xffffffffull);
    if (cap <= g->arena_cap) return fail(GEQ_ERR_OUT_OF_MEMORY, "energy arena exceeds 2^32 entries");
    uint16_t* p = nullptr;
    hipError_t e = device_malloc(reinterpret_cast<void**>(&p), 2ull * cap);
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


namespace geq {

// Launch shape of one round.  Kernels use grid-stride loops over the device
// worklist, so the shape only has to fill the chip: a few blocks per CU.
RoundShape round_shape(const geq_game* g) {
    const uint32_t ncu = (uint32_t)std::max(1, g->ctx->n_cu);
    RoundShape r;
    if (g->conf.generic) {
        // list kernels: GEN_BLOCK threads, one node per wave at a time, and a separate k_expand
        r.threads = GEN_BLOCK;
        r.fused = false;
        const uint32_t per_cu = std::min(g->tune.eval_blocks_per_cu, list_blocks_resident(g->conf.esz));
        r.eval_blocks = std::max(1u, std::min(ncu * per_cu, (g->V + GEN_BLOCK / 64 - 1) / (GEN_BLOCK / 64)));
        r.eval_blocks = std::min(r.eval_blocks, STAT_SLOTS);
        const uint32_t xpc = g->tune.expand_blocks_per_cu ? g->tune.expand_blocks_per_cu : 4u;
        r.expand_blocks = std::max(1u, std::min(ncu * xpc, (g->V + GEN_BLOCK - 1) / GEN_BLOCK));
        r.rounds_per_launch = std::max(1u, g->tune.rounds_per_launch);
        r.ordered = false;
        r.collect_blocks = 0;
        return r;
    }
    const uint32_t w = std::max(1u, std::min<uint32_t>(g->tune.waves_per_block, MAX_WAVES));
    r.threads = 64u * w;
    r.fused = g->tune.expand_blocks_per_cu == 0;
    r.eval_blocks = std::max(1u, std::min(ncu * g->tune.eval_blocks_per_cu, (g->V + w - 1) / w));
    // (the seeding launch uses this grid in the fused shape too)
    const uint32_t xpc = r.fused ? 4u : g->tune.expand_blocks_per_cu;
    r.expand_blocks = std::max(1u, std::min(ncu * xpc, (g->V + 64u * w - 1) / (64u * w)));
    r.rounds_per_launch = std::max(1u, g->tune.rounds_per_launch);
    r.ordered = r.fused && (g->tune.ordered == 1 || (g->tune.ordered == 2 && g->V >= ORDERED_MIN_V));
    r.collect_blocks = (uint32_t)std::max<uint64_t>(
        1, ((uint64_t)queue_words(g->V) + COLLECT_THREADS * COLLECT_WORDS - 1) / (COLLECT_THREADS * COLLECT_WORDS));
    // k_eval's grid-stride loop covers any worklist: a grid larger than what
    // is resident at once only adds a second, mostly idle wave of blocks
    r.eval_blocks = std::min(r.eval_blocks, ncu * eval_blocks_resident(r.fused, w));
    r.eval_blocks = std::min(r.eval_blocks, STAT_SLOTS);
    return r;
}

// Blocks of k_eval (w waves each, fused or not) that one CU holds at once.
// (Cached per shape; host threads may race to fill an entry, with the same value.)
uint32_t eval_blocks_resident(bool fused, uint32_t w) {
    static std::atomic<int> cache[2][MAX_WAVES + 1];
    std::atomic<int>& c = cache[fused ? 1 : 0][w];
    int v = c.load(std::memory_order_relaxed);
    if (v == 0) {
        int per_cu = 0;
        const void* fn = fused ? reinterpret_cast<const void*>(&k_eval<true>) : reinterpret_cast<const void*>(&k_eval<false>);
        const size_t lds = (size_t)w * (fused ? EVAL_FUSED_LDS_PER_WAVE : EVAL_LDS_PER_WAVE);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)(64 * w), lds) != hipSuccess || per_cu <= 0) per_cu = 64;
        v = per_cu;
        c.store(v, std::memory_order_relaxed);
    }
    return (uint32_t)v;
}

// The k_eval_list instantiation for energies of esz words.
const void* list_kernel(uint32_t esz) {
    return esz == 1 ? reinterpret_cast<const void*>(&k_eval_list<1>)
                    : esz <= 4 ? reinterpret_cast<const void*>(&k_eval_list<4>) : reinterpret_cast<const void*>(&k_eval_list<32>);
}

uint32_t list_blocks_resident(uint32_t esz) {
    static std::atomic<int> cache[3];
    std::atomic<int>& c = cache[esz == 1 ? 0 : esz <= 4 ? 1 : 2];
    int v = c.load(std::memory_order_relaxed);
    if (v == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, list_kernel(esz), (int)GEN_BLOCK, 0) != hipSuccess || per_cu <= 0)
            per_cu = 8;
        v = per_cu;
        c.store(v, std::memory_order_relaxed);
    }
    return (uint32_t)v;
}

void launch_list_eval(const RoundShape& shape, const GenArgs& ga, uint32_t k, hipStream_t s) {
    const uint32_t esz = ga.c.esz;
    if (esz == 1) k_eval_list<1><<<shape.eval_blocks, GEN_BLOCK, 0, s>>>(ga, k);
    else if (esz <= 4) k_eval_list<4><<<shape.eval_blocks, GEN_BLOCK, 0, s>>>(ga, k);
    else k_eval_list<32><<<shape.eval_blocks, GEN_BLOCK, 0, s>>>(ga, k);
}

GenArgs gen_args(geq_game* g) {
    GenArgs a{};
    a.wl[0] = g->wl0.p;
    a.wl[1] = g->wl1.p;
    a.edges = g->edges.p;
    a.wt = g->gwt.p;
    a.st = g->st.p;
    a.arena = g->garena.p;
    a.ctl = g->ctl.p;
    a.arena_cap = g->garena_cap;
    a.chg = g->chg.p;
    a.stats = g->stats.p;
    a.queued = g->queued.p;
    a.qwords = queue_words(g->V);
    a.scr = g->gscr.p;
    a.scr_words = g->scr_words;
    a.c = gconf_of(g->conf);
    return a;
}

// The list kernels' scratch: scr_words words for every wave of the grid.
geq_status ensure_scratch(geq_game* g, const RoundShape& shape, uint64_t words) {
    words = std::max<uint64_t>(std::max<uint64_t>(words, 4096), g->scr_words);
    if (words > 0x7fffffffull) return fail(GEQ_ERR_OUT_OF_MEMORY, "a node's candidate lists need more than 2^31 words");
    const uint64_t total = words * shape.eval_blocks * (GEN_BLOCK / 64);
    if (g->gscr.p && g->gscr.n >= total && g->scr_words >= words) return GEQ_OK;
    g->gscr.release();
    hipError_t e = g->gscr.alloc(total);
    if (e != hipSuccess)
        return fail(GEQ_ERR_OUT_OF_MEMORY, "allocating " + std::to_string(4 * total) + " bytes of list scratch: " + hipGetErrorString(e));
    g->scr_words = (uint32_t)words;
    return GEQ_OK;
}

EvalArgs eval_args(geq_game* g) {
    return EvalArgs{{g->wl0.p, g->wl1.p}, g->row.p, g->edges.p, g->atk.p, g->st.p, g->arena.p, g->ctl.p, g->arena_cap,
                    g->chg.p, g->conf.max, g->stats.p, g->rrow.p, g->rcol.p, g->queued.p, queue_words(g->V),
                    round_shape(g).ordered ? 1u : 0u};
}

uint32_t worklist_order_default() {
    if (const char* env = std::getenv("GEQ_WORKLIST_ORDER")) {
        const int v = std::atoi(env);
        if (v == 0 || v == 1) return (uint32_t)v;
    }
    return 2;
}
ExpandArgs expand_args(geq_game* g) {
    return ExpandArgs{{g->wl0.p, g->wl1.p}, nullptr, 0u, g->chg.p, g->rrow.p, g->rcol.p, g->queued.p, queue_words(g->V), g->ctl.p};
}

// The rounds are captured on the game's own stream, never on the context's
// shared one: host threads solving other games may enqueue on that stream
// meanwhile, and a capture would take their work in.  (Capturing records;
// nothing runs on the capture stream.)
// One batch of rounds_per_launch rounds and the k_advance after them, on s:
// what a round graph holds.
void launch_rounds(geq_game* g, const EvalArgs& ea, const ExpandArgs& xa, const RoundShape& shape, hipStream_t s) {
    const GenArgs ga = g->conf.generic ? gen_args(g) : GenArgs{};
    for (uint32_t k = 0; k < shape.rounds_per_launch; k++) {
        if (g->conf.generic) {
            launch_list_eval(shape, ga, k, s);
            k_expand<<<shape.expand_blocks, shape.threads, 0, s>>>(xa, k);
        } else if (shape.fused) {
            k_eval<true><<<shape.eval_blocks, shape.threads, (shape.threads / 64) * EVAL_FUSED_LDS_PER_WAVE, s>>>(ea, k);
            if (shape.ordered)
                k_collect<<<shape.collect_blocks, COLLECT_THREADS, 0, s>>>(CollectArgs{{g->wl0.p, g->wl1.p}, g->queued.p,
                                                                                        queue_words(g->V), g->ctl.p}, k);
        } else {
            k_eval<false><<<shape.eval_blocks, shape.threads, (shape.threads / 64) * EVAL_LDS_PER_WAVE, s>>>(ea, k);
            k_expand<<<shape.expand_blocks, shape.threads, 0, s>>>(xa, k);
        }
    }
    k_advance<<<1, 64, 0, s>>>(g->ctl.p, shape.rounds_per_launch);
}

geq_status capture_rounds(geq_game* g, const EvalArgs& ea, const ExpandArgs& xa, const RoundShape& shape) {
    if (!g->cap_stream) GEQ_HIP(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking), GEQ_ERR_NO_DEVICE);
    hipStream_t s = g->cap_stream;
    if (g->graph_exec) {
        (void)hipGraphExecDestroy(g->graph_exec);
        g->graph_exec = nullptr;
    }
    GEQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), GEQ_ERR_NO_DEVICE);
    launch_rounds(g, ea, xa, shape, s);
    hipGraph_t graph = nullptr;
    GEQ_HIP(hipStreamEndCapture(s, &graph), GEQ_ERR_NO_DEVICE);
    hipError_t e = hipGraphInstantiate(&g->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(GEQ_ERR_NO_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    g->graph_arena = g->conf.generic ? (void*)g->garena.p : (void*)g->arena.p;
    g->graph_scr = g->gscr.p;
    g->graph_scr_words = g->scr_words;
    return GEQ_OK;
}

// After a batch of rounds stopped at an arena overflow (ctl->round is the
// round that overflowed, nothing in flight): grow the arena and reset what
// the interrupted round left behind, so that it runs again from its start.
// Its half-written slots carry its own tag and are ignored by the rerun; its
// appends to the next worklist are dropped with their queued bits (the rerun
// clears the bits of its own list again as it evaluates it).
geq_status recover_overflow(geq_game* g) {
    hipStream_t s = g->ctx->stream;
    geq_status st = GEQ_OK;
    if (!g->conf.generic || g->h_ctl->arena_top > g->garena_cap) st = grow_arena(g, g->h_ctl->arena_top);
    if (st != GEQ_OK) return st;
    if (g->conf.generic && g->h_ctl->scr_need > g->scr_words) {
        // a node's candidates did not fit its wave's scratch
        st = ensure_scratch(g, round_shape(g), std::max<uint64_t>(g->h_ctl->scr_need, 2ull * g->scr_words));
        if (st != GEQ_OK) return st;
    }
    Ctl fix = *g->h_ctl;
    fix.overflow = 0;
    fix.scr_need = 0;
    fix.wl_n[(fix.round + 1u) % 3u] = 0;
    GEQ_HIP(hipMemsetAsync(g->queued.p, 0, 8ull * queue_words(g->V), s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipMemcpyAsync(g->ctl.p, &fix, sizeof(Ctl), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    *g->h_ctl = fix;
    return GEQ_OK;
}

// Between batches of rounds: re-tag every node's slots once RETAG_ROUNDS
// rounds have passed since the last re-tag, so slot ages stay below 2^16.
geq_status maybe_retag(geq_game* g) {
    const uint32_t r = g->h_ctl->round;
    if (r - g->retag_round < retag_interval()) return GEQ_OK;
    if (g->V) k_retag<<<(g->V + 255) / 256, 256, 0, g->ctx->stream>>>(g->st.p, g->V, round_tag(r));
    GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
    g->retag_round = r;
    return GEQ_OK;
}

geq_status solve_games(geq_game* g, EvalArgs ea, const uint32_t* seed, uint32_t n_reach, uint64_t* launches, hipEvent_t ev_end) {
    hipStream_t s = g->ctx->stream;
    const bool batch = g->game_base.size() >= 2;
    const uint32_t G = batch ? (uint32_t)g->game_base.size() - 1 : 1u;
    hipError_t e;
    if (!g->pg_ready) {
        std::vector<uint32_t> base = batch ? g->game_base : std::vector<uint32_t>{0u, g->V};
        std::vector<uint32_t> order(G);
        for (uint32_t i = 0; i < G; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return base[x + 1] - base[x] > base[y + 1] - base[y]; });
        if ((e = g->pg_base.alloc(G + 1)) != hipSuccess || (e = g->pg_order.alloc(G)) != hipSuccess ||
            (e = g->pg_soff.alloc(G + 1)) != hipSuccess || (e = g->pg_cnt.alloc(G)) != hipSuccess ||
            (e = g->pg_next.alloc(1)) != hipSuccess || (e = g->pg_ctl.alloc(G)) != hipSuccess)
            return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("per-game solve state: ") + hipGetErrorString(e));
        GEQ_HIP(hipMemcpyAsync(g->pg_base.p, base.data(), 4ull * (G + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipMemcpyAsync(g->pg_order.p, order.data(), 4ull * G, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        g->pg_ready = true;
    }
    // the seeds grouped by game (in any order within a game: a list's order
    // changes neither the values nor the counts of a round); the grouping of
    // the game's standard reach is kept for its next solves
    const bool std_seeds = n_reach && seed == g->std_seed.p && n_reach == g->std_n;
    geq_status st = GEQ_OK;
    if (!(std_seeds && g->pg_std_grouped)) {
        g->pg_std_grouped = false;
        if (n_reach && (!g->pg_seed.p || g->pg_seed.n < n_reach)) {
            g->pg_seed.release();
            if ((e = g->pg_seed.alloc(n_reach)) != hipSuccess)
                return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("per-game seeds: ") + hipGetErrorString(e));
        }
        const uint32_t tb = 256;
        GEQ_HIP(hipMemsetAsync(g->pg_cnt.p, 0, 4ull * G, s), GEQ_ERR_TRANSFER);
        if (n_reach) k_seed_count<<<(n_reach + tb - 1) / tb, tb, 0, s>>>(seed, n_reach, g->pg_base.p, G, g->pg_cnt.p);
        st = exclusive_scan_dev<uint32_t, SumU32>(g->pg_cnt.p, g->pg_soff.p, G, g->pg_scan, s, SumU32{});
        if (st != GEQ_OK) return st;
        if (n_reach) {
            GEQ_HIP(hipMemsetAsync(g->pg_cnt.p, 0, 4ull * G, s), GEQ_ERR_TRANSFER);
            k_seed_scatter<<<(n_reach + tb - 1) / tb, tb, 0, s>>>(seed, n_reach, g->pg_base.p, G, g->pg_soff.p, g->pg_cnt.p,
                                                                  g->pg_seed.p);
        }
        g->pg_std_grouped = std_seeds;
    }
    static_assert(GAME_NEW == 0, "the per-game state starts zeroed");
    GEQ_HIP(hipMemsetAsync(g->pg_ctl.p, 0, sizeof(GameCtl) * G, s), GEQ_ERR_TRANSFER);
    const uint32_t w = std::max(1u, std::min<uint32_t>(g->tune.waves_per_block, MAX_WAVES));
    const size_t lds = (size_t)w * EVAL_FUSED_LDS_PER_WAVE;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_solve_games), (int)(64 * w), lds) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    // no block waits for another, so the grid need not be resident at once;
    // more blocks than games would find the queue empty
    const uint32_t grid = std::max(1u, std::min({G, (uint32_t)std::max(1, g->ctx->n_cu) * (uint32_t)per_cu, STAT_SLOTS}));
    ea.ordered = 0;   // the expansion appends; no k_collect
    GamesArgs ga{g->pg_base.p, g->pg_order.p, g->pg_seed.p ? g->pg_seed.p : g->pg_soff.p, g->pg_soff.p, g->pg_ctl.p,
                 g->pg_next.p, G, retag_interval()};
    while (true) {
        GEQ_HIP(hipMemsetAsync(g->pg_next.p, 0, 4, s), GEQ_ERR_TRANSFER);
        k_solve_games<<<grid, 64 * w, lds, s>>>(ea, ga);
        GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
        (*launches)++;
        // (the solve's end, waited for with the control block just below)
        GEQ_HIP(hipEventRecord(ev_end, s), GEQ_ERR_NO_DEVICE);
        k_reduce_stats<<<1, 256, 0, s>>>(g->stats.p, grid, g->ctl.p->totals);
        GEQ_HIP(hipMemcpyAsync(g->h_ctl, g->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
        if (!g->h_ctl->overflow) {
            // every game ran rounds of its own: the download reads each node's
            // younger slot at round min(rounds, retag_interval) (games that
            // passed retag_interval re-tagged themselves to round 0)
            g->h_ctl->round = std::min<uint32_t>(g->h_ctl->last_nonempty, retag_interval());
            return GEQ_OK;
        }
        // the games stopped at a round boundary (the round that saw the
        // overflow runs again); a larger arena, cleared queued bitmaps, and
        // the games not done yet go on
        st = recover_overflow(g);
        if (st != GEQ_OK) return st;
        ea.arena = g->arena.p;
        ea.arena_cap = g->arena_cap;
    }
}

uint32_t retag_interval() {
    // GEQ_RETAG_ROUNDS (1 .. RETAG_ROUNDS) re-tags more often, for tests
    if (const char* env = std::getenv("GEQ_RETAG_ROUNDS")) {
        const long x = std::atol(env);
        if (x >= 1 && x <= (long)RETAG_ROUNDS) return (uint32_t)x;
    }
    return RETAG_ROUNDS;
}

}  // namespace geq

extern "C" geq_status geq_game_solve(geq_game* g, const uint32_t* to_reach, uint32_t n_reach, const geq_energies* initial,
                                     int profile_kernels, geq_stats* stats) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "game is NULL");
    if (n_reach && !to_reach) return fail(GEQ_ERR_INVALID_ARG, "to_reach is NULL");
    for (uint32_t i = 0; i < n_reach; i++)
        if (to_reach[i] >= g->V) return fail(GEQ_ERR_INVALID_ARG, "to_reach id out of range");
    (void)hipSetDevice(g->ctx->device);
    std::vector<uint32_t> mapped;     // the internal ids of a renumbered game
    if (n_reach && !g->inv.empty()) {
        mapped.resize(n_reach);
        for (uint32_t i = 0; i < n_reach; i++) mapped[i] = g->inv[to_reach[i]];
        to_reach = mapped.data();
    }
    if (n_reach) {
        if (!g->seed.p || g->seed.n < n_reach) GEQ_HIP(g->seed.alloc(n_reach), GEQ_ERR_OUT_OF_MEMORY);
        GEQ_HIP(hipMemcpyAsync(g->seed.p, to_reach, 4ull * n_reach, hipMemcpyHostToDevice, g->ctx->stream), GEQ_ERR_TRANSFER);
        if (!mapped.empty()) GEQ_HIP(hipStreamSynchronize(g->ctx->stream), GEQ_ERR_TRANSFER);   // `mapped` is freed on return
    }
    return geq::solve_core(g, n_reach, initial, profile_kernels, stats);
}

extern "C" geq_status geq_game_solve_standard(geq_game* g, int profile_kernels, geq_stats* stats) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "game is NULL");
    (void)hipSetDevice(g->ctx->device);
    const geq_status st = geq::standard_seed(g);
    if (st != GEQ_OK) return st;
    return geq::solve_core(g, g->std_n, nullptr, profile_kernels, stats, g->std_seed.p);
}

// The fixpoint on a resident game whose n_reach nodes to reach are already in g->seed.
namespace {
// Defender positions without successors.  A block of 256 threads takes
// 256 * REACH_PER consecutive nodes, counts its matches, reserves their room
// with one atomic and writes them in node order.
constexpr uint32_t REACH_PER = 16;
__global__ __launch_bounds__(256) void k_standard_reach(const uint32_t* row, const uint8_t* atk, uint32_t V, uint32_t* out,
                                                        uint32_t* n) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t v0 = (blockIdx.x * 256u + threadIdx.x) * REACH_PER;
    auto hit = [&](uint32_t v) { return v < V && !atk[v] && row[v] == row[v + 1]; };
    uint32_t c = 0;
    for (uint32_t k = 0; k < REACH_PER; k++) c += hit(v0 + k) ? 1u : 0u;
    const uint32_t incl = wave_incl_scan(c, lane);
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(n, tot) : 0u;
    }
    __syncthreads();
    uint32_t off = s_base + incl - c;
    for (uint32_t w = 0; w < wid; w++) off += s_wave[w];
    for (uint32_t k = 0; k < REACH_PER; k++)
        if (hit(v0 + k)) out[off++] = v0 + k;
}
}  // namespace

geq_status geq::standard_seed(geq_game* g) {
    if (g->std_ready) return GEQ_OK;
    g->pg_std_grouped = false;   // a new list: solve mode 3 groups it again
    hipStream_t s = g->ctx->stream;
    const uint32_t V = g->V;
    if (g->std_seed.n < std::max<uint32_t>(V, 1) + 1ull)
        GEQ_HIP(g->std_seed.alloc(std::max<uint32_t>(V, 1) + 1), GEQ_ERR_OUT_OF_MEMORY);
    uint32_t* cnt = g->std_seed.p + std::max<uint32_t>(V, 1);   // the count sits after the list
    GEQ_HIP(hipMemsetAsync(cnt, 0, 4, s), GEQ_ERR_TRANSFER);
    const uint32_t Vr = std::min(V, g->reach_limit);   // the nodes that may be reached
    if (Vr) k_standard_reach<<<(Vr + 256 * REACH_PER - 1) / (256 * REACH_PER), 256, 0, s>>>(g->row.p, g->atk.p, Vr, g->std_seed.p, cnt);
    GEQ_HIP(hipMemcpyAsync(&g->std_n, cnt, 4, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    g->std_ready = true;
    return GEQ_OK;
}

geq_status geq::solve_core(geq_game* g, uint32_t n_reach, const geq_energies* initial, int profile_kernels, geq_stats* stats,
                           const uint32_t* seed) {
    if (!seed) seed = g->seed.p;
    const uint32_t V = g->V;
    geq_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;

    // ---- initial state: EnergyGame::with_reach (src/energygame.rs:355-365) or explicit lists
    std::vector<uint16_t> h_arena;
    std::vector<uint32_t> h_words;   // generic confs: the initial lists' words
    if (initial && g->conf.generic) {
        const uint32_t esz = g->conf.esz;
        if (initial->n_nodes != V || !initial->offsets || (initial->offsets[V] && !initial->data))
            return fail(GEQ_ERR_INVALID_ARG, "initial energies must cover every node");
        if (initial->energy_size != esz)
            return fail(GEQ_ERR_INVALID_ARG, "initial energies: energy_size must be " + std::to_string(esz) + " for this conf");
        std::vector<uint4> h_st(2ull * V, make_uint4(0, 0, 0, 0));   // slot 0 current (tags 0), slot 1 empty
        for (uint32_t v = 0; v < V; v++) {
            const uint32_t b = initial->offsets[v], e = initial->offsets[v + 1];
            if (e < b || e - b > GLEN_MASK) return fail(GEQ_ERR_INVALID_ARG, "initial energies: bad offsets");
            const uint64_t off = GEN_ZERO_WORDS + h_words.size();
            if (off + (uint64_t)(e - b) * esz > 0x7fffffffull) return fail(GEQ_ERR_OUT_OF_MEMORY, "initial energies too large");
            for (uint32_t j = b; j < e; j++) {
                const uint32_t* w = initial->data + (size_t)j * esz;
                if (!energy_valid(g->conf, w)) return fail(GEQ_ERR_INVALID_ARG, "initial energies: invalid energy");
                h_words.insert(h_words.end(), w, w + esz);
            }
            h_st[2ull * (g->inv.empty() ? v : g->inv[v])] = make_uint4(e - b, (uint32_t)off, 0u, 0u);   // the node's internal id
        }
        if (GEN_ZERO_WORDS + h_words.size() + (1u << 16) > g->garena_cap) {
            geq_status st = grow_arena(g, (uint32_t)(GEN_ZERO_WORDS + h_words.size()));
            if (st != GEQ_OK) return st;
        }
        if (V)
            GEQ_HIP(hipMemcpy2DAsync(g->st.p + 1, 16ull * NODE_U4, h_st.data(), 32, 32, V, hipMemcpyHostToDevice, s),
                    GEQ_ERR_TRANSFER);
        if (!h_words.empty())
            GEQ_HIP(hipMemcpyAsync(g->garena.p + GEN_ZERO_WORDS, h_words.data(), 4ull * h_words.size(), hipMemcpyHostToDevice, s),
                    GEQ_ERR_TRANSFER);
    } else if (initial) {
        if (initial->n_nodes != V || !initial->offsets || (initial->offsets[V] && !initial->data))
            return fail(GEQ_ERR_INVALID_ARG, "initial energies must cover every node");
        if (initial->energy_size != 1) return fail(GEQ_ERR_INVALID_ARG, "initial energies: energy_size must be 1 for this conf");
        std::vector<uint4> h_st(2ull * V, make_uint4(0, 0, 0, 0));   // slot 0 current (tags 0), slot 1 empty
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
                h[2] = (uint16_t)(off & 0xffff);   // second dword of the record = arena offset
                h[3] = (uint16_t)(off >> 16);
            }
            std::memcpy(&h_st[2ull * (g->inv.empty() ? v : g->inv[v])], h, 16);   // the node's internal id
        }
        if (h_arena.size() + (1u << 20) > g->arena_cap) {
            geq_status st = grow_arena(g, (uint32_t)h_arena.size());
            if (st != GEQ_OK) return st;
        }
        // both slots of every node, leaving the headers in place
        if (V)
            GEQ_HIP(hipMemcpy2DAsync(g->st.p + 1, 16ull * NODE_U4, h_st.data(), 32, 32, V, hipMemcpyHostToDevice, s),
                    GEQ_ERR_TRANSFER);
        if (!h_arena.empty())
            GEQ_HIP(hipMemcpyAsync(g->arena.p, h_arena.data(), 2ull * h_arena.size(), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    Ctl c0{};
    c0.arena_top = g->conf.generic ? (uint32_t)(GEN_ZERO_WORDS + h_words.size()) : (uint32_t)h_arena.size();
    g->retag_round = 0;
    {
        const uint32_t q2 = 2u * queue_words(V), ns = 4u * STAT_SLOTS + 4u;
        const uint32_t most = std::max(std::max(V, q2), ns);
        k_solve_init<<<std::min(65536u, (most + 255) / 256), 256, 0, s>>>(g->st.p, V, initial ? 0 : 1, g->queued.p, q2, g->stats.p,
                                                                         ns, g->ctl.p, c0.arena_top);
    }
    if (!initial && n_reach) k_seed_zero<<<(n_reach + 255) / 256, 256, 0, s>>>(g->st.p, seed, n_reach);

    if (!g->ev_start) GEQ_HIP(hipEventCreate(&g->ev_start), GEQ_ERR_NO_DEVICE);
    if (!g->ev_end) GEQ_HIP(hipEventCreate(&g->ev_end), GEQ_ERR_NO_DEVICE);
    hipEvent_t ev_start = g->ev_start, ev_end = g->ev_end;
    std::vector<hipEvent_t> prof;   // 4 events per round: around k_eval, around k_expand
    GEQ_HIP(hipEventRecord(ev_start, s), GEQ_ERR_NO_DEVICE);

    const RoundShape shape = round_shape(g);
    EvalArgs ea = eval_args(g);
    ExpandArgs xa = expand_args(g);
    if (g->conf.generic) {
        const geq_status st = ensure_scratch(g, shape, g->scr_words ? g->scr_words : 16384u);
        if (st != GEQ_OK) return st;
    }

    const bool games_mode = g->tune.mode == SOLVE_MODE_GAMES && !profile_kernels;
    if (games_mode && g->conf.generic)
        return fail(GEQ_ERR_UNSUPPORTED, "the per-game persistent solve (GEQ_SOLVE_MODE=3) covers the bitmap energy layout only");
    // seed: predecessors of the nodes to reach (initialize_visit_lists, :1461-1465)
    if (n_reach && !games_mode) {
        ExpandArgs sx = xa;
        sx.seed = seed;
        sx.n_seed = n_reach;
        k_expand<<<std::max(1u, std::min(shape.expand_blocks, (n_reach + 64u * WAVES - 1) / (64u * WAVES))), BLOCK, 0, s>>>(sx, 0u);
    }
    // No readback here: with seeds the first batch of rounds runs (a batch
    // over an empty worklist only costs its empty launches), without them
    // nothing does.
    *g->h_ctl = c0;
    g->h_ctl->wl_n[0] = n_reach ? 1u : 0u;

    uint64_t launches = 0;
    geq_status result = GEQ_OK;
    bool stats_in_ctl = false;   // h_ctl->totals hold the statistics of every round that ran
    if (games_mode) {
        result = solve_games(g, ea, seed, n_reach, &launches, ev_end);
        stats_in_ctl = result == GEQ_OK;
    }
    // GEQ_SOLVE_TRACE=1 with profiled kernels: one stderr line per round
    // (worklist size, k_eval and expansion times)
    const char* trace_env = std::getenv("GEQ_SOLVE_TRACE");
    const bool trace = profile_kernels && trace_env && std::atoi(trace_env) != 0;
    while (!games_mode && g->h_ctl->wl_n[g->h_ctl->round % 3u] != 0) {
        stats_in_ctl = false;
        const uint32_t trace_round = g->h_ctl->round, trace_n = g->h_ctl->wl_n[g->h_ctl->round % 3u];
        if (profile_kernels) {
            // one round per launch; each kernel's dispatch carries its own
            // start / end events (hipExtLaunchKernelGGL), so the times are the
            // kernels' own, without the gaps between dispatches
            hipEvent_t ev[4];
            for (hipEvent_t& e : ev) GEQ_HIP(hipEventCreate(&e), GEQ_ERR_NO_DEVICE);
            if (g->conf.generic) {
                const GenArgs ga = gen_args(g);
                if (ga.c.esz == 1)
                    hipExtLaunchKernelGGL(k_eval_list<1>, dim3(shape.eval_blocks), dim3(GEN_BLOCK), 0, s, ev[0], ev[1], 0, ga, 0u);
                else if (ga.c.esz <= 4)
                    hipExtLaunchKernelGGL(k_eval_list<4>, dim3(shape.eval_blocks), dim3(GEN_BLOCK), 0, s, ev[0], ev[1], 0, ga, 0u);
                else
                    hipExtLaunchKernelGGL(k_eval_list<32>, dim3(shape.eval_blocks), dim3(GEN_BLOCK), 0, s, ev[0], ev[1], 0, ga, 0u);
                hipExtLaunchKernelGGL(k_expand, dim3(shape.expand_blocks), dim3(shape.threads), 0, s, ev[2], ev[3], 0, xa, 0u);
            } else if (shape.fused) {
                hipExtLaunchKernelGGL(k_eval<true>, dim3(shape.eval_blocks), dim3(shape.threads),
                                      (shape.threads / 64) * EVAL_FUSED_LDS_PER_WAVE, s, ev[0], ev[1], 0, ea, 0u);
                GEQ_HIP(hipEventRecord(ev[2], s), GEQ_ERR_NO_DEVICE);
                if (shape.ordered)   // (its time goes with the expansion's)
                    k_collect<<<shape.collect_blocks, COLLECT_THREADS, 0, s>>>(
                        CollectArgs{{g->wl0.p, g->wl1.p}, g->queued.p, queue_words(g->V), g->ctl.p}, 0u);
                GEQ_HIP(hipEventRecord(ev[3], s), GEQ_ERR_NO_DEVICE);
            } else {
                hipExtLaunchKernelGGL(k_eval<false>, dim3(shape.eval_blocks), dim3(shape.threads),
                                      (shape.threads / 64) * EVAL_LDS_PER_WAVE, s, ev[0], ev[1], 0, ea, 0u);
                hipExtLaunchKernelGGL(k_expand, dim3(shape.expand_blocks), dim3(shape.threads), 0, s, ev[2], ev[3], 0, xa, 0u);
            }
            GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
            k_advance<<<1, 64, 0, s>>>(g->ctl.p, 1u);
            for (hipEvent_t e : ev) prof.push_back(e);
            launches++;
        } else if (g->tune.mode == 0) {
            // The cooperative one-launch solve is switched off: its grid of
            // min(occupancy API, 8) blocks per CU could exceed the waves per
            // SIMD its register use allows, so its grid barrier could wait for
            // blocks that are never resident.
            result = fail(GEQ_ERR_UNSUPPORTED, "the cooperative solve mode (GEQ_SOLVE_MODE=0) is disabled");
            break;
        } else {
            // rounds_per_launch captured rounds per graph launch.  A game's
            // first solve launches the same kernels directly: a game solved
            // once does not pay for a capture; its next solve captures.
            if (!g->graph_exec && g->solves == 0) {
                launch_rounds(g, ea, xa, shape, s);
                GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
            } else {
                const bool stale = g->conf.generic ? (g->graph_arena != g->garena.p || g->graph_scr != g->gscr.p ||
                                                      g->graph_scr_words != g->scr_words)
                                                   : g->graph_arena != g->arena.p;
                if (!g->graph_exec || stale) {
                    result = capture_rounds(g, ea, xa, shape);
                    if (result != GEQ_OK) break;
                }
                GEQ_HIP(hipGraphLaunch(g->graph_exec, s), GEQ_ERR_NO_DEVICE);
            }
            launches += shape.rounds_per_launch;
            // the statistics so far travel with the control block: if this
            // batch was the last, the solve needs no further wait
            GEQ_HIP(hipEventRecord(ev_end, s), GEQ_ERR_NO_DEVICE);
            k_reduce_stats<<<1, 256, 0, s>>>(g->stats.p, shape.eval_blocks, g->ctl.p->totals);
            stats_in_ctl = true;
        }
        // the control block comes back through the mapped copy and its signal word
        const uint32_t seq = ++g->sig_seq;
        k_ctl_out<<<1, 64, 0, s>>>(g->ctl.p, g->d_ctl_host, seq);
        GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
        result = wait_signal(g, seq);
        if (result != GEQ_OK) break;
        if (trace && prof.size() >= 4) {
            float te = 0, tx = 0;
            (void)hipEventElapsedTime(&te, prof[prof.size() - 4], prof[prof.size() - 3]);
            (void)hipEventElapsedTime(&tx, prof[prof.size() - 2], prof[prof.size() - 1]);
            std::fprintf(stderr, "solve: round %u: %u nodes, eval %.1f us, expand %.1f us\n", trace_round, trace_n, 1e3 * te,
                         1e3 * tx);
        }
        if (g->h_ctl->overflow) {
            result = recover_overflow(g);
            if (result != GEQ_OK) break;
            ea.arena = g->arena.p;
            ea.arena_cap = g->arena_cap;
        }
        result = maybe_retag(g);
        if (result != GEQ_OK) break;
    }
    const uint64_t rounds = g->h_ctl->last_nonempty;
    if (result == GEQ_OK) g->solves++;
    unsigned long long h_stats[4] = {0, 0, 0, 0};
    if (stats_in_ctl && result == GEQ_OK) {
        for (int k = 0; k < 4; k++) h_stats[k] = g->h_ctl->totals[k];
    } else {
        // the statistics come back with one more synchronisation
        GEQ_HIP(hipEventRecord(ev_end, s), GEQ_ERR_NO_DEVICE);
        if (result == GEQ_OK) {
            k_reduce_stats<<<1, 256, 0, s>>>(g->stats.p, shape.eval_blocks, g->stats.p + 4ull * STAT_SLOTS);
            GEQ_HIP(hipMemcpyAsync(h_stats, g->stats.p + 4ull * STAT_SLOTS, sizeof(h_stats), hipMemcpyDeviceToHost, s),
                    GEQ_ERR_TRANSFER);
        }
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev_start, ev_end);
    double eval_ms = 0, expand_ms = 0;
    for (size_t i = 0; i + 3 < prof.size(); i += 4) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, prof[i], prof[i + 1]);
        (void)hipEventElapsedTime(&b, prof[i + 2], prof[i + 3]);
        eval_ms += a;
        expand_ms += b;
    }
    for (hipEvent_t ev : prof) (void)hipEventDestroy(ev);
    if (result != GEQ_OK) return result;
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
        stats->solve_kernel_ms = 0;
        stats->mode = profile_kernels ? 2u : g->tune.mode;
    }
    return GEQ_OK;
}

extern "C" geq_status geq_game_energies(geq_game* g, geq_energies* out) {
    if (!g || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    return geq::download_energies(g, 0, g->V, out);
}

namespace {
// download_energies for a generic conf: every node's list in the crate's
// packing, esz words per energy, energies in ascending lexicographic order
// of their words (word 0 first).
geq_status download_lists(geq_game* g, uint32_t first, uint32_t V, uint32_t t, geq_energies* out) {
    hipStream_t s = g->ctx->stream;
    const uint32_t esz = g->conf.esz;
    const uint4* stb = g->inv_map.p ? g->st.p : g->st.p + (size_t)NODE_U4 * first;
    const uint32_t* map = g->inv_map.p ? g->inv_map.p + first : nullptr;
    if (V) k_list_lens_g<<<(V + 255) / 256, 256, 0, s>>>(stb, V, t, g->lens.p, map);
    std::vector<uint32_t> lens(V), offs(V + 1);
    if (V) GEQ_HIP(hipMemcpyAsync(lens.data(), g->lens.p, 4ull * V, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    uint64_t tot = 0;
    for (uint32_t v = 0; v < V; v++) {
        offs[v] = (uint32_t)tot;
        tot += lens[v];
    }
    if (tot > 0xffffffffull || tot * esz > 0xffffffffull) return fail(GEQ_ERR_OUT_OF_MEMORY, "result exceeds 2^32 words");
    offs[V] = (uint32_t)tot;
    if (!g->offs.p || g->offs.n < V + 1ull) GEQ_HIP(g->offs.alloc(V + 1), GEQ_ERR_OUT_OF_MEMORY);
    if (!g->out.p || g->out.n < tot * esz) GEQ_HIP(g->out.alloc(std::max<uint64_t>(tot * esz, 1)), GEQ_ERR_OUT_OF_MEMORY);
    GEQ_HIP(hipMemcpyAsync(g->offs.p, offs.data(), 4ull * (V + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    if (V) k_gather_g<<<(V + 255) / 256, 256, 0, s>>>(stb, g->garena.p, V, t, esz, g->offs.p, g->out.p, map);
    uint32_t* data = (uint32_t*)std::malloc(4ull * std::max<uint64_t>(tot * esz, 1));
    uint32_t* ooffs = (uint32_t*)std::malloc(4ull * (V + 1));
    if (!data || !ooffs) {
        std::free(data);
        std::free(ooffs);
        return fail(GEQ_ERR_OUT_OF_MEMORY, "host allocation of the result");
    }
    hipError_t e = tot ? hipMemcpyAsync(data, g->out.p, 4ull * tot * esz, hipMemcpyDeviceToHost, s) : hipSuccess;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        std::free(data);
        std::free(ooffs);
        return fail(GEQ_ERR_TRANSFER, std::string("downloading energies: ") + hipGetErrorString(e));
    }
    std::memcpy(ooffs, offs.data(), 4ull * (V + 1));
    // ascending order per node (rows compared word by word)
    std::vector<uint32_t> idx, tmp;
    for (uint32_t v = 0; v < V; v++) {
        const uint32_t n = lens[v];
        if (n < 2) continue;
        uint32_t* rows = data + (size_t)offs[v] * esz;
        idx.resize(n);
        for (uint32_t i = 0; i < n; i++) idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
            return std::lexicographical_compare(rows + (size_t)x * esz, rows + (size_t)(x + 1) * esz, rows + (size_t)y * esz,
                                                rows + (size_t)(y + 1) * esz);
        });
        tmp.assign(rows, rows + (size_t)n * esz);
        for (uint32_t i = 0; i < n; i++) std::copy(tmp.begin() + (size_t)idx[i] * esz, tmp.begin() + (size_t)(idx[i] + 1) * esz,
                                                  rows + (size_t)i * esz);
    }
    out->n_nodes = V;
    out->energy_size = esz;
    out->offsets = ooffs;
    out->data = data;
    return GEQ_OK;
}
}  // namespace

// Energies of the nodes [first, first + V) as packed words, ascending per node.
// The game's mapped pinned staging, with room for `words` words.
geq_status geq::ensure_stage(geq_game* g, size_t words) {
    if (g->h_stage && g->h_stage_n >= words) return GEQ_OK;
    if (g->h_stage) (void)hipHostFree(g->h_stage);
    g->h_stage = g->d_stage = nullptr;
    g->h_stage_n = 0;
    const size_t cap = std::max<size_t>(words, 1024);
    GEQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->h_stage), 4 * cap, hipHostMallocMapped | hipHostMallocCoherent),
            GEQ_ERR_OUT_OF_MEMORY);
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&g->d_stage), g->h_stage, 0) != hipSuccess || !g->d_stage) {
        (void)hipHostFree(g->h_stage);
        g->h_stage = g->d_stage = nullptr;
        return fail(GEQ_ERR_INTERNAL, "no device view of the download staging");
    }
    g->h_stage_n = cap;
    return GEQ_OK;
}

// Preorder verdicts of nodes [first, first + count) against energies[0 .. n_eqs)
// (EnergyArray::test_equivalence, src/energy.rs:307-318, on each node's list),
// computed on the device: only the verdict bytes come back.
geq_status geq::game_preorders(geq_game* g, uint32_t first, uint32_t count, const uint32_t* energies, uint32_t n_eqs,
                               uint8_t* out) {
    if (!g || (count && n_eqs && !out) || (n_eqs && !energies)) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if ((uint64_t)first + count > g->V) return fail(GEQ_ERR_INVALID_ARG, "node range out of bounds");
    if (n_eqs > EQ_WORDS_MAX) return fail(GEQ_ERR_INVALID_ARG, "at most 64 energies per call");
    if (g->conf.generic)
        return fail(GEQ_ERR_UNSUPPORTED, "preorder verdicts on the device take confs of at most 6 elements and max 3");
    EqWords ew{};
    ew.n = n_eqs;
    for (uint32_t k = 0; k < n_eqs; k++)
        if (!energy_valid(g->conf, energies + k) || !energy_to_canon(g->conf, energies[k], ew.w[k]))
            return fail(GEQ_ERR_INVALID_ARG, "energy " + std::to_string(k) + " is not packed for the game's EnergyConf");
    if (!count || !n_eqs) return GEQ_OK;
    (void)hipSetDevice(g->ctx->device);
    hipStream_t s = g->ctx->stream;
    const uint32_t t = round_tag(g->h_ctl->round);   // the lists after the last round that ran
    const uint4* stb = g->inv_map.p ? g->st.p : g->st.p + (size_t)NODE_U4 * first;
    const uint32_t* map = g->inv_map.p ? g->inv_map.p + first : nullptr;
    const size_t bytes = (size_t)count * n_eqs;
    geq_status est = ensure_stage(g, (bytes + 3) / 4);
    if (est != GEQ_OK) return est;
    k_preorders<<<(count + 255) / 256, 256, 0, s>>>(stb, g->arena.p, count, t, map, ew, g->conf.elements,
                                                   reinterpret_cast<uint8_t*>(g->d_stage));
    GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    std::memcpy(out, g->h_stage, bytes);
    return GEQ_OK;
}

geq_status geq::download_energies(geq_game* g, uint32_t first, uint32_t V, geq_energies* out) {
    if (!g || !out || (uint64_t)first + V > g->V) return fail(GEQ_ERR_INVALID_ARG, "node range out of bounds");
    (void)hipSetDevice(g->ctx->device);
    if (!g->lens.p || g->lens.n < V) GEQ_HIP(g->lens.alloc(std::max<uint32_t>(V, 1)), GEQ_ERR_OUT_OF_MEMORY);
    const uint32_t t = round_tag(g->h_ctl->round);   // the lists after the last round that ran
    if (g->conf.generic) return download_lists(g, first, V, t, out);
    // renumbered games: output entry v is internal node inv_map[first + v]
    const uint4* stb = g->inv_map.p ? g->st.p : g->st.p + (size_t)NODE_U4 * first;
    const uint32_t* map = g->inv_map.p ? g->inv_map.p + first : nullptr;
    return download_mapped(g, stb, map, V, t, out);
}

// The lists of nodes map[0 .. V) (of nodes stb[0 .. V) without a map) at round tag t.
// Lengths -> offsets (device scan) -> lists gathered, ascending; offsets and
// lists come back through the game's mapped pinned staging, after one wait.
geq_status geq::download_mapped(geq_game* g, const uint4* stb, const uint32_t* map, uint32_t V, uint32_t t, geq_energies* out) {
    hipStream_t s = g->ctx->stream;
    if (!g->lens.p || g->lens.n < V) GEQ_HIP(g->lens.alloc(std::max<uint32_t>(V, 1)), GEQ_ERR_OUT_OF_MEMORY);
    if (!g->offs.p || g->offs.n < V + 1ull) GEQ_HIP(g->offs.alloc(V + 1), GEQ_ERR_OUT_OF_MEMORY);
    if (V) k_list_lens<<<(V + 255) / 256, 256, 0, s>>>(stb, V, t, g->lens.p, map);
    geq_status sst = exclusive_scan_dev<uint32_t, SumU32>(g->lens.p, g->offs.p, V, g->scan_tmp, s, SumU32{});
    if (sst != GEQ_OK) return sst;
    // room for the gathered lists: every list is inline (<= INLINE_CAP) or in the arena
    const uint64_t bound = (uint64_t)INLINE_CAP * V + std::min<uint64_t>((uint64_t)g->h_ctl->arena_top, (uint64_t)g->arena_cap);
    if (!g->out.p || g->out.n < bound) GEQ_HIP(g->out.alloc(std::max<uint64_t>(bound, 1)), GEQ_ERR_OUT_OF_MEMORY);
    if (V) k_gather<<<(V + 255) / 256, 256, 0, s>>>(stb, g->arena.p, V, t, g->offs.p, g->out.p, bound, map);
    // offsets and lists written by one copy kernel into mapped pinned memory
    // (room for the bound), then one wait
    const size_t words = V + 1ull + bound;
    geq_status est = ensure_stage(g, words);
    if (est != GEQ_OK) return est;
    k_stage_out<<<std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)((words + 255) / 256), 2048)), 256, 0, s>>>(
        g->offs.p, V, g->out.p, bound, g->d_stage, g->d_stage + V + 1);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    const uint64_t tot = g->h_stage[V];
    if (tot > bound) return fail(GEQ_ERR_INTERNAL, "energy download: lists exceed their bound");
    uint32_t* ooffs = (uint32_t*)std::malloc(4ull * (V + 1));
    uint32_t* data = (uint32_t*)std::malloc(4ull * (tot ? tot : 1));
    if (!data || !ooffs) {
        std::free(data);
        std::free(ooffs);
        return fail(GEQ_ERR_OUT_OF_MEMORY, "host allocation of the result");
    }
    std::memcpy(ooffs, g->h_stage, 4ull * (V + 1));
    if (tot) std::memcpy(data, g->h_stage + V + 1, 4ull * tot);
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

extern "C" geq_status geq_batch_solve(geq_game* b, const uint32_t* const* to_reach, const uint32_t* n_reach, int profile_kernels,
                                      geq_stats* stats) {
    if (b && !to_reach && !n_reach) {
        // every game's standard_reach, kept on the device
        (void)hipSetDevice(b->ctx->device);
        geq_status st = geq::standard_seed(b);
        if (st != GEQ_OK) return st;
        return geq::solve_core(b, b->std_n, nullptr, profile_kernels, stats, b->std_seed.p);
    }
    if (!b || !n_reach) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    const uint32_t n_games = (uint32_t)b->game_base.size() - 1;
    std::vector<uint32_t> reach;
    for (uint32_t i = 0; i < n_games; i++) {
        if (n_reach[i] && (!to_reach || !to_reach[i])) return fail(GEQ_ERR_INVALID_ARG, "to_reach list is NULL");
        const uint32_t Vi = b->game_base[i + 1] - b->game_base[i];
        for (uint32_t k = 0; k < n_reach[i]; k++) {
            if (to_reach[i][k] >= Vi) return fail(GEQ_ERR_INVALID_ARG, "to_reach id out of range in game " + std::to_string(i));
            reach.push_back(b->game_base[i] + to_reach[i][k]);
        }
    }
    return geq_game_solve(b, reach.data(), (uint32_t)reach.size(), nullptr, profile_kernels, stats);
}

extern "C" geq_status geq_batch_energies(geq_game* b, geq_energies* outs) {
    if (!b || !outs) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    geq_energies all{};
    geq_status st = geq_game_energies(b, &all);
    if (st != GEQ_OK) return st;
    const uint32_t n_games = (uint32_t)b->game_base.size() - 1;
    const uint32_t esz = all.energy_size;
    for (uint32_t i = 0; i < n_games; i++) {
        const uint32_t v0 = b->game_base[i], v1 = b->game_base[i + 1], n = v1 - v0;
        const uint32_t e0 = all.offsets[v0], e1 = all.offsets[v1];
        outs[i].n_nodes = n;
        outs[i].energy_size = esz;
        outs[i].offsets = (uint32_t*)std::malloc(4ull * (n + 1));
        outs[i].data = (uint32_t*)std::malloc(4ull * std::max<uint64_t>(1, (uint64_t)(e1 - e0) * esz));
        if (!outs[i].offsets || !outs[i].data) {
            for (uint32_t j = 0; j <= i; j++) geq_energies_free(&outs[j]);
            geq_energies_free(&all);
            return fail(GEQ_ERR_OUT_OF_MEMORY, "host allocation of batch results");
        }
        for (uint32_t v = 0; v <= n; v++) outs[i].offsets[v] = all.offsets[v0 + v] - e0;
        if (e1 > e0) std::memcpy(outs[i].data, all.data + (size_t)e0 * esz, 4ull * (e1 - e0) * esz);
    }
    geq_energies_free(&all);
    return GEQ_OK;
}
