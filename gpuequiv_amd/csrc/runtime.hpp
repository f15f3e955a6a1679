// Host-side runtime of the solver: context, error reporting, device games.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpuequiv.h"
#include "solve_kernels.hpp"

namespace geq {

void set_error(const std::string& msg);
geq_status fail(geq_status s, const std::string& msg);

#define GEQ_HIP(expr, code)                                                                  \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return ::geq::fail((code), std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

// Energy configuration of a game.  elements <= 6 and max <= 3 run on the
// bitmap kernels (2-bit canonical fields); every other conf whose energies
// take at most GEN_MAX_WORDS words runs on the list kernels in the crate's
// own packing (`generic`).
struct ConfInfo {
    uint32_t elements, max;
    uint32_t ebits, emask, ubits, umask;
    uint32_t esz = 1, usz = 1;   // words per energy / per update (EnergyConf::energy_size, update_size)
    bool generic = false;
};
geq_status check_conf(const geq_conf& c, ConfInfo& out);
// the list kernels' view of a generic conf
GConf gconf_of(const ConfInfo& c);
// a packed energy / update of esz / usz words is well formed for the conf
// (fields within range, unused bits zero)
bool energy_valid(const ConfInfo& c, const uint32_t* words);
bool update_valid(const ConfInfo& c, const uint32_t* words);
// conf packing <-> canonical (2-bit energies, 4-bit updates)
bool energy_to_canon(const ConfInfo& c, uint32_t packed, uint32_t& canon);
uint32_t energy_from_canon(const ConfInfo& c, uint32_t canon);
bool update_to_canon(const ConfInfo& c, uint32_t packed, uint32_t& canon);

// hipMalloc that, when the device is out of memory, frees the game buffers
// every context on the current device keeps for reuse (release_built_game)
// and tries once more.  Every device allocation of the library goes through it.
hipError_t device_malloc(void** p, size_t bytes);
// Free the cached spare game of every context on `device`; true if one was freed.
bool drop_spare_games(int device);
// geq_solve_noalloc's set-up on the current device (at context creation):
// the device address of its static scratch array, once per device.
geq_status noalloc_setup(int device);

// Device buffers kept for reuse by later builds of one build scratch.
struct SpareBufs {
    struct Buf {
        void* p;
        size_t bytes;
    };
    std::vector<Buf> bufs;
    SpareBufs() = default;
    SpareBufs(const SpareBufs&) = delete;
    SpareBufs& operator=(const SpareBufs&) = delete;
    ~SpareBufs() { drop(); }
    // a buffer with room for `bytes` (and at most 4x as much), or NULL; its size in *got
    void* take(size_t bytes, size_t* got);
    bool drop();   // frees them all; true if there were any
    void trim();   // frees those past the limits below
};
// Spares kept: the largest buffers whose bytes fit SPARE_BYTES (at most
// SPARE_BUFS of them), and besides those up to SPARE_SMALL_BUFS buffers of at
// most SPARE_SMALL bytes each.  Every buffer freed is a hipFree, which waits
// for the whole device, so the many small buffers a growing build outgrows
// are kept rather than freed.
constexpr size_t SPARE_BUFS = 24;
constexpr size_t SPARE_BYTES = 4ull << 30;
constexpr size_t SPARE_SMALL_BUFS = 256;
constexpr size_t SPARE_SMALL = 16ull << 20;

// Buffers outgrown inside a build.  hipFree waits for every stream of the
// device, so a buffer that a build's queued kernels may still read is not
// freed when it is outgrown: while a RetireScope is open on the thread it is
// kept, and when the scope closes (after a wait on the scope's stream) it
// joins the scope's spare buffers, which the build allocations of later
// scopes on the same pool take before calling hipMalloc (spare_take).  Device
// memory exhaustion frees the open scope's spares.  Scopes nest: an inner
// scope hands its buffers to the outermost one.
class RetireScope {
  public:
    RetireScope(hipStream_t s, SpareBufs* pool);
    ~RetireScope();
    RetireScope(const RetireScope&) = delete;
    RetireScope& operator=(const RetireScope&) = delete;
    // true: p (of `bytes` bytes) is kept until the outermost scope closes;
    // false: no scope is open
    static bool take(void* p, size_t bytes);
    static bool active();

  private:
    hipStream_t s_;
    bool outer_;
};
// A spare buffer of the open scope's pool with room for `bytes` (NULL: none, or no scope).
void* spare_take(size_t bytes, size_t* got);
// Frees the open scope's spare buffers; true if there were any.
bool spare_drop();

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // release for a buffer that queued kernels may still use (RetireScope)
    void retire() {
        if (p && RetireScope::take(p, sizeof(T) * n)) {
            p = nullptr;
            n = 0;
        }
        release();
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return device_malloc(reinterpret_cast<void**>(&p), sizeof(T) * (count ? count : 1));
    }
};

// The fixpoint on a resident game seeded with the n_reach device node ids
// `seed` (g->seed when NULL).
geq_status solve_core(geq_game* g, uint32_t n_reach, const geq_energies* initial, int profile_kernels, geq_stats* stats,
                      const uint32_t* seed = nullptr);
// EnergyGame::standard_reach (src/energygame.rs:347-352) of a resident game:
// the defender positions without successors, found on the device once and
// kept in g->std_seed / g->std_n.
geq_status standard_seed(geq_game* g);
// A built game's buffers go to its context's cache (one game) for the next
// built game, or are freed when the cache is taken.
void release_built_game(geq_game* g);
// The game's pinned control block with its device view (alloc_game_buffers).
hipError_t alloc_host_ctl(geq_game* g);
// Waits until the game's signal word holds `seq` (written last by k_ctl_out
// on the game's stream): a spin of up to 2 ms, then a stream synchronisation.
geq_status wait_signal(geq_game* g, uint32_t seq);
namespace gb {
struct BuildScratch;
}
// Frees the spare buffers of a build scratch (true if there were any).
bool drop_build_spares(gb::BuildScratch* ws);
geq_status download_energies(geq_game* g, uint32_t first, uint32_t count, geq_energies* out);
// The game's mapped pinned staging with room for `words` words.
geq_status ensure_stage(geq_game* g, size_t words);
// Preorder verdicts of nodes [first, first + count) against n_eqs packed
// energies (bitmap-kernel confs): out[i * n_eqs + k] = 1 when no energy of
// node first + i lies below energy k, else 0.
geq_status game_preorders(geq_game* g, uint32_t first, uint32_t count, const uint32_t* energies, uint32_t n_eqs,
                          uint8_t* out);
// The same for any nodes of a game without renumbering and in the bitmap
// layout: entry v of the result is node map[v] (device ids); t is the tag of
// the round after the last one that ran.
geq_status download_mapped(geq_game* g, const uint4* stb, const uint32_t* map, uint32_t V, uint32_t t, geq_energies* out);
// Device-resident game from device arrays (STANDARD conf): rows, columns,
// packed updates, attacker flags; the reverse CSR is derived on the device.
geq_status game_from_device(geq_ctx* ctx, uint32_t V, uint32_t E, const uint32_t* d_row, const uint32_t* d_col,
                            const uint32_t* d_wt, const uint8_t* d_atk, geq_game** out);
// Energy-arena entries (u16) a game starts with: `want`, or one per two edges
// (at least 2^20) when 0;
// the environment variable GEQ_ARENA_ENTRIES overrides both (small values
// force the regrow path).
uint32_t initial_arena_entries(uint64_t E, uint64_t want);
// INVALID_ARG checks of geq_gamebuild's inputs (LTS shape, label order, pairs)
geq_status validate_job(const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs);
// Solver buffers for V positions / E edges (contents uninitialised).
geq_status alloc_game_buffers(geq_ctx* ctx, uint32_t V, uint32_t E, geq_game** out, uint64_t arena_entries = 0);
// Launch shape of one round and the hipGraph of `rounds_per_launch` captured rounds.
struct RoundShape {
    uint32_t eval_blocks, expand_blocks, rounds_per_launch;
    uint32_t threads;                 // block size of k_eval / k_expand
    bool fused;                       // k_eval expands too: one kernel per round
    bool ordered;                     // (fused only) k_collect lists the next round in id order
    uint32_t collect_blocks;
};
// games from this many positions list their rounds in id order by default
constexpr uint32_t ORDERED_MIN_V = 1u << 20;
// the worklist order of a new game: GEQ_WORKLIST_ORDER (0 / 1), else by size
uint32_t worklist_order_default();
RoundShape round_shape(const geq_game* g);
uint32_t eval_blocks_resident(bool fused, uint32_t waves_per_block);
// list kernels: the k_eval_list instantiation for esz-word energies, its
// blocks per CU, one launch of it, and the per-wave scratch (words per wave)
const void* list_kernel(uint32_t esz);
uint32_t list_blocks_resident(uint32_t esz);
void launch_list_eval(const RoundShape& shape, const GenArgs& ga, uint32_t k, hipStream_t s);
geq_status ensure_scratch(geq_game* g, const RoundShape& shape, uint64_t words);
EvalArgs eval_args(geq_game* g);
ExpandArgs expand_args(geq_game* g);
geq_status capture_rounds(geq_game* g, const EvalArgs& ea, const ExpandArgs& xa, const RoundShape& shape);
// Grow the energy arena (u16 entries, or words for a generic conf) past needed_top, keeping its content.
geq_status grow_arena(geq_game* g, uint32_t needed_top);
// words of the generic arena kept for the zero energy (its start)
constexpr uint32_t GEN_ZERO_WORDS = GEN_MAX_WORDS;
GenArgs gen_args(geq_game* g);
// ctl->overflow was raised: grow the arena and make the round run again
geq_status recover_overflow(geq_game* g);
// between batches of rounds: k_retag when due (slot tags stay unambiguous)
geq_status maybe_retag(geq_game* g);
uint32_t retag_interval();
// The persistent batched solve (solve mode 3, k_solve_games): the seeds
// grouped by game, then launches until no arena overflow is left.  *h_ctl
// holds the statistics and the round count afterwards.
geq_status solve_games(geq_game* g, EvalArgs ea, const uint32_t* seed, uint32_t n_reach, uint64_t* launches, hipEvent_t ev_end);
constexpr uint32_t SOLVE_MODE_GRAPH = 1, SOLVE_MODE_PROFILED = 2, SOLVE_MODE_GAMES = 3;

}  // namespace geq

struct geq_game;
namespace geq {
namespace gb {
struct BuildScratch;
}
}  // namespace geq

struct geq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    int n_cu = 0;
    // geq_gamebuild's device scratch, kept between builds (one build at a
    // time uses it; a concurrent build takes scratch of its own), and the
    // sizes of the last build, which size the next one's arrays up front
    std::mutex build_mu;
    std::shared_ptr<geq::gb::BuildScratch> build_ws;
    std::atomic<uint64_t> hint_positions{0}, hint_edges{0}, hint_words{0};
    // the device buffers of the last freed built game, for the next built game
    // that fits in them (game_from_device); guarded by game_mu
    std::mutex game_mu;
    geq_game* spare_game = nullptr;
};

struct geq_game {
    geq_ctx* ctx = nullptr;
    geq::ConfInfo conf{};
    uint32_t V = 0, E = 0;
    geq::DevBuf<uint32_t> row, rrow, rcol, wl0, wl1, seed, lens, offs, out;
    geq::DevBuf<uint32_t> queued;     // two bitmaps of queue_words(V) words
    geq::DevBuf<uint32_t> std_seed;   // standard_reach, once computed
    uint32_t std_n = 0;
    bool std_ready = false;
    uint32_t reach_limit = 0xffffffffu;   // standard_reach looks at the nodes below this id only
    geq::DevBuf<uint2> edges;
    geq::DevBuf<uint8_t> atk, chg;
    geq::DevBuf<uint4> st;            // two list slots per node
    geq::DevBuf<uint32_t> rev_cnt, rev_fill;   // reverse-CSR scratch of game_from_device (in-degrees; edge ranks)
    geq::DevBuf<uint32_t> rev_tmp;
    geq::DevBuf<uint16_t> arena;
    // generic confs (list kernels): update words per edge, the word arena (its
    // first GEN_ZERO_WORDS words hold the zero energy), the per-wave scratch
    geq::DevBuf<uint32_t> gwt, garena, gscr;
    uint32_t garena_cap = 0;          // words
    uint32_t scr_words = 0;           // scratch words per wave
    void* graph_scr = nullptr;        // scratch pointer and size baked into graph_exec
    uint32_t graph_scr_words = 0;
    geq::DevBuf<geq::Ctl> ctl;
    geq::DevBuf<unsigned long long> stats;
    // pinned, mapped and coherent: the control block (read back after each
    // batch of rounds), then at byte 64 the number the last readback kernel
    // wrote (the host spins on it); d_ctl_host is the device view
    geq::Ctl* h_ctl = nullptr;
    geq::Ctl* d_ctl_host = nullptr;
    uint32_t sig_seq = 0;             // the last number asked for
    geq::DevBuf<uint32_t> scan_tmp;   // block totals of the download's offset scan
    uint32_t* h_stage = nullptr;      // mapped pinned staging of the energy download (host view)
    uint32_t* d_stage = nullptr;      // its device view
    size_t h_stage_n = 0;             // its words
    uint32_t arena_cap = 0;
    uint32_t retag_round = 0;         // round of the last k_retag (or of the initialisation)
    std::vector<uint8_t> h_atk;       // kept for standard_reach / reach validation
    // Positions renumbered at upload (reverse breadth-first order from the
    // defender dead ends, per game): internal id = inv[caller's id]; empty when
    // the ids are the caller's.  inv_map is the same map on the device.
    std::vector<uint32_t> inv;
    geq::DevBuf<uint32_t> inv_map;
    struct Tune {                     // launch shape, set by the autotuner or defaults
        uint32_t mode = 1;            // 1: hipGraph of captured rounds; 3: one persistent launch, a block per game
                                      // (0, the cooperative launch, is switched off)
        uint32_t eval_blocks_per_cu = 8;     // capped at the blocks a CU holds at once
        uint32_t expand_blocks_per_cu = 0;   // 0: the expansion runs inside k_eval
        uint32_t rounds_per_launch = 8;
        uint32_t waves_per_block = 8;     // k_eval / k_expand workgroup: 1..8 waves
        // worklist order: 0 append order, 1 id order (k_collect after each
        // k_eval), 2 id order for games of at least ORDERED_MIN_V positions
        uint32_t ordered = 2;
    } tune;
    std::vector<uint32_t> game_base;      // batched games: first node of every game (+ total)
    // solve mode 3: game bounds, queue order (largest first), seeds grouped by
    // game, per-game state; pg_base / pg_order are uploaded once
    geq::DevBuf<uint32_t> pg_base, pg_order, pg_seed, pg_soff, pg_cnt, pg_next, pg_scan;
    geq::DevBuf<geq::GameCtl> pg_ctl;
    bool pg_ready = false;
    bool pg_std_grouped = false;          // pg_seed / pg_soff hold std_seed grouped by game
    hipGraphExec_t graph_exec = nullptr;  // rounds_per_launch captured rounds
    uint64_t solves = 0;                  // solves finished (the first launches its rounds directly)
    hipEvent_t ev_start = nullptr, ev_end = nullptr;   // around each solve's rounds
    hipStream_t cap_stream = nullptr;                  // graph capture only
    void* graph_arena = nullptr;          // arena pointer baked into graph_exec
    ~geq_game();
};


// A game built on the device (geq_gamebuild): positions, set pool, CSR.
struct geq_built {
    geq_ctx* ctx = nullptr;
    uint32_t V = 0, E = 0, n_start = 0;
    uint64_t spool_top = 0;
    geq::DevBuf<uint8_t> kind, atk;
    geq::DevBuf<uint32_t> p, q_off, q_len, x_off, x_len, hash, spool, row, col, wt;
    geq_game* game = nullptr;      // created on the first solve
    uint32_t n_reach = 0;
    // max_positions stopped the build: the positions from `frontier` on were
    // never expanded (their rows are empty although they may have successors)
    bool truncated = false;
    uint32_t frontier = 0;
    // the largest position, edge and set-pool capacities the build asked for
    // (above the final counts: a level's candidates include its duplicates);
    // the next build of the context starts at these sizes
    uint64_t peak_nodes = 0, peak_edges = 0, peak_words = 0;
    ~geq_built() {
        if (game) geq::release_built_game(game);
    }
};
