// Kernels of the energy-game fixpoint (one round = k_eval + k_expand, or
// k_eval alone with the expansion folded in).
//
// Round structure follows GPURunner::execute_gpu (src/energygame.rs:1481-1523):
// every node on the round's worklist is evaluated against the state at the
// start of the round (Jacobi), then the predecessors of every changed node form
// the next worklist.  The host gather / upload / readback of the reference is
// replaced by device-resident energy lists and a device worklist.
#pragma once
#include "geq_device.hpp"

namespace geq {

constexpr int WAVES = 4;            // default waves per block of the eval/expand kernels
constexpr int BLOCK = 64 * WAVES;
constexpr int MAX_WAVES = 8;        // k_eval / k_expand take any block of 1..MAX_WAVES waves
constexpr int MAX_BLOCK = 64 * MAX_WAVES;
constexpr uint32_t STAT_SLOTS = 8192;   // max grid size of the statistics-writing kernels
constexpr uint32_t STAGE = 128;         // successor records staged per wave in k_eval
constexpr uint32_t APPEND_STAGE = 256;  // worklist appends staged per wave before a flush
// k_eval's dynamic LDS per wave: staged records + updates, run offsets and first edges, three 4096-bit sets
constexpr uint32_t EVAL_LDS_PER_WAVE = STAGE * 16 + STAGE * 4 + 2 * 64 * 4 + 3 * 128 * 4;
// ... plus the staged worklist appends when k_eval also expands
constexpr uint32_t EVAL_FUSED_LDS_PER_WAVE = EVAL_LDS_PER_WAVE + APPEND_STAGE * 4;

struct Ctl {                        // small device control block, read back between round batches
    uint32_t wl_n[3];               // worklist sizes: round r reads wl_n[r % 3] and appends to wl_n[(r + 1) % 3]
    uint32_t overflow;              // energy arena exhausted during k_eval: that round + 1 (0: none)
    uint32_t arena_top;             // bump pointer of the energy arena (u16 units)
    uint32_t round;                 // current round; its worklist is wl[round & 1]
    uint32_t last_nonempty;         // 1 + the last round whose worklist was not empty
    uint32_t scr_need;              // list kernels: scratch words per wave a node needed (0: enough)
    unsigned long long totals[4];   // k_reduce_stats after a batch of rounds: node evaluations, edge
                                    // updates, algorithmic bytes, changed nodes (read back with the block)
};
static_assert(sizeof(Ctl) == 64, "Ctl is one 64-byte block");

// Kernels read the round's worklist and size from Ctl, so a round needs no
// host-side state and can be captured once in a hipGraph and replayed.
struct EvalArgs {
    uint32_t* wl[2];
    const uint32_t* row;
    const uint2* edges;             // (column index, packed update) per edge
    const uint8_t* atk;
    uint4* st;                      // node records (header + two list slots)
    uint16_t* arena;
    Ctl* ctl;
    uint32_t arena_cap;
    uint8_t* chg;                   // per worklist entry: node changed (read by a separate k_expand)
    uint32_t maxv;                  // EnergyConf::max (<= 3)
    unsigned long long* stats;      // per block: node evals, edge updates, bytes, changed
    const uint32_t* rrow;           // reverse CSR: the expansion folded into k_eval
    const uint32_t* rcol;
    uint32_t* queued;               // two bitmaps of qwords words: node queued on the list of a round parity
    uint32_t qwords;
    uint32_t ordered;               // 1: the expansion only sets queued bits; k_collect builds the next list in id order
};

// The list a round of k_eval's evaluation phase reads, where its appends go,
// and the waves that share it: the whole grid (k_eval), or one block that runs
// one game's rounds (k_solve_games).
struct PhaseList {
    const uint32_t* wl;             // the round's worklist
    uint32_t* out;                  // the next round's worklist
    uint32_t* out_n;                // its size (global memory or LDS)
    uint32_t wave0, waves;          // this wave's index among the sharing waves, their number
    bool grid_wide;                 // the grid shares the list: it may clear the queued bitmap with plain stores
};

// Persistent batched solve (k_solve_games): one launch for a whole batch of
// games.  Each block takes games from a queue and runs all of a game's rounds
// itself, with block barriers between the rounds; no block ever waits for
// another, so the grid may be of any size.  Game g: nodes [base[g],
// base[g + 1]), its worklists in the same ranges of wl[0] and wl[1], its
// reach seeds at seed[soff[g] .. soff[g + 1]).
struct GameCtl {                    // per game, 16 bytes
    uint32_t state;                 // GAME_NEW, GAME_PAUSED (before `round`) or GAME_DONE
    uint32_t round;                 // paused: the round that runs next
    uint32_t n;                     // paused: its worklist size
    uint32_t retag;                 // the round of the game's last re-tag
};
constexpr uint32_t GAME_NEW = 0, GAME_PAUSED = 1, GAME_DONE = 2;
struct GamesArgs {
    const uint32_t* base;           // n_games + 1 entries
    const uint32_t* order;          // the queue: game indices, largest game first
    const uint32_t* seed;
    const uint32_t* soff;           // n_games + 1 entries
    GameCtl* gctl;
    uint32_t* next;                 // the queue's head (zero before the launch)
    uint32_t n_games;
    uint32_t retag_interval;        // a game re-tags its nodes every this many rounds
};
__global__ void k_solve_games(EvalArgs a, GamesArgs ga);
__global__ void k_seed_count(const uint32_t* seed, uint32_t n, const uint32_t* base, uint32_t n_games, uint32_t* cnt);
__global__ void k_seed_scatter(const uint32_t* seed, uint32_t n, const uint32_t* base, uint32_t n_games, const uint32_t* soff,
                               uint32_t* fill, uint32_t* out);

// The next round's worklist from its queued bitmap, in id order (each block a
// contiguous range of bitmap words, one atomic per block).
struct CollectArgs {
    uint32_t* wl[2];
    const uint32_t* queued;
    uint32_t qwords;
    Ctl* ctl;
};
__global__ void k_collect(CollectArgs a, uint32_t k);
constexpr uint32_t COLLECT_THREADS = 256, COLLECT_WORDS = 4;   // bitmap words per thread

struct StatAcc {                    // per-lane statistics, flushed once per kernel
    uint64_t evals = 0, edges = 0, bytes = 0, changed = 0;
};

struct ExpandArgs {
    uint32_t* wl[2];
    const uint32_t* seed;           // seeding mode: these nodes count as changed
    uint32_t n_seed;
    const uint8_t* chg;
    const uint32_t* rrow;
    const uint32_t* rcol;
    uint32_t* queued;
    uint32_t qwords;
    Ctl* ctl;
};

// Words of one "queued" bitmap for V nodes.  Bit v of bitmap (r & 1) is set
// while node v is on round r's worklist: set by the append that queues it,
// cleared by round r's evaluation of v (so the bitmap of a parity is empty
// again before the round two later starts appending to it).
__host__ __device__ inline uint32_t queue_words(uint32_t V) { return (V + 31u) / 32u; }

// Energy configurations beyond the bitmap layout (elements > 6 or max > 3):
// the crate's packing, `ebits`-wide fields (a power of two), at most
// GEN_MAX_WORDS words per energy.  E masks the even fields of a word, G holds
// the lowest bit of every odd field (the guard bits of word_le / word_max).
constexpr uint32_t GEN_MAX_WORDS = 32;
constexpr uint32_t GEN_BLOCK = 256;           // k_eval_list: 4 waves, one node per wave at a time
constexpr uint32_t GLEN_MASK = 0x7fffffffu;   // list length in a list kernel's slot (x)
struct GConf {
    uint32_t elements, maxv, ebits, emask, ubits, umask, esz, usz;
    uint32_t E, G;
};
// List slot of the list kernels (16 B, two per node like the bitmap slots):
// x = length, y = word offset into the arena, z = 0, w = round tag << 16.
struct GenArgs {
    uint32_t* wl[2];
    const uint2* edges;             // (column index, first update word) per edge
    const uint32_t* wt;             // usz update words per edge
    uint4* st;                      // node records (header + two list slots)
    uint32_t* arena;                // energies, esz words each; words [0, esz) hold the zero energy
    Ctl* ctl;
    uint32_t arena_cap;             // words
    uint8_t* chg;                   // per worklist entry: node changed (read by k_expand)
    unsigned long long* stats;
    uint32_t* queued;
    uint32_t qwords;
    uint32_t* scr;                  // per wave scr_words words of candidate scratch
    uint32_t scr_words;
    GConf c;
};
template <int ESZ>
__global__ void k_eval_list(GenArgs a, uint32_t k);
__global__ void k_list_lens_g(const uint4* st, uint32_t n, uint32_t t, uint32_t* lens, const uint32_t* map);
__global__ void k_gather_g(const uint4* st, const uint32_t* arena, uint32_t n, uint32_t t, uint32_t esz, const uint32_t* offs,
                           uint32_t* out, const uint32_t* map);

// The round of a kernel at position k of a batch, its worklist size (0 after
// an overflow) and the overflow flag, from one scalar load of the control
// block's first 32 bytes (round and counts share them).
struct RoundCtl {
    uint32_t r, n;
    bool ovf;
};
__device__ __forceinline__ RoundCtl read_round(const Ctl* ctl, uint32_t k) {
    const uint4* p = reinterpret_cast<const uint4*>(ctl);
    const uint4 w0 = p[0], w1 = p[1];   // w0: wl_n[0..2], overflow; w1: arena_top, round, last_nonempty, pad
    const uint32_t r = uniform(w1.y) + k;
    const uint32_t i = r % 3u;
    const uint32_t n = uniform(i == 0u ? w0.x : i == 1u ? w0.y : w0.z);
    const bool ovf = uniform(w0.w) != 0u;
    return RoundCtl{r, ovf ? 0u : n, ovf};
}

// Per-block statistics slot (no cross-block atomics): slot b is only ever
// written by block b of the fixed-size grid, rounds are stream-ordered.
__device__ __forceinline__ void flush_stats(const StatAcc& acc, unsigned long long* slots) {
    __shared__ unsigned long long s_red[4];
    if (threadIdx.x < 4) s_red[threadIdx.x] = 0ull;
    __syncthreads();
    if (acc.bytes) atomicAdd(&s_red[2], (unsigned long long)acc.bytes);
    if ((threadIdx.x & 63u) == 0) {
        if (acc.evals) atomicAdd(&s_red[0], (unsigned long long)acc.evals);
        if (acc.edges) atomicAdd(&s_red[1], (unsigned long long)acc.edges);
        if (acc.changed) atomicAdd(&s_red[3], (unsigned long long)acc.changed);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_red[threadIdx.x]) slots[4ull * blockIdx.x + threadIdx.x] += s_red[threadIdx.x];
}


// FUSED: k_eval also appends the predecessors of its changed nodes (no k_expand).
template <bool FUSED>
__global__ void k_eval(EvalArgs a, uint32_t k);
__global__ void k_expand(ExpandArgs a, uint32_t k);
__global__ void k_advance(Ctl* ctl, uint32_t rounds);
__global__ void k_retag(uint4* st, uint32_t n, uint32_t t);
__global__ void k_node_hdr(const uint32_t* row, const uint32_t* rrow, const uint8_t* atk, uint32_t n, uint4* st);
__global__ void k_reduce_stats(const unsigned long long* slots, uint32_t nslots, unsigned long long* out);
// Output download: entry v of the output is node map[v] (node v when map is NULL).
__global__ void k_list_lens(const uint4* st, uint32_t n, uint32_t t, uint32_t* lens, const uint32_t* map);
// The equivalences' energies of a preorder check, packed (2-bit fields).
constexpr uint32_t EQ_WORDS_MAX = 64;
struct EqWords {
    uint32_t w[EQ_WORDS_MAX];
    uint32_t n;
};
__global__ void k_preorders(const uint4* st, const uint16_t* arena, uint32_t n, uint32_t t, const uint32_t* map, EqWords eq,
                            uint32_t fields, uint8_t* out);
// out holds cap words; a list that would pass it is skipped (the caller has
// checked offs[n] <= cap)
__global__ void k_gather(const uint4* st, const uint16_t* arena, uint32_t n, uint32_t t, const uint32_t* offs, uint32_t* out,
                         uint64_t cap, const uint32_t* map);

}  // namespace geq
