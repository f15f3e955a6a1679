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
    uint32_t pad;
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
};

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

// FUSED: k_eval also appends the predecessors of its changed nodes (no k_expand).
template <bool FUSED>
__global__ void k_eval(EvalArgs a, uint32_t k);
__global__ void k_expand(ExpandArgs a, uint32_t k);
__global__ void k_advance(Ctl* ctl, uint32_t rounds);
__global__ void k_retag(uint4* st, uint32_t n, uint32_t t);
__global__ void k_node_hdr(const uint32_t* row, const uint32_t* rrow, const uint8_t* atk, uint32_t n, uint4* st);
__global__ void k_reduce_stats(const unsigned long long* slots, uint32_t nslots, unsigned long long* out);
__global__ void k_list_lens(const uint4* st, uint32_t n, uint32_t t, uint32_t* lens);
__global__ void k_gather(const uint4* st, const uint16_t* arena, uint32_t n, uint32_t t, const uint32_t* offs, uint32_t* out);

}  // namespace geq
