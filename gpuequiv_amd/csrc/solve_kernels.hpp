// Kernels of the energy-game fixpoint (one round = k_eval + k_expand).
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
// k_eval's dynamic LDS per wave: staged records + updates, run offsets and first edges, three 4096-bit sets
constexpr uint32_t EVAL_LDS_PER_WAVE = STAGE * 16 + STAGE * 4 + 2 * 64 * 4 + 3 * 128 * 4;

struct Ctl {                        // small device control block, read back between round batches
    uint32_t wl_n[2];               // worklist sizes (ping-pong by round parity)
    uint32_t overflow;              // energy arena exhausted during k_eval: that round + 1 (0: none)
    uint32_t arena_top;             // bump pointer of the energy arena (u16 units)
    uint32_t round;                 // current round; its worklist is wl[round & 1]
    uint32_t rounds_done;           // rounds that evaluated at least one node
    uint32_t error;                 // k_solve: a grid barrier timed out
    uint32_t pad0;
    uint32_t bar_count, bar_gen;    // k_solve grid barrier (own 64-byte line)
    uint32_t pad1[14];
};

// Kernels read the round's worklist and size from Ctl, so a round needs no
// host-side state and can be captured once in a hipGraph and replayed.
struct EvalArgs {
    const uint32_t* wl[2];
    const uint32_t* row;
    const uint2* edges;             // (column index, packed update) per edge
    const uint8_t* atk;
    const uint4* cur;               // NodeList per node: state at the start of the round
    uint4* nxt;                     // NodeList per node: new state of changed nodes
    uint16_t* arena;
    Ctl* ctl;
    uint32_t arena_cap;
    uint8_t* chg;                   // per worklist entry: node changed
    uint32_t maxv;                  // EnergyConf::max (<= 3)
    unsigned long long* stats;      // per block: node evals, edge updates, bytes, changed
};

struct StatAcc {                    // per-lane statistics, flushed once per kernel
    uint64_t evals = 0, edges = 0, bytes = 0, changed = 0;
};

struct ExpandArgs {
    uint32_t* wl[2];
    const uint32_t* seed;           // seeding mode: these nodes count as changed, no commit
    uint32_t n_seed;
    const uint8_t* chg;
    uint4* cur;
    const uint4* nxt;
    const uint32_t* rrow;
    const uint32_t* rcol;
    uint32_t* stamp;
    Ctl* ctl;
};

__global__ void k_eval(EvalArgs a, uint32_t k);
__global__ void k_expand(ExpandArgs a, uint32_t k);
__global__ void k_advance(Ctl* ctl, uint32_t rounds);
__global__ void k_solve(EvalArgs ea, ExpandArgs xa, uint32_t max_rounds, uint32_t spin_limit);
__global__ void k_reduce_stats(const unsigned long long* slots, uint32_t nslots, unsigned long long* out);
__global__ void k_list_lens(const uint4* cur, uint32_t n, uint32_t* lens);
__global__ void k_gather(const uint4* cur, const uint16_t* arena, uint32_t n, const uint32_t* offs, uint32_t* out);

}  // namespace geq
