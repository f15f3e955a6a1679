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

constexpr int WAVES = 4;            // waves per block of the eval/expand kernels
constexpr int BLOCK = 64 * WAVES;

struct Ctl {                        // small device control block, read back once per round
    uint32_t wl_n[2];               // worklist sizes (ping-pong)
    uint32_t overflow;              // energy arena exhausted during k_eval
    uint32_t arena_top;             // bump pointer of the energy arena (u16 units)
};

struct EvalArgs {
    const uint32_t* wl;
    uint32_t n;
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
    unsigned long long* stats;      // node evals, edge updates, bytes, changed
};

struct ExpandArgs {
    const uint32_t* list;           // worklist of the round (or the seed list)
    uint32_t n;
    const uint8_t* chg;             // nullptr: every entry counts as changed
    uint4* cur;
    const uint4* nxt;               // nullptr: no commit (seeding)
    const uint32_t* rrow;
    const uint32_t* rcol;
    uint32_t* stamp;
    uint32_t stamp_val;
    uint32_t* out;
    Ctl* ctl;
    uint32_t out_slot;              // which ctl->wl_n counter receives the appends
};

__global__ void k_eval(EvalArgs a);
__global__ void k_expand(ExpandArgs a);
__global__ void k_list_lens(const uint4* cur, uint32_t n, uint32_t* lens);
__global__ void k_gather(const uint4* cur, const uint16_t* arena, uint32_t n, const uint32_t* offs, uint32_t* out);

}  // namespace geq
