// Device gamebuild: data layout shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "runtime.hpp"
#include "sort.hpp"

namespace geq {
namespace gb {

constexpr unsigned long long EMPTY_ENTRY = ~0ull;   // free slot of the position table

// Label-sorted LTS in CSR form plus, per state, its sorted distinct labels
// (the enabledness sets compared by compare_enabled).
struct LtsArgs {
    const uint32_t* row;
    const uint32_t* dst;
    const int32_t* lbl;
    const uint32_t* en_row;
    const int32_t* en_lbl;
    // merged builds of several jobs: the job of each state and the jobs that
    // stopped growing (NULL for a single game)
    const uint32_t* job_of_state;
    const uint8_t* job_dead;
    // sizes for a copy into LDS (k_generate): states, transitions, enabled labels
    uint32_t n, m, n_en;
    // when every label is in [0, 64): per state the bit set of its labels
    // (bit a: label a), else NULL
    const unsigned long long* en_mask;
};
// k_generate copies the LTS (row, dst, lbl, en_row, en_lbl) into LDS when it
// takes at most this many words.
constexpr uint32_t LTS_LDS_WORDS = 12 * 1024;
__host__ __device__ inline uint32_t lts_words(uint32_t n, uint32_t m, uint32_t n_en) { return 2u * (n + 1u) + 2u * m + n_en; }

// Position records (structure of arrays).  Sets Q / Q* live in the set pool:
// Attack(p, Q) = (ATTACK, p, Q, -); Clause(p, q) = (CLAUSE, p, [q], -);
// Defend(p, Q, Q*) = (DEFEND, p, Q, Q*).
struct NodeArgs {
    uint8_t* kind;
    uint32_t* p;
    uint32_t* q_off;
    uint32_t* q_len;
    uint32_t* x_off;
    uint32_t* x_len;
    uint32_t* hash;
    uint32_t* spool;
};

struct Cand {               // one successor candidate of a frontier position
    uint32_t kind_p;        // kind << 30 | p, or INVALID for an unused slot
    uint32_t q_off, q_len, x_off, x_len;
    uint32_t weight;
    uint32_t hash;
    uint32_t res;           // after k_insert: existing node id, or PENDING | table slot
    uint32_t parent;        // index in the frontier
    uint32_t slot;          // successor index within the parent
};

// Heavy frontier positions: attack positions too large for the kernel that
// generates the level (see bounds_of) are generated by one block each
// (k_generate_heavy) or one wave each (k_generate_wave).  HV_CAP bounds their |Q| and sum of member degrees
// (the block's LDS).
constexpr uint32_t HV_CAP = 8192;
// k_generate_sub: lanes per position, threads per block, and the largest |Q|
// and sum of member degrees its LDS holds; k_generate_wave (a wave per
// position): the same limits.
constexpr uint32_t SUB_LANES = 8, SUB_BLOCK = 256, SUB_QMAX = 32, SUB_TMAX = 64;
constexpr uint32_t WAVE_QMAX = 256, WAVE_TMAX = 1024;
// Which kernels generate a level: k_generate with k_generate_heavy
// (GEN_BY_THREAD), or k_generate_sub with k_generate_wave and k_generate_heavy
// for the positions past its limits (GEN_BY_SUB); GEN_BY_WAVE and GEN_BY_BLOCK send
// every attack position with a non-empty Q that fits to the wave / block
// kernel (tests).
constexpr uint32_t GEN_BY_THREAD = 0, GEN_BY_SUB = 1, GEN_BY_WAVE = 2, GEN_BY_BLOCK = 3;
// List entries of k_generate_wave's positions carry this bit (frontier
// indices stay below 2^31).
constexpr uint32_t WAVE_TAG = 0x80000000u;
struct HeavyArgs {
    uint8_t* flag;       // per frontier position: 1 if listed
    uint32_t* list;      // the listed positions' frontier indices (any order), WAVE_TAG on k_generate_wave's
    uint32_t* count;     // their number (zero before k_bounds)
    uint32_t heavy_min;  // GEN_BY_THREAD: weight from which a position is listed; any mode: 0 lists none
    uint32_t mode;       // GEN_BY_*
};

// Per frontier position: upper bounds on its candidates (x) and on the set-pool
// words its successor sets take (y); entry n_front is zero.  Also the heavy flags and list.
__global__ void k_bounds(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L, uint2* ub, HeavyArgs H);
// The same for a frontier of *n_new + add positions from `lo` whose size only
// the device knows: entries [n_front, n_slots) are zero, so the exclusive scan
// of all n_slots entries has the frontier's offsets in front and the totals
// from entry n_front on.  The heavy flags and list cover the frontier's first
// min(n_front, chunk) positions.
__global__ void k_bounds_next(NodeArgs N, uint32_t lo, const uint32_t* n_new, uint32_t add, uint32_t n_slots, uint32_t chunk,
                              LtsArgs L, uint2* ub, HeavyArgs H);
// `off`: the exclusive scan of k_bounds' pairs (candidate and set-pool offsets).
// Positions flagged in `heavy` are left to k_generate_heavy.
__global__ void k_generate(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L, const uint2* off, const uint8_t* heavy,
                           uint64_t spool_base, Cand* cands, uint32_t* nsucc);
__global__ void k_generate_heavy(NodeArgs N, uint32_t lo, LtsArgs L, const uint2* off, uint64_t spool_base, HeavyArgs H,
                                 Cand* cands, uint32_t* nsucc);
// The same candidates as k_generate, SUB_LANES lanes per position (grid of
// (n_front + 1) * SUB_LANES threads in blocks of SUB_BLOCK).
__global__ void k_generate_sub(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L, const uint2* off, const uint8_t* heavy,
                               uint64_t spool_base, Cand* cands, uint32_t* nsucc);
// The listed positions tagged WAVE_TAG, a wave each (blocks of SUB_BLOCK, any grid).
__global__ void k_generate_wave(NodeArgs N, uint32_t lo, LtsArgs L, const uint2* off, uint64_t spool_base, HeavyArgs H,
                                Cand* cands, uint32_t* nsucc);
__global__ void k_insert(Cand* cands, uint32_t n_cand, NodeArgs N, unsigned long long* table, uint32_t* keymin, uint32_t mask);
// rank[0..n_cand] := exclusive scan of the candidates' winner flags (the
// smallest candidate index of each new position), written to flag[0..n_cand)
// first.
geq_status scan_marks(const Cand* cands, uint32_t n_cand, const uint32_t* keymin, uint32_t* rank, uint32_t* flag,
                      DevBuf<uint32_t>& tmp, hipStream_t s);
__global__ void k_finalize(const Cand* cands, uint32_t n_cand, const uint32_t* rank, uint32_t v_base, NodeArgs N,
                           unsigned long long* table, uint32_t* keymin, uint32_t* clear);
__global__ void k_edges(const Cand* cands, uint32_t n_cand, const uint32_t* rowl, uint32_t e_base, const unsigned long long* table,
                        uint32_t* col, uint32_t* wt);
// k_edges, the rows of k_rows and (n_slots > 0) k_bounds_next in one launch of
// max(n_cand, nf, n_slots) threads, for builds without per-job counts; the
// heavy count was cleared by k_finalize.
__global__ void k_level_out(const Cand* cands, uint32_t n_cand, const uint32_t* rowl, uint32_t e_base, const unsigned long long* table,
                            uint32_t* col, uint32_t* wt, uint32_t nf, uint32_t lo, uint32_t* row, NodeArgs N, uint32_t hi,
                            const uint32_t* n_new, uint32_t add, uint32_t n_slots, uint32_t chunk, LtsArgs L, uint2* ub, HeavyArgs H);
// Work that follows a chunk's rows, done by k_rows' first block: the heavy
// count is cleared for the next k_bounds_next (NULL: not), and merged builds
// mark the jobs whose position counts passed the cap (dead NULL: none).
struct RowsTail {
    uint32_t* heavy_count;
    const uint32_t* cnt;
    uint32_t n_jobs;
    uint64_t cap;
    uint8_t* dead;
};
__global__ void k_rows(const uint32_t* rowl, uint32_t n_front, uint32_t lo, uint32_t e_base, uint32_t* row, RowsTail tail);
__global__ void k_starts(const uint32_t* pairs, uint32_t n, const uint8_t* insert, NodeArgs N, unsigned long long* table,
                         uint32_t mask);
__global__ void k_rehash(const uint32_t* hash, uint32_t n, const uint8_t* insert, uint32_t n_start, unsigned long long* table,
                         uint32_t mask);
__global__ void k_fill_u64(unsigned long long* a, uint64_t n, unsigned long long v);
__global__ void k_fill_u32(uint32_t* a, uint64_t n, uint32_t v);
__global__ void k_attacker_flags(const uint8_t* kind, uint32_t n, uint8_t* atk);

// Words of the pinned readback block: a chunk's totals, the heavy positions
// of the chunk bounded last, then the per-job position counts.
// RB_SEQ is written last, after a system-scope fence, with the number the
// host waits for (see rb_publish).
constexpr uint32_t RB_EDGES = 0, RB_NEW = 1, RB_CAND = 2, RB_WORDS = 3, RB_HEAVY = 4, RB_SEQ = 5;
// k_level_tail's further levels: how many, and the last one's frontier start
// and size, candidates, set-pool words, first new id, first edge and set-pool base
constexpr uint32_t RB_LOOP = 6, RB_LOOP_LO = 7, RB_LOOP_NF = 8, RB_LOOP_CAND = 9, RB_LOOP_WORDS = 10, RB_LOOP_V = 11,
                   RB_LOOP_E = 12, RB_LOOP_SP = 13;
constexpr uint32_t RB_JOBS = 14;   // then the per-job counts

// Sum of (candidates, set-pool words) bound pairs, saturating at `lim`
// (saturating addition of non-negative values is associative, so the scan
// stays exact below lim).  A level whose running total reaches lim is refused
// before any candidate is written: its offsets no longer fit 32 bits, and a
// wrapped total would size the buffers below the offsets k_generate writes at.
struct PairSum {
    uint32_t lim;
    __host__ __device__ uint2 operator()(const uint2& a, const uint2& b) const {
        const uint64_t x = (uint64_t)a.x + b.x, y = (uint64_t)a.y + b.y;
        return make_uint2(x < lim ? (uint32_t)x : lim, y < lim ? (uint32_t)y : lim);
    }
    __host__ __device__ static uint2 identity() { return make_uint2(0u, 0u); }
};

// The count words at the end of a chunk, in pinned host memory: its edges, its
// new positions and the per-job position counts cnt.  With `next` (the offsets
// k_bounds_next scanned for the next level, whose *n_new + add positions are
// taken in chunks of `chunk`), also the candidate and set-pool totals and the
// heavy positions (*heavy) of the next level's first chunk.  `seq` goes to
// out[RB_SEQ] last.
__global__ void k_chunk_counts(const uint32_t* n_edges, const uint32_t* n_new, uint32_t add, uint32_t chunk, const uint2* next,
                               const uint32_t* heavy, const uint32_t* cnt, uint32_t n_cnt, uint32_t* out, uint32_t seq);

// The last step of a kernel that fills count words for the host, called by
// the block's first wave, which has written all of them: its system-scope
// fence orders those stores before `seq` goes to out[RB_SEQ].  The host, which
// waits for that number, then reads every word (the memory is coherent).
// One wave fences: each fence writes back the L2.
__device__ __forceinline__ void rb_publish(uint32_t* out, uint32_t seq) {
    __threadfence_system();
    if (threadIdx.x == 0) __hip_atomic_store(out + RB_SEQ, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A chunk whose frontier, candidates and next-level bounds each fit TAIL_MAX
// items runs everything after k_generate(_heavy) as one block of TAIL_BLOCK
// threads, its phases separated by block barriers: the row scan, inserts,
// marks, the rank scan, finalize, edges, per-job counts, rows, dead jobs, the
// next level's bounds and their scan, and the count words.  The same per-item
// steps as the kernels of larger chunks, so the same positions, ids and edges.
constexpr uint32_t TAIL_BLOCK = 1024, TAIL_MAX = 4 * TAIL_BLOCK;
// What k_level_tail's further levels may use: the capacities (entries) of the
// frontier arrays (nsucc, rowl), candidate arrays (cands, flag, rank), bound
// arrays (ub, off, heavy flags and list), position records, edges and set
// pool; the table's slots (a level may fill half); the position cap (0: none);
// the launched level's set-pool base and words; the largest candidate count
// of a chunk the block runs (at most TAIL_MAX); and the largest level it
// generates itself (at most TAIL_GEN positions).
struct TailRoom {
    uint32_t front, cand, slots;
    uint64_t nodes, edges, words, table, max_positions;
    uint32_t spool_top, n_words;
    uint32_t tail_max, gen_max;
};
constexpr uint32_t TAIL_GEN = 512;   // positions of a level k_level_tail_loop generates itself

struct TailArgs {
    NodeArgs N;
    LtsArgs L;
    HeavyArgs H;
    Cand* cands;
    uint32_t* nsucc;
    uint32_t* rowl;
    uint32_t* flag;
    uint32_t* rank;
    unsigned long long* table;
    uint32_t* keymin;
    uint32_t* col;
    uint32_t* wt;
    uint32_t* row;
    uint2* ub;                 // last chunk: the next level's bounds ...
    uint2* off;                // ... and their scan
    uint32_t* job_cnt;         // merged builds: per-job counts (NULL otherwise)
    uint8_t* job_dead;
    uint32_t n_jobs;
    uint64_t cap;              // merged builds: the position cap per job (0: none)
    uint32_t* rb;              // pinned count words
    uint32_t lo, nf, n_cand, v_base, e_base, mask;
    uint32_t hi, add, slots, chunk, last, lim;
    uint32_t seq;              // written to rb[RB_SEQ] last
    uint32_t loop;             // further levels the block may build (0: the launched chunk only)
    TailRoom room;
};
__global__ void k_level_tail(TailArgs T);
__global__ void k_level_tail_loop(TailArgs T);

// Pinned host words that kernels write directly (the level loop's counts).
struct HostWords {
    uint32_t* h = nullptr;   // host view
    uint32_t* d = nullptr;   // device view of the same memory
    size_t n = 0;
    HostWords() = default;
    HostWords(const HostWords&) = delete;
    HostWords& operator=(const HostWords&) = delete;
    ~HostWords() { release(); }
    void release() {
        if (h) (void)hipHostFree(h);
        h = d = nullptr;
        n = 0;
    }
};

// Pinned host memory a build packs its uploads into.
struct PinnedStage {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// Device buffers of one build (level loop and job extraction), kept by a
// caller that builds repeatedly so that steady-state builds allocate nothing.
struct BuildScratch {
    DevBuf<uint32_t> nsucc, rowl, flag, rank, keymin, d_jcnt;
    // the build's uploads in one buffer: the LTS (rows, targets, labels), its
    // distinct labels per state, label bit sets, start pairs and their insert
    // flags, and for merged builds the job of each state and the job flags
    DevBuf<uint32_t> l_pack;
    PinnedStage stage;                // their pinned copy on the host
    const uint32_t* job_map = nullptr;   // merged builds: the job of each state (in l_pack)
    DevBuf<uint2> ub, off;
    DevBuf<uint8_t> hv_flag;          // heavy positions of a frontier (k_bounds)
    DevBuf<uint32_t> hv_list, hv_count;
    HostWords rb;
    uint32_t rb_seq = 0;              // the last number written to rb[RB_SEQ]
    DevBuf<uint32_t> scan32;          // block totals of the device scans
    DevBuf<uint2> scan2;
    RadixScratch sort;                // job extraction's sort
    DevBuf<Cand> cands;
    DevBuf<unsigned long long> table;
    DevBuf<uint32_t> key, key_s, val, order, dsort, escan, local, d_ps, d_es;
    SpareBufs spare;                  // buffers outgrown by earlier builds (RetireScope)
};

}  // namespace gb

// Games of several jobs built together (gamebuild_jobs): job j's positions are
// the contiguous range [pos_start[j], pos_start[j] + n_pos[j]) with local ids
// in the order its own build gives them; its rows are
// row[pos_start[j] + j .. + n_pos[j]] (n_pos[j] + 1 entries, relative to
// edge_start[j]), its columns (local ids) and updates col / wt[edge_start[j] ..).
struct BuiltJobs {
    DevBuf<uint32_t> row, col, wt;
    DevBuf<uint8_t> atk;
    std::vector<uint32_t> pos_start, n_pos, edge_start, n_edges;
    std::vector<uint8_t> dead;       // grew past the position cap
    uint32_t levels = 0;
    double build_ms = 0;
    double prep_ms = 0;       // the merged LTS on the host
    double extract_ms = 0;    // cutting the per-job games out of the merged build
};
// Builds the games of jobs (validated, each with at least one pair) in one
// level loop over the disjoint union of their LTSs.  `ws` and `store` (the
// merged positions) are reused across calls; `out`'s buffers only grow.
geq_status gamebuild_jobs(geq_ctx* ctx, const std::vector<geq_lts>& ltss, const std::vector<const uint32_t*>& pairs,
                          const std::vector<uint32_t>& n_pairs, uint64_t cap, gb::BuildScratch& ws, geq_built& store,
                          BuiltJobs* out);

}  // namespace geq
