/*
 * gpuequiv on MI355X (gfx950): C ABI of the energy-game solver.
 *
 * This is the drop-in boundary for the spectroscopy path of the Rust crate
 * Gobbel2000/gpuequiv v1.0.0.  It replaces the crate's wgpu device / buffer /
 * pipeline layer (src/energygame/gpuutils.rs, the private GPURunner,
 * AttackShader and DefendIterShader of src/energygame.rs and the four WGSL
 * kernels).  A Rust GPURunner replacement binds these functions through an
 * `extern "C"` block with the same signatures and #[repr(C)] structs.
 *
 * Conventions
 *  - All arrays are plain host pointers + sizes; the caller keeps ownership of
 *    inputs.  Outputs allocated by the library are released with the matching
 *    geq_*_free function.
 *  - Every function returns a geq_status; on failure geq_last_error() returns
 *    a thread-local message.  No function aborts or throws across the ABI.
 *  - Energies and updates use the crate's bit packing (src/energy.rs:36-78,
 *    :143-162, :543-565).  Energies are returned sorted ascending per node as
 *    packed words (the crate's EnergyArray compares order-insensitively,
 *    src/energy.rs:334-369, so this is its canonical form).
 *  - Energy configurations: every EnergyConf whose energies take at most 32
 *    words (EnergyConf::energy_size).  elements <= 6 and max <= 3 (which
 *    includes EnergyConf::STANDARD, the conf gamebuild uses,
 *    src/gamebuild.rs:152) run on bitmap kernels; the others, among them
 *    SILENT_STEP and ONE_DIMENSION (src/energy.rs:45-46), on list kernels in
 *    the crate's own packing.  Energies of several words come back in
 *    ascending lexicographic order of their words (word 0 first).  More than
 *    32 words, or elements == 0, returns GEQ_ERR_UNSUPPORTED.
 */
#ifndef GPUEQUIV_H
#define GPUEQUIV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  1..3 mirror gpuequiv::Error::{NoAdapter, NoDevice, BufferMap}
 * (src/error.rs:3-8); the others replace panics of the reference on invalid
 * input or resource exhaustion. */
typedef enum {
    GEQ_OK = 0,
    GEQ_ERR_NO_ADAPTER = 1,    /* no GPU / invalid device ordinal          */
    GEQ_ERR_NO_DEVICE = 2,     /* device present but could not be set up   */
    GEQ_ERR_TRANSFER = 3,      /* host<->device copy failed (BufferMap)    */
    GEQ_ERR_INVALID_ARG = 4,   /* malformed graph / energies / updates     */
    GEQ_ERR_OUT_OF_MEMORY = 5, /* device allocation or budget exhausted    */
    GEQ_ERR_UNSUPPORTED = 6,   /* energy configuration of more than 32 words */
    GEQ_ERR_INTERNAL = 7
} geq_status;

typedef struct geq_ctx geq_ctx;
typedef struct geq_game geq_game;

/* EnergyConf (src/energy.rs:36-47). */
typedef struct {
    uint32_t elements;
    uint32_t max;
} geq_conf;

/* GameGraph public fields (src/energygame.rs:120-138), borrowed for the call.
 * weights holds n_edges * update_size packed words.  rev_* may be NULL, in
 * which case the reverse CSR is derived (GameGraph::make_reverse, :235-250). */
typedef struct {
    geq_conf conf;
    uint32_t n_vertices;
    uint32_t n_edges;
    const uint32_t* row_offsets;      /* n_vertices + 1 */
    const uint32_t* column_indices;   /* n_edges        */
    const uint32_t* weights;          /* n_edges * update_size */
    const uint8_t* attacker_pos;      /* n_vertices, 0 / 1 */
    const uint32_t* rev_row_offsets;  /* n_vertices + 1 or NULL */
    const uint32_t* rev_column_indices; /* n_edges or NULL */
} geq_graph;

/* Per-node energy lists (EnergyGame::energies, src/energygame.rs:329-338):
 * node v owns data[offsets[v] * energy_size .. offsets[v+1] * energy_size). */
typedef struct {
    uint32_t n_nodes;
    uint32_t energy_size;
    uint32_t* offsets;   /* n_nodes + 1, in energies */
    uint32_t* data;      /* packed words */
} geq_energies;

/* Counters of one solve.  A node evaluation is one attacker update or one
 * defender combination in some round; an edge update is one edge relaxed by
 * such an evaluation.  alg_bytes follows the per-evaluation byte count
 * 9 + sum_v (4 + 4*update_size + 4*|E(v)|) + 4*(|E_old| + |E_new| if written). */
typedef struct {
    uint64_t rounds;
    uint64_t node_evals;
    uint64_t edge_updates;
    uint64_t alg_bytes;
    uint64_t changed;
    double solve_ms;        /* device time of the fixpoint loop (hipEvents)        */
    double eval_kernel_ms;  /* sum of k_eval durations, when profiling was asked   */
    double expand_kernel_ms;/* sum of k_expand durations, when profiling was asked */
    uint64_t eval_launches;
    double solve_kernel_ms; /* reserved (0) */
    uint32_t mode;          /* 1 hipGraph rounds, 2 profiled per-kernel launches, 3 one persistent launch (a block per game) */
    uint32_t pad;
} geq_stats;

/* ---- context (GPUCommon::get_gpu, src/energygame/gpuutils.rs:9-61) ---- */
geq_status geq_ctx_create(int device, geq_ctx** out);
void geq_ctx_destroy(geq_ctx* ctx);
const char* geq_last_error(void);
/* HIP device name / CU count of the context's device. */
geq_status geq_ctx_info(geq_ctx* ctx, char* name, size_t name_len, int* n_cu);
/* Free and total device memory of the context's device, in bytes (for sizing
 * a serving budget).  No counterpart in the reference. */
geq_status geq_ctx_memory(geq_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- one-shot solve: EnergyGame::run (src/energygame.rs:379-383) ----
 * to_reach: the nodes initialised with the zero energy (EnergyGame::with_reach,
 * :355-365; standard_reach :347-352 picks the defender dead ends).
 * initial: optional explicit starting energies (NULL = with_reach semantics).
 * The result is written to *out (free with geq_energies_free). */
geq_status geq_solve(geq_ctx* ctx, const geq_graph* graph, const uint32_t* to_reach, uint32_t n_reach,
                     const geq_energies* initial, geq_energies* out, geq_stats* stats);

void geq_energies_free(geq_energies* e);

/* geq_solve without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every device
 * buffer of the solve lives in one static device array of 64 MiB per device
 * (held from library load to process end); games that do not fit it (about
 * 350,000 positions at 4 edges each) return GEQ_ERR_OUT_OF_MEMORY, as does an
 * energy arena that runs out.  Energy configurations: elements <= 6 and
 * max <= 3 (others GEQ_ERR_UNSUPPORTED).  Inputs as geq_solve (no initial
 * energies: EnergyGame::with_reach); the caller passes the output arrays:
 * out_offsets (n_vertices + 1 entries, in energies) and out_data (out_cap
 * packed energies, one word each, ascending per node).  *out_n (may be NULL)
 * receives the result's energy count; a result larger than out_cap returns
 * GEQ_ERR_INVALID_ARG with *out_n set.  Calls on one device share its array:
 * the library serialises them with a per-device lock. */
geq_status geq_solve_noalloc(geq_ctx* ctx, const geq_graph* graph, const uint32_t* to_reach, uint32_t n_reach,
                             uint32_t* out_offsets, uint32_t* out_data, uint64_t out_cap, uint64_t* out_n, geq_stats* stats);

/* ---- device-resident game (the pieces of geq_solve, for repeated solves) ----
 * geq_game_upload copies the graph to HBM; geq_game_solve resets the energies
 * (to_reach / initial as above) and runs the fixpoint with every input already
 * resident; geq_game_energies downloads the result. */
geq_status geq_game_upload(geq_ctx* ctx, const geq_graph* graph, geq_game** out);
geq_status geq_game_solve(geq_game* game, const uint32_t* to_reach, uint32_t n_reach, const geq_energies* initial,
                          int profile_kernels, geq_stats* stats);
/* The same with EnergyGame::standard_reach (src/energygame.rs:347-352): the
 * defender positions without successors start at {0}.  They are found on the
 * device at the first such call and kept with the game, so repeated solves
 * upload nothing. */
geq_status geq_game_solve_standard(geq_game* game, int profile_kernels, geq_stats* stats);
geq_status geq_game_energies(geq_game* game, geq_energies* out);
void geq_game_free(geq_game* game);

/* ---- batched games (many independent games, one device-resident solve) ----
 * The games (same EnergyConf) are laid out back to back in HBM; one shared
 * device worklist holds the positions of every game, so each round evaluates
 * the frontier of all games together.  Games cannot influence each other and
 * every round is evaluated per node against the previous round's state, so
 * each game's result is exactly its standalone result.  There is no
 * counterpart in the reference, which solves one game per EnergyGame::run. */
geq_status geq_batch_upload(geq_ctx* ctx, uint32_t n_games, const geq_graph* graphs, geq_game** out);
/* reach lists per game (node ids local to the game); to_reach and n_reach both
 * NULL: every game's standard_reach (defender positions without successors,
 * EnergyGame::standard_reach src/energygame.rs:347-352), found on the device
 * once and kept with the batch (also for geq_batch_autotune) */
geq_status geq_batch_solve(geq_game* batch, const uint32_t* const* to_reach, const uint32_t* n_reach, int profile_kernels,
                           geq_stats* stats);
/* outs: n_games results, each freed with geq_energies_free */
geq_status geq_batch_energies(geq_game* batch, geq_energies* outs);

/* ---- launch parameters and the autotuner ----
 * The fixpoint's launch shape: workgroup size of k_eval / k_expand (waves per
 * block), their grid sizes (blocks per CU), the rounds captured per replayed
 * hipGraph, and the energy-arena capacity (u16 entries; it only grows).
 * geq_*_autotune times candidate shapes on the game itself (median device time
 * of `repeats` solves each, coordinate search from the current shape with a
 * separate k_expand and from the same shape with the expansion inside k_eval,
 * then the persistent solve mode 3 with 4 and 8 waves per block for the bitmap
 * layout), keeps the fastest (geq_game_get_solve_mode tells the mode) and
 * sizes the arena for the largest solve seen.  Every
 * shape computes the same result.  No counterpart in the reference (fixed
 * WGSL workgroup sizes). */
typedef struct {
    uint32_t waves_per_block;       /* 1..8 (default 8) */
    uint32_t eval_blocks_per_cu;    /* 1..64 (default 8; capped at the blocks a CU holds at once) */
    uint32_t expand_blocks_per_cu;  /* 0..64 (default 0: no k_expand, k_eval queues the predecessors itself) */
    uint32_t rounds_per_launch;     /* 1..256 (default 8) */
    uint64_t arena_entries;
} geq_tune_params;
typedef struct {
    geq_tune_params best;
    double default_ms, best_ms;     /* median device time per solve: starting shape, chosen shape */
    uint32_t candidates, repeats;
    uint64_t arena_peak;            /* arena entries the solves used */
} geq_tune_report;
/* How geq_game_solve / geq_batch_solve run the rounds.  1 (default): batches
 * of rounds captured in a hipGraph and replayed, every round over the whole
 * batch's worklist.  3: one persistent launch; each block takes whole games
 * from a device queue (largest first) and runs all of a game's rounds with
 * block barriers in between, so no block waits for another.  Both give the
 * same energies, node evaluations and edge updates (a game's rounds inside
 * the disjoint union are its rounds alone); the round count is the largest
 * game's.  Mode 3 covers the bitmap energy layout (elements <= 6, max <= 3)
 * and returns GEQ_ERR_UNSUPPORTED for other confs.  GEQ_SOLVE_MODE sets the
 * mode of new games.  No counterpart in the reference (one dispatch per
 * shader pass, src/energygame.rs:1481-1523). */
geq_status geq_game_set_solve_mode(geq_game* game, uint32_t mode);
geq_status geq_game_get_solve_mode(geq_game* game, uint32_t* mode);
geq_status geq_game_get_tune(geq_game* game, geq_tune_params* out);
geq_status geq_game_set_tune(geq_game* game, const geq_tune_params* params);
geq_status geq_game_autotune(geq_game* game, const uint32_t* to_reach, uint32_t n_reach, uint32_t repeats,
                             geq_tune_report* out);
geq_status geq_batch_autotune(geq_game* batch, const uint32_t* const* to_reach, const uint32_t* n_reach, uint32_t repeats,
                              geq_tune_report* out);

/* ---- gamebuild: LTS -> spectroscopy energy game, built on the device ----
 * Replaces GameBuild::{compare, compare_multiple, compare_all} +
 * build_internal (src/gamebuild.rs:165-265) and the position successor rules
 * (src/gamebuild/gamepos.rs:17-198).  The game is seeded with the attack
 * positions (p, {q}) of `pairs` (ids 0..n_pairs-1, in the given order) and
 * grown breadth-first; positions get the ids the reference's FIFO walk gives
 * them.  The result stays in HBM: geq_built_solve runs the fixpoint on it
 * (EnergyGame::standard_reach + run), geq_built_download copies graph and
 * positions to the host. */
typedef struct {
    uint32_t n_states;
    uint32_t n_transitions;
    const uint32_t* row_offsets;   /* n_states + 1; each state's transitions sorted by label (src/lib.rs:489-493) */
    const uint32_t* targets;       /* n_transitions */
    const int32_t* labels;         /* n_transitions */
} geq_lts;

typedef struct {
    uint64_t positions;
    uint64_t edges;
    uint64_t levels;
    uint64_t set_words;            /* words of Q / Q* set data written */
    double build_ms;               /* wall time of the device build */
    uint32_t truncated;            /* stopped after max_positions: the last level has no successors */
    uint32_t pad;
} geq_build_stats;

/* Host copy of a built game.  Position i: kind[i] (0 attack, 1 clause,
 * 2 defend), process p[i], Q = pool[q_off[i] .. +q_len[i]) (the clause's q for
 * a clause), Q* = pool[x_off[i] .. +x_len[i]) (defend only).  weights use the
 * STANDARD packing (6 x 4 bits). */
typedef struct {
    uint32_t n_positions, n_edges, n_start;
    uint32_t* row_offsets;
    uint32_t* column_indices;
    uint32_t* weights;
    uint8_t* attacker_pos;
    uint8_t* kind;
    uint32_t *p, *q_off, *q_len, *x_off, *x_len;
    uint32_t* pool;
    uint64_t pool_len;
} geq_built_game;

typedef struct geq_built geq_built;

/* max_positions: 0 = unbounded; otherwise the build stops after the level that
 * crosses it (stats->truncated = 1). */
geq_status geq_gamebuild(geq_ctx* ctx, const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs, uint64_t max_positions,
                         geq_built** out, geq_build_stats* stats);
geq_status geq_built_download(geq_built* built, geq_built_game* out);
void geq_built_game_free(geq_built_game* g);
/* standard_reach solve of the built game on the device.  A build stopped by
 * max_positions (geq_build_stats.truncated) is refused with
 * GEQ_ERR_INVALID_ARG: its last level's positions were never expanded, and
 * solving it as a complete game would count them as defender dead ends the
 * attacker wins at zero energy. */
geq_status geq_built_solve(geq_built* built, int profile_kernels, geq_stats* stats);
/* The solve of a capped build as a game of its own: the unexpanded frontier
 * positions are left out of the positions to reach (the attacker cannot win
 * through them), every other rule as geq_built_solve.  Each budget set is then
 * contained in the complete game's up-closure, so these are NOT the reference's
 * answers; for a complete build it equals geq_built_solve.  No counterpart in
 * the reference, which always builds the whole game. */
geq_status geq_built_solve_capped(geq_built* built, int profile_kernels, geq_stats* stats);
/* truncated: 1 if max_positions stopped the build; frontier: the first
 * unexpanded position (the position count when complete) */
geq_status geq_built_frontier(geq_built* built, uint32_t* truncated, uint32_t* frontier);
/* energies of positions [first, first + count) after geq_built_solve */
geq_status geq_built_energies(geq_built* built, uint32_t first, uint32_t count, geq_energies* out);
/* Replaces Equivalence::preorder's test (src/equivalence.rs:125-132,
 * EnergyArray::test_equivalence src/energy.rs:307-318) for positions
 * [first, first + count) after geq_built_solve, on the device: out[i * n_energies
 * + k] = 1 when no budget of position first + i lies below energies[k] in every
 * component (the preorder holds for that starting point), else 0.  energies:
 * n_energies (at most 64) energies packed in the game's EnergyConf.  Only the
 * verdict bytes cross to the host.  GEQ_ERR_UNSUPPORTED for confs of more than
 * 6 elements or max above 3 (download the energies instead). */
geq_status geq_built_preorders(geq_built* built, uint32_t first, uint32_t count, const uint32_t* energies,
                               uint32_t n_energies, uint8_t* out);
void geq_built_free(geq_built* built);

/* ---- serving: continuous batching of spectroscopy jobs ----
 * Jobs (an LTS and the (p, q) pairs to compare; the game of
 * TransitionSystem::compare_unminimized seeded with every pair) arrive at any
 * time.  Builder threads build the queued jobs' games together (one level
 * loop over the union of their LTSs, on the builder's own stream); at the
 * next round boundary the game is placed into shared HBM pools (admission:
 * its positions and edges must fit the pools' free ranges, i.e. the HBM
 * budget) and its frontier joins the shared device worklist, so resident jobs
 * are solved together, round by round.  A job retires as soon as none of its
 * positions is on the worklist; its result is the same as when solved alone.
 * No counterpart in the reference (one game per compare / equivalences call). */
typedef struct geq_server geq_server;
typedef struct {
    uint64_t budget_bytes;         /* HBM: 7/8 for the position / edge pools, 1/8 for the energy
                                      arena's initial size (default 4 GiB) */
    double edges_per_position;     /* pool shape (default 4) */
    uint32_t max_jobs;             /* resident jobs at most (default 4096) */
    uint32_t rounds_per_step;      /* rounds per geq_server_step (default 4) */
    uint64_t compact_entries;      /* after a retirement, pack the energy arena once its top passes this
                                      many u16 entries (default 0: three quarters of its size) */
    uint64_t max_positions;        /* a job whose game grows past this many positions fails with
                                      GEQ_ERR_OUT_OF_MEMORY (0: no limit but the budget) */
    uint32_t builders;             /* game-building threads, each with its own stream (default 3) */
    uint32_t pad;
} geq_server_opts;
typedef struct {
    uint32_t state;                /* 0 queued, 1 building, 2 built, 3 resident, 4 done, 5 failed */
    uint32_t status;               /* geq_status of a failed job */
    uint32_t positions, edges;
    double t_submit_ms, t_built_ms, t_admit_ms, t_done_ms;   /* since geq_server_create */
} geq_job_info;
geq_status geq_server_create(geq_ctx* ctx, const geq_server_opts* opts, geq_server** out);
void geq_server_destroy(geq_server* server);
/* copies the inputs; returns immediately */
geq_status geq_server_submit(geq_server* server, const geq_lts* lts, const uint32_t* pairs, uint32_t n_pairs,
                             uint64_t* job_id);
/* admit built jobs that fit, run rounds_per_step rounds, retire finished jobs */
geq_status geq_server_step(geq_server* server, uint32_t* n_resident, uint32_t* n_waiting);
/* state of a job; for a finished job, `out` (may be NULL) receives the budgets
 * of its starting positions (one per pair, in submission order) */
geq_status geq_server_poll(geq_server* server, uint64_t job_id, geq_job_info* info, geq_energies* out);
/* ids of jobs that finished (done or failed) since the last call, oldest first;
 * at most `cap` are taken */
geq_status geq_server_finished(geq_server* server, uint64_t* ids, uint32_t cap, uint32_t* n);
/* forget a finished job */
geq_status geq_server_release(geq_server* server, uint64_t job_id);
/* counters of the serving loop */
typedef struct {
    uint64_t steps, rounds, compactions;      /* steps that ran rounds, rounds, energy-arena compactions */
    uint64_t node_slots, node_slots_used;     /* position pool: size, slots held by resident jobs */
    uint64_t edge_slots, edge_slots_used;     /* edge pool: size, slots held by resident jobs */
    uint64_t arena_entries, arena_top;        /* energy arena: size, entries in use (u16) */
    uint64_t jobs_resident, jobs_waiting, jobs_done;
    uint64_t builds;                          /* merged builds (jobs built together) so far */
    uint64_t split_builds;                    /* merged builds that failed and were rebuilt one job at a time */
    uint64_t deferred_admissions;             /* admissions sent back to wait: no device memory for the placement */
} geq_server_stats;
geq_status geq_server_get_stats(geq_server* server, geq_server_stats* out);

/* ---- strong-bisimulation minimisation on the device ----
 * The same results as geq_signature_refinement / geq_bisimilar_minimize below
 * (TransitionSystem::{signature_refinement, bisimilar_minimize},
 * src/lib.rs:413-478), computed on the device: per refinement round a
 * segmented sort of (label, class of target) keys, a hash table of the
 * signatures and a scan numbering the classes by first occurrence.  rounds
 * (may be NULL) receives the number of refinement rounds. */
geq_status geq_signature_refinement_gpu(geq_ctx* ctx, const geq_lts* lts, uint32_t* partition, uint32_t* n_classes,
                                        uint32_t* rounds);
geq_status geq_bisimilar_minimize_gpu(geq_ctx* ctx, const geq_lts* lts, uint32_t* partition, uint32_t* n_classes,
                                      uint32_t* out_row, uint32_t* out_targets, int32_t* out_labels,
                                      uint32_t* out_n_transitions);

/* ---- helpers (no device needed) ---- */
/* TransitionSystem::signature_refinement (src/lib.rs:444-478): partition[s] is
 * the strong-bisimulation class of s, numbered by first occurrence. */
geq_status geq_signature_refinement(const geq_lts* lts, uint32_t* partition, uint32_t* n_classes);
/* TransitionSystem::bisimilar_minimize (src/lib.rs:413-432): the quotient LTS
 * (out_row: n_classes + 1 entries; out_targets / out_labels: at most
 * n_transitions entries each) and the partition used. */
geq_status geq_bisimilar_minimize(const geq_lts* lts, uint32_t* partition, uint32_t* n_classes, uint32_t* out_row,
                                  uint32_t* out_targets, int32_t* out_labels, uint32_t* out_n_transitions);
/* standard_reach: defender positions without successors (:347-352). Writes at
 * most n_vertices ids into out, returns the count through *n_out. */
geq_status geq_standard_reach(const geq_graph* graph, uint32_t* out, uint32_t* n_out);

#ifdef __cplusplus
}
#endif

#endif /* GPUEQUIV_H */
