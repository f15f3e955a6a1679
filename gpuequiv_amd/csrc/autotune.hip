// Autotuner of the solver's launch parameters: workgroup size (waves per
// block), grid sizes of k_eval / k_expand (blocks per CU; no k_expand grid:
// the expansion runs inside k_eval), rounds per replayed hipGraph, the solve
// mode (replayed rounds, or one persistent launch with a block per game) and
// the energy-arena capacity.  Candidates are timed on the game
// itself (median device time of repeated solves), searched coordinate by
// coordinate from the current parameters.  Every candidate computes the same
// fixpoint; only the launch shape changes.  No counterpart in the reference,
// whose workgroup sizes are fixed in the WGSL sources.
#include <algorithm>
#include <string>
#include <vector>

#include "runtime.hpp"

using namespace geq;

namespace {

void apply(geq_game* g, const geq_game::Tune& t) {
    g->tune = t;
    if (g->graph_exec) {   // the captured rounds bake in the launch shape
        (void)hipGraphExecDestroy(g->graph_exec);
        g->graph_exec = nullptr;
    }
}

geq_tune_params params_of(const geq_game* g) {
    geq_tune_params p{};
    p.waves_per_block = g->tune.waves_per_block;
    p.eval_blocks_per_cu = g->tune.eval_blocks_per_cu;
    p.expand_blocks_per_cu = g->tune.expand_blocks_per_cu;
    p.rounds_per_launch = g->tune.rounds_per_launch;
    p.arena_entries = g->conf.generic ? g->garena_cap : g->arena_cap;   // u16 entries, or words of a generic conf
    return p;
}

// median device time (ms) of `repeats` solves after one warm-up solve;
// reach NULL: the game's standard_reach, kept on the device
geq_status measure(geq_game* g, const std::vector<uint32_t>* reach, uint32_t repeats, double* ms, uint64_t* arena_peak) {
    std::vector<double> t;
    for (uint32_t r = 0; r <= repeats; r++) {
        geq_stats st{};
        geq_status s = reach ? geq_game_solve(g, reach->data(), (uint32_t)reach->size(), nullptr, 0, &st)
                             : solve_core(g, g->std_n, nullptr, 0, &st, g->std_seed.p);
        if (s != GEQ_OK) return s;
        *arena_peak = std::max<uint64_t>(*arena_peak, g->h_ctl->arena_top);
        if (r) t.push_back(st.solve_ms);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return GEQ_OK;
}

geq_status autotune(geq_game* g, const std::vector<uint32_t>* reach, uint32_t repeats, geq_tune_report* rep) {
    repeats = std::max(1u, std::min(repeats, 50u));
    if (!reach) {
        geq_status s = standard_seed(g);
        if (s != GEQ_OK) return s;
    }
    using Tune = geq_game::Tune;
    struct Dim {
        uint32_t Tune::*field;
        std::vector<uint32_t> values;
    };
    // the game's own round count is a candidate batch of rounds too: one
    // replay then covers the whole fixpoint with no empty rounds after it
    std::vector<uint32_t> batches = {1, 2, 4, 8, 16, 32};
    {
        geq_stats st{};
        const geq_status s = reach ? geq_game_solve(g, reach->data(), (uint32_t)reach->size(), nullptr, 0, &st)
                                   : solve_core(g, g->std_n, nullptr, 0, &st, g->std_seed.p);
        if (s != GEQ_OK) return s;
        if (st.rounds >= 1 && st.rounds <= 256 && std::find(batches.begin(), batches.end(), (uint32_t)st.rounds) == batches.end())
            batches.push_back((uint32_t)st.rounds);
    }
    const std::vector<Dim> dims = {
        {&Tune::waves_per_block, {1, 2, 4, 8}},
        {&Tune::eval_blocks_per_cu, {1, 2, 4, 8, 16, 32}},
        {&Tune::expand_blocks_per_cu, {0, 1, 2, 4, 8, 16}},   // 0: expansion inside k_eval
        {&Tune::rounds_per_launch, batches},
    };
    uint64_t peak = 0;
    uint32_t candidates = 0;
    // coordinate search from `start`; returns the best shape and its time
    auto search = [&](Tune start, Tune& best, double& best_ms, double& start_ms) -> geq_status {
        best = start;
        best.mode = 1;
        apply(g, best);
        geq_status s = measure(g, reach, repeats, &best_ms, &peak);
        if (s != GEQ_OK) return s;
        start_ms = best_ms;
        candidates++;
        for (int pass = 0; pass < 2; pass++) {
            bool improved = false;
            for (const Dim& d : dims) {
                for (uint32_t v : d.values) {
                    if (best.*d.field == v) continue;
                    Tune c = best;
                    c.*d.field = v;
                    apply(g, c);
                    double ms = 0;
                    s = measure(g, reach, repeats, &ms, &peak);
                    if (s != GEQ_OK) return s;
                    candidates++;
                    if (ms < 0.97 * best_ms) {   // keep a change only for a clear gain
                        best = c;
                        best_ms = ms;
                        improved = true;
                    }
                }
            }
            if (!improved) break;
        }
        return GEQ_OK;
    };
    // two starts: the current shape with a separate k_expand, and the same
    // shape with the expansion inside k_eval (whose best grid tends to differ)
    const Tune start = g->tune;
    Tune sep = start, fus = start;
    if (sep.expand_blocks_per_cu == 0) sep.expand_blocks_per_cu = 4;
    fus.expand_blocks_per_cu = 0;
    Tune best_a{}, best_b{};
    double ms_a = 0, ms_b = 0, start_a = 0, start_b = 0;
    geq_status s = search(sep, best_a, ms_a, start_a);
    if (s == GEQ_OK) s = search(fus, best_b, ms_b, start_b);
    const double default_ms = start.expand_blocks_per_cu == 0 ? start_b : start_a;
    if (s != GEQ_OK) {
        apply(g, start);
        return s;
    }
    Tune best = ms_b < ms_a ? best_b : best_a;
    double best_ms = std::min(ms_a, ms_b);
    // the persistent launch with a block per game (bitmap layout only): wins
    // when the games are small next to the whole batch, loses when one game
    // alone outlasts the rest on its single block
    if (!g->conf.generic)
        for (uint32_t w : {4u, 8u}) {
            Tune c = best;
            c.mode = SOLVE_MODE_GAMES;
            c.waves_per_block = w;
            apply(g, c);
            double ms = 0;
            s = measure(g, reach, repeats, &ms, &peak);
            if (s != GEQ_OK) {
                apply(g, start);
                return s;
            }
            candidates++;
            if (ms < 0.97 * best_ms) {
                best = c;
                best_ms = ms;
            }
        }
    apply(g, best);
    // antichain capacity: the arena holds the largest solve seen without regrowing
    const uint64_t want = peak + peak / 4 + 1024;
    if (want > (g->conf.generic ? g->garena_cap : g->arena_cap)) {
        s = grow_arena(g, (uint32_t)std::min<uint64_t>(want, 0xffffffffull));
        if (s != GEQ_OK) return s;
    }
    if (rep) {
        rep->best = params_of(g);
        rep->default_ms = default_ms;
        rep->best_ms = best_ms;
        rep->candidates = candidates;
        rep->repeats = repeats;
        rep->arena_peak = peak;
    }
    return GEQ_OK;
}

}  // namespace

extern "C" {

geq_status geq_game_get_tune(geq_game* g, geq_tune_params* out) {
    if (!g || !out) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    *out = params_of(g);
    return GEQ_OK;
}

geq_status geq_game_set_tune(geq_game* g, const geq_tune_params* p) {
    if (!g || !p) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    if (p->waves_per_block < 1 || p->waves_per_block > (uint32_t)MAX_WAVES || p->eval_blocks_per_cu < 1 ||
        p->eval_blocks_per_cu > 64 || p->expand_blocks_per_cu > 64 || p->rounds_per_launch < 1 || p->rounds_per_launch > 256)
        return fail(GEQ_ERR_INVALID_ARG,
                    "launch parameters out of range (waves 1..8, eval blocks per CU 1..64, expand blocks per CU 0..64, rounds 1..256)");
    (void)hipSetDevice(g->ctx->device);
    geq_game::Tune t = g->tune;
    t.waves_per_block = p->waves_per_block;
    t.eval_blocks_per_cu = p->eval_blocks_per_cu;
    t.expand_blocks_per_cu = p->expand_blocks_per_cu;
    t.rounds_per_launch = p->rounds_per_launch;
    apply(g, t);
    if (p->arena_entries > (g->conf.generic ? g->garena_cap : g->arena_cap)) {
        geq_status s = grow_arena(g, (uint32_t)std::min<uint64_t>(p->arena_entries, 0xffffffffull));
        if (s != GEQ_OK) return s;
    }
    return GEQ_OK;
}

geq_status geq_game_set_solve_mode(geq_game* g, uint32_t mode) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "game is NULL");
    if (mode != SOLVE_MODE_GRAPH && mode != SOLVE_MODE_GAMES)
        return fail(GEQ_ERR_INVALID_ARG, "solve mode: 1 (replayed round graphs) or 3 (one persistent launch, a block per game)");
    if (mode == SOLVE_MODE_GAMES && g->conf.generic)
        return fail(GEQ_ERR_UNSUPPORTED, "the per-game persistent solve covers the bitmap energy layout only");
    g->tune.mode = mode;
    return GEQ_OK;
}

geq_status geq_game_get_solve_mode(geq_game* g, uint32_t* mode) {
    if (!g || !mode) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    *mode = g->tune.mode;
    return GEQ_OK;
}

geq_status geq_game_autotune(geq_game* g, const uint32_t* to_reach, uint32_t n_reach, uint32_t repeats, geq_tune_report* out) {
    if (!g) return fail(GEQ_ERR_INVALID_ARG, "game is NULL");
    if (n_reach && !to_reach) return fail(GEQ_ERR_INVALID_ARG, "to_reach is NULL");
    (void)hipSetDevice(g->ctx->device);
    const std::vector<uint32_t> reach(to_reach, to_reach + n_reach);
    return autotune(g, &reach, repeats, out);
}

geq_status geq_batch_autotune(geq_game* b, const uint32_t* const* to_reach, const uint32_t* n_reach, uint32_t repeats,
                              geq_tune_report* out) {
    if (b && !to_reach && !n_reach) {   // every game's standard_reach
        (void)hipSetDevice(b->ctx->device);
        return autotune(b, nullptr, repeats, out);
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
    (void)hipSetDevice(b->ctx->device);
    return autotune(b, &reach, repeats, out);
}

}  // extern "C"
