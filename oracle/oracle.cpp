// CPU oracle for the gpuequiv spectroscopy path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file; it is used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, and only as the checker / reported CPU baseline.
//
// It is a literal, sequential restatement of the reference crate's algorithms
// (Gobbel2000/gpuequiv v1.0.0, Rust + WGSL).  Every function cites the file:line
// of the reference it follows.  It deliberately uses the reference's own
// list-based algorithms (pairwise suprema + O(n^2) Pareto filter, chunked the
// way the WGSL kernels chunk them) so that it is independent of the GPU path,
// which uses 4096-bit up-closure bitmaps instead.
//
// Parity pins (see tests/test_oracle.py):
//   * energy/update packing: src/energy.rs:805-937 known answers;
//   * inv_update, attacker/defender steps, the round loop: all five games of
//     tests/energygame.rs and the kernel fixtures of src/energygame/tests.rs;
//   * gamebuild: tests/complete.rs:24-75 (exact position order + adjacency);
//   * verdicts / bisimulation: tests/complete.rs:140-201, src/lib.rs doctests;
//   * an independent Kleene least-fixpoint over up-sets (numpy, tests/) on
//     random games, brute force over all 4096 points for the set operations.

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// EnergyConf  (src/energy.rs:36-78)
// ---------------------------------------------------------------------------
static uint32_t next_pow2(uint32_t x) {
    if (x <= 1) return 1;
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
static uint32_t bitlen(uint32_t x) { return x == 0 ? 0 : 32u - (uint32_t)__builtin_clz(x); }

struct Conf {
    uint32_t elements, max;
    uint32_t energy_bits() const { return next_pow2(bitlen(max)); }                 // :56-59
    uint32_t energy_size() const { return (energy_bits() * elements + 31) / 32; }    // :49-53
    uint32_t energy_mask() const { return 0xffffffffu >> (32 - energy_bits()); }     // :61-63
    uint32_t update_bits() const { return next_pow2(bitlen(elements + 2)); }         // :70-73
    uint32_t update_size() const { return (update_bits() * elements + 31) / 32; }    // :65-68
    uint32_t update_mask() const { return 0xffffffffu >> (32 - update_bits()); }     // :75-77
};

// Unpack `n` fields of `bits` width (src/energy.rs:164-179, shadergen.rs:87-101)
static void unpack_fields(const uint32_t* w, uint32_t n, uint32_t bits, uint32_t mask, uint32_t* out) {
    uint32_t shift = 0, word = 0;
    for (uint32_t i = 0; i < n; i++) {
        out[i] = (w[word] >> shift) & mask;
        shift += bits;
        if (shift >= 32) { shift = 0; word++; }
    }
}

// Pack fields, advancing to the next word when the next field would not fit
// (src/energy.rs:143-162 / :543-565).
static void pack_fields(const uint32_t* v, uint32_t n, uint32_t bits, uint32_t nwords, uint32_t* out) {
    for (uint32_t i = 0; i < nwords; i++) out[i] = 0;
    uint32_t shift = 0, word = 0;
    for (uint32_t i = 0; i < n; i++) {
        out[word] |= v[i] << shift;
        shift += bits;
        if (shift + bits > 32) { shift = 0; word++; }
    }
}

constexpr int MAXW = 32;     // max energy words handled by the oracle
constexpr int MAXEL = 1024;  // max elements

struct Energy {
    uint32_t w[MAXW];
};

struct Ctx {
    Conf c;
    uint32_t esz, usz, ebits, umask, emask, ubits;
    explicit Ctx(Conf conf) : c(conf) {
        esz = c.energy_size(); usz = c.update_size();
        ebits = c.energy_bits(); emask = c.energy_mask();
        ubits = c.update_bits(); umask = c.update_mask();
    }
    // less_eq: component-wise <= on masked packed fields (shadergen.rs:55-73)
    bool le(const uint32_t* a, const uint32_t* b) const {
        uint32_t shift = 0, word = 0;
        for (uint32_t i = 0; i < c.elements; i++) {
            uint32_t m = emask << shift;
            if ((a[word] & m) > (b[word] & m)) return false;
            shift += ebits;
            if (shift >= 32) { shift = 0; word++; }
        }
        return true;
    }
    // energy_eq: word-wise equality (shadergen.rs:75-85)
    bool eq(const uint32_t* a, const uint32_t* b) const {
        for (uint32_t i = 0; i < esz; i++) if (a[i] != b[i]) return false;
        return true;
    }
    // inv_update (shadergen.rs:103-128 as instantiated in attack_update.wgsl:45-61):
    // first all 1-updates e[i] = min(e[i] + (u[i]==1), max) with u32 wrap-around,
    // then, sequentially for i = 0..elements, Min-updates e[u[i]-2] = max(e[u[i]-2], e[i]).
    void inv_update(const uint32_t* e, const uint32_t* u, uint32_t* out) const {
        uint32_t en[MAXEL], up[MAXEL];
        unpack_fields(e, c.elements, ebits, emask, en);
        unpack_fields(u, c.elements, ubits, umask, up);
        for (uint32_t i = 0; i < c.elements; i++) {
            uint32_t s = en[i] + (up[i] == 1u ? 1u : 0u);
            en[i] = std::min(s, c.max);
        }
        for (uint32_t i = 0; i < c.elements; i++) {
            uint32_t upd = up[i];
            if (upd > 1u) {
                upd -= 2u;
                if (upd < c.elements) en[upd] = std::max(en[upd], en[i]);
            }
        }
        // PACK_ENERGY (shadergen.rs:130-149): fields are or-ed in without masking
        uint32_t i = 0;
        for (uint32_t word = 0; word < esz; word++) {
            out[word] = 0;
            for (uint32_t shift = 0; shift < 32; shift += ebits) {
                out[word] |= en[i] << shift;
                i++;
                if (i >= c.elements) break;
            }
        }
    }
    // component-wise max of two packed energies (shadergen.rs:151-157)
    void sup(const uint32_t* a, const uint32_t* b, uint32_t* out) const {
        uint32_t ea[MAXEL], eb[MAXEL];
        unpack_fields(a, c.elements, ebits, emask, ea);
        unpack_fields(b, c.elements, ebits, emask, eb);
        for (uint32_t k = 0; k < c.elements; k++) ea[k] = std::max(ea[k], eb[k]);
        uint32_t i = 0;
        for (uint32_t word = 0; word < esz; word++) {
            out[word] = 0;
            for (uint32_t shift = 0; shift < 32; shift += ebits) {
                out[word] |= ea[i] << shift;
                i++;
                if (i >= c.elements) break;
            }
        }
    }
};

using List = std::vector<uint32_t>;   // flattened energies, n * esz words

// ---------------------------------------------------------------------------
// Attacker step: update + minimize
// (attack_update.wgsl:77-113, attack_minimize.wgsl:67-122,
//  collect_data src/energygame.rs:1258-1314, process_results :1372-1426)
// ---------------------------------------------------------------------------
struct AtkResult { bool changed; List list; };

static AtkResult attack_eval(const Ctx& X, uint32_t u, const uint32_t* row, const uint32_t* col,
                             const uint32_t* wt, const std::vector<List>& E) {
    const uint32_t esz = X.esz, usz = X.usz;
    List items;
    // gathered energies: each successor's list with its edge weight applied,
    // then the node's own list with the all-zero update.
    for (uint32_t k = row[u]; k < row[u + 1]; k++) {
        const List& s = E[col[k]];
        size_t n = s.size() / esz;
        for (size_t j = 0; j < n; j++) {
            uint32_t o[MAXW];
            X.inv_update(&s[j * esz], &wt[(size_t)k * usz], o);
            items.insert(items.end(), o, o + esz);
        }
    }
    size_t own_start = items.size() / esz;
    {
        uint32_t zero[MAXW] = {0};
        const List& s = E[u];
        size_t n = s.size() / esz;
        for (size_t j = 0; j < n; j++) {
            uint32_t o[MAXW];
            X.inv_update(&s[j * esz], zero, o);
            items.insert(items.end(), o, o + esz);
        }
    }
    size_t n = items.size() / esz;
    AtkResult r{false, {}};
    std::vector<char> keep(n, 1);
    for (size_t i = 0; i < n; i++) {
        const uint32_t* ei = &items[i * esz];
        for (size_t j = 0; j < n; j++) {
            if (j == i) continue;
            const uint32_t* ej = &items[j * esz];
            bool e = X.eq(ej, ei);
            if ((e && i < j) || (!e && X.le(ej, ei))) { keep[i] = 0; break; }
        }
        if (keep[i] && i < own_start) r.changed = true;
    }
    if (r.changed) {
        for (size_t i = 0; i < n; i++)
            if (keep[i]) r.list.insert(r.list.end(), &items[i * esz], &items[i * esz] + esz);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Defender step (defend_update.wgsl:73-101 + defend_iterative.wgsl:182-267)
// Simulates the single-workgroup fold sequentially, with the same 64-wide
// chunking, element order and memory accounting as the kernel.
// ---------------------------------------------------------------------------
struct SupMem {
    std::vector<uint32_t> m;   // suprema region, esz words per energy
    uint32_t esz;
    uint32_t* at(size_t i) { return &m[i * esz]; }
};

// minimize(sbase, extra_shift, size) of defend_iterative.wgsl:123-180
static uint32_t def_minimize(const Ctx& X, SupMem& S, size_t sbase, size_t extra_shift, size_t size) {
    const uint32_t esz = X.esz;
    size_t n_minimized = 0;
    size_t nbase = sbase + extra_shift;
    for (size_t chunk = 0; chunk < size; chunk += 64) {
        uint32_t energy[64][MAXW];
        int is_min[64];
        for (size_t l = 0; l < 64; l++) {
            size_t i = chunk + l;
            size_t extra_i = i + extra_shift;
            is_min[l] = 0;
            if (i < size) {
                std::memcpy(energy[l], S.at(nbase + i), esz * 4);
                is_min[l] = 1;
                for (size_t j = 0; j < n_minimized; j++) {
                    const uint32_t* e2 = S.at(sbase + j);
                    bool e = X.eq(e2, energy[l]);
                    if (j != extra_i && ((e && extra_i > j) || (!e && X.le(e2, energy[l])))) is_min[l] = 0;
                }
                if (is_min[l]) {
                    for (size_t j = chunk; j < size; j++) {
                        const uint32_t* e2 = S.at(nbase + j);
                        bool e = X.eq(e2, energy[l]);
                        if (j != i && ((e && i > j) || (!e && X.le(e2, energy[l])))) is_min[l] = 0;
                    }
                }
            }
        }
        // exclusive prefix sum + stable compaction (prefix_sum :81-113)
        size_t pre = 0;
        for (size_t l = 0; l < 64; l++) {
            if (is_min[l]) {
                std::memcpy(S.at(sbase + n_minimized + pre), energy[l], esz * 4);
                pre++;
            }
        }
        n_minimized += pre;
    }
    return (uint32_t)n_minimized;
}

struct DefResult { int32_t status; List list; };

static DefResult defend_eval(const Ctx& X, uint32_t d, const uint32_t* row, const uint32_t* col,
                             const uint32_t* wt, const std::vector<List>& E, uint64_t mem) {
    const uint32_t esz = X.esz, usz = X.usz;
    const uint32_t deg = row[d + 1] - row[d];
    // gathered + inverse-updated successor energies (defend_update.wgsl:73-101)
    std::vector<List> succ(deg);
    for (uint32_t k = 0; k < deg; k++) {
        const List& s = E[col[row[d] + k]];
        size_t n = s.size() / esz;
        for (size_t j = 0; j < n; j++) {
            uint32_t o[MAXW];
            X.inv_update(&s[j * esz], &wt[(size_t)(row[d] + k) * usz], o);
            succ[k].insert(succ[k].end(), o, o + esz);
        }
    }
    DefResult r{0, {}};
    const uint64_t available = mem;
    size_t slen = succ[0].size() / esz;
    if (slen > available) { r.status = -(int32_t)slen; return r; }
    // working region, grown on demand; the kernel's bound is `available`
    SupMem S{std::vector<uint32_t>((slen + 64) * esz, 0), esz};
    auto ensure = [&](size_t n) { if (S.m.size() < n * esz) S.m.resize(n * esz, 0); };
    std::memcpy(S.at(0), succ[0].data(), slen * esz * 4);
    for (uint32_t k = 1; k < deg; k++) {
        size_t w = succ[k].size() / esz;
        size_t comb = slen * w;
        size_t n_minimized = 0;
        for (size_t chunk = 0; chunk < comb; chunk += 64) {
            uint64_t required = slen + n_minimized + std::min<size_t>(comb - chunk, 64);
            if (required > available) { r.status = -(int32_t)required; return r; }
            ensure(slen + n_minimized + 64);
            for (size_t l = 0; l < 64; l++) {
                size_t i = chunk + l;
                if (i < comb) {
                    size_t prev_pick = i % slen, next_pick = i / slen;
                    uint32_t o[MAXW];
                    X.sup(S.at(prev_pick), &succ[k][next_pick * esz], o);
                    std::memcpy(S.at(slen + n_minimized + l), o, esz * 4);
                }
            }
            if (chunk + 64 >= comb) {
                size_t size = n_minimized + comb - chunk;
                slen = def_minimize(X, S, 0, slen, size);
            } else {
                n_minimized = def_minimize(X, S, slen, 0, n_minimized + 64);
            }
        }
    }
    r.status = (int32_t)slen;
    r.list.assign(S.m.begin(), S.m.begin() + slen * esz);
    return r;
}

// Order-insensitive comparison of src/energy.rs:334-369 as used at
// src/energygame.rs:929: equal iff same row count and every new row is in old.
static bool lists_equal_ref(const Ctx& X, const List& a_new, const List& b_old) {
    if (a_new.size() != b_old.size()) return false;
    const uint32_t esz = X.esz;
    size_t n = b_old.size() / esz;
    for (size_t i = 0; i < n; i++) {
        bool found = false;
        for (size_t j = 0; j < n && !found; j++) found = X.eq(&a_new[i * esz], &b_old[j * esz]);
        if (!found) return false;
    }
    return true;
}

// Threads for the evaluations of one round (orc_set_threads).  Rounds are
// Jacobi, so a round's evaluations are independent and any split over threads
// gives the same result; the merge stays sequential, in worklist order.
static int g_threads = 1;
// ORACLE_TRACE=1: one stderr line per round of the fixpoint loop (worklist sizes)
static const bool g_trace = std::getenv("ORACLE_TRACE") && std::atoi(std::getenv("ORACLE_TRACE")) != 0;

struct SolveStats {
    uint64_t rounds = 0, node_evals = 0, atk_evals = 0, def_evals = 0, edge_updates = 0, bytes = 0,
             def_retries = 0;
};

// The host fixpoint loop: GPURunner::execute_gpu (src/energygame.rs:1481-1523),
// initialize_visit_lists / changed_nodes (:1461-1479), the defender skip rule
// (:789-792), retry with max(128, next_pow2) memory (:919-923).
static void solve(const Ctx& X, uint32_t V, const uint32_t* row, const uint32_t* col, const uint32_t* wt,
                  const uint8_t* atk, const uint32_t* rrow, const uint32_t* rcol, std::vector<List>& E,
                  const uint32_t* reach, uint32_t n_reach, int emulate_mem, SolveStats& st) {
    std::set<uint32_t> atk_visit;
    std::map<uint32_t, uint64_t> def_visit;
    const uint64_t DEFAULT_MEM = 64;
    const uint64_t UNBOUNDED = ~0ull;
    auto changed_nodes = [&](const std::vector<uint32_t>& nodes) {
        for (uint32_t v : nodes)
            for (uint32_t k = rrow[v]; k < rrow[v + 1]; k++) {
                uint32_t w = rcol[k];
                if (atk[w]) atk_visit.insert(w);
                else def_visit.emplace(w, emulate_mem ? DEFAULT_MEM : UNBOUNDED);
            }
    };
    changed_nodes(std::vector<uint32_t>(reach, reach + n_reach));
    const uint32_t esz = X.esz;
    while (true) {
        std::vector<uint32_t> atk_jobs(atk_visit.begin(), atk_visit.end());
        atk_visit.clear();
        std::vector<std::pair<uint32_t, uint64_t>> def_jobs;
        for (auto& kv : def_visit) {
            uint32_t node = kv.first;
            bool skip = false;
            for (uint32_t k = row[node]; k < row[node + 1]; k++)
                if (E[col[k]].empty()) { skip = true; break; }
            if (!skip) def_jobs.push_back(kv);
        }
        def_visit.clear();
        // Jacobi: all evaluations read the state from the start of the round.
        std::vector<AtkResult> ar(atk_jobs.size());
        const long na = (long)atk_jobs.size(), nd = (long)def_jobs.size();
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
        for (long i = 0; i < na; i++) ar[i] = attack_eval(X, atk_jobs[i], row, col, wt, E);
        std::vector<DefResult> dr(def_jobs.size());
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
        for (long i = 0; i < nd; i++)
            dr[i] = defend_eval(X, def_jobs[i].first, row, col, wt, E, def_jobs[i].second);
        // accounting: edge updates and algorithmic bytes (attempted evaluations)
        auto account = [&](uint32_t u, const List& old_list, const List* new_list) {
            uint64_t b = 9;
            for (uint32_t k = row[u]; k < row[u + 1]; k++)
                b += 4 + 4ull * X.usz + 4ull * E[col[k]].size();
            b += 4ull * old_list.size() + (new_list ? 4ull * new_list->size() : 0);
            st.bytes += b;
            st.edge_updates += row[u + 1] - row[u];
            st.node_evals++;
        };
        for (size_t i = 0; i < atk_jobs.size(); i++) {
            account(atk_jobs[i], E[atk_jobs[i]], ar[i].changed ? &ar[i].list : nullptr);
            st.atk_evals++;
        }
        // the new list counts only when it is written (the node changed)
        std::vector<char> def_changed(def_jobs.size(), 0);
        for (size_t i = 0; i < def_jobs.size(); i++) {
            def_changed[i] = dr[i].status >= 0 && !lists_equal_ref(X, dr[i].list, E[def_jobs[i].first]);
            account(def_jobs[i].first, E[def_jobs[i].first], def_changed[i] ? &dr[i].list : nullptr);
            st.def_evals++;
        }
        st.rounds++;
        if (g_trace)
            std::fprintf(stderr, "oracle: round %llu: %zu attacker, %zu defender evaluations\n",
                         (unsigned long long)st.rounds, atk_jobs.size(), def_jobs.size());
        // process attacker results, then defender results (:1510-1513)
        std::vector<uint32_t> changed;
        for (size_t i = 0; i < atk_jobs.size(); i++)
            if (ar[i].changed) { E[atk_jobs[i]] = std::move(ar[i].list); changed.push_back(atk_jobs[i]); }
        changed_nodes(changed);
        changed.clear();
        for (size_t i = 0; i < def_jobs.size(); i++) {
            uint32_t node = def_jobs[i].first;
            if (dr[i].status < 0) {
                uint64_t need = (uint64_t)(-(int64_t)dr[i].status);
                def_visit[node] = std::max<uint64_t>(128, next_pow2((uint32_t)need));
                st.def_retries++;
                continue;
            }
            if (def_changed[i]) {
                E[node] = std::move(dr[i].list);
                changed.push_back(node);
            }
        }
        changed_nodes(changed);
        (void)esz;
        if (atk_visit.empty() && def_visit.empty()) break;
    }
}

// ---------------------------------------------------------------------------
// LTS + gamebuild (src/lib.rs, src/gamebuild.rs, src/gamebuild/gamepos.rs)
// ---------------------------------------------------------------------------
struct Trans { uint32_t process; int32_t label; };
struct Lts { std::vector<std::vector<Trans>> adj; };

static void sort_labels(Lts& l) {   // src/lib.rs:489-493 (stable)
    for (auto& r : l.adj)
        std::stable_sort(r.begin(), r.end(), [](const Trans& a, const Trans& b) { return a.label < b.label; });
}

enum Kind : uint8_t { ATTACK = 0, CLAUSE = 1, DEFEND = 2 };
struct Pos {
    uint8_t kind;
    uint32_t p;
    uint32_t q1;                // clause partner
    std::vector<uint32_t> q, qx;
    bool operator==(const Pos& o) const {
        return kind == o.kind && p == o.p && q1 == o.q1 && q == o.q && qx == o.qx;
    }
};
struct PosHash {
    size_t operator()(const Pos& x) const {
        uint64_t h = 1469598103934665603ull ^ (x.kind * 0x9e3779b97f4a7c15ull);
        auto mix = [&](uint64_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
        mix(x.p); mix(x.q1);
        mix(x.q.size()); for (auto v : x.q) mix(v);
        mix(0xffffffffu); mix(x.qx.size()); for (auto v : x.qx) mix(v);
        return (size_t)h;
    }
};

static Pos mk_attack(uint32_t p, std::vector<uint32_t> q) { Pos r{ATTACK, p, 0, std::move(q), {}}; return r; }
static Pos mk_clause(uint32_t p, uint32_t q) { Pos r{CLAUSE, p, q, {}, {}}; return r; }
static Pos mk_defend(uint32_t p, std::vector<uint32_t> q, std::vector<uint32_t> qx) {
    Pos r{DEFEND, p, 0, std::move(q), std::move(qx)}; return r;
}

// compare_enabled (src/gamebuild.rs:23-98)
static void compare_enabled(const Lts& L, uint32_t p, uint32_t q, bool& subset, bool& superset) {
    subset = true; superset = true;
    if (p == q) return;
    const auto& pl = L.adj[p];
    const auto& ql = L.adj[q];
    size_t pp = 0, qp = 0, pm = pl.size(), qm = ql.size();
    while (true) {
        if (!(subset || superset)) break;
        if (pp == pm && qp == qm) break;
        if (pp == pm) { subset = false; break; }
        if (qp == qm) { superset = false; break; }
        int32_t pa = pl[pp].label, qa = ql[qp].label;
        bool ap = false, aq = false;
        if (pa == qa) { ap = aq = true; }
        else if (pa < qa) { superset = false; ap = true; }
        else { subset = false; aq = true; }
        if (ap) { pp++; while (pp < pm && pl[pp].label == pa) pp++; }
        if (aq) { qp++; while (qp < qm && ql[qp].label == qa) qp++; }
    }
}

// Update words for the STANDARD conf (4-bit fields): observation [-1],
// challenge [0,-1], clause [Min(4)] / [Min(5),0,0,0,0,-1], answer [0,0,0,Min(3)],
// revival [Min(3)] (gamepos.rs:18-24, :139-150, :180-186)
constexpr uint32_t W_OBS = 0x1, W_CHAL = 0x10, W_CL1 = 0x5, W_CL2 = 0x100006, W_ANS = 0x4000, W_REV = 0x4;

static void successors(const Lts& L, const Pos& x, std::vector<Pos>& out, std::vector<uint32_t>& w) {
    out.clear(); w.clear();
    if (x.kind == ATTACK) {                           // gamepos.rs:17-106
        if (x.q.empty()) {
            out.push_back(mk_defend(x.p, {}, {}));
            w.push_back(W_CHAL);
            return;
        }
        for (const Trans& t : L.adj[x.p]) {
            std::vector<uint32_t> q2;                 // set_transition :111-120
            for (uint32_t q : x.q)
                for (const Trans& u : L.adj[q])
                    if (u.label == t.label) q2.push_back(u.process);
            std::sort(q2.begin(), q2.end());
            q2.erase(std::unique(q2.begin(), q2.end()), q2.end());
            if (std::find(q2.begin(), q2.end(), t.process) == q2.end()) {
                out.push_back(mk_attack(t.process, std::move(q2)));
                w.push_back(W_OBS);
            }
        }
        size_t cstart = out.size();
        out.push_back(mk_defend(x.p, x.q, {}));
        std::vector<uint32_t> q_sub, qx_sub, q_sup, qx_sup, q_eq, qx_eq;
        for (uint32_t q : x.q) {
            bool sub, sup;
            compare_enabled(L, x.p, q, sub, sup);
            (sub ? qx_sub : q_sub).push_back(q);
            (sup ? qx_sup : q_sup).push_back(q);
            ((sub && sup) ? qx_eq : q_eq).push_back(q);
        }
        auto contained = [&](const Pos& y) {
            for (size_t i = cstart + 1; i < out.size(); i++) if (out[i] == y) return true;
            return false;
        };
        if (!qx_sub.empty() && !q_sub.empty()) out.push_back(mk_defend(x.p, q_sub, qx_sub));
        if (!qx_sup.empty() && !q_sup.empty()) {
            Pos y = mk_defend(x.p, q_sup, qx_sup);
            if (!contained(y)) out.push_back(y);
        }
        if (!qx_eq.empty() && !q_eq.empty()) {
            Pos y = mk_defend(x.p, q_eq, qx_eq);
            if (!contained(y)) out.push_back(y);
        }
        while (w.size() < out.size()) w.push_back(W_CHAL);
    } else if (x.kind == CLAUSE) {                    // gamepos.rs:139-158
        out.push_back(mk_attack(x.p, {x.q1})); w.push_back(W_CL1);
        out.push_back(mk_attack(x.q1, {x.p})); w.push_back(W_CL2);
    } else {                                          // gamepos.rs:179-198
        for (uint32_t q : x.q) { out.push_back(mk_clause(x.p, q)); w.push_back(W_ANS); }
        if (!x.qx.empty()) { out.push_back(mk_attack(x.p, x.qx)); w.push_back(W_REV); }
    }
}

struct Build {
    std::vector<Pos> nodes;
    std::unordered_map<Pos, uint32_t, PosHash> map;
    std::vector<uint32_t> row, col, wt;
    std::vector<uint8_t> atk;
    uint32_t n_start = 0;
    uint32_t new_node(const Pos& x) {                 // gamebuild.rs:255-265
        uint32_t idx = (uint32_t)nodes.size();
        atk.push_back(x.kind != DEFEND);
        nodes.push_back(x);
        map[x] = idx;
        return idx;
    }
    // max_positions (0: none) is this build's cap, not the reference's: the
    // walk stops after the breadth-first level (the nodes created by the
    // previous level) that leaves more than max_positions nodes; the nodes
    // created by that level keep empty rows (truncated = true).
    bool truncated = false;
    void build_internal(const Lts& L, uint64_t max_positions = 0) {   // gamebuild.rs:229-239, :242-252
        n_start = (uint32_t)nodes.size();
        std::vector<Pos> succ;
        std::vector<uint32_t> w;
        size_t level_end = nodes.size();
        for (size_t idx = 0; idx < nodes.size(); idx++) {
            if (idx == level_end) {
                if (max_positions && nodes.size() > max_positions) {
                    truncated = true;
                    for (; idx < nodes.size(); idx++) row.push_back((uint32_t)col.size());
                    break;
                }
                level_end = nodes.size();
            }
            Pos cur = nodes[idx];
            successors(L, cur, succ, w);
            row.push_back((uint32_t)col.size());
            for (size_t i = 0; i < succ.size(); i++) {
                auto it = map.find(succ[i]);
                uint32_t s = (it != map.end()) ? it->second : new_node(succ[i]);
                col.push_back(s);
                wt.push_back(w[i]);
            }
        }
        row.push_back((uint32_t)col.size());
    }
};

// signature refinement + minimisation (src/lib.rs:413-478)
static uint32_t signature_refinement(const Lts& L, std::vector<uint32_t>& part) {
    size_t n = L.adj.size();
    part.assign(n, 0);
    size_t prev = 0, cnt = 1;
    while (prev != cnt) {
        prev = cnt;
        cnt = 0;
        std::vector<std::vector<std::pair<int32_t, uint32_t>>> sigs(n);
        for (size_t i = 0; i < n; i++) {
            for (const Trans& t : L.adj[i]) sigs[i].push_back({t.label, part[t.process]});
            std::sort(sigs[i].begin(), sigs[i].end());
            sigs[i].erase(std::unique(sigs[i].begin(), sigs[i].end()), sigs[i].end());
        }
        std::map<std::vector<std::pair<int32_t, uint32_t>>, uint32_t> m;
        std::vector<uint32_t> np(n);
        for (size_t i = 0; i < n; i++) {
            auto it = m.find(sigs[i]);
            if (it == m.end()) it = m.emplace(sigs[i], (uint32_t)cnt++).first;
            np[i] = it->second;
        }
        part = np;
    }
    return (uint32_t)cnt;
}

static Lts bisimilar_minimize(const Lts& L, const std::vector<uint32_t>& bisim, uint32_t count) {
    Lts R;
    R.adj.assign(count, {});
    std::vector<char> rep(count, 0);
    for (size_t p = 0; p < L.adj.size(); p++) {
        uint32_t c = bisim[p];
        if (!rep[c]) {
            for (const Trans& t : L.adj[p]) R.adj[c].push_back({bisim[t.process], t.label});
            rep[c] = 1;
        }
    }
    sort_labels(R);
    return R;
}

static Lts lts_from_csr(uint32_t n, const uint32_t* row, const uint32_t* dst, const int32_t* lbl, int sort) {
    Lts L;
    L.adj.resize(n);
    for (uint32_t s = 0; s < n; s++)
        for (uint32_t k = row[s]; k < row[s + 1]; k++) L.adj[s].push_back({dst[k], lbl[k]});
    if (sort) sort_labels(L);
    return L;
}

static void make_reverse(uint32_t V, const std::vector<uint32_t>& row, const std::vector<uint32_t>& col,
                         std::vector<uint32_t>& rrow, std::vector<uint32_t>& rcol) {
    // GameGraph::make_reverse (src/energygame.rs:235-250): predecessors in
    // increasing source order.
    rrow.assign(V + 1, 0);
    for (uint32_t v = 0; v < V; v++)
        for (uint32_t k = row[v]; k < row[v + 1]; k++) rrow[col[k] + 1]++;
    for (uint32_t v = 0; v < V; v++) rrow[v + 1] += rrow[v];
    rcol.assign(rrow[V], 0);
    std::vector<uint32_t> fill(rrow.begin(), rrow.end() - 1);
    for (uint32_t v = 0; v < V; v++)
        for (uint32_t k = row[v]; k < row[v + 1]; k++) rcol[fill[col[k]]++] = v;
}

}  // namespace

// ===========================================================================
// C ABI for ctypes (tests/ and bench.py cpu_baseline only)
// ===========================================================================
extern "C" {

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

struct orc_lists {          // per-node energy lists: offsets in energies (not words)
    uint32_t n_nodes;
    uint32_t esz;
    uint32_t* offsets;      // n_nodes + 1
    uint32_t* data;         // offsets[n] * esz words
};

void orc_conf_info(uint32_t elements, uint32_t max, uint32_t* out6) {
    Conf c{elements, max};
    out6[0] = c.energy_size(); out6[1] = c.energy_bits(); out6[2] = c.energy_mask();
    out6[3] = c.update_size(); out6[4] = c.update_bits(); out6[5] = c.update_mask();
}

// values -> packed words; returns 0 or -1 when a value exceeds max / too many values
int orc_pack_energy(uint32_t elements, uint32_t max, const uint32_t* vals, uint32_t n, uint32_t* out) {
    Conf c{elements, max};
    if (n > elements) return -1;
    for (uint32_t i = 0; i < n; i++) if (vals[i] > max) return -1;
    pack_fields(vals, n, c.energy_bits(), c.energy_size(), out);
    return 0;
}

void orc_unpack_energy(uint32_t elements, uint32_t max, const uint32_t* w, uint32_t* out) {
    Conf c{elements, max};
    unpack_fields(w, elements, c.energy_bits(), c.energy_mask(), out);
}

// update ints: 0 -> Zero, -1 -> Decrement, k >= 1 -> Min(k)  (src/energy.rs:474-490)
int orc_pack_update(uint32_t elements, uint32_t max, const int32_t* upd, uint32_t n, uint32_t* out) {
    Conf c{elements, max};
    if (n > elements) return -1;
    uint32_t v[MAXEL];
    for (uint32_t i = 0; i < n; i++) {
        if (upd[i] < -1) return -1;
        if (upd[i] == 0) v[i] = 0;
        else if (upd[i] == -1) v[i] = 1;
        else { if ((uint32_t)upd[i] > elements) return -1; v[i] = (uint32_t)upd[i] + 1; }
    }
    pack_fields(v, n, c.update_bits(), c.update_size(), out);
    return 0;
}

void orc_inv_update(uint32_t elements, uint32_t max, const uint32_t* e, const uint32_t* u, uint32_t* out) {
    Ctx X(Conf{elements, max});
    X.inv_update(e, u, out);
}

// Defender fold on explicit, already-gathered successor lists (the kernel-level
// fixture form of src/energygame/tests.rs): deg lists given by offsets into
// `energies` (energy units). Weights are applied first (defend_update.wgsl),
// with node 0's adjacency weights `wt` (deg * usz words).
int orc_defend_fold(uint32_t elements, uint32_t max, uint32_t deg, const uint32_t* succ_off,
                    const uint32_t* energies, const uint32_t* wt, uint64_t mem, uint32_t* out, uint32_t out_cap,
                    int32_t* status) {
    Ctx X(Conf{elements, max});
    std::vector<uint32_t> row = {0, deg}, col(deg);
    std::vector<List> E(deg + 1);
    for (uint32_t k = 0; k < deg; k++) {
        col[k] = k + 1;
        E[k + 1].assign(energies + (size_t)succ_off[k] * X.esz, energies + (size_t)succ_off[k + 1] * X.esz);
    }
    DefResult r = defend_eval(X, 0, row.data(), col.data(), wt, E, mem);
    *status = r.status;
    if (r.status < 0) return 0;
    if ((uint32_t)r.status > out_cap) return -1;
    std::memcpy(out, r.list.data(), r.list.size() * 4);
    return 0;
}

// Solve an energy game. Initial lists given by init_off (V+1, energy units) +
// init_data.  Returns per-node lists in the reference's element order.
orc_lists* orc_solve(uint32_t elements, uint32_t max, uint32_t V, const uint32_t* row, const uint32_t* col,
                     const uint32_t* wt, const uint8_t* atk, const uint32_t* init_off, const uint32_t* init_data,
                     const uint32_t* reach, uint32_t n_reach, int emulate_mem, uint64_t* stats7) {
    Conf c{elements, max};
    Ctx X(c);
    if (X.esz > (uint32_t)MAXW || elements > (uint32_t)MAXEL) return nullptr;
    std::vector<uint32_t> rowv(row, row + V + 1), colv(col, col + row[V]), rrow, rcol;
    make_reverse(V, rowv, colv, rrow, rcol);
    std::vector<List> E(V);
    for (uint32_t v = 0; v < V; v++)
        E[v].assign(init_data + (size_t)init_off[v] * X.esz, init_data + (size_t)init_off[v + 1] * X.esz);
    SolveStats st;
    solve(X, V, row, col, wt, atk, rrow.data(), rcol.data(), E, reach, n_reach, emulate_mem, st);
    orc_lists* r = (orc_lists*)std::malloc(sizeof(orc_lists));
    r->n_nodes = V; r->esz = X.esz;
    r->offsets = (uint32_t*)std::malloc(sizeof(uint32_t) * (V + 1));
    size_t tot = 0;
    for (uint32_t v = 0; v < V; v++) { r->offsets[v] = (uint32_t)(tot / X.esz); tot += E[v].size(); }
    r->offsets[V] = (uint32_t)(tot / X.esz);
    r->data = (uint32_t*)std::malloc(sizeof(uint32_t) * (tot ? tot : 1));
    size_t o = 0;
    for (uint32_t v = 0; v < V; v++) { if (!E[v].empty()) std::memcpy(r->data + o, E[v].data(), E[v].size() * 4); o += E[v].size(); }
    if (stats7) {
        stats7[0] = st.rounds; stats7[1] = st.node_evals; stats7[2] = st.atk_evals; stats7[3] = st.def_evals;
        stats7[4] = st.edge_updates; stats7[5] = st.bytes; stats7[6] = st.def_retries;
    }
    return r;
}

void orc_lists_free(orc_lists* r) {
    if (!r) return;
    std::free(r->offsets); std::free(r->data); std::free(r);
}

struct orc_game {
    uint32_t V, E, n_start;
    uint32_t *row, *col, *wt;
    uint8_t* atk;
    // positions: kind, p, q1 per node; q / qx sets as offsets into a pool
    uint8_t* kind;
    uint32_t *p, *q1, *q_off, *qx_off, *pool;
};

// GameBuild seeded with an explicit list of (p, q) attack positions
// (GameBuild::compare :165-170 for one pair; compare_multiple / compare_all
// :176-215 with the caller choosing the seed order).
orc_game* orc_gamebuild_capped(uint32_t n_states, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl,
                               uint32_t n_pairs, const uint32_t* pairs, uint64_t max_positions, int* truncated) {
    Lts L = lts_from_csr(n_states, lrow, ldst, llbl, 1);
    Build B;
    for (uint32_t i = 0; i < n_pairs; i++) B.new_node(mk_attack(pairs[2 * i], {pairs[2 * i + 1]}));
    B.build_internal(L, max_positions);
    if (truncated) *truncated = B.truncated ? 1 : 0;
    orc_game* g = (orc_game*)std::calloc(1, sizeof(orc_game));
    uint32_t V = (uint32_t)B.nodes.size(), E = (uint32_t)B.col.size();
    g->V = V; g->E = E; g->n_start = B.n_start;
    auto dup32 = [](const std::vector<uint32_t>& v) {
        uint32_t* p = (uint32_t*)std::malloc(4 * (v.size() ? v.size() : 1));
        if (!v.empty()) std::memcpy(p, v.data(), 4 * v.size());
        return p;
    };
    g->row = dup32(B.row); g->col = dup32(B.col); g->wt = dup32(B.wt);
    g->atk = (uint8_t*)std::malloc(V ? V : 1);
    if (V) std::memcpy(g->atk, B.atk.data(), V);
    g->kind = (uint8_t*)std::malloc(V ? V : 1);
    std::vector<uint32_t> p(V), q1(V), qo(V + 1), xo(V + 1), pool;
    for (uint32_t i = 0; i < V; i++) {
        const Pos& x = B.nodes[i];
        g->kind[i] = x.kind; p[i] = x.p; q1[i] = x.q1;
        qo[i] = (uint32_t)pool.size(); pool.insert(pool.end(), x.q.begin(), x.q.end());
        xo[i] = (uint32_t)pool.size(); pool.insert(pool.end(), x.qx.begin(), x.qx.end());
    }
    qo[V] = xo[V] = (uint32_t)pool.size();
    g->p = dup32(p); g->q1 = dup32(q1); g->q_off = dup32(qo); g->qx_off = dup32(xo); g->pool = dup32(pool);
    return g;
}

orc_game* orc_gamebuild(uint32_t n_states, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl,
                        uint32_t n_pairs, const uint32_t* pairs) {
    return orc_gamebuild_capped(n_states, lrow, ldst, llbl, n_pairs, pairs, 0, nullptr);
}

void orc_game_free(orc_game* g) {
    if (!g) return;
    std::free(g->row); std::free(g->col); std::free(g->wt); std::free(g->atk); std::free(g->kind);
    std::free(g->p); std::free(g->q1); std::free(g->q_off); std::free(g->qx_off); std::free(g->pool);
    std::free(g);
}

// signature refinement: partition (n) + count
uint32_t orc_signature_refinement(uint32_t n, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl,
                                  uint32_t* part_out) {
    Lts L = lts_from_csr(n, lrow, ldst, llbl, 1);
    std::vector<uint32_t> part;
    uint32_t cnt = signature_refinement(L, part);
    std::memcpy(part_out, part.data(), 4 * n);
    return cnt;
}

// bisimilar_minimize: returns the reduced LTS as CSR (label-sorted rows).
// out_row must hold count+1 entries; out_dst/out_lbl hold at most row[n] entries.
uint32_t orc_bisimilar_minimize(uint32_t n, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl,
                                uint32_t* part_out, uint32_t* out_row, uint32_t* out_dst, int32_t* out_lbl) {
    Lts L = lts_from_csr(n, lrow, ldst, llbl, 1);
    std::vector<uint32_t> part;
    uint32_t cnt = signature_refinement(L, part);
    std::memcpy(part_out, part.data(), 4 * n);
    Lts R = bisimilar_minimize(L, part, cnt);
    uint32_t k = 0;
    for (uint32_t s = 0; s < cnt; s++) {
        out_row[s] = k;
        for (const Trans& t : R.adj[s]) { out_dst[k] = t.process; out_lbl[k] = t.label; k++; }
    }
    out_row[cnt] = k;
    return cnt;
}

}  // extern "C"
