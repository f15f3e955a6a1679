// gpuequiv on MI355X: C++17 host API over the C ABI of include/gpuequiv.h.
//
// Mirrors the public types of the Rust crate Gobbel2000/gpuequiv v1.0.0 with
// the same names, argument meaning and error behaviour:
//   EnergyConf, Energy, Upd, Update, EnergyArray, UpdateArray  (src/energy.rs)
//   GameGraph, EnergyGame                                      (src/energygame.rs)
//   Transition, TransitionSystem                               (src/lib.rs)
//   Position, GameBuild                                        (src/gamebuild.rs, src/gamebuild/gamepos.rs)
//   StartInfo, Equivalence, EquivalenceRelation               (src/equivalence.rs)
//   std_equivalences                                           (src/energy.rs:198-225)
// Rust's `new` constructors are spelled `make` (a C++ keyword otherwise);
// constructors that return Option return std::optional; the crate's Error
// variants are gpuequiv::Error with an ErrorKind.  The crate's async `run`,
// `compare` and `equivalences` are synchronous here.
//
// Header-only; link the program with -lgpuequiv_hip.  Every compute call goes
// to the GPU library: there is no host fallback.
#pragma once

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <optional>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <variant>
#include <vector>

#include "gpuequiv.h"

namespace gpuequiv {

// ---------------------------------------------------------------- errors (src/error.rs:3-8)
enum class ErrorKind : int {
    NoAdapter = GEQ_ERR_NO_ADAPTER,
    NoDevice = GEQ_ERR_NO_DEVICE,
    BufferMap = GEQ_ERR_TRANSFER,
    InvalidArg = GEQ_ERR_INVALID_ARG,
    OutOfMemory = GEQ_ERR_OUT_OF_MEMORY,
    Unsupported = GEQ_ERR_UNSUPPORTED,
    Internal = GEQ_ERR_INTERNAL,
};

class Error : public std::runtime_error {
public:
    Error(ErrorKind k, const std::string& msg) : std::runtime_error(msg), kind(k) {}
    ErrorKind kind;
};

namespace detail {
inline void check(geq_status st) {
    if (st != GEQ_OK) throw Error(static_cast<ErrorKind>(st), geq_last_error());
}
inline uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
inline uint32_t bit_len(uint32_t x) {
    uint32_t n = 0;
    while (x) {
        n++;
        x >>= 1;
    }
    return n;
}
// field packing that moves to the next word when a field would not fit (src/energy.rs:143-162)
inline std::vector<uint32_t> pack(const std::vector<uint32_t>& values, uint32_t bits, uint32_t words) {
    std::vector<uint32_t> out(words, 0);
    uint32_t shift = 0, w = 0;
    for (uint32_t v : values) {
        out[w] |= (uint32_t)((uint64_t)v << shift);
        shift += bits;
        if (shift + bits > 32) {
            shift = 0;
            w++;
        }
    }
    return out;
}
inline std::vector<uint32_t> unpack(const uint32_t* words, uint32_t n, uint32_t bits, uint32_t mask) {
    std::vector<uint32_t> out;
    uint32_t shift = 0, w = 0;
    for (uint32_t i = 0; i < n; i++) {
        out.push_back((words[w] >> shift) & mask);
        shift += bits;
        if (shift >= 32) {
            shift = 0;
            w++;
        }
    }
    return out;
}

// The process-wide device context (GPUCommon::get_gpu keeps one in a OnceLock).
inline geq_ctx* context() {
    struct Holder {
        geq_ctx* ctx = nullptr;
        geq_status st = GEQ_OK;
        std::string err;
        Holder() {
            st = geq_ctx_create(0, &ctx);
            if (st != GEQ_OK) err = geq_last_error();
        }
        ~Holder() {
            if (ctx) geq_ctx_destroy(ctx);
        }
    };
    static Holder h;
    if (h.st != GEQ_OK) throw Error(static_cast<ErrorKind>(h.st), h.err);
    return h.ctx;
}
}  // namespace detail

// ---------------------------------------------------------------- EnergyConf (src/energy.rs:36-78)
struct EnergyConf {
    uint32_t elements;
    uint32_t max;

    uint32_t energy_bits() const { return detail::next_pow2(detail::bit_len(max)); }
    uint32_t energy_size() const { return (energy_bits() * elements + 31) / 32; }
    uint32_t energy_mask() const { return 0xffffffffu >> (32 - energy_bits()); }
    uint32_t update_bits() const { return detail::next_pow2(detail::bit_len(elements + 2)); }
    uint32_t update_size() const { return (update_bits() * elements + 31) / 32; }
    uint32_t update_mask() const { return 0xffffffffu >> (32 - update_bits()); }
    bool operator==(const EnergyConf& o) const { return elements == o.elements && max == o.max; }
    bool operator!=(const EnergyConf& o) const { return !(*this == o); }

    static const EnergyConf STANDARD;
    static const EnergyConf SILENT_STEP;
    static const EnergyConf ONE_DIMENSION;
};
inline const EnergyConf EnergyConf::STANDARD{6, 3};
inline const EnergyConf EnergyConf::SILENT_STEP{8, 2};
inline const EnergyConf EnergyConf::ONE_DIMENSION{1, 0xffffffffu};

// ---------------------------------------------------------------- Energy (src/energy.rs:81-141)
class Energy {
public:
    Energy(std::vector<uint32_t> data, EnergyConf conf) : data_(std::move(data)), conf_(conf) {}
    // Energy::new: None when too many values or a value above conf.max
    static std::optional<Energy> make(const std::vector<uint32_t>& values, EnergyConf conf) {
        if (values.size() > conf.elements) return std::nullopt;
        for (uint32_t v : values)
            if (v > conf.max) return std::nullopt;
        return Energy(detail::pack(values, conf.energy_bits(), conf.energy_size()), conf);
    }
    static Energy from_raw_data(const std::vector<uint32_t>& data, EnergyConf conf) { return Energy(data, conf); }
    static Energy zero(EnergyConf conf) { return Energy(std::vector<uint32_t>(conf.energy_size(), 0), conf); }
    const std::vector<uint32_t>& raw_data() const { return data_; }
    EnergyConf get_conf() const { return conf_; }
    std::vector<uint32_t> to_vec() const {
        return detail::unpack(data_.data(), conf_.elements, conf_.energy_bits(), conf_.energy_mask());
    }
    // PartialOrd (:121-141): -1 less, 0 equal, 1 greater, nullopt incomparable
    std::optional<int> partial_cmp(const Energy& o) const {
        if (conf_.elements != o.conf_.elements) throw Error(ErrorKind::InvalidArg, "Incompatible Energy configurations");
        const auto a = to_vec(), b = o.to_vec();
        bool less = true, greater = true;
        for (size_t i = 0; i < a.size(); i++) {
            if (a[i] < b[i]) greater = false;
            if (a[i] > b[i]) less = false;
        }
        if (less && greater) return 0;
        if (less) return -1;
        if (greater) return 1;
        return std::nullopt;
    }
    bool operator<=(const Energy& o) const {
        auto c = partial_cmp(o);
        return c && *c <= 0;
    }
    bool operator==(const Energy& o) const { return conf_ == o.conf_ && data_ == o.data_; }
    bool operator!=(const Energy& o) const { return !(*this == o); }

private:
    std::vector<uint32_t> data_;
    EnergyConf conf_;
};

// ---------------------------------------------------------------- Upd / Update (src/energy.rs:444-613)
struct Upd {
    enum Kind { Zero, Decrement, Min } kind = Zero;
    uint32_t idx = 0;   // Min(idx), 1-based
    static Upd zero() { return {Zero, 0}; }
    static Upd decrement() { return {Decrement, 0}; }
    static Upd min(uint32_t k) { return {Min, k}; }
    // the update! macro's integer form: 0 Zero, -1 Decrement, k >= 1 Min(k)
    static Upd from_int(int v) {
        if (v == 0) return zero();
        if (v == -1) return decrement();
        if (v < -1) throw Error(ErrorKind::InvalidArg, "Invalid update entry");
        return min((uint32_t)v);
    }
    int to_int() const { return kind == Zero ? 0 : kind == Decrement ? -1 : (int)idx; }
    uint32_t packed_repr() const { return kind == Zero ? 0u : kind == Decrement ? 1u : idx + 1u; }
    bool operator==(const Upd& o) const { return kind == o.kind && idx == o.idx; }
};

class Update {
public:
    Update(std::vector<uint32_t> data, EnergyConf conf) : data_(std::move(data)), conf_(conf) {}
    static std::optional<Update> make(const std::vector<Upd>& values, EnergyConf conf) {
        if (values.size() > conf.elements) return std::nullopt;
        std::vector<uint32_t> raw;
        for (const Upd& u : values) {
            if (u.kind == Upd::Min && u.idx > conf.elements) return std::nullopt;
            raw.push_back(u.packed_repr());
        }
        return Update(detail::pack(raw, conf.update_bits(), conf.update_size()), conf);
    }
    static std::optional<Update> from_ints(const std::vector<int>& values, EnergyConf conf) {
        std::vector<Upd> u;
        for (int v : values) u.push_back(Upd::from_int(v));
        return make(u, conf);
    }
    static Update zero(EnergyConf conf) { return Update(std::vector<uint32_t>(conf.update_size(), 0), conf); }
    const std::vector<uint32_t>& raw_data() const { return data_; }
    std::vector<Upd> to_vec() const {
        std::vector<Upd> out;
        for (uint32_t v : detail::unpack(data_.data(), conf_.elements, conf_.update_bits(), conf_.update_mask()))
            out.push_back(v == 0 ? Upd::zero() : v == 1 ? Upd::decrement() : Upd::min(v - 1));
        return out;
    }
    bool operator==(const Update& o) const { return conf_ == o.conf_ && data_ == o.data_; }

private:
    std::vector<uint32_t> data_;
    EnergyConf conf_;
};

// ---------------------------------------------------------------- EnergyArray (src/energy.rs:227-441)
class EnergyArray {
public:
    EnergyArray() : conf_(EnergyConf::STANDARD) {}
    EnergyArray(std::vector<uint32_t> data, EnergyConf conf) : data_(std::move(data)), conf_(conf) {
        if (data_.size() % conf_.energy_size()) throw Error(ErrorKind::InvalidArg, "array width does not match the energy configuration");
    }
    static EnergyArray empty(EnergyConf conf) { return EnergyArray({}, conf); }
    static EnergyArray zero(size_t n, EnergyConf conf) { return EnergyArray(std::vector<uint32_t>(n * conf.energy_size(), 0), conf); }
    // FromEnergyConf<&[Vec<u32>]> (:431-441)
    static EnergyArray from_conf(const std::vector<std::vector<uint32_t>>& rows, EnergyConf conf) {
        std::vector<uint32_t> data;
        for (const auto& r : rows) {
            auto e = Energy::make(r, conf);
            if (!e) throw Error(ErrorKind::InvalidArg, "Energies don't match configuration");
            data.insert(data.end(), e->raw_data().begin(), e->raw_data().end());
        }
        return EnergyArray(std::move(data), conf);
    }
    size_t n_energies() const { return data_.size() / conf_.energy_size(); }
    bool is_empty() const { return data_.empty(); }
    std::optional<Energy> get(size_t idx) const {
        if (idx >= n_energies()) return std::nullopt;
        const size_t w = conf_.energy_size();
        return Energy(std::vector<uint32_t>(data_.begin() + idx * w, data_.begin() + (idx + 1) * w), conf_);
    }
    std::vector<Energy> iter() const {
        std::vector<Energy> out;
        for (size_t i = 0; i < n_energies(); i++) out.push_back(*get(i));
        return out;
    }
    std::vector<std::vector<uint32_t>> to_vecs() const {
        std::vector<std::vector<uint32_t>> out;
        for (const Energy& e : iter()) out.push_back(e.to_vec());
        return out;
    }
    const std::vector<uint32_t>& data() const { return data_; }
    EnergyConf get_conf() const { return conf_; }
    // true iff no minimal budget lies at or below the equivalence's coordinates (:307-318)
    bool test_equivalence(const Energy& equivalence) const {
        if (conf_ != equivalence.get_conf()) throw Error(ErrorKind::InvalidArg, "Incompatible Energy configurations");
        for (const Energy& e : iter())
            if (e <= equivalence) return false;
        return true;
    }
    // order-insensitive, as the crate's PartialEq (:334-369)
    bool operator==(const EnergyArray& o) const {
        if (conf_ != o.conf_ || data_.size() != o.data_.size()) return false;
        std::set<std::vector<uint32_t>> rows;
        for (const Energy& e : iter()) rows.insert(e.raw_data());
        for (const Energy& e : o.iter())
            if (!rows.count(e.raw_data())) return false;
        return true;
    }
    bool operator!=(const EnergyArray& o) const { return !(*this == o); }

private:
    std::vector<uint32_t> data_;
    EnergyConf conf_;
};

// ---------------------------------------------------------------- UpdateArray (src/energy.rs:623-711)
class UpdateArray {
public:
    UpdateArray() : conf_(EnergyConf::STANDARD) {}
    UpdateArray(std::vector<uint32_t> data, EnergyConf conf) : data_(std::move(data)), conf_(conf) {}
    static UpdateArray empty(EnergyConf conf) { return UpdateArray({}, conf); }
    static UpdateArray from_updates(const std::vector<Update>& ups, EnergyConf conf) {
        std::vector<uint32_t> data;
        for (const Update& u : ups) data.insert(data.end(), u.raw_data().begin(), u.raw_data().end());
        return UpdateArray(std::move(data), conf);
    }
    size_t n_updates() const { return data_.size() / conf_.update_size(); }
    std::optional<Update> get(size_t idx) const {
        if (idx >= n_updates()) return std::nullopt;
        const size_t w = conf_.update_size();
        return Update(std::vector<uint32_t>(data_.begin() + idx * w, data_.begin() + (idx + 1) * w), conf_);
    }
    const std::vector<uint32_t>& data() const { return data_; }
    EnergyConf get_conf() const { return conf_; }

private:
    std::vector<uint32_t> data_;
    EnergyConf conf_;
};

// ---------------------------------------------------------------- GameGraph (src/energygame.rs:120-250)
class GameGraph {
public:
    std::vector<uint32_t> column_indices, row_offsets;
    UpdateArray weights;
    std::vector<uint32_t> rev_column_indices, rev_row_offsets;
    std::vector<bool> attacker_pos;

    GameGraph() = default;
    // GameGraph::new (:141-194): edges keep their per-source input order
    GameGraph(uint32_t n_vertices, const std::vector<std::tuple<uint32_t, uint32_t, Update>>& edges,
              std::vector<bool> attacker, EnergyConf conf)
        : attacker_pos(std::move(attacker)), conf_(conf) {
        std::vector<std::vector<uint32_t>> adj(n_vertices), rev(n_vertices);
        std::vector<std::vector<const Update*>> w(n_vertices);
        for (const auto& [from, to, u] : edges) {
            if (from >= n_vertices || to >= n_vertices) throw Error(ErrorKind::InvalidArg, "edge endpoint out of range");
            adj[from].push_back(to);
            rev[to].push_back(from);
            w[from].push_back(&u);
        }
        std::vector<Update> flat;
        row_offsets.push_back(0);
        rev_row_offsets.push_back(0);
        for (uint32_t v = 0; v < n_vertices; v++) {
            column_indices.insert(column_indices.end(), adj[v].begin(), adj[v].end());
            for (const Update* u : w[v]) flat.push_back(*u);
            row_offsets.push_back((uint32_t)column_indices.size());
            rev_column_indices.insert(rev_column_indices.end(), rev[v].begin(), rev[v].end());
            rev_row_offsets.push_back((uint32_t)rev_column_indices.size());
        }
        weights = UpdateArray::from_updates(flat, conf);
    }
    static GameGraph empty(EnergyConf conf) {
        GameGraph g;
        g.conf_ = conf;
        g.weights = UpdateArray::empty(conf);
        return g;
    }
    uint32_t n_vertices() const { return (uint32_t)attacker_pos.size(); }
    EnergyConf get_conf() const { return conf_; }
    std::vector<uint32_t> adj(uint32_t v) const {
        return std::vector<uint32_t>(column_indices.begin() + row_offsets[v], column_indices.begin() + row_offsets[v + 1]);
    }
    std::vector<uint32_t> reverse(uint32_t v) const {
        return std::vector<uint32_t>(rev_column_indices.begin() + rev_row_offsets[v],
                                     rev_column_indices.begin() + rev_row_offsets[v + 1]);
    }
    void make_reverse() {
        std::vector<std::vector<uint32_t>> rev(n_vertices());
        for (uint32_t f = 0; f < n_vertices(); f++)
            for (uint32_t k = row_offsets[f]; k < row_offsets[f + 1]; k++) rev[column_indices[k]].push_back(f);
        rev_row_offsets.assign(1, 0);
        rev_column_indices.clear();
        for (auto& r : rev) {
            rev_column_indices.insert(rev_column_indices.end(), r.begin(), r.end());
            rev_row_offsets.push_back((uint32_t)rev_column_indices.size());
        }
    }

private:
    friend class EnergyGame;
    EnergyConf conf_ = EnergyConf::STANDARD;
};

namespace detail {
inline std::vector<EnergyArray> to_arrays(const geq_energies& out, EnergyConf conf) {
    std::vector<EnergyArray> res;
    const uint32_t es = out.energy_size;
    for (uint32_t v = 0; v < out.n_nodes; v++)
        res.emplace_back(std::vector<uint32_t>(out.data + (size_t)out.offsets[v] * es, out.data + (size_t)out.offsets[v + 1] * es),
                         conf);
    return res;
}
struct EnergiesGuard {
    geq_energies e{};
    ~EnergiesGuard() { geq_energies_free(&e); }
};
}  // namespace detail

// ---------------------------------------------------------------- EnergyGame (src/energygame.rs:325-384)
class EnergyGame {
public:
    GameGraph graph;
    std::vector<EnergyArray> energies;
    std::vector<uint32_t> to_reach;

    // defender positions without successors are the attacker's targets (:347-352)
    static EnergyGame standard_reach(GameGraph graph) {
        std::vector<uint32_t> reach;
        for (uint32_t v = 0; v < graph.n_vertices(); v++)
            if (!graph.attacker_pos[v] && graph.row_offsets[v] == graph.row_offsets[v + 1]) reach.push_back(v);
        return with_reach(std::move(graph), std::move(reach));
    }
    static EnergyGame with_reach(GameGraph graph, std::vector<uint32_t> to_reach) {
        EnergyGame g;
        const EnergyConf conf = graph.get_conf();
        g.energies.assign(graph.n_vertices(), EnergyArray::empty(conf));
        for (uint32_t v : to_reach) g.energies.at(v) = EnergyArray::zero(1, conf);
        g.graph = std::move(graph);
        g.to_reach = std::move(to_reach);
        return g;
    }
    // Process the game on the GPU; returns the winning budgets of every node (:379-383)
    const std::vector<EnergyArray>& run() {
        const EnergyConf conf = graph.get_conf();
        std::vector<uint8_t> atk(graph.n_vertices());
        for (uint32_t v = 0; v < graph.n_vertices(); v++) atk[v] = graph.attacker_pos[v] ? 1 : 0;
        if (graph.row_offsets.size() != (size_t)graph.n_vertices() + 1)
            throw Error(ErrorKind::InvalidArg, "row_offsets must have n_vertices + 1 entries");
        geq_graph g{};
        g.conf = geq_conf{conf.elements, conf.max};
        g.n_vertices = graph.n_vertices();
        g.n_edges = (uint32_t)graph.column_indices.size();
        g.row_offsets = graph.row_offsets.data();
        g.column_indices = graph.column_indices.data();
        g.weights = graph.weights.data().data();
        g.attacker_pos = atk.data();
        const bool have_rev = graph.rev_row_offsets.size() == graph.row_offsets.size();
        g.rev_row_offsets = have_rev ? graph.rev_row_offsets.data() : nullptr;
        g.rev_column_indices = have_rev ? graph.rev_column_indices.data() : nullptr;
        // starting energies: the current energies (with_reach's zeros, or a caller's)
        std::vector<uint32_t> offs(1, 0), data;
        for (const EnergyArray& e : energies) {
            data.insert(data.end(), e.data().begin(), e.data().end());
            offs.push_back((uint32_t)(data.size() / conf.energy_size()));
        }
        if (data.empty()) data.push_back(0);
        geq_energies init{graph.n_vertices(), conf.energy_size(), offs.data(), data.data()};
        detail::EnergiesGuard out;
        geq_stats st{};
        detail::check(geq_solve(detail::context(), &g, to_reach.data(), (uint32_t)to_reach.size(), &init, &out.e, &st));
        energies = detail::to_arrays(out.e, conf);
        last_stats = st;
        return energies;
    }
    geq_stats last_stats{};
};

// ---------------------------------------------------------------- positions (src/gamebuild/gamepos.rs)
struct AttackPosition {
    uint32_t p;
    std::vector<uint32_t> q;
    bool operator==(const AttackPosition& o) const { return p == o.p && q == o.q; }
};
struct SingletonPosition {
    uint32_t p, q;
    bool operator==(const SingletonPosition& o) const { return p == o.p && q == o.q; }
};
struct DefendPosition {
    uint32_t p;
    std::vector<uint32_t> q, qx;
    bool operator==(const DefendPosition& o) const { return p == o.p && q == o.q && qx == o.qx; }
};
struct Position {
    std::variant<AttackPosition, SingletonPosition, DefendPosition> v;
    static Position attack(uint32_t p, std::vector<uint32_t> q) { return {AttackPosition{p, std::move(q)}}; }
    static Position clause(uint32_t p, uint32_t q) { return {SingletonPosition{p, q}}; }
    static Position defend(uint32_t p, std::vector<uint32_t> q, std::vector<uint32_t> qx) {
        return {DefendPosition{p, std::move(q), std::move(qx)}};
    }
    bool is_attack() const { return !std::holds_alternative<DefendPosition>(v); }
    bool operator==(const Position& o) const { return v == o.v; }
};

// ---------------------------------------------------------------- Equivalence (src/equivalence.rs)
class EquivalenceRelation {
public:
    explicit EquivalenceRelation(size_t n = 0) : parent_(n) {
        for (size_t i = 0; i < n; i++) parent_[i] = i;
    }
    size_t find(size_t x) const {
        while (parent_[x] != x) x = parent_[x];
        return x;
    }
    void unite(size_t a, size_t b) {
        a = find(a);
        b = find(b);
        if (a != b) parent_[std::max(a, b)] = std::min(a, b);
    }
    size_t len() const { return parent_.size(); }
    std::vector<std::vector<size_t>> get_classes() const {
        std::map<size_t, size_t> idx;
        std::vector<std::vector<size_t>> out;
        for (size_t i = 0; i < len(); i++) {
            const size_t r = find(i);
            auto it = idx.find(r);
            if (it == idx.end()) {
                idx[r] = out.size();
                out.push_back({i});
            } else {
                out[it->second].push_back(i);
            }
        }
        return out;
    }
    uint32_t count_classes() const {
        std::set<size_t> roots;
        for (size_t i = 0; i < len(); i++) roots.insert(find(i));
        return (uint32_t)roots.size();
    }
    std::vector<size_t> class_of(size_t proc) const {
        std::vector<size_t> out;
        const size_t r = find(proc);
        for (size_t i = 0; i < len(); i++)
            if (find(i) == r) out.push_back(i);
        return out;
    }
    EquivalenceRelation with_mapping(const std::vector<size_t>& mapping) const {
        EquivalenceRelation out(mapping.size());
        std::map<size_t, size_t> rep;
        for (size_t proc = 0; proc < mapping.size(); proc++) {
            const size_t part = find(mapping[proc]);
            auto it = rep.find(part);
            if (it == rep.end()) rep[part] = proc;
            else out.unite(proc, it->second);
        }
        return out;
    }

private:
    std::vector<size_t> parent_;
};

struct StartInfo {
    std::vector<AttackPosition> starting_points;
    size_t n_processes = 0;
};

class Equivalence {
public:
    Equivalence(StartInfo info, std::vector<EnergyArray> energies) : info_(std::move(info)) {
        if (energies.size() < info_.starting_points.size()) throw Error(ErrorKind::InvalidArg, "Not enough energies");
        energies.resize(info_.starting_points.size());
        energies_ = std::move(energies);
        for (size_t i = 0; i < info_.starting_points.size(); i++)
            pos_to_idx_[{info_.starting_points[i].p, info_.starting_points[i].q.at(0)}] = i;
    }
    void set_minimization(std::vector<size_t> m) { minimization_ = std::move(m); }
    // budgets of the pair, after the minimisation mapping (:96-98)
    const EnergyArray* energies(uint32_t p, uint32_t q) const { return minimized(mapped(p), mapped(q)); }
    bool preorder(uint32_t p, uint32_t q, const Energy& eq) const {
        p = mapped(p);
        q = mapped(q);
        if (p == q) return true;
        const EnergyArray* e = minimized(p, q);
        return e && e->test_equivalence(eq);
    }
    bool equiv(uint32_t p, uint32_t q, const Energy& eq) const { return preorder(p, q, eq) && preorder(q, p, eq); }
    EquivalenceRelation relation(const Energy& eq) const {
        EquivalenceRelation uf(info_.n_processes);
        for (size_t i = 0; i < info_.starting_points.size(); i++) {
            const uint32_t p = info_.starting_points[i].p, q = info_.starting_points[i].q.at(0);
            const EnergyArray* back = minimized(q, p);
            if (!back) throw Error(ErrorKind::InvalidArg, "Symmetric position not included in starting points for energy game");
            if (energies_[i].test_equivalence(eq) && back->test_equivalence(eq)) uf.unite(p, q);
        }
        return minimization_ ? uf.with_mapping(*minimization_) : uf;
    }

private:
    uint32_t mapped(uint32_t p) const { return minimization_ ? (uint32_t)minimization_->at(p) : p; }
    const EnergyArray* minimized(uint32_t p, uint32_t q) const {
        if (p >= info_.n_processes || q >= info_.n_processes) throw Error(ErrorKind::InvalidArg, "process out of range");
        auto it = pos_to_idx_.find({p, q});
        return it == pos_to_idx_.end() ? nullptr : &energies_[it->second];
    }
    StartInfo info_;
    std::vector<EnergyArray> energies_;
    std::map<std::pair<uint32_t, uint32_t>, size_t> pos_to_idx_;
    std::optional<std::vector<size_t>> minimization_;
};

// ---------------------------------------------------------------- TransitionSystem (src/lib.rs)
struct Transition {
    uint32_t process;
    int32_t label;
};
using Edge = std::tuple<uint32_t, uint32_t, int32_t>;

class GameBuild;

class TransitionSystem {
public:
    TransitionSystem() = default;
    // From<Vec<Vec<Transition>>> (:508-515): each adjacency list stably sorted by label
    static TransitionSystem from_adj(std::vector<std::vector<Transition>> adj) {
        for (auto& r : adj)
            std::stable_sort(r.begin(), r.end(), [](const Transition& a, const Transition& b) { return a.label < b.label; });
        return from_adj_unchecked(std::move(adj));
    }
    static TransitionSystem from_adj_unchecked(std::vector<std::vector<Transition>> adj) {
        TransitionSystem t;
        t.row_.assign(1, 0);
        for (const auto& r : adj) {
            for (const Transition& x : r) {
                t.dst_.push_back(x.process);
                t.lbl_.push_back(x.label);
            }
            t.row_.push_back((uint32_t)t.dst_.size());
        }
        return t;
    }
    // From<Vec<Edge>> (:517-535)
    static TransitionSystem from_edges(const std::vector<Edge>& edges) {
        uint32_t n = 0;
        for (const auto& [a, b, l] : edges) n = std::max({n, a + 1, b + 1});
        std::vector<std::vector<Transition>> adj(n);
        for (const auto& [a, b, l] : edges) adj[a].push_back({b, l});
        return from_adj(std::move(adj));
    }
    // CSV lines `from,to,label`; label "i" is 0, others numbered by first appearance (:372-406)
    static TransitionSystem from_csv_lines(std::istream& in) {
        std::unordered_map<std::string, int32_t> labels{{"i", 0}};
        std::vector<Edge> edges;
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            const size_t c1 = line.find(','), c2 = c1 == std::string::npos ? c1 : line.find(',', c1 + 1);
            if (c2 == std::string::npos) throw Error(ErrorKind::InvalidArg, "insufficient number of values in line");
            const uint32_t a = (uint32_t)std::stoul(line.substr(0, c1)), b = (uint32_t)std::stoul(line.substr(c1 + 1, c2 - c1 - 1));
            std::string name = line.substr(c2 + 1);
            if (name.size() >= 2 && name.front() == '"' && name.back() == '"') name = name.substr(1, name.size() - 2);
            auto it = labels.find(name);
            const int32_t id = it == labels.end() ? (labels[name] = (int32_t)labels.size()) : it->second;
            edges.emplace_back(a, b, id);
        }
        return from_edges(edges);
    }
    static TransitionSystem from_csv_file(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw Error(ErrorKind::InvalidArg, "cannot open " + path);
        return from_csv_lines(f);
    }

    uint32_t n_vertices() const { return row_.empty() ? 0 : (uint32_t)row_.size() - 1; }
    std::vector<Transition> adj(uint32_t p) const {
        std::vector<Transition> out;
        for (uint32_t k = row_.at(p); k < row_.at(p + 1); k++) out.push_back({dst_[k], lbl_[k]});
        return out;
    }
    geq_lts c_lts() const {
        return geq_lts{n_vertices(), (uint32_t)dst_.size(), row_.data(), dst_.empty() ? nullptr : dst_.data(),
                       lbl_.empty() ? nullptr : lbl_.data()};
    }

    // (partition, class count) of strong bisimilarity (:444-478), refined on the device
    // (on_device = false: the library's host routine, same result)
    std::pair<std::vector<size_t>, size_t> signature_refinement(bool on_device = true) const {
        std::vector<uint32_t> part(std::max<uint32_t>(n_vertices(), 1));
        uint32_t cnt = 0;
        const geq_lts l = c_lts();
        if (on_device) detail::check(geq_signature_refinement_gpu(detail::context(), &l, part.data(), &cnt, nullptr));
        else detail::check(geq_signature_refinement(&l, part.data(), &cnt));
        return {std::vector<size_t>(part.begin(), part.begin() + n_vertices()), cnt};
    }
    // (quotient LTS, partition) (:413-432)
    std::pair<TransitionSystem, std::vector<size_t>> bisimilar_minimize(bool on_device = true) const {
        const uint32_t n = n_vertices(), m = (uint32_t)dst_.size();
        std::vector<uint32_t> part(std::max(n, 1u)), orow(n + 1), odst(std::max(m, 1u));
        std::vector<int32_t> olbl(std::max(m, 1u));
        uint32_t cnt = 0, om = 0;
        const geq_lts l = c_lts();
        if (on_device)
            detail::check(geq_bisimilar_minimize_gpu(detail::context(), &l, part.data(), &cnt, orow.data(), odst.data(), olbl.data(), &om));
        else
            detail::check(geq_bisimilar_minimize(&l, part.data(), &cnt, orow.data(), odst.data(), olbl.data(), &om));
        TransitionSystem t;
        t.row_.assign(orow.begin(), orow.begin() + cnt + 1);
        t.dst_.assign(odst.begin(), odst.begin() + om);
        t.lbl_.assign(olbl.begin(), olbl.begin() + om);
        return {t, std::vector<size_t>(part.begin(), part.begin() + n)};
    }
    // processes grouped by their set of enabled labels, groups ordered by first member (src/gamebuild.rs:104-133)
    std::vector<std::vector<uint32_t>> enabledness_partitions(const std::vector<uint32_t>* processes = nullptr) const {
        std::map<std::vector<int32_t>, size_t> idx;
        std::vector<std::vector<uint32_t>> out;
        auto add = [&](uint32_t p) {
            std::vector<int32_t> key(lbl_.begin() + row_[p], lbl_.begin() + row_[p + 1]);
            std::sort(key.begin(), key.end());
            key.erase(std::unique(key.begin(), key.end()), key.end());
            auto it = idx.find(key);
            if (it == idx.end()) {
                idx[key] = out.size();
                out.push_back({p});
            } else {
                out[it->second].push_back(p);
            }
        };
        if (processes)
            for (uint32_t p : *processes) add(p);
        else
            for (uint32_t p = 0; p < n_vertices(); p++) add(p);
        return out;
    }

    EnergyArray compare(uint32_t p, uint32_t q) const;
    EnergyArray compare_unminimized(uint32_t p, uint32_t q) const;
    Equivalence equivalences() const;
    Equivalence equivalences_unminimized() const;

private:
    std::vector<uint32_t> row_{0}, dst_;
    std::vector<int32_t> lbl_;
};

// ---------------------------------------------------------------- GameBuild (src/gamebuild.rs:139-265)
class GameBuild {
public:
    GameBuild(const TransitionSystem& lts, std::vector<std::pair<uint32_t, uint32_t>> pairs, uint64_t max_positions = 0)
        : pairs_(std::move(pairs)) {
        std::vector<uint32_t> flat;
        for (auto [p, q] : pairs_) {
            flat.push_back(p);
            flat.push_back(q);
        }
        if (flat.empty()) flat.assign(2, 0);
        const geq_lts l = lts.c_lts();
        geq_built* b = nullptr;
        detail::check(geq_gamebuild(detail::context(), &l, flat.data(), (uint32_t)pairs_.size(), max_positions, &b, &stats_));
        built_ = std::shared_ptr<geq_built>(b, Free{});
        n_processes_ = lts.n_vertices();
    }
    static GameBuild compare(const TransitionSystem& lts, uint32_t p, uint32_t q) { return GameBuild(lts, {{p, q}}); }
    static std::pair<GameBuild, StartInfo> compare_multiple(const TransitionSystem& lts, const std::vector<uint32_t>& processes) {
        return seeded(lts, lts.enabledness_partitions(&processes));
    }
    static std::pair<GameBuild, StartInfo> compare_all(const TransitionSystem& lts) {
        return seeded(lts, lts.enabledness_partitions());
    }

    size_t n_starting_points() const { return pairs_.size(); }
    std::vector<AttackPosition> starting_points() const {
        std::vector<AttackPosition> out;
        for (auto [p, q] : pairs_) out.push_back({p, {q}});
        return out;
    }
    const geq_build_stats& build_stats() const { return stats_; }
    uint64_t n_positions() const { return stats_.positions; }

    // EnergyGame::standard_reach(game).run() on the device-resident game; a
    // build stopped by max_positions throws unless `capped` asks for the solve
    // of the capped game (its unexpanded frontier never reached)
    geq_stats solve(bool capped = false) {
        geq_stats st{};
        detail::check(capped ? geq_built_solve_capped(built_.get(), 0, &st) : geq_built_solve(built_.get(), 0, &st));
        return st;
    }
    bool truncated() const { return stats_.truncated != 0; }
    std::vector<EnergyArray> energies(uint32_t first, uint32_t count) const {
        detail::EnergiesGuard out;
        detail::check(geq_built_energies(built_.get(), first, count, &out.e));
        return detail::to_arrays(out.e, EnergyConf::STANDARD);
    }
    // the game graph and the positions, downloaded from the device
    GameGraph game() const {
        download();
        GameGraph g = GameGraph::empty(EnergyConf::STANDARD);
        g.row_offsets = host_->row;
        g.column_indices = host_->col;
        g.weights = UpdateArray(host_->wt, EnergyConf::STANDARD);
        g.attacker_pos.assign(host_->atk.begin(), host_->atk.end());
        g.make_reverse();
        return g;
    }
    std::vector<Position> nodes() const {
        download();
        std::vector<Position> out;
        const Host& h = *host_;
        for (size_t i = 0; i < h.kind.size(); i++) {
            std::vector<uint32_t> q(h.pool.begin() + h.q_off[i], h.pool.begin() + h.q_off[i] + h.q_len[i]);
            if (h.kind[i] == 0) out.push_back(Position::attack(h.p[i], q));
            else if (h.kind[i] == 1) out.push_back(Position::clause(h.p[i], q.at(0)));
            else out.push_back(Position::defend(h.p[i], q, std::vector<uint32_t>(h.pool.begin() + h.x_off[i],
                                                                                h.pool.begin() + h.x_off[i] + h.x_len[i])));
        }
        return out;
    }

private:
    static std::pair<GameBuild, StartInfo> seeded(const TransitionSystem& lts, const std::vector<std::vector<uint32_t>>& parts) {
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (const auto& part : parts)
            for (uint32_t p : part)
                for (uint32_t q : part)
                    if (p != q) pairs.emplace_back(p, q);
        GameBuild b(lts, pairs);
        StartInfo info{b.starting_points(), lts.n_vertices()};
        return {std::move(b), std::move(info)};
    }
    struct Host {
        std::vector<uint32_t> row, col, wt, p, q_off, q_len, x_off, x_len, pool;
        std::vector<uint8_t> atk, kind;
    };
    void download() const {
        if (host_) return;
        geq_built_game g{};
        detail::check(geq_built_download(built_.get(), &g));
        auto h = std::make_shared<Host>();
        const size_t V = g.n_positions, E = g.n_edges;
        h->row.assign(g.row_offsets, g.row_offsets + V + 1);
        h->col.assign(g.column_indices, g.column_indices + E);
        h->wt.assign(g.weights, g.weights + E);
        h->atk.assign(g.attacker_pos, g.attacker_pos + V);
        h->kind.assign(g.kind, g.kind + V);
        h->p.assign(g.p, g.p + V);
        h->q_off.assign(g.q_off, g.q_off + V);
        h->q_len.assign(g.q_len, g.q_len + V);
        h->x_off.assign(g.x_off, g.x_off + V);
        h->x_len.assign(g.x_len, g.x_len + V);
        h->pool.assign(g.pool, g.pool + g.pool_len);
        geq_built_game_free(&g);
        host_ = h;
    }
    struct Free {
        void operator()(geq_built* b) const { geq_built_free(b); }
    };
    std::vector<std::pair<uint32_t, uint32_t>> pairs_;
    std::shared_ptr<geq_built> built_;
    geq_build_stats stats_{};
    uint32_t n_processes_ = 0;
    mutable std::shared_ptr<Host> host_;
};

// compare (src/lib.rs:275-283): minimise first; bisimilar processes get no budgets
inline EnergyArray TransitionSystem::compare(uint32_t p, uint32_t q) const {
    auto [reduced, part] = bisimilar_minimize();
    const uint32_t pm = (uint32_t)part.at(p), qm = (uint32_t)part.at(q);
    if (pm == qm) return EnergyArray::empty(EnergyConf::STANDARD);
    return reduced.compare_unminimized(pm, qm);
}
inline EnergyArray TransitionSystem::compare_unminimized(uint32_t p, uint32_t q) const {
    GameBuild b = GameBuild::compare(*this, p, q);
    b.solve();
    return b.energies(0, 1).at(0);
}
// equivalences (src/lib.rs:329-343)
inline Equivalence TransitionSystem::equivalences() const {
    auto [reduced, part] = bisimilar_minimize();
    Equivalence eq = reduced.equivalences_unminimized();
    eq.set_minimization(part);
    return eq;
}
inline Equivalence TransitionSystem::equivalences_unminimized() const {
    auto [b, info] = GameBuild::compare_all(*this);
    b.solve();
    return Equivalence(info, b.energies(0, (uint32_t)b.n_starting_points()));
}

// ---------------------------------------------------------------- std_equivalences (src/energy.rs:198-225)
namespace std_equivalences {
inline Energy make_(std::vector<uint32_t> v) { return *Energy::make(v, EnergyConf::STANDARD); }
inline Energy enabledness() { return make_({1, 1}); }
inline Energy traces() { return make_({3, 1}); }
inline Energy failures() { return make_({3, 2, 0, 0, 1, 1}); }
inline Energy failure_traces() { return make_({3, 3, 3, 0, 1, 1}); }
inline Energy readiness() { return make_({3, 2, 1, 1, 1, 1}); }
inline Energy readiness_traces() { return make_({3, 3, 3, 1, 1, 1}); }
inline Energy revivals() { return make_({3, 2, 1, 0, 1, 1}); }
inline Energy impossible_futures() { return make_({3, 2, 0, 0, 3, 1}); }
inline Energy possible_futures() { return make_({3, 2, 3, 3, 3, 1}); }
inline Energy simulation() { return make_({3, 3, 3, 3, 0, 0}); }
inline Energy ready_simulation() { return make_({3, 3, 3, 3, 1, 1}); }
inline Energy nested_2simulation() { return make_({3, 3, 3, 3, 3, 1}); }
inline Energy bisimulation() { return make_({3, 3, 3, 3, 3, 3}); }
inline std::vector<std::pair<std::string, Energy>> all() {
    return {{"enabledness", enabledness()},
            {"traces", traces()},
            {"failures", failures()},
            {"failure_traces", failure_traces()},
            {"readiness", readiness()},
            {"readiness_traces", readiness_traces()},
            {"revivals", revivals()},
            {"impossible_futures", impossible_futures()},
            {"possible_futures", possible_futures()},
            {"simulation", simulation()},
            {"ready_simulation", ready_simulation()},
            {"nested_2simulation", nested_2simulation()},
            {"bisimulation", bisimulation()}};
}
inline Energy by_name(const std::string& name) {
    for (auto& [n, e] : all())
        if (n == name) return e;
    throw Error(ErrorKind::InvalidArg, "unknown equivalence " + name);
}
}  // namespace std_equivalences

}  // namespace gpuequiv
