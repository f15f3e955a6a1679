// Tests of the C++ host API (include/gpuequiv.hpp), driven by a case file
// written from the golden fixtures of the reference's own tests
// (tests/energygame.rs, tests/complete.rs, src/lib.rs doctests, src/energy.rs).
//
//   test_api --cpu <cases>   packing known answers + the no-device error path
//   test_api <cases>         everything, on the GPU
//
// Prints one line per case and "ALL OK" at the end; exits non-zero on the
// first failure.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <set>
#include <string>
#include <vector>

#include "gpuequiv.hpp"

using namespace gpuequiv;

namespace {

struct Reader {
    std::ifstream in;
    explicit Reader(const char* path) : in(path) {
        if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    }
    bool next(std::string& tok) { return bool(in >> tok); }
    std::string word() {
        std::string s;
        if (!(in >> s)) throw std::runtime_error("case file ended early");
        return s;
    }
    long num() { return std::stol(word()); }
    std::vector<uint32_t> list() {
        std::vector<uint32_t> v(num());
        for (auto& x : v) x = (uint32_t)num();
        return v;
    }
    std::vector<std::vector<uint32_t>> energies() {
        std::vector<std::vector<uint32_t>> out(num());
        for (auto& e : out) e = list();
        return out;
    }
};

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        std::fflush(stdout);
        failures++;
    }
}

std::string show(const EnergyArray& a) {
    std::string s = "[";
    for (const auto& v : a.to_vecs()) {
        s += "(";
        for (uint32_t x : v) s += std::to_string(x);
        s += ")";
    }
    return s + "]";
}

Position read_position(Reader& r) {
    const std::string k = r.word();
    const uint32_t p = (uint32_t)r.num();
    if (k == "A") return Position::attack(p, r.list());
    if (k == "C") return Position::clause(p, (uint32_t)r.num());
    auto q = r.list();
    return Position::defend(p, q, r.list());
}

}  // namespace

int main(int argc, char** argv) {
    bool cpu_only = false;
    int ai = 1;
    if (argc > 1 && std::string(argv[1]) == "--cpu") {
        cpu_only = true;
        ai = 2;
    }
    if (ai >= argc) {
        std::fprintf(stderr, "usage: test_api [--cpu] <cases>\n");
        return 2;
    }
    try {
        Reader r(argv[ai]);
        std::string tok;
        while (r.next(tok)) {
            if (tok == "pack_energy" || tok == "pack_update") {
                const EnergyConf conf{(uint32_t)r.num(), (uint32_t)r.num()};
                std::vector<long> vals(r.num());
                for (auto& v : vals) v = r.num();
                const auto words = r.list();
                if (tok == "pack_energy") {
                    std::vector<uint32_t> u(vals.begin(), vals.end());
                    auto e = Energy::make(u, conf);
                    expect(e && e->raw_data() == words, "Energy::make packing");
                    std::vector<uint32_t> back = e->to_vec();
                    u.resize(conf.elements, 0);
                    expect(back == u, "Energy::to_vec round trip");
                } else {
                    std::vector<int> iv(vals.begin(), vals.end());
                    auto u = Update::from_ints(iv, conf);
                    expect(u && u->raw_data() == words, "Update packing");
                }
            } else if (tok == "no_device") {
                // on a machine without a GPU every compute call reports NoAdapter
                bool raised = false;
                try {
                    GameGraph g(2, {{0, 1, *Update::from_ints({-1}, EnergyConf::STANDARD)}}, {true, false}, EnergyConf::STANDARD);
                    EnergyGame::standard_reach(g).run();
                } catch (const Error& e) {
                    raised = e.kind == ErrorKind::NoAdapter;
                }
                expect(raised, "EnergyGame::run without a device raises NoAdapter");
                std::printf("no_device ok\n");
            } else if (tok == "game") {
                const std::string name = r.word();
                const EnergyConf conf{(uint32_t)r.num(), (uint32_t)r.num()};
                const uint32_t n = (uint32_t)r.num();
                std::vector<bool> atk(n, false);
                for (uint32_t a : r.list()) atk[a] = true;
                std::vector<std::tuple<uint32_t, uint32_t, Update>> edges;
                for (long m = r.num(); m > 0; m--) {
                    const uint32_t f = (uint32_t)r.num(), t = (uint32_t)r.num();
                    std::vector<int> u(r.num());
                    for (auto& x : u) x = (int)r.num();
                    edges.emplace_back(f, t, *Update::from_ints(u, conf));
                }
                std::vector<EnergyArray> want;
                for (uint32_t v = 0; v < n; v++) want.push_back(EnergyArray::from_conf(r.energies(), conf));
                if (cpu_only) continue;
                EnergyGame game = EnergyGame::standard_reach(GameGraph(n, edges, atk, conf));
                const auto& got = game.run();
                for (uint32_t v = 0; v < n; v++)
                    expect(got[v] == want[v], "game " + name + " node " + std::to_string(v) + ": got " + show(got[v]) +
                                                  " want " + show(want[v]));
                std::printf("game %s ok (%u nodes)\n", name.c_str(), n);
            } else if (tok == "lts") {
                const std::string name = r.word();
                std::vector<Edge> edges;
                for (long m = r.num(); m > 0; m--) {
                    const uint32_t a = (uint32_t)r.num(), b = (uint32_t)r.num();
                    edges.emplace_back(a, b, (int32_t)r.num());
                }
                const TransitionSystem lts = TransitionSystem::from_edges(edges);
                {
                    // the CSR holds exactly the input transitions, each state's sorted by label
                    uint32_t n = 0;
                    for (const auto& [a, b, l] : edges) n = std::max({n, a + 1, b + 1});
                    expect(lts.n_vertices() == n, name + " n_vertices");
                    std::multiset<std::tuple<uint32_t, uint32_t, int32_t>> want(edges.begin(), edges.end()), got;
                    for (uint32_t p = 0; p < lts.n_vertices(); p++) {
                        const auto adj = lts.adj(p);
                        for (size_t k = 0; k < adj.size(); k++) {
                            got.insert({p, adj[k].process, adj[k].label});
                            if (k) expect(adj[k - 1].label <= adj[k].label, name + " adjacency sorted by label");
                        }
                    }
                    expect(got == want, name + " transitions");
                    // strong bisimulation classes are numbered by first occurrence
                    const auto [part, cnt] = lts.signature_refinement(!cpu_only);
                    expect(part.size() == n && (n == 0 || part[0] == 0), name + " signature_refinement");
                    for (size_t v : part) expect(v < cnt, name + " partition id range");
                }
                std::string sub;
                while ((sub = r.word()) != "end") {
                    if (sub == "compare") {
                        // budgets of compare(p, q) + the verdicts that hold / fail
                        const uint32_t p = (uint32_t)r.num(), q = (uint32_t)r.num();
                        const auto holds = r.num();
                        std::vector<std::string> h, f;
                        for (long i = 0; i < holds; i++) h.push_back(r.word());
                        const auto fails = r.num();
                        for (long i = 0; i < fails; i++) f.push_back(r.word());
                        const long has_e = r.num();
                        std::vector<std::vector<uint32_t>> e;
                        if (has_e) e = r.energies();
                        if (cpu_only) continue;
                        const EnergyArray got = lts.compare(p, q);
                        if (has_e) expect(got == EnergyArray::from_conf(e, EnergyConf::STANDARD), name + " compare budgets " + show(got));
                        for (auto& s : h) expect(got.test_equivalence(std_equivalences::by_name(s)), name + " holds " + s);
                        for (auto& s : f) expect(!got.test_equivalence(std_equivalences::by_name(s)), name + " fails " + s);
                    } else if (sub == "game") {
                        // GameBuild::compare: exact positions, adjacency and energies
                        const uint32_t p = (uint32_t)r.num(), q = (uint32_t)r.num();
                        std::vector<Position> pos(r.num());
                        for (auto& x : pos) x = read_position(r);
                        std::vector<std::vector<uint32_t>> adj(pos.size());
                        for (auto& a : adj) a = r.list();
                        std::vector<EnergyArray> en;
                        for (size_t v = 0; v < pos.size(); v++) en.push_back(EnergyArray::from_conf(r.energies(), EnergyConf::STANDARD));
                        if (cpu_only) continue;
                        GameBuild b = GameBuild::compare(lts, p, q);
                        expect(b.nodes() == pos, name + " positions");
                        const GameGraph g = b.game();
                        for (uint32_t v = 0; v < pos.size(); v++) {
                            auto s = g.adj(v);
                            std::sort(s.begin(), s.end());
                            expect(s == adj[v], name + " adjacency of " + std::to_string(v));
                        }
                        b.solve();
                        const auto got = b.energies(0, (uint32_t)b.n_positions());
                        for (uint32_t v = 0; v < pos.size(); v++)
                            expect(got[v] == en[v], name + " energies of " + std::to_string(v) + ": " + show(got[v]));
                    } else if (sub == "preorder" || sub == "equiv") {
                        const uint32_t p = (uint32_t)r.num(), q = (uint32_t)r.num();
                        const std::string eqn = r.word();
                        const bool want = r.num() != 0;
                        if (cpu_only) continue;
                        const Equivalence eq = lts.equivalences();
                        const bool got = sub == "preorder" ? eq.preorder(p, q, std_equivalences::by_name(eqn))
                                                           : eq.equiv(p, q, std_equivalences::by_name(eqn));
                        expect(got == want, name + " " + sub + " " + std::to_string(p) + " " + std::to_string(q) + " " + eqn);
                    } else if (sub == "classes") {
                        const std::string eqn = r.word();
                        const long count = r.num();
                        const auto cls0 = r.list();
                        if (cpu_only) continue;
                        const EquivalenceRelation rel = lts.equivalences().relation(std_equivalences::by_name(eqn));
                        expect(rel.count_classes() == (uint32_t)count, name + " count_classes");
                        const auto c = rel.class_of(0);
                        expect(std::vector<uint32_t>(c.begin(), c.end()) == cls0, name + " class_of(0)");
                    } else {
                        throw std::runtime_error("unknown lts item " + sub);
                    }
                }
                if (!cpu_only) std::printf("lts %s ok\n", name.c_str());
            } else {
                throw std::runtime_error("unknown case " + tok);
            }
        }
    } catch (const Error& e) {
        std::printf("FAIL gpuequiv::Error %d: %s\n", (int)e.kind, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
