// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
// the C ABI's no-device entry points (signature refinement, minimisation,
// standard_reach, argument checks, the no-adapter path of a context) from a
// host-instrumented build of the library, and the CPU oracle (packing,
// inverse updates, the fixpoint for several energy configurations, gamebuild,
// minimisation) compiled into this program.  Random inputs from fixed seeds;
// the two minimisations must agree.  Prints "ALL OK" and exits 0.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "gpuequiv.h"

extern "C" {
struct orc_lists {
    uint32_t n_nodes, esz;
    uint32_t* offsets;
    uint32_t* data;
};
struct orc_game {
    uint32_t V, E, n_start;
    uint32_t *row, *col, *wt;
    uint8_t* atk;
    uint8_t* kind;
    uint32_t *p, *q1, *q_off, *qx_off, *pool;
};
void orc_conf_info(uint32_t elements, uint32_t max, uint32_t* out6);
int orc_pack_energy(uint32_t elements, uint32_t max, const uint32_t* vals, uint32_t n, uint32_t* out);
void orc_unpack_energy(uint32_t elements, uint32_t max, const uint32_t* w, uint32_t* out);
int orc_pack_update(uint32_t elements, uint32_t max, const int32_t* upd, uint32_t n, uint32_t* out);
void orc_inv_update(uint32_t elements, uint32_t max, const uint32_t* e, const uint32_t* u, uint32_t* out);
orc_lists* orc_solve(uint32_t elements, uint32_t max, uint32_t V, const uint32_t* row, const uint32_t* col, const uint32_t* wt,
                     const uint8_t* atk, const uint32_t* init_off, const uint32_t* init_data, const uint32_t* reach,
                     uint32_t n_reach, int emulate_mem, uint64_t* stats7);
void orc_lists_free(orc_lists* r);
orc_game* orc_gamebuild(uint32_t n_states, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl, uint32_t n_pairs,
                        const uint32_t* pairs);
void orc_game_free(orc_game* g);
uint32_t orc_signature_refinement(uint32_t n, const uint32_t* lrow, const uint32_t* ldst, const int32_t* llbl, uint32_t* part);
}

namespace {

int failures = 0;
void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        failures++;
    }
}

struct Lts {
    uint32_t n;
    std::vector<uint32_t> row, dst;
    std::vector<int32_t> lbl;
    geq_lts view() const { return geq_lts{n, (uint32_t)dst.size(), row.data(), dst.data(), lbl.data()}; }
};

// label-sorted CSR of a random LTS
Lts random_lts(uint32_t n, uint32_t m, int labels, std::mt19937& rng) {
    std::vector<std::vector<std::pair<int32_t, uint32_t>>> adj(n);
    for (uint32_t k = 0; k < m; k++) adj[rng() % n].push_back({(int32_t)(rng() % labels), (uint32_t)(rng() % n)});
    Lts L{n, {0}, {}, {}};
    for (auto& a : adj) {
        std::stable_sort(a.begin(), a.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (auto& [l, t] : a) {
            L.dst.push_back(t);
            L.lbl.push_back(l);
        }
        L.row.push_back((uint32_t)L.dst.size());
    }
    return L;
}

void check_abi_without_device(std::mt19937& rng) {
    geq_ctx* ctx = nullptr;
    const geq_status st = geq_ctx_create(0, &ctx);
    expect(st == GEQ_ERR_NO_ADAPTER || st == GEQ_OK, "ctx_create reports no adapter without a device");
    expect(std::strlen(geq_last_error()) > 0 || st == GEQ_OK, "an error message");
    if (ctx) geq_ctx_destroy(ctx);
    expect(geq_ctx_create(0, nullptr) == GEQ_ERR_INVALID_ARG, "ctx_create(NULL)");
    expect(geq_solve(nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr) == GEQ_ERR_INVALID_ARG, "solve(NULL)");
    geq_energies e{};
    geq_energies_free(&e);
    geq_energies_free(nullptr);
    for (int rep = 0; rep < 20; rep++) {
        const uint32_t n = 2 + rng() % 300, m = rng() % (4 * n);
        Lts L = random_lts(n, m, 1 + rng() % 8, rng);
        const geq_lts v = L.view();
        std::vector<uint32_t> part(n), part_orc(n);
        uint32_t classes = 0;
        expect(geq_signature_refinement(&v, part.data(), &classes) == GEQ_OK, "signature_refinement");
        const uint32_t c2 = orc_signature_refinement(n, L.row.data(), L.dst.data(), L.lbl.data(), part_orc.data());
        expect(classes == c2 && part == part_orc, "partition equals the oracle's");
        std::vector<uint32_t> orow(n + 1), odst(m + 1);
        std::vector<int32_t> olbl(m + 1);
        uint32_t om = 0, c3 = 0;
        expect(geq_bisimilar_minimize(&v, part.data(), &c3, orow.data(), odst.data(), olbl.data(), &om) == GEQ_OK,
               "bisimilar_minimize");
        expect(c3 == classes && om <= m && orow[c3] == om, "quotient shape");
        // malformed input is refused, not read out of bounds
        std::vector<uint32_t> bad_row = L.row;
        bad_row[n] = m + 5;
        geq_lts bad{n, m, bad_row.data(), L.dst.data(), L.lbl.data()};
        expect(geq_signature_refinement(&bad, part.data(), &classes) == GEQ_ERR_INVALID_ARG, "malformed LTS refused");
    }
    // standard_reach on a random game graph
    const uint32_t V = 500;
    std::vector<uint32_t> row(V + 1, 0), col;
    std::vector<uint8_t> atk(V);
    for (uint32_t v = 0; v < V; v++) {
        atk[v] = rng() & 1;
        const uint32_t d = rng() % 4;
        for (uint32_t k = 0; k < d; k++) col.push_back(rng() % V);
        row[v + 1] = (uint32_t)col.size();
    }
    std::vector<uint32_t> wt(col.size(), 0), out(V);
    geq_graph g{{6, 3}, V, (uint32_t)col.size(), row.data(), col.data(), wt.data(), atk.data(), nullptr, nullptr};
    uint32_t nr = 0;
    expect(geq_standard_reach(&g, out.data(), &nr) == GEQ_OK, "standard_reach");
    uint32_t want = 0;
    for (uint32_t v = 0; v < V; v++) want += (!atk[v] && row[v] == row[v + 1]) ? 1u : 0u;
    expect(nr == want, "standard_reach count");
}

void check_oracle(std::mt19937& rng) {
    // packing known answers (src/energy.rs:846-871)
    const uint32_t vals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 13, 14, 15};
    uint32_t w[32] = {0};
    expect(orc_pack_energy(12, 15, vals, 12, w) == 0 && w[0] == 0x76543210u && w[1] == 0xfed8u, "pack_energy KAT");
    uint32_t back[12];
    orc_unpack_energy(12, 15, w, back);
    expect(std::memcmp(back, vals, sizeof vals) == 0, "unpack_energy KAT");
    const int32_t upd[12] = {0, -1, 1, 2, 3, 4, 5, 6, 9, 10, 11, 12};
    uint32_t uw[32] = {0};
    expect(orc_pack_update(12, 3, upd, 12, uw) == 0 && uw[0] == 0x76543210u && uw[1] == 0xdcbau, "pack_update KAT");
    uint32_t info[6], e1[32] = {0xffffffffu}, u1[32] = {1}, o1[32] = {0};
    orc_conf_info(1, 0xffffffffu, info);
    orc_inv_update(1, 0xffffffffu, e1, u1, o1);
    expect(o1[0] == 0, "ONE_DIMENSION decrement wraps");
    // the fixpoint on random games of several confs
    const uint32_t confs[5][2] = {{6, 3}, {8, 2}, {1, 0xffffffffu}, {20, 3}, {10, 70000}};
    for (const auto& c : confs) {
        orc_conf_info(c[0], c[1], info);
        const uint32_t usz = info[3], ubits = info[4];
        for (int rep = 0; rep < 3; rep++) {
            const uint32_t V = 50 + rng() % 400;
            std::vector<uint32_t> row(V + 1, 0), col, wt, reach;
            std::vector<uint8_t> atk(V);
            for (uint32_t v = 0; v < V; v++) {
                atk[v] = rng() & 1;
                const uint32_t d = rng() % 6;
                for (uint32_t k = 0; k < d; k++) {
                    col.push_back(rng() % V);
                    std::vector<uint32_t> u(usz, 0);
                    const uint32_t i = rng() % c[0], r = rng() % 10;
                    const uint32_t f = r < 5 ? 0u : r < 8 ? 1u : 2u + rng() % c[0];
                    u[(i * ubits) / 32] |= f << ((i * ubits) % 32);
                    wt.insert(wt.end(), u.begin(), u.end());
                }
                row[v + 1] = (uint32_t)col.size();
                if (!atk[v] && d == 0) reach.push_back(v);
            }
            std::vector<uint32_t> ioff(V + 1, 0);
            for (uint32_t v = 0, k = 0; v < V; v++) {
                if (k < reach.size() && reach[k] == v) k++;
                ioff[v + 1] = k;
            }
            std::vector<uint32_t> idata((size_t)reach.size() * info[0] + 1, 0);
            uint64_t stats[7];
            orc_lists* r = orc_solve(c[0], c[1], V, row.data(), col.data(), wt.data(), atk.data(), ioff.data(), idata.data(),
                                     reach.data(), (uint32_t)reach.size(), rep == 2, stats);
            expect(r != nullptr && r->n_nodes == V, "oracle solve");
            orc_lists_free(r);
        }
    }
    // gamebuild of a random LTS, all pairs of the first states
    Lts L = random_lts(60, 180, 4, rng);
    std::vector<uint32_t> pairs;
    for (uint32_t p = 0; p < 6; p++)
        for (uint32_t q = 0; q < 6; q++)
            if (p != q) {
                pairs.push_back(p);
                pairs.push_back(q);
            }
    orc_game* g = orc_gamebuild(L.n, L.row.data(), L.dst.data(), L.lbl.data(), (uint32_t)pairs.size() / 2, pairs.data());
    expect(g && g->V >= pairs.size() / 2, "oracle gamebuild");
    orc_game_free(g);
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    check_abi_without_device(rng);
    check_oracle(rng);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
