// Host-side LTS utilities of the C ABI: strong-bisimulation minimisation by
// signature refinement (TransitionSystem::{signature_refinement,
// bisimilar_minimize}, src/lib.rs:413-478).
//
// Classes are numbered in order of first occurrence over the processes, as the
// reference's sigmap does, so the partition and the minimised LTS are the
// reference's exactly.  Signatures are kept as sorted (label, class) runs in
// one flat buffer and deduplicated with an open-addressing table of hashes.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpuequiv.h"

namespace geq {
geq_status fail(geq_status s, const std::string& msg);
// quotient of an LTS by a partition with classes numbered by first occurrence
void quotient_lts(const geq_lts* lts, const uint32_t* partition, uint32_t c, uint32_t* out_row, uint32_t* out_targets,
                  int32_t* out_labels, uint32_t* out_n_transitions);
}

namespace {

struct SigTable {
    std::vector<uint64_t> keys;     // hash << 32 | class id + 1, 0 = empty
    uint64_t mask = 0;
    explicit SigTable(size_t n) {
        size_t cap = 16;
        while (cap < 2 * n + 2) cap <<= 1;
        keys.assign(cap, 0);
        mask = cap - 1;
    }
};

uint64_t hash_sig(const uint64_t* a, size_t n) {
    uint64_t h = 1469598103934665603ull ^ n;
    for (size_t i = 0; i < n; i++) {
        h ^= a[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0x100000001b3ull;
    }
    return h;
}

}  // namespace

extern "C" {

geq_status geq_signature_refinement(const geq_lts* lts, uint32_t* partition, uint32_t* n_classes) {
    if (!lts || !partition || !n_classes || !lts->row_offsets) return geq::fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    const uint32_t n = lts->n_states, m = lts->n_transitions;
    // the CSR shape first: every row range must lie inside the transition arrays
    if (m && (!lts->targets || !lts->labels)) return geq::fail(GEQ_ERR_INVALID_ARG, "missing LTS transition arrays");
    if (lts->row_offsets[0] != 0 || lts->row_offsets[n] != m)
        return geq::fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets must start at 0 and end at n_transitions");
    for (uint32_t s = 0; s < n; s++)
        if (lts->row_offsets[s + 1] < lts->row_offsets[s]) return geq::fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets not monotone");
    for (uint32_t k = 0; k < m; k++)
        if (lts->targets[k] >= n) return geq::fail(GEQ_ERR_INVALID_ARG, "LTS target out of range");
    std::vector<uint32_t> part(n, 0), next(n);
    std::vector<uint64_t> sig;                 // (label << 32 | class) runs, per state sorted + unique
    std::vector<uint32_t> off(n + 1);
    uint32_t prev = 0, cnt = 1;
    while (prev != cnt) {
        prev = cnt;
        sig.clear();
        for (uint32_t s = 0; s < n; s++) {
            off[s] = (uint32_t)sig.size();
            for (uint32_t k = lts->row_offsets[s]; k < lts->row_offsets[s + 1]; k++)
                sig.push_back(((uint64_t)(uint32_t)(lts->labels[k] ^ 0x80000000) << 32) | part[lts->targets[k]]);
            std::sort(sig.begin() + off[s], sig.end());
            sig.erase(std::unique(sig.begin() + off[s], sig.end()), sig.end());
        }
        off[n] = (uint32_t)sig.size();
        SigTable t(n);
        std::vector<uint32_t> rep;            // first state of every class
        cnt = 0;
        for (uint32_t s = 0; s < n; s++) {
            const uint64_t* a = sig.data() + off[s];
            const size_t len = off[s + 1] - off[s];
            const uint64_t h = hash_sig(a, len);
            uint64_t i = h & t.mask;
            while (true) {
                const uint64_t e = t.keys[i];
                if (e == 0) {
                    t.keys[i] = (h & ~0xffffffffull) | (uint64_t)(cnt + 1);
                    rep.push_back(s);
                    next[s] = cnt++;
                    break;
                }
                if ((e & ~0xffffffffull) == (h & ~0xffffffffull)) {
                    const uint32_t c = (uint32_t)(e & 0xffffffffull) - 1;
                    const uint32_t r = rep[c];
                    const size_t rl = off[r + 1] - off[r];
                    if (rl == len && std::equal(a, a + len, sig.data() + off[r])) {
                        next[s] = c;
                        break;
                    }
                }
                i = (i + 1) & t.mask;
            }
        }
        part.swap(next);
    }
    if (n) std::memcpy(partition, part.data(), 4ull * n);
    *n_classes = n ? cnt : 0;
    return GEQ_OK;
}

}  // extern "C"

// The minimised LTS: class c takes the transitions of its first process, with
// targets mapped to classes, then each row is stably sorted by label.
void geq::quotient_lts(const geq_lts* lts, const uint32_t* partition, uint32_t c, uint32_t* out_row, uint32_t* out_targets,
                       int32_t* out_labels, uint32_t* out_n_transitions) {
    const uint32_t n = lts->n_states;
    std::vector<int64_t> first(c, -1);
    for (uint32_t s = 0; s < n; s++)
        if (first[partition[s]] < 0) first[partition[s]] = s;
    uint32_t k = 0;
    std::vector<std::pair<int32_t, uint32_t>> row;
    for (uint32_t cl = 0; cl < c; cl++) {
        out_row[cl] = k;
        const uint32_t s = (uint32_t)first[cl];
        row.clear();
        for (uint32_t e = lts->row_offsets[s]; e < lts->row_offsets[s + 1]; e++)
            row.push_back({lts->labels[e], partition[lts->targets[e]]});
        std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& t : row) {
            out_targets[k] = t.second;
            out_labels[k] = t.first;
            k++;
        }
    }
    out_row[c] = k;
    *out_n_transitions = k;
}

extern "C" geq_status geq_bisimilar_minimize(const geq_lts* lts, uint32_t* partition, uint32_t* n_classes, uint32_t* out_row,
                                             uint32_t* out_targets, int32_t* out_labels, uint32_t* out_n_transitions) {
    geq_status st = geq_signature_refinement(lts, partition, n_classes);
    if (st != GEQ_OK) return st;
    geq::quotient_lts(lts, partition, *n_classes, out_row, out_targets, out_labels, out_n_transitions);
    return GEQ_OK;
}
