// Strong-bisimulation minimisation on the device: the signature refinement of
// TransitionSystem::signature_refinement (src/lib.rs:444-478, after Blom and
// Orzan), one device pass per refinement round.
//
//   signature(s) = the sorted set of (label, class of target) over s's transitions
//
// Per round: keys (label, class[target]) per transition; a segmented sort
// orders every state's keys (one thread per short row, one block per long row); one thread per state drops duplicates and
// hashes the run; a hash table with full comparison of the runs finds, for
// every state, the first state with the same signature (atomicMin); a scan of
// the "I am first" flags numbers the classes in order of first occurrence,
// exactly as the reference's sigmap does.  Rounds repeat until the class count
// stops changing.
#include <string>
#include <vector>

#include "runtime.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace geq {
void quotient_lts(const geq_lts* lts, const uint32_t* partition, uint32_t c, uint32_t* out_row, uint32_t* out_targets,
                  int32_t* out_labels, uint32_t* out_n_transitions);
}

using namespace geq;

namespace {

constexpr unsigned long long SIG_EMPTY = ~0ull;

__global__ void k_sig_keys(const uint32_t* dst, const int32_t* lbl, const uint32_t* part, uint32_t m, unsigned long long* keys) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) keys[k] = ((unsigned long long)((uint32_t)lbl[k] ^ 0x80000000u) << 32) | part[dst[k]];
}

// drop repeated keys of each state's sorted run; hash the run
__global__ void k_sig_dedup(const uint32_t* row, unsigned long long* keys, uint32_t n, uint32_t* len, uint32_t* hash) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t b = row[s], e = row[s + 1];
    uint32_t l = 0;
    unsigned long long h = 1469598103934665603ull;
    for (uint32_t k = b; k < e; k++) {
        const unsigned long long x = keys[k];
        if (l && keys[b + l - 1] == x) continue;
        keys[b + l] = x;
        l++;
        h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0x100000001b3ull;
    }
    h ^= l;
    len[s] = l;
    uint32_t h32 = (uint32_t)(h ^ (h >> 32));
    hash[s] = h32 == 0xffffffffu ? 0x7fffffffu : h32;
}

__device__ bool sig_equal(const uint32_t* row, const unsigned long long* keys, const uint32_t* len, uint32_t a, uint32_t b) {
    const uint32_t l = len[a];
    if (len[b] != l) return false;
    const unsigned long long* x = keys + row[a];
    const unsigned long long* y = keys + row[b];
    for (uint32_t k = 0; k < l; k++)
        if (x[k] != y[k]) return false;
    return true;
}

// every state finds the table slot of its signature; keymin[slot] = first state with it
__global__ void k_sig_insert(const uint32_t* row, const unsigned long long* keys, const uint32_t* len, const uint32_t* hash,
                             uint32_t n, unsigned long long* table, uint32_t* keymin, uint32_t mask, uint32_t* slot_of) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t h = hash[s];
    uint32_t i = h & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        unsigned long long e = __hip_atomic_load(&table[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e == SIG_EMPTY) {
            const unsigned long long mine = ((unsigned long long)h << 32) | s;
            const unsigned long long old = atomicCAS(&table[i], SIG_EMPTY, mine);
            if (old == SIG_EMPTY) {
                atomicMin(&keymin[i], s);
                slot_of[s] = i;
                return;
            }
            e = old;
        }
        if ((uint32_t)(e >> 32) == h && sig_equal(row, keys, len, (uint32_t)e, s)) {
            atomicMin(&keymin[i], s);
            slot_of[s] = i;
            return;
        }
        i = (i + 1) & mask;
    }
}

__global__ void k_sig_first(const uint32_t* slot_of, const uint32_t* keymin, uint32_t n, uint32_t* rep, uint32_t* flag) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) {
        const uint32_t r = keymin[slot_of[s]];
        rep[s] = r;
        flag[s] = r == s ? 1u : 0u;
    } else if (s == n) {
        flag[n] = 0u;
    }
}

__global__ void k_sig_assign(const uint32_t* rep, const uint32_t* cls, uint32_t n, uint32_t* part) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) part[s] = cls[rep[s]];
}

__global__ void k_fill64(unsigned long long* a, uint32_t n, unsigned long long v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill32(uint32_t* a, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

geq_status refine(geq_ctx* ctx, const geq_lts* lts, uint32_t* partition, uint32_t* n_classes, uint32_t* rounds) {
    const uint32_t n = lts->n_states, m = lts->n_transitions;
    hipStream_t s = ctx->stream;
    const uint32_t tb = 256;
    DevBuf<uint32_t> d_row, d_dst, d_part, d_len, d_hash, d_keymin, d_slot, d_rep, d_flag, d_cls, scan_tmp;
    DevBuf<int32_t> d_lbl;
    DevBuf<unsigned long long> keys, table;
    SegSortPlan plan;
    uint32_t cap = 16;
    while (cap < 2 * n + 2) cap <<= 1;
    hipError_t e;
    if ((e = d_row.alloc(n + 1)) != hipSuccess || (e = d_dst.alloc(std::max(m, 1u))) != hipSuccess ||
        (e = d_lbl.alloc(std::max(m, 1u))) != hipSuccess || (e = d_part.alloc(n)) != hipSuccess || (e = d_len.alloc(n)) != hipSuccess ||
        (e = d_hash.alloc(n)) != hipSuccess || (e = d_slot.alloc(n)) != hipSuccess || (e = d_rep.alloc(n)) != hipSuccess ||
        (e = d_flag.alloc(n + 1)) != hipSuccess || (e = d_cls.alloc(n + 1)) != hipSuccess ||
        (e = keys.alloc(std::max(m, 1u))) != hipSuccess ||
        (e = table.alloc(cap)) != hipSuccess || (e = d_keymin.alloc(cap)) != hipSuccess)
        return fail(GEQ_ERR_OUT_OF_MEMORY, std::string("bisimulation buffers: ") + hipGetErrorString(e));
    GEQ_HIP(hipMemcpyAsync(d_row.p, lts->row_offsets, 4ull * (n + 1), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    if (m) {
        GEQ_HIP(hipMemcpyAsync(d_dst.p, lts->targets, 4ull * m, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipMemcpyAsync(d_lbl.p, lts->labels, 4ull * m, hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    }
    k_fill32<<<(n + tb - 1) / tb, tb, 0, s>>>(d_part.p, n, 0u);
    geq_status pst = seg_sort_plan(lts->row_offsets, n, plan, s);
    if (pst != GEQ_OK) return pst;
    uint32_t prev = 0, cnt = 1, r = 0;
    while (prev != cnt) {
        prev = cnt;
        r++;
        if (m) {
            k_sig_keys<<<(m + tb - 1) / tb, tb, 0, s>>>(d_dst.p, d_lbl.p, d_part.p, m, keys.p);
            geq_status sst = seg_sort_u64(d_row.p, n, keys.p, plan, s);
            if (sst != GEQ_OK) return sst;
        }
        k_sig_dedup<<<(n + tb - 1) / tb, tb, 0, s>>>(d_row.p, keys.p, n, d_len.p, d_hash.p);
        k_fill64<<<(cap + tb - 1) / tb, tb, 0, s>>>(table.p, cap, SIG_EMPTY);
        k_fill32<<<(cap + tb - 1) / tb, tb, 0, s>>>(d_keymin.p, cap, 0xffffffffu);
        k_sig_insert<<<(n + tb - 1) / tb, tb, 0, s>>>(d_row.p, keys.p, d_len.p, d_hash.p, n, table.p, d_keymin.p, cap - 1, d_slot.p);
        k_sig_first<<<(n + 1 + tb - 1) / tb, tb, 0, s>>>(d_slot.p, d_keymin.p, n, d_rep.p, d_flag.p);
        geq_status cst = exclusive_scan_dev<uint32_t, SumU32>(d_flag.p, d_cls.p, n, scan_tmp, s, SumU32{});
        if (cst != GEQ_OK) return cst;
        k_sig_assign<<<(n + tb - 1) / tb, tb, 0, s>>>(d_rep.p, d_cls.p, n, d_part.p);
        GEQ_HIP(hipMemcpyAsync(&cnt, d_cls.p + n, 4, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
        GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    }
    GEQ_HIP(hipMemcpyAsync(partition, d_part.p, 4ull * n, hipMemcpyDeviceToHost, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    *n_classes = cnt;
    if (rounds) *rounds = r;
    return GEQ_OK;
}

geq_status check_lts(const geq_lts* lts) {
    if (!lts || !lts->row_offsets || (lts->n_transitions && (!lts->targets || !lts->labels)))
        return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    const uint32_t n = lts->n_states;
    if (lts->row_offsets[0] != 0 || lts->row_offsets[n] != lts->n_transitions)
        return fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets malformed");
    for (uint32_t st = 0; st < n; st++) {
        if (lts->row_offsets[st + 1] < lts->row_offsets[st]) return fail(GEQ_ERR_INVALID_ARG, "LTS row_offsets not monotone");
        for (uint32_t k = lts->row_offsets[st]; k < lts->row_offsets[st + 1]; k++)
            if (lts->targets[k] >= n) return fail(GEQ_ERR_INVALID_ARG, "LTS target out of range");
    }
    if (n >= (1u << 30)) return fail(GEQ_ERR_INVALID_ARG, "at most 2^30 states");
    return GEQ_OK;
}

}  // namespace

extern "C" {

geq_status geq_signature_refinement_gpu(geq_ctx* ctx, const geq_lts* lts, uint32_t* partition, uint32_t* n_classes,
                                        uint32_t* rounds) {
    if (!ctx || !partition || !n_classes) return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    geq_status st = check_lts(lts);
    if (st != GEQ_OK) return st;
    (void)hipSetDevice(ctx->device);
    if (lts->n_states == 0) {
        *n_classes = 0;
        if (rounds) *rounds = 0;
        return GEQ_OK;
    }
    return refine(ctx, lts, partition, n_classes, rounds);
}

geq_status geq_bisimilar_minimize_gpu(geq_ctx* ctx, const geq_lts* lts, uint32_t* partition, uint32_t* n_classes,
                                      uint32_t* out_row, uint32_t* out_targets, int32_t* out_labels, uint32_t* out_n_transitions) {
    if (!out_row || !out_n_transitions || (lts && lts->n_transitions && (!out_targets || !out_labels)))
        return fail(GEQ_ERR_INVALID_ARG, "NULL argument");
    geq_status st = geq_signature_refinement_gpu(ctx, lts, partition, n_classes, nullptr);
    if (st != GEQ_OK) return st;
    quotient_lts(lts, partition, *n_classes, out_row, out_targets, out_labels, out_n_transitions);
    return GEQ_OK;
}

}  // extern "C"
