// Gamebuild on the device: labelled transition system -> spectroscopy energy
// game (GameBuild::build_internal, src/gamebuild.rs:229-265, with the position
// rules of src/gamebuild/gamepos.rs:17-198).
//
// The reference grows the game by a FIFO walk over positions in creation order:
// node idx's successors, in slot order, become new nodes in order of first
// occurrence.  Here every BFS level (the contiguous id range created by the
// previous level) is expanded at once:
//   1. k_bounds   - per frontier position an upper bound of its successor
//                   count and of the set words they need;
//   2. scans      - exclusive sums give every parent a private region, so the
//                   candidate index order is exactly (parent, slot) order;
//   3. k_generate - each parent writes its successor candidates in the
//                   reference's slot order (observation moves in transition
//                   order, then the conjunction challenges);
//   4. k_insert   - a device hash table with full-equality checks finds known
//                   positions or claims a pending entry; among duplicates of a
//                   new position the smallest candidate index wins;
//   5. k_mark + scan - winners get ranks, i.e. new ids in first-occurrence
//                   order, which is the reference's numbering;
//   6. k_finalize / k_edges - node records, hash entries, CSR rows, columns,
//                   weights (k_level_out: the edges, rows and the next level's
//                   bounds in one launch when no per-job counts are kept).
#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "gamebuild.hpp"
#include "scan.hpp"

namespace geq {
namespace gb {

constexpr uint32_t K_ATTACK = 0, K_CLAUSE = 1, K_DEFEND = 2;
constexpr uint32_t INVALID = 0xffffffffu;
constexpr uint32_t PENDING = 0x80000000u;
constexpr unsigned long long EMPTY = EMPTY_ENTRY;

// Updates in the STANDARD packing (4-bit fields): observation [-1], challenge
// [0,-1], clause [Min(4)] and [Min(5),0,0,0,0,-1], answer [0,0,0,Min(3)],
// revival [Min(3)] (gamepos.rs:18-24, :139-150, :180-186).
constexpr uint32_t W_OBS = 0x1, W_CHAL = 0x10, W_CL1 = 0x5, W_CL2 = 0x100006, W_ANS = 0x4000, W_REV = 0x4;

__device__ __forceinline__ uint32_t mix(uint32_t h, uint32_t v) {
    h ^= v + 0x9e3779b9u + (h << 6) + (h >> 2);
    return h * 0x85ebca6bu;
}

// Position hashes key the position table only: every match is confirmed by
// comparing the positions, so no id depends on them.  A position's hash is
// fmix32(head + S_Q + S_X): the head mixes kind, p and the two set sizes, and
// each set adds one term per member (fmix32 of the member, salted apart for
// Q and Q*).  The terms of a set sum in any order, so lanes or a block add a
// set's part by a reduction, and all candidates of a label group share it.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
constexpr uint32_t SALT_Q = 0x2545f491u, SALT_X = 0x68e31da4u;
__device__ __forceinline__ uint32_t set_term(uint32_t v, uint32_t salt) { return fmix32(v * 0x9e3779b1u + salt); }
__device__ __forceinline__ uint32_t hash_head(uint32_t kind, uint32_t p, uint32_t ql, uint32_t xl) {
    return mix(mix(mix(mix(0x811c9dc5u, kind), p), ql), xl);
}
__device__ __forceinline__ uint32_t hash_of(uint32_t head, uint32_t sets) {
    const uint32_t h = fmix32(head + sets);
    return h == 0xffffffffu ? 0x7fffffffu : h;
}

__device__ uint32_t pos_hash(uint32_t kind, uint32_t p, const uint32_t* __restrict__ spool, uint32_t qo, uint32_t ql,
                             uint32_t xo, uint32_t xl) {
    uint32_t sum = 0;
    for (uint32_t k = 0; k < ql; k++) sum += set_term(spool[qo + k], SALT_Q);
    for (uint32_t k = 0; k < xl; k++) sum += set_term(spool[xo + k], SALT_X);
    return hash_of(hash_head(kind, p, ql, xl), sum);
}

// The same hash of a position with set Q in LDS (entry k at q[k * st]) and no Q*.
__device__ uint32_t pos_hash_lds(uint32_t kind, uint32_t p, const uint32_t* q, uint32_t ql, uint32_t st) {
    uint32_t sum = 0;
    for (uint32_t k = 0; k < ql; k++) sum += set_term(q[k * st], SALT_Q);
    return hash_of(hash_head(kind, p, ql, 0), sum);
}

// the sets of len words at ao and bo in the pool are equal; four words of each
// per step, read at clamped (in-range) indices so the eight loads are issued
// together instead of one compare waiting on the next pair (eight words a step
// made the large levels' inserts slower: most sets have one or two members)
__device__ bool sets_equal(const uint32_t* __restrict__ spool, uint32_t ao, uint32_t bo, uint32_t len) {
    for (uint32_t k = 0; k < len; k += 4) {
        uint32_t d = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            const uint32_t x = min(k + u, len - 1u);
            const uint32_t v = spool[ao + x] ^ spool[bo + x];
            d |= k + u < len ? v : 0u;
        }
        if (d) return false;
    }
    return true;
}

// label range [lo, hi) of `a` in state s's label-sorted transitions
__device__ __forceinline__ void label_range(const LtsArgs& L, uint32_t s, int32_t a, uint32_t& lo, uint32_t& hi) {
    if (L.en_mask && !((L.en_mask[s] >> (a & 63)) & 1ull)) {   // s has no a-transition
        lo = hi = 0;
        return;
    }
    uint32_t b = L.row[s], e = L.row[s + 1];
    uint32_t l = b, r = e;
    while (l < r) {
        uint32_t m = (l + r) >> 1;
        if (L.lbl[m] < a) l = m + 1; else r = m;
    }
    lo = l;
    r = e;
    while (l < r) {
        uint32_t m = (l + r) >> 1;
        if (L.lbl[m] <= a) l = m + 1; else r = m;
    }
    hi = l;
}

// compare_enabled (src/gamebuild.rs:23-98) on the sorted distinct label lists
__device__ void compare_enabled(const LtsArgs& L, uint32_t p, uint32_t q, bool& subset, bool& superset) {
    subset = true;
    superset = true;
    if (p == q) return;
    if (L.en_mask) {        // the label bit sets: two words instead of the list walk
        const unsigned long long mp = L.en_mask[p], mq = L.en_mask[q];
        subset = (mq & ~mp) == 0ull;
        superset = (mp & ~mq) == 0ull;
        return;
    }
    uint32_t pi = L.en_row[p], pe = L.en_row[p + 1], qi = L.en_row[q], qe = L.en_row[q + 1];
    while (subset || superset) {
        if (pi == pe && qi == qe) break;
        if (pi == pe) { subset = false; break; }
        if (qi == qe) { superset = false; break; }
        int32_t pa = L.en_lbl[pi], qa = L.en_lbl[qi];
        if (pa == qa) { pi++; qi++; }
        else if (pa < qa) { superset = false; pi++; }
        else { subset = false; qi++; }
    }
}

// in-place ascending sort of a[0..n) (insertion sort for short runs, heap sort otherwise)
__device__ void sort_u32(uint32_t* a, uint32_t n) {
    if (n <= 24) {
        for (uint32_t i = 1; i < n; i++) {
            uint32_t x = a[i], j = i;
            while (j > 0 && a[j - 1] > x) { a[j] = a[j - 1]; j--; }
            a[j] = x;
        }
        return;
    }
    auto sift = [&](uint32_t root, uint32_t end) {
        while (2 * root + 1 < end) {
            uint32_t c = 2 * root + 1;
            if (c + 1 < end && a[c] < a[c + 1]) c++;
            if (a[root] >= a[c]) return;
            uint32_t t = a[root]; a[root] = a[c]; a[c] = t;
            root = c;
        }
    };
    for (uint32_t i = n / 2; i-- > 0;) sift(i, n);
    for (uint32_t end = n - 1; end > 0; end--) {
        uint32_t t = a[0]; a[0] = a[end]; a[end] = t;
        sift(0, end);
    }
}

__device__ __forceinline__ bool contains_sorted(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t l = 0, r = n;
    while (l < r) {
        uint32_t m = (l + r) >> 1;
        if (a[m] < x) l = m + 1; else r = m;
    }
    return l < n && a[l] == x;
}

__device__ __forceinline__ uint32_t deg_of(const LtsArgs& L, uint32_t s) { return L.row[s + 1] - L.row[s]; }

// The bounds of position u, and whether it is listed for a kernel of its own
// (`heavy`; `wave`: k_generate_wave, else k_generate_heavy).  An attack
// position with a non-empty Q is listed (none when heavy_min is 0):
//   - under k_generate (GEN_BY_THREAD), when its serial walk (one binary search
//     per member of Q and label group of p, the successor sets gathered and
//     sorted) weighs at least heavy_min and its sets fit the block's LDS;
//   - under k_generate_sub (GEN_BY_SUB), when it is past that kernel's LDS
//     limits: for a wave when it fits the wave's, else for a block when it
//     fits the block's (past both: the eight-lane kernel's first lane walks
//     it); GEN_BY_WAVE / GEN_BY_BLOCK list every one that fits for a wave / block.
__device__ __forceinline__ uint2 bounds_of(const NodeArgs& N, const LtsArgs& L, uint32_t u, const HeavyArgs& H, bool& heavy,
                                           bool& wave) {
    const uint32_t kind = N.kind[u], p = N.p[u], ql = N.q_len[u], xl = N.x_len[u];
    uint32_t c = 0, s = 0;
    heavy = wave = false;
    if (L.job_dead && L.job_dead[L.job_of_state[p]]) {
        // a job past its position cap: no successors
    } else if (kind == K_CLAUSE) {
        c = 2; s = 1;
    } else if (kind == K_DEFEND) {
        c = ql + (xl ? 1u : 0u);
    } else if (ql == 0) {
        c = 1;
    } else {
        c = deg_of(L, p) + 4;
        // the members' degrees, eight members a step: their ids, then their rows,
        // each read at clamped (in-range) indices so the loads go out together
        uint32_t sd = 0;
        const uint32_t qo = N.q_off[u];
        for (uint32_t k = 0; k < ql; k += 8) {
            uint32_t m[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) m[j] = N.spool[qo + min(k + j, ql - 1u)];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t d = deg_of(L, m[j]);
                sd += k + j < ql ? d : 0u;
            }
        }
        s = sd + 4 * ql;
        const bool fits_block = sd <= HV_CAP && ql <= HV_CAP;
        const bool fits_wave = sd <= WAVE_TMAX && ql <= WAVE_QMAX;
        if (H.heavy_min == 0u) {
        } else if (H.mode == GEN_BY_THREAD) {
            const uint32_t ng = L.en_row[p + 1] - L.en_row[p];
            heavy = fits_block && (uint64_t)ql * (ng + 1u) + sd >= H.heavy_min;
        } else if (H.mode == GEN_BY_WAVE) {
            heavy = wave = fits_wave;
        } else if (H.mode == GEN_BY_BLOCK) {
            heavy = fits_block;
        } else if (sd > SUB_TMAX || ql > SUB_QMAX) {
            wave = fits_wave;
            heavy = fits_wave || fits_block;
        }
    }
    return make_uint2(c, s);
}

__global__ void k_bounds(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L, uint2* ub, HeavyArgs H) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_front) return;
    bool heavy = false, wave = false;
    // trailing zero so the exclusive scan yields the totals
    ub[i] = i == n_front ? make_uint2(0u, 0u) : bounds_of(N, L, lo + i, H, heavy, wave);
    if (i < n_front) {
        H.flag[i] = heavy ? 1 : 0;
        if (heavy) H.list[atomicAdd(H.count, 1u)] = i | (wave ? WAVE_TAG : 0u);
    }
}

// bound entry i of the next level (frontier of n_front positions from lo)
__device__ __forceinline__ void bounds_next_at(const NodeArgs& N, uint32_t lo, uint32_t n_front, uint32_t chunk, const LtsArgs& L,
                                               uint2* ub, const HeavyArgs& H, uint32_t i) {
    bool heavy = false, wave = false;
    ub[i] = i < n_front ? bounds_of(N, L, lo + i, H, heavy, wave) : make_uint2(0u, 0u);
    if (i < min(n_front, chunk)) {      // the flags and list cover the first chunk
        H.flag[i] = heavy ? 1 : 0;
        if (heavy) H.list[atomicAdd(H.count, 1u)] = i | (wave ? WAVE_TAG : 0u);
    }
}

__global__ void k_bounds_next(NodeArgs N, uint32_t lo, const uint32_t* n_new, uint32_t add, uint32_t n_slots, uint32_t chunk,
                              LtsArgs L, uint2* ub, HeavyArgs H) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slots) return;
    bounds_next_at(N, lo, *n_new + add, chunk, L, ub, H, i);
}

// Dynamic LDS: 0, or 4 * lts_words(...) bytes for a copy of the LTS, which
// the block then reads instead of global memory.  Each thread also keeps the
// successor set Q' of one label group in LDS while it has at most GEN_LDS
// entries (entry k of thread t at qbuf[k * GEN_BLOCK + t]): it is built
// sorted and duplicate-free there, probed and hashed there, and written to the
// set pool once.  Larger sets take the same steps in the set pool.
constexpr uint32_t GEN_BLOCK = 256, GEN_LDS = 16;

__device__ __forceinline__ void gen_walk(const NodeArgs& N, uint32_t lo, uint32_t i, const LtsArgs& L, const uint2* off, uint64_t spool_base,
                         Cand* cands, uint32_t* nsucc, uint32_t* qb, uint32_t st);

__global__ __launch_bounds__(GEN_BLOCK) void k_generate(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L,
                                                        const uint2* off, const uint8_t* heavy, uint64_t spool_base,
                                                        Cand* cands, uint32_t* nsucc) {
    extern __shared__ uint32_t lts_lds[];
    __shared__ uint32_t s_qbuf[GEN_LDS * GEN_BLOCK];
    uint32_t* const qb = s_qbuf + threadIdx.x;
    if (lts_words(L.n, L.m, L.n_en) <= LTS_LDS_WORDS) {   // the host then passed LDS for the copy
        uint32_t* r = lts_lds;
        uint32_t* d = r + (L.n + 1);
        int32_t* lb = reinterpret_cast<int32_t*>(d + L.m);
        uint32_t* er = reinterpret_cast<uint32_t*>(lb + L.m);
        int32_t* el = reinterpret_cast<int32_t*>(er + (L.n + 1));
        for (uint32_t k = threadIdx.x; k <= L.n; k += blockDim.x) {
            r[k] = L.row[k];
            er[k] = L.en_row[k];
        }
        for (uint32_t k = threadIdx.x; k < L.m; k += blockDim.x) {
            d[k] = L.dst[k];
            lb[k] = L.lbl[k];
        }
        for (uint32_t k = threadIdx.x; k < L.n_en; k += blockDim.x) el[k] = L.en_lbl[k];
        __syncthreads();
        L.row = r; L.dst = d; L.lbl = lb; L.en_row = er; L.en_lbl = el;   // en_mask stays in global memory
    }
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == n_front) nsucc[i] = 0;   // trailing zero: the exclusive scan yields the edge total
    if (i >= n_front || heavy[i]) return;   // k_generate_heavy writes the heavy positions
    gen_walk(N, lo, i, L, off, spool_base, cands, nsucc, qb, GEN_BLOCK);
}

// One position's successor candidates by one thread (the walk described above
// k_generate); qb: GEN_LDS words of LDS at stride st.
__device__ __forceinline__ void gen_walk(const NodeArgs& N, uint32_t lo, uint32_t i, const LtsArgs& L, const uint2* off, uint64_t spool_base,
                         Cand* cands, uint32_t* nsucc, uint32_t* qb, uint32_t st) {
    const uint32_t u = lo + i;
    const uint32_t kind = N.kind[u], p = N.p[u], qo = N.q_off[u], ql = N.q_len[u], xo = N.x_off[u], xl = N.x_len[u];
    uint32_t* spool = N.spool;
    const uint2 o0 = off[i];
    const uint32_t cbase = o0.x, cend = off[i + 1].x;
    const uint64_t sbase = spool_base + o0.y;
    uint32_t slot = 0;
    if (L.job_dead && L.job_dead[L.job_of_state[p]]) {   // k_bounds gave it no room
        nsucc[i] = 0;
        return;
    }
    auto emit = [&](uint32_t k, uint32_t pp, uint32_t a_off, uint32_t a_len, uint32_t b_off, uint32_t b_len, uint32_t w) {
        Cand c;
        c.kind_p = (k << 30) | pp;
        c.q_off = a_off; c.q_len = a_len; c.x_off = b_off; c.x_len = b_len;
        c.weight = w;
        c.hash = pos_hash(k, pp, spool, a_off, a_len, b_off, b_len);
        c.res = 0;
        c.parent = i;
        c.slot = slot;
        cands[cbase + slot] = c;
        slot++;
    };
    if (kind == K_CLAUSE) {
        // Clause(p, q) -> Attack(p, [q]) [Min(4)], Attack(q, [p]) [Min(5),0,0,0,0,-1]
        const uint32_t q = spool[qo];
        emit(K_ATTACK, p, qo, 1, 0, 0, W_CL1);
        spool[sbase] = p;
        emit(K_ATTACK, q, (uint32_t)sbase, 1, 0, 0, W_CL2);
    } else if (kind == K_DEFEND) {
        // Defend(p, Q, Q*) -> Clause(p, q) per q in Q [0,0,0,Min(3)], Attack(p, Q*) [Min(3)]
        for (uint32_t k = 0; k < ql; k++) emit(K_CLAUSE, p, qo + k, 1, 0, 0, W_ANS);
        if (xl) emit(K_ATTACK, p, xo, xl, 0, 0, W_REV);
    } else if (ql == 0) {
        emit(K_DEFEND, p, 0, 0, 0, 0, W_CHAL);        // Attack(p, {}) -> Defend(p, {}, {})
    } else {
        // observation moves, one label group of p's transitions at a time
        uint64_t sp = sbase;
        uint32_t sd = 0;
        for (uint32_t k = 0; k < ql; k++) sd += deg_of(L, spool[qo + k]);
        const uint32_t t_end = L.row[p + 1];
        uint32_t t = L.row[p];
        while (t < t_end) {
            const int32_t a = L.lbl[t];
            uint32_t g_end = t;
            while (g_end < t_end && L.lbl[g_end] == a) g_end++;
            // Q' = the union of the label-a successors of Q: sorted and
            // duplicate-free in LDS while it fits, else gathered in the pool
            uint32_t nq = 0, T = 0;
            bool spill = false;
            for (uint32_t k = 0; k < ql; k++) {
                uint32_t lo_q, hi_q;
                label_range(L, spool[qo + k], a, lo_q, hi_q);
                for (uint32_t x = lo_q; x < hi_q; x++) {
                    const uint32_t v = L.dst[x];
                    if (spill) {
                        spool[sp + T++] = v;
                        continue;
                    }
                    uint32_t j = nq;
                    while (j > 0 && qb[(j - 1) * st] > v) j--;
                    if (j > 0 && qb[(j - 1) * st] == v) continue;
                    if (nq == GEN_LDS) {          // full: continue in the pool
                        for (uint32_t y = 0; y < nq; y++) spool[sp + y] = qb[y * st];
                        T = nq;
                        spool[sp + T++] = v;
                        spill = true;
                        continue;
                    }
                    for (uint32_t y = nq; y > j; y--) qb[y * st] = qb[(y - 1) * st];
                    qb[j * st] = v;
                    nq++;
                }
            }
            if (!spill) {
                for (uint32_t y = 0; y < nq; y++) spool[sp + y] = qb[y * st];
                for (uint32_t tt = t; tt < g_end; tt++) {
                    const uint32_t pp = L.dst[tt];
                    bool in = false;
                    for (uint32_t y = 0; y < nq; y++) in |= qb[y * st] == pp;
                    if (!in) {
                        Cand c;
                        c.kind_p = (K_ATTACK << 30) | pp;
                        c.q_off = (uint32_t)sp; c.q_len = nq; c.x_off = 0; c.x_len = 0;
                        c.weight = W_OBS;
                        c.hash = pos_hash_lds(K_ATTACK, pp, qb, nq, st);
                        c.res = 0;
                        c.parent = i;
                        c.slot = slot;
                        cands[cbase + slot] = c;
                        slot++;
                    }
                }
                sp += nq;
            } else {
                uint32_t* qa = spool + sp;
                sort_u32(qa, T);
                uint32_t Tu = 0;
                for (uint32_t x = 0; x < T; x++)
                    if (Tu == 0 || qa[Tu - 1] != qa[x]) qa[Tu++] = qa[x];
                for (uint32_t tt = t; tt < g_end; tt++) {
                    const uint32_t pp = L.dst[tt];
                    if (!contains_sorted(qa, Tu, pp)) emit(K_ATTACK, pp, (uint32_t)sp, Tu, 0, 0, W_OBS);
                }
                sp += Tu;
            }
            t = g_end;
        }
        emit(K_DEFEND, p, qo, ql, 0, 0, W_CHAL);     // Q* = {}
        // conjunction challenges with partitioned Q* (gamepos.rs:329-375)
        // per q: subset / superset flags, in two registers for |Q| <= 32, else in the pool
        const uint64_t fl = sbase + sd + 3ull * ql;  // per q: bit0 subset, bit1 superset
        uint32_t sub_m = 0, sup_m = 0;
        for (uint32_t k = 0; k < ql; k++) {
            bool sub, sup;
            compare_enabled(L, p, spool[qo + k], sub, sup);
            if (ql <= 32) {
                sub_m |= (sub ? 1u : 0u) << k;
                sup_m |= (sup ? 1u : 0u) << k;
            } else {
                spool[fl + k] = (sub ? 1u : 0u) | (sup ? 2u : 0u);
            }
        }
        auto in_x = [&](uint32_t which, uint32_t k) -> bool {
            uint32_t f;
            if (ql <= 32) f = ((sub_m >> k) & 1u) | (((sup_m >> k) & 1u) << 1);
            else f = spool[fl + k];
            return which == 0 ? (f & 1u) : which == 1 ? (f & 2u) : (f == 3u);
        };
        uint64_t pb = sbase + sd;
        uint32_t emitted[3];
        uint32_t n_emitted = 0;
        for (uint32_t which = 0; which < 3; which++) {
            uint32_t cnt = 0;
            for (uint32_t k = 0; k < ql; k++) cnt += in_x(which, k) ? 1u : 0u;
            if (cnt == 0 || cnt == ql) continue;
            bool dup = false;
            for (uint32_t e = 0; e < n_emitted && !dup; e++) {
                bool same = true;
                for (uint32_t k = 0; k < ql && same; k++) same = in_x(which, k) == in_x(emitted[e], k);
                dup = same;
            }
            if (dup) continue;
            uint32_t a = 0, b = ql - cnt;
            for (uint32_t k = 0; k < ql; k++) {
                if (in_x(which, k)) spool[pb + b++] = spool[qo + k];
                else spool[pb + a++] = spool[qo + k];
            }
            emit(K_DEFEND, p, (uint32_t)pb, ql - cnt, (uint32_t)(pb + ql - cnt), cnt, W_CHAL);
            emitted[n_emitted++] = which;
            pb += ql;
        }
    }
    nsucc[i] = slot;
    for (uint32_t j = cbase + slot; j < cend; j++) cands[j].kind_p = INVALID;
}

namespace {

constexpr uint32_t HV_BLOCK = 256;


// Exclusive prefix sum of one value per thread over the block (threads in
// order); `total` gets the block's sum.  ws: HV_BLOCK / 64 words of LDS.
__device__ uint32_t block_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63u) ws[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t k = 0; k < HV_BLOCK / 64u; k++) {
        const uint32_t s = ws[k];
        pre += k < w ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return pre + x - v;
}

// Thread t's contiguous share [a, b) of n items (thread order = item order).
__device__ __forceinline__ void share(uint32_t n, uint32_t& a, uint32_t& b) {
    const uint32_t per = (n + HV_BLOCK - 1u) / HV_BLOCK;
    a = min(n, threadIdx.x * per);
    b = min(n, a + per);
}

// Ascending bitonic sort of a[0..n) in LDS, n a power of two, by the block.
__device__ void bitonic_lds(uint32_t* a, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t x = threadIdx.x; x < n; x += HV_BLOCK) {
                const uint32_t y = x ^ j;
                if (y > x) {
                    const uint32_t u = a[x], v = a[y];
                    if ((u > v) == ((x & k) == 0u)) {
                        a[x] = v;
                        a[y] = u;
                    }
                }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ bool in_part(uint32_t which, uint8_t f) {
    return which == 0u ? (f & 1u) != 0u : which == 1u ? (f & 2u) != 0u : f == 3u;
}

}  // namespace

// One block per heavy attack position (H.list): the same candidates, set-pool
// layout and hashes as k_generate's walk, with the work of each label group
// spread over the block: the members of Q search their label ranges in
// parallel, Q' is sorted and deduplicated in LDS, the group's transitions are
// tested and emitted in parallel (ranks from a block scan keep the slot order),
// and the conjunction challenges partition Q with block scans.
__global__ __launch_bounds__(HV_BLOCK) void k_generate_heavy(NodeArgs N, uint32_t lo, LtsArgs L, const uint2* off,
                                                             uint64_t spool_base, HeavyArgs H, Cand* cands, uint32_t* nsucc) {
    __shared__ uint32_t qs[HV_CAP];
    __shared__ uint8_t fl[HV_CAP];
    __shared__ uint32_t ws[HV_BLOCK / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t n_heavy = *H.count;
    uint32_t* spool = N.spool;
    for (uint32_t h = blockIdx.x; h < n_heavy; h += gridDim.x) {
        const uint32_t i = H.list[h];
        if (i & WAVE_TAG) continue;   // k_generate_wave's (the same for the whole block)
        const uint32_t u = lo + i;
        const uint32_t p = N.p[u], qo = N.q_off[u], ql = N.q_len[u];
        const uint2 o0 = off[i];
        const uint32_t cbase = o0.x, cend = off[i + 1].x;
        const uint64_t sbase = spool_base + o0.y;
        uint32_t a0, a1;
        share(ql, a0, a1);
        uint32_t part = 0;
        for (uint32_t k = a0; k < a1; k++) part += deg_of(L, spool[qo + k]);
        uint32_t sd;
        (void)block_scan(part, ws, sd);
        uint32_t slot = 0;
        uint64_t sp = sbase;
        // observation moves, one label group of p's transitions at a time (p's
        // distinct labels in order)
        for (uint32_t gi = L.en_row[p]; gi < L.en_row[p + 1]; gi++) {
            const int32_t a = L.en_lbl[gi];
            uint32_t t, g_end;
            label_range(L, p, a, t, g_end);
            uint32_t cnt = 0;
            for (uint32_t k = a0; k < a1; k++) {
                uint32_t x0, x1;
                label_range(L, spool[qo + k], a, x0, x1);
                cnt += x1 - x0;
            }
            uint32_t T;
            uint32_t w = block_scan(cnt, ws, T);   // T <= sd <= HV_CAP
            for (uint32_t k = a0; k < a1; k++) {
                uint32_t x0, x1;
                label_range(L, spool[qo + k], a, x0, x1);
                for (uint32_t x = x0; x < x1; x++) qs[w++] = L.dst[x];
            }
            uint32_t n2 = 1;
            while (n2 < T) n2 <<= 1;
            for (uint32_t x = T + tid; x < n2; x += HV_BLOCK) qs[x] = 0xffffffffu;
            __syncthreads();
            bitonic_lds(qs, n2);
            // Q' = the distinct values in order, written to the pool at sp
            uint32_t b0, b1;
            share(T, b0, b1);
            uint32_t nd = 0, part_sum = 0;
            for (uint32_t x = b0; x < b1; x++)
                if (x == 0u || qs[x] != qs[x - 1]) {
                    nd++;
                    part_sum += set_term(qs[x], SALT_Q);
                }
            uint32_t Tu, sq;
            uint32_t r = block_scan(nd, ws, Tu);
            (void)block_scan(part_sum, ws, sq);   // Q''s part of its candidates' hashes
            for (uint32_t x = b0; x < b1; x++)
                if (x == 0u || qs[x] != qs[x - 1]) spool[sp + r++] = qs[x];
            // an observation move to every target of the group outside Q', in transition order
            uint32_t c0, c1;
            share(g_end - t, c0, c1);
            uint32_t ne = 0;
            for (uint32_t x = c0; x < c1; x++) ne += contains_sorted(qs, T, L.dst[t + x]) ? 0u : 1u;
            uint32_t n_emit;
            uint32_t rr = block_scan(ne, ws, n_emit);
            for (uint32_t x = c0; x < c1; x++) {
                const uint32_t pp = L.dst[t + x];
                if (contains_sorted(qs, T, pp)) continue;
                Cand c;
                c.kind_p = (K_ATTACK << 30) | pp;
                c.q_off = (uint32_t)sp; c.q_len = Tu; c.x_off = 0; c.x_len = 0;
                c.weight = W_OBS;
                c.hash = hash_of(hash_head(K_ATTACK, pp, Tu, 0), sq);
                c.res = 0;
                c.parent = i;
                c.slot = slot + rr;
                cands[cbase + slot + rr] = c;
                rr++;
            }
            slot += n_emit;
            sp += Tu;
            __syncthreads();   // qs is rewritten by the next group
        }
        // Q and its enabledness flags (bit 0 subset, bit 1 superset) in LDS
        uint32_t q_sum = 0;
        for (uint32_t k = a0; k < a1; k++) {
            const uint32_t q = spool[qo + k];
            bool sub, sup;
            compare_enabled(L, p, q, sub, sup);
            qs[k] = q;
            fl[k] = (sub ? 1u : 0u) | (sup ? 2u : 0u);
            q_sum += set_term(q, SALT_Q);
        }
        uint32_t sq;
        (void)block_scan(q_sum, ws, sq);   // its barriers also order the writes above
        if (tid == 0) {                 // Defend(p, Q, {})
            Cand c;
            c.kind_p = (K_DEFEND << 30) | p;
            c.q_off = qo; c.q_len = ql; c.x_off = 0; c.x_len = 0;
            c.weight = W_CHAL;
            c.hash = hash_of(hash_head(K_DEFEND, p, ql, 0), sq);
            c.res = 0;
            c.parent = i;
            c.slot = slot;
            cands[cbase + slot] = c;
        }
        slot++;
        // conjunction challenges with partitioned Q* (gamepos.rs:329-375)
        uint64_t pb = sbase + sd;
        uint32_t emitted[3];
        uint32_t n_emitted = 0;
        for (uint32_t which = 0; which < 3u; which++) {
            uint32_t local = 0;
            for (uint32_t k = a0; k < a1; k++) local += in_part(which, fl[k]) ? 1u : 0u;
            uint32_t cnt;
            uint32_t rin = block_scan(local, ws, cnt);
            if (cnt == 0u || cnt == ql) continue;
            bool dup = false;
            for (uint32_t e = 0; e < n_emitted; e++) {
                uint32_t diff = 0;
                for (uint32_t k = a0; k < a1; k++) diff += in_part(which, fl[k]) != in_part(emitted[e], fl[k]) ? 1u : 0u;
                uint32_t n_diff;
                (void)block_scan(diff, ws, n_diff);
                dup |= n_diff == 0u;
            }
            if (dup) continue;
            // stable partition: the members outside the part, then those inside
            const uint32_t na = ql - cnt;
            uint32_t h_sum = 0;   // the members outside the part are Q, those inside Q*
            for (uint32_t k = a0; k < a1; k++) {
                if (in_part(which, fl[k])) {
                    spool[pb + na + rin++] = qs[k];
                    h_sum += set_term(qs[k], SALT_X);
                } else {
                    spool[pb + (k - rin)] = qs[k];
                    h_sum += set_term(qs[k], SALT_Q);
                }
            }
            uint32_t sqx;
            (void)block_scan(h_sum, ws, sqx);
            if (tid == 0) {
                Cand c;
                c.kind_p = (K_DEFEND << 30) | p;
                c.q_off = (uint32_t)pb; c.q_len = na; c.x_off = (uint32_t)(pb + na); c.x_len = cnt;
                c.weight = W_CHAL;
                c.hash = hash_of(hash_head(K_DEFEND, p, na, cnt), sqx);
                c.res = 0;
                c.parent = i;
                c.slot = slot;
                cands[cbase + slot] = c;
            }
            slot++;
            emitted[n_emitted++] = which;
            pb += ql;
        }
        if (tid == 0) nsucc[i] = slot;
        for (uint32_t j = cbase + slot + tid; j < cend; j += HV_BLOCK) cands[j].kind_p = INVALID;
        __syncthreads();   // qs and fl are rewritten for the next position
    }
}

namespace {

constexpr uint32_t SUB_RC = 8;   // transitions of a member kept in registers

// Lanes [first, first + G) of a wave work on one position.
template <uint32_t G>
struct Coop {
    static __device__ __forceinline__ uint32_t lane() { return threadIdx.x & (G - 1u); }
    static __device__ __forceinline__ uint32_t first() { return threadIdx.x & 63u & ~(G - 1u); }
    // exclusive prefix sum over the lanes; `total` gets their sum
    static __device__ __forceinline__ uint32_t scan(uint32_t v, uint32_t& total) {
        const uint32_t r = lane();
        uint32_t x = v;
        for (uint32_t d = 1; d < G; d <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)x, d, (int)G);
            if (r >= d) x += y;
        }
        total = (uint32_t)__shfl((int)x, (int)G - 1, (int)G);
        return x - v;
    }
    static __device__ __forceinline__ uint32_t sum(uint32_t v) {
        uint32_t t;
        (void)scan(v, t);
        return t;
    }
    // the lanes' bits of a wave ballot (bit r: lane r)
    static __device__ __forceinline__ uint64_t ballot(bool f) {
        const uint64_t b = __ballot(f);
        if constexpr (G == 64u) return b;
        else return (b >> first()) & ((1ull << G) - 1ull);
    }
    static __device__ __forceinline__ uint32_t below(uint64_t m) { return (uint32_t)__popcll(m & ((1ull << lane()) - 1ull)); }
};

// One position's LDS: Q's members, their enabledness flags (bit 0 subset,
// bit 1 superset), a label group's successors as gathered, and the rank
// sort's output for groups of at most SUB_SORT successors (larger groups of
// a wave are sorted in place).
constexpr uint32_t SUB_SORT = 64;
template <uint32_t QMAX, uint32_t TMAX>
struct CoopLds {
    uint32_t mem[QMAX];
    uint32_t gath[TMAX];
    uint32_t sort[SUB_SORT];
    uint8_t flag[QMAX];
};

}  // namespace

// Position i's candidates by G lanes, with the candidates, set-pool layout and
// hashes of gen_walk.  Attack positions with a non-empty Q of at most QMAX
// members whose member degrees sum to at most TMAX:
//   - lane r takes member r (and r + G, ...); member r's transitions stay in
//     registers when it has at most SUB_RC of them, so the label groups of p
//     read no further memory for it;
//   - per label group of p, the members' successors are gathered in LDS
//     (offsets from a scan over the lanes), sorted (rank counting, or a
//     bitonic sort for large groups of a wave), deduplicated into Q' (written
//     to the pool), and the group's transitions are tested and emitted in
//     parallel, ranks from a ballot keeping the slot order;
//   - the conjunction challenges partition Q with ballots.
// Defend positions emit their clause moves in parallel.  Clause positions,
// attack positions with an empty Q and attack positions past the limits run
// gen_walk on the first lane.
template <uint32_t G, uint32_t QMAX, uint32_t TMAX>
__device__ __forceinline__ void gen_coop(const NodeArgs& N, uint32_t lo, uint32_t i, const LtsArgs& L, const uint2* off,
                                         uint64_t spool_base, Cand* cands, uint32_t* nsucc, CoopLds<QMAX, TMAX>& S) {
    using C = Coop<G>;
    const uint32_t r = C::lane();
    uint32_t* const sm = S.mem;
    uint8_t* const sf = S.flag;
    const uint32_t u = lo + i;
    const uint32_t kind = N.kind[u], p = N.p[u], qo = N.q_off[u], ql = N.q_len[u], xl = N.x_len[u];
    uint32_t* spool = N.spool;
    const uint2 o0 = off[i], o1 = off[i + 1];
    const uint32_t cbase = o0.x, cend = o1.x;
    const uint64_t sbase = spool_base + o0.y;
    if (L.job_dead && L.job_dead[L.job_of_state[p]]) {   // k_bounds gave it no room
        if (r == 0) nsucc[i] = 0;
        return;
    }
    if (kind == K_DEFEND) {
        // Defend(p, Q, Q*) -> Clause(p, q) per q in Q [0,0,0,Min(3)], Attack(p, Q*) [Min(3)]
        for (uint32_t k = r; k < ql; k += G) {
            Cand c;
            c.kind_p = (K_CLAUSE << 30) | p;
            c.q_off = qo + k; c.q_len = 1; c.x_off = 0; c.x_len = 0;
            c.weight = W_ANS;
            c.hash = pos_hash(K_CLAUSE, p, spool, qo + k, 1, 0, 0);
            c.res = 0;
            c.parent = i;
            c.slot = k;
            cands[cbase + k] = c;
        }
        if (r == 0) {
            if (xl) {
                const uint32_t xo = N.x_off[u];
                Cand c;
                c.kind_p = (K_ATTACK << 30) | p;
                c.q_off = xo; c.q_len = xl; c.x_off = 0; c.x_len = 0;
                c.weight = W_REV;
                c.hash = pos_hash(K_ATTACK, p, spool, xo, xl, 0, 0);
                c.res = 0;
                c.parent = i;
                c.slot = ql;
                cands[cbase + ql] = c;
            }
            nsucc[i] = ql + (xl ? 1u : 0u);   // all the room k_bounds gave it
        }
        return;
    }
    const uint32_t sd = (o1.y - o0.y) - 4u * ql;   // attack positions: set-pool words = sum of member degrees + 4 |Q|
    if (kind == K_CLAUSE || ql == 0 || ql > QMAX || sd > TMAX) {
        if (r == 0) gen_walk(N, lo, i, L, off, spool_base, cands, nsucc, S.sort, 1);
        return;
    }
    // Q's members in LDS; member r and, when they fit, its transitions in registers
    for (uint32_t k = r; k < ql; k += G) sm[k] = spool[qo + k];
    const uint32_t q0 = r < ql ? spool[qo + r] : 0u;
    uint32_t q0b = 0, q0d = 0;
    if (r < ql) {
        q0b = L.row[q0];
        q0d = L.row[q0 + 1] - q0b;
    }
    const bool cached = r < ql && q0d <= SUB_RC;
    int32_t cl[SUB_RC];
    uint32_t cd[SUB_RC];
#pragma unroll
    for (uint32_t j = 0; j < SUB_RC; j++) {
        const bool in = cached && j < q0d;
        cl[j] = in ? L.lbl[q0b + j] : INT32_MIN;   // no label equals INT32_MIN
        cd[j] = in ? L.dst[q0b + j] : 0u;
    }
    // p's label groups: from transitions held one per lane when p has at most
    // G of them, else from its distinct labels and label ranges
    const uint32_t pr0 = L.row[p], pdeg = L.row[p + 1] - pr0;
    const bool pc = pdeg <= G;
    int32_t plb = INT32_MIN;
    uint32_t pdst = 0;
    if (pc && r < pdeg) {
        plb = L.lbl[pr0 + r];
        pdst = L.dst[pr0 + r];
    }
    const int32_t prev = __shfl_up(plb, 1, (int)G);
    uint64_t rest = C::ballot(pc && r < pdeg && (r == 0 || prev != plb));   // group starts
    const uint32_t ng = pc ? (uint32_t)__popcll(rest) : L.en_row[p + 1] - L.en_row[p];
    wave_sync();   // the members in LDS
    uint32_t slot = 0;
    uint64_t sp = sbase;
    for (uint32_t g = 0; g < ng; g++) {
        int32_t a;
        uint32_t t, te;
        if (pc) {
            const uint32_t s0 = (uint32_t)__ffsll((unsigned long long)rest) - 1u;
            rest &= rest - 1ull;
            const uint32_t s1 = rest ? (uint32_t)__ffsll((unsigned long long)rest) - 1u : pdeg;
            t = pr0 + s0;
            te = pr0 + s1;
            a = __shfl(plb, (int)s0, (int)G);
        } else {
            a = L.en_lbl[L.en_row[p] + g];
            label_range(L, p, a, t, te);
        }
        // the members' label-a successors, gathered (in any order: they are sorted next)
        uint32_t c = 0;
        if (cached) {
#pragma unroll
            for (uint32_t j = 0; j < SUB_RC; j++) c += cl[j] == a ? 1u : 0u;
        } else if (r < ql) {
            uint32_t x0, x1;
            label_range(L, q0, a, x0, x1);
            c = x1 - x0;
        }
        for (uint32_t k = r + G; k < ql; k += G) {
            uint32_t x0, x1;
            label_range(L, sm[k], a, x0, x1);
            c += x1 - x0;
        }
        uint32_t T;
        uint32_t w = C::scan(c, T);   // T <= sd <= TMAX
        uint32_t* const sg = S.gath;
        if (cached) {
#pragma unroll
            for (uint32_t j = 0; j < SUB_RC; j++)
                if (cl[j] == a) sg[w++] = cd[j];
        } else if (r < ql) {
            uint32_t x0, x1;
            label_range(L, q0, a, x0, x1);
            for (uint32_t x = x0; x < x1; x++) sg[w++] = L.dst[x];
        }
        for (uint32_t k = r + G; k < ql; k += G) {
            uint32_t x0, x1;
            label_range(L, sm[k], a, x0, x1);
            for (uint32_t x = x0; x < x1; x++) sg[w++] = L.dst[x];
        }
        wave_sync();
        // sorted: by rank counting (ties by gather index) into S.sort, or for a
        // large group of a wave by a bitonic sort in place (padding sorts last:
        // state ids are below 2^30); Q' then goes to sg (in place for the
        // bitonic sort: each round reads its values before it moves them down)
        const uint32_t* srt = S.sort;
        uint32_t* qs = sg;   // Q'
        if (G == 64u && T > SUB_SORT) {
            uint32_t n2 = 1;
            while (n2 < T) n2 <<= 1;   // <= TMAX, a power of two
            for (uint32_t x = T + r; x < n2; x += G) sg[x] = 0xffffffffu;
            wave_sync();
            for (uint32_t k = 2; k <= n2; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t x = r; x < n2; x += G) {
                        const uint32_t y = x ^ j;
                        if (y > x) {
                            const uint32_t va = sg[x], vb = sg[y];
                            if ((va > vb) == ((x & k) == 0u)) {
                                sg[x] = vb;
                                sg[y] = va;
                            }
                        }
                    }
                    wave_sync();
                }
            srt = sg;
        } else {
            for (uint32_t x = r; x < T; x += G) {
                const uint32_t v = sg[x];
                uint32_t rank = 0;
                for (uint32_t y = 0; y < T; y++) {
                    const uint32_t z = sg[y];
                    rank += (z < v || (z == v && y < x)) ? 1u : 0u;
                }
                S.sort[rank] = v;
            }
            wave_sync();
        }
        // Q' = the distinct values in order: into qs and the pool at sp; sq is
        // its part of the group's candidate hashes
        uint32_t nq = 0, part_sum = 0;
        for (uint32_t x0 = 0; x0 < T; x0 += G) {
            const uint32_t x = x0 + r;
            const uint32_t v = x < T ? srt[x] : 0u;
            const bool keep = x < T && (x == 0u || v != srt[x - 1u]);
            const uint64_t m = C::ballot(keep);
            wave_sync();   // every read of this round before the moves (qs may be srt)
            if (keep) {
                const uint32_t d = nq + C::below(m);   // d <= x
                qs[d] = v;
                spool[sp + d] = v;
                part_sum += set_term(v, SALT_Q);
            }
            nq += (uint32_t)__popcll(m);
            wave_sync();
        }
        const uint32_t sq = C::sum(part_sum);
        wave_sync();
        // an observation move to every target of the group outside Q', in transition order
        for (uint32_t b0 = t; b0 < te; b0 += G) {
            uint32_t pp = 0;
            bool in;
            if (pc) {   // lane r holds transition pr0 + r
                in = r < pdeg && pr0 + r >= t && pr0 + r < te;
                pp = pdst;
            } else {
                in = b0 + r < te;
                if (in) pp = L.dst[b0 + r];
            }
            const bool e = in && !contains_sorted(qs, nq, pp);
            const uint64_t m = C::ballot(e);
            if (e) {
                Cand cc;
                cc.kind_p = (K_ATTACK << 30) | pp;
                cc.q_off = (uint32_t)sp; cc.q_len = nq; cc.x_off = 0; cc.x_len = 0;
                cc.weight = W_OBS;
                cc.hash = hash_of(hash_head(K_ATTACK, pp, nq, 0), sq);
                cc.res = 0;
                cc.parent = i;
                cc.slot = slot + C::below(m);
                cands[cbase + cc.slot] = cc;
            }
            slot += (uint32_t)__popcll(m);
            if (pc) break;   // one pass covers p's transitions
        }
        sp += nq;
        wave_sync();   // the buffers are rewritten by the next group
    }
    uint32_t q_sum = 0;
    for (uint32_t k = r; k < ql; k += G) q_sum += set_term(sm[k], SALT_Q);
    const uint32_t sq_all = C::sum(q_sum);
    if (r == 0) {   // Defend(p, Q, {})
        Cand cc;
        cc.kind_p = (K_DEFEND << 30) | p;
        cc.q_off = qo; cc.q_len = ql; cc.x_off = 0; cc.x_len = 0;
        cc.weight = W_CHAL;
        cc.hash = hash_of(hash_head(K_DEFEND, p, ql, 0), sq_all);
        cc.res = 0;
        cc.parent = i;
        cc.slot = slot;
        cands[cbase + slot] = cc;
    }
    slot++;
    // conjunction challenges with partitioned Q* (gamepos.rs:329-375)
    for (uint32_t k = r; k < ql; k += G) {
        bool sub, sup;
        compare_enabled(L, p, sm[k], sub, sup);
        sf[k] = (uint8_t)((sub ? 1u : 0u) | (sup ? 2u : 0u));
    }
    wave_sync();
    uint64_t pb = sbase + sd;
    uint32_t emitted[3];
    uint32_t n_emitted = 0;
    for (uint32_t which = 0; which < 3u; which++) {
        uint32_t local = 0;
        for (uint32_t k = r; k < ql; k += G) local += in_part(which, sf[k]) ? 1u : 0u;
        const uint32_t cnt = C::sum(local);
        if (cnt == 0u || cnt == ql) continue;
        bool dup = false;
        for (uint32_t e = 0; e < n_emitted; e++) {
            uint32_t diff = 0;
            for (uint32_t k = r; k < ql; k += G) diff += in_part(which, sf[k]) != in_part(emitted[e], sf[k]) ? 1u : 0u;
            dup |= C::sum(diff) == 0u;
        }
        if (dup) continue;
        // stable partition: the members outside the part, then those inside
        const uint32_t na = ql - cnt;
        uint32_t run = 0, h_sum = 0;   // the members outside the part are Q, those inside Q*
        for (uint32_t k0 = 0; k0 < ql; k0 += G) {
            const uint32_t k = k0 + r;
            const bool f = k < ql && in_part(which, sf[k]);
            const uint64_t m = C::ballot(f);
            const uint32_t before = run + C::below(m);
            if (k < ql) {
                if (f) spool[pb + na + before] = sm[k];
                else spool[pb + (k - before)] = sm[k];
                h_sum += set_term(sm[k], f ? SALT_X : SALT_Q);
            }
            run += (uint32_t)__popcll(m);
        }
        const uint32_t sqx = C::sum(h_sum);
        if (r == 0) {
            Cand cc;
            cc.kind_p = (K_DEFEND << 30) | p;
            cc.q_off = (uint32_t)pb; cc.q_len = na; cc.x_off = (uint32_t)(pb + na); cc.x_len = cnt;
            cc.weight = W_CHAL;
            cc.hash = hash_of(hash_head(K_DEFEND, p, na, cnt), sqx);
            cc.res = 0;
            cc.parent = i;
            cc.slot = slot;
            cands[cbase + slot] = cc;
        }
        slot++;
        emitted[n_emitted++] = which;
        pb += ql;
    }
    if (r == 0) nsucc[i] = slot;
    for (uint32_t j = cbase + slot + r; j < cend; j += G) cands[j].kind_p = INVALID;
}

// SUB_LANES lanes per frontier position; the listed (heavy) positions are left
// to k_generate_wave and k_generate_heavy.
__global__ __launch_bounds__(SUB_BLOCK) void k_generate_sub(NodeArgs N, uint32_t lo, uint32_t n_front, LtsArgs L,
                                                            const uint2* off, const uint8_t* heavy, uint64_t spool_base,
                                                            Cand* cands, uint32_t* nsucc) {
    constexpr uint32_t PER_BLOCK = SUB_BLOCK / SUB_LANES;
    __shared__ CoopLds<SUB_QMAX, SUB_TMAX> lds[PER_BLOCK];
    const uint32_t gs = threadIdx.x / SUB_LANES;
    const uint32_t i = blockIdx.x * PER_BLOCK + gs;
    if (i > n_front) return;
    if (i == n_front) {   // trailing zero: the exclusive scan yields the edge total
        if ((threadIdx.x & (SUB_LANES - 1u)) == 0) nsucc[i] = 0;
        return;
    }
    if (heavy[i]) return;
    gen_coop<SUB_LANES, SUB_QMAX, SUB_TMAX>(N, lo, i, L, off, spool_base, cands, nsucc, lds[gs]);
}

// One wave per listed position tagged WAVE_TAG (grid-stride over the list).
__global__ __launch_bounds__(SUB_BLOCK) void k_generate_wave(NodeArgs N, uint32_t lo, LtsArgs L, const uint2* off,
                                                             uint64_t spool_base, HeavyArgs H, Cand* cands, uint32_t* nsucc) {
    constexpr uint32_t PER_BLOCK = SUB_BLOCK / 64u;
    __shared__ CoopLds<WAVE_QMAX, WAVE_TMAX> lds[PER_BLOCK];
    const uint32_t w = threadIdx.x / 64u;
    const uint32_t n = *H.count;
    for (uint32_t h = blockIdx.x * PER_BLOCK + w; h < n; h += gridDim.x * PER_BLOCK) {
        const uint32_t e = H.list[h];
        if (!(e & WAVE_TAG)) continue;
        gen_coop<64u, WAVE_QMAX, WAVE_TMAX>(N, lo, e & ~WAVE_TAG, L, off, spool_base, cands, nsucc, lds[w]);
        wave_sync();   // the LDS is rewritten by the next position
    }
}

__device__ bool cand_equals_cand(const Cand& a, const Cand& b, const uint32_t* spool) {
    return a.kind_p == b.kind_p && a.q_len == b.q_len && a.x_len == b.x_len && sets_equal(spool, a.q_off, b.q_off, a.q_len) &&
           sets_equal(spool, a.x_off, b.x_off, a.x_len);
}

__device__ bool cand_equals_node(const Cand& a, const NodeArgs& N, uint32_t id) {
    // every field read before the first comparison (no load waits on another)
    const uint32_t kind = N.kind[id], p = N.p[id], ql = N.q_len[id], xl = N.x_len[id], qo = N.q_off[id], xo = N.x_off[id];
    if (a.kind_p != ((kind << 30) | p) || a.q_len != ql || a.x_len != xl) return false;
    return sets_equal(N.spool, a.q_off, qo, ql) && sets_equal(N.spool, a.x_off, xo, xl);
}

// candidate j finds its position in the table, or claims a pending entry
__device__ __forceinline__ void insert_cand(Cand* cands, uint32_t j, const NodeArgs& N, unsigned long long* table, uint32_t* keymin,
                                            uint32_t mask) {
    Cand c = cands[j];
    if (c.kind_p == INVALID) return;
    const uint32_t h = c.hash;
    uint32_t s = h & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        unsigned long long e = __hip_atomic_load(&table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e == EMPTY) {
            const unsigned long long mine = ((unsigned long long)h << 32) | (PENDING | j);
            unsigned long long old = atomicCAS(&table[s], EMPTY, mine);
            if (old == EMPTY) {
                atomicMin(&keymin[s], j);
                cands[j].res = PENDING | s;
                return;
            }
            e = old;
        }
        if ((uint32_t)(e >> 32) == h) {
            const uint32_t ref = (uint32_t)e;
            if (ref & PENDING) {
                if (cand_equals_cand(c, cands[ref & ~PENDING], N.spool)) {
                    atomicMin(&keymin[s], j);
                    cands[j].res = PENDING | s;
                    return;
                }
            } else if (cand_equals_node(c, N, ref)) {
                cands[j].res = ref;
                return;
            }
        }
        s = (s + 1) & mask;
    }
}

__global__ void k_insert(Cand* cands, uint32_t n_cand, NodeArgs N, unsigned long long* table, uint32_t* keymin, uint32_t mask) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_cand) insert_cand(cands, j, N, table, keymin, mask);
}

// 1 if candidate j is the first of a new position (j == n_cand: the trailing 0)
__device__ __forceinline__ uint32_t mark_cand(const Cand* cands, uint32_t j, uint32_t n_cand, const uint32_t* keymin) {
    if (j == n_cand) return 0u;
    const Cand& c = cands[j];
    // keymin was set by atomics, which act in the L2: read it there (a block
    // that runs several levels may hold an older copy of its line in the L1)
    return (c.kind_p != INVALID && (c.res & PENDING) &&
            __hip_atomic_load(&keymin[c.res & ~PENDING], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == j)
               ? 1u
               : 0u;
}

__global__ void k_mark(const Cand* cands, uint32_t n_cand, const uint32_t* keymin, uint32_t* flag) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_cand) flag[j] = mark_cand(cands, j, n_cand, keymin);
}

geq_status scan_marks(const Cand* cands, uint32_t n_cand, const uint32_t* keymin, uint32_t* rank, uint32_t* flag,
                      DevBuf<uint32_t>& tmp, hipStream_t s) {
    // the flags first: one candidate per thread (computing them inside the
    // scan made each thread chase 16 records and their table slots in turn)
    if (n_cand) k_mark<<<(n_cand + 255) / 256, 256, 0, s>>>(cands, n_cand, keymin, flag);
    return exclusive_scan_dev<uint32_t, SumU32>(flag, rank, n_cand, tmp, s, SumU32{});
}


// a new position's record and table entry (candidate j, flagged)
__device__ __forceinline__ void finalize_cand(const Cand* cands, uint32_t j, const uint32_t* rank, uint32_t v_base, const NodeArgs& N,
                                              unsigned long long* table, uint32_t* keymin) {
    const Cand c = cands[j];
    const uint32_t id = v_base + rank[j];
    const uint32_t s = c.res & ~PENDING;
    N.kind[id] = (uint8_t)(c.kind_p >> 30);
    N.p[id] = c.kind_p & 0x3fffffffu;
    N.q_off[id] = c.q_off; N.q_len[id] = c.q_len;
    N.x_off[id] = c.x_off; N.x_len[id] = c.x_len;
    N.hash[id] = c.hash;
    table[s] = ((unsigned long long)c.hash << 32) | id;
    keymin[s] = 0xffffffffu;
}

// rank: the exclusive scan of the winner flags (a winner is where it steps)
// (clear: the heavy count, zeroed for a following k_level_out; NULL: none)
__global__ void k_finalize(const Cand* cands, uint32_t n_cand, const uint32_t* rank, uint32_t v_base, NodeArgs N,
                           unsigned long long* table, uint32_t* keymin, uint32_t* clear) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear && j == 0) *clear = 0u;
    if (j < n_cand && rank[j + 1] != rank[j]) finalize_cand(cands, j, rank, v_base, N, table, keymin);
}

// candidate j's edge: its parent's row, its slot, the target's id
__device__ __forceinline__ void edge_cand(const Cand* cands, uint32_t j, const uint32_t* rowl, uint32_t e_base,
                                          const unsigned long long* table, uint32_t* col, uint32_t* wt) {
    const Cand c = cands[j];
    if (c.kind_p == INVALID) return;
    const uint32_t target = (c.res & PENDING) ? (uint32_t)table[c.res & ~PENDING] : c.res;
    const uint32_t e = e_base + rowl[c.parent] + c.slot;
    col[e] = target;
    wt[e] = c.weight;
}

__global__ void k_edges(const Cand* cands, uint32_t n_cand, const uint32_t* rowl, uint32_t e_base, const unsigned long long* table,
                        uint32_t* col, uint32_t* wt) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_cand) edge_cand(cands, j, rowl, e_base, table, col, wt);
}

__global__ void k_level_out(const Cand* cands, uint32_t n_cand, const uint32_t* rowl, uint32_t e_base, const unsigned long long* table,
                            uint32_t* col, uint32_t* wt, uint32_t nf, uint32_t lo, uint32_t* row, NodeArgs N, uint32_t hi,
                            const uint32_t* n_new, uint32_t add, uint32_t n_slots, uint32_t chunk, LtsArgs L, uint2* ub, HeavyArgs H) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_cand) edge_cand(cands, g, rowl, e_base, table, col, wt);
    if (g < nf) row[lo + g] = e_base + rowl[g];
    if (g < n_slots) bounds_next_at(N, hi, *n_new + add, chunk, L, ub, H, g);
}

__global__ void k_rows(const uint32_t* rowl, uint32_t n_front, uint32_t lo, uint32_t e_base, uint32_t* row, RowsTail T) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_front) row[lo + i] = e_base + rowl[i];
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0 && T.heavy_count) *T.heavy_count = 0u;
        if (T.dead)
            for (uint32_t j = threadIdx.x; j < T.n_jobs; j += blockDim.x)
                if (T.cnt[j] > T.cap) T.dead[j] = 1;
    }
}

// count word i of a chunk (see k_chunk_counts)
__device__ __forceinline__ uint32_t chunk_word(uint32_t i, const uint32_t* n_edges, const uint32_t* n_new, uint32_t add, uint32_t chunk,
                                               const uint2* next, const uint32_t* heavy, const uint32_t* cnt) {
    if (i == RB_EDGES) return *n_edges;
    if (i == RB_NEW) return *n_new;
    if (i == RB_CAND || i == RB_WORDS) {
        if (!next) return 0u;
        const uint2 t = next[min(*n_new + add, chunk)];
        return i == RB_CAND ? t.x : t.y;
    }
    if (i == RB_HEAVY) return next ? __hip_atomic_load(heavy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;   // counted by atomics
    if (i < RB_JOBS) return 0u;   // the words of k_level_tail's own levels
    return __hip_atomic_load(&cnt[i - RB_JOBS], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // counted by atomics
}

__global__ void k_chunk_counts(const uint32_t* n_edges, const uint32_t* n_new, uint32_t add, uint32_t chunk, const uint2* next,
                               const uint32_t* heavy, const uint32_t* cnt, uint32_t n_cnt, uint32_t* out, uint32_t seq) {
    if (threadIdx.x >= 64) return;   // the first wave writes and publishes every word
    for (uint32_t i = threadIdx.x; i < RB_JOBS + n_cnt; i += 64)
        if (i != RB_SEQ) out[i] = chunk_word(i, n_edges, n_new, add, chunk, next, heavy, cnt);
    rb_publish(out, seq);
}

// Exclusive scan of in[0..n) into out[0..n], out[n] the total, by one block of
// TAIL_BLOCK threads (four consecutive items each; n <= TAIL_MAX).
template <class T, class Op>
__device__ void tail_scan(const T* in, T* out, uint32_t n, Op op) {
    const uint32_t b = 4u * threadIdx.x;
    T v[4];
    T acc = Op::identity();
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        v[k] = b + k < n ? in[b + k] : Op::identity();
        acc = op(acc, v[k]);
    }
    T total;
    T run = detail::block_scan_excl<T, Op, TAIL_BLOCK>(acc, op, total);
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        if (b + k < n) out[b + k] = run;
        run = op(run, v[k]);
    }
    if (threadIdx.x == 0) out[n] = total;
}

// One chunk after its generation (the body of k_level_tail).
__device__ void tail_chunk(const TailArgs& T, uint32_t lo, uint32_t nf, uint32_t n_cand, uint32_t v_base, uint32_t e_base,
                           uint32_t hi, uint32_t add, uint32_t slots, uint32_t last) {
    const uint32_t t = threadIdx.x;
    // edge offsets of the chunk's positions (the generation wrote their successor counts)
    tail_scan<uint32_t, SumU32>(T.nsucc, T.rowl, nf, SumU32{});
    for (uint32_t j = t; j < n_cand; j += TAIL_BLOCK) insert_cand(T.cands, j, T.N, T.table, T.keymin, T.mask);
    __syncthreads();
    for (uint32_t j = t; j <= n_cand; j += TAIL_BLOCK) T.flag[j] = mark_cand(T.cands, j, n_cand, T.keymin);
    __syncthreads();
    tail_scan<uint32_t, SumU32>(T.flag, T.rank, n_cand, SumU32{});
    __syncthreads();
    for (uint32_t j = t; j < n_cand; j += TAIL_BLOCK)
        if (T.flag[j]) finalize_cand(T.cands, j, T.rank, v_base, T.N, T.table, T.keymin);
    __syncthreads();   // edges read the table entries finalize wrote, and the row offsets
    for (uint32_t j = t; j < n_cand; j += TAIL_BLOCK) {
        edge_cand(T.cands, j, T.rowl, e_base, T.table, T.col, T.wt);
        if (T.job_cnt && T.flag[j]) atomicAdd(&T.job_cnt[T.L.job_of_state[T.cands[j].kind_p & 0x3fffffffu]], 1u);
    }
    for (uint32_t i = t; i < nf; i += TAIL_BLOCK) T.row[lo + i] = e_base + T.rowl[i];
    __syncthreads();   // the per-job counts are complete
    if (T.job_dead && T.cap)   // the counts come from atomics: read in the L2
        for (uint32_t j = t; j < T.n_jobs; j += TAIL_BLOCK)
            if (__hip_atomic_load(&T.job_cnt[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > T.cap) T.job_dead[j] = 1;
    if (last) {
        if (t == 0) *T.H.count = 0u;
        __syncthreads();   // the dead jobs and the cleared heavy count before the bounds
        const uint32_t n_front = T.rank[n_cand] + add;   // the positions this level created
        for (uint32_t i = t; i < slots; i += TAIL_BLOCK) bounds_next_at(T.N, hi, n_front, T.chunk, T.L, T.ub, T.H, i);
        __syncthreads();
        tail_scan<uint2, PairSum>(T.ub, T.off, slots - 1u, PairSum{T.lim});
    }
    __syncthreads();
}

// The count words of the last level done (first wave), RB_SEQ published last.
__device__ void tail_publish(const TailArgs& T, uint32_t done, uint32_t lo, uint32_t nf, uint32_t n_cand, uint32_t n_words,
                             uint32_t v_base, uint32_t e_base, uint32_t sp, uint32_t add, uint32_t last) {
    const uint32_t t = threadIdx.x;
    const uint32_t n_cnt = T.job_cnt ? T.n_jobs : 0u;
    if (t >= 64) return;   // the first wave writes and publishes every word
    for (uint32_t i = t; i < RB_JOBS + n_cnt; i += 64) {
        uint32_t w;
        switch (i) {
            case RB_SEQ: continue;
            case RB_LOOP: w = done; break;
            case RB_LOOP_LO: w = lo; break;
            case RB_LOOP_NF: w = nf; break;
            case RB_LOOP_CAND: w = n_cand; break;
            case RB_LOOP_WORDS: w = n_words; break;
            case RB_LOOP_V: w = v_base; break;
            case RB_LOOP_E: w = e_base; break;
            case RB_LOOP_SP: w = sp; break;
            default:
                w = chunk_word(i, T.rowl + nf, T.rank + n_cand, add, T.chunk, last ? T.off : nullptr, T.H.count, T.job_cnt);
        }
        T.rb[i] = w;
    }
    rb_publish(T.rb, T.seq);
}

__global__ __launch_bounds__(TAIL_BLOCK) void k_level_tail(TailArgs T) {
    tail_chunk(T, T.lo, T.nf, T.n_cand, T.v_base, T.e_base, T.hi, T.add, T.slots, T.last);
    tail_publish(T, 0u, T.lo, T.nf, T.n_cand, T.room.n_words, T.v_base, T.e_base, T.room.spool_top, T.add, T.last);
}

// The launched chunk (a level's only chunk), then up to T.loop further levels
// in the same block: when the next level has at most T.room.gen_max positions
// (at most TAIL_GEN) and TAIL_MAX candidates, lists no position for
// k_generate_wave / k_generate_heavy, and fits every buffer's capacity and the
// table's load (T.room), the block generates it itself (SUB_LANES lanes per
// position, as k_generate_sub) and goes on.  The count words describe the
// last level done; RB_LOOP* give how many levels were added and where the
// last one started.
__global__ __launch_bounds__(TAIL_BLOCK) void k_level_tail_loop(TailArgs T) {
    constexpr uint32_t GROUPS = TAIL_BLOCK / SUB_LANES;   // positions generated at once
    __shared__ CoopLds<SUB_QMAX, SUB_TMAX> glds[GROUPS];
    const uint32_t t = threadIdx.x;
    uint32_t lo = T.lo, nf = T.nf, n_cand = T.n_cand, v_base = T.v_base, e_base = T.e_base, hi = T.hi;
    uint32_t sp = T.room.spool_top, n_words = T.room.n_words, done = 0;
    for (;;) {
        tail_chunk(T, lo, nf, n_cand, v_base, e_base, hi, 0u, n_cand + 1u, 1u);
        if (done >= T.loop) break;
        // the next level: the positions this one created, [hi, hi + n_new)
        const uint32_t n_new = T.rank[n_cand];
        const uint32_t e_next = e_base + T.rowl[nf];
        const uint2 tot = T.off[min(n_new, T.chunk)];
        const uint32_t v_next = hi + n_new;
        const uint64_t sp_next = (uint64_t)sp + n_words;
        const TailRoom& R = T.room;
        const bool go = n_new > 0 && n_new <= min(TAIL_GEN, R.gen_max) && n_new <= T.chunk && tot.x <= min(TAIL_MAX, R.tail_max) &&
                        __hip_atomic_load(T.H.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u &&
                        (!R.max_positions || v_next <= R.max_positions) && n_new + 1u <= R.front &&
                        tot.x + 1u <= R.cand && tot.x + 1u <= R.slots && (uint64_t)v_next + tot.x + 1u <= R.nodes &&
                        2ull * ((uint64_t)v_next + tot.x) <= R.table && (uint64_t)e_next + tot.x + 1u <= R.edges &&
                        sp_next + tot.y + 1u <= R.words && tot.x < T.lim && tot.y < T.lim;
        if (!go) break;
        __syncthreads();   // every thread has read the values above before the generation rewrites them
        for (uint32_t i = t / SUB_LANES; i < n_new; i += GROUPS)
            gen_coop<SUB_LANES, SUB_QMAX, SUB_TMAX>(T.N, hi, i, T.L, T.off, sp_next, T.cands, T.nsucc, glds[t / SUB_LANES]);
        if (t == 0) T.nsucc[n_new] = 0u;   // trailing zero: the scan yields the edge total
        __syncthreads();
        lo = hi; nf = n_new; n_cand = tot.x; v_base = v_next; e_base = e_next; hi = v_next;
        sp = (uint32_t)sp_next; n_words = tot.y;
        done++;
    }
    tail_publish(T, done, lo, nf, n_cand, n_words, v_base, e_base, sp, 0u, 1u);
}

// start positions Attack(p, [q]) at ids [0, n): node records + hash entries
__global__ void k_starts(const uint32_t* pairs, uint32_t n, const uint8_t* insert, NodeArgs N, unsigned long long* table,
                         uint32_t mask) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    N.kind[i] = K_ATTACK;
    N.p[i] = pairs[2 * i];
    N.spool[i] = pairs[2 * i + 1];
    N.q_off[i] = i; N.q_len[i] = 1;
    N.x_off[i] = 0; N.x_len[i] = 0;
    // pairs were written by the host before the launch; spool[i] is this thread's own store
    const uint32_t h = pos_hash(K_ATTACK, pairs[2 * i], pairs, 2 * i + 1, 1, 0, 0);
    N.hash[i] = h;
    if (!insert[i]) return;
    uint32_t s = h & mask;
    const unsigned long long mine = ((unsigned long long)h << 32) | i;
    while (atomicCAS(&table[s], EMPTY, mine) != EMPTY) s = (s + 1) & mask;
}

// re-insert every node into a larger table, except the start positions that
// k_starts left out (an earlier copy of a repeated start pair: the table keeps
// the last id, as GameBuild::new_node does)
__global__ void k_rehash(const uint32_t* hash, uint32_t n, const uint8_t* insert, uint32_t n_start, unsigned long long* table,
                         uint32_t mask) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (i < n_start && !insert[i])) return;
    const uint32_t h = hash[i];
    uint32_t s = h & mask;
    const unsigned long long mine = ((unsigned long long)h << 32) | i;
    while (atomicCAS(&table[s], EMPTY, mine) != EMPTY) s = (s + 1) & mask;
}

__global__ void k_fill_u64(unsigned long long* a, uint64_t n, unsigned long long v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_u32(uint32_t* a, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_attacker_flags(const uint8_t* kind, uint32_t n, uint8_t* atk) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atk[i] = kind[i] != K_DEFEND ? 1 : 0;
}

}  // namespace gb
}  // namespace geq
