// The fixpoint for energy configurations outside the bitmap kernels' layout
// (any EnergyConf whose energies take at most 32 words, e.g. SILENT_STEP
// {elements 8, max 2} and ONE_DIMENSION {1, u32::MAX}, src/energy.rs:44-46).
//
// Energies and updates stay in the crate's packing (src/energy.rs:143-179,
// :543-588): `ebits`-wide fields, 32 / ebits of them per word.  Lists are
// antichains of such energies in a word arena; a node's two list slots, their
// round tags, the worklist, the queued bitmaps and the Jacobi round structure
// are those of the bitmap kernels (one k_eval_list + one k_expand per round).
//
// One wave evaluates one worklist node, with the reference's list algorithms
// (attack_minimize.wgsl:67-122, defend_iterative.wgsl:182-267): the attacker
// gathers every successor's inverse-updated list and its own list and keeps
// the Pareto-minimal ones (equal energies: the later copy, so its own
// energies win); the defender folds the successors' images pairwise by
// component-wise maximum, minimising after each successor, and copies a single
// successor's images unminimised.  Candidate lists live in a per-wave region
// of a global scratch buffer; a node that needs more room than the region
// holds records the size it needs and the round runs again after the host
// has grown the buffer (as after an exhausted arena).
#include "solve_kernels.hpp"

namespace geq {

namespace {

__device__ __forceinline__ void gsync() {
    // writes of this wave to the global scratch become visible to its other lanes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// word w of a packed energy held in registers (ESZ words, w < ESZ)
template <int ESZ>
__device__ __forceinline__ uint32_t getw(const uint32_t (&e)[ESZ], uint32_t w) {
    uint32_t r = e[0];
#pragma unroll
    for (int k = 1; k < ESZ; k++) r = (uint32_t)k == w ? e[k] : r;
    return r;
}
template <int ESZ>
__device__ __forceinline__ void setw(uint32_t (&e)[ESZ], uint32_t w, uint32_t v) {
#pragma unroll
    for (int k = 0; k < ESZ; k++)
        if ((uint32_t)k == w) e[k] = v;
}

template <int ESZ>
__device__ __forceinline__ void load_e(uint32_t (&e)[ESZ], const uint32_t* __restrict__ p, uint32_t esz) {
#pragma unroll
    for (int k = 0; k < ESZ; k++) e[k] = (uint32_t)k < esz ? p[k] : 0u;
}
template <int ESZ>
__device__ __forceinline__ void store_e(uint32_t* __restrict__ p, const uint32_t (&e)[ESZ], uint32_t esz) {
#pragma unroll
    for (int k = 0; k < ESZ; k++)
        if ((uint32_t)k < esz) p[k] = e[k];
}

// inv_update (shadergen.rs:103-128 as instantiated in attack_update.wgsl:45-61):
// first every Decrement field i: e[i] = min(e[i] + 1, max) in u32 arithmetic
// (a 32-bit field at u32::MAX wraps to 0 before the min, as in WGSL); then, in
// order i = 0..elements, every Min field (value v >= 2): e[v-2] = max(e[v-2],
// e[i]), reading values already updated earlier in the sequence.  Fields sit
// whole inside their word (field widths are powers of two).
template <int ESZ>
__device__ void inv_update_g(uint32_t (&e)[ESZ], const uint32_t* __restrict__ u, const GConf& c) {
    for (uint32_t i = 0; i < c.elements; i++) {
        const uint32_t ub = i * c.ubits;
        if (((u[ub >> 5] >> (ub & 31u)) & c.umask) != 1u) continue;
        const uint32_t eb = i * c.ebits, w = eb >> 5, sh = eb & 31u;
        const uint32_t word = getw(e, w);
        const uint32_t v = min(((word >> sh) & c.emask) + 1u, c.maxv);
        setw(e, w, (word & ~(c.emask << sh)) | (v << sh));
    }
    for (uint32_t i = 0; i < c.elements; i++) {
        const uint32_t ub = i * c.ubits;
        const uint32_t uf = (u[ub >> 5] >> (ub & 31u)) & c.umask;
        if (uf < 2u) continue;
        const uint32_t j = uf - 2u;          // validated on the host: j < elements
        const uint32_t ei = i * c.ebits, ej = j * c.ebits;
        const uint32_t vi = (getw(e, ei >> 5) >> (ei & 31u)) & c.emask;
        const uint32_t wj = getw(e, ej >> 5), sj = ej & 31u;
        const uint32_t vj = (wj >> sj) & c.emask;
        if (vi > vj) setw(e, ej >> 5, (wj & ~(c.emask << sj)) | (vi << sj));
    }
}

// Every field of word a <= the same field of word b.  Fields narrower than 32
// bits are compared two lanes at a time: the even fields (mask E) and then
// the odd ones, each placed in a lane of twice the field width whose upper
// half holds a guard bit (G); the guard survives the subtraction exactly
// where b's field >= a's field.
__device__ __forceinline__ bool word_le(uint32_t a, uint32_t b, const GConf& c) {
    if (c.ebits == 32u) return a <= b;
    const uint32_t ae = a & c.E, be = b & c.E;
    const uint32_t ao = (a >> c.ebits) & c.E, bo = (b >> c.ebits) & c.E;
    return ((((be | c.G) - ae) & c.G) == c.G) && ((((bo | c.G) - ao) & c.G) == c.G);
}
// field-wise maximum of two words (the same lanes; a set guard selects b's field)
__device__ __forceinline__ uint32_t word_max(uint32_t a, uint32_t b, const GConf& c) {
    if (c.ebits == 32u) return max(a, b);
    const uint32_t ae = a & c.E, be = b & c.E;
    const uint32_t ao = (a >> c.ebits) & c.E, bo = (b >> c.ebits) & c.E;
    const uint32_t ge = ((be | c.G) - ae) & c.G, go = ((bo | c.G) - ao) & c.G;
    const uint32_t me = ge - (ge >> c.ebits), mo = go - (go >> c.ebits);
    const uint32_t xe = (be & me) | (ae & ~me), xo = (bo & mo) | (ao & ~mo);
    return xe | (xo << c.ebits);
}

template <int ESZ>
__device__ __forceinline__ bool e_le(const uint32_t (&a)[ESZ], const uint32_t* __restrict__ b, uint32_t esz, const GConf& c) {
    bool r = true;
#pragma unroll
    for (int k = 0; k < ESZ; k++)
        if ((uint32_t)k < esz) r &= word_le(b[k], a[k], c);   // b <= a
    return r;
}
template <int ESZ>
__device__ __forceinline__ bool e_eq(const uint32_t (&a)[ESZ], const uint32_t* __restrict__ b, uint32_t esz) {
    bool r = true;
#pragma unroll
    for (int k = 0; k < ESZ; k++)
        if ((uint32_t)k < esz) r &= a[k] == b[k];
    return r;
}

// Per-wave LDS: the chunk of up to 64 successors whose lists are being
// gathered (exclusive offsets of their images, arena offsets, edge ids).
struct WaveLds {
    uint32_t ex[64], off[64], edge[64];
};

// Images of successor lists into scr[(base + s) * esz ..]: the n successors
// of the chunk (lane k: list length len, arena offset off, edge id), laid out
// one after the other.  Returns the chunk's total.
template <int ESZ>
__device__ uint32_t gather_images(const GenArgs& a, WaveLds& L, uint32_t lane, uint32_t len, uint32_t off, uint32_t edge,
                                  uint32_t* __restrict__ scr, uint32_t base) {
    const uint32_t esz = a.c.esz;
    const uint32_t incl = wave_incl_scan(len, lane);
    const uint32_t tot = lane_val(incl, 63);
    L.ex[lane] = incl - len;
    L.off[lane] = off;
    L.edge[lane] = edge;
    wave_sync();
    for (uint32_t s = lane; s < tot; s += 64) {
        uint32_t lo = 0, hi = 63;        // the largest lane with ex <= s (lanes without items repeat offsets)
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (L.ex[mid] <= s) lo = mid; else hi = mid - 1;
        }
        uint32_t e[ESZ];
        load_e(e, a.arena + L.off[lo] + (s - L.ex[lo]) * esz, esz);
        inv_update_g(e, a.wt + (size_t)L.edge[lo] * a.c.usz, a.c);
        store_e(scr + (size_t)(base + s) * esz, e, esz);
    }
    gsync();
    return tot;
}

// keep[i] for the n items at scr[0 ..): no other item lies strictly below,
// and of equal items only one is kept (the last one when `keep_last`, else
// the first).  Flags are written to flags[i]; returns the number kept.
template <int ESZ>
__device__ uint32_t pareto_flags(const uint32_t* __restrict__ scr, uint32_t n, uint32_t esz, const GConf& c, bool keep_last,
                                 uint32_t* __restrict__ flags, uint32_t lane, uint32_t own_start, bool& from_succ) {
    uint32_t kept = 0;
    bool fs = false;
    for (uint32_t b0 = 0; b0 < n; b0 += 64) {
        const uint32_t i = b0 + lane;
        uint32_t e[ESZ];
        bool keep = i < n;
        if (keep) load_e(e, scr + (size_t)i * esz, esz);
        for (uint32_t j = 0; j < n; j++) {
            if (__ballot(keep) == 0ull) break;
            const uint32_t* ej = scr + (size_t)j * esz;
            if (keep && j != i) {
                const bool eq = e_eq(e, ej, esz);
                if (eq ? (keep_last ? i < j : j < i) : e_le(e, ej, esz, c)) keep = false;
            }
        }
        if (i < n) flags[i] = keep ? 1u : 0u;
        fs |= keep && i < own_start;
        kept += (uint32_t)__popcll(__ballot(keep));
    }
    from_succ = __ballot(fs) != 0ull;
    return kept;
}

// The kept items (flags) of scr[src ..) written to dst[0 ..) in order.  dst
// may overlap the items below src: each 64-item piece is read completely
// before it is written, and lands at or below its own position.
template <int ESZ>
__device__ void compact_kept(const uint32_t* src, const uint32_t* __restrict__ flags, uint32_t n, uint32_t esz, uint32_t* dst,
                             uint32_t lane) {
    uint32_t pos = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += 64) {
        const uint32_t i = b0 + lane;
        const bool keep = i < n && flags[i] != 0u;
        uint32_t e[ESZ];
        if (keep) load_e(e, src + (size_t)i * esz, esz);
        const uint64_t bal = __ballot(keep);
        gsync();
        if (keep) store_e(dst + (size_t)(pos + mbcnt(bal)) * esz, e, esz);
        pos += (uint32_t)__popcll(bal);
        gsync();
    }
}

// Whether the list src[0 .. n) (energies) equals node u's current list own
// (len own_len at arena offset own_off) as the reference compares
// (src/energygame.rs:929, src/energy.rs:334-369): same length and every new
// energy found in the old list.
template <int ESZ>
__device__ bool same_as_own(const GenArgs& a, const uint32_t* src, uint32_t n, uint32_t own_len, uint32_t own_off, uint32_t lane) {
    if (n != own_len) return false;
    const uint32_t esz = a.c.esz;
    bool all_in = true;
    for (uint32_t b0 = 0; b0 < n; b0 += 64) {
        const uint32_t i = b0 + lane;
        bool found = i >= n;
        uint32_t e[ESZ];
        if (!found) load_e(e, src + (size_t)i * esz, esz);
        for (uint32_t j = 0; j < own_len; j++) {
            if (__ballot(!found) == 0ull) break;
            if (!found) found = e_eq(e, a.arena + own_off + (size_t)j * esz, esz);
        }
        all_in &= found;
    }
    return __ballot(!all_in) == 0ull;
}

// A list of n energies from src into the arena and the node's slot `out`
// tagged with round r; false (after flagging the overflow) when the arena is full.
template <int ESZ>
__device__ bool write_list(const GenArgs& a, const uint32_t* src, uint32_t n, uint4* out, uint32_t r, uint32_t lane) {
    const uint32_t esz = a.c.esz;
    const uint64_t words = (uint64_t)n * esz;
    uint32_t base = 0;
    if (lane == 0) base = words ? atomicAdd(&a.ctl->arena_top, (uint32_t)min<uint64_t>(words, 0xffffffffull)) : 0u;
    base = lane_val(base, 0);
    if ((uint64_t)base + words > a.arena_cap) {
        if (lane == 0) atomicCAS(&a.ctl->overflow, 0u, r + 1u);
        return false;
    }
    for (uint32_t k = lane; k < words; k += 64) a.arena[base + k] = src[k];
    if (lane == 0) *out = make_uint4(n, base, 0u, round_tag(r) << 16);
    return true;
}

// The scratch a node's evaluation needs does not fit the wave's region:
// record the words needed and make the round run again.
__device__ void need_scratch(const GenArgs& a, uint64_t words, uint32_t r, uint32_t lane) {
    if (lane == 0) {
        atomicMax(&a.ctl->scr_need, (uint32_t)min<uint64_t>(words, 0xffffffffull));
        atomicCAS(&a.ctl->overflow, 0u, r + 1u);
    }
}

}  // namespace

template <int ESZ>
__global__ __launch_bounds__(GEN_BLOCK) void k_eval_list(GenArgs a, uint32_t k) {
    __shared__ WaveLds s_lds[GEN_BLOCK / 64];
    StatAcc acc;
    const RoundCtl rc = read_round(a.ctl, k);
    const uint32_t r = rc.r, n = rc.n;
    if (!rc.ovf && blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl->wl_n[(r + 2u) % 3u] = 0u;      // round r - 1's count is consumed (see k_eval)
        if (n) a.ctl->last_nonempty = r + 1u;
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    WaveLds& L = s_lds[wid];
    const uint32_t gw = blockIdx.x * nw + wid, waves = gridDim.x * nw;
    uint32_t* const scr = a.scr + (size_t)gw * a.scr_words;
    const uint32_t esz = a.c.esz;
    const uint32_t t = round_tag(r);
    // this round's list leaves its queued bitmap (as in k_eval)
    uint32_t* const qcur = a.queued + (r & 1u) * a.qwords;
    const bool clear_all = (uint64_t)a.qwords < 8ull * n;
    if (clear_all)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.qwords; i += gridDim.x * blockDim.x) qcur[i] = 0u;
    const uint32_t* __restrict__ wl = a.wl[r & 1u];
    for (uint32_t idx = gw; idx < n; idx += waves) {
        const uint32_t u = uniform(wl[idx]);
        const uint4 hdr = a.st[(size_t)NODE_U4 * u];
        uint32_t slot = 0;
        const uint4 own = load_cur(a.st, u, t, slot);
        const uint32_t r0 = uniform(hdr.x), deg = uniform(hdr.y & ~ATK_BIT);
        const bool isatk = (hdr.y & ATK_BIT) != 0u;
        const uint32_t own_len = uniform(own.x & GLEN_MASK), own_off = uniform(own.y);
        uint4* const out = a.st + slot_index(u, uniform(slot) ^ 1u);
        if (!clear_all && lane == 0) atomicAnd(&qcur[u >> 5], ~(1u << (u & 31u)));
        uint64_t bytes = lane == 0 ? 9u + 4ull * esz * own_len : 0ull;   // this lane's share of the node's bytes
        bool evaluated = true, changed = false, abort = false;
        uint32_t new_len = 0;
        if (isatk) {
            // ---- attacker: successor images, then the own list; Pareto-minimal ones kept
            uint32_t C = 0;
            for (uint32_t b0 = 0; b0 < deg && !abort; b0 += 64) {
                const uint32_t j = b0 + lane;
                uint32_t len = 0, off = 0;
                if (j < deg) {
                    const uint4 d = load_cur(a.st, a.edges[r0 + j].x, t);
                    len = d.x & GLEN_MASK;
                    off = d.y;
                    bytes += 4u + 4ull * a.c.usz + 4ull * esz * len;
                }
                const uint64_t tot = lane_val(wave_incl_scan(len, lane), 63);
                const uint64_t need = ((uint64_t)C + tot + own_len) * (esz + 1u);
                if (need > a.scr_words) {
                    need_scratch(a, need, r, lane);
                    abort = true;
                    break;
                }
                C += gather_images<ESZ>(a, L, lane, len, off, r0 + j, scr, C);
            }
            if (!abort) {
                if ((uint64_t)(C + own_len) * (esz + 1u) > a.scr_words) {   // (a node without successors)
                    need_scratch(a, (uint64_t)(C + own_len) * (esz + 1u), r, lane);
                    abort = true;
                }
            }
            if (!abort) {
                const uint32_t own_start = C;
                for (uint32_t j = lane; j < own_len * esz; j += 64) scr[(size_t)C * esz + j] = a.arena[own_off + j];
                C += own_len;
                gsync();
                uint32_t* const flags = scr + (size_t)C * esz;
                bool from_succ = false;
                const uint32_t kept = pareto_flags<ESZ>(scr, C, esz, a.c, true, flags, lane, own_start, from_succ);
                gsync();
                // changed iff a minimal energy came from a successor (attack_minimize.wgsl:106-121)
                changed = from_succ;
                if (changed) {
                    compact_kept<ESZ>(scr, flags, C, esz, scr, lane);
                    if (!write_list<ESZ>(a, scr, kept, out, r, lane)) abort = true;
                    new_len = kept;
                }
            }
        } else {
            // ---- defender: skipped while a successor has no energies (src/energygame.rs:789-792)
            bool any_empty = deg == 0;
            for (uint32_t b0 = 0; b0 < deg; b0 += 64) {
                const uint32_t j = b0 + lane;
                uint32_t len = 1;
                if (j < deg) len = load_cur(a.st, a.edges[r0 + j].x, t).x & GLEN_MASK;
                any_empty |= __ballot(len == 0u) != 0ull;
            }
            if (any_empty) {
                evaluated = false;
            } else {
                // the fold: acc = images of successor 0 at scr[0 ..), then for each
                // further successor its images B after acc and the suprema after B
                uint32_t A = 0;
                for (uint32_t kk = 0; kk < deg && !abort; kk++) {
                    const uint4 d = load_cur(a.st, a.edges[r0 + kk].x, t);
                    const uint32_t len = uniform(d.x & GLEN_MASK), off = uniform(d.y);
                    bytes += lane == 0 ? 4u + 4ull * a.c.usz + 4ull * esz * len : 0u;
                    if (kk == 0) {
                        if ((uint64_t)len * (esz + 1u) > a.scr_words) {
                            need_scratch(a, (uint64_t)len * (esz + 1u), r, lane);
                            abort = true;
                            break;
                        }
                        A = gather_images<ESZ>(a, L, lane, lane == 0 ? len : 0u, off, r0, scr, 0);
                        continue;
                    }
                    const uint64_t comb = (uint64_t)A * len;
                    const uint64_t need = ((uint64_t)A + len + comb) * esz + comb;
                    if (need > a.scr_words) {
                        need_scratch(a, need, r, lane);
                        abort = true;
                        break;
                    }
                    const uint32_t Bb = A;                           // B at item A
                    gather_images<ESZ>(a, L, lane, lane == 0 ? len : 0u, off, r0 + kk, scr, Bb);
                    const uint32_t S0 = A + len;                     // suprema at item A + len
                    for (uint32_t s = lane; s < comb; s += 64) {
                        const uint32_t ia = s % A, ib = s / A;
                        const uint32_t* ea = scr + (size_t)ia * esz;
                        const uint32_t* eb = scr + (size_t)(Bb + ib) * esz;
                        uint32_t* o = scr + (size_t)(S0 + s) * esz;
                        for (uint32_t w = 0; w < esz; w++) o[w] = word_max(ea[w], eb[w], a.c);
                    }
                    gsync();
                    uint32_t* const flags = scr + (size_t)(S0 + comb) * esz;
                    bool unused = false;
                    const uint32_t kept =
                        pareto_flags<ESZ>(scr + (size_t)S0 * esz, (uint32_t)comb, esz, a.c, false, flags, lane, 0u, unused);
                    gsync();
                    compact_kept<ESZ>(scr + (size_t)S0 * esz, flags, (uint32_t)comb, esz, scr, lane);
                    A = kept;
                }
                if (!abort) {
                    // a single successor's images are copied unminimised
                    // (defend_iterative.wgsl:301-313: the fold loop does not run)
                    changed = !same_as_own<ESZ>(a, scr, A, own_len, own_off, lane);
                    if (changed) {
                        if (!write_list<ESZ>(a, scr, A, out, r, lane)) abort = true;
                        new_len = A;
                    }
                }
            }
        }
        if (abort) {
            changed = false;      // the round runs again
            evaluated = false;
        }
        if (lane == 0) {
            a.chg[idx] = changed ? 1 : 0;
            if (evaluated) {
                acc.evals++;
                acc.edges += deg;
                acc.bytes += changed ? 4ull * esz * new_len : 0ull;
                acc.changed += changed ? 1 : 0;
            }
        }
        if (evaluated) acc.bytes += bytes;
        gsync();
    }
    flush_stats(acc, a.stats);
}
template __global__ void k_eval_list<1>(GenArgs, uint32_t);
template __global__ void k_eval_list<4>(GenArgs, uint32_t);
template __global__ void k_eval_list<32>(GenArgs, uint32_t);

// t: tag of the round after the last one that ran
__global__ void k_list_lens_g(const uint4* st, uint32_t n, uint32_t t, uint32_t* lens, const uint32_t* map) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) lens[v] = load_cur(st, map ? map[v] : v, t).x & GLEN_MASK;
}

// Every node's current list, esz words per energy, at offs[v] (in energies);
// the host puts each list in ascending order.
__global__ void k_gather_g(const uint4* st, const uint32_t* arena, uint32_t n, uint32_t t, uint32_t esz, const uint32_t* offs,
                           uint32_t* out, const uint32_t* map) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint4 d = load_cur(st, map ? map[v] : v, t);
    const uint64_t words = (uint64_t)(d.x & GLEN_MASK) * esz;
    uint32_t* o = out + (size_t)offs[v] * esz;
    for (uint64_t k = 0; k < words; k++) o[k] = arena[d.y + k];
}

}  // namespace geq
