// Fixpoint kernels of the energy-game solver: one wave per worklist node.
#include "solve_kernels.hpp"

namespace geq {

// Write the members of a 4096-point bitmap (one u64 word per lane) as an
// ascending list: inline into the node's record (slot) when it fits, else
// into the arena; the record is tagged with round r.  Returns the list length
// (uniform).  On arena exhaustion sets ctl->overflow to the round + 1 and
// returns ~0u.
__device__ __forceinline__ uint32_t write_sorted(uint64_t M, uint32_t lane, uint4* rec, uint16_t* arena, Ctl* ctl,
                                                 uint32_t arena_cap, uint32_t r) {
    const uint32_t c = (uint32_t)__popcll(M);
    const uint32_t incl = wave_incl_scan(c, lane);
    const uint32_t total = lane_val(incl, 63);
    uint32_t p = incl - c;
    if (total <= INLINE_CAP) {
        uint16_t* h = reinterpret_cast<uint16_t*>(rec);
        uint64_t m = M;
        while (m) {
            uint32_t b = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            h[1 + p] = (uint16_t)(64u * lane + b);
            p++;
        }
        if (lane == 0) {
            h[0] = (uint16_t)total;
            h[7] = (uint16_t)round_tag(r);
        }
        return total;
    }
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&ctl->arena_top, total);
    base = lane_val(base, 0);
    if ((uint64_t)base + total > arena_cap) {
        if (lane == 0) atomicCAS(&ctl->overflow, 0u, r + 1u);
        return ~0u;
    }
    uint64_t m = M;
    while (m) {
        uint32_t b = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        arena[base + p] = (uint16_t)(64u * lane + b);
        p++;
    }
    if (lane == 0) *rec = make_uint4(total, base, 0u, round_tag(r) << 16);
    return total;
}

// Successors of the node being evaluated, read straight from HBM (nodes with
// more successors than the wave's staging area) ...
struct SuccGlobal {
    const uint2* __restrict__ edges;
    const uint4* __restrict__ st;
    uint32_t r0;
    uint32_t t;                     // tag of the round
    __device__ __forceinline__ void get(uint32_t j, uint4& d, uint32_t& w) const {
        const uint2 ed = edges[r0 + j];
        w = ed.y;
        d = load_cur(st, ed.x, t);
    }
};
// ... or from the wave's LDS staging area, gathered for a run of worklist nodes at once.
struct SuccStaged {
    const uint4* rec;
    const uint32_t* w;
    uint32_t off;
    __device__ __forceinline__ void get(uint32_t j, uint4& d, uint32_t& ww) const {
        d = rec[off + j];
        ww = w[off + j];
    }
};

// Evaluate node u (worklist slot idx): the attacker merge or the defender
// combination, writing a changed list into `out` (u's non-current slot) and,
// unless FUSED, the changed flag chg[idx].  Returns whether u changed
// (uniform).  Successor j's record and update come from `src`; everything
// else is the same for both sources.
template <bool FUSED, class Src>
__device__ __forceinline__ bool eval_node(const EvalArgs& a, uint32_t r, uint32_t* A, uint32_t* B, uint32_t* O, uint32_t lane, uint32_t idx,
                                          uint32_t deg, bool isatk, const uint4 own, uint4* out, const Src& src, StatAcc& acc,
                                          uint64_t& st_bytes_lane) {
    const uint32_t maxrep = a.maxv * 0x555u;
    const uint16_t* __restrict__ arena = a.arena;
    const uint32_t own_len = uniform(nl_len(own));

    A[lane] = 0u; A[lane + 64] = 0u;
    O[lane] = 0u; O[lane + 64] = 0u;
    wave_sync();
    for (uint32_t j = lane; j < own_len; j += 64) {
        uint32_t e = nl_get(own, arena, j);
        bm_set(O, e);
        if (isatk) bm_set(A, e);   // own slice: the identity update
    }

    bool changed = false;
    bool evaluated = true;
    uint32_t new_len = 0;
    if (isatk) {
        // ---- attacker: OR of the inverse-updated successor lists + own list
        for (uint32_t base = 0; base < deg; base += 64) {
            const uint32_t k = base + lane;
            uint32_t len = 0, w = 0;
            uint4 d = make_uint4(0, 0, 0, 0);
            if (k < deg) {
                src.get(k, d, w);
                len = nl_len(d);
                st_bytes_lane += 8u + 4u * len;
            }
            const InvUpd pu = prep_update(w);
            if (len <= INLINE_CAP) {
                for (uint32_t j = 0; j < len; j++) bm_set(A, inv_apply(nl_inline(d, j), pu, maxrep));
            }
            uint64_t longm = __ballot(len > INLINE_CAP);
            while (longm) {
                const uint32_t s = (uint32_t)__builtin_ctzll(longm);
                longm &= longm - 1;
                const InvUpd lpu = prep_update(lane_val(w, s));
                const uint32_t loff = lane_val(d.y, s);
                const uint32_t llen = lane_val(len, s);
                for (uint32_t j = lane; j < llen; j += 64) bm_set(A, inv_apply(arena[loff + j], lpu, maxrep));
            }
        }
        wave_sync();
        const uint64_t W = up_close(bm_word(A, lane), lane);
        const uint64_t M = minimal_of(W, lane);
        const uint64_t Ow = bm_word(O, lane);
        // changed iff a minimal budget is not already one of the node's own
        // energies (attack_minimize.wgsl:208-214: a survivor from a successor slice)
        changed = __ballot((M & ~Ow) != 0ull) != 0ull;
        if (changed) new_len = write_sorted(M, lane, out, a.arena, a.ctl, a.arena_cap, r);
    } else {
        // ---- defender: skipped while any successor has no energies (src/energygame.rs:789-792)
        bool any_empty = false;
        for (uint32_t base = 0; base < deg; base += 64) {
            const uint32_t k = base + lane;
            uint32_t len = 1;
            if (k < deg) {
                uint4 d;
                uint32_t w;
                src.get(k, d, w);
                len = nl_len(d);
            }
            any_empty |= (__ballot(len == 0) != 0ull);
        }
        if (any_empty || deg == 0) {
            evaluated = false;
        } else if (deg == 1) {
            // A single successor is copied, not minimised
            // (defend_iterative.wgsl:301-313 and the empty fold loop).
            uint4 d;
            uint32_t w;
            src.get(0, d, w);
            const InvUpd pu = prep_update(uniform(w));
            const uint32_t len = uniform(nl_len(d));
            st_bytes_lane += (lane == 0) ? 8u + 4u * len : 0u;
            bool all_in = true;
            for (uint32_t j = lane; j < len; j += 64) all_in &= bm_test(O, inv_apply(nl_get(d, arena, j), pu, maxrep));
            const bool same = (len == own_len) && (__ballot(!all_in) == 0ull);
            changed = !same;
            if (changed) {
                uint4* rec = out;
                if (len <= INLINE_CAP) {
                    uint16_t* h = reinterpret_cast<uint16_t*>(rec);
                    if (lane < len) h[1 + lane] = (uint16_t)inv_apply(nl_inline(d, lane), pu, maxrep);
                    if (lane == 0) {
                        h[0] = (uint16_t)(len | (len > 1 ? UNSORTED : 0u));
                        h[7] = (uint16_t)round_tag(r);
                    }
                    new_len = len;
                } else {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&a.ctl->arena_top, len);
                    base = lane_val(base, 0);
                    if ((uint64_t)base + len > a.arena_cap) {
                        if (lane == 0) atomicCAS(&a.ctl->overflow, 0u, r + 1u);
                        new_len = ~0u;
                    } else {
                        for (uint32_t j = lane; j < len; j += 64)
                            a.arena[base + j] = (uint16_t)inv_apply(arena[nl_off(d) + j], pu, maxrep);
                        if (lane == 0) *rec = make_uint4(len | UNSORTED, base, 0u, round_tag(r) << 16);
                        new_len = len;
                    }
                }
            }
        } else {
            // AND of the up-closures of every successor's inverse-updated list
            uint64_t acc_bm = ~0ull;
            for (uint32_t base = 0; base < deg; base += 64) {
                const uint32_t k = base + lane;
                uint32_t w = 0;
                uint4 d = make_uint4(0, 0, 0, 0);
                if (k < deg) {
                    src.get(k, d, w);
                    st_bytes_lane += 8u + 4u * nl_len(d);
                }
                const uint32_t nk = min(64u, deg - base);
                for (uint32_t s = 0; s < nk; s++) {
                    const InvUpd spu = prep_update(lane_val(w, s));
                    uint4 ds;
                    ds.x = lane_val(d.x, s);
                    ds.y = lane_val(d.y, s);
                    ds.z = lane_val(d.z, s);
                    ds.w = lane_val(d.w, s);
                    const uint32_t ls = nl_len(ds);
                    B[lane] = 0u; B[lane + 64] = 0u;
                    wave_sync();
                    for (uint32_t j = lane; j < ls; j += 64) bm_set(B, inv_apply(nl_get(ds, arena, j), spu, maxrep));
                    wave_sync();
                    acc_bm &= up_close(bm_word(B, lane), lane);
                    wave_sync();
                }
            }
            const uint64_t M = minimal_of(acc_bm, lane);
            const uint64_t Ow = bm_word(O, lane);
            const uint32_t total = uniform(wave_sum((uint32_t)__popcll(M)));
            // equality test of src/energygame.rs:929 (same size, new within old)
            const bool same = (total == own_len) && (__ballot((M & ~Ow) != 0ull) == 0ull);
            changed = !same;
            if (changed) new_len = write_sorted(M, lane, out, a.arena, a.ctl, a.arena_cap, r);
        }
    }
    if (new_len == ~0u) changed = false;   // arena overflow: the round runs again
    if (lane == 0) {
        if (!FUSED) a.chg[idx] = changed ? 1 : 0;
        if (evaluated) {
            acc.evals++;
            acc.edges += deg;
            acc.bytes += 9u + 4u * own_len + (changed ? 4u * new_len : 0u);
            acc.changed += changed ? 1 : 0;
        }
    }
    if (evaluated) acc.bytes += st_bytes_lane;
    st_bytes_lane = 0;
    wave_sync();
    return changed;
}

// Append the not-yet-queued predecessors of the lanes' nodes
// (GPURunner::changed_nodes, src/energygame.rs:1467-1479): lane k's node has
// its pc predecessors at rcol[p0 ..] (pc = 0: nothing to queue).  The
// predecessor slots of all lanes are walked 64 at a time (one lane per slot).
// A predecessor is queued once per list: the lane whose atomicOr sets its bit
// in the target list's bitmap q appends it.  Appends go to the wave's LDS
// stage; a stage past 192 entries is flushed by the wave with one atomic on
// *out_n, the rest waits for flush_appends.  off / rs: 64-entry LDS scratch.
__device__ __forceinline__ void expand_nodes(const uint32_t* __restrict__ rcol, uint32_t* q, uint32_t p0, uint32_t pc, uint32_t lane,
                                             uint32_t* off, uint32_t* rs, uint32_t* stage, uint32_t& cnt, uint32_t* out,
                                             uint32_t* out_n, bool bits_only = false) {
    const uint32_t incl = wave_incl_scan(pc, lane);
    const uint32_t total = lane_val(incl, 63);
    if (total == 0) return;
    off[lane] = incl - pc;
    rs[lane] = p0;
    wave_sync();
    for (uint32_t s0 = 0; s0 < total; s0 += 64) {
        const uint32_t s = s0 + lane;
        bool win = false;
        uint32_t w = 0;
        if (s < total) {
            uint32_t lo = 0, hi = 63;   // off[0] = 0 <= s: the largest lane with off <= s owns slot s
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (off[mid] <= s) lo = mid; else hi = mid - 1;
            }
            w = rcol[rs[lo] + (s - off[lo])];
            const uint32_t bit = 1u << (w & 31u);
            win = (atomicOr(&q[w >> 5], bit) & bit) == 0u;
        }
        if (bits_only) continue;   // k_collect lists the queued nodes after the round
        const uint64_t bal = __ballot(win);
        if (win) stage[cnt + mbcnt(bal)] = w;
        cnt += (uint32_t)__popcll(bal);
        if (cnt > APPEND_STAGE - 64) {
            wave_sync();
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(out_n, cnt);
            base = lane_val(base, 0);
            for (uint32_t j = lane; j < cnt; j += 64) out[base + j] = stage[j];
            cnt = 0;
            wave_sync();
        }
    }
    wave_sync();   // off / rs are rewritten by the next use
}

// Block-level flush of every wave's remaining appends: one atomic per block.
// Every thread of the block calls it.
__device__ __forceinline__ void flush_appends(const uint32_t* stage, uint32_t cnt, uint32_t* s_cnt, uint32_t* out, uint32_t* out_n) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    if (lane == 0) s_cnt[wid] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; w++) tot += s_cnt[w];
        s_cnt[MAX_WAVES] = tot ? atomicAdd(out_n, tot) : 0u;
    }
    __syncthreads();
    uint32_t base = s_cnt[MAX_WAVES];
    for (uint32_t w = 0; w < wid; w++) base += s_cnt[w];
    for (uint32_t j = lane; j < cnt; j += 64) out[base + j] = stage[j];
    __syncthreads();
}

// One round's evaluations.  A wave takes up to 64 worklist entries at a time:
// one lane per node loads the node's row, flag and current record, then the
// successor records of a run of these nodes are gathered into the wave's LDS
// staging area with all lanes' loads in flight together, and the nodes are
// evaluated one after the other from LDS.  A node with more successors than
// the staging area holds reads them from HBM directly.  FUSED: the wave then
// appends the predecessors of its changed nodes to the next worklist.
// r: the round; n: its worklist size (0 after an arena overflow).
// lds: the block's dynamic LDS, EVAL_(FUSED_)LDS_PER_WAVE bytes per wave.
// P: the round's list, where its appends go and which waves share it.
template <bool FUSED>
__device__ __forceinline__ void eval_phase(const EvalArgs& a, unsigned char* lds, uint32_t* s_cnt, StatAcc& acc, uint32_t r,
                                           uint32_t n, const PhaseList& P) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = threadIdx.x >> 6;
    unsigned char* mine = lds + (size_t)wid * (FUSED ? EVAL_FUSED_LDS_PER_WAVE : EVAL_LDS_PER_WAVE);
    uint4* rec = reinterpret_cast<uint4*>(mine);                          // staged successor records
    uint32_t* wts = reinterpret_cast<uint32_t*>(mine + STAGE * 16);        // staged successor updates
    uint32_t* ex = wts + STAGE;                                           // first staging slot of each node of the run
    uint32_t* rs = ex + 64;                                               // first edge of each node of the run
    uint32_t* A = rs + 64;        // union (attacker) bitmap
    uint32_t* B = A + 128;        // one successor's images (defender)
    uint32_t* O = B + 128;        // old list of the node
    uint32_t* stage = O + 128;    // FUSED: staged appends (APPEND_STAGE entries)

    const uint32_t t = round_tag(r);
    // This round's list leaves its queued bitmap: a large list clears the
    // whole bitmap with plain stores (grid-stride), a short one clears its
    // nodes' bits one atomic each.
    uint32_t* const qcur = a.queued + (r & 1u) * a.qwords;
    const bool clear_all = P.grid_wide && (uint64_t)a.qwords < 8ull * n;
    if (clear_all)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.qwords; i += gridDim.x * blockDim.x) qcur[i] = 0u;
    uint32_t* out = P.out;
    uint32_t* out_n = P.out_n;
    uint32_t app = 0;             // FUSED: appends staged by this wave
    uint64_t st_bytes_lane = 0;
    const uint32_t* __restrict__ wl = P.wl;
    // entries per wave visit: up to 64, fewer when the worklist is too short
    // to give every wave that shares it a full set
    const uint32_t waves = P.waves;
    const uint32_t chunk = max(1u, min(64u, (n + waves - 1) / max(waves, 1u)));
    for (uint32_t base = uniform(P.wave0 * chunk); base < n; base += waves * chunk) {
        const uint32_t cnt = min(chunk, n - base);
        // lane k: worklist entry base + k
        uint32_t u = 0, r0 = 0, deg = 0, slot = 0, p0 = 0, pc = 0;
        bool isatk = false, chg = false;
        uint4 own = make_uint4(0, 0, 0, 0);
        if (lane < cnt) {
            u = wl[base + lane];
            // header and both slots: one 64-byte record
            const uint4 hdr = a.st[(size_t)NODE_U4 * u];
            own = load_cur(a.st, u, t, slot);
            r0 = hdr.x;
            deg = hdr.y & ~ATK_BIT;
            isatk = (hdr.y & ATK_BIT) != 0u;
            p0 = hdr.z;             // the predecessor range, in case u changes
            pc = hdr.w;
            if (!clear_all) atomicAnd(&qcur[u >> 5], ~(1u << (u & 31u)));
        }
        uint32_t k0 = 0;
        while (k0 < cnt) {
            // the run [k0, k1): consecutive nodes whose successors fit the staging area
            const uint32_t d = (lane >= k0 && lane < cnt) ? deg : 0u;
            const uint32_t incl = wave_incl_scan(d, lane);
            const uint64_t over = __ballot(lane >= k0 && lane < cnt && incl > STAGE);
            const uint32_t k1 = over ? (uint32_t)__builtin_ctzll(over) : cnt;
            if (k1 == k0) {
                // node k0 alone has more successors than the staging area
                const uint32_t vr0 = lane_val(r0, k0);
                const uint32_t vdeg = lane_val(deg, k0);
                const bool vatk = lane_val(isatk ? 1u : 0u, k0) != 0u;
                uint4 vown;
                vown.x = lane_val(own.x, k0); vown.y = lane_val(own.y, k0); vown.z = lane_val(own.z, k0); vown.w = lane_val(own.w, k0);
                // the node's new list goes to its non-current slot
                uint4* vout = a.st + slot_index(lane_val(u, k0), lane_val(slot, k0) ^ 1u);
                const bool c = eval_node<FUSED>(a, r, A, B, O, lane, base + k0, vdeg, vatk, vown, vout,
                                                SuccGlobal{a.edges, a.st, vr0, t}, acc, st_bytes_lane);
                chg |= c && lane == k0;
                k0++;
                continue;
            }
            const uint32_t total = lane_val(incl, k1 - 1);
            if (lane >= k0 && lane < k1) {
                ex[lane] = incl - d;
                rs[lane] = r0;
            }
            wave_sync();
            // gather: slot s belongs to the last node of the run whose first slot is <= s;
            // both slots of a lane issue their loads before either is waited for
            static_assert(STAGE == 128, "the gather below handles two slots per lane");
            auto edge_of = [&](uint32_t s) -> uint2 {
                uint32_t lo = k0, hi = k1 - 1;   // ex[lo] <= s: find the largest such index
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (ex[mid] <= s) lo = mid; else hi = mid - 1;
                }
                return a.edges[rs[lo] + (s - ex[lo])];   // (LDS, not a lane shuffle: lane lo may be inactive here)
            };
            const uint32_t s0 = lane, s1 = 64u + lane;
            const uint2 ed0 = s0 < total ? edge_of(s0) : make_uint2(0u, 0u);
            const uint2 ed1 = s1 < total ? edge_of(s1) : make_uint2(0u, 0u);
            const uint4 rc0 = s0 < total ? load_cur(a.st, ed0.x, t) : make_uint4(0, 0, 0, 0);
            const uint4 rc1 = s1 < total ? load_cur(a.st, ed1.x, t) : make_uint4(0, 0, 0, 0);
            if (s0 < total) {
                rec[s0] = rc0;
                wts[s0] = ed0.y;
            }
            if (s1 < total) {
                rec[s1] = rc1;
                wts[s1] = ed1.y;
            }
            wave_sync();
            for (uint32_t k = k0; k < k1; k++) {
                const uint32_t vdeg = lane_val(deg, k);
                const bool vatk = lane_val(isatk ? 1u : 0u, k) != 0u;
                uint4 vown;
                vown.x = lane_val(own.x, k); vown.y = lane_val(own.y, k); vown.z = lane_val(own.z, k); vown.w = lane_val(own.w, k);
                uint4* vout = a.st + slot_index(lane_val(u, k), lane_val(slot, k) ^ 1u);
                const uint32_t off = uniform(ex[k]);
                const bool c = eval_node<FUSED>(a, r, A, B, O, lane, base + k, vdeg, vatk, vown, vout, SuccStaged{rec, wts, off},
                                                acc, st_bytes_lane);
                chg |= c && lane == k;
            }
            k0 = k1;
        }
        if (FUSED) {
            wave_sync();   // ex / rs of the last run are reused as expansion scratch
            expand_nodes(a.rcol, a.queued + ((r + 1u) & 1u) * a.qwords, p0, chg ? pc : 0u, lane, ex, rs, stage, app, out, out_n,
                         a.ordered != 0u);
        }
    }
    if (FUSED) flush_appends(stage, app, s_cnt, out, out_n);
}

// k: the kernel's round within the current batch of rounds (round = ctl->round + k).
template <bool FUSED>
__global__ __launch_bounds__(MAX_BLOCK) void k_eval(EvalArgs a, uint32_t k) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ uint32_t s_cnt[MAX_WAVES + 1];
    StatAcc acc;
    const RoundCtl rc = read_round(a.ctl, k);
    const uint32_t r = rc.r, n = rc.n;
    const bool ovf = rc.ovf;
    if (!ovf && blockIdx.x == 0 && threadIdx.x == 0) {
        // round r - 1's list count is consumed (its k_eval and k_expand are
        // done): it becomes the count of round r + 1's appends
        a.ctl->wl_n[(r + 2u) % 3u] = 0u;
        if (n) a.ctl->last_nonempty = r + 1u;
    }
    const uint32_t nw = blockDim.x >> 6;
    const PhaseList P{a.wl[r & 1u], a.wl[(r + 1u) & 1u], &a.ctl->wl_n[(r + 1u) % 3u], blockIdx.x * nw + (threadIdx.x >> 6),
                      gridDim.x * nw, true};
    eval_phase<FUSED>(a, dyn_lds, s_cnt, acc, r, n, P);
    flush_stats(acc, a.stats);
}
template __global__ void k_eval<false>(EvalArgs, uint32_t);
template __global__ void k_eval<true>(EvalArgs, uint32_t);

namespace {

// thread 0 reads a word that other blocks may write; every thread gets its
// value (one barrier on each side, so the block's control flow stays uniform)
__device__ __forceinline__ uint32_t block_read(const uint32_t* p, uint32_t* s_word) {
    if (threadIdx.x == 0) *s_word = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t v = *s_word;
    __syncthreads();
    return v;
}

// A game's first worklist: the predecessors of its reach seeds (as k_expand's
// seeding mode), queued on the parity-0 bitmap.  Uses each wave's expansion
// scratch and append stage of the fused k_eval layout.
__device__ void seed_game(const EvalArgs& a, unsigned char* lds, uint32_t* s_cnt, const uint32_t* seeds, uint32_t ns,
                          uint32_t* out, uint32_t* out_n) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned char* mine = lds + (size_t)wid * EVAL_FUSED_LDS_PER_WAVE;
    uint32_t* ex = reinterpret_cast<uint32_t*>(mine + STAGE * 16) + STAGE;
    uint32_t* rs = ex + 64;
    uint32_t* stage = rs + 64 + 3 * 128;
    uint32_t app = 0;
    for (uint32_t b = wid * 64u; b < ns; b += nw * 64u) {
        uint32_t p0 = 0, pc = 0;
        if (b + lane < ns) {
            const uint4 hdr = a.st[(size_t)NODE_U4 * seeds[b + lane]];
            p0 = hdr.z;
            pc = hdr.w;
        }
        expand_nodes(a.rcol, a.queued, p0, pc, lane, ex, rs, stage, app, out, out_n);
    }
    flush_appends(stage, app, s_cnt, out, out_n);
}

}  // namespace

// One launch for a batch: blocks take games from the queue (ga.order) and run
// each game's rounds to its fixpoint, one block barrier between rounds.  The
// rounds of a game are the Jacobi rounds of the game alone, which are also its
// rounds inside the disjoint union that k_eval runs: same lists, same
// evaluations.  After an arena overflow (set by any block) a block stops its
// game at the end of the current round, records that round to run again, and
// starts no further game; the host grows the arena and launches again.
__global__ __launch_bounds__(MAX_BLOCK) void k_solve_games(EvalArgs a, GamesArgs ga) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ uint32_t s_cnt[MAX_WAVES + 1];
    __shared__ uint32_t s_word, s_next_n;
    StatAcc acc;
    const uint32_t nw = blockDim.x >> 6, wid = threadIdx.x >> 6;
    uint32_t last = 0;   // 1 + the last non-empty round of this block's games
    while (true) {
        if (threadIdx.x == 0) s_word = atomicAdd(ga.next, 1u);
        __syncthreads();
        const uint32_t q = s_word;
        __syncthreads();
        if (q >= ga.n_games) break;
        const uint32_t g = ga.order[q];
        GameCtl c = ga.gctl[g];
        if (c.state == GAME_DONE) continue;
        if (block_read(&a.ctl->overflow, &s_word)) continue;   // left for the next launch
        const uint32_t nb = ga.base[g], ne = ga.base[g + 1];
        uint32_t* const wl0 = a.wl[0] + nb;
        uint32_t* const wl1 = a.wl[1] + nb;
        uint32_t r = 0, n = 0;
        if (c.state == GAME_NEW) {
            if (threadIdx.x == 0) s_next_n = 0u;
            __syncthreads();
            seed_game(a, dyn_lds, s_cnt, ga.seed + ga.soff[g], ga.soff[g + 1] - ga.soff[g], wl0, &s_next_n);
            n = s_next_n;   // flush_appends ended with a barrier
            __syncthreads();   // every thread has n before s_next_n is reset
            c.retag = 0;
        } else {
            r = c.round;
            n = c.n;
        }
        bool paused = false;
        while (n) {
            if (r - c.retag >= ga.retag_interval) {
                // k_retag over the game's nodes, between two of its rounds
                const uint32_t t = round_tag(r);
                for (uint32_t v = nb + threadIdx.x; v < ne; v += blockDim.x) {
                    const uint4 x = a.st[slot_index(v, 0)], y = a.st[slot_index(v, 1)];
                    const uint32_t cs = cur_slot(x, y, t);
                    a.st[slot_index(v, cs)] = nl_with_tag(cs ? y : x, t - 1u);
                    a.st[slot_index(v, cs ^ 1u)] = nl_with_tag(cs ? x : y, t - 2u);
                }
                c.retag = r;
            }
            if (threadIdx.x == 0) s_next_n = 0u;
            __syncthreads();
            const PhaseList P{(r & 1u) ? wl1 : wl0, (r & 1u) ? wl0 : wl1, &s_next_n, wid, nw, false};
            eval_phase<true>(a, dyn_lds, s_cnt, acc, r, n, P);   // ends with a barrier
            last = max(last, r + 1u);
            const uint32_t next_n = s_next_n;   // (block_read's barriers order this read before the reset)
            if (block_read(&a.ctl->overflow, &s_word)) {
                paused = true;   // round r runs again
                break;
            }
            n = next_n;
            r++;
        }
        if (!paused && r >= ga.retag_interval) {
            // a long game: its slots re-tagged to the epoch of round 0 (current
            // slot tag 0, the other 0xffff), so that the download's round tag,
            // min(rounds, retag_interval) + 1, reads every game's last lists
            const uint32_t t = round_tag(0);
            for (uint32_t v = nb + threadIdx.x; v < ne; v += blockDim.x) {
                const uint4 x = a.st[slot_index(v, 0)], y = a.st[slot_index(v, 1)];
                const uint32_t cs = cur_slot(x, y, round_tag(r));
                a.st[slot_index(v, cs)] = nl_with_tag(cs ? y : x, t - 1u);
                a.st[slot_index(v, cs ^ 1u)] = nl_with_tag(cs ? x : y, t - 2u);
            }
        }
        if (threadIdx.x == 0) {
            if (paused) {
                c.state = GAME_PAUSED;
                c.round = r;
                c.n = n;
            } else {
                c.state = GAME_DONE;
            }
            ga.gctl[g] = c;
        }
    }
    if (threadIdx.x == 0 && last) atomicMax(&a.ctl->last_nonempty, last);
    flush_stats(acc, a.stats);
}

__device__ __forceinline__ uint32_t game_of(const uint32_t* base, uint32_t n_games, uint32_t v) {
    uint32_t lo = 0, hi = n_games - 1;   // the last game whose first node is <= v
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (base[mid] <= v) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Seeds of one game tend to sit next to each other: lanes group into runs of
// equal games (a run starts where the game differs from the previous lane's)
// and the first lane of a run does the run's one atomic.  Lanes past n get
// game ~0u, a run of their own that does nothing.
struct SeedRun {
    uint32_t g, head, len;   // this lane's game, its run's first lane and length
};
__device__ __forceinline__ SeedRun seed_run(const uint32_t* seed, uint32_t n, const uint32_t* base, uint32_t n_games, uint32_t i) {
    const uint32_t lane = threadIdx.x & 63u;
    SeedRun s;
    s.g = i < n ? game_of(base, n_games, seed[i]) : ~0u;
    const uint32_t prev = (uint32_t)__shfl_up((int)s.g, 1u, 64);
    const uint64_t heads = __ballot(lane == 0u || prev != s.g);
    const uint64_t upto = heads & (lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull));
    s.head = 63u - (uint32_t)__builtin_clzll(upto);
    const uint64_t after = lane == 63u ? 0ull : heads >> (s.head + 1u);
    s.len = after ? (uint32_t)__builtin_ctzll(after) + 1u : 64u - s.head;
    return s;
}

__global__ void k_seed_count(const uint32_t* seed, uint32_t n, const uint32_t* base, uint32_t n_games, uint32_t* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const SeedRun s = seed_run(seed, n, base, n_games, i);
    if (s.g != ~0u && (threadIdx.x & 63u) == s.head) atomicAdd(&cnt[s.g], s.len);
}

__global__ void k_seed_scatter(const uint32_t* seed, uint32_t n, const uint32_t* base, uint32_t n_games, const uint32_t* soff,
                               uint32_t* fill, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const SeedRun s = seed_run(seed, n, base, n_games, i);
    uint32_t at = 0;
    if (s.g != ~0u && lane == s.head) at = soff[s.g] + atomicAdd(&fill[s.g], s.len);
    at = (uint32_t)__shfl((int)at, s.head, 64);   // the run's first slot, from its head
    if (s.g != ~0u) out[at + (lane - s.head)] = seed[i];
}

// The predecessors of round r's changed nodes form round r + 1's worklist; in
// seeding mode the seeds' predecessors join round r's own worklist.  A wave
// takes 64 worklist entries at a time.
// k: the kernel's round within the current batch of rounds (round = ctl->round + k).
__global__ __launch_bounds__(MAX_BLOCK) void k_expand(ExpandArgs a, uint32_t k) {
    __shared__ uint32_t s_stage[MAX_WAVES][APPEND_STAGE];
    __shared__ uint32_t s_off[MAX_WAVES][64], s_rs[MAX_WAVES][64];
    __shared__ uint32_t s_cnt[MAX_WAVES + 1];
    const RoundCtl rc = read_round(a.ctl, k);
    const uint32_t r = rc.r;
    const bool seeding = a.seed != nullptr;
    const uint32_t n = rc.ovf ? 0u : (seeding ? a.n_seed : rc.n);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t* list = seeding ? a.seed : a.wl[r & 1u];
    const uint32_t tr = seeding ? r : r + 1u;   // the round whose worklist receives the appends
    uint32_t* out = a.wl[tr & 1u];
    uint32_t* out_n = &a.ctl->wl_n[tr % 3u];
    uint32_t cnt = 0;
    const uint32_t nw = blockDim.x >> 6;
    const uint32_t gstride = gridDim.x * nw * 64;
    for (uint32_t b = (blockIdx.x * nw + wid) * 64; b < n; b += gstride) {
        const uint32_t i = b + lane;
        uint32_t p0 = 0, pc = 0;
        if (i < n) {
            const uint32_t u = list[i];
            if (seeding || a.chg[i] != 0) {
                p0 = a.rrow[u];
                pc = a.rrow[u + 1] - p0;
            }
        }
        expand_nodes(a.rcol, a.queued + (tr & 1u) * a.qwords, p0, pc, lane, s_off[wid], s_rs[wid], s_stage[wid], cnt, out, out_n);
    }
    flush_appends(s_stage[wid], cnt, s_cnt, out, out_n);
}

// After round r's evaluations (ordered worklists): the nodes queued for round
// r + 1 (set bits of bitmap (r + 1) & 1) become its worklist in id order.
// Block b lists the nodes of bitmap words [b * W, (b + 1) * W), W =
// COLLECT_THREADS * COLLECT_WORDS, with one atomic for the block's room.
__global__ __launch_bounds__(COLLECT_THREADS) void k_collect(CollectArgs a, uint32_t k) {
    __shared__ uint32_t s_w[COLLECT_THREADS / 64 + 1];
    const RoundCtl rc = read_round(a.ctl, k);
    if (rc.ovf) return;
    const uint32_t tr = rc.r + 1u;
    const uint32_t* __restrict__ q = a.queued + (tr & 1u) * a.qwords;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const size_t w0 = ((size_t)blockIdx.x * COLLECT_THREADS + threadIdx.x) * COLLECT_WORDS;
    uint32_t bits[COLLECT_WORDS];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < COLLECT_WORDS; i++) {
        bits[i] = w0 + i < a.qwords ? q[w0 + i] : 0u;
        c += (uint32_t)__popc(bits[i]);
    }
    const uint32_t incl = wave_incl_scan(c, lane);
    if (lane == 63u) s_w[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < COLLECT_THREADS / 64; w++) tot += s_w[w];
        s_w[COLLECT_THREADS / 64] = tot ? atomicAdd(&a.ctl->wl_n[tr % 3u], tot) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_w[COLLECT_THREADS / 64] + incl - c;
    for (uint32_t w = 0; w < wid; w++) pos += s_w[w];
    uint32_t* out = a.wl[tr & 1u];
#pragma unroll
    for (uint32_t i = 0; i < COLLECT_WORDS; i++) {
        uint32_t m = bits[i];
        while (m) {
            const uint32_t b = (uint32_t)__builtin_ctz(m);
            m &= m - 1u;
            out[pos++] = (uint32_t)((w0 + i) * 32u + b);
        }
    }
}

// End of a batch of `rounds` rounds: the round advances past them, or, after
// an arena overflow, stops at the round that overflowed (it runs again).
__global__ void k_advance(Ctl* ctl, uint32_t rounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t o = ctl->overflow;
    ctl->round = o ? o - 1u : ctl->round + rounds;
}

// Between rounds (nothing in flight): give every node's current slot age 1
// and its other slot age 2 at round tag t, keeping which one is current.
__global__ void k_retag(uint4* st, uint32_t n, uint32_t t) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint4 a = st[slot_index(v, 0)], b = st[slot_index(v, 1)];
    const uint32_t c = cur_slot(a, b, t);
    st[slot_index(v, c)] = nl_with_tag(c ? b : a, t - 1u);
    st[slot_index(v, c ^ 1u)] = nl_with_tag(c ? a : b, t - 2u);
}

// Node headers from the CSR arrays (after an upload or a placement).
__global__ void k_node_hdr(const uint32_t* row, const uint32_t* rrow, const uint8_t* atk, uint32_t n, uint4* st) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t r0 = row[v], p0 = rrow[v];
    st[(size_t)NODE_U4 * v] = make_uint4(r0, (row[v + 1] - r0) | (atk[v] ? ATK_BIT : 0u), p0, rrow[v + 1] - p0);
    st[(size_t)NODE_U4 * v + 3] = make_uint4(0, 0, 0, 0);
}


// Sum the per-block statistics slots into out[0..3].
__global__ void k_reduce_stats(const unsigned long long* slots, uint32_t nslots, unsigned long long* out) {
    __shared__ unsigned long long s[4][256];
    for (int c = 0; c < 4; c++) {
        unsigned long long v = 0;
        for (uint32_t b = threadIdx.x; b < nslots; b += blockDim.x) v += slots[4ull * b + c];
        s[c][threadIdx.x] = v;
    }
    __syncthreads();
    for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int c = 0; c < 4; c++) s[c][threadIdx.x] += s[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}

// t: tag of the round after the last one that ran
__global__ void k_list_lens(const uint4* st, uint32_t n, uint32_t t, uint32_t* lens, const uint32_t* map) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) lens[v] = nl_len(load_cur(st, map ? map[v] : v, t));
}

// Preorder verdicts (src/equivalence.rs:125-132, EnergyGame's budgets of the
// starting points against each equivalence's energy): out[v * n_eqs + k] = 1
// when no energy of node v's list lies below eq.w[k] in every field (2-bit
// fields, `fields` of them), else 0.  Entry v is node map[v] (v without a map).
__global__ void k_preorders(const uint4* st, const uint16_t* arena, uint32_t n, uint32_t t, const uint32_t* map, EqWords eq,
                            uint32_t fields, uint8_t* out) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint4 d = load_cur(st, map ? map[v] : v, t);
    const uint32_t len = nl_len(d);
    for (uint32_t k = 0; k < eq.n; k++) {
        const uint32_t e = eq.w[k];
        uint32_t holds = 1;
        for (uint32_t j = 0; j < len && holds; j++) {
            const uint32_t b = nl_get(d, arena, j);
            bool below = true;
            for (uint32_t f = 0; f < fields; f++) below = below && ((b >> (2 * f)) & 3u) <= ((e >> (2 * f)) & 3u);
            if (below) holds = 0;
        }
        out[(size_t)v * eq.n + k] = (uint8_t)holds;
    }
}

// Copy every node's list into the packed output, ascending.  Lists produced
// by the single-successor defender copy keep the successor's order and are
// sorted here (insertion sort; such lists are short antichain images).
__global__ void k_gather(const uint4* st, const uint16_t* arena, uint32_t n, uint32_t t, const uint32_t* offs, uint32_t* out,
                         uint64_t cap, const uint32_t* map) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint4 d = load_cur(st, map ? map[v] : v, t);
    const uint32_t len = nl_len(d);
    if ((uint64_t)offs[v] + len > cap) return;   // the caller checks the total against cap
    uint32_t* o = out + offs[v];
    for (uint32_t j = 0; j < len; j++) o[j] = nl_get(d, arena, j);
    if (d.x & UNSORTED) {
        for (uint32_t i = 1; i < len; i++) {
            uint32_t x = o[i];
            uint32_t j = i;
            while (j > 0 && o[j - 1] > x) { o[j] = o[j - 1]; j--; }
            o[j] = x;
        }
    }
}

}  // namespace geq
