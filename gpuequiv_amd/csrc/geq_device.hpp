// Device-side primitives of the energy-game solver (gfx950, wave64).
//
// Energies of every supported EnergyConf (elements <= 6, max <= 3) are
// handled in one canonical layout: 6 fields of 2 bits (12-bit point, stored as
// u16), updates as 6 nibbles (u32).  The 4096 points of {0..3}^6 form one
// 4096-bit set = 64 lanes x 64 bits: lane l holds the points 64*l .. 64*l+63,
// i.e. dimensions 0..2 live inside a lane's u64 word and dimensions 3..5 are
// spread across lanes (strides of 1, 4 and 16 lanes).
//
// Upward-closed sets are the solver's currency:
//  * the attacker merge E(u) <- min(E(u) U U_v inv(E(v), w)) is the minimal
//    set of the OR of the up-closures (inv_update is monotone);
//  * the defender combination min{ sup(e_1..e_k) } of
//    src/energygame/defend_iterative.wgsl:182-267 has the up-closure
//    AND_k up(inv(E(s_k), w_k)), because x >= sup(a, b) iff x >= a and x >= b.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace geq {

// ---------------------------------------------------------------------------
// Per-node energy list descriptor ("slot"): 16 bytes.
//   h[0]: bits 0..14 = number of energies, bit 15 = list is not sorted
//   len <= 6 : h[1..6] hold the energies inline
//   len  > 6 : the record's second dword (h[2] | h[3] << 16) is the offset
//              into the energy arena (u16 units)
//   h[7]: tag of the round that wrote the slot (round + 1, mod 2^16)
//
// Every node owns two slots (32 bytes, in its node record below).  A round
// writes a node's new list into the slot that is not current, so the other
// evaluations of the same round keep reading the list as of the round's
// start (Jacobi) and no commit pass is needed.  At round r (tag t):
//   * a slot tagged t was written during round r itself: not current yet;
//   * otherwise the younger slot (smaller age t - tag, mod 2^16) is current,
//     slot 0 on a tie (both slots untouched since the initialisation).
// Ages stay far below 2^16: the host re-tags every node (k_retag) at least
// every RETAG_ROUNDS rounds.
// ---------------------------------------------------------------------------
constexpr uint32_t INLINE_CAP = 6;
constexpr uint32_t LEN_MASK = 0x7fffu;
constexpr uint32_t UNSORTED = 0x8000u;
constexpr uint32_t RETAG_ROUNDS = 16384;

struct __align__(16) NodeList {
    uint16_t h[8];
};

// Node record: 64 bytes = 4 x uint4, one per node, so that everything a
// node's own evaluation reads comes in one cache-line piece:
//   [0] header {first edge, degree | attacker << 31, first reverse edge, reverse degree}
//   [1], [2] the two list slots
//   [3] unused
constexpr uint32_t NODE_U4 = 4;
constexpr uint32_t ATK_BIT = 0x80000000u;
__host__ __device__ __forceinline__ size_t slot_index(uint32_t v, uint32_t s) { return (size_t)NODE_U4 * v + 1u + s; }

__host__ __device__ __forceinline__ uint32_t round_tag(uint32_t r) { return (r + 1u) & 0xffffu; }
__device__ __forceinline__ uint32_t nl_len(const uint4& d) { return d.x & LEN_MASK; }
__device__ __forceinline__ uint32_t nl_off(const uint4& d) { return d.y; }  // h[2] | h[3] << 16
__device__ __forceinline__ uint32_t nl_tag(const uint4& d) { return d.w >> 16; }
// the record with its tag field replaced
__device__ __forceinline__ uint4 nl_with_tag(uint4 d, uint32_t tag) {
    d.w = (d.w & 0xffffu) | ((tag & 0xffffu) << 16);
    return d;
}
// which of a node's two slots is current at round tag t (see above)
__device__ __forceinline__ uint32_t cur_slot(const uint4& a, const uint4& b, uint32_t t) {
    const uint32_t age_a = (t - nl_tag(a)) & 0xffffu, age_b = (t - nl_tag(b)) & 0xffffu;
    return (age_a == 0u || (age_b != 0u && age_b < age_a)) ? 1u : 0u;
}
// node v's current list slot at round tag t (both slots are loaded together)
__device__ __forceinline__ uint4 load_cur(const uint4* __restrict__ nodes, uint32_t v, uint32_t t) {
    const uint4 a = nodes[slot_index(v, 0)], b = nodes[slot_index(v, 1)];
    return cur_slot(a, b, t) ? b : a;
}
__device__ __forceinline__ uint4 load_cur(const uint4* __restrict__ nodes, uint32_t v, uint32_t t, uint32_t& slot) {
    const uint4 a = nodes[slot_index(v, 0)], b = nodes[slot_index(v, 1)];
    slot = cur_slot(a, b, t);
    return slot ? b : a;
}
// inline energy j (0..5) of a descriptor
__device__ __forceinline__ uint32_t nl_inline(const uint4& d, uint32_t j) {
    // energies sit at u16 index 1 + j of the 16-byte record
    uint32_t idx = 1 + j;
    uint32_t w = (idx < 2) ? d.x : (idx < 4) ? d.y : (idx < 6) ? d.z : d.w;
    return (idx & 1) ? (w >> 16) : (w & 0xffffu);
}
__device__ __forceinline__ uint32_t nl_get(const uint4& d, const uint16_t* __restrict__ arena, uint32_t j) {
    return nl_len(d) <= INLINE_CAP ? nl_inline(d, j) : (uint32_t)arena[nl_off(d) + j];
}

// ---------------------------------------------------------------------------
// inv_update on the canonical layout (shadergen.rs:103-128, attack_update.wgsl:45-61):
// first every Decrement field i: e[i] = min(e[i] + 1, max); then, in order
// i = 0..5, every Min field (nibble v >= 2): e[v-2] = max(e[v-2], e[i]),
// reading values already updated earlier in the sequence.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t inv_update(uint32_t e, uint32_t u, uint32_t maxv) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        uint32_t ui = (u >> (4 * i)) & 0xfu;
        uint32_t ei = (e >> (2 * i)) & 3u;
        uint32_t ni = (ui == 1u) ? min(ei + 1u, maxv) : min(ei, maxv);
        e = (e & ~(3u << (2 * i))) | (ni << (2 * i));
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        uint32_t ui = (u >> (4 * i)) & 0xfu;
        if (ui > 1u) {
            uint32_t j = ui - 2u;                 // validated on the host: j < 6
            uint32_t ei = (e >> (2 * i)) & 3u;
            uint32_t ej = (e >> (2 * j)) & 3u;
            uint32_t m = ei > ej ? ei : ej;
            e = (e & ~(3u << (2 * j))) | (m << (2 * j));
        }
    }
    return e;
}

// The same map with the per-update work done once per edge: prep_update
// derives the fields to increment (bit 2i of `dec` for a Decrement in field i)
// and the Min fields (bit 4i of `mins`); inv_apply then needs a saturating
// add and one max per Min field, in field order.
struct InvUpd {
    uint32_t dec;
    uint32_t mins;
    uint32_t u;
};
__device__ __forceinline__ InvUpd prep_update(uint32_t u) {
    const uint32_t t = u ^ 0x111111u;                                  // nibble i == 1  <=>  t's nibble i == 0
    const uint32_t nonzero = (((t & 0x777777u) + 0x777777u) | t) & 0x888888u;
    const uint32_t one = ~nonzero & 0x888888u;                          // bit 4i+3: nibble i is Decrement
    uint32_t dec = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) dec |= ((one >> (4 * i + 3)) & 1u) << (2 * i);
    return InvUpd{dec, ((u >> 1) | (u >> 2) | (u >> 3)) & 0x111111u, u};   // bit 4i: nibble i >= 2 (Min)
}
// maxrep: maxv in every 2-bit field (maxv * 0x555); every field of e is <= maxv
__device__ __forceinline__ uint32_t inv_apply(uint32_t e, const InvUpd& p, uint32_t maxrep) {
    const uint32_t x = e ^ maxrep;
    const uint32_t at_max = ~(x | (x >> 1)) & 0x555u;                   // bit 2i: field i == maxv
    e += p.dec & ~at_max;                                               // no carries: those fields are < maxv
    uint32_t mm = p.mins;
    while (mm) {
        const uint32_t b = (uint32_t)__builtin_ctz(mm);                 // 4i
        mm &= mm - 1u;
        const uint32_t i2 = b >> 1;
        const uint32_t j2 = 2u * (((p.u >> b) & 0xfu) - 2u);             // Min(j + 1) is stored as j + 2
        const uint32_t ei = (e >> i2) & 3u, ej = (e >> j2) & 3u;
        if (ei > ej) e += (ei - ej) << j2;
    }
    return e;
}

// ---------------------------------------------------------------------------
// Wave-level helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// lanes below this one that are set in `mask`
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Memory ordering between the lanes of one wave through LDS / global memory.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// value of lane (lane_id - delta); garbage for lanes below delta (masked by callers)
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, uint32_t delta, uint32_t lane) {
    int src = ((int)lane - (int)delta) & 63;
    uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t shfl32(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

// the value of lane `src` (uniform) in every lane, read by the scalar unit
__device__ __forceinline__ uint32_t lane_val(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}

// DPP-moved value (the control below), 0 where the source lane falls outside
// the row; as the operand of an OR it folds into one v_or_b32_dpp.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_z(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
constexpr int DPP_QUAD_SHR1 = 0x90;    // quad_perm [0,0,1,2]: lane i of a quad reads lane max(i - 1, 0)
constexpr int DPP_QUAD_SHR2 = 0x40;    // quad_perm [0,0,0,1]: lane i of a quad reads lane max(i - 2, 0)
constexpr int DPP_ROW_SHR1 = 0x111;    // row_shr:1
constexpr int DPP_ROW_SHR4 = 0x114;    // row_shr:4
constexpr int DPP_ROW_SHR8 = 0x118;    // row_shr:8

// Value of lane (lane - n) within the lane's row of 16 (DPP row_shr:n, no
// LDS traffic); 0 for the first n lanes of a row.  n in {1, 2, 4, 8}.
template <int N>
__device__ __forceinline__ uint32_t row_shr(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 | N, 0xF, 0xF, true);
}
template <int N>
__device__ __forceinline__ uint64_t row_shr64(uint64_t v) {
    return ((uint64_t)row_shr<N>((uint32_t)(v >> 32)) << 32) | row_shr<N>((uint32_t)v);
}

// inclusive wave prefix sum: within rows of 16 by DPP, then across rows
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
    v += row_shr<1>(v);
    v += row_shr<2>(v);
    v += row_shr<4>(v);
    v += row_shr<8>(v);
    // row sums: lane 15 of each row holds it
    const uint32_t r0 = lane_val(v, 15), r1 = lane_val(v, 31), r2 = lane_val(v, 47);
    const uint32_t row = lane >> 4;
    if (row >= 1) v += r0;
    if (row >= 2) v += r1;
    if (row >= 3) v += r2;
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, (int)d, 64);
    return v;
}

// ---------------------------------------------------------------------------
// 4096-point bitmap operations on one u64 word per lane.
// Point p = 64 * lane + b has coordinates e_k = (p >> 2k) & 3.
// ---------------------------------------------------------------------------

// Upward closure: every point >= some member becomes a member.  Each step
// ORs in the set shifted up by one or two values of one dimension.
//  * dims 0 and 1 stay inside 16-bit groups: the two 32-bit halves are shifted
//    independently;
//  * dim 2 moves whole 16-bit groups (lo.low, lo.high, hi.low, hi.high);
//  * dim 3 (lanes of a quad) and dim 4 (quads of a row of 16) are DPP quad
//    permutations and zero-filling row shifts, folded into the OR (a lane that
//    reads itself adds nothing);
//  * dim 5 (rows) crosses rows: ds_bpermute, masked below row 1 / row 2.
__device__ __forceinline__ uint64_t up_close(uint64_t w, uint32_t lane) {
    uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
    lo |= (lo & 0x77777777u) << 1;  hi |= (hi & 0x77777777u) << 1;
    lo |= (lo & 0x33333333u) << 2;  hi |= (hi & 0x33333333u) << 2;
    lo |= (lo & 0x0FFF0FFFu) << 4;  hi |= (hi & 0x0FFF0FFFu) << 4;
    lo |= (lo & 0x00FF00FFu) << 8;  hi |= (hi & 0x00FF00FFu) << 8;
    const uint32_t lo1 = lo | (lo << 16);                  // group 1 |= group 0
    const uint32_t hi1 = hi | (hi << 16) | (lo >> 16);     // groups 2, 3 |= groups 1, 2
    lo = lo1;
    hi = hi1 | lo1;                                        // groups 2, 3 |= groups 0, 1
    lo |= dpp_z<DPP_QUAD_SHR1>(lo);  hi |= dpp_z<DPP_QUAD_SHR1>(hi);
    lo |= dpp_z<DPP_QUAD_SHR2>(lo);  hi |= dpp_z<DPP_QUAD_SHR2>(hi);
    lo |= dpp_z<DPP_ROW_SHR4>(lo);   hi |= dpp_z<DPP_ROW_SHR4>(hi);
    lo |= dpp_z<DPP_ROW_SHR8>(lo);   hi |= dpp_z<DPP_ROW_SHR8>(hi);
    w = ((uint64_t)hi << 32) | lo;
    const uint32_t e5 = lane >> 4;
    uint64_t t;
    t = shfl_up64(w, 16, lane); if (e5 >= 1) w |= t;
    t = shfl_up64(w, 32, lane); if (e5 >= 2) w |= t;
    return w;
}

// Minimal points of an upward-closed set: x is minimal iff no x - e_k is a member.
__device__ __forceinline__ uint64_t minimal_of(uint64_t w, uint32_t lane) {
    const uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
    uint32_t nlo = ((lo & 0x77777777u) << 1) | ((lo & 0x0FFF0FFFu) << 4) | (lo << 16);
    uint32_t nhi = ((hi & 0x77777777u) << 1) | ((hi & 0x0FFF0FFFu) << 4) | (hi << 16) | (lo >> 16);
    // dim 3: the lane below, within the quad (read by every lane, kept by lanes 1..3 of a quad)
    const uint32_t tlo = dpp_z<DPP_ROW_SHR1>(lo), thi = dpp_z<DPP_ROW_SHR1>(hi);
    nlo |= (lane & 3u) ? tlo : 0u;
    nhi |= (lane & 3u) ? thi : 0u;
    nlo |= dpp_z<DPP_ROW_SHR4>(lo);  // dim 4: four lanes below, 0 in the row's first quad
    nhi |= dpp_z<DPP_ROW_SHR4>(hi);
    uint64_t nm = ((uint64_t)nhi << 32) | nlo;
    const uint64_t t = shfl_up64(w, 16, lane);
    if (lane >= 16u) nm |= t;        // dim 5: the row below
    return w & ~nm;
}

__device__ __forceinline__ void bm_set(uint32_t* bm, uint32_t p) {
    atomicOr(&bm[p >> 5], 1u << (p & 31u));
}
__device__ __forceinline__ uint64_t bm_word(const uint32_t* bm, uint32_t lane) {
    return (uint64_t)bm[2 * lane] | ((uint64_t)bm[2 * lane + 1] << 32);
}
__device__ __forceinline__ bool bm_test(const uint32_t* bm, uint32_t p) {
    return (bm[p >> 5] >> (p & 31u)) & 1u;
}

}  // namespace geq
