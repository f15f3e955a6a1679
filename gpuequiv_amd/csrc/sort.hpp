// Device sorts of the builder and of bisimulation minimisation.
//
//   radix_sort_pairs : stable LSD radix sort of (key, value) u32 pairs, 8-bit
//                      digits, one histogram + scan + scatter per digit.  The
//                      scatter ranks equal digits of a wave with eight ballots
//                      (lanes whose digits agree bit by bit) and of the block
//                      with per-wave digit counts in LDS, so the order of equal
//                      keys is kept.
//   seg_sort_u64     : every row r of a CSR (keys[row[r]..row[r+1])) sorted
//                      ascending in place; short rows by one thread each
//                      (insertion sort), long rows by one block each (bitonic
//                      network in LDS, or in global memory past the LDS tile).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "runtime.hpp"
#include "scan.hpp"

namespace geq {

constexpr uint32_t RADIX_THREADS = 256;   // = the number of digit values
constexpr uint32_t RADIX_ROUNDS = 16;     // sub-tiles of 256 items per block
constexpr uint32_t RADIX_TILE = RADIX_THREADS * RADIX_ROUNDS;

// per-block digit counts, digit-major: hist[d * nb + block]
static __global__ __launch_bounds__(RADIX_THREADS) void k_radix_hist(const uint32_t* key, size_t n, uint32_t shift, uint32_t nb,
                                                              uint32_t* hist) {
    __shared__ uint32_t cnt[RADIX_THREADS];
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * RADIX_TILE;
    for (uint32_t r = 0; r < RADIX_ROUNDS; r++) {
        const size_t i = b0 + (size_t)r * RADIX_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// offs = the exclusive scan of hist: where this block's items of each digit go
static __global__ __launch_bounds__(RADIX_THREADS) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                                 uint32_t* vout, size_t n, uint32_t shift, uint32_t nb,
                                                                 const uint32_t* offs) {
    constexpr uint32_t NW = RADIX_THREADS / 64;
    __shared__ uint32_t run[RADIX_THREADS];      // next output slot of each digit
    __shared__ uint32_t wcnt[NW][RADIX_THREADS]; // this sub-tile's digit counts per wave
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    run[t] = offs[(size_t)t * nb + blockIdx.x];
    for (uint32_t q = 0; q < NW; q++) wcnt[q][t] = 0u;
    __syncthreads();
    const uint64_t below_me = (1ull << lane) - 1ull;
    const size_t b0 = (size_t)blockIdx.x * RADIX_TILE;
    for (uint32_t r = 0; r < RADIX_ROUNDS; r++) {
        const size_t i = b0 + (size_t)r * RADIX_THREADS + t;
        const bool ok = i < n;
        uint32_t k = 0u, v = 0u, d = 0u;
        if (ok) {
            k = kin[i];
            v = vin[i];
            d = (k >> shift) & 255u;
        }
        // lanes holding the same digit: agree on each of its 8 bits
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t ones = __ballot(bit);
            peers &= bit ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below_me);
        if (ok && rank == 0u) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (ok) {
            uint32_t pos = run[d] + rank;
            for (uint32_t q = 0; q < w; q++) pos += wcnt[q][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0u;
        for (uint32_t q = 0; q < NW; q++) {
            add += wcnt[q][t];
            wcnt[q][t] = 0u;
        }
        run[t] += add;
        __syncthreads();
    }
}

struct RadixScratch {
    DevBuf<uint32_t> hist, scan;
};

// Stable sort of (key_in, val_in)[0..n) by the low `bits` bits of the keys into
// (key_out, val_out).  key_alt / val_alt (n entries) hold the middle passes when
// bits > 8; the inputs are left as they were.
static inline geq_status radix_sort_pairs(const uint32_t* key_in, const uint32_t* val_in, uint32_t* key_out, uint32_t* val_out,
                                   uint32_t* key_alt, uint32_t* val_alt, size_t n, uint32_t bits, RadixScratch& sc,
                                   hipStream_t s) {
    if (n == 0) return GEQ_OK;
    const uint32_t passes = bits <= 8 ? 1u : (std::min<uint32_t>(bits, 32u) + 7u) / 8u;
    if (passes > 1 && (!key_alt || !val_alt)) return fail(GEQ_ERR_INTERNAL, "radix sort: no buffers for the middle passes");
    const size_t nb = (n + RADIX_TILE - 1) / RADIX_TILE;
    const size_t nh = nb * RADIX_THREADS;
    if (nh >= 0xffffffffull) return fail(GEQ_ERR_INVALID_ARG, "radix sort of more than 2^32 items");
    if (!sc.hist.p || sc.hist.n < nh + 1) {
        sc.hist.retire();
        GEQ_HIP(sc.hist.alloc(nh + 1), GEQ_ERR_OUT_OF_MEMORY);
    }
    const uint32_t* sk = key_in;
    const uint32_t* sv = val_in;
    for (uint32_t p = 0; p < passes; p++) {
        // the last pass writes the output; the ones before alternate so that holds
        const bool to_out = ((passes - 1 - p) & 1u) == 0u;
        uint32_t* dk = to_out ? key_out : key_alt;
        uint32_t* dv = to_out ? val_out : val_alt;
        k_radix_hist<<<(uint32_t)nb, RADIX_THREADS, 0, s>>>(sk, n, 8u * p, (uint32_t)nb, sc.hist.p);
        geq_status st = exclusive_scan_dev<uint32_t, SumU32>(sc.hist.p, sc.hist.p, nh, sc.scan, s, SumU32{});
        if (st != GEQ_OK) return st;
        k_radix_scatter<<<(uint32_t)nb, RADIX_THREADS, 0, s>>>(sk, sv, dk, dv, n, 8u * p, (uint32_t)nb, sc.hist.p);
        sk = dk;
        sv = dv;
    }
    GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
    return GEQ_OK;
}

constexpr uint32_t SEG_SMALL = 32;     // rows up to this length: one thread
constexpr uint32_t SEG_LDS = 4096;     // rows up to this length: sorted in LDS
constexpr uint32_t SEG_THREADS = 512;

static __global__ void k_seg_sort_small(const uint32_t* row, uint32_t n, unsigned long long* keys) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t b = row[r], e = row[r + 1];
    if (e - b < 2u || e - b > SEG_SMALL) return;
    for (uint32_t i = b + 1; i < e; i++) {
        const unsigned long long x = keys[i];
        uint32_t j = i;
        while (j > b) {
            const unsigned long long y = keys[j - 1];
            if (y <= x) break;
            keys[j] = y;
            j--;
        }
        keys[j] = x;
    }
}

// Bitonic network with every comparator ascending (the first step of each merge
// compares mirrored positions): a[len..p2) act as +infinity, so a comparator
// that reaches past len never swaps and is skipped.
static __device__ void bitonic_sort_block(unsigned long long* a, uint32_t len, uint32_t p2) {
    const uint32_t half = p2 >> 1;
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        const uint32_t hk = k >> 1;
        for (uint32_t c = threadIdx.x; c < half; c += blockDim.x) {
            const uint32_t i = (c / hk) * k + (c % hk);
            const uint32_t j = i ^ (k - 1u);
            if (j < len) {
                const unsigned long long x = a[i], y = a[j];
                if (x > y) {
                    a[i] = y;
                    a[j] = x;
                }
            }
        }
        __syncthreads();
        for (uint32_t h = hk >> 1; h >= 1u; h >>= 1) {
            for (uint32_t c = threadIdx.x; c < half; c += blockDim.x) {
                const uint32_t i = (c / h) * 2u * h + (c % h);
                const uint32_t j = i + h;
                if (j < len) {
                    const unsigned long long x = a[i], y = a[j];
                    if (x > y) {
                        a[i] = y;
                        a[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

static __global__ __launch_bounds__(SEG_THREADS) void k_seg_sort_long(const uint32_t* row, const uint32_t* long_rows,
                                                               unsigned long long* keys) {
    __shared__ unsigned long long tile[SEG_LDS];
    const uint32_t r = long_rows[blockIdx.x];
    const uint32_t b = row[r], len = row[r + 1] - b;
    uint32_t p2 = 1;
    while (p2 < len) p2 <<= 1;
    if (len <= SEG_LDS) {
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) tile[i] = keys[b + i];
        __syncthreads();
        bitonic_sort_block(tile, len, p2);
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) keys[b + i] = tile[i];
    } else {
        // past the tile the network runs on the row in global memory; the block's
        // barriers order its own global reads and writes
        bitonic_sort_block(keys + b, len, p2);
    }
}

// The rows longer than SEG_SMALL, found once on the host (a refinement sorts
// the same rows every round).
struct SegSortPlan {
    DevBuf<uint32_t> long_rows;
    uint32_t n_long = 0;
};

static inline geq_status seg_sort_plan(const uint32_t* h_row, uint32_t n, SegSortPlan& plan, hipStream_t s) {
    std::vector<uint32_t> rows;
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t len = h_row[r + 1] - h_row[r];
        if (len > (1u << 31)) return fail(GEQ_ERR_INVALID_ARG, "a row of more than 2^31 keys");
        if (len > SEG_SMALL) rows.push_back(r);
    }
    plan.n_long = (uint32_t)rows.size();
    if (rows.empty()) return GEQ_OK;
    GEQ_HIP(plan.long_rows.alloc(rows.size()), GEQ_ERR_OUT_OF_MEMORY);
    GEQ_HIP(hipMemcpyAsync(plan.long_rows.p, rows.data(), 4ull * rows.size(), hipMemcpyHostToDevice, s), GEQ_ERR_TRANSFER);
    GEQ_HIP(hipStreamSynchronize(s), GEQ_ERR_TRANSFER);
    return GEQ_OK;
}

static inline geq_status seg_sort_u64(const uint32_t* d_row, uint32_t n, unsigned long long* keys, const SegSortPlan& plan,
                               hipStream_t s) {
    const uint32_t tb = 256;
    if (n) k_seg_sort_small<<<(n + tb - 1) / tb, tb, 0, s>>>(d_row, n, keys);
    if (plan.n_long) k_seg_sort_long<<<plan.n_long, SEG_THREADS, 0, s>>>(d_row, plan.long_rows.p, keys);
    GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
    return GEQ_OK;
}

}  // namespace geq
