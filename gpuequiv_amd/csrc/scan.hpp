// Device-wide exclusive scan (reduce-then-scan): used for the energy
// download's offsets, gamebuild's per-level offsets, the builder's and the
// server's reverse CSRs, and the radix sort's digit offsets.
//   k_scan_reduce : every block of SCAN_ITEMS consecutive items -> its total
//   k_scan_tops   : one block scans the block totals in place (and the grand total)
//   k_scan_apply  : every block scans its items again from its block's offset
// An input of at most SCAN_ITEMS items (4-byte items: SCAN_ONE_ITEMS) takes
// k_scan_apply alone, in one block (of SCAN_THREADS threads, or
// SCAN_ONE_THREADS past SCAN_ITEMS items): one launch; one of at most
// SCAN_RAW_MAX blocks skips k_scan_tops, each
// k_scan_apply block combining the totals of the blocks before it itself (two
// launches).
// Global memory is read and written striped (lane-consecutive, coalesced); the
// apply step transposes its tile through LDS so that each thread scans
// SCAN_PER_THREAD consecutive items.  The combine operator is a template
// (plain sums, or the saturating pair sums of the gamebuild bounds); it must
// be associative and commutative (the block totals combine in striped order).
// The input is read through an accessor (an array, or values computed from
// other arrays, as gamebuild's candidate flags).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "runtime.hpp"

namespace geq {

constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_PER_THREAD = 16;
constexpr uint32_t SCAN_ITEMS = SCAN_THREADS * SCAN_PER_THREAD;   // items per block
constexpr uint32_t SCAN_TOP_THREADS = 1024;
constexpr uint32_t SCAN_RAW_MAX = 1024;   // block totals an apply block combines itself
// the one-block scan of 4-byte items: 1,024 threads, 16,384 items (a 66 KB LDS
// tile; for 8-byte items this block took longer than the two launches)
constexpr uint32_t SCAN_ONE_THREADS = 1024;
constexpr uint32_t SCAN_ONE_ITEMS = SCAN_ONE_THREADS * SCAN_PER_THREAD;

struct SumU32 {
    __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
    __host__ __device__ static uint32_t identity() { return 0u; }
};

namespace detail {

__device__ __forceinline__ uint32_t shfl_up_t(uint32_t v, uint32_t d) { return (uint32_t)__shfl_up((int)v, d, 64); }
__device__ __forceinline__ uint2 shfl_up_t(uint2 v, uint32_t d) {
    return make_uint2((uint32_t)__shfl_up((int)v.x, d, 64), (uint32_t)__shfl_up((int)v.y, d, 64));
}

// inclusive scan of one value per lane over the wave (op applied by shuffles)
template <class T, class Op>
__device__ __forceinline__ T wave_scan_incl(T v, uint32_t lane, Op op) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const T y = shfl_up_t(v, d);
        if (lane >= d) v = op(y, v);
    }
    return v;
}

// exclusive scan of one value per thread over the block; `total` gets the block's combination
template <class T, class Op, uint32_t NT>
__device__ __forceinline__ T block_scan_excl(T v, Op op, T& total) {
    __shared__ T s_w[NT / 64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const T incl = wave_scan_incl(v, lane, op);
    if (lane == 63u) s_w[w] = incl;
    __syncthreads();
    T pre = Op::identity(), tot = Op::identity();
    for (uint32_t k = 0; k < NT / 64; k++) {
        if (k < w) pre = op(pre, s_w[k]);
        tot = op(tot, s_w[k]);
    }
    __syncthreads();
    total = tot;
    // exclusive value of this thread: pre (+) the wave's exclusive part
    const T wex = shfl_up_t(incl, 1);
    return lane == 0 ? pre : op(pre, wex);
}

}  // namespace detail

// Input accessor of an array.
template <class T>
struct ScanPtr {
    const T* p;
    __device__ __forceinline__ T operator()(size_t i) const { return p[i]; }
};

template <class T, class Op, class In>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(In in, size_t n, T* tops, Op op) {
    const size_t b0 = (size_t)blockIdx.x * SCAN_ITEMS + threadIdx.x;
    T acc = Op::identity();
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; k++) {
        const size_t i = b0 + (size_t)k * SCAN_THREADS;
        if (i < n) acc = op(acc, in(i));
    }
    T total;
    (void)detail::block_scan_excl<T, Op, SCAN_THREADS>(acc, op, total);
    if (threadIdx.x == 0) tops[blockIdx.x] = total;
}

// tops[0..nb) scanned in place (exclusive); tops[nb] = the grand total
template <class T, class Op>
__global__ __launch_bounds__(SCAN_TOP_THREADS) void k_scan_tops(T* tops, uint32_t nb, Op op) {
    const uint32_t per = (nb + SCAN_TOP_THREADS - 1) / SCAN_TOP_THREADS;
    const uint32_t b = min(nb, threadIdx.x * per), e = min(nb, b + per);
    T acc = Op::identity();
    for (uint32_t i = b; i < e; i++) acc = op(acc, tops[i]);
    T total;
    T run = detail::block_scan_excl<T, Op, SCAN_TOP_THREADS>(acc, op, total);
    for (uint32_t i = b; i < e; i++) {
        const T x = tops[i];
        tops[i] = run;
        run = op(run, x);
    }
    if (threadIdx.x == 0) tops[nb] = total;
}

// LDS slot of tile item i: one pad slot per 32 items spreads the per-thread
// runs of SCAN_PER_THREAD consecutive items over the banks
__device__ __forceinline__ uint32_t scan_slot(uint32_t i) { return i + (i >> 5); }

// How k_scan_apply finds its block's offset: none (one block), from tops
// scanned by k_scan_tops, or by combining the raw block totals before it.
constexpr uint32_t SCAN_TOPS_NONE = 0, SCAN_TOPS_SCANNED = 1, SCAN_TOPS_RAW = 2;

template <class T, class Op, class In, uint32_t NT = SCAN_THREADS>
__global__ __launch_bounds__(NT) void k_scan_apply(In in, T* out, size_t n, const T* tops, uint32_t mode, Op op) {
    constexpr uint32_t ITEMS = NT * SCAN_PER_THREAD;
    __shared__ T tile[ITEMS + ITEMS / 32];
    const size_t base = (size_t)blockIdx.x * ITEMS;
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; k++) {
        const uint32_t j = k * NT + t;
        tile[scan_slot(j)] = base + j < n ? in(base + j) : Op::identity();
    }
    T pre = Op::identity();
    if (mode == SCAN_TOPS_SCANNED) {
        pre = tops[blockIdx.x];
    } else if (mode == SCAN_TOPS_RAW) {   // the totals of the blocks before this one
        T a = Op::identity();
        for (uint32_t k = t; k < blockIdx.x; k += NT) a = op(a, tops[k]);
        (void)detail::block_scan_excl<T, Op, NT>(a, op, pre);
    }
    __syncthreads();
    T v[SCAN_PER_THREAD];
    T acc = Op::identity();
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; k++) {
        v[k] = tile[scan_slot(t * SCAN_PER_THREAD + k)];
        acc = op(acc, v[k]);
    }
    T total;
    T run = op(pre, detail::block_scan_excl<T, Op, NT>(acc, op, total));
    // each thread rewrites only the slots it read; the barrier inside the block
    // scan has ordered every read of the tile before these writes
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; k++) {
        tile[scan_slot(t * SCAN_PER_THREAD + k)] = run;
        run = op(run, v[k]);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; k++) {
        const uint32_t j = k * NT + t;
        if (base + j < n) out[base + j] = tile[scan_slot(j)];
    }
    if (blockIdx.x == gridDim.x - 1 && t == 0)   // the grand total
        out[n] = mode == SCAN_TOPS_SCANNED ? tops[gridDim.x] : op(pre, total);
}

// Exclusive scan of in(0..n) into out[0..n), the total into out[n] (out holds
// n + 1 entries; an array input may be out itself).  tmp: block totals, grown
// as needed.
template <class T, class Op, class In>
geq_status exclusive_scan_in(In in, T* out, size_t n, DevBuf<T>& tmp, hipStream_t s, Op op) {
    const size_t nb = std::max<size_t>(1, (n + SCAN_ITEMS - 1) / SCAN_ITEMS);
    if (nb > 0x7fffffffull) return fail(GEQ_ERR_INVALID_ARG, "scan of more than 2^31 blocks");
    if (nb == 1) {   // one block holds the input: one launch
        k_scan_apply<T, Op, In><<<1, SCAN_THREADS, 0, s>>>(in, out, n, nullptr, SCAN_TOPS_NONE, op);
        GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
        return GEQ_OK;
    }
    if (sizeof(T) == 4 && n <= SCAN_ONE_ITEMS) {   // one large block: one launch instead of two (8-byte pairs: slower)
        k_scan_apply<T, Op, In, SCAN_ONE_THREADS><<<1, SCAN_ONE_THREADS, 0, s>>>(in, out, n, nullptr, SCAN_TOPS_NONE, op);
        GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
        return GEQ_OK;
    }
    if (!tmp.p || tmp.n < nb + 1) {
        tmp.retire();
        GEQ_HIP(tmp.alloc(nb + 1), GEQ_ERR_OUT_OF_MEMORY);
    }
    k_scan_reduce<T, Op, In><<<(uint32_t)nb, SCAN_THREADS, 0, s>>>(in, n, tmp.p, op);
    if (nb <= SCAN_RAW_MAX) {
        k_scan_apply<T, Op, In><<<(uint32_t)nb, SCAN_THREADS, 0, s>>>(in, out, n, tmp.p, SCAN_TOPS_RAW, op);
    } else {
        k_scan_tops<T, Op><<<1, SCAN_TOP_THREADS, 0, s>>>(tmp.p, (uint32_t)nb, op);
        k_scan_apply<T, Op, In><<<(uint32_t)nb, SCAN_THREADS, 0, s>>>(in, out, n, tmp.p, SCAN_TOPS_SCANNED, op);
    }
    GEQ_HIP(hipGetLastError(), GEQ_ERR_NO_DEVICE);
    return GEQ_OK;
}

template <class T, class Op>
geq_status exclusive_scan_dev(const T* in, T* out, size_t n, DevBuf<T>& tmp, hipStream_t s, Op op) {
    return exclusive_scan_in<T, Op, ScanPtr<T>>(ScanPtr<T>{in}, out, n, tmp, s, op);
}

}  // namespace geq
