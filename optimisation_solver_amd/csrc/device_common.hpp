// Device building blocks shared by the PDLP kernels (gfx950, wave64).
//
// Row-block SpMV: the host cuts the rows of one matrix orientation (CSR rows
// of A, or CSC columns of A = rows of A^T) into blocks of at most kTileNnz
// nonzeros and kTileRows rows; a row longer than kTileNnz gets a block of its
// own.  One 256-thread workgroup per block streams the block's nonzeros with
// fully coalesced loads, multiplies by the gathered vector entry and stages
// the products in LDS; then one thread per row sums its contiguous LDS slice
// in order and hands the finished row to a fused epilogue (prox step,
// residual terms, ...).  Long rows are reduced across the whole workgroup.
// Per-block partial reductions are combined by a separate fixed-order pass, so
// every result is run-to-run deterministic.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pdlpx_dev {

constexpr int kThreads = 256;          // workgroup size of every row-block kernel
constexpr int kWaves = kThreads / 64;
constexpr int kTileNnz = 1024;         // nonzeros staged per block
constexpr int kTileRows = kThreads;    // rows per block (one thread each)
constexpr int kReduceThreads = 1024;   // single-workgroup final reductions

// Matrix orientation as the kernels see it: rows of this orientation.
template <typename T>
struct Orient {
    const int32_t* ptr;     // nrows + 1
    const int32_t* idx;     // nnz
    const T* val;           // nnz
    const int32_t* blk;     // nblk + 1 row-block starts
    const int32_t* binst;   // nblk: batch instance owning each block
    int32_t nblk;
    int32_t nrows;
};

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sums N per-thread values over the workgroup in a fixed tree order; the
// results are valid in thread 0.  `sh` needs N * kWaves doubles.
template <int N>
__device__ inline void block_sums(double (&v)[N], double* sh) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) sh[q * kWaves + w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            double s = 0.0;
            for (int i = 0; i < kWaves; ++i) s += sh[q * kWaves + i];
            v[q] = s;
        }
    }
    __syncthreads();
}

template <typename T>
__device__ inline T block_total(T v, T* sh) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const T s = wave_sum(v);
    if (lane == 0) sh[w] = s;
    __syncthreads();
    T total = T(0);
    if (threadIdx.x == 0)
        for (int i = 0; i < kWaves; ++i) total += sh[i];
    __syncthreads();
    return total;
}

// One row block of y = M v with a per-row epilogue epi(row, sum).  `lds` holds
// kTileNnz elements of T, `shT` kWaves elements of T.
template <typename T, typename Epi>
__device__ inline void row_block(const Orient<T>& M, const T* __restrict__ v, T* lds, T* shT,
                                 Epi& epi) {
    const int b = blockIdx.x;
    const int r0 = M.blk[b];
    const int r1 = M.blk[b + 1];
    const int k0 = M.ptr[r0];
    const int k1 = M.ptr[r1];
    const int t = threadIdx.x;
    if (k1 - k0 > kTileNnz) {
        T s = T(0);
        for (int k = k0 + t; k < k1; k += kThreads) s += M.val[k] * v[M.idx[k]];
        const T total = block_total(s, shT);
        if (t == 0) epi(r0, total);
        return;
    }
    for (int k = k0 + t; k < k1; k += kThreads) lds[k - k0] = M.val[k] * v[M.idx[k]];
    __syncthreads();
    const int r = r0 + t;
    if (r < r1) {
        const int a = M.ptr[r] - k0;
        const int e = M.ptr[r + 1] - k0;
        T s = T(0);
        for (int q = a; q < e; ++q) s += lds[q];
        epi(r, s);
    }
}

// Two vectors through the same block (termination checks score the current
// and the averaged iterate in one matrix pass).  `lds` holds 2 * kTileNnz.
template <typename T, typename Epi>
__device__ inline void row_block2(const Orient<T>& M, const T* __restrict__ v0,
                                  const T* __restrict__ v1, T* lds, T* shT, Epi& epi) {
    const int b = blockIdx.x;
    const int r0 = M.blk[b];
    const int r1 = M.blk[b + 1];
    const int k0 = M.ptr[r0];
    const int k1 = M.ptr[r1];
    const int t = threadIdx.x;
    if (k1 - k0 > kTileNnz) {
        T s0 = T(0), s1 = T(0);
        for (int k = k0 + t; k < k1; k += kThreads) {
            const T a = M.val[k];
            const int c = M.idx[k];
            s0 += a * v0[c];
            s1 += a * v1[c];
        }
        const T t0 = block_total(s0, shT);
        const T t1 = block_total(s1, shT);
        if (t == 0) epi(r0, t0, t1);
        return;
    }
    for (int k = k0 + t; k < k1; k += kThreads) {
        const T a = M.val[k];
        const int c = M.idx[k];
        lds[k - k0] = a * v0[c];
        lds[kTileNnz + k - k0] = a * v1[c];
    }
    __syncthreads();
    const int r = r0 + t;
    if (r < r1) {
        const int a = M.ptr[r] - k0;
        const int e = M.ptr[r + 1] - k0;
        T s0 = T(0), s1 = T(0);
        for (int q = a; q < e; ++q) {
            s0 += lds[q];
            s1 += lds[kTileNnz + q];
        }
        epi(r, s0, s1);
    }
}

// std::max / std::min / clip with the reference's argument order, so NaN and
// signed-zero behaviour match the CPU code exactly.
template <typename T>
__device__ inline T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__device__ inline T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T>
__device__ inline T clip(T v, T lo, T hi) { return smax(lo, smin(v, hi)); }

}  // namespace pdlpx_dev
