// Device-wide exclusive scan of uint32 values (reduce-then-scan).
//
// The reference's prefix_sum (src/utils.rs:42-57) is inclusive and
// sequential; this is the device-wide building block the builders use where
// an exclusive prefix over a large array is needed.
#include "block_scan.hpp"
#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kThreads = 256;
constexpr int kPer = 8;
constexpr uint32_t kTile = kThreads * kPer;  // values per block

__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t* total) {
    return block_exclusive_n<kThreads>(v, sh, total);
}

__global__ void __launch_bounds__(kThreads)
    tile_sums_kernel(const uint32_t* data, uint32_t count, uint32_t* sums) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t s = 0;
    for (int k = 0; k < kPer; ++k)
        if (base + k < count) s += data[base + k];
    uint32_t tot = 0;
    block_exclusive(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kThreads)
    tile_scan_kernel(uint32_t* data, uint32_t count, const uint32_t* offsets) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t s = 0;
    for (int k = 0; k < kPer; ++k) {
        v[k] = base + k < count ? data[base + k] : 0u;
        s += v[k];
    }
    uint32_t run = block_exclusive(s, sh, nullptr) + offsets[blockIdx.x];
    for (int k = 0; k < kPer; ++k) {
        if (base + k < count) data[base + k] = run;
        run += v[k];
    }
}

// Single-block exclusive scan for the tile sums; writes the grand total.
__global__ void __launch_bounds__(1024) sums_scan_kernel(uint32_t* data, uint32_t count,
                                                         uint32_t* total) {
    __shared__ uint32_t sh[1024 / 64];
    const uint32_t per = (count + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = b + per < count ? b + per : count;
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    uint32_t all = 0;
    uint32_t run = block_exclusive_n<1024>(s, sh, &all);
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = all;
}

// Device-count variants (scan_exclusive_dev): K arrays scanned together,
// each of its own length, read on the device; a block past every length
// returns at once, and a block past one array's length skips that array
// (a uniform branch: the block's barriers stay matched).
template <int K>
__device__ __forceinline__ uint32_t load_counts(const ScanCounts& c, uint32_t (&cnt)[K]) {
    uint32_t most = 0;
    for (int k = 0; k < K; ++k) {
        cnt[k] = *c.p[k];
        most = cnt[k] > most ? cnt[k] : most;
    }
    return most;
}

template <int K>
__global__ void __launch_bounds__(kThreads)
    tile_sums_dev_kernel(ScanArrays a, ScanCounts counts, uint32_t* sums, uint32_t stride) {
    __shared__ uint32_t sh[kThreads / 64];
    uint32_t cnt[K];
    if (blockIdx.x * kTile >= load_counts<K>(counts, cnt)) return;  // whole block
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    for (int k = 0; k < K; ++k) {
        const uint32_t count = cnt[k];
        if (blockIdx.x * kTile >= count) continue;
        uint32_t s = 0;
        for (int q = 0; q < kPer; ++q)
            if (base + q < count) s += a.p[k][base + q];
        uint32_t tot = 0;
        block_exclusive(s, sh, &tot);
        if (threadIdx.x == 0) sums[k * stride + blockIdx.x] = tot;
    }
}

template <int K>
__global__ void __launch_bounds__(1024)
    sums_scan_dev_kernel(uint32_t* sums, uint32_t stride, ScanCounts counts, ScanTotals totals) {
    __shared__ uint32_t sh[1024 / 64];
    uint32_t cnt[K];
    load_counts<K>(counts, cnt);
    for (int k = 0; k < K; ++k) {
        const uint32_t tiles = (cnt[k] + kTile - 1) / kTile;
        const uint32_t per = (tiles + 1023) / 1024;
        const uint32_t b = threadIdx.x * per;
        const uint32_t e = b + per < tiles ? b + per : tiles;
        uint32_t* data = sums + k * stride;
        uint32_t s = 0;
        for (uint32_t i = b; i < e; ++i) s += data[i];
        uint32_t all = 0;
        uint32_t run = block_exclusive_n<1024>(s, sh, &all);
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t v = data[i];
            data[i] = run;
            run += v;
        }
        if (threadIdx.x == 0 && totals.p[k]) *totals.p[k] = all;
    }
}

template <int K>
__global__ void __launch_bounds__(kThreads)
    tile_scan_dev_kernel(ScanArrays a, ScanCounts counts, const uint32_t* sums, uint32_t stride) {
    __shared__ uint32_t sh[kThreads / 64];
    uint32_t cnt[K];
    if (blockIdx.x * kTile >= load_counts<K>(counts, cnt)) return;  // whole block
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    for (int k = 0; k < K; ++k) {
        const uint32_t count = cnt[k];
        if (blockIdx.x * kTile >= count) continue;
        uint32_t v[kPer];
        uint32_t s = 0;
        for (int q = 0; q < kPer; ++q) {
            v[q] = base + q < count ? a.p[k][base + q] : 0u;
            s += v[q];
        }
        uint32_t run = block_exclusive(s, sh, nullptr) + sums[k * stride + blockIdx.x];
        for (int q = 0; q < kPer; ++q) {
            if (base + q < count) a.p[k][base + q] = run;
            run += v[q];
        }
    }
}

template <int K>
hipError_t launch_scan_dev(const ScanArrays& a, const ScanCounts& counts, uint32_t max_count,
                           const ScanTotals& totals, uint32_t* sums, hipStream_t stream) {
    const uint32_t tiles = (max_count + kTile - 1) / kTile;
    const uint32_t stride = tiles + 1;
    hipLaunchKernelGGL(tile_sums_dev_kernel<K>, dim3(tiles), dim3(kThreads), 0, stream, a, counts,
                       sums, stride);
    hipLaunchKernelGGL(sums_scan_dev_kernel<K>, dim3(1), dim3(1024), 0, stream, sums, stride,
                       counts, totals);
    hipLaunchKernelGGL(tile_scan_dev_kernel<K>, dim3(tiles), dim3(kThreads), 0, stream, a, counts,
                       sums, stride);
    return hipGetLastError();
}

}  // namespace

size_t scan_scratch_bytes(uint32_t count) {
    const size_t tiles = (size_t(count) + kTile - 1) / kTile;
    return (tiles + 1) * sizeof(uint32_t);
}

// In-place exclusive scan of data[0..count); *total (device- or
// host-accessible) receives the sum.  scratch holds scan_scratch_bytes(count).
hipError_t scan_exclusive_u32(uint32_t* data, uint32_t count, uint32_t* total, void* scratch,
                              size_t scratch_bytes, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t tiles = (count + kTile - 1) / kTile;
    if (scratch_bytes < scan_scratch_bytes(count)) return hipErrorInvalidValue;
    uint32_t* sums = static_cast<uint32_t*>(scratch);
    hipLaunchKernelGGL(tile_sums_kernel, dim3(tiles), dim3(kThreads), 0, stream, data, count, sums);
    hipLaunchKernelGGL(sums_scan_kernel, dim3(1), dim3(1024), 0, stream, sums, tiles, total);
    hipLaunchKernelGGL(tile_scan_kernel, dim3(tiles), dim3(kThreads), 0, stream, data, count, sums);
    return hipGetLastError();
}

size_t scan_dev_scratch_bytes(uint32_t max_count, int k) {
    const size_t tiles = (size_t(max_count) + kTile - 1) / kTile;
    return size_t(k) * (tiles + 1) * sizeof(uint32_t);
}

// In-place exclusive scans of the first k (1..kScanArraysMax) arrays of a,
// array j of *counts.p[j] values (read on the device, at most max_count);
// *totals.p[j] (device memory, or null) receives array j's sum.  scratch
// holds scan_dev_scratch_bytes(max_count, k).  Nothing waits for the host, so
// a sequence of them can be enqueued (or recorded) with counts still to come.
hipError_t scan_exclusive_dev(const ScanArrays& a, int k, const ScanCounts& counts,
                              uint32_t max_count, const ScanTotals& totals, void* scratch,
                              size_t scratch_bytes, hipStream_t stream) {
    if (max_count == 0) return hipSuccess;
    if (k < 1 || k > kScanArraysMax || scratch_bytes < scan_dev_scratch_bytes(max_count, k))
        return hipErrorInvalidValue;
    for (int j = 0; j < k; ++j)
        if (counts.p[j] == nullptr || a.p[j] == nullptr) return hipErrorInvalidValue;
    uint32_t* sums = static_cast<uint32_t*>(scratch);
    switch (k) {
        case 1: return launch_scan_dev<1>(a, counts, max_count, totals, sums, stream);
        case 2: return launch_scan_dev<2>(a, counts, max_count, totals, sums, stream);
        case 3: return launch_scan_dev<3>(a, counts, max_count, totals, sums, stream);
        case 4: return launch_scan_dev<4>(a, counts, max_count, totals, sums, stream);
        case 5: return launch_scan_dev<5>(a, counts, max_count, totals, sums, stream);
        default: return launch_scan_dev<6>(a, counts, max_count, totals, sums, stream);
    }
}

}  // namespace rtbvh
