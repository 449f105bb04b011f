// Ahead-of-time gfx950 kernels that do not depend on a query's row code:
// the device hash index of reffine's build_index
// (include/reffine/arrow/table.h:107-138), its point lookup (vector_locate,
// include/reffine/vinstr/vinstr.h:137-146), fills used instead of
// hipMemset so that launches stay capturable, and a copy kernel of known
// size used to calibrate HBM byte counters.
#include <hip/hip_runtime.h>

#include "rf_device.h"

extern "C" __global__ void rf_fill_u64(uint64_t* p, uint64_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Inserts every valid key of column 0 with its row; equal keys keep the
// smallest row (the first occurrence).
extern "C" __global__ void rf_index_build(const int64_t* __restrict__ keys,
                                          const uint16_t* __restrict__ valid,
                                          int64_t n, unsigned long long* slot_keys,
                                          unsigned long long* slot_rows, uint64_t mask,
                                          unsigned long long* min_key_row) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (valid && !rf_valid_bit(valid, i)) continue;
        const int64_t key = keys[i];
        if (key == RF_EMPTY_KEY) {
            atomicMin(min_key_row, (unsigned long long)i);
            continue;
        }
        uint64_t s = rf_hash64((uint64_t)key) & mask;
        for (;;) {
            const unsigned long long prev =
                atomicCAS(&slot_keys[s], (unsigned long long)RF_EMPTY_KEY,
                          (unsigned long long)key);
            if (prev == (unsigned long long)RF_EMPTY_KEY || prev == (unsigned long long)key) {
                atomicMin(&slot_rows[s], (unsigned long long)i);
                break;
            }
            s = (s + 1) & mask;
        }
    }
}

extern "C" __global__ void rf_index_lookup(RfProbeDev h, const int64_t* __restrict__ keys,
                                           int64_t n, int64_t* __restrict__ rows) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        rows[i] = rf_probe(h, keys[i]);
}

// Device copy in the streaming shape of the scan kernels (contiguous range
// per wave, 8 nontemporal 16-byte loads per lane in flight): the measured
// HBM ceiling, and a kernel of known size to calibrate FETCH/WRITE_SIZE.
typedef unsigned int rf_u4 __attribute__((ext_vector_type(4)));

extern "C" __global__ void rf_copy_calib(const rf_u4* __restrict__ src, rf_u4* __restrict__ dst,
                                         int64_t n16) {
    constexpr int U = 8;
    constexpr int64_t CHUNK = 64 * U;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / 64);
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int64_t lane = threadIdx.x & 63;
    const int64_t nchunks = n16 / CHUNK;
    const int64_t per = (nchunks + nw - 1) / nw;
    const int64_t c0 = gw * per, c1 = (c0 + per < nchunks) ? c0 + per : nchunks;
    for (int64_t c = c0; c < c1; c++) {
        rf_u4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(src + c * CHUNK + u * 64 + lane);
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], dst + c * CHUNK + u * 64 + lane);
    }
    for (int64_t i = nchunks * CHUNK + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------- host launchers

namespace rf {

static int grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

hipError_t aot_fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, (uint64_t*)p, v, n);
    return hipGetLastError();
}

hipError_t aot_index_build(const int64_t* keys, const uint16_t* valid, int64_t n,
                           void* slot_keys, void* slot_rows, uint64_t mask, void* min_key_row,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_index_build, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, valid, n,
                       (unsigned long long*)slot_keys, (unsigned long long*)slot_rows, mask,
                       (unsigned long long*)min_key_row);
    return hipGetLastError();
}

hipError_t aot_index_lookup(const RfProbeDev& h, const int64_t* keys, int64_t n, int64_t* rows,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_index_lookup, dim3(grid_for(n, 256)), dim3(256), 0, s, h, keys, n, rows);
    return hipGetLastError();
}

hipError_t aot_copy(const void* src, void* dst, int64_t n16, hipStream_t s) {
    if (n16 <= 0) return hipSuccess;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rf_copy_calib, dim3(cus * 4), dim3(256), 0, s, (const rf_u4*)src,
                       (rf_u4*)dst, n16);
    return hipGetLastError();
}

}  // namespace rf
