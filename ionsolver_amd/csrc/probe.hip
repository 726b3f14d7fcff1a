// HBM copy probe: the bandwidth a plain streaming kernel reaches on this GPU,
// reported by the benchmark next to the published peak, and the known-size
// reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "device_lattice.hpp"
#include "ionsolver.h"

namespace {
typedef float v4f __attribute__((ext_vector_type(4)));

// 16-byte nontemporal loads and stores, grid-stride
__global__ void __launch_bounds__(256) k_copy_probe(const v4f* __restrict__ a, v4f* __restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

// stream_collide's access shape without its arithmetic: one work-item per
// cell reads S structure-of-arrays streams a[k * n + cell] (all loads first)
// and writes them to b[k * n + cell], nontemporal, elements of type T
template <int S, class T>
__global__ void __launch_bounds__(256) k_stream_probe(const T* __restrict__ a, T* __restrict__ b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[S];
#pragma unroll
    for (int k = 0; k < S; k++) v[k] = __builtin_nontemporal_load(&a[(uint64_t)k * n + i]);
#pragma unroll
    for (int k = 0; k < S; k++) __builtin_nontemporal_store(v[k], &b[(uint64_t)k * n + i]);
}

// stream_collide's in-place esoteric-pull addressing without its arithmetic,
// on an NX x NY x NZ lattice: S SoA arrays (D3Q19 fi, or fi, ei and the D3Q7
// fqi), each read and rewritten by one work-item per cell at the cell (arrays
// 0 and the even directions) or at the periodic neighbour n + c_i (the odd
// directions i), the layout of an even step. One 64-bit scalar base per array
// and a 32-bit byte offset per lane, as in the kernels.
template <int S, class T>
__global__ void __launch_bounds__(256) k_pull_probe(T* __restrict__ a, uint32_t NX, uint32_t NY, uint32_t NZ,
                                                    uint64_t N) {
    const uint64_t n64 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n64 >= N) return;
    const uint32_t n = (uint32_t)n64, nxy = NX * NY;
    const uint32_t z = n / nxy, y = (n - z * nxy) / NX, x = n - z * nxy - y * NX;
    T v[S];
    uint32_t off[S];
#pragma unroll
    for (int k = 0; k < S; k++) {
        const int m = k < 19 ? k : (k < 38 ? k - 19 : k - 38);
        const bool nb = (m & 1) != 0;
        const int cx = nb ? ion::VS<19>::cx[m] : 0, cy = nb ? ion::VS<19>::cy[m] : 0, cz = nb ? ion::VS<19>::cz[m] : 0;
        const uint32_t xs = cx > 0 ? (x + 1u == NX ? 0u : x + 1u) : (cx < 0 ? (x == 0u ? NX - 1u : x - 1u) : x);
        const uint32_t ys = cy > 0 ? (y + 1u == NY ? 0u : y + 1u) : (cy < 0 ? (y == 0u ? NY - 1u : y - 1u) : y);
        const uint32_t zs = cz > 0 ? (z + 1u == NZ ? 0u : z + 1u) : (cz < 0 ? (z == 0u ? NZ - 1u : z - 1u) : z);
        off[k] = (xs + ys * NX + zs * nxy) * (uint32_t)sizeof(T);
        v[k] = __builtin_nontemporal_load((const T*)((const char*)(a + (uint64_t)k * N) + off[k]));
    }
#pragma unroll
    for (int k = 0; k < S; k++) __builtin_nontemporal_store((T)(v[k] + 1u), (T*)((char*)(a + (uint64_t)k * N) + off[k]));
}

template <int S, class T>
void launch_stream_probe(const void* a, void* b, uint32_t n) {
    hipLaunchKernelGGL((k_stream_probe<S, T>), dim3((n + 255) / 256), dim3(256), 0, 0, (const T*)a, (T*)b, n);
}
}  // namespace

extern "C" __attribute__((visibility("hidden"))) int ion_internal_set_error(int code, const char* msg);

extern "C" int ion_hbm_copy_probe(uint64_t bytes, uint32_t reps, double* best_ms, double* gb_per_s) {
    if (!best_ms || !gb_per_s || bytes < 16 || reps == 0) {
        return ion_internal_set_error(ION_ERR_INVALID, "ion_hbm_copy_probe: bytes >= 16, reps >= 1 and two outputs required");
    }
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemset(a, 0, n * 16);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 1e30f;
    for (uint32_t r = 0; e == hipSuccess && r < reps + 1; r++) {
        e = hipEventRecord(e0, 0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_copy_probe, dim3(65536), dim3(256), 0, 0, (const v4f*)a, (v4f*)b, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0 && ms < best) best = ms;  // the first run warms up
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) {
        return ion_internal_set_error(ION_ERR_DEVICE, (std::string("ion_hbm_copy_probe: ") + hipGetErrorString(e)).c_str());
    }
    *best_ms = best;
    *gb_per_s = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
    return ION_OK;
}

extern "C" int ion_hbm_stream_probe(uint32_t elem_bytes, uint32_t streams, uint32_t cells, uint32_t reps,
                                    double* best_ms, double* gb_per_s) {
    if (!best_ms || !gb_per_s || cells == 0 || reps == 0 || (elem_bytes != 2 && elem_bytes != 4) ||
        (streams != 19 && streams != 45)) {
        return ion_internal_set_error(ION_ERR_INVALID,
                                      "ion_hbm_stream_probe: elem_bytes 2 or 4, streams 19 or 45, cells and reps >= 1");
    }
    const uint64_t bytes = (uint64_t)elem_bytes * streams * cells;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 1e30f;
    for (uint32_t r = 0; e == hipSuccess && r < reps + 1; r++) {
        e = hipEventRecord(e0, 0);
        if (e == hipSuccess) {
            if (elem_bytes == 2) {
                if (streams == 19) launch_stream_probe<19, uint16_t>(a, b, cells);
                else launch_stream_probe<45, uint16_t>(a, b, cells);
            } else {
                if (streams == 19) launch_stream_probe<19, uint32_t>(a, b, cells);
                else launch_stream_probe<45, uint32_t>(a, b, cells);
            }
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0 && ms < best) best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) {
        return ion_internal_set_error(ION_ERR_DEVICE, (std::string("ion_hbm_stream_probe: ") + hipGetErrorString(e)).c_str());
    }
    *best_ms = best;
    *gb_per_s = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    return ION_OK;
}

extern "C" int ion_hbm_pull_probe(uint32_t elem_bytes, uint32_t streams, uint32_t nx, uint32_t ny, uint32_t nz,
                                  uint32_t reps, double* best_ms, double* gb_per_s) {
    // nx * ny first, so that the product of three 32-bit sides cannot wrap
    const uint64_t nxy = (uint64_t)nx * ny;
    const uint64_t N = nxy > (1ull << 30) ? 0 : nxy * nz;
    if (!best_ms || !gb_per_s || N == 0 || N > (1ull << 30) || reps == 0 || (elem_bytes != 2 && elem_bytes != 4) ||
        (streams != 19 && streams != 45)) {
        return ion_internal_set_error(ION_ERR_INVALID, "ion_hbm_pull_probe: elem_bytes 2 or 4, streams 19 or 45, "
                                                       "1 <= nx ny nz <= 2^30, reps >= 1");
    }
    const uint64_t bytes = (uint64_t)elem_bytes * streams * N;
    void* a = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 1e30f;
    const dim3 g((uint32_t)((N + 255) / 256)), b(256);
    for (uint32_t r = 0; e == hipSuccess && r < reps + 1; r++) {
        e = hipEventRecord(e0, 0);
        if (e == hipSuccess) {
            if (elem_bytes == 2) {
                if (streams == 19) hipLaunchKernelGGL((k_pull_probe<19, uint16_t>), g, b, 0, 0, (uint16_t*)a, nx, ny, nz, N);
                else hipLaunchKernelGGL((k_pull_probe<45, uint16_t>), g, b, 0, 0, (uint16_t*)a, nx, ny, nz, N);
            } else {
                if (streams == 19) hipLaunchKernelGGL((k_pull_probe<19, uint32_t>), g, b, 0, 0, (uint32_t*)a, nx, ny, nz, N);
                else hipLaunchKernelGGL((k_pull_probe<45, uint32_t>), g, b, 0, 0, (uint32_t*)a, nx, ny, nz, N);
            }
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0 && ms < best) best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (e != hipSuccess) {
        return ion_internal_set_error(ION_ERR_DEVICE, (std::string("ion_hbm_pull_probe: ") + hipGetErrorString(e)).c_str());
    }
    *best_ms = best;
    *gb_per_s = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    return ION_OK;
}
