/*
 * pt_membench.hip -- the bandwidth ceilings the tracking kernels are measured against.
 *
 * No counterpart in the reference (a measurement aid).  The step kernel's dominant memory
 * traffic is a cell's three 64-coefficient blocks (1536 bytes, twelve 128-byte lines) gathered
 * from a random cell by each lane; this file measures what the device sustains for exactly
 * that pattern (each lane reads its own contiguous block of block_bytes, 512 or 1536, from
 * a uniformly random block of a table of table_bytes, a line's eight 16-byte loads issued
 * together), and for a plain streaming copy.  Tables that fit an XCD's 4 MB L2 are served
 * from there, tables that fit the 256 MB Infinity Cache from it, larger ones from HBM.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/pentrack_gpu.h"
#include "pt_adjoint.h"

namespace ptm_bench {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* each thread gathers `gathers` random blocks of 16 * LINES double2 words (LINES 128-byte
   lines, the step kernel's cell block of three 512-byte components is 12) and keeps a sum;
   the eight 16-byte loads of one line are issued together, as the tricubic evaluation
   issues them */
template <int LINES>
__global__ void __launch_bounds__(256) gather_kernel(const double2 *__restrict__ table, uint64_t nblocks,
                                                     int gathers, uint64_t seed, double *__restrict__ sink) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    for (int g = 0; g < gathers; ++g) {
        /* uniform in [0, nblocks) by a multiply-high of 32 random bits (nblocks < 2^32) */
        const uint64_t b = ((mix64(seed ^ (tid * 0x100000001b3ull + (uint64_t)g)) >> 32) * nblocks) >> 32;
        const double2 *p = table + b * (uint64_t)(8 * LINES);
#pragma unroll
        for (int l = 0; l < LINES; ++l) {
            double2 v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = p[8 * l + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) s += v[q].x + v[q].y;
        }
    }
    sink[tid] = s;
}

/* the step kernels' cooperative form of the same gather: the 64 lanes of a wave load each
   other's blocks in units of one half 128-byte line per lane (lane l loads chunk l & 3 of
   the half line of lane 16j + (l >> 2), j = 0..3: four lanes read 64 contiguous bytes) and
   pass the chunks to their owners through 4 KB of LDS per wave, two units' loads in flight
   ahead of the one being staged; `extra_lds` dynamic bytes per block set the step kernels'
   occupancy (two 256-thread blocks of 80 KB per CU) */
#define PTB_FENCE() asm volatile("" ::: "memory")
struct PtbChunks { double2 a, b, c, d; };
__device__ __forceinline__ PtbChunks ptb_load(const double2 *__restrict__ t, const uint32_t *src, uint32_t off) {
    PtbChunks u;
    u.a = t[src[0] + off]; u.b = t[src[1] + off]; u.c = t[src[2] + off]; u.d = t[src[3] + off];
    return u;
}
__global__ void __launch_bounds__(256) coop_gather_kernel(const double2 *__restrict__ table, uint64_t nblocks,
                                                          int gathers, uint64_t seed, double *__restrict__ sink) {
    extern __shared__ double extra[];
    __shared__ double2 stg[4][4][64];
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, qq = lane & 3, grp = lane >> 2;
    char *S = reinterpret_cast<char *>(stg[threadIdx.x >> 6]);
    const uint32_t wb = (qq << 10) | (((grp ^ (2 * qq)) & 15) << 4);
    const uint32_t rb = lane << 4;
    double s = 0.0;
    for (int g = 0; g < gathers; ++g) {
        const uint64_t b = ((mix64(seed ^ (tid * 0x100000001b3ull + (uint64_t)g)) >> 32) * nblocks) >> 32;
        const uint32_t mine = (uint32_t)b * 96u;
        uint32_t src[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) src[k] = (uint32_t)__shfl((int)mine, 16 * k + (int)grp, 64) + qq;
        auto off = [](int u) -> uint32_t { return (uint32_t)(8 * (11 - u / 2) + 4 * (1 - (u & 1))); };
        PtbChunks buf[3];
        buf[0] = ptb_load(table, src, off(0));
        buf[1] = ptb_load(table, src, off(1));
#pragma unroll
        for (int u = 0; u < 24; ++u) {
            if (u + 2 < 24) buf[(u + 2) % 3] = ptb_load(table, src, off(u + 2));
            const PtbChunks &vc = buf[u % 3];
            *reinterpret_cast<double2 *>(S + wb) = vc.a;
            *reinterpret_cast<double2 *>(S + wb + 256) = vc.b;
            *reinterpret_cast<double2 *>(S + wb + 512) = vc.c;
            *reinterpret_cast<double2 *>(S + wb + 768) = vc.d;
            PTB_FENCE();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 v = *reinterpret_cast<const double2 *>(S + ((rb ^ (uint32_t)(32 * q)) + 1024u * q));
                s += v.x + v.y;
            }
            PTB_FENCE();
        }
    }
    if (s == -1.0) extra[threadIdx.x] = s;   /* never true for the table's values; keeps the allocation */
    sink[tid] = s;
}

__global__ void __launch_bounds__(256) copy_kernel(const double2 *__restrict__ a, double2 *__restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void __launch_bounds__(256) fill_kernel(double2 *__restrict__ a, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = make_double2(1e-3 * (double)(i & 1023), 1.0);
}

constexpr int PTB_COOP_EXTRA_LDS_V = 64 * 1024;

} // namespace ptm_bench

#define PTB_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) { rc = pt_fail(PT_EDEVICE, (std::string(#call ": ") + hipGetErrorString(e_)).c_str()); goto done; } \
    } while (0)

extern "C" int pt_gather_bandwidth(uint64_t table_bytes, int block_bytes, int threads, int gathers, int reps,
                                   double *gather_gbs, double *copy_gbs) {
    if ((block_bytes != 512 && block_bytes != 1536) || table_bytes < (uint64_t)block_bytes ||
        table_bytes / (uint64_t)block_bytes >= (1ull << 32) || threads < 256 ||
        threads % 256 != 0 || gathers < 1 || reps < 1)
        return pt_fail(PT_EINVAL, "pt_gather_bandwidth: bad arguments");
    int rc = PT_OK;
    const uint64_t nblocks = table_bytes / (uint64_t)block_bytes;
    const uint64_t n2 = nblocks * (uint64_t)block_bytes / 16;
    double2 *table = nullptr, *copy = nullptr;
    double *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    PTB_HIP(hipMalloc(&table, n2 * 16));
    PTB_HIP(hipMalloc(&copy, n2 * 16));
    PTB_HIP(hipMalloc(&sink, (size_t)threads * sizeof(double)));
    PTB_HIP(hipEventCreate(&e0));
    PTB_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(ptm_bench::fill_kernel, dim3(4096), dim3(256), 0, 0, table, n2);
    PTB_HIP(hipGetLastError());
    /* warm-up, then the timed repetitions of the gather */
    for (int r = -1; r < reps; ++r) {
        if (r == 0) PTB_HIP(hipEventRecord(e0, 0));
        const uint64_t seed = r < 0 ? 12345ull : 777ull + (uint64_t)r;
        if (block_bytes == 512)
            hipLaunchKernelGGL(ptm_bench::gather_kernel<4>, dim3(threads / 256), dim3(256), 0, 0, table, nblocks,
                               gathers, seed, sink);
        else
            hipLaunchKernelGGL(ptm_bench::gather_kernel<12>, dim3(threads / 256), dim3(256), 0, 0, table, nblocks,
                               gathers, seed, sink);
        PTB_HIP(hipGetLastError());
    }
    PTB_HIP(hipEventRecord(e1, 0));
    PTB_HIP(hipEventSynchronize(e1));
    PTB_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (gather_gbs) *gather_gbs = (double)threads * gathers * block_bytes * reps / (ms * 1e-3) / 1e9;
    /* streaming copy of the table (read + write bytes) */
    hipLaunchKernelGGL(ptm_bench::copy_kernel, dim3(8192), dim3(256), 0, 0, table, copy, n2);
    PTB_HIP(hipGetLastError());
    PTB_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(ptm_bench::copy_kernel, dim3(8192), dim3(256), 0, 0, table, copy, n2);
    PTB_HIP(hipEventRecord(e1, 0));
    PTB_HIP(hipEventSynchronize(e1));
    PTB_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (copy_gbs) *copy_gbs = 2.0 * (double)n2 * 16 * reps / (ms * 1e-3) / 1e9;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (copy) (void)hipFree(copy);
    if (table) (void)hipFree(table);
    return rc;
}

extern "C" int pt_gather_bandwidth_coop(uint64_t table_bytes, int threads, int gathers, int reps, double *gather_gbs) {
    if (table_bytes < 1536 || table_bytes / 1536 >= (1ull << 32) / 96 || threads < 256 || threads % 256 != 0 ||
        gathers < 1 || reps < 1)
        return pt_fail(PT_EINVAL, "pt_gather_bandwidth_coop: bad arguments");
    int rc = PT_OK;
    const uint64_t nblocks = table_bytes / 1536;
    const uint64_t n2 = nblocks * 96;
    double2 *table = nullptr;
    double *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    PTB_HIP(hipFuncSetAttribute((const void *)ptm_bench::coop_gather_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                ptm_bench::PTB_COOP_EXTRA_LDS_V));
    PTB_HIP(hipMalloc(&table, n2 * 16));
    PTB_HIP(hipMalloc(&sink, (size_t)threads * sizeof(double)));
    PTB_HIP(hipEventCreate(&e0));
    PTB_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(ptm_bench::fill_kernel, dim3(4096), dim3(256), 0, 0, table, n2);
    PTB_HIP(hipGetLastError());
    for (int r = -1; r < reps; ++r) {
        if (r == 0) PTB_HIP(hipEventRecord(e0, 0));
        const uint64_t seed = r < 0 ? 12345ull : 777ull + (uint64_t)r;
        hipLaunchKernelGGL(ptm_bench::coop_gather_kernel, dim3(threads / 256), dim3(256), ptm_bench::PTB_COOP_EXTRA_LDS_V,
                           0, table, nblocks, gathers, seed, sink);
        PTB_HIP(hipGetLastError());
    }
    PTB_HIP(hipEventRecord(e1, 0));
    PTB_HIP(hipEventSynchronize(e1));
    PTB_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (gather_gbs) *gather_gbs = (double)threads * gathers * 1536.0 * reps / (ms * 1e-3) / 1e9;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (table) (void)hipFree(table);
    return rc;
}
