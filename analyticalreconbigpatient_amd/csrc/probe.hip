// Bandwidth probes for the backprojector's roofs (diagnostics; no counterpart
// in the reference).
//
// The LDS-staged backprojector reads every voxel-update's samples from LDS
// (ds_read2_b32 / ds_read2_b64: 8 or 16 bytes per lane), after the footprint
// of each view has been staged from L2/HBM.  Its algorithmic sample traffic
// (16 B per voxel-update in FP32, SURVEY's gather roof) is therefore bounded by
// LDS read bandwidth, and the gathering z-column kernel's by L2 gathers.  These
// kernels measure what the device delivers for those two access kinds:
//
//  probe_lds<B>  every lane reads B bytes (8: ds_read_b64, 16: ds_read_b128)
//                at lane-consecutive addresses of a 16 KB LDS buffer (no bank
//                conflicts), 8 workgroups of 256 lanes per CU;
//  probe_l2      every lane reads 16 bytes from a device buffer of `span`
//                bytes (small enough to stay in L2) at scattered 16-byte
//                aligned offsets, as the gathering kernel's row-pair loads do.
//
// Each returns the bytes it reads; the caller times the launch with events.
#include "arecon_internal.h"

#include <algorithm>

namespace arecon {
namespace {

template <int B>
__global__ __launch_bounds__(256) void probe_lds(float* __restrict__ sink, int iters) {
    __shared__ __attribute__((aligned(16))) float buf[4096];      // 16 KB
    for (int i = threadIdx.x; i < 4096; i += 256) buf[i] = (float)(i & 255) * 0.5f;
    __syncthreads();
    const char* base = reinterpret_cast<const char*>(buf);
    unsigned a = threadIdx.x * B;
    float acc0 = 0.f, acc1 = 0.f;
    for (int it = 0; it < iters; it += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned off = (a + (unsigned)u * 256u * B) & 16383u;
            if constexpr (B == 8) {
                const float2 v = *reinterpret_cast<const float2*>(base + off);
                acc0 += v.x;
                acc1 += v.y;
            } else {
                const float4 v = *reinterpret_cast<const float4*>(base + off);
                acc0 += v.x + v.z;
                acc1 += v.y + v.w;
            }
        }
        a += 1024u * B + 64u;
    }
    sink[(long long)blockIdx.x * blockDim.x + threadIdx.x] = acc0 + acc1;
}

__global__ __launch_bounds__(256) void probe_l2(float* __restrict__ sink,
                                                const float4* __restrict__ src,
                                                unsigned span16, int iters) {
    unsigned h = (blockIdx.x * 256u + threadIdx.x) * 2654435761u;
    float acc = 0.f;
    for (int it = 0; it < iters; it += 4) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            h = h * 1664525u + 1013904223u;
            v[u] = src[(h >> 8) % span16];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    sink[(long long)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace
}  // namespace arecon

using namespace arecon;

extern "C" int arecon_probe_lds(float* d_sink, int bytes_per_lane, int iters, int blocks,
                                double* bytes_read, void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(d_sink && bytes_read, "arecon_probe_lds: null pointer");
        ARECON_REQUIRE(bytes_per_lane == 8 || bytes_per_lane == 16,
                       "arecon_probe_lds: 8 or 16 bytes per lane");
        ARECON_REQUIRE(iters > 0 && iters % 4 == 0 && blocks > 0,
                       "arecon_probe_lds: iters a positive multiple of 4, blocks > 0");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (bytes_per_lane == 8)
            hipLaunchKernelGGL(probe_lds<8>, dim3(blocks), dim3(256), 0, s, d_sink, iters);
        else
            hipLaunchKernelGGL(probe_lds<16>, dim3(blocks), dim3(256), 0, s, d_sink, iters);
        ARECON_HIP(hipGetLastError());
        *bytes_read = (double)blocks * 256.0 * iters * bytes_per_lane;
    });
}

extern "C" int arecon_probe_l2(float* d_sink, const void* d_src, long long span_bytes,
                               int iters, int blocks, double* bytes_read, void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(d_sink && d_src && bytes_read, "arecon_probe_l2: null pointer");
        ARECON_REQUIRE(span_bytes >= 16 && span_bytes <= (1LL << 32),
                       "arecon_probe_l2: span between 16 bytes and 4 GiB");
        ARECON_REQUIRE(iters > 0 && iters % 4 == 0 && blocks > 0,
                       "arecon_probe_l2: iters a positive multiple of 4, blocks > 0");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(probe_l2, dim3(blocks), dim3(256), 0, s, d_sink,
                           static_cast<const float4*>(d_src), (unsigned)(span_bytes / 16), iters);
        ARECON_HIP(hipGetLastError());
        *bytes_read = (double)blocks * 256.0 * iters * 16.0;
    });
}

namespace arecon {
namespace {

// Streaming copy, 16 bytes per lane per step: calibrates the HBM byte counters
// (a copy of n bytes reads n and writes n).
__global__ __launch_bounds__(256) void probe_copy(float4* __restrict__ dst,
                                                  const float4* __restrict__ src, long long n16) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace
}  // namespace arecon

extern "C" int arecon_probe_copy(void* d_dst, const void* d_src, long long bytes, void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(d_dst && d_src && bytes > 0 && bytes % 16 == 0,
                       "arecon_probe_copy: non-null buffers and a multiple of 16 bytes");
        int dev = 0, cus = 0;
        ARECON_HIP(hipGetDevice(&dev));
        ARECON_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long long n16 = bytes / 16;
        const long long grid = std::max<long long>(1, std::min<long long>((n16 + 255) / 256,
                                                                          (long long)cus * 32));
        hipLaunchKernelGGL(probe_copy, dim3((unsigned)grid), dim3(256), 0,
                           static_cast<hipStream_t>(stream), static_cast<float4*>(d_dst),
                           static_cast<const float4*>(d_src), n16);
        ARECON_HIP(hipGetLastError());
    });
}
