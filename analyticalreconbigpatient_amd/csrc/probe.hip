// Bandwidth probes for the backprojector's roofs (diagnostics; no counterpart
// in the reference).
//
// The LDS-staged backprojector reads every voxel-update's samples from LDS
// (ds_read2_b32 / ds_read2_b64: 8 or 16 bytes per lane), after the footprint
// of each view has been staged from L2/HBM.  Its algorithmic sample traffic
// (16 B per voxel-update in FP32, the gather roof) is therefore bounded by
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
//
//  probe_copy    a streaming copy (calibrates the HBM byte counters);
//  probe_valu<F> issue rate of one VALU instruction form: every lane runs
//                eight independent chains of it, so the rate is the form's
//                throughput (wavefront instructions per CU per clock) with
//                enough waves to hide its latency.  The forms are the ones the
//                backprojector's inner loop is built from or was tried with:
//                FP32 FMA with VGPR operands and with a scalar operand (the
//                per-slice table value), the mixed FMA that blends FP16
//                records, its FP16-result forms (mixlo / mixhi), the packed FP16
//                dot product and FMA, the floor conversion, the shift-add of
//                the LDS address, and the reciprocal of the 3D weight.
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

namespace arecon {
namespace {

enum { VF_FMAC = 0, VF_FMAC_SGPR, VF_FMA_MIX, VF_FMA_MIXLO, VF_DOT2, VF_PK_FMA_F16, VF_CVT_FLR,
       VF_LSHL_ADD, VF_RCP, VF_COUNT };
constexpr int kValuIters = 4096;

template <int F>
__global__ __launch_bounds__(256) void probe_valu(float* __restrict__ sink, float seed, float sv) {
    float a[8];
    unsigned u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = seed + threadIdx.x * 0.001f + i;
        u[i] = __float_as_uint(a[i]) ^ 0x3c003c00u;
    }
    const float fr = seed * 0.5f;
    const unsigned w = 0x3c003800u;
    for (int it = 0; it < kValuIters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (F == VF_FMAC)
                asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(a[i]) : "v"(fr), "v"(a[(i + 1) & 7]));
            if constexpr (F == VF_FMAC_SGPR)
                asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(a[i]) : "s"(sv), "v"(a[(i + 1) & 7]));
            if constexpr (F == VF_FMA_MIX)
                asm volatile("v_fma_mix_f32 %0, %1, %2, %2 op_sel:[0,1,0] op_sel_hi:[0,1,1]"
                             : "=v"(a[i]) : "v"(fr), "v"(u[i]));
            if constexpr (F == VF_FMA_MIXLO)
                asm volatile("v_fma_mixlo_f16 %0, %1, %2, %2 op_sel:[0,1,0] op_sel_hi:[0,1,1]"
                             : "+v"(u[i]) : "v"(fr), "v"(u[(i + 3) & 7]));
            if constexpr (F == VF_DOT2)
                asm volatile("v_dot2_f32_f16 %0, %1, %2, %0" : "+v"(a[i]) : "v"(u[i]), "v"(w));
            if constexpr (F == VF_PK_FMA_F16)
                asm volatile("v_pk_fma_f16 %0, %1, %2, %0" : "+v"(u[i]) : "v"(w), "v"(u[(i + 5) & 7]));
            if constexpr (F == VF_CVT_FLR)
                asm volatile("v_cvt_flr_i32_f32 %0, %1" : "=v"(u[i]) : "v"(a[i]));
            if constexpr (F == VF_LSHL_ADD)
                asm volatile("v_lshl_add_u32 %0, %1, 7, %2"
                             : "=v"(u[i]) : "v"(u[(i + 1) & 7]), "v"(u[(i + 2) & 7]));
            if constexpr (F == VF_RCP)
                asm volatile("v_rcp_f32 %0, %1" : "=v"(a[i]) : "v"(a[(i + 1) & 7]));
        }
    }
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) t += a[i] + __uint_as_float(u[i] & 0x3fffffffu);
    sink[(long long)blockIdx.x * blockDim.x + threadIdx.x] = t;
}

}  // namespace
}  // namespace arecon

extern "C" int arecon_probe_valu(float* d_sink, int form, int blocks, double* instructions,
                                 void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(d_sink && instructions, "arecon_probe_valu: null pointer");
        ARECON_REQUIRE(form >= 0 && form < VF_COUNT && blocks > 0,
                       "arecon_probe_valu: unknown form or no blocks");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const dim3 g(blocks), b(256);
        switch (form) {
        case VF_FMAC: hipLaunchKernelGGL(probe_valu<VF_FMAC>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_FMAC_SGPR: hipLaunchKernelGGL(probe_valu<VF_FMAC_SGPR>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_FMA_MIX: hipLaunchKernelGGL(probe_valu<VF_FMA_MIX>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_FMA_MIXLO: hipLaunchKernelGGL(probe_valu<VF_FMA_MIXLO>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_DOT2: hipLaunchKernelGGL(probe_valu<VF_DOT2>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_PK_FMA_F16: hipLaunchKernelGGL(probe_valu<VF_PK_FMA_F16>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_CVT_FLR: hipLaunchKernelGGL(probe_valu<VF_CVT_FLR>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        case VF_LSHL_ADD: hipLaunchKernelGGL(probe_valu<VF_LSHL_ADD>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        default: hipLaunchKernelGGL(probe_valu<VF_RCP>, g, b, 0, s, d_sink, 1.f, 0.5f); break;
        }
        ARECON_HIP(hipGetLastError());
        *instructions = (double)blocks * 4.0 * kValuIters * 8.0;   // wavefront instructions
    });
}
