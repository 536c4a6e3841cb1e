// Fan-to-parallel rebinning on the curved detector (gfx950).
//
// Behaviour follows the reference's CPU path, src/ParallelRebinningCBCurve.cpp:22-97:
// for parallel view i and channel j,
//   Theta = i*DeltaTheta, t = (j-PLC)*DeltaT, Beta = asin(t/SO),
//   FaiIndex = (Theta+Beta)/DeltaFai, UIndex = atan(t/sqrt(SO^2-t^2))/dYA + YLC,
// a 2x2 linear blend of fan views ceil(FaiIndex)-1/ceil(FaiIndex) and channels
// ceil(UIndex)-1/ceil(UIndex), with the CPU path's ">" edge clamps (:55-85), and
// rows passed straight through (:87-94).
//
// Design.  Everything that depends on the channel only (Beta, the channel pair
// and its weights) is computed once per call on the host, in the reference's
// own arithmetic (double libm rounded to float), and read from a small table.
// The kernels do the view-dependent part in the same order with FMA contraction
// switched off, so their result is bit-identical to the float CPU path.
//
//   rebin_vyz  output [ViewN][YL][ZL] (reference layout): one thread per
//              4 consecutive rows of one (view, channel), float4 loads/stores.
//   rebin_vzy  output [ViewN][ZL][YL] (the filter's layout): one workgroup per
//              (view, block of JB channels); the ZL-contiguous input vectors
//              are blended with coalesced loads into an LDS tile [ZL][JB + 1]
//              and written out transposed with coalesced stores.  JB = 128
//              (512-byte row stores) while the tile fits 48 KB (ZL <= 95),
//              else 64 (ZL <= 242, the 64 KB workgroup limit with the static
//              blend tables).  Workgroups are remapped so each XCD walks a
//              contiguous run of views (neighbouring output views share fan
//              views, whose second read then hits the same L2).
// HBM roofline: 8 B per sample (one read, one write); the 4-tap gather is
// served by L2 (neighbouring channels share fan views).
#include "arecon_internal.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace arecon {
namespace {

struct ChanCoef {
    float beta;   // asin(t/SO), rounded to float
    float cyb;    // coeYB
    float cyu;    // coeYU
    int iyb;      // IndexYB
    int iyu;      // IndexYU
    int pad[3];
};
static_assert(sizeof(ChanCoef) == 32, "ChanCoef layout");

// Host: channel table in the reference's arithmetic (:37-48, :71-85).
std::vector<ChanCoef> channel_table(int YL, float YLC, float dYA, float PLC,
                                    float DeltaT, float SO) {
    std::vector<ChanCoef> c(YL);
    for (int j = 0; j < YL; ++j) {
        volatile float t = ((float)j - PLC) * DeltaT;
        volatile float q = t / SO;
        volatile float beta = (float)std::asin((double)q);
        volatile float r2 = SO * SO - t * t;
        volatile float a = (float)std::atan((double)t / std::sqrt((double)r2));
        volatile float ad = a / dYA;
        volatile float uindex = ad + YLC;
        const int UI = (int)std::ceil((double)uindex);
        volatile float cyb = (float)UI - uindex;
        volatile float cyu = (float)(1.0 - (double)cyb);
        ChanCoef e{};
        e.beta = beta;
        e.cyb = cyb;
        e.cyu = cyu;
        if (UI <= 0) { e.iyb = 0; e.iyu = 0; }
        else if (UI > YL - 1) { e.iyb = YL - 1; e.iyu = YL - 1; }
        else { e.iyb = UI - 1; e.iyu = UI; }
        c[j] = e;
    }
    return c;
}

// Device copies of channel tables, keyed by (device, geometry bits).  A table
// is uploaded once with a blocking copy and kept for the life of the process:
// captured HIP graphs and plans hold its address, so no table is ever freed
// (about 24 bytes per channel and geometry).
std::mutex g_chan_mu;
std::map<std::array<uint32_t, 7>, ChanCoef*> g_chan_cache;

uint32_t bits(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
}

const ChanCoef* channel_table_dev(int YL, float YLC, float dYA, float PLC, float DeltaT,
                                  float SO) {
    int dev = 0;
    ARECON_HIP(hipGetDevice(&dev));
    const std::array<uint32_t, 7> key{(uint32_t)dev, (uint32_t)YL, bits(YLC), bits(dYA),
                                      bits(PLC), bits(DeltaT), bits(SO)};
    std::lock_guard<std::mutex> lk(g_chan_mu);
    auto it = g_chan_cache.find(key);
    if (it != g_chan_cache.end()) return it->second;
    const std::vector<ChanCoef> tab = channel_table(YL, YLC, dYA, PLC, DeltaT, SO);
    ChanCoef* d = nullptr;
    ARECON_HIP(hipMalloc(&d, sizeof(ChanCoef) * YL));
    ARECON_HIP(hipMemcpy(d, tab.data(), sizeof(ChanCoef) * YL, hipMemcpyHostToDevice));
    g_chan_cache.emplace(key, d);
    return d;
}

// View-dependent part of the blend weights (:34,:39,:41,:43,:45-46,:55-69).
// FMA contraction is off here so every operation rounds as the CPU's does.
struct Blend {
    float c00, c10, c01, c11;   // (XB,YB) (XU,YB) (XB,YU) (XU,YU)
    int ixb, ixu;
};

__device__ __forceinline__ Blend view_blend(int i, const ChanCoef& cc, int ViewN,
                                            float DeltaTheta, float DeltaFai) {
#pragma clang fp contract(off)
    const float theta = (float)i * DeltaTheta;
    const float fai = theta + cc.beta;
    const float fidx = fai / DeltaFai;          // IEEE division (correctly rounded)
    const float fic = ceilf(fidx);
    const int FI = (int)fic;
    const float cxb = fic - fidx;
    const float cxu = 1.0f - cxb;
    Blend b;
    if (FI <= 0) { b.ixb = 0; b.ixu = 0; }
    else if (FI > ViewN - 1) { b.ixb = ViewN - 1; b.ixu = ViewN - 1; }
    else { b.ixb = FI - 1; b.ixu = FI; }
    b.c00 = cxb * cc.cyb;
    b.c10 = cxu * cc.cyb;
    b.c01 = cxb * cc.cyu;
    b.c11 = cxu * cc.cyu;
    return b;
}

// ((c00*p0 + c10*p1) + c01*p2) + c11*p3, each operation rounded (:89-93).
__device__ __forceinline__ float blend4(const Blend& b, float p0, float p1, float p2,
                                        float p3) {
#pragma clang fp contract(off)
    return b.c00 * p0 + b.c10 * p1 + b.c01 * p2 + b.c11 * p3;
}

// [ViewN][YL][ZL] output, ZL % 4 == 0: one thread per float4 of rows.
__global__ __launch_bounds__(256) void rebin_vyz_v4(
    float4* __restrict__ out, const float4* __restrict__ in,
    const ChanCoef* __restrict__ chan, int YL, int ZL4, int ViewN, float DeltaTheta,
    float DeltaFai, long long total) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int kq = (int)(gid % ZL4);
    const long long ij = gid / ZL4;
    const int j = (int)(ij % YL);
    const int i = (int)(ij / YL);
    const ChanCoef cc = chan[j];
    const Blend b = view_blend(i, cc, ViewN, DeltaTheta, DeltaFai);
    const float4 p0 = in[((long long)b.ixb * YL + cc.iyb) * ZL4 + kq];
    const float4 p1 = in[((long long)b.ixu * YL + cc.iyb) * ZL4 + kq];
    const float4 p2 = in[((long long)b.ixb * YL + cc.iyu) * ZL4 + kq];
    const float4 p3 = in[((long long)b.ixu * YL + cc.iyu) * ZL4 + kq];
    float4 o;
    o.x = blend4(b, p0.x, p1.x, p2.x, p3.x);
    o.y = blend4(b, p0.y, p1.y, p2.y, p3.y);
    o.z = blend4(b, p0.z, p1.z, p2.z, p3.z);
    o.w = blend4(b, p0.w, p1.w, p2.w, p3.w);
    out[gid] = o;
}

// [ViewN][YL][ZL] output, any ZL: one thread per sample.
__global__ __launch_bounds__(256) void rebin_vyz_s(
    float* __restrict__ out, const float* __restrict__ in,
    const ChanCoef* __restrict__ chan, int YL, int ZL, int ViewN, float DeltaTheta,
    float DeltaFai, long long total) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int k = (int)(gid % ZL);
    const long long ij = gid / ZL;
    const int j = (int)(ij % YL);
    const int i = (int)(ij / YL);
    const ChanCoef cc = chan[j];
    const Blend b = view_blend(i, cc, ViewN, DeltaTheta, DeltaFai);
    const float p0 = in[((long long)b.ixb * YL + cc.iyb) * ZL + k];
    const float p1 = in[((long long)b.ixu * YL + cc.iyb) * ZL + k];
    const float p2 = in[((long long)b.ixb * YL + cc.iyu) * ZL + k];
    const float p3 = in[((long long)b.ixu * YL + cc.iyu) * ZL + k];
    out[gid] = blend4(b, p0, p1, p2, p3);
}

// [ViewN][ZL][YL] output.  Block = (view i, channels j0..j0+JB-1).
// LDS tile[ZL][JB+1] (padded against bank conflicts on the transposed read).
// One-dimensional grid, channel blocks fastest, remapped so that each XCD (hardware
// workgroup b runs on XCD b % 8) takes a contiguous run of views: a fan view
// feeds two neighbouring output views, and with the plain round-robin order
// those ran on different XCDs, so each L2 fetched it from HBM again.
template <int JB>
__global__ __launch_bounds__(256) void rebin_vzy(
    float* __restrict__ out, const float* __restrict__ in,
    const ChanCoef* __restrict__ chan, int YL, int ZL, int ViewN, float DeltaTheta,
    float DeltaFai, int view_begin, int view_count, int vec4) {
    extern __shared__ float tile[];             // ZL * (JB + 1)
    __shared__ Blend sb[JB];
    __shared__ int sbase[JB][4];
    const int nb = gridDim.x, hb = blockIdx.x;
    const int qx = nb >> 3, rx = nb & 7, xcd = hb & 7;
    const int lb = xcd * qx + min(xcd, rx) + (hb >> 3);
    const int ncb = (YL + JB - 1) / JB;
    const int vi = lb / ncb;
    const int i = view_begin + vi;
    const int j0 = (lb - vi * ncb) * JB;
    const int nj = min(JB, YL - j0);
    const int tid = threadIdx.x;
    if (tid < nj) {
        const ChanCoef cc = chan[j0 + tid];
        const Blend b = view_blend(i, cc, ViewN, DeltaTheta, DeltaFai);
        sb[tid] = b;
        sbase[tid][0] = b.ixb * YL + cc.iyb;
        sbase[tid][1] = b.ixu * YL + cc.iyb;
        sbase[tid][2] = b.ixb * YL + cc.iyu;
        sbase[tid][3] = b.ixu * YL + cc.iyu;
    }
    __syncthreads();
    if (vec4) {   // ZL % 4 == 0 and a 16-byte aligned input
        // four rows per thread: 16-byte taps (a quarter of the address work of
        // 4-byte loads; the 4-tap gather is otherwise bound by it)
        const int ZL4 = ZL >> 2;
        const int n = nj * ZL4;
        const float4* in4 = reinterpret_cast<const float4*>(in);
        for (int e = tid; e < n; e += blockDim.x) {
            const int jj = e / ZL4;
            const int k4 = e - jj * ZL4;
            const Blend b = sb[jj];
            const float4 p0 = in4[(long long)sbase[jj][0] * ZL4 + k4];
            const float4 p1 = in4[(long long)sbase[jj][1] * ZL4 + k4];
            const float4 p2 = in4[(long long)sbase[jj][2] * ZL4 + k4];
            const float4 p3 = in4[(long long)sbase[jj][3] * ZL4 + k4];
            float* t = tile + 4 * k4 * (JB + 1) + jj;
            t[0] = blend4(b, p0.x, p1.x, p2.x, p3.x);
            t[JB + 1] = blend4(b, p0.y, p1.y, p2.y, p3.y);
            t[2 * (JB + 1)] = blend4(b, p0.z, p1.z, p2.z, p3.z);
            t[3 * (JB + 1)] = blend4(b, p0.w, p1.w, p2.w, p3.w);
        }
    } else {
        const int n = nj * ZL;
        for (int e = tid; e < n; e += blockDim.x) {
            const int jj = e / ZL;
            const int k = e - jj * ZL;
            const Blend b = sb[jj];
            const float p0 = in[(long long)sbase[jj][0] * ZL + k];
            const float p1 = in[(long long)sbase[jj][1] * ZL + k];
            const float p2 = in[(long long)sbase[jj][2] * ZL + k];
            const float p3 = in[(long long)sbase[jj][3] * ZL + k];
            tile[k * (JB + 1) + jj] = blend4(b, p0, p1, p2, p3);
        }
    }
    __syncthreads();
    for (int e = tid; e < ZL * JB; e += blockDim.x) {
        const int k = e / JB;
        const int jj = e - k * JB;
        if (jj < nj)
            out[((long long)i * ZL + k) * YL + j0 + jj] = tile[k * (JB + 1) + jj];
    }
}

}  // namespace

void launch_rebin(float* out, const float* in, int YL, int ZL, int ViewN, float YLC,
                  float dYA, float DeltaTheta, float PLC, float DeltaT, float DeltaFai,
                  float SO, int out_layout, hipStream_t s, int view_begin, int view_count) {
    ARECON_REQUIRE(YL > 0 && ZL > 0 && ViewN > 0, "rebin: YL, ZL and ViewN must be positive");
    if (view_count < 0) view_count = ViewN - view_begin;
    ARECON_REQUIRE(view_begin >= 0 && view_count > 0 && view_begin + view_count <= ViewN,
                   "rebin: view range outside [0, ViewN)");
    ARECON_REQUIRE(out_layout == ARECON_LAYOUT_VZY || (view_begin == 0 && view_count == ViewN),
                   "rebin: view ranges need the transposed output layout");
    ARECON_REQUIRE(out != nullptr && in != nullptr, "rebin: null buffer");
    ARECON_REQUIRE(out != in, "rebin: input and output must not alias");
    ARECON_REQUIRE((long long)YL * ViewN < (1LL << 31),
                   "rebin: YL*ViewN must fit in 31 bits");
    ARECON_REQUIRE(out_layout == ARECON_LAYOUT_VYZ || out_layout == ARECON_LAYOUT_VZY,
                   "rebin: unknown output layout");
    const ChanCoef* chan = channel_table_dev(YL, YLC, dYA, PLC, DeltaT, SO);
    if (out_layout == ARECON_LAYOUT_VYZ) {
        if (ZL % 4 == 0) {
            const long long total = (long long)ViewN * YL * (ZL / 4);
            const unsigned grid = (unsigned)((total + 255) / 256);
            hipLaunchKernelGGL(rebin_vyz_v4, dim3(grid), dim3(256), 0, s,
                               reinterpret_cast<float4*>(out),
                               reinterpret_cast<const float4*>(in), chan,
                               YL, ZL / 4, ViewN, DeltaTheta, DeltaFai, total);
        } else {
            const long long total = (long long)ViewN * YL * ZL;
            const unsigned grid = (unsigned)((total + 255) / 256);
            hipLaunchKernelGGL(rebin_vyz_s, dim3(grid), dim3(256), 0, s, out, in,
                               chan, YL, ZL, ViewN, DeltaTheta, DeltaFai,
                               total);
        }
    } else {
        // 128-channel blocks (512-byte row stores, 1.21 ms against 1.26 ms for
        // 64 on the 888 x 64 x 11,808 sinogram) while the dynamic tile stays
        // within 48 KB of LDS (ZL <= 95), else 64-channel blocks.  The static
        // blend tables add JB * 40 bytes (5 KB at 128, 2.5 KB at 64); a
        // workgroup may hold 64 KB in all, so the 64-channel blocks take
        // ZL <= 242.
        const int vec4 = (ZL % 4 == 0) && ((reinterpret_cast<uintptr_t>(in) & 15) == 0);
        auto launch = [&](auto jb) {
            constexpr int JB = decltype(jb)::value;
            const size_t lds = sizeof(float) * (size_t)ZL * (JB + 1);
            const long long nblocks = (long long)view_count * ((YL + JB - 1) / JB);
            ARECON_REQUIRE(nblocks < (1LL << 31), "rebin: too many workgroups");
            hipLaunchKernelGGL(rebin_vzy<JB>, dim3((unsigned)nblocks), dim3(256), lds, s, out,
                               in, chan, YL, ZL, ViewN, DeltaTheta, DeltaFai, view_begin,
                               view_count, vec4);
        };
        if (sizeof(float) * (size_t)ZL * 129 <= 48 * 1024) {
            launch(std::integral_constant<int, 128>{});
        } else {
            ARECON_REQUIRE(sizeof(float) * (size_t)ZL * 65 + sizeof(Blend) * 64 +
                                   sizeof(int) * 4 * 64 <= 64 * 1024,
                           "rebin: ZL above 242 does not fit the transposing kernel's LDS "
                           "(output layout [ViewN][ZL][YL])");
            launch(std::integral_constant<int, 64>{});
        }
    }
    ARECON_HIP(hipGetLastError());
}

}  // namespace arecon
