// 3D-weighted helical FDK backprojection in the cone-parallel geometry (gfx950).
//
// Reference behaviour: src/Parallel_FDK_Helical_3DWeighting.cpp:42-113 (CPU path;
// the view loop covers the voxel's window [s1, s2], :71).  For voxel (x, y, z):
//   s0 = ceil((2 pi z / h - BetaS) / DeltaFai - 0.5), window s0 -+ W,
//   W = ceil(N_pi HSCoef), clamped to [0, PN-1], skipped unless s0-W < PN and
//   s0+W > 0 (:63-70).  For each view p of the window, with V = BetaS + p DeltaFai:
//     UU = -x cosV - y sinV,  Yr = -x sinV + y cosV,  temp1 = sqrt(SO^2 - Yr^2),
//     temp2 = z - h V / 2pi,  Zr = temp2 SO SD / (UU temp1 + SO^2 - Yr^2),
//     bilinear sample at (Yr/dYL + YLC, Zr/dZL + ZLC) if 0 < ceil(.) < YL / ZL,
//     weight = w1 Gc^k1 / (w2 G^k1 + w1 Gc^k1),  G = |temp2/(temp1+UU)|,
//     Gc = |(temp2 -+ 0.5)/(temp1-UU)|  (- before s0, + from s0 on),
//     w1 = w[N_pi - (s0 - p)], w2 = w[w1 index +- N_pi]   (:73-104).
//
// Design (MI355X-first).
//  * One wavefront = an 8x8 patch of (x, y) columns at the same slices, so the
//    window, the per-view table entries and the weight-table entries are
//    wavefront-uniform (scalar loads, scalar branches); the 64 gathers of a load
//    instruction fall on a few neighbouring detector cache lines.
//  * Each thread owns NZ consecutive slices of its column.  Everything that
//    depends only on (column, view) -- UU, Yr, temp1, the channel and its blend
//    weight, the magnification and the ratio (temp1-UU)/(temp1+UU) -- is computed
//    once per view and reused by the NZ slices; a slice whose window does not
//    contain the view skips it with a uniform branch, so every voxel still sees
//    exactly its own window.
//  * FP64 on the host for everything that cancels over thousands of views: the
//    per-view cos/sin table, the per-slice s0 and the slice's offset from its
//    central view's source height, z - h V(s0) / 2pi.  In the kernel
//    temp2 = offset - (p - s0) h DeltaFai / 2pi, both terms small.
//  * The 3D weight uses one ratio: weight = w1 / (w1 + w2 (G/Gc)^k1), with
//    (G/Gc)^k1 = exp2(k1 log2 ratio).  Inside the window's centre (no conjugate
//    view in the window) w2 = 0 and the weight is exactly 1.
//  * Workgroups are remapped so that each of the 8 XCDs walks a contiguous run of
//    the logical (slice set, column tile) sequence and keeps its own L2 warm.
// Roofline: 16 B of gathered FP32 samples (8 B FP16/BF16) per voxel-update,
// served from L2/MALL; compulsory HBM traffic is the filtered sinogram once and
// the volume once.  VALU (transcendentals) is the co-bound.
#include "arecon_internal.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <type_traits>
#include <vector>

namespace arecon {

namespace {

struct BPZ {        // per slice
    int s0;         // central view
    int s1, s2;     // clamped window; s1 > s2 when the slice gets no view
    float dz;       // z - h (BetaS + s0 DeltaFai) / 2pi
};

struct BPArgs {
    int XN, YN, ZN, YL, ZL, PN, W;
    int tilesx, ntiles, nzsets, order, k1zero;
    int lb_offset;                    // first logical workgroup of a partial launch
    int vec4;                         // float4 stores allowed (ZN % 4 == 0, aligned output)
    float XNC, YNC, dx, dy;
    float YLC, ZLC, invdYL, KSD;      // KSD = SO SD / dZL
    float YLm1, ZLm1;                 // YL-1, ZL-1 as floats
    float SO2, hstep, k1, DeltaFai;   // hstep = h DeltaFai / 2pi
    long long view_stride;            // ZL * YL samples
    unsigned view_bytes;              // ZL * YL * sizeof(sample)
    unsigned long long* stats;        // LDS-staged kernel: workgroups staged / direct
    int inside;                       // every column well inside the source circle and
                                      // the detector (z-column kernels' INSIDE path)
    // LDS-staged kernel: the plan's footprint table (per slice block and tile, one
    // entry per view of the block, padded to a multiple of four), its per-block
    // offsets and per-(block, tile) flags (1: every footprint fits the slots)
    const uint2* zs_tab;
    const long long* zs_boff;
    const int* zs_fits;
    // LDS-staged kernel: one record per (block, view), at the view's mask index:
    // the slice masks, the view's cos / sin and the block's ZB (temp2, a) entries
    // (4 + 2 ZB words), so a view's scalar operands come in one load from an
    // offset that is a constant within a chunk
    const unsigned* zs_rec;
};

// Logical workgroup lb -> slice block zs and 16x16-column tile (txb, tyb).
// order 0: slice-major, the tiles of a block row by row; 1: tile-major;
// 2: slice-major, the tiles of a block column by column within bands of
// kTileBand tile rows.  The LDS-staged kernel is sensitive to the order (L2
// reuse of the staged footprints between the ~160 workgroups resident on an
// XCD): on the 512x512x400 clinical grid, bands of 6 rows take 29.9 ms against
// 31.2 ms row by row (bands of 3 to 12 rows: 30.2-30.4 ms; tile-major: 42 ms).
constexpr int kTileBand = 6;
__device__ __forceinline__ void wg_tile(int lb, const BPArgs& A, int& zs, int& tyb, int& txb) {
    int tile;
    if (A.order == 1) { tile = lb / A.nzsets; zs = lb - tile * A.nzsets; }
    else { zs = lb / A.ntiles; tile = lb - zs * A.ntiles; }
    if (A.order != 2) {
        tyb = tile / A.tilesx;
        txb = tile - tyb * A.tilesx;
        return;
    }
    const int tilesy = A.ntiles / A.tilesx;
    const int band = tile / (kTileBand * A.tilesx);
    const int y0 = band * kTileBand;
    const int bh = min(kTileBand, tilesy - y0);           // rows in this band
    const int idx = tile - band * kTileBand * A.tilesx;
    txb = idx / bh;
    tyb = y0 + idx - txb * bh;
}

// Two adjacent detector samples (channels U-1, U of one row) as floats.
template <class T> __device__ __forceinline__ float2 load2(const T* p);

typedef float f32x2_u __attribute__((ext_vector_type(2), aligned(4)));
template <> __device__ __forceinline__ float2 load2<float>(const float* p) {
    const f32x2_u v = *reinterpret_cast<const f32x2_u*>(p);
    return make_float2(v.x, v.y);
}
template <> __device__ __forceinline__ float2 load2<__half>(const __half* p) {
    const unsigned short* q = reinterpret_cast<const unsigned short*>(p);
    return make_float2(__half2float(__ushort_as_half(q[0])), __half2float(__ushort_as_half(q[1])));
}
template <> __device__ __forceinline__ float2 load2<__hip_bfloat16>(const __hip_bfloat16* p) {
    const unsigned short* q = reinterpret_cast<const unsigned short*>(p);
    return make_float2(__uint_as_float((unsigned)q[0] << 16), __uint_as_float((unsigned)q[1] << 16));
}

// Buffer loads: one descriptor per view (wavefront-uniform, in SGPRs) and a
// 32-bit byte offset per lane; the second detector row comes through the
// instruction's scalar offset.  Out-of-range buffer reads return 0 instead of
// faulting.
typedef float f32x2_v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t view_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// Samples at byte offset `off` and the next element, as floats.
template <class T> __device__ __forceinline__ float2 bload2(__amdgpu_buffer_rsrc_t r,
                                                           unsigned off, unsigned soff);
template <> __device__ __forceinline__ float2 bload2<float>(__amdgpu_buffer_rsrc_t r,
                                                           unsigned off, unsigned soff) {
    const f32x2_v v = __builtin_bit_cast(f32x2_v, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    return make_float2(v.x, v.y);
}
// 16-bit samples: one 8-byte load from the 4-byte-aligned address below `off`
// covers both elements whether `off` is 4-aligned or 2 past it.
__device__ __forceinline__ unsigned pair16(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    const unsigned long long v = __builtin_bit_cast(
        unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, off & ~3u, soff, 0));
    return (unsigned)(v >> ((off & 2u) * 8u));
}
template <> __device__ __forceinline__ float2 bload2<__half>(__amdgpu_buffer_rsrc_t r,
                                                            unsigned off, unsigned soff) {
    const unsigned b = pair16(r, off, soff);
    return make_float2(__half2float(__ushort_as_half((unsigned short)(b & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(b >> 16))));
}
template <> __device__ __forceinline__ float2 bload2<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r,
                                                                    unsigned off, unsigned soff) {
    const unsigned b = pair16(r, off, soff);
    return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xffff0000u));
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Read-only tables reached through the constant address space: loads with a
// uniform address become scalar loads even when the pointer itself came from
// memory (the batched launch's slab descriptors), where the compiler cannot
// otherwise prove that the kernel's stores leave them unchanged.
template <class P> using cptr = const __attribute__((address_space(4))) P*;
template <class P> __device__ __forceinline__ cptr<P> const_ptr(const P* p) {
    return (cptr<P>)p;
}
// Element i of a float2 / uint2 table read through the constant address space.
__device__ __forceinline__ float2 cld_f2(cptr<float> p, long long i) {
    return make_float2(p[2 * i], p[2 * i + 1]);
}
__device__ __forceinline__ uint2 cld_u2(cptr<unsigned> p, long long i) {
    return make_uint2(p[2 * i], p[2 * i + 1]);
}


template <int NZ, class T>
__global__ __launch_bounds__(256) void bp_fdk3dw(float* __restrict__ img,
                                                 const T* __restrict__ fp,
                                                 const float2* __restrict__ vtab,
                                                 const float4* __restrict__ ktab,
                                                 const BPZ* __restrict__ ztab, BPArgs A) {
    // XCD-aware remap: hardware workgroup b runs on XCD b % 8; give each XCD a
    // contiguous run of logical workgroups.
    const int nb = gridDim.x, b = blockIdx.x;
    const int qx = nb >> 3, rx = nb & 7, xcd = b & 7;
    const int lb = xcd * qx + min(xcd, rx) + (b >> 3) + A.lb_offset;
    int zs, tyb, txb;
    wg_tile(lb, A, zs, tyb, txb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int xi = txb * 16 + (wave & 1) * 8 + (lane & 7);
    const int yi = tyb * 16 + (wave >> 1) * 8 + (lane >> 3);
    const float x = ((float)xi - A.XNC) * A.dx;
    const float y = ((float)yi - A.YNC) * A.dy;
    const int z0 = zs * NZ;

    int s0[NZ], s1[NZ], s2[NZ];
    float dzs[NZ], acc[NZ];
    int pmin = INT_MAX, pmax = INT_MIN;
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        acc[j] = 0.f;
        if (z0 + j < A.ZN) {
            const BPZ t = ztab[z0 + j];
            s0[j] = t.s0; s1[j] = t.s1; s2[j] = t.s2; dzs[j] = t.dz;
        } else {
            s0[j] = 0; s1[j] = 1; s2[j] = 0; dzs[j] = 0.f;
        }
        if (s1[j] <= s2[j]) { pmin = min(pmin, s1[j]); pmax = max(pmax, s2[j]); }
    }

    const long long view_stride = (long long)A.ZL * A.YL;
    for (int p = pmin; p <= pmax; ++p) {
        const float2 cs = vtab[p];
        const float UU = -x * cs.x - y * cs.y;
        const float Yr = -x * cs.y + y * cs.x;
        const float t1sq = A.SO2 - Yr * Yr;
        const float temp1 = fast_sqrt(t1sq);
        const float U1 = Yr * A.invdYL + A.YLC;
        const bool uok = (U1 > 0.f) && (U1 <= A.YLm1);   // 0 < ceil(U1) < YL
        const float Uc = ceilf(U1);
        const float Dey = Uc - U1;
        const int U = uok ? (int)Uc : 1;
        const float K = A.KSD * fast_rcp(fmaf(UU, temp1, t1sq));
        const float R = fabsf((temp1 - UU) * fast_rcp(temp1 + UU));
        const T* col = fp + (long long)p * view_stride + (U - 1);
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
            if (p >= s1[j] && p <= s2[j]) {
                const int k = p - s0[j];
                const float4 kt = ktab[k + A.W];   // w1, w2, -+0.5, w1/(w1+w2)
                const float temp2 = fmaf(-(float)k, A.hstep, dzs[j]);
                const float V1 = fmaf(temp2, K, A.ZLC);
                if (uok && V1 > 0.f && V1 <= A.ZLm1) {       // 0 < ceil(V1) < ZL
                    const float Vc = ceilf(V1);
                    const float Dez = Vc - V1;
                    const T* q0 = col + ((int)Vc - 1) * A.YL;
                    const float2 r0 = load2<T>(q0);          // row V-1: U-1, U
                    const float2 r1 = load2<T>(q0 + A.YL);   // row V:   U-1, U
                    const float top = fmaf(Dey, r0.x - r0.y, r0.y);
                    const float bot = fmaf(Dey, r1.x - r1.y, r1.y);
                    const float tou = fmaf(Dez, top - bot, bot);
                    float weight;
                    if (kt.y == 0.f) {
                        weight = 1.f;
                    } else if (A.k1zero) {
                        weight = kt.w;
                    } else {
                        const float ratio = fabsf(temp2) * R * fast_rcp(fabsf(temp2 + kt.z));
                        const float m = fast_exp2(A.k1 * fast_log2(ratio));
                        weight = kt.x * fast_rcp(fmaf(kt.y, m, kt.x));
                    }
                    acc[j] = fmaf(weight, tou, acc[j]);
                }
            }
        }
    }
    if (xi < A.XN && yi < A.YN) {
        float* o = img + ((long long)yi * A.XN + xi) * A.ZN + z0;
        bool done = false;
        if constexpr (NZ == 4) {
            if (A.vec4 && z0 + 4 <= A.ZN) {
                *reinterpret_cast<float4*>(o) =
                    make_float4(acc[0] * A.DeltaFai, acc[1] * A.DeltaFai,
                                acc[2] * A.DeltaFai, acc[3] * A.DeltaFai);
                done = true;
            }
        }
        if (!done) {
#pragma unroll
            for (int j = 0; j < NZ; ++j)
                if (z0 + j < A.ZN) o[j] = acc[j] * A.DeltaFai;
        }
    }
}

// View-table kernel.  temp2 = z - h V / 2pi depends only on (slice, view), so the
// row offset and the slice-dependent part of the 3D weight,
//   a = 2^Q,  Q = log2(w2/w1) + k1 log2(|temp2| / |temp2 -+ 0.5|)   (a = 0 when w2 = 0),
// come from a table built in FP64 on the host, laid out per (slice set, view) so
// that one scalar load brings the NZ slices' (temp2, a) for a view.  A slice
// whose window does not contain the view has temp2 = 1e30 there (its row is then
// off the detector), so the branch-free update adds nothing and no window tests
// run in the loop.  With the column term B = ((t1-UU)/(t1+UU))^k1,
//   weight = w1 Gc^k1 / (w2 G^k1 + w1 Gc^k1) = 1 / (1 + a B),
// and the views where every slice of the set has a = 0 (no conjugate view in
// the window) run a second loop body without the weight at all.
struct ZSet {
    int pmin, pmax;     // union of the set's windows (pmin > pmax: empty)
    int clo, chi;       // views with weight 1 for every slice of the set
    long long off;      // first table entry (view pmin, slice 0)
    long long pad;
};

template <int NZ, class T, bool WEIGHTED>
__device__ __forceinline__ void bp_view(int p, float x, float y, const T* __restrict__ fp,
                                        const float2* __restrict__ vtab,
                                        const float2* __restrict__ trow, const BPArgs& A,
                                        unsigned rowb, float (&acc)[NZ]) {
    struct Ent { float2 e[NZ]; };
    const Ent ent = *reinterpret_cast<const Ent*>(trow + (long long)p * NZ);   // scalar load
    const float2 cs = vtab[p];
    const __amdgpu_buffer_rsrc_t rs = view_rsrc(fp + (long long)p * A.view_stride, A.view_bytes);
    const float UU = -x * cs.x - y * cs.y;
    const float Yr = -x * cs.y + y * cs.x;
    const float t1sq = A.SO2 - Yr * Yr;
    const float temp1 = fast_sqrt(t1sq);
    const float U1 = Yr * A.invdYL + A.YLC;
    const bool uok = (U1 > 0.f) && (U1 <= A.YLm1);           // 0 < ceil(U1) < YL
    const float Uc = ceilf(U1);
    const float Dey = Uc - U1;
    // byte offset of channel U-1 in row V-1 is (V-1)*rowb + (U-1)*sizeof(T); the
    // -rowb folds the row's -1 into the per-slice multiply-add below
    const unsigned ucolb = (uok ? (unsigned)Uc - 1u : 0u) * (unsigned)sizeof(T) - rowb;
    const float inv = fast_rcp(fmaf(UU, temp1, t1sq));     // 1 / (t1 (t1 + UU))
    const float K = A.KSD * inv;
    float B = 1.f;
    if (WEIGHTED)   // B = ((t1-UU)/(t1+UU))^k1, kept finite and nonzero
        B = fminf(fmaxf(fast_exp2(A.k1 * fast_log2(fabsf((temp1 - UU) * temp1 * inv))), 1e-30f),
                  1e30f);
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const float V1 = fmaf(ent.e[j].x, K, A.ZLC);     // huge for an inactive slice
        const bool ok = uok && V1 > 0.f && V1 <= A.ZLm1;  // 0 < ceil(V1) < ZL
        const float Vc = ceilf(V1);
        const float Dez = Vc - V1;
        // off-detector lanes read past the descriptor's range: 0, no access
        const unsigned ob = ok ? __umul24((unsigned)Vc, rowb) + ucolb : 0xfffffff0u;
        const float2 r0 = bload2<T>(rs, ob, 0u);          // row V-1: U-1, U
        const float2 r1 = bload2<T>(rs, ob, rowb);        // row V:   U-1, U
        const float top = fmaf(Dey, r0.x - r0.y, r0.y);
        const float bot = fmaf(Dey, r1.x - r1.y, r1.y);
        const float tou = ok ? fmaf(Dez, top - bot, bot) : 0.f;
        if (WEIGHTED)   // w1 Gc^k1 / (w2 G^k1 + w1 Gc^k1) = 1 / (1 + 2^Q B)
            acc[j] = fmaf(fast_rcp(fmaf(ent.e[j].y, B, 1.f)), tou, acc[j]);
        else
            acc[j] += tou;
    }
}

template <int NZ, class T>
__global__ __launch_bounds__(256) void bp_fdk3dw_vt(float* __restrict__ img,
                                                    const T* __restrict__ fp,
                                                    const float2* __restrict__ vtab,
                                                    const ZSet* __restrict__ zsets,
                                                    const float2* __restrict__ table, BPArgs A) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int qx = nb >> 3, rx = nb & 7, xcd = b & 7;
    const int lb = xcd * qx + min(xcd, rx) + (b >> 3) + A.lb_offset;
    int zs, tyb, txb;
    wg_tile(lb, A, zs, tyb, txb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int xi = txb * 16 + (wave & 1) * 8 + (lane & 7);
    const int yi = tyb * 16 + (wave >> 1) * 8 + (lane >> 3);
    const float x = ((float)xi - A.XNC) * A.dx;
    const float y = ((float)yi - A.YNC) * A.dy;
    const int z0 = zs * NZ;
    const ZSet h = zsets[zs];
    const float2* trow = table + h.off - (long long)h.pmin * NZ;
    const unsigned rowb = (unsigned)A.YL * (unsigned)sizeof(T);   // bytes per detector row
    // (V, rowb) < 2^24 for every supported detector, so 24-bit multiplies suffice
    float acc[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) acc[j] = 0.f;
    const int c0 = max(h.clo, h.pmin), c1 = min(h.chi, h.pmax);
    if (c0 <= c1) {
        for (int p = h.pmin; p < c0; ++p) bp_view<NZ, T, true>(p, x, y, fp, vtab, trow, A, rowb, acc);
        for (int p = c0; p <= c1; ++p) bp_view<NZ, T, false>(p, x, y, fp, vtab, trow, A, rowb, acc);
        for (int p = c1 + 1; p <= h.pmax; ++p)
            bp_view<NZ, T, true>(p, x, y, fp, vtab, trow, A, rowb, acc);
    } else {
        for (int p = h.pmin; p <= h.pmax; ++p)
            bp_view<NZ, T, true>(p, x, y, fp, vtab, trow, A, rowb, acc);
    }
    if (xi < A.XN && yi < A.YN) {
        float* o = img + ((long long)yi * A.XN + xi) * A.ZN + z0;
        if (NZ % 4 == 0 && A.vec4 && z0 + NZ <= A.ZN) {
#pragma unroll
            for (int j = 0; j < NZ; j += 4)
                *reinterpret_cast<float4*>(o + j) =
                    make_float4(acc[j] * A.DeltaFai, acc[(j + 1) % NZ] * A.DeltaFai,
                                acc[(j + 2) % NZ] * A.DeltaFai, acc[(j + 3) % NZ] * A.DeltaFai);
        } else {
#pragma unroll
            for (int j = 0; j < NZ; ++j)
                if (z0 + j < A.ZN) o[j] = acc[j] * A.DeltaFai;
        }
    }
}

// Z-column kernel.  Each thread owns a block of ZB consecutive slices of its
// column and walks the union of their windows once.  Consecutive slices' central
// views are s0(z+dz) - s0(z) = N_2pi dz / h apart, so one view lies in the windows
// of ~L / (N_2pi dz / h) consecutive slices: the column terms (UU, Yr, temp1, the
// channel, the magnification and B) are computed once per (column, view) and
// shared by all of them.  Per (block, view) the table holds two wavefront-uniform
// bitmasks -- the slices whose window contains the view, and those of them with
// a nonzero conjugate weight (a != 0) -- and each slice's (temp2, a) as in the
// view-table kernel.  Inactive slices are skipped with scalar branches, so every
// voxel still sees exactly its own window and no work is spent outside it.
struct ZBlock {
    int pmin, pmax;     // union of the block's windows (pmin > pmax: empty)
    long long off;      // first (temp2, a) entry of view pmin; masks at off / ZB
};

// Bilinear sample of the z-column kernel at row offset fr (= 1 - Dez) and channel
// blend weight Dey (on channel U-1), from the byte offset `ob` of sample
// (row V-1, channel U-1).  load() issues the gathers and lerp() blends them, so
// a view's gathers can all be in flight before the first blend.  Rows layout: two
// gathers, rows V-1 and V (the second through the instruction's scalar offset).
// Row-pair layout: one gather of the four samples (V-1, U-1), (V, U-1), (V-1, U),
// (V, U).
template <class T, int LAYOUT> struct Gather;

template <class T> struct Gather<T, ARECON_FP_ROWS> {
    static constexpr unsigned kElem = sizeof(T);            // bytes per channel step
    static constexpr bool kPair = false;
    struct Raw { float2 r0, r1; };
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t rs, unsigned ob,
                                               unsigned rowb) {
        return Raw{bload2<T>(rs, ob, 0u), bload2<T>(rs, ob, rowb)};   // rows V-1, V
    }
    __device__ static __forceinline__ float lerp(const Raw& q, float fr, float Dey, float) {
        const float top = fmaf(Dey, q.r0.x - q.r0.y, q.r0.y);
        const float bot = fmaf(Dey, q.r1.x - q.r1.y, q.r1.y);
        return fmaf(fr, bot - top, top);
    }
};

typedef float f32x4_v __attribute__((ext_vector_type(4)));

template <class T> struct Quad;           // four samples as loaded
template <> struct Quad<float> {
    typedef f32x4_v Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(f32x4_v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0u, 0));
    }
    __device__ static __forceinline__ float4 get(const Raw& v) {
        return make_float4(v.x, v.y, v.z, v.w);
    }
};
template <> struct Quad<__half> {
    typedef unsigned long long Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(unsigned long long,
                                  __builtin_amdgcn_raw_buffer_load_b64(r, off, 0u, 0));
    }
    __device__ static __forceinline__ float4 get(const Raw& v) {
        const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
        return make_float4(__half2float(__ushort_as_half((unsigned short)(lo & 0xffffu))),
                           __half2float(__ushort_as_half((unsigned short)(lo >> 16))),
                           __half2float(__ushort_as_half((unsigned short)(hi & 0xffffu))),
                           __half2float(__ushort_as_half((unsigned short)(hi >> 16))));
    }
};
template <> struct Quad<__hip_bfloat16> {
    typedef unsigned long long Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(unsigned long long,
                                  __builtin_amdgcn_raw_buffer_load_b64(r, off, 0u, 0));
    }
    __device__ static __forceinline__ float4 get(const Raw& v) {
        const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
        return make_float4(__uint_as_float(lo << 16), __uint_as_float(lo & 0xffff0000u),
                           __uint_as_float(hi << 16), __uint_as_float(hi & 0xffff0000u));
    }
};

// Pair layouts (row pairs, row differences): chan() gives the two channels'
// row-blended values (c0 on channel U-1, c1 on channel U); lerp() blends them,
// and an unweighted update adds them to the accumulator directly (acc_pair).
template <class T> struct Gather<T, ARECON_FP_ROWPAIRS> {
    static constexpr unsigned kElem = 2 * sizeof(T);
    static constexpr bool kPair = true;
    typedef typename Quad<T>::Raw Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t rs, unsigned ob, unsigned) {
        return Quad<T>::load(rs, ob);
    }
    __device__ static __forceinline__ float2 chan(const Raw& raw, float fr) {
        const float4 q = Quad<T>::get(raw);
        return make_float2(fmaf(fr, q.y - q.x, q.x),      // channel U-1, rows blended
                           fmaf(fr, q.w - q.z, q.z));     // channel U
    }
    __device__ static __forceinline__ float lerp(const Raw& raw, float fr, float Dey, float) {
        const float2 c = chan(raw, fr);
        return fmaf(Dey, c.x - c.y, c.y);
    }
};

// Row differences: (f, f(r+1) - f) for channels U-1 and U.  The differences are
// the ones the row-pair blend forms, in the same FP32 rounding, so the sample is
// bit-identical to it in four instructions instead of six.
template <class T> struct Gather<T, ARECON_FP_ROWDIFF> {
    static constexpr unsigned kElem = 2 * sizeof(T);
    static constexpr bool kPair = true;
    typedef typename Quad<T>::Raw Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t rs, unsigned ob, unsigned) {
        return Quad<T>::load(rs, ob);
    }
    __device__ static __forceinline__ float2 chan(const Raw& raw, float fr) {
        const float4 q = Quad<T>::get(raw);
        return make_float2(fmaf(fr, q.y, q.x),            // channel U-1, rows blended
                           fmaf(fr, q.w, q.z));           // channel U
    }
    __device__ static __forceinline__ float lerp(const Raw& raw, float fr, float Dey, float) {
        const float2 c = chan(raw, fr);
        return fmaf(Dey, c.x - c.y, c.y);
    }
};

// Unweighted update of a pair layout: acc + Dey c0 + wu c1 in two FMAs (the
// blend and the add would take three).
__device__ __forceinline__ float acc_pair(float acc, float2 c, float Dey, float wu) {
    return fmaf(wu, c.y, fmaf(Dey, c.x, acc));
}

// Difference quads: (a, dr, dc, drc) with the sample at (row V-1 + fr, channel
// U-1 + wu) equal to a + fr dr + wu (dc + fr drc), wu = 1 - Dey.
template <class T> struct Gather<T, ARECON_FP_QUADS> {
    static constexpr unsigned kElem = 4 * sizeof(T);
    static constexpr bool kPair = false;
    typedef typename Quad<T>::Raw Raw;
    __device__ static __forceinline__ Raw load(__amdgpu_buffer_rsrc_t rs, unsigned ob, unsigned) {
        return Quad<T>::load(rs, ob);
    }
    __device__ static __forceinline__ float lerp(const Raw& raw, float fr, float, float wu) {
        const float4 q = Quad<T>::get(raw);
        return fmaf(wu, fmaf(fr, q.w, q.z), fmaf(fr, q.y, q.x));
    }
};

// Row index and blend weight of one slice: vi = floor(V1) (= ceil(V1) - 1 of the
// reference for every non-integer V1; for an integer V1 both pick the same
// sample) and fr = fract(V1) = 1 - Dez.  Returns the gather offset, or one past
// any view when the row is off the detector (0 < ceil(V1) < ZL fails): with the
// row-pair and quad layouts the view's buffer ends after row ZL-2, so clamping
// vi to ZL-1 is enough; with rows, the second row must be tested explicitly.
template <int LAYOUT>
__device__ __forceinline__ unsigned zc_offset(float V1, unsigned rowb, unsigned ucolb,
                                              unsigned zlm1, float& fr) {
    // floor and convert in one instruction.  V1 comes from an FMA (no
    // transcendental-result hazard for the asm to miss).
    int vf;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(vf) : "v"(V1));
    fr = __builtin_amdgcn_fractf(V1);
    if (LAYOUT != ARECON_FP_ROWS)
        return __umul24(min((unsigned)vf, zlm1), rowb) + ucolb;
    return (unsigned)vf < zlm1 ? __umul24((unsigned)vf, rowb) + ucolb : 0x80000000u;
}

// One group of G consecutive slices of the block, branch-free: the slices of the
// group outside the view's window carry temp2 = 1e30 in the table, so their row
// is off the detector and their gathers read 0.
template <int G, class T, int LAYOUT, bool WEIGHTED>
__device__ __forceinline__ void zc_group(cptr<float> e, float K, float zlc,
                                         float B, unsigned ucolb, unsigned rowb,
                                         unsigned zlm1, __amdgpu_buffer_rsrc_t rs, float Dey,
                                         float wu, float* acc) {
    typedef Gather<T, LAYOUT> Gt;
    typename Gt::Raw q[G];
    float fr[G];
#pragma unroll
    for (int i = 0; i < G; ++i)
        q[i] = Gt::load(rs, zc_offset<LAYOUT>(fmaf(e[2 * i], K, zlc), rowb, ucolb, zlm1, fr[i]),
                        rowb);
#pragma unroll
    for (int i = 0; i < G; ++i) {
        if constexpr (WEIGHTED)   // 1 / (1 + 2^Q B); exactly 1 where a = 0
            acc[i] = fmaf(fast_rcp(fmaf(e[2 * i + 1], B, 1.f)), Gt::lerp(q[i], fr[i], Dey, wu),
                          acc[i]);
        else if constexpr (Gt::kPair)
            acc[i] = acc_pair(acc[i], Gt::chan(q[i], fr[i]), Dey, wu);
        else
            acc[i] += Gt::lerp(q[i], fr[i], Dey, wu);
    }
}

// In-place 3D-weighted accumulation of one slice's sample t:
//   acc = fma(rcp(1 + a B), t, acc)   when bit J of the wavefront-uniform mask
//                                     `wm` is set,
//   acc += t                          otherwise (no conjugate view: weight 1).
// One asm block with the uniform branch inside, and the accumulator tied to the
// destination: written as C++, the two paths give the accumulator two new
// registers and the compiler copies every accumulator back at every view.  The
// s_nop is the one wait state that a VALU read of a transcendental's result
// (v_rcp_f32) needs; the compiler does not insert it for inline asm.
template <int J>
__device__ __forceinline__ void acc_weighted(float& acc, float t, unsigned wm, float a, float B) {
    float w;
    asm volatile(
        "s_bitcmp1_b32 %[wm], %[j]\n\t"
        "s_cbranch_scc0 1f\n\t"
        "v_fma_f32 %[w], %[a], %[B], 1.0\n\t"
        "v_rcp_f32 %[w], %[w]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[acc], %[w], %[t]\n\t"
        "s_branch 2f\n"
        "1:\n\t"
        "v_add_f32 %[acc], %[acc], %[t]\n"
        "2:"
        : [acc] "+v"(acc), [w] "=&v"(w)
        : [t] "v"(t), [wm] "s"(wm), [j] "i"(J), [a] "s"(a), [B] "v"(B)
        : "scc");
}

// acc_weighted for a pair layout, from the two channel values c = (c0, c1):
// weighted, t = fma(Dey, c0 - c1, c1) (the layout's lerp) and the same weighted
// add; unweighted, acc_pair's two FMAs.  c is clobbered (it holds t).
template <int J>
__device__ __forceinline__ void acc_weighted_p(float& acc, float2 c, float Dey, float wu,
                                               unsigned wm, float a, float B) {
    float w, u = c.x;
    asm volatile(
        "s_bitcmp1_b32 %[wm], %[j]\n\t"
        "s_cbranch_scc0 1f\n\t"
        "v_sub_f32 %[u], %[u], %[v]\n\t"
        "v_fma_f32 %[w], %[a], %[B], 1.0\n\t"
        "v_fma_f32 %[u], %[Dey], %[u], %[v]\n\t"
        "v_rcp_f32 %[w], %[w]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[acc], %[w], %[u]\n\t"
        "s_branch 2f\n"
        "1:\n\t"
        "v_fmac_f32 %[acc], %[Dey], %[u]\n\t"
        "v_fmac_f32 %[acc], %[wu], %[v]\n"
        "2:"
        : [acc] "+v"(acc), [w] "=&v"(w), [u] "+v"(u)
        : [v] "v"(c.y), [Dey] "v"(Dey), [wu] "v"(wu), [wm] "s"(wm), [j] "i"(J), [a] "s"(a),
          [B] "v"(B)
        : "scc");
}

// 3D-weighted accumulation of slices G..G+3 (a group of four; samples t0..t3)
// with one uniform branch per group instead of one per slice.  `mw` is the view
// record's weighted word: bit j = slice j weighted, bit 16 + G/4 = some slice of
// the group weighted, bit 20 + G/4 = every active slice of the group weighted.
// No weighted slice: four adds.  Every active slice weighted: acc_i =
// fma(rcp(1 + a_i B), t_i, acc_i) for all four (an inactive slice has t_i = 0
// and a_i = 0, so it adds exactly zero).  Mixed: slice by slice as
// acc_weighted.  The transcendental results are read three or more
// instructions after they are written, so no wait states are needed there.
template <int G>
__device__ __forceinline__ void acc_weighted_4(float& acc0, float& acc1, float& acc2,
                                               float& acc3, float t0, float t1, float t2,
                                               float t3, unsigned mw, float a0, float a1,
                                               float a2, float a3, float B) {
    float w0, w1, w2, w3;
    asm volatile(
        "s_bitcmp1_b32 %[mw], %[bany]\n\t"
        "s_cbranch_scc0 13f\n\t"
        "s_bitcmp1_b32 %[mw], %[ball]\n\t"
        "s_cbranch_scc1 12f\n\t"
        // mixed
        "s_bitcmp1_b32 %[mw], %[b0]\n\t"
        "s_cbranch_scc0 20f\n\t"
        "v_fma_f32 %[w0], %[a0], %[B], 1.0\n\t"
        "v_rcp_f32 %[w0], %[w0]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c0], %[w0], %[t0]\n\t"
        "s_branch 21f\n"
        "20:\n\t"
        "v_add_f32 %[c0], %[c0], %[t0]\n"
        "21:\n\t"
        "s_bitcmp1_b32 %[mw], %[b1]\n\t"
        "s_cbranch_scc0 22f\n\t"
        "v_fma_f32 %[w1], %[a1], %[B], 1.0\n\t"
        "v_rcp_f32 %[w1], %[w1]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c1], %[w1], %[t1]\n\t"
        "s_branch 23f\n"
        "22:\n\t"
        "v_add_f32 %[c1], %[c1], %[t1]\n"
        "23:\n\t"
        "s_bitcmp1_b32 %[mw], %[b2]\n\t"
        "s_cbranch_scc0 24f\n\t"
        "v_fma_f32 %[w2], %[a2], %[B], 1.0\n\t"
        "v_rcp_f32 %[w2], %[w2]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c2], %[w2], %[t2]\n\t"
        "s_branch 25f\n"
        "24:\n\t"
        "v_add_f32 %[c2], %[c2], %[t2]\n"
        "25:\n\t"
        "s_bitcmp1_b32 %[mw], %[b3]\n\t"
        "s_cbranch_scc0 26f\n\t"
        "v_fma_f32 %[w3], %[a3], %[B], 1.0\n\t"
        "v_rcp_f32 %[w3], %[w3]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c3], %[w3], %[t3]\n\t"
        "s_branch 14f\n"
        "26:\n\t"
        "v_add_f32 %[c3], %[c3], %[t3]\n\t"
        "s_branch 14f\n"
        // every active slice weighted
        "12:\n\t"
        "v_fma_f32 %[w0], %[a0], %[B], 1.0\n\t"
        "v_fma_f32 %[w1], %[a1], %[B], 1.0\n\t"
        "v_fma_f32 %[w2], %[a2], %[B], 1.0\n\t"
        "v_fma_f32 %[w3], %[a3], %[B], 1.0\n\t"
        "v_rcp_f32 %[w0], %[w0]\n\t"
        "v_rcp_f32 %[w1], %[w1]\n\t"
        "v_rcp_f32 %[w2], %[w2]\n\t"
        "v_rcp_f32 %[w3], %[w3]\n\t"
        "v_fmac_f32 %[c0], %[w0], %[t0]\n\t"
        "v_fmac_f32 %[c1], %[w1], %[t1]\n\t"
        "v_fmac_f32 %[c2], %[w2], %[t2]\n\t"
        "v_fmac_f32 %[c3], %[w3], %[t3]\n\t"
        "s_branch 14f\n"
        // no weighted slice
        "13:\n\t"
        "v_add_f32 %[c0], %[c0], %[t0]\n\t"
        "v_add_f32 %[c1], %[c1], %[t1]\n\t"
        "v_add_f32 %[c2], %[c2], %[t2]\n\t"
        "v_add_f32 %[c3], %[c3], %[t3]\n"
        "14:"
        : [c0] "+v"(acc0), [c1] "+v"(acc1), [c2] "+v"(acc2), [c3] "+v"(acc3),
          [w0] "=&v"(w0), [w1] "=&v"(w1), [w2] "=&v"(w2), [w3] "=&v"(w3)
        : [t0] "v"(t0), [t1] "v"(t1), [t2] "v"(t2), [t3] "v"(t3), [mw] "s"(mw),
          [bany] "i"(16 + G / 4), [ball] "i"(20 + G / 4), [b0] "i"(G), [b1] "i"(G + 1),
          [b2] "i"(G + 2), [b3] "i"(G + 3), [a0] "s"(a0), [a1] "s"(a1), [a2] "s"(a2),
          [a3] "s"(a3), [B] "v"(B)
        : "scc");
}

// acc_weighted_4 for a pair layout, from the channel values u_i = c0, v_i = c1
// of the four slices: no weighted slice, acc_i = fma(wu, v_i, fma(Dey, u_i,
// acc_i)); every active slice weighted, t_i = fma(Dey, u_i - v_i, v_i) and
// acc_i = fma(rcp(1 + a_i B), t_i, acc_i) (an inactive slice has u_i = v_i = 0
// and a_i = 0 and adds exactly zero); mixed, slice by slice.  The u_i are
// clobbered (they hold t_i).
template <int G>
__device__ __forceinline__ void acc_weighted_4p(float& acc0, float& acc1, float& acc2,
                                                float& acc3, float2 c0, float2 c1, float2 c2,
                                                float2 c3, float Dey, float wu, unsigned mw,
                                                float a0, float a1, float a2, float a3,
                                                float B) {
    float w0, w1, w2, w3, u0 = c0.x, u1 = c1.x, u2 = c2.x, u3 = c3.x;
    asm volatile(
        "s_bitcmp1_b32 %[mw], %[bany]\n\t"
        "s_cbranch_scc0 13f\n\t"
        "s_bitcmp1_b32 %[mw], %[ball]\n\t"
        "s_cbranch_scc1 12f\n\t"
        // mixed
        "s_bitcmp1_b32 %[mw], %[b0]\n\t"
        "s_cbranch_scc0 20f\n\t"
        "v_sub_f32 %[u0], %[u0], %[v0]\n\t"
        "v_fma_f32 %[w0], %[a0], %[B], 1.0\n\t"
        "v_fma_f32 %[u0], %[Dey], %[u0], %[v0]\n\t"
        "v_rcp_f32 %[w0], %[w0]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c0], %[w0], %[u0]\n\t"
        "s_branch 21f\n"
        "20:\n\t"
        "v_fmac_f32 %[c0], %[Dey], %[u0]\n\t"
        "v_fmac_f32 %[c0], %[wu], %[v0]\n"
        "21:\n\t"
        "s_bitcmp1_b32 %[mw], %[b1]\n\t"
        "s_cbranch_scc0 22f\n\t"
        "v_sub_f32 %[u1], %[u1], %[v1]\n\t"
        "v_fma_f32 %[w1], %[a1], %[B], 1.0\n\t"
        "v_fma_f32 %[u1], %[Dey], %[u1], %[v1]\n\t"
        "v_rcp_f32 %[w1], %[w1]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c1], %[w1], %[u1]\n\t"
        "s_branch 23f\n"
        "22:\n\t"
        "v_fmac_f32 %[c1], %[Dey], %[u1]\n\t"
        "v_fmac_f32 %[c1], %[wu], %[v1]\n"
        "23:\n\t"
        "s_bitcmp1_b32 %[mw], %[b2]\n\t"
        "s_cbranch_scc0 24f\n\t"
        "v_sub_f32 %[u2], %[u2], %[v2]\n\t"
        "v_fma_f32 %[w2], %[a2], %[B], 1.0\n\t"
        "v_fma_f32 %[u2], %[Dey], %[u2], %[v2]\n\t"
        "v_rcp_f32 %[w2], %[w2]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c2], %[w2], %[u2]\n\t"
        "s_branch 25f\n"
        "24:\n\t"
        "v_fmac_f32 %[c2], %[Dey], %[u2]\n\t"
        "v_fmac_f32 %[c2], %[wu], %[v2]\n"
        "25:\n\t"
        "s_bitcmp1_b32 %[mw], %[b3]\n\t"
        "s_cbranch_scc0 26f\n\t"
        "v_sub_f32 %[u3], %[u3], %[v3]\n\t"
        "v_fma_f32 %[w3], %[a3], %[B], 1.0\n\t"
        "v_fma_f32 %[u3], %[Dey], %[u3], %[v3]\n\t"
        "v_rcp_f32 %[w3], %[w3]\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32 %[c3], %[w3], %[u3]\n\t"
        "s_branch 14f\n"
        "26:\n\t"
        "v_fmac_f32 %[c3], %[Dey], %[u3]\n\t"
        "v_fmac_f32 %[c3], %[wu], %[v3]\n"
        "s_branch 14f\n"
        // every active slice weighted
        "12:\n\t"
        "v_sub_f32 %[u0], %[u0], %[v0]\n\t"
        "v_sub_f32 %[u1], %[u1], %[v1]\n\t"
        "v_sub_f32 %[u2], %[u2], %[v2]\n\t"
        "v_sub_f32 %[u3], %[u3], %[v3]\n\t"
        "v_fma_f32 %[w0], %[a0], %[B], 1.0\n\t"
        "v_fma_f32 %[w1], %[a1], %[B], 1.0\n\t"
        "v_fma_f32 %[w2], %[a2], %[B], 1.0\n\t"
        "v_fma_f32 %[w3], %[a3], %[B], 1.0\n\t"
        "v_rcp_f32 %[w0], %[w0]\n\t"
        "v_rcp_f32 %[w1], %[w1]\n\t"
        "v_rcp_f32 %[w2], %[w2]\n\t"
        "v_rcp_f32 %[w3], %[w3]\n\t"
        "v_fma_f32 %[u0], %[Dey], %[u0], %[v0]\n\t"
        "v_fma_f32 %[u1], %[Dey], %[u1], %[v1]\n\t"
        "v_fma_f32 %[u2], %[Dey], %[u2], %[v2]\n\t"
        "v_fma_f32 %[u3], %[Dey], %[u3], %[v3]\n\t"
        "v_fmac_f32 %[c0], %[w0], %[u0]\n\t"
        "v_fmac_f32 %[c1], %[w1], %[u1]\n\t"
        "v_fmac_f32 %[c2], %[w2], %[u2]\n\t"
        "v_fmac_f32 %[c3], %[w3], %[u3]\n\t"
        "s_branch 14f\n"
        // no weighted slice
        "13:\n\t"
        "v_fmac_f32 %[c0], %[Dey], %[u0]\n\t"
        "v_fmac_f32 %[c1], %[Dey], %[u1]\n\t"
        "v_fmac_f32 %[c2], %[Dey], %[u2]\n\t"
        "v_fmac_f32 %[c3], %[Dey], %[u3]\n\t"
        "v_fmac_f32 %[c0], %[wu], %[v0]\n\t"
        "v_fmac_f32 %[c1], %[wu], %[v1]\n\t"
        "v_fmac_f32 %[c2], %[wu], %[v2]\n\t"
        "v_fmac_f32 %[c3], %[wu], %[v3]\n"
        "14:"
        : [c0] "+v"(acc0), [c1] "+v"(acc1), [c2] "+v"(acc2), [c3] "+v"(acc3),
          [w0] "=&v"(w0), [w1] "=&v"(w1), [w2] "=&v"(w2), [w3] "=&v"(w3),
          [u0] "+v"(u0), [u1] "+v"(u1), [u2] "+v"(u2), [u3] "+v"(u3)
        : [v0] "v"(c0.y), [v1] "v"(c1.y), [v2] "v"(c2.y), [v3] "v"(c3.y), [Dey] "v"(Dey),
          [wu] "v"(wu), [mw] "s"(mw), [bany] "i"(16 + G / 4), [ball] "i"(20 + G / 4),
          [b0] "i"(G), [b1] "i"(G + 1), [b2] "i"(G + 2), [b3] "i"(G + 3), [a0] "s"(a0),
          [a1] "s"(a1), [a2] "s"(a2), [a3] "s"(a3), [B] "v"(B)
        : "scc");
}

// Gives a sample register an (unspecified) value without an instruction.  The
// samples of slices outside a view's window are never read, but left undefined
// the compiler merges them with the previous view's registers and copies every
// one of them around the loop.
template <class R> __device__ __forceinline__ void undef_reg(R& r) { asm volatile("" : "=v"(r)); }

// Blend and accumulate slices J..ZB-1 of a view whose gathers are in flight:
// each active slice (bit of m.x) blends its sample and adds it with its 3D
// weight (bit of m.y: conjugate view present).
template <int J, int ZB, class Gt>
__device__ __forceinline__ void zc_blend(const typename Gt::Raw* q, const float* fr,
                                         const float2* ent, uint2 m, float B, float Dey,
                                         float wu, float* acc) {
    if constexpr (J < ZB) {
        if ((m.x >> J) & 1u) {
            if constexpr (Gt::kPair)
                acc_weighted_p<J>(acc[J], Gt::chan(q[J], fr[J]), Dey, wu, m.y, ent[J].y, B);
            else
                acc_weighted<J>(acc[J], Gt::lerp(q[J], fr[J], Dey, wu), m.y, ent[J].y, B);
        }
        zc_blend<J + 1, ZB, Gt>(q, fr, ent, m, B, Dey, wu, acc);
    }
}

// All ZB slices of the block for one view, each behind its own uniform window
// test: first every active slice's gather, then every blend.
template <int ZB, class T, int LAYOUT>
__device__ __forceinline__ void zc_split(cptr<float> e, uint2 m, float K,
                                         float zlc, float B, unsigned ucolb, unsigned rowb,
                                         unsigned zlm1, __amdgpu_buffer_rsrc_t rs, float Dey,
                                         float wu, float* acc) {
    typedef Gather<T, LAYOUT> Gt;
    float2 ent[ZB];                                              // scalar loads
#pragma unroll
    for (int j = 0; j < ZB; ++j) ent[j] = cld_f2(e, j);
    typename Gt::Raw q[ZB];
    float fr[ZB];
#pragma unroll
    for (int j = 0; j < ZB; ++j)
        if ((m.x >> j) & 1u)
            q[j] = Gt::load(rs, zc_offset<LAYOUT>(fmaf(ent[j].x, K, zlc), rowb, ucolb, zlm1,
                                                  fr[j]), rowb);
    zc_blend<0, ZB, Gt>(q, fr, ent, m, B, Dey, wu, acc);
}

// Per-(column, view) terms of the z-column kernels: the channel blend (channel
// ub = U-1 of the reference, weight Dey on it, wu = 1 - Dey on channel U), the
// row scale K = SO SD / (dZL t1 (t1 + UU)) and B = ((t1-UU)/(t1+UU))^k1 (only
// when the view has a weighted slice).  uok: 0 < ceil(U1) < YL and a finite
// magnification -- the reference's row test rejects the NaN / infinite Zr of a
// voxel outside the source circle; K = 0 off the detector.
// INSIDE: the host has shown that every column of the launch lies well inside
// the source circle (t1 + UU > 0) and projects strictly inside (1, YL-2), so the
// tests and clamps drop out; floor / fract pick the same two samples with the
// same weights as ceil(U1) - 1 and ceil(U1) - U1 (for an integer U1 both give
// the sample at U1).
struct ColTerms {
    float Dey, wu, K, B;
    int ub;
    bool uok;
};
// wmask: the view's weighted slices (B is computed only when it is nonzero);
// ylc: YLC in a VGPR, so the channel FMA takes invdYL as its one scalar operand.
template <bool INSIDE>
__device__ __forceinline__ ColTerms column_terms(float x, float y, float2 cs, const BPArgs& A,
                                                 unsigned wmask, float ylc) {
    ColTerms c;
    const float UU = -x * cs.x - y * cs.y;
    const float Yr = -x * cs.y + y * cs.x;
    const float t1sq = A.SO2 - Yr * Yr;
    const float temp1 = fast_sqrt(t1sq);
    const float U1 = fmaf(Yr, A.invdYL, ylc);
    const float den = fmaf(UU, temp1, t1sq);              // t1 (t1 + UU)
    const float inv = fast_rcp(den);
    c.B = 1.f;
    if constexpr (INSIDE) {
        int ub;   // U1 comes from an FMA: no transcendental-result hazard to miss
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(ub) : "v"(U1));
        c.ub = ub;
        c.wu = __builtin_amdgcn_fractf(U1);
        c.Dey = 1.f - c.wu;
        c.uok = true;
        c.K = A.KSD * inv;
        // B = (t1 (t1 - UU) / den)^k1 = exp2(k1 log2(num inv)) behind a uniform
        // branch on wmask (as C++ the compiler computes it for every view and
        // selects).  Unset for a view without weighted slices, which never
        // reads it.  The s_nops are the wait states after v_log / v_exp.
        const float num = fmaf(-UU, temp1, t1sq);          // t1 (t1 - UU)
        float B;
        asm("s_cmp_eq_u32 %[wm], 0\n\t"
            "s_cbranch_scc1 1f\n\t"
            "v_mul_f32 %[B], %[num], %[inv]\n\t"
            "v_log_f32 %[B], %[B]\n\t"
            "s_nop 0\n\t"
            "v_mul_f32 %[B], %[k1], %[B]\n\t"
            "v_exp_f32 %[B], %[B]\n\t"
            "s_nop 0\n"
            "1:"
            : [B] "=&v"(B)
            : [wm] "s"(wmask), [num] "v"(num), [inv] "v"(inv), [k1] "s"(A.k1)
            : "scc");
        c.B = B;
    } else {
        const bool weighted = wmask != 0u;
        c.uok = (U1 > 0.f) && (U1 <= A.YLm1) && (t1sq >= 0.f) && (den != 0.f);
        const float Uc = ceilf(U1);
        c.Dey = Uc - U1;
        c.wu = 1.f - c.Dey;
        c.ub = (int)Uc - 1;
        c.K = c.uok ? A.KSD * inv : 0.f;
        // the INSIDE path's num, so both give the same B where it is in range
        if (weighted)   // kept finite and nonzero
            c.B = fminf(fmaxf(fast_exp2(A.k1 * fast_log2(fabsf(fmaf(-UU, temp1, t1sq) * inv))),
                              1e-30f), 1e30f);
    }
    return c;
}

// XCD-aware remap: hardware workgroup b runs on XCD b % 8; give each XCD a
// contiguous run of logical workgroups.
__device__ __forceinline__ int xcd_logical_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int qx = nb >> 3, rx = nb & 7, xcd = b & 7;
    return xcd * qx + min(xcd, rx) + (b >> 3);
}

// Values the whole workgroup shares, re-marked as such: loaded from a slab
// descriptor in global memory they may reach the loop in VGPRs, and a buffer
// descriptor built from VGPRs would make the compiler wrap every gather in a
// readfirstlane loop.
template <class P> __device__ __forceinline__ P* uniform_ptr(P* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<P*>(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ long long uniform_i64(long long x) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)x);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)x >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// f(integral_constant<int, I>) for I = B .. E-1, each with I a constant
// expression (asm immediates, template arguments).
template <int B, int E, class F> __device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Body of the z-column kernel for logical workgroup lb of one plan.
template <int ZB, int G, class T, int LAYOUT, bool INSIDE>
__device__ __forceinline__ void zc_run(int lb, const BPArgs& A, float* __restrict__ img,
                                       const T* __restrict__ fp,
                                       const float2* __restrict__ vtab,
                                       const ZBlock* __restrict__ blocks,
                                       const uint2* __restrict__ masks,
                                       const float2* __restrict__ table) {
    static_assert(G == 0 ? ZB <= 16 : (ZB % G == 0 && ZB <= 32 && (G == 1 || G == 2 || G == 4)),
                  "bad block shape");
    lb += A.lb_offset;
    fp = uniform_ptr(fp);
    vtab = uniform_ptr(vtab);
    blocks = uniform_ptr(blocks);
    masks = uniform_ptr(masks);
    table = uniform_ptr(table);
    const long long view_stride = uniform_i64(A.view_stride);
    const unsigned view_bytes = __builtin_amdgcn_readfirstlane(A.view_bytes);
    int zs, tyb, txb;
    wg_tile(lb, A, zs, tyb, txb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int xi = txb * 16 + (wave & 1) * 8 + (lane & 7);
    const int yi = tyb * 16 + (wave >> 1) * 8 + (lane >> 3);
    const float x = ((float)xi - A.XNC) * A.dx;
    const float y = ((float)yi - A.YNC) * A.dy;
    const int z0 = zs * ZB;
    const cptr<int> hbp = (cptr<int>)const_ptr(blocks + zs);
    ZBlock hb;
    hb.pmin = hbp[0];
    hb.pmax = hbp[1];
    hb.off = *(cptr<long long>)(hbp + 2);
    const cptr<unsigned> mrow = (cptr<unsigned>)const_ptr(masks + (hb.off / ZB - hb.pmin));
    const cptr<float> trow = (cptr<float>)const_ptr(table + (hb.off - (long long)hb.pmin * ZB));
    const cptr<float> vt = (cptr<float>)const_ptr(vtab);
    constexpr unsigned kElem = Gather<T, LAYOUT>::kElem;
    // bytes per detector row: YL elements, or YL-1 quads (one per channel pair)
    const unsigned rowb = (unsigned)(LAYOUT == ARECON_FP_QUADS ? A.YL - 1 : A.YL) * kElem;
    const unsigned zlm1 = (unsigned)A.ZL - 1u;
    // ZLC as a per-lane value, so the per-slice fma takes the table's temp2 as
    // its one scalar operand.
    float zlc, ylc;
    asm("v_mov_b32 %0, %1" : "=v"(zlc) : "s"(A.ZLC));
    asm("v_mov_b32 %0, %1" : "=v"(ylc) : "s"(A.YLC));
    constexpr unsigned gmask = (1u << (G == 0 ? 1 : G)) - 1u;
    float acc[ZB];
#pragma unroll
    for (int j = 0; j < ZB; ++j) acc[j] = 0.f;

    for (int p = hb.pmin; p <= hb.pmax; ++p) {
        const uint2 m = cld_u2(mrow, p);               // scalar: active, weighted
        const float2 cs = cld_f2(vt, p);
        const __amdgpu_buffer_rsrc_t rs = view_rsrc(fp + (long long)p * view_stride, view_bytes);
        const ColTerms ct = column_terms<INSIDE>(x, y, cs, A, m.y, ylc);
        const float Dey = ct.Dey, wu = ct.wu, K = ct.K, B = ct.B;
        // byte offset of channel U-1; off the detector, past any view (< 2 GiB) so
        // both gathers read 0 whatever the row
        const unsigned ucolb = ct.uok ? (unsigned)ct.ub * kElem : 0x80000000u;
        const cptr<float> erow = trow + 2LL * p * ZB;
        if constexpr (G == 0) {
            zc_split<ZB, T, LAYOUT>(erow, m, K, zlc, B, ucolb, rowb, zlm1, rs, Dey, wu, acc);
            continue;
        }
#pragma unroll
        for (int g = 0; g < (G == 0 ? 0 : ZB); g += (G == 0 ? 1 : G)) {
            if (((m.x >> g) & gmask) == 0u) continue;
            constexpr int GG = G == 0 ? 1 : G;
            if ((m.y >> g) & gmask)
                zc_group<GG, T, LAYOUT, true>(erow + 2 * g, K, zlc, B, ucolb, rowb, zlm1, rs, Dey,
                                              wu, acc + g);
            else
                zc_group<GG, T, LAYOUT, false>(erow + 2 * g, K, zlc, B, ucolb, rowb, zlm1, rs, Dey,
                                               wu, acc + g);
        }
    }
    if (xi < A.XN && yi < A.YN) {
        float* o = img + ((long long)yi * A.XN + xi) * A.ZN + z0;
        if (A.vec4 && z0 + ZB <= A.ZN) {
#pragma unroll
            for (int j = 0; j < ZB; j += 4)
                *reinterpret_cast<float4*>(o + j) =
                    make_float4(acc[j] * A.DeltaFai, acc[j + 1] * A.DeltaFai,
                                acc[j + 2] * A.DeltaFai, acc[j + 3] * A.DeltaFai);
        } else {
#pragma unroll
            for (int j = 0; j < ZB; ++j)
                if (z0 + j < A.ZN) o[j] = acc[j] * A.DeltaFai;
        }
    }
}

template <int ZB, int G, class T, int LAYOUT, bool INSIDE>
__global__ __launch_bounds__(256) void bp_fdk3dw_zc(float* __restrict__ img,
                                                    const T* __restrict__ fp,
                                                    const float2* __restrict__ vtab,
                                                    const ZBlock* __restrict__ blocks,
                                                    const uint2* __restrict__ masks,
                                                    const float2* __restrict__ table, BPArgs A) {
    zc_run<ZB, G, T, LAYOUT, INSIDE>(xcd_logical_block(), A, img, fp, vtab, blocks, masks, table);
}

// ---------------------------------------------------------------------------
// LDS-staged z-column kernel.  The z-column kernel issues one 64-lane gather per
// voxel-update, and the texture addresser takes 4 lane addresses per clock, so
// its gathers alone cost 16 clocks per wavefront-update.  Here the workgroup's
// detector footprint of each view is copied once into LDS with coalesced
// direct-to-LDS buffer loads (64 lanes x 16 B per instruction, no VGPRs), and
// every voxel-update reads LDS instead.  The arithmetic is that of the
// z-column kernel step for step, so the result is bit-identical.
//
// Footprint of view p for a 16x16-column tile and the block's active slices
// (bounds from the tile's corners, with margins for FP32 rounding):
//  * channels: Yr = -x sinV + y cosV is affine in (x, y), so its range over the
//    tile is centre +- (7.5 dx |sinV| + 7.5 dy |cosV|); CW = 32 records
//    (channels) per staged row starting at c0 cover it with a margin.
//  * rows: V1 = temp2 K + ZLC with K = SO SD / (dZL t1 (t1 + UU)),
//    t1 = sqrt(SO^2 - Yr^2).  t1 (t1 + UU) grows with UU and with t1 (when
//    2 t1 + UU > 0), so the box of (|Yr|, UU) over the tile bounds K, and the
//    first and last active slices' temp2 bound temp2 (affine in the slice).
//    Record rows [rA, rB] are staged, clipped to [-1, ZL-1]; rows -1 and ZL-1
//    are off the detector and load as zeros (out-of-range buffer reads).
//  * a workgroup with a view whose footprint does not fit (rows > RMAX,
//    channels > CW, or a tile reaching the source circle) runs the z-column
//    kernel's body, reading its samples straight from global memory; a view
//    whose rows all lie off the detector is skipped (every sample it would
//    read is zero).  The footprints come from a table the plan builds once
//    (bp_zs_table).
// Records are 8 bytes (FP32 row pairs or row differences, 16-bit difference
// quads; a staged row of CW records is 256 B) or 4 bytes (16-bit row
// differences; 128 B rows, half the LDS per view, so more workgroups per CU).
// With 8-byte records each wave stages one view of a chunk of four, then every
// wave consumes the four views; with 4-byte records the chunks hold two views
// and the next one is staged while this one is read.  A view's scalar operands come from one per-(block, view) record
// of the plan (slice masks with per-group class bits, cos / sin, the block's ZB
// (temp2, a)), loaded whole behind one wait; its footprint entry also carries
// the LDS row term and which groups of slices need their rows clamped.  Split
// over the separate mask, view and slice tables, the same operands took three
// dependent load waits and ~10 scalar address instructions per view.
// ---------------------------------------------------------------------------
// Records per staged row: 28 for 8-byte records (224-byte rows, so that 25-row
// slots of four views fit seven workgroups per CU), 32 for 4-byte records.  A
// 16x16-column tile spans at most 16 sqrt(2) dx / channel pitch channels (21.8
// on the 512^2 clinical grid) plus the blend's second channel and a channel
// of margin on each side.
__host__ __device__ constexpr int zs_cw(int rec_bytes) { return rec_bytes == 8 ? 28 : 32; }
// Rows staged for a footprint of rows [rA, rB]: rB - rA + 1 rounded up to an even
// count for 8-byte records (16-byte loads of four rows, the last possibly of
// two: ~1.5 fewer rows staged per view, FP32 backprojection 0.2-0.7% faster on
// two boxes), to a multiple of four for 4-byte records (loads of two rows; with
// even counts more 16-slice FP16 footprints fit the slots, yet the kernel ran
// 0.7% slower).
__host__ __device__ constexpr int zs_rows(int rA, int rB, int rec_bytes) {
    return rec_bytes == 8 ? 2 * ((rB - rA + 2) / 2) : 4 * ((rB - rA + 4) / 4);
}
enum { ZS_SKIP = 0, ZS_LDS = 1, ZS_DIRECT = 2 };

typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(3))) void lds_void;

// Sample record(s) of one update at LDS address `a`.
template <class T, int LAYOUT> struct LdsGather;
template <> struct LdsGather<float, ARECON_FP_ROWPAIRS> {   // records c, c+1 (16 B)
    typedef Gather<float, ARECON_FP_ROWPAIRS>::Raw Raw;
    __device__ static __forceinline__ Raw load(const lds_char* a) {
        typedef const __attribute__((address_space(3))) f32x2_v lds_f2;
        const f32x2_v lo = *reinterpret_cast<lds_f2*>(a);
        const f32x2_v hi = *reinterpret_cast<lds_f2*>(a + 8);
        return Raw{lo.x, lo.y, hi.x, hi.y};
    }
};
template <> struct LdsGather<float, ARECON_FP_ROWDIFF> : LdsGather<float, ARECON_FP_ROWPAIRS> {};
template <class T> struct LdsGather<T, ARECON_FP_ROWDIFF> {   // 16-bit: records c, c+1 (8 B)
    typedef typename Gather<T, ARECON_FP_ROWDIFF>::Raw Raw;
    __device__ static __forceinline__ Raw load(const lds_char* a) {
        // 4-byte records: the pair is only 4-byte aligned, so two 4-byte reads
        typedef const __attribute__((address_space(3))) unsigned lds_u32;
        const unsigned lo = *reinterpret_cast<lds_u32*>(a);
        const unsigned hi = *reinterpret_cast<lds_u32*>(a + 4);
        return ((unsigned long long)hi << 32) | lo;
    }
};
template <class T> struct LdsGather<T, ARECON_FP_QUADS> {         // one quad (8 B)
    typedef typename Gather<T, ARECON_FP_QUADS>::Raw Raw;
    __device__ static __forceinline__ Raw load(const lds_char* a) {
        return *reinterpret_cast<const __attribute__((address_space(3))) unsigned long long*>(a);
    }
};

struct ZsView {      // footprint of one staged view (LDS)
    int mode, rA, rB, c0;
};

// Footprint of view p for the tile centred at (xc, yc) and the active slices m
// of the block (temp2 of slice j at erow[2 j]).  nrec: records per detector row
// of the layout.
//
// PAIRS (8-byte records): staging moves records in 16-byte pairs, and a pair
// whose first byte lies outside the view's buffer loads as zeros as a whole.  A
// pair that straddles the start (records -1 and 0) or the end of the buffer
// would therefore zero the view's first or last record.  An update inside the
// detector (INSIDE: channels in (1, YL-2)) never reads either; otherwise c0 is
// made even (row 0's pairs then start at even records and never straddle record
// 0), and a footprint whose pairs would straddle the last record runs the
// gathering body.  4-byte records are staged one per lane and need neither.
template <int RMAX, bool INSIDE, bool PAIRS>
__device__ __forceinline__ ZsView zs_footprint(const BPArgs& A, float2 cs, unsigned m,
                                               cptr<float> erow, float xc, float yc,
                                               float ex, float ey, int nrec) {
    ZsView v{ZS_SKIP, 0, -1, 0};
    if (m == 0u) return v;
    v.mode = ZS_DIRECT;
    const float Yrc = -xc * cs.y + yc * cs.x;
    const float UUc = -xc * cs.x - yc * cs.y;
    const float eY = ex * fabsf(cs.y) + ey * fabsf(cs.x);
    const float eU = ex * fabsf(cs.x) + ey * fabsf(cs.y);
    const float ua = (Yrc - eY) * A.invdYL + A.YLC, ub = (Yrc + eY) * A.invdYL + A.YLC;
    // a lane's channel pair (floor(U1), floor(U1) + 1) lies in [c0, ceil(ub) + 1]
    // with a margin of kM - 1 channels on each side
    constexpr int kCW = zs_cw(PAIRS ? 8 : 4), kM = kCW == 32 ? 2 : 1;
    int c0 = (int)floorf(fminf(ua, ub)) - kM;
    if constexpr (!INSIDE && PAIRS) c0 &= ~1;           // round down to even
    if ((int)ceilf(fmaxf(ua, ub)) + kM > c0 + kCW - 1) return v;
    const float ya = fmaxf(fabsf(Yrc) - eY, 0.f), yb = fabsf(Yrc) + eY;
    const float t1lo2 = A.SO2 - yb * yb;
    if (!(t1lo2 > 0.f)) return v;
    const float t1lo = sqrtf(t1lo2), t1hi = sqrtf(A.SO2 - ya * ya);
    const float uulo = UUc - eU, uuhi = UUc + eU;
    if (!(t1lo + uulo > 1e-3f * t1lo)) return v;       // den > 0 and 2 t1 + UU > 0
    const float Klo = A.KSD / (t1hi * (t1hi + uuhi)) * (1.f - 1e-5f);
    const float Khi = A.KSD / (t1lo * (t1lo + uulo)) * (1.f + 1e-5f);
    const int jlo = __builtin_ctz(m), jhi = 31 - __builtin_clz(m);
    const float ta = erow[2 * jlo], tb = erow[2 * jhi];
    const float vlo = fminf(fminf(ta * Klo, ta * Khi), fminf(tb * Klo, tb * Khi)) + A.ZLC;
    const float vhi = fmaxf(fmaxf(ta * Klo, ta * Khi), fmaxf(tb * Klo, tb * Khi)) + A.ZLC;
    if (!(vlo > -1e6f && vhi < 1e6f)) return v;
    const int zlm1 = A.ZL - 1;
    const int rA = max((int)floorf(vlo) - 1, -1);
    const int rB = min((int)floorf(vhi) + 1, zlm1);
    if (rB <= -1 || rA >= zlm1) { v.mode = ZS_SKIP; return v; }   // only zero rows
    // rows [rA, rB] staged in pairs (8-byte records) or fours (4-byte records),
    // fewer than RMAX: the slot's last row stays zero
    const int nl = zs_rows(rA, rB, PAIRS ? 8 : 4);
    if (nl >= RMAX) return v;
    if constexpr (!INSIDE && PAIRS) {
        // last record of the view: channel nrec-1 of row ZL-2, staged as the first
        // of a pair (an even distance from c0) when both lie in the footprint
        const int last = zlm1 - 1;
        if (rB + 1 - nl <= last && last <= rB && nrec - 1 <= c0 + kCW - 1 &&
            ((nrec - 1 - c0) & 1) == 0)
            return v;
    }
    v.mode = ZS_LDS; v.rA = rA; v.rB = rB; v.c0 = c0;
    return v;
}

// Footprint table entry, two words.  The first: c0 (16 bits), rA + 1 and rB + 1
// (8 bits each); a view with rB < rA is skipped.  The second: 16 cterm + clamp
// bits, cterm = -c0 kRec + (RMAX - 2 - rB) kRowB the LDS offset of record c0 of
// row 0 when row rB is staged at slot row RMAX - 2, and bit k set when group k
// of four slices needs its rows clamped.  Rows need no clamp in a group whose
// four slices are all active when the view's row range was not cut at the
// detector's first or last row and every lane is on the detector (INSIDE): the
// footprint then bounds every active slice's row with a row of margin on each
// side.
template <int ZB, int RMAX, bool INSIDE, bool PAIRS>
__device__ __forceinline__ uint2 zs_pack(const ZsView& v, unsigned act, int zl) {
    if (v.mode != ZS_LDS) return make_uint2(1u << 16, 0u);   // rA = 0, rB = -1: skip
    constexpr int kRec = PAIRS ? 8 : 4, kRowB = zs_cw(kRec) * kRec;
    const int cterm = -v.c0 * kRec + (RMAX - 2 - v.rB) * kRowB;
    const bool clipped = !INSIDE || v.rA < 0 || v.rB > zl - 2;
    unsigned cl = 0u;
    for (int k = 0; k < ZB / 4; ++k)
        if (clipped || ((act >> (4 * k)) & 0xfu) != 0xfu) cl |= 1u << k;
    return make_uint2(((unsigned)v.c0 & 0xffffu) | ((unsigned)(v.rA + 1) << 16) |
                          ((unsigned)(v.rB + 1) << 24),
                      ((unsigned)cterm << 4) | cl);
}

// Views of slice block zs in the footprint table: pmax - pmin + 1 rounded up to a
// multiple of four, so every (block, tile) row starts 16-byte aligned.
__host__ __device__ __forceinline__ int zs_table_views(int pmin, int pmax) {
    return (pmax - pmin + 4) & ~3;
}

// The footprint table of the LDS-staged kernel, built once per plan and record
// layout: workgroup (zs, tile) writes one entry per view of block zs (zs_pack;
// padding entries are skipped views) and whether every footprint fits the
// slots.  nrec: records per detector row; PAIRS: 8-byte records staged in
// 16-byte pairs.
template <int ZB, int RMAX, bool INSIDE, bool PAIRS>
__global__ __launch_bounds__(256) void bp_zs_table(BPArgs A, const float2* __restrict__ vtab,
                                                   const ZBlock* __restrict__ blocks,
                                                   const uint2* __restrict__ masks,
                                                   const float2* __restrict__ table,
                                                   uint2* __restrict__ tab,
                                                   const long long* __restrict__ boff,
                                                   int* __restrict__ fits, int nrec) {
    const int b = blockIdx.x;
    const int zs = b / A.ntiles, tile = b - zs * A.ntiles;
    const int tyb = tile / A.tilesx, txb = tile - tyb * A.tilesx;
    const ZBlock hb = blocks[zs];
    const int pmin = hb.pmin, pmax = hb.pmax;
    const unsigned* mrow = reinterpret_cast<const unsigned*>(masks + (hb.off / ZB - pmin));
    const float* trow = reinterpret_cast<const float*>(table + (hb.off - (long long)pmin * ZB));
    const float xc = ((float)(txb * 16) + 7.5f - A.XNC) * A.dx;
    const float yc = ((float)(tyb * 16) + 7.5f - A.YNC) * A.dy;
    const float ex = 7.5f * fabsf(A.dx), ey = 7.5f * fabsf(A.dy);
    const int nvp = zs_table_views(pmin, pmax);
    uint2* out = tab + boff[zs] + (long long)tile * nvp;
    int ok = 1;
    for (int i = threadIdx.x; i < nvp; i += 256) {
        const int p = pmin + i;
        uint2 e = make_uint2(1u << 16, 0u);                    // a skipped view
        if (p <= pmax) {
            const unsigned act = mrow[2LL * p];
            const ZsView v = zs_footprint<RMAX, INSIDE, PAIRS>(
                A, vtab[p], act, const_ptr(trow + 2LL * p * ZB), xc, yc, ex, ey, nrec);
            if (v.mode == ZS_DIRECT) ok = 0;
            e = zs_pack<ZB, RMAX, INSIDE, PAIRS>(v, act, A.ZL);
        }
        out[i] = e;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) fits[b] = ok;
}

template <int ZB, class T, int LAYOUT, int RMAX, bool INSIDE>
__device__ __forceinline__ void zs_run(int lb, const BPArgs& A, float* __restrict__ img,
                                       const T* __restrict__ fp,
                                       const float2* __restrict__ vtab,
                                       const ZBlock* __restrict__ blocks,
                                       const uint2* __restrict__ masks,
                                       const float2* __restrict__ table) {
    constexpr int kRec = (int)Gather<T, LAYOUT>::kElem;   // bytes per record
    static_assert(kRec == 8 || kRec == 4, "LDS staging takes 8- or 4-byte records");
    static_assert(ZB <= 16 && RMAX >= 5, "bad block shape");
    constexpr int kCW = zs_cw(kRec);
    constexpr int kRowB = kCW * kRec;                      // bytes per staged row
    constexpr int kRecShift = kRec == 8 ? 3 : 2;
    constexpr int kRowShift = kRowB == 256 ? 8 : 7;       // used when kRowB is a power of 2
    static_assert(kRec == 1 << kRecShift, "record shift");
    constexpr int kSlot = RMAX * kRowB;
    __shared__ __attribute__((aligned(16))) char stage0[4 * kSlot];
    typedef Gather<T, LAYOUT> Gt;
    typedef LdsGather<T, LAYOUT> Lg;

    const int lb0 = lb;
    lb += A.lb_offset;
    int zs, tyb, txb;
    wg_tile(lb, A, zs, tyb, txb);
    const cptr<int> hbp = (cptr<int>)const_ptr(uniform_ptr(blocks) + zs);
    const int pmin = hbp[0], pmax = hbp[1];
    const long long hoff = *(cptr<long long>)(hbp + 2);
    // the block's view records (masks, cos / sin, ZB entries), from view pmin
    constexpr int kRecW = 4 + 2 * ZB;
    const cptr<unsigned> vrec =
        const_ptr(uniform_ptr(A.zs_rec) + (hoff / ZB) * kRecW);

    // Every view's footprint comes from the plan's table (bp_zs_table, built once
    // per plan and layout; kept in LDS it cost the sixth or seventh workgroup per
    // CU, and recomputed here about two VALU per update).  A workgroup with a
    // view that does not fit runs the z-column kernel's body instead.
    const int tile = tyb * A.tilesx + txb;
    // the tile's footprint entries (two words per view), from view pmin
    const cptr<unsigned> ftab = (cptr<unsigned>)const_ptr(
        uniform_ptr(A.zs_tab) + *(cptr<long long>)const_ptr(uniform_ptr(A.zs_boff) + zs) +
        (long long)tile * zs_table_views(pmin, pmax));
    const int staged = *(cptr<int>)const_ptr(uniform_ptr(A.zs_fits) + (zs * A.ntiles + tile));
    if (threadIdx.x == 0 && A.stats) atomicAdd(A.stats + (staged ? 0 : 1), 1ull);
    if (!staged) {
        zc_run<ZB, 0, T, LAYOUT, INSIDE>(lb0, A, img, fp, vtab, blocks, masks, table);
        return;
    }
    // the last row of every slot is zero and stays zero (a view fits only when
    // its loaded rows (zs_rows) are fewer than RMAX): rows past the
    // view's last staged row (slices outside the view's window, temp2 = 1e30,
    // and rows off the detector's end) clamp onto it and add exactly zero, and
    // lanes off the detector read it.  The first chunk's barrier orders these
    // writes before any read.
    for (int i = threadIdx.x; i < 4 * kRowB / 4; i += 256) {
        const int slot = i / (kRowB / 4), w = i - slot * (kRowB / 4);
        reinterpret_cast<__attribute__((address_space(3))) unsigned*>(
            (lds_char*)stage0 + slot * kSlot + (RMAX - 1) * kRowB)[w] = 0u;
    }

    fp = uniform_ptr(fp);
    const long long view_stride = uniform_i64(A.view_stride);
    const unsigned view_bytes = __builtin_amdgcn_readfirstlane(A.view_bytes);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xi = txb * 16 + (wave & 1) * 8 + (lane & 7);
    const int yi = tyb * 16 + (wave >> 1) * 8 + (lane >> 3);
    const float x = ((float)xi - A.XNC) * A.dx;
    const float y = ((float)yi - A.YNC) * A.dy;
    const int z0 = zs * ZB;
    constexpr unsigned kElem = Gt::kElem;
    const int rowb = (LAYOUT == ARECON_FP_QUADS ? A.YL - 1 : A.YL) * (int)kElem;
    lds_char* const lstage = (lds_char*)stage0;
    float zlc, zoff, ylc;
    asm("v_mov_b32 %0, %1" : "=v"(zlc) : "s"(A.ZLC));
    asm("v_mov_b32 %0, %1" : "=v"(ylc) : "s"(A.YLC));
    // a lane off the detector gets row -1e9, clamped to rA, and reads the slot's
    // zero row
    asm("v_mov_b32 %0, %1" : "=v"(zoff) : "s"(-1e9f));
    float acc[ZB];
#pragma unroll
    for (int j = 0; j < ZB; ++j) acc[j] = 0.f;

    // a staging lane's row and record pair within a four-row load of 8-byte
    // records
    const int srow = lane / (kCW / 2), spair = lane - srow * (kCW / 2);
    // Chunks.  8-byte records: four views in the four slots, staged then read
    // (seven workgroups per CU cover each chunk's copies).  4-byte records: two
    // views per chunk, double-buffered in the same slots (chunk buffer b in
    // slots 2b, 2b + 1), the next chunk's copies in flight while this one is
    // read: with six workgroups per CU this gains 2.8% (16-slice FP16 blocks;
    // the same form runs FP32 no faster).
    constexpr bool kPrefetch = kRec == 4;
    constexpr int kViews = kPrefetch ? 2 : 4;        // views per chunk
    // stage the chunk of views p0 .. p0 + kViews - 1 into buffer buf (their
    // table entries e0 .. e3, the first kViews used): one wave per view
    // (four-view chunks), or two waves per view alternating its loads (two-view
    // chunks).  The entry is selected, not indexed: a local array indexed by
    // the wave would live in scratch memory.
    auto stage_chunk = [&](int p0, int buf, unsigned e0, unsigned e1v, unsigned e2,
                           unsigned e3) {
        const int v = kPrefetch ? wave >> 1 : wave, half = kPrefetch ? wave & 1 : 0;
        constexpr int kStep = kPrefetch ? 2 : 1;
        const int p = p0 + v;
        if (p <= pmax) {
            const unsigned e = v == 0 ? e0 : v == 1 ? e1v : v == 2 ? e2 : e3;
            const int rA = (int)((e >> 16) & 0xffu) - 1, rB = (int)(e >> 24) - 1;
            const int c0 = (int)(short)(e & 0xffffu);
            if (rB >= rA) {
                const __amdgpu_buffer_rsrc_t rs =
                    view_rsrc(fp + (long long)p * view_stride, view_bytes);
                // nl rows [rB + 1 - nl, rB] (zs_rows) go to slot rows
                // [RMAX - 1 - nl, RMAX - 2], so row rB + 1 lands on the slot's zero
                // last row.  Rows before the view's first or after its last record
                // (-1, ZL-1 and beyond) load as zeros.
                const int nl = zs_rows(rA, rB, kRec);
                lds_char* dst =
                    lstage + (kViews * buf + v) * kSlot + (RMAX - 1 - nl) * kRowB;
                if constexpr (kRec == 8) {
                    // 16-byte loads, four rows each (4 kCW / 2 lanes: the rest
                    // idle; in a last load of two rows, the lanes of rows 2 and 3
                    // too, so nothing lands on the zero row or past the slot).
                    // Lane: row rB + 1 - nl + i + srow, records c0 + 2 spair + {0, 1}
                    const int off = (rB + 1 - nl + srow) * rowb + (c0 + 2 * spair) * kRec;
                    if (lane < 2 * kCW)
                        for (int i = 4 * half; i < nl; i += 4 * kStep)
                            if (srow < nl - i)
                                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                    rs, (lds_void*)(dst + i * kRowB), 16,
                                    (unsigned)(off + i * rowb), 0, 0, 0);
                } else {
                    // 4-byte loads (one record per lane), two rows each.  Lane: row
                    // rB + 1 - nl + i + lane / 32, record c0 + lane % 32
                    static_assert(kCW == 32, "one 64-lane load: two rows of 4-byte records");
                    const int off = (rB + 1 - nl + (lane >> 5)) * rowb + (c0 + (lane & 31)) * kRec;
                    for (int i = 2 * half; i < nl; i += 2 * kStep)
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(
                            rs, (lds_void*)(dst + i * kRowB), 4, (unsigned)(off + i * rowb), 0,
                            0, 0);
                }
            }
        }
    };
    // consume the chunk of views p0 .. p0 + kViews - 1 from buffer buf (table
    // entries e8: two words per view), each view's record at a constant offset
    // from the chunk's
    typedef unsigned u32x4_v __attribute__((ext_vector_type(4)));
    typedef unsigned u32x16_v __attribute__((ext_vector_type(16)));
    typedef unsigned u32x8_v __attribute__((ext_vector_type(8)));
    auto consume_chunk = [&](int p0, int buf, const u32x8_v& e8) {
        const cptr<unsigned> rc = vrec + (p0 - pmin) * kRecW;
#pragma unroll
        for (int s = 0; s < kViews; ++s) {
            const int p = p0 + s;
            if (p > pmax) break;
            // the view's whole record in flight at once, and one wait for it
            // (and the chunk's table entries): left to the compiler, each part
            // is loaded where it is first used, behind a wait of its own
            const cptr<unsigned> r = rc + s * kRecW;
            const u32x4_v hd = *(cptr<u32x4_v>)r;       // active, weighted, cos, sin
            u32x16_v er[ZB / 8];                        // (temp2, a) of the ZB slices
#pragma unroll
            for (int k = 0; k < ZB / 8; ++k) er[k] = *(cptr<u32x16_v>)(r + 4 + 16 * k);
            asm volatile("" ::"s"(hd));
#pragma unroll
            for (int k = 0; k < ZB / 8; ++k) asm volatile("" ::"s"(er[k]));
            // table entry: (c0, rA + 1, rB + 1), and e1 = cterm * 16 + clamp bits
            const unsigned e = e8[2 * s], e1 = e8[2 * s + 1];
            const int rA = (int)((e >> 16) & 0xffu) - 1, rB = (int)(e >> 24) - 1;
            if (rB < rA) continue;                      // every sample would be zero
            const uint2 m = make_uint2(hd.x, hd.y);
            const float2 cs = make_float2(__uint_as_float(hd.z), __uint_as_float(hd.w));
            const ColTerms ct = column_terms<INSIDE>(x, y, cs, A, m.y, ylc);
            const float Dey = ct.Dey, wu = ct.wu, K = ct.K, B = ct.B;
            const bool uok = ct.uok;
            float2 ent[ZB];
#pragma unroll
            for (int j = 0; j < ZB; ++j)
                ent[j] = make_float2(__uint_as_float(er[j / 8][2 * (j % 8)]),
                                     __uint_as_float(er[j / 8][2 * (j % 8) + 1]));
            // byte offset in the slot = row * kRowB + colterm, row clamped to
            // [rA, rBv]; a lane off the detector gets row -1e9 (clamped to rA) and
            // the slot's zero row
            const float zl = uok ? zlc : zoff;
            // slot s's offset s * kSlot goes in the loads' immediate offset; the
            // rest of the address is one per-view VGPR, kept whole (opaque to the
            // compiler, which otherwise splits its scalar part off and adds it
            // back at every update), so an update's address is one v_lshl_add_u32
            // -c0 kRec + (RMAX - 2 - rB) kRowB: row rB at RMAX - 2 of the buffer's slot
            const int cterm = ((int)e1 >> 4) + kViews * buf * kSlot;
            int colterm;
            if constexpr (INSIDE) {
                asm("v_lshl_add_u32 %0, %1, %2, %3"
                    : "=v"(colterm) : "v"(ct.ub), "i"(kRecShift), "s"(cterm));
            } else {
                // the slot's zero row (row rBv): r clamps to rA off the detector
                const int zterm = (RMAX - 1 - rA) * kRowB + kViews * buf * kSlot;
                const int c = uok ? ct.ub * kRec + cterm : zterm;
                asm("v_mov_b32 %0, %1" : "=v"(colterm) : "v"(c));
            }
            int rBv;   // clamp bound: row rB + 1, the slot's last row, kept zero
            asm("v_mov_b32 %0, %1" : "=v"(rBv) : "s"(rB + 1));
            // groups of four slices, each skipped when none of its slices has the
            // view in its window (bit 16 + g/4 of the record's active word), and
            // branch-free inside: a slice outside the window has temp2 = 1e30, so
            // its row clamps onto the zero row (and a lane off the detector reads
            // the zero record): it adds exactly zero
            static_for<0, ZB / 4>([&](auto gi) {
                constexpr int g = 4 * decltype(gi)::value;
                if (!((m.x >> (16 + g / 4)) & 1u)) return;
                float V1[4], fr[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    V1[i] = fmaf(ent[g + i].x, K, zl);
                    fr[i] = __builtin_amdgcn_fractf(V1[i]);
                }
                // rows (floor, clamped to [rA, rBv] unless bit g/4 of the table
                // entry says the group needs no clamp; always clamped off the
                // INSIDE path) and LDS offsets of the four slices in one block: as
                // separate blocks the clamp's two paths would give every
                // accumulator a copy at the join.  row * kRowB + colterm: a
                // shift-add for power-of-two rows, else a signed 24-bit
                // multiply-add (the row may be -1)
                int a[4];
#define ZS_ROWS_                                                                       \
    "v_cvt_flr_i32_f32 %[a0], %[v0]\n\t"                                                \
    "v_cvt_flr_i32_f32 %[a1], %[v1]\n\t"                                                \
    "v_cvt_flr_i32_f32 %[a2], %[v2]\n\t"                                                \
    "v_cvt_flr_i32_f32 %[a3], %[v3]\n\t"
#define ZS_CLAMP_                                                                      \
    "v_med3_i32 %[a0], %[a0], %[rA], %[rB]\n\t"                                         \
    "v_med3_i32 %[a1], %[a1], %[rA], %[rB]\n\t"                                         \
    "v_med3_i32 %[a2], %[a2], %[rA], %[rB]\n\t"                                         \
    "v_med3_i32 %[a3], %[a3], %[rA], %[rB]\n"
#define ZS_SKIP_ "s_bitcmp1_b32 %[cl], %[cb]\n\ts_cbranch_scc0 1f\n\t"
#define ZS_ADDR_(OP)                                                                   \
    "1:\n\t" OP " %[a0], %[a0], %[rw], %[ct]\n\t" OP " %[a1], %[a1], %[rw], %[ct]\n\t" OP  \
    " %[a2], %[a2], %[rw], %[ct]\n\t" OP " %[a3], %[a3], %[rw], %[ct]"
#define ZS_OPS_(RW)                                                                    \
    : [a0] "=&v"(a[0]), [a1] "=&v"(a[1]), [a2] "=&v"(a[2]), [a3] "=&v"(a[3])             \
    : [v0] "v"(V1[0]), [v1] "v"(V1[1]), [v2] "v"(V1[2]), [v3] "v"(V1[3]), [cl] "s"(e1), \
      [cb] "i"(g / 4), [rA] "s"(rA), [rB] "v"(rBv), [ct] "v"(colterm), [rw] RW          \
    : "scc"
                if constexpr ((kRowB & (kRowB - 1)) == 0) {
                    if constexpr (INSIDE)
                        asm(ZS_ROWS_ ZS_SKIP_ ZS_CLAMP_ ZS_ADDR_("v_lshl_add_u32")
                            ZS_OPS_("i"(kRowShift)));
                    else
                        asm(ZS_ROWS_ ZS_CLAMP_ ZS_ADDR_("v_lshl_add_u32") ZS_OPS_("i"(kRowShift)));
                } else {
                    if constexpr (INSIDE)
                        asm(ZS_ROWS_ ZS_SKIP_ ZS_CLAMP_ ZS_ADDR_("v_mad_i32_i24")
                            ZS_OPS_("s"(kRowB)));
                    else
                        asm(ZS_ROWS_ ZS_CLAMP_ ZS_ADDR_("v_mad_i32_i24") ZS_OPS_("s"(kRowB)));
                }
#undef ZS_ROWS_
#undef ZS_CLAMP_
#undef ZS_SKIP_
#undef ZS_ADDR_
#undef ZS_OPS_
                typename Gt::Raw q[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) q[i] = Lg::load(lstage + s * kSlot + a[i]);
                // all four reads in flight before the first blend waits on one
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (Gt::kPair) {
                    float2 c[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[i] = Gt::chan(q[i], fr[i]);
                    acc_weighted_4p<g>(acc[g], acc[g + 1], acc[g + 2], acc[g + 3], c[0], c[1],
                                       c[2], c[3], Dey, wu, m.y, ent[g].y, ent[g + 1].y,
                                       ent[g + 2].y, ent[g + 3].y, B);
                } else {
                    float t[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) t[i] = Gt::lerp(q[i], fr[i], Dey, wu);
                    acc_weighted_4<g>(acc[g], acc[g + 1], acc[g + 2], acc[g + 3], t[0], t[1],
                                      t[2], t[3], m.y, ent[g].y, ent[g + 1].y, ent[g + 2].y,
                                      ent[g + 3].y, B);
                }
            });
        }
    };
    // the chunk's table entries in one scalar load: views p0 .. p0 + 3 (a row
    // of the table holds a multiple of four views, and the table ends with four
    // spare entries, so a two-view chunk's load of four stays inside it)
    if constexpr (kPrefetch) {
        // one barrier per chunk: it makes chunk p0's copies (issued a chunk
        // earlier) visible, and ends every read of the other buffer before the
        // next chunk's copies overwrite it
        {
            const u32x4_v e0 = *(cptr<u32x4_v>)ftab;
            stage_chunk(pmin, 0, e0.x, e0.z, 0u, 0u);
        }
        int buf = 0;
        for (int p0 = pmin; p0 <= pmax; p0 += 2) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const u32x8_v e8 = *(cptr<u32x8_v>)(ftab + 2 * (p0 - pmin));
            if (p0 + 2 <= pmax) stage_chunk(p0 + 2, buf ^ 1, e8[4], e8[6], 0u, 0u);
            consume_chunk(p0, buf, e8);
            buf ^= 1;
        }
    } else {
        for (int p0 = pmin; p0 <= pmax; p0 += 4) {
            const u32x8_v e8 = *(cptr<u32x8_v>)(ftab + 2 * (p0 - pmin));
            stage_chunk(p0, 0, e8[0], e8[2], e8[4], e8[6]);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            consume_chunk(p0, 0, e8);
            __syncthreads();
        }
    }
    if (xi < A.XN && yi < A.YN) {
        float* o = img + ((long long)yi * A.XN + xi) * A.ZN + z0;
        if (A.vec4 && z0 + ZB <= A.ZN) {
#pragma unroll
            for (int j = 0; j < ZB; j += 4)
                *reinterpret_cast<float4*>(o + j) =
                    make_float4(acc[j] * A.DeltaFai, acc[j + 1] * A.DeltaFai,
                                acc[j + 2] * A.DeltaFai, acc[j + 3] * A.DeltaFai);
        } else {
#pragma unroll
            for (int j = 0; j < ZB; ++j)
                if (z0 + j < A.ZN) o[j] = acc[j] * A.DeltaFai;
        }
    }
}

template <int ZB, class T, int LAYOUT, int RMAX, bool INSIDE>
__global__ __launch_bounds__(256) void bp_fdk3dw_zs(float* __restrict__ img,
                                                    const T* __restrict__ fp,
                                                    const float2* __restrict__ vtab,
                                                    const ZBlock* __restrict__ blocks,
                                                    const uint2* __restrict__ masks,
                                                    const float2* __restrict__ table, BPArgs A) {
    zs_run<ZB, T, LAYOUT, RMAX, INSIDE>(xcd_logical_block(), A, img, fp, vtab, blocks, masks, table);
}

// Batched z-column launch: several plans (z-slabs of one or more scans, each
// with its own geometry, volume and filtered sinogram) in one grid.  Slab s owns
// logical workgroups [wg_begin, wg_begin + its grid); the XCD remap runs over
// the whole grid, so each XCD walks a contiguous run across slab boundaries.
struct SlabDesc {
    BPArgs A;
    float* img;
    const void* fp;
    const float2* vtab;
    const ZBlock* blocks;
    const uint2* masks;
    const float2* table;
    int wg_begin;
    int pad;
};

template <int ZB, int G, class T, int LAYOUT>
__global__ __launch_bounds__(256) void bp_fdk3dw_zc_batch(const SlabDesc* __restrict__ slabs,
                                                          int nslabs) {
    const int lb = xcd_logical_block();
    int s = 0;
    while (s + 1 < nslabs && lb >= slabs[s + 1].wg_begin) ++s;
    const SlabDesc& d = slabs[s];
    zc_run<ZB, G, T, LAYOUT, false>(lb - d.wg_begin, d.A, d.img, static_cast<const T*>(d.fp), d.vtab,
                             d.blocks, d.masks, d.table);
}

// The LDS-staged kernel over a batch of slabs: each slab's descriptor carries
// its plan's footprint table (built when the batch was created).  INSIDE when
// every plan of the batch is: the slabs then run exactly the code of their
// single launches (bit-identical volumes).
template <int ZB, class T, int LAYOUT, int RMAX, bool INSIDE>
__global__ __launch_bounds__(256) void bp_fdk3dw_zs_batch(const SlabDesc* __restrict__ slabs,
                                                          int nslabs) {
    const int lb = xcd_logical_block();
    int s = 0;
    while (s + 1 < nslabs && lb >= slabs[s + 1].wg_begin) ++s;
    const SlabDesc& d = slabs[s];
    zs_run<ZB, T, LAYOUT, RMAX, INSIDE>(lb - d.wg_begin, d.A, d.img, static_cast<const T*>(d.fp),
                                        d.vtab, d.blocks, d.masks, d.table);
}

// FP64 path: the same math as the reference, one thread per voxel, in double.
__global__ __launch_bounds__(256) void bp_fdk3dw_f64(double* __restrict__ img,
                                                     const double* __restrict__ fp,
                                                     const double* __restrict__ w,
                                                     arecon_bp_geom g, double invPI,
                                                     double pi_ref) {
    const long long n = (long long)g.XN * g.YN * g.ZN;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int zi = (int)(id % g.ZN);
    const long long xy = id / g.ZN;
    const int xi = (int)(xy % g.XN);
    const int yi = (int)(xy / g.XN);
    const double x = (xi - g.XNC) * g.dx;
    const double y = (yi - g.YNC) * g.dy;
    const double z = (zi - g.ZNC) * g.dz;
    const double Beta0 = 2.0 * pi_ref * z / g.h;
    const int s0 = (int)ceil((Beta0 - g.BetaS) / g.DeltaFai - 0.5);
    int s1 = (int)(s0 - ceil(g.N_pi * g.HSCoef));
    int s2 = (int)(s0 + ceil(g.N_pi * g.HSCoef));
    double res = 0;
    if ((s1 < g.PN) && (s2 > 0)) {
        s1 = (s1 < 0) ? 0 : s1;
        s2 = (s2 > g.PN - 1) ? g.PN - 1 : s2;
        for (int p = s1; p <= s2; ++p) {
            const double View = g.BetaS + p * g.DeltaFai;
            const int d1 = (int)(g.N_pi - (s0 - p));
            const int d2 = (int)((p < s0) ? (d1 + g.N_pi) : (d1 - g.N_pi));
            double sv, cv;
            sincos(View, &sv, &cv);
            const double UU = -x * cv - y * sv;
            const double Yr = -x * sv + y * cv;
            const double temp1 = sqrt(g.SO * g.SO - Yr * Yr);
            const double temp2 = z - g.h * View * invPI;
            const double Zr = temp2 * (g.SO * g.SD) / (UU * temp1 + g.SO * g.SO - Yr * Yr);
            const double U1 = Yr / g.dYL + g.YLC;
            const int U = (int)ceil(U1);
            const double V1 = Zr / g.dZL + g.ZLC;
            const int V = (int)ceil(V1);
            const double Dey = U - U1, Dez = V - V1;
            if ((U > 0) && (U < g.YL) && (V > 0) && (V < g.ZL)) {
                const double* f = fp + (long long)p * g.ZL * g.YL;
                const double tou = Dey * Dez * f[(V - 1) * g.YL + (U - 1)] +
                                   Dey * (1.0 - Dez) * f[V * g.YL + (U - 1)] +
                                   (1.0 - Dey) * Dez * f[(V - 1) * g.YL + U] +
                                   (1.0 - Dey) * (1.0 - Dez) * f[V * g.YL + U];
                const double w1 = w[d1], w2 = w[d2];
                const double G = fabs(temp2 / (temp1 + UU));
                const double Gc = (p < s0) ? fabs((temp2 - 0.5) / (temp1 - UU))
                                           : fabs((temp2 + 0.5) / (temp1 - UU));
                const double m1 = pow(G, g.k1), m2 = pow(Gc, g.k1);
                res += (w1 * m2) / (w2 * m1 + w1 * m2) * tou * g.DeltaFai;
            }
        }
    }
    img[((long long)yi * g.XN + xi) * g.ZN + zi] = res;
}

// File-level pi of the reference (src/Parallel_FDK_Helical_3DWeighting.cpp:16),
// used for Beta0 (:62); invPI uses the full constant (:51).
constexpr double kRefPi = 3.14159265358979;
constexpr double kInvTwoPi = 1.0 / (2.0 * 3.141592653589793);

// Index arithmetic of :74-75 in the reference's types (double N_pi, int result).
inline int weight_index1(double N_pi, int k) { return (int)(N_pi + k); }
inline int weight_index2(double N_pi, int k) {
    const int d1 = weight_index1(N_pi, k);
    return (int)((k < 0) ? (d1 + N_pi) : (d1 - N_pi));
}

void validate_geom(const arecon_bp_geom& g) {
    ARECON_REQUIRE(g.XN > 0 && g.YN > 0 && g.ZN > 0, "bp: XN, YN and ZN must be positive");
    ARECON_REQUIRE(g.YL >= 2 && g.ZL >= 2, "bp: YL and ZL must be at least 2");
    ARECON_REQUIRE(g.PN > 0 && g.N_2pi > 0, "bp: PN and N_2pi must be positive");
    ARECON_REQUIRE(g.h != 0.0, "bp: helical pitch h must be nonzero");
    ARECON_REQUIRE(g.DeltaFai > 0.0, "bp: DeltaFai must be positive");
    ARECON_REQUIRE(g.HSCoef >= 0.0, "bp: HSCoef must be non-negative");
    ARECON_REQUIRE(g.SO > 0.0 && g.dYL != 0.0 && g.dZL != 0.0, "bp: SO, dYL, dZL invalid");
    const int W = (int)std::ceil(g.N_pi * g.HSCoef);
    for (int k = -W; k <= W; ++k) {
        const int d1 = weight_index1(g.N_pi, k), d2 = weight_index2(g.N_pi, k);
        ARECON_REQUIRE(d1 >= 0 && d1 < g.N_2pi && d2 >= 0 && d2 < g.N_2pi,
                       "bp: N_pi*HSCoef puts a weight index outside w[0, N_2pi) "
                       "(the reference reads out of bounds there)");
    }
}

int auto_nz(const std::vector<BPZ>& z, int W) {
    // Largest NZ whose union of NZ consecutive windows stays within ~5% of one
    // window (slices further apart share too few views to amortise the column
    // work).
    const int n = (int)z.size();
    double ds = 0.0;
    if (n > 1) ds = std::fabs((double)(z[n - 1].s0 - z[0].s0)) / (n - 1);
    const double L = 2.0 * W + 1.0;
    for (int nz : {16, 8, 4, 2})
        if ((nz - 1) * ds <= 0.05 * L && n >= nz) return nz;
    return 1;
}

}  // namespace

struct BPPlan {
    arecon_bp_geom g;
    int device = 0;
    float2* vtab = nullptr;
    float4* ktab = nullptr;
    ZSet* zsets = nullptr;      // per slice set: view ranges and table offset
    float2* vt_table = nullptr; // per (slice set or block, view, slice): temp2, a
    ZBlock* zblocks = nullptr;  // z-column kernel: per block view range and offset
    uint2* zmasks = nullptr;    // z-column kernel: per (block, view) slice masks
    unsigned* zrec = nullptr;   // LDS-staged kernel: per (block, view) records
    BPZ* ztab = nullptr;
    unsigned long long* stats = nullptr;   // LDS-staged kernel: workgroups staged / direct
    BPArgs a{};
    int variant = 0;
    int nz = 1;
    int group = 1;      // z-column kernel: slices per branch-free group
    int grid = 0;
    std::vector<int> blk_pmin, blk_pmax;   // per slice set: union of the windows
    // LDS-staged kernel: footprint tables, built at the first launch of each
    // record kind (8-byte records over YL channels, quads over YL-1, 4-byte
    // records over YL)
    struct ZsTable {
        uint2* tab = nullptr;
        long long* boff = nullptr;
        int* fits = nullptr;
    };
    ZsTable zs_tables[3];
    std::mutex zs_mu;
};

// One slice's entry for one view, FP64 on the host: temp2 = z - h V / 2pi (:79) and
// a = 2^Q, Q = log2(w2/w1) + k1 log2(|temp2|/|temp2 -+ 0.5|) (:95-102); a = 0 where
// w2 = 0.  Returns false (entry (1e30, 0)) when the view is outside the slice's
// clamped window (:63-71); `weighted` tells whether w2 != 0.
static bool slice_view_entry(const arecon_bp_geom& g, const double* w, const std::vector<BPZ>& zt,
                             int zi, int p, float2& e, bool& weighted) {
    e = make_float2(1e30f, 0.f);
    weighted = false;
    if (zi >= g.ZN || p < zt[zi].s1 || p > zt[zi].s2) return false;
    const int k = p - zt[zi].s0;
    const double z = (zi - g.ZNC) * g.dz;
    const double V = g.BetaS + p * g.DeltaFai;
    const double temp2 = z - g.h * V * kInvTwoPi;
    const double w1 = w[weight_index1(g.N_pi, k)], w2 = w[weight_index2(g.N_pi, k)];
    double Q = -INFINITY;
    if (w2 != 0.0) {
        Q = std::log2(w2) - std::log2(w1);
        if (g.k1 != 0.0)
            Q += g.k1 * (std::log2(std::fabs(temp2)) -
                         std::log2(std::fabs(temp2 + (k < 0 ? -0.5 : 0.5))));
        weighted = true;
    }
    // 2^Q in FP32: 0 for Q = -inf, +inf for Q = +inf (weight 0)
    e = make_float2((float)temp2, (float)std::exp2(Q));
    return true;
}

// Per-(slice set, view, slice) table of the view-table kernel.  A slice whose
// window does not contain the view gets temp2 = 1e30 (its row is off the
// detector).
static void build_view_table(BPPlan* P, const double* w, const std::vector<BPZ>& zt) {
    const arecon_bp_geom& g = P->g;
    const int NZ = P->nz, nsets = P->a.nzsets;
    std::vector<ZSet> sets(nsets);
    std::vector<float2> tab;
    for (int zs = 0; zs < nsets; ++zs) {
        ZSet h{};
        h.pmin = INT_MAX; h.pmax = INT_MIN;
        for (int j = 0; j < NZ; ++j) {
            const int zi = zs * NZ + j;
            if (zi < g.ZN && zt[zi].s1 <= zt[zi].s2) {
                h.pmin = std::min(h.pmin, zt[zi].s1);
                h.pmax = std::max(h.pmax, zt[zi].s2);
            }
        }
        h.off = (long long)tab.size();
        h.clo = 1; h.chi = 0;
        if (h.pmin > h.pmax) { h.pmin = 1; h.pmax = 0; sets[zs] = h; continue; }
        int run_lo = -1, best_lo = 1, best_hi = 0;
        for (int p = h.pmin; p <= h.pmax; ++p) {
            bool unweighted = true;
            for (int j = 0; j < NZ; ++j) {
                float2 e;
                bool wt;
                slice_view_entry(g, w, zt, zs * NZ + j, p, e, wt);
                if (wt) unweighted = false;
                tab.push_back(e);
            }
            if (unweighted) {
                if (run_lo < 0) run_lo = p;
                if (p - run_lo > best_hi - best_lo) { best_lo = run_lo; best_hi = p; }
            } else {
                run_lo = -1;
            }
        }
        h.clo = best_lo; h.chi = best_hi;
        sets[zs] = h;
    }
    if (tab.empty()) tab.push_back(make_float2(1e30f, 0.f));
    ARECON_HIP(hipMalloc(&P->zsets, sizeof(ZSet) * sets.size()));
    ARECON_HIP(hipMalloc(&P->vt_table, sizeof(float2) * tab.size()));
    ARECON_HIP(hipMemcpy(P->zsets, sets.data(), sizeof(ZSet) * sets.size(), hipMemcpyHostToDevice));
    ARECON_HIP(hipMemcpy(P->vt_table, tab.data(), sizeof(float2) * tab.size(), hipMemcpyHostToDevice));
}

// Per-(block, view) masks and per-(block, view, slice) entries of the z-column
// kernel.  Bit j of masks.x: slice j's window contains the view; bit j of
// masks.y: that slice's weight has a conjugate term (a != 0).
static void build_block_table(BPPlan* P, const double* w, const std::vector<BPZ>& zt,
                              const std::vector<float2>& vt) {
    const arecon_bp_geom& g = P->g;
    const int ZB = P->nz, nblk = P->a.nzsets;
    std::vector<ZBlock> blks(nblk);
    std::vector<float2> tab;
    std::vector<uint2> msk;
    for (int zb = 0; zb < nblk; ++zb) {
        ZBlock h{};
        h.pmin = INT_MAX; h.pmax = INT_MIN;
        for (int j = 0; j < ZB; ++j) {
            const int zi = zb * ZB + j;
            if (zi < g.ZN && zt[zi].s1 <= zt[zi].s2) {
                h.pmin = std::min(h.pmin, zt[zi].s1);
                h.pmax = std::max(h.pmax, zt[zi].s2);
            }
        }
        h.off = (long long)tab.size();
        if (h.pmin > h.pmax) { h.pmin = 1; h.pmax = 0; blks[zb] = h; continue; }
        for (int p = h.pmin; p <= h.pmax; ++p) {
            uint2 m = make_uint2(0u, 0u);
            for (int j = 0; j < ZB; ++j) {
                float2 e;
                bool wt;
                if (slice_view_entry(g, w, zt, zb * ZB + j, p, e, wt)) {
                    m.x |= 1u << j;
                    if (wt) m.y |= 1u << j;
                }
                tab.push_back(e);
            }
            msk.push_back(m);
        }
        blks[zb] = h;
    }
    if (tab.empty()) {
        tab.assign(ZB, make_float2(1e30f, 0.f));
        msk.push_back(make_uint2(0u, 0u));
    }
    ARECON_HIP(hipMalloc(&P->zblocks, sizeof(ZBlock) * blks.size()));
    ARECON_HIP(hipMalloc(&P->zmasks, sizeof(uint2) * msk.size()));
    ARECON_HIP(hipMalloc(&P->vt_table, sizeof(float2) * tab.size()));
    ARECON_HIP(hipMemcpy(P->zblocks, blks.data(), sizeof(ZBlock) * blks.size(), hipMemcpyHostToDevice));
    ARECON_HIP(hipMemcpy(P->zmasks, msk.data(), sizeof(uint2) * msk.size(), hipMemcpyHostToDevice));
    ARECON_HIP(hipMemcpy(P->vt_table, tab.data(), sizeof(float2) * tab.size(), hipMemcpyHostToDevice));
    if (P->variant != 3) return;
    // records of the LDS-staged kernel: record i (mask index i, view p of block
    // zb at blks[zb].off / ZB + p - pmin) = masks, cos / sin of view p, the
    // block's ZB entries
    const int RW = 4 + 2 * ZB;
    std::vector<unsigned> rec(msk.size() * RW, 0u);
    for (int zb = 0; zb < nblk; ++zb) {
        const ZBlock& h = blks[zb];
        for (int p = h.pmin; p <= h.pmax; ++p) {
            const long long i = h.off / ZB + (p - h.pmin);
            unsigned* r = rec.data() + i * RW;
            // bits 16 + k (active word: some slice of group k of four active;
            // weighted word: some slice weighted) and 20 + k (weighted word: every
            // active slice of the group weighted)
            r[0] = msk[i].x;
            r[1] = msk[i].y;
            for (int k = 0; k < ZB / 4; ++k) {
                const unsigned act = (msk[i].x >> (4 * k)) & 0xfu;
                const unsigned wt = (msk[i].y >> (4 * k)) & 0xfu;
                if (act) r[0] |= 1u << (16 + k);
                if (wt) r[1] |= 1u << (16 + k);
                if (wt && wt == act) r[1] |= 1u << (20 + k);
            }
            std::memcpy(r + 2, &vt[p], sizeof(float2));
            std::memcpy(r + 4, &tab[i * ZB], sizeof(float2) * ZB);
        }
    }
    ARECON_HIP(hipMalloc(&P->zrec, sizeof(unsigned) * rec.size()));
    ARECON_HIP(hipMemcpy(P->zrec, rec.data(), sizeof(unsigned) * rec.size(), hipMemcpyHostToDevice));
    P->a.zs_rec = P->zrec;
}

BPPlan* bp_plan_create(const arecon_bp_geom& g, const double* w) {
    validate_geom(g);
    ARECON_REQUIRE(w != nullptr, "bp: null weight table");
    const int W = (int)std::ceil(g.N_pi * g.HSCoef);
    // Per-view table (FP64 -> FP32).
    std::vector<float2> vt(g.PN);
    for (int p = 0; p < g.PN; ++p) {
        const double V = g.BetaS + p * g.DeltaFai;
        vt[p] = make_float2((float)std::cos(V), (float)std::sin(V));
    }
    // Per-offset table k = p - s0 in [-W, W].
    std::vector<float4> kt(2 * W + 1);
    for (int k = -W; k <= W; ++k) {
        const double w1 = w[weight_index1(g.N_pi, k)], w2 = w[weight_index2(g.N_pi, k)];
        const double r = (w1 + w2 != 0.0) ? w1 / (w1 + w2) : 0.0;
        kt[k + W] = make_float4((float)w1, (float)w2, k < 0 ? -0.5f : 0.5f, (float)r);
    }
    // Per-slice table.
    std::vector<BPZ> zt(g.ZN);
    for (int zi = 0; zi < g.ZN; ++zi) {
        const double z = (zi - g.ZNC) * g.dz;
        const double Beta0 = 2.0 * kRefPi * z / g.h;
        const int s0 = (int)std::ceil((Beta0 - g.BetaS) / g.DeltaFai - 0.5);
        int s1 = s0 - W, s2 = s0 + W;
        BPZ e;
        e.s0 = s0;
        if ((s1 < g.PN) && (s2 > 0)) {
            e.s1 = std::max(s1, 0);
            e.s2 = std::min(s2, g.PN - 1);
        } else {
            e.s1 = 1;
            e.s2 = 0;
        }
        e.dz = (float)(z - g.h * (g.BetaS + s0 * g.DeltaFai) * kInvTwoPi);
        zt[zi] = e;
    }
    ARECON_REQUIRE(g.variant >= 0 && g.variant <= 3, "bp: variant must be 0, 1, 2 or 3");
    ARECON_REQUIRE(g.order >= 0 && g.order <= 2, "bp: order must be 0, 1 or 2");
    int nz;
    int group = 1;
    if (g.variant == 3) {
        nz = g.nz > 0 ? g.nz : 8;
        ARECON_REQUIRE(nz == 8 || nz == 16,
                       "bp: nz of the LDS-staged z-column kernel must be 0, 8 or 16");
        group = 8;
    } else if (g.variant == 2) {
        nz = g.nz > 0 ? g.nz : 32;
        ARECON_REQUIRE(nz == 8 || nz == 16 || nz == 32,
                       "bp: nz of the z-column kernel must be 0, 8, 16 or 32");
        group = g.group > 0 ? g.group : 2;
        ARECON_REQUIRE(group == 1 || group == 2 || group == 4 || (group == 8 && nz <= 16),
                       "bp: group of the z-column kernel must be 0, 1, 2, 4, or 8 (split "
                       "gathers and blends, nz <= 16)");
    } else {
        nz = g.nz > 0 ? g.nz : auto_nz(zt, W);
        ARECON_REQUIRE(nz == 1 || nz == 2 || nz == 4 || nz == 8 || nz == 16,
                       "bp: nz must be 0, 1, 2, 4, 8 or 16");
    }
    BPPlan* P = new BPPlan();
    P->g = g;
    P->variant = g.variant;
    P->nz = nz;
    P->group = group;
    BPArgs& A = P->a;
    A.XN = g.XN; A.YN = g.YN; A.ZN = g.ZN; A.YL = g.YL; A.ZL = g.ZL; A.PN = g.PN; A.W = W;
    A.tilesx = (g.XN + 15) / 16;
    const int tilesy = (g.YN + 15) / 16;
    A.ntiles = A.tilesx * tilesy;
    A.nzsets = (g.ZN + P->nz - 1) / P->nz;
    A.order = g.order;
    A.k1zero = (g.k1 == 0.0);
    A.XNC = (float)g.XNC; A.YNC = (float)g.YNC; A.dx = (float)g.dx; A.dy = (float)g.dy;
    A.YLC = (float)g.YLC; A.ZLC = (float)g.ZLC; A.invdYL = (float)(1.0 / g.dYL);
    A.KSD = (float)(g.SO * g.SD / g.dZL);
    A.YLm1 = (float)(g.YL - 1); A.ZLm1 = (float)(g.ZL - 1);
    A.SO2 = (float)(g.SO * g.SO); A.hstep = (float)(g.h * g.DeltaFai * kInvTwoPi);
    A.k1 = (float)g.k1; A.DeltaFai = (float)g.DeltaFai;
    A.view_stride = (long long)g.ZL * g.YL;
    const long long grid = (long long)A.ntiles * A.nzsets;
    P->grid = (int)grid;
    {
        // Largest radius of any column a workgroup computes (tiles padded to 16),
        // in FP64.  Inside when t1 + UU >= sqrt(SO^2 - R^2) - R stays well above 0
        // (R < SO / sqrt 2 with a margin) and every U1 = Yr / dYL + YLC lies in
        // (1, YL - 2).
        double R2 = 0.0;
        const double xs[2] = {(0 - g.XNC) * g.dx, (A.tilesx * 16 - 1 - g.XNC) * g.dx};
        const double ys[2] = {(0 - g.YNC) * g.dy, ((A.ntiles / A.tilesx) * 16 - 1 - g.YNC) * g.dy};
        for (double xv : xs)
            for (double yv : ys) R2 = std::max(R2, xv * xv + yv * yv);
        const double R = std::sqrt(R2);
        const double u = R / std::fabs(g.dYL);
        A.inside = (R < 0.99 * g.SO / std::sqrt(2.0)) && (g.YLC - u > 1.0) &&
                   (g.YLC + u < g.YL - 2.0) && std::isfinite(g.k1) ? 1 : 0;
    }
    P->blk_pmin.assign(A.nzsets, 1);
    P->blk_pmax.assign(A.nzsets, 0);
    for (int zs = 0; zs < A.nzsets; ++zs) {
        int lo = INT_MAX, hi = INT_MIN;
        for (int j = 0; j < nz; ++j) {
            const int zi = zs * nz + j;
            if (zi < g.ZN && zt[zi].s1 <= zt[zi].s2) {
                lo = std::min(lo, zt[zi].s1);
                hi = std::max(hi, zt[zi].s2);
            }
        }
        if (lo <= hi) { P->blk_pmin[zs] = lo; P->blk_pmax[zs] = hi; }
    }
    try {
        // one view in any layout (rows, row pairs at twice and quads at four times
        // the size) stays below 2 GiB, and rows and row strides fit the kernels'
        // 24-bit multiplies
        ARECON_REQUIRE(A.view_stride * 16 < (1LL << 31), "bp: one view exceeds 128 MiB");
        ARECON_REQUIRE((long long)g.YL * 16 < (1LL << 24) && g.ZL < (1 << 24),
                       "bp: detector too large for 24-bit row offsets");
        ARECON_REQUIRE(grid < (1LL << 31), "bp: volume too large for one launch");
        ARECON_HIP(hipGetDevice(&P->device));
        ARECON_HIP(hipMalloc(&P->vtab, sizeof(float2) * vt.size()));
        ARECON_HIP(hipMalloc(&P->ktab, sizeof(float4) * kt.size()));
        ARECON_HIP(hipMalloc(&P->ztab, sizeof(BPZ) * zt.size()));
        if (P->variant == 0) build_view_table(P, w, zt);
        if (P->variant >= 2) build_block_table(P, w, zt, vt);
        if (P->variant == 3) {
            ARECON_REQUIRE(g.ZL <= 250, "bp: variant 3 takes at most 250 detector rows");
            // the footprint table stores the first channel in 16 signed bits
            ARECON_REQUIRE(g.YL < 32768, "bp: variant 3 takes at most 32767 detector channels");
            ARECON_HIP(hipMalloc(&P->stats, 2 * sizeof(unsigned long long)));
            ARECON_HIP(hipMemset(P->stats, 0, 2 * sizeof(unsigned long long)));
            A.stats = P->stats;
        }
        ARECON_HIP(hipMemcpy(P->vtab, vt.data(), sizeof(float2) * vt.size(), hipMemcpyHostToDevice));
        ARECON_HIP(hipMemcpy(P->ktab, kt.data(), sizeof(float4) * kt.size(), hipMemcpyHostToDevice));
        ARECON_HIP(hipMemcpy(P->ztab, zt.data(), sizeof(BPZ) * zt.size(), hipMemcpyHostToDevice));
    } catch (...) {
        bp_plan_destroy(P);
        throw;
    }
    return P;
}

int bp_plan_nz(const BPPlan* P) { return P ? P->nz : 0; }

int bp_plan_nblocks(const BPPlan* P) { return P ? P->a.nzsets : 0; }

int bp_plan_inside(const BPPlan* P) { return P ? P->a.inside : 0; }

void bp_plan_stats(const BPPlan* P, unsigned long long* staged, unsigned long long* direct) {
    ARECON_REQUIRE(P != nullptr && staged != nullptr && direct != nullptr, "bp: null argument");
    unsigned long long h[2] = {0ull, 0ull};
    if (P->stats) {
        int cur = 0;
        ARECON_HIP(hipGetDevice(&cur));
        ARECON_HIP(hipSetDevice(P->device));
        const hipError_t e = hipMemcpy(h, P->stats, sizeof(h), hipMemcpyDeviceToHost);
        (void)hipSetDevice(cur);
        ARECON_HIP(e);
    }
    *staged = h[0];
    *direct = h[1];
}

void bp_plan_block_views(const BPPlan* P, int block, int* pmin, int* pmax) {
    ARECON_REQUIRE(P != nullptr && pmin != nullptr && pmax != nullptr, "bp: null argument");
    ARECON_REQUIRE(block >= 0 && block < P->a.nzsets, "bp: slice block out of range");
    *pmin = P->blk_pmin[block];
    *pmax = P->blk_pmax[block];
}

void bp_plan_destroy(BPPlan* P) {
    if (!P) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(P->device);
    if (P->vtab) (void)hipFree(P->vtab);
    if (P->ktab) (void)hipFree(P->ktab);
    if (P->ztab) (void)hipFree(P->ztab);
    if (P->zsets) (void)hipFree(P->zsets);
    if (P->vt_table) (void)hipFree(P->vt_table);
    if (P->zblocks) (void)hipFree(P->zblocks);
    if (P->zmasks) (void)hipFree(P->zmasks);
    if (P->zrec) (void)hipFree(P->zrec);
    if (P->stats) (void)hipFree(P->stats);
    for (BPPlan::ZsTable& z : P->zs_tables) {
        if (z.tab) (void)hipFree(z.tab);
        if (z.boff) (void)hipFree(z.boff);
        if (z.fits) (void)hipFree(z.fits);
    }
    (void)hipSetDevice(cur);
    delete P;
}

// Kernel arguments of one launch of plan P: the view stride of the filtered
// layout and whether the volume takes float4 stores.
static BPArgs launch_args(const BPPlan* P, const float* img, int layout, size_t elem) {
    ARECON_REQUIRE(layout == ARECON_FP_ROWS || P->variant >= 2,
                   "bp: the row-pair, row-difference and quad layouts need a z-column kernel "
                   "(variant 2 or 3)");
    ARECON_REQUIRE(P->variant != 3 || layout != ARECON_FP_ROWS,
                   "bp: the LDS-staged z-column kernel (variant 3) reads row pairs, row "
                   "differences or quads");
    BPArgs a = P->a;
    a.vec4 = (a.ZN % 4 == 0) && ((reinterpret_cast<uintptr_t>(img) & 15) == 0);
    if (layout == ARECON_FP_ROWPAIRS || layout == ARECON_FP_ROWDIFF)
        a.view_stride = 2LL * (a.ZL - 1) * a.YL;
    if (layout == ARECON_FP_QUADS) a.view_stride = 4LL * (a.ZL - 1) * (a.YL - 1);
    a.view_bytes = (unsigned)(a.view_stride * (long long)elem);
    return a;
}

// The plan's footprint table for the LDS-staged kernel reading layout LY, built
// on stream s at its first use (allocates, then synchronizes s once, so that a
// launch on another stream never reads a table still being written).
template <int NZ, class T, int LY, int RMAX>
static const BPPlan::ZsTable& zs_table(const BPPlan* P, hipStream_t s) {
    constexpr int kRec = (int)Gather<T, LY>::kElem;
    const int kind = kRec == 4 ? 2 : (LY == ARECON_FP_QUADS ? 1 : 0);
    BPPlan* M = const_cast<BPPlan*>(P);
    std::lock_guard<std::mutex> lock(M->zs_mu);
    BPPlan::ZsTable& z = M->zs_tables[kind];
    if (z.tab) return z;
    const BPArgs& A = P->a;
    std::vector<long long> boff(A.nzsets);
    long long total = 0;
    for (int zs = 0; zs < A.nzsets; ++zs) {
        boff[zs] = total;
        total += (long long)A.ntiles * zs_table_views(P->blk_pmin[zs], P->blk_pmax[zs]);
    }
    BPPlan::ZsTable t;
    try {
        ARECON_HIP(hipMalloc(&t.boff, sizeof(long long) * A.nzsets));
        ARECON_HIP(hipMalloc(&t.fits, sizeof(int) * (size_t)A.nzsets * A.ntiles));
        ARECON_HIP(hipMalloc(&t.tab, sizeof(uint2) * (size_t)(total + 4)));
        ARECON_HIP(hipMemcpy(t.boff, boff.data(), sizeof(long long) * A.nzsets,
                             hipMemcpyHostToDevice));
        const int nrec = LY == ARECON_FP_QUADS ? A.YL - 1 : A.YL;
        const dim3 grid((unsigned)(A.nzsets * A.ntiles)), block(256);
        if (A.inside)
            hipLaunchKernelGGL((bp_zs_table<NZ, RMAX, true, kRec == 8>), grid, block, 0, s, A,
                               P->vtab, P->zblocks, P->zmasks, P->vt_table, t.tab, t.boff,
                               t.fits, nrec);
        else
            hipLaunchKernelGGL((bp_zs_table<NZ, RMAX, false, kRec == 8>), grid, block, 0, s, A,
                               P->vtab, P->zblocks, P->zmasks, P->vt_table, t.tab, t.boff,
                               t.fits, nrec);
        ARECON_HIP(hipGetLastError());
        ARECON_HIP(hipStreamSynchronize(s));
    } catch (...) {
        if (t.tab) (void)hipFree(t.tab);
        if (t.boff) (void)hipFree(t.boff);
        if (t.fits) (void)hipFree(t.fits);
        throw;
    }
    z = t;
    return z;
}

// Slot rows of the LDS-staged kernel (a view's footprint loads in multiples of
// four rows, fewer than RMAX; the slot's last row is the zero row).  8-slice
// blocks of 8-byte records: 25 rows of 224 bytes, 21.9 KB of LDS per
// workgroup, seven workgroups per CU (21 rows of 256 bytes, the same LDS, left
// 10% of the clinical workgroups with a footprint too tall to stage, gathering
// from global memory; 25 rows, 1%).  16-slice blocks of 4-byte records: 48
// rows, 24.6 KB, six workgroups per CU (as the VGPRs allow): 40 rows leave 10%
// of the clinical workgroups to gather, 48 rows 1.4% (backprojection 23.7 ->
// 23.1 ms).  8-slice blocks of 4-byte records: 21 rows; 16-slice blocks of
// 8-byte records: 40 rows (35 KB, four workgroups per CU).
template <int NZ, class T, int LY> constexpr int zs_rmax() {
    return Gather<T, LY>::kElem == 4 ? (NZ == 8 ? 21 : 48) : (NZ == 8 ? 25 : 40);
}

// One launch of the LDS-staged kernel for layout LY.
template <int NZ, class T, int LY, int RMAX>
static void launch_zs(const BPPlan* P, BPArgs a, float* img, const T* fp, hipStream_t s,
                      int grid) {
    const BPPlan::ZsTable& t = zs_table<NZ, T, LY, RMAX>(P, s);
    a.zs_tab = t.tab;
    a.zs_boff = t.boff;
    a.zs_fits = t.fits;
    if (a.inside)
        hipLaunchKernelGGL((bp_fdk3dw_zs<NZ, T, LY, RMAX, true>), dim3(grid), dim3(256), 0, s,
                           img, fp, P->vtab, P->zblocks, P->zmasks, P->vt_table, a);
    else
        hipLaunchKernelGGL((bp_fdk3dw_zs<NZ, T, LY, RMAX, false>), dim3(grid), dim3(256), 0, s,
                           img, fp, P->vtab, P->zblocks, P->zmasks, P->vt_table, a);
}

template <int NZ, class T>
static void launch_nz(const BPPlan* P, float* img, const T* fp, int layout, hipStream_t s,
                      int lb_offset, int grid) {
    BPArgs a = launch_args(P, img, layout, sizeof(T));
    a.lb_offset = lb_offset;
    if constexpr (NZ <= 16) {
        if (P->variant == 1) {
            hipLaunchKernelGGL((bp_fdk3dw<NZ, T>), dim3(grid), dim3(256), 0, s, img, fp,
                               P->vtab, P->ktab, P->ztab, a);
            return;
        }
        if (P->variant == 0) {
            hipLaunchKernelGGL((bp_fdk3dw_vt<NZ, T>), dim3(grid), dim3(256), 0, s, img, fp,
                               P->vtab, P->zsets, P->vt_table, a);
            return;
        }
    }
    if constexpr (NZ == 8 || NZ == 16) {
        if (P->variant == 3) {
#define ARECON_ZS_(LY) launch_zs<NZ, T, LY, zs_rmax<NZ, T, LY>()>(P, a, img, fp, s, grid)
            if constexpr (std::is_same<T, float>::value) {
                ARECON_REQUIRE(layout == ARECON_FP_ROWPAIRS || layout == ARECON_FP_ROWDIFF,
                               "bp: variant 3 reads FP32 row pairs or row differences");
                if (layout == ARECON_FP_ROWDIFF) ARECON_ZS_(ARECON_FP_ROWDIFF);
                else ARECON_ZS_(ARECON_FP_ROWPAIRS);
            } else {
                ARECON_REQUIRE(layout == ARECON_FP_QUADS || layout == ARECON_FP_ROWDIFF,
                               "bp: variant 3 reads 16-bit quads or row differences");
                if (layout == ARECON_FP_ROWDIFF) ARECON_ZS_(ARECON_FP_ROWDIFF);
                else ARECON_ZS_(ARECON_FP_QUADS);
            }
#undef ARECON_ZS_
            return;
        }
    }
    if constexpr (NZ % 8 == 0 && NZ <= 32) {
        if (P->variant == 2) {
#define ARECON_ZC_(G, LY, IN)                                                                 \
    hipLaunchKernelGGL((bp_fdk3dw_zc<NZ, G, T, LY, IN>), dim3(grid), dim3(256), 0, s, img, fp, \
                       P->vtab, P->zblocks, P->zmasks, P->vt_table, a)
// split mode (G = 0) has the INSIDE specialisation
#define ARECON_ZC(G, LY)                                                                      \
    do {                                                                                      \
        if constexpr (G == 0) {                                                               \
            if (a.inside) ARECON_ZC_(0, LY, true);                                           \
            else ARECON_ZC_(0, LY, false);                                                   \
        } else {                                                                              \
            ARECON_ZC_(G, LY, false);                                                        \
        }                                                                                     \
    } while (0)
            if (layout == ARECON_FP_ROWS) {
                if (P->group == 1) ARECON_ZC(1, ARECON_FP_ROWS);
                else if (P->group == 2) ARECON_ZC(2, ARECON_FP_ROWS);
                else if (P->group == 4) ARECON_ZC(4, ARECON_FP_ROWS);
                else if constexpr (NZ <= 16) ARECON_ZC(0, ARECON_FP_ROWS);
            } else if (layout == ARECON_FP_ROWPAIRS) {
                if (P->group == 1) ARECON_ZC(1, ARECON_FP_ROWPAIRS);
                else if (P->group == 2) ARECON_ZC(2, ARECON_FP_ROWPAIRS);
                else if (P->group == 4) ARECON_ZC(4, ARECON_FP_ROWPAIRS);
                else if constexpr (NZ <= 16) ARECON_ZC(0, ARECON_FP_ROWPAIRS);
            } else if (layout == ARECON_FP_ROWDIFF) {
                if (P->group == 1) ARECON_ZC(1, ARECON_FP_ROWDIFF);
                else if (P->group == 2) ARECON_ZC(2, ARECON_FP_ROWDIFF);
                else if (P->group == 4) ARECON_ZC(4, ARECON_FP_ROWDIFF);
                else if constexpr (NZ <= 16) ARECON_ZC(0, ARECON_FP_ROWDIFF);
            } else {
                if (P->group == 1) ARECON_ZC(1, ARECON_FP_QUADS);
                else if (P->group == 2) ARECON_ZC(2, ARECON_FP_QUADS);
                else if (P->group == 4) ARECON_ZC(4, ARECON_FP_QUADS);
                else if constexpr (NZ <= 16) ARECON_ZC(0, ARECON_FP_QUADS);
            }
#undef ARECON_ZC
#undef ARECON_ZC_
            return;
        }
    }
    fail(ARECON_ERR_INTERNAL, "bp: no kernel for this variant and nz");
}

template <class T>
static void launch_t(const BPPlan* P, float* img, const T* fp, int layout, hipStream_t s,
                     int lb_offset, int grid) {
    switch (P->nz) {
    case 1: launch_nz<1, T>(P, img, fp, layout, s, lb_offset, grid); break;
    case 2: launch_nz<2, T>(P, img, fp, layout, s, lb_offset, grid); break;
    case 4: launch_nz<4, T>(P, img, fp, layout, s, lb_offset, grid); break;
    case 8: launch_nz<8, T>(P, img, fp, layout, s, lb_offset, grid); break;
    case 16: launch_nz<16, T>(P, img, fp, layout, s, lb_offset, grid); break;
    default: launch_nz<32, T>(P, img, fp, layout, s, lb_offset, grid); break;
    }
}

void bp_plan_launch(const BPPlan* P, float* img, const void* fp, int fp_dtype, int fp_layout,
                    hipStream_t s, int block_begin, int block_end) {
    ARECON_REQUIRE(P != nullptr, "bp: null plan");
    ARECON_REQUIRE(img != nullptr && fp != nullptr, "bp: null buffer");
    ARECON_REQUIRE(fp_layout == ARECON_FP_ROWS || fp_layout == ARECON_FP_ROWPAIRS ||
                       fp_layout == ARECON_FP_QUADS || fp_layout == ARECON_FP_ROWDIFF,
                   "bp: unknown fp layout");
    if (block_end < 0) block_end = P->a.nzsets;
    ARECON_REQUIRE(block_begin >= 0 && block_begin < block_end && block_end <= P->a.nzsets,
                   "bp: slice-block range outside the plan");
    ARECON_REQUIRE(P->a.order != 1 || (block_begin == 0 && block_end == P->a.nzsets),
                   "bp: slice-block ranges need the slice-major workgroup order");
    const int lb_offset = block_begin * P->a.ntiles;
    const int grid = (block_end - block_begin) * P->a.ntiles;
    switch (fp_dtype) {
    case ARECON_F32:
        launch_t<float>(P, img, static_cast<const float*>(fp), fp_layout, s, lb_offset, grid);
        break;
    case ARECON_F16:
        launch_t<__half>(P, img, static_cast<const __half*>(fp), fp_layout, s, lb_offset, grid);
        break;
    case ARECON_BF16:
        launch_t<__hip_bfloat16>(P, img, static_cast<const __hip_bfloat16*>(fp), fp_layout, s,
                                 lb_offset, grid);
        break;
    default: fail(ARECON_ERR_ARG, "bp: unknown fp dtype");
    }
    ARECON_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Batches of z-column plans bound to their volumes and filtered sinograms.
// ---------------------------------------------------------------------------
struct BPBatch {
    int device = 0;
    int variant = 2, nz = 0, group = 0, dtype = 0, layout = 0;
    int inside = 0;     // every plan INSIDE (the LDS-staged batch kernel's specialisation)
    int nslabs = 0;
    long long grid = 0;
    SlabDesc* d_slabs = nullptr;
};

// A plan's footprint table for the LDS-staged batch kernel (built on the null
// stream, which it synchronizes); the row-pair / row-difference / quad layouts
// of the single launches.
template <int NZ, class T>
static const BPPlan::ZsTable& batch_zs_table_t(const BPPlan* P, int layout) {
    if (layout == ARECON_FP_ROWDIFF)
        return zs_table<NZ, T, ARECON_FP_ROWDIFF, zs_rmax<NZ, T, ARECON_FP_ROWDIFF>()>(P, nullptr);
    if constexpr (std::is_same<T, float>::value) {
        ARECON_REQUIRE(layout == ARECON_FP_ROWPAIRS,
                       "bp batch: variant 3 reads FP32 row pairs or row differences");
        return zs_table<NZ, T, ARECON_FP_ROWPAIRS, zs_rmax<NZ, T, ARECON_FP_ROWPAIRS>()>(P,
                                                                                     nullptr);
    } else {
        ARECON_REQUIRE(layout == ARECON_FP_QUADS,
                       "bp batch: variant 3 reads 16-bit quads or row differences");
        return zs_table<NZ, T, ARECON_FP_QUADS, zs_rmax<NZ, T, ARECON_FP_QUADS>()>(P, nullptr);
    }
}
template <int NZ>
static const BPPlan::ZsTable& batch_zs_table_nz(const BPPlan* P, int dtype, int layout) {
    if (dtype == ARECON_F32) return batch_zs_table_t<NZ, float>(P, layout);
    if (dtype == ARECON_F16) return batch_zs_table_t<NZ, __half>(P, layout);
    return batch_zs_table_t<NZ, __hip_bfloat16>(P, layout);
}

BPBatch* bp_batch_create(const BPPlan* const* plans, float* const* imgs, const void* const* fps,
                         int n, int fp_dtype, int fp_layout) {
    ARECON_REQUIRE(n > 0 && plans && imgs && fps, "bp batch: empty or null arguments");
    ARECON_REQUIRE(fp_dtype == ARECON_F32 || fp_dtype == ARECON_F16 || fp_dtype == ARECON_BF16,
                   "bp batch: unknown fp dtype");
    ARECON_REQUIRE(fp_layout == ARECON_FP_ROWS || fp_layout == ARECON_FP_ROWPAIRS ||
                       fp_layout == ARECON_FP_QUADS || fp_layout == ARECON_FP_ROWDIFF,
                   "bp batch: unknown fp layout");
    std::vector<SlabDesc> slabs(n);
    long long wg = 0;
    for (int i = 0; i < n; ++i) {
        const BPPlan* P = plans[i];
        ARECON_REQUIRE(P && imgs[i] && fps[i], "bp batch: null plan or buffer");
        ARECON_REQUIRE((P->variant == 2 || P->variant == 3) && P->variant == plans[0]->variant,
                       "bp batch: every plan must use the same z-column kernel (variant 2 or 3)");
        ARECON_REQUIRE(P->nz == plans[0]->nz && P->group == plans[0]->group &&
                           P->device == plans[0]->device,
                       "bp batch: plans must share nz, group and device");
        SlabDesc d{};
        d.A = launch_args(P, imgs[i], fp_layout, dtype_size(fp_dtype));
        if (P->variant == 3) {
            ARECON_REQUIRE(P->nz == 8 || P->nz == 16, "bp batch: variant 3 needs nz 8 or 16");
            int cur = 0;
            ARECON_HIP(hipGetDevice(&cur));
            ARECON_HIP(hipSetDevice(P->device));
            const BPPlan::ZsTable* t = nullptr;
            try {
                t = P->nz == 8 ? &batch_zs_table_nz<8>(P, fp_dtype, fp_layout)
                               : &batch_zs_table_nz<16>(P, fp_dtype, fp_layout);
            } catch (...) {
                (void)hipSetDevice(cur);
                throw;
            }
            ARECON_HIP(hipSetDevice(cur));
            d.A.zs_tab = t->tab;
            d.A.zs_boff = t->boff;
            d.A.zs_fits = t->fits;
        }
        d.img = imgs[i];
        d.fp = fps[i];
        d.vtab = P->vtab;
        d.blocks = P->zblocks;
        d.masks = P->zmasks;
        d.table = P->vt_table;
        d.wg_begin = (int)wg;
        wg += P->grid;
        ARECON_REQUIRE(wg < (1LL << 31), "bp batch: too many workgroups for one launch");
        slabs[i] = d;
    }
    BPBatch* B = new BPBatch();
    B->device = plans[0]->device;
    B->variant = plans[0]->variant;
    B->inside = 1;
    for (int i = 0; i < n; ++i) B->inside &= plans[i]->a.inside;
    B->nz = plans[0]->nz;
    B->group = plans[0]->group;
    B->dtype = fp_dtype;
    B->layout = fp_layout;
    B->nslabs = n;
    B->grid = wg;
    int cur = 0;
    ARECON_HIP(hipGetDevice(&cur));
    try {
        ARECON_HIP(hipSetDevice(B->device));
        ARECON_HIP(hipMalloc(&B->d_slabs, sizeof(SlabDesc) * slabs.size()));
        ARECON_HIP(hipMemcpy(B->d_slabs, slabs.data(), sizeof(SlabDesc) * slabs.size(),
                             hipMemcpyHostToDevice));
        ARECON_HIP(hipSetDevice(cur));
    } catch (...) {
        (void)hipSetDevice(cur);
        bp_batch_destroy(B);
        throw;
    }
    return B;
}

void bp_batch_destroy(BPBatch* B) {
    if (!B) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(B->device);
    if (B->d_slabs) (void)hipFree(B->d_slabs);
    (void)hipSetDevice(cur);
    delete B;
}

template <int NZ, int G, class T>
static void batch_launch_t(const BPBatch* B, hipStream_t s) {
    const dim3 grid((unsigned)B->grid), block(256);
    switch (B->layout) {
    case ARECON_FP_ROWS:
        hipLaunchKernelGGL((bp_fdk3dw_zc_batch<NZ, G, T, ARECON_FP_ROWS>), grid, block, 0, s,
                           B->d_slabs, B->nslabs);
        break;
    case ARECON_FP_ROWPAIRS:
        hipLaunchKernelGGL((bp_fdk3dw_zc_batch<NZ, G, T, ARECON_FP_ROWPAIRS>), grid, block, 0,
                           s, B->d_slabs, B->nslabs);
        break;
    case ARECON_FP_ROWDIFF:
        hipLaunchKernelGGL((bp_fdk3dw_zc_batch<NZ, G, T, ARECON_FP_ROWDIFF>), grid, block, 0, s,
                           B->d_slabs, B->nslabs);
        break;
    default:
        hipLaunchKernelGGL((bp_fdk3dw_zc_batch<NZ, G, T, ARECON_FP_QUADS>), grid, block, 0, s,
                           B->d_slabs, B->nslabs);
        break;
    }
}

template <int NZ, int G>
static void batch_launch_g(const BPBatch* B, hipStream_t s) {
    switch (B->dtype) {
    case ARECON_F32: batch_launch_t<NZ, G, float>(B, s); break;
    case ARECON_F16: batch_launch_t<NZ, G, __half>(B, s); break;
    default: batch_launch_t<NZ, G, __hip_bfloat16>(B, s); break;
    }
}

template <int NZ>
static void batch_launch_nz(const BPBatch* B, hipStream_t s) {
    switch (B->group) {
    case 1: batch_launch_g<NZ, 1>(B, s); break;
    case 2: batch_launch_g<NZ, 2>(B, s); break;
    case 4: batch_launch_g<NZ, 4>(B, s); break;
    default:
        if constexpr (NZ <= 16) batch_launch_g<NZ, 0>(B, s);
        else fail(ARECON_ERR_INTERNAL, "bp batch: split mode needs nz <= 16");
        break;
    }
}

template <int NZ, class T, bool INSIDE>
static void batch_launch_zs_in(const BPBatch* B, hipStream_t s) {
    const dim3 grid((unsigned)B->grid), block(256);
    constexpr int kD = ARECON_FP_ROWDIFF, kP = ARECON_FP_ROWPAIRS, kQ = ARECON_FP_QUADS;
    if (B->layout == ARECON_FP_ROWDIFF)
        hipLaunchKernelGGL((bp_fdk3dw_zs_batch<NZ, T, kD, zs_rmax<NZ, T, kD>(), INSIDE>), grid,
                           block, 0, s, B->d_slabs, B->nslabs);
    else if constexpr (std::is_same<T, float>::value)
        hipLaunchKernelGGL((bp_fdk3dw_zs_batch<NZ, T, kP, zs_rmax<NZ, T, kP>(), INSIDE>), grid,
                           block, 0, s, B->d_slabs, B->nslabs);
    else
        hipLaunchKernelGGL((bp_fdk3dw_zs_batch<NZ, T, kQ, zs_rmax<NZ, T, kQ>(), INSIDE>), grid,
                           block, 0, s, B->d_slabs, B->nslabs);
}
template <int NZ, class T>
static void batch_launch_zs(const BPBatch* B, hipStream_t s) {
    if (B->inside) batch_launch_zs_in<NZ, T, true>(B, s);
    else batch_launch_zs_in<NZ, T, false>(B, s);
}
template <int NZ>
static void batch_launch_zs_nz(const BPBatch* B, hipStream_t s) {
    switch (B->dtype) {
    case ARECON_F32: batch_launch_zs<NZ, float>(B, s); break;
    case ARECON_F16: batch_launch_zs<NZ, __half>(B, s); break;
    default: batch_launch_zs<NZ, __hip_bfloat16>(B, s); break;
    }
}

void bp_batch_launch(const BPBatch* B, hipStream_t s) {
    ARECON_REQUIRE(B != nullptr, "bp batch: null batch");
    if (B->variant == 3) {
        if (B->nz == 8) batch_launch_zs_nz<8>(B, s);
        else batch_launch_zs_nz<16>(B, s);
        ARECON_HIP(hipGetLastError());
        return;
    }
    switch (B->nz) {
    case 8: batch_launch_nz<8>(B, s); break;
    case 16: batch_launch_nz<16>(B, s); break;
    default: batch_launch_nz<32>(B, s); break;
    }
    ARECON_HIP(hipGetLastError());
}

void bp_launch_f64(double* img, const double* fp, const arecon_bp_geom& g, const double* w,
                   hipStream_t s) {
    validate_geom(g);
    ARECON_REQUIRE(img != nullptr && fp != nullptr && w != nullptr, "bp: null buffer");
    const long long n = (long long)g.XN * g.YN * g.ZN;
    const long long grid = (n + 255) / 256;
    ARECON_REQUIRE(grid < (1LL << 31), "bp: volume too large for one launch");
    hipLaunchKernelGGL(bp_fdk3dw_f64, dim3((unsigned)grid), dim3(256), 0, s, img, fp, w, g,
                       kInvTwoPi, kRefPi);
    ARECON_HIP(hipGetLastError());
}

}  // namespace arecon
