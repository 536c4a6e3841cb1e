// Cone/row preweighting fused with the ramp filter (gfx950).
//
// Reference behaviour (src/filtering.cu:187-229 and :86-175; CPU twin
// src/analyticalReconBigPatient.m:107-134): every detector row of the
// [ViewN][ZL][YL] array is multiplied by
//     SO^2 / sqrt(SO^2 (SO^2 + b^2) - b^2 t^2),  t = (j-PLC) dYL, b = (k-ZLC) dZL,
// and then linearly convolved with h(n) = -2 / (pi^2 (4n^2 - 1)) / dYL
// (src/filtering.cu:23).  The reference zero-pads each row to 2^ceil(log2 3YL)
// complex samples in a separate HBM buffer and runs batched C2C FFTs; the
// result equals the exact linear convolution restricted to the first YL outputs.
//
// Design.  Output j of a row only sees taps |n| <= YL-1, so any transform length
// N >= 2YL gives the same linear convolution; this kernel uses
// N = 2^ceil(log2 2YL) (2048 for YL = 888).  The tap sequence is real and even,
// so its spectrum H is real: two detector rows are packed as the real and
// imaginary parts of one complex row, and H * FFT(a + ib) transforms back to
// (a*h) + i(b*h) with no cross talk.  Per pair of rows the whole transform lives
// in LDS:
//   load (global, preweight fused) -> radix-4 Stockham passes (+1 radix-2 pass
//   when log2 N is odd) -> multiply by H/N and conjugate (fused into the last
//   forward pass) -> the same forward passes again (conj(FFT(conj(Y))) is the
//   inverse) -> store of the first YL outputs (fused into the last pass),
//   in FP32, FP16 or BF16.
// The first pass reads only the nonzero half of the padded row (N >= 2YL means
// inputs N/2..N-1 are zero).  Twiddles exp(-2 pi i m/N) and H are computed in
// FP64 on the host, cached per (device, YL, dYL), and staged in LDS once per
// workgroup; workgroups loop over row pairs.
// HBM roofline: 8 B per sample FP32 -> FP32 (6 B with a 16-bit output).
#include "arecon_internal.h"

#include <array>

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

namespace arecon {
namespace {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// A product that is never contracted into an FMA with the sum it feeds.
__device__ __forceinline__ float mul_nc(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
// Complex products with their FMAs spelled out: left to contraction, which of
// the two products of a component is fused depends on the code around the
// call, so the same transform inlined in two kernels could round differently.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -mul_nc(a.y, b.y)), fmaf(a.x, b.y, mul_nc(a.y, b.x)));
}

// Radix-4 DFT butterfly (forward sign).
__device__ __forceinline__ void bfly4(float2& v0, float2& v1, float2& v2, float2& v3) {
    const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
    const float2 d = csub(v1, v3);
    const float2 a3 = make_float2(d.y, -d.x);   // -i * d
    v0 = cadd(a0, a2);
    v1 = cadd(a1, a3);
    v2 = csub(a0, a2);
    v3 = csub(a1, a3);
}

template <class T> __device__ __forceinline__ T to_out(float v);
template <> __device__ __forceinline__ float to_out<float>(float v) { return v; }
template <> __device__ __forceinline__ __half to_out<__half>(float v) { return __float2half(v); }
template <> __device__ __forceinline__ __hip_bfloat16 to_out<__hip_bfloat16>(float v) {
    return __float2bfloat16(v);
}

// Destinations of the two filtered rows a, b of a transform.  Rows layout: row
// r of view v at ((v ZL + r) YL + c).  Row-pair layout ([ViewN][ZL-1][YL][2]):
// element (v, r, c) holds rows r and r+1 of channel c, so row k goes to slot 0
// of pair k (k < ZL-1) and to slot 1 of pair k-1 (k > 0).  Null pointers are
// uniform across the workgroup.
template <class OutT> struct RowOut {
    OutT* a0; OutT* a1; OutT* b0; OutT* b1;
    int stride;
    __device__ __forceinline__ void put(int d, float va, float vb) const {
        const int i = d * stride;
        if (a0) a0[i] = to_out<OutT>(va);
        if (a1) a1[i] = to_out<OutT>(va);
        if (b0) b0[i] = to_out<OutT>(vb);
        if (b1) b1[i] = to_out<OutT>(vb);
    }
};

template <class OutT>
__device__ __forceinline__ void row_dest(OutT* out, int layout, long long row, int YL, int ZL,
                                         OutT*& p0, OutT*& p1) {
    if (layout == ARECON_FP_ROWS) {
        p0 = out + row * YL;
        p1 = nullptr;
        return;
    }
    const long long v = row / ZL;
    const int k = (int)(row - v * ZL);
    OutT* base = out + v * (long long)(ZL - 1) * YL * 2;
    p0 = (k < ZL - 1) ? base + (long long)k * YL * 2 : nullptr;
    p1 = (k > 0) ? base + (long long)(k - 1) * YL * 2 + 1 : nullptr;
}

template <class OutT>
__device__ __forceinline__ RowOut<OutT> row_out(OutT* out, int layout, long long ra, bool has_b,
                                                int YL, int ZL) {
    RowOut<OutT> o;
    row_dest(out, layout, ra, YL, ZL, o.a0, o.a1);
    o.b0 = o.b1 = nullptr;
    if (has_b) row_dest(out, layout, ra + 1, YL, ZL, o.b0, o.b1);
    o.stride = (layout == ARECON_FP_ROWS) ? 1 : 2;
    return o;
}

// Epilogue of a pass: 0 = LDS store, 1 = LDS store of H[d]*conj(v),
// 2 = store of the first YL outputs of both rows to global memory.
template <int EPI, class OutT>
__device__ __forceinline__ void emit(float2* dst, int d, float2 v, const float* H,
                                     const RowOut<OutT>& o, int YL) {
    if constexpr (EPI == 0) {
        dst[d] = v;
    } else if constexpr (EPI == 1) {
        const float hs = H[d];
        dst[d] = make_float2(mul_nc(hs, v.x), mul_nc(-hs, v.y));
    } else {
        if (d < YL) o.put(d, v.x, -v.y);
    }
}

template <int EPI, class OutT>
__device__ __forceinline__ void pass4(const float2* src, float2* dst, int N, int Ns,
                                      const float2* tw, const float* H,
                                      const RowOut<OutT>& o, int YL) {
    const int q = N >> 2;
    const int tstride = N / (Ns * 4);
    for (int j = threadIdx.x; j < q; j += blockDim.x) {
        const int jm = j & (Ns - 1);
        float2 v0 = src[j], v1 = src[j + q], v2 = src[j + 2 * q], v3 = src[j + 3 * q];
        const int t = jm * tstride;
        v1 = cmul(v1, tw[t]);
        v2 = cmul(v2, tw[2 * t]);
        v3 = cmul(v3, tw[3 * t]);
        bfly4(v0, v1, v2, v3);
        const int d = (j - jm) * 4 + jm;
        emit<EPI, OutT>(dst, d, v0, H, o, YL);
        emit<EPI, OutT>(dst, d + Ns, v1, H, o, YL);
        emit<EPI, OutT>(dst, d + 2 * Ns, v2, H, o, YL);
        emit<EPI, OutT>(dst, d + 3 * Ns, v3, H, o, YL);
    }
}

// Final radix-2 pass (Ns = N/2).
template <int EPI, class OutT>
__device__ __forceinline__ void pass2_last(const float2* src, float2* dst, int N,
                                           const float2* tw, const float* H,
                                           const RowOut<OutT>& o, int YL) {
    const int hN = N >> 1;
    for (int j = threadIdx.x; j < hN; j += blockDim.x) {
        const float2 v0 = src[j];
        const float2 v1 = cmul(src[j + hN], tw[j]);
        emit<EPI, OutT>(dst, j, cadd(v0, v1), H, o, YL);
        emit<EPI, OutT>(dst, j + hN, csub(v0, v1), H, o, YL);
    }
}

// Passes 2.. of one transform (the first pass has been done into `cur`).
// Returns the buffer holding the result when EPI_LAST writes to LDS.
template <int EPI_LAST, class OutT>
__device__ __forceinline__ float2* passes_after_first(float2* cur, float2* nxt, int N,
                                                      const float2* tw, const float* H,
                                                      const RowOut<OutT>& o, int YL) {
    int Ns = 4;
    while (Ns * 4 <= N) {
        const bool last = (Ns * 4 == N);
        __syncthreads();
        if (last)
            pass4<EPI_LAST, OutT>(cur, nxt, N, Ns, tw, H, o, YL);
        else
            pass4<0, OutT>(cur, nxt, N, Ns, tw, H, o, YL);
        float2* t = cur; cur = nxt; nxt = t;
        Ns *= 4;
    }
    if (Ns < N) {   // N = 2 * 4^k
        __syncthreads();
        pass2_last<EPI_LAST, OutT>(cur, nxt, N, tw, H, o, YL);
        float2* t = cur; cur = nxt; nxt = t;
    }
    return cur;
}

__device__ __forceinline__ float preweight(int j, float b, float PLC, float dYL, float SO) {
    // src/filtering.cu:202-204 (FP32, as the reference's GPU path computes it)
    const float t = ((float)j - PLC) * dYL;
    return SO * SO / sqrtf(SO * SO * (SO * SO + b * b) - b * t * b * t);
}

template <class OutT>
__global__ __launch_bounds__(256) void filter_rows(
    OutT* out, int layout, const float* in, const float2* __restrict__ twg,
    const float* __restrict__ Hg, int YL, int ZL, long long nrows, int N, float PLC,
    float ZLC, float dYL, float dZL, float SO) {
    extern __shared__ float2 smem[];
    float2* bufA = smem;
    float2* bufB = smem + N;
    float2* tw = smem + 2 * N;
    float* H = reinterpret_cast<float*>(smem + 3 * N);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        tw[i] = twg[i];
        H[i] = Hg[i];
    }
    const long long npairs = (nrows + 1) >> 1;
    const int q = N >> 2;
    for (long long pair = blockIdx.x; pair < npairs; pair += gridDim.x) {
        const long long ra = 2 * pair;
        const bool has_b = (ra + 1) < nrows;
        const float* ia = in + ra * YL;
        const float* ib = ia + YL;
        const int ka = (int)(ra % ZL);
        const int kb = (ka + 1 == ZL) ? 0 : ka + 1;
        const float ba = ((float)ka - ZLC) * dZL;
        const float bb = ((float)kb - ZLC) * dZL;
        __syncthreads();   // previous pair's readers of bufA/bufB are done
        // First forward pass (Ns = 1) straight from global memory.  Inputs
        // n >= N/2 >= YL are zero, so v2 = v3 = 0.
        for (int j = threadIdx.x; j < q; j += blockDim.x) {
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            if (j < YL) {
                v0.x = mul_nc(ia[j], preweight(j, ba, PLC, dYL, SO));
                if (has_b) v0.y = mul_nc(ib[j], preweight(j, bb, PLC, dYL, SO));
            }
            const int j1 = j + q;
            if (j1 < YL) {
                v1.x = mul_nc(ia[j1], preweight(j1, ba, PLC, dYL, SO));
                if (has_b) v1.y = mul_nc(ib[j1], preweight(j1, bb, PLC, dYL, SO));
            }
            const float2 nv1 = make_float2(v1.y, -v1.x);   // -i * v1
            bufA[4 * j + 0] = cadd(v0, v1);
            bufA[4 * j + 1] = cadd(v0, nv1);
            bufA[4 * j + 2] = csub(v0, v1);
            bufA[4 * j + 3] = csub(v0, nv1);
        }
        const RowOut<OutT> o = row_out(out, layout, ra, has_b, YL, ZL);
        float2* spec = passes_after_first<1, OutT>(bufA, bufB, N, tw, H, o, YL);
        float2* other = (spec == bufA) ? bufB : bufA;
        // Inverse: forward passes over conj(H X / N); first pass from LDS.
        __syncthreads();
        for (int j = threadIdx.x; j < q; j += blockDim.x) {
            float2 v0 = spec[j], v1 = spec[j + q], v2 = spec[j + 2 * q], v3 = spec[j + 3 * q];
            bfly4(v0, v1, v2, v3);
            other[4 * j + 0] = v0;
            other[4 * j + 1] = v1;
            other[4 * j + 2] = v2;
            other[4 * j + 3] = v3;
        }
        passes_after_first<2, OutT>(other, spec, N, tw, H, o, YL);
    }
}

// ---------------------------------------------------------------------------
// Wavefront-per-transform kernel (N = 256 .. 2048).  Each thread keeps N/T
// points in registers (T = 64, or 128 at N = 2048); a pass is a register DFT of
// radix 16, 8 or 4 (Stockham order), and LDS only carries the data between
// passes, in place, in a workgroup of one or two wavefronts.  Twiddles of a
// butterfly come from one table read w^1 and a short power recurrence; the
// spectrum H is read through L1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cscale(float2 a, float c, float s) {   // a * (c + i s)
    return make_float2(fmaf(a.x, c, -mul_nc(a.y, s)), fmaf(a.x, s, mul_nc(a.y, c)));
}

template <int R> __device__ __forceinline__ void dft(float2* v);
template <> __device__ __forceinline__ void dft<4>(float2* v) { bfly4(v[0], v[1], v[2], v[3]); }
template <> __device__ __forceinline__ void dft<8>(float2* v) {
    // n = n2 + 2 n1, k = k1 + 4 k2: X[k1 + 4k2] = y0[k1] + (-1)^k2 W8^k1 y1[k1]
    float2 a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
    float2 b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
    bfly4(a0, a1, a2, a3);
    bfly4(b0, b1, b2, b3);
    const float h = 0.70710678118654752f;
    b1 = cscale(b1, h, -h);
    b2 = make_float2(b2.y, -b2.x);                        // * -i
    b3 = cscale(b3, -h, -h);
    v[0] = cadd(a0, b0); v[4] = csub(a0, b0);
    v[1] = cadd(a1, b1); v[5] = csub(a1, b1);
    v[2] = cadd(a2, b2); v[6] = csub(a2, b2);
    v[3] = cadd(a3, b3); v[7] = csub(a3, b3);
}
template <> __device__ __forceinline__ void dft<16>(float2* v) {
    // n = n2 + 4 n1, k = k1 + 4 k2: inner DFT4 over n1, twiddle W16^(n2 k1), outer DFT4 over n2
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 a0 = v[n2], a1 = v[n2 + 4], a2 = v[n2 + 8], a3 = v[n2 + 12];
        bfly4(a0, a1, a2, a3);
        y[n2][0] = a0; y[n2][1] = a1; y[n2][2] = a2; y[n2][3] = a3;
    }
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
    // W16^m = cos(pi m/8) - i sin(pi m/8), m = n2 * k1
    y[1][1] = cscale(y[1][1], c1, -s1);
    y[1][2] = cscale(y[1][2], h, -h);
    y[1][3] = cscale(y[1][3], s1, -c1);
    y[2][1] = cscale(y[2][1], h, -h);
    y[2][2] = make_float2(y[2][2].y, -y[2][2].x);
    y[2][3] = cscale(y[2][3], -h, -h);
    y[3][1] = cscale(y[3][1], s1, -c1);
    y[3][2] = cscale(y[3][2], -h, -h);
    y[3][3] = cscale(y[3][3], -c1, s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
        bfly4(b0, b1, b2, b3);
        v[k1] = b0; v[k1 + 4] = b1; v[k1 + 8] = b2; v[k1 + 12] = b3;
    }
}

// v[r] *= w^r for r = 1 .. R-1 (powers by repeated squaring/products).
template <int R> __device__ __forceinline__ void twiddle(float2* v, float2 w1) {
    float2 wp[R];
    wp[1] = w1;
#pragma unroll
    for (int r = 2; r < R; ++r) wp[r] = (r & 1) ? cmul(wp[r - 1], w1) : cmul(wp[r / 2], wp[r / 2]);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], wp[r]);
}

// Pass modes: first forward pass reads global rows (preweighted); last forward
// pass multiplies by H and conjugates; last inverse pass stores rows, to global
// memory or, as two real rows, to LDS (floats [0, YL) and [N, N + YL)).
// SRC_STAGED: the fused filter's next-pair buffer in dynamic LDS (fstage): rows
// a and b at [0, YL) and [YL, 2 YL), their preweights at [2 YL, 3 YL) and
// [3 YL, 4 YL), copied there by fstage_copy while the previous pair ran.
enum { SRC_LDS = 0, SRC_GLOBAL = 1, SRC_STAGED = 2 };
extern __shared__ __attribute__((aligned(16))) float fstage[];
// DST_ROWDIFF: the fused filter's row differences straight from the last
// inverse pass (RowCtx's rowdiff fields).
enum { DST_LDS = 0, DST_LDS_SPEC = 1, DST_GLOBAL = 2, DST_LDS_ROWS = 3, DST_ROWDIFF = 4 };

struct RowCtx {
    const float* ia; const float* ib;      // input rows (ib valid when has_b)
    const float* wa; const float* wb;      // their preweights (table rows ka, kb)
    bool has_b;
    int YL;
    float PLC, dYL, SO;
    const float* H;
    // DST_ROWDIFF: the view's output (OutT), the pair's first row k, the
    // filtered row k - 1 (LDS, k > 0) and where row k + 1 goes for the next pair
    void* ov;
    int krow;
    const float* pr;
    float* pn;
};

// Two values as one 8-byte (FP32) or 4-byte (16-bit) store.
template <class OutT> struct Pair2;
template <> struct Pair2<float> {
    __device__ static __forceinline__ void put(float* o, float a, float b) {
        *reinterpret_cast<float2*>(o) = make_float2(a, b);
    }
};
template <> struct Pair2<__half> {
    __device__ static __forceinline__ void put(__half* o, float a, float b) {
        *reinterpret_cast<unsigned*>(o) = (unsigned)__half_as_ushort(__float2half(a)) |
                                          ((unsigned)__half_as_ushort(__float2half(b)) << 16);
    }
};
template <> struct Pair2<__hip_bfloat16> {
    __device__ static __forceinline__ void put(__hip_bfloat16* o, float a, float b) {
        auto bits = [](float v) { return (unsigned)__bfloat16_as_ushort(__float2bfloat16(v)); };
        *reinterpret_cast<unsigned*>(o) = bits(a) | (bits(b) << 16);
    }
};

// Threads per transform: one wavefront up to N = 1024.  At N = 2048, schedule
// S = 0 uses two wavefronts, one radix-16 butterfly per thread (16, 16, 8);
// S = 1 four wavefronts, one radix-8 butterfly per thread (8, 8, 8, 4): one
// pass more, but fewer registers per thread and so more wavefronts per SIMD.
template <int N, int S = 0> constexpr int tpt() {
    return N >= 2048 ? (S == 1 ? N / 8 : N / 16) : 64;
}

// LDS slot of element k of a transform: one float2 of padding after every 32.
// gfx950's LDS serves 64 banks of 4 bytes, so 32 lanes of 64-bit accesses per
// cycle, conflict-free when their slots differ mod 32.  Reads (consecutive k in
// aligned runs of 32) then never conflict, nor do the first pass's writes
// (k = R j + r: slot R j + r + j R / 32); the second radix-8 pass's writes
// (k = 64 q + j % 8 + 8 r) do.  Measured on the 888 x 64 x 11,808 sinogram, the
// 2048-point FP32 filter takes 2.83 ms against 2.97 ms with a pad after every 8
// (2-way conflicts on every read), 2.96 ms after every 64 and 3.49 ms unpadded.
// An LDS transform buffer holds N + N / 32 float2.
constexpr int kPadShift = 5;
__device__ __forceinline__ int lpad(int k) { return k + (k >> kPadShift); }
template <int N> constexpr int lds_fft_len() { return N + (N >> kPadShift); }

// WAITV: wait for this wavefront's outstanding memory operations (the staging
// copies of the next pair) before the pass's read/write barrier, which then
// makes them visible to the workgroup.
template <int N, int R, int Ns, int SRC, int DST, class OutT, int S, bool WAITV = false>
__device__ __forceinline__ void wave_pass(float2* lds, const float2* __restrict__ twg,
                                          const RowCtx& c, const RowOut<OutT>& o) {
    constexpr int T = tpt<N, S>();
    constexpr int NB = N / R, PER = NB / T;
    static_assert(PER >= 1 && PER * T == NB, "radix schedule does not fit the thread count");
    const int lane = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = lane + T * b;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = j + r * NB;
            if (SRC == SRC_LDS) {
                // lpad(n) = lpad(j) + lpad(r NB) (NB a multiple of the pad
                // period): one base register and immediate offsets
                static_assert(NB % (1 << kPadShift) == 0, "padded reads need NB % 32 == 0");
                v[b][r] = lds[lpad(j) + lpad(r * NB)];
            } else if (r < R / 2) {                          // n < N/2; n >= YL reads zero
                float2 x = make_float2(0.f, 0.f);
                if (n < c.YL) {
                    // preweighted samples, rounded before the butterflies add them
                    if (SRC == SRC_STAGED) {
                        x.x = mul_nc(fstage[n], fstage[2 * c.YL + n]);
                        if (c.has_b) x.y = mul_nc(fstage[c.YL + n], fstage[3 * c.YL + n]);
                    } else {
                        x.x = mul_nc(c.ia[n], c.wa[n]);
                        if (c.has_b) x.y = mul_nc(c.ib[n], c.wb[n]);
                    }
                }
                v[b][r] = x;
            } else {
                v[b][r] = make_float2(0.f, 0.f);
            }
        }
    }
    static_assert(!WAITV || (SRC == SRC_LDS && DST != DST_GLOBAL), "the wait needs the pass's barrier");
    if (WAITV) __builtin_amdgcn_s_waitcnt(0);
    if (SRC == SRC_LDS && DST != DST_GLOBAL) __syncthreads();   // all reads before in-place writes
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = lane + T * b;
        const int jm = j & (Ns - 1);
        if (Ns > 1) twiddle<R>(v[b], twg[jm * (N / (Ns * R))]);
        dft<R>(v[b]);
        const int d = (j - jm) * R + jm;
        // lpad(d + r Ns) = lpad(d) + lpad(r Ns) when the low kPadShift bits of d
        // and r Ns do not carry: Ns a multiple of the pad period P, or Ns = 1 with
        // d = R j and R a multiple or a divisor of P (r < R).  Then one base
        // register serves every r.
        constexpr int kP = 1 << kPadShift;
        constexpr bool kSplit = Ns % kP == 0 || (Ns == 1 && (R % kP == 0 || kP % R == 0));
        const int dp = lpad(d);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = d + r * Ns;
            const int kp = kSplit ? dp + lpad(r * Ns) : lpad(k);
            if (DST == DST_LDS) {
                lds[kp] = v[b][r];
            } else if (DST == DST_LDS_SPEC) {
                const float hs = c.H[k];
                lds[kp] = make_float2(mul_nc(hs, v[b][r].x), mul_nc(-hs, v[b][r].y));
            } else if (DST == DST_LDS_ROWS) {   // unpadded rows (2 N floats)
                if (k < c.YL) {
                    float* rows = reinterpret_cast<float*>(lds);
                    rows[k] = v[b][r].x;
                    rows[N + k] = -v[b][r].y;
                }
            } else if (DST == DST_ROWDIFF) {    // rows krow, krow + 1 of channel k
                if (k < c.YL) {
                    const float a0 = v[b][r].x, b0 = -v[b][r].y;
                    OutT* ov = static_cast<OutT*>(c.ov);
                    if (c.krow > 0) {
                        const float p0 = c.pr[k];
                        Pair2<OutT>::put(ov + ((long long)(c.krow - 1) * c.YL + k) * 2, p0,
                                         a0 - p0);
                    }
                    if (c.has_b) {
                        c.pn[k] = b0;
                        Pair2<OutT>::put(ov + ((long long)c.krow * c.YL + k) * 2, a0, b0 - a0);
                    }
                }
            } else if (k < c.YL) {
                o.put(k, v[b][r].x, -v[b][r].y);
            }
        }
    }
    __syncthreads();
}

// The whole forward (MODE 0: global rows -> spectrum in LDS) or inverse
// (MODE 1: LDS -> global rows; MODE 2: LDS -> the two real rows in LDS; MODE 3:
// LDS -> row differences in global memory) transform of one pair of rows.  PF (the fused filter's prefetch): the forward
// transform reads its rows from the staging buffer and calls after_first() once
// its first pass is done with them; the inverse transform's first pass waits
// for the staging copies before its barrier.
template <int N, int MODE, class OutT, int S, bool PF = false, class F = void (*)()>
__device__ __forceinline__ void wave_fft(float2* lds, const float2* __restrict__ twg,
                                         const RowCtx& c, const RowOut<OutT>& o,
                                         F&& after_first = [] {}) {
    constexpr int S0 = (MODE == 0) ? (PF ? SRC_STAGED : SRC_GLOBAL) : SRC_LDS;
    constexpr int DL = MODE == 0 ? DST_LDS_SPEC
                       : MODE == 1 ? DST_GLOBAL
                       : MODE == 2 ? DST_LDS_ROWS : DST_ROWDIFF;
    constexpr bool W0 = PF && MODE >= 2;
    if constexpr (N == 2048 && S == 1) {
        wave_pass<N, 8, 1, S0, DST_LDS, OutT, S, W0>(lds, twg, c, o);
        if constexpr (MODE == 0) after_first();
        wave_pass<N, 8, 8, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 8, 64, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 4, 512, SRC_LDS, DL, OutT, S>(lds, twg, c, o);
    } else if constexpr (N == 2048) {
        wave_pass<N, 16, 1, S0, DST_LDS, OutT, S, W0>(lds, twg, c, o);
        if constexpr (MODE == 0) after_first();
        wave_pass<N, 16, 16, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 8, 256, SRC_LDS, DL, OutT, S>(lds, twg, c, o);
    } else if constexpr (N == 1024) {
        wave_pass<N, 16, 1, S0, DST_LDS, OutT, S, W0>(lds, twg, c, o);
        if constexpr (MODE == 0) after_first();
        wave_pass<N, 16, 16, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 4, 256, SRC_LDS, DL, OutT, S>(lds, twg, c, o);
    } else if constexpr (N == 512) {
        wave_pass<N, 8, 1, S0, DST_LDS, OutT, S, W0>(lds, twg, c, o);
        if constexpr (MODE == 0) after_first();
        wave_pass<N, 8, 8, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 8, 64, SRC_LDS, DL, OutT, S>(lds, twg, c, o);
    } else {
        static_assert(N == 256, "wave FFT supports N = 256 .. 2048");
        wave_pass<N, 4, 1, S0, DST_LDS, OutT, S, W0>(lds, twg, c, o);
        if constexpr (MODE == 0) after_first();
        wave_pass<N, 4, 4, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 4, 16, SRC_LDS, DST_LDS, OutT, S>(lds, twg, c, o);
        wave_pass<N, 4, 64, SRC_LDS, DL, OutT, S>(lds, twg, c, o);
    }
}

// A pair's forward transform (its first pass then after_first()) and inverse
// transform MODE_INV (2 or 3 of wave_fft).  (Fusing the forward's last pass
// with the inverse's first, whose elements are the same threads' at N = 2048,
// saves an LDS round trip and two barriers per pair but measured slower: 2.53
// against 2.35 ms on the clinical sinogram.)
template <int N, int MODE_INV, class OutT, int S, bool PF, class F>
__device__ __forceinline__ void wave_fft_pair(float2* lds, const float2* __restrict__ twg,
                                              const RowCtx& c, const RowOut<OutT>& o,
                                              F&& after_first) {
    wave_fft<N, 0, OutT, S, PF>(lds, twg, c, o, after_first);
    wave_fft<N, MODE_INV, OutT, S, PF>(lds, twg, c, o);
}

// Copies rows k, k + 1 of a view (`rows` points at row k) and their preweights
// (`w` at row k of the preweight table) into fstage with direct-to-LDS 16-byte
// buffer loads: no registers, no wait here.  A row of YL floats (YL % 4 == 0,
// 16-byte aligned) is ceil(YL / 256) loads; the tail's lanes are switched off,
// so nothing lands past the row.  Row k + 1 and its weights only when has_b.
template <int T>
__device__ __forceinline__ void fstage_copy(const float* rows, const float* w, bool has_b,
                                            int YL) {
    typedef __attribute__((address_space(3))) float lds_float;
    typedef __attribute__((address_space(3))) void lds_void;
    constexpr int NW = T / 64;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned rowb = 4u * (unsigned)YL;
    for (int r = wave; r < 4; r += NW) {       // 0, 1: rows k, k + 1; 2, 3: their weights
        if ((r & 1) && !has_b) continue;
        const float* src = (r < 2 ? rows : w) + (r & 1) * YL;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(src), (short)0, (int)rowb, 0x00020000);
        lds_float* dst = (lds_float*)fstage + r * YL;
        for (unsigned off = 0; off < rowb; off += 1024u)
            if (off + 16u * lane < rowb)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void*)(dst + off / 4u), 16,
                                                        off + 16u * lane, 0, 0, 0);
    }
}

template <int N, class OutT, int S>
__global__ __launch_bounds__((tpt<N, S>())) void filter_rows_wave(
    OutT* out, int layout, const float* in, const float2* __restrict__ twg,
    const float* __restrict__ Hg, const float* __restrict__ Wg, int YL, int ZL,
    long long nrows, float PLC, float ZLC, float dYL, float dZL, float SO) {
    __shared__ float2 lds[lds_fft_len<N>()];
    const long long npairs = (nrows + 1) >> 1;
    for (long long pair = blockIdx.x; pair < npairs; pair += gridDim.x) {
        const long long ra = 2 * pair;
        RowCtx c;
        c.has_b = (ra + 1) < nrows;
        c.ia = in + ra * YL;
        c.ib = c.ia + YL;
        const int ka = (int)(ra % ZL);
        const int kb = (ka + 1 == ZL) ? 0 : ka + 1;
        c.wa = Wg + (long long)ka * YL;
        c.wb = Wg + (long long)kb * YL;
        c.YL = YL; c.PLC = PLC; c.dYL = dYL; c.SO = SO; c.H = Hg;
        const RowOut<OutT> o = row_out(out, layout, ra, c.has_b, YL, ZL);
        wave_fft<N, 0, OutT, S>(lds, twg, c, o);
        wave_fft<N, 1, OutT, S>(lds, twg, c, o);
    }
}

struct FilterTables {
    float2* tw = nullptr;
    float* H = nullptr;
    int N = 0;
};

std::mutex g_filter_mu;
std::map<std::tuple<int, int, uint64_t>, FilterTables> g_filter_cache;

int filter_length(int YL) {
    int N = 64;
    while (N < 2 * YL) N <<= 1;
    return N;
}

// Twiddles and the real spectrum of the tap sequence, FP64 on the host.
const FilterTables& filter_tables(int device, int YL, float dYL_f) {
    const double dYL = (double)dYL_f;
    uint64_t key_bits;
    std::memcpy(&key_bits, &dYL, sizeof(key_bits));
    std::lock_guard<std::mutex> lk(g_filter_mu);
    auto key = std::make_tuple(device, YL, key_bits);
    auto it = g_filter_cache.find(key);
    if (it != g_filter_cache.end()) return it->second;
    const int N = filter_length(YL);
    std::vector<double> c(N), s(N);
    for (int m = 0; m < N; ++m) {
        c[m] = std::cos(2.0 * M_PI * m / N);
        s[m] = std::sin(2.0 * M_PI * m / N);
    }
    std::vector<float2> tw(N);
    for (int m = 0; m < N; ++m) tw[m] = make_float2((float)c[m], (float)-s[m]);
    // h(n), src/filtering.cu:23 (pi^2 literal of the reference), |n| <= YL-1.
    std::vector<double> h(YL);
    for (int n = 0; n < YL; ++n)
        h[n] = -2.0 / (9.869604401089358 * (4.0 * n * n - 1.0)) / dYL;
    std::vector<float> H(N);
    for (int f = 0; f < N; ++f) {
        double acc = h[0];
        long long idx = 0;
        for (int n = 1; n < YL; ++n) {
            idx += f;
            if (idx >= N) idx -= N;
            acc += 2.0 * h[n] * c[idx];
        }
        H[f] = (float)(acc / N);
    }
    FilterTables t;
    t.N = N;
    ARECON_HIP(hipMalloc(&t.tw, sizeof(float2) * N));
    ARECON_HIP(hipMalloc(&t.H, sizeof(float) * N));
    ARECON_HIP(hipMemcpy(t.tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
    ARECON_HIP(hipMemcpy(t.H, H.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    return g_filter_cache.emplace(key, t).first->second;
}

// Cone and row preweight SO^2 / sqrt(SO^2 (SO^2 + b^2) - b^2 t^2) with
// t = (j - PLC) dYL and b = (k - ZLC) dZL (src/filtering.cu:202-205), for every
// row k and channel j: [ZL][YL], FP64 on the host, rounded to FP32.  The
// reference's GPU path evaluates it per sample in FP32; its CPU twin
// (src/analyticalReconBigPatient.m:109-115) in FP64.  Like the twiddle tables,
// these are kept for the life of the process (captured graphs hold their
// addresses): 4 bytes per detector sample of each geometry.
std::mutex g_pw_mu;
std::map<std::array<uint32_t, 8>, float*> g_pw_cache;

uint32_t fbits(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
}

const float* preweight_table(int device, int YL, int ZL, float PLC, float ZLC, float dYL,
                             float dZL, float SO) {
    const std::array<uint32_t, 8> key{(uint32_t)device, (uint32_t)YL, (uint32_t)ZL,
                                      fbits(PLC), fbits(ZLC), fbits(dYL), fbits(dZL),
                                      fbits(SO)};
    std::lock_guard<std::mutex> lk(g_pw_mu);
    auto it = g_pw_cache.find(key);
    if (it != g_pw_cache.end()) return it->second;
    std::vector<float> w((size_t)ZL * YL);
    const double so2 = (double)SO * SO;
    for (int k = 0; k < ZL; ++k) {
        const double b = ((double)k - ZLC) * dZL;
        for (int j = 0; j < YL; ++j) {
            const double t = ((double)j - PLC) * dYL;
            w[(size_t)k * YL + j] = (float)(so2 / std::sqrt(so2 * (so2 + b * b) - b * b * t * t));
        }
    }
    float* d = nullptr;
    ARECON_HIP(hipMalloc(&d, sizeof(float) * w.size()));
    ARECON_HIP(hipMemcpy(d, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    g_pw_cache.emplace(key, d);
    return d;
}

}  // namespace

// Difference quads from filtered FP32 rows: element (v, r, c) of
// [ViewN][ZL-1][YL-1][4] holds a = f(r,c), f(r+1,c)-a, f(r,c+1)-a and
// (f(r+1,c+1)-f(r,c+1)) - (f(r+1,c)-a).  One thread per element; consecutive
// threads walk the channels, so reads and writes are coalesced.
template <class OutT> struct Quad4;
template <> struct Quad4<float> {
    __device__ static __forceinline__ void put(float* o, float a, float b, float c, float d) {
        *reinterpret_cast<float4*>(o) = make_float4(a, b, c, d);
    }
};
template <> struct Quad4<__half> {
    __device__ static __forceinline__ void put(__half* o, float a, float b, float c, float d) {
        const unsigned lo = (unsigned)__half_as_ushort(__float2half(a)) |
                            ((unsigned)__half_as_ushort(__float2half(b)) << 16);
        const unsigned hi = (unsigned)__half_as_ushort(__float2half(c)) |
                            ((unsigned)__half_as_ushort(__float2half(d)) << 16);
        *reinterpret_cast<uint2*>(o) = make_uint2(lo, hi);
    }
};
template <> struct Quad4<__hip_bfloat16> {
    __device__ static __forceinline__ void put(__hip_bfloat16* o, float a, float b, float c,
                                               float d) {
        auto bits = [](float v) { return (unsigned)__bfloat16_as_ushort(__float2bfloat16(v)); };
        *reinterpret_cast<uint2*>(o) =
            make_uint2(bits(a) | (bits(b) << 16), bits(c) | (bits(d) << 16));
    }
};

// The two values of a row-difference element, f(r, c) and f(r+1, c) - f(r, c),
// in one store (8 bytes in FP32, 4 in FP16 / BF16).
template <class OutT>
__global__ __launch_bounds__(256) void expand_quads(OutT* __restrict__ out,
                                                    const float* __restrict__ rows, int YL,
                                                    int ZL, long long n) {
    const int nc = YL - 1, nr = ZL - 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long t = i / nc;
        const int c = (int)(i - t * nc);
        const long long v = t / nr;
        const int r = (int)(t - v * nr);
        const float* p = rows + (v * ZL + r) * YL + c;
        const float a = p[0], ac = p[1], b = p[YL], bc = p[YL + 1];
        const float dr = b - a, dc = ac - a;
        Quad4<OutT>::put(out + 4 * i, a, dr, dc, (bc - ac) - dr);
    }
}

// Filter and difference quads (OUTL = ARECON_FP_QUADS) or row differences
// (ARECON_FP_ROWDIFF, FP32) in one pass (N = 256 .. 2048).  A workgroup takes
// whole views and walks their rows in pairs (k, k+1): the inverse transform
// leaves both filtered rows in LDS, where they give output rows k-1 (with row
// k-1, kept from the previous pair) and k.  The differences are the same FP32
// ones that expand_quads / expand_rowdiff form from stored rows, rounded once to
// OutT, so the output is identical; the FP32 rows never reach HBM.
// PF (YL % 4 == 0): the next pair's rows and preweights are copied into the
// dynamic LDS buffer fstage (4 YL floats) while this pair's transforms run, and
// the first pass reads them from there: its global loads (rows from HBM,
// preweights from L2) otherwise wait once per pair.
template <int N, class OutT, int S, int OUTL, bool PF>
__global__ __launch_bounds__((tpt<N, S>())) void filter_quads_wave(
    OutT* __restrict__ out, const float* in, const float2* __restrict__ twg,
    const float* __restrict__ Hg, const float* __restrict__ Wg, int YL, int ZL, int ViewN,
    float PLC, float dYL, float SO) {
    constexpr int T = tpt<N, S>();
    __shared__ float2 lds[lds_fft_len<N>()];
    __shared__ float prev[2][N / 2];             // filtered row k-1, double-buffered
    const float* la = reinterpret_cast<const float*>(lds);
    const float* lb = la + N;
    const int nc = YL - 1;
    RowOut<OutT> o{};                            // the rows stay in LDS
    for (int v = blockIdx.x; v < ViewN; v += gridDim.x) {
        OutT* ov = out + (long long)v * (ZL - 1) * (OUTL == ARECON_FP_QUADS ? nc * 4 : YL * 2);
        if constexpr (PF) {
            fstage_copy<T>(in + (long long)v * ZL * YL, Wg, true, YL);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        for (int k = 0; k < ZL; k += 2) {
            RowCtx c;
            c.has_b = k + 1 < ZL;
            c.ia = in + ((long long)v * ZL + k) * YL;
            c.ib = c.ia + YL;
            c.wa = Wg + (long long)k * YL;
            c.wb = Wg + (long long)(c.has_b ? k + 1 : 0) * YL;
            c.YL = YL; c.PLC = PLC; c.dYL = dYL; c.SO = SO; c.H = Hg;
            const float* pr = prev[(k >> 1) & 1];
            float* pn = prev[((k >> 1) & 1) ^ 1];
            // Row differences come from the last inverse pass: rows k and k + 1
            // of a channel are one thread's, row k - 1 comes from prev, and row
            // k + 1 goes to the other prev buffer for the next pair (which reads
            // it only after the barriers of its transforms).  The next pair's
            // copies start once the first pass has read this pair's (its barrier
            // is behind).
            c.ov = ov; c.krow = k; c.pr = pr; c.pn = pn;
            wave_fft_pair<N, OUTL == ARECON_FP_ROWDIFF ? 3 : 2, OutT, S, PF>(
                lds, twg, c, o, [&] {
                    if constexpr (PF)
                        if (k + 2 < ZL)
                            fstage_copy<T>(in + ((long long)v * ZL + k + 2) * YL,
                                           Wg + (long long)(k + 2) * YL, k + 3 < ZL, YL);
                });
            if constexpr (OUTL == ARECON_FP_ROWDIFF) continue;
            // rows k (la) and k+1 (lb) are in LDS, row k-1 in prev[(k/2) & 1]
            // (k > 0); row k+1 goes to the other buffer for the next pair, which
            // reads it only after the barriers of its transforms
            for (int ch = threadIdx.x; ch < YL; ch += T) {
                const float a0 = la[ch];
                const float b0 = c.has_b ? lb[ch] : 0.f;
                if (c.has_b) pn[ch] = b0;
                if (ch >= nc) continue;
                const float a1 = la[ch + 1];
                if (k > 0) {
                    const float p0 = pr[ch], p1 = pr[ch + 1];
                    const float dr = a0 - p0;
                    Quad4<OutT>::put(ov + ((long long)(k - 1) * nc + ch) * 4, p0, dr, p1 - p0,
                                     (a1 - p1) - dr);
                }
                if (c.has_b) {
                    const float b1 = lb[ch + 1];
                    const float dr = b0 - a0;
                    Quad4<OutT>::put(ov + ((long long)k * nc + ch) * 4, a0, dr, a1 - a0,
                                     (b1 - a1) - dr);
                }
            }
            __syncthreads();     // every LDS read of this pair before the next writes
        }
    }
}

// FFT schedule of the 2048-point kernels (see tpt): 1 (radix 8, four
// wavefronts; 3.05 ms against 3.63 ms for the 888 x 64 x 11,808 sinogram) unless
// ARECON_FILTER_SCHEDULE=0 asks for the radix-16 schedule.  Both give
// the same result up to FP32 rounding of the passes.  Read at every launch.
static int filter_schedule() {
    const char* e = std::getenv("ARECON_FILTER_SCHEDULE");
    return (e && e[0] == '0') ? 0 : 1;
}

// Row differences from filtered FP32 rows: element (v, r, c) of
// [ViewN][ZL-1][YL][2] holds f(r,c) and f(r+1,c)-f(r,c).
template <class OutT>
__global__ __launch_bounds__(256) void expand_rowdiff(OutT* __restrict__ out,
                                                      const float* __restrict__ rows, int YL,
                                                      int ZL, long long n) {
    const int nr = ZL - 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long t = i / YL;
        const int c = (int)(i - t * YL);
        const long long v = t / nr;
        const int r = (int)(t - v * nr);
        const float* p = rows + (v * ZL + r) * YL + c;
        const float a = p[0];
        Pair2<OutT>::put(out + 2 * i, a, p[YL] - a);
    }
}

// The device tables a launch reads: twiddles and spectrum (N entries) and the
// preweight [ZL][YL] (needed for N <= 2048).  Either the per-device caches of
// filter_tables() / preweight_table(), or the scratch of
// arecon_filter_dev_noalloc.
struct TableSet {
    const float2* tw = nullptr;
    const float* H = nullptr;
    const float* W = nullptr;
    int N = 0;
};

static TableSet cached_tables(int dev, int YL, int ZL, float PLC, float ZLC, float dYL,
                              float dZL, float SO) {
    const FilterTables& t = filter_tables(dev, YL, dYL);
    TableSet ts;
    ts.tw = t.tw;
    ts.H = t.H;
    ts.N = t.N;
    if (t.N >= 256 && t.N <= 2048) ts.W = preweight_table(dev, YL, ZL, PLC, ZLC, dYL, dZL, SO);
    return ts;
}

static void launch_filter_rows(void* out, int out_dtype, int out_layout, const float* in,
                               int YL, int ZL, int ViewN, float PLC, float ZLC, float dYL,
                               float dZL, float SO, hipStream_t s, const TableSet* fixed);

// Launch of the fused per-view filter for layout OUTL (PF: rows staged in LDS
// a pair ahead, dynamic LDS of 4 YL floats).
template <int OUTL, class OT, bool PF>
static void launch_filter_fused_pf(OT* outp, const float* in, const TableSet& t, const float* W,
                                   int N, int YL, int ZL, int ViewN, float PLC, float dYL,
                                   float SO, hipStream_t s) {
    const dim3 grid((unsigned)ViewN);
    const unsigned dyn = PF ? 16u * (unsigned)YL : 0u;
    switch (N) {
    case 256: hipLaunchKernelGGL((filter_quads_wave<256, OT, 0, OUTL, PF>), grid, dim3(tpt<256>()),
                                 dyn, s, outp, in, t.tw, t.H, W, YL, ZL, ViewN, PLC, dYL, SO); break;
    case 512: hipLaunchKernelGGL((filter_quads_wave<512, OT, 0, OUTL, PF>), grid, dim3(tpt<512>()),
                                 dyn, s, outp, in, t.tw, t.H, W, YL, ZL, ViewN, PLC, dYL, SO); break;
    case 1024: hipLaunchKernelGGL((filter_quads_wave<1024, OT, 0, OUTL, PF>), grid,
                                  dim3(tpt<1024>()), dyn, s, outp, in, t.tw, t.H, W, YL, ZL, ViewN,
                                  PLC, dYL, SO); break;
    default:
        if (filter_schedule() == 1)
            hipLaunchKernelGGL((filter_quads_wave<2048, OT, 1, OUTL, PF>), grid,
                               dim3(tpt<2048, 1>()), dyn, s, outp, in, t.tw, t.H, W, YL, ZL,
                               ViewN, PLC, dYL, SO);
        else
            hipLaunchKernelGGL((filter_quads_wave<2048, OT, 0, OUTL, PF>), grid,
                               dim3(tpt<2048, 0>()), dyn, s, outp, in, t.tw, t.H, W, YL, ZL,
                               ViewN, PLC, dYL, SO);
        break;
    }
}

template <int OUTL, class OT>
static void launch_filter_fused(OT* outp, const float* in, const TableSet& t, const float* W,
                                int N, int YL, int ZL, int ViewN, float PLC, float dYL, float SO,
                                hipStream_t s) {
    // the staging copies move 16-byte pieces of 16-byte aligned rows
    const bool pf = YL % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(W) & 15) == 0;
    if (pf)
        launch_filter_fused_pf<OUTL, OT, true>(outp, in, t, W, N, YL, ZL, ViewN, PLC, dYL, SO, s);
    else
        launch_filter_fused_pf<OUTL, OT, false>(outp, in, t, W, N, YL, ZL, ViewN, PLC, dYL, SO, s);
}

// fixed: the tables to read (the allocation-free entry point's scratch), or
// null for the per-device caches.
static void launch_filter_impl(void* out, int out_dtype, int out_layout, const float* in,
                               int YL, int ZL, int ViewN, float PLC, float ZLC, float dYL,
                               float dZL, float SO, hipStream_t s, const TableSet* fixed) {
    if (out_layout != ARECON_FP_QUADS && out_layout != ARECON_FP_ROWDIFF) {
        launch_filter_rows(out, out_dtype, out_layout, in, YL, ZL, ViewN, PLC, ZLC, dYL, dZL,
                           SO, s, fixed);
        return;
    }
    const bool quads = out_layout == ARECON_FP_QUADS;
    ARECON_REQUIRE(YL >= 2 && ZL >= 2,
                   "filter: the quad and row-difference layouts need YL >= 2 and ZL >= 2");
    ARECON_REQUIRE(out != nullptr && out != (const void*)in,
                   "filter: the quad or row-difference output cannot alias the input");
    ARECON_REQUIRE(ViewN > 0, "filter: YL, ZL and ViewN must be positive");
    ARECON_REQUIRE(YL <= 2048, "filter: YL above 2048 is not supported");
    ARECON_REQUIRE(dYL != 0.f && SO > 0.f, "filter: dYL must be nonzero and SO positive");
    ARECON_REQUIRE(out_dtype == ARECON_F32 || out_dtype == ARECON_F16 || out_dtype == ARECON_BF16,
                   "filter: unknown output dtype");
    int dev = 0;
    ARECON_HIP(hipGetDevice(&dev));
    const int N = filter_length(YL);
    if (N >= 256 && N <= 2048) {      // fused: filtered rows stay in LDS
        const TableSet t = fixed ? *fixed : cached_tables(dev, YL, ZL, PLC, ZLC, dYL, dZL, SO);
        const float* W = t.W;
        if (!quads && out_dtype == ARECON_F32)
            launch_filter_fused<ARECON_FP_ROWDIFF>(static_cast<float*>(out), in, t, W, N, YL, ZL,
                                                   ViewN, PLC, dYL, SO, s);
        else if (!quads && out_dtype == ARECON_F16)
            launch_filter_fused<ARECON_FP_ROWDIFF>(static_cast<__half*>(out), in, t, W, N, YL,
                                                   ZL, ViewN, PLC, dYL, SO, s);
        else if (!quads)
            launch_filter_fused<ARECON_FP_ROWDIFF>(static_cast<__hip_bfloat16*>(out), in, t, W,
                                                   N, YL, ZL, ViewN, PLC, dYL, SO, s);
        else if (out_dtype == ARECON_F32)
            launch_filter_fused<ARECON_FP_QUADS>(static_cast<float*>(out), in, t, W, N, YL, ZL,
                                                 ViewN, PLC, dYL, SO, s);
        else if (out_dtype == ARECON_F16)
            launch_filter_fused<ARECON_FP_QUADS>(static_cast<__half*>(out), in, t, W, N, YL, ZL,
                                                 ViewN, PLC, dYL, SO, s);
        else
            launch_filter_fused<ARECON_FP_QUADS>(static_cast<__hip_bfloat16*>(out), in, t, W, N,
                                                 YL, ZL, ViewN, PLC, dYL, SO, s);
        ARECON_HIP(hipGetLastError());
        return;
    }
    // other transform lengths: filter in place, then derive the layout
    float* rows = const_cast<float*>(in);
    launch_filter_rows(rows, ARECON_F32, ARECON_FP_ROWS, in, YL, ZL, ViewN, PLC, ZLC, dYL, dZL,
                       SO, s, fixed);
    const long long n = (long long)ViewN * (ZL - 1) * (quads ? YL - 1 : YL);
    int cus = 0;
    ARECON_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const long long grid = std::max<long long>(1, std::min<long long>((n + 255) / 256,
                                                                      (long long)cus * 64));
    if (!quads) {
        if (out_dtype == ARECON_F32)
            hipLaunchKernelGGL(expand_rowdiff<float>, dim3((unsigned)grid), dim3(256), 0, s,
                               static_cast<float*>(out), rows, YL, ZL, n);
        else if (out_dtype == ARECON_F16)
            hipLaunchKernelGGL(expand_rowdiff<__half>, dim3((unsigned)grid), dim3(256), 0, s,
                               static_cast<__half*>(out), rows, YL, ZL, n);
        else
            hipLaunchKernelGGL(expand_rowdiff<__hip_bfloat16>, dim3((unsigned)grid), dim3(256),
                               0, s, static_cast<__hip_bfloat16*>(out), rows, YL, ZL, n);
        ARECON_HIP(hipGetLastError());
        return;
    }
    switch (out_dtype) {
    case ARECON_F32:
        hipLaunchKernelGGL(expand_quads<float>, dim3((unsigned)grid), dim3(256), 0, s,
                           static_cast<float*>(out), rows, YL, ZL, n);
        break;
    case ARECON_F16:
        hipLaunchKernelGGL(expand_quads<__half>, dim3((unsigned)grid), dim3(256), 0, s,
                           static_cast<__half*>(out), rows, YL, ZL, n);
        break;
    case ARECON_BF16:
        hipLaunchKernelGGL(expand_quads<__hip_bfloat16>, dim3((unsigned)grid), dim3(256), 0, s,
                           static_cast<__hip_bfloat16*>(out), rows, YL, ZL, n);
        break;
    default: fail(ARECON_ERR_ARG, "filter: unknown output dtype");
    }
    ARECON_HIP(hipGetLastError());
}

static void launch_filter_rows(void* out, int out_dtype, int out_layout, const float* in,
                               int YL, int ZL, int ViewN, float PLC, float ZLC, float dYL,
                               float dZL, float SO, hipStream_t s, const TableSet* fixed) {
    ARECON_REQUIRE(YL > 0 && ZL > 0 && ViewN > 0, "filter: YL, ZL and ViewN must be positive");
    ARECON_REQUIRE(YL <= 2048, "filter: YL above 2048 is not supported");
    ARECON_REQUIRE(out != nullptr && in != nullptr, "filter: null buffer");
    ARECON_REQUIRE(out_dtype == ARECON_F32 || out_dtype == ARECON_F16 || out_dtype == ARECON_BF16,
                   "filter: unknown output dtype");
    ARECON_REQUIRE(out_layout == ARECON_FP_ROWS || out_layout == ARECON_FP_ROWPAIRS,
                   "filter: unknown output layout");
    ARECON_REQUIRE((out_dtype == ARECON_F32 && out_layout == ARECON_FP_ROWS) ||
                       out != (const void*)in,
                   "filter: only the FP32 rows layout can be written in place");
    ARECON_REQUIRE(out_layout == ARECON_FP_ROWS || ZL >= 2,
                   "filter: the row-pair layout needs ZL >= 2");
    ARECON_REQUIRE(dYL != 0.f && SO > 0.f, "filter: dYL must be nonzero and SO positive");
    int dev = 0;
    ARECON_HIP(hipGetDevice(&dev));
    const TableSet t = fixed ? *fixed : cached_tables(dev, YL, ZL, PLC, ZLC, dYL, dZL, SO);
    const int N = t.N;
    const size_t lds = sizeof(float2) * 3 * (size_t)N + sizeof(float) * (size_t)N;
    const long long nrows = (long long)ZL * ViewN;
    const long long npairs = (nrows + 1) / 2;
    int cus = 0;
    ARECON_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
    const long long grid = std::min<long long>(npairs, (long long)cus * per_cu);
    static std::once_flag lds_once;
    std::call_once(lds_once, [] {
        // Transforms of 4096 points need more than the default 64 KB of LDS.
        const int max_lds = 160 * 1024;
        (void)hipFuncSetAttribute((const void*)filter_rows<float>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        (void)hipFuncSetAttribute((const void*)filter_rows<__half>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        (void)hipFuncSetAttribute((const void*)filter_rows<__hip_bfloat16>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    });
    if (N >= 256 && N <= 2048) {
        const float* W = t.W;
        const long long wgrid = std::min<long long>(npairs, (long long)cus * 16);
        auto go = [&](auto outp) {
            using OT = std::remove_pointer_t<decltype(outp)>;
            switch (N) {
            case 256: hipLaunchKernelGGL((filter_rows_wave<256, OT, 0>), dim3((unsigned)wgrid), dim3(tpt<256>()), 0, s,
                                         outp, out_layout, in, t.tw, t.H, W, YL, ZL, nrows, PLC, ZLC, dYL, dZL, SO); break;
            case 512: hipLaunchKernelGGL((filter_rows_wave<512, OT, 0>), dim3((unsigned)wgrid), dim3(tpt<512>()), 0, s,
                                         outp, out_layout, in, t.tw, t.H, W, YL, ZL, nrows, PLC, ZLC, dYL, dZL, SO); break;
            case 1024: hipLaunchKernelGGL((filter_rows_wave<1024, OT, 0>), dim3((unsigned)wgrid), dim3(tpt<1024>()), 0, s,
                                          outp, out_layout, in, t.tw, t.H, W, YL, ZL, nrows, PLC, ZLC, dYL, dZL, SO); break;
            default:
                if (filter_schedule() == 1)
                    hipLaunchKernelGGL((filter_rows_wave<2048, OT, 1>), dim3((unsigned)wgrid), dim3(tpt<2048, 1>()), 0, s,
                                       outp, out_layout, in, t.tw, t.H, W, YL, ZL, nrows, PLC, ZLC, dYL, dZL, SO);
                else
                    hipLaunchKernelGGL((filter_rows_wave<2048, OT, 0>), dim3((unsigned)wgrid), dim3(tpt<2048, 0>()), 0, s,
                                       outp, out_layout, in, t.tw, t.H, W, YL, ZL, nrows, PLC, ZLC, dYL, dZL, SO);
                break;
            }
        };
        switch (out_dtype) {
        case ARECON_F32: go(static_cast<float*>(out)); break;
        case ARECON_F16: go(static_cast<__half*>(out)); break;
        default: go(static_cast<__hip_bfloat16*>(out)); break;
        }
        ARECON_HIP(hipGetLastError());
        return;
    }
    switch (out_dtype) {
    case ARECON_F32:
        hipLaunchKernelGGL(filter_rows<float>, dim3((unsigned)grid), dim3(256), lds, s,
                           static_cast<float*>(out), out_layout, in, t.tw, t.H, YL, ZL, nrows, N, PLC,
                           ZLC, dYL, dZL, SO);
        break;
    case ARECON_F16:
        hipLaunchKernelGGL(filter_rows<__half>, dim3((unsigned)grid), dim3(256), lds, s,
                           static_cast<__half*>(out), out_layout, in, t.tw, t.H, YL, ZL, nrows, N, PLC,
                           ZLC, dYL, dZL, SO);
        break;
    default:
        hipLaunchKernelGGL(filter_rows<__hip_bfloat16>, dim3((unsigned)grid), dim3(256), lds,
                           s, static_cast<__hip_bfloat16*>(out), out_layout, in, t.tw, t.H,
                           YL, ZL, nrows, N, PLC, ZLC, dYL, dZL, SO);
        break;
    }
    ARECON_HIP(hipGetLastError());
}

void launch_filter(void* out, int out_dtype, int out_layout, const float* in, int YL, int ZL,
                   int ViewN, float PLC, float ZLC, float dYL, float dZL, float SO,
                   hipStream_t s) {
    launch_filter_impl(out, out_dtype, out_layout, in, YL, ZL, ViewN, PLC, ZLC, dYL, dZL, SO, s,
                       nullptr);
}

// ---------------------------------------------------------------------------
// Allocation-free variant (arecon_filter_dev_noalloc).  The tables that
// launch_filter reads from its per-device caches (allocated on first use) are
// built instead, by a kernel on the caller's stream, in a fixed device array:
//
//   [0, 32 KB)         twiddles, N float2 (N <= 4096)
//   [32 KB, 48 KB)     spectrum H/N, N floats
//   [48 KB, 1 MB)      preweight [ZL][YL] floats: ZL * YL <= 249,856
//
// Each device holds one copy for the life of the process.  Every call on a
// device rewrites it, so calls that overlap on one device (two streams, two
// threads) race; calls on one stream are ordered.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(1) << 20;
constexpr size_t kScratchTwOff = 0;
constexpr size_t kScratchHOff = 32 * 1024;
constexpr size_t kScratchWOff = 48 * 1024;
constexpr long long kScratchMaxW = (long long)((kScratchBytes - kScratchWOff) / sizeof(float));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current (the symbol's host-side name is not a device address).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static unsigned char* scratch_ptr(int dev) {
    ARECON_REQUIRE(dev >= 0 && dev < kMaxDevices, "filter (no-alloc): device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ARECON_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ARECON_REQUIRE(p != nullptr, "filter (no-alloc): scratch symbol has no device address");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

// The same tables filter_tables() and preweight_table() compute on the host,
// computed here in FP64 with the same expressions in the same order (FMA
// contraction off) and rounded to FP32.  The spectrum sums the taps against the
// cosine table kept in LDS, as the host does; only the libm behind cos/sin
// differs (OCML vs the host's), by at most an ulp of FP64.
__global__ __launch_bounds__(256) void filter_tables_dev(float2* __restrict__ tw,
                                                         float* __restrict__ H,
                                                         float* __restrict__ W, int N, int YL,
                                                         int ZL, float PLC, float ZLC,
                                                         float dYL_f, float dZL, float SO) {
#pragma clang fp contract(off)
    extern __shared__ double tabs[];     // c[N], then h[YL]
    double* c = tabs;
    double* h = tabs + N;
    const double dYL = (double)dYL_f;
    for (int m = threadIdx.x; m < N; m += blockDim.x) c[m] = cos(2.0 * M_PI * m / N);
    for (int n = threadIdx.x; n < YL; n += blockDim.x)
        h[n] = -2.0 / (9.869604401089358 * (4.0 * n * n - 1.0)) / dYL;
    __syncthreads();
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < N; f += gridDim.x * blockDim.x) {
        double acc = h[0];
        long long idx = 0;
        for (int n = 1; n < YL; ++n) {
            idx += f;
            if (idx >= N) idx -= N;
            acc += 2.0 * h[n] * c[idx];
        }
        H[f] = (float)(acc / N);
        tw[f] = make_float2((float)c[f], (float)-sin(2.0 * M_PI * f / N));
    }
    const double so2 = (double)SO * SO;
    const long long nw = (long long)ZL * YL;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nw;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / YL);
        const int j = (int)(i - (long long)k * YL);
        const double b = ((double)k - ZLC) * dZL;
        const double t = ((double)j - PLC) * dYL;
        W[i] = (float)(so2 / sqrt(so2 * (so2 + b * b) - b * b * t * t));
    }
}

void launch_filter_noalloc(void* out, int out_dtype, int out_layout, const float* in, int YL,
                           int ZL, int ViewN, float PLC, float ZLC, float dYL, float dZL,
                           float SO, hipStream_t s) {
    ARECON_REQUIRE(YL > 0 && ZL > 0 && ViewN > 0, "filter: YL, ZL and ViewN must be positive");
    ARECON_REQUIRE(YL <= 2048, "filter: YL above 2048 is not supported");
    ARECON_REQUIRE((long long)ZL * YL <= kScratchMaxW,
                   "filter (no-alloc): ZL * YL exceeds the 249,856 detector samples per view "
                   "that its fixed 1 MiB scratch holds; use arecon_filter_dev");
    ARECON_REQUIRE(dYL != 0.f && SO > 0.f, "filter: dYL must be nonzero and SO positive");
    int dev = 0;
    ARECON_HIP(hipGetDevice(&dev));
    unsigned char* base = scratch_ptr(dev);
    const int N = filter_length(YL);
    TableSet t;
    t.tw = reinterpret_cast<const float2*>(base + kScratchTwOff);
    t.H = reinterpret_cast<const float*>(base + kScratchHOff);
    t.W = reinterpret_cast<const float*>(base + kScratchWOff);
    t.N = N;
    const size_t lds = sizeof(double) * ((size_t)N + (size_t)YL);
    const unsigned blocks = (unsigned)std::max(1, N / 256);
    hipLaunchKernelGGL(filter_tables_dev, dim3(blocks), dim3(256), lds, s,
                       reinterpret_cast<float2*>(base + kScratchTwOff),
                       reinterpret_cast<float*>(base + kScratchHOff),
                       reinterpret_cast<float*>(base + kScratchWOff), N, YL, ZL, PLC, ZLC, dYL,
                       dZL, SO);
    ARECON_HIP(hipGetLastError());
    launch_filter_impl(out, out_dtype, out_layout, in, YL, ZL, ViewN, PLC, ZLC, dYL, dZL, SO, s,
                       &t);
}

}  // namespace arecon
