// Kernels of the learned artifact-correction network (3-D residual
// encoder-decoder, 3x3x3 convolutions): BF16 activations and weights on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation.  No
// counterpart in the reference (it has no learned part).
//
// Layouts.
//  * Activations: channels-last BF16, [B][D][H][W][C] with C a multiple of 8
//    (the one-channel input is padded to 8 with zeros).
//  * Forward weights: BF16 [Cout][Kp], k = tap * Cin + ci, tap = 9 (dz+1) +
//    3 (dy+1) + (dx+1), Kp = 27 Cin rounded up to 32 (zero tail).
//  * Weight gradients: FP32 [Cout][27][Cin].
//
// MFMA 16x16x32 BF16 lane layout (checked on the device): A (16 x 32): lane l
// holds A[l%16][8(l/16) .. +8]; B (32 x 16): lane l holds B[8(l/16) .. +8][l%16];
// D (16 x 16, FP32): lane l holds D[4(l/16) + v][l%16], v = 0..3.
//
// Convolution as implicit GEMM: out[co][n] = sum_k W[co][k] X[nb(n, tap(k))][ci(k)].
// A lane's 8 consecutive k are 8 consecutive channels of one tap, i.e. one
// 16-byte load from the channels-last tensor (zero outside the volume).  The
// data gradient is the same convolution with flipped taps and transposed
// weights (packed that way by pack_weights, so it runs as a plain forward
// convolution); the weight gradient reduces over voxels, for which the activations
// and output gradients are first rewritten channel-major (the input also in
// three copies shifted by dx = -1, 0, +1, so every 8-voxel run is one aligned
// 16-byte load).
#include "arecon_internal.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace arecon {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(3))) void lds_void;

struct Vol {            // batch of volumes, channels-last
    int B, D, H, W;
    __host__ __device__ long long nvox() const { return (long long)B * D * H * W; }
};

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ u16 f2bf(float f) {            // round to nearest even
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return (u16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 zero8() {
    return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
}
__device__ __forceinline__ bf16x8 load8(const u16* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// ---------------------------------------------------------------------------
// Forward convolution (also the data gradient).  A wavefront computes 32
// voxels (two 16-voxel tiles) x all output channels (MT tiles of 16); a
// workgroup is 4 wavefronts.  Epilogue: + bias, + skip (pre-activation), ReLU.
// ---------------------------------------------------------------------------
template <int MT>
__global__ __launch_bounds__(256) void conv3_fwd(u16* __restrict__ y, const u16* __restrict__ x,
                                                 const u16* __restrict__ w,
                                                 const float* __restrict__ bias,
                                                 const u16* __restrict__ skip, Vol v, int cin,
                                                 int cout, int kp, int relu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long nv = v.nvox();
    const long long n0 = ((long long)blockIdx.x * 4 + wave) * 32;
    if (n0 >= nv) return;
    const int col = lane & 15, kg = lane >> 4;
    // the two voxels this lane feeds as B columns
    int zc[2], yc[2], xc[2];
    long long nb0[2];
    bool inb[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const long long n = n0 + 16 * u + col;
        inb[u] = n < nv;
        const long long nn = inb[u] ? n : 0;
        xc[u] = (int)(nn % v.W);
        const long long t = nn / v.W;
        yc[u] = (int)(t % v.H);
        const long long t2 = t / v.H;
        zc[u] = (int)(t2 % v.D);
        nb0[u] = nn;
    }
    f32x4 acc[MT][2];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunks = kp / 32;
    for (int c = 0; c < nchunks; ++c) {
        // K steps in the tiled kernel's order: Cin a multiple of 32, 32-channel
        // passes, each over the 27 taps (step c: tap c % 27, channels
        // 32 (c / 27) + ..); otherwise 32 consecutive k
        int k0, tap, ci0;
        if (cin % 32 == 0) {
            tap = c % 27;
            ci0 = 32 * (c / 27) + kg * 8;
            k0 = tap * cin + ci0;
        } else {
            k0 = c * 32 + kg * 8;
            tap = k0 / cin;
            ci0 = k0 - tap * cin;
        }
        bf16x8 b[2];
        if (tap < 27) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int zz = zc[u] + dz, yy = yc[u] + dy, xx = xc[u] + dx;
                const bool ok = inb[u] && zz >= 0 && zz < v.D && yy >= 0 && yy < v.H &&
                                xx >= 0 && xx < v.W;
                const long long nbr = nb0[u] + ((long long)dz * v.H + dy) * v.W + dx;
                b[u] = ok ? load8(x + nbr * cin + ci0) : zero8();
            }
        } else {
            b[0] = b[1] = zero8();
        }
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const bf16x8 a = load8(w + (long long)(16 * t + col) * kp + k0);
#pragma unroll
            for (int u = 0; u < 2; ++u)
                acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[u], acc[t][u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const long long n = n0 + 16 * u + col;
        if (n >= nv) continue;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const int co0 = 16 * t + 4 * kg;
            if (co0 >= cout) continue;
            u16 o[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float r = acc[t][u][q] + (bias ? bias[co0 + q] : 0.f);
                if (skip) r += bf2f(skip[n * cout + co0 + q]);
                if (relu) r = fmaxf(r, 0.f);
                o[q] = f2bf(r);
            }
            *reinterpret_cast<uint2*>(y + n * cout + co0) =
                make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16),
                           (unsigned)o[2] | ((unsigned)o[3] << 16));
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled forward convolution (also the data gradient) for Cin <= 64.  The
// untiled kernel above fetches every operand through the texture addresser:
// each activation is read again for each of the 27 taps, and that, not the
// matrix cores, bounds it.  Here a workgroup (4 wavefronts) computes an output
// tile of 4 (z) x NV (y) x 16 (x) voxels of one volume: wavefront w the NV rows
// of 16 voxels at z0 + w, i.e. NV B tiles, times all MT output-channel tiles.
// The input tile with its one-voxel halo, 6 x (NV+2) x 18 voxels x Cin BF16
// channels, is staged in LDS once (zeros outside the volume), and every B
// operand is one 16-byte LDS read.  A (weights) comes from global memory, the
// same 16-byte pieces for every wavefront (L1 hits), once per K chunk and reused
// over the NV tiles.  The K loop is the untiled kernel's (k = tap Cin + ci in
// chunks of 32), so every output sees the same MFMA sequence and the result is
// identical.
// ---------------------------------------------------------------------------
template <int MT, int NV, int CIN>
__global__ __launch_bounds__(256, 2) void conv3_fwd_tiled(u16* __restrict__ y,
                                                       const u16* __restrict__ x,
                                                       const u16* __restrict__ w,
                                                       const float* __restrict__ bias,
                                                       const u16* __restrict__ skip, Vol v,
                                                       int cout, int relu,
                                                       int tilesx, int tilesy, int tilesz) {
    constexpr int cin = CIN, kp = (27 * CIN + 31) / 32 * 32;
    // channels staged per pass (CS) and passes: all of them up to 32, else 32 at
    // a time, so the tile takes 69 KB and two workgroups share a CU (one stages
    // while the other computes)
    constexpr int CS = CIN < 32 ? CIN : 32, NPASS = CIN / CS;
    constexpr int nchunks = CIN < 32 ? kp / 32 : 27;   // K steps per pass
    extern __shared__ uint4 tile[];          // [6][NV+2][18][CS/8] 16-byte pieces
    constexpr int TY = NV + 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int t = blockIdx.x;
    const int tx = t % tilesx; t /= tilesx;
    const int ty = t % tilesy; t /= tilesy;
    const int tz = t % tilesz;
    const int b = t / tilesz;
    const int x0 = tx * 16, y0 = ty * NV, z0 = tz * 4;
    constexpr int c8 = CS / 8;
    // the tile's 16-byte slots, padded to whole 64-lane staging instructions
    constexpr int NSLOT = 6 * TY * 18 * c8, NINS = (NSLOT + 63) / 64, NSLOTP = NINS * 64;
    const long long zbase = (long long)b * v.D;
    const int col = lane & 15, kg = lane >> 4;
    f32x4 acc[MT][NV];
#pragma unroll
    for (int tt = 0; tt < MT; ++tt)
#pragma unroll
        for (int r = 0; r < NV; ++r) acc[tt][r] = f32x4{0.f, 0.f, 0.f, 0.f};
    const u16* wl = w + (long long)col * kp + kg * 8;
    // Staging: direct global -> LDS loads (no VGPRs, asynchronous), 64 slots per
    // instruction, the wavefronts taking every fourth instruction.  Lane slot s
    // of the tile (voxel-major) loads its voxel's piece through a
    // buffer descriptor based at the halo's origin voxel; a voxel outside the
    // volume (or a padding slot) gets an offset past the descriptor's range and
    // loads zeros.  One buffer (69 KB at most): two workgroups share a CU, and
    // one stages while the other computes.  (Double-buffered, pass h + 1 loading
    // while pass h computes, the 139 KB tile left one workgroup per CU: 64-channel
    // layer 16.3 -> 22.7 ms.)
    const long long org = ((zbase + z0 - 1) * v.H + (y0 - 1)) * (long long)v.W + (x0 - 1);
    auto stage = [&](int pass, int buf) {
        const unsigned long long bp =
            reinterpret_cast<unsigned long long>(x) + (unsigned long long)((org * cin + pass * CS) * 2);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)bp);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(bp >> 32));
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo), (short)0, 0x7fffffff,
            0x00020000);
        for (int j = wave; j < NINS; j += 4) {
            const int sl = 64 * j + lane;
            unsigned off = 0x80000000u;
            if (sl < NSLOT) {
                const int pp = sl % c8;
                int r = sl / c8;
                const int hx = r % 18;
                r /= 18;
                const int hy = r % TY, hz = r / TY;
                const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
                if (gz >= 0 && gz < v.D && gy >= 0 && gy < v.H && gx >= 0 && gx < v.W)
                    off = (unsigned)(((((long long)hz * v.H + hy) * v.W + hx) * cin +
                                      8 * pp) * 2);
            }
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rs, (lds_void*)((lds_char*)tile + ((long long)buf * NSLOTP + 64 * j) * 16), 16, off,
                0, 0, 0);
        }
    };
#pragma unroll 1
    for (int pass = 0; pass < NPASS; ++pass) {
        if (pass > 0) __syncthreads();       // every read of the last pass's tile
        stage(pass, 0);
        __builtin_amdgcn_s_waitcnt(0);       // this wavefront's loads landed
        __syncthreads();                     // and every other wavefront's
        const int bufo = 0;
        // K step c of this pass: k0 = its first k (the lane adds 8 kg), and the
        // tap and staged piece its B reads.  Cin >= 32: one tap's 32 channels
        // of the pass; Cin < 32: 32 consecutive k of the whole K.  The padded
        // K tail (k >= 27 Cin, only for Cin = 8) has zero weights, so its B may
        // be any finite samples: the tap is clamped to 26 for the address.
        auto kstep = [&](int c, int& k0, int& tap, int& piece) {
            if constexpr (CIN >= 32) {
                tap = c;
                k0 = c * CIN + pass * 32;
                piece = kg;
            } else {
                const int kl = c * 32 + kg * 8;
                const int tk = kl / CIN;
                tap = min(tk, 26);
                k0 = c * 32;
                piece = (kl - tk * CIN) >> 3;
            }
        };
        auto aaddr = [&](int c) {
            int k0, tap, piece;
            kstep(c, k0, tap, piece);
            return wl + k0;
        };
        auto bload = [&](int c, bf16x8* bv) {
            int k0, tap, piece;
            kstep(c, k0, tap, piece);
            const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;     // halo-tile offsets
            const int base = bufo + (((wave + dz) * TY + dy) * 18 + col + dx) * c8 +
                             piece;
#pragma unroll
            for (int r = 0; r < NV; ++r)
                bv[r] = __builtin_bit_cast(bf16x8, tile[base + r * 18 * c8]);
        };
        // Cin >= 32: the pass's 27 steps are one tap each, unrolled, so every LDS
        // and weight address is a per-lane base plus an immediate.  A (weights,
        // from L1/L2) is loaded kPA steps ahead in a ring of registers; B (the NV
        // LDS reads of a step) one step ahead into the other of two register
        // sets.  A scheduling barrier closes each step, so the compiler keeps the
        // loads where they are issued (left free, it moved each next to its use;
        // with the steps computed at run time, their address arithmetic cost
        // about 90 VALU per step).  MT = 8 (128 output channels) has no room for
        // the rings beside its 128 accumulator registers and takes the run-time
        // loop below.
        if constexpr (CIN >= 32 && MT <= 4) {
            int lb[3];                        // per-lane base of a B read, by dx
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
                lb[dx] = bufo + (wave * TY * 18 + col + dx) * c8 + kg;
            const u16* wlp = wl + pass * 32;
            constexpr int kPA = MT <= 2 ? 4 : 2;
            bf16x8 ar[kPA][MT];
#pragma unroll
            for (int i = 0; i < kPA; ++i)
#pragma unroll
                for (int tt = 0; tt < MT; ++tt)
                    ar[i][tt] = load8(wlp + (long long)16 * tt * kp + i * CIN);
            bf16x8 bb[2][NV];
#pragma unroll
            for (int r = 0; r < NV; ++r)
                bb[0][r] = __builtin_bit_cast(bf16x8, tile[lb[0] + r * 18 * c8]);
#pragma unroll
            for (int c = 0; c < 27; ++c) {
                if (c + 1 < 27) {
                    const int tn = c + 1, dz = tn / 9, dy = (tn / 3) % 3, dx = tn % 3;
#pragma unroll
                    for (int r = 0; r < NV; ++r)
                        bb[(c + 1) & 1][r] = __builtin_bit_cast(
                            bf16x8, tile[lb[dx] + ((dz * TY + dy) * 18 + r * 18) * c8]);
                }
#pragma unroll
                for (int r = 0; r < NV; ++r)
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt)
                        acc[tt][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            ar[c % kPA][tt], bb[c & 1][r], acc[tt][r], 0, 0, 0);
                if (c + kPA < 27) {
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt)
                        ar[c % kPA][tt] = load8(wlp + (long long)16 * tt * kp + (c + kPA) * CIN);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else if constexpr (MT <= 2) {
            // Cin < 32, MT <= 2: the same pipeline over run-time steps (a step
            // spans taps: 32 consecutive k)
            constexpr int kPA = 4;
            bf16x8 ar[kPA][MT];
#pragma unroll
            for (int i = 0; i < kPA; ++i)
                if (i < nchunks) {
                    const u16* pa = aaddr(i);
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt) ar[i][tt] = load8(pa + (long long)16 * tt * kp);
                }
            bf16x8 bb[2][NV];
            bload(0, bb[0]);
            // not unrolled: each turn's loads for the next turn stay a turn ahead
            // (fully unrolled, the scheduler moved every load next to its use)
#pragma unroll 1
            for (int c0 = 0; c0 < nchunks; c0 += kPA) {
#pragma unroll
                for (int i = 0; i < kPA; ++i) {  // kPA even: step c's B set is c & 1 = i & 1
                    const int c = c0 + i;
                    if (c < nchunks) {
                        if (c + 1 < nchunks) bload(c + 1, bb[(i + 1) & 1]);
#pragma unroll
                        for (int r = 0; r < NV; ++r)
#pragma unroll
                            for (int tt = 0; tt < MT; ++tt)
                                acc[tt][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    ar[i][tt], bb[i & 1][r], acc[tt][r], 0, 0, 0);
                        if (c + kPA < nchunks) {
                            const u16* pa = aaddr(c + kPA);
#pragma unroll
                            for (int tt = 0; tt < MT; ++tt)
                                ar[i][tt] = load8(pa + (long long)16 * tt * kp);
                        }
                    }
                }
            }
        } else {
            // MT >= 4 (Cin < 32, or MT = 8): A one step ahead, B at the step's
            // start (the accumulators leave no room for the rings)
            bf16x8 a[MT], an[MT];
            {
                const u16* pa = aaddr(0);
#pragma unroll
                for (int tt = 0; tt < MT; ++tt) a[tt] = load8(pa + (long long)16 * tt * kp);
            }
            for (int c = 0; c < nchunks; ++c) {
                if (c + 1 < nchunks) {
                    const u16* pa = aaddr(c + 1);
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt) an[tt] = load8(pa + (long long)16 * tt * kp);
                }
                bf16x8 bv[NV];
                bload(c, bv);
#pragma unroll
                for (int r = 0; r < NV; ++r)
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt)
                        acc[tt][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[tt], bv[r],
                                                                              acc[tt][r], 0, 0, 0);
                if (c + 1 < nchunks) {
#pragma unroll
                    for (int tt = 0; tt < MT; ++tt) a[tt] = an[tt];
                }
            }
        }
    }
    const int zz = z0 + wave, xx = x0 + col;
    if (zz >= v.D || xx >= v.W) return;
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int yy = y0 + r;
        if (yy >= v.H) continue;
        const long long n = ((zbase + zz) * v.H + yy) * v.W + xx;
#pragma unroll
        for (int tt = 0; tt < MT; ++tt) {
            const int co0 = 16 * tt + 4 * kg;
            if (co0 >= cout) continue;
            u16 o[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float rr = acc[tt][r][q] + (bias ? bias[co0 + q] : 0.f);
                if (skip) rr += bf2f(skip[n * cout + co0 + q]);
                if (relu) rr = fmaxf(rr, 0.f);
                o[q] = f2bf(rr);
            }
            *reinterpret_cast<uint2*>(y + n * cout + co0) =
                make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16),
                           (unsigned)o[2] | ((unsigned)o[3] << 16));
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] += sum_n dY[co][n] X_dx[ci][nb(n, dz, dy)],
// with dY (already masked by the ReLU derivative) and X channel-major.  One
// wavefront: 16 co x 16 ci of one tap over one range (split) of the voxels;
// the K dimension of the MFMA is 32 voxels of one row segment (W is a multiple
// of 8, so each lane's 8 voxels lie in one row).  Each wavefront writes its
// tile's partial sums to part[split][co][tap][ci] (every entry has exactly one
// writer); wgrad_reduce then adds the splits in order, so the gradient is the
// same bits on every run.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv3_wgrad(float* __restrict__ part,
                                                   const u16* __restrict__ dyT,
                                                   const u16* __restrict__ xT3, Vol v, int cin,
                                                   int cout, long long vox_per_split,
                                                   int nsplit) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ct = (cout + 15) / 16, it = (cin + 15) / 16;
    long long job = (long long)blockIdx.x * 4 + wave;
    const long long njobs = 27LL * ct * it * nsplit;
    if (job >= njobs) return;
    const int split = (int)(job % nsplit);
    job /= nsplit;
    const int ti = (int)(job % it);
    job /= it;
    const int tc = (int)(job % ct);
    const int tap = (int)(job / ct);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const long long nv = v.nvox();
    const long long nbeg = split * vox_per_split;
    const long long nend = std::min(nv, nbeg + vox_per_split);
    const int col = lane & 15, kg = lane >> 4;
    const int co = 16 * tc + col, ci = 16 * ti + col;
    const u16* xs = xT3 + (long long)(dx + 1) * cin * nv;    // copy shifted by dx
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long n = nbeg; n < nend; n += 32) {
        const long long m = n + 8 * kg;                       // this lane's 8 voxels
        bf16x8 a = zero8(), b = zero8();
        if (m < nend) {
            if (co < cout) a = load8(dyT + (long long)co * nv + m);
            if (ci < cin) {
                const long long t = m / v.W;
                const int yy = (int)(t % v.H) + dy;
                const int zz = (int)((t / v.H) % v.D) + dz;
                if (yy >= 0 && yy < v.H && zz >= 0 && zz < v.D)
                    b = load8(xs + (long long)ci * nv + m + ((long long)dz * v.H + dy) * v.W);
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D[i = co in tile][j = ci in tile]
    float* ps = part + (long long)split * cout * 27 * cin;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c_o = 16 * tc + 4 * kg + q, c_i = 16 * ti + col;
        if (c_o < cout && c_i < cin) ps[((long long)c_o * 27 + tap) * cin + c_i] = acc[q];
    }
}

// ---------------------------------------------------------------------------
// Weight gradient with the input staged in LDS.  The gathering kernel above
// issues two 16-byte global loads per MFMA, so the texture addresser (a 64-lane
// load every 16 clocks per CU), not the matrix cores, bounds it.  Here a
// workgroup takes one 16-channel input tile and NCO 16-channel output tiles and
// walks voxel tiles of 2 (z) x 4 (y) x 32 (x).  Per voxel tile it stages the
// input tile's three dx-shifted channel-major copies with their z/y halo,
// [3][16][4][6][32] BF16 (one 16-byte pad per channel plane, so the 16 lanes of
// a B read cover all 64 banks), 74.5 KB: two workgroups per CU.  A K step is one
// x row of 32 voxels: each wavefront loads its output tile's dY for the row once
// (16 bytes per lane, channel-major) and reads the shifted X rows of its taps
// from LDS, one MFMA per tap.  S = 1: a wavefront owns all 27 taps of its output
// tile (27 accumulators); S = 3: one dz plane of them (9), three wavefronts per
// output tile.  Partial sums per (split, co, tap, ci) go to part and sum_parts
// adds the splits in order, so the gradient is deterministic.
// ---------------------------------------------------------------------------
constexpr int kWgTz = 2, kWgTy = 4, kWgTx = 32;
constexpr int kWgZh = kWgTz + 2, kWgYh = kWgTy + 2;
constexpr int kWgPlane = kWgZh * kWgYh * kWgTx * 2 + 16;        // bytes per channel plane
constexpr int kWgCopy = 16 * kWgPlane;                          // bytes per dx copy
constexpr int kWgLds = 3 * kWgCopy;                             // 74,496 bytes

template <int NCO, int S>
__global__ __launch_bounds__(64 * NCO * S) void conv3_wgrad_lds(
    float* __restrict__ part, const u16* __restrict__ dyT, const u16* __restrict__ xT3, Vol v,
    int cin, int cout, int tiles_per_split, int ntiles) {
    constexpr int NT = S == 1 ? 27 : 9;                  // taps per wavefront
    constexpr int NTHR = 64 * NCO * S;
    extern __shared__ uint4 xs4[];
    char* xs = reinterpret_cast<char*>(xs4);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int split = blockIdx.x, ct = blockIdx.y;
    const int cot = blockIdx.z * NCO + wave / S;         // this wavefront's output tile
    const int dzw = S == 1 ? 0 : wave % S;               // its dz plane (S = 3)
    const int col = lane & 15, g = lane >> 4;
    const long long nv = v.nvox();
    const int TX = (v.W + kWgTx - 1) / kWgTx, TY = (v.H + kWgTy - 1) / kWgTy;
    const int TZ = (v.D + kWgTz - 1) / kWgTz;
    const int co = cot * 16 + col;
    const bool co_ok = co < cout;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int tb = split * tiles_per_split, te = min(ntiles, tb + tiles_per_split);
    for (int tile = tb; tile < te; ++tile) {
        int r = tile;
        const int tx = r % TX; r /= TX;
        const int ty = r % TY; r /= TY;
        const int tz = r % TZ;
        const int b = r / TZ;
        const int x0 = tx * kWgTx, y0 = ty * kWgTy, z0 = tz * kWgTz;
        // stage: 3 copies x 16 channels x 4 x 6 rows x 4 pieces of 8 voxels
        constexpr int kPieces = 3 * 16 * kWgZh * kWgYh * 4;
        for (int i = threadIdx.x; i < kPieces; i += NTHR) {
            const int q = i & 3;
            int rr = i >> 2;
            const int yh = rr % kWgYh; rr /= kWgYh;
            const int zh = rr % kWgZh; rr /= kWgZh;
            const int ci = rr & 15, c = rr >> 4;
            const int z = z0 + zh - 1, y = y0 + yh - 1, x = x0 + 8 * q;
            const int cg = ct * 16 + ci;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (z >= 0 && z < v.D && y >= 0 && y < v.H && x < v.W && cg < cin)
                val = *reinterpret_cast<const uint4*>(
                    xT3 + ((long long)c * cin + cg) * nv +
                    (((long long)b * v.D + z) * v.H + y) * v.W + x);
            *reinterpret_cast<uint4*>(xs + c * kWgCopy + ci * kWgPlane +
                                      ((zh * kWgYh + yh) * kWgTx + 8 * q) * 2) = val;
        }
        __syncthreads();
        if (cot * 16 < cout) {
            // this lane's B rows: channel col, voxels 8g .. 8g+7 of a row
            const char* bl = xs + col * kWgPlane + 16 * g + dzw * kWgYh * kWgTx * 2;
            // dY of row k (zero outside the volume), loaded a row ahead
            auto load_a = [&](int k) {
                const int z = z0 + k / kWgTy, y = y0 + k % kWgTy, x = x0 + 8 * g;
                bf16x8 a = zero8();
                if (co_ok && z < v.D && y < v.H && x < v.W)
                    a = load8(dyT + (long long)co * nv +
                              (((long long)b * v.D + z) * v.H + y) * v.W + x);
                return a;
            };
            bf16x8 an = load_a(0);
#pragma unroll 1
            for (int k = 0; k < kWgTz * kWgTy; ++k) {
                const int zz = k / kWgTy, yy = k % kWgTy;
                const bf16x8 a = an;
                if (k + 1 < kWgTz * kWgTy) an = load_a(k + 1);
                const char* bk = bl + (zz * kWgYh + yy) * kWgTx * 2;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int dz = S == 1 ? t / 9 : 0, dy = (t / 3) % 3, dx = t % 3;
                    const bf16x8 bv = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4*>(
                                    bk + dx * kWgCopy + (dz * kWgYh + dy) * kWgTx * 2));
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();                  // every read of the tile before the next staging
    }
    if (cot * 16 >= cout) return;
    // D[co = 4g + q][ci = col] of tap t
    float* ps = part + (long long)split * cout * 27 * cin;
    const int ci = ct * 16 + col;
    if (ci >= cin) return;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int tap = S == 1 ? t : dzw * 9 + t;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c_o = cot * 16 + 4 * g + q;
            if (c_o < cout) ps[((long long)c_o * 27 + tap) * cin + ci] = acc[t][q];
        }
    }
}

// dst[i] += sum over k = 0 .. nparts-1, in order, of part[k * n + i].
__global__ __launch_bounds__(256) void sum_parts(float* __restrict__ dst,
                                                 const float* __restrict__ part, long long n,
                                                 int nparts) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int k = 0; k < nparts; ++k) acc += part[(long long)k * n + i];
        dst[i] += acc;
    }
}

// First level of a two-level ordered sum: part2[gi][i] = sum over the parts
// k = 32 gi .. 32 gi + 31 (< nparts), in order, of part[k * n + i].
__global__ __launch_bounds__(256) void sum_part_groups(float* __restrict__ part2,
                                                       const float* __restrict__ part,
                                                       long long n, int nparts) {
    const int gi = blockIdx.y, k1 = min(nparts, 32 * gi + 32);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int k = 32 * gi; k < k1; ++k) acc += part[(long long)k * n + i];
        part2[(long long)gi * n + i] = acc;
    }
}

// ---------------------------------------------------------------------------
// Elementwise and layout kernels.
// ---------------------------------------------------------------------------
#define GRID_STRIDE(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (long long)gridDim.x * blockDim.x)

// dYp = dY * (Y > 0) (or dY when relu == 0), channels-last, and its
// channel-major copy.
__global__ void relu_bwd(u16* __restrict__ dyp, u16* __restrict__ dypT,
                         const u16* __restrict__ dy, const u16* __restrict__ y, long long nv,
                         int c, int relu) {
    GRID_STRIDE(i, nv * c) {
        const long long n = i / c;
        const int ch = (int)(i - n * c);
        u16 g = dy[i];
        if (relu && !(bf2f(y[i]) > 0.f)) g = 0;
        dyp[i] = g;
        dypT[(long long)ch * nv + n] = g;
    }
}

// Bias gradient, first pass: part[b][ch] = sum of dYT[ch][n] over block b's
// share of the voxels (blockIdx.y = channel); sum_parts adds the blocks in order.
__global__ __launch_bounds__(256) void channel_sum(float* __restrict__ part,
                                                   const u16* __restrict__ dyT, long long nv) {
    const int ch = blockIdx.y;
    float s = 0.f;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nv;
         n += (long long)gridDim.x * blockDim.x)
        s += bf2f(dyT[(long long)ch * nv + n]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        part[(long long)blockIdx.x * gridDim.y + ch] = (red[0] + red[1]) + (red[2] + red[3]);
}

// X channels-last -> three channel-major copies shifted by dx = -1, 0, +1
// (zero where x + dx leaves the row).
__global__ void to_channel_major3(u16* __restrict__ xT3, const u16* __restrict__ x, Vol v,
                                  int c) {
    const long long nv = v.nvox();
    GRID_STRIDE(i, nv * c) {
        const int ch = (int)(i / nv);
        const long long n = i - (long long)ch * nv;
        const int xx = (int)(n % v.W);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int dx = s - 1;
            const bool ok = xx + dx >= 0 && xx + dx < v.W;
            xT3[(long long)s * c * nv + i] = ok ? x[(n + dx) * c + ch] : (u16)0;
        }
    }
}

// The same two layouts when W is a multiple of 8 (every 8-voxel run lies in one
// row): a thread takes one channel of one 8-voxel run and writes 16-byte pieces
// of the channel-major copies; consecutive threads take consecutive channels, so
// each of its loads reads one voxel's channels, contiguous across the
// wavefront.  (The per-element kernels above read with a stride of c elements
// and run about 7x below the copy rate.)
__device__ __forceinline__ uint4 pack_u16x8(const u16* e) {
    return make_uint4((unsigned)e[0] | ((unsigned)e[1] << 16), (unsigned)e[2] | ((unsigned)e[3] << 16),
                      (unsigned)e[4] | ((unsigned)e[5] << 16), (unsigned)e[6] | ((unsigned)e[7] << 16));
}

__global__ void relu_bwd8(u16* __restrict__ dyp, u16* __restrict__ dypT,
                          const u16* __restrict__ dy, const u16* __restrict__ y, long long nv,
                          int c, int relu) {
    GRID_STRIDE(i, nv / 8 * c) {
        const long long grp = i / c;
        const int ch = (int)(i - grp * c);
        const long long n0 = grp * 8;
        u16 o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long long idx = (n0 + k) * c + ch;
            u16 gv = dy[idx];
            if (relu && !(bf2f(y[idx]) > 0.f)) gv = 0;
            dyp[idx] = gv;
            o[k] = gv;
        }
        *reinterpret_cast<uint4*>(dypT + (long long)ch * nv + n0) = pack_u16x8(o);
    }
}

__global__ void to_channel_major3_8(u16* __restrict__ xT3, const u16* __restrict__ x, Vol v,
                                    int c) {
    const long long nv = v.nvox();
    GRID_STRIDE(i, nv / 8 * c) {
        const long long grp = i / c;
        const int ch = (int)(i - grp * c);
        const long long n0 = grp * 8;
        const int x0 = (int)(n0 % v.W);
        u16 e[10];                                   // voxels n0 - 1 .. n0 + 8 of the row
        e[0] = x0 > 0 ? x[(n0 - 1) * c + ch] : (u16)0;
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k + 1] = x[(n0 + k) * c + ch];
        e[9] = x0 + 8 < v.W ? x[(n0 + 8) * c + ch] : (u16)0;
#pragma unroll
        for (int sft = 0; sft < 3; ++sft)            // dx = sft - 1: copy[n] = x[n + dx]
            *reinterpret_cast<uint4*>(xT3 + ((long long)sft * c + ch) * nv + n0) =
                pack_u16x8(e + sft);
    }
}

// The same two layouts through LDS, when W is a multiple of 8 and C of 8 and at
// most 128: a workgroup takes 64 consecutive voxels (with one voxel of halo on
// each side for the shifted copies), loads their channels-last block (one
// contiguous run of memory) with 16-byte loads, and writes the channel-major
// rows in 16-byte pieces of 8 voxels, eight lanes per 128-byte row segment.
// (The 8-voxel-run kernels above write 16 bytes per lane to 64 different rows
// per instruction: 1.5 TB/s.)
constexpr int kLtV = 64;                                  // voxels per workgroup

__global__ __launch_bounds__(256) void to_channel_major3_lds(u16* __restrict__ xT3,
                                                            const u16* __restrict__ x, Vol v,
                                                            int c) {
    __shared__ uint4 blk4[(kLtV + 2) * 128 / 8];
    u16* blk = reinterpret_cast<u16*>(blk4);            // [kLtV + 2][c], voxel n0 - 1 first
    const long long nv = v.nvox();
    const long long n0 = (long long)blockIdx.x * kLtV;
    const int c8 = c / 8;
    // voxels n0 - 1 .. n0 + kLtV (those outside [0, nv) load as zeros)
    for (int i = threadIdx.x; i < (kLtV + 2) * c8; i += 256) {
        const int vv = i / c8, p = i - vv * c8;
        const long long n = n0 - 1 + vv;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (n >= 0 && n < nv) val = *reinterpret_cast<const uint4*>(x + n * c + 8 * p);
        blk4[i] = val;
    }
    __syncthreads();
    // pieces: (shift, channel, 8-voxel run q)
    for (int i = threadIdx.x; i < 3 * c * (kLtV / 8); i += 256) {
        const int q = i & 7, r = i >> 3;
        const int ch = r % c, sft = r / c;
        const long long nq = n0 + 8 * q;
        if (nq >= nv) continue;
        const int xq = (int)(nq % v.W);                  // the run lies in one row
        u16 e[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int xx = xq + k + sft - 1;              // x of the voxel copied
            e[k] = (xx >= 0 && xx < v.W) ? blk[(8 * q + k + sft) * c + ch] : (u16)0;
        }
        *reinterpret_cast<uint4*>(xT3 + ((long long)sft * c + ch) * nv + nq) = pack_u16x8(e);
    }
}

__global__ __launch_bounds__(256) void relu_bwd_lds(u16* __restrict__ dyp, u16* __restrict__ dypT,
                                                   const u16* __restrict__ dy,
                                                   const u16* __restrict__ y, long long nv, int c,
                                                   int relu) {
    __shared__ uint4 blk4[kLtV * 128 / 8];
    u16* blk = reinterpret_cast<u16*>(blk4);            // [kLtV][c]: the masked gradient
    const long long n0 = (long long)blockIdx.x * kLtV;
    const int c8 = c / 8;
    for (int i = threadIdx.x; i < kLtV * c8; i += 256) {
        const long long e0 = n0 * c + 8LL * i;           // element offset of the piece
        if (e0 >= nv * c) continue;
        uint4 g = *reinterpret_cast<const uint4*>(dy + e0);
        if (relu) {
            const uint4 yv = *reinterpret_cast<const uint4*>(y + e0);
            unsigned gw[4] = {g.x, g.y, g.z, g.w};
            const unsigned yw[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned lo = bf2f((u16)(yw[k] & 0xffffu)) > 0.f ? 0x0000ffffu : 0u;
                const unsigned hi = bf2f((u16)(yw[k] >> 16)) > 0.f ? 0xffff0000u : 0u;
                gw[k] &= lo | hi;
            }
            g = make_uint4(gw[0], gw[1], gw[2], gw[3]);
        }
        *reinterpret_cast<uint4*>(dyp + e0) = g;
        blk4[i] = g;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c * (kLtV / 8); i += 256) {
        const int q = i & 7, ch = i >> 3;
        const long long nq = n0 + 8 * q;
        if (nq >= nv) continue;
        u16 e[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k] = blk[(8 * q + k) * c + ch];
        *reinterpret_cast<uint4*>(dypT + (long long)ch * nv + nq) = pack_u16x8(e);
    }
}

// 2x2x2 average pooling and its adjoint; nearest 2x upsampling and its adjoint.
// One thread per voxel and 8 channels (one 16-byte piece; C is a multiple of 8).
__device__ __forceinline__ void unpack8(const uint4& q, float* f) {
    const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = bf2f((u16)(u[i] & 0xffffu));
        f[2 * i + 1] = bf2f((u16)(u[i] >> 16));
    }
}
__device__ __forceinline__ uint4 pack8(const float* f) {
    unsigned u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = (unsigned)f2bf(f[2 * i]) | ((unsigned)f2bf(f[2 * i + 1]) << 16);
    return make_uint4(u[0], u[1], u[2], u[3]);
}

__global__ void avgpool2(u16* __restrict__ y, const u16* __restrict__ x, Vol vo, int c) {
    const int c8 = c >> 3;
    const uint4* x4 = reinterpret_cast<const uint4*>(x);
    uint4* y4 = reinterpret_cast<uint4*>(y);
    GRID_STRIDE(i, vo.nvox() * c8) {
        const int cg = (int)(i % c8);
        long long n = i / c8;
        const int xo = (int)(n % vo.W); n /= vo.W;
        const int yo = (int)(n % vo.H); n /= vo.H;
        const int zo = (int)(n % vo.D);
        const long long b = n / vo.D;
        const int Wi = 2 * vo.W, Hi = 2 * vo.H, Di = 2 * vo.D;
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int d = 0; d < 8; ++d) {
            const long long ni = ((b * Di + 2 * zo + (d >> 2)) * Hi + 2 * yo + ((d >> 1) & 1)) * Wi +
                                 2 * xo + (d & 1);
            float f[8];
            unpack8(x4[ni * c8 + cg], f);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += f[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] *= 0.125f;
        y4[i] = pack8(s);
    }
}
__global__ void avgpool2_bwd(u16* __restrict__ dx, const u16* __restrict__ dy, Vol vi, int c,
                             int accumulate) {
    const int c8 = c >> 3;
    const uint4* dy4 = reinterpret_cast<const uint4*>(dy);
    uint4* dx4 = reinterpret_cast<uint4*>(dx);
    GRID_STRIDE(i, vi.nvox() * c8) {
        const int cg = (int)(i % c8);
        long long n = i / c8;
        const int xi = (int)(n % vi.W); n /= vi.W;
        const int yi = (int)(n % vi.H); n /= vi.H;
        const int zi = (int)(n % vi.D);
        const long long b = n / vi.D;
        const long long no = ((b * (vi.D / 2) + zi / 2) * (vi.H / 2) + yi / 2) * (vi.W / 2) + xi / 2;
        float g[8];
        unpack8(dy4[no * c8 + cg], g);
#pragma unroll
        for (int k = 0; k < 8; ++k) g[k] *= 0.125f;
        if (accumulate) {
            float f[8];
            unpack8(dx4[i], f);
#pragma unroll
            for (int k = 0; k < 8; ++k) g[k] += f[k];
        }
        dx4[i] = pack8(g);
    }
}
__global__ void upsample2(u16* __restrict__ y, const u16* __restrict__ x, Vol vo, int c) {
    const int c8 = c >> 3;
    const uint4* x4 = reinterpret_cast<const uint4*>(x);
    uint4* y4 = reinterpret_cast<uint4*>(y);
    GRID_STRIDE(i, vo.nvox() * c8) {
        const int cg = (int)(i % c8);
        long long n = i / c8;
        const int xo = (int)(n % vo.W); n /= vo.W;
        const int yo = (int)(n % vo.H); n /= vo.H;
        const int zo = (int)(n % vo.D);
        const long long b = n / vo.D;
        const long long ni = ((b * (vo.D / 2) + zo / 2) * (vo.H / 2) + yo / 2) * (vo.W / 2) + xo / 2;
        y4[i] = x4[ni * c8 + cg];
    }
}
__global__ void upsample2_bwd(u16* __restrict__ dx, const u16* __restrict__ dy, Vol vi, int c) {
    const int c8 = c >> 3;
    const uint4* dy4 = reinterpret_cast<const uint4*>(dy);
    uint4* dx4 = reinterpret_cast<uint4*>(dx);
    GRID_STRIDE(i, vi.nvox() * c8) {
        const int cg = (int)(i % c8);
        long long n = i / c8;
        const int xi = (int)(n % vi.W); n /= vi.W;
        const int yi = (int)(n % vi.H); n /= vi.H;
        const int zi = (int)(n % vi.D);
        const long long b = n / vi.D;
        const int Wo = 2 * vi.W, Ho = 2 * vi.H, Do = 2 * vi.D;
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int d = 0; d < 8; ++d) {
            const long long no = ((b * Do + 2 * zi + (d >> 2)) * Ho + 2 * yi + ((d >> 1) & 1)) * Wo +
                                 2 * xi + (d & 1);
            float f[8];
            unpack8(dy4[no * c8 + cg], f);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += f[k];
        }
        dx4[i] = pack8(s);
    }
}

// dst += src (BF16, elementwise)
__global__ void add_bf16(u16* __restrict__ dst, const u16* __restrict__ src, long long n) {
    GRID_STRIDE(i, n) dst[i] = f2bf(bf2f(dst[i]) + bf2f(src[i]));
}

// FP32 [Cout][27][Cin] master weights -> BF16 forward layout [Cout_p][Kp] and
// the flipped, transposed data-gradient layout [Cin_p][Kp2] (k = tap' Cout + co,
// tap' = 26 - tap); padding rows / tails are zero.
__global__ void pack_weights(u16* __restrict__ wf, u16* __restrict__ wd,
                             const float* __restrict__ w, int cout, int cin, int cout_p,
                             int cin_p, int kp, int kp2) {
    GRID_STRIDE(i, (long long)cout_p * kp) {
        const int co = (int)(i / kp), k = (int)(i - (long long)co * kp);
        const int tap = k / cin, ci = k - tap * cin;
        wf[i] = (co < cout && tap < 27) ? f2bf(w[((long long)co * 27 + tap) * cin + ci]) : (u16)0;
    }
    if (wd) {
        GRID_STRIDE(i, (long long)cin_p * kp2) {
            const int ci = (int)(i / kp2), k = (int)(i - (long long)ci * kp2);
            const int tp = k / cout, co = k - tp * cout;
            wd[i] = (ci < cin && tp < 27)
                        ? f2bf(w[((long long)co * 27 + (26 - tp)) * cin + ci]) : (u16)0;
        }
    }
}

// Residual output and loss gradient: pred = x0 + out[ch 0] (x0 in FP32);
// dout[ch 0] = 2 (pred - target) / N (other channels 0); each block writes its
// partial sum of the squared error / N to part[blockIdx.x] (sum_parts adds
// them in order).
__global__ __launch_bounds__(256) void mse_residual(u16* __restrict__ dout,
                                                    float* __restrict__ part,
                                                    const u16* __restrict__ out, int cout,
                                                    const float* __restrict__ x0,
                                                    const float* __restrict__ target,
                                                    long long nv, float scale) {
    float local = 0.f;
    GRID_STRIDE(n, nv) {
        const float pred = x0[n] + bf2f(out[n * cout]);
        const float e = pred - target[n];
        local += e * e;
        dout[n * cout] = f2bf(2.f * e * scale);
        for (int ch = 1; ch < cout; ++ch) dout[n * cout + ch] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local * scale;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Adam on FP32 master weights.
__global__ void adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v2,
                     const float* __restrict__ g, long long n, float lr, float b1, float b2,
                     float eps, float c1, float c2) {
    GRID_STRIDE(i, n) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v2[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v2[i] = vi;
        w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

// FP32 -> BF16 channels-last input with the channel padded to c (zeros), and
// BF16 channel 0 -> FP32.
__global__ void to_bf16_padded(u16* __restrict__ y, const float* __restrict__ x, long long nv,
                               int c) {
    GRID_STRIDE(i, nv * c) {
        const long long n = i / c;
        y[i] = (i - n * c) == 0 ? f2bf(x[n]) : (u16)0;
    }
}
__global__ void residual_out(float* __restrict__ y, const u16* __restrict__ out, int cout,
                             const float* __restrict__ x0, long long nv) {
    GRID_STRIDE(n, nv) y[n] = x0[n] + bf2f(out[n * cout]);
}

unsigned grid_for(long long n, int per = 256) {
    return (unsigned)std::max<long long>(1, std::min<long long>((n + per - 1) / per, 1 << 20));
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers (device pointers, stream-ordered).
// ---------------------------------------------------------------------------
static Vol mkvol(const int* dims) {
    ARECON_REQUIRE(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0 && dims[3] > 0,
                   "cnn: volume dimensions must be positive");
    return Vol{dims[0], dims[1], dims[2], dims[3]};
}

void cnn_conv3_fwd(void* y, const void* x, const void* w, const float* bias, const void* skip,
                   const int* dims, int cin, int cout, int relu, hipStream_t s) {
    const Vol v = mkvol(dims);
    ARECON_REQUIRE(cin % 8 == 0 && cin >= 8, "cnn: input channels must be a multiple of 8");
    ARECON_REQUIRE(cout % 4 == 0 && cout >= 4 && cout <= 128,
                   "cnn: output channels must be a multiple of 4 in [4, 128]");
    const int kp = (27 * cin + 31) / 32 * 32;
    const int mt = (cout + 15) / 16;
    const char* tiled_env = std::getenv("ARECON_CNN_CONV_TILED");
    if ((cin == 8 || cin == 16 || cin == 32 || cin == 64 || cin == 128) &&
        !(tiled_env && tiled_env[0] == '0')) {
        const int nv = mt == 8 ? 4 : 8;            // rows of 16 voxels per wavefront
        const int tilesx = (v.W + 15) / 16, tilesy = (v.H + nv - 1) / nv, tilesz = (v.D + 3) / 4;
        const long long nblk = (long long)v.B * tilesz * tilesy * tilesx;
        ARECON_REQUIRE(nblk < (1LL << 31), "cnn: volume too large for one launch");
        // the tile's 16-byte slots padded to whole 64-slot staging instructions
        const size_t slots = ((size_t)6 * (nv + 2) * 18 * (std::min(cin, 32) / 8) + 63) / 64 * 64;
        const size_t lds = slots * 16;
        ARECON_REQUIRE((long long)6 * v.H * v.W * cin * 2 < (1LL << 31),
                       "cnn: a 6-plane tile of this volume exceeds the staging descriptor's range");
        const dim3 grid((unsigned)nblk), block(256);
        auto* Y = static_cast<u16*>(y);
        auto* X = static_cast<const u16*>(x);
        auto* Wt = static_cast<const u16*>(w);
        auto* S = static_cast<const u16*>(skip);
        auto go = [&](auto kern) {
            (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
            hipLaunchKernelGGL(kern, grid, block, lds, s, Y, X, Wt, bias, S, v, cout, relu, tilesx,
                               tilesy, tilesz);
        };
#define ARECON_CONV_C(C)                                                                      \
    switch (mt) {                                                                             \
    case 1: go(conv3_fwd_tiled<1, 8, C>); break;                                              \
    case 2: go(conv3_fwd_tiled<2, 8, C>); break;                                              \
    case 4: go(conv3_fwd_tiled<4, 8, C>); break;                                              \
    case 8: go(conv3_fwd_tiled<8, 4, C>); break;                                              \
    default: fail(ARECON_ERR_ARG, "cnn: output channels must be 16, 32, 64 or 128 (or <= 16)"); \
    }
        if (cin == 8) { ARECON_CONV_C(8) }
        else if (cin == 16) { ARECON_CONV_C(16) }
        else if (cin == 32) { ARECON_CONV_C(32) }
        else if (cin == 64) { ARECON_CONV_C(64) }
        else if (cin == 128) { ARECON_CONV_C(128) }
        else fail(ARECON_ERR_ARG, "cnn: tiled convolution takes 8, 16, 32, 64 or 128 input channels");
#undef ARECON_CONV_C
        ARECON_HIP(hipGetLastError());
        return;
    }
    const long long waves = (v.nvox() + 31) / 32;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    auto* Y = static_cast<u16*>(y);
    auto* X = static_cast<const u16*>(x);
    auto* Wt = static_cast<const u16*>(w);
    auto* S = static_cast<const u16*>(skip);
    switch (mt) {
    case 1: hipLaunchKernelGGL(conv3_fwd<1>, grid, block, 0, s, Y, X, Wt, bias, S, v, cin, cout, kp, relu); break;
    case 2: hipLaunchKernelGGL(conv3_fwd<2>, grid, block, 0, s, Y, X, Wt, bias, S, v, cin, cout, kp, relu); break;
    case 4: hipLaunchKernelGGL(conv3_fwd<4>, grid, block, 0, s, Y, X, Wt, bias, S, v, cin, cout, kp, relu); break;
    case 8: hipLaunchKernelGGL(conv3_fwd<8>, grid, block, 0, s, Y, X, Wt, bias, S, v, cin, cout, kp, relu); break;
    default: fail(ARECON_ERR_ARG, "cnn: output channels must be 16, 32, 64 or 128 (or <= 16)");
    }
    ARECON_HIP(hipGetLastError());
}

void cnn_conv3_wgrad(float* dw, const void* dyT, const void* xT3, const int* dims, int cin,
                     int cout, hipStream_t s) {
    const Vol v = mkvol(dims);
    ARECON_REQUIRE(v.W % 8 == 0, "cnn: weight gradient needs W to be a multiple of 8");
    ARECON_REQUIRE(cin % 8 == 0 && cout % 4 == 0, "cnn: channel counts");
    const long long nv = v.nvox();
    const int ct = (cout + 15) / 16, it = (cin + 15) / 16;
    const char* lds_env = std::getenv("ARECON_CNN_WGRAD_LDS");
    if (!(lds_env && lds_env[0] == '0')) {
        // the LDS-staged kernel: NCO output tiles per workgroup, S wavefronts each
        // (S = 3 for one or two output tiles, so a workgroup has 3 or 6 wavefronts)
        const int nco = std::min(4, ct);
        const int cog = (ct + nco - 1) / nco;
        const long long ntiles = (long long)v.B * ((v.D + kWgTz - 1) / kWgTz) *
                                 ((v.H + kWgTy - 1) / kWgTy) * ((v.W + kWgTx - 1) / kWgTx);
        ARECON_REQUIRE(ntiles < (1LL << 31), "cnn: volume too large for one launch");
        long long nsplit = std::max<long long>(
            1, std::min<long long>(ntiles, (1024 + (long long)it * cog - 1) / ((long long)it * cog)));
        const long long per = (ntiles + nsplit - 1) / nsplit;
        nsplit = (ntiles + per - 1) / per;
        const long long nw = 27LL * cout * cin;
        DevBuf part(sizeof(float) * (size_t)(nw * nsplit), s);
        const dim3 grid((unsigned)nsplit, (unsigned)it, (unsigned)cog);
        auto go = [&](auto kern, int threads) {
            (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      kWgLds);
            hipLaunchKernelGGL(kern, grid, dim3(threads), kWgLds, s, part.as<float>(),
                               static_cast<const u16*>(dyT), static_cast<const u16*>(xT3), v, cin,
                               cout, (int)per, (int)ntiles);
        };
        switch (nco) {
        case 4: go(conv3_wgrad_lds<4, 1>, 256); break;
        case 3: go(conv3_wgrad_lds<3, 1>, 192); break;
        case 2: go(conv3_wgrad_lds<2, 3>, 384); break;
        default: go(conv3_wgrad_lds<1, 3>, 192); break;
        }
        ARECON_HIP(hipGetLastError());
        // the splits' partial sums, added in a fixed order: in groups of 32, then the groups
        // (one thread per weight adding hundreds of splits in sequence was latency-bound)
        const int ngroups = (int)((nsplit + 31) / 32);
        DevBuf part2(sizeof(float) * (size_t)(nw * ngroups), s);
        hipLaunchKernelGGL(sum_part_groups, dim3(grid_for(nw), (unsigned)ngroups), dim3(256), 0,
                           s, part2.as<float>(), part.as<float>(), nw, (int)nsplit);
        ARECON_HIP(hipGetLastError());
        hipLaunchKernelGGL(sum_parts, dim3(grid_for(nw)), dim3(256), 0, s, dw, part2.as<float>(),
                           nw, ngroups);
        ARECON_HIP(hipGetLastError());
        return;
    }
    const long long tiles = 27LL * ct * it;
    // enough waves to fill the GPU: split the voxel range (multiples of 32)
    long long nsplit = std::max<long long>(1, std::min<long long>((nv + 2047) / 2048,
                                                                  (8192 + tiles - 1) / tiles));
    long long per = (nv + nsplit - 1) / nsplit;
    per = (per + 31) / 32 * 32;
    nsplit = (nv + per - 1) / per;
    const long long waves = tiles * nsplit;
    const long long nw = 27LL * cout * cin;
    DevBuf part(sizeof(float) * (size_t)(nw * nsplit), s);
    hipLaunchKernelGGL(conv3_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s,
                       part.as<float>(), static_cast<const u16*>(dyT),
                       static_cast<const u16*>(xT3), v, cin, cout, per, (int)nsplit);
    ARECON_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_parts, dim3(grid_for(nw)), dim3(256), 0, s, dw, part.as<float>(), nw,
                       (int)nsplit);
    ARECON_HIP(hipGetLastError());
}

void cnn_relu_bwd(void* dyp, void* dypT, float* db, const void* dy, const void* y,
                  const int* dims, int c, int relu, hipStream_t s) {
    const Vol v = mkvol(dims);
    if (v.W % 8 == 0 && c % 8 == 0 && c <= 128)
        hipLaunchKernelGGL(relu_bwd_lds, dim3((unsigned)((v.nvox() + kLtV - 1) / kLtV)), dim3(256), 0,
                           s, static_cast<u16*>(dyp), static_cast<u16*>(dypT),
                           static_cast<const u16*>(dy), static_cast<const u16*>(y), v.nvox(), c,
                           relu);
    else if (v.W % 8 == 0)
        hipLaunchKernelGGL(relu_bwd8, dim3(grid_for(v.nvox() / 8 * c)), dim3(256), 0, s,
                           static_cast<u16*>(dyp), static_cast<u16*>(dypT),
                           static_cast<const u16*>(dy), static_cast<const u16*>(y), v.nvox(), c,
                           relu);
    else
        hipLaunchKernelGGL(relu_bwd, dim3(grid_for(v.nvox() * c)), dim3(256), 0, s,
                           static_cast<u16*>(dyp), static_cast<u16*>(dypT),
                           static_cast<const u16*>(dy), static_cast<const u16*>(y), v.nvox(), c,
                           relu);
    ARECON_HIP(hipGetLastError());
    if (db) {
        const unsigned bx = (unsigned)std::max<long long>(1, std::min<long long>(
                                                                 (v.nvox() + 4095) / 4096, 256));
        DevBuf part(sizeof(float) * (size_t)bx * (size_t)c, s);
        hipLaunchKernelGGL(channel_sum, dim3(bx, (unsigned)c), dim3(256), 0, s,
                           part.as<float>(), static_cast<const u16*>(dypT), v.nvox());
        ARECON_HIP(hipGetLastError());
        hipLaunchKernelGGL(sum_parts, dim3(1), dim3(256), 0, s, db, part.as<float>(),
                           (long long)c, (int)bx);
        ARECON_HIP(hipGetLastError());
    }
}

void cnn_channel_major3(void* xT3, const void* x, const int* dims, int c, hipStream_t s) {
    const Vol v = mkvol(dims);
    if (v.W % 8 == 0 && c % 8 == 0 && c <= 128)
        hipLaunchKernelGGL(to_channel_major3_lds, dim3((unsigned)((v.nvox() + kLtV - 1) / kLtV)),
                           dim3(256), 0, s, static_cast<u16*>(xT3), static_cast<const u16*>(x), v,
                           c);
    else if (v.W % 8 == 0)
        hipLaunchKernelGGL(to_channel_major3_8, dim3(grid_for(v.nvox() / 8 * c)), dim3(256), 0, s,
                           static_cast<u16*>(xT3), static_cast<const u16*>(x), v, c);
    else
        hipLaunchKernelGGL(to_channel_major3, dim3(grid_for(v.nvox() * c)), dim3(256), 0, s,
                           static_cast<u16*>(xT3), static_cast<const u16*>(x), v, c);
    ARECON_HIP(hipGetLastError());
}

void cnn_pool(void* y, const void* x, const int* dims_out, int c, hipStream_t s) {
    const Vol v = mkvol(dims_out);
    ARECON_REQUIRE(c % 8 == 0, "cnn: pooling and upsampling take a multiple of 8 channels");
    hipLaunchKernelGGL(avgpool2, dim3(grid_for(v.nvox() * (c / 8))), dim3(256), 0, s,
                       static_cast<u16*>(y), static_cast<const u16*>(x), v, c);
    ARECON_HIP(hipGetLastError());
}
void cnn_pool_bwd(void* dx, const void* dy, const int* dims_in, int c, int accumulate,
                  hipStream_t s) {
    const Vol v = mkvol(dims_in);
    ARECON_REQUIRE(v.D % 2 == 0 && v.H % 2 == 0 && v.W % 2 == 0, "cnn: pooling needs even dims");
    ARECON_REQUIRE(c % 8 == 0, "cnn: pooling and upsampling take a multiple of 8 channels");
    hipLaunchKernelGGL(avgpool2_bwd, dim3(grid_for(v.nvox() * (c / 8))), dim3(256), 0, s,
                       static_cast<u16*>(dx), static_cast<const u16*>(dy), v, c, accumulate);
    ARECON_HIP(hipGetLastError());
}
void cnn_upsample(void* y, const void* x, const int* dims_out, int c, hipStream_t s) {
    const Vol v = mkvol(dims_out);
    ARECON_REQUIRE(c % 8 == 0, "cnn: pooling and upsampling take a multiple of 8 channels");
    hipLaunchKernelGGL(upsample2, dim3(grid_for(v.nvox() * (c / 8))), dim3(256), 0, s,
                       static_cast<u16*>(y), static_cast<const u16*>(x), v, c);
    ARECON_HIP(hipGetLastError());
}
void cnn_upsample_bwd(void* dx, const void* dy, const int* dims_in, int c, hipStream_t s) {
    const Vol v = mkvol(dims_in);
    ARECON_REQUIRE(c % 8 == 0, "cnn: pooling and upsampling take a multiple of 8 channels");
    hipLaunchKernelGGL(upsample2_bwd, dim3(grid_for(v.nvox() * (c / 8))), dim3(256), 0, s,
                       static_cast<u16*>(dx), static_cast<const u16*>(dy), v, c);
    ARECON_HIP(hipGetLastError());
}
void cnn_add(void* dst, const void* src, long long n, hipStream_t s) {
    hipLaunchKernelGGL(add_bf16, dim3(grid_for(n)), dim3(256), 0, s, static_cast<u16*>(dst),
                       static_cast<const u16*>(src), n);
    ARECON_HIP(hipGetLastError());
}
void cnn_pack_weights(void* wf, void* wd, const float* w, int cout, int cin, hipStream_t s) {
    ARECON_REQUIRE(cin % 8 == 0 && cout % 4 == 0, "cnn: channel counts");
    const int cout_p = (cout + 15) / 16 * 16, cin_p = (cin + 15) / 16 * 16;
    const int kp = (27 * cin + 31) / 32 * 32, kp2 = (27 * cout + 31) / 32 * 32;
    hipLaunchKernelGGL(pack_weights, dim3(grid_for((long long)std::max(cout_p * kp, cin_p * kp2))),
                       dim3(256), 0, s, static_cast<u16*>(wf), static_cast<u16*>(wd), w, cout,
                       cin, cout_p, cin_p, kp, kp2);
    ARECON_HIP(hipGetLastError());
}
void cnn_mse_residual(void* dout, float* loss, const void* out, int cout, const float* x0,
                      const float* target, long long nv, hipStream_t s) {
    const unsigned blocks = std::min<unsigned>(grid_for(nv), 1024u);
    DevBuf part(sizeof(float) * blocks, s);
    hipLaunchKernelGGL(mse_residual, dim3(blocks), dim3(256), 0, s, static_cast<u16*>(dout),
                       part.as<float>(), static_cast<const u16*>(out), cout, x0, target, nv,
                       1.f / (float)nv);
    ARECON_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_parts, dim3(1), dim3(256), 0, s, loss, part.as<float>(), 1LL,
                       (int)blocks);
    ARECON_HIP(hipGetLastError());
}
void cnn_adam(float* w, float* m, float* v, const float* g, long long n, float lr, float b1,
              float b2, float eps, int step, hipStream_t s) {
    ARECON_REQUIRE(step >= 1, "cnn: adam step counts from 1");
    const float c1 = 1.f / (1.f - std::pow(b1, (float)step));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)step));
    hipLaunchKernelGGL(adam, dim3(grid_for(n)), dim3(256), 0, s, w, m, v, g, n, lr, b1, b2, eps,
                       c1, c2);
    ARECON_HIP(hipGetLastError());
}
void cnn_to_bf16(void* y, const float* x, long long nv, int c, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_padded, dim3(grid_for(nv * c)), dim3(256), 0, s,
                       static_cast<u16*>(y), x, nv, c);
    ARECON_HIP(hipGetLastError());
}
void cnn_residual_out(float* y, const void* out, int cout, const float* x0, long long nv,
                      hipStream_t s) {
    hipLaunchKernelGGL(residual_out, dim3(grid_for(nv)), dim3(256), 0, s, y,
                       static_cast<const u16*>(out), cout, x0, nv);
    ARECON_HIP(hipGetLastError());
}

}  // namespace arecon

// ---------------------------------------------------------------------------
// C ABI (include/arecon_cnn.h).
// ---------------------------------------------------------------------------
extern "C" {

using namespace arecon;

int arecon_cnn_conv3(void* y, const void* x, const void* w, const float* bias, const void* skip,
                     const int* dims, int cin, int cout, int relu, void* stream) {
    return guarded([&] {
        cnn_conv3_fwd(y, x, w, bias, skip, dims, cin, cout, relu,
                      static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_conv3_wgrad(float* dw, const void* dyT, const void* xT3, const int* dims,
                           int cin, int cout, void* stream) {
    return guarded([&] {
        cnn_conv3_wgrad(dw, dyT, xT3, dims, cin, cout, static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_relu_bwd(void* dyp, void* dypT, float* db, const void* dy, const void* y,
                        const int* dims, int c, int relu, void* stream) {
    return guarded([&] {
        cnn_relu_bwd(dyp, dypT, db, dy, y, dims, c, relu, static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_channel_major3(void* xT3, const void* x, const int* dims, int c, void* stream) {
    return guarded([&] { cnn_channel_major3(xT3, x, dims, c, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_pool(void* y, const void* x, const int* dims_out, int c, void* stream) {
    return guarded([&] { cnn_pool(y, x, dims_out, c, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_pool_bwd(void* dx, const void* dy, const int* dims_in, int c, int accumulate,
                        void* stream) {
    return guarded([&] {
        cnn_pool_bwd(dx, dy, dims_in, c, accumulate, static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_upsample(void* y, const void* x, const int* dims_out, int c, void* stream) {
    return guarded([&] { cnn_upsample(y, x, dims_out, c, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_upsample_bwd(void* dx, const void* dy, const int* dims_in, int c, void* stream) {
    return guarded([&] { cnn_upsample_bwd(dx, dy, dims_in, c, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_add(void* dst, const void* src, long long n, void* stream) {
    return guarded([&] { cnn_add(dst, src, n, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_pack_weights(void* wf, void* wd, const float* w, int cout, int cin, void* stream) {
    return guarded([&] { cnn_pack_weights(wf, wd, w, cout, cin, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_mse_residual(void* dout, float* loss, const void* out, int cout, const float* x0,
                            const float* target, long long nv, void* stream) {
    return guarded([&] {
        cnn_mse_residual(dout, loss, out, cout, x0, target, nv, static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_adam(float* w, float* m, float* v, const float* g, long long n, float lr, float b1,
                    float b2, float eps, int step, void* stream) {
    return guarded([&] {
        cnn_adam(w, m, v, g, n, lr, b1, b2, eps, step, static_cast<hipStream_t>(stream));
    });
}
int arecon_cnn_to_bf16(void* y, const float* x, long long nv, int c, void* stream) {
    return guarded([&] { cnn_to_bf16(y, x, nv, c, static_cast<hipStream_t>(stream)); });
}
int arecon_cnn_residual_out(float* y, const void* out, int cout, const float* x0, long long nv,
                            void* stream) {
    return guarded([&] { cnn_residual_out(y, out, cout, x0, nv, static_cast<hipStream_t>(stream)); });
}

}  // extern "C"
