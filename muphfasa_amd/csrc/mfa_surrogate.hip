// mfa_surrogate.hip -- BF16 MFMA kernels of the learned warm start: a
// convolutional encoder-decoder (3x3 convolutions, stride 1 or 2, ReLU,
// nearest upsampling, additive skips) trained and evaluated on gfx950.
//
// Activations are NHWC bfloat16; weights [Cout][3][3][Cin] bfloat16 copied
// from an FP32 master each step (plus the [3][3][Cin][Cout] transpose used by
// the data gradient).  Every convolution is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   forward        Y[p][co]  = sum_k  X(p,k)  W[co][k],  k = tap*Cin + ci
//   data grad      dX[p][ci] = sum_k dY(p,k) WT[k][ci],  k = tap*Cout + co
//   weight grad    dW[co][k] = sum_p dY[p][co] X(p,k)   (split over pixels;
//                  operands transposed in registers by a selector MFMA)
// MFMA operand layout (wave64): lane l holds A[row l%16][k 8*(l/16)..+7] and
// B[k 8*(l/16)..+7][col l%16]; the 16x16 FP32 result lands as
// C[4*(l/16)+i][l%16], i = 0..3.  A wave computes a 32x32 tile with 2x2
// MFMAs per 32-deep slice of k.  Channel counts are multiples of 8 so each
// lane's 8 consecutive k of every operand load is one 16-byte load.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <string>
#include <type_traits>

#include "../../include/mfa_surrogate.h"

namespace mfa_nn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static thread_local std::string g_nn_err;
static int nn_fail(int code, const std::string& m)
{
    g_nn_err = m;
    return code;
}
// Tensors the kernels read or write 16 bytes at a time.
static bool aligned16(std::initializer_list<const void*> ptrs)
{
    for (const void* p : ptrs)
        if ((uintptr_t)p % 16) return false;
    return true;
}

#define NN_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return nn_fail(-3, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
    return z;
}

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *(const bf16x8*)p; }

// 16 bytes at element offset off (32-bit) from a wave-uniform base.
__device__ __forceinline__ bf16x8 load8_at(const bf16* base, unsigned off)
{
    return *(const bf16x8*)((const char*)base + off * 2u);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct ConvShape {
    int N, H, W, Cin, Cout, stride, Ho, Wo;
};

// ---------------------------------------------------------------- forward --
// Block: 4 waves.  With KS = 1, WC waves along output channels (1 or 2) and
// 4/WC along pixels; a wave computes 16*TI pixels x 16*TJ channels (TI x TJ
// MFMA tiles per 32-deep slice of k).  TI = 4, TJ = 1 serves layers with 16
// or fewer output channels without idle MFMA columns.  With KS = 4 (WC = 1)
// the four waves share one 16*TI x 16*TJ tile and each takes every fourth
// slice of k; their partial sums meet in LDS.  That serves the deep layers
// on small maps, whose few output tiles would otherwise leave most CUs idle
// while each wave walks all of k = 9 * Cin (up to 1080).  The next slice's
// operands are loaded before the current slice's MFMAs.
template <int TI, int TJ>
struct ConvAcc {
    f32x4 v[TI][TJ];
};

// Sum of the KS waves' accumulators into wave 0's (KS = 4 only).
template <int KS, int TI, int TJ>
__device__ __forceinline__ void reduce_waves(f32x4 (&acc)[TI][TJ], int wave, int lane)
{
    if constexpr (KS > 1) {
        __shared__ ConvAcc<TI, TJ> red[KS - 1][64];
        if (wave > 0) {
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) red[wave - 1][lane].v[i][j] = acc[i][j];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int w = 0; w < KS - 1; ++w)
#pragma unroll
                for (int i = 0; i < TI; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j) acc[i][j] += red[w][lane].v[i][j];
        }
    }
}

// The k loop of one wave: slices 32 wk, 32 (wk + KS), ...  With KS > 1
// (few waves per CU, each walking a long k) the next slice's operands are
// loaded before the current slice's MFMAs; with KS = 1 the many resident
// waves hide the loads and the second set of operand registers would cost
// occupancy (measured: 36 -> 41 us for a 16-channel 256^2 layer).
template <int KS, int TI, int TJ, class Load, class Mma>
__device__ __forceinline__ void k_loop(int K, int wk, Load&& load, Mma&& mma)
{
    constexpr int STEP = 32 * KS;
    if constexpr (KS == 1) {
        for (int k0 = 0; k0 < K; k0 += 32) {
            bf16x8 a[TI], b[TJ];
            load(k0, a, b);
            mma(a, b);
        }
    } else {
        bf16x8 a0[TI], b0[TJ], a1[TI], b1[TJ];
        int k0 = 32 * wk;
        if (k0 < K) load(k0, a0, b0);
        for (; k0 < K; k0 += 2 * STEP) {
            if (k0 + STEP < K) load(k0 + STEP, a1, b1);
            mma(a0, b0);
            if (k0 + STEP >= K) break;
            if (k0 + 2 * STEP < K) load(k0 + 2 * STEP, a0, b0);
            mma(a1, b1);
        }
    }
}

// Epilogue of one wave's tile (16 TI pixels x 16 TJ channels, FP32, MFMA
// layout: lane l holds rows 4(l/16)+e, column l%16 of each 16 x 16 tile).
// The tile goes through the wave's own LDS slab [pixel][channel] and comes
// back as 8 consecutive channels of one pixel per lane and vector, so the
// global accesses of the epilogue are 16 bytes per lane instead of 2.
// emit(pixel in tile, first channel in tile, float[8]) runs once per vector.
template <int TI, int TJ>
struct EpiSlab {
    static constexpr int NP = 16 * TI, NC = 16 * TJ, LD = NC + 4;
    float v[NP * LD];
};

template <int TI, int TJ, class Emit>
__device__ __forceinline__ void tile_out(const f32x4 (&acc)[TI][TJ], EpiSlab<TI, TJ>& slab, int lane,
                                         Emit&& emit)
{
    using S = EpiSlab<TI, TJ>;
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) slab.v[(16 * i + 4 * g + e) * S::LD + 16 * j + r] = acc[i][j][e];
    // lanes of this wave read what others wrote
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int CV = S::NC / 8, NV = S::NP * CV;
#pragma unroll
    for (int v = lane; v < NV; v += 64) {
        const int px = v / CV, c8 = 8 * (v % CV);
        const f32x4 lo = *(const f32x4*)&slab.v[px * S::LD + c8];
        const f32x4 hi = *(const f32x4*)&slab.v[px * S::LD + c8 + 4];
        const float o[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        emit(px, c8, o);
    }
}

// Per 8-column group of k (k = tap * C + c): the tap, and the element
// offset of the group's source from the pixel's own base offset in the
// gathered tensor.  Built once per block in LDS, so the k loop needs no
// division by the channel count.
struct KEntry {
    int delta;
    short ky, kx;
};
constexpr int kMaxKGroups = 9 * 128 / 8;    // k = 9 * C, C <= 128

template <int WC, int TI, int TJ, int KS>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                  const float* __restrict__ bias, bf16* __restrict__ y,
                                                  float* __restrict__ yf, const ConvShape s, const int relu)
{
    static_assert(KS == 1 || WC == 1, "split k: the waves share one tile");
    constexpr int WP = KS > 1 ? 1 : 4 / WC;
    __shared__ KEntry tbl[kMaxKGroups];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wp = KS > 1 ? 0 : wave % WP, wc = KS > 1 ? 0 : wave / WP, wk = KS > 1 ? wave : 0;
    const unsigned P = (unsigned)s.N * (unsigned)s.Ho * (unsigned)s.Wo;
    const unsigned p0 = blockIdx.x * (16 * TI * WP) + 16 * TI * wp;
    const int c0 = blockIdx.y * (16 * TJ * WC) + 16 * TJ * wc;
    const int K = 9 * s.Cin;
    const int r = lane & 15, kg = lane >> 4;
    for (int t = threadIdx.x; t < K / 8; t += 256) {
        const int tap = (8 * t) / s.Cin, ci = (8 * t) % s.Cin;
        const int ky = tap / 3, kx = tap % 3;
        tbl[t] = KEntry{((ky - 1) * s.W + (kx - 1)) * s.Cin + ci, (short)ky, (short)kx};
    }
    // the A rows this lane feeds: source base offset and corner (y s - 1, x s - 1)
    unsigned pb[TI];
    int ys[TI], xs[TI];
    bool pv[TI];
#pragma unroll
    for (int i = 0; i < TI; ++i) {
        const unsigned p = p0 + 16 * i + r;
        pv[i] = p < P;
        const unsigned pp = pv[i] ? p : 0u, t = pp / (unsigned)s.Wo;
        const unsigned xo = pp - t * (unsigned)s.Wo, yo = t % (unsigned)s.Ho, n = t / (unsigned)s.Ho;
        pb[i] = ((n * (unsigned)s.H + yo * s.stride) * (unsigned)s.W + xo * s.stride) * (unsigned)s.Cin;
        ys[i] = (int)yo * s.stride - 1;
        xs[i] = (int)xo * s.stride - 1;
    }
    unsigned wrow[TJ];
    bool cv[TJ];
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
        const int co = c0 + 16 * j + r;
        cv[j] = co < s.Cout;
        wrow[j] = cv[j] ? (unsigned)co * (unsigned)K : 0u;
    }
    __syncthreads();
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    auto load = [&](int k0, bf16x8 (&a)[TI], bf16x8 (&b)[TJ]) {
        const int k = k0 + 8 * kg;
        const bool kv = k < K;
        const KEntry e = tbl[kv ? k / 8 : 0];
#pragma unroll
        for (int i = 0; i < TI; ++i) {
            const bool ok = kv && pv[i] && (unsigned)(ys[i] + e.ky) < (unsigned)s.H &&
                            (unsigned)(xs[i] + e.kx) < (unsigned)s.W;
            a[i] = ok ? load8_at(x, pb[i] + e.delta) : zero8();
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j) b[j] = (kv && cv[j]) ? load8_at(w, wrow[j] + k) : zero8();
    };
    auto mma = [&](const bf16x8 (&a)[TI], const bf16x8 (&b)[TJ]) {
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    };
    k_loop<KS, TI, TJ>(K, wk, load, mma);
    reduce_waves<KS>(acc, wave, lane);
    if (KS > 1 && wave != 0) return;
    __shared__ EpiSlab<TI, TJ> slab[KS > 1 ? 1 : 4];
    tile_out(acc, slab[KS > 1 ? 0 : wave], lane, [&](int px, int c8, const float (&o)[8]) {
        const unsigned p = p0 + px;
        const int co = c0 + c8;
        if (p >= P || co >= s.Cout) return;
        const long long idx = (long long)p * s.Cout + co;
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v[e] = o[e] + (bias ? bias[co + e] : 0.f);
            if (relu) v[e] = v[e] > 0.f ? v[e] : 0.f;
        }
        if (y) {
            bf16x8 b;
#pragma unroll
            for (int e = 0; e < 8; ++e) b[e] = (bf16)v[e];
            *(bf16x8*)(y + idx) = b;
        }
        if (yf) {
            *(f32x4*)(yf + idx) = f32x4{v[0], v[1], v[2], v[3]};
            *(f32x4*)(yf + idx + 4) = f32x4{v[4], v[5], v[6], v[7]};
        }
    });
}

// ---------------------------------------------------------- data gradient --
// dX[p_in][ci] = sum over (tap, co) dY[p_out][co] * WT[tap][ci][co], with
// p_out = ((y_in + 1 - ky) / stride, (x_in + 1 - kx) / stride) when exact.
// Epilogue: dX = (accumulate ? dX : 0) + value * (mask ? (mask > 0) : 1).
// Wave tiles, split k and the tap table as in the forward kernel (pixels x
// input channels); with stride 1 the source is a fixed offset from the
// pixel, with stride 2 it is found per pixel and tap.
template <int WC, int TI, int TJ, int KS, bool S1>
__global__ __launch_bounds__(256) void k_conv_bwd_data(const bf16* __restrict__ dy, const bf16* __restrict__ wt,
                                                       bf16* __restrict__ dx, const bf16* __restrict__ mask,
                                                       const ConvShape s, const int accumulate)
{
    static_assert(KS == 1 || WC == 1, "split k: the waves share one tile");
    constexpr int WP = KS > 1 ? 1 : 4 / WC;
    __shared__ KEntry tbl[kMaxKGroups];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wp = KS > 1 ? 0 : wave % WP, wc = KS > 1 ? 0 : wave / WP, wk = KS > 1 ? wave : 0;
    const unsigned P = (unsigned)s.N * (unsigned)s.H * (unsigned)s.W;
    const unsigned p0 = blockIdx.x * (16 * TI * WP) + 16 * TI * wp;
    const int c0 = blockIdx.y * (16 * TJ * WC) + 16 * TJ * wc;
    const int K = 9 * s.Cout;
    const int r = lane & 15, kg = lane >> 4;
    for (int t = threadIdx.x; t < K / 8; t += 256) {
        const int tap = (8 * t) / s.Cout, co = (8 * t) % s.Cout;
        const int ky = tap / 3, kx = tap % 3;
        // stride 1: source (y + 1 - ky, x + 1 - kx); stride 2: the channel only
        const int d = S1 ? ((1 - ky) * s.Wo + (1 - kx)) * s.Cout + co : co;
        tbl[t] = KEntry{d, (short)ky, (short)kx};
    }
    unsigned pb[TI];
    int py[TI], px[TI];
    bool pv[TI];
#pragma unroll
    for (int i = 0; i < TI; ++i) {
        const unsigned p = p0 + 16 * i + r;
        pv[i] = p < P;
        const unsigned pp = pv[i] ? p : 0u, t = pp / (unsigned)s.W;
        const unsigned xi = pp - t * (unsigned)s.W, yi = t % (unsigned)s.H, n = t / (unsigned)s.H;
        py[i] = (int)yi;
        px[i] = (int)xi;
        // stride 1: the pixel's own offset in dY; stride 2: its image's
        pb[i] = S1 ? ((n * (unsigned)s.Ho + yi) * (unsigned)s.Wo + xi) * (unsigned)s.Cout
                   : n * (unsigned)s.Ho * (unsigned)s.Wo * (unsigned)s.Cout;
    }
    unsigned wcol[TJ];
    bool cv[TJ];
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
        const int ci = c0 + 16 * j + r;
        cv[j] = ci < s.Cin;
        wcol[j] = cv[j] ? (unsigned)ci * (unsigned)s.Cout : 0u;
    }
    __syncthreads();
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto load = [&](int k0, bf16x8 (&a)[TI], bf16x8 (&b)[TJ]) {
        const int k = k0 + 8 * kg;
        const bool kv = k < K;
        const KEntry e = tbl[kv ? k / 8 : 0];
#pragma unroll
        for (int i = 0; i < TI; ++i) {
            const int ty = py[i] + 1 - e.ky, tx = px[i] + 1 - e.kx;
            if constexpr (S1) {
                const bool ok = kv && pv[i] && (unsigned)ty < (unsigned)s.Ho && (unsigned)tx < (unsigned)s.Wo;
                a[i] = ok ? load8_at(dy, pb[i] + e.delta) : zero8();
            } else {
                const bool ok = kv && pv[i] && ((ty | tx) & 1) == 0 && (unsigned)(ty >> 1) < (unsigned)s.Ho &&
                                (unsigned)(tx >> 1) < (unsigned)s.Wo;
                const unsigned off = pb[i] + ((unsigned)(ty >> 1) * (unsigned)s.Wo + (unsigned)(tx >> 1)) *
                                                 (unsigned)s.Cout + e.delta;
                a[i] = ok ? load8_at(dy, off) : zero8();
            }
        }
        // WT[tap][ci][co] at (tap * Cin + ci) * Cout + co = wcol + tap * Cin * Cout + co
        const unsigned tb = (unsigned)(e.ky * 3 + e.kx) * (unsigned)s.Cin * (unsigned)s.Cout +
                            (unsigned)(S1 ? (k - (e.ky * 3 + e.kx) * s.Cout) : e.delta);
#pragma unroll
        for (int j = 0; j < TJ; ++j) b[j] = (kv && cv[j]) ? load8_at(wt, wcol[j] + tb) : zero8();
    };
    auto mma = [&](const bf16x8 (&a)[TI], const bf16x8 (&b)[TJ]) {
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    };
    k_loop<KS, TI, TJ>(K, wk, load, mma);
    reduce_waves<KS>(acc, wave, lane);
    if (KS > 1 && wave != 0) return;
    __shared__ EpiSlab<TI, TJ> slab[KS > 1 ? 1 : 4];
    tile_out(acc, slab[KS > 1 ? 0 : wave], lane, [&](int pxl, int c8, const float (&o)[8]) {
        const unsigned p = p0 + pxl;
        const int ci = c0 + c8;
        if (p >= P || ci >= s.Cin) return;
        const long long idx = (long long)p * s.Cin + ci;
        const bf16x8 m = mask ? load8(mask + idx) : zero8();
        const bf16x8 a = accumulate ? load8(dx + idx) : zero8();
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float v = o[e];
            if (mask && !((float)m[e] > 0.f)) v = 0.f;
            if (accumulate) v += (float)a[e];
            b[e] = (bf16)v;
        }
        *(bf16x8*)(dx + idx) = b;
    });
}

// -------------------------------------------------------- weight gradient --
// dW[co][k] = sum_p dY[p][co] X(p,k) and dbias[co] = sum_p dY[p][co]: the
// MFMA's reduction runs over pixels, so each lane must hold 8 pixels of one
// channel, while NHWC storage gives 8 channels of one pixel per 16-byte load.
// The operands are transposed in registers by an MFMA against a selector
// matrix: with lane l holding D[pixel l%16][channel 8(l/16)..+7] (16 pixels
// x 32 channels, one natural load), D x S_h, S_h[k][n] = (k == n + 16h),
// leaves lane l holding channel 16h + l%16 of pixels 4(l/16)..+3 in FP32.
// Each result is one bf16 value times 1.0 plus exact zeros, so it is the
// stored value itself.  Two such results (pixels 0-15 and 16-31 of a chunk)
// give a lane the 8 pixels of one channel that the reduction MFMA takes; the
// k index of that MFMA names pixel 4(l/16)+j (j < 4) or 16+4(l/16)+j-4, the
// same order for both operands, so the sum covers each pixel once.
//
// Block: 4 waves over one tile set (16*MB output channels x 32*NB columns of
// k) and one range of 32-pixel chunks (blockIdx.z), the waves taking every
// fourth chunk; the waves' FP32 partials are summed in LDS and added to dW
// (and dbias) atomically, once per block.
__device__ __forceinline__ bf16x8 selector(int lane, int h)
{
    bf16x8 sel;
    const int g = lane >> 4, n = lane & 15;
#pragma unroll
    for (int j = 0; j < 8; ++j) sel[j] = (bf16)((8 * g + j == n + 16 * h) ? 1.0f : 0.0f);
    return sel;
}

__device__ __forceinline__ bf16x8 pack8(const f32x4 a, const f32x4 b)
{
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        v[e] = (bf16)a[e];
        v[4 + e] = (bf16)b[e];
    }
    return v;
}

// Blocks per launch the pixel split aims at, and whether a wave loads its
// next chunk before computing the current one.
constexpr int kWgBlocks = 512;
constexpr bool kWgPrefetch = false;

// ROW: the output width is a multiple of 32, so a chunk's 32 pixels lie in
// one row and its (image, row, first column) are wave-uniform, computed once
// per chunk in scalar registers.  Offsets are 32-bit (the host checks that
// the tensors hold fewer than 2^31 elements).
template <int MB, int NB, bool ROW>
__global__ __launch_bounds__(256) void k_conv_bwd_weight(const bf16* __restrict__ x,
                                                         const bf16* __restrict__ dy,
                                                         float* __restrict__ dw, float* __restrict__ db,
                                                         const ConvShape s, const int chunks_per_block)
{
    constexpr int MJ = (MB + 1) / 2;          // 32-channel loads of dY per half chunk
    constexpr int NT = 2 * NB;                // 16-column tiles of k
    __shared__ f32x4 red[3][MB][NT][64];
    __shared__ float bred[4][MB][16];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, r = lane & 15;
    const int co0 = blockIdx.x * 16 * MB, kc0 = blockIdx.y * 32 * NB;
    const int K = 9 * s.Cin;
    const unsigned P = (unsigned)s.N * (unsigned)s.Ho * (unsigned)s.Wo;
    const unsigned nchunks = (P + 31) / 32;
    const unsigned cbeg = blockIdx.z * (unsigned)chunks_per_block;
    const unsigned cend = cbeg + chunks_per_block < nchunks ? cbeg + chunks_per_block : nchunks;
    // this lane's 8 columns of each 32-column block: tap, and the element
    // offset of the column's source from the pixel (y s, x s) of X
    int ky[NB], kx[NB], dk[NB];
    bool kv[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int kk = kc0 + 32 * b + 8 * g;
        kv[b] = kk < K;
        const int tap = kv[b] ? kk / s.Cin : 0, ci = kv[b] ? kk % s.Cin : 0;
        ky[b] = tap / 3;
        kx[b] = tap % 3;
        dk[b] = ((ky[b] - 1) * s.W + (kx[b] - 1)) * s.Cin + ci;
    }
    const bf16x8 sel0 = selector(lane, 0), sel1 = selector(lane, 1);
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4 acc[MB][NT];
    float bsum[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        bsum[m] = 0.f;
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = z4;
    }
    // operands of one chunk: dY (32 channels per load) and X columns, for
    // pixels 0-15 (half 0) and 16-31 (half 1) of the chunk
    auto load = [&](unsigned c, bf16x8 (&yv)[MJ][2], bf16x8 (&xv)[NB][2]) {
        unsigned n0 = 0, yo0 = 0, xo0 = 0;
        if constexpr (ROW) {
            const unsigned cpr = (unsigned)s.Wo / 32, row = c / cpr;
            xo0 = (c - row * cpr) * 32;
            yo0 = row % (unsigned)s.Ho;
            n0 = row / (unsigned)s.Ho;
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const unsigned p = 32 * c + 16 * half + r;
            bool pv;
            unsigned n, yo, xo;
            if constexpr (ROW) {
                pv = true;
                n = n0;
                yo = yo0;
                xo = xo0 + 16 * half + r;
            } else {
                pv = p < P;
                const unsigned pp = pv ? p : 0u, t = pp / (unsigned)s.Wo;
                xo = pp - t * (unsigned)s.Wo;
                yo = t % (unsigned)s.Ho;
                n = t / (unsigned)s.Ho;
            }
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                const int co = co0 + 32 * j + 8 * g;
                yv[j][half] = (pv && co < s.Cout) ? load8_at(dy, p * (unsigned)s.Cout + co) : zero8();
            }
            const int ys = (int)yo * s.stride - 1, xs = (int)xo * s.stride - 1;
            const unsigned xb = ((n * (unsigned)s.H + yo * s.stride) * (unsigned)s.W + xo * s.stride) *
                                (unsigned)s.Cin;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const bool ok = pv && kv[b] && (unsigned)(ys + ky[b]) < (unsigned)s.H &&
                                (unsigned)(xs + kx[b]) < (unsigned)s.W;
                xv[b][half] = ok ? load8_at(x, xb + dk[b]) : zero8();
            }
        }
    };
    auto compute = [&](const bf16x8 (&yv)[MJ][2], const bf16x8 (&xv)[NB][2]) {
        bf16x8 opy[MB], opx[NT];
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const bf16x8 sl = (m & 1) ? sel1 : sel0;
            const f32x4 t0 = mfma(yv[m / 2][0], sl, z4), t1 = mfma(yv[m / 2][1], sl, z4);
            opy[m] = pack8(t0, t1);
            bsum[m] += ((t0[0] + t0[1]) + (t0[2] + t0[3])) + ((t1[0] + t1[1]) + (t1[2] + t1[3]));
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 sl = (t & 1) ? sel1 : sel0;
            opx[t] = pack8(mfma(xv[t / 2][0], sl, z4), mfma(xv[t / 2][1], sl, z4));
        }
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[m][t] = mfma(opy[m], opx[t], acc[m][t]);
    };
    if constexpr (kWgPrefetch) {
        // the next chunk's loads are issued before the current chunk's MFMAs
        bf16x8 y0[MJ][2], x0[NB][2], y1[MJ][2], x1[NB][2];
        unsigned c = cbeg + wave;
        if (c < cend) load(c, y0, x0);
        for (; c < cend; c += 8) {
            if (c + 4 < cend) load(c + 4, y1, x1);
            compute(y0, x0);
            if (c + 4 >= cend) break;
            if (c + 8 < cend) load(c + 8, y0, x0);
            compute(y1, x1);
        }
    } else {
        for (unsigned c = cbeg + wave; c < cend; c += 4) {
            bf16x8 y0[MJ][2], x0[NB][2];
            load(c, y0, x0);
            compute(y0, x0);
        }
    }
    // bias partials: lane l holds channel l%16 of each tile; sum the four
    // lane groups
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        float v = bsum[m];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (g == 0) bred[wave][m][r] = v;
    }
    if (wave > 0) {
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) red[wave - 1][m][t][lane] = acc[m][t];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const f32x4 a = acc[m][t], b1 = red[0][m][t][lane], b2 = red[1][m][t][lane],
                            b3 = red[2][m][t][lane];
                const int k = kc0 + 16 * t + r;
                if (k >= K) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int co = co0 + 16 * m + 4 * g + e;
                    if (co < s.Cout) atomicAdd(&dw[(long long)co * K + k], (a[e] + b1[e]) + (b2[e] + b3[e]));
                }
            }
    }
    if (db && blockIdx.y == 0 && threadIdx.x < 16 * MB) {
        const int m = threadIdx.x / 16, ch = threadIdx.x % 16;
        const int co = co0 + threadIdx.x;
        if (co < s.Cout)
            atomicAdd(&db[co], (bred[0][m][ch] + bred[1][m][ch]) + (bred[2][m][ch] + bred[3][m][ch]));
    }
}

// ------------------------------------------------------------ elementwise --
__global__ void k_fill_f32(float* __restrict__ p, long long n, float v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// FP32 master [Cout][9][Cin] -> bf16 same layout and bf16 transpose [9][Cin][Cout]
__global__ void k_prep_weights(const float* __restrict__ w32, bf16* __restrict__ w, bf16* __restrict__ wt,
                               int Cout, int Cin)
{
    const long long n = (long long)Cout * 9 * Cin;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = (int)(i % Cin);
    const long long t = i / Cin;
    const int tap = (int)(t % 9);
    const int co = (int)(t / 9);
    const bf16 v = (bf16)w32[i];
    w[i] = v;
    if (wt) wt[((long long)tap * Cin + ci) * Cout + co] = v;
}

// y[n][2i+a][2j+b][c] = x[n][i][j][c] + (skip ? skip[...] : 0)
__global__ void k_upsample_add(const bf16* __restrict__ x, const bf16* __restrict__ skip, bf16* __restrict__ y,
                               int N, int H, int W, int C)
{
    const long long n = (long long)N * 4 * H * W * C;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int xo = (int)(t % (2 * W));
    t /= 2 * W;
    const int yo = (int)(t % (2 * H));
    const int nn = (int)(t / (2 * H));
    float v = (float)x[(((long long)nn * H + yo / 2) * W + xo / 2) * C + c];
    if (skip) v += (float)skip[i];
    y[i] = (bf16)v;
}

// dx[n][i][j][c] = sum_ab dy[n][2i+a][2j+b][c], masked by (mask > 0)
__global__ void k_upsample_bwd(const bf16* __restrict__ dy, bf16* __restrict__ dx, const bf16* __restrict__ mask,
                               int N, int H, int W, int C, int accumulate)
{
    const long long n = (long long)N * H * W * C;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int xx = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    const int nn = (int)(t / H);
    float v = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            v += (float)dy[(((long long)nn * 2 * H + 2 * yy + a) * 2 * W + 2 * xx + b) * C + c];
    if (mask && !((float)mask[i] > 0.f)) v = 0.f;
    if (accumulate) v += (float)dx[i];
    dx[i] = (bf16)v;
}

// The same three element-wise kernels over 8 channels per thread (16-byte
// loads and stores), for C % 8 == 0 and 16-byte aligned tensors: eight
// times fewer memory instructions for the same bytes.
__device__ __forceinline__ void store8(bf16* p, const bf16x8 v) { *(bf16x8*)p = v; }

__global__ void k_upsample_add8(const bf16* __restrict__ x, const bf16* __restrict__ skip, bf16* __restrict__ y,
                                int N, int H, int W, int C)
{
    const int C8 = C / 8;
    const long long n = (long long)N * 4 * H * W * C8;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C8) * 8;
    long long t = i / C8;
    const int xo = (int)(t % (2 * W));
    t /= 2 * W;
    const int yo = (int)(t % (2 * H));
    const int nn = (int)(t / (2 * H));
    const bf16x8 a = load8(x + (((long long)nn * H + yo / 2) * W + xo / 2) * C + c);
    bf16x8 v = a;
    if (skip) {
        const bf16x8 b = load8(skip + i * 8);
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (bf16)((float)a[e] + (float)b[e]);
    }
    store8(y + i * 8, v);
}

__global__ void k_upsample_bwd8(const bf16* __restrict__ dy, bf16* __restrict__ dx, const bf16* __restrict__ mask,
                                int N, int H, int W, int C, int accumulate)
{
    const int C8 = C / 8;
    const long long n = (long long)N * H * W * C8;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C8) * 8;
    long long t = i / C8;
    const int xx = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    const int nn = (int)(t / H);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const bf16x8 d = load8(dy + (((long long)nn * 2 * H + 2 * yy + a) * 2 * W + 2 * xx + b) * C + c);
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] += (float)d[e];
        }
    bf16x8 m = zero8(), o = zero8();
    if (mask) m = load8(mask + i * 8);
    if (accumulate) o = load8(dx + i * 8);
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float u = v[e];
        if (mask && !((float)m[e] > 0.f)) u = 0.f;
        if (accumulate) u += (float)o[e];
        r[e] = (bf16)u;
    }
    store8(dx + i * 8, r);
}

__global__ void k_masked_add8(const bf16* __restrict__ src, bf16* __restrict__ dst, const bf16* __restrict__ mask,
                              long long n8, int accumulate)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const bf16x8 a = load8(src + i * 8);
    bf16x8 m = zero8(), o = zero8();
    if (mask) m = load8(mask + i * 8);
    if (accumulate) o = load8(dst + i * 8);
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float v = (float)a[e];
        if (mask && !((float)m[e] > 0.f)) v = 0.f;
        if (accumulate) v += (float)o[e];
        r[e] = (bf16)v;
    }
    store8(dst + i * 8, r);
}

// dst = (accumulate ? dst : 0) + src * (mask > 0)
__global__ void k_masked_add(const bf16* __restrict__ src, bf16* __restrict__ dst, const bf16* __restrict__ mask,
                             long long n, int accumulate)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = (float)src[i];
    if (mask && !((float)mask[i] > 0.f)) v = 0.f;
    if (accumulate) v += (float)dst[i];
    dst[i] = (bf16)v;
}

// Masked MSE over the first `nt` channels of an FP32 prediction with C
// channels: loss partial per block; gradient (bf16, all C channels) =
// 2 (pred - target) * scale[c] / count on fluid cells, 0 elsewhere.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                                             const unsigned char* __restrict__ solid, bf16* __restrict__ grad,
                                             float* __restrict__ part, long long npix, int C, int nt,
                                             float inv_count)
{
    __shared__ float sh[4];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < npix) {
        const bool fluid = !(solid && solid[i]);
        for (int c = 0; c < C; ++c) {
            float g = 0.f;
            if (c < nt && fluid) {
                const float d = pred[i * C + c] - target[i * nt + c];
                l += d * d;
                g = 2.f * d * inv_count;
            }
            if (grad) grad[i * C + c] = (bf16)g;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Adam (FP32), bias-corrected.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Adam over every parameter tensor in one launch: the tensors lie in flat
// FP32 buffers (parameters, gradients, both moments) as segments; a thread
// finds its segment by binary search over the table, updates its element and,
// for a weight segment, writes the bf16 working copy [Cout][9][Cin] and its
// transpose [9][Cin][Cout] the convolutions read.
constexpr int kMaxSegments = 64;
struct SegTable {
    long long off[kMaxSegments + 1];   // first element of each segment, then the total
    int cout[kMaxSegments], cin[kMaxSegments];
    float wd[kMaxSegments];
    bf16* w[kMaxSegments];
    bf16* wt[kMaxSegments];
    int n;
};

__global__ void k_adam_multi(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, const SegTable tab, float lr, float b1, float b2,
                             float eps, float c1, float c2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tab.off[tab.n]) return;
    int lo = 0, hi = tab.n - 1;            // last segment whose start is <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const float wi = w[i];
    const float gi = g[i] + tab.wd[lo] * wi;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wn = wi - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wn;
    if (tab.w[lo]) {
        const long long e = i - tab.off[lo];
        const int Cin = tab.cin[lo], Cout = tab.cout[lo];
        const int ci = (int)(e % Cin);
        const long long t = e / Cin;
        const int tap = (int)(t % 9), co = (int)(t / 9);
        const bf16 b = (bf16)wn;
        tab.w[lo][e] = b;
        if (tab.wt[lo]) tab.wt[lo][((long long)tap * Cin + ci) * Cout + co] = b;
    }
}

// Network input: per pixel [solid, log-Re feature, x, y, 1, 0, 0, 0] (bf16)
__global__ void k_make_input(const unsigned char* __restrict__ solid, const float* __restrict__ re_feat,
                             bf16* __restrict__ out, int N, int H, int W)
{
    const long long n = (long long)N * H * W;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int xx = (int)(i % W);
    const long long t = i / W;
    const int yy = (int)(t % H);
    const int nn = (int)(t / H);
    bf16x8 v = zero8();
    v[0] = (bf16)(float)(solid ? solid[i] : 0);
    v[1] = (bf16)re_feat[nn];
    v[2] = (bf16)((xx + 0.5f) / W);
    v[3] = (bf16)((yy + 0.5f) / H);
    v[4] = (bf16)1.0f;
    *(bf16x8*)(out + i * 8) = v;
}

static dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace mfa_nn

using namespace mfa_nn;

extern "C" {

const char* mfa_nn_error(void) { return g_nn_err.c_str(); }

int mfa_nn_conv3x3(const void* x, const void* w, const float* bias, void* y, float* y_f32, int N, int H,
                   int W, int Cin, int Cout, int stride, int relu, void* stream)
{
    if (!x || !w || (!y && !y_f32)) return nn_fail(-1, "mfa_nn_conv3x3: null pointer");
    if (Cin % 8 || Cout % 8 || (stride != 1 && stride != 2) || N < 1 || H < 1 || W < 1)
        return nn_fail(-1, "mfa_nn_conv3x3: channels must be multiples of 8, stride 1 or 2");
    if (!aligned16({x, w, y, y_f32})) return nn_fail(-1, "mfa_nn_conv3x3: tensors must be 16-byte aligned");
    if (Cin > 128 || (long long)N * H * W * Cin > 0x7fffffffLL || (long long)Cout * 9 * Cin > 0x7fffffffLL)
        return nn_fail(-1, "mfa_nn_conv3x3: at most 128 input channels and 2^31 elements per tensor");
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * s.Ho * s.Wo;
    // block tiles (pixels x channels): 256 x 16, 128 x 32, 64 x 64; layers
    // whose 64 x 64 tiles number fewer than 512 (the deep ones on small
    // maps) take 32 x 32 tiles with the four waves splitting k
    hipStream_t st = (hipStream_t)stream;
    const bf16* xb = (const bf16*)x;
    const bf16* wb = (const bf16*)w;
    if (Cout <= 16) {
        dim3 grid((unsigned)((P + 255) / 256), (Cout + 15) / 16);
        hipLaunchKernelGGL((k_conv_fwd<1, 4, 1, 1>), grid, dim3(256), 0, st, xb, wb, bias, (bf16*)y, y_f32, s, relu);
    } else if (Cout <= 32) {
        dim3 grid((unsigned)((P + 127) / 128), (Cout + 31) / 32);
        hipLaunchKernelGGL((k_conv_fwd<1, 2, 2, 1>), grid, dim3(256), 0, st, xb, wb, bias, (bf16*)y, y_f32, s, relu);
    } else if (((P + 63) / 64) * ((Cout + 63) / 64) >= 512) {
        dim3 grid((unsigned)((P + 63) / 64), (Cout + 63) / 64);
        hipLaunchKernelGGL((k_conv_fwd<2, 2, 2, 1>), grid, dim3(256), 0, st, xb, wb, bias, (bf16*)y, y_f32, s, relu);
    } else {
        dim3 grid((unsigned)((P + 31) / 32), (Cout + 31) / 32);
        hipLaunchKernelGGL((k_conv_fwd<1, 2, 2, 4>), grid, dim3(256), 0, st, xb, wb, bias, (bf16*)y, y_f32, s, relu);
    }
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_conv3x3_bwd_data(const void* dy, const void* wt, void* dx, const void* mask, int N, int H,
                            int W, int Cin, int Cout, int stride, int accumulate, void* stream)
{
    if (!dy || !wt || !dx) return nn_fail(-1, "mfa_nn_conv3x3_bwd_data: null pointer");
    if (Cin % 8 || Cout % 8 || (stride != 1 && stride != 2))
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_data: channels must be multiples of 8, stride 1 or 2");
    if (!aligned16({dy, wt, dx, mask}))
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_data: tensors must be 16-byte aligned");
    if (Cout > 128 || (long long)N * H * W * Cin > 0x7fffffffLL || (long long)N * H * W * Cout > 0x7fffffffLL)
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_data: at most 128 output channels and 2^31 elements per tensor");
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * H * W;
    hipStream_t st = (hipStream_t)stream;
    const bf16* dyb = (const bf16*)dy;
    const bf16* wtb = (const bf16*)wt;
    const bf16* mb = (const bf16*)mask;
    auto go = [&](auto s1) {
        constexpr bool S1 = decltype(s1)::value;
        if (Cin <= 16) {
            dim3 grid((unsigned)((P + 255) / 256), (Cin + 15) / 16);
            hipLaunchKernelGGL((k_conv_bwd_data<1, 4, 1, 1, S1>), grid, dim3(256), 0, st, dyb, wtb, (bf16*)dx, mb, s, accumulate);
        } else if (Cin <= 32) {
            dim3 grid((unsigned)((P + 127) / 128), (Cin + 31) / 32);
            hipLaunchKernelGGL((k_conv_bwd_data<1, 2, 2, 1, S1>), grid, dim3(256), 0, st, dyb, wtb, (bf16*)dx, mb, s, accumulate);
        } else if (((P + 63) / 64) * ((Cin + 63) / 64) >= 512) {
            dim3 grid((unsigned)((P + 63) / 64), (Cin + 63) / 64);
            hipLaunchKernelGGL((k_conv_bwd_data<2, 2, 2, 1, S1>), grid, dim3(256), 0, st, dyb, wtb, (bf16*)dx, mb, s, accumulate);
        } else {
            dim3 grid((unsigned)((P + 31) / 32), (Cin + 31) / 32);
            hipLaunchKernelGGL((k_conv_bwd_data<1, 2, 2, 4, S1>), grid, dim3(256), 0, st, dyb, wtb, (bf16*)dx, mb, s, accumulate);
        }
    };
    if (stride == 1) go(std::true_type{});
    else go(std::false_type{});
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_conv3x3_bwd_weight(const void* x, const void* dy, float* dw, float* dbias, int N, int H,
                              int W, int Cin, int Cout, int stride, void* stream)
{
    if (!x || !dy || !dw) return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: null pointer");
    if (Cin % 8 || Cout % 8 || (stride != 1 && stride != 2))
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: channels must be multiples of 8, stride 1 or 2");
    if (!aligned16({x, dy})) return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: tensors must be 16-byte aligned");
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * s.Ho * s.Wo;
    if (P * Cout > 0x7fffffffLL || (long long)N * H * W * Cin > 0x7fffffffLL)
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: a tensor holds 2^31 elements or more");
    const int K = 9 * Cin;
    // tiles: 16 or 32 output channels (MB 1 or 2) x 32 NB columns of k.
    // Thin layers (K <= 160) take all their columns in one block; wider ones
    // 96-column blocks (K = 288, 576, 864 split evenly).
    const int MB = Cout <= 16 ? 1 : 2;
    const int NB = K <= 96 ? 3 : (K <= 160 ? 5 : 3);
    const int gx = (Cout + 16 * MB - 1) / (16 * MB), gy = (K + 32 * NB - 1) / (32 * NB);
    // about kWgBlocks blocks over the pixel chunks
    const long long nchunks = (P + 31) / 32;
    long long nz = kWgBlocks / (gx * gy);
    if (nz < 1) nz = 1;
    if (nz > nchunks) nz = nchunks;
    long long per = (nchunks + nz - 1) / nz;
    nz = (nchunks + per - 1) / per;
    dim3 grid(gx, gy, (unsigned)nz);
    hipStream_t st = (hipStream_t)stream;
    const bool row = s.Wo % 32 == 0;
    auto go = [&](auto mb, auto nb, auto rw) {
        hipLaunchKernelGGL((k_conv_bwd_weight<decltype(mb)::value, decltype(nb)::value, decltype(rw)::value>),
                           grid, dim3(256), 0, st, (const bf16*)x, (const bf16*)dy, dw, dbias, s, (int)per);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I5 = std::integral_constant<int, 5>;
    auto by_row = [&](auto mb, auto nb) {
        if (row) go(mb, nb, std::true_type{});
        else go(mb, nb, std::false_type{});
    };
    if (MB == 1 && NB == 3) by_row(I1{}, I3{});
    else if (MB == 1 && NB == 5) by_row(I1{}, I5{});
    else if (MB == 2 && NB == 3) by_row(I2{}, I3{});
    else by_row(I2{}, I5{});
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_fill(float* p, long long n, float v, void* stream)
{
    if (!p || n < 0) return nn_fail(-1, "mfa_nn_fill: bad argument");
    if (n) hipLaunchKernelGGL(k_fill_f32, g1(n), dim3(256), 0, (hipStream_t)stream, p, n, v);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_prep_weights(const float* w32, void* w, void* wt, int Cout, int Cin, void* stream)
{
    if (!w32 || !w) return nn_fail(-1, "mfa_nn_prep_weights: null pointer");
    const long long n = (long long)Cout * 9 * Cin;
    hipLaunchKernelGGL(k_prep_weights, g1(n), dim3(256), 0, (hipStream_t)stream, w32, (bf16*)w, (bf16*)wt, Cout, Cin);
    NN_HIP(hipGetLastError());
    return 0;
}

static bool vec8(int C, std::initializer_list<const void*> ptrs) { return C % 8 == 0 && aligned16(ptrs); }

int mfa_nn_upsample_add(const void* x, const void* skip, void* y, int N, int H, int W, int C, void* stream)
{
    if (!x || !y) return nn_fail(-1, "mfa_nn_upsample_add: null pointer");
    if (vec8(C, {x, skip, y}))
        hipLaunchKernelGGL(k_upsample_add8, g1((long long)N * 4 * H * W * (C / 8)), dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)skip, (bf16*)y, N, H, W, C);
    else
        hipLaunchKernelGGL(k_upsample_add, g1((long long)N * 4 * H * W * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)skip, (bf16*)y, N, H, W, C);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_upsample_bwd(const void* dy, void* dx, const void* mask, int N, int H, int W, int C,
                        int accumulate, void* stream)
{
    if (!dy || !dx) return nn_fail(-1, "mfa_nn_upsample_bwd: null pointer");
    if (vec8(C, {dy, dx, mask}))
        hipLaunchKernelGGL(k_upsample_bwd8, g1((long long)N * H * W * (C / 8)), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (bf16*)dx, (const bf16*)mask, N, H, W, C, accumulate);
    else
        hipLaunchKernelGGL(k_upsample_bwd, g1((long long)N * H * W * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (bf16*)dx, (const bf16*)mask, N, H, W, C, accumulate);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_masked_add(const void* src, void* dst, const void* mask, long long n, int accumulate,
                      void* stream)
{
    if (!src || !dst) return nn_fail(-1, "mfa_nn_masked_add: null pointer");
    if (n % 8 == 0 && vec8(8, {src, dst, mask}))
        hipLaunchKernelGGL(k_masked_add8, g1(n / 8), dim3(256), 0, (hipStream_t)stream, (const bf16*)src, (bf16*)dst, (const bf16*)mask, n / 8, accumulate);
    else
        hipLaunchKernelGGL(k_masked_add, g1(n), dim3(256), 0, (hipStream_t)stream, (const bf16*)src, (bf16*)dst, (const bf16*)mask, n, accumulate);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_mse(const float* pred, const float* target, const unsigned char* solid, void* grad,
               float* partial, long long npix, int C, int nt, float inv_count, void* stream)
{
    if (!pred || !target || !partial) return nn_fail(-1, "mfa_nn_mse: null pointer");
    hipLaunchKernelGGL(k_mse, g1(npix), dim3(256), 0, (hipStream_t)stream, pred, target, solid, (bf16*)grad, partial, npix, C, nt, inv_count);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float beta1,
                float beta2, float eps, int t, float weight_decay, void* stream)
{
    if (!w || !g || !m || !v || t < 1) return nn_fail(-1, "mfa_nn_adam: bad argument");
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_adam_multi(float* w, const float* g, float* m, float* v, const mfa_nn_segment* seg,
                      int nseg, float lr, float beta1, float beta2, float eps, int t, void* stream)
{
    if (!w || !g || !m || !v || !seg || t < 1) return nn_fail(-1, "mfa_nn_adam_multi: bad argument");
    if (nseg < 1 || nseg > kMaxSegments)
        return nn_fail(-1, "mfa_nn_adam_multi: 1 to " + std::to_string(kMaxSegments) + " segments");
    SegTable tab;
    tab.n = nseg;
    long long end = 0;
    for (int k = 0; k < nseg; ++k) {
        const mfa_nn_segment& sg = seg[k];
        if (sg.offset != end || sg.n < 1)
            return nn_fail(-1, "mfa_nn_adam_multi: segments must be non-empty and follow each other");
        if (sg.w_bf16 && (sg.cout < 1 || sg.cin < 1 || (long long)sg.cout * 9 * sg.cin != sg.n))
            return nn_fail(-1, "mfa_nn_adam_multi: a weight segment holds cout * 9 * cin elements");
        tab.off[k] = sg.offset;
        tab.cout[k] = sg.cout;
        tab.cin[k] = sg.cin;
        tab.wd[k] = sg.weight_decay;
        tab.w[k] = (bf16*)sg.w_bf16;
        tab.wt[k] = sg.w_bf16 ? (bf16*)sg.wt_bf16 : nullptr;
        end += sg.n;
    }
    tab.off[nseg] = end;
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam_multi, g1(end), dim3(256), 0, (hipStream_t)stream, w, g, m, v, tab, lr, beta1,
                       beta2, eps, c1, c2);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_make_input(const unsigned char* solid, const float* re_feat, void* out, int N, int H, int W,
                      void* stream)
{
    if (!re_feat || !out) return nn_fail(-1, "mfa_nn_make_input: null pointer");
    hipLaunchKernelGGL(k_make_input, g1((long long)N * H * W), dim3(256), 0, (hipStream_t)stream, solid, re_feat, (bf16*)out, N, H, W);
    NN_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
