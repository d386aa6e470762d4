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
//   weight grad    dW[co][k] = sum_p dY[p][co] X(p,k)   (split over pixels)
// MFMA operand layout (wave64): lane l holds A[row l%16][k 8*(l/16)..+7] and
// B[k 8*(l/16)..+7][col l%16]; the 16x16 FP32 result lands as
// C[4*(l/16)+i][l%16], i = 0..3.  A wave computes a 32x32 tile with 2x2
// MFMAs per 32-deep slice of k.  Channel counts are multiples of 8 so each
// lane's 8 consecutive k of the forward / data-gradient operands are one
// 16-byte load; the weight gradient stages both operands through LDS
// transposed so its pixel-major reduction reads 16 bytes per lane as well.
#include <hip/hip_runtime.h>

#include <string>

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

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct ConvShape {
    int N, H, W, Cin, Cout, stride, Ho, Wo;
};

// ---------------------------------------------------------------- forward --
// Block: 4 waves.  WC waves along output channels (1 or 2), 4/WC along pixels;
// a wave computes 16*TI pixels x 16*TJ channels (TI x TJ MFMA tiles per
// 32-deep slice of k).  TI = 4, TJ = 1 serves layers with 16 or fewer output
// channels without idle MFMA columns.
template <int WC, int TI, int TJ>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                  const float* __restrict__ bias, bf16* __restrict__ y,
                                                  float* __restrict__ yf, const ConvShape s, const int relu)
{
    constexpr int WP = 4 / WC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wp = wave % WP, wc = wave / WP;
    const long long P = (long long)s.N * s.Ho * s.Wo;
    const long long p0 = (long long)blockIdx.x * (16 * TI * WP) + 16 * TI * wp;
    const int c0 = blockIdx.y * (16 * TJ * WC) + 16 * TJ * wc;
    const int K = 9 * s.Cin;
    const int r = lane & 15, kg = lane >> 4;

    // pixel coordinates of the two A rows this lane feeds
    int pn[TI], py[TI], px[TI];
    bool pv[TI];
#pragma unroll
    for (int i = 0; i < TI; ++i) {
        const long long p = p0 + 16 * i + r;
        pv[i] = p < P;
        const long long pp = pv[i] ? p : 0;
        px[i] = (int)(pp % s.Wo);
        const long long t = pp / s.Wo;
        py[i] = (int)(t % s.Ho);
        pn[i] = (int)(t / s.Ho);
    }
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += 32) {
        const int k = k0 + 8 * kg;
        bf16x8 a[TI], b[TJ];
        const bool kv = k < K;
        const int tap = kv ? k / s.Cin : 0, ci = kv ? k % s.Cin : 0;
        const int ky = tap / 3, kx = tap % 3;
#pragma unroll
        for (int i = 0; i < TI; ++i) {
            const int yy = py[i] * s.stride + ky - 1, xx = px[i] * s.stride + kx - 1;
            const bool ok = kv && pv[i] && yy >= 0 && yy < s.H && xx >= 0 && xx < s.W;
            a[i] = ok ? load8(x + (((long long)pn[i] * s.H + yy) * s.W + xx) * s.Cin + ci) : zero8();
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            const int co = c0 + 16 * j + r;
            b[j] = (kv && co < s.Cout) ? load8(w + (long long)co * K + k) : zero8();
        }
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            const int co = c0 + 16 * j + r;
            if (co >= s.Cout) continue;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const long long p = p0 + 16 * i + 4 * kg + e;
                if (p >= P) continue;
                float v = acc[i][j][e] + bv;
                if (relu) v = v > 0.f ? v : 0.f;
                if (y) y[p * s.Cout + co] = (bf16)v;
                if (yf) yf[p * s.Cout + co] = v;
            }
        }
}

// ---------------------------------------------------------- data gradient --
// dX[p_in][ci] = sum over (tap, co) dY[p_out][co] * WT[tap][ci][co], with
// p_out = ((y_in + 1 - ky) / stride, (x_in + 1 - kx) / stride) when exact.
// Epilogue: dX = (accumulate ? dX : 0) + value * (mask ? (mask > 0) : 1).
// Wave tile as in the forward kernel: 16*TI pixels x 16*TJ input channels.
template <int WC, int TI, int TJ>
__global__ __launch_bounds__(256) void k_conv_bwd_data(const bf16* __restrict__ dy, const bf16* __restrict__ wt,
                                                       bf16* __restrict__ dx, const bf16* __restrict__ mask,
                                                       const ConvShape s, const int accumulate)
{
    constexpr int WP = 4 / WC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wp = wave % WP, wc = wave / WP;
    const long long P = (long long)s.N * s.H * s.W;
    const long long p0 = (long long)blockIdx.x * (16 * TI * WP) + 16 * TI * wp;
    const int c0 = blockIdx.y * (16 * TJ * WC) + 16 * TJ * wc;
    const int K = 9 * s.Cout;
    const int r = lane & 15, kg = lane >> 4;
    int pn[TI], py[TI], px[TI];
    bool pv[TI];
#pragma unroll
    for (int i = 0; i < TI; ++i) {
        const long long p = p0 + 16 * i + r;
        pv[i] = p < P;
        const long long pp = pv[i] ? p : 0;
        px[i] = (int)(pp % s.W);
        const long long t = pp / s.W;
        py[i] = (int)(t % s.H);
        pn[i] = (int)(t / s.H);
    }
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int k = k0 + 8 * kg;
        const bool kv = k < K;
        const int tap = kv ? k / s.Cout : 0, co = kv ? k % s.Cout : 0;
        const int ky = tap / 3, kx = tap % 3;
        bf16x8 a[TI], b[TJ];
#pragma unroll
        for (int i = 0; i < TI; ++i) {
            const int ty = py[i] + 1 - ky, tx = px[i] + 1 - kx;
            bool ok = kv && pv[i] && ty >= 0 && tx >= 0;
            int yo = 0, xo = 0;
            if (ok) {
                if (s.stride == 1) { yo = ty; xo = tx; }
                else { ok = ((ty | tx) & 1) == 0; yo = ty >> 1; xo = tx >> 1; }
                ok = ok && yo < s.Ho && xo < s.Wo;
            }
            a[i] = ok ? load8(dy + (((long long)pn[i] * s.Ho + yo) * s.Wo + xo) * s.Cout + co) : zero8();
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            const int ci = c0 + 16 * j + r;
            b[j] = (kv && ci < s.Cin) ? load8(wt + ((long long)tap * s.Cin + ci) * s.Cout + co) : zero8();
        }
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            const int ci = c0 + 16 * j + r;
            if (ci >= s.Cin) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const long long p = p0 + 16 * i + 4 * kg + e;
                if (p >= P) continue;
                const long long idx = p * s.Cin + ci;
                float v = acc[i][j][e];
                if (mask && !((float)mask[idx] > 0.f)) v = 0.f;
                if (accumulate) v += (float)dx[idx];
                dx[idx] = (bf16)v;
            }
        }
}

// -------------------------------------------------------- weight gradient --
// Block tile: 64 output channels x 64 columns of k; each block reduces one
// contiguous range of output pixels (split over blockIdx.z) and adds its
// FP32 partial into dW (and dbias) atomically.  Per stage of kWgStage
// pixels the dY and im2col(X) tiles are staged in LDS transposed
// ([row][pixel]) so each lane's 8 pixels are one 16-byte LDS read; with 128
// pixels a stage each wave issues 16 MFMAs between two barriers.
constexpr int kWgStage = 128;

__global__ __launch_bounds__(256) void k_conv_bwd_weight(const bf16* __restrict__ x,
                                                         const bf16* __restrict__ dy,
                                                         float* __restrict__ dw, float* __restrict__ db,
                                                         const ConvShape s, const int pix_per_block)
{
    constexpr int PS = kWgStage;
    __shared__ __attribute__((aligned(16))) bf16 sA[64][PS + 8];  // [co][p]  (+ pad)
    __shared__ __attribute__((aligned(16))) bf16 sB[64][PS + 8];  // [k][p]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave & 1, wcl = wave >> 1;
    const int co0 = blockIdx.x * 64, kk0 = blockIdx.y * 64;
    const int K = 9 * s.Cin;
    const long long P = (long long)s.N * s.Ho * s.Wo;
    const long long pbeg = (long long)blockIdx.z * pix_per_block;
    const long long pend = (pbeg + pix_per_block < P) ? pbeg + pix_per_block : P;
    const int r = lane & 15, kg = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (long long ps = pbeg; ps < pend; ps += PS) {
        // staging: PS pixels x 8 groups of 8 rows per operand, PS*8/256 items per thread
#pragma unroll
        for (int m = 0; m < PS * 8 / 256; ++m) {
            const int item = threadIdx.x + 256 * m;
            const int sp = item >> 3;           // pixel within the stage
            const int sg = item & 7;            // group of 8 rows
            const long long p = ps + sp;
            const bool pv = p < pend;
            int n = 0, yo = 0, xo = 0;
            if (pv) {
                xo = (int)(p % s.Wo);
                const long long t = p / s.Wo;
                yo = (int)(t % s.Ho);
                n = (int)(t / s.Ho);
            }
            // dY slice -> sA[co][p]
            {
                const int co = co0 + 8 * sg;
                const bf16x8 v = (pv && co < s.Cout) ? load8(dy + p * s.Cout + co) : zero8();
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[8 * sg + e][sp] = v[e];
            }
            // im2col(X) slice -> sB[k][p]
            {
                const int k = kk0 + 8 * sg;
                bf16x8 v = zero8();
                if (pv && k < K) {
                    const int tap = k / s.Cin, ci = k % s.Cin;
                    const int yy = yo * s.stride + tap / 3 - 1, xx = xo * s.stride + tap % 3 - 1;
                    if (yy >= 0 && yy < s.H && xx >= 0 && xx < s.W)
                        v = load8(x + (((long long)n * s.H + yy) * s.W + xx) * s.Cin + ci);
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[8 * sg + e][sp] = v[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < PS / 32; ++kk) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8*)&sA[32 * wr + 16 * i + r][32 * kk + 8 * kg];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8*)&sB[32 * wcl + 16 * j + r][32 * kk + 8 * kg];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
        if (db && blockIdx.y == 0 && threadIdx.x < 64) {
#pragma unroll 8
            for (int q = 0; q < PS; ++q) bsum += (float)sA[threadIdx.x][q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int k = kk0 + 32 * wcl + 16 * j + r;
            if (k >= K) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int co = co0 + 32 * wr + 16 * i + 4 * kg + e;
                if (co < s.Cout) atomicAdd(&dw[(long long)co * K + k], acc[i][j][e]);
            }
        }
    if (db && blockIdx.y == 0 && threadIdx.x < 64 && co0 + (int)threadIdx.x < s.Cout)
        atomicAdd(&db[co0 + threadIdx.x], bsum);
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
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * s.Ho * s.Wo;
    // block tiles (pixels x channels): 256 x 16, 128 x 32, 64 x 64
    if (Cout <= 16) {
        dim3 grid((unsigned)((P + 255) / 256), (Cout + 15) / 16);
        hipLaunchKernelGGL((k_conv_fwd<1, 4, 1>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)w, bias, (bf16*)y, y_f32, s, relu);
    } else if (Cout <= 32) {
        dim3 grid((unsigned)((P + 127) / 128), (Cout + 31) / 32);
        hipLaunchKernelGGL((k_conv_fwd<1, 2, 2>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)w, bias, (bf16*)y, y_f32, s, relu);
    } else {
        dim3 grid((unsigned)((P + 63) / 64), (Cout + 63) / 64);
        hipLaunchKernelGGL((k_conv_fwd<2, 2, 2>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)w, bias, (bf16*)y, y_f32, s, relu);
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
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * H * W;
    if (Cin <= 16) {
        dim3 grid((unsigned)((P + 255) / 256), (Cin + 15) / 16);
        hipLaunchKernelGGL((k_conv_bwd_data<1, 4, 1>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (const bf16*)wt, (bf16*)dx, (const bf16*)mask, s, accumulate);
    } else if (Cin <= 32) {
        dim3 grid((unsigned)((P + 127) / 128), (Cin + 31) / 32);
        hipLaunchKernelGGL((k_conv_bwd_data<1, 2, 2>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (const bf16*)wt, (bf16*)dx, (const bf16*)mask, s, accumulate);
    } else {
        dim3 grid((unsigned)((P + 63) / 64), (Cin + 63) / 64);
        hipLaunchKernelGGL((k_conv_bwd_data<2, 2, 2>), grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (const bf16*)wt, (bf16*)dx, (const bf16*)mask, s, accumulate);
    }
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_conv3x3_bwd_weight(const void* x, const void* dy, float* dw, float* dbias, int N, int H,
                              int W, int Cin, int Cout, int stride, void* stream)
{
    if (!x || !dy || !dw) return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: null pointer");
    if (Cin % 8 || Cout % 8 || (stride != 1 && stride != 2))
        return nn_fail(-1, "mfa_nn_conv3x3_bwd_weight: channels must be multiples of 8, stride 1 or 2");
    ConvShape s{N, H, W, Cin, Cout, stride, (H + stride - 1) / stride, (W + stride - 1) / stride};
    const long long P = (long long)N * s.Ho * s.Wo;
    const int tiles = ((Cout + 63) / 64) * ((9 * Cin + 63) / 64);
    // enough pixel slices to fill the chip (~2048 blocks), whole LDS stages
    long long per = (P * tiles + 2047) / 2048;
    per = ((per + kWgStage - 1) / kWgStage) * kWgStage;
    if (per < 2 * kWgStage) per = 2 * kWgStage;
    const int nz = (int)((P + per - 1) / per);
    dim3 grid((Cout + 63) / 64, (9 * Cin + 63) / 64, nz);
    hipLaunchKernelGGL(k_conv_bwd_weight, grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)dy, dw, dbias, s, (int)per);
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

int mfa_nn_upsample_add(const void* x, const void* skip, void* y, int N, int H, int W, int C, void* stream)
{
    if (!x || !y) return nn_fail(-1, "mfa_nn_upsample_add: null pointer");
    hipLaunchKernelGGL(k_upsample_add, g1((long long)N * 4 * H * W * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (const bf16*)skip, (bf16*)y, N, H, W, C);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_upsample_bwd(const void* dy, void* dx, const void* mask, int N, int H, int W, int C,
                        int accumulate, void* stream)
{
    if (!dy || !dx) return nn_fail(-1, "mfa_nn_upsample_bwd: null pointer");
    hipLaunchKernelGGL(k_upsample_bwd, g1((long long)N * H * W * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, (bf16*)dx, (const bf16*)mask, N, H, W, C, accumulate);
    NN_HIP(hipGetLastError());
    return 0;
}

int mfa_nn_masked_add(const void* src, void* dst, const void* mask, long long n, int accumulate,
                      void* stream)
{
    if (!src || !dst) return nn_fail(-1, "mfa_nn_masked_add: null pointer");
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

int mfa_nn_make_input(const unsigned char* solid, const float* re_feat, void* out, int N, int H, int W,
                      void* stream)
{
    if (!re_feat || !out) return nn_fail(-1, "mfa_nn_make_input: null pointer");
    hipLaunchKernelGGL(k_make_input, g1((long long)N * H * W), dim3(256), 0, (hipStream_t)stream, solid, re_feat, (bf16*)out, N, H, W);
    NN_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
