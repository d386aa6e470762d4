/* mfa_surrogate.h -- C ABI of the learned warm start's network kernels:
 * BF16 MFMA (v_mfma_f32_16x16x32_bf16) 3x3 convolutions with FP32
 * accumulation, their data and weight gradients, and the element-wise,
 * loss and optimiser kernels of training.  No counterpart in the reference
 * (it has no learned component).
 *
 * Tensors are NHWC; activations and gradients bfloat16 unless named f32;
 * channel counts are multiples of 8; weights [Cout][3][3][Cin] (bf16 copy of
 * an FP32 master) and their transpose [3][3][Cin][Cout] for the data
 * gradient.  Padding 1; stride 1 or 2 (output (H+1)/2 x (W+1)/2).  The
 * convolutions' tensors must be 16-byte aligned (refused otherwise).
 * Functions return 0 or a negative code (message: mfa_nn_error()).
 */
#ifndef MFA_SURROGATE_H
#define MFA_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

const char* mfa_nn_error(void);

/* y = act(conv3x3(x, w) + bias): writes y (bf16) and/or y_f32 */
int mfa_nn_conv3x3(const void* x, const void* w, const float* bias, void* y, float* y_f32, int N,
                   int H, int W, int Cin, int Cout, int stride, int relu, void* stream);
/* dx = (accumulate ? dx : 0) + conv3x3^T(dy, wt) * (mask > 0 if mask) */
int mfa_nn_conv3x3_bwd_data(const void* dy, const void* wt, void* dx, const void* mask, int N,
                            int H, int W, int Cin, int Cout, int stride, int accumulate,
                            void* stream);
/* dw[Cout][9*Cin] += sum_p dy x_col, dbias[Cout] += sum_p dy (FP32, atomic) */
int mfa_nn_conv3x3_bwd_weight(const void* x, const void* dy, float* dw, float* dbias, int N, int H,
                              int W, int Cin, int Cout, int stride, void* stream);
int mfa_nn_fill(float* p, long long n, float v, void* stream);
int mfa_nn_prep_weights(const float* w32, void* w, void* wt, int Cout, int Cin, void* stream);
/* y = nearest-upsample(x) (2H x 2W) + skip (skip may be NULL) */
int mfa_nn_upsample_add(const void* x, const void* skip, void* y, int N, int H, int W, int C,
                        void* stream);
int mfa_nn_upsample_bwd(const void* dy, void* dx, const void* mask, int N, int H, int W, int C,
                        int accumulate, void* stream);
int mfa_nn_masked_add(const void* src, void* dst, const void* mask, long long n, int accumulate,
                      void* stream);
/* masked MSE of the first nt of C FP32 prediction channels; partial[block]
 * sums of squared errors; grad (bf16, C channels) = 2 d inv_count */
int mfa_nn_mse(const float* pred, const float* target, const unsigned char* solid, void* grad,
               float* partial, long long npix, int C, int nt, float inv_count, void* stream);
int mfa_nn_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float beta1,
                float beta2, float eps, int t, float weight_decay, void* stream);
/* One tensor of a flat parameter buffer: elements [offset, offset + n).  A
 * weight segment (w_bf16 not NULL) holds cout * 9 * cin elements [cout][9][cin]
 * and gets its bf16 copy and, if wt_bf16 is not NULL, the transpose
 * [9][cin][cout] written after the update; a bias segment has w_bf16 NULL. */
typedef struct {
    long long offset;
    long long n;
    int cout, cin;
    float weight_decay;
    void* w_bf16;
    void* wt_bf16;
} mfa_nn_segment;
/* Adam (FP32, bias-corrected, step t >= 1) over every segment of the flat
 * buffers w, g, m, v in one launch; segments follow each other from offset 0
 * (at most 64) */
int mfa_nn_adam_multi(float* w, const float* g, float* m, float* v, const mfa_nn_segment* seg,
                      int nseg, float lr, float beta1, float beta2, float eps, int t, void* stream);
/* per pixel [solid, re_feat[n], (x+.5)/W, (y+.5)/H, 1, 0, 0, 0] */
int mfa_nn_make_input(const unsigned char* solid, const float* re_feat, void* out, int N, int H,
                      int W, void* stream);

#ifdef __cplusplus
}
#endif
#endif
