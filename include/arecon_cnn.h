/* Kernels of the learned artifact-correction network (3-D residual
 * encoder-decoder with 3x3x3 convolutions, BF16 on the gfx950 matrix cores,
 * FP32 accumulation).  No counterpart in the reference: it is an extension
 * beside the drop-in path of arecon.h.  All pointers are device pointers;
 * every call is stream-ordered on `stream` and returns ARECON_OK or an error
 * code (message via arecon_last_error).
 *
 * Layouts: activations channels-last BF16 [B][D][H][W][C], C a multiple of 8;
 * dims = {B, D, H, W}.  Forward weights BF16 [round16(Cout)][Kp], k = tap*Cin
 * + ci, tap = 9(dz+1) + 3(dy+1) + (dx+1), Kp = round32(27*Cin); data-gradient
 * weights [round16(Cin)][round32(27*Cout)] (taps flipped, transposed).  Master
 * weights and weight gradients FP32 [Cout][27][Cin]. */
#ifndef ARECON_CNN_H_
#define ARECON_CNN_H_

#ifdef __cplusplus
extern "C" {
#endif

/* y = act(conv3x3x3(x, w) + bias + skip); act = ReLU when relu != 0.  bias and
 * skip ([B][D][H][W][Cout]) may be NULL.  Cout a multiple of 4, at most 128.
 * With the data-gradient weights this computes dX from dY. */
int arecon_cnn_conv3(void* y, const void* x, const void* w, const float* bias, const void* skip,
                     const int* dims, int cin, int cout, int relu, void* stream);

/* dw[co][tap][ci] += sum over voxels of dyT[co][n] * x[nb(n, tap)][ci], from
 * the channel-major output gradient dyT [Cout][nvox] and the three
 * channel-major input copies xT3 [3][Cin][nvox] (arecon_cnn_channel_major3).
 * W must be a multiple of 8. */
int arecon_cnn_conv3_wgrad(float* dw, const void* dyT, const void* xT3, const int* dims,
                           int cin, int cout, void* stream);

/* dyp = dy * (y > 0) (relu != 0) or dy, and its channel-major copy dypT; db
 * (may be NULL) += per-channel sums (the bias gradient). */
int arecon_cnn_relu_bwd(void* dyp, void* dypT, float* db, const void* dy, const void* y,
                        const int* dims, int c, int relu, void* stream);

/* xT3[s][c][n] = x[n + s - 1][c] along W (zero past the row ends), s = 0..2. */
int arecon_cnn_channel_major3(void* xT3, const void* x, const int* dims, int c, void* stream);

/* 2x2x2 average pooling (dims of the output) and its adjoint (dims of the
 * input; accumulate != 0 adds into dx); nearest 2x upsampling (dims of the
 * output) and its adjoint (dims of the input). */
int arecon_cnn_pool(void* y, const void* x, const int* dims_out, int c, void* stream);
int arecon_cnn_pool_bwd(void* dx, const void* dy, const int* dims_in, int c, int accumulate,
                        void* stream);
int arecon_cnn_upsample(void* y, const void* x, const int* dims_out, int c, void* stream);
int arecon_cnn_upsample_bwd(void* dx, const void* dy, const int* dims_in, int c, void* stream);

/* dst += src, n BF16 elements. */
int arecon_cnn_add(void* dst, const void* src, long long n, void* stream);

/* FP32 master weights [Cout][27][Cin] -> BF16 forward weights wf and (if wd is
 * not NULL) data-gradient weights wd. */
int arecon_cnn_pack_weights(void* wf, void* wd, const float* w, int cout, int cin, void* stream);

/* Residual prediction x0 + out[ch 0] (x0 FP32 [nvox]) against target (FP32
 * [nvox]): dout = d(mean squared error)/d(out) (BF16, channels other than 0
 * zeroed), *loss += the mean squared error. */
int arecon_cnn_mse_residual(void* dout, float* loss, const void* out, int cout, const float* x0,
                            const float* target, long long nv, void* stream);

/* One Adam step (bias-corrected, step counts from 1) on FP32 weights. */
int arecon_cnn_adam(float* w, float* m, float* v, const float* g, long long n, float lr, float b1,
                    float b2, float eps, int step, void* stream);

/* FP32 [nvox] -> BF16 channels-last [nvox][c] (channel 0, others zero). */
int arecon_cnn_to_bf16(void* y, const float* x, long long nv, int c, void* stream);

/* y = x0 + out[ch 0] (FP32 [nvox]). */
int arecon_cnn_residual_out(float* y, const void* out, int cout, const float* x0, long long nv,
                            void* stream);

#ifdef __cplusplus
}
#endif

#endif  /* ARECON_CNN_H_ */
