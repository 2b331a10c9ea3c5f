// Kernels of the learned occlusion pass: a U-Net over the G-buffer whose 3x3
// convolutions run as implicit GEMMs on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_f16 forward, v_mfma_f32_16x16x16_f16 backward: FP16
// operands, FP32 accumulation).
//
// Tensors are channels-last: element (n, y, x, c) at ((n*H + y)*W + x)*C + c.
// Forward convolution weights are [Cout][Kp] FP16 with k = tap*Cin + ci (tap =
// ky*3 + kx reads pixel (y + ky - 1, x + kx - 1), zero padding), Kp = 9*Cin
// rounded up to a multiple of 32 with zero columns.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace isaac_hip
{
    // Input of a convolution: channels [0, ca) of pixel (y, x) come from `up` at
    // (y/2, x/2) (nearest-neighbour upsampling of an (h/2, w/2, ca) tensor) and
    // channels [ca, cin) from `x` (h, w, cin - ca); with up == nullptr all cin
    // channels come from x.  This fuses the decoder's upsample + concat.
    struct ConvArgs
    {
        const _Float16* x; // [n][h][w][cin - ca]
        const _Float16* up; // [n][h/2][w/2][ca] or nullptr
        int32_t ca; // channels from `up` (a multiple of 4; 0 without `up`)
        const _Float16* w; // [cout][kp]
        const float* bias; // [cout] (forward only)
        void* y; // forward: _Float16 [n][h][w][cout] = relu(conv + bias) (nullable); backward data: float, +=
        int32_t n, h, w_, cin, kp; // cin a multiple of 4
        // forward only (launch_conv_fwd), all optional:
        _Float16* pool; // [n][h/2][w/2][cout]: 2x2 average of the FP16 activation (h, w even)
        float* head; // cout 16: [n][imgH][imgW] = clamp(act . headW[0..15] + headW[16], 0, 1)
        const float* headW;
        int32_t imgW, imgH;
    };

    // Backward data of a layer: y (FP32, +=) = 3x3 convolution of x with the layer's
    // flipped, transposed weights ([cin][kp], kp = 9*cout); `cout` here is the
    // layer's cin, in {16, 32, 48, 64, 96, 128}.  max_lds: the device's LDS bytes per
    // workgroup (picks the rows per workgroup).
    hipError_t launch_conv3x3_bwd(const ConvArgs& a, int cout, size_t max_lds, hipStream_t s);

    // Forward layer y = relu(conv(x) + bias) on v_mfma_f32_16x16x32_f16 for the
    // network's (cin, cout, ca) shapes: (8,16,0) (16,16,0) (16,32,0) (32,32,0)
    // (32,64,0) (64,64,0) (128,64,64) (96,32,64) (48,16,32); weights [cout][kp] with
    // kp = 9*cin rounded up to a multiple of 32 (zero columns).  The fused pool and
    // head outputs equal launch_pool2 / launch_head_infer applied to the activation.
    hipError_t launch_conv_fwd(const ConvArgs& a, int cout, size_t max_lds, hipStream_t s);

    // dW[co][k] += sum_p dz[p][co] * col(x)[p][k] for k < 9*cin (FP32 atomics into dw,
    // rows of 9*cin floats: the master weight layout [cout][9][cin])
    hipError_t launch_conv3x3_dw(
        const _Float16* x,
        const _Float16* up,
        int ca,
        const _Float16* dz,
        float* dw,
        int n,
        int h,
        int w,
        int cin,
        int cout,
        hipStream_t s);
    // db[c] += sum_p dz[p][c]
    hipError_t launch_bias_grad(const _Float16* dz, float* db, int64_t npx, int c, hipStream_t s);
    // dz = half(dy * (y > 0))
    hipError_t launch_relu_bwd(const float* dy, const _Float16* y, _Float16* dz, int64_t count, hipStream_t s);
    // backward of the 2x2 average pool (h, w even): dx += dy / 4
    hipError_t launch_pool2_bwd(const float* dy, float* dx, int n, int h, int w, int c, hipStream_t s);
    // backward of y = concat(nearest-upsample(a), skip): da += 2x2 sums of dy[.., :ca],
    // dskip += dy[.., ca:]
    hipError_t launch_upcat_bwd(const float* dy, float* da, float* dskip, int n, int h, int w, int ca, int cs, hipStream_t s);

    // network input from a G-buffer window: 8 channels (depth / scale or 0, background
    // flag, normal xyz, 0, 0, 0); pixels outside the image are background.  With `ao`
    // and `target` set, also copies the occlusion term as the training target.
    struct FeatureArgs
    {
        const float* depth; // [imgH][imgW]
        const float* normal; // [imgH][imgW][3]
        const float* ao; // [imgH][imgW] or nullptr
        int32_t imgW, imgH;
        float invScale;
        const int32_t* origin; // per batch item: x0, y0 of the window (device), or nullptr = (0, 0)
        _Float16* x; // [n][h][w][8]
        float* target; // [n][h][w] or nullptr
        int32_t n, h, w;
        int64_t imgStride; // elements between the images of consecutive items (0: one image)
        const float* invScaleItem; // per item 1 / scale (device), or nullptr = invScale
    };
    hipError_t launch_features(const FeatureArgs& a, hipStream_t s);

    // Training: out = d0 . wh + bh (wh[16] = bh); loss += sum (out - t)^2;
    // gwh[0..15] += sum dout * d0, gwh[16] += sum dout; dd0 = dout * wh with
    // dout = 2 (out - t); `out` (nullable) receives the raw outputs
    hipError_t launch_head_train(
        const _Float16* d0,
        const float* wh,
        const float* target,
        float* loss,
        float* gwh,
        float* dd0,
        float* out,
        int64_t npx,
        hipStream_t s);

    // Adam over the flat FP32 parameter vector (gradients scaled by gscale first)
    hipError_t launch_adam(
        float* w,
        const float* g,
        float* m,
        float* v,
        int64_t count,
        float lr,
        float b1,
        float b2,
        float eps,
        float c1,
        float c2,
        float gscale,
        hipStream_t s);
    // FP16 forward weights [cout][kp] and flipped-transposed backward weights
    // [cin][9*cout] (w'[ci][t][co] = w[co][8-t][ci]) from FP32 master [cout][9][cin]
    hipError_t launch_pack_weights(
        const float* master,
        _Float16* wf,
        _Float16* wb,
        int cin,
        int cout,
        int kp,
        hipStream_t s);
} // namespace isaac_hip
