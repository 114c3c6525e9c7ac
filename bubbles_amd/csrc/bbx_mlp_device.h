// bbx_mlp_device.h -- device side of the warm-start MLP shared by the MLP
// kernels and the fused feature + inference gather: parameter layout, BF16
// packing, GELU, the MFMA tile and the inference forward pass.
//
// Layout.  A wave processes tiles of 16 samples.  Every layer is computed
// transposed, H^T = W . X^T, with v_mfma_f32_16x16x16_bf16: for an MFMA
// tile the A operand (weights) has lane l holding row l%16 and columns
// 4*(l/16)+0..3, the B operand (activations) has lane l holding column
// (sample) l%16 and rows 4*(l/16)+0..3, and the FP32 result has lane l
// holding column l%16 and rows 4*(l/16)+0..3, so one layer's result layout
// is the next layer's B layout.
#pragma once

#include "bbx_device.h"

namespace bbx {

constexpr int kH = BBX_MLP_HIDDEN;  // 64
constexpr int kF = BBX_NFEAT;       // 16
constexpr int OFF_W1 = 0;                 // [64][16]
constexpr int OFF_B1 = OFF_W1 + kH * kF;  // [64]
constexpr int OFF_W2 = OFF_B1 + kH;       // [64][64]
constexpr int OFF_B2 = OFF_W2 + kH * kH;  // [64]
constexpr int OFF_W3 = OFF_B2 + kH;       // [64]
constexpr int OFF_B3 = OFF_W3 + kH;       // [1]
constexpr int kTrainable = OFF_B3 + 1;    // 5313
constexpr int OFF_MEAN = kTrainable;      // feature mean [16]
constexpr int OFF_ISTD = OFF_MEAN + kF;   // feature 1/std [16]
constexpr int OFF_OSCALE = OFF_ISTD + kF; // target scale [1]
constexpr int OFF_ALPHA = OFF_OSCALE + 1; // warm start = alpha * clamp(prediction, 0, pcap)
constexpr int OFF_PCAP = OFF_ALPHA + 1;
constexpr int OFF_RESID = OFF_PCAP + 1;   // prediction = net * out_scale + resid * x[PPREV]
constexpr int kParams = OFF_RESID + 1;    // 5349

// Normalised features are clipped to +-kFeatClip standard deviations: rare
// states far outside the training data (an impact, a splash) then enter the
// network at the edge of what it was trained on instead of as extreme inputs
// it would extrapolate from.
constexpr float kFeatClip = 8.f;
__device__ __forceinline__ float norm_feature(float x, float mean, float istd) {
    return fminf(fmaxf((x - mean) * istd, -kFeatClip), kFeatClip);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 round-to-nearest-even conversions (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short f2bf(float f) {
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    const bf4 v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    return __builtin_bit_cast(bf16x4, v);
}

// GELU, tanh form: 0.5 z (1 + tanh(u)) = z s with s = 1 / (1 + exp(-2u)) and
// u = sqrt(2/pi) (z + 0.044715 z^3); exp and the reciprocal are the hardware
// approximations (about 1 ulp), far below the BF16 rounding that follows.
// The exponent is folded into one polynomial for the hardware exp2:
// -2 u log2(e) = z (a + b z^2), a = -2 sqrt(2/pi) log2(e), b = 0.044715 a.
__device__ __forceinline__ float gelu_sig(float z) {
    const float a = -2.f * 0.7978845608028654f * 1.4426950408889634f;
    const float b = a * 0.044715f;
    const float t = z * fmaf(b, z * z, a);
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}

// value and derivative: gelu'(z) = s + 2 z s (1 - s) u'(z)
__device__ __forceinline__ void gelu_fwd(float z, float &g, float &dg) {
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float sg = gelu_sig(z);
    g = z * sg;
    dg = sg + 2.f * z * sg * (1.f - sg) * k0 * (1.f + 3.f * k1 * (z * z));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// MFMA A operands of both layers and the feature normalisation of lane
// `lane` (rows 16m + lane%16, columns 4*(lane/16)..+3)
__device__ __forceinline__ void load_infer_operands(const float *__restrict__ P, int lane,
                                                    bf16x4 (&w1)[4], bf16x4 (&w2)[4][4],
                                                    float (&mean)[4], float (&istd)[4]) {
    const int r = lane & 15, q = lane >> 4;
    for (int m = 0; m < 4; m++) {
        const float *row = P + OFF_W1 + (16 * m + r) * kF + 4 * q;
        w1[m] = pack4(row[0], row[1], row[2], row[3]);
        for (int kc = 0; kc < 4; kc++) {
            const float *r2 = P + OFF_W2 + (16 * m + r) * kH + 16 * kc + 4 * q;
            w2[m][kc] = pack4(r2[0], r2[1], r2[2], r2[3]);
        }
    }
    for (int v = 0; v < 4; v++) {
        mean[v] = P[OFF_MEAN + 4 * q + v];
        istd[v] = P[OFF_ISTD + 4 * q + v];
    }
}

// The BF16 weights of both layers staged once per workgroup in LDS.  Rows
// are padded by four entries (8 bytes), so the sixteen rows that a wave
// reads at once start in different banks.
constexpr int kW1Row = kF + 4;                          // 20
constexpr int kW2Row = kH + 4;                          // 68
constexpr int kLdsWeights = kH * kW1Row + kH * kW2Row;  // BF16 entries

// Staging for a 512-thread workgroup, split in two: the global loads are
// issued early (before a long loop, to hide their latency) and the LDS
// stores later.  P is 16-byte aligned (its own allocation), and so are both
// weight blocks.  Thread t holds W1 float4 t (t < 256) and W2 float4s t and
// t + 512, the layer biases and output weight of unit t (t < 64), and the
// feature normalisation of its MFMA lane.
struct InferPrefetch {
    float4 w1, w2a, w2b;
    float b1, b2, w3;
    float mean[4], istd[4];
};
constexpr int kPrefetchThreads = 512;
static_assert(kH * kF / 4 <= kPrefetchThreads && kH * kH / 4 == 2 * kPrefetchThreads && kH <= kPrefetchThreads,
              "prefetch split");

__device__ __forceinline__ void prefetch_infer(const float *__restrict__ P, int tid, InferPrefetch &f) {
    const float4 *w1 = reinterpret_cast<const float4 *>(P + OFF_W1);
    const float4 *w2 = reinterpret_cast<const float4 *>(P + OFF_W2);
    f.w1 = tid < kH * kF / 4 ? w1[tid] : make_float4(0.f, 0.f, 0.f, 0.f);
    f.w2a = w2[tid];
    f.w2b = w2[tid + kPrefetchThreads];
    f.b1 = f.b2 = f.w3 = 0.f;
    if (tid < kH) {
        f.b1 = P[OFF_B1 + tid];
        f.b2 = P[OFF_B2 + tid];
        f.w3 = P[OFF_W3 + tid];
    }
    const int q = (tid & 63) >> 4;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        f.mean[v] = P[OFF_MEAN + 4 * q + v];
        f.istd[v] = P[OFF_ISTD + 4 * q + v];
    }
}

// LDS side: BF16 weights into lw (one rounding of the FP32 parameters), b1 | b2 | w3 into lb
__device__ __forceinline__ void store_infer_prefetch(const InferPrefetch &f, unsigned short *lw, float *lb,
                                                     int tid) {
    if (tid < kH * kF / 4)
        *reinterpret_cast<bf16x4 *>(lw + (tid / (kF / 4)) * kW1Row + 4 * (tid % (kF / 4))) =
            pack4(f.w1.x, f.w1.y, f.w1.z, f.w1.w);
    unsigned short *lw2 = lw + kH * kW1Row;
    const int ka = tid, kb = tid + kPrefetchThreads;
    *reinterpret_cast<bf16x4 *>(lw2 + (ka / (kH / 4)) * kW2Row + 4 * (ka % (kH / 4))) =
        pack4(f.w2a.x, f.w2a.y, f.w2a.z, f.w2a.w);
    *reinterpret_cast<bf16x4 *>(lw2 + (kb / (kH / 4)) * kW2Row + 4 * (kb % (kH / 4))) =
        pack4(f.w2b.x, f.w2b.y, f.w2b.z, f.w2b.w);
    if (tid < kH) {
        lb[tid] = f.b1;
        lb[kH + tid] = f.b2;
        lb[2 * kH + tid] = f.w3;
    }
}

// The same LDS image written by a block of any size straight from the FP32
// parameters (each weight rounded to BF16 once): W1, W2, then b1 | b2 | w3.
__device__ __forceinline__ void stage_infer_weights(const float *__restrict__ P, unsigned short *lw, float *lb,
                                                    int tid, int nthreads) {
    const float4 *w1 = reinterpret_cast<const float4 *>(P + OFF_W1);
    const float4 *w2 = reinterpret_cast<const float4 *>(P + OFF_W2);
    unsigned short *lw2 = lw + kH * kW1Row;
    for (int k = tid; k < kH * kF / 4; k += nthreads) {
        const float4 v = w1[k];
        *reinterpret_cast<bf16x4 *>(lw + (k / (kF / 4)) * kW1Row + 4 * (k % (kF / 4))) = pack4(v.x, v.y, v.z, v.w);
    }
    for (int k = tid; k < kH * kH / 4; k += nthreads) {
        const float4 v = w2[k];
        *reinterpret_cast<bf16x4 *>(lw2 + (k / (kH / 4)) * kW2Row + 4 * (k % (kH / 4))) = pack4(v.x, v.y, v.z, v.w);
    }
    for (int k = tid; k < kH; k += nthreads) {
        lb[k] = P[OFF_B1 + k];
        lb[kH + k] = P[OFF_B2 + k];
        lb[2 * kH + k] = P[OFF_W3 + k];
    }
}

// MFMA weight operands of lane `lane` from the staged LDS copy
__device__ __forceinline__ void load_weight_operands_lds(const unsigned short *lw, int lane, bf16x4 (&w1)[4],
                                                         bf16x4 (&w2)[4][4]) {
    const int r = lane & 15, q = lane >> 4;
    const unsigned short *lw2 = lw + kH * kW1Row;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        w1[m] = *reinterpret_cast<const bf16x4 *>(lw + (16 * m + r) * kW1Row + 4 * q);
#pragma unroll
        for (int kc = 0; kc < 4; kc++)
            w2[m][kc] = *reinterpret_cast<const bf16x4 *>(lw2 + (16 * m + r) * kW2Row + 16 * kc + 4 * q);
    }
}

// Inference forward of T 16-sample tiles at once (the MFMA and GELU chains
// of one tile overlap those of the others).  x[t] = raw features 4q..4q+3
// of sample lane%16 of tile t; sb1, sb2, sw3 = layer biases and output
// weights in LDS; y[t] = network output of sample lane%16 (every lane).
template <int T>
__device__ __forceinline__ void forward_infer_tiles(const bf16x4 (&w1)[4], const bf16x4 (&w2)[4][4],
                                                    const float *sb1, const float *sb2,
                                                    const float *sw3, float b3, const float (&x)[T][4],
                                                    const float (&mean)[4], const float (&istd)[4],
                                                    int q, float (&y)[T]) {
    bf16x4 xb[T];
#pragma unroll
    for (int t = 0; t < T; t++)
        xb[t] = pack4(norm_feature(x[t][0], mean[0], istd[0]), norm_feature(x[t][1], mean[1], istd[1]),
                      norm_feature(x[t][2], mean[2], istd[2]), norm_feature(x[t][3], mean[3], istd[3]));
    bf16x4 h1b[T][4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const float4 b = *reinterpret_cast<const float4 *>(sb1 + 16 * m + 4 * q);
        f32x4 acc[T];
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = mfma(w1[m], xb[t], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int t = 0; t < T; t++) {
            const float z0 = acc[t][0] + b.x, z1 = acc[t][1] + b.y, z2 = acc[t][2] + b.z,
                        z3 = acc[t][3] + b.w;
            h1b[t][m] = pack4(z0 * gelu_sig(z0), z1 * gelu_sig(z1), z2 * gelu_sig(z2), z3 * gelu_sig(z3));
        }
    }
    float ys[T];
#pragma unroll
    for (int t = 0; t < T; t++) ys[t] = 0.f;
#pragma unroll
    for (int m2 = 0; m2 < 4; m2++) {
        const float4 b = *reinterpret_cast<const float4 *>(sb2 + 16 * m2 + 4 * q);
        const float4 o = *reinterpret_cast<const float4 *>(sw3 + 16 * m2 + 4 * q);
        f32x4 acc[T];
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < 4; kc++)
#pragma unroll
            for (int t = 0; t < T; t++) acc[t] = mfma(w2[m2][kc], h1b[t][kc], acc[t]);
#pragma unroll
        for (int t = 0; t < T; t++) {
            const float z0 = acc[t][0] + b.x, z1 = acc[t][1] + b.y, z2 = acc[t][2] + b.z,
                        z3 = acc[t][3] + b.w;
            ys[t] += o.x * (z0 * gelu_sig(z0)) + o.y * (z1 * gelu_sig(z1)) +
                     o.z * (z2 * gelu_sig(z2)) + o.w * (z3 * gelu_sig(z3));
        }
    }
#pragma unroll
    for (int t = 0; t < T; t++) {
        float v = ys[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[t] = v + b3;
    }
}

}  // namespace bbx
