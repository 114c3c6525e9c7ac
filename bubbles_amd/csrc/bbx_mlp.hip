// bbx_mlp.hip -- learned PCISPH pressure warm start: a 16 -> 64 -> 64 -> 1
// MLP (GELU, tanh form) with BF16 weights/activations on the MFMA units and
// FP32 accumulation, its backward pass, and Adam -- all on gfx950.
//
// The reference has no learned component; this file is this library's own.
//
// Layout.  A wave processes tiles of 16 samples.  Every layer is computed
// transposed, H^T = W . X^T, with v_mfma_f32_16x16x16_bf16: for an MFMA
// tile the A operand (weights) has lane l holding row l%16 and columns
// 4*(l/16)+0..3, the B operand (activations) has lane l holding column
// (sample) l%16 and rows 4*(l/16)+0..3, and the FP32 result has lane l
// holding column l%16 and rows 4*(l/16)+0..3.  The result layout of one
// layer is therefore the B layout of the next, so activations never leave
// registers in the forward pass.  Weight gradients contract over samples;
// their operands are transposed through a per-wave LDS tile.
#include "bbx_mlp_device.h"

namespace bbx {

constexpr int kWaves = 4;         // waves per workgroup
constexpr int kTilesPerWave = 4;  // training tiles per wave per step

// Weights in MFMA A layout, loaded once per wave.
struct Weights {
    bf16x4 w1[4];       // W1 rows 16m + l%16, cols 4q..4q+3
    bf16x4 w2[4][4];    // W2 block (m2, kc): rows 16m2 + l%16, cols 16kc + 4q..
    float b1[4][4], b2[4][4], w3[4][4];
    float b3;
    float mean[4], istd[4];  // features 4q..4q+3
};

__device__ __forceinline__ void load_weights(const float *__restrict__ P, Weights &w, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int m = 0; m < 4; m++) {
        const float *row = P + OFF_W1 + (16 * m + r) * kF + 4 * q;
        w.w1[m] = pack4(row[0], row[1], row[2], row[3]);
        for (int kc = 0; kc < 4; kc++) {
            const float *r2 = P + OFF_W2 + (16 * m + r) * kH + 16 * kc + 4 * q;
            w.w2[m][kc] = pack4(r2[0], r2[1], r2[2], r2[3]);
        }
        for (int v = 0; v < 4; v++) {
            int h = 16 * m + 4 * q + v;
            w.b1[m][v] = P[OFF_B1 + h];
            w.b2[m][v] = P[OFF_B2 + h];
            w.w3[m][v] = P[OFF_W3 + h];
        }
    }
    w.b3 = P[OFF_B3];
    for (int v = 0; v < 4; v++) {
        w.mean[v] = P[OFF_MEAN + 4 * q + v];
        w.istd[v] = P[OFF_ISTD + 4 * q + v];
    }
}

struct Acts {
    f32x4 z1[4], z2[4];     // pre-activations (result layout)
    float g1[4][4], d1[4][4], g2[4][4], d2[4][4];
    float xn[4];            // normalised features 4q..4q+3 of sample l%16
    float y;                // output of sample l%16 (all lanes)
};

// forward of one 16-sample tile; x4 = raw features 4q..4q+3 of sample l%16
__device__ __forceinline__ void forward_tile(const Weights &w, const float x4[4], Acts &a, int lane) {
    for (int v = 0; v < 4; v++) a.xn[v] = norm_feature(x4[v], w.mean[v], w.istd[v]);
    bf16x4 xb = pack4(a.xn[0], a.xn[1], a.xn[2], a.xn[3]);
    bf16x4 h1b[4];
    for (int m = 0; m < 4; m++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(w.w1[m], xb, acc);
        for (int v = 0; v < 4; v++) {
            float z = acc[v] + w.b1[m][v];
            a.z1[m][v] = z;
            gelu_fwd(z, a.g1[m][v], a.d1[m][v]);
        }
        h1b[m] = pack4(a.g1[m][0], a.g1[m][1], a.g1[m][2], a.g1[m][3]);
    }
    float ys = 0.f;
    for (int m2 = 0; m2 < 4; m2++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < 4; kc++) acc = mfma(w.w2[m2][kc], h1b[kc], acc);
        for (int v = 0; v < 4; v++) {
            float z = acc[v] + w.b2[m2][v];
            a.z2[m2][v] = z;
            gelu_fwd(z, a.g2[m2][v], a.d2[m2][v]);
            ys += w.w3[m2][v] * a.g2[m2][v];
        }
    }
    ys += __shfl_xor(ys, 16, 64);
    ys += __shfl_xor(ys, 32, 64);
    a.y = ys + w.b3;
}

// ---------------------------------------------------------------------------
// inference: prediction_i = y_i * out_scale + resid * x_i[PPREV] (target
// units).  Solver path (clamp = 1): out_i = scale * alpha * clamp(prediction,
// 0, pcap); raw path (clamp = 0): out_i = scale * prediction.
// Features of sample i, feature f at x[f * fstride + i * sstride].
// Each wave keeps the MFMA weight operands in registers (biases and the
// output layer in LDS) and runs two 16-sample tiles at a time, so the MFMA
// and GELU chains of one tile overlap those of the other.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * kWaves) k_mlp_infer(const float *__restrict__ P,
                                                           const float *__restrict__ x, int n,
                                                           int fstride, int sstride, float scale,
                                                           int clamp, float *__restrict__ out) {
    __shared__ __align__(16) float sb1[kH];
    __shared__ __align__(16) float sb2[kH];
    __shared__ __align__(16) float sw3[kH];
    if (threadIdx.x < kH) {
        sb1[threadIdx.x] = P[OFF_B1 + threadIdx.x];
        sb2[threadIdx.x] = P[OFF_B2 + threadIdx.x];
        sw3[threadIdx.x] = P[OFF_W3 + threadIdx.x];
    }
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * kWaves;
    const int q = lane >> 4, j = lane & 15;
    bf16x4 w1[4], w2[4][4];
    float mean[4], istd[4];
    load_infer_operands(P, lane, w1, w2, mean, istd);
    const float b3 = P[OFF_B3];
    const float oscale = P[OFF_OSCALE], alpha = P[OFF_ALPHA], pcap = P[OFF_PCAP], resid = P[OFF_RESID];
    __syncthreads();
    const int tiles = (n + 15) / 16;
    for (int t0 = 2 * wave; t0 < tiles; t0 += 2 * nwaves) {
        float xv[2][4];
        int idx[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            idx[t] = (t0 + t) * 16 + j;
#pragma unroll
            for (int v = 0; v < 4; v++)
                xv[t][v] = idx[t] < n ? x[(size_t)(4 * q + v) * fstride + (size_t)idx[t] * sstride] : 0.f;
        }
        float yn[2];
        forward_infer_tiles<2>(w1, w2, sb1, sb2, sw3, b3, xv, mean, istd, q, yn);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int i = idx[t];
            if (lane < 16 && i < n) {
                float y = yn[t] * oscale;
                if (resid != 0.f) y += resid * x[(size_t)BBX_F_PPREV * fstride + (size_t)i * sstride];
                if (clamp) y = alpha * fminf(fmaxf(y, 0.f), pcap);
                out[i] = y * scale;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// training: forward + backward over kTilesPerWave tiles per wave; per-wave
// gradient partials, reduced in a fixed order by k_mlp_reduce.
// Loss = 0.5 / B * sum (y_j - t_j / out_scale)^2 over the batch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(64 * kWaves) k_mlp_grad(const float *__restrict__ P,
                                                          const float *__restrict__ xs,
                                                          const float *__restrict__ ts, int n,
                                                          int batch, int random, unsigned long long seed,
                                                          float expectile,
                                                          float *__restrict__ gpart,
                                                          float *__restrict__ lpart) {
    __shared__ unsigned short lds[kWaves][3][kH * 16];  // dz2, h1, dz1 tiles [hidden][sample]
    __shared__ unsigned short ldx[kWaves][kF * 16];     // x tile [feature][sample]
    const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
    const int wave = blockIdx.x * kWaves + wl;
    const int q = lane >> 4, j = lane & 15;
    Weights w;
    load_weights(P, w, lane);
    // W2^T in A layout for dH1 = W2^T dZ2: block (kc, m2), rows 16kc + l%16
    bf16x4 w2t[4][4];
    for (int kc = 0; kc < 4; kc++)
        for (int m2 = 0; m2 < 4; m2++) {
            float e[4];
            for (int v = 0; v < 4; v++) e[v] = P[OFF_W2 + (16 * m2 + 4 * q + v) * kH + 16 * kc + j];
            w2t[kc][m2] = pack4(e[0], e[1], e[2], e[3]);
        }
    const float inv_os = 1.f / P[OFF_OSCALE];
    const float inv_b = 1.f / (float)batch;
    f32x4 gW2[4][4], gW1[4];
    for (int a = 0; a < 4; a++) {
        gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < 4; b++) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1[4][4] = {}, gb2[4][4] = {}, gw3[4][4] = {};
    float gb3 = 0.f, loss = 0.f;
    unsigned short *T2 = lds[wl][0], *T1 = lds[wl][1], *T3 = lds[wl][2], *TX = ldx[wl];
    for (int tt = 0; tt < kTilesPerWave; tt++) {
        const int bpos = (wave * kTilesPerWave + tt) * 16 + j;  // position in the batch
        const bool valid = bpos < batch;
        int idx = 0;
        if (valid) idx = random ? (int)(mix64(seed ^ ((unsigned long long)bpos << 20)) % (unsigned long long)n) : bpos;
        float x4[4];
        for (int v = 0; v < 4; v++) x4[v] = valid ? xs[(size_t)idx * kF + 4 * q + v] : 0.f;
        Acts a;
        forward_tile(w, x4, a, lane);
        const float tj = valid ? ts[idx] * inv_os : 0.f;
        const float diff = valid ? (a.y - tj) : 0.f;
        // expectile weight: 2 (1 - tau) above the target, 2 tau below (1 at tau = 0.5)
        const float wgt = diff > 0.f ? 2.f * (1.f - expectile) : 2.f * expectile;
        const float dy = wgt * diff * inv_b;
        if (lane < 16) {
            loss += 0.5f * wgt * diff * diff * inv_b;
            gb3 += dy;
        }
        // layer 3 and layer-2 activations
        float dz2[4][4];
        for (int m2 = 0; m2 < 4; m2++)
            for (int v = 0; v < 4; v++) {
                gw3[m2][v] += dy * a.g2[m2][v];
                dz2[m2][v] = w.w3[m2][v] * dy * a.d2[m2][v];
                gb2[m2][v] += dz2[m2][v];
            }
        // dH1 = W2^T dZ2 (dz2 already in B layout)
        bf16x4 dz2b[4];
        for (int m2 = 0; m2 < 4; m2++) dz2b[m2] = pack4(dz2[m2][0], dz2[m2][1], dz2[m2][2], dz2[m2][3]);
        float dz1[4][4];
        for (int kc = 0; kc < 4; kc++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int m2 = 0; m2 < 4; m2++) acc = mfma(w2t[kc][m2], dz2b[m2], acc);
            for (int v = 0; v < 4; v++) {
                dz1[kc][v] = acc[v] * a.d1[kc][v];
                gb1[kc][v] += dz1[kc][v];
            }
        }
        // transpose dz2, h1, dz1, x through LDS: tile[hidden][sample]
        __syncthreads();
        for (int m = 0; m < 4; m++)
            for (int v = 0; v < 4; v++) {
                const int hrow = 16 * m + 4 * q + v;
                T2[hrow * 16 + j] = f2bf(dz2[m][v]);
                T1[hrow * 16 + j] = f2bf(a.g1[m][v]);
                T3[hrow * 16 + j] = f2bf(dz1[m][v]);
            }
        for (int v = 0; v < 4; v++) TX[(4 * q + v) * 16 + j] = f2bf(a.xn[v]);
        __syncthreads();
        // A operands: rows (hidden) 16m + l%16, cols (samples) 4q..4q+3
        bf16x4 A2[4], A3[4], B1[4];
        for (int m = 0; m < 4; m++) {
            const int hrow = 16 * m + j;
            for (int v = 0; v < 4; v++) {
                A2[m][v] = (short)T2[hrow * 16 + 4 * q + v];
                A3[m][v] = (short)T3[hrow * 16 + 4 * q + v];
                B1[m][v] = (short)T1[hrow * 16 + 4 * q + v];  // B: rows samples 4q.., col hidden 16m + l%16
            }
        }
        bf16x4 BX;
        for (int v = 0; v < 4; v++) BX[v] = (short)TX[j * 16 + 4 * q + v];
        // dW2[i][k] += sum_j dz2[i][j] h1[k][j]
        for (int m2 = 0; m2 < 4; m2++)
            for (int kc = 0; kc < 4; kc++) gW2[m2][kc] = mfma(A2[m2], B1[kc], gW2[m2][kc]);
        // dW1[i][f] += sum_j dz1[i][j] x[f][j]
        for (int m = 0; m < 4; m++) gW1[m] = mfma(A3[m], BX, gW1[m]);
    }
    // per-wave partials
    float *g = gpart + (size_t)wave * kTrainable;
    for (int m = 0; m < 4; m++)
        for (int v = 0; v < 4; v++) {
            const int row = 16 * m + 4 * q + v;
            g[OFF_W1 + row * kF + j] = gW1[m][v];
            for (int kc = 0; kc < 4; kc++) g[OFF_W2 + row * kH + 16 * kc + j] = gW2[m][kc][v];
        }
    // reduce per-hidden-unit sums over the 16 samples (lanes sharing q)
    for (int m = 0; m < 4; m++)
        for (int v = 0; v < 4; v++) {
            float s1 = gb1[m][v], s2 = gb2[m][v], s3 = gw3[m][v];
            for (int off = 1; off < 16; off <<= 1) {
                s1 += __shfl_xor(s1, off, 64);
                s2 += __shfl_xor(s2, off, 64);
                s3 += __shfl_xor(s3, off, 64);
            }
            if (j == 0) {
                const int h = 16 * m + 4 * q + v;
                g[OFF_B1 + h] = s1;
                g[OFF_B2 + h] = s2;
                g[OFF_W3 + h] = s3;
            }
        }
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        loss += __shfl_xor(loss, off, 64);
    }
    if (lane == 0) {
        g[OFF_B3] = gb3;
        lpart[wave] = loss;
    }
}

// grad[p] = sum over waves in index order (deterministic); optional Adam step
__global__ void k_mlp_reduce(const float *__restrict__ gpart, const float *__restrict__ lpart,
                             int nwaves, float *__restrict__ grad, float *__restrict__ loss_out,
                             float *__restrict__ P, float *__restrict__ m1, float *__restrict__ m2,
                             int adam, float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float l = 0.f;
        for (int w = 0; w < nwaves; w++) l += lpart[w];
        *loss_out = l;
    }
    if (p >= kTrainable) return;
    float s = 0.f;
    for (int w = 0; w < nwaves; w++) s += gpart[(size_t)w * kTrainable + p];
    grad[p] = s;
    if (adam) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        float a = b1 * m1[p] + (1.f - b1) * s;
        float b = b2 * m2[p] + (1.f - b2) * s * s;
        m1[p] = a;
        m2[p] = b;
        P[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    }
}

}  // namespace bbx

using namespace bbx;

struct bbx_mlp {
    float *params = nullptr;   // kParams FP32 (device)
    float *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *loss = nullptr;
    float *gpart = nullptr, *lpart = nullptr;
    int gpart_waves = 0;
    float *data_x = nullptr, *data_t = nullptr;  // resident training set (sample-major)
    int data_n = 0;
    float *tmp_x = nullptr, *tmp_y = nullptr;
    int tmp_n = 0;
    int adam_t = 0;
    float expectile = 0.5f;    // training loss: 0.5 = mean squared error
    hipStream_t stream = nullptr;
    int refs = 0;          // solvers this network is attached to
    bool dead = false;     // destroyed while attached: freed at the last release
};

namespace bbx {

static int ensure_tmp(bbx_mlp *m, int n) {
    if (n <= m->tmp_n) return BBX_OK;
    if (m->tmp_x) (void)hipFree(m->tmp_x);
    if (m->tmp_y) (void)hipFree(m->tmp_y);
    m->tmp_x = m->tmp_y = nullptr;
    m->tmp_n = 0;
    BBX_HIP_CHECK(hipMalloc((void **)&m->tmp_x, sizeof(float) * (size_t)n * kF));
    BBX_HIP_CHECK(hipMalloc((void **)&m->tmp_y, sizeof(float) * (size_t)n));
    m->tmp_n = n;
    return BBX_OK;
}

static int ensure_gpart(bbx_mlp *m, int waves) {
    if (waves <= m->gpart_waves) return BBX_OK;
    if (m->gpart) (void)hipFree(m->gpart);
    if (m->lpart) (void)hipFree(m->lpart);
    m->gpart = m->lpart = nullptr;
    m->gpart_waves = 0;
    BBX_HIP_CHECK(hipMalloc((void **)&m->gpart, sizeof(float) * (size_t)waves * kTrainable));
    BBX_HIP_CHECK(hipMalloc((void **)&m->lpart, sizeof(float) * (size_t)waves));
    m->gpart_waves = waves;
    return BBX_OK;
}

static int infer_launch(bbx_mlp *m, const float *x, int n, int fstride, int sstride, float scale,
                        int clamp0, float *out, hipStream_t st) {
    if (n <= 0) return BBX_OK;
    // two tiles per wave per pass; about four passes per wave at the largest
    // sizes so the per-wave weight load is amortised
    const int tiles = (n + 15) / 16;
    int blocks = (tiles + 2 * kWaves - 1) / (2 * kWaves);
    if (blocks > 1024) blocks = 1024;
    k_mlp_infer<<<blocks, 64 * kWaves, 0, st>>>(m->params, x, n, fstride, sstride, scale, clamp0, out);
    BBX_HIP_CHECK(hipGetLastError());
    return BBX_OK;
}

// one gradient (+ optional Adam) pass over a batch of samples resident on the device
static int grad_launch(bbx_mlp *m, const float *xs, const float *ts, int n, int batch, int random,
                       unsigned long long seed, int adam, float lr) {
    const int tiles = (batch + 15) / 16;
    const int waves_needed = (tiles + kTilesPerWave - 1) / kTilesPerWave;
    const int blocks = (waves_needed + kWaves - 1) / kWaves;
    const int nwaves = blocks * kWaves;
    BBX_CHECK(ensure_gpart(m, nwaves));
    k_mlp_grad<<<blocks, 64 * kWaves, 0, m->stream>>>(m->params, xs, ts, n, batch, random, seed,
                                                       m->expectile,
                                                       m->gpart, m->lpart);
    BBX_HIP_CHECK(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (adam) {
        m->adam_t++;
        bc1 = 1.f - powf(0.9f, (float)m->adam_t);
        bc2 = 1.f - powf(0.999f, (float)m->adam_t);
    }
    k_mlp_reduce<<<(kTrainable + 255) / 256, 256, 0, m->stream>>>(m->gpart, m->lpart, nwaves, m->grad,
                                                                  m->loss, m->params, m->m1, m->m2,
                                                                  adam, lr, bc1, bc2);
    BBX_HIP_CHECK(hipGetLastError());
    return BBX_OK;
}


}  // namespace bbx

extern "C" {

int bbx_mlp_param_count(int *count) {
    if (!count) return fail(BBX_E_INVALID, "null count");
    *count = kParams;
    return BBX_OK;
}

int bbx_mlp_set_expectile(bbx_mlp *m, float tau) {
    if (!m) return fail(BBX_E_INVALID, "null network");
    if (!(tau > 0.f && tau < 1.f)) return fail(BBX_E_INVALID, "expectile must lie in (0, 1)");
    m->expectile = tau;
    return BBX_OK;
}

int bbx_mlp_create(int seed, bbx_mlp **out) {
    if (!out) return fail(BBX_E_INVALID, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(BBX_E_NODEVICE, "no HIP device visible");
    }
    bbx_mlp *m = new bbx_mlp();
    // deterministic Glorot-uniform init (SplitMix64 stream)
    std::vector<float> p(kParams, 0.f);
    unsigned long long st = (unsigned long long)seed * 0x9E3779B97F4A7C15ull + 12345ull;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int cnt, int fan_in, int fan_out) {
        double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < cnt; i++) p[off + i] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(OFF_W1, kH * kF, kF, kH);
    fill(OFF_W2, kH * kH, kH, kH);
    fill(OFF_W3, kH, kH, 1);
    for (int f = 0; f < kF; f++) p[OFF_ISTD + f] = 1.f;
    p[OFF_OSCALE] = 1.f;
    p[OFF_ALPHA] = 1.f;
    p[OFF_PCAP] = 3.0e38f;
    p[OFF_RESID] = 0.f;
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&m->params, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc((void **)&m->m1, sizeof(float) * kTrainable);
    if (e == hipSuccess) e = hipMalloc((void **)&m->m2, sizeof(float) * kTrainable);
    if (e == hipSuccess) e = hipMalloc((void **)&m->grad, sizeof(float) * kTrainable);
    if (e == hipSuccess) e = hipMalloc((void **)&m->loss, sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(m->params, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m1, 0, sizeof(float) * kTrainable);
    if (e == hipSuccess) e = hipMemset(m->m2, 0, sizeof(float) * kTrainable);
    if (e != hipSuccess) {
        bbx_mlp_destroy(m);
        return fail(BBX_E_HIP, std::string("mlp create: ") + hipGetErrorString(e));
    }
    *out = m;
    return BBX_OK;
}

static void mlp_free(bbx_mlp *m) {
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float *bufs[] = {m->params, m->m1, m->m2, m->grad, m->loss, m->gpart, m->lpart,
                     m->data_x, m->data_t, m->tmp_x, m->tmp_y};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

// A network attached to a solver stays alive until the last solver releases
// it, whatever order the caller destroys them in.
int bbx_mlp_destroy(bbx_mlp *m) {
    if (!m) return BBX_OK;
    if (m->refs > 0) {
        m->dead = true;
        return BBX_OK;
    }
    mlp_free(m);
    return BBX_OK;
}

int bbx_mlp_get_params(bbx_mlp *m, float *out) {
    if (!m || !out) return fail(BBX_E_INVALID, "null argument");
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    BBX_HIP_CHECK(hipMemcpy(out, m->params, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    return BBX_OK;
}

int bbx_mlp_set_params(bbx_mlp *m, const float *in) {
    if (!m || !in) return fail(BBX_E_INVALID, "null argument");
    if (!(in[OFF_OSCALE] != 0.f)) return fail(BBX_E_INVALID, "target scale must be non-zero");
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    BBX_HIP_CHECK(hipMemcpy(m->params, in, sizeof(float) * kParams, hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemset(m->m1, 0, sizeof(float) * kTrainable));
    BBX_HIP_CHECK(hipMemset(m->m2, 0, sizeof(float) * kTrainable));
    m->adam_t = 0;
    return BBX_OK;
}

int bbx_mlp_infer(bbx_mlp *m, const float *x, int n, float *y) {
    if (!m || n < 0 || (n > 0 && (!x || !y))) return fail(BBX_E_INVALID, "bad arguments");
    if (n == 0) return BBX_OK;
    BBX_CHECK(ensure_tmp(m, n));
    BBX_HIP_CHECK(hipMemcpyAsync(m->tmp_x, x, sizeof(float) * (size_t)n * kF, hipMemcpyHostToDevice, m->stream));
    // raw network output times the target scale (no clamp)
    BBX_CHECK(infer_launch(m, m->tmp_x, n, 1, kF, 1.f, 0, m->tmp_y, m->stream));
    BBX_HIP_CHECK(hipMemcpyAsync(y, m->tmp_y, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, m->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    return BBX_OK;
}

static int upload_batch(bbx_mlp *m, const float *x, const float *t, int n) {
    BBX_CHECK(ensure_tmp(m, n));
    BBX_HIP_CHECK(hipMemcpyAsync(m->tmp_x, x, sizeof(float) * (size_t)n * kF, hipMemcpyHostToDevice, m->stream));
    BBX_HIP_CHECK(hipMemcpyAsync(m->tmp_y, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream));
    return BBX_OK;
}

int bbx_mlp_gradients(bbx_mlp *m, const float *x, const float *t, int n, float *grad, float *loss) {
    if (!m || n <= 0 || !x || !t || !grad) return fail(BBX_E_INVALID, "bad arguments");
    BBX_CHECK(upload_batch(m, x, t, n));
    BBX_CHECK(grad_launch(m, m->tmp_x, m->tmp_y, n, n, 0, 0, 0, 0.f));
    std::vector<float> g(kTrainable);
    BBX_HIP_CHECK(hipMemcpyAsync(g.data(), m->grad, sizeof(float) * kTrainable, hipMemcpyDeviceToHost, m->stream));
    float l = 0.f;
    BBX_HIP_CHECK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    for (int i = 0; i < kParams; i++) grad[i] = i < kTrainable ? g[i] : 0.f;
    if (loss) *loss = l;
    return BBX_OK;
}

int bbx_mlp_train_step(bbx_mlp *m, const float *x, const float *t, int n, float lr, float *loss) {
    if (!m || n <= 0 || !x || !t) return fail(BBX_E_INVALID, "bad arguments");
    BBX_CHECK(upload_batch(m, x, t, n));
    BBX_CHECK(grad_launch(m, m->tmp_x, m->tmp_y, n, n, 0, 0, 1, lr));
    float l = 0.f;
    BBX_HIP_CHECK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
    return BBX_OK;
}

int bbx_mlp_train(bbx_mlp *m, const float *x, const float *t, int n, int steps, int batch, float lr,
                  int seed, float *final_loss) {
    if (!m || n <= 0 || !x || !t || steps < 0 || batch <= 0) return fail(BBX_E_INVALID, "bad arguments");
    if (m->data_x) (void)hipFree(m->data_x);
    if (m->data_t) (void)hipFree(m->data_t);
    m->data_x = m->data_t = nullptr;
    BBX_HIP_CHECK(hipMalloc((void **)&m->data_x, sizeof(float) * (size_t)n * kF));
    BBX_HIP_CHECK(hipMalloc((void **)&m->data_t, sizeof(float) * (size_t)n));
    BBX_HIP_CHECK(hipMemcpy(m->data_x, x, sizeof(float) * (size_t)n * kF, hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemcpy(m->data_t, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    m->data_n = n;
    for (int s = 0; s < steps; s++) {
        unsigned long long sd = ((unsigned long long)(unsigned)seed << 32) ^ (unsigned long long)s * 0x2545F4914F6CDD1Dull;
        BBX_CHECK(grad_launch(m, m->data_x, m->data_t, n, batch, 1, sd, 1, lr));
    }
    float l = 0.f;
    BBX_HIP_CHECK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(m->stream));
    if (final_loss) *final_loss = l;
    return BBX_OK;
}

}  // extern "C"

namespace bbx {
const float *mlp_params_device(const bbx_mlp *m) { return m ? m->params : nullptr; }
void mlp_retain(bbx_mlp *m) {
    if (m) m->refs++;
}
void mlp_release(bbx_mlp *m) {
    if (!m) return;
    if (--m->refs <= 0 && m->dead) mlp_free(m);
}
}  // namespace bbx
