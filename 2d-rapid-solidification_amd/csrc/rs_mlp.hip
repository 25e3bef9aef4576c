// rs_mlp.hip -- MFMA training (BF16) and inference (BF16, FP16) kernels of the CALPHAD
// surrogate (layouts and conventions in rs_mlp.hpp), and its exact
// training-sample generator.
#include <math.h>

#include "rs_mlp.hpp"
#include "rs_physics.hpp"

namespace rs {

using namespace mlp;

namespace {

// hidden unit of (K block kb, lane group g, element e) of a K32 fragment
__device__ __forceinline__ int sigma(int kb, int g, int e) {
    return 32 * kb + 16 * (e >> 2) + 4 * g + (e & 3);
}
__device__ __forceinline__ uint2 pack_k16(const float* v) {
    return make_uint2(cvt2(v[0], v[1]), cvt2(v[2], v[3]));
}
__device__ __forceinline__ uint4 pack_k32(const float* v) {
    return make_uint4(cvt2(v[0], v[1]), cvt2(v[2], v[3]), cvt2(v[4], v[5]), cvt2(v[6], v[7]));
}
// the BF16 values of v held in FP16 (exact for |v| >= 2^-14; below, FP16
// subnormals round them)
__device__ __forceinline__ uint4 pack_k32_h(const float* v) {
    float r[8];
    for (int e = 0; e < 8; ++e) r[e] = bf_lo(cvt2(v[e], 0.f));
    return make_uint4(cvt2h(r[0], r[1]), cvt2h(r[2], r[3]), cvt2h(r[4], r[5]), cvt2h(r[6], r[7]));
}

template <typename T>
__device__ __forceinline__ void copy_words(T* dst, const T* src, int tid, int nthr) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    constexpr int n = sizeof(T) / 16;
    static_assert(sizeof(T) % 16 == 0, "fragment blocks are whole 16-byte words");
    for (int q = tid; q < n; q += nthr) d[q] = s[q];
}

}  // namespace

// ---------------------------------------------------------------------------
// weight fragments from the FP32 master vector (one wavefront)
// ---------------------------------------------------------------------------
__global__ void mlp_pack_kernel(const float* P, MlpFrags* F) {
    const int lane = threadIdx.x;
    const int m16 = lane & 15, g = lane >> 4;
    const float* W1 = P + kOffW1;
    const float* W2 = P + kOffW2;
    const float* W3 = P + kOffW3;
    float v[8];
    for (int mt = 0; mt < 4; ++mt) {
        const int m = 16 * mt + m16;
        const bool on = g == 0;
        v[0] = v[1] = v[2] = on ? W1[m * 2 + 0] : 0.f;
        v[3] = v[4] = v[5] = on ? W1[m * 2 + 1] : 0.f;
        v[6] = v[7] = 0.f;
        F->f.w1[mt][lane] = pack_k32(v);
    }
    for (int t = 0; t < 4; ++t)
        for (int mt = 0; mt < 4; ++mt) {
            const int m = 16 * mt + m16;
            const bool on = g == t;
            for (int e = 0; e < 4; ++e) v[e] = on ? W3[e * kH + m] : 0.f;
            F->w3t[t][mt][lane] = pack_k16(v);
        }
    for (int mo = 0; mo < 4; ++mo)
        for (int kb = 0; kb < 2; ++kb) {
            const int m = 16 * mo + m16;
            for (int e = 0; e < 8; ++e) v[e] = W2[m * kH + sigma(kb, g, e)];
            F->f.w2[mo][kb][lane] = pack_k32(v);
            F->h.w2[mo][kb][lane] = pack_k32_h(v);
            for (int e = 0; e < 8; ++e) v[e] = W2[sigma(kb, g, e) * kH + m];
            F->w2t[mo][kb][lane] = pack_k32(v);
        }
    for (int t = 0; t < 4; ++t)
        for (int kb = 0; kb < 2; ++kb) {
            const bool on = m16 >= 4 * t && m16 < 4 * t + 4;
            const int o = m16 - 4 * t;
            for (int e = 0; e < 8; ++e) v[e] = on ? W3[o * kH + sigma(kb, g, e)] : 0.f;
            F->f.w3[t][kb][lane] = pack_k32(v);
            F->h.w3[t][kb][lane] = pack_k32_h(v);
        }
    for (int mt = 0; mt < 4; ++mt) F->h.w1[mt][lane] = F->f.w1[mt][lane];
    F->f.b1[lane] = F->h.b1[lane] = P[kOffB1 + lane];
    F->f.b2[lane] = F->h.b2[lane] = P[kOffB2 + lane];
    if (lane < 4) F->f.b3[lane] = F->h.b3[lane] = P[kOffB3 + lane];
}

hipError_t launch_mlp_pack(const float* P, MlpFrags* F, hipStream_t s) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3(1), dim3(64), 0, s, P, F);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// training: per 16-sample tile, forward (the inference layers), backward,
// and weight gradients by MFMA over the tile's samples (operands transposed
// through LDS); one wavefront = 64 samples per workgroup, partial sums per
// workgroup in fixed order
// ---------------------------------------------------------------------------
template <bool SPLIT>
__global__ __launch_bounds__(64) void mlp_train_kernel(const float* X, const float* Yn, int64_t n,
                                                       int64_t row0, int64_t mult, int64_t off,
                                                       float gscale, const MlpFrags* Fg,
                                                       float* partial, float* loss_part) {
    __shared__ MlpTrain F;
    // transposed BF16 tiles [feature][sample of the tile]
    __shared__ __align__(16) uint16_t sH1[kH][16], sH2[kH][16], sZ1[kH][16], sZ2[kH][16];
    __shared__ __align__(16) uint16_t sX[16][16], sY[16][16];
    const int lane = threadIdx.x;
    copy_words<MlpTrain>(&F, Fg, lane, 64);
    for (int q = lane; q < 16 * 16; q += 64) {
        (&sX[0][0])[q] = 0;
        (&sY[0][0])[q] = 0;
    }
    const int m16 = lane & 15, g = lane >> 4;

    const int64_t r = row0 + (int64_t)blockIdx.x * 64 + lane;
    const int64_t idx = (r * mult + off) % n;
    const uint4 xin = input_frag(X[2 * idx], X[2 * idx + 1]);
    const f4 yt = {Yn[4 * idx], Yn[4 * idx + 1], Yn[4 * idx + 2], Yn[4 * idx + 3]};
    __syncthreads();

    f4 gW2[4][4], gW1[4], gW3[4];
    float gb1[4][4], gb2[4][4], gb3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f4{0.f, 0.f, 0.f, 0.f};
        gW3[a] = f4{0.f, 0.f, 0.f, 0.f};
        gb3[a] = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gW2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            gb1[a][b] = 0.f;
            gb2[a][b] = 0.f;
        }
    }
    float loss = 0.f;

    for (int t = 0; t < 4; ++t) {
        // forward of tile t (its outputs land in lanes 16t..16t+15)
        f4 z1[4], z2[4];
        Acts<SPLIT ? kActSplit : kActBf16> h1, h2;
        layer1(F.f, xin, t, lane, z1);
        activate(z1, h1);
        layer2(F.f, h1, lane, z2);
        activate(z2, h2);
        const f4 y = layer3(F.f, h2, t, lane, f4{F.f.b3[0], F.f.b3[1], F.f.b3[2], F.f.b3[3]});
        // dL/dy of the mean over samples x outputs, tile-t lanes only
        f4 dy = {0.f, 0.f, 0.f, 0.f};
        if (g == t) {
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const float e = y[o] - yt[o];
                loss += e * e;
                dy[o] = gscale * e;
            }
        }
        const uint2 dyb = pack4(dy);
        // dH2 = W3^T dY, dZ2 = dH2 act'(Z2)
        f4 dz2[4];
        uint2 dz2b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const f4 dh = mfma16(F.w3t[t][mt][lane], dyb, f4{0.f, 0.f, 0.f, 0.f});
            dz2[mt] = dh * dact4(z2[mt]);
            dz2b[mt] = pack4(dz2[mt]);
        }
        // dH1 = W2^T dZ2, dZ1 = dH1 act'(Z1)
        const uint4 k0 = cat(dz2b[0], dz2b[1]), k1 = cat(dz2b[2], dz2b[3]);
        f4 dz1[4];
        uint2 dz1b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 acc = mfma32(F.w2t[mt][0][lane], k0, f4{0.f, 0.f, 0.f, 0.f});
            acc = mfma32(F.w2t[mt][1][lane], k1, acc);
            dz1[mt] = acc * dact4(z1[mt]);
            dz1b[mt] = pack4(dz1[mt]);
        }
        // bias gradients (FP32, per lane; summed over samples below)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                gb1[mt][q] += dz1[mt][q];
                gb2[mt][q] += dz2[mt][q];
            }
#pragma unroll
        for (int o = 0; o < 4; ++o) gb3[o] += dy[o];
        // transposes: D element q of lane (m16, g) of m-tile mt is feature
        // 16 mt + 4 g + q of sample m16 of the tile; the weight gradients
        // take the BF16 high part of split activations
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int f = 16 * mt + 4 * g;
            const uint2 a1 = h1.hi[mt], a2 = h2.hi[mt], a3 = dz1b[mt], a4 = dz2b[mt];
            sH1[f + 0][m16] = (uint16_t)a1.x; sH1[f + 1][m16] = (uint16_t)(a1.x >> 16);
            sH1[f + 2][m16] = (uint16_t)a1.y; sH1[f + 3][m16] = (uint16_t)(a1.y >> 16);
            sH2[f + 0][m16] = (uint16_t)a2.x; sH2[f + 1][m16] = (uint16_t)(a2.x >> 16);
            sH2[f + 2][m16] = (uint16_t)a2.y; sH2[f + 3][m16] = (uint16_t)(a2.y >> 16);
            sZ1[f + 0][m16] = (uint16_t)a3.x; sZ1[f + 1][m16] = (uint16_t)(a3.x >> 16);
            sZ1[f + 2][m16] = (uint16_t)a3.y; sZ1[f + 3][m16] = (uint16_t)(a3.y >> 16);
            sZ2[f + 0][m16] = (uint16_t)a4.x; sZ2[f + 1][m16] = (uint16_t)(a4.x >> 16);
            sZ2[f + 2][m16] = (uint16_t)a4.y; sZ2[f + 3][m16] = (uint16_t)(a4.y >> 16);
        }
        if (g == t) {
            sX[0][m16] = (uint16_t)xin.x; sX[1][m16] = (uint16_t)(xin.x >> 16);
            sX[2][m16] = (uint16_t)xin.y; sX[3][m16] = (uint16_t)(xin.y >> 16);
            sX[4][m16] = (uint16_t)xin.z; sX[5][m16] = (uint16_t)(xin.z >> 16);
            sY[0][m16] = (uint16_t)dyb.x; sY[1][m16] = (uint16_t)(dyb.x >> 16);
            sY[2][m16] = (uint16_t)dyb.y; sY[3][m16] = (uint16_t)(dyb.y >> 16);
        }
        __syncthreads();
        // K16 over the tile's samples: A rows = output features with
        // k = samples 4g..4g+3, B rows k = samples, columns = input features
        auto rd = [&](const uint16_t (*M)[16], int row) -> uint2 {
            return *reinterpret_cast<const uint2*>(&M[row][4 * g]);
        };
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
            const uint2 a2 = rd(sZ2, 16 * mo + m16);
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) gW2[mo][kt] = mfma16(a2, rd(sH1, 16 * kt + m16), gW2[mo][kt]);
            gW1[mo] = mfma16(rd(sZ1, 16 * mo + m16), rd(sX, m16), gW1[mo]);
        }
        const uint2 a3 = rd(sY, m16);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) gW3[kt] = mfma16(a3, rd(sH2, 16 * kt + m16), gW3[kt]);
        __syncthreads();
    }

    float* out = partial + (int64_t)blockIdx.x * kNQ;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = 4 * g + q;
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                out[kQW2 + (16 * mo + rr) * kH + 16 * kt + m16] = gW2[mo][kt][q];
            out[kQW1 + (16 * mo + rr) * 16 + m16] = gW1[mo][q];
        }
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) out[kQW3 + rr * kH + 16 * kt + m16] = gW3[kt][q];
    }
    // bias sums over the 16 samples (lanes) of each group, then over groups
    // for b3 (its terms sit in the home lanes)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float b1v = gb1[mt][q], b2v = gb2[mt][q];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                b1v += __shfl_xor(b1v, o, 64);
                b2v += __shfl_xor(b2v, o, 64);
            }
            if (m16 == 0) {
                out[kQB1 + 16 * mt + 4 * g + q] = b1v;
                out[kQB2 + 16 * mt + 4 * g + q] = b2v;
            }
        }
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        float b3v = gb3[o];
#pragma unroll
        for (int s = 1; s < 64; s <<= 1) b3v += __shfl_xor(b3v, s, 64);
        if (lane == 0) out[kQB3 + o] = b3v;
    }
    if (lane < 12) out[kQB3 + 4 + lane] = 0.f;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) loss += __shfl_xor(loss, s, 64);
    if (lane == 0) loss_part[blockIdx.x] = loss;
}

// gradient in master layout: fixed-order sum over the workgroup partials
__device__ __forceinline__ float grad_of(const float* partial, int nwg, int p) {
    int q0, nq = 1;
    if (p < kOffB1) {
        const int h = p / 2, k = p % 2;
        q0 = kQW1 + h * 16 + 3 * k;  // W1 multiplies the three parts of its input
        nq = 3;
    } else if (p < kOffW2) {
        q0 = kQB1 + (p - kOffB1);
    } else if (p < kOffB2) {
        q0 = kQW2 + (p - kOffW2);
    } else if (p < kOffW3) {
        q0 = kQB2 + (p - kOffB2);
    } else if (p < kOffB3) {
        q0 = kQW3 + (p - kOffW3);
    } else {
        q0 = kQB3 + (p - kOffB3);
    }
    float g = 0.f;
    for (int w = 0; w < nwg; ++w) {
        const float* pw = partial + (int64_t)w * kNQ;
        g += nq == 3 ? (pw[q0] + pw[q0 + 1]) + pw[q0 + 2] : pw[q0];
    }
    return g;
}

__global__ void mlp_adam_kernel(const float* partial, int nwg, float* P, float* m, float* v,
                                float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam) return;
    const float g = grad_of(partial, nwg, p);
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

__global__ void mlp_grad_kernel(const float* partial, int nwg, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNParam) grad[p] = grad_of(partial, nwg, p);
}

hipError_t launch_mlp_train_step(const float* X, const float* Yn, int64_t n, int64_t row0, int nb,
                                 int64_t mult, int64_t off, MlpFrags* F, float* partial,
                                 float* loss_part, float* P, float* m, float* v, float lr, float b1,
                                 float b2, float eps, int t, bool split, hipStream_t s) {
    const int nwg = nb / 64;
    const float gscale = 2.0f / (float)((double)nb * kOut);
    if (split)
        hipLaunchKernelGGL(mlp_train_kernel<true>, dim3(nwg), dim3(64), 0, s, X, Yn, n, row0, mult,
                           off, gscale, F, partial, loss_part);
    else
        hipLaunchKernelGGL(mlp_train_kernel<false>, dim3(nwg), dim3(64), 0, s, X, Yn, n, row0,
                           mult, off, gscale, F, partial, loss_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kNParam + 255) / 256), dim3(256), 0, s, partial, nwg,
                       P, m, v, lr, b1, b2, eps, bc1, bc2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_mlp_pack(P, F, s);
}

hipError_t launch_mlp_grad(const float* X, const float* Yn, int n, const MlpFrags* F,
                           float* partial, float* loss_part, float* grad, bool split,
                           hipStream_t s) {
    const int nwg = n / 64;
    const float gscale = 2.0f / (float)((double)n * kOut);
    if (split)
        hipLaunchKernelGGL(mlp_train_kernel<true>, dim3(nwg), dim3(64), 0, s, X, Yn, (int64_t)n,
                           (int64_t)0, (int64_t)1, (int64_t)0, gscale, F, partial, loss_part);
    else
        hipLaunchKernelGGL(mlp_train_kernel<false>, dim3(nwg), dim3(64), 0, s, X, Yn, (int64_t)n,
                           (int64_t)0, (int64_t)1, (int64_t)0, gscale, F, partial, loss_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((kNParam + 255) / 256), dim3(256), 0, s, partial, nwg,
                       grad);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// inference over rows (tests, trust table, rs_surrogate_time): the fused
// stepper's forward; 4 wavefronts per SIMD for the FP16 formats (<= 128
// registers without spills), 2 for the BF16 ones
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256, fp16_fmt(FMT) ? 4 : 2) void mlp_eval_kernel(
    const float* X, int64_t n, const MlpFrags* Fg, float* out) {
    __shared__ MlpInfer F;
    copy_words(&F, fp16_fmt(FMT) ? &Fg->h : &Fg->f, threadIdx.x, blockDim.x);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t r = ((int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * 64 + lane;
    const int64_t rc = r < n ? r : n - 1;
    const f4 y = forward<FMT>(F, input_frag(X[2 * rc], X[2 * rc + 1]), lane);
    if (r < n)
        for (int o = 0; o < 4; ++o) out[4 * r + o] = y[o];
}

__global__ void mlp_spread_inputs_kernel(float* X, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double u = (double)q * 0.6180339887498949, v = (double)q * 0.7548776662466927;
    X[2 * q] = (float)(-0.95 + 1.9 * (u - floor(u)));
    X[2 * q + 1] = (float)(-0.95 + 1.9 * (v - floor(v)));
}

hipError_t launch_mlp_spread_inputs(float* X, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(mlp_spread_inputs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       X, n);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(const float* X, int64_t n, const MlpFrags* F, float* out, int fmt,
                           hipStream_t s) {
    const int64_t waves = (n + 63) / 64;
    const dim3 grid((unsigned)((waves + 3) / 4));
    if (fmt == kActFp16)
        hipLaunchKernelGGL(mlp_eval_kernel<kActFp16>, grid, dim3(256), 0, s, X, n, F, out);
    else if (fmt == kActFp16Pk)
        hipLaunchKernelGGL(mlp_eval_kernel<kActFp16Pk>, grid, dim3(256), 0, s, X, n, F, out);
    else if (fmt == kActSplit)
        hipLaunchKernelGGL(mlp_eval_kernel<kActSplit>, grid, dim3(256), 0, s, X, n, F, out);
    else if (fmt == kActBf16)
        hipLaunchKernelGGL(mlp_eval_kernel<kActBf16>, grid, dim3(256), 0, s, X, n, F, out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// training samples with exact targets (FP64 CALPHAD)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t q, int k) {
    const uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ULL * (2 * q + k + 1));
    return (double)(h >> 11) * 0x1.0p-53;
}

// y0 = f_L - f_S, y1 = f_S - ideal, y2 = mu_L - RT ln(c/(1-c)), y3 = mu_S - mu_L
// (fun.cu:1671-1720 without the ideal-mixing terms)
__device__ __forceinline__ void exact_targets(double L0s, double L0l, double L1s, double L1l,
                                              double T, double c, double* y) {
    const Unaries u = sgte_unaries(T, log(T));
    const double om = 1.0 - c;
    const double om2 = 1.0 - 2.0 * c;
    const double six = 1.0 - 6.0 * c + 6.0 * (c * c);
    const double fl = om * u.nb_liq + c * u.ti_liq + c * om * (L0l + L1l * om2);
    const double fs = om * u.nb_bcc + c * u.ti_bcc + c * om * (L0s + L1s * om2);
    const double ml = u.ti_liq - u.nb_liq + L0l * om2 + L1l * six;
    const double ms = u.ti_bcc - u.nb_bcc + L0s * om2 + L1s * six;
    y[0] = fl - fs;
    y[1] = fs;
    y[2] = ml;
    y[3] = ms - ml;
}

__global__ void mlp_sample_kernel(SampleArgs a) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = a.n_uni + a.n_near;
    if (q >= n) return;
    const double T = a.T_lo + (a.T_hi - a.T_lo) * u01(a.seed, q, 0);
    double xnb;
    if (q < a.n_uni) {
        xnb = 0.4 * u01(a.seed, q, 1);
    } else {
        const double s = (T - a.T_lo) / (a.T_hi - a.T_lo) * (a.n_curve - 1);
        int k0 = (int)s;
        if (k0 > a.n_curve - 2) k0 = a.n_curve - 2;
        const double w = s - k0;
        const double* cur = ((q - a.n_uni) & 1) ? a.cs : a.cl;
        const double cb = cur[k0] + w * (cur[k0 + 1] - cur[k0]);
        xnb = (1.0 - cb) + 0.05 * (2.0 * u01(a.seed, q, 1) - 1.0);
        xnb = fmin(fmax(xnb, 0.0), 0.4);
    }
    const double c = 1.0 - xnb;
    double y[4];
    exact_targets(a.L0_bcc, a.L0_liq, a.L1_bcc, a.L1_liq, T, c, y);
    a.X[2 * q] = (float)((T - a.t_mid) / a.t_half);
    a.X[2 * q + 1] = (float)((xnb - 0.2) / 0.2);
    for (int k = 0; k < 4; ++k) a.Y[4 * q + k] = y[k];
    a.Tc[2 * q] = T;
    a.Tc[2 * q + 1] = c;
}

hipError_t launch_mlp_samples(const SampleArgs& a, hipStream_t s) {
    const int64_t n = a.n_uni + a.n_near;
    hipLaunchKernelGGL(mlp_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void mlp_targets_kernel(const double* Tc, int64_t n, double L0s, double L0l, double L1s,
                                   double L1l, double* Y) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    exact_targets(L0s, L0l, L1s, L1l, Tc[2 * q], Tc[2 * q + 1], Y + 4 * q);
}

hipError_t launch_mlp_targets(const double* Tc, int64_t n, double L0_bcc, double L0_liq,
                              double L1_bcc, double L1_liq, double* Y, hipStream_t s) {
    hipLaunchKernelGGL(mlp_targets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Tc,
                       n, L0_bcc, L0_liq, L1_bcc, L1_liq, Y);
    return hipGetLastError();
}

__global__ void mlp_inputs_kernel(const double* Tc, int64_t n, double t_mid, double t_half,
                                  float* X) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    X[2 * q] = (float)((Tc[2 * q] - t_mid) / t_half);
    X[2 * q + 1] = (float)(((1.0 - Tc[2 * q + 1]) - 0.2) / 0.2);
}

hipError_t launch_mlp_inputs(const double* Tc, int64_t n, double t_mid, double t_half, float* X,
                             hipStream_t s) {
    hipLaunchKernelGGL(mlp_inputs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Tc,
                       n, t_mid, t_half, X);
    return hipGetLastError();
}

__global__ void mlp_normalise_kernel(const double* Y, float* Yn, int64_t n, MlpNorm nm) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int k = 0; k < 4; ++k)
        Yn[4 * q + k] = (float)((Y[4 * q + k] - nm.y_mean[k]) / nm.y_std[k]);
}

hipError_t launch_mlp_normalise(const double* Y, float* Yn, int64_t n, const MlpNorm& nm,
                                hipStream_t s) {
    hipLaunchKernelGGL(mlp_normalise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Y,
                       Yn, n, nm);
    return hipGetLastError();
}

}  // namespace rs
