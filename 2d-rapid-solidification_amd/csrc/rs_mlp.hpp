// rs_mlp.hpp -- learned CALPHAD surrogate: a 2 -> 64 -> 64 -> 4 MLP trained
// with BF16 MFMA and evaluated with BF16 or FP16 MFMA (FP32 accumulate) on
// gfx950.
//
// Inputs  (a, b): a = (T - t_mid)/t_half, b = (x_Nb - 0.2)/0.2 with
//                 x_Nb = 1 - c (c = Ti fraction), i.e. x_Nb in [0, 0.4].
// Outputs: the four CALPHAD quantities (f_L, f_S, mu_L, mu_S) of
// fun.cu:1671-1720 in the basis the stepper consumes, each normalised by a
// training-set mean/std:
//   y0 = f_L - f_S            (the phi driving force; ideal mixing cancels)
//   y1 = f_S - RT[c ln c + (1-c) ln(1-c)]
//   y2 = mu_L - RT ln(c/(1-c))
//   y3 = mu_S - mu_L          (ideal mixing cancels)
// which is an invertible linear map of (f_L, f_S, mu_L, mu_S); the ideal
// RT x ln x terms are added back exactly in FP64 by the user.
// Activation: SiLU in its hard form, act(x) = x * clamp(x/6 + 1/2, 0, 1)
// (two VALU instructions: FMA with the output clamp, multiply).
// Activation formats (FMT) of the hidden layers, computed in FP32 and
// rounded once where they enter the next product:
//   kActBf16  BF16 (8 significant bits), BF16 MFMA;
//   kActSplit two BF16 parts, h = hi + lo, lo = bf16(h - hi) (~16 bits),
//             two BF16 MFMAs per product;
//   kActFp16  FP16 (11 bits; one v_cvt_pk_f16_f32 per pair), FP16 MFMA for
//             layers 2 and 3 (weights: their BF16 values held in FP16);
//   kActFp16Pk as kActFp16, but the layer output is rounded to FP16 first
//             and the activation computed in packed FP16 (v_pk_fma_f16 with
//             the clamp, v_pk_mul_f16): 1.5 VALU per activation, 3 roundings.
// Weights are BF16 values in every format, layer 1 is BF16 MFMA in every
// format, and training runs the BF16 formats.
//
// Wavefront layout ("home lanes"): each lane owns one sample (a stencil
// cell). The 64 samples of a wavefront form 4 tiles of 16: tile t is lanes
// 16t..16t+15. MFMA 16x16xK operands: A[m][k] in lane m + 16*kg(k), B[k][n]
// in lane n + 16*kg(k), D[m][n] in lane n + 16*(m/4), element m%4, with the
// same (lane group, element) -> k map for A and B. Hence:
//  * layer 1 (K32): each input is split into three BF16 parts, (a_hi,
//    a_mid, a_lo, b_hi, b_mid, b_lo, 0, 0), which carry the full FP32 value;
//    the inputs of tile t are moved into lane group 0 (ds_bpermute; the
//    other groups get copies that meet zero weights), so every tile uses
//    the same k slots and a cell's result does not depend on its lane (two
//    wavefronts that both hold a halo cell must compute the same value for
//    the solute fluxes to stay antisymmetric);
//  * layer 2 (K32): two D tiles of layer 1 concatenated are a B fragment; the
//    32 hidden units of K block kb sit at (group g, element e) ->
//    sigma(kb, g, e) = 32 kb + 16 (e/4) + 4 g + e%4, and W2 is packed with the
//    same map;
//  * layer 3 (K32): the weight fragment of tile t places W3 in rows
//    4t..4t+3, so the outputs of a sample land in its own lane (elements
//    0..3); all four tiles accumulate into one register quad.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rs {

constexpr int kH = 64;  // hidden width
constexpr int kIn = 2;
constexpr int kOut = 4;
// FP32 master parameter vector: W1[64][2] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kNParam = kOffB3 + kOut;  // 4612
// per-workgroup gradient partial layout (MFMA-shaped): dW1' 64x16 (input
// parts a_hi, a_mid, a_lo, b_hi, b_mid, b_lo in columns 0..5), db1 64,
// dW2 64x64, db2 64, dW3' 16x64, db3 16
constexpr int kQW1 = 0;
constexpr int kQB1 = kQW1 + kH * 16;
constexpr int kQW2 = kQB1 + kH;
constexpr int kQB2 = kQW2 + kH * kH;
constexpr int kQW3 = kQB2 + kH;
constexpr int kQB3 = kQW3 + 16 * kH;
constexpr int kNQ = kQB3 + 16;  // 6288

// trust table: 64 x 64 bins over (a, b) in [-1, 1]^2, bit ia of word ib
constexpr int kTrustN = 64;

enum { kActBf16 = 0, kActSplit = 1, kActFp16 = 2, kActFp16Pk = 3 };
// formats whose layers 2 and 3 run FP16 MFMA on the MlpFrags::h set
constexpr bool fp16_fmt(int fmt) { return fmt == kActFp16 || fmt == kActFp16Pk; }

// weight fragments for inference (20.8 KB, staged in LDS); w1 is BF16, w2
// and w3 BF16 (f) or FP16 (h) bits
struct MlpInfer {
    uint4 w1[4][64];     // [m-tile]            K32, k slots of lane group 0
    uint4 w2[4][2][64];  // [m-tile][k-block]   K32
    uint4 w3[4][2][64];  // [tile][k-block]     K32
    float b1[kH], b2[kH], b3[4];
};
// training adds the transposed fragments of the backward pass
struct MlpTrain {
    MlpInfer f;
    uint2 w3t[4][4][64];  // [tile][m-tile]     K16: W3^T, k-group = tile
    uint4 w2t[4][2][64];  // [m-tile][k-block]  K32: W2^T
};
// all fragments; h is the inference set of the FP16 format
struct MlpFrags : MlpTrain {
    MlpInfer h;
};

struct MlpNorm {
    float t_mid, t_half;  // T normalisation (K)
    float y_mean[kOut], y_std[kOut];
};

// training sample generator: n_uni uniform samples in [T_lo, T_hi] x
// [x_Nb 0, 0.4], then n_near samples within +-0.05 of the liquidus (even)
// and solidus (odd) compositions, tables cl[T], cs[T] of n_curve points on
// [T_lo, T_hi]
struct SampleArgs {
    float* X;   // n x 2 (a, b)
    double* Y;  // n x 4 exact targets (J/mol), basis y0..y3 above
    double* Tc; // n x 2 (T, c) raw
    int64_t n_uni, n_near;
    uint64_t seed;
    double T_lo, T_hi, t_mid, t_half;
    const double* cl;  // Ti fraction on the liquidus side, n_curve points
    const double* cs;
    int n_curve;
    double L0_bcc, L0_liq, L1_bcc, L1_liq, RR;
};

hipError_t launch_mlp_samples(const SampleArgs& a, hipStream_t s);
// exact targets (basis y0..y3, J/mol) at given (T, c) pairs
hipError_t launch_mlp_targets(const double* Tc, int64_t n, double L0_bcc, double L0_liq,
                              double L1_bcc, double L1_liq, double* Y, hipStream_t s);
// (T, c) -> network inputs (a, b)
hipError_t launch_mlp_inputs(const double* Tc, int64_t n, double t_mid, double t_half, float* X,
                             hipStream_t s);
hipError_t launch_mlp_normalise(const double* Y, float* Yn, int64_t n, const MlpNorm& nm,
                                hipStream_t s);
hipError_t launch_mlp_pack(const float* P, MlpFrags* F, hipStream_t s);
// one optimiser step over the permuted rows [row0, row0 + nb) of an n-row
// set (nb a multiple of 64; row r reads sample (r*mult + off) mod n):
// per-workgroup partial gradients, then a fixed-order reduce + Adam on the
// FP32 master weights, then the weight fragments are repacked
hipError_t launch_mlp_train_step(const float* X, const float* Yn, int64_t n, int64_t row0, int nb,
                                 int64_t mult, int64_t off, MlpFrags* F, float* partial,
                                 float* loss_part, float* P, float* m, float* v, float lr, float b1,
                                 float b2, float eps, int t, bool split, hipStream_t s);
// gradient of the mean squared error over rows [0, n) (n a multiple of 64)
// without the update; grad in master layout, loss_part per workgroup
hipError_t launch_mlp_grad(const float* X, const float* Yn, int n, const MlpFrags* F,
                           float* partial, float* loss_part, float* grad, bool split,
                           hipStream_t s);
// forward over n rows -> normalised outputs (n x 4)
hipError_t launch_mlp_eval(const float* X, int64_t n, const MlpFrags* F, float* out, int fmt,
                           hipStream_t s);
// n network inputs spread over [-0.95, 0.95]^2 (additive recurrences with
// the golden ratios), for timing the network alone
hipError_t launch_mlp_spread_inputs(float* X, int64_t n, hipStream_t s);

// ---------------------------------------------------------------------------
// device helpers shared by the training kernels and the fused stepper
// ---------------------------------------------------------------------------
#if defined(__HIPCC__)
namespace mlp {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef __bf16 b2 __attribute__((ext_vector_type(2)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// two floats -> packed BF16 pair, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
    const f2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, b2));
}
__device__ __forceinline__ uint2 pack4(f4 v) {
    return make_uint2(cvt2(v[0], v[1]), cvt2(v[2], v[3]));
}
// two floats -> packed FP16 pair, round to nearest even (v_cvt_pk_f16_f32)
__device__ __forceinline__ uint32_t cvt2h(float a, float b) {
    const f2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, h2));
}
template <int FMT>
__device__ __forceinline__ uint2 pack4f(f4 v) {
    if constexpr (FMT == kActFp16)
        return make_uint2(cvt2h(v[0], v[1]), cvt2h(v[2], v[3]));
    else
        return pack4(v);
}
// hard SiLU of a pair in packed FP16: p = fp16(z), u = clamp(p/6 + 1/2)
// (one FMA with the clamp, rounded), h = p u (rounded)
__device__ __forceinline__ uint32_t act_pk2(float z0, float z1) {
    const f2 v = {z0, z1};
    const h2 p = __builtin_convertvector(v, h2);
    const h2 c6 = {(_Float16)(1.0f / 6.0f), (_Float16)(1.0f / 6.0f)};
    const h2 c05 = {(_Float16)0.5f, (_Float16)0.5f};
    const h2 zero = {(_Float16)0.0f, (_Float16)0.0f}, one = {(_Float16)1.0f, (_Float16)1.0f};
    h2 u = __builtin_elementwise_fma(p, c6, c05);
    u = __builtin_elementwise_min(__builtin_elementwise_max(u, zero), one);
    return __builtin_bit_cast(uint32_t, p * u);
}
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

__device__ __forceinline__ f4 mfma16(uint2 a, uint2 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a),
                                                     __builtin_bit_cast(s4, b), c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32(uint4 a, uint4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                   __builtin_bit_cast(b8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32h(uint4 a, uint4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                  __builtin_bit_cast(h8, b), c, 0, 0, 0);
}
// the product of a hidden layer in format FMT
template <int FMT>
__device__ __forceinline__ f4 mfma32f(uint4 a, uint4 b, f4 c) {
    if constexpr (fp16_fmt(FMT))
        return mfma32h(a, b, c);
    else
        return mfma32(a, b, c);
}
__device__ __forceinline__ uint4 cat(uint2 lo, uint2 hi) { return make_uint4(lo.x, lo.y, hi.x, hi.y); }

__device__ __forceinline__ float act(float x) {
    return x * fminf(fmaxf(fmaf(x, 1.0f / 6.0f, 0.5f), 0.0f), 1.0f);
}
__device__ __forceinline__ float dact(float x) {
    return x <= -3.0f ? 0.0f : (x >= 3.0f ? 1.0f : fmaf(x, 1.0f / 3.0f, 0.5f));
}
__device__ __forceinline__ f4 act4(f4 z) { return f4{act(z[0]), act(z[1]), act(z[2]), act(z[3])}; }
__device__ __forceinline__ f4 dact4(f4 z) {
    return f4{dact(z[0]), dact(z[1]), dact(z[2]), dact(z[3])};
}
// bias of D rows 4g..4g+3 of m-tile mt: one 16-byte load (the bias arrays
// are 16-byte aligned), whose constant part rides in the load's offset
// field (two 4-byte loads took a vector add for each address)
__device__ __forceinline__ f4 bias4(const float* b, int mt, int lane) {
    return *reinterpret_cast<const f4*>(b + 16 * mt + 4 * (lane >> 4));
}

// the lane's own input as a K32 fragment (a_hi, a_mid, a_lo, b_hi, b_mid,
// b_lo, 0, 0): each part is the BF16 rounding of what the previous parts
// leave, so the three carry the FP32 value
__device__ __forceinline__ uint4 input_frag(float a, float b) {
    const uint32_t h = cvt2(a, b);
    const float ra = a - bf_lo(h), rb = b - bf_hi(h);
    const uint32_t m = cvt2(ra, rb);
    const uint32_t l = cvt2(ra - bf_lo(m), rb - bf_hi(m));
    // element order: x = (a_hi, a_mid), y = (a_lo, b_hi), z = (b_mid, b_lo)
    return make_uint4((h & 0xFFFFu) | (m << 16), (l & 0xFFFFu) | (h & 0xFFFF0000u),
                      (m >> 16) | (l & 0xFFFF0000u), 0u);
}

// the input fragments of tile t in every lane group; the layer-1 weight
// fragment is zero outside the k slots of group 0, so the copies in groups
// 1..3 (finite BF16 values) contribute exact zeros
__device__ __forceinline__ uint4 tile_inputs(uint4 xin, int t, int lane) {
    const int src = (16 * t + (lane & 15)) << 2;
    uint4 r;
    r.x = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)xin.x);
    r.y = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)xin.y);
    r.z = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)xin.z);
    r.w = 0u;
    return r;
}
// layer 1 of tile t: z1[mt] = W1' x + b1 for the 16 samples of tile t
__device__ __forceinline__ void layer1(const MlpInfer& F, uint4 xin, int t, int lane, f4* z1) {
    const uint4 xt = tile_inputs(xin, t, lane);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) z1[mt] = mfma32(F.w1[mt][lane], xt, bias4(F.b1, mt, lane));
}
__device__ __forceinline__ f4 unpack4(uint2 w) {
    return f4{bf_lo(w.x), bf_hi(w.x), bf_lo(w.y), bf_hi(w.y)};
}

// activations of one layer as 4 D tiles (hi: the BF16 or FP16 value; lo:
// the BF16 remainder of the split format)
template <int FMT>
struct Acts {
    uint2 hi[4];
    uint2 lo[FMT == kActSplit ? 4 : 1];
};
template <int FMT>
__device__ __forceinline__ void activate(const f4* z, Acts<FMT>& h) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        if constexpr (FMT == kActFp16Pk) {
            h.hi[mt] = make_uint2(act_pk2(z[mt][0], z[mt][1]), act_pk2(z[mt][2], z[mt][3]));
        } else {
            const f4 a = act4(z[mt]);
            h.hi[mt] = pack4f<FMT>(a);
            if constexpr (FMT == kActSplit) h.lo[mt] = pack4(a - unpack4(h.hi[mt]));
        }
    }
}

// layer 2: z2[mo] = W2 h1 + b2
template <int FMT>
__device__ __forceinline__ void layer2(const MlpInfer& F, const Acts<FMT>& h1, int lane, f4* z2) {
    const uint4 k0 = cat(h1.hi[0], h1.hi[1]), k1 = cat(h1.hi[2], h1.hi[3]);
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) {
        f4 acc = bias4(F.b2, mo, lane);
        acc = mfma32f<FMT>(F.w2[mo][0][lane], k0, acc);
        acc = mfma32f<FMT>(F.w2[mo][1][lane], k1, acc);
        if constexpr (FMT == kActSplit) {
            acc = mfma32(F.w2[mo][0][lane], cat(h1.lo[0], h1.lo[1]), acc);
            acc = mfma32(F.w2[mo][1][lane], cat(h1.lo[2], h1.lo[3]), acc);
        }
        z2[mo] = acc;
    }
}
// layer 3 of tile t accumulated into the home lanes of tile t
template <int FMT>
__device__ __forceinline__ f4 layer3(const MlpInfer& F, const Acts<FMT>& h2, int t, int lane,
                                     f4 acc) {
    acc = mfma32f<FMT>(F.w3[t][0][lane], cat(h2.hi[0], h2.hi[1]), acc);
    acc = mfma32f<FMT>(F.w3[t][1][lane], cat(h2.hi[2], h2.hi[3]), acc);
    if constexpr (FMT == kActSplit) {
        acc = mfma32(F.w3[t][0][lane], cat(h2.lo[0], h2.lo[1]), acc);
        acc = mfma32(F.w3[t][1][lane], cat(h2.lo[2], h2.lo[3]), acc);
    }
    return acc;
}

// full forward of the wavefront's 64 samples; returns the lane's own
// normalised outputs y0..y3 (F: the set of format FMT, MlpFrags::h for
// the FP16 formats)
template <int FMT>
__device__ __forceinline__ f4 forward(const MlpInfer& F, uint4 xin, int lane) {
    f4 y = {F.b3[0], F.b3[1], F.b3[2], F.b3[3]};
    // two tiles in flight: one tile's activations fill the other's MFMA
    // latency (the tiles are independent until layer 3, which accumulates
    // in tile order)
#pragma unroll
    for (int t = 0; t < 4; t += 2) {
        f4 za[4], zb[4];
        Acts<FMT> ha, hb;
        layer1(F, xin, t, lane, za);
        layer1(F, xin, t + 1, lane, zb);
        activate(za, ha);
        activate(zb, hb);
        layer2(F, ha, lane, za);
        layer2(F, hb, lane, zb);
        activate(za, ha);
        activate(zb, hb);
        y = layer3(F, ha, t, lane, y);
        y = layer3(F, hb, t + 1, lane, y);
    }
    return y;
}

}  // namespace mlp
#endif

}  // namespace rs
