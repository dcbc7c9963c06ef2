// Learned troubled-cell sensor on the matrix cores (gfx950): an MLP 16 -> 32 -> 32 -> 1 (ReLU)
// from the 16 Legendre modal density coefficients of a p = 3 element to its sensor value sigma.
// It stands in for the Persson sigma of sense! (Interfaces.jl:79-87, PerssonAV2D.jl:20-59); the
// element viscosity formula that consumes sigma is unchanged.
//
// One wave evaluates 16 elements with v_mfma_f32_16x16x16bf16_1k: bfloat16 operands, FP32
// accumulation.  Operand layout of that instruction (checked on the device): lane l holds
//   A[i = l % 16][k = 4 (l / 16) + r]   (r = 0..3, four bf16),
//   B[k = 4 (l / 16) + r][n = l % 16],
//   D[i = 4 (l / 16) + r][n = l % 16]   (four FP32).
// Rows i are elements, k the inputs of a layer, n its outputs.  Between layers the D layout is
// turned into the A layout through 1 KB of LDS per wave.  The packed weights (MlpPacked) hold
// every B operand already in lane order, written by frforge_amd.learned.MLPWeights.pack.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace frf {

struct MlpPacked {
    uint16_t w[8][64][4];   // W1 n-blocks 0,1; W2 (k-block, n-block) 00,01,10,11; W3 k-blocks 0,1
    float b1[32], b2[32], b3[4];
};
static_assert(sizeof(MlpPacked) == 8 * 64 * 4 * 2 + 68 * 4, "packed MLP layout");

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short mlp_bf16(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ mlp_s4 mlp_wload(const MlpPacked* W, int m, int lane) {
    const uint2 v = *reinterpret_cast<const uint2*>(&W->w[m][lane][0]);
    mlp_s4 r;
    r[0] = (short)(v.x & 0xffffu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffffu);
    r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ void mlp_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hidden layer output (D layout) -> bias, ReLU, bf16 -> scratch rows [16][32].
__device__ __forceinline__ void mlp_store_hidden(short* sc, const mlp_f4& d, const float* bias, int nb,
                                                 int lane) {
    const int col = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float h = fmaxf(d[r] + bias[nb * 16 + col], 0.0f);
        sc[(4 * kq + r) * 32 + nb * 16 + col] = mlp_bf16(h);
    }
}

__device__ __forceinline__ mlp_s4 mlp_load_a(const short* sc, int kb, int lane) {
    const int row = lane & 15, kq = lane >> 4;
    return *reinterpret_cast<const mlp_s4*>(&sc[row * 32 + kb * 16 + 4 * kq]);
}

// Evaluates the network for the 16 elements of this wave.  mrow: the 16 modal coefficients of
// this lane's element (row lane % 16), k = a + 4 b.  sc: this wave's 16 x 32 bf16 scratch.
// Returns in every lane the raw outputs of rows 4 (lane / 16) + r, valid in lanes with
// lane % 16 == 0; not clamped, so a non-finite output reaches the caller's fallback.
__device__ __forceinline__ mlp_f4 mlp_wave16(const MlpPacked* __restrict__ W, const float* mrow,
                                             float eps_floor, short* sc) {
    const int lane = threadIdx.x & 63;
    const int kq = lane >> 4;
    float m[4];
    float e = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        m[r] = mrow[4 * kq + r];
        e += m[r] * m[r];
    }
    e += __shfl_xor(e, 16, 64);   // the four lanes of a row: l, l^16, l^32, l^48
    e += __shfl_xor(e, 32, 64);
    const float ie = 1.0f / (e + eps_floor);
    mlp_s4 a;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        // x = clip(log10(m^2 / E + 1e-10), -10, 0) / 10 + 1: the floor keeps the FP32 rounding
        // of modes that vanish in exact arithmetic (energy fraction ~1e-14) out of the features
        float x = __builtin_amdgcn_logf(m[r] * m[r] * ie + 1e-10f) * 0.30102999566398120f;
        x = fminf(fmaxf(x, -10.0f), 0.0f) * 0.1f + 1.0f;
        a[r] = mlp_bf16(x);
    }
    const mlp_f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    // layer 1: 16 inputs -> 32
    mlp_f4 d0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, mlp_wload(W, 0, lane), z, 0, 0, 0);
    mlp_f4 d1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, mlp_wload(W, 1, lane), z, 0, 0, 0);
    mlp_store_hidden(sc, d0, W->b1, 0, lane);
    mlp_store_hidden(sc, d1, W->b1, 1, lane);
    mlp_wave_sync();
    // layer 2: 32 -> 32
    mlp_s4 a0 = mlp_load_a(sc, 0, lane), a1 = mlp_load_a(sc, 1, lane);
    d0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, mlp_wload(W, 2, lane), z, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, mlp_wload(W, 4, lane), d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, mlp_wload(W, 3, lane), z, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, mlp_wload(W, 5, lane), d1, 0, 0, 0);
    mlp_wave_sync();   // every lane has read layer 1 before it is overwritten
    mlp_store_hidden(sc, d0, W->b2, 0, lane);
    mlp_store_hidden(sc, d1, W->b2, 1, lane);
    mlp_wave_sync();
    // layer 3: 32 -> 1 (output column 0 of a 16-wide block)
    a0 = mlp_load_a(sc, 0, lane);
    a1 = mlp_load_a(sc, 1, lane);
    mlp_f4 o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, mlp_wload(W, 6, lane), z, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, mlp_wload(W, 7, lane), o, 0, 0, 0);
    mlp_wave_sync();   // scratch free for the caller
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] += W->b3[0];
    return o;
}

// Clamp of a finite network output to [0, 1]; a non-finite output selects the fallback.
__device__ __forceinline__ float mlp_sigma_or(float out, float fallback) {
    return isfinite(out) ? fminf(fmaxf(out, 0.0f), 1.0f) : fallback;
}

}  // namespace frf
