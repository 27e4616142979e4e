// Device pieces of the learned-proposal network shared by the batched forward kernel and the fused
// learned move: BF16 packing, SiLU, and one hidden layer on the BF16 matrix cores.
#pragma once
#include "qmi_kernels.hpp"

namespace qmi
{
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ inline uint16_t bf16_bits(float x)
{
  uint32_t b = __float_as_uint(x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z)
{
  const float sg = 1.0f / (1.0f + __expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// 8 BF16 values from 16 bytes at p (LDS or global)
__device__ inline bf8 load_bf8(const uint16_t* p)
{
  const uint4 u = *reinterpret_cast<const uint4*>(p);
  bf8 r;
  __builtin_memcpy(&r, &u, sizeof(r));
  return r;
}

constexpr int MLP_LDA = MLP_HID + 8; // LDS row stride (BF16 elements) of the activations

// One hidden layer for the 16 samples of this workgroup: in [16][MLP_LDA] (BF16, LDS) ->
// out [16][MLP_LDA]; wave w computes neurons [32 w, 32 w + 32).  z (pre-activations, FP32) is
// stored to zout[16][128] when non-null.
template<int K>
__device__ inline void mlp_layer(const uint16_t* in, const uint16_t* __restrict__ W, const float* __restrict__ b,
                                 uint16_t* out, float* __restrict__ zout, int nvalid)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, li = lane & 15;
  f4 acc[2] = {f4{0, 0, 0, 0}, f4{0, 0, 0, 0}};
#pragma unroll
  for (int ks = 0; ks < K / 32; ks++)
  {
    const bf8 a = load_bf8(in + li * MLP_LDA + ks * 32 + 8 * q);
#pragma unroll
    for (int t = 0; t < 2; t++)
    {
      const int n  = wave * 32 + t * 16 + li;
      const bf8 bw = load_bf8(W + (size_t)n * K + ks * 32 + 8 * q);
      acc[t]       = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 2; t++)
  {
    const int n    = wave * 32 + t * 16 + li;
    const float bn = b[n];
#pragma unroll
    for (int v = 0; v < 4; v++)
    {
      const int row = 4 * q + v;
      const float z = acc[t][v] + bn;
      out[row * MLP_LDA + n] = bf16_bits(silu(z));
      if (zout && row < nvalid)
        zout[row * MLP_HID + n] = z;
    }
  }
}

// Output layer for the 16 rows of `in` (wave 0 only): out[row][0..4) = W4 h + b4 for row < nvalid.
__device__ inline void mlp_out_layer(const uint16_t* in, const MlpParams& P, float* out, int ostride, int nvalid)
{
  const int lane = threadIdx.x & 63, q = lane >> 4, li = lane & 15;
  const float* bias = P.master;
  f4 acc = f4{0, 0, 0, 0};
#pragma unroll
  for (int ks = 0; ks < MLP_HID / 32; ks++)
  {
    const bf8 a  = load_bf8(in + li * MLP_LDA + ks * 32 + 8 * q);
    const bf8 bw = load_bf8(P.w4b + (size_t)li * MLP_HID + ks * 32 + 8 * q);
    acc          = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc, 0, 0, 0);
  }
  if (li < MLP_OUT)
#pragma unroll
    for (int v = 0; v < 4; v++)
    {
      const int row = 4 * q + v;
      if (row < nvalid)
        out[(size_t)row * ostride + li] = acc[v] + bias[MLP_OFF_B4 + li];
    }
}

} // namespace qmi
