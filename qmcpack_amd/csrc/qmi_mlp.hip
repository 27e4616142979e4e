// Learned per-electron proposal: the MLP 68 -> 128 -> 128 -> 128 -> 4 (SiLU), its forward pass on the
// BF16 matrix cores with FP32 accumulation, its backward pass and the Adam update.  There is no
// counterpart in the reference.
//
// Forward: one workgroup (4 waves) per 16 samples.  v_mfma_f32_16x16x32_bf16 operand layout
// (measured): lane l holds A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16], e < 8; result
// register v of lane l holds D[4 (l / 16) + v][l % 16].  A = layer input (samples x K, from LDS),
// B[k][n] = W[n][k] (8 consecutive BF16 of a weight row per lane), so each wave computes
// 16 samples x 32 neurons of a hidden layer (two 16-column tiles); the layer output goes back to
// LDS (BF16) as the next layer's input.
#include "qmi_mlp.hpp"

namespace qmi
{
// pack the FP32 master weights into the padded BF16 layouts read by the forward kernel
__global__ void k_mlp_pack(MlpParams P)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // W1 [128][96] (K padded), W2/W3 [128][128], W4 [16][128] (N padded)
  if (i < MLP_HID * MLP_K1)
  {
    const int n = i / MLP_K1, k = i % MLP_K1;
    P.w1b[i]    = k < MLP_IN ? bf16_bits(P.master[MLP_OFF_W1 + n * MLP_IN + k]) : 0;
  }
  if (i < MLP_HID * MLP_HID)
  {
    P.w2b[i] = bf16_bits(P.master[MLP_OFF_W2 + i]);
    P.w3b[i] = bf16_bits(P.master[MLP_OFF_W3 + i]);
  }
  if (i < 16 * MLP_HID)
  {
    const int n = i / MLP_HID;
    P.w4b[i]    = n < MLP_OUT ? bf16_bits(P.master[MLP_OFF_W4 + i]) : 0;
  }
}

// X [B][68] FP32 -> out [B][4]; pre-activations of layer l for sample b at
// z[(l * zstride + zoff + b) * 128] when z != nullptr (training).
__global__ __launch_bounds__(WG) void k_mlp_forward(MlpParams P, int B, const float* __restrict__ X,
                                                    float* __restrict__ out, float* __restrict__ z, int zstride,
                                                    int zoff)
{
  __shared__ __align__(16) uint16_t sA[16 * MLP_LDA];
  __shared__ __align__(16) uint16_t sB[16 * MLP_LDA];
  const int b0 = blockIdx.x * 16, t = threadIdx.x;
  const int nvalid = min(16, B - b0);
  for (int e = t; e < 16 * MLP_K1; e += WG)
  {
    const int r = e / MLP_K1, k = e % MLP_K1;
    const float x = (r < nvalid && k < MLP_IN) ? X[(size_t)(b0 + r) * MLP_IN + k] : 0.0f;
    sA[r * MLP_LDA + k] = bf16_bits(x);
  }
  __syncthreads();
  const float* bias = P.master;
  const size_t zb = (size_t)zoff + b0;
  mlp_layer<MLP_K1>(sA, P.w1b, bias + MLP_OFF_B1, sB, z ? z + zb * MLP_HID : nullptr, nvalid);
  __syncthreads();
  mlp_layer<MLP_HID>(sB, P.w2b, bias + MLP_OFF_B2, sA, z ? z + ((size_t)zstride + zb) * MLP_HID : nullptr, nvalid);
  __syncthreads();
  mlp_layer<MLP_HID>(sA, P.w3b, bias + MLP_OFF_B3, sB, z ? z + ((size_t)2 * zstride + zb) * MLP_HID : nullptr,
                     nvalid);
  __syncthreads();
  if ((t >> 6) == 0)
    mlp_out_layer(sB, P, out + (size_t)b0 * MLP_OUT, MLP_OUT, nvalid);
}

//------------------------------------------------------------------------------------------------
// Backward (FP32): activations h_l = SiLU(z_l) recomputed from the stored z.
// dZ3 = (dOut W4) .* SiLU'(z3); dZ2 = (dZ3 W3) .* SiLU'(z2); dZ1 = (dZ2 W2) .* SiLU'(z1);
// dW_l = dZ_l^T h_{l-1}, db_l = sum_b dZ_l.

// dH[b][k] = sum_n dZ[b][n] W[n][k]  then  dZnext[b][k] = dH[b][k] SiLU'(z[b][k]).  One thread per (b, k).
__global__ void k_mlp_back_act(int B, int nout, const float* __restrict__ dZ, const float* __restrict__ W,
                               const float* __restrict__ z, float* __restrict__ dZnext)
{
  const int k = threadIdx.x, b = blockIdx.x;
  if (b >= B || k >= MLP_HID)
    return;
  const float* dz = dZ + (size_t)b * nout;
  float s         = 0;
  for (int n = 0; n < nout; n++)
    s += dz[n] * W[(size_t)n * MLP_HID + k];
  dZnext[(size_t)b * MLP_HID + k] = s * dsilu(z[(size_t)b * MLP_HID + k]);
}

// dW[n][k] += sum_b dZ[b][n] H[b][k] (H = SiLU(z) when silu_in, else the input itself; row stride
// ldh), db[n] += sum_b dZ[b][n].  Tile 32 (n) x 32 (k) per workgroup of 256 threads (4 outputs
// per thread), the batch in chunks of 32 staged in LDS.
__global__ __launch_bounds__(256) void k_mlp_back_weight(int B, int nout, int nin, const float* __restrict__ dZ,
                                                         const float* __restrict__ H, int ldh, int silu_in,
                                                         float* __restrict__ dW, float* __restrict__ db)
{
  __shared__ float sZ[32][33];
  __shared__ float sH[32][33];
  const int n0 = blockIdx.y * 32, k0 = blockIdx.x * 32, t = threadIdx.x;
  const int tn = t / 8, tk = (t % 8) * 4;
  float acc[4] = {0, 0, 0, 0}, accb = 0;
  for (int b0 = 0; b0 < B; b0 += 32)
  {
    for (int e = t; e < 32 * 32; e += 256)
    {
      const int bb = e / 32, c = e % 32, b = b0 + bb;
      sZ[bb][c]    = (b < B && n0 + c < nout) ? dZ[(size_t)b * nout + n0 + c] : 0.0f;
      float h      = (b < B && k0 + c < nin) ? H[(size_t)b * ldh + k0 + c] : 0.0f;
      sH[bb][c]    = silu_in ? silu(h) : h;
    }
    __syncthreads();
    for (int bb = 0; bb < 32; bb++)
    {
      const float zn = sZ[bb][tn];
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[j] += zn * sH[bb][tk + j];
      if (tk == 0)
        accb += zn;
    }
    __syncthreads();
  }
  const int n = n0 + tn;
  if (n < nout)
  {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k0 + tk + j < nin)
        dW[(size_t)n * nin + k0 + tk + j] += acc[j];
    if (tk == 0 && blockIdx.x == 0)
      db[n] += accb;
  }
}

__global__ void k_fill(float* p, size_t n, float v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Adam (Kingma & Ba 2015) on the flat FP32 parameter vector; g is the gradient of the loss to
// minimise.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i]           = mi;
  v[i]           = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t launch_mlp_pack(const MlpParams& P, hipStream_t stream)
{
  k_mlp_pack<<<(MLP_HID * MLP_HID + 255) / 256, 256, 0, stream>>>(P);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpParams& P, int B, const float* X, float* out, float* z, int zstride, int zoff,
                              hipStream_t stream)
{
  if (B <= 0)
    return hipSuccess;
  k_mlp_forward<<<(B + 15) / 16, WG, 0, stream>>>(P, B, X, out, z, zstride, zoff);
  return hipGetLastError();
}

hipError_t launch_mlp_backward(const MlpParams& P, int B, const float* X, const float* z, const float* dOut,
                               float* grad, float* dz_a, float* dz_b, hipStream_t stream)
{
  if (B <= 0)
    return hipSuccess;
  const float* Wm = P.master;
  k_fill<<<(MLP_NPARAM + 255) / 256, 256, 0, stream>>>(grad, MLP_NPARAM, 0.0f);
  const float* z1 = z;
  const float* z2 = z + (size_t)B * MLP_HID;
  const float* z3 = z + (size_t)2 * B * MLP_HID;
  // output layer
  k_mlp_back_weight<<<dim3(MLP_HID / 32, 1), 256, 0, stream>>>(B, MLP_OUT, MLP_HID, dOut, z3, MLP_HID, 1,
                                                               grad + MLP_OFF_W4, grad + MLP_OFF_B4);
  k_mlp_back_act<<<B, MLP_HID, 0, stream>>>(B, MLP_OUT, dOut, Wm + MLP_OFF_W4, z3, dz_a);
  // layer 3
  k_mlp_back_weight<<<dim3(MLP_HID / 32, MLP_HID / 32), 256, 0, stream>>>(B, MLP_HID, MLP_HID, dz_a, z2, MLP_HID, 1,
                                                                          grad + MLP_OFF_W3, grad + MLP_OFF_B3);
  k_mlp_back_act<<<B, MLP_HID, 0, stream>>>(B, MLP_HID, dz_a, Wm + MLP_OFF_W3, z2, dz_b);
  // layer 2
  k_mlp_back_weight<<<dim3(MLP_HID / 32, MLP_HID / 32), 256, 0, stream>>>(B, MLP_HID, MLP_HID, dz_b, z1, MLP_HID, 1,
                                                                          grad + MLP_OFF_W2, grad + MLP_OFF_B2);
  k_mlp_back_act<<<B, MLP_HID, 0, stream>>>(B, MLP_HID, dz_b, Wm + MLP_OFF_W2, z1, dz_a);
  // layer 1 (input X)
  k_mlp_back_weight<<<dim3((MLP_IN + 31) / 32, MLP_HID / 32), 256, 0, stream>>>(B, MLP_HID, MLP_IN, dz_a, X, MLP_IN,
                                                                                0, grad + MLP_OFF_W1, grad + MLP_OFF_B1);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t stream)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  k_adam<<<(n + 255) / 256, 256, 0, stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

} // namespace qmi
