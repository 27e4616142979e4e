// Delayed-update flush (DelayedUpdateBatched::mw_updateInvMat, DelayedUpdateBatched.h:675-738):
//   T    = Ainv U^T             (n x d)
//   T[delay_list[c]][c] -= 1    (applyW, detail/OMPTarget/AccelMatrixUpdateOMPTarget.hpp:167-189)
//   BV   = Binv V               (d x n; stored orbital-major as BV_T[o][c], written over U_T)
//   Ainv -= T BV                (n x n)
// The delayed rows are stored orbital-major (U_T[o][c], V_T[o][c]).  LDS-tiled FP64 register
// blocking; one workgroup per (64-row block, walker).
#include "qmi_kernels.hpp"

namespace qmi
{
constexpr int FT_ROWS = 64; // rows of Ainv per workgroup
constexpr int FT_KC   = 32; // contraction chunk

// T = Ainv U^T - W for rows [i0, i0+64).  Thread (row = t/4, cq = t%4) owns columns c = cq + 4 q.
__global__ __launch_bounds__(WG) void k_flush_T(SysC s, State st, int g, int d, double* __restrict__ T)
{
  __shared__ double As[FT_ROWS][FT_KC + 1];
  __shared__ double Us[FT_KC][MAX_K + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * FT_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, row = t >> 2, cq = t & 3;
  const double* A = st.Ainv[g] + (size_t)w * n * ld;
  const double* U = st.U[g] + (size_t)w * ld * k;
  double acc[MAX_K / 4];
#pragma unroll
  for (int q = 0; q < MAX_K / 4; q++)
    acc[q] = 0;
  for (int j0 = 0; j0 < n; j0 += FT_KC)
  {
    for (int e = t; e < FT_ROWS * FT_KC; e += WG)
    {
      const int rr = e / FT_KC, jj = e % FT_KC;
      const int i = i0 + rr, j = j0 + jj;
      As[rr][jj] = (i < n && j < n) ? A[(size_t)i * ld + j] : 0.0;
    }
    for (int e = t; e < FT_KC * d; e += WG)
    {
      const int jj = e / d, c = e % d;
      const int j  = j0 + jj;
      Us[jj][c]    = j < n ? U[(size_t)j * k + c] : 0.0;
    }
    __syncthreads();
    for (int jj = 0; jj < FT_KC; jj++)
    {
      const double a = As[row][jj];
#pragma unroll
      for (int q = 0; q < MAX_K / 4; q++)
      {
        const int c = cq + 4 * q;
        if (c < d)
          acc[q] += a * Us[jj][c];
      }
    }
    __syncthreads();
  }
  const int i = i0 + row;
  if (i < n)
  {
    const int* dl = st.dlist[g] + (size_t)w * k;
    double* Tw    = T + (size_t)w * n * k;
#pragma unroll
    for (int q = 0; q < MAX_K / 4; q++)
    {
      const int c = cq + 4 * q;
      if (c < d)
        Tw[(size_t)i * k + c] = acc[q] - ((dl[c] == i) ? 1.0 : 0.0);
    }
  }
}

// BV_T[o][c] = sum_e Binv[c][e] V_T[o][e] over the first d delayed rows, written into U_T.
__global__ __launch_bounds__(WG) void k_flush_BV(SysC s, State st, int g, int d)
{
  __shared__ double Bs[MAX_K * MAX_K];
  const int w = blockIdx.x;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const double* B = st.Binv[g] + (size_t)w * k * k;
  const double* V = st.V[g] + (size_t)w * ld * k;
  double* U       = st.U[g] + (size_t)w * ld * k;
  for (int e = threadIdx.x; e < d * d; e += WG)
    Bs[(e / d) * MAX_K + e % d] = B[(e / d) * k + e % d];
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += WG)
  {
    const double* vo = V + (size_t)o * k;
    for (int c0 = 0; c0 < d; c0 += 16)
    {
      double acc[16];
#pragma unroll
      for (int cc = 0; cc < 16; cc++)
        acc[cc] = 0;
      for (int e = 0; e < d; e++)
      {
        const double v = vo[e];
#pragma unroll
        for (int cc = 0; cc < 16; cc++)
          if (c0 + cc < d)
            acc[cc] += Bs[(c0 + cc) * MAX_K + e] * v;
      }
#pragma unroll
      for (int cc = 0; cc < 16; cc++)
        if (c0 + cc < d)
          U[(size_t)o * k + c0 + cc] = acc[cc];
    }
  }
}

// Ainv[i][o] = -sum_c T[i][c] BV_T[o][c] + Ainv[i][o] for rows [i0, i0+64); 64x64 output tiles,
// 4x4 register block per thread.
__global__ __launch_bounds__(WG) void k_flush_update(SysC s, State st, int g, int d, const double* __restrict__ T)
{
  __shared__ double Ts[FT_ROWS][MAX_K + 1];
  __shared__ double Bv[MAX_K][64 + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * FT_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  double* A        = st.Ainv[g] + (size_t)w * n * ld;
  const double* BV = st.U[g] + (size_t)w * ld * k;
  const double* Tw = T + (size_t)w * n * k;
  for (int e = t; e < FT_ROWS * d; e += WG)
  {
    const int rr = e / d, c = e % d;
    Ts[rr][c]    = (i0 + rr < n) ? Tw[(size_t)(i0 + rr) * k + c] : 0.0;
  }
  for (int o0 = 0; o0 < n; o0 += 64)
  {
    __syncthreads();
    for (int e = t; e < 64 * d; e += WG)
    {
      const int oo = e / d, c = e % d;
      Bv[c][oo]    = (o0 + oo < n) ? BV[(size_t)(o0 + oo) * k + c] : 0.0;
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        acc[a][b] = 0;
    for (int c = 0; c < d; c++)
    {
      double tv[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; a++)
        tv[a] = Ts[tr + 16 * a][c];
#pragma unroll
      for (int b = 0; b < 4; b++)
        bv[b] = Bv[c][tc + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a][b] += tv[a] * bv[b];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
    {
      const int i = i0 + tr + 16 * a;
      if (i >= n)
        continue;
#pragma unroll
      for (int b = 0; b < 4; b++)
      {
        const int o = o0 + tc + 16 * b;
        if (o < n)
          A[(size_t)i * ld + o] = -1.0 * acc[a][b] + A[(size_t)i * ld + o];
      }
    }
  }
}

hipError_t launch_flush(const SysC& s, const State& st, int g, int d, int nw, double* scratchT, hipStream_t stream)
{
  if (d <= 0)
    return hipSuccess;
  if (d > MAX_K)
    return hipErrorInvalidValue;
  const int n = s.n[g];
  dim3 grid((n + FT_ROWS - 1) / FT_ROWS, nw);
  k_flush_T<<<grid, WG, 0, stream>>>(s, st, g, d, scratchT);
  k_flush_BV<<<nw, WG, 0, stream>>>(s, st, g, d);
  k_flush_update<<<grid, WG, 0, stream>>>(s, st, g, d, scratchT);
  return hipGetLastError();
}

} // namespace qmi
