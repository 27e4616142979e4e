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

// BV_T[o][c] = sum_e Binv[c][e] V_T[o][e] over the first d delayed rows, written to BVout
// (orbital-major, row stride k, walker stride ld * k).
__global__ __launch_bounds__(WG) void k_flush_BV(SysC s, State st, int g, int d, double* __restrict__ BVout)
{
  __shared__ double Bs[MAX_K * MAX_K];
  const int w = blockIdx.x;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const double* B = st.Binv[g] + (size_t)w * k * k;
  const double* V = st.V[g] + (size_t)w * ld * k;
  double* U       = BVout + (size_t)w * ld * k;
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

// Fused flush for n <= 256 on the FP64 matrix cores.  v_mfma_f64_16x16x4f64 operand layout
// (measured): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; result register v of lane l
// holds D[l / 16 + 4 v][l % 16].  One wave owns 16 rows of Ainv and keeps them in registers for
// both products, so Ainv is read and written once:
//   lane (li = l % 16, q = l / 16) holds a[4 jt + v] = Ainv[row li][16 jt + 4 q + v], jt < 16.
//   1. T (16 x d) = Ainv_rows U^T: the K index of MFMA step kk is column 16 (kk / 4) + 4 q + kk % 4,
//      which is exactly a[kk]; then T[i][c] -= 1 where delay_list[c] == i (applyW).
//   2. T goes through LDS to the B-operand layout, and D^T = BV_T T^T is formed per 16-column
//      tile jt with the BV row of A-operand row m taken as column 16 jt + 4 (m % 4) + m / 4, so
//      that result register v of lane (li, q) is Delta[row li][16 jt + 4 q + v]: the register
//      a[4 jt + v] that it updates.
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int MF_COLS = 256; // n limit of the register-resident path

__global__ __launch_bounds__(WG, 2) void k_flush_mfma(SysC s, State st, int g, int d, const double* __restrict__ BV)
{
  __shared__ double sT[NWAVE][16][MAX_K + 1];
  const int w = blockIdx.y;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, q = lane >> 4, li = lane & 15;
  const int rbase = blockIdx.x * 64 + wave * 16;
  const int row   = rbase + li;
  const bool rok  = row < n;
  double* A        = st.Ainv[g] + (size_t)w * n * ld;
  const double* U  = st.U[g] + (size_t)w * ld * k;
  const double* Bw = BV + (size_t)w * ld * k;
  const int* dl    = st.dlist[g] + (size_t)w * k;
  const int njt    = (n + 15) >> 4;

  double a[64];
#pragma unroll
  for (int jt = 0; jt < 16; jt++)
  {
    const int j0     = jt * 16 + 4 * q;
    const double* ap = A + (size_t)row * ld + j0;
    if (rok && j0 + 3 < n)
    {
      const double2 x0 = reinterpret_cast<const double2*>(ap)[0];
      const double2 x1 = reinterpret_cast<const double2*>(ap)[1];
      a[4 * jt + 0]    = x0.x;
      a[4 * jt + 1]    = x0.y;
      a[4 * jt + 2]    = x1.x;
      a[4 * jt + 3]    = x1.y;
    }
    else
    {
#pragma unroll
      for (int v = 0; v < 4; v++)
        a[4 * jt + v] = (rok && j0 + v < n) ? ap[v] : 0.0;
    }
  }

  // 1. T = Ainv_rows U^T - W
  const int nct = (d + 15) >> 4;
  dbl4 acc[MAX_K / 16];
#pragma unroll
  for (int ct = 0; ct < MAX_K / 16; ct++)
    acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 64; kk++)
  {
    if ((kk >> 2) < njt)
    {
      const int j = (kk >> 2) * 16 + 4 * q + (kk & 3);
#pragma unroll
      for (int ct = 0; ct < MAX_K / 16; ct++)
        if (ct < nct)
        {
          const int c    = ct * 16 + li;
          const double b = (j < n && c < d) ? U[(size_t)j * k + c] : 0.0;
          acc[ct]        = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], b, acc[ct], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int ct = 0; ct < MAX_K / 16; ct++)
    if (ct < nct)
    {
      const int c = ct * 16 + li;
      if (c < d)
      {
        const int dc = dl[c];
#pragma unroll
        for (int v = 0; v < 4; v++)
          sT[wave][q + 4 * v][c] = acc[ct][v] - ((dc == rbase + q + 4 * v) ? 1.0 : 0.0);
      }
    }
  __syncthreads();

  // 2. Ainv_rows -= T BV
  const int ncc = (d + 3) >> 2;
#pragma unroll
  for (int jt = 0; jt < 16; jt++)
  {
    if (jt < njt)
    {
      const int jrow = jt * 16 + 4 * (li & 3) + (li >> 2);
      dbl4 o         = dbl4{0.0, 0.0, 0.0, 0.0};
      for (int cc = 0; cc < ncc; cc++)
      {
        const int c     = 4 * cc + q;
        const double av = (jrow < n && c < d) ? Bw[(size_t)jrow * k + c] : 0.0;
        const double bv = (c < d) ? sT[wave][li][c] : 0.0;
        o               = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, o, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++)
        a[4 * jt + v] = a[4 * jt + v] - o[v];
    }
  }

#pragma unroll
  for (int jt = 0; jt < 16; jt++)
  {
    const int j0 = jt * 16 + 4 * q;
    double* ap   = A + (size_t)row * ld + j0;
    if (rok && j0 + 3 < n)
    {
      reinterpret_cast<double2*>(ap)[0] = double2{a[4 * jt + 0], a[4 * jt + 1]};
      reinterpret_cast<double2*>(ap)[1] = double2{a[4 * jt + 2], a[4 * jt + 3]};
    }
    else
    {
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (rok && j0 + v < n)
          ap[v] = a[4 * jt + v];
    }
  }
}

hipError_t launch_flush(const SysC& s, const State& st, int g, int d, int nw, double* scratch, hipStream_t stream)
{
  if (d <= 0)
    return hipSuccess;
  if (d > MAX_K)
    return hipErrorInvalidValue;
  const int n = s.n[g];
  dim3 grid((n + FT_ROWS - 1) / FT_ROWS, nw);
  if (n <= MF_COLS)
  {
    // scratch holds BV_T (n x k per walker, walker stride ld * k)
    k_flush_BV<<<nw, WG, 0, stream>>>(s, st, g, d, scratch);
    k_flush_mfma<<<grid, WG, 0, stream>>>(s, st, g, d, scratch);
    return hipGetLastError();
  }
  // n > 256: three passes; scratch holds T (n x k per walker)
  k_flush_T<<<grid, WG, 0, stream>>>(s, st, g, d, scratch);
  k_flush_BV<<<nw, WG, 0, stream>>>(s, st, g, d, st.U[g]);
  k_flush_update<<<grid, WG, 0, stream>>>(s, st, g, d, scratch);
  return hipGetLastError();
}

} // namespace qmi
