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
template<typename DT>
__global__ __launch_bounds__(WG) void k_flush_T(SysC s, State st, int g, int d, double* __restrict__ T)
{
  __shared__ double As[FT_ROWS][FT_KC + 1];
  __shared__ double Us[FT_KC][MAX_K + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * FT_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, row = t >> 2, cq = t & 3;
  const DT* A = det_ptr<DT>(st.Ainv[g]) + (size_t)w * n * ld;
  const DT* U = det_ptr<DT>(st.U[g]) + (size_t)w * ld * k;
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
template<typename DT>
__global__ __launch_bounds__(WG) void k_flush_BV(SysC s, State st, int g, int d, DT* __restrict__ BVout)
{
  __shared__ double Bs[MAX_K * MAX_K];
  const int w = blockIdx.x;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const DT* B = det_ptr<DT>(st.Binv[g]) + (size_t)w * k * k;
  const DT* V = det_ptr<DT>(st.V[g]) + (size_t)w * ld * k;
  DT* U       = BVout + (size_t)w * ld * k;
  for (int e = threadIdx.x; e < d * d; e += WG)
    Bs[(e / d) * MAX_K + e % d] = B[(e / d) * k + e % d];
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += WG)
  {
    const DT* vo = V + (size_t)o * k;
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
          U[(size_t)o * k + c0 + cc] = (DT)acc[cc];
    }
  }
}

// Ainv[i][o] = -sum_c T[i][c] BV_T[o][c] + Ainv[i][o] for rows [i0, i0+64); 64x64 output tiles,
// 4x4 register block per thread.
template<typename DT>
__global__ __launch_bounds__(WG) void k_flush_update(SysC s, State st, int g, int d, const double* __restrict__ T)
{
  __shared__ double Ts[FT_ROWS][MAX_K + 1];
  __shared__ double Bv[MAX_K][64 + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * FT_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  DT* A        = det_ptr<DT>(st.Ainv[g]) + (size_t)w * n * ld;
  const DT* BV = det_ptr<DT>(st.U[g]) + (size_t)w * ld * k;
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
          A[(size_t)i * ld + o] = (DT)(-1.0 * acc[a][b] + (double)A[(size_t)i * ld + o]);
      }
    }
  }
}

// Fused flush for n <= 512 on the FP64 / FP32 matrix cores, using the association
//   Ainv -= ((Ainv U^T - W) Binv) V
// (the same product as (Ainv U^T - W)(Binv V), without materialising Binv V).
// Matrix-core operand layout (measured, see Mfma below): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; result register v of lane l holds row drow(l / 16, v) of D, column l % 16
// (for FP64 drow = l / 16 + 4 v, the case written out below).
// One wave owns 16 rows of Ainv and keeps a 256-column block of them in registers (for n <= 256 the
// whole row, so Ainv is read and written once):
//   lane (li = l % 16, q = l / 16) holds a[4 jt + v] = Ainv[row li][16 jt + 4 q + v], jt < 16.
//   1. T (16 x d) = Ainv_rows U^T: the K index of MFMA step kk is column 16 (kk / 4) + 4 q + kk % 4,
//      which is exactly a[kk]; U^T rows are staged in LDS 64 at a time.  T[i][c] -= 1 where
//      delay_list[c] == i (applyW).
//   2. T' = T Binv (Binv staged in LDS), through LDS into the B-operand layout.
//   3. D^T = V_T T'^T per 16-column tile jt, with the V row of A-operand row m taken as column
//      16 jt + 4 (m % 4) + m / 4, so that result register v of lane (li, q) is
//      Delta[row li][16 jt + 4 q + v]: the register a[4 jt + v] that it updates.
// LDS (doubles, LT = 16 ceil(k / 16) + 1): a 64-row chunk of U^T or V^T [64][LT], T per wave
// [4][16][LT], Binv zero-padded [16 ceil(k / 16)][LT].  Grid: 8 walkers per group of blocks, the
// row blocks of one walker on one XCD (block b runs on XCD b % 8) so that they share its L2.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int MF_COLS = 256; // n limit of the register-resident path

// 16x16x4 matrix-core step in the determinant storage type.  Both forms take A[l % 16][l / 16] and
// B[l / 16][l % 16]; their results differ (measured): result register v of lane l (q = l / 16) is
// row drow(q, v) = q + 4 v (FP64) or 4 q + v (FP32) of D, column l % 16.  acol(m) = 4 q + v for the
// (q, v) with drow(q, v) = m, the offset that maps A-operand row m of step 3 onto the register
// a[4 jt + v] of lane (li, q).
template<typename DT>
struct Mfma;
template<>
struct Mfma<double>
{
  using acc_t = dbl4;
  __device__ static acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int drow(int q, int v) { return q + 4 * v; }
  __device__ static int acol(int m) { return 4 * (m & 3) + (m >> 2); }
};
template<>
struct Mfma<float>
{
  using acc_t = flt4;
  __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int drow(int q, int v) { return 4 * q + v; }
  __device__ static int acol(int m) { return m; }
};

__host__ __device__ inline int flush_lt(int k) { return 16 * ((k + 15) / 16) + 1; }
__host__ __device__ inline size_t flush_lds_bytes(int k, size_t esize)
{
  const int lt = flush_lt(k);
  return (size_t)(64 * lt + 64 * lt + (lt - 1) * lt) * esize;
}

// Stage rows [j0, j0 + 64) x columns [0, 16 NCT) of an orbital-major delayed block X (row stride
// k) into sX[64][LT], zero outside j < n, c < d.  Thread (cl = t % 16, r0 = t / 16) loads
// columns ct * 16 + cl of rows r0 + 16 i; the loads are straight-line code from clamped in-range
// addresses, so that all of them are in flight together, and the padding is applied by a 0/1
// factor (the delayed blocks hold finite values only: zeroed at allocation, finite writes).
template<typename DT, int NCT>
__device__ inline void stage_rows(const DT* __restrict__ X, int k, int n, int d, int j0, DT* sX, int LT)
{
  const int t = threadIdx.x, cl = t & 15, r0 = t >> 4;
  DT tmp[NCT][4];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
      const int r = r0 + 16 * i, c = ct * 16 + cl, j = j0 + r;
      const DT x = X[(size_t)min(j, n - 1) * k + min(c, d - 1)];
      tmp[ct][i] = x * ((j < n && c < d) ? DT(1) : DT(0));
    }
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
#pragma unroll
    for (int i = 0; i < 4; i++)
      sX[(r0 + 16 * i) * LT + ct * 16 + cl] = tmp[ct][i];
}

// NCT = ceil(d / 16) column tiles of the delayed block; columns are zero-padded to 16 NCT, so the
// padded MFMA steps add exact zeros.  NCB = ceil(n / COLS) column blocks of COLS (256, or 128 for
// FP64 at n <= 256: 160 instead of 234 registers, occupancy 3, 0.358 instead of 0.379 ms at
// n = 256, k = 32, though Ainv is read 1.5 times; FP32 keeps 256 columns, 0.185 vs 0.207 ms): the
// register slab holds one block; step 1 accumulates T over the blocks in order, step 3 updates the
// last block (still in registers) first and reloads the earlier ones (Ainv read 2 - 1/NCB times,
// written once).
template<typename DT, int NCT, int NCB, int COLS>
__global__ __launch_bounds__(WG, 2) void k_flush_mfma(SysC s, State st, int g, int d, int nw)
{
  using MF    = Mfma<DT>;
  using acc_t = typename MF::acc_t;
  extern __shared__ double lds_raw[];
  DT* lds = reinterpret_cast<DT*>(lds_raw);
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int nblk = (n + 63) >> 6;
  const int b = blockIdx.x, xcd = b & 7, rest = b >> 3;
  const int w = (rest / nblk) * 8 + xcd, bx = rest % nblk;
  if (w >= nw)
    return;
  constexpr int DP = 16 * NCT; // padded delay columns
  const int LT     = flush_lt(k);
  DT* sU           = lds;           // [64][LT]
  DT* sT           = lds + 64 * LT; // [NWAVE][16][LT]
  DT* sB           = sT + 64 * LT;  // [DP][LT]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, q = lane >> 4, li = lane & 15;
  const int rbase = bx * 64 + wave * 16;
  const int row   = rbase + li;
  const bool rok  = row < n;
  DT* A          = det_ptr<DT>(st.Ainv[g]) + (size_t)w * n * ld;
  const DT* U    = det_ptr<DT>(st.U[g]) + (size_t)w * ld * k;
  const DT* V    = det_ptr<DT>(st.V[g]) + (size_t)w * ld * k;
  const DT* Binv = det_ptr<DT>(st.Binv[g]) + (size_t)w * k * k;
  const int* dl  = st.dlist[g] + (size_t)w * k;
  DT* sTw        = sT + wave * 16 * LT;
  const int ncb  = min(NCB, (n + COLS - 1) / COLS);

  constexpr int NJT = COLS / 16; // 16-column tiles of the register slab
  DT a[4 * NJT];
  auto load_block = [&](int cb) {
#pragma unroll
    for (int jt = 0; jt < NJT; jt++)
    {
      const int j0 = cb * COLS + jt * 16 + 4 * q;
      const DT* ap = A + (size_t)row * ld + j0;
      if (rok && j0 + 3 < n)
      {
#pragma unroll
        for (int v = 0; v < 4; v++)
          a[4 * jt + v] = ap[v];
      }
      else
      {
#pragma unroll
        for (int v = 0; v < 4; v++)
          a[4 * jt + v] = (rok && j0 + v < n) ? ap[v] : DT(0);
      }
    }
  };
  auto store_block = [&](int cb) {
#pragma unroll
    for (int jt = 0; jt < NJT; jt++)
    {
      const int j0 = cb * COLS + jt * 16 + 4 * q;
      DT* ap       = A + (size_t)row * ld + j0;
      if (rok && j0 + 3 < n)
      {
#pragma unroll
        for (int v = 0; v < 4; v++)
          ap[v] = a[4 * jt + v];
      }
      else
      {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (rok && j0 + v < n)
            ap[v] = a[4 * jt + v];
      }
    }
  };
  {
    constexpr int NB = (DP * DP + WG - 1) / WG;
    DT tb[NB];
#pragma unroll
    for (int i = 0; i < NB; i++)
    {
      const int e = t + WG * i, r = e / DP, c = e - r * DP;
      tb[i]       = Binv[min(r, d - 1) * k + min(c, d - 1)] * ((r < d && c < d) ? DT(1) : DT(0));
    }
#pragma unroll
    for (int i = 0; i < NB; i++)
    {
      const int e = t + WG * i, r = e / DP, c = e - r * DP;
      if (e < DP * DP)
        sB[r * LT + c] = tb[i];
    }
  }

  // 1. T = Ainv_rows U^T over the column blocks
  acc_t acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
    acc[ct] = acc_t{0, 0, 0, 0};
  for (int cb = 0; cb < ncb; cb++)
  {
    load_block(cb);
#pragma unroll
    for (int ch = 0; ch < COLS / 64; ch++)
    {
      const int j0 = cb * COLS + ch * 64;
      if (j0 < n)
      {
        __syncthreads();
        stage_rows<DT, NCT>(U, k, n, d, j0, sU, LT);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; i++)
        {
          const int jl = (i >> 2) * 16 + 4 * q + (i & 3);
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
            acc[ct] = MF::mma(a[ch * 16 + i], sU[jl * LT + ct * 16 + li], acc[ct]);
        }
      }
    }
  }
  // applyW; T rows q + 4 v of this wave, columns ct * 16 + li (zero beyond d)
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
  {
    const int c  = ct * 16 + li;
    const int dc = c < d ? dl[c] : -2;
#pragma unroll
    for (int v = 0; v < 4; v++)
      sTw[MF::drow(q, v) * LT + c] = c < d ? acc[ct][v] - ((dc == rbase + MF::drow(q, v)) ? DT(1) : DT(0)) : DT(0);
  }
  __syncthreads();

  // 2. T' = T Binv (K index e = 4 ee + q)
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
    acc[ct] = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int ee = 0; ee < DP / 4; ee++)
  {
    const DT av = sTw[li * LT + 4 * ee + q];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
      acc[ct] = MF::mma(av, sB[(4 * ee + q) * LT + ct * 16 + li], acc[ct]);
  }
  __syncthreads();
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
#pragma unroll
    for (int v = 0; v < 4; v++)
      sTw[MF::drow(q, v) * LT + ct * 16 + li] = acc[ct][v];

  // 3. Ainv_rows -= T' V, last column block first (its values are still in a[])
  for (int cb = ncb - 1; cb >= 0; cb--)
  {
    if (cb != ncb - 1)
      load_block(cb);
#pragma unroll
    for (int ch = 0; ch < COLS / 64; ch++)
    {
      const int j0 = cb * COLS + ch * 64;
      if (j0 < n)
      {
        __syncthreads();
        stage_rows<DT, NCT>(V, k, n, d, j0, sU, LT);
        __syncthreads();
#pragma unroll
        for (int j4 = 0; j4 < 4; j4++)
        {
          const int jt = ch * 4 + j4;
          const int jl = j4 * 16 + MF::acol(li);
          acc_t o      = acc_t{0, 0, 0, 0};
#pragma unroll
          for (int cc = 0; cc < DP / 4; cc++)
            o = MF::mma(sU[jl * LT + 4 * cc + q], sTw[li * LT + 4 * cc + q], o);
#pragma unroll
          for (int v = 0; v < 4; v++)
            a[4 * jt + v] = a[4 * jt + v] - o[v];
        }
      }
    }
    store_block(cb);
  }
}

hipError_t launch_flush(const SysC& s, const State& st, int g, int d, int nw, double* scratch, hipStream_t stream,
                        int variant)
{
  if (d <= 0)
    return hipSuccess;
  if (d > MAX_K)
    return hipErrorInvalidValue;
  const int n = s.n[g];
  dim3 grid((n + FT_ROWS - 1) / FT_ROWS, nw);
  return dispatch_prec(s.prec, [&](auto, auto dtv) {
    using DT = decltype(dtv);
    if (n <= 2 * MF_COLS && variant == 0)
    {
      const int nblk = (n + 63) / 64, ngrp = (nw + 7) / 8;
      const dim3 grid1(nblk * ngrp * 8);
      const size_t lds = flush_lds_bytes(s.k, sizeof(DT));
      auto go          = [&](auto cbt, auto colt) {
        constexpr int CB = decltype(cbt)::value, COLS = decltype(colt)::value;
        switch ((d + 15) / 16)
        {
        case 1: k_flush_mfma<DT, 1, CB, COLS><<<grid1, WG, lds, stream>>>(s, st, g, d, nw); break;
        case 2: k_flush_mfma<DT, 2, CB, COLS><<<grid1, WG, lds, stream>>>(s, st, g, d, nw); break;
        case 3: k_flush_mfma<DT, 3, CB, COLS><<<grid1, WG, lds, stream>>>(s, st, g, d, nw); break;
        default: k_flush_mfma<DT, 4, CB, COLS><<<grid1, WG, lds, stream>>>(s, st, g, d, nw); break;
        }
      };
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using C128 = std::integral_constant<int, MF_COLS / 2>;
      using C256 = std::integral_constant<int, MF_COLS>;
      if (std::is_same<DT, double>::value && n > MF_COLS / 2 && n <= MF_COLS)
        go(I2{}, C128{});
      else if (n <= MF_COLS)
        go(I1{}, C256{});
      else
        go(I2{}, C256{});
      return hipGetLastError();
    }
    // three passes; scratch holds T (n x k per walker)
    k_flush_T<DT><<<grid, WG, 0, stream>>>(s, st, g, d, scratch);
    k_flush_BV<DT><<<nw, WG, 0, stream>>>(s, st, g, d, det_ptr<DT>(st.U[g]));
    k_flush_update<DT><<<grid, WG, 0, stream>>>(s, st, g, d, scratch);
    return hipGetLastError();
  });
}

} // namespace qmi
