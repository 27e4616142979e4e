// Batched LU factorisation with partial pivoting, inverse and log determinant of Slater matrices
// (FP64 real and complex): the recompute of DiracDeterminantBatched (mw_recompute / mw_invertPsiM,
// DiracDeterminantBatched.cpp:117-146, 1121-1198) through DiracMatrix::invert_transpose /
// computeInvertAndLog / computeLogDet (DiracMatrix.h:101-106, 151-173, 189-220) and the batched
// inverter DiracMatrixInverter::mw_invert_transpose (DiracMatrixInverter.hpp:41-44), whose CUDA form
// calls cuBLAS getrf/getri and computeLogDet_kernel (detail/CUDA/cuBLAS_LU.cu:62-108).
//
// Semantics (LAPACK getrf on the matrix A(i, c), stored column-major as a[c * ld + i]):
//  * column j: the pivot is the first row i >= j with the largest |A(i, j)| (|re| + |im| for complex
//    values, as izamax); rows j and p are interchanged across the whole matrix; the entries below the
//    diagonal are multiplied by 1 / pivot; the trailing matrix receives A(i, c) -= A(i, j) A(j, c) for
//    every i, c > j with A(i, j) != 0;
//  * ipiv[j] = p + 1; info = first j + 1 with a zero pivot;
//  * log det = sum_i log(+-U_ii), the sign flipped where ipiv[i] != i + 1;
//  * inverse X = A^-1 column by column: b = P e_k, L y = b (forward, y_i accumulated over j = 0..i-1
//    in increasing order), U x = y (backward, x_i accumulated over j = n-1..i+1 in decreasing order,
//    then divided by U_ii); the output is X^T row-major (= X column-major), i.e. row k of the result is
//    column k of A^-1 (the inverse transposed that the determinant engine keeps, Ainv = (psiM^T)^-1).
//
// Every entry undergoes the same sequence of fused multiply-adds as in the unblocked algorithm above,
// only the order across independent entries changes: the panel, the rows of U right of it and the
// trailing update are blocked for the LDS, the triangular solves by row blocks of BR rows with one lane
// per right-hand side. The results are therefore bit-identical to a sequential restatement of the same
// arithmetic (complex products and quotients written out with the fused forms below).
#include "../../include/qmcmi.h"
#include "qmi_kernels.hpp"

#include <mutex>

namespace qmi
{
namespace
{
constexpr int LU_WG  = 256; // factorisation: one workgroup per matrix
constexpr int INV_WG = 64;  // inverse: one lane per column of the inverse
constexpr int INV_BR = 16;  // rows per block of the triangular solves (real; complex: 8)
constexpr int INV_JC = 64;  // factor columns staged in LDS at a time

// value arithmetic, real and complex (double2 = re, im); the complex forms fix the order of the fused
// operations so that a CPU restatement with std::fma gives the same bits
__device__ inline double lu_abs1(double v) { return fabs(v); }
__device__ inline double lu_abs1(double2 v) { return fabs(v.x) + fabs(v.y); }
__device__ inline bool lu_zero(double v) { return v == 0.0; }
__device__ inline bool lu_zero(double2 v) { return v.x == 0.0 && v.y == 0.0; }
template<typename V>
__device__ inline V lu_val(double re);
template<>
__device__ inline double lu_val<double>(double re) { return re; }
template<>
__device__ inline double2 lu_val<double2>(double re) { return make_double2(re, 0.0); }
// a - l u
__device__ inline double lu_msub(double a, double l, double u) { return __fma_rn(-l, u, a); }
__device__ inline double2 lu_msub(double2 a, double2 l, double2 u)
{
  return make_double2(__fma_rn(-l.x, u.x, __fma_rn(l.y, u.y, a.x)), __fma_rn(-l.x, u.y, __fma_rn(-l.y, u.x, a.y)));
}
// 1 / d
__device__ inline double lu_recip(double d) { return 1.0 / d; }
__device__ inline double2 lu_recip(double2 d)
{
  const double den = __fma_rn(d.x, d.x, d.y * d.y);
  return make_double2(d.x / den, -d.y / den);
}
// a r
__device__ inline double lu_mul(double a, double r) { return a * r; }
__device__ inline double2 lu_mul(double2 a, double2 r)
{
  return make_double2(__fma_rn(a.x, r.x, -(a.y * r.y)), __fma_rn(a.x, r.y, a.y * r.x));
}
// s / d
__device__ inline double lu_div(double s, double d) { return s / d; }
__device__ inline double2 lu_div(double2 s, double2 d)
{
  const double den = __fma_rn(d.x, d.x, d.y * d.y);
  return make_double2(__fma_rn(s.x, d.x, s.y * d.y) / den, __fma_rn(s.y, d.x, -(s.x * d.y)) / den);
}
// an empty asm statement that takes x as input and output: the value is complete at this point and
// the loads of the next row are not hoisted above it (bounds the registers of the unrolled solves)
__device__ inline void lu_pin(double& x) { asm volatile("" : "+v"(x)::"memory"); }
__device__ inline void lu_pin(double2& x) { asm volatile("" : "+v"(x.x), "+v"(x.y)::"memory"); }
// result in the determinant storage type
__device__ inline void lu_put(double* o, double v) { *o = v; }
__device__ inline void lu_put(float* o, double v) { *o = (float)v; }
__device__ inline void lu_put(double2* o, double2 v) { *o = v; }

// ------------------------------------------------------------------------------------------------
// Factorisation, one workgroup per matrix. Panels of NB columns are factored in LDS (P[q][i] =
// A(i, j0 + q) for rows i >= j0); then the panel's interchanges are applied to the other columns, the
// NB rows of U right of the panel are solved (one thread per column, kept in LDS as Ur[q][c]), and the
// trailing matrix is updated with one lane per row and a wave per group of four columns.
template<typename V, int NB>
__global__ __launch_bounds__(LU_WG) void k_lu_factor(int n, int ld, V* __restrict__ Aall, int* __restrict__ pivall,
                                                     int* __restrict__ permall, int* __restrict__ infoall)
{
  extern __shared__ __align__(16) unsigned char lu_lds[];
  V* P      = reinterpret_cast<V*>(lu_lds);             // [NB][n]
  V* Ur     = P + (size_t)NB * n;                       // [NB][n]
  int* perm = reinterpret_cast<int*>(Ur + (size_t)NB * n); // [n]: original row now at row r
  __shared__ double red_v[LU_WG / 64];
  __shared__ int red_i[LU_WG / 64];
  __shared__ int s_piv[NB];
  __shared__ int s_info;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const size_t w = blockIdx.x;
  V* A           = Aall + w * (size_t)n * ld;
  for (int i = t; i < n; i += LU_WG)
    perm[i] = i;
  if (t == 0)
    s_info = 0;
  __syncthreads();
  for (int j0 = 0; j0 < n; j0 += NB)
  {
    const int nb = min(NB, n - j0);
    const int h  = n - j0;
    for (int e = t; e < nb * h; e += LU_WG)
    {
      const int q = e / h, i = j0 + e % h;
      P[q * n + i] = A[(size_t)(j0 + q) * ld + i];
    }
    __syncthreads();
    // --- unblocked factorisation of the panel
    for (int jj = 0; jj < nb; jj++)
    {
      const int j  = j0 + jj;
      const V* col = P + jj * n;
      double bv    = -1.0;
      int bi       = n;
      for (int i = j + t; i < n; i += LU_WG)
      {
        const double v = lu_abs1(col[i]);
        if (v > bv)
        {
          bv = v;
          bi = i;
        }
      }
      // largest magnitude, ties to the smaller row: the first maximum of a sequential scan
      for (int off = 32; off > 0; off >>= 1)
      {
        const double ov = __shfl_xor(bv, off);
        const int oi    = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi))
        {
          bv = ov;
          bi = oi;
        }
      }
      if (lane == 0)
      {
        red_v[wv] = bv;
        red_i[wv] = bi;
      }
      __syncthreads();
      bv = red_v[0];
      bi = red_i[0];
      for (int q = 1; q < LU_WG / 64; q++)
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bi))
        {
          bv = red_v[q];
          bi = red_i[q];
        }
      const int p   = bi < n ? bi : j; // bi == n only when every candidate is NaN
      const V pv    = col[p];
      const bool nz = !lu_zero(pv);
      __syncthreads(); // every thread holds p and pv before the interchange
      if (nz && p != j)
      {
        for (int q = t; q < nb; q += LU_WG)
        {
          const V a    = P[q * n + p];
          P[q * n + p] = P[q * n + j];
          P[q * n + j] = a;
        }
        if (t == 0)
        {
          const int a = perm[p];
          perm[p]     = perm[j];
          perm[j]     = a;
        }
      }
      if (t == 0)
      {
        s_piv[jj]                 = nz ? p : j;
        pivall[w * (size_t)n + j] = p + 1;
        if (!nz && s_info == 0)
          s_info = j + 1;
      }
      __syncthreads();
      if (nz)
      {
        const V rp = lu_recip(pv);
        for (int i = j + 1 + t; i < n; i += LU_WG)
          P[jj * n + i] = lu_mul(P[jj * n + i], rp);
      }
      __syncthreads();
      const int rows = n - j - 1, cols = nb - jj - 1;
      for (int e = t; e < rows * cols; e += LU_WG)
      {
        const int q = jj + 1 + e / rows, i = j + 1 + e % rows;
        const V l   = P[jj * n + i];
        if (!lu_zero(l))
          P[q * n + i] = lu_msub(P[q * n + i], l, P[q * n + j]);
      }
      __syncthreads();
    }
    for (int e = t; e < nb * h; e += LU_WG)
    {
      const int q = e / h, i = j0 + e % h;
      A[(size_t)(j0 + q) * ld + i] = P[q * n + i];
    }
    // --- interchanges on the columns outside the panel; rows j0..j0+nb-1 of U right of it
    for (int c = t; c < n - nb; c += LU_WG)
    {
      const int cc = c < j0 ? c : c + nb;
      V* Ac        = A + (size_t)cc * ld;
      for (int q = 0; q < nb; q++)
      {
        const int p = s_piv[q];
        if (p != j0 + q)
        {
          const V a  = Ac[p];
          Ac[p]      = Ac[j0 + q];
          Ac[j0 + q] = a;
        }
      }
      if (cc >= j0 + nb)
      {
        V x[NB];
#pragma unroll
        for (int r = 0; r < NB; r++)
          x[r] = r < nb ? Ac[j0 + r] : lu_val<V>(0.0);
#pragma unroll
        for (int r = 1; r < NB; r++)
        {
#pragma unroll
          for (int q = 0; q < r; q++)
          {
            const V l = P[q * n + j0 + min(r, nb - 1)];
            if (r < nb && !lu_zero(l))
              x[r] = lu_msub(x[r], l, x[q]);
          }
        }
#pragma unroll
        for (int r = 0; r < NB; r++)
          if (r < nb)
          {
            Ac[j0 + r]     = x[r];
            Ur[r * n + cc] = x[r];
          }
      }
    }
    __syncthreads();
    // --- trailing update: A(i, c) -= sum_q L(i, j0 + q) U(j0 + q, c), q in increasing order
    const int r0 = j0 + nb, R = n - r0;
    if (R > 0)
    {
      const int nrc = (R + 63) / 64, ncg = (R + 3) / 4;
      int cur       = -1;
      V l[NB];
      for (int u = wv; u < nrc * ncg; u += LU_WG / 64)
      {
        const int rc = u / ncg, cg = u % ncg;
        const int i  = r0 + rc * 64 + lane;
        if (rc != cur)
        {
          cur = rc;
#pragma unroll
          for (int q = 0; q < NB; q++)
            l[q] = (i < n && q < nb) ? P[q * n + i] : lu_val<V>(0.0);
        }
        if (i < n)
        {
          V a[4];
          int cidx[4];
#pragma unroll
          for (int k = 0; k < 4; k++)
          {
            cidx[k] = min(r0 + cg * 4 + k, n - 1);
            a[k]    = A[(size_t)cidx[k] * ld + i];
          }
#pragma unroll
          for (int q = 0; q < NB; q++)
          {
            if (q < nb && !lu_zero(l[q]))
            {
#pragma unroll
              for (int k = 0; k < 4; k++)
                a[k] = lu_msub(a[k], l[q], Ur[q * n + cidx[k]]);
            }
          }
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (r0 + cg * 4 + k < n)
              A[(size_t)cidx[k] * ld + i] = a[k];
        }
      }
    }
    __syncthreads();
  }
  // perm[r] = original row now at row r; b = P e_k has its 1 at the row r with perm[r] == k
  for (int r = t; r < n; r += LU_WG)
    permall[w * (size_t)n + perm[r]] = r;
  if (t == 0)
    infoall[w] = s_info;
}

// ------------------------------------------------------------------------------------------------
// Inverse from the factors: one lane per column k of X = A^-1, blocks of BR rows; y and x are kept in
// the row-major work array W (W[i][k], coalesced across the lanes), the result is written as X^T
// row-major (out[k][i] = X(i, k)). Factor entries are staged in LDS INV_JC columns at a time.
template<typename V, typename O, int BR>
__global__ __launch_bounds__(INV_WG, 4) void k_lu_inverse(int n, int ld, const V* __restrict__ LUall,
                                                        const int* __restrict__ permall, V* __restrict__ Wall,
                                                        O* __restrict__ Outall)
{
  __shared__ V Ls[INV_JC][BR]; // Ls[jj][r] = factor entry (i0 + r, jc + jj)
  __shared__ V Ds[BR][BR];     // diagonal block (i0 + r, i0 + q)
  const int lane = threadIdx.x;
  const int k    = blockIdx.x * INV_WG + lane;
  const size_t w = blockIdx.y;
  const V* LU    = LUall + w * (size_t)n * ld;
  V* W           = Wall + w * (size_t)n * ld;
  O* Out         = Outall + w * (size_t)n * ld;
  const bool act = k < n;
  const int kc   = act ? k : 0;
  const int rk   = act ? permall[w * (size_t)n + k] : n;
  const V zero   = lu_val<V>(0.0);
  const int nblk = (n + BR - 1) / BR;

  // forward: L y = P e_k (unit lower triangle)
  for (int b = 0; b < nblk; b++)
  {
    const int i0 = b * BR, nr = min(BR, n - i0);
    V s[BR];
#pragma unroll
    for (int r = 0; r < BR; r++)
      s[r] = lu_val<V>(i0 + r == rk ? 1.0 : 0.0);
    for (int jc = 0; jc < i0; jc += INV_JC)
    {
      const int nj = min(INV_JC, i0 - jc);
      __syncthreads();
      for (int e = lane; e < nj * BR; e += INV_WG)
      {
        const int jj = e / BR, r = e % BR;
        Ls[jj][r]    = r < nr ? LU[(size_t)(jc + jj) * ld + i0 + r] : zero;
      }
      __syncthreads();
      int jj = 0;
#pragma unroll 1
      for (; jj + 2 <= nj; jj += 2)
      {
        V y[2];
#pragma unroll
        for (int u = 0; u < 2; u++)
          y[u] = W[(size_t)(jc + jj + u) * ld + kc];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int r = 0; r < BR; r++)
            s[r] = lu_msub(s[r], Ls[jj + u][r], y[u]);
      }
      for (; jj < nj; jj++)
      {
        const V y = W[(size_t)(jc + jj) * ld + kc];
#pragma unroll
        for (int r = 0; r < BR; r++)
          s[r] = lu_msub(s[r], Ls[jj][r], y);
      }
    }
    __syncthreads();
    for (int e = lane; e < BR * BR; e += INV_WG)
    {
      const int r = e / BR, q = e % BR;
      Ds[r][q]    = (r < nr && q < nr) ? LU[(size_t)(i0 + q) * ld + i0 + r] : zero;
    }
    __syncthreads();
#pragma unroll
    for (int r = 1; r < BR; r++)
    {
#pragma unroll
      for (int q = 0; q < r; q++)
        s[r] = lu_msub(s[r], Ds[r][q], s[q]);
      lu_pin(s[r]);
    }
    if (act)
    {
#pragma unroll
      for (int r = 0; r < BR; r++)
        if (r < nr)
          W[(size_t)(i0 + r) * ld + k] = s[r];
    }
  }

  // backward: U x = y
  for (int b = nblk - 1; b >= 0; b--)
  {
    const int i0 = b * BR, nr = min(BR, n - i0), i1 = i0 + nr;
    V s[BR];
#pragma unroll
    for (int r = 0; r < BR; r++)
      s[r] = r < nr ? W[(size_t)(i0 + r) * ld + kc] : zero;
    for (int jhi = n; jhi > i1; jhi -= INV_JC)
    {
      const int jlo = max(i1, jhi - INV_JC), nj = jhi - jlo;
      __syncthreads();
      for (int e = lane; e < nj * BR; e += INV_WG)
      {
        const int jj = e / BR, r = e % BR;
        Ls[jj][r]    = r < nr ? LU[(size_t)(jlo + jj) * ld + i0 + r] : zero;
      }
      __syncthreads();
      int jj = nj - 1;
#pragma unroll 1
      for (; jj >= 1; jj -= 2)
      {
        V x[2];
#pragma unroll
        for (int u = 0; u < 2; u++)
          x[u] = W[(size_t)(jlo + jj - u) * ld + kc];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int r = 0; r < BR; r++)
            s[r] = lu_msub(s[r], Ls[jj - u][r], x[u]);
      }
      for (; jj >= 0; jj--)
      {
        const V x = W[(size_t)(jlo + jj) * ld + kc];
#pragma unroll
        for (int r = 0; r < BR; r++)
          s[r] = lu_msub(s[r], Ls[jj][r], x);
      }
    }
    __syncthreads();
    for (int e = lane; e < BR * BR; e += INV_WG)
    {
      const int r = e / BR, q = e % BR;
      Ds[r][q]    = (r < nr && q < nr) ? LU[(size_t)(i0 + q) * ld + i0 + r] : lu_val<V>(1.0);
    }
    __syncthreads();
#pragma unroll
    for (int r = BR - 1; r >= 0; r--)
    {
#pragma unroll
      for (int q = BR - 1; q > r; q--)
        if (q < nr)
          s[r] = lu_msub(s[r], Ds[r][q], s[q]);
      s[r] = lu_div(s[r], Ds[r][r]);
      lu_pin(s[r]);
    }
    if (act)
    {
#pragma unroll
      for (int r = 0; r < BR; r++)
        if (r < nr)
          W[(size_t)(i0 + r) * ld + k] = s[r];
#pragma unroll
      for (int r = 0; r < BR; r++)
        if (r < nr)
          lu_put(Out + (size_t)k * ld + i0 + r, s[r]);
    }
  }
}

// log det from the diagonal and the pivots: log|v| and the phase (pi per negative real factor,
// arg(v) per complex factor) summed in row order
__global__ void k_lu_logdet_r(int n, int ld, int nmat, const double* __restrict__ LUall, const int* __restrict__ pivall,
                              double* __restrict__ out, int ostride)
{
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nmat)
    return;
  const double* A = LUall + (size_t)w * n * ld;
  const int* p    = pivall + (size_t)w * n;
  double re = 0, im = 0;
  for (int i = 0; i < n; i++)
  {
    const double dg = A[(size_t)i * ld + i];
    const double v  = (p[i] == i + 1) ? dg : -dg;
    re += log(fabs(v));
    if (v < 0)
      im += M_PI;
  }
  out[(size_t)w * ostride + 0] = re;
  out[(size_t)w * ostride + 1] = im;
}

__global__ void k_lu_logdet_c(int n, int ld, int nmat, const double2* __restrict__ LUall,
                              const int* __restrict__ pivall, double* __restrict__ out, int ostride)
{
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nmat)
    return;
  const double2* A = LUall + (size_t)w * n * ld;
  const int* p     = pivall + (size_t)w * n;
  double re = 0, im = 0;
  for (int i = 0; i < n; i++)
  {
    const double2 dg = A[(size_t)i * ld + i];
    const double2 v  = (p[i] == i + 1) ? dg : make_double2(-dg.x, -dg.y);
    re += 0.5 * log(v.x * v.x + v.y * v.y);
    im += atan2(v.y, v.x);
  }
  out[(size_t)w * ostride + 0] = re;
  out[(size_t)w * ostride + 1] = im;
}

template<typename V, int NB>
hipError_t launch_factor_nb(int n, int ld, int nmat, V* a, int* piv, int* perm, int* info, hipStream_t stream)
{
  const size_t lds = (size_t)2 * NB * n * sizeof(V) + (size_t)n * sizeof(int);
  if (lds > 160 * 1024)
    return hipErrorInvalidValue;
  if (lds > 64 * 1024)
  {
    const hipError_t e = hipFuncSetAttribute((const void*)k_lu_factor<V, NB>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
      return e;
  }
  k_lu_factor<V, NB><<<nmat, LU_WG, lds, stream>>>(n, ld, a, piv, perm, info);
  return hipGetLastError();
}

// the widest panel whose two LDS blocks fit in 80 KB (two workgroups per compute unit)
template<typename V>
hipError_t launch_factor_v(int n, int ld, int nmat, V* a, int* piv, int* perm, int* info, hipStream_t stream)
{
  const size_t per = (size_t)2 * n * sizeof(V);
  if (32 * per <= 80 * 1024)
    return launch_factor_nb<V, 32>(n, ld, nmat, a, piv, perm, info, stream);
  if (16 * per <= 80 * 1024)
    return launch_factor_nb<V, 16>(n, ld, nmat, a, piv, perm, info, stream);
  if (8 * per <= 80 * 1024)
    return launch_factor_nb<V, 8>(n, ld, nmat, a, piv, perm, info, stream);
  return launch_factor_nb<V, 4>(n, ld, nmat, a, piv, perm, info, stream);
}
// row-major a[m][i][j] (vw doubles per value) -> column-major t[m][j][i]: the storage of A(i, c) =
// a[i][c] that the factorisation works on.  Grid (ceil(n / 64), n, nmat), 64 threads: row i of the
// output column j.
__global__ __launch_bounds__(64) void k_lu_transpose_in(int n, int vw, const double* __restrict__ a,
                                                       double* __restrict__ t)
{
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y;
  const size_t per = (size_t)n * n * vw, m = blockIdx.z;
  if (i >= n)
    return;
  for (int q = 0; q < vw; q++)
    t[m * per + ((size_t)j * n + i) * vw + q] = a[m * per + ((size_t)i * n + j) * vw + q];
}

// The allocation-free inverter's device scratch: the LU factors, the solve work array, pivots and
// permutations of one call, carved by the host.  One copy per GPU, held until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns for that GPU.
static __device__ __align__(256) unsigned char scratch[QMI_NOALLOC_SCRATCH_BYTES];
} // namespace

hipError_t lu_noalloc_scratch(int device, void** out)
{
  static std::mutex mu;
  static void* addr[QMI_MAX_DEVICES] = {};
  if (device < 0 || device >= QMI_MAX_DEVICES || !out)
    return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device])
  {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
      return e;
    if (cur != device) // resolved with that device current: each GPU holds its own copy
      return hipErrorInvalidDevice;
    void* p = nullptr;
    e       = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return e;
    addr[device] = p;
  }
  *out = addr[device];
  return hipSuccess;
}

hipError_t launch_lu_transpose_in(int n, int nmat, int vw, const double* a, double* t, hipStream_t stream)
{
  if (n <= 0 || nmat <= 0 || (vw != 1 && vw != 2))
    return hipErrorInvalidValue;
  k_lu_transpose_in<<<dim3((n + 63) / 64, n, nmat), 64, 0, stream>>>(n, vw, a, t);
  return hipGetLastError();
}

hipError_t launch_lu_factor(bool complex_values, int n, int ld, int nmat, double* a, int* piv, int* perm, int* info,
                            hipStream_t stream)
{
  if (n <= 0 || nmat <= 0 || ld < n || n > MAX_N)
    return hipErrorInvalidValue;
  return complex_values ? launch_factor_v<double2>(n, ld, nmat, reinterpret_cast<double2*>(a), piv, perm, info, stream)
                        : launch_factor_v<double>(n, ld, nmat, a, piv, perm, info, stream);
}

hipError_t launch_lu_logdet(bool complex_values, int n, int ld, int nmat, const double* a, const int* piv, double* out,
                            int ostride, hipStream_t stream)
{
  if (n <= 0 || nmat <= 0)
    return hipErrorInvalidValue;
  if (complex_values)
    k_lu_logdet_c<<<(nmat + 63) / 64, 64, 0, stream>>>(n, ld, nmat, reinterpret_cast<const double2*>(a), piv, out,
                                                       ostride);
  else
    k_lu_logdet_r<<<(nmat + 63) / 64, 64, 0, stream>>>(n, ld, nmat, a, piv, out, ostride);
  return hipGetLastError();
}

hipError_t launch_lu_inverse(int out_kind, int n, int ld, int nmat, const double* lu, const int* perm, double* work,
                             void* out, hipStream_t stream)
{
  if (n <= 0 || nmat <= 0 || ld < n)
    return hipErrorInvalidValue;
  const dim3 grid((n + INV_WG - 1) / INV_WG, nmat);
  switch (out_kind)
  {
  case LU_OUT_F64:
    k_lu_inverse<double, double, INV_BR><<<grid, INV_WG, 0, stream>>>(n, ld, lu, perm, work, static_cast<double*>(out));
    break;
  case LU_OUT_F32:
    k_lu_inverse<double, float, INV_BR><<<grid, INV_WG, 0, stream>>>(n, ld, lu, perm, work, static_cast<float*>(out));
    break;
  case LU_OUT_C128:
    k_lu_inverse<double2, double2, INV_BR / 2><<<grid, INV_WG, 0, stream>>>(
        n, ld, reinterpret_cast<const double2*>(lu), perm, reinterpret_cast<double2*>(work),
        static_cast<double2*>(out));
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace qmi
