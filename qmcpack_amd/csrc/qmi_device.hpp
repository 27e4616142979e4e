// Device-side building blocks for the batched walker-move path on gfx950.
//
// One workgroup of 256 threads (4 wave64) owns one walker.  Orbital-indexed work is distributed
// as orbital o = t + 256 m (coalesced over the orbital-fastest spline table and the per-walker
// inverse rows); electron-indexed work as j = t + 256 m.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmi
{
// threadIdx.x passed through an empty asm, so that the compiler cannot treat values computed from it
// (LDS and global offsets, masks) as invariant across the moves of a looped kernel: hoisted ahead of
// k_sweep_segment's move loop they kept ~30 registers live through it, and the loop kernel spilled
// under its 128-register bound (552 bytes per lane; 20 with this and no machine LICM)
__device__ __forceinline__ int thread_index()
{
  int t = __builtin_amdgcn_workitem_id_x();
  asm volatile("" : "+v"(t));
  return t;
}

constexpr int WG         = 256;
constexpr int NWAVE      = WG / 64;
constexpr int MAX_JCOEF  = 68; // functor coefficients (NumParams + 4)
constexpr int MAX_K      = 64; // delay rank
constexpr int MAX_N      = 512; // electrons per spin (orbitals per determinant)
constexpr double DBL_EPS = 2.220446049250313e-16;
constexpr int TRACE_FIELDS = 8; // ratio, accepted, proposed position[3], gradient at the current position[3]

// Storage precision of a context (SysC::prec).  Arithmetic of the move phases, the ratio and
// gradient dot products, the Jastrows, positions, the LU re-inversion and log det stay FP64.
enum Precision
{
  PREC_FP64  = 0, // FP64 spline table and determinant engine
  PREC_FP32  = 1, // FP32 spline table, FP32 determinant engine (Ainv, dpsi, U, V, Binv; FP32 MFMA flush)
  PREC_BF16  = 2, // BF16 spline table, FP32 determinant engine
  PREC_T32   = 3, // FP32 spline table, FP64 determinant engine
  PREC_TBF16 = 4  // BF16 spline table, FP64 determinant engine
};

// BF16 storage (upper half of an IEEE float); read as double through the float it encodes.
struct bf16s
{
  uint16_t u;
  __device__ operator double() const { return (double)__uint_as_float((uint32_t)u << 16); }
};

template<typename T>
__device__ inline T from_double(double x);
template<>
__device__ inline double from_double<double>(double x)
{
  return x;
}
template<>
__device__ inline float from_double<float>(double x)
{
  return (float)x;
}
template<>
__device__ inline bf16s from_double<bf16s>(double x)
{
  // round to nearest even on the float encoding (finite inputs)
  uint32_t b = __float_as_uint((float)x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return bf16s{(uint16_t)(b >> 16)};
}

// Radial B-spline functor (QMCWaveFunctions/Jastrow/BsplineFunctor.h:117-133 coefficients with
// the cusp folded into coefs[0]).
struct FunctorC
{
  double c[MAX_JCOEF];
  double DeltaRInv;
  double rcut;
  int max_index;
  int np;
};

// System constants, passed by value to every kernel.
struct SysC
{
  int N, nup, ndn, nion;
  int first[2], n[2];
  int Npad; // padded electron count of per-walker SoA arrays
  int ld;   // leading dimension of determinant rows (>= max n, multiple of 8)
  int k;    // delay rank
  int has_det, has_j1, has_j2;
  int prec; // Precision
  int cplx; // complex (C2C) orbitals: complex FP64 determinant engine, table of 2 norb splines
  int kld;  // row stride of the k-point table (>= norb)
  double L[3], Linv[3];
  int periodic[3];
  double G[9], symGG[6];
  // spline table [Mx+3][My+3][Mz+3][npad]
  int M[3];
  int npad;
  long long xs, ys, zs;
  double dinv[3];
  FunctorC j1;
  FunctorC j2[2]; // [0] like spins, [1] unlike spins
};

// Per-walker device state (walker-major arrays).  The determinant-engine arrays (Ainv, dpsi, U, V,
// Binv) hold elements of the context's determinant type (double, or float for the mixed
// precisions); kernels view them through det_ptr<DT>.
struct State
{
  double* R;        // [nw][3][Npad]
  double* Ainv[2];  // [nw][n][ld]        rows = electrons, Ainv = (psiM^T)^-1
  double* dpsi[2];  // [nw][n][4][ld]     gx, gy, gz, lapl of orbitals at each electron
  double* U[2];     // [nw][ld][k]  delayed orbital rows, stored transposed (orbital-major)
  double* V[2];     // [nw][ld][k]  delayed stale inverse rows, stored transposed
  double* Binv[2];  // [nw][k][k]
  int* dlist[2];    // [nw][k]
  double* jas;      // [nw][10][Npad]  Vat, J1 gx gy gz, J1 lap, Uat, dUat x y z, d2Uat
  double* logdet;   // [nw][2][2]      (log|det|, phase) per determinant
  double* stats;    // [nw][2]         accepted, rejected
  // per-move scratch (API calls and the two-kernel crowd path)
  double* invRow;   // [nw][ld]
  double* phi;      // [nw][5][ld]
  double* mv;       // [nw][MV_SIZE]
  const double* ions; // [3][nion]
  const double* kpts; // [3][kld] Cartesian k of each orbital (complex orbitals)
};

template<typename DT>
__host__ __device__ inline DT* det_ptr(double* p)
{
  return reinterpret_cast<DT*>(p);
}
template<typename DT>
__host__ __device__ inline const DT* det_ptr(const double* p)
{
  return reinterpret_cast<const DT*>(p);
}

enum JasRow
{
  J_VAT = 0,
  J_G1X = 1,
  J_LAP1 = 4,
  J_UAT = 5,
  J_DUX = 6,
  J_D2U = 9,
  J_ROWS = 10
};

// move-state fields (per walker, st.mv)
enum MoveField
{
  MV_POS   = 0, // 3: proposed position
  MV_RATIO = 3, // total ratio
  MV_PROB  = 4, // |ratio|^2 * exp(log_gb - log_gf)
  MV_FLAG  = 5, // 1 if |ratio|^2 >= eps (a uniform is drawn)
  MV_C     = 8, // 13: J1 (curAt, curGrad[3], curLap), J2 (cur_Uat, grad[3]), det (ratio, grad numerators[3])
  MV_GNOW  = 21, // 3: gradient at the current position (drift input)
  MV_CDET  = 24, // 8: complex orbitals: det (ratio, grad numerators[3]) as (re, im)
  MV_SIZE  = 32
};

//------------------------------------------------------------------------------------------------
// Numerics/SplineBound.hpp:36-61
__device__ inline void spline_bound(double x, int nmax, int& ind, double& dx)
{
  if (x < 0)
  {
    ind = 0;
    dx  = 0.0;
  }
  else
  {
    double ipart;
    dx  = modf(x, &ipart);
    ind = static_cast<int>(ipart);
    if (ind > nmax)
    {
      ind = nmax;
      dx  = 1.0 - DBL_EPS;
    }
  }
}

// minimum-image displacement src - pos (Particle/Lattice/ParticleBConds3DSoa.h:110-139)
__device__ inline double min_image(const SysC& s, double sx, double sy, double sz, const double pos[3], double dr[3])
{
  const double src[3] = {sx, sy, sz};
#pragma unroll
  for (int d = 0; d < 3; d++)
  {
    if (s.periodic[d])
    {
      const double u = (src[d] - pos[d]) * s.Linv[d];
      dr[d]          = s.L[d] * (u - round(u));
    }
    else
      dr[d] = src[d] - pos[d];
  }
  return sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
}

// BsplineFunctor::evaluateVGL element (BsplineFunctor.h:771-838): u, du/dr / r, d2u/dr2, with the
// coefficient table at c (kernel argument or an LDS copy).
__device__ inline void functor_vgl_c(const double* c, double DeltaRInv, int max_index, double r, double& u,
                                     double& du, double& d2u)
{
  const double dSquareDeltaRinv = DeltaRInv * DeltaRInv;
  const double rinv             = 1.0 / r;
  r *= DeltaRInv;
  int i;
  double t;
  spline_bound(r, max_index, i, t);
  const double s0 = c[i], s1 = c[i + 1], s2 = c[i + 2], s3 = c[i + 3];
  d2u = dSquareDeltaRinv *
      (s0 * (-1.0 * t + 1.0) + s1 * (3.0 * t + -2.0) + s2 * (-3.0 * t + 1.0) + s3 * (1.0 * t + 0.0));
  du = DeltaRInv * rinv *
      (s0 * ((-0.5 * t + 1.0) * t + -0.5) + s1 * ((1.5 * t + -2.0) * t + 0.0) + s2 * ((-1.5 * t + 1.0) * t + 0.5) +
       s3 * ((0.5 * t + 0.0) * t + 0.0));
  u = (s0 * (((-1.0 / 6.0 * t + 3.0 / 6.0) * t + -3.0 / 6.0) * t + 1.0 / 6.0) +
       s1 * (((3.0 / 6.0 * t + -6.0 / 6.0) * t + 0.0 / 6.0) * t + 4.0 / 6.0) +
       s2 * (((-3.0 / 6.0 * t + 3.0 / 6.0) * t + 3.0 / 6.0) * t + 1.0 / 6.0) +
       s3 * (((1.0 / 6.0 * t + 0.0 / 6.0) * t + 0.0 / 6.0) * t + 0.0 / 6.0));
}

// value only of the same element (BsplineFunctor::evaluate, BsplineFunctor.h:117-133)
__device__ inline double functor_v_c(const double* c, double DeltaRInv, int max_index, double r)
{
  r *= DeltaRInv;
  int i;
  double t;
  spline_bound(r, max_index, i, t);
  const double s0 = c[i], s1 = c[i + 1], s2 = c[i + 2], s3 = c[i + 3];
  return s0 * (((-1.0 / 6.0 * t + 3.0 / 6.0) * t + -3.0 / 6.0) * t + 1.0 / 6.0) +
      s1 * (((3.0 / 6.0 * t + -6.0 / 6.0) * t + 0.0 / 6.0) * t + 4.0 / 6.0) +
      s2 * (((-3.0 / 6.0 * t + 3.0 / 6.0) * t + 3.0 / 6.0) * t + 1.0 / 6.0) +
      s3 * (((1.0 / 6.0 * t + 0.0 / 6.0) * t + 0.0 / 6.0) * t + 0.0 / 6.0);
}

__device__ inline void functor_vgl(const FunctorC& f, double r, double& u, double& du, double& d2u)
{
  functor_vgl_c(f.c, f.DeltaRInv, f.max_index, r, u, du, d2u);
}

// spline2/MultiBsplineData.hpp:50-76
__device__ inline void bspline_prefactors(double tx, double a[4], double da[4], double d2a[4])
{
  a[0]   = ((-1.0 / 6.0 * tx + 3.0 / 6.0) * tx + -3.0 / 6.0) * tx + 1.0 / 6.0;
  a[1]   = ((3.0 / 6.0 * tx + -6.0 / 6.0) * tx + 0.0 / 6.0) * tx + 4.0 / 6.0;
  a[2]   = ((-3.0 / 6.0 * tx + 3.0 / 6.0) * tx + 3.0 / 6.0) * tx + 1.0 / 6.0;
  a[3]   = ((1.0 / 6.0 * tx + 0.0 / 6.0) * tx + 0.0 / 6.0) * tx + 0.0 / 6.0;
  da[0]  = (-0.5 * tx + 1.0) * tx + -0.5;
  da[1]  = (1.5 * tx + -2.0) * tx + 0.0;
  da[2]  = (-1.5 * tx + 1.0) * tx + 0.5;
  da[3]  = (0.5 * tx + 0.0) * tx + 0.0;
  d2a[0] = -1.0 * tx + 1.0;
  d2a[1] = 3.0 * tx + -2.0;
  d2a[2] = -3.0 * tx + 1.0;
  d2a[3] = 1.0 * tx + 0.0;
}

// Location of a Cartesian point in the table: SplineR2R::convertPos (SplineR2R.h:156-170) and
// computeLocationAndFractional (spline2/MultiBsplineEval_helper.hpp:60-96).
struct SplineLoc
{
  long long off; // element offset of the (ix, iy, iz) corner
  double a[4], da[4], d2a[4], b[4], db[4], d2b[4], c[4], dc[4], d2c[4];
};

__device__ inline void spline_locate(const SysC& s, const double r[3], SplineLoc& L)
{
  double ru[3];
#pragma unroll
  for (int i = 0; i < 3; i++)
    ru[i] = r[0] * s.G[0 * 3 + i] + r[1] * s.G[1 * 3 + i] + r[2] * s.G[2 * 3 + i];
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (-DBL_EPS < ru[i] && ru[i] < 0)
      ru[i] = 0.0;
    else
      ru[i] -= floor(ru[i]);
  int ix, iy, iz;
  double tx, ty, tz;
  spline_bound(ru[0] * s.dinv[0], s.M[0] - 1, ix, tx);
  spline_bound(ru[1] * s.dinv[1], s.M[1] - 1, iy, ty);
  spline_bound(ru[2] * s.dinv[2], s.M[2] - 1, iz, tz);
  bspline_prefactors(tx, L.a, L.da, L.d2a);
  bspline_prefactors(ty, L.b, L.db, L.d2b);
  bspline_prefactors(tz, L.c, L.dc, L.d2c);
  L.off = ix * s.xs + iy * s.ys + iz * s.zs;
}

// 4x4x4 stencil for one orbital o, then assign_vgl (SplineR2R.cpp:338-374):
// out = psi, dpsi x/y/z, d2psi (bc sign +1 for a real gamma-point set).
// Operation order follows spline2/MultiBsplineVGLH.hpp:168-228.
template<typename ST>
__device__ inline void spline_vgl_orbital(const SysC& s, const ST* __restrict__ coefs, const SplineLoc& L, int o,
                                          double out[5])
{
  double v = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hxy = 0, hxz = 0, hyy = 0, hyz = 0, hzz = 0;
  const ST* base = coefs + L.off + o;
  const long long zs = s.zs;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
      const ST* p        = base + i * s.xs + j * s.ys;
      const double c0    = p[0];
      const double c1    = p[zs];
      const double c2    = p[2 * zs];
      const double c3    = p[3 * zs];
      const double pre20 = L.d2a[i] * L.b[j];
      const double pre10 = L.da[i] * L.b[j];
      const double pre00 = L.a[i] * L.b[j];
      const double pre11 = L.da[i] * L.db[j];
      const double pre01 = L.a[i] * L.db[j];
      const double pre02 = L.a[i] * L.d2b[j];
      const double sum0  = L.c[0] * c0 + L.c[1] * c1 + L.c[2] * c2 + L.c[3] * c3;
      const double sum1  = L.dc[0] * c0 + L.dc[1] * c1 + L.dc[2] * c2 + L.dc[3] * c3;
      const double sum2  = L.d2c[0] * c0 + L.d2c[1] * c1 + L.d2c[2] * c2 + L.d2c[3] * c3;
      hxx += pre20 * sum0;
      hxy += pre11 * sum0;
      hxz += pre10 * sum1;
      hyy += pre02 * sum0;
      hyz += pre01 * sum1;
      hzz += pre00 * sum2;
      gx += pre10 * sum0;
      gy += pre01 * sum0;
      gz += pre00 * sum1;
      v += pre00 * sum0;
    }
  const double dxInv = s.dinv[0], dyInv = s.dinv[1], dzInv = s.dinv[2];
  gx *= dxInv;
  gy *= dyInv;
  gz *= dzInv;
  hxx *= dxInv * dxInv;
  hyy *= dyInv * dyInv;
  hzz *= dzInv * dzInv;
  hxy *= dxInv * dyInv;
  hxz *= dxInv * dzInv;
  hyz *= dyInv * dzInv;
  const double* G = s.G;
  out[0]          = v;
  out[1]          = G[0] * gx + G[1] * gy + G[2] * gz;
  out[2]          = G[3] * gx + G[4] * gy + G[5] * gz;
  out[3]          = G[6] * gx + G[7] * gy + G[8] * gz;
  out[4] = hxx * s.symGG[0] + hxy * s.symGG[1] + hxz * s.symGG[2] + hyy * s.symGG[3] + hyz * s.symGG[4] +
      hzz * s.symGG[5];
}

// Value only of orbital o (SplineR2R::evaluateValue / assign_v, SplineR2R.cpp:162-202):
// sum_ij a_i b_j sum_k c_k coef.
template<typename ST>
__device__ inline double spline_v_orbital(const SysC& s, const ST* __restrict__ coefs, const SplineLoc& L, int o)
{
  double val     = 0;
  const ST* base = coefs + L.off + o;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
    {
      const ST* q     = base + i * s.xs + j * s.ys;
      const double s0 = L.c[0] * (double)q[0] + L.c[1] * (double)q[s.zs] + L.c[2] * (double)q[2 * s.zs] +
          L.c[3] * (double)q[3 * s.zs];
      val += (L.a[i] * L.b[j]) * s0;
    }
  return val;
}

// Uniform per-position spline factors staged in LDS: the 16 (i,j) products
// {d2a_i b_j, da_i b_j, a_i b_j, da_i db_j, a_i db_j, a_i d2b_j} and the z factors c, dc, d2c.
struct SplineLDS
{
  double pre[16][6];
  double cz[3][4];
  long long off;
};

// Fill SplineLDS from a located point (one thread; the caller synchronises before use).
__device__ inline void spline_stage(const SplineLoc& L, SplineLDS& S)
{
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
      double* pr = S.pre[i * 4 + j];
      pr[0]      = L.d2a[i] * L.b[j];
      pr[1]      = L.da[i] * L.b[j];
      pr[2]      = L.a[i] * L.b[j];
      pr[3]      = L.da[i] * L.db[j];
      pr[4]      = L.a[i] * L.db[j];
      pr[5]      = L.a[i] * L.d2b[j];
    }
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    S.cz[0][q] = L.c[q];
    S.cz[1][q] = L.dc[q];
    S.cz[2][q] = L.d2c[q];
  }
  S.off = L.off;
}

// Same stencil and assign_vgl as spline_vgl_orbital, with the uniform factors read from LDS.
template<typename ST>
__device__ inline void spline_vgl_orbital_lds(const SysC& s, const ST* __restrict__ coefs, const SplineLDS& S, int o,
                                              double out[5])
{
  double v = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hxy = 0, hxz = 0, hyy = 0, hyz = 0, hzz = 0;
  double c[4], dc[4], d2c[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    c[q]   = S.cz[0][q];
    dc[q]  = S.cz[1][q];
    d2c[q] = S.cz[2][q];
  }
  const ST* base = coefs + S.off + o;
  const long long zs = s.zs;
#pragma unroll 2
  for (int i = 0; i < 4; i++)
  {
    const ST* pi = base + i * s.xs;
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
      const ST* p        = pi + j * s.ys;
      const double c0    = p[0];
      const double c1    = p[zs];
      const double c2    = p[2 * zs];
      const double c3    = p[3 * zs];
      const double* pre  = S.pre[i * 4 + j];
      const double pre20 = pre[0], pre10 = pre[1], pre00 = pre[2], pre11 = pre[3], pre01 = pre[4], pre02 = pre[5];
      const double sum0  = c[0] * c0 + c[1] * c1 + c[2] * c2 + c[3] * c3;
      const double sum1  = dc[0] * c0 + dc[1] * c1 + dc[2] * c2 + dc[3] * c3;
      const double sum2  = d2c[0] * c0 + d2c[1] * c1 + d2c[2] * c2 + d2c[3] * c3;
      hxx += pre20 * sum0;
      hxy += pre11 * sum0;
      hxz += pre10 * sum1;
      hyy += pre02 * sum0;
      hyz += pre01 * sum1;
      hzz += pre00 * sum2;
      gx += pre10 * sum0;
      gy += pre01 * sum0;
      gz += pre00 * sum1;
      v += pre00 * sum0;
    }
  }
  const double dxInv = s.dinv[0], dyInv = s.dinv[1], dzInv = s.dinv[2];
  gx *= dxInv;
  gy *= dyInv;
  gz *= dzInv;
  hxx *= dxInv * dxInv;
  hyy *= dyInv * dyInv;
  hzz *= dzInv * dzInv;
  hxy *= dxInv * dyInv;
  hxz *= dxInv * dzInv;
  hyz *= dyInv * dzInv;
  const double* G = s.G;
  out[0]          = v;
  out[1]          = G[0] * gx + G[1] * gy + G[2] * gz;
  out[2]          = G[3] * gx + G[4] * gy + G[5] * gz;
  out[3]          = G[6] * gx + G[7] * gy + G[8] * gz;
  out[4] = hxx * s.symGG[0] + hxy * s.symGG[1] + hxz * s.symGG[2] + hyy * s.symGG[3] + hyz * s.symGG[4] +
      hzz * s.symGG[5];
}

//------------------------------------------------------------------------------------------------
// reductions over a 256-thread workgroup
__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// Sum K (<= 16) per-thread values over the workgroup through LDS (no cross-lane shuffles):
// every thread stores its K values, K*16 threads add 16 each, K threads add the 16 partials.
// scratch holds K*WG doubles; the result lands in out[0..K) (LDS) for all threads.
template<int K>
__device__ inline void block_sum(double (&v)[K], double* scratch, double* out)
{
  static_assert(K * 16 <= WG, "block_sum handles at most 16 values");
  const int t = thread_index();
#pragma unroll
  for (int k = 0; k < K; k++)
    scratch[k * WG + t] = v[k];
  __syncthreads();
  double s = 0;
  if (t < K * 16)
  {
    const double* src = scratch + (t >> 4) * WG + (t & 15) * 16;
#pragma unroll
    for (int i = 0; i < 16; i++)
      s += src[i];
  }
  __syncthreads();
  if (t < K * 16)
    scratch[t] = s;
  __syncthreads();
  if (t < K)
  {
    double r = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      r += scratch[t * 16 + i];
    out[t] = r;
  }
  __syncthreads();
}

// DriftModifierUNR::getDrift (GreenFunctionModifiers/DriftModifierUNR.cpp), a = 1
__device__ inline void unr_drift(double tau, const double g[3], double drift[3])
{
  const double vsq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  const double sc  = vsq < DBL_EPS ? tau : ((-1.0 + sqrt(1.0 + 2.0 * 1.0 * tau * vsq)) / (1.0 * vsq));
  drift[0]         = g[0] * sc;
  drift[1]         = g[1] * sc;
  drift[2]         = g[2] * sc;
}

} // namespace qmi
