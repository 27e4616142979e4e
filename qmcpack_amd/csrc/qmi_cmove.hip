// Complex (C2C) orbitals: the walker-move kernels, the determinant recompute pieces, the
// delayed-update flush and mw_evaluateGL with a complex FP64 determinant engine.
//
// Orbitals psi_j(r) = exp(-i k_j . r) u_j(r), u_j = a_j + i b_j from a table of 2 norb real splines
// with (a_j, b_j) adjacent (SplineC2C::assign_vgl, QMCWaveFunctions/BsplineFactory/SplineC2C.cpp:200-279).
// The determinant engine (Ainv, dpsi, U, V, Binv) holds complex values as (re, im) double pairs;
// the algebra is that of the real engine without conjugation (DelayedUpdateBatched.h, BLAS 'T'
// products).  A move's acceptance uses |ratio|^2 (VMCBatched.cpp:152) and its drift the real part
// of the gradient (DriftOperators.h:36-44); the log determinant accumulates log|ratio| and arg(ratio).
#include "qmi_move_common.hpp"

#include <algorithm>

namespace qmi
{

// ---- complex arithmetic on double2 (x: real part, y: imaginary part)
__host__ __device__ inline double2 zmk(double re, double im) { return make_double2(re, im); }
__device__ inline double2 zadd(double2 a, double2 b) { return zmk(a.x + b.x, a.y + b.y); }
__device__ inline double2 zsub(double2 a, double2 b) { return zmk(a.x - b.x, a.y - b.y); }
__device__ inline double2 zmul(double2 a, double2 b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zscale(double2 a, double s) { return zmk(a.x * s, a.y * s); }
// c + a b
__device__ inline double2 zfma(double2 a, double2 b, double2 c)
{
  return zmk(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}
__device__ inline double znorm(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double2 zinv(double2 a)
{
  const double m = 1.0 / znorm(a);
  return zmk(a.x * m, -a.y * m);
}
__device__ inline double2 zdiv(double2 a, double2 b) { return zmul(a, zinv(b)); }

// LDS of the complex move kernels (same roles as MoveLDS)
struct CMoveLDS
{
  double2 row[MAX_N]; // Ainv row / SM-1 temp vector
  double2 vec[MAX_N]; // proposed orbital values
  double2 p[MAX_K];
  double2 y[MAX_K]; // row d of Binv (from phase A), zero beyond d
  double2 Y[MAX_K]; // new column d of Binv
  double red[9 * WG];
  double bc[64];
  double jc[3][MAX_JCOEF];
  SplineLDS spl;
};

enum CBcField
{
  BCZ_GRAD = 48, // 6: determinant gradient at the current position (complex)
  BCZ_DET  = 54  // 8: determinant ratio and gradient numerators at the proposed position (complex)
};
static_assert(BCZ_DET + 8 <= 64, "complex broadcast slots");

// ---- C2C orbital j at the staged spline factors S: u from splines 2j, 2j+1 (one 16-byte load per
// grid point), then psi = e^{-i k.r}(u), grad psi, lap psi with the k terms of assign_vgl.
// pos: the particle position (the phase uses the unfolded Cartesian position).
__device__ inline void spline_vgl_orbital_c(const SysC& s, const double* __restrict__ coefs, const SplineLDS& S, int o,
                                            const double* __restrict__ kp, const double pos[3], double2 out[5])
{
  double v[2] = {0, 0}, gx[2] = {0, 0}, gy[2] = {0, 0}, gz[2] = {0, 0}, hxx[2] = {0, 0}, hxy[2] = {0, 0},
         hxz[2] = {0, 0}, hyy[2] = {0, 0}, hyz[2] = {0, 0}, hzz[2] = {0, 0};
  double c[4], dc[4], d2c[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    c[q]   = S.cz[0][q];
    dc[q]  = S.cz[1][q];
    d2c[q] = S.cz[2][q];
  }
  const double* base = coefs + S.off + 2 * o;
  const long long zs = s.zs;
#pragma unroll 1
  for (int i = 0; i < 4; i++)
  {
    const double* pi = base + i * s.xs;
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
      const double* p   = pi + j * s.ys;
      const double2 c0  = *reinterpret_cast<const double2*>(p);
      const double2 c1  = *reinterpret_cast<const double2*>(p + zs);
      const double2 c2  = *reinterpret_cast<const double2*>(p + 2 * zs);
      const double2 c3  = *reinterpret_cast<const double2*>(p + 3 * zs);
      const double* pre = S.pre[i * 4 + j];
      const double pre20 = pre[0], pre10 = pre[1], pre00 = pre[2], pre11 = pre[3], pre01 = pre[4], pre02 = pre[5];
      const double cr[4] = {c0.x, c1.x, c2.x, c3.x};
      const double ci[4] = {c0.y, c1.y, c2.y, c3.y};
#pragma unroll
      for (int h = 0; h < 2; h++)
      {
        const double* cc  = h == 0 ? cr : ci;
        const double sum0 = c[0] * cc[0] + c[1] * cc[1] + c[2] * cc[2] + c[3] * cc[3];
        const double sum1 = dc[0] * cc[0] + dc[1] * cc[1] + dc[2] * cc[2] + dc[3] * cc[3];
        const double sum2 = d2c[0] * cc[0] + d2c[1] * cc[1] + d2c[2] * cc[2] + d2c[3] * cc[3];
        hxx[h] += pre20 * sum0;
        hxy[h] += pre11 * sum0;
        hxz[h] += pre10 * sum1;
        hyy[h] += pre02 * sum0;
        hyz[h] += pre01 * sum1;
        hzz[h] += pre00 * sum2;
        gx[h] += pre10 * sum0;
        gy[h] += pre01 * sum0;
        gz[h] += pre00 * sum1;
        v[h] += pre00 * sum0;
      }
    }
  }
  const double dxInv = s.dinv[0], dyInv = s.dinv[1], dzInv = s.dinv[2];
  const double* G    = s.G;
  double d[2][3], lc[2];
#pragma unroll
  for (int h = 0; h < 2; h++)
  {
    const double ax = gx[h] * dxInv, ay = gy[h] * dyInv, az = gz[h] * dzInv;
    d[h][0] = G[0] * ax + G[1] * ay + G[2] * az;
    d[h][1] = G[3] * ax + G[4] * ay + G[5] * az;
    d[h][2] = G[6] * ax + G[7] * ay + G[8] * az;
    lc[h]   = hxx[h] * (dxInv * dxInv) * s.symGG[0] + hxy[h] * (dxInv * dyInv) * s.symGG[1] +
        hxz[h] * (dxInv * dzInv) * s.symGG[2] + hyy[h] * (dyInv * dyInv) * s.symGG[3] +
        hyz[h] * (dyInv * dzInv) * s.symGG[4] + hzz[h] * (dzInv * dzInv) * s.symGG[5];
  }
  const int kl      = s.kld;
  const double k[3] = {kp[o], kp[kl + o], kp[2 * kl + o]};
  const double mKK  = -(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  double sn, cs;
  sincos(-(pos[0] * k[0] + pos[1] * k[1] + pos[2] * k[2]), &sn, &cs);
  const double vr = v[0], vi = v[1];
  out[0] = zmk(cs * vr - sn * vi, cs * vi + sn * vr);
#pragma unroll
  for (int q = 0; q < 3; q++)
  {
    const double g_r = d[0][q] + vi * k[q];
    const double g_i = d[1][q] - vr * k[q];
    out[1 + q]       = zmk(cs * g_r - sn * g_i, cs * g_i + sn * g_r);
  }
  const double lap_r = lc[0] + mKK * vr + 2.0 * (k[0] * d[1][0] + k[1] * d[1][1] + k[2] * d[1][2]);
  const double lap_i = lc[1] + mKK * vi - 2.0 * (k[0] * d[0][0] + k[1] * d[0][1] + k[2] * d[0][2]);
  out[4]             = zmk(cs * lap_r - sn * lap_i, cs * lap_i + sn * lap_r);
}

// Value only (SplineC2C::evaluateValue / assign_v, SplineC2C.cpp:140-198)
__device__ inline double2 spline_v_orbital_c(const SysC& s, const double* __restrict__ coefs, const SplineLoc& L, int o,
                                             const double* __restrict__ kp, const double pos[3])
{
  double vr = 0, vi = 0;
  const double* base = coefs + L.off + 2 * o;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
    {
      const double* q = base + i * s.xs + j * s.ys;
      double sr = 0, si = 0;
#pragma unroll
      for (int z = 0; z < 4; z++)
      {
        const double2 cv = *reinterpret_cast<const double2*>(q + z * s.zs);
        sr += L.c[z] * cv.x;
        si += L.c[z] * cv.y;
      }
      const double ab = L.a[i] * L.b[j];
      vr += ab * sr;
      vi += ab * si;
    }
  const int kl = s.kld;
  double sn, cs;
  sincos(-(pos[0] * kp[o] + pos[1] * kp[kl + o] + pos[2] * kp[2 * kl + o]), &sn, &cs);
  return zmk(cs * vr - sn * vi, cs * vi + sn * vr);
}

// ---- reductions of complex values: K complex per thread as 2K doubles
template<int K>
__device__ inline void block_sum_z(const double2 (&v)[K], double* scratch, double* out)
{
  double a[2 * K];
#pragma unroll
  for (int i = 0; i < K; i++)
  {
    a[2 * i]     = v[i].x;
    a[2 * i + 1] = v[i].y;
  }
  block_sum<2 * K>(a, scratch, out);
}

__device__ inline double2 bcz(const double* bc, int slot) { return zmk(bc[slot], bc[slot + 1]); }

// out[c] = scale * sum_o X[o][c] vec[o], c < d (complex form of delayed_gemv).  Ends with a barrier.
__device__ inline void delayed_gemv_c(const double2* __restrict__ X, const double2* vec, int n, int k, int d,
                                      double* red, double2* out, double scale)
{
  const int t = threadIdx.x, G = WG / k, c = t % k, q = t / k;
  double2 a0 = zmk(0, 0);
  if (q < G && c < d)
  {
    const double2* xp = X + c;
#pragma unroll 4
    for (int o = q; o < n; o += G)
      a0 = zfma(xp[(size_t)o * k], vec[o], a0);
  }
  red[t]      = a0.x;
  red[WG + t] = a0.y;
  __syncthreads();
  if (t < d)
  {
    double sr = 0, si = 0;
    for (int j = 0; j < G; j++)
    {
      sr += red[j * k + t];
      si += red[WG + j * k + t];
    }
    out[t] = zmk(scale * sr, scale * si);
  }
  __syncthreads();
}

__device__ inline double2 row_dot_c(const double2* __restrict__ v, const double2* y, int d)
{
  double2 a = zmk(0, 0);
#pragma unroll 4
  for (int c = 0; c < d; c++)
    a = zfma(v[c], y[c], a);
  return a;
}

struct CWalkerView
{
  double* R;
  double2 *Ainv, *dpsi, *U, *V, *Binv;
  double *jas, *logdet, *stats;
  int* dl;
};

__device__ inline CWalkerView cwalker_view(const SysC& s, const State& st, int g, int w)
{
  const WalkerView<double2> v = walker_view<double2>(s, st, g, w);
  return CWalkerView{v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.jas, v.logdet, v.stats, v.dl};
}

//------------------------------------------------------------------------------------------------
// Phase A: inverse row through the delayed rows (mw_prepareInvRow, DelayedUpdateBatched.h:174-235)
template<int MAXM>
__device__ inline void phaseA_c(int n, int ld, int k, int d, int r, const double2* __restrict__ Ainv,
                                const double2* __restrict__ U, const double2* __restrict__ V, double2* __restrict__ Binv,
                                CMoveLDS& sh, double2* sB, double2 (&ainv_r)[MAXM], double2 (&inv)[MAXM])
{
  const int t         = threadIdx.x;
  const double2* arow = Ainv + (size_t)r * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
    ainv_r[m]   = o < n ? arow[o] : zmk(0, 0);
    inv[m]      = ainv_r[m];
  }
  if (k > 1 && d > 0)
  {
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + WG * m;
      if (o < n)
        sh.row[o] = ainv_r[m];
    }
    stage_binv(Binv, k, d, sB);
    __syncthreads();
    delayed_gemv_c(U, sh.row, n, k, d, sh.red, sh.p, 1.0);
    if (t < k)
    {
      double2 acc = zmk(0, 0);
      if (t < d)
        for (int j = 0; j < d; j++)
          acc = zfma(sB[j * (k + 1) + t], sh.p[j], acc);
      const double2 y = zmk(-acc.x, -acc.y);
      sh.y[t]         = t < d ? y : zmk(0, 0);
      if (t < d)
        Binv[d * k + t] = y;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + WG * m;
      if (o < n)
        inv[m] = zadd(row_dot_c(V + (size_t)o * k, sh.y, d), ainv_r[m]);
    }
  }
}

// determinant gradient invRow . dpsiM[r] -> sh.bc[BCZ_GRAD] (complex); real part in sh.bc[BC_GRAD]
template<int MAXM>
__device__ inline void det_grad_c(int n, int ld, int r, const double2* __restrict__ dpsi, const double2 (&inv)[MAXM],
                                  CMoveLDS& sh)
{
  const int t         = threadIdx.x;
  double2 gv[3]       = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
  const double2* drow = dpsi + (size_t)r * 4 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
    if (o < n)
    {
      gv[0] = zfma(inv[m], drow[o], gv[0]);
      gv[1] = zfma(inv[m], drow[ld + o], gv[1]);
      gv[2] = zfma(inv[m], drow[2 * ld + o], gv[2]);
    }
  }
  block_sum_z<3>(gv, sh.red, sh.bc + BCZ_GRAD);
  if (t < 3)
    sh.bc[BC_GRAD + t] = sh.bc[BCZ_GRAD + 2 * t];
  __syncthreads();
}

// Phase C: Jastrow sums, then the C2C orbitals at the proposed position and their dot products
// with the inverse row -> sh.bc[BCZ_DET] (ratio, gradient numerators).  All threads.
template<int MAXM>
__device__ inline void phaseC_c(const SysC& s, int g, int iat, int n, const double* __restrict__ R,
                                const double* __restrict__ ions, const double* __restrict__ coefs,
                                const double* __restrict__ kp, const double2 (&inv)[MAXM], double2 (&phv)[MAXM][5],
                                CMoveLDS& sh)
{
  const int t         = threadIdx.x;
  const double pos[3] = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
  if (s.has_det && t == 64)
  {
    SplineLoc L;
    spline_locate(s, pos, L);
    spline_stage(L, sh.spl);
  }
  jastrow_ratio_sums<2 * MAXM>(s, g, iat, R, ions, pos, sh);
  double2 acc[4] = {zmk(0, 0), zmk(0, 0), zmk(0, 0), zmk(0, 0)};
  if (s.has_det)
  {
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + WG * m;
      if (o < n)
      {
        spline_vgl_orbital_c(s, coefs, sh.spl, o, kp, pos, phv[m]);
#pragma unroll
        for (int q = 0; q < 4; q++)
          acc[q] = zfma(inv[m], phv[m][q], acc[q]);
      }
      else
      {
#pragma unroll
        for (int q = 0; q < 5; q++)
          phv[m][q] = zmk(0, 0);
      }
    }
  }
  block_sum_z<4>(acc, sh.red, sh.bc + BCZ_DET);
}

// Total complex ratio and gradient at the proposed position from the phase-C sums (one thread).
// vat, uat: the electron's current J1 and J2 values
__device__ inline void combine_ratio_cv(const SysC& s, double vat, double uat, const double* bc, double2& ratio,
                                        double2 gnew[3])
{
  const double* C = bc + BC_C;
  ratio           = zmk(1, 0);
  gnew[0] = gnew[1] = gnew[2] = zmk(0, 0);
  if (s.has_det)
  {
    const double2 rd = bcz(bc, BCZ_DET);
    ratio            = rd;
    const double2 ri = zinv(rd);
    for (int q = 0; q < 3; q++)
      gnew[q] = zmul(bcz(bc, BCZ_DET + 2 + 2 * q), ri);
  }
  if (s.has_j1)
  {
    for (int q = 0; q < 3; q++)
      gnew[q].x += C[1 + q];
    ratio = zscale(ratio, exp(vat - C[0]));
  }
  if (s.has_j2)
  {
    for (int q = 0; q < 3; q++)
      gnew[q].x += C[6 + q];
    ratio = zscale(ratio, exp(uat - C[5]));
  }
}

__device__ inline void combine_ratio_c(const SysC& s, int iat, const double* __restrict__ jas, const double* bc,
                                       double2& ratio, double2 gnew[3])
{
  const int Np = s.Npad;
  combine_ratio_cv(s, s.has_j1 ? jas[J_VAT * Np + iat] : 0.0, s.has_j2 ? jas[J_UAT * Np + iat] : 0.0, bc, ratio,
                   gnew);
}

// Phase D core (thread 0): returns prob_eff; flag = |ratio|^2 >= eps; |ratio| in ratio_abs.
__device__ inline double mh_prob_cv(const SysC& s, double vat, double uat, const MoveArgs& a, CMoveLDS& sh,
                                    double& ratio_abs, bool& flag, double& gn2)
{
  const double tauovermass = a.tau;
  const double oneover2tau = 0.5 / tauovermass;
  double2 ratio, gz[3];
  combine_ratio_cv(s, vat, uat, sh.bc, ratio, gz);
  gn2 = znorm(gz[0]) + znorm(gz[1]) + znorm(gz[2]);
  const double gnew[3] = {gz[0].x, gz[1].x, gz[2].x};
  double drev[3];
  unr_drift(tauovermass, gnew, drev);
  for (int q = 0; q < 3; q++)
    drev[q] += sh.bc[BC_MOVE + q];
  const double log_gb = -oneover2tau * (drev[0] * drev[0] + drev[1] * drev[1] + drev[2] * drev[2]);
  const double prob   = znorm(ratio);
  flag                = prob >= DBL_EPS;
  ratio_abs           = sqrt(prob);
  sh.bc[BC_RATIO]     = ratio_abs;
  return prob * exp(log_gb - sh.bc[BC_LOGGF]);
}

__device__ inline double mh_prob_c(const SysC& s, int iat, const MoveArgs& a, const double* __restrict__ jas,
                                   CMoveLDS& sh, double& ratio_abs, bool& flag, double& gn2)
{
  const double tauovermass = a.tau;
  const double oneover2tau = 0.5 / tauovermass;
  double2 ratio, gz[3];
  combine_ratio_c(s, iat, jas, sh.bc, ratio, gz);
  gn2 = znorm(gz[0]) + znorm(gz[1]) + znorm(gz[2]);
  const double gnew[3] = {gz[0].x, gz[1].x, gz[2].x};
  double drev[3];
  unr_drift(tauovermass, gnew, drev);
  for (int q = 0; q < 3; q++)
    drev[q] += sh.bc[BC_MOVE + q];
  const double log_gb = -oneover2tau * (drev[0] * drev[0] + drev[1] * drev[1] + drev[2] * drev[2]);
  const double prob   = znorm(ratio);
  flag                = prob >= DBL_EPS;
  ratio_abs           = sqrt(prob);
  sh.bc[BC_RATIO]     = ratio_abs;
  return prob * exp(log_gb - sh.bc[BC_LOGGF]);
}

// Phase E: accept (sh.bc[BC_ACC] != 0) or reject.  All threads.
template<int MAXM>
__device__ inline void phaseE_c(const SysC& s, int g, int iat, int n, int ld, int k, int d, int r, const CWalkerView& v,
                                const double2 (&ainv_r)[MAXM], const double2 (&phv)[MAXM][5], CMoveLDS& sh, double2* sB,
                                bool stage)
{
  const int t         = threadIdx.x;
  const bool accepted = sh.bc[BC_ACC] != 0.0;
  if (s.has_det)
  {
    const double2 ratio_det = bcz(sh.bc, BCZ_DET);
    double2* drow           = v.dpsi + (size_t)r * 4 * ld;
    if (k > 1)
    {
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + WG * m;
        if (o < n)
          v.V[(size_t)o * k + d] = ainv_r[m];
      }
      if (accepted)
      {
        const double2 sigma = zinv(ratio_det);
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            sh.vec[o] = phv[m][0];
        }
        if (stage)
          stage_binv(v.Binv, k, d, sB);
        __syncthreads();
        delayed_gemv_c(v.V, sh.vec, n, k, d, sh.red, sh.p, -1.0);
        if (t < d)
        {
          double2 acc = zmk(0, 0);
          for (int j = 0; j < d; j++)
            acc = zfma(sB[t * (k + 1) + j], sh.p[j], acc);
          const double2 Yv  = zmul(sigma, acc);
          v.Binv[t * k + d] = Yv;
          sh.Y[t]           = Yv;
        }
        __syncthreads();
        for (int idx = t; idx < d * d; idx += WG)
        {
          const int j = idx / d, i = idx - j * d;
          v.Binv[j * k + i] = zfma(sh.y[i], sh.Y[j], sB[j * (k + 1) + i]);
        }
        if (t < d)
          v.Binv[d * k + t] = zmul(sh.y[t], sigma);
        if (t == 0)
        {
          v.Binv[d * k + d] = sigma;
          v.dl[d]           = r;
        }
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
          {
            v.U[(size_t)o * k + d] = phv[m][0];
            drow[o]                = phv[m][1];
            drow[ld + o]           = phv[m][2];
            drow[2 * ld + o]       = phv[m][3];
            drow[3 * ld + o]       = phv[m][4];
          }
        }
      }
      else
      {
        if (t < d)
        {
          v.Binv[d * k + t] = zmk(0, 0);
          v.Binv[t * k + d] = zmk(0, 0);
        }
        if (t == 0)
        {
          v.Binv[d * k + d] = zmk(1, 0);
          v.dl[d]           = -1;
        }
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            v.U[(size_t)o * k + d] = zmk(0, 0);
        }
      }
    }
    else if (accepted)
    {
      // Sherman-Morrison (Fahy): temp = Ainv phi, temp[r] -= 1, Ainv[i] += (-temp[i] / ratio) Ainv_old[r]
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + WG * m;
        if (o < n)
          sh.vec[o] = phv[m][0];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int i = t + WG * m;
        if (i < n)
        {
          double2 acc       = zmk(0, 0);
          const double2* ai = v.Ainv + (size_t)i * ld;
          for (int o = 0; o < n; o++)
            acc = zfma(ai[o], sh.vec[o], acc);
          sh.row[i] = acc;
        }
      }
      __syncthreads();
      if (t == 0)
        sh.row[r].x -= 1.0;
      __syncthreads();
      const double2 cri = zscale(zinv(ratio_det), -1.0);
      for (int i = 0; i < n; i++)
      {
        const double2 ci = zmul(cri, sh.row[i]);
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            v.Ainv[(size_t)i * ld + o] = zfma(ci, ainv_r[m], v.Ainv[(size_t)i * ld + o]);
        }
      }
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + WG * m;
        if (o < n)
        {
          drow[o]          = phv[m][1];
          drow[ld + o]     = phv[m][2];
          drow[2 * ld + o] = phv[m][3];
          drow[3 * ld + o] = phv[m][4];
        }
      }
    }
    if (accepted && t == 0)
    {
      // convertValueToLog: log|ratio| + i arg(ratio)
      v.logdet[g * 2 + 0] += 0.5 * log(znorm(ratio_det));
      v.logdet[g * 2 + 1] += atan2(ratio_det.y, ratio_det.x);
    }
  }
  if (accepted)
    jastrow_accept<2 * MAXM>(s, g, iat, v.R, v.jas, sh);
  __syncthreads();
}

//------------------------------------------------------------------------------------------------
// Kernels.  Dynamic LDS: the d x d block of Binv (complex), binv_lds_bytes(k, 16).
template<int MAXM>
__global__ __launch_bounds__(WG) void k_move_one_c(SysC s, State st, const double* __restrict__ coefs, MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = a.w0 + blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const CWalkerView v = cwalker_view(s, st, g, w);
  int ucount          = a.ubase[(size_t)w * (N + 1) + iat];
  double2 ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    ainv_r[m] = inv[m] = zmk(0, 0);
  prefetch_move(s, iat, w, ucount, a, v.R, v.jas, sh);
  if (s.has_det)
  {
    phaseA_c<MAXM>(n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sBz, ainv_r, inv);
    det_grad_c<MAXM>(n, ld, r, v.dpsi, inv, sh);
  }
  else
    __syncthreads(); // publishes the prefetched scalars
  if (threadIdx.x == 0)
    propose_pf(s, a, sh);
  __syncthreads();
  phaseC_c<MAXM>(s, g, iat, n, v.R, st.ions, coefs, st.kpts, inv, phv, sh);
  if (threadIdx.x == 0)
  {
    double ratio_abs;
    bool flag;
    double gn2;
    const double pe = mh_prob_cv(s, sh.bc[BC_PF + PF_VAT], sh.bc[BC_PF + PF_UAT], a, sh, ratio_abs, flag, gn2);
    nan_guard(a, w, iat, ratio_abs * ratio_abs, gn2, gnow_norm2(sh.bc));
    if (a.dmc) // complex wave function: no node check (SFNBranch.h:161-169), prob_eff against eps
      flag = pe >= DBL_EPS;
    bool acc = false;
    if (flag)
    {
      const double u = sh.bc[BC_PF + PF_UNIF]; // a.unif[w N + ucount]
      ucount++;
      acc = u < pe;
    }
    sh.bc[BC_ACC] = acc ? 1.0 : 0.0;
    if (a.dmc)
      dmc_record(a, iat, w, acc);
    v.stats[acc ? 0 : 1] += 1;
    a.ubase[(size_t)w * (N + 1) + iat + 1] = ucount;
    if (a.trace)
    {
      double* tr = a.trace + ((size_t)iat * a.nw + w) * TRACE_FIELDS;
      tr[0]      = ratio_abs;
      tr[1]      = acc ? 1.0 : 0.0;
      for (int q = 0; q < 3; q++)
      {
        tr[2 + q] = sh.bc[BC_POS + q];
        tr[5 + q] = sh.bc[BC_GNOW + q];
      }
    }
  }
  __syncthreads();
  phaseE_c<MAXM>(s, g, iat, n, ld, k, d, r, v, ainv_r, phv, sh, sBz, false);
}

// Crowd path, first kernel: phases A-C and the MH probability; orbitals to st.phi (complex
// [nw][5][ld]), scalars to st.mv.
template<int MAXM>
__global__ __launch_bounds__(WG) void k_propose_c(SysC s, State st, const double* __restrict__ coefs, MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const CWalkerView v = cwalker_view(s, st, g, w);
  double2 ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    ainv_r[m] = inv[m] = zmk(0, 0);
  if (s.has_det)
  {
    phaseA_c<MAXM>(n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sBz, ainv_r, inv);
    det_grad_c<MAXM>(n, ld, r, v.dpsi, inv, sh);
  }
  if (threadIdx.x == 0)
    propose(s, iat, w, a, v.R, v.jas, sh);
  __syncthreads();
  phaseC_c<MAXM>(s, g, iat, n, v.R, st.ions, coefs, st.kpts, inv, phv, sh);
  double* mv = st.mv + (size_t)w * MV_SIZE;
  if (threadIdx.x == 0)
  {
    double ratio_abs;
    bool flag;
    double gn2;
    const double pe = mh_prob_c(s, iat, a, v.jas, sh, ratio_abs, flag, gn2);
    nan_guard(a, w, iat, ratio_abs * ratio_abs, gn2, gnow_norm2(sh.bc));
    mv[MV_RATIO]    = ratio_abs;
    mv[MV_PROB]     = pe;
    mv[MV_FLAG]     = flag ? 1.0 : 0.0;
    for (int q = 0; q < 3; q++)
    {
      mv[MV_POS + q]  = sh.bc[BC_POS + q];
      mv[MV_GNOW + q] = sh.bc[BC_GNOW + q];
    }
    for (int q = 0; q < 13; q++)
      mv[MV_C + q] = sh.bc[BC_C + q];
    for (int q = 0; q < 8; q++)
      mv[MV_CDET + q] = sh.bc[BCZ_DET + q];
  }
  double2* phi = reinterpret_cast<double2*>(st.phi) + (size_t)w * 5 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = threadIdx.x + WG * m;
    if (o < n)
      for (int q = 0; q < 5; q++)
        phi[q * ld + o] = phv[m][q];
  }
}

// restore the move scalars of st.mv into sh.bc and load the row/orbitals for phase E
template<int MAXM>
__device__ inline void load_accept_state(const SysC& s, const State& st, int w, int n, int ld, int k, int d, int r,
                                         const CWalkerView& v, CMoveLDS& sh, double2 (&ainv_r)[MAXM],
                                         double2 (&phv)[MAXM][5])
{
  const double* mv = st.mv + (size_t)w * MV_SIZE;
  if (threadIdx.x == 0)
  {
    for (int q = 0; q < 3; q++)
      sh.bc[BC_POS + q] = mv[MV_POS + q];
    for (int q = 0; q < 13; q++)
      sh.bc[BC_C + q] = mv[MV_C + q];
    for (int q = 0; q < 8; q++)
      sh.bc[BCZ_DET + q] = mv[MV_CDET + q];
  }
  if (k > 1 && threadIdx.x < d)
    sh.y[threadIdx.x] = v.Binv[d * k + threadIdx.x];
  const double2* phi = reinterpret_cast<const double2*>(st.phi) + (size_t)w * 5 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = threadIdx.x + WG * m;
    ainv_r[m]   = o < n ? v.Ainv[(size_t)r * ld + o] : zmk(0, 0);
    for (int q = 0; q < 5; q++)
      phv[m][q] = o < n ? phi[q * ld + o] : zmk(0, 0);
  }
  __syncthreads();
}

// Crowd path, second kernel: MH with the crowd's RNG stream, then phase E.
template<int MAXM>
__global__ __launch_bounds__(WG) void k_accept_crowd_c(SysC s, State st, MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const int crowd = w / a.wpc, w0 = crowd * a.wpc, w1 = min(a.nw, w0 + a.wpc);
  const CWalkerView v = cwalker_view(s, st, g, w);
  const double* mv    = st.mv + (size_t)w * MV_SIZE;
  double cnt[2]       = {0, 0};
  for (int j = w0 + threadIdx.x; j < w1; j += WG)
  {
    const double f = st.mv[(size_t)j * MV_SIZE + MV_FLAG];
    cnt[1] += f;
    if (j < w)
      cnt[0] += f;
  }
  block_sum<2>(cnt, sh.red, sh.bc + BC_OLD);
  if (threadIdx.x == 0)
  {
    const int base = a.ubase[(size_t)crowd * (N + 1) + iat];
    bool acc       = false;
    if (mv[MV_FLAG] != 0.0)
    {
      const double u = a.unif[(size_t)crowd * N * a.wpc + base + (int)sh.bc[BC_OLD]];
      acc            = u < mv[MV_PROB];
    }
    if (w == w1 - 1)
      a.ubase[(size_t)crowd * (N + 1) + iat + 1] = base + (int)sh.bc[BC_OLD + 1];
    sh.bc[BC_ACC] = acc ? 1.0 : 0.0;
    v.stats[acc ? 0 : 1] += 1;
    if (a.trace)
    {
      double* tr = a.trace + ((size_t)iat * a.nw + w) * TRACE_FIELDS;
      tr[0]      = mv[MV_RATIO];
      tr[1]      = acc ? 1.0 : 0.0;
      for (int q = 0; q < 3; q++)
      {
        tr[2 + q] = mv[MV_POS + q];
        tr[5 + q] = mv[MV_GNOW + q];
      }
    }
  }
  double2 ainv_r[MAXM], phv[MAXM][5];
  load_accept_state<MAXM>(s, st, w, n, ld, k, d, r, v, sh, ainv_r, phv);
  phaseE_c<MAXM>(s, g, iat, n, ld, k, d, r, v, ainv_r, phv, sh, sBz, true);
}

// API: mw_evalGrad -> out[nw][3] complex; inverse row to st.invRow (complex)
template<int MAXM>
__global__ __launch_bounds__(WG) void k_api_eval_grad_c(SysC s, State st, MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, Np = s.Npad;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const CWalkerView v = cwalker_view(s, st, g, w);
  double2 ainv_r[MAXM], inv[MAXM];
  if (s.has_det)
  {
    phaseA_c<MAXM>(n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sBz, ainv_r, inv);
    det_grad_c<MAXM>(n, ld, r, v.dpsi, inv, sh);
    double2* ir = reinterpret_cast<double2*>(st.invRow) + (size_t)w * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = threadIdx.x + WG * m;
      if (o < n)
        ir[o] = inv[m];
    }
  }
  if (threadIdx.x == 0)
  {
    for (int q = 0; q < 3; q++)
    {
      double2 gq = s.has_det ? bcz(sh.bc, BCZ_GRAD + 2 * q) : zmk(0, 0);
      if (s.has_j1)
        gq.x += v.jas[(J_G1X + q) * Np + iat];
      if (s.has_j2)
        gq.x += v.jas[(J_DUX + q) * Np + iat];
      a.out[((size_t)w * 3 + q) * 2]     = gq.x;
      a.out[((size_t)w * 3 + q) * 2 + 1] = gq.y;
    }
  }
}

// API: mw_calcRatioGrad at st.mv[MV_POS] -> out: ratios [nw] complex, then grads [nw][3] complex
template<int MAXM>
__global__ __launch_bounds__(WG) void k_api_ratio_grad_c(SysC s, State st, const double* __restrict__ coefs,
                                                         MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const CWalkerView v = cwalker_view(s, st, g, w);
  double* mv          = st.mv + (size_t)w * MV_SIZE;
  double2 ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    ainv_r[m] = inv[m] = zmk(0, 0);
  if (s.has_det)
  {
    double2* ir = reinterpret_cast<double2*>(st.invRow) + (size_t)w * ld;
    if (a.prepare)
    {
      phaseA_c<MAXM>(n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sBz, ainv_r, inv);
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = threadIdx.x + WG * m;
        if (o < n)
          ir[o] = inv[m];
      }
    }
    else
    {
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = threadIdx.x + WG * m;
        inv[m]      = o < n ? ir[o] : zmk(0, 0);
      }
    }
  }
  if (threadIdx.x < 3)
    sh.bc[BC_POS + threadIdx.x] = mv[MV_POS + threadIdx.x];
  __syncthreads();
  phaseC_c<MAXM>(s, g, iat, n, v.R, st.ions, coefs, st.kpts, inv, phv, sh);
  if (threadIdx.x == 0)
  {
    double2 ratio, gnew[3];
    combine_ratio_c(s, iat, v.jas, sh.bc, ratio, gnew);
    mv[MV_RATIO] = sqrt(znorm(ratio));
    for (int q = 0; q < 13; q++)
      mv[MV_C + q] = sh.bc[BC_C + q];
    for (int q = 0; q < 8; q++)
      mv[MV_CDET + q] = sh.bc[BCZ_DET + q];
    a.out[2 * w]     = ratio.x;
    a.out[2 * w + 1] = ratio.y;
    for (int q = 0; q < 3; q++)
    {
      a.out[2 * (a.nw + (size_t)w * 3 + q)]     = gnew[q].x;
      a.out[2 * (a.nw + (size_t)w * 3 + q) + 1] = gnew[q].y;
    }
  }
  double2* phi = reinterpret_cast<double2*>(st.phi) + (size_t)w * 5 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = threadIdx.x + WG * m;
    if (o < n)
      for (int q = 0; q < 5; q++)
        phi[q * ld + o] = phv[m][q];
  }
}

// API: mw_accept_rejectMove with host decisions a.host_acc[nw]
template<int MAXM>
__global__ __launch_bounds__(WG) void k_api_accept_c(SysC s, State st, MoveArgs a)
{
  __shared__ CMoveLDS sh;
  extern __shared__ double2 sBz[];
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const CWalkerView v = cwalker_view(s, st, g, w);
  if (threadIdx.x == 0)
  {
    const bool acc = a.host_acc[w] != 0;
    sh.bc[BC_ACC]  = acc ? 1.0 : 0.0;
    v.stats[acc ? 0 : 1] += 1;
  }
  double2 ainv_r[MAXM], phv[MAXM][5];
  load_accept_state<MAXM>(s, st, w, n, ld, k, d, r, v, sh, ainv_r, phv);
  phaseE_c<MAXM>(s, g, iat, n, ld, k, d, r, v, ainv_r, phv, sh, sBz, true);
}

hipError_t launch_move_c(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                         hipStream_t stream)
{
  const int maxn   = move_maxm(s, a.g) * WG;
  if (maxn == 0)
    return hipErrorInvalidValue;
  if ((a.w0 != 0 || a.wn != 0) && which != MK_MOVE_ONE)
    return hipErrorInvalidValue; // only the fused kernel takes a walker partition
  if (a.w0 < 0 || a.wn < 0 || a.w0 + a.wn > a.nw)
    return hipErrorInvalidValue;
  const dim3 grid(a.wn > 0 ? a.wn : a.nw), block(WG);
  const size_t dyn = binv_lds_bytes(s.k, sizeof(double2));
  const double* c  = static_cast<const double*>(coefs);
  auto run         = [&](auto mtag) -> hipError_t {
    constexpr int M = decltype(mtag)::value;
    switch (which)
    {
    case MK_MOVE_ONE: k_move_one_c<M><<<grid, block, dyn, stream>>>(s, st, c, a); break;
    case MK_PROPOSE: k_propose_c<M><<<grid, block, dyn, stream>>>(s, st, c, a); break;
    case MK_ACCEPT_CROWD: k_accept_crowd_c<M><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_GRAD: k_api_eval_grad_c<M><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_RATIO: k_api_ratio_grad_c<M><<<grid, block, dyn, stream>>>(s, st, c, a); break;
    case MK_API_ACCEPT: k_api_accept_c<M><<<grid, block, dyn, stream>>>(s, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  };
  if (maxn <= WG)
    return run(std::integral_constant<int, 1>{});
  if (maxn <= 2 * WG)
    return run(std::integral_constant<int, 2>{});
  return hipErrorInvalidValue;
}

//------------------------------------------------------------------------------------------------
// Recompute: C2C orbitals at every electron of determinant g, values transposed into the complex
// LU workspace lu ([nw][n][ld] complex), gradients/Laplacians into dpsi.
__global__ __launch_bounds__(WG) void k_spo_rows_c(SysC s, State st, const double* __restrict__ coefs, int g,
                                                   double2* __restrict__ lu)
{
  __shared__ SplineLDS spl;
  const int i = blockIdx.x, w = blockIdx.y;
  const int n = s.n[g], ld = s.ld, Np = s.Npad;
  const int iat       = s.first[g] + i;
  const double* R     = st.R + (size_t)w * 3 * Np;
  const double pos[3] = {R[iat], R[Np + iat], R[2 * Np + iat]};
  if (threadIdx.x == 0)
  {
    SplineLoc L;
    spline_locate(s, pos, L);
    spline_stage(L, spl);
  }
  __syncthreads();
  double2* A    = lu + (size_t)w * n * ld;
  double2* drow = det_ptr<double2>(st.dpsi[g]) + ((size_t)w * n + i) * 4 * ld;
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double2 out[5];
    spline_vgl_orbital_c(s, coefs, spl, o, st.kpts, pos, out);
    A[(size_t)o * ld + i] = out[0];
    drow[o]               = out[1];
    drow[ld + o]          = out[2];
    drow[2 * ld + o]      = out[3];
    drow[3 * ld + o]      = out[4];
  }
}

hipError_t launch_spo_rows_c(const SysC& s, const State& st, const void* coefs, int g, int nw, double* lu,
                             hipStream_t stream)
{
  k_spo_rows_c<<<dim3(s.n[g], nw), WG, 0, stream>>>(s, st, static_cast<const double*>(coefs), g,
                                                    reinterpret_cast<double2*>(lu));
  return hipGetLastError();
}

// mw_evaluateGL with complex determinants: G, L complex ([nw][N][3][2], [nw][N][2]);
// out3 = {ekin = -1/2 Re(sum G.G + sum L) (BareKineticEnergy.cpp:114), log|psi|, phase}.
// Lane 0 of the wave that reduces an electron's determinant row completes that electron.
__device__ inline void gl_finish_c(const SysC& s, const double* __restrict__ jas, int w, int iat, double2 g0, double2 g1,
                                   double2 g2, double2 l, double* __restrict__ Gout, double* __restrict__ Lout,
                                   double& e_gg, double& e_l)
{
  const int Np = s.Npad, N = s.N;
  if (s.has_j1)
  {
    g0.x += jas[(J_G1X + 0) * Np + iat];
    g1.x += jas[(J_G1X + 1) * Np + iat];
    g2.x += jas[(J_G1X + 2) * Np + iat];
    l.x -= jas[J_LAP1 * Np + iat];
  }
  if (s.has_j2)
  {
    g0.x += jas[(J_DUX + 0) * Np + iat];
    g1.x += jas[(J_DUX + 1) * Np + iat];
    g2.x += jas[(J_DUX + 2) * Np + iat];
    l.x += jas[J_D2U * Np + iat];
  }
  e_gg += zmul(g0, g0).x + zmul(g1, g1).x + zmul(g2, g2).x;
  e_l += l.x;
  if (Gout)
  {
    double* G = Gout + ((size_t)w * N + iat) * 6;
    G[0]      = g0.x;
    G[1]      = g0.y;
    G[2]      = g1.x;
    G[3]      = g1.y;
    G[4]      = g2.x;
    G[5]      = g2.y;
  }
  if (Lout)
  {
    Lout[((size_t)w * N + iat) * 2]     = l.x;
    Lout[((size_t)w * N + iat) * 2 + 1] = l.y;
  }
}

__global__ __launch_bounds__(WG) void k_eval_gl_c(SysC s, State st, double* __restrict__ out3, double* __restrict__ Gout,
                                                  double* __restrict__ Lout)
{
  __shared__ double red[4 * WG];
  __shared__ double bc[4];
  const int w = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int N = s.N, Np = s.Npad, ld = s.ld;
  const double* jas = st.jas + (size_t)w * J_ROWS * Np;
  double acc[4]     = {0, 0, 0, 0}; // Re(G.G), Re(sum L), sum Vat, sum Uat
  if (s.has_det)
  {
    for (int g = 0; g < 2; g++)
    {
      const int n      = s.n[g];
      const double2* A = det_ptr<double2>(st.Ainv[g]) + (size_t)w * n * ld;
      const double2* D = det_ptr<double2>(st.dpsi[g]) + (size_t)w * n * 4 * ld;
      for (int i = wave; i < n; i += NWAVE)
      {
        double2 dq[4] = {zmk(0, 0), zmk(0, 0), zmk(0, 0), zmk(0, 0)};
        const double2* a  = A + (size_t)i * ld;
        const double2* dd = D + (size_t)i * 4 * ld;
        for (int o = lane; o < n; o += 64)
        {
          const double2 av = a[o];
#pragma unroll
          for (int q = 0; q < 4; q++)
            dq[q] = zfma(av, dd[q * ld + o], dq[q]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          dq[q] = zmk(wave_sum(dq[q].x), wave_sum(dq[q].y));
        if (lane == 0)
        {
          // L += lap - dot(g, g) without conjugation (DiracDeterminantBatched.cpp:593-604)
          double2 gg = zmk(0, 0);
          for (int q = 0; q < 3; q++)
            gg = zfma(dq[q], dq[q], gg);
          gl_finish_c(s, jas, w, s.first[g] + i, dq[0], dq[1], dq[2], zsub(dq[3], gg), Gout, Lout, acc[0], acc[1]);
        }
      }
    }
  }
  else
    for (int i = t; i < N; i += WG)
      gl_finish_c(s, jas, w, i, zmk(0, 0), zmk(0, 0), zmk(0, 0), zmk(0, 0), Gout, Lout, acc[0], acc[1]);
  for (int i = t; i < N; i += WG)
  {
    if (s.has_j1)
      acc[2] += jas[J_VAT * Np + i];
    if (s.has_j2)
      acc[3] += jas[J_UAT * Np + i];
  }
  block_sum<4>(acc, red, bc);
  if (t == 0)
  {
    const double* ld4 = st.logdet + (size_t)w * 4;
    double logpsi = 0, phase = 0;
    if (s.has_det)
    {
      logpsi += ld4[0] + ld4[2];
      phase += ld4[1] + ld4[3];
    }
    if (s.has_j1)
      logpsi += -bc[2];
    if (s.has_j2)
      logpsi += -0.5 * bc[3];
    out3[(size_t)w * 3 + 0] = -0.5 * (bc[0] + bc[1]);
    out3[(size_t)w * 3 + 1] = logpsi;
    out3[(size_t)w * 3 + 2] = phase;
  }
}

hipError_t launch_eval_gl_c(const SysC& s, const State& st, int nw, double* out3, double* G, double* L,
                            hipStream_t stream)
{
  if (s.N > 1024)
    return hipErrorInvalidValue;
  k_eval_gl_c<<<nw, WG, 0, stream>>>(s, st, out3, G, L);
  return hipGetLastError();
}

//------------------------------------------------------------------------------------------------
// Complex delayed-update flush (DelayedUpdateBatched::mw_updateInvMat): T = Ainv U^T - W (n x d),
// BV_T[o][c] = sum_e Binv[c][e] V_T[o][e] (over U), Ainv -= T BV.  The two n x n x d products are
// register-blocked: a workgroup owns a 64 x 64 output tile, thread (tr = t / 16, tc = t % 16) the
// 4 x 4 block of rows 4 tr + a and columns tc + 16 b, with 32-wide K chunks of both operands in LDS.
constexpr int CF_ROWS = 64;
constexpr int CF_KC   = 32;

// T = Ainv U^T - W for rows [i0, i0 + 64), columns c < d (<= 64)
__global__ __launch_bounds__(WG) void k_flush_T_c(SysC s, State st, int g, int d, double2* __restrict__ T)
{
  __shared__ double2 As[CF_ROWS][CF_KC + 1];
  __shared__ double2 Us[CF_KC][MAX_K + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * CF_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  const double2* A = det_ptr<double2>(st.Ainv[g]) + (size_t)w * n * ld;
  const double2* U = det_ptr<double2>(st.U[g]) + (size_t)w * ld * k;
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
      acc[a][b] = zmk(0, 0);
  for (int j0 = 0; j0 < n; j0 += CF_KC)
  {
    __syncthreads();
    for (int e = t; e < CF_ROWS * CF_KC; e += WG)
    {
      const int rr = e / CF_KC, jj = e % CF_KC;
      const int i = i0 + rr, j = j0 + jj;
      As[rr][jj]  = (i < n && j < n) ? A[(size_t)i * ld + j] : zmk(0, 0);
    }
    for (int e = t; e < CF_KC * MAX_K; e += WG)
    {
      const int jj = e / MAX_K, c = e % MAX_K;
      const int j  = j0 + jj;
      Us[jj][c]    = (j < n && c < d) ? U[(size_t)j * k + c] : zmk(0, 0);
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < CF_KC; jj++)
    {
      double2 av[4], uv[4];
#pragma unroll
      for (int a = 0; a < 4; a++)
        av[a] = As[4 * tr + a][jj];
#pragma unroll
      for (int b = 0; b < 4; b++)
        uv[b] = Us[jj][tc + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a][b] = zfma(av[a], uv[b], acc[a][b]);
    }
  }
  const int* dl = st.dlist[g] + (size_t)w * k;
  double2* Tw   = T + (size_t)w * n * k;
#pragma unroll
  for (int a = 0; a < 4; a++)
  {
    const int i = i0 + 4 * tr + a;
    if (i >= n)
      continue;
#pragma unroll
    for (int b = 0; b < 4; b++)
    {
      const int c = tc + 16 * b;
      if (c < d)
        Tw[(size_t)i * k + c] = zmk(acc[a][b].x - ((dl[c] == i) ? 1.0 : 0.0), acc[a][b].y);
    }
  }
}

__global__ __launch_bounds__(WG) void k_flush_BV_c(SysC s, State st, int g, int d)
{
  __shared__ double2 Bs[MAX_K * MAX_K];
  const int w = blockIdx.x;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const double2* B = det_ptr<double2>(st.Binv[g]) + (size_t)w * k * k;
  const double2* V = det_ptr<double2>(st.V[g]) + (size_t)w * ld * k;
  double2* U       = det_ptr<double2>(st.U[g]) + (size_t)w * ld * k;
  for (int e = threadIdx.x; e < d * d; e += WG)
    Bs[(e / d) * MAX_K + e % d] = B[(e / d) * k + e % d];
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += WG)
  {
    const double2* vo = V + (size_t)o * k;
    for (int c0 = 0; c0 < d; c0 += 8)
    {
      double2 acc[8];
#pragma unroll
      for (int cc = 0; cc < 8; cc++)
        acc[cc] = zmk(0, 0);
      for (int e = 0; e < d; e++)
      {
        const double2 v = vo[e];
#pragma unroll
        for (int cc = 0; cc < 8; cc++)
          if (c0 + cc < d)
            acc[cc] = zfma(Bs[(c0 + cc) * MAX_K + e], v, acc[cc]);
      }
#pragma unroll
      for (int cc = 0; cc < 8; cc++)
        if (c0 + cc < d)
          U[(size_t)o * k + c0 + cc] = acc[cc];
    }
  }
}

// Ainv[i][o] -= sum_c T[i][c] BV_T[o][c] for rows [i0, i0 + 64), every 64-column tile of the row band
__global__ __launch_bounds__(WG) void k_flush_update_c(SysC s, State st, int g, int d, const double2* __restrict__ T)
{
  __shared__ double2 Ts[CF_ROWS][CF_KC + 1];
  __shared__ double2 Bv[CF_KC][64 + 1];
  const int w = blockIdx.y, i0 = blockIdx.x * CF_ROWS;
  const int n = s.n[g], ld = s.ld, k = s.k;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  double2* A        = det_ptr<double2>(st.Ainv[g]) + (size_t)w * n * ld;
  const double2* BV = det_ptr<double2>(st.U[g]) + (size_t)w * ld * k;
  const double2* Tw = T + (size_t)w * n * k;
  for (int o0 = 0; o0 < n; o0 += 64)
  {
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        acc[a][b] = zmk(0, 0);
    for (int c0 = 0; c0 < d; c0 += CF_KC)
    {
      __syncthreads();
      for (int e = t; e < CF_ROWS * CF_KC; e += WG)
      {
        const int rr = e / CF_KC, cc = e % CF_KC;
        const int i = i0 + rr, c = c0 + cc;
        Ts[rr][cc]  = (i < n && c < d) ? Tw[(size_t)i * k + c] : zmk(0, 0);
      }
      for (int e = t; e < 64 * CF_KC; e += WG)
      {
        const int oo = e / CF_KC, cc = e % CF_KC;
        const int o = o0 + oo, c = c0 + cc;
        Bv[cc][oo]  = (o < n && c < d) ? BV[(size_t)o * k + c] : zmk(0, 0);
      }
      __syncthreads();
#pragma unroll 4
      for (int cc = 0; cc < CF_KC; cc++)
      {
        double2 tv[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++)
          tv[a] = Ts[4 * tr + a][cc];
#pragma unroll
        for (int b = 0; b < 4; b++)
          bv[b] = Bv[cc][tc + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++)
            acc[a][b] = zfma(tv[a], bv[b], acc[a][b]);
      }
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
    {
      const int i = i0 + 4 * tr + a;
      if (i >= n)
        continue;
#pragma unroll
      for (int b = 0; b < 4; b++)
      {
        const int o = o0 + tc + 16 * b;
        if (o < n)
        {
          double2* ap = A + (size_t)i * ld + o;
          *ap         = zsub(*ap, acc[a][b]);
        }
      }
    }
  }
}

hipError_t launch_flush_c(const SysC& s, const State& st, int g, int d, int nw, double* scratch, hipStream_t stream)
{
  if (d <= 0)
    return hipSuccess;
  if (d > MAX_K)
    return hipErrorInvalidValue;
  const int n = s.n[g];
  const dim3 grid((n + CF_ROWS - 1) / CF_ROWS, nw);
  double2* T = reinterpret_cast<double2*>(scratch);
  k_flush_T_c<<<grid, WG, 0, stream>>>(s, st, g, d, T);
  k_flush_BV_c<<<nw, WG, 0, stream>>>(s, st, g, d);
  k_flush_update_c<<<grid, WG, 0, stream>>>(s, st, g, d, T);
  return hipGetLastError();
}

//------------------------------------------------------------------------------------------------
// API helpers: SPOSet::mw_evaluateVGL (out [nw][5][n] complex) and mw_evaluateDetRatios
// (inv [nw][n] complex, out [nw][nvp] complex)
__global__ __launch_bounds__(WG) void k_api_spo_vgl_c(SysC s, const double* __restrict__ coefs, const double* kpts,
                                                      int n, const double* __restrict__ pos, double2* __restrict__ out)
{
  __shared__ SplineLDS spl;
  const int w       = blockIdx.x;
  const double p[3] = {pos[w * 3], pos[w * 3 + 1], pos[w * 3 + 2]};
  if (threadIdx.x == 0)
  {
    SplineLoc L;
    spline_locate(s, p, L);
    spline_stage(L, spl);
  }
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double2 v[5];
    spline_vgl_orbital_c(s, coefs, spl, o, kpts, p, v);
    for (int q = 0; q < 5; q++)
      out[((size_t)w * 5 + q) * n + o] = v[q];
  }
}

hipError_t launch_api_spo_vgl_c(const SysC& s, const void* coefs, const double* kpts, int nw, int n,
                                const double* pos, double* out, hipStream_t stream)
{
  k_api_spo_vgl_c<<<nw, WG, 0, stream>>>(s, static_cast<const double*>(coefs), kpts, n, pos,
                                          reinterpret_cast<double2*>(out));
  return hipGetLastError();
}

__global__ __launch_bounds__(WG) void k_api_det_ratios_c(SysC s, const double* __restrict__ coefs,
                                                         const double* __restrict__ kpts, int nvp, int n,
                                                         const double* __restrict__ pos,
                                                         const double2* __restrict__ inv, double2* __restrict__ out)
{
  __shared__ double red[2 * WG];
  __shared__ double bc[2];
  const int v = blockIdx.x, w = blockIdx.y;
  const double* pp  = pos + ((size_t)w * nvp + v) * 3;
  const double p[3] = {pp[0], pp[1], pp[2]};
  SplineLoc L;
  spline_locate(s, p, L);
  double2 acc[1] = {zmk(0, 0)};
  for (int o = threadIdx.x; o < n; o += WG)
    acc[0] = zfma(spline_v_orbital_c(s, coefs, L, o, kpts, p), inv[(size_t)w * n + o], acc[0]);
  block_sum_z<1>(acc, red, bc);
  if (threadIdx.x == 0)
    out[(size_t)w * nvp + v] = zmk(bc[0], bc[1]);
}

// TrialWaveFunction::mw_evaluateRatios with complex orbitals: out[w][nvp] complex
struct VpLDSc
{
  alignas(16) double red[9 * WG];
  double bc[64];
  double jc[3][MAX_JCOEF];
};

template<int NJ>
__global__ __launch_bounds__(WG) void k_api_vp_ratios_c(SysC s, State st, const double* __restrict__ coefs, int iat,
                                                        int nvp, const double* __restrict__ pos, double2* __restrict__ out)
{
  __shared__ VpLDSc sh;
  stage_jastrow(s, sh);
  const int v = blockIdx.x, w = blockIdx.y, t = threadIdx.x;
  const int g = iat < s.nup ? 0 : 1, n = s.n[g], ld = s.ld, Np = s.Npad;
  const int r       = iat - s.first[g];
  const double* R   = st.R + (size_t)w * 3 * Np;
  const double* jas = st.jas + (size_t)w * J_ROWS * Np;
  const double* pp  = pos + ((size_t)w * nvp + v) * 3;
  const double p[3] = {pp[0], pp[1], pp[2]};
  jastrow_ratio_sums<NJ>(s, g, iat, R, st.ions, p, sh);
  double2 acc[1] = {zmk(0, 0)};
  if (s.has_det)
  {
    SplineLoc L;
    spline_locate(s, p, L);
    const double2* arow = det_ptr<double2>(st.Ainv[g]) + ((size_t)w * n + r) * ld;
    for (int o = t; o < n; o += WG)
      acc[0] = zfma(spline_v_orbital_c(s, coefs, L, o, st.kpts, p), arow[o], acc[0]);
  }
  block_sum_z<1>(acc, sh.red, sh.bc + 56);
  if (t == 0)
  {
    const double* C = sh.bc + BC_C;
    double2 ratio   = s.has_det ? bcz(sh.bc, 56) : zmk(1, 0);
    if (s.has_j1)
      ratio = zscale(ratio, exp(jas[J_VAT * Np + iat] - C[0]));
    if (s.has_j2)
      ratio = zscale(ratio, exp(jas[J_UAT * Np + iat] - C[5]));
    out[(size_t)w * nvp + v] = ratio;
  }
}

hipError_t launch_api_vp_ratios_c(const SysC& s, const State& st, const void* coefs, int nw, int iat, int nvp,
                                  const double* pos, double* out, hipStream_t stream)
{
  const int mm = move_maxm(s, iat < s.nup ? 0 : 1);
  if (mm == 0 || nvp <= 0)
    return hipErrorInvalidValue;
  const dim3 grid(nvp, nw);
  const double* c = static_cast<const double*>(coefs);
  double2* o      = reinterpret_cast<double2*>(out);
  if (mm == 1)
    k_api_vp_ratios_c<2><<<grid, WG, 0, stream>>>(s, st, c, iat, nvp, pos, o);
  else
    k_api_vp_ratios_c<4><<<grid, WG, 0, stream>>>(s, st, c, iat, nvp, pos, o);
  return hipGetLastError();
}

hipError_t launch_api_det_ratios_c(const SysC& s, const void* coefs, const double* kpts, int nw, int nvp, int n,
                                   const double* pos, const double* inv, double* out, hipStream_t stream)
{
  k_api_det_ratios_c<<<dim3(nvp, nw), WG, 0, stream>>>(s, static_cast<const double*>(coefs), kpts, nvp, n, pos,
                                                        reinterpret_cast<const double2*>(inv),
                                                        reinterpret_cast<double2*>(out));
  return hipGetLastError();
}

// SPOSet::mw_evaluateVGLandDetRatioGrads with complex orbitals (SplineC2C.cpp:140-279 for the orbital
// VGL): phi[5][nw][n] complex, rg[w][4] complex = ratio, gradient / ratio
__global__ __launch_bounds__(WG) void k_api_vgl_ratio_grads_c(SysC s, const double* __restrict__ coefs,
                                                              const double* __restrict__ kpts, int nw, int n,
                                                              const double* __restrict__ pos,
                                                              const double2* __restrict__ inv,
                                                              double2* __restrict__ phi, double2* __restrict__ rg)
{
  __shared__ SplineLDS spl;
  __shared__ double red[8 * WG];
  __shared__ double bc[8];
  const int w       = blockIdx.x;
  const double p[3] = {pos[w * 3], pos[w * 3 + 1], pos[w * 3 + 2]};
  if (threadIdx.x == 0)
  {
    SplineLoc L;
    spline_locate(s, p, L);
    spline_stage(L, spl);
  }
  __syncthreads();
  double2 acc[4] = {zmk(0, 0), zmk(0, 0), zmk(0, 0), zmk(0, 0)};
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double2 v[5];
    spline_vgl_orbital_c(s, coefs, spl, o, kpts, p, v);
    if (phi)
      for (int q = 0; q < 5; q++)
        phi[((size_t)q * nw + w) * n + o] = v[q];
    const double2 a = inv[(size_t)w * n + o];
    for (int q = 0; q < 4; q++)
      acc[q] = zfma(v[q], a, acc[q]);
  }
  block_sum_z<4>(acc, red, bc);
  if (threadIdx.x == 0)
  {
    const double2 r = bcz(bc, 0);
    rg[w * 4]       = r;
    for (int q = 1; q < 4; q++)
      rg[w * 4 + q] = zdiv(bcz(bc, 2 * q), r);
  }
}

hipError_t launch_api_vgl_ratio_grads_c(const SysC& s, const void* coefs, const double* kpts, int nw, int n,
                                        const double* pos, const double* inv, double* phi, double* rg,
                                        hipStream_t stream)
{
  k_api_vgl_ratio_grads_c<<<nw, WG, 0, stream>>>(s, static_cast<const double*>(coefs), kpts, nw, n, pos,
                                                  reinterpret_cast<const double2*>(inv),
                                                  reinterpret_cast<double2*>(phi), reinterpret_cast<double2*>(rg));
  return hipGetLastError();
}

} // namespace qmi
