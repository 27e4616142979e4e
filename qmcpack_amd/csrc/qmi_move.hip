// Per-walker move kernels: one 256-thread workgroup per walker.
//
// Phases of one proposed single-electron move (VMCBatched::advanceWalkers inner loop,
// QMCDrivers/VMC/VMCBatched.cpp:119-172):
//   A  inverse row of the moving electron brought up to date through the delayed rows
//      (DelayedUpdateBatched::mw_prepareInvRow, DelayedUpdateBatched.h:174-235) and the
//      determinant gradient (mw_evalGrad, :353-400)
//   B  drift-diffusion proposal (DriftModifierUNR, VMCBatched.cpp:121-134)
//   C  J1/J2 ratio+gradient from fresh distance rows, orbital VGL at the proposed position and
//      determinant ratio/gradient (mw_calcRatioGrad, TrialWaveFunction.cpp:685-741)
//   D  Green's functions and Metropolis-Hastings (VMCBatched.cpp:141-167)
//   E  accept (rank-1 growth of the delayed-update block, DelayedUpdateBatched.h:542-670, or the
//      Sherman-Morrison update for delay rank 1, :252-340; Jastrow accumulators,
//      TwoBodyJastrow.cpp:581-628, J1OrbitalSoA.h:456-468) or pseudo-accept on reject.
#include "qmi_mlp.hpp"
#include "qmi_move_common.hpp"

namespace qmi
{

// out[c] = scale * sum_o X[o][c] vec[o] for c < d over an orbital-major delayed block X (row
// stride k).  With dd = d rounded up to a power of two (<= 64), thread t works on column
// c = t % dd for the orbitals o = q, q + G, ... of its group q = t / dd (G = WG / dd groups), so
// every thread of the workgroup is busy at any d and each runs n dd / WG independent loads.  The
// group partials of a column are added by xor shuffles inside a wave (lanes of equal c differ by
// multiples of dd) and across the four waves through LDS.  Ends with a barrier.
template<typename DT>
__device__ inline void delayed_gemv(const DT* __restrict__ X, const double* vec, int n, int k, int d,
                                    double* red, double* out, double scale)
{
  const int t = thread_index();
  int dd      = 1;
  while (dd < d)
    dd <<= 1;
  const int G = WG / dd, c = t & (dd - 1), q = t / dd;
  double a0 = 0;
  if (c < d)
  {
    const DT* xp = X + c;
#pragma unroll 8
    for (int o = q; o < n; o += G)
      a0 += (double)xp[(size_t)o * k] * vec[o];
  }
  for (int off = 32; off >= dd; off >>= 1)
    a0 += __shfl_xor(a0, off, 64);
  const int lane = t & 63, wave = t >> 6;
  if (lane < dd)
    red[wave * 64 + lane] = a0;
  __syncthreads();
  if (t < d)
    out[t] = scale * (red[t] + red[64 + t] + red[128 + t] + red[192 + t]);
  __syncthreads();
}

// sum_{c<d} v[c] y[c] over one orbital-major row; y is zero-padded to k so that an even-k row can
// be read in pairs (the column past an odd d is a finite stored value times zero).
template<typename DT>
struct Pair2;
template<>
struct Pair2<double>
{
  using T = double2;
};
template<>
struct Pair2<float>
{
  using T = float2;
};

template<typename DT>
__device__ inline double row_dot(const DT* __restrict__ v, const double* y, int d, int k)
{
  double a0 = 0, a1 = 0;
  if ((k & 1) == 0)
  {
    using P2     = typename Pair2<DT>::T;
    const P2* v2 = reinterpret_cast<const P2*>(v);
    const int h  = (d + 1) >> 1;
#pragma unroll 4
    for (int c = 0; c < h; c++)
    {
      const P2 x = v2[c];
      a0 += (double)x.x * y[2 * c];
      a1 += (double)x.y * y[2 * c + 1];
    }
  }
  else
  {
#pragma unroll 4
    for (int c = 0; c < d; c++)
      a0 += (double)v[c] * y[c];
  }
  return a0 + a1;
}

//------------------------------------------------------------------------------------------------
template<typename DT, int MAXM>
__device__ inline void phaseA(const SysC& s, int n, int ld, int k, int d, int r, const DT* __restrict__ Ainv,
                              const DT* __restrict__ U, const DT* __restrict__ V, DT* __restrict__ Binv, MoveLDS& sh,
                              DT* sB, double (&ainv_r)[MAXM], double (&inv)[MAXM])
{
  const int t    = thread_index();
  const DT* arow = Ainv + (size_t)r * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
    ainv_r[m]   = o < n ? arow[o] : 0.0;
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
    // p = U invRow over the d delayed rows
    delayed_gemv(U, sh.row, n, k, d, sh.red, sh.p, 1.0);
    // Binv[d][i] = -sum_j Binv[j][i] p[j]
    if (t < k)
    {
      double acc = 0;
      if (t < d)
        for (int j = 0; j < d; j++)
          acc += sB[j * (k + 1) + t] * sh.p[j];
      const double y = -1.0 * acc;
      sh.y[t]        = t < d ? y : 0.0;
      if (t < d)
        Binv[d * k + t] = (DT)y;
    }
    __syncthreads();
    // invRow = Ainv[r] + V^T y
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + WG * m;
      if (o < n)
        inv[m] = 1.0 * row_dot(V + (size_t)o * k, sh.y, d, k) + 1.0 * ainv_r[m];
    }
  }
}

// the orbital gradients of row r, loaded before phase A so that their latency overlaps it
template<typename DT, int MAXM>
__device__ inline void load_dpsi_row(int n, int ld, int r, const DT* __restrict__ dpsi, double (&dg)[MAXM][3])
{
  const int t    = thread_index();
  const DT* drow = dpsi + (size_t)r * 4 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
#pragma unroll
    for (int q = 0; q < 3; q++)
      dg[m][q] = o < n ? (double)drow[q * ld + o] : 0.0;
  }
}

// det_grad from preloaded gradients (the same sums)
template<int MAXM>
__device__ inline void det_grad_pre(int n, const double (&dg)[MAXM][3], const double (&inv)[MAXM], MoveLDS& sh)
{
  const int t  = thread_index();
  double gv[3] = {0, 0, 0};
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
    if (o < n)
    {
      gv[0] += inv[m] * dg[m][0];
      gv[1] += inv[m] * dg[m][1];
      gv[2] += inv[m] * dg[m][2];
    }
  }
  block_sum<3>(gv, sh.red, sh.bc + BC_GRAD);
}

// determinant gradient at the current position: invRow . dpsiM[r]  -> sh.bc[BC_GRAD..+3)
template<typename DT, int MAXM>
__device__ inline void det_grad(int n, int ld, int r, const DT* __restrict__ dpsi, const double (&inv)[MAXM],
                                MoveLDS& sh)
{
  const int t    = thread_index();
  double gv[3]   = {0, 0, 0};
  const DT* drow = dpsi + (size_t)r * 4 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = t + WG * m;
    if (o < n)
    {
      gv[0] += inv[m] * (double)drow[o];
      gv[1] += inv[m] * (double)drow[ld + o];
      gv[2] += inv[m] * (double)drow[2 * ld + o];
    }
  }
  block_sum<3>(gv, sh.red, sh.bc + BC_GRAD);
}

// Phase C: all threads; position in sh.bc[BC_POS].  The Jastrow sums are reduced before the
// spline gather so that they are not live across it.
template<typename ST, int MAXM>
__device__ inline void phaseC(const SysC& s, int g, int iat, int n, const double* __restrict__ R,
                              const double* __restrict__ ions, const ST* __restrict__ coefs,
                              const double (&inv)[MAXM], double (&phv)[MAXM][5], MoveLDS& sh)
{
  const int t = thread_index();
  const int N = s.N, Np = s.Npad;
  const double pos[3] = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
  if (s.has_det && t == 64) // locate once; the Jastrow reduction's barriers publish it
  {
    SplineLoc L;
    spline_locate(s, pos, L);
    spline_stage(L, sh.spl);
  }
  jastrow_ratio_sums<2 * MAXM>(s, g, iat, R, ions, pos, sh);
  double acc[4] = {0, 0, 0, 0};
  if (s.has_det)
  {
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + WG * m;
      if (o < n)
      {
        spline_vgl_orbital_lds<ST>(s, coefs, sh.spl, o, phv[m]);
        acc[0] += inv[m] * phv[m][0];
        acc[1] += inv[m] * phv[m][1];
        acc[2] += inv[m] * phv[m][2];
        acc[3] += inv[m] * phv[m][3];
      }
      else
      {
#pragma unroll
        for (int q = 0; q < 5; q++)
          phv[m][q] = 0.0;
      }
    }
  }
  block_sum<4>(acc, sh.red, sh.bc + BC_C + 9);
}


// Phase D helper (one thread): total ratio, gradient at the proposed position, log G_b; vat and
// uat are the electron's current J1 and J2 values.
__device__ inline void combine_ratio_v(const SysC& s, double vat, double uat, const MoveLDS& sh, double& ratio,
                                       double gnew[3])
{
  const double* C = sh.bc + BC_C;
  ratio           = 1.0;
  gnew[0] = gnew[1] = gnew[2] = 0.0;
  if (s.has_det)
  {
    ratio *= C[9];
    gnew[0] += C[10] / C[9];
    gnew[1] += C[11] / C[9];
    gnew[2] += C[12] / C[9];
  }
  if (s.has_j1)
  {
    gnew[0] += C[1];
    gnew[1] += C[2];
    gnew[2] += C[3];
    ratio *= exp(vat - C[0]);
  }
  if (s.has_j2)
  {
    gnew[0] += C[6];
    gnew[1] += C[7];
    gnew[2] += C[8];
    ratio *= exp(uat - C[5]);
  }
}

__device__ inline void combine_ratio(const SysC& s, int iat, const double* __restrict__ jas, const MoveLDS& sh,
                                     double& ratio, double gnew[3])
{
  const int Np = s.Npad;
  combine_ratio_v(s, s.has_j1 ? jas[J_VAT * Np + iat] : 0.0, s.has_j2 ? jas[J_UAT * Np + iat] : 0.0, sh, ratio,
                  gnew);
}

// Phase E: accept (sh.bc[BC_ACC] != 0) or reject.  All threads.
template<typename DT, int MAXM>
__device__ inline void phaseE(const SysC& s, int g, int iat, int n, int ld, int k, int d, int r, double* __restrict__ R,
                              DT* __restrict__ Ainv, DT* __restrict__ dpsi, DT* __restrict__ U, DT* __restrict__ V,
                              DT* __restrict__ Binv, int* __restrict__ dl, double* __restrict__ jas,
                              double* __restrict__ logdet, const double (&ainv_r)[MAXM], const double (&phv)[MAXM][5],
                              MoveLDS& sh, DT* sB, bool stage)
{
  const int t = thread_index();
  const int N = s.N, Np = s.Npad;
  const bool accepted = sh.bc[BC_ACC] != 0.0;
  const double* C     = sh.bc + BC_C;
  if (s.has_det)
  {
    const double ratio_det = C[9];
    if (k > 1)
    {
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + WG * m;
        if (o < n)
          V[(size_t)o * k + d] = (DT)ainv_r[m];
      }
      if (accepted)
      {
        const double sigma = 1.0 / ratio_det;
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            sh.vec[o] = phv[m][0];
        }
        if (stage)
          stage_binv(Binv, k, d, sB);
        __syncthreads();
        // p = -V phi over the d delayed rows (column d of V, written above, is not included)
        delayed_gemv(V, sh.vec, n, k, d, sh.red, sh.p, -1.0);
        if (t < d)
        {
          double acc = 0;
          for (int j = 0; j < d; j++)
            acc += sB[t * (k + 1) + j] * sh.p[j];
          const double Yv = sigma * acc;
          Binv[t * k + d] = (DT)Yv;
          sh.Y[t]         = Yv;
        }
        __syncthreads();
        for (int idx = t; idx < d * d; idx += WG)
        {
          const int j = idx / d, i = idx - j * d;
          Binv[j * k + i] = (DT)((double)sB[j * (k + 1) + i] + 1.0 * sh.y[i] * sh.Y[j]);
        }
        if (t < d)
          Binv[d * k + t] = (DT)(sh.y[t] * sigma);
        if (t == 0)
        {
          Binv[d * k + d] = (DT)sigma;
          dl[d]           = r;
        }
        DT* drow = dpsi + (size_t)r * 4 * ld;
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
          {
            U[(size_t)o * k + d] = (DT)phv[m][0];
            drow[o]              = (DT)phv[m][1];
            drow[ld + o]         = (DT)phv[m][2];
            drow[2 * ld + o]     = (DT)phv[m][3];
            drow[3 * ld + o]     = (DT)phv[m][4];
          }
        }
      }
      else
      {
        if (t < d)
        {
          Binv[d * k + t] = 0.0;
          Binv[t * k + d] = 0.0;
        }
        if (t == 0)
        {
          Binv[d * k + d] = 1.0;
          dl[d]           = -1;
        }
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            U[(size_t)o * k + d] = 0.0;
        }
      }
    }
    else if (accepted)
    {
      // Fahy's Sherman-Morrison update: temp = Ainv phi, temp[r] -= 1,
      // Ainv[i][:] += (-1/ratio * temp[i]) * Ainv_old[r][:]
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
          double acc      = 0;
          const DT* ai = Ainv + (size_t)i * ld;
          for (int o = 0; o < n; o++)
            acc += (double)ai[o] * sh.vec[o];
          sh.row[i] = acc;
        }
      }
      __syncthreads();
      if (t == 0)
        sh.row[r] = sh.row[r] - 1.0;
      __syncthreads();
      const double c_ratio_inv = -1.0 / ratio_det;
      for (int i = 0; i < n; i++)
      {
        const double ci = c_ratio_inv * sh.row[i];
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const int o = t + WG * m;
          if (o < n)
            Ainv[(size_t)i * ld + o] = (DT)((double)Ainv[(size_t)i * ld + o] + ci * ainv_r[m]);
        }
      }
      DT* drow = dpsi + (size_t)r * 4 * ld;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + WG * m;
        if (o < n)
        {
          drow[o]          = (DT)phv[m][1];
          drow[ld + o]     = (DT)phv[m][2];
          drow[2 * ld + o] = (DT)phv[m][3];
          drow[3 * ld + o] = (DT)phv[m][4];
        }
      }
    }
    if (accepted && t == 0)
    {
      logdet[g * 2 + 0] += log(fabs(ratio_det));
      if (ratio_det < 0)
        logdet[g * 2 + 1] += M_PI;
    }
  }
  if (accepted)
    jastrow_accept<2 * MAXM>(s, g, iat, R, jas, sh);
  __syncthreads();
}

//------------------------------------------------------------------------------------------------
// Phase D core (thread 0): returns prob_eff; flag = |ratio|^2 >= eps.
__device__ inline double mh_prob_v(const SysC& s, double vat, double uat, const MoveArgs& a, MoveLDS& sh,
                                   double& ratio, bool& flag, double& gn2)
{
  const double tauovermass = a.tau;
  const double oneover2tau = 0.5 / tauovermass;
  double gnew[3];
  combine_ratio_v(s, vat, uat, sh, ratio, gnew);
  gn2 = gnew[0] * gnew[0] + gnew[1] * gnew[1] + gnew[2] * gnew[2];
  double drev[3];
  unr_drift(tauovermass, gnew, drev);
  for (int q = 0; q < 3; q++)
    drev[q] += sh.bc[BC_MOVE + q];
  const double log_gb = -oneover2tau * (drev[0] * drev[0] + drev[1] * drev[1] + drev[2] * drev[2]);
  const double prob   = ratio * ratio;
  flag                = prob >= DBL_EPS;
  sh.bc[BC_RATIO]     = ratio;
  return prob * exp(log_gb - sh.bc[BC_LOGGF]);
}

__device__ inline double mh_prob(const SysC& s, int iat, const MoveArgs& a, const double* __restrict__ jas,
                                 MoveLDS& sh, double& ratio, bool& flag, double& gn2)
{
  const double tauovermass = a.tau;
  const double oneover2tau = 0.5 / tauovermass;
  double gnew[3];
  combine_ratio(s, iat, jas, sh, ratio, gnew);
  gn2 = gnew[0] * gnew[0] + gnew[1] * gnew[1] + gnew[2] * gnew[2];
  double drev[3];
  unr_drift(tauovermass, gnew, drev);
  for (int q = 0; q < 3; q++)
    drev[q] += sh.bc[BC_MOVE + q];
  const double log_gb = -oneover2tau * (drev[0] * drev[0] + drev[1] * drev[1] + drev[2] * drev[2]);
  const double prob   = ratio * ratio;
  flag                = prob >= DBL_EPS;
  sh.bc[BC_RATIO]     = ratio;
  return prob * exp(log_gb - sh.bc[BC_LOGGF]);
}

//------------------------------------------------------------------------------------------------
// One fused move (phases A-E) of electron iat for walker w at delay d; the walker's uniform
// counter is carried in ucount.
template<typename ST, typename DT, int MAXM>
__device__ inline void fused_move(const SysC& s, const State& st, const ST* __restrict__ coefs, const MoveArgs& a,
                                  int w, int iat, int d, MoveLDS& sh, DT* sB, int& ucount, double& nacc,
                                  double& nrej)
{
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  const int r = iat - s.first[g];
  double ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    ainv_r[m] = inv[m] = 0.0;
  }
  prefetch_move(s, iat, w, ucount, a, v.R, v.jas, sh);
  if (s.has_det)
  {
    double dg[MAXM][3];
    load_dpsi_row<DT, MAXM>(n, ld, r, v.dpsi, dg);
    phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
    det_grad_pre<MAXM>(n, dg, inv, sh);
  }
  else
    __syncthreads(); // publishes the prefetched scalars
  if (thread_index() == 0)
    propose_pf(s, a, sh);
  __syncthreads();
  phaseC<ST, MAXM>(s, g, iat, n, v.R, st.ions, coefs, inv, phv, sh);
  if (thread_index() == 0)
  {
    double ratio;
    bool flag;
    double gn2;
    const double pe = mh_prob_v(s, sh.bc[BC_PF + PF_VAT], sh.bc[BC_PF + PF_UAT], a, sh, ratio, flag, gn2);
    nan_guard(a, w, iat, ratio * ratio, gn2, gnow_norm2(sh.bc));
    if (a.dmc) // a node crossing (ratio <= 0) is rejected before the draw; prob_eff against eps
      flag = ratio > 0.0 && pe >= DBL_EPS;
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
    // arithmetic selects (a branch on acc made the pair a dynamically indexed private array)
    nacc += acc ? 1.0 : 0.0;
    nrej += acc ? 0.0 : 1.0;
    if (a.trace)
    {
      double* tr = a.trace + ((size_t)iat * a.nw + w) * TRACE_FIELDS;
      tr[0]      = ratio;
      tr[1]      = acc ? 1.0 : 0.0;
      for (int q = 0; q < 3; q++)
      {
        tr[2 + q] = sh.bc[BC_POS + q];
        tr[5 + q] = sh.bc[BC_GNOW + q];
      }
    }
  }
  __syncthreads();
  phaseE<DT, MAXM>(s, g, iat, n, ld, k, d, r, v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.dl, v.jas, v.logdet, ainv_r, phv,
               sh, sB, false);
}

// Fused sweep segment: moves [iat0, iat1) of spin group g for every walker (one walker per
// workgroup, one RNG stream per walker), delay count d0 at iat0 with d0 + (iat1 - iat0) <= k.
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG, 4) void k_sweep_segment(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = a.w0 + blockIdx.x;
  if (w >= a.nw)
    return;
  const int N = s.N;
  int ucount  = a.ubase[(size_t)w * (N + 1) + a.iat0];
  double nacc = 0, nrej = 0;
#pragma unroll 1
  for (int iat = a.iat0; iat < a.iat1; ++iat)
  {
    const int d = s.k > 1 ? a.d0 + (iat - a.iat0) : 0;
    fused_move<ST, DT, MAXM>(s, st, coefs, a, w, iat, d, sh, sB, ucount, nacc, nrej);
  }
  if (thread_index() == 0)
  {
    a.ubase[(size_t)w * (N + 1) + a.iat1] = ucount;
    st.stats[(size_t)w * 2 + 0] += nacc;
    st.stats[(size_t)w * 2 + 1] += nrej;
  }
}

// The same fused move for a single electron a.iat0 (no move loop in the kernel: the register
// allocation of the loop version is about twice that of one move, which halves occupancy).
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_move_one(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = a.w0 + blockIdx.x;
  if (w >= a.nw)
    return;
  const int N = s.N;
  int ucount  = a.ubase[(size_t)w * (N + 1) + a.iat0];
  double nacc = 0, nrej = 0;
  fused_move<ST, DT, MAXM>(s, st, coefs, a, w, a.iat0, a.d0, sh, sB, ucount, nacc, nrej);
  if (thread_index() == 0)
  {
    a.ubase[(size_t)w * (N + 1) + a.iat0 + 1] = ucount;
    st.stats[(size_t)w * 2 + 0] += nacc;
    st.stats[(size_t)w * 2 + 1] += nrej;
  }
}

// Crowd path, first kernel: phases A-C and the MH probability for one move; the proposed orbitals
// go to st.phi and the scalars to st.mv.
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_propose(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
  if (s.has_det)
  {
    phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
    det_grad<DT, MAXM>(n, ld, r, v.dpsi, inv, sh);
  }
  if (thread_index() == 0)
    propose(s, iat, w, a, v.R, v.jas, sh);
  __syncthreads();
  phaseC<ST, MAXM>(s, g, iat, n, v.R, st.ions, coefs, inv, phv, sh);
  double* mv = st.mv + (size_t)w * MV_SIZE;
  if (thread_index() == 0)
  {
    double ratio;
    bool flag;
    double gn2;
    const double pe = mh_prob(s, iat, a, v.jas, sh, ratio, flag, gn2);
    nan_guard(a, w, iat, ratio * ratio, gn2, gnow_norm2(sh.bc));
    mv[MV_RATIO]    = ratio;
    mv[MV_PROB]     = pe;
    mv[MV_FLAG]     = flag ? 1.0 : 0.0;
    for (int q = 0; q < 3; q++)
    {
      mv[MV_POS + q]  = sh.bc[BC_POS + q];
      mv[MV_GNOW + q] = sh.bc[BC_GNOW + q];
    }
    for (int q = 0; q < 13; q++)
      mv[MV_C + q] = sh.bc[BC_C + q];
  }
  if (s.has_det) // the proposed orbitals (no orbital buffers without determinants)
  {
    double* phi = st.phi + (size_t)w * 5 * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = thread_index() + WG * m;
      if (o < n)
        for (int q = 0; q < 5; q++)
          phi[q * ld + o] = phv[m][q];
    }
  }
}

// Crowd path, second kernel: Metropolis-Hastings with the crowd's shared RNG stream (uniform index =
// crowd base + number of earlier walkers of the crowd that draw), then phase E.
template<typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_accept_crowd(SysC s, State st, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  const int crowd = w / a.wpc, w0 = crowd * a.wpc, w1 = min(a.nw, w0 + a.wpc);
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  const double* mv = st.mv + (size_t)w * MV_SIZE;
  // prefix count of drawing walkers before w in the crowd, and the crowd total
  double cnt[2] = {0, 0};
  for (int j = w0 + thread_index(); j < w1; j += WG)
  {
    const double f = st.mv[(size_t)j * MV_SIZE + MV_FLAG];
    cnt[1] += f;
    if (j < w)
      cnt[0] += f;
  }
  block_sum<2>(cnt, sh.red, sh.bc + BC_OLD);
  if (thread_index() == 0)
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
    for (int q = 0; q < 3; q++)
      sh.bc[BC_POS + q] = mv[MV_POS + q];
    for (int q = 0; q < 13; q++)
      sh.bc[BC_C + q] = mv[MV_C + q];
  }
  if (s.has_det && k > 1 && thread_index() < d)
    sh.y[thread_index()] = v.Binv[d * k + thread_index()];
  __syncthreads();
  double ainv_r[MAXM], phv[MAXM][5];
  const double* phi = st.phi + (size_t)w * 5 * ld;
  const int nd      = s.has_det ? n : 0; // no determinant buffers without determinants
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = thread_index() + WG * m;
    ainv_r[m]   = o < nd ? v.Ainv[(size_t)r * ld + o] : 0.0;
    for (int q = 0; q < 5; q++)
      phv[m][q] = o < nd ? phi[q * ld + o] : 0.0;
  }
  __syncthreads();
  phaseE<DT, MAXM>(s, g, iat, n, ld, k, d, r, v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.dl, v.jas, v.logdet, ainv_r, phv,
               sh, sB, true);
}

//------------------------------------------------------------------------------------------------
// API kernels (reference call protocol, host-side Metropolis-Hastings)

// mw_evalGrad: inverse row (stored to st.invRow) and total gradient at R[iat] -> out[nw][3]
template<typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_api_eval_grad(SysC s, State st, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, Np = s.Npad;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double ainv_r[MAXM], inv[MAXM];
  if (s.has_det)
  {
    phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
    det_grad<DT, MAXM>(n, ld, r, v.dpsi, inv, sh);
    double* ir = st.invRow + (size_t)w * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = thread_index() + WG * m;
      if (o < n)
        ir[o] = inv[m];
    }
  }
  if (thread_index() == 0)
  {
    double gnow[3] = {0, 0, 0};
    for (int q = 0; q < 3; q++)
    {
      if (s.has_det)
        gnow[q] += sh.bc[BC_GRAD + q];
      if (s.has_j1)
        gnow[q] += v.jas[(J_G1X + q) * Np + iat];
      if (s.has_j2)
        gnow[q] += v.jas[(J_DUX + q) * Np + iat];
      a.out[(size_t)w * 3 + q] = gnow[q];
    }
  }
}

// mw_calcRatioGrad at the position stored by mw_makeMove (st.mv[MV_POS]); prepares the inverse
// row first when `a.d0 >= 0` and the row is not prepared (a.flag_prepare).
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_api_ratio_grad(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double* mv = st.mv + (size_t)w * MV_SIZE;
  double ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
  if (s.has_det)
  {
    if (a.prepare)
    {
      phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
      double* ir = st.invRow + (size_t)w * ld;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = thread_index() + WG * m;
        if (o < n)
          ir[o] = inv[m];
      }
    }
    else
    {
      const double* ir = st.invRow + (size_t)w * ld;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = thread_index() + WG * m;
        inv[m]      = o < n ? ir[o] : 0.0;
      }
    }
  }
  if (thread_index() < 3)
    sh.bc[BC_POS + thread_index()] = mv[MV_POS + thread_index()];
  __syncthreads();
  phaseC<ST, MAXM>(s, g, iat, n, v.R, st.ions, coefs, inv, phv, sh);
  if (thread_index() == 0)
  {
    double ratio, gnew[3];
    combine_ratio(s, iat, v.jas, sh, ratio, gnew);
    mv[MV_RATIO] = ratio;
    for (int q = 0; q < 13; q++)
      mv[MV_C + q] = sh.bc[BC_C + q];
    a.out[w] = ratio;
    for (int q = 0; q < 3; q++)
      a.out[a.nw + (size_t)w * 3 + q] = gnew[q];
  }
  if (s.has_det) // the proposed orbitals (no orbital buffers without determinants)
  {
    double* phi = st.phi + (size_t)w * 5 * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = thread_index() + WG * m;
      if (o < n)
        for (int q = 0; q < 5; q++)
          phi[q * ld + o] = phv[m][q];
    }
  }
}

// mw_accept_rejectMove with host decisions a.host_acc[nw]
template<typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_api_accept(SysC s, State st, MoveArgs a)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  const double* mv = st.mv + (size_t)w * MV_SIZE;
  if (thread_index() == 0)
  {
    const bool acc = a.host_acc[w] != 0;
    sh.bc[BC_ACC]  = acc ? 1.0 : 0.0;
    v.stats[acc ? 0 : 1] += 1;
    for (int q = 0; q < 3; q++)
      sh.bc[BC_POS + q] = mv[MV_POS + q];
    for (int q = 0; q < 13; q++)
      sh.bc[BC_C + q] = mv[MV_C + q];
  }
  if (s.has_det && k > 1 && thread_index() < d)
    sh.y[thread_index()] = v.Binv[d * k + thread_index()];
  __syncthreads();
  double ainv_r[MAXM], phv[MAXM][5];
  const double* phi = st.phi + (size_t)w * 5 * ld;
  const int nd      = s.has_det ? n : 0; // no determinant buffers without determinants
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = thread_index() + WG * m;
    ainv_r[m]   = o < nd ? v.Ainv[(size_t)r * ld + o] : 0.0;
    for (int q = 0; q < 5; q++)
      phv[m][q] = o < nd ? phi[q * ld + o] : 0.0;
  }
  __syncthreads();
  phaseE<DT, MAXM>(s, g, iat, n, ld, k, d, r, v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.dl, v.jas, v.logdet, ainv_r, phv,
               sh, sB, true);
}

// LDS of the virtual-move ratio kernels: the Jastrow reductions only
struct VpLDS
{
  alignas(16) double red[9 * WG];
  double bc[48];
  double jc[3][MAX_JCOEF];
};

// TrialWaveFunction::mw_evaluateRatios (TrialWaveFunction.cpp:1079-1110) for virtual moves of
// electron iat (one VirtualParticleSet per walker, positions pos[w][nvp][3]): psi with r_iat at
// each virtual position over psi, as the product of the determinant ratio (DiracDeterminantBatched::
// mw_evaluateRatios, DiracDeterminantBatched.cpp:812-846: value-only orbitals dotted with the stored
// inverse row, so no delayed update may be pending), the J1 ratio (J1OrbitalSoA.h:295-299) and the
// J2 ratio (TwoBodyJastrow.cpp:166-171, the pair with iat itself excluded).  Grid (nvp, nw).
template<typename ST, typename DT, int NJ>
__global__ __launch_bounds__(WG) void k_api_vp_ratios(SysC s, State st, const ST* __restrict__ coefs, int iat, int nvp,
                                                      const double* __restrict__ pos, double* __restrict__ out)
{
  __shared__ VpLDS sh;
  stage_jastrow(s, sh);
  const int v = blockIdx.x, w = blockIdx.y, t = thread_index();
  const int g = iat < s.nup ? 0 : 1, n = s.n[g], ld = s.ld, Np = s.Npad;
  const int r       = iat - s.first[g];
  const double* R   = st.R + (size_t)w * 3 * Np;
  const double* jas = st.jas + (size_t)w * J_ROWS * Np;
  const double* pp  = pos + ((size_t)w * nvp + v) * 3;
  const double p[3] = {pp[0], pp[1], pp[2]};
  jastrow_ratio_sums<NJ>(s, g, iat, R, st.ions, p, sh);
  double acc[1] = {0};
  if (s.has_det)
  {
    SplineLoc L;
    spline_locate(s, p, L);
    const DT* arow = det_ptr<DT>(st.Ainv[g]) + ((size_t)w * n + r) * ld;
    for (int o = t; o < n; o += WG)
      acc[0] += spline_v_orbital<ST>(s, coefs, L, o) * (double)arow[o];
  }
  block_sum<1>(acc, sh.red, sh.bc + 40);
  if (t == 0)
  {
    const double* C = sh.bc + BC_C;
    double ratio    = 1.0;
    if (s.has_det)
      ratio *= sh.bc[40];
    if (s.has_j1)
      ratio *= exp(jas[J_VAT * Np + iat] - C[0]);
    if (s.has_j2)
      ratio *= exp(jas[J_UAT * Np + iat] - C[5]);
    out[(size_t)w * nvp + v] = ratio;
  }
}

hipError_t launch_api_vp_ratios(const SysC& s, const State& st, const void* coefs, int nw, int iat, int nvp,
                                const double* pos, double* out, hipStream_t stream)
{
  const int mm = move_maxm(s, iat < s.nup ? 0 : 1);
  if (mm == 0 || nvp <= 0)
    return hipErrorInvalidValue;
  return dispatch_prec(s.prec, [&](auto stv, auto dtv) {
    using ST = decltype(stv);
    using DT = decltype(dtv);
    const dim3 grid(nvp, nw);
    if (mm == 1)
      k_api_vp_ratios<ST, DT, 2><<<grid, WG, 0, stream>>>(s, st, static_cast<const ST*>(coefs), iat, nvp, pos, out);
    else
      k_api_vp_ratios<ST, DT, 4><<<grid, WG, 0, stream>>>(s, st, static_cast<const ST*>(coefs), iat, nvp, pos, out);
    return hipGetLastError();
  });
}

// mw_makeMove: proposed position R[iat] + displ
__global__ void k_api_make_move(SysC s, State st, int iat, int nw, const double* __restrict__ displ)
{
  const int w = blockIdx.x * blockDim.x + thread_index();
  if (w >= nw)
    return;
  const double* R = st.R + (size_t)w * 3 * s.Npad;
  double* mv      = st.mv + (size_t)w * MV_SIZE;
  for (int q = 0; q < 3; q++)
    mv[MV_POS + q] = R[q * s.Npad + iat] + displ[(size_t)w * 3 + q];
}

//------------------------------------------------------------------------------------------------
// Learned proposal (no counterpart in the reference).

// Indices of the K smallest values of r[0, count) (ties by the lower index) into sel[0, K) (-1
// where fewer than K are finite).  The largest finite value sets the range of a 64-bin histogram;
// the bin holding the K-th value bounds the candidates, which are compacted and ranked among
// themselves (a full ranking when more than SEL_CAP fall below the bound).  scratch: at least
// SEL_CAP + 68 ints of LDS.  Call with all threads; ends with a barrier.
constexpr int SEL_BINS = 64, SEL_CAP = 256;
__device__ inline void select_nearest(const double* r, int count, int K, int* sel, int* scratch)
{
  const int t   = thread_index();
  int* hist     = scratch;              // [SEL_BINS]
  int* cnt      = scratch + SEL_BINS;   // candidate count
  int* rmaxbits = scratch + SEL_BINS + 1;
  int* bstar    = scratch + SEL_BINS + 2;
  int* list     = scratch + SEL_BINS + 4; // [SEL_CAP]
  if (t < SEL_BINS)
    hist[t] = 0;
  if (t == 0)
  {
    *cnt      = 0;
    *rmaxbits = 0;
  }
  if (t < K)
    sel[t] = -1;
  __syncthreads();
  for (int j = t; j < count; j += WG)
    if (r[j] < INFINITY)
      atomicMax(rmaxbits, __float_as_int((float)r[j]));
  __syncthreads();
  const double rcap = (double)__int_as_float(*rmaxbits) * (1.0 + 1e-6) + 1e-30;
  for (int j = t; j < count; j += WG)
    if (r[j] < INFINITY)
      atomicAdd(&hist[min(SEL_BINS - 1, (int)(r[j] / rcap * SEL_BINS))], 1);
  __syncthreads();
  if (t == 0)
  {
    int c = 0, b = 0;
    for (; b < SEL_BINS; b++)
    {
      c += hist[b];
      if (c >= K)
        break;
    }
    *bstar = min(b, SEL_BINS - 1);
  }
  __syncthreads();
  const int bs = *bstar;
  for (int j = t; j < count; j += WG)
    if (r[j] < INFINITY && min(SEL_BINS - 1, (int)(r[j] / rcap * SEL_BINS)) <= bs)
    {
      const int slot = atomicAdd(cnt, 1);
      if (slot < SEL_CAP)
        list[slot] = j;
    }
  __syncthreads();
  const int m = *cnt;
  if (m <= SEL_CAP)
  {
    if (t < m)
    {
      const int j     = list[t];
      const double rj = r[j];
      int rank        = 0;
      for (int i = 0; i < m; i++)
      {
        const int ji    = list[i];
        const double ri = r[ji];
        rank += (ri < rj) || (ri == rj && ji < j);
      }
      if (rank < K)
        sel[rank] = j;
    }
  }
  else
  {
    for (int j = t; j < count; j += WG)
    {
      const double rj = r[j];
      if (!(rj < INFINITY))
        continue;
      int rank = 0;
      for (int i = 0; i < count; i++)
      {
        const double ri = r[i];
        rank += (ri < rj) || (ri == rj && i < j);
      }
      if (rank < K)
        sel[rank] = j;
    }
  }
  __syncthreads();
}

// Features of electron iat at pos with drift vd: the 4 nearest ions (dx, dy, dz, |d|), the 16
// nearest other electrons (dx, dy, dz) (d = minimum image of source - pos), the drift and tau,
// written to x[68] (FP32).  Neighbours are ranked by minimum-image distance, ties by the lower
// index; missing neighbours are zero.  sh.red holds the distances (electrons at [0, N), ions at
// [1024, 1024 + nion)); sh.row / sh.vec serve as the selection scratch.  Starts and ends with a
// barrier.
__device__ inline void proposal_features(const SysC& s, const double* __restrict__ R, const double* __restrict__ ions,
                                         int iat, const double pos[3], const double vd[3], double tau, MoveLDS& sh,
                                         float* __restrict__ x)
{
  const int t = thread_index(), N = s.N, Np = s.Npad, ni = s.nion;
  double* de = sh.red;
  double* di = sh.red + 1024;
  __syncthreads();
  for (int j = t; j < N; j += WG)
  {
    double dr[3];
    const double r = min_image(s, R[j], R[Np + j], R[2 * Np + j], pos, dr);
    de[j]          = j == iat ? INFINITY : r;
  }
  for (int j = t; j < ni; j += WG)
  {
    double dr[3];
    di[j] = min_image(s, ions[j], ions[ni + j], ions[2 * ni + j], pos, dr);
  }
  select_nearest(de, N, N_EL_NB, sh.sel, reinterpret_cast<int*>(sh.row));
  select_nearest(di, ni, N_ION_NB, sh.sel + N_EL_NB, reinterpret_cast<int*>(sh.vec));
  if (t < MLP_IN)
  {
    double val = 0.0;
    if (t < 4 * N_ION_NB)
    {
      const int j = sh.sel[N_EL_NB + (t >> 2)], c = t & 3;
      if (j >= 0)
      {
        double dr[3];
        const double r = min_image(s, ions[j], ions[ni + j], ions[2 * ni + j], pos, dr);
        val            = c < 3 ? dr[c] : r;
      }
    }
    else if (t < 4 * N_ION_NB + 3 * N_EL_NB)
    {
      const int u = t - 4 * N_ION_NB, j = sh.sel[u / 3];
      if (j >= 0)
      {
        double dr[3];
        min_image(s, R[j], R[Np + j], R[2 * Np + j], pos, dr);
        val = dr[u % 3];
      }
    }
    else if (t < MLP_IN - 1)
      val = vd[t - (MLP_IN - 4)];
    else
      val = tau;
    x[t] = (float)val;
  }
  __syncthreads();
}

// total gradient at the current position of iat (determinant part from sh.bc[BC_GRAD]), summed in
// the component order of the plain proposal
__device__ inline void grad_now(const SysC& s, int iat, const double* __restrict__ jas, const MoveLDS& sh, double g[3])
{
  const int Np = s.Npad;
  for (int q = 0; q < 3; q++)
    g[q] = 0.0;
  if (s.has_det)
    for (int q = 0; q < 3; q++)
      g[q] += sh.bc[BC_GRAD + q];
  if (s.has_j1)
    for (int q = 0; q < 3; q++)
      g[q] += jas[(J_G1X + q) * Np + iat];
  if (s.has_j2)
    for (int q = 0; q < 3; q++)
      g[q] += jas[(J_DUX + q) * Np + iat];
}

// Inverse row (to st.invRow), gradient and drift at the current position, features at r (rows
// [0, nw) of la.X).
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_learn_pre(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a,
                                                  LearnArgs la)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, Np = s.Npad;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double ainv_r[MAXM], inv[MAXM];
  if (s.has_det)
  {
    phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
    det_grad<DT, MAXM>(n, ld, r, v.dpsi, inv, sh);
    double* ir = st.invRow + (size_t)w * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = thread_index() + WG * m;
      if (o < n)
        ir[o] = inv[m];
    }
  }
  __syncthreads();
  if (thread_index() == 0)
  {
    double* ls = la.ls + (size_t)w * LS_SIZE;
    double gn[3], dv[3];
    grad_now(s, iat, v.jas, sh, gn);
    unr_drift(a.tau, gn, dv);
    for (int q = 0; q < 3; q++)
    {
      ls[LS_R + q]    = v.R[q * Np + iat];
      ls[LS_V + q]    = dv[q];
      ls[LS_GNOW + q] = gn[q];
      sh.bc[BC_POS + q]  = ls[LS_R + q];
      sh.bc[BC_MOVE + q] = dv[q];
    }
  }
  __syncthreads();
  const double pos[3] = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
  const double vd[3]  = {sh.bc[BC_MOVE], sh.bc[BC_MOVE + 1], sh.bc[BC_MOVE + 2]};
  proposal_features(s, v.R, st.ions, iat, pos, vd, a.tau, sh, la.X + (size_t)w * MLP_IN);
}

// Proposal r' = r + v + mu + sqrt(tau) e^s xi from the network output at r; ratio and gradient at r'
// (phase C with the stored inverse row); drift and features at r' (rows [nw, 2 nw) of la.X).
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_learn_prop(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a,
                                                   LearnArgs la)
{
  __shared__ MoveLDS sh;
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld;
  const int iat = a.iat0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double* ls       = la.ls + (size_t)w * LS_SIZE;
  double* mv       = st.mv + (size_t)w * MV_SIZE;
  if (thread_index() == 0)
  {
    const float* of  = la.out + (size_t)w * MLP_OUT;
    const double* gz = a.gauss + ((size_t)iat * a.nw + w) * 3;
    const double sig = sqrt(a.tau) * exp((double)of[3]);
    double xx        = 0;
    for (int q = 0; q < 3; q++)
    {
      const double xi    = gz[q];
      const double move  = (ls[LS_V + q] + (double)of[q]) + xi * sig;
      const double rp    = ls[LS_R + q] + move;
      ls[LS_RP + q]      = rp;
      ls[LS_XI + q]      = xi;
      sh.bc[BC_POS + q]  = rp;
      xx += xi * xi;
    }
    ls[LS_LQF] = -0.5 * xx - 3.0 * (double)of[3];
  }
  __syncthreads();
  double inv[MAXM], phv[MAXM][5];
  const double* ir = st.invRow + (size_t)w * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = thread_index() + WG * m;
    inv[m]      = (s.has_det && o < n) ? ir[o] : 0.0;
  }
  phaseC<ST, MAXM>(s, g, iat, n, v.R, st.ions, coefs, inv, phv, sh);
  if (thread_index() == 0)
  {
    double ratio, gnew[3], vr[3];
    combine_ratio(s, iat, v.jas, sh, ratio, gnew);
    nan_guard(a, w, iat, ratio * ratio, gnew[0] * gnew[0] + gnew[1] * gnew[1] + gnew[2] * gnew[2],
              ls[LS_GNOW] * ls[LS_GNOW] + ls[LS_GNOW + 1] * ls[LS_GNOW + 1] + ls[LS_GNOW + 2] * ls[LS_GNOW + 2]);
    unr_drift(a.tau, gnew, vr);
    mv[MV_RATIO] = ratio;
    for (int q = 0; q < 3; q++)
    {
      mv[MV_POS + q]  = ls[LS_RP + q];
      mv[MV_GNOW + q] = ls[LS_GNOW + q];
      ls[LS_GNEW + q] = gnew[q];
      ls[LS_VR + q]   = vr[q];
      sh.bc[BC_MOVE + q] = vr[q];
    }
    for (int q = 0; q < 13; q++)
      mv[MV_C + q] = sh.bc[BC_C + q];
  }
  double* phi = st.phi + (size_t)w * 5 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = thread_index() + WG * m;
    if (o < n)
      for (int q = 0; q < 5; q++)
        phi[q * ld + o] = phv[m][q];
  }
  __syncthreads();
  const double pos[3] = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
  const double vd[3]  = {sh.bc[BC_MOVE], sh.bc[BC_MOVE + 1], sh.bc[BC_MOVE + 2]};
  proposal_features(s, v.R, st.ions, iat, pos, vd, a.tau, sh, la.X + ((size_t)a.nw + w) * MLP_IN);
}

// Metropolis-Hastings with the asymmetric-proposal correction (the walker's own uniform stream),
// the training objective f = min(1, e^L) |r' - r|^2 and its gradient with respect to both network
// outputs (loss -mean f), then accept / reject (phase E).
template<typename DT, int MAXM>
__global__ __launch_bounds__(WG) void k_learn_accept(SysC s, State st, MoveArgs a, LearnArgs la)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  const double* mv = st.mv + (size_t)w * MV_SIZE;
  const double* ls = la.ls + (size_t)w * LS_SIZE;
  if (thread_index() == 0)
  {
    const float* of = la.out + (size_t)w * MLP_OUT;
    const float* orr = la.out + ((size_t)a.nw + w) * MLP_OUT;
    const double tau = a.tau, ratio = mv[MV_RATIO];
    const double s_r = orr[3], sig_r2 = tau * exp(2.0 * s_r);
    double e[3], D[3], ee = 0, D2 = 0;
    for (int q = 0; q < 3; q++)
    {
      e[q] = ls[LS_R + q] - ((ls[LS_RP + q] + ls[LS_VR + q]) + (double)orr[q]);
      D[q] = ls[LS_RP + q] - ls[LS_R + q];
      ee += e[q] * e[q];
      D2 += D[q] * D[q];
    }
    const double lqf = ls[LS_LQF];
    const double lqr = -ee / (2.0 * sig_r2) - 3.0 * s_r;
    const double pr2 = ratio * ratio;
    const bool flag  = pr2 >= DBL_EPS;
    const double prob = pr2 * exp(lqr - lqf);
    int ucount        = a.ubase[(size_t)w * (N + 1) + iat];
    bool acc          = false;
    if (flag)
    {
      const double u = a.unif[(size_t)w * N + ucount];
      ucount++;
      acc = u < prob;
    }
    a.ubase[(size_t)w * (N + 1) + iat + 1] = ucount;
    v.stats[acc ? 0 : 1] += 1;
    sh.bc[BC_ACC] = acc ? 1.0 : 0.0;
    // objective and its output gradients
    const double A     = prob < 1.0 ? prob : 1.0;
    const bool below   = prob < 1.0;
    const double f     = A * D2;
    const double sig_f = sqrt(tau) * exp((double)of[3]);
    double gf[4], gr[4], dfD[3], dot = 0;
    for (int q = 0; q < 3; q++)
    {
      const double dL = 2.0 * ls[LS_GNEW + q] + e[q] / sig_r2;
      dfD[q]          = 2.0 * A * D[q] + (below ? A * D2 * dL : 0.0);
      gf[q]           = dfD[q];
      gr[q]           = below ? A * D2 * e[q] / sig_r2 : 0.0;
      dot += dfD[q] * sig_f * ls[LS_XI + q];
    }
    gf[3] = dot + (below ? 3.0 * A * D2 : 0.0);
    gr[3] = below ? A * D2 * (ee / sig_r2 - 3.0) : 0.0;
    if (la.dout)
      for (int q = 0; q < 4; q++)
      {
        la.dout[(size_t)w * MLP_OUT + q]          = (float)(-gf[q] / a.nw);
        la.dout[((size_t)a.nw + w) * MLP_OUT + q] = (float)(-gr[q] / a.nw);
      }
    if (la.fsum)
    {
      la.fsum[(size_t)w * 2 + 0] += f;
      la.fsum[(size_t)w * 2 + 1] += A;
    }
    if (a.trace)
    {
      double* tr = a.trace + ((size_t)iat * a.nw + w) * TRACE_FIELDS;
      tr[0]      = ratio;
      tr[1]      = acc ? 1.0 : 0.0;
      for (int q = 0; q < 3; q++)
      {
        tr[2 + q] = ls[LS_RP + q];
        tr[5 + q] = ls[LS_GNOW + q];
      }
    }
    if (la.ltrace)
    {
      double* lt = la.ltrace + ((size_t)iat * a.nw + w) * LTRACE_FIELDS;
      for (int q = 0; q < 4; q++)
      {
        lt[q]     = of[q];
        lt[4 + q] = orr[q];
      }
      lt[8]  = lqf;
      lt[9]  = lqr;
      lt[10] = prob;
      lt[11] = f;
      for (int q = 0; q < 3; q++)
        lt[12 + q] = ls[LS_VR + q];
    }
    for (int q = 0; q < 3; q++)
      sh.bc[BC_POS + q] = mv[MV_POS + q];
    for (int q = 0; q < 13; q++)
      sh.bc[BC_C + q] = mv[MV_C + q];
  }
  if (k > 1 && thread_index() < d)
    sh.y[thread_index()] = v.Binv[d * k + thread_index()];
  __syncthreads();
  double ainv_r[MAXM], phv[MAXM][5];
  const double* phi = st.phi + (size_t)w * 5 * ld;
#pragma unroll
  for (int m = 0; m < MAXM; m++)
  {
    const int o = thread_index() + WG * m;
    ainv_r[m]   = o < n ? (double)v.Ainv[(size_t)r * ld + o] : 0.0;
    for (int q = 0; q < 5; q++)
      phv[m][q] = o < n ? phi[q * ld + o] : 0.0;
  }
  __syncthreads();
  phaseE<DT, MAXM>(s, g, iat, n, ld, k, d, r, v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.dl, v.jas, v.logdet, ainv_r, phv,
                   sh, sB, true);
}

// The frozen network on one sample per workgroup: x[68] (FP32, LDS) -> out[4].  The 16-row
// matrix-core tiles of the batched forward kernel carry the sample in row 0 (rows 1..15 zero), so
// row 0 of every layer is computed exactly as there.  Activations in sh.red; all threads, ends with
// a barrier.
__device__ inline void mlp_single(const MlpParams& P, const float* x, MoveLDS& sh, float* out)
{
  uint16_t* sA = reinterpret_cast<uint16_t*>(sh.red);
  uint16_t* sB = sA + 16 * MLP_LDA;
  const int t  = thread_index();
  for (int e = t; e < 16 * MLP_LDA; e += WG)
    sA[e] = (e < MLP_IN) ? bf16_bits(x[e]) : (uint16_t)0;
  __syncthreads();
  const float* bias = P.master;
  mlp_layer<MLP_K1>(sA, P.w1b, bias + MLP_OFF_B1, sB, nullptr, 1);
  __syncthreads();
  mlp_layer<MLP_HID>(sB, P.w2b, bias + MLP_OFF_B2, sA, nullptr, 1);
  __syncthreads();
  mlp_layer<MLP_HID>(sA, P.w3b, bias + MLP_OFF_B3, sB, nullptr, 1);
  __syncthreads();
  if ((t >> 6) == 0)
    mlp_out_layer(sB, P, out, MLP_OUT, 1);
  __syncthreads();
}

// One learned move with a frozen network, start to end in one workgroup per walker: the same
// steps and arithmetic as k_learn_pre, k_mlp_forward, k_learn_prop, k_mlp_forward and
// k_learn_accept (without the objective's gradient), with the inverse row, the proposed orbitals
// and the move scalars kept in registers and LDS instead of global memory.
template<typename ST, typename DT, int MAXM>
__global__ __launch_bounds__(WG, 4) void k_learn_move(SysC s, State st, const ST* __restrict__ coefs, MoveArgs a,
                                                   LearnArgs la, MlpParams P)
{
  __shared__ MoveLDS sh;
  extern __shared__ double sB_raw[];
  DT* sB = reinterpret_cast<DT*>(sB_raw);
  stage_jastrow(s, sh);
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const int g = a.g, n = s.n[g], ld = s.ld, k = s.k, N = s.N, Np = s.Npad;
  const int iat = a.iat0, r = iat - s.first[g], d = a.d0;
  // move scalars in sh.bc beyond the plain move's fields
  constexpr int L_R = 32, L_V = 35, L_XI = 38, L_LQF = 41, L_GNEW = 42, L_VR = 45, L_PF = 48;
  static_assert(L_VR + 3 <= L_PF && L_PF + 15 <= 64, "broadcast slots");
  WalkerView<DT> v = walker_view<DT>(s, st, g, w);
  double ainv_r[MAXM], inv[MAXM], phv[MAXM][5];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    ainv_r[m] = inv[m] = 0.0;
  // the scalars thread 0 needs later, loaded now (published by the barrier after phase A)
  int ucount = a.ubase[(size_t)w * (N + 1) + iat];
  prefetch_move(s, iat, w, ucount, a, v.R, v.jas, sh, L_PF);
  const double* pf = sh.bc + L_PF;
  if (s.has_det)
  {
    phaseA<DT, MAXM>(s, n, ld, k, d, r, v.Ainv, v.U, v.V, v.Binv, sh, sB, ainv_r, inv);
    det_grad<DT, MAXM>(n, ld, r, v.dpsi, inv, sh);
  }
  __syncthreads();
  if (thread_index() == 0)
  {
    double gn[3] = {0, 0, 0}, dv[3];
    if (s.has_det) // grad_now from the prefetched Jastrow gradients
      for (int q = 0; q < 3; q++)
        gn[q] += sh.bc[BC_GRAD + q];
    if (s.has_j1)
      for (int q = 0; q < 3; q++)
        gn[q] += pf[PF_G1 + q];
    if (s.has_j2)
      for (int q = 0; q < 3; q++)
        gn[q] += pf[PF_G2 + q];
    unr_drift(a.tau, gn, dv);
    for (int q = 0; q < 3; q++)
    {
      sh.bc[L_R + q]     = pf[PF_R + q];
      sh.bc[L_V + q]     = dv[q];
      sh.bc[BC_GNOW + q] = gn[q];
    }
  }
  __syncthreads();
  float* xf = reinterpret_cast<float*>(sh.Y); // features (68 FP32)
  float* mo = reinterpret_cast<float*>(sh.p); // network outputs at r [0, 4) and at r' [4, 8)
  {
    const double pos[3] = {sh.bc[L_R], sh.bc[L_R + 1], sh.bc[L_R + 2]};
    const double vd[3]  = {sh.bc[L_V], sh.bc[L_V + 1], sh.bc[L_V + 2]};
    proposal_features(s, v.R, st.ions, iat, pos, vd, a.tau, sh, xf);
  }
  mlp_single(P, xf, sh, mo);
  if (thread_index() == 0)
  {
    const double* gz = pf + PF_GAUSS;
    const double sig = sqrt(a.tau) * exp((double)mo[3]);
    double xx        = 0;
    for (int q = 0; q < 3; q++)
    {
      const double xi   = gz[q];
      const double move = (sh.bc[L_V + q] + (double)mo[q]) + xi * sig;
      sh.bc[BC_POS + q] = sh.bc[L_R + q] + move;
      sh.bc[L_XI + q]   = xi;
      xx += xi * xi;
    }
    sh.bc[L_LQF] = -0.5 * xx - 3.0 * (double)mo[3];
  }
  __syncthreads();
  phaseC<ST, MAXM>(s, g, iat, n, v.R, st.ions, coefs, inv, phv, sh);
  if (thread_index() == 0)
  {
    double ratio, gnew[3], vr[3];
    combine_ratio_v(s, pf[PF_VAT], pf[PF_UAT], sh, ratio, gnew);
    nan_guard(a, w, iat, ratio * ratio, gnew[0] * gnew[0] + gnew[1] * gnew[1] + gnew[2] * gnew[2],
              gnow_norm2(sh.bc));
    unr_drift(a.tau, gnew, vr);
    sh.bc[BC_RATIO] = ratio;
    for (int q = 0; q < 3; q++)
    {
      sh.bc[L_GNEW + q]  = gnew[q];
      sh.bc[L_VR + q]    = vr[q];
      sh.bc[BC_MOVE + q] = vr[q];
    }
  }
  __syncthreads();
  {
    const double pos[3] = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
    const double vd[3]  = {sh.bc[BC_MOVE], sh.bc[BC_MOVE + 1], sh.bc[BC_MOVE + 2]};
    proposal_features(s, v.R, st.ions, iat, pos, vd, a.tau, sh, xf);
  }
  mlp_single(P, xf, sh, mo + MLP_OUT);
  if (thread_index() == 0)
  {
    const float* of  = mo;
    const float* orr = mo + MLP_OUT;
    const double tau = a.tau, ratio = sh.bc[BC_RATIO];
    const double s_r = orr[3], sig_r2 = tau * exp(2.0 * s_r);
    double e[3], D[3], ee = 0, D2 = 0;
    for (int q = 0; q < 3; q++)
    {
      e[q] = sh.bc[L_R + q] - ((sh.bc[BC_POS + q] + sh.bc[L_VR + q]) + (double)orr[q]);
      D[q] = sh.bc[BC_POS + q] - sh.bc[L_R + q];
      ee += e[q] * e[q];
      D2 += D[q] * D[q];
    }
    const double lqf  = sh.bc[L_LQF];
    const double lqr  = -ee / (2.0 * sig_r2) - 3.0 * s_r;
    const double pr2  = ratio * ratio;
    const bool flag   = pr2 >= DBL_EPS;
    const double prob = pr2 * exp(lqr - lqf);
    bool acc          = false;
    if (flag)
    {
      const double u = pf[PF_UNIF]; // a.unif[w N + ucount]
      ucount++;
      acc = u < prob;
    }
    a.ubase[(size_t)w * (N + 1) + iat + 1] = ucount;
    v.stats[acc ? 0 : 1] += 1;
    sh.bc[BC_ACC]    = acc ? 1.0 : 0.0;
    const double A   = prob < 1.0 ? prob : 1.0;
    const double f   = A * D2;
    if (la.fsum)
    {
      la.fsum[(size_t)w * 2 + 0] += f;
      la.fsum[(size_t)w * 2 + 1] += A;
    }
    if (a.trace)
    {
      double* tr = a.trace + ((size_t)iat * a.nw + w) * TRACE_FIELDS;
      tr[0]      = ratio;
      tr[1]      = acc ? 1.0 : 0.0;
      for (int q = 0; q < 3; q++)
      {
        tr[2 + q] = sh.bc[BC_POS + q];
        tr[5 + q] = sh.bc[BC_GNOW + q];
      }
    }
    if (la.ltrace)
    {
      double* lt = la.ltrace + ((size_t)iat * a.nw + w) * LTRACE_FIELDS;
      for (int q = 0; q < 4; q++)
      {
        lt[q]     = of[q];
        lt[4 + q] = orr[q];
      }
      lt[8]  = lqf;
      lt[9]  = lqr;
      lt[10] = prob;
      lt[11] = f;
      for (int q = 0; q < 3; q++)
        lt[12 + q] = sh.bc[L_VR + q];
    }
  }
  __syncthreads();
  phaseE<DT, MAXM>(s, g, iat, n, ld, k, d, r, v.R, v.Ainv, v.dpsi, v.U, v.V, v.Binv, v.dl, v.jas, v.logdet, ainv_r, phv,
                   sh, sB, false);
}

// features of electron a.iat0 at the position LS_R with the drift LS_V of each walker (test hook)
__global__ __launch_bounds__(WG) void k_learn_features(SysC s, State st, MoveArgs a, LearnArgs la)
{
  __shared__ MoveLDS sh;
  const int w = blockIdx.x;
  if (w >= a.nw)
    return;
  const double* ls    = la.ls + (size_t)w * LS_SIZE;
  const double pos[3] = {ls[LS_R], ls[LS_R + 1], ls[LS_R + 2]};
  const double vd[3]  = {ls[LS_V], ls[LS_V + 1], ls[LS_V + 2]};
  proposal_features(s, st.R + (size_t)w * 3 * s.Npad, st.ions, a.iat0, pos, vd, a.tau, sh, la.X + (size_t)w * MLP_IN);
}

template<typename ST, typename DT>
static hipError_t launch_learn_t(int which, const SysC& s, const State& st, const ST* coefs, const MoveArgs& a,
                                 const LearnArgs& la, hipStream_t stream, const MlpParams* P)
{
  const int maxn   = move_maxm(s, a.g) * WG;
  if (maxn == 0)
    return hipErrorInvalidValue;
  const dim3 grid(a.nw), block(WG);
  const size_t dyn = binv_lds_bytes(s.k, sizeof(DT));
  if (which == LK_FEATURES)
  {
    k_learn_features<<<grid, block, 0, stream>>>(s, st, a, la);
    return hipGetLastError();
  }
  if (maxn <= WG)
  {
    switch (which)
    {
    case LK_PRE: k_learn_pre<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a, la); break;
    case LK_PROP: k_learn_prop<ST, DT, 1><<<grid, block, 0, stream>>>(s, st, coefs, a, la); break;
    case LK_ACCEPT: k_learn_accept<DT, 1><<<grid, block, dyn, stream>>>(s, st, a, la); break;
    case LK_FUSED: k_learn_move<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a, la, *P); break;
    default: return hipErrorInvalidValue;
    }
  }
  else if (maxn <= 2 * WG)
  {
    switch (which)
    {
    case LK_PRE: k_learn_pre<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a, la); break;
    case LK_PROP: k_learn_prop<ST, DT, 2><<<grid, block, 0, stream>>>(s, st, coefs, a, la); break;
    case LK_ACCEPT: k_learn_accept<DT, 2><<<grid, block, dyn, stream>>>(s, st, a, la); break;
    case LK_FUSED: k_learn_move<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a, la, *P); break;
    default: return hipErrorInvalidValue;
    }
  }
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_learn(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                        const LearnArgs& la, hipStream_t stream, const MlpParams* P)
{
  if (which == LK_FUSED && !P)
    return hipErrorInvalidValue;
  return dispatch_prec(s.prec, [&](auto stv, auto dtv) {
    using ST = decltype(stv);
    using DT = decltype(dtv);
    return launch_learn_t<ST, DT>(which, s, st, static_cast<const ST*>(coefs), a, la, stream, P);
  });
}

//------------------------------------------------------------------------------------------------
// launchers
template<typename ST, typename DT>
static hipError_t launch_move_t(int which, const SysC& s, const State& st, const ST* coefs, const MoveArgs& a,
                                hipStream_t stream)
{
  const int maxn = move_maxm(s, a.g) * WG;
  if (maxn == 0)
    return hipErrorInvalidValue;
  if ((a.w0 != 0 || a.wn != 0) && which != MK_SEGMENT && which != MK_MOVE_ONE)
    return hipErrorInvalidValue; // only the fused kernels take a walker partition
  if (a.w0 < 0 || a.wn < 0 || a.w0 + a.wn > a.nw)
    return hipErrorInvalidValue;
  dim3 grid(a.wn > 0 ? a.wn : a.nw), block(WG);
  const size_t dyn = binv_lds_bytes(s.k, sizeof(DT));
  if (maxn <= WG)
  {
    switch (which)
    {
    case MK_SEGMENT: k_sweep_segment<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_MOVE_ONE: k_move_one<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_PROPOSE: k_propose<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_ACCEPT_CROWD: k_accept_crowd<DT, 1><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_GRAD: k_api_eval_grad<DT, 1><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_RATIO: k_api_ratio_grad<ST, DT, 1><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_API_ACCEPT: k_api_accept<DT, 1><<<grid, block, dyn, stream>>>(s, st, a); break;
    default: return hipErrorInvalidValue;
    }
  }
  else if (maxn <= 2 * WG)
  {
    switch (which)
    {
    case MK_SEGMENT: k_sweep_segment<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_MOVE_ONE: k_move_one<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_PROPOSE: k_propose<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_ACCEPT_CROWD: k_accept_crowd<DT, 2><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_GRAD: k_api_eval_grad<DT, 2><<<grid, block, dyn, stream>>>(s, st, a); break;
    case MK_API_RATIO: k_api_ratio_grad<ST, DT, 2><<<grid, block, dyn, stream>>>(s, st, coefs, a); break;
    case MK_API_ACCEPT: k_api_accept<DT, 2><<<grid, block, dyn, stream>>>(s, st, a); break;
    default: return hipErrorInvalidValue;
    }
  }
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_move(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                       hipStream_t stream)
{
  return dispatch_prec(s.prec, [&](auto stv, auto dtv) {
    using ST = decltype(stv);
    using DT = decltype(dtv);
    return launch_move_t<ST, DT>(which, s, st, static_cast<const ST*>(coefs), a, stream);
  });
}

hipError_t launch_make_move(const SysC& s, const State& st, int iat, int nw, const double* displ, hipStream_t stream)
{
  k_api_make_move<<<(nw + 255) / 256, 256, 0, stream>>>(s, st, iat, nw, displ);
  return hipGetLastError();
}

} // namespace qmi
