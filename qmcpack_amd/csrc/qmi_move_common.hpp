// Pieces of the per-walker move kernels shared by the real and the complex-orbital forms:
// LDS layouts, the Jastrow ratio sums and accept updates, and the drift-diffusion proposal.
#pragma once
#include "qmi_kernels.hpp"

namespace qmi
{

struct MoveLDS
{
  double row[MAX_N]; // Ainv row / SM-1 temp vector
  double vec[MAX_N]; // proposed orbital values
  double p[MAX_K];
  double y[MAX_K]; // row d of Binv (from phase A), zero beyond d
  double Y[MAX_K]; // new column d of Binv
  alignas(16) double red[9 * WG]; // block_sum / delayed_gemv scratch (up to 9 values per thread)
  double bc[64];      // broadcast scalars
  double jc[3][MAX_JCOEF]; // Jastrow functor tables: J1, J2 like spins, J2 unlike spins
  SplineLDS spl;      // uniform spline factors of the proposed position
  int sel[N_EL_NB + N_ION_NB]; // learned proposal: selected neighbours (electrons, then ions)
};

// Copy the functor tables to LDS (published by the caller's next barrier).
template<typename SH>
__device__ inline void stage_jastrow(const SysC& s, SH& sh)
{
  for (int e = thread_index(); e < 3 * MAX_JCOEF; e += WG)
  {
    const int f = e / MAX_JCOEF, i = e - f * MAX_JCOEF;
    sh.jc[f][i] = f == 0 ? s.j1.c[i] : s.j2[f - 1].c[i];
  }
}

// Electrons per thread when every electron of a walker is visited once: a kernel instantiated for
// MAXM orbitals per thread runs systems with N <= 2 MAXM WG electrons (move_maxm below), so
// 2 MAXM electrons per thread cover them all.
constexpr int JPT = (2 * MAX_N + WG - 1) / WG;

// orbitals per thread (1 or 2; 0 if unsupported) of the move kernels for the moving spin group g:
// enough for its n orbitals and for 2 MAXM electrons per thread in the Jastrow loops
__host__ __device__ inline int move_maxm(const SysC& s, int g)
{
  const int need = s.n[g] > (s.N + 1) / 2 ? s.n[g] : (s.N + 1) / 2;
  return need <= WG ? 1 : need <= 2 * WG ? 2 : 0;
}

// The leading d x d block of Binv is staged in dynamically sized LDS with row stride k + 1
// (column reads by consecutive threads then fall in different banks).
__host__ __device__ constexpr size_t binv_lds_bytes(int k, size_t esize) { return (size_t)k * (k + 1) * esize; }

template<typename DT>
__device__ inline void stage_binv(const DT* __restrict__ Binv, int k, int d, DT* sB)
{
  for (int e = thread_index(); e < d * d; e += WG)
  {
    const int j = e / d, i = e - j * d;
    sB[j * (k + 1) + i] = Binv[j * k + i];
  }
}

enum BcField
{
  BC_GRAD = 0,  // 3 determinant gradient at the current position
  BC_POS  = 3,  // 3 proposed position
  BC_MOVE = 6,  // 3 displacement
  BC_LOGGF = 9,
  BC_C    = 10, // 13 phase-C sums: J1 (u, g3, lap), J2 (u, g3), det (ratio, g3)
  BC_ACC  = 23,
  BC_RATIO = 24,
  BC_D2U  = 25,
  BC_OLD  = 26, // 2 crowd prefix counts
  BC_GNOW = 29  // 3 gradient at the current position
};

// squared norm of the gradient at the current position (NaN guard)
__device__ inline double gnow_norm2(const double* bc)
{
  return bc[BC_GNOW] * bc[BC_GNOW] + bc[BC_GNOW + 1] * bc[BC_GNOW + 1] + bc[BC_GNOW + 2] * bc[BC_GNOW + 2];
}


// J1/J2 ratio sums at the proposed position pos -> sh.bc[BC_C..BC_C+9): J1 (u, g3, lap), J2 (u, g3)
// (J1OrbitalSoA.h:136-186 computeU3; TwoBodyJastrow.cpp:420-446).  All threads; ends with a barrier.
// positions of this thread's electrons t + u WG (zero past N)
template<int NJ>
__device__ inline void load_electrons(const SysC& s, const double* __restrict__ R, double (&px)[NJ], double (&py)[NJ],
                                      double (&pz)[NJ])
{
  const int t = thread_index(), N = s.N, Np = s.Npad;
#pragma unroll
  for (int u = 0; u < NJ; u++)
  {
    const int j   = t + u * WG;
    const bool ok = j < N;
    px[u]         = ok ? R[j] : 0.0;
    py[u]         = ok ? R[Np + j] : 0.0;
    pz[u]         = ok ? R[2 * Np + j] : 0.0;
  }
}

// the same sums with this thread's electron positions already in registers
template<int NJ, typename SH>
__device__ inline void jastrow_ratio_sums_p(const SysC& s, int g, int iat, const double (&px)[NJ],
                                            const double (&py)[NJ], const double (&pz)[NJ],
                                            const double* __restrict__ ions, const double pos[3], SH& sh);

template<int NJ, typename SH>
__device__ inline void jastrow_ratio_sums(const SysC& s, int g, int iat, const double* __restrict__ R,
                                          const double* __restrict__ ions, const double pos[3], SH& sh)
{
  double px[NJ], py[NJ], pz[NJ];
  if (s.has_j2)
    load_electrons<NJ>(s, R, px, py, pz);
  jastrow_ratio_sums_p<NJ>(s, g, iat, px, py, pz, ions, pos, sh);
}

template<int NJ, typename SH>
__device__ inline void jastrow_ratio_sums_p(const SysC& s, int g, int iat, const double (&px)[NJ],
                                            const double (&py)[NJ], const double (&pz)[NJ],
                                            const double* __restrict__ ions, const double pos[3], SH& sh)
{
  const int t = thread_index();
  const int N = s.N;
  double acc[9];
#pragma unroll
  for (int i = 0; i < 9; i++)
    acc[i] = 0;
  if (s.has_j1)
    for (int j = t; j < s.nion; j += WG)
    {
      double dr[3];
      const double r = min_image(s, ions[j], ions[s.nion + j], ions[2 * s.nion + j], pos, dr);
      if (r < s.j1.rcut)
      {
        double u, du, d2u;
        functor_vgl_c(sh.jc[0], s.j1.DeltaRInv, s.j1.max_index, r, u, du, d2u);
        acc[0] += u;
        acc[1] += du * dr[0];
        acc[2] += du * dr[1];
        acc[3] += du * dr[2];
        acc[4] += d2u + 2.0 * du;
      }
    }
  if (s.has_j2)
  {
#pragma unroll
    for (int u = 0; u < NJ; u++)
    {
      const int j = t + u * WG;
      if (j >= N || j == iat)
        continue;
      double dr[3];
      const double r    = min_image(s, px[u], py[u], pz[u], pos, dr);
      const int fi      = ((j < s.nup) ? 0 : 1) == g ? 0 : 1;
      const FunctorC& f = s.j2[fi];
      if (r < f.rcut)
      {
        double u2, du, d2u;
        functor_vgl_c(sh.jc[1 + fi], f.DeltaRInv, f.max_index, r, u2, du, d2u);
        acc[5] += u2;
        acc[6] += du * dr[0];
        acc[7] += du * dr[1];
        acc[8] += du * dr[2];
      }
    }
  }
  block_sum<9>(acc, sh.red, sh.bc + BC_C);
}

// Accepted move, Jastrow side and the particle position (J1OrbitalSoA.h:456-468;
// TwoBodyJastrow::acceptMove, TwoBodyJastrow.cpp:581-628; ParticleSet::acceptMove).  Reads the
// phase-C sums at sh.bc[BC_C]; all threads.
template<int NJ, typename SH>
__device__ inline void jastrow_accept(const SysC& s, int g, int iat, double* __restrict__ R,
                                      double* __restrict__ jas, SH& sh)
{
  const int t = thread_index();
  const int N = s.N, Np = s.Npad;
  const double* C = sh.bc + BC_C;
  if (s.has_j1 && t == 0)
  {
    jas[J_VAT * Np + iat]      = C[0];
    jas[(J_G1X + 0) * Np + iat] = C[1];
    jas[(J_G1X + 1) * Np + iat] = C[2];
    jas[(J_G1X + 2) * Np + iat] = C[3];
    jas[J_LAP1 * Np + iat]     = C[4];
  }
  if (s.has_j2)
  {
    const double pos[3]  = {sh.bc[BC_POS], sh.bc[BC_POS + 1], sh.bc[BC_POS + 2]};
    const double rold[3] = {R[iat], R[Np + iat], R[2 * Np + iat]};
    double cd2[1]        = {0.0};
    // positions and the five accumulator rows of every electron of this thread first
    double px[NJ], py[NJ], pz[NJ], ja[NJ][5];
#pragma unroll
    for (int u = 0; u < NJ; u++)
    {
      const int j   = t + u * WG;
      const bool ok = j < N;
      px[u]         = ok ? R[j] : 0.0;
      py[u]         = ok ? R[Np + j] : 0.0;
      pz[u]         = ok ? R[2 * Np + j] : 0.0;
      ja[u][0]      = ok ? jas[J_UAT * Np + j] : 0.0;
      ja[u][1]      = ok ? jas[J_D2U * Np + j] : 0.0;
#pragma unroll
      for (int q = 0; q < 3; q++)
        ja[u][2 + q] = ok ? jas[(J_DUX + q) * Np + j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NJ; u++)
    {
      const int j = t + u * WG;
      if (j >= N || j == iat)
        continue;
      double drn[3], dro[3];
      const double rn   = min_image(s, px[u], py[u], pz[u], pos, drn);
      const double ro   = min_image(s, px[u], py[u], pz[u], rold, dro);
      const int fi      = ((j < s.nup) ? 0 : 1) == g ? 0 : 1;
      const FunctorC& f = s.j2[fi];
      const double* fc  = sh.jc[1 + fi];
      double un = 0, dun = 0, d2un = 0, uo = 0, duo = 0, d2uo = 0;
      if (rn < f.rcut)
        functor_vgl_c(fc, f.DeltaRInv, f.max_index, rn, un, dun, d2un);
      if (ro < f.rcut)
        functor_vgl_c(fc, f.DeltaRInv, f.max_index, ro, uo, duo, d2uo);
      const double newl = d2un + 2.0 * dun;
      jas[J_UAT * Np + j] = ja[u][0] + (un - uo);
      jas[J_D2U * Np + j] = ja[u][1] + (d2uo + 2.0 * duo - newl);
      cd2[0] -= newl;
#pragma unroll
      for (int q = 0; q < 3; q++)
      {
        const double newg          = dun * drn[q];
        jas[(J_DUX + q) * Np + j] = ja[u][2 + q] - (newg - duo * dro[q]);
      }
    }
    block_sum<1>(cd2, sh.red, sh.bc + BC_D2U);
    if (t == 0)
    {
      jas[J_UAT * Np + iat]       = C[5];
      jas[(J_DUX + 0) * Np + iat] = C[6];
      jas[(J_DUX + 1) * Np + iat] = C[7];
      jas[(J_DUX + 2) * Np + iat] = C[8];
      jas[J_D2U * Np + iat]       = sh.bc[BC_D2U];
    }
  }
  __syncthreads();
  if (t < 3)
    R[t * Np + iat] = sh.bc[BC_POS + t];
}

//------------------------------------------------------------------------------------------------
template<typename DT>
struct WalkerView
{
  double* R;
  DT *Ainv, *dpsi, *U, *V, *Binv;
  double *jas, *logdet, *stats;
  int* dl;
};

template<typename DT>
__device__ inline WalkerView<DT> walker_view(const SysC& s, const State& st, int g, int w)
{
  WalkerView<DT> v;
  const int n = s.n[g], ld = s.ld, k = s.k;
  v.R      = st.R + (size_t)w * 3 * s.Npad;
  v.Ainv   = det_ptr<DT>(st.Ainv[g]) + (size_t)w * n * ld;
  v.dpsi   = det_ptr<DT>(st.dpsi[g]) + (size_t)w * n * 4 * ld;
  v.U      = det_ptr<DT>(st.U[g]) + (size_t)w * ld * k;
  v.V      = det_ptr<DT>(st.V[g]) + (size_t)w * ld * k;
  v.Binv   = det_ptr<DT>(st.Binv[g]) + (size_t)w * k * k;
  v.dl     = st.dlist[g] + (size_t)w * k;
  v.jas    = st.jas + (size_t)w * J_ROWS * s.Npad;
  v.logdet = st.logdet + (size_t)w * 4;
  v.stats  = st.stats + (size_t)w * 2;
  return v;
}

// Phase B (thread 0): proposal from the gradient at the current position.
// Scalars that thread 0 needs later in the move, loaded at its start by 15 threads into
// sh.bc[BC_PF, BC_PF + 15) so that their latency overlaps the inverse-row phase: the walker's next
// uniform, the normal deviate, the electron's position, its J1 and J2 gradients and values
// (published by the caller's next barrier).
constexpr int BC_PF = 33;
enum PrefetchSlot
{
  PF_UNIF = 0,
  PF_GAUSS = 1,
  PF_R = 4,
  PF_G1 = 7,
  PF_G2 = 10,
  PF_VAT = 13,
  PF_UAT = 14
};
template<typename SH>
__device__ inline void prefetch_move(const SysC& s, int iat, int w, int ucount, const MoveArgs& a,
                                     const double* __restrict__ R, const double* __restrict__ jas, SH& sh,
                                     int base = BC_PF)
{
  const int t = thread_index(), Np = s.Npad;
  if (t >= 15)
    return;
  double v = 0.0;
  if (t == PF_UNIF)
    v = a.unif[(size_t)w * s.N + ucount];
  else if (t < PF_R)
    v = a.gauss[((size_t)iat * a.nw + w) * 3 + (t - PF_GAUSS)];
  else if (t < PF_G1)
    v = R[(t - PF_R) * Np + iat];
  else if (t < PF_G2)
    v = s.has_j1 ? jas[(J_G1X + t - PF_G1) * Np + iat] : 0.0;
  else if (t < PF_VAT)
    v = s.has_j2 ? jas[(J_DUX + t - PF_G2) * Np + iat] : 0.0;
  else if (t == PF_VAT)
    v = s.has_j1 ? jas[J_VAT * Np + iat] : 0.0;
  else
    v = s.has_j2 ? jas[J_UAT * Np + iat] : 0.0;
  sh.bc[base + t] = v;
}

// propose() from the prefetched scalars (thread 0; the same arithmetic)
template<typename SH>
__device__ inline void propose_pf(const SysC& s, const MoveArgs& a, SH& sh, int base = BC_PF)
{
  const double tauovermass = a.tau;
  const double sqrttau     = sqrt(tauovermass);
  const double oneover2tau = 0.5 / tauovermass;
  const double* pf         = sh.bc + base;
  double gnow[3]           = {0, 0, 0};
  if (s.has_det)
    for (int q = 0; q < 3; q++)
      gnow[q] += sh.bc[BC_GRAD + q];
  if (s.has_j1)
    for (int q = 0; q < 3; q++)
      gnow[q] += pf[PF_G1 + q];
  if (s.has_j2)
    for (int q = 0; q < 3; q++)
      gnow[q] += pf[PF_G2 + q];
  double delta[3], drift[3];
  for (int q = 0; q < 3; q++)
    delta[q] = pf[PF_GAUSS + q] * sqrttau;
  unr_drift(tauovermass, gnow, drift);
  for (int q = 0; q < 3; q++)
  {
    sh.bc[BC_GNOW + q] = gnow[q];
    drift[q] += delta[q];
    sh.bc[BC_MOVE + q] = drift[q];
    sh.bc[BC_POS + q]  = pf[PF_R + q] + drift[q];
  }
  sh.bc[BC_LOGGF] = -oneover2tau * (delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
}

template<typename SH>
__device__ inline void propose(const SysC& s, int iat, int w, const MoveArgs& a, const double* __restrict__ R,
                               const double* __restrict__ jas, SH& sh)
{
  const int Np = s.Npad;
  const double tauovermass = a.tau;
  const double sqrttau     = sqrt(tauovermass);
  const double oneover2tau = 0.5 / tauovermass;
  double gnow[3] = {0, 0, 0};
  if (s.has_det)
    for (int q = 0; q < 3; q++)
      gnow[q] += sh.bc[BC_GRAD + q];
  if (s.has_j1)
    for (int q = 0; q < 3; q++)
      gnow[q] += jas[(J_G1X + q) * Np + iat];
  if (s.has_j2)
    for (int q = 0; q < 3; q++)
      gnow[q] += jas[(J_DUX + q) * Np + iat];
  double delta[3], drift[3];
  const double* gz = a.gauss + ((size_t)iat * a.nw + w) * 3;
  for (int q = 0; q < 3; q++)
    delta[q] = gz[q] * sqrttau;
  unr_drift(tauovermass, gnow, drift);
  for (int q = 0; q < 3; q++)
  {
    sh.bc[BC_GNOW + q] = gnow[q];
    drift[q] += delta[q];
    sh.bc[BC_MOVE + q] = drift[q];
    sh.bc[BC_POS + q]  = R[q * Np + iat] + drift[q];
  }
  sh.bc[BC_LOGGF] = -oneover2tau * (delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
}

// DMC bookkeeping of one move of walker w (DMCBatched.cpp:172-175, 225): tau |xi|^2 of the
// unscaled normal xi, added to the proposed sum and, when accepted, to the accepted sum
__device__ inline void dmc_record(const MoveArgs& a, int iat, int w, bool acc)
{
  const double* gz = a.gauss + ((size_t)iat * a.nw + w) * 3;
  const double rr  = a.tau * (gz[0] * gz[0] + gz[1] * gz[1] + gz[2] * gz[2]);
  a.rr[2 * (size_t)w] += rr;
  if (acc)
    a.rr[2 * (size_t)w + 1] += rr;
}


} // namespace qmi
