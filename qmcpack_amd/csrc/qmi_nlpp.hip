// Non-local pseudopotential energy of every walker: NonLocalECPotential::evaluate over the
// electron-ion pairs within the cutoff (QMCHamiltonians/NonLocalECPotential.cpp:200-228) with
// NonLocalECPComponent::evaluateOne (NonLocalECPComponent.cpp:109-166, 826-888): for electron e at
// distance r < r_cut from ion I (dr = r_e - r_I, minimum image), the quadrature points
// r_e + (r u_j - dr) on the sphere of radius r around the ion (u_j the rotated grid) give the
// wave-function ratios psi_j / psi (determinant from value-only orbitals and the stored inverse
// row, J1, J2), and
//   V_eI = sum_j w_j [sum_l (2 l + 1) v_l(r) P_l(dr . u_j / r)] psi_j / psi.
// One workgroup per (electron, walker): the ions within the cutoff in ion order (wave ballots),
// the inverse row in registers for every quadrature point.  Per-electron sums to part[w][N], then
// one thread per walker adds them in electron order.
#include "qmi_move_common.hpp"

namespace qmi
{

template<int NL_KC, int TM>
struct NlLDS
{
  alignas(16) double red[4][3 * NL_KC]; // per-wave partials: det, J1, J2 of each point
  double jc[3][MAX_JCOEF];
  unsigned long long mask[NLPP_MAX_IONS / 64];
  double vp[NL_KC][3];  // virtual positions
  double ab[NL_KC][16]; // spline factors a_i b_j of each point
  double cz[NL_KC][4];  // and c_k
  long long off[NL_KC]; // table offset of each point's stencil corner
  double kp[NL_KC];     // knot partial products w_j sum_l (2l+1) v_l(r) P_l(cos)
  double sum[3 * NL_KC];
  double dv[TM ? NL_KC : 1][3]; // deltaV_j = r u_j - dr (the non-local move of each point)
};

// v_l(r) from the per-interval cubic of channel ch (r inside the grid)
__device__ inline double nlpp_vl(const NlppC& c, int ch, double r)
{
  const int loc   = min((int)(r * c.drinv), c.ng - 2);
  const double t  = (r - loc * c.dr) * c.drinv;
  const double* q = c.vpoly + ((size_t)ch * (c.ng - 1) + loc) * 4;
  return q[0] + t * (q[1] + t * (q[2] + t * q[3]));
}

// One electron-ion pair: the quadrature points of electron iel (at pe) on the sphere around ion
// `ion`, in passes of NL_KC points; returns sum_j w_j [sum_l (2 l + 1) v_l(r) P_l] psi_j / psi on
// thread 0 (0 on the others).  Called by every thread of the workgroup (it synchronises).  TM != 0
// continues the scan of the non-local moves from `scan` (thread 0); TM == 2 records the selected
// point in sel.  wout (thread 0, optional): the weight of every point.
template<typename ST, typename DT, int NJ, int MAXM, int NL_KC, int TM>
__device__ __forceinline__ double nlpp_pair(const SysC& s, const State& st, const ST* __restrict__ coefs,
                                            const NlppC& c, NlLDS<NL_KC, TM>& sh, const double* rw,
                                            const double* jas, int w, int iel, int g, int n, int ion,
                                            const double (&pe)[3], const double (&arow)[MAXM],
                                            const double (&ex)[NJ], const double (&ey)[NJ], const double (&ez)[NJ],
                                            double& scan, double (&sel)[4], const TmoveArgs& tm,
                                            double* __restrict__ wout = nullptr)
{
  const int ni = s.nion, N = s.N, Np = s.Npad;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  double dI[3];
  const double r = min_image(s, st.ions[ion], st.ions[ni + ion], st.ions[2 * ni + ion], pe, dI);
  const double dr[3] = {-dI[0], -dI[1], -dI[2]}; // r_e - r_I
  double pairpot     = 0.0;
  for (int j0 = 0; j0 < c.nknot; j0 += NL_KC)
  {
    const int kc = min(NL_KC, c.nknot - j0);
    if (t < kc)
    {
      // NonLocalECPComponent: rotated point, deltaV, knot partial product
      const double* k = c.knots + 4 * (j0 + t);
      double u[3], p[3];
      for (int a = 0; a < 3; a++)
        u[a] = rw[3 * a] * k[0] + rw[3 * a + 1] * k[1] + rw[3 * a + 2] * k[2];
      for (int a = 0; a < 3; a++)
      {
        const double dv = r * u[a] - dr[a];
        if constexpr (TM != 0)
          sh.dv[t][a] = dv;
        p[a]        = pe[a] + dv;
        sh.vp[t][a] = p[a];
      }
      const double zz = (dr[0] * u[0] + dr[1] * u[1] + dr[2] * u[2]) * (1.0 / r);
      double lpol[NLPP_MAX_L + 1];
      lpol[0]         = 1.0;
      double lpolprev = 0.0;
      for (int l = 0; l < c.lmax; l++)
      {
        lpol[l + 1] = ((double)(2 * l + 1) * zz * lpol[l] - l * lpolprev) * (1.0 / (double)(l + 1));
        lpolprev    = lpol[l];
      }
      double lsum = 0.0;
      for (int ch = 0; ch < c.nch; ch++)
        lsum += (nlpp_vl(c, ch, r) * (double)(2 * c.l[ch] + 1)) * lpol[c.l[ch]];
      sh.kp[t] = lsum * k[3];
      if (s.has_det)
      {
        SplineLoc L;
        spline_locate(s, p, L);
        sh.off[t] = L.off;
        for (int i = 0; i < 4; i++)
        {
          sh.cz[t][i] = L.c[i];
          for (int jj = 0; jj < 4; jj++)
            sh.ab[t][i * 4 + jj] = L.a[i] * L.b[jj];
        }
      }
    }
    __syncthreads();
    double adet[NL_KC], aj1[NL_KC], aj2[NL_KC];
#pragma unroll
    for (int q = 0; q < NL_KC; q++)
      adet[q] = aj1[q] = aj2[q] = 0.0;
    if (s.has_det)
    {
#pragma unroll
      for (int mm = 0; mm < MAXM; mm++)
      {
        const int o = t + mm * WG;
        if (o < n)
        {
#pragma unroll
          for (int q = 0; q < NL_KC; q++)
            if (q < kc)
            {
              const ST* base = coefs + sh.off[q] + o;
              double val     = 0.0;
#pragma unroll
              for (int i = 0; i < 4; i++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                {
                  const ST* qq    = base + i * s.xs + jj * s.ys;
                  const double s0 = sh.cz[q][0] * (double)qq[0] + sh.cz[q][1] * (double)qq[s.zs] +
                      sh.cz[q][2] * (double)qq[2 * s.zs] + sh.cz[q][3] * (double)qq[3 * s.zs];
                  val += sh.ab[q][i * 4 + jj] * s0;
                }
              adet[q] += val * arow[mm];
            }
        }
      }
    }
    if (s.has_j1)
      for (int j = t; j < ni; j += WG)
      {
        const double ix = st.ions[j], iy = st.ions[ni + j], iz = st.ions[2 * ni + j];
#pragma unroll
        for (int q = 0; q < NL_KC; q++)
          if (q < kc)
          {
            double d[3];
            const double rq = min_image(s, ix, iy, iz, sh.vp[q], d);
            if (rq < s.j1.rcut)
              aj1[q] += functor_v_c(sh.jc[0], s.j1.DeltaRInv, s.j1.max_index, rq);
          }
      }
    if (s.has_j2)
    {
#pragma unroll
      for (int u = 0; u < NJ; u++)
      {
        const int j = t + u * WG;
        if (j >= N || j == iel)
          continue;
        const int fi      = ((j < s.nup) ? 0 : 1) == g ? 0 : 1;
        const FunctorC& f = s.j2[fi];
#pragma unroll
        for (int q = 0; q < NL_KC; q++)
          if (q < kc)
          {
            double d[3];
            const double rq = min_image(s, ex[u], ey[u], ez[u], sh.vp[q], d);
            if (rq < f.rcut)
              aj2[q] += functor_v_c(sh.jc[1 + fi], f.DeltaRInv, f.max_index, rq);
          }
      }
    }
    // reduce the 3 kc sums: xor shuffles within each wave, then the four wave partials
#pragma unroll
    for (int q = 0; q < NL_KC; q++)
    {
      double v0 = adet[q], v1 = aj1[q], v2 = aj2[q];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1)
      {
        v0 += __shfl_xor(v0, off, 64);
        v1 += __shfl_xor(v1, off, 64);
        v2 += __shfl_xor(v2, off, 64);
      }
      if (lane == 0)
      {
        sh.red[wave][q]             = v0;
        sh.red[wave][NL_KC + q]     = v1;
        sh.red[wave][2 * NL_KC + q] = v2;
      }
    }
    __syncthreads();
    if (t < 3 * NL_KC)
      sh.sum[t] = sh.red[0][t] + sh.red[1][t] + sh.red[2][t] + sh.red[3][t];
    __syncthreads();
    if (t == 0)
      for (int q = 0; q < kc; q++)
      {
        double ratio = 1.0;
        if (s.has_det)
          ratio *= sh.sum[q];
        if (s.has_j1)
          ratio *= exp(jas[J_VAT * Np + iel] - sh.sum[NL_KC + q]);
        if (s.has_j2)
          ratio *= exp(jas[J_UAT * Np + iel] - sh.sum[2 * NL_KC + q]);
        const double wgt = sh.kp[q] * ratio;
        pairpot += wgt;
        if (wout)
          wout[j0 + q] = wgt;
        if constexpr (TM != 0)
        {
          if (TM == 2 && scan <= tm.sel_target[w])
          {
            sel[0] = 1.0;
            sel[1] = sh.dv[q][0];
            sel[2] = sh.dv[q][1];
            sel[3] = sh.dv[q][2];
          }
          scan += wgt > 0.0 ? wgt * tm.plus : wgt * tm.minus;
        }
      }
    __syncthreads(); // the point tables are rewritten by the next chunk
  }
  return pairpot;
}

// Up to NL_KC (6 or 12) quadrature points at once (6 at occupancy 3 is the faster form): each thread accumulates, for every point, its orbitals'
// values dotted with the inverse row and its ions' / electrons' Jastrow values, so that the
// stencil gathers of all points are in flight together and one reduction serves them all.
// TM: 0 energy only, 1 energy and the per-electron scan weights, 2 selection (TmoveArgs)
template<typename ST, typename DT, int NJ, int MAXM, int NL_KC, int TM>
__device__ __forceinline__ void nlpp_body(SysC s, State st, const ST* __restrict__ coefs, NlppC c,
                                          const double* __restrict__ rot, double* __restrict__ part, TmoveArgs tm)
{
  __shared__ NlLDS<NL_KC, TM> sh;
  const int w = blockIdx.y, t = threadIdx.x;
  const int iel = TM == 2 ? tm.sel_iel[w] : (int)blockIdx.x;
  if constexpr (TM == 2)
    if (iel < 0) // no move selected for this walker
      return;
  const int ni = s.nion, N = s.N, Np = s.Npad;
  const double* R    = st.R + (size_t)w * 3 * Np;
  const double pe[3] = {R[iel], R[Np + iel], R[2 * Np + iel]};
  // ions within the cutoff, one bit per ion in ion order
  const int wave = t >> 6, lane = t & 63;
  for (int j0 = 0; j0 < ni; j0 += WG)
  {
    const int j = j0 + t;
    bool in     = false;
    if (j < ni)
    {
      double d[3];
      in = min_image(s, st.ions[j], st.ions[ni + j], st.ions[2 * ni + j], pe, d) < c.rcut;
    }
    const unsigned long long b = __ballot(in);
    if (lane == 0)
      sh.mask[(j0 >> 6) + wave] = b;
  }
  __syncthreads();
  const int nmask = (ni + 63) >> 6;
  bool any        = false;
  for (int m = 0; m < nmask; m++)
    any |= sh.mask[m] != 0ull;
  if (!any)
  {
    if (t == 0)
    {
      if constexpr (TM == 2)
        tm.sel_out[(size_t)w * 4] = 0.0;
      else
        part[(size_t)w * N + iel] = 0.0;
      if constexpr (TM == 1)
        tm.tsum[(size_t)w * N + iel] = -1.0;
    }
    return;
  }
  stage_jastrow(s, sh);
  const int g = iel < s.nup ? 0 : 1, n = s.n[g], ld = s.ld, r0 = iel - s.first[g];
  const double* jas = st.jas + (size_t)w * J_ROWS * Np;
  double arow[MAXM];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    arow[m] = 0.0;
  if (s.has_det)
  {
    const DT* a = det_ptr<DT>(st.Ainv[g]) + ((size_t)w * n + r0) * ld;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int o = t + m * WG;
      arow[m]     = o < n ? (double)a[o] : 0.0;
    }
  }
  // this thread's electrons for the J2 sums
  double ex[NJ], ey[NJ], ez[NJ];
#pragma unroll
  for (int u = 0; u < NJ; u++)
  {
    const int j = t + u * WG;
    const bool ok = j < N;
    ex[u] = ok ? R[j] : 0.0;
    ey[u] = ok ? R[Np + j] : 0.0;
    ez[u] = ok ? R[2 * Np + j] : 0.0;
  }
  const double* rw   = rot + (size_t)w * 9;
  double value       = 0.0; // thread 0
  double scan        = TM == 2 ? tm.sel_base[w] : 0.0; // thread 0: scan weight
  double sel[4]      = {0.0, 0.0, 0.0, 0.0};
  __syncthreads();
  for (int m = 0; m < nmask; m++)
  {
    unsigned long long bits = sh.mask[m];
    while (bits)
    {
      const int ion = m * 64 + __ffsll((long long)bits) - 1;
      bits &= bits - 1;
      const double pairpot = nlpp_pair<ST, DT, NJ, MAXM, NL_KC, TM>(s, st, coefs, c, sh, rw, jas, w, iel, g, n, ion, pe,
                                                                    arow, ex, ey, ez, scan, sel, tm);
      if (t == 0)
        value += pairpot;
    }
  }
  if (t == 0)
  {
    if constexpr (TM == 2)
      for (int a = 0; a < 4; a++)
        tm.sel_out[(size_t)w * 4 + a] = sel[a];
    else
      part[(size_t)w * N + iel] = value;
    if constexpr (TM == 1)
      tm.tsum[(size_t)w * N + iel] = scan;
  }
}

// the energy alone keeps the plain signature; the T-move modes take the move arguments
template<typename ST, typename DT, int NJ, int MAXM, int NL_KC>
__global__ __launch_bounds__(WG, NL_KC == 6 ? 3 : 1) void k_nlpp(SysC s, State st, const ST* __restrict__ coefs, NlppC c,
                                                               const double* __restrict__ rot,
                                                               double* __restrict__ part)
{
  nlpp_body<ST, DT, NJ, MAXM, NL_KC, 0>(s, st, coefs, c, rot, part, TmoveArgs{});
}

template<typename ST, typename DT, int NJ, int MAXM, int NL_KC, int TM>
__global__ __launch_bounds__(WG, NL_KC == 6 ? 3 : 1) void k_nlpp_tm(SysC s, State st, const ST* __restrict__ coefs,
                                                                  NlppC c, const double* __restrict__ rot,
                                                                  double* __restrict__ part, TmoveArgs tm)
{
  nlpp_body<ST, DT, NJ, MAXM, NL_KC, TM>(s, st, coefs, c, rot, part, tm);
}

__global__ void k_nlpp_sum(int nw, int N, const double* __restrict__ part, double* __restrict__ out)
{
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw)
    return;
  double s = 0.0;
  for (int i = 0; i < N; i++)
    s += part[(size_t)w * N + i];
  out[w] = s;
}

// ---- pair-list form (modes 0 and 1) ----------------------------------------------------------------

__global__ void k_nlpp_pairs_zero(int ni, int* __restrict__ icnt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ni)
    icnt[i] = 0;
}

// one thread per (walker, electron): the ions within the cutoff in ion order (the same distance test
// as the per-electron form), their count per ion
__global__ void k_nlpp_pairs_find(SysC s, State st, NlppC c, NlppPairs pl, int nN)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nN)
    return;
  const int w = idx / s.N, e = idx - w * s.N, ni = s.nion, Np = s.Npad;
  const double* R    = st.R + (size_t)w * 3 * Np;
  const double pe[3] = {R[e], R[Np + e], R[2 * Np + e]};
  int k              = 0;
  for (int ion = 0; ion < ni; ion++)
  {
    double d[3];
    if (min_image(s, st.ions[ion], st.ions[ni + ion], st.ions[2 * ni + ion], pe, d) < c.rcut)
    {
      if (k < pl.ke)
      {
        pl.pion[(size_t)idx * pl.ke + k] = ion;
        atomicAdd(&pl.icnt[ion], 1);
      }
      k++;
    }
  }
  pl.ecnt[idx] = k;
}

// one workgroup: exclusive scan of the per-ion counts (they become the scatter cursors) and the total
__global__ __launch_bounds__(WG) void k_nlpp_pairs_scan(int ni, NlppPairs pl)
{
  __shared__ int part[WG];
  const int t = threadIdx.x, per = (ni + WG - 1) / WG;
  const int lo = min(ni, t * per), hi = min(ni, lo + per);
  int sum = 0;
  for (int i = lo; i < hi; i++)
    sum += pl.icnt[i];
  part[t] = sum;
  __syncthreads();
  if (t == 0)
  {
    int acc = 0;
    for (int i = 0; i < WG; i++)
    {
      const int v = part[i];
      part[i]     = acc;
      acc += v;
    }
    *pl.total = acc;
  }
  __syncthreads();
  int acc = part[t];
  for (int i = lo; i < hi; i++)
  {
    const int v = pl.icnt[i];
    pl.icnt[i]  = acc;
    acc += v;
  }
}

// one thread per (walker, electron): its slots into the ion-sorted list (order within one ion's
// range is arbitrary; each slot's result does not depend on it)
__global__ void k_nlpp_pairs_scatter(NlppPairs pl, int nN)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nN)
    return;
  const int k = min(pl.ecnt[idx], pl.ke);
  for (int j = 0; j < k; j++)
  {
    const int slot = idx * pl.ke + j;
    pl.list[atomicAdd(&pl.icnt[pl.pion[slot]], 1)] = slot;
  }
}

// Persistent grid over the ion-sorted pairs: workgroup b runs on XCD b % 8 (round-robin dispatch)
// and takes pairs lo + b / 8, lo + b / 8 + grid / 8, ... of that XCD's contiguous eighth of the
// list, so each XCD sweeps its ions in order with all its workgroups on one or two ions at a time.
// Every workgroup reaches the end of its range (the loop bound is uniform across the workgroup).
// PTS: the weights of every point to wpts (the non-local moves' scan), else the pair sums to wres.
template<typename ST, typename DT, int NJ, int MAXM, int NL_KC, bool PTS>
__global__ __launch_bounds__(WG, NL_KC == 6 ? 4 : 1) void k_nlpp_pairs(SysC s, State st, const ST* __restrict__ coefs,
                                                                     NlppC c, const double* __restrict__ rot,
                                                                     NlppPairs pl)
{
  __shared__ NlLDS<NL_KC, 0> sh;
  const int t = threadIdx.x, N = s.N, Np = s.Npad;
  const int total = *pl.total;
  const int per = gridDim.x >> 3, j = blockIdx.x >> 3, chunk = (total + 7) >> 3;
  const int lo = (blockIdx.x & 7) * chunk, hi = min(total, lo + chunk);
  stage_jastrow(s, sh);
  double scan   = 0.0;
  double sel[4] = {0.0, 0.0, 0.0, 0.0};
  const TmoveArgs tm{};
  for (int i = lo + j; i < hi; i += per)
  {
    const int slot = pl.list[i];
    const int ion  = pl.pion[slot];
    const int we   = slot / pl.ke;
    const int w = we / N, iel = we - w * N;
    const double* R    = st.R + (size_t)w * 3 * Np;
    const double pe[3] = {R[iel], R[Np + iel], R[2 * Np + iel]};
    const int g = iel < s.nup ? 0 : 1, n = s.n[g], ld = s.ld, r0 = iel - s.first[g];
    const double* jas = st.jas + (size_t)w * J_ROWS * Np;
    double arow[MAXM];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      arow[m] = 0.0;
    if (s.has_det)
    {
      const DT* a = det_ptr<DT>(st.Ainv[g]) + ((size_t)w * n + r0) * ld;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int o = t + m * WG;
        arow[m]     = o < n ? (double)a[o] : 0.0;
      }
    }
    double ex[NJ], ey[NJ], ez[NJ];
#pragma unroll
    for (int u = 0; u < NJ; u++)
    {
      const int jj  = t + u * WG;
      const bool ok = jj < N;
      ex[u]         = ok ? R[jj] : 0.0;
      ey[u]         = ok ? R[Np + jj] : 0.0;
      ez[u]         = ok ? R[2 * Np + jj] : 0.0;
    }
    const double pairpot =
        nlpp_pair<ST, DT, NJ, MAXM, NL_KC, 0>(s, st, coefs, c, sh, rot + (size_t)w * 9, jas, w, iel, g, n, ion, pe, arow,
                                              ex, ey, ez, scan, sel, tm,
                                              PTS && t == 0 ? pl.wpts + (size_t)slot * c.nknot : nullptr);
    if (!PTS && t == 0)
      pl.wres[slot] = pairpot;
  }
}

// one thread per (walker, electron): its pair sums in ion order, as the per-electron form adds them
__global__ void k_nlpp_pairs_esum(NlppPairs pl, int nN, double* __restrict__ part)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nN)
    return;
  const int k = pl.ecnt[idx];
  double v    = 0.0;
  if (k > pl.ke)
    v = __builtin_nan("");
  else
    for (int j = 0; j < k; j++)
      v += pl.wres[(size_t)idx * pl.ke + j];
  part[idx] = v;
}

// mode 1: from the point weights, the energy and the scan weight of each electron with the
// per-electron form's operations in its order (points of a pair from 0, pairs in ion order; the scan
// running over all points of the electron); -1 for an electron without pairs
__global__ void k_nlpp_pairs_esum_tm(NlppPairs pl, int nN, int nknot, double plus, double minus,
                                     double* __restrict__ part, double* __restrict__ tsum)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nN)
    return;
  const int k = pl.ecnt[idx];
  double v = 0.0, scan = 0.0;
  if (k > pl.ke)
    v = scan = __builtin_nan("");
  else
    for (int j = 0; j < k; j++)
    {
      const double* wq = pl.wpts + ((size_t)idx * pl.ke + j) * nknot;
      double pairpot   = 0.0;
      for (int q = 0; q < nknot; q++)
      {
        const double wgt = wq[q];
        pairpot += wgt;
        scan += wgt > 0.0 ? wgt * plus : wgt * minus;
      }
      v += pairpot;
    }
  part[idx] = v;
  tsum[idx] = k == 0 ? -1.0 : scan;
}

hipError_t launch_nlpp(const SysC& s, const State& st, const void* coefs, const NlppC& c, int nw, const double* rot,
                       double* part, double* out, const TmoveArgs& tm, hipStream_t stream,
                       const NlppPairs* pairs)
{
  if (s.nion > NLPP_MAX_IONS || c.nch < 1 || c.nch > NLPP_MAX_CH || c.lmax > NLPP_MAX_L || c.ng < 2 || nw < 1)
    return hipErrorInvalidValue;
  const int mm = std::max(move_maxm(s, 0), move_maxm(s, 1));
  if (mm == 0)
    return hipErrorInvalidValue;
  if (pairs && (tm.mode == 0 || tm.mode == 1))
  {
    const NlppPairs& pl = *pairs;
    const int nN        = nw * s.N;
    const bool pts      = tm.mode == 1;
    if (pl.ke < 1 || pl.grid < 8 || pl.grid % 8 != 0 || (long long)nN * pl.ke > (1LL << 31) - 1 ||
        (pts && (!pl.wpts || !tm.tsum)))
      return hipErrorInvalidValue;
    const int tb = 256, nb = (nN + tb - 1) / tb;
    k_nlpp_pairs_zero<<<std::max(1, (s.nion + tb - 1) / tb), tb, 0, stream>>>(s.nion, pl.icnt); // no ions: one idle block
    k_nlpp_pairs_find<<<nb, tb, 0, stream>>>(s, st, c, pl, nN);
    k_nlpp_pairs_scan<<<1, WG, 0, stream>>>(s.nion, pl);
    k_nlpp_pairs_scatter<<<nb, tb, 0, stream>>>(pl, nN);
    const hipError_t e = dispatch_prec(s.prec, [&](auto stv, auto dtv) {
      using ST     = decltype(stv);
      using DT     = decltype(dtv);
      const ST* cf = static_cast<const ST*>(coefs);
      auto go = [&](auto ptsv) {
        constexpr bool P = decltype(ptsv)::value;
        if (mm == 1 && c.kc == 6)
          k_nlpp_pairs<ST, DT, 2, 1, 6, P><<<pl.grid, WG, 0, stream>>>(s, st, cf, c, rot, pl);
        else if (mm == 1)
          k_nlpp_pairs<ST, DT, 2, 1, 12, P><<<pl.grid, WG, 0, stream>>>(s, st, cf, c, rot, pl);
        else if (c.kc == 6)
          k_nlpp_pairs<ST, DT, 4, 2, 6, P><<<pl.grid, WG, 0, stream>>>(s, st, cf, c, rot, pl);
        else
          k_nlpp_pairs<ST, DT, 4, 2, 12, P><<<pl.grid, WG, 0, stream>>>(s, st, cf, c, rot, pl);
      };
      if (pts)
        go(std::true_type{});
      else
        go(std::false_type{});
      return hipGetLastError();
    });
    if (e != hipSuccess)
      return e;
    if (pts)
      k_nlpp_pairs_esum_tm<<<nb, tb, 0, stream>>>(pl, nN, c.nknot, tm.plus, tm.minus, part, tm.tsum);
    else
      k_nlpp_pairs_esum<<<nb, tb, 0, stream>>>(pl, nN, part);
    k_nlpp_sum<<<(nw + 63) / 64, 64, 0, stream>>>(nw, s.N, part, out);
    return hipGetLastError();
  }
  auto launch = [&](auto tmv) {
    constexpr int TM = decltype(tmv)::value;
    return dispatch_prec(s.prec, [&](auto stv, auto dtv) {
      using ST = decltype(stv);
      using DT = decltype(dtv);
      const dim3 grid(TM == 2 ? 1 : s.N, nw);
      const ST* cf = static_cast<const ST*>(coefs);
      if constexpr (TM == 0)
      {
        if (mm == 1 && c.kc == 6)
          k_nlpp<ST, DT, 2, 1, 6><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part);
        else if (mm == 1)
          k_nlpp<ST, DT, 2, 1, 12><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part);
        else if (c.kc == 6)
          k_nlpp<ST, DT, 4, 2, 6><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part);
        else
          k_nlpp<ST, DT, 4, 2, 12><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part);
      }
      else
      {
        if (mm == 1 && c.kc == 6)
          k_nlpp_tm<ST, DT, 2, 1, 6, TM><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part, tm);
        else if (mm == 1)
          k_nlpp_tm<ST, DT, 2, 1, 12, TM><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part, tm);
        else if (c.kc == 6)
          k_nlpp_tm<ST, DT, 4, 2, 6, TM><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part, tm);
        else
          k_nlpp_tm<ST, DT, 4, 2, 12, TM><<<grid, WG, 0, stream>>>(s, st, cf, c, rot, part, tm);
      }
      return hipGetLastError();
    });
  };
  const hipError_t e = tm.mode == 2   ? launch(std::integral_constant<int, 2>{})
                       : tm.mode == 1 ? launch(std::integral_constant<int, 1>{})
                                      : launch(std::integral_constant<int, 0>{});
  if (e != hipSuccess || tm.mode == 2)
    return e;
  k_nlpp_sum<<<(nw + 63) / 64, 64, 0, stream>>>(nw, s.N, part, out);
  return hipGetLastError();
}

} // namespace qmi
