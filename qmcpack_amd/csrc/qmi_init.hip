// Initialisation, recompute and gradient/Laplacian kernels.
#include "qmi_kernels.hpp"

#include <algorithm>

namespace qmi
{

// Padded table from the unique periodic coefficients: table[i][j][k][o] = u[(i-1)%Mx][(j-1)%My][(k-1)%Mz][o]
// (einspline/bspline_create.c:1027-1034 periodic wrap; MultiBsplineBase.hpp:75-121 layout).
template<typename ST>
__global__ void k_wrap_coefs(SysC s, const double* __restrict__ unique, int norb, ST* __restrict__ table)
{
  const long long Nx = s.M[0] + 3, Ny = s.M[1] + 3, Nz = s.M[2] + 3;
  const long long pt = blockIdx.x;
  if (pt >= Nx * Ny * Nz)
    return;
  const int i = int(pt / (Ny * Nz)), j = int((pt / Nz) % Ny), kk = int(pt % Nz);
  const int ui = (i - 1 + s.M[0]) % s.M[0], uj = (j - 1 + s.M[1]) % s.M[1], uk = (kk - 1 + s.M[2]) % s.M[2];
  const double* src = unique + (((long long)ui * s.M[1] + uj) * s.M[2] + uk) * norb;
  ST* dst           = table + pt * s.npad;
  for (int o = threadIdx.x; o < s.npad; o += blockDim.x)
    dst[o] = from_double<ST>(o < norb ? src[o] : 0.0);
}

hipError_t launch_wrap_coefs(const SysC& s, const double* unique, int norb, void* table, hipStream_t stream)
{
  const long long npts = (long long)(s.M[0] + 3) * (s.M[1] + 3) * (s.M[2] + 3);
  return dispatch_prec(s.prec, [&](auto stv, auto) {
    using ST = decltype(stv);
    k_wrap_coefs<ST><<<dim3((unsigned)npts), 256, 0, stream>>>(s, unique, norb, static_cast<ST*>(table));
    return hipGetLastError();
  });
}

// Orbital VGL at every electron of determinant g (BsplineSet::mw_evaluate_notranspose,
// BsplineSet.h:142-189).  Values are written transposed into the FP64 LU workspace
// (lu[w][j][i] = phi_j(r_i), the column-major psiM handed to the LU; DiracMatrix.h:204-220 upcasts
// to FP64 the same way in mixed precision), gradients/Laplacians into dpsi[w][i][0..3][j].
template<typename ST, typename DT>
__global__ __launch_bounds__(WG) void k_spo_rows(SysC s, State st, const ST* __restrict__ coefs, int g,
                                                 double* __restrict__ lu)
{
  const int i = blockIdx.x, w = blockIdx.y;
  const int n = s.n[g], ld = s.ld, Np = s.Npad;
  const int iat   = s.first[g] + i;
  const double* R = st.R + (size_t)w * 3 * Np;
  const double pos[3] = {R[iat], R[Np + iat], R[2 * Np + iat]};
  SplineLoc L;
  spline_locate(s, pos, L);
  double* A = lu + (size_t)w * n * ld;
  DT* drow  = det_ptr<DT>(st.dpsi[g]) + ((size_t)w * n + i) * 4 * ld;
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double out[5];
    spline_vgl_orbital<ST>(s, coefs, L, o, out);
    A[(size_t)o * ld + i] = out[0];
    drow[o]               = (DT)out[1];
    drow[ld + o]          = (DT)out[2];
    drow[2 * ld + o]      = (DT)out[3];
    drow[3 * ld + o]      = (DT)out[4];
  }
}

hipError_t launch_spo_rows(const SysC& s, const State& st, const void* coefs, int g, int nw, double* lu,
                           hipStream_t stream)
{
  return dispatch_prec(s.prec, [&](auto stv, auto dtv) {
    using ST = decltype(stv);
    using DT = decltype(dtv);
    k_spo_rows<ST, DT><<<dim3(s.n[g], nw), WG, 0, stream>>>(s, st, static_cast<const ST*>(coefs), g, lu);
    return hipGetLastError();
  });
}

// J1 and J2 accumulators from scratch (J1OrbitalSoA.h:240-249; TwoBodyJastrow.cpp:667-713, summed
// over the full row instead of the lower triangle).  One thread per electron.
__global__ __launch_bounds__(WG) void k_jastrow_recompute(SysC s, State st)
{
  const int w = blockIdx.y;
  const int i = blockIdx.x * WG + threadIdx.x;
  const int N = s.N, Np = s.Npad;
  if (i >= N)
    return;
  const double* R = st.R + (size_t)w * 3 * Np;
  double* jas     = st.jas + (size_t)w * J_ROWS * Np;
  const double pos[3] = {R[i], R[Np + i], R[2 * Np + i]};
  if (s.has_j1)
  {
    double v = 0, gx = 0, gy = 0, gz = 0, lap = 0;
    for (int j = 0; j < s.nion; j++)
    {
      double dr[3];
      const double r = min_image(s, st.ions[j], st.ions[s.nion + j], st.ions[2 * s.nion + j], pos, dr);
      if (r < s.j1.rcut)
      {
        double u, du, d2u;
        functor_vgl(s.j1, r, u, du, d2u);
        v += u;
        gx += du * dr[0];
        gy += du * dr[1];
        gz += du * dr[2];
        lap += d2u + 2.0 * du;
      }
    }
    jas[J_VAT * Np + i]       = v;
    jas[(J_G1X + 0) * Np + i] = gx;
    jas[(J_G1X + 1) * Np + i] = gy;
    jas[(J_G1X + 2) * Np + i] = gz;
    jas[J_LAP1 * Np + i]      = lap;
  }
  if (s.has_j2)
  {
    const int gi = i < s.nup ? 0 : 1;
    double v = 0, gx = 0, gy = 0, gz = 0, lap = 0;
    for (int j = 0; j < N; j++)
    {
      if (j == i)
        continue;
      double dr[3];
      const double r    = min_image(s, R[j], R[Np + j], R[2 * Np + j], pos, dr);
      const FunctorC& f = s.j2[((j < s.nup) ? 0 : 1) == gi ? 0 : 1];
      if (r < f.rcut)
      {
        double u, du, d2u;
        functor_vgl(f, r, u, du, d2u);
        v += u;
        gx += du * dr[0];
        gy += du * dr[1];
        gz += du * dr[2];
        lap += d2u + 2.0 * du;
      }
    }
    jas[J_UAT * Np + i]       = v;
    jas[(J_DUX + 0) * Np + i] = gx;
    jas[(J_DUX + 1) * Np + i] = gy;
    jas[(J_DUX + 2) * Np + i] = gz;
    jas[J_D2U * Np + i]       = -lap;
  }
}

hipError_t launch_jastrow_recompute(const SysC& s, const State& st, int nw, hipStream_t stream)
{
  k_jastrow_recompute<<<dim3((s.N + WG - 1) / WG, nw), WG, 0, stream>>>(s, st);
  return hipGetLastError();
}

// TrialWaveFunction::mw_evaluateGL (TrialWaveFunction.cpp:869-918): determinant rows
// (DiracDeterminantBatched::computeGL, :593-604), J1 (J1OrbitalSoA.h:115-123), J2
// (TwoBodyJastrow.cpp:769-778); kinetic energy -1/2 (sum G.G + sum L) (BareKineticEnergy.cpp:109-118).
template<typename DT>
__global__ __launch_bounds__(WG) void k_eval_gl(SysC s, State st, double* __restrict__ out3, double* __restrict__ Gout,
                                                double* __restrict__ Lout)
{
  __shared__ double sG[3 * 1024];
  __shared__ double sL[1024];
  __shared__ double red[4 * WG];
  __shared__ double bc[4];
  const int w = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int N = s.N, Np = s.Npad, ld = s.ld;
  const double* jas = st.jas + (size_t)w * J_ROWS * Np;
  for (int i = t; i < N; i += WG)
  {
    sG[3 * i] = sG[3 * i + 1] = sG[3 * i + 2] = 0.0;
    sL[i]                                     = 0.0;
  }
  __syncthreads();
  if (s.has_det)
    for (int g = 0; g < 2; g++)
    {
      const int n     = s.n[g];
      const DT* A = det_ptr<DT>(st.Ainv[g]) + (size_t)w * n * ld;
      const DT* D = det_ptr<DT>(st.dpsi[g]) + (size_t)w * n * 4 * ld;
      for (int i = wave; i < n; i += NWAVE)
      {
        double gx = 0, gy = 0, gz = 0, lp = 0;
        const DT* a  = A + (size_t)i * ld;
        const DT* dd = D + (size_t)i * 4 * ld;
        for (int o = lane; o < n; o += 64)
        {
          const double av = a[o];
          gx += av * (double)dd[o];
          gy += av * (double)dd[ld + o];
          gz += av * (double)dd[2 * ld + o];
          lp += av * (double)dd[3 * ld + o];
        }
        gx = wave_sum(gx);
        gy = wave_sum(gy);
        gz = wave_sum(gz);
        lp = wave_sum(lp);
        if (lane == 0)
        {
          const int iat = s.first[g] + i;
          sG[3 * iat]     = gx;
          sG[3 * iat + 1] = gy;
          sG[3 * iat + 2] = gz;
          sL[iat]         = lp - (gx * gx + gy * gy + gz * gz);
        }
      }
    }
  __syncthreads();
  double acc[4] = {0, 0, 0, 0}; // G.G, sum L, sum Vat, sum Uat
  for (int i = t; i < N; i += WG)
  {
    double g0 = sG[3 * i], g1 = sG[3 * i + 1], g2 = sG[3 * i + 2], l = sL[i];
    if (s.has_j1)
    {
      g0 += jas[(J_G1X + 0) * Np + i];
      g1 += jas[(J_G1X + 1) * Np + i];
      g2 += jas[(J_G1X + 2) * Np + i];
      l -= jas[J_LAP1 * Np + i];
      acc[2] += jas[J_VAT * Np + i];
    }
    if (s.has_j2)
    {
      g0 += jas[(J_DUX + 0) * Np + i];
      g1 += jas[(J_DUX + 1) * Np + i];
      g2 += jas[(J_DUX + 2) * Np + i];
      l += jas[J_D2U * Np + i];
      acc[3] += jas[J_UAT * Np + i];
    }
    acc[0] += g0 * g0 + g1 * g1 + g2 * g2;
    acc[1] += l;
    if (Gout)
    {
      double* G = Gout + ((size_t)w * N + i) * 3;
      G[0]      = g0;
      G[1]      = g1;
      G[2]      = g2;
    }
    if (Lout)
      Lout[(size_t)w * N + i] = l;
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

hipError_t launch_eval_gl(const SysC& s, const State& st, int nw, double* out3, double* G, double* L,
                          hipStream_t stream)
{
  if (s.N > 1024)
    return hipErrorInvalidValue;
  return dispatch_prec(s.prec, [&](auto, auto dtv) {
    using DT = decltype(dtv);
    k_eval_gl<DT><<<nw, WG, 0, stream>>>(s, st, out3, G, L);
    return hipGetLastError();
  });
}

//------------------------------------------------------------------------------------------------
// API helpers

// SPOSet::mw_evaluateVGL: out[w][5][n] at pos[w]
template<typename ST>
__global__ __launch_bounds__(WG) void k_api_spo_vgl(SysC s, const ST* __restrict__ coefs, int n,
                                                    const double* __restrict__ pos, double* __restrict__ out)
{
  const int w = blockIdx.x;
  const double p[3] = {pos[w * 3], pos[w * 3 + 1], pos[w * 3 + 2]};
  SplineLoc L;
  spline_locate(s, p, L);
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double v[5];
    spline_vgl_orbital<ST>(s, coefs, L, o, v);
    for (int q = 0; q < 5; q++)
      out[((size_t)w * 5 + q) * n + o] = v[q];
  }
}

hipError_t launch_api_spo_vgl(const SysC& s, const void* coefs, int nw, int n, const double* pos, double* out,
                              hipStream_t stream)
{
  return dispatch_prec(s.prec, [&](auto stv, auto) {
    using ST = decltype(stv);
    k_api_spo_vgl<ST><<<nw, WG, 0, stream>>>(s, static_cast<const ST*>(coefs), n, pos, out);
    return hipGetLastError();
  });
}

// SPOSet::mw_evaluateDetRatios (SplineR2R.cpp:204-335): value-only orbitals at virtual positions
// dotted with an inverse row.  Grid (nvp, nw).
template<typename ST>
__global__ __launch_bounds__(WG) void k_api_det_ratios(SysC s, const ST* __restrict__ coefs, int nvp, int n,
                                                       const double* __restrict__ pos, const double* __restrict__ inv,
                                                       double* __restrict__ out)
{
  __shared__ double red[WG];
  __shared__ double bc[1];
  const int v = blockIdx.x, w = blockIdx.y;
  const double* pp  = pos + ((size_t)w * nvp + v) * 3;
  const double p[3] = {pp[0], pp[1], pp[2]};
  SplineLoc L;
  spline_locate(s, p, L);
  double acc[1] = {0};
  for (int o = threadIdx.x; o < n; o += WG)
    acc[0] += spline_v_orbital<ST>(s, coefs, L, o) * inv[(size_t)w * n + o];
  block_sum<1>(acc, red, bc);
  if (threadIdx.x == 0)
    out[(size_t)w * nvp + v] = bc[0];
}

hipError_t launch_api_det_ratios(const SysC& s, const void* coefs, int nw, int nvp, int n, const double* pos,
                                 const double* inv, double* out, hipStream_t stream)
{
  return dispatch_prec(s.prec, [&](auto stv, auto) {
    using ST = decltype(stv);
    k_api_det_ratios<ST><<<dim3(nvp, nw), WG, 0, stream>>>(s, static_cast<const ST*>(coefs), nvp, n, pos, inv, out);
    return hipGetLastError();
  });
}

// SPOSet::mw_evaluateVGLandDetRatioGrads (SplineR2R.cpp:414-580): orbital VGL at pos[w] written as
// phi[5][nw][n] (value, gradient x/y/z, Laplacian), the determinant ratio sum_o inv[w][o] phi_o and
// the gradient sum_o inv[w][o] dphi_o / ratio.  Grid nw.
template<typename ST>
__global__ __launch_bounds__(WG) void k_api_vgl_ratio_grads(SysC s, const ST* __restrict__ coefs, int nw, int n,
                                                            const double* __restrict__ pos,
                                                            const double* __restrict__ inv, double* __restrict__ phi,
                                                            double* __restrict__ rg)
{
  __shared__ double red[4 * WG];
  __shared__ double bc[4];
  const int w       = blockIdx.x;
  const double p[3] = {pos[w * 3], pos[w * 3 + 1], pos[w * 3 + 2]};
  SplineLoc L;
  spline_locate(s, p, L);
  double acc[4] = {0, 0, 0, 0};
  for (int o = threadIdx.x; o < n; o += WG)
  {
    double v[5];
    spline_vgl_orbital<ST>(s, coefs, L, o, v);
    if (phi)
      for (int q = 0; q < 5; q++)
        phi[((size_t)q * nw + w) * n + o] = v[q];
    const double a = inv[(size_t)w * n + o];
    for (int q = 0; q < 4; q++)
      acc[q] = fma(v[q], a, acc[q]);
  }
  block_sum<4>(acc, red, bc);
  if (threadIdx.x == 0)
  {
    rg[w * 4] = bc[0];
    for (int q = 1; q < 4; q++)
      rg[w * 4 + q] = bc[q] / bc[0];
  }
}

hipError_t launch_api_vgl_ratio_grads(const SysC& s, const void* coefs, int nw, int n, const double* pos,
                                      const double* inv, double* phi, double* rg, hipStream_t stream)
{
  return dispatch_prec(s.prec, [&](auto stv, auto) {
    using ST = decltype(stv);
    k_api_vgl_ratio_grads<ST><<<nw, WG, 0, stream>>>(s, static_cast<const ST*>(coefs), nw, n, pos, inv, phi, rg);
    return hipGetLastError();
  });
}

// DistanceTable::mw_move rows: electrons and ions relative to pos[w] (source - target)
__global__ void k_api_distances(SysC s, State st, const double* __restrict__ pos, double* __restrict__ ee_r,
                                double* __restrict__ ee_dr, double* __restrict__ ei_r, double* __restrict__ ei_dr)
{
  const int w  = blockIdx.x;
  const int Np = s.Npad;
  const double* R   = st.R + (size_t)w * 3 * Np;
  const double p[3] = {pos[w * 3], pos[w * 3 + 1], pos[w * 3 + 2]};
  for (int j = threadIdx.x; j < s.N; j += blockDim.x)
  {
    double dr[3];
    const double r = min_image(s, R[j], R[Np + j], R[2 * Np + j], p, dr);
    if (ee_r)
      ee_r[(size_t)w * s.N + j] = r;
    if (ee_dr)
      for (int q = 0; q < 3; q++)
        ee_dr[((size_t)w * s.N + j) * 3 + q] = dr[q];
  }
  for (int j = threadIdx.x; j < s.nion; j += blockDim.x)
  {
    double dr[3];
    const double r = min_image(s, st.ions[j], st.ions[s.nion + j], st.ions[2 * s.nion + j], p, dr);
    if (ei_r)
      ei_r[(size_t)w * s.nion + j] = r;
    if (ei_dr)
      for (int q = 0; q < 3; q++)
        ei_dr[((size_t)w * s.nion + j) * 3 + q] = dr[q];
  }
}

hipError_t launch_api_distances(const SysC& s, const State& st, int nw, const double* pos, double* ee_r,
                                double* ee_dr, double* ei_r, double* ei_dr, hipStream_t stream)
{
  k_api_distances<<<nw, 256, 0, stream>>>(s, st, pos, ee_r, ee_dr, ei_r, ei_dr);
  return hipGetLastError();
}

__global__ void k_fill_i32(int* __restrict__ p, size_t n, int v)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Non-local moves: acc[w] = chosen[w] and a non-zero ratio; count[0] += the accepted moves,
// count[1] += the chosen moves whose ratio is NaN (not accepted)
__global__ void k_tmove_accept_mask(int nw, const double* __restrict__ ratio, const int* __restrict__ chosen,
                                    int* __restrict__ acc, int* __restrict__ count)
{
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw)
    return;
  const double r = ratio[w];
  const bool nan = r != r;
  const int a    = chosen[w] != 0 && r != 0.0 && !nan;
  acc[w]         = a;
  if (a)
    atomicAdd(count, 1);
  if (chosen[w] != 0 && nan)
    atomicAdd(count + 1, 1);
}

hipError_t launch_tmove_accept_mask(int nw, const double* ratio, const int* chosen, int* acc, int* count,
                                    hipStream_t stream)
{
  k_tmove_accept_mask<<<(nw + 255) / 256, 256, 0, stream>>>(nw, ratio, chosen, acc, count);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t stream)
{
  if (n == 0)
    return hipSuccess;
  k_fill_i32<<<(unsigned)std::min<size_t>(1024, (n + 255) / 256), 256, 0, stream>>>(p, n, v);
  return hipGetLastError();
}

// Walker copies of DMC branching: workgroup (chunk, pair, array) copies 64 KB of one array's slice
// with four independent 16-byte loads per thread in flight
constexpr size_t COPY_CHUNK = 65536;
__global__ __launch_bounds__(WG) void k_copy_walkers(WalkerArrays t, const int* __restrict__ pairs, int npairs)
{
  const int p = blockIdx.y, a = blockIdx.z;
  if (p >= npairs || a >= t.n)
    return;
  const size_t bytes = t.bytes[a];
  const size_t c0    = (size_t)blockIdx.x * COPY_CHUNK;
  if (c0 >= bytes)
    return;
  const size_t n16 = (min(bytes - c0, COPY_CHUNK)) / 16;
  const uint4* src = reinterpret_cast<const uint4*>(t.base[a] + (size_t)pairs[2 * p] * bytes + c0);
  uint4* dst       = reinterpret_cast<uint4*>(t.base[a] + (size_t)pairs[2 * p + 1] * bytes + c0);
  size_t i = threadIdx.x;
  for (; i + 3 * WG < n16; i += 4 * WG)
  {
    const uint4 v0 = src[i], v1 = src[i + WG], v2 = src[i + 2 * WG], v3 = src[i + 3 * WG];
    dst[i]          = v0;
    dst[i + WG]     = v1;
    dst[i + 2 * WG] = v2;
    dst[i + 3 * WG] = v3;
  }
  for (; i < n16; i += WG)
    dst[i] = src[i];
}

hipError_t launch_copy_walkers(const WalkerArrays& t, const int* pairs, int npairs, hipStream_t stream)
{
  if (npairs <= 0)
    return hipSuccess;
  if (t.n < 1 || t.n > WA_MAX || npairs > 65535)
    return hipErrorInvalidValue;
  size_t maxb = 0;
  for (int a = 0; a < t.n; a++)
  {
    if (t.bytes[a] % 16 != 0 || reinterpret_cast<uintptr_t>(t.base[a]) % 16 != 0)
      return hipErrorInvalidValue;
    maxb = std::max(maxb, t.bytes[a]);
  }
  const unsigned chunks = (unsigned)((maxb + COPY_CHUNK - 1) / COPY_CHUNK);
  k_copy_walkers<<<dim3(chunks, npairs, t.n), WG, 0, stream>>>(t, pairs, npairs);
  return hipGetLastError();
}

} // namespace qmi
