// Periodic Coulomb energy of every walker with the Ewald breakup: CoulombPBCAA (electron-electron)
// and CoulombPBCAB (electron-ion) of QMCHamiltonians/CoulombPBCAA.cpp:568-840 and
// CoulombPBCAB.cpp:373-453 with the handler of Particle/LongRange/EwaldHandler3D.{h,cpp}:
//   V = sum_{i<j} q_i q_j v_sr(r_ij) + 1/2 sum_k F_k |rho_k|^2 + constants,
// v_sr(r) = erfc(sigma r) / r through a clamped cubic spline of r v_sr on [0, r_c] (zero beyond),
// F_k = 4 pi / V exp(-k^2 / 4 sigma^2) / k^2 over |k| <= k_c.  The host builds the tables
// (qmi_host.cpp: coulomb_setup); one workgroup per walker evaluates its energies.
#include "qmi_kernels.hpp"

namespace qmi
{

// v_sr(r) for r^2 = r2 < r_c^2: the spline interval's cubic in t = (r - i dr) / dr, divided by r
__device__ inline double ewald_vsr(const CoulombC& c, const double4* __restrict__ sp, double r2)
{
  const double r    = sqrt(r2);
  const int loc     = min((int)(r * c.drinv), c.ng - 2); // r < r_c; rounding can reach ng - 1
  const double t    = (r - loc * c.dr) * c.drinv;
  const double4 q   = sp[loc];
  return (q.x + t * (q.y + t * (q.z + t * q.w))) / r;
}

__device__ inline double min_image2(const SysC& s, double ax, double ay, double az, double bx, double by, double bz)
{
  const double d[3] = {ax - bx, ay - by, az - bz};
  double r2         = 0;
#pragma unroll
  for (int q = 0; q < 3; q++)
  {
    const double u  = d[q] * s.Linv[q];
    const double dd = s.L[q] * (u - rint(u));
    r2 += dd * dd;
  }
  return r2;
}

// out[w][3] = {V_ee, V_eI, V_II}; electrons carry charge -1, ions c.zion.  LDS: electron positions
// (N <= 1024) and the spline (ng <= COUL_MAX_NG).  Periodic cells only (the host checks).
__global__ __launch_bounds__(WG) void k_coulomb(SysC s, State st, CoulombC c, double* __restrict__ out)
{
  __shared__ double px[1024], py[1024], pz[1024];
  __shared__ double4 sp[COUL_MAX_NG - 1];
  __shared__ double red[4 * WG];
  __shared__ double bc[4];
  const int w = blockIdx.x, t = threadIdx.x;
  const int N = s.N, Np = s.Npad, ni = s.nion;
  const double* R = st.R + (size_t)w * 3 * Np;
  for (int i = t; i < N; i += WG)
  {
    px[i] = R[i];
    py[i] = R[Np + i];
    pz[i] = R[2 * Np + i];
  }
  const double4* gp = reinterpret_cast<const double4*>(c.poly);
  for (int i = t; i < c.ng - 1; i += WG)
    sp[i] = gp[i];
  __syncthreads();
  const double rc2 = c.rmax * c.rmax;
  double acc[4]    = {0, 0, 0, 0}; // short range e-e, e-I; long range e-e, e-I
  // e-e pairs j < i, rows i and N - 1 - i together (equal work per thread)
  const int half = (N + 1) / 2;
  for (int rr = t; rr < half; rr += WG)
  {
    for (int pass = 0; pass < 2; pass++)
    {
      const int i = pass == 0 ? rr : N - 1 - rr;
      if (pass == 1 && i == rr)
        break;
      const double xi = px[i], yi = py[i], zi = pz[i];
      double esum     = 0;
      for (int j = 0; j < i; j++)
      {
        const double r2 = min_image2(s, px[j], py[j], pz[j], xi, yi, zi);
        if (r2 < rc2)
          esum += ewald_vsr(c, sp, r2);
      }
      acc[0] += esum; // q_i q_j = 1
    }
  }
  // e-I pairs
  for (int i = t; i < N; i += WG)
  {
    const double xi = px[i], yi = py[i], zi = pz[i];
    double esum     = 0;
    for (int j = 0; j < ni; j++)
    {
      const double r2 = min_image2(s, st.ions[j], st.ions[ni + j], st.ions[2 * ni + j], xi, yi, zi);
      if (r2 < rc2)
        esum += ewald_vsr(c, sp, r2);
    }
    acc[1] += -c.zion * esum;
  }
  // long range: rho_k of the electrons (charge -1) against itself and the ions
  for (int k = t; k < c.nk; k += WG)
  {
    const double kx = c.kvec[k], ky = c.kvec[c.nk + k], kz = c.kvec[2 * c.nk + k];
    double re = 0, im = 0;
    for (int i = 0; i < N; i++)
    {
      double sn, cs;
      sincos(kx * px[i] + ky * py[i] + kz * pz[i], &sn, &cs);
      re -= cs;
      im -= sn;
    }
    acc[2] += c.fk[k] * (re * re + im * im);
    acc[3] += c.fk[k] * (re * c.rhoi[k] + im * c.rhoi[c.nk + k]);
  }
  block_sum<4>(acc, red, bc);
  if (t == 0)
  {
    out[(size_t)w * 3 + 0] = bc[0] + 0.5 * bc[2] + c.const_ee;
    out[(size_t)w * 3 + 1] = bc[1] + bc[3] + c.const_ei;
    out[(size_t)w * 3 + 2] = c.vii;
  }
}

// Structure factor of each spin group for nk wave vectors (StructFact::computeRhok,
// Particle/LongRange/StructFact.cpp:138-205): rhok[w][g][k] = (sum over the electrons i of group g,
// in order, of cos(k . r_i), and of sin(k . r_i)).  Grid (ceil(nk / WG), nw): one thread per k,
// the walker's positions staged in LDS.
__global__ __launch_bounds__(WG) void k_rhok(SysC s, State st, int nk, const double* __restrict__ kv,
                                             double* __restrict__ out)
{
  __shared__ double sx[COUL_MAX_N], sy[COUL_MAX_N], sz[COUL_MAX_N];
  const int w = blockIdx.y, t = threadIdx.x;
  const int N = s.N, Np = s.Npad;
  const double* R = st.R + (size_t)w * 3 * Np;
  for (int i = t; i < N; i += WG)
  {
    sx[i] = R[i];
    sy[i] = R[Np + i];
    sz[i] = R[2 * Np + i];
  }
  __syncthreads();
  const int k = blockIdx.x * WG + t;
  if (k >= nk)
    return;
  const double kx = kv[3 * (size_t)k], ky = kv[3 * (size_t)k + 1], kz = kv[3 * (size_t)k + 2];
  for (int g = 0; g < 2; g++)
  {
    double re = 0, im = 0;
    const int i0 = s.first[g], i1 = s.first[g] + s.n[g];
    for (int i = i0; i < i1; i++)
    {
      double sn, cs;
      sincos(kx * sx[i] + ky * sy[i] + kz * sz[i], &sn, &cs);
      re += cs;
      im += sn;
    }
    double* o = out + (((size_t)w * 2 + g) * nk + k) * 2;
    o[0]      = re;
    o[1]      = im;
  }
}

hipError_t launch_rhok(const SysC& s, const State& st, int nw, int nk, const double* kv, double* out, hipStream_t stream)
{
  if (s.N > COUL_MAX_N || nk <= 0 || nw <= 0 || nw > 65535)
    return hipErrorInvalidValue;
  const dim3 grid((nk + WG - 1) / WG, nw);
  k_rhok<<<grid, WG, 0, stream>>>(s, st, nk, kv, out);
  return hipGetLastError();
}

hipError_t launch_coulomb(const SysC& s, const State& st, const CoulombC& c, int nw, double* out, hipStream_t stream)
{
  if (s.N > 1024 || c.ng > COUL_MAX_NG || c.ng < 2)
    return hipErrorInvalidValue;
  k_coulomb<<<nw, WG, 0, stream>>>(s, st, c, out);
  return hipGetLastError();
}

} // namespace qmi
