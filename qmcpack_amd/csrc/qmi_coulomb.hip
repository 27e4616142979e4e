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

// v_sr(r) from the spline values y / second derivatives y2 on the grid of spacing dr (ng points)
__device__ inline double ewald_vsr(const CoulombC& c, const double* __restrict__ y, const double* __restrict__ y2,
                                   double r)
{
  if (r >= c.rmax)
    return 0.0;
  const int loc     = (int)(r * c.drinv);
  const double dist = r - loc * c.dr;
  const double cl = dist * c.drinv, cr = (c.dr - dist) * c.drinv, h6 = c.dr / 6.0;
  const double q1 = cr * (cr * cr - 1.0) * h6 * c.dr, q2 = cl * (cl * cl - 1.0) * h6 * c.dr;
  return (cr * y[loc] + cl * y[loc + 1] + q1 * y2[loc] + q2 * y2[loc + 1]) / r;
}

// out[w][3] = {V_ee, V_eI, V_II}; electrons carry charge -1, ions c.zion.  LDS: electron positions
// (N <= 1024) and the spline (ng <= COUL_MAX_NG).
__global__ __launch_bounds__(WG) void k_coulomb(SysC s, State st, CoulombC c, double* __restrict__ out)
{
  __shared__ double px[1024], py[1024], pz[1024];
  __shared__ double sy[COUL_MAX_NG], sy2[COUL_MAX_NG];
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
  for (int i = t; i < c.ng; i += WG)
  {
    sy[i]  = c.y[i];
    sy2[i] = c.y2[i];
  }
  __syncthreads();
  double acc[4] = {0, 0, 0, 0}; // short range e-e, e-I; long range e-e, e-I
  // e-e pairs j < i, rows i and N - 1 - i together (equal work per thread)
  const int half = (N + 1) / 2;
  for (int rr = t; rr < half; rr += WG)
  {
    for (int pass = 0; pass < 2; pass++)
    {
      const int i = pass == 0 ? rr : N - 1 - rr;
      if (pass == 1 && i == rr)
        break;
      const double ri[3] = {px[i], py[i], pz[i]};
      double esum        = 0;
      for (int j = 0; j < i; j++)
      {
        double dr[3];
        const double r = min_image(s, px[j], py[j], pz[j], ri, dr);
        esum += ewald_vsr(c, sy, sy2, r);
      }
      acc[0] += esum; // q_i q_j = 1
    }
  }
  // e-I pairs
  for (int i = t; i < N; i += WG)
  {
    const double ri[3] = {px[i], py[i], pz[i]};
    double esum        = 0;
    for (int j = 0; j < ni; j++)
    {
      double dr[3];
      const double r = min_image(s, st.ions[j], st.ions[ni + j], st.ions[2 * ni + j], ri, dr);
      esum += ewald_vsr(c, sy, sy2, r);
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

hipError_t launch_coulomb(const SysC& s, const State& st, const CoulombC& c, int nw, double* out, hipStream_t stream)
{
  if (s.N > 1024 || c.ng > COUL_MAX_NG)
    return hipErrorInvalidValue;
  k_coulomb<<<nw, WG, 0, stream>>>(s, st, c, out);
  return hipGetLastError();
}

} // namespace qmi
