// Stateful multi-walker distance tables: DistanceTable::mw_evaluate / mw_move / update /
// mw_updatePartial (Particle/DistanceTable.h:104-183) with the dense SoA semantics of
// SoaDistanceTableAA (Particle/SoaDistanceTableAA.h:45-127: the lower triangle, row i holds
// electrons j < i) and SoaDistanceTableAB (Particle/SoaDistanceTableAB.h:41-86: row i holds every
// ion), minimum image of the orthorhombic PPPO branch (Particle/Lattice/ParticleBConds3DSoa.h:110-139).
//
// Layout per walker (walker-major, as the rest of the walker state): table r[N][ns] and
// displacements dr[N][3][ns] (ns = N for AA, the ion count for AB; the AA upper triangle is not
// maintained, as in the reference), proposal rows temp_r[ns], temp_dr[3][ns] and, for AA, the rows of
// the current position old_r[ns], old_dr[3][ns].  Displacements are source - target.
#include "qmi_kernels.hpp"
#include <cfloat>

namespace qmi
{
namespace
{
constexpr int DT_WG = 256;

__device__ inline const double* dt_source(const SysC& s, const State& st, int kind, int w)
{
  return kind == DT_AA ? st.R + (size_t)w * 3 * s.Npad : st.ions;
}

// full table: grid (N, nw); row i of walker w from the current positions
__global__ __launch_bounds__(DT_WG) void k_dt_evaluate(SysC s, State st, DTableArrays t)
{
  const int i = blockIdx.x, w = blockIdx.y;
  const int ns = t.ns, sst = t.kind == DT_AA ? s.Npad : s.nion;
  const double* R   = st.R + (size_t)w * 3 * s.Npad;
  const double* src = dt_source(s, st, t.kind, w);
  const double p[3] = {R[i], R[s.Npad + i], R[2 * s.Npad + i]};
  const int last    = t.kind == DT_AA ? i : ns;
  double* r         = t.r + ((size_t)w * s.N + i) * ns;
  double* dr        = t.dr + ((size_t)w * s.N + i) * 3 * ns;
  for (int j = threadIdx.x; j < last; j += DT_WG)
  {
    double d[3];
    r[j] = min_image(s, src[j], src[sst + j], src[2 * sst + j], p, d);
    for (int q = 0; q < 3; q++)
      dr[q * ns + j] = d[q];
  }
}

// proposal rows from rnew[w] (SoaDistanceTableAA.h:57-75, SoaDistanceTableAB.h:60-71); grid nw
__global__ __launch_bounds__(DT_WG) void k_dt_move(SysC s, State st, DTableArrays t, int iat,
                                                   const double* __restrict__ rnew, int prepare_old)
{
  const int w  = blockIdx.x;
  const int ns = t.ns, sst = t.kind == DT_AA ? s.Npad : s.nion;
  const double* R   = st.R + (size_t)w * 3 * s.Npad;
  const double* src = dt_source(s, st, t.kind, w);
  const double pn[3] = {rnew[w * 3], rnew[w * 3 + 1], rnew[w * 3 + 2]};
  const double po[3] = {R[iat], R[s.Npad + iat], R[2 * s.Npad + iat]};
  double* tr  = t.temp_r + (size_t)w * ns;
  double* tdr = t.temp_dr + (size_t)w * 3 * ns;
  for (int j = threadIdx.x; j < ns; j += DT_WG)
  {
    double d[3];
    tr[j] = min_image(s, src[j], src[sst + j], src[2 * sst + j], pn, d);
    for (int q = 0; q < 3; q++)
      tdr[q * ns + j] = d[q];
    if (prepare_old)
    {
      // AA: the old row with a large number at the moving particle; AB: the table row itself is
      // rewritten from the current position (no consumer needs the full table at all times)
      double* orow  = t.kind == DT_AA ? t.old_r + (size_t)w * ns : t.r + ((size_t)w * s.N + iat) * ns;
      double* odrow = t.kind == DT_AA ? t.old_dr + (size_t)w * 3 * ns : t.dr + ((size_t)w * s.N + iat) * 3 * ns;
      const double ro = min_image(s, src[j], src[sst + j], src[2 * sst + j], po, d);
      orow[j]         = (t.kind == DT_AA && j == iat) ? DBL_MAX : ro;
      for (int q = 0; q < 3; q++)
        odrow[q * ns + j] = d[q];
    }
  }
}

// update (accepted walkers; partial = 0) or updatePartial (from temp where accepted, else from the
// old row; partial = 1): SoaDistanceTableAA.h:80-120, SoaDistanceTableAB.h:74-80.  Grid nw.
__global__ __launch_bounds__(DT_WG) void k_dt_update(SysC s, DTableArrays t, int iat, const int* __restrict__ acc,
                                                     int partial)
{
  const int w  = blockIdx.x;
  const int ns = t.ns;
  const bool from_temp = acc[w] != 0;
  if (!from_temp && (!partial || t.kind == DT_AB))
    return;
  const double* sr  = (from_temp ? t.temp_r : t.old_r) + (size_t)w * ns;
  const double* sdr = (from_temp ? t.temp_dr : t.old_dr) + (size_t)w * 3 * ns;
  const int nrow    = t.kind == DT_AA ? iat : ns;
  double* r         = t.r + ((size_t)w * s.N + iat) * ns;
  double* dr        = t.dr + ((size_t)w * s.N + iat) * 3 * ns;
  for (int j = threadIdx.x; j < nrow; j += DT_WG)
  {
    r[j] = sr[j];
    for (int q = 0; q < 3; q++)
      dr[q * ns + j] = sdr[q * ns + j];
  }
  // the column below the diagonal (regular-mode AA update only)
  if (t.kind == DT_AA && !partial)
    for (int i = iat + 1 + threadIdx.x; i < s.N; i += DT_WG)
    {
      double* ri  = t.r + ((size_t)w * s.N + i) * ns;
      double* dri = t.dr + ((size_t)w * s.N + i) * 3 * ns;
      ri[iat]     = sr[i];
      for (int q = 0; q < 3; q++)
        dri[q * ns + iat] = -sdr[q * ns + i];
    }
}
} // namespace

hipError_t launch_dt_evaluate(const SysC& s, const State& st, const DTableArrays& t, int nw, hipStream_t stream)
{
  if (nw <= 0 || s.N <= 0 || t.ns <= 0)
    return hipErrorInvalidValue;
  k_dt_evaluate<<<dim3(s.N, nw), DT_WG, 0, stream>>>(s, st, t);
  return hipGetLastError();
}

hipError_t launch_dt_move(const SysC& s, const State& st, const DTableArrays& t, int nw, int iat, const double* rnew,
                          int prepare_old, hipStream_t stream)
{
  if (nw <= 0 || iat < 0 || iat >= s.N || t.ns <= 0)
    return hipErrorInvalidValue;
  k_dt_move<<<nw, DT_WG, 0, stream>>>(s, st, t, iat, rnew, prepare_old);
  return hipGetLastError();
}

hipError_t launch_dt_update(const SysC& s, const DTableArrays& t, int nw, int iat, const int* acc, int partial,
                            hipStream_t stream)
{
  if (nw <= 0 || iat < 0 || iat >= s.N || t.ns <= 0)
    return hipErrorInvalidValue;
  k_dt_update<<<nw, DT_WG, 0, stream>>>(s, t, iat, acc, partial);
  return hipGetLastError();
}

} // namespace qmi
