// Optimized long-range breakup of the Coulomb potential on a piecewise-quintic Hermite basis
// (the reference's default long-range handler, "opt_breakup": Particle/LongRange/LRHandlerTemp.h
// with LPQHIBasis.{h,cpp} and LRBreakup.h). Host-side setup: it produces the tables that the
// device's Coulomb kernel reads (F_k per k vector, r V_s(r) on the spline grid, the constants).
//
// 1/r = V_l(r) + V_s(r) with V_l(r) = sum_n t_n h_n(r) below r_c and 1/r beyond; the t_n minimise
// the weighted misfit of V_l's Fourier transform over k > k_c (lattice shells up to 60 pi V^-1/3,
// then 4000 continuum shells up to 6000 / r_c), so the k-space sum stops at k_c.
#pragma once
#include <vector>

namespace qmi
{

struct LrBreakup
{
  double L[3]   = {0, 0, 0};
  double volume = 0, rc = 0, kc = 0, delta = 0;
  int nknots    = 15;      // LRHandlerTemp.h:197
  std::vector<double> t;   // 3 * nknots coefficients, n = 3 * knot + kind (value, slope, curvature)
  double chi2 = 0;         // weighted misfit of the fit
  int nfit    = 0;         // k points of the fit
  double kc2_lr = 0;       // |k|^2 bound of the long-range sum: max(k_c^2, first lattice shell)

  // basis function n at r (LPQHIBasis.h:65-86)
  double h(int n, double r) const;
  // int_0^rc r^2 h_n(r) dr
  double r2_integral(int n) const;
  // c_n(k) = 4 pi / (k V) int_0^rc r sin(k r) h_n(r) dr
  double fourier(int n, double k) const;
  // F_k of the long-range sum: 4 pi cos(k r_c) / (V k^2) + sum_n t_n c_n(k) (LRHandlerTemp.h:253-259)
  double fk(double k) const;
  // v_l(r = 0) and v_s(k = 0) of the constant terms (LRHandlerTemp.h:153-167)
  double vlr_r0() const;
  double vsr_k0() const;
  // r V_s(r) = 1 - r V_l(r) below r_c, 0 beyond (LRHandlerTemp.h:98-108)
  double rvsr(double r) const;
};

// The fitted breakup of an orthorhombic cell with sides L (r_c = min(L) / 2, k_c = dim_cutoff / r_c).
// Throws std::runtime_error when the fit finds no usable singular direction.
LrBreakup lr_breakup(const double L[3], double dim_cutoff);

} // namespace qmi
