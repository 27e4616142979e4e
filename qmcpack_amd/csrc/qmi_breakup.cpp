// Optimized long-range breakup of 1/r (host setup). The basis, the k points of the fit and the
// objective follow the reference (LPQHIBasis.h:65-86, 149-166; LRBreakup.h:71-186, 316-386;
// LRHandlerTemp.h:195-227; LRCoulombSingleton.cpp:32-56). The integrals and the solver are this
// file's own: Fourier integrals of each polynomial piece by Gauss-Legendre quadrature where the
// piece spans under 2.5 radians of k r and by the closed-form integration-by-parts antiderivative
// above; r^2 integrals by Gauss-Legendre (exact for these degrees); the normal equations solved
// through a cyclic Jacobi eigendecomposition with directions below eps * max |lambda| dropped
// (for the symmetric positive semi-definite normal matrix these are its singular directions).
#include "qmi_breakup.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>

namespace qmi
{
namespace
{

// shape polynomials in x = |r - r_i| / delta: value, slope and curvature functions (the Hermite
// quintics with unit value / first / second derivative at x = 0 and all three zero at x = 1)
constexpr double SHAPE[3][6] = {{1.0, 0.0, 0.0, -10.0, 15.0, -6.0},
                                {0.0, 1.0, 0.0, -6.0, 8.0, -3.0},
                                {0.0, 0.0, 0.5, -1.5, 1.5, -0.5}};
// the slope function is odd about its knot, value and curvature functions even
constexpr double MIRROR[3] = {1.0, -1.0, 1.0};

// n-point Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n from the Chebyshev-like guess)
struct GaussLegendre
{
  std::vector<double> x, w;
  explicit GaussLegendre(int n) : x(n), w(n)
  {
    for (int i = 0; i < n; i++)
    {
      double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
      for (int it = 0; it < 100; it++)
      {
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; k++)
        {
          const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
          p0              = p1;
          p1              = p2;
        }
        dp              = n * (z * p1 - p0) / (z * z - 1.0);
        const double dz = p1 / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-16)
          break;
      }
      x[i] = z;
      w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
  }
};

const GaussLegendre& gl16()
{
  static const GaussLegendre g(16);
  return g;
}

// One polynomial piece of basis function n: r in [r0, r1] around knot r_i, where
// h(r) = m * sum_j S_j (s (r - r_i) / delta)^j; as a polynomial in u = r - r_i its coefficients
// are hc[j] = m S_j s^j / delta^j.
struct Piece
{
  double r0, r1, ri;
  double hc[6];
};

int pieces(const LrBreakup& b, int n, Piece out[2])
{
  const int i = n / 3, kind = n % 3;
  const double ri = b.delta * i;
  int np          = 0;
  for (int side = -1; side <= 1; side += 2)
  {
    const double lo = side < 0 ? std::max(0.0, ri - b.delta) : ri;
    const double hi = side < 0 ? ri : std::min(b.rc, ri + b.delta);
    if (!(hi > lo))
      continue;
    Piece& p = out[np++];
    p.r0     = lo;
    p.r1     = hi;
    p.ri     = ri;
    const double m = side < 0 ? MIRROR[kind] : 1.0;
    double sc      = 1.0;
    for (int j = 0; j < 6; j++)
    {
      p.hc[j] = m * SHAPE[kind][j] * sc;
      sc *= side / b.delta;
    }
  }
  return np;
}

double horner(const double* c, int deg, double u)
{
  double v = c[deg];
  for (int j = deg - 1; j >= 0; j--)
    v = v * u + c[j];
  return v;
}

// int over the piece of r h(r) sin(k r) dr
double piece_sine_integral(const Piece& p, double k)
{
  // p(u) = (r_i + u) h(u), degree 6 in u
  double pc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 6; j++)
  {
    pc[j] += p.ri * p.hc[j];
    pc[j + 1] += p.hc[j];
  }
  const double u0 = p.r0 - p.ri, u1 = p.r1 - p.ri;
  if (k * (u1 - u0) < 2.5)
  {
    const GaussLegendre& g = gl16();
    const double half = 0.5 * (u1 - u0), mid = 0.5 * (u1 + u0);
    double s = 0;
    for (size_t q = 0; q < g.x.size(); q++)
    {
      const double u = mid + half * g.x[q];
      s += g.w[q] * horner(pc, 6, u) * std::sin(k * (p.ri + u));
    }
    return half * s;
  }
  // antiderivative sum_m p^(m)(u) T_m(k r) / k^(m+1) with T = -cos, sin, cos, -sin repeating
  auto G = [&](double u) {
    double d[7];
    std::copy(pc, pc + 7, d);
    const double c = std::cos(k * (p.ri + u)), s = std::sin(k * (p.ri + u));
    double acc = 0, kp = 1.0 / k;
    for (int m = 0; m <= 6; m++)
    {
      const double v = horner(d, 6 - m, u);
      switch (m & 3)
      {
      case 0: acc -= v * c * kp; break;
      case 1: acc += v * s * kp; break;
      case 2: acc += v * c * kp; break;
      default: acc -= v * s * kp; break;
      }
      kp /= k;
      for (int j = 0; j < 6 - m; j++) // differentiate
        d[j] = (j + 1) * d[j + 1];
    }
    return acc;
  };
  return G(u1) - G(u0);
}

// eigenpairs of the symmetric n x n matrix a (row-major, destroyed) by cyclic Jacobi rotations:
// eigenvalues in lam, eigenvectors in the columns of v
void jacobi_eigen(int n, std::vector<double>& a, std::vector<double>& lam, std::vector<double>& v)
{
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    v[(size_t)i * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  for (int sweep = 0; sweep < 100; sweep++)
  {
    bool rotated = false;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++)
      {
        const double apq = A(p, q);
        if (apq == 0.0)
          continue;
        // negligible against the geometric mean of the two diagonal entries: drop it
        if (std::fabs(apq) <= 1e-3 * DBL_EPSILON * std::sqrt(std::fabs(A(p, p) * A(q, q))))
        {
          A(p, q) = A(q, p) = 0.0;
          continue;
        }
        rotated            = true;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t     = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) // columns p, q
        {
          const double akp = A(k, p), akq = A(k, q);
          A(k, p)          = c * akp - s * akq;
          A(k, q)          = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) // rows p, q
        {
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k)          = c * apk - s * aqk;
          A(q, k)          = s * apk + c * aqk;
        }
        A(p, q) = A(q, p) = 0.0;
        for (int k = 0; k < n; k++)
        {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
    if (!rotated)
      break;
  }
  lam.resize(n);
  for (int i = 0; i < n; i++)
    lam[i] = A(i, i);
}

} // namespace

double LrBreakup::h(int n, double r) const
{
  const int i = n / 3, kind = n % 3;
  const double ri = delta * i;
  if (r < ri - delta || r > std::min(rc, ri + delta))
    return 0.0;
  const double x = std::fabs(r - ri) / delta;
  const double v = horner(SHAPE[kind], 5, x);
  return r <= ri ? MIRROR[kind] * v : v;
}

double LrBreakup::r2_integral(int n) const
{
  Piece ps[2];
  const int np = pieces(*this, n, ps);
  const GaussLegendre& g = gl16();
  double s = 0;
  for (int k = 0; k < np; k++)
  {
    const double half = 0.5 * (ps[k].r1 - ps[k].r0), mid = 0.5 * (ps[k].r1 + ps[k].r0);
    for (size_t q = 0; q < g.x.size(); q++)
    {
      const double r = mid + half * g.x[q];
      s += half * g.w[q] * r * r * horner(ps[k].hc, 5, r - ps[k].ri);
    }
  }
  return s;
}

double LrBreakup::fourier(int n, double k) const
{
  Piece ps[2];
  const int np = pieces(*this, n, ps);
  double s     = 0;
  for (int q = 0; q < np; q++)
    s += piece_sine_integral(ps[q], k);
  return 4.0 * M_PI / (k * volume) * s;
}

double LrBreakup::fk(double k) const
{
  double f = 4.0 * M_PI / (volume * k * k) * std::cos(k * rc);
  for (size_t n = 0; n < t.size(); n++)
    f += t[n] * fourier((int)n, k);
  return f;
}

double LrBreakup::vlr_r0() const
{
  double v = 0;
  for (size_t n = 0; n < t.size(); n++)
    v += t[n] * h((int)n, 0.0);
  return v;
}

double LrBreakup::vsr_k0() const
{
  double v = 0.5 * rc * rc;
  for (size_t n = 0; n < t.size(); n++)
    v -= t[n] * r2_integral((int)n);
  return v * 4.0 * M_PI / volume;
}

double LrBreakup::rvsr(double r) const
{
  if (r >= rc)
    return 0.0;
  double vl = 0;
  // only the (at most six) functions of the two knots around r are nonzero
  const int i0 = std::max(0, (int)std::floor(r / delta) - 1), i1 = std::min(nknots - 1, i0 + 3);
  for (int i = i0; i <= i1; i++)
    for (int kind = 0; kind < 3; kind++)
      vl += t[3 * i + kind] * h(3 * i + kind, r);
  return 1.0 - r * vl;
}

LrBreakup lr_breakup(const double L[3], double dim_cutoff)
{
  if (!(L[0] > 0 && L[1] > 0 && L[2] > 0) || !(dim_cutoff > 0))
    throw std::runtime_error("breakup: cell sides and dim_cutoff must be positive");
  LrBreakup b;
  std::copy(L, L + 3, b.L);
  b.volume = L[0] * L[1] * L[2];
  b.rc     = 0.5 * std::min(L[0], std::min(L[1], L[2])); // Wigner-Seitz radius, orthorhombic cell
  b.kc     = dim_cutoff / b.rc;
  b.delta  = b.rc / (b.nknots - 1);
  const int nb = 3 * b.nknots;

  // lattice shells up to kcont, keyed by floor(|k|^2 * 1e7) in ascending order, each with the
  // |k|^2 of its first vector in enumeration order and its size
  const double kcont = 60.0 * M_PI * std::pow(b.volume, -1.0 / 3.0), kmax = 6000.0 / b.rc;
  int mmax[3];
  for (int d = 0; d < 3; d++)
    mmax[d] = (int)std::floor(L[d] * kcont / (2 * M_PI)) + 1;
  std::map<int64_t, std::pair<double, int>> shells;
  for (int i = -mmax[0]; i <= mmax[0]; i++)
    for (int j = -mmax[1]; j <= mmax[1]; j++)
      for (int k = -mmax[2]; k <= mmax[2]; k++)
      {
        if (i == 0 && j == 0 && k == 0)
          continue;
        const double x = 2 * M_PI * i / L[0], y = 2 * M_PI * j / L[1], z = 2 * M_PI * k / L[2];
        const double k2 = x * x + y * y + z * z;
        if (k2 > kcont * kcont)
          continue;
        auto it = shells.emplace((int64_t)(k2 * 1e7), std::make_pair(k2, 0)).first;
        it->second.second++;
      }
  if (shells.empty())
    throw std::runtime_error("breakup: no lattice k vector below the continuum cutoff");
  // the long-range sum keeps at least the first shell
  b.kc2_lr = std::max(b.kc * b.kc, shells.begin()->second.first);
  std::vector<double> kf, wf;
  for (const auto& s : shells)
    if (s.second.first > b.kc2_lr)
    {
      kf.push_back(std::sqrt(s.second.first));
      wf.push_back((double)s.second.second);
    }
  // continuum shells of equal width from kcont to kmax at their mid |k|, weighted by volume over
  // the volume per k point (2 pi)^3 / V; all lie above every lattice shell, so none coincides
  const int ncont = 4000;
  const double kel = 8.0 * M_PI * M_PI * M_PI / b.volume, dk = (kmax - kcont) / ncont;
  for (int i = 0; i < ncont; i++)
  {
    const double k1 = kcont + dk * i, k2 = k1 + dk;
    kf.push_back(0.5 * (k1 + k2));
    wf.push_back(4.0 * M_PI * (k2 * k2 * k2 - k1 * k1 * k1) / 3.0 / kel);
  }
  b.nfit = (int)kf.size();

  // normal equations of min sum_k w_k (X_k - sum_n t_n c_n(k))^2 with X_k = -4 pi cos(k r_c) / (V k^2)
  // (the transform of 1/r beyond r_c, negated), accumulated k by k in list order
  std::vector<double> cnk((size_t)nb), A((size_t)nb * nb, 0.0), rhs(nb, 0.0), xk(kf.size());
  std::vector<double> call((size_t)nb * kf.size());
  for (size_t q = 0; q < kf.size(); q++)
  {
    const double k = kf[q];
    xk[q]          = -4.0 * M_PI / (b.volume * k * k) * std::cos(k * b.rc);
    for (int n = 0; n < nb; n++)
      cnk[n] = call[q * nb + n] = b.fourier(n, k);
    for (int l = 0; l < nb; l++)
    {
      const double wl = wf[q] * cnk[l];
      rhs[l] += wl * xk[q];
      for (int n = 0; n < nb; n++)
        A[(size_t)l * nb + n] += wl * cnk[n];
    }
  }
  std::vector<double> lam, V;
  jacobi_eigen(nb, A, lam, V);
  double lmax = 0;
  for (double l : lam)
    lmax = std::max(lmax, std::fabs(l));
  if (!(lmax > 0))
    throw std::runtime_error("breakup: singular normal equations");
  b.t.assign(nb, 0.0);
  for (int i = 0; i < nb; i++)
  {
    if (std::fabs(lam[i]) < DBL_EPSILON * lmax)
      continue;
    double proj = 0;
    for (int j = 0; j < nb; j++)
      proj += V[(size_t)j * nb + i] * rhs[j];
    proj /= lam[i];
    for (int j = 0; j < nb; j++)
      b.t[j] += proj * V[(size_t)j * nb + i];
  }
  double chi2 = 0;
  for (size_t q = 0; q < kf.size(); q++)
  {
    double y = xk[q];
    for (int n = 0; n < nb; n++)
      y -= call[q * nb + n] * b.t[n];
    chi2 += wf[q] * y * y;
  }
  b.chi2 = chi2;
  return b;
}

} // namespace qmi
