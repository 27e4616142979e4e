// Host random numbers of the VMC driver.  Each crowd owns a std::mt19937 engine; a uniform is one
// 32-bit draw scaled into [0, 1), and normal deviates come in pairs from the Box-Muller transform
// (G. E. P. Box and M. E. Muller, Ann. Math. Statist. 29 (1958) 610-611): radius
// sqrt(-2 ln(1 - u1)), angle 2 pi u2, the pair (radius cos angle, radius sin angle).
//
// The reference's streams fix the operation order, which is kept here so that a chain can be
// replayed against them number for number: the uniform divides (draw - min) by (max - min + 1)
// and then applies the [0, 1) affine map (Utilities/StdRandom.h:26-44); the transform scales u1
// by the largest double below one before taking 1 - u1 (so the logarithm never sees zero), fills
// pairs in order, and an odd last entry takes the cosine term of one more pair
// (Particle/ParticleBase/RandomSeqGenerator.h:31-50).
//
// Compiled by the host C++ compiler with floating-point contraction disabled, so the sequence does
// not depend on FMA fusion choices of the compiler.
#include "qmi_rng.hpp"

#include <cmath>
#include <limits>

namespace qmi
{
double rng_uniform01(std::mt19937& e)
{
  constexpr double lo = 0.0, hi = 1.0;
  const double span   = static_cast<double>(e.max() - e.min()) + 1;
  const double draw   = static_cast<double>(e() - e.min());
  return draw / span * (hi - lo) + lo;
}

void rng_fill_uniform(double* a, size_t n, std::mt19937& e)
{
  for (size_t i = 0; i < n; i++)
    a[i] = rng_uniform01(e);
}

namespace
{
// radius of a Box-Muller pair from its first uniform
inline double bm_radius(double u)
{
  const double below_one = 1.0 - std::numeric_limits<double>::epsilon();
  return std::sqrt(-2.0 * std::log(1.0 - below_one * u));
}
inline double bm_angle(double u) { return 2.0 * M_PI * u; }
} // namespace

void rng_fill_gauss(double* a, unsigned n, std::mt19937& e)
{
  const int pairs_end = int(n) - 1; // pairs fill [0, n - 1) two at a time
  for (int i = 0; i < pairs_end; i += 2)
  {
    const double rad = bm_radius(rng_uniform01(e));
    const double ang = bm_angle(rng_uniform01(e));
    a[i]             = rad * std::cos(ang);
    a[i + 1]         = rad * std::sin(ang);
  }
  if (n % 2 == 1)
  {
    const double rad = bm_radius(rng_uniform01(e));
    const double ang = bm_angle(rng_uniform01(e));
    a[n - 1]         = rad * std::cos(ang);
  }
}
} // namespace qmi
