// Host random numbers of the VMC driver, in the reference's order and arithmetic:
// std::mt19937 -> boost-style uniform in [0,1) (Utilities/StdRandom.h:26-85) and Box-Muller
// Gaussians (Particle/ParticleBase/RandomSeqGenerator.h:31-50).
//
// Compiled by the host C++ compiler with floating-point contraction disabled, so the
// sequence does not depend on FMA fusion choices of the compiler.
#include "qmi_rng.hpp"

#include <cmath>
#include <limits>

namespace qmi
{
double rng_uniform01(std::mt19937& e)
{
  return static_cast<double>(e() - e.min()) / (static_cast<double>(e.max() - e.min()) + 1) * (1.0 - 0.0) + 0.0;
}

void rng_fill_uniform(double* a, size_t n, std::mt19937& e)
{
  for (size_t i = 0; i < n; i++)
    a[i] = rng_uniform01(e);
}

void rng_fill_gauss(double* a, unsigned n, std::mt19937& e)
{
  const double slightly_less_than_one = 1.0 - std::numeric_limits<double>::epsilon();
  const int nm1                       = int(n) - 1;
  double temp1, temp2;
  for (int i = 0; i < nm1; i += 2)
  {
    temp1    = std::sqrt(-2.0 * std::log(1.0 - slightly_less_than_one * rng_uniform01(e)));
    temp2    = 2.0 * M_PI * rng_uniform01(e);
    a[i]     = temp1 * std::cos(temp2);
    a[i + 1] = temp1 * std::sin(temp2);
  }
  if (n % 2 == 1)
  {
    temp1  = std::sqrt(-2.0 * std::log(1.0 - slightly_less_than_one * rng_uniform01(e)));
    temp2  = 2.0 * M_PI * rng_uniform01(e);
    a[nm1] = temp1 * std::cos(temp2);
  }
}
} // namespace qmi
