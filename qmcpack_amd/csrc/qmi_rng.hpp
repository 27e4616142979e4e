// Host RNG streams of the VMC driver (see qmi_rng.cpp).
#pragma once
#include <cstddef>
#include <random>

namespace qmi
{
double rng_uniform01(std::mt19937& e);
void rng_fill_uniform(double* a, size_t n, std::mt19937& e);
void rng_fill_gauss(double* a, unsigned n, std::mt19937& e);
} // namespace qmi
