// Bit-exact host helpers (compiled with g++; see host_init.cpp).
#pragma once
#include <vector>

namespace qmin {
void director_q(double nx, double ny, double nz, double S0, double* q);
void random_nematic(long long n, const int* types, int seed, double S0, double* q);
void random_director(int seed, double* n);
void normalize_director(double* n);
void sphere_sites(double cx, double cy, double cz, double radius, int btype, double S0, std::vector<int>& xyz,
                  std::vector<double>& qb);
}  // namespace qmin
