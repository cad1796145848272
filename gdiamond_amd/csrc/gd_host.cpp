// Host-side problem builders of libgdiamond_hip.so (no device needed):
// the reference default problem, the seeded sphere medium and source tables.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/gdiamond/gd_capi.h"

namespace {

// Constants of the reference default problem (gdiamond.hpp:894-906,
// utils.h:17-22). All float, as declared there.
struct DefaultConsts {
  float pi = 3.14159265359f;
  float eps0 = 1.0f, mu0 = 1.0f;
  float um = 1.0f;
  float c0;      // 1/sqrt(mu0*eps0)
  float lambda;  // source wavelength = 1 um
  float freq;    // c0 / lambda
  float omega;   // 2*pi*freq
  float dx;      // lambda / 10
  float dt = 0.05f;
  DefaultConsts() {
    c0 = 1.0f / std::sqrt(mu0 * eps0);
    lambda = 1 * um;
    freq = c0 / lambda;
    omega = 2 * pi * freq;
    dx = lambda / 10;
  }
};

// Complex float with the operation order of utils.h:31-116.
struct CF {
  float re, im;
};
inline CF operator*(CF a, CF b) {
  return CF{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline CF operator/(CF a, CF b) {
  const float d = b.re * b.re + b.im * b.im;
  return CF{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// Stretch factor of one axis in the graded absorbing layer
// (utils.h:602-633): distance into the layer, then
// s = (1 + a_max*d^p) + i*sigma_max*d^p/(omega*eps0), mixed float/double as
// written there.
CF stretch(int i, int n, int t_pml, float a_max, int p, float sigma_max, float omega,
           float eps0) {
  float d = 0;
  if (i < t_pml) d = 1.0 - float(i) / t_pml;
  if (i + t_pml >= n) d = 1.0 - float(n - i - 1) / t_pml;
  const double dp = std::pow((double)d, (double)p);
  CF s;
  s.re = 1 + a_max * dp;
  s.im = sigma_max * dp / (omega * eps0);
  return s;
}

// Ca/Cb (or Da/Db) of one component from its complex permittivity
// (utils.h:655-686).
void lossy_pair(CF e, float omega, float eps0, float dt, float dx, float* ca, float* cb) {
  float er = e.re;
  if (er < 1) er = 1;
  const float sigma = std::abs(omega * eps0 * e.im);
  const float tmp = sigma * dt / (2.0 * er * eps0);
  *ca = (1.0 - tmp) / (1.0 + tmp);
  *cb = (dt / (er * eps0)) / (1.0 + tmp) / dx;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

}  // namespace

extern "C" int gd_default_problem_coefficients(size_t nx, size_t ny, size_t nz,
                                               float* const coeff[12]) {
  if (nx < 1 || ny < 1 || nz < 1 || !coeff) return GD_ERR_ARG;
  for (int a = 0; a < 12; ++a)
    if (!coeff[a]) return GD_ERR_ARG;
  const DefaultConsts K;
  // t_PML = ceil(1 um / dx) (gdiamond.hpp:63-64); sigma_max (utils.h:569-573)
  const int t_pml = (int)std::ceil(1 * K.um / K.dx);
  const float a_max = 2.0f;
  const int p = 3;
  const float eta0 = std::sqrt(K.mu0 / K.eps0);
  const float sigma_max = -(p + 1) * std::log(1e-5) / (2 * eta0 * t_pml * K.dx);
  const int Nx = (int)nx, Ny = (int)ny, Nz = (int)nz;
  // The default structure mask is all false (gdiamond.hpp:52-54), so the
  // relative permittivity is the air value 1 + 0i in every cell.
  const CF eps_r{1.0f, 0.0f}, mu_r{1.0f, 0.0f};
  std::vector<CF> sx(Nx), sy(Ny), sz(Nz);
  for (int i = 0; i < Nx; ++i) sx[i] = stretch(i, Nx, t_pml, a_max, p, sigma_max, K.omega, K.eps0);
  for (int j = 0; j < Ny; ++j) sy[j] = stretch(j, Ny, t_pml, a_max, p, sigma_max, K.omega, K.eps0);
  for (int k = 0; k < Nz; ++k) sz[k] = stretch(k, Nz, t_pml, a_max, p, sigma_max, K.omega, K.eps0);
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t idx = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
        const CF ex = (eps_r * sy[j] * sz[k]) / sx[i];
        const CF ey = (eps_r * sx[i] * sz[k]) / sy[j];
        const CF ez = (eps_r * sx[i] * sy[j]) / sz[k];
        const CF mx = (mu_r * sy[j] * sz[k]) / sx[i];
        const CF my = (mu_r * sx[i] * sz[k]) / sy[j];
        const CF mz = (mu_r * sx[i] * sy[j]) / sz[k];
        lossy_pair(ex, K.omega, K.eps0, K.dt, K.dx, &coeff[0][idx], &coeff[1][idx]);
        lossy_pair(ey, K.omega, K.eps0, K.dt, K.dx, &coeff[2][idx], &coeff[3][idx]);
        lossy_pair(ez, K.omega, K.eps0, K.dt, K.dx, &coeff[4][idx], &coeff[5][idx]);
        lossy_pair(mx, K.omega, K.mu0, K.dt, K.dx, &coeff[6][idx], &coeff[7][idx]);
        lossy_pair(my, K.omega, K.mu0, K.dt, K.dx, &coeff[8][idx], &coeff[9][idx]);
        lossy_pair(mz, K.omega, K.mu0, K.dt, K.dx, &coeff[10][idx], &coeff[11][idx]);
      }
  return GD_OK;
}

extern "C" int gd_default_mz_table(size_t nsteps, float* out) {
  if (!out && nsteps) return GD_ERR_ARG;
  const DefaultConsts K;
  const float eta0 = std::sqrt(K.mu0 / K.eps0);
  const float j_amp = 5e4;
  const float m_amp = j_amp * std::pow(eta0, 3.0);
  for (size_t t = 0; t < nsteps; ++t) out[t] = m_amp * std::sin(K.omega * t * K.dt);
  return GD_OK;
}

extern "C" int gd_sphere_medium(size_t nx, size_t ny, size_t nz, uint64_t seed, int nspheres,
                                double rmin, double rmax, double emin, double emax,
                                float* eps) {
  if (!eps || nx < 1 || ny < 1 || nz < 1 || nspheres < 0 || rmax < rmin || emax < emin)
    return GD_ERR_ARG;
  const size_t N = nx * ny * nz;
  for (long long q = 0; q < (long long)N; ++q) eps[q] = 1.0f;
  uint64_t s = seed;
  for (int sp = 0; sp < nspheres; ++sp) {
    const double cx = uniform01(s) * (double)nx;
    const double cy = uniform01(s) * (double)ny;
    const double cz = uniform01(s) * (double)nz;
    const double r = rmin + uniform01(s) * (rmax - rmin);
    const float e = (float)(emin + uniform01(s) * (emax - emin));
    const double r2 = r * r;
    const long long i0 = std::max(0LL, (long long)std::ceil(cx - r));
    const long long i1 = std::min((long long)nx - 1, (long long)std::floor(cx + r));
    const long long j0 = std::max(0LL, (long long)std::ceil(cy - r));
    const long long j1 = std::min((long long)ny - 1, (long long)std::floor(cy + r));
    const long long k0 = std::max(0LL, (long long)std::ceil(cz - r));
    const long long k1 = std::min((long long)nz - 1, (long long)std::floor(cz + r));
    for (long long k = k0; k <= k1; ++k)
      for (long long j = j0; j <= j1; ++j)
        for (long long i = i0; i <= i1; ++i) {
          const double ddx = (double)i - cx, ddy = (double)j - cy, ddz = (double)k - cz;
          if (ddx * ddx + ddy * ddy + ddz * ddz <= r2)
            eps[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny] = e;
        }
  }
  return GD_OK;
}

extern "C" int gd_gaussian_pulse_table(size_t nsteps, double t0, double tau, double f_cps,
                                       float* out) {
  if ((!out && nsteps) || tau == 0.0) return GD_ERR_ARG;
  const double two_pi = 6.283185307179586476925286766559;
  for (size_t n = 0; n < nsteps; ++n) {
    const double u = ((double)n - t0) / tau;
    out[n] = (float)(std::exp(-(u * u)) * std::sin(two_pi * f_cps * (double)n));
  }
  return GD_OK;
}
