// Host-side problem builders of libgdiamond_hip.so (no device needed):
// the reference default problem, the seeded sphere medium and source tables.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
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

// ---- figure output ------------------------------------------------------------
//
// Pixel rule of the reference's field figures (save_field_png, utils.h:210-269):
// value = u / vmax in float, clamped to [-1, 1]; value >= 0 interpolates white
// -> red (r = 255, g = b = (uint8)(255 * (1 - value))), value < 0 white -> blue
// (b = 255, r = g = (uint8)(255 * (1 + value))); pixel (i, j) is u[i + j*nx],
// image nx wide and ny high, row j. Mask figures (save_mask_png, utils.h:170-207):
// black where the mask is set, white elsewhere. The PNG encoder is this
// library's own: 8-bit RGB, unfiltered rows, zlib stream of stored deflate
// blocks (valid for every PNG reader; larger files than the reference's
// compressed ones, identical pixels).

namespace {

uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  for (size_t q = 0; q < n; ++q) crc = table[(crc ^ p[q]) & 0xffu] ^ (crc >> 8);
  return crc;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}

void put_chunk(std::vector<unsigned char>& out, const char type[4],
               const std::vector<unsigned char>& data) {
  put_be32(out, (uint32_t)data.size());
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  const uint32_t crc =
      crc32_update(0xffffffffu, out.data() + start, data.size() + 4) ^ 0xffffffffu;
  put_be32(out, crc);
}

}  // namespace

extern "C" int gd_field_rgb(const float* u, int nx, int ny, float vmax, uint8_t* rgb) {
  if (!u || !rgb || nx < 1 || ny < 1) return GD_ERR_ARG;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      const int q = i + j * nx;
      float value = u[q] / vmax;
      unsigned char r, g, b;
      if (value >= 0) {
        if (value > 1) value = 1;
        r = 255;
        g = b = (unsigned char)(255 * (1 - value));
      } else {
        if (value < -1) value = -1;
        b = 255;
        r = g = (unsigned char)(255 * (1 + value));
      }
      rgb[3 * q] = r;
      rgb[3 * q + 1] = g;
      rgb[3 * q + 2] = b;
    }
  return GD_OK;
}

extern "C" int gd_write_png_rgb(const char* filename, int w, int h, const uint8_t* rgb) {
  if (!filename || !rgb || w < 1 || h < 1) return GD_ERR_ARG;
  std::vector<unsigned char> raw;  // filter byte 0 + row
  raw.reserve((size_t)h * (3 * (size_t)w + 1));
  for (int y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb + (size_t)y * 3 * w, rgb + (size_t)(y + 1) * 3 * w);
  }
  std::vector<unsigned char> z = {0x78, 0x01};
  size_t pos = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - pos);
    z.push_back(pos + n == raw.size() ? 1 : 0);  // BFINAL, BTYPE 00 (stored)
    z.push_back((unsigned char)(n & 0xff));
    z.push_back((unsigned char)(n >> 8));
    z.push_back((unsigned char)(~n & 0xff));
    z.push_back((unsigned char)((~n >> 8) & 0xff));
    z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
    pos += n;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (unsigned char c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  put_be32(z, (b << 16) | a);
  std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  std::vector<unsigned char> ihdr;
  put_be32(ihdr, (uint32_t)w);
  put_be32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter, no interlace
  put_chunk(png, "IHDR", ihdr);
  put_chunk(png, "IDAT", z);
  put_chunk(png, "IEND", {});
  FILE* f = std::fopen(filename, "wb");
  if (!f) return GD_ERR_ARG;
  const size_t wrote = std::fwrite(png.data(), 1, png.size(), f);
  const int closed = std::fclose(f);
  return wrote == png.size() && closed == 0 ? GD_OK : GD_ERR_ARG;
}

extern "C" int gd_save_field_png(const float* u, const char* filename, int nx, int ny,
                                 float vmax) {
  if (!u || !filename || nx < 1 || ny < 1) return GD_ERR_ARG;
  std::vector<uint8_t> rgb((size_t)nx * ny * 3);
  const int rc = gd_field_rgb(u, nx, ny, vmax, rgb.data());
  return rc != GD_OK ? rc : gd_write_png_rgb(filename, nx, ny, rgb.data());
}

extern "C" int gd_save_mask_png(const uint8_t* mask, const char* filename, int nx, int ny) {
  if (!mask || !filename || nx < 1 || ny < 1) return GD_ERR_ARG;
  std::vector<uint8_t> rgb((size_t)nx * ny * 3);
  for (size_t q = 0; q < (size_t)nx * ny; ++q) {
    const uint8_t v = mask[q] ? 0 : 255;
    rgb[3 * q] = rgb[3 * q + 1] = rgb[3 * q + 2] = v;
  }
  return gd_write_png_rgb(filename, nx, ny, rgb.data());
}
