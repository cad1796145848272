// Thin C entry points over the reference's own CPU path, compiled from the
// reference headers where they lie (see oracle/Makefile, target `ref`).
// TEST INFRASTRUCTURE ONLY, used where the reference sources are present
// (never on a GPU host: oracle/_ref is neither committed nor copied there):
// it writes the golden vectors under tests/golden/ and cross-checks the
// oracle restatement in the CPU tests. The product, the GPU tests and the
// benchmark never load it.
//
// The reference keeps its arrays private (gdiamond.hpp:798-891); the standard
// headers are included first and `private` is widened only for the one
// reference header so the results can be copied out unchanged.
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>
#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#define private public
#include "gdiamond.hpp"
#undef private

extern "C" {

// gDiamond(Nx,Ny,Nz) (gdiamond.hpp:20-73): copy the 12 coefficient arrays
// Cax Cbx Cay Cby Caz Cbz Dax Dbx Day Dby Daz Dbz into out[12*N].
int ref_default_coefficients(int nx, int ny, int nz, float* out) {
  gdiamond::gDiamond g(nx, ny, nz);
  const size_t N = (size_t)nx * ny * nz;
  const std::vector<float>* c[12] = {&g._Cax, &g._Cbx, &g._Cay, &g._Cby, &g._Caz, &g._Cbz,
                                     &g._Dax, &g._Dbx, &g._Day, &g._Dby, &g._Daz, &g._Dbz};
  for (int a = 0; a < 12; ++a) memcpy(out + a * N, c[a]->data(), N * sizeof(float));
  return 0;
}

// update_FDTD_seq(T) (gdiamond.hpp:1164-1263): copy _Ex_seq.._Hz_seq into out[6*N].
int ref_seq(int nx, int ny, int nz, int T, float* out) {
  gdiamond::gDiamond g(nx, ny, nz);
  g.update_FDTD_seq((size_t)T);
  const size_t N = (size_t)nx * ny * nz;
  const std::vector<float>* f[6] = {&g._Ex_seq, &g._Ey_seq, &g._Ez_seq,
                                    &g._Hx_seq, &g._Hy_seq, &g._Hz_seq};
  for (int a = 0; a < 6; ++a) memcpy(out + a * N, f[a]->data(), N * sizeof(float));
  return 0;
}

// update_FDTD_seq(T) on the default problem, returning the rate the reference
// itself measures and prints (gdiamond.hpp:1186, 1251-1253: Mcells/s over the
// step loop); its console output is captured instead of printed.
int ref_seq_rate(int nx, int ny, int nz, int T, double* mcells_per_s) {
  gdiamond::gDiamond g(nx, ny, nz);
  std::ostringstream out;
  std::streambuf* old = std::cout.rdbuf(out.rdbuf());
  g.update_FDTD_seq((size_t)T);
  std::cout.rdbuf(old);
  const std::string key = "seq performance (excluding figures output): ";
  const std::string s = out.str();
  const size_t at = s.find(key);
  if (at == std::string::npos) return -1;
  *mcells_per_s = std::strtod(s.c_str() + at + key.size(), nullptr);
  return 0;
}

// The per-step source value update_FDTD_seq writes (gdiamond.hpp:1189-1191),
// evaluated by the reference object's own members.
float ref_default_mz(int t) {
  gdiamond::gDiamond g(3, 3, 3);
  return g.M_source_amp * std::sin(g.SOURCE_OMEGA * (size_t)t * g.dt);
}

// The reference's figure writers (utils.h:168-269, stb_image_write as the
// reference vendors it): the PNG a field or a mask becomes.
int ref_save_field_png(const float* u, const char* filename, int nx, int ny, float vmax) {
  std::vector<float> v(u, u + (size_t)nx * ny);
  save_field_png(v.data(), filename, nx, ny, vmax);
  return 0;
}

int ref_save_mask_png(const unsigned char* mask, const char* filename, int nx, int ny) {
  std::vector<char> b((size_t)nx * ny);
  bool* m = reinterpret_cast<bool*>(b.data());
  for (size_t q = 0; q < b.size(); ++q) m[q] = mask[q] != 0;
  save_mask_png(m, filename, nx, ny);
  return 0;
}

}  // extern "C"
