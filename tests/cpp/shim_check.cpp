// Exercises the C++ drop-in include/gdiamond/gdiamond.hpp the way the
// reference's own test and examples drive gdiamond::gDiamond:
//   cpu  : seq engine output for the golden-vector comparison, omp == seq
//   seqfig: update_FDTD_seq_figures in the current directory (figure files)
//   gpufig: update_FDTD_gpu_figures(2100) in the current directory
//   gpu  : unittests/checker_gpu.cu:9-22 (seq vs gpu_2D at 100^3 x 100),
//          examples/gpu_mm_ver2.cu / gpu_mm_ver3.cu (naive vs tiled, constant
//          source), examples/gpu_figures.cu (DFT monitor)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "gdiamond/gdiamond.hpp"

static int fail(const char* what) {
  std::fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 2) return fail("usage: shim_check cpu NX NY NZ T OUT | gpu");
  const std::string mode = argv[1];
  try {
    if (mode == "cpu") {
      if (argc != 7) return fail("usage: shim_check cpu NX NY NZ T OUT");
      const size_t nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
      const size_t T = std::atoi(argv[5]);
      gdiamond::gDiamond exp(nx, ny, nz);
      exp.update_FDTD_seq(T);
      exp.update_FDTD_omp(T);
      if (!exp.check_correctness_omp()) return fail("omp != seq");
      FILE* f = std::fopen(argv[6], "wb");
      if (!f) return fail("cannot open output");
      for (int c = 0; c < 6; ++c) {
        const auto& v = exp.field(gdiamond::gDiamond::SEQ, c);
        std::fwrite(v.data(), sizeof(float), v.size(), f);
      }
      std::fclose(f);
      std::printf("cpu ok\n");
      return 0;
    }
    if (mode == "seqfig") {
      if (argc != 7) return fail("usage: shim_check seqfig NX NY NZ T OUT");
      const size_t nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
      const size_t T = std::atoi(argv[5]);
      gdiamond::gDiamond exp(nx, ny, nz);
      exp.update_FDTD_seq_figures(T);
      FILE* f = std::fopen(argv[6], "wb");
      if (!f) return fail("cannot open output");
      const auto& hz = exp.field(gdiamond::gDiamond::SEQ, 5);
      std::fwrite(hz.data(), sizeof(float), hz.size(), f);
      std::fclose(f);
      std::printf("seqfig ok\n");
      return 0;
    }
    if (mode == "gpufig") {
      gdiamond::gDiamond exp(60, 50, 40);
      exp.update_FDTD_gpu_figures(2100);
      std::printf("gpufig ok\n");
      return 0;
    }
    if (mode == "gpu") {
      {  // unittests/checker_gpu.cu
        gdiamond::gDiamond exp(100, 100, 100);
        exp.update_FDTD_seq(100);
        exp.update_FDTD_gpu_2D(100);
        if (!exp.check_correctness_gpu_2D()) return fail("checker_gpu: seq vs gpu_2D");
        std::printf("checker_gpu ok\n");
      }
      {  // examples/gpu_mm_ver2.cu with Tx=Ty=Tz=3: N = 3*(16+11)+16-2+11
        const size_t nx = 3 * 27 + 25, ny = 3 * 11 + 9, nz = 3 * 11 + 9;
        gdiamond::gDiamond exp(nx, ny, nz);
        exp.update_FDTD_gpu_3D_warp_underutilization_fix(31);
        exp.update_FDTD_mix_mapping_gpu_ver2(31, 3, 3, 3);
        if (!exp.check_correctness_gpu()) return fail("gpu_mm_ver2: naive vs tiled");
        std::printf("gpu_mm_ver2 ok\n");
      }
      {  // examples/gpu_mm_ver3.cu with Tx=Ty=Tz=5: N = T*VALLEY = (65, 25, 25)
        gdiamond::gDiamond exp(65, 25, 25);
        exp.update_FDTD_gpu_3D_warp_underutilization_fix(20);
        exp.update_FDTD_mix_mapping_gpu_ver3(20, 5, 5, 5);
        if (!exp.check_correctness_gpu()) return fail("gpu_mm_ver3: naive vs tiled");
        std::printf("gpu_mm_ver3 ok\n");
      }
      {  // examples/gpu_figures.cu
        gdiamond::gDiamond exp(60, 50, 40);
        exp.update_FDTD_gpu_figures(200);
        const auto& re = exp.dft_real();
        if (exp.dft_frequencies().size() != 26 || re.size() != 26u * 60 * 50)
          return fail("gpu_figures: monitor shape");
        double mx = 0;
        for (double v : re) mx = std::fmax(mx, std::fabs(v));
        if (!(mx > 0) || !std::isfinite(mx)) return fail("gpu_figures: empty monitor");
        std::printf("gpu_figures ok (max |re| %.3e)\n", mx);
      }
      return 0;
    }
  } catch (const gdiamond::gd_error& e) {
    std::fprintf(stderr, "gd_error %d: %s\n", e.code, e.what());
    return 2;
  }
  return fail("unknown mode");
}
