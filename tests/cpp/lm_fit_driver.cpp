// lm_fit_driver.cpp -- C++ caller of the GPU model through the ModelObject-shaped
// host class (imfit_amd/host/model_object_gpu.hpp), set up the way the reference's
// SetupModelObject does (core/setup_model_object.cpp:58-153) for the
// imfit_textout3 regression fit (tests/imfit_reference/imfit_config_ic3478_64x64b.dat:
// Sersic, GAIN 4.725, READNOISE 4.3, ORIGINAL_SKY 130.1, ftol 1e-8), then fitted
// with mpfit (finite-difference Jacobian), mpfit with the analytic Jacobian, and the
// normal-equation LM.
//
// usage: lm_fit_driver <data.f64> <nColumns> <nRows>
//   data.f64: the 64x64 data image as raw little-endian doubles (row-major)
// prints one line per fit: "<method> status <s> niter <n> nfev <n> chi2 <x> params ..."
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../imfit_amd/host/model_object_gpu.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <data.f64> <nColumns> <nRows>\n", argv[0]);
    return 2;
  }
  const int nc = std::atoi(argv[2]), nr = std::atoi(argv[3]);
  std::vector<double> data((size_t)nc * nr);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(data.data(), sizeof(double), data.size(), f) != data.size()) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  std::fclose(f);

  const double start[7] = {32, 32, 18, 0.2, 1.5, 15, 25};   // X0 Y0 PA ell n I_e r_e
  const double lim[7][2] = {{1, 64}, {1, 64}, {0, 90}, {0, 1}, {0, 5}, {0, 500}, {0, 100}};
  std::vector<ifg_param_info> pinfo(7);
  for (int i = 0; i < 7; i++) {
    pinfo[i].fixed = 0;
    pinfo[i].limited[0] = pinfo[i].limited[1] = 1;
    pinfo[i].limits[0] = lim[i][0];
    pinfo[i].limits[1] = lim[i][1];
  }
  ifg_lm_config cfg = {1e-8, 0.0, 0.0, 1000, 0};   // imfit's ftol and maxiter (solvers/levmar_fit.cpp:38,99-101)

  int failures = 0;
  for (int method = 0; method < 3; method++) {
    imfit_gpu::ModelObjectGPU m(0);
    if (!m.Valid()) return 3;
    if (m.AddFunction("Sersic", true) < 0 || m.AddImageDataVector(data.data(), nc, nr) < 0) return 3;
    m.AddImageCharacteristics(4.725, 4.3, 1.0, 1, 130.1);
    if (m.FinalSetupForFitting() < 0) return 3;
    double p[7];
    for (int i = 0; i < 7; i++) p[i] = start[i];
    const double chi2_start = m.GetFitStatistic(p);
    ifg_lm_result res;
    int st;
    const char* name;
    if (method == 0) {
      name = "mpfit-fd";
      st = m.FitLM(p, pinfo.data(), cfg, IFG_LM_JACOBIAN_FD, &res);
    } else if (method == 1) {
      name = "mpfit-analytic";
      st = m.FitLM(p, pinfo.data(), cfg, IFG_LM_JACOBIAN_ANALYTIC, &res);
    } else {
      name = "normal-analytic";
      ifg_lm_config c2 = cfg;
      c2.ftol = 1e-12;
      st = m.FitLMNormal(p, pinfo.data(), c2, &res);
    }
    const double chi2 = m.GetFitStatistic(p);
    std::printf("%s status %d niter %d nfev %d chi2_start %.6f chi2 %.9f params", name, st, res.niter, res.nfev,
                chi2_start, chi2);
    for (int i = 0; i < 7; i++) std::printf(" %.9g", p[i]);
    std::printf("\n");
    if (st <= 0) failures++;
  }
  return failures ? 1 : 0;
}
