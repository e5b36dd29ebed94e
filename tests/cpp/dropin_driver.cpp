// dropin_driver.cpp -- the call sequences a drop-in ModelObject subclass makes, through
// the ModelObject-shaped host class (imfit_amd/host/model_object_gpu.hpp):
//
//   weights <functions> <data.f64> <mask.f64|-> <sigma.f64|-> <psf.f64|-> <psfSide> <nColumns>
//           <nRows> <gain> <readNoise> <originalSky> <nParams> <params...>
//     imfit's fitting setup after the reference's FinalSetupForFitting has run: its
//     internal weight vector is 1/sigma times the 0/1 mask (core/model_object.cpp:573-619,
//     642-673, 796-816).  The subclass hands those final weights to the library squared,
//     as "weights" (1/sigma^2, which AddErrorVector square-roots back), with the mask.
//     Prints "chi2 <value>".
//   makeimage <functions> <psf.f64|-> <psfSide> <osPsf.f64|-> <osSide> <scale> <x1> <x2> <y1>
//             <y2> <nColumns> <nRows> <out.f64> <nParams> <params...>
//     makeimage's sequence (core/makeimage_main.cpp:239-276): functions, PSF, image size,
//     oversampled region, CreateModelImage -- no data and no FinalSetupForFitting.
//     Writes the model image; prints "expanded ok" or "expanded unsupported".
// <functions>: comma-separated names, a leading '+' starting a new function set
// (e.g. "+Sersic,Exponential").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../imfit_amd/host/model_object_gpu.hpp"

namespace {

int AddFunctions(imfit_gpu::ModelObjectGPU& m, const std::string& list) {
  size_t pos = 0;
  while (pos <= list.size()) {
    size_t end = list.find(',', pos);
    if (end == std::string::npos) end = list.size();
    std::string name = list.substr(pos, end - pos);
    const bool newSet = !name.empty() && name[0] == '+';
    if (newSet) name = name.substr(1);
    if (m.AddFunction(name, newSet) < 0) return -1;
    pos = end + 1;
  }
  return 0;
}

std::vector<double> ReadRaw(const char* path, size_t n) {
  std::vector<double> v;
  if (std::strcmp(path, "-") == 0) return v;
  v.resize(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

int Weights(int argc, char** argv) {
  if (argc < 15) return 2;
  const int side = std::atoi(argv[7]), nc = std::atoi(argv[8]), nr = std::atoi(argv[9]);
  const double gain = std::atof(argv[10]), rn = std::atof(argv[11]), sky = std::atof(argv[12]);
  const int np = std::atoi(argv[13]);
  if (argc != 14 + np) return 2;
  std::vector<double> params(np);
  for (int i = 0; i < np; i++) params[i] = std::atof(argv[14 + i]);
  const long n = (long)nc * nr;
  std::vector<double> data = ReadRaw(argv[3], n), mask = ReadRaw(argv[4], n), sigma = ReadRaw(argv[5], n);
  std::vector<double> psf = ReadRaw(argv[6], (size_t)side * side);

  // what the reference leaves in weightVector after FinalSetupForFitting (exposure
  // time 1, one combined image): 1/sigma from the error image or from the data, times
  // the 0/1 mask (1 = good, as AddMaskVector converts the input mask)
  std::vector<double> good(n, 1.0), w(n);
  for (long z = 0; z < n; z++) {
    if (!mask.empty()) good[z] = (mask[z] > 0.0) ? 0.0 : 1.0;   // input mask: zero is good
    double s;
    if (!sigma.empty()) {
      s = sigma[z];
    } else {
      double flux = data[z] + sky;
      if (flux < 0.0) flux = 0.0;
      s = std::sqrt(flux / gain + rn * rn / (gain * gain));
    }
    w[z] = good[z] * (1.0 / s);
  }

  imfit_gpu::ModelObjectGPU m(0);
  if (!m.Valid() || AddFunctions(m, argv[2]) < 0) return 3;
  if (m.GetNParams() != np) return 2;
  if (!psf.empty() && m.AddPSFVector((long)side * side, side, side, psf.data(), true) < 0) return 3;
  if (m.AddImageDataVector(data.data(), nc, nr) < 0) return 3;
  m.AddImageCharacteristics(gain, rn, 1.0, 1, sky);
  std::vector<double> w2(n);
  for (long z = 0; z < n; z++) w2[z] = w[z] * w[z];   // AddErrorVector takes the square root back
  if (m.AddErrorVector(n, nc, nr, w2.data(), IFG_WEIGHTS_ARE_WEIGHTS) < 0) return 3;
  if (m.AddMaskVector(n, nc, nr, good.data(), IFG_MASK_ZERO_IS_BAD) < 0) return 3;
  if (m.FinalSetupForFitting() < 0) return 3;
  std::printf("chi2 %.17g\n", m.GetFitStatistic(params.data()));
  return 0;
}

int Makeimage(int argc, char** argv) {
  if (argc < 17) return 2;
  const int side = std::atoi(argv[4]), osSide = std::atoi(argv[6]), scale = std::atoi(argv[7]);
  const int x1 = std::atoi(argv[8]), x2 = std::atoi(argv[9]), y1 = std::atoi(argv[10]), y2 = std::atoi(argv[11]);
  const int nc = std::atoi(argv[12]), nr = std::atoi(argv[13]);
  const char* out = argv[14];
  const int np = std::atoi(argv[15]);
  if (argc != 16 + np) return 2;
  std::vector<double> params(np);
  for (int i = 0; i < np; i++) params[i] = std::atof(argv[16 + i]);
  std::vector<double> psf = ReadRaw(argv[3], (size_t)side * side);
  std::vector<double> osPsf = ReadRaw(argv[5], (size_t)osSide * osSide);

  imfit_gpu::ModelObjectGPU m(0);
  if (!m.Valid() || AddFunctions(m, argv[2]) < 0) return 3;
  if (m.GetNParams() != np) return 2;
  if (!psf.empty() && m.AddPSFVector((long)side * side, side, side, psf.data(), true) < 0) return 3;
  if (m.SetupModelImage(nc, nr) < 0) return 3;
  if (!osPsf.empty()) {
    imfit_gpu::OversampledRegionInfo o;
    o.psfPixels = osPsf;
    o.nColumns = o.nRows = osSide;
    o.scale = scale;
    o.x1 = x1;
    o.x2 = x2;
    o.y1 = y1;
    o.y2 = y2;
    if (m.AddOversampledPsfInfo(o) < 0) return 3;
  }
  m.CreateModelImage(params.data());
  const double* img = m.GetModelImageVector();
  FILE* f = std::fopen(out, "wb");
  if (!f || std::fwrite(img, sizeof(double), (size_t)nc * nr, f) != (size_t)nc * nr) return 4;
  std::fclose(f);
  std::printf("expanded %s\n", m.GetExpandedModelImageVector() ? "ok" : "unsupported");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "weights") return Weights(argc, argv);
  if (argc > 1 && std::string(argv[1]) == "makeimage") return Makeimage(argc, argv);
  std::fprintf(stderr, "usage: %s weights|makeimage ...\n", argv[0]);
  return 2;
}
