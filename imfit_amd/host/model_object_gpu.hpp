// model_object_gpu.hpp -- C++ host class over the C ABI in include/imfit_gpu.h,
// with the method names, argument meaning and error behaviour of the reference's
// ModelObject (core/model_object.h:38-347), so that C++ callers written against
// ModelObject (mpfit's callback, the DE/NLopt/DREAM objective functions) can drive
// the GPU model unchanged.
//
// Error behaviour as ModelObject: setup methods return an int status, negative on
// error (the message is kept and available from LastError(), and is printed to
// stderr like the reference's messages); the evaluation entry points return
// void/double and yield NaN results when the library reports an error.
//
// Header-only; link with -limfit_gpu (imfit_amd/libimfit_gpu.so).
#pragma once

#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "imfit_gpu.h"

namespace imfit_gpu {

// One oversampled region: what PsfOversamplingInfo carries
// (core/psf_oversampling_info.h:29-37), with 1-based inclusive data coordinates.
struct OversampledRegionInfo {
  std::vector<double> psfPixels;
  int nColumns = 0, nRows = 0, scale = 1;
  int x1 = 0, x2 = 0, y1 = 0, y2 = 0;
  bool normalizePSF = true;
};

class ModelObjectGPU {
 public:
  explicit ModelObjectGPU(int device = 0, int precision = IFG_PRECISION_FP64) : h_(ifg_create(device)) {
    if (h_ && precision != IFG_PRECISION_FP64) Check(ifg_set_precision(h_, precision), "SetPrecision");
  }
  ~ModelObjectGPU() { ifg_destroy(h_); }
  ModelObjectGPU(const ModelObjectGPU&) = delete;
  ModelObjectGPU& operator=(const ModelObjectGPU&) = delete;

  bool Valid() const { return h_ != nullptr; }
  const char* LastError() const { return h_ ? ifg_last_error(h_) : "no device"; }

  // ---- model definition (AddFunctions / DefineFunctionSets, core/add_functions.cpp)
  int AddFunction(const std::string& name, bool startsNewFunctionSet) {
    return Check(ifg_add_function(h_, name.c_str(), startsNewFunctionSet ? 1 : 0), "AddFunction");
  }
  void SetSubsampling(bool on) { Check(ifg_set_subsampling(h_, on ? 1 : 0), "SetSubsampling"); }
  int GetNParams() const { return ifg_get_nparams(h_); }

  // ---- image setup (core/model_object.h:82,92,115,119)
  int AddPSFVector(long nPixels_psf, int nColumns_psf, int nRows_psf, const double* psfPixels,
                   bool normalizePSF = true) {
    if (nPixels_psf != (long)nColumns_psf * nRows_psf) return Fail("AddPSFVector: size mismatch");
    return Check(ifg_add_psf(h_, psfPixels, nColumns_psf, nRows_psf, normalizePSF ? 1 : 0), "AddPSFVector");
  }
  int SetupModelImage(int nImageColumns, int nImageRows) {
    nColumns_ = nImageColumns;
    nRows_ = nImageRows;
    return Check(ifg_setup_model_image(h_, nImageColumns, nImageRows), "SetupModelImage");
  }
  int AddImageDataVector(const double* pixels, int nImageColumns, int nImageRows) {
    nColumns_ = nImageColumns;
    nRows_ = nImageRows;
    return Check(ifg_add_image_data(h_, pixels, nImageColumns, nImageRows), "AddImageDataVector");
  }
  void AddImageCharacteristics(double imageGain, double readoutNoise, double expTime, int nCombinedImages,
                               double originalSkyBackground) {
    Check(ifg_add_image_characteristics(h_, imageGain, readoutNoise, expTime, nCombinedImages,
                                        originalSkyBackground),
          "AddImageCharacteristics");
  }
  int AddErrorVector(long nDataValues, int nImageColumns, int nImageRows, const double* pixels, int inputType) {
    if (nDataValues != (long)nImageColumns * nImageRows) return Fail("AddErrorVector: size mismatch");
    return Check(ifg_add_error_vector(h_, pixels, nDataValues, inputType), "AddErrorVector");
  }
  int AddMaskVector(long nDataValues, int nImageColumns, int nImageRows, const double* pixels, int inputType) {
    if (nDataValues != (long)nImageColumns * nImageRows) return Fail("AddMaskVector: size mismatch");
    return Check(ifg_add_mask_vector(h_, pixels, nDataValues, inputType), "AddMaskVector");
  }
  int AddOversampledPsfInfo(const OversampledRegionInfo& o) {
    return Check(ifg_add_oversampled_psf(h_, o.psfPixels.data(), o.nColumns, o.nRows, o.scale, o.x1, o.x2, o.y1,
                                         o.y2, o.normalizePSF ? 1 : 0),
                 "AddOversampledPsfInfo");
  }
  int UseCashStatistic() { return Check(ifg_use_cash_statistic(h_), "UseCashStatistic"); }
  void UsePoissonMLR() { Check(ifg_use_poisson_mlr(h_), "UsePoissonMLR"); }
  int UseModelErrors() { return Check(ifg_use_model_errors(h_), "UseModelErrors"); }
  int FinalSetupForFitting() { return Check(ifg_final_setup(h_), "FinalSetupForFitting"); }
  int WhichFitStatistic() const { return ifg_which_fit_statistic(h_); }
  long GetNDataValues() const { return ifg_get_ndata(h_); }
  long GetNValidPixels() const { return ifg_get_nvalid(h_); }

  // bootstrap resampling (core/model_object.cpp:2036-2087); the generator is seeded
  // process-wide with ifg_init_genrand, as imfit seeds init_genrand
  int UseBootstrap() { return Check(ifg_use_bootstrap(h_), "UseBootstrap"); }
  int MakeBootstrapSample() { return Check(ifg_make_bootstrap_sample(h_), "MakeBootstrapSample"); }

  // ---- hot path (core/model_object.h:132,141,153,155,157,223)
  void CreateModelImage(const double params[]) {
    model_.clear();
    Check(ifg_create_model_image(h_, params), "CreateModelImage");
  }
  void ComputeDeviates(double yResults[], const double params[]) {
    if (Check(ifg_compute_deviates(h_, params, yResults), "ComputeDeviates") < 0)
      for (long i = 0; i < GetNDataValues(); i++) yResults[i] = NaN();
  }
  double GetFitStatistic(const double params[]) {
    double s = NaN();
    if (Check(ifg_fit_statistic(h_, params, &s), "GetFitStatistic") < 0) return NaN();
    return s;
  }
  double ChiSquared(const double params[]) { return GetFitStatistic(params); }
  double CashStatistic(const double params[]) { return GetFitStatistic(params); }
  // Model image of the last evaluation (nColumns x nRows, row-major), owned here.
  const double* GetModelImageVector() {
    model_.assign((size_t)nColumns_ * nRows_, NaN());
    Check(ifg_get_model_image(h_, model_.data()), "GetModelImageVector");
    return model_.data();
  }

  // GetExpandedModelImageVector (core/model_object.cpp:2297-2305): the model on the
  // PSF-padded grid; null (with the library's message) where the GPU does not compute it
  const double* GetExpandedModelImageVector() {
    model_.assign((size_t)nColumns_ * nRows_, NaN());
    if (Check(ifg_get_expanded_model_image(h_, model_.data()), "GetExpandedModelImageVector") < 0) return nullptr;
    return model_.data();
  }

  // ---- extensions over the reference interface
  // population form of GetFitStatistic: params is nVectors x nParams, row-major
  void GetFitStatisticBatch(const double* params, int nVectors, double* out) {
    if (Check(ifg_fit_statistic_batch(h_, params, nVectors, out), "GetFitStatisticBatch") < 0)
      for (int i = 0; i < nVectors; i++) out[i] = NaN();
  }
  // exact Jacobian columns (mpfit's side == 3 layout: column j at fjac + j*nData)
  int Jacobian(const double* params, int nFree, const int* ifree, double* fvec, double* fjac) {
    return Check(ifg_jacobian(h_, params, nFree, ifree, fvec, fjac), "Jacobian");
  }
  int NormalEquations(const double* params, int nFree, const int* ifree, double* jtj, double* jtr, double* stat) {
    return Check(ifg_normal_equations(h_, params, nFree, ifree, jtj, jtr, stat), "NormalEquations");
  }
  int FitStatisticGradient(const double* params, double* grad, double* stat) {
    return Check(ifg_fit_statistic_gradient(h_, params, grad, stat), "FitStatisticGradient");
  }
  // mpfit on the GPU model (jacobianMode IFG_LM_JACOBIAN_FD or _ANALYTIC); params in/out
  int FitLM(double* params, const ifg_param_info* pinfo, const ifg_lm_config& cfg, int jacobianMode,
            ifg_lm_result* res) {
    return Check(ifg_fit_lm(h_, params, pinfo, &cfg, jacobianMode, res), "FitLM");
  }
  int FitLMNormal(double* params, const ifg_param_info* pinfo, const ifg_lm_config& cfg, ifg_lm_result* res) {
    return Check(ifg_fit_lm_normal(h_, params, pinfo, &cfg, res), "FitLMNormal");
  }

 private:
  static double NaN() { return std::numeric_limits<double>::quiet_NaN(); }
  int Fail(const char* msg) {
    std::fprintf(stderr, "** ERROR: %s\n", msg);
    return -1;
  }
  int Check(int status, const char* what) {
    if (!h_) return Fail("ModelObjectGPU: no device handle");
    if (status < 0) std::fprintf(stderr, "** ERROR: %s: %s\n", what, ifg_last_error(h_));
    return status;
  }

  ifg_model* h_;
  std::vector<double> model_;
  int nColumns_ = 0, nRows_ = 0;
};

}  // namespace imfit_gpu
