/* imfit_gpu.h -- C ABI of the MI355X (gfx950) model-image library.
 *
 * One `ifg_model` plays the role of one reference ModelObject
 * (core/model_object.h:38-347): the setup calls mirror ModelObject's setup
 * methods in the same order (core/setup_model_object.cpp:58-153: PSF, data,
 * oversampled regions, mask, statistic, FinalSetupForFitting), and the
 * evaluation calls replace its hot entry points (CreateModelImage,
 * ComputeDeviates, GetFitStatistic).  Plain pointers and sizes only; every call
 * returns an int status, negative on error (message via ifg_last_error), like
 * the reference's int-status setup methods (e.g. core/model_object.cpp:517-523,
 * 839-856).  No C++ exception leaves any entry point, and a null handle is
 * IFG_ERR_ARG (the getters of counts return 0 for it).  Caller buffers are
 * copied, never modified.
 *
 * Pixel layout everywhere: row-major, x (column) fastest, as the reference's
 * pixel vectors.  Parameter vectors use the reference layout: X0,Y0 at the start
 * of each function set, then each component's parameters in order
 * (core/model_object.cpp:1126-1135).
 *
 * Threading: one host thread per ifg_model at a time (the reference ModelObject
 * is not re-entrant either).  All device work of a handle runs on its own HIP
 * stream.
 */
#ifndef IMFIT_GPU_H
#define IMFIT_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifg_model ifg_model;

/* status codes */
#define IFG_OK 0
#define IFG_ERR_ARG (-1)         /* bad argument (size mismatch, non-finite PSF, ...) */
#define IFG_ERR_STATE (-2)       /* call out of order (e.g. PSF after SetupModelImage) */
#define IFG_ERR_DATA (-3)        /* FinalSetupForFitting found bad data / weights */
#define IFG_ERR_HIP (-4)         /* HIP runtime error (message has the HIP error string) */
#define IFG_ERR_UNSUPPORTED (-5) /* configuration outside what the library implements */
#define IFG_ERR_NOMEM (-6)       /* host memory exhausted inside the call */
#define IFG_ERR_INTERNAL (-7)    /* any other failure inside the library (message via ifg_last_error) */

/* input-type codes: same values as core/definitions.h:75-82 */
#define IFG_WEIGHTS_ARE_SIGMAS 100
#define IFG_WEIGHTS_ARE_VARIANCES 110
#define IFG_WEIGHTS_ARE_WEIGHTS 120
#define IFG_MASK_ZERO_IS_GOOD 10
#define IFG_MASK_ZERO_IS_BAD 20

/* fit statistics: same values as core/definitions.h:54-59 */
#define IFG_FITSTAT_CHISQUARE 1
#define IFG_FITSTAT_CASH 2
#define IFG_FITSTAT_POISSON_MLR 3

/* pixel arithmetic: FP64 (parity mode) or FP32 pixels + FFT with FP64 reductions */
#define IFG_PRECISION_FP64 0
#define IFG_PRECISION_FP32 1

/* ---- lifetime ------------------------------------------------------------- */
/* ModelObject::ModelObject (core/model_object.cpp:96-140); device = HIP ordinal */
ifg_model* ifg_create(int device);
/* ModelObject::~ModelObject (core/model_object.cpp:180-232) */
void ifg_destroy(ifg_model* m);
/* last error message of this handle ("" if none) */
const char* ifg_last_error(const ifg_model* m);
/* pixel arithmetic (IFG_PRECISION_*); must precede the first evaluation */
int ifg_set_precision(ifg_model* m, int precision);

/* ---- model definition ----------------------------------------------------- */
/* AddFunction + DefineFunctionSets (core/model_object.cpp:277-340; name -> class
 * as in core/add_functions.cpp:163-385).  Supported names: "Sersic",
 * "Exponential", "BrokenExponential", "EdgeOnDisk", "FlatSky", "Gaussian", "Moffat",
 * "ExponentialDisk3D", "Sersic_GenEllipse", "PointSource", "PointSourceRot".
 * PointSource (I_tot) and PointSourceRot (PA, I_tot) are the PSF itself, interpolated
 * at the source's position and added after the convolution (function_objects/
 * func_pointsource.cpp, core/model_object.cpp:1183-1218); they need a PSF of at least
 * 4 x 4 pixels (checked at the first evaluation). */
int ifg_add_function(ifg_model* m, const char* name, int starts_new_function_set);
/* ModelObject::SetupPsfInterpolation (core/model_object.cpp:343-380): how point sources
 * interpolate the PSF -- IFG_INTERP_BICUBIC (default; GSL's 2-D bicubic spline in the
 * reference), IFG_INTERP_LANCZOS2 or IFG_INTERP_LANCZOS3 (the "method" optional
 * parameter of the first PointSource in a configuration file).  Oversampled regions
 * always interpolate their PSF bicubically (core/oversampled_region.cpp:181-185).
 * Before the first evaluation. */
#define IFG_INTERP_BICUBIC 1
#define IFG_INTERP_LANCZOS2 2
#define IFG_INTERP_LANCZOS3 3
int ifg_set_psf_interpolation(ifg_model* m, int method);
/* FunctionObject::SetSubsampling for every component (imfit --no-subsampling) */
int ifg_set_subsampling(ifg_model* m, int do_subsampling);
/* GetNParams (core/model_object.cpp:2398) */
int ifg_get_nparams(const ifg_model* m);

/* ---- image setup ---------------------------------------------------------- */
/* AddPSFVector (core/model_object.cpp:827-876): must precede image setup */
int ifg_add_psf(ifg_model* m, const double* psf, int n_columns, int n_rows, int normalize);
/* SetupModelImage (core/model_object.cpp:488-527): make-image mode, no data */
int ifg_setup_model_image(ifg_model* m, int n_columns, int n_rows);
/* AddImageDataVector (core/model_object.cpp:462-477) */
int ifg_add_image_data(ifg_model* m, const double* data, int n_columns, int n_rows);
/* AddImageCharacteristics (core/model_object.cpp:532-547) */
int ifg_add_image_characteristics(ifg_model* m, double gain, double read_noise, double exp_time, int n_combined,
                                  double original_sky);
/* AddErrorVector (core/model_object.cpp:573-619) */
int ifg_add_error_vector(ifg_model* m, const double* errors, long n_values, int input_type);
/* AddMaskVector (core/model_object.cpp:721-776) */
int ifg_add_mask_vector(ifg_model* m, const double* mask, long n_values, int input_type);
/* AddOversampledPsfInfo (core/model_object.cpp:881-955) for one region
 * x1:x2, y1:y2 (1-based, inclusive, data-image coordinates) */
int ifg_add_oversampled_psf(ifg_model* m, const double* psf, int n_columns, int n_rows, int scale, int x1, int x2,
                            int y1, int y2, int normalize);
/* UseCashStatistic / UsePoissonMLR / UseModelErrors (core/model_object.cpp:1500-1598) */
int ifg_use_cash_statistic(ifg_model* m);
int ifg_use_poisson_mlr(ifg_model* m);
int ifg_use_model_errors(ifg_model* m);
/* FinalSetupForFitting (core/model_object.cpp:967-1079); returns 0, or
 * IFG_ERR_DATA for bad weights (-1 in the reference), non-finite unmasked data
 * (-2) or no valid pixels (-3) */
int ifg_final_setup(ifg_model* m);
/* WhichFitStatistic (core/model_object.cpp:1601-1617), non-verbose form */
int ifg_which_fit_statistic(const ifg_model* m);
/* GetNDataValues / GetNValidPixels (core/model_object.cpp:2405-2420) */
long ifg_get_ndata(const ifg_model* m);
long ifg_get_nvalid(const ifg_model* m);

/* ---- bootstrap resampling ------------------------------------------------- */
/* init_genrand (core/mersenne_twister.cpp:60-76): seed the process-wide MT19937
 * that bootstrap samples are drawn from, as core/bootstrap_errors.cpp:174-178 does
 * with imfit's --seed. */
void ifg_init_genrand(unsigned long seed);
/* UseBootstrap (core/model_object.cpp:2042-2053): bootstrap mode on, with a first
 * sample drawn.  After ifg_final_setup; IFG_ERR_UNSUPPORTED with model errors. */
int ifg_use_bootstrap(ifg_model* m);
/* MakeBootstrapSample (core/model_object.cpp:2061-2087): nValid draws of
 * floor(genrand_real2() * nData), masked pixels redrawn.  The statistic then sums
 * each drawn pixel's term once per draw (as ChiSquared / CashStatistic and
 * ComputeDeviates do over the index list, core/model_object.cpp:1443-1454,
 * 1651-1683, 1715-1760).  ComputeDeviates and the Jacobian stay per data pixel:
 * each pixel's deviate is multiplied by sqrt(times drawn), so their sums of
 * squares, J^T J and the fits are those of the resampled data. */
int ifg_make_bootstrap_sample(ifg_model* m);
/* The same with a caller's sample: n = nValid data-pixel indices, each unmasked.
 * No counterpart in the reference (its samples come from its own generator). */
int ifg_set_bootstrap_indices(ifg_model* m, const long* indices, long n);
/* The current sample's nValid indices, in draw order.  No counterpart in the reference. */
int ifg_get_bootstrap_indices(const ifg_model* m, long* indices);
/* Back to the full data set.  No counterpart in the reference (a ModelObject stays
 * in bootstrap mode once UseBootstrap is called). */
int ifg_stop_bootstrap(ifg_model* m);

/* ---- hot path ------------------------------------------------------------- */
/* CreateModelImage (core/model_object.cpp:1103-1229) */
int ifg_create_model_image(ifg_model* m, const double* params);
/* GetModelImageVector (core/model_object.cpp:1310-1335): data-size image of the
 * last CreateModelImage / ComputeDeviates / GetFitStatistic call */
int ifg_get_model_image(ifg_model* m, double* out);
/* GetExpandedModelImageVector (core/model_object.cpp:2297-2305): the model on the
 * reference's PSF-padded grid (nColumns + 2 P_c) x (nRows + 2 P_r).  Without a PSF that
 * grid is the data grid and this equals ifg_get_model_image.  With a PSF the padded
 * border is not computed on the GPU (its planes hold the data grid and the PSF halo
 * only): IFG_ERR_UNSUPPORTED, and the caller keeps the reference's CPU path for it. */
int ifg_get_expanded_model_image(ifg_model* m, double* out);
/* ComputeDeviates (core/model_object.cpp:1413-1496): y[nData] */
int ifg_compute_deviates(ifg_model* m, const double* params, double* y);
/* GetFitStatistic (core/model_object.cpp:1620-1626): chi^2, Cash or Poisson-MLR */
int ifg_fit_statistic(ifg_model* m, const double* params, double* out);
/* Population form of GetFitStatistic for the callers that evaluate many
 * parameter vectors per step: DE candidates (solvers/DESolver.cpp:222-252),
 * DREAM chains (cdream/dream.cpp:268-300), LM finite-difference columns
 * (solvers/mpfit.cpp:1241-1343).  params: n_vectors x nParams, row-major. */
int ifg_fit_statistic_batch(ifg_model* m, const double* params, int n_vectors, double* out);
/* GetFitStatistic that allocates nothing.  No counterpart in the reference; it is
 * for callers that cannot allocate device or pinned memory in the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  The
 * per-call memory of ifg_fit_statistic (complex planes, sub-pixel values, partial
 * sums, call tables, results) comes from one fixed array per GPU of
 * IFG_NOALLOC_SCRATCH_BYTES that the library holds until the process ends; the
 * kernels are the same and launch eagerly, and the result equals
 * ifg_fit_statistic's.  The handle's resident state (data, weights, PSF spectra)
 * must exist: one ordinary ifg_fit_statistic call on the handle creates it
 * (IFG_ERR_STATE otherwise).  An evaluation whose work memory exceeds the array
 * returns IFG_ERR_UNSUPPORTED.  All calls on one GPU share the array: calls that
 * overlap on one GPU race. */
#define IFG_NOALLOC_SCRATCH_BYTES (32u << 20)
int ifg_fit_statistic_noalloc(ifg_model* m, const double* params, double* out);

/* ---- derivatives (no counterpart in the reference: mpfit differentiates by
 * finite differences, solvers/mpfit.cpp:1102-1356) ----------------------------
 * Exact derivatives of the deviates w.r.t. the parameters: the analytic
 * derivative of every image function (averaged over the same sub-pixel samples
 * as the value), pushed through the PSF convolution and the oversampled
 * regions by linearity.  ifree[j] lists the parameter of column j (NULL: the
 * first n_free parameters); n_free <= 48. */
/* Jacobian for mpfit's user-derivative hook (mp_par side == 3,
 * solvers/mpfit.cpp:1208-1232): fjac column j at fjac + j*nData holds
 * d(deviate)/d(param ifree[j]); fvec (nData, may be NULL) gets the deviates. */
int ifg_jacobian(ifg_model* m, const double* params, int n_free, const int* ifree, double* fvec, double* fjac);
/* Gauss-Newton normal equations reduced on the GPU in FP64: jtj = J^T J
 * (n_free x n_free, row-major), jtr = J^T r with r the deviates, stat = the
 * fit statistic at params. */
int ifg_normal_equations(ifg_model* m, const double* params, int n_free, const int* ifree, double* jtj, double* jtr,
                         double* stat);
/* Gradient of GetFitStatistic w.r.t. all nParams parameters (and the statistic).
 * Default method: reverse mode (adjoint) -- one forward evaluation that keeps
 * d(statistic)/d(model) per data pixel, then one pass backwards through the model:
 * the oversampled regions' block means spread over their sub-pixels, the PSF
 * convolution's adjoint (correlation with the PSF: the FFT passes with the
 * conjugate PSF spectrum, core/convolver.cpp:372-379 transposed), main-grid pixels
 * that a region replaces contributing nothing (core/downsample.cpp:96-104), and the
 * analytic derivatives of every component summed against the result.  Its cost does
 * not grow with the number of parameters. */
int ifg_fit_statistic_gradient(ifg_model* m, const double* params, double* grad, double* stat);
/* Method of ifg_fit_statistic_gradient: IFG_GRADIENT_ADJOINT (default) or
 * IFG_GRADIENT_TANGENT (one derivative plane per parameter through the whole
 * forward path, as ifg_jacobian).  Both are exact; they differ in rounding only. */
#define IFG_GRADIENT_ADJOINT 0
#define IFG_GRADIENT_TANGENT 1
int ifg_set_gradient_method(ifg_model* m, int method);
/* Vector-Jacobian product of ComputeDeviates: grad[k] = sum_i u[i] d(deviate_i)/d(param k)
 * over the nData deviates, by the adjoint pass; fvec (nData, may be NULL) gets the
 * deviates.  For torch's backward of the deviates (J^T u without forming J). */
int ifg_deviates_vjp(ifg_model* m, const double* params, const double* u, double* grad, double* fvec);
/* Vector-Jacobian product of CreateModelImage (the nData-pixel model image that
 * ifg_get_model_image returns): grad[k] = sum_i u[i] d(model_i)/d(param k). */
int ifg_model_image_vjp(ifg_model* m, const double* params, const double* u, double* grad);

/* ---- Levenberg-Marquardt callers ----------------------------------------------
 * Parameter constraints: the fixed / limited / limits fields of mpfit's mp_par
 * (core/param_struct.h:21-32).  NULL: every parameter free and unbounded. */
typedef struct {
  int fixed;
  int limited[2];
  double limits[2];
} ifg_param_info;
/* Tolerances (<= 0: mpfit's defaults 1e-10, solvers/mpfit.cpp:377-388; imfit passes
 * ftol = 1e-8 and maxiter = 1000, solvers/levmar_fit.cpp:38,99-101). */
typedef struct {
  double ftol, xtol, gtol;
  int maxiter, maxfev;
} ifg_lm_config;
/* mp_result subset (solvers/mpfit.h:152-166) */
typedef struct {
  double bestnorm, orignorm;
  int status, niter, nfev, nfree, npegged;
} ifg_lm_result;
/* status values: mpfit's (solvers/mpfit.h:109-129) */
#define IFG_LM_OK_CHI 1
#define IFG_LM_OK_PAR 2
#define IFG_LM_OK_BOTH 3
#define IFG_LM_OK_DIR 4
#define IFG_LM_MAXITER 5
#define IFG_LM_FTOL 6
#define IFG_LM_XTOL 7
#define IFG_LM_GTOL 8
#define IFG_LM_ERR_NAN (-16)
#define IFG_LM_ERR_NFREE (-19)
#define IFG_LM_ERR_INITBOUNDS (-21)
#define IFG_LM_ERR_BOUNDS (-22)
#define IFG_LM_ERR_DOF (-24)
#define IFG_LM_JACOBIAN_FD 0        /* one GPU ComputeDeviates per free parameter (mp_fdjac2) */
#define IFG_LM_JACOBIAN_ANALYTIC 1  /* ifg_jacobian (mpfit's side == 3 hook) */
/* mpfit (solvers/mpfit.cpp:352-1097) on the GPU model: MINPACK LMDIF with QR on
 * the host; params (nParams) is updated in place; returns the mpfit status. */
int ifg_fit_lm(ifg_model* m, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
               int jacobian_mode, ifg_lm_result* res);
/* mpfit on deviates supplied by the caller: fn(ctx, npar, params, m, fvec) fills the m
 * deviates at params and returns a negative value on failure.  Forward-difference
 * Jacobian (mp_fdjac2, solvers/mpfit.cpp:1241-1300), the same code as ifg_fit_lm.  For
 * callers whose deviates combine several models, such as ModelObjectMultImage's
 * ComputeDeviates (core/model_object_multimage.cpp:411-438) in multimfit's L-M
 * (core/multimfit_main.cpp calling mpfit through solvers/levmar_fit.cpp:54-60). */
typedef int (*ifg_deviates_fn)(void* ctx, int npar, const double* params, long m, double* fvec);
int ifg_mpfit(long m, int npar, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
              ifg_deviates_fn fn, void* ctx, ifg_lm_result* res);
/* Levenberg-Marquardt on the Gauss-Newton normal equations J^T J, J^T r reduced
 * on the GPU in FP64 (ifg_normal_equations); no counterpart in the reference. */
int ifg_fit_lm_normal(ifg_model* m, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                      ifg_lm_result* res);
/* ifg_fit_lm_normal with mpfit's forward-difference Jacobian in place of the exact one
 * (mp_fdjac2's steps, solvers/mpfit.cpp:1241-1300): the nFree + 1 evaluations of each
 * Jacobian run as one batched launch and J^T J, J^T r are reduced on the GPU.  The
 * plain-GPU finite-difference LM that the exact Jacobian is measured against; no
 * counterpart in the reference. */
int ifg_fit_lm_normal_fd(ifg_model* m, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                         ifg_lm_result* res);

/* ---- Differential Evolution (DiffEvolnFit, solvers/diff_evoln_fit.cpp:95-168, and
 * DESolver, solvers/DESolver.cpp) ------------------------------------------------
 * imfit's DE: rand-to-best/1/exp, F 0.7, CR 1, 8 x nFree population uniform within the
 * limits (both limits required for every free parameter), MT19937 seeded with seed
 * (0: the time), convergence when the best statistic's relative change, taken every
 * 10 generations from generation 20, is below ftol three times in a row.
 * IFG_DE_SEQUENTIAL: DESolver::Solve's loop, one evaluation per trial, a better trial
 * replacing its candidate at once.  IFG_DE_GENERATIONAL (default): each generation's
 * trials are bred from the population as the generation found it (the same draws in
 * the same order) and evaluated together by one ifg_fit_statistic_batch launch.
 * Returns the status (IFG_DE_CONVERGED or IFG_DE_MAXGEN) or an error; params receive
 * the best vector. */
#define IFG_DE_GENERATIONAL 0
#define IFG_DE_SEQUENTIAL 1
#define IFG_DE_CONVERGED 1
#define IFG_DE_MAXGEN 5
#define IFG_DE_ERR_LIMITS (-12)   /* a free parameter without both limits (DiffEvolnFit's -2) */
typedef struct {
  double ftol;           /* <= 0: 1e-8 */
  unsigned long seed;    /* 0: the time */
  int max_generations;   /* <= 0: 600 */
  int pop_per_param;     /* <= 0: 8 */
  int mode;              /* IFG_DE_GENERATIONAL or IFG_DE_SEQUENTIAL */
} ifg_de_config;
typedef struct {
  double best;           /* the best statistic */
  long nfev;             /* generations x population, as imfit reports it */
  int status, generations, population;
} ifg_de_result;
int ifg_fit_de(ifg_model* m, double* params, const ifg_param_info* pinfo, const ifg_de_config* cfg,
               ifg_de_result* res);

/* ---- batches of independent models (no counterpart in the reference, which fits
 * one image per run) ------------------------------------------------------------
 * Many ifg_models -- e.g. cutouts of different sizes, each set up through the calls
 * above up to ifg_final_setup -- evaluated together: every merged launch covers the
 * grids of all of them.  Parameters of all models are concatenated in the order the
 * models were added.  The models share the device, the pixel precision, the fit
 * statistic and the sub-pixel switch; the batch uses them but does not own them
 * (destroy the batch before its models). */
typedef struct ifg_batch ifg_batch;
ifg_batch* ifg_batch_create(int device);
void ifg_batch_destroy(ifg_batch* b);
const char* ifg_batch_last_error(const ifg_batch* b);
/* append a set-up model; returns its index in the batch */
int ifg_batch_add(ifg_batch* b, ifg_model* m);
int ifg_batch_size(const ifg_batch* b);
/* GetFitStatistic of every model: out[nModels] */
int ifg_batch_fit_statistic(ifg_batch* b, const double* params, double* out);
/* Normal equations of every model w.r.t. its n_free[k] parameters listed in ifree
 * (lists concatenated; NULL: every parameter of every model), per model the packed
 * upper triangle of J^T J (row-major) then J^T r, concatenated; stat[nModels]. */
int ifg_batch_normal_equations(ifg_batch* b, const double* params, const int* n_free, const int* ifree,
                               double* normal, double* stat);
/* ifg_fit_lm_normal on every model at once, each step of all models that are still
 * iterating in one batched evaluation: pinfo concatenated per model (NULL: all free
 * and unbounded), params updated in place, res[nModels].  Returns the number of
 * models that ended without a convergence status (0: all converged), or an error.
 * Replaces mpfit's per-model loop (solvers/mpfit.cpp:610-746) for many models.  By
 * default the host drives the rounds (the LM steps on the host, one synchronisation
 * per round); ifg_batch_set_lm_resident keeps the fits resident on the GPU instead. */
int ifg_batch_fit_lm(ifg_batch* b, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                     ifg_lm_result* res);
/* ifg_batch_fit_lm's driver: 0 (default) the host-driven rounds (the LM step on the
 * host, one synchronisation per round), on != 0 the GPU-resident fits: staging,
 * evaluation and every LM step (trust region, Levenberg-Marquardt parameter, Cholesky
 * solve, limits) on the device, with one host wait per 8 rounds.  Both take the same
 * steps; the resident fits fall back to the host rounds when a model's sub-pixel
 * boxes outgrow the batch's layout.  The resident driver's LM step runs in one
 * wavefront per model (about 0.1 ms a round): it gains where the host's per-round
 * work over many models outweighs that, and loses where the rounds serve a few slowly
 * converging fits. */
int ifg_batch_set_lm_resident(ifg_batch* b, int on);
/* as ifg_set_graph_mode, for the batch's own launches */
int ifg_batch_set_graph_mode(ifg_batch* b, int on);

/* ---- instrumentation (no counterpart in the reference, whose only timing is
 * wall-clock, core/makeimage_main.cpp:351-363) -------------------------------- */
/* Record HIP events around every kernel launch of this handle (on its stream);
 * switching on clears the totals. */
int ifg_set_kernel_timing(ifg_model* m, int on);
/* Replay each evaluation shape from a captured HIP graph (default on; the
 * environment variable IFG_NO_GRAPH=1 switches the default off).  Off: launch
 * the same kernels eagerly every call. */
int ifg_set_graph_mode(ifg_model* m, int on);
/* Per-kernel totals: fills name / total milliseconds / launch count of entry
 * `index` and returns the number of entries (0 when index is out of range). */
int ifg_get_kernel_timing(const ifg_model* m, int index, char* name, int name_len, double* total_ms,
                          long* launches);

/* ---- launch-shape autotuning (no counterpart in the reference; the FFT convolver
 * and the per-pixel loops of core/model_object.cpp have no launch shapes).
 * Each convolution grid (main image, oversampled regions) has two launch shapes: the
 * column pass's strip width (columns per workgroup: 1, 2, 4 or 8) and the row passes'
 * rows per workgroup.  ifg_autotune times this model's GetFitStatistic of n_vectors
 * vectors at params under nearby shapes (one grid and one shape at a time), keeps the
 * fastest, and records the choice for the geometry in this process: a later call for
 * a model with the same grids applies the record without timing.  Returns the number
 * of timed trials (0: taken from the record) or an error.  Results depend on the
 * shapes only through the grouping of the statistic's reduction (last bits). */
int ifg_autotune(ifg_model* m, const double* params, int n_vectors);
/* The current shapes of grid `grid` (0: main image, 1..: oversampled regions) on the
 * single-evaluation grids (batch_grids 0) or the population grids (1). */
int ifg_get_launch_shape(const ifg_model* m, int grid, int batch_grids, int* cols_per_block,
                         int* rows_per_block);
/* Set the launch shape of a grid as ifg_get_launch_shape reports it, e.g. a shape
 * found by ifg_autotune in an earlier process.  The grid must exist (one evaluation
 * with the same n_vectors class first).  IFG_ERR_ARG if the shape does not fit.
 * No counterpart in the reference (its CPU path has no launch shapes). */
int ifg_set_launch_shape(ifg_model* m, int grid, int batch_grids, int cols_per_block, int rows_per_block);
/* As ifg_autotune for a batch's ifg_batch_fit_statistic: one strip width and one
 * multiplier of each grid's default rows per workgroup for every model (grids where
 * the choice does not fit keep their defaults).  Returns the number of trials. */
int ifg_batch_autotune(ifg_batch* b, const double* params, int* cols_per_block, double* rows_factor);

#ifdef __cplusplus
}
#endif
#endif /* IMFIT_GPU_H */
