/* qlx -- MI355X-native pixellated-source log-evidence (drop-in for QLens's inversion block).
 *
 * C ABI: plain pointers and sizes, no C++ or HIP types.  One context owns one HIP stream and
 * all device buffers; a context is not thread-safe, many contexts may coexist.  Every entry
 * point returns a qlx_status (0 = ok, negative = error) and never exits the process (the
 * reference's die() -> exit(1), errors.cpp:25-54, becomes a status code here).
 *
 * The path replaced is the Cartesian / DENSE / curvature / 3-point-interpolate block of
 * QLens::invert_image_surface_brightness_map (lens.cpp:12364-12615) plus its chi-square loop
 * (lens.cpp:12950-13015); a QLens-side adapter calls qlx_eval() in its place.
 * Conventions are QLens's: lens emode 1, lens position angle measured from +y
 * (profile.cpp:20, 1625-1637), shear angle theta_eff = theta + 90 deg (models.cpp:2212-2224),
 * images stored x-major (value of pixel (i,j) at [i*ny + j]) like ImagePixelData's [i][j].
 */
#ifndef QLX_H
#define QLX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    QLX_OK = 0,
    QLX_ERR_ARG = -1,      /* invalid argument (null pointer, bad size) */
    QLX_ERR_STATE = -2,    /* call out of order (e.g. eval before set_image) */
    QLX_ERR_HIP = -3,      /* HIP runtime error or no device */
    QLX_ERR_NOMEM = -4,    /* device allocation failed / admission refused */
    QLX_ERR_NOT_PD = -5,   /* F or R not positive definite (reference: warn + die, pixelgrid.cpp:17095-17099, 17384) */
    QLX_ERR_NOSOURCE = -6, /* no source pixel mapped (reference sentinel 2e30, pixelgrid.cpp:12807) */
    QLX_ERR_RANGE = -7,    /* size above a compiled-in maximum */
    QLX_ERR_REFINE = -8,   /* MIXED item: the FP64 refinement stopped above refine_tol (value kept) */
    QLX_ERR_IO = -9,       /* a file could not be read or is not in the expected format */
    QLX_ITEMS_FAILED = 1   /* return of qlx_eval*: the call completed, some items carry a non-ok status */
} qlx_status;

/* lens types */
enum { QLX_LENS_SIE = 1, QLX_LENS_NFW = 2, QLX_LENS_SHEAR = 3, QLX_LENS_SPLE = 4, QLX_LENS_SHEAR_COMP = 5 };
/* how a numerical profile's deflection is obtained */
enum { QLX_DEF_QUADRATURE = 0, QLX_DEF_SURROGATE = 1 };
/* precision modes */
enum {
    QLX_PREC_FP64 = 0,  /* parity mode: FP64 everywhere (MFMA f64 for F and Cholesky) */
    QLX_PREC_MIXED = 1  /* F from an FP16 pool on MFMA with FP32 accumulation, FP32 Cholesky + FP64
                           iterative refinement, FP64 log-dets and everything after the solve */
};
/* ray-shooting arithmetic */
enum { QLX_RAY_FP64 = 0, QLX_RAY_FP32 = 1 };

#define QLX_MAX_LENSES 8

/* One lens in QLens parameters (replaces LensProfile + its subclass, profile.h / models.cpp):
 *   SIE   p = {b, q, theta_deg, xc, yc, s}     SPLE_Lens with alpha = 1, cored when s > 0
 *                                              (models.cpp:101-230)
 *   NFW   p = {ks, rs, q, theta_deg, xc, yc}   NFW pmode 0 (models.cpp:953-1010)
 *   SHEAR p = {gamma, theta_deg, xc, yc}       Shear (models.cpp:2140-2272)
 *   SPLE  p = {b, alpha, q, theta_deg, xc, yc} SPLE_Lens ('lens alpha' / 'lens sple') without a
 *                                              core: Tessore et al. series for q < 1
 *                                              (models.cpp:244-351), the spherical kappa average
 *                                              for q = 1 (models.cpp:169-172); alpha in (0, 2);
 *                                              alpha = 1 is the SIE formula, as the reference
 *                                              dispatches it (models.cpp:141-147)
 *   SHEAR_COMP p = {shear1, shear2, xc, yc}    Shear with 'shear_components on' (models.cpp:2215-2217)
 * def_mode applies to NFW only: Gauss-Patterson quadrature (profile.cpp:2359-2416) or the
 * trained surrogate (falls back to quadrature outside its training box). */
typedef struct {
    int type;
    int def_mode;
    double p[6];
} qlx_lens;

/* A full lens model: the reference's lens_list (qlens.h), one lens plane, zfactor 1. */
typedef struct {
    int nlens;
    qlx_lens lens[QLX_MAX_LENSES];
} qlx_lens_model;

typedef struct {
    int device;             /* HIP device ordinal */
    size_t hbm_budget;      /* bytes this context may use for batch admission (0 = 85% of free) */
} qlx_opts;

/* Inversion settings (the QLens fields the block reads, qlens.h:750-893):
 * nsplit = default_imgpixel_nsplit (0 = split_imgpixels off: one ray at each pixel centre,
 * pixelgrid.cpp:9659-9666, 13277-13285); source grid = srcgrid_x/ymin/max and srcgrid_npixels_x/y;
 * regparam = the pixsrc regparam; integral_tol = integral_tolerance; curvature regularization,
 * 3-point interpolation, dense inversion. */
typedef struct {
    int nsplit;
    double src_xmin, src_xmax, src_ymin, src_ymax;
    int src_nx, src_ny;
    double regparam;
    double integral_tol;
    int precision;          /* QLX_PREC_* */
    int ray_precision;      /* QLX_RAY_* */
    int refine_iters;       /* MIXED: max FP64 iterative-refinement steps (<= 0: default 6) */
    double refine_tol;      /* MIXED: relative residual target (<= 0: default 1e-10) */
    /* Regularization-parameter optimisation (optimize_regparam on; optimize_regularization_parameter,
     * pixelgrid.cpp:15865-15979): per item, Brent's method over log10(lambda) in
     * [regopt_minlog, regopt_maxlog] minimising chi^2 over the active pixels + lambda Es - N ln lambda
     * - ln det R + ln det F (chisq_regparam_dense, 16526-16699; brents_min_method, 16849-16932), then
     * the evaluation at the optimum.  FP64 precision only.  Defaults as the reference's
     * (lens.cpp:423-426): bounds -3 and 5 when regopt_maxlog <= regopt_minlog, tolerance 0.01 when
     * regopt_tol <= 0, at most 20 iterations when regopt_max_iter <= 0. */
    int optimize_regparam;
    double regopt_minlog, regopt_maxlog, regopt_tol;
    int regopt_max_iter;
    /* Automatic source grid, per item from its own rays (lens.cpp:12442-12458):
     * auto_srcgrid (auto_srcgrid on; ImagePixelGrid::find_optimal_sourcegrid,
     *   pixelgrid.cpp:10998-11080): the bounding box of the mask pixels' mean subpixel source
     *   points, widened by 10 % per axis, replaces src_x/ymin/max.  With noise_threshold > 0,
     *   pixels whose data is at most noise_threshold x their noise (1/sqrt(covinv)) are left out.
     * auto_src_npixels (auto_src_npixels on; find_optimal_sourcegrid_npixels, 11469-11486):
     *   with count = mask pixels whose centre ray lies strictly inside the box,
     *   nx = 2 (int)sqrt(pixel_fraction count dx/dy), ny = 2 (int)(nx/2 dy/dx), replacing
     *   src_nx/src_ny (and the per-item grids of qlx_eval_batch_grids).  A grid below 2 cells per
     *   axis fails its item with QLX_ERR_NOSOURCE (lens.cpp:12453-12458); above 512 per axis, or
     *   beyond QLX_MAX_LD active cells, with QLX_ERR_RANGE.
     * pixel_fraction <= 0 means the reference's default 0.3 (pixelgrid.cpp:559). */
    int auto_srcgrid, auto_src_npixels;
    double pixel_fraction, noise_threshold;
    /* MIXED: panel width of the FP32 Cholesky in columns (512, 1024 or 2048).  0 = autotune: the
     * first wave of each size class (F's padded dimension) times the three widths on its own
     * matrices and keeps the fastest for the context's later waves; the factor, hence every
     * result, is bitwise the same for every width. */
    int chol_panel;
    /* The factorizations' launch sequences (F's Cholesky, R's banded one) are captured in a HIP
     * graph per size class on first use and replayed (one launch instead of hundreds); no_graphs = 1
     * launches them one by one instead.  The results are bitwise the same either way. */
    int no_graphs;
} qlx_settings;

/* Per-evaluation result (the scalars QLens fills, qlens.h:750-761, 891). */
typedef struct {
    double minus2lnE;   /* -2 ln E without noise constants (lens.cpp:13033-13043 off), or the
                           sentinel 2e30 for an item whose status is an error: no or too small a
                           source grid (lens.cpp:12453-12458, pixelgrid.cpp:12807), invalid lens
                           parameters, a source grid beyond QLX_MAX_LD, or F / R not positive
                           definite (where the reference dies, pixelgrid.cpp:17384).  Items with
                           QLX_ERR_REFINE keep their value (accurate to refine_resid). */
    double chisq0;      /* chi-square of the best-fit model over mask pixels */
    double logdetF, logdetR, Es; /* ln det F, ln det R, s^T R s */
    double refine_resid;/* MIXED: final relative residual of the refined solve (FP64: 0) */
    int n_img_active;   /* image_npixels */
    int n_src_active;   /* source_npixels */
    int status;         /* qlx_status of this item */
    int refine_steps;
    double regparam;    /* lambda of the evaluation: the settings' value, or the optimised one */
    int regopt_evals;   /* objective evaluations of the regparam optimisation (0 without it) */
    int src_nx, src_ny; /* the item's source grid: cells per axis ... */
    int pad_;
    double src_xmin, src_xmax, src_ymin, src_ymax; /* ... and its bounds (the settings' or the
                           automatic ones; the reference's sourcegrid_x/ymin/max after the call) */
} qlx_result;

/* Sizes of the structs above as this library was compiled, in the order qlx_lens, qlx_lens_model,
 * qlx_opts, qlx_settings, qlx_result, qlx_fit_param, qlx_surrogate_cfg, qlx_stats, for bindings to
 * check their mirrors against (no counterpart in the reference, whose callers link C++).  Writes
 * min(n, 8) sizes and returns 8. */
int qlx_struct_sizes(size_t *sizes, int n);

/* ---- context ---------------------------------------------------------------------------- */
int qlx_ctx_create(const qlx_opts *opts, void **ctx);
int qlx_ctx_destroy(void *ctx);
const char *qlx_last_error(void *ctx);

/* ---- inputs (copied at call time) ----------------------------------------------------------
 * qlx_set_image replaces ImagePixelData (pixelgrid.h:513-641, loaded by load_data
 * pixelgrid.cpp:5797-5904): corners xvals[nx+1], yvals[ny+1]; data, covinv (1/sigma^2, the
 * noise map's covinv_map) and mask (1 = in mask; NULL = all) are x-major [nx*ny].
 * Limits: nx*ny <= 2^24 and nx, ny <= 65535. */
int qlx_set_image(void *ctx, int nx, int ny, const double *xvals, const double *yvals,
                  const double *data, const double *covinv, const unsigned char *mask);
/* PSF matrix psf[i*pny + j] = QLens psf_matrix[i][j] (use_input_psf_matrix, qlens.h:877-880). */
int qlx_set_psf(void *ctx, int pnx, int pny, const double *psf);
/* Gaussian PSF built as QLens::generate_PSF_matrix (pixelgrid.cpp:14601-14668). */
int qlx_set_psf_gaussian(void *ctx, double sigma_x, double sigma_y, double threshold);
int qlx_set_settings(void *ctx, const qlx_settings *s);

/* ---- evaluation ----------------------------------------------------------------------------
 * Replaces one call of QLens::invert_image_surface_brightness_map (lens.cpp:12293) per model:
 * nbatch independent lens models (e.g. sampler proposals) are evaluated on the same image and
 * settings.  Synchronous.  Returns QLX_OK when every item succeeded, QLX_ITEMS_FAILED (> 0) when
 * the call completed but some items carry an error status in out[b].status (their minus2lnE
 * holds the 2e30 sentinel), or a negative qlx_status when the call as a whole failed (bad
 * arguments, an unknown lens type, no surrogate loaded, HIP or memory errors).  Lens parameters
 * that are not finite, or outside the profile's domain (q in (0, 1], rs > 0), fail their own item
 * with QLX_ERR_ARG.  Items are processed in waves sized to the HBM budget; an item's result does
 * not depend on the other items or on the wave it lands in. */
int qlx_eval(void *ctx, int nbatch, const qlx_lens_model *models, qlx_result *out);

/* Batched evaluation with per-item source grids (uneven source sizes in one submission).
 * src_n[2*b], src_n[2*b+1] = source cells of item b (bounds from the settings; ignored, and src_n
 * may be NULL, when the settings ask for auto_src_npixels). */
int qlx_eval_batch_grids(void *ctx, int nbatch, const qlx_lens_model *models, const int *src_n,
                         qlx_result *out);

/* Read back item `item` (0 .. nbatch-1) of the last qlx_eval / qlx_eval_batch_grids call.  Every
 * successful item of that call is kept on the device until the next evaluation starts.
 *   src_nx, src_ny  (out, may be NULL) the item's own source grid
 *   src_cells       amplitude per cell, x-major [src_nx*src_ny], 0 for inactive cells
 *                   (update_source_amplitudes, pixelgrid.cpp:17850-17881); src_cap = its length
 *                   in doubles, which must be at least src_nx*src_ny
 *   active_index    the reference's active index of each cell (x-major, -1 = inactive;
 *                   SourcePixel::assign_active_indices, pixelgrid.cpp:3586-3626), same capacity
 *   model_image     the model image, x-major [nx*ny] (store_image_pixel_surface_brightness,
 *                   pixelgrid.cpp:18483-18498); model_cap = its length, at least nx*ny
 * Any array pointer may be NULL (its capacity is then ignored).  QLX_ERR_ARG for a capacity
 * below the item's size or an item out of range; QLX_ERR_STATE if the item has no solution (its
 * status was an error, or no evaluation has run since the image, PSF or settings changed). */
int qlx_get_source(void *ctx, int item, int *src_nx, int *src_ny, double *src_cells,
                   int *active_index, size_t src_cap, double *model_image, size_t model_cap);

/* Source-plane points of every subpixel ray of item `item` of the last evaluation, as the
 * reference keeps them in ImagePixelGrid::subpixel_center_sourcepts after an inversion
 * (calculate_sourcepts_and_areas, pixelgrid.cpp:10434-10441): beta[((i*ny + j)*nsplit^2 + k)*2 + c]
 * for pixel (i, j), subpixel k = ii*nsplit + jj, component c; 0 for pixels outside the mask.
 * Available for the items of the call's last wave (every item unless HBM admission split the
 * call); cap = length of beta in doubles, at least nx*ny*nsplit^2*2. */
int qlx_get_rays(void *ctx, int item, double *beta, size_t cap);

/* The PSF matrix in use (x-major [pnx*pny], as passed to qlx_set_psf or built by
 * qlx_set_psf_gaussian); psf may be NULL to query the size only. */
int qlx_get_psf(void *ctx, int *pnx, int *pny, double *psf, size_t cap);

/* ---- on-disk inputs (host only, no context or device needed) -------------------------------
 * Two-call pattern: with NULL arrays the sizes are returned.  Errors: QLX_ERR_IO (unreadable or
 * malformed file), QLX_ERR_ARG, QLX_ERR_RANGE; the message is qlx_io_last_error(). */
const char *qlx_io_last_error(void);
/* Text image, replaces ImagePixelData::load_data (pixelgrid.cpp:5797-5904): root.x and root.y
 * hold the nx+1 / ny+1 pixel corners, root.dat the nx*ny surface brightness values in x-major
 * order (i outer, j inner), all whitespace-separated.  The mask is every pixel. */
int qlx_read_image_text(const char *root, int *nx, int *ny, double *xvals, double *yvals,
                        double *sb, size_t sb_cap);
/* FITS primary image as cfitsio's fits_read_pix into doubles reads it for load_data_fits,
 * load_mask_fits and load_psf_fits (pixelgrid.cpp:6016-6280, 6512-6651, 6737-6876): BITPIX 8,
 * 16, 32, 64, -32 or -64, NAXIS 1 or 2 (further axes: the first plane), BSCALE / BZERO applied;
 * pix[i*naxis2 + j] = pixel (i+1, j+1), axis 1 along x.  pxsize / pxnoise (may be NULL) receive
 * the reference's PXSIZE / PXNOISE header keywords, NaN when absent (pixelgrid.cpp:6094-6110). */
int qlx_read_fits(const char *path, int *naxis1, int *naxis2, double *pix, size_t cap,
                  double *pxsize, double *pxnoise);
/* Pixel corners of an nx x ny image with square pixels of pixel_size centred on the origin, as
 * load_data_fits builds them with a pixel size (pixelgrid.cpp:6198-6213): x from -nx s/2 in
 * accumulated steps of s.  xvals[nx+1], yvals[ny+1]. */
int qlx_pixel_grid(int nx, int ny, double pixel_size, double *xvals, double *yvals);
/* The PSF matrix QLens::load_psf_fits keeps from an input image img[nx*ny] (x-major)
 * (pixelgrid.cpp:6790-6840): the box of pixels above threshold x peak (psf_threshold), a window
 * of 2 half + 1 pixels per axis centred on (nx/2, ny/2), normalised to unit sum.  QLX_ERR_RANGE
 * when that window would reach past the input. */
int qlx_psf_from_image(const double *img, int nx, int ny, double threshold, int *pnx, int *pny,
                       double *psf, size_t cap);
/* Applies a mask image m[mnx*mny] (x-major; nonzero = in the mask) to mask[nx*ny] as
 * load_mask_fits does for the primary mask (pixelgrid.cpp:6532-6620): a smaller mask is centred
 * with offsets ((nx-mnx)/2, (ny-mny)/2) and leaves the pixels it does not cover as they were;
 * a larger one is refused (QLX_ERR_ARG). */
int qlx_mask_from_image(const double *m, int mnx, int mny, int nx, int ny, unsigned char *mask);

/* ---- fit model and likelihood (sampler interface) --------------------------------------------
 * The reference's fit parameters (vary flags on lens parameters with penalty limits, transforms and
 * priors: ParamSettings, params.h:722-893) and QLens::fitmodel_loglike_extended_source
 * (lens.cpp:10657-10748), which every sampler and optimiser calls through LogLikePtr
 * (mcmchdr.cpp:22).  Here one call takes a whole population of sampler points. */
enum { QLX_TRANSFORM_NONE = 0, QLX_TRANSFORM_LOG10 = 1, QLX_TRANSFORM_LINEAR = 2 };
enum { QLX_PRIOR_UNIFORM = 0, QLX_PRIOR_LOG = 1, QLX_PRIOR_GAUSS = 2 };
typedef struct {
    int lens, param;        /* lens index in the base model, parameter index in its p[] */
    double lo, hi;          /* penalty limits on the physical value (lo >= hi: none) */
    int transform;          /* physical = x, 10^x, or (x - tb) / ta (inverse_transform_parameters) */
    double ta, tb;
    int include_jacobian;   /* LOG10: add -ln(physical) (add_jacobian_terms_to_loglike) */
    int prior;              /* on the sampler coordinate x (add_prior_terms_to_loglike):
                               UNIFORM nothing; LOG ln(norm) + ln x; GAUSS ln(norm) + ((x - mu)/sig)^2/2 */
    double prior_mu, prior_sig, prior_norm;
} qlx_fit_param;
/* Set the model the sampler varies: the base lens model and nparam varied parameters. */
int qlx_set_fit_model(void *ctx, const qlx_lens_model *base, int nparam, const qlx_fit_param *params);
/* -ln L for nbatch sampler points x[b*nparam + k]: 1e30 when a physical value lies outside its
 * penalty limits or the item fails (the reference's sentinels), else (-2 ln E)/2 plus the prior and
 * Jacobian terms.  The points inside their limits are evaluated in one qlx_eval; out (may be NULL)
 * receives each item's result (a point outside its limits: status QLX_ERR_ARG, minus2lnE 2e30). */
int qlx_loglike(void *ctx, int nbatch, const double *x, double *loglike, qlx_result *out);

/* ---- ray shooting and data (building blocks, host buffers) --------------------------------
 * beta = theta - alpha at n points (QLens::find_sourcept, lenscalc.cpp:318-330). */
int qlx_raytrace(void *ctx, const qlx_lens_model *model, int n, const double *xy, double *beta);
/* qlx_raytrace without any device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  The points and results are staged in one file-scope device array per GPU (8 MiB,
 * held until the process ends), so n is at most 262144 (QLX_ERR_RANGE beyond).  Calls on one GPU
 * share that array: calls that overlap in time on the same GPU (two contexts or threads) race. */
int qlx_raytrace_noalloc(void *ctx, const qlx_lens_model *model, int n, const double *xy, double *beta);
/* Lens-frame NFW deflection + Hessian (alpha_x, alpha_y, psi_xx, psi_xy, psi_yy) by quadrature
 * (LensProfile::deflection_and_hessian_numerical, profile.cpp:2532-2560): surrogate targets. */
int qlx_nfw_def_hess(void *ctx, const double *nfw_params, double tol, int n, const double *xy,
                     double *out5);
/* Mock image: ray-traced Sersic sources averaged over nsplit^2 subpixels, PSF-convolved, plus
 * Gaussian noise from the NR3 generator with `seed` (lens.cpp:11882-12066 path;
 * add_pixel_noise pixelgrid.cpp:10987-10996).  sersic[7*k] = {s0, Reff, n, q, theta, xc, yc}. */
int qlx_mock_image(void *ctx, const qlx_lens_model *model, int nsrc, const double *sersic,
                   double sigma, unsigned long long seed, double *img_out);

/* ---- learned deflection surrogate (replaces the tab/qtab option, models.cpp:4054-5148) ----
 * MLP 4 -> 128 x4 (SiLU) -> 5 in the NFW lens frame with ks factored out.  Inputs
 * (ln r, phi, q, ln rs); the five outputs encode (ln J0, ln J1, 2x^2 K0/J0, 2|xy| K1/sqrt(J0 J1),
 * 2y^2 K2/J1) of the deflection/Hessian integrals (profile.cpp:2532-2560), standardized, so that
 * (alpha_x, alpha_y, psi_xx, psi_xy, psi_yy)/ks follow exactly.  BF16 weights, FP32 accumulation.
 * Lenses with def_mode QLX_DEF_SURROGATE use it (one such lens per model); rays outside the
 * training box (r in [1e-3, 20], q in [0.3, 1], rs in [1, 20]) fall back to quadrature. */
typedef struct {
    int n_samples;       /* training samples generated by the quadrature kernel */
    int steps;           /* optimizer steps */
    int batch;           /* minibatch size */
    double lr;           /* Adam learning rate (cosine decay to lr/100 over the steps) */
    double r_min, r_max, q_min, q_max, rs_min, rs_max; /* sampling box */
    unsigned long long seed;
    double tol;          /* quadrature tolerance for the targets */
    int graph;           /* 1: capture one optimizer step in a HIP graph and replay it */
} qlx_surrogate_cfg;
int qlx_surrogate_train(void *ctx, const qlx_surrogate_cfg *cfg, double *final_loss);
int qlx_surrogate_save(void *ctx, const char *path);
int qlx_surrogate_load(void *ctx, const char *path);
/* Forward pass on n lens-frame inputs x[4n] = (ln r, phi, q, ln rs) ->
 * y[5n] = (alpha_x, alpha_y, psi_xx, psi_xy, psi_yy) / ks (FP32). */
int qlx_surrogate_eval(void *ctx, int n, const float *x, float *y);
/* Parameters: W1[128][4] b1[128] W2[128][128] b2[128] W3 b3 W4 b4 W5[5][128] b5[5] (FP32). */
int qlx_surrogate_param_count(void *ctx);
int qlx_surrogate_get_params(void *ctx, float *params);
int qlx_surrogate_set_params(void *ctx, const float *params);
/* One forward+backward of the network itself on a batch: x[4n] are the standardized features
 * (each in [-1, 1]), y[5n] the standardized targets; writes the mean-squared-error loss over the
 * n x 5 outputs and its FP32 gradient with respect to every parameter (the gradient check). */
int qlx_surrogate_grad(void *ctx, int n, const float *x, const float *y, float *loss, float *grad);

/* ---- timing support for benchmarks ---------------------------------------------------------
 * Average device time (ms) per launch of the named stage over the last qlx_eval, from HIP events
 * on the context's stream; one launch covers one wave (every item of the wave); names:
 * "raytrace", "lbuild" (L rows, support masks and the blurred-L pool), "lbkernel" (the pool
 * build kernel alone), "syrk", "cholesky", "solve", "total", "surrogate". */
int qlx_kernel_time(void *ctx, const char *name, double *ms, int *launches);

/* Work counters of the last wave of the last qlx_eval (no reference counterpart; benchmark
 * accounting).  Sums over the wave's items unless noted:
 *   syrk_dense_flops   2*64^3 per (lower tile, 64-row chunk) pair of a dense SYRK
 *   syrk_exec_flops    MFMA flops executed: 2*64*CB^2 per pair of nonzero 64 x CB pool blocks
 *                      of one 64-row chunk (CB = 64 in FP64, 128 in MIXED; one product each)
 *   syrk_alg_flops     N_img*N_src*(N_src+1) + 2*N_img*N_src (the F and D arithmetic)
 *   lb_bytes           bytes of the blurred lensing matrix blocks written
 *   integrand_evals    NFW integrand evaluations of the ray shooting
 *   chol_flops         N_src^3 / 3 (F's factorization; R's banded one is not counted)
 *   n_items            items of the wave
 *   n_img, n_src, ld_src, rows_pad: sizes of the wave's last item
 *   n_waves            waves the last call was split into (HBM admission)
 *   chol_panel         columns per panel of the MIXED factorization of the last wave (the
 *                      settings' chol_panel, or the autotuner's choice for the wave's size class)
 *   graph_captures, graph_replays  factorization graphs captured / replayed by the context so far */
typedef struct {
    double syrk_dense_flops, syrk_exec_flops, syrk_alg_flops;
    double lb_bytes, integrand_evals, chol_flops;
    int n_items;
    int n_img, n_src, ld_src, rows_pad;
    int n_waves;
    int chol_panel;
    int graph_captures, graph_replays;
} qlx_stats;
int qlx_get_stats(void *ctx, qlx_stats *out);

/* Diagnostic (no reference counterpart): one MFMA on fixed operands, out[lane*4 + r] = result
 * register r of lane.  kind 0: v_mfma_f64_16x16x4f64 with A[i][k] = i + 1 + 0.01k,
 * B[k][j] = 1 + k + 0.1j (register r of lane l = D[l/16 + 4r][l%16] on gfx950); kind 1: the same
 * operands on v_mfma_f32_16x16x4f32 (register r of lane l = D[4(l/16) + r][l%16]); kind 2:
 * v_mfma_f32_16x16x32_bf16 with A[i][k] = (i+1)(k%4+1)/4, B[k][j] = (j+1)((k/4)%8+1)/8 where lane
 * l holds row/column l%16 and k = 8(l/16) .. 8(l/16)+7 (output layout as kind 1); kind 3: the
 * kind 2 operands on v_mfma_f32_16x16x32_f16.  Lets a test confirm the operand and accumulator
 * layouts the tile kernels assume. */
int qlx_selftest_mfma(void *ctx, int kind, double *out256);

#ifdef __cplusplus
}
#endif
#endif
