// Internal declarations shared by the qlx host orchestration (qlx_api.cpp) and the HIP kernel
// translation units.  Not part of the public ABI (include/qlx.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qlx.h"

namespace qlx {

// Per-lens meta-parameters, computed on the host with the host libm exactly as the reference's
// update_meta_parameters does, so the device evaluates only per-ray arithmetic.
struct LensPrep {
    int type;
    int def_mode;
    double xc, yc;
    double costh, sinth;       // after the orient-north swap (profile.cpp:1625-1637)
    int rotate;                // sintheta != 0 (profile.cpp:1791)
    double q, qsq, bprime, sprime, ssq;                 // SIE
    double sie_u, sie_fac;                              // SIE: sqrt(1-q^2), bprime*q/u
    double ks, rs, rs2, eps, fsqinv, sqrt1meps;         // NFW
    double shear1, shear2;                              // shear
};

struct ModelPrep {
    int nlens;
    LensPrep lens[QLX_MAX_LENSES];
};

void prep_model(const qlx_lens_model& in, ModelPrep& out);

// ---- image / source grid geometry (device pointers) ----
struct GridDev {
    int nx, ny, nsplit, nsub;
    const double* xvals;   // nx+1
    const double* yvals;   // ny+1
    const unsigned char* mask;  // nx*ny
    int snx, sny;
    const double* sxv;     // snx+1 accumulated source corners
    const double* syv;     // sny+1
    const double* scx;     // snx centres
    const double* scy;     // sny centres
};

// per-item work buffers (device)
struct ItemBufs {
    double* beta;        // nx*ny*nsub*2
    int* map_cell;       // nx*ny*nsub*3
    double* map_w;       // nx*ny*nsub*3
    int* pmapped;        // nx*ny
    int* cmapped;        // snx*sny
    int* sidx;           // snx*sny active index or -1
    int* pidx;           // nx*ny active row or -1
    int* act_pix;        // active rows -> pixel id
    int* counts;         // [0] n_src, [1] n_img, [2] status flags
    int* ucol;           // n_img*12 merged L columns
    double* uval;        // n_img*12
    int* ucnt;           // n_img
    long long* nevals;   // integrand evaluation counter
    int* key;            // n_img: source cell (active index) of each active pixel's first ray
    int* bucket_cnt;     // n_src+1
    int* bucket_off;     // n_src+1
    int* bucket_fill;    // n_src+1
    int* perm;           // n_img: sorted row -> active row
};

// ---- kernel launchers (all asynchronous on `stream`) ----
void launch_raytrace_map(hipStream_t stream, const GridDev& g, const ModelPrep& m, double tol, int ray_fp32,
                         const ItemBufs& b, const void* surrogate);
void launch_fill_i32(hipStream_t stream, int* p, int v, size_t n);
void launch_fill_f64(hipStream_t stream, double* p, double v, size_t n);
void launch_compact(hipStream_t stream, const GridDev& g, const ItemBufs& b);
void launch_merge_lrows(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_img);
void launch_build_lb(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_img, int n_src,
                     const double* psf, int pnx, int pny, const double* data, double* Lb, int ldL,
                     unsigned int* nzmask, int ncb);
void launch_build_lb_bf16(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_img, int n_src,
                          const double* psf, int pnx, int pny, const double* data, const double* covinv,
                          uint16_t* Lb16, int ldL, unsigned int* nzmask, int ncb, float* dcol);
void launch_syrk_f64(hipStream_t stream, const double* Lb, int ldL, const double* cvec, int nrows, int ncols,
                     double* F, int ldF, const unsigned int* nzmask, int ncb);
void launch_syrk_bf16(hipStream_t stream, const uint16_t* Lb16, int ldL, int nrows, int ncols, float* F32,
                      int ldF, const unsigned int* nzmask, int ncb);
void launch_gather_cvec(hipStream_t stream, const ItemBufs& b, const double* covinv, const double* data,
                        int n_img, double* cvec, double* dvec);
void launch_add_curvature(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_src, double lambda,
                          double* F, int ldF, double* R, int ldR);
void launch_add_curvature_f32(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_src, double lambda,
                              float* F, int ldF);
// Cholesky (lower, in place, row-major ld) of an n x n matrix padded to a multiple of 64.
int cholesky_f64(hipStream_t stream, double* A, int ld, int n, double* diag, int* status, double* work);
int cholesky_f32(hipStream_t stream, float* A, int ld, int n, float* diag, int* status, float* work);
void launch_trisolve_f64(hipStream_t stream, const double* L, int ld, int n, const double* b, double* x);
void launch_trisolve_f32(hipStream_t stream, const float* L, int ld, int n, const float* b, float* x);
void launch_logdet(hipStream_t stream, const double* diag, int n, double* out);
void launch_logdet_f32diag(hipStream_t stream, const float* diag, int n, double* out);
// model image, chi-square, regularization terms, evidence
void launch_evidence(hipStream_t stream, const GridDev& g, const ItemBufs& b, int n_img, int n_src,
                     const double* s, const double* psf, int pnx, int pny, const double* data,
                     const double* covinv, double lambda, const double* logdets, double* model, double* out);
// mixed precision: r = D - (F s) with F = symmetric lower (FP64 F built from FP32 pieces)
void launch_symv_resid(hipStream_t stream, const double* F, int ldF, int n, const double* s, const double* D,
                       double* r);
void launch_symv_resid_f32F(hipStream_t stream, const float* F, int ldF, int n, const double* s,
                            const double* D, double* r);
void launch_axpy_f64(hipStream_t stream, int n, const double* dx, double* x);
void launch_norm2_pair(hipStream_t stream, int n, const double* a, const double* b, double* out2);
void launch_f64_to_f32(hipStream_t stream, const double* a, float* b, int n);
void launch_f32_to_f64(hipStream_t stream, const float* a, double* b, int n);
void launch_extract_col_f32(hipStream_t stream, const float* F, int ldF, int col, int n, double* out);
void launch_extract_col_f64(hipStream_t stream, const double* F, int ldF, int col, int n, double* out);

// mock image
void launch_mock_sb(hipStream_t stream, const GridDev& g, const ModelPrep& m, double tol, int nsrc,
                    const double* sersic, double* sb);
void launch_conv_image(hipStream_t stream, int nx, int ny, const double* psf, int pnx, int pny,
                       const double* in, double* out);
void launch_raytrace_points(hipStream_t stream, const ModelPrep& m, double tol, int n, const double* xy,
                            double* beta);
void launch_nfw_def_hess(hipStream_t stream, const LensPrep& L, double tol, int n, const double* xy,
                         double* out5);

// block-sparse blurred L and SYRK (blocksparse.hip)
void launch_sort_rows(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int n_src);
void launch_row_support(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int n_src, int pnx, int pny, unsigned* nzmask,
                        int ncbw);
void launch_build_lb_blocks_f64(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int nrows, int n_src, const double* psf,
                                int pnx, int pny, const double* data, double* pool, int ncb, const unsigned* nzmask, int ncbw);
void launch_build_lb_blocks_bf16(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int nrows, int n_src, const double* psf,
                                 int pnx, int pny, const double* data, const double* covinv, uint16_t* pool, int ncb,
                                 const unsigned* nzmask, int ncbw);
void launch_gather_cvec_sorted(hipStream_t s, const ItemBufs& b, const double* covinv, int n_img, int nrows, double* cvec);
void launch_syrk_blocks_f64(hipStream_t s, const double* pool, int ncb, int nkt, const double* cvec, double* F, int ldF,
                            const unsigned* nz, int ncbw);

// surrogate (MLP) kernels
struct MlpDev;
int surrogate_kernels_forward(hipStream_t stream, const MlpDev* mlp, int n, const float* x, float* y);

}  // namespace qlx
