// Internal declarations shared by the qlx host orchestration (qlx_api.cpp) and the HIP kernel
// translation units.  Not part of the public ABI (include/qlx.h).
//
// Every evaluation kernel works on a batch of items (lens-model proposals): the item index is
// blockIdx.y (blockIdx.y = matrix index for the Cholesky kernels), per-item work arrays sit at
// fixed strides, and per-item sizes / pool offsets come from a device array of ItemMeta filled by
// the host once the active sets are known.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qlx.h"

namespace qlx {

#define QLX_MAX_SRC_AXIS 512
#define QLX_MAX_LD 16384          // padded source dimension limit of the dense F (n_src <= 16383)
#define QLX_MAX_CBW 8             // 32-bit support-mask words per 64-row chunk (QLX_MAX_LD / 64 / 32)
#define QLX_CSC_SEG 512           // row positions per segment of L's column-ordered (CSC) build
__host__ __device__ inline long long csc_segments(long long s_pix) { return (s_pix + QLX_CSC_SEG - 1) / QLX_CSC_SEG; }
// per-item source grid block: sxv[513], syv[513], scx[512], scy[512]
#define QLX_SG_STRIDE 2052
#define QLX_SG_SYV 513
#define QLX_SG_SCX 1026
#define QLX_SG_SCY 1538
// per-item output slots
#define QLX_OUT_STRIDE 10         // -2lnE, chisq, logdetF, logdetR, Es, regterms, refine_resid, refine_steps,
                                  // regparam objective (chi^2 over the active pixels + regterms + ln det F), spare
#define QLX_VEC_SLOTS 8           // per item, ldN doubles each: s (F order), diagF, diagR, Fdiag, resid,
                                  // partial sums, s (reference order), the residual's row-strip sums
// internal source order: active cells ranked in 8 (x) by 16 (y) cell tiles, tiles y-major
#define QLX_TILE_W 8
#define QLX_TILE_H 16

// Per-lens meta-parameters, computed on the host with the host libm exactly as the reference's
// update_meta_parameters does, so the device evaluates only per-ray arithmetic.
// per batch item: QLX_MAX_LENSES lenses x 511 Patterson nodes x 4 NFW node constants
#define QLX_NFWTAB_DOUBLES (QLX_MAX_LENSES * 511 * 4)

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
    double alpha, sple_ff, sple_beta, sple_fac;         // SPLE: slope, (1-q)/(1+q), 2/(2-alpha), 2 b' q/(1+q)
    double sple_balpha, sple_e;                         // SPLE, q = 1: b^alpha, 1 - alpha/2
};

struct ModelPrep {
    int nlens;
    int sur_lens;              // index of the lens whose deflection comes from sur_alpha, or -1
    LensPrep lens[QLX_MAX_LENSES];
};

void prep_model(const qlx_lens_model& in, ModelPrep& out);

// rays per pixel: nsplit^2 subpixels, or one pixel-centre ray when nsplit is 0 (split_imgpixels off)
__host__ __device__ inline int rays_per_pixel(int nsplit) { return nsplit ? nsplit * nsplit : 1; }

// sizes and pool offsets of one item
struct ItemMeta {
    int valid;                       // item takes part in the dense stages
    int n_src, n_img, ldN, nrows;    // active counts; padded source dim (>= n_src+1); padded rows
    int ncb, ncbw, nkt;              // column blocks, mask words per chunk, 64-row chunks
    int snx, sny, bwb, nblk;         // source grid; R band in blocks; Lb support blocks
    int nsup_max, cbs;               // most support blocks of one chunk; log2 of the Lb column-block width
    long long lb_off;                // Lb pool (elements: double, or 2 x nblk x 4096 bf16)
    long long f_off;                 // F (elements of F or F32)
    long long r_off;                 // banded R tiles (doubles)
    long long cvec_off, vec_off, nz_off, bb_off, w_off;
    long long keep_off;              // byte offset of the item's solution block in the keep buffer
};

struct BatchDev {
    int B;
    // image (shared by all items)
    int nx, ny, nsplit, nsub;
    const double* xvals;
    const double* yvals;
    const unsigned char* mask;  // nullptr = all pixels
    const double* data;
    const double* covinv;
    const double* psf;
    int pnx, pny;
    double tol, lambda;
    const double* lam;          // per-item lambda (regparam optimisation), or null: lambda for every item
    const double* Fraw;         // FP64 F as the SYRK formed it (regparam optimisation), or null
    // per-item inputs
    const ModelPrep* models;
    const double* sgrid;        // QLX_SG_STRIDE doubles per item
    const int* src_n;           // 2 ints per item (snx, sny)
    const float* sur_alpha;     // surrogate deflections (2 floats per ray and item), or null
    double* nfwtab;             // NFW quadrature node constants: QLX_NFWTAB_DOUBLES per item
    // per-item work arrays (strides from s_pix, s_cell)
    double* beta;               // s_pix*nsub*2
    int* map_cell;              // s_pix*nsub*3
    double* map_w;              // s_pix*nsub*3
    int* pmapped;               // s_pix
    int* cmapped;               // s_cell
    int* sidx;                  // s_cell: reference (j-major) active index of each cell, or -1
    int* tix;                   // s_cell: internal (tiled) active index of each cell, or -1
    int* pisrc;                 // s_cell: reference index of internal index p
    unsigned long long* rkey;   // 1 per item: hash of the grid shape and active-index map (R's identity)
    int* rrep;                  // 1 per item: the first item of the wave with the same R (itself if none)
    int* pidx;                  // s_pix
    int* act_pix;               // s_pix
    int* counts;                // 4 per item: n_src, n_img, support blocks, largest chunk support
    int* ucol;                  // s_pix*3*nsub
    double* uval;               // s_pix*3*nsub
    int* ucnt;                  // s_pix
    int* apre;                  // s_pix+1: active pixels before pixel t = j*nx + i (image order j-outer)
    int* eptr;                  // s_pix+1: first row-order position of each 8 x 8 pixel tile
    int* csh;                   // csc_segments(s_pix)*s_cell: per segment of QLX_CSC_SEG row positions,
                                // the entries per L column, then the segment's first slot in it
    unsigned* cpix;             // s_pix*3*nsub: L by columns, ordered: pixel (i << 16 | j) of each entry
    double* cval;               // s_pix*3*nsub: L by columns, ordered: value of each entry
    unsigned* pmask;            // s_pix*QLX_MAX_CBW: column blocks touched by each active pixel's L row
    unsigned* hmask;            // s_pix*QLX_MAX_CBW per grid pixel (y-major): pmask ORed over the PSF's x window
    double s16;                 // mixed path: power of two that scales the FP16 pool entries (|x s16| <= 2^14)
    int* key;                   // s_pix: row-order position of each active row
    int* bcnt;                  // s_cell+2: entries per L column
    int* boff;                  // s_cell+2: first entry of each L column (n_src + 1 bounds)
    int* bfill;                 // s_cell+2: scatter cursors
    int* perm;                  // s_pix: active row at each row-order position
    unsigned long long* nevals; // 1 per item
    long long s_pix, s_cell;
    // pools and outputs
    const ItemMeta* meta;
    void* lb;                   // blurred L support blocks (double, or bf16 hi/lo when mixed)
    double* F;                  // FP64 F (parity mode)
    float* F32;                 // FP32 F (mixed mode), factored in place
    float* Ff32;                // FP32 F as formed by the mixed SYRK (lower 128-tiles), kept for the
                                // refinement residual
    double* R;                  // banded R tiles
    double* cvec;
    double* vec;                // QLX_VEC_SLOTS * ldN doubles per item
    double* zres;               // mixed: the refinement residual's transposed tile products, ldN^2 / 128
                                // doubles per item (at f_off / 128)
    unsigned* nz;               // nkt * ncbw support words per item
    int* bbase;                 // nkt + 1 slot bases per item
    double* winv;               // per item 2 * ldN*64 doubles (F then R block inverses)
    double* model;              // per item 2*s_pix
    double* out;                // QLX_OUT_STRIDE per item
    int* status;                // per item 4
};

// one matrix of a batched Cholesky (lower, row-major tiles of 64x64)
//   bwb == 0: dense, tile (I,J) at A + I*64*ld + J*64, row stride ld
//   bwb  > 0: banded, tile (I,J) (I-J <= bwb) at A + (I*(bwb+1) + I-J)*4096, row stride 64
struct MatDesc {
    void* A;
    int ld, n, nb, f32, bwb, pad;
    void* diag;                 // nb*64 pivots (element type of A)
    int* status;                // bit 0 set on a negative pivot of a row < n
    void* W;                    // nb block inverses (64x64 each, element type of A)
    void* A0;                   // dense FP32: F as formed; the kernels that touch a block first (the
                                // first part's diagonal and panel kernels, the first strip and
                                // trailing updates) read it here and write A (null: in place)
};

void upload_patterson(int device);

// ---- raytrace.hip ----
void launch_raytrace_map_b(hipStream_t s, const BatchDev& bd, int ray_fp32, int kappa_table, int map);
void launch_map_rays_b(hipStream_t s, const BatchDev& bd);
// automatic source grid: crays (2 s_pix doubles per item) receive the pixel-centre rays when
// count_centres; res (8 doubles per item) receives xmin, xmax, ymin, ymax, centre count; box (host
// memory: xmin, xmax, ymin, ymax) is the given grid box, used when auto_box is 0
void launch_srcgrid_auto_b(hipStream_t s, const BatchDev& bd, int auto_box, int count_centres, double nt, const double* box,
                           double* crays, double* res);
void launch_raytrace_points(hipStream_t stream, const ModelPrep& m, double tol, int n, const double* xy, double* beta);
// the file-scope staging array of the allocation-free ray entry point (raytrace.hip): its address on
// `device` (which must be current), taken once per device; 0 on success
#define QLX_RT_SCRATCH_POINTS (1 << 18)
int raytrace_scratch(int device, double** out);
void launch_nfw_def_hess(hipStream_t stream, const LensPrep& L, double tol, int n, const double* xy, double* out5);
void launch_mock_sb(hipStream_t stream, int nx, int ny, int nsplit, const double* xvals, const double* yvals, const ModelPrep& m,
                    double tol, int nsrc, const double* sersic, double* sb);
void launch_conv_image(hipStream_t stream, int nx, int ny, const double* psf, int pnx, int pny, const double* in, double* out);

// ---- surrogate.hip (learned NFW deflection; host driver in surrogate_host.cpp) ----
struct SurParams {            // device state of one network
    float* P;                 // FP32 master parameters (QLX_SUR_NPARAM)
    float* m;                 // Adam moments
    float* v;
    uint16_t* Wb;             // BF16 W2, W3, W4 ([out][in], 128x128 each)
    uint16_t* WTb;            // their transposes
    float* W1r;               // BF16-rounded W1 [128][4] and W5 [5][128] as floats
    float* W5r;
    float* norm;              // output standardization: mean[5], std[5]
    int* step;                // Adam step counter
};
struct SurWork {              // per-minibatch buffers of the training kernels
    float* zs;                // pre-activations of the 4 hidden layers, [n][4*128]
    uint16_t* hs;             // BF16 activations of the 4 hidden layers, [n][4*128]
    uint16_t* dzA;            // BF16 dz of a hidden layer, [n][128] (two buffers)
    uint16_t* dzB;
    float* dz1;               // FP32 dz of the first hidden layer, [n][128]
    float* loss_part;         // per workgroup
    float* g5_part;           // per workgroup: W5, b5 gradients (645)
    float* gW_part;           // [3][nwg][128*128]
    float* gb_part;           // [3][nwg][128]
    float* g1_part;           // per workgroup: W1 (512), b1 (128)
    float* grad;              // reduced gradient (QLX_SUR_NPARAM)
    float* loss;              // reduced loss
};
struct SurDev {               // what the inference and training kernels read
    const float* P;
    const uint16_t* Wb;
    const uint16_t* WTb;
    const float* W1r;
    const float* W5r;
    const float* norm;
};
#define QLX_SUR_NPARAM 50821
#define QLX_SUR_SB 64         // samples per workgroup of the surrogate kernels
SurDev sur_dev(const SurParams& sp);
void launch_sur_eval(hipStream_t s, const SurDev& net, int n, const float* in4, float* out5);
void launch_sur_rays(hipStream_t s, const SurDev& net, const BatchDev& bd, float* sur_alpha);
void launch_sur_gather(hipStream_t s, int n, int nsamp, unsigned long long seed, const int* step, const float* feat, const float* targ,
                       float* bf, float* bt);
void launch_sur_fwdbwd(hipStream_t s, const SurDev& net, int n, const float* bf, const float* bt, SurWork& w);
void launch_sur_adam(hipStream_t s, SurParams& sp, const float* grad, int total_steps, float lr0, float lr_min);
void launch_sur_refresh(hipStream_t s, SurParams& sp);
void launch_sur_targets(hipStream_t s, int n, const double* in4, const double* j5, float* feat, float* targ);
void launch_nfw_samples(hipStream_t stream, int n, const double* in4, double tol, double* out5);

// context services used by the surrogate driver (qlx_api.cpp)
hipStream_t ctx_stream(void* ctx);
int ctx_fail(void* ctx, int code, const char* msg);
void* ctx_surrogate(void* ctx);                  // the context's surrogate state (created on demand)
const SurDev* ctx_surrogate_net(void* ctx);      // trained network, or nullptr
void* sur_state_new();
void sur_state_destroy(void* state);
const SurDev* sur_state_net(void* state);

// ---- batch.hip ----
void launch_fill_i32(hipStream_t s, int* p, int v, size_t n);
void launch_fill_f64(hipStream_t s, double* p, double v, size_t n);
void launch_compact_b(hipStream_t s, const BatchDev& bd);
void launch_merge_lrows_b(hipStream_t s, const BatchDev& bd, int max_img);
void launch_sort_rows_b(hipStream_t s, const BatchDev& bd, int max_img, int max_src);
void launch_lcols_b(hipStream_t s, const BatchDev& bd, int max_img, int max_src);
void launch_row_support_b(hipStream_t s, const BatchDev& bd, int max_nkt);
int launch_build_lb_b(hipStream_t s, const BatchDev& bd, int max_rows, int max_nsup, int mixed, int cbw);
void launch_gather_cvec_b(hipStream_t s, const BatchDev& bd, int max_rows);
void launch_syrk_f64_b(hipStream_t s, const BatchDev& bd, int max_ncb, int max_nkt);
void launch_syrk_f16_b(hipStream_t s, const BatchDev& bd, int max_ncb, int max_nkt);
void launch_curvature_b(hipStream_t s, const BatchDev& bd, int max_cells, int max_rtiles);
// items whose R equals an earlier item's (same grid shape and active-index map) get rrep = that item,
// others rrep = themselves: one build and one factorization of R serve them all (the host reads rrep
// back and gives R's factorization the representatives' descriptors only; launch_curvature_b
// builds the representatives' R only)
void launch_rdup_b(hipStream_t s, const BatchDev& bd);
void launch_prep_fr_b(hipStream_t s, const BatchDev& bd, int max_ld, int f32);
// panel_blocks: width of the FP32 dense schedule's panels in 64-blocks (a multiple of 4; 0 = 16);
// the factor is bitwise the same for every width
void cholesky_batched(hipStream_t s, const MatDesc* d_mats, int n32, int n64, int max_nb, int dense, int max_bwb,
                      int panel_blocks = 0);
void launch_logdets_b(hipStream_t s, const BatchDev& bd, int f32);
void launch_trisolve_b(hipStream_t s, const BatchDev& bd, int max_ld, int f32);
void launch_refine_b(hipStream_t s, const BatchDev& bd, int max_ld, int iters, double tol);
void launch_evidence_b(hipStream_t s, const BatchDev& bd, int max_img, int max_src);
// copy each valid item's source amplitudes (reference order), model image and cell map into its
// block of `keep` at meta.keep_off: [cells doubles] [s_pix doubles] [cells ints]
void launch_keep_b(hipStream_t s, const BatchDev& bd, void* keep, int max_cells);
void launch_selftest_mfma(hipStream_t s, double* out);

}  // namespace qlx
