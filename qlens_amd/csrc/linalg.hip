// Dense inversion kernels for gfx950 (FP64 path): active-set compaction, merged lensing-matrix
// rows, PSF-blurred dense Lb, F = Lb^T C Lb on FP64 MFMA, curvature regularization, blocked
// Cholesky + triangular solves, and the evidence terms.
//
// Layouts: Lb is row-major [rows = active image pixels (padded to 64)] x [ldL columns]; column
// n_src holds the data vector so one SYRK also yields D = Lb^T C d (row n_src of F) and d^T C d.
// F and R are row-major ldF x ldF, lower triangle used; padding rows/cols carry an identity.
//
// Reference behaviour followed:
//   active sets              assign_active_indices / assign_pixel_mappings  pixelgrid.cpp:3586-3626, 12879-12914
//   dense L (merged in CSR)  assign_Lmatrix + convert_Lmatrix_to_dense      pixelgrid.cpp:13196-13321, 13717-13747
//   PSF on L                 PSF_convolution_Lmatrix_dense (direct)         pixelgrid.cpp:14092-14176
//   D, F                     create_lensing_matrices_from_Lmatrix_dense     pixelgrid.cpp:15590-15711
//   curvature R              generate_hmatrices / Rmatrix_from_hmatrices    pixelgrid.cpp:2968-3240
//   F += lambda R            add_regularization_term_to_dense_Fmatrix      pixelgrid.cpp:15808-15818
//   Cholesky, solve, logdet  Cholesky_dcmp/solve/logdet_lower_packed        pixelgrid.cpp:17071-17246
//   model, chi^2, Es         lens.cpp:12950-13015, pixelgrid.cpp:15821-15863, 18268-18307
#include "qlx_internal.h"

namespace qlx {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------------
// small utilities
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int v, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}
__global__ void k_fill_f64(double* p, double v, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}
static inline unsigned grid_for(size_t n, unsigned bs = 256, unsigned cap = 65535u * 4u)
{
    size_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}
// gfx950 has 160 KB of LDS per CU; allow kernels with dynamic LDS to use it (set once per kernel).
template <typename K>
static void allow_big_lds(K kernel)
{
    static bool done = false;
    if (!done) {
        hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 48 * 1024);
        done = true;
    }
}

void launch_fill_i32(hipStream_t s, int* p, int v, size_t n)
{
    if (n) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}
void launch_fill_f64(hipStream_t s, double* p, double v, size_t n)
{
    if (n) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}

// wave-wide inclusive prefix sum of an int (64 lanes)
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block (1024 threads) exclusive scan of flags over a logical sequence; returns the running base
// through `base` and writes the exclusive index of the current element.
__device__ __forceinline__ int block_excl_scan(int flag, int* sh_wave, int& total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = wave_incl_scan(flag);
    if (lane == 63) sh_wave[w] = incl;
    __syncthreads();
    if (w == 0) {
        int v = (lane < 16) ? sh_wave[lane] : 0;
        int s = wave_incl_scan(v);
        if (lane < 16) sh_wave[16 + lane] = s - v;  // exclusive wave offsets
        if (lane == 15) sh_wave[32] = s;
    }
    __syncthreads();
    int excl = sh_wave[16 + w] + incl - flag;
    total = sh_wave[32];
    __syncthreads();
    return excl;
}

// Active source cells in j-major order and active image pixels in j-outer/i-inner order.
__global__ void __launch_bounds__(1024) k_compact(GridDev g, ItemBufs b)
{
    __shared__ int sh[40];
    int base = 0;
    const int ncell = g.snx * g.sny;
    for (int t0 = 0; t0 < ncell; t0 += 1024) {
        int t = t0 + threadIdx.x;
        int c = -1, flag = 0;
        if (t < ncell) {
            int i = t % g.snx, j = t / g.snx;
            c = i * g.sny + j;
            flag = b.cmapped[c] ? 1 : 0;
        }
        int tot;
        int ex = block_excl_scan(flag, sh, tot);
        if (c >= 0) b.sidx[c] = flag ? base + ex : -1;
        base += tot;
    }
    if (threadIdx.x == 0) b.counts[0] = base;
    base = 0;
    const int npix = g.nx * g.ny;
    for (int t0 = 0; t0 < npix; t0 += 1024) {
        int t = t0 + threadIdx.x;
        int p = -1, flag = 0;
        if (t < npix) {
            int i = t % g.nx, j = t / g.nx;
            p = i * g.ny + j;
            flag = ((!g.mask || g.mask[p]) && b.pmapped[p]) ? 1 : 0;
        }
        int tot;
        int ex = block_excl_scan(flag, sh, tot);
        if (p >= 0) {
            b.pidx[p] = flag ? base + ex : -1;
            if (flag) b.act_pix[base + ex] = p;
        }
        base += tot;
    }
    if (threadIdx.x == 0) b.counts[1] = base;
}

void launch_compact(hipStream_t s, const GridDev& g, const ItemBufs& b)
{
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, s, g, b);
}

// Merged row of L for each active pixel: entries in CSR order (subpixel, then cell/x-nbr/y-nbr),
// duplicate columns summed in that order starting from 0, as the dense conversion does.
__global__ void __launch_bounds__(256) k_merge_lrows(GridDev g, ItemBufs b, int n_img)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_img) return;
    const int U = 3 * g.nsub;
    const long long p = b.act_pix[r];
    int* uc = b.ucol + (size_t)r * U;
    double* uv = b.uval + (size_t)r * U;
    int cnt = 0;
    for (int k = 0; k < g.nsub; k++) {
        const int* mc = b.map_cell + (p * g.nsub + k) * 3;
        const double* mw = b.map_w + (p * g.nsub + k) * 3;
        if (mc[0] < 0) continue;
        for (int e = 0; e < 3; e++) {
            int col = b.sidx[mc[e]];
            int q = 0;
            while (q < cnt && uc[q] != col) q++;
            if (q == cnt) { uc[cnt] = col; uv[cnt] = 0.0 + mw[e]; cnt++; }
            else uv[q] += mw[e];
        }
    }
    b.ucnt[r] = cnt;
}

void launch_merge_lrows(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img)
{
    if (n_img > 0) hipLaunchKernelGGL(k_merge_lrows, dim3((n_img + 255) / 256), dim3(256), 0, s, g, b, n_img);
}

// One wavefront per active pixel builds its blurred row Lb[r][:] in LDS: taps in the reference's
// order (psf_k outer, psf_l inner), each adding psf * (merged L row of the tap pixel); within a
// tap every merged column is distinct, so lanes never collide.  Column n_src receives the data
// value.  The row is then streamed out and the 64x64 tile-occupancy bitmap updated.
__global__ void __launch_bounds__(64) k_build_lb(GridDev g, ItemBufs b, int n_img, int n_src, const double* psf, int pnx, int pny,
                                                 const double* data, double* Lb, int ldL, unsigned int* nzmask, int ncbw)
{
    extern __shared__ double row[];
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    for (int c = lane; c < ldL; c += 64) row[c] = 0.0;
    __syncthreads();
    const int U = 3 * g.nsub;
    const int p = b.act_pix[r];
    const int k = p / g.ny, l = p % g.ny;
    const int nxh = pnx / 2, nyh = pny / 2;
    for (int pk = 0; pk < pnx; pk++) {
        const int ii = k + nxh - pk;
        if (ii < 0 || ii >= g.nx) continue;
        for (int pl = 0; pl < pny; pl++) {
            const int jj = l + nyh - pl;
            if (jj < 0 || jj >= g.ny) continue;
            const int r2 = b.pidx[ii * g.ny + jj];
            if (r2 < 0) continue;
            const double pv = psf[pk * pny + pl];
            const int cnt = b.ucnt[r2];
            if (lane < cnt) {
                const int c = b.ucol[(size_t)r2 * U + lane];
                row[c] += pv * b.uval[(size_t)r2 * U + lane];
            }
            __syncthreads();
        }
    }
    if (lane == 0) row[n_src] = data[p];
    __syncthreads();
    double* out = Lb + (size_t)r * ldL;
    for (int c = lane; c < ldL; c += 64) out[c] = row[c];
    // occupancy of 64-column blocks for this row's 64-row chunk
    const int ncb = ldL / 64;
    for (int cb = lane; cb < ncb; cb += 64) {
        bool nz = false;
        for (int c = cb * 64; c < cb * 64 + 64; c++) nz |= (row[c] != 0.0);
        if (nz) atomicOr(&nzmask[(size_t)(r / 64) * ncbw + (cb >> 5)], 1u << (cb & 31));
    }
}

void launch_build_lb(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int n_src, const double* psf, int pnx, int pny,
                     const double* data, double* Lb, int ldL, unsigned int* nzmask, int ncbw)
{
    if (n_img <= 0) return;
    size_t lds = (size_t)ldL * sizeof(double);
    allow_big_lds(k_build_lb);
    hipLaunchKernelGGL(k_build_lb, dim3(n_img), dim3(64), lds, s, g, b, n_img, n_src, psf, pnx, pny, data, Lb, ldL, nzmask, ncbw);
}

// covinv of each active row (padding rows get 0)
__global__ void k_gather_cvec(ItemBufs b, const double* covinv, int n_img, int nrows_pad, double* cvec)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows_pad) return;
    cvec[r] = (r < n_img) ? covinv[b.act_pix[r]] : 0.0;
}

void launch_gather_cvec(hipStream_t s, const ItemBufs& b, const double* covinv, const double* /*data*/, int n_img, double* cvec,
                        double* /*dvec*/)
{
    int pad = (n_img + 63) / 64 * 64;
    if (pad == 0) pad = 64;
    hipLaunchKernelGGL(k_gather_cvec, dim3((pad + 255) / 256), dim3(256), 0, s, b, covinv, n_img, pad, cvec);
}

// ------------------------------------------------------------------------------------------------
// FP64 MFMA tile engine: acc(64x64) = sum_k A(k, i) * B(k, j) [* cvec[k]]
// A(k,i) = A[k*sAk + i*sAi], B(k,j) = B[k*sBk + j*sBj].  256 threads = 4 waves, each a 32x32
// quadrant as 2x2 v_mfma_f64_16x16x4f64 tiles.  K is processed in 64-wide chunks through LDS.
// ------------------------------------------------------------------------------------------------
#define TS 64
#define LDSP 66  // padded LDS row (doubles) to spread banks

__device__ __forceinline__ void load_tile(double (*S)[LDSP], const double* P, long long sk, long long si, int k0, const double* cvec)
{
    const int tid = threadIdx.x;
    if (si == 1) {
#pragma unroll 4
        for (int e = tid; e < TS * TS; e += 256) {
            int kk = e >> 6, ii = e & 63;
            double v = P[(long long)(k0 + kk) * sk + ii];
            if (cvec) v *= cvec[k0 + kk];
            S[kk][ii] = v;
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < TS * TS; e += 256) {
            int ii = e >> 6, kk = e & 63;
            double v = P[(long long)(k0 + kk) * sk + (long long)ii * si];
            if (cvec) v *= cvec[k0 + kk];
            S[kk][ii] = v;
        }
    }
}

struct TileJob {
    const double* A;  // already offset to this tile's first i
    long long sAk, sAi;
    const double* B;  // already offset to this tile's first j
    long long sBk, sBj;
    const double* cvec;
    int K;
};

__device__ __forceinline__ void tile_mma(const TileJob& jb, double (*As)[LDSP], double (*Bs)[LDSP], double4_t acc[2][2],
                                         const unsigned int* nz, int ncbw, int I, int J)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (double4_t){0, 0, 0, 0};
    for (int k0 = 0; k0 < jb.K; k0 += TS) {
        if (nz) {
            const unsigned* row = nz + (size_t)(k0 / TS) * ncbw;
            bool hasI = (row[I >> 5] >> (I & 31)) & 1u;
            bool hasJ = (row[J >> 5] >> (J & 31)) & 1u;
            if (!(hasI && hasJ)) continue;  // uniform across the workgroup
        }
        __syncthreads();
        load_tile(As, jb.A, jb.sAk, jb.sAi, k0, nullptr);
        load_tile(Bs, jb.B, jb.sBk, jb.sBj, k0, jb.cvec);
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < TS; kk += 4) {
            const int kq = kk + (lane >> 4);
            double a0 = As[kq][wr * 32 + (lane & 15)];
            double a1 = As[kq][wr * 32 + 16 + (lane & 15)];
            double b0 = Bs[kq][wc * 32 + (lane & 15)];
            double b1 = Bs[kq][wc * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
}

// v_mfma_f64_16x16x4f64 accumulator layout: register v of lane l holds D[l/16 + 4v][l%16]
// (measured on gfx950 with qlx_selftest_mfma; tests/test_gpu_parity.py keeps checking it).
__device__ __forceinline__ void tile_store(double* C, long long ldC, const double4_t acc[2][2], int mode)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int rr = wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                int cc = wc * 32 + c * 16 + (lane & 15);
                double* dst = C + (long long)rr * ldC + cc;
                if (mode == 0) *dst = acc[a][c][v];
                else *dst -= acc[a][c][v];
            }
}

__device__ __forceinline__ void tri_decode(int t, int& I, int& J)
{
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    while (i * (i + 1) / 2 > t) i--;
    I = i;
    J = t - i * (i + 1) / 2;
}

// F[I-block][J-block] = sum_r Lb[r][I cols] * c_r * Lb[r][J cols] over lower-triangular tiles
__global__ void __launch_bounds__(256) k_syrk_f64(const double* Lb, int ldL, const double* cvec, int K, double* F, int ldF,
                                                  const unsigned int* nz, int ncbw)
{
    __shared__ double As[TS][LDSP];
    __shared__ double Bs[TS][LDSP];
    int I, J;
    tri_decode(blockIdx.x, I, J);
    TileJob jb;
    jb.A = Lb + I * TS;
    jb.sAk = ldL;
    jb.sAi = 1;
    jb.B = Lb + J * TS;
    jb.sBk = ldL;
    jb.sBj = 1;
    jb.cvec = cvec;
    jb.K = K;
    double4_t acc[2][2];
    tile_mma(jb, As, Bs, acc, nz, ncbw, I, J);
    tile_store(F + (long long)I * TS * ldF + J * TS, ldF, acc, 0);
}

void launch_syrk_f64(hipStream_t s, const double* Lb, int ldL, const double* cvec, int nrows, int ncols, double* F, int ldF,
                     const unsigned int* nzmask, int ncbw)
{
    int nb = ncols / TS;
    int ntiles = nb * (nb + 1) / 2;
    int K = (nrows + TS - 1) / TS * TS;
    hipLaunchKernelGGL(k_syrk_f64, dim3(ntiles), dim3(256), 0, s, Lb, ldL, cvec, K, F, ldF, nzmask, ncbw);
}

// ------------------------------------------------------------------------------------------------
// curvature regularization (exact small integers) and F preparation
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_curvature(GridDev g, ItemBufs b, double* R, int ldR)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.snx * g.sny) return;
    const int a = b.sidx[c];
    if (a < 0) return;
    const int i = c / g.sny, j = c % g.sny;
    for (int l = 0; l < 2; l++) {
        int ids[3];
        double h[3];
        int m = 0;
        ids[m] = a;
        h[m++] = -2.0;
        int nb0, nb1;
        if (l == 0) {
            nb0 = (i < g.snx - 1) ? b.sidx[(i + 1) * g.sny + j] : -1;
            nb1 = (i > 0) ? b.sidx[(i - 1) * g.sny + j] : -1;
        } else {
            nb0 = (j < g.sny - 1) ? b.sidx[i * g.sny + j + 1] : -1;
            nb1 = (j > 0) ? b.sidx[i * g.sny + j - 1] : -1;
        }
        if (nb0 >= 0) { ids[m] = nb0; h[m++] = 1.0; }
        if (nb1 >= 0) { ids[m] = nb1; h[m++] = 1.0; }
        for (int x = 0; x < m; x++)
            for (int y = x; y < m; y++) {
                int r0 = ids[x] > ids[y] ? ids[x] : ids[y];
                int c0 = ids[x] > ids[y] ? ids[y] : ids[x];
                double e = h[x] * h[y];
                if (x != y && ids[x] == ids[y]) e *= 2.0;  // cannot occur on a uniform grid
                atomicAdd(&R[(long long)r0 * ldR + c0], e);
            }
    }
}

// F_lower += lambda*R for the n x n source block.  Row n of F keeps the data row (D^T and d^T C d
// from the SYRK) so the Cholesky of the augmented matrix [[F, D], [D^T, d^T C d]] delivers
// y = L^-1 D as its last row; rows/cols > n (and >= n in R) become identity padding, which leaves
// the log-determinants of the leading n x n blocks unchanged.
__global__ void __launch_bounds__(256) k_prep_fr(double* F, double* R, int ld, int n, double lambda)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)ld * ld;
    for (; t < total; t += (long long)gridDim.x * blockDim.x) {
        int r = (int)(t / ld), c = (int)(t % ld);
        if (c > r) continue;
        if (r < n) {
            F[t] += lambda * R[t];
        } else {
            double v = (r == c) ? 1.0 : 0.0;
            if (r > n) F[t] = v;
            R[t] = v;
        }
    }
}

void launch_add_curvature(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_src, double lambda, double* F, int ldF, double* R,
                          int ldR)
{
    // R must be zero on entry; D is written to R's padding-free scratch provided by the caller
    hipLaunchKernelGGL(k_curvature, dim3((g.snx * g.sny + 255) / 256), dim3(256), 0, s, g, b, R, ldR);
    (void)n_src;
    (void)lambda;
    (void)F;
    (void)ldF;
}

void launch_prep_fr(hipStream_t s, double* F, double* R, int ld, int n, double lambda)
{
    size_t total = (size_t)ld * ld;
    hipLaunchKernelGGL(k_prep_fr, dim3(grid_for(total, 256, 8192)), dim3(256), 0, s, F, R, ld, n, lambda);
}

// ------------------------------------------------------------------------------------------------
// Blocked Cholesky (lower, row-major), 64x64 blocks: diagonal factor + inverse, panel solve by
// multiplication with the inverse (MFMA), trailing update (MFMA).
// ------------------------------------------------------------------------------------------------
// A[ib][kb] <- A[ib][kb] * Winv^T for ib > kb (one workgroup per row block), Winv = L[kb][kb]^-1
__global__ void __launch_bounds__(256) k_chol_panel(double* A, int ld, int kb, const double* Winv)
{
    __shared__ double As[TS][LDSP];
    __shared__ double Bs[TS][LDSP];
    const int ib = kb + 1 + blockIdx.x;
    TileJob jb;
    jb.A = A + (long long)ib * TS * ld + (long long)kb * TS;  // A(k=q, i=r) = A[(ib*64+r)*ld + kb*64 + q]
    jb.sAk = 1;
    jb.sAi = ld;
    jb.B = Winv + (size_t)kb * TS * TS;  // B(k=q, j=c) = Winv[c][q]
    jb.sBk = 1;
    jb.sBj = TS;
    jb.cvec = nullptr;
    jb.K = TS;
    double4_t acc[2][2];
    tile_mma(jb, As, Bs, acc, nullptr, 0, 0, 0);
    __syncthreads();
    tile_store(A + (long long)ib * TS * ld + (long long)kb * TS, ld, acc, 0);
}

// trailing update A[I][J] -= P[I] P[J]^T over lower tiles I >= J > kb
__global__ void __launch_bounds__(256) k_chol_update(double* A, int ld, int kb)
{
    __shared__ double As[TS][LDSP];
    __shared__ double Bs[TS][LDSP];
    int I, J;
    tri_decode(blockIdx.x, I, J);
    I += kb + 1;
    J += kb + 1;
    TileJob jb;
    jb.A = A + (long long)I * TS * ld + (long long)kb * TS;
    jb.sAk = 1;
    jb.sAi = ld;
    jb.B = A + (long long)J * TS * ld + (long long)kb * TS;
    jb.sBk = 1;
    jb.sBj = ld;
    jb.cvec = nullptr;
    jb.K = TS;
    double4_t acc[2][2];
    tile_mma(jb, As, Bs, acc, nullptr, 0, 0, 0);
    tile_store(A + (long long)I * TS * ld + (long long)J * TS, ld, acc, 1);
}

void launch_chol_diag(hipStream_t s, double* A, int ld, int kb, double* diag, int* status, double* Winv, int nflag);

int cholesky_f64(hipStream_t s, double* A, int ld, int n, double* diag, int* status, double* work)
{
    const int nb = ld / TS;
    for (int kb = 0; kb < nb; kb++) {
        const int rest = nb - kb - 1;
        launch_chol_diag(s, A, ld, kb, diag, status, work, n);
        if (rest > 0) {
            hipLaunchKernelGGL(k_chol_panel, dim3(rest), dim3(256), 0, s, A, ld, kb, work);
            hipLaunchKernelGGL(k_chol_update, dim3(rest * (rest + 1) / 2), dim3(256), 0, s, A, ld, kb);
        }
    }
    return 0;
}

// Forward + backward substitution with the stored block inverses (single workgroup).
// Forward: rhs_k = b_k - L[k, <k] y_<k (16 threads per row read consecutive columns), then
// y_k = Winv_k rhs_k.  Backward: rhs_k = y_k - L[>k, k]^T x_>k (threads over columns), then
// x_k = Winv_k^T rhs_k.  Each block inverse is staged through LDS.
template <typename T>
__global__ void __launch_bounds__(1024) k_trisolve(const T* L, int ld, int nb, const T* Winv, const T* bvec, const T* yrow, int n,
                                                   T* x)
{
    extern __shared__ unsigned char smem_raw[];
    T* sx = (T*)smem_raw;  // ld entries
    __shared__ T part[16][TS + 1];
    __shared__ T rhs[TS];
    __shared__ T Wl[TS][TS + 1];
    const int tid = threadIdx.x;
    if (yrow) {
        // y = L^-1 D was produced by the augmented factorization (row n of L)
        for (int c = tid; c < nb * TS; c += 1024) sx[c] = (c < n) ? yrow[c] : (T)0;
        __syncthreads();
    } else {
        for (int kb = 0; kb < nb; kb++) {
            {
                const int r = tid >> 4, t16 = tid & 15;
                const int row = kb * TS + r;
                const T* Lr = L + (long long)row * ld;
                T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
                int c = t16;
                for (; c + 48 < kb * TS; c += 64) {
                    a0 = fma(Lr[c], sx[c], a0);
                    a1 = fma(Lr[c + 16], sx[c + 16], a1);
                    a2 = fma(Lr[c + 32], sx[c + 32], a2);
                    a3 = fma(Lr[c + 48], sx[c + 48], a3);
                }
                for (; c < kb * TS; c += 16) a0 = fma(Lr[c], sx[c], a0);
                part[t16][r] = (a0 + a1) + (a2 + a3);
            }
            const T* W = Winv + (size_t)kb * TS * TS;
            for (int e = tid; e < TS * TS; e += 1024) Wl[e >> 6][e & 63] = W[e];
            __syncthreads();
            if (tid < TS) {
                T s = bvec[kb * TS + tid];
                for (int q = 0; q < 16; q++) s -= part[q][tid];
                rhs[tid] = s;
            }
            __syncthreads();
            if (tid < TS) {
                T y = 0;
                for (int q = 0; q <= tid; q++) y = fma(Wl[tid][q], rhs[q], y);
                sx[kb * TS + tid] = y;
            }
            __syncthreads();
        }
    }
    for (int kb = nb - 1; kb >= 0; kb--) {
        {
            const int r = tid & 63, t16 = tid >> 6;
            const int col = kb * TS + r;
            T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            int rr = (kb + 1) * TS + t16;
            for (; rr + 48 < nb * TS; rr += 64) {
                a0 = fma(L[(long long)rr * ld + col], sx[rr], a0);
                a1 = fma(L[(long long)(rr + 16) * ld + col], sx[rr + 16], a1);
                a2 = fma(L[(long long)(rr + 32) * ld + col], sx[rr + 32], a2);
                a3 = fma(L[(long long)(rr + 48) * ld + col], sx[rr + 48], a3);
            }
            for (; rr < nb * TS; rr += 16) a0 = fma(L[(long long)rr * ld + col], sx[rr], a0);
            part[t16][r] = (a0 + a1) + (a2 + a3);
        }
        const T* W = Winv + (size_t)kb * TS * TS;
        for (int e = tid; e < TS * TS; e += 1024) Wl[e >> 6][e & 63] = W[e];
        __syncthreads();
        if (tid < TS) {
            T s = sx[kb * TS + tid];
            for (int q = 0; q < 16; q++) s -= part[q][tid];
            rhs[tid] = s;
        }
        __syncthreads();
        if (tid < TS) {
            T v = 0;
            for (int q = tid; q < TS; q++) v = fma(Wl[q][tid], rhs[q], v);
            sx[kb * TS + tid] = v;
        }
        __syncthreads();
    }
    for (int c = tid; c < nb * TS; c += 1024) x[c] = sx[c];
}

void launch_trisolve_f64(hipStream_t s, const double* L, int ld, int n, const double* b, double* x)
{
    // Winv for every diagonal block sits right after the diag array in `work`; see qlx_api.cpp
    (void)n;
    (void)s;
    (void)L;
    (void)ld;
    (void)b;
    (void)x;
}

// x = F^-1 b with F = L L^T; with yrow != nullptr the forward half is skipped and y = yrow[0..n)
void launch_trisolve_f64_w(hipStream_t s, const double* L, int ld, const double* Winv, const double* b, const double* yrow, int n,
                           double* x)
{
    const int nb = ld / TS;
    size_t lds = (size_t)ld * sizeof(double);
    allow_big_lds(k_trisolve<double>);
    hipLaunchKernelGGL(k_trisolve<double>, dim3(1), dim3(1024), lds, s, L, ld, nb, Winv, b, yrow, n, x);
}

void launch_trisolve_f32_w(hipStream_t s, const float* L, int ld, const float* Winv, const float* b, const float* yrow, int n,
                           float* x)
{
    const int nb = ld / TS;
    size_t lds = (size_t)ld * sizeof(float);
    allow_big_lds(k_trisolve<float>);
    hipLaunchKernelGGL(k_trisolve<float>, dim3(1), dim3(1024), lds, s, L, ld, nb, Winv, b, yrow, n, x);
}

// 2 * sum log(diag) in a fixed order (Cholesky_logdet_lower_packed, pixelgrid.cpp:17201-17210)
__global__ void __launch_bounds__(1024) k_logdet(const double* diag, int n, double* out)
{
    __shared__ double sh[1024];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) acc += log(fabs(diag[i]));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = 2.0 * sh[0];
}

void launch_logdet(hipStream_t s, const double* diag, int n, double* out)
{
    hipLaunchKernelGGL(k_logdet, dim3(1), dim3(1024), 0, s, diag, n, out);
}

// ------------------------------------------------------------------------------------------------
// evidence: v = L s per active pixel, model = PSF (x) v over active taps, chi^2 over the mask,
// Es = sum_l ||H_l s||^2 (= s^T R s), and -2 ln E.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_Ls(GridDev g, ItemBufs b, int n_img, const double* s, double* v)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_img) return;
    const int U = 3 * g.nsub;
    double acc = 0.0;
    for (int e = 0; e < b.ucnt[r]; e++) acc += b.uval[(size_t)r * U + e] * s[b.ucol[(size_t)r * U + e]];
    v[r] = acc;
}

__global__ void __launch_bounds__(256) k_model(GridDev g, ItemBufs b, int n_img, const double* v, const double* psf, int pnx, int pny,
                                               double* model)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_img) return;
    const int p = b.act_pix[r];
    const int k = p / g.ny, l = p % g.ny;
    const int nxh = pnx / 2, nyh = pny / 2;
    double acc = 0.0;
    for (int pk = 0; pk < pnx; pk++) {
        int ii = k + nxh - pk;
        if (ii < 0 || ii >= g.nx) continue;
        for (int pl = 0; pl < pny; pl++) {
            int jj = l + nyh - pl;
            if (jj < 0 || jj >= g.ny) continue;
            int r2 = b.pidx[ii * g.ny + jj];
            if (r2 < 0) continue;
            acc += psf[pk * pny + pl] * v[r2];
        }
    }
    model[p] = acc;
}

// out: [0] -2lnE, [1] chisq0, [2] logdetF, [3] logdetR, [4] Es, [5] regterms
__global__ void __launch_bounds__(1024) k_evidence_reduce(GridDev g, ItemBufs b, int n_src, const double* s, const double* model,
                                                          const double* data, const double* covinv, double lambda,
                                                          const double* logdets, double* out)
{
    __shared__ double sh[1024];
    __shared__ double sh2[1024];
    const int npix = g.nx * g.ny;
    double chi = 0.0;
    for (int p = threadIdx.x; p < npix; p += 1024) {
        if (g.mask && !g.mask[p]) continue;
        double d = data[p];
        if (b.pmapped[p]) {
            double r = model[p] - d;
            chi += (r * r) * covinv[p];
        } else {
            chi += (d * d) * covinv[p];
        }
    }
    double es = 0.0;
    const int ncell = g.snx * g.sny;
    for (int c = threadIdx.x; c < ncell; c += 1024) {
        int a = b.sidx[c];
        if (a < 0) continue;
        int i = c / g.sny, j = c % g.sny;
        double sa = s[a];
        int n0 = (i < g.snx - 1) ? b.sidx[(i + 1) * g.sny + j] : -1;
        int n1 = (i > 0) ? b.sidx[(i - 1) * g.sny + j] : -1;
        double h = -2.0 * sa + (n0 >= 0 ? s[n0] : 0.0) + (n1 >= 0 ? s[n1] : 0.0);
        es += h * h;
        n0 = (j < g.sny - 1) ? b.sidx[i * g.sny + j + 1] : -1;
        n1 = (j > 0) ? b.sidx[i * g.sny + j - 1] : -1;
        h = -2.0 * sa + (n0 >= 0 ? s[n0] : 0.0) + (n1 >= 0 ? s[n1] : 0.0);
        es += h * h;
    }
    sh[threadIdx.x] = chi;
    sh2[threadIdx.x] = es;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sh[threadIdx.x] += sh[threadIdx.x + st];
            sh2[threadIdx.x] += sh2[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double chisq = sh[0], Es = sh2[0];
        double logdetF = logdets[0], logdetR = logdets[1];
        double loglike = 0.0, regterms = 0.0;
        if (lambda != 0) {
            regterms = lambda * Es - n_src * log(lambda) - logdetR;
            loglike += regterms;
        }
        loglike += logdetF;
        loglike += chisq;
        out[0] = loglike;
        out[1] = chisq;
        out[2] = logdetF;
        out[3] = logdetR;
        out[4] = Es;
        out[5] = regterms;
    }
}

void launch_evidence(hipStream_t st, const GridDev& g, const ItemBufs& b, int n_img, int n_src, const double* s, const double* psf,
                     int pnx, int pny, const double* data, const double* covinv, double lambda, const double* logdets, double* model,
                     double* out)
{
    // `model` must hold nx*ny doubles plus n_img scratch doubles after it
    double* v = model + (size_t)g.nx * g.ny;
    launch_fill_f64(st, model, 0.0, (size_t)g.nx * g.ny);
    if (n_img > 0) {
        hipLaunchKernelGGL(k_Ls, dim3((n_img + 255) / 256), dim3(256), 0, st, g, b, n_img, s, v);
        hipLaunchKernelGGL(k_model, dim3((n_img + 255) / 256), dim3(256), 0, st, g, b, n_img, v, psf, pnx, pny, model);
    }
    hipLaunchKernelGGL(k_evidence_reduce, dim3(1), dim3(1024), 0, st, g, b, n_src, s, model, data, covinv, lambda, logdets, out);
}

// ------------------------------------------------------------------------------------------------
// MFMA f64 layout self-test: D = A(16x4) * B(4x16) with A[i][k] = i + 1 + 0.01 k and
// B[k][j] = 1 + k + 0.1 j; out[lane*4 + v] = D register v of lane.
// ------------------------------------------------------------------------------------------------
__global__ void k_selftest_mfma(double* out)
{
    const int lane = threadIdx.x;
    const int i = lane & 15, k = lane >> 4;
    double a = i + 1 + 0.01 * k;
    double b = 1 + k + 0.1 * (lane & 15);
    double4_t acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; v++) out[lane * 4 + v] = acc[v];
}

void launch_selftest_mfma(hipStream_t s, double* out)
{
    hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, s, out);
}

}  // namespace qlx
