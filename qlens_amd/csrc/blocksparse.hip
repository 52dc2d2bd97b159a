// Block-sparse blurred lensing matrix and F = Lb^T C Lb for gfx950.
//
// The reference builds a dense N_img x N_src blurred matrix and a dense packed F
// (PSF_convolution_Lmatrix_dense pixelgrid.cpp:14092-14176; create_lensing_matrices_from_Lmatrix_dense
// 15590-15711).  Each blurred row is non-zero only over the source cells its PSF footprint maps to,
// so here:
//   1. active image pixels are ordered by the source cell their first ray lands in (a stable
//      counting sort; rows of one 64-row chunk then cover a compact source region),
//   2. each chunk's support is the union of its rows' 64-column blocks (bit masks from the merged L
//      rows of the PSF taps, no values needed),
//   3. rows are blurred in LDS exactly as before and only the chunk's support blocks are written,
//      block-major (block (chunk, colblock) at (chunk*ncb + colblock)*64*64, rows inside row-major),
//   4. every lower output tile of F sums only the chunks whose support holds both of its blocks.
// F is a sum over rows, so the row order changes nothing but the floating-point summation order.
#include "qlx_internal.h"

namespace qlx {

typedef double double4_t __attribute__((ext_vector_type(4)));
#define TS 64
#define LDSP 66

template <typename K>
static void allow_big_lds_bs(K kernel)
{
    static bool done = false;
    if (!done) {
        hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 48 * 1024);
        done = true;
    }
}

// ---- 1. ordering ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_row_keys(GridDev g, ItemBufs b, int n_img)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_img) return;
    const long long p = b.act_pix[r];
    int key = 0;
    for (int k = 0; k < g.nsub; k++) {
        int c0 = b.map_cell[(p * g.nsub + k) * 3];
        if (c0 >= 0) {
            key = b.sidx[c0];
            break;
        }
    }
    b.key[r] = key;
    atomicAdd(&b.bucket_cnt[key], 1);
}

__global__ void __launch_bounds__(1024) k_bucket_scan(ItemBufs b, int n_src)
{
    __shared__ int sh[1024];
    int base = 0;
    for (int t0 = 0; t0 < n_src; t0 += 1024) {
        int t = t0 + threadIdx.x;
        int v = (t < n_src) ? b.bucket_cnt[t] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (t < n_src) {
            b.bucket_off[t] = base + sh[threadIdx.x] - v;
            b.bucket_fill[t] = 0;
        }
        int tot = sh[1023];
        __syncthreads();
        base += tot;
    }
    if (threadIdx.x == 0) b.bucket_off[n_src] = base;
}

__global__ void __launch_bounds__(256) k_row_scatter(ItemBufs b, int n_img)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_img) return;
    int key = b.key[r];
    int pos = b.bucket_off[key] + atomicAdd(&b.bucket_fill[key], 1);
    b.perm[pos] = r;
}

// rows inside one bucket in ascending original order (makes the result independent of atomics)
__global__ void __launch_bounds__(256) k_bucket_sort(ItemBufs b, int n_src)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_src) return;
    int lo = b.bucket_off[k], hi = b.bucket_off[k + 1];
    for (int i = lo + 1; i < hi; i++) {
        int v = b.perm[i];
        int j = i - 1;
        while (j >= lo && b.perm[j] > v) {
            b.perm[j + 1] = b.perm[j];
            j--;
        }
        b.perm[j + 1] = v;
    }
}

// ---- 2. chunk support ---------------------------------------------------------------------
// one wavefront per sorted row: lanes split the PSF taps, OR the 64-column blocks of the tap
// pixels' merged L rows, then one lane ORs the row's mask into its chunk's mask.
__global__ void __launch_bounds__(64) k_row_support(GridDev g, ItemBufs b, int n_img, int n_src, int pnx, int pny, unsigned* nzmask,
                                                    int ncbw)
{
    const int rs = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = b.perm[rs];
    const int p = b.act_pix[r];
    const int k = p / g.ny, l = p % g.ny;
    const int nxh = pnx / 2, nyh = pny / 2;
    const int U = 3 * g.nsub;
    unsigned long long m0 = 0, m1 = 0, m2 = 0, m3 = 0;  // up to 256 column blocks
    for (int t = lane; t < pnx * pny; t += 64) {
        const int pk = t / pny, pl = t % pny;
        const int ii = k + nxh - pk, jj = l + nyh - pl;
        if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny) continue;
        const int r2 = b.pidx[ii * g.ny + jj];
        if (r2 < 0) continue;
        const int cnt = b.ucnt[r2];
        for (int e = 0; e < cnt; e++) {
            int cb = b.ucol[(size_t)r2 * U + e] >> 6;
            unsigned long long bit = 1ull << (cb & 63);
            if (cb < 64) m0 |= bit;
            else if (cb < 128) m1 |= bit;
            else if (cb < 192) m2 |= bit;
            else m3 |= bit;
        }
    }
    // data column block
    {
        int cb = n_src >> 6;
        unsigned long long bit = 1ull << (cb & 63);
        if (lane == 0) {
            if (cb < 64) m0 |= bit;
            else if (cb < 128) m1 |= bit;
            else if (cb < 192) m2 |= bit;
            else m3 |= bit;
        }
    }
    // wave OR-reduce
    for (int off = 32; off > 0; off >>= 1) {
        m0 |= __shfl_xor(m0, off, 64);
        m1 |= __shfl_xor(m1, off, 64);
        m2 |= __shfl_xor(m2, off, 64);
        m3 |= __shfl_xor(m3, off, 64);
    }
    if (lane < ncbw) {
        unsigned long long m = (lane < 2) ? m0 : (lane < 4) ? m1 : (lane < 6) ? m2 : m3;
        unsigned w = (lane & 1) ? (unsigned)(m >> 32) : (unsigned)(m & 0xffffffffu);
        if (w) atomicOr(&nzmask[(size_t)(rs / TS) * ncbw + lane], w);
    }
}

// ---- 3. blurred rows into the chunk's support blocks ------------------------------------
template <typename T>
__device__ __forceinline__ T to_elem(double v);
template <>
__device__ __forceinline__ double to_elem<double>(double v) { return v; }
template <>
__device__ __forceinline__ uint16_t to_elem<uint16_t>(double v)
{
    // round-to-nearest-even float -> bf16
    float f = (float)v;
    unsigned u = __float_as_uint(f);
    unsigned lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

template <typename T>
__global__ void __launch_bounds__(64) k_build_lb_blocks(GridDev g, ItemBufs b, int n_img, int n_src, const double* psf, int pnx, int pny,
                                                        const double* data, const double* covinv, double rowscale_mode, T* pool,
                                                        int ncb, const unsigned* nzmask, int ncbw)
{
    extern __shared__ double row[];
    const int rs = blockIdx.x;
    const int lane = threadIdx.x;
    const int ld = ncb * TS;
    for (int c = lane; c < ld; c += 64) row[c] = 0.0;
    __syncthreads();
    const int U = 3 * g.nsub;
    const int kt = rs / TS, rr = rs % TS;
    if (rs < n_img) {
        const int p = b.act_pix[b.perm[rs]];
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
        if (rowscale_mode != 0.0) {
            // mixed precision stores sqrt(c_r) * row so the BF16 SYRK needs no per-row scale
            const double sc = sqrt(covinv[p]);
            for (int c = lane; c < ld; c += 64) row[c] *= sc;
            __syncthreads();
        }
    }
    const unsigned* m = nzmask + (size_t)kt * ncbw;
    for (int cb = 0; cb < ncb; cb++) {
        if (!((m[cb >> 5] >> (cb & 31)) & 1u)) continue;
        T* dst = pool + ((size_t)kt * ncb + cb) * TS * TS + (size_t)rr * TS;
        dst[lane] = to_elem<T>(row[cb * TS + lane]);
    }
}

__global__ void k_gather_cvec_sorted(ItemBufs b, const double* covinv, int n_img, int nrows_pad, double* cvec)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows_pad) return;
    cvec[r] = (r < n_img) ? covinv[b.act_pix[b.perm[r]]] : 0.0;
}

void launch_sort_rows(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int n_src)
{
    if (n_img <= 0) return;
    launch_fill_i32(s, b.bucket_cnt, 0, (size_t)n_src + 1);
    hipLaunchKernelGGL(k_row_keys, dim3((n_img + 255) / 256), dim3(256), 0, s, g, b, n_img);
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, b, n_src);
    hipLaunchKernelGGL(k_row_scatter, dim3((n_img + 255) / 256), dim3(256), 0, s, b, n_img);
    hipLaunchKernelGGL(k_bucket_sort, dim3((n_src + 255) / 256), dim3(256), 0, s, b, n_src);
}

void launch_row_support(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int n_src, int pnx, int pny, unsigned* nzmask,
                        int ncbw)
{
    if (n_img > 0) hipLaunchKernelGGL(k_row_support, dim3(n_img), dim3(64), 0, s, g, b, n_img, n_src, pnx, pny, nzmask, ncbw);
}

void launch_build_lb_blocks_f64(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int nrows, int n_src, const double* psf,
                                int pnx, int pny, const double* data, double* pool, int ncb, const unsigned* nzmask, int ncbw)
{
    size_t lds = (size_t)ncb * TS * sizeof(double);
    allow_big_lds_bs(k_build_lb_blocks<double>);
    hipLaunchKernelGGL(k_build_lb_blocks<double>, dim3(nrows), dim3(64), lds, s, g, b, n_img, n_src, psf, pnx, pny, data,
                       (const double*)nullptr, 0.0, pool, ncb, nzmask, ncbw);
}

void launch_build_lb_blocks_bf16(hipStream_t s, const GridDev& g, const ItemBufs& b, int n_img, int nrows, int n_src, const double* psf,
                                 int pnx, int pny, const double* data, const double* covinv, uint16_t* pool, int ncb,
                                 const unsigned* nzmask, int ncbw)
{
    size_t lds = (size_t)ncb * TS * sizeof(double);
    allow_big_lds_bs(k_build_lb_blocks<uint16_t>);
    hipLaunchKernelGGL(k_build_lb_blocks<uint16_t>, dim3(nrows), dim3(64), lds, s, g, b, n_img, n_src, psf, pnx, pny, data, covinv, 1.0,
                       pool, ncb, nzmask, ncbw);
}

void launch_gather_cvec_sorted(hipStream_t s, const ItemBufs& b, const double* covinv, int n_img, int nrows, double* cvec)
{
    hipLaunchKernelGGL(k_gather_cvec_sorted, dim3((nrows + 255) / 256), dim3(256), 0, s, b, covinv, n_img, nrows, cvec);
}

// ---- 4. block-sparse SYRK on FP64 MFMA --------------------------------------------------
// one workgroup per lower tile (I, J); 4 waves as 2x2 quadrants of 2x2 v_mfma_f64_16x16x4f64.
// Accumulator layout on gfx950: register v of lane l holds D[l/16 + 4v][l%16].
__global__ void __launch_bounds__(256) k_syrk_blocks_f64(const double* pool, int ncb, int nkt, const double* cvec, double* F, int ldF,
                                                         const unsigned* nz, int ncbw)
{
    __shared__ double As[TS][LDSP];
    __shared__ double Bs[TS][LDSP];
    int t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int J = t - I * (I + 1) / 2;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (double4_t){0, 0, 0, 0};
    for (int kt = 0; kt < nkt; kt++) {
        const unsigned* m = nz + (size_t)kt * ncbw;
        if (!(((m[I >> 5] >> (I & 31)) & 1u) && ((m[J >> 5] >> (J & 31)) & 1u))) continue;
        const double* Ablk = pool + ((size_t)kt * ncb + I) * TS * TS;
        const double* Bblk = pool + ((size_t)kt * ncb + J) * TS * TS;
        const double* cv = cvec + (size_t)kt * TS;
        __syncthreads();
#pragma unroll 4
        for (int e = tid; e < TS * TS; e += 256) {
            const int kk = e >> 6, ii = e & 63;
            As[kk][ii] = Ablk[e];
            Bs[kk][ii] = Bblk[e] * cv[kk];
        }
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
    double* C = F + (size_t)I * TS * ldF + (size_t)J * TS;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int rr = wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                int cc = wc * 32 + c * 16 + (lane & 15);
                C[(size_t)rr * ldF + cc] = acc[a][c][v];
            }
}

void launch_syrk_blocks_f64(hipStream_t s, const double* pool, int ncb, int nkt, const double* cvec, double* F, int ldF,
                            const unsigned* nz, int ncbw)
{
    int ntiles = ncb * (ncb + 1) / 2;
    hipLaunchKernelGGL(k_syrk_blocks_f64, dim3(ntiles), dim3(256), 0, s, pool, ncb, nkt, cvec, F, ldF, nz, ncbw);
}

}  // namespace qlx
