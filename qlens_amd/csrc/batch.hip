// Batched inversion kernels for gfx950: active sets, merged lensing rows, row ordering, block-sparse
// blurred L, F = Lb^T C Lb (FP64 MFMA, or FP16 MFMA with FP32 accumulation), curvature R
// (banded FP64 factorization for ln det R), blocked dense Cholesky of F (FP64 or FP32), solves,
// FP64 iterative refinement and the evidence terms.
//
// Item b of a batch is blockIdx.y (matrix index for the Cholesky kernels); per-item sizes and pool
// offsets come from BatchDev::meta.  Layouts: Lb blocks are 64x64 row-major at
// (chunk*ncb + colblock)*4096 of the item's pool; column n_src carries the data vector so the SYRK
// also yields D = Lb^T C d (row n_src of F).  F and R are row-major ldN x ldN (lower triangle used),
// padding beyond the source block is identity.
//
// Reference behaviour followed:
//   active sets              assign_active_indices / assign_pixel_mappings  pixelgrid.cpp:3586-3626, 12879-12914
//   dense L (merged in CSR)  assign_Lmatrix + convert_Lmatrix_to_dense      pixelgrid.cpp:13196-13321, 13717-13747
//   PSF on L                 PSF_convolution_Lmatrix_dense (direct)         pixelgrid.cpp:14092-14176
//   D, F                     create_lensing_matrices_from_Lmatrix_dense     pixelgrid.cpp:15590-15711
//   curvature R              generate_hmatrices / Rmatrix_from_hmatrices    pixelgrid.cpp:2968-3240
//   F += lambda R            add_regularization_term_to_dense_Fmatrix      pixelgrid.cpp:15808-15818
//   ln det R                 Rmatrix_determinant_dense                      pixelgrid.cpp:18176-18191
//   Cholesky, solve, logdet  Cholesky_dcmp/solve/logdet_lower_packed        pixelgrid.cpp:17071-17246
//   model, chi^2, Es         lens.cpp:12950-13015, pixelgrid.cpp:15821-15863, 18268-18307
#include "qlx_internal.h"

namespace qlx {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

// 16-byte store marked non-temporal (streamed past the caches): for buffers written once and read
// much later, so that they do not push out lines that are still being re-read
__device__ __forceinline__ void store_nt(uint4* p, const uint4& v)
{
    __builtin_nontemporal_store((u32x4_t){v.x, v.y, v.z, v.w}, (u32x4_t*)p);
}
#define TS 64
#define LDSP 66

// raise the dynamic-LDS limit of a kernel once per process (kernels are told apart by address)
static void allow_big_lds(const void* kernel)
{
    static const void* done[32];
    static int ndone = 0;
    for (int i = 0; i < ndone; i++)
        if (done[i] == kernel) return;
    hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 16 * 1024);
    if (ndone < 32) done[ndone++] = kernel;
}

static inline unsigned grid_for(size_t n, unsigned bs = 256, unsigned cap = 65535u * 4u)
{
    size_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// ------------------------------------------------------------------------------------------------
// utilities
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int v, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}
__global__ void k_fill_f64(double* p, double v, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}
void launch_fill_i32(hipStream_t s, int* p, int v, size_t n)
{
    if (n) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}
void launch_fill_f64(hipStream_t s, double* p, double v, size_t n)
{
    if (n) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}

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

// 1024-thread exclusive scan of one flag per thread; returns this thread's exclusive index
__device__ __forceinline__ int block_excl_scan(int flag, int* sh, int& total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = wave_incl_scan(flag);
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    if (w == 0) {
        int v = (lane < 16) ? sh[lane] : 0;
        int s = wave_incl_scan(v);
        if (lane < 16) sh[16 + lane] = s - v;
        if (lane == 15) sh[32] = s;
    }
    __syncthreads();
    int excl = sh[16 + w] + incl - flag;
    total = sh[32];
    __syncthreads();
    return excl;
}

// ------------------------------------------------------------------------------------------------
// active sets: source cells in j-major order, image pixels in j-outer / i-inner order
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_compact_b(BatchDev bd)
{
    __shared__ int sh[40];
    const int item = blockIdx.y;
    const int snx = bd.src_n[2 * item], sny = bd.src_n[2 * item + 1];
    const int* cm = bd.cmapped + item * bd.s_cell;
    int* sidx = bd.sidx + item * bd.s_cell;
    const int* pm = bd.pmapped + item * bd.s_pix;
    int* pidx = bd.pidx + item * bd.s_pix;
    int* act = bd.act_pix + item * bd.s_pix;
    int base = 0;
    const int ncell = snx * sny;
    for (int t0 = 0; t0 < ncell; t0 += 1024) {
        int t = t0 + threadIdx.x, c = -1, flag = 0;
        if (t < ncell) {
            c = (t % snx) * sny + t / snx;
            flag = cm[c] ? 1 : 0;
        }
        int tot, ex = block_excl_scan(flag, sh, tot);
        if (c >= 0) sidx[c] = flag ? base + ex : -1;
        base += tot;
    }
    if (threadIdx.x == 0) bd.counts[4 * item] = base;
    __syncthreads();
    // internal order of F's source columns: the same active cells ranked tile by tile
    // (QLX_TILE_W x QLX_TILE_H cells, tiles and cells inside a tile y-major), so that a compact
    // source-plane patch occupies a compact index range
    {
        int* tix = bd.tix + item * bd.s_cell;
        int* pis = bd.pisrc + item * bd.s_cell;
        const int ntw = (snx + QLX_TILE_W - 1) / QLX_TILE_W, nth = (sny + QLX_TILE_H - 1) / QLX_TILE_H;
        const int tcells = QLX_TILE_W * QLX_TILE_H;
        const int total = ntw * nth * tcells;
        int tbase = 0;
        for (int t0 = 0; t0 < total; t0 += 1024) {
            const int t = t0 + threadIdx.x;
            int c = -1, flag = 0;
            if (t < total) {
                const int tile = t / tcells, wv = t - tile * tcells;
                const int i = (tile % ntw) * QLX_TILE_W + wv % QLX_TILE_W;
                const int j = (tile / ntw) * QLX_TILE_H + wv / QLX_TILE_W;
                if (i < snx && j < sny) {
                    c = i * sny + j;
                    flag = cm[c] ? 1 : 0;
                }
            }
            int tot, ex = block_excl_scan(flag, sh, tot);
            if (c >= 0) tix[c] = flag ? tbase + ex : -1;
            if (flag) pis[tbase + ex] = sidx[c];
            tbase += tot;
        }
    }
    base = 0;
    const int npix = bd.nx * bd.ny;
    int* apre = bd.apre + item * (bd.s_pix + 1);
    for (int t0 = 0; t0 < npix; t0 += 1024) {
        int t = t0 + threadIdx.x, p = -1, flag = 0;
        if (t < npix) {
            p = (t % bd.nx) * bd.ny + t / bd.nx;
            flag = ((!bd.mask || bd.mask[p]) && pm[p]) ? 1 : 0;
        }
        int tot, ex = block_excl_scan(flag, sh, tot);
        if (p >= 0) {
            pidx[p] = flag ? base + ex : -1;
            apre[t] = base + ex;
            if (flag) act[base + ex] = p;
        }
        base += tot;
    }
    if (threadIdx.x == 0) {
        bd.counts[4 * item + 1] = base;
        apre[npix] = base;
    }
}

void launch_compact_b(hipStream_t s, const BatchDev& bd)
{
    hipLaunchKernelGGL(k_compact_b, dim3(1, bd.B), dim3(1024), 0, s, bd);
}

// merged row of L per active pixel: CSR order, duplicates summed from 0 in that order.  Up to
// nsplit 2 (12 entries) the row is built in registers (predicated, no dynamic indexing) and stored
// once; larger splits build it in place.
__global__ void __launch_bounds__(256) k_merge_lrows_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || r >= M.n_img) return;
    const int nsub = bd.nsub, U = 3 * nsub;
    const long long p = bd.act_pix[item * bd.s_pix + r];
    const int* mcb = bd.map_cell + (size_t)item * bd.s_pix * nsub * 3;
    const double* mwb = bd.map_w + (size_t)item * bd.s_pix * nsub * 3;
    const int* tix = bd.tix + item * bd.s_cell;
    int* uc = bd.ucol + ((size_t)item * bd.s_pix + r) * U;
    double* uv = bd.uval + ((size_t)item * bd.s_pix + r) * U;
    if (nsub <= 4) {
        constexpr int MU = 12;
        int cols[MU];
        double vals[MU];
#pragma unroll
        for (int j = 0; j < MU; j++) {
            cols[j] = -1;
            vals[j] = 0.0;
        }
        // the pixel's nsub x 3 ray entries; with nsub = 4 (2 x 2 split) they are 48 + 96 contiguous,
        // 16-byte aligned bytes, read with six vector loads
        int mcv[MU];
        double mwv[MU];
        if (nsub == 4) {
            const int4* mc4 = (const int4*)(mcb + p * MU);
            const double2* mw2 = (const double2*)(mwb + p * MU);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int4 t = mc4[j];
                mcv[4 * j] = t.x;
                mcv[4 * j + 1] = t.y;
                mcv[4 * j + 2] = t.z;
                mcv[4 * j + 3] = t.w;
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const double2 t = mw2[j];
                mwv[2 * j] = t.x;
                mwv[2 * j + 1] = t.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < MU; j++) {
                mcv[j] = (j < 3 * nsub) ? mcb[p * nsub * 3 + j] : -1;
                mwv[j] = (j < 3 * nsub) ? mwb[p * nsub * 3 + j] : 0.0;
            }
        }
        int n = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (mcv[3 * k] < 0) continue;
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const int col = tix[mcv[3 * k + e]];
                const double w = mwv[3 * k + e];
                bool found = false;
#pragma unroll
                for (int j = 0; j < MU; j++)
                    if (!found && j < n && cols[j] == col) {
                        vals[j] += w;
                        found = true;
                    }
                if (!found) {
#pragma unroll
                    for (int j = 0; j < MU; j++)
                        if (j == n) {
                            cols[j] = col;
                            vals[j] = 0.0 + w;
                        }
                    n++;
                }
            }
        }
        unsigned m[QLX_MAX_CBW];
#pragma unroll
        for (int w = 0; w < QLX_MAX_CBW; w++) m[w] = 0;
#pragma unroll
        for (int j = 0; j < MU; j++)
            if (j < n) {
                const int cb = cols[j] >> M.cbs;
#pragma unroll
                for (int w = 0; w < QLX_MAX_CBW; w++) m[w] |= (w == (cb >> 5)) ? (1u << (cb & 31)) : 0u;
            }
        if (nsub == 4) {  // the whole 12-slot row with vector stores (slots >= n: col -1, value 0, never read)
#pragma unroll
            for (int j = 0; j < 3; j++) ((int4*)uc)[j] = make_int4(cols[4 * j], cols[4 * j + 1], cols[4 * j + 2], cols[4 * j + 3]);
#pragma unroll
            for (int j = 0; j < 6; j++) ((double2*)uv)[j] = make_double2(vals[2 * j], vals[2 * j + 1]);
        } else {
#pragma unroll
            for (int j = 0; j < MU; j++)
                if (j < n) {
                    uc[j] = cols[j];
                    uv[j] = vals[j];
                }
        }
        bd.ucnt[item * bd.s_pix + r] = n;
        unsigned* pm = bd.pmask + ((size_t)item * bd.s_pix + r) * QLX_MAX_CBW;
#pragma unroll
        for (int w = 0; w < QLX_MAX_CBW; w++)
            if (w < M.ncbw) pm[w] = m[w];
        return;
    }
    int cnt = 0;
    for (int k = 0; k < nsub; k++) {
        const int* mc = mcb + (p * nsub + k) * 3;
        const double* mw = mwb + (p * nsub + k) * 3;
        if (mc[0] < 0) continue;
        for (int e = 0; e < 3; e++) {
            int col = tix[mc[e]];
            int q = 0;
            while (q < cnt && uc[q] != col) q++;
            if (q == cnt) {
                uc[cnt] = col;
                uv[cnt] = 0.0 + mw[e];
                cnt++;
            } else {
                uv[q] += mw[e];
            }
        }
    }
    bd.ucnt[item * bd.s_pix + r] = cnt;
    unsigned* pm = bd.pmask + ((size_t)item * bd.s_pix + r) * QLX_MAX_CBW;
    for (int w = 0; w < M.ncbw; w++) pm[w] = 0;
    for (int e = 0; e < cnt; e++) {
        const int cb = uc[e] >> M.cbs;
        pm[cb >> 5] |= 1u << (cb & 31);
    }
}

void launch_merge_lrows_b(hipStream_t s, const BatchDev& bd, int max_img)
{
    if (max_img > 0) hipLaunchKernelGGL(k_merge_lrows_b, dim3((max_img + 255) / 256, bd.B), dim3(256), 0, s, bd);
}

// a merged 12-slot L row (2 x 2 split) as 3 + 6 16-byte loads
__device__ __forceinline__ void load_row12(const int* uc, const double* uv, int (&c)[12], double (&w)[12])
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int4 t = ((const int4*)uc)[j];
        c[4 * j] = t.x;
        c[4 * j + 1] = t.y;
        c[4 * j + 2] = t.z;
        c[4 * j + 3] = t.w;
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double2 t = ((const double2*)uv)[j];
        w[2 * j] = t.x;
        w[2 * j + 1] = t.y;
    }
}

// ------------------------------------------------------------------------------------------------
// Row order of the blurred lensing matrix: active pixels in image tiles of 8 x 8 pixels (tiles and
// pixels inside a tile y-major), so that the 64 rows of a chunk are image-compact.  Their PSF
// windows then overlap (the Lb build applies each L entry to every row of the chunk it reaches) and
// their blurred supports are one compact source-plane patch (fewer pool blocks and SYRK block pairs
// than rows ordered by source position: 2 780 against 3 767 blocks and 14.6k against 25.8k pairs on
// the 200 x 200 / 60 x 60 workload).  One wave per tile; the order is a pure function of the active
// set, so an item's rows do not depend on its batch.
//   perm[pos] = active row of tile-order position pos, key[r] = tile-order position of row r,
//   tbase[t] = first position of tile t (the eptr buffer, s_pix + 1 ints per item)
// ------------------------------------------------------------------------------------------------
#define RT 8  // row tile side in pixels
__global__ void __launch_bounds__(64) k_tile_count_b(BatchDev bd, int ntx, int nty)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int t = blockIdx.x;
    if (!M.valid || t >= ntx * nty) return;
    const int lane = threadIdx.x;
    const int i = (t % ntx) * RT + (lane % RT), j = (t / ntx) * RT + lane / RT;
    const bool act = i < bd.nx && j < bd.ny && bd.pidx[item * bd.s_pix + (size_t)i * bd.ny + j] >= 0;
    const unsigned long long m = __ballot(act);
    if (lane == 0) bd.eptr[item * (bd.s_pix + 1) + t] = __popcll(m);
}

__global__ void __launch_bounds__(1024) k_tile_scan_b(BatchDev bd, int ntiles)
{
    __shared__ int sh[40];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    int* tb = bd.eptr + item * (bd.s_pix + 1);
    int base = 0;
    for (int t0 = 0; t0 < ntiles; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        const int v = t < ntiles ? tb[t] : 0;
        int tot, ex = block_excl_scan(v, sh, tot);
        if (t < ntiles) tb[t] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) tb[ntiles] = base;
}

__global__ void __launch_bounds__(64) k_tile_scatter_b(BatchDev bd, int ntx, int nty)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int t = blockIdx.x;
    if (!M.valid || t >= ntx * nty) return;
    const int lane = threadIdx.x;
    const int i = (t % ntx) * RT + (lane % RT), j = (t / ntx) * RT + lane / RT;
    const int r = (i < bd.nx && j < bd.ny) ? bd.pidx[item * bd.s_pix + (size_t)i * bd.ny + j] : -1;
    const unsigned long long m = __ballot(r >= 0);
    if (r < 0) return;
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    const int pos = bd.eptr[item * (bd.s_pix + 1) + t] + rank;
    bd.perm[item * bd.s_pix + pos] = r;
    bd.key[item * bd.s_pix + r] = pos;
}

void launch_sort_rows_b(hipStream_t s, const BatchDev& bd, int max_img, int max_src)
{
    (void)max_src;
    if (max_img <= 0) return;
    const int ntx = (bd.nx + RT - 1) / RT, nty = (bd.ny + RT - 1) / RT;
    hipLaunchKernelGGL(k_tile_count_b, dim3(ntx * nty, bd.B), dim3(64), 0, s, bd, ntx, nty);
    hipLaunchKernelGGL(k_tile_scan_b, dim3(1, bd.B), dim3(1024), 0, s, bd, ntx * nty);
    hipLaunchKernelGGL(k_tile_scatter_b, dim3(ntx * nty, bd.B), dim3(64), 0, s, bd, ntx, nty);
}

// ------------------------------------------------------------------------------------------------
// L by columns (CSC) for the Lb build: for every internal source column c, the entries (pixel, value)
// of the merged L rows, ordered by the rows' tile-order position (a fixed order, so every sum over a
// column's entries is reproducible): cpix (pixel i << 16 | j) and cval, with the column sizes and
// first entries in bcnt and boff (s_cell + 2 ints per item).
// ------------------------------------------------------------------------------------------------
// The columns are built in row-position order directly (a stable counting sort of the entries by
// column): the positions are cut into segments of QLX_CSC_SEG; k_csc_hist_b counts each segment's
// entries per column, k_csc_base_b turns the counts into each segment's first slot within the
// column (segments in order) and the column's size, k_csc_scan_b places the columns, and
// k_csc_place_b lets one wave per segment walk its rows in position order, each entry taking the
// next slot of its column.
__global__ void __launch_bounds__(256) k_csc_hist_b(BatchDev bd)
{
    extern __shared__ int hs[];  // n_src counters
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int seg = blockIdx.x, k0 = seg * QLX_CSC_SEG;
    if (!M.valid || k0 >= M.n_img) return;
    const int n = M.n_src, U = 3 * bd.nsub, tid = threadIdx.x;
    for (int c = tid; c < n; c += 256) hs[c] = 0;
    __syncthreads();
    const int k1 = min(k0 + QLX_CSC_SEG, M.n_img);
    const int* perm = bd.perm + item * bd.s_pix;
    const int* ucnt = bd.ucnt + item * bd.s_pix;
    const int* ucol = bd.ucol + (size_t)item * bd.s_pix * U;
    for (int k = k0 + tid; k < k1; k += 256) {
        const int r = perm[k], nu = ucnt[r];
        for (int e = 0; e < nu; e++) atomicAdd(&hs[ucol[(size_t)r * U + e]], 1);
    }
    __syncthreads();
    int* H = bd.csh + ((size_t)item * csc_segments(bd.s_pix) + seg) * bd.s_cell;
    for (int c = tid; c < n; c += 256) H[c] = hs[c];
}

__global__ void __launch_bounds__(256) k_csc_base_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || c >= M.n_src) return;
    const int nsg = (M.n_img + QLX_CSC_SEG - 1) / QLX_CSC_SEG;
    int* H = bd.csh + (size_t)item * csc_segments(bd.s_pix) * bd.s_cell + c;
    int run = 0;
    for (int sg = 0; sg < nsg; sg++) {
        const int v = H[(size_t)sg * bd.s_cell];
        H[(size_t)sg * bd.s_cell] = run;
        run += v;
    }
    bd.bcnt[item * (bd.s_cell + 2) + c] = run;
}

__global__ void __launch_bounds__(1024) k_csc_scan_b(BatchDev bd)
{
    __shared__ int sh[40];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const int* cnt = bd.bcnt + item * (bd.s_cell + 2);
    int* off = bd.boff + item * (bd.s_cell + 2);
    int* fill = bd.bfill + item * (bd.s_cell + 2);
    int base = 0;
    for (int t0 = 0; t0 < M.n_src; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        const int v = t < M.n_src ? cnt[t] : 0;
        int tot, ex = block_excl_scan(v, sh, tot);
        if (t < M.n_src) {
            off[t] = base + ex;
            fill[t] = 0;
        }
        base += tot;
    }
    if (threadIdx.x == 0) off[M.n_src] = base;
}

// one wave per segment: the column cursors start at the segment's first slots; 64 rows at a time
// are staged in LDS (lane l: row k + l's entries and pixel), then the rows are walked in order, lane
// e placing entry e of the row (a merged row's columns are distinct)
__global__ void __launch_bounds__(64) k_csc_place_b(BatchDev bd)
{
    // dynamic LDS: the staged rows' values (64 x U doubles), columns (64 x U ints), then the n_src
    // column cursors
    extern __shared__ double place_raw[];
    __shared__ int scnt[64];
    __shared__ unsigned spk[64];
    const int U = 3 * bd.nsub;
    double* sval = place_raw;
    int* scol = (int*)(place_raw + 64 * U);
    int* cur = scol + 64 * U;
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int seg = blockIdx.x, k0 = seg * QLX_CSC_SEG;
    if (!M.valid || k0 >= M.n_img) return;
    const int n = M.n_src, lane = threadIdx.x;
    const int* H = bd.csh + ((size_t)item * csc_segments(bd.s_pix) + seg) * bd.s_cell;
    const int* off = bd.boff + item * (bd.s_cell + 2);
    for (int c = lane; c < n; c += 64) cur[c] = off[c] + H[c];
    const int k1 = min(k0 + QLX_CSC_SEG, M.n_img);
    const int* perm = bd.perm + item * bd.s_pix;
    const int* act = bd.act_pix + item * bd.s_pix;
    const int* ucnt = bd.ucnt + item * bd.s_pix;
    const int* ucol = bd.ucol + (size_t)item * bd.s_pix * U;
    const double* uval = bd.uval + (size_t)item * bd.s_pix * U;
    const size_t eb = (size_t)item * bd.s_pix * U;
    unsigned* cpix = bd.cpix + eb;
    double* cval = bd.cval + eb;
    for (int kb = k0; kb < k1; kb += 64) {
        const int k = kb + lane;
        int nu = 0;
        unsigned pk = 0;
        if (k < k1) {
            const int r = perm[k], p = act[r];
            nu = ucnt[r];
            pk = ((unsigned)(p / bd.ny) << 16) | (unsigned)(p - (p / bd.ny) * bd.ny);
            for (int e = 0; e < nu; e++) {
                scol[lane * U + e] = ucol[(size_t)r * U + e];
                sval[lane * U + e] = uval[(size_t)r * U + e];
            }
        }
        scnt[lane] = nu;
        spk[lane] = pk;
        __syncthreads();
        const int nrow = min(64, k1 - kb);
        for (int l = 0; l < nrow; l++) {
            for (int e = lane; e < scnt[l]; e += 64) {  // U <= 48 entries per row
                const int c = scol[l * U + e];
                const int slot = cur[c];
                cur[c] = slot + 1;
                cpix[slot] = spk[l];
                cval[slot] = sval[l * U + e];
            }
        }
        __syncthreads();
    }
}

void launch_lcols_b(hipStream_t s, const BatchDev& bd, int max_img, int max_src)
{
    if (max_img <= 0 || max_src <= 0) return;
    const int nsg = (max_img + QLX_CSC_SEG - 1) / QLX_CSC_SEG;
    const size_t lds = (size_t)bd.s_cell * 4;
    allow_big_lds((const void*)k_csc_hist_b);
    allow_big_lds((const void*)k_csc_place_b);
    hipLaunchKernelGGL(k_csc_hist_b, dim3(nsg, bd.B), dim3(256), lds, s, bd);
    hipLaunchKernelGGL(k_csc_base_b, dim3((max_src + 255) / 256, bd.B), dim3(256), 0, s, bd);
    hipLaunchKernelGGL(k_csc_scan_b, dim3(1, bd.B), dim3(1024), 0, s, bd);
    const size_t U = 3 * (size_t)bd.nsub;
    hipLaunchKernelGGL(k_csc_place_b, dim3(nsg, bd.B), dim3(64), 64 * U * 12 + lds, s, bd);
}

// ------------------------------------------------------------------------------------------------
// chunk support: union of the 64-column blocks touched by the PSF taps of a chunk's 64 rows
// (one workgroup per chunk), then per-item slot bases of the compact block pool.
// ------------------------------------------------------------------------------------------------

// m[o .. o+3] |= the four words of v.  Mask rows hold QLX_MAX_CBW = 8 words; words from ncbw on may
// hold anything (they are never written to the support masks)
static_assert(QLX_MAX_CBW == 8, "mask rows are read as two 16-byte vectors");
__device__ __forceinline__ void or_mask4(unsigned (&m)[QLX_MAX_CBW], int o, uint4 v)
{
    m[o] |= v.x;
    m[o + 1] |= v.y;
    m[o + 2] |= v.z;
    m[o + 3] |= v.w;
}

// PSF x window, per grid pixel t = y*nx + x: hmask[t] = OR of pmask over the active pixels x' in
// [x + nxh - pnx + 1, x + nxh] of row y (one contiguous active-index range [apre[..i0], apre[..i1+1]))
__global__ void __launch_bounds__(256) k_psf_hdilate_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int npix = bd.nx * bd.ny;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (!M.valid || t >= npix) return;
    const int y = t / bd.nx, x = t - y * bd.nx;
    const int nxh = bd.pnx / 2;
    const int i0 = max(0, x + nxh - bd.pnx + 1), i1 = min(bd.nx - 1, x + nxh);
    const int* apre = bd.apre + item * (bd.s_pix + 1);
    const unsigned* pm = bd.pmask + (size_t)item * bd.s_pix * QLX_MAX_CBW;
    unsigned m[QLX_MAX_CBW];
#pragma unroll
    for (int w = 0; w < QLX_MAX_CBW; w++) m[w] = 0;
    const int a0 = apre[y * bd.nx + i0], a1 = apre[y * bd.nx + i1 + 1];
    const bool wide = M.ncbw > 4;  // mask rows are 32-byte aligned: one or two 16-byte loads per row
    for (int a = a0; a < a1; a++) {
        const uint4* p4 = (const uint4*)(pm + (size_t)a * QLX_MAX_CBW);
        or_mask4(m, 0, p4[0]);
        if (wide) or_mask4(m, 4, p4[1]);
    }
    uint4* h4 = (uint4*)(bd.hmask + ((size_t)item * bd.s_pix + t) * QLX_MAX_CBW);
    h4[0] = make_uint4(m[0], m[1], m[2], m[3]);
    if (wide) h4[1] = make_uint4(m[4], m[5], m[6], m[7]);
}

__global__ void __launch_bounds__(256) k_row_support_b(BatchDev bd)
{
    __shared__ unsigned msk[QLX_MAX_CBW];
    __shared__ int rk[TS], rl[TS];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int kt = blockIdx.x;
    if (!M.valid || kt >= M.nkt) return;
    const int tid = threadIdx.x;
    if (tid < QLX_MAX_CBW) msk[tid] = 0;
    const int rows = min(TS, M.n_img - kt * TS);
    if (tid < rows) {
        const int p = bd.act_pix[item * bd.s_pix + bd.perm[item * bd.s_pix + kt * TS + tid]];
        rk[tid] = p / bd.ny;
        rl[tid] = p - (p / bd.ny) * bd.ny;
    }
    __syncthreads();
    unsigned m[QLX_MAX_CBW];
#pragma unroll
    for (int w = 0; w < QLX_MAX_CBW; w++) m[w] = 0;
    {
        const int cbd = M.n_src >> M.cbs;  // the data column (thread 0)
#pragma unroll
        for (int w = 0; w < QLX_MAX_CBW; w++) m[w] |= (tid == 0 && w == (cbd >> 5)) ? (1u << (cbd & 31)) : 0u;
    }
    const int nyh = bd.pny / 2;
    const unsigned* hm = bd.hmask + (size_t)item * bd.s_pix * QLX_MAX_CBW;
    // (row, window row) pairs: the PSF window of row (k, l) at y = l + nyh - pl is the x-window
    // union already held by hmask[y*nx + k]
    for (int q = tid; q < rows * bd.pny; q += 256) {
        const int rr = q / bd.pny, pl = q - rr * bd.pny;
        const int jj = rl[rr] + nyh - pl;
        if (jj < 0 || jj >= bd.ny) continue;
        const uint4* h4 = (const uint4*)(hm + ((size_t)jj * bd.nx + rk[rr]) * QLX_MAX_CBW);
        or_mask4(m, 0, h4[0]);
        if (M.ncbw > 4) or_mask4(m, 4, h4[1]);
    }
#pragma unroll
    for (int w = 0; w < QLX_MAX_CBW; w++)
        if (m[w]) atomicOr(&msk[w], m[w]);
    __syncthreads();
    if (tid < M.ncbw) bd.nz[M.nz_off + (size_t)kt * M.ncbw + tid] = msk[tid];
}

__global__ void __launch_bounds__(1024) k_block_scan_b(BatchDev bd)
{
    __shared__ int sh[40];
    __shared__ int smax[1024];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const unsigned* nz = bd.nz + M.nz_off;
    int* bb = bd.bbase + M.bb_off;
    int base = 0, best = 0;
    for (int t0 = 0; t0 < M.nkt; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        int v = 0;
        if (t < M.nkt)
            for (int w = 0; w < M.ncbw; w++) v += __popc(nz[(size_t)t * M.ncbw + w]);
        best = max(best, v);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int incl = wave_incl_scan(v);
        if (lane == 63) sh[wv] = incl;
        __syncthreads();
        if (wv == 0) {
            int x = (lane < 16) ? sh[lane] : 0;
            int sx = wave_incl_scan(x);
            if (lane < 16) sh[16 + lane] = sx - x;
            if (lane == 15) sh[32] = sx;
        }
        __syncthreads();
        if (t < M.nkt) bb[t] = base + sh[16 + wv] + incl - v;
        const int tot = sh[32];
        __syncthreads();
        base += tot;
    }
    smax[threadIdx.x] = best;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bb[M.nkt] = base;
        bd.counts[4 * item + 2] = base;
        bd.counts[4 * item + 3] = smax[0];
    }
}

void launch_row_support_b(hipStream_t s, const BatchDev& bd, int max_nkt)
{
    if (max_nkt <= 0) return;
    const int npix = bd.nx * bd.ny;
    hipLaunchKernelGGL(k_psf_hdilate_b, dim3((npix + 255) / 256, bd.B), dim3(256), 0, s, bd);
    hipLaunchKernelGGL(k_row_support_b, dim3(max_nkt, bd.B), dim3(256), 0, s, bd);
    hipLaunchKernelGGL(k_block_scan_b, dim3(1, bd.B), dim3(1024), 0, s, bd);
}

// rank of block cb among the set bits of a chunk's support mask
__device__ __forceinline__ int mask_rank(const unsigned* m, int cb)
{
    int s = 0;
    for (int w = 0; w < (cb >> 5); w++) s += __popc(m[w]);
    return s + __popc(m[cb >> 5] & ((1u << (cb & 31)) - 1u));
}
__device__ __forceinline__ bool mask_has(const unsigned* m, int cb) { return (m[cb >> 5] >> (cb & 31)) & 1u; }

// ------------------------------------------------------------------------------------------------
// Blurred rows of one 64-row chunk into the chunk's support blocks (PSF_convolution_Lmatrix_dense,
// direct branch, pixelgrid.cpp:14092-14176): Lb[r][c] = sum over pixels (i, j) of the chunk's PSF
// reach of psf[k_r + hx - i][l_r + hy - j] * L[(i, j)][c], taps inside the PSF only, (i, j) active.
// One workgroup per chunk; lane l of every wave owns chunk row l (the rows are image-compact, so an
// L entry of a pixel near the chunk reaches most of its rows).  Per support column c (columns of a
// block dealt to the four waves), the wave walks column c's entries in L's CSC form, keeps those
// whose pixel lies in the chunk's PSF reach (a ballot over 64 entries at a time), broadcasts each
// kept entry and every lane adds its own tap weight times the value: the output element is summed
// in registers in a fixed order and written once, with no atomics.  A block's 64 x CB values go to
// an LDS tile and from there to the pool with 16-byte stores.
//   T = double: FP64 pool, 64-column blocks, row-major, unscaled (the FP64 SYRK applies C);
//   T = float:  FP16 pool, 128-column blocks, k4-interleaved (element (k, i) at
//               ((k >> 2) * CB + i) * 4 + (k & 3), so the SYRK stages blocks with 16-byte copies),
//               entries x = Lb sqrt(c_r) s16 rounded once (s16 a power of two, |x| <= 2^14).
// Column n_src carries the data: d (FP64) or d sqrt(c_r) s16 (FP16).
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f32_to_f16_rne(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float readlane_f32(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
template <typename T> __device__ __forceinline__ T readlane_t(T v, int l);
template <> __device__ __forceinline__ double readlane_t<double>(double v, int l) { return readlane_f64(v, l); }
template <> __device__ __forceinline__ float readlane_t<float>(float v, int l) { return readlane_f32(v, l); }

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Tap table in LDS: tap (tp, tq) of the PSF at (tp + toff) * tpc + tq.  Padded form (toff =
// LB_SPAN + 1, tpc >= pny + LB_SPAN + 1, pnx + 2 LB_SPAN + 2 rows, zero outside the PSF): for a chunk
// whose rows span at most LB_SPAN + 1 pixels in each direction, every kept L entry (i, j) of the
// chunk's reach and every row (k, l) of it give tp = k + hx - i in [-LB_SPAN, pnx - 1 + LB_SPAN] and
// tq = l + hy - j in [-LB_SPAN, pny - 1 + LB_SPAN], so the flat index K(k, l) - E(i, j), with
// K = (k + hx + toff) tpc + l + hy and E = i tpc + j, lands on the tap when it lies in the PSF and
// on a zero otherwise (a tq past either edge falls on the zero columns of that row or the one
// before): no bounds test, and the zero products add exact zeros, so the sums are the bounds-tested
// ones bit for bit.  Other chunks (and PSFs too large for the padded table: compact form, toff = 0,
// tpc = pny) test the bounds per entry.
#define LB_SPAN 15

template <typename T>
struct LbEnt {
    int e;
    T v;
};

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Half h (CB/2 columns) of support block cb of a chunk by columns: wave w takes the half's columns
// w + 4 l; lane = chunk row.  Per column the wave walks the column's entries in L's CSC form, keeps
// those whose pixel lies in the chunk's PSF reach (a ballot over 64 entries at a time) and every lane
// adds its own tap weight times each kept value, in the entries' order.  Fast chunks pass the kept
// entries (E, value) through the wave's LDS slots (ent) and read the padded table at K - E; the
// others broadcast each entry by readlane and test the tap's bounds.  Writes the half block's image
// into `tile` (no barrier inside).
template <typename T>
__device__ __forceinline__ void lb_half(const BatchDev& bd, const ItemMeta& M, const T* tab, int toff, int tpc, LbEnt<T>* ent,
                                        const unsigned* colbits, char* tile, int cb, int h, int w, int lane, int p, int kr2,
                                        int lr2, int K, bool fast, int i_lo, unsigned i_span, int j_lo, unsigned j_span,
                                        double sc, double dv, const unsigned* cpix, const double* cval, const int* cptr)
{
    constexpr bool F64 = sizeof(T) == 8;
    const int cbs = M.cbs, CB = 1 << cbs, HB = CB >> 1;
    const int pnx = bd.pnx, pny = bd.pny;
    const int ncol = HB / 4;
    const int c0h = (cb << cbs) + h * HB;
    const int cl = c0h + w + 4 * lane;
    int lb0 = 0, lb1 = 0;
    if (lane < ncol && cl < M.n_src) {
        lb0 = cptr[cl];
        lb1 = cptr[cl + 1];
    }
    // the wave's columns with an entry in the chunk's reach (and the data column n_src): the others
    // sum nothing, and their outputs are written directly
    const bool live = lane < ncol && (cl == M.n_src || (cl < M.n_src && ((colbits[cl >> 5] >> (cl & 31)) & 1u)));
    unsigned long long lmask = __ballot(live);
    int nxt = lmask ? __builtin_ctzll(lmask) : ncol;
    int ce0 = 0, ce1 = 0;
    unsigned pk = 0;
    double vd = 0.0;
    if (nxt < ncol) {
        ce0 = __builtin_amdgcn_readlane(lb0, nxt);
        ce1 = __builtin_amdgcn_readlane(lb1, nxt);
        if (ce0 + lane < ce1) {
            pk = cpix[ce0 + lane];
            vd = cval[ce0 + lane];
        }
    }
    for (int l = 0; l < ncol; l++) {
        const int ii = w + 4 * l;
        if (l != nxt) {  // no entry in reach: acc = 0
            const int c_ = c0h + ii;
            if constexpr (F64) {
                ((double*)tile)[lane * (HB + 2) + ii] = c_ == M.n_src ? dv : 0.0;
            } else {
                const double x = c_ == M.n_src ? dv : 0.0 * sc;
                ((uint16_t*)tile)[((lane >> 2) * (HB + 2) + ii) * 4 + (lane & 3)] = f32_to_f16_rne((float)x);
            }
            continue;
        }
        lmask &= lmask - 1;
        nxt = lmask ? __builtin_ctzll(lmask) : ncol;
        const int e0c = ce0, e1c = ce1;
        unsigned pkc = pk;
        double vc = vd;
        pk = 0;
        vd = 0.0;
        if (nxt < ncol) {
            ce0 = __builtin_amdgcn_readlane(lb0, nxt);
            ce1 = __builtin_amdgcn_readlane(lb1, nxt);
            if (ce0 + lane < ce1) {
                pk = cpix[ce0 + lane];
                vd = cval[ce0 + lane];
            }
        }
        T acc = (T)0;
        for (int e0 = e0c; e0 < e1c; e0 += 64) {
            if (e0 != e0c) {
                pkc = 0;
                vc = 0.0;
                if (e0 + lane < e1c) {
                    pkc = cpix[e0 + lane];
                    vc = cval[e0 + lane];
                }
            }
            const bool rel = e0 + lane < e1c && (unsigned)((int)(pkc >> 16) - i_lo) <= i_span &&
                             (unsigned)((int)(pkc & 0xffffu) - j_lo) <= j_span;
            const T v = (T)vc;
            unsigned long long bal = __ballot(rel);
            if (fast) {
                if (rel) {
                    const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                    ent[pos].e = ((int)(pkc >> 16) * tpc + (int)(pkc & 0xffffu)) * (int)sizeof(T);
                    ent[pos].v = v;
                }
                wave_lds_sync();
                const int n = __popcll(bal);
                const char* tb = (const char*)tab + K * (int)sizeof(T);  // byte offsets: tap at tb - E
                int k = 0;
                if constexpr (sizeof(T) == 4) {
                    // two entries per 16-byte LDS read (the same value in every lane)
                    for (; k + 4 <= n; k += 4) {
                        const uint4 q0 = *(const uint4*)&ent[k], q1 = *(const uint4*)&ent[k + 2];
                        const T t0 = *(const T*)(tb - (int)q0.x), t1 = *(const T*)(tb - (int)q0.z);
                        const T t2 = *(const T*)(tb - (int)q1.x), t3 = *(const T*)(tb - (int)q1.z);
                        acc += t0 * __uint_as_float(q0.y);
                        acc += t1 * __uint_as_float(q0.w);
                        acc += t2 * __uint_as_float(q1.y);
                        acc += t3 * __uint_as_float(q1.w);
                    }
                }
                for (; k + 4 <= n; k += 4) {  // four entries' reads in flight together, summed in order
                    const LbEnt<T> x0 = ent[k], x1 = ent[k + 1], x2 = ent[k + 2], x3 = ent[k + 3];
                    const T t0 = *(const T*)(tb - x0.e), t1 = *(const T*)(tb - x1.e);
                    const T t2 = *(const T*)(tb - x2.e), t3 = *(const T*)(tb - x3.e);
                    acc += t0 * x0.v;
                    acc += t1 * x1.v;
                    acc += t2 * x2.v;
                    acc += t3 * x3.v;
                }
                for (; k < n; k++) acc += *(const T*)(tb - ent[k].e) * ent[k].v;
                wave_lds_sync();
            } else {
                while (bal) {
                    const int b = __builtin_ctzll(bal);
                    bal &= bal - 1;
                    const unsigned pkb = (unsigned)__builtin_amdgcn_readlane((int)pkc, b);
                    const T vb = readlane_t<T>(v, b);
                    const int tp = kr2 - (int)(pkb >> 16), tq = lr2 - (int)(pkb & 0xffffu);
                    if (p >= 0 && (unsigned)tp < (unsigned)pnx && (unsigned)tq < (unsigned)pny) acc += tab[(tp + toff) * tpc + tq] * vb;
                }
            }
        }
        if (p < 0) acc = (T)0;  // rows past the image: the fast path read row 0's taps for them
        const int c_ = c0h + ii;
        if constexpr (F64) {
            ((double*)tile)[lane * (HB + 2) + ii] = c_ == M.n_src ? dv : (double)acc;
        } else {
            const double x = c_ == M.n_src ? dv : (double)acc * sc;
            ((uint16_t*)tile)[((lane >> 2) * (HB + 2) + ii) * 4 + (lane & 3)] = f32_to_f16_rne((float)x);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_build_lb_b(BatchDev bd, int tab_bytes, int toff, int tpc, int trows, int nbw)
{
    extern __shared__ double lds_raw[];
    const int item = blockIdx.y, kt = blockIdx.x;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid || kt >= M.nkt) return;
    constexpr bool F64 = sizeof(T) == 8;
    const int cbs = M.cbs, CB = 1 << cbs, HB = CB >> 1;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int pnx = bd.pnx, pny = bd.pny;
    T* tab = (T*)lds_raw;
    LbEnt<T>* ents = (LbEnt<T>*)((char*)lds_raw + tab_bytes);
    unsigned* colbits = (unsigned*)(ents + 4 * 64);  // nbw words: columns with an entry in the reach
    char* tile = (char*)(colbits + nbw);
    for (int t = tid; t < nbw; t += 256) colbits[t] = 0u;
    for (int t = tid; t < trows * tpc; t += 256) {
        const int tr = t / tpc, tq = t - tr * tpc, tp = tr - toff;
        tab[t] = (tp >= 0 && tp < pnx && tq < pny) ? (T)bd.psf[tp * pny + tq] : (T)0;
    }
    // this lane's row
    const int rs = kt * TS + lane;
    const int p = rs < M.n_img ? bd.act_pix[item * bd.s_pix + bd.perm[item * bd.s_pix + rs]] : -1;
    const int kr = p >= 0 ? p / bd.ny : 0, lr = p >= 0 ? p - (p / bd.ny) * bd.ny : 0;
    // the chunk's PSF reach: pixels (i, j) with i in [kmin + hx - pnx + 1, kmax + hx], likewise j
    const int hx = pnx / 2, hy = pny / 2;
    const int kmin = wave_min_i(p >= 0 ? kr : 1 << 30), kmax = wave_max_i(p >= 0 ? kr : -(1 << 30));
    const int lmin = wave_min_i(p >= 0 ? lr : 1 << 30), lmax = wave_max_i(p >= 0 ? lr : -(1 << 30));
    const int i_lo = kmin + hx - pnx + 1, j_lo = lmin + hy - pny + 1;
    const unsigned i_span = (unsigned)(kmax - kmin + pnx - 1), j_span = (unsigned)(lmax - lmin + pny - 1);
    const int kr2 = kr + hx, lr2 = lr + hy;  // tap (kr2 - i, lr2 - j) of pixel (i, j)
    const bool fast = toff > 0 && kmax - kmin <= LB_SPAN && lmax - lmin <= LB_SPAN;
    int K = (kr2 + toff) * tpc + lr2;
    K = p >= 0 ? K : __builtin_amdgcn_readlane(K, 0);  // lane 0 holds a row of every chunk
    double sc = 0.0, dv = 0.0;
    if (p >= 0) {
        if constexpr (F64) {
            dv = bd.data[p];
        } else {
            sc = sqrt(bd.covinv[p]) * bd.s16;
            dv = bd.data[p] * sc;
        }
    }
    const unsigned* m = bd.nz + M.nz_off + (size_t)kt * M.ncbw;
    const int sbase = bd.bbase[M.bb_off + kt];
    const size_t bsz = (size_t)TS * CB;  // elements of one pool block
    const size_t eb = (size_t)item * bd.s_pix * 3 * bd.nsub;
    const unsigned* cpix = bd.cpix + eb;
    const double* cval = bd.cval + eb;
    const int* cptr = bd.boff + item * (bd.s_cell + 2);
    LbEnt<T>* ent = ents + 64 * w;
    __syncthreads();  // colbits cleared
    // the columns of the L rows of the active pixels in the chunk's reach (the rectangle the CSC
    // filter below keeps): exactly the columns with at least one kept entry
    {
        const int U = 3 * bd.nsub, wj = (int)j_span + 1;
        const int nrect = ((int)i_span + 1) * wj;
        const int* pidx = bd.pidx + item * bd.s_pix;
        const int* ucnt = bd.ucnt + item * bd.s_pix;
        const int* ucol = bd.ucol + (size_t)item * bd.s_pix * U;
        for (int t = tid; t < nrect; t += 256) {
            const int i = i_lo + t / wj, j = j_lo + t % wj;
            if (i < 0 || i >= bd.nx || j < 0 || j >= bd.ny) continue;
            const int r = pidx[(size_t)i * bd.ny + j];
            if (r < 0) continue;
            const int nu = ucnt[r];
            for (int e = 0; e < nu; e++) {
                const int c = ucol[(size_t)r * U + e];
                atomicOr(&colbits[c >> 5], 1u << (c & 31));
            }
        }
    }
    __syncthreads();
    // a half block's image: runs of HB * 8 bytes (FP64: 64 rows; FP16: 16 groups of 4 rows, the
    // k4-interleaved layout), each landing at byte 8 (run * CB + h * HB) of the pool block.  In LDS
    // the runs are HB * 8 + 16 bytes apart: at HB * 8 bytes (a multiple of 256) every lane's store of
    // a column fell on the same two banks
    const int nrun = F64 ? TS : TS / 4, upr = HB / 2;
    int q = 0;
    for (int wd = 0; wd < M.ncbw; wd++) {
        unsigned bits = m[wd];
        while (bits) {
            const int cb = wd * 32 + __builtin_ctz(bits);
            bits &= bits - 1;
            uint4* dst = (uint4*)((char*)bd.lb + (M.lb_off + (size_t)(sbase + q) * bsz) * (F64 ? 8 : 2));
            for (int h = 0; h < 2; h++) {
                lb_half<T>(bd, M, tab, toff, tpc, ent, colbits, tile, cb, h, w, lane, p, kr2, lr2, K, fast, i_lo, i_span,
                           j_lo, j_span, sc, dv, cpix, cval, cptr);
                __syncthreads();
                const uint4* src = (const uint4*)tile;
                for (int u = tid; u < nrun * upr; u += 256) {
                    const int run = u >> (cbs - 2), o = u & (upr - 1);  // upr = 2^(cbs - 2)
                    store_nt(&dst[(run * CB + h * HB) / 2 + o], src[run * (upr + 1) + o]);
                }
                __syncthreads();
            }
            q++;
        }
    }
}

int launch_build_lb_b(hipStream_t s, const BatchDev& bd, int max_rows, int max_nsup, int mixed, int cbw)
{
    if (max_rows <= 0 || max_nsup <= 0) return 0;
    const int max_nkt = (max_rows + TS - 1) / TS;
    const size_t esz = mixed ? 4 : 8;
    // padded tap table when it fits in 32 KB, else the compact PSF (bounds tested everywhere)
    int tpc = bd.pny + LB_SPAN + 1;
    tpc = (tpc + 7) / 16 * 16 + 8;  // odd multiple of 8: the rows of an 8 x 8 pixel tile's taps on distinct banks
    int toff = LB_SPAN + 1, trows = bd.pnx + 2 * LB_SPAN + 2;
    if ((size_t)trows * tpc * esz > 32 * 1024) {
        tpc = bd.pny;
        toff = 0;
        trows = bd.pnx;
    }
    const size_t tab_bytes = ((size_t)trows * tpc * esz + 15) / 16 * 16;
    if (tab_bytes > 64 * 1024) return -1;  // PSF beyond the LDS budget
    const size_t tile_bytes = (size_t)TS * (cbw / 2) * (mixed ? 2 : 8) + (mixed ? TS / 4 : TS) * 16;  // + 16 B a run
    const int nbw = (int)((bd.s_cell + 32) / 32 + 3) / 4 * 4;  // column bitmap words (16-byte multiple)
    const size_t lds = tab_bytes + 4 * 64 * (mixed ? sizeof(LbEnt<float>) : sizeof(LbEnt<double>)) + (size_t)nbw * 4 + tile_bytes;
    if (mixed) {
        allow_big_lds((const void*)k_build_lb_b<float>);
        hipLaunchKernelGGL(k_build_lb_b<float>, dim3(max_nkt, bd.B), dim3(256), lds, s, bd, (int)tab_bytes, toff, tpc, trows, nbw);
    } else {
        allow_big_lds((const void*)k_build_lb_b<double>);
        hipLaunchKernelGGL(k_build_lb_b<double>, dim3(max_nkt, bd.B), dim3(256), lds, s, bd, (int)tab_bytes, toff, tpc, trows, nbw);
    }
    return 0;
}

__global__ void k_gather_cvec_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || r >= M.nrows) return;
    bd.cvec[M.cvec_off + r] =
        (r < M.n_img) ? bd.covinv[bd.act_pix[item * bd.s_pix + bd.perm[item * bd.s_pix + r]]] : 0.0;
}

void launch_gather_cvec_b(hipStream_t s, const BatchDev& bd, int max_rows)
{
    if (max_rows > 0) hipLaunchKernelGGL(k_gather_cvec_b, dim3((max_rows + 255) / 256, bd.B), dim3(256), 0, s, bd);
}

__device__ __forceinline__ void tri_decode(int t, int& I, int& J)
{
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    while (i * (i + 1) / 2 > t) i--;
    I = i;
    J = t - i * (i + 1) / 2;
}

// Tile order of the SYRK and Cholesky-update grids.  Workgroups are dealt to the 8 XCDs
// round-robin by linear index, and each XCD has its own L2.  The lower-triangular tile set is cut
// into SUPER x SUPER super-tiles (all slots of a super-tile are launched; slots above the
// diagonal or past the matrix exit at once) and the grid is arranged so that every super-tile
// lands whole on one XCD: its SUPER^2 workgroups run together there and share SUPER row blocks
// and SUPER column blocks in that L2.  Super-tiles go to XCDs round-robin, starting at an XCD
// rotated by the item (blockIdx.y), so small triangles of many items still spread over all XCDs.
#define SUPER 8
#define NXCD 8
// Extents below one super-tile get a plain triangular grid (g -> tri_decode(g)); larger ones whole
// super-tile rounds.  syrk_tile tells the two apart by the grid size.
static unsigned syrk_grid(int max_ncb)
{
    if (max_ncb < SUPER) return (unsigned)std::max(1, max_ncb * (max_ncb + 1) / 2);
    const unsigned nS = (max_ncb + SUPER - 1) / SUPER;
    const unsigned nsup = nS * (nS + 1) / 2;
    const unsigned rounds = (nsup + NXCD - 1) / NXCD;
    return rounds * NXCD * SUPER * SUPER;
}

__device__ __forceinline__ bool syrk_tile(int ncb, int& I, int& J)
{
    const unsigned g = blockIdx.x;
    if (gridDim.x < NXCD * SUPER * SUPER) {
        tri_decode((int)g, I, J);
        return I < ncb;
    }
    const unsigned x = g % NXCD, k = g / NXCD;
    const unsigned w = k % (SUPER * SUPER), r = k / (SUPER * SUPER);
    const unsigned vx = (x + NXCD - (blockIdx.y % NXCD)) % NXCD;
    int SI, SJ;
    tri_decode((int)(r * NXCD + vx), SI, SJ);
    I = SI * SUPER + (int)(w / SUPER);
    J = SJ * SUPER + (int)(w % SUPER);
    return I < ncb && J <= I;
}

// ------------------------------------------------------------------------------------------------
// F = Lb^T C Lb, FP64 MFMA (v_mfma_f64_16x16x4f64; accumulator register v of lane l holds
// D[l/16 + 4v][l%16] on gfx950).  One workgroup per lower tile; only chunks holding both blocks.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_syrk_f64_b(BatchDev bd)
{
    __shared__ double As[TS][LDSP];
    __shared__ double Bs[TS][LDSP];
    extern __shared__ unsigned long long cmask[];  // chunks holding both blocks, 64 a word
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    int I, J;
    if (!syrk_tile(M.ncb, I, J)) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const double* pool = (const double*)bd.lb + M.lb_off;
    const double* cvec = bd.cvec + M.cvec_off;
    const unsigned* nz = bd.nz + M.nz_off;
    const int* bb = bd.bbase + M.bb_off;
    const int nwd = (M.nkt + 63) >> 6;
    for (int q = w; q < nwd; q += 4) {
        const int k = 64 * q + lane;
        bool b = false;
        if (k < M.nkt) {
            const unsigned* m = nz + (size_t)k * M.ncbw;
            b = mask_has(m, I) && mask_has(m, J);
        }
        const unsigned long long bal = __ballot(b);
        if (lane == 0) cmask[q] = bal;
    }
    __syncthreads();
    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (double4_t){0, 0, 0, 0};
    // the next chunk after k holding both blocks (M.nkt if none)
    auto next = [&](int k) {
        ++k;
        for (int q = k >> 6; q < nwd; q++) {
            unsigned long long v = cmask[q];
            if (q == (k >> 6)) v &= ~0ULL << (k & 63);
            if (v) return 64 * q + __builtin_ctzll(v);
        }
        return M.nkt;
    };
    // the next chunk's blocks and weights are loaded into registers while this one is multiplied:
    // thread element i is e = tid + 256 i (row kk = e >> 6 = tid / 64 + 4 i, column tid % 64)
    double ra[16], rb[16], rc[16];
    auto load = [&](int kt) {
        const unsigned* m = nz + (size_t)kt * M.ncbw;
        const double* Ab = pool + (size_t)(bb[kt] + mask_rank(m, I)) * TS * TS;
        const double* Bb = pool + (size_t)(bb[kt] + mask_rank(m, J)) * TS * TS;
        const double* cv = cvec + (size_t)kt * TS;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            ra[i] = Ab[tid + 256 * i];
            rb[i] = Bb[tid + 256 * i];
            rc[i] = cv[(tid >> 6) + 4 * i];
        }
    };
    int kt = next(-1);
    if (kt < M.nkt) load(kt);
    while (kt < M.nkt) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int kk = (tid >> 6) + 4 * i, ii = tid & 63;
            As[kk][ii] = ra[i];
            Bs[kk][ii] = rb[i] * rc[i];
        }
        __syncthreads();
        const int nk = next(kt);
        if (nk < M.nkt) load(nk);
        kt = nk;
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
    double* C = bd.F + M.f_off + (size_t)I * TS * M.ldN + (size_t)J * TS;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int cc = wc * 32 + c * 16 + (lane & 15);
                C[(size_t)rr * M.ldN + cc] = acc[a][c][v];
            }
}

void launch_syrk_f64_b(hipStream_t s, const BatchDev& bd, int max_ncb, int max_nkt)
{
    hipLaunchKernelGGL(k_syrk_f64_b, dim3(syrk_grid(max_ncb), bd.B), dim3(256), (size_t)((max_nkt + 63) / 64) * 8, s, bd);
}

// ------------------------------------------------------------------------------------------------
// F32 = Lb^T Lb on the mixed path: one FP16 pool of sqrt(c) s16-scaled rows in k4-interleaved
// 64 x 128 blocks, v_mfma_f32_16x16x32_f16 accumulated in FP32, scaled back by s16^-2 (exact) at
// the store.  One workgroup per 128x128 lower tile of F; wave (wr, wc) computes the 64x64 quadrant
// (wr, wc).  A operand lane l holds A[i = l%16][8 consecutive k], B the same with j; accumulator
// register v of lane l holds D[4*(l/16) + v][l%16].  The next chunk's two blocks are loaded into
// registers while the current chunk is multiplied from LDS.
// ------------------------------------------------------------------------------------------------
#define SCB 128  // column-block width of the mixed path
#define BP2 544  // LDS elements per k-group: 128 columns x 4 k + 32, so the four lane groups of a
                 // 64-bit operand read fall on alternating bank halves
__device__ __forceinline__ f16x8_t lds_k8(const uint16_t* arr, int g, int i)
{
    const uint2 a = *(const uint2*)&arr[g * BP2 + i * 4];
    const uint2 b = *(const uint2*)&arr[(g + 1) * BP2 + i * 4];
    return __builtin_bit_cast(f16x8_t, make_uint4(a.x, a.y, b.x, b.y));
}

__device__ __forceinline__ double item_lambda(const BatchDev& bd, int item) { return bd.lam ? bd.lam[item] : bd.lambda; }

// R entry (r, c), c <= r, of the banded curvature tiles (reference order, |r - c| within the band)
__device__ __forceinline__ double* rtile(const BatchDev& bd, const ItemMeta& M, int r, int c)
{
    const int I = r >> 6, J = c >> 6;
    return bd.R + M.r_off + ((size_t)I * (M.bwb + 1) + (I - J)) * TS * TS + (r & 63) * TS + (c & 63);
}
// the same entry of an item's R as read: R is built once per distinct R, by the item rrep names
__device__ __forceinline__ double rget(const BatchDev& bd, int item, int r, int c)
{
    return *rtile(bd, bd.meta[bd.rrep[item]], r, c);
}

// F's preparation, fused into the FP16 SYRK's store of tile (I, J): scale back, F += lambda R over
// the source block (R is held in the reference order: internal indices map through pisrc), the data
// row n kept with a large pivot (the refined solve never reads it), identity beyond n and the FP64
// diagonal into Fdiag.  A diagonal tile's upper half gets the same symmetric formula.  The tiles go to
// Ff32 (F as formed, which the refinement residual reads); the Cholesky kernels that touch a block
// first (the first part's diagonal and panel kernels, the first strip and trailing updates) read it
// there and write the factor buffer F32.  Wave (wr, wc) holds quadrant (wr, wc).
__device__ __forceinline__ void syrk_f16_store(const BatchDev& bd, const ItemMeta& M, int item, int I, int J,
                                               const float4_t (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
    const float unscale = (float)(1.0 / (bd.s16 * bd.s16));
    const int ld = M.ldN, n = M.n_src;
    const double lam = item_lambda(bd, item);
    const int* pis = bd.pisrc + item * bd.s_cell;
    double* fdiag = bd.vec + M.vec_off + 3 * (size_t)ld;
    float* C = bd.Ff32 + M.f_off + (size_t)I * SCB * ld + (size_t)J * SCB;
    const int r0 = I * SCB + wr * 64 + 4 * (lane >> 4), c0 = J * SCB + wc * 64 + (lane & 15);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int cg = c0 + c * 16;
        const int pb = cg < n ? pis[cg] : -1;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = r0 + a * 16 + v;
                float x = acc[a][c][v] * unscale;
                if (r < n && cg < n) {
                    const int pa = pis[r];
                    const int hi = pa > pb ? pa : pb, lo = pa > pb ? pb : pa;
                    if ((hi >> 6) - (lo >> 6) <= M.bwb) x = (float)((double)x + lam * rget(bd, item, hi, lo));
                    if (r == cg) fdiag[r] = (double)x;
                } else if (r > n || cg > n) {
                    x = (r == cg) ? 1.0f : 0.0f;
                } else if (r == cg) {
                    x = 1e30f;  // (n, n); the rest of row n (D) is kept
                }
                __builtin_nontemporal_store(x, &C[(size_t)(r - I * SCB) * ld + (cg - J * SCB)]);
            }
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 4))) k_syrk_f16_b(BatchDev bd)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t lds2[];  // A, B: 16 x BP2 each
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    int I, J;
    if (!syrk_tile(M.ncb, I, J)) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const size_t bsz = (size_t)TS * SCB;
    const uint16_t* pool = (const uint16_t*)bd.lb + M.lb_off;
    const unsigned* nz = bd.nz + M.nz_off;
    const int* bb = bd.bbase + M.bb_off;
    const int nkt = M.nkt, ncbw = M.ncbw;
    uint16_t* const as = lds2;
    uint16_t* const bs = lds2 + 16 * BP2;
    const bool dg = I == J;  // a diagonal tile reads its one block once, as both operands
    const uint16_t* const bsrc = dg ? as : bs;
    // thread piece h (of 4) of every block: 16 bytes at element (tid + 256 h) * 8
    const int e0 = tid * 8;
    const int o0 = (e0 >> 9) * BP2 + (e0 & 511);                    // pieces h = 0..3 are 2048
    const int o1 = ((e0 + 2048) >> 9) * BP2 + ((e0 + 2048) & 511);  // elements apart: 4 k-groups
    const int o2 = ((e0 + 4096) >> 9) * BP2 + ((e0 + 4096) & 511);
    const int o3 = ((e0 + 6144) >> 9) * BP2 + ((e0 + 6144) & 511);
    // the chunks holding both blocks as a bit mask in LDS (64 chunks a word, one ballot each), so
    // that finding the next one is a word test rather than a scan of the chunks' supports
    unsigned long long* const cmask = (unsigned long long*)(lds2 + 32 * BP2);
    const int nwd = (nkt + 63) >> 6;
    for (int q = w; q < nwd; q += 4) {
        const int k = 64 * q + lane;
        bool b = false;
        if (k < nkt) {
            const unsigned* m = nz + (size_t)k * ncbw;
            b = mask_has(m, I) && mask_has(m, J);
        }
        const unsigned long long bal = __ballot(b);
        if (lane == 0) cmask[q] = bal;
    }
    __syncthreads();
    // the next chunk after k holding both blocks (nkt if none)
    auto next = [&](int k) {
        ++k;
        for (int q = k >> 6; q < nwd; q++) {
            unsigned long long v = cmask[q];
            if (q == (k >> 6)) v &= ~0ULL << (k & 63);
            if (v) return 64 * q + __builtin_ctzll(v);
        }
        return nkt;
    };
#define SYRK_LOAD(kk_, A0, A1, A2, A3, B0, B1, B2, B3)                                                  \
    {                                                                                                    \
        const unsigned* m_ = nz + (size_t)(kk_) * ncbw;                                                  \
        const uint4* a_ = (const uint4*)(pool + (size_t)(bb[kk_] + mask_rank(m_, I)) * bsz) + tid;       \
        A0 = a_[0]; A1 = a_[256]; A2 = a_[512]; A3 = a_[768];                                            \
        if (!dg) {                                                                                       \
            const uint4* b_ = (const uint4*)(pool + (size_t)(bb[kk_] + mask_rank(m_, J)) * bsz) + tid;   \
            B0 = b_[0]; B1 = b_[256]; B2 = b_[512]; B3 = b_[768];                                        \
        }                                                                                                \
    }
#define SYRK_STAGE(A0, A1, A2, A3, B0, B1, B2, B3)                                                       \
    {                                                                                                    \
        __syncthreads();                                                                                 \
        *(uint4*)&as[o0] = A0; *(uint4*)&as[o1] = A1; *(uint4*)&as[o2] = A2; *(uint4*)&as[o3] = A3;      \
        if (!dg) { *(uint4*)&bs[o0] = B0; *(uint4*)&bs[o1] = B1; *(uint4*)&bs[o2] = B2; *(uint4*)&bs[o3] = B3; } \
        __syncthreads();                                                                                 \
    }
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = (float4_t){0, 0, 0, 0};
    auto mma = [&]() {
#pragma unroll
        for (int k0 = 0; k0 < TS; k0 += 32) {
            const int g = (k0 >> 2) + 2 * (lane >> 4);
            f16x8_t av[4], bv[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                av[t] = lds_k8(as, g, wr * 64 + t * 16 + (lane & 15));
                bv[t] = lds_k8(bsrc, g, wc * 64 + t * 16 + (lane & 15));
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[a], bv[c], acc[a][c], 0, 0, 0);
        }
    };
    // the next chunk's blocks are loaded into registers while this one is multiplied (loading two
    // chunks ahead needs 176 registers, two waves per SIMD, and took 23.3 ms against 17.2)
    uint4 pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3;
    int kt = next(-1);
    if (kt < nkt) SYRK_LOAD(kt, pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3);
    while (kt < nkt) {
        SYRK_STAGE(pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3);
        const int nk = next(kt);
        if (nk < nkt) SYRK_LOAD(nk, pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3);
        mma();
        kt = nk;
    }
#undef SYRK_LOAD
#undef SYRK_STAGE
    syrk_f16_store(bd, M, item, I, J, acc);
}

void launch_syrk_f16_b(hipStream_t s, const BatchDev& bd, int max_ncb, int max_nkt)
{
    const size_t lds = (size_t)2 * 16 * BP2 * sizeof(uint16_t) + (size_t)((max_nkt + 63) / 64) * 8;
    allow_big_lds((const void*)k_syrk_f16_b);
    hipLaunchKernelGGL(k_syrk_f16_b, dim3(syrk_grid(max_ncb), bd.B), dim3(256), lds, s, bd);
}

// ------------------------------------------------------------------------------------------------
// curvature R (exact small integers) in banded tiles.  In the active (j-major) order an R entry
// couples cells at most two source rows apart, |a - b| <= 2*snx, so the lower band in 64-blocks is
// bwb = (2*snx)/64 + 1 and the Cholesky factor stays inside it.
// ------------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_zero_r_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid || bd.rrep[item] != item) return;  // an earlier item holds this R
    // 16-byte stores: the band holds whole 64 x 64 tiles and starts on a tile boundary
    double2* R2 = (double2*)(bd.R + M.r_off);
    const size_t n2 = (size_t)(M.ldN / TS) * (M.bwb + 1) * TS * TS / 2;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
        R2[t] = make_double2(0.0, 0.0);
}

__global__ void __launch_bounds__(256) k_curvature_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || bd.rrep[item] != item) return;  // an earlier item holds this R
    if (c >= M.snx * M.sny) {
        // identity on the padding rows (the data row of F's augmented matrix has no R entries)
        const int r = c - M.snx * M.sny + M.n_src;
        if (r < M.ldN && c - M.snx * M.sny < M.ldN - M.n_src) *rtile(bd, M, r, r) = 1.0;
        return;
    }
    const int* sidx = bd.sidx + item * bd.s_cell;
    const int a = sidx[c];
    if (a < 0) return;
    const int snx = M.snx, sny = M.sny;
    const int i = c / sny, j = c % sny;
    for (int l = 0; l < 2; l++) {
        int ids[3];
        double h[3];
        int m = 0;
        ids[m] = a;
        h[m++] = -2.0;
        int nb0, nb1;
        if (l == 0) {
            nb0 = (i < snx - 1) ? sidx[(i + 1) * sny + j] : -1;
            nb1 = (i > 0) ? sidx[(i - 1) * sny + j] : -1;
        } else {
            nb0 = (j < sny - 1) ? sidx[i * sny + j + 1] : -1;
            nb1 = (j > 0) ? sidx[i * sny + j - 1] : -1;
        }
        if (nb0 >= 0) { ids[m] = nb0; h[m++] = 1.0; }
        if (nb1 >= 0) { ids[m] = nb1; h[m++] = 1.0; }
        for (int x = 0; x < m; x++)
            for (int y = x; y < m; y++) {
                const int r0 = ids[x] > ids[y] ? ids[x] : ids[y];
                const int c0 = ids[x] > ids[y] ? ids[y] : ids[x];
                atomicAdd(rtile(bd, M, r0, c0), h[x] * h[y]);
            }
    }
}

// R is a function of the grid's shape (snx, sny) and the item's active-index map (sidx): its entries
// are sums of small integers, exact in any order.  Items with equal shapes and maps share one
// factorization: k_rhash_b keys each item (an order-free sum of mixed (cell, index) words),
// k_rdup_b compares an item with the earlier items of equal key, entry by entry, and takes the first
// equal one as its representative (whose own representative is itself).
__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

__global__ void __launch_bounds__(256) k_rhash_b(BatchDev bd)
{
    __shared__ unsigned long long sh[256];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int tid = threadIdx.x;
    unsigned long long h = 0;
    if (M.valid) {
        const int* sidx = bd.sidx + item * bd.s_cell;
        const int n = M.snx * M.sny;
        for (int c = tid; c < n; c += 256) h += mix64(((unsigned long long)c << 32) ^ (unsigned)(sidx[c] + 1));
    }
    sh[tid] = h;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) sh[tid] += sh[tid + st];
        __syncthreads();
    }
    if (tid == 0) bd.rkey[item] = M.valid ? sh[0] ^ mix64(((unsigned long long)M.snx << 32) | (unsigned)M.sny) : 0ULL;
}

__global__ void __launch_bounds__(256) k_rdup_b(BatchDev bd)
{
    __shared__ int eq;
    const int b = blockIdx.y;
    const ItemMeta& M = bd.meta[b];
    int rep = b;
    if (M.valid) {
        const int n = M.snx * M.sny;
        const int* sb = bd.sidx + b * bd.s_cell;
        for (int a = 0; a < b; a++) {
            const ItemMeta& A = bd.meta[a];
            if (!A.valid || bd.rkey[a] != bd.rkey[b] || A.snx != M.snx || A.sny != M.sny || A.n_src != M.n_src) continue;
            const int* sa = bd.sidx + a * bd.s_cell;
            if (threadIdx.x == 0) eq = 1;
            __syncthreads();
            for (int c = threadIdx.x; c < n; c += 256)
                if (sa[c] != sb[c]) eq = 0;
            __syncthreads();
            const int same = eq;
            __syncthreads();
            if (same) {
                rep = a;
                break;
            }
        }
    }
    if (threadIdx.x == 0) bd.rrep[b] = rep;
}

void launch_rdup_b(hipStream_t s, const BatchDev& bd)
{
    hipLaunchKernelGGL(k_rhash_b, dim3(1, bd.B), dim3(256), 0, s, bd);
    hipLaunchKernelGGL(k_rdup_b, dim3(1, bd.B), dim3(256), 0, s, bd);
}

void launch_curvature_b(hipStream_t s, const BatchDev& bd, int max_cells, int max_rtiles)
{
    hipLaunchKernelGGL(k_zero_r_b, dim3(grid_for((size_t)max_rtiles * TS * TS, 256, 2048), bd.B), dim3(256), 0, s, bd);
    const int n = max_cells + QLX_MAX_LD;
    hipLaunchKernelGGL(k_curvature_b, dim3((n + 255) / 256, bd.B), dim3(256), 0, s, bd);
}

// FP64 path: F_lower += lambda R over the n x n source block; row n (the data row D^T | d^T C d of
// the augmented factorization) is kept, rows > n are identity.  (The mixed path does this in the
// FP16 SYRK's epilogue, syrk_f16_store.)
__global__ void __launch_bounds__(256) k_prep_f_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const int ld = M.ldN, n = M.n_src, nb = ld / TS;
    if ((int)blockIdx.x >= nb * (nb + 1) / 2) return;
    int I, J;
    tri_decode(blockIdx.x, I, J);
    double* F = bd.F + M.f_off;
    const double* F0 = bd.Fraw ? bd.Fraw + M.f_off : F;  // out of place when F as formed is kept
    const int* pis = bd.pisrc + item * bd.s_cell;
    const double lam = item_lambda(bd, item);
    // 16 elements per thread, unrolled so that their gathers (pisrc, R tile, F) are in flight together
#pragma unroll 8
    for (int e = threadIdx.x; e < TS * TS; e += 256) {
        const int rr = e >> 6, cc = e & 63;
        const int r = I * TS + rr, c = J * TS + cc;
        if (c <= r) {
            const size_t t = (size_t)r * ld + c;
            if (r < n) {
                // R is held in the reference order: F's internal indices map through pisrc
                const int a = pis[r], b = pis[c];
                const int hi = a > b ? a : b, lo = a > b ? b : a;
                const double rv = ((hi >> 6) - (lo >> 6) <= M.bwb) ? rget(bd, item, hi, lo) : 0.0;
                F[t] = F0[t] + lam * rv;
            } else if (r > n) {
                F[t] = (r == c) ? 1.0 : 0.0;
            } else {
                F[t] = F0[t];  // the data row
            }
        }
    }
}

void launch_prep_fr_b(hipStream_t s, const BatchDev& bd, int max_ld, int f32)
{
    const int nb = max_ld / TS;
    if (f32) return;  // the mixed path's FP16 SYRK prepares F in its epilogue
    const unsigned g = nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_prep_f_b, dim3(g, bd.B), dim3(256), 0, s, bd);
}

// ------------------------------------------------------------------------------------------------
// Batched blocked Cholesky (lower, 64x64 tiles) of many matrices, FP64 or FP32 per matrix, dense
// or banded per matrix: per block step a diagonal kernel (factor + block inverse), a panel kernel
// (A_ib,kb <- A_ib,kb W^T on MFMA) and a trailing-update kernel (A_IJ -= P_I P_J^T on MFMA).
// Matrices with fewer blocks skip the step; the upper triangle is never written.
// Restates Cholesky_dcmp_packed (pixelgrid.cpp:17071-17105) blockwise: a pivot of a row < n that
// is not positive (negative, zero or NaN) sets *status and the factorization continues with
// sqrt(|a|) as that routine does; the item is then reported as not positive definite (the
// reference's non-MKL caller dies there, pixelgrid.cpp:17384).
// ------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T* tile_ptr(const MatDesc& D, int I, int J, int& ld)
{
    if (D.bwb > 0) {
        ld = TS;
        return (T*)D.A + ((size_t)I * (D.bwb + 1) + (I - J)) * TS * TS;
    }
    ld = D.ld;
    return (T*)D.A + (size_t)I * TS * D.ld + (size_t)J * TS;
}

__device__ __forceinline__ int band_last(const MatDesc& D, int kb)
{
    return D.bwb > 0 ? min(D.nb - 1, kb + D.bwb) : D.nb - 1;
}

// MFMA operand/accumulator types: FP64 v_mfma_f64_16x16x4f64 (accumulator register v of lane l
// holds D[l/16 + 4v][l%16]) and FP32 v_mfma_f32_16x16x4f32 (D[4(l/16) + v][l%16])
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double4_t acc_t;
    typedef double2 vec_t;
    __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma16<float> {
    typedef float4_t acc_t;
    typedef float4 vec_t;
    __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// value of lane `src` in every lane (src is a constant of the unrolled loops that call this)
__device__ __forceinline__ float lane_bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// 64x64x64 MFMA product acc = sum_q A(i, q) B(j, q) with A and B stored "row holds the k-vector"
// (the panel blocks as they lie in the factor), staged into LDS unchanged ([i][k], row pitch
// LDT) and read as MFMA operands A[i = lane%16][k = kk + lane/16].
#define LDT 68
template <typename T>
__device__ __forceinline__ void stage64(const T* P, int sp, T (*S)[LDT])
{
    typedef typename Mfma16<T>::vec_t V;
    constexpr int VW = sizeof(V) / sizeof(T);
    for (int e = threadIdx.x * VW; e < TS * TS; e += 256 * VW) {
        const int i = e >> 6, k = e & 63;
        *(V*)&S[i][k] = *(const V*)&P[(size_t)i * sp + k];
    }
}

template <typename T>
__device__ __forceinline__ void mma64_add(T (*As)[LDT], T (*Bs)[LDT], typename Mfma16<T>::acc_t acc[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll 4
    for (int kk = 0; kk < TS; kk += 4) {
        const int kq = kk + (lane >> 4);
        T a0 = As[wr * 32 + (lane & 15)][kq];
        T a1 = As[wr * 32 + 16 + (lane & 15)][kq];
        T b0 = Bs[wc * 32 + (lane & 15)][kq];
        T b1 = Bs[wc * 32 + 16 + (lane & 15)][kq];
        acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
        acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
        acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
        acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
    }
}

template <typename T>
__device__ __forceinline__ void mma64(T (*As)[LDT], T (*Bs)[LDT], typename Mfma16<T>::acc_t acc[2][2])
{
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (typename Mfma16<T>::acc_t){0, 0, 0, 0};
    mma64_add<T>(As, Bs, acc);
}

// C (64x64 at row pitch ldc) = acc (mode 0), C -= acc (mode 1), lower triangle only (mode 2)
template <typename T>
__device__ __forceinline__ void store64(T* C, int ldc, const typename Mfma16<T>::acc_t acc[2][2], int mode)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 32 + a * 16 + Mfma16<T>::row(lane, v);
                const int cc = wc * 32 + c * 16 + (lane & 15);
                T* d = C + (size_t)rr * ldc + cc;
                if (mode == 0) *d = acc[a][c][v];
                else if (mode == 1 || cc <= rr) *d -= acc[a][c][v];
            }
}

// the 16 C values of a lane (store64's layout), loaded before the staging barrier so their latency
// overlaps the operand loads; store64c then writes C - acc (mode 1) or its lower triangle (mode 2)
template <typename T>
__device__ __forceinline__ void load64c(const T* C, int ldc, T (&cv)[16])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 32 + a * 16 + Mfma16<T>::row(lane, v);
                const int cc = wc * 32 + c * 16 + (lane & 15);
                cv[(a * 2 + c) * 4 + v] = C[(size_t)rr * ldc + cc];
            }
}

template <typename T>
__device__ __forceinline__ void store64c(T* C, int ldc, const typename Mfma16<T>::acc_t acc[2][2], const T (&cv)[16], int mode)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 32 + a * 16 + Mfma16<T>::row(lane, v);
                const int cc = wc * 32 + c * 16 + (lane & 15);
                if (mode == 1 || cc <= rr) C[(size_t)rr * ldc + cc] = cv[(a * 2 + c) * 4 + v] - acc[a][c][v];
            }
}

#define MM_LDS (2 * TS * LDT)
template <typename T>
__device__ void chol_panel_body(const MatDesc& D, int kb, int ib, double* raw)
{
    T(*As)[LDT] = (T(*)[LDT])raw;
    T(*Bs)[LDT] = (T(*)[LDT])((T*)raw + TS * LDT);
    int ld;
    T* blk = tile_ptr<T>(D, ib, kb, ld);
    const T* W = (const T*)D.W + (size_t)kb * TS * TS;
    stage64<T>(blk, ld, As);
    stage64<T>(W, TS, Bs);
    __syncthreads();
    typename Mfma16<T>::acc_t acc[2][2];
    mma64<T>(As, Bs, acc);  // C(i,j) = sum_q A_ib[i][q] W[j][q]
    store64<T>(blk, ld, acc, 0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_chol_panel_m(const MatDesc* mats, int kb)
{
    __shared__ T raw[MM_LDS];
    const MatDesc D = mats[blockIdx.y];
    if (kb >= D.nb) return;
    const int ib = kb + 1 + blockIdx.x;
    if (ib > band_last(D, kb)) return;
    chol_panel_body<T>(D, kb, ib, (double*)raw);
}

// trailing update A_IJ -= P_I P_J^T for one element type (matrices of one launch share it)
template <typename T>
__global__ void __launch_bounds__(256) k_chol_update_m(const MatDesc* mats, int kb)
{
    __shared__ T As[TS][LDT];
    __shared__ T Bs[TS][LDT];
    const MatDesc D = mats[blockIdx.y];
    if (kb >= D.nb) return;
    const int rest = band_last(D, kb) - kb;
    int I, J;
    if (!syrk_tile(rest, I, J)) return;  // XCD-aware super-tile order: neighbours share panel blocks in L2
    I += kb + 1;
    J += kb + 1;
    int lda, ldb, ldc;
    const T* Pa = tile_ptr<T>(D, I, kb, lda);
    const T* Pb = tile_ptr<T>(D, J, kb, ldb);
    T* C = tile_ptr<T>(D, I, J, ldc);
    T cv[16];
    load64c<T>(C, ldc, cv);
    stage64<T>(Pa, lda, As);
    stage64<T>(Pb, ldb, Bs);
    __syncthreads();
    typename Mfma16<T>::acc_t acc[2][2];
    mma64<T>(As, Bs, acc);
    store64c<T>(C, ldc, acc, cv, I == J ? 2 : 1);
}

// rank-64 update from column block kb of the tiles (I, J), kb < J < kend, J <= I < nb: the
// columns of the current 256-wide panel that are still to be factored (dense matrices only)
template <typename T>
__global__ void __launch_bounds__(256) k_chol_upanel_m(const MatDesc* mats, int kb, int kend)
{
    __shared__ T As[TS][LDT];
    __shared__ T Bs[TS][LDT];
    const MatDesc D = mats[blockIdx.y];
    if (kb >= D.nb) return;
    const int je = min(kend, D.nb);
    int t = blockIdx.x, I = -1, J = -1;
    for (int j = kb + 1; j < je; j++) {
        const int cnt = D.nb - j;
        if (t < cnt) {
            J = j;
            I = j + t;
            break;
        }
        t -= cnt;
    }
    if (J < 0) return;
    int lda, ldb, ldc;
    const T* Pa = tile_ptr<T>(D, I, kb, lda);
    const T* Pb = tile_ptr<T>(D, J, kb, ldb);
    T* C = tile_ptr<T>(D, I, J, ldc);
    T cv[16];
    load64c<T>(C, ldc, cv);
    stage64<T>(Pa, lda, As);
    stage64<T>(Pb, ldb, Bs);
    __syncthreads();
    typename Mfma16<T>::acc_t acc[2][2];
    mma64<T>(As, Bs, acc);
    store64c<T>(C, ldc, acc, cv, I == J ? 2 : 1);
}

// Left-looking step of a 256-wide panel (columns k0 .. kend-1, dense matrices): column block kb.
// Workgroup x takes row block ib = kb + 1 + x and forms
//   A_ib,kb - sum_{k0 <= k' < kb} L_ib,k' L_kb,k'^T  (the earlier columns of the panel, summed in
//   the MFMA accumulators while the block waits in registers),  L_ib,kb = that times W_kb^T,
// so each block is read and written once per panel column.  The workgroup of ib = kb + 1 then
// brings the next diagonal block up to date for the diagonal kernel, when it lies in the panel:
//   A_ib,ib -= sum_{k0 <= k' <= kb} L_ib,k' L_ib,k'^T  (lower half stored).
template <typename T>
__global__ void __launch_bounds__(256) k_chol_lpanel_m(const MatDesc* mats, int kb, int k0, int kend)
{
    __shared__ T As[TS][LDT];
    __shared__ T Bs[TS][LDT];
    typedef typename Mfma16<T>::acc_t Acc;
    const MatDesc D = mats[blockIdx.y];
    if (kb >= D.nb) return;
    const int ib = kb + 1 + blockIdx.x;
    if (ib > band_last(D, kb)) return;
    int ld, lp;
    T* blk = tile_ptr<T>(D, ib, kb, ld);
    // the part that starts at column 0 reads its blocks from F as formed (A0), when the matrix has
    // it: each is read there once and written to A
    const bool first = (k0 == 0 && D.A0);
    T cv[16];
    load64c<T>(first ? (const T*)D.A0 + (blk - (T*)D.A) : blk, ld, cv);
    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (Acc){0, 0, 0, 0};
    for (int kp = k0; kp < kb; kp++) {
        const T* Pa = tile_ptr<T>(D, ib, kp, lp);
        const T* Pb = tile_ptr<T>(D, kb, kp, lp);
        __syncthreads();
        stage64<T>(Pa, lp, As);
        stage64<T>(Pb, lp, Bs);
        __syncthreads();
        mma64_add<T>(As, Bs, acc);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[a][c][v] = cv[(a * 2 + c) * 4 + v] - acc[a][c][v];
    __syncthreads();
    store64<T>(&As[0][0], LDT, acc, 0);
    stage64<T>((const T*)D.W + (size_t)kb * TS * TS, TS, Bs);
    __syncthreads();
    mma64<T>(As, Bs, acc);  // L_ib,kb(i, j) = sum_q A'(i, q) W_kb(j, q)
    store64<T>(blk, ld, acc, 0);
    if (ib != kb + 1 || ib >= kend) return;
    T* dg = tile_ptr<T>(D, ib, ib, ld);
    load64c<T>(first ? (const T*)D.A0 + (dg - (T*)D.A) : dg, ld, cv);
    __syncthreads();
    store64<T>(&As[0][0], LDT, acc, 0);
    __syncthreads();
    mma64<T>(As, As, acc);
    for (int kp = k0; kp < kb; kp++) {
        const T* Pa = tile_ptr<T>(D, ib, kp, lp);
        __syncthreads();
        stage64<T>(Pa, lp, As);
        __syncthreads();
        mma64_add<T>(As, As, acc);
    }
    store64c<T>(dg, ld, acc, cv, 2);
}

// acc += As(rows of quadrant wr) * Bs(rows of quadrant wc)^T over the 64 staged columns.  Lane group
// g = lane/16 reads VW consecutive k of its row with one LDS vector read; MFMA s of the group takes
// component s, so the VW MFMAs of a step sum k = kk + VW*g + s over all 4*VW values of the step.
template <typename T>
__device__ __forceinline__ void trail_mma(T (*As)[LDT], T (*Bs)[LDT], typename Mfma16<T>::acc_t (&acc)[4][4])
{
    typedef typename Mfma16<T>::vec_t V;
    constexpr int VW = sizeof(V) / sizeof(T);
    typedef T P __attribute__((ext_vector_type(VW)));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll 1
    for (int kk = 0; kk < TS; kk += 4 * VW) {
        const int kq = kk + VW * (lane >> 4);
        P av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) av[a] = *(const P*)&As[wr * 64 + a * 16 + (lane & 15)][kq];
#pragma unroll
        for (int c = 0; c < 4; c++) bv[c] = *(const P*)&Bs[wc * 64 + c * 16 + (lane & 15)][kq];
#pragma unroll
        for (int s2 = 0; s2 < VW; s2++)
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[a][c] = Mfma16<T>::op(av[a][s2], bv[c][s2], acc[a][c]);
    }
}

// Trailing update after a 256-wide panel: C(I, J) -= sum over the panel's nq column blocks of
// L(I, q) L(J, q)^T, on 128 x 128 tiles of the trailing matrix (first row block k0 + nq, XCD-aware
// tile order).  Wave (wr, wc) owns the 64 x 64 quadrant (wr, wc): 4 x 4 MFMA tiles of 16 x 16.
#define TT 128
// Two 128 x 68 staging tiles per WG (70 KB in FP32) allow two WGs per CU, so two waves per SIMD
// is all the occupancy this kernel gets: the register budget is sized for that.
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 1 : 2, sizeof(T) == 8 ? 1 : 2)))
k_chol_trail_m(const MatDesc* mats, int k0, int nq)
{
    extern __shared__ double trail_raw[];
    T(*As)[LDT] = (T(*)[LDT])trail_raw;
    T(*Bs)[LDT] = (T(*)[LDT])((T*)trail_raw + TT * LDT);
    const MatDesc D = mats[blockIdx.y];
    const int r0b = k0 + nq;
    if (r0b >= D.nb) return;
    const int nt = (D.nb - r0b) / 2;
    int I2, J2;
    if (!syrk_tile(nt, I2, J2)) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const long long ld = D.ld;
    T* A = (T*)D.A;
    const long long rI = (long long)(r0b + 2 * I2) * TS, rJ = (long long)(r0b + 2 * J2) * TS;
    typename Mfma16<T>::acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = (typename Mfma16<T>::acc_t){0, 0, 0, 0};
    typedef typename Mfma16<T>::vec_t V;
    constexpr int VW = sizeof(V) / sizeof(T);
    constexpr int NPF = (TT * TS) / (256 * VW);  // pieces per thread and matrix
    typedef T P __attribute__((ext_vector_type(VW)));
    P pa[NPF], pb[NPF];
    T cv[64];  // the C tile values of this lane (loaded during the last block)
    // the next column block is loaded into registers while the current one is multiplied (FP64: one
    // workgroup per CU for its 139 KB of LDS, so the registers are there)
#define TRAIL_LOAD(q)                                                    \
    {                                                                    \
        const long long c0 = (long long)(k0 + (q)) * TS;                 \
        _Pragma("unroll") for (int h = 0; h < NPF; h++)                  \
        {                                                                \
            const int e = (tid + 256 * h) * VW;                          \
            const int i = e >> 6, k = e & 63;                            \
            pa[h] = *(const P*)&A[(rI + i) * ld + c0 + k];               \
            pb[h] = *(const P*)&A[(rJ + i) * ld + c0 + k];               \
        }                                                                \
    }
    {
        TRAIL_LOAD(0);
        for (int q = 0; q < nq; q++) {
            __syncthreads();
#pragma unroll
            for (int h = 0; h < NPF; h++) {
                const int e = (tid + 256 * h) * VW;
                const int i = e >> 6, k = e & 63;
                *(P*)&As[i][k] = pa[h];
                *(P*)&Bs[i][k] = pb[h];
            }
            __syncthreads();
            if (q + 1 == nq) break;  // the last block is multiplied below, with the C tile loads
            TRAIL_LOAD(q + 1);
            trail_mma<T>(As, Bs, acc);
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = wr * 64 + a * 16 + Mfma16<T>::row(lane, v);
                    const int cc = wc * 64 + c * 16 + (lane & 15);
                    cv[(a * 4 + c) * 4 + v] = A[(rI + rr) * ld + rJ + cc];
                }
        trail_mma<T>(As, Bs, acc);
    }
    const bool diag = (I2 == J2);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 64 + a * 16 + Mfma16<T>::row(lane, v);
                const int cc = wc * 64 + c * 16 + (lane & 15);
                T* d = &A[(rI + rr) * ld + rJ + cc];
                const T old = cv[(a * 4 + c) * 4 + v];
                if (!diag || cc <= rr) *d = old - acc[a][c][v];
            }
}
#undef TRAIL_LOAD

// FP32 matrices (mixed path): the same trailing update with the panel operands split x = hi + lo
// into BF16 while they are staged (hi = bf16(x), lo = bf16(x - hi): 2^-17 of each entry, with the
// hardware round-to-nearest conversion) and all four products hi*hi + hi*lo + lo*hi + lo*lo on
// v_mfma_f32_16x16x32_bf16, accumulated in FP32.  The accumulators start from the C tile and the A
// operand is negated while staging, so the MFMAs form C - A B^T directly.  Loads go through a buffer
// resource (32-bit offsets: one VGPR per lane, row offsets in SGPRs).  Operands are staged 32 k at
// a time (LDS: A hi, A lo, B hi, B lo as 128 rows x TBP BF16, 40 KB; 80-byte rows put the 16 rows
// of an operand read on distinct bank groups), which with 150 registers allows three waves per SIMD.
// The update covers the 128 x 128 tiles of rows and columns >= k0 + nq (64-blocks) from the nq panel
// blocks k0 .. k0+nq-1; ncolt > 0 restricts it to the first ncolt tile columns (a strip).
#define TBP 40
static_assert((long long)QLX_MAX_LD * QLX_MAX_LD * 4 < (1LL << 31), "trailing-update byte offsets must fit in 31 bits");
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void chol_diag_f32_4w(const MatDesc& D, int kb, float* tile, int ld, float (&a)[16],
                                                 float (*Lt)[TS], float (*X)[TS], float* Rd);

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_chol_trail_bf16(const MatDesc* mats, int k0, int nq, int ncolt, int fdiag)
{
    typedef float P __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) float gfloat;
    extern __shared__ double trail_raw[];
    __bf16* Ah = (__bf16*)trail_raw;
    __bf16* Al = Ah + TT * TBP;
    __bf16* Bh = Al + TT * TBP;
    __bf16* Bl = Bh + TT * TBP;
    const MatDesc D = mats[blockIdx.y];
    const int r0b = k0 + nq;
    if (r0b >= D.nb) return;
    const int nt = (D.nb - r0b) / 2;
    int I2 = -1, J2 = -1;
    if (ncolt == 0) {
        if (!syrk_tile(nt, I2, J2)) return;
    } else {
        int t = blockIdx.x;
        for (int j = 0; j < ncolt && j < nt; j++) {
            if (t < nt - j) {
                J2 = j;
                I2 = j + t;
                break;
            }
            t -= nt - j;
        }
        if (J2 < 0) return;
    }
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const int ld = D.ld;  // ld <= QLX_MAX_LD: byte offsets within the matrix fit in 31 bits
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(D.A, 0, ld * ld * 4, 0x00020000);
    // C tiles: from F as formed (A0) in the first update, which thereby moves them into A
    const __amdgpu_buffer_rsrc_t rc = (k0 == 0 && D.A0) ? __builtin_amdgcn_make_buffer_rsrc(D.A0, 0, ld * ld * 4, 0x00020000) : rs;
    const int rI = (r0b + 2 * I2) * TS, rJ = (r0b + 2 * J2) * TS;
    // byte offsets: the lane's part in a VGPR, the uniform part in an SGPR.  Staging step st covers
    // columns 32 st .. 32 st + 31 of the panel; piece h of a thread: row srow + 32 h, columns
    // 4 (tid & 7) .. +3 of the step.  The 16 lanes of each group of tid stage rows r and r + 4,
    // whose 64-byte LDS pieces (row pitch 80 bytes) fall on disjoint banks
    const int g16 = tid >> 4;
    const int srow = (g16 & 3) + 8 * (g16 >> 2) + 4 * ((tid >> 3) & 1);
    const int vp = (srow * ld + 4 * (tid & 7)) * 4;
    const int sA = (rI * ld + k0 * TS) * 4, sB = (rJ * ld + k0 * TS) * 4;
    const int vc = ((wr * 64 + 4 * (lane >> 4)) * ld + wc * 64 + (lane & 15)) * 4;
    const int sC = (rI * ld + rJ) * 4;
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                acc[a][c][v] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rc, vc + c * 64, sC + (a * 16 + v) * ld * 4, 0));
    P pa[4], pb[4];
#define TRAIL_LOAD_BF(st)                                                                                          \
    {                                                                                                              \
        _Pragma("unroll") for (int h = 0; h < 4; h++)                                                              \
        {                                                                                                          \
            pa[h] = __builtin_bit_cast(P, __builtin_amdgcn_raw_buffer_load_b128(rs, vp, sA + (32 * h * ld + (st) * 32) * 4, 0)); \
            pb[h] = __builtin_bit_cast(P, __builtin_amdgcn_raw_buffer_load_b128(rs, vp, sB + (32 * h * ld + (st) * 32) * 4, 0)); \
        }                                                                                                          \
    }
    TRAIL_LOAD_BF(0);
    const int nst = 2 * nq;
    for (int st = 0; st < nst; st++) {
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 4; h++) {
            const int i = srow + 32 * h, k = 4 * (tid & 7);
            const P xa = -pa[h], xb = pb[h];
            const bf16x4_t ah = __builtin_convertvector(xa, bf16x4_t);
            const bf16x4_t bh = __builtin_convertvector(xb, bf16x4_t);
            const bf16x4_t al = __builtin_convertvector(xa - __builtin_convertvector(ah, P), bf16x4_t);
            const bf16x4_t bl = __builtin_convertvector(xb - __builtin_convertvector(bh, P), bf16x4_t);
            *(bf16x4_t*)&Ah[i * TBP + k] = ah;
            *(bf16x4_t*)&Al[i * TBP + k] = al;
            *(bf16x4_t*)&Bh[i * TBP + k] = bh;
            *(bf16x4_t*)&Bl[i * TBP + k] = bl;
        }
        __syncthreads();
        if (st + 1 < nst) TRAIL_LOAD_BF(st + 1);
        const int kq = 8 * (lane >> 4);
        bf16x8_t bhv[4], blv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int jb = (wc * 64 + t * 16 + (lane & 15)) * TBP + kq;
            bhv[t] = *(const bf16x8_t*)&Bh[jb];
            blv[t] = *(const bf16x8_t*)&Bl[jb];
        }
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int ia = (wr * 64 + a * 16 + (lane & 15)) * TBP + kq;
            const bf16x8_t ahv = *(const bf16x8_t*)&Ah[ia];
            const bf16x8_t alv = *(const bf16x8_t*)&Al[ia];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bhv[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, blv[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alv, bhv[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alv, blv[c], acc[a][c], 0, 0, 0);
            }
        }
    }
#undef TRAIL_LOAD_BF
    // plain global stores (the buffer-store form stored one register for all four rows of a
    // 16 x 16 block under the diagonal predicate with this compiler)
    const bool diag = (I2 == J2);
    gfloat* Cg = (gfloat*)D.A + (size_t)rI * ld + rJ;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 64 + a * 16 + 4 * (lane >> 4) + v;
                const int cc = wc * 64 + c * 16 + (lane & 15);
                if (!diag || cc <= rr) Cg[(size_t)rr * ld + cc] = acc[a][c][v];
            }
    // fdiag: the workgroup of tile (0, 0) also factors diagonal block r0b, now up to date (wave 0's
    // quadrant), so the part that starts there needs no diagonal launch; through LDS (S, 64 x 65,
    // over the operand tiles) into rows per lane, as the panel kernel does
    if (fdiag && I2 == 0 && J2 == 0) {
        float* S = (float*)trail_raw;
        float(*Lt)[TS] = (float(*)[TS])(S + TS * 65);
        float* Rd = S + TS * 65 + TS * TS;
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 4; c++)
#pragma unroll
                    for (int v = 0; v < 4; v++) S[(a * 16 + 4 * (lane >> 4) + v) * 65 + c * 16 + (lane & 15)] = acc[a][c][v];
        }
        __syncthreads();
        float av[16];
#pragma unroll
        for (int j = 0; j < 16; j++) av[j] = S[lane * 65 + 16 * w + j];
        chol_diag_f32_4w(D, r0b, (float*)D.A + (size_t)rI * ld + rJ, ld, av, Lt, (float(*)[TS])S, Rd);  // X over S: read before its first barrier
    }
}

// The left-looking panel step for FP32 matrices (k_chol_lpanel_m's schedule and operands) with the
// products on v_mfma_f32_16x16x32_bf16: every operand is split x = hi + lo into BF16 as it is
// staged (as in k_chol_trail_bf16) and the four products hi*hi + hi*lo + lo*hi + lo*lo are summed
// in FP32 (1.9 ms faster than FP32 MFMA per batch of 256 3712 x 3712 matrices; |delta lnE| against
// FP64 unchanged).  The workgroup of row block kb + 1, having brought diagonal block kb + 1 up to
// date, also factors it (chol_diag_f32_4w: the same bits as the diagonal kernel), so within a part
// only the first column block needs a diagonal launch; the grid is (matrices, row blocks) so that
// these workgroups are dispatched first (dispatched last, their serial factorization prolonged
// every launch).  LDS: A hi, A lo, B hi, B lo, 64 rows x LPB BF16 each (144-byte rows: the 16 rows
// of an operand read start on distinct 16-byte bank groups).
#define LPB 72
typedef float f32x4_t __attribute__((ext_vector_type(4)));
// a 64 x 64 FP32 block (row pitch sp) into registers: thread piece h holds row e >> 6, columns
// e & 63 .. +3 of e = (tid + 256 h) * 4
__device__ __forceinline__ void load64_regs(const float* P, int sp, f32x4_t (&r)[4])
{
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int e = (threadIdx.x + 256 * h) * 4;
        r[h] = *(const f32x4_t*)&P[(size_t)(e >> 6) * sp + (e & 63)];
    }
}
// those registers split x = hi + lo into BF16 rows of an operand
__device__ __forceinline__ void split64_store(const f32x4_t (&r)[4], __bf16* H, __bf16* L)
{
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int e = (threadIdx.x + 256 * h) * 4;
        const int i = e >> 6, k = e & 63;
        const bf16x4_t xh = __builtin_convertvector(r[h], bf16x4_t);
        *(bf16x4_t*)&H[i * LPB + k] = xh;
        *(bf16x4_t*)&L[i * LPB + k] = __builtin_convertvector(r[h] - __builtin_convertvector(xh, f32x4_t), bf16x4_t);
    }
}
__device__ __forceinline__ void stage64_split(const float* P, int sp, __bf16* H, __bf16* L)
{
    f32x4_t r[4];
    load64_regs(P, sp, r);
    split64_store(r, H, L);
}

// an accumulator tile (mma64's layout: wave (wr, wc) holds quadrant rows wr*32.., columns wc*32..)
// split into H / L as the rows of an operand
__device__ __forceinline__ void acc64_split(const float4_t (&acc)[2][2], __bf16* H, __bf16* L)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rr = wr * 32 + a * 16 + 4 * (lane >> 4) + v;
                const int cc = wc * 32 + c * 16 + (lane & 15);
                const float x = acc[a][c][v];
                const __bf16 xh = (__bf16)x;
                H[rr * LPB + cc] = xh;
                L[rr * LPB + cc] = (__bf16)(x - (float)xh);
            }
}

// acc(i, j) += sum_q A(i, q) B(j, q) over the 64 staged q, A = Ah + Al, B = Bh + Bl
__device__ __forceinline__ void mma64_bf16(const __bf16* Ah, const __bf16* Al, const __bf16* Bh, const __bf16* Bl,
                                           float4_t (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
#pragma unroll
    for (int ks = 0; ks < TS; ks += 32) {
        const int kq = ks + 8 * (lane >> 4);
        bf16x8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int ia = (wr * 32 + t * 16 + (lane & 15)) * LPB + kq;
            const int ib = (wc * 32 + t * 16 + (lane & 15)) * LPB + kq;
            ah[t] = *(const bf16x8_t*)&Ah[ia];
            al[t] = *(const bf16x8_t*)&Al[ia];
            bh[t] = *(const bf16x8_t*)&Bh[ib];
            bl[t] = *(const bf16x8_t*)&Bl[ib];
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bl[c], acc[a][c], 0, 0, 0);
            }
    }
}

// The FP32 diagonal block factored by the four waves of a workgroup (one wave per 16-column panel;
// the earlier one-wave form held 256 registers and its 64 column steps ran serially): lane r of wave
// w holds a[j] = A(r, 16 w + j).  d = |p| rsqrt|p| and l = a rsqrt|p| for the pivot p.  Column c: its pivot is lane c's entry (readlane), the
// column goes to Lt[c] (Lt[c][i] = L(i, c)) and 1/sqrt of the pivot to Rd[c], and every entry right
// of c takes the rank-1 update.  L^-1 by the right-looking substitution, lane r of wave w holding
// rows 16 w .. 16 w + 15 of column r (row q scaled by Rd[q], published in X[q], then the rows below
// updated).  Both blocked by 16-column (row) panels, one per wave, one barrier per panel: wave p
// does its own panel's steps, the waves after it then apply the panel's 16 updates in order.  Writes
// L's lower triangle into the block (row pitch ld), the pivots and W.  LDS: Lt and X 64 x 64 each,
// Rd 64.
__device__ __forceinline__ void chol_diag_f32_4w(const MatDesc& D, int kb, float* tile, int ld, float (&a)[16],
                                                 float (*Lt)[TS], float (*X)[TS], float* Rd)
{
    typedef float T;
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long base = (long long)kb * TS;
    bool bad = false;
    // factor, by 16-column panels: wave p factors its own panel (its columns' pivots, the columns into
    // Lt, the rank-1 updates of its later columns), then the waves right of it apply the panel's 16
    // updates in column order: each element sees the updates of columns 0, 1, ... in turn
#pragma unroll
    for (int p = 0; p < 4; p++) {
        if (w == p) {
#pragma unroll
            for (int cc = 0; cc < 16; cc++) {
                const int c = 16 * p + cc;
                const T piv = lane_bcast(a[cc], c);
                bad |= (!(piv > 0) && base + c < D.n);  // negative, zero or NaN pivot
                const T pa = fabs(piv);
                const T rd = rsqrtf(pa);
                const T d = (pa > 0) ? pa * rd : (T)0;
                const T l = (r > c) ? a[cc] * rd : d;
                a[cc] = l;
                Lt[c][r] = l;
                if (r == c) Rd[c] = rd;
                wave_lds_sync();
#pragma unroll
                for (int j = cc + 1; j < 16; j++) a[j] = fma(-l, Lt[c][16 * p + j], a[j]);
            }
        }
        __syncthreads();
        if (w > p) {
#pragma unroll
            for (int cc = 0; cc < 16; cc++) {
                const int c = 16 * p + cc;
                const T l = Lt[c][r];
#pragma unroll
                for (int j = 0; j < 16; j++) a[j] = fma(-l, Lt[c][16 * w + j], a[j]);
            }
        }
    }
    if (r == 0 && bad) atomicOr(D.status, 1);  // bad is uniform in a wave (the pivots are broadcast)
    // row r of the lower triangle
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
        const int q = 16 * w + j;
        T* dst = tile + (size_t)r * ld + q;
        if (q + 3 <= r) {
            *(float4*)dst = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
        } else if (q <= r) {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (q + u <= r) dst[u] = a[j + u];
        }
    }
    if (w == 0) ((T*)D.diag)[base + r] = Lt[r][r];
    // L^-1 by 16-row panels in the same way: wave p finalizes its rows (scale by 1/L(q, q), then the
    // updates of its later rows), publishes them in X, and the waves below apply them in row order
    T x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = (16 * w + j == r) ? (T)1 : (T)0;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        if (w == p) {
#pragma unroll
            for (int qq = 0; qq < 16; qq++) {
                const int q = 16 * p + qq;
                x[qq] *= Rd[q];
                X[q][r] = x[qq];
#pragma unroll
                for (int j = qq + 1; j < 16; j++) x[j] = fma(-Lt[q][16 * p + j], x[qq], x[j]);
            }
        }
        __syncthreads();
        if (w > p) {
#pragma unroll
            for (int qq = 0; qq < 16; qq++) {
                const int q = 16 * p + qq;
                const T xq = X[q][r];
#pragma unroll
                for (int j = 0; j < 16; j++) x[j] = fma(-Lt[q][16 * w + j], xq, x[j]);
            }
        }
    }
    T* W = (T*)D.W + (size_t)kb * TS * TS;
#pragma unroll
    for (int j = 0; j < 16; j++) W[(16 * w + j) * TS + r] = x[j];
}

// The FP64 diagonal block by the same four-wave panel scheme: d = sqrt|p|, l = a / d; row q of L^-1
// divided by L(q, q), its updates applied in row order (the left-looking substitution's fused
// multiply-adds in the same order, so the same bits as that form).  LDS: Lt
// 64 x 64, X 2 x 16 x 64 (panel p's rows in X[p & 1]: the barrier after panel p + 1 orders their
// readers before panel p + 2 rewrites them).
__device__ __forceinline__ void chol_diag_f64_4w(const MatDesc& D, int kb, double* tile, int ld, double (&a)[16],
                                                 double (*Lt)[TS], double (*X)[16][TS])
{
    typedef double T;
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long base = (long long)kb * TS;
    bool bad = false;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        if (w == p) {
#pragma unroll
            for (int cc = 0; cc < 16; cc++) {
                const int c = 16 * p + cc;
                const T piv = readlane_f64(a[cc], c);
                bad |= (!(piv > 0) && base + c < D.n);  // negative, zero or NaN pivot
                const T d = sqrt(fabs(piv));
                const T l = (r > c) ? a[cc] / d : d;
                a[cc] = l;
                Lt[c][r] = l;
                wave_lds_sync();
#pragma unroll
                for (int j = cc + 1; j < 16; j++) a[j] = fma(-l, Lt[c][16 * p + j], a[j]);
            }
        }
        __syncthreads();
        if (w > p) {
#pragma unroll
            for (int cc = 0; cc < 16; cc++) {
                const int c = 16 * p + cc;
                const T l = Lt[c][r];
#pragma unroll
                for (int j = 0; j < 16; j++) a[j] = fma(-l, Lt[c][16 * w + j], a[j]);
            }
        }
    }
    if (r == 0 && bad) atomicOr(D.status, 1);
#pragma unroll
    for (int j = 0; j < 16; j += 2) {
        const int q = 16 * w + j;
        T* dst = tile + (size_t)r * ld + q;
        if (q + 1 <= r) {
            *(double2*)dst = make_double2(a[j], a[j + 1]);
        } else if (q <= r) {
            dst[0] = a[j];
        }
    }
    if (w == 0) ((T*)D.diag)[base + r] = Lt[r][r];
    T x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = (16 * w + j == r) ? (T)1 : (T)0;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        if (w == p) {
#pragma unroll
            for (int qq = 0; qq < 16; qq++) {
                const int q = 16 * p + qq;
                x[qq] = (r <= q) ? x[qq] / Lt[q][q] : (T)0;
                X[p & 1][qq][r] = x[qq];
#pragma unroll
                for (int j = qq + 1; j < 16; j++) x[j] = fma(-Lt[q][16 * p + j], x[qq], x[j]);
            }
        }
        __syncthreads();
        if (w > p) {
#pragma unroll
            for (int qq = 0; qq < 16; qq++) {
                const int q = 16 * p + qq;
                const T xq = X[p & 1][qq][r];
#pragma unroll
                for (int j = 0; j < 16; j++) x[j] = fma(-Lt[q][16 * w + j], xq, x[j]);
            }
        }
    }
    T* W = (T*)D.W + (size_t)kb * TS * TS;
#pragma unroll
    for (int j = 0; j < 16; j++) W[(16 * w + j) * TS + r] = x[j];
}

// one diagonal block per matrix (blockIdx.y), four waves; FP32 blocks of the first column read F as
// formed (A0), when the matrix has it
template <typename T>
__global__ void __launch_bounds__(256) k_chol_diag4_m(const MatDesc* mats, int kb)
{
    const MatDesc D = mats[blockIdx.y];
    if (kb >= D.nb) return;
    int ld;
    T* tile = tile_ptr<T>(D, kb, kb, ld);
    const T* src = (sizeof(T) == 4 && kb == 0 && D.A0) ? (const T*)D.A0 + (tile - (T*)D.A) : tile;
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
    typedef typename Mfma16<T>::vec_t V;
    constexpr int VW = sizeof(V) / sizeof(T);
    T a[16];
#pragma unroll
    for (int j = 0; j < 16; j += VW) {
        const V v = *(const V*)&src[(size_t)r * ld + 16 * w + j];
        a[j] = v.x;
        a[j + 1] = v.y;
        if constexpr (VW == 4) {
            a[j + 2] = v.z;
            a[j + 3] = v.w;
        }
    }
    if constexpr (sizeof(T) == 8) {
        __shared__ double Lt[TS][TS];
        __shared__ double X[2][16][TS];
        chol_diag_f64_4w(D, kb, tile, ld, a, Lt, X);
    } else {
        __shared__ __align__(16) float Lt[TS][TS];
        __shared__ float X[TS][TS];
        __shared__ float Rd[TS];
        chol_diag_f32_4w(D, kb, tile, ld, a, Lt, X, Rd);
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_chol_lpanel_bf16(const MatDesc* mats, int kb, int k0, int kend)
{
    __shared__ __attribute__((aligned(16))) __bf16 sm[4 * TS * LPB];
    __bf16* const Ah = sm;
    __bf16* const Al = sm + TS * LPB;
    __bf16* const Bh = sm + 2 * TS * LPB;
    __bf16* const Bl = sm + 3 * TS * LPB;
    typedef float T;
    // grid (matrices, row blocks): the workgroups of the first row block, which also factor the next
    // diagonal block, are dispatched first
    const MatDesc D = mats[blockIdx.x];
    if (kb >= D.nb) return;
    const int ib = kb + 1 + blockIdx.y;
    if (ib > band_last(D, kb)) return;
    int ld, lp;
    T* blk = tile_ptr<T>(D, ib, kb, ld);
    const bool first = (k0 == 0 && D.A0);
    // operand blocks are loaded into registers one step ahead of their use (the first with the C
    // block, W_kb during the last step), so their latency overlaps the barriers and products
    T cv[16];
    load64c<T>(first ? (const T*)D.A0 + (blk - (T*)D.A) : blk, ld, cv);
    f32x4_t ra[4], rb[4];
    const T* const Wkb = (const T*)D.W + (size_t)kb * TS * TS;
    if (k0 < kb) {
        load64_regs(tile_ptr<T>(D, ib, k0, lp), lp, ra);
        load64_regs(tile_ptr<T>(D, kb, k0, lp), lp, rb);
    } else {
        load64_regs(Wkb, TS, ra);
    }
    float4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (float4_t){0, 0, 0, 0};
    for (int kp = k0; kp < kb; kp++) {
        __syncthreads();
        split64_store(ra, Ah, Al);
        split64_store(rb, Bh, Bl);
        __syncthreads();
        if (kp + 1 < kb) {
            load64_regs(tile_ptr<T>(D, ib, kp + 1, lp), lp, ra);
            load64_regs(tile_ptr<T>(D, kb, kp + 1, lp), lp, rb);
        } else {
            load64_regs(Wkb, TS, ra);
        }
        mma64_bf16(Ah, Al, Bh, Bl, acc);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[a][c][v] = cv[(a * 2 + c) * 4 + v] - acc[a][c][v];
    __syncthreads();
    acc64_split(acc, Ah, Al);
    split64_store(ra, Bh, Bl);  // W_kb
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (float4_t){0, 0, 0, 0};
    mma64_bf16(Ah, Al, Bh, Bl, acc);  // L_ib,kb(i, j) = sum_q A'(i, q) W_kb(j, q)
    store64<T>(blk, ld, acc, 0);
    if (ib != kb + 1 || ib >= kend) return;
    T* dg = tile_ptr<T>(D, ib, ib, ld);
    load64c<T>(first ? (const T*)D.A0 + (dg - (T*)D.A) : dg, ld, cv);
    if (k0 < kb) load64_regs(tile_ptr<T>(D, ib, k0, lp), lp, ra);
    __syncthreads();
    acc64_split(acc, Ah, Al);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = (float4_t){0, 0, 0, 0};
    mma64_bf16(Ah, Al, Ah, Al, acc);
    for (int kp = k0; kp < kb; kp++) {
        __syncthreads();
        split64_store(ra, Ah, Al);
        __syncthreads();
        if (kp + 1 < kb) load64_regs(tile_ptr<T>(D, ib, kp + 1, lp), lp, ra);
        mma64_bf16(Ah, Al, Ah, Al, acc);
    }
    // the block, now up to date, is factored here, so the next column step needs no diagonal
    // launch: through LDS (S, 64 x 65, over the operand tiles) into rows per lane
    float* S = (float*)sm;
    float(*Lt)[TS] = (float(*)[TS])((float*)sm + TS * 65);
    float* Rd = (float*)sm + TS * 65 + TS * TS;
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int wr = w >> 1, wc = w & 1;
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = wr * 32 + a * 16 + 4 * (lane >> 4) + v;
                    const int cc = wc * 32 + c * 16 + (lane & 15);
                    S[rr * 65 + cc] = cv[(a * 2 + c) * 4 + v] - acc[a][c][v];
                }
        __syncthreads();
        float av[16];
#pragma unroll
        for (int j = 0; j < 16; j++) av[j] = S[lane * 65 + 16 * w + j];
        chol_diag_f32_4w(D, ib, dg, ld, av, Lt, (float(*)[TS])S, Rd);  // X over S: read before the first barrier
    }
}

template <typename T>
static void chol_step(hipStream_t s, const MatDesc* mats, int nmat, int kb, int rest)
{
    hipLaunchKernelGGL(k_chol_diag4_m<T>, dim3(1, nmat), dim3(256), 0, s, mats, kb);
    if (rest > 0) {  // rest: the widest trailing extent of any matrix in the launch
        hipLaunchKernelGGL(k_chol_panel_m<T>, dim3(rest, nmat), dim3(256), 0, s, mats, kb);
        hipLaunchKernelGGL(k_chol_update_m<T>, dim3(syrk_grid(rest), nmat), dim3(256), 0, s, mats, kb);
    }
}

// one column block after another of the panel [k0, kend): the diagonal kernel, then the panel kernel
template <typename T>
static void chol_subpanel(hipStream_t s, const MatDesc* mats, int nmat, int max_nb, int k0, int kend, bool left,
                          bool first_done = false)
{
    for (int kb = k0; kb < kend && kb < max_nb; kb++) {
        const int rest = max_nb - kb - 1;
        const bool f32 = sizeof(T) == 4;
        // FP32 left-looking parts: the panel kernel of column block kb - 1 factored block kb, and
        // the update before the part factored block k0 (first_done)
        if (!(left && f32 && (kb > k0 || first_done))) hipLaunchKernelGGL(k_chol_diag4_m<T>, dim3(1, nmat), dim3(256), 0, s, mats, kb);
        if (rest <= 0) continue;
        if (left) {
            if (f32)
                hipLaunchKernelGGL(k_chol_lpanel_bf16, dim3(nmat, rest), dim3(256), 0, s, mats, kb, k0, kend);
            else
                hipLaunchKernelGGL(k_chol_lpanel_m<T>, dim3(rest, nmat), dim3(256), 0, s, mats, kb, k0, kend);
            continue;
        }
        hipLaunchKernelGGL(k_chol_panel_m<T>, dim3(rest, nmat), dim3(256), 0, s, mats, kb);
        int cnt = 0;
        for (int j = kb + 1; j < kend && j < max_nb; j++) cnt += max_nb - j;
        if (cnt > 0) hipLaunchKernelGGL(k_chol_upanel_m<T>, dim3(cnt, nmat), dim3(256), 0, s, mats, kb, kend);
    }
}

// dense matrices (padded to a multiple of 128), factored panel by panel, each panel column block
// by column block and then one trailing update of the rest.  FP32 (the mixed path, evaluated in
// batches): 1024-wide panels of four 256-wide left-looking parts (the left-looking panel kernel
// reads and writes each block once per column: 0.8 ms faster per batch of 64 3600 x 3600
// matrices than the right-looking pair), each part after the first brought up to date by a strip
// update from the earlier parts, then one rank-1024 BF16-split trailing update.  Per batch of 256
// 3600 x 3600 matrices the factorization took 33.7 ms with a rank-256 update per 256 columns,
// 30.0 ms with 512-wide panels and 29.4 ms with 1024-wide ones (2048: 29.1, one panel: 30.1; FP32
// MFMA in the panel kernel then); the arithmetic, and so the result, is the same bit for bit.  FP64: 256-wide panels factored with the
// right-looking panel and panel-update kernels, whose three times more workgroups fill the GPU
// better when one matrix is factored alone (0.25 ms faster for one 3600 x 3600 matrix), and the
// FP64 trailing update.  The schedule depends on the element type only, never on the batch, so an
// item's arithmetic is the same alone and in any batch.
#define QLX_CHOL_SUB 4  // 256-wide left-looking parts per FP32 panel
template <typename T>
static void chol_two_level(hipStream_t s, const MatDesc* mats, int nmat, int max_nb, int panel_blocks)
{
    if (sizeof(T) == 4) {
        const size_t lds_bf = (size_t)4 * TT * TBP * sizeof(uint16_t);
        allow_big_lds((const void*)k_chol_trail_bf16);
        const int NB = panel_blocks > 0 ? panel_blocks : 4 * QLX_CHOL_SUB;  // panel width in 64-blocks
        for (int k0 = 0; k0 < max_nb; k0 += NB) {
            for (int ks = k0; ks < k0 + NB && ks < max_nb; ks += 4) {
                if (ks > k0) {  // the half's columns, from all earlier halves of the panel (a strip)
                    const int nts = (max_nb - ks) / 2;
                    hipLaunchKernelGGL(k_chol_trail_bf16, dim3(2 * nts - 1, nmat), dim3(256), lds_bf, s, mats, k0, ks - k0, 2, 1);
                }
                // the part's first diagonal block was factored by the strip or trailing update before it
                chol_subpanel<T>(s, mats, nmat, max_nb, ks, ks + 4, true, ks > 0);
            }
            if (k0 + NB >= max_nb) break;
            const int nt = (max_nb - k0 - NB) / 2;
            hipLaunchKernelGGL(k_chol_trail_bf16, dim3(syrk_grid(nt), nmat), dim3(256), lds_bf, s, mats, k0, NB, 0, 1);
        }
        return;
    }
    const size_t lds = (size_t)2 * TT * LDT * sizeof(T);
    allow_big_lds((const void*)k_chol_trail_m<T>);
    for (int k0 = 0; k0 < max_nb; k0 += 4) {
        chol_subpanel<T>(s, mats, nmat, max_nb, k0, k0 + 4, false);
        if (k0 + 4 < max_nb) {
            const int nt = (max_nb - k0 - 4) / 2;
            hipLaunchKernelGGL(k_chol_trail_m<T>, dim3(syrk_grid(nt), nmat), dim3(256), lds, s, mats, k0, 4);
        }
    }
}

// d_mats holds n32 FP32 matrices first, then n64 FP64 matrices.  Dense matrices (all of the
// launch) use the two-level schedule; banded ones (lower band <= max_bwb blocks) the 64-block
// right-looking one.
void cholesky_batched(hipStream_t s, const MatDesc* d_mats, int n32, int n64, int max_nb, int dense, int max_bwb, int panel_blocks)
{
    if (dense) {
        if (n32) chol_two_level<float>(s, d_mats, n32, max_nb, panel_blocks);
        if (n64) chol_two_level<double>(s, d_mats + n32, n64, max_nb, 0);
        return;
    }
    for (int kb = 0; kb < max_nb; kb++) {
        const int rest = std::min(max_nb - kb - 1, max_bwb);  // banded: at most max_bwb blocks below
        if (n32) chol_step<float>(s, d_mats, n32, kb, rest);
        if (n64) chol_step<double>(s, d_mats + n32, n64, kb, rest);
    }
}

// 2 sum log(diag) in a fixed order (Cholesky_logdet_lower_packed, pixelgrid.cpp:17201-17210)
template <typename T>
__device__ double block_sum_log(const T* diag, int n, double* sh)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += log(fabs((double)diag[i]));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    double r = 2.0 * sh[0];
    __syncthreads();
    return r;
}

template <typename T>
__global__ void __launch_bounds__(256) k_logdets_b(BatchDev bd)
{
    __shared__ double sh[256];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const double* v = bd.vec + M.vec_off;
    const T* diagF = (const T*)(v + M.ldN);
    // R's pivots: those of the item whose factorization serves this one (the same R)
    const int rr = bd.rrep[item];
    const double* diagR = bd.vec + bd.meta[rr].vec_off + 2 * (size_t)M.ldN;
    const double lf = block_sum_log<T>(diagF, M.n_src, sh);
    const double lr = block_sum_log<double>(diagR, M.n_src, sh);
    if (threadIdx.x == 0) {
        bd.out[QLX_OUT_STRIDE * item + 2] = lf;
        bd.out[QLX_OUT_STRIDE * item + 3] = lr;
        if (rr != item) bd.status[4 * item + 1] = bd.status[4 * rr + 1];
    }
}

void launch_logdets_b(hipStream_t s, const BatchDev& bd, int f32)
{
    if (f32) hipLaunchKernelGGL(k_logdets_b<float>, dim3(1, bd.B), dim3(256), 0, s, bd);
    else hipLaunchKernelGGL(k_logdets_b<double>, dim3(1, bd.B), dim3(256), 0, s, bd);
}

// ------------------------------------------------------------------------------------------------
// Triangular solves with the stored block inverses (one workgroup per item).
// y = L^-1 D is row n of the augmented factor, so the first solve is backward only: x = L^-T y.
// ------------------------------------------------------------------------------------------------
// Each block step: a strip product over the solved part of the vector (16 B loads, eight in flight
// per lane), then the 64x64 block inverse applied by all 1024 threads, 16 lanes per output.
template <typename T>
__device__ __forceinline__ T sum16(T v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// forward: x_kb = W_kb (b_kb - L[kb, 0:kb] x_{0:kb}); wave w owns rows w, w+16, w+32, w+48 of the block
template <typename T>
__device__ void tri_fwd(const T* L, int ld, int nb, const T* W, T* sx, T (*part)[TS + 1], T* rhs)
{
    constexpr int VW = 16 / sizeof(T);
    typedef T P __attribute__((ext_vector_type(VW)));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int orow = tid >> 4, oq = 4 * (tid & 15);
    for (int kb = 0; kb < nb; kb++) {
        const int r0 = kb * TS;
        T wv[4];
        {
            const T* Wr = W + (size_t)kb * TS * TS + orow * TS + oq;
            P a = *(const P*)Wr;
            if constexpr (VW == 4) {
                wv[0] = a.x; wv[1] = a.y; wv[2] = a.z; wv[3] = a.w;
            } else {
                const P b = *(const P*)(Wr + 2);
                wv[0] = a.x; wv[1] = a.y; wv[2] = b.x; wv[3] = b.y;
            }
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (oq + j > orow) wv[j] = (T)0;
        }
        const T* L0 = L + (size_t)(r0 + w) * ld;
        const size_t s16 = (size_t)16 * ld;
        P acc[4];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = (P)0;
        int c = VW * lane;
        for (; c + 64 * VW < r0; c += 128 * VW) {
            P v[8];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                v[q] = *(const P*)(L0 + q * s16 + c);
                v[4 + q] = *(const P*)(L0 + q * s16 + c + 64 * VW);
            }
            const P sa = *(const P*)(sx + c), sb = *(const P*)(sx + c + 64 * VW);
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = __builtin_elementwise_fma(v[4 + q], sb, __builtin_elementwise_fma(v[q], sa, acc[q]));
        }
        if (c < r0) {
            const P sa = *(const P*)(sx + c);
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = __builtin_elementwise_fma(*(const P*)(L0 + q * s16 + c), sa, acc[q]);
        }
        T a[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            a[q] = acc[q].x + acc[q].y;
            if constexpr (VW == 4) a[q] += acc[q].z + acc[q].w;
        }
        // 4 row sums over 64 lanes: halve the set twice, then a 16-lane sum
        const bool h5 = lane & 32, h4 = lane & 16;
        const T k0 = (h5 ? a[2] : a[0]) + __shfl_xor(h5 ? a[0] : a[2], 32, 64);
        const T k1 = (h5 ? a[3] : a[1]) + __shfl_xor(h5 ? a[1] : a[3], 32, 64);
        const T k = sum16<T>((h4 ? k1 : k0) + __shfl_xor(h4 ? k0 : k1, 16, 64));
        if ((lane & 15) == 0) {
            const int r = w + 16 * (lane >> 4);
            rhs[r] = sx[r0 + r] - k;
        }
        __syncthreads();
        T y = wv[0] * rhs[oq];
#pragma unroll
        for (int j = 1; j < 4; j++) y = fma(wv[j], rhs[oq + j], y);
        y = sum16<T>(y);
        if ((tid & 15) == 0) sx[r0 + orow] = y;
        __syncthreads();
    }
}

// backward: x_kb = W_kb^T (b_kb - L[kb+1:, kb]^T x_{kb+1:}); lanes cover the block's 64 columns with
// 16 B loads, the rest of the wave and the 16 waves split the rows
template <typename T>
__device__ void tri_bwd(const T* L, int ld, int nb, const T* W, T* sx, T (*part)[TS + 1], T* rhs)
{
    constexpr int VW = 16 / sizeof(T);
    constexpr int LPR = TS / VW;   // lanes per row
    constexpr int RPW = 64 / LPR;  // rows per wave per pass
    constexpr int RPP = 16 * RPW;  // rows per pass over the workgroup
    typedef T P __attribute__((ext_vector_type(VW)));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ocol = tid >> 4, oq = 4 * (tid & 15);
    const int cv = VW * (lane % LPR), rsub = lane / LPR;
    const int rend = nb * TS;
    for (int kb = nb - 1; kb >= 0; kb--) {
        const int r0 = kb * TS;
        T wv[4];
        {
            const T* Wc = W + (size_t)kb * TS * TS + (size_t)oq * TS + ocol;
#pragma unroll
            for (int j = 0; j < 4; j++) wv[j] = (oq + j >= ocol) ? Wc[j * TS] : (T)0;
        }
        P acc = (P)0;
        int rr = r0 + TS + RPW * w + rsub;
        for (; rr + 7 * RPP < rend; rr += 8 * RPP) {
            P v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = *(const P*)(L + (size_t)(rr + u * RPP) * ld + r0 + cv);
#pragma unroll
            for (int u = 0; u < 8; u++) acc = __builtin_elementwise_fma(v[u], (P)sx[rr + u * RPP], acc);
        }
        for (; rr < rend; rr += RPP) acc = __builtin_elementwise_fma(*(const P*)(L + (size_t)rr * ld + r0 + cv), (P)sx[rr], acc);
#pragma unroll
        for (int off = LPR; off < 64; off <<= 1) {
            acc.x += __shfl_xor(acc.x, off, 64);
            acc.y += __shfl_xor(acc.y, off, 64);
            if constexpr (VW == 4) {
                acc.z += __shfl_xor(acc.z, off, 64);
                acc.w += __shfl_xor(acc.w, off, 64);
            }
        }
        if (rsub == 0) {
            part[w][cv] = acc.x;
            part[w][cv + 1] = acc.y;
            if constexpr (VW == 4) {
                part[w][cv + 2] = acc.z;
                part[w][cv + 3] = acc.w;
            }
        }
        __syncthreads();
        if (tid < TS) {
            T s = sx[r0 + tid];
#pragma unroll
            for (int q = 0; q < 16; q++) s -= part[q][tid];
            rhs[tid] = s;
        }
        __syncthreads();
        T y = wv[0] * rhs[oq];
#pragma unroll
        for (int j = 1; j < 4; j++) y = fma(wv[j], rhs[oq + j], y);
        y = sum16<T>(y);
        if ((tid & 15) == 0) sx[r0 + ocol] = y;
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_trisolve_b(BatchDev bd)
{
    extern __shared__ double sx_raw[];
    __shared__ T part[16][TS + 1];
    __shared__ T rhs[TS];
    T* sx = (T*)sx_raw;
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const int ld = M.ldN, nb = ld / TS, n = M.n_src;
    const T* L = (sizeof(T) == 8) ? (const T*)(bd.F + M.f_off) : (const T*)(bd.F32 + M.f_off);
    const T* W = (const T*)(bd.winv + M.w_off);
    for (int c = threadIdx.x; c < ld; c += 1024) sx[c] = (c < n) ? L[(size_t)n * ld + c] : (T)0;
    __syncthreads();
    tri_bwd<T>(L, ld, nb, W, sx, part, rhs);
    double* s = bd.vec + M.vec_off;
    for (int c = threadIdx.x; c < ld; c += 1024) s[c] = (c < n) ? (double)sx[c] : 0.0;
}

void launch_trisolve_b(hipStream_t s, const BatchDev& bd, int max_ld, int f32)
{
    if (f32) {
        allow_big_lds((const void*)k_trisolve_b<float>);
        hipLaunchKernelGGL(k_trisolve_b<float>, dim3(1, bd.B), dim3(1024), (size_t)max_ld * 4, s, bd);
    } else {
        allow_big_lds((const void*)k_trisolve_b<double>);
        hipLaunchKernelGGL(k_trisolve_b<double>, dim3(1, bd.B), dim3(1024), (size_t)max_ld * 8, s, bd);
    }
}

// ------------------------------------------------------------------------------------------------
// FP64 iterative refinement of the mixed-precision solve (classical mixed-precision refinement,
// Wilkinson / Moler): r = D - F s in FP64 arithmetic on the FP32 F as formed (Ff32 + Fdiag),
// dx = F^-1 r with the FP32 factor, s += dx.  The residual kernel writes per-row-block partial
// sums of r^2 and D^2; the refine kernel stops an item once ||r|| <= tol ||D||.
// status[4*item+2] counts steps; status[4*item+3] marks a finished item: 1 converged, 2 stopped after
// the last allowed step with the residual still above the tolerance.
// ------------------------------------------------------------------------------------------------
// sum of a[k] over the 64 lanes for 16 values at once: four halvings, then a 4-lane sum; lane l ends
// with the total of a[(l >> 2) & 15]
__device__ __forceinline__ double lane_sum16(const double* a)
{
    const int lane = threadIdx.x & 63;
    double b[8], c[4], d[2];
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = a[(h5 ? 8 : 0) + j] + __shfl_xor(a[(h5 ? 0 : 8) + j], 32, 64);
#pragma unroll
    for (int j = 0; j < 4; j++) c[j] = b[(h4 ? 4 : 0) + j] + __shfl_xor(b[(h4 ? 0 : 4) + j], 16, 64);
#pragma unroll
    for (int j = 0; j < 2; j++) d[j] = c[(h3 ? 2 : 0) + j] + __shfl_xor(c[(h3 ? 0 : 2) + j], 8, 64);
    double e = d[h2 ? 1 : 0] + __shfl_xor(d[h2 ? 0 : 1], 4, 64);
    e += __shfl_xor(e, 2, 64);
    e += __shfl_xor(e, 1, 64);
    return e;
}

// r = D - F s on F as formed (Ff32, lower 128-tiles; a diagonal tile is held whole), reading each
// lower element once: k_resid_rows_b takes row block I (64 rows): its row strip (columns c < 64 I
// of rows i) gives the rows' sums, and the same loads give the strip's transposed products
// z_I(c) = sum_i F(i, c) s(i), stored per row block (zres); with the diagonal tile (through LDS, its
// upper half) the rows' partial sums go to the slot Yp.  k_resid_cols_b adds, for the rows of block
// J, the column strip sum_{I > J} z_I in order of I, then r = D - (Fdiag s + that + Yp); D(i) =
// F(n, i).  Products are formed and summed in FP64; rows >= n take no part.
__global__ void __launch_bounds__(256) k_resid_rows_b(BatchDev bd)
{
    __shared__ float tile[TS][TS + 1];
    __shared__ double dacc[4][TS];
    __shared__ double racc[TS];
    __shared__ double zs[2][4][256];  // column partials of passes it (even / odd)
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid || bd.status[4 * item + 3]) return;
    const int ld = M.ldN, n = M.n_src, nb = ld / TS;
    const int I = (int)gridDim.x - 1 - (int)blockIdx.x;  // the longest row strips are dispatched first
    if (I >= nb || I * TS >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const float* U = bd.Ff32 + M.f_off;
    double* v = bd.vec + M.vec_off;
    const double* s = v;
    double* yp = v + 7 * (size_t)ld;
    double* Z = bd.zres + M.f_off / 128 + (size_t)TS * I * (I - 1) / 2;  // z_I: c0 values
    const int c0 = I * TS;
    // diagonal tile
#pragma unroll
    for (int e = tid; e < TS * TS / 4; e += 256) {
        const int row = e >> 4, cc = 4 * (e & 15);
        const float4_t t = *(const float4_t*)(U + (size_t)(c0 + row) * ld + c0 + cc);
        tile[row][cc] = t.x;
        tile[row][cc + 1] = t.y;
        tile[row][cc + 2] = t.z;
        tile[row][cc + 3] = t.w;
    }
    // row strip: F(i,c) = U[i][c] for c < c0; wave w owns rows c0 + w + 4k, k = 0..15, lane the four
    // columns 256 it + 4 lane .. + 3 of pass it (every thread runs every pass: the passes' barrier).
    // A pass's loads are issued before the previous pass's barrier and column sums; the column
    // partials alternate between two LDS buffers, so one barrier per pass orders them
    {
        double ra[16], si[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            ra[k] = 0.0;
            const int row = c0 + w + 4 * k;
            si[k] = row < n ? s[row] : 0.0;
        }
        const float* Ur = U + (size_t)(c0 + w) * ld;
        const int npass = (c0 + 255) / 256;
        float4_t t[16];
        if (4 * lane < c0) {  // c0 is a multiple of 64, so c + 3 < c0 for every c < c0 below
#pragma unroll
            for (int k = 0; k < 16; k++) t[k] = *(const float4_t*)(Ur + (size_t)(4 * k) * ld + 4 * lane);
        }
        for (int it = 0; it < npass; it++) {
            const int c = 256 * it + 4 * lane;
            double z0 = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
            if (c < c0) {
                const double s0 = s[c], s1 = s[c + 1], s2 = s[c + 2], s3 = s[c + 3];
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    ra[k] = fma((double)t[k].x, s0, fma((double)t[k].y, s1, fma((double)t[k].z, s2, fma((double)t[k].w, s3, ra[k]))));
                    z0 = fma((double)t[k].x, si[k], z0);
                    z1 = fma((double)t[k].y, si[k], z1);
                    z2 = fma((double)t[k].z, si[k], z2);
                    z3 = fma((double)t[k].w, si[k], z3);
                }
            }
            double* zb = &zs[it & 1][0][0];
            zb[w * 256 + 4 * lane] = z0;
            zb[w * 256 + 4 * lane + 1] = z1;
            zb[w * 256 + 4 * lane + 2] = z2;
            zb[w * 256 + 4 * lane + 3] = z3;
            if (c + 256 < c0) {
#pragma unroll
                for (int k = 0; k < 16; k++) t[k] = *(const float4_t*)(Ur + (size_t)(4 * k) * ld + c + 256);
            }
            __syncthreads();
            if (256 * it + tid < c0) Z[256 * it + tid] = ((zb[tid] + zb[256 + tid]) + zb[512 + tid]) + zb[768 + tid];
        }
        const double tot = lane_sum16(ra);
        if ((lane & 3) == 0) racc[w + 4 * (lane >> 2)] = tot;
    }
    __syncthreads();
    // diagonal tile: thread (i, g) sums c in [16 g, 16 g + 16), c != i, c0 + c < n
    {
        const int i = tid & 63, g = tid >> 6;
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = 16 * g + q;
            if (c != i && c0 + c < n) a = fma((double)(c > i ? tile[i][c] : tile[c][i]), s[c0 + c], a);
        }
        dacc[g][i] = a;
    }
    __syncthreads();
    if (tid < TS && c0 + tid < n)
        yp[c0 + tid] = ((dacc[0][tid] + dacc[1][tid]) + (dacc[2][tid] + dacc[3][tid])) + racc[tid];
}

__global__ void __launch_bounds__(64) k_resid_cols_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid || bd.status[4 * item + 3]) return;
    const int ld = M.ldN, n = M.n_src, nb = ld / TS;
    const int J = blockIdx.x;
    if (J >= nb || J * TS >= n) return;
    const int nbr = (n + TS - 1) / TS;
    const int tid = threadIdx.x;
    const float* U = bd.Ff32 + M.f_off;
    double* v = bd.vec + M.vec_off;
    const double* s = v;
    const double* fdiag = v + 3 * (size_t)ld;
    double* r = v + 4 * (size_t)ld;
    double* part = v + 5 * (size_t)ld;
    const double* yp = v + 7 * (size_t)ld;
    const double* Zb = bd.zres + M.f_off / 128;
    const int i = J * TS + tid;
    double rr = 0.0, dd = 0.0;
    if (i < n) {
        double cs = 0.0;
        for (int I = J + 1; I < nbr; I++) cs += Zb[(size_t)TS * I * (I - 1) / 2 + i];
        const double D = (double)U[(size_t)n * ld + i];
        const double Fs = fdiag[i] * s[i] + cs + yp[i];
        rr = D - Fs;
        r[i] = rr;
        rr = rr * rr;
        dd = D * D;
    }
    for (int off = 32; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off, 64);
        dd += __shfl_xor(dd, off, 64);
    }
    if (tid == 0) {
        part[J] = rr;
        part[nb + J] = dd;
    }
}

// one refinement step per item: check convergence from the residual partials, else correct s
__global__ void __launch_bounds__(1024) k_refine_b(BatchDev bd, int final_check, double tol)
{
    extern __shared__ double sx_raw[];
    __shared__ float part[16][TS + 1];
    __shared__ float rhs[TS];
    __shared__ double red[2][64];
    __shared__ int done;
    float* sx = (float*)sx_raw;
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid || bd.status[4 * item + 3]) return;
    const int ld = M.ldN, nb = ld / TS, n = M.n_src;
    const int tid = threadIdx.x;
    double* v = bd.vec + M.vec_off;
    double* s = v;
    const double* r = v + 4 * (size_t)ld;
    const double* pr = v + 5 * (size_t)ld;
    const int nbr = (n + TS - 1) / TS;
    if (tid < 64) {
        double a = 0.0, b = 0.0;
        for (int k = tid; k < nbr; k += 64) {
            a += pr[k];
            b += pr[nb + k];
        }
        red[0][tid] = a;
        red[1][tid] = b;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 64; k++) {
            a += red[0][k];
            b += red[1][k];
        }
        const double rel = (b > 0) ? sqrt(a / b) : 0.0;
        bd.out[QLX_OUT_STRIDE * item + 6] = rel;
        done = (rel <= tol || final_check) ? 1 : 0;
        if (done) bd.status[4 * item + 3] = (rel <= tol) ? 1 : 2;  // 2: stopped above the tolerance
        else bd.status[4 * item + 2] += 1;
    }
    __syncthreads();
    if (done) return;
    const float* L = bd.F32 + M.f_off;
    const float* W = (const float*)(bd.winv + M.w_off);
    for (int c = tid; c < ld; c += 1024) sx[c] = (c < n) ? (float)r[c] : 0.0f;
    __syncthreads();
    tri_fwd<float>(L, ld, nb, W, sx, part, rhs);
    tri_bwd<float>(L, ld, nb, W, sx, part, rhs);
    for (int c = tid; c < n; c += 1024) s[c] += (double)sx[c];
}

void launch_refine_b(hipStream_t s, const BatchDev& bd, int max_ld, int iters, double tol)
{
    const int nb = max_ld / TS;
    allow_big_lds((const void*)k_refine_b);
    for (int it = 0; it <= iters; it++) {
        hipLaunchKernelGGL(k_resid_rows_b, dim3(nb, bd.B), dim3(256), 0, s, bd);
        hipLaunchKernelGGL(k_resid_cols_b, dim3(nb, bd.B), dim3(64), 0, s, bd);
        hipLaunchKernelGGL(k_refine_b, dim3(1, bd.B), dim3(1024), (size_t)max_ld * 4, s, bd, it == iters ? 1 : 0, tol);
    }
}

// ------------------------------------------------------------------------------------------------
// evidence: v = L s per active pixel, model = PSF (x) v over active taps, chi^2 over the mask,
// Es = sum_l ||H_l s||^2 (= s^T R s), -2 ln E = ((0 + regterms) + ln det F) + chi^2
// ------------------------------------------------------------------------------------------------
// s in the reference order: s_ref[pisrc[p]] = s[p]
__global__ void __launch_bounds__(256) k_unpermute_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || p >= M.n_src) return;
    const double* s = bd.vec + M.vec_off;
    double* sref = bd.vec + M.vec_off + 6 * (size_t)M.ldN;
    sref[bd.pisrc[item * bd.s_cell + p]] = s[p];
}

__global__ void __launch_bounds__(256) k_Ls_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || r >= M.n_img) return;
    const int U = 3 * bd.nsub;
    const int* ucol = bd.ucol + ((size_t)item * bd.s_pix + r) * U;
    const double* uval = bd.uval + ((size_t)item * bd.s_pix + r) * U;
    const double* s = bd.vec + M.vec_off;
    double acc = 0.0;
    const int cnt = bd.ucnt[item * bd.s_pix + r];
    if (U == 12) {  // 2 x 2 split: all 12 slots written by the merge, same summation order
        int c[12];
        double w[12];
        load_row12(ucol, uval, c, w);
#pragma unroll
        for (int e = 0; e < 12; e++)
            if (e < cnt) acc += w[e] * s[c[e]];
    } else {
        for (int e = 0; e < cnt; e++) acc += uval[e] * s[ucol[e]];
    }
    // the unblurred model as a dense image (pixel index; inactive pixels keep the zero fill)
    bd.model[(size_t)item * 2 * bd.s_pix + bd.s_pix + bd.act_pix[item * bd.s_pix + r]] = acc;
}

// blurred model of every active pixel: the PSF taps over the dense unblurred image (zero outside
// the active set, which is what skipping inactive neighbours gives).  A lane takes MKB pixels
// (k0 .. k0+MKB-1, l) and lanes of a wave consecutive l, so the lanes of one load read consecutive
// pixels; each image value it loads serves every one of its pixels that the tap reaches
// ((pnx + MKB - 1) pny loads for MKB pixels instead of MKB pnx pny).  Image rows go in descending
// order, so each pixel still sums its taps pk ascending, pl ascending: the result is the
// one-pixel-per-lane gather's, bit for bit.
#define MKB 4
__global__ void __launch_bounds__(256) k_model_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    const int nx = bd.nx, ny = bd.ny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (!M.valid || t >= (nx + MKB - 1) / MKB * ny) return;
    const int kb = t / ny, l = t - kb * ny, k0 = kb * MKB;
    const int* pidx = bd.pidx + item * bd.s_pix;
    bool any = false;
#pragma unroll
    for (int a = 0; a < MKB; a++) any |= (k0 + a < nx) && pidx[(k0 + a) * ny + l] >= 0;
    if (!any) return;
    const int nxh = bd.pnx / 2, nyh = bd.pny / 2, pnx = bd.pnx, pny = bd.pny;
    double* model = bd.model + (size_t)item * 2 * bd.s_pix;
    const double* v = model + bd.s_pix;
    double acc[MKB];
#pragma unroll
    for (int a = 0; a < MKB; a++) acc[a] = 0.0;
    for (int ii = k0 + MKB - 1 + nxh; ii >= k0 + nxh - pnx + 1; ii--) {
        if (ii < 0 || ii >= nx) continue;
        const double* vr = v + (size_t)ii * ny;
        for (int pl = 0; pl < pny; pl++) {
            const int jj = l + nyh - pl;
            if (jj < 0 || jj >= ny) continue;
            const double x = vr[jj];
#pragma unroll
            for (int a = 0; a < MKB; a++) {
                const int pk = k0 + a + nxh - ii;
                if (pk >= 0 && pk < pnx) acc[a] += bd.psf[pk * pny + pl] * x;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < MKB; a++) {
        const int pa = (k0 + a) * ny + l;
        if (k0 + a < nx && pidx[pa] >= 0) model[pa] = acc[a];
    }
}

__global__ void __launch_bounds__(1024) k_evidence_b(BatchDev bd)
{
    __shared__ double sh[1024];
    __shared__ double sh2[1024];
    __shared__ double sh3[1024];
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const int npix = bd.nx * bd.ny;
    const int* pm = bd.pmapped + item * bd.s_pix;
    const double* model = bd.model + (size_t)item * 2 * bd.s_pix;
    const int* sidx = bd.sidx + item * bd.s_cell;
    const double* s = bd.vec + M.vec_off + 6 * (size_t)M.ldN;  // reference order
    double chi = 0.0, chia = 0.0;  // all mask pixels; active pixels only (chisq_regparam_dense's Ed)
    for (int p = threadIdx.x; p < npix; p += 1024) {
        if (bd.mask && !bd.mask[p]) continue;
        const double d = bd.data[p];
        if (pm[p]) {
            const double r = model[p] - d;
            chi += (r * r) * bd.covinv[p];
            chia += (r * r) * bd.covinv[p];
        } else {
            chi += (d * d) * bd.covinv[p];
        }
    }
    double es = 0.0;
    const int snx = M.snx, sny = M.sny;
    for (int c = threadIdx.x; c < snx * sny; c += 1024) {
        const int a = sidx[c];
        if (a < 0) continue;
        const int i = c / sny, j = c % sny;
        const double sa = s[a];
        int n0 = (i < snx - 1) ? sidx[(i + 1) * sny + j] : -1;
        int n1 = (i > 0) ? sidx[(i - 1) * sny + j] : -1;
        double h = -2.0 * sa + (n0 >= 0 ? s[n0] : 0.0) + (n1 >= 0 ? s[n1] : 0.0);
        es += h * h;
        n0 = (j < sny - 1) ? sidx[i * sny + j + 1] : -1;
        n1 = (j > 0) ? sidx[i * sny + j - 1] : -1;
        h = -2.0 * sa + (n0 >= 0 ? s[n0] : 0.0) + (n1 >= 0 ? s[n1] : 0.0);
        es += h * h;
    }
    sh[threadIdx.x] = chi;
    sh2[threadIdx.x] = es;
    sh3[threadIdx.x] = chia;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sh[threadIdx.x] += sh[threadIdx.x + st];
            sh2[threadIdx.x] += sh2[threadIdx.x + st];
            sh3[threadIdx.x] += sh3[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = bd.out + QLX_OUT_STRIDE * item;
        const double chisq = sh[0], Es = sh2[0];
        const double logdetF = o[2], logdetR = o[3];
        const double lam = item_lambda(bd, item);
        double loglike = 0.0, regterms = 0.0;
        if (lam != 0) {
            regterms = lam * Es - M.n_src * log(lam) - logdetR;
            loglike += regterms;
        }
        loglike += logdetF;
        loglike += chisq;
        o[0] = loglike;
        o[1] = chisq;
        o[4] = Es;
        o[5] = regterms;
        o[7] = (double)bd.status[4 * item + 2];
        o[8] = (sh3[0] + regterms) + logdetF;  // chisq_regparam_dense: Ed + loglike_reg, then + logdet
    }
}

__global__ void __launch_bounds__(256) k_keep_b(BatchDev bd, char* keep)
{
    const int item = blockIdx.y;
    const ItemMeta& M = bd.meta[item];
    if (!M.valid) return;
    const int ncell = M.snx * M.sny;
    double* ks = (double*)(keep + M.keep_off);
    double* km = ks + ncell;
    int* ki = (int*)(km + bd.s_pix);
    const double* sref = bd.vec + M.vec_off + 6 * (size_t)M.ldN;
    const double* model = bd.model + (size_t)item * 2 * bd.s_pix;
    const int* sidx = bd.sidx + (size_t)item * bd.s_cell;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < bd.s_pix; t += stride) {
        km[t] = model[t];
        if (t < ncell) ki[t] = sidx[t];
        if (t < M.n_src) ks[t] = sref[t];
    }
    for (long long t = bd.s_pix + (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) {
        ki[t] = sidx[t];
        if (t < M.n_src) ks[t] = sref[t];
    }
}

void launch_keep_b(hipStream_t s, const BatchDev& bd, void* keep, int max_cells)
{
    const long long n = std::max<long long>(bd.s_pix, max_cells);
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 64);
    hipLaunchKernelGGL(k_keep_b, dim3(nb, bd.B), dim3(256), 0, s, bd, (char*)keep);
}

void launch_evidence_b(hipStream_t st, const BatchDev& bd, int max_img, int max_src)
{
    launch_fill_f64(st, bd.model, 0.0, (size_t)bd.B * 2 * bd.s_pix);
    if (max_src > 0) hipLaunchKernelGGL(k_unpermute_b, dim3((max_src + 255) / 256, bd.B), dim3(256), 0, st, bd);
    if (max_img > 0) {
        hipLaunchKernelGGL(k_Ls_b, dim3((max_img + 255) / 256, bd.B), dim3(256), 0, st, bd);
        const long long nlane = (long long)(bd.nx + MKB - 1) / MKB * bd.ny;
        hipLaunchKernelGGL(k_model_b, dim3((unsigned)((nlane + 255) / 256), bd.B), dim3(256), 0, st, bd);
    }
    hipLaunchKernelGGL(k_evidence_b, dim3(1, bd.B), dim3(1024), 0, st, bd);
}

// ------------------------------------------------------------------------------------------------
// MFMA layout self-tests (fixed operands, results per lane/register)
// ------------------------------------------------------------------------------------------------
__global__ void k_selftest_mfma(double* out)
{
    const int lane = threadIdx.x;
    // f64 16x16x4: A[i][k] = i + 1 + 0.01 k, B[k][j] = 1 + k + 0.1 j
    {
        const int i = lane & 15, k = lane >> 4;
        double a = i + 1 + 0.01 * k, b = 1 + k + 0.1 * (lane & 15);
        double4_t acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; v++) out[lane * 4 + v] = acc[v];
    }
    // f32 16x16x4: same operands in float
    {
        const int i = lane & 15, k = lane >> 4;
        float a = i + 1 + 0.01f * k, b = 1 + k + 0.1f * (lane & 15);
        float4_t acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; v++) out[256 + lane * 4 + v] = acc[v];
    }
    // bf16 16x16x32: lane holds k = 8*(lane/16) .. +8 of row/col lane%16;
    // A[i][k] = (i+1) * (k%4+1) / 4, B[k][j] = (j+1) * ((k/4)%8 + 1) / 8
    {
        bf16x8_t a, b;
        for (int t = 0; t < 8; t++) {
            const int k = 8 * (lane >> 4) + t;
            a[t] = (__bf16)((float)((lane & 15) + 1) * (float)(k % 4 + 1) * 0.25f);
            b[t] = (__bf16)((float)((lane & 15) + 1) * (float)((k / 4) % 8 + 1) * 0.125f);
        }
        float4_t acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; v++) out[512 + lane * 4 + v] = acc[v];
    }
    // f16 16x16x32: the same operand values and lane layout in FP16
    {
        f16x8_t a, b;
        for (int t = 0; t < 8; t++) {
            const int k = 8 * (lane >> 4) + t;
            a[t] = (_Float16)((float)((lane & 15) + 1) * (float)(k % 4 + 1) * 0.25f);
            b[t] = (_Float16)((float)((lane & 15) + 1) * (float)((k / 4) % 8 + 1) * 0.125f);
        }
        float4_t acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; v++) out[768 + lane * 4 + v] = acc[v];
    }
}

void launch_selftest_mfma(hipStream_t s, double* out)
{
    hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, s, out);
}

}  // namespace qlx
