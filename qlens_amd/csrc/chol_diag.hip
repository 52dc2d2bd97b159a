// Diagonal-block step of the blocked FP64 Cholesky (gfx950).
//
// 256 threads factor one 64x64 block: thread (r, g) holds row r, columns 16g..16g+15 in
// registers.  Each column step: the owner of the pivot writes sqrt(a_cc) to LDS, the column-c
// owners scale and publish l_rc to LDS, and every thread applies the rank-1 update to its 16
// entries reading l_qc from LDS (one barrier-separated broadcast per step).  The block inverse
// L^-1 (for the panel multiply and the triangular solves) is then formed with 4 threads per
// column, each summing a quarter of every dot product.
// Restates Cholesky_dcmp_packed (pixelgrid.cpp:17071-17105) for one diagonal block: negative
// pivots are flagged in *status and continued with sqrt(|a|) as the reference does.
#include "qlx_internal.h"

namespace qlx {

#define TS 64

template <typename T>
__global__ void __launch_bounds__(256) k_chol_diag(T* A, int ld, int kb, T* diag, int* status, T* Winv, int nflag)
{
    __shared__ T colbuf[2][TS];     // l_qc of the current step, double-buffered by step parity
    __shared__ T Ls[TS][TS + 1];    // final L (row-major)
    __shared__ T Xs[TS][TS + 1];    // L^-1
    __shared__ T part[4][TS];
    __shared__ int bad;
    const int tid = threadIdx.x;
    const int r = tid >> 2, g = tid & 3;
    const long long base = (long long)kb * TS;
    T a[16];
    const T* arow = A + (base + r) * ld + base + 16 * g;
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = (16 * g + j <= r) ? arow[j] : (T)0;
    if (tid == 0) bad = 0;
    for (int c = 0; c < TS; c++) {
        T* col = colbuf[c & 1];  // step c+1 writes the other buffer while step c may still be read
        const int gc = c >> 4, jc = c & 15;
        // owner of (c, c) publishes the pivot
        T v = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) v = (j == jc) ? a[j] : v;  // this thread's column-c entry
        if (r == c && g == gc) {
            if (v < 0 && base + c < nflag) bad = 1;
            col[c] = sqrt(fabs(v));
        }
        __syncthreads();
        const T d = col[c];
        // column owners scale l_rc = a_rc / d (one division), the pivot stays on the diagonal
        T lrc = (r > c) ? v / d : d;
        if (g == gc && r >= c) {
#pragma unroll
            for (int j = 0; j < 16; j++) a[j] = (j == jc) ? lrc : a[j];
        }
        if (g == gc && r > c) col[r] = lrc;
        __syncthreads();
        if (r > c) {
            const T l = col[r];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int q = 16 * g + j;
                if (q > c && q <= r) a[j] = fma(-l, col[q], a[j]);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && bad) atomicOr(status, 1);
    T* drow = A + (base + r) * ld + base + 16 * g;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int q = 16 * g + j;
        T v = (q <= r) ? a[j] : (T)0;
        drow[j] = v;
        Ls[r][q] = v;
    }
    if (g == (r >> 4)) diag[base + r] = a[r & 15];
    __syncthreads();
    // X = L^-1 by rows: X[i][c] = (delta_ic - sum_{q<i} L[i][q] X[q][c]) / L[i][i]; thread (c, g)
    // sums q = g, g+4, ... and the four partial sums are combined in LDS.
    const int cidx = tid >> 2;
    for (int i = 0; i < TS; i++) {
        T acc = 0;
        for (int q = g; q < i; q += 4) acc = fma(Ls[i][q], Xs[q][cidx], acc);
        part[g][cidx] = acc;
        __syncthreads();
        if (g == 0) {
            T s = (i == cidx ? (T)1 : (T)0) - (part[0][cidx] + part[1][cidx] + part[2][cidx] + part[3][cidx]);
            Xs[i][cidx] = (cidx <= i) ? s / Ls[i][i] : (T)0;
        }
        __syncthreads();
    }
    T* W = Winv + (size_t)kb * TS * TS;
    for (int e = tid; e < TS * TS; e += 256) W[e] = Xs[e >> 6][e & 63];
}

// nflag: only pivots of rows < nflag are checked (rows past the source pixels hold the data row
// of the augmented factorization and identity padding)
void launch_chol_diag(hipStream_t s, double* A, int ld, int kb, double* diag, int* status, double* Winv, int nflag)
{
    hipLaunchKernelGGL(k_chol_diag<double>, dim3(1), dim3(256), 0, s, A, ld, kb, diag, status, Winv, nflag);
}

void launch_chol_diag_f32(hipStream_t s, float* A, int ld, int kb, float* diag, int* status, float* Winv, int nflag)
{
    hipLaunchKernelGGL(k_chol_diag<float>, dim3(1), dim3(256), 0, s, A, ld, kb, diag, status, Winv, nflag);
}

}  // namespace qlx
