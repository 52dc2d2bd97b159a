// Learned NFW deflection surrogate: MLP 4 -> 128 x4 (SiLU) -> 5 with BF16 weights and FP32
// accumulation on v_mfma_f32_16x16x32_bf16, trained on the GPU (forward, backward, Adam).
//
// It replaces the reference's tabulate-and-interpolate option for numerical profiles
// (Tabulated_Model / QTabulated_Model, models.cpp:4054-5148): the reference grid-samples the
// deflection and Hessian of a profile in (log r, phi[, q]) and interpolates; here a small network
// is fitted to quadrature samples of the NFW integrals instead and evaluated in the ray shooting.
//
// Network (lens frame, ks factored out; x = r cos phi, y = r sin phi):
//   inputs  f = (ln r, phi folded into [0, pi/2], q, ln rs), each mapped affinely to [-1, 1]
//   outputs o = (ln J0, ln J1, 2 x^2 K0 / J0, 2 |x y| K1 / sqrt(J0 J1), 2 y^2 K2 / J1), each
//           standardized by the training-set mean and standard deviation
// from which alpha_x = x J0, alpha_y = y J1, psi_xx = J0 + 2 x^2 K0, psi_xy = 2 x y K1,
// psi_yy = J1 + 2 y^2 K2 (profile.cpp:2532-2560) follow exactly (times ks).
//
// Parameter vector (FP32 master copy, the order of qlx_surrogate_get_params):
//   W1[128][4] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][128] b4[128]
//   W5[5][128] b5[5]                                                      = 50821 values
// Every weight enters the arithmetic rounded to BF16 (round-to-nearest-even); biases are FP32.
// Activations are rounded to BF16 where they feed an MFMA; every sum accumulates in FP32.
#include "qlx_internal.h"

namespace qlx {

typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

#define NH 128
#define HP 136   // LDS/row pitch (elements) of 128-wide BF16 rows: 272 bytes spreads 16 rows over all banks
#define SB 64    // samples per workgroup

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
// inference SiLU on the hardware exp2 / reciprocal (about 1 ulp each; the result is rounded to
// BF16 right after, 2^-9 relative)
__device__ __forceinline__ float silu_hw(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------------------------------------
// 64 x 128 x 128 BF16 product on MFMA: C[s][o] = sum_k A[s][k] B[o][k] (rows hold k, pitch HP).
// Wave w computes columns [32w, 32w+32) for all 64 rows: acc[st][ot], st = 16-row tile, ot = 16-col
// tile; accumulator register v of lane l holds C[16 st + 4(l/16) + v][32w + 16 ot + l%16].
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void mm_64x128(const uint16_t* A, const uint16_t* B, float4_t acc[4][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int st = 0; st < 4; st++)
#pragma unroll
        for (int ot = 0; ot < 2; ot++) acc[st][ot] = (float4_t){0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < NH; k0 += 32) {
        const int kq = k0 + 8 * (lane >> 4);
        bf16x8_t a[4], b[2];
#pragma unroll
        for (int st = 0; st < 4; st++) a[st] = *(const bf16x8_t*)&A[(16 * st + (lane & 15)) * HP + kq];
#pragma unroll
        for (int ot = 0; ot < 2; ot++) b[ot] = *(const bf16x8_t*)&B[(32 * w + 16 * ot + (lane & 15)) * HP + kq];
#pragma unroll
        for (int st = 0; st < 4; st++)
#pragma unroll
            for (int ot = 0; ot < 2; ot++) acc[st][ot] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[st], b[ot], acc[st][ot], 0, 0, 0);
    }
}

// stage a 128 x 128 BF16 matrix (row-major, dense) into LDS rows of pitch HP
__device__ __forceinline__ void stage_w(const uint16_t* W, uint16_t* Ws)
{
    for (int e = threadIdx.x; e < NH * NH / 8; e += blockDim.x) {
        const int r = e >> 4, c8 = (e & 15) * 8;
        *(uint4*)&Ws[r * HP + c8] = *(const uint4*)&W[r * NH + c8];
    }
}

// ------------------------------------------------------------------------------------------------
// fused forward of SB samples whose standardized features sit in fs[SB][4] (LDS); network
// outputs (standardized) to out[SB][5] (LDS).  Optionally stores z of every hidden layer and the
// BF16 activations to global (training).
// ------------------------------------------------------------------------------------------------

#define P_W1 0
#define P_B1 512
#define P_W2 640
#define P_B2 17024
#define P_W3 17152
#define P_B3 33536
#define P_W4 33664
#define P_B4 50048
#define P_W5 50176
#define P_B5 50816
#define P_TOTAL 50821

__device__ void mlp_forward(const SurDev& net, const float (*fs)[4], float (*out)[5], uint16_t* Ha, uint16_t* Hb, uint16_t* Ws,
                            float* zsave, uint16_t* hsave, int s0, int nvalid)
{
    const int tid = threadIdx.x;
    // layer 1 (4 -> 128) on the VALU: thread (s = tid % 64, quarter = tid / 64) forms 32 outputs
    {
        const int s = tid & 63, q4 = tid >> 6;
        for (int o = 32 * q4; o < 32 * q4 + 32; o++) {
            const float* w = net.W1r + o * 4;
            float z = net.P[P_B1 + o];
            z = fmaf(w[0], fs[s][0], z);
            z = fmaf(w[1], fs[s][1], z);
            z = fmaf(w[2], fs[s][2], z);
            z = fmaf(w[3], fs[s][3], z);
            const uint16_t h = bf16_rne(silu(z));
            Ha[s * HP + o] = h;
            if (zsave && s < nvalid) {
                zsave[(size_t)(s0 + s) * (4 * NH) + o] = z;
                hsave[((size_t)(s0 + s)) * (4 * NH) + o] = h;
            }
        }
    }
    uint16_t* Hin = Ha;
    uint16_t* Hout = Hb;
    const int lane = tid & 63, w = tid >> 6;
    for (int layer = 0; layer < 3; layer++) {
        __syncthreads();
        stage_w(net.Wb + layer * NH * NH, Ws);
        __syncthreads();
        float4_t acc[4][2];
        mm_64x128(Hin, Ws, acc);
        const float* b = net.P + (layer == 0 ? P_B2 : layer == 1 ? P_B3 : P_B4);
#pragma unroll
        for (int st = 0; st < 4; st++)
#pragma unroll
            for (int ot = 0; ot < 2; ot++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int s = 16 * st + 4 * (lane >> 4) + v;
                    const int o = 32 * w + 16 * ot + (lane & 15);
                    const float z = acc[st][ot][v] + b[o];
                    const uint16_t h = bf16_rne(silu(z));
                    Hout[s * HP + o] = h;
                    if (zsave && s < nvalid) {
                        zsave[(size_t)(s0 + s) * (4 * NH) + (layer + 1) * NH + o] = z;
                        hsave[(size_t)(s0 + s) * (4 * NH) + (layer + 1) * NH + o] = h;
                    }
                }
        uint16_t* t = Hin;
        Hin = Hout;
        Hout = t;
    }
    __syncthreads();
    // output layer (128 -> 5) on the VALU
    for (int e = tid; e < SB * 5; e += blockDim.x) {
        const int s = e / 5, o = e - 5 * s;
        const float* w5 = net.W5r + o * NH;
        float y = net.P[P_B5 + o];
        for (int k = 0; k < NH; k++) y = fmaf(w5[k], bf16f(Hin[s * HP + k]), y);
        out[s][o] = y;
    }
    __syncthreads();
}

// feature mapping of raw (ln r, phi, q, ln rs) (phi already folded to [0, pi/2]): training box
// ln r in [ln 1e-3, ln 20], q in [0.3, 1], ln rs in [ln 1, ln 20]
#define SUR_LNR_LO (-6.907755278982137)
#define SUR_LNR_HI (2.995732273553991)
#define SUR_Q_LO 0.3
#define SUR_Q_HI 1.0
#define SUR_LRS_LO 0.0
#define SUR_LRS_HI 2.995732273553991
__device__ __forceinline__ void features(double lnr, double phi, double q, double lnrs, float* f)
{
    f[0] = (float)(2.0 * (lnr - SUR_LNR_LO) / (SUR_LNR_HI - SUR_LNR_LO) - 1.0);
    f[1] = (float)(phi * (4.0 / 3.14159265358979323846) - 1.0);
    f[2] = (float)(2.0 * (q - SUR_Q_LO) / (SUR_Q_HI - SUR_Q_LO) - 1.0);
    f[3] = (float)(2.0 * (lnrs - SUR_LRS_LO) / (SUR_LRS_HI - SUR_LRS_LO) - 1.0);
}

// FP32 feature mapping for the ray path (the network's own error is ~1e-3 relative)
__device__ __forceinline__ void features_f(float x, float y, float q, float lnrs, float* f)
{
    const float lnr = 0.5f * __logf(x * x + y * y);
    const float phi = atan2f(fabsf(y), fabsf(x));
    f[0] = (float)(2.0 / (SUR_LNR_HI - SUR_LNR_LO)) * (lnr - (float)SUR_LNR_LO) - 1.0f;
    f[1] = phi * (float)(4.0 / 3.14159265358979323846) - 1.0f;
    f[2] = (float)(2.0 / (SUR_Q_HI - SUR_Q_LO)) * (q - (float)SUR_Q_LO) - 1.0f;
    f[3] = (float)(2.0 / (SUR_LRS_HI - SUR_LRS_LO)) * (lnrs - (float)SUR_LRS_LO) - 1.0f;
}

// network outputs (standardized) -> (alpha_x, alpha_y, psi_xx, psi_xy, psi_yy) / ks
__device__ __forceinline__ void decode(const float* o, const float* norm, double x, double y, double* d5)
{
    double t[5];
    for (int k = 0; k < 5; k++) t[k] = (double)o[k] * (double)norm[5 + k] + (double)norm[k];
    const double j0 = exp(t[0]), j1 = exp(t[1]);
    const double sxy = (x * y >= 0) ? 1.0 : -1.0;
    d5[0] = x * j0;
    d5[1] = y * j1;
    d5[2] = j0 * (1.0 + t[2]);
    d5[3] = sxy * sqrt(j0 * j1) * t[3];
    d5[4] = j1 * (1.0 + t[4]);
}

// ------------------------------------------------------------------------------------------------
// Inference: 8 waves run tiles of SBI samples through all layers.  Wave w produces outputs
// [16w, 16w+16) of every hidden layer, so it only ever needs those 16 rows of W2..W4: it keeps its
// MFMA B operands of the three layers and of the output layer (5 rows, padded to 16) in registers
// for the whole kernel, and LDS holds only the two activation tiles.
// ------------------------------------------------------------------------------------------------
#define SBI 128  // samples per inference tile
#define INF_LDS_ELEMS (2 * SBI * HP)
struct InferRegs {
    bf16x8_t w[3][4];  // W2..W4: row 16w + lane%16, k = 32 ks + 8 (lane/16) .. + 7
    bf16x8_t w5[4];    // output layer row lane%16 (zero for rows >= 5)
};

__device__ __forceinline__ void infer_load_regs(const SurDev& net, InferRegs& R)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int o = 16 * w + (lane & 15), o5 = lane & 15;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        const int k = 32 * ks + 8 * (lane >> 4);
#pragma unroll
        for (int l = 0; l < 3; l++) R.w[l][ks] = *(const bf16x8_t*)&net.Wb[l * NH * NH + o * NH + k];
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 8; e++) h[e] = (o5 < 5) ? bf16_rne(net.W5r[o5 * NH + k + e]) : (uint16_t)0;
        R.w5[ks] = __builtin_bit_cast(bf16x8_t, make_uint4(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16),
                                                           h[4] | ((unsigned)h[5] << 16), h[6] | ((unsigned)h[7] << 16)));
    }
}

// one tile of SBI samples with 512 threads (8 waves); fs: standardized features; out: standardized
// network outputs; bias: b1..b4 (4 x 128) and b5 staged in LDS
__device__ void infer_tile(const SurDev& net, const InferRegs& R, const float* bias, uint16_t* Ha, uint16_t* Hb,
                           const float (*fs)[4], float (*out)[5])
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        // first layer: thread (sample tid % SBI, outputs 32 (tid / SBI) .. + 31)
        const int s = tid & (SBI - 1), oq = tid / SBI;
        const float f0 = fs[s][0], f1 = fs[s][1], f2 = fs[s][2], f3 = fs[s][3];
        for (int o = 32 * oq; o < 32 * oq + 32; o++) {
            const float* wr = net.W1r + o * 4;
            float z = bias[o];
            z = fmaf(wr[0], f0, z);
            z = fmaf(wr[1], f1, z);
            z = fmaf(wr[2], f2, z);
            z = fmaf(wr[3], f3, z);
            Ha[s * HP + o] = bf16_rne(silu_hw(z));
        }
    }
    uint16_t* Hin = Ha;
    uint16_t* Hout = Hb;
#pragma unroll
    for (int layer = 0; layer < 3; layer++) {
        __syncthreads();
        float4_t acc[SBI / 16];
#pragma unroll
        for (int st = 0; st < SBI / 16; st++) acc[st] = (float4_t){0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kq = 32 * ks + 8 * (lane >> 4);
#pragma unroll
            for (int st = 0; st < SBI / 16; st++) {
                const bf16x8_t a = *(const bf16x8_t*)&Hin[(16 * st + (lane & 15)) * HP + kq];
                acc[st] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, R.w[layer][ks], acc[st], 0, 0, 0);
            }
        }
        const int o = 16 * w + (lane & 15);
        const float bo = bias[(layer + 1) * NH + o];
#pragma unroll
        for (int st = 0; st < SBI / 16; st++)
#pragma unroll
            for (int v = 0; v < 4; v++) Hout[(16 * st + 4 * (lane >> 4) + v) * HP + o] = bf16_rne(silu_hw(acc[st][v] + bo));
        uint16_t* t = Hin;
        Hin = Hout;
        Hout = t;
    }
    __syncthreads();
    // output layer: wave w takes samples [16w, 16w+16), 16 (5 used) outputs, K = 128
    {
        float4_t acc = (float4_t){0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kq = 32 * ks + 8 * (lane >> 4);
            const bf16x8_t a = *(const bf16x8_t*)&Hin[(16 * w + (lane & 15)) * HP + kq];
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, R.w5[ks], acc, 0, 0, 0);
        }
        const int o = lane & 15;
        if (o < 5)
#pragma unroll
            for (int v = 0; v < 4; v++) out[16 * w + 4 * (lane >> 4) + v][o] = acc[v] + bias[4 * NH + o];
    }
    __syncthreads();
}

__device__ void infer_stage_bias(const SurDev& net, float* bias)
{
    for (int e = threadIdx.x; e < 4 * NH + 5; e += blockDim.x) {
        const int l = e / NH, o = e - l * NH;
        bias[e] = (e >= 4 * NH) ? net.P[P_B5 + (e - 4 * NH)] : net.P[(l == 0 ? P_B1 : l == 1 ? P_B2 : l == 2 ? P_B3 : P_B4) + o];
    }
}

// ------------------------------------------------------------------------------------------------
// inference on raw lens-frame samples: in4 = (ln r, phi, q, ln rs), out5 = (alpha, psi) / ks
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(512) k_sur_eval(SurDev net, int n, const float* in4, float* out5)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_inf[];
    uint16_t* Ha = lds_inf;
    uint16_t* Hb = Ha + SBI * HP;
    __shared__ float fs[SBI][4];
    __shared__ float os[SBI][5];
    __shared__ double raw[SBI][2];
    __shared__ float bias[4 * NH + 5];
    const int tid = threadIdx.x;
    InferRegs R;
    infer_load_regs(net, R);
    infer_stage_bias(net, bias);
    const int ntile = (n + SBI - 1) / SBI;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int s0 = tile * SBI;
        __syncthreads();
        if (tid < SBI) {
            const int i = s0 + tid;
            double lnr = 0, phi = 0, q = 1, lnrs = 0;
            if (i < n) {
                lnr = in4[4 * i];
                phi = in4[4 * i + 1];
                q = in4[4 * i + 2];
                lnrs = in4[4 * i + 3];
            }
            const double r = exp(lnr);
            const double x = r * cos(phi), y = r * sin(phi);
            raw[tid][0] = x;
            raw[tid][1] = y;
            features(lnr, atan2(fabs(y), fabs(x)), q, lnrs, fs[tid]);
        }
        __syncthreads();
        infer_tile(net, R, bias, Ha, Hb, fs, os);
        if (tid < SBI && s0 + tid < n) {
            double d5[5];
            decode(os[tid], net.norm, raw[tid][0], raw[tid][1], d5);
            for (int k = 0; k < 5; k++) out5[5 * (s0 + tid) + k] = (float)d5[k];
        }
    }
}

static void sur_allow_lds(const void* k)
{
    static const void* done[4];
    static int nd = 0;
    for (int i = 0; i < nd; i++)
        if (done[i] == k) return;
    hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, INF_LDS_ELEMS * 2);
    if (nd < 4) done[nd++] = k;
}

void launch_sur_eval(hipStream_t s, const SurDev& net, int n, const float* in4, float* out5)
{
    if (n <= 0) return;
    sur_allow_lds((const void*)k_sur_eval);
    const int ntile = (n + SBI - 1) / SBI;
    hipLaunchKernelGGL(k_sur_eval, dim3(ntile < 1024 ? ntile : 1024), dim3(512), INF_LDS_ELEMS * 2, s, net, n, in4, out5);
}

// ------------------------------------------------------------------------------------------------
// deflection of every ray for the items whose model has a surrogate lens: image-frame FP32 values
// in sur_alpha (NaN outside the training box, where the ray shooting uses quadrature).
// Shift and rotation follow LensProfile::deflection (profile.cpp:1782-1805).
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(512) k_sur_rays(SurDev net, BatchDev bd, float* sur_alpha)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_inf[];
    uint16_t* Ha = lds_inf;
    uint16_t* Hb = Ha + SBI * HP;
    __shared__ float fs[SBI][4];
    __shared__ float os[SBI][5];
    __shared__ double raw[SBI][2];
    __shared__ int inbox[SBI];
    __shared__ float bias[4 * NH + 5];
    const int tid = threadIdx.x;
    InferRegs R;
    infer_load_regs(net, R);
    infer_stage_bias(net, bias);
    const long long nray = (long long)bd.nx * bd.ny * bd.nsub;
    const long long tpi = (nray + SBI - 1) / SBI;  // tiles per item
    for (long long tile = blockIdx.x; tile < tpi * bd.B; tile += gridDim.x) {
        const int item = (int)(tile / tpi);
        const ModelPrep& m = bd.models[item];
        if (m.sur_lens < 0) continue;  // uniform over the workgroup
        const LensPrep& L = m.lens[m.sur_lens];
        const long long t0 = (tile - (long long)item * tpi) * SBI;
        __syncthreads();
        if (tid < SBI) {
            const int t = (int)(t0 + tid);  // ray index within the item (the image holds < 2^24 pixels x 16 rays)
            int ok = 0;
            double x = 0, y = 0;
            if (t < nray) {
                const int k = t % bd.nsub;
                const int p = t / bd.nsub;
                const int i = p / bd.ny, j = p - i * bd.ny;
                if (!bd.mask || bd.mask[p]) {
                    double sx, sy;
                    if (bd.nsplit == 0) {  // the pixel-centre ray of the ray shooting kernel
                        sx = bd.xvals[i] + 0.5 * ((bd.xvals[bd.nx] - bd.xvals[0]) / bd.nx);
                        sy = bd.yvals[j] + 0.5 * ((bd.yvals[bd.ny] - bd.yvals[0]) / bd.ny);
                    } else {
                        const int ii = k / bd.nsplit, jj = k % bd.nsplit;
                        const double u0 = ((double)(1 + 2 * ii)) / (2 * bd.nsplit);
                        const double w0 = ((double)(1 + 2 * jj)) / (2 * bd.nsplit);
                        sx = (1 - u0) * bd.xvals[i] + u0 * bd.xvals[i + 1];
                        sy = (1 - w0) * bd.yvals[j] + w0 * bd.yvals[j + 1];
                    }
                    x = sx - L.xc;
                    y = sy - L.yc;
                    if (L.rotate) {
                        const double xp = x * L.costh + y * L.sinth;
                        y = -x * L.sinth + y * L.costh;
                        x = xp;
                    }
                    const double r2 = x * x + y * y;
                    ok = (r2 >= 1e-6 && r2 <= 400.0 && L.q >= SUR_Q_LO && L.q <= SUR_Q_HI && L.rs >= 1.0 && L.rs <= 20.0) ? 1 : 0;
                    if (ok) features_f((float)x, (float)y, (float)L.q, logf((float)L.rs), fs[tid]);
                }
            }
            if (!ok) fs[tid][0] = fs[tid][1] = fs[tid][2] = fs[tid][3] = 0.0f;
            raw[tid][0] = x;
            raw[tid][1] = y;
            inbox[tid] = ok;
        }
        __syncthreads();
        infer_tile(net, R, bias, Ha, Hb, fs, os);
        if (tid < SBI && t0 + tid < nray) {
            float ax = __int_as_float(0x7fc00000), ay = ax;  // quiet NaN: quadrature fallback
            if (inbox[tid]) {
                const float j0 = __expf(os[tid][0] * net.norm[5] + net.norm[0]);
                const float j1 = __expf(os[tid][1] * net.norm[6] + net.norm[1]);
                double a0 = L.ks * (raw[tid][0] * (double)j0), a1 = L.ks * (raw[tid][1] * (double)j1);
                if (L.rotate) {
                    const double xp = a0 * L.costh - a1 * L.sinth;
                    a1 = a0 * L.sinth + a1 * L.costh;
                    a0 = xp;
                }
                ax = (float)a0;
                ay = (float)a1;
            }
            const size_t ro = (size_t)item * bd.s_pix * bd.nsub + t0 + tid;
            sur_alpha[2 * ro] = ax;
            sur_alpha[2 * ro + 1] = ay;
        }
    }
}

void launch_sur_rays(hipStream_t s, const SurDev& net, const BatchDev& bd, float* sur_alpha)
{
    sur_allow_lds((const void*)k_sur_rays);
    const long long nray = (long long)bd.nx * bd.ny * bd.nsub;
    const long long tiles = (nray + SBI - 1) / SBI * bd.B;
    const unsigned grid = (unsigned)(tiles < 1024 ? tiles : 1024);
    hipLaunchKernelGGL(k_sur_rays, dim3(grid), dim3(512), INF_LDS_ELEMS * 2, s, net, bd, sur_alpha);
}

// ------------------------------------------------------------------------------------------------
// training
// ------------------------------------------------------------------------------------------------
// minibatch gather: sample i of step `step` (device counter) is a hash of (seed, step, i)
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ void k_sur_gather(int n, int nsamp, unsigned long long seed, const int* step, const float* feat, const float* targ,
                             float* bf, float* bt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long h = mix64(seed ^ mix64(((unsigned long long)(*step) << 32) | (unsigned)i));
    const int k = (int)(h % (unsigned long long)nsamp);
    for (int c = 0; c < 4; c++) bf[4 * i + c] = feat[4 * k + c];
    for (int c = 0; c < 5; c++) bt[5 * i + c] = targ[5 * k + c];
}

// forward + loss + output-layer backward for SB samples per workgroup.  Saves z and BF16 h of
// the hidden layers, writes dz4 (BF16, the input of the first MFMA backward layer), the per
// workgroup loss and the per-workgroup partial gradients of W5 and b5.
__global__ void __launch_bounds__(256) k_sur_fwd(SurDev net, int n, const float* bf, const float* bt, float* zs, uint16_t* hs,
                                                 uint16_t* dz, float* loss_part, float* g5_part)
{
    __shared__ __attribute__((aligned(16))) uint16_t Ha[SB * HP], Hb[SB * HP], Ws[NH * HP];
    __shared__ float fs[SB][4];
    __shared__ float os[SB][5];
    __shared__ float dy[SB][5];
    __shared__ float red[256];
    const int s0 = blockIdx.x * SB, tid = threadIdx.x;
    const int nvalid = min(SB, n - s0);
    if (tid < SB) {
        for (int c = 0; c < 4; c++) fs[tid][c] = (tid < nvalid) ? bf[4 * (s0 + tid) + c] : 0.0f;
    }
    __syncthreads();
    mlp_forward(net, fs, os, Ha, Hb, Ws, zs, hs, s0, nvalid);
    // mean-squared error over n samples x 5 outputs
    const float scale = 2.0f / (5.0f * (float)n);
    float lsum = 0.0f;
    for (int e = tid; e < SB * 5; e += 256) {
        const int s = e / 5, o = e - 5 * s;
        float d = 0.0f;
        if (s < nvalid) {
            const float r = os[s][o] - bt[5 * (s0 + s) + o];
            lsum += r * r;
            d = scale * r;
        }
        dy[s][o] = d;
    }
    red[tid] = lsum;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) red[tid] += red[tid + st];
        __syncthreads();
    }
    if (tid == 0) loss_part[blockIdx.x] = red[0] / (5.0f * (float)n);
    // hidden layer 4's activations are in Hb after three swaps (Ha -> Hb -> Ha -> Hb)
    const uint16_t* H4 = Hb;
    // partial dW5[o][k] = sum_s dy[s][o] h4[s][k], db5[o] = sum_s dy[s][o]
    for (int e = tid; e < 5 * NH + 5; e += 256) {
        float g = 0.0f;
        if (e < 5 * NH) {
            const int o = e / NH, k = e - NH * o;
            for (int s = 0; s < SB; s++) g = fmaf(dy[s][o], bf16f(H4[s * HP + k]), g);
        } else {
            const int o = e - 5 * NH;
            for (int s = 0; s < SB; s++) g += dy[s][o];
        }
        g5_part[(size_t)blockIdx.x * (5 * NH + 5) + e] = g;
    }
    // dh4 = W5^T dy, dz4 = dh4 * silu'(z4) -> BF16
    for (int e = tid; e < SB * NH; e += 256) {
        const int s = e / NH, k = e - NH * s;
        if (s >= nvalid) continue;
        float g = 0.0f;
        for (int o = 0; o < 5; o++) g = fmaf(net.W5r[o * NH + k], dy[s][o], g);
        const float z = zs[(size_t)(s0 + s) * (4 * NH) + 3 * NH + k];
        dz[(size_t)(s0 + s) * NH + k] = bf16_rne(g * dsilu(z));
    }
}

// backward through one 128x128 layer (W = W2, W3 or W4 by index wi = 0, 1, 2) for SB samples:
//   partial dW[o][i] = sum_s dz[s][o] h_prev[s][i], db[o] = sum_s dz[s][o]
//   dh_prev = dz W  ->  dz_prev = dh_prev * silu'(z_prev)  (BF16, or FP32 for the first layer)
__global__ void __launch_bounds__(256) k_sur_bwd(SurDev net, int n, int wi, const float* zs, const uint16_t* hs, const uint16_t* dz_in,
                                                 uint16_t* dz_out, float* dz1_out, float* gW_part, float* gb_part)
{
    __shared__ __attribute__((aligned(16))) uint16_t Dz[SB * HP];    // [s][o]
    __shared__ __attribute__((aligned(16))) uint16_t DzT[NH * 72];   // [o][s], pitch 72
    __shared__ __attribute__((aligned(16))) uint16_t HT[NH * 72];    // [i][s], pitch 72
    __shared__ __attribute__((aligned(16))) uint16_t Ws[NH * HP];    // W^T rows: [i][o]
    const int s0 = blockIdx.x * SB, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nvalid = min(SB, n - s0);
    const int prev = wi;  // hidden layer index (0..3) of h_prev within the saved rows
    for (int e = tid; e < SB * NH; e += 256) {
        const int s = e / NH, k = e - NH * s;
        const uint16_t d = (s < nvalid) ? dz_in[(size_t)(s0 + s) * NH + k] : (uint16_t)0;
        const uint16_t h = (s < nvalid) ? hs[(size_t)(s0 + s) * (4 * NH) + prev * NH + k] : (uint16_t)0;
        Dz[s * HP + k] = d;
        DzT[k * 72 + s] = d;
        HT[k * 72 + s] = h;
    }
    stage_w(net.WTb + wi * NH * NH, Ws);
    __syncthreads();
    // weight gradient: 128 x 128 output, K = 64 samples; wave w takes rows o in [32w, 32w+32)
    {
        float4_t acc[2][8];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 8; c++) acc[a][c] = (float4_t){0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < SB; k0 += 32) {
            const int kq = k0 + 8 * (lane >> 4);
            bf16x8_t av[2], bv[8];
#pragma unroll
            for (int a = 0; a < 2; a++) av[a] = *(const bf16x8_t*)&DzT[(32 * w + 16 * a + (lane & 15)) * 72 + kq];
#pragma unroll
            for (int c = 0; c < 8; c++) bv[c] = *(const bf16x8_t*)&HT[(16 * c + (lane & 15)) * 72 + kq];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int c = 0; c < 8; c++) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[c], acc[a][c], 0, 0, 0);
        }
        float* g = gW_part + (size_t)blockIdx.x * NH * NH;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int o = 32 * w + 16 * a + 4 * (lane >> 4) + v;
                    const int i = 16 * c + (lane & 15);
                    g[o * NH + i] = acc[a][c][v];
                }
    }
    if (tid < NH) {
        float b = 0.0f;
        for (int s = 0; s < SB; s++) b += bf16f(Dz[s * HP + tid]);
        gb_part[(size_t)blockIdx.x * NH + tid] = b;
    }
    // data gradient: dh_prev[s][i] = sum_o dz[s][o] W[o][i] = sum_o Dz[s][o] WT[i][o]
    float4_t acc[4][2];
    mm_64x128(Dz, Ws, acc);
#pragma unroll
    for (int st = 0; st < 4; st++)
#pragma unroll
        for (int ot = 0; ot < 2; ot++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int s = 16 * st + 4 * (lane >> 4) + v;
                const int i = 32 * w + 16 * ot + (lane & 15);
                if (s >= nvalid) continue;
                const float z = zs[(size_t)(s0 + s) * (4 * NH) + prev * NH + i];
                const float g = acc[st][ot][v] * dsilu(z);
                if (dz1_out) dz1_out[(size_t)(s0 + s) * NH + i] = g;
                else dz_out[(size_t)(s0 + s) * NH + i] = bf16_rne(g);
            }
}

// first layer weight gradient: partial dW1[o][c] = sum_s dz1[s][o] f[s][c], db1[o] = sum_s dz1[s][o]
__global__ void __launch_bounds__(128) k_sur_bwd1(int n, const float* bf, const float* dz1, float* g1_part)
{
    const int s0 = blockIdx.x * SB, o = threadIdx.x;
    const int nvalid = min(SB, n - s0);
    float g[5] = {0, 0, 0, 0, 0};
    for (int s = 0; s < nvalid; s++) {
        const float d = dz1[(size_t)(s0 + s) * NH + o];
        for (int c = 0; c < 4; c++) g[c] = fmaf(d, bf[4 * (s0 + s) + c], g[c]);
        g[4] += d;
    }
    float* p = g1_part + (size_t)blockIdx.x * (5 * NH);
    for (int c = 0; c < 4; c++) p[o * 4 + c] = g[c];
    p[4 * NH + o] = g[4];
}

// gradient reduction over workgroup partials (fixed order) into grad[P_TOTAL]; loss sum
__global__ void __launch_bounds__(256) k_sur_reduce(int nwg, const float* g1_part, const float* gW_part, const float* gb_part,
                                                    const float* g5_part, const float* loss_part, float* grad, float* loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float l = 0.0f;
        for (int b = 0; b < nwg; b++) l += loss_part[b];
        *loss = l;
    }
    if (p >= P_TOTAL) return;
    float g = 0.0f;
    if (p < P_W2) {  // W1, b1
        const int off = (p < P_B1) ? p : 4 * NH + (p - P_B1);
        for (int b = 0; b < nwg; b++) g += g1_part[(size_t)b * 5 * NH + off];
    } else if (p < P_W5) {
        const int rel = p - P_W2, layer = rel / (NH * NH + NH), r = rel % (NH * NH + NH);
        if (r < NH * NH)
            for (int b = 0; b < nwg; b++) g += gW_part[((size_t)layer * nwg + b) * NH * NH + r];
        else
            for (int b = 0; b < nwg; b++) g += gb_part[((size_t)layer * nwg + b) * NH + (r - NH * NH)];
    } else {
        const int r = p - P_W5;
        for (int b = 0; b < nwg; b++) g += g5_part[(size_t)b * (5 * NH + 5) + r];
    }
    grad[p] = g;
}

// Adam (Kingma & Ba 2015) with a cosine learning-rate decay; refreshes the BF16 weight copies
// and advances the device step counter.
__global__ void __launch_bounds__(256) k_sur_adam(float* P, float* m, float* v, const float* grad, int* step, int total_steps, float lr0,
                                                  float lr_min, uint16_t* Wb, uint16_t* WTb, float* W1r, float* W5r)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_TOTAL) return;
    const int t = *step + 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float frac = fminf((float)(t - 1) / (float)max(total_steps, 1), 1.0f);
    const float lr = lr_min + 0.5f * (lr0 - lr_min) * (1.0f + cosf(3.14159265f * frac));
    const float g = grad[p];
    const float mm = b1 * m[p] + (1 - b1) * g;
    const float vv = b2 * v[p] + (1 - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1 - powf(b1, (float)t)), vh = vv / (1 - powf(b2, (float)t));
    const float val = P[p] - lr * mh / (sqrtf(vh) + eps);
    P[p] = val;
    const uint16_t hb = bf16_rne(val);
    if (p < P_B1) W1r[p] = bf16f(hb);
    else if (p >= P_W5 && p < P_B5) W5r[p - P_W5] = bf16f(hb);
    else if (p >= P_W2 && p < P_W5) {
        const int rel = p - P_W2, layer = rel / (NH * NH + NH), r = rel % (NH * NH + NH);
        if (r < NH * NH) {
            const int o = r / NH, i = r % NH;
            Wb[layer * NH * NH + o * NH + i] = hb;
            WTb[layer * NH * NH + i * NH + o] = hb;
        }
    }
}

// after set_params / load: BF16 copies of the weights from the master parameters
__global__ void __launch_bounds__(256) k_sur_refresh(const float* P, uint16_t* Wb, uint16_t* WTb, float* W1r, float* W5r)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_TOTAL) return;
    const uint16_t hb = bf16_rne(P[p]);
    if (p < P_B1) W1r[p] = bf16f(hb);
    else if (p >= P_W5 && p < P_B5) W5r[p - P_W5] = bf16f(hb);
    else if (p >= P_W2 && p < P_W5) {
        const int rel = p - P_W2, layer = rel / (NH * NH + NH), r = rel % (NH * NH + NH);
        if (r < NH * NH) {
            const int o = r / NH, i = r % NH;
            Wb[layer * NH * NH + o * NH + i] = hb;
            WTb[layer * NH * NH + i * NH + o] = hb;
        }
    }
}

__global__ void k_sur_step_advance(int* step)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// training-set preparation on the GPU: quadrature integrals -> standardized network targets
// (means/std applied on the host afterwards) and features
__global__ void k_sur_targets(int n, const double* in4, const double* j5, float* feat, float* targ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in4[4 * i], y = in4[4 * i + 1], q = in4[4 * i + 2], rs = in4[4 * i + 3];
    const double r = sqrt(x * x + y * y);
    features(log(r), atan2(fabs(y), fabs(x)), q, log(rs), feat + 4 * i);
    const double J0 = j5[5 * i], J1 = j5[5 * i + 1], K0 = j5[5 * i + 2], K1 = j5[5 * i + 3], K2 = j5[5 * i + 4];
    targ[5 * i + 0] = (float)log(J0);
    targ[5 * i + 1] = (float)log(J1);
    targ[5 * i + 2] = (float)(2 * x * x * K0 / J0);
    targ[5 * i + 3] = (float)(2 * fabs(x * y) * K1 / sqrt(J0 * J1));
    targ[5 * i + 4] = (float)(2 * y * y * K2 / J1);
}

// ---- launchers used by the host training driver ----
void launch_sur_gather(hipStream_t s, int n, int nsamp, unsigned long long seed, const int* step, const float* feat, const float* targ,
                       float* bf, float* bt)
{
    hipLaunchKernelGGL(k_sur_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, nsamp, seed, step, feat, targ, bf, bt);
}

void launch_sur_fwdbwd(hipStream_t s, const SurDev& net, int n, const float* bf, const float* bt, SurWork& w)
{
    const int nwg = (n + SB - 1) / SB;
    hipLaunchKernelGGL(k_sur_fwd, dim3(nwg), dim3(256), 0, s, net, n, bf, bt, w.zs, w.hs, w.dzA, w.loss_part, w.g5_part);
    // layer 4 (W4, h_prev = h3), layer 3 (W3, h2), layer 2 (W2, h1 -> dz1 in FP32)
    hipLaunchKernelGGL(k_sur_bwd, dim3(nwg), dim3(256), 0, s, net, n, 2, (const float*)w.zs, (const uint16_t*)w.hs, (const uint16_t*)w.dzA,
                       w.dzB, (float*)nullptr, w.gW_part + (size_t)2 * nwg * NH * NH, w.gb_part + (size_t)2 * nwg * NH);
    hipLaunchKernelGGL(k_sur_bwd, dim3(nwg), dim3(256), 0, s, net, n, 1, (const float*)w.zs, (const uint16_t*)w.hs, (const uint16_t*)w.dzB,
                       w.dzA, (float*)nullptr, w.gW_part + (size_t)1 * nwg * NH * NH, w.gb_part + (size_t)1 * nwg * NH);
    hipLaunchKernelGGL(k_sur_bwd, dim3(nwg), dim3(256), 0, s, net, n, 0, (const float*)w.zs, (const uint16_t*)w.hs, (const uint16_t*)w.dzA,
                       (uint16_t*)nullptr, w.dz1, w.gW_part, w.gb_part);
    hipLaunchKernelGGL(k_sur_bwd1, dim3(nwg), dim3(NH), 0, s, n, bf, (const float*)w.dz1, w.g1_part);
    hipLaunchKernelGGL(k_sur_reduce, dim3((P_TOTAL + 255) / 256), dim3(256), 0, s, nwg, (const float*)w.g1_part, (const float*)w.gW_part,
                       (const float*)w.gb_part, (const float*)w.g5_part, (const float*)w.loss_part, w.grad, w.loss);
}

void launch_sur_adam(hipStream_t s, SurParams& sp, const float* grad, int total_steps, float lr0, float lr_min)
{
    hipLaunchKernelGGL(k_sur_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, s, sp.P, sp.m, sp.v, grad, sp.step, total_steps, lr0, lr_min,
                       sp.Wb, sp.WTb, sp.W1r, sp.W5r);
    hipLaunchKernelGGL(k_sur_step_advance, dim3(1), dim3(64), 0, s, sp.step);
}

void launch_sur_refresh(hipStream_t s, SurParams& sp)
{
    hipLaunchKernelGGL(k_sur_refresh, dim3((P_TOTAL + 255) / 256), dim3(256), 0, s, (const float*)sp.P, sp.Wb, sp.WTb, sp.W1r, sp.W5r);
}

void launch_sur_targets(hipStream_t s, int n, const double* in4, const double* j5, float* feat, float* targ)
{
    if (n > 0) hipLaunchKernelGGL(k_sur_targets, dim3((n + 255) / 256), dim3(256), 0, s, n, in4, j5, feat, targ);
}

SurDev sur_dev(const SurParams& sp)
{
    SurDev d;
    d.P = sp.P;
    d.Wb = sp.Wb;
    d.WTb = sp.WTb;
    d.W1r = sp.W1r;
    d.W5r = sp.W5r;
    d.norm = sp.norm;
    return d;
}

}  // namespace qlx
