// Ray shooting and source-plane mapping kernels for gfx950.
//
// One thread per (supersampled) image ray: deflection through the lens model (SIE and shear in
// closed form, NFW by nested Gauss-Patterson quadrature), beta = theta - alpha, then the
// 3-point interpolation mapping onto the Cartesian source grid.  Compiled with
// -ffp-contract=off so the FP64 path keeps the reference's operation order.
//
// Reference behaviour followed:
//   subpixel centres        ImagePixelGrid::set_nsplits            pixelgrid.cpp:10898-10938
//   lens equation / sum     QLens::find_sourcept, deflection       lenscalc.cpp:81-113, 318-330
//   frame shift + rotation  LensProfile::deflection, rotate        profile.cpp:1743-1805
//   SIE                     SPLE_Lens::deflection_elliptical_iso   models.cpp:208-216
//   shear                   Shear::deflection                      models.cpp:2266-2272
//   NFW                     deflection_numerical + J integrals     profile.cpp:2359-2376, 2703-2783
//   kappa                   NFW::kappa_rsq                         models.cpp:978-999
//   quadrature              LensIntegral::PattersonIntegrate       profile.cpp:3830-3881
//   mapping                 assign_source_mapping_flags_interpolate pixelgrid.cpp:2291-2432
//   L weights               calculate_Lmatrix_interpolate          pixelgrid.cpp:2501-2527
#include <map>
#include <mutex>

#include "qlx_internal.h"
#include "patterson_table.h"
#include "nfw_gtab.h"

namespace qlx {

// Patterson nodes and a dense per-level weight table WL[level][node] (0 when the node is not
// part of that level's rule).  Instead of keeping all 511 integrand values per lane (the
// reference's pat_funcs array), each new value is added at once into the running sums of every
// level that contains it, so a lane holds 9 partial sums per integral.
__constant__ double c_pat_x[QLX_PAT_N];
__constant__ double c_pat_wl[QLX_PAT_NLEV][QLX_PAT_N];
__constant__ float c_pat_x32[QLX_PAT_N];
__constant__ float c_pat_wl32[QLX_PAT_NLEV][QLX_PAT_N];

static bool g_pat_uploaded[64] = {false};

void upload_patterson(int device)
{
    if (device >= 0 && device < 64 && g_pat_uploaded[device]) return;
    static double wl[QLX_PAT_NLEV][QLX_PAT_N];
    static float wl32[QLX_PAT_NLEV][QLX_PAT_N];
    static float x32[QLX_PAT_N];
    for (int L = 0; L < QLX_PAT_NLEV; L++) {
        for (int i = 0; i < QLX_PAT_N; i++) wl[L][i] = 0.0;
        int order = (1 << (L + 1)) - 1;
        int step = (QLX_PAT_N + 1) / (order + 1);
        for (int k = 0; k < order; k++) wl[L][(k + 1) * step - 1] = QLX_PAT_W[QLX_PAT_WOFF[L] + k];
        for (int i = 0; i < QLX_PAT_N; i++) wl32[L][i] = (float)wl[L][i];
    }
    for (int i = 0; i < QLX_PAT_N; i++) x32[i] = (float)QLX_PAT_X[i];
    hipMemcpyToSymbol(HIP_SYMBOL(c_pat_x), QLX_PAT_X, sizeof(QLX_PAT_X));
    hipMemcpyToSymbol(HIP_SYMBOL(c_pat_wl), wl, sizeof(wl));
    hipMemcpyToSymbol(HIP_SYMBOL(c_pat_x32), x32, sizeof(x32));
    hipMemcpyToSymbol(HIP_SYMBOL(c_pat_wl32), wl32, sizeof(wl32));
    if (device >= 0 && device < 64) g_pat_uploaded[device] = true;
}

template <typename T> struct PatTab;
template <> struct PatTab<double> {
    __device__ static double x(int i) { return c_pat_x[i]; }
    __device__ static double w(int L, int i) { return c_pat_wl[L][i]; }
};
template <> struct PatTab<float> {
    __device__ static float x(int i) { return c_pat_x32[i]; }
    __device__ static float w(int L, int i) { return c_pat_wl32[L][i]; }
};

// kappa of the NFW at x^2 = r^2 / rs^2 (models.cpp:978-999)
template <typename T>
__device__ __forceinline__ T nfw_kappa_x(const LensPrep& L, T xsq)
{
    const T ks = (T)L.ks;
    if (xsq < (T)1e-6) return -ks * ((T)2 + log(xsq / (T)4));
    if (fabs(xsq - (T)1) < (T)1e-5) return (T)2 * ks * ((T)0.3333333333333333 - (xsq - (T)1) / (T)5.0);
    T F;
    if (xsq > (T)1.0) {
        T s = sqrt(xsq - (T)1);
        F = atan(s) / s;
    } else {
        T s = sqrt((T)1 - xsq);
        F = atanh(s) / s;
    }
    return (T)2 * ks * ((T)1 - F) / (xsq - (T)1);
}

// Mixed path: kappa = 2 ks g(x^2) with g(t) = (1 - F(t)) / (t - 1) from the piecewise polynomials
// of nfw_gtab.h (relative error below 5e-13; gen_nfw_gtab.py): the piece comes from the exponent
// and the top five mantissa bits of t, u in [0, 1) from the remaining bits, then five FMAs.  Below
// 1e-6 (the reference's logarithmic form) and outside the table the closed form is used.
__device__ __forceinline__ double nfw_kappa_gtab(const LensPrep& L, double t)
{
    if (t < 1e-6) return nfw_kappa_x<double>(L, t);
    const long long b = __double_as_longlong(t);
    const int e = (int)((b >> 52) & 0x7ff) - 1023;
    if (e < QLX_GTAB_EMIN || e >= QLX_GTAB_EMAX) return nfw_kappa_x<double>(L, t);
    const int p = (int)((b >> 47) & (QLX_GTAB_NP - 1));
    const double u = __longlong_as_double(((b & ((1LL << 47) - 1)) << 5) | (1023LL << 52)) - 1.0;
    const double* c = qlx_gtab + ((e - QLX_GTAB_EMIN) * QLX_GTAB_NP + p) * QLX_GTAB_NC;
    double g = c[5];
    g = fma(g, u, c[4]);
    g = fma(g, u, c[3]);
    g = fma(g, u, c[2]);
    g = fma(g, u, c[1]);
    g = fma(g, u, c[0]);
    return 2.0 * L.ks * g;
}
static_assert(QLX_GTAB_NP == 32 && QLX_GTAB_NC == 6, "nfw_kappa_gtab assumes 32 pieces per octave, degree 5");

template <typename T>
__device__ __forceinline__ T nfw_kappa_rsq(const LensPrep& L, T rsq)
{
    return nfw_kappa_x<T>(L, rsq / (T)L.rs2);
}

template <typename T>
__device__ __forceinline__ T nfw_kappa_rsq_deriv(const LensPrep& L, T rsq)
{
    const T ks = (T)L.ks;
    T xsq = rsq / (T)L.rs2;
    if (xsq < (T)1e-6) return -ks / rsq;
    if (fabs(xsq - (T)1.0) < (T)1e-5) return (T)2 * ks / (T)L.rs2 * ((T)-0.2 + (T)2 * (xsq - (T)1) / (T)7.0);
    T F;
    if (xsq > (T)1.0) {
        T s = sqrt(xsq - (T)1);
        F = atan(s) / s;
    } else {
        T s = sqrt((T)1 - xsq);
        F = atanh(s) / s;
    }
    return -(ks / rsq) * ((xsq * ((T)2.0 - (T)3 * F) + (T)1) / ((xsq - (T)1) * (xsq - (T)1)));
}

// J0 and J1 of one ray together: both integrands share kappa at every node; each integral keeps
// its own stop level (profile.cpp:3858: level > 1 and |I_k - I_{k-1}| < tol |I_k|), and the
// 511-point value is returned when a level never converges.  Returns the number of integrand
// evaluations this lane performed (2 per node, like two separate reference integrals would when
// their levels agree).
template <typename T>
__device__ __forceinline__ int nfw_j01(const LensPrep& L, T x, T y, T tol, T& j0, T& j1)
{
    const T xsq = x * x, ysq = y * y;
    const T eps = (T)L.eps, fsqinv = (T)L.fsqinv;
    T S0[QLX_PAT_NLEV], S1[QLX_PAT_NLEV];
#pragma unroll
    for (int k = 0; k < QLX_PAT_NLEV; k++) { S0[k] = 0; S1[k] = 0; }
    bool d0 = false, d1 = false;
    T r0 = 0, r1 = 0;
    int nev = 0;
#pragma unroll
    for (int m = 0; m < QLX_PAT_NLEV; m++) {
        const int order = (1 << (m + 1)) - 1;
        const int step = (QLX_PAT_N + 1) / (order + 1);
        for (int i = step - 1; i < QLX_PAT_N; i += 2 * step) {
            T w = (T)0.5 + (T)0.5 * PatTab<T>::x(i);
            T u = w * w;
            T qfac = (T)1 - eps * u;
            T kap = nfw_kappa_rsq<T>(L, u * (xsq + ysq / qfac) * fsqinv);
            T t = (T)2 * w * kap;
            T f0 = t / sqrt(qfac);          // pow(qfac, 0.5): both correctly rounded
            T f1 = t / pow(qfac, (T)1.5);
#pragma unroll
            for (int k = m; k < QLX_PAT_NLEV; k++) {
                T wk = PatTab<T>::w(k, i);
                S0[k] += wk * f0;
                S1[k] += wk * f1;
            }
            nev += 2;
        }
        if (m > 1) {
            if (!d0 && fabs(S0[m] - S0[m - 1]) < tol * fabs(S0[m])) { d0 = true; r0 = S0[m]; }
            if (!d1 && fabs(S1[m] - S1[m - 1]) < tol * fabs(S1[m])) { d1 = true; r1 = S1[m]; }
            if (d0 && d1) break;
        }
    }
    if (!d0) r0 = S0[QLX_PAT_NLEV - 1];
    if (!d1) r1 = S1[QLX_PAT_NLEV - 1];
    j0 = (T)L.sqrt1meps * ((T)0.5 * r0);
    j1 = (T)L.sqrt1meps * ((T)0.5 * r1);
    return nev;
}

// Node constants of one NFW lens (every ray of a batch item shares them): with w = (1 + x_i)/2,
// u = w^2 and qfac = 1 - eps u, node i holds A = u fsqinv / rs^2, B = A / qfac,
// C = 2 w / sqrt(qfac) and D = C / qfac, so that the kappa argument is A x^2 + B y^2 and the J0 and
// J1 integrands are C kappa and D kappa (profile.cpp:2703-2734 with the per-node factors hoisted).
static_assert(QLX_NFWTAB_DOUBLES == QLX_MAX_LENSES * QLX_PAT_N * 4, "NFW node table size");
__global__ void __launch_bounds__(256) k_nfw_tables(BatchDev bd)
{
    const int item = blockIdx.y / QLX_MAX_LENSES, l = blockIdx.y % QLX_MAX_LENSES;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= QLX_PAT_N) return;
    const ModelPrep& m = bd.models[item];
    if (l >= m.nlens || m.lens[l].type != QLX_LENS_NFW) return;
    const LensPrep& L = m.lens[l];
    const double w = 0.5 + 0.5 * c_pat_x[i];
    const double u = w * w;
    const double qfac = 1.0 - L.eps * u;
    const double a = u * L.fsqinv / L.rs2;
    const double c = 2.0 * w / sqrt(qfac);
    double* t = bd.nfwtab + ((size_t)(item * QLX_MAX_LENSES + l) * QLX_PAT_N + i) * 4;
    t[0] = a;
    t[1] = a / qfac;
    t[2] = c;
    t[3] = c / qfac;
}

// nfw_j01 with the node constants of k_nfw_tables (tab: this lens's QLX_PAT_N x 4 doubles).  The
// sums of the levels are formed as the nodes come in (a node of level m adds its weighted values to
// every level's sum from m on); with NL < QLX_PAT_NLEV only the first NL levels are formed and a ray
// not converged by then reports -1, to be integrated again with all levels (the sums of the first
// NL levels take the same operations in the same order either way).
template <typename T, bool KT = false, int NL = QLX_PAT_NLEV>
__device__ __forceinline__ int nfw_j01_tab(const LensPrep& L, const double* tab, T x, T y, T tol, T& j0, T& j1)
{
    const T xsq = x * x, ysq = y * y;
    T S0[NL], S1[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) { S0[k] = 0; S1[k] = 0; }
    bool d0 = false, d1 = false;
    T r0 = 0, r1 = 0;
    int nev = 0;
#pragma unroll
    for (int m = 0; m < NL; m++) {
        const int order = (1 << (m + 1)) - 1;
        const int step = (QLX_PAT_N + 1) / (order + 1);
        for (int i = step - 1; i < QLX_PAT_N; i += 2 * step) {
            const double* tn = tab + 4 * i;
            const T xs = (T)tn[0] * xsq + (T)tn[1] * ysq;
            T kap;
            if constexpr (KT) kap = nfw_kappa_gtab(L, xs);
            else kap = nfw_kappa_x<T>(L, xs);
            const T f0 = (T)tn[2] * kap;
            const T f1 = (T)tn[3] * kap;
#pragma unroll
            for (int k = m; k < NL; k++) {
                T wk = PatTab<T>::w(k, i);
                S0[k] += wk * f0;
                S1[k] += wk * f1;
            }
            nev += 2;
        }
        if (m > 1) {
            if (!d0 && fabs(S0[m] - S0[m - 1]) < tol * fabs(S0[m])) { d0 = true; r0 = S0[m]; }
            if (!d1 && fabs(S1[m] - S1[m - 1]) < tol * fabs(S1[m])) { d1 = true; r1 = S1[m]; }
            if (d0 && d1) break;
        }
    }
    if constexpr (NL < QLX_PAT_NLEV) {
        if (!(d0 && d1)) return -1;
    }
    if (!d0) r0 = S0[NL - 1];
    if (!d1) r1 = S1[NL - 1];
    j0 = (T)L.sqrt1meps * ((T)0.5 * r0);
    j1 = (T)L.sqrt1meps * ((T)0.5 * r1);
    return nev;
}

// the nine-level quadrature as a called function: a ray that needs more than seven levels is rare,
// and inlining this second copy of the unrolled level loop doubles the kernel's code (19.7 k
// instructions, beyond what the instruction cache holds)
template <typename T, bool KT>
__device__ __attribute__((noinline)) int nfw_j01_tab_all(const LensPrep& L, const double* tab, T x, T y, T tol, T& j0, T& j1)
{
    return nfw_j01_tab<T, KT>(L, tab, x, y, tol, j0, j1);
}

template <typename T>
__device__ __forceinline__ T patterson_k(const LensPrep& L, T x, T y, T tol, T nph)
{
    // single K integral (profile.cpp:2735-2764, 2785-2790) for the surrogate's Hessian targets
    const T xsq = x * x, ysq = y * y;
    const T eps = (T)L.eps, fsqinv = (T)L.fsqinv;
    T S[QLX_PAT_NLEV];
#pragma unroll
    for (int k = 0; k < QLX_PAT_NLEV; k++) S[k] = 0;
    T r = 0;
    bool done = false;
#pragma unroll
    for (int m = 0; m < QLX_PAT_NLEV; m++) {
        const int order = (1 << (m + 1)) - 1;
        const int step = (QLX_PAT_N + 1) / (order + 1);
        for (int i = step - 1; i < QLX_PAT_N; i += 2 * step) {
            T w = (T)0.5 + (T)0.5 * PatTab<T>::x(i);
            T u = w * w;
            T qfac = (T)1 - eps * u;
            T f = fsqinv * ((T)2 * w * u * nfw_kappa_rsq_deriv<T>(L, u * (xsq + ysq / qfac) * fsqinv) / pow(qfac, nph));
#pragma unroll
            for (int k = m; k < QLX_PAT_NLEV; k++) S[k] += PatTab<T>::w(k, i) * f;
        }
        if (m > 1 && fabs(S[m] - S[m - 1]) < tol * fabs(S[m])) { done = true; r = S[m]; break; }
    }
    if (!done) r = S[QLX_PAT_NLEV - 1];
    return (T)L.sqrt1meps * ((T)0.5 * r);
}

// Deflection of one lens at (x, y) (profile.cpp:1782-1805; shear models.cpp:2266-2272).  TAB: the
// NFW integrals use the node constants tab of k_nfw_tables; otherwise they are formed per node.
template <typename T, bool TAB = false, bool KT = false>
__device__ __forceinline__ void lens_deflection(const LensPrep& L, T x, T y, T tol, T& d0, T& d1, int& nev,
                                                const double* tab = nullptr)
{
    x -= (T)L.xc;
    y -= (T)L.yc;
    if (L.type == QLX_LENS_SHEAR) {
        d0 = x * (T)L.shear1 + y * (T)L.shear2;
        d1 = -y * (T)L.shear1 + x * (T)L.shear2;
        return;
    }
    if (L.rotate) {
        T xp = x * (T)L.costh + y * (T)L.sinth;
        y = -x * (T)L.sinth + y * (T)L.costh;
        x = xp;
    }
    T a0, a1;
    if (L.type == QLX_LENS_SIE) {
        if (L.q != 1.0) {
            // u = sqrt(1-q^2) and bprime*q/u are per-lens constants computed on the host
            T psi = sqrt((T)L.qsq * ((T)L.ssq + x * x) + y * y);
            T u = (T)L.sie_u;
            T fac = (T)L.sie_fac;
            a0 = fac * atan(u * x / (psi + (T)L.sprime));
            a1 = fac * atanh(u * y / (psi + (T)L.qsq * (T)L.sprime));
        } else {
            T rsq = x * x + y * y;
            T kapavg = (T)L.bprime * (sqrt((T)L.sprime * (T)L.sprime + rsq) - (T)L.sprime) / rsq;
            a0 = kapavg * x;
            a1 = kapavg * y;
        }
    } else if (L.type == QLX_LENS_SPLE) {
        if (L.q != 1.0) {
            // SPLE_Lens::deflection_elliptical_nocore (models.cpp:244-262): elliptical radius and
            // angle, then the series of deflection_angular_factor (models.cpp:334-351), omega_i =
            // -f (beta i - 1)/(beta i + 1) e^{2 i phi} omega_{i-1}, summed while |omega|^2 exceeds
            // 1e-16 |sum|^2 (1e-13 in FP32, whose rounding sits above 1e-16); at most 4096 terms
            const T R = sqrt(x * x + y * y / (T)L.qsq);
            T phi = atan(fabs(y / ((T)L.q * x)));
            if (x < 0) phi = (y < 0) ? phi - (T)M_PI : (T)M_PI - phi;
            else if (y < 0) phi = -phi;
            const T c2 = cos(2 * phi), s2 = sin(2 * phi);
            T wr = cos(phi), wi = sin(phi);
            T fr = wr, fi = wi;
            const T stol = sizeof(T) == 4 ? (T)1e-13 : (T)1e-16;
            const T ff = (T)L.sple_ff, beta = (T)L.sple_beta;
            for (int i = 1; i <= 4096; i++) {
                const T r = ff * (beta * i - 1) / (beta * i + 1);
                const T pr = -(r * c2), pi = -(r * s2);
                const T nr = pr * wr - pi * wi;
                wi = pr * wi + pi * wr;
                wr = nr;
                fr += wr;
                fi += wi;
                if (!(wr * wr + wi * wi > stol * (fr * fr + fi * fi))) break;
            }
            const T amp = (T)L.sple_fac * pow((T)L.bprime / R, (T)L.alpha - 1);
            a0 = amp * fr;
            a1 = amp * fi;
        } else {
            // spherical default with kapavg_spherical_rsq, s = 0 (profile.cpp:2281-2290, models.cpp:169-172)
            const T rsq = x * x + y * y;
            const T kapavg = (T)L.sple_balpha * pow(rsq, (T)L.sple_e) / rsq;
            a0 = kapavg * x;
            a1 = kapavg * y;
        }
    } else {  // NFW
        if (L.q != 1.0) {
            if ((fabs(x) < (T)1e-14) && (fabs(y) < (T)1e-14)) {
                a0 = 0;
                a1 = 0;
            } else {
                T j0, j1;
                if constexpr (TAB) {
                    // seven levels (127 nodes) first: the rest only for the rays that need them
                    int ne = nfw_j01_tab<T, KT, 7>(L, tab, x, y, tol, j0, j1);
                    if (ne < 0) {
                        // called in the tabulated-kappa kernel (half its code; 9.6 -> 9.0 ms), inlined
                        // in the closed-form one (called there, that kernel spilled more: 16.6 -> 19.1 ms)
                        if constexpr (KT) ne = nfw_j01_tab_all<T, KT>(L, tab, x, y, tol, j0, j1);
                        else ne = nfw_j01_tab<T, KT>(L, tab, x, y, tol, j0, j1);
                    }
                    nev += ne;
                }
                else nev += nfw_j01<T>(L, x, y, tol, j0, j1);
                a0 = x * j0;
                a1 = y * j1;
            }
        } else {
            T rsq = x * x + y * y;
            T xsq = rsq / (T)L.rs2;
            T kapavg;
            if (xsq > (T)1e-5) {
                T F;
                if (xsq > (T)1.0) { T s = sqrt(xsq - (T)1); F = atan(s) / s; }
                else if (xsq < (T)1.0) { T s = sqrt((T)1 - xsq); F = atanh(s) / s; }
                else F = (T)1.0;
                kapavg = (T)2 * (T)L.ks * ((T)2 * F + log(xsq / (T)4)) / xsq;
            } else {
                kapavg = -(T)L.ks * ((T)1 + log(xsq / (T)4));
            }
            a0 = kapavg * x;
            a1 = kapavg * y;
        }
    }
    if (L.rotate) {
        T xp = a0 * (T)L.costh - a1 * (T)L.sinth;
        a1 = a0 * (T)L.sinth + a1 * (T)L.costh;
        a0 = xp;
    }
    d0 = a0;
    d1 = a1;
}

template <typename T>
__device__ __forceinline__ void total_deflection(const ModelPrep& m, T x, T y, T tol, T& d0, T& d1, int& nev)
{
    T s0 = 0, s1 = 0;
    for (int l = 0; l < m.nlens; l++) {
        T a0, a1;
        lens_deflection<T>(m.lens[l], x, y, tol, a0, a1, nev);
        s0 += a0;
        s1 += a1;
    }
    d0 = s0;
    d1 = s1;
}

__device__ __forceinline__ int upper_index(const double* v, int n, double x)
{
    // largest i in [0, n) with v[i] <= x, for v monotone increasing and v[0] <= x < v[n]
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (v[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// One thread per (item, ray): item = blockIdx.y.  A lens listed as sur_lens takes its deflection
// from sur_alpha (image-frame FP32 values written by the surrogate kernel for the same rays).
// Three waves per SIMD: at its natural 195 FP64 registers the kernel ran two and was latency-bound
// (5.0 ms per 64 evaluations); held to 168 it spills 36 registers and runs in 4.5 ms.
// The 3-point mapping of ray (item, pixel p, slot ro) with source-plane point (bx, by) onto the
// item's Cartesian grid (bisection_search_interpolate + assign_source_mapping_flags_interpolate,
// pixelgrid.cpp:2291-2432; weights calculate_Lmatrix_interpolate 2501-2527).
__device__ __forceinline__ void map_ray(const BatchDev& bd, int item, size_t ro, long long p, double bx, double by)
{
    const double* sg = bd.sgrid + (size_t)item * QLX_SG_STRIDE;
    const double* sxv = sg;
    const double* syv = sg + QLX_SG_SYV;
    const double* scx = sg + QLX_SG_SCX;
    const double* scy = sg + QLX_SG_SCY;
    const int snx = bd.src_n[2 * item], sny = bd.src_n[2 * item + 1];
    const int nsub = bd.nsub;
    int* mc = bd.map_cell + ro * 3;
    double* mw = bd.map_w + ro * 3;
    if (!(bx >= sxv[0] && bx < sxv[snx] && by >= syv[0] && by < syv[sny])) {
        mc[0] = mc[1] = mc[2] = -1;
        return;
    }
    const int ci = upper_index(sxv, snx, bx);
    const int cj = upper_index(syv, sny, by);
    const int xi = (((bx > scx[ci]) && (ci < snx - 1)) || (ci == 0)) ? ci + 1 : ci - 1;
    const int yj = (((by > scy[cj]) && (cj < sny - 1)) || (cj == 0)) ? cj + 1 : cj - 1;
    const int c0 = ci * sny + cj, c1 = xi * sny + cj, c2 = ci * sny + yj;
    mc[0] = c0;
    mc[1] = c1;
    mc[2] = c2;
    const double weight = 1.0 / nsub;
    const double p0x = scx[ci], p0y = scy[cj];
    const double p1x = scx[xi], p1y = scy[cj];
    const double p2x = scx[ci], p2y = scy[yj];
    const double d = (p0x - p1x) * (p1y - p2y) - (p1x - p2x) * (p0y - p1y);
    mw[0] = weight * (bx * (p1y - p2y) + by * (p2x - p1x) + p1x * p2y - p1y * p2x) / d;
    mw[1] = weight * (bx * (p2y - p0y) + by * (p0x - p2x) + p0y * p2x - p0x * p2y) / d;
    mw[2] = weight * (bx * (p0y - p1y) + by * (p1x - p0x) + p0x * p1y - p0y * p1x) / d;
    // the flags are set by many rays: a store only where the flag is still clear keeps the
    // repeated partial-line writes out of the memory traffic
    int* cm = bd.cmapped + (size_t)item * bd.s_cell;
    if (!cm[c0]) cm[c0] = 1;
    if (!cm[c1]) cm[c1] = 1;
    if (!cm[c2]) cm[c2] = 1;
    int* pm = bd.pmapped + (size_t)item * bd.s_pix + p;
    if (!*pm) *pm = 1;
}

// Origin of ray k of pixel (i, j): the centre of subpixel (k / nsplit, k % nsplit) of an
// nsplit x nsplit split (ImagePixelGrid::set_nsplits, pixelgrid.cpp:10898-10938), or with nsplit = 0
// (split_imgpixels off) the pixel centre, corner plus half the pixel size (pixelgrid.cpp:9643-9666)
__device__ __forceinline__ void ray_origin(const double* xv, const double* yv, int nx, int ny, int nsplit, int i, int j, int k,
                                           double& sx, double& sy)
{
    if (nsplit == 0) {
        sx = xv[i] + 0.5 * ((xv[nx] - xv[0]) / nx);
        sy = yv[j] + 0.5 * ((yv[ny] - yv[0]) / ny);
        return;
    }
    const int ii = k / nsplit, jj = k % nsplit;
    const double u0 = ((double)(1 + 2 * ii)) / (2 * nsplit);
    const double w0 = ((double)(1 + 2 * jj)) / (2 * nsplit);
    sx = (1 - u0) * xv[i] + u0 * xv[i + 1];
    sy = (1 - w0) * yv[j] + w0 * yv[j + 1];
}

// map = 0: rays only (the source grid is not known yet; k_map_rays_b maps them afterwards)
template <typename T, bool KT = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_raytrace_map_b(BatchDev bd, T tol, int map)
{
    const int item = blockIdx.y;
    const int nsub = bd.nsub;
    const long long nray = (long long)bd.nx * bd.ny * nsub;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nray) return;
    const ModelPrep& m = bd.models[item];
    const size_t ro = (size_t)item * bd.s_pix * nsub + t;  // ray slot of this item
    const int k = (int)(t % nsub);
    const long long p = t / nsub;  // pixel id i*ny + j
    const int i = (int)(p / bd.ny), j = (int)(p % bd.ny);
    int* mc = bd.map_cell + ro * 3;
    double* mw = bd.map_w + ro * 3;
    if (bd.mask && !bd.mask[p]) {
        mc[0] = mc[1] = mc[2] = -1;
        bd.beta[ro * 2] = 0;
        bd.beta[ro * 2 + 1] = 0;
        return;
    }
    double sx, sy;
    ray_origin(bd.xvals, bd.yvals, bd.nx, bd.ny, bd.nsplit, i, j, k, sx, sy);
    T a0 = 0, a1 = 0;
    int nev = 0;
    for (int l = 0; l < m.nlens; l++) {
        T d0, d1;
        const float sa0 = (l == m.sur_lens && bd.sur_alpha) ? bd.sur_alpha[ro * 2] : __int_as_float(0x7fc00000);
        if (sa0 == sa0) {  // surrogate value present (NaN marks rays outside its training box)
            d0 = (T)sa0;
            d1 = (T)bd.sur_alpha[ro * 2 + 1];
        } else {
            lens_deflection<T, true, KT>(m.lens[l], (T)sx, (T)sy, tol, d0, d1, nev,
                                     bd.nfwtab + (size_t)(item * QLX_MAX_LENSES + l) * QLX_PAT_N * 4);
        }
        a0 += d0;
        a1 += d1;
    }
    const double bx = (double)((T)sx - a0), by = (double)((T)sy - a1);
    bd.beta[ro * 2] = bx;
    bd.beta[ro * 2 + 1] = by;
    if (nev) atomicAdd(bd.nevals + item, (unsigned long long)nev);
    if (map) map_ray(bd, item, ro, p, bx, by);
}

// the mapping of rays traced by a map = 0 pass, once each item's grid is known
__global__ void __launch_bounds__(256) k_map_rays_b(BatchDev bd)
{
    const int item = blockIdx.y;
    const long long nray = (long long)bd.nx * bd.ny * bd.nsub;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nray) return;
    const long long p = t / bd.nsub;
    if (bd.mask && !bd.mask[p]) return;  // set unmapped by the ray pass
    const size_t ro = (size_t)item * bd.s_pix * bd.nsub + t;
    map_ray(bd, item, ro, p, bd.beta[ro * 2], bd.beta[ro * 2 + 1]);
}

// Pixel-centre rays of the mask pixels (find_sourcept of center_pts, which
// calculate_sourcepts_and_areas(true) traces, pixelgrid.cpp:10323-10330, 10457-10464), for the
// automatic source-grid size; the centre is the corner plus half the pixel size
// (pixelgrid.cpp:9659-9666).  Written to out[2 (item s_pix + p)].
__global__ void __launch_bounds__(256) k_centre_rays_b(BatchDev bd, double* out)
{
    const int item = blockIdx.y;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)bd.nx * bd.ny) return;
    if (bd.mask && !bd.mask[p]) return;
    const ModelPrep& m = bd.models[item];
    const int i = (int)(p / bd.ny), j = (int)(p % bd.ny);
    const double pxl = (bd.xvals[bd.nx] - bd.xvals[0]) / bd.nx, pyl = (bd.yvals[bd.ny] - bd.yvals[0]) / bd.ny;
    const double x = bd.xvals[i] + 0.5 * pxl, y = bd.yvals[j] + 0.5 * pyl;
    double a0 = 0, a1 = 0;
    int nev = 0;
    for (int l = 0; l < m.nlens; l++) {
        double d0, d1;
        lens_deflection<double, true, false>(m.lens[l], x, y, bd.tol, d0, d1, nev,
                                             bd.nfwtab + (size_t)(item * QLX_MAX_LENSES + l) * QLX_PAT_N * 4);
        a0 += d0;
        a1 += d1;
    }
    out[2 * ((size_t)item * bd.s_pix + p)] = x - a0;
    out[2 * ((size_t)item * bd.s_pix + p) + 1] = y - a1;
}

// Automatic source grid of each item (find_optimal_sourcegrid, pixelgrid.cpp:10998-11080, with
// noise_threshold nt; find_optimal_sourcegrid_npixels, 11469-11486): the bounding box of the mask
// pixels' mean subpixel source points (pixels with data <= nt * noise skipped when nt > 0) widened
// by 10 %, or the given box; then the count of mask pixels whose centre ray lies strictly inside.
// res[8 item ..]: xmin, xmax, ymin, ymax, count.  One workgroup of 1024 per item.
__global__ void __launch_bounds__(1024) k_srcgrid_auto_b(BatchDev bd, int auto_box, double nt, double gx0, double gx1, double gy0,
                                                         double gy1, const double* crays, double* res)
{
    __shared__ double sm[4][1024];
    const int item = blockIdx.y;
    const int npix = bd.nx * bd.ny, nsub = bd.nsub;
    double xmin = 1e30, xmax = -1e30, ymin = 1e30, ymax = -1e30;
    if (auto_box) {
        for (int p = threadIdx.x; p < npix; p += 1024) {
            if (bd.mask && !bd.mask[p]) continue;
            if (nt > 0 && bd.data[p] <= nt * (1.0 / sqrt(bd.covinv[p]))) continue;
            const double* b = bd.beta + ((size_t)item * bd.s_pix * nsub + (size_t)p * nsub) * 2;
            double xs = 0, ys = 0;
            for (int k = 0; k < nsub; k++) {
                xs += b[2 * k];
                ys += b[2 * k + 1];
            }
            xs /= nsub;
            ys /= nsub;
            xmin = fmin(xmin, xs);
            xmax = fmax(xmax, xs);
            ymin = fmin(ymin, ys);
            ymax = fmax(ymax, ys);
        }
    }
    sm[0][threadIdx.x] = xmin;
    sm[1][threadIdx.x] = xmax;
    sm[2][threadIdx.x] = ymin;
    sm[3][threadIdx.x] = ymax;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + st]);
            sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + st]);
            sm[2][threadIdx.x] = fmin(sm[2][threadIdx.x], sm[2][threadIdx.x + st]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + st]);
        }
        __syncthreads();
    }
    double bx0, bx1, by0, by1;
    if (auto_box) {
        bx0 = sm[0][0];
        bx1 = sm[1][0];
        by0 = sm[2][0];
        by1 = sm[3][0];
        const double xwidth_adj = 0.1 * (bx1 - bx0), ywidth_adj = 0.1 * (by1 - by0);
        bx0 -= xwidth_adj / 2;
        bx1 += xwidth_adj / 2;
        by0 -= ywidth_adj / 2;
        by1 += ywidth_adj / 2;
    } else {
        bx0 = gx0;
        bx1 = gx1;
        by0 = gy0;
        by1 = gy1;
    }
    __syncthreads();
    int cnt = 0;
    if (crays)
        for (int p = threadIdx.x; p < npix; p += 1024) {
            if (bd.mask && !bd.mask[p]) continue;
            const double cx = crays[2 * ((size_t)item * bd.s_pix + p)], cy = crays[2 * ((size_t)item * bd.s_pix + p) + 1];
            if ((cx > bx0) && (cx < bx1) && (cy > by0) && (cy < by1)) cnt++;
        }
    sm[0][threadIdx.x] = (double)cnt;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sm[0][threadIdx.x] += sm[0][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* r = res + 8 * item;
        r[0] = bx0;
        r[1] = bx1;
        r[2] = by0;
        r[3] = by1;
        r[4] = sm[0][0];
    }
}

void launch_srcgrid_auto_b(hipStream_t s, const BatchDev& bd, int auto_box, int count_centres, double nt, const double* box,
                           double* crays, double* res)
{
    // box: the given bounds (host values), used when auto_box is 0
    if (count_centres) {
        const long long npix = (long long)bd.nx * bd.ny;
        hipLaunchKernelGGL(k_centre_rays_b, dim3((unsigned)((npix + 255) / 256), bd.B), dim3(256), 0, s, bd, crays);
    }
    hipLaunchKernelGGL(k_srcgrid_auto_b, dim3(1, bd.B), dim3(1024), 0, s, bd, auto_box, nt, box[0], box[1], box[2], box[3],
                       count_centres ? crays : nullptr, res);
}

void launch_map_rays_b(hipStream_t s, const BatchDev& bd)
{
    const long long total = (long long)bd.nx * bd.ny * bd.nsub;
    hipLaunchKernelGGL(k_map_rays_b, dim3((unsigned)((total + 255) / 256), bd.B), dim3(256), 0, s, bd);
}

// ray_fp32: FP32 ray shooting; kappa_table: FP64 rays with the NFW convergence from nfw_gtab.h
// (the mixed path); otherwise FP64 with the closed-form convergence (the parity path)
void launch_raytrace_map_b(hipStream_t stream, const BatchDev& bd, int ray_fp32, int kappa_table, int map)
{
    const long long total = (long long)bd.nx * bd.ny * bd.nsub;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_nfw_tables, dim3((QLX_PAT_N + 255) / 256, bd.B * QLX_MAX_LENSES), dim3(256), 0, stream, bd);
    if (ray_fp32)
        hipLaunchKernelGGL(k_raytrace_map_b<float>, dim3(blocks, bd.B), dim3(256), 0, stream, bd, (float)bd.tol, map);
    else if (kappa_table)
        hipLaunchKernelGGL((k_raytrace_map_b<double, true>), dim3(blocks, bd.B), dim3(256), 0, stream, bd, bd.tol, map);
    else
        hipLaunchKernelGGL(k_raytrace_map_b<double>, dim3(blocks, bd.B), dim3(256), 0, stream, bd, bd.tol, map);
}

__global__ void __launch_bounds__(256) k_raytrace_points(ModelPrep m, double tol, int n, const double* xy, double* beta)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double x = xy[2 * t], y = xy[2 * t + 1];
    double a0, a1;
    int nev = 0;
    total_deflection<double>(m, x, y, tol, a0, a1, nev);
    beta[2 * t] = x - a0;
    beta[2 * t + 1] = y - a1;
}

void launch_raytrace_points(hipStream_t stream, const ModelPrep& m, double tol, int n, const double* xy, double* beta)
{
    hipLaunchKernelGGL(k_raytrace_points, dim3((n + 255) / 256), dim3(256), 0, stream, m, tol, n, xy, beta);
}

// Staging for the allocation-free ray entry point (qlx_raytrace_noalloc): the points in
// [0, 2 * QLX_RT_SCRATCH_POINTS) and the results in [2 * QLX_RT_SCRATCH_POINTS, 4 * ...) doubles.
// Each GPU holds its own copy for the life of the process; host code must use the address
// hipGetSymbolAddress returns for the current device, never the array's name.
static __device__ __align__(256) double scratch[4 * QLX_RT_SCRATCH_POINTS];

// the device address of `scratch` on `device` (taken once per device, with that device current)
int raytrace_scratch(int device, double** out)
{
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(device);
    if (it == ptrs.end()) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return -1;
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        it = ptrs.emplace(device, (double*)p).first;
    }
    *out = it->second;
    return 0;
}

// Lens-frame NFW deflection and Hessian (profile.cpp:2532-2560): (ax, ay, psi_xx, psi_xy, psi_yy)
__global__ void __launch_bounds__(256) k_nfw_def_hess(LensPrep L, double tol, int n, const double* xy, double* out5)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double x = xy[2 * t], y = xy[2 * t + 1];
    double j0, j1;
    nfw_j01<double>(L, x, y, tol, j0, j1);
    double k0 = patterson_k<double>(L, x, y, tol, 0.5);
    double k1 = patterson_k<double>(L, x, y, tol, 1.5);
    double k2 = patterson_k<double>(L, x, y, tol, 2.5);
    out5[5 * t + 0] = x * j0;
    out5[5 * t + 1] = y * j1;
    out5[5 * t + 2] = 2 * x * x * k0 + j0;
    out5[5 * t + 3] = 2 * x * y * k1;
    out5[5 * t + 4] = 2 * y * y * k2 + j1;
}

void launch_nfw_def_hess(hipStream_t stream, const LensPrep& L, double tol, int n, const double* xy, double* out5)
{
    hipLaunchKernelGGL(k_nfw_def_hess, dim3((n + 255) / 256), dim3(256), 0, stream, L, tol, n, xy, out5);
}

// Training targets of the deflection surrogate: per sample (x, y, q, rs) in the lens frame with
// ks = 1, the integrals J0, J1, K0, K1, K2 of LensProfile::deflection_and_hessian_numerical
// (profile.cpp:2532-2560; k_integral 2735-2764), by the same Patterson quadrature at `tol`.
__global__ void __launch_bounds__(256) k_nfw_samples(int n, const double* in4, double tol, double* out5)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double x = in4[4 * t], y = in4[4 * t + 1], q = in4[4 * t + 2], rs = in4[4 * t + 3];
    LensPrep L;
    L.type = QLX_LENS_NFW;
    L.ks = 1.0;
    L.rs = rs;
    L.rs2 = rs * rs;
    L.q = q;
    L.eps = 1 - q * q;
    L.fsqinv = q;
    L.sqrt1meps = sqrt(1 - L.eps);
    double j0, j1;
    nfw_j01<double>(L, x, y, tol, j0, j1);
    out5[5 * t + 0] = j0;
    out5[5 * t + 1] = j1;
    out5[5 * t + 2] = patterson_k<double>(L, x, y, tol, 0.5);
    out5[5 * t + 3] = patterson_k<double>(L, x, y, tol, 1.5);
    out5[5 * t + 4] = patterson_k<double>(L, x, y, tol, 2.5);
}

void launch_nfw_samples(hipStream_t stream, int n, const double* in4, double tol, double* out5)
{
    if (n > 0) hipLaunchKernelGGL(k_nfw_samples, dim3((n + 255) / 256), dim3(256), 0, stream, n, in4, tol, out5);
}

// ---- mock surface brightness (find_surface_brightness split branch, pixelgrid.cpp:11835-11896;
// SB_Profile::surface_brightness + Sersic::sb_rsq, sbprofile.cpp:1346-1366, 2472-2516) ----
__global__ void __launch_bounds__(256) k_mock_sb(int nx, int ny, int nsplit, const double* xvals, const double* yvals, ModelPrep m,
                                                 double tol, int nsrc, const double* sersic, double* sb)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nx * ny) return;
    const int i = (int)(p / ny), j = (int)(p % ny);
    double sbtot = 0;
    const int nr = rays_per_pixel(nsplit);  // nsplit 0: the unsplit branch, pixelgrid.cpp:11896-11925
    for (int k = 0; k < nr; k++) {
            double sx, sy;
            ray_origin(xvals, yvals, nx, ny, nsplit, i, j, k, sx, sy);
            double a0, a1;
            int nev = 0;
            total_deflection<double>(m, sx, sy, tol, a0, a1, nev);
            double bx = sx - a0, by = sy - a1;
            for (int k = 0; k < nsrc; k++) {
                const double* S = sersic + 11 * k;  // s0, Reff, n, q, theta, xc, yc, cos, sin, b, rot
                double x = bx - S[5], y = by - S[6];
                if (S[10] != 0) {
                    double xp = x * S[7] + y * S[8];
                    y = -x * S[8] + y * S[7];
                    x = xp;
                }
                double q = S[3];
                double rsq_ell = x * x + y * y / (q * q);
                rsq_ell *= q;
                double v = S[0] * exp(-S[9] * pow(rsq_ell / (S[1] * S[1]), 0.5 / S[2]));
                if (v != 0) sbtot += v;
            }
        }
    sb[p] = 0 + sbtot / nr;
}

void launch_mock_sb(hipStream_t stream, int nx, int ny, int nsplit, const double* xvals, const double* yvals, const ModelPrep& m,
                    double tol, int nsrc, const double* sersic, double* sb)
{
    long long n = (long long)nx * ny;
    hipLaunchKernelGGL(k_mock_sb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, nx, ny, nsplit, xvals, yvals, m, tol, nsrc,
                       sersic, sb);
}

// PSF_convolution_pixel_vector direct branch over every pixel (pixelgrid.cpp:14452-14490)
__global__ void __launch_bounds__(256) k_conv_image(int nx, int ny, const double* psf, int pnx, int pny, const double* in, double* out)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nx * ny) return;
    const int k = (int)(p / ny), l = (int)(p % ny);
    const int nxh = pnx / 2, nyh = pny / 2;
    double acc = 0;
    for (int pk = 0; pk < pnx; pk++) {
        int ii = k + nxh - pk;
        if (ii < 0 || ii >= nx) continue;
        for (int pl = 0; pl < pny; pl++) {
            int jj = l + nyh - pl;
            if (jj < 0 || jj >= ny) continue;
            acc += psf[pk * pny + pl] * in[(size_t)ii * ny + jj];
        }
    }
    out[p] = acc;
}

void launch_conv_image(hipStream_t stream, int nx, int ny, const double* psf, int pnx, int pny, const double* in, double* out)
{
    long long n = (long long)nx * ny;
    hipLaunchKernelGGL(k_conv_image, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, nx, ny, psf, pnx, pny, in, out);
}

}  // namespace qlx
