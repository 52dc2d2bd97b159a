// Host side of the qlx C ABI: context, device buffers, and the batched evaluation pipeline that
// replaces the Cartesian/DENSE/curvature block of QLens::invert_image_surface_brightness_map
// (lens.cpp:12364-12615) and its chi-square loop (lens.cpp:12950-13015).
//
// A call evaluates its items in waves.  Per wave: ray shooting + active sets for every item,
// one small device->host copy of the active counts (sizes of each item's dense problem), the
// lensing rows and block-support masks, a second copy of the support-block counts (sizes of the
// compact Lb pools), then the dense stages for all items in shared launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "qlx_internal.h"

using namespace qlx;

namespace {

const double DEG2RAD = 0.017453292519943295;  // mathexpr.h:27
const double HALFPI_REF = 1.57079632679;       // mathexpr.h:10 (the reference's truncated pi/2)

std::atomic<size_t> g_dev_bytes{0};  // device bytes held by every DevBuf of this process

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) hipFree(p);
        g_dev_bytes -= bytes;
        p = nullptr;
        bytes = 0;
    }
    // At least n bytes.  A buffer is given an eighth more than asked when that fits: the sizes of
    // later calls vary a little with the proposals, and each growth frees and allocates again (a
    // large hipFree / hipMalloc pair stalls the device).  Without room for the margin, exactly n.
    bool ensure(size_t n)
    {
        if (n <= bytes && p) return true;
        release();
        if (n == 0) n = 8;
        for (size_t m : {n + n / 8, n}) {
            m = (m + 255) & ~(size_t)255;
            if (hipMalloc(&p, m) == hipSuccess) {
                bytes = m;
                g_dev_bytes += m;
                return true;
            }
            p = nullptr;
        }
        return false;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    double total_ms = 0;
    int launches = 0;
    bool pending = false;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // R's factorization beside F's (same default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    size_t hbm_budget = 0;
    // image
    int nx = 0, ny = 0;
    bool image_set = false;
    std::vector<double> h_xvals, h_yvals, h_data, h_covinv;
    double max_sqrtc = 0.0, max_dsc = 0.0;  // max sqrt(covinv) and max |data| sqrt(covinv) over the image
    std::vector<unsigned char> h_mask;
    bool has_mask = false;
    DevBuf xvals, yvals, data, covinv, mask;
    // psf
    int pnx = 0, pny = 0;
    bool psf_set = false;
    std::vector<double> h_psf;
    DevBuf psf;
    // settings
    qlx_settings set{};
    bool settings_set = false;
    // per-item inputs
    DevBuf d_models, d_sgrid, d_srcn, d_meta, d_mats;
    // per-item work (stride s_pix / s_cell)
    DevBuf beta, map_cell, map_w, pmapped, cmapped, sidx, tix, pisrc, pidx, act_pix, counts, ucol, uval, ucnt, nevals;
    DevBuf apre, eptr, csh, cpix, cval, nfwtab, pmask, hmask, rdup;
    DevBuf key, bcnt, boff, bfill, perm;
    // pools
    DevBuf lb, F, Ff, Fz, R, Rf, cvec, vec, nz, bbase, winv, model, out, status;  // Fz: residual partials  // Rf: R's copy for its factorization
    DevBuf Fraw, d_lam;  // regparam optimisation: F as formed and the per-item lambda
    DevBuf agrid;        // automatic source grids: per item box and centre-ray count
    // solutions of the last call, kept on the device for qlx_get_source: per caller item a block
    // at kept[item].off of [cells doubles: source amplitudes in the reference order]
    // [s_pix doubles: model image] [cells ints: reference active index of each cell]
    DevBuf keep;
    struct Kept {
        int ok, n_src, snx, sny;
        size_t off;
    };
    std::vector<Kept> kept;
    size_t keep_npix = 0;
    double* rt_scratch = nullptr;  // this device's staging array of qlx_raytrace_noalloc
    // fit model of qlx_loglike: base lens model and the varied parameters
    qlx_lens_model fit_base{};
    std::vector<qlx_fit_param> fit_params;
    bool fit_set = false;
    std::vector<int> last_wave;  // caller indices of the last wave's items, in wave order (their rays
                                 // stay in `beta` until the next evaluation)
    std::vector<unsigned> nzh;  // host copy of the support masks (work accounting)
    std::map<int, int> chol_tuned;  // autotuned FP32 Cholesky panel (64-blocks) per padded block count
    int last_panel = 0;
    // captured factorization sequences: (stream, MatDesc array, counts, sizes, panel) -> executable
    std::map<std::vector<long long>, hipGraphExec_t> graphs;
    int graph_captures = 0, graph_replays = 0;
    // timers
    std::map<std::string, Timer> timers;
    qlx_stats stats{};
    int stats_waves = 0;  // waves of the last evaluation
    // learned deflection surrogate (surrogate_host.cpp) and its per-ray outputs
    void* sur = nullptr;
    DevBuf sur_alpha;
};

thread_local std::string g_err;

int fail(Ctx* c, int code, const char* msg)
{
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            char buf_[512];                                                                 \
            snprintf(buf_, sizeof(buf_), "%s: %s", #expr, hipGetErrorString(e_));           \
            return fail(c, QLX_ERR_HIP, buf_);                                              \
        }                                                                                   \
    } while (0)

void timer_begin(Ctx* c, const char* name)
{
    Timer& t = c->timers[name];
    if (!t.a) {
        hipEventCreate(&t.a);
        hipEventCreate(&t.b);
    }
    hipEventRecord(t.a, c->stream);
}

void timer_end(Ctx* c, const char* name)
{
    Timer& t = c->timers[name];
    hipEventRecord(t.b, c->stream);
    t.pending = true;
}

void timers_collect(Ctx* c)
{
    for (auto& kv : c->timers) {
        Timer& t = kv.second;
        if (!t.pending) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            t.total_ms += ms;
            t.launches++;
        }
        t.pending = false;
    }
}

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

}  // namespace

namespace qlx {

hipStream_t ctx_stream(void* ctx) { return ((Ctx*)ctx)->stream; }
int ctx_fail(void* ctx, int code, const char* msg) { return fail((Ctx*)ctx, code, msg); }
void* ctx_surrogate(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c->sur) c->sur = sur_state_new();
    return c->sur;
}
const SurDev* ctx_surrogate_net(void* ctx) { return sur_state_net(((Ctx*)ctx)->sur); }

void prep_model(const qlx_lens_model& in, ModelPrep& out)
{
    memset(&out, 0, sizeof(out));
    out.nlens = in.nlens;
    out.sur_lens = -1;
    for (int l = 0; l < in.nlens; l++) {
        const qlx_lens& s = in.lens[l];
        LensPrep& L = out.lens[l];
        L.type = s.type;
        L.def_mode = s.def_mode;
        if (s.type == QLX_LENS_SIE || s.type == QLX_LENS_NFW || s.type == QLX_LENS_SPLE) {
            // LensProfile::set_angle + orient_major_axis_north (profile.cpp:1625-1637)
            double th = DEG2RAD * (s.type == QLX_LENS_SIE ? s.p[2] : s.p[3]);
            double cth = cos(th), sth = sin(th);
            L.sinth = cth;
            L.costh = -sth;
            L.rotate = (L.sinth != 0);
        }
        if (s.type == QLX_LENS_SIE || (s.type == QLX_LENS_SPLE && s.p[1] == 1.0)) {
            // SPLE_Lens::update_meta_parameters, emode 1 (models.cpp:101-110, profile.cpp:1690-1693);
            // alpha = 1 takes the isothermal formulas (models.cpp:141-147)
            const bool sple = s.type == QLX_LENS_SPLE;
            L.type = QLX_LENS_SIE;
            double b = s.p[0], q = sple ? s.p[2] : s.p[1], core = sple ? 0.0 : s.p[5];
            double fmaj = 1.0 / sqrt(q);
            L.q = q;
            L.bprime = b * fmaj;
            L.sprime = core * fmaj;
            L.qsq = q * q;
            L.ssq = L.sprime * L.sprime;
            L.sie_u = sqrt(1 - L.qsq);
            L.sie_fac = L.bprime * q / L.sie_u;
            L.xc = sple ? s.p[4] : s.p[3];
            L.yc = sple ? s.p[5] : s.p[4];
        } else if (s.type == QLX_LENS_SPLE) {
            // SPLE_Lens::update_meta_parameters (models.cpp:101-110) and the constants of
            // deflection_elliptical_nocore / deflection_angular_factor (models.cpp:244-262, 334-351)
            double b = s.p[0], alpha = s.p[1], q = s.p[2];
            L.q = q;
            L.qsq = q * q;
            L.bprime = b * (1.0 / sqrt(q));
            L.alpha = alpha;
            L.sple_ff = (1 - q) / (1 + q);
            L.sple_beta = 2.0 / (2 - alpha);
            L.sple_fac = 2 * L.bprime * q / (1 + q);
            L.sple_balpha = pow(b, alpha);  // kapavg_spherical_rsq (models.cpp:169-172), s = 0
            L.sple_e = 1 - alpha / 2;
            L.xc = s.p[4];
            L.yc = s.p[5];
        } else if (s.type == QLX_LENS_NFW) {
            if (s.def_mode == QLX_DEF_SURROGATE && out.sur_lens < 0) out.sur_lens = l;
            double q = s.p[2];
            L.ks = s.p[0];
            L.rs = s.p[1];
            L.rs2 = L.rs * L.rs;
            L.q = q;
            L.eps = 1 - q * q;  // profile.h:446
            L.fsqinv = q;       // emode 1, profile.h:448
            L.sqrt1meps = sqrt(1 - L.eps);
            L.xc = s.p[4];
            L.yc = s.p[5];
        } else if (s.type == QLX_LENS_SHEAR) {
            // Shear::update_meta_parameters (models.cpp:2212-2224)
            double th = DEG2RAD * s.p[1];
            double theta_eff = th + HALFPI_REF;
            L.shear1 = s.p[0] * cos(2 * theta_eff);
            L.shear2 = s.p[0] * sin(2 * theta_eff);
            L.xc = s.p[2];
            L.yc = s.p[3];
        } else if (s.type == QLX_LENS_SHEAR_COMP) {
            // use_shear_component_params: the components are the parameters (models.cpp:2215-2217)
            L.type = QLX_LENS_SHEAR;
            L.shear1 = s.p[0];
            L.shear2 = s.p[1];
            L.xc = s.p[2];
            L.yc = s.p[3];
        }
    }
}

}  // namespace qlx

namespace {

// Parameter values inside each profile's domain: every parameter finite, SIE q in (0, 1], NFW
// q in (0, 1] and rs > 0.  The reference itself has no such check (a sampler's prior limits keep
// proposals in range, lens.cpp:10657-10748); here a bad proposal fails only its own item.
const char* lens_params_error(const qlx_lens_model& m)
{
    for (int l = 0; l < m.nlens; l++) {
        const qlx_lens& L = m.lens[l];
        for (int i = 0; i < 6; i++)
            if (!std::isfinite(L.p[i])) return "lens parameter is not finite";
        if (L.type == QLX_LENS_SIE && !(L.p[1] > 0 && L.p[1] <= 1)) return "SIE q must be in (0,1]";
        if (L.type == QLX_LENS_NFW && !(L.p[2] > 0 && L.p[2] <= 1 && L.p[1] > 0)) return "NFW needs rs > 0 and q in (0,1]";
        if (L.type == QLX_LENS_SPLE && !(L.p[2] > 0 && L.p[2] <= 1 && L.p[1] > 0 && L.p[1] < 2 && L.p[0] > 0))
            return "SPLE needs b > 0, alpha in (0,2) and q in (0,1]";
    }
    return nullptr;
}

// Structure of a model (lens count and types, deflection modes, a loaded surrogate where one is
// asked for); with `values`, also the parameter domain (lens_params_error).
int check_model(Ctx* c, const qlx_lens_model* m, bool values = true)
{
    if (!m || m->nlens < 0 || m->nlens > QLX_MAX_LENSES) return fail(c, QLX_ERR_ARG, "invalid lens model");
    for (int l = 0; l < m->nlens; l++) {
        int t = m->lens[l].type;
        if (t != QLX_LENS_SIE && t != QLX_LENS_NFW && t != QLX_LENS_SHEAR && t != QLX_LENS_SPLE && t != QLX_LENS_SHEAR_COMP)
            return fail(c, QLX_ERR_ARG, "unknown lens type");
        if (m->lens[l].def_mode != QLX_DEF_QUADRATURE && m->lens[l].def_mode != QLX_DEF_SURROGATE)
            return fail(c, QLX_ERR_ARG, "unknown deflection mode");
    }
    if (values) {
        const char* e = lens_params_error(*m);
        if (e) return fail(c, QLX_ERR_ARG, e);
    }
    int nsur = 0;
    for (int l = 0; l < m->nlens; l++)
        if (m->lens[l].type == QLX_LENS_NFW && m->lens[l].def_mode == QLX_DEF_SURROGATE) nsur++;
    if (nsur > 1) return fail(c, QLX_ERR_ARG, "at most one surrogate lens per model");
    if (nsur && !ctx_surrogate_net(c)) return fail(c, QLX_ERR_STATE, "surrogate deflection requested but no surrogate is trained or loaded");
    return QLX_OK;
}

// Source grid corners accumulated x += xstep and 4-corner-mean centres
// (SourcePixelGrid::create_pixel_grid pixelgrid.cpp:441-455, SourcePixel ctor 761-783).
void fill_source_grid(const double* box, int snx, int sny, double* g)
{
    double* sx = g;
    double* sy = g + QLX_SG_SYV;
    double* cx = g + QLX_SG_SCX;
    double* cy = g + QLX_SG_SCY;
    const double xstep = (box[1] - box[0]) / snx, ystep = (box[3] - box[2]) / sny;
    double x = box[0], y = box[2];
    for (int i = 0; i <= snx; i++, x += xstep) sx[i] = x;
    for (int j = 0; j <= sny; j++, y += ystep) sy[j] = y;
    for (int i = 0; i < snx; i++) cx[i] = (sx[i] + sx[i] + sx[i + 1] + sx[i + 1]) / 4.0;
    for (int j = 0; j < sny; j++) cy[j] = (sy[j] + sy[j + 1] + sy[j] + sy[j + 1]) / 4.0;
}

// bytes of the per-item ray/row work arrays
size_t stage1_bytes(size_t s_pix, size_t s_cell, size_t nsub)
{
    const size_t U = 3 * nsub;
    return s_pix * nsub * (16 + 12 + 24) + s_pix * U * 24 + (size_t)csc_segments(s_pix) * s_cell * 4 + s_pix * 4 * 7 + (s_pix + 1) * 12 + s_cell * 4 * 5 + 64 + s_pix * 16 + QLX_NFWTAB_DOUBLES * 8 + s_pix * QLX_MAX_CBW * 4;
}

// upper bound of the dense bytes of one item with at most ncell active source cells
size_t dense_bytes_bound(size_t ncell, size_t snx)
{
    const size_t ldN = round_up(ncell + 1, 128), nb = ldN / 64;
    const size_t bwb = (2 * snx) / 64 + 1;
    const size_t f = ldN * ldN * 8;                     // FP64 F, or FP32 F and its formed copy
    const size_t r = 2 * nb * (bwb + 1) * 4096 * 8;     // R and the copy that is factored
    return f + r + ldN * ldN / 16 + ldN * 8 * (QLX_VEC_SLOTS + 128);  // + the residual's tile partials
}

struct WaveItem {
    int idx;       // index in the caller's arrays
    int snx, sny;  // source cells per axis (0 until an automatic size is found)
    size_t cells;  // upper bound of snx * sny (the work arrays are sized to it)
};

// The automatic cells per axis of find_optimal_sourcegrid_npixels (pixelgrid.cpp:11481-11484,
// doubled as lens.cpp:12449-12450 does): nx0 = (int)sqrt(f count dx/dy), ny0 = (int)(nx0 dy/dx).
// Returns false when the box or the count gives no finite size.
bool auto_npixels(double f, double count, const double* box, int& snx, int& sny)
{
    const double dx = box[1] - box[0], dy = box[3] - box[2];
    if (!(dx > 0) || !(dy > 0) || !std::isfinite(dx) || !std::isfinite(dy)) return false;
    const double v = sqrt(f * count * dx / dy);
    if (!(v < 1e6)) return false;
    const int nx0 = (int)v;
    const double w = nx0 * dy / dx;
    if (!(w < 1e6)) return false;
    const int ny0 = (int)w;
    snx = 2 * nx0;
    sny = 2 * ny0;
    return true;
}

double pixel_fraction(const qlx_settings& S) { return S.pixel_fraction > 0 ? S.pixel_fraction : 0.3; }

// Brent's method over log10(lambda) as the reference runs it (brents_min_method,
// pixelgrid.cpp:16849-16932), restated as a resumable state machine so that every item of a wave
// advances by one objective evaluation per batched factorization: point() is where the objective
// is wanted next, feed() takes its value.  The arithmetic and the order of the tests follow the
// reference line by line (the returned point is x, the best point so far).
struct RegoptBrent {
    static constexpr double CGOLD = 0.3819660, ZEPS = 1.0e-10, ROOTPREC = 1.0e-8;
    double ax = 0, bx = 0, tol = 0.01;
    int max_iter = 20;
    double a = 0, b = 0, xstep = 0.0, fu = 0, fwprev = 0, fw = 0, fx = 0, u = 0, wprev = 0, w = 0, x = 0, e = 0.0;
    int iter = 0, evals = 0, state = 0;  // 0: f(x) wanted, 1: f(u) wanted, 2: converged or out of iterations
    void start(double ax_, double bx_, double tol_, int max_iter_)
    {
        ax = ax_;
        bx = bx_;
        tol = tol_;
        max_iter = max_iter_;
        a = ax;
        b = bx;
        x = w = wprev = bx - CGOLD * (bx - ax);
        xstep = 0.0;
        e = 0.0;
        iter = evals = state = 0;
    }
    double point() const { return state == 1 ? u : x; }
    void feed(double f)
    {
        evals++;
        if (state == 0) {
            fw = fwprev = fx = f;
            advance();
            return;
        }
        fu = f;
        if (fu <= fx) {
            if (u >= x) a = x;
            else b = x;
            wprev = w;
            w = x;
            x = u;
            fwprev = fw;
            fw = fx;
            fx = fu;
        } else {
            if (u < x) a = u;
            else b = u;
            if (fu <= fw or w == x) {
                wprev = w;
                fwprev = fw;
                w = u;
                fw = fu;
            } else if ((fu <= fwprev) or (wprev == x) or (wprev == w)) {
                wprev = u;
                fwprev = fu;
            }
        }
        iter++;
        advance();
    }
    // the top of the reference's loop: stop, or choose the next point u
    void advance()
    {
        if (iter >= max_iter) {
            state = 2;
            return;
        }
        double tol1, tol2, p, q, r, etemp;
        const double xmid = 0.5 * (a + b);
        tol2 = 2.0 * ((tol1 = tol * fabs(x)) + ZEPS);
        if (fabs(x - xmid) <= (tol2 - 0.5 * (b - a))) {
            state = 2;
            return;
        }
        if ((w != wprev) and (fabs(e) > tol1)) {
            r = (x - w) * (fx - fwprev);
            q = (x - wprev) * (fx - fw);
            p = (x - wprev) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            etemp = e;
            e = xstep;
            if ((fabs(p) >= fabs(0.5 * q * etemp)) or (p <= q * (a - x)) or (p >= q * (b - x))) {
                xstep = CGOLD * (e = (x >= xmid ? a - x : b - x));
            } else {
                xstep = p / q;
                u = x + xstep;
                if ((u - a < tol2) or (b - u < tol2)) xstep = ((xmid - x) >= 0 ? (tol1 >= 0 ? tol1 : -tol1) : (tol1 >= 0 ? -tol1 : tol1));
            }
        } else {
            xstep = CGOLD * (e = (x >= xmid ? a - x : b - x));
        }
        if (fabs(xstep) >= ROOTPREC) u = x + xstep;
        else u = x + (xstep >= 0 ? ROOTPREC : (-ROOTPREC));
        state = 1;
    }
};

// A factorization's launch sequence (cholesky_batched) through a HIP graph: captured on the stream
// on first use of its arguments (the kernels read the matrix descriptors from device memory, so a
// replay serves new contents of the same descriptor array), then replayed as one launch.  With
// no_graphs set, or if capture fails, the launches go to the stream one by one.
int factor_on(Ctx* c, hipStream_t s, const MatDesc* mats, int n32, int n64, int max_nb, int dense, int max_bwb, int panel)
{
    if (c->set.no_graphs) {
        cholesky_batched(s, mats, n32, n64, max_nb, dense, max_bwb, panel);
        return QLX_OK;
    }
    const std::vector<long long> key = {(long long)(uintptr_t)s, (long long)(uintptr_t)mats, n32, n64, max_nb, dense, max_bwb, panel};
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        if (c->graphs.size() >= 32) {  // bound the cache: drop everything (sizes changed a lot)
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream2));
            for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
            c->graphs.clear();
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        cholesky_batched(s, mats, n32, n64, max_nb, dense, max_bwb, panel);
        HIPCHK(c, hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, QLX_ERR_HIP, "hipGraphInstantiate of a factorization failed");
        it = c->graphs.emplace(key, ex).first;
        c->graph_captures++;
    } else {
        c->graph_replays++;
    }
    HIPCHK(c, hipGraphLaunch(it->second, s));
    return QLX_OK;
}

// run_wave's answer when the wave's exact dense and Lb-pool bytes exceed the budget: the caller
// evaluates the same items again in smaller waves (an item's result does not depend on its wave)
const int WAVE_TOO_BIG = 100;

// One wave of items through the whole pipeline.  Results go to out[item.idx]; each valid item's
// solution is copied to its block of the keep buffer.  With more than one item, a wave whose
// exact dense + Lb bytes exceed `budget` returns WAVE_TOO_BIG before allocating them.
int run_wave(Ctx* c, const qlx_lens_model* models, const std::vector<WaveItem>& items_in, qlx_result* out, size_t budget)
{
    std::vector<WaveItem> items(items_in);  // automatic grids fill in their sizes
    const int B = (int)items.size();
    if (B == 0) return QLX_OK;
    const qlx_settings& S = c->set;
    const bool mixed = S.precision == QLX_PREC_MIXED;
    const int nsub = rays_per_pixel(S.nsplit);
    const size_t s_pix = (size_t)c->nx * c->ny;
    size_t s_cell = 4;
    for (auto& it : items) s_cell = std::max(s_cell, it.cells);
    const bool auto_box = S.auto_srcgrid != 0, auto_n = S.auto_src_npixels != 0;
    const size_t U = 3 * (size_t)nsub;
    hipStream_t st = c->stream;

    // ---- inputs ----
    std::vector<ModelPrep> mp(B);
    std::vector<double> sg((size_t)B * QLX_SG_STRIDE, 0.0);
    std::vector<int> srcn(2 * B);
    std::vector<double> boxes((size_t)4 * B);
    std::vector<int> grid_status(B, QLX_OK);  // an automatic grid that failed its item
    for (int b = 0; b < B; b++) {
        prep_model(models[items[b].idx], mp[b]);
        double* bx = boxes.data() + 4 * b;
        bx[0] = S.src_xmin;
        bx[1] = S.src_xmax;
        bx[2] = S.src_ymin;
        bx[3] = S.src_ymax;
        if (!auto_box && !auto_n) fill_source_grid(bx, items[b].snx, items[b].sny, sg.data() + (size_t)b * QLX_SG_STRIDE);
        srcn[2 * b] = items[b].snx;
        srcn[2 * b + 1] = items[b].sny;
    }
    bool ok = c->d_models.ensure(sizeof(ModelPrep) * B) && c->d_sgrid.ensure(sg.size() * 8) && c->d_srcn.ensure(srcn.size() * 4) &&
              c->d_meta.ensure(sizeof(ItemMeta) * B) && c->d_mats.ensure(sizeof(MatDesc) * 2 * B) &&
              c->beta.ensure(B * s_pix * nsub * 16) && c->map_cell.ensure(B * s_pix * nsub * 12) &&
              c->map_w.ensure(B * s_pix * nsub * 24) && c->pmapped.ensure(B * s_pix * 4) && c->cmapped.ensure(B * s_cell * 4) &&
              c->sidx.ensure(B * s_cell * 4) && c->tix.ensure(B * s_cell * 4) && c->pisrc.ensure(B * s_cell * 4) &&
              c->pidx.ensure(B * s_pix * 4) && c->act_pix.ensure(B * s_pix * 4) &&
              c->counts.ensure((size_t)B * 16) && c->ucol.ensure(B * s_pix * U * 4) && c->uval.ensure(B * s_pix * U * 8) &&
              c->ucnt.ensure(B * s_pix * 4) && c->nevals.ensure((size_t)B * 8) && c->key.ensure(B * s_pix * 4) &&
              c->apre.ensure(B * (s_pix + 1) * 4) && c->eptr.ensure(B * (s_pix + 1) * 4) &&
              c->csh.ensure(B * (size_t)csc_segments(s_pix) * s_cell * 4) && c->cpix.ensure(B * s_pix * U * 4) && c->cval.ensure(B * s_pix * U * 8) &&
              c->nfwtab.ensure((size_t)B * QLX_NFWTAB_DOUBLES * 8) &&
              c->pmask.ensure((size_t)B * s_pix * QLX_MAX_CBW * 4) && c->hmask.ensure((size_t)B * s_pix * QLX_MAX_CBW * 4) &&
              c->bcnt.ensure(B * (s_cell + 2) * 4) && c->boff.ensure(B * (s_cell + 2) * 4) && c->bfill.ensure(B * (s_cell + 2) * 4) &&
              c->perm.ensure(B * s_pix * 4) && c->model.ensure(B * s_pix * 16) && c->out.ensure((size_t)B * QLX_OUT_STRIDE * 8) &&
              c->status.ensure((size_t)B * 16) && c->agrid.ensure((size_t)B * 64) &&
              c->rdup.ensure((size_t)B * 12);
    if (!ok) return fail(c, QLX_ERR_NOMEM, "work buffer allocation failed");
    c->last_wave.clear();
    for (int b = 0; b < B; b++) c->last_wave.push_back(items[b].idx);
    HIPCHK(c, hipMemcpyAsync(c->d_models.p, mp.data(), sizeof(ModelPrep) * B, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->d_sgrid.p, sg.data(), sg.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->d_srcn.p, srcn.data(), srcn.size() * 4, hipMemcpyHostToDevice, st));

    BatchDev bd{};
    bd.B = B;
    bd.nx = c->nx;
    bd.ny = c->ny;
    bd.nsplit = S.nsplit;
    bd.nsub = nsub;
    bd.xvals = c->xvals.as<double>();
    bd.yvals = c->yvals.as<double>();
    bd.mask = c->has_mask ? c->mask.as<unsigned char>() : nullptr;
    bd.data = c->data.as<double>();
    bd.covinv = c->covinv.as<double>();
    bd.psf = c->psf.as<double>();
    bd.pnx = c->pnx;
    bd.pny = c->pny;
    bd.tol = S.integral_tol;
    bd.lambda = S.regparam;
    bd.models = c->d_models.as<ModelPrep>();
    bd.sgrid = c->d_sgrid.as<double>();
    bd.src_n = c->d_srcn.as<int>();
    bd.sur_alpha = nullptr;
    bd.beta = c->beta.as<double>();
    bd.map_cell = c->map_cell.as<int>();
    bd.map_w = c->map_w.as<double>();
    bd.pmapped = c->pmapped.as<int>();
    bd.cmapped = c->cmapped.as<int>();
    bd.sidx = c->sidx.as<int>();
    bd.tix = c->tix.as<int>();
    bd.pisrc = c->pisrc.as<int>();
    bd.rkey = c->rdup.as<unsigned long long>();
    bd.rrep = (int*)(c->rdup.as<char>() + (size_t)B * 8);
    bd.pidx = c->pidx.as<int>();
    bd.act_pix = c->act_pix.as<int>();
    bd.counts = c->counts.as<int>();
    bd.ucol = c->ucol.as<int>();
    bd.uval = c->uval.as<double>();
    bd.ucnt = c->ucnt.as<int>();
    bd.apre = c->apre.as<int>();
    bd.eptr = c->eptr.as<int>();
    bd.csh = c->csh.as<int>();
    bd.cpix = c->cpix.as<unsigned>();
    bd.cval = c->cval.as<double>();
    bd.nfwtab = c->nfwtab.as<double>();
    bd.pmask = c->pmask.as<unsigned>();
    bd.hmask = c->hmask.as<unsigned>();
    {
        // mixed-path pool entries are x = (blurred L entry) sqrt(c) or data sqrt(c), with
        // |blurred L entry| <= sum|psf| * max_r ||L_r||_1 <= 3 sum|psf| (per subpixel the 3-point
        // weights are 1/nsub times barycentric coordinates with |w1|, |w2| <= 1/2, 0 <= w0 <= 2); the
        // power of two s16 puts the bound on |x| at or below 2^14, inside FP16's range with a factor
        // 4 to spare
        double apsf = 0.0;
        for (double v : c->h_psf) apsf += fabs(v);
        const double xmax = std::max(3.0 * apsf * c->max_sqrtc, c->max_dsc);
        int e = 0;
        if (xmax > 0 && std::isfinite(xmax)) e = (int)std::floor(std::log2(16384.0 / xmax));
        bd.s16 = std::ldexp(1.0, std::min(60, std::max(-60, e)));
    }
    bd.key = c->key.as<int>();
    bd.bcnt = c->bcnt.as<int>();
    bd.boff = c->boff.as<int>();
    bd.bfill = c->bfill.as<int>();
    bd.perm = c->perm.as<int>();
    bd.nevals = c->nevals.as<unsigned long long>();
    bd.s_pix = (long long)s_pix;
    bd.s_cell = (long long)s_cell;
    bd.meta = c->d_meta.as<ItemMeta>();
    bd.model = c->model.as<double>();
    bd.out = c->out.as<double>();
    bd.status = c->status.as<int>();

    launch_fill_i32(st, bd.pmapped, 0, B * s_pix);
    launch_fill_i32(st, bd.cmapped, 0, B * s_cell);
    launch_fill_i32(st, bd.status, 0, (size_t)B * 4);
    launch_fill_i32(st, (int*)bd.nevals, 0, (size_t)B * 2);
    launch_fill_f64(st, bd.out, 0.0, (size_t)B * QLX_OUT_STRIDE);

    timer_begin(c, "total");
    bool any_sur = false;
    for (int b = 0; b < B; b++) any_sur = any_sur || mp[b].sur_lens >= 0;
    if (any_sur) {
        const SurDev* net = ctx_surrogate_net(c);
        if (!net) return fail(c, QLX_ERR_STATE, "surrogate not trained or loaded");
        if (!c->sur_alpha.ensure((size_t)B * s_pix * nsub * 2 * sizeof(float))) return fail(c, QLX_ERR_NOMEM, "surrogate buffer allocation failed");
        bd.sur_alpha = c->sur_alpha.as<float>();
        timer_begin(c, "surrogate");
        launch_sur_rays(st, *net, bd, c->sur_alpha.as<float>());
        timer_end(c, "surrogate");
    }
    timer_begin(c, "raytrace");
    launch_raytrace_map_b(st, bd, S.ray_precision == QLX_RAY_FP32, mixed ? 1 : 0, (auto_box || auto_n) ? 0 : 1);
    timer_end(c, "raytrace");
    if (auto_box || auto_n) {
        // each item's grid from its own rays, then the mapping of the rays onto it
        launch_srcgrid_auto_b(st, bd, auto_box ? 1 : 0, auto_n ? 1 : 0, S.noise_threshold, boxes.data(), bd.model,
                              c->agrid.as<double>());
        std::vector<double> ag((size_t)B * 8);
        HIPCHK(c, hipMemcpyAsync(ag.data(), c->agrid.p, ag.size() * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        for (int b = 0; b < B; b++) {
            double* bx = boxes.data() + 4 * b;
            if (auto_box)
                for (int k = 0; k < 4; k++) bx[k] = ag[8 * b + k];
            int snx = items[b].snx, sny = items[b].sny;
            bool good = std::isfinite(bx[0]) && std::isfinite(bx[1]) && std::isfinite(bx[2]) && std::isfinite(bx[3]) &&
                        bx[1] > bx[0] && bx[3] > bx[2];
            if (good && auto_n) good = auto_npixels(pixel_fraction(S), ag[8 * b + 4], bx, snx, sny);
            if (!good || snx < 2 || sny < 2) grid_status[b] = QLX_ERR_NOSOURCE;  // lens.cpp:12453-12458
            else if (snx > QLX_MAX_SRC_AXIS || sny > QLX_MAX_SRC_AXIS || (size_t)snx * sny > items[b].cells)
                grid_status[b] = QLX_ERR_RANGE;
            items[b].snx = grid_status[b] == QLX_OK ? snx : 0;  // no cells: nothing maps, n_src = 0
            items[b].sny = grid_status[b] == QLX_OK ? sny : 0;
            srcn[2 * b] = items[b].snx;
            srcn[2 * b + 1] = items[b].sny;
            if (grid_status[b] == QLX_OK) fill_source_grid(bx, snx, sny, sg.data() + (size_t)b * QLX_SG_STRIDE);
        }
        HIPCHK(c, hipMemcpyAsync(c->d_sgrid.p, sg.data(), sg.size() * 8, hipMemcpyHostToDevice, st));
        HIPCHK(c, hipMemcpyAsync(c->d_srcn.p, srcn.data(), srcn.size() * 4, hipMemcpyHostToDevice, st));
        launch_map_rays_b(st, bd);
    }
    launch_compact_b(st, bd);
    std::vector<int> counts((size_t)B * 4);
    HIPCHK(c, hipMemcpyAsync(counts.data(), bd.counts, (size_t)B * 16, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));

    // ---- per-item dense sizes ----
    std::vector<ItemMeta> meta(B);
    size_t f_tot = 0, r_tot = 0, cv_tot = 0, vec_tot = 0, nz_tot = 0, bb_tot = 0, w_tot = 0;
    int max_img = 0, max_src = 0, max_rows = 0, max_nkt = 0, max_ncb = 0, max_ld = 0, max_cells = 0, max_rtiles = 0;
    for (int b = 0; b < B; b++) {
        ItemMeta& M = meta[b];
        memset(&M, 0, sizeof(M));
        const int n_src = counts[4 * b], n_img = counts[4 * b + 1];
        qlx_result* r = &out[items[b].idx];
        memset(r, 0, sizeof(*r));
        r->n_src_active = n_src;
        r->n_img_active = n_img;
        M.n_src = n_src;
        M.n_img = n_img;
        M.snx = items[b].snx;
        M.sny = items[b].sny;
        r->src_nx = items[b].snx;
        r->src_ny = items[b].sny;
        r->src_xmin = boxes[4 * b];
        r->src_xmax = boxes[4 * b + 1];
        r->src_ymin = boxes[4 * b + 2];
        r->src_ymax = boxes[4 * b + 3];
        if (grid_status[b] != QLX_OK) {
            r->minus2lnE = 2e30;
            r->status = grid_status[b];
            continue;
        }
        if (n_src == 0) {  // pixelgrid.cpp:12807 -> lens.cpp:12528-12530
            r->minus2lnE = 2e30;
            r->status = QLX_ERR_NOSOURCE;
            continue;
        }
        if (n_src + 1 > QLX_MAX_LD) {
            r->minus2lnE = 2e30;
            r->status = QLX_ERR_RANGE;
            continue;
        }
        M.valid = 1;
        M.keep_off = (long long)c->kept[items[b].idx].off;
        M.cbs = mixed ? 7 : 6;  // Lb column blocks of 128 (mixed) or 64 (FP64) source pixels
        M.ldN = (int)round_up((size_t)n_src + 1, 128);  // 128: whole tiles of the trailing Cholesky update
        M.nrows = (int)std::max<size_t>(64, round_up((size_t)n_img, 64));
        M.ncb = M.ldN >> M.cbs;
        M.ncbw = (M.ncb + 31) / 32;
        M.nkt = M.nrows / 64;
        M.bwb = (2 * M.snx) / 64 + 1;
        const size_t nb = M.ldN / 64;
        M.f_off = (long long)f_tot;
        f_tot += (size_t)M.ldN * M.ldN;
        M.r_off = (long long)r_tot;
        r_tot += nb * (M.bwb + 1) * 4096;
        M.cvec_off = (long long)cv_tot;
        cv_tot += M.nrows;
        M.vec_off = (long long)vec_tot;
        vec_tot += (size_t)QLX_VEC_SLOTS * M.ldN;
        M.nz_off = (long long)nz_tot;
        nz_tot += (size_t)M.nkt * M.ncbw;
        M.bb_off = (long long)bb_tot;
        bb_tot += M.nkt + 1;
        M.w_off = (long long)w_tot;
        w_tot += (size_t)2 * M.ldN * 64;
        max_img = std::max(max_img, n_img);
        max_src = std::max(max_src, n_src);
        max_rows = std::max(max_rows, M.nrows);
        max_nkt = std::max(max_nkt, M.nkt);
        max_ncb = std::max(max_ncb, M.ncb);
        max_ld = std::max(max_ld, M.ldN);
        max_cells = std::max(max_cells, M.snx * M.sny);
        max_rtiles = std::max(max_rtiles, (int)(nb * (M.bwb + 1)));
    }
    if (max_src == 0) {
        timer_end(c, "total");
        HIPCHK(c, hipStreamSynchronize(st));
        timers_collect(c);
        return QLX_OK;
    }
    const bool opt = S.optimize_regparam != 0 && !mixed;
    const size_t fixed_bytes = (size_t)B * stage1_bytes(s_pix, s_cell, nsub);
    const size_t dense_bytes = f_tot * (opt ? 16 : 8) + r_tot * 16 + cv_tot * 8 + vec_tot * 8 + nz_tot * 4 + bb_tot * 4 + w_tot * 8 +
                               (mixed ? f_tot / 16 : 0);
    if (B > 1 && fixed_bytes + dense_bytes > budget) return WAVE_TOO_BIG;
    ok = c->R.ensure(r_tot * 8) && c->Rf.ensure(r_tot * 8) && c->cvec.ensure(cv_tot * 8) && c->vec.ensure(vec_tot * 8) && c->nz.ensure(nz_tot * 4) &&
         c->bbase.ensure(bb_tot * 4) && c->winv.ensure(w_tot * 8) && c->F.ensure(f_tot * (mixed ? 4 : 8)) &&
         (!mixed || (c->Ff.ensure(f_tot * 4) && c->Fz.ensure(f_tot / 16)));
    if (!ok) return fail(c, QLX_ERR_NOMEM, "dense buffer allocation failed (F/R)");
    bd.F = mixed ? nullptr : c->F.as<double>();
    bd.F32 = mixed ? c->F.as<float>() : nullptr;
    bd.Ff32 = mixed ? c->Ff.as<float>() : nullptr;
    bd.zres = mixed ? c->Fz.as<double>() : nullptr;
    bd.R = c->R.as<double>();
    bd.cvec = c->cvec.as<double>();
    bd.vec = c->vec.as<double>();
    bd.nz = c->nz.as<unsigned>();
    bd.bbase = c->bbase.as<int>();
    bd.winv = c->winv.as<double>();
    HIPCHK(c, hipMemcpyAsync(c->d_meta.p, meta.data(), sizeof(ItemMeta) * B, hipMemcpyHostToDevice, st));

    timer_begin(c, "lbuild");
    launch_merge_lrows_b(st, bd, max_img);
    launch_sort_rows_b(st, bd, max_img, max_src);
    launch_lcols_b(st, bd, max_img, max_src);
    launch_row_support_b(st, bd, max_nkt);
    launch_rdup_b(st, bd);
    HIPCHK(c, hipMemcpyAsync(counts.data(), bd.counts, (size_t)B * 16, hipMemcpyDeviceToHost, st));
    c->nzh.resize(nz_tot);
    HIPCHK(c, hipMemcpyAsync(c->nzh.data(), bd.nz, nz_tot * 4, hipMemcpyDeviceToHost, st));
    std::vector<int> rrep(B);
    HIPCHK(c, hipMemcpyAsync(rrep.data(), bd.rrep, (size_t)B * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    // ---- factorization descriptors: F (dense, FP64 or FP32) of every item and R (banded FP64) of
    // every distinct R: an item whose R equals an earlier item's (rrep, k_rdup_b) reads that one's
    // log-determinant and status, so R's chain launches cover the distinct matrices only ----
    std::vector<MatDesc> fm, rm;
    for (int b = 0; b < B; b++) {
        const ItemMeta& M = meta[b];
        if (!M.valid) continue;
        double* v = bd.vec + M.vec_off;
        MatDesc f{};
        f.A = mixed ? (void*)(bd.F32 + M.f_off) : (void*)(bd.F + M.f_off);
        f.ld = M.ldN;
        f.n = M.n_src;
        f.nb = M.ldN / 64;
        f.f32 = mixed ? 1 : 0;
        f.bwb = 0;
        f.diag = (void*)(v + M.ldN);
        f.status = bd.status + 4 * b;
        f.W = (void*)(bd.winv + M.w_off);
        f.A0 = mixed ? (void*)(bd.Ff32 + M.f_off) : nullptr;
        fm.push_back(f);
        if (rrep[b] != b) continue;
        MatDesc r{};
        r.A = (void*)(c->Rf.as<double>() + M.r_off);
        r.ld = 64;
        r.n = M.n_src;
        r.nb = M.ldN / 64;
        r.f32 = 0;
        r.bwb = M.bwb;
        r.diag = (void*)(v + 2 * (size_t)M.ldN);
        r.status = bd.status + 4 * b + 1;
        r.W = (void*)(bd.winv + M.w_off + (size_t)M.ldN * 64);
        rm.push_back(r);
    }
    std::vector<MatDesc> mats(fm);
    mats.insert(mats.end(), rm.begin(), rm.end());
    HIPCHK(c, hipMemcpyAsync(c->d_mats.p, mats.data(), sizeof(MatDesc) * mats.size(), hipMemcpyHostToDevice, st));
    const int nF = (int)fm.size(), nR = (int)rm.size();
    size_t lb_tot = 0;
    int max_nsup = 0;
    for (int b = 0; b < B; b++) {
        ItemMeta& M = meta[b];
        if (!M.valid) continue;
        M.nblk = counts[4 * b + 2];
        M.nsup_max = counts[4 * b + 3];
        M.lb_off = (long long)lb_tot;
        lb_tot += (size_t)M.nblk * ((size_t)64 << M.cbs);
        max_nsup = std::max(max_nsup, M.nsup_max);
    }
    if (B > 1 && fixed_bytes + dense_bytes + lb_tot * (mixed ? 2 : 8) > budget) return WAVE_TOO_BIG;
    if (!c->lb.ensure(lb_tot * (mixed ? 2 : 8))) return fail(c, QLX_ERR_NOMEM, "Lb pool allocation failed");
    bd.lb = c->lb.p;
    HIPCHK(c, hipMemcpyAsync(c->d_meta.p, meta.data(), sizeof(ItemMeta) * B, hipMemcpyHostToDevice, st));
    timer_begin(c, "lbkernel");
    if (launch_build_lb_b(st, bd, max_rows, max_nsup, mixed ? 1 : 0, mixed ? 128 : 64) != 0)
        return fail(c, QLX_ERR_RANGE, "PSF too large for the Lb build (at most 8192 taps in FP64, 16384 mixed)");
    timer_end(c, "lbkernel");
    timer_end(c, "lbuild");
    // R depends only on the active cells: it is built now and factored (ln det R only) on the second
    // stream, from a copy, while the main stream forms F (the SYRK, then k_prep_f_b, which reads R
    // itself); the streams join before the log-determinants.  (Started before the Lb build instead,
    // its small launches slow that latency-bound kernel by more than they save.)  R's many small
    // launches are enqueued after the SYRK's: the host's time spent enqueueing them then overlaps
    // the SYRK on the GPU instead of leaving the main stream idle.
    // regparam optimisation: per-item lambda on the device, F as formed kept beside the factor
    std::vector<RegoptBrent> brent(opt ? B : 0);
    std::vector<double> lamh(B, S.regparam);
    if (opt) {
        double lo = S.regopt_minlog, hi = S.regopt_maxlog;
        if (!(hi > lo)) {  // lens.cpp:424-425
            lo = -3;
            hi = 5;
        }
        const double tol = S.regopt_tol > 0 ? S.regopt_tol : 0.01;
        const int maxit = S.regopt_max_iter > 0 ? S.regopt_max_iter : 20;
        for (int b = 0; b < B; b++) {
            brent[b].start(lo, hi, tol, maxit);
            lamh[b] = pow(10, brent[b].point());  // chisq_regparam_dense: regparam = 10^logreg
        }
        if (!c->Fraw.ensure(f_tot * 8) || !c->d_lam.ensure((size_t)B * 8)) return fail(c, QLX_ERR_NOMEM, "regparam buffers");
        HIPCHK(c, hipMemcpyAsync(c->d_lam.p, lamh.data(), (size_t)B * 8, hipMemcpyHostToDevice, st));
        bd.lam = c->d_lam.as<double>();
    }
    launch_curvature_b(st, bd, max_cells, max_rtiles);  // R of the items that hold one
    HIPCHK(c, hipEventRecord(c->ev_fork, st));
    if (!mixed) launch_gather_cvec_b(st, bd, max_rows);
    timer_begin(c, "syrk");
    if (mixed) launch_syrk_f16_b(st, bd, max_ncb, max_nkt);
    else launch_syrk_f64_b(st, bd, max_ncb, max_nkt);
    timer_end(c, "syrk");
    if (opt) {
        HIPCHK(c, hipMemcpyAsync(c->Fraw.p, c->F.p, f_tot * 8, hipMemcpyDeviceToDevice, st));
        bd.Fraw = c->Fraw.as<double>();
    }
    launch_prep_fr_b(st, bd, max_ld, mixed ? 1 : 0);
    HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    HIPCHK(c, hipMemcpyAsync(c->Rf.p, c->R.p, r_tot * 8, hipMemcpyDeviceToDevice, c->stream2));
    int max_bwb = 1;
    for (const MatDesc& r : rm) max_bwb = std::max(max_bwb, r.bwb);
    {
        const int rc = factor_on(c, c->stream2, c->d_mats.as<MatDesc>() + nF, 0, nR, max_ld / 64, 0, max_bwb, 0);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));

    int panel = 0;  // FP32 panel in 64-blocks
    if (mixed) {
        const int nbk = max_ld / 64;
        if (S.chol_panel > 0) {
            panel = S.chol_panel / 64;
        } else {
            auto it = c->chol_tuned.find(nbk);
            if (it != c->chol_tuned.end()) {
                panel = it->second;
            } else {
                // autotune: each candidate width factors this wave's matrices (a repeat reads F as
                // formed again, so the runs are interchangeable); the fastest is kept
                hipEvent_t e0 = nullptr, e1 = nullptr;
                HIPCHK(c, hipEventCreate(&e0));
                HIPCHK(c, hipEventCreate(&e1));
                float best = 0.f;
                for (int cand : {8, 16, 32}) {
                    HIPCHK(c, hipEventRecord(e0, st));
                    cholesky_batched(st, c->d_mats.as<MatDesc>(), nF, 0, nbk, 1, 0, cand);
                    HIPCHK(c, hipEventRecord(e1, st));
                    HIPCHK(c, hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
                    if (panel == 0 || ms < best) {
                        best = ms;
                        panel = cand;
                    }
                }
                hipEventDestroy(e0);
                hipEventDestroy(e1);
                c->chol_tuned[nbk] = panel;
            }
        }
        c->last_panel = panel * 64;
    }
    timer_begin(c, "cholesky");
    {
        const int rc = factor_on(c, st, c->d_mats.as<MatDesc>(), mixed ? nF : 0, mixed ? 0 : nF, max_ld / 64, 1, 0, panel);
        if (rc) return rc;
    }
    timer_end(c, "cholesky");
    HIPCHK(c, hipStreamWaitEvent(st, c->ev_join, 0));
    launch_logdets_b(st, bd, mixed ? 1 : 0);
    timer_begin(c, "solve");
    launch_trisolve_b(st, bd, max_ld, mixed ? 1 : 0);
    if (mixed) {
        const int iters = S.refine_iters > 0 ? S.refine_iters : 6;
        const double rtol = S.refine_tol > 0 ? S.refine_tol : 1e-10;
        launch_refine_b(st, bd, max_ld, iters, rtol);
    }
    timer_end(c, "solve");
    launch_evidence_b(st, bd, max_img, max_src);
    if (opt) {
        // one objective evaluation per item per pass; items that converged are evaluated at their
        // optimum again, and a last pass evaluates every item there (the state the reference leaves:
        // regparam = 10^x, the source and ln det F of that evaluation)
        std::vector<double> ob((size_t)B * QLX_OUT_STRIDE);
        bool final_pass = false;
        for (;;) {
            HIPCHK(c, hipMemcpyAsync(ob.data(), bd.out, ob.size() * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
            if (final_pass) break;
            bool all_done = true;
            for (int b = 0; b < B; b++) {
                if (!meta[b].valid) continue;
                if (brent[b].state != 2) brent[b].feed(ob[(size_t)b * QLX_OUT_STRIDE + 8]);
                all_done = all_done && brent[b].state == 2;
            }
            for (int b = 0; b < B; b++) lamh[b] = meta[b].valid ? pow(10, brent[b].point()) : S.regparam;
            final_pass = all_done;
            HIPCHK(c, hipMemcpyAsync(c->d_lam.p, lamh.data(), (size_t)B * 8, hipMemcpyHostToDevice, st));
            launch_prep_fr_b(st, bd, max_ld, 0);
            timer_begin(c, "cholesky");
            {
                const int rc = factor_on(c, st, c->d_mats.as<MatDesc>(), 0, nF, max_ld / 64, 1, 0, 0);
                if (rc) return rc;
            }
            timer_end(c, "cholesky");
            launch_logdets_b(st, bd, 0);
            launch_trisolve_b(st, bd, max_ld, 0);
            launch_evidence_b(st, bd, max_img, max_src);
        }
    }
    timer_end(c, "total");
    launch_keep_b(st, bd, c->keep.p, max_cells);

    std::vector<double> o((size_t)B * QLX_OUT_STRIDE);
    std::vector<int> stat((size_t)B * 4);
    std::vector<unsigned long long> nev(B);
    HIPCHK(c, hipMemcpyAsync(o.data(), bd.out, o.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(stat.data(), bd.status, stat.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(nev.data(), bd.nevals, (size_t)B * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    timers_collect(c);
    for (int b = 0; b < B; b++) {
        if (!meta[b].valid) continue;
        qlx_result* r = &out[items[b].idx];
        const double* ob = o.data() + (size_t)b * QLX_OUT_STRIDE;
        r->minus2lnE = ob[0];
        r->chisq0 = ob[1];
        r->logdetF = ob[2];
        r->logdetR = ob[3];
        r->Es = ob[4];
        r->refine_resid = ob[6];
        r->refine_steps = (int)ob[7];
        r->regparam = lamh[b];
        r->regopt_evals = opt ? brent[b].evals : 0;
        r->status = QLX_OK;
        if (stat[4 * b + 3] == 2) r->status = QLX_ERR_REFINE;  // value kept, accurate to refine_resid
        if (stat[4 * b] || stat[4 * b + 1] || !std::isfinite(r->minus2lnE)) {
            r->status = QLX_ERR_NOT_PD;  // the reference dies here (pixelgrid.cpp:17384)
            r->minus2lnE = 2e30;
        }
        Ctx::Kept& K = c->kept[items[b].idx];
        K.ok = (r->status == QLX_OK || r->status == QLX_ERR_REFINE) ? 1 : 0;
        K.n_src = meta[b].n_src;
        K.snx = meta[b].snx;
        K.sny = meta[b].sny;
    }

    // ---- work accounting over the wave (the launches timed above) ----
    {
        const std::vector<unsigned>& nzh = c->nzh;  // read back with the support counts
        qlx_stats& T = c->stats;
        memset(&T, 0, sizeof(T));
        for (int b = 0; b < B; b++) {
            const ItemMeta& M = meta[b];
            if (!M.valid) continue;
            const double cbw = (double)(1 << M.cbs);
            const double tile = 2.0 * 64 * cbw * cbw;
            double exec_pairs = 0, nblk = 0;
            for (int kt = 0; kt < M.nkt; kt++) {
                double bits = 0;
                for (int w = 0; w < M.ncbw; w++) bits += __builtin_popcount(nzh[M.nz_off + (size_t)kt * M.ncbw + w]);
                exec_pairs += bits * (bits + 1) / 2;
                nblk += bits;
            }
            T.syrk_exec_flops += exec_pairs * tile;
            T.syrk_dense_flops += (double)M.ncb * (M.ncb + 1) / 2 * M.nkt * tile;
            T.syrk_alg_flops += (double)M.n_img * M.n_src * (M.n_src + 1) + 2.0 * M.n_img * M.n_src;
            T.lb_bytes += nblk * 64 * cbw * (mixed ? 2 : 8);
            T.integrand_evals += (double)nev[b];
            T.chol_flops += (double)M.n_src * M.n_src * M.n_src / 3.0;
            T.n_items++;
            T.n_img = M.n_img;
            T.n_src = M.n_src;
            T.ld_src = M.ldN;
            T.rows_pad = M.nrows;
        }
    }

    return QLX_OK;
}

// Evaluate nbatch items (per-item source grids) in waves sized to the HBM budget.
int eval_items(Ctx* c, int nbatch, const qlx_lens_model* models, const int* src_n, qlx_result* out)
{
    for (auto& kv : c->timers) {
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
    c->kept.assign(nbatch, Ctx::Kept{0, 0, 0, 0, 0});
    c->last_wave.clear();
    c->stats_waves = 0;
    const bool mixed = c->set.precision == QLX_PREC_MIXED;
    const int nsub = rays_per_pixel(c->set.nsplit);
    const size_t s_pix = (size_t)c->nx * c->ny;
    const qlx_settings& S = c->set;
    const bool auto_n = S.auto_src_npixels != 0;
    // automatic sizes: 2 nx0 x 2 ny0 cells with nx0 ny0 <= pixel_fraction x (mask pixels), as
    // nx0^2 <= f count dx/dy and ny0 <= nx0 dy/dx (+4 for rounding)
    size_t auto_cells = 0;
    if (auto_n) {
        size_t nmask = s_pix;
        if (c->has_mask) {
            nmask = 0;
            for (unsigned char m : c->h_mask) nmask += m ? 1 : 0;
        }
        auto_cells = std::min<size_t>((size_t)QLX_MAX_SRC_AXIS * QLX_MAX_SRC_AXIS, (size_t)(4.0 * pixel_fraction(S) * nmask) + 4);
    }
    std::vector<WaveItem> todo;
    for (int b = 0; b < nbatch; b++) {
        int rc = check_model(c, &models[b], false);
        if (rc) return rc;
        const int snx = auto_n ? 0 : src_n[2 * b], sny = auto_n ? 0 : src_n[2 * b + 1];
        if (snx > QLX_MAX_SRC_AXIS || sny > QLX_MAX_SRC_AXIS) return fail(c, QLX_ERR_RANGE, "source grid larger than 512 per axis");
        memset(&out[b], 0, sizeof(out[b]));
        out[b].src_nx = snx;
        out[b].src_ny = sny;
        out[b].src_xmin = S.src_xmin;
        out[b].src_xmax = S.src_xmax;
        out[b].src_ymin = S.src_ymin;
        out[b].src_ymax = S.src_ymax;
        if (!auto_n && (snx < 2 || sny < 2)) {  // lens.cpp:12453-12458
            out[b].minus2lnE = 2e30;
            out[b].status = QLX_ERR_NOSOURCE;
            continue;
        }
        if (lens_params_error(models[b])) {
            out[b].minus2lnE = 2e30;
            out[b].status = QLX_ERR_ARG;
            continue;
        }
        todo.push_back({b, snx, sny, auto_n ? auto_cells : (size_t)snx * sny});
    }
    // every item's solution block in the keep buffer (upper bound: n_src <= cells)
    size_t keep_tot = 0;
    for (auto& it : todo) {
        c->kept[it.idx].off = keep_tot;
        keep_tot += round_up(it.cells * 12 + s_pix * 8, 256);
    }
    c->keep_npix = s_pix;
    if (!c->keep.ensure(keep_tot)) return fail(c, QLX_ERR_NOMEM, "solution buffer allocation failed");
    // admission: items per wave from the budget (ray/row arrays + a dense bound + an Lb estimate);
    // run_wave checks the exact dense and Lb bytes and the wave is halved when they do not fit
    size_t max_cell = 4, max_snx = 2;
    for (auto& it : todo) {
        max_cell = std::max(max_cell, it.cells);
        max_snx = std::max(max_snx, auto_n ? std::min<size_t>(QLX_MAX_SRC_AXIS, it.cells) : (size_t)it.snx);
    }
    const size_t ncell_bound = std::min<size_t>(max_cell, QLX_MAX_LD - 1);
    const size_t ldN = round_up(ncell_bound + 1, 64);
    const size_t lb_est = round_up(std::max<size_t>(s_pix, 64), 64) * std::min<size_t>(ldN, 32 * 64) * (mixed ? 2 : 8);
    const size_t per_item = stage1_bytes(s_pix, max_cell, nsub) + dense_bytes_bound(ncell_bound, max_snx) + lb_est;
    size_t budget = c->hbm_budget;
    if (budget == 0) {
        // 85 % of what is free plus what this process's buffers already hold (they are reused)
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = (size_t)(0.85 * (double)(fr + g_dev_bytes.load()));
        else budget = (size_t)16 << 30;
    }
    budget = budget > keep_tot ? budget - keep_tot : 0;
    size_t wave = std::max<size_t>(1, std::min<size_t>(budget / std::max<size_t>(per_item, 1), 1024));
    size_t w0 = 0;
    while (w0 < todo.size()) {
        const size_t n = std::min(wave, todo.size() - w0);
        std::vector<WaveItem> items(todo.begin() + w0, todo.begin() + w0 + n);
        int rc = run_wave(c, models, items, out, budget);
        if (rc == WAVE_TOO_BIG) {
            wave = std::max<size_t>(1, n / 2);
            continue;
        }
        if (rc) return rc;
        c->stats_waves++;
        w0 += n;
    }
    for (int b = 0; b < nbatch; b++)
        if (out[b].status != QLX_OK) return QLX_ITEMS_FAILED;
    return QLX_OK;
}

}  // namespace

extern "C" {

int qlx_struct_sizes(size_t* sizes, int n)
{
    const size_t s[8] = {sizeof(qlx_lens),     sizeof(qlx_lens_model), sizeof(qlx_opts),          sizeof(qlx_settings),
                         sizeof(qlx_result),   sizeof(qlx_fit_param),  sizeof(qlx_surrogate_cfg), sizeof(qlx_stats)};
    for (int k = 0; k < n && k < 8 && sizes; k++) sizes[k] = s[k];
    return 8;
}

int qlx_ctx_create(const qlx_opts* opts, void** out)
{
    if (!out) return fail(nullptr, QLX_ERR_ARG, "null ctx pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, QLX_ERR_HIP, "no HIP device");
    Ctx* c = new Ctx();
    c->device = opts ? opts->device : 0;
    c->hbm_budget = opts ? opts->hbm_budget : 0;
    if (c->device < 0 || c->device >= ndev) {
        delete c;
        return fail(nullptr, QLX_ERR_ARG, "invalid device ordinal");
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return fail(nullptr, QLX_ERR_HIP, "stream creation failed");
    }
    upload_patterson(c->device);
    if (raytrace_scratch(c->device, &c->rt_scratch) != 0) {
        qlx_ctx_destroy(c);
        return fail(nullptr, QLX_ERR_HIP, "hipGetSymbolAddress of the ray staging array failed");
    }
    *out = c;
    return QLX_OK;
}

int qlx_ctx_destroy(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& kv : c->timers) {
        if (kv.second.a) hipEventDestroy(kv.second.a);
        if (kv.second.b) hipEventDestroy(kv.second.b);
    }
    hipStreamSynchronize(c->stream2);
    for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
    if (c->sur) sur_state_destroy(c->sur);
    hipEventDestroy(c->ev_fork);
    hipEventDestroy(c->ev_join);
    hipStreamDestroy(c->stream2);
    hipStreamDestroy(c->stream);
    delete c;
    return QLX_OK;
}

const char* qlx_last_error(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    return c ? c->err.c_str() : g_err.c_str();
}

int qlx_set_image(void* ctx, int nx, int ny, const double* xvals, const double* yvals, const double* data, const double* covinv,
                  const unsigned char* mask)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    c->kept.clear();  // solutions of earlier calls no longer match the inputs
    c->last_wave.clear();
    if (nx < 1 || ny < 1 || !xvals || !yvals || !data || !covinv) return fail(c, QLX_ERR_ARG, "qlx_set_image: bad arguments");
    if ((long long)nx * ny > (1LL << 24)) return fail(c, QLX_ERR_RANGE, "image too large");
    if (nx > 65535 || ny > 65535) return fail(c, QLX_ERR_RANGE, "image wider or taller than 65535 pixels");
    hipSetDevice(c->device);
    size_t npix = (size_t)nx * ny;
    c->nx = nx;
    c->ny = ny;
    c->h_xvals.assign(xvals, xvals + nx + 1);
    c->h_yvals.assign(yvals, yvals + ny + 1);
    c->h_data.assign(data, data + npix);
    c->h_covinv.assign(covinv, covinv + npix);
    c->max_sqrtc = 0.0;
    c->max_dsc = 0.0;
    for (size_t p = 0; p < npix; p++) {
        const double sc = covinv[p] > 0 ? sqrt(covinv[p]) : 0.0;
        c->max_sqrtc = std::max(c->max_sqrtc, sc);
        c->max_dsc = std::max(c->max_dsc, fabs(data[p]) * sc);
    }
    c->has_mask = (mask != nullptr);
    if (mask) c->h_mask.assign(mask, mask + npix);
    if (!c->xvals.ensure((nx + 1) * 8) || !c->yvals.ensure((ny + 1) * 8) || !c->data.ensure(npix * 8) || !c->covinv.ensure(npix * 8) ||
        !c->mask.ensure(npix))
        return fail(c, QLX_ERR_NOMEM, "image allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->xvals.p, xvals, (nx + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->yvals.p, yvals, (ny + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->data.p, data, npix * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->covinv.p, covinv, npix * 8, hipMemcpyHostToDevice, c->stream));
    if (mask) HIPCHK(c, hipMemcpyAsync(c->mask.p, mask, npix, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->image_set = true;
    return QLX_OK;
}

int qlx_set_psf(void* ctx, int pnx, int pny, const double* psf)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    c->kept.clear();  // solutions of earlier calls no longer match the inputs
    c->last_wave.clear();
    if (pnx < 1 || pny < 1 || !psf || (pnx % 2) == 0 || (pny % 2) == 0) return fail(c, QLX_ERR_ARG, "PSF must be odd-sized and non-empty");
    if ((long long)pnx * pny > 4096 * 4096) return fail(c, QLX_ERR_RANGE, "PSF too large");
    hipSetDevice(c->device);
    c->pnx = pnx;
    c->pny = pny;
    c->h_psf.assign(psf, psf + (size_t)pnx * pny);
    if (!c->psf.ensure((size_t)pnx * pny * 8)) return fail(c, QLX_ERR_NOMEM, "psf allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->psf.p, psf, (size_t)pnx * pny * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->psf_set = true;
    return QLX_OK;
}

int qlx_set_psf_gaussian(void* ctx, double sigx, double sigy, double threshold)
{
    // QLens::generate_PSF_matrix without supersampling (pixelgrid.cpp:14601-14668)
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!c->image_set) return fail(c, QLX_ERR_STATE, "set the image before a generated PSF");
    if (!(threshold > 0 && threshold < 1) || sigx <= 0 || sigy <= 0) return fail(c, QLX_ERR_ARG, "bad PSF width/threshold");
    double xstep = (c->h_xvals[c->nx] - c->h_xvals[0]) / c->nx;
    double ystep = (c->h_yvals[c->ny] - c->h_yvals[0]) / c->ny;
    double frac = sqrt(-2 * log(threshold));
    double nxd = frac * sigx / xstep, nyd = frac * sigy / ystep;
    int nxh = (int)nxd, nyh = (int)nyd;
    if ((nxd - nxh) > 0.5) nxh++;
    if ((nyd - nyh) > 0.5) nyh++;
    int nx = 2 * nxh + 1, ny = 2 * nyh + 1;
    if ((long long)nx * ny > 4096 * 4096) return fail(c, QLX_ERR_RANGE, "PSF too large");
    double xmax_sup = ((2 * nxh + 1) * xstep - xstep) / 2;
    double ymax_sup = ((2 * nyh + 1) * ystep - ystep) / 2;
    std::vector<double> psf((size_t)nx * ny);
    double norm = 0, x, y;
    int i, j;
    for (i = 0, x = -xmax_sup; i < nx; i++, x += xstep)
        for (j = 0, y = -ymax_sup; j < ny; j++, y += ystep) {
            double a = x / sigx, bb = y / sigy;
            psf[(size_t)i * ny + j] = exp(-0.5 * (a * a + bb * bb));
            norm += psf[(size_t)i * ny + j];
        }
    for (auto& v : psf) v /= norm;
    return qlx_set_psf(ctx, nx, ny, psf.data());
}

int qlx_set_settings(void* ctx, const qlx_settings* s)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    c->kept.clear();  // solutions of earlier calls no longer match the inputs
    c->last_wave.clear();
    if (!s) return fail(c, QLX_ERR_ARG, "null settings");
    if (s->nsplit < 0 || s->nsplit > 4) return fail(c, QLX_ERR_RANGE, "nsplit must be 0..4");
    if (!s->auto_srcgrid && (!(s->src_xmax > s->src_xmin) || !(s->src_ymax > s->src_ymin)))
        return fail(c, QLX_ERR_ARG, "empty source grid bounds");
    if (!(s->noise_threshold >= 0) || !(s->pixel_fraction < 1e6)) return fail(c, QLX_ERR_ARG, "bad noise threshold or pixel fraction");
    if (!s->auto_src_npixels && (s->src_nx > QLX_MAX_SRC_AXIS || s->src_ny > QLX_MAX_SRC_AXIS))
        return fail(c, QLX_ERR_RANGE, "source grid larger than 512 per axis");
    if (s->precision != QLX_PREC_FP64 && s->precision != QLX_PREC_MIXED) return fail(c, QLX_ERR_ARG, "unknown precision mode");
    if (s->ray_precision != QLX_RAY_FP64 && s->ray_precision != QLX_RAY_FP32) return fail(c, QLX_ERR_ARG, "unknown ray precision");
    if (s->integral_tol <= 0) return fail(c, QLX_ERR_ARG, "integral tolerance must be > 0");
    if (s->optimize_regparam && s->precision != QLX_PREC_FP64) return fail(c, QLX_ERR_ARG, "regparam optimisation runs in FP64 precision only");
    if (s->optimize_regparam && s->regopt_max_iter > 1000) return fail(c, QLX_ERR_RANGE, "at most 1000 regparam iterations");
    if (s->chol_panel != 0 && s->chol_panel != 512 && s->chol_panel != 1024 && s->chol_panel != 2048)
        return fail(c, QLX_ERR_ARG, "chol_panel must be 0 (autotune), 512, 1024 or 2048");
    c->set = *s;
    c->settings_set = true;
    return QLX_OK;
}

int qlx_eval(void* ctx, int nbatch, const qlx_lens_model* models, qlx_result* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (nbatch < 0 || (nbatch > 0 && (!models || !out))) return fail(c, QLX_ERR_ARG, "qlx_eval: bad arguments");
    if (!c->image_set || !c->settings_set || !c->psf_set) return fail(c, QLX_ERR_STATE, "image, PSF and settings must be set first");
    hipSetDevice(c->device);
    std::vector<int> srcn(2 * (size_t)nbatch);
    for (int b = 0; b < nbatch; b++) {
        srcn[2 * b] = c->set.src_nx;
        srcn[2 * b + 1] = c->set.src_ny;
    }
    return eval_items(c, nbatch, models, srcn.data(), out);
}

int qlx_eval_batch_grids(void* ctx, int nbatch, const qlx_lens_model* models, const int* src_n, qlx_result* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (nbatch < 0 || (nbatch > 0 && (!models || !out || (!src_n && !c->set.auto_src_npixels)))) return fail(c, QLX_ERR_ARG, "bad arguments");
    if (!c->image_set || !c->settings_set || !c->psf_set) return fail(c, QLX_ERR_STATE, "image, PSF and settings must be set first");
    hipSetDevice(c->device);
    return eval_items(c, nbatch, models, src_n, out);
}

int qlx_get_source(void* ctx, int item, int* src_nx, int* src_ny, double* src_cells, int* active_index, size_t src_cap,
                   double* model_image, size_t model_cap)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (item < 0 || item >= (int)c->kept.size()) return fail(c, QLX_ERR_ARG, "qlx_get_source: item out of range of the last call");
    const Ctx::Kept& K = c->kept[item];
    if (!K.ok) return fail(c, QLX_ERR_STATE, "qlx_get_source: the item has no solution (its status was an error)");
    const size_t ncell = (size_t)K.snx * K.sny, npix = c->keep_npix;
    if (src_nx) *src_nx = K.snx;
    if (src_ny) *src_ny = K.sny;
    if ((src_cells || active_index) && src_cap < ncell) return fail(c, QLX_ERR_ARG, "qlx_get_source: source buffer smaller than the item's grid");
    if (model_image && model_cap < npix) return fail(c, QLX_ERR_ARG, "qlx_get_source: model buffer smaller than the image");
    hipSetDevice(c->device);
    const char* k = (const char*)c->keep.p + K.off;
    std::vector<double> s((size_t)K.n_src);
    std::vector<int> sidx(ncell);
    if (src_cells) HIPCHK(c, hipMemcpyAsync(s.data(), k, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c->stream));
    if (src_cells || active_index)
        HIPCHK(c, hipMemcpyAsync(sidx.data(), k + sizeof(double) * (ncell + npix), sizeof(int) * ncell, hipMemcpyDeviceToHost, c->stream));
    if (model_image) HIPCHK(c, hipMemcpyAsync(model_image, k + sizeof(double) * ncell, sizeof(double) * npix, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (src_cells)
        for (size_t q = 0; q < ncell; q++) src_cells[q] = sidx[q] >= 0 ? s[sidx[q]] : 0.0;
    if (active_index) memcpy(active_index, sidx.data(), sizeof(int) * ncell);
    return QLX_OK;
}

int qlx_get_rays(void* ctx, int item, double* beta, size_t cap)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !beta) return QLX_ERR_ARG;
    const size_t n = (size_t)c->nx * c->ny * rays_per_pixel(c->set.nsplit) * 2;
    if (cap < n) return fail(c, QLX_ERR_ARG, "qlx_get_rays: buffer smaller than nx*ny*(rays per pixel)*2");
    int pos = -1;
    for (size_t w = 0; w < c->last_wave.size(); w++)
        if (c->last_wave[w] == item) pos = (int)w;
    if (pos < 0) return fail(c, QLX_ERR_STATE, "qlx_get_rays: the item was not in the last wave of the last evaluation");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(beta, c->beta.as<double>() + (size_t)pos * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

int qlx_get_psf(void* ctx, int* pnx, int* pny, double* psf, size_t cap)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!c->psf_set) return fail(c, QLX_ERR_STATE, "no PSF set");
    if (pnx) *pnx = c->pnx;
    if (pny) *pny = c->pny;
    if (psf) {
        if (cap < c->h_psf.size()) return fail(c, QLX_ERR_ARG, "qlx_get_psf: buffer smaller than the PSF");
        memcpy(psf, c->h_psf.data(), c->h_psf.size() * sizeof(double));
    }
    return QLX_OK;
}

int qlx_set_fit_model(void* ctx, const qlx_lens_model* base, int nparam, const qlx_fit_param* params)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!base || nparam < 0 || (nparam > 0 && !params)) return fail(c, QLX_ERR_ARG, "qlx_set_fit_model: bad arguments");
    int rc = check_model(c, base, false);
    if (rc) return rc;
    for (int k = 0; k < nparam; k++) {
        const qlx_fit_param& f = params[k];
        if (f.lens < 0 || f.lens >= base->nlens || f.param < 0 || f.param >= 6) return fail(c, QLX_ERR_ARG, "fit parameter outside the lens model");
        if (f.transform < QLX_TRANSFORM_NONE || f.transform > QLX_TRANSFORM_LINEAR || (f.transform == QLX_TRANSFORM_LINEAR && f.ta == 0))
            return fail(c, QLX_ERR_ARG, "unknown or singular parameter transform");
        if (f.prior < QLX_PRIOR_UNIFORM || f.prior > QLX_PRIOR_GAUSS || (f.prior == QLX_PRIOR_GAUSS && !(f.prior_sig > 0)) ||
            (f.prior != QLX_PRIOR_UNIFORM && !(f.prior_norm > 0)))
            return fail(c, QLX_ERR_ARG, "unknown prior, or a Gaussian prior without a positive width and norm");
    }
    c->fit_base = *base;
    c->fit_params.assign(params, params + nparam);
    c->fit_set = true;
    return QLX_OK;
}

int qlx_loglike(void* ctx, int nbatch, const double* x, double* loglike, qlx_result* out)
{
    // QLens::fitmodel_loglike_extended_source (lens.cpp:10657-10748) over a batch of sampler points
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!c->fit_set) return fail(c, QLX_ERR_STATE, "qlx_loglike: no fit model set");
    const int np = (int)c->fit_params.size();
    if (nbatch < 0 || (nbatch > 0 && (!x || !loglike))) return fail(c, QLX_ERR_ARG, "qlx_loglike: bad arguments");
    if (!c->image_set || !c->settings_set || !c->psf_set) return fail(c, QLX_ERR_STATE, "image, PSF and settings must be set first");
    hipSetDevice(c->device);
    std::vector<qlx_lens_model> models;
    std::vector<int> idx, srcn;
    std::vector<double> phys((size_t)nbatch * np);
    std::vector<qlx_result> res(nbatch);
    for (int b = 0; b < nbatch; b++) {
        const double* xb = x + (size_t)b * np;
        double* tb = phys.data() + (size_t)b * np;
        bool inside = true;
        for (int k = 0; k < np; k++) {  // inverse_transform_parameters (params.h:722-743)
            const qlx_fit_param& f = c->fit_params[k];
            if (f.transform == QLX_TRANSFORM_NONE) tb[k] = xb[k];
            else if (f.transform == QLX_TRANSFORM_LOG10) tb[k] = pow(10.0, xb[k]);
            else tb[k] = (xb[k] - f.tb) / f.ta;
            if (f.lo < f.hi && ((tb[k] < f.lo) or (tb[k] > f.hi))) inside = false;
        }
        memset(&res[b], 0, sizeof(res[b]));
        if (!inside) {  // lens.cpp:10673-10677
            loglike[b] = 1e30;
            res[b].minus2lnE = 2e30;
            res[b].status = QLX_ERR_ARG;
            continue;
        }
        qlx_lens_model m = c->fit_base;  // update_model
        for (int k = 0; k < np; k++) m.lens[c->fit_params[k].lens].p[c->fit_params[k].param] = tb[k];
        models.push_back(m);
        idx.push_back(b);
        srcn.push_back(c->set.src_nx);
        srcn.push_back(c->set.src_ny);
    }
    std::vector<qlx_result> ev(models.size());
    if (!models.empty()) {
        const int rc = eval_items(c, (int)models.size(), models.data(), srcn.data(), ev.data());
        if (rc < 0) return rc;
        // the kept solutions and rays are indexed by the caller's point: remap them
        std::vector<Ctx::Kept> kept(nbatch, Ctx::Kept{0, 0, 0, 0, 0});
        for (size_t k = 0; k < idx.size(); k++) kept[idx[k]] = c->kept[k];
        c->kept.swap(kept);
        for (int& w : c->last_wave) w = idx[w];
    } else {
        c->kept.assign(nbatch, Ctx::Kept{0, 0, 0, 0, 0});
        c->last_wave.clear();
    }
    int any_bad = 0;
    for (size_t k = 0; k < idx.size(); k++) {
        const int b = idx[k];
        res[b] = ev[k];
        if (ev[k].status != QLX_OK && ev[k].status != QLX_ERR_REFINE) {
            loglike[b] = 1e30;  // a failed inversion: the sentinel, as update_model's huge penalty
            any_bad = 1;
            continue;
        }
        double L = 0.0;
        L += ev[k].minus2lnE / 2;  // lens.cpp:10719
        const double* xb = x + (size_t)b * np;
        const double* tb = phys.data() + (size_t)b * np;
        double dl = 0.0;  // add_prior_terms_to_loglike on the sampler coordinates (params.h:810-850)
        for (int q = 0; q < np; q++) {
            const qlx_fit_param& f = c->fit_params[q];
            if (f.prior == QLX_PRIOR_UNIFORM) continue;
            dl += log(f.prior_norm);
            if (f.prior == QLX_PRIOR_LOG) dl += log(xb[q]);
            else dl += ((xb[q] - f.prior_mu) / f.prior_sig) * ((xb[q] - f.prior_mu) / f.prior_sig) / 2.0;
        }
        L += dl;
        for (int q = 0; q < np; q++)  // add_jacobian_terms_to_loglike on the physical values (params.h:884-893)
            if (c->fit_params[q].include_jacobian && c->fit_params[q].transform == QLX_TRANSFORM_LOG10) L -= log(tb[q]);
        loglike[b] = L;
    }
    if (out) memcpy(out, res.data(), sizeof(qlx_result) * nbatch);
    for (int b = 0; b < nbatch; b++)
        if (res[b].status != QLX_OK) any_bad = 1;
    return any_bad ? QLX_ITEMS_FAILED : QLX_OK;
}

int qlx_raytrace(void* ctx, const qlx_lens_model* model, int n, const double* xy, double* beta)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (n < 0 || (n > 0 && (!xy || !beta))) return fail(c, QLX_ERR_ARG, "bad arguments");
    int rc = check_model(c, model);
    if (rc) return rc;
    if (n == 0) return QLX_OK;
    hipSetDevice(c->device);
    double tol = c->settings_set ? c->set.integral_tol : 1e-3;
    ModelPrep mp;
    prep_model(*model, mp);
    DevBuf dxy, db;
    if (!dxy.ensure((size_t)n * 16) || !db.ensure((size_t)n * 16)) return fail(c, QLX_ERR_NOMEM, "allocation failed");
    HIPCHK(c, hipMemcpyAsync(dxy.p, xy, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    launch_raytrace_points(c->stream, mp, tol, n, dxy.as<double>(), db.as<double>());
    HIPCHK(c, hipMemcpyAsync(beta, db.p, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

int qlx_raytrace_noalloc(void* ctx, const qlx_lens_model* model, int n, const double* xy, double* beta)
{
    // qlx_raytrace with the device staging in the file-scope array of raytrace.hip: no allocation
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (n < 0 || (n > 0 && (!xy || !beta))) return fail(c, QLX_ERR_ARG, "bad arguments");
    if (n > QLX_RT_SCRATCH_POINTS) return fail(c, QLX_ERR_RANGE, "qlx_raytrace_noalloc: more than 262144 points");
    int rc = check_model(c, model);
    if (rc) return rc;
    if (n == 0) return QLX_OK;
    hipSetDevice(c->device);
    double tol = c->settings_set ? c->set.integral_tol : 1e-3;
    ModelPrep mp;
    prep_model(*model, mp);
    double* in = c->rt_scratch;
    double* res = c->rt_scratch + 2 * (size_t)QLX_RT_SCRATCH_POINTS;
    HIPCHK(c, hipMemcpyAsync(in, xy, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    launch_raytrace_points(c->stream, mp, tol, n, in, res);
    HIPCHK(c, hipMemcpyAsync(beta, res, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

int qlx_nfw_def_hess(void* ctx, const double* nfw_params, double tol, int n, const double* xy, double* out5)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!nfw_params || n < 0 || (n > 0 && (!xy || !out5)) || tol <= 0) return fail(c, QLX_ERR_ARG, "bad arguments");
    if (n == 0) return QLX_OK;
    hipSetDevice(c->device);
    qlx_lens_model m{};
    m.nlens = 1;
    m.lens[0].type = QLX_LENS_NFW;
    for (int i = 0; i < 6; i++) m.lens[0].p[i] = nfw_params[i];
    int rc = check_model(c, &m);
    if (rc) return rc;
    ModelPrep mp;
    prep_model(m, mp);
    DevBuf dxy, dout;
    if (!dxy.ensure((size_t)n * 16) || !dout.ensure((size_t)n * 40)) return fail(c, QLX_ERR_NOMEM, "allocation failed");
    HIPCHK(c, hipMemcpyAsync(dxy.p, xy, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    launch_nfw_def_hess(c->stream, mp.lens[0], tol, n, dxy.as<double>(), dout.as<double>());
    HIPCHK(c, hipMemcpyAsync(out5, dout.p, (size_t)n * 40, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

// NR3 "Ran" generator and polar Box-Muller (Press et al. 2007; the reference's rand.h:19-76)
struct Nr3 {
    uint64_t u, v, w;
    bool first = true;
    double gset = 0;
    uint64_t next()
    {
        u = u * 2862933555777941757ULL + 7046029254386353087ULL;
        v ^= v >> 17;
        v ^= v << 31;
        v ^= v >> 8;
        w = 4294957665U * (w & 0xffffffff) + (w >> 32);
        uint64_t x = u ^ (u << 21);
        x ^= x >> 35;
        x ^= x << 4;
        return (x + v) ^ w;
    }
    explicit Nr3(uint64_t j)
    {
        v = 4101842887655102017ULL;
        w = 1;
        u = j ^ v;
        next();
        v = u;
        next();
        next();
        w = v;
        next();
        next();
    }
    double normal()
    {
        if (first) {
            double v1, v2, rsq;
            do {
                v1 = 2.0 * (5.42101086242752217E-20 * next()) - 1.0;
                v2 = 2.0 * (5.42101086242752217E-20 * next()) - 1.0;
                rsq = v1 * v1 + v2 * v2;
            } while ((rsq >= 1.0) || (rsq == 0.0));
            double fac = sqrt(-2.0 * log(rsq) / rsq);
            gset = v1 * fac;
            first = false;
            return v2 * fac;
        }
        first = true;
        return gset;
    }
};

int qlx_mock_image(void* ctx, const qlx_lens_model* model, int nsrc, const double* sersic, double sigma, unsigned long long seed,
                   double* img_out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (!c->image_set || !c->psf_set) return fail(c, QLX_ERR_STATE, "set the image grid and PSF first");
    if (nsrc < 0 || (nsrc > 0 && !sersic) || !img_out || sigma < 0) return fail(c, QLX_ERR_ARG, "bad arguments");
    int rc = check_model(c, model);
    if (rc) return rc;
    hipSetDevice(c->device);
    const int nsplit = c->settings_set ? c->set.nsplit : 2;
    const double tol = c->settings_set ? c->set.integral_tol : 1e-3;
    // Sersic meta-parameters on the host (Sersic::update_meta_parameters sbprofile.cpp:2472-2478;
    // source angle from +x, SB_Profile::set_angle with orient_major_axis_north = false)
    std::vector<double> sp((size_t)std::max(nsrc, 1) * 11);
    for (int k = 0; k < nsrc; k++) {
        const double* S = sersic + 7 * k;
        double n = S[2];
        double th = DEG2RAD * S[4];
        double* o = sp.data() + 11 * k;
        for (int i = 0; i < 7; i++) o[i] = S[i];
        o[7] = cos(th);
        o[8] = sin(th);
        o[9] = 2 * n - 0.33333333333333 + 4.0 / (405 * n) + 46.0 / (25515 * n * n) + 131.0 / (1148175 * n * n * n);
        o[10] = (th != 0) ? 1.0 : 0.0;
    }
    ModelPrep mp;
    prep_model(*model, mp);
    size_t npix = (size_t)c->nx * c->ny;
    DevBuf dsp, dsb, dimg;
    if (!dsp.ensure(sp.size() * 8) || !dsb.ensure(npix * 8) || !dimg.ensure(npix * 8)) return fail(c, QLX_ERR_NOMEM, "allocation failed");
    HIPCHK(c, hipMemcpyAsync(dsp.p, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, c->stream));
    launch_mock_sb(c->stream, c->nx, c->ny, nsplit, c->xvals.as<double>(), c->yvals.as<double>(), mp, tol, nsrc, dsp.as<double>(),
                   dsb.as<double>());
    launch_conv_image(c->stream, c->nx, c->ny, c->psf.as<double>(), c->pnx, c->pny, dsb.as<double>(), dimg.as<double>());
    HIPCHK(c, hipMemcpyAsync(img_out, dimg.p, npix * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (sigma > 0) {  // ImagePixelGrid::add_pixel_noise: j outer, i inner (pixelgrid.cpp:10987-10996)
        Nr3 r(seed);
        for (int j = 0; j < c->ny; j++)
            for (int i = 0; i < c->nx; i++) img_out[(size_t)i * c->ny + j] += sigma * r.normal();
    }
    return QLX_OK;
}

int qlx_kernel_time(void* ctx, const char* name, double* ms, int* launches)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !name || !ms) return QLX_ERR_ARG;
    auto it = c->timers.find(name);
    if (it == c->timers.end() || it->second.launches == 0) {
        *ms = 0;
        if (launches) *launches = 0;
        return QLX_ERR_STATE;
    }
    *ms = it->second.total_ms / it->second.launches;
    if (launches) *launches = it->second.launches;
    return QLX_OK;
}

int qlx_get_stats(void* ctx, qlx_stats* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !out) return QLX_ERR_ARG;
    *out = c->stats;
    out->n_waves = c->stats_waves;
    out->chol_panel = c->last_panel;
    out->graph_captures = c->graph_captures;
    out->graph_replays = c->graph_replays;
    return QLX_OK;
}

int qlx_selftest_mfma(void* ctx, int kind, double* out256)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !out256 || kind < 0 || kind > 3) return QLX_ERR_ARG;
    hipSetDevice(c->device);
    DevBuf d;
    if (!d.ensure(1024 * 8)) return QLX_ERR_NOMEM;
    launch_selftest_mfma(c->stream, d.as<double>());
    HIPCHK(c, hipMemcpyAsync(out256, d.as<double>() + 256 * kind, 256 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

}  // extern "C"
