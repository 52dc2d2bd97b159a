// Host side of the qlx C ABI: context, device buffers, and the per-evaluation pipeline that
// replaces the Cartesian/DENSE/curvature block of QLens::invert_image_surface_brightness_map
// (lens.cpp:12364-12615) and its chi-square loop (lens.cpp:12950-13015).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "qlx_internal.h"

namespace qlx {
void upload_patterson(int device);
void launch_prep_fr(hipStream_t s, double* F, double* R, int ld, int n, double lambda);
void launch_trisolve_f64_w(hipStream_t s, const double* L, int ld, const double* Winv, const double* b, const double* yrow, int n,
                           double* x);
void launch_selftest_mfma(hipStream_t s, double* out);
}  // namespace qlx

using namespace qlx;

namespace {

const double DEG2RAD = 0.017453292519943295;  // mathexpr.h:27
const double HALFPI_REF = 1.57079632679;       // mathexpr.h:10 (the reference's truncated pi/2)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    bool ensure(size_t n)
    {
        if (n <= bytes && p) return true;
        release();
        if (n == 0) n = 8;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = n;
        return true;
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
    hipStream_t stream2 = nullptr;        // runs the R factorization beside F's
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    // image
    int nx = 0, ny = 0;
    bool image_set = false;
    std::vector<double> h_xvals, h_yvals, h_data, h_covinv;
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
    DevBuf sxv, syv, scx, scy;
    // per-item work
    DevBuf beta, map_cell, map_w, pmapped, cmapped, sidx, pidx, act_pix, counts, ucol, uval, ucnt, nevals;
    DevBuf key, bucket_cnt, bucket_off, bucket_fill, perm;
    // dense
    DevBuf Lb, cvec, F, R, D, s, diagF, diagR, work, work2, nz, model, out, logdets, status;
    // last solution
    std::vector<double> last_src, last_model;
    int last_snx = 0, last_sny = 0;
    // timers
    std::map<std::string, Timer> timers;
    qlx_stats stats{};
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

void prep_model(const qlx_lens_model& in, ModelPrep& out)
{
    memset(&out, 0, sizeof(out));
    out.nlens = in.nlens;
    for (int l = 0; l < in.nlens; l++) {
        const qlx_lens& s = in.lens[l];
        LensPrep& L = out.lens[l];
        L.type = s.type;
        L.def_mode = s.def_mode;
        if (s.type == QLX_LENS_SIE || s.type == QLX_LENS_NFW) {
            // LensProfile::set_angle + orient_major_axis_north (profile.cpp:1625-1637)
            double th = DEG2RAD * (s.type == QLX_LENS_SIE ? s.p[2] : s.p[3]);
            double cth = cos(th), sth = sin(th);
            L.sinth = cth;
            L.costh = -sth;
            L.rotate = (L.sinth != 0);
        }
        if (s.type == QLX_LENS_SIE) {
            // SPLE_Lens::update_meta_parameters, emode 1 (models.cpp:101-110, profile.cpp:1690-1693)
            double b = s.p[0], q = s.p[1], core = s.p[5];
            double fmaj = 1.0 / sqrt(q);
            L.q = q;
            L.bprime = b * fmaj;
            L.sprime = core * fmaj;
            L.qsq = q * q;
            L.ssq = L.sprime * L.sprime;
            L.sie_u = sqrt(1 - L.qsq);
            L.sie_fac = L.bprime * q / L.sie_u;
            L.xc = s.p[3];
            L.yc = s.p[4];
        } else if (s.type == QLX_LENS_NFW) {
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
        }
    }
}

}  // namespace qlx

namespace {

int check_model(Ctx* c, const qlx_lens_model* m)
{
    if (!m || m->nlens < 0 || m->nlens > QLX_MAX_LENSES) return fail(c, QLX_ERR_ARG, "invalid lens model");
    for (int l = 0; l < m->nlens; l++) {
        int t = m->lens[l].type;
        if (t != QLX_LENS_SIE && t != QLX_LENS_NFW && t != QLX_LENS_SHEAR) return fail(c, QLX_ERR_ARG, "unknown lens type");
        if (t == QLX_LENS_SIE && !(m->lens[l].p[1] > 0 && m->lens[l].p[1] <= 1)) return fail(c, QLX_ERR_ARG, "SIE q must be in (0,1]");
        if (t == QLX_LENS_NFW && !(m->lens[l].p[2] > 0 && m->lens[l].p[2] <= 1 && m->lens[l].p[1] > 0))
            return fail(c, QLX_ERR_ARG, "NFW needs rs > 0 and q in (0,1]");
    }
    return QLX_OK;
}

GridDev make_grid(Ctx* c, int snx, int sny)
{
    GridDev g;
    g.nx = c->nx;
    g.ny = c->ny;
    g.nsplit = c->set.nsplit;
    g.nsub = c->set.nsplit * c->set.nsplit;
    g.xvals = c->xvals.as<double>();
    g.yvals = c->yvals.as<double>();
    g.mask = c->has_mask ? c->mask.as<unsigned char>() : nullptr;
    g.snx = snx;
    g.sny = sny;
    g.sxv = c->sxv.as<double>();
    g.syv = c->syv.as<double>();
    g.scx = c->scx.as<double>();
    g.scy = c->scy.as<double>();
    return g;
}

// Source grid corners accumulated x += xstep and 4-corner-mean centres
// (SourcePixelGrid::create_pixel_grid pixelgrid.cpp:441-455, SourcePixel ctor 761-783).
int upload_source_grid(Ctx* c, int snx, int sny)
{
    const qlx_settings& s = c->set;
    std::vector<double> sx(snx + 1), sy(sny + 1), cx(snx), cy(sny);
    double xstep = (s.src_xmax - s.src_xmin) / snx, ystep = (s.src_ymax - s.src_ymin) / sny;
    double x = s.src_xmin, y = s.src_ymin;
    for (int i = 0; i <= snx; i++, x += xstep) sx[i] = x;
    for (int j = 0; j <= sny; j++, y += ystep) sy[j] = y;
    for (int i = 0; i < snx; i++) cx[i] = (sx[i] + sx[i] + sx[i + 1] + sx[i + 1]) / 4.0;
    for (int j = 0; j < sny; j++) cy[j] = (sy[j] + sy[j + 1] + sy[j] + sy[j + 1]) / 4.0;
    if (!c->sxv.ensure(sizeof(double) * (snx + 1)) || !c->syv.ensure(sizeof(double) * (sny + 1)) ||
        !c->scx.ensure(sizeof(double) * snx) || !c->scy.ensure(sizeof(double) * sny))
        return fail(c, QLX_ERR_NOMEM, "source grid allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->sxv.p, sx.data(), sizeof(double) * (snx + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->syv.p, sy.data(), sizeof(double) * (sny + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->scx.p, cx.data(), sizeof(double) * snx, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->scy.p, cy.data(), sizeof(double) * sny, hipMemcpyHostToDevice, c->stream));
    return QLX_OK;
}

int ensure_item_bufs(Ctx* c, int snx, int sny)
{
    size_t npix = (size_t)c->nx * c->ny;
    size_t nsub = (size_t)c->set.nsplit * c->set.nsplit;
    size_t ncell = (size_t)snx * sny;
    size_t U = 3 * nsub;
    bool ok = c->beta.ensure(npix * nsub * 2 * 8) && c->map_cell.ensure(npix * nsub * 3 * 4) && c->map_w.ensure(npix * nsub * 3 * 8) &&
              c->pmapped.ensure(npix * 4) && c->cmapped.ensure(ncell * 4) && c->sidx.ensure(ncell * 4) && c->pidx.ensure(npix * 4) &&
              c->act_pix.ensure(npix * 4) && c->counts.ensure(16 * 4) && c->ucol.ensure(npix * U * 4) && c->uval.ensure(npix * U * 8) &&
              c->ucnt.ensure(npix * 4) && c->nevals.ensure(8) && c->out.ensure(16 * 8) && c->logdets.ensure(4 * 8) &&
              c->status.ensure(4 * 4) && c->model.ensure((npix * 2 + 64) * 8) && c->key.ensure(npix * 4) &&
              c->bucket_cnt.ensure((ncell + 2) * 4) && c->bucket_off.ensure((ncell + 2) * 4) &&
              c->bucket_fill.ensure((ncell + 2) * 4) && c->perm.ensure(npix * 4);
    if (!ok) return fail(c, QLX_ERR_NOMEM, "work buffer allocation failed");
    return QLX_OK;
}

ItemBufs item_bufs(Ctx* c)
{
    ItemBufs b;
    b.beta = c->beta.as<double>();
    b.map_cell = c->map_cell.as<int>();
    b.map_w = c->map_w.as<double>();
    b.pmapped = c->pmapped.as<int>();
    b.cmapped = c->cmapped.as<int>();
    b.sidx = c->sidx.as<int>();
    b.pidx = c->pidx.as<int>();
    b.act_pix = c->act_pix.as<int>();
    b.counts = c->counts.as<int>();
    b.ucol = c->ucol.as<int>();
    b.uval = c->uval.as<double>();
    b.ucnt = c->ucnt.as<int>();
    b.nevals = (long long*)c->nevals.p;
    b.key = c->key.as<int>();
    b.bucket_cnt = c->bucket_cnt.as<int>();
    b.bucket_off = c->bucket_off.as<int>();
    b.bucket_fill = c->bucket_fill.as<int>();
    b.perm = c->perm.as<int>();
    return b;
}

// One evaluation in FP64 (parity) mode.
int eval_one_fp64(Ctx* c, const qlx_lens_model& model, int snx, int sny, qlx_result* res)
{
    memset(res, 0, sizeof(*res));
    int rc = upload_source_grid(c, snx, sny);
    if (rc) return rc;
    rc = ensure_item_bufs(c, snx, sny);
    if (rc) return rc;
    ModelPrep mp;
    prep_model(model, mp);
    GridDev g = make_grid(c, snx, sny);
    ItemBufs b = item_bufs(c);
    hipStream_t st = c->stream;
    const size_t npix = (size_t)c->nx * c->ny;
    launch_fill_i32(st, b.pmapped, 0, npix);
    launch_fill_i32(st, b.cmapped, 0, (size_t)snx * sny);
    launch_fill_i32(st, c->status.as<int>(), 0, 4);
    launch_fill_i32(st, (int*)c->nevals.p, 0, 2);

    timer_begin(c, "total");
    timer_begin(c, "raytrace");
    launch_raytrace_map(st, g, mp, c->set.integral_tol, c->set.ray_precision == QLX_RAY_FP32, b, nullptr);
    timer_end(c, "raytrace");
    launch_compact(st, g, b);
    int counts[2];
    HIPCHK(c, hipMemcpyAsync(counts, b.counts, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int n_src = counts[0], n_img = counts[1];
    res->n_src_active = n_src;
    res->n_img_active = n_img;
    if (n_src == 0) {
        timer_end(c, "total");
        res->minus2lnE = 2e30;  // pixelgrid.cpp:12807 -> lens.cpp:12528-12530
        res->status = QLX_ERR_NOSOURCE;
        return QLX_OK;
    }
    launch_merge_lrows(st, g, b, n_img);

    const int ldN = (int)round_up((size_t)n_src + 1, 64);
    const int nrows = (int)std::max<size_t>(64, round_up((size_t)n_img, 64));
    const int ncb = ldN / 64, ncbw = (ncb + 31) / 32, nkt = nrows / 64;
    if (!c->Lb.ensure((size_t)nrows * ldN * 8) || !c->cvec.ensure((size_t)nrows * 8) || !c->F.ensure((size_t)ldN * ldN * 8) ||
        !c->R.ensure((size_t)ldN * ldN * 8) || !c->D.ensure((size_t)ldN * 8) || !c->s.ensure((size_t)ldN * 8) ||
        !c->diagF.ensure((size_t)ldN * 8) || !c->diagR.ensure((size_t)ldN * 8) || !c->work.ensure((size_t)ldN * 64 * 8) || !c->work2.ensure((size_t)ldN * 64 * 8) ||
        !c->nz.ensure((size_t)nkt * ncbw * 4))
        return fail(c, QLX_ERR_NOMEM, "dense buffer allocation failed (Lb/F/R)");
    double* Lb = c->Lb.as<double>();
    launch_fill_i32(st, c->nz.as<int>(), 0, (size_t)nkt * ncbw);
    timer_begin(c, "lbuild");
    launch_sort_rows(st, g, b, n_img, n_src);
    launch_row_support(st, g, b, n_img, n_src, c->pnx, c->pny, c->nz.as<unsigned int>(), ncbw);
    launch_build_lb_blocks_f64(st, g, b, n_img, nrows, n_src, c->psf.as<double>(), c->pnx, c->pny, c->data.as<double>(), Lb, ncb,
                               c->nz.as<unsigned int>(), ncbw);
    timer_end(c, "lbuild");
    launch_gather_cvec_sorted(st, b, c->covinv.as<double>(), n_img, nrows, c->cvec.as<double>());
    timer_begin(c, "syrk");
    launch_syrk_blocks_f64(st, Lb, ncb, nkt, c->cvec.as<double>(), c->F.as<double>(), ldN, c->nz.as<unsigned int>(), ncbw);
    timer_end(c, "syrk");
    launch_fill_f64(st, c->R.as<double>(), 0.0, (size_t)ldN * ldN);
    launch_add_curvature(st, g, b, n_src, c->set.regparam, c->F.as<double>(), ldN, c->R.as<double>(), ldN);
    launch_prep_fr(st, c->F.as<double>(), c->R.as<double>(), ldN, n_src, c->set.regparam);
    // R (ln det R only) and F (solve + ln det F) factor concurrently on two streams
    timer_begin(c, "cholesky");
    HIPCHK(c, hipEventRecord(c->ev_fork, st));
    HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    cholesky_f64(c->stream2, c->R.as<double>(), ldN, n_src, c->diagR.as<double>(), c->status.as<int>() + 1, c->work2.as<double>());
    launch_logdet(c->stream2, c->diagR.as<double>(), n_src, c->logdets.as<double>() + 1);
    HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));
    cholesky_f64(st, c->F.as<double>(), ldN, n_src, c->diagF.as<double>(), c->status.as<int>(), c->work.as<double>());
    launch_logdet(st, c->diagF.as<double>(), n_src, c->logdets.as<double>());
    HIPCHK(c, hipStreamWaitEvent(st, c->ev_join, 0));
    timer_end(c, "cholesky");
    // backward solve only: y = L^-1 D is row n_src of the factored augmented matrix
    launch_trisolve_f64_w(st, c->F.as<double>(), ldN, c->work.as<double>(), nullptr, c->F.as<double>() + (size_t)n_src * ldN, n_src,
                          c->s.as<double>());
    launch_evidence(st, g, b, n_img, n_src, c->s.as<double>(), c->psf.as<double>(), c->pnx, c->pny, c->data.as<double>(),
                    c->covinv.as<double>(), c->set.regparam, c->logdets.as<double>(), c->model.as<double>(), c->out.as<double>());
    timer_end(c, "total");
    double out[6];
    int status[2];
    long long nev = 0;
    HIPCHK(c, hipMemcpyAsync(out, c->out.p, sizeof(out), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(status, c->status.p, sizeof(status), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&nev, c->nevals.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    timers_collect(c);
    res->minus2lnE = out[0];
    res->chisq0 = out[1];
    res->logdetF = out[2];
    res->logdetR = out[3];
    res->Es = out[4];
    res->status = (status[0] || status[1]) ? QLX_ERR_NOT_PD : QLX_OK;

    // work accounting: executed vs dense SYRK tile-chunks from the occupancy bitmap
    {
        std::vector<unsigned int> nzh((size_t)nkt * ncbw);
        HIPCHK(c, hipMemcpy(nzh.data(), c->nz.p, nzh.size() * 4, hipMemcpyDeviceToHost));
        double exec_pairs = 0;
        for (int kt = 0; kt < nkt; kt++) {
            long long bits = 0;
            for (int w = 0; w < ncbw; w++) bits += __builtin_popcount(nzh[(size_t)kt * ncbw + w]);
            exec_pairs += (double)bits * (bits + 1) / 2;
        }
        const double tile = 2.0 * 64 * 64 * 64;
        qlx_stats& S = c->stats;
        S.syrk_exec_flops = exec_pairs * tile;
        S.syrk_dense_flops = (double)ncb * (ncb + 1) / 2 * nkt * tile;
        S.syrk_alg_flops = (double)n_img * n_src * (n_src + 1) + 2.0 * n_img * n_src;
        double nblk = 0;
        for (size_t w = 0; w < nzh.size(); w++) nblk += __builtin_popcount(nzh[w]);
        S.lb_bytes = nblk * 64 * 64 * 8;
        S.integrand_evals = (double)nev;
        S.chol_flops = 2.0 * (double)n_src * n_src * n_src / 3.0;
        S.n_img = n_img;
        S.n_src = n_src;
        S.ld_src = ldN;
        S.rows_pad = nrows;
    }

    // keep the solution of the last item for qlx_get_solution
    std::vector<double> s(n_src);
    std::vector<int> sidx((size_t)snx * sny);
    c->last_model.resize(npix);
    HIPCHK(c, hipMemcpyAsync(s.data(), c->s.p, sizeof(double) * n_src, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(sidx.data(), b.sidx, sizeof(int) * sidx.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(c->last_model.data(), c->model.p, sizeof(double) * npix, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    c->last_src.assign((size_t)snx * sny, 0.0);
    for (size_t k = 0; k < sidx.size(); k++)
        if (sidx[k] >= 0) c->last_src[k] = s[sidx[k]];
    c->last_snx = snx;
    c->last_sny = sny;
    return QLX_OK;
}

}  // namespace

extern "C" {

int qlx_ctx_create(const qlx_opts* opts, void** out)
{
    if (!out) return fail(nullptr, QLX_ERR_ARG, "null ctx pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, QLX_ERR_HIP, "no HIP device");
    Ctx* c = new Ctx();
    c->device = opts ? opts->device : 0;
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
    if (nx < 1 || ny < 1 || !xvals || !yvals || !data || !covinv) return fail(c, QLX_ERR_ARG, "qlx_set_image: bad arguments");
    if ((long long)nx * ny > (1LL << 26)) return fail(c, QLX_ERR_RANGE, "image too large");
    hipSetDevice(c->device);
    size_t npix = (size_t)nx * ny;
    c->nx = nx;
    c->ny = ny;
    c->h_xvals.assign(xvals, xvals + nx + 1);
    c->h_yvals.assign(yvals, yvals + ny + 1);
    c->h_data.assign(data, data + npix);
    c->h_covinv.assign(covinv, covinv + npix);
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
    if (pnx < 1 || pny < 1 || !psf || (pnx % 2) == 0 || (pny % 2) == 0) return fail(c, QLX_ERR_ARG, "PSF must be odd-sized and non-empty");
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
    if (nx * ny > 4096 * 4096) return fail(c, QLX_ERR_RANGE, "PSF too large");
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
    if (!s) return fail(c, QLX_ERR_ARG, "null settings");
    if (s->nsplit < 1 || s->nsplit > 4) return fail(c, QLX_ERR_RANGE, "nsplit must be 1..4");
    if (!(s->src_xmax > s->src_xmin) || !(s->src_ymax > s->src_ymin)) return fail(c, QLX_ERR_ARG, "empty source grid bounds");
    if (s->src_nx > 512 || s->src_ny > 512) return fail(c, QLX_ERR_RANGE, "source grid larger than 512 per axis");
    if (s->precision != QLX_PREC_FP64 && s->precision != QLX_PREC_MIXED) return fail(c, QLX_ERR_ARG, "unknown precision mode");
    if (s->integral_tol <= 0) return fail(c, QLX_ERR_ARG, "integral tolerance must be > 0");
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
    for (auto& kv : c->timers) {
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
    int first_bad = QLX_OK;
    for (int b = 0; b < nbatch; b++) {
        int rc = check_model(c, &models[b]);
        if (rc) return rc;
        qlx_result* r = &out[b];
        if (c->set.src_nx < 2 || c->set.src_ny < 2) {  // lens.cpp:12453-12458
            memset(r, 0, sizeof(*r));
            r->minus2lnE = 2e30;
            r->status = QLX_ERR_NOSOURCE;
        } else {
            rc = eval_one_fp64(c, models[b], c->set.src_nx, c->set.src_ny, r);
            if (rc) return rc;
        }
        if (r->status != QLX_OK && first_bad == QLX_OK) first_bad = r->status;
    }
    return first_bad;
}

int qlx_eval_batch_grids(void* ctx, int nbatch, const qlx_lens_model* models, const int* src_n, qlx_result* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (nbatch < 0 || (nbatch > 0 && (!models || !out || !src_n))) return fail(c, QLX_ERR_ARG, "bad arguments");
    if (!c->image_set || !c->settings_set || !c->psf_set) return fail(c, QLX_ERR_STATE, "image, PSF and settings must be set first");
    hipSetDevice(c->device);
    int first_bad = QLX_OK;
    for (int b = 0; b < nbatch; b++) {
        int rc = check_model(c, &models[b]);
        if (rc) return rc;
        int snx = src_n[2 * b], sny = src_n[2 * b + 1];
        if (snx > 512 || sny > 512) return fail(c, QLX_ERR_RANGE, "source grid larger than 512 per axis");
        qlx_result* r = &out[b];
        if (snx < 2 || sny < 2) {
            memset(r, 0, sizeof(*r));
            r->minus2lnE = 2e30;
            r->status = QLX_ERR_NOSOURCE;
        } else {
            rc = eval_one_fp64(c, models[b], snx, sny, r);
            if (rc) return rc;
        }
        if (r->status != QLX_OK && first_bad == QLX_OK) first_bad = r->status;
    }
    return first_bad;
}

int qlx_get_solution(void* ctx, double* src_cells, double* model_image)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return QLX_ERR_ARG;
    if (c->last_src.empty()) return fail(c, QLX_ERR_STATE, "no evaluation yet");
    if (src_cells) memcpy(src_cells, c->last_src.data(), sizeof(double) * c->last_src.size());
    if (model_image) memcpy(model_image, c->last_model.data(), sizeof(double) * c->last_model.size());
    return QLX_OK;
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
    qlx_settings saved = c->set;
    c->set.nsplit = nsplit;
    GridDev g = make_grid(c, 2, 2);
    c->set = saved;
    g.nsplit = nsplit;
    g.nsub = nsplit * nsplit;
    g.mask = nullptr;
    size_t npix = (size_t)c->nx * c->ny;
    DevBuf dsp, dsb, dimg;
    if (!dsp.ensure(sp.size() * 8) || !dsb.ensure(npix * 8) || !dimg.ensure(npix * 8)) return fail(c, QLX_ERR_NOMEM, "allocation failed");
    HIPCHK(c, hipMemcpyAsync(dsp.p, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, c->stream));
    launch_mock_sb(c->stream, g, mp, tol, nsrc, dsp.as<double>(), dsb.as<double>());
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
    return QLX_OK;
}

int qlx_selftest_mfma(void* ctx, double* out256)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !out256) return QLX_ERR_ARG;
    hipSetDevice(c->device);
    DevBuf d;
    if (!d.ensure(256 * 8)) return QLX_ERR_NOMEM;
    launch_selftest_mfma(c->stream, d.as<double>());
    HIPCHK(c, hipMemcpyAsync(out256, d.p, 256 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QLX_OK;
}

}  // extern "C"
