// Host driver of the learned NFW deflection surrogate (C ABI qlx_surrogate_*): parameter
// storage, training-set generation by the GPU quadrature kernel, the training loop (one Adam
// step captured in a HIP graph and replayed), save/load and evaluation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "qlx_internal.h"

using namespace qlx;

namespace {

struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    bool ensure(size_t bytes)
    {
        if (p && bytes <= n) return true;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = bytes;
        return true;
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct SurState {
    bool alloc = false, trained = false;
    DBuf P, m, v, Wb, WTb, W1r, W5r, norm, step;
    SurParams sp{};
    SurDev net{};
    float h_norm[10];
    // minibatch work and training data
    DBuf zs, hs, dzA, dzB, dz1, loss_part, g5_part, gW_part, gb_part, g1_part, grad, loss, bf, bt;
    SurWork w{};
    ~SurState()
    {
        for (DBuf* b : {&P, &m, &v, &Wb, &WTb, &W1r, &W5r, &norm, &step, &zs, &hs, &dzA, &dzB, &dz1, &loss_part, &g5_part, &gW_part,
                        &gb_part, &g1_part, &grad, &loss, &bf, &bt})
            b->release();
    }
};

// SplitMix64 (Steele, Lea & Flood 2014): host-side seeded uniforms for sampling and init
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

#define SURCHK(c, expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            char b_[300];                                                                \
            snprintf(b_, sizeof(b_), "%s: %s", #expr, hipGetErrorString(e_));            \
            return ctx_fail(c, QLX_ERR_HIP, b_);                                         \
        }                                                                                \
    } while (0)

int ensure_params(void* ctx, SurState* S)
{
    if (S->alloc) return QLX_OK;
    const size_t np = QLX_SUR_NPARAM;
    if (!S->P.ensure(np * 4) || !S->m.ensure(np * 4) || !S->v.ensure(np * 4) || !S->Wb.ensure(3 * 128 * 128 * 2) ||
        !S->WTb.ensure(3 * 128 * 128 * 2) || !S->W1r.ensure(512 * 4) || !S->W5r.ensure(640 * 4) || !S->norm.ensure(10 * 4) ||
        !S->step.ensure(4))
        return ctx_fail(ctx, QLX_ERR_NOMEM, "surrogate parameter allocation failed");
    S->sp.P = S->P.as<float>();
    S->sp.m = S->m.as<float>();
    S->sp.v = S->v.as<float>();
    S->sp.Wb = S->Wb.as<uint16_t>();
    S->sp.WTb = S->WTb.as<uint16_t>();
    S->sp.W1r = S->W1r.as<float>();
    S->sp.W5r = S->W5r.as<float>();
    S->sp.norm = S->norm.as<float>();
    S->sp.step = S->step.as<int>();
    S->net = sur_dev(S->sp);
    for (int k = 0; k < 5; k++) {
        S->h_norm[k] = 0.0f;
        S->h_norm[5 + k] = 1.0f;
    }
    hipStream_t st = ctx_stream(ctx);
    SURCHK(ctx, hipMemcpyAsync(S->norm.p, S->h_norm, 40, hipMemcpyHostToDevice, st));
    std::vector<float> zero(np, 0.0f);
    SURCHK(ctx, hipMemcpyAsync(S->P.p, zero.data(), np * 4, hipMemcpyHostToDevice, st));
    SURCHK(ctx, hipMemcpyAsync(S->m.p, zero.data(), np * 4, hipMemcpyHostToDevice, st));
    SURCHK(ctx, hipMemcpyAsync(S->v.p, zero.data(), np * 4, hipMemcpyHostToDevice, st));
    SURCHK(ctx, hipMemcpyAsync(S->step.p, zero.data(), 4, hipMemcpyHostToDevice, st));
    launch_sur_refresh(st, S->sp);
    SURCHK(ctx, hipStreamSynchronize(st));
    S->alloc = true;
    return QLX_OK;
}

int ensure_work(void* ctx, SurState* S, int n)
{
    const size_t nwg = (n + QLX_SUR_SB - 1) / QLX_SUR_SB, N = nwg * QLX_SUR_SB;
    bool ok = S->zs.ensure(N * 512 * 4) && S->hs.ensure(N * 512 * 2) && S->dzA.ensure(N * 128 * 2) && S->dzB.ensure(N * 128 * 2) &&
              S->dz1.ensure(N * 128 * 4) && S->loss_part.ensure(nwg * 4) && S->g5_part.ensure(nwg * 645 * 4) &&
              S->gW_part.ensure(3 * nwg * 128 * 128 * 4) && S->gb_part.ensure(3 * nwg * 128 * 4) && S->g1_part.ensure(nwg * 640 * 4) &&
              S->grad.ensure(QLX_SUR_NPARAM * 4) && S->loss.ensure(4) && S->bf.ensure(N * 4 * 4) && S->bt.ensure(N * 5 * 4);
    if (!ok) return ctx_fail(ctx, QLX_ERR_NOMEM, "surrogate work allocation failed");
    S->w.zs = S->zs.as<float>();
    S->w.hs = S->hs.as<uint16_t>();
    S->w.dzA = S->dzA.as<uint16_t>();
    S->w.dzB = S->dzB.as<uint16_t>();
    S->w.dz1 = S->dz1.as<float>();
    S->w.loss_part = S->loss_part.as<float>();
    S->w.g5_part = S->g5_part.as<float>();
    S->w.gW_part = S->gW_part.as<float>();
    S->w.gb_part = S->gb_part.as<float>();
    S->w.g1_part = S->g1_part.as<float>();
    S->w.grad = S->grad.as<float>();
    S->w.loss = S->loss.as<float>();
    return QLX_OK;
}

SurState* state(void* ctx) { return (SurState*)ctx_surrogate(ctx); }

}  // namespace

namespace qlx {

void* sur_state_new() { return new SurState(); }
void sur_state_destroy(void* s) { delete (SurState*)s; }
const SurDev* sur_state_net(void* s)
{
    SurState* S = (SurState*)s;
    return (S && S->trained) ? &S->net : nullptr;
}

}  // namespace qlx

extern "C" {

int qlx_surrogate_param_count(void* ctx) { return ctx ? QLX_SUR_NPARAM : 0; }

int qlx_surrogate_get_params(void* ctx, float* params)
{
    if (!ctx || !params) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    int rc = ensure_params(ctx, S);
    if (rc) return rc;
    SURCHK(ctx, hipMemcpy(params, S->P.p, QLX_SUR_NPARAM * 4, hipMemcpyDeviceToHost));
    return QLX_OK;
}

int qlx_surrogate_set_params(void* ctx, const float* params)
{
    if (!ctx || !params) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    int rc = ensure_params(ctx, S);
    if (rc) return rc;
    hipStream_t st = ctx_stream(ctx);
    SURCHK(ctx, hipMemcpyAsync(S->P.p, params, QLX_SUR_NPARAM * 4, hipMemcpyHostToDevice, st));
    launch_sur_refresh(st, S->sp);
    SURCHK(ctx, hipStreamSynchronize(st));
    S->trained = true;
    return QLX_OK;
}

int qlx_surrogate_train(void* ctx, const qlx_surrogate_cfg* cfg, double* final_loss)
{
    if (!ctx || !cfg) return QLX_ERR_ARG;
    if (cfg->n_samples < 64 || cfg->steps < 1 || cfg->batch < 64 || cfg->batch > (1 << 20) || !(cfg->lr > 0) || !(cfg->tol > 0))
        return ctx_fail(ctx, QLX_ERR_ARG, "surrogate cfg: need n_samples >= 64, steps >= 1, 64 <= batch <= 2^20, lr > 0, tol > 0");
    if (!(cfg->r_min > 0 && cfg->r_max > cfg->r_min && cfg->q_min > 0 && cfg->q_max <= 1 && cfg->q_max >= cfg->q_min && cfg->rs_min > 0 &&
          cfg->rs_max >= cfg->rs_min))
        return ctx_fail(ctx, QLX_ERR_ARG, "surrogate cfg: bad sampling box");
    SurState* S = state(ctx);
    int rc = ensure_params(ctx, S);
    if (rc) return rc;
    rc = ensure_work(ctx, S, cfg->batch);
    if (rc) return rc;
    hipStream_t st = ctx_stream(ctx);
    const int ns = cfg->n_samples;
    Rng rng(cfg->seed);

    // initial weights: He-normal for the SiLU layers, small for the output layer; biases 0
    {
        std::vector<float> P(QLX_SUR_NPARAM, 0.0f);
        for (int i = 0; i < 512; i++) P[i] = (float)(rng.normal() * sqrt(2.0 / 4));
        for (int l = 0; l < 3; l++)
            for (int i = 0; i < 128 * 128; i++) P[640 + l * (16384 + 128) + i] = (float)(rng.normal() * sqrt(2.0 / 128));
        for (int i = 0; i < 640; i++) P[50176 + i] = (float)(rng.normal() * sqrt(1.0 / 128));
        std::vector<float> zero(QLX_SUR_NPARAM, 0.0f);
        SURCHK(ctx, hipMemcpyAsync(S->P.p, P.data(), P.size() * 4, hipMemcpyHostToDevice, st));
        SURCHK(ctx, hipMemcpyAsync(S->m.p, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, st));
        SURCHK(ctx, hipMemcpyAsync(S->v.p, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, st));
        SURCHK(ctx, hipMemcpyAsync(S->step.p, zero.data(), 4, hipMemcpyHostToDevice, st));
        launch_sur_refresh(st, S->sp);
        SURCHK(ctx, hipStreamSynchronize(st));
    }

    // training set: lens-frame points (ln r uniform, phi uniform, q uniform, ln rs uniform), ks = 1,
    // integrals by the GPU quadrature kernel, then features and network targets
    DBuf in4, j5, feat, targ;
    if (!in4.ensure((size_t)ns * 32) || !j5.ensure((size_t)ns * 40) || !feat.ensure((size_t)ns * 16) || !targ.ensure((size_t)ns * 20))
        return ctx_fail(ctx, QLX_ERR_NOMEM, "surrogate training-set allocation failed");
    {
        std::vector<double> h((size_t)ns * 4);
        const double lr0 = log(cfg->r_min), lr1 = log(cfg->r_max), ls0 = log(cfg->rs_min), ls1 = log(cfg->rs_max);
        for (int i = 0; i < ns; i++) {
            const double r = exp(lr0 + (lr1 - lr0) * rng.uniform());
            const double phi = 6.283185307179586 * rng.uniform();
            h[4 * i] = r * cos(phi);
            h[4 * i + 1] = r * sin(phi);
            h[4 * i + 2] = cfg->q_min + (cfg->q_max - cfg->q_min) * rng.uniform();
            h[4 * i + 3] = exp(ls0 + (ls1 - ls0) * rng.uniform());
        }
        SURCHK(ctx, hipMemcpyAsync(in4.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, st));
        launch_nfw_samples(st, ns, in4.as<double>(), cfg->tol, j5.as<double>());
        launch_sur_targets(st, ns, in4.as<double>(), j5.as<double>(), feat.as<float>(), targ.as<float>());
        std::vector<float> t((size_t)ns * 5);
        SURCHK(ctx, hipMemcpyAsync(t.data(), targ.p, t.size() * 4, hipMemcpyDeviceToHost, st));
        SURCHK(ctx, hipStreamSynchronize(st));
        for (int k = 0; k < 5; k++) {
            double m = 0, v = 0;
            for (int i = 0; i < ns; i++) m += t[5 * i + k];
            m /= ns;
            for (int i = 0; i < ns; i++) v += (t[5 * i + k] - m) * (t[5 * i + k] - m);
            const double sd = sqrt(v / ns) > 1e-12 ? sqrt(v / ns) : 1.0;
            S->h_norm[k] = (float)m;
            S->h_norm[5 + k] = (float)sd;
            for (int i = 0; i < ns; i++) t[5 * i + k] = (float)((t[5 * i + k] - m) / sd);
        }
        SURCHK(ctx, hipMemcpyAsync(targ.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, st));
        SURCHK(ctx, hipMemcpyAsync(S->norm.p, S->h_norm, 40, hipMemcpyHostToDevice, st));
        SURCHK(ctx, hipStreamSynchronize(st));
    }

    // one optimizer step = gather -> forward/backward -> reduce -> Adam; captured once in a HIP
    // graph (kernel nodes only) and replayed, or launched directly when cfg->graph == 0
    const float lr = (float)cfg->lr, lr_min = (float)(cfg->lr * 0.01);
    auto step_launches = [&]() {
        launch_sur_gather(st, cfg->batch, ns, cfg->seed, S->sp.step, feat.as<float>(), targ.as<float>(), S->bf.as<float>(),
                          S->bt.as<float>());
        launch_sur_fwdbwd(st, S->net, cfg->batch, S->bf.as<float>(), S->bt.as<float>(), S->w);
        launch_sur_adam(st, S->sp, S->w.grad, cfg->steps, lr, lr_min);
    };
    if (cfg->graph) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        SURCHK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        step_launches();
        SURCHK(ctx, hipStreamEndCapture(st, &g));
        SURCHK(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        hipError_t e = hipSuccess;
        for (int it = 0; it < cfg->steps && e == hipSuccess; it++) e = hipGraphLaunch(ge, st);
        hipGraphExecDestroy(ge);
        hipGraphDestroy(g);
        SURCHK(ctx, e);
    } else {
        for (int it = 0; it < cfg->steps; it++) step_launches();
    }
    float l = 0.0f;
    SURCHK(ctx, hipMemcpyAsync(&l, S->w.loss, 4, hipMemcpyDeviceToHost, st));
    SURCHK(ctx, hipStreamSynchronize(st));
    in4.release();
    j5.release();
    feat.release();
    targ.release();
    if (final_loss) *final_loss = l;
    S->trained = true;
    return QLX_OK;
}

int qlx_surrogate_save(void* ctx, const char* path)
{
    if (!ctx || !path) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    if (!S->trained) return ctx_fail(ctx, QLX_ERR_STATE, "no trained surrogate to save");
    std::vector<float> P(QLX_SUR_NPARAM);
    SURCHK(ctx, hipMemcpy(P.data(), S->P.p, P.size() * 4, hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) return ctx_fail(ctx, QLX_ERR_ARG, "cannot open surrogate file for writing");
    const char magic[8] = {'Q', 'L', 'X', 'S', 'U', 'R', '1', 0};
    const int n = QLX_SUR_NPARAM;
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&n, 4, 1, f) == 1 && fwrite(S->h_norm, 4, 10, f) == 10 &&
              fwrite(P.data(), 4, P.size(), f) == P.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? QLX_OK : ctx_fail(ctx, QLX_ERR_ARG, "surrogate file write failed");
}

int qlx_surrogate_load(void* ctx, const char* path)
{
    if (!ctx || !path) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    int rc = ensure_params(ctx, S);
    if (rc) return rc;
    FILE* f = fopen(path, "rb");
    if (!f) return ctx_fail(ctx, QLX_ERR_ARG, "cannot open surrogate file");
    char magic[8];
    int n = 0;
    float norm[10];
    std::vector<float> P(QLX_SUR_NPARAM);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "QLXSUR1", 8) == 0 && fread(&n, 4, 1, f) == 1 && n == QLX_SUR_NPARAM &&
              fread(norm, 4, 10, f) == 10 && fread(P.data(), 4, P.size(), f) == P.size();
    fclose(f);
    if (!ok) return ctx_fail(ctx, QLX_ERR_ARG, "not a qlx surrogate file (or wrong size)");
    memcpy(S->h_norm, norm, sizeof(norm));
    hipStream_t st = ctx_stream(ctx);
    SURCHK(ctx, hipMemcpyAsync(S->norm.p, S->h_norm, 40, hipMemcpyHostToDevice, st));
    SURCHK(ctx, hipMemcpyAsync(S->P.p, P.data(), P.size() * 4, hipMemcpyHostToDevice, st));
    launch_sur_refresh(st, S->sp);
    SURCHK(ctx, hipStreamSynchronize(st));
    S->trained = true;
    return QLX_OK;
}

int qlx_surrogate_eval(void* ctx, int n, const float* x, float* y)
{
    if (!ctx || n < 0 || (n > 0 && (!x || !y))) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    if (!S->trained) return ctx_fail(ctx, QLX_ERR_STATE, "surrogate not trained or loaded");
    if (n == 0) return QLX_OK;
    DBuf dx, dy;
    if (!dx.ensure((size_t)n * 16) || !dy.ensure((size_t)n * 20)) return ctx_fail(ctx, QLX_ERR_NOMEM, "allocation failed");
    hipStream_t st = ctx_stream(ctx);
    SURCHK(ctx, hipMemcpyAsync(dx.p, x, (size_t)n * 16, hipMemcpyHostToDevice, st));
    launch_sur_eval(st, S->net, n, dx.as<float>(), dy.as<float>());
    SURCHK(ctx, hipMemcpyAsync(y, dy.p, (size_t)n * 20, hipMemcpyDeviceToHost, st));
    SURCHK(ctx, hipStreamSynchronize(st));
    dx.release();
    dy.release();
    return QLX_OK;
}

int qlx_surrogate_grad(void* ctx, int n, const float* x, const float* y, float* loss, float* grad)
{
    if (!ctx || n < 1 || !x || !y || !loss || !grad) return QLX_ERR_ARG;
    SurState* S = state(ctx);
    int rc = ensure_params(ctx, S);
    if (rc) return rc;
    rc = ensure_work(ctx, S, n);
    if (rc) return rc;
    hipStream_t st = ctx_stream(ctx);
    SURCHK(ctx, hipMemcpyAsync(S->bf.p, x, (size_t)n * 16, hipMemcpyHostToDevice, st));
    SURCHK(ctx, hipMemcpyAsync(S->bt.p, y, (size_t)n * 20, hipMemcpyHostToDevice, st));
    launch_sur_fwdbwd(st, S->net, n, S->bf.as<float>(), S->bt.as<float>(), S->w);
    SURCHK(ctx, hipMemcpyAsync(loss, S->w.loss, 4, hipMemcpyDeviceToHost, st));
    SURCHK(ctx, hipMemcpyAsync(grad, S->w.grad, QLX_SUR_NPARAM * 4, hipMemcpyDeviceToHost, st));
    SURCHK(ctx, hipStreamSynchronize(st));
    return QLX_OK;
}

}  // extern "C"
