// Host side of the learned surrogate: pair features on a loaded KSG plan and
// the MLP object (training, FP32 forward, BF16 inference) of include/gmxcorr.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "dense_internal.h"
#include "gmxcorr.h"
#include "ksg_internal.h"
#include "surrogate_internal.h"

namespace gmxcorr {
int set_error(int code, const std::string& msg);
int set_hip_error(hipError_t e, const char* where);
bool device_ok();
bool plan_view(const gmxcorr_ksg_plan* pl, PlanView* v);
}  // namespace gmxcorr

namespace {

#define HIPCHK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return gmxcorr::set_hip_error(e_, #expr);       \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(T)); }
};

}  // namespace

struct gmxcorr_mlp {
    int cap = 0;                  // samples per device batch
    hipStream_t s = nullptr;
    float* params = nullptr;
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    int step = 0;
    float* X = nullptr;           // [cap][16]
    float* Y = nullptr;           // [cap]
    float* out = nullptr;         // [cap][2]
    gmxcorr::MlpScratch sc{};
    __bf16* W1b = nullptr;
    __bf16* W2b = nullptr;
    __bf16* W3b = nullptr;
    float* bias = nullptr;
    bool bf16_stale = true;
};

extern "C" {

int gmxcorr_ksg_plan_features(gmxcorr_ksg_plan* plan, long long npairs, const int* a1,
                              const int* a2, int nsample, float* feat, double* ms)
{
    gmxcorr::PlanView pv;
    if (!gmxcorr::plan_view(plan, &pv)) return gmxcorr::set_error(GMXCORR_EINVAL, "plan not loaded");
    if (npairs <= 0 || !a1 || !a2 || !feat || nsample < 1)
        return gmxcorr::set_error(GMXCORR_EINVAL, "features: bad arguments");
    for (long long p = 0; p < npairs; ++p)
        if (a1[p] < 0 || a2[p] < 0 || a1[p] >= pv.natoms || a2[p] >= pv.natoms || a1[p] == a2[p])
            return gmxcorr::set_error(GMXCORR_EINVAL, "pair index out of range or a1 == a2");
    const int n = pv.natoms, N = pv.N, M = 3 * n;
    const int Kp = (N + 63) / 64 * 64;
    hipStream_t s = pv.stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    DevBuf<double> Vc, kurt;
    DevBuf<_Float16> hi, lo;
    DevBuf<float> C, Cp, rad, fk, q50, q90, dfeat;
    DevBuf<int> d1, d2;
    HIPCHK(Vc.alloc((size_t)M * N));
    HIPCHK(kurt.alloc(M));
    HIPCHK(hi.alloc((size_t)M * Kp));
    HIPCHK(lo.alloc((size_t)M * Kp));
    HIPCHK(C.alloc((size_t)M * M));
    const int nsplit = gmxcorr::dense_cov_splits(M, Kp);
    if (nsplit > 1) HIPCHK(Cp.alloc((size_t)nsplit * M * M));
    HIPCHK(rad.alloc((size_t)n * nsample));
    HIPCHK(fk.alloc(n));
    HIPCHK(q50.alloc(n));
    HIPCHK(q90.alloc(n));
    HIPCHK(dfeat.alloc((size_t)npairs * 16));
    HIPCHK(d1.alloc(npairs));
    HIPCHK(d2.alloc(npairs));
    HIPCHK(hipMemcpyAsync(d1.p, a1, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d2.p, a2, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(gmxcorr::launch_row_stats(pv.X, M, N, Vc.p, kurt.p, s));
    // |z| of a standardised sample is below sqrt(N); scale the FP16 split to 2^10
    const double scale = std::ldexp(1.0, 9 - std::ilogb(std::sqrt((double)N) + 1.0));
    HIPCHK(gmxcorr::launch_dense_split(Vc.p, M, N, N, Kp, scale, hi.p, lo.p, s));
    HIPCHK(gmxcorr::launch_dense_cov(hi.p, lo.p, M, Kp, (float)(1.0 / (scale * scale * N)),
                                     nsplit > 1 ? Cp.p : C.p, nsplit, C.p, s));
    HIPCHK(gmxcorr::launch_atom_knn_radius(pv.D, n, N, pv.P, pv.K, nsample, rad.p, s));
    std::vector<double> hk(M);
    std::vector<float> hr((size_t)n * nsample);
    HIPCHK(hipMemcpyAsync(hk.data(), kurt.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hr.data(), rad.p, hr.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<float> hfk(n), hq50(n), hq90(n), tmp(nsample);
    for (int a = 0; a < n; ++a) {
        hfk[a] = (float)((hk[3 * a] + hk[3 * a + 1] + hk[3 * a + 2]) / 3.0);
        std::copy(hr.begin() + (size_t)a * nsample, hr.begin() + (size_t)(a + 1) * nsample, tmp.begin());
        std::sort(tmp.begin(), tmp.end());
        hq50[a] = tmp[nsample / 2];
        hq90[a] = tmp[std::min(nsample - 1, (int)(0.9 * nsample))];
    }
    HIPCHK(hipMemcpyAsync(fk.p, hfk.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(q50.p, hq50.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(q90.p, hq90.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(gmxcorr::launch_pair_features(C.p, n, d1.p, d2.p, npairs, fk.p, q50.p, q90.p, dfeat.p, s));
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipMemcpyAsync(feat, dfeat.p, (size_t)npairs * 16 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return GMXCORR_OK;
}

int gmxcorr_mlp_num_params(void) { return gmxcorr::kMlpParams; }

void gmxcorr_mlp_destroy(gmxcorr_mlp* m)
{
    if (!m) return;
    if (m->s) (void)hipStreamSynchronize(m->s);
    for (void* p : {(void*)m->params, (void*)m->grad, (void*)m->m, (void*)m->v, (void*)m->X,
                    (void*)m->Y, (void*)m->out, (void*)m->sc.H1, (void*)m->sc.G1, (void*)m->sc.H2,
                    (void*)m->sc.G2, (void*)m->sc.D1, (void*)m->sc.D2, (void*)m->sc.D3,
                    (void*)m->sc.loss, (void*)m->W1b, (void*)m->W2b, (void*)m->W3b,
                    (void*)m->bias})
        if (p) (void)hipFree(p);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
}

int gmxcorr_mlp_create(int max_batch, gmxcorr_mlp** out)
{
    if (!out || max_batch < 1) return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_create: bad arguments");
    *out = nullptr;
    if (!gmxcorr::device_ok()) return gmxcorr::set_error(GMXCORR_ENODEV, "no gfx950 device");
    auto* m = new gmxcorr_mlp();
    m->cap = max_batch;
    auto bail = [&](hipError_t e, const char* w) {
        gmxcorr_mlp_destroy(m);
        return gmxcorr::set_hip_error(e, w);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
    const int P = gmxcorr::kMlpParams;
    const size_t B = (size_t)max_batch;
    float** fl[] = {&m->params, &m->grad, &m->m, &m->v};
    for (float** f : fl)
        if ((e = hipMalloc(f, P * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(params)");
    if ((e = hipMalloc(&m->X, B * 16 * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(X)");
    if ((e = hipMalloc(&m->Y, B * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(Y)");
    if ((e = hipMalloc(&m->out, B * 2 * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(out)");
    float** sb[] = {&m->sc.H1, &m->sc.G1, &m->sc.H2, &m->sc.G2, &m->sc.D1, &m->sc.D2};
    for (float** f : sb)
        if ((e = hipMalloc(f, B * 64 * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(scratch)");
    if ((e = hipMalloc(&m->sc.D3, B * 2 * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(D3)");
    if ((e = hipMalloc(&m->sc.loss, B * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(loss)");
    if ((e = hipMalloc(&m->W1b, 1024 * sizeof(__bf16))) != hipSuccess) return bail(e, "hipMalloc(W1b)");
    if ((e = hipMalloc(&m->W2b, 4096 * sizeof(__bf16))) != hipSuccess) return bail(e, "hipMalloc(W2b)");
    if ((e = hipMalloc(&m->W3b, 2048 * sizeof(__bf16))) != hipSuccess) return bail(e, "hipMalloc(W3b)");
    if ((e = hipMalloc(&m->bias, 130 * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(bias)");
    std::vector<float> zero(P, 0.0f);
    (void)hipMemcpy(m->m, zero.data(), P * sizeof(float), hipMemcpyHostToDevice);
    (void)hipMemcpy(m->v, zero.data(), P * sizeof(float), hipMemcpyHostToDevice);
    (void)hipMemcpy(m->params, zero.data(), P * sizeof(float), hipMemcpyHostToDevice);
    *out = m;
    return GMXCORR_OK;
}

int gmxcorr_mlp_set_params(gmxcorr_mlp* m, const float* params)
{
    if (!m || !params) return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_set_params: null");
    HIPCHK(hipMemcpy(m->params, params, gmxcorr::kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> zero(gmxcorr::kMlpParams, 0.0f);
    HIPCHK(hipMemcpy(m->m, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->v, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice));
    m->step = 0;
    m->bf16_stale = true;
    return GMXCORR_OK;
}

int gmxcorr_mlp_get_params(gmxcorr_mlp* m, float* params)
{
    if (!m || !params) return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_get_params: null");
    HIPCHK(hipMemcpy(params, m->params, gmxcorr::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return GMXCORR_OK;
}

// Loss and gradient of one batch without an update (grad may be NULL).
int gmxcorr_mlp_grad(gmxcorr_mlp* m, const float* x, const float* y, int B, float* grad,
                     float* loss)
{
    if (!m || !x || !y || B < 1 || B > m->cap) return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_grad: bad batch");
    HIPCHK(hipMemcpyAsync(m->X, x, (size_t)B * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(m->Y, y, (size_t)B * sizeof(float), hipMemcpyHostToDevice, m->s));
    HIPCHK(gmxcorr::launch_mlp_fwd_bwd(m->params, m->X, m->Y, B, m->sc, nullptr, m->s));
    HIPCHK(gmxcorr::launch_mlp_grad(m->X, m->sc, B, m->grad, m->s));
    if (grad)
        HIPCHK(hipMemcpyAsync(grad, m->grad, gmxcorr::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->s));
    if (loss) {
        std::vector<float> l(B);
        HIPCHK(hipMemcpyAsync(l.data(), m->sc.loss, B * sizeof(float), hipMemcpyDeviceToHost, m->s));
        HIPCHK(hipStreamSynchronize(m->s));
        double acc = 0.0;
        for (float t : l) acc += t;
        *loss = (float)(acc / B);
    }
    HIPCHK(hipStreamSynchronize(m->s));
    return GMXCORR_OK;
}

// One pass over n samples in consecutive batches of `batch` (Adam, lr);
// mean_loss receives the mean of the batch losses.
int gmxcorr_mlp_train_epoch(gmxcorr_mlp* m, const float* x, const float* y, long long n, int batch,
                            float lr, float* mean_loss)
{
    if (!m || !x || !y || n < 1 || batch < 1 || batch > m->cap)
        return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_train_epoch: bad arguments");
    double acc = 0.0;
    int nb = 0;
    std::vector<float> l(batch);
    for (long long s0 = 0; s0 + batch <= n; s0 += batch) {
        HIPCHK(hipMemcpyAsync(m->X, x + s0 * 16, (size_t)batch * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
        HIPCHK(hipMemcpyAsync(m->Y, y + s0, (size_t)batch * sizeof(float), hipMemcpyHostToDevice, m->s));
        HIPCHK(gmxcorr::launch_mlp_fwd_bwd(m->params, m->X, m->Y, batch, m->sc, nullptr, m->s));
        HIPCHK(gmxcorr::launch_mlp_grad(m->X, m->sc, batch, m->grad, m->s));
        HIPCHK(gmxcorr::launch_adam(m->params, m->grad, m->m, m->v, lr, ++m->step, m->s));
        if (mean_loss) {
            HIPCHK(hipMemcpyAsync(l.data(), m->sc.loss, batch * sizeof(float), hipMemcpyDeviceToHost, m->s));
            HIPCHK(hipStreamSynchronize(m->s));
            for (float t : l) acc += t;
            ++nb;
        }
    }
    HIPCHK(hipStreamSynchronize(m->s));
    if (mean_loss) *mean_loss = nb ? (float)(acc / ((double)nb * batch)) : 0.0f;
    m->bf16_stale = true;
    return GMXCORR_OK;
}

// Predictions out[n][2] = (mu, log sigma^2); bf16 != 0 uses the BF16 MFMA
// kernel, else the FP32 training forward.  ms (optional): device time.
int gmxcorr_mlp_forward(gmxcorr_mlp* m, const float* x, long long n, int bf16, float* out, double* ms)
{
    if (!m || !x || !out || n < 1) return gmxcorr::set_error(GMXCORR_EINVAL, "mlp_forward: bad arguments");
    DevBuf<float> dx, dout;
    HIPCHK(dx.alloc((size_t)n * 16));
    HIPCHK(dout.alloc((size_t)n * 2));
    HIPCHK(hipMemcpyAsync(dx.p, x, (size_t)n * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
    if (bf16 && m->bf16_stale) {
        HIPCHK(gmxcorr::launch_mlp_to_bf16(m->params, m->W1b, m->W2b, m->W3b, m->bias, m->s));
        m->bf16_stale = false;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, m->s));
    if (bf16) {
        HIPCHK(gmxcorr::launch_mlp_infer_bf16(m->W1b, m->W2b, m->W3b, m->bias, dx.p, n, dout.p, m->s));
    } else {
        for (long long s0 = 0; s0 < n; s0 += m->cap) {
            const int b = (int)std::min<long long>(m->cap, n - s0);
            gmxcorr::MlpScratch none{};
            HIPCHK(gmxcorr::launch_mlp_fwd_bwd(m->params, dx.p + s0 * 16, nullptr, b, none,
                                               dout.p + s0 * 2, m->s));
        }
    }
    HIPCHK(hipEventRecord(e1, m->s));
    HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return GMXCORR_OK;
}

}  // extern "C"
