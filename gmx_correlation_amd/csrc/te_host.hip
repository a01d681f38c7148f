// Host side of the transfer-entropy matrix (gmxcorr_te_matrix / _pairs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gmxcorr.h"
#include "te_internal.h"

namespace gmxcorr {
int set_error(int code, const std::string& msg);
int set_hip_error(hipError_t e, const char* where);
bool device_ok();
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

// Pass-1 kernel of later calls (gmxcorr_te_set_path): 0 automatic, 1 heap, 2 select.
int g_path = 0;

// mat != nullptr: full matrix over all ordered pairs (diagonal 0);
// otherwise te_out[p] for the given list.
int te_run(const double* traj, int natoms, int nframes, int k, int lag, long long npairs,
           const int* a1, const int* a2, double* mat, double* te_out, double* ms)
{
    if (!traj || natoms < 1 || nframes < 1 || k < 1 || lag < 0)
        return gmxcorr::set_error(GMXCORR_EINVAL, "te: bad arguments");
    // the count kernel addresses an atom's frames by 32-bit byte offsets
    if ((long long)nframes * 24 >= (1LL << 32))
        return gmxcorr::set_error(GMXCORR_EUNSUP, "te: more than 178,956,970 frames");
    const int Nv = nframes - lag;
    if (mat) {
        for (size_t t = 0; t < (size_t)natoms * natoms; ++t) mat[t] = 0.0;
        if (Nv <= k) {
            std::fprintf(stderr, "Warning: TE: N_valid=%d <= k=%d; skipping.\n", Nv, k);
            return GMXCORR_OK;
        }
    } else if (Nv <= k) {
        return gmxcorr::set_error(GMXCORR_EINVAL, "te: need nframes - lag > k");
    }
    if (g_path == 1 && k > gmxcorr::kTeHeapMaxK)
        return gmxcorr::set_error(GMXCORR_EUNSUP, "te: the heap kernel takes k <= 128");
    for (size_t t = 0; t < (size_t)natoms * nframes * 3; ++t)
        if (!std::isfinite(traj[t])) return gmxcorr::set_error(GMXCORR_EINVAL, "te: non-finite coordinate");
    if (!gmxcorr::device_ok()) return gmxcorr::set_error(GMXCORR_ENODEV, "no gfx950 device");
    std::vector<int> h1, h2;
    if (mat) {
        for (int x = 0; x < natoms; ++x)
            for (int y = 0; y < natoms; ++y)
                if (x != y) { h1.push_back(x); h2.push_back(y); }
        npairs = (long long)h1.size();
        a1 = h1.data();
        a2 = h2.data();
    }
    for (long long p = 0; p < npairs; ++p)
        if (a1[p] < 0 || a2[p] < 0 || a1[p] >= natoms || a2[p] >= natoms || a1[p] == a2[p])
            return gmxcorr::set_error(GMXCORR_EINVAL, "te: pair index out of range or a1 == a2");
    // a list goes to the device in target-major order (stable), so that
    // pairs with one target atom sit next to each other and the count
    // kernel takes them together; order[i] is the caller's index of the
    // i-th pair launched
    std::vector<long long> order;
    if (!mat) {
        order.resize((size_t)npairs);
        for (long long p = 0; p < npairs; ++p) order[(size_t)p] = p;
        std::stable_sort(order.begin(), order.end(),
                         [&](long long x, long long y) { return a1[x] < a1[y]; });
        h1.resize((size_t)npairs);
        h2.resize((size_t)npairs);
        for (long long i = 0; i < npairs; ++i) {
            h1[(size_t)i] = a1[order[(size_t)i]];
            h2[(size_t)i] = a2[order[(size_t)i]];
        }
        a1 = h1.data();
        a2 = h2.data();
    }
    // digamma table of size Nv+1 as lib/transfer_entropy.cpp:41-48
    std::vector<double> psi(Nv + 1, 0.0);
    psi[1] = -0.57721566490153;
    for (int i = 1; i < Nv; ++i) psi[i + 1] = psi[i] + 1.0 / i;
    const int nqb = (Nv + 63) / 64;
    DevBuf<double> dTraj, dPsi, dPart, dMat;
    DevBuf<float> dRoot;
    DevBuf<int> d1, d2;
    HIPCHK(dTraj.alloc((size_t)natoms * nframes * 3));
    HIPCHK(dPsi.alloc(psi.size()));
    HIPCHK(dPart.alloc((size_t)npairs * nqb));
    const long long chunk = gmxcorr::te_chunk_pairs(nqb);
    HIPCHK(dRoot.alloc((size_t)(npairs < chunk ? npairs : chunk) * nqb * 64));
    HIPCHK(d1.alloc(npairs));
    HIPCHK(d2.alloc(npairs));
    HIPCHK(dMat.alloc(mat ? (size_t)natoms * natoms : (size_t)npairs));
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemcpyAsync(dTraj.p, traj, (size_t)natoms * nframes * 3 * sizeof(double),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPsi.p, psi.data(), psi.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d1.p, a1, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d2.p, a2, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    gmxcorr::TeArgs ta{dTraj.p, nframes, lag, k, g_path, d1.p, d2.p, npairs, nqb, dPsi.p, dPart.p, dRoot.p};
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(gmxcorr::launch_te(ta, s));
    HIPCHK(hipEventRecord(e1, s));
    if (mat) {
        std::vector<double> zero((size_t)natoms * natoms, 0.0);
        HIPCHK(hipMemcpyAsync(dMat.p, zero.data(), zero.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(gmxcorr::launch_te_finalize(ta, dMat.p, natoms, s));
        HIPCHK(hipMemcpyAsync(mat, dMat.p, zero.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    } else {
        // te_out[p]: per-pair sums of the query blocks / Nv on the host
        std::vector<double> part((size_t)npairs * nqb);
        HIPCHK(hipMemcpyAsync(part.data(), dPart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (long long p = 0; p < npairs; ++p) {
            double acc = 0.0;
            for (int b = 0; b < nqb; ++b) acc += part[(size_t)p * nqb + b];
            te_out[order[(size_t)p]] = acc / Nv;
        }
    }
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return GMXCORR_OK;
}

}  // namespace

extern "C" {

int gmxcorr_te_matrix(const double* traj, int natoms, int nframes, int k, int lag, double* mat,
                      double* ms)
{
    if (!mat) return gmxcorr::set_error(GMXCORR_EINVAL, "te_matrix: null output");
    return te_run(traj, natoms, nframes, k, lag, 0, nullptr, nullptr, mat, nullptr, ms);
}

int gmxcorr_te_pairs(const double* traj, int natoms, int nframes, int k, int lag,
                     long long npairs, const int* a1, const int* a2, double* te, double* ms)
{
    if (!te || !a1 || !a2 || npairs < 1) return gmxcorr::set_error(GMXCORR_EINVAL, "te_pairs: bad arguments");
    return te_run(traj, natoms, nframes, k, lag, npairs, a1, a2, nullptr, te, ms);
}

int gmxcorr_te_set_path(int path)
{
    if (path < 0 || path > 2) return gmxcorr::set_error(GMXCORR_EINVAL, "te_set_path: path must be 0, 1 or 2");
    g_path = path;
    return GMXCORR_OK;
}

}  // extern "C"
