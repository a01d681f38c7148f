// Host side of the Gaussian (-linear) and Gaussian-copula (-gcmi) matrices:
// gmxcorr_gauss_matrix / gmxcorr_gcmi_matrix of include/gmxcorr.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "dense_internal.h"
#include "gmxcorr.h"

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

struct DenseBuffers {
    double* traj = nullptr;
    double* rows = nullptr;
    double* z = nullptr;
    double* ztab = nullptr;   // copula: normal score of each rank
    _Float16* hi = nullptr;
    _Float16* lo = nullptr;
    float* C = nullptr;
    float* Cp = nullptr;   // split-K partial matrices (nullptr: one split into C)
    double* mat = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev[6] = {};
    ~DenseBuffers()
    {
        if (s) (void)hipStreamSynchronize(s);
        for (void* p : {(void*)traj, (void*)rows, (void*)z, (void*)ztab, (void*)hi, (void*)lo,
                        (void*)C, (void*)Cp, (void*)mat})
            if (p) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }
};

float ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Frames per GEMM slab set by gmxcorr_dense_set_slab_frames (0: the kernel's
// own limit).
int g_slab_frames = 0;

// mode 0 Gaussian, 1 GCMI.
int dense_matrix(int mode, const double* traj, int natoms, int nframes, double* mat,
                 double* cov_out, double* z_out, gmxcorr_dense_stats* st)
{
    if (!traj || !mat) return gmxcorr::set_error(GMXCORR_EINVAL, "dense: null argument");
    if (natoms < 2) return gmxcorr::set_error(GMXCORR_EINVAL, "need at least 2 atoms");
    if (mode == 1 && nframes < 6)
        return gmxcorr::set_error(GMXCORR_EINVAL,
                                  "gcmi: need at least 6 frames for a non-singular 6x6 covariance");
    if (nframes < 1) return gmxcorr::set_error(GMXCORR_EINVAL, "need at least 1 frame");
    if (!gmxcorr::device_ok()) return gmxcorr::set_error(GMXCORR_ENODEV, "no gfx950 device");
    const int M = 3 * natoms, N = nframes;
    // frame slabs of Ks frames (a multiple of 64) whose M x Ks operands stay
    // within the GEMM kernel's 32-bit offsets; one slab for most inputs
    const long long Kall = (N + 63) / 64 * 64;
    long long Kmax = gmxcorr::kGemmMaxElems / M / 64 * 64;
    if (g_slab_frames > 0) Kmax = std::min<long long>(Kmax, (g_slab_frames + 63) / 64 * 64);
    if (Kmax < 64) return gmxcorr::set_error(GMXCORR_EUNSUP, "dense: more than 2^25 atoms");
    const int Kp = (int)std::min(Kall, Kmax);
    const int nslab = (int)((Kall + Kp - 1) / Kp);
    const size_t tbytes = (size_t)natoms * N * 3 * sizeof(double);
    DenseBuffers b;
    HIPCHK(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
    for (auto& e : b.ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipMalloc(&b.traj, tbytes));
    HIPCHK(hipMalloc(&b.rows, tbytes));
    if (mode == 1) {
        HIPCHK(hipMalloc(&b.z, tbytes));
        HIPCHK(hipMalloc(&b.ztab, (size_t)N * sizeof(double)));
    }
    HIPCHK(hipMalloc(&b.hi, (size_t)M * Kp * sizeof(_Float16)));
    HIPCHK(hipMalloc(&b.lo, (size_t)M * Kp * sizeof(_Float16)));
    HIPCHK(hipMalloc(&b.C, (size_t)M * M * sizeof(float)));
    const int nsplit = gmxcorr::dense_cov_splits(M, Kp);
    if (nsplit * nslab > 1)
        HIPCHK(hipMalloc(&b.Cp, (size_t)nsplit * nslab * M * M * sizeof(float)));
    HIPCHK(hipMalloc(&b.mat, (size_t)natoms * natoms * sizeof(double)));

    // power-of-two scale putting the largest |value| near 2^10 for FP16
    double vmax = 0.0;
    if (mode == 0) {
        for (size_t t = 0; t < (size_t)natoms * N * 3; ++t) vmax = std::fmax(vmax, std::fabs(traj[t]));
    } else {
        vmax = 6.0;  // copula scores are normal scores, |z| < 6 for N < 5e8
    }
    if (!std::isfinite(vmax)) return gmxcorr::set_error(GMXCORR_EINVAL, "non-finite coordinates");
    const double scale = vmax > 0.0 ? std::ldexp(1.0, 10 - std::ilogb(vmax) - 1) : 1.0;
    const float out_scale = (float)(1.0 / (scale * scale * N));

    HIPCHK(hipEventRecord(b.ev[0], b.s));
    HIPCHK(hipMemcpyAsync(b.traj, traj, tbytes, hipMemcpyHostToDevice, b.s));
    HIPCHK(gmxcorr::launch_dense_rows(b.traj, natoms, N, b.rows, b.s));
    HIPCHK(hipEventRecord(b.ev[1], b.s));
    const double* V = b.rows;
    if (mode == 1) {
        HIPCHK(gmxcorr::launch_dense_copula(b.rows, M, N, N, b.z, N, b.ztab, b.s));
        V = b.z;
    }
    HIPCHK(hipEventRecord(b.ev[2], b.s));
    float split_ms = 0.f, gemm_ms = 0.f;
    if (nslab == 1) {
        HIPCHK(gmxcorr::launch_dense_split(V, M, N, N, Kp, scale, b.hi, b.lo, b.s));
        HIPCHK(hipEventRecord(b.ev[3], b.s));
        HIPCHK(gmxcorr::launch_dense_cov(b.hi, b.lo, M, Kp, out_scale, nsplit > 1 ? b.Cp : b.C,
                                         nsplit, b.C, b.s));
        HIPCHK(hipEventRecord(b.ev[4], b.s));
    } else {
        // each slab's split, then its GEMM into its own nsplit partial
        // matrices; one reduce over all of them
        for (int sl = 0; sl < nslab; ++sl) {
            const long long f0 = (long long)sl * Kp;
            const int nf = (int)std::min<long long>(Kp, N - f0);
            HIPCHK(hipEventRecord(b.ev[2], b.s));
            HIPCHK(gmxcorr::launch_dense_split(V + f0, M, nf, N, Kp, scale, b.hi, b.lo, b.s));
            HIPCHK(hipEventRecord(b.ev[3], b.s));
            HIPCHK(gmxcorr::launch_dense_cov(b.hi, b.lo, M, Kp, out_scale,
                                             b.Cp + (size_t)sl * nsplit * M * M, nsplit, b.C, b.s,
                                             0));
            HIPCHK(hipEventRecord(b.ev[4], b.s));
            HIPCHK(hipEventSynchronize(b.ev[4]));
            split_ms += ms_between(b.ev[2], b.ev[3]);
            gemm_ms += ms_between(b.ev[3], b.ev[4]);
        }
        HIPCHK(gmxcorr::launch_dense_reduce(b.Cp, M, nsplit * nslab, b.C, b.s));
        HIPCHK(hipEventRecord(b.ev[4], b.s));
    }
    HIPCHK(gmxcorr::launch_dense_pairs(b.C, natoms, mode, b.mat, b.s));
    HIPCHK(hipEventRecord(b.ev[5], b.s));
    HIPCHK(hipMemcpyAsync(mat, b.mat, (size_t)natoms * natoms * sizeof(double),
                          hipMemcpyDeviceToHost, b.s));
    HIPCHK(hipStreamSynchronize(b.s));
    if (cov_out) {
        std::vector<float> c((size_t)M * M);
        HIPCHK(hipMemcpy(c.data(), b.C, c.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < c.size(); ++t) cov_out[t] = c[t];
    }
    if (z_out && mode == 1)
        HIPCHK(hipMemcpy(z_out, b.z, tbytes, hipMemcpyDeviceToHost));
    if (st) {
        st->ms_load = ms_between(b.ev[0], b.ev[1]);
        st->ms_copula = ms_between(b.ev[1], b.ev[2]);
        st->ms_split = nslab == 1 ? ms_between(b.ev[2], b.ev[3]) : split_ms;
        st->ms_gemm = nslab == 1 ? ms_between(b.ev[3], b.ev[4]) : gemm_ms;
        st->ms_pairs = ms_between(b.ev[4], b.ev[5]);
        st->gemm_flops = 2.0 * M * (double)M * N;
        st->natoms = natoms;
        st->nframes = N;
    }
    return GMXCORR_OK;
}

}  // namespace

extern "C" {

int gmxcorr_dense_set_slab_frames(int frames)
{
    if (frames < 0) return gmxcorr::set_error(GMXCORR_EINVAL, "slab frames must be >= 0");
    g_slab_frames = frames;
    return GMXCORR_OK;
}

int gmxcorr_gauss_matrix(const double* traj, int natoms, int nframes, double* mat,
                         double* cov_out, gmxcorr_dense_stats* st)
{
    return dense_matrix(0, traj, natoms, nframes, mat, cov_out, nullptr, st);
}

int gmxcorr_gcmi_matrix(const double* traj, int natoms, int nframes, double* mat,
                        double* cov_out, double* z_out, gmxcorr_dense_stats* st)
{
    return dense_matrix(1, traj, natoms, nframes, mat, cov_out, z_out, st);
}

}  // extern "C"
