// Host side of libgmxcorr_hip.so: device plan for the KSG matrix, the
// extern "C" ABI (include/gmxcorr.h) and the reference's C++ GPU boundary
// (include/kraskov_gpu.h, replacing lib/kraskov_gpu.cu:210-308).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gmxcorr.h"
#include "kraskov_gpu.h"
#include "ksg_internal.h"

using gmxcorr::FixEntry;
using gmxcorr::KsgArgs;

namespace {
thread_local std::string g_err;
}  // namespace

namespace gmxcorr {

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int set_hip_error(hipError_t e, const char* where)
{
    return set_error(e == hipErrorOutOfMemory ? GMXCORR_ENOMEM : GMXCORR_EHIP,
                     std::string(where) + ": " + hipGetErrorString(e));
}

bool device_ok();

int check_rows(const double* xrows, long long nrows, int N, const char* who)
{
    for (long long r = 0; r < nrows; ++r) {
        const double* x = xrows + r * N;
        double lo = x[0], hi = x[0];
        for (int j = 0; j < N; ++j) {
            if (!std::isfinite(x[j]))
                return set_error(GMXCORR_EINVAL,
                                 std::string(who) +
                                     ": non-finite normalised coordinate (constant component: zero SD)");
            lo = std::min(lo, x[j]);
            hi = std::max(hi, x[j]);
        }
        if (!(hi - lo < kMaxRowSpread))
            return set_error(GMXCORR_EINVAL,
                             std::string(who) +
                                 ": normalised coordinate spread >= 2048 (rows must be z-scores, as "
                                 "kraskov_prepare makes them)");
    }
    return GMXCORR_OK;
}

// Blocks of four pairs for the all-pairs run (pair index of (a1, a2), a2 <
// a1, is a1*(a1-1)/2 + a2): 2x2 atom tiles {2t, 2t+1} x {2s, 2s+1}, s < t,
// ordered by t then s, so two waves share each A row and two each B row;
// the pairs outside full tiles follow in runs of four that share a1.
std::vector<int> pair_quads(int n)
{
    auto pid = [](int a1, int a2) { return (int)((long long)a1 * (a1 - 1) / 2 + a2); };
    std::vector<int> q, rest;
    for (int t = 0; 2 * t + 1 < n; ++t) {
        for (int s = 0; s < t; ++s) {
            q.push_back(pid(2 * t, 2 * s));
            q.push_back(pid(2 * t, 2 * s + 1));
            q.push_back(pid(2 * t + 1, 2 * s));
            q.push_back(pid(2 * t + 1, 2 * s + 1));
        }
        rest.push_back(pid(2 * t + 1, 2 * t));
    }
    if (n % 2)
        for (int a2 = 0; a2 < n - 1; ++a2) rest.push_back(pid(n - 1, a2));
    for (size_t r = 0; r < rest.size(); r += 4)
        for (size_t c = 0; c < 4; ++c) q.push_back(r + c < rest.size() ? rest[r + c] : -1);
    return q;
}

}  // namespace gmxcorr

namespace {

int fail(int code, const std::string& msg) { return gmxcorr::set_error(code, msg); }

int hip_fail(hipError_t e, const char* where) { return gmxcorr::set_hip_error(e, where); }

#define HIPCHK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
    } while (0)

bool device_is_gfx950()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

bool gmxcorr::device_ok() { return device_is_gfx950(); }

namespace {

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct gmxcorr_ksg_plan {
    int natoms = 0, N = 0, K = 0;
    long long P = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    double* dX = nullptr;
    double* dPsi = nullptr;
    float* dD = nullptr;
    int* dQuads = nullptr;  // all-pairs run: [nquads][4] pair indices (2x2 atom tiles)
    long long nquads = 0;
    float* dSD = nullptr;   // sorted rows, summaries, prefix indices (N <= kSortedMaxFrames)
    float* dSM = nullptr;
    uint16_t* dSI = nullptr;
    float* dLV = nullptr;   // prefix stores (kSortedMaxFrames < N <= kMaxFrames, k + 1 <= 256)
    uint16_t* dLI = nullptr;
    float* dLS = nullptr;
    int* dA1 = nullptr;
    int* dA2 = nullptr;
    long long pair_cap = 0;
    double* dPartial = nullptr;
    double* dFixsum = nullptr;
    double* dMI = nullptr;
    FixEntry* dFix = nullptr;
    unsigned fixcap = 0;
    FixEntry* dLong = nullptr;   // rows beyond the sorted prefix (kind 2)
    unsigned longcap = 0;
    unsigned* dCounters = nullptr;
    int* dRcNx = nullptr;
    int* dRcNy = nullptr;
    unsigned char* dRcFlag = nullptr;
    long long rc_cap = 0;
    double psiN_phiK = 0.0;
    bool loaded = false;
    bool all_pairs = false;
    long long last_npairs = 0;
    int var_kind = -1, var_s4 = 0;
    gmxcorr::TrajDesc* dTraj = nullptr;   // one descriptor for the recount/finalise kernels
    gmxcorr_ksg_stats st{};
};

namespace {

int ensure_pairs(gmxcorr_ksg_plan* pl, long long npairs)
{
    if (npairs <= pl->pair_cap) return GMXCORR_OK;
    dfree(pl->dA1); dfree(pl->dA2); dfree(pl->dPartial); dfree(pl->dFixsum); dfree(pl->dMI);
    pl->pair_cap = 0;   // a failed allocation below leaves no stale capacity
    HIPCHK(hipMalloc(&pl->dA1, npairs * sizeof(int)));
    HIPCHK(hipMalloc(&pl->dA2, npairs * sizeof(int)));
    HIPCHK(hipMalloc(&pl->dPartial, npairs * gmxcorr::kRowSplit * sizeof(double)));
    HIPCHK(hipMalloc(&pl->dFixsum, npairs * sizeof(double)));
    HIPCHK(hipMalloc(&pl->dMI, npairs * sizeof(double)));
    pl->pair_cap = npairs;
    return GMXCORR_OK;
}

int ensure_fix(gmxcorr_ksg_plan* pl, unsigned cap)
{
    if (cap <= pl->fixcap) return GMXCORR_OK;
    dfree(pl->dFix);
    pl->fixcap = 0;
    HIPCHK(hipMalloc(&pl->dFix, (size_t)cap * sizeof(FixEntry)));
    pl->fixcap = cap;
    return GMXCORR_OK;
}

int ensure_long(gmxcorr_ksg_plan* pl, unsigned cap)
{
    if (cap <= pl->longcap) return GMXCORR_OK;
    dfree(pl->dLong);
    pl->longcap = 0;
    HIPCHK(hipMalloc(&pl->dLong, (size_t)cap * sizeof(FixEntry)));
    pl->longcap = cap;
    return GMXCORR_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}


// Pair kernel + FP64 recount + finalise on the plan stream.  If the recount
// queue overflowed, the run is repeated with a queue large enough.
int run_loaded(gmxcorr_ksg_plan* pl, long long npairs, bool detail, bool tiles)
{
    KsgArgs a{};
    if (tiles) { a.quads = pl->dQuads; a.nquads = pl->nquads; }
    a.X = pl->dX; a.D = pl->dD; a.SD = pl->dSD; a.SM = pl->dSM; a.SI = pl->dSI; a.LV = pl->dLV; a.LI = pl->dLI; a.LS = pl->dLS; a.psi = pl->dPsi; a.a1 = pl->dA1; a.a2 = pl->dA2;
    a.partial = pl->dPartial; a.fixsum = pl->dFixsum; a.mi = pl->dMI;
    a.counters = pl->dCounters;
    a.npairs = npairs; a.P = pl->P; a.N = pl->N; a.K = pl->K; a.psiN_phiK = pl->psiN_phiK;
    if (detail) { a.rc_nx = pl->dRcNx; a.rc_ny = pl->dRcNy; a.rc_flag = pl->dRcFlag; }
    // Queues sized up front, as an overflow re-runs the whole launch: the
    // FP32-tie recount for 1% of the rows (runs measured here queue
    // 0.04-0.2%), the long-row queue for 4% (a prefix path keeps its share
    // of rows beyond the prefix below that; 1.8% at 400 atoms x 5000
    // frames, k = 100, 768 frames), at most 64M and 256M entries (1 and 4
    // GiB).  Both keep their size for later runs.
    {
        const long long rows = npairs * (long long)pl->N;
        const long long want_fix = std::min<long long>(std::max<long long>(65536, rows / 100),
                                                       64LL << 20);
        const long long want_long = std::min<long long>(std::max<long long>(65536, rows / 25),
                                                        256LL << 20);
        int rc = ensure_fix(pl, (unsigned)want_fix);
        if (!rc) {
            rc = ensure_long(pl, (unsigned)want_long);
            if (rc == GMXCORR_ENOMEM && want_long > want_fix) {
                // a plan near the HBM limit: the 1% size, overflow re-runs if needed
                (void)hipGetLastError();
                rc = ensure_long(pl, (unsigned)want_fix);
            }
        }
        if (rc) return rc;
    }
    gmxcorr::FixupArgs f{};
    f.traj = pl->dTraj;
    f.item_traj = nullptr;
    f.a1 = pl->dA1;
    f.a2 = pl->dA2;
    f.fixsum = pl->dFixsum;
    f.counters = pl->dCounters;
    if (detail) { f.rc_nx = pl->dRcNx; f.rc_ny = pl->dRcNy; f.rc_flag = pl->dRcFlag; }
    for (int attempt = 0; attempt < 3; ++attempt) {
        a.fixq = pl->dFix; a.fixcap = pl->fixcap;
        a.longq = pl->dLong; a.longcap = pl->longcap;
        f.fixq = pl->dFix;
        HIPCHK(gmxcorr::launch_ksg_clear(pl->dFixsum, npairs, pl->dCounters, nullptr, pl->stream));
        HIPCHK(hipEventRecord(pl->ev[2], pl->stream));
        int nsplit = gmxcorr::kRowSplit;
        HIPCHK(gmxcorr::launch_ksg_pairs(a, pl->stream, pl->var_kind, pl->var_s4, &nsplit));
        HIPCHK(hipEventRecord(pl->ev[3], pl->stream));
        unsigned counters[3] = {0, 0, 0};
        HIPCHK(hipMemcpyAsync(counters, pl->dCounters, sizeof(counters), hipMemcpyDeviceToHost,
                              pl->stream));
        HIPCHK(hipStreamSynchronize(pl->stream));
        if (counters[2] > pl->longcap) {
            int rc = ensure_long(pl, counters[2] + counters[2] / 4 + 1024);
            if (rc) return rc;
            continue;
        }
        const unsigned nlong = counters[2];
        if (nlong) {
            // the streaming path may queue FP32-ambiguous rows of its own
            HIPCHK(gmxcorr::launch_ksg_long_rows(a, nlong, pl->stream));
            HIPCHK(hipMemcpyAsync(counters, pl->dCounters, sizeof(counters),
                                  hipMemcpyDeviceToHost, pl->stream));
            HIPCHK(hipStreamSynchronize(pl->stream));
        }
        if (counters[0] > pl->fixcap) {
            int rc = ensure_fix(pl, counters[0] + counters[0] / 4 + 1024);
            if (rc) return rc;
            continue;
        }
        HIPCHK(gmxcorr::launch_ksg_fixup(f, counters[0], pl->stream));
        HIPCHK(gmxcorr::launch_ksg_finalize(f, npairs, pl->dPartial, nsplit,
                                                  pl->dMI, pl->stream));
        HIPCHK(hipEventRecord(pl->ev[4], pl->stream));
        HIPCHK(hipMemcpyAsync(counters, pl->dCounters, sizeof(counters), hipMemcpyDeviceToHost,
                              pl->stream));
        HIPCHK(hipStreamSynchronize(pl->stream));
        pl->st.ms_pairs = elapsed(pl->ev[2], pl->ev[3]);
        pl->st.ms_fixup = elapsed(pl->ev[3], pl->ev[4]);
        pl->st.npairs = npairs;
        pl->st.fixup_rows = counters[0];
        pl->st.long_rows = nlong;
        pl->st.ties = counters[1];
        pl->last_npairs = npairs;
        return GMXCORR_OK;
    }
    return fail(GMXCORR_EHIP, "recount queues overflowed on every attempt");
}

}  // namespace

namespace gmxcorr {
bool plan_view(const gmxcorr_ksg_plan* pl, PlanView* v)
{
    if (!pl || !pl->loaded) return false;
    v->X = pl->dX;
    v->D = pl->dD;
    v->natoms = pl->natoms;
    v->N = pl->N;
    v->K = pl->K;
    v->P = pl->P;
    v->stream = pl->stream;
    return true;
}
}  // namespace gmxcorr

extern "C" {

const char* gmxcorr_last_error(void) { return g_err.c_str(); }

int gmxcorr_probe(void) { return device_is_gfx950() ? 1 : 0; }

int gmxcorr_ksg_max_frames(void) { return gmxcorr::kMaxFrames; }

int gmxcorr_selftest_wave(int* out)
{
    if (!out) return fail(GMXCORR_EINVAL, "selftest: null output");
    if (!device_is_gfx950()) return fail(GMXCORR_ENODEV, "no gfx950 device");
    int* d = nullptr;
    HIPCHK(hipMalloc(&d, 67 * sizeof(int)));
    hipError_t e = gmxcorr::launch_selftest(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, d, 67 * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "selftest");
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_create(int natoms, int nframes, int k, gmxcorr_ksg_plan** out)
{
    if (!out) return fail(GMXCORR_EINVAL, "plan_create: null output");
    *out = nullptr;
    if (natoms < 2) return fail(GMXCORR_EINVAL, "need at least 2 atoms");
    if (k < 1 || nframes <= k) return fail(GMXCORR_EINVAL, "need 1 <= k < nframes");
    if (nframes > gmxcorr::kMaxFrames)
        return fail(GMXCORR_EUNSUP, "nframes above the register-resident kernel limit");
    if (!device_is_gfx950()) return fail(GMXCORR_ENODEV, "no gfx950 device");
    auto* pl = new gmxcorr_ksg_plan();
    pl->natoms = natoms; pl->N = nframes; pl->K = k;
    pl->P = gmxcorr::row_pitch(nframes);
    auto bail = [&](int rc) { gmxcorr_ksg_plan_destroy(pl); return rc; };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    for (auto& ev : pl->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
    const size_t xbytes = (size_t)3 * natoms * nframes * sizeof(double);
    const size_t dbytes = (size_t)natoms * nframes * pl->P * sizeof(float);
    if ((e = hipMalloc(&pl->dX, xbytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(X)"));
    if ((e = hipMalloc(&pl->dPsi, (nframes + 1) * sizeof(double))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(psi)"));
    if ((e = hipMalloc(&pl->dD, dbytes)) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(distance rows)"));
    size_t sbytes = 0;
    if (nframes <= gmxcorr::kSortedMaxFrames) {
        const size_t rows = (size_t)natoms * nframes;
        sbytes = rows * ((gmxcorr::sorted_pitch(nframes) + gmxcorr::kSumPitch) * sizeof(float) +
                         gmxcorr::kPrefix * sizeof(uint16_t));
        if ((e = hipMalloc(&pl->dSD, rows * gmxcorr::sorted_pitch(nframes) * sizeof(float))) !=
            hipSuccess)
            return bail(hip_fail(e, "hipMalloc(sorted rows)"));
        if ((e = hipMalloc(&pl->dSM, rows * gmxcorr::kSumPitch * sizeof(float))) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc(sorted-row summaries)"));
        if ((e = hipMalloc(&pl->dSI, rows * gmxcorr::kPrefix * sizeof(uint16_t))) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc(sorted-row prefix indices)"));
    }
    if ((e = hipMalloc(&pl->dCounters, 4 * sizeof(unsigned))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(counters)"));
    if ((e = hipMalloc(&pl->dTraj, sizeof(gmxcorr::TrajDesc))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(descriptor)"));
    int rc = ensure_fix(pl, 65536);
    if (rc) return bail(rc);
    // the recount and long-row queues of an all-pairs run (1% of its rows
    // each, as run_loaded sizes them), before the prefix stores take HBM
    {
        const long double run_rows = (long double)natoms * (natoms - 1) / 2 * nframes;
        const long long want = (long long)std::min<long double>(
            std::max<long double>(65536, run_rows / 100), (long double)(64LL << 20));
        int rc2 = ensure_fix(pl, (unsigned)want);
        if (!rc2 && gmxcorr::variant_fits(3, 4, nframes, k)) rc2 = ensure_long(pl, (unsigned)want);
        if (rc2) return bail(rc2);
    }
    // prefix stores for the long-trajectory path when they fit with 4 GB to
    // spare; otherwise the plan runs the cooperative path
    if (gmxcorr::variant_fits(3, 4, nframes, k)) {
        const size_t rows = (size_t)natoms * nframes;
        const size_t lbytes = rows * (gmxcorr::kLongPitch * sizeof(float) +
                                      gmxcorr::kLongPrefix * sizeof(uint16_t) +
                                      gmxcorr::kLongSum * sizeof(float));
        const size_t reserve = 4ull << 30;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > lbytes + reserve) {
            if (hipMalloc(&pl->dLV, rows * gmxcorr::kLongPitch * sizeof(float)) != hipSuccess ||
                hipMalloc(&pl->dLI, rows * gmxcorr::kLongPrefix * sizeof(uint16_t)) != hipSuccess ||
                hipMalloc(&pl->dLS, rows * gmxcorr::kLongSum * sizeof(float)) != hipSuccess) {
                (void)hipGetLastError();
                dfree(pl->dLV); dfree(pl->dLI); dfree(pl->dLS);
            } else {
                sbytes += lbytes;
            }
        }
    }
    pl->st.row_bytes = (long long)(dbytes + sbytes);
    pl->st.nframes = nframes; pl->st.k = k; pl->st.natoms = natoms; pl->st.pitch = (int)pl->P;
    *out = pl;
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_load(gmxcorr_ksg_plan* pl, const double* xrows, const double* psi)
{
    if (!pl || !xrows || !psi) return fail(GMXCORR_EINVAL, "plan_load: null argument");
    const size_t nx = (size_t)3 * pl->natoms * pl->N;
    if (int rc = gmxcorr::check_rows(xrows, 3LL * pl->natoms, pl->N, "plan_load")) return rc;
    HIPCHK(hipEventRecord(pl->ev[0], pl->stream));
    HIPCHK(hipMemcpyAsync(pl->dX, xrows, nx * sizeof(double), hipMemcpyHostToDevice, pl->stream));
    HIPCHK(hipMemcpyAsync(pl->dPsi, psi, (pl->N + 1) * sizeof(double), hipMemcpyHostToDevice,
                          pl->stream));
    HIPCHK(hipEventRecord(pl->ev[1], pl->stream));
    HIPCHK(gmxcorr::launch_build_rows(pl->dX, pl->natoms, pl->N, pl->P, pl->dD, pl->dSD, pl->dSM,
                                      pl->dSI, pl->stream));
    if (pl->dLV)
        HIPCHK(gmxcorr::launch_build_prefix(pl->dD, pl->natoms, pl->N, pl->P, pl->dLV, pl->dLI,
                                            pl->dLS, pl->stream));
    HIPCHK(hipEventRecord(pl->ev[5], pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    pl->st.ms_load = elapsed(pl->ev[0], pl->ev[1]);
    pl->st.ms_build = elapsed(pl->ev[1], pl->ev[5]);
    pl->psiN_phiK = psi[pl->N] + (psi[pl->K] - 1 / (1.0 * pl->K));
    gmxcorr::TrajDesc td{};
    td.X = pl->dX;
    td.D = pl->dD;
    td.SD = pl->dSD;
    td.SM = pl->dSM;
    td.SI = pl->dSI;
    td.LV = pl->dLV;
    td.LI = pl->dLI;
    td.LS = pl->dLS;
    td.psi = pl->dPsi;
    td.P = pl->P;
    td.N = pl->N;
    td.K = pl->K;
    td.psiN_phiK = pl->psiN_phiK;
    td.kind = -1;
    td.s4 = 0;
    HIPCHK(hipMemcpy(pl->dTraj, &td, sizeof(td), hipMemcpyHostToDevice));
    pl->loaded = true;
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_run_all(gmxcorr_ksg_plan* pl)
{
    if (!pl || !pl->loaded) return fail(GMXCORR_EINVAL, "plan not loaded");
    const long long n = pl->natoms;
    const long long npairs = n * (n - 1) / 2;
    if (!pl->all_pairs || pl->pair_cap < npairs) {
        int rc = ensure_pairs(pl, npairs);
        if (rc) return rc;
        std::vector<int> h1(npairs), h2(npairs);
        long long idx = 0;
        for (int a1 = 1; a1 < n; ++a1)
            for (int a2 = 0; a2 < a1; ++a2, ++idx) { h1[idx] = a1; h2[idx] = a2; }
        HIPCHK(hipMemcpy(pl->dA1, h1.data(), npairs * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(pl->dA2, h2.data(), npairs * sizeof(int), hipMemcpyHostToDevice));
        if (!pl->dQuads && npairs < (1LL << 31)) {
            const std::vector<int> q = gmxcorr::pair_quads((int)n);
            HIPCHK(hipMalloc(&pl->dQuads, q.size() * sizeof(int)));
            HIPCHK(hipMemcpy(pl->dQuads, q.data(), q.size() * sizeof(int), hipMemcpyHostToDevice));
            pl->nquads = (long long)q.size() / 4;
        }
        pl->all_pairs = true;
    }
    return run_loaded(pl, npairs, false, pl->dQuads != nullptr);
}

int gmxcorr_ksg_plan_run_pairs(gmxcorr_ksg_plan* pl, long long npairs, const int* a1,
                               const int* a2, int* nx, int* ny, unsigned char* flags)
{
    if (!pl || !pl->loaded) return fail(GMXCORR_EINVAL, "plan not loaded");
    if (npairs <= 0 || !a1 || !a2) return fail(GMXCORR_EINVAL, "empty pair list");
    for (long long p = 0; p < npairs; ++p)
        if (a1[p] < 0 || a2[p] < 0 || a1[p] >= pl->natoms || a2[p] >= pl->natoms || a1[p] == a2[p])
            return fail(GMXCORR_EINVAL, "pair index out of range or a1 == a2");
    const bool detail = nx || ny || flags;
    int rc = ensure_pairs(pl, npairs);
    if (rc) return rc;
    pl->all_pairs = false;
    HIPCHK(hipMemcpy(pl->dA1, a1, npairs * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pl->dA2, a2, npairs * sizeof(int), hipMemcpyHostToDevice));
    const long long rows = npairs * pl->N;
    if (detail && rows > pl->rc_cap) {
        dfree(pl->dRcNx); dfree(pl->dRcNy); dfree(pl->dRcFlag);
        pl->rc_cap = 0;   // a failed allocation below leaves no stale capacity
        HIPCHK(hipMalloc(&pl->dRcNx, rows * sizeof(int)));
        HIPCHK(hipMalloc(&pl->dRcNy, rows * sizeof(int)));
        HIPCHK(hipMalloc(&pl->dRcFlag, rows));
        pl->rc_cap = rows;
    }
    rc = run_loaded(pl, npairs, detail, false);
    if (rc) return rc;
    if (nx) HIPCHK(hipMemcpy(nx, pl->dRcNx, rows * sizeof(int), hipMemcpyDeviceToHost));
    if (ny) HIPCHK(hipMemcpy(ny, pl->dRcNy, rows * sizeof(int), hipMemcpyDeviceToHost));
    if (flags) HIPCHK(hipMemcpy(flags, pl->dRcFlag, rows, hipMemcpyDeviceToHost));
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_result(gmxcorr_ksg_plan* pl, double* mi, long long npairs)
{
    if (!pl || !mi) return fail(GMXCORR_EINVAL, "plan_result: null argument");
    if (npairs != pl->last_npairs) return fail(GMXCORR_EINVAL, "pair count differs from last run");
    HIPCHK(hipMemcpy(mi, pl->dMI, npairs * sizeof(double), hipMemcpyDeviceToHost));
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_matrix(gmxcorr_ksg_plan* pl, double* mat)
{
    if (!pl || !mat) return fail(GMXCORR_EINVAL, "plan_matrix: null argument");
    const int n = pl->natoms;
    const long long npairs = (long long)n * (n - 1) / 2;
    if (!pl->all_pairs || pl->last_npairs != npairs)
        return fail(GMXCORR_EINVAL, "last run was not run_all");
    std::vector<double> mi(npairs);
    HIPCHK(hipMemcpy(mi.data(), pl->dMI, npairs * sizeof(double), hipMemcpyDeviceToHost));
    // diagonal 2000 and symmetric fill, as lib/kraskov_gpu.cu:294-303
    for (int i = 0; i < n; ++i) mat[(size_t)i * n + i] = 2000.0;
    long long idx = 0;
    for (int a1 = 1; a1 < n; ++a1)
        for (int a2 = 0; a2 < a1; ++a2, ++idx)
            mat[(size_t)a1 * n + a2] = mat[(size_t)a2 * n + a1] = mi[idx];
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_set_variant(gmxcorr_ksg_plan* pl, int kind, int s4)
{
    if (!pl) return fail(GMXCORR_EINVAL, "set_variant: null plan");
    if (kind >= 0 && !gmxcorr::variant_fits(kind, s4, pl->N, pl->K))
        return fail(GMXCORR_EUNSUP, "kernel variant does not cover this frame count and k");
    pl->var_kind = kind;
    pl->var_s4 = s4;
    return GMXCORR_OK;
}

int gmxcorr_ksg_variant_fits(int kind, int s4, int nframes, int k)
{
    return gmxcorr::variant_fits(kind, s4, nframes, k) ? 1 : 0;
}

int gmxcorr_ksg_plan_stats(const gmxcorr_ksg_plan* pl, gmxcorr_ksg_stats* st)
{
    if (!pl || !st) return fail(GMXCORR_EINVAL, "plan_stats: null argument");
    *st = pl->st;
    return GMXCORR_OK;
}

void gmxcorr_ksg_plan_destroy(gmxcorr_ksg_plan* pl)
{
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    dfree(pl->dX); dfree(pl->dPsi); dfree(pl->dD); dfree(pl->dSD); dfree(pl->dSM); dfree(pl->dSI); dfree(pl->dLV); dfree(pl->dLI); dfree(pl->dLS); dfree(pl->dQuads); dfree(pl->dA1); dfree(pl->dA2);
    dfree(pl->dPartial); dfree(pl->dFixsum); dfree(pl->dMI); dfree(pl->dFix); dfree(pl->dLong);
    dfree(pl->dCounters); dfree(pl->dRcNx); dfree(pl->dRcNy); dfree(pl->dRcFlag);
    dfree(pl->dTraj);
    for (auto& ev : pl->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

int gmxcorr_ksg_matrix(const double* xrows, int natoms, int nframes, const double* psi, int k,
                       double* mat)
{
    if (!xrows || !psi || !mat) return fail(GMXCORR_EINVAL, "ksg_matrix: null argument");
    gmxcorr_ksg_plan* pl = nullptr;
    int rc = gmxcorr_ksg_plan_create(natoms, nframes, k, &pl);
    if (rc) return rc;
    rc = gmxcorr_ksg_plan_load(pl, xrows, psi);
    if (!rc) rc = gmxcorr_ksg_plan_run_all(pl);
    if (!rc) rc = gmxcorr_ksg_plan_matrix(pl, mat);
    gmxcorr_ksg_plan_destroy(pl);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The reference's C++ boundary (lib/kraskov_gpu.h:20,32).
// ---------------------------------------------------------------------------
bool kraskov_gpu_probe() { return device_is_gfx950(); }

void kraskov_corrmatrix_gpu(const t_kraskov* kr, int natoms, double* mat, int k)
{
    const int N = kr->N;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess)
            std::fprintf(stderr, "GPU: %s  (%s, %.0f MB)\n", prop.name, prop.gcnArchName,
                         prop.totalGlobalMem / 1e6);
    }
    // flatten the 3*natoms separately allocated rows (lib/kraskov_gpu.cu:240-246)
    std::vector<double> rows((size_t)3 * natoms * N);
    for (int c = 0; c < 3 * natoms; ++c)
        std::memcpy(rows.data() + (size_t)c * N, kr->x[c], N * sizeof(double));
    std::fprintf(stderr, "GPU Kraskov (HIP gfx950): %lld pairs, N=%d, k=%d\n",
                 (long long)natoms * (natoms - 1) / 2, N, k);
    const int rc = gmxcorr_ksg_matrix(rows.data(), natoms, N, kr->psi, k, mat);
    if (rc != GMXCORR_OK) {
        std::fprintf(stderr, "HIP error in kraskov_corrmatrix_gpu: %s\n", gmxcorr_last_error());
        std::exit(1);
    }
}
