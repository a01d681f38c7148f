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

// Pass-1 kernel of later calls (gmxcorr_te_set_path): 0 automatic, 1 heap,
// 2 select, 3 prefix stores (with the heap kernel for the queries they
// leave).
int g_path = 0;
// The last call: queries, queries resolved by the prefix select, queries
// left to the heap kernel, whether the prefix path ran.
long long g_stats[4] = {0, 0, 0, 0};

// The prefix path over pairs a.a1 / a.a2 (target-major; h1 is the host copy
// of a.a1): per launch chunk of at most te_chunk_pairs pairs and maxT
// distinct target atoms, the targets' prefix stores, the select (exact FP64
// radii), the heap kernel for the queries it lists (their roots and exact
// radii), and the count pass in FP64 over every query.
int run_prefix(const gmxcorr::TeArgs& a, const int* h1, hipStream_t s)
{
    using namespace gmxcorr;
    const int Nv = a.N - a.lag;
    const int nqb = a.nqblocks;
    const long long npairs = a.npairs;
    const long long chunk = te_chunk_pairs(nqb);
    const size_t per_target = (size_t)Nv * kTePrefix * (sizeof(uint32_t) + sizeof(uint16_t));
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const size_t budget = std::min(free_b / 2, (size_t)32 << 30);
    // chunks of at most `chunk` pairs and `maxT` distinct targets (pairs
    // are target-major); the stores are sized for the largest chunk's targets
    long long maxT = (long long)(budget / per_target);
    if (maxT < 1) return gmxcorr::set_error(GMXCORR_ENOMEM, "te: no room for one target's prefix stores");
    auto chunk_end = [&](long long p0, long long& ntargets) {
        long long p1 = p0;
        ntargets = 0;
        while (p1 < npairs && p1 - p0 < chunk) {
            if (p1 == p0 || h1[p1] != h1[p1 - 1]) {
                if (ntargets == maxT) break;
                ++ntargets;
            }
            ++p1;
        }
        return p1;
    };
    long long most = 0;
    for (long long p0 = 0, nt = 0; p0 < npairs; most = std::max(most, nt)) p0 = chunk_end(p0, nt);
    maxT = most;
    const long long npc = std::min(npairs, chunk);
    DevBuf<uint32_t> tv;
    DevBuf<uint16_t> ti, qlist;
    DevBuf<int> dTargets, dSlot;
    DevBuf<unsigned> dCount;
    DevBuf<int2> dBlocks;
    DevBuf<double> dEps;
    HIPCHK(dEps.alloc((size_t)npc * nqb * 64));
    HIPCHK(tv.alloc((size_t)maxT * Nv * kTePrefix));
    HIPCHK(ti.alloc((size_t)maxT * Nv * kTePrefix));
    HIPCHK(qlist.alloc((size_t)npc * Nv));
    HIPCHK(dTargets.alloc((size_t)maxT));
    HIPCHK(dSlot.alloc((size_t)npc));
    HIPCHK(dCount.alloc((size_t)npc));
    HIPCHK(dBlocks.alloc((size_t)npc * nqb));
    std::vector<int> targets, tslot;
    std::vector<unsigned> cnt;
    std::vector<int2> blocks;
    for (long long p0 = 0; p0 < npairs;) {
        long long nt = 0;
        const long long p1 = chunk_end(p0, nt);
        targets.clear();
        tslot.clear();
        for (long long p = p0; p < p1; ++p) {
            if (p == p0 || h1[p] != h1[p - 1]) targets.push_back(h1[p]);
            tslot.push_back((int)targets.size() - 1);
        }
        const long long np = p1 - p0;
        TeArgs c = a;
        c.a1 = a.a1 + p0;
        c.a2 = a.a2 + p0;
        c.partial = a.partial + p0 * nqb;
        c.npairs = np;
        c.eps64 = dEps.p;
        const TePrefixArgs pa{dTargets.p, (long long)targets.size(), dSlot.p, tv.p, ti.p, dCount.p,
                              qlist.p};
        HIPCHK(hipMemcpyAsync(dTargets.p, targets.data(), targets.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dSlot.p, tslot.data(), tslot.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(dCount.p, 0, (size_t)np * sizeof(unsigned), s));
        HIPCHK(launch_te_prefix_build(c, pa, s));
        HIPCHK(launch_te_prefix_select(c, pa, s));
        cnt.resize((size_t)np);
        HIPCHK(hipMemcpyAsync(cnt.data(), dCount.p, (size_t)np * sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        blocks.clear();
        long long left = 0;
        for (long long p = 0; p < np; ++p) {
            if (cnt[(size_t)p] > (unsigned)Nv)
                return gmxcorr::set_error(GMXCORR_EHIP, "te: prefix select listed more queries than a pair has");
            left += cnt[(size_t)p];
            for (unsigned f = 0; f < cnt[(size_t)p]; f += 64) blocks.push_back(make_int2((int)p, (int)f));
        }
        if (!blocks.empty()) {
            HIPCHK(hipMemcpyAsync(dBlocks.p, blocks.data(), blocks.size() * sizeof(int2),
                                  hipMemcpyHostToDevice, s));
            TeArgs l = c;
            l.qblocks = dBlocks.p;
            l.qcount = dCount.p;
            l.qlist = qlist.p;
            // the listed queries' roots and exact radii
            HIPCHK(launch_te_heap_list(l, (long long)blocks.size(), s));
        }
        // the counts from the exact FP64 radii
        HIPCHK(launch_te_count64(c, s));
        // the host vectors are refilled for the next chunk
        HIPCHK(hipStreamSynchronize(s));
        g_stats[0] += np * Nv;
        g_stats[1] += np * Nv - left;
        g_stats[2] += left;
        p0 = p1;
    }
    g_stats[3] = 1;
    return GMXCORR_OK;
}

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
    if ((g_path == 1 || g_path == 3) && k > gmxcorr::kTeHeapMaxK)
        return gmxcorr::set_error(GMXCORR_EUNSUP, "te: the heap and prefix paths take k <= 128");
    if (g_path == 3 && Nv > gmxcorr::kTePrefixMaxFrames)
        return gmxcorr::set_error(GMXCORR_EUNSUP, "te: the prefix path takes nframes - lag <= 16384");
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
    // pass 1 through the prefix stores: forced, or automatic for series
    // longer than kTePrefixAutoFrames whose pairs share their targets (the
    // stores of a target serve all its pairs)
    bool prefix = g_path == 3;
    if (g_path == 0 && k <= gmxcorr::kTeHeapMaxK && Nv <= gmxcorr::kTePrefixMaxFrames &&
        Nv > gmxcorr::kTePrefixAutoFrames) {
        long long distinct = 0;
        for (long long p = 0; p < npairs; ++p) distinct += p == 0 || a1[p] != a1[p - 1];
        prefix = npairs >= (long long)gmxcorr::kTePrefixMinPairs * distinct;
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
    // the stream and events are released on every return
    struct Handles {
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Handles()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        }
    } h;
    HIPCHK(hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&h.e0));
    HIPCHK(hipEventCreate(&h.e1));
    const hipStream_t s = h.s;
    const hipEvent_t e0 = h.e0, e1 = h.e1;
    HIPCHK(hipMemcpyAsync(dTraj.p, traj, (size_t)natoms * nframes * 3 * sizeof(double),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPsi.p, psi.data(), psi.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d1.p, a1, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d2.p, a2, npairs * sizeof(int), hipMemcpyHostToDevice, s));
    gmxcorr::TeArgs ta{dTraj.p, nframes, lag, k, g_path, d1.p, d2.p, npairs, nqb, dPsi.p, dPart.p, dRoot.p};
    for (auto& v : g_stats) v = 0;
    HIPCHK(hipEventRecord(e0, s));
    if (prefix) {
        if (int rc = run_prefix(ta, a1, s)) return rc;
    } else {
        HIPCHK(gmxcorr::launch_te(ta, s));
        g_stats[0] = npairs * Nv;
    }
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
    if (path < 0 || path > 3)
        return gmxcorr::set_error(GMXCORR_EINVAL, "te_set_path: path must be 0, 1, 2 or 3");
    g_path = path;
    return GMXCORR_OK;
}

int gmxcorr_te_stats(long long* out)
{
    if (!out) return gmxcorr::set_error(GMXCORR_EINVAL, "te_stats: null output");
    for (int i = 0; i < 4; ++i) out[i] = g_stats[i];
    return GMXCORR_OK;
}

}  // extern "C"
