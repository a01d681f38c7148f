// Host side of the batched KSG run: many trajectories of uneven atom and
// frame counts, all their atom pairs processed by one persistent kernel
// launch per memory wave (distance rows of all trajectories of a wave
// resident in HBM together).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "gmxcorr.h"
#include "ksg_internal.h"

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

struct HostTraj {
    int natoms = 0, N = 0, K = 0, kind = -1, s4 = 0;
    int chunk = gmxcorr::kLongChunk;   // frame rows per kind-3 work item
    long long P = 0;
    std::vector<double> rows;   // [3n][N]
    std::vector<double> psi;    // [N+1]
    std::vector<double> mi;     // [npairs] after run
    double psiN_phiK = 0.0;
    int wave = -1;
    double ms_done = 0.0;       // batch-relative completion time of its wave
    bool sorted() const { return gmxcorr::variant_sorted(kind); }
    long long d_bytes() const { return (long long)natoms * N * P * 4; }
    long long sd_bytes() const { return sorted() ? (long long)natoms * N * gmxcorr::sorted_pitch(N) * 4 : 0; }
    long long sm_bytes() const { return sorted() ? (long long)natoms * N * gmxcorr::kSumPitch * 4 : 0; }
    long long si_bytes() const { return sorted() ? (long long)natoms * N * gmxcorr::kPrefix * 2 : 0; }
    bool stored() const { return gmxcorr::variant_prefix_store(kind); }
    long long lv_bytes() const { return stored() ? (long long)natoms * N * gmxcorr::kLongPitch * 4 : 0; }
    long long li_bytes() const { return stored() ? (long long)natoms * N * gmxcorr::kLongPrefix * 2 : 0; }
    long long ls_bytes() const { return stored() ? (long long)natoms * N * gmxcorr::kLongSum * 4 : 0; }
    // device bytes of the distance rows, sorted rows, summaries and prefix
    // indices (256-aligned parts)
    long long row_bytes() const
    {
        auto al = [](long long v) { return (v + 255) / 256 * 256; };
        return al(d_bytes()) + al(sd_bytes()) + al(sm_bytes()) + al(si_bytes()) + al(lv_bytes()) +
               al(li_bytes()) + al(ls_bytes());
    }
    long long npairs() const { return (long long)natoms * (natoms - 1) / 2; }
};

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct gmxcorr_ksg_batch {
    std::vector<HostTraj> traj;
    gmxcorr_batch_stats st{};
};

extern "C" {

int gmxcorr_ksg_batch_create(gmxcorr_ksg_batch** out)
{
    if (!out) return gmxcorr::set_error(GMXCORR_EINVAL, "batch_create: null output");
    *out = new gmxcorr_ksg_batch();
    return GMXCORR_OK;
}

void gmxcorr_ksg_batch_destroy(gmxcorr_ksg_batch* b) { delete b; }

int gmxcorr_ksg_batch_add(gmxcorr_ksg_batch* b, const double* xrows, const double* psi,
                          int natoms, int nframes, int k, int kind, int s4)
{
    if (!b || !xrows || !psi) return gmxcorr::set_error(GMXCORR_EINVAL, "batch_add: null argument");
    if (natoms < 2 || k < 1 || nframes <= k)
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_add: need natoms >= 2 and 1 <= k < nframes");
    if (nframes > gmxcorr::kMaxFrames)
        return gmxcorr::set_error(GMXCORR_EUNSUP, "batch_add: nframes above the kernel limit");
    if (kind < 0) {
        // the sorted-prefix path where it applies, else the smallest compiled
        // batch path that covers the frame count
        static const int single[] = {2, 4, 6, 8};
        static const int coop[] = {3, 4, 6, 8};
        kind = -1;
        if (gmxcorr::batch_variant_ok(2, 1, nframes, k)) { kind = 2; s4 = 1; }
        if (kind < 0 && gmxcorr::batch_variant_ok(3, 2, nframes, k)) {
            kind = 3;
            s4 = gmxcorr::long_prefix_chunks(nframes, k);
        }
        if (kind < 0)
            for (int v : single)
                if (nframes <= 256 * v) { kind = 0; s4 = v; break; }
        if (kind < 0)
            for (int v : coop)
                if (nframes <= 1024 * v) { kind = 1; s4 = v; break; }
    }
    if (!gmxcorr::batch_variant_ok(kind, s4, nframes, k))
        return gmxcorr::set_error(GMXCORR_EUNSUP, "batch_add: variant not compiled into the batch kernel");
    const size_t nx = (size_t)3 * natoms * nframes;
    if (int rc = gmxcorr::check_rows(xrows, 3LL * natoms, nframes, "batch_add")) return rc;
    HostTraj h;
    h.natoms = natoms;
    h.N = nframes;
    h.K = k;
    h.kind = kind;
    h.s4 = s4;
    h.P = gmxcorr::row_pitch(nframes);
    h.rows.assign(xrows, xrows + nx);
    h.psi.assign(psi, psi + nframes + 1);
    h.psiN_phiK = psi[nframes] + (psi[k] - 1 / (1.0 * k));
    b->traj.push_back(std::move(h));
    return (int)b->traj.size() - 1;
}

int gmxcorr_ksg_batch_set_chunk(gmxcorr_ksg_batch* b, int t, int chunk)
{
    if (!b || t < 0 || t >= (int)b->traj.size())
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_set_chunk: no such trajectory");
    if (chunk < 0 || chunk > 1024)
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_set_chunk: chunk must be in [0, 1024]");
    b->traj[t].chunk = chunk > 0 ? chunk : gmxcorr::kLongChunk;
    return GMXCORR_OK;
}

// Run every pair of every trajectory.  Trajectories are packed greedily (in
// order of decreasing row bytes) into memory waves of at most
// max_row_bytes of distance rows; each wave is one row build per trajectory
// and ONE persistent pair launch of nblocks blocks over all its pairs,
// largest frame counts first.
int gmxcorr_ksg_batch_run(gmxcorr_ksg_batch* b, long long max_row_bytes, int nblocks)
{
    if (!b || b->traj.empty()) return gmxcorr::set_error(GMXCORR_EINVAL, "batch_run: empty batch");
    if (!gmxcorr::device_ok()) return gmxcorr::set_error(GMXCORR_ENODEV, "no gfx950 device");
    if (nblocks < 1) nblocks = 2048;
    const int nt = (int)b->traj.size();
    std::vector<int> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return b->traj[x].row_bytes() > b->traj[y].row_bytes(); });
    std::vector<std::vector<int>> waves;
    {
        std::vector<long long> used;
        for (int t : order) {
            const long long need = b->traj[t].row_bytes();
            if (need > max_row_bytes)
                return gmxcorr::set_error(GMXCORR_ENOMEM, "batch_run: one trajectory exceeds the row budget");
            bool placed = false;
            for (size_t w = 0; w < waves.size() && !placed; ++w)
                if (used[w] + need <= max_row_bytes) {
                    waves[w].push_back(t);
                    used[w] += need;
                    placed = true;
                }
            if (!placed) {
                waves.push_back({t});
                used.push_back(need);
            }
        }
    }
    // one arena for the distance rows (and one for rows + psi), reused by
    // every wave: allocating and freeing ~10^11 bytes per wave costs seconds
    auto align = [](long long v) { return (v + 255) / 256 * 256; };
    long long dbytes = 0, xbytes = 0;
    for (const auto& ts : waves) {
        long long d = 0, x = 0;
        for (int t : ts) {
            d += align(b->traj[t].row_bytes());
            x += align((long long)b->traj[t].rows.size() * 8) + align((long long)b->traj[t].psi.size() * 8);
        }
        dbytes = std::max(dbytes, d);
        xbytes = std::max(xbytes, x);
    }
    char* dArena = nullptr;
    char* xArena = nullptr;
    const auto ta0 = std::chrono::steady_clock::now();
    HIPCHK(hipMalloc(&dArena, (size_t)dbytes));
    if (hipMalloc(&xArena, (size_t)xbytes) != hipSuccess) {
        (void)hipFree(dArena);
        return gmxcorr::set_error(GMXCORR_ENOMEM, "batch_run: cannot allocate the coordinate arena");
    }
    const double ms_alloc =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta0).count();
    std::fprintf(stderr, "gmxcorr batch: %zu waves, row arena %.1f GB allocated in %.1f ms\n",
                 waves.size(), dbytes / 1e9, ms_alloc);
    std::fflush(stderr);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t ev[6];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(ev[0], s));
    double ms_total = 0.0, ms_pairs_total = 0.0, ms_build_total = 0.0, ms_fix_total = 0.0;
    long long fix_rows = 0, ties = 0, long_rows = 0, items_total = 0;
    int rc = GMXCORR_OK;
    for (size_t wv = 0; wv < waves.size() && rc == GMXCORR_OK; ++wv) {
        // trajectories by decreasing frame count: the work order
        std::vector<int> ts = waves[wv];
        std::stable_sort(ts.begin(), ts.end(),
                         [&](int x, int y) { return b->traj[x].N > b->traj[y].N; });
        std::vector<double*> dX(ts.size(), nullptr), dPsi(ts.size(), nullptr);
        std::vector<float*> dD(ts.size(), nullptr);
        gmxcorr::TrajDesc* dDesc = nullptr;
        int *dItemTraj = nullptr, *dA1 = nullptr, *dA2 = nullptr, *dQuads = nullptr;
        double *dPartial = nullptr, *dMI = nullptr;
        gmxcorr::FixSum* dFixsum = nullptr;
        gmxcorr::FixEntry *dFix = nullptr, *dLong = nullptr;
        unsigned* dCnt = nullptr;
        unsigned long long* dQueue = nullptr;
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(s);
            dfree(dDesc); dfree(dItemTraj); dfree(dA1); dfree(dA2); dfree(dPartial);
            dfree(dFixsum); dfree(dMI); dfree(dFix); dfree(dLong); dfree(dCnt); dfree(dQueue);
            dfree(dQuads);
        };
        auto body = [&]() -> int {
            std::fprintf(stderr, "gmxcorr batch: wave %zu/%zu: building rows of %zu trajectories\n",
                         wv + 1, waves.size(), ts.size());
            std::fflush(stderr);
            std::vector<gmxcorr::TrajDesc> desc(ts.size());
            int kinds = 0;   // code paths the wave's trajectories take
            for (size_t q = 0; q < ts.size(); ++q) kinds |= 1 << b->traj[ts[q]].kind;
            HIPCHK(hipEventRecord(ev[1], s));
            long long doff = 0, xoff = 0;
            for (size_t q = 0; q < ts.size(); ++q) {
                HostTraj& h = b->traj[ts[q]];
                dD[q] = reinterpret_cast<float*>(dArena + doff);
                float* dSD = h.sorted() ? reinterpret_cast<float*>(dArena + doff + align(h.d_bytes()))
                                        : nullptr;
                float* dSM = h.sorted() ? reinterpret_cast<float*>(dArena + doff + align(h.d_bytes()) +
                                                                  align(h.sd_bytes()))
                                        : nullptr;
                uint16_t* dSI =
                    h.sorted() ? reinterpret_cast<uint16_t*>(dArena + doff + align(h.d_bytes()) +
                                                             align(h.sd_bytes()) + align(h.sm_bytes()))
                               : nullptr;
                const long long loff = doff + align(h.d_bytes()) + align(h.sd_bytes()) +
                                       align(h.sm_bytes()) + align(h.si_bytes());
                float* dLV = h.stored() ? reinterpret_cast<float*>(dArena + loff) : nullptr;
                uint16_t* dLI = h.stored()
                                    ? reinterpret_cast<uint16_t*>(dArena + loff + align(h.lv_bytes()))
                                    : nullptr;
                float* dLS = h.stored() ? reinterpret_cast<float*>(dArena + loff + align(h.lv_bytes()) +
                                                                   align(h.li_bytes()))
                                        : nullptr;
                doff += align(h.row_bytes());
                dX[q] = reinterpret_cast<double*>(xArena + xoff);
                xoff += align((long long)h.rows.size() * 8);
                dPsi[q] = reinterpret_cast<double*>(xArena + xoff);
                xoff += align((long long)h.psi.size() * 8);
                HIPCHK(hipMemcpyAsync(dX[q], h.rows.data(), h.rows.size() * sizeof(double),
                                      hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(dPsi[q], h.psi.data(), h.psi.size() * sizeof(double),
                                      hipMemcpyHostToDevice, s));
                HIPCHK(gmxcorr::launch_build_rows(dX[q], h.natoms, h.N, h.P, dD[q], dSD, dSM, dSI,
                                                  s));
                if (h.stored())
                    HIPCHK(gmxcorr::launch_build_prefix(dD[q], h.natoms, h.N, h.P, dLV, dLI, dLS, s));
                desc[q] = gmxcorr::TrajDesc{dX[q], dD[q], dSD, dSM, dSI, dLV, dLI, dLS, dPsi[q],
                                            h.P, h.N, h.K, h.psiN_phiK, h.kind, h.s4, h.chunk};
            }
            // pairs a1 > a2 of each trajectory, in the work order
            std::vector<int> tq(ts.size());
            std::iota(tq.begin(), tq.end(), 0);
            std::vector<int> it, i1, i2;
            std::vector<long long> first(ts.size());
            for (int q : tq) {
                const HostTraj& h = b->traj[ts[q]];
                first[q] = (long long)it.size();
                for (int a1 = 1; a1 < h.natoms; ++a1)
                    for (int a2 = 0; a2 < a1; ++a2) {
                        it.push_back(q);
                        i1.push_back(a1);
                        i2.push_back(a2);
                    }
            }
            const long long nitems = (long long)it.size();   // pairs of the wave
            items_total += nitems;
            // work items of each trajectory, and the tile tables of the kind-3 ones
            std::vector<int> hq;
            std::vector<long long> qoff(ts.size(), -1);
            for (size_t q = 0; q < ts.size(); ++q) {
                const HostTraj& h = b->traj[ts[q]];
                desc[q].pair0 = first[q];
                desc[q].quads = nullptr;
                if (h.kind == 3) {
                    const std::vector<int> t4 = gmxcorr::pair_quads(h.natoms);
                    qoff[q] = (long long)hq.size();
                    hq.insert(hq.end(), t4.begin(), t4.end());
                    desc[q].nquads = (long long)t4.size() / 4;
                    desc[q].nitems = desc[q].nquads * ((h.N + h.chunk - 1) / h.chunk);
                } else {
                    desc[q].nquads = 0;
                    desc[q].nitems = h.npairs();
                }
            }
            if (!hq.empty()) {
                HIPCHK(hipMalloc(&dQuads, hq.size() * sizeof(int)));
                HIPCHK(hipMemcpyAsync(dQuads, hq.data(), hq.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
                for (size_t q = 0; q < ts.size(); ++q)
                    if (qoff[q] >= 0) desc[q].quads = dQuads + qoff[q];
            }
            HIPCHK(hipMalloc(&dDesc, desc.size() * sizeof(gmxcorr::TrajDesc)));
            HIPCHK(hipMalloc(&dItemTraj, nitems * sizeof(int)));
            HIPCHK(hipMalloc(&dA1, nitems * sizeof(int)));
            HIPCHK(hipMalloc(&dA2, nitems * sizeof(int)));
            HIPCHK(hipMalloc(&dPartial, nitems * sizeof(double)));
            HIPCHK(hipMalloc(&dFixsum, nitems * sizeof(gmxcorr::FixSum)));
            HIPCHK(hipMalloc(&dMI, nitems * sizeof(double)));
            HIPCHK(hipMalloc(&dCnt, 4 * sizeof(unsigned)));
            HIPCHK(hipMalloc(&dQueue, 8 * sizeof(unsigned long long)));
            // recount queue for 1% of the wave's rows (0.04-0.2% measured),
            // long-row queue for 6% (the tuner keeps the long-row share of a
            // trajectory below 4%); an overflow re-runs the wave
            long long rows_w = 0;
            for (size_t q = 0; q < ts.size(); ++q)
                rows_w += b->traj[ts[q]].npairs() * (long long)b->traj[ts[q]].N;
            unsigned fixcap = (unsigned)std::min<long long>(std::max<long long>(1 << 18, rows_w / 100),
                                                            64LL << 20);
            unsigned longcap = (unsigned)std::min<long long>(
                std::max<long long>(1 << 18, rows_w / 16), 256LL << 20);
            HIPCHK(hipMalloc(&dFix, (size_t)fixcap * sizeof(gmxcorr::FixEntry)));
            HIPCHK(hipMalloc(&dLong, (size_t)longcap * sizeof(gmxcorr::FixEntry)));
            HIPCHK(hipMemcpyAsync(dDesc, desc.data(), desc.size() * sizeof(gmxcorr::TrajDesc),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(dItemTraj, it.data(), nitems * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(dA1, i1.data(), nitems * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(dA2, i2.data(), nitems * sizeof(int), hipMemcpyHostToDevice, s));
            gmxcorr::KsgBatchArgs ba{dDesc, dItemTraj, dA1, dA2, nitems, dPartial, dFix, dCnt, fixcap,
                                     dQueue, dLong, longcap, dFixsum, (int)ts.size()};
            gmxcorr::FixupArgs fa{};
            fa.traj = dDesc;
            fa.item_traj = dItemTraj;
            fa.a1 = dA1;
            fa.a2 = dA2;
            fa.fixsum = dFixsum;
            fa.counters = dCnt;
            HIPCHK(hipEventRecord(ev[5], s));
            unsigned counters[3] = {0, 0, 0};
            for (int attempt = 0; attempt < 3; ++attempt) {
                ba.fixq = dFix;
                ba.fixcap = fixcap;
                ba.longq = dLong;
                ba.longcap = longcap;
                fa.fixq = dFix;
                HIPCHK(gmxcorr::launch_ksg_clear(dFixsum, nitems, dCnt, dQueue, s, dPartial, 8));
                HIPCHK(hipEventRecord(ev[2], s));
                HIPCHK(gmxcorr::launch_ksg_batch(ba, nblocks, kinds, s));
                HIPCHK(hipEventRecord(ev[3], s));
                HIPCHK(hipMemcpyAsync(counters, dCnt, sizeof(counters), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (counters[2] <= longcap) {
                    // the long rows may queue tie rows to the recount
                    HIPCHK(gmxcorr::launch_ksg_batch_long_rows(ba, dFixsum, counters[2], s));
                    HIPCHK(hipMemcpyAsync(counters, dCnt, sizeof(counters), hipMemcpyDeviceToHost, s));
                    HIPCHK(hipStreamSynchronize(s));
                    if (counters[0] <= fixcap) break;
                }
                if (attempt == 2) return gmxcorr::set_error(GMXCORR_EHIP, "recount queue overflow");
                std::fprintf(stderr, "gmxcorr batch: queue overflow (recount %u, long %u rows), "
                             "re-running the wave\n", counters[0], counters[2]);
                if (counters[0] > fixcap) {
                    dfree(dFix);
                    fixcap = counters[0] + counters[0] / 4 + 1024;
                    HIPCHK(hipMalloc(&dFix, (size_t)fixcap * sizeof(gmxcorr::FixEntry)));
                }
                if (counters[2] > longcap) {
                    dfree(dLong);
                    longcap = counters[2] + counters[2] / 4 + 1024;
                    HIPCHK(hipMalloc(&dLong, (size_t)longcap * sizeof(gmxcorr::FixEntry)));
                }
            }
            HIPCHK(gmxcorr::launch_ksg_fixup(fa, counters[0], s));
            HIPCHK(gmxcorr::launch_ksg_finalize(fa, nitems, dPartial, 1, dMI, s));
            HIPCHK(hipEventRecord(ev[4], s));
            HIPCHK(hipMemcpyAsync(counters, dCnt, sizeof(counters), hipMemcpyDeviceToHost, s));
            std::vector<double> mi(nitems);
            HIPCHK(hipMemcpyAsync(mi.data(), dMI, nitems * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            float tb = 0.f, tp = 0.f, tf = 0.f, tt = 0.f;
            (void)hipEventElapsedTime(&tb, ev[1], ev[5]);
            (void)hipEventElapsedTime(&tp, ev[2], ev[3]);
            (void)hipEventElapsedTime(&tf, ev[3], ev[4]);
            (void)hipEventElapsedTime(&tt, ev[0], ev[4]);
            ms_build_total += tb;
            ms_pairs_total += tp;
            ms_fix_total += tf;
            ms_total = tt;
            fix_rows += counters[0];
            ties += counters[1];
            long_rows += counters[2];
            std::fprintf(stderr,
                         "gmxcorr batch: wave %zu/%zu: %zu trajectories, %lld pairs, rows %.1f ms, "
                         "pairs %.1f ms, long rows + recount %.1f ms (%u + %u rows)\n",
                         wv + 1, waves.size(), ts.size(), nitems, tb, tp, tf, counters[2],
                         counters[0]);
            std::fflush(stderr);
            for (size_t q = 0; q < ts.size(); ++q) {
                HostTraj& h = b->traj[ts[q]];
                h.mi.assign(mi.begin() + first[q], mi.begin() + first[q] + h.npairs());
                h.wave = (int)wv;
                h.ms_done = tt;
            }
            return GMXCORR_OK;
        };
        rc = body();
        cleanup();
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s);
    (void)hipFree(dArena);
    (void)hipFree(xArena);
    if (rc != GMXCORR_OK) return rc;
    b->st.nwaves = (int)waves.size();
    b->st.ms_total = ms_total;
    b->st.ms_build = ms_build_total;
    b->st.ms_pairs = ms_pairs_total;
    b->st.ms_fixup = ms_fix_total;
    b->st.npairs = items_total;
    b->st.fixup_rows = fix_rows;
    b->st.ties = ties;
    b->st.long_rows = long_rows;
    return GMXCORR_OK;
}

long long gmxcorr_ksg_batch_traj_bytes(const gmxcorr_ksg_batch* b, int t)
{
    if (!b || t < 0 || t >= (int)b->traj.size())
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_traj_bytes: bad trajectory index");
    return b->traj[t].row_bytes();
}

int gmxcorr_ksg_batch_result(gmxcorr_ksg_batch* b, int t, double* mi, long long npairs,
                             double* ms_done, int* wave)
{
    if (!b || t < 0 || t >= (int)b->traj.size())
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_result: bad trajectory index");
    const HostTraj& h = b->traj[t];
    if (h.mi.empty() || npairs != h.npairs())
        return gmxcorr::set_error(GMXCORR_EINVAL, "batch_result: not run or wrong pair count");
    if (mi) std::memcpy(mi, h.mi.data(), npairs * sizeof(double));
    if (ms_done) *ms_done = h.ms_done;
    if (wave) *wave = h.wave;
    return GMXCORR_OK;
}

int gmxcorr_ksg_batch_stats(const gmxcorr_ksg_batch* b, gmxcorr_batch_stats* st)
{
    if (!b || !st) return gmxcorr::set_error(GMXCORR_EINVAL, "batch_stats: null argument");
    *st = b->st;
    return GMXCORR_OK;
}

}  // extern "C"
