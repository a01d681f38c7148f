// Host side of libgmxcorr_hip.so: device plan for the KSG matrix, the
// extern "C" ABI (include/gmxcorr.h) and the reference's C++ GPU boundary
// (include/kraskov_gpu.h, replacing lib/kraskov_gpu.cu:210-308).
//
// A plan holds the FP64 coordinates of all atoms and, when the path needs
// them, the FP32 distance rows of some atoms ("slots"):
//   resident     slots = natoms: every atom's rows built once at load;
//   memory waves slots = 2B < natoms: the atoms are cut into blocks of B,
//                and for each pair of blocks (I >= J) block I's rows sit in
//                slots [0, B) and block J's in [B, 2B) while the pairs
//                between them run (rows rebuilt per wave);
//   coordinates  slots = 0: kind 4 only, every row in FP64 from the
//                coordinates (frame counts above kMaxFrames, or rows of even
//                two atoms larger than the row budget).
// The row budget is the free device memory less what the pair buffers and
// recount queues need, or the caller's max_row_bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
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

// Row ranges per pair for the coordinate path: enough work items to fill the
// device (about 8192 blocks of four waves), at least 16 rows per range.
int coord_split(long long npairs, int N)
{
    long long r = (8192 + npairs - 1) / npairs;
    r = std::min<long long>(r, std::max(1, N / 16));
    return (int)std::max<long long>(1, std::min<long long>(r, kMaxRowSplit));
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

// Bytes of one atom's distance rows and of its sorted rows / prefix stores.
long long d_atom_bytes(int N) { return (long long)N * gmxcorr::row_pitch(N) * 4; }
long long sorted_atom_bytes(int N)
{
    if (N > gmxcorr::kSortedMaxFrames) return 0;
    return (long long)N * (gmxcorr::sorted_pitch(N) * 4 + gmxcorr::kSumPitch * 4 +
                           gmxcorr::kPrefix * 2);
}
long long store_atom_bytes(int N)
{
    return (long long)N * (gmxcorr::kLongPitch * 4 + gmxcorr::kLongPrefix * 2 +
                           gmxcorr::kLongSum * 4);
}

// Default row ranges per pair of kinds 0 and 2: kRowSplit, fewer when the
// partial sums (8 bytes per range and pair) would pass 8 GB.
int default_split(long long npairs)
{
    const long long cap = (8LL << 30) / 8;
    const long long r = std::max<long long>(1, cap / std::max<long long>(1, npairs));
    return (int)std::min<long long>(gmxcorr::kRowSplit, r);
}

}  // namespace

struct gmxcorr_ksg_plan {
    int natoms = 0, N = 0, K = 0;
    long long P = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    double* dX = nullptr;    // [3n][N] all atoms
    double* dXs = nullptr;   // [3 slots][N] coordinates of the atoms in the row slots (== dX when resident)
    double* dPsi = nullptr;
    // row storage for `slots` atoms
    int slots = 0;           // natoms (resident), 2 * wave_atoms (waves) or 0 (coordinates only)
    int wave_atoms = 0;      // B: atoms per block in memory waves (0: resident or coordinates only)
    long long row_budget = 0;
    float* dD = nullptr;
    int* dQuads = nullptr;  // all-pairs run: [nquads][4] pair indices (2x2 atom tiles)
    long long nquads = 0;
    float* dSD = nullptr;   // sorted rows, summaries, prefix indices (N <= kSortedMaxFrames)
    float* dSM = nullptr;
    uint16_t* dSI = nullptr;
    float* dLV = nullptr;   // prefix stores (kSortedMaxFrames < N <= kMaxFrames, k + 1 <= 256)
    uint16_t* dLI = nullptr;
    float* dLS = nullptr;
    int block_in_slot[2] = {-1, -1};   // which block's rows the two wave slots hold
    int* dA1 = nullptr;
    int* dA2 = nullptr;
    long long pair_cap = 0;
    double* dPartial = nullptr;
    long long partial_cap = 0;   // doubles
    gmxcorr::FixSum* dFixsum = nullptr;
    double* dMI = nullptr;
    FixEntry* dFix = nullptr;
    unsigned fixcap = 0;
    FixEntry* dLong = nullptr;   // rows beyond the sorted prefix (kinds 2 and 3)
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
    std::vector<double> host_mi;   // results of the last run when it ran in memory waves
    bool host_result = false;
    int var_kind = -1, var_s4 = 0;
    int nsplit = 0;          // row ranges per pair (0: default for the pair count)
    int chunk = gmxcorr::kLongChunk;
    bool tiles = true;       // all-pairs blocks: 2x2 atom tiles, or runs of four pairs
    gmxcorr::TrajDesc* dTraj = nullptr;   // one descriptor for the recount/finalise kernels
    unsigned char* hStage = nullptr;      // pinned staging for the upload: two halves
    hipEvent_t stage_ev[2] = {};          // copy of each half done
    gmxcorr_ksg_stats st{};
};

namespace {

// Code path of a run: the caller's variant, else the default for N and k,
// with kind 3 replaced by the cooperative path when no prefix stores were
// built, and kind 4 whenever the plan holds no rows.
void plan_kind(const gmxcorr_ksg_plan* pl, int& kind, int& s4)
{
    kind = pl->var_kind;
    s4 = pl->var_s4;
    if (pl->slots == 0) { kind = 4; s4 = 1; return; }
    if (kind >= 0) return;
    gmxcorr::default_variant(pl->N, pl->K, kind, s4);
    if (kind == 3 && !pl->dLV) {
        kind = 1;
        s4 = std::max(2, (pl->N + 1023) / 1024);
    }
}

int plan_split(const gmxcorr_ksg_plan* pl, int kind, long long npairs)
{
    if (kind == 4) return pl->nsplit > 0 ? pl->nsplit : gmxcorr::coord_split(npairs, pl->N);
    return pl->nsplit > 0 ? pl->nsplit : default_split(npairs);
}

int ensure_pairs(gmxcorr_ksg_plan* pl, long long npairs, int slots_per_pair)
{
    const long long want_partial = npairs * std::max(1, slots_per_pair);
    if (npairs <= pl->pair_cap && want_partial <= pl->partial_cap) return GMXCORR_OK;
    if (npairs > pl->pair_cap) {
        dfree(pl->dA1); dfree(pl->dA2); dfree(pl->dFixsum); dfree(pl->dMI);
        pl->pair_cap = 0;   // a failed allocation below leaves no stale capacity
        HIPCHK(hipMalloc(&pl->dA1, npairs * sizeof(int)));
        HIPCHK(hipMalloc(&pl->dA2, npairs * sizeof(int)));
        HIPCHK(hipMalloc(&pl->dFixsum, npairs * sizeof(gmxcorr::FixSum)));
        HIPCHK(hipMalloc(&pl->dMI, npairs * sizeof(double)));
        pl->pair_cap = npairs;
        pl->all_pairs = false;
    }
    if (want_partial > pl->partial_cap) {
        dfree(pl->dPartial);
        pl->partial_cap = 0;
        HIPCHK(hipMalloc(&pl->dPartial, want_partial * sizeof(double)));
        pl->partial_cap = want_partial;
    }
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

// Recount and long-row queue sizes for a run of `rows` frame rows: the
// FP32-tie recount for 1% of the rows (runs measured here queue 0.04-0.2%),
// the long-row queue for 4% (a prefix path keeps its share of rows beyond
// the prefix below that; 1.8% at 400 atoms x 5000 frames, k = 100, 768
// frames), at most 64M and 256M entries (1 and 4 GiB).
long long want_fix_entries(long long rows)
{
    return std::min<long long>(std::max<long long>(65536, rows / 100), 64LL << 20);
}
long long want_long_entries(long long rows)
{
    return std::min<long long>(std::max<long long>(65536, rows / 25), 256LL << 20);
}

// Pair kernel + long rows + FP64 recount + finalise on the plan stream, for
// npairs pairs already in dA1/dA2 (slot indices).  If a queue overflowed,
// the run is repeated with a larger one.  Times add to the plan's stats.
int run_loaded(gmxcorr_ksg_plan* pl, long long npairs, bool detail, bool tiles)
{
    int kind, s4;
    plan_kind(pl, kind, s4);
    const int nsplit = plan_split(pl, kind, npairs);
    if (int rc = ensure_pairs(pl, npairs, gmxcorr::variant_slots(kind, nsplit))) return rc;
    KsgArgs a{};
    if (tiles) { a.quads = pl->dQuads; a.nquads = pl->nquads; }
    a.X = pl->dXs; a.D = pl->dD; a.SD = pl->dSD; a.SM = pl->dSM; a.SI = pl->dSI; a.LV = pl->dLV;
    a.LI = pl->dLI; a.LS = pl->dLS; a.psi = pl->dPsi; a.a1 = pl->dA1; a.a2 = pl->dA2;
    a.partial = pl->dPartial; a.fixsum = pl->dFixsum; a.mi = pl->dMI;
    a.counters = pl->dCounters;
    a.npairs = npairs; a.P = pl->P; a.N = pl->N; a.K = pl->K; a.psiN_phiK = pl->psiN_phiK;
    a.nsplit = nsplit; a.chunk = pl->chunk;
    if (detail) { a.rc_nx = pl->dRcNx; a.rc_ny = pl->dRcNy; a.rc_flag = pl->dRcFlag; }
    // queues sized up front, as an overflow re-runs the whole launch; both
    // keep their size for later runs
    if (kind != 4) {
        const long long rows = npairs * (long long)pl->N;
        const long long want_fix = want_fix_entries(rows);
        const long long want_long = want_long_entries(rows);
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
        int slots = nsplit;
        HIPCHK(gmxcorr::launch_ksg_pairs(a, pl->stream, kind, s4, &slots));
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
        HIPCHK(gmxcorr::launch_ksg_finalize(f, npairs, pl->dPartial, slots, pl->dMI, pl->stream));
        HIPCHK(hipEventRecord(pl->ev[4], pl->stream));
        HIPCHK(hipMemcpyAsync(counters, pl->dCounters, sizeof(counters), hipMemcpyDeviceToHost,
                              pl->stream));
        HIPCHK(hipStreamSynchronize(pl->stream));
        pl->st.ms_pairs += elapsed(pl->ev[2], pl->ev[3]);
        pl->st.ms_fixup += elapsed(pl->ev[3], pl->ev[4]);
        pl->st.npairs += npairs;
        pl->st.fixup_rows += counters[0];
        pl->st.long_rows += nlong;
        pl->st.ties += counters[1];
        pl->st.kind = kind;
        pl->st.s4 = s4;
        pl->st.nsplit = nsplit;
        pl->st.chunk = pl->chunk;
        return GMXCORR_OK;
    }
    return fail(GMXCORR_EHIP, "recount queues overflowed on every attempt");
}

void reset_run_stats(gmxcorr_ksg_plan* pl)
{
    if (pl->wave_atoms > 0) pl->st.ms_build = 0.0;   // waves rebuild rows inside the run
    pl->st.ms_pairs = pl->st.ms_fixup = 0.0;
    pl->st.npairs = pl->st.fixup_rows = pl->st.long_rows = pl->st.ties = 0;
    pl->st.waves = 0;
}

// Pinned staging for uploads: kStageBytes per half.  Copying pageable host
// memory straight to the device runs at about 1.3 GB/s here; through two
// alternating pinned halves the host copy of one chunk overlaps the DMA of
// the previous one.
constexpr size_t kStageBytes = 4u << 20;

int upload(gmxcorr_ksg_plan* pl, void* dst, const void* src, size_t bytes)
{
    const unsigned char* s = static_cast<const unsigned char*>(src);
    unsigned char* d = static_cast<unsigned char*>(dst);
    int half = 0;
    for (size_t off = 0; off < bytes; off += kStageBytes, half ^= 1) {
        const size_t n = std::min(kStageBytes, bytes - off);
        unsigned char* st = pl->hStage + half * kStageBytes;
        HIPCHK(hipEventSynchronize(pl->stage_ev[half]));   // this half's last DMA is done
        std::memcpy(st, s + off, n);
        HIPCHK(hipMemcpyAsync(d + off, st, n, hipMemcpyHostToDevice, pl->stream));
        HIPCHK(hipEventRecord(pl->stage_ev[half], pl->stream));
    }
    return GMXCORR_OK;
}

// Builds the rows of `nat` atoms whose coordinates sit at slot s0 of dXs.
int build_slots(gmxcorr_ksg_plan* pl, int s0, int nat)
{
    const size_t soff = (size_t)s0 * pl->N;
    HIPCHK(gmxcorr::launch_build_rows(
        pl->dXs + 3 * soff, nat, pl->N, pl->P, pl->dD + soff * pl->P,
        pl->dSD ? pl->dSD + soff * gmxcorr::sorted_pitch(pl->N) : nullptr,
        pl->dSM ? pl->dSM + soff * gmxcorr::kSumPitch : nullptr,
        pl->dSI ? pl->dSI + soff * gmxcorr::kPrefix : nullptr, pl->stream));
    if (pl->dLV)
        HIPCHK(gmxcorr::launch_build_prefix(pl->dD + soff * pl->P, nat, pl->N, pl->P,
                                            pl->dLV + soff * gmxcorr::kLongPitch,
                                            pl->dLI + soff * gmxcorr::kLongPrefix,
                                            pl->dLS + soff * gmxcorr::kLongSum, pl->stream));
    return GMXCORR_OK;
}

// Puts block b (atoms [b*B, min(n, (b+1)*B))) into wave slot w: its
// coordinates into dXs, then its rows.
int load_block(gmxcorr_ksg_plan* pl, int w, int b)
{
    if (pl->block_in_slot[w] == b) return GMXCORR_OK;
    const int B = pl->wave_atoms;
    const int a0 = b * B, nat = std::min(pl->natoms, a0 + B) - a0;
    const int s0 = w * B;
    HIPCHK(hipEventRecord(pl->ev[0], pl->stream));
    HIPCHK(hipMemcpyAsync(pl->dXs + (size_t)3 * s0 * pl->N, pl->dX + (size_t)3 * a0 * pl->N,
                          (size_t)3 * nat * pl->N * sizeof(double), hipMemcpyDeviceToDevice,
                          pl->stream));
    if (int rc = build_slots(pl, s0, nat)) return rc;
    HIPCHK(hipEventRecord(pl->ev[1], pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    pl->st.ms_build += elapsed(pl->ev[0], pl->ev[1]);
    pl->block_in_slot[w] = b;
    return GMXCORR_OK;
}

// Runs a global pair list (atom indices) in memory waves: pairs grouped by
// their blocks (I >= J), block I's rows in slot 0 and J's in slot 1.  MI,
// and the per-row outputs when asked, come back in the caller's order.
int run_waves(gmxcorr_ksg_plan* pl, long long npairs, const int* a1, const int* a2, double* mi,
              int* nx, int* ny, unsigned char* flags)
{
    const int B = pl->wave_atoms;
    const int nb = (pl->natoms + B - 1) / B;
    const bool detail = nx || ny || flags;
    // pair p into bucket (I, J), I >= J, as slot-local indices
    std::vector<std::vector<long long>> bucket((size_t)nb * nb);
    for (long long p = 0; p < npairs; ++p) {
        int I = a1[p] / B, J = a2[p] / B;
        if (I < J) std::swap(I, J);
        bucket[(size_t)I * nb + J].push_back(p);
    }
    std::vector<int> h1, h2;
    const int N = pl->N;
    for (int I = 0; I < nb; ++I) {
        for (int J = 0; J <= I; ++J) {
            const std::vector<long long>& ps = bucket[(size_t)I * nb + J];
            if (ps.empty()) continue;
            if (int rc = load_block(pl, 0, I)) return rc;
            if (J != I)
                if (int rc = load_block(pl, 1, J)) return rc;
            const long long m = (long long)ps.size();
            h1.resize(m);
            h2.resize(m);
            auto slot = [&](int a) { return a / B == I ? a - I * B : B + a - J * B; };
            for (long long t = 0; t < m; ++t) {
                h1[t] = slot(a1[ps[t]]);
                h2[t] = slot(a2[ps[t]]);
            }
            if (int rc = ensure_pairs(pl, m, 1)) return rc;
            HIPCHK(hipMemcpy(pl->dA1, h1.data(), m * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(pl->dA2, h2.data(), m * sizeof(int), hipMemcpyHostToDevice));
            if (detail && m * N > pl->rc_cap) {
                dfree(pl->dRcNx); dfree(pl->dRcNy); dfree(pl->dRcFlag);
                pl->rc_cap = 0;
                HIPCHK(hipMalloc(&pl->dRcNx, m * N * sizeof(int)));
                HIPCHK(hipMalloc(&pl->dRcNy, m * N * sizeof(int)));
                HIPCHK(hipMalloc(&pl->dRcFlag, m * N));
                pl->rc_cap = m * N;
            }
            if (int rc = run_loaded(pl, m, detail, false)) return rc;
            pl->st.waves += 1;
            std::vector<double> got(m);
            HIPCHK(hipMemcpy(got.data(), pl->dMI, m * sizeof(double), hipMemcpyDeviceToHost));
            for (long long t = 0; t < m; ++t) mi[ps[t]] = got[t];
            if (detail) {
                std::vector<int> vx(m * N), vy(m * N);
                std::vector<unsigned char> vf(m * N);
                HIPCHK(hipMemcpy(vx.data(), pl->dRcNx, m * N * sizeof(int), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(vy.data(), pl->dRcNy, m * N * sizeof(int), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(vf.data(), pl->dRcFlag, m * N, hipMemcpyDeviceToHost));
                for (long long t = 0; t < m; ++t) {
                    const size_t dst = (size_t)ps[t] * N, src = (size_t)t * N;
                    if (nx) std::memcpy(nx + dst, vx.data() + src, N * sizeof(int));
                    if (ny) std::memcpy(ny + dst, vy.data() + src, N * sizeof(int));
                    if (flags) std::memcpy(flags + dst, vf.data() + src, N);
                }
            }
        }
    }
    return GMXCORR_OK;
}

void set_descriptor(gmxcorr_ksg_plan* pl, gmxcorr::TrajDesc& td)
{
    td = gmxcorr::TrajDesc{};
    td.X = pl->dXs;
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
    td.chunk = pl->chunk;
}

}  // namespace

namespace gmxcorr {
bool plan_view(const gmxcorr_ksg_plan* pl, PlanView* v)
{
    // the surrogate's features read every atom's distance rows
    if (!pl || !pl->loaded || pl->slots != pl->natoms) return false;
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

int gmxcorr_ksg_plan_layout(int natoms, int nframes, int k, long long max_row_bytes,
                            int* slots, int* wave_atoms, int* stores)
{
    if (!slots || !wave_atoms || !stores) return fail(GMXCORR_EINVAL, "plan_layout: null output");
    if (natoms < 2 || k < 1 || nframes <= k) return fail(GMXCORR_EINVAL, "need natoms >= 2 and 1 <= k < nframes");
    if (max_row_bytes <= 0) return fail(GMXCORR_EINVAL, "plan_layout: need a row budget");
    *slots = 0;
    *wave_atoms = 0;
    *stores = 0;
    if (nframes > gmxcorr::kMaxFrames) return GMXCORR_OK;   // coordinates only
    const long long base = d_atom_bytes(nframes) + sorted_atom_bytes(nframes);
    const long long store = gmxcorr::variant_fits(3, 4, nframes, k) ? store_atom_bytes(nframes) : 0;
    const long long n = natoms;
    if (n * (base + store) <= max_row_bytes) {
        *slots = natoms;
        *stores = store > 0;
    } else if (n * base <= max_row_bytes && store == 0) {
        *slots = natoms;
    } else {
        // memory waves of two blocks; prefix stores kept if a block still
        // holds at least two atoms with them
        long long B = max_row_bytes / (2 * (base + store));
        bool st = store > 0;
        if (B < 2 && st) {
            B = max_row_bytes / (2 * base);
            st = false;
        }
        if (B >= 1) {
            B = std::min<long long>(B, (n + 1) / 2);
            *wave_atoms = (int)B;
            *slots = (int)(2 * B);
            *stores = st;
        }
    }
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_create_budget(int natoms, int nframes, int k, long long max_row_bytes,
                                   gmxcorr_ksg_plan** out)
{
    if (!out) return fail(GMXCORR_EINVAL, "plan_create: null output");
    *out = nullptr;
    if (natoms < 2) return fail(GMXCORR_EINVAL, "need at least 2 atoms");
    if (k < 1 || nframes <= k) return fail(GMXCORR_EINVAL, "need 1 <= k < nframes");
    if (max_row_bytes < 0) return fail(GMXCORR_EINVAL, "negative row budget");
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
    if ((e = hipMalloc(&pl->dX, xbytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(X)"));
    if ((e = hipMalloc(&pl->dPsi, (nframes + 1) * sizeof(double))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(psi)"));
    if ((e = hipMalloc(&pl->dCounters, 4 * sizeof(unsigned))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(counters)"));
    if ((e = hipMalloc(&pl->dTraj, sizeof(gmxcorr::TrajDesc))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(descriptor)"));
    if ((e = hipHostMalloc((void**)&pl->hStage, 2 * kStageBytes, hipHostMallocDefault)) != hipSuccess)
        return bail(hip_fail(e, "hipHostMalloc(upload staging)"));
    for (auto& ev : pl->stage_ev) {
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return bail(hip_fail(e, "hipEventCreate"));
        if ((e = hipEventRecord(ev, pl->stream)) != hipSuccess)
            return bail(hip_fail(e, "hipEventRecord"));
    }
    // Row budget.  Automatic: free memory less the all-pairs run's pair
    // buffers (ids, recount sum, MI, partial sums) and its two queues, and
    // 2 GB for the rest.
    const long double apairs = (long double)natoms * (natoms - 1) / 2;
    const long long npairs = (long long)apairs;
    if (max_row_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        if ((e = hipMemGetInfo(&free_b, &total_b)) != hipSuccess)
            return bail(hip_fail(e, "hipMemGetInfo"));
        const long long rows = npairs * (long long)nframes;
        const long long pair_bytes =
            npairs * (4 + 4 + 8 + 8 + 8 * (long long)default_split(npairs));
        const long long queue_bytes =
            (want_fix_entries(rows) + want_long_entries(rows)) * (long long)sizeof(FixEntry);
        max_row_bytes = (long long)free_b - pair_bytes - queue_bytes - (2LL << 30);
        if (max_row_bytes < (1LL << 20)) max_row_bytes = 1LL << 20;
    }
    pl->row_budget = max_row_bytes;
    int stores = 0;
    if (int rc = gmxcorr_ksg_plan_layout(natoms, nframes, k, max_row_bytes, &pl->slots,
                                         &pl->wave_atoms, &stores))
        return bail(rc);
    if (pl->slots == natoms || pl->slots == 0) {
        pl->dXs = pl->dX;   // pairs index atoms directly
    } else {
        if ((e = hipMalloc(&pl->dXs, (size_t)3 * pl->slots * nframes * sizeof(double))) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc(slot coordinates)"));
    }
    const size_t srows = (size_t)pl->slots * nframes;
    long long row_bytes = 0;
    if (pl->slots > 0) {
        if ((e = hipMalloc(&pl->dD, srows * pl->P * sizeof(float))) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc(distance rows)"));
        row_bytes += (long long)(srows * pl->P * sizeof(float));
        if (nframes <= gmxcorr::kSortedMaxFrames) {
            if ((e = hipMalloc(&pl->dSD, srows * gmxcorr::sorted_pitch(nframes) * sizeof(float))) !=
                hipSuccess)
                return bail(hip_fail(e, "hipMalloc(sorted rows)"));
            if ((e = hipMalloc(&pl->dSM, srows * gmxcorr::kSumPitch * sizeof(float))) != hipSuccess)
                return bail(hip_fail(e, "hipMalloc(sorted-row summaries)"));
            if ((e = hipMalloc(&pl->dSI, srows * gmxcorr::kPrefix * sizeof(uint16_t))) != hipSuccess)
                return bail(hip_fail(e, "hipMalloc(sorted-row prefix indices)"));
            row_bytes += pl->slots * sorted_atom_bytes(nframes);
        }
        if (stores) {
            if (hipMalloc(&pl->dLV, srows * gmxcorr::kLongPitch * sizeof(float)) != hipSuccess ||
                hipMalloc(&pl->dLI, srows * gmxcorr::kLongPrefix * sizeof(uint16_t)) != hipSuccess ||
                hipMalloc(&pl->dLS, srows * gmxcorr::kLongSum * sizeof(float)) != hipSuccess) {
                // the plan runs the cooperative path instead
                (void)hipGetLastError();
                dfree(pl->dLV); dfree(pl->dLI); dfree(pl->dLS);
            } else {
                row_bytes += pl->slots * store_atom_bytes(nframes);
            }
        }
    }
    pl->st.row_bytes = row_bytes;
    pl->st.nframes = nframes; pl->st.k = k; pl->st.natoms = natoms; pl->st.pitch = (int)pl->P;
    pl->st.slots = pl->slots;
    pl->st.wave_atoms = pl->wave_atoms;
    *out = pl;
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_create(int natoms, int nframes, int k, gmxcorr_ksg_plan** out)
{
    return gmxcorr_ksg_plan_create_budget(natoms, nframes, k, 0, out);
}

int gmxcorr_ksg_plan_load(gmxcorr_ksg_plan* pl, const double* xrows, const double* psi)
{
    if (!pl || !xrows || !psi) return fail(GMXCORR_EINVAL, "plan_load: null argument");
    const size_t nx = (size_t)3 * pl->natoms * pl->N;
    if (int rc = gmxcorr::check_rows(xrows, 3LL * pl->natoms, pl->N, "plan_load")) return rc;
    pl->loaded = false;
    pl->block_in_slot[0] = pl->block_in_slot[1] = -1;
    HIPCHK(hipEventRecord(pl->ev[0], pl->stream));
    if (int rc = upload(pl, pl->dX, xrows, nx * sizeof(double))) return rc;
    if (int rc = upload(pl, pl->dPsi, psi, (pl->N + 1) * sizeof(double))) return rc;
    HIPCHK(hipEventRecord(pl->ev[1], pl->stream));
    if (pl->slots == pl->natoms)
        if (int rc = build_slots(pl, 0, pl->natoms)) return rc;
    HIPCHK(hipEventRecord(pl->ev[5], pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    pl->st.ms_load = elapsed(pl->ev[0], pl->ev[1]);
    pl->st.ms_build = elapsed(pl->ev[1], pl->ev[5]);
    pl->psiN_phiK = psi[pl->N] + (psi[pl->K] - 1 / (1.0 * pl->K));
    gmxcorr::TrajDesc td;
    set_descriptor(pl, td);
    HIPCHK(hipMemcpy(pl->dTraj, &td, sizeof(td), hipMemcpyHostToDevice));
    pl->loaded = true;
    return GMXCORR_OK;
}

// Pair buffers sized for the current path and shape, the all-pairs list
// (a1 > a2, row by row) and its 2x2 atom tiles on the device, once per
// plan until another pair list replaces them.
static int setup_all_pairs(gmxcorr_ksg_plan* pl)
{
    const long long n = pl->natoms;
    const long long npairs = n * (n - 1) / 2;
    int kind, s4;
    plan_kind(pl, kind, s4);
    if (int rc = ensure_pairs(pl, npairs, gmxcorr::variant_slots(kind, plan_split(pl, kind, npairs))))
        return rc;
    if (!pl->all_pairs) {
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
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_run_all(gmxcorr_ksg_plan* pl)
{
    if (!pl || !pl->loaded) return fail(GMXCORR_EINVAL, "plan not loaded");
    const long long n = pl->natoms;
    const long long npairs = n * (n - 1) / 2;
    reset_run_stats(pl);
    if (pl->wave_atoms > 0) {
        std::vector<int> h1(npairs), h2(npairs);
        long long idx = 0;
        for (int a1 = 1; a1 < n; ++a1)
            for (int a2 = 0; a2 < a1; ++a2, ++idx) { h1[idx] = a1; h2[idx] = a2; }
        pl->host_mi.assign(npairs, 0.0);
        pl->all_pairs = false;
        if (int rc = run_waves(pl, npairs, h1.data(), h2.data(), pl->host_mi.data(), nullptr,
                               nullptr, nullptr))
            return rc;
        pl->host_result = true;
        pl->last_npairs = npairs;
        pl->st.all_pairs = 1;
        return GMXCORR_OK;
    }
    if (int rc = setup_all_pairs(pl)) return rc;
    pl->host_result = false;
    const int rc = run_loaded(pl, npairs, false, pl->tiles && pl->dQuads != nullptr);
    if (!rc) {
        pl->last_npairs = npairs;
        pl->st.all_pairs = 1;
    }
    return rc;
}

int gmxcorr_ksg_plan_probe(gmxcorr_ksg_plan* pl, double fraction, double* ms)
{
    if (!pl || !pl->loaded || !ms) return fail(GMXCORR_EINVAL, "plan_probe: plan not loaded");
    if (!(fraction > 0.0 && fraction <= 1.0)) return fail(GMXCORR_EINVAL, "plan_probe: fraction in (0, 1]");
    if (pl->wave_atoms > 0) return fail(GMXCORR_EUNSUP, "plan_probe: plan runs in memory waves");
    // the all-pairs list and tiles, as run_all sets them up (without a run)
    if (int rc = setup_all_pairs(pl)) return rc;
    const long long npairs = (long long)pl->natoms * (pl->natoms - 1) / 2;
    int kind, s4;
    plan_kind(pl, kind, s4);
    KsgArgs a{};
    if (pl->tiles && pl->dQuads) { a.quads = pl->dQuads; a.nquads = pl->nquads; }
    a.X = pl->dXs; a.D = pl->dD; a.SD = pl->dSD; a.SM = pl->dSM; a.SI = pl->dSI; a.LV = pl->dLV;
    a.LI = pl->dLI; a.LS = pl->dLS; a.psi = pl->dPsi; a.a1 = pl->dA1; a.a2 = pl->dA2;
    a.partial = pl->dPartial; a.fixsum = pl->dFixsum; a.mi = pl->dMI; a.counters = pl->dCounters;
    a.fixq = pl->dFix; a.fixcap = pl->fixcap; a.longq = pl->dLong; a.longcap = pl->longcap;
    a.npairs = npairs; a.P = pl->P; a.N = pl->N; a.K = pl->K; a.psiN_phiK = pl->psiN_phiK;
    a.nsplit = plan_split(pl, kind, npairs); a.chunk = pl->chunk;
    const long long items = gmxcorr::pair_items(kind, a);
    a.item_limit = std::max<long long>(1, (long long)std::ceil(fraction * (double)items));
    HIPCHK(gmxcorr::launch_ksg_clear(pl->dFixsum, npairs, pl->dCounters, nullptr, pl->stream));
    HIPCHK(hipEventRecord(pl->ev[2], pl->stream));
    HIPCHK(gmxcorr::launch_ksg_pairs(a, pl->stream, kind, s4, nullptr));
    HIPCHK(hipEventRecord(pl->ev[3], pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    *ms = elapsed(pl->ev[2], pl->ev[3]);
    pl->last_npairs = -1;   // the probe leaves no result
    pl->st.all_pairs = 0;
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_run_pairs(gmxcorr_ksg_plan* pl, long long npairs, const int* a1,
                               const int* a2, int* nx, int* ny, unsigned char* flags)
{
    if (!pl || !pl->loaded) return fail(GMXCORR_EINVAL, "plan not loaded");
    if (npairs <= 0 || !a1 || !a2) return fail(GMXCORR_EINVAL, "empty pair list");
    for (long long p = 0; p < npairs; ++p)
        if (a1[p] < 0 || a2[p] < 0 || a1[p] >= pl->natoms || a2[p] >= pl->natoms || a1[p] == a2[p])
            return fail(GMXCORR_EINVAL, "pair index out of range or a1 == a2");
    reset_run_stats(pl);
    pl->all_pairs = false;
    pl->st.all_pairs = 0;
    if (pl->wave_atoms > 0) {
        pl->host_mi.assign(npairs, 0.0);
        if (int rc = run_waves(pl, npairs, a1, a2, pl->host_mi.data(), nx, ny, flags)) return rc;
        pl->host_result = true;
        pl->last_npairs = npairs;
        return GMXCORR_OK;
    }
    const bool detail = nx || ny || flags;
    int kind, s4;
    plan_kind(pl, kind, s4);
    int rc = ensure_pairs(pl, npairs, gmxcorr::variant_slots(kind, plan_split(pl, kind, npairs)));
    if (rc) return rc;
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
    pl->host_result = false;
    rc = run_loaded(pl, npairs, detail, false);
    if (rc) return rc;
    pl->last_npairs = npairs;
    if (nx) HIPCHK(hipMemcpy(nx, pl->dRcNx, rows * sizeof(int), hipMemcpyDeviceToHost));
    if (ny) HIPCHK(hipMemcpy(ny, pl->dRcNy, rows * sizeof(int), hipMemcpyDeviceToHost));
    if (flags) HIPCHK(hipMemcpy(flags, pl->dRcFlag, rows, hipMemcpyDeviceToHost));
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_result(gmxcorr_ksg_plan* pl, double* mi, long long npairs)
{
    if (!pl || !mi) return fail(GMXCORR_EINVAL, "plan_result: null argument");
    if (npairs != pl->last_npairs) return fail(GMXCORR_EINVAL, "pair count differs from last run");
    if (pl->host_result) {
        std::memcpy(mi, pl->host_mi.data(), npairs * sizeof(double));
        return GMXCORR_OK;
    }
    HIPCHK(hipMemcpy(mi, pl->dMI, npairs * sizeof(double), hipMemcpyDeviceToHost));
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_matrix(gmxcorr_ksg_plan* pl, double* mat)
{
    if (!pl || !mat) return fail(GMXCORR_EINVAL, "plan_matrix: null argument");
    const int n = pl->natoms;
    const long long npairs = (long long)n * (n - 1) / 2;
    if (!pl->st.all_pairs || pl->last_npairs != npairs)
        return fail(GMXCORR_EINVAL, "last run was not run_all");
    std::vector<double> mi(npairs);
    if (int rc = gmxcorr_ksg_plan_result(pl, mi.data(), npairs)) return rc;
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
    if (kind >= 0 && kind != 4 && pl->slots == 0)
        return fail(GMXCORR_EUNSUP, "plan holds no distance rows: only the coordinate path (kind 4)");
    pl->var_kind = kind;
    pl->var_s4 = s4;
    return GMXCORR_OK;
}

int gmxcorr_ksg_plan_set_shape(gmxcorr_ksg_plan* pl, int nsplit, int chunk, int tile)
{
    if (!pl) return fail(GMXCORR_EINVAL, "set_shape: null plan");
    if (nsplit < 0 || nsplit > gmxcorr::kMaxRowSplit)
        return fail(GMXCORR_EINVAL, "set_shape: nsplit must be in [0, 256] (0 = default)");
    if (chunk < 0 || chunk > 1024) return fail(GMXCORR_EINVAL, "set_shape: chunk must be in [0, 1024] (0 = default)");
    if (tile < 0 || tile > 2) return fail(GMXCORR_EINVAL, "set_shape: tile must be 0, 1 or 2");
    pl->nsplit = nsplit;
    pl->chunk = chunk > 0 ? chunk : gmxcorr::kLongChunk;
    pl->tiles = tile != 2;
    if (pl->loaded) {   // the recount kernels read the descriptor
        gmxcorr::TrajDesc td;
        set_descriptor(pl, td);
        HIPCHK(hipMemcpy(pl->dTraj, &td, sizeof(td), hipMemcpyHostToDevice));
    }
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
    if (pl->dXs != pl->dX) dfree(pl->dXs);
    pl->dXs = nullptr;
    dfree(pl->dX); dfree(pl->dPsi); dfree(pl->dD); dfree(pl->dSD); dfree(pl->dSM); dfree(pl->dSI);
    dfree(pl->dLV); dfree(pl->dLI); dfree(pl->dLS); dfree(pl->dQuads); dfree(pl->dA1);
    dfree(pl->dA2); dfree(pl->dPartial); dfree(pl->dFixsum); dfree(pl->dMI); dfree(pl->dFix);
    dfree(pl->dLong); dfree(pl->dCounters); dfree(pl->dRcNx); dfree(pl->dRcNy);
    dfree(pl->dRcFlag); dfree(pl->dTraj);
    if (pl->hStage) (void)hipHostFree(pl->hStage);
    for (auto& e : pl->stage_ev)
        if (e) (void)hipEventDestroy(e);
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

// ---- Device-pointer pair list on the coordinate path ---------------------
int gmxcorr_ksg_pairs_device(const double* xrows, int natoms, int nframes, const double* psi,
                             int k, long long npairs, const int* a1, const int* a2, double* mi,
                             void* stream)
{
    if (!xrows || !psi || !a1 || !a2 || !mi) return fail(GMXCORR_EINVAL, "pairs_device: null argument");
    if (natoms < 2 || k < 1 || nframes <= k || npairs < 1)
        return fail(GMXCORR_EINVAL, "pairs_device: need natoms >= 2, 1 <= k < nframes, npairs >= 1");
    if (!device_is_gfx950()) return fail(GMXCORR_ENODEV, "no gfx950 device");
    const hipStream_t s = (hipStream_t)stream;
    const int nsplit = gmxcorr::coord_split(npairs, nframes);
    double* part = nullptr;
    unsigned* counters = nullptr;
    HIPCHK(hipMalloc(&part, (size_t)npairs * nsplit * sizeof(double)));
    hipError_t e = hipMalloc(&counters, 4 * sizeof(unsigned));
    if (e == hipSuccess)
        e = gmxcorr::launch_ksg_coord_pairs(xrows, nframes, psi, k, npairs, a1, a2, nsplit, part,
                                            counters, mi, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(part);
    if (counters) (void)hipFree(counters);
    if (e != hipSuccess) return hip_fail(e, "pairs_device");
    return GMXCORR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free variant of gmxcorr_ksg_pairs_device.  Its scratch (the
// coordinate path's counters and per-range partial sums) is this file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel: its device address is looked
// up once per device with hipGetSymbolAddress and kept in g_scratch_dev.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = 32u << 20;   // 32 MiB
constexpr size_t kScratchHead = 256;          // counters[4] at the start
constexpr long long kScratchPartials = (long long)((kScratchBytes - kScratchHead) / sizeof(double));
constexpr int kScratchMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
void* g_scratch_dev[kScratchMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once).
int scratch_address(void** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices) return fail(GMXCORR_EUNSUP, "device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GMXCORR_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return GMXCORR_OK;
}
}  // namespace

extern "C" int gmxcorr_ksg_pairs_device_noalloc(const double* xrows, int natoms, int nframes,
                                                const double* psi, int k, long long npairs,
                                                const int* a1, const int* a2, double* mi,
                                                void* stream)
{
    if (!xrows || !psi || !a1 || !a2 || !mi)
        return fail(GMXCORR_EINVAL, "pairs_device_noalloc: null argument");
    if (natoms < 2 || k < 1 || nframes <= k || npairs < 1)
        return fail(GMXCORR_EINVAL,
                    "pairs_device_noalloc: need natoms >= 2, 1 <= k < nframes, npairs >= 1");
    if (npairs > kScratchPartials)
        return fail(GMXCORR_EINVAL, "pairs_device_noalloc: npairs above the scratch maximum of " +
                                        std::to_string(kScratchPartials) + " pairs");
    if (!device_is_gfx950()) return fail(GMXCORR_ENODEV, "no gfx950 device");
    void* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    // the allocating entry point's row ranges when they fit, else fewer
    int nsplit = gmxcorr::coord_split(npairs, nframes);
    if (npairs * nsplit > kScratchPartials) nsplit = (int)(kScratchPartials / npairs);
    unsigned* counters = reinterpret_cast<unsigned*>(base);
    double* part = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kScratchHead);
    HIPCHK(gmxcorr::launch_ksg_coord_pairs(xrows, nframes, psi, k, npairs, a1, a2, nsplit, part,
                                           counters, mi, (hipStream_t)stream));
    return GMXCORR_OK;
}

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
