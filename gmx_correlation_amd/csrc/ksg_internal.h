// Internal interface between the KSG kernels (ksg_device.hip) and the host
// plan code (ksg_host.hip).  Not installed; the public ABI is gmxcorr.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

namespace gmxcorr {

// Sums that blocks and recount waves add into in any order (kind 3's row
// chunks, recounted rows): each term is rounded to a multiple of 2^-44 and
// added as a 64-bit two's-complement integer, so the total does not depend
// on the order of the adds and MI is reproducible bit for bit.  A pair's
// total stays below 2^19 in magnitude (at most 8192 rows of psi[nx] +
// psi[ny] < 19 each), inside the 2^63 / 2^44 range; the rounding moves MI
// by less than 3e-14.
using FixSum = unsigned long long;
constexpr double kFixScale = 17592186044416.0;   // 2^44
__device__ inline void fix_add(FixSum* s, double v)
{
    atomicAdd(s, (FixSum)__double2ll_rn(v * kFixScale));
}
__host__ __device__ inline double fix_value(FixSum s) { return (double)(long long)s / kFixScale; }


// Largest accepted max - min of one normalised coordinate row.  z-scores of
// N frames lie within sqrt(N) of zero, far inside it; the pair kernels rely
// on every distance being below 2^11 (padding 3072, keys below 2^12).
constexpr double kMaxRowSpread = 2048.0;

// Rejects rows with a non-finite value or a spread >= kMaxRowSpread.
int check_rows(const double* xrows, long long nrows, int N, const char* who);

// A frame row whose FP32 boundary is ambiguous, queued for the FP64 recount.
struct FixEntry {
    int      pair;     // index into the pair list of the run
    int      row;      // query frame i
    uint32_t eps_bits; // FP32 bits of the (k+1)-th smallest joint distance (self included)
    int      pad;
};

// Device buffers and shape of one KSG run.
struct KsgArgs {
    const double* X;      // [3n][N] normalised coordinates (FP64)
    const float*  D;      // [n][N][P] per-atom max-norm distance rows (FP32)
    const float*  SD;     // [n][N][sorted_pitch(N)] sorted rows (N <= kSortedMaxFrames)
    const float*  SM;     // [n][N][kSumPitch] their summaries
    const uint16_t* SI;   // [n][N][kPrefix] frame indices of the kPrefix smallest values
    const float*  LV;     // [n][N][kLongPitch] prefix stores (kind 3, N > kSortedMaxFrames):
    const uint16_t* LI;   //   the kLongPrefix smallest values (64 leading zeros), their
    const float*  LS;     //   frames, and every 64th of those values
    const double* psi;    // [N+1]
    const int*    a1;     // [npairs]
    const int*    a2;     // [npairs]
    double*       partial;   // [npairs][nsplit] sums of psi[nx]+psi[ny] over unflagged rows
    FixSum*       fixsum;    // [npairs] same over FP64-recounted rows (fix_add)
    double*       mi;        // [npairs]
    FixEntry*     fixq;      // [fixcap]
    unsigned*     counters;  // [0] queued rows, [1] exact ties, [2] long rows
    unsigned      fixcap;
    FixEntry*     longq;     // optional [longcap]: rows beyond the sorted prefix (kind 2)
    unsigned      longcap;
    int*          rc_nx;     // optional [npairs][N]
    int*          rc_ny;     // optional [npairs][N]
    unsigned char* rc_flag;  // optional [npairs][N]
    long long     npairs;
    const int*    quads;     // optional [nquads][4] pair indices per block (-1: none);
    long long     nquads;    // nullptr: pairs 4q..4q+3
    long long     P;         // row pitch in floats, row_pitch(N)
    int           N;
    int           K;
    double        psiN_phiK; // psi[N] + (psi[K] - 1/K)
    long long     vb0;       // first virtual block of this launch (set by the launcher)
    int           nsplit;    // row ranges per pair (kinds 0, 2 and 4; partial slots per pair)
    int           chunk;     // frame rows per block of kind 3
    long long     item_limit;  // > 0: run only the first item_limit work items (timing probe)
};

// One trajectory's device data and shape.  kind/s4 select the pair-kernel
// code path for it in the batched kernel (see variant_fits).
struct TrajDesc {
    const double* X;     // [3n][N] normalised rows (FP64)
    const float* D;      // [n][N][P] distance rows (FP32)
    const float* SD;     // sorted rows, summaries and prefix indices (kinds 0 and 2)
    const float* SM;
    const uint16_t* SI;
    const float* LV;     // prefix stores (kind 3)
    const uint16_t* LI;
    const float* LS;
    const double* psi;   // [N+1]
    long long P;
    int N;
    int K;
    double psiN_phiK;
    int kind;
    int s4;
    int chunk;           // frame rows per kind-3 work item
    // batched launch: this trajectory's work items (kind 3: 2x2 atom tiles
    // times chunk-row chunks, chunk-major; otherwise one per pair), the
    // global index of its first pair, and its tile table
    long long nitems;
    long long pair0;
    long long nquads;
    const int* quads;    // [nquads][4] local pair indices, -1 for none (kind 3)
};

// Arguments of the FP64 recount and finalise kernels.  item_traj == nullptr
// means every item belongs to traj[0].
struct FixupArgs {
    const TrajDesc* traj;
    const int* item_traj;
    const int* a1;
    const int* a2;
    const FixEntry* fixq;
    FixSum* fixsum;
    unsigned* counters;
    int* rc_nx;
    int* rc_ny;
    unsigned char* rc_flag;
};

// Arguments of the persistent batched pair kernel.
struct KsgBatchArgs {
    const TrajDesc* traj;
    const int* item_traj;
    const int* a1;
    const int* a2;
    long long nitems;
    double* partial;
    FixEntry* fixq;
    unsigned* counters;
    unsigned fixcap;
    unsigned long long* queue;   // [8] work counters, one per XCD
    FixEntry* longq;   // rows whose (k+1)-th neighbour lies beyond the prefix
    unsigned longcap;
    FixSum* fixsum;    // [pairs] kind 3 adds its row sums here (fix_add)
    int ntraj;         // traj[] in work order (decreasing frame count)
};

// Largest N handled by the distance-row pair kernels (cooperative kernel,
// 4 x 2048 frames).  Longer trajectories take the coordinate path (kind 4),
// which has no frame limit.
constexpr int kMaxFrames = 8192;

// The single-trajectory pair kernels of kinds 0, 2 and 4 split a pair's
// frame rows into nsplit ranges (KsgArgs::nsplit) and write one partial sum
// per range, summed in a fixed order by the finalise kernel; partial holds
// npairs * nsplit.  Kind 1 writes one partial per pair and kind 3 adds its
// row sums into fixsum.  Short ranges keep the blocks in flight on one XCD
// on the same few rows of neighbouring tiles, whose shared atoms' rows then
// come from L2: at 247 atoms x 2000 frames, k = 100, the sorted-prefix
// kernel ran 57.4 ms with 16 ranges, 56.2 with 32, 53.9 with 64 and 49.0
// with 128.  kRowSplit is the default; the autotuner may choose another
// value up to kMaxRowSplit, and plans with very many pairs take fewer so
// that the partial sums stay small against the distance rows.
constexpr int kRowSplit = 128;
constexpr int kMaxRowSplit = 256;

// Distance-row pitch in floats: a multiple of 256 frames, so the single-wave
// kernels read whole 64-lane chunks of 4 frames without a per-lane bound.
__host__ __device__ inline long long row_pitch(int N) { return ((long long)N + 255) / 256 * 256; }

// Sorted distance rows (single-wave paths, N <= kSortedMaxFrames): row pitch
// sorted_pitch(N) floats (64 front pad, N values, padding up to at least
// kPrefix values and 64 more), a summary of kSumPitch floats per row (every
// 64th sorted value) and the frame indices of the kPrefix smallest values
// (ties in ascending frame order; 0 past N).
constexpr int kSortedMaxFrames = 2048;
constexpr int kSumPitch = 64;
constexpr int kPrefix = 512;
// The sorted-prefix path (kind 2) is used for k + 1 <= kPrefixMaxT.
constexpr int kPrefixMaxT = 256;
__host__ __device__ inline long long sorted_pitch(int N)
{
    const long long n = N > kPrefix ? N : kPrefix;
    return 64 + (n + 64 + 63) / 64 * 64;
}

// Prefix stores of the long-trajectory path (kind 3, kSortedMaxFrames < N <=
// kMaxFrames): per row the kLongPrefix smallest values in order after 64
// zeros (pitch kLongPitch floats), their frames (u16), and kLongSum summary
// values (value 64m + 63).  6.4 KB per row instead of a full sorted row.
constexpr int kLongPrefix = 1024;
// Frame rows per block of the long-trajectory pair kernel (KsgArgs::chunk,
// TrajDesc::chunk; this is the default, the autotuner may choose another).
// Short chunks keep the blocks in flight on one XCD on the same rows (as the
// row ranges do for the sorted-prefix kernel): 16 rows ran 4-11% faster
// than 32 on 400 atoms x 5000 frames (k = 100) and x 8000 frames (k = 10),
// and the batch autotuner, timing 4 to 64 rows, picks 4 on each of the 29
// trajectories of 94-587 atoms x 2244-7783 frames it tunes.
constexpr int kLongChunk = 4;
constexpr int kLongPitch = 64 + kLongPrefix;
constexpr int kLongSum = kLongPrefix / 64;

// Read-only view of a loaded plan's device data (for the surrogate features).
struct PlanView {
    const double* X;   // [3n][N] normalised rows
    const float* D;    // [n][N][P] distance rows
    int natoms, N, K;
    long long P;
    hipStream_t stream;
};

// SD/SM/SI: sorted rows, their summaries and prefix indices for the
// single-wave paths, or all nullptr (required for N > kSortedMaxFrames).
hipError_t launch_build_rows(const double* X, int natoms, int N, long long P, float* D,
                             float* SD, float* SM, uint16_t* SI, hipStream_t s);
// LV/LI/LS from the distance rows D (N in (kSortedMaxFrames, kMaxFrames]).
hipError_t launch_build_prefix(const float* D, int natoms, int N, long long P, float* LV,
                               uint16_t* LI, float* LS, hipStream_t s);
// 2x2 atom tiles of the all-pairs list of n atoms (pair index of (a1, a2),
// a2 < a1, is a1*(a1-1)/2 + a2), four pair indices per tile, -1 for none.
std::vector<int> pair_quads(int n);

// Zeroes fixsum[0..n), partial[0..n) (if given), the counters and nqueue
// work counters.
hipError_t launch_ksg_clear(FixSum* fixsum, long long n, unsigned* counters,
                            unsigned long long* queue, hipStream_t s, double* partial = nullptr,
                            int nqueue = 1);
// kind 0: one wave per frame row streaming both distance rows, N <= 256*s4
//         (s4 in 1,2,3,4,5,6,8);
// kind 1: four waves per row, N <= 1024*s4 (s4 in 2..8);
// kind 2: one wave per row over the sorted prefix of the sparser atom,
//         N <= kSortedMaxFrames and K + 1 <= kPrefixMaxT (s4 ignored, 1);
// kind 3: the same over the first 256*s4 frames of the kLongPrefix prefix
//         stores (s4 in 2, 3, 4), gathering the other atom's distances from
//         its row, kSortedMaxFrames < N <= kMaxFrames and K + 1 <= kPrefixMaxT;
// kind 4: every row in FP64 from the coordinates (no distance rows, any N
//         and K; s4 ignored);
// kind < 0: default.
bool variant_fits(int kind, int s4, int N, int K);
void default_variant(int N, int K, int& kind, int& s4);
// Partial-sum slots per pair that a launch of this kind writes (the
// finalise step's nsplit; 0: all of it goes to fixsum).
int variant_slots(int kind, int nsplit);
// Work items (blocks' worth of work) of a pair launch of this kind, in the
// launch's block order (row-range-major, chunk-major or pair-major).
long long pair_items(int kind, const KsgArgs& a);
// nsplit (optional): partial-sum slots per pair the launch wrote, for the
// finalise step (0: all of it went to fixsum).  a.nsplit / a.chunk: row
// ranges per pair (kinds 0, 2, 4) and rows per block (kind 3).
hipError_t launch_ksg_pairs(const KsgArgs& a, hipStream_t s, int kind = -1, int s4 = 0,
                            int* nsplit = nullptr);
hipError_t launch_ksg_fixup(const FixupArgs& f, unsigned nqueued, hipStream_t s);
// Coordinate path (kind 4) over a device pair list with caller-provided
// scratch: partial [npairs * nsplit] doubles and counters [4]; MI to mi.
// Clears the counters, runs the pairs and the finalise on stream s.
hipError_t launch_ksg_coord_pairs(const double* X, int N, const double* psi, int K,
                                 long long npairs, const int* a1, const int* a2, int nsplit,
                                 double* partial, unsigned* counters, double* mi, hipStream_t s);
// Row ranges per pair the coordinate path uses for npairs pairs of N frames.
int coord_split(long long npairs, int N);
// Rows that the sorted-prefix path queued to a.longq (k-th neighbour beyond
// the prefix): the streaming one-wave path per row, psi sums into a.fixsum.
hipError_t launch_ksg_long_rows(const KsgArgs& a, unsigned nlong, hipStream_t s);
hipError_t launch_ksg_finalize(const FixupArgs& f, long long nitems, const double* partial,
                               int nsplit, double* mi, hipStream_t s);
// Code paths compiled into the batched kernel.
bool batch_variant_ok(int kind, int s4, int N, int K);
// Whether a variant reads the sorted rows (and needs SD, SM, SI built), or
// the prefix stores (LV, LI, LS).
inline bool variant_sorted(int kind) { return kind == 0 || kind == 2; }
inline bool variant_prefix_store(int kind) { return kind == 3; }
// Prefix chunks (256 frames) the long-trajectory path reads by default.  For
// two independent atoms the (k+1)-th joint neighbour lies about
// sqrt((k+1) N) frames deep in either atom's sorted row; correlated atoms
// need fewer.  Rows beyond the prefix run the streaming path, about 4x the
// cost of a prefix row.  Measured on 120-atom trajectories of 20 modes with
// switch and rotor couplings: 512 frames is fastest or ties at (N, k) =
// (2500, 100), (3000, 50), (4000, 50), (7000, 30), (8000, 10), all with
// (k+1) N <= 2^18; 768 frames is fastest, 4-9% ahead of 1024, at (4000, 100),
// (5000, 100), (6000, 50), (8000, 50), (k+1) N up to 505,000.  Above 2^19
// the whole 1024-frame store.
inline int long_prefix_chunks(int N, int K)
{
    const long long depth2 = (long long)(K + 1) * N;
    return depth2 <= (1LL << 18) ? 2 : depth2 <= (1LL << 19) ? 3 : 4;
}
// kinds: bit k set when some trajectory of the launch takes kind k.
hipError_t launch_ksg_batch(const KsgBatchArgs& b, int nblocks, int kinds, hipStream_t s);
// The batch's long rows (b.longq, nlong entries): cooperative streaming row,
// psi sum added to fixsum[item]; tie rows go to b.fixq.
hipError_t launch_ksg_batch_long_rows(const KsgBatchArgs& b, FixSum* fixsum, unsigned nlong,
                                      hipStream_t s);
hipError_t launch_selftest(int* out, hipStream_t s);

}  // namespace gmxcorr
