// Internal interface of the transfer-entropy kernels (te_device.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace gmxcorr {

struct TeArgs {
    const double* traj;   // [n][N][3] mean-centred coordinates
    int N;
    int lag;
    int k;                // >= 1; the heap kernel takes k <= kTeHeapMaxK
    int path;             // pass 1: 0 heap for k <= kTeHeapMaxK, else select; 1 heap; 2 select
    const int* a1;        // target atom of ordered pair p (TE(a2 -> a1))
    const int* a2;        // source atom
    long long npairs;
    int nqblocks;         // ceil((N - lag) / 64)
    const double* psi;    // [N - lag + 1]
    double* partial;      // [npairs][nqblocks]
    float* root;          // [te_chunk_pairs(nqblocks) * nqblocks * 64] per-query fl32(eps)
};

// Largest k of the heap kernel: its LDS holds (k - 9) * 256 bytes per wave.
constexpr int kTeHeapMaxK = 128;
// Heap positions per lane held in registers by the heap kernel (the root
// and its 9 children); positions beyond them are in LDS.
constexpr int kTeHeapRegs = 10;
// The heap kernel's cap on the root: the kTeSampleRank-th smallest key of
// every stride-th frame, stride = max(kTeSampleStride, kTeSampleStride *
// (k + 1) / 101), so that on average at most 3.2 of the k + 1 smallest keys
// are sampled; the cap then holds fewer than k + 1 keys (and the wave
// repeats its frame loop without it) with probability about 5e-6 per query
// for frames in random order.
constexpr int kTeSampleStride = 32;
// Frames per LDS stage chunk of the heap kernel (3 * 3 * 8 bytes each, after
// the heap's slots; even).
constexpr int kTeStage = 32;
// Ordered pairs per wave of the count kernel when they share their target
// atom (consecutive pairs of a launch; the full matrix is target-major).
constexpr int kTeCountGroup = 4;
constexpr int kTeSampleRank = 14;
// Queries (64 per query block) per launch pair of the heap and count
// kernels: each dispatch stays far below 2^32 work-items, and the root
// buffer at 4 bytes per query below 2 GiB.
constexpr long long kTeChunkQueries = 1LL << 29;
// Pairs per launch pair for nqblocks query blocks per pair.
long long te_chunk_pairs(int nqblocks);
hipError_t launch_te(const TeArgs& a, hipStream_t s);
hipError_t launch_te_finalize(const TeArgs& a, double* mat, int natoms, hipStream_t s);

}  // namespace gmxcorr
