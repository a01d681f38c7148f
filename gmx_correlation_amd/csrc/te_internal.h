// Internal interface of the transfer-entropy kernels (te_device.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmxcorr {

struct TeArgs {
    const double* traj;   // [n][N][3] mean-centred coordinates
    int N;
    int lag;
    int k;                // >= 1; the heap kernel takes k <= kTeHeapMaxK
    int path;             // pass 1: 0 heap for k <= kTeHeapMaxK, else select; 1 heap; 2 select
                          // (the prefix path is chosen on the host)
    const int* a1;        // target atom of ordered pair p (TE(a2 -> a1))
    const int* a2;        // source atom
    long long npairs;
    int nqblocks;         // ceil((N - lag) / 64)
    const double* psi;    // [N - lag + 1]
    double* partial;      // [npairs][nqblocks]
    float* root;          // [te_chunk_pairs(nqblocks) * nqblocks * 64] per-query fl32(eps)
    // List mode of the heap kernel (qblocks != nullptr): block b takes the
    // queries qlist[p * (N - lag) + first + lane], lane < qcount[p] - first,
    // of pair p, where {p, first} = qblocks[b]; otherwise block b is query
    // block b % nqblocks of pair b / nqblocks.
    const int2* qblocks = nullptr;
    const unsigned* qcount = nullptr;
    const uint16_t* qlist = nullptr;
    // Prefix path: the exact FP64 radius of each query ([chunk pairs][nqblocks
    // * 64]), from the select kernel or, for the queries it lists, the heap
    // kernel's list mode
    double* eps64 = nullptr;
};

// Prefix path of pass 1 (k <= kTeHeapMaxK, N - lag <= kTePrefixMaxFrames).
// For target atom a1 and query q, M(q, j) = max(|X_a1(q+L) - X_a1(j+L)|,
// |X_a1(q) - X_a1(j)|) (max norms) does not depend on the source atom.  The
// build stores, per target slot and query, the kTePrefix frames of smallest
// fl32(M) (the query's own frame first among them), sorted by (key, frame).
// A source's joint key is max(fl32(M), fl32(|Y(q) - Y(j)|)) >= fl32(M), so
// every frame outside the prefix has a key >= H, the prefix's largest: when
// the (k+1)-th smallest joint key of the prefix is below H it is the root.
// Other queries go to a per-pair list that the heap kernel finishes.
struct TePrefixArgs {
    const int* targets;   // [ntargets] target atoms of the slots
    long long ntargets;
    const int* tslot;     // [npairs] slot of each pair's target
    uint32_t* tv;         // [ntargets][N - lag][kTePrefix] fl32(M) bits, ascending
    uint16_t* ti;         // [ntargets][N - lag][kTePrefix] their frames
    unsigned* qcount;     // [npairs] queries left to the heap kernel
    uint16_t* qlist;      // [npairs][N - lag] those queries
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
// Frames of the prefix path's stores per target and query.
constexpr int kTePrefix = 1024;
// Largest N - lag of the prefix path: the build holds a row's keys in
// registers, at most 32 per thread of 512 (frames are u16 in the stores).
constexpr int kTePrefixMaxFrames = 512 * 32;
// Pairs per distinct target atom from which the automatic path builds
// prefix stores (their build is shared by the target's pairs).
constexpr int kTePrefixMinPairs = 8;
// Valid frames above which the automatic path takes the prefix stores (it
// is faster from 400 frames on: 1.7x at 80 atoms x 400 frames, 2.1x at
// 80 x 1000, 1.7x at 40 x 5000).
constexpr int kTePrefixAutoFrames = 256;
// Queries (64 per query block) per launch pair of the heap and count
// kernels: each dispatch stays far below 2^32 work-items, and the root
// buffer at 4 bytes per query below 2 GiB.
constexpr long long kTeChunkQueries = 1LL << 29;
// Pairs per launch pair for nqblocks query blocks per pair.
long long te_chunk_pairs(int nqblocks);
hipError_t launch_te(const TeArgs& a, hipStream_t s);
// Prefix path pieces, for the pairs of one launch chunk: the stores of the
// chunk's target slots, the select over them (roots, or the query appended
// to its pair's list), the heap kernel over the listed queries (nblocks
// blocks of a.qblocks), then the count kernel over every query.
hipError_t launch_te_prefix_build(const TeArgs& a, const TePrefixArgs& pa, hipStream_t s);
hipError_t launch_te_prefix_select(const TeArgs& a, const TePrefixArgs& pa, hipStream_t s);
hipError_t launch_te_heap_list(const TeArgs& a, long long nblocks, hipStream_t s);
hipError_t launch_te_count(const TeArgs& a, hipStream_t s);
// The counts of the prefix path from the exact radii a.eps64.
hipError_t launch_te_count64(const TeArgs& a, hipStream_t s);
hipError_t launch_te_finalize(const TeArgs& a, double* mat, int natoms, hipStream_t s);

}  // namespace gmxcorr
