// Internal interface of the Gaussian / GCMI kernels (dense_device.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace gmxcorr {

// Largest frame count whose component row is staged whole in LDS by the
// copula kernel (8 bytes per frame within 160 KB); above it the kernel reads
// the row from global memory.
constexpr int kCopulaMaxFrames = 19456;
// Frames per GEMM launch: the covariance kernel indexes M x Kp operands with
// 32-bit offsets, so longer trajectories run in frame slabs whose partial
// matrices are summed by the reduce kernel.
constexpr long long kGemmMaxElems = (1LL << 31) - 1;
// Up to this frame count the copula sorts each row in LDS (64-bit keys).
constexpr int kCopulaSortFrames = 16384;

hipError_t launch_dense_rows(const double* traj, int natoms, int N, double* rows, hipStream_t s);
hipError_t launch_dense_split(const double* V, int M, int N, long long ldv, int Kp, double scale,
                              _Float16* hi, _Float16* lo, hipStream_t s);
// K-splits of the covariance GEMM for an M x M output (enough blocks to fill
// the chip, at least four 32-frame steps per split).
int dense_cov_splits(int M, int Kp);
// C (M x M) from the hi/lo rows [M][Kp] (Kp a multiple of 32).  Cp holds
// nsplit partial matrices; with nsplit == 1 it may be C itself.  nreduce:
// partial matrices summed into C afterwards (-1: the nsplit just written; 0:
// none, the caller reduces later, e.g. after several frame slabs).
hipError_t launch_dense_cov(const _Float16* hi, const _Float16* lo, int M, int Kp,
                            float out_scale, float* Cp, int nsplit, float* C, hipStream_t s,
                            int nreduce = -1);
// C = sum of nparts partial M x M matrices at Cp.
hipError_t launch_dense_reduce(const float* Cp, int M, int nparts, float* C, hipStream_t s);
hipError_t launch_dense_pairs(const float* C, int natoms, int mode, double* mat, hipStream_t s);
// Copula scores Z of the M component rows of V; ztab: N doubles of
// workspace (the normal score of each rank, shared by all rows).
hipError_t launch_dense_copula(const double* V, int M, int N, long long ldv, double* Z,
                               long long ldz, double* ztab, hipStream_t s);

}  // namespace gmxcorr
