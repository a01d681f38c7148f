// Internal interface of the Gaussian / GCMI kernels (dense_device.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace gmxcorr {

// Largest frame count whose component row is staged whole in LDS by the
// copula kernel (8 bytes per frame within 160 KB).
constexpr int kCopulaMaxFrames = 19456;
// Up to this frame count the copula sorts each row in LDS (64-bit keys).
constexpr int kCopulaSortFrames = 16384;

hipError_t launch_dense_rows(const double* traj, int natoms, int N, double* rows, hipStream_t s);
hipError_t launch_dense_split(const double* V, int M, int N, long long ldv, int Kp, double scale,
                              _Float16* hi, _Float16* lo, hipStream_t s);
// K-splits of the covariance GEMM for an M x M output (enough blocks to fill
// the chip, at least four 32-frame steps per split).
int dense_cov_splits(int M, int Kp);
// C (M x M) from the hi/lo rows [M][Kp] (Kp a multiple of 32).  Cp holds
// nsplit partial matrices; with nsplit == 1 it may be C itself.
hipError_t launch_dense_cov(const _Float16* hi, const _Float16* lo, int M, int Kp,
                            float out_scale, float* Cp, int nsplit, float* C, hipStream_t s);
hipError_t launch_dense_pairs(const float* C, int natoms, int mode, double* mat, hipStream_t s);
hipError_t launch_dense_copula(const double* V, int M, int N, long long ldv, double* Z,
                               long long ldz, hipStream_t s);

}  // namespace gmxcorr
