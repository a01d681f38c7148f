// Internal interface of the surrogate kernels (surrogate_device.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace gmxcorr {

// W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]
constexpr int kMlpParams = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2;

// Per-sample training buffers (B rows each).
struct MlpScratch {
    float* H1;    // [B][64] layer-1 activations
    float* G1;    // [B][64] GELU'(z1)
    float* H2;    // [B][64]
    float* G2;    // [B][64]
    float* D1;    // [B][64] layer-1 deltas
    float* D2;    // [B][64]
    float* D3;    // [B][2]
    float* loss;  // [B]
};

hipError_t launch_row_stats(const double* X, int M, int N, double* Vc, double* kurt, hipStream_t s);
hipError_t launch_atom_knn_radius(const float* D, int natoms, int N, long long P, int K, int S,
                                  float* rad, hipStream_t s);
hipError_t launch_pair_features(const float* C, int natoms, const int* a1, const int* a2,
                                long long npairs, const float* kurt, const float* q50,
                                const float* q90, float* feat, hipStream_t s);
hipError_t launch_mlp_fwd_bwd(const float* params, const float* X, const float* Y, int B,
                              MlpScratch& w, float* out, hipStream_t s);
hipError_t launch_mlp_grad(const float* X, const MlpScratch& w, int B, float* grad, hipStream_t s);
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, int step,
                       hipStream_t s);
hipError_t launch_mlp_to_bf16(const float* params, __bf16* W1b, __bf16* W2b, __bf16* W3b,
                              float* bias, hipStream_t s);
hipError_t launch_mlp_infer_bf16(const __bf16* W1b, const __bf16* W2b, const __bf16* W3b,
                                 const float* bias, const float* X, long long n, float* out,
                                 hipStream_t s);

}  // namespace gmxcorr
