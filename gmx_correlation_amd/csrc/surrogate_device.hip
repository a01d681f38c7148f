// Learned surrogate of the KSG r(MI) with exact-KSG fallback: feature
// kernels, FP32 training kernels (forward, backward, gradient reduction,
// Adam) and the BF16 matrix-core inference kernel, for gfx950.
//
// Network: 16 inputs (14 features + 2 zero pads) -> 64 -> 64 -> 2, exact-erf
// GELU on both hidden layers; outputs mu (predicted r) and s = log sigma^2.
// Loss per sample: 0.5 * (s + (y - mu)^2 * exp(-s)), mean over the batch.
// Parameter vector (FP32): W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "surrogate_internal.h"

namespace gmxcorr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr int kWave = 64;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

template <int n>
__device__ double lu_det_s(double* m)
{
    double sign = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = fabs(m[col * n + col]);
        for (int r = col + 1; r < n; ++r) {
            const double v = fabs(m[r * n + col]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) return 0.0;
        if (piv != col) {
            for (int j = 0; j < n; ++j) {
                const double t = m[col * n + j];
                m[col * n + j] = m[piv * n + j];
                m[piv * n + j] = t;
            }
            sign = -sign;
        }
        const double diag = m[col * n + col];
        for (int r = col + 1; r < n; ++r) {
            const double f = m[r * n + col] / diag;
            for (int j = col + 1; j < n; ++j) m[r * n + j] -= f * m[col * n + j];
        }
    }
    double det = sign;
    for (int i = 0; i < n; ++i) det *= m[i * n + i];
    return det;
}

// Eigenvalues of a symmetric 6x6 matrix by cyclic Jacobi rotations (FP64).
__device__ void jacobi6(double* a, double* ev)
{
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < 6; ++p)
            for (int q = p + 1; q < 6; ++q) off += a[p * 6 + q] * a[p * 6 + q];
        if (off < 1e-30) break;
        for (int p = 0; p < 6; ++p)
            for (int q = p + 1; q < 6; ++q) {
                const double apq = a[p * 6 + q];
                if (fabs(apq) < 1e-300) continue;
                const double theta = (a[q * 6 + q] - a[p * 6 + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; ++k) {
                    const double akp = a[k * 6 + p], akq = a[k * 6 + q];
                    a[k * 6 + p] = c * akp - s * akq;
                    a[k * 6 + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; ++k) {
                    const double apk = a[p * 6 + k], aqk = a[q * 6 + k];
                    a[p * 6 + k] = c * apk - s * aqk;
                    a[q * 6 + k] = s * apk + c * aqk;
                }
            }
    }
    for (int i = 0; i < 6; ++i) ev[i] = a[i * 6 + i];
    for (int i = 1; i < 6; ++i)  // insertion sort, ascending
        for (int j = i; j > 0 && ev[j] < ev[j - 1]; --j) {
            const double t = ev[j];
            ev[j] = ev[j - 1];
            ev[j - 1] = t;
        }
}

}  // namespace

// ---------------------------------------------------------------------------
// Row statistics: block per component row of X [M][N]: mean-centred copy
// into Vc and the excess kurtosis m4/m2^2 - 3.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_stats_kernel(const double* __restrict__ X, int N,
                                                        double* __restrict__ Vc,
                                                        double* __restrict__ kurt)
{
    __shared__ double red[256];
    const long long c = blockIdx.x;
    const double* row = X + c * N;
    double s = 0.0;
    for (int f = threadIdx.x; f < N; f += 256) s += row[f];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double mean = red[0] / N;
    __syncthreads();
    double m2 = 0.0, m4 = 0.0;
    for (int f = threadIdx.x; f < N; f += 256) {
        const double d = row[f] - mean;
        Vc[c * N + f] = d;
        m2 += d * d;
        m4 += d * d * d * d;
    }
    red[threadIdx.x] = m2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double M2 = red[0] / N;
    __syncthreads();
    red[threadIdx.x] = m4;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) kurt[c] = M2 > 0.0 ? (red[0] / N) / (M2 * M2) - 3.0 : 0.0;
}

// ---------------------------------------------------------------------------
// Per-atom marginal k-NN radius at S sampled frames: wave per (atom, sample),
// the k-th smallest of D[a][i][j] over j != i by an LDS histogram over the
// FP32 bit pattern (bits 30..20), refined inside the bin by 10 more bits;
// the result is the lower edge of the 2^10-ulp sub-bin (a feature, not an
// exact count).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void atom_knn_radius_kernel(const float* __restrict__ D, int N,
                                                              long long P, int K, int S,
                                                              int natoms, float* __restrict__ rad)
{
    __shared__ uint32_t lds[4][2048];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long task = (long long)blockIdx.x * 4 + w;
    if (task >= (long long)natoms * S) return;
    uint32_t* hist = lds[w];
    const int a = (int)(task / S), s = (int)(task % S);
    const int i = (int)((long long)s * N / S);
    const float* row = D + ((size_t)a * N + i) * P;
    for (int q = lane; q < 2048; q += 64) hist[q] = 0u;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    for (int j = lane; j < N; j += 64)
        if (j != i) atomicAdd(&hist[__float_as_uint(row[j]) >> 20], 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // sequential scan by lane 0 over 2048 bins is cheap for one wave per task
    int bin = 0, rank = K;
    if (lane == 0) {
        int cum = 0;
        for (int q = 0; q < 2048; ++q) {
            const int v = (int)hist[q];
            if (cum + v >= K) { bin = q; rank = K - cum; break; }
            cum += v;
        }
    }
    bin = __shfl(bin, 0, 64);
    rank = __shfl(rank, 0, 64);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int q = lane; q < 1024; q += 64) hist[q] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < N; j += 64) {
        const uint32_t key = __float_as_uint(row[j]);
        if (j != i && (key >> 20) == (uint32_t)bin) atomicAdd(&hist[(key >> 10) & 1023u], 1u);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        int cum = 0, sub = 0;
        for (int q = 0; q < 1024; ++q) {
            const int v = (int)hist[q];
            if (cum + v >= rank) { sub = q; break; }
            cum += v;
        }
        rad[(size_t)a * S + s] = __uint_as_float(((uint32_t)bin << 20) | ((uint32_t)sub << 10));
    }
}

// ---------------------------------------------------------------------------
// Pair features from the covariance C [3n][3n] (FP32) of the normalised
// coordinates and per-atom arrays; thread per pair of the list.
//   0      Gaussian MI of the pair (FP64 LU determinants)
//   1..6   eigenvalues of the pair's 6x6 correlation matrix, ascending
//   7, 8   min / max of the atoms' mean excess kurtosis
//   9, 10  min / max of the atoms' median k-NN radius
//   11, 12 min / max of the atoms' 90th-percentile k-NN radius
//   13     r of the Gaussian MI: sqrt(1 - exp(-2/3 MI))
//   14, 15 zero
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void pair_features_kernel(
    const float* __restrict__ C, int natoms, const int* __restrict__ a1,
    const int* __restrict__ a2, long long npairs, const float* __restrict__ atom_kurt,
    const float* __restrict__ atom_q50, const float* __restrict__ atom_q90,
    float* __restrict__ feat)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const int M = 3 * natoms;
    const int A = a1[p], B = a2[p];
    int idx[6];
    for (int d = 0; d < 3; ++d) { idx[d] = 3 * A + d; idx[3 + d] = 3 * B + d; }
    double cj[36], tmp[36], sd[6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) cj[r * 6 + c] = C[(size_t)idx[r] * M + idx[c]];
    double small[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) small[r * 3 + c] = cj[r * 6 + c];
    const double d1 = lu_det_s<3>(small);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) small[r * 3 + c] = cj[(r + 3) * 6 + c + 3];
    const double d2 = lu_det_s<3>(small);
    for (int t = 0; t < 36; ++t) tmp[t] = cj[t];
    const double dj = lu_det_s<6>(tmp);
    auto H = [](double det, int n) {
        return det <= 0.0 ? 0.0 : 0.5 * (n * (1.0 + log(2.0 * kPi)) + log(det));
    };
    const double gmi = H(d1, 3) + H(d2, 3) - H(dj, 6);
    for (int r = 0; r < 6; ++r) sd[r] = sqrt(fmax(cj[r * 6 + r], 1e-300));
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) tmp[r * 6 + c] = cj[r * 6 + c] / (sd[r] * sd[c]);
    double ev[6];
    jacobi6(tmp, ev);
    float* f = feat + p * 16;
    f[0] = (float)gmi;
    for (int t = 0; t < 6; ++t) f[1 + t] = (float)ev[t];
    f[7] = fminf(atom_kurt[A], atom_kurt[B]);
    f[8] = fmaxf(atom_kurt[A], atom_kurt[B]);
    f[9] = fminf(atom_q50[A], atom_q50[B]);
    f[10] = fmaxf(atom_q50[A], atom_q50[B]);
    f[11] = fminf(atom_q90[A], atom_q90[B]);
    f[12] = fmaxf(atom_q90[A], atom_q90[B]);
    f[13] = gmi > 0.0 ? (float)sqrt(1.0 - exp(-2.0 / 3.0 * gmi)) : 0.0f;
    f[14] = 0.0f;
    f[15] = 0.0f;
}

// ---------------------------------------------------------------------------
// FP32 training: thread per sample.  Writes the loss, the per-sample
// activations (x is the input itself, h1, h2) and deltas (d1, d2, d3) that
// the reduction kernel turns into parameter gradients.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ X,
                                                          const float* __restrict__ Y, int B,
                                                          float inv_b, MlpScratch sc,
                                                          float* __restrict__ out)
{
    __shared__ float w[kMlpParams];
    for (int t = threadIdx.x; t < kMlpParams; t += blockDim.x) w[t] = params[t];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    const float* W1 = w;
    const float* b1 = W1 + 64 * 16;
    const float* W2 = b1 + 64;
    const float* b2 = W2 + 64 * 64;
    const float* W3 = b2 + 64;
    const float* b3 = W3 + 2 * 64;
    const bool train = Y != nullptr;
    float x[16], h[64];
    for (int k = 0; k < 16; ++k) x[k] = X[(size_t)s * 16 + k];
    for (int n = 0; n < 64; ++n) {
        float acc = b1[n];
        for (int k = 0; k < 16; ++k) acc += W1[n * 16 + k] * x[k];
        h[n] = gelu(acc);
        if (train) {
            sc.H1[(size_t)s * 64 + n] = h[n];
            sc.G1[(size_t)s * 64 + n] = gelu_grad(acc);
        }
    }
    float mu = b3[0], ls = b3[1];
    for (int n = 0; n < 64; ++n) {
        float acc = b2[n];
        for (int k = 0; k < 64; ++k) acc += W2[n * 64 + k] * h[k];
        const float h2 = gelu(acc);
        mu += W3[n] * h2;
        ls += W3[64 + n] * h2;
        if (train) {
            sc.H2[(size_t)s * 64 + n] = h2;
            sc.G2[(size_t)s * 64 + n] = gelu_grad(acc);
        }
    }
    if (out) {
        out[(size_t)s * 2 + 0] = mu;
        out[(size_t)s * 2 + 1] = ls;
    }
    if (!train) return;
    const float y = Y[s];
    const float e = expf(-ls), r = y - mu;
    if (sc.loss) sc.loss[s] = 0.5f * (ls + r * r * e);
    const float dmu = -r * e * inv_b, dls = 0.5f * (1.0f - r * r * e) * inv_b;
    sc.D3[(size_t)s * 2 + 0] = dmu;
    sc.D3[(size_t)s * 2 + 1] = dls;
    // h now holds the layer-2 deltas
    for (int k = 0; k < 64; ++k) {
        h[k] = (W3[k] * dmu + W3[64 + k] * dls) * sc.G2[(size_t)s * 64 + k];
        sc.D2[(size_t)s * 64 + k] = h[k];
    }
    for (int j = 0; j < 64; ++j) {
        float acc = 0.0f;
        for (int n = 0; n < 64; ++n) acc += W2[n * 64 + j] * h[n];
        sc.D1[(size_t)s * 64 + j] = acc * sc.G1[(size_t)s * 64 + j];
    }
}

// Gradient of each parameter: a fixed-order sum over the batch (deterministic).
__global__ void mlp_grad_reduce_kernel(const float* __restrict__ X, const float* __restrict__ H1,
                                       const float* __restrict__ H2,
                                       const float* __restrict__ D1,
                                       const float* __restrict__ D2,
                                       const float* __restrict__ D3, int B,
                                       float* __restrict__ grad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kMlpParams) return;
    float g = 0.0f;
    if (t < 64 * 16) {                                   // W1[n][k]
        const int n = t / 16, k = t % 16;
        for (int s = 0; s < B; ++s) g += D1[(size_t)s * 64 + n] * X[(size_t)s * 16 + k];
    } else if (t < 64 * 16 + 64) {                       // b1
        const int n = t - 64 * 16;
        for (int s = 0; s < B; ++s) g += D1[(size_t)s * 64 + n];
    } else if (t < 64 * 16 + 64 + 4096) {                // W2[n][k]
        const int u = t - (64 * 16 + 64), n = u / 64, k = u % 64;
        for (int s = 0; s < B; ++s) g += D2[(size_t)s * 64 + n] * H1[(size_t)s * 64 + k];
    } else if (t < 64 * 16 + 64 + 4096 + 64) {           // b2
        const int n = t - (64 * 16 + 64 + 4096);
        for (int s = 0; s < B; ++s) g += D2[(size_t)s * 64 + n];
    } else if (t < 64 * 16 + 64 + 4096 + 64 + 128) {     // W3[o][k]
        const int u = t - (64 * 16 + 64 + 4096 + 64), o = u / 64, k = u % 64;
        for (int s = 0; s < B; ++s) g += D3[(size_t)s * 2 + o] * H2[(size_t)s * 64 + k];
    } else {                                             // b3
        const int o = t - (64 * 16 + 64 + 4096 + 64 + 128);
        for (int s = 0; s < B; ++s) g += D3[(size_t)s * 2 + o];
    }
    grad[t] = g;
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grad,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kMlpParams) return;
    const float g = grad[t];
    const float mt = b1 * m[t] + (1.0f - b1) * g;
    const float vt = b2 * v[t] + (1.0f - b2) * g * g;
    m[t] = mt;
    v[t] = vt;
    params[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ---------------------------------------------------------------------------
// BF16 inference on the matrix cores: wave per 32 samples, three layers of
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation; GELU and biases in FP32;
// activations pass between layers through LDS in BF16.  MFMA layout: lane l
// supplies A row l%32 / B column l%32 and the 8 consecutive k at 8*(l/32) of
// each 16-wide k step; result register q of lane l is row
// 8*(q/4) + 4*(l/32) + q%4, column l%32.
// Weights: W1b [64][16], W2b [64][64], W3b [32][64] (rows >= 2 zero), BF16.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_infer_bf16_kernel(const __bf16* __restrict__ W1b,
                                                             const __bf16* __restrict__ W2b,
                                                             const __bf16* __restrict__ W3b,
                                                             const float* __restrict__ bias,
                                                             const float* __restrict__ X,
                                                             long long n, float* __restrict__ out)
{
    __shared__ __bf16 act[4][32][64 + 8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long s0 = ((long long)blockIdx.x * 4 + w) * 32;
    if (s0 >= n) return;
    const float* b1 = bias;
    const float* b2 = bias + 64;
    const float* b3 = bias + 128;
    const int r = lane & 31, kq = (lane >> 5) * 8;
    // layer 1: A = X tile [32][16]
    bf16x8 xa;
    {
        const long long s = s0 + r;
        for (int t = 0; t < 8; ++t)
            xa[t] = (__bf16)(s < n ? X[s * 16 + kq + t] : 0.0f);
    }
    for (int tile = 0; tile < 2; ++tile) {
        const int col = tile * 32 + r;
        bf16x8 wb;
        for (int t = 0; t < 8; ++t) wb[t] = W1b[col * 16 + kq + t];
        float16v acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa, wb, acc, 0, 0, 0);
        for (int q = 0; q < 16; ++q) {
            const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            act[w][row][col] = (__bf16)gelu(acc[q] + b1[col]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // layer 2: A = H1 [32][64], four k steps
    float16v acc2[2] = {{}, {}};
    for (int ks = 0; ks < 4; ++ks) {
        bf16x8 ha;
        for (int t = 0; t < 8; ++t) ha[t] = act[w][r][ks * 16 + kq + t];
        for (int tile = 0; tile < 2; ++tile) {
            const int col = tile * 32 + r;
            bf16x8 wb;
            for (int t = 0; t < 8; ++t) wb[t] = W2b[col * 64 + ks * 16 + kq + t];
            acc2[tile] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha, wb, acc2[tile], 0, 0, 0);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int tile = 0; tile < 2; ++tile) {
        const int col = tile * 32 + r;
        for (int q = 0; q < 16; ++q) {
            const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            act[w][row][col] = (__bf16)gelu(acc2[tile][q] + b2[col]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // layer 3: out [32][32] (columns 0, 1 used)
    float16v acc3 = {};
    for (int ks = 0; ks < 4; ++ks) {
        bf16x8 ha, wb;
        for (int t = 0; t < 8; ++t) {
            ha[t] = act[w][r][ks * 16 + kq + t];
            wb[t] = W3b[r * 64 + ks * 16 + kq + t];
        }
        acc3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha, wb, acc3, 0, 0, 0);
    }
    if (r < 2) {
        for (int q = 0; q < 16; ++q) {
            const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const long long s = s0 + row;
            if (s < n) out[s * 2 + r] = acc3[q] + b3[r];
        }
    }
}

// FP32 master weights -> BF16 copies for inference (W3 padded to 32 rows).
__global__ void mlp_to_bf16_kernel(const float* __restrict__ params, __bf16* __restrict__ W1b,
                                   __bf16* __restrict__ W2b, __bf16* __restrict__ W3b,
                                   float* __restrict__ bias)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = params;
    const float* b1 = W1 + 1024;
    const float* W2 = b1 + 64;
    const float* b2 = W2 + 4096;
    const float* W3 = b2 + 64;
    const float* b3 = W3 + 128;
    if (t < 1024) W1b[t] = (__bf16)W1[t];
    if (t < 4096) W2b[t] = (__bf16)W2[t];
    if (t < 2048) W3b[t] = (__bf16)(t < 128 ? W3[t] : 0.0f);
    if (t < 64) {
        bias[t] = b1[t];
        bias[64 + t] = b2[t];
    }
    if (t < 2) bias[128 + t] = b3[t];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_row_stats(const double* X, int M, int N, double* Vc, double* kurt, hipStream_t s)
{
    hipLaunchKernelGGL(row_stats_kernel, dim3(M), dim3(256), 0, s, X, N, Vc, kurt);
    return hipGetLastError();
}

hipError_t launch_atom_knn_radius(const float* D, int natoms, int N, long long P, int K, int S,
                                  float* rad, hipStream_t s)
{
    const long long tasks = (long long)natoms * S;
    hipLaunchKernelGGL(atom_knn_radius_kernel, dim3((unsigned)((tasks + 3) / 4)), dim3(256), 0, s,
                       D, N, P, K, S, natoms, rad);
    return hipGetLastError();
}

hipError_t launch_pair_features(const float* C, int natoms, const int* a1, const int* a2,
                                long long npairs, const float* kurt, const float* q50,
                                const float* q90, float* feat, hipStream_t s)
{
    hipLaunchKernelGGL(pair_features_kernel, dim3((unsigned)((npairs + 127) / 128)), dim3(128), 0,
                       s, C, natoms, a1, a2, npairs, kurt, q50, q90, feat);
    return hipGetLastError();
}

hipError_t launch_mlp_fwd_bwd(const float* params, const float* X, const float* Y, int B,
                              MlpScratch& w, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3((B + 255) / 256), dim3(256), 0, s, params, X, Y, B,
                       1.0f / B, w, out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* X, const MlpScratch& w, int B, float* grad, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, X,
                       w.H1, w.H2, w.D1, w.D2, w.D3, B, grad);
    return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, int step,
                       hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, grad,
                       m, v, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_to_bf16(const float* params, __bf16* W1b, __bf16* W2b, __bf16* W3b,
                              float* bias, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(16), dim3(256), 0, s, params, W1b, W2b, W3b, bias);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_bf16(const __bf16* W1b, const __bf16* W2b, const __bf16* W3b,
                                 const float* bias, const float* X, long long n, float* out,
                                 hipStream_t s)
{
    const long long waves = (n + 31) / 32;
    hipLaunchKernelGGL(mlp_infer_bf16_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s,
                       W1b, W2b, W3b, bias, X, n, out);
    return hipGetLastError();
}

}  // namespace gmxcorr
