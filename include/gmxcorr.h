/*
 * gmxcorr.h -- extern "C" ABI of libgmxcorr_hip.so (MI355X / gfx950).
 *
 * Plain pointers and sizes only.  Host arrays are passed in and out; device
 * memory is owned by the library.  Every function returns GMXCORR_OK (0) or a
 * negative status; gmxcorr_last_error() gives the message of the last failure
 * on the calling thread.
 *
 * The reference's own boundary for this path is the C++ pair in
 * kraskov_gpu.h (lib/kraskov_gpu.h:20,32), which this library also exports.
 * The functions below expose the same computation to FFI callers (ctypes) and
 * add the device-resident plan that the benchmark times.
 */
#ifndef GMXCORR_H
#define GMXCORR_H

#ifdef __cplusplus
extern "C" {
#endif

#define GMXCORR_OK       0
#define GMXCORR_EINVAL  -1   /* bad argument (sizes, k, non-finite input) */
#define GMXCORR_ENODEV  -2   /* no usable gfx950 device */
#define GMXCORR_ENOMEM  -3   /* device allocation failed */
#define GMXCORR_EHIP    -4   /* HIP runtime error */
#define GMXCORR_EUNSUP  -5   /* shape outside what the kernels support */
#define GMXCORR_EOF      1   /* gmxcorr_xtc_read: no further frame */

const char* gmxcorr_last_error(void);

/* 1 if device 0 is a gfx950 GPU, else 0.  Same test as kraskov_gpu_probe()
 * (lib/kraskov_gpu.h:20). */
int gmxcorr_probe(void);

/* Runs the wave-level scan/reduction primitives of the kernels once on
 * device 0 and copies their 67 results to out (test hook). */
int gmxcorr_selftest_wave(int* out);

/* Largest frame count of the distance-row pair kernels (kinds 0-3).  Longer
 * trajectories run on the coordinate path (kind 4), which has no limit. */
int gmxcorr_ksg_max_frames(void);

/* One-shot KSG-2 matrix: the flat-array form of kraskov_corrmatrix_gpu()
 * (lib/kraskov_gpu.h:32).  xrows: [3*natoms][nframes] normalised rows as
 * kraskov_prepare() leaves them (lib/kraskov.cpp:2168-2220); psi:
 * [nframes+1]; mat: [natoms*natoms], written with the 2000 diagonal and the
 * symmetric MI (nats).  1 <= k < nframes, natoms >= 2. */
int gmxcorr_ksg_matrix(const double* xrows, int natoms, int nframes, const double* psi, int k,
                       double* mat);

/* ---- Device-resident plan ------------------------------------------------
 * create -> load (host rows -> HBM, build the per-atom distance rows) ->
 * run_all / run_pairs (any number of times) -> result / stats -> destroy.
 * Pair order of run_all is the reference's: for a1 = 1..n-1, a2 = 0..a1-1
 * (lib/kraskov_gpu.cu:248-256, lib/correlation_core.cpp:403-405). */
typedef struct gmxcorr_ksg_plan gmxcorr_ksg_plan;

typedef struct {
    double ms_load;        /* host->device copy of rows and psi */
    double ms_build;       /* per-atom distance-row build (in memory waves: all of
                              the last run's rebuilds) */
    double ms_pairs;       /* main pair kernel (HIP events on the plan stream;
                              summed over memory waves) */
    double ms_fixup;       /* long rows + FP64 recount of flagged rows + finalise */
    long long npairs;      /* pairs in the last run */
    long long fixup_rows;  /* rows recounted in FP64 (FP32 near-ties) */
    long long long_rows;   /* rows whose k-th neighbour lay beyond the sorted
                              prefix, rerun on the streaming path */
    long long ties;        /* rows with an exact FP64 tie at the k-th distance */
    long long row_bytes;   /* bytes of distance rows (and sorted rows / prefix
                              stores) the plan holds in HBM */
    int nframes, k, natoms, pitch;
    int kind, s4;          /* code path of the last run (kind 4: coordinates) */
    int nsplit, chunk;     /* its row ranges per pair and kind-3 rows per block */
    int slots;             /* atoms whose rows fit at once (natoms: resident;
                              0: coordinate path only) */
    int wave_atoms;        /* atoms per block in memory waves (0: no waves) */
    int waves;             /* block pairs the last run went through */
    int all_pairs;         /* 1 if the last run was run_all */
} gmxcorr_ksg_stats;

/* Layout of a plan: rows resident for every atom, memory waves of two
 * blocks of wave_atoms atoms, or the coordinate path only (slots = 0; also
 * for nframes > gmxcorr_ksg_max_frames()), for a budget of max_row_bytes of
 * distance rows, sorted rows and prefix stores.  Host-only arithmetic (no
 * device needed).  stores: 1 if the long-trajectory prefix stores are
 * included. */
int  gmxcorr_ksg_plan_layout(int natoms, int nframes, int k, long long max_row_bytes,
                             int* slots, int* wave_atoms, int* stores);
/* Plan with the automatic row budget (free device memory less the pair
 * buffers and recount queues of an all-pairs run and 2 GB). */
int  gmxcorr_ksg_plan_create(int natoms, int nframes, int k, gmxcorr_ksg_plan** out);
/* Same with an explicit budget (0: automatic). */
int  gmxcorr_ksg_plan_create_budget(int natoms, int nframes, int k, long long max_row_bytes,
                                    gmxcorr_ksg_plan** out);
int  gmxcorr_ksg_plan_load(gmxcorr_ksg_plan* plan, const double* xrows, const double* psi);
int  gmxcorr_ksg_plan_run_all(gmxcorr_ksg_plan* plan);
/* MI of an explicit pair list (a1[p] != a2[p]), e.g. the exact-KSG fallback
 * of the surrogate.  When nx/ny/flags are non-NULL they receive per-frame
 * marginal counts [npairs][nframes] (the reference's nx2/ny2,
 * lib/kraskov.cpp:1311-1314) and flags (1 = recounted in FP64, 2 = exact tie). */
int  gmxcorr_ksg_plan_run_pairs(gmxcorr_ksg_plan* plan, long long npairs, const int* a1,
                                const int* a2, int* nx, int* ny, unsigned char* flags);
/* MI (nats) of the last run, in its pair order. */
int  gmxcorr_ksg_plan_result(gmxcorr_ksg_plan* plan, double* mi, long long npairs);
/* Fill mat[natoms*natoms] from the last run_all (diagonal 2000). */
int  gmxcorr_ksg_plan_matrix(gmxcorr_ksg_plan* plan, double* mat);
int  gmxcorr_ksg_plan_stats(const gmxcorr_ksg_plan* plan, gmxcorr_ksg_stats* st);
/* Pair-kernel variant (the autotuner's knob): kind 0 = one wavefront per
 * frame row streaming both distance rows, 256*s4 frames (s4 in
 * 1,2,3,4,5,6,8); kind 1 = four wavefronts per row holding 1024*s4 frames
 * (s4 in 2..8); kind 2 = one wavefront per row over the 512 nearest frames
 * of the sparser atom (nframes <= 2048, k + 1 <= 256; s4 ignored); kind 3 =
 * the same over the first 256*s4 frames of 1024-frame prefix stores
 * (2048 < nframes <= 8192, k + 1 <= 256, s4 = 2, 3 or 4; built when they fit
 * in HBM); kind 4 = every row in FP64 from the coordinates (any nframes and
 * k, no distance rows; the default above 8192 frames); kind -1 =
 * the default for the plan's frame count and k.  variant_fits: 1 if
 * (kind, s4) is compiled and covers nframes at this k. */
int  gmxcorr_ksg_plan_set_variant(gmxcorr_ksg_plan* plan, int kind, int s4);
/* Row-block and tile shape (the autotuner's second knob): nsplit row ranges
 * per pair for kinds 0, 2 and 4 (1..256; 0 = default: 128 for kinds 0 and 2,
 * fewer when the partial sums would pass 8 GB, and enough to fill the device
 * for kind 4), chunk frame rows per block of kind 3 (1..1024; 0 = default
 * 4), tile: the four pairs of a block in the all-pairs run (0 or 1: a 2x2
 * atom tile, 2: four consecutive pairs sharing their first atom). */
int  gmxcorr_ksg_plan_set_shape(gmxcorr_ksg_plan* plan, int nsplit, int chunk, int tile);
int  gmxcorr_ksg_variant_fits(int kind, int s4, int nframes, int k);
/* Timing probe of the autotuner: the all-pairs pair kernel with the plan's
 * current variant and shape over only the first `fraction` of its work items
 * (the same block order, so the same rows shared in L2 by the blocks in
 * flight, over fewer frame rows); ms receives its time (HIP events).  Leaves
 * no result: plan_result fails until the next run.  Resident plans only. */
int  gmxcorr_ksg_plan_probe(gmxcorr_ksg_plan* plan, double fraction, double* ms);
void gmxcorr_ksg_plan_destroy(gmxcorr_ksg_plan* plan);

/* ---- Device-pointer pair list on the coordinate path -----------------------
 * MI of npairs pairs (a1[p] != a2[p], both < natoms) computed from the
 * coordinates alone: every frame row in FP64 (kind 4), no distance rows, any
 * nframes and k < nframes, device memory O(natoms * nframes) as the
 * reference's kernel (lib/kraskov_gpu.cu:107-196).  xrows [3*natoms][nframes],
 * psi [nframes+1], a1, a2 [npairs] and mi [npairs] are DEVICE pointers; the
 * work runs on `stream` (a hipStream_t, NULL = default stream).  Allocates its
 * scratch (8 * npairs * nsplit bytes, nsplit <= 256) and returns when mi is
 * written.  Indices are not checked on the device: the caller guarantees
 * them. */
int gmxcorr_ksg_pairs_device(const double* xrows, int natoms, int nframes, const double* psi,
                             int k, long long npairs, const int* a1, const int* a2, double* mi,
                             void* stream);

/* Allocation-free variant of gmxcorr_ksg_pairs_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; it calls no hipMalloc/hipFree and
 * does not synchronise: the work is enqueued on `stream` and mi is valid once
 * the stream reaches it.  Scratch is a 32 MiB device array inside the library,
 * one per GPU, held until the process ends, so npairs is limited to 4,194,272
 * (larger lists are rejected with GMXCORR_EINVAL).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams or host
 * threads) race and must be serialised by the caller. */
int gmxcorr_ksg_pairs_device_noalloc(const double* xrows, int natoms, int nframes,
                                     const double* psi, int k, long long npairs, const int* a1,
                                     const int* a2, double* mi, void* stream);

/* ---- Gaussian and Gaussian-copula matrices --------------------------------
 * traj: mean-centred fluctuations [natoms][nframes][3] (t_traj.x flattened).
 * mat:  [natoms*natoms], diagonal 2000, symmetric MI in nats.
 * The second-moment matrix V^T V / N (3n x 3n) is a split-precision FP16
 * MFMA GEMM with FP32 accumulation; log-determinants are FP64.
 * cov_out (optional, [3n*3n]) receives that matrix; z_out (optional,
 * [3n][nframes]) the copula scores. */
typedef struct {
    double ms_load;      /* host->device copy + layout change */
    double ms_copula;    /* copula transform (GCMI only) */
    double ms_split;     /* FP16 hi/lo split */
    double ms_gemm;      /* MFMA second-moment GEMM */
    double ms_pairs;     /* per-pair FP64 log-determinants */
    double gemm_flops;   /* 2 * (3n)^2 * N, counted once */
    int natoms, nframes;
} gmxcorr_dense_stats;

/* Frames per slab of the second-moment GEMM (0, the default: as many as the
 * kernel's 32-bit operand offsets allow, 2^31 / (3 natoms)).  Longer
 * trajectories run slab by slab and the slabs' partial matrices are summed;
 * smaller slabs bound the FP16 operand buffers.  Process-wide setting. */
int gmxcorr_dense_set_slab_frames(int frames);

/* lib/correlation_core.cpp:280-311 (gauss_corrmatrix). */
int gmxcorr_gauss_matrix(const double* traj, int natoms, int nframes, double* mat,
                         double* cov_out, gmxcorr_dense_stats* st);
/* lib/gcmi.cpp:246-317 (gcmi_corrmatrix); the reference's GPU signature for
 * this estimator is gcmi_corrmatrix_metal (lib/gcmi_metal.h), which takes the
 * copula scores; here the transform runs on the GPU too.  nframes >= 6, no
 * upper limit (rows above 19,456 frames are ranked from global memory). */
int gmxcorr_gcmi_matrix(const double* traj, int natoms, int nframes, double* mat,
                        double* cov_out, double* z_out, gmxcorr_dense_stats* st);

/* ---- Transfer entropy -------------------------------------------------------
 * traj: mean-centred fluctuations [natoms][nframes][3] (raw, as the reference
 * CPU path uses them).  mat[a1*n+a2] = TE(a2 -> a1) at `lag`, diagonal 0,
 * all zero (with a warning) when nframes - lag <= k
 * (lib/transfer_entropy.cpp:63-221; the reference's GPU signature is
 * te_corrmatrix_metal, lib/te_metal.h).  Any 1 <= k < nframes - lag;
 * non-finite coordinates are rejected.  ms (optional): kernel time.
 * te_pairs: the same for an explicit list of ordered pairs. */
int gmxcorr_te_matrix(const double* traj, int natoms, int nframes, int k, int lag, double* mat,
                      double* ms);
int gmxcorr_te_pairs(const double* traj, int natoms, int nframes, int k, int lag,
                     long long npairs, const int* a1, const int* a2, double* te, double* ms);
/* Kernel that finds each query's k-th neighbour distance in later TE calls
 * of this process (no counterpart in the reference): 0 = automatic (the
 * prefix stores for k <= 128 and 256 < nframes - lag <= 16384 when the
 * pairs average 8 or more per target atom, else a per-lane heap for
 * k <= 128, a radix select above), 1 = the heap (k <= 128 only), 2 = the
 * radix select (any k), 3 = the prefix stores (k <= 128, nframes - lag <=
 * 16384): per target atom and query the 1024 frames nearest in the target's
 * own two 3-D distances, shared by the target's pairs, with the heap for
 * the queries whose neighbour lies beyond them.  Results are identical. */
int gmxcorr_te_set_path(int path);
/* Counts of the last TE call (no counterpart in the reference): out[0]
 * queries, out[1] resolved from the prefix stores, out[2] left to the heap,
 * out[3] 1 when the prefix path ran. */
int gmxcorr_te_stats(long long* out);

/* ---- Batched uneven trajectories (no counterpart in the reference) --------
 * Many independent trajectories, each with its own atom count, frame count
 * and k.  run() packs them into memory waves of at most max_row_bytes of
 * distance rows; each wave is one persistent pair-kernel launch over all its
 * pairs, each trajectory on the code path (kind, s4) given at add() (from
 * the autotuner; -1 = smallest path that fits). */
typedef struct gmxcorr_ksg_batch gmxcorr_ksg_batch;
typedef struct {
    double ms_total;      /* first upload to last result, all waves */
    double ms_build;      /* uploads + row builds */
    double ms_pairs;      /* persistent pair launches */
    double ms_fixup;      /* long rows + FP64 recount + finalise */
    long long npairs;
    long long fixup_rows;
    long long ties;
    int nwaves;
    long long long_rows;  /* rows beyond the sorted prefix (streaming pass) */
} gmxcorr_batch_stats;
int  gmxcorr_ksg_batch_create(gmxcorr_ksg_batch** out);
/* Returns the trajectory's index (>= 0) or a negative status. */
int  gmxcorr_ksg_batch_add(gmxcorr_ksg_batch* b, const double* xrows, const double* psi,
                           int natoms, int nframes, int k, int kind, int s4);
/* Device bytes that trajectory t's distance rows (and, on the one-wave-per-
 * row path, sorted rows) take in a memory wave; max_row_bytes of
 * gmxcorr_ksg_batch_run must be at least the largest of these. */
long long gmxcorr_ksg_batch_traj_bytes(const gmxcorr_ksg_batch* b, int t);
int  gmxcorr_ksg_batch_run(gmxcorr_ksg_batch* b, long long max_row_bytes, int nblocks);
/* Frame rows per work item of trajectory t when it runs on kind 3 (1..1024;
 * 0 = default 4); the autotuner's shape knob for the batch. */
int  gmxcorr_ksg_batch_set_chunk(gmxcorr_ksg_batch* b, int t, int chunk);
/* MI of trajectory t's pairs (a1 = 1.., a2 < a1); ms_done: batch-relative
 * completion time of its wave; wave: its wave index (both optional). */
int  gmxcorr_ksg_batch_result(gmxcorr_ksg_batch* b, int t, double* mi, long long npairs,
                              double* ms_done, int* wave);
int  gmxcorr_ksg_batch_stats(const gmxcorr_ksg_batch* b, gmxcorr_batch_stats* st);
void gmxcorr_ksg_batch_destroy(gmxcorr_ksg_batch* b);

/* ---- Learned surrogate (no counterpart in the reference) -------------------
 * Features of a pair list on a loaded KSG plan: feat [npairs][16] holds the
 * Gaussian MI, the six eigenvalues of the pair's 6x6 correlation matrix,
 * min/max of the atoms' marginal excess kurtosis, min/max of the atoms'
 * median and 90th-percentile k-NN radius (nsample frames per atom), r of the
 * Gaussian MI, and two zeros.  ms (optional): device time. */
int gmxcorr_ksg_plan_features(gmxcorr_ksg_plan* plan, long long npairs, const int* a1,
                              const int* a2, int nsample, float* feat, double* ms);

/* MLP 16 -> 64 -> 64 -> 2 (GELU), outputs (mu, log sigma^2) of r(MI); FP32
 * master weights trained on the GPU with Adam; BF16 matrix-core inference. */
typedef struct gmxcorr_mlp gmxcorr_mlp;
int  gmxcorr_mlp_num_params(void);
int  gmxcorr_mlp_create(int max_batch, gmxcorr_mlp** out);
int  gmxcorr_mlp_set_params(gmxcorr_mlp* m, const float* params);
int  gmxcorr_mlp_get_params(gmxcorr_mlp* m, float* params);
int  gmxcorr_mlp_grad(gmxcorr_mlp* m, const float* x, const float* y, int batch, float* grad,
                      float* loss);
int  gmxcorr_mlp_train_epoch(gmxcorr_mlp* m, const float* x, const float* y, long long n,
                             int batch, float lr, float* mean_loss);
int  gmxcorr_mlp_forward(gmxcorr_mlp* m, const float* x, long long n, int bf16, float* out,
                         double* ms);
void gmxcorr_mlp_destroy(gmxcorr_mlp* m);

/* ---- XTC trajectories (GROMACS compressed coordinates) ----------------------
 * Frame ingest without GROMACS: the reference's frames come from the GROMACS
 * analysis framework (src/gmx_correlation.cpp:158-182), which reads XTC.
 * Host-only code, no device needed.  write: 0 = read, 1 = write.  read fills
 * natoms (always), step, time, box [9] and precision when non-NULL, and x
 * [natoms][3] in nm (max_atoms: the capacity of x); returns GMXCORR_EOF after
 * the last frame.  With x NULL or too small it sets natoms, leaves the file
 * at the same frame and returns GMXCORR_EINVAL (size the buffer, call again).  write: precision is the XTC's (1000 = 0.001 nm);
 * natoms <= 9 frames are stored uncompressed. */
typedef struct gmxcorr_xtc gmxcorr_xtc;
int  gmxcorr_xtc_open(const char* path, int write, gmxcorr_xtc** out);
int  gmxcorr_xtc_read(gmxcorr_xtc* f, int* natoms, int* step, float* time, float* box, float* x,
                      int max_atoms, float* precision);
int  gmxcorr_xtc_write(gmxcorr_xtc* f, int natoms, int step, float time, const float* box,
                       const float* x, float precision);
void gmxcorr_xtc_close(gmxcorr_xtc* f);

#ifdef __cplusplus
}
#endif

#endif
