/*
 * ksg_oracle.c -- CPU oracle for the all-pairs KSG mutual-information matrix.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference's float64 CPU path.  It is loaded by the test-suite, by the smoke
 * check and by the CPU-baseline leg of the benchmark as the *checker*; the
 * product library never links it and never falls back to it.
 *
 * What it restates (reference file:line):
 *   - per-atom mean-centring of float coordinates promoted to double
 *       src/gmx_correlation.cpp:184-223            -> oracle_center()
 *   - kraskov_prepare: z-score, shift by the minimum, digamma table
 *       lib/kraskov.cpp:2168-2220                  -> oracle_prepare()
 *   - kraskov_wrap + mir_xnyn (KSG algorithm 2, 3-D x 3-D) with the 2-D
 *     wrap-around hash grid and layered neighbour search
 *       lib/kraskov.cpp:2222-2282, 1228-1331       -> oracle_ksg_pair_grid()
 *       lib/kraskov.cpp:56-91   (grid build)       -> grid_build()
 *       lib/kraskov.cpp:257-343 (exact K-NN)       -> grid_knn()
 *       lib/kraskov.cpp:113-177 (marginal count)   -> grid_count()
 *   - the OpenMP pair loop with the 2000 diagonal sentinel
 *       lib/correlation_core.cpp:324-459           -> oracle_ksg_matrix()
 *   - pearsify  lib/correlation_core.cpp:461-469   -> oracle_pearsify()
 *
 * The grid restatement keeps the reference's traversal order (cells visited
 * layer by layer, points inside a cell in ascending frame index) because that
 * order decides which of several equidistant points enter the K-neighbour set.
 * oracle_ksg_pair_brute() is an independent FP64 brute-force recount of the
 * same quantities that reports ties at the K-th joint distance.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Mean-centring and normalisation                                          */
/* ------------------------------------------------------------------------ */

/* coords: float32 [nframes][natoms][3] (frame-major, as frames arrive).
 * traj:   double  [natoms][nframes][3] = coords - per-atom mean (t_traj.x).
 * The mean is a sequential double sum over frames divided by nframes, as in
 * src/gmx_correlation.cpp:200-215. */
void oracle_center(const float *coords, int nframes, int natoms, double *traj)
{
    for (int a = 0; a < natoms; ++a) {
        double mean[3] = {0.0, 0.0, 0.0};
        for (int f = 0; f < nframes; ++f)
            for (int d = 0; d < 3; ++d)
                mean[d] += (double)coords[((size_t)f * natoms + a) * 3 + d];
        for (int d = 0; d < 3; ++d) mean[d] /= nframes;
        for (int f = 0; f < nframes; ++f)
            for (int d = 0; d < 3; ++d)
                traj[((size_t)a * nframes + f) * 3 + d] =
                    (double)coords[((size_t)f * natoms + a) * 3 + d] - mean[d];
    }
}

/* traj: [natoms][N][3]  ->  kx: [3*natoms][N] component rows (c = 3*atom+d),
 * psi: [N+1].  Per component: mean, sample SD with N-1, z-score while tracking
 * the minimum, then subtract the minimum (lib/kraskov.cpp:2197-2210).  The
 * digamma table is psi[0]=0, psi[1]=-0.57721566490153,
 * psi[m+1]=psi[m]+1/m (lib/kraskov.cpp:2216-2219).  A constant component
 * (s == 0) is left to produce the same non-finite values as the reference. */
void oracle_prepare(const double *traj, int natoms, int N, double *kx, double *psi)
{
    const int dim = 3 * natoms;
    for (int c = 0; c < dim; ++c) {
        const int a = c / 3, d = c % 3;
        double *row = kx + (size_t)c * N;
        for (int f = 0; f < N; ++f) row[f] = traj[((size_t)a * N + f) * 3 + d];
        double me = 0.0, s = 0.0;
        for (int f = 0; f < N; ++f) me += row[f];
        me /= N;
        for (int f = 0; f < N; ++f) s += (row[f] - me) * (row[f] - me);
        s /= (N - 1);
        s = sqrt(s);
        /* the reference starts its running minimum at 1e6 (its DBL_MAX/2) */
        double lo = 1000000.0;
        for (int f = 0; f < N; ++f) {
            row[f] = (row[f] - me) / s;
            if (row[f] < lo) lo = row[f];
        }
        for (int f = 0; f < N; ++f) row[f] = row[f] - lo;
    }
    psi[0] = 0.0;
    psi[1] = -0.57721566490153;
    for (int m = 1; m < N; ++m) psi[m + 1] = psi[m] + 1 / (1.0 * m);
}

/* ------------------------------------------------------------------------ */
/* 2-D wrap-around hash grid                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nbox;       /* cells per side, a power of two */
    int *start;     /* [nbox*nbox+1] offsets into pts, cell-major (ix, iy) */
    int *pts;       /* point indices, ascending inside each cell */
} grid2d;

/* Cell of a value for grid construction: (int)(v * inveps) & (nbox-1) with
 * the integer inveps = nbox/4 (lib/kraskov.cpp:69-70, 1260-1263). */
static inline int build_cell(double v, int inveps, int mask)
{
    return ((int)(v * inveps)) & mask;
}

/* Points grouped by cell.  Inside a cell the reference's linked list, after
 * its reorder-and-rebuild step, yields ascending original index, which is
 * what a stable counting sort gives here. */
static void grid_build(grid2d *g, const double *u, const double *v, int N, int nbox)
{
    const int ncell = nbox * nbox, mask = nbox - 1, inveps = nbox / 4;
    g->nbox = nbox;
    g->start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    g->pts = (int *)malloc((size_t)N * sizeof(int));
    int *cellof = (int *)malloc((size_t)N * sizeof(int));
    for (int i = 0; i < N; ++i) {
        cellof[i] = build_cell(u[i], inveps, mask) * nbox + build_cell(v[i], inveps, mask);
        g->start[cellof[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) g->start[c + 1] += g->start[c];
    int *fill = (int *)malloc((size_t)ncell * sizeof(int));
    memcpy(fill, g->start, (size_t)ncell * sizeof(int));
    for (int i = 0; i < N; ++i) g->pts[fill[cellof[i]]++] = i;
    free(fill);
    free(cellof);
}

static void grid_free(grid2d *g)
{
    free(g->start);
    free(g->pts);
}

/* Visit the cells of one search layer in the reference's order and call
 * visit(cell) for each.  Layer jj (< nbox/2) is the square ring at Chebyshev
 * cell distance jj: first the two horizontal edges (x from ix-jj..ix+jj,
 * y in {iy-jj, iy+jj}), then the two vertical edges without corners.  The
 * half layer at jj == nbox/2 covers the cells not yet visited on the torus.
 * lib/kraskov.cpp:273-307 (full layers), 310-341 (half layer). */
#define LAYER_CELLS(ix, iy, jj, mask, nbox, BODY)                                   \
    do {                                                                            \
        const int step_ = (jj) ? 2 * (jj) : 1;                                      \
        for (int x1_ = (ix) - (jj); x1_ <= (ix) + (jj); ++x1_)                      \
            for (int y1_ = (iy) - (jj); y1_ <= (iy) + (jj); y1_ += step_) {         \
                const int cell_ = (x1_ & (mask)) * (nbox) + (y1_ & (mask));         \
                BODY                                                                \
            }                                                                       \
        for (int x1_ = (ix) - (jj); x1_ <= (ix) + (jj); x1_ += step_)               \
            for (int y1_ = (iy) - (jj) + 1; y1_ <= (iy) + (jj) - 1; ++y1_) {        \
                const int cell_ = (x1_ & (mask)) * (nbox) + (y1_ & (mask));         \
                BODY                                                                \
            }                                                                       \
    } while (0)

#define HALF_LAYER_CELLS(ix, iy, jj, mask, nbox, BODY)                              \
    do {                                                                            \
        for (int x1_ = (ix) - (jj); x1_ < (ix) + (jj); ++x1_) {                     \
            const int cell_ = (x1_ & (mask)) * (nbox) + (((iy) - (jj)) & (mask));   \
            BODY                                                                    \
        }                                                                           \
        for (int y1_ = (iy) - (jj) + 1; y1_ <= (iy) + (jj) - 1; ++y1_) {            \
            const int cell_ = ((((ix) - (jj))) & (mask)) * (nbox) + (y1_ & (mask)); \
            BODY                                                                    \
        }                                                                           \
    } while (0)

/* Max-norm distance between frames i and j over `dim` component rows. */
static inline double maxnorm(const double *const *rows, int dim, int i, int j)
{
    double dd = fabs(rows[0][i] - rows[0][j]);
    for (int d = 1; d < dim; ++d) {
        const double t = fabs(rows[d][i] - rows[d][j]);
        if (t > dd) dd = t;
    }
    return dd;
}

/* Exact K nearest neighbours of frame q (self excluded) in the max-norm over
 * `dim` rows, searching the grid built on rows gu/gv.  Sorted insertion with
 * a strict '<' keeps, among equidistant points, the one visited first.
 * nn[1..K] receives the neighbour indices, dn[1..K] their distances.
 * lib/kraskov.cpp:257-343. */
static void grid_knn(const grid2d *g, const double *const *rows, int dim, int gu, int gv,
                     int q, int K, int *nn, double *dn)
{
    const int nbox = g->nbox, mask = nbox - 1;
    const double cellw = 4.0 / nbox;
    dn[0] = 0.0;
    for (int k = 1; k <= K; ++k) dn[k] = 1e30;
    const int ix = ((int)(rows[gu][q] / cellw)) & mask;
    const int iy = ((int)(rows[gv][q] / cellw)) & mask;

#define KNN_VISIT                                                            \
    for (int p_ = g->start[cell_]; p_ < g->start[cell_ + 1]; ++p_) {         \
        const int el_ = g->pts[p_];                                          \
        if (el_ == q) continue;                                              \
        const double dd_ = maxnorm(rows, dim, q, el_);                       \
        if (dd_ < dn[K]) {                                                   \
            int k_ = K;                                                      \
            while (dd_ < dn[k_]) {                                           \
                if (k_ < K) { dn[k_ + 1] = dn[k_]; nn[k_ + 1] = nn[k_]; }    \
                k_--;                                                        \
            }                                                                \
            dn[k_ + 1] = dd_;                                                \
            nn[k_ + 1] = el_;                                                \
        }                                                                    \
    }

    int jj = 0;
    while (dn[K] > cellw * (jj - 1)) {
        LAYER_CELLS(ix, iy, jj, mask, nbox, KNN_VISIT);
        jj++;
        if (jj == nbox / 2) break;
    }
    if (jj == nbox / 2) HALF_LAYER_CELLS(ix, iy, jj, mask, nbox, KNN_VISIT);
#undef KNN_VISIT
}

/* Number of frames j != q with max-norm distance <= eps over `dim` rows.
 * Duplicates of q count.  lib/kraskov.cpp:113-177 (returns nx-1 for self). */
static int grid_count(const grid2d *g, const double *const *rows, int dim, int gu, int gv,
                      int q, double eps)
{
    const int nbox = g->nbox, mask = nbox - 1;
    const double cellw = 4.0 / nbox;
    const int ix = ((int)(rows[gu][q] / cellw)) & mask;
    const int iy = ((int)(rows[gv][q] / cellw)) & mask;
    int n = 0;
#define CNT_VISIT                                                            \
    for (int p_ = g->start[cell_]; p_ < g->start[cell_ + 1]; ++p_)           \
        if (maxnorm(rows, dim, q, g->pts[p_]) <= eps) n++;

    int jj = 0;
    while (eps > cellw * (jj - 1)) {
        LAYER_CELLS(ix, iy, jj, mask, nbox, CNT_VISIT);
        jj++;
        if (jj == nbox / 2) break;
    }
    if (jj == nbox / 2) HALF_LAYER_CELLS(ix, iy, jj, mask, nbox, CNT_VISIT);
#undef CNT_VISIT
    return n - 1;
}

/* Grid size: smallest power of two with 0.5*B*B*K >= N (lib/kraskov.cpp:1260). */
static int grid_side(int N, int K)
{
    int b = 1;
    while (0.5 * b * b * K < N) b *= 2;
    return b;
}

/* ------------------------------------------------------------------------ */
/* KSG-2 for one atom pair                                                  */
/* ------------------------------------------------------------------------ */

/* MI (nats) of atoms a1, a2 from the normalised rows kx[3*natoms][N].
 * Per frame i (original order): the K nearest in the joint 6-D max-norm,
 * epsx/epsy = largest marginal distance among them, nx/ny = marginal counts
 * within those radii, sum psi[nx]+psi[ny]; MI = psi[N] + (psi[K]-1/K) - sum/N.
 * Optional per-frame outputs (may be NULL): nx, ny, epsx, epsy, epsk.
 * lib/kraskov.cpp:1228-1331 (mir_xnyn), 2222-2282 (kraskov_wrap). */
double oracle_ksg_pair_grid(const double *kx, int N, int a1, int a2, int K, const double *psi,
                            int *nx_out, int *ny_out, double *epsx_out, double *epsy_out,
                            double *epsk_out)
{
    const double *rows[6];
    for (int d = 0; d < 3; ++d) {
        rows[d] = kx + (size_t)(3 * a1 + d) * N;
        rows[3 + d] = kx + (size_t)(3 * a2 + d) * N;
    }
    const int nbox = grid_side(N, K);
    grid2d gj, gx, gy;
    grid_build(&gj, rows[0], rows[3], N, nbox);   /* joint grid: comps 0 and 3 */
    grid_build(&gx, rows[0], rows[2], N, nbox);   /* x marginal: comps 0 and 2 */
    grid_build(&gy, rows[3], rows[5], N, nbox);   /* y marginal: comps 3 and 5 */

    int *nn = (int *)malloc((size_t)(K + 1) * sizeof(int));
    double *dn = (double *)malloc((size_t)(K + 1) * sizeof(double));
    const double phiK = psi[K] - 1 / (1.0 * K);
    double acc = 0.0;
    for (int i = 0; i < N; ++i) {
        grid_knn(&gj, rows, 6, 0, 3, i, K, nn, dn);
        double ex = 0.0, ey = 0.0;
        for (int d = 0; d < 3; ++d)
            for (int k = 1; k <= K; ++k) {
                const double t = fabs(rows[d][i] - rows[d][nn[k]]);
                if (t > ex) ex = t;
            }
        for (int d = 3; d < 6; ++d)
            for (int k = 1; k <= K; ++k) {
                const double t = fabs(rows[d][i] - rows[d][nn[k]]);
                if (t > ey) ey = t;
            }
        const int nx = grid_count(&gx, rows, 3, 0, 2, i, ex);
        const int ny = grid_count(&gy, rows + 3, 3, 0, 2, i, ey);
        acc += psi[nx] + psi[ny];
        if (nx_out) nx_out[i] = nx;
        if (ny_out) ny_out[i] = ny;
        if (epsx_out) epsx_out[i] = ex;
        if (epsy_out) epsy_out[i] = ey;
        if (epsk_out) epsk_out[i] = dn[K];
    }
    acc /= N;
    free(nn);
    free(dn);
    grid_free(&gj);
    grid_free(&gx);
    grid_free(&gy);
    return psi[N] + phiK - acc;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* Independent FP64 brute-force recount.  For each frame i: eps_k = K-th
 * smallest joint distance over j != i; epsx/epsy = largest marginal distance
 * over ALL j with joint distance <= eps_k; nx/ny = #{j != i: marginal <= eps}.
 * When more than K points lie within eps_k the K-set is ambiguous; such
 * frames are counted in *nties and flagged in tie_out[i] (may be NULL).
 * Without ties this equals the grid result exactly. */
double oracle_ksg_pair_brute(const double *kx, int N, int a1, int a2, int K, const double *psi,
                             int *nx_out, int *ny_out, double *epsx_out, double *epsy_out,
                             double *epsk_out, unsigned char *tie_out, int *nties)
{
    const double *rows[6];
    for (int d = 0; d < 3; ++d) {
        rows[d] = kx + (size_t)(3 * a1 + d) * N;
        rows[3 + d] = kx + (size_t)(3 * a2 + d) * N;
    }
    double *dA = (double *)malloc((size_t)N * sizeof(double));
    double *dB = (double *)malloc((size_t)N * sizeof(double));
    double *dJ = (double *)malloc((size_t)N * sizeof(double));
    double *srt = (double *)malloc((size_t)N * sizeof(double));
    double acc = 0.0;
    int ties = 0;
    for (int i = 0; i < N; ++i) {
        int m = 0;
        for (int j = 0; j < N; ++j) {
            dA[j] = maxnorm(rows, 3, i, j);
            dB[j] = maxnorm(rows + 3, 3, i, j);
            dJ[j] = dA[j] > dB[j] ? dA[j] : dB[j];
            if (j != i) srt[m++] = dJ[j];
        }
        qsort(srt, (size_t)m, sizeof(double), cmp_double);
        const double ek = srt[K - 1];
        int inside = 0;
        double ex = 0.0, ey = 0.0;
        for (int j = 0; j < N; ++j) {
            if (j == i || dJ[j] > ek) continue;
            inside++;
            if (dA[j] > ex) ex = dA[j];
            if (dB[j] > ey) ey = dB[j];
        }
        int nx = 0, ny = 0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            nx += dA[j] <= ex;
            ny += dB[j] <= ey;
        }
        const int tie = inside > K;
        ties += tie;
        acc += psi[nx] + psi[ny];
        if (nx_out) nx_out[i] = nx;
        if (ny_out) ny_out[i] = ny;
        if (epsx_out) epsx_out[i] = ex;
        if (epsy_out) epsy_out[i] = ey;
        if (epsk_out) epsk_out[i] = ek;
        if (tie_out) tie_out[i] = (unsigned char)tie;
    }
    acc /= N;
    free(dA);
    free(dB);
    free(dJ);
    free(srt);
    if (nties) *nties = ties;
    return psi[N] + (psi[K] - 1 / (1.0 * K)) - acc;
}

/* Full matrix: diagonal 2000 sentinel, symmetric MI, pairs a1 > a2 shared out
 * to OpenMP threads with schedule(dynamic,1) over a1 as in
 * lib/correlation_core.cpp:383-439.  nthreads <= 0 uses all threads.
 * method 0 = grid restatement, 1 = brute-force recount. */
void oracle_ksg_matrix(const double *kx, int natoms, int N, int K, const double *psi,
                       double *mat, int nthreads, int method)
{
    for (int i = 0; i < natoms * natoms; ++i) mat[i] = 0.0;
    for (int i = 0; i < natoms; ++i) mat[(size_t)i * natoms + i] = 2000.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int a1 = 1; a1 < natoms; ++a1)
        for (int a2 = 0; a2 < a1; ++a2) {
            const double v = method == 0
                ? oracle_ksg_pair_grid(kx, N, a1, a2, K, psi, NULL, NULL, NULL, NULL, NULL)
                : oracle_ksg_pair_brute(kx, N, a1, a2, K, psi, NULL, NULL, NULL, NULL, NULL,
                                        NULL, NULL);
            mat[(size_t)a1 * natoms + a2] = v;
            mat[(size_t)a2 * natoms + a1] = v;
        }
}

/* A list of pairs (a1[p], a2[p]) -> out[p]; used to time bounded samples. */
void oracle_ksg_pairs(const double *kx, int N, int K, const double *psi, int npairs,
                      const int *a1, const int *a2, double *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int p = 0; p < npairs; ++p)
        out[p] = oracle_ksg_pair_grid(kx, N, a1[p], a2[p], K, psi, NULL, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Timed CPU baseline: the same KSG-2 pair with the reference's memory       */
/* layout and allocation pattern                                            */
/* ------------------------------------------------------------------------ */

/* The grid restatement above groups frame indices by cell and reads the
 * coordinates in place; it runs about twice as fast as the reference's own
 * code.  The baseline below restates the reference's cost structure as well
 * as its arithmetic, so that its time stands for the reference's:
 *   - each pair copies its six rows, and each marginal its three, into fresh
 *     allocations (lib/kraskov.cpp:2253-2262, 1282-1287);
 *   - each grid physically reorders its copy into cell order and chains the
 *     points of a cell in a linked list, built, reordered and rebuilt as in
 *     lib/kraskov.cpp:56-91; cells are rows of a pointer array;
 *   - every neighbour query allocates its scratch (lib/kraskov.cpp:264-268,
 *     120-121), and the marginal count stops evaluating components once one
 *     exceeds the radius (lib/kraskov.cpp:130-134);
 *   - the grid build, the K-nearest search and the count are out-of-line
 *     functions that take the dimension at run time, as the reference's
 *     make_box2ind, neiK and neiE are for every estimator of lib/kraskov.cpp:
 *     noipa keeps gcc from specialising them for dim 6 and 3 (unrolled
 *     component loops), which alone made this restatement 15-25% faster than
 *     the reference's compiled code.
 * The traversal order, and hence the neighbour set and the digamma sum, is
 * the grid restatement's: both return the same MI bit for bit. */

#define BASELINE_OUT_OF_LINE __attribute__((noipa))

typedef struct {
    int bs;
    int **head;   /* [bs][bs] first point of each cell, -1 when empty */
    int *next;    /* [N] next point of the same cell, -1 at the end */
} listgrid;

static int lg_cell(double v, int inveps, int mask) { return ((int)(v * inveps)) & mask; }

/* Reorders x[0..dim) in place into cell order of components c1, c2 and
 * builds the lists; ind[orig] = new index. */
BASELINE_OUT_OF_LINE static void lg_build(double **x, int dim, int N, int c1, int c2, int bs, int inveps, int *ind,
                     listgrid *g)
{
    const int mask = bs - 1;
    g->bs = bs;
    g->head = (int **)calloc((size_t)bs, sizeof(int *));
    for (int r = 0; r < bs; ++r) g->head[r] = (int *)calloc((size_t)bs, sizeof(int));
    g->next = (int *)calloc((size_t)N, sizeof(int));
    double **tmp = (double **)calloc((size_t)dim, sizeof(double *));
    for (int d = 0; d < dim; ++d) {
        tmp[d] = (double *)calloc((size_t)N, sizeof(double));
        memcpy(tmp[d], x[d], (size_t)N * sizeof(double));
    }
    for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c) g->head[r][c] = -1;
    for (int i = 0; i < N; ++i) {
        const int cx = lg_cell(x[c1][i], inveps, mask), cy = lg_cell(x[c2][i], inveps, mask);
        g->next[i] = g->head[cx][cy];
        g->head[cx][cy] = i;
    }
    int m = -1;
    for (int cx = 0; cx < bs; ++cx)
        for (int cy = 0; cy < bs; ++cy) {
            for (int el = g->head[cx][cy]; el >= 0; el = g->next[el]) {
                ++m;
                for (int d = 0; d < dim; ++d) x[d][m] = tmp[d][el];
                ind[el] = m;
            }
            g->head[cx][cy] = -1;
        }
    for (int i = 0; i < N; ++i) {
        const int cx = lg_cell(x[c1][i], inveps, mask), cy = lg_cell(x[c2][i], inveps, mask);
        g->next[i] = g->head[cx][cy];
        g->head[cx][cy] = i;
    }
    for (int d = 0; d < dim; ++d) free(tmp[d]);
    free(tmp);
}

static void lg_free(listgrid *g)
{
    for (int r = 0; r < g->bs; ++r) free(g->head[r]);
    free(g->head);
    free(g->next);
}

/* K nearest of point q (reordered index) in the max-norm over dim rows. */
BASELINE_OUT_OF_LINE static void lg_knn(double **x, int dim, int c1, int c2, int q, const listgrid *g, double cellw,
                   int K, int *nn)
{
    const int bs = g->bs, mask = bs - 1;
    double *dn = (double *)calloc((size_t)K + 1, sizeof(double));
    double *xq = (double *)calloc((size_t)dim, sizeof(double));
    for (int d = 0; d < dim; ++d) xq[d] = x[d][q];
    dn[0] = 0.0;
    for (int k = 1; k <= K; ++k) dn[k] = 1e30;
    const int ix = ((int)(xq[c1] / cellw)) & mask, iy = ((int)(xq[c2] / cellw)) & mask;
    /* REPEAT1: the reference's vertical-edge loop evaluates component 1 a
     * second time (lib/kraskov.cpp:297); the max is unchanged, the work is
     * kept so that the time is the reference's */
#define LG_KNN_VISIT(REPEAT1)                                                        \
    for (int el = g->head[x1_ & mask][y1_ & mask]; el != -1; el = g->next[el]) {     \
        if (el == q) continue;                                                       \
        double dd = fabs(xq[0] - x[0][el]);                                          \
        for (int d = 1; d < dim; ++d) {                                              \
            const double t = fabs(xq[d] - x[d][el]);                                 \
            if (t > dd) dd = t;                                                      \
        }                                                                            \
        if (REPEAT1) {                                                               \
            const double t = fabs(xq[1] - x[1][el]);                                 \
            if (t > dd) dd = t;                                                      \
        }                                                                            \
        if (dd < dn[K]) {                                                            \
            int k = K;                                                               \
            while (dd < dn[k]) {                                                     \
                if (k < K) { dn[k + 1] = dn[k]; nn[k + 1] = nn[k]; }                 \
                k--;                                                                 \
            }                                                                        \
            dn[k + 1] = dd;                                                          \
            nn[k + 1] = el;                                                          \
        }                                                                            \
    }
    int jj = 0;
    while (dn[K] > cellw * (jj - 1)) {
        const int step = jj ? 2 * jj : 1;
        for (int x1_ = ix - jj; x1_ <= ix + jj; ++x1_)
            for (int y1_ = iy - jj; y1_ <= iy + jj; y1_ += step) { LG_KNN_VISIT(0) }
        for (int x1_ = ix - jj; x1_ <= ix + jj; x1_ += step)
            for (int y1_ = iy - jj + 1; y1_ <= iy + jj - 1; ++y1_) { LG_KNN_VISIT(1) }
        jj++;
        if (jj == bs / 2) break;
    }
    if (jj == bs / 2) {
        for (int x1_ = ix - jj; x1_ < ix + jj; ++x1_) {
            const int y1_ = iy - jj;
            LG_KNN_VISIT(0)
        }
        for (int y1_ = iy - jj + 1; y1_ <= iy + jj - 1; ++y1_) {
            const int x1_ = ix - jj;
            LG_KNN_VISIT(0)
        }
    }
#undef LG_KNN_VISIT
    free(dn);
    free(xq);
}

/* #{j != q : max-norm distance over dim rows <= eps}, duplicates counted. */
BASELINE_OUT_OF_LINE static int lg_count(double **x, int dim, int c1, int c2, int q, const listgrid *g, double cellw,
                    double eps)
{
    const int bs = g->bs, mask = bs - 1;
    double *xq = (double *)calloc((size_t)dim, sizeof(double));
    for (int d = 0; d < dim; ++d) xq[d] = x[d][q];
    const int ix = ((int)(xq[c1] / cellw)) & mask, iy = ((int)(xq[c2] / cellw)) & mask;
    int n = 0;
#define LG_CNT_VISIT                                                                 \
    for (int el = g->head[x1_ & mask][y1_ & mask]; el != -1; el = g->next[el]) {     \
        double dd = fabs(xq[0] - x[0][el]);                                          \
        for (int d = 1; d < dim; ++d) {                                              \
            const double t = fabs(xq[d] - x[d][el]);                                 \
            if (t > dd && (dd = t) > eps) break;                                     \
        }                                                                            \
        if (dd <= eps) n++;                                                          \
    }
    int jj = 0;
    while (eps > cellw * (jj - 1)) {
        const int step = jj ? 2 * jj : 1;
        for (int x1_ = ix - jj; x1_ <= ix + jj; ++x1_)
            for (int y1_ = iy - jj; y1_ <= iy + jj; y1_ += step) { LG_CNT_VISIT }
        for (int x1_ = ix - jj; x1_ <= ix + jj; x1_ += step)
            for (int y1_ = iy - jj + 1; y1_ <= iy + jj - 1; ++y1_) { LG_CNT_VISIT }
        jj++;
        if (jj == bs / 2) break;
    }
    if (jj == bs / 2) {
        for (int x1_ = ix - jj; x1_ < ix + jj; ++x1_) {
            const int y1_ = iy - jj;
            LG_CNT_VISIT
        }
        for (int y1_ = iy - jj + 1; y1_ <= iy + jj - 1; ++y1_) {
            const int x1_ = ix - jj;
            LG_CNT_VISIT
        }
    }
#undef LG_CNT_VISIT
    free(xq);
    return n - 1;
}

double oracle_ksg_pair_baseline(const double *kx, int N, int a1, int a2, int K, const double *psi)
{
    double **x = (double **)calloc(6, sizeof(double *));
    for (int d = 0; d < 3; ++d) {
        x[d] = (double *)calloc((size_t)N, sizeof(double));
        x[3 + d] = (double *)calloc((size_t)N, sizeof(double));
        memcpy(x[d], kx + (size_t)(3 * a1 + d) * N, (size_t)N * sizeof(double));
        memcpy(x[3 + d], kx + (size_t)(3 * a2 + d) * N, (size_t)N * sizeof(double));
    }
    double **xx = (double **)calloc(3, sizeof(double *));
    double **yy = (double **)calloc(3, sizeof(double *));
    for (int d = 0; d < 3; ++d) {
        xx[d] = (double *)calloc((size_t)N, sizeof(double));
        yy[d] = (double *)calloc((size_t)N, sizeof(double));
        memcpy(xx[d], x[d], (size_t)N * sizeof(double));
        memcpy(yy[d], x[3 + d], (size_t)N * sizeof(double));
    }
    const int bs = grid_side(N, K), inveps = bs / 4;
    const double cellw = 4.0 / bs;
    int *ind = (int *)calloc((size_t)N, sizeof(int));
    int *indx = (int *)calloc((size_t)N, sizeof(int));
    int *indy = (int *)calloc((size_t)N, sizeof(int));
    int *nn = (int *)calloc((size_t)K + 1, sizeof(int));
    listgrid gj, gx, gy;
    lg_build(x, 6, N, 0, 3, bs, inveps, ind, &gj);
    lg_build(xx, 3, N, 0, 2, bs, inveps, indx, &gx);
    lg_build(yy, 3, N, 0, 2, bs, inveps, indy, &gy);
    double acc = 0.0, xc[6];
    for (int i = 0; i < N; ++i) {
        const int q = ind[i];
        for (int d = 0; d < 6; ++d) xc[d] = x[d][q];
        lg_knn(x, 6, 0, 3, q, &gj, cellw, K, nn);
        double ex = 0.0, ey = 0.0;
        for (int d = 0; d < 3; ++d)
            for (int k = 1; k <= K; ++k) {
                const double t = fabs(xc[d] - x[d][nn[k]]);
                if (t > ex) ex = t;
            }
        for (int d = 3; d < 6; ++d)
            for (int k = 1; k <= K; ++k) {
                const double t = fabs(xc[d] - x[d][nn[k]]);
                if (t > ey) ey = t;
            }
        const int nx = lg_count(xx, 3, 0, 2, indx[i], &gx, cellw, ex);
        const int ny = lg_count(yy, 3, 0, 2, indy[i], &gy, cellw, ey);
        acc += psi[nx] + psi[ny];
    }
    acc /= N;
    lg_free(&gj);
    lg_free(&gx);
    lg_free(&gy);
    free(ind);
    free(indx);
    free(indy);
    free(nn);
    for (int d = 0; d < 3; ++d) {
        free(xx[d]);
        free(yy[d]);
    }
    free(xx);
    free(yy);
    for (int d = 0; d < 6; ++d) free(x[d]);
    free(x);
    return psi[N] + (psi[K] - 1 / (1.0 * K)) - acc;
}

/* The baseline over a pair list, OpenMP over pairs as the reference's loop
 * (lib/correlation_core.cpp:403, schedule(dynamic, 1)). */
void oracle_ksg_pairs_baseline(const double *kx, int N, int K, const double *psi, int npairs,
                               const int *a1, const int *a2, double *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int p = 0; p < npairs; ++p)
        out[p] = oracle_ksg_pair_baseline(kx, N, a1[p], a2[p], K, psi);
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* r = sqrt(1 - exp(-2/dim * MI)) for MI > 0, else 0, element-wise including
 * the diagonal (2000 -> 1).  lib/correlation_core.cpp:461-469. */
void oracle_pearsify(double *mat, int n, int dim)
{
    for (int i = 0; i < n * n; ++i)
        mat[i] = mat[i] > 0.0 ? sqrt(1.0 - exp(-2.0 / dim * mat[i])) : 0.0;
}
