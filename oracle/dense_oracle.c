/*
 * dense_oracle.c -- CPU oracle for the Gaussian, Gaussian-copula (GCMI) and
 * transfer-entropy matrices.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference's float64
 * CPU code, used by the tests and the benchmark's CPU leg as the checker.
 * The product library never links or calls it.
 *
 *   oracle_gauss_matrix   lib/correlation_core.cpp:49-166, 280-311
 *   oracle_copula         lib/gcmi.cpp:26-79 (erfinv), 95-146 (ranks, midranks)
 *   oracle_gcmi_matrix    lib/gcmi.cpp:152-231, 246-317
 *   oracle_te_matrix      lib/transfer_entropy.cpp:41-221
 *
 * All inputs are the mean-centred trajectory traj[natoms][N][3] (t_traj.x).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Determinant by LU with partial pivoting on a private copy; 0 for a zero
 * pivot column.  lib/correlation_core.cpp:49-98 and lib/gcmi.cpp:152-184
 * perform the same row operations. */
double oracle_lu_det(const double *m_in, int n)
{
    double m[36];
    memcpy(m, m_in, (size_t)n * n * sizeof(double));
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

/* Gaussian differential entropy 0.5*(n(1+log 2pi) + log det), 0 if det<=0.
 * lib/correlation_core.cpp:100-111. */
static double gauss_entropy(const double *cov, int n)
{
    const double det = oracle_lu_det(cov, n);
    if (det <= 0.0) return 0.0;
    return 0.5 * (n * (1.0 + log(2.0 * 3.14159265358979323846)) + log(det));
}

/* Covariance C = X^T X / N of the mean-centred coordinates, C[3n][3n],
 * each entry a sequential sum over frames.  lib/correlation_core.cpp:140-166. */
void oracle_covariance(const double *traj, int natoms, int N, double *cov)
{
    const int nc = 3 * natoms;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < nc; ++i) {
        const int a1 = i / 3, d1 = i % 3;
        for (int j = 0; j < nc; ++j) {
            const int a2 = j / 3, d2 = j % 3;
            double s = 0.0;
            for (int f = 0; f < N; ++f)
                s += traj[((size_t)a1 * N + f) * 3 + d1] * traj[((size_t)a2 * N + f) * 3 + d2];
            cov[(size_t)i * nc + j] = s / N;
        }
    }
}

/* Pair MI from a covariance matrix: H(a1) + H(a2) - H(a1,a2) with the 6x6
 * joint block [[C11, C12], [C12^T, C22]].  lib/correlation_core.cpp:280-311. */
double oracle_gauss_pair(const double *cov, int natoms, int a1, int a2)
{
    const int nc = 3 * natoms;
    double c1[9], c2[9], cj[36];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            c1[r * 3 + c] = cov[(size_t)(3 * a1 + r) * nc + 3 * a1 + c];
            c2[r * 3 + c] = cov[(size_t)(3 * a2 + r) * nc + 3 * a2 + c];
            const double x12 = cov[(size_t)(3 * a1 + r) * nc + 3 * a2 + c];
            cj[r * 6 + c] = c1[r * 3 + c];
            cj[(r + 3) * 6 + (c + 3)] = c2[r * 3 + c];
            cj[r * 6 + (c + 3)] = x12;
            cj[(c + 3) * 6 + r] = x12;
        }
    return gauss_entropy(c1, 3) + gauss_entropy(c2, 3) - gauss_entropy(cj, 6);
}

void oracle_gauss_matrix_from_cov(const double *cov, int natoms, double *mat)
{
    for (int a1 = 0; a1 < natoms; ++a1) {
        mat[(size_t)a1 * natoms + a1] = 2000.0;
        for (int a2 = a1 + 1; a2 < natoms; ++a2) {
            const double mi = oracle_gauss_pair(cov, natoms, a1, a2);
            mat[(size_t)a1 * natoms + a2] = mi;
            mat[(size_t)a2 * natoms + a1] = mi;
        }
    }
}

void oracle_gauss_matrix(const double *traj, int natoms, int N, double *mat)
{
    const int nc = 3 * natoms;
    double *cov = (double *)malloc((size_t)nc * nc * sizeof(double));
    oracle_covariance(traj, natoms, N, cov);
    oracle_gauss_matrix_from_cov(cov, natoms, mat);
    free(cov);
}

/* ------------------------------------------------------------------------ */
/* GCMI                                                                      */
/* ------------------------------------------------------------------------ */

/* Inverse error function: Acklam's published rational approximation of the
 * inverse normal CDF at p=(x+1)/2, divided by sqrt 2, then one Newton step on
 * erf (Acklam 2003; lib/gcmi.cpp:26-79 uses the same published method). */
double oracle_erfinv(double x)
{
    if (x >= 1.0) return 1e30;
    if (x <= -1.0) return -1e30;
    if (x == 0.0) return 0.0;
    static const double A[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double B[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double C[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                -2.400758277161838e+00, -2.549732539343734e+00,
                                4.374664141464968e+00, 2.938163982698783e+00};
    static const double D[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    const double p = (x + 1.0) * 0.5;
    double r;
    if (p < 0.02425) {
        const double q = sqrt(-2.0 * log(p));
        r = (((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5]) /
            ((((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0);
    } else if (p <= 0.97575) {
        const double q = p - 0.5, s = q * q;
        r = (((((A[0] * s + A[1]) * s + A[2]) * s + A[3]) * s + A[4]) * s + A[5]) * q /
            (((((B[0] * s + B[1]) * s + B[2]) * s + B[3]) * s + B[4]) * s + 1.0);
    } else {
        const double q = sqrt(-2.0 * log(1.0 - p));
        r = -(((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5]) /
            ((((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0);
    }
    r /= sqrt(2.0);
    r -= (erf(r) - x) * sqrt(3.14159265358979323846) * 0.5 * exp(r * r);
    return r;
}

typedef struct { double v; int idx; } vi_t;
static int cmp_vi(const void *a, const void *b)
{
    const double x = ((const vi_t *)a)->v, y = ((const vi_t *)b)->v;
    return (x > y) - (x < y);
}

/* Van der Waerden copula scores z[comp][N] (comp = 3*atom+d): frames sorted
 * by value, rank t -> sqrt(2)*erfinv(2(t+0.5)/N - 1), tied runs receive the
 * mean score of their rank range.  lib/gcmi.cpp:95-146. */
void oracle_copula(const double *traj, int natoms, int N, double *z)
{
    const int nc = 3 * natoms;
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        vi_t *buf = (vi_t *)malloc((size_t)N * sizeof(vi_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int c = 0; c < nc; ++c) {
            const int a = c / 3, d = c % 3;
            for (int f = 0; f < N; ++f) {
                buf[f].v = traj[((size_t)a * N + f) * 3 + d];
                buf[f].idx = f;
            }
            qsort(buf, (size_t)N, sizeof(vi_t), cmp_vi);
            int r = 0;
            while (r < N) {
                int s = r + 1;
                while (s < N && buf[s].v == buf[r].v) ++s;
                double avg = 0.0;
                for (int t = r; t < s; ++t) {
                    const double p = (t + 0.5) / N;
                    avg += sqrt(2.0) * oracle_erfinv(2.0 * p - 1.0);
                }
                avg /= (s - r);
                for (int t = r; t < s; ++t) z[(size_t)c * N + buf[t].idx] = avg;
                r = s;
            }
        }
        free(buf);
    }
}

/* MI of one pair from copula scores: 6x6 second-moment matrix (no mean
 * subtraction), 0.5*(log d1 + log d2 - log dJ), 0 if any det <= 0.
 * lib/gcmi.cpp:190-231. */
double oracle_gcmi_pair(const double *z, int N, int a1, int a2)
{
    double m[36] = {0};
    const double invN = 1.0 / N;
    for (int f = 0; f < N; ++f) {
        double v[6];
        for (int d = 0; d < 3; ++d) {
            v[d] = z[(size_t)(3 * a1 + d) * N + f];
            v[d + 3] = z[(size_t)(3 * a2 + d) * N + f];
        }
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) m[r * 6 + c] += v[r] * v[c];
    }
    for (int i = 0; i < 36; ++i) m[i] *= invN;
    double c1[9], c2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            c1[r * 3 + c] = m[r * 6 + c];
            c2[r * 3 + c] = m[(r + 3) * 6 + c + 3];
        }
    const double d1 = oracle_lu_det(c1, 3), d2 = oracle_lu_det(c2, 3), dj = oracle_lu_det(m, 6);
    if (d1 <= 0.0 || d2 <= 0.0 || dj <= 0.0) return 0.0;
    return 0.5 * (log(d1) + log(d2) - log(dj));
}

/* Returns 0, or -1 for natoms < 2 / N < 6 (the reference throws there). */
int oracle_gcmi_matrix(const double *traj, int natoms, int N, double *mat, int nthreads)
{
    if (natoms < 2 || N < 6) return -1;
    memset(mat, 0, (size_t)natoms * natoms * sizeof(double));
    double *z = (double *)malloc((size_t)3 * natoms * N * sizeof(double));
    oracle_copula(traj, natoms, N, z);
    for (int i = 0; i < natoms; ++i) mat[(size_t)i * natoms + i] = 2000.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int a1 = 1; a1 < natoms; ++a1)
        for (int a2 = 0; a2 < a1; ++a2) {
            const double mi = oracle_gcmi_pair(z, N, a1, a2);
            mat[(size_t)a1 * natoms + a2] = mi;
            mat[(size_t)a2 * natoms + a1] = mi;
        }
    free(z);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Transfer entropy (Frenzel-Pompe conditional MI)                           */
/* ------------------------------------------------------------------------ */

/* TE(a2 -> a1) at `lag` on the raw mean-centred coordinates, for one ordered
 * pair.  Query q = [X_a1(t+lag), X_a1(t), X_a2(t)] in the 9-D max-norm over
 * t < Nv = N - lag; eps = k-th smallest distance (self excluded); strict '<'
 * counts in (Xf,Xp), (Xp,Yp) and Xp; mean of
 * psi[k] + psi[n_xp+1] - psi[n_xfxp+1] - psi[n_ypxp+1] with the digamma table
 * of size Nv+1.  lib/transfer_entropy.cpp:41-48, 114-204. */
double oracle_te_pair(const double *traj, int N, int a1, int a2, int k, int lag, const double *psi)
{
    const int Nv = N - lag;
    const double *X = traj + (size_t)a1 * N * 3, *Y = traj + (size_t)a2 * N * 3;
    double *dist = (double *)malloc((size_t)Nv * sizeof(double));
    double *buf = (double *)malloc((size_t)k * sizeof(double));
    double sum = 0.0;
    for (int qi = 0; qi < Nv; ++qi) {
        for (int j = 0; j < Nv; ++j) {
            double d9 = 0.0;
            for (int d = 0; d < 3; ++d) {
                d9 = fmax(d9, fabs(X[(qi + lag) * 3 + d] - X[(j + lag) * 3 + d]));
                d9 = fmax(d9, fabs(X[qi * 3 + d] - X[j * 3 + d]));
                d9 = fmax(d9, fabs(Y[qi * 3 + d] - Y[j * 3 + d]));
            }
            dist[j] = d9;
        }
        /* k smallest via a max-buffer: eps = its maximum */
        for (int t = 0; t < k; ++t) buf[t] = 1e30;
        double bmax = 1e30;
        int bidx = 0;
        for (int j = 0; j < Nv; ++j) {
            if (j == qi) continue;
            if (dist[j] < bmax) {
                buf[bidx] = dist[j];
                bmax = buf[0];
                bidx = 0;
                for (int t = 1; t < k; ++t)
                    if (buf[t] > bmax) { bmax = buf[t]; bidx = t; }
            }
        }
        const double eps = bmax;
        int nxfxp = 0, nypxp = 0, nxp = 0;
        for (int j = 0; j < Nv; ++j) {
            if (j == qi) continue;
            double dxf = 0.0, dxp = 0.0, dyp = 0.0;
            for (int d = 0; d < 3; ++d) {
                dxf = fmax(dxf, fabs(X[(qi + lag) * 3 + d] - X[(j + lag) * 3 + d]));
                dxp = fmax(dxp, fabs(X[qi * 3 + d] - X[j * 3 + d]));
                dyp = fmax(dyp, fabs(Y[qi * 3 + d] - Y[j * 3 + d]));
            }
            nxfxp += fmax(dxf, dxp) < eps;
            nypxp += fmax(dyp, dxp) < eps;
            nxp += dxp < eps;
        }
        if (nxfxp > Nv) nxfxp = Nv;
        if (nypxp > Nv) nypxp = Nv;
        if (nxp > Nv) nxp = Nv;
        sum += psi[k] + psi[nxp + 1] - psi[nxfxp + 1] - psi[nypxp + 1];
    }
    free(dist);
    free(buf);
    return sum / Nv;
}

/* Digamma table of size n+1 as lib/transfer_entropy.cpp:41-48. */
void oracle_te_psi(int n, double *psi)
{
    for (int i = 0; i <= n; ++i) psi[i] = 0.0;
    psi[1] = -0.57721566490153;
    for (int i = 1; i < n; ++i) psi[i + 1] = psi[i] + 1.0 / i;
}

/* mat[a1*n+a2] = TE(a2 -> a1); diagonal 0; all zero when N-lag <= k. */
void oracle_te_matrix(const double *traj, int natoms, int N, int k, int lag, double *mat,
                      int nthreads)
{
    memset(mat, 0, (size_t)natoms * natoms * sizeof(double));
    const int Nv = N - lag;
    if (Nv <= k) return;
    double *psi = (double *)malloc((size_t)(Nv + 1) * sizeof(double));
    oracle_te_psi(Nv, psi);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
#endif
    for (int a1 = 0; a1 < natoms; ++a1)
        for (int a2 = 0; a2 < natoms; ++a2) {
            if (a1 == a2) continue;
            mat[(size_t)a1 * natoms + a2] = oracle_te_pair(traj, N, a1, a2, k, lag, psi);
        }
    free(psi);
}
