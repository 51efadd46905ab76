/*
 * ORACLE (test infrastructure only) -- FP64 band LU with partial pivoting,
 * restating the published LAPACK algorithm DGBTF2 / DGBTRS (unblocked,
 * row interchanges within the band, fill-in up to kl extra super-diagonals).
 *
 * The reference hands its tangent to MKL PARDISO (src/SparseMatrix.cpp:74-166);
 * any backward-stable direct solve of the same FP64 system agrees with it to
 * rounding.  This one is pinned in tests/test_oracle_linalg.py against
 * numpy.linalg.solve and scipy.linalg.solve_banded on random band systems
 * (including ones that need row interchanges).
 */
#include <math.h>
#include "gfo_internal.h"

#define AB(i, j) ab[(size_t)(kv + (i) - (j)) + (size_t)(j) * (size_t)ldab]

int gfo_gbtrf(int n, int kl, int ku, double *ab, int ldab, int *ipiv)
{
    int kv = ku + kl;
    int info = 0;
    /* zero the fill-in rows of the first kv columns */
    for (int j = ku + 1; j < kv && j < n; j++)
        for (int i = kv - j; i < kl; i++) ab[(size_t)i + (size_t)j * ldab] = 0.0;
    int ju = 0;
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = 0; i < kl; i++) ab[(size_t)i + (size_t)(j + kv) * ldab] = 0.0;
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0;
        double best = fabs(AB(j, j));
        for (int i = 1; i <= km; i++) {
            double v = fabs(AB(j + i, j));
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = j + jp;
        if (AB(j + jp, j) != 0.0) {
            int cand = j + ku + jp;
            if (cand > n - 1) cand = n - 1;
            if (cand > ju) ju = cand;
            if (jp != 0)
                for (int c = j; c <= ju; c++) {
                    double t = AB(j + jp, c);
                    AB(j + jp, c) = AB(j, c);
                    AB(j, c) = t;
                }
            if (km > 0) {
                double r = 1.0 / AB(j, j);
                for (int i = 1; i <= km; i++) AB(j + i, j) *= r;
                for (int c = j + 1; c <= ju; c++) {
                    double u = AB(j, c);
                    if (u != 0.0)
                        for (int i = 1; i <= km; i++) AB(j + i, c) -= AB(j + i, j) * u;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
    return info;
}

void gfo_gbtrs(int n, int kl, int ku, const double *ab, int ldab, const int *ipiv, double *b)
{
    int kv = ku + kl;
    /* L solve with the row interchanges */
    if (kl > 0)
        for (int j = 0; j < n - 1; j++) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = ipiv[j];
            if (l != j) { double t = b[l]; b[l] = b[j]; b[j] = t; }
            double bj = b[j];
            for (int i = 1; i <= lm; i++) b[j + i] -= AB(j + i, j) * bj;
        }
    /* U solve, U has kv super-diagonals */
    for (int j = n - 1; j >= 0; j--) {
        if (b[j] != 0.0) {
            b[j] /= AB(j, j);
            double t = b[j];
            int i0 = j - kv > 0 ? j - kv : 0;
            for (int i = i0; i < j; i++) b[i] -= AB(i, j) * t;
        }
    }
}
