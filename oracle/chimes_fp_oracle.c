/*
 * ORACLE -- test infrastructure only.
 *
 * Plain-C, CPU, FP64 restatement of the cluster fingerprint histogram of the
 * reference (chimesFF/src/FP/calc_histograms.cpp) on one process:
 *   oc_fp_edges  read_flat_clusters (140-249): each dumped cluster row
 *                (pair distances, then member types) becomes its sorted
 *                Morse-transformed edge lengths, with the cutoffs and Morse
 *                lambdas of the cluster's pairs;
 *   oc_fp_hist   gen_flat_hists (351-489) with get_bin (251-259): the
 *                histogram of the distances between every cluster of one set
 *                and every cluster of another (or every later cluster of the
 *                same set).
 * In gen_flat_hists the edge and type vectors handed to
 * compute_composition_weight_{2,3,4}b are built with n zeros and then
 * appended to, so those functions read only the zeros and every
 * composition weight is 0: the cluster distance is the Euclidean distance of
 * the sorted edges.  This restatement keeps that behaviour.
 *
 * Parity pins (tests/test_fp_oracle.py): the transform's closed forms (the
 * inner cutoff maps to 1, the outer to -1, monotone in between) at every
 * cluster slot including special cutoffs; the histogram against a
 * vectorised brute force (full distance matrix, floor, bincount); exact
 * sample counts n1 n2 and n (n - 1) / 2.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chimes_oracle.h"

/* transform (calc_histograms.cpp:126-134) */
static double fp_transform(double rcin, double rcout, double lambda, double rij)
{
    double x_min = exp(-1 * rcin / lambda);
    double x_max = exp(-1 * rcout / lambda);
    double x_avg = 0.5 * (x_max + x_min);
    double x_diff = 0.5 * (x_max - x_min);
    x_diff *= -1.0;
    return (exp(-1 * rij / lambda) - x_avg) / x_diff;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int oc_fp_edges(const oc_params *P, int nb, long n, const double *dist, const int *types, double *edges)
{
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    static const int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    const int np = nb * (nb - 1) / 2;
    for (long c = 0; c < n; c++) {
        const int *t = types + c * nb;
        const double *d = dist + c * np;
        double *e = edges + c * np;
        for (int p = 0; p < np; p++) {
            const int a = nb == 3 ? pa3[p] : pa[p], b = nb == 3 ? pb3[p] : pb[p];
            double rin, rout, lam;
            if (oc_cluster_pair_cutoffs(P, nb, t, p, &rin, &rout) != 0) return 1;
            lam = oc_morse_lambda(P, t[a], t[b]);
            e[p] = fp_transform(rin, rout, lam, d[p]);
        }
        qsort(e, (size_t)np, sizeof(double), cmp_double);
    }
    return 0;
}

/* get_bin (calc_histograms.cpp:251-259) */
static int fp_bin(double binw, double maxval, double dist)
{
    int bin = (int)floor(dist / binw);
    if (dist == maxval) return bin - 1;
    return bin;
}

int oc_fp_hist(int np, long n1, const double *e1, long n2, const double *e2, int same, int nbin, double binw,
               double maxd, long long *counts, long long *nsamples)
{
    memset(counts, 0, sizeof(long long) * nbin);
    *nsamples = 0;
    for (long i = 0; i < n1; i++) {
        for (long j = same ? i + 1 : 0; j < n2; j++) {
            double s = 0;
            for (int k = 0; k < np; k++) s += pow(e1[i * np + k] - e2[j * np + k], 2.0);
            const double total = sqrt(s) + fabs(0.0 - 0.0);
            const int bin = fp_bin(binw, maxd, total);
            if (bin < 0 || bin >= nbin) return 1;
            counts[bin]++;
            (*nsamples)++;
        }
    }
    return 0;
}
