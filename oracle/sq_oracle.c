/*
 * sq_oracle.c -- CPU FP64 restatement of SQMiniApp's spectral-quadrature hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
 * entry of __graft_entry__.py and the cpu_baseline leg of bench.py load it.  The
 * product (sqminiapp_amd/, libsqamd.so) never links, loads or calls it.
 *
 * Every function follows the reference file:line cited above it, in the same
 * order of floating-point operations where that order is visible in the
 * reference source.  Where the reference delegates arithmetic to PETSc/BLAS
 * (VecDot, VecNorm, the SeqAIJ MatMult row sum) the oracle uses the plain
 * sequential sum; the PETSc version is unpinned, so that rounding is
 * "parity unpinned" at the ulp level (see the notes next to each use).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "sq_oracle_internal.h"

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }

/* ------------------------------------------------------------------ tools */

/* fract(n,k) = (n!)^2/((n-k)!(n+k)!)   -- lsdft_tools.cc:519-531 */
double sqo_fract(int n, int k)
{
    double Nr = 1, Dr = 1;
    for (int i = n - k + 1; i <= n; i++) Nr *= i;
    for (int i = n + 1; i <= n + k; i++) Dr *= i;
    return Nr / Dr;
}

/* tridiag_gen: Gauss elimination, lsdft_tools.cc:122-151 */
static int tridiag_gen(double *A, double *B, double *C, double *D, int len)
{
    double b, *F = dalloc(len);
    b = B[0];
    D[0] = D[0] / b;
    for (int i = 1; i < len; i++) {
        F[i] = C[i - 1] / b;
        b = B[i] - A[i] * F[i];
        if (b == 0) { free(F); return SQO_ERR_RANGE; }
        D[i] = (D[i] - D[i - 1] * A[i]) / b;
    }
    for (int i = len - 2; i >= 0; i--) D[i] -= (D[i + 1] * F[i + 1]);
    free(F);
    return SQO_OK;
}

/* getYD_gen: not-a-knot cubic-spline slopes (Wolberg), lsdft_tools.cc:75-115 */
int sqo_spline_yd(const double *X, const double *Y, double *YD, int len)
{
    int i;
    double h0, h1, r0, r1;
    double *A = dalloc(len), *B = dalloc(len), *C = dalloc(len);
    h0 = X[1] - X[0];
    h1 = X[2] - X[1];
    r0 = (Y[1] - Y[0]) / h0;
    r1 = (Y[2] - Y[1]) / h1;
    B[0] = h1 * (h0 + h1);
    C[0] = (h0 + h1) * (h0 + h1);
    YD[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
    for (i = 1; i < len - 1; i++) {
        h0 = X[i] - X[i - 1];
        h1 = X[i + 1] - X[i];
        r0 = (Y[i] - Y[i - 1]) / h0;
        r1 = (Y[i + 1] - Y[i]) / h1;
        A[i] = h1;
        B[i] = 2 * (h0 + h1);
        C[i] = h0;
        YD[i] = 3 * (r0 * h1 + r1 * h0);
    }
    A[i] = (h0 + h1) * (h0 + h1);
    B[i] = h0 * (h0 + h1);
    YD[i] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
    int rc = tridiag_gen(A, B, C, YD, len);
    free(A); free(B); free(C);
    return rc;
}

/* SplineEvaluate for one abscissa, lsdft_tools.cc:15-68.  The reference finds
 * the interval with a linear scan from j=0 ("for(;j<len1 && p2>X1[j];j++);
 * if(p2<X1[j]) j--;"); the first X1[j] >= p2 is found here by bisection, which
 * selects the same j.  Out-of-range abscissae are an error (the reference
 * calls exit(1), lsdft_tools.cc:20-26). */
static int spline_eval_d(const double *X1, const double *Y1, int len1, double p2, const double *YD, double *out,
                         double *dout);
int sqo_spline_eval(const double *X1, const double *Y1, int len1, double p2, const double *YD, double *out)
{
    return spline_eval_d(X1, Y1, len1, p2, YD, out, NULL);
}

/* The same cubic and, if dout is given, its derivative (3 A3 x + 2 A2) x + A1:
 * the exact derivative of the interpolant the value comes from. */
static int spline_eval_d(const double *X1, const double *Y1, int len1, double p2, const double *YD, double *out,
                         double *dout)
{
    if (p2 < X1[0] || p2 > X1[len1 - 1]) return SQO_ERR_RANGE;
    int lo = 0, hi = len1; /* first index with X1[idx] >= p2 */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (p2 > X1[mid]) lo = mid + 1; else hi = mid;
    }
    int j = lo;
    if (p2 < X1[j]) j--;
    double p1 = X1[j];
    double dx = 1.0 / (X1[j + 1] - X1[j]);
    double dy = (Y1[j + 1] - Y1[j]) * dx;
    double A0 = Y1[j];
    double A1 = YD[j];
    double A2 = dx * (3.0 * dy - 2.0 * YD[j] - YD[j + 1]);
    double A3 = dx * dx * (-2.0 * dy + YD[j] + YD[j + 1]);
    double x = p2 - p1;
    *out = ((A3 * x + A2) * x + A1) * x + A0;
    if (dout) *dout = (3.0 * A3 * x + 2.0 * A2) * x + A1;
    return SQO_OK;
}

/* RealSphericalHarmonic, lsdft_tools.cc:156-367, for l<=3 (the .pot reader
 * stores l=0..3 only, lsdft_readfiles.cc:949-1096, so l=4..6 are unreachable on
 * the path).  The coefficients are the reference's 15-digit truncations. */
double sqo_ylm(double x, double y, double z, int l, int m, int *err)
{
    const double C00 = 0.282094791773878;
    const double C1m1 = 0.488602511902920, C10 = 0.488602511902920, C1p1 = 0.488602511902920;
    const double C2m2 = 1.092548430592079, C2m1 = 1.092548430592079, C20 = 0.315391565252520;
    const double C2p1 = 1.092548430592079, C2p2 = 0.546274215296040;
    const double C3m3 = 0.590043589926644, C3m2 = 2.890611442640554, C3m1 = 0.457045799464466;
    const double C30 = 0.373176332590115, C3p1 = 0.457045799464466, C3p2 = 1.445305721320277;
    const double C3p3 = 0.590043589926644;
    double r = sqrt(x * x + y * y + z * z);
    double SH = 0.0;
    if (err) *err = 0;
    if (l == 0) SH = C00;
    if (r >= 1e-10) {
        if (l == 0) SH = C00;
        else if (l == 1) {
            if (m == -1) SH = C1m1 * (y / r);
            else if (m == 0) SH = C10 * (z / r);
            else if (m == 1) SH = C1p1 * (x / r);
            else if (err) *err = SQO_ERR_ARG;
        } else if (l == 2) {
            if (m == -2) SH = C2m2 * (x * y) / (r * r);
            else if (m == -1) SH = C2m1 * (y * z) / (r * r);
            else if (m == 0) SH = C20 * (-x * x - y * y + 2.0 * z * z) / (r * r);
            else if (m == 1) SH = C2p1 * (z * x) / (r * r);
            else if (m == 2) SH = C2p2 * (x * x - y * y) / (r * r);
            else if (err) *err = SQO_ERR_ARG;
        } else if (l == 3) {
            if (m == -3) SH = C3m3 * (3 * x * x - y * y) * y / (r * r * r);
            else if (m == -2) SH = C3m2 * (x * y * z) / (r * r * r);
            else if (m == -1) SH = C3m1 * y * (4 * z * z - x * x - y * y) / (r * r * r);
            else if (m == 0) SH = C30 * z * (2 * z * z - 3 * x * x - 3 * y * y) / (r * r * r);
            else if (m == 1) SH = C3p1 * x * (4 * z * z - x * x - y * y) / (r * r * r);
            else if (m == 2) SH = C3p2 * z * (x * x - y * y) / (r * r * r);
            else if (m == 3) SH = C3p3 * x * (x * x - 3 * y * y) / (r * r * r);
            else if (err) *err = SQO_ERR_ARG;
        } else if (err) *err = SQO_ERR_ARG;
    }
    return SH;
}

/* TridiagEigenVecSolve_NodesAndWeights, lsdft_tools.cc:912-1026: implicit QL
 * (Numerical Recipes tqli) with the full eigenvector matrix z (z = I at the
 * start); lambda_i = d_i (unsorted), w_i = z[0][i]^2. */
/* The QL iteration itself: d (diagonal, n), e (subdiagonal, n; e[n-1] is
 * overwritten), z (n*n, the identity on entry) -> eigenvalues in d,
 * eigenvectors in the columns of z. */
static int ql_full(double *d, double *e, double *z, int n)
{
    int m, l, iter, i, k;
    double s, r, p, g, f, dd, c, b;
#define Z(a, bb) z[(size_t)(a) * n + (bb)]
    e[n - 1] = 0.0;
    for (l = 0; l <= n - 1; l++) {
        iter = 0;
        do {
            for (m = l; m <= n - 2; m++) {
                dd = fabs(d[m]) + fabs(d[m + 1]);
                if ((double)(fabs(e[m]) + dd) == dd) break;
            }
            if (m != l) {
                if (iter++ == 50) return SQO_ERR_QL;
                g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                r = sqrt(g * g + 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                s = c = 1.0;
                p = 0.0;
                for (i = m - 1; i >= l; i--) {
                    f = s * e[i];
                    b = c * e[i];
                    e[i + 1] = (r = sqrt(g * g + f * f));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (k = 0; k < n; k++) {
                        f = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * f;
                        Z(k, i) = c * Z(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
#undef Z
    return SQO_OK;
}

int sqo_tridiag_nodes_weights(const double *diag, const double *subdiag, int n,
                              double *lambda, double *w, double *lmin, double *lmax)
{
    double *d = dalloc(n), *e = dalloc(n), *z = dalloc((size_t)n * n);
    for (int i = 0; i < n; i++) {
        d[i] = diag[i];
        e[i] = subdiag[i];
        z[(size_t)i * n + i] = 1.0;
    }
    int rc = ql_full(d, e, z, n);
    if (rc) { free(d); free(e); free(z); return rc; }
    double emin = d[0], emax = d[0];
    for (int i = 0; i < n; i++) {
        lambda[i] = d[i];
        w[i] = z[i] * z[i];   /* z[0][i]^2 */
        if (d[i] > emax) emax = d[i];
        else if (d[i] < emin) emin = d[i];
    }
    if (lmin) *lmin = emin;
    if (lmax) *lmax = emax;
    free(d); free(e); free(z);
    return SQO_OK;
}

/* ------------------------------------------------------------ input files */

/* Read_Parameters, lsdft_readfiles.cc:21-672 (the keywords the path uses). */
static int read_inpt(sqo_sys *S, const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return SQO_ERR_IO;
    S->ntype = -1; S->order = 6; S->beta = 1000; S->bc = 1; S->nqp = 50; S->rcut = 12.0;
    S->tol_psc = 1e-6;
    S->mix_beta = 0.3; S->mix_history = 7; S->pulay_frequency = 1; S->tol_scf = 1e-6; S->tol_poisson = 1e-8;
    S->maxit_scf = 100;
    char str[256];
    do {
        if (fscanf(f, "%255s", str) != 1) break;
        if (!strcmp(str, "FD_ORDER:")) { if (fscanf(f, "%d", &S->order) != 1) goto bad; S->order /= 2; }
        else if (!strcmp(str, "BOUNDARY_CONDITION:")) { if (fscanf(f, "%d", &S->bc) != 1) goto bad; }
        else if (!strcmp(str, "CELL:")) { if (fscanf(f, "%lf %lf %lf", &S->range[0], &S->range[1], &S->range[2]) != 3) goto bad; }
        else if (!strcmp(str, "FD_GRID:")) { if (fscanf(f, "%d %d %d", &S->npts[0], &S->npts[1], &S->npts[2]) != 3) goto bad; }
        else if (!strcmp(str, "BETA:")) { if (fscanf(f, "%lf", &S->beta) != 1) goto bad; }
        else if (!strcmp(str, "RADIUS_TRUNC:")) { if (fscanf(f, "%lf", &S->rcut) != 1) goto bad; }
        else if (!strcmp(str, "LANCZOS_DEGREE:")) { if (fscanf(f, "%d", &S->nqp) != 1) goto bad; }
        else if (!strcmp(str, "TOL_PSEUDOCHARGE:")) { if (fscanf(f, "%lf", &S->tol_psc) != 1) goto bad; }
        else if (!strcmp(str, "MIXING_PARAMETER:")) { if (fscanf(f, "%lf", &S->mix_beta) != 1) goto bad; }
        else if (!strcmp(str, "MIXING_HISTORY:")) { if (fscanf(f, "%d", &S->mix_history) != 1) goto bad; }
        else if (!strcmp(str, "PULAY_FREQUENCY:")) { if (fscanf(f, "%d", &S->pulay_frequency) != 1) goto bad; }
        else if (!strcmp(str, "TOL_SCF:")) { if (fscanf(f, "%lf", &S->tol_scf) != 1) goto bad; }
        else if (!strcmp(str, "TOL_POISSON:")) { if (fscanf(f, "%lf", &S->tol_poisson) != 1) goto bad; }
        else if (!strcmp(str, "MAXIT_SCF:")) { if (fscanf(f, "%d", &S->maxit_scf) != 1) goto bad; }
        else if (!strcmp(str, "NTYPE:")) {
            if (fscanf(f, "%d", &S->ntype) != 1 || S->ntype < 1 || S->ntype > SQO_MAXTYPE) goto bad;
        } else if (!strcmp(str, "PSEUDOPOTENTIAL_LOCAL:")) {
            if (S->ntype <= 0) goto bad;
            for (int i = 0; i < S->ntype; i++) if (fscanf(f, "%d", &S->localPsd[i]) != 1) goto bad;
        } else if (!strcmp(str, "PSEUDOPOTENTIAL_FILE:")) {
            if (S->ntype <= 0) goto bad;
            for (int i = 0; i < S->ntype; i++) if (fscanf(f, "%63s", S->psdName[i]) != 1) goto bad;
        }
    } while (!feof(f));
    fclose(f);
    if (S->ntype <= 0 || S->order < 1 || S->order > SQO_MAXORDER) return SQO_ERR_FORMAT;
    if (S->bc != 1) return SQO_ERR_ARG; /* BC=1 is the only case the shipped deck and the path use */
    /* grid spacing, lsdft_readfiles.cc:356-360 */
    for (int d = 0; d < 3; d++) S->delta[d] = 2.0 * S->range[d] / (S->npts[d] - 1);
    /* cube edge, lsdft_readfiles.cc:442-444 */
    for (int d = 0; d < 3; d++) S->nloc[d] = 2 * (int)ceil(S->rcut / S->delta[d]) + 1;
    S->dV = S->delta[0] * S->delta[1] * S->delta[2];
    /* -(1/2) Laplacian FD weights, lsdft_readfiles.cc:453-476 */
    double c0 = 0;
    for (int a = 1; a <= S->order; a++) c0 += ((double)1.0 / (a * a));
    for (int d = 0; d < 3; d++) {
        S->coeffs[d][0] = c0;
        S->coeffs[d][0] *= ((double)1.0 / (S->delta[d] * S->delta[d]));
        for (int p = 1; p <= S->order; p++)
            S->coeffs[d][p] = (double)(-1 * pow(-1, p + 1) * sqo_fract(S->order, p) /
                                       (p * p * S->delta[d] * S->delta[d]));
        for (int p = 0; p <= S->order; p++) S->stencil_coeffs[d][p] = S->coeffs[d][p] / (2.0 * M_PI);
    }
    return SQO_OK;
bad:
    fclose(f);
    return SQO_ERR_FORMAT;
}

/* Read_Ion, lsdft_readfiles.cc:677-789 */
static int read_ion(sqo_sys *S, const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return SQO_ERR_IO;
    if (fscanf(f, "%d", &S->nAtoms) != 1 || S->nAtoms < 0) { fclose(f); return SQO_ERR_FORMAT; }
    S->atompos = dalloc(3 * (size_t)S->nAtoms + 3);
    int at = 0, index = 0;
    S->noetot = 0;
    for (;;) {
        if (at >= SQO_MAXTYPE) break;
        if (fscanf(f, "%15s", S->atomType[at]) != 1) break;
        if (fscanf(f, "%d", &S->noe[at]) != 1) break;
        if (fscanf(f, "%d", &S->noa[at]) != 1) break;
        S->startPos[at] = index;
        for (int k = 0; k < S->noa[at]; k++) {
            S->noetot += S->noe[at];
            if (index + 3 > 3 * S->nAtoms) { fclose(f); return SQO_ERR_FORMAT; }
            if (fscanf(f, "%lf %lf %lf", &S->atompos[index], &S->atompos[index + 1], &S->atompos[index + 2]) != 3) {
                fclose(f); return SQO_ERR_FORMAT;
            }
            index += 3;
        }
        S->endPos[at] = index - 1;
        at++;
    }
    fclose(f);
    if (at != S->ntype) return SQO_ERR_FORMAT; /* lsdft_readfiles.cc:741-745 */
    return SQO_OK;
}

/* Read_Pseudopotential (Troullier-Martins format), lsdft_readfiles.cc:858-1101.
 * Units: V_l = 0.5*(r V_l in Ry)/r (Ry->Ha and r*V -> V), U_l = u_l/r,
 * uu = rho/(4 pi r^2); r=0 is prepended with V(0)=V(r_1), U(0)=U(r_1). */
static int read_pot(sqo_psd *P, const char *path, const char *atomType, int lloc)
{
    char str[60];
    double value, rc;
    int l, count = 0;
    FILE *f = fopen(path, "rb");
    if (!f) return SQO_ERR_IO;
    if (fscanf(f, "%59s", str) != 1 || strcmp(str, atomType)) { fclose(f); return SQO_ERR_FORMAT; }
    do { if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; } } while (strcmp(str, " Radial grid follows\n"));
    do {
        if (fscanf(f, "%59s", str) != 1) { fclose(f); return SQO_ERR_FORMAT; }
        count++;
    } while (strcmp(str, "Pseudopotential") != 0);
    fclose(f);
    count = count - 1;
    P->size = count + 1;
    P->r = dalloc(P->size);
    P->uu = dalloc(P->size);
    P->Vloc = dalloc(P->size);
    for (int i = 0; i < 4; i++) { P->V[i] = dalloc(P->size); P->U[i] = dalloc(P->size); P->rc[i] = 0.0; }
    /* rc[3] is never written for a 3-channel potential in the reference (UB,
     * lsdft_readfiles.cc:914); it is taken as 0 here. */
    f = fopen(path, "rb");
    if (!f) return SQO_ERR_IO;
    do { if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; } } while (strcmp(str, " Radial grid follows\n"));
    P->r[0] = 0.0;
    for (int i = 1; i <= count; i++) { if (fscanf(f, "%lf", &value) != 1) { fclose(f); return SQO_ERR_FORMAT; } P->r[i] = value; }
    int lineCtr = 0;
    while (strcmp(str, " Pseudopotential follows (l on next line)\n")) {
        if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; }
        lineCtr++;
    }
    P->lmax = -1;
    while (strcmp(str, " Pseudopotential follows (l on next line)\n") == 0) {
        if (fscanf(f, "%d", &l) != 1 || l < 0 || l > 3) { fclose(f); return SQO_ERR_FORMAT; }
        P->lmax = l;
        for (int i = 1; i <= count; i++) {
            if (fscanf(f, "%lf", &value) != 1) { fclose(f); return SQO_ERR_FORMAT; }
            value = 0.5 * value / (P->r[i]);
            P->V[l][i] = value;
            if (lloc == l) P->Vloc[i] = value;
        }
        P->V[l][0] = P->V[l][1];
        if (lloc == l) P->Vloc[0] = P->V[l][1];
        for (int i = 0; i < lineCtr; i++) if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; }
    }
    while (strcmp(str, " Valence charge follows\n")) if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; }
    while (strcmp(str, " Valence charge follows\n") == 0) {
        for (int i = 1; i <= count; i++) {
            if (fscanf(f, "%lf", &value) != 1) { fclose(f); return SQO_ERR_FORMAT; }
            value = value / (4 * M_PI * P->r[i] * P->r[i]);
            P->uu[i] = value;
        }
        P->uu[0] = 0.0;
        for (int i = 0; i < lineCtr; i++) if (!fgets(str, 60, f)) { fclose(f); return SQO_ERR_FORMAT; }
    }
    while (strcmp(str, " Pseudo-wave-function follows (l, zelect, rc)\n") == 0) {
        if (fscanf(f, "%d %lf %lf", &l, &value, &rc) != 3 || l < 0 || l > 3) { fclose(f); return SQO_ERR_FORMAT; }
        P->rc[l] = rc;
        for (int i = 1; i <= count; i++) {
            if (fscanf(f, "%lf", &value) != 1) { fclose(f); return SQO_ERR_FORMAT; }
            value = value / P->r[i];
            P->U[l][i] = value;
        }
        P->U[l][0] = P->U[l][1];
        for (int i = 0; i < lineCtr; i++) {
            if (feof(f)) break;
            if (!fgets(str, 60, f)) break;
        }
    }
    fclose(f);
    if (P->lmax < 0 || lloc > P->lmax) return SQO_ERR_FORMAT;
    return SQO_OK;
}

/* Calculate_SplineFit, lsdft_initObjs.cc:576-622 */
static int spline_fit(sqo_psd *P)
{
    int rc = 0;
    P->YD_Vloc = dalloc(P->size);
    P->YD_uu = dalloc(P->size);
    rc |= sqo_spline_yd(P->r, P->Vloc, P->YD_Vloc, P->size);
    rc |= sqo_spline_yd(P->r, P->uu, P->YD_uu, P->size);
    for (int l = 0; l <= P->lmax; l++) {
        P->UDeltaV[l] = dalloc(P->size);
        P->YD_UDeltaV[l] = dalloc(P->size);
        P->YD_U[l] = dalloc(P->size);
        for (int i = 0; i < P->size; i++) P->UDeltaV[l][i] = P->U[l][i] * (P->V[l][i] - P->Vloc[i]);
        rc |= sqo_spline_yd(P->r, P->UDeltaV[l], P->YD_UDeltaV[l], P->size);
        rc |= sqo_spline_yd(P->r, P->U[l], P->YD_U[l], P->size);
    }
    return rc ? SQO_ERR_RANGE : SQO_OK;
}

/* Calculate_NonlocalCoefficients, lsdft_initObjs.cc:822-869: Gamma_l = 1/D_l
 * with D_l a rectangle rule (dr = 1e-6) of U_l*dV_l*U_l*r^2 up to rc_l + max h,
 * minus half the last sample, divided by dV. */
static int nonlocal_gamma(sqo_sys *S, sqo_psd *P, int lloc)
{
    const double dr = 1e-6;
    double a = S->delta[0], b = S->delta[1], c = S->delta[2];
    double delta = (a > b && a > c ? a : b > c ? b : c);
    for (int l = 0; l <= P->lmax; l++) {
        P->Gamma[l] = 0.0;
        double rtemp = dr, Dexact = 0.0, UlDelVl = 0.0, Ul = 0.0;
        if (l != lloc) {
            double Rcut = P->rc[l];
            while (rtemp <= (Rcut + delta)) {
                if (sqo_spline_eval(P->r, P->UDeltaV[l], P->size, rtemp, P->YD_UDeltaV[l], &UlDelVl)) return SQO_ERR_RANGE;
                if (sqo_spline_eval(P->r, P->U[l], P->size, rtemp, P->YD_U[l], &Ul)) return SQO_ERR_RANGE;
                Dexact += UlDelVl * Ul * rtemp * rtemp * dr;
                rtemp = rtemp + dr;
            }
            Dexact = (Dexact - 0.5 * UlDelVl * Ul * (rtemp - dr) * (rtemp - dr) * dr) / (S->dV);
            P->Gamma[l] = 1.0 / Dexact;
        }
    }
    return SQO_OK;
}

/* Calculate_PseudochargeCutoff, lsdft_initObjs.cc:628-817: the pseudocharge
 * b = -(1/4pi) Lap_h V_loc on an octant grid around an atom at the origin;
 * the radius grows by 1 bohr from rc[lloc] until |int b + Z|/Z <= tol or 15. */
static int pseudocharge_cutoff(sqo_sys *S, int at)
{
    const double MaxRadius = 15.0;
    sqo_psd *P = &S->psd[at];
    int o = S->order;
    int Np[3];
    for (int d = 0; d < 3; d++) Np[d] = (int)ceil(MaxRadius / S->delta[d]);
    int Ex = Np[0] + 2 * o, Ey = Np[1] + 2 * o, Ez = Np[2] + 2 * o;
    double *Vps = dalloc((size_t)Ex * Ey * Ez);
    double *BJ = dalloc((size_t)Np[0] * Np[1] * Np[2]);
#define VPS(k, j, i) Vps[((size_t)(k) * Ey + (j)) * Ex + (i)]
#define BJA(k, j, i) BJ[((size_t)(k) * Np[1] + (j)) * Np[0] + (i)]
    for (int k = 0; k < Ez; k++)
        for (int j = 0; j < Ey; j++)
            for (int i = 0; i < Ex; i++) {
                double x = (i - o) * S->delta[0], y = (j - o) * S->delta[1], z = (k - o) * S->delta[2];
                double r = sqrt(x * x + y * y + z * z);
                if (sqo_spline_eval(P->r, P->Vloc, P->size, r, P->YD_Vloc, &VPS(k, j, i))) {
                    free(Vps); free(BJ); return SQO_ERR_RANGE;
                }
            }
    const double *scx = S->stencil_coeffs[0], *scy = S->stencil_coeffs[1], *scz = S->stencil_coeffs[2];
    for (int k = o; k < Np[2] + o; k++)
        for (int j = o; j < Np[1] + o; j++)
            for (int i = o; i < Np[0] + o; i++) {
                double v = VPS(k, j, i) * (scx[0] + scy[0] + scz[0]);
                for (int a = 1; a <= o; a++)
                    v += (VPS(k, j, i - a) + VPS(k, j, i + a)) * scx[a] +
                         (VPS(k, j - a, i) + VPS(k, j + a, i)) * scy[a] +
                         (VPS(k - a, j, i) + VPS(k + a, j, i)) * scz[a];
                BJA(k - o, j - o, i - o) = v;
            }
    double Rcut = P->rc[S->localPsd[at]], error;
    int Ncx, Ncy, Ncz;
    do {
        Ncx = (int)ceil(Rcut / S->delta[0]);
        Ncy = (int)ceil(Rcut / S->delta[1]);
        Ncz = (int)ceil(Rcut / S->delta[2]);
        double Bint = 0.0;
        for (int k = 0; k < Ncz; k++)
            for (int j = 0; j < Ncy; j++)
                for (int i = 0; i < Ncx; i++) {
                    double val = BJA(k, j, i);
                    if (k == 0) val = 0.5 * val;
                    if (k == Ncz - 1) val = 0.5 * val;
                    if (j == 0) val = 0.5 * val;
                    if (j == Ncy - 1) val = 0.5 * val;
                    if (i == 0) val = 0.5 * val;
                    if (i == Ncx - 1) val = 0.5 * val;
                    Bint += val;
                }
        Bint = Bint * 8.0 * S->dV;
        error = fabs(Bint + S->noe[at]) / fabs((double)S->noe[at]);
        Rcut = Rcut + 1.0;
    } while (Rcut <= 15.0 && error > S->tol_psc);
    S->cutoff[at][0] = Ncx * S->delta[0];
    S->cutoff[at][1] = Ncy * S->delta[1];
    S->cutoff[at][2] = Ncz * S->delta[2];
#undef VPS
#undef BJA
    free(Vps); free(BJ);
    return SQO_OK;
}

/* Check_Overlap, lsdft_initObjs.cc:2612-2663 */
static int check_overlap(int xcor, int ycor, int zcor, int lxdim, int lydim, int lzdim, int xs, int ys, int zs,
                         int xl, int yl, int zl, int *xstart, int *ystart, int *zstart, int *xend, int *yend, int *zend)
{
    *xstart = *ystart = *zstart = *xend = *yend = *zend = -1000;
    if (xs >= xcor && xs <= xcor + lxdim - 1) *xstart = xs;
    else if (xcor >= xs && xcor <= xl) *xstart = xcor;
    if (xl >= xcor && xl <= xcor + lxdim - 1) *xend = xl;
    else if (xcor + lxdim - 1 >= xs && xcor + lxdim - 1 <= xl) *xend = xcor + lxdim - 1;
    if (ys >= ycor && ys <= ycor + lydim - 1) *ystart = ys;
    else if (ycor >= ys && ycor <= yl) *ystart = ycor;
    if (yl >= ycor && yl <= ycor + lydim - 1) *yend = yl;
    else if (ycor + lydim - 1 >= ys && ycor + lydim - 1 <= yl) *yend = ycor + lydim - 1;
    if (zs >= zcor && zs <= zcor + lzdim - 1) *zstart = zs;
    else if (zcor >= zs && zcor <= zl) *zstart = zcor;
    if (zl >= zcor && zl <= zcor + lzdim - 1) *zend = zl;
    else if (zcor + lzdim - 1 >= zs && zcor + lzdim - 1 <= zl) *zend = zcor + lzdim - 1;
    return (*xstart != -1000) && (*xend != -1000) && (*ystart != -1000) && (*yend != -1000) &&
           (*zstart != -1000) && (*zend != -1000);
}

/* Atom position -> nearest grid index as the reference computes it:
 * roundf((x0 + R)/h) rounds a float (lsdft_initObjs.cc:1534-1536, 1800-1802). */
static int grid_round(double v) { return (int)roundf((float)v); }

/* ------------------------------------------------------------- public API */

void sqo_free(sqo_sys *S);

/* Initialize (lsdft_initObjs.cc:36-93) minus the PETSc objects: Read_Parameters,
 * Read_Ion, Read_Pseudopotential, Calculate_SplineFit,
 * Calculate_NonlocalCoefficients, Calculate_PseudochargeCutoff, and
 * Create_SupercellAtomList + Calculate_NonlocalIndex with no phantom atom.
 * Reads <dir>/<name>.inpt, <dir>/<name>.ion, <dir>/pseudopotential/<file>. */
int sqo_load(const char *dir, const char *name, sqo_sys **out)
{
    char path[1024];
    int rc;
    sqo_sys *S = (sqo_sys *)calloc(1, sizeof(sqo_sys));
    if (!S) return SQO_ERR_ALLOC;
    snprintf(path, sizeof path, "%s/%s.inpt", dir, name);
    if ((rc = read_inpt(S, path))) goto fail;
    snprintf(path, sizeof path, "%s/%s.ion", dir, name);
    if ((rc = read_ion(S, path))) goto fail;
    for (int at = 0; at < S->ntype; at++) {
        snprintf(path, sizeof path, "%s/pseudopotential/%s", dir, S->psdName[at]);
        if ((rc = read_pot(&S->psd[at], path, S->atomType[at], S->localPsd[at]))) goto fail;
        if ((rc = spline_fit(&S->psd[at]))) goto fail;
        if ((rc = nonlocal_gamma(S, &S->psd[at], S->localPsd[at]))) goto fail;
        if ((rc = pseudocharge_cutoff(S, at))) goto fail;
    }
    extern int sqo_set_phantom(sqo_sys *, int, double, double, double);
    if ((rc = sqo_set_phantom(S, 0, 0, 0, 0))) goto fail;
    *out = S;
    return SQO_OK;
fail:
    sqo_free(S);
    *out = NULL;
    return rc;
}

/* Create_SupercellAtomList (lsdft_initObjs.cc:3031-3041) + Calculate_NonlocalIndex
 * (875-967).  The reference sets endPosSC[0] = 3*nAtoms, one past the last
 * coordinate, so every atom loop visits a phantom atom whose coordinates are
 * read past the end of the position array (undefined behaviour).  enable=0
 * drops it (the loops visit the real atoms only); enable=1 places it at
 * (x,y,z), as if the memory after the array held those values.  Only Ntype=1
 * reaches the phantom: the reference leaves startPosSC/endPosSC of other types
 * unset, and here they take the .ion ranges. */
int sqo_set_phantom(sqo_sys *S, int enable, double x, double y, double z)
{
    free(S->atomposSC);
    S->atomposSC = dalloc(3 * (size_t)S->nAtoms + 3);
    memcpy(S->atomposSC, S->atompos, sizeof(double) * 3 * S->nAtoms);
    S->atomposSC[3 * S->nAtoms] = x;
    S->atomposSC[3 * S->nAtoms + 1] = y;
    S->atomposSC[3 * S->nAtoms + 2] = z;
    for (int at = 0; at < S->ntype; at++) {
        S->startPosSC[at] = S->startPos[at];
        S->endPosSC[at] = S->endPos[at];
    }
    if (enable) {
        if (S->ntype != 1) return SQO_ERR_ARG;
        S->startPosSC[0] = 0;
        S->endPosSC[0] = 3 * S->nAtoms;
    }
    /* Calculate_NonlocalIndex: IP_index[p] = first alpha slot of atom p */
    int natoms_sc = 0, len = 0;
    for (int at = 0; at < S->ntype; at++) {
        int start = S->startPosSC[at] / 3, end = S->endPosSC[at] / 3;
        natoms_sc += end - start + 1;
        int lmax = S->psd[at].lmax;
        for (int p = start; p <= end; p++)
            for (int l = 0; l <= lmax; l++)
                if (l != S->localPsd[at]) len += 2 * l + 1;
    }
    free(S->IP_index);
    S->IP_index = ialloc(natoms_sc);
    S->nIP = natoms_sc;
    S->IP_length = len;
    int index = 0, posIdx = 0;
    for (int at = 0; at < S->ntype; at++) {
        int start = S->startPosSC[at] / 3, end = S->endPosSC[at] / 3;
        for (int p = start; p <= end; p++) {
            S->IP_index[index++] = posIdx;
            for (int l = 0; l <= S->psd[at].lmax; l++)
                if (l != S->localPsd[at]) posIdx += 2 * l + 1;
        }
    }
    S->have_nl = 0;
    return SQO_OK;
}

/* Image range of the periodic map (BC=1), lsdft_initObjs.cc:1490-1499 / 1763-1770 */
static void image_range(const sqo_sys *S, int at, int Imax[3])
{
    for (int d = 0; d < 3; d++) Imax[d] = (int)ceil(S->cutoff[at][d] / S->range[d]);
}

/* GetInfluencingAtoms_Local count, lsdft_initObjs.cc:1409-1665: number of atom
 * images whose pseudocharge box overlaps the rank box (xcor,ycor,zcor,lx,ly,lz). */
int sqo_count_local_overlaps(sqo_sys *S, const int box[6], int *counts)
{
    int index = 0;
    for (int at = 0; at < S->ntype; at++) {
        int off[3], Imax[3], count = 0;
        for (int d = 0; d < 3; d++) off[d] = (int)ceil(S->cutoff[at][d] / S->delta[d] + 0.5);
        image_range(S, at, Imax);
        int start = S->startPosSC[at] / 3, end = S->endPosSC[at] / 3;
        for (int p = start; p <= end; p++) {
            double X0 = S->atomposSC[index++], Y0 = S->atomposSC[index++], Z0 = S->atomposSC[index++];
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        double x0 = X0 + PP * 2.0 * S->range[0];
                        double y0 = Y0 + QQ * 2.0 * S->range[1];
                        double z0 = Z0 + RR * 2.0 * S->range[2];
                        int xi = grid_round((x0 + S->range[0]) / S->delta[0]);
                        int yj = grid_round((y0 + S->range[1]) / S->delta[1]);
                        int zk = grid_round((z0 + S->range[2]) / S->delta[2]);
                        int xs, ys, zs, xe, ye, ze;
                        count += check_overlap(box[0], box[1], box[2], box[3], box[4], box[5], xi - off[0],
                                               yj - off[1], zk - off[2], xi + off[0], yj + off[1], zk + off[2],
                                               &xs, &ys, &zs, &xe, &ye, &ze);
                    }
        }
        counts[at] = count;
    }
    return SQO_OK;
}

static void free_nl(sqo_sys *S)
{
    for (int at = 0; at < SQO_MAXTYPE; at++) {
        sqo_overlap *O = &S->ovl[at];
        for (int p = 0; p < O->n; p++) {
            if (O->chi) free(O->chi[p]);
            if (O->lId) free(O->lId[p]);
            if (O->mId) free(O->mId[p]);
            if (O->kId) free(O->kId[p]);
            if (O->jId) free(O->jId[p]);
            if (O->iId) free(O->iId[p]);
        }
        free(O->chi); free(O->lId); free(O->mId); free(O->kId); free(O->jId); free(O->iId); free(O->ndim);
        free(O->X0); free(O->Y0); free(O->Z0); free(O->xs); free(O->ys); free(O->zs);
        free(O->xe); free(O->ye); free(O->ze); free(O->xindex);
        memset(O, 0, sizeof *O);
    }
    S->have_nl = 0;
}

/* GetInfluencingAtoms_Nonlocal (lsdft_initObjs.cc:1671-1946) for the rank box
 * box = (xcor,ycor,zcor,lx,ly,lz) extended by ceil(Rcut/h) on every side, then
 * CalculateNonlocalProjectors (lsdft_nl.cc:141-208) and the flattening of
 * NonlocalProjectors (lsdft_nl.cc:6-130): k,j,i outer, (l,m) inner. */
int sqo_build_nonlocal(sqo_sys *S, const int box[6])
{
    free_nl(S);
    memcpy(S->box, box, sizeof(int) * 6);
    int nr[3];
    for (int d = 0; d < 3; d++) nr[d] = (int)ceil(S->rcut / S->delta[d]);
    int xcor = box[0] - nr[0], ycor = box[1] - nr[1], zcor = box[2] - nr[2];
    int lxdim = box[3] + 2 * nr[0], lydim = box[4] + 2 * nr[1], lzdim = box[5] + 2 * nr[2];
    int index = 0;
    for (int at = 0; at < S->ntype; at++) {
        sqo_psd *P = &S->psd[at];
        sqo_overlap *O = &S->ovl[at];
        double max1 = P->rc[0] > P->rc[1] ? P->rc[0] : P->rc[1];
        double max2 = P->rc[2] > P->rc[3] ? P->rc[2] : P->rc[3];
        double cut = max1 > max2 ? max1 : max2;
        int off[3], Imax[3];
        for (int d = 0; d < 3; d++) off[d] = (int)ceil(cut / S->delta[d]);
        image_range(S, at, Imax);
        int start = S->startPosSC[at] / 3, end = S->endPosSC[at] / 3;
        /* the reference counts first, then stores; one pass with a growing list is equivalent */
        int cap = 64;
        O->X0 = dalloc(cap); O->Y0 = dalloc(cap); O->Z0 = dalloc(cap);
        O->xs = ialloc(cap); O->ys = ialloc(cap); O->zs = ialloc(cap);
        O->xe = ialloc(cap); O->ye = ialloc(cap); O->ze = ialloc(cap); O->xindex = ialloc(cap);
        for (int p = start; p <= end; p++) {
            double X0 = S->atomposSC[index++], Y0 = S->atomposSC[index++], Z0 = S->atomposSC[index++];
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        if (P->lmax == 0) continue;
                        double x0 = X0 + PP * 2.0 * S->range[0];
                        double y0 = Y0 + QQ * 2.0 * S->range[1];
                        double z0 = Z0 + RR * 2.0 * S->range[2];
                        int xi = grid_round((x0 + S->range[0] + 0 * S->rcut) / S->delta[0]);
                        int yj = grid_round((y0 + S->range[1] + 0 * S->rcut) / S->delta[1]);
                        int zk = grid_round((z0 + S->range[2] + 0 * S->rcut) / S->delta[2]);
                        int xs, ys, zs, xe, ye, ze;
                        if (!check_overlap(xcor, ycor, zcor, lxdim, lydim, lzdim, xi - off[0], yj - off[1],
                                           zk - off[2], xi + off[0], yj + off[1], zk + off[2], &xs, &ys, &zs,
                                           &xe, &ye, &ze))
                            continue;
                        if (O->n == cap) {
                            cap *= 2;
                            O->X0 = realloc(O->X0, cap * sizeof(double)); O->Y0 = realloc(O->Y0, cap * sizeof(double));
                            O->Z0 = realloc(O->Z0, cap * sizeof(double));
                            O->xs = realloc(O->xs, cap * sizeof(int)); O->ys = realloc(O->ys, cap * sizeof(int));
                            O->zs = realloc(O->zs, cap * sizeof(int)); O->xe = realloc(O->xe, cap * sizeof(int));
                            O->ye = realloc(O->ye, cap * sizeof(int)); O->ze = realloc(O->ze, cap * sizeof(int));
                            O->xindex = realloc(O->xindex, cap * sizeof(int));
                        }
                        int c = O->n++;
                        O->X0[c] = x0; O->Y0[c] = y0; O->Z0[c] = z0;
                        O->xs[c] = xs; O->ys[c] = ys; O->zs[c] = zs;
                        O->xe[c] = xe; O->ye[c] = ye; O->ze[c] = ze;
                        O->xindex[c] = index - 3;
                    }
        }
        /* projectors */
        int lloc = S->localPsd[at];
        int Nlm = P->lmax * P->lmax + 2 * (P->lmax - lloc);
        O->ndim = ialloc(O->n);
        O->chi = calloc(O->n ? O->n : 1, sizeof(double *));
        O->lId = calloc(O->n ? O->n : 1, sizeof(int *)); O->mId = calloc(O->n ? O->n : 1, sizeof(int *));
        O->kId = calloc(O->n ? O->n : 1, sizeof(int *)); O->jId = calloc(O->n ? O->n : 1, sizeof(int *));
        O->iId = calloc(O->n ? O->n : 1, sizeof(int *));
        for (int c = 0; c < O->n; c++) {
            int nz = O->ze[c] - O->zs[c] + 1, ny = O->ye[c] - O->ys[c] + 1, nx = O->xe[c] - O->xs[c] + 1;
            int nd = Nlm * nz * ny * nx;
            O->ndim[c] = nd;
            O->chi[c] = dalloc(nd);
            O->lId[c] = ialloc(nd); O->mId[c] = ialloc(nd); O->kId[c] = ialloc(nd);
            O->jId[c] = ialloc(nd); O->iId[c] = ialloc(nd);
            int LI = 0;
            for (int k = O->zs[c]; k <= O->ze[c]; k++) {
                double z = S->delta[2] * k - S->range[2] - O->Z0[c];
                double zz = z * z;
                for (int j = O->ys[c]; j <= O->ye[c]; j++) {
                    double y = S->delta[1] * j - S->range[1] - O->Y0[c];
                    double yy = y * y;
                    for (int i = O->xs[c]; i <= O->xe[c]; i++) {
                        double x = S->delta[0] * i - S->range[0] - O->X0[c];
                        double r = sqrt(x * x + yy + zz);
                        for (int l = 0; l <= P->lmax; l++) {
                            if (l == lloc) continue;
                            double ul;
                            if (sqo_spline_eval(P->r, P->UDeltaV[l], P->size, r, P->YD_UDeltaV[l], &ul)) {
                                free_nl(S); return SQO_ERR_RANGE;
                            }
                            for (int m = -l; m <= l; m++) {
                                int e2;
                                double sh = sqo_ylm(x, y, z, l, m, &e2);
                                if (e2) { free_nl(S); return e2; }
                                O->chi[c][LI] = ul * sh;
                                O->lId[c][LI] = l; O->mId[c][LI] = m;
                                O->kId[c][LI] = k; O->jId[c][LI] = j; O->iId[c][LI] = i;
                                LI++;
                            }
                        }
                    }
                }
            }
        }
    }
    S->have_nl = 1;
    return SQO_OK;
}
/* Note on the projector coordinate: the reference writes
 * x = delta_x*(i + xstart) - range_x - X0 with i counted from xstart
 * (lsdft_nl.cc:189-190); the loop above counts the global index directly, the
 * same integer, so x is bit-identical. */

/* Create_LaplacianLocal (lsdft_initObjs.cc:393-475): the 6*order+1 point star of
 * -(1/2) Lap_h on the n_x*n_y*n_z cube with zero-Dirichlet truncation
 * (DM_BOUNDARY_GHOSTED drops columns outside the cube).  Stored as AIJ rows
 * with ascending column index, the order SeqAIJ keeps. */
static int build_cube_laplacian(sqo_sys *S)
{
    if (S->rowptr) return SQO_OK;
    int nx = S->nloc[0], ny = S->nloc[1], nz = S->nloc[2], o = S->order;
    int nc = nx * ny * nz;
    S->nc = nc;
    S->rowptr = ialloc(nc + 1);
    S->colidx = ialloc((size_t)nc * (6 * o + 1));
    S->val = dalloc((size_t)nc * (6 * o + 1));
    int nnz = 0;
    double cdiag = S->coeffs[0][0] + S->coeffs[1][0] + S->coeffs[2][0];
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int row = (k * ny + j) * nx + i;
                S->rowptr[row] = nnz;
#define PUT(kk, jj, ii, v) do { S->colidx[nnz] = ((kk) * ny + (jj)) * nx + (ii); S->val[nnz] = (v); nnz++; } while (0)
                for (int l = o; l >= 1; l--) if (k - l >= 0) PUT(k - l, j, i, S->coeffs[2][l]);
                for (int l = o; l >= 1; l--) if (j - l >= 0) PUT(k, j - l, i, S->coeffs[1][l]);
                for (int l = o; l >= 1; l--) if (i - l >= 0) PUT(k, j, i - l, S->coeffs[0][l]);
                PUT(k, j, i, cdiag);
                for (int l = 1; l <= o; l++) if (i + l < nx) PUT(k, j, i + l, S->coeffs[0][l]);
                for (int l = 1; l <= o; l++) if (j + l < ny) PUT(k, j + l, i, S->coeffs[1][l]);
                for (int l = 1; l <= o; l++) if (k + l < nz) PUT(k + l, j, i, S->coeffs[2][l]);
#undef PUT
            }
    S->rowptr[nc] = nnz;
    S->center = (((nz - 1) / 2) * ny + (ny - 1) / 2) * nx + (nx - 1) / 2; /* Nmid, lsdft_scf.cc:53-57 */
    return SQO_OK;
}

/* per-thread scratch of one node's Lanczos */
typedef struct {
    double *dval;   /* diagonal-updated copy of the AIJ values (MatCopy + MatDiagonalSet) */
    double *vk, *vkm1, *vkp1, *alpha, *veff;
} node_ws;

static int ws_init(node_ws *W, const sqo_sys *S)
{
    size_t nnz = S->rowptr[S->nc];
    W->dval = dalloc(nnz);
    W->vk = dalloc(S->nc); W->vkm1 = dalloc(S->nc); W->vkp1 = dalloc(S->nc); W->veff = dalloc(S->nc);
    W->alpha = dalloc(S->IP_length > 0 ? S->IP_length : 1);
    return (W->dval && W->vk && W->vkm1 && W->vkp1 && W->alpha && W->veff) ? SQO_OK : SQO_ERR_ALLOC;
}
static void ws_free(node_ws *W)
{
    free(W->dval); free(W->vk); free(W->vkm1); free(W->vkp1); free(W->alpha); free(W->veff);
}

/* V_eff on the node's cube: the global field (natural order, i fastest) wrapped
 * periodically in grid index, with period N or, for fields on the distinct
 * points of the periodic cell, N-1 (sqo_set_cell_period); NULL means the
 * constant veff_const everywhere,
 * which is what the reference uses (VecSet(VeffNodal, 0.333), lsdft_scf.cc:71). */
static void cube_veff(const sqo_sys *S, const double *field, double veff_const, int Kp, int Jp, int Ip, double *out)
{
    int nx = S->nloc[0], ny = S->nloc[1], nz = S->nloc[2];
    if (!field) { for (int p = 0; p < S->nc; p++) out[p] = veff_const; return; }
    int Nx = S->npts[0] - S->cell_period, Ny = S->npts[1] - S->cell_period, Nz = S->npts[2] - S->cell_period;
    int x0 = Ip - (nx - 1) / 2, y0 = Jp - (ny - 1) / 2, z0 = Kp - (nz - 1) / 2;
    for (int k = 0; k < nz; k++) {
        int K = ((z0 + k) % Nz + Nz) % Nz;
        for (int j = 0; j < ny; j++) {
            int J = ((y0 + j) % Ny + Ny) % Ny;
            for (int i = 0; i < nx; i++) {
                int I = ((x0 + i) % Nx + Nx) % Nx;
                out[(k * ny + j) * nx + i] = field[((size_t)K * Ny + J) * Nx + I];
            }
        }
    }
}

/* MatCopy(LaplacianOprloc) + MatDiagonalSet(+VeffNodal, ADD_VALUES), lsdft_scf.cc:73-75 */
static void rebuild_operator(const sqo_sys *S, node_ws *W)
{
    memcpy(W->dval, S->val, sizeof(double) * S->rowptr[S->nc]);
    for (int row = 0; row < S->nc; row++)
        for (int q = S->rowptr[row]; q < S->rowptr[row + 1]; q++)
            if (S->colidx[q] == row) { W->dval[q] = W->dval[q] + W->veff[row]; break; }
}

/* y = A x for the AIJ cube operator (PETSc MatMult_SeqAIJ row sum, taken as
 * the sequential sum over ascending columns; PETSc's unrolling is unpinned). */
static void spmv(const sqo_sys *S, const double *dval, const double *x, double *y)
{
    for (int row = 0; row < S->nc; row++) {
        double sum = 0.0;
        for (int q = S->rowptr[row]; q < S->rowptr[row + 1]; q++) sum += dval[q] * x[S->colidx[q]];
        y[row] = sum;
    }
}

/* V2 += V_nl V1 on the cube of node (Kp,Jp,Ip).
 * mode 0 = Mult_NonlocalVector (lsdft_matvec.cc:25-149): every entry of atom p
 *          uses the single slot alpha[IP_index[p]] (all (l,m) channels and all
 *          images of the atom collapse into one coefficient).
 * mode 1 = Mult_HamiltonianVector (lsdft_matvec.cc:155-293): slot
 *          IP_index[p] + (l,m) channel, the Kleinman-Bylander operator.
 * Entries are scanned over every image of the rank box, with the cube test of
 * lsdft_matvec.cc:95,134.  V2 == NULL computes alpha only (the first pass). */
static void nonlocal_apply(const sqo_sys *S, int mode, double *alpha, int Kp, int Jp, int Ip,
                           const double *V1, double *V2)
{
    int nx = S->nloc[0], ny = S->nloc[1];
    int hx = (int)ceil(S->rcut / S->delta[0]), hy = (int)ceil(S->rcut / S->delta[1]), hz = (int)ceil(S->rcut / S->delta[2]);
    int Xstart = Ip - hx, Xend = Ip + hx, Ystart = Jp - hy, Yend = Jp + hy, Zstart = Kp - hz, Zend = Kp + hz;
    memset(alpha, 0, sizeof(double) * (S->IP_length > 0 ? S->IP_length : 1));
    for (int pass = 0; pass < (V2 ? 2 : 1); pass++) {
        for (int at = 0; at < S->ntype; at++) {
            const sqo_overlap *O = &S->ovl[at];
            const double *G = S->psd[at].Gamma;
            int lloc = S->localPsd[at];
            for (int c = 0; c < O->n; c++) {
                int p = O->xindex[c] / 3;
                int base = S->IP_index[p];
                const double *chi = O->chi[c];
                const int *lI = O->lId[c], *mI = O->mId[c], *kI = O->kId[c], *jI = O->jId[c], *iI = O->iId[c];
                for (int LI = 0; LI < O->ndim[c]; LI++) {
                    int k = kI[LI], j = jI[LI], i = iI[LI];
                    if ((k <= Zend && k >= Zstart) && (j <= Yend && j >= Ystart) && (i <= Xend && i >= Xstart)) {
                        int slot = base;
                        if (mode == 1) {
                            int l = lI[LI], ch = 0;
                            for (int ll = 0; ll < l; ll++) if (ll != lloc) ch += 2 * ll + 1;
                            slot = base + ch + (mI[LI] + l);
                        }
                        size_t q = ((size_t)(k - Zstart) * ny + (j - Ystart)) * nx + (i - Xstart);
                        if (pass == 0) alpha[slot] += V1[q] * chi[LI] * G[lI[LI]];
                        else V2[q] += chi[LI] * alpha[slot];
                    }
                }
            }
        }
    }
}

/* The nonlocal part alone, accumulated into y: Mult_NonlocalVector (mode 0)
 * or Mult_HamiltonianVector (mode 1) as called at lsdft_scf.cc:123,140. */
int sqo_apply_nonlocal(sqo_sys *S, int mode, int Kp, int Jp, int Ip, const double *x, double *y)
{
    if (!S->have_nl || (mode != 0 && mode != 1)) return SQO_ERR_ARG;
    double *alpha = (double *)calloc(S->IP_length > 0 ? S->IP_length : 1, sizeof(double));
    if (!alpha) return SQO_ERR_ALLOC;
    nonlocal_apply(S, mode, alpha, Kp, Jp, Ip, x, y);
    free(alpha);
    return SQO_OK;
}

/* One H application on node (Kp,Jp,Ip): y = (L + diag V_eff) x + V_nl x. */
int sqo_apply_h(sqo_sys *S, int mode, const double *veff_field, double veff_const, int Kp, int Jp, int Ip,
                const double *x, double *y)
{
    if (!S->have_nl || (mode != 0 && mode != 1)) return SQO_ERR_ARG;
    build_cube_laplacian(S);
    node_ws W;
    if (ws_init(&W, S)) { ws_free(&W); return SQO_ERR_ALLOC; }
    cube_veff(S, veff_field, veff_const, Kp, Jp, Ip, W.veff);
    rebuild_operator(S, &W);
    spmv(S, W.dval, x, y);
    nonlocal_apply(S, mode, W.alpha, Kp, Jp, Ip, x, y);
    ws_free(&W);
    return SQO_OK;
}

static double vdot(const double *x, const double *y, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += y[i] * x[i];
    return s;
}
static double vnorm(const double *x, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* CalculateSpectralNodesAndWeights, lsdft_scf.cc:87-164 (Lanczos part).
 * a[0..nqp], b[0..nqp] are written (a[nqp], b[nqp] are not used by the QL).
 * Q, if given, receives the Lanczos vectors q_0 .. q_{nqp-1} (nqp rows of nc). */
static int lanczos_node(const sqo_sys *S, int mode, node_ws *W, int Kp, int Jp, int Ip, double *a, double *b,
                        double *Q)
{
    int n = S->nc, N_qp = S->nqp;
    memset(W->vk, 0, sizeof(double) * n);
    memset(W->vkm1, 0, sizeof(double) * n);
    memset(W->vkp1, 0, sizeof(double) * n);
    W->vkm1[S->center] = 1.0;
    if (Q) memcpy(Q, W->vkm1, sizeof(double) * n);
    spmv(S, W->dval, W->vkm1, W->vk);
    nonlocal_apply(S, mode, W->alpha, Kp, Jp, Ip, W->vkm1, W->vk);
    a[0] = vdot(W->vkm1, W->vk, n);
    for (int i = 0; i < n; i++) W->vk[i] = W->vk[i] + (-a[0]) * W->vkm1[i];
    b[0] = vnorm(W->vk, n);
    if (b[0] == 0.0) return SQO_ERR_BREAKDOWN;
    double s = 1.0 / b[0];
    for (int i = 0; i < n; i++) W->vk[i] = s * W->vk[i];
    if (Q && N_qp > 1) memcpy(Q + n, W->vk, sizeof(double) * n);
    for (int k = 0; k < N_qp; k++) {
        spmv(S, W->dval, W->vk, W->vkp1);
        nonlocal_apply(S, mode, W->alpha, Kp, Jp, Ip, W->vk, W->vkp1);
        a[k + 1] = vdot(W->vk, W->vkp1, n);
        for (int i = 0; i < n; i++) W->vkp1[i] = W->vkp1[i] + (-a[k + 1]) * W->vk[i];
        for (int i = 0; i < n; i++) W->vkp1[i] = W->vkp1[i] + (-b[k]) * W->vkm1[i];
        memcpy(W->vkm1, W->vk, sizeof(double) * n);
        b[k + 1] = vnorm(W->vkp1, n);
        memcpy(W->vk, W->vkp1, sizeof(double) * n);
        if (b[k + 1] == 0.0) return SQO_ERR_BREAKDOWN;
        s = 1.0 / b[k + 1];
        for (int i = 0; i < n; i++) W->vk[i] = s * W->vk[i];
        if (Q && k + 2 < N_qp) memcpy(Q + (size_t)(k + 2) * n, W->vk, sizeof(double) * n);
    }
    return SQO_OK;
}

/* Lanczos coefficients a[0..nqp], b[0..nqp] of one node */
int sqo_lanczos(sqo_sys *S, int mode, const double *veff_field, double veff_const, int Kp, int Jp, int Ip,
                double *a, double *b)
{
    if (!S->have_nl || (mode != 0 && mode != 1)) return SQO_ERR_ARG;
    build_cube_laplacian(S);
    node_ws W;
    if (ws_init(&W, S)) { ws_free(&W); return SQO_ERR_ALLOC; }
    cube_veff(S, veff_field, veff_const, Kp, Jp, Ip, W.veff);
    rebuild_operator(S, &W);
    int rc = lanczos_node(S, mode, &W, Kp, Jp, Ip, a, b, NULL);
    ws_free(&W);
    return rc;
}

/* NodesAndWeights_LSSGQ, lsdft_scf.cc:32-82, for the nodes of the rank box
 * given to sqo_build_nonlocal.  node_ids (or NULL for all, in LIpos order
 * (k-zcor)*lx*ly + (j-ycor)*lx + (i-xcor)) selects which nodes to evaluate;
 * row r of lam/w (nqp entries each) and lmin/lmax belong to node_ids[r].
 * nthreads > 1 evaluates nodes in parallel (one scratch set per thread). */
int sqo_nodes_and_weights(sqo_sys *S, int mode, const double *veff_field, double veff_const,
                          const int *node_ids, int n_nodes, double *lam, double *w, double *lmin, double *lmax,
                          int nthreads)
{
    if (!S->have_nl || (mode != 0 && mode != 1)) return SQO_ERR_ARG;
    build_cube_laplacian(S);
    int lx = S->box[3], ly = S->box[4], lz = S->box[5];
    int total = lx * ly * lz;
    if (!node_ids) n_nodes = total;
    int err = SQO_OK;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        node_ws W;
        int e = ws_init(&W, S);
        double *a = dalloc(S->nqp + 1), *b = dalloc(S->nqp + 1);
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < n_nodes; r++) {
            if (e) continue;
            int id = node_ids ? node_ids[r] : r;
            if (id < 0 || id >= total) { e = SQO_ERR_ARG; continue; }
            int i = S->box[0] + id % lx, j = S->box[1] + (id / lx) % ly, k = S->box[2] + id / (lx * ly);
            cube_veff(S, veff_field, veff_const, k, j, i, W.veff);
            rebuild_operator(S, &W);
            int rc = lanczos_node(S, mode, &W, k, j, i, a, b, NULL);
            if (!rc) rc = sqo_tridiag_nodes_weights(a, b, S->nqp, lam + (size_t)r * S->nqp, w + (size_t)r * S->nqp,
                                                    lmin + r, lmax + r);
            if (rc) e = rc;
        }
        if (e) {
#pragma omp critical
            err = e;
        }
        free(a); free(b);
        ws_free(&W);
    }
    return err;
}

/* ------------------------------------------------ density, Fermi, energy
 * The reference declares these only (lsdft.h:98-102, 209).  Written from SQ
 * theory: with the unit vector e_p, rho_p = (2/dV) sum_i w_pi f(lambda_pi),
 * the Fermi level solves 2 sum_p sum_i w_pi f(lambda_pi) = N_e, the band energy
 * is 2 sum w lambda f and -TS = (2/beta) sum w [f ln f + (1-f) ln(1-f)],
 * f(l) = 1/(1+exp(beta (l - lambda_f))).  The root is found by bisection on
 * [min lambda_min - 1, max lambda_max + 1] to a bracket of 1e-14 (relative)
 * or 200 halvings. */
static double fermi(double x) /* 1/(1+e^x) */
{
    if (x > 0) { double t = exp(-x); return t / (1.0 + t); }
    return 1.0 / (1.0 + exp(x));
}
/* f ln f + (1-f) ln(1-f) at x = beta (lambda - lambda_f):
 * = -log1p(t) - |x| t/(1+t), t = exp(-|x|). */
static double fermi_entropy_term(double x)
{
    double ax = fabs(x), t = exp(-ax);
    return -log1p(t) - ax * t / (1.0 + t);
}

double sqo_electron_count(const double *lam, const double *w, long nq, double beta, double lf)
{
    double s = 0.0;
    for (long q = 0; q < nq; q++) s += w[q] * fermi(beta * (lam[q] - lf));
    return 2.0 * s;
}

int sqo_density_fermi_energy(const double *lam, const double *w, int n_nodes, int nqp, double beta, double nelec,
                             double dV, double *lambda_f, double *rho, double *eband, double *mts)
{
    long nq = (long)n_nodes * nqp;
    if (nq <= 0) return SQO_ERR_ARG;
    double lo = lam[0], hi = lam[0];
    for (long q = 0; q < nq; q++) { if (lam[q] < lo) lo = lam[q]; if (lam[q] > hi) hi = lam[q]; }
    lo -= 1.0; hi += 1.0;
    double glo = sqo_electron_count(lam, w, nq, beta, lo) - nelec;
    double ghi = sqo_electron_count(lam, w, nq, beta, hi) - nelec;
    if (glo > 0 || ghi < 0) return SQO_ERR_RANGE;
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (hi - lo <= 1e-14 * (fabs(mid) > 1.0 ? fabs(mid) : 1.0)) break;
        double g = sqo_electron_count(lam, w, nq, beta, mid) - nelec;
        if (g > 0) hi = mid; else lo = mid;
    }
    double lf = 0.5 * (lo + hi);
    *lambda_f = lf;
    double eb = 0.0, ts = 0.0;
    for (int p = 0; p < n_nodes; p++) {
        double s = 0.0;
        for (int i = 0; i < nqp; i++) {
            long q = (long)p * nqp + i;
            double x = beta * (lam[q] - lf);
            double f = fermi(x);
            s += w[q] * f;
            eb += w[q] * lam[q] * f;
            ts += w[q] * fermi_entropy_term(x);
        }
        if (rho) rho[p] = 2.0 * s / dV;
    }
    *eband = 2.0 * eb;
    *mts = 2.0 * ts / beta;
    return SQO_OK;
}

/* ------------------------------------------------------------- accessors */
typedef struct {
    int npts[3], nloc[3], order, nqp, ntype, nAtoms, noetot, IP_length, nIP, nc, center, bc;
    double range[3], delta[3], dV, beta, rcut;
    double coeffs[3][SQO_MAXORDER + 1];
    double cutoff[SQO_MAXTYPE][3];
    double gamma[SQO_MAXTYPE][4], rc[SQO_MAXTYPE][4];
    int lmax[SQO_MAXTYPE], lloc[SQO_MAXTYPE], psd_size[SQO_MAXTYPE], nimages[SQO_MAXTYPE];
} sqo_info_t;

/* ---------------------------------------------------------- SCF fields
 *
 * The reference builds the pieces of its SCF but never assembles them: the
 * periodic Poisson operator (Create_Laplacian, lsdft_initObjs.cc:309-388),
 * the pseudocharge construction b = -(1/4pi) Lap_h V_loc
 * (Calculate_PseudochargeCutoff, 628-817), the local-overlap boxes of every
 * atom image (GetInfluencingAtoms_Local, 1409-1665) and the isolated-atom
 * valence density of the .pot file (lsdft_readfiles.cc:1026-1040, spline at
 * lsdft_initObjs.cc:590).  The functions below assemble them on the
 * periodic cell: the (N0-1) x (N1-1) x (N2-1) distinct grid points (the
 * atom images repeat every 2R = (N-1) h, so grid plane N-1 repeats plane 0).
 */
void sqo_set_cell_period(sqo_sys *S, int enable) { S->cell_period = enable ? 1 : 0; }

/* beta, history, frequency, TOL_SCF, TOL_POISSON, MAXIT_SCF */
void sqo_scf_parameters(sqo_sys *S, double *out)
{
    out[0] = S->mix_beta; out[1] = S->mix_history; out[2] = S->pulay_frequency;
    out[3] = S->tol_scf; out[4] = S->tol_poisson; out[5] = S->maxit_scf;
}

/* V_loc of type at at radius r: the spline inside the radial table, -Z/r past it. */
static int vloc_at(const sqo_sys *S, int at, double r, double *out)
{
    const sqo_psd *P = &S->psd[at];
    if (r > P->r[P->size - 1]) { *out = -(double)S->noe[at] / r; return SQO_OK; }
    return sqo_spline_eval(P->r, P->Vloc, P->size, r, P->YD_Vloc, out);
}

/* Pseudocharge box of one atom image, GetInfluencingAtoms_Local
 * (lsdft_initObjs.cc:1409-1665): centre index roundf((x0 + R)/h), half-width
 * ceil(cutoff/h + 0.5). */
static void psc_box(const sqo_sys *S, int at, const double X[3], int lo[3], int hi[3])
{
    for (int d = 0; d < 3; d++) {
        int off = (int)ceil(S->cutoff[at][d] / S->delta[d] + 0.5);
        int c = grid_round((X[d] + S->range[d]) / S->delta[d]);
        lo[d] = c - off;
        hi[d] = c + off;
    }
}

/* b_J = -(1/4pi) Lap_h V_J on the box lo..hi of the image at X, with the
 * expression and order of lsdft_initObjs.cc:725-735; out is the box, i fastest. */
static int psc_image(const sqo_sys *S, int at, const double X[3], const int lo[3], const int hi[3], double *out,
                     double *vout)
{
    int o = S->order, n[3], e[3];
    for (int d = 0; d < 3; d++) { n[d] = hi[d] - lo[d] + 1; e[d] = n[d] + 2 * o; }
    double *V = dalloc((size_t)e[0] * e[1] * e[2]);
#define VV(k, j, i) V[((size_t)(k) * e[1] + (j)) * e[0] + (i)]
    for (int k = 0; k < e[2]; k++)
        for (int j = 0; j < e[1]; j++)
            for (int i = 0; i < e[0]; i++) {
                double x = S->delta[0] * (lo[0] - o + i) - S->range[0] - X[0];
                double y = S->delta[1] * (lo[1] - o + j) - S->range[1] - X[1];
                double z = S->delta[2] * (lo[2] - o + k) - S->range[2] - X[2];
                double r = sqrt(x * x + y * y + z * z);
                if (vloc_at(S, at, r, &VV(k, j, i))) { free(V); return SQO_ERR_RANGE; }
            }
    const double *scx = S->stencil_coeffs[0], *scy = S->stencil_coeffs[1], *scz = S->stencil_coeffs[2];
    for (int k = o; k < n[2] + o; k++)
        for (int j = o; j < n[1] + o; j++)
            for (int i = o; i < n[0] + o; i++) {
                double v = VV(k, j, i) * (scx[0] + scy[0] + scz[0]);
                for (int a = 1; a <= o; a++)
                    v += (VV(k, j, i - a) + VV(k, j, i + a)) * scx[a] +
                         (VV(k, j - a, i) + VV(k, j + a, i)) * scy[a] +
                         (VV(k - a, j, i) + VV(k + a, j, i)) * scz[a];
                size_t q = ((size_t)(k - o) * n[1] + (j - o)) * n[0] + (i - o);
                out[q] = v;
                if (vout) vout[q] = VV(k, j, i);
            }
#undef VV
    free(V);
    return SQO_OK;
}

/* The pseudocharge on the periodic cell, b[(k*(N1-1)+j)*(N0-1)+i]: the sum
 * over atoms (no phantom) and their images (shifts 2R, image range
 * ceil(cutoff/R) as lsdft_initObjs.cc:1490-1499) of b_J on the part of each
 * box inside the cell, then scaled by -N_e/(sum b dV) so the cell is neutral
 * (the truncation at the cutoff leaves |sum b dV + N_e|/N_e <= TOL_PSEUDOCHARGE).
 * eself = (scale/2) sum_J sum_box b_J V_J dV over each atom's whole box. */
int sqo_pseudocharge(sqo_sys *S, double *b, double *eself, double *scale)
{
    const int G[3] = {S->npts[0] - 1, S->npts[1] - 1, S->npts[2] - 1};
    const size_t ng = (size_t)G[0] * G[1] * G[2];
    memset(b, 0, sizeof(double) * ng);
    double es = 0.0;
    for (int at = 0; at < S->ntype; at++) {
        int Imax[3];
        image_range(S, at, Imax);
        for (int p = S->startPos[at] / 3; p <= S->endPos[at] / 3; p++) {
            const double *A = &S->atompos[3 * p];
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        double X[3] = {A[0] + PP * 2.0 * S->range[0], A[1] + QQ * 2.0 * S->range[1],
                                       A[2] + RR * 2.0 * S->range[2]};
                        int lo[3], hi[3], clo[3], chi[3], empty = 0;
                        psc_box(S, at, X, lo, hi);
                        for (int d = 0; d < 3; d++) {
                            clo[d] = lo[d] > 0 ? lo[d] : 0;
                            chi[d] = hi[d] < G[d] - 1 ? hi[d] : G[d] - 1;
                            empty |= clo[d] > chi[d];
                        }
                        const int self = PP == 0 && QQ == 0 && RR == 0;
                        if (empty && !self) continue;
                        int n[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
                        size_t nb = (size_t)n[0] * n[1] * n[2];
                        double *bj = dalloc(nb), *vj = dalloc(nb);
                        if (psc_image(S, at, X, lo, hi, bj, vj)) { free(bj); free(vj); return SQO_ERR_RANGE; }
                        if (self)
                            for (size_t q = 0; q < nb; q++) es += bj[q] * vj[q];
                        if (!empty)
                            for (int k = clo[2]; k <= chi[2]; k++)
                                for (int j = clo[1]; j <= chi[1]; j++)
                                    for (int i = clo[0]; i <= chi[0]; i++)
                                        b[((size_t)k * G[1] + j) * G[0] + i] +=
                                            bj[((size_t)(k - lo[2]) * n[1] + (j - lo[1])) * n[0] + (i - lo[0])];
                        free(bj);
                        free(vj);
                    }
        }
    }
    double tot = 0.0;
    for (size_t q = 0; q < ng; q++) tot += b[q];
    tot *= S->dV;
    const double sc = -(double)S->noetot / tot;
    for (size_t q = 0; q < ng; q++) b[q] *= sc;
    *scale = sc;
    *eself = 0.5 * es * S->dV * sc;
    return SQO_OK;
}

/* Superposition of the isolated-atom valence densities (the .pot "Valence
 * charge" block over 4 pi r^2, lsdft_readfiles.cc:1026-1040) on the same
 * boxes, scaled to N_e electrons: the initial SCF density. */
int sqo_atomic_density(sqo_sys *S, double *rho, double *scale)
{
    const int G[3] = {S->npts[0] - 1, S->npts[1] - 1, S->npts[2] - 1};
    const size_t ng = (size_t)G[0] * G[1] * G[2];
    memset(rho, 0, sizeof(double) * ng);
    for (int at = 0; at < S->ntype; at++) {
        const sqo_psd *P = &S->psd[at];
        int Imax[3];
        image_range(S, at, Imax);
        for (int p = S->startPos[at] / 3; p <= S->endPos[at] / 3; p++) {
            const double *A = &S->atompos[3 * p];
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        double X[3] = {A[0] + PP * 2.0 * S->range[0], A[1] + QQ * 2.0 * S->range[1],
                                       A[2] + RR * 2.0 * S->range[2]};
                        int lo[3], hi[3];
                        psc_box(S, at, X, lo, hi);
                        for (int d = 0; d < 3; d++) {
                            if (lo[d] < 0) lo[d] = 0;
                            if (hi[d] > G[d] - 1) hi[d] = G[d] - 1;
                        }
                        for (int k = lo[2]; k <= hi[2]; k++)
                            for (int j = lo[1]; j <= hi[1]; j++)
                                for (int i = lo[0]; i <= hi[0]; i++) {
                                    double x = S->delta[0] * i - S->range[0] - X[0];
                                    double y = S->delta[1] * j - S->range[1] - X[1];
                                    double z = S->delta[2] * k - S->range[2] - X[2];
                                    double r = sqrt(x * x + y * y + z * z), v = 0.0;
                                    if (r <= P->r[P->size - 1] &&
                                        sqo_spline_eval(P->r, P->uu, P->size, r, P->YD_uu, &v))
                                        return SQO_ERR_RANGE;
                                    rho[((size_t)k * G[1] + j) * G[0] + i] += v;
                                }
                    }
        }
    }
    double tot = 0.0;
    for (size_t q = 0; q < ng; q++) tot += rho[q];
    tot *= S->dV;
    const double sc = (double)S->noetot / tot;
    for (size_t q = 0; q < ng; q++) rho[q] *= sc;
    *scale = sc;
    return SQO_OK;
}

/* ------------------------------------------------------------ forces
 * The reference declares its force routines only (Calculate_localforce,
 * Calculate_nonlocalforce, CalculateGradientDensityMatrixColumn,
 * CalculateNonlocalForceContribution, lsdft.h:188-191,223-227); it builds the
 * projectors on the force overlaps (CalculateNonlocalProjectorsForces,
 * lsdft_initObjs.cc:1044-1114) and the FD gradient weights
 * (lsdft_readfiles.cc:490-497).  Written here as the derivative, with respect
 * to an atom's position R_J, of the energy the SCF reports,
 *   E = E_band + E_xc - int V_xc rho + (1/2) int phi (b - rho) - E_self - TS,
 * at self-consistency (Hellmann-Feynman: only the explicit R dependence of b,
 * E_self and V_nl counts; the density terms are stationary):
 *
 * local    F_J = f dV sum_{J'} (B g_J')^T phi  -  (f dV / 2) [(B g_s)^T V_s + b_s^T g_s]
 *          with B = -(1/4pi) Lap_h (the pseudocharge stencil), g = grad V_loc
 *          (analytic: the V_loc spline's derivative, Z/r^2 past the table)
 *          sampled on each image box J' of atom J (the part inside the cell,
 *          as b is placed), and the self term over the atom's own whole box
 *          (b_s, V_s unscaled, f the neutrality scale of b);
 * nonlocal F_J = 4 sum_c sum_{J' of J, c in box J'} sum_lm grad chi_J'lm(c) *
 *                gamma_l <chi_J,lm, D_c>
 *          from E_nl = 2 tr(V_nl D), V_nl = sum gamma chi chi^T (gamma carries
 *          dV), the gradient taken on the node side of the symmetric D; D_c =
 *          Q f(T) e_1 is the node's Gauss-quadrature density-matrix column (Q
 *          the Lanczos vectors, T the tridiagonal of its rule), so
 *          <chi, D_c>_c = rho_c dV / 2 for chi = e_c.  chi_J sums every image
 *          of atom J in the cube, as the Hamiltonian's alpha does. */

/* V_loc'(r) (x, y, z)/r; zero at r = 0. */
static int vloc_grad(const sqo_sys *S, int at, double x, double y, double z, double g[3])
{
    const sqo_psd *P = &S->psd[at];
    double r = sqrt(x * x + y * y + z * z), v, dv;
    if (r > P->r[P->size - 1]) dv = (double)S->noe[at] / (r * r);
    else if (spline_eval_d(P->r, P->Vloc, P->size, r, P->YD_Vloc, &v, &dv)) return SQO_ERR_RANGE;
    if (r < 1e-10) { g[0] = g[1] = g[2] = 0.0; return SQO_OK; }
    g[0] = dv * x / r; g[1] = dv * y / r; g[2] = dv * z / r;
    return SQO_OK;
}

/* Gradient of the real harmonic C P_lm(x,y,z) / r^l of sqo_ylm:
 * C (grad P / r^l - l P x / r^(l+2)); zero for r < 1e-10 (where the
 * reference sets Y_lm = 0 for l > 0). */
static void ylm_grad(double x, double y, double z, int l, int m, double g[3])
{
    double r2 = x * x + y * y + z * z, r = sqrt(r2);
    g[0] = g[1] = g[2] = 0.0;
    if (l == 0 || r < 1e-10) return;
    double C = 0.0, P = 0.0, d[3] = {0.0, 0.0, 0.0};
    double xx = x * x, yy = y * y, zz = z * z;
    switch (l * 10 + m + 3) {
    case 12: C = 0.488602511902920; P = y; d[1] = 1; break;
    case 13: C = 0.488602511902920; P = z; d[2] = 1; break;
    case 14: C = 0.488602511902920; P = x; d[0] = 1; break;
    case 21: C = 1.092548430592079; P = x * y; d[0] = y; d[1] = x; break;
    case 22: C = 1.092548430592079; P = y * z; d[1] = z; d[2] = y; break;
    case 23: C = 0.315391565252520; P = -xx - yy + 2 * zz; d[0] = -2 * x; d[1] = -2 * y; d[2] = 4 * z; break;
    case 24: C = 1.092548430592079; P = z * x; d[0] = z; d[2] = x; break;
    case 25: C = 0.546274215296040; P = xx - yy; d[0] = 2 * x; d[1] = -2 * y; break;
    case 30: C = 0.590043589926644; P = (3 * xx - yy) * y; d[0] = 6 * x * y; d[1] = 3 * xx - 3 * yy; break;
    case 31: C = 2.890611442640554; P = x * y * z; d[0] = y * z; d[1] = x * z; d[2] = x * y; break;
    case 32: C = 0.457045799464466; P = y * (4 * zz - xx - yy); d[0] = -2 * x * y; d[1] = 4 * zz - xx - 3 * yy; d[2] = 8 * y * z; break;
    case 33: C = 0.373176332590115; P = z * (2 * zz - 3 * xx - 3 * yy); d[0] = -6 * x * z; d[1] = -6 * y * z; d[2] = 6 * zz - 3 * xx - 3 * yy; break;
    case 34: C = 0.457045799464466; P = x * (4 * zz - xx - yy); d[0] = 4 * zz - 3 * xx - yy; d[1] = -2 * x * y; d[2] = 8 * x * z; break;
    case 35: C = 1.445305721320277; P = z * (xx - yy); d[0] = 2 * x * z; d[1] = -2 * y * z; d[2] = xx - yy; break;
    case 36: C = 0.590043589926644; P = x * (xx - 3 * yy); d[0] = 3 * xx - 3 * yy; d[1] = -6 * x * y; break;
    default: return;
    }
    double rl = 1.0;
    for (int k = 0; k < l; k++) rl *= r;
    double X[3] = {x, y, z};
    for (int k = 0; k < 3; k++) g[k] = C * (d[k] / rl - l * P * X[k] / (rl * r2));
}

void sqo_ylm_grad(double x, double y, double z, int l, int m, double *g) { ylm_grad(x, y, z, l, m, g); }

/* chi_lm = U_l dV_l(r) Y_lm at (x, y, z) from the atom and its gradient
 * S'(r) (x/r) Y + S grad Y (zero at r < 1e-10). */
int sqo_chi_grad(sqo_sys *S, int at, int l, int m, double x, double y, double z, double *val, double *g)
{
    const sqo_psd *P = &S->psd[at];
    double r = sqrt(x * x + y * y + z * z), u, du, gy[3];
    int e2 = 0;
    if (spline_eval_d(P->r, P->UDeltaV[l], P->size, r, P->YD_UDeltaV[l], &u, &du)) return SQO_ERR_RANGE;
    double Y = sqo_ylm(x, y, z, l, m, &e2);
    if (e2) return e2;
    *val = u * Y;
    g[0] = g[1] = g[2] = 0.0;
    if (r < 1e-10) return SQO_OK;
    ylm_grad(x, y, z, l, m, gy);
    double X[3] = {x, y, z};
    for (int k = 0; k < 3; k++) g[k] = du * (X[k] / r) * Y + u * gy[k];
    return SQO_OK;
}

/* c = f(T) e_1 = Z f(Lambda) Z^T e_1 for the n x n tridiagonal (diag, subdiag)
 * with f(l) = 1/(1 + exp(beta (l - lf))): column 0 of the Gauss-quadrature
 * density matrix in the Lanczos basis. */
int sqo_fermi_column(const double *diag, const double *subdiag, int n, double beta, double lf, double *c)
{
    double *d = dalloc(n), *e = dalloc(n), *z = dalloc((size_t)n * n);
    for (int i = 0; i < n; i++) { d[i] = diag[i]; e[i] = subdiag[i]; z[(size_t)i * n + i] = 1.0; }
    int rc = ql_full(d, e, z, n);
    if (!rc)
        for (int k = 0; k < n; k++) {
            double s = 0.0;
            for (int j = 0; j < n; j++) s += z[(size_t)k * n + j] * fermi(beta * (d[j] - lf)) * z[j];
            c[k] = s;
        }
    free(d); free(e); free(z);
    return rc;
}

/* Local (pseudocharge) forces, F[3*nAtoms] (formula above).  phi is the
 * electrostatic potential on the periodic cell, scale the factor f of b. */
int sqo_local_forces(sqo_sys *S, const double *phi, double scale, double *F)
{
    const int G[3] = {S->npts[0] - 1, S->npts[1] - 1, S->npts[2] - 1};
    const int o = S->order;
    const double *scx = S->stencil_coeffs[0], *scy = S->stencil_coeffs[1], *scz = S->stencil_coeffs[2];
    memset(F, 0, sizeof(double) * 3 * S->nAtoms);
    for (int at = 0; at < S->ntype; at++) {
        int Imax[3];
        image_range(S, at, Imax);
        for (int p = S->startPos[at] / 3; p <= S->endPos[at] / 3; p++) {
            const double *A = &S->atompos[3 * p];
            double fphi[3] = {0, 0, 0}, fself[3] = {0, 0, 0};
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        double X[3] = {A[0] + PP * 2.0 * S->range[0], A[1] + QQ * 2.0 * S->range[1],
                                       A[2] + RR * 2.0 * S->range[2]};
                        int lo[3], hi[3], clo[3], chi[3], empty = 0, n[3], e[3];
                        psc_box(S, at, X, lo, hi);
                        for (int d = 0; d < 3; d++) {
                            clo[d] = lo[d] > 0 ? lo[d] : 0;
                            chi[d] = hi[d] < G[d] - 1 ? hi[d] : G[d] - 1;
                            empty |= clo[d] > chi[d];
                            n[d] = hi[d] - lo[d] + 1;
                            e[d] = n[d] + 2 * o;
                        }
                        const int self = PP == 0 && QQ == 0 && RR == 0;
                        if (empty && !self) continue;
                        size_t ne = (size_t)e[0] * e[1] * e[2], nb = (size_t)n[0] * n[1] * n[2];
                        double *g = dalloc(3 * ne), *bj = NULL, *vj = NULL;
                        for (int k = 0; k < e[2]; k++)
                            for (int j = 0; j < e[1]; j++)
                                for (int i = 0; i < e[0]; i++) {
                                    double x = S->delta[0] * (lo[0] - o + i) - S->range[0] - X[0];
                                    double y = S->delta[1] * (lo[1] - o + j) - S->range[1] - X[1];
                                    double z = S->delta[2] * (lo[2] - o + k) - S->range[2] - X[2];
                                    size_t q = ((size_t)k * e[1] + j) * e[0] + i;
                                    if (vloc_grad(S, at, x, y, z, g + 3 * q)) { free(g); return SQO_ERR_RANGE; }
                                }
                        if (self) {
                            bj = dalloc(nb); vj = dalloc(nb);
                            if (psc_image(S, at, X, lo, hi, bj, vj)) { free(g); free(bj); free(vj); return SQO_ERR_RANGE; }
                        }
#define GG(k, j, i, d) g[3 * (((size_t)(k) * e[1] + (j)) * e[0] + (i)) + (d)]
                        for (int k = o; k < n[2] + o; k++)
                            for (int j = o; j < n[1] + o; j++)
                                for (int i = o; i < n[0] + o; i++) {
                                    int gi = lo[0] - o + i, gj = lo[1] - o + j, gk = lo[2] - o + k;
                                    int inside = !empty && gi >= clo[0] && gi <= chi[0] && gj >= clo[1] && gj <= chi[1] &&
                                                 gk >= clo[2] && gk <= chi[2];
                                    if (!inside && !self) continue;
                                    size_t q = ((size_t)(k - o) * n[1] + (j - o)) * n[0] + (i - o);
                                    for (int d = 0; d < 3; d++) {
                                        double v = GG(k, j, i, d) * (scx[0] + scy[0] + scz[0]);
                                        for (int a = 1; a <= o; a++)
                                            v += (GG(k, j, i - a, d) + GG(k, j, i + a, d)) * scx[a] +
                                                 (GG(k, j - a, i, d) + GG(k, j + a, i, d)) * scy[a] +
                                                 (GG(k - a, j, i, d) + GG(k + a, j, i, d)) * scz[a];
                                        if (inside) fphi[d] += v * phi[((size_t)gk * G[1] + gj) * G[0] + gi];
                                        if (self) fself[d] += v * vj[q] + bj[q] * GG(k, j, i, d);
                                    }
                                }
#undef GG
                        free(g); free(bj); free(vj);
                    }
            for (int d = 0; d < 3; d++) F[3 * p + d] = scale * S->dV * fphi[d] - 0.5 * scale * S->dV * fself[d];
        }
    }
    return SQO_OK;
}

/* Nonlocal forces, F[3*nAtoms] (formula above), KB operator, from the nodes
 * node_ids (NULL: every node of the rank box) with V_eff as for
 * sqo_nodes_and_weights.  Fnode (optional) receives each node's
 * contribution [n_nodes][nAtoms][3]; F sums them in node order. */
int sqo_nonlocal_forces(sqo_sys *S, const double *veff_field, double veff_const, double beta, double lf,
                        const int *node_ids, int n_nodes, int nthreads, double *F, double *Fnode)
{
    if (!S->have_nl) return SQO_ERR_ARG;
    build_cube_laplacian(S);
    int lx = S->box[3], ly = S->box[4], lz = S->box[5];
    int total = lx * ly * lz, na = S->nAtoms, nq = S->nqp, nc = S->nc;
    if (!node_ids) n_nodes = total;
    double *own = NULL;
    if (!Fnode) own = Fnode = dalloc((size_t)n_nodes * na * 3);
    else memset(Fnode, 0, sizeof(double) * (size_t)n_nodes * na * 3);
    int err = SQO_OK;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        node_ws W;
        int e = ws_init(&W, S);
        double *a = dalloc(nq + 1), *b = dalloc(nq + 1), *cf = dalloc(nq), *Q = dalloc((size_t)nq * nc), *dcol = dalloc(nc);
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < n_nodes; r++) {
            if (e) continue;
            int id = node_ids ? node_ids[r] : r;
            if (id < 0 || id >= total) { e = SQO_ERR_ARG; continue; }
            int I = S->box[0] + id % lx, J = S->box[1] + (id / lx) % ly, K = S->box[2] + id / (lx * ly);
            /* images whose box holds the node */
            int near = 0;
            for (int at = 0; at < S->ntype && !near; at++)
                for (int c = 0; c < S->ovl[at].n && !near; c++) {
                    const sqo_overlap *O = &S->ovl[at];
                    near = I >= O->xs[c] && I <= O->xe[c] && J >= O->ys[c] && J <= O->ye[c] && K >= O->zs[c] && K <= O->ze[c];
                }
            if (!near) continue;
            cube_veff(S, veff_field, veff_const, K, J, I, W.veff);
            rebuild_operator(S, &W);
            int rc = lanczos_node(S, 1, &W, K, J, I, a, b, Q);
            if (!rc) rc = sqo_fermi_column(a, b, nq, beta, lf, cf);
            if (rc) { e = rc; continue; }
            for (int i = 0; i < nc; i++) {
                double s = 0.0;
                for (int k = 0; k < nq; k++) s += cf[k] * Q[(size_t)k * nc + i];
                dcol[i] = s;
            }
            nonlocal_apply(S, 1, W.alpha, K, J, I, dcol, NULL);   /* alpha = gamma <chi_atom, D_c> */
            double *fn = Fnode + (size_t)r * na * 3;

            for (int at = 0; at < S->ntype; at++) {
                const sqo_overlap *O = &S->ovl[at];
                int lloc = S->localPsd[at];
                for (int c = 0; c < O->n; c++) {
                    if (!(I >= O->xs[c] && I <= O->xe[c] && J >= O->ys[c] && J <= O->ye[c] && K >= O->zs[c] && K <= O->ze[c]))
                        continue;
                    int p = O->xindex[c] / 3, slot = S->IP_index[p];
                    double x = S->delta[0] * I - S->range[0] - O->X0[c];
                    double y = S->delta[1] * J - S->range[1] - O->Y0[c];
                    double z = S->delta[2] * K - S->range[2] - O->Z0[c];
                    for (int l = 0; l <= S->psd[at].lmax; l++) {
                        if (l == lloc) continue;
                        for (int m = -l; m <= l; m++, slot++) {
                            double v, g[3];
                            if (sqo_chi_grad(S, at, l, m, x, y, z, &v, g)) { e = SQO_ERR_RANGE; continue; }
                            for (int d = 0; d < 3; d++) fn[3 * p + d] += 4.0 * g[d] * W.alpha[slot];
                        }
                    }
                }
            }
        }
        if (e) {
#pragma omp critical
            err = e;
        }
        free(a); free(b); free(cf); free(Q); free(dcol);
        ws_free(&W);
    }
    memset(F, 0, sizeof(double) * 3 * na);
    for (int r = 0; r < n_nodes; r++)
        for (int q = 0; q < 3 * na; q++) F[q] += Fnode[(size_t)r * na * 3 + q];
    free(own);
    return err;
}

void sqo_info(sqo_sys *S, sqo_info_t *I)
{
    memset(I, 0, sizeof *I);
    build_cube_laplacian(S);
    for (int d = 0; d < 3; d++) {
        I->npts[d] = S->npts[d]; I->nloc[d] = S->nloc[d]; I->range[d] = S->range[d]; I->delta[d] = S->delta[d];
        for (int p = 0; p <= S->order; p++) I->coeffs[d][p] = S->coeffs[d][p];
    }
    I->order = S->order; I->nqp = S->nqp; I->ntype = S->ntype; I->nAtoms = S->nAtoms; I->noetot = S->noetot;
    I->IP_length = S->IP_length; I->nIP = S->nIP; I->nc = S->nc; I->center = S->center; I->bc = S->bc;
    I->dV = S->dV; I->beta = S->beta; I->rcut = S->rcut;
    for (int at = 0; at < S->ntype; at++) {
        for (int d = 0; d < 3; d++) I->cutoff[at][d] = S->cutoff[at][d];
        for (int l = 0; l < 4; l++) { I->gamma[at][l] = S->psd[at].Gamma[l]; I->rc[at][l] = S->psd[at].rc[l]; }
        I->lmax[at] = S->psd[at].lmax; I->lloc[at] = S->localPsd[at]; I->psd_size[at] = S->psd[at].size;
        I->nimages[at] = S->have_nl ? S->ovl[at].n : -1;
    }
}

int sqo_info_size(void) { return (int)sizeof(sqo_info_t); }

/* Radial tables of type at: r, Vloc, and UDeltaV/U/YD for channel l (any may be NULL). */
int sqo_get_radial(sqo_sys *S, int at, int l, double *r, double *vloc, double *udv, double *yd_udv, double *u,
                   double *yd_u)
{
    if (at < 0 || at >= S->ntype || l < 0 || l > S->psd[at].lmax) return SQO_ERR_ARG;
    sqo_psd *P = &S->psd[at];
    size_t n = sizeof(double) * P->size;
    if (r) memcpy(r, P->r, n);
    if (vloc) memcpy(vloc, P->Vloc, n);
    if (udv) memcpy(udv, P->UDeltaV[l], n);
    if (yd_udv) memcpy(yd_udv, P->YD_UDeltaV[l], n);
    if (u) memcpy(u, P->U[l], n);
    if (yd_u) memcpy(yd_u, P->YD_U[l], n);
    return SQO_OK;
}

/* Image c of type at: center (3), box (xs,ys,zs,xe,ye,ze), atom index p, ndim */
int sqo_get_image(sqo_sys *S, int at, int c, double *center, int *box, int *atom, int *ndim)
{
    if (!S->have_nl || at < 0 || at >= S->ntype || c < 0 || c >= S->ovl[at].n) return SQO_ERR_ARG;
    sqo_overlap *O = &S->ovl[at];
    center[0] = O->X0[c]; center[1] = O->Y0[c]; center[2] = O->Z0[c];
    box[0] = O->xs[c]; box[1] = O->ys[c]; box[2] = O->zs[c];
    box[3] = O->xe[c]; box[4] = O->ye[c]; box[5] = O->ze[c];
    *atom = O->xindex[c] / 3;
    *ndim = O->ndim[c];
    return SQO_OK;
}

int sqo_get_chi(sqo_sys *S, int at, int c, double *chi, int *lId, int *mId, int *kId, int *jId, int *iId)
{
    if (!S->have_nl || at < 0 || at >= S->ntype || c < 0 || c >= S->ovl[at].n) return SQO_ERR_ARG;
    sqo_overlap *O = &S->ovl[at];
    size_t n = O->ndim[c];
    if (chi) memcpy(chi, O->chi[c], n * sizeof(double));
    if (lId) memcpy(lId, O->lId[c], n * sizeof(int));
    if (mId) memcpy(mId, O->mId[c], n * sizeof(int));
    if (kId) memcpy(kId, O->kId[c], n * sizeof(int));
    if (jId) memcpy(jId, O->jId[c], n * sizeof(int));
    if (iId) memcpy(iId, O->iId[c], n * sizeof(int));
    return SQO_OK;
}

int sqo_get_ip_index(sqo_sys *S, int *out) { memcpy(out, S->IP_index, sizeof(int) * S->nIP); return S->nIP; }

/* Atom positions of the .ion file (3*nAtoms) */
int sqo_get_valence_density(sqo_sys *S, int at, double *out)
{
    if (at < 0 || at >= S->ntype) return -1;
    memcpy(out, S->psd[at].uu, sizeof(double) * S->psd[at].size);
    return S->psd[at].size;
}

int sqo_get_atoms(sqo_sys *S, double *pos) { memcpy(pos, S->atompos, sizeof(double) * 3 * S->nAtoms); return S->nAtoms; }

/* Dense cube Laplacian values (CSR) for brute-force checks. */
int sqo_get_cube_csr(sqo_sys *S, int *rowptr, int *col, double *val)
{
    build_cube_laplacian(S);
    int nnz = S->rowptr[S->nc];
    if (rowptr) memcpy(rowptr, S->rowptr, sizeof(int) * (S->nc + 1));
    if (col) memcpy(col, S->colidx, sizeof(int) * nnz);
    if (val) memcpy(val, S->val, sizeof(double) * nnz);
    return nnz;
}

/* Override the Lanczos degree (LANCZOS_DEGREE) after loading. */
void sqo_set_nqp(sqo_sys *S, int nqp) { S->nqp = nqp; }

void sqo_free(sqo_sys *S)
{
    if (!S) return;
    free_nl(S);
    free(S->atompos); free(S->atomposSC); free(S->IP_index);
    free(S->rowptr); free(S->colidx); free(S->val);
    for (int at = 0; at < SQO_MAXTYPE; at++) {
        sqo_psd *P = &S->psd[at];
        free(P->r); free(P->uu); free(P->Vloc); free(P->YD_Vloc); free(P->YD_uu);
        for (int l = 0; l < 4; l++) { free(P->V[l]); free(P->U[l]); free(P->UDeltaV[l]); free(P->YD_UDeltaV[l]); free(P->YD_U[l]); }
    }
    free(S);
}
