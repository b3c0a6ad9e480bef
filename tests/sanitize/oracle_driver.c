/* Drives the CPU oracle (oracle/sq_oracle.c) over one deck for an
 * AddressSanitizer + UndefinedBehaviorSanitizer build: load, projector overlap
 * of the whole grid and of the periodic cell, with and without the phantom
 * atom, Lanczos + QL on a few nodes in both nonlocal modes, the pseudocharge,
 * the atomic density, local and nonlocal forces.  Prints checksums. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/sq_oracle_internal.h"

int sqo_load(const char *dir, const char *name, sqo_sys **out);
void sqo_free(sqo_sys *S);
int sqo_set_phantom(sqo_sys *S, int enable, double x, double y, double z);
int sqo_build_nonlocal(sqo_sys *S, const int box[6]);
void sqo_set_cell_period(sqo_sys *S, int enable);
int sqo_nodes_and_weights(sqo_sys *S, int mode, const double *veff_field, double veff_const, const int *node_ids,
                          int n_nodes, double *lam, double *w, double *lmin, double *lmax, int nthreads);
int sqo_pseudocharge(sqo_sys *S, double *b, double *eself, double *scale);
int sqo_atomic_density(sqo_sys *S, double *rho, double *scale);
int sqo_local_forces(sqo_sys *S, const double *phi, double scale, double *F);
int sqo_nonlocal_forces(sqo_sys *S, const double *veff_field, double veff_const, double beta, double lf,
                        const int *node_ids, int n_nodes, int nthreads, double *F, double *Fnode);

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s deck_dir name\n", argv[0]);
        return 2;
    }
    for (int variant = 0; variant < 2; variant++) {
        sqo_sys *S = NULL;
        int rc = sqo_load(argv[1], argv[2], &S);
        if (rc) { fprintf(stderr, "sqo_load: %d\n", rc); return 1; }
        sqo_set_phantom(S, variant == 1, 0.0, 0.0, 0.0);
        const int G[3] = {S->npts[0] - 1, S->npts[1] - 1, S->npts[2] - 1};
        const int cell[6] = {0, 0, 0, G[0], G[1], G[2]};
        sqo_set_cell_period(S, 1);
        if ((rc = sqo_build_nonlocal(S, cell))) { fprintf(stderr, "sqo_build_nonlocal: %d\n", rc); return 1; }
        const long ng = (long)G[0] * G[1] * G[2];
        const int nq = S->nqp, na = S->nAtoms;
        /* the first node, the cell's centre node (the phantom atom's place), the last */
        int ids[3] = {0, (int)(((long)(G[2] / 2) * G[1] + G[1] / 2) * G[0] + G[0] / 2), (int)(ng - 1)};
        double *lam = calloc((size_t)3 * nq, sizeof(double)), *w = calloc((size_t)3 * nq, sizeof(double));
        double lmin[3], lmax[3], chk = 0;
        double *veff = malloc((size_t)ng * sizeof(double));
        for (long q = 0; q < ng; q++) veff[q] = 0.3 + 0.05 * sin(0.11 * (double)q);
        for (int mode = 0; mode < 2; mode++) {
            if ((rc = sqo_nodes_and_weights(S, mode, veff, 0.0, ids, 3, lam, w, lmin, lmax, 2))) {
                fprintf(stderr, "sqo_nodes_and_weights: %d\n", rc);
                return 1;
            }
            for (int k = 0; k < 3 * nq; k++) chk += lam[k] * w[k];
        }
        double *b = malloc((size_t)ng * sizeof(double)), *rho = malloc((size_t)ng * sizeof(double));
        double eself = 0, bs = 0, rs = 0;
        if ((rc = sqo_pseudocharge(S, b, &eself, &bs)) || (rc = sqo_atomic_density(S, rho, &rs))) {
            fprintf(stderr, "pseudocharge / atomic density: %d\n", rc);
            return 1;
        }
        double *F = calloc((size_t)3 * na, sizeof(double)), *Fn = calloc((size_t)3 * na, sizeof(double));
        double *Fnode = calloc((size_t)3 * 3 * na, sizeof(double));
        if ((rc = sqo_local_forces(S, veff, bs, F)) ||
            (rc = sqo_nonlocal_forces(S, veff, 0.0, S->beta, 0.6, ids, 3, 2, Fn, Fnode))) {
            fprintf(stderr, "forces: %d\n", rc);
            return 1;
        }
        double sb = 0, sr = 0, sf = 0;
        for (long q = 0; q < ng; q++) sb += b[q], sr += rho[q];
        for (int k = 0; k < 3 * na; k++) sf += F[k] + Fn[k];
        printf("variant %d: %ld cell points, quadrature checksum %.12e, int b dV %.10f, int rho dV %.10f, "
               "E_self %.10f, sum F %.6e\n", variant, ng, chk, sb * S->dV, sr * S->dV, eself, sf);
        free(lam); free(w); free(veff); free(b); free(rho); free(F); free(Fn); free(Fnode);
        sqo_free(S);
    }
    printf("oracle driver done\n");
    return 0;
}
