/*
 * sq_oracle_internal.h -- the CPU oracle's data structures, shared by its
 * restatement (sq_oracle.c) and its optimised CPU baseline (sq_cpu_fast.c).
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/.
 */
#ifndef SQ_ORACLE_INTERNAL_H
#define SQ_ORACLE_INTERNAL_H

#define SQO_MAXTYPE 8
#define SQO_MAXORDER 10

enum { SQO_OK = 0, SQO_ERR_IO = 1, SQO_ERR_FORMAT = 2, SQO_ERR_RANGE = 3,
       SQO_ERR_ARG = 4, SQO_ERR_QL = 5, SQO_ERR_ALLOC = 6, SQO_ERR_BREAKDOWN = 7 };

typedef struct {
    int size, lmax;
    double *r, *V[4], *U[4], *uu, *Vloc;
    double rc[4];
    double *UDeltaV[4], *YD_UDeltaV[4], *YD_U[4], *YD_Vloc, *YD_uu;
    double Gamma[4];
} sqo_psd;

typedef struct {
    int n;
    double *X0, *Y0, *Z0;
    int *xs, *ys, *zs, *xe, *ye, *ze, *xindex;
    int *ndim;
    double **chi;
    int **lId, **mId, **kId, **jId, **iId;
} sqo_overlap;

typedef struct sqo_sys {
    /* .inpt (lsdft_readfiles.cc:62-102 defaults, 148-331 keywords) */
    int order, bc, nqp, ntype;
    double range[3], beta, rcut, tol_psc;
    /* SCF keywords (lsdft_readfiles.cc:62-102 defaults, 219-240) */
    double mix_beta, tol_scf, tol_poisson;
    int mix_history, pulay_frequency, maxit_scf;
    int npts[3];
    int localPsd[SQO_MAXTYPE];
    char psdName[SQO_MAXTYPE][64];
    /* derived (lsdft_readfiles.cc:356-497) */
    double delta[3], dV;
    int nloc[3];
    double coeffs[3][SQO_MAXORDER + 1], stencil_coeffs[3][SQO_MAXORDER + 1];
    /* .ion (lsdft_readfiles.cc:677-789) */
    int nAtoms, noetot;
    char atomType[SQO_MAXTYPE][16];
    int noe[SQO_MAXTYPE], noa[SQO_MAXTYPE], startPos[SQO_MAXTYPE], endPos[SQO_MAXTYPE];
    double *atompos;
    sqo_psd psd[SQO_MAXTYPE];
    double cutoff[SQO_MAXTYPE][3];
    /* supercell list (lsdft_initObjs.cc:3031-3041) */
    double *atomposSC;
    int startPosSC[SQO_MAXTYPE], endPosSC[SQO_MAXTYPE];
    int IP_length, nIP, *IP_index;
    /* nonlocal overlap for the current rank box */
    int box[6], have_nl;
    sqo_overlap ovl[SQO_MAXTYPE];
    /* V_eff fields wrap with period N (the DMDA grid, 0) or N-1 (the
     * distinct points of the periodic cell, 1) */
    int cell_period;
    /* cube Laplacian in CSR (lsdft_initObjs.cc:393-475) */
    int nc, *rowptr, *colidx;
    double *val;
    int center;
} sqo_sys;

/* shared entry points of sq_oracle.c used by sq_cpu_fast.c */
int sqo_tridiag_nodes_weights(const double *diag, const double *subdiag, int n, double *lambda, double *w,
                              double *lmin, double *lmax);

#endif
