/*
 * sq_cpu_fast.c -- an optimised CPU form of the SQ pass, the second CPU
 * baseline beside the reference-faithful restatement in sq_oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu baseline leg times it and tests/
 * check it against the restatement.  The product never links it.
 *
 * Same mathematics as NodesAndWeights_LSSGQ (lsdft_scf.cc:32-82) with the
 * Kleinman-Bylander operator (Mult_HamiltonianVector, lsdft_matvec.cc:155-293),
 * without the reference's per-node costs:
 *   - the stencil is applied matrix-free on a zero-padded copy of the vector
 *     (no MatCopy + MatDiagonalSet of the AIJ operator per node, no SpMV);
 *   - each node lists once the projector images whose box meets its cube and
 *     visits only those (the reference scans every image of the rank domain
 *     with a cube test per entry at every application);
 *   - nodes run in parallel (OpenMP), one scratch set per thread.
 * Built with -O3 for x86-64-v3 and FMA contraction, so its sums round
 * differently from the restatement's (tests compare to rounding).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "sq_oracle_internal.h"

typedef struct {
    int at, c, slot;                 /* type, image, first alpha slot of its atom */
    int lo[3], hi[3];                /* image box (inclusive) clipped to the cube, cube coordinates */
    int bx, by, bnz, box_lo[3];      /* the image's stored box: extents and origin (global) */
} img_ref;

static void apply_h(const sqo_sys *S, int nx, int ny, int nz, int o, const double *cw, double cdiag,
                    const double *veff, const img_ref *imgs, int nimg, double *pad, double *alpha,
                    const double (*gch)[16], const double *x, double *y)
{
    const int px = nx + 2 * o, py = ny + 2 * o;
    const size_t pplane = (size_t)px * py;
    /* x into the zero halo */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            memcpy(pad + ((size_t)(k + o) * py + (j + o)) * px + o, x + ((size_t)k * ny + j) * nx, nx * sizeof(double));
    const double *cx = cw, *cy = cw + SQO_MAXORDER + 1, *cz = cw + 2 * (SQO_MAXORDER + 1);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            const double *pc = pad + ((size_t)(k + o) * py + (j + o)) * px + o;
            const size_t q0 = ((size_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) {
                const double *p = pc + i;
                double s = (cdiag + veff[q0 + i]) * p[0];
                for (int a = 1; a <= o; a++)
                    s += cx[a] * (p[-a] + p[a]) + cy[a] * (p[-a * px] + p[a * px]) +
                         cz[a] * (p[-(long)a * pplane] + p[(long)a * pplane]);
                y[q0 + i] = s;
            }
        }
    /* Kleinman-Bylander: alpha per (atom, l, m) over the images in the cube, then the scatter */
    for (int g = 0; g < nimg; g++) {
        const img_ref *r = &imgs[g];
        const sqo_overlap *O = &S->ovl[r->at];
        const int nlm = O->ndim[r->c] / (r->bx * r->by * r->bnz);
        memset(alpha + r->slot, 0, nlm * sizeof(double));
    }
    for (int g = 0; g < nimg; g++) {
        const img_ref *r = &imgs[g];
        const sqo_overlap *O = &S->ovl[r->at];
        const double *chi = O->chi[r->c];
        const int nlm = O->ndim[r->c] / (r->bx * r->by * r->bnz);
        double acc[16] = {0};
        for (int k = r->lo[2]; k <= r->hi[2]; k++)
            for (int j = r->lo[1]; j <= r->hi[1]; j++)
                for (int i = r->lo[0]; i <= r->hi[0]; i++) {
                    const double v = x[((size_t)k * ny + j) * nx + i];
                    const size_t e = (((size_t)(k - (r->box_lo[2])) * r->by + (j - r->box_lo[1])) * r->bx +
                                      (i - r->box_lo[0])) * nlm;
                    for (int ch = 0; ch < nlm; ch++) acc[ch] += v * chi[e + ch];
                }
        for (int ch = 0; ch < nlm; ch++) alpha[r->slot + ch] += acc[ch] * gch[r->at][ch];
    }
    for (int g = 0; g < nimg; g++) {
        const img_ref *r = &imgs[g];
        const sqo_overlap *O = &S->ovl[r->at];
        const double *chi = O->chi[r->c];
        const int nlm = O->ndim[r->c] / (r->bx * r->by * r->bnz);
        const double *al = alpha + r->slot;
        for (int k = r->lo[2]; k <= r->hi[2]; k++)
            for (int j = r->lo[1]; j <= r->hi[1]; j++)
                for (int i = r->lo[0]; i <= r->hi[0]; i++) {
                    const size_t e = (((size_t)(k - r->box_lo[2]) * r->by + (j - r->box_lo[1])) * r->bx +
                                      (i - r->box_lo[0])) * nlm;
                    double s = 0.0;
                    for (int ch = 0; ch < nlm; ch++) s += chi[e + ch] * al[ch];
                    y[((size_t)k * ny + j) * nx + i] += s;
                }
    }
}

/* As sqo_nodes_and_weights with mode 1 (KB), for the same node box and V_eff
 * conventions. */
int sqo_fast_nodes_and_weights(sqo_sys *S, const double *veff_field, double veff_const, const int *node_ids, int n_nodes,
                               double *lam, double *w, double *lmin, double *lmax, int nthreads)
{
    if (!S->have_nl) return SQO_ERR_ARG;
    const int nx = S->nloc[0], ny = S->nloc[1], nz = S->nloc[2], o = S->order, nq = S->nqp;
    const int hx = (nx - 1) / 2, hy = (ny - 1) / 2, hz = (nz - 1) / 2;
    const size_t nc = (size_t)nx * ny * nz;
    const int lx = S->box[3], ly = S->box[4], lz = S->box[5];
    const int total = lx * ly * lz;
    if (!node_ids) n_nodes = total;
    const int center = (int)((((size_t)hz) * ny + hy) * nx + hx);
    double cw[3 * (SQO_MAXORDER + 1)];
    for (int d = 0; d < 3; d++)
        for (int a = 0; a <= SQO_MAXORDER; a++) cw[d * (SQO_MAXORDER + 1) + a] = a <= o ? S->coeffs[d][a] : 0.0;
    const double cdiag = S->coeffs[0][0] + S->coeffs[1][0] + S->coeffs[2][0];
    double gch[SQO_MAXTYPE][16];
    for (int at = 0; at < S->ntype; at++) {
        int ch = 0;
        for (int l = 0; l <= S->psd[at].lmax; l++) {
            if (l == S->localPsd[at]) continue;
            for (int m = -l; m <= l && ch < 16; m++) gch[at][ch++] = S->psd[at].Gamma[l];
        }
    }
    int nimg_all = 0;
    for (int at = 0; at < S->ntype; at++) nimg_all += S->ovl[at].n;
    const int Nper[3] = {S->npts[0] - S->cell_period, S->npts[1] - S->cell_period, S->npts[2] - S->cell_period};
    int err = SQO_OK;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *v = calloc(nc, sizeof(double)), *vp = calloc(nc, sizeof(double)), *u = calloc(nc, sizeof(double));
        double *veff = calloc(nc, sizeof(double));
        double *pad = calloc((size_t)(nx + 2 * o) * (ny + 2 * o) * (nz + 2 * o), sizeof(double));
        double *alpha = calloc(S->IP_length > 0 ? S->IP_length : 1, sizeof(double));
        double *a = calloc(nq + 1, sizeof(double)), *b = calloc(nq + 1, sizeof(double));
        img_ref *imgs = calloc(nimg_all > 0 ? nimg_all : 1, sizeof(img_ref));
        int e = (v && vp && u && veff && pad && alpha && a && b && imgs) ? SQO_OK : SQO_ERR_ALLOC;
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < n_nodes; r++) {
            if (e) continue;
            const int id = node_ids ? node_ids[r] : r;
            if (id < 0 || id >= total) { e = SQO_ERR_ARG; continue; }
            const int I = S->box[0] + id % lx, J = S->box[1] + (id / lx) % ly, K = S->box[2] + id / (lx * ly);
            const int c0[3] = {I - hx, J - hy, K - hz};
            /* V_eff on the cube (the restatement's cube_veff) */
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++)
                    for (int i = 0; i < nx; i++) {
                        double val = veff_const;
                        if (veff_field) {
                            const int gi = ((c0[0] + i) % Nper[0] + Nper[0]) % Nper[0];
                            const int gj = ((c0[1] + j) % Nper[1] + Nper[1]) % Nper[1];
                            const int gk = ((c0[2] + k) % Nper[2] + Nper[2]) % Nper[2];
                            val = veff_field[((size_t)gk * Nper[1] + gj) * Nper[0] + gi];
                        }
                        veff[((size_t)k * ny + j) * nx + i] = val;
                    }
            /* images meeting the cube */
            int nimg = 0;
            for (int at = 0; at < S->ntype; at++) {
                const sqo_overlap *O = &S->ovl[at];
                for (int c = 0; c < O->n; c++) {
                    const int blo[3] = {O->xs[c], O->ys[c], O->zs[c]}, bhi[3] = {O->xe[c], O->ye[c], O->ze[c]};
                    const int ext[3] = {nx, ny, nz};
                    img_ref rr;
                    int ok = 1;
                    for (int d = 0; d < 3; d++) {
                        const int lo = blo[d] > c0[d] ? blo[d] : c0[d];
                        const int hi = bhi[d] < c0[d] + ext[d] - 1 ? bhi[d] : c0[d] + ext[d] - 1;
                        if (lo > hi) { ok = 0; break; }
                        rr.lo[d] = lo - c0[d];
                        rr.hi[d] = hi - c0[d];
                        rr.box_lo[d] = blo[d] - c0[d];
                    }
                    if (!ok) continue;
                    rr.at = at;
                    rr.c = c;
                    rr.slot = S->IP_index[O->xindex[c] / 3];
                    rr.bx = bhi[0] - blo[0] + 1;
                    rr.by = bhi[1] - blo[1] + 1;
                    rr.bnz = bhi[2] - blo[2] + 1;
                    imgs[nimg++] = rr;
                }
            }
            /* Lanczos (lsdft_scf.cc:115-156) from e_center */
            memset(vp, 0, nc * sizeof(double));
            vp[center] = 1.0;
            apply_h(S, nx, ny, nz, o, cw, cdiag, veff, imgs, nimg, pad, alpha, (const double(*)[16])gch, vp, v);
            double s = 0.0;
            for (size_t q = 0; q < nc; q++) s += vp[q] * v[q];
            a[0] = s;
            double bb = 0.0;
            for (size_t q = 0; q < nc; q++) { v[q] -= a[0] * vp[q]; bb += v[q] * v[q]; }
            b[0] = sqrt(bb);
            if (b[0] == 0.0) { e = SQO_ERR_BREAKDOWN; continue; }
            for (size_t q = 0; q < nc; q++) v[q] *= 1.0 / b[0];
            for (int k = 0; k < nq && !e; k++) {
                apply_h(S, nx, ny, nz, o, cw, cdiag, veff, imgs, nimg, pad, alpha, (const double(*)[16])gch, v, u);
                double aa = 0.0;
                for (size_t q = 0; q < nc; q++) aa += v[q] * u[q];
                a[k + 1] = aa;
                bb = 0.0;
                for (size_t q = 0; q < nc; q++) {
                    const double t = u[q] - aa * v[q] - b[k] * vp[q];
                    vp[q] = v[q];
                    u[q] = t;
                    bb += t * t;
                }
                b[k + 1] = sqrt(bb);
                if (b[k + 1] == 0.0) { e = SQO_ERR_BREAKDOWN; break; }
                const double sc = 1.0 / b[k + 1];
                for (size_t q = 0; q < nc; q++) v[q] = sc * u[q];
            }
            if (e) continue;
            int rc = sqo_tridiag_nodes_weights(a, b, nq, lam + (size_t)r * nq, w + (size_t)r * nq, lmin + r, lmax + r);
            if (rc) e = rc;
        }
        if (e) {
#pragma omp critical
            err = e;
        }
        free(v); free(vp); free(u); free(veff); free(pad); free(alpha); free(a); free(b); free(imgs);
    }
    return err;
}
