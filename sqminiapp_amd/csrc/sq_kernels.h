// sq_kernels.h -- launch parameters and launchers of the SQ device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sqamd {

constexpr int kMaxChannels = 16;  // (l,m) channels per type for lmax <= 3
constexpr int kMaxNqp = 256;      // Lanczos degree limit of the QL kernel
constexpr int kMatsubara = 8;     // Matsubara points of the early-exit test
constexpr int kMaxTypesDev = 8;
constexpr int kRecMax = 32;       // image boxes holding one node, force pass

// Everything a Lanczos / apply kernel reads; passed by value.
struct KParams {
    // cube
    int nloc[3], nc, center, order, hw[3];
    // rank box and its extension
    int box0[3], box_n[3], ext0[3], extn[3];
    // run
    long node_begin, node_count;
    int nqp, mode, nlm_max, n_images, img_cap;
    // LDS layout (byte offsets)
    int lds_A, lds_alpha, lds_red, lds_local;
    int lds_tlo, lds_tdim, lds_toff, lds_tcen, lds_plane;  // z-march kernel tables and plane
    int kmax;                                     // images per cube column (diagnostic)
    double cw64[3][7];                            // z-march stencil weights (x,y,z) by value: scalar registers
    float cw32[3][7];
    // factored projectors of the z-march kernel: chi_lm = T_l * C_lm P_lm(x,y,z)
    const void *rad;                              // radial factors (element type = precision)
    const int64_t *img_rad;                       // first radial factor per image
    const double4 *img_cen;                       // image position (X0,Y0,Z0,-)
    double hgrid[3], range[3];                    // x = h*gx - R - X0 (lsdft_nl.cc:174-190)
    signed char ch_l[kMaxTypesDev][kMaxChannels], ch_m[kMaxTypesDev][kMaxChannels], ch_r[kMaxTypesDev][kMaxChannels];
    int layout;                                   // 1: every type has nonlocal l = {0, 2} (s+d channels)
    int zwg;                                      // z-march workgroup: 512 (1-2 columns per thread) or 1024 (1)
    // stencil weights cx[0..order] per axis (element type = precision)
    const void *cx_ptr, *cy_ptr, *cz_ptr;
    // cube diagonal fl(cdiag + V_eff) on the extended box (element type = precision)
    const void *diag_ext;
    // projector images: lo (x,y,z,-), dims (nx,ny,nz,nlm), chi offset, atom, type
    const int4 *img_lo, *img_dim;
    const int64_t *img_off;
    const int *img_atom, *img_type;
    const void *chi;             // element type = precision
    const double *gamma_ch;      // [type][kMaxChannels]
    // images covering each extended-box point, ascending
    const int *cov_off, *cov_img;
    // outputs and scratch
    void *scratch;               // 3*nc elements per workgroup
    double *out_a, *out_b;       // [node][nqp+1]
    int *status;                 // [node]
    // work queue of the z-march kernel: one counter per XCD, zeroed before a launch
    int *queue;
    // per-node early exit: stop when the lowest kMatsubara Matsubara terms
    // (2/beta) g(lambda_f + i w_m) of the node's quadrature change by < exit_tol
    // between checks at n = exit_min, exit_min + exit_every, ... (exit_tol <= 0: off)
    double exit_tol, exit_lf, exit_beta;
    int exit_min, exit_every;
    int *nsteps;                 // [node] Gauss nodes of the node's rule
    // force pass: nodes from a list, and per Lanczos step the alpha of the
    // images whose box holds the node (alpha = gamma <chi_atom, q_k>)
    const int *node_list;        // [node_count] rank-box node ids (nullptr: node_begin + r)
    double *rec;                 // [node_count][nqp][rec_slots][nlm_max]
    int *rec_img;                // [node_count][rec_slots] image id, -1 past the node's count
    int rec_slots;
    // batched kernel: node_count batches of up to B nodes (i, j, k .. k+nb-1),
    // each {id of its lowest node, nb}
    const int2 *batch;
    const void *basis;           // FP32: (t0, t2, t2 z, t2 z^2) per image box point (FP64 reads rad)
    int lds_run, lds_share;      // batched kernel: per-image atom run and type; per-atom shares of <v, V_nl v>
    int plane_px;                // z-plane row pitch in LDS: the cube's nx plus halo and bank padding
};

hipError_t launch_lanczos(const KParams &P, int precision, int grid, size_t lds, hipStream_t st);
hipError_t launch_zmarch(const KParams &P, int precision, int grid, size_t lds, hipStream_t st);
int zmarch_occupancy(const KParams &P, int precision, size_t lds);
// batched kernel (sq_zbatch.hip): B nodes stacked in z per 1024-thread workgroup
bool zbatch_supported(int precision, int B, int ug);
hipError_t launch_zbatch(const KParams &P, int precision, int B, int ug, int grid, size_t lds, hipStream_t st);
int zbatch_occupancy(const KParams &P, int precision, int B, int ug, size_t lds);
hipError_t launch_basis_tables(const double *rad, int n_images, const int4 *lo, const int4 *dim, const int64_t *roff,
                               const double4 *cen, double hz, double rz, float *b32, hipStream_t st);
hipError_t launch_apply(const KParams &P, long id, const double *x, double *y, int nonlocal_only, size_t lds,
                        hipStream_t st);
hipError_t launch_ql(const double *a, const double *b, int n, int stride, long count, double *lam, double *w,
                     double *lmin, double *lmax, int *status, double *ws, hipStream_t st, const int *nsteps = nullptr);
// QL of host-supplied device tridiagonals with its workspace in the library's
// static device array (no allocation): the array's per-device address
// (noalloc_scratch_init, with the device current), the bytes a call needs, and
// the launch, which returns the device address of the first-failure index.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch_init(int device);
void *noalloc_scratch(int device);
size_t noalloc_ql_bytes(int n, long count);
hipError_t launch_ql_noalloc(void *ws, const double *a, const double *b, int n, long count, double *lam, double *w,
                             double *lmin, double *lmax, long long **first_bad, hipStream_t st);
hipError_t launch_occupation(const double *lam, const double *w, long nq, double beta, double lf, double *partial,
                             int nblocks, double *out, hipStream_t st);
hipError_t launch_density_energy(const double *lam, const double *w, long n_nodes, int nqp, double beta, double lf,
                                 double two_over_dv, double *rho, double *part_e, double *part_s, int nblocks,
                                 double *out2, hipStream_t st);
hipError_t launch_lam_range(const double *lmin, const double *lmax, long n, double *out, hipStream_t st);
// the same with the Fermi level read from device memory (and a skip flag)
hipError_t launch_occupation_partial_dev(const double *lam, const double *w, long nq, double beta, const double *lf_dev,
                                         const int *skip, double *partial, int nblocks, hipStream_t st);
hipError_t launch_density_energy_dev(const double *lam, const double *w, long n_nodes, int nqp, double beta,
                                     const double *lf_dev, double two_over_dv, double *rho, double *part_e, double *part_s,
                                     int nblocks, double *out2, hipStream_t st);

// ---- forces (sq_forces.hip)
// coef[r][k] = (f(T) e_1)_k for node ids[r]'s tridiagonal (a, b at id * stride,
// nsteps[id] rows when given), f the Fermi function; zws holds one n x n
// eigenvector matrix per workgroup (min(count, kFermiGrid) workgroups).
constexpr int kFermiGrid = 2048;
hipError_t launch_fermi_column(const double *a, const double *b, int stride, const int *ids, const int *nsteps, long count,
                               int nmax, double beta, double lf, double *coef, double *zws, int *status, hipStream_t st);
struct TabType;
struct ForceParams {
    const int *ids;              // [count] node ids
    long count;
    int slots, nlm, nqp;         // rec layout [count][nqp][slots][nlm]
    const double *coef;          // [count][nqp]
    const double *rec;
    const int *rec_img;
    const int *nsteps;           // [node] or nullptr
    const double4 *img_cen;
    const int *img_type;
    const TabType *types;
    const double *tab;
    double h[3], range[3];
    int box0[3], box_n[3];
    double *out;                 // [count][slots][3]
    int *status;                 // [node]
};
hipError_t launch_nl_force(const ForceParams &F, hipStream_t st);

// ---- initial-density network (sq_mlp.hip): K descriptors, H hidden units
constexpr int kMlpK = 32, kMlpH = 64;
constexpr double kMlpRf = 8.0;   // descriptor cutoff, bohr
int mlp_param_count(int K, int H);
// overflow: device int counting points with more neighbours than the kernel holds
hipError_t launch_mlp_features(const double4 *img, int nimg, const int G[3], const double h[3], const double R[3], int K,
                               double rf, float *out, int *overflow, hipStream_t st);
// Z1, Z2 (may be null): pre-activations of the two hidden layers, for the deltas
hipError_t launch_mlp_forward(const float *X, long N, const float *P, const float *mean, const float *rstd, float *H1,
                              float *H2, float *Z1, float *Z2, float *Y, int bf16, hipStream_t st);
hipError_t launch_mlp_deltas(const float *Z1, const float *Z2, const float *Y, const float *T, long N, const float *P,
                             float *D1, float *D2, float *D3, hipStream_t st);
// out[i][j] = sum_n A[n][i] B[n][j] (B == nullptr: 1), in nchunk double partials summed in order
hipError_t launch_outer_sum(const float *A, int a, const float *B, int b, long N, double *partial, int nchunk, double *out,
                            hipStream_t st);
hipError_t launch_adam(float *p, const double *g, float *m, float *v, int n, float lr, int t, hipStream_t st);

// FP64 -> FP32 copy of a table (sq_setup.hip)
hipError_t launch_to_float(const double *a, float *b, long n, hipStream_t st);

// ---- V_eff pipeline of the SCF (sq_scf.hip); fields on the periodic cell G
struct StencilCoeffs {
    double w[3][11];   // -(1/2) Lap_h weights per axis, w[d][0] the centre
};
hipError_t launch_xc(const double *rho, long n, double *eps, double *vxc, hipStream_t st);
// skip (device int, may be null): a kernel whose *skip is nonzero returns at once, so a
// captured sequence can hold steps that a device-side test has made unnecessary
hipError_t launch_lap_periodic(const double *x, double *y, const int G[3], int order, const StencilCoeffs &c, hipStream_t st,
                               const int *skip = nullptr);
hipError_t launch_dot(const double *a, const double *b, long n, double scale, double *partial, int nblocks, double *out,
                      hipStream_t st, const int *skip = nullptr);
hipError_t launch_axpbypgz(double alpha, const double *x, double beta, double *y, double gamma, const double *z, long n,
                           hipStream_t st, const int *skip = nullptr);
hipError_t launch_cg_xr(double *x, double *r, const double *p, const double *Ap, const double *s, long n, hipStream_t st,
                        const int *skip = nullptr);
hipError_t launch_cg_p(double *p, const double *r, double *s, long n, hipStream_t st, const int *skip = nullptr);
// m_dev (device int, may be null) replaces m
hipError_t launch_pulay_update(const double *x, const double *f, const double *dX, const double *dF, const double *g, int m,
                               double beta, long n, double *xn, hipStream_t st, const int *m_dev = nullptr,
                               const int *skip = nullptr);
hipError_t launch_pulay_split(const double *x, const double *f, const double *dX, const double *dF, const double *g, int m,
                              long n, double *xa, double *u, hipStream_t st, const int *m_dev = nullptr,
                              const int *skip = nullptr);
// ---- device-side decisions of one SCF step (sq_scfstep.hip)
constexpr int kMaxHist = 16;   // Pulay history limit of the device solve
struct ScfCtl {
    double pstop, kstop;                 // CG thresholds: Poisson, Kerker
    double lo, hi, mid, lf, glo, ghi;    // Fermi bisection
    double gram[kMaxHist * kMaxHist], rhs[kMaxHist], g[kMaxHist];   // Pulay Gram system, coefficients
    int pdone, piters, pfail;            // Poisson CG: stopped, iterations, not converged
    int kdone, kiters, kfail;            // Kerker CG
    int fdone, fbad, fcount;             // bisection: stopped, level not bracketed, counts
    int it, head, nhist, mu;             // SCF iteration; Pulay history slot, length, coefficients used
    int conv, nopush, pulay, bad;        // SCF stopped; no history push; Pulay step; first bad node (-1)
    int gskip[kMaxHist];                 // Gram row a not needed this step
};
hipError_t launch_cg_begin(const double *s, double *stop, int *done, int *iters, double scale, int sidx, const int *skip,
                           hipStream_t st);
hipError_t launch_cg_check(const double *s, const double *stop, int *done, int *iters, int maxit, hipStream_t st);
hipError_t launch_cg_end(const double *s, const double *stop, int *fail, const int *skip, hipStream_t st);
hipError_t launch_axpb_mean(double alpha, const double *x, double *y, double mscale, const double *sum, long n,
                            const int *skip, hipStream_t st);
hipError_t launch_copy(double *dst, const double *src, long n, const int *skip, hipStream_t st);
hipError_t launch_zero(double *y, long n, const int *skip, hipStream_t st);
hipError_t launch_zero_int(int *y, int n, hipStream_t st);
hipError_t launch_bis_begin(const double *range, ScfCtl *c, hipStream_t st);
hipError_t launch_bis_store(const double *partial, int nblocks, double *out, hipStream_t st);
hipError_t launch_bis_first(ScfCtl *c, double nelec, hipStream_t st);
hipError_t launch_bis_step(const double *partial, int nblocks, ScfCtl *c, double nelec, hipStream_t st);
hipError_t launch_bis_end(ScfCtl *c, hipStream_t st);
hipError_t launch_status_any(const int *status, long n, ScfCtl *c, hipStream_t st);
hipError_t launch_conv(const double *s, ScfCtl *c, double tol, hipStream_t st);
hipError_t launch_hist_push(const double *x, const double *xp, const double *f, const double *fp, double *dX, double *dF,
                            long n, const ScfCtl *c, hipStream_t st);
hipError_t launch_hist_rotate(ScfCtl *c, int m, int pf, hipStream_t st);
hipError_t launch_pulay_solve(ScfCtl *c, int m, hipStream_t st);
hipError_t launch_it_next(ScfCtl *c, hipStream_t st);
hipError_t launch_step_begin(ScfCtl *c, hipStream_t st);

hipError_t launch_veff_to_diag(const double *veff, double cdiag, const int ext0[3], const int extn[3], const int G[3],
                               double *d64, float *d32, hipStream_t st);

}  // namespace sqamd
