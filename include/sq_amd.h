/*
 * sq_amd.h -- C ABI of the MI355X spectral-quadrature (SQ) engine, libsqamd.so.
 *
 * The engine replaces SQMiniApp's per-grid-node Lanczos hot path
 * (NodesAndWeights_LSSGQ and what it calls).  Plain pointers, sizes and error
 * codes only; no PETSc, torch or HIP types cross this boundary.  Unless a
 * function says "device", every pointer argument is host memory.
 *
 * Conventions taken from the reference:
 *  - grid node (K,J,I) is a global FD-grid index, i fastest;
 *  - a node's cube is the n_x*n_y*n_z box with origin (I-ceil(Rcut/h_x), ...),
 *    n_d = 2*ceil(Rcut/h_d)+1 (lsdft_readfiles.cc:442-444,
 *    lsdft_matvec.cc:38-45), i fastest;
 *  - per-node outputs are rows in LIpos order (k-zcor)*lx*ly + (j-ycor)*lx +
 *    (i-xcor) over the rank box (lsdft_scf.cc:67).
 * Errors: every int-returning function returns SQ_OK (0) or an sq_status code
 * and sets a thread-local message readable with sq_last_error(); nothing calls
 * exit() (the reference exits on these conditions, e.g. lsdft_tools.cc:20-26,
 * 950-953).
 */
#ifndef SQ_AMD_H
#define SQ_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sq_ctx sq_ctx;

enum sq_status {
    SQ_OK = 0,
    SQ_ERR_IO = 1,        /* a deck file could not be opened */
    SQ_ERR_FORMAT = 2,    /* a deck file is malformed */
    SQ_ERR_RANGE = 3,     /* spline abscissa outside the radial table, or a value out of range */
    SQ_ERR_ARG = 4,       /* bad argument */
    SQ_ERR_QL = 5,        /* more than 50 QL sweeps for one eigenvalue (lsdft_tools.cc:950-953) */
    SQ_ERR_ALLOC = 6,     /* host or device allocation failed */
    SQ_ERR_BREAKDOWN = 7, /* Lanczos b_k == 0: the reference divides by it (lsdft_scf.cc:131,147) */
    SQ_ERR_DEVICE = 8,    /* a HIP call failed */
    SQ_ERR_STATE = 9      /* call out of order (e.g. results requested before a run) */
};

/* Nonlocal operator applied inside H. */
enum sq_nonlocal_mode {
    SQ_NL_REF_COLLAPSED = 0, /* Mult_NonlocalVector (lsdft_matvec.cc:25-149): one alpha per atom */
    SQ_NL_KB = 1             /* Mult_HamiltonianVector (lsdft_matvec.cc:155-293): alpha per (atom,l,m) */
};

/* Arithmetic of the Lanczos recurrence. */
enum sq_precision {
    SQ_FP64 = 0, /* plain form: FP64 vectors, FP64 dot products */
    SQ_FP32 = 1  /* FP32 vectors and stencil, compensated dot products; FP64 QL and quadrature */
};

typedef struct {
    int phantom_atom;        /* 1: emulate the reference's extra atom read past the position
                                array (endPosSC = 3*nAtoms, lsdft_initObjs.cc:3038) at phantom_pos */
    double phantom_pos[3];
    int rank_box[6];         /* owned node box (xcor,ycor,zcor,lx,ly,lz); lx<=0 means the whole grid */
    int nqp_override;        /* >0 replaces LANCZOS_DEGREE */
    int device;              /* HIP device ordinal; < 0 builds the host tables only (no device
                                calls; the run entry points then return SQ_ERR_STATE) */
} sq_options;

typedef struct {
    int npts[3], nloc[3], halfwidth[3], order, nqp, ntype, nAtoms, noetot;
    int n_images, n_sc_atoms, ip_length, n_nodes, nc;
    int rank_box[6], ext_origin[3], ext_dims[3];
    double range[3], delta[3], dV, beta, rcut;
    double coeffs[3][11];    /* -(1/2) Laplacian FD weights, lsdft_readfiles.cc:453-476 */
    double gamma[8][4];      /* Gamma_l per type, lsdft_initObjs.cc:822-869 */
    double cutoff[8][3];     /* pseudocharge cutoff, lsdft_initObjs.cc:628-817 */
    long long chi_entries;   /* total projector entries over all images */
    int kernel;              /* Lanczos kernel in use: 1 column-marching, 0 generic */
    int kmax;                /* most images meeting one (x,y) column of a cube */
    int workgroup[2];        /* Lanczos workgroup size per precision (SQ_FP64, SQ_FP32), chosen by
                                the autotuner at creation */
    float tune_ms[2][2];     /* autotuner probe times per precision for 512 and 1024 threads
                                (0 where not timed) */
    int batch[2];            /* nodes per Lanczos workgroup per precision: 1 (one-node kernel) or
                                2, 4 (batched kernel, nodes stacked in z) */
    int gather_trip[2];      /* batched kernel's gather trip per precision (0: one-node kernel) */
    float tune_candidates_ms[2][6]; /* autotuner probe times per precision: one-node 512 / 1024
                                threads, B=2 trips a / b, B=4 trips a / b (0: not timed) */
    int union_images[2];     /* most images meeting one union cube, B = 2 and B = 4 */
} sq_info;

/* Message of the last failing call on this thread ("" if none). */
const char *sq_last_error(void);

/* Default options: no phantom atom, whole grid, deck Lanczos degree, device 0. */
void sq_default_options(sq_options *opt);

/* Host setup + device upload.  Replaces Initialize (lsdft_initObjs.cc:36-93),
 * GetInfluencingAtoms_Nonlocal (lsdft_initObjs.cc:1671-1946),
 * CalculateNonlocalProjectors (lsdft_nl.cc:141-208) and NonlocalProjectors
 * (lsdft_nl.cc:6-130).  Reads <dir>/<name>.inpt, <dir>/<name>.ion and
 * <dir>/pseudopotential/<PSEUDOPOTENTIAL_FILE>.  V_eff starts as the
 * reference's constant 0.333 (lsdft_scf.cc:71). */
int sq_create_from_deck(const char *dir, const char *name, const sq_options *opt, sq_ctx **out);

int sq_destroy(sq_ctx *ctx);

int sq_get_info(sq_ctx *ctx, sq_info *info);

/* V_eff on the global grid (npts[0]*npts[1]*npts[2], i fastest), wrapped
 * periodically in grid index on each cube; field == NULL sets the constant
 * veff_const everywhere.  Replaces VecSet(VeffNodal) + MatDiagonalSet
 * (lsdft_scf.cc:71-75). */
int sq_set_veff(sq_ctx *ctx, const double *field, double veff_const);

/* The batched hot path: NodesAndWeights_LSSGQ (lsdft_scf.cc:32-82) for the
 * rank-box nodes [node_begin, node_begin+node_count): N_qp+1 H applications
 * per node (CalculateSpectralNodesAndWeights, lsdft_scf.cc:87-164) and the QL
 * nodes/weights (TridiagEigenVecSolve_NodesAndWeights, lsdft_tools.cc:912-1026).
 * Results stay on the device (see sq_get_nodes_and_weights). */
int sq_run(sq_ctx *ctx, int nonlocal_mode, int precision, long node_begin, long node_count);

/* Copy the last sq_run results for its nodes: lambda/w [node_count][nqp],
 * lmin/lmax [node_count] (any pointer may be NULL). */
int sq_get_nodes_and_weights(sq_ctx *ctx, double *lambda, double *w, double *lmin, double *lmax);

/* Lanczos coefficients a[node][0..nqp], b[node][0..nqp] of the last sq_run. */
int sq_get_lanczos(sq_ctx *ctx, double *a, double *b);

/* sq_run + sq_get_nodes_and_weights. */
int sq_nodes_and_weights(sq_ctx *ctx, int nonlocal_mode, int precision, long node_begin, long node_count,
                         double *lambda, double *w, double *lmin, double *lmax);

/* Device time of the last sq_run's Lanczos kernel and QL kernel, in ms,
 * measured with HIP events on the context's stream. */
int sq_last_timing(sq_ctx *ctx, double *lanczos_ms, double *ql_ms);

/* One H application on node (Kp,Jp,Ip), on the device:
 * y = (L + diag V_eff) x + V_nl x.  Replaces MatMult(LapPlusVeffOprloc) +
 * Mult_NonlocalVector / Mult_HamiltonianVector (lsdft_scf.cc:121-123,
 * lsdft_matvec.cc:25-293).  x, y: host vectors of nc entries. */
int sq_apply_h(sq_ctx *ctx, int nonlocal_mode, int Kp, int Jp, int Ip, const double *x, double *y);

/* The nonlocal part alone, accumulated: y += V_nl x on node (Kp,Jp,Ip), each
 * point's projector terms added onto its incoming y in ascending image and
 * channel order.  Replaces Mult_NonlocalVector (nonlocal_mode
 * SQ_NL_REF_COLLAPSED, lsdft_matvec.cc:25-149) and Mult_HamiltonianVector
 * (SQ_NL_KB, lsdft_matvec.cc:155-293), which add into V2.  x, y: host
 * vectors of nc entries. */
int sq_apply_nonlocal(sq_ctx *ctx, int nonlocal_mode, int Kp, int Jp, int Ip, const double *x, double *y);

/* Batched QL nodes and weights on the device for host tridiagonals:
 * diag[count][n], subdiag[count][n] (subdiag[.][n-1] ignored), outputs
 * lambda/w [count][n], lmin/lmax [count].  TridiagEigenVecSolve_NodesAndWeights,
 * lsdft_tools.cc:912-1026. */
int sq_tridiag_nodes_weights(sq_ctx *ctx, const double *diag, const double *subdiag, int n, long count,
                             double *lambda, double *w, double *lmin, double *lmax);

/* sq_tridiag_nodes_weights for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  diag, subdiag, lambda,
 * w [count][n] and lmin, lmax [count] are DEVICE buffers of the context's
 * device, all required, the inputs complete before the call (it orders
 * nothing against other streams).  The QL workspace is the library's static
 * device array of 32 MiB per device, 3*8*n*count + 4*count bytes plus
 * alignment (n = 80: count <= 17439); larger batches return SQ_ERR_ARG.  The
 * call allocates nothing, runs on the context's stream and waits for it.
 * Calls that overlap on one device race on that array. */
int sq_tridiag_nodes_weights_noalloc(sq_ctx *ctx, const double *diag, const double *subdiag, int n, long count,
                                     double *lambda, double *w, double *lmin, double *lmax);

/* Density, Fermi level and band/entropy energy from the last sq_run, which must
 * have covered every node of the rank box.  The reference declares these only
 * (CalculateDensity_lssgq, constraint_lssgq, FermiDirac, findRootBrent,
 * SystemEnergy_Calc_lssgq: lsdft.h:98-102,209).  rho[n_nodes] may be NULL.
 * nelec <= 0 uses the deck's electron count; beta <= 0 uses the deck's BETA. */
int sq_density_fermi_energy(sq_ctx *ctx, double nelec, double beta, double *lambda_f, double *rho,
                            double *e_band, double *minus_ts);

/* Per-node early exit of the batched Lanczos (the column-marching kernel): a
 * node stops at n Lanczos steps once the lowest 8 Matsubara terms
 * (2/beta) g(lambda_f + i (2m+1) pi / beta), g(z) = e1^T (z - T_n)^{-1} e1,
 * of its Gauss rule change by less than tol between checks at n = min_steps,
 * min_steps + every, ... (the Fermi quadrature is 1/2 + the sum of their real
 * parts over all m), up to the deck's LANCZOS_DEGREE.  Its rule then has n
 * nodes; the rest of its lambda/w row is padded with weight zero.  tol <= 0
 * turns it off; beta <= 0 uses the deck's BETA.  The reference has no
 * counterpart (it runs N_qp steps for every node, lsdft_scf.cc:137). */
int sq_set_early_exit(sq_ctx *ctx, double tol, double lambda_f, double beta, int min_steps, int every);

/* Gauss nodes per node of the last sq_run (N_qp unless it stopped early). */
int sq_get_rule_sizes(sq_ctx *ctx, int *n);

/* The autotuner's rule (no counterpart in the reference), as sq_create_from_deck
 * applies it per precision to its probe times: candidates 0/1 = the one-node
 * kernel with 512/1024 threads, 2/3 = 2 nodes a workgroup with gather trip a/b,
 * 4/5 = 4 nodes with trip a/b (planes a trip: 1/2 in FP64 on the vector ALU;
 * 4 on the matrix cores / 2 on the vector ALU in FP32).  available[6]:
 * which can launch; probe_ms[6]: their probe times; fixed_* the values of
 * SQ_ZMARCH_WG, SQ_BATCH and SQ_BATCH_TRIP (NULL when unset), which narrow the
 * set.  *candidate = the fastest remaining (lower index on a tie); candidate 0
 * when the narrowing leaves nothing and it can launch; -1 when nothing can. */
int sq_tune_select(int precision, const int *available, const float *probe_ms, const char *fixed_workgroup,
                   const char *fixed_batch, const char *fixed_trip, int *candidate);

/* Wait for all device work of the context's device (hipDeviceSynchronize). */
int sq_synchronize(sq_ctx *ctx);

/* Projector image c of the setup: center (3), box (xs,ys,zs,xe,ye,ze),
 * supercell atom index, type, entries (Nlm*box points), and the flattened
 * chi (k,j,i outer, (l,m) inner, as lsdft_nl.cc:81-123) if chi != NULL. */
int sq_get_image(sq_ctx *ctx, int c, double *center, int *box, int *atom, int *type, int *entries, double *chi);

/* ------------------------------------------------------------------ SCF
 * The reference declares its SCF but does not build it (SCF_LSSGQ,
 * SolvePoisson, Vxc_Calc/Exc_Calc, mix, FormVeffNodal: lsdft.h:93,124,138,
 * 183,197-204); it builds the pieces: the periodic Poisson operator and KSP
 * (lsdft_initObjs.cc:249-276,309-388), the pseudocharge (628-817), the
 * Periodic-Pulay vectors (233-245) and the SCF keywords (lsdft_readfiles.cc:
 * 219-240).  These entry points assemble them on the periodic cell: the
 * (N0-1) x (N1-1) x (N2-1) distinct grid points (the atom images repeat every
 * 2R = (N-1) h), i fastest.  A context used for the SCF must be created with
 * rank_box = (0, 0, 0, N0-1, N1-1, N2-1), so its nodes are the cell's points. */

typedef struct {
    int nonlocal_mode;       /* SQ_NL_KB (default) or SQ_NL_REF_COLLAPSED */
    int precision;           /* SQ_FP64 (default) or SQ_FP32 Lanczos */
    int max_iter;            /* <= 0: MAXIT_SCF of the deck */
    double tol;              /* <= 0: TOL_SCF (relative density residual) */
    double mixing;           /* <= 0: MIXING_PARAMETER */
    int history;             /* <= 0: MIXING_HISTORY */
    int pulay_frequency;     /* <= 0: PULAY_FREQUENCY */
    double tol_poisson;      /* <= 0: TOL_POISSON */
    const double *rho0;      /* initial density on the cell; NULL: superposed atomic densities */
    int verbose;             /* 1: one line per iteration on stderr */
    double exit_tol;         /* > 0: per-node early exit (sq_set_early_exit) from the second
                                iteration on, at the previous iteration's Fermi level */
    int exit_min, exit_every;   /* its first check and spacing (<= 0: 30 and 5) */
    double kerker;           /* > 0: Kerker-preconditioned mixing with this k0 (1/bohr), the new
                                density (x - dX g) + beta P (f - dF g), P = q^2/(q^2 + k0^2) */
    double tol_report;       /* > 0: also record the wall time, iteration and energy at which the
                                residual first falls below this value (a looser stop of the same run) */
} sq_scf_options;

typedef struct {
    int iterations, converged, poisson_iterations;
    double residual;         /* last ||rho_out - rho_in|| / ||rho_out|| */
    double lambda_f, e_band, minus_ts, e_xc, e_vxc, e_el, e_self, e_total;
    /* e_total = e_band + e_xc - e_vxc + e_el - e_self + minus_ts (Ha), with
       e_el = 1/2 int phi (b - rho); energies of the last input density */
    double wall_s, sq_s, poisson_s;   /* SCF wall time; SQ passes + Fermi level; Poisson solves */
    double mean_rule_size;   /* mean Gauss nodes per node over the SQ passes */
    int graph_steps;         /* steps replayed from the captured graph of one SCF step (the first
                                step runs launch by launch; SQ_NO_GRAPHS or early exit: none) */
    int report_iterations;   /* tol_report: iteration of the first residual below it (0: not reached) */
    double report_wall_s, report_e_total;   /* wall time and total energy at that iteration */
} sq_scf_result;

enum sq_field {
    SQ_FIELD_RHO = 0,         /* input density of the last SCF iteration (the converged one on success) */
    SQ_FIELD_RHO_OUT = 1,     /* mixed density after the last iteration (or rho_out when converged) */
    SQ_FIELD_PHI = 2,         /* electrostatic potential */
    SQ_FIELD_VXC = 3,         /* exchange-correlation potential */
    SQ_FIELD_VEFF = 4,        /* phi + V_xc */
    SQ_FIELD_B = 5,           /* pseudocharge density (setup) */
    SQ_FIELD_RHO_ATOMIC = 6   /* superposed isolated-atom density, the default initial guess (setup) */
};

void sq_default_scf_options(sq_scf_options *opt);

/* SCF_LSSGQ: the SCF loop on the device (V_eff pipeline, SQ pass, density,
 * Periodic Pulay mixing) until the density residual is below tol. */
int sq_scf(sq_ctx *ctx, const sq_scf_options *opt, sq_scf_result *result);

/* A field of the cell (sq_field), out[(N0-1)*(N1-1)*(N2-1)]. */
int sq_get_cell_field(sq_ctx *ctx, int which, double *out);

/* V_eff on the cell (period N-1 in every direction) for the following sq_run:
 * FormVeffNodal / VeffGlobalToLocal (declared, lsdft.h:197-204). */
int sq_set_veff_cell(sq_ctx *ctx, const double *veff);

/* Vxc_Calc / Exc_Calc (declared, lsdft.h:138-150) on the device: LDA
 * (Slater exchange + Perdew-Wang 1992 correlation), eps_xc and v_xc per point. */
int sq_xc(sq_ctx *ctx, const double *rho, long n, double *eps, double *vxc);

/* SolvePoisson (declared, lsdft.h:93): -(1/4pi) Lap_h phi = f on the cell by
 * conjugate gradients on the periodic order-2*order stencil to tol (<= 0:
 * TOL_POISSON), mean(phi) = 0; f must be neutral up to rounding. */
int sq_poisson(sq_ctx *ctx, const double *f, double tol, double *phi, int *iterations);

/* ------------------------------------------------------------ forces
 * Calculate_localforce and Calculate_nonlocalforce (declared, lsdft.h:188-189)
 * with CalculateGradientDensityMatrixColumn / CalculateNonlocalForceContribution
 * (declared, lsdft.h:224-226): the forces on the atoms (Ha/bohr) at the last SQ
 * pass of sq_scf on this context, as the derivative of the SCF energy at that
 * pass's input density.
 *   local    : the pseudocharge term, from phi (host);
 *   nonlocal : F_J = 4 sum_c sum_lm grad chi_J,lm(c) gamma <chi_J,lm, D_c>, D_c the
 *              node's Gauss-quadrature density-matrix column Q f(T) e_1, from a
 *              repeat of the pass over the nodes inside projector boxes.
 * forces, f_local, f_nonlocal: 3 * nAtoms each (x, y, z per atom, .ion order);
 * any may be NULL, and a part is computed only when forces or its own array
 * is given (f_local alone skips the nonlocal pass).  Needs the
 * Kleinman-Bylander operator (SQ_NL_KB). */
typedef struct {
    double lambda_f;         /* Fermi level of the pass */
    double wall_s;           /* whole call */
    double nonlocal_s;       /* the repeated pass, Fermi columns and contraction */
    long nodes;              /* nodes inside a projector box */
    int slots;               /* most projector boxes holding one node */
    int chunks;              /* node chunks the record was split into */
} sq_force_info;

int sq_forces(sq_ctx *ctx, double *forces, double *f_local, double *f_nonlocal, sq_force_info *info);

/* ------------------------------------------------------------ initial-density network
 * No counterpart in the reference, whose SCF (declared only) would start from
 * the superposed atomic densities (lsdft_readfiles.cc:1026-1040).  A small
 * network predicts the SCF's correction to that superposition at every cell
 * point from 32 Behler-type symmetry functions of the Mg atoms within 8 bohr
 * of the point: 16 radial Gaussians (centres 8k/15, width 8/15) and 16
 * angular functions 2^(1-zeta) sum_{J<L} (1 + lam cos t_JL)^zeta
 * exp(-eta (r_J^2 + r_L^2)) over eta in {0.005, 0.02}, zeta in {1, 2, 4, 16},
 * lam = +-1, all with the cutoff (cos(pi r/8) + 1)/2; 32 -> 64 -> 64 -> 1,
 * SiLU.  It is trained on densities the SQ SCF itself converged
 * (sq_mlp_add_sample after sq_scf) by the library's own forward and backward
 * kernels (Adam, full batch, FP32), and its guess (BF16 weights and
 * activations, FP32 accumulation, or FP32) goes to sq_scf_options.rho0.  The
 * superposition stays the fallback (rho0 = NULL). */
typedef struct sq_mlp sq_mlp;

/* parameter count (6337) and the normalization block: mean[32], 1/std[32], target scale */
#define SQ_MLP_PARAMS 6337
#define SQ_MLP_NORM 65

int sq_mlp_create(int device, unsigned seed, sq_mlp **model);
int sq_mlp_destroy(sq_mlp *model);
/* params: SQ_MLP_PARAMS floats (W1[64][32], b1, W2[64][64], b2, w3, b3); norm:
 * SQ_MLP_NORM floats; either may be NULL.  set_params resets the optimizer. */
int sq_mlp_get_params(sq_mlp *model, float *params, float *norm);
int sq_mlp_set_params(sq_mlp *model, const float *params, const float *norm);
/* descriptors of the context's periodic cell, out[(N0-1)(N1-1)(N2-1)][32] */
int sq_mlp_descriptors(sq_ctx *ctx, float *out);
/* appends the context's cell (descriptors, target rho - rho_atomic) to the
 * training set; the context holds a converged sq_scf */
int sq_mlp_add_sample(sq_mlp *model, sq_ctx *ctx);
/* the same with every stride-th cell point only (smaller training sets) */
int sq_mlp_add_sample_every(sq_mlp *model, sq_ctx *ctx, int stride);
/* epochs of Adam (learning rate lr) over the whole training set; loss = the
 * final (1/2N) sum (y - t)^2 in scaled units.  The first call fixes the
 * descriptor standardization and target scale from the set. */
int sq_mlp_train(sq_mlp *model, int epochs, double lr, double *loss);
/* the network on N descriptor rows (normalize: apply the stored
 * standardization first); bf16: weights and activations rounded to bfloat16,
 * FP32 accumulation */
int sq_mlp_forward(sq_mlp *model, const float *x, long n, int bf16, int normalize, float *y);
/* loss and gradient (SQ_MLP_PARAMS doubles) on standardized rows x, scaled targets t */
int sq_mlp_loss_grad(sq_mlp *model, const float *x, const float *t, long n, double *loss, double *grad);
/* initial density for the context's cell: max(0, rho_atomic + network),
 * scaled to N_e; rho0[(N0-1)(N1-1)(N2-1)] */
int sq_mlp_guess(sq_mlp *model, sq_ctx *ctx, int bf16, double *rho0);

/* The Kerker preconditioner of the mixing on a cell field: y = A (A + k0^2/2)^-1 u,
 * A = -(1/2) Lap_h, by conjugate gradients (no reference counterpart: its mix
 * is declared only, lsdft.h:183). */
int sq_kerker(sq_ctx *ctx, const double *u, double k0, double *y, int *iterations);

#ifdef __cplusplus
}
#endif
#endif
