/*
 * sq_lsdft.h -- SQMiniApp's host entry points for the spectral-quadrature hot
 * path, with the reference's names and argument meaning, implemented over the
 * C ABI of sq_amd.h (libsqamd.so exports both).
 *
 * C++ linkage, as the reference declares them (src/lsdft.h).  The reference's
 * god-struct LSDFT_OBJ (src/ilsdft.h) becomes SQ_LSDFT: the fields the path
 * reads or writes, under the reference's names, plus the sq_ctx holding the
 * device state.
 *
 * Error behaviour.  The reference returns 0 (or nothing) and calls exit(1) on
 * fatal conditions (e.g. src/lsdft_tools.cc:950-953, lsdft_readfiles.cc:1121-
 * 1125).  Here every function returns 0 on success or the sq_status code, with
 * the message in sq_last_error(); sq_main, the driver, prints it and exits(1)
 * where the reference would.  Functions the reference declares void return int
 * for that reason.
 */
#ifndef SQ_LSDFT_H
#define SQ_LSDFT_H

#include <string>
#include <vector>

#include "sq_amd.h"

struct SQ_LSDFT {
    sq_ctx *ctx = nullptr;
    sq_info info{};
    /* Nonlocal operator inside H.  The reference's hot loop calls
     * Mult_NonlocalVector (lsdft_scf.cc:123,140): SQ_NL_REF_COLLAPSED. */
    int nonlocal_mode = SQ_NL_REF_COLLAPSED;
    int precision = SQ_FP64;

    /* src/ilsdft.h fields */
    int numPoints_x = 0, numPoints_y = 0, numPoints_z = 0;  /* :86-88 */
    int nx_loc = 0, ny_loc = 0, nz_loc = 0;                 /* :94-96 */
    int order = 0;                                          /* :100, half the FD order */
    int nAtoms = 0, noetot = 0, N_qp = 0, IP_length = 0;    /* :101,106,109,116 */
    double range_x = 0, range_y = 0, range_z = 0;           /* :145-147 */
    double delta_x = 0, delta_y = 0, delta_z = 0;           /* :148-150 */
    double delVol = 0, Rcut = 0, Beta = 0;                  /* :171,173,191 */
    double coeffs_x[11] = {}, coeffs_y[11] = {}, coeffs_z[11] = {};  /* :182-184 */
    double Eband = 0, Entropy = 0;                          /* :158-159; Entropy holds -TS */
    double Exc = 0, Eself = 0, Etot = 0;                    /* :152-154 and the SCF total energy */
    int MAXITSCF = 0, SCFIterations = 0;                    /* :111; iterations of the last SCF */
    double TOLSCF = 0, MixingParameter = 0;                 /* :156,200 (0: the deck's values) */

    /* DMDAGetCorners(da) of this rank (lsdft_scf.cc:49) */
    int xcor = 0, ycor = 0, zcor = 0, lxdim = 0, lydim = 0, lzdim = 0;

    /* lssgq_lambda / lssgq_weights [LIpos][N_qp] and lambda_min / lambda_max
     * [LIpos] (ilsdft.h:176-179, allocated at lsdft_initObjs.cc:284-293),
     * stored row-major in one block each. */
    std::vector<double> lssgq_lambda, lssgq_weights, lambda_min, lambda_max;
    /* electron density on the owned nodes, LIpos order (elecDensRho) */
    std::vector<double> elecDensRho;
    double lambda_f = 0;

    /* forces (ilsdft.h:163-164,299-300): x, y, z per atom in .ion order, and
     * their local / nonlocal parts */
    std::vector<double> forces, localforces, nonlocalforces;
    double MaxForce = 0, NormForce = 0;   /* largest component; mean |F_J| */

    int rank = 0;
    std::string DenFilename;     /* "./<name>" (lsdft_readfiles.cc:118,136) */
    std::string ReadDenFilename;
    std::string ForceFilename;   /* "./<name>.force" (lsdft_readfiles.cc:115,124-125) */
};

/* Get_Input + Initialize + GetInfluencingAtoms_Local/_Nonlocal (main.cc:31-38;
 * lsdft_readfiles.cc:148-331, lsdft_initObjs.cc:36-93,1671-1946): reads
 * <dir>/<name>.inpt/.ion/pseudopotential, builds the node box of this rank,
 * the projector images and tables, and uploads them to opt->device. */
int Initialize(SQ_LSDFT *pLsdft, const char *dir, const char *name, const sq_options *opt);

/* Releases the context (Objects_Destroy, which the reference never calls). */
int Objects_Destroy(SQ_LSDFT *pLsdft);

/* CalculateNonlocalProjectors (lsdft_nl.cc:141-208) and NonlocalProjectors
 * (lsdft_nl.cc:6-130).  Initialize has already built chi and its radial
 * factors -- on the device, one workgroup per image, when opt->device >= 0;
 * on the host for a device-less context -- so these check that the tables
 * exist and return 0. */
double CalculateNonlocalProjectors(SQ_LSDFT *pLsdft);
int NonlocalProjectors(SQ_LSDFT *pLsdft);

/* NodesAndWeights_LSSGQ (lsdft_scf.cc:32-82): every owned node, nodes and
 * weights into lssgq_lambda/lssgq_weights, extremes into lambda_min/max. */
double NodesAndWeights_LSSGQ(SQ_LSDFT *pLsdft);

/* CalculateSpectralNodesAndWeights (lsdft_scf.cc:87-164): one node.  p must
 * be the cube centre Nmid (lsdft_scf.cc:53-57) and LIp the LIpos of
 * (Kp,Jp,Ip) (lsdft_scf.cc:67), otherwise SQ_ERR_ARG. */
double CalculateSpectralNodesAndWeights(SQ_LSDFT *pLsdft, int p, int LIp, int Kp, int Jp, int Ip);

/* Mult_NonlocalVector (lsdft_matvec.cc:25-149) and Mult_HamiltonianVector
 * (lsdft_matvec.cc:155-293): V2 += V_nl V1 on the cube of node (Kp,Jp,Ip).
 * V1, V2: nc host values, cube-natural order (i fastest). */
int Mult_NonlocalVector(SQ_LSDFT *pLsdft, const double *V1, double *V2, int Kp, int Jp, int Ip);
int Mult_HamiltonianVector(SQ_LSDFT *pLsdft, const double *V1, double *V2, int Kp, int Jp, int Ip);

/* TridiagEigenVecSolve_NodesAndWeights (lsdft_tools.cc:912-1026): QL of
 * diag[0..n-1], subdiag[0..n-2] into row LIp of the lssgq arrays. */
int TridiagEigenVecSolve_NodesAndWeights(SQ_LSDFT *pLsdft, double diag[], double subdiag[], int n, int LIp);

/* FermiDirac (declared lsdft.h:99): 1/(1+exp(bet*(lambda-lambdaf))). */
double FermiDirac(double bet, double lambda, double lambdaf);

/* CalculateDensity_lssgq (declared lsdft.h:100): Fermi level from the
 * electron-count constraint (constraint_lssgq / findRootBrent, declared
 * lsdft.h:98,102), elecDensRho, Eband and Entropy (-TS), on the device. */
int CalculateDensity_lssgq(SQ_LSDFT *pLsdft);

/* SystemEnergy_Calc_lssgq (declared lsdft.h:209): prints the band and
 * entropy energies that CalculateDensity_lssgq computed. */
int SystemEnergy_Calc_lssgq(SQ_LSDFT *pLsdft);

/* SCF_LSSGQ (declared lsdft.h:124) on the periodic cell: the context must
 * have been initialized with rank_box = (0,0,0,N0-1,N1-1,N2-1).  Runs
 * sq_scf with the deck's SCF keywords (or MAXITSCF / TOLSCF / MixingParameter
 * when set), fills elecDensRho with the converged density and lambda_f,
 * Eband, Entropy, Exc, Eself, Etot; returns 0 or the sq_status code
 * (SQ_ERR_RANGE when the loop ends unconverged). */
double SCF_LSSGQ(SQ_LSDFT *pLsdft);

/* Calculate_localforce / Calculate_nonlocalforce (declared, lsdft.h:188-189):
 * the pseudocharge and Kleinman-Bylander parts of the forces at the last pass
 * of SCF_LSSGQ (sq_forces) into localforces / nonlocalforces; forces is their
 * sum once both have run, with MaxForce and NormForce. */
int Calculate_localforce(SQ_LSDFT *pLsdft);
int Calculate_nonlocalforce(SQ_LSDFT *pLsdft);

/* Display_force (declared, lsdft.h:191): prints the forces and writes them to
 * ForceFilename, one line "Fx Fy Fz" (Ha/bohr) per atom. */
int Display_force(SQ_LSDFT *pLsdft);

/* The reference's per-rank text files (lsdft_readfiles.cc:1106-1380):
 * <DenFilename>_<rank>.den holds "xcor ycor zcor lx ly lz" then "k j i rho"
 * per owned node; .lambdas holds the same header then "LIpos lmin lmax".
 * The readers reject a file written for another node box (the reference
 * exits there). */
int Print_ElecDens(SQ_LSDFT *pLsdft);
int Read_ElecDens(SQ_LSDFT *pLsdft);
int Print_MinMaxLambda(SQ_LSDFT *pLsdft);
int Read_MinMaxLambda(SQ_LSDFT *pLsdft);

#endif
