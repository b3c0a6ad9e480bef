// sq_lsdft.cpp -- the reference's host entry points (include/sq_lsdft.h) over
// the C ABI.  Each forwards to sq_* and keeps the reference's outputs in the
// SQ_LSDFT arrays; no arithmetic of the path runs here.
#include "../../include/sq_lsdft.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace sqamd {
int record_error(int code, const std::string &msg);  // sq_api.cpp: the sq_last_error() slot
}

namespace {

int fail(int code, const std::string &msg) { return sqamd::record_error(code, msg); }

bool ready(const SQ_LSDFT *p) { return p && p->ctx; }

long n_owned(const SQ_LSDFT *p) { return (long)p->lxdim * p->lydim * p->lzdim; }

std::string rank_file(const std::string &base, int rank, const char *ext)
{
    // DenFilename + "_" + rank + ext (lsdft_readfiles.cc:1113-1117, 1301-1305;
    // the reference prefixes "./" to a DenFilename that already starts with it)
    return base + "_" + std::to_string(rank) + ext;
}

}  // namespace

int Initialize(SQ_LSDFT *p, const char *dir, const char *name, const sq_options *opt)
{
    if (!p || !dir || !name) return fail(SQ_ERR_ARG, "Initialize: null argument");
    sq_options o;
    sq_default_options(&o);
    if (opt) o = *opt;
    int rc = sq_create_from_deck(dir, name, &o, &p->ctx);
    if (rc) return rc;
    rc = sq_get_info(p->ctx, &p->info);
    if (rc) return rc;
    const sq_info &I = p->info;
    p->numPoints_x = I.npts[0];
    p->numPoints_y = I.npts[1];
    p->numPoints_z = I.npts[2];
    p->nx_loc = I.nloc[0];
    p->ny_loc = I.nloc[1];
    p->nz_loc = I.nloc[2];
    p->order = I.order;
    p->nAtoms = I.nAtoms;
    p->noetot = I.noetot;
    p->N_qp = I.nqp;
    p->IP_length = I.ip_length;
    p->range_x = I.range[0];
    p->range_y = I.range[1];
    p->range_z = I.range[2];
    p->delta_x = I.delta[0];
    p->delta_y = I.delta[1];
    p->delta_z = I.delta[2];
    p->delVol = I.dV;
    p->Rcut = I.rcut;
    p->Beta = I.beta;
    for (int q = 0; q < 11; q++) {
        p->coeffs_x[q] = I.coeffs[0][q];
        p->coeffs_y[q] = I.coeffs[1][q];
        p->coeffs_z[q] = I.coeffs[2][q];
    }
    p->xcor = I.rank_box[0];
    p->ycor = I.rank_box[1];
    p->zcor = I.rank_box[2];
    p->lxdim = I.rank_box[3];
    p->lydim = I.rank_box[4];
    p->lzdim = I.rank_box[5];
    const size_t n = (size_t)n_owned(p);
    p->lssgq_lambda.assign(n * p->N_qp, 0.0);
    p->lssgq_weights.assign(n * p->N_qp, 0.0);
    p->lambda_min.assign(n, 0.0);
    p->lambda_max.assign(n, 0.0);
    p->elecDensRho.assign(n, 0.0);
    if (p->DenFilename.empty()) p->DenFilename = std::string("./") + name;
    if (p->ForceFilename.empty()) p->ForceFilename = std::string("./") + name + ".force";
    p->forces.assign(3 * (size_t)p->nAtoms, 0.0);
    p->localforces.assign(3 * (size_t)p->nAtoms, 0.0);
    p->nonlocalforces.assign(3 * (size_t)p->nAtoms, 0.0);
    return SQ_OK;
}

int Objects_Destroy(SQ_LSDFT *p)
{
    if (!p) return fail(SQ_ERR_ARG, "Objects_Destroy: null argument");
    int rc = p->ctx ? sq_destroy(p->ctx) : SQ_OK;
    p->ctx = nullptr;
    return rc;
}

double CalculateNonlocalProjectors(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "CalculateNonlocalProjectors: Initialize first");
    return p->info.chi_entries > 0 || p->info.n_images == 0 ? 0 : fail(SQ_ERR_STATE, "projector tables missing");
}

int NonlocalProjectors(SQ_LSDFT *p)
{
    return (int)CalculateNonlocalProjectors(p);
}

double NodesAndWeights_LSSGQ(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "NodesAndWeights_LSSGQ: Initialize first");
    const long n = n_owned(p);
    return sq_nodes_and_weights(p->ctx, p->nonlocal_mode, p->precision, 0, n, p->lssgq_lambda.data(),
                                p->lssgq_weights.data(), p->lambda_min.data(), p->lambda_max.data());
}

double CalculateSpectralNodesAndWeights(SQ_LSDFT *p, int pmid, int LIp, int Kp, int Jp, int Ip)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "CalculateSpectralNodesAndWeights: Initialize first");
    // the Lanczos start vector is the unit vector at the cube centre (lsdft_scf.cc:53-57)
    const int Nmid = ((p->nz_loc - 1) / 2 * p->ny_loc + (p->ny_loc - 1) / 2) * p->nx_loc + (p->nx_loc - 1) / 2;
    if (pmid != Nmid) return fail(SQ_ERR_ARG, "CalculateSpectralNodesAndWeights: p is not the cube centre");
    if (Ip < p->xcor || Ip >= p->xcor + p->lxdim || Jp < p->ycor || Jp >= p->ycor + p->lydim || Kp < p->zcor ||
        Kp >= p->zcor + p->lzdim)
        return fail(SQ_ERR_ARG, "CalculateSpectralNodesAndWeights: node outside the rank box");
    const long LIpos = ((long)(Kp - p->zcor) * p->lydim + (Jp - p->ycor)) * p->lxdim + (Ip - p->xcor);
    if (LIp != LIpos) return fail(SQ_ERR_ARG, "CalculateSpectralNodesAndWeights: LIp differs from the LIpos of (Kp,Jp,Ip)");
    const size_t q = (size_t)LIp * p->N_qp;
    return sq_nodes_and_weights(p->ctx, p->nonlocal_mode, p->precision, LIp, 1, &p->lssgq_lambda[q],
                                &p->lssgq_weights[q], &p->lambda_min[LIp], &p->lambda_max[LIp]);
}

int Mult_NonlocalVector(SQ_LSDFT *p, const double *V1, double *V2, int Kp, int Jp, int Ip)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Mult_NonlocalVector: Initialize first");
    return sq_apply_nonlocal(p->ctx, SQ_NL_REF_COLLAPSED, Kp, Jp, Ip, V1, V2);
}

int Mult_HamiltonianVector(SQ_LSDFT *p, const double *V1, double *V2, int Kp, int Jp, int Ip)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Mult_HamiltonianVector: Initialize first");
    return sq_apply_nonlocal(p->ctx, SQ_NL_KB, Kp, Jp, Ip, V1, V2);
}

int TridiagEigenVecSolve_NodesAndWeights(SQ_LSDFT *p, double diag[], double subdiag[], int n, int LIp)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "TridiagEigenVecSolve_NodesAndWeights: Initialize first");
    if (n != p->N_qp) return fail(SQ_ERR_ARG, "TridiagEigenVecSolve_NodesAndWeights: n differs from N_qp");
    if (LIp < 0 || LIp >= n_owned(p)) return fail(SQ_ERR_ARG, "TridiagEigenVecSolve_NodesAndWeights: LIp out of range");
    const size_t q = (size_t)LIp * n;
    return sq_tridiag_nodes_weights(p->ctx, diag, subdiag, n, 1, &p->lssgq_lambda[q], &p->lssgq_weights[q],
                                    &p->lambda_min[LIp], &p->lambda_max[LIp]);
}

double FermiDirac(double bet, double lambda, double lambdaf)
{
    return 1.0 / (1.0 + exp(bet * (lambda - lambdaf)));
}

int CalculateDensity_lssgq(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "CalculateDensity_lssgq: Initialize first");
    return sq_density_fermi_energy(p->ctx, 0.0, 0.0, &p->lambda_f, p->elecDensRho.data(), &p->Eband, &p->Entropy);
}

int SystemEnergy_Calc_lssgq(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "SystemEnergy_Calc_lssgq: Initialize first");
    printf("Fermi level      = %.15f (Ha)\n", p->lambda_f);
    printf("Band energy      = %.15f (Ha)\n", p->Eband);
    printf("Entropy (-TS)    = %.15f (Ha)\n", p->Entropy);
    return SQ_OK;
}

double SCF_LSSGQ(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "SCF_LSSGQ: Initialize first");
    sq_scf_options o;
    sq_default_scf_options(&o);
    o.nonlocal_mode = p->nonlocal_mode;
    o.precision = p->precision;
    o.max_iter = p->MAXITSCF;
    o.tol = p->TOLSCF;
    o.mixing = p->MixingParameter;
    sq_scf_result r;
    int rc = sq_scf(p->ctx, &o, &r);
    if (rc) return rc;
    p->SCFIterations = r.iterations;
    p->lambda_f = r.lambda_f;
    p->Eband = r.e_band;
    p->Entropy = r.minus_ts;
    p->Exc = r.e_xc;
    p->Eself = r.e_self;
    p->Etot = r.e_total;
    if ((rc = sq_get_cell_field(p->ctx, SQ_FIELD_RHO, p->elecDensRho.data()))) return rc;
    // nodes, weights and extremes of the last SQ pass of the loop
    if ((rc = sq_get_nodes_and_weights(p->ctx, p->lssgq_lambda.data(), p->lssgq_weights.data(), p->lambda_min.data(),
                                       p->lambda_max.data())))
        return rc;
    return r.converged ? SQ_OK : fail(SQ_ERR_RANGE, "SCF_LSSGQ: not converged in MAXIT_SCF iterations");
}

// forces = local + nonlocal, MaxForce = the largest |component|, NormForce =
// the mean over atoms of |F_J|
static void sum_forces(SQ_LSDFT *p)
{
    double mx = 0.0, nrm = 0.0;
    for (int a = 0; a < p->nAtoms; a++) {
        double s2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double f = p->localforces[3 * a + d] + p->nonlocalforces[3 * a + d];
            p->forces[3 * a + d] = f;
            mx = std::max(mx, std::fabs(f));
            s2 += f * f;
        }
        nrm += std::sqrt(s2);
    }
    p->MaxForce = mx;
    p->NormForce = p->nAtoms > 0 ? nrm / p->nAtoms : 0.0;
}

int Calculate_localforce(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Calculate_localforce: Initialize first");
    int rc = sq_forces(p->ctx, nullptr, p->localforces.data(), nullptr, nullptr);
    if (!rc) sum_forces(p);
    return rc;
}

int Calculate_nonlocalforce(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Calculate_nonlocalforce: Initialize first");
    int rc = sq_forces(p->ctx, nullptr, nullptr, p->nonlocalforces.data(), nullptr);
    if (!rc) sum_forces(p);
    return rc;
}

int Display_force(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Display_force: Initialize first");
    printf("Atomic forces (Ha/Bohr):\n");
    for (int a = 0; a < p->nAtoms; a++)
        printf("%4d  % .10E  % .10E  % .10E\n", a, p->forces[3 * a], p->forces[3 * a + 1], p->forces[3 * a + 2]);
    printf("Maximum force component: %.10E   mean |F|: %.10E\n", p->MaxForce, p->NormForce);
    FILE *fp = fopen(p->ForceFilename.c_str(), "w");
    if (!fp) return fail(SQ_ERR_IO, "Couldn't create output file for Forces " + p->ForceFilename);
    for (int a = 0; a < p->nAtoms; a++)
        fprintf(fp, "%.16E %.16E %.16E\n", p->forces[3 * a], p->forces[3 * a + 1], p->forces[3 * a + 2]);
    return fclose(fp) == 0 ? SQ_OK : fail(SQ_ERR_IO, "write failed: " + p->ForceFilename);
}

int Print_ElecDens(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Print_ElecDens: Initialize first");
    const std::string f = rank_file(p->DenFilename, p->rank, ".den");
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return fail(SQ_ERR_IO, "Couldn't create density file " + f);
    fprintf(fp, "%d %d %d %d %d %d\n", p->xcor, p->ycor, p->zcor, p->lxdim, p->lydim, p->lzdim);
    for (int k = p->zcor; k < p->zcor + p->lzdim; k++)
        for (int j = p->ycor; j < p->ycor + p->lydim; j++)
            for (int i = p->xcor; i < p->xcor + p->lxdim; i++) {
                const long LIpos = ((long)(k - p->zcor) * p->lydim + (j - p->ycor)) * p->lxdim + (i - p->xcor);
                fprintf(fp, "%d %d %d %0.16lf\n", k, j, i, p->elecDensRho[LIpos]);
            }
    return fclose(fp) == 0 ? SQ_OK : fail(SQ_ERR_IO, "write failed: " + f);
}

static int read_header(FILE *fp, const SQ_LSDFT *p, const std::string &f)
{
    int h[6];
    if (fscanf(fp, "%d %d %d %d %d %d", &h[0], &h[1], &h[2], &h[3], &h[4], &h[5]) != 6)
        return fail(SQ_ERR_FORMAT, "bad header in " + f);
    const int mine[6] = {p->xcor, p->ycor, p->zcor, p->lxdim, p->lydim, p->lzdim};
    for (int d = 0; d < 6; d++)
        if (h[d] != mine[d]) return fail(SQ_ERR_FORMAT, "Different domain decomposition used in " + f);
    return SQ_OK;
}

int Read_ElecDens(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Read_ElecDens: Initialize first");
    const std::string f = rank_file(p->ReadDenFilename.empty() ? p->DenFilename : p->ReadDenFilename, p->rank, ".den");
    FILE *fp = fopen(f.c_str(), "r");
    if (!fp) return fail(SQ_ERR_IO, "Couldn't open density file " + f);
    int rc = read_header(fp, p, f);
    for (long n = 0; !rc && n < n_owned(p); n++) {
        int K, J, I;
        double v;
        if (fscanf(fp, "%d %d %d %lf", &K, &J, &I, &v) != 4) { rc = fail(SQ_ERR_FORMAT, "short density file " + f); break; }
        if (I < p->xcor || I >= p->xcor + p->lxdim || J < p->ycor || J >= p->ycor + p->lydim || K < p->zcor ||
            K >= p->zcor + p->lzdim) { rc = fail(SQ_ERR_FORMAT, "node outside the rank box in " + f); break; }
        p->elecDensRho[((long)(K - p->zcor) * p->lydim + (J - p->ycor)) * p->lxdim + (I - p->xcor)] = v;
    }
    fclose(fp);
    return rc;
}

int Print_MinMaxLambda(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Print_MinMaxLambda: Initialize first");
    const std::string f = rank_file(p->DenFilename, p->rank, ".lambdas");
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return fail(SQ_ERR_IO, "Couldn't create lambdas file " + f);
    fprintf(fp, "%d %d %d %d %d %d\n", p->xcor, p->ycor, p->zcor, p->lxdim, p->lydim, p->lzdim);
    for (long LIpos = 0; LIpos < n_owned(p); LIpos++)
        fprintf(fp, "%ld %0.16lf %0.16lf\n", LIpos, p->lambda_min[LIpos], p->lambda_max[LIpos]);
    return fclose(fp) == 0 ? SQ_OK : fail(SQ_ERR_IO, "write failed: " + f);
}

int Read_MinMaxLambda(SQ_LSDFT *p)
{
    if (!ready(p)) return fail(SQ_ERR_STATE, "Read_MinMaxLambda: Initialize first");
    const std::string f = rank_file(p->ReadDenFilename.empty() ? p->DenFilename : p->ReadDenFilename, p->rank, ".lambdas");
    FILE *fp = fopen(f.c_str(), "r");
    if (!fp) return fail(SQ_ERR_IO, "Couldn't open lambdas file " + f);
    int rc = read_header(fp, p, f);
    for (long n = 0; !rc && n < n_owned(p); n++) {
        long LIpos;
        double mn, mx;
        if (fscanf(fp, "%ld %lf %lf", &LIpos, &mn, &mx) != 3) { rc = fail(SQ_ERR_FORMAT, "short lambdas file " + f); break; }
        if (LIpos < 0 || LIpos >= n_owned(p)) { rc = fail(SQ_ERR_FORMAT, "LIpos out of range in " + f); break; }
        p->lambda_min[LIpos] = mn;
        p->lambda_max[LIpos] = mx;
    }
    fclose(fp);
    return rc;
}
