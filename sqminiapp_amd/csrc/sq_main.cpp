// sq_main.cpp -- command-line driver with the steps of the reference's main
// (src/main.cc:17-75) over the reference-named host layer:
//
//   sq_main -name <dir>/<name> [-mode collapsed|kb] [-fp32] [-phantom [x y z]]
//           [-rank_box xcor ycor zcor lx ly lz] [-rank r] [-nqp n] [-device d]
//           [-out <DenFilename>] [-density] [-verify_files] [-scf [-forces]]
//
// It prints the reference's progress lines and its two SQ timers, then writes
// <DenFilename>_<rank>.lambdas (Print_MinMaxLambda) and, with -density, the
// Fermi level, energies and <DenFilename>_<rank>.den (Print_ElecDens).  Any
// failing step prints sq_last_error() and exits with status 1, where the
// reference exits too.  -scf runs SCF_LSSGQ on the periodic cell instead of
// one SQ pass (the node box becomes the cell) and prints its energies; with
// -forces (Kleinman-Bylander mode, -mode kb) it then runs Calculate_localforce,
// Calculate_nonlocalforce and Display_force, which writes ./<name>.force.
// -verify_files reads the written files back with
// Read_MinMaxLambda / Read_ElecDens and checks them against the results.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "../../include/sq_lsdft.h"

static void die(const char *step, int rc)
{
    fprintf(stderr, "%s failed (status %d): %s\n", step, rc, sq_last_error());
    exit(1);
}

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv)
{
    std::string path;
    SQ_LSDFT lsdft;
    sq_options opt;
    sq_default_options(&opt);
    bool density = false, verify = false, scf = false, forces = false;
    std::string out;
    for (int a = 1; a < argc; a++) {
        const std::string s = argv[a];
        auto need = [&](int n) {
            if (a + n >= argc) { fprintf(stderr, "%s needs %d value(s)\n", s.c_str(), n); exit(1); }
        };
        if (s == "-name") { need(1); path = argv[++a]; }
        else if (s == "-mode") {
            need(1);
            const std::string m = argv[++a];
            if (m == "collapsed") lsdft.nonlocal_mode = SQ_NL_REF_COLLAPSED;
            else if (m == "kb") lsdft.nonlocal_mode = SQ_NL_KB;
            else { fprintf(stderr, "-mode is collapsed or kb\n"); return 1; }
        }
        else if (s == "-fp32") lsdft.precision = SQ_FP32;
        else if (s == "-phantom") {
            opt.phantom_atom = 1;
            auto is_num = [](const char *t) { char *e; strtod(t, &e); return e != t && *e == 0; };
            if (a + 3 < argc && is_num(argv[a + 1]) && is_num(argv[a + 2]) && is_num(argv[a + 3])) {
                for (int d = 0; d < 3; d++) opt.phantom_pos[d] = atof(argv[++a]);
            }
        }
        else if (s == "-rank_box") { need(6); for (int d = 0; d < 6; d++) opt.rank_box[d] = atoi(argv[++a]); }
        else if (s == "-rank") { need(1); lsdft.rank = atoi(argv[++a]); }
        else if (s == "-nqp") { need(1); opt.nqp_override = atoi(argv[++a]); }
        else if (s == "-device") { need(1); opt.device = atoi(argv[++a]); }
        else if (s == "-out") { need(1); out = argv[++a]; }
        else if (s == "-density") density = true;
        else if (s == "-verify_files") verify = true;
        else if (s == "-scf") scf = true;
        else if (s == "-forces") forces = true;
        else { fprintf(stderr, "unknown option %s\n", s.c_str()); return 1; }
    }
    if (path.empty()) { fprintf(stderr, "usage: sq_main -name <dir>/<name> [options]\n"); return 1; }
    if (forces && !scf) { fprintf(stderr, "-forces needs -scf\n"); return 1; }
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "." : path.substr(0, slash);
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    lsdft.DenFilename = out.empty() ? "./" + name : out;

    if (scf) {
        // the SCF runs on the distinct points of the periodic cell
        FILE *fp = fopen((dir + "/" + name + ".inpt").c_str(), "r");
        char tok[256];
        int N[3] = {0, 0, 0};
        while (fp && fscanf(fp, "%255s", tok) == 1)
            if (!strcmp(tok, "FD_GRID:") && fscanf(fp, "%d %d %d", &N[0], &N[1], &N[2]) == 3) break;
        if (fp) fclose(fp);
        if (N[0] < 2) { fprintf(stderr, "FD_GRID missing in the deck\n"); return 1; }
        int box[6] = {0, 0, 0, N[0] - 1, N[1] - 1, N[2] - 1};
        memcpy(opt.rank_box, box, sizeof box);
    }
    const clock_t begin0 = clock();
    const double w1 = now();
    int rc = Initialize(&lsdft, dir.c_str(), name.c_str(), &opt);
    if (rc) die("Initialize", rc);
    printf("done influencing atoms local\n");
    printf("done influencing atoms nonlocal\n");
    if ((rc = (int)CalculateNonlocalProjectors(&lsdft))) die("CalculateNonlocalProjectors", rc);
    printf("Nonlocal projectors calculated\n");
    if ((rc = NonlocalProjectors(&lsdft))) die("NonlocalProjectors", rc);
    printf("rank=%d,NEW Nonlocal projectors calculated\n", lsdft.rank);
    printf("rank=%d, number of FD points=%d \n", lsdft.rank, lsdft.lxdim * lsdft.lydim * lsdft.lzdim);

    const clock_t begin = clock();
    const double v1 = now();
    if (scf) {
        if ((rc = (int)SCF_LSSGQ(&lsdft))) die("SCF_LSSGQ", rc);
        printf("SCF converged in %d iterations\n", lsdft.SCFIterations);
        printf("Total energy     = %.12f (Ha)\n", lsdft.Etot);
        printf("Total energy/atom= %.12f (Ha)\n", lsdft.Etot / lsdft.nAtoms);
        printf("Exchange-corr.   = %.12f (Ha)\n", lsdft.Exc);
        printf("Self energy      = %.12f (Ha)\n", lsdft.Eself);
        SystemEnergy_Calc_lssgq(&lsdft);
        density = false;   // rho of the SCF is already in elecDensRho
        if ((rc = Print_ElecDens(&lsdft))) die("Print_ElecDens", rc);
        if (forces) {
            const double f1 = now();
            if ((rc = Calculate_localforce(&lsdft))) die("Calculate_localforce", rc);
            if ((rc = Calculate_nonlocalforce(&lsdft))) die("Calculate_nonlocalforce", rc);
            printf("rank=%d, Time spent in forces=%lf \n", lsdft.rank, now() - f1);
            if ((rc = Display_force(&lsdft))) die("Display_force", rc);
        }
    } else if ((rc = (int)NodesAndWeights_LSSGQ(&lsdft))) die("NodesAndWeights_LSSGQ", rc);
    const double v2 = now();
    const clock_t end = clock();
    printf("rank=%d, Time spent in SQ=%lf \n", lsdft.rank, (double)(end - begin) / CLOCKS_PER_SEC);
    printf("rank=%d, Petsc Time spent in SQ=%lf \n", lsdft.rank, v2 - v1);
    printf("Done LSSGQ \n");
    if ((rc = Print_MinMaxLambda(&lsdft))) die("Print_MinMaxLambda", rc);
    if (density) {
        if ((rc = CalculateDensity_lssgq(&lsdft))) die("CalculateDensity_lssgq", rc);
        SystemEnergy_Calc_lssgq(&lsdft);
        if ((rc = Print_ElecDens(&lsdft))) die("Print_ElecDens", rc);
    }
    if (verify) {
        // %0.16lf keeps 16 decimals: values of magnitude < 100 come back within 1e-14
        const auto lmin = lsdft.lambda_min, lmax = lsdft.lambda_max, rho = lsdft.elecDensRho;
        lsdft.lambda_min.assign(lmin.size(), 0.0);
        lsdft.lambda_max.assign(lmax.size(), 0.0);
        lsdft.elecDensRho.assign(rho.size(), 0.0);
        if ((rc = Read_MinMaxLambda(&lsdft))) die("Read_MinMaxLambda", rc);
        if (density && (rc = Read_ElecDens(&lsdft))) die("Read_ElecDens", rc);
        double err = 0;
        for (size_t q = 0; q < lmin.size(); q++) {
            err = std::max(err, std::fabs(lmin[q] - lsdft.lambda_min[q]));
            err = std::max(err, std::fabs(lmax[q] - lsdft.lambda_max[q]));
            if (density) err = std::max(err, std::fabs(rho[q] - lsdft.elecDensRho[q]));
        }
        if (!(err <= 1e-14)) { fprintf(stderr, "file round trip differs by %.3e\n", err); return 1; }
        printf("files read back: max difference %.3e\n", err);
    }
    const double w2 = now();
    printf("rank=%d, Total time spent=%lf \n", lsdft.rank, (double)(clock() - begin0) / CLOCKS_PER_SEC);
    printf("rank=%d, Petsc Total time spent=%lf \n", lsdft.rank, w2 - w1);
    Objects_Destroy(&lsdft);
    return 0;
}
