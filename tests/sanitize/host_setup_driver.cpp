// Drives the host setup of the product (sqminiapp_amd/csrc/sq_setup.cpp) over
// one deck for an AddressSanitizer + UndefinedBehaviorSanitizer build: the
// system with host projector tables (whole grid, the periodic cell, one rank
// box, with and without the phantom atom), the pseudocharge, the atomic
// density and the local forces of a synthetic potential.  Prints checksums.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../sqminiapp_amd/csrc/sq_system.h"

using namespace sqamd;

static double sum(const std::vector<double> &v)
{
    double s = 0;
    for (double x : v) s += x;
    return s;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s deck_dir name\n", argv[0]);
        return 2;
    }
    for (int variant = 0; variant < 4; variant++) {
        SetupOptions opt;
        opt.host_tables = true;
        opt.phantom_atom = variant == 1;
        System S;
        if (variant >= 2) {
            // the periodic cell, then one interior rank box
            System G;
            Status g = build_system(argv[1], argv[2], SetupOptions(), G);
            if (!g.ok()) { fprintf(stderr, "%s\n", g.msg.c_str()); return 1; }
            const int cell[6] = {0, 0, 0, G.npts[0] - 1, G.npts[1] - 1, G.npts[2] - 1};
            const int part[6] = {G.npts[0] / 4, G.npts[1] / 3, G.npts[2] / 2, G.npts[0] / 2, G.npts[1] / 2, 3};
            for (int d = 0; d < 6; d++) opt.rank_box[d] = variant == 2 ? cell[d] : part[d];
        }
        Status st = build_system(argv[1], argv[2], opt, S);
        if (!st.ok()) { fprintf(stderr, "build_system: %s\n", st.msg.c_str()); return 1; }
        printf("variant %d: %ld nodes, %zu images, chi %ld, rad %ld, chi sum %.12e\n", variant, (long)S.n_nodes(),
               S.images.size(), (long)S.chi_total, (long)S.rad_total, sum(S.chi));
        if (variant != 2) continue;
        std::vector<double> b, rho, F;
        double eself = 0, bscale = 0, rscale = 0;
        if (!(st = build_pseudocharge(S, b, eself, bscale)).ok()) { fprintf(stderr, "%s\n", st.msg.c_str()); return 1; }
        if (!(st = build_atomic_density(S, rho, rscale)).ok()) { fprintf(stderr, "%s\n", st.msg.c_str()); return 1; }
        std::vector<double> phi(b.size());
        for (size_t q = 0; q < phi.size(); q++) phi[q] = 0.1 * std::sin(0.37 * (double)q) + 0.01 * (double)(q % 13);
        if (!(st = local_forces(S, phi.data(), bscale, F)).ok()) { fprintf(stderr, "%s\n", st.msg.c_str()); return 1; }
        printf("  int b dV %.12f, E_self %.12f, int rho dV %.12f, sum F %.6e\n", sum(b) * S.dV, eself,
               sum(rho) * S.dV, sum(F));
    }
    printf("host setup driver done\n");
    return 0;
}
