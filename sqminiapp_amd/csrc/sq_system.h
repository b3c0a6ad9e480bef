// sq_system.h -- host-side model of one SQ problem (deck, pseudopotentials,
// projector images, coverage map), built without PETSc by sq_setup.cpp and
// uploaded to the device by sq_api.cpp.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sqamd {

constexpr int kMaxType = 8;
constexpr int kMaxOrder = 10;

struct Status {
    int code = 0;
    std::string msg;
    bool ok() const { return code == 0; }
};

// Natural cubic spline with not-a-knot ends on a radial table.
struct RadialSpline {
    const double *x = nullptr;  // abscissae (owned by the Pseudopotential)
    std::vector<double> y, yd;
    int n = 0;
    // Evaluates at r (and, with dout, the cubic's derivative); returns false
    // when r is outside [x0, x_{n-1}].
    bool eval(double r, double *out, double *dout = nullptr) const;
};

struct Pseudopotential {
    std::string element;
    int size = 0, lmax = -1, lloc = 0, valence = 0;
    std::vector<double> r, uu, vloc;
    std::vector<double> V[4], U[4];
    double rc[4] = {0, 0, 0, 0};
    RadialSpline s_vloc, s_udv[4], s_u[4], s_uu;  // s_uu: isolated-atom valence density
    double gamma[4] = {0, 0, 0, 0};
    // (l,m) channel table of the nonlocal part: channel c -> l, m; l != lloc.
    std::vector<int> ch_l, ch_m;
    int nlm() const { return (int)ch_l.size(); }
};

struct Image {
    int type = 0, atom = 0;           // atom = index into the supercell atom list
    double X0 = 0, Y0 = 0, Z0 = 0;    // image position
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // box, inclusive, global unwrapped indices
    int64_t chi_off = 0;              // first entry in System::chi
    int64_t rad_off = 0;              // first entry in System::rad
    int nlm = 0, nl = 0;              // (l,m) channels; nonlocal l values
    int dim(int d) const { return hi[d] - lo[d] + 1; }
    int64_t points() const { return (int64_t)dim(0) * dim(1) * dim(2); }
};

struct System {
    // ---- deck
    int npts[3] = {0, 0, 0};
    double range[3] = {0, 0, 0};
    int order = 6, bc = 1, nqp = 50, ntype = -1;
    double beta = 1000, rcut = 12.0, tol_psc = 1e-6;
    // SCF keywords (Read_Parameters defaults, lsdft_readfiles.cc:62-102, 219-240)
    double mix_beta = 0.3, tol_scf = 1e-6, tol_poisson = 1e-8;
    int mix_history = 7, pulay_frequency = 1, maxit_scf = 100;
    std::string xc = "LDA";
    int localPsd[kMaxType] = {0};
    std::string psdFile[kMaxType];
    // ---- derived grid quantities
    double h[3] = {0, 0, 0}, dV = 0;
    int nloc[3] = {0, 0, 0}, hw[3] = {0, 0, 0};
    double coeff[3][kMaxOrder + 1] = {};
    double cdiag = 0;  // coeff_x[0] + coeff_y[0] + coeff_z[0]
    // ---- ions
    int nAtoms = 0, noetot = 0;
    std::string atomType[kMaxType];
    int noe[kMaxType] = {0}, noa[kMaxType] = {0};
    std::vector<double> atompos;  // 3*nAtoms
    std::vector<int> atomtype;    // per atom
    std::vector<Pseudopotential> psd;
    double cutoff[kMaxType][3] = {};
    // ---- supercell list and alpha slots
    std::vector<double> scpos;    // 3 * n_sc
    std::vector<int> sctype;
    std::vector<int> ip_index;    // first alpha slot per supercell atom
    int ip_length = 0;
    // ---- rank box and its extension by the cube half-width
    int box[6] = {0, 0, 0, 0, 0, 0};
    int ext0[3] = {0, 0, 0}, extn[3] = {0, 0, 0};
    // ---- projector images overlapping the extended box
    std::vector<Image> images;
    std::vector<double> chi;      // per image: (k,j,i) box points outer, channel inner
    // Radial factors of the same projectors: per image and box point, one value
    // per nonlocal l (ascending): T_0 = C00 S_0(r), T_l = S_l(r)/r^l (l >= 1,
    // 0 at r < 1e-10), so chi_lm = T_l C_lm P_lm(x,y,z) with P_lm the
    // homogeneous polynomial of the real harmonic.
    std::vector<double> rad;
    int64_t chi_total = 0, rad_total = 0;   // sizes of chi and rad (also when only the device holds them)
    // ---- per extended-box point: images whose box contains it, ascending
    std::vector<int32_t> cov_off, cov_img;
    int max_cov = 0, max_nlm = 0;

    int64_t nc() const { return (int64_t)nloc[0] * nloc[1] * nloc[2]; }
    int64_t n_nodes() const { return (int64_t)box[3] * box[4] * box[5]; }
    int64_t ext_points() const { return (int64_t)extn[0] * extn[1] * extn[2]; }
};

struct SetupOptions {
    bool phantom_atom = false;
    double phantom_pos[3] = {0, 0, 0};
    int rank_box[6] = {0, 0, 0, 0, 0, 0};  // lx <= 0: whole grid
    int nqp_override = 0;
    bool host_tables = true;   // false: chi and rad are left to the device (build_projector_tables)
};

// Reads the deck and builds every host table of the path.
Status build_system(const std::string &dir, const std::string &name, const SetupOptions &opt, System &sys);

// SCF fields on the periodic cell: the (N0-1) x (N1-1) x (N2-1) distinct grid
// points, i fastest.  b: pseudocharge -(1/4pi) Lap_h V_loc of every atom image
// on its local-overlap box, scaled to -N_e (scale returned), with the self
// energy (1/2) sum_J int b_J V_J; rho: superposed isolated-atom valence
// densities scaled to N_e.
Status build_pseudocharge(const System &S, std::vector<double> &b, double &eself, double &scale);
Status build_atomic_density(const System &S, std::vector<double> &rho, double &scale);
// Pseudocharge forces F[3 * nAtoms] for the potential phi on the periodic cell
// (the derivative of (1/2) int phi (b - rho) - E_self at fixed rho; scale is
// the neutrality factor build_pseudocharge returned).
Status local_forces(const System &S, const double *phi, double scale, std::vector<double> &F);

// chi and rad of every image on the device (sq_setup.hip), with the host
// fill's arithmetic: the tables come out bit-identical.  chi/rad: device
// arrays of S.chi_total / S.rad_total doubles; status: device int, set to
// nonzero when a radius falls off a radial table.
int build_projector_tables(const System &S, double *chi, double *rad, int *status, void *stream);

// Pieces exposed for the reference-named host entry points.
double real_spherical_harmonic(double x, double y, double z, int l, int m, bool *bad);
double fd_fract(int n, int k);

}  // namespace sqamd
