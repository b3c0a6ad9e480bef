// sq_setup.cpp -- PETSc-free host setup of the SQ hot path.
//
// Builds, from an SQMiniApp input deck, everything the device path consumes:
// grid and FD weights, pseudopotential splines and KB coefficients, the
// pseudocharge cutoff that fixes the periodic image range, the projector
// images with their dense chi boxes, and, per point of the extended rank box,
// the ascending list of images whose box holds it.  The arithmetic of every
// quantity follows the reference function cited at it, so chi is bit-identical
// to the reference's (the cube, image and spline steps have to be alike).
#include "sq_system.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

namespace sqamd {

namespace {

Status fail(int code, const std::string &msg) { return Status{code, msg}; }

// ------------------------------------------------------------------ splines
// Not-a-knot slopes of a cubic spline (Wolberg's formulation, the one of
// getYD_gen/tridiag_gen, lsdft_tools.cc:75-151): a tridiagonal system solved
// by forward elimination and back substitution, in the same operation order.
bool notaknot_slopes(const double *X, const double *Y, int n, std::vector<double> &YD)
{
    YD.assign(n, 0.0);
    std::vector<double> A(n, 0.0), B(n, 0.0), Cc(n, 0.0), F(n, 0.0);
    double h0 = X[1] - X[0], h1 = X[2] - X[1];
    double r0 = (Y[1] - Y[0]) / h0, r1 = (Y[2] - Y[1]) / h1;
    B[0] = h1 * (h0 + h1);
    Cc[0] = (h0 + h1) * (h0 + h1);
    YD[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
    int i = 1;
    for (; i < n - 1; i++) {
        h0 = X[i] - X[i - 1];
        h1 = X[i + 1] - X[i];
        r0 = (Y[i] - Y[i - 1]) / h0;
        r1 = (Y[i + 1] - Y[i]) / h1;
        A[i] = h1;
        B[i] = 2 * (h0 + h1);
        Cc[i] = h0;
        YD[i] = 3 * (r0 * h1 + r1 * h0);
    }
    A[i] = (h0 + h1) * (h0 + h1);
    B[i] = h0 * (h0 + h1);
    YD[i] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
    double b = B[0];
    YD[0] = YD[0] / b;
    for (int k = 1; k < n; k++) {
        F[k] = Cc[k - 1] / b;
        b = B[k] - A[k] * F[k];
        if (b == 0) return false;
        YD[k] = (YD[k] - YD[k - 1] * A[k]) / b;
    }
    for (int k = n - 2; k >= 0; k--) YD[k] -= (YD[k + 1] * F[k + 1]);
    return true;
}

}  // namespace

// Cubic Hermite evaluation on the interval holding r (SplineEvaluate,
// lsdft_tools.cc:15-68).  The interval is the one the reference's linear scan
// stops at: the first knot >= r, stepped back unless r sits on it.
bool RadialSpline::eval(double r, double *out, double *dout) const
{
    if (!(r >= x[0] && r <= x[n - 1])) return false;
    int j = (int)(std::lower_bound(x, x + n, r) - x);
    if (r < x[j]) j--;
    if (j > n - 2) j = n - 2;   // r on the last knot: the last interval
    double p1 = x[j];
    double dx = 1.0 / (x[j + 1] - x[j]);
    double dy = (y[j + 1] - y[j]) * dx;
    double A0 = y[j], A1 = yd[j];
    double A2 = dx * (3.0 * dy - 2.0 * yd[j] - yd[j + 1]);
    double A3 = dx * dx * (-2.0 * dy + yd[j] + yd[j + 1]);
    double t = r - p1;
    *out = ((A3 * t + A2) * t + A1) * t + A0;
    if (dout) *dout = (3.0 * A3 * t + 2.0 * A2) * t + A1;
    return true;
}

// (n!)^2/((n-k)!(n+k)!)  (fract, lsdft_tools.cc:519-531)
double fd_fract(int n, int k)
{
    double num = 1, den = 1;
    for (int i = n - k + 1; i <= n; i++) num *= i;
    for (int i = n + 1; i <= n + k; i++) den *= i;
    return num / den;
}

// Real spherical harmonics for l <= 3 with the reference's coefficient table
// (RealSphericalHarmonic, lsdft_tools.cc:156-240): the same 15-digit
// constants and the same expressions, so projector values match bit for bit.
double real_spherical_harmonic(double x, double y, double z, int l, int m, bool *bad)
{
    static const double c00 = 0.282094791773878;
    static const double c1 = 0.488602511902920;
    static const double c2a = 1.092548430592079, c20 = 0.315391565252520, c22 = 0.546274215296040;
    static const double c33 = 0.590043589926644, c32n = 2.890611442640554, c31 = 0.457045799464466;
    static const double c30 = 0.373176332590115, c32p = 1.445305721320277;
    *bad = false;
    double r = std::sqrt(x * x + y * y + z * z);
    if (l == 0) return c00;
    if (l > 3 || m < -l || m > l) { *bad = true; return 0.0; }
    if (!(r >= 1e-10)) return 0.0;
    switch (l * 10 + (m + 3)) {
    case 12: return c1 * (y / r);
    case 13: return c1 * (z / r);
    case 14: return c1 * (x / r);
    case 21: return c2a * (x * y) / (r * r);
    case 22: return c2a * (y * z) / (r * r);
    case 23: return c20 * (-x * x - y * y + 2.0 * z * z) / (r * r);
    case 24: return c2a * (z * x) / (r * r);
    case 25: return c22 * (x * x - y * y) / (r * r);
    case 30: return c33 * (3 * x * x - y * y) * y / (r * r * r);
    case 31: return c32n * (x * y * z) / (r * r * r);
    case 32: return c31 * y * (4 * z * z - x * x - y * y) / (r * r * r);
    case 33: return c30 * z * (2 * z * z - 3 * x * x - 3 * y * y) / (r * r * r);
    case 34: return c31 * x * (4 * z * z - x * x - y * y) / (r * r * r);
    case 35: return c32p * z * (x * x - y * y) / (r * r * r);
    case 36: return c33 * x * (x * x - 3 * y * y) / (r * r * r);
    }
    *bad = true;
    return 0.0;
}

namespace {

// ------------------------------------------------------------------- decks
bool read_file(const std::string &path, std::string &out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// .inpt keyword scan (Read_Parameters, lsdft_readfiles.cc:148-331, with the
// defaults of 62-102); derived h, n_loc and FD weights (356-476).
Status parse_inpt(const std::string &path, System &S)
{
    std::string text;
    if (!read_file(path, text)) return fail(1, "cannot open " + path);
    std::istringstream in(text);
    std::string tok;
    auto need = [&](auto &v) -> bool { return (bool)(in >> v); };
    while (in >> tok) {
        bool ok = true;
        if (tok == "FD_ORDER:") { ok = need(S.order); S.order /= 2; }
        else if (tok == "BOUNDARY_CONDITION:") ok = need(S.bc);
        else if (tok == "CELL:") ok = need(S.range[0]) && need(S.range[1]) && need(S.range[2]);
        else if (tok == "FD_GRID:") ok = need(S.npts[0]) && need(S.npts[1]) && need(S.npts[2]);
        else if (tok == "BETA:") ok = need(S.beta);
        else if (tok == "RADIUS_TRUNC:") ok = need(S.rcut);
        else if (tok == "LANCZOS_DEGREE:") ok = need(S.nqp);
        else if (tok == "TOL_PSEUDOCHARGE:") ok = need(S.tol_psc);
        else if (tok == "MIXING_PARAMETER:") ok = need(S.mix_beta);
        else if (tok == "MIXING_HISTORY:") ok = need(S.mix_history);
        else if (tok == "PULAY_FREQUENCY:") ok = need(S.pulay_frequency);
        else if (tok == "TOL_SCF:") ok = need(S.tol_scf);
        else if (tok == "TOL_POISSON:") ok = need(S.tol_poisson);
        else if (tok == "MAXIT_SCF:") ok = need(S.maxit_scf);
        else if (tok == "EXCHANGE_CORRELATION:") ok = need(S.xc);
        else if (tok == "NTYPE:") ok = need(S.ntype) && S.ntype >= 1 && S.ntype <= kMaxType;
        else if (tok == "PSEUDOPOTENTIAL_LOCAL:") {
            ok = S.ntype > 0;
            for (int i = 0; ok && i < S.ntype; i++) ok = need(S.localPsd[i]);
        } else if (tok == "PSEUDOPOTENTIAL_FILE:") {
            ok = S.ntype > 0;
            for (int i = 0; ok && i < S.ntype; i++) ok = need(S.psdFile[i]);
        }
        if (!ok) return fail(2, "malformed value after " + tok + " in " + path);
    }
    if (S.ntype <= 0) return fail(2, "NTYPE missing in " + path);
    if (S.order < 1 || S.order > kMaxOrder) return fail(2, "FD_ORDER out of range in " + path);
    if (S.bc != 1) return fail(4, "only BOUNDARY_CONDITION 1 is on the path");
    for (int d = 0; d < 3; d++)
        if (S.npts[d] < 2 || !(S.range[d] > 0)) return fail(2, "bad CELL/FD_GRID in " + path);
    if (S.nqp < 1) return fail(2, "LANCZOS_DEGREE must be positive");
    for (int d = 0; d < 3; d++) {
        S.h[d] = 2.0 * S.range[d] / (S.npts[d] - 1);
        S.hw[d] = (int)std::ceil(S.rcut / S.h[d]);
        S.nloc[d] = 2 * S.hw[d] + 1;
    }
    S.dV = S.h[0] * S.h[1] * S.h[2];
    double c0 = 0;
    for (int a = 1; a <= S.order; a++) c0 += ((double)1.0 / (a * a));
    for (int d = 0; d < 3; d++) {
        S.coeff[d][0] = c0;
        S.coeff[d][0] *= ((double)1.0 / (S.h[d] * S.h[d]));
        for (int p = 1; p <= S.order; p++)
            S.coeff[d][p] = (double)(-1 * std::pow(-1, p + 1) * fd_fract(S.order, p) / (p * p * S.h[d] * S.h[d]));
    }
    S.cdiag = S.coeff[0][0] + S.coeff[1][0] + S.coeff[2][0];
    return {};
}

// .ion blocks "<type> <valence> <count>" + positions (Read_Ion, lsdft_readfiles.cc:677-789)
Status parse_ion(const std::string &path, System &S)
{
    std::string text;
    if (!read_file(path, text)) return fail(1, "cannot open " + path);
    std::istringstream in(text);
    if (!(in >> S.nAtoms) || S.nAtoms < 0) return fail(2, "atom count missing in " + path);
    int at = 0;
    S.noetot = 0;
    std::string name;
    while (at < kMaxType && (in >> name)) {
        if (!(in >> S.noe[at] >> S.noa[at])) break;
        S.atomType[at] = name;
        for (int k = 0; k < S.noa[at]; k++) {
            double x, y, z;
            if (!(in >> x >> y >> z)) return fail(2, "truncated positions in " + path);
            S.atompos.push_back(x); S.atompos.push_back(y); S.atompos.push_back(z);
            S.atomtype.push_back(at);
            S.noetot += S.noe[at];
        }
        at++;
    }
    if ((int)S.atomtype.size() != S.nAtoms) return fail(2, "atom count mismatch in " + path);
    if (at != S.ntype) return fail(2, "number of atom blocks differs from NTYPE in " + path);
    return {};
}

// Troullier-Martins .pot (Read_Pseudopotential, lsdft_readfiles.cc:858-1101):
// radial grid with r=0 prepended; V_l = 0.5 (r V_l)/r (Ry -> Ha), U_l = u_l/r,
// rho = q/(4 pi r^2); at r=0, V and U repeat their first value and rho is 0.
Status parse_pot(const std::string &path, const std::string &element, int lloc, Pseudopotential &P)
{
    std::string text;
    if (!read_file(path, text)) return fail(1, "cannot open " + path);
    // Token stream in which the block header lines become marker tokens.
    enum Kind { NUM, RADIAL, PSP, CORE, VALENCE, WAVE, OTHER };
    struct Tok { Kind k; std::string s; };
    std::vector<Tok> toks;
    std::istringstream lines(text);
    std::string line;
    bool first = true;
    std::string head;
    while (std::getline(lines, line)) {
        std::string t = line;
        t.erase(0, t.find_first_not_of(" \t\r"));
        if (first) { head = t.substr(0, t.find_first_of(" \t\r")); first = false; continue; }
        if (t.rfind("Radial grid follows", 0) == 0) { toks.push_back({RADIAL, ""}); continue; }
        if (t.rfind("Pseudopotential follows", 0) == 0) { toks.push_back({PSP, ""}); continue; }
        if (t.rfind("Core charge follows", 0) == 0) { toks.push_back({CORE, ""}); continue; }
        if (t.rfind("Valence charge follows", 0) == 0) { toks.push_back({VALENCE, ""}); continue; }
        if (t.rfind("Pseudo-wave-function follows", 0) == 0) { toks.push_back({WAVE, ""}); continue; }
        std::istringstream ls(line);
        std::string w;
        while (ls >> w) toks.push_back({NUM, w});
    }
    if (head != element) return fail(2, path + " is not a pseudopotential of " + element);
    size_t i = 0;
    while (i < toks.size() && toks[i].k != RADIAL) i++;
    if (i == toks.size()) return fail(2, "no radial grid in " + path);
    i++;
    std::vector<double> rv;
    while (i < toks.size() && toks[i].k == NUM) rv.push_back(std::strtod(toks[i++].s.c_str(), nullptr));
    int count = (int)rv.size();
    if (count < 3) return fail(2, "radial grid too short in " + path);
    P.element = element;
    P.size = count + 1;
    P.r.assign(P.size, 0.0);
    for (int k = 0; k < count; k++) P.r[k + 1] = rv[k];
    P.uu.assign(P.size, 0.0);
    P.vloc.assign(P.size, 0.0);
    for (int l = 0; l < 4; l++) { P.V[l].assign(P.size, 0.0); P.U[l].assign(P.size, 0.0); P.rc[l] = 0.0; }
    auto take = [&](double &v) -> bool {
        if (i >= toks.size() || toks[i].k != NUM) return false;
        v = std::strtod(toks[i++].s.c_str(), nullptr);
        return true;
    };
    bool have_valence = false;
    P.lmax = -1;
    while (i < toks.size()) {
        Kind k = toks[i++].k;
        if (k == PSP || k == WAVE) {
            double lv, zel = 0, rc = 0;
            if (!take(lv)) return fail(2, "missing l in " + path);
            int l = (int)lv;
            if (l < 0 || l > 3) return fail(2, "l out of range in " + path);
            if (k == WAVE && !(take(zel) && take(rc))) return fail(2, "missing zelect/rc in " + path);
            for (int q = 1; q <= count; q++) {
                double v;
                if (!take(v)) return fail(2, "truncated block in " + path);
                if (k == PSP) {
                    v = 0.5 * v / (P.r[q]);
                    P.V[l][q] = v;
                    if (lloc == l) P.vloc[q] = v;
                } else {
                    v = v / P.r[q];
                    P.U[l][q] = v;
                }
            }
            if (k == PSP) {
                P.V[l][0] = P.V[l][1];
                if (lloc == l) P.vloc[0] = P.V[l][1];
                P.lmax = std::max(P.lmax, l);
            } else {
                P.rc[l] = rc;
                P.U[l][0] = P.U[l][1];
            }
        } else if (k == VALENCE) {
            for (int q = 1; q <= count; q++) {
                double v;
                if (!take(v)) return fail(2, "truncated valence charge in " + path);
                P.uu[q] = v / (4 * M_PI * P.r[q] * P.r[q]);
            }
            P.uu[0] = 0.0;
            have_valence = true;
        } else if (k == CORE) {
            while (i < toks.size() && toks[i].k == NUM) i++;
        }
    }
    if (P.lmax < 0 || lloc > P.lmax || !have_valence) return fail(2, "incomplete pseudopotential " + path);
    P.lloc = lloc;
    return {};
}

// Spline fits (Calculate_SplineFit, lsdft_initObjs.cc:576-622) and Gamma_l
// (Calculate_NonlocalCoefficients, lsdft_initObjs.cc:822-869): a rectangle
// rule, dr = 1e-6, of U dV U r^2 to rc_l + max h, half of the last sample
// removed, divided by dV, inverted.
Status fit_pseudopotential(const System &S, Pseudopotential &P)
{
    auto fit = [&](RadialSpline &s, const std::vector<double> &y) {
        s.x = P.r.data();
        s.n = P.size;
        s.y = y;
        return notaknot_slopes(P.r.data(), y.data(), P.size, s.yd);
    };
    if (!fit(P.s_vloc, P.vloc)) return fail(3, "singular spline system (Vloc)");
    if (!fit(P.s_uu, P.uu)) return fail(3, "singular spline system (valence density)");
    for (int l = 0; l <= P.lmax; l++) {
        std::vector<double> udv(P.size);
        for (int q = 0; q < P.size; q++) udv[q] = P.U[l][q] * (P.V[l][q] - P.vloc[q]);
        if (!fit(P.s_udv[l], udv) || !fit(P.s_u[l], P.U[l])) return fail(3, "singular spline system");
    }
    const double dr = 1e-6;
    double a = S.h[0], b = S.h[1], c = S.h[2];
    double delta = (a > b && a > c ? a : b > c ? b : c);
    for (int l = 0; l <= P.lmax; l++) {
        P.gamma[l] = 0.0;
        if (l == P.lloc) continue;
        double rt = dr, D = 0.0, f1 = 0.0, f2 = 0.0;
        while (rt <= (P.rc[l] + delta)) {
            if (!P.s_udv[l].eval(rt, &f1) || !P.s_u[l].eval(rt, &f2)) return fail(3, "Gamma quadrature off the radial grid");
            D += f1 * f2 * rt * rt * dr;
            rt = rt + dr;
        }
        D = (D - 0.5 * f1 * f2 * (rt - dr) * (rt - dr) * dr) / (S.dV);
        P.gamma[l] = 1.0 / D;
    }
    P.ch_l.clear();
    P.ch_m.clear();
    for (int l = 0; l <= P.lmax; l++) {
        if (l == P.lloc) continue;
        for (int m = -l; m <= l; m++) { P.ch_l.push_back(l); P.ch_m.push_back(m); }
    }
    return {};
}

// Pseudocharge cutoff (Calculate_PseudochargeCutoff, lsdft_initObjs.cc:628-817).
// Only its result is used here: ceil(cutoff/R) bounds the periodic images.
Status pseudocharge_cutoff(System &S, int at)
{
    const Pseudopotential &P = S.psd[at];
    const int o = S.order;
    int Np[3];
    for (int d = 0; d < 3; d++) Np[d] = (int)std::ceil(15.0 / S.h[d]);
    const int Ex = Np[0] + 2 * o, Ey = Np[1] + 2 * o, Ez = Np[2] + 2 * o;
    std::vector<double> vps((size_t)Ex * Ey * Ez), bj((size_t)Np[0] * Np[1] * Np[2]);
    auto VP = [&](int k, int j, int i) -> double & { return vps[((size_t)k * Ey + j) * Ex + i]; };
    for (int k = 0; k < Ez; k++)
        for (int j = 0; j < Ey; j++)
            for (int i = 0; i < Ex; i++) {
                double x = (i - o) * S.h[0], y = (j - o) * S.h[1], z = (k - o) * S.h[2];
                double r = std::sqrt(x * x + y * y + z * z);
                if (!P.s_vloc.eval(r, &VP(k, j, i))) return fail(3, "pseudocharge grid beyond the radial table");
            }
    double sc[3][kMaxOrder + 1];
    for (int d = 0; d < 3; d++)
        for (int p = 0; p <= o; p++) sc[d][p] = S.coeff[d][p] / (2.0 * M_PI);
    for (int k = o; k < Np[2] + o; k++)
        for (int j = o; j < Np[1] + o; j++)
            for (int i = o; i < Np[0] + o; i++) {
                double v = VP(k, j, i) * (sc[0][0] + sc[1][0] + sc[2][0]);
                for (int a = 1; a <= o; a++)
                    v += (VP(k, j, i - a) + VP(k, j, i + a)) * sc[0][a] + (VP(k, j - a, i) + VP(k, j + a, i)) * sc[1][a] +
                         (VP(k - a, j, i) + VP(k + a, j, i)) * sc[2][a];
                bj[((size_t)(k - o) * Np[1] + (j - o)) * Np[0] + (i - o)] = v;
            }
    double Rc = P.rc[P.lloc], err;
    int Nc[3];
    do {
        for (int d = 0; d < 3; d++) Nc[d] = (int)std::ceil(Rc / S.h[d]);
        double B = 0.0;
        for (int k = 0; k < Nc[2]; k++)
            for (int j = 0; j < Nc[1]; j++)
                for (int i = 0; i < Nc[0]; i++) {
                    double v = bj[((size_t)k * Np[1] + j) * Np[0] + i];
                    if (k == 0) v = 0.5 * v;
                    if (k == Nc[2] - 1) v = 0.5 * v;
                    if (j == 0) v = 0.5 * v;
                    if (j == Nc[1] - 1) v = 0.5 * v;
                    if (i == 0) v = 0.5 * v;
                    if (i == Nc[0] - 1) v = 0.5 * v;
                    B += v;
                }
        B = B * 8.0 * S.dV;
        err = std::fabs(B + S.noe[at]) / std::fabs((double)S.noe[at]);
        Rc = Rc + 1.0;
    } while (Rc <= 15.0 && err > S.tol_psc);
    for (int d = 0; d < 3; d++) S.cutoff[at][d] = Nc[d] * S.h[d];
    return {};
}

// Clip [s,e] to the window [w0, w0+n-1] the way Check_Overlap
// (lsdft_initObjs.cc:2612-2663) does; false when an end stays unset.
bool clip_axis(int w0, int n, int s, int e, int *cs, int *ce)
{
    const int unset = -1000;
    int a = unset, b = unset;
    if (s >= w0 && s <= w0 + n - 1) a = s;
    else if (w0 >= s && w0 <= e) a = w0;
    if (e >= w0 && e <= w0 + n - 1) b = e;
    else if (w0 + n - 1 >= s && w0 + n - 1 <= e) b = w0 + n - 1;
    *cs = a;
    *ce = b;
    return a != unset && b != unset;
}

// Nearest grid index of a coordinate, rounded in single precision as the
// reference's roundf((x0 + R)/h) does (lsdft_initObjs.cc:1800-1802).
int nearest_index(double v) { return (int)roundf((float)v); }

// Images of every supercell atom inside the periodic range that overlap the
// extended rank box, in the reference's order (atoms, then P, Q, R;
// GetInfluencingAtoms_Nonlocal, lsdft_initObjs.cc:1671-1946).
void collect_images(System &S)
{
    S.images.clear();
    const int nsc = (int)S.sctype.size();
    for (int at = 0; at < S.ntype; at++) {
        const Pseudopotential &P = S.psd[at];
        if (P.lmax == 0) continue;
        double m1 = P.rc[0] > P.rc[1] ? P.rc[0] : P.rc[1];
        double m2 = P.rc[2] > P.rc[3] ? P.rc[2] : P.rc[3];
        double cut = m1 > m2 ? m1 : m2;
        int off[3], Im[3];
        for (int d = 0; d < 3; d++) {
            off[d] = (int)std::ceil(cut / S.h[d]);
            Im[d] = (int)std::ceil(S.cutoff[at][d] / S.range[d]);
        }
        for (int p = 0; p < nsc; p++) {
            if (S.sctype[p] != at) continue;
            const double X0 = S.scpos[3 * p], Y0 = S.scpos[3 * p + 1], Z0 = S.scpos[3 * p + 2];
            for (int PP = -Im[0]; PP <= Im[0]; PP++)
                for (int QQ = -Im[1]; QQ <= Im[1]; QQ++)
                    for (int RR = -Im[2]; RR <= Im[2]; RR++) {
                        double x0 = X0 + PP * 2.0 * S.range[0];
                        double y0 = Y0 + QQ * 2.0 * S.range[1];
                        double z0 = Z0 + RR * 2.0 * S.range[2];
                        int c[3] = {nearest_index((x0 + S.range[0] + 0 * S.rcut) / S.h[0]),
                                    nearest_index((y0 + S.range[1] + 0 * S.rcut) / S.h[1]),
                                    nearest_index((z0 + S.range[2] + 0 * S.rcut) / S.h[2])};
                        Image im;
                        bool ok = true;
                        for (int d = 0; d < 3 && ok; d++)
                            ok = clip_axis(S.ext0[d], S.extn[d], c[d] - off[d], c[d] + off[d], &im.lo[d], &im.hi[d]);
                        if (!ok) continue;
                        im.type = at;
                        im.atom = p;
                        im.X0 = x0; im.Y0 = y0; im.Z0 = z0;
                        im.nlm = P.nlm();
                        S.images.push_back(im);
                    }
        }
    }
}

// chi_lm(x) = S_l(|x - R|) Y_lm(x - R) on each image box, stored with the box
// point outer and channel inner (CalculateNonlocalProjectors,
// lsdft_nl.cc:141-208, flattened as NonlocalProjectors, lsdft_nl.cc:81-123).
void layout_projectors(System &S)
{
    int64_t total = 0;
    S.max_nlm = 0;
    for (Image &im : S.images) {
        im.chi_off = total;
        total += im.points() * im.nlm;
        S.max_nlm = std::max(S.max_nlm, im.nlm);
    }
    S.chi_total = total;
    int64_t rtotal = 0;
    for (Image &im : S.images) {
        const Pseudopotential &P = S.psd[im.type];
        im.nl = 0;
        for (int l = 0; l <= P.lmax; l++)
            if (l != P.lloc) im.nl++;
        im.rad_off = rtotal;
        rtotal += im.points() * im.nl;
    }
    S.rad_total = rtotal;
}

Status fill_projectors(System &S)
{
    S.chi.assign((size_t)S.chi_total, 0.0);
    S.rad.assign((size_t)S.rad_total, 0.0);
    std::atomic<int> next{0}, bad{0};
    auto worker = [&]() {
        for (;;) {
            int c = next.fetch_add(1);
            if (c >= (int)S.images.size() || bad.load()) return;
            const Image &im = S.images[c];
            const Pseudopotential &P = S.psd[im.type];
            double *out = S.chi.data() + im.chi_off;
            double *rout = S.rad.data() + im.rad_off;
            for (int k = im.lo[2]; k <= im.hi[2]; k++) {
                double z = S.h[2] * k - S.range[2] - im.Z0, zz = z * z;
                for (int j = im.lo[1]; j <= im.hi[1]; j++) {
                    double y = S.h[1] * j - S.range[1] - im.Y0, yy = y * y;
                    for (int i = im.lo[0]; i <= im.hi[0]; i++) {
                        double x = S.h[0] * i - S.range[0] - im.X0;
                        double r = std::sqrt(x * x + yy + zz);
                        double radial[4] = {0, 0, 0, 0};
                        for (int l = 0; l <= P.lmax; l++) {
                            if (l == P.lloc) continue;
                            if (!P.s_udv[l].eval(r, &radial[l])) { bad = 3; return; }
                        }
                        for (int ch = 0; ch < im.nlm; ch++) {
                            bool b = false;
                            double sh = real_spherical_harmonic(x, y, z, P.ch_l[ch], P.ch_m[ch], &b);
                            if (b) { bad = 4; return; }
                            *out++ = radial[P.ch_l[ch]] * sh;
                        }
                        for (int l = 0; l <= P.lmax; l++) {
                            if (l == P.lloc) continue;
                            double t;
                            if (l == 0) t = radial[0] * 0.282094791773878;
                            else if (!(r >= 1e-10)) t = 0.0;
                            else {
                                double rl = r;
                                for (int q = 1; q < l; q++) rl *= r;
                                t = radial[l] / rl;
                            }
                            *rout++ = t;
                        }
                    }
                }
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    if (bad) return fail(bad.load(), "projector evaluation failed (radius off the table or bad l)");
    return {};
}

// For every point of the extended box, the ascending list of images whose box
// holds it: the order in which the reference adds chi*alpha into a point
// (lsdft_matvec.cc:110-142).
void build_coverage(System &S)
{
    const int64_t np = S.ext_points();
    std::vector<int32_t> cnt(np + 1, 0);
    auto eidx = [&](int i, int j, int k) {
        return ((int64_t)(k - S.ext0[2]) * S.extn[1] + (j - S.ext0[1])) * S.extn[0] + (i - S.ext0[0]);
    };
    for (const Image &im : S.images)
        for (int k = im.lo[2]; k <= im.hi[2]; k++)
            for (int j = im.lo[1]; j <= im.hi[1]; j++)
                for (int i = im.lo[0]; i <= im.hi[0]; i++) cnt[eidx(i, j, k) + 1]++;
    S.max_cov = 0;
    for (int64_t p = 0; p < np; p++) S.max_cov = std::max(S.max_cov, cnt[p + 1]);
    for (int64_t p = 0; p < np; p++) cnt[p + 1] += cnt[p];
    S.cov_off = cnt;
    S.cov_img.assign(cnt[np], 0);
    std::vector<int32_t> fillp(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < (int)S.images.size(); c++) {
        const Image &im = S.images[c];
        for (int k = im.lo[2]; k <= im.hi[2]; k++)
            for (int j = im.lo[1]; j <= im.hi[1]; j++)
                for (int i = im.lo[0]; i <= im.hi[0]; i++) S.cov_img[fillp[eidx(i, j, k)]++] = c;
    }
}

}  // namespace

// ------------------------------------------------------------ SCF fields
// The reference builds these pieces but never assembles them (its SCF is
// declared only): the pseudocharge -(1/4pi) Lap_h V_loc of
// Calculate_PseudochargeCutoff (lsdft_initObjs.cc:628-817) placed on each
// atom image's local-overlap box (GetInfluencingAtoms_Local, 1409-1665), and
// the isolated-atom valence density of the .pot file (lsdft_readfiles.cc:
// 1026-1040).  The grid is the periodic cell of distinct points: the images
// repeat every 2R = (N-1) h, so plane N-1 repeats plane 0.  Threads split the
// cell into z-slabs and each walks every image in the same order, so every
// point sums its images in one fixed order whatever the thread count.

namespace {

bool vloc_at(const System &S, int at, double r, double *out)
{
    const Pseudopotential &P = S.psd[at];
    if (r > P.r[P.size - 1]) { *out = -(double)S.noe[at] / r; return true; }  // Coulomb tail
    return P.s_vloc.eval(r, out);
}

struct CellImage {
    int at, atom;       // type, atom index
    double X[3];
    int lo[3], hi[3];   // box, unwrapped global indices
    bool self;          // image 0 of an atom
};

// Atom images in the loop order atom type, atom, shift (x, y, z), whose box
// meets the cell or that are an atom's own (unshifted) image.
std::vector<CellImage> cell_images(const System &S, const int G[3])
{
    std::vector<CellImage> out;
    for (int at = 0; at < S.ntype; at++) {
        int Imax[3], off[3];
        for (int d = 0; d < 3; d++) {
            Imax[d] = (int)std::ceil(S.cutoff[at][d] / S.range[d]);
            off[d] = (int)std::ceil(S.cutoff[at][d] / S.h[d] + 0.5);
        }
        for (int p = 0; p < S.nAtoms; p++) {
            if (S.atomtype[p] != at) continue;
            const double *A = &S.atompos[3 * p];
            for (int PP = -Imax[0]; PP <= Imax[0]; PP++)
                for (int QQ = -Imax[1]; QQ <= Imax[1]; QQ++)
                    for (int RR = -Imax[2]; RR <= Imax[2]; RR++) {
                        CellImage im;
                        im.at = at;
                        im.atom = p;
                        im.X[0] = A[0] + PP * 2.0 * S.range[0];
                        im.X[1] = A[1] + QQ * 2.0 * S.range[1];
                        im.X[2] = A[2] + RR * 2.0 * S.range[2];
                        bool meets = true;
                        for (int d = 0; d < 3; d++) {
                            const int c = nearest_index((im.X[d] + S.range[d]) / S.h[d]);
                            im.lo[d] = c - off[d];
                            im.hi[d] = c + off[d];
                            meets = meets && im.lo[d] <= G[d] - 1 && im.hi[d] >= 0;
                        }
                        im.self = PP == 0 && QQ == 0 && RR == 0;
                        if (meets || im.self) out.push_back(im);
                    }
        }
    }
    return out;
}

// b_J = -(1/4pi) Lap_h V_J on the sub-box lo..hi (inclusive) of an image, in
// the expression and order of lsdft_initObjs.cc:725-735; V sampled on the
// sub-box widened by the stencil half-width.  bj (and vj, if given) are the
// sub-box, i fastest.
bool image_pseudocharge(const System &S, const CellImage &im, const int lo[3], const int hi[3], double sc[3][kMaxOrder + 1],
                        std::vector<double> &bj, std::vector<double> *vj)
{
    const int o = S.order;
    int n[3], e[3];
    for (int d = 0; d < 3; d++) { n[d] = hi[d] - lo[d] + 1; e[d] = n[d] + 2 * o; }
    std::vector<double> V((size_t)e[0] * e[1] * e[2]);
    auto VV = [&](int k, int j, int i) -> double & { return V[((size_t)k * e[1] + j) * e[0] + i]; };
    for (int k = 0; k < e[2]; k++)
        for (int j = 0; j < e[1]; j++)
            for (int i = 0; i < e[0]; i++) {
                double x = S.h[0] * (lo[0] - o + i) - S.range[0] - im.X[0];
                double y = S.h[1] * (lo[1] - o + j) - S.range[1] - im.X[1];
                double z = S.h[2] * (lo[2] - o + k) - S.range[2] - im.X[2];
                double r = std::sqrt(x * x + y * y + z * z);
                if (!vloc_at(S, im.at, r, &VV(k, j, i))) return false;
            }
    bj.assign((size_t)n[0] * n[1] * n[2], 0.0);
    if (vj) vj->assign(bj.size(), 0.0);
    for (int k = o; k < n[2] + o; k++)
        for (int j = o; j < n[1] + o; j++)
            for (int i = o; i < n[0] + o; i++) {
                double v = VV(k, j, i) * (sc[0][0] + sc[1][0] + sc[2][0]);
                for (int a = 1; a <= o; a++)
                    v += (VV(k, j, i - a) + VV(k, j, i + a)) * sc[0][a] + (VV(k, j - a, i) + VV(k, j + a, i)) * sc[1][a] +
                         (VV(k - a, j, i) + VV(k + a, j, i)) * sc[2][a];
                const size_t q = ((size_t)(k - o) * n[1] + (j - o)) * n[0] + (i - o);
                bj[q] = v;
                if (vj) (*vj)[q] = VV(k, j, i);
            }
    return true;
}

unsigned setup_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

}  // namespace

Status build_pseudocharge(const System &S, std::vector<double> &b, double &eself, double &scale)
{
    const int G[3] = {S.npts[0] - 1, S.npts[1] - 1, S.npts[2] - 1};
    for (int d = 0; d < 3; d++)
        if (G[d] < 1) return fail(4, "the periodic cell needs at least two grid points per direction");
    const size_t ng = (size_t)G[0] * G[1] * G[2];
    b.assign(ng, 0.0);
    double sc[3][kMaxOrder + 1];
    for (int d = 0; d < 3; d++)
        for (int p = 0; p <= S.order; p++) sc[d][p] = S.coeff[d][p] / (2.0 * M_PI);
    const std::vector<CellImage> imgs = cell_images(S, G);
    const unsigned nt = setup_threads();
    std::atomic<int> bad{0};
    // slabs of z-planes: each thread adds every image to its own planes, in image order
    auto slab_worker = [&](unsigned t) {
        const int z0 = (int)((long)G[2] * t / nt), z1 = (int)((long)G[2] * (t + 1) / nt) - 1;
        if (z0 > z1) return;
        std::vector<double> bj;
        for (const CellImage &im : imgs) {
            int lo[3], hi[3];
            bool empty = false;
            for (int d = 0; d < 3; d++) {
                lo[d] = std::max(im.lo[d], 0);
                hi[d] = std::min(im.hi[d], G[d] - 1);
            }
            lo[2] = std::max(lo[2], z0);
            hi[2] = std::min(hi[2], z1);
            for (int d = 0; d < 3; d++) empty = empty || lo[d] > hi[d];
            if (empty) continue;
            if (!image_pseudocharge(S, im, lo, hi, sc, bj, nullptr)) { bad = 3; return; }
            const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++)
                        b[((size_t)k * G[1] + j) * G[0] + i] += bj[((size_t)(k - lo[2]) * ny + (j - lo[1])) * nx + (i - lo[0])];
        }
    };
    // self energy of each atom over its whole box, summed afterwards in atom order
    std::vector<const CellImage *> selves;
    for (const CellImage &im : imgs)
        if (im.self) selves.push_back(&im);
    std::vector<std::vector<double>> prod(selves.size());
    std::atomic<size_t> next{0};
    auto self_worker = [&]() {
        std::vector<double> bj, vj;
        for (size_t a; (a = next++) < selves.size();) {
            if (!image_pseudocharge(S, *selves[a], selves[a]->lo, selves[a]->hi, sc, bj, &vj)) { bad = 3; return; }
            prod[a].resize(bj.size());
            for (size_t q = 0; q < bj.size(); q++) prod[a][q] = bj[q] * vj[q];
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++) pool.emplace_back(slab_worker, t);
    for (auto &t : pool) t.join();
    pool.clear();
    for (unsigned t = 0; t < nt; t++) pool.emplace_back(self_worker);
    for (auto &t : pool) t.join();
    if (bad) return fail(3, "pseudocharge: V_loc spline evaluation failed");
    double es = 0.0;
    for (const auto &pr : prod)
        for (double v : pr) es += v;
    double tot = 0.0;
    for (size_t q = 0; q < ng; q++) tot += b[q];
    tot *= S.dV;
    const double f = -(double)S.noetot / tot;
    for (size_t q = 0; q < ng; q++) b[q] *= f;
    scale = f;
    eself = 0.5 * es * S.dV * f;
    return {};
}

// Local forces.  With B = -(1/4pi) Lap_h (the pseudocharge stencil) and
// g = grad V_loc sampled on an image's box widened by the stencil, the
// derivative of b_J = B V_J with respect to the atom is -B g, so
//   F_J = f dV sum_{J' of J} (B g_J')^T phi  -  (f dV / 2) [(B g_s)^T V_s + b_s^T g_s],
// the first sum over the part of each image box inside the cell (where b is
// placed), the self term over the atom's own whole box (b_s, V_s unscaled,
// as E_self is formed).  Images are shared out to threads; each image's
// partial sums are kept and added per atom in image order.
Status local_forces(const System &S, const double *phi, double scale, std::vector<double> &F)
{
    const int G[3] = {S.npts[0] - 1, S.npts[1] - 1, S.npts[2] - 1};
    const int o = S.order;
    double sc[3][kMaxOrder + 1];
    for (int d = 0; d < 3; d++)
        for (int p = 0; p <= o; p++) sc[d][p] = S.coeff[d][p] / (2.0 * M_PI);
    const std::vector<CellImage> imgs = cell_images(S, G);
    std::vector<std::array<double, 6>> part(imgs.size());
    std::atomic<size_t> next{0};
    std::atomic<int> bad{0};
    auto worker = [&]() {
        std::vector<double> g, bj, vj;
        for (size_t q; (q = next++) < imgs.size();) {
            const CellImage &im = imgs[q];
            int clo[3], chi[3], n[3], e[3];
            bool empty = false;
            for (int d = 0; d < 3; d++) {
                clo[d] = std::max(im.lo[d], 0);
                chi[d] = std::min(im.hi[d], G[d] - 1);
                empty = empty || clo[d] > chi[d];
                n[d] = im.hi[d] - im.lo[d] + 1;
                e[d] = n[d] + 2 * o;
            }
            part[q].fill(0.0);
            if (empty && !im.self) continue;
            const Pseudopotential &P = S.psd[im.at];
            g.assign((size_t)3 * e[0] * e[1] * e[2], 0.0);
            for (int k = 0; k < e[2]; k++)
                for (int j = 0; j < e[1]; j++)
                    for (int i = 0; i < e[0]; i++) {
                        const double x = S.h[0] * (im.lo[0] - o + i) - S.range[0] - im.X[0];
                        const double y = S.h[1] * (im.lo[1] - o + j) - S.range[1] - im.X[1];
                        const double z = S.h[2] * (im.lo[2] - o + k) - S.range[2] - im.X[2];
                        const double r = std::sqrt(x * x + y * y + z * z);
                        double v, dv;
                        if (r > P.r[P.size - 1]) dv = (double)S.noe[im.at] / (r * r);
                        else if (!P.s_vloc.eval(r, &v, &dv)) { bad = 3; return; }
                        if (r < 1e-10) continue;
                        double *gq = &g[3 * (((size_t)k * e[1] + j) * e[0] + i)];
                        gq[0] = dv * x / r;
                        gq[1] = dv * y / r;
                        gq[2] = dv * z / r;
                    }
            if (im.self && !image_pseudocharge(S, im, im.lo, im.hi, sc, bj, &vj)) { bad = 3; return; }
            auto GG = [&](int k, int j, int i, int d) { return g[3 * (((size_t)k * e[1] + j) * e[0] + i) + d]; };
            double fphi[3] = {0, 0, 0}, fself[3] = {0, 0, 0};
            for (int k = o; k < n[2] + o; k++)
                for (int j = o; j < n[1] + o; j++)
                    for (int i = o; i < n[0] + o; i++) {
                        const int gi = im.lo[0] - o + i, gj = im.lo[1] - o + j, gk = im.lo[2] - o + k;
                        const bool inside = !empty && gi >= clo[0] && gi <= chi[0] && gj >= clo[1] && gj <= chi[1] &&
                                            gk >= clo[2] && gk <= chi[2];
                        if (!inside && !im.self) continue;
                        const size_t qb = ((size_t)(k - o) * n[1] + (j - o)) * n[0] + (i - o);
                        for (int d = 0; d < 3; d++) {
                            double v = GG(k, j, i, d) * (sc[0][0] + sc[1][0] + sc[2][0]);
                            for (int a = 1; a <= o; a++)
                                v += (GG(k, j, i - a, d) + GG(k, j, i + a, d)) * sc[0][a] +
                                     (GG(k, j - a, i, d) + GG(k, j + a, i, d)) * sc[1][a] +
                                     (GG(k - a, j, i, d) + GG(k + a, j, i, d)) * sc[2][a];
                            if (inside) fphi[d] += v * phi[((size_t)gk * G[1] + gj) * G[0] + gi];
                            if (im.self) fself[d] += v * vj[qb] + bj[qb] * GG(k, j, i, d);
                        }
                    }
            for (int d = 0; d < 3; d++) { part[q][d] = fphi[d]; part[q][3 + d] = fself[d]; }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < setup_threads(); t++) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    if (bad) return fail(3, "local forces: V_loc spline evaluation failed");
    std::vector<double> fphi((size_t)3 * S.nAtoms, 0.0), fself((size_t)3 * S.nAtoms, 0.0);
    for (size_t q = 0; q < imgs.size(); q++)
        for (int d = 0; d < 3; d++) {
            fphi[3 * imgs[q].atom + d] += part[q][d];
            fself[3 * imgs[q].atom + d] += part[q][3 + d];
        }
    F.assign((size_t)3 * S.nAtoms, 0.0);
    for (size_t q = 0; q < F.size(); q++) F[q] = scale * S.dV * fphi[q] - 0.5 * scale * S.dV * fself[q];
    return {};
}

Status build_atomic_density(const System &S, std::vector<double> &rho, double &scale)
{
    const int G[3] = {S.npts[0] - 1, S.npts[1] - 1, S.npts[2] - 1};
    for (int d = 0; d < 3; d++)
        if (G[d] < 1) return fail(4, "the periodic cell needs at least two grid points per direction");
    const size_t ng = (size_t)G[0] * G[1] * G[2];
    rho.assign(ng, 0.0);
    const std::vector<CellImage> imgs = cell_images(S, G);
    const unsigned nt = setup_threads();
    std::atomic<int> bad{0};
    auto worker = [&](unsigned t) {
        const int z0 = (int)((long)G[2] * t / nt), z1 = (int)((long)G[2] * (t + 1) / nt) - 1;
        for (const CellImage &im : imgs) {
            const Pseudopotential &P = S.psd[im.at];
            int lo[3], hi[3];
            for (int d = 0; d < 3; d++) {
                lo[d] = std::max(im.lo[d], 0);
                hi[d] = std::min(im.hi[d], G[d] - 1);
            }
            lo[2] = std::max(lo[2], z0);
            hi[2] = std::min(hi[2], z1);
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        double x = S.h[0] * i - S.range[0] - im.X[0];
                        double y = S.h[1] * j - S.range[1] - im.X[1];
                        double z = S.h[2] * k - S.range[2] - im.X[2];
                        double r = std::sqrt(x * x + y * y + z * z), v = 0.0;
                        if (r <= P.r[P.size - 1] && !P.s_uu.eval(r, &v)) { bad = 3; return; }
                        rho[((size_t)k * G[1] + j) * G[0] + i] += v;
                    }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++) pool.emplace_back(worker, t);
    for (auto &t : pool) t.join();
    if (bad) return fail(3, "atomic density: spline evaluation failed");
    double tot = 0.0;
    for (size_t q = 0; q < ng; q++) tot += rho[q];
    tot *= S.dV;
    const double f = (double)S.noetot / tot;
    for (size_t q = 0; q < ng; q++) rho[q] *= f;
    scale = f;
    return {};
}

Status build_system(const std::string &dir, const std::string &name, const SetupOptions &opt, System &S)
{
    S = System();
    Status st = parse_inpt(dir + "/" + name + ".inpt", S);
    if (!st.ok()) return st;
    if (opt.nqp_override > 0) S.nqp = opt.nqp_override;
    if (!(st = parse_ion(dir + "/" + name + ".ion", S)).ok()) return st;
    S.psd.resize(S.ntype);
    for (int at = 0; at < S.ntype; at++) {
        if (!(st = parse_pot(dir + "/pseudopotential/" + S.psdFile[at], S.atomType[at], S.localPsd[at], S.psd[at])).ok())
            return st;
        S.psd[at].valence = S.noe[at];
        if (!(st = fit_pseudopotential(S, S.psd[at])).ok()) return st;
        if (!(st = pseudocharge_cutoff(S, at)).ok()) return st;
    }
    // Supercell list (Create_SupercellAtomList, lsdft_initObjs.cc:3031-3041) and
    // alpha slots (Calculate_NonlocalIndex, lsdft_initObjs.cc:875-967).
    S.scpos = S.atompos;
    S.sctype = S.atomtype;
    if (opt.phantom_atom) {
        if (S.ntype != 1) return fail(4, "the phantom atom of the reference exists for NTYPE 1 only");
        for (int d = 0; d < 3; d++) S.scpos.push_back(opt.phantom_pos[d]);
        S.sctype.push_back(0);
    }
    S.ip_index.assign(S.sctype.size(), 0);
    S.ip_length = 0;
    for (int at = 0; at < S.ntype; at++)
        for (size_t p = 0; p < S.sctype.size(); p++)
            if (S.sctype[p] == at) { S.ip_index[p] = S.ip_length; S.ip_length += S.psd[at].nlm(); }
    // rank box and its extension by the cube half-width
    if (opt.rank_box[3] > 0) {
        for (int d = 0; d < 3; d++) {
            if (opt.rank_box[d] < 0 || opt.rank_box[3 + d] < 1 || opt.rank_box[d] + opt.rank_box[3 + d] > S.npts[d])
                return fail(4, "rank box outside the grid");
            S.box[d] = opt.rank_box[d];
            S.box[3 + d] = opt.rank_box[3 + d];
        }
    } else {
        for (int d = 0; d < 3; d++) { S.box[d] = 0; S.box[3 + d] = S.npts[d]; }
    }
    for (int d = 0; d < 3; d++) {
        S.ext0[d] = S.box[d] - S.hw[d];
        S.extn[d] = S.box[3 + d] + 2 * S.hw[d];
    }
    collect_images(S);
    layout_projectors(S);
    if (opt.host_tables && !(st = fill_projectors(S)).ok()) return st;
    build_coverage(S);
    return {};
}

}  // namespace sqamd
