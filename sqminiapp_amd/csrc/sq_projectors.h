// sq_projectors.h -- the projector functions on the device: the radial
// splines U_l dV_l of each type packed into one table, SplineEvaluate
// (lsdft_tools.cc:15-68) with its derivative, RealSphericalHarmonic
// (lsdft_tools.cc:156-240, l <= 3) and its gradient.  Shared by the table
// build (sq_setup.hip) and the force contraction (sq_forces.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sq_system.h"

namespace sqamd {

struct TabType {
    int r_off, n, lmax, lloc, nlm;
    int y_off[4], yd_off[4];   // U_l dV_l values and not-a-knot slopes in the packed table
    int ch_l[16], ch_m[16];
};

// Packs the radial grids and U_l dV_l splines of every type (host).
// Returns 0, or 4 when a type has more than 16 channels.
int pack_projector_splines(const System &S, std::vector<double> &tab, std::vector<TabType> &types);

// SplineEvaluate on the interval holding r (first knot >= r, stepped back
// unless r sits on it; the last knot takes the last interval); dout, when
// given, receives the cubic's derivative.
__device__ __forceinline__ bool spline_at(const double *x, const double *y, const double *yd, int n, double r, double *out,
                                          double *dout = nullptr)
{
    if (!(r >= x[0] && r <= x[n - 1])) return false;
    int lo = 0, hi = n;   // lower_bound
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    int j = lo;
    if (r < x[j]) j--;
    if (j > n - 2) j = n - 2;
    const double p1 = x[j];
    const double dx = 1.0 / (x[j + 1] - x[j]);
    const double dy = (y[j + 1] - y[j]) * dx;
    const double A0 = y[j], A1 = yd[j];
    const double A2 = dx * (3.0 * dy - 2.0 * yd[j] - yd[j + 1]);
    const double A3 = dx * dx * (-2.0 * dy + yd[j] + yd[j + 1]);
    const double t = r - p1;
    *out = ((A3 * t + A2) * t + A1) * t + A0;
    if (dout) *dout = (3.0 * A3 * t + 2.0 * A2) * t + A1;
    return true;
}

// RealSphericalHarmonic for l <= 3, the reference's 15-digit constants and
// expressions; *bad is set for l > 3 or |m| > l.
__device__ __forceinline__ double ylm(double x, double y, double z, int l, int m, bool *bad)
{
    const double c00 = 0.282094791773878;
    const double c1 = 0.488602511902920;
    const double c2a = 1.092548430592079, c20 = 0.315391565252520, c22 = 0.546274215296040;
    const double c33 = 0.590043589926644, c32n = 2.890611442640554, c31 = 0.457045799464466;
    const double c30 = 0.373176332590115, c32p = 1.445305721320277;
    *bad = false;
    const double r = sqrt(x * x + y * y + z * z);
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

// Gradient of Y_lm = C P_lm(x,y,z) / r^l: C (grad P / r^l - l P (x,y,z) / r^(l+2));
// zero for l = 0 and for r < 1e-10 (where Y_lm, l > 0, is set to zero).
__device__ __forceinline__ void ylm_grad(double x, double y, double z, int l, int m, double g[3])
{
    const double r2 = x * x + y * y + z * z, r = sqrt(r2);
    g[0] = g[1] = g[2] = 0.0;
    if (l == 0 || !(r >= 1e-10)) return;
    const double xx = x * x, yy = y * y, zz = z * z;
    double C = 0.0, P = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
    switch (l * 10 + (m + 3)) {
    case 12: C = 0.488602511902920; P = y; d1 = 1; break;
    case 13: C = 0.488602511902920; P = z; d2 = 1; break;
    case 14: C = 0.488602511902920; P = x; d0 = 1; break;
    case 21: C = 1.092548430592079; P = x * y; d0 = y; d1 = x; break;
    case 22: C = 1.092548430592079; P = y * z; d1 = z; d2 = y; break;
    case 23: C = 0.315391565252520; P = -xx - yy + 2 * zz; d0 = -2 * x; d1 = -2 * y; d2 = 4 * z; break;
    case 24: C = 1.092548430592079; P = z * x; d0 = z; d2 = x; break;
    case 25: C = 0.546274215296040; P = xx - yy; d0 = 2 * x; d1 = -2 * y; break;
    case 30: C = 0.590043589926644; P = (3 * xx - yy) * y; d0 = 6 * x * y; d1 = 3 * xx - 3 * yy; break;
    case 31: C = 2.890611442640554; P = x * y * z; d0 = y * z; d1 = x * z; d2 = x * y; break;
    case 32: C = 0.457045799464466; P = y * (4 * zz - xx - yy); d0 = -2 * x * y; d1 = 4 * zz - xx - 3 * yy; d2 = 8 * y * z; break;
    case 33: C = 0.373176332590115; P = z * (2 * zz - 3 * xx - 3 * yy); d0 = -6 * x * z; d1 = -6 * y * z; d2 = 6 * zz - 3 * xx - 3 * yy; break;
    case 34: C = 0.457045799464466; P = x * (4 * zz - xx - yy); d0 = 4 * zz - 3 * xx - yy; d1 = -2 * x * y; d2 = 8 * x * z; break;
    case 35: C = 1.445305721320277; P = z * (xx - yy); d0 = 2 * x * z; d1 = -2 * y * z; d2 = xx - yy; break;
    case 36: C = 0.590043589926644; P = x * (xx - 3 * yy); d0 = 3 * xx - 3 * yy; d1 = -6 * x * y; break;
    default: return;
    }
    double rl = 1.0;
    for (int k = 0; k < l; k++) rl *= r;
    const double q = l * P / (rl * r2);
    g[0] = C * (d0 / rl - q * x);
    g[1] = C * (d1 / rl - q * y);
    g[2] = C * (d2 / rl - q * z);
}

}  // namespace sqamd
