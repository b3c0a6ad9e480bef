// sq_setup.hip -- projector tables on the device: chi_lm = S_l(r) Y_lm and the
// radial factors T_l of every image box point (CalculateNonlocalProjectors,
// lsdft_nl.cc:141-208, flattened as NonlocalProjectors, lsdft_nl.cc:81-123),
// one workgroup per image.  Every value is formed with the operations and
// order of the host fill (sq_setup.cpp: RadialSpline::eval,
// real_spherical_harmonic, fill_projectors) -- only +, -, *, / and sqrt, all
// correctly rounded here -- so the tables are bit-identical to it.  The spline
// and harmonic functions are those of sq_projectors.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sq_projectors.h"
#include "sq_system.h"

namespace sqamd {

namespace {

constexpr int kTabThreads = 256;

struct TabImage {
    int lo[3], n[3];
    double X[3];
    int type, nlm, nl;
    long long chi_off, rad_off;
};

__global__ void __launch_bounds__(kTabThreads) projector_tables(const TabImage *__restrict__ imgs,
                                                                const TabType *__restrict__ types,
                                                                const double *__restrict__ tab, double h0, double h1,
                                                                double h2, double R0, double R1, double R2,
                                                                double *__restrict__ chi, double *__restrict__ rad,
                                                                int *__restrict__ status)
{
    const TabImage im = imgs[blockIdx.x];
    const TabType &T = types[im.type];
    const double *rx = tab + T.r_off;
    const long np = (long)im.n[0] * im.n[1] * im.n[2];
    for (long p = threadIdx.x; p < np; p += blockDim.x) {
        const int i = (int)(p % im.n[0]);
        const int j = (int)((p / im.n[0]) % im.n[1]);
        const int k = (int)(p / ((long)im.n[0] * im.n[1]));
        const double z = h2 * (im.lo[2] + k) - R2 - im.X[2], zz = z * z;
        const double y = h1 * (im.lo[1] + j) - R1 - im.X[1], yy = y * y;
        const double x = h0 * (im.lo[0] + i) - R0 - im.X[0];
        const double r = sqrt(x * x + yy + zz);
        double radial[4] = {0.0, 0.0, 0.0, 0.0};
        for (int l = 0; l <= T.lmax; l++) {
            if (l == T.lloc) continue;
            if (!spline_at(rx, tab + T.y_off[l], tab + T.yd_off[l], T.n, r, &radial[l])) { *status = 3; return; }
        }
        double *out = chi + im.chi_off + p * im.nlm;
        for (int ch = 0; ch < im.nlm; ch++) {
            bool b = false;
            const double sh = ylm(x, y, z, T.ch_l[ch], T.ch_m[ch], &b);
            if (b) { *status = 4; return; }
            out[ch] = radial[T.ch_l[ch]] * sh;
        }
        double *rout = rad + im.rad_off + p * im.nl;
        int q = 0;
        for (int l = 0; l <= T.lmax; l++) {
            if (l == T.lloc) continue;
            double t;
            if (l == 0) t = radial[0] * 0.282094791773878;
            else if (!(r >= 1e-10)) t = 0.0;
            else {
                double rl = r;
                for (int e = 1; e < l; e++) rl *= r;
                t = radial[l] / rl;
            }
            rout[q++] = t;
        }
    }
}

__global__ void to_float(const double *__restrict__ a, float *__restrict__ b, long n)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) b[q] = (float)a[q];
}

}  // namespace

hipError_t launch_to_float(const double *a, float *b, long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const long blocks = std::min<long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(to_float, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

int pack_projector_splines(const System &S, std::vector<double> &tab, std::vector<TabType> &types)
{
    tab.clear();
    types.assign(S.ntype, TabType{});
    for (int at = 0; at < S.ntype; at++) {
        const Pseudopotential &P = S.psd[at];
        TabType &T = types[at];
        T.r_off = (int)tab.size();
        T.n = P.size;
        T.lmax = P.lmax;
        T.lloc = P.lloc;
        T.nlm = P.nlm();
        if (T.nlm > 16) return 4;
        tab.insert(tab.end(), P.r.begin(), P.r.end());
        for (int l = 0; l < 4; l++) {
            T.y_off[l] = T.yd_off[l] = 0;
            if (l > P.lmax || l == P.lloc) continue;
            T.y_off[l] = (int)tab.size();
            tab.insert(tab.end(), P.s_udv[l].y.begin(), P.s_udv[l].y.end());
            T.yd_off[l] = (int)tab.size();
            tab.insert(tab.end(), P.s_udv[l].yd.begin(), P.s_udv[l].yd.end());
        }
        for (int c = 0; c < 16; c++) {
            T.ch_l[c] = c < T.nlm ? P.ch_l[c] : 0;
            T.ch_m[c] = c < T.nlm ? P.ch_m[c] : 0;
        }
    }
    return 0;
}

int build_projector_tables(const System &S, double *chi, double *rad, int *status, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (S.images.empty()) return 0;
    std::vector<double> tab;
    std::vector<TabType> types;
    if (int rc = pack_projector_splines(S, tab, types)) return rc;
    std::vector<TabImage> imgs(S.images.size());
    for (size_t g = 0; g < S.images.size(); g++) {
        const Image &src = S.images[g];
        TabImage &im = imgs[g];
        for (int d = 0; d < 3; d++) {
            im.lo[d] = src.lo[d];
            im.n[d] = src.dim(d);
        }
        im.X[0] = src.X0;
        im.X[1] = src.Y0;
        im.X[2] = src.Z0;
        im.type = src.type;
        im.nlm = src.nlm;
        im.nl = src.nl;
        im.chi_off = src.chi_off;
        im.rad_off = src.rad_off;
    }
    TabImage *dimg = nullptr;
    TabType *dtyp = nullptr;
    double *dtab = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&dimg, imgs.size() * sizeof(TabImage)) != hipSuccess ||
        hipMalloc((void **)&dtyp, types.size() * sizeof(TabType)) != hipSuccess ||
        hipMalloc((void **)&dtab, tab.size() * sizeof(double)) != hipSuccess)
        rc = 6;
    if (!rc && (hipMemcpyAsync(dimg, imgs.data(), imgs.size() * sizeof(TabImage), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dtyp, types.data(), types.size() * sizeof(TabType), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess))
        rc = 8;
    if (!rc) {
        hipLaunchKernelGGL(projector_tables, dim3((unsigned)imgs.size()), dim3(kTabThreads), 0, st, dimg, dtyp, dtab, S.h[0],
                           S.h[1], S.h[2], S.range[0], S.range[1], S.range[2], chi, rad, status);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = 8;
    }
    if (dimg) (void)hipFree(dimg);
    if (dtyp) (void)hipFree(dtyp);
    if (dtab) (void)hipFree(dtab);
    return rc;
}

}  // namespace sqamd
