// sq_forces.hip -- the nonlocal force of the SQ density matrix.
//
// The reference declares the force routines only (Calculate_nonlocalforce,
// CalculateGradientDensityMatrixColumn, CalculateNonlocalForceContribution,
// lsdft.h:189,223-226) and builds projectors on force overlaps
// (CalculateNonlocalProjectorsForces, lsdft_initObjs.cc:1044-1114).  Here the
// force is the derivative of E_nl = 2 tr(V_nl D), V_nl = sum gamma chi chi^T:
//   F_J = 4 sum_c sum_{J' of J, c in box J'} sum_lm grad chi_J'lm(c) gamma <chi_J,lm, D_c>
// with D_c = Q f(T) e_1 the node's Gauss-quadrature density-matrix column.
// The Lanczos pass records alpha_k = gamma <chi_J, q_k> for the images whose
// box holds the node (the gather computes these sums anyway), so
//   gamma <chi_J, D_c> = sum_k c_k alpha_k,   c = f(T) e_1,
// and no vector beyond the recurrence's own is formed.
//   fermi_column : c per node, implicit QL with the full eigenvector matrix
//                  (the reference's TridiagEigenVecSolve rotations), one
//                  wavefront per node, the rows of z spread over the lanes
//   nl_force     : one thread per (node, record slot): grad chi at the node
//                  from the radial spline's derivative and grad Y_lm, times
//                  sum_k c_k alpha_k
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sq_kernels.h"
#include "sq_projectors.h"

namespace sqamd {

namespace {

__device__ __forceinline__ double fermi_fn(double x)  // 1/(1+e^x)
{
    if (x > 0) { const double t = exp(-x); return t / (1.0 + t); }
    return 1.0 / (1.0 + exp(x));
}

// z is column-major per workgroup: column j (eigenvector j) at z + j * n, so the
// lanes of a rotation touch consecutive rows.  Every lane runs the scalar
// recurrence on the same values (d, e in LDS, written identically by all).
__global__ void __launch_bounds__(64) fermi_column(const double *a, const double *bsub, int stride, const int *ids,
                                                   const int *nsteps, long count, int nmax, double beta, double lf,
                                                   double *coef, double *zws, int *status)
{
    __shared__ double sd[kMaxNqp], se[kMaxNqp];
    const int lane = threadIdx.x;
    double *z = zws + (size_t)blockIdx.x * nmax * nmax;
    for (long r = blockIdx.x; r < count; r += gridDim.x) {
        const long id = ids[r];
        const int n = nsteps ? max(1, min(nmax, nsteps[id])) : nmax;
        for (int i = lane; i < n; i += 64) {
            sd[i] = a[id * stride + i];
            se[i] = bsub[id * stride + i];
        }
        for (int q = lane; q < n * n; q += 64) z[q] = 0.0;
        __syncthreads();
        for (int i = lane; i < n; i += 64) z[(size_t)i * n + i] = 1.0;
        se[n - 1] = 0.0;
        __syncthreads();
#define Z(row, col) z[(size_t)(col) * n + (row)]
        int m, l, iter, i, bad = 0;
        double s, rr, p, g, f, dd, c, b;
        for (l = 0; l <= n - 1 && !bad; l++) {
            iter = 0;
            do {
                for (m = l; m <= n - 2; m++) {
                    dd = fabs(sd[m]) + fabs(sd[m + 1]);
                    if ((double)(fabs(se[m]) + dd) == dd) break;
                }
                if (m != l) {
                    if (iter++ == 50) { bad = 1; break; }
                    g = (sd[l + 1] - sd[l]) / (2.0 * se[l]);
                    rr = sqrt(g * g + 1.0);
                    g = sd[m] - sd[l] + se[l] / (g + (g >= 0.0 ? fabs(rr) : -fabs(rr)));
                    s = c = 1.0;
                    p = 0.0;
                    for (i = m - 1; i >= l; i--) {
                        f = s * se[i];
                        b = c * se[i];
                        se[i + 1] = (rr = sqrt(g * g + f * f));
                        if (rr == 0.0) {
                            sd[i + 1] -= p;
                            se[m] = 0.0;
                            break;
                        }
                        s = f / rr;
                        c = g / rr;
                        g = sd[i + 1] - p;
                        rr = (sd[i] - g) * s + 2.0 * c * b;
                        sd[i + 1] = g + (p = s * rr);
                        g = c * rr - b;
                        for (int k = lane; k < n; k += 64) {
                            const double zt = Z(k, i + 1);
                            Z(k, i + 1) = s * Z(k, i) + c * zt;
                            Z(k, i) = c * Z(k, i) - s * zt;
                        }
                    }
                    if (rr == 0.0 && i >= l) continue;
                    sd[l] -= p;
                    se[l] = g;
                    se[m] = 0.0;
                }
            } while (m != l);
        }
        __syncthreads();
        if (bad) {
            if (lane == 0) status[id] = 5;
        } else {
            for (int k = lane; k < nmax; k += 64) {
                double acc = 0.0;
                if (k < n)
                    for (int j = 0; j < n; j++) acc += Z(k, j) * fermi_fn(beta * (sd[j] - lf)) * Z(0, j);
                coef[r * nmax + k] = acc;
            }
        }
#undef Z
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) nl_force(ForceParams F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F.count * F.slots) return;
    const long r = t / F.slots;
    const int s = (int)(t - r * F.slots);
    double *out = F.out + (size_t)t * 3;
    const int g = F.rec_img[t];
    out[0] = out[1] = out[2] = 0.0;
    if (g < 0) return;
    const long id = F.ids[r];
    const int lx = F.box_n[0], ly = F.box_n[1];
    const int I = F.box0[0] + (int)(id % lx), J = F.box0[1] + (int)((id / lx) % ly), K = F.box0[2] + (int)(id / ((long)lx * ly));
    const double4 cen = F.img_cen[g];
    const double x = F.h[0] * I - F.range[0] - cen.x;
    const double y = F.h[1] * J - F.range[1] - cen.y;
    const double z = F.h[2] * K - F.range[2] - cen.z;
    const double rad = sqrt(x * x + y * y + z * z);
    if (!(rad >= 1e-10)) return;   // the gradient of chi is set to zero at the atom
    const TabType &T = F.types[F.img_type[g]];
    const double *rx = F.tab + T.r_off;
    double S[4] = {0, 0, 0, 0}, dS[4] = {0, 0, 0, 0};
    for (int l = 0; l <= T.lmax; l++) {
        if (l == T.lloc) continue;
        if (!spline_at(rx, F.tab + T.y_off[l], F.tab + T.yd_off[l], T.n, rad, &S[l], &dS[l])) { F.status[id] = 3; return; }
    }
    const int n = F.nsteps ? max(1, min(F.nqp, F.nsteps[id])) : F.nqp;
    const double *cf = F.coef + (size_t)r * F.nqp;
    const double *rec = F.rec + (size_t)r * F.nqp * F.slots * F.nlm + (size_t)s * F.nlm;
    const double ux = x / rad, uy = y / rad, uz = z / rad;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int ch = 0; ch < T.nlm; ch++) {
        const int l = T.ch_l[ch], m = T.ch_m[ch];
        bool bad = false;
        const double Y = ylm(x, y, z, l, m, &bad);
        double gy[3];
        ylm_grad(x, y, z, l, m, gy);
        double pr = 0.0;
        for (int k = 0; k < n; k++) pr += cf[k] * rec[(size_t)k * F.slots * F.nlm + ch];
        const double w = 4.0 * pr;
        fx += w * (dS[l] * ux * Y + S[l] * gy[0]);
        fy += w * (dS[l] * uy * Y + S[l] * gy[1]);
        fz += w * (dS[l] * uz * Y + S[l] * gy[2]);
    }
    out[0] = fx;
    out[1] = fy;
    out[2] = fz;
}

}  // namespace

hipError_t launch_fermi_column(const double *a, const double *b, int stride, const int *ids, const int *nsteps, long count,
                               int nmax, double beta, double lf, double *coef, double *zws, int *status, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    if (nmax < 1 || nmax > kMaxNqp) return hipErrorInvalidValue;
    const int grid = (int)std::min<long>(count, kFermiGrid);
    hipLaunchKernelGGL(fermi_column, dim3(grid), dim3(64), 0, st, a, b, stride, ids, nsteps, count, nmax, beta, lf, coef,
                       zws, status);
    return hipGetLastError();
}

hipError_t launch_nl_force(const ForceParams &F, hipStream_t st)
{
    const long n = F.count * F.slots;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nl_force, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, F);
    return hipGetLastError();
}

}  // namespace sqamd
