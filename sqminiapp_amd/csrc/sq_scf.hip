// sq_scf.hip -- device kernels of the V_eff pipeline that closes the SCF loop
// around the SQ pass: LDA exchange-correlation, the periodic -(1/2) Lap_h
// stencil of the Poisson solve, deterministic dot products, the conjugate-
// gradient and Periodic-Pulay vector updates, and V_eff -> the diagonal of the
// node cubes.  All fields live on the periodic cell (the distinct grid points,
// i fastest); every reduction sums fixed per-block ranges and then the block
// partials in block order, so results do not depend on scheduling.
#include <algorithm>

#include "sq_kernels.h"

namespace sqamd {

namespace {
constexpr int kT = 256;          // threads of the elementwise and reduction kernels
constexpr double kCx = 0.7385587663820224;   // (3/4)(3/pi)^(1/3): eps_x = -kCx rho^(1/3)
// Perdew-Wang 1992 correlation of the unpolarised gas (Phys. Rev. B 45, 13244)
constexpr double kA = 0.031091, kA1 = 0.21370, kB1 = 7.5957, kB2 = 3.5876, kB3 = 1.6382, kB4 = 0.49294;

__device__ __forceinline__ double block_reduce(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w];
    return s;   // valid in thread 0
}
}  // namespace

// eps_xc and v_xc of the unpolarised LDA (Slater exchange + PW92 correlation);
// both zero where rho <= 0.
__global__ void __launch_bounds__(kT) xc_lda_pw(const double *__restrict__ rho, long n, double *__restrict__ eps,
                                                 double *__restrict__ vxc)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const double r = rho[q];
        double e = 0.0, v = 0.0;
        if (r > 0.0) {
            const double ex = -kCx * cbrt(r);
            const double rs = cbrt(3.0 / (4.0 * M_PI * r));
            const double srs = sqrt(rs);
            const double Q = 2.0 * kA * (kB1 * srs + kB2 * rs + kB3 * rs * srs + kB4 * rs * rs);
            const double dQ = 2.0 * kA * (0.5 * kB1 / srs + kB2 + 1.5 * kB3 * srs + 2.0 * kB4 * rs);
            const double L = log1p(1.0 / Q);
            const double ec = -2.0 * kA * (1.0 + kA1 * rs) * L;
            const double dec = -2.0 * kA * kA1 * L + 2.0 * kA * (1.0 + kA1 * rs) * dQ / (Q * (Q + 1.0));
            e = ex + ec;
            v = 4.0 / 3.0 * ex + ec - rs / 3.0 * dec;
        }
        eps[q] = e;
        vxc[q] = v;
    }
}

// y = -(1/2) Lap_h x on the periodic cell G0 x G1 x G2 (the reference's
// Poisson operator, Create_Laplacian, lsdft_initObjs.cc:309-388, on a periodic
// DMDA): y = (c0x+c0y+c0z) x + sum_a c_a (x(-a) + x(+a)) per axis.
__global__ void __launch_bounds__(kT) lap_periodic(const double *__restrict__ x, double *__restrict__ y, int G0, int G1,
                                                    int G2, int order, StencilCoeffs c, const int *skip)
{
    if (skip && *skip) return;
    const long n = (long)G0 * G1 * G2;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % G0);
        const int j = (int)((q / G0) % G1);
        const int k = (int)(q / ((long)G0 * G1));
        double s = (c.w[0][0] + c.w[1][0] + c.w[2][0]) * x[q];
        for (int a = 1; a <= order; a++) {
            const int im = ((i - a) % G0 + G0) % G0, ip = (i + a) % G0;
            const int jm = ((j - a) % G1 + G1) % G1, jp = (j + a) % G1;
            const int km = ((k - a) % G2 + G2) % G2, kp = (k + a) % G2;
            const long row = ((long)k * G1 + j) * G0;
            s += c.w[0][a] * (x[row + im] + x[row + ip]) + c.w[1][a] * (x[((long)k * G1 + jm) * G0 + i] + x[((long)k * G1 + jp) * G0 + i]) +
                 c.w[2][a] * (x[((long)km * G1 + j) * G0 + i] + x[((long)kp * G1 + j) * G0 + i]);
        }
        y[q] = s;
    }
}

// Block partials of sum a*b over fixed contiguous chunks (b == nullptr: sum a).
__global__ void __launch_bounds__(kT) dot_partial(const double *__restrict__ a, const double *__restrict__ b, long n,
                                                  double *__restrict__ partial, const int *skip)
{
    if (skip && *skip) return;
    __shared__ double sh[kT / 64];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long p0 = blockIdx.x * chunk, p1 = min(n, p0 + chunk);
    double s = 0.0;
    for (long q = p0 + threadIdx.x; q < p1; q += blockDim.x) s += b ? a[q] * b[q] : a[q];
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[slot] = scale * sum of the partials, in block order.
__global__ void sum_partials(const double *__restrict__ partial, int n, double scale, double *__restrict__ out,
                             const int *skip)
{
    if (skip && *skip) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += partial[k];
        out[0] = scale * s;
    }
}

// y = alpha x + gamma z + beta y; with z null the constant gamma is added;
// y is not read when beta is 0 (x may alias y)
__global__ void __launch_bounds__(kT) axpbypgz(double alpha, const double *x, double beta, double *y, double gamma,
                                               const double *__restrict__ z, long n, const int *skip)
{
    if (skip && *skip) return;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double v = alpha * x[q] + (z ? gamma * z[q] : gamma);
        if (beta != 0.0) v = v + beta * y[q];
        y[q] = v;
    }
}

// Conjugate-gradient step with the scalars on the device: s = {rr, pAp, rr_new, ...}.
// x += (rr/pAp) p ; r -= (rr/pAp) Ap
__global__ void __launch_bounds__(kT) cg_xr(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                                            const double *__restrict__ Ap, const double *__restrict__ s, long n,
                                            const int *skip)
{
    if (skip && *skip) return;
    const double alpha = s[0] / s[1];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        x[q] = fma(alpha, p[q], x[q]);
        r[q] = fma(-alpha, Ap[q], r[q]);
    }
}

// p = r + (rr_new/rr) p, then rr := rr_new
__global__ void __launch_bounds__(kT) cg_p(double *__restrict__ p, const double *__restrict__ r, double *__restrict__ s, long n,
                                           const int *skip)
{
    if (skip && *skip) return;
    const double beta = s[2] / s[0];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        p[q] = fma(beta, p[q], r[q]);
}
__global__ void cg_roll(double *s, const int *skip)
{
    if (skip && *skip) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) s[0] = s[2];
}

// Periodic Pulay update x_new = x + beta f - sum_j g_j (dX_j + beta dF_j);
// dX/dF are m rows of n values, g holds m coefficients (m = 0: simple mixing).
__global__ void __launch_bounds__(kT) pulay_update(const double *__restrict__ x, const double *__restrict__ f,
                                                   const double *__restrict__ dX, const double *__restrict__ dF,
                                                   const double *__restrict__ g, int m, double beta, long n,
                                                   double *__restrict__ xn, const int *m_dev, const int *skip)
{
    if (skip && *skip) return;
    if (m_dev) m = *m_dev;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double v = x[q] + beta * f[q];
        for (int j = 0; j < m; j++) v -= g[j] * (dX[(long)j * n + q] + beta * dF[(long)j * n + q]);
        xn[q] = v;
    }
}

// The two halves of a Pulay step for a preconditioned residual:
// xa = x - sum_j g_j dX_j (the averaged iterate), u = f - sum_j g_j dF_j (the
// averaged residual); the new iterate is xa + beta P u.
__global__ void __launch_bounds__(kT) pulay_split(const double *__restrict__ x, const double *__restrict__ f,
                                                  const double *__restrict__ dX, const double *__restrict__ dF,
                                                  const double *__restrict__ g, int m, long n, double *__restrict__ xa,
                                                  double *__restrict__ u, const int *m_dev, const int *skip)
{
    if (skip && *skip) return;
    if (m_dev) m = *m_dev;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double a = x[q], b = f[q];
        for (int j = 0; j < m; j++) {
            a -= g[j] * dX[(long)j * n + q];
            b -= g[j] * dF[(long)j * n + q];
        }
        xa[q] = a;
        u[q] = b;
    }
}

// cube diagonal c_diag + V_eff over the extended node box; V_eff is a field on
// the periodic cell G, wrapped with period G in every direction.
__global__ void __launch_bounds__(kT) veff_to_diag(const double *__restrict__ veff, double cdiag, int e0x, int e0y, int e0z,
                                                   int enx, int eny, int enz, int G0, int G1, int G2,
                                                   double *__restrict__ d64, float *__restrict__ d32)
{
    const long n = (long)enx * eny * enz;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % enx), j = (int)((e / enx) % eny), k = (int)(e / ((long)enx * eny));
        const int I = ((e0x + i) % G0 + G0) % G0, J = ((e0y + j) % G1 + G1) % G1, K = ((e0z + k) % G2 + G2) % G2;
        const double v = cdiag + veff[((long)K * G1 + J) * G0 + I];
        d64[e] = v;
        d32[e] = (float)v;
    }
}

// ---------------------------------------------------------------- launchers
static dim3 grid_for(long n) { return dim3((unsigned)std::min<long>(4096, std::max<long>(1, (n + kT - 1) / kT))); }

hipError_t launch_xc(const double *rho, long n, double *eps, double *vxc, hipStream_t st)
{
    hipLaunchKernelGGL(xc_lda_pw, grid_for(n), dim3(kT), 0, st, rho, n, eps, vxc);
    return hipGetLastError();
}

hipError_t launch_lap_periodic(const double *x, double *y, const int G[3], int order, const StencilCoeffs &c, hipStream_t st,
                               const int *skip)
{
    const long n = (long)G[0] * G[1] * G[2];
    hipLaunchKernelGGL(lap_periodic, grid_for(n), dim3(kT), 0, st, x, y, G[0], G[1], G[2], order, c, skip);
    return hipGetLastError();
}

hipError_t launch_dot(const double *a, const double *b, long n, double scale, double *partial, int nblocks, double *out,
                      hipStream_t st, const int *skip)
{
    hipLaunchKernelGGL(dot_partial, dim3(nblocks), dim3(kT), 0, st, a, b, n, partial, skip);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(64), 0, st, partial, nblocks, scale, out, skip);
    return hipGetLastError();
}

hipError_t launch_axpbypgz(double alpha, const double *x, double beta, double *y, double gamma, const double *z, long n,
                           hipStream_t st, const int *skip)
{
    hipLaunchKernelGGL(axpbypgz, grid_for(n), dim3(kT), 0, st, alpha, x, beta, y, gamma, z, n, skip);
    return hipGetLastError();
}

hipError_t launch_cg_xr(double *x, double *r, const double *p, const double *Ap, const double *s, long n, hipStream_t st,
                        const int *skip)
{
    hipLaunchKernelGGL(cg_xr, grid_for(n), dim3(kT), 0, st, x, r, p, Ap, s, n, skip);
    return hipGetLastError();
}

hipError_t launch_cg_p(double *p, const double *r, double *s, long n, hipStream_t st, const int *skip)
{
    hipLaunchKernelGGL(cg_p, grid_for(n), dim3(kT), 0, st, p, r, s, n, skip);
    hipLaunchKernelGGL(cg_roll, dim3(1), dim3(64), 0, st, s, skip);
    return hipGetLastError();
}

hipError_t launch_pulay_update(const double *x, const double *f, const double *dX, const double *dF, const double *g, int m,
                               double beta, long n, double *xn, hipStream_t st, const int *m_dev, const int *skip)
{
    hipLaunchKernelGGL(pulay_update, grid_for(n), dim3(kT), 0, st, x, f, dX, dF, g, m, beta, n, xn, m_dev, skip);
    return hipGetLastError();
}

hipError_t launch_pulay_split(const double *x, const double *f, const double *dX, const double *dF, const double *g, int m,
                              long n, double *xa, double *u, hipStream_t st, const int *m_dev, const int *skip)
{
    hipLaunchKernelGGL(pulay_split, grid_for(n), dim3(kT), 0, st, x, f, dX, dF, g, m, n, xa, u, m_dev, skip);
    return hipGetLastError();
}

hipError_t launch_veff_to_diag(const double *veff, double cdiag, const int ext0[3], const int extn[3], const int G[3],
                               double *d64, float *d32, hipStream_t st)
{
    const long n = (long)extn[0] * extn[1] * extn[2];
    hipLaunchKernelGGL(veff_to_diag, grid_for(n), dim3(kT), 0, st, veff, cdiag, ext0[0], ext0[1], ext0[2], extn[0], extn[1],
                       extn[2], G[0], G[1], G[2], d64, d32);
    return hipGetLastError();
}

}  // namespace sqamd
