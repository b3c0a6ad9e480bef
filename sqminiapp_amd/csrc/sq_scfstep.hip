// sq_scfstep.hip -- the decisions of one SCF step, taken on the device.
//
// The step (V_eff pipeline, SQ pass, Fermi level, energies, Pulay mixing) is a
// fixed sequence of launches on fixed buffers: every test that decides what
// comes next -- a conjugate-gradient solve reaching its tolerance, the Fermi
// bisection bracket and stop, the Pulay history slot and its small
// least-squares solve, the stop of the SCF itself -- is a one-thread kernel
// writing the control block ScfCtl, and the launches that a test makes
// unnecessary read a skip flag from it and return.  So the same sequence runs
// launch by launch or replayed from one captured graph, and both do exactly
// the arithmetic of the host-driven loop it replaces.
#include <cmath>

#include "sq_kernels.h"

namespace sqamd {

namespace {
constexpr int kT = 256;

__global__ void k_cg_begin(const double *s, double *stop, int *done, int *iters, double scale, int sidx,
                           const int *skip)
{
    if (threadIdx.x || blockIdx.x) return;
    if (skip && *skip) { *done = 1; *iters = 0; return; }
    *stop = scale * s[sidx];
    *done = !(s[0] > *stop);
    *iters = 0;
}

__global__ void k_cg_check(const double *s, const double *stop, int *done, int *iters, int maxit)
{
    if (threadIdx.x || blockIdx.x || *done) return;
    *iters += 4;
    if (!(s[0] > *stop) || *iters >= maxit) *done = 1;
}

__global__ void k_cg_end(const double *s, const double *stop, int *fail, const int *skip)
{
    if (threadIdx.x || blockIdx.x) return;
    *fail = (skip && *skip) ? 0 : !(s[0] <= *stop);
}

// y = alpha x + mscale * (sum / n), the mean of a field taken from a dot slot
__global__ void __launch_bounds__(kT) k_axpb_mean(double alpha, const double *x, double *y, double mscale, const double *sum,
                                                  long n, const int *skip)
{
    if (skip && *skip) return;
    const double gamma = mscale * (sum[0] / (double)n);
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        y[q] = alpha * x[q] + gamma;
}

__global__ void __launch_bounds__(kT) k_copy(double *dst, const double *src, long n, const int *skip)
{
    if (skip && *skip) return;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) dst[q] = src[q];
}

__global__ void __launch_bounds__(kT) k_zero(double *y, long n, const int *skip)
{
    if (skip && *skip) return;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) y[q] = 0.0;
}

__global__ void k_zero_int(int *y, int n)
{
    for (int q = threadIdx.x; q < n; q += blockDim.x) y[q] = 0;
}

// ---- Fermi level: bisection on [lambda_min - 1, lambda_max + 1] to 1e-14
// (relative), at most 200 occupation counts (the host loop it replaces)
__global__ void k_bis_begin(const double *range, ScfCtl *c)
{
    if (threadIdx.x || blockIdx.x) return;
    c->lo = range[0] - 1.0;
    c->hi = range[1] + 1.0;
    c->fdone = 0;
    c->fbad = 0;
    c->fcount = 0;
}

// N(lf) = 2 sum of the occupation partials, in block order
__global__ void k_bis_store(const double *partial, int nblocks, double *out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int k = 0; k < nblocks; k++) s += partial[k];
    *out = 2.0 * s;
}

__device__ __forceinline__ void bis_mid(ScfCtl *c)
{
    c->mid = 0.5 * (c->lo + c->hi);
    if (c->hi - c->lo <= 1e-14 * (fabs(c->mid) > 1.0 ? fabs(c->mid) : 1.0)) c->fdone = 1;
}

__global__ void k_bis_first(ScfCtl *c, double nelec)
{
    if (threadIdx.x || blockIdx.x) return;
    if (c->glo - nelec > 0 || c->ghi - nelec < 0) {
        c->fbad = 1;
        c->fdone = 1;
        return;
    }
    bis_mid(c);
}

__global__ void k_bis_step(const double *partial, int nblocks, ScfCtl *c, double nelec)
{
    if (threadIdx.x || blockIdx.x || c->fdone) return;
    double s = 0.0;
    for (int k = 0; k < nblocks; k++) s += partial[k];
    const double g = 2.0 * s;
    if (g - nelec > 0) c->hi = c->mid;
    else c->lo = c->mid;
    if (++c->fcount >= 200) { c->fdone = 1; return; }
    bis_mid(c);
}

__global__ void k_bis_end(ScfCtl *c)
{
    if (threadIdx.x || blockIdx.x) return;
    c->lf = 0.5 * (c->lo + c->hi);
}

// first node (lowest index) whose status is not 0, or -1
__global__ void k_status_any(const int *status, long n, ScfCtl *c)
{
    __shared__ long sm[kT];
    long first = n;
    for (long p = threadIdx.x; p < n; p += blockDim.x)
        if (status[p] != 0) { first = p; break; }
    sm[threadIdx.x] = first;
    __syncthreads();
    if (threadIdx.x == 0) {
        long m = n;
        for (int k = 0; k < (int)blockDim.x; k++) m = sm[k] < m ? sm[k] : m;
        if (c->bad < 0 && m < n) c->bad = (int)m;
    }
}

// ---- SCF stop and Pulay history
__global__ void k_conv(const double *s, ScfCtl *c, double tol)
{
    if (threadIdx.x || blockIdx.x) return;
    const double res = sqrt(s[10] / s[11]);
    c->conv = res < tol;
    c->nopush = c->conv || c->it <= 1;
}

// dX[head] = x - x_prev, dF[head] = f - f_prev (as the host loop's axpy pair)
__global__ void __launch_bounds__(kT) k_hist_push(const double *x, const double *xp, const double *f, const double *fp,
                                                  double *dX, double *dF, long n, const ScfCtl *c)
{
    if (c->nopush) return;
    double *dx = dX + (size_t)c->head * n, *df = dF + (size_t)c->head * n;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        dx[q] = 1.0 * x[q] + -1.0 * xp[q];
        df[q] = 1.0 * f[q] + -1.0 * fp[q];
    }
}

__global__ void k_hist_rotate(ScfCtl *c, int m, int pf)
{
    if (threadIdx.x || blockIdx.x) return;
    if (!c->nopush) {
        c->head = (c->head + 1) % m;
        c->nhist = c->nhist + 1 < m ? c->nhist + 1 : m;
    }
    c->pulay = !c->conv && c->it % pf == 0 && c->nhist > 0;
    for (int a = 0; a < kMaxHist; a++) c->gskip[a] = !(c->pulay && a < c->nhist);
}

// Gaussian elimination with partial pivoting on the nhist x nhist Gram system
// (the host's small_solve, operation for operation); mu = nhist on success
__global__ void k_pulay_solve(ScfCtl *c, int m)
{
    if (threadIdx.x || blockIdx.x) return;
    c->mu = 0;
    if (!c->pulay) return;
    const int nh = c->nhist;
    double A[kMaxHist * kMaxHist], x[kMaxHist];
    for (int a = 0; a < nh; a++) {
        for (int b = 0; b <= a; b++) A[a * nh + b] = A[b * nh + a] = c->gram[a * m + b];
        x[a] = c->rhs[a];
    }
    for (int col = 0; col < nh; col++) {
        int piv = col;
        for (int r = col + 1; r < nh; r++)
            if (fabs(A[r * nh + col]) > fabs(A[piv * nh + col])) piv = r;
        if (A[piv * nh + col] == 0.0) return;
        if (piv != col) {
            for (int k = 0; k < nh; k++) { const double t = A[col * nh + k]; A[col * nh + k] = A[piv * nh + k]; A[piv * nh + k] = t; }
            const double t = x[col]; x[col] = x[piv]; x[piv] = t;
        }
        for (int r = col + 1; r < nh; r++) {
            const double f = A[r * nh + col] / A[col * nh + col];
            for (int k = col; k < nh; k++) A[r * nh + k] -= f * A[col * nh + k];
            x[r] -= f * x[col];
        }
    }
    for (int r = nh - 1; r >= 0; r--) {
        double v = x[r];
        for (int k = r + 1; k < nh; k++) v -= A[r * nh + k] * x[k];
        x[r] = v / A[r * nh + r];
    }
    for (int a = 0; a < nh; a++) c->g[a] = x[a];
    c->mu = nh;
}

__global__ void k_step_begin(ScfCtl *c)
{
    if (threadIdx.x || blockIdx.x) return;
    c->bad = -1;
    c->pfail = c->kfail = c->fbad = 0;
    c->conv = 0;
}

__global__ void k_it_next(ScfCtl *c)
{
    if (threadIdx.x || blockIdx.x) return;
    c->it++;
}

dim3 grid_for(long n) { return dim3((unsigned)std::min<long>(4096, std::max<long>(1, (n + kT - 1) / kT))); }
}  // namespace

hipError_t launch_cg_begin(const double *s, double *stop, int *done, int *iters, double scale, int sidx, const int *skip,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(64), 0, st, s, stop, done, iters, scale, sidx, skip);
    return hipGetLastError();
}
hipError_t launch_cg_check(const double *s, const double *stop, int *done, int *iters, int maxit, hipStream_t st)
{
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(64), 0, st, s, stop, done, iters, maxit);
    return hipGetLastError();
}
hipError_t launch_cg_end(const double *s, const double *stop, int *fail, const int *skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_cg_end, dim3(1), dim3(64), 0, st, s, stop, fail, skip);
    return hipGetLastError();
}
hipError_t launch_axpb_mean(double alpha, const double *x, double *y, double mscale, const double *sum, long n,
                            const int *skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpb_mean, grid_for(n), dim3(kT), 0, st, alpha, x, y, mscale, sum, n, skip);
    return hipGetLastError();
}
hipError_t launch_copy(double *dst, const double *src, long n, const int *skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, grid_for(n), dim3(kT), 0, st, dst, src, n, skip);
    return hipGetLastError();
}
hipError_t launch_zero(double *y, long n, const int *skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(kT), 0, st, y, n, skip);
    return hipGetLastError();
}
hipError_t launch_zero_int(int *y, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, y, n);
    return hipGetLastError();
}
hipError_t launch_bis_begin(const double *range, ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_bis_begin, dim3(1), dim3(64), 0, st, range, c);
    return hipGetLastError();
}
hipError_t launch_bis_store(const double *partial, int nblocks, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_bis_store, dim3(1), dim3(64), 0, st, partial, nblocks, out);
    return hipGetLastError();
}
hipError_t launch_bis_first(ScfCtl *c, double nelec, hipStream_t st)
{
    hipLaunchKernelGGL(k_bis_first, dim3(1), dim3(64), 0, st, c, nelec);
    return hipGetLastError();
}
hipError_t launch_bis_step(const double *partial, int nblocks, ScfCtl *c, double nelec, hipStream_t st)
{
    hipLaunchKernelGGL(k_bis_step, dim3(1), dim3(64), 0, st, partial, nblocks, c, nelec);
    return hipGetLastError();
}
hipError_t launch_bis_end(ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_bis_end, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}
hipError_t launch_status_any(const int *status, long n, ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_status_any, dim3(1), dim3(kT), 0, st, status, n, c);
    return hipGetLastError();
}
hipError_t launch_conv(const double *s, ScfCtl *c, double tol, hipStream_t st)
{
    hipLaunchKernelGGL(k_conv, dim3(1), dim3(64), 0, st, s, c, tol);
    return hipGetLastError();
}
hipError_t launch_hist_push(const double *x, const double *xp, const double *f, const double *fp, double *dX, double *dF,
                            long n, const ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_hist_push, grid_for(n), dim3(kT), 0, st, x, xp, f, fp, dX, dF, n, c);
    return hipGetLastError();
}
hipError_t launch_hist_rotate(ScfCtl *c, int m, int pf, hipStream_t st)
{
    hipLaunchKernelGGL(k_hist_rotate, dim3(1), dim3(64), 0, st, c, m, pf);
    return hipGetLastError();
}
hipError_t launch_pulay_solve(ScfCtl *c, int m, hipStream_t st)
{
    hipLaunchKernelGGL(k_pulay_solve, dim3(1), dim3(64), 0, st, c, m);
    return hipGetLastError();
}
hipError_t launch_step_begin(ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}
hipError_t launch_it_next(ScfCtl *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_it_next, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

}  // namespace sqamd
