// sq_kernels.hip -- gfx950 kernels of the SQ hot path.
//
// One workgroup runs the whole Lanczos recurrence of one grid node's cube
// (persistent over the N_qp+1 H applications) and then takes the next node;
// the Lanczos vectors live in a per-workgroup slice of HBM that stays resident
// in the L2/Infinity cache while the node is processed.  Per H application:
//   gather  : A[img][ch] = sum_{cube & box} v * chi   (one wavefront per image,
//             DPP/shuffle reduction, no atomics -> deterministic)
//   alpha   : per atom, from the A of its images (collapsed or KB slots)
//   apply   : u = sum over the 6*order+1 stencil columns in ascending column
//             order, + diag(L)+V_eff, then + chi*alpha over the images covering
//             the point in ascending image order (the reference's order)
//   reduce  : a = <v,u>, then w = u - a v - b vprev, b = ||w||
// This file is compiled with -ffp-contract=off: in the FP64 form every
// product and sum rounds as written.  The FP32 form uses explicit fmaf.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>
#include <mutex>

#include "sq_device_common.h"

namespace sqamd {

template <typename T, int NLM>
__global__ void __launch_bounds__(kWG) lanczos_kernel(KParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int *img_list = reinterpret_cast<int *>(smem);
    double *A = reinterpret_cast<double *>(smem + P.lds_A);
    double *alpha = reinterpret_cast<double *>(smem + P.lds_alpha);
    double *red = reinterpret_cast<double *>(smem + P.lds_red);
    short *local_of = reinterpret_cast<short *>(smem + P.lds_local);
    __shared__ int s_count;

    const int nx = P.nloc[0], ny = P.nloc[1];
    const int nc = P.nc;
    const float invx = 1.0f / (float)nx, invy = 1.0f / (float)ny;
    // three vectors per workgroup: v, vprev (raw, with scales) and u = H v
    T *buf0 = reinterpret_cast<T *>(P.scratch) + (size_t)blockIdx.x * 3 * nc;
    T *buf1 = buf0 + nc;
    T *Bu = buf1 + nc;
    const T *diag_ext = reinterpret_cast<const T *>(P.diag_ext);
    const T *chi = reinterpret_cast<const T *>(P.chi);

    __shared__ int s_near[kRecMax], s_nnear;
    for (long r = blockIdx.x; r < P.node_count; r += gridDim.x) {
        const long id = P.node_list ? (long)P.node_list[r] : P.node_begin + r;
        int node[3];
        node_coords(P, id, node);
        int cube_lo[3] = {node[0] - P.hw[0], node[1] - P.hw[1], node[2] - P.hw[2]};
        // images whose box meets this cube, ascending (ordered compaction)
        if (threadIdx.x == 0) s_count = 0;
        __syncthreads();
        for (int base = 0; base < P.n_images; base += kWG) {
            const int g = base + threadIdx.x;
            bool hit = false;
            if (g < P.n_images) {
                const int4 lo4 = P.img_lo[g], dm4 = P.img_dim[g];
                hit = lo4.x <= cube_lo[0] + P.nloc[0] - 1 && lo4.x + dm4.x - 1 >= cube_lo[0] &&
                      lo4.y <= cube_lo[1] + P.nloc[1] - 1 && lo4.y + dm4.y - 1 >= cube_lo[1] &&
                      lo4.z <= cube_lo[2] + P.nloc[2] - 1 && lo4.z + dm4.z - 1 >= cube_lo[2];
            }
            const unsigned long long bal = __ballot(hit);
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const int before = __popcll(bal & ((1ull << lane) - 1ull));
            int *wcnt = reinterpret_cast<int *>(red);
            if (lane == 0) wcnt[wave] = __popcll(bal);
            __syncthreads();
            int off = s_count;
            for (int w = 0; w < wave; w++) off += wcnt[w];
            if (hit) {
                const int li = off + before;
                if (li < P.img_cap) { img_list[li] = g; local_of[g] = (short)li; }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int t = s_count;
                for (int w = 0; w < kWaves; w++) t += wcnt[w];
                s_count = t;
            }
            __syncthreads();
        }
        const int nimg = s_count;
        if (nimg > P.img_cap) {  // the host sized img_cap from the geometry; flag the node
            if (threadIdx.x == 0) P.status[id] = 4;
            __syncthreads();
            continue;
        }
        if (P.rec) {
            if (record_near(P, r, img_list, nimg, node, s_near, &s_nnear) > P.rec_slots) {
                if (threadIdx.x == 0) P.status[id] = 4;
                __syncthreads();
                continue;
            }
        }
        // Lanczos start: v = e_center, vprev = 0
        for (int p = threadIdx.x; p < nc; p += kWG) { buf0[p] = T(0); buf1[p] = T(0); }
        __syncthreads();
        if (threadIdx.x == 0) buf0[P.center] = T(1);
        __syncthreads();
        T *Bv = buf0, *Bvp = buf1;
        T sv = T(1), svp = T(0);
        double bprev = 0.0;
        double *out_a = P.out_a + id * (P.nqp + 1), *out_b = P.out_b + id * (P.nqp + 1);
        int status = 0;
        for (int step = 0; step <= P.nqp; step++) {
            gather_images<T, NLM>(P, img_list, nimg, cube_lo, Bv, sv, chi, A);
            __syncthreads();
            images_alpha<NLM>(P, img_list, nimg, A, alpha);
            __syncthreads();
            if (P.rec) record_step(P, r, step, s_near, s_nnear, alpha, NLM);
            double pa = 0.0;
            CSum ca{0.f, 0.f};
            for (int p = threadIdx.x; p < nc; p += kWG) {
                const int jk = fdiv(p, nx, invx);
                const int li = p - jk * nx;
                const int lk = fdiv(jk, ny, invy);
                const int lj = jk - lk * ny;
                const int64_t E = ((int64_t)(cube_lo[2] + lk - P.ext0[2]) * P.extn[1] + (cube_lo[1] + lj - P.ext0[1])) *
                                      P.extn[0] + (cube_lo[0] + li - P.ext0[0]);
                const T u = apply_point<T, NLM>(P, p, li, lj, lk, cube_lo, Bv, sv, diag_ext, chi, local_of, alpha, E);
                Bu[p] = u;
                const T vc = sv * Bv[p];
                if constexpr (sizeof(T) == 8) pa += vc * u;
                else ca = csum_add_prod(ca, vc, u);
            }
            double a;
            if constexpr (sizeof(T) == 8) a = block_sum(pa, red);
            else a = block_csum(ca, red);
            // w = (u - a v) - bprev vprev, written over vprev's storage
            const T ta = T(-a), tb = T(-bprev);
            double pb = 0.0;
            CSum cb{0.f, 0.f};
            for (int p = threadIdx.x; p < nc; p += kWG) {
                const T vc = sv * Bv[p];
                const T vpc = svp * Bvp[p];
                const T u = Bu[p];
                T w;
                if constexpr (sizeof(T) == 8) {
                    w = u + ta * vc;
                    w = w + tb * vpc;
                    pb += w * w;
                } else {
                    w = fmaf(ta, vc, u);
                    w = fmaf(tb, vpc, w);
                    cb = csum_add_prod(cb, w, w);
                }
                Bvp[p] = w;
            }
            double bb;
            if constexpr (sizeof(T) == 8) bb = block_sum(pb, red);
            else bb = block_csum(cb, red);
            const double b = sqrt(bb);
            if (threadIdx.x == 0) { out_a[step] = a; out_b[step] = b; }
            if (b == 0.0) { status = 7; break; }
            // rotate: v <- w / b, vprev <- v
            T *t = Bv; Bv = Bvp; Bvp = t;
            svp = sv;
            sv = T(1.0 / b);
            bprev = b;
        }
        if (threadIdx.x == 0) P.status[id] = status;
        __syncthreads();
    }
}

// One H application on one node (device vectors x, y of nc entries):
// y = H x, or with nonlocal_only y += V_nl x, each point's projector terms
// added onto its incoming y in ascending image order as Mult_NonlocalVector
// accumulates into V2 (lsdft_matvec.cc:110-142).
template <typename T, int NLM>
__global__ void __launch_bounds__(kWG) apply_kernel(KParams P, long id, const T *x, T *y, int nonlocal_only)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int *img_list = reinterpret_cast<int *>(smem);
    double *A = reinterpret_cast<double *>(smem + P.lds_A);
    double *alpha = reinterpret_cast<double *>(smem + P.lds_alpha);
    short *local_of = reinterpret_cast<short *>(smem + P.lds_local);
    const int nx = P.nloc[0], ny = P.nloc[1];
    const float invx = 1.0f / (float)nx, invy = 1.0f / (float)ny;
    int node[3];
    node_coords(P, id, node);
    int cube_lo[3] = {node[0] - P.hw[0], node[1] - P.hw[1], node[2] - P.hw[2]};
    if (threadIdx.x == 0) {
        int n = 0;
        for (int g = 0; g < P.n_images; g++) {
            const int4 lo4 = P.img_lo[g], dm4 = P.img_dim[g];
            bool hit = lo4.x <= cube_lo[0] + P.nloc[0] - 1 && lo4.x + dm4.x - 1 >= cube_lo[0] &&
                       lo4.y <= cube_lo[1] + P.nloc[1] - 1 && lo4.y + dm4.y - 1 >= cube_lo[1] &&
                       lo4.z <= cube_lo[2] + P.nloc[2] - 1 && lo4.z + dm4.z - 1 >= cube_lo[2];
            if (hit && n < P.img_cap) { img_list[n] = g; local_of[g] = (short)n; }
            n += hit;
        }
        reinterpret_cast<int *>(smem + P.lds_red)[0] = n;
    }
    __syncthreads();
    const int nimg = reinterpret_cast<int *>(smem + P.lds_red)[0];
    if (nimg > P.img_cap) { if (threadIdx.x == 0) P.status[0] = 4; return; }
    __syncthreads();
    gather_images<T, NLM>(P, img_list, nimg, cube_lo, x, T(1), reinterpret_cast<const T *>(P.chi), A);
    __syncthreads();
    images_alpha<NLM>(P, img_list, nimg, A, alpha);
    __syncthreads();
    for (int p = threadIdx.x; p < P.nc; p += blockDim.x) {
        const int jk = fdiv(p, nx, invx);
        const int li = p - jk * nx;
        const int lk = fdiv(jk, ny, invy);
        const int lj = jk - lk * ny;
        const int64_t E = ((int64_t)(cube_lo[2] + lk - P.ext0[2]) * P.extn[1] + (cube_lo[1] + lj - P.ext0[1])) * P.extn[0] +
                          (cube_lo[0] + li - P.ext0[0]);
        y[p] = apply_point<T, NLM>(P, p, li, lj, lk, cube_lo, x, T(1), reinterpret_cast<const T *>(P.diag_ext),
                                   reinterpret_cast<const T *>(P.chi), local_of, alpha, E, !nonlocal_only,
                                   nonlocal_only ? y[p] : T(0));
    }
}

// Implicit QL of one tridiagonal per thread (TridiagEigenVecSolve_NodesAndWeights,
// lsdft_tools.cc:912-1026).  Only row 0 of the eigenvector matrix is carried:
// the reference's rotations update each row of z independently, so row 0 -- the
// one the weights w_i = z[0][i]^2 use -- is the same number sequence.
__global__ void ql_kernel(const double *a, const double *bsub, int nmax, int stride, long count, double *lam, double *w,
                          double *lmin, double *lmax, int *status, double *ws, const int *nsteps)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    // a node that stopped early has a Gauss rule of nsteps[t] nodes; the rest of
    // its row is padded with weight zero at its largest node
    const int n = nsteps ? max(1, min(nmax, nsteps[t])) : nmax;
    // d, e and row 0 of z live in a workspace laid out [i][thread] so that the
    // threads of a wavefront touch consecutive addresses.
    double *d = ws + t, *e = ws + (long)nmax * count + t, *z = ws + 2L * nmax * count + t;
#define D(i) d[(long)(i) * count]
#define EE(i) e[(long)(i) * count]
#define Z(i) z[(long)(i) * count]
    for (int i = 0; i < n; i++) {
        D(i) = a[t * stride + i];
        EE(i) = bsub[t * stride + i];
        Z(i) = (i == 0) ? 1.0 : 0.0;
    }
    EE(n - 1) = 0.0;
    int m, l, iter, i;
    double s, r, p, g, f, dd, c, b;
    for (l = 0; l <= n - 1; l++) {
        iter = 0;
        do {
            for (m = l; m <= n - 2; m++) {
                dd = fabs(D(m)) + fabs(D(m + 1));
                if ((double)(fabs(EE(m)) + dd) == dd) break;
            }
            if (m != l) {
                if (iter++ == 50) { status[t] = 5; return; }
                g = (D(l + 1) - D(l)) / (2.0 * EE(l));
                r = sqrt(g * g + 1.0);
                g = D(m) - D(l) + EE(l) / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                s = c = 1.0;
                p = 0.0;
                for (i = m - 1; i >= l; i--) {
                    f = s * EE(i);
                    b = c * EE(i);
                    EE(i + 1) = (r = sqrt(g * g + f * f));
                    if (r == 0.0) {
                        D(i + 1) -= p;
                        EE(m) = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = D(i + 1) - p;
                    r = (D(i) - g) * s + 2.0 * c * b;
                    D(i + 1) = g + (p = s * r);
                    g = c * r - b;
                    f = Z(i + 1);
                    Z(i + 1) = s * Z(i) + c * f;
                    Z(i) = c * Z(i) - s * f;
                }
                if (r == 0.0 && i >= l) continue;
                D(l) -= p;
                EE(l) = g;
                EE(m) = 0.0;
            }
        } while (m != l);
    }
    double emin = D(0), emax = D(0);
    for (i = 0; i < n; i++) {
        const double di = D(i), zi = Z(i);
        lam[t * nmax + i] = di;
        w[t * nmax + i] = zi * zi;
        if (di > emax) emax = di;
        else if (di < emin) emin = di;
    }
    for (i = n; i < nmax; i++) {
        lam[t * nmax + i] = emax;
        w[t * nmax + i] = 0.0;
    }
#undef D
#undef EE
#undef Z
    lmin[t] = emin;
    lmax[t] = emax;
    status[t] = 0;
}

// ------------------------------------------------ density / Fermi / energy
__device__ __forceinline__ double fermi_occ(double x)  // 1/(1+e^x)
{
    if (x > 0) { double t = exp(-x); return t / (1.0 + t); }
    return 1.0 / (1.0 + exp(x));
}
__device__ __forceinline__ double fermi_ent(double x)  // f ln f + (1-f) ln(1-f)
{
    double ax = fabs(x), t = exp(-ax);
    return -log1p(t) - ax * t / (1.0 + t);
}

constexpr int kRedThreads = 256;

// partial[block] = sum over the block's chunk of w * f(beta (lam - lf))
__global__ void __launch_bounds__(kRedThreads) occupation_partial(const double *lam, const double *w, long nq, double beta,
                                                                  double lf, double *partial, const double *lf_dev,
                                                                  const int *skip)
{
    if (skip && *skip) return;
    if (lf_dev) lf = *lf_dev;
    __shared__ double red[kRedThreads / 64];
    const long chunk = (nq + gridDim.x - 1) / gridDim.x;
    const long q0 = blockIdx.x * chunk, q1 = min(nq, q0 + chunk);
    double s = 0.0;
    for (long q = q0 + threadIdx.x; q < q1; q += blockDim.x) s += w[q] * fermi_occ(beta * (lam[q] - lf));
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kRedThreads / 64; k++) t += red[k];
        partial[blockIdx.x] = t;
    }
}

// sum of partials in index order -> out[0]
__global__ void ordered_sum(const double *partial, int n, double scale, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += partial[k];
        out[0] = scale * s;
    }
}

// per node: rho = (2/dV) sum_i w f; partials of sum w lam f and sum w S(f)
__global__ void __launch_bounds__(kRedThreads) density_energy(const double *lam, const double *w, long n_nodes, int nqp,
                                                              double beta, double lf, double two_over_dv, double *rho,
                                                              double *part_e, double *part_s, const double *lf_dev)
{
    if (lf_dev) lf = *lf_dev;
    __shared__ double re[kRedThreads / 64], rs[kRedThreads / 64];
    const long chunk = (n_nodes + gridDim.x - 1) / gridDim.x;
    const long p0 = blockIdx.x * chunk, p1 = min(n_nodes, p0 + chunk);
    double se = 0.0, ss = 0.0;
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        double sr = 0.0;
        for (int i = 0; i < nqp; i++) {
            const long q = p * nqp + i;
            const double x = beta * (lam[q] - lf);
            const double f = fermi_occ(x);
            sr += w[q] * f;
            se += w[q] * lam[q] * f;
            ss += w[q] * fermi_ent(x);
        }
        if (rho) rho[p] = two_over_dv * sr;
    }
    se = wave_sum(se);
    ss = wave_sum(ss);
    if ((threadIdx.x & 63) == 0) { re[threadIdx.x >> 6] = se; rs[threadIdx.x >> 6] = ss; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kRedThreads / 64; k++) { a += re[k]; b += rs[k]; }
        part_e[blockIdx.x] = a;
        part_s[blockIdx.x] = b;
    }
}

// min/max of lambda over all nodes (bracket of the Fermi level)
__global__ void lam_range(const double *lmin, const double *lmax, long n, double *out)
{
    __shared__ double smin[kRedThreads], smax[kRedThreads];
    double a = 1e300, b = -1e300;
    for (long p = threadIdx.x; p < n; p += blockDim.x) { a = fmin(a, lmin[p]); b = fmax(b, lmax[p]); }
    smin[threadIdx.x] = a;
    smax[threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)blockDim.x; k++) { a = fmin(a, smin[k]); b = fmax(b, smax[k]); }
        out[0] = a;
        out[1] = b;
    }
}

// ------------------------------------------------------------ launchers
template <typename T, int NLM>
static hipError_t launch_lanczos_nlm(const KParams &P, int grid, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL((lanczos_kernel<T, NLM>), dim3(grid), dim3(kWG), lds, st, P);
    return hipGetLastError();
}

hipError_t launch_lanczos(const KParams &P, int precision, int grid, size_t lds, hipStream_t st)
{
    if (P.nlm_max <= 6) {
        if (precision == 0) return launch_lanczos_nlm<double, 6>(P, grid, lds, st);
        return launch_lanczos_nlm<float, 6>(P, grid, lds, st);
    }
    if (P.nlm_max <= kMaxChannels) {
        if (precision == 0) return launch_lanczos_nlm<double, kMaxChannels>(P, grid, lds, st);
        return launch_lanczos_nlm<float, kMaxChannels>(P, grid, lds, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_apply(const KParams &P, long id, const double *x, double *y, int nonlocal_only, size_t lds,
                        hipStream_t st)
{
    if (P.nlm_max <= 6)
        hipLaunchKernelGGL((apply_kernel<double, 6>), dim3(1), dim3(kWG), lds, st, P, id, x, y, nonlocal_only);
    else if (P.nlm_max <= kMaxChannels)
        hipLaunchKernelGGL((apply_kernel<double, kMaxChannels>), dim3(1), dim3(kWG), lds, st, P, id, x, y,
                           nonlocal_only);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ql(const double *a, const double *b, int n, int stride, long count, double *lam, double *w,
                     double *lmin, double *lmax, int *status, double *ws, hipStream_t st, const int *nsteps)
{
    if (n < 1 || n > kMaxNqp || count <= 0) return hipErrorInvalidValue;
    const int threads = 64;
    const long blocks = (count + threads - 1) / threads;
    hipLaunchKernelGGL(ql_kernel, dim3((unsigned)blocks), dim3(threads), 0, st, a, b, n, stride, count, lam, w, lmin,
                       lmax, status, ws, nsteps);
    return hipGetLastError();
}

hipError_t launch_occupation(const double *lam, const double *w, long nq, double beta, double lf, double *partial,
                             int nblocks, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(occupation_partial, dim3(nblocks), dim3(kRedThreads), 0, st, lam, w, nq, beta, lf, partial,
                       (const double *)nullptr, (const int *)nullptr);
    hipLaunchKernelGGL(ordered_sum, dim3(1), dim3(64), 0, st, partial, nblocks, 2.0, out);
    return hipGetLastError();
}

hipError_t launch_density_energy(const double *lam, const double *w, long n_nodes, int nqp, double beta, double lf,
                                 double two_over_dv, double *rho, double *part_e, double *part_s, int nblocks,
                                 double *out2, hipStream_t st)
{
    hipLaunchKernelGGL(density_energy, dim3(nblocks), dim3(kRedThreads), 0, st, lam, w, n_nodes, nqp, beta, lf,
                       two_over_dv, rho, part_e, part_s, (const double *)nullptr);
    hipLaunchKernelGGL(ordered_sum, dim3(1), dim3(64), 0, st, part_e, nblocks, 2.0, out2);
    hipLaunchKernelGGL(ordered_sum, dim3(1), dim3(64), 0, st, part_s, nblocks, 2.0 / beta, out2 + 1);
    return hipGetLastError();
}

hipError_t launch_occupation_partial_dev(const double *lam, const double *w, long nq, double beta, const double *lf_dev,
                                         const int *skip, double *partial, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(occupation_partial, dim3(nblocks), dim3(kRedThreads), 0, st, lam, w, nq, beta, 0.0, partial, lf_dev,
                       skip);
    return hipGetLastError();
}

hipError_t launch_density_energy_dev(const double *lam, const double *w, long n_nodes, int nqp, double beta,
                                     const double *lf_dev, double two_over_dv, double *rho, double *part_e, double *part_s,
                                     int nblocks, double *out2, hipStream_t st)
{
    hipLaunchKernelGGL(density_energy, dim3(nblocks), dim3(kRedThreads), 0, st, lam, w, n_nodes, nqp, beta, 0.0,
                       two_over_dv, rho, part_e, part_s, lf_dev);
    hipLaunchKernelGGL(ordered_sum, dim3(1), dim3(64), 0, st, part_e, nblocks, 2.0, out2);
    hipLaunchKernelGGL(ordered_sum, dim3(1), dim3(64), 0, st, part_s, nblocks, 2.0 / beta, out2 + 1);
    return hipGetLastError();
}

hipError_t launch_lam_range(const double *lmin, const double *lmax, long n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(lam_range, dim3(1), dim3(kRedThreads), 0, st, lmin, lmax, n, out);
    return hipGetLastError();
}

// ------------------------------------------------- QL without allocation
// Workspace of sq_tridiag_nodes_weights_noalloc: a static device array of
// kNoAllocScratchBytes, one copy per GPU at its own address, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory: its address on each device comes from hipGetSymbolAddress
// (noalloc_scratch_init, once per device, with that device current) and only
// that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_dev[kNoAllocMaxDevices];

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_dev[device] = p;
    return hipSuccess;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}

// workspace layout: QL d, e and z row 0 [3][n][count] doubles | status[count] |
// the index of the first failed tridiagonal (long long)
size_t noalloc_ql_bytes(int n, long count)
{
    return align256((size_t)3 * n * count * sizeof(double)) + align256((size_t)count * sizeof(int)) + 256;
}

// Index of the first tridiagonal whose QL did not converge, -1 when none
// (one workgroup; each thread scans its stride in ascending order).
__global__ void __launch_bounds__(256) first_failure(const int *status, long count, long long *out)
{
    __shared__ long long m[256];
    long long best = -1;
    for (long i = threadIdx.x; i < count; i += 256)
        if (status[i]) { best = i; break; }
    m[threadIdx.x] = best < 0 ? LLONG_MAX : best;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) m[threadIdx.x] = min(m[threadIdx.x], m[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = m[0] == LLONG_MAX ? -1 : m[0];
}

hipError_t launch_ql_noalloc(void *ws, const double *a, const double *b, int n, long count, double *lam, double *w,
                             double *lmin, double *lmax, long long **first_bad, hipStream_t st)
{
    if (!ws || n < 1 || n > kMaxNqp || count <= 0 || noalloc_ql_bytes(n, count) > kNoAllocScratchBytes)
        return hipErrorInvalidValue;
    unsigned char *base = static_cast<unsigned char *>(ws);
    double *d = reinterpret_cast<double *>(base);
    int *status = reinterpret_cast<int *>(base + align256((size_t)3 * n * count * sizeof(double)));
    long long *bad = reinterpret_cast<long long *>(reinterpret_cast<unsigned char *>(status) +
                                                   align256((size_t)count * sizeof(int)));
    hipError_t e = launch_ql(a, b, n, n, count, lam, w, lmin, lmax, status, d, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(first_failure, dim3(1), dim3(256), 0, st, status, count, bad);
    *first_bad = bad;
    return hipGetLastError();
}

}  // namespace sqamd
