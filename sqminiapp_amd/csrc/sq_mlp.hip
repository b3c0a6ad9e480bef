// sq_mlp.hip -- the initial-density network: descriptors, forward (FP32 and
// BF16-storage), deltas, gradient sums and the Adam step.
//
// No counterpart in the reference (its SCF, declared only, would start from
// the superposed atomic densities, lsdft_readfiles.cc:1026-1040).  The network
// predicts the SCF's correction to that superposition at every cell point
// from Behler-type symmetry functions of the Mg atoms within rf = 8 bohr of
// the point (the point is the centre; fc(r) = (cos(pi r/rf) + 1)/2):
//   radial   f_k = sum_J exp(-((r_J - mu_k)/s)^2) fc(r_J),  mu_k = k rf/15, s = rf/15, k < 16
//   angular  g_a = 2^(1-zeta) sum_{J<L} (1 + lam cos t_JL)^zeta exp(-eta (r_J^2 + r_L^2)) fc(r_J) fc(r_L)
//            over eta in {0.005, 0.02}, zeta in {1, 2, 4, 16}, lam = +-1 (16 functions)
//   y = w3 . silu(W2 silu(W1 xhat + b1) + b2) + b3,  xhat = (f - mean) * rstd,  silu(z) = z/(1 + e^-z)
// Parameters packed W1[H][K], b1[H], W2[H][H], b2[H], w3[H], b3 (6337 floats
// for K = 32, H = 64): they sit in LDS; one thread carries one point through
// the layers with FMA chains.  The layers are 32x64 and 64x64 -- far below an
// MFMA tile's worth of reuse per point -- and the whole forward over a cell is
// microseconds, so plain FMA is the right tool here.
// Training (FP32, full batch): deltas per point, then the weight gradients as
// sums over points of outer products, split over point chunks with double
// partial sums and a fixed-order second pass (deterministic), then Adam.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sq_kernels.h"

namespace sqamd {

namespace {

template <int K, int H>
constexpr int param_count() { return H * K + H + H * H + H + H + 1; }

// float -> bfloat16 -> float, round to nearest even (finite inputs)
__device__ __forceinline__ float to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(u);
}

constexpr int kRad = 16, kAng = 16, kMaxNb = 96;   // descriptor split; neighbours within rf of one point

__global__ void __launch_bounds__(256) mlp_features(const double4 *__restrict__ img, int nimg, int G0, int G1, int G2,
                                                    double h0, double h1, double h2, double R0, double R1, double R2,
                                                    double rf, float *__restrict__ out, int *__restrict__ overflow)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)G0 * G1 * G2) return;
    const int i = (int)(p % G0), j = (int)((p / G0) % G1), k = (int)(p / ((long)G0 * G1));
    const double x = h0 * i - R0, y = h1 * j - R1, z = h2 * k - R2;
    const double s = rf / (kRad - 1);
    double f[kRad + kAng];
#pragma unroll
    for (int q = 0; q < kRad + kAng; q++) f[q] = 0.0;
    double nx[kMaxNb], ny[kMaxNb], nz[kMaxNb], nr[kMaxNb], nf[kMaxNb];
    int nn = 0;
    for (int a = 0; a < nimg; a++) {
        const double4 X = img[a];
        const double dx = X.x - x, dy = X.y - y, dz = X.z - z;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (!(r < rf)) continue;
        const double fc = 0.5 * (cos(M_PI * r / rf) + 1.0);
#pragma unroll
        for (int q = 0; q < kRad; q++) {
            const double u = (r - q * rf / (kRad - 1)) / s;
            f[q] += exp(-(u * u)) * fc;
        }
        if (nn < kMaxNb) { nx[nn] = dx; ny[nn] = dy; nz[nn] = dz; nr[nn] = r; nf[nn] = fc; }
        nn++;
    }
    if (nn > kMaxNb) { atomicAdd(overflow, 1); nn = kMaxNb; }
    const double eta[2] = {0.005, 0.02};
    for (int a = 0; a < nn; a++)
        for (int b = a + 1; b < nn; b++) {
            const double c = (nx[a] * nx[b] + ny[a] * ny[b] + nz[a] * nz[b]) / (nr[a] * nr[b]);
            const double rr = nr[a] * nr[a] + nr[b] * nr[b];
            const double w = nf[a] * nf[b];
#pragma unroll
            for (int ie = 0; ie < 2; ie++) {
                const double e = exp(-eta[ie] * rr) * w;
#pragma unroll
                for (int il = 0; il < 2; il++) {
                    const double p1 = 1.0 + (il ? -c : c), p2 = p1 * p1, p4 = p2 * p2, p16 = (p4 * p4) * (p4 * p4);
                    // zeta = 1, 2, 4, 16 with 2^(1-zeta)
                    f[kRad + (ie * 4 + 0) * 2 + il] += p1 * e;
                    f[kRad + (ie * 4 + 1) * 2 + il] += 0.5 * p2 * e;
                    f[kRad + (ie * 4 + 2) * 2 + il] += 0.125 * p4 * e;
                    f[kRad + (ie * 4 + 3) * 2 + il] += 0x1p-15 * p16 * e;
                }
            }
        }
#pragma unroll
    for (int q = 0; q < kRad + kAng; q++) out[p * (kRad + kAng) + q] = (float)f[q];
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float sg = 1.0f / (1.0f + expf(-z));
    return sg * (1.0f + z * (1.0f - sg));
}

template <int K, int H, bool B16>
__global__ void __launch_bounds__(256) mlp_forward(const float *__restrict__ X, long N, const float *__restrict__ P,
                                                   const float *__restrict__ mean, const float *__restrict__ rstd,
                                                   float *__restrict__ H1, float *__restrict__ H2, float *__restrict__ Z1,
                                                   float *__restrict__ Z2, float *__restrict__ Y)
{
    constexpr int NP = param_count<K, H>();
    __shared__ float w[NP];
    for (int q = threadIdx.x; q < NP; q += blockDim.x) w[q] = B16 ? to_bf16(P[q]) : P[q];
    __syncthreads();
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float *W1 = w, *b1 = W1 + H * K, *W2 = b1 + H, *b2 = W2 + H * H, *w3 = b2 + H, *b3 = w3 + H;
    float xh[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        float v = X[n * K + q];
        if (mean) v = (v - mean[q]) * rstd[q];
        xh[q] = B16 ? to_bf16(v) : v;
    }
    float a1[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
        float s = b1[i];
#pragma unroll
        for (int q = 0; q < K; q++) s = fmaf(W1[i * K + q], xh[q], s);
        const float t = silu(s);
        a1[i] = B16 ? to_bf16(t) : t;
        if (H1) H1[n * H + i] = a1[i];
        if (Z1) Z1[n * H + i] = s;
    }
    float y = b3[0];
    for (int j = 0; j < H; j++) {
        float s = b2[j];
#pragma unroll
        for (int i = 0; i < H; i++) s = fmaf(W2[j * H + i], a1[i], s);
        float t = silu(s);
        if (B16) t = to_bf16(t);
        if (H2) H2[n * H + j] = t;
        if (Z2) Z2[n * H + j] = s;
        y = fmaf(w3[j], t, y);
    }
    Y[n] = y;
}

// L = (1/2N) sum (y - t)^2: d3 = (y - t)/N, d2 = d3 w3 silu'(z2), d1 = (W2^T d2) silu'(z1)
template <int K, int H>
__global__ void __launch_bounds__(256) mlp_deltas(const float *__restrict__ Z1, const float *__restrict__ Z2,
                                                  const float *__restrict__ Y, const float *__restrict__ T, long N,
                                                  const float *__restrict__ P, float *__restrict__ D1,
                                                  float *__restrict__ D2, float *__restrict__ D3)
{
    __shared__ float w2[H * H], w3[H];
    const float *W2g = P + H * K + H, *w3g = W2g + H * H + H;
    for (int q = threadIdx.x; q < H * H; q += blockDim.x) w2[q] = W2g[q];
    for (int q = threadIdx.x; q < H; q += blockDim.x) w3[q] = w3g[q];
    __syncthreads();
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float d3 = (Y[n] - T[n]) / (float)N;
    D3[n] = d3;
    float d2[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
        d2[j] = d3 * w3[j] * dsilu(Z2[n * H + j]);
        D2[n * H + j] = d2[j];
    }
    for (int i = 0; i < H; i++) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < H; j++) s = fmaf(w2[j * H + i], d2[j], s);
        D1[n * H + i] = s * dsilu(Z1[n * H + i]);
    }
}

// partial[c][i][j] = sum over the points of chunk c of A[n][i] * B[n][j] (B == nullptr: 1)
__global__ void __launch_bounds__(256) outer_partial(const float *__restrict__ A, int a, const float *__restrict__ B, int b,
                                                     long N, long chunk, double *__restrict__ partial)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a * b) return;
    const int i = e / b, j = e - i * b;
    const long n0 = blockIdx.y * chunk, n1 = min(N, n0 + chunk);
    double s = 0.0;
    for (long n = n0; n < n1; n++) s += (double)A[n * a + i] * (B ? (double)B[n * b + j] : 1.0);
    partial[(size_t)blockIdx.y * a * b + e] = s;
}

__global__ void sum_chunks(const double *__restrict__ partial, int nchunk, int m, double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += partial[(size_t)c * m + e];
    out[e] = s;
}

__global__ void adam_step(float *__restrict__ p, const double *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                          int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = (float)g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace

int mlp_param_count(int K, int H) { return H * K + H + H * H + H + H + 1; }

hipError_t launch_mlp_features(const double4 *img, int nimg, const int G[3], const double h[3], const double R[3], int K,
                               double rf, float *out, int *overflow, hipStream_t st)
{
    if (K != kMlpK || kRad + kAng != kMlpK) return hipErrorInvalidValue;
    const long n = (long)G[0] * G[1] * G[2];
    hipLaunchKernelGGL(mlp_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, img, nimg, G[0], G[1],
                       G[2], h[0], h[1], h[2], R[0], R[1], R[2], rf, out, overflow);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *X, long N, const float *P, const float *mean, const float *rstd, float *H1,
                              float *H2, float *Z1, float *Z2, float *Y, int bf16, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    const dim3 g((unsigned)((N + 255) / 256)), b(256);
    if (bf16) hipLaunchKernelGGL((mlp_forward<kMlpK, kMlpH, true>), g, b, 0, st, X, N, P, mean, rstd, H1, H2, Z1, Z2, Y);
    else hipLaunchKernelGGL((mlp_forward<kMlpK, kMlpH, false>), g, b, 0, st, X, N, P, mean, rstd, H1, H2, Z1, Z2, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_deltas(const float *Z1, const float *Z2, const float *Y, const float *T, long N, const float *P,
                             float *D1, float *D2, float *D3, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL((mlp_deltas<kMlpK, kMlpH>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, Z1, Z2, Y, T, N, P,
                       D1, D2, D3);
    return hipGetLastError();
}

hipError_t launch_outer_sum(const float *A, int a, const float *B, int b, long N, double *partial, int nchunk, double *out,
                            hipStream_t st)
{
    const long chunk = (N + nchunk - 1) / nchunk;
    const int m = a * b;
    hipLaunchKernelGGL(outer_partial, dim3((unsigned)((m + 255) / 256), (unsigned)nchunk), dim3(256), 0, st, A, a, B, b, N,
                       chunk, partial);
    hipLaunchKernelGGL(sum_chunks, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, partial, nchunk, m, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, const double *g, float *m, float *v, int n, float lr, int t, hipStream_t st)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_step, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

}  // namespace sqamd
