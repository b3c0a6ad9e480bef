// sq_device_common.h -- device helpers shared by the Lanczos kernels:
// wavefront / workgroup reductions (plain and compensated), small integer
// division, node coordinates, the projector gather of the generic kernel and
// the per-image alpha coefficients.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sq_kernels.h"

namespace sqamd {

constexpr int kWG = 1024;          // threads of the generic Lanczos workgroup
constexpr int kWaves = kWG / 64;

// Lanczos vectors are stored raw (w before scaling) with their scale factor
// s = 1/b; v_i = s * w_i is formed at every read, which is the value VecScale
// (x_i <- (1/b) x_i) stores in the reference (lsdft_scf.cc:131,147).

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Compensated pair used by the FP32 form's dot products: TwoSum for the
// additions, TwoProduct (fma) for the products (Ogita-Rump-Oishi Dot2).
struct CSum {
    float hi, lo;
};
__device__ __forceinline__ CSum two_sum(float a, float b)
{
    float s = a + b;
    float bb = s - a;
    float err = (a - (s - bb)) + (b - bb);
    return {s, err};
}
__device__ __forceinline__ CSum csum_add(CSum x, float y)
{
    CSum t = two_sum(x.hi, y);
    t.lo += x.lo;
    return t;
}
// Dot2 accumulation of a product: its rounding error (fma) joins the
// compensation term with the TwoSum error of the addition
__device__ __forceinline__ CSum csum_add_prod(CSum x, float a, float b)
{
    const float p = a * b;
    const float e = fmaf(a, b, -p);
    CSum t = two_sum(x.hi, p);
    t.lo += x.lo + e;
    return t;
}
__device__ __forceinline__ CSum csum_merge(CSum x, CSum y)
{
    CSum t = two_sum(x.hi, y.hi);
    t.lo += x.lo + y.lo;
    return t;
}
__device__ __forceinline__ CSum wave_csum(CSum v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        CSum o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
        v = csum_merge(v, o);
    }
    return v;
}

// Block-wide deterministic sum of one double per thread (fixed tree).
template <int NW = kWaves>
__device__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NW; w++) s += red[w];
        red[NW] = s;
    }
    __syncthreads();
    s = red[NW];
    __syncthreads();
    return s;
}

// Block-wide compensated sum (FP32 form); result widened to double.
template <int NW = kWaves>
__device__ double block_csum(CSum v, double *red)
{
    v = wave_csum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = (double)v.hi + (double)v.lo;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NW; w++) s += red[w];
        red[NW] = s;
    }
    __syncthreads();
    s = red[NW];
    __syncthreads();
    return s;
}

// q / d for 0 <= q < 2^22, 1 <= d < 2^10 via a float reciprocal (exact after
// the two corrections).
__device__ __forceinline__ int fdiv(int q, int d, float inv)
{
    int r = (int)((float)q * inv);
    if (r * d > q) r--;
    if ((r + 1) * d <= q) r++;
    return r;
}

// ------------------------------------------------------------------ gather
// A[li][ch] = sum over the points of image li's box inside the cube of
// v(pt) * chi_ch(pt); one wavefront per image.
template <typename T, int NLM>
__device__ void gather_images(const KParams &P, const int *img_list, int nimg, const int cube_lo[3], const T *vraw,
                              T vs, const T *chi, double *A)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nx = P.nloc[0], nxy = P.nloc[0] * P.nloc[1];
    for (int li = wave; li < nimg; li += kWaves) {
        const int g = img_list[li];
        const int4 lo4 = P.img_lo[g];
        const int4 dm4 = P.img_dim[g];
        const int ilo[3] = {lo4.x, lo4.y, lo4.z};
        const int idm[3] = {dm4.x, dm4.y, dm4.z};
        int a[3], n[3];
        for (int d = 0; d < 3; d++) {
            int s = max(ilo[d], cube_lo[d]);
            int e = min(ilo[d] + idm[d] - 1, cube_lo[d] + P.nloc[d] - 1);
            a[d] = s;
            n[d] = e - s + 1;
        }
        const int npt = n[0] * n[1] * n[2];
        const float inv0 = 1.0f / (float)n[0], inv1 = 1.0f / (float)n[1];
        const T *c = chi + P.img_off[g];
        const int nlm = dm4.w & 255;
        T acc[NLM];
#pragma unroll
        for (int ch = 0; ch < NLM; ch++) acc[ch] = T(0);
        for (int q = lane; q < npt; q += 64) {
            int jk = fdiv(q, n[0], inv0);
            int qi = q - jk * n[0];
            int qk = fdiv(jk, n[1], inv1);
            int qj = jk - qk * n[1];
            int gx = a[0] + qi, gy = a[1] + qj, gz = a[2] + qk;
            int p = (gz - cube_lo[2]) * nxy + (gy - cube_lo[1]) * nx + (gx - cube_lo[0]);
            T v = vs * vraw[p];
            int64_t b = ((int64_t)((gz - ilo[2]) * idm[1] + (gy - ilo[1])) * idm[0] + (gx - ilo[0])) * nlm;
#pragma unroll
            for (int ch = 0; ch < NLM; ch++)
                if (ch < nlm) {
                    if constexpr (sizeof(T) == 8) acc[ch] = acc[ch] + v * c[b + ch];
                    else acc[ch] = fmaf(v, c[b + ch], acc[ch]);
                }
        }
#pragma unroll
        for (int ch = 0; ch < NLM; ch++) {
            double s = wave_sum((double)acc[ch]);
            if (lane == 0 && ch < nlm) A[li * NLM + ch] = s;
        }
    }
}

// alpha of every image in the node list from the A of all images of the same
// atom (they are contiguous in the ascending list).
template <int NLM>
__device__ void images_alpha(const KParams &P, const int *img_list, int nimg, const double *A, double *alpha)
{
    for (int li = threadIdx.x; li < nimg; li += blockDim.x) {
        const int g = img_list[li];
        const int atom = P.img_atom[g];
        const int type = P.img_type[g];
        const int nlm = P.img_dim[g].w & 255;
        int s = li, e = li;
        while (s > 0 && P.img_atom[img_list[s - 1]] == atom) s--;
        while (e + 1 < nimg && P.img_atom[img_list[e + 1]] == atom) e++;
        const double *gam = P.gamma_ch + type * kMaxChannels;
        if (P.mode == 0) {  // one coefficient per atom
            double al = 0.0;
            for (int q = s; q <= e; q++)
                for (int ch = 0; ch < nlm; ch++) al += gam[ch] * A[q * NLM + ch];
            for (int ch = 0; ch < NLM; ch++) alpha[li * NLM + ch] = al;
        } else {            // per (l,m) channel
            for (int ch = 0; ch < NLM; ch++) {
                double al = 0.0;
                if (ch < nlm) {
                    for (int q = s; q <= e; q++) al += A[q * NLM + ch];
                    al = gam[ch] * al;
                }
                alpha[li * NLM + ch] = al;
            }
        }
    }
}

// u(p) = stencil + diag + nonlocal scatter for cube point p.
template <typename T, int NLM>
__device__ __forceinline__ T apply_point(const KParams &P, int p, int li, int lj, int lk, const int cube_lo[3],
                                         const T *vraw, T vs, const T *diag_ext, const T *chi, const short *local_of,
                                         const double *alpha, int64_t E, bool with_local = true, T init = T(0))
{
    const int nx = P.nloc[0], ny = P.nloc[1], nz = P.nloc[2], nxy = nx * ny, o = P.order;
    T sum = init;
    auto term = [&](T cf, int q) {
        if constexpr (sizeof(T) == 8) sum = sum + cf * (vs * vraw[q]);
        else sum = fmaf(cf, vs * vraw[q], sum);
    };
    const T *cx = reinterpret_cast<const T *>(P.cx_ptr), *cy = reinterpret_cast<const T *>(P.cy_ptr),
            *cz = reinterpret_cast<const T *>(P.cz_ptr);
    if (with_local) {
        for (int l = o; l >= 1; l--) if (lk - l >= 0) term(cz[l], p - l * nxy);
        for (int l = o; l >= 1; l--) if (lj - l >= 0) term(cy[l], p - l * nx);
        for (int l = o; l >= 1; l--) if (li - l >= 0) term(cx[l], p - l);
        term(diag_ext[E], p);
        for (int l = 1; l <= o; l++) if (li + l < nx) term(cx[l], p + l);
        for (int l = 1; l <= o; l++) if (lj + l < ny) term(cy[l], p + l * nx);
        for (int l = 1; l <= o; l++) if (lk + l < nz) term(cz[l], p + l * nxy);
    }
    // nonlocal scatter, images in ascending order, channels ascending
    const int e0 = P.cov_off[E], e1 = P.cov_off[E + 1];
    const int gx = cube_lo[0] + li, gy = cube_lo[1] + lj, gz = cube_lo[2] + lk;
    for (int e = e0; e < e1; e++) {
        const int g = P.cov_img[e];
        const int4 lo4 = P.img_lo[g];
        const int4 dm4 = P.img_dim[g];
        const int nlm = dm4.w & 255;
        const int64_t b = ((int64_t)((gz - lo4.z) * dm4.y + (gy - lo4.y)) * dm4.x + (gx - lo4.x)) * nlm;
        const T *c = chi + P.img_off[g] + b;
        const double *al = alpha + local_of[g] * NLM;
#pragma unroll
        for (int ch = 0; ch < NLM; ch++)
            if (ch < nlm) {
                if constexpr (sizeof(T) == 8) sum = sum + c[ch] * al[ch];
                else sum = fmaf(c[ch], (float)al[ch], sum);
            }
    }
    return sum;
}

// ---- force pass records
// The images of the node's list whose box holds the node itself (the only
// ones whose projectors can be nonzero there), in list order, into near[];
// their ids go to rec_img.  Returns the count (> rec_slots: the record
// overflows and the node is flagged by the caller).
__device__ __forceinline__ int record_near(const KParams &P, long r, const int *img_list, int nimg, const int node[3],
                                           int *near, int *s_n)
{
    if (threadIdx.x == 0) {
        int n = 0;
        for (int li = 0; li < nimg; li++) {
            const int g = img_list[li];
            const int4 lo = P.img_lo[g], dm = P.img_dim[g];
            if (node[0] >= lo.x && node[0] < lo.x + dm.x && node[1] >= lo.y && node[1] < lo.y + dm.y &&
                node[2] >= lo.z && node[2] < lo.z + dm.z) {
                if (n < P.rec_slots) near[n] = li;
                n++;
            }
        }
        *s_n = n;
        for (int s = 0; s < P.rec_slots; s++) P.rec_img[r * P.rec_slots + s] = s < n ? img_list[near[s]] : -1;
    }
    __syncthreads();
    return *s_n;
}

// alpha of the near images at Lanczos step `step` (q_step) into the record
// (their count read from LDS, so no register carries it through the steps).
__device__ __forceinline__ void record_step(const KParams &P, long r, int step, const int *near, const int &nnear_lds,
                                            const double *alpha, int stride)
{
    if (step >= P.nqp) return;
    const int nnear = nnear_lds;
    const int L = P.nlm_max;
    double *dst = P.rec + ((size_t)r * P.nqp + step) * P.rec_slots * L;
    for (int e = threadIdx.x; e < nnear * L; e += blockDim.x) {
        const int s = e / L, ch = e - s * L;
        dst[s * L + ch] = alpha[near[s] * stride + ch];
    }
}

// Node (K,J,I) of rank-box linear id (LIpos order).
__device__ __forceinline__ void node_coords(const KParams &P, long id, int out[3])
{
    const int lx = P.box_n[0], ly = P.box_n[1];
    out[0] = P.box0[0] + (int)(id % lx);
    out[1] = P.box0[1] + (int)((id / lx) % ly);
    out[2] = P.box0[2] + (int)(id / ((long)lx * ly));
}


// chi_lm = T_l * C_lm P_lm(x,y,z) for the channels of one image: T holds the
// radial factors of the type's nonlocal l values (T_0 carries C00), P_lm is
// the homogeneous polynomial of the real harmonic with the reference's
// coefficients (RealSphericalHarmonic, lsdft_tools.cc:156-240).
template <typename T, int NLM, int LAYOUT = 0>
__device__ __forceinline__ void chi_channels(const KParams &P, int type, int nlm, const T *t, T x, T y, T z, T out[NLM])
{
    const T xx = x * x, yy = y * y, zz = z * z;
    if constexpr (LAYOUT == 1) {
        // nonlocal l = 0 and 2 (local l = 1): channels (0,0), (2,-2..2)
        out[0] = t[0];
        out[1] = t[1] * (T(1.092548430592079) * (x * y));
        out[2] = t[1] * (T(1.092548430592079) * (y * z));
        out[3] = t[1] * (T(0.315391565252520) * (-xx - yy + T(2) * zz));
        out[4] = t[1] * (T(1.092548430592079) * (z * x));
        out[5] = t[1] * (T(0.546274215296040) * (xx - yy));
#pragma unroll
        for (int ch = 6; ch < NLM; ch++) out[ch] = T(0);
        return;
    }
#pragma unroll
    for (int ch = 0; ch < NLM; ch++) {
        out[ch] = T(0);
        if (ch >= nlm) continue;
        const int l = P.ch_l[type][ch], m = P.ch_m[type][ch];
        const T tl = t[P.ch_r[type][ch]];
        T poly;
        switch (l * 10 + m + 3) {
        case 3: poly = T(1); break;
        case 12: poly = T(0.488602511902920) * y; break;
        case 13: poly = T(0.488602511902920) * z; break;
        case 14: poly = T(0.488602511902920) * x; break;
        case 21: poly = T(1.092548430592079) * (x * y); break;
        case 22: poly = T(1.092548430592079) * (y * z); break;
        case 23: poly = T(0.315391565252520) * (-xx - yy + T(2) * zz); break;
        case 24: poly = T(1.092548430592079) * (z * x); break;
        case 25: poly = T(0.546274215296040) * (xx - yy); break;
        case 30: poly = T(0.590043589926644) * (T(3) * xx - yy) * y; break;
        case 31: poly = T(2.890611442640554) * (x * y * z); break;
        case 32: poly = T(0.457045799464466) * y * (T(4) * zz - xx - yy); break;
        case 33: poly = T(0.373176332590115) * z * (T(2) * zz - T(3) * xx - T(3) * yy); break;
        case 34: poly = T(0.457045799464466) * x * (T(4) * zz - xx - yy); break;
        case 35: poly = T(1.445305721320277) * z * (xx - yy); break;
        default: poly = T(0.590043589926644) * x * (xx - T(3) * yy); break;
        }
        out[ch] = tl * poly;
    }
}

}  // namespace sqamd
