// sq_zbatch.hip -- the batched Lanczos kernel: one workgroup runs the Lanczos
// recurrences of B grid nodes stacked along z, (i, j, k .. k+B-1), at once.
//
// The B cubes share their (x,y) columns and differ by one plane in z, so the
// batch lives on a union cube of nx x ny x (nz + B - 1) points.  Each thread
// owns one (x,y) column and carries the B nodes' values of every plane
// together (one 16-byte load or store moves a plane point of the whole
// batch).  What the nodes share is read once for all of them:
//  * the diagonal fl(c_diag + V_eff) of a union plane point;
//  * the projector radial factors of an image point, in the gather and the
//    scatter;
//  * the image list, the column masks, the stencil weights.
// Vector layout in the workgroup's scratch: [union plane][column][node], node
// b's local plane z sitting at union plane z + b; the planes of the union
// outside node b's cube hold exact zeros, which is the zero-Dirichlet
// truncation of its cube in z.
//
// The projector gather is a batched small GEMM: for one image and one (x,y)
// column of its box, S[m][n] = sum_z basis_m(z) v_n(z) with the four column
// basis functions t0, t2, t2 z, t2 z^2 (the s+d channels are t0 and t2 times
// a quadratic in z with column constants) and n the batch's nodes.  Two
// engines, chosen per deck by the autotuner through the gather trip: FP32
// trips of 4 planes run it on the matrix cores, v_mfma_f32_4x4x1_16b (16
// columns per instruction, one plane a step); trips of 1 or 2 planes (FP64,
// and FP32 trip 2) on the vector ALU, one column per lane.  The FP64 matrix
// instruction (v_mfma_f64_4x4x4_4b) runs at the FP64 vector rate; it was
// measured 10-35 % slower than the vector form and is not built.  The six
// channel sums of a node follow from S and the column constants, then a
// butterfly over the columns.  FP32 reads the four basis values of a point
// from a precomputed table (16 bytes); FP64 reads the point's radial pair
// (t0, t2) (16 bytes, half the FP64 table) and forms t2 z and t2 z^2 in
// registers with the table builder's own operations, so the values, and every
// result, are the same bit for bit.
//
// Per H application: gather -> alpha per node -> local march (stencil
// from a 13-deep register ring per node and an LDS plane of the batch) ->
// scatter of chi alpha per column and node + three-term update.
#include <type_traits>

#include "sq_device_common.h"

namespace sqamd {

namespace {

// B values of one plane point, one per node of the batch (16 bytes)
template <typename T, int B>
struct VecOf { typedef T type __attribute__((ext_vector_type(B))); };

template <typename V, typename T>
__device__ __forceinline__ T vget(const V &v, int b) { return v[b]; }
template <typename V, typename T>
__device__ __forceinline__ void vset(V &v, int b, T x) { v[b] = x; }

typedef float f32x4 __attribute__((ext_vector_type(4)));

}  // namespace

// Sum of B values over the workgroup (fixed tree, deterministic); every
// thread gets the B totals.
template <int B, int NW>
__device__ __forceinline__ void block_sum_vec(double v[B], double *red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < B; b++) v[b] = wave_sum(v[b]);
    if (lane == 0)
#pragma unroll
        for (int b = 0; b < B; b++) red[b * NW + wave] = v[b];
    __syncthreads();
    if (threadIdx.x < B) {
        double s = 0.0;
        for (int w = 0; w < NW; w++) s += red[threadIdx.x * NW + w];
        red[B * NW + threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; b++) v[b] = red[B * NW + b];
    __syncthreads();
}

// One workgroup of 1024 threads per batch, one (x,y) column per thread.
// UG: z-planes per gather trip,
// UC: planes per scatter/update chunk.
template <typename T, int B, int MW, int UG, int UC>
__global__ void __launch_bounds__(1024) lanczos_zbatch(KParams P)
{
    constexpr int WG = 1024, NW = WG / 64, O = 6, W = 2 * O + 1, NLM = 6;
    using VB = typename VecOf<T, B>::type;
    typedef T V4 __attribute__((ext_vector_type(4)));
    typedef T V2 __attribute__((ext_vector_type(2)));
    extern __shared__ __align__(16) unsigned char smem[];
    int *img_list = reinterpret_cast<int *>(smem);
    int4 *tlo = reinterpret_cast<int4 *>(smem + P.lds_tlo);
    int4 *tdim = reinterpret_cast<int4 *>(smem + P.lds_tdim);
    int64_t *toff = reinterpret_cast<int64_t *>(smem + P.lds_toff);
    double4 *tcen = reinterpret_cast<double4 *>(smem + P.lds_tcen);
    double *A = reinterpret_cast<double *>(smem + P.lds_A);          // [img][B][NLM]
    double *alpha = reinterpret_cast<double *>(smem + P.lds_alpha);  // [img][B][NLM]
    double *red = reinterpret_cast<double *>(smem + P.lds_red);
    double *s_ab = reinterpret_cast<double *>(smem + P.lds_local);   // [B][2][nqp+1] (early exit)
    VB *plane = reinterpret_cast<VB *>(smem + P.lds_plane);
    int4 *trun = reinterpret_cast<int4 *>(smem + P.lds_run);          // [img]: first, last image of its atom; type
    double *pshare = reinterpret_cast<double *>(smem + P.lds_share);  // [img][B]: an atom's share of <v, V_nl v>
    __shared__ int s_count;
    __shared__ T s_w[3][8];
    __shared__ long s_next;
    __shared__ double s_g[4][kMatsubara][2];
    __shared__ int s_flag[4];
    // per-node state of the batch (uniform): vector scale 1/b of v and of
    // vprev, b of the last step, Gauss nodes, status; bit b of s_alive: node b
    // still runs
    __shared__ double s_sv[4], s_svp[4], s_bprev[4];
    __shared__ int s_nq[4], s_st[4], s_alive;
    // per scatter chunk of UC union planes: the images whose z range meets it
    constexpr int kMaxChunks = 16;
    __shared__ unsigned s_zmask[kMaxChunks][MW];

    const int nx = P.nloc[0], ny = P.nloc[1], nz = P.nloc[2];
    const int ncol = nx * ny;
    const int nzs = nz + B - 1;                       // union planes of the storage
    const int px = P.plane_px, pplane = px * (ny + 2 * O);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = tid;
    const bool has = col < ncol;
    const int cj = has ? col / nx : 0, ci = has ? col - cj * nx : 0;
    const int pcol = (cj + O) * px + ci + O;
    for (int q = tid; q < 2 * pplane; q += WG) plane[q] = VB{};   // the halo stays zero
    const size_t vlen = (size_t)nzs * ncol * B;
    T *__restrict__ buf0 = reinterpret_cast<T *>(P.scratch) + (size_t)blockIdx.x * 3 * vlen;
    T *__restrict__ buf1 = buf0 + vlen;
    T *__restrict__ Bu = buf1 + vlen;
    const T *__restrict__ diag_ext = reinterpret_cast<const T *>(P.diag_ext);
    // FP32: (t0, t2, t2 z, t2 z^2) per image point; FP64: the radial pairs (t0, t2)
    const T *__restrict__ bas = reinterpret_cast<const T *>(P.basis);
    const T *__restrict__ rad = reinterpret_cast<const T *>(P.rad);
    const double hz = P.hgrid[2], rz = P.range[2];
    if (tid < 21) {
        if constexpr (sizeof(T) == 8) s_w[tid / 7][tid % 7] = P.cw64[tid / 7][tid % 7];
        else s_w[tid / 7][tid % 7] = P.cw32[tid / 7][tid % 7];
    }
    __syncthreads();
    const long bstride = (long)P.box_n[0] * P.box_n[1];   // node-id step of one plane in z
    const int ccen = P.center % ncol, zcen = P.center / ncol;

    // work queue over batches: eight contiguous parts, one per XCD
    const long part = (P.node_count + 7) / 8;
    const int home = blockIdx.x & 7;
    int q_at = 0;
    for (;;) {
        if (tid == 0) {
            long r = -1;
            for (; q_at < 8; q_at++) {
                const int x = (home + q_at) & 7;
                const long lo = x * part, hi = min(P.node_count, lo + part);
                if (lo >= hi) continue;
                const long k = atomicAdd(&P.queue[x], 1);
                if (lo + k < hi) { r = lo + k; break; }
            }
            s_next = r;
        }
        __syncthreads();
        const long r = s_next;
        __syncthreads();
        if (r < 0) break;
        const int2 bt = P.batch[r];
        const long id0 = bt.x;
        const int nb = bt.y;                            // live nodes of the batch (1..B)
        const int NZ = nz + nb - 1;                     // union planes the live nodes span
        int node[3];
        node_coords(P, id0, node);
        const int cube_lo[3] = {node[0] - P.hw[0], node[1] - P.hw[1], node[2] - P.hw[2]};
        // ---- images whose box meets the union cube, ascending
        if (tid == 0) s_count = 0;
        __syncthreads();
        for (int base = 0; base < P.n_images; base += WG) {
            const int g = base + tid;
            bool hit = false;
            int4 lo4 = make_int4(0, 0, 0, 0), dm4 = make_int4(0, 0, 0, 0);
            if (g < P.n_images) {
                lo4 = P.img_lo[g];
                dm4 = P.img_dim[g];
                hit = lo4.x <= cube_lo[0] + nx - 1 && lo4.x + dm4.x - 1 >= cube_lo[0] &&
                      lo4.y <= cube_lo[1] + ny - 1 && lo4.y + dm4.y - 1 >= cube_lo[1] &&
                      lo4.z <= cube_lo[2] + NZ - 1 && lo4.z + dm4.z - 1 >= cube_lo[2];
            }
            const unsigned long long bal = __ballot(hit);
            const int before = __popcll(bal & ((1ull << lane) - 1ull));
            int *wcnt = reinterpret_cast<int *>(red);
            if (lane == 0) wcnt[wave] = __popcll(bal);
            __syncthreads();
            int off = s_count;
            for (int w = 0; w < wave; w++) off += wcnt[w];
            if (hit) {
                const int li = off + before;
                if (li < P.img_cap) {
                    img_list[li] = g;
                    tlo[li] = lo4;
                    tdim[li] = dm4;
                    toff[li] = P.img_rad[g];
                    tcen[li] = P.img_cen[g];
                }
            }
            __syncthreads();
            if (tid == 0) {
                int t = s_count;
                for (int w = 0; w < NW; w++) t += wcnt[w];
                s_count = t;
            }
            __syncthreads();
        }
        const int nimg = s_count;
        if (nimg > P.img_cap) {
            if (tid < nb) P.status[id0 + tid * bstride] = 4;
            __syncthreads();
            continue;
        }
        // ---- per image: the run of list entries holding the same atom (the list is
        // ascending, so an atom's images are contiguous) and its type
        for (int li = tid; li < nimg; li += WG) {
            const int atom = P.img_atom[img_list[li]];
            int s0 = li, e0 = li;
            while (s0 > 0 && P.img_atom[img_list[s0 - 1]] == atom) s0--;
            while (e0 + 1 < nimg && P.img_atom[img_list[e0 + 1]] == atom) e0++;
            trun[li] = make_int4(s0, e0, P.img_type[img_list[li]], 0);
        }
        // ---- per column: bitmask of the images whose x/y box holds it
        unsigned cmask[MW];
        const int gxc = cube_lo[0] + ci, gyc = cube_lo[1] + cj;
        const int64_t ecol = ((int64_t)(cube_lo[2] - P.ext0[2]) * P.extn[1] + (gyc - P.ext0[1])) * P.extn[0] +
                             (gxc - P.ext0[0]);
#pragma unroll
        for (int w = 0; w < MW; w++) cmask[w] = 0u;
        for (int li = 0; li < nimg; li++) {
            const int4 lo = tlo[li], dm = tdim[li];
            const bool in = has && gxc >= lo.x && gxc < lo.x + dm.x && gyc >= lo.y && gyc < lo.y + dm.y;
#pragma unroll
            for (int w = 0; w < MW; w++)
                if (w == (li >> 5) && in) cmask[w] |= 1u << (li & 31);
        }
        const int64_t eplane = (int64_t)P.extn[1] * P.extn[0];
        // ---- per scatter chunk: bitmask of the images whose z range meets it, so a
        // column's chunk visits only those (all images when the cube is too tall)
        const int nch = (NZ + UC - 1) / UC;
        const bool zmasked = nch <= kMaxChunks;
        if (zmasked)
            for (int e = tid; e < nch * MW; e += WG) {
                const int c = e / MW, w = e - c * MW;
                const int z0 = c * UC, z1 = z0 + UC - 1;
                unsigned bits = 0u;
                for (int k = 0; k < 32; k++) {
                    const int li = w * 32 + k;
                    if (li >= nimg) break;
                    const int4 lo = tlo[li], dm = tdim[li];
                    const int zlo = max(lo.z, cube_lo[2]) - cube_lo[2];
                    const int zhi = min(lo.z + dm.z - 1, cube_lo[2] + NZ - 1) - cube_lo[2];
                    if (zhi >= z0 && zlo <= z1) bits |= 1u << k;
                }
                s_zmask[c][w] = bits;
            }
        // ---- Lanczos start: v_b = e_center of node b (union plane zcen + b), vprev = 0
        for (size_t p = tid; p < 2 * vlen; p += WG) buf0[p] = T(0);
        __syncthreads();
        if (tid < nb) buf0[((size_t)(zcen + tid) * ncol + ccen) * B + tid] = T(1);
        __syncthreads();
        T *Bv = buf0, *Bvp = buf1;
        if (tid < B) {
            s_sv[tid] = tid < nb ? 1.0 : 0.0;
            s_svp[tid] = 0.0;
            s_bprev[tid] = 0.0;
            s_nq[tid] = P.nqp;
            s_st[tid] = 0;
        }
        if (tid == 0) s_alive = (1 << nb) - 1;
        __syncthreads();
        int checks = 0;
        for (int step = 0; step <= P.nqp && s_alive; step++) {
            // ---- gather on the matrix cores: A[li][n][ch] = sum_{box & cube} v_n chi_ch
            for (int li = wave; li < nimg; li += NW) {
                const int4 lo = tlo[li], dm = tdim[li];
                const int nlm = dm.w & 255;
                const double4 cen = tcen[li];
                const int ilo[3] = {lo.x, lo.y, lo.z}, idm[3] = {dm.x, dm.y, dm.z};
                const int ext[3] = {nx, ny, NZ};
                int a0[3], n[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int s0 = max(ilo[d], cube_lo[d]);
                    const int e0 = min(ilo[d] + idm[d] - 1, cube_lo[d] + ext[d] - 1);
                    a0[d] = s0;
                    n[d] = e0 - s0 + 1;
                }
                const int nxy = n[0] * n[1];
                const float inv0 = 1.0f / (float)n[0];
                const int zs2 = dm.x * dm.y;
                // the basis tables hold z relative to the image centre, so a node's
                // values depend neither on where the batch's union starts nor on how
                // a rank box clips the image box
                const int j = lane & 3;                 // basis index (A) and node index (B, D columns)
                T acc[NLM];
#pragma unroll
                for (int ch = 0; ch < NLM; ch++) acc[ch] = T(0);
                if constexpr (sizeof(T) == 4 && UG >= 4) {
                    // matrix cores: 16 columns per instruction (block = lane / 4), one plane per step
                    const int g = lane >> 2;
                    for (int q0 = 0; q0 < nxy; q0 += 16) {
                        const int q = q0 + g;
                        const bool ok = q < nxy;
                        const int yi = ok ? fdiv(q, n[0], inv0) : 0;
                        const int xi = ok ? q - yi * n[0] : 0;
                        const int gx = a0[0] + xi, gy = a0[1] + yi;
                        const int c = (gy - cube_lo[1]) * nx + (gx - cube_lo[0]);
                        const T *vp = Bv + ((size_t)(a0[2] - cube_lo[2]) * ncol + c) * B + (j < B ? j : 0);
                        const T *rp = bas + 2 * toff[li] +
                                      ((int64_t)((a0[2] - lo.z) * dm.y + (gy - lo.y)) * dm.x + (gx - lo.x)) * 4 + j;
                        f32x4 D = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
                        for (int zi = 0; zi < n[2]; zi += UG) {
                            T tv[UG], vv[UG];
#pragma unroll
                            for (int u = 0; u < UG; u++) {
                                const bool in = ok && zi + u < n[2];
                                tv[u] = in ? rp[(int64_t)(zi + u) * zs2 * 4] : T(0);
                                vv[u] = (in && j < B) ? vp[(size_t)(zi + u) * ncol * B] : T(0);
                            }
#pragma unroll
                            for (int u = 0; u < UG; u++) D = __builtin_amdgcn_mfma_f32_4x4x1f32(tv[u], vv[u], D, 0, 0, 0);
                        }
                        // D[m] = S[basis m][node j] of column q
                        if (ok && j < B) {
                            const T x = (T)(P.hgrid[0] * gx - P.range[0] - cen.x);
                            const T y = (T)(P.hgrid[1] * gy - P.range[1] - cen.y);
                            const T C = T(1.092548430592079), C20 = T(0.315391565252520), C22 = T(0.546274215296040);
                            acc[0] += D[0];
                            acc[1] = fma(C * (x * y), D[1], acc[1]);
                            acc[2] = fma(C * y, D[2], acc[2]);
                            acc[3] = fma(C20, fma(-(x * x) - y * y, D[1], T(2) * D[3]), acc[3]);
                            acc[4] = fma(C * x, D[2], acc[4]);
                            acc[5] = fma(C22 * (x * x - y * y), D[1], acc[5]);
                        }
                    }
                } else {
                    // vector ALU: one column of the box per lane, UG planes a trip.  Each
                    // node's four basis sums run over z in ascending order from the box's
                    // first plane in the union; planes outside a node's own cube hold
                    // exact zeros, so a node's sums do not depend on the batch it runs in.
                    constexpr int UZ = UG;
                    T accb[B][NLM];
#pragma unroll
                    for (int b = 0; b < B; b++)
#pragma unroll
                        for (int ch = 0; ch < NLM; ch++) accb[b][ch] = T(0);
                    for (int q0 = 0; q0 < nxy; q0 += 64) {
                        const int q = q0 + lane;
                        const bool ok = q < nxy;
                        const int yi = ok ? fdiv(q, n[0], inv0) : 0;
                        const int xi = ok ? q - yi * n[0] : 0;
                        const int gx = a0[0] + xi, gy = a0[1] + yi;
                        const int c = (gy - cube_lo[1]) * nx + (gx - cube_lo[0]);
                        const VB *vp = reinterpret_cast<const VB *>(Bv) + (size_t)(a0[2] - cube_lo[2]) * ncol + c;
                        const int64_t pix = ((int64_t)(a0[2] - lo.z) * dm.y + (gy - lo.y)) * dm.x + (gx - lo.x);
                        T sm[B][4];
#pragma unroll
                        for (int b = 0; b < B; b++)
#pragma unroll
                            for (int m = 0; m < 4; m++) sm[b][m] = T(0);
#pragma unroll 1
                        for (int zi = 0; zi < n[2]; zi += UZ) {
                            V4 tb[UZ];
                            VB vv[UZ];
#pragma unroll
                            for (int u = 0; u < UZ; u++) {
                                const bool in = ok && zi + u < n[2];
                                const int64_t p = pix + (int64_t)(zi + u) * zs2;
                                if constexpr (sizeof(T) == 8) {
                                    // z of the plane: uniform across the wave
                                    const V2 r = in ? reinterpret_cast<const V2 *>(rad + toff[li])[p] : V2{};
                                    const double z = hz * (a0[2] + zi + u) - rz - cen.z;
                                    const double t2z = r.y * z;
                                    tb[u] = V4{r.x, r.y, t2z, t2z * z};
                                } else {
                                    tb[u] = in ? reinterpret_cast<const V4 *>(bas + 2 * toff[li])[p] : V4{};
                                }
                                vv[u] = in ? vp[(size_t)(zi + u) * ncol] : VB{};
                            }
#pragma unroll
                            for (int u = 0; u < UZ; u++)
#pragma unroll
                                for (int b = 0; b < B; b++) {
                                    const T vb = vv[u][b];
#pragma unroll
                                    for (int m = 0; m < 4; m++) sm[b][m] = fma(tb[u][m], vb, sm[b][m]);
                                }
                        }
                        if (ok) {
                            const T x = (T)(P.hgrid[0] * gx - P.range[0] - cen.x);
                            const T y = (T)(P.hgrid[1] * gy - P.range[1] - cen.y);
                            const T C = T(1.092548430592079), C20 = T(0.315391565252520), C22 = T(0.546274215296040);
#pragma unroll
                            for (int b = 0; b < B; b++) {
                                accb[b][0] += sm[b][0];
                                accb[b][1] = fma(C * (x * y), sm[b][1], accb[b][1]);
                                accb[b][2] = fma(C * y, sm[b][2], accb[b][2]);
                                accb[b][3] = fma(C20, fma(-(x * x) - y * y, sm[b][1], T(2) * sm[b][3]), accb[b][3]);
                                accb[b][4] = fma(C * x, sm[b][2], accb[b][4]);
                                accb[b][5] = fma(C22 * (x * x - y * y), sm[b][1], accb[b][5]);
                            }
                        }
                    }
                    // butterfly over all lanes, node by node, in the working precision (one
                    // lane exchange a level in FP32); lane 0 writes
#pragma unroll
                    for (int b = 0; b < B; b++)
#pragma unroll
                        for (int ch = 0; ch < NLM; ch++) {
                            T sv = accb[b][ch];
#pragma unroll
                            for (int off = 1; off < 64; off <<= 1) sv += __shfl_xor(sv, off, 64);
                            if (lane == 0 && ch < nlm) A[(li * B + b) * NLM + ch] = (double)sv * (double)(T)s_sv[b];
                        }
                    continue;
                }
                // butterfly over the lanes of one node (lane bits 2..5) in the working
                // precision; lanes 0..B-1 end with node lane
#pragma unroll
                for (int ch = 0; ch < NLM; ch++) {
                    T s = acc[ch];
#pragma unroll
                    for (int off = 4; off < 64; off <<= 1) s += __shfl_xor(s, off, 64);
                    if (lane < B && ch < nlm) A[(li * B + lane) * NLM + ch] = (double)s * (double)(T)s_sv[lane];   // the sums used the stored v
                }
            }
            __syncthreads();
            // ---- per atom (the first list entry of its run of images) and node:
            // alpha of the atom's images and the atom's share of <v_b, V_nl v_b>;
            // then per node the shares in ascending atom order.  Images or atoms
            // outside a node's cube add exact zeros, so neither sum depends on the
            // batch the node runs in.
            for (int e = tid; e < nimg * B; e += WG) {
                const int li = e / B, b = e - li * B;
                const int4 run = trun[li];
                if (run.x != li) continue;
                const int s = run.x, t = run.y;
                const int nlm = tdim[li].w & 255;
                const double *gam = P.gamma_ch + run.z * kMaxChannels;
                double share = 0.0;
                double a6[NLM];
                if (P.mode == 0) {   // one coefficient per atom
                    double al = 0.0, sa = 0.0;
                    for (int q = s; q <= t; q++)
                        for (int ch = 0; ch < nlm; ch++) {
                            const double x = A[(q * B + b) * NLM + ch];
                            al += gam[ch] * x;
                            sa += x;
                        }
                    for (int ch = 0; ch < NLM; ch++) a6[ch] = al;
                    share = al * sa;
                } else {             // per (l,m) channel
                    for (int ch = 0; ch < NLM; ch++) {
                        double sa = 0.0, al = 0.0;
                        if (ch < nlm) {
                            for (int q = s; q <= t; q++) sa += A[(q * B + b) * NLM + ch];
                            al = gam[ch] * sa;
                            share = fma(al, sa, share);
                        }
                        a6[ch] = al;
                    }
                }
                // the scatter's column-independent coefficients of the atom's images
                // for node b: chi alpha = t0 K0 + t2 (xy Kxy + xx Kxx + yy Kyy)
                // + t2 z (y Ky + x Kx) + t2 z^2 (-(Kxx + Kyy)), (x, y) the column
                // relative to the image centre
                {
                    const T C = T(1.092548430592079), C20 = T(0.315391565252520), C22 = T(0.546274215296040);
                    const T a1 = (T)a6[1], a2 = (T)a6[2], a3 = (T)a6[3], a4 = (T)a6[4], a5 = (T)a6[5];
                    const T k6[NLM] = {(T)a6[0], C * a1, fma(C22, a5, -(C20 * a3)), fma(-C22, a5, -(C20 * a3)),
                                       C * a2, C * a4};
                    T *coef = reinterpret_cast<T *>(alpha);
                    for (int q = s; q <= t; q++)
#pragma unroll
                        for (int k = 0; k < NLM; k++) coef[(q * B + b) * NLM + k] = k6[k];
                }
                pshare[li * B + b] = share;
            }
            __syncthreads();
            if (tid < B) {
                double sn = 0.0;
                for (int li = 0; li < nimg; li++)
                    if (trun[li].x == li) sn += pshare[li * B + tid];
                red[(B + 1) * NW + tid] = sn;
            }
            __syncthreads();
            double pn[B];
#pragma unroll
            for (int b = 0; b < B; b++) pn[b] = red[(B + 1) * NW + b];
            __syncthreads();
            // ---- local part, marching the union planes: u_b = (L + diag) v_b into Bu and <v_b, u_b>
            // <v_b, u_b> and later ||w_b||^2: in FP32 each product is exact in FP64
            // (24 x 24 bits), so the sums are FP64 sums of exact products
            double pa[B];
#pragma unroll
            for (int b = 0; b < B; b++) pa[b] = 0.0;
            T win[B][W];
            T svr[B];
#pragma unroll
            for (int b = 0; b < B; b++) svr[b] = (T)s_sv[b];
#pragma unroll
            for (int q = 0; q < W; q++) {
                const int z = q - O;
                VB x{};
                if (has && z >= 0 && z < NZ) x = *reinterpret_cast<const VB *>(Bv + ((size_t)z * ncol + col) * B);
#pragma unroll
                for (int b = 0; b < B; b++) win[b][q] = svr[b] * vget<VB, T>(x, b);
            }
            T dgn = has ? diag_ext[ecol] : T(0);
            // Plane offsets are carried as opaque scalars (one add per plane):
            // left to itself the compiler hoists a 64-bit address per unrolled
            // plane out of the loop and spills them.
            const int ncolB = ncol * B, colB = col * B;
            int zo = 0;                       // z * ncolB
            int64_t eo = eplane;              // (z + 1) * eplane
            for (int zb = 0; zb < NZ; zb += W) {
#pragma unroll
                for (int s = 0; s < W; s++) {
                    const int z = zb + s;
                    if (z < NZ) {   // uniform
                        VB *pl = plane + (z & 1) * pplane;
                        VB nx4{};
                        const T dg = dgn;
                        if (has) {
                            VB cur;
#pragma unroll
                            for (int b = 0; b < B; b++) vset<VB, T>(cur, b, win[b][(s + O) % W]);
                            pl[pcol] = cur;
                            if (z + O + 1 < NZ) nx4 = *reinterpret_cast<const VB *>(Bv + (colB + zo + (O + 1) * ncolB));
                            if (z + 1 < NZ) dgn = diag_ext[ecol + eo];
                        }
                        __syncthreads();
                        if (has) {
                            // one fma chain per node (the batch supplies the independent work)
                            const VB *pc = pl + pcol;
                            T sm[B];
#pragma unroll
                            for (int b = 0; b < B; b++) sm[b] = dg * win[b][(s + O) % W];
#pragma unroll
                            for (int l = 1; l <= O; l++) {
                                const T wz = s_w[2][l], wy = s_w[1][l], wx = s_w[0][l];
#pragma unroll
                                for (int b = 0; b < B; b++)
                                    sm[b] = fma(wz, win[b][(s + O - l) % W] + win[b][(s + O + l) % W], sm[b]);
                                {
                                    const VB m1 = pc[-l], p1 = pc[l];
#pragma unroll
                                    for (int b = 0; b < B; b++) sm[b] = fma(wx, vget<VB, T>(m1, b) + vget<VB, T>(p1, b), sm[b]);
                                }
                                {
                                    const VB m1 = pc[-l * px], p1 = pc[l * px];
#pragma unroll
                                    for (int b = 0; b < B; b++) sm[b] = fma(wy, vget<VB, T>(m1, b) + vget<VB, T>(p1, b), sm[b]);
                                }
                            }
                            VB uo;
#pragma unroll
                            for (int b = 0; b < B; b++) {
                                const T vc = win[b][(s + O) % W];
                                vset<VB, T>(uo, b, sm[b]);
                                pa[b] = fma((double)vc, (double)sm[b], pa[b]);
                            }
                            *reinterpret_cast<VB *>(Bu + (colB + zo)) = uo;
                        }
                        // plane z - O leaves the window; plane z + O + 1 takes its slot
#pragma unroll
                        for (int b = 0; b < B; b++) win[b][s % W] = svr[b] * vget<VB, T>(nx4, b);
                        zo += ncolB;
                        eo += eplane;
                        asm volatile("" : "+s"(zo), "+s"(eo));
                    }
                }
            }
            double a[B];
#pragma unroll
            for (int b = 0; b < B; b++) a[b] = pa[b];
            block_sum_vec<B, NW>(a, red);
#pragma unroll
            for (int b = 0; b < B; b++) a[b] += pn[b];
            // ---- per column, chunks of UC union planes in registers: local part
            // of u from the march + chi alpha of the column's images (ascending),
            // then w = (u - a v) - bprev vprev over vprev's storage and ||w||^2
            T ta[B], tb[B];
#pragma unroll
            for (int b = 0; b < B; b++) { ta[b] = T(-a[b]) * (T)s_sv[b]; tb[b] = T(-s_bprev[b]) * (T)s_svp[b]; }
            const int alive = s_alive;
            double pb[B];
#pragma unroll
            for (int b = 0; b < B; b++) pb[b] = 0.0;
            if (has) {
                const T x0 = (T)(P.hgrid[0] * gxc - P.range[0]);
                const T y0 = (T)(P.hgrid[1] * gyc - P.range[1]);
#pragma unroll 1
                for (int z0 = 0; z0 < NZ; z0 += UC) {
                    T uu[B][UC];
#pragma unroll
                    for (int u = 0; u < UC; u++) {
                        VB x{};
                        if (z0 + u < NZ) x = *reinterpret_cast<const VB *>(Bu + ((size_t)(z0 + u) * ncol + col) * B);
#pragma unroll
                        for (int b = 0; b < B; b++) uu[b][u] = vget<VB, T>(x, b);
                    }
#pragma unroll 1
                    for (int wd = 0; wd < MW; wd++) {
                        unsigned m = cmask[0];
#pragma unroll
                        for (int q = 1; q < MW; q++) m = (wd == q) ? cmask[q] : m;
                        if (zmasked) m &= s_zmask[z0 / UC][wd];
                        while (m) {
                            const int li = (wd << 5) + __ffs(m) - 1;
                            m &= m - 1u;
                            const int4 lo = tlo[li], dm = tdim[li];
                            const int zlo = max(lo.z, cube_lo[2]) - cube_lo[2];
                            const int zhi = min(lo.z + dm.z - 1, cube_lo[2] + NZ - 1) - cube_lo[2];
                            if (zhi < z0 || zlo > z0 + UC - 1) continue;
                            const double4 cen = tcen[li];
                            // the point's radial pair; its basis quadruple sits at 2 rb0 in the table
                            const int64_t rb0 = toff[li] +
                                                ((int64_t)((cube_lo[2] + z0 - lo.z) * dm.y + (gyc - lo.y)) * dm.x + (gxc - lo.x)) * 2;
                            const T xc = x0 - (T)cen.x, yc = y0 - (T)cen.y;
                            const T xy = xc * yc, xx = xc * xc, yy = yc * yc;
                            const T *coef = reinterpret_cast<const T *>(alpha) + li * B * NLM;
                            T al0[B], p0[B], p1[B], p2[B];
#pragma unroll
                            for (int b = 0; b < B; b++) {
                                const T *k6 = coef + b * NLM;
                                al0[b] = k6[0];
                                p0[b] = fma(xy, k6[1], fma(xx, k6[2], yy * k6[3]));
                                p1[b] = fma(yc, k6[4], xc * k6[5]);
                                p2[b] = -(k6[2] + k6[3]);
                            }
                            V4 t4[UC];
#pragma unroll
                            for (int u = 0; u < UC; u++) {
                                const bool ok = z0 + u >= zlo && z0 + u <= zhi;
                                if constexpr (sizeof(T) == 8) {
                                    const V2 r = ok ? reinterpret_cast<const V2 *>(rad + rb0)[u * (dm.x * dm.y)] : V2{};
                                    const double z = hz * (cube_lo[2] + z0 + u) - rz - cen.z;
                                    const double t2z = r.y * z;
                                    t4[u] = V4{r.x, r.y, t2z, t2z * z};
                                } else {
                                    t4[u] = ok ? reinterpret_cast<const V4 *>(bas + 2 * rb0)[u * (dm.x * dm.y)] : V4{};
                                }
                            }
#pragma unroll
                            for (int u = 0; u < UC; u++) {
                                if (z0 + u >= zlo && z0 + u <= zhi)
#pragma unroll
                                    for (int b = 0; b < B; b++)
                                        uu[b][u] = fma(t4[u][3], p2[b], fma(t4[u][2], p1[b], fma(t4[u][1], p0[b], fma(t4[u][0], al0[b], uu[b][u]))));
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UC; u++) {
                        const int z = z0 + u;
                        if (z >= NZ) break;
                        const size_t p = ((size_t)z * ncol + col) * B;
                        const VB vr = *reinterpret_cast<const VB *>(Bv + p);
                        const VB vpr = *reinterpret_cast<const VB *>(Bvp + p);
                        VB wo;
#pragma unroll
                        for (int b = 0; b < B; b++) {
                            const bool in = z >= b && z < b + nz && ((alive >> b) & 1);
                            T w;
                            if constexpr (sizeof(T) == 8) {
                                w = fma(ta[b], vget<VB, T>(vr, b), uu[b][u]);
                                w = fma(tb[b], vget<VB, T>(vpr, b), w);
                                w = in ? w : T(0);
                                pb[b] = fma(w, w, pb[b]);
                            } else {
                                w = fmaf(ta[b], vget<VB, T>(vr, b), uu[b][u]);
                                w = fmaf(tb[b], vget<VB, T>(vpr, b), w);
                                w = in ? w : T(0);
                                pb[b] = fma((double)w, (double)w, pb[b]);
                            }
                            vset<VB, T>(wo, b, w);
                        }
                        *reinterpret_cast<VB *>(Bvp + p) = wo;
                    }
                }
            }
            double bb[B];
#pragma unroll
            for (int b = 0; b < B; b++) bb[b] = pb[b];
            block_sum_vec<B, NW>(bb, red);
            double bv[B];
#pragma unroll
            for (int b = 0; b < B; b++) bv[b] = sqrt(bb[b]);
            if (tid == 0) {
                int al = s_alive;
#pragma unroll
                for (int b = 0; b < B; b++)
                    if ((al >> b) & 1) {
                        const long id = id0 + b * bstride;
                        P.out_a[id * (P.nqp + 1) + step] = a[b];
                        P.out_b[id * (P.nqp + 1) + step] = bv[b];
                        if (bv[b] == 0.0) { s_st[b] = 7; al &= ~(1 << b); }
                    }
                s_alive = al;
            }
            if (P.exit_tol > 0.0) {
                // per-node early exit on the Matsubara terms of the node's Gauss rule
                // (the z-march kernel's test, one group of kMatsubara lanes per node)
                const int ns = P.nqp + 1;
                if (tid == 0)
#pragma unroll
                    for (int b = 0; b < B; b++) { s_ab[(b * 2) * ns + step] = a[b]; s_ab[(b * 2 + 1) * ns + step] = bv[b]; }
                const int n = step + 1;
                if (n >= P.exit_min && (n - P.exit_min) % P.exit_every == 0 && n < P.nqp) {
                    if (tid < B) s_flag[tid] = 0;
                    __syncthreads();
                    if (tid < B * kMatsubara) {
                        const int b = tid / kMatsubara, mm = tid - b * kMatsubara;
                        const double *ab = s_ab + (b * 2) * ns, *bb2 = s_ab + (b * 2 + 1) * ns;
                        const double wm = (2 * mm + 1) * M_PI / P.exit_beta;
                        double tr = P.exit_lf - ab[n - 1], ti = wm;
                        for (int jj = n - 2; jj >= 0; jj--) {
                            const double den = tr * tr + ti * ti;
                            const double q2 = bb2[jj] * bb2[jj];
                            tr = (P.exit_lf - ab[jj]) - q2 * tr / den;
                            ti = wm + q2 * ti / den;
                        }
                        const double den = tr * tr + ti * ti;
                        const double gr = tr / den, gi = -ti / den;
                        const double dr = gr - s_g[b][mm][0], di = gi - s_g[b][mm][1];
                        const double d = (2.0 / P.exit_beta) * sqrt(dr * dr + di * di);
                        if (checks == 0 || !(d < P.exit_tol)) s_flag[b] = 1;
                        s_g[b][mm][0] = gr;
                        s_g[b][mm][1] = gi;
                    }
                    __syncthreads();
                    if (tid == 0) {
                        int al = s_alive;
                        for (int b = 0; b < B; b++)
                            if (((al >> b) & 1) && s_flag[b] == 0) { s_nq[b] = n; al &= ~(1 << b); }
                        s_alive = al;
                    }
                    checks++;
                }
            }
            __syncthreads();
            T *t = Bv; Bv = Bvp; Bvp = t;
            if (tid < B) {
                double bt = bv[0];
#pragma unroll
                for (int b = 1; b < B; b++) bt = tid == b ? bv[b] : bt;
                s_svp[tid] = s_sv[tid];
                s_sv[tid] = ((s_alive >> tid) & 1) ? (double)T(1.0 / bt) : 0.0;
                s_bprev[tid] = bt;
            }
            __syncthreads();
        }
        if (tid < nb) {
            const long id = id0 + tid * bstride;
            P.status[id] = s_st[tid];
            if (P.nsteps) P.nsteps[id] = s_nq[tid];
        }
        __syncthreads();
    }
}

// Column basis table of the batched kernel's FP32 form: for every point of an
// s+d image box, (t0, t2, t2 z, t2 z^2) with t0, t2 the radial factors and z
// the point's height above the image centre (lsdft_nl.cc:174-190), formed in
// FP64 and rounded once.  Entry 4 p of image g starts at 2 img_rad[g].  (The
// FP64 form forms the same values in the kernel from the radial pairs.)
__global__ void basis_tables(const double *__restrict__ rad, const int4 *__restrict__ lo, const int4 *__restrict__ dim,
                             const int64_t *__restrict__ roff, const double4 *__restrict__ cen, double hz, double rz,
                             float *__restrict__ b32)
{
    const int g = blockIdx.x;
    const int4 l = lo[g], d = dim[g];
    const long npt = (long)d.x * d.y * d.z;
    const int64_t o = roff[g];
    for (long p = threadIdx.x; p < npt; p += blockDim.x) {
        const int zi = (int)(p / ((long)d.x * d.y));
        const double z = hz * (l.z + zi) - rz - cen[g].z;
        const double t0 = rad[o + 2 * p], t2 = rad[o + 2 * p + 1];
        const double v[4] = {t0, t2, t2 * z, (t2 * z) * z};
        for (int k = 0; k < 4; k++) b32[2 * o + 4 * p + k] = (float)v[k];
    }
}

hipError_t launch_basis_tables(const double *rad, int n_images, const int4 *lo, const int4 *dim, const int64_t *roff,
                               const double4 *cen, double hz, double rz, float *b32, hipStream_t st)
{
    if (n_images <= 0) return hipSuccess;
    hipLaunchKernelGGL(basis_tables, dim3(n_images), dim3(256), 0, st, rad, lo, dim, roff, cen, hz, rz, b32);
    return hipGetLastError();
}

// planes of a scatter/update chunk: as many as the registers hold beside the
// batch's rows (B values a plane) and the basis quadruples
template <typename T, int B>
constexpr int scatter_chunk() { return sizeof(T) == 8 || B == 4 ? 4 : 8; }

template <typename T, int B, int UG>
static hipError_t launch_b(const KParams &P, int grid, size_t lds, hipStream_t st)
{
    constexpr int UC = scatter_chunk<T, B>();
    if (P.img_cap <= 128) hipLaunchKernelGGL((lanczos_zbatch<T, B, 4, UG, UC>), dim3(grid), dim3(1024), lds, st, P);
    else if (P.img_cap <= 256) hipLaunchKernelGGL((lanczos_zbatch<T, B, 8, UG, UC>), dim3(grid), dim3(1024), lds, st, P);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T, int B, int UG>
static int occ_b(const KParams &P, size_t lds)
{
    constexpr int UC = scatter_chunk<T, B>();
    int n = 0;
    hipError_t e = P.img_cap <= 128
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lanczos_zbatch<T, B, 4, UG, UC>, 1024, lds)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lanczos_zbatch<T, B, 8, UG, UC>, 1024, lds);
    return e == hipSuccess ? n : 0;
}

// The tuned shapes: FP64 batches of 2 nodes, FP32 batches of 2 or 4; the
// gather trip (UG) of 1 or 2 planes on the vector ALU (FP64), 4 planes on the
// matrix cores or 2 on the vector ALU (FP32).
bool zbatch_supported(int precision, int B, int ug)
{
    if (precision == 0) return B == 2 && (ug == 1 || ug == 2);
    return (B == 2 || B == 4) && (ug == 4 || ug == 2);
}

hipError_t launch_zbatch(const KParams &P, int precision, int B, int ug, int grid, size_t lds, hipStream_t st)
{
    if (P.nloc[0] * P.nloc[1] > 1024 || P.layout != 1 || P.nlm_max > 6) return hipErrorInvalidValue;
    if (precision == 0) {
        if (B == 2 && ug == 1) return launch_b<double, 2, 1>(P, grid, lds, st);
        if (B == 2 && ug == 2) return launch_b<double, 2, 2>(P, grid, lds, st);
    } else {
        if (B == 2 && ug == 4) return launch_b<float, 2, 4>(P, grid, lds, st);
        if (B == 2 && ug == 2) return launch_b<float, 2, 2>(P, grid, lds, st);
        if (B == 4 && ug == 4) return launch_b<float, 4, 4>(P, grid, lds, st);
        if (B == 4 && ug == 2) return launch_b<float, 4, 2>(P, grid, lds, st);
    }
    return hipErrorInvalidValue;
}

int zbatch_occupancy(const KParams &P, int precision, int B, int ug, size_t lds)
{
    if (P.nloc[0] * P.nloc[1] > 1024 || P.layout != 1 || P.nlm_max > 6) return 0;
    if (precision == 0) {
        if (B == 2 && ug == 1) return occ_b<double, 2, 1>(P, lds);
        if (B == 2 && ug == 2) return occ_b<double, 2, 2>(P, lds);
    } else {
        if (B == 2 && ug == 4) return occ_b<float, 2, 4>(P, lds);
        if (B == 2 && ug == 2) return occ_b<float, 2, 2>(P, lds);
        if (B == 4 && ug == 4) return occ_b<float, 4, 4>(P, lds);
        if (B == 4 && ug == 2) return occ_b<float, 4, 2>(P, lds);
    }
    return 0;
}

}  // namespace sqamd
