// sq_zmarch.hip -- the fast Lanczos kernel: one workgroup per node cube; each
// thread owns CPT (x,y) columns of the cube and marches them along z.
//
// Per H application:
//  * gather: one wavefront per projector image; lanes take the (x,y) columns
//    of the image box inside the cube and walk z, U planes per trip, so the
//    loads of v and of the projector radial factors are independent and in
//    flight together;
//  * alpha: per atom from the gathered sums (collapsed or KB slots);
//  * apply: the 12 z-neighbours of the stencil come from a 13-deep register
//    window of each column, the 24 x/y-neighbours from the current z-plane
//    staged in LDS (double-buffered, one barrier per plane) inside a zero halo
//    of the stencil's half-width; neighbours outside the cube read the halo's
//    exact zeros at fixed offsets, so the sum runs over the reference's
//    ascending AIJ column order with every product and sum rounded as written;
//  * nonlocal: each column then adds chi*alpha of the images covering it (a
//    register bitmask over the node's image list, built once per node), in
//    ascending image order and channel order -- after the stencil, as the
//    reference adds Mult_NonlocalVector after MatMult -- and forms its part of
//    <v,u>; chi_lm = T_l * C_lm P_lm(x,y,z) is rebuilt from the radial factor;
//    the projector sums use fused multiply-adds (their order already differs
//    from the reference's sequential scan), the stencil does not;
//  * update: w = (u - a v) - b vprev and ||w|| with the same column mapping.
// 512 threads leave each lane 256 registers, which the two columns' windows,
// the projector registers and the load batches need.
#include <type_traits>

#include "sq_device_common.h"

namespace sqamd {

// s+d channels of one (x,y) column of an image: the z-independent parts of
// C_lm P_lm computed once, so a point needs only its z terms.
template <typename T>
struct SdColumn {
    T cxy, cy, cx, q, cxxyy;  // C xy, C y, C x, -xx-yy, C22 (xx-yy)
};
template <typename T>
__device__ __forceinline__ SdColumn<T> sd_column(T x, T y)
{
    const T C = T(1.092548430592079), C22 = T(0.546274215296040);
    return {C * (x * y), C * y, C * x, -(x * x) - y * y, C22 * (x * x - y * y)};
}
// chi of channels (0,0), (2,-2), (2,-1), (2,0), (2,1), (2,2) at height z
template <typename T, int NLM>
__device__ __forceinline__ void sd_channels(const SdColumn<T> &c, T t0, T t2, T z, T out[NLM])
{
    out[0] = t0;
    out[1] = t2 * c.cxy;
    out[2] = t2 * (c.cy * z);
    out[3] = t2 * (T(0.315391565252520) * (c.q + T(2) * (z * z)));
    out[4] = t2 * (c.cx * z);
    out[5] = t2 * c.cxxyy;
#pragma unroll
    for (int ch = 6; ch < NLM; ch++) out[ch] = T(0);
}

// The FP32 form of the 512-thread shape is held to 128 VGPRs (four waves per
// SIMD), so two workgroups, two nodes, share a CU and cover each other's
// barrier and load waits (804 against 867 ms for the 16-atom pass).  The FP64
// form would spill three times as much there and keeps 256 VGPRs.
template <typename T, int NLM, int CPT, int MW, int LAYOUT, int WG>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(WG == 512 && sizeof(T) == 4 ? 4 : 1)))
lanczos_zmarch(KParams P)
{
    constexpr int NW = WG / 64;
    constexpr int O = 6;
    // z-planes per trip of the scatter (U) and gather (UG) loops: as many loads
    // in flight as the register budget allows (two waves per SIMD for the FP64
    // 512-thread shape; four for the others, where the FP32 512-thread shape
    // measured 804, 781 and 767 ms at 16, 8 and 4 planes)
    constexpr bool WL = WG == 1024;   // register-lean shape: 4 waves per SIMD
    constexpr int U = WL ? (sizeof(T) == 8 ? 4 : 8) : (sizeof(T) == 8 ? 8 : 4);
    constexpr int UG = U;
    constexpr int U0 = 4;   // the generic channel layout holds up to four radial factors per point
    constexpr int UC = 8;   // planes of a column chunk in the scatter and update
    extern __shared__ __align__(16) unsigned char smem[];
    int *img_list = reinterpret_cast<int *>(smem);
    int4 *tlo = reinterpret_cast<int4 *>(smem + P.lds_tlo);
    int4 *tdim = reinterpret_cast<int4 *>(smem + P.lds_tdim);
    int64_t *toff = reinterpret_cast<int64_t *>(smem + P.lds_toff);   // radial-factor offset per image
    double4 *tcen = reinterpret_cast<double4 *>(smem + P.lds_tcen);   // image position
    double *A = reinterpret_cast<double *>(smem + P.lds_A);
    double *alpha = reinterpret_cast<double *>(smem + P.lds_alpha);
    double *red = reinterpret_cast<double *>(smem + P.lds_red);
    T *plane = reinterpret_cast<T *>(smem + P.lds_plane);
    __shared__ int s_count;
    __shared__ T s_w[3][8];
    const T hzt = (T)P.hgrid[2];   // plane u of a trip sits at z0 + u h_z

    const int nx = P.nloc[0], ny = P.nloc[1], nz = P.nloc[2];
    const int ncol = nx * ny, nc = P.nc;
    // z-planes in LDS with a zero halo of O cells on every side in x and y
    const int px = P.plane_px, pplane = px * (ny + 2 * O);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int col[CPT], ci[CPT], cj[CPT];
    bool has[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        col[k] = tid + k * WG;
        has[k] = col[k] < ncol;
        cj[k] = has[k] ? col[k] / nx : 0;
        ci[k] = has[k] ? col[k] - cj[k] * nx : 0;
    }
    int pcol[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) pcol[k] = (cj[k] + O) * px + ci[k] + O;
    for (int q = tid; q < 2 * pplane; q += WG) plane[q] = T(0);  // halo stays zero: columns write the interior only
    T *__restrict__ buf0 = reinterpret_cast<T *>(P.scratch) + (size_t)blockIdx.x * 3 * nc;
    T *__restrict__ buf1 = buf0 + nc;
    T *__restrict__ Bu = buf1 + nc;
    const T *__restrict__ diag_ext = reinterpret_cast<const T *>(P.diag_ext);
    const T *__restrict__ rad = reinterpret_cast<const T *>(P.rad);
    // stencil weights, staged through LDS into vector registers (the scalar
    // register file is taken by the node and image bookkeeping)
    if (tid < 21) {
        if constexpr (sizeof(T) == 8) s_w[tid / 7][tid % 7] = P.cw64[tid / 7][tid % 7];
        else s_w[tid / 7][tid % 7] = P.cw32[tid / 7][tid % 7];
    }
    __syncthreads();

    // Nodes are taken in passes of gridDim.x; within a pass the workgroups of
    // one XCD (blockIdx % 8 under round-robin dispatch) take a contiguous run,
    // so the projector tables they share stay in that XCD's L2.
    // Work queue: the node range is cut into eight contiguous parts, one per
    // XCD (blockIdx % 8 under round-robin dispatch); a workgroup takes the next
    // node of its XCD's part with an atomic counter, so the projector tables its
    // neighbours share stay in that XCD's L2, and moves on to the other parts
    // when its own is done -- nodes that stop early leave no workgroup idle.
    const long part = (P.node_count + 7) / 8;
    const int home = blockIdx.x & 7;
    int q_at = 0;                       // parts tried so far (thread 0)
    __shared__ long s_next;
    __shared__ double s_ab[2][kMaxNqp + 1];
    __shared__ double s_g[kMatsubara][2];
    __shared__ int s_flag;
    __shared__ int s_near[kRecMax], s_nnear;
    // per scatter chunk of UC planes: the images whose z range meets it
    constexpr int kMaxChunks = 16;
    __shared__ unsigned s_zmask[kMaxChunks][MW];
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
        const long id = P.node_list ? (long)P.node_list[r] : P.node_begin + r;
        int node[3];
        node_coords(P, id, node);
        const int cube_lo[3] = {node[0] - P.hw[0], node[1] - P.hw[1], node[2] - P.hw[2]};
        // ---- images whose box meets the cube, ascending, with their tables in LDS
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
                      lo4.z <= cube_lo[2] + nz - 1 && lo4.z + dm4.z - 1 >= cube_lo[2];
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
            if (tid == 0) P.status[id] = 4;
            __syncthreads();
            continue;
        }
        if (P.rec) {
            if (record_near(P, r, img_list, nimg, node, s_near, &s_nnear) > P.rec_slots) {
                if (tid == 0) P.status[id] = 4;
                __syncthreads();
                continue;
            }
        }
        // ---- per column: bitmask of the node's images whose x/y box holds it
        unsigned cmask[CPT][MW];
        int gxc[CPT], gyc[CPT];
        int64_t ecol[CPT];
#pragma unroll
        for (int k = 0; k < CPT; k++) {
            gxc[k] = cube_lo[0] + ci[k];
            gyc[k] = cube_lo[1] + cj[k];
            ecol[k] = ((int64_t)(cube_lo[2] - P.ext0[2]) * P.extn[1] + (gyc[k] - P.ext0[1])) * P.extn[0] +
                      (gxc[k] - P.ext0[0]);
#pragma unroll
            for (int w = 0; w < MW; w++) cmask[k][w] = 0u;
        }
        for (int li = 0; li < nimg; li++) {
            const int4 lo = tlo[li], dm = tdim[li];
#pragma unroll
            for (int k = 0; k < CPT; k++) {
                const bool in = has[k] && gxc[k] >= lo.x && gxc[k] < lo.x + dm.x && gyc[k] >= lo.y && gyc[k] < lo.y + dm.y;
#pragma unroll
                for (int w = 0; w < MW; w++)
                    if (w == (li >> 5) && in) cmask[k][w] |= 1u << (li & 31);
            }
        }
        const int64_t eplane = (int64_t)P.extn[1] * P.extn[0];
        // ---- per scatter chunk: bitmask of the images whose z range meets it, so a
        // column's chunk visits only those (all images when the cube is too tall)
        const int nch = (nz + UC - 1) / UC;
        const bool zmasked = nch <= kMaxChunks;
        if (zmasked)
            for (int e = tid; e < nch * MW; e += WG) {
                const int c = e / MW, w = e - c * MW;
                const int z0 = c * UC, z1 = z0 + UC - 1;
                unsigned bits = 0u;
                for (int q = 0; q < 32; q++) {
                    const int li = w * 32 + q;
                    if (li >= nimg) break;
                    const int4 lo = tlo[li], dm = tdim[li];
                    const int zlo = max(lo.z, cube_lo[2]) - cube_lo[2];
                    const int zhi = min(lo.z + dm.z - 1, cube_lo[2] + nz - 1) - cube_lo[2];
                    if (zhi >= z0 && zlo <= z1) bits |= 1u << q;
                }
                s_zmask[c][w] = bits;
            }
        // ---- Lanczos start: v = e_center, vprev = 0
        for (int p = tid; p < nc; p += WG) { buf0[p] = T(0); buf1[p] = T(0); }
        __syncthreads();
        if (tid == 0) buf0[P.center] = T(1);
        __syncthreads();
        T *Bv = buf0, *Bvp = buf1;
        T sv = T(1), svp = T(0);
        double bprev = 0.0;
        double *out_a = P.out_a + id * (P.nqp + 1), *out_b = P.out_b + id * (P.nqp + 1);
        int status = 0;
        int nq = P.nqp, checks = 0;   // Gauss nodes of the node's rule; early-exit checks made
        for (int step = 0; step <= P.nqp; step++) {
            // ---- gather: A[li][ch] = sum_{box & cube} v chi
            for (int li = wave; li < nimg; li += NW) {
                const int4 lo = tlo[li], dm = tdim[li];
                const int nlm = dm.w & 255, nl = (dm.w >> 8) & 255, type = dm.w >> 16;
                const double4 cen = tcen[li];
                int a0[3], n[3];
                const int ilo[3] = {lo.x, lo.y, lo.z}, idm[3] = {dm.x, dm.y, dm.z};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int s0 = max(ilo[d], cube_lo[d]);
                    const int e0 = min(ilo[d] + idm[d] - 1, cube_lo[d] + P.nloc[d] - 1);
                    a0[d] = s0;
                    n[d] = e0 - s0 + 1;
                }
                const int nxy = n[0] * n[1];
                const float inv0 = 1.0f / (float)n[0];
                const int zs = dm.x * dm.y * nl;
                T acc[NLM];
#pragma unroll
                for (int ch = 0; ch < NLM; ch++) acc[ch] = T(0);
                for (int q = lane; q < nxy; q += 64) {
                    const int yi = fdiv(q, n[0], inv0);
                    const int xi = q - yi * n[0];
                    const int gx = a0[0] + xi, gy = a0[1] + yi;
                    const int p0 = ((a0[2] - cube_lo[2]) * ny + (gy - cube_lo[1])) * nx + (gx - cube_lo[0]);
                    const int64_t rb = toff[li] + ((int64_t)((a0[2] - lo.z) * dm.y + (gy - lo.y)) * dm.x + (gx - lo.x)) * nl;
                    const T x = (T)(P.hgrid[0] * gx - P.range[0] - cen.x);
                    const T y = (T)(P.hgrid[1] * gy - P.range[1] - cen.y);
                    const T *vb = Bv + p0;
                    if constexpr (LAYOUT == 1) {
                        // Two radial factors per point, one 16-byte load.  Along the
                        // column the d channels are t2 * (c0 + c1 z + c2 z^2) with
                        // column constants, so the z loop keeps four sums and the
                        // six channels are formed once per column.
                        using V2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
                        const V2 *r2 = reinterpret_cast<const V2 *>(rad + rb);
                        const int zs2 = dm.x * dm.y;
                        T s0 = T(0), st = T(0), s1 = T(0), s2 = T(0);
                        // whole trips, the last one masked: every load of a trip is
                        // in flight together
#pragma unroll 1
                        for (int zi = 0; zi < n[2]; zi += UG) {
                            T v4[UG];
                            V2 t4[UG];
#pragma unroll
                            for (int u = 0; u < UG; u++) {
                                const bool ok = zi + u < n[2];
                                v4[u] = ok ? vb[(zi + u) * ncol] : T(0);
                                t4[u] = ok ? r2[(zi + u) * zs2] : V2{T(0), T(0)};
                            }
                            const T zb = (T)(P.hgrid[2] * (a0[2] + zi) - P.range[2] - cen.z);
#pragma unroll
                            for (int u = 0; u < UG; u++) {
                                const T z = fma(hzt, T(u), zb);
                                s0 = fma(v4[u], t4[u].x, s0);
                                const T vt = v4[u] * t4[u].y;
                                st = st + vt;
                                const T tz = vt * z;
                                s1 = s1 + tz;
                                s2 = fma(tz, z, s2);
                            }
                        }
                        const SdColumn<T> pz = sd_column(x, y);
                        acc[0] = acc[0] + s0;
                        acc[1] = fma(pz.cxy, st, acc[1]);
                        acc[2] = fma(pz.cy, s1, acc[2]);
                        acc[3] = fma(T(0.315391565252520), fma(pz.q, st, T(2) * s2), acc[3]);
                        acc[4] = fma(pz.cx, s1, acc[4]);
                        acc[5] = fma(pz.cxxyy, st, acc[5]);
                    } else {
                        const T *r0 = rad + rb;
                        for (int zi = 0; zi < n[2]; zi += U0) {
                            T v4[U0], t4[U0][4];
#pragma unroll
                            for (int u = 0; u < U0; u++) {
                                const bool ok = zi + u < n[2];
                                v4[u] = ok ? vb[(zi + u) * ncol] : T(0);
#pragma unroll
                                for (int k = 0; k < 4; k++) t4[u][k] = (ok && k < nl) ? r0[(int64_t)(zi + u) * zs + k] : T(0);
                            }
#pragma unroll
                            for (int u = 0; u < U0; u++) {
                                if (zi + u < n[2]) {
                                    const T z = (T)(P.hgrid[2] * (a0[2] + zi + u) - P.range[2] - cen.z);
                                    T c[NLM];
                                    chi_channels<T, NLM, LAYOUT>(P, type, nlm, t4[u], x, y, z, c);
                                    const T v = sv * v4[u];
#pragma unroll
                                    for (int ch = 0; ch < NLM; ch++) acc[ch] = fma(v, c[ch], acc[ch]);
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int ch = 0; ch < NLM; ch++) {
                    double sm = wave_sum((double)acc[ch]);
                    if constexpr (LAYOUT == 1) sm *= (double)sv;   // the sums used the stored (unscaled) v
                    if (lane == 0 && ch < nlm) A[li * NLM + ch] = sm;
                }
            }
            __syncthreads();
            images_alpha<NLM>(P, img_list, nimg, A, alpha);
            __syncthreads();
            if (P.rec) record_step(P, r, step, s_near, s_nnear, alpha, NLM);
            // ---- nonlocal part of <v, H v> from the gathered sums: the scatter adds
            // sum_ch chi_ch alpha_ch per image, so <v, V_nl v> = sum alpha . A
            double pn = 0.0;
            for (int e = tid; e < nimg * NLM; e += WG) {
                const int li = e / NLM, ch = e - li * NLM;
                if (ch < (tdim[li].w & 255)) pn += alpha[e] * A[e];
            }
            // ---- local part, marching along z: u = (L + diag) v into Bu and <v, u>.
            // The z loop is unrolled by the window length so the window is a ring
            // with fixed register slots: plane z + d sits in slot (z + d + O) % W.
            constexpr int W = 2 * O + 1;
            // stencil weights in vector registers for the march only (read after
            // the barriers above, so they are not held through the projector loops)
            // (the 1024-thread shape has 128 registers a lane: it reads them from LDS)
            T wreg[3][O + 1];
            if constexpr (!WL) {
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int l = 1; l <= O; l++) wreg[d][l] = s_w[d][l];
            }
            auto wgt = [&](int d, int l) -> T {
                if constexpr (WL) return s_w[d][l];
                else return wreg[d][l];
            };
            double pa = 0.0;
            CSum ca{0.f, 0.f};
            T win[CPT][W];
#pragma unroll
            for (int k = 0; k < CPT; k++)
#pragma unroll
                for (int q = 0; q < W; q++) {
                    const int z = q - O;   // slot q holds plane q - O before the first step
                    win[k][q] = (has[k] && z >= 0 && z < nz) ? sv * Bv[col[k] + z * ncol] : T(0);
                }
            T dgn[CPT];
#pragma unroll
            for (int k = 0; k < CPT; k++) dgn[k] = has[k] ? diag_ext[ecol[k]] : T(0);
            for (int zb = 0; zb < nz; zb += W) {
#pragma unroll
                for (int s = 0; s < W; s++) {
                    const int z = zb + s;
                    if (z < nz) {   // uniform: the last block is partial
                    T *pl = plane + (z & 1) * pplane;
                    T next[CPT], dg[CPT];
#pragma unroll
                    for (int k = 0; k < CPT; k++) {
                        next[k] = T(0);
                        dg[k] = dgn[k];
                        if (has[k]) {
                            pl[pcol[k]] = win[k][(s + O) % W];
                            if (z + O + 1 < nz) next[k] = sv * Bv[col[k] + (z + O + 1) * ncol];
                            if (z + 1 < nz) dgn[k] = diag_ext[ecol[k] + (z + 1) * eplane];
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < CPT; k++) {
                        if (has[k]) {
                            const T *pc = pl + pcol[k];
                            // three independent fma chains; the plane's zero halo
                            // supplies the out-of-cube neighbours
                            T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
                            for (int l = 1; l <= O; l++) {
                                sz = fma(wgt(2, l), win[k][(s + O - l) % W], sz);
                                sy = fma(wgt(1, l), pc[-l * px], sy);
                                sx = fma(wgt(0, l), pc[-l], sx);
                            }
#pragma unroll
                            for (int l = 1; l <= O; l++) {
                                sz = fma(wgt(2, l), win[k][(s + O + l) % W], sz);
                                sy = fma(wgt(1, l), pc[l * px], sy);
                                sx = fma(wgt(0, l), pc[l], sx);
                            }
                            const T vc = win[k][(s + O) % W];
                            const T u = fma(dg[k], vc, sx) + (sy + sz);
                            Bu[col[k] + z * ncol] = u;
                            if constexpr (sizeof(T) == 8) pa = fma(vc, u, pa);
                            else ca = csum_add_prod(ca, vc, u);
                        }
                        // plane z - O leaves the window; plane z + O + 1 takes its slot
                        win[k][s % W] = next[k];
                    }
                    }
                }
            }
            double a;
            if constexpr (sizeof(T) == 8) a = block_sum<NW>(pa + pn, red);
            else a = block_csum<NW>(ca, red) + block_sum<NW>(pn, red);
            // ---- per column, in chunks of UC planes held in registers: the local
            // part of u from the march, plus chi*alpha of every image of the column
            // that meets the chunk (ascending images), then w = (u - a v) - bprev
            // vprev over vprev's storage and ||w||^2.  u is never written back.
            const T ta = T(-a), tb = T(-bprev);
            double pb = 0.0;
            CSum cb{0.f, 0.f};
#pragma unroll
            for (int k = 0; k < CPT; k++) {
                if (!has[k]) continue;
                const int c = col[k];
                const T x = (T)(P.hgrid[0] * gxc[k] - P.range[0]);   // minus the image centre below
                const T y = (T)(P.hgrid[1] * gyc[k] - P.range[1]);
#pragma unroll 1
                for (int z0 = 0; z0 < nz; z0 += UC) {
                    T uu[UC];
#pragma unroll
                    for (int u = 0; u < UC; u++) uu[u] = (z0 + u < nz) ? Bu[c + (z0 + u) * ncol] : T(0);
#pragma unroll 1
                    for (int wd = 0; wd < MW; wd++) {
                        // mask word wd by a select chain: no dynamic register indexing
                        unsigned m = cmask[k][0];
#pragma unroll
                        for (int q = 1; q < MW; q++) m = (wd == q) ? cmask[k][q] : m;
                        if (zmasked) m &= s_zmask[z0 / UC][wd];
                        while (m) {
                            const int li = (wd << 5) + __ffs(m) - 1;
                            m &= m - 1u;
                            const int4 lo = tlo[li], dm = tdim[li];
                            const int zlo = max(lo.z, cube_lo[2]) - cube_lo[2];
                            const int zhi = min(lo.z + dm.z - 1, cube_lo[2] + nz - 1) - cube_lo[2];
                            if (zhi < z0 || zlo > z0 + UC - 1) continue;
                            const int nlm = dm.w & 255, nl = (dm.w >> 8) & 255, type = dm.w >> 16;
                            const double4 cen = tcen[li];
                            // radial factors of this column at plane z0 (entries outside the
                            // image's z range are never loaded)
                            const int64_t rb0 = toff[li] +
                                                ((int64_t)((cube_lo[2] + z0 - lo.z) * dm.y + (gyc[k] - lo.y)) * dm.x +
                                                 (gxc[k] - lo.x)) * nl;
                            const T xc = x - (T)cen.x, yc = y - (T)cen.y;
                            T al[NLM];
#pragma unroll
                            for (int ch = 0; ch < NLM; ch++) al[ch] = ch < nlm ? (T)alpha[li * NLM + ch] : T(0);
                            const T zb = (T)(P.hgrid[2] * (cube_lo[2] + z0) - P.range[2] - cen.z);
                            if constexpr (LAYOUT == 1) {
                                // chi . alpha along the column = t0 al0 + t2 (p0 + p1 z + p2 z^2)
                                using V2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
                                const SdColumn<T> pz = sd_column(xc, yc);
                                const T C20 = T(0.315391565252520);
                                const T p0 = fma(pz.cxy, al[1], fma(pz.cxxyy, al[5], C20 * (pz.q * al[3])));
                                const T p1 = fma(pz.cy, al[2], pz.cx * al[4]);
                                const T p2 = T(2) * C20 * al[3];
                                const V2 *r2 = reinterpret_cast<const V2 *>(rad + rb0);
                                const int zs2 = dm.x * dm.y;
                                V2 t4[UC];
#pragma unroll
                                for (int u = 0; u < UC; u++) {
                                    const bool ok = z0 + u >= zlo && z0 + u <= zhi;
                                    t4[u] = ok ? r2[u * zs2] : V2{T(0), T(0)};
                                }
#pragma unroll
                                for (int u = 0; u < UC; u++) {
                                    const T z = fma(hzt, T(u), zb);
                                    const T h = fma(z, fma(z, p2, p1), p0);
                                    if (z0 + u >= zlo && z0 + u <= zhi) uu[u] = fma(t4[u].y, h, fma(t4[u].x, al[0], uu[u]));
                                }
                            } else {
                                const int64_t zs = (int64_t)dm.x * dm.y * nl;
#pragma unroll 1
                                for (int u = 0; u < UC; u++) {
                                    if (z0 + u < zlo || z0 + u > zhi) continue;
                                    T tt[4];
#pragma unroll
                                    for (int q = 0; q < 4; q++) tt[q] = q < nl ? rad[rb0 + u * zs + q] : T(0);
                                    const T zc = fma(hzt, T(u), zb);
                                    T cv[NLM];
                                    chi_channels<T, NLM, LAYOUT>(P, type, nlm, tt, xc, yc, zc, cv);
                                    T sum = uu[0];
#pragma unroll
                                    for (int q = 0; q < UC; q++) sum = (q == u) ? uu[q] : sum;
#pragma unroll
                                    for (int ch = 0; ch < NLM; ch++)
                                        if (ch < nlm) sum = fma(cv[ch], al[ch], sum);
#pragma unroll
                                    for (int q = 0; q < UC; q++) uu[q] = (q == u) ? sum : uu[q];
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UC; u++) {
                        if (z0 + u >= nz) break;
                        const int p = c + (z0 + u) * ncol;
                        const T vc = sv * Bv[p];
                        const T vpc = svp * Bvp[p];
                        T w;
                        if constexpr (sizeof(T) == 8) {
                            w = fma(ta, vc, uu[u]);
                            w = fma(tb, vpc, w);
                            pb = fma(w, w, pb);
                        } else {
                            w = fmaf(ta, vc, uu[u]);
                            w = fmaf(tb, vpc, w);
                            cb = csum_add_prod(cb, w, w);
                        }
                        Bvp[p] = w;
                    }
                }
            }
            double bb;
            if constexpr (sizeof(T) == 8) bb = block_sum<NW>(pb, red);
            else bb = block_csum<NW>(cb, red);
            const double b = sqrt(bb);
            if (tid == 0) { out_a[step] = a; out_b[step] = b; }
            if (b == 0.0) { status = 7; break; }
            if (P.exit_tol > 0.0) {
                // Early exit: the node's Gauss rule of n nodes is J_n (a_0..a_{n-1},
                // b_0..b_{n-2}); its resolvent g(z) = e1^T (z - J_n)^{-1} e1 is the
                // continued fraction 1/(z - a_0 - b_0^2/(z - a_1 - ...)).  The
                // Fermi quadrature is 1/2 + (2/beta) sum_m Re g(lambda_f + i w_m)
                // with w_m = (2m+1) pi / beta; the lowest terms converge last.
                if (tid == 0) { s_ab[0][step] = a; s_ab[1][step] = b; }
                const int n = step + 1;
                if (n >= P.exit_min && (n - P.exit_min) % P.exit_every == 0 && n < P.nqp) {
                    if (tid == 0) s_flag = 0;
                    __syncthreads();
                    if (tid < kMatsubara) {
                        const double wm = (2 * tid + 1) * M_PI / P.exit_beta;
                        double tr = P.exit_lf - s_ab[0][n - 1], ti = wm;
                        for (int j = n - 2; j >= 0; j--) {
                            const double den = tr * tr + ti * ti;   // >= wm^2 > 0
                            const double bb = s_ab[1][j] * s_ab[1][j];
                            tr = (P.exit_lf - s_ab[0][j]) - bb * tr / den;
                            ti = wm + bb * ti / den;
                        }
                        const double den = tr * tr + ti * ti;
                        const double gr = tr / den, gi = -ti / den;
                        const double dr = gr - s_g[tid][0], di = gi - s_g[tid][1];
                        const double d = (2.0 / P.exit_beta) * sqrt(dr * dr + di * di);
                        if (checks == 0 || !(d < P.exit_tol)) s_flag = 1;
                        s_g[tid][0] = gr;
                        s_g[tid][1] = gi;
                    }
                    __syncthreads();
                    const bool stop = s_flag == 0;
                    checks++;
                    __syncthreads();
                    if (stop) { nq = n; break; }
                }
            }
            T *t = Bv; Bv = Bvp; Bvp = t;
            svp = sv;
            sv = T(1.0 / b);
            bprev = b;
        }
        if (tid == 0) {
            P.status[id] = status;
            if (P.nsteps) P.nsteps[id] = nq;
        }
        __syncthreads();
    }
}

template <typename T, int NLM, int CPT, int LAYOUT, int WG>
static hipError_t launch_z(const KParams &P, int grid, size_t lds, hipStream_t st)
{
    if (P.img_cap <= 128) hipLaunchKernelGGL((lanczos_zmarch<T, NLM, CPT, 4, LAYOUT, WG>), dim3(grid), dim3(WG), lds, st, P);
    else if (P.img_cap <= 256) hipLaunchKernelGGL((lanczos_zmarch<T, NLM, CPT, 8, LAYOUT, WG>), dim3(grid), dim3(WG), lds, st, P);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T, int NLM, int CPT, int LAYOUT, int WG>
static int occ_z(const KParams &P, size_t lds)
{
    int n = 0;
    hipError_t e = P.img_cap <= 128
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lanczos_zmarch<T, NLM, CPT, 4, LAYOUT, WG>, WG, lds)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lanczos_zmarch<T, NLM, CPT, 8, LAYOUT, WG>, WG, lds);
    return e == hipSuccess ? n : 0;
}

// Calls f.template run<T, NLM, CPT, LAYOUT, WG>() for the precision, the
// channel table, the workgroup shape (P.zwg threads; columns per thread from
// it) and the channel layout of P.
template <typename R, typename F>
static R dispatch(const KParams &P, int precision, F &&f)
{
    const bool wide = P.zwg == 1024;                  // one column per thread
    const bool two = !wide && P.nloc[0] * P.nloc[1] > 512;
    const bool sd = P.layout == 1 && P.nlm_max <= 6;
#define SQ_Z_SHAPES(T, NLM, L)                                              \
    (wide ? f.template run<T, NLM, 1, L, 1024>()                              \
          : two ? f.template run<T, NLM, 2, L, 512>() : f.template run<T, NLM, 1, L, 512>())
    if (precision == 0) {
        if (sd) return SQ_Z_SHAPES(double, 6, 1);
        if (P.nlm_max <= 6) return SQ_Z_SHAPES(double, 6, 0);
        return SQ_Z_SHAPES(double, kMaxChannels, 0);
    }
    if (sd) return SQ_Z_SHAPES(float, 6, 1);
    if (P.nlm_max <= 6) return SQ_Z_SHAPES(float, 6, 0);
    return SQ_Z_SHAPES(float, kMaxChannels, 0);
#undef SQ_Z_SHAPES
}

struct LaunchOp {
    const KParams &P;
    int grid;
    size_t lds;
    hipStream_t st;
    template <typename T, int NLM, int CPT, int LAYOUT, int WG>
    hipError_t run() { return launch_z<T, NLM, CPT, LAYOUT, WG>(P, grid, lds, st); }
};

struct OccOp {
    const KParams &P;
    size_t lds;
    template <typename T, int NLM, int CPT, int LAYOUT, int WG>
    int run() { return occ_z<T, NLM, CPT, LAYOUT, WG>(P, lds); }
};

hipError_t launch_zmarch(const KParams &P, int precision, int grid, size_t lds, hipStream_t st)
{
    if (P.nloc[0] * P.nloc[1] > 1024) return hipErrorInvalidValue;
    return dispatch<hipError_t>(P, precision, LaunchOp{P, grid, lds, st});
}


int zmarch_occupancy(const KParams &P, int precision, size_t lds)
{
    if (P.nloc[0] * P.nloc[1] > 1024) return 0;
    return dispatch<int>(P, precision, OccOp{P, lds});
}

}  // namespace sqamd
