// sq_api.cpp -- the C ABI of include/sq_amd.h: context creation (host setup +
// upload), the batched nodes-and-weights run, single-node H application,
// batched QL, and the density / Fermi level / energy reduction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "../../include/sq_amd.h"
#include "sq_kernels.h"
#include "sq_projectors.h"
#include "sq_system.h"

using namespace sqamd;

namespace {
thread_local std::string g_err;

int set_err(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_err(SQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

constexpr int kRedBlocks = 512;
}  // namespace

namespace sqamd {
// Error entry for the reference-named host layer (sq_lsdft.cpp): same message slot.
int record_error(int code, const std::string &msg) { return set_err(code, msg); }
}  // namespace sqamd

struct sq_ctx {
    System sys;
    int device = 0, n_cu = 0, grid = 0, img_cap = 0, nlm_tmpl = 6;
    size_t lds = 0;
    int lds_A = 0, lds_alpha = 0, lds_red = 0, lds_local = 0;
    // z-march kernel: eligibility, column-list length, LDS layout, grid per precision
    bool zmarch = false;
    int kmax = 0, zgrid[2] = {0, 0};
    int zwg[2] = {512, 512};      // z-march workgroup size per precision (512 or 1024 threads)
    float tune_ms[2][2] = {};     // autotuner probe time per precision and shape (0: 512, 1: 1024)
    size_t zlds = 0;
    int z_tlo = 0, z_tdim = 0, z_toff = 0, z_tcen = 0, z_A = 0, z_alpha = 0, z_red = 0, z_plane = 0;
    int plane_px = 0;   // LDS row pitch of the z-march kernels' planes
    // batched kernel (sq_zbatch.hip): B = 2 or 4 nodes stacked in z per workgroup.
    // Per B: images per union cube, LDS size and offsets (tlo, tdim, toff, tcen,
    // A, alpha, red, local, plane), resident grid.
    bool zbatch = false;
    int zb_cap[5] = {}, zb_grid[2][5] = {};   // grid per precision and B
    size_t zb_lds[5] = {};
    int zb_off[5][11] = {};
    // the configuration the autotuner chose per precision: nodes per workgroup
    // (1: the one-node z-march kernel in its zwg shape) and gather trip
    int cfg_B[2] = {1, 1}, cfg_ug[2] = {0, 0};
    float tune_cand[2][6] = {};   // probe ms: one-node 512 / 1024 threads, B=2 trips a/b, B=4 trips a/b
    // batch list of the last node range
    int2 *batches = nullptr;
    long batch_cap = 0, n_batches = 0, bat_begin = -1, bat_count = -1;
    int bat_B = 0;
    int *probe_ids = nullptr;     // the autotuner's node set (one-node kernel form)
    float *bas32 = nullptr;       // column basis table of the batched kernel's FP32 form
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // device tables
    double *chi64 = nullptr;
    float *chi32 = nullptr;
    int4 *img_lo = nullptr, *img_dim = nullptr, *img_dimp = nullptr;
    int64_t *img_off = nullptr;
    int *img_atom = nullptr, *img_type = nullptr;
    double *gamma = nullptr;
    int *cov_off = nullptr, *cov_img = nullptr;
    double *diag64 = nullptr;
    float *diag32 = nullptr;
    double *rad64 = nullptr;
    float *rad32 = nullptr;
    int64_t *img_rad = nullptr;
    double4 *img_cen = nullptr;
    double *c64 = nullptr;  // 3 x (kMaxOrder+1)
    float *c32 = nullptr;
    void *scratch = nullptr;
    double *a = nullptr, *b = nullptr, *lam = nullptr, *w = nullptr, *lmin = nullptr, *lmax = nullptr, *qlws = nullptr;
    int *status = nullptr;
    int *queue = nullptr, *nsteps = nullptr;   // z-march work queue (8 counters); Gauss nodes per node
    double *partial = nullptr, *partial2 = nullptr, *out = nullptr;
    // per-node early exit of the z-march kernel (exit_tol <= 0: off)
    double exit_tol = 0, exit_lf = 0, exit_beta = 0;
    int exit_min = 30, exit_every = 5;
    bool last_exit = false;   // the last sq_run used it
    // last run
    long last_begin = 0, last_count = 0;
    bool have_run = false;
    double t_lanczos = 0, t_ql = 0;
    // SCF state on the periodic cell (allocated by the first SCF call)
    struct Scf {
        long ng = 0;
        int hist = 0;
        double *rho = nullptr, *rho_out = nullptr, *b = nullptr, *phi = nullptr, *eps = nullptr, *vxc = nullptr,
               *veff = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *f = nullptr, *x_prev = nullptr,
               *f_prev = nullptr, *dX = nullptr, *dF = nullptr, *g = nullptr, *s = nullptr, *part = nullptr,
               *ku = nullptr, *ky = nullptr;   // Kerker: averaged residual, its preconditioned form
        bool have_b = false;
        double eself = 0, b_scale = 0;
        std::vector<double> host_b;
        // four CG steps of the Poisson solve, captured once (the buffers above are fixed)
        hipGraph_t cg_graph = nullptr;
        hipGraphExec_t cg_exec = nullptr;
        // the device-side control block of an SCF step, its timing events and the
        // step captured as one graph (per sq_scf call)
        ScfCtl *ctl = nullptr;
        hipEvent_t sev[4] = {nullptr, nullptr, nullptr, nullptr};
        hipGraph_t step_graph = nullptr;
        hipGraphExec_t step_exec = nullptr;
        // the last SQ pass of the loop, which sq_forces repeats: its input
        // density's V_eff (left in the diagonal), phi, Fermi level and settings
        bool force_ready = false;
        double lambda_f = 0;
        int last_mode = 1, last_prec = 0;
        double last_exit_tol = 0, last_exit_lf = 0, last_exit_beta = 0;
        int last_exit_min = 0, last_exit_every = 0;
    } scf;
};

namespace {

template <typename T>
int dev_upload(T **dst, const T *src, size_t n)
{
    HIP_TRY(hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return SQ_OK;
}

void free_ctx(sq_ctx *c)
{
    if (!c) return;
    void *ptrs[] = {c->img_dimp, c->rad64, c->rad32, c->img_rad, c->img_cen, c->chi64, c->chi32, c->img_lo, c->img_dim, c->img_off, c->img_atom, c->img_type, c->gamma,
                    c->cov_off, c->cov_img, c->diag64, c->diag32, c->c64, c->c32, c->scratch, c->a, c->b,
                    c->lam, c->w, c->lmin, c->lmax, c->qlws, c->status, c->partial, c->partial2, c->out,
                    c->queue, c->nsteps, c->batches, c->probe_ids, c->bas32};
    void *scf_ptrs[] = {c->scf.rho, c->scf.rho_out, c->scf.b, c->scf.phi, c->scf.eps, c->scf.vxc, c->scf.veff, c->scf.r,
                        c->scf.p, c->scf.Ap, c->scf.f, c->scf.x_prev, c->scf.f_prev, c->scf.dX, c->scf.dF, c->scf.g,
                        c->scf.s, c->scf.part, c->scf.ku, c->scf.ky};
    if (c->scf.cg_exec) (void)hipGraphExecDestroy(c->scf.cg_exec);
    if (c->scf.cg_graph) (void)hipGraphDestroy(c->scf.cg_graph);
    if (c->scf.step_exec) (void)hipGraphExecDestroy(c->scf.step_exec);
    if (c->scf.step_graph) (void)hipGraphDestroy(c->scf.step_graph);
    if (c->scf.ctl) (void)hipFree(c->scf.ctl);
    for (hipEvent_t e : c->scf.sev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : scf_ptrs)
        if (p) (void)hipFree(p);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Largest number of images any cube of the rank box meets: a 3-D difference
// array over nodes, one box per image expanded by the cube half-width.
int max_images_per_cube(const System &S)
{
    const int lx = S.box[3], ly = S.box[4], lz = S.box[5];
    std::vector<int> diff((size_t)(lx + 1) * (ly + 1) * (lz + 1), 0);
    auto at = [&](int i, int j, int k) -> int & { return diff[((size_t)k * (ly + 1) + j) * (lx + 1) + i]; };
    for (const Image &im : S.images) {
        int lo[3], hi[3];
        bool empty = false;
        for (int d = 0; d < 3; d++) {
            lo[d] = std::max(im.lo[d] - S.hw[d], S.box[d]) - S.box[d];
            hi[d] = std::min(im.hi[d] + S.hw[d], S.box[d] + S.box[3 + d] - 1) - S.box[d];
            if (hi[d] < lo[d]) empty = true;
        }
        if (empty) continue;
        for (int s = 0; s < 8; s++) {
            int i = (s & 1) ? hi[0] + 1 : lo[0], j = (s & 2) ? hi[1] + 1 : lo[1], k = (s & 4) ? hi[2] + 1 : lo[2];
            int sign = ((s & 1) ? -1 : 1) * ((s & 2) ? -1 : 1) * ((s & 4) ? -1 : 1);
            at(i, j, k) += sign;
        }
    }
    for (int k = 0; k <= lz; k++)
        for (int j = 0; j <= ly; j++)
            for (int i = 1; i <= lx; i++) at(i, j, k) += at(i - 1, j, k);
    for (int k = 0; k <= lz; k++)
        for (int j = 1; j <= ly; j++)
            for (int i = 0; i <= lx; i++) at(i, j, k) += at(i, j - 1, k);
    for (int k = 1; k <= lz; k++)
        for (int j = 0; j <= ly; j++)
            for (int i = 0; i <= lx; i++) at(i, j, k) += at(i, j, k - 1);
    int mx = 0;
    for (int k = 0; k < lz; k++)
        for (int j = 0; j < ly; j++)
            for (int i = 0; i < lx; i++) mx = std::max(mx, at(i, j, k));
    return mx;
}

// Largest number of images any union cube of a batch of B nodes stacked in z
// (lowest node k: planes k - hw .. k + hw + B - 1) meets.
int max_images_per_union(const System &S, int B)
{
    const int lx = S.box[3], ly = S.box[4], lz = S.box[5];
    std::vector<int> diff((size_t)(lx + 1) * (ly + 1) * (lz + 1), 0);
    auto at = [&](int i, int j, int k) -> int & { return diff[((size_t)k * (ly + 1) + j) * (lx + 1) + i]; };
    for (const Image &im : S.images) {
        int lo[3], hi[3];
        bool empty = false;
        for (int d = 0; d < 3; d++) {
            const int down = d == 2 ? S.hw[d] + B - 1 : S.hw[d];
            lo[d] = std::max(im.lo[d] - down, S.box[d]) - S.box[d];
            hi[d] = std::min(im.hi[d] + S.hw[d], S.box[d] + S.box[3 + d] - 1) - S.box[d];
            if (hi[d] < lo[d]) empty = true;
        }
        if (empty) continue;
        for (int s = 0; s < 8; s++) {
            int i = (s & 1) ? hi[0] + 1 : lo[0], j = (s & 2) ? hi[1] + 1 : lo[1], k = (s & 4) ? hi[2] + 1 : lo[2];
            int sign = ((s & 1) ? -1 : 1) * ((s & 2) ? -1 : 1) * ((s & 4) ? -1 : 1);
            at(i, j, k) += sign;
        }
    }
    for (int k = 0; k <= lz; k++)
        for (int j = 0; j <= ly; j++)
            for (int i = 1; i <= lx; i++) at(i, j, k) += at(i - 1, j, k);
    for (int k = 0; k <= lz; k++)
        for (int j = 1; j <= ly; j++)
            for (int i = 0; i <= lx; i++) at(i, j, k) += at(i, j - 1, k);
    for (int k = 1; k <= lz; k++)
        for (int j = 0; j <= ly; j++)
            for (int i = 0; i <= lx; i++) at(i, j, k) += at(i, j, k - 1);
    int mx = 0;
    for (int k = 0; k < lz; k++)
        for (int j = 0; j < ly; j++)
            for (int i = 0; i < lx; i++) mx = std::max(mx, at(i, j, k));
    return mx;
}

// Batches of up to B nodes (i, j, k .. k+nb-1) covering the rank-box ids
// [begin, begin + count): for fixed (i, j) the ids of a range are a run of
// consecutive k, so each batch is {lowest id, nb}.  Blocks of B planes from
// the range's first plane, (i, j) fastest within a block.
void build_batches(const System &S, long begin, long count, int B, std::vector<int2> &out)
{
    out.clear();
    const long plane = (long)S.box[3] * S.box[4], end = begin + count;
    const long k_lo = begin / plane, k_hi = (end - 1) / plane;
    for (long kb = k_lo; kb <= k_hi; kb += B)
        for (long ij = 0; ij < plane; ij++) {
            long first = -1;
            int nb = 0;
            for (long k = kb; k < std::min(kb + B, k_hi + 1); k++) {
                const long id = k * plane + ij;
                if (id >= begin && id < end) {
                    if (first < 0) first = id;
                    nb++;
                }
            }
            if (nb) out.push_back(make_int2((int)first, nb));
        }
}

// Largest number of images any (x,y) column of any cube meets: per image, the
// columns of its x/y box, for every node z whose cube z-range meets its box.
int max_images_per_column(const System &S)
{
    const int ex = S.extn[0], ey = S.extn[1], lz = S.box[5];
    std::vector<int> cnt((size_t)ex * ey * lz, 0);
    int mx = 0;
    for (const Image &im : S.images) {
        const int k0 = std::max(im.lo[2] - S.hw[2], S.box[2]), k1 = std::min(im.hi[2] + S.hw[2], S.box[2] + lz - 1);
        for (int k = k0; k <= k1; k++)
            for (int j = im.lo[1]; j <= im.hi[1]; j++)
                for (int i = im.lo[0]; i <= im.hi[0]; i++) {
                    int &v = cnt[((size_t)(k - S.box[2]) * ey + (j - S.ext0[1])) * ex + (i - S.ext0[0])];
                    mx = std::max(mx, ++v);
                }
    }
    return mx;
}

int upload_diag(sq_ctx *c, const double *field, double veff_const)
{
    const System &S = c->sys;
    const int64_t ne = S.ext_points();
    std::vector<double> d64(ne);
    std::vector<float> d32(ne);
    for (int k = 0; k < S.extn[2]; k++) {
        int K = (((S.ext0[2] + k) % S.npts[2]) + S.npts[2]) % S.npts[2];
        for (int j = 0; j < S.extn[1]; j++) {
            int J = (((S.ext0[1] + j) % S.npts[1]) + S.npts[1]) % S.npts[1];
            for (int i = 0; i < S.extn[0]; i++) {
                int I = (((S.ext0[0] + i) % S.npts[0]) + S.npts[0]) % S.npts[0];
                double v = field ? field[((size_t)K * S.npts[1] + J) * S.npts[0] + I] : veff_const;
                size_t e = ((size_t)k * S.extn[1] + j) * S.extn[0] + i;
                d64[e] = S.cdiag + v;  // MatDiagonalSet(ADD_VALUES): diag + V_eff
                d32[e] = (float)d64[e];
            }
        }
    }
    if (!c->diag64) {
        HIP_TRY(hipMalloc((void **)&c->diag64, ne * sizeof(double)));
        HIP_TRY(hipMalloc((void **)&c->diag32, ne * sizeof(float)));
    }
    HIP_TRY(hipMemcpy(c->diag64, d64.data(), ne * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->diag32, d32.data(), ne * sizeof(float), hipMemcpyHostToDevice));
    return SQ_OK;
}

int create_device(sq_ctx *c, int device)
{
    System &S = c->sys;
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
    // the static workspace of sq_tridiag_nodes_weights_noalloc on this device
    if (noalloc_scratch_init(device) != hipSuccess) return set_err(SQ_ERR_DEVICE, "static workspace symbol not found");
    // projector tables
    const int M = (int)S.images.size();
    std::vector<int4> lo(M), dim(M);
    std::vector<int64_t> off(M);
    std::vector<int> atom(M), type(M);
    for (int g = 0; g < M; g++) {
        const Image &im = S.images[g];
        lo[g] = make_int4(im.lo[0], im.lo[1], im.lo[2], 0);
        dim[g] = make_int4(im.dim(0), im.dim(1), im.dim(2), im.nlm);
        off[g] = im.chi_off;
        atom[g] = im.atom;
        type[g] = im.type;
    }
    int rc;
    if ((rc = dev_upload(&c->img_lo, lo.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_dim, dim.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_off, off.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_atom, atom.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_type, type.data(), M))) return rc;
    if (S.chi.empty() && S.chi_total > 0) {
        // projector tables built on the device (bit-identical to the host fill)
        HIP_TRY(hipMalloc((void **)&c->chi64, S.chi_total * sizeof(double)));
        HIP_TRY(hipMalloc((void **)&c->rad64, std::max<int64_t>(1, S.rad_total) * sizeof(double)));
        HIP_TRY(hipMalloc((void **)&c->chi32, S.chi_total * sizeof(float)));
        HIP_TRY(hipMalloc((void **)&c->rad32, std::max<int64_t>(1, S.rad_total) * sizeof(float)));
        int *dst = nullptr, hst = 0;
        HIP_TRY(hipMalloc((void **)&dst, sizeof(int)));
        HIP_TRY(hipMemset(dst, 0, sizeof(int)));
        const int brc = build_projector_tables(S, c->chi64, c->rad64, dst, c->stream);
        hipError_t e = hipMemcpy(&hst, dst, sizeof(int), hipMemcpyDeviceToHost);
        (void)hipFree(dst);
        if (brc || e != hipSuccess) return set_err(SQ_ERR_DEVICE, "device projector tables failed");
        if (hst) return set_err(SQ_ERR_RANGE, "projector evaluation failed (radius off the table or bad l)");
        HIP_TRY(launch_to_float(c->chi64, c->chi32, S.chi_total, c->stream));
        HIP_TRY(launch_to_float(c->rad64, c->rad32, S.rad_total, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        if ((rc = dev_upload(&c->chi64, S.chi.data(), S.chi.size()))) return rc;
        std::vector<float> chi32(S.chi.begin(), S.chi.end());
        if ((rc = dev_upload(&c->chi32, chi32.data(), chi32.size()))) return rc;
        if ((rc = dev_upload(&c->rad64, S.rad.data(), S.rad.size()))) return rc;
        std::vector<float> rad32(S.rad.begin(), S.rad.end());
        if ((rc = dev_upload(&c->rad32, rad32.data(), rad32.size()))) return rc;
    }
    std::vector<int64_t> roff(M);
    std::vector<double4> cen(M);
    for (int g = 0; g < M; g++) {
        roff[g] = S.images[g].rad_off;
        cen[g] = make_double4(S.images[g].X0, S.images[g].Y0, S.images[g].Z0, 0.0);
        // dims.w packs channels (8 bits), nonlocal l count (8 bits) and type (8 bits)
        dim[g].w = S.images[g].nlm | (S.images[g].nl << 8) | (S.images[g].type << 16);
    }
    if ((rc = dev_upload(&c->img_dimp, dim.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_rad, roff.data(), M))) return rc;
    if ((rc = dev_upload(&c->img_cen, cen.data(), M))) return rc;
    std::vector<double> gam((size_t)kMaxTypesDev * kMaxChannels, 0.0);
    for (int t = 0; t < S.ntype; t++)
        for (int ch = 0; ch < S.psd[t].nlm(); ch++) gam[t * kMaxChannels + ch] = S.psd[t].gamma[S.psd[t].ch_l[ch]];
    if ((rc = dev_upload(&c->gamma, gam.data(), gam.size()))) return rc;
    if ((rc = dev_upload(&c->cov_off, S.cov_off.data(), S.cov_off.size()))) return rc;
    if ((rc = dev_upload(&c->cov_img, S.cov_img.data(), S.cov_img.size()))) return rc;
    std::vector<double> cf(3 * (kMaxOrder + 1), 0.0);
    std::vector<float> cf32(3 * (kMaxOrder + 1), 0.f);
    for (int d = 0; d < 3; d++)
        for (int p = 0; p <= S.order; p++) {
            cf[d * (kMaxOrder + 1) + p] = S.coeff[d][p];
            cf32[d * (kMaxOrder + 1) + p] = (float)S.coeff[d][p];
        }
    if ((rc = dev_upload(&c->c64, cf.data(), cf.size()))) return rc;
    if ((rc = dev_upload(&c->c32, cf32.data(), cf32.size()))) return rc;
    if ((rc = upload_diag(c, nullptr, 0.333))) return rc;
    // per-node outputs
    const int64_t nn = S.n_nodes();
    HIP_TRY(hipMalloc((void **)&c->a, nn * (S.nqp + 1) * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->b, nn * (S.nqp + 1) * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->lam, nn * S.nqp * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->w, nn * S.nqp * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->lmin, nn * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->lmax, nn * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->status, std::max<int64_t>(nn, 1) * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&c->qlws, 3 * nn * S.nqp * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->partial, kRedBlocks * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->partial2, kRedBlocks * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->out, 4 * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->queue, 8 * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&c->nsteps, std::max<int64_t>(1, nn) * sizeof(int)));
    // LDS layout of the Lanczos workgroup
    c->img_cap = std::max(1, max_images_per_cube(S));
    c->nlm_tmpl = S.max_nlm <= 6 ? 6 : kMaxChannels;
    if (S.max_nlm > kMaxChannels) return set_err(SQ_ERR_ARG, "more (l,m) channels than the kernels support");
    auto align16 = [](size_t x) { return (x + 15) & ~size_t(15); };
    size_t o = align16((size_t)c->img_cap * sizeof(int));
    c->lds_A = (int)o;
    o = align16(o + (size_t)c->img_cap * c->nlm_tmpl * sizeof(double));
    c->lds_alpha = (int)o;
    o = align16(o + (size_t)c->img_cap * c->nlm_tmpl * sizeof(double));
    c->lds_red = (int)o;
    o = align16(o + 32 * sizeof(double));
    c->lds_local = (int)o;
    o = align16(o + (size_t)std::max(M, 1) * sizeof(short));
    c->lds = o;
    if (c->lds > 150 * 1024) return set_err(SQ_ERR_ARG, "projector tables exceed the LDS budget of a workgroup");
    if (c->img_cap > 32767) return set_err(SQ_ERR_ARG, "too many images per cube");
    // one resident workgroup per CU slot; the grid strides over nodes
    c->grid = c->n_cu * 2;
    // z-march kernel: FD order 12, columns fit a workgroup, z and image counts
    // fit the packed column entries, chi offsets fit 31 bits
    const int ncol = S.nloc[0] * S.nloc[1];
    c->kmax = max_images_per_column(S);
    c->zmarch = S.order == 6 && ncol <= 1024 && c->img_cap <= 256 && getenv("SQ_FORCE_GENERIC") == nullptr;
    // Row pitch of the LDS planes: the stencil needs a halo of 6 on each side
    // (nx + 12).  Padding the pitch to nx + 32 keeps lanes that cross a row end
    // on consecutive bank slots; on the 128-atom cell it gains 3 % (FP64
    // one-node 16.48k -> 16.97k nodes/s, FP32 batched 34.4k -> 35.3k), the
    // results identical bit for bit.  SQ_PLANE_PAD (12..64) overrides it.
    {
        int pad = 32;
        if (const char *e = getenv("SQ_PLANE_PAD")) pad = std::max(12, std::min(64, atoi(e)));
        c->plane_px = S.nloc[0] + pad;
    }
    if (c->zmarch) {
        size_t z = align16((size_t)c->img_cap * sizeof(int));
        c->z_tlo = (int)z;
        z = align16(z + (size_t)c->img_cap * sizeof(int4));
        c->z_tdim = (int)z;
        z = align16(z + (size_t)c->img_cap * sizeof(int4));
        c->z_toff = (int)z;
        z = align16(z + (size_t)c->img_cap * sizeof(int64_t));
        c->z_tcen = (int)z;
        z = align16(z + (size_t)c->img_cap * sizeof(double4));
        c->z_A = (int)z;
        z = align16(z + (size_t)c->img_cap * c->nlm_tmpl * sizeof(double));
        c->z_alpha = (int)z;
        z = align16(z + (size_t)c->img_cap * c->nlm_tmpl * sizeof(double));
        c->z_red = (int)z;
        z = align16(z + 32 * sizeof(double));
        c->z_plane = (int)z;
        z = align16(z + 2 * (size_t)c->plane_px * (S.nloc[1] + 12) * sizeof(double));
        c->zlds = z;
        if (c->zlds > 150 * 1024) c->zmarch = false;
    }
    // batched kernel: the one-node kernel's conditions, s+d channels only
    bool sd = true;
    for (int t = 0; t < S.ntype; t++) {
        const Pseudopotential &ps = S.psd[t];
        if (!(ps.nlm() == 6 && ps.ch_l[0] == 0 && ps.ch_l[1] == 2 && ps.ch_l[5] == 2)) sd = false;
    }
    c->zbatch = c->zmarch && sd && getenv("SQ_NO_BATCH") == nullptr;
    for (int B = 2; B <= 4 && c->zbatch; B += 2) {
        const int cap = std::max(1, max_images_per_union(S, B));
        size_t z = align16((size_t)cap * sizeof(int));
        int *off = c->zb_off[B];
        off[0] = (int)z; z = align16(z + (size_t)cap * sizeof(int4));      // tlo
        off[1] = (int)z; z = align16(z + (size_t)cap * sizeof(int4));      // tdim
        off[2] = (int)z; z = align16(z + (size_t)cap * sizeof(int64_t));   // toff
        off[3] = (int)z; z = align16(z + (size_t)cap * sizeof(double4));   // tcen
        off[4] = (int)z; z = align16(z + (size_t)cap * B * 6 * sizeof(double));   // A
        off[5] = (int)z; z = align16(z + (size_t)cap * B * 6 * sizeof(double));   // alpha
        off[6] = (int)z; z = align16(z + (size_t)(B + 1) * 16 * sizeof(double) + 64);   // red
        off[7] = (int)z; z = align16(z + (size_t)B * 2 * (S.nqp + 1) * sizeof(double));   // early-exit rule
        off[8] = (int)z; z = align16(z + 2 * (size_t)c->plane_px * (S.nloc[1] + 12) * 16);   // plane (16 B a point)
        off[9] = (int)z; z = align16(z + (size_t)cap * sizeof(int4));      // atom runs
        off[10] = (int)z; z = align16(z + (size_t)cap * B * sizeof(double));   // per-atom shares
        c->zb_cap[B] = cap;
        c->zb_lds[B] = cap <= 256 && z <= 150 * 1024 ? z : 0;
    }
    if (c->zbatch) {
        // FP32: (t0, t2, t2 z, t2 z^2) per image box point, from the radial
        // factors (the FP64 form forms them in the kernel)
        const int64_t n4 = std::max<int64_t>(1, 2 * S.rad_total);
        HIP_TRY(hipMalloc((void **)&c->bas32, n4 * sizeof(float)));
        HIP_TRY(launch_basis_tables(c->rad64, M, c->img_lo, c->img_dim, c->img_rad, c->img_cen, S.h[2], S.range[2],
                                    c->bas32, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return SQ_OK;
}

KParams make_params(const sq_ctx *c, int mode, int precision, long begin, long count, bool zmarch = false)
{
    const System &S = c->sys;
    KParams P;
    memset(&P, 0, sizeof P);
    for (int d = 0; d < 3; d++) {
        P.nloc[d] = S.nloc[d];
        P.hw[d] = S.hw[d];
        P.box0[d] = S.box[d];
        P.box_n[d] = S.box[3 + d];
        P.ext0[d] = S.ext0[d];
        P.extn[d] = S.extn[d];
    }
    P.nc = (int)S.nc();
    P.center = (((S.nloc[2] - 1) / 2) * S.nloc[1] + (S.nloc[1] - 1) / 2) * S.nloc[0] + (S.nloc[0] - 1) / 2;
    P.order = S.order;
    P.node_begin = begin;
    P.node_count = count;
    P.nqp = S.nqp;
    P.mode = mode;
    P.nlm_max = c->nlm_tmpl;
    P.n_images = (int)S.images.size();
    P.img_cap = c->img_cap;
    P.lds_A = c->lds_A;
    P.lds_alpha = c->lds_alpha;
    P.lds_red = c->lds_red;
    P.lds_local = c->lds_local;
    const int st = kMaxOrder + 1;
    if (precision == SQ_FP64) {
        P.cx_ptr = c->c64; P.cy_ptr = c->c64 + st; P.cz_ptr = c->c64 + 2 * st;
        P.diag_ext = c->diag64;
        P.chi = c->chi64;
    } else {
        P.cx_ptr = c->c32; P.cy_ptr = c->c32 + st; P.cz_ptr = c->c32 + 2 * st;
        P.diag_ext = c->diag32;
        P.chi = c->chi32;
    }
    P.img_lo = c->img_lo;
    P.img_dim = c->img_dim;
    P.img_off = c->img_off;
    P.img_atom = c->img_atom;
    P.img_type = c->img_type;
    P.gamma_ch = c->gamma;
    P.cov_off = c->cov_off;
    P.cov_img = c->cov_img;
    for (int d = 0; d < 3; d++)
        for (int l = 0; l < 7 && l <= S.order; l++) {
            P.cw64[d][l] = S.coeff[d][l];
            P.cw32[d][l] = (float)S.coeff[d][l];
        }
    P.rad = precision == SQ_FP64 ? (const void *)c->rad64 : (const void *)c->rad32;
    P.img_rad = c->img_rad;
    P.img_cen = c->img_cen;
    for (int d = 0; d < 3; d++) {
        P.hgrid[d] = S.h[d];
        P.range[d] = S.range[d];
    }
    for (int t = 0; t < S.ntype && t < kMaxTypesDev; t++) {
        const Pseudopotential &ps = S.psd[t];
        int rl[4] = {-1, -1, -1, -1}, k = 0;
        for (int l = 0; l <= ps.lmax; l++)
            if (l != ps.lloc) rl[l] = k++;
        for (int ch = 0; ch < ps.nlm() && ch < kMaxChannels; ch++) {
            P.ch_l[t][ch] = (signed char)ps.ch_l[ch];
            P.ch_m[t][ch] = (signed char)ps.ch_m[ch];
            P.ch_r[t][ch] = (signed char)rl[ps.ch_l[ch]];
        }
    }
    P.layout = 1;
    for (int t = 0; t < S.ntype; t++) {
        const Pseudopotential &ps = S.psd[t];
        const bool sd = ps.nlm() == 6 && ps.ch_l[0] == 0 && ps.ch_l[1] == 2 && ps.ch_l[5] == 2;
        if (!sd) P.layout = 0;
    }
    if (zmarch) {
        P.img_dim = c->img_dimp;
        P.lds_A = c->z_A;
        P.lds_alpha = c->z_alpha;
        P.lds_red = c->z_red;
        P.lds_tlo = c->z_tlo;
        P.lds_tdim = c->z_tdim;
        P.lds_toff = c->z_toff;
        P.lds_tcen = c->z_tcen;
        P.lds_plane = c->z_plane;
        P.plane_px = c->plane_px;
        P.kmax = c->kmax;
        P.zwg = c->zwg[precision];
        P.queue = c->queue;
        P.nsteps = c->nsteps;
        if (c->exit_tol > 0) {
            P.exit_tol = c->exit_tol;
            P.exit_lf = c->exit_lf;
            P.exit_beta = c->exit_beta > 0 ? c->exit_beta : S.beta;
            P.exit_min = std::max(2, c->exit_min);
            P.exit_every = std::max(1, c->exit_every);
        }
    }
    P.scratch = c->scratch;
    P.out_a = c->a;
    P.out_b = c->b;
    P.status = c->status;
    return P;
}

// Parameters of the batched kernel for B nodes a workgroup over a batch list.
KParams make_batch_params(const sq_ctx *c, int mode, int precision, int B, const int2 *batch, long n_batch)
{
    KParams P = make_params(c, mode, precision, 0, n_batch, true);
    const int *off = c->zb_off[B];
    P.lds_tlo = off[0];
    P.lds_tdim = off[1];
    P.lds_toff = off[2];
    P.lds_tcen = off[3];
    P.lds_A = off[4];
    P.lds_alpha = off[5];
    P.lds_red = off[6];
    P.lds_local = off[7];
    P.lds_plane = off[8];
    P.lds_run = off[9];
    P.lds_share = off[10];
    P.basis = precision == SQ_FP64 ? nullptr : (const void *)c->bas32;
    P.img_cap = c->zb_cap[B];
    P.batch = batch;
    P.node_count = n_batch;
    P.node_begin = 0;
    return P;
}

// Resident grid of the one-node z-march kernel for a workgroup shape (a
// multiple of 8 so the XCD split is even); 0 if it cannot launch.
static int zmarch_grid(sq_ctx *c, int prec, int wg)
{
    KParams P = make_params(c, SQ_NL_KB, prec, 0, 1, true);
    P.zwg = wg;
    const int occ = zmarch_occupancy(P, prec, c->zlds);
    return occ <= 0 ? 0 : std::max(8, (c->n_cu * occ) & ~7);
}

// The gather trips the batched kernel is built with, per precision: a, b
// (FP64: 1 or 2 planes on the vector ALU; FP32: 4 planes on the matrix cores
// or 2 on the vector ALU).
static int batch_trip(int prec, int which) { return prec == 0 ? 1 + which : 4 - 2 * which; }

static int zbatch_grid(sq_ctx *c, int prec, int B, int ug)
{
    if (!c->zb_lds[B] || !zbatch_supported(prec, B, ug)) return 0;
    KParams P = make_batch_params(c, SQ_NL_KB, prec, B, nullptr, 1);
    const int occ = zbatch_occupancy(P, prec, B, ug, c->zb_lds[B]);
    return occ <= 0 ? 0 : std::max(8, (c->n_cu * occ) & ~7);
}

// Uploads a batch list into the context's buffer (grown as needed).
static int upload_batches(sq_ctx *c, const std::vector<int2> &v)
{
    if ((long)v.size() > c->batch_cap) {
        if (c->batches) HIP_TRY(hipFree(c->batches));
        c->batches = nullptr;
        c->batch_cap = std::max<long>((long)v.size(), 1024);
        HIP_TRY(hipMalloc((void **)&c->batches, c->batch_cap * sizeof(int2)));
    }
    if (!v.empty()) HIP_TRY(hipMemcpy(c->batches, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice));
    c->n_batches = (long)v.size();
    return SQ_OK;
}

// The batch list of rank-box ids [begin, begin + count) for B, cached.
static int ensure_batches(sq_ctx *c, int B, long begin, long count)
{
    if (c->bat_B == B && c->bat_begin == begin && c->bat_count == count) return SQ_OK;
    std::vector<int2> v;
    build_batches(c->sys, begin, count, B, v);
    int rc = upload_batches(c, v);
    if (rc) return rc;
    c->bat_B = B;
    c->bat_begin = begin;
    c->bat_count = count;
    return SQ_OK;
}

// One Lanczos launch of configuration (B, shape or trip) on the probe's node
// set, timed by HIP events on the context's stream.
static int probe_launch(sq_ctx *c, int prec, int B, int shape, int nqp, long nprobe, float *ms)
{
    hipError_t e = hipSuccess;
    for (int rep = 0; rep < 2; rep++) {   // the first launch loads the code object
        HIP_TRY(hipMemsetAsync(c->queue, 0, 8 * sizeof(int), c->stream));
        HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        if (B == 1) {
            KParams P = make_params(c, SQ_NL_KB, prec, 0, nprobe, true);
            P.zwg = shape;
            P.node_list = c->probe_ids;
            P.nqp = nqp;
            P.exit_tol = 0;
            const int grid = (int)std::min<long>(zmarch_grid(c, prec, shape), std::max<long>(8, (nprobe + 7) & ~7L));
            e = launch_zmarch(P, prec, grid, c->zlds, c->stream);
        } else {
            KParams P = make_batch_params(c, SQ_NL_KB, prec, B, c->batches, c->n_batches);
            P.nqp = nqp;
            P.exit_tol = 0;
            const int grid = (int)std::min<long>(c->zb_grid[prec][B], std::max<long>(8, (c->n_batches + 7) & ~7L));
            e = launch_zbatch(P, prec, B, shape, grid, c->zb_lds[B], c->stream);
        }
        if (e != hipSuccess) return set_err(SQ_ERR_DEVICE, std::string("autotuner probe: ") + hipGetErrorString(e));
        HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    }
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    HIP_TRY(hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
    return SQ_OK;
}

// The autotuner's rule.  Candidates: 0/1 the one-node kernel with 512/1024
// threads, 2/3 B = 2 with gather trip a/b, 4/5 B = 4 with trip a/b.
// available[k]: candidate k can launch on this deck.  The environment strings
// (NULL when unset) narrow the set: SQ_ZMARCH_WG keeps the one-node kernel in
// that shape, SQ_BATCH the batch size, SQ_BATCH_TRIP the trip of a batched
// candidate.  Nothing left: candidate 0 when it can launch.
static int tune_allowed(int prec, const int *available, const char *fixed_wg, const char *fixed_b,
                        const char *fixed_t, bool allow[6])
{
    int n = 0;
    for (int k = 0; k < 6; k++) {
        const int B = k < 2 ? 1 : (k < 4 ? 2 : 4);
        allow[k] = available[k] != 0;
        if (fixed_wg) allow[k] = B == 1 && (atoi(fixed_wg) == 1024) == (k == 1) && available[k];
        if (fixed_b && atoi(fixed_b) != B) allow[k] = false;
        if (fixed_t && B > 1 && atoi(fixed_t) != batch_trip(prec, k & 1)) allow[k] = false;
        n += allow[k];
    }
    if (n == 0 && available[0]) {
        allow[0] = true;
        n = 1;
    }
    return n;
}

// The fastest allowed candidate by its probe time (the lower index on a tie);
// -1 when none is allowed.
static int tune_pick(const bool allow[6], const float ms[6])
{
    int best = -1;
    for (int k = 0; k < 6; k++)
        if (allow[k] && (best < 0 || ms[k] < ms[best])) best = k;
    return best;
}

// Grids, the Lanczos scratch sized for the largest configuration, and the
// autotuner.  Per precision it times every configuration on one node set --
// four planes of the rank box's middle (fewer if the box is thinner), the
// first min(plane, 1024) nodes of each, 20 Lanczos steps (fewer if the deck
// runs fewer) -- and keeps the fastest:
//   the one-node z-march kernel with 512 threads (two columns a thread) or
//   1024 (one); the batched kernel with B = 2 (FP64, FP32) or 4 (FP32) nodes
//   a workgroup, each with two gather trips.
// The same nodes and steps for every candidate, so the times compare
// directly.  SQ_ZMARCH_WG=512|1024 fixes the one-node kernel in that shape;
// SQ_BATCH=1|2|4 (and SQ_BATCH_TRIP) fixes the batch size (and trip).
int finish_device_setup(sq_ctx *c)
{
    const System &S = c->sys;
    int maxgrid = c->grid;
    int grids[2][2] = {{0, 0}, {0, 0}};
    if (c->zmarch) {
        for (int prec = 0; prec < 2 && c->zmarch; prec++)
            for (int sh = 0; sh < 2; sh++) {
                grids[prec][sh] = zmarch_grid(c, prec, sh ? 1024 : 512);
                if (sh == 0 && grids[prec][0] == 0) c->zmarch = false;
                maxgrid = std::max(maxgrid, grids[prec][sh]);
            }
    }
    size_t scratch = (size_t)maxgrid * 3 * S.nc() * sizeof(double);
    if (c->zmarch && c->zbatch) {
        const size_t ncol = (size_t)S.nloc[0] * S.nloc[1];
        for (int B = 2; B <= 4; B += 2)
            for (int prec = 0; prec < 2; prec++)
                for (int w = 0; w < 2; w++) {
                    const int g = zbatch_grid(c, prec, B, batch_trip(prec, w));
                    if (!g) continue;
                    c->zb_grid[prec][B] = std::max(c->zb_grid[prec][B], g);
                    const size_t el = prec == 0 ? sizeof(double) : sizeof(float);
                    scratch = std::max(scratch, (size_t)g * 3 * (S.nloc[2] + B - 1) * ncol * B * el);
                }
    }
    HIP_TRY(hipMalloc(&c->scratch, scratch));
    if (!c->zmarch) return SQ_OK;
    for (int prec = 0; prec < 2; prec++) {
        c->zwg[prec] = 512;
        c->zgrid[prec] = grids[prec][0];
        c->cfg_B[prec] = 1;
        c->cfg_ug[prec] = 0;
    }
    const char *fixed_wg = getenv("SQ_ZMARCH_WG");
    const char *fixed_b = getenv("SQ_BATCH");
    const char *fixed_t = getenv("SQ_BATCH_TRIP");
    // the probe's node set
    const int lz = S.box[5];
    const long plane = (long)S.box[3] * S.box[4];
    const int np = std::min(4, lz), k0 = std::max(0, lz / 2 - 2);
    const long M = std::min<long>(plane, 1024);
    std::vector<int> ids;
    std::vector<int2> bl[5];
    for (long ij = 0; ij < M; ij++)
        for (int k = k0; k < k0 + np; k++) ids.push_back((int)((long)k * plane + ij));
    for (int B = 2; B <= 4; B += 2)
        for (long ij = 0; ij < M; ij++)
            for (int kb = k0; kb < k0 + np; kb += B)
                bl[B].push_back(make_int2((int)((long)kb * plane + ij), std::min(B, k0 + np - kb)));
    int rc;
    if ((rc = dev_upload(&c->probe_ids, ids.data(), ids.size()))) return rc;
    const int nqp = std::min(S.nqp, 20);
    for (int prec = 0; prec < 2; prec++) {
        int available[6];
        for (int k = 0; k < 6; k++) {
            const int B = k < 2 ? 1 : (k < 4 ? 2 : 4);
            available[k] = B == 1 ? grids[prec][k] > 0
                                  : (c->zbatch && zbatch_grid(c, prec, B, batch_trip(prec, k & 1)) > 0);
        }
        bool allow[6];
        const int n_allowed = tune_allowed(prec, available, fixed_wg, fixed_b, fixed_t, allow);
        float ms[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 6 && n_allowed > 1; k++) {
            if (!allow[k]) continue;
            const int B = k < 2 ? 1 : (k < 4 ? 2 : 4);
            if (B > 1 && (rc = upload_batches(c, bl[B]))) return rc;
            if ((rc = probe_launch(c, prec, B, B == 1 ? (k ? 1024 : 512) : batch_trip(prec, k & 1), nqp, (long)ids.size(),
                                   &ms[k])))
                return rc;
            c->tune_cand[prec][k] = ms[k];
            if (k < 2) c->tune_ms[prec][k] = ms[k];
        }
        int best = tune_pick(allow, ms);
        if (best < 0) best = 0;
        if (best < 2) {
            c->zwg[prec] = best ? 1024 : 512;
            c->zgrid[prec] = grids[prec][best];
        } else {
            const int B = best < 4 ? 2 : 4;
            c->cfg_B[prec] = B;
            c->cfg_ug[prec] = batch_trip(prec, best & 1);
            // the one-node kernel stays available (force pass, node lists) in the 1024 shape when it fits
            c->zwg[prec] = grids[prec][1] > 0 ? 1024 : 512;
            c->zgrid[prec] = grids[prec][grids[prec][1] > 0 ? 1 : 0];
        }
    }
    c->bat_B = 0;   // the probe's list is not a node range
    c->bat_begin = c->bat_count = -1;
    return SQ_OK;
}

int check_status(sq_ctx *c, long begin, long count, const char *what)
{
    std::vector<int> st(count);
    HIP_TRY(hipMemcpy(st.data(), c->status + begin, count * sizeof(int), hipMemcpyDeviceToHost));
    for (long r = 0; r < count; r++)
        if (st[r]) {
            char msg[160];
            snprintf(msg, sizeof msg, "%s: node %ld (rank-box id) status %d", what, begin + r, st[r]);
            int code = st[r] == 7 ? SQ_ERR_BREAKDOWN : st[r] == 5 ? SQ_ERR_QL : SQ_ERR_ARG;
            return set_err(code, msg);
        }
    return SQ_OK;
}

}  // namespace

extern "C" {

const char *sq_last_error(void) { return g_err.c_str(); }

int sq_tune_select(int precision, const int *available, const float *probe_ms, const char *fixed_workgroup,
                   const char *fixed_batch, const char *fixed_trip, int *candidate)
{
    if (!available || !probe_ms || !candidate) return set_err(SQ_ERR_ARG, "null argument");
    if (precision != SQ_FP64 && precision != SQ_FP32) return set_err(SQ_ERR_ARG, "precision must be SQ_FP64 or SQ_FP32");
    bool allow[6];
    tune_allowed(precision, available, fixed_workgroup, fixed_batch, fixed_trip, allow);
    *candidate = tune_pick(allow, probe_ms);
    return SQ_OK;
}

void sq_default_options(sq_options *opt)
{
    memset(opt, 0, sizeof *opt);
}

int sq_create_from_deck(const char *dir, const char *name, const sq_options *opt, sq_ctx **out)
{
    g_err.clear();
    if (!dir || !name || !out) return set_err(SQ_ERR_ARG, "null argument");
    *out = nullptr;
    sq_options defaults;
    sq_default_options(&defaults);
    if (!opt) opt = &defaults;
    SetupOptions so;
    so.phantom_atom = opt->phantom_atom != 0;
    for (int d = 0; d < 3; d++) so.phantom_pos[d] = opt->phantom_pos[d];
    for (int d = 0; d < 6; d++) so.rank_box[d] = opt->rank_box[d];
    so.nqp_override = opt->nqp_override;
    // a device context builds chi and rad on the device; SQ_HOST_TABLES=1 builds them on the host
    so.host_tables = opt->device < 0 || getenv("SQ_HOST_TABLES") != nullptr;
    sq_ctx *c = new sq_ctx();
    Status st = build_system(dir, name, so, c->sys);
    if (!st.ok()) {
        free_ctx(c);
        return set_err(st.code, st.msg);
    }
    if (c->sys.nqp > kMaxNqp) {
        free_ctx(c);
        return set_err(SQ_ERR_ARG, "LANCZOS_DEGREE above the supported maximum");
    }
    int rc = opt->device >= 0 ? create_device(c, opt->device) : SQ_OK;
    if (!rc && opt->device >= 0) rc = finish_device_setup(c);
    if (rc) {
        std::string keep = g_err;
        free_ctx(c);
        g_err = keep;
        return rc;
    }
    *out = c;
    return SQ_OK;
}

int sq_destroy(sq_ctx *ctx)
{
    if (ctx) {
        if (ctx->stream) (void)hipSetDevice(ctx->device);
        free_ctx(ctx);
    }
    return SQ_OK;
}

int sq_get_info(sq_ctx *c, sq_info *I)
{
    if (!c || !I) return set_err(SQ_ERR_ARG, "null argument");
    const System &S = c->sys;
    memset(I, 0, sizeof *I);
    for (int d = 0; d < 3; d++) {
        I->npts[d] = S.npts[d];
        I->nloc[d] = S.nloc[d];
        I->halfwidth[d] = S.hw[d];
        I->range[d] = S.range[d];
        I->delta[d] = S.h[d];
        I->ext_origin[d] = S.ext0[d];
        I->ext_dims[d] = S.extn[d];
        for (int p = 0; p <= S.order && p < 11; p++) I->coeffs[d][p] = S.coeff[d][p];
    }
    for (int d = 0; d < 6; d++) I->rank_box[d] = S.box[d];
    I->order = S.order;
    I->nqp = S.nqp;
    I->ntype = S.ntype;
    I->nAtoms = S.nAtoms;
    I->noetot = S.noetot;
    I->n_images = (int)S.images.size();
    I->n_sc_atoms = (int)S.sctype.size();
    I->ip_length = S.ip_length;
    I->n_nodes = (int)S.n_nodes();
    I->nc = (int)S.nc();
    I->dV = S.dV;
    I->beta = S.beta;
    I->rcut = S.rcut;
    for (int t = 0; t < S.ntype && t < 8; t++) {
        for (int l = 0; l < 4; l++) I->gamma[t][l] = S.psd[t].gamma[l];
        for (int d = 0; d < 3; d++) I->cutoff[t][d] = S.cutoff[t][d];
    }
    I->chi_entries = (long long)S.chi_total;
    I->kernel = c->zmarch ? 1 : 0;
    I->kmax = c->kmax;
    for (int prec = 0; prec < 2; prec++) {
        I->workgroup[prec] = c->zmarch ? (c->cfg_B[prec] > 1 ? 1024 : c->zwg[prec]) : 1024;
        I->tune_ms[prec][0] = c->tune_ms[prec][0];
        I->tune_ms[prec][1] = c->tune_ms[prec][1];
        I->batch[prec] = c->zmarch ? c->cfg_B[prec] : 1;
        I->gather_trip[prec] = c->cfg_ug[prec];
        for (int k = 0; k < 6; k++) I->tune_candidates_ms[prec][k] = c->tune_cand[prec][k];
    }
    I->union_images[0] = c->zb_cap[2];
    I->union_images[1] = c->zb_cap[4];
    return SQ_OK;
}

int sq_set_veff(sq_ctx *c, const double *field, double veff_const)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    HIP_TRY(hipSetDevice(c->device));
    return upload_diag(c, field, veff_const);
}

int sq_run(sq_ctx *c, int mode, int precision, long begin, long count)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    const System &S = c->sys;
    if (mode != SQ_NL_REF_COLLAPSED && mode != SQ_NL_KB) return set_err(SQ_ERR_ARG, "unknown nonlocal mode");
    if (precision != SQ_FP64 && precision != SQ_FP32) return set_err(SQ_ERR_ARG, "unknown precision");
    if (begin < 0 || count <= 0 || begin + count > S.n_nodes()) return set_err(SQ_ERR_ARG, "node range outside the rank box");
    HIP_TRY(hipSetDevice(c->device));
    c->have_run = false;
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    c->last_exit = false;
    if (c->zmarch && c->cfg_B[precision] > 1) {
        const int B = c->cfg_B[precision];
        int rc = ensure_batches(c, B, begin, count);
        if (rc) return rc;
        KParams P = make_batch_params(c, mode, precision, B, c->batches, c->n_batches);
        int grid = c->zb_grid[precision][B];
        if (c->n_batches < grid) grid = std::max<long>(8, (c->n_batches + 7) & ~7L);
        c->last_exit = P.exit_tol > 0;
        HIP_TRY(hipMemsetAsync(c->queue, 0, 8 * sizeof(int), c->stream));
        HIP_TRY(launch_zbatch(P, precision, B, c->cfg_ug[precision], grid, c->zb_lds[B], c->stream));
    } else if (c->zmarch) {
        KParams P = make_params(c, mode, precision, begin, count, true);
        int grid = c->zgrid[precision];
        if (count < grid) grid = std::max<long>(8, (count + 7) & ~7L);
        c->last_exit = P.exit_tol > 0;
        HIP_TRY(hipMemsetAsync(c->queue, 0, 8 * sizeof(int), c->stream));
        HIP_TRY(launch_zmarch(P, precision, grid, c->zlds, c->stream));
    } else {
        KParams P = make_params(c, mode, precision, begin, count);
        const int grid = (int)std::min<long>(c->grid, count);
        HIP_TRY(launch_lanczos(P, precision, grid, c->lds, c->stream));
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    int rc = check_status(c, begin, count, "Lanczos");
    if (rc) return rc;
    HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(launch_ql(c->a + begin * (S.nqp + 1), c->b + begin * (S.nqp + 1), S.nqp, S.nqp + 1, count,
                      c->lam + begin * S.nqp, c->w + begin * S.nqp, c->lmin + begin, c->lmax + begin, c->status + begin,
                      c->qlws, c->stream, c->last_exit ? c->nsteps + begin : nullptr));
    HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[3]));
    if ((rc = check_status(c, begin, count, "QL"))) return rc;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->t_lanczos = ms;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->t_ql = ms;
    c->last_begin = begin;
    c->last_count = count;
    c->have_run = true;
    return SQ_OK;
}

int sq_get_nodes_and_weights(sq_ctx *c, double *lambda, double *w, double *lmin, double *lmax)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->have_run) return set_err(SQ_ERR_STATE, "no completed sq_run");
    const int nqp = c->sys.nqp;
    const long b = c->last_begin, n = c->last_count;
    HIP_TRY(hipSetDevice(c->device));
    if (lambda) HIP_TRY(hipMemcpy(lambda, c->lam + b * nqp, n * nqp * sizeof(double), hipMemcpyDeviceToHost));
    if (w) HIP_TRY(hipMemcpy(w, c->w + b * nqp, n * nqp * sizeof(double), hipMemcpyDeviceToHost));
    if (lmin) HIP_TRY(hipMemcpy(lmin, c->lmin + b, n * sizeof(double), hipMemcpyDeviceToHost));
    if (lmax) HIP_TRY(hipMemcpy(lmax, c->lmax + b, n * sizeof(double), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_get_lanczos(sq_ctx *c, double *a, double *b)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->have_run) return set_err(SQ_ERR_STATE, "no completed sq_run");
    const int s = c->sys.nqp + 1;
    const long b0 = c->last_begin, n = c->last_count;
    HIP_TRY(hipSetDevice(c->device));
    if (a) HIP_TRY(hipMemcpy(a, c->a + b0 * s, n * s * sizeof(double), hipMemcpyDeviceToHost));
    if (b) HIP_TRY(hipMemcpy(b, c->b + b0 * s, n * s * sizeof(double), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_nodes_and_weights(sq_ctx *c, int mode, int precision, long begin, long count, double *lambda, double *w,
                         double *lmin, double *lmax)
{
    int rc = sq_run(c, mode, precision, begin, count);
    if (rc) return rc;
    return sq_get_nodes_and_weights(c, lambda, w, lmin, lmax);
}

int sq_last_timing(sq_ctx *c, double *lanczos_ms, double *ql_ms)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->have_run) return set_err(SQ_ERR_STATE, "no completed sq_run");
    if (lanczos_ms) *lanczos_ms = c->t_lanczos;
    if (ql_ms) *ql_ms = c->t_ql;
    return SQ_OK;
}

static int apply_common(sq_ctx *c, int mode, int Kp, int Jp, int Ip, const double *x, double *y, int nonlocal_only)
{
    if (!c || !x || !y) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    const System &S = c->sys;
    if (mode != SQ_NL_REF_COLLAPSED && mode != SQ_NL_KB) return set_err(SQ_ERR_ARG, "unknown nonlocal mode");
    int rel[3] = {Ip - S.box[0], Jp - S.box[1], Kp - S.box[2]};
    for (int d = 0; d < 3; d++)
        if (rel[d] < 0 || rel[d] >= S.box[3 + d]) return set_err(SQ_ERR_ARG, "node outside the rank box");
    const long id = ((long)rel[2] * S.box[4] + rel[1]) * S.box[3] + rel[0];
    HIP_TRY(hipSetDevice(c->device));
    const size_t nc = S.nc();
    double *dx = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc((void **)&dx, nc * sizeof(double)));
    hipError_t e = hipMalloc((void **)&dy, nc * sizeof(double));
    if (e != hipSuccess) { (void)hipFree(dx); return set_err(SQ_ERR_DEVICE, "hipMalloc failed"); }
    int rc = SQ_OK;
    KParams P = make_params(c, mode, SQ_FP64, id, 1);
    int zero = 0;
    if (hipMemcpy(dx, x, nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        (nonlocal_only && hipMemcpy(dy, y, nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(c->status, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        launch_apply(P, id, dx, dy, nonlocal_only, c->lds, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(y, dy, nc * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "apply device step failed");
    if (!rc) {
        int st = 0;
        (void)hipMemcpy(&st, c->status, sizeof(int), hipMemcpyDeviceToHost);
        if (st) rc = set_err(SQ_ERR_ARG, "apply: image list overflow");
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    c->have_run = false;  // status slot 0 was reused
    return rc;
}

int sq_apply_h(sq_ctx *c, int mode, int Kp, int Jp, int Ip, const double *x, double *y)
{
    return apply_common(c, mode, Kp, Jp, Ip, x, y, 0);
}

int sq_apply_nonlocal(sq_ctx *c, int mode, int Kp, int Jp, int Ip, const double *x, double *y)
{
    return apply_common(c, mode, Kp, Jp, Ip, x, y, 1);
}

int sq_tridiag_nodes_weights(sq_ctx *c, const double *diag, const double *subdiag, int n, long count, double *lambda,
                             double *w, double *lmin, double *lmax)
{
    if (!c || !diag || !subdiag) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    if (n < 1 || n > kMaxNqp || count <= 0) return set_err(SQ_ERR_ARG, "bad tridiagonal size");
    HIP_TRY(hipSetDevice(c->device));
    const size_t m = (size_t)n * count;
    double *buf = nullptr;
    int *st = nullptr;
    HIP_TRY(hipMalloc((void **)&buf, (2 * m + 2 * m + 2 * count + 3 * m) * sizeof(double)));
    hipError_t e = hipMalloc((void **)&st, count * sizeof(int));
    if (e != hipSuccess) { (void)hipFree(buf); return set_err(SQ_ERR_DEVICE, "hipMalloc failed"); }
    double *da = buf, *db = da + m, *dl = db + m, *dw = dl + m, *dmin = dw + m, *dmax = dmin + count, *ws = dmax + count;
    int rc = SQ_OK;
    if (hipMemcpy(da, diag, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, subdiag, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        launch_ql(da, db, n, n, count, dl, dw, dmin, dmax, st, ws, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "QL device step failed");
    if (!rc) {
        std::vector<int> hs(count);
        (void)hipMemcpy(hs.data(), st, count * sizeof(int), hipMemcpyDeviceToHost);
        for (long r = 0; r < count && !rc; r++)
            if (hs[r]) rc = set_err(SQ_ERR_QL, "QL: more than 50 sweeps for tridiagonal " + std::to_string(r));
    }
    if (!rc) {
        if ((lambda && hipMemcpy(lambda, dl, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
            (w && hipMemcpy(w, dw, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
            (lmin && hipMemcpy(lmin, dmin, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
            (lmax && hipMemcpy(lmax, dmax, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
            rc = set_err(SQ_ERR_DEVICE, "QL copy-back failed");
    }
    (void)hipFree(buf);
    (void)hipFree(st);
    return rc;
}

// A caller's device buffer: device memory of the context's device, with at
// least `bytes` from p to the end of its allocation.
static int check_device_buffer(const sq_ctx *c, const void *p, size_t bytes, const char *name)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != c->device) {
        (void)hipGetLastError();
        return set_err(SQ_ERR_ARG, std::string(name) + ": not device memory of the context's device");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(SQ_ERR_ARG, std::string(name) + ": allocation range unknown");
    }
    if ((const char *)p + bytes > (const char *)base + size)
        return set_err(SQ_ERR_ARG, std::string(name) + ": smaller than the tridiagonal batch");
    return SQ_OK;
}

int sq_tridiag_nodes_weights_noalloc(sq_ctx *c, const double *diag, const double *subdiag, int n, long count,
                                     double *lambda, double *w, double *lmin, double *lmax)
{
    if (!c || !diag || !subdiag || !lambda || !w || !lmin || !lmax) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    if (n < 1 || n > kMaxNqp || count <= 0) return set_err(SQ_ERR_ARG, "bad tridiagonal size");
    const size_t need = noalloc_ql_bytes(n, count);
    if (need > kNoAllocScratchBytes) {
        long lo = 0, hi = (long)(kNoAllocScratchBytes / sizeof(double));
        while (lo < hi) {
            const long mid = (lo + hi + 1) / 2;
            if (noalloc_ql_bytes(n, mid) <= kNoAllocScratchBytes) lo = mid;
            else hi = mid - 1;
        }
        return set_err(SQ_ERR_ARG, "tridiagonal batch needs " + std::to_string(need) + " bytes of workspace; the "
                       "static workspace holds " + std::to_string(kNoAllocScratchBytes) + ": at n = " +
                       std::to_string(n) + ", count <= " + std::to_string(lo));
    }
    HIP_TRY(hipSetDevice(c->device));
    void *ws = noalloc_scratch(c->device);
    if (!ws) return set_err(SQ_ERR_STATE, "static workspace address not resolved on this device");
    const size_t m = (size_t)n * count * sizeof(double);
    int rc;
    if ((rc = check_device_buffer(c, diag, m, "diag")) || (rc = check_device_buffer(c, subdiag, m, "subdiag")) ||
        (rc = check_device_buffer(c, lambda, m, "lambda")) || (rc = check_device_buffer(c, w, m, "w")) ||
        (rc = check_device_buffer(c, lmin, count * sizeof(double), "lmin")) ||
        (rc = check_device_buffer(c, lmax, count * sizeof(double), "lmax")))
        return rc;
    long long *dbad = nullptr, bad = -1;
    HIP_TRY(launch_ql_noalloc(ws, diag, subdiag, n, count, lambda, w, lmin, lmax, &dbad, c->stream));
    HIP_TRY(hipMemcpyAsync(&bad, dbad, sizeof bad, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (bad >= 0) return set_err(SQ_ERR_QL, "QL: more than 50 sweeps for tridiagonal " + std::to_string(bad));
    return SQ_OK;
}

// Fermi level by bisection on the device electron count, then rho (into the
// device array drho when given), E_band and -TS.
static int density_device(sq_ctx *c, double nelec, double beta, double *lambda_f, double *drho, double *e_band,
                          double *minus_ts)
{
    const System &S = c->sys;
    const long nn = S.n_nodes();
    if (!c->have_run || c->last_begin != 0 || c->last_count != nn)
        return set_err(SQ_ERR_STATE, "density needs a completed sq_run over every node of the rank box");
    if (nelec <= 0) nelec = S.noetot;
    if (beta <= 0) beta = S.beta;
    HIP_TRY(hipSetDevice(c->device));
    const long nq = nn * S.nqp;
    double range[2];
    HIP_TRY(launch_lam_range(c->lmin, c->lmax, nn, c->out, c->stream));
    HIP_TRY(hipMemcpyAsync(range, c->out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double lo = range[0] - 1.0, hi = range[1] + 1.0;
    auto count_at = [&](double lf, double *res) -> int {
        HIP_TRY(launch_occupation(c->lam, c->w, nq, beta, lf, c->partial, kRedBlocks, c->out, c->stream));
        HIP_TRY(hipMemcpyAsync(res, c->out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return SQ_OK;
    };
    double glo, ghi;
    int rc;
    if ((rc = count_at(lo, &glo)) || (rc = count_at(hi, &ghi))) return rc;
    if (glo - nelec > 0 || ghi - nelec < 0) return set_err(SQ_ERR_RANGE, "Fermi level not bracketed by the spectrum");
    // bisection on [lambda_min - 1, lambda_max + 1] to 1e-14 (relative), at most 200 halvings
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (hi - lo <= 1e-14 * (std::fabs(mid) > 1.0 ? std::fabs(mid) : 1.0)) break;
        double g;
        if ((rc = count_at(mid, &g))) return rc;
        if (g - nelec > 0) hi = mid;
        else lo = mid;
    }
    const double lf = 0.5 * (lo + hi);
    double res[2];
    HIP_TRY(launch_density_energy(c->lam, c->w, nn, S.nqp, beta, lf, 2.0 / S.dV, drho, c->partial, c->partial2,
                                  kRedBlocks, c->out, c->stream));
    HIP_TRY(hipMemcpyAsync(res, c->out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *lambda_f = lf;
    *e_band = res[0];
    *minus_ts = res[1];
    return SQ_OK;
}

int sq_density_fermi_energy(sq_ctx *c, double nelec, double beta, double *lambda_f, double *rho, double *e_band,
                            double *minus_ts)
{
    if (!c || !lambda_f || !e_band || !minus_ts) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    const long nn = c->sys.n_nodes();
    double *drho = nullptr;
    if (rho) HIP_TRY(hipMalloc((void **)&drho, nn * sizeof(double)));
    int rc = density_device(c, nelec, beta, lambda_f, drho, e_band, minus_ts);
    if (!rc && rho && hipMemcpy(rho, drho, nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "density copy failed");
    if (drho) (void)hipFree(drho);
    return rc;
}

// ================================================================= SCF
// The reference declares its SCF (SCF_LSSGQ, SolvePoisson, Vxc_Calc/Exc_Calc,
// mix: src/lsdft.h:93,124,138,183,197-204) and builds only the pieces; the
// loop here runs on the periodic cell -- the (N-1)^3 distinct grid points,
// the context's rank box -- with every field on the device:
//   V_xc, eps_xc  <- LDA (Slater + PW92) of rho_in
//   phi           <- -(1/4pi) Lap_h phi = rho_in + b by conjugate gradients on
//                    the periodic -(1/2) Lap_h (warm start, mean zero)
//   V_eff = phi + V_xc -> the node cubes' diagonal -> SQ pass -> rho_out
//   rho_in        <- Periodic Pulay mix of (rho_in, rho_out - rho_in)
// until ||rho_out - rho_in|| / ||rho_out|| < TOL_SCF, with the Harris-Foulkes
// energy E = E_band + E_xc - int V_xc rho + 1/2 int phi (b - rho) - E_self - TS.

int sq_get_rule_sizes(sq_ctx *c, int *n);
int sq_set_early_exit(sq_ctx *c, double tol, double lambda_f, double beta, int min_steps, int every);

static int cell_dims(const sq_ctx *c, int G[3])
{
    const System &S = c->sys;
    for (int d = 0; d < 3; d++) {
        G[d] = S.npts[d] - 1;
        if (S.box[d] != 0 || S.box[3 + d] != G[d])
            return set_err(SQ_ERR_STATE, "the SCF runs on the periodic cell: create the context with rank_box = "
                                         "(0, 0, 0, N0-1, N1-1, N2-1)");
    }
    return SQ_OK;
}

static StencilCoeffs stencil_of(const System &S)
{
    StencilCoeffs sc;
    memset(&sc, 0, sizeof sc);
    for (int d = 0; d < 3; d++)
        for (int a = 0; a <= S.order && a < 11; a++) sc.w[d][a] = S.coeff[d][a];
    return sc;
}

static int scf_alloc(sq_ctx *c, int hist)
{
    auto &F = c->scf;
    const long ng = c->sys.n_nodes();
    if (F.rho && F.ng == ng && F.hist >= hist) return SQ_OK;
    // the captured CG block names the old buffers
    if (F.cg_exec) (void)hipGraphExecDestroy(F.cg_exec);
    if (F.cg_graph) (void)hipGraphDestroy(F.cg_graph);
    F.cg_exec = nullptr;
    F.cg_graph = nullptr;
    double **all[] = {&F.rho, &F.rho_out, &F.b, &F.phi, &F.eps, &F.vxc, &F.veff, &F.r, &F.p, &F.Ap, &F.f, &F.x_prev,
                      &F.f_prev, &F.dX, &F.dF, &F.g, &F.s, &F.part, &F.ku, &F.ky};
    for (double **q : all) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    double **vecs[] = {&F.rho, &F.rho_out, &F.b, &F.phi, &F.eps, &F.vxc, &F.veff, &F.r, &F.p, &F.Ap, &F.f, &F.x_prev, &F.f_prev,
                       &F.ku, &F.ky};
    for (double **v : vecs) {
        HIP_TRY(hipMalloc((void **)v, ng * sizeof(double)));
        HIP_TRY(hipMemset(*v, 0, ng * sizeof(double)));
    }
    HIP_TRY(hipMalloc((void **)&F.dX, std::max(1, hist) * ng * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&F.dF, std::max(1, hist) * ng * sizeof(double)));
    HIP_TRY(hipMemset(F.dX, 0, std::max(1, hist) * ng * sizeof(double)));
    HIP_TRY(hipMemset(F.dF, 0, std::max(1, hist) * ng * sizeof(double)));
    if (!F.ctl) HIP_TRY(hipMalloc((void **)&F.ctl, sizeof(ScfCtl)));
    for (hipEvent_t &e : F.sev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipMalloc((void **)&F.g, std::max(1, hist) * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&F.s, 16 * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&F.part, kRedBlocks * sizeof(double)));
    F.ng = ng;
    F.hist = hist;
    if (F.have_b) HIP_TRY(hipMemcpy(F.b, F.host_b.data(), ng * sizeof(double), hipMemcpyHostToDevice));
    return SQ_OK;
}

static int scf_pseudocharge(sq_ctx *c)
{
    auto &F = c->scf;
    if (!F.have_b) {
        Status st = build_pseudocharge(c->sys, F.host_b, F.eself, F.b_scale);
        if (!st.ok()) return set_err(st.code, st.msg);
        F.have_b = true;
        if (F.b) HIP_TRY(hipMemcpy(F.b, F.host_b.data(), F.ng * sizeof(double), hipMemcpyHostToDevice));
    }
    return SQ_OK;
}

// <a, b> (or sum a when b is null) into slot of F.s, read back to host when out != nullptr
static int scf_dot(sq_ctx *c, const double *a, const double *b, int slot, double *out)
{
    auto &F = c->scf;
    HIP_TRY(launch_dot(a, b, F.ng, 1.0, F.part, kRedBlocks, F.s + slot, c->stream));
    if (out) {
        HIP_TRY(hipMemcpyAsync(out, F.s + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return SQ_OK;
}

// Solve -(1/4pi) Lap_h phi = f (f on the device, neutral up to rounding) into
// F.phi by conjugate gradients on A = -(1/2) Lap_h, A phi = 2 pi f, from the
// current F.phi; stops at ||r|| <= tol ||2 pi f|| (checked every 4 steps, the
// scalars of a step stay on the device) and removes the mean of phi.
static int scf_poisson(sq_ctx *c, const double *f, double tol, int max_it, int *iters)
{
    auto &F = c->scf;
    const System &S = c->sys;
    int G[3];
    if (int rc = cell_dims(c, G)) return rc;
    const StencilCoeffs sc = stencil_of(S);
    const long n = F.ng;
    int rc;
    double mean = 0.0;
    // right-hand side without its (rounding-level) mean: F.r = 2 pi (f - mean f)
    if ((rc = scf_dot(c, f, nullptr, 4, &mean))) return rc;
    mean /= (double)n;
    HIP_TRY(launch_axpbypgz(2.0 * M_PI, f, 0.0, F.r, -2.0 * M_PI * mean, nullptr, n, c->stream));
    double bnorm2 = 0.0;
    if ((rc = scf_dot(c, F.r, F.r, 5, &bnorm2))) return rc;
    const double stop = tol * tol * bnorm2;
    // r = 2 pi f - A phi ; p = r ; s0 = <r,r>
    HIP_TRY(launch_lap_periodic(F.phi, F.Ap, G, S.order, sc, c->stream));
    HIP_TRY(launch_axpbypgz(-1.0, F.Ap, 1.0, F.r, 0.0, nullptr, n, c->stream));
    HIP_TRY(hipMemcpyAsync(F.p, F.r, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if ((rc = scf_dot(c, F.r, F.r, 0, nullptr))) return rc;
    int it = 0;
    double rr = 0.0;
    HIP_TRY(hipMemcpyAsync(&rr, F.s, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // One CG step is seven launches (stencil, two two-stage dots, two updates),
    // all on fixed buffers with their scalars on the device: four steps are
    // captured once as a graph and replayed between the host's convergence
    // checks (SQ_NO_GRAPHS set: launched one by one, the same kernels in the
    // same order).
    auto cg_step = [&]() -> int {
        HIP_TRY(launch_lap_periodic(F.p, F.Ap, G, S.order, sc, c->stream));
        if (int r2 = scf_dot(c, F.p, F.Ap, 1, nullptr)) return r2;
        HIP_TRY(launch_cg_xr(F.phi, F.r, F.p, F.Ap, F.s, n, c->stream));
        if (int r2 = scf_dot(c, F.r, F.r, 2, nullptr)) return r2;
        HIP_TRY(launch_cg_p(F.p, F.r, F.s, n, c->stream));
        return SQ_OK;
    };
    const bool graphs = getenv("SQ_NO_GRAPHS") == nullptr;
    while (rr > stop && it < max_it) {
        if (graphs && it + 4 <= max_it) {
            if (!F.cg_exec) {
                HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int rc2 = SQ_OK;
                for (int k = 0; k < 4 && !rc2; k++) rc2 = cg_step();
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc2) { if (g) (void)hipGraphDestroy(g); return rc2; }
                if (e != hipSuccess) return set_err(SQ_ERR_DEVICE, "Poisson: graph capture failed");
                F.cg_graph = g;
                HIP_TRY(hipGraphInstantiate(&F.cg_exec, g, nullptr, nullptr, 0));
            }
            HIP_TRY(hipGraphLaunch(F.cg_exec, c->stream));
            it += 4;
        } else {
            for (int k = 0; k < 4 && it < max_it; k++, it++)
                if ((rc = cg_step())) return rc;
        }
        HIP_TRY(hipMemcpyAsync(&rr, F.s, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (!(rr <= stop)) return set_err(SQ_ERR_RANGE, "Poisson: conjugate gradients did not reach TOL_POISSON");
    if ((rc = scf_dot(c, F.phi, nullptr, 4, &mean))) return rc;
    mean /= (double)n;
    HIP_TRY(launch_axpbypgz(1.0, F.phi, 0.0, F.phi, -mean, nullptr, n, c->stream));
    if (iters) *iters = it;
    return SQ_OK;
}

// Kerker preconditioner y = A (A + k0^2/2)^-1 u, A = -(1/2) Lap_h on the cell
// (symbol q^2 / (q^2 + k0^2)): (A + k0^2/2) y = A u by conjugate gradients from
// y = 0 to ||r|| <= 1e-10 ||A u||, checked every 4 steps.  A u has zero mean,
// so y does: a preconditioned residual moves no charge.  Uses the CG buffers
// of the Poisson solve (not in use meanwhile) and runs eagerly.
static int scf_kerker(sq_ctx *c, const double *u, double *y, double k0, int *iters)
{
    auto &F = c->scf;
    const System &S = c->sys;
    int G[3];
    if (int rc = cell_dims(c, G)) return rc;
    const StencilCoeffs sc = stencil_of(S);
    const long n = F.ng;
    const double shift = 0.5 * k0 * k0;
    int rc;
    HIP_TRY(launch_lap_periodic(u, F.r, G, S.order, sc, c->stream));       // r = A u - (A + s) 0
    HIP_TRY(hipMemsetAsync(y, 0, n * sizeof(double), c->stream));
    HIP_TRY(hipMemcpyAsync(F.p, F.r, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    double bnorm2 = 0.0;
    if ((rc = scf_dot(c, F.r, F.r, 0, &bnorm2))) return rc;
    const double stop = 1e-20 * bnorm2;
    double rr = bnorm2;
    int it = 0;
    while (rr > stop && it < 2000) {
        for (int k = 0; k < 4; k++, it++) {
            HIP_TRY(launch_lap_periodic(F.p, F.Ap, G, S.order, sc, c->stream));
            HIP_TRY(launch_axpbypgz(shift, F.p, 1.0, F.Ap, 0.0, nullptr, n, c->stream));   // (A + s) p
            if ((rc = scf_dot(c, F.p, F.Ap, 1, nullptr))) return rc;
            HIP_TRY(launch_cg_xr(y, F.r, F.p, F.Ap, F.s, n, c->stream));
            if ((rc = scf_dot(c, F.r, F.r, 2, nullptr))) return rc;
            HIP_TRY(launch_cg_p(F.p, F.r, F.s, n, c->stream));
        }
        HIP_TRY(hipMemcpyAsync(&rr, F.s, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (!(rr <= stop)) return set_err(SQ_ERR_RANGE, "Kerker: conjugate gradients did not converge");
    if (iters) *iters = it;
    return SQ_OK;
}

// ------------------------------------------------------------ the SCF step on the device
// One SCF step as a fixed sequence of launches (sq_scfstep.hip): no host read
// inside it, so it can be captured as one graph and replayed.  eager: the
// conjugate-gradient loops are extended from the host until they stop (the
// first step's cold Poisson solve); otherwise they hold kPoissonBlocks /
// kKerkerBlocks blocks of four steps whose launches past the stop return at once.
constexpr int kPoissonBlocks = 192, kKerkerBlocks = 64;

struct StepPlan {
    int mode, prec;
    double beta, tol, tolp, kerker;
    int m, pf;
    bool eager;
};

static int cg_blocks(sq_ctx *c, bool eager, int nblocks, int *done, const std::function<int()> &block)
{
    if (!eager) {
        for (int k = 0; k < nblocks; k++)
            if (int rc = block()) return rc;
        return SQ_OK;
    }
    for (;;) {
        for (int k = 0; k < 16; k++)
            if (int rc = block()) return rc;
        int d = 0;
        HIP_TRY(hipMemcpyAsync(&d, done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (d) return SQ_OK;
    }
}

// Poisson: phi from f as scf_poisson, the scalars and the stop on the device
static int enqueue_poisson(sq_ctx *c, const double *f, double tol, bool eager)
{
    auto &F = c->scf;
    const System &S = c->sys;
    int G[3];
    if (int rc = cell_dims(c, G)) return rc;
    const StencilCoeffs sc = stencil_of(S);
    const long n = F.ng;
    hipStream_t st = c->stream;
    ScfCtl *ctl = F.ctl;
    HIP_TRY(launch_dot(f, nullptr, n, 1.0, F.part, kRedBlocks, F.s + 4, st));
    HIP_TRY(launch_axpb_mean(2.0 * M_PI, f, F.r, -2.0 * M_PI, F.s + 4, n, nullptr, st));
    HIP_TRY(launch_dot(F.r, F.r, n, 1.0, F.part, kRedBlocks, F.s + 5, st));
    HIP_TRY(launch_lap_periodic(F.phi, F.Ap, G, S.order, sc, st));
    HIP_TRY(launch_axpbypgz(-1.0, F.Ap, 1.0, F.r, 0.0, nullptr, n, st));
    HIP_TRY(launch_copy(F.p, F.r, n, nullptr, st));
    HIP_TRY(launch_dot(F.r, F.r, n, 1.0, F.part, kRedBlocks, F.s, st));
    HIP_TRY(launch_cg_begin(F.s, &ctl->pstop, &ctl->pdone, &ctl->piters, tol * tol, 5, nullptr, st));
    const int *sk = &ctl->pdone;
    auto block = [&]() -> int {
        for (int k = 0; k < 4; k++) {
            HIP_TRY(launch_lap_periodic(F.p, F.Ap, G, S.order, sc, st, sk));
            HIP_TRY(launch_dot(F.p, F.Ap, n, 1.0, F.part, kRedBlocks, F.s + 1, st, sk));
            HIP_TRY(launch_cg_xr(F.phi, F.r, F.p, F.Ap, F.s, n, st, sk));
            HIP_TRY(launch_dot(F.r, F.r, n, 1.0, F.part, kRedBlocks, F.s + 2, st, sk));
            HIP_TRY(launch_cg_p(F.p, F.r, F.s, n, st, sk));
        }
        HIP_TRY(launch_cg_check(F.s, &ctl->pstop, &ctl->pdone, &ctl->piters, 20000, st));
        return SQ_OK;
    };
    if (int rc = cg_blocks(c, eager, kPoissonBlocks, &ctl->pdone, block)) return rc;
    HIP_TRY(launch_cg_end(F.s, &ctl->pstop, &ctl->pfail, nullptr, st));
    HIP_TRY(launch_dot(F.phi, nullptr, n, 1.0, F.part, kRedBlocks, F.s + 4, st));
    HIP_TRY(launch_axpb_mean(1.0, F.phi, F.phi, -1.0, F.s + 4, n, nullptr, st));
    return SQ_OK;
}

// Kerker preconditioner as scf_kerker, skipped when the step has stopped the SCF
static int enqueue_kerker(sq_ctx *c, const double *u, double *y, double k0, bool eager)
{
    auto &F = c->scf;
    const System &S = c->sys;
    int G[3];
    if (int rc = cell_dims(c, G)) return rc;
    const StencilCoeffs sc = stencil_of(S);
    const long n = F.ng;
    const double shift = 0.5 * k0 * k0;
    hipStream_t st = c->stream;
    ScfCtl *ctl = F.ctl;
    const int *cv = &ctl->conv;
    HIP_TRY(launch_lap_periodic(u, F.r, G, S.order, sc, st, cv));
    HIP_TRY(launch_zero(y, n, cv, st));
    HIP_TRY(launch_copy(F.p, F.r, n, cv, st));
    HIP_TRY(launch_dot(F.r, F.r, n, 1.0, F.part, kRedBlocks, F.s, st, cv));
    HIP_TRY(launch_cg_begin(F.s, &ctl->kstop, &ctl->kdone, &ctl->kiters, 1e-20, 0, cv, st));
    const int *sk = &ctl->kdone;
    auto block = [&]() -> int {
        for (int k = 0; k < 4; k++) {
            HIP_TRY(launch_lap_periodic(F.p, F.Ap, G, S.order, sc, st, sk));
            HIP_TRY(launch_axpbypgz(shift, F.p, 1.0, F.Ap, 0.0, nullptr, n, st, sk));
            HIP_TRY(launch_dot(F.p, F.Ap, n, 1.0, F.part, kRedBlocks, F.s + 1, st, sk));
            HIP_TRY(launch_cg_xr(y, F.r, F.p, F.Ap, F.s, n, st, sk));
            HIP_TRY(launch_dot(F.r, F.r, n, 1.0, F.part, kRedBlocks, F.s + 2, st, sk));
            HIP_TRY(launch_cg_p(F.p, F.r, F.s, n, st, sk));
        }
        HIP_TRY(launch_cg_check(F.s, &ctl->kstop, &ctl->kdone, &ctl->kiters, 2000, st));
        return SQ_OK;
    };
    if (int rc = cg_blocks(c, eager, kKerkerBlocks, &ctl->kdone, block)) return rc;
    HIP_TRY(launch_cg_end(F.s, &ctl->kstop, &ctl->kfail, cv, st));
    return SQ_OK;
}

// The SQ pass of sq_run over every node, without its host checks (the first
// bad node status lands in ctl->bad)
static int enqueue_sq_pass(sq_ctx *c, int mode, int precision)
{
    const System &S = c->sys;
    const long count = S.n_nodes();
    hipStream_t st = c->stream;
    c->last_exit = false;
    if (c->zmarch && c->cfg_B[precision] > 1) {
        const int B = c->cfg_B[precision];
        KParams P = make_batch_params(c, mode, precision, B, c->batches, c->n_batches);
        int grid = c->zb_grid[precision][B];
        if (c->n_batches < grid) grid = std::max<long>(8, (c->n_batches + 7) & ~7L);
        c->last_exit = P.exit_tol > 0;
        HIP_TRY(launch_zero_int(c->queue, 8, st));
        HIP_TRY(launch_zbatch(P, precision, B, c->cfg_ug[precision], grid, c->zb_lds[B], st));
    } else if (c->zmarch) {
        KParams P = make_params(c, mode, precision, 0, count, true);
        int grid = c->zgrid[precision];
        if (count < grid) grid = std::max<long>(8, (count + 7) & ~7L);
        c->last_exit = P.exit_tol > 0;
        HIP_TRY(launch_zero_int(c->queue, 8, st));
        HIP_TRY(launch_zmarch(P, precision, grid, c->zlds, st));
    } else {
        KParams P = make_params(c, mode, precision, 0, count);
        HIP_TRY(launch_lanczos(P, precision, (int)std::min<long>(c->grid, count), c->lds, st));
    }
    HIP_TRY(launch_status_any(c->status, count, c->scf.ctl, st));
    HIP_TRY(launch_ql(c->a, c->b, S.nqp, S.nqp + 1, count, c->lam, c->w, c->lmin, c->lmax, c->status, c->qlws, st,
                      c->last_exit ? c->nsteps : nullptr));
    HIP_TRY(launch_status_any(c->status, count, c->scf.ctl, st));
    return SQ_OK;
}

// Fermi level by bisection, then density, band energy and -TS (density_device's steps)
static int enqueue_fermi_density(sq_ctx *c, double *rho_out)
{
    const System &S = c->sys;
    const long nn = S.n_nodes(), nq = nn * S.nqp;
    const double beta = S.beta, nelec = S.noetot;
    hipStream_t st = c->stream;
    ScfCtl *ctl = c->scf.ctl;
    HIP_TRY(launch_lam_range(c->lmin, c->lmax, nn, c->out, st));
    HIP_TRY(launch_bis_begin(c->out, ctl, st));
    HIP_TRY(launch_occupation_partial_dev(c->lam, c->w, nq, beta, &ctl->lo, nullptr, c->partial, kRedBlocks, st));
    HIP_TRY(launch_bis_store(c->partial, kRedBlocks, &ctl->glo, st));
    HIP_TRY(launch_occupation_partial_dev(c->lam, c->w, nq, beta, &ctl->hi, nullptr, c->partial, kRedBlocks, st));
    HIP_TRY(launch_bis_store(c->partial, kRedBlocks, &ctl->ghi, st));
    HIP_TRY(launch_bis_first(ctl, nelec, st));
    for (int k = 0; k < 200; k++) {
        HIP_TRY(launch_occupation_partial_dev(c->lam, c->w, nq, beta, &ctl->mid, &ctl->fdone, c->partial, kRedBlocks, st));
        HIP_TRY(launch_bis_step(c->partial, kRedBlocks, ctl, nelec, st));
    }
    HIP_TRY(launch_bis_end(ctl, st));
    HIP_TRY(launch_density_energy_dev(c->lam, c->w, nn, S.nqp, beta, &ctl->lf, 2.0 / S.dV, rho_out, c->partial,
                                      c->partial2, kRedBlocks, c->out, st));
    return SQ_OK;
}

// One whole SCF step: V_eff from the input density, SQ pass, Fermi level and
// density, energies of the input density, residual and stop, Pulay history,
// mixing into F.rho.
static int enqueue_scf_step(sq_ctx *c, const StepPlan &pl)
{
    auto &F = c->scf;
    const System &S = c->sys;
    int G[3];
    if (int rc = cell_dims(c, G)) return rc;
    const long n = F.ng;
    hipStream_t st = c->stream;
    ScfCtl *ctl = F.ctl;
    const int ext0[3] = {S.ext0[0], S.ext0[1], S.ext0[2]}, extn[3] = {S.extn[0], S.extn[1], S.extn[2]};
    int rc;
    HIP_TRY(launch_step_begin(ctl, st));
    HIP_TRY(launch_xc(F.rho, n, F.eps, F.vxc, st));
    HIP_TRY(launch_axpbypgz(1.0, F.rho, 0.0, F.veff, 1.0, F.b, n, st));   // rho + b
    HIP_TRY(hipEventRecord(F.sev[0], st));
    if ((rc = enqueue_poisson(c, F.veff, pl.tolp, pl.eager))) return rc;
    HIP_TRY(hipEventRecord(F.sev[1], st));
    HIP_TRY(launch_axpbypgz(1.0, F.phi, 0.0, F.veff, 1.0, F.vxc, n, st));  // phi + V_xc
    HIP_TRY(launch_veff_to_diag(F.veff, S.cdiag, ext0, extn, G, c->diag64, c->diag32, st));
    HIP_TRY(hipEventRecord(F.sev[2], st));
    if ((rc = enqueue_sq_pass(c, pl.mode, pl.prec)) || (rc = enqueue_fermi_density(c, F.rho_out))) return rc;
    HIP_TRY(hipEventRecord(F.sev[3], st));
    // energies of the input density; residual f = rho_out - rho_in; stop
    HIP_TRY(launch_dot(F.rho, F.eps, n, 1.0, F.part, kRedBlocks, F.s + 6, st));
    HIP_TRY(launch_dot(F.rho, F.vxc, n, 1.0, F.part, kRedBlocks, F.s + 7, st));
    HIP_TRY(launch_dot(F.phi, F.b, n, 1.0, F.part, kRedBlocks, F.s + 8, st));
    HIP_TRY(launch_dot(F.phi, F.rho, n, 1.0, F.part, kRedBlocks, F.s + 9, st));
    HIP_TRY(launch_axpbypgz(1.0, F.rho_out, 0.0, F.f, -1.0, F.rho, n, st));
    HIP_TRY(launch_dot(F.f, F.f, n, 1.0, F.part, kRedBlocks, F.s + 10, st));
    HIP_TRY(launch_dot(F.rho_out, F.rho_out, n, 1.0, F.part, kRedBlocks, F.s + 11, st));
    HIP_TRY(launch_conv(F.s, ctl, pl.tol, st));
    // Periodic Pulay: history of iterate and residual differences, Gram system, solve
    const int *cv = &ctl->conv;
    HIP_TRY(launch_hist_push(F.rho, F.x_prev, F.f, F.f_prev, F.dX, F.dF, n, ctl, st));
    HIP_TRY(launch_hist_rotate(ctl, pl.m, pl.pf, st));
    HIP_TRY(launch_copy(F.x_prev, F.rho, n, cv, st));
    HIP_TRY(launch_copy(F.f_prev, F.f, n, cv, st));
    for (int a = 0; a < pl.m; a++) {
        for (int b2 = 0; b2 <= a; b2++)
            HIP_TRY(launch_dot(F.dF + (size_t)a * n, F.dF + (size_t)b2 * n, n, 1.0, F.part, kRedBlocks,
                               &ctl->gram[a * pl.m + b2], st, &ctl->gskip[a]));
        HIP_TRY(launch_dot(F.dF + (size_t)a * n, F.f, n, 1.0, F.part, kRedBlocks, &ctl->rhs[a], st, &ctl->gskip[a]));
    }
    HIP_TRY(launch_pulay_solve(ctl, pl.m, st));
    if (pl.kerker > 0) {
        // x_new = (x - dX g) + beta P (f - dF g)
        HIP_TRY(launch_pulay_split(F.rho, F.f, F.dX, F.dF, ctl->g, 0, n, F.rho_out, F.ku, st, &ctl->mu, cv));
        if ((rc = enqueue_kerker(c, F.ku, F.ky, pl.kerker, pl.eager))) return rc;
        HIP_TRY(launch_axpbypgz(pl.beta, F.ky, 1.0, F.rho_out, 0.0, nullptr, n, st, cv));
    } else {
        HIP_TRY(launch_pulay_update(F.rho, F.f, F.dX, F.dF, ctl->g, 0, pl.beta, n, F.rho_out, st, &ctl->mu, cv));
    }
    HIP_TRY(launch_copy(F.rho, F.rho_out, n, cv, st));
    HIP_TRY(launch_it_next(ctl, st));
    return SQ_OK;
}

void sq_default_scf_options(sq_scf_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->nonlocal_mode = SQ_NL_KB;
    o->precision = SQ_FP64;
}

int sq_scf(sq_ctx *c, const sq_scf_options *opt_in, sq_scf_result *res)
{
    if (!c || !res) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    sq_scf_options opt;
    sq_default_scf_options(&opt);
    if (opt_in) opt = *opt_in;
    const System &S = c->sys;
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    if (S.xc != "LDA" && S.xc != "LDA_PW") return set_err(SQ_ERR_ARG, "EXCHANGE_CORRELATION " + S.xc + " is not LDA_PW");
    const int maxit = opt.max_iter > 0 ? opt.max_iter : S.maxit_scf;
    const double tol = opt.tol > 0 ? opt.tol : S.tol_scf;
    const double beta = opt.mixing > 0 ? opt.mixing : S.mix_beta;
    const int m = opt.history > 0 ? opt.history : S.mix_history;
    const int pf = opt.pulay_frequency > 0 ? opt.pulay_frequency : std::max(1, S.pulay_frequency);
    const double tolp = opt.tol_poisson > 0 ? opt.tol_poisson : S.tol_poisson;
    HIP_TRY(hipSetDevice(c->device));
    // the loop sets the early exit per pass; the caller's setting comes back after it
    struct ExitGuard {
        sq_ctx *c;
        double tol, lf, beta;
        int mn, ev;
        ~ExitGuard() { c->exit_tol = tol; c->exit_lf = lf; c->exit_beta = beta; c->exit_min = mn; c->exit_every = ev; }
    } guard{c, c->exit_tol, c->exit_lf, c->exit_beta, c->exit_min, c->exit_every};
    c->exit_tol = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = scf_alloc(c, m)) || (rc = scf_pseudocharge(c))) return rc;
    auto &F = c->scf;
    const long n = F.ng;
    const double dV = S.dV;
    if (opt.rho0) HIP_TRY(hipMemcpy(F.rho, opt.rho0, n * sizeof(double), hipMemcpyHostToDevice));
    else {
        std::vector<double> rho0;
        double sc;
        Status st = build_atomic_density(S, rho0, sc);
        if (!st.ok()) return set_err(st.code, st.msg);
        HIP_TRY(hipMemcpy(F.rho, rho0.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemset(F.phi, 0, n * sizeof(double)));
    memset(res, 0, sizeof *res);
    if (m > kMaxHist) return set_err(SQ_ERR_ARG, "MIXING_HISTORY above the supported maximum (16)");
    double t_sq = 0.0, t_poisson = 0.0, rule_sum = 0.0;
    // the control block: iteration 1, empty Pulay history
    {
        ScfCtl init;
        memset(&init, 0, sizeof init);
        init.it = 1;
        init.bad = -1;
        HIP_TRY(hipMemcpy(F.ctl, &init, sizeof init, hipMemcpyHostToDevice));
    }
    // The step is captured as one graph after the first (eager) step, unless
    // SQ_NO_GRAPHS is set or the per-node early exit (which takes the previous
    // Fermi level from the host) is on.
    bool use_graph = getenv("SQ_NO_GRAPHS") == nullptr && !(opt.exit_tol > 0);
    if (c->zmarch && c->cfg_B[opt.precision] > 1) {
        if ((rc = ensure_batches(c, c->cfg_B[opt.precision], 0, n))) return rc;
    }
    auto drop_graph = [&]() {
        if (F.step_exec) (void)hipGraphExecDestroy(F.step_exec);
        if (F.step_graph) (void)hipGraphDestroy(F.step_graph);
        F.step_exec = nullptr;
        F.step_graph = nullptr;
    };
    drop_graph();
    struct GraphGuard {
        std::function<void()> f;
        ~GraphGuard() { f(); }
    } gguard{drop_graph};
    int graph_steps = 0;
    for (int it = 1; it <= maxit; it++) {
        if (opt.exit_tol > 0 && it > 1)
            sq_set_early_exit(c, opt.exit_tol, res->lambda_f, 0.0, opt.exit_min > 0 ? opt.exit_min : 30,
                              opt.exit_every > 0 ? opt.exit_every : 5);
        F.force_ready = false;
        F.last_exit_tol = c->exit_tol;
        F.last_exit_lf = c->exit_lf;
        F.last_exit_beta = c->exit_beta;
        F.last_exit_min = c->exit_min;
        F.last_exit_every = c->exit_every;
        StepPlan pl{opt.nonlocal_mode, opt.precision, beta, tol, tolp, opt.kerker, m, pf, !(use_graph && it > 1)};
        if (!pl.eager && !F.step_exec) {
            // capture once; a capture the runtime refuses leaves the loop eager
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_scf_step(c, pl);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (!rc && e == hipSuccess && g && hipGraphInstantiate(&F.step_exec, g, nullptr, nullptr, 0) == hipSuccess) {
                F.step_graph = g;
            } else {
                if (g) (void)hipGraphDestroy(g);
                F.step_exec = nullptr;
                (void)hipGetLastError();
                use_graph = false;
                pl.eager = true;
                rc = SQ_OK;
            }
        }
        if (pl.eager) {
            rc = enqueue_scf_step(c, pl);
        } else {
            HIP_TRY(hipGraphLaunch(F.step_exec, c->stream));
            graph_steps++;
        }
        c->exit_tol = 0.0;
        if (rc) return rc;
        // the step's scalars: control block, dot slots, band energy and -TS
        ScfCtl h;
        double sl[16], eo[2];
        HIP_TRY(hipMemcpyAsync(&h, F.ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(sl, F.s, sizeof sl, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(eo, c->out, sizeof eo, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, F.sev[0], F.sev[1]));
        t_poisson += ms * 1e-3;
        HIP_TRY(hipEventElapsedTime(&ms, F.sev[2], F.sev[3]));
        t_sq += ms * 1e-3;
        if (h.bad >= 0) {
            int stv = 0;
            HIP_TRY(hipMemcpy(&stv, c->status + h.bad, sizeof(int), hipMemcpyDeviceToHost));
            char msg[160];
            snprintf(msg, sizeof msg, "SCF pass: node %d (rank-box id) status %d", h.bad, stv);
            return set_err(stv == 7 ? SQ_ERR_BREAKDOWN : stv == 5 ? SQ_ERR_QL : SQ_ERR_ARG, msg);
        }
        if (h.pfail) return set_err(SQ_ERR_RANGE, "Poisson: conjugate gradients did not reach TOL_POISSON");
        if (h.fbad) return set_err(SQ_ERR_RANGE, "Fermi level not bracketed by the spectrum");
        if (h.kfail) return set_err(SQ_ERR_RANGE, "Kerker: conjugate gradients did not converge");
        c->have_run = true;
        c->last_begin = 0;
        c->last_count = n;
        if (c->last_exit) {
            std::vector<int> ns(n);
            if ((rc = sq_get_rule_sizes(c, ns.data()))) return rc;
            long tot = 0;
            for (int v : ns) tot += v;
            rule_sum += (double)tot / n;
        } else rule_sum += S.nqp;
        const double lf = h.lf, eb = eo[0], mts = eo[1];
        F.lambda_f = lf;
        F.last_mode = opt.nonlocal_mode;
        F.last_prec = opt.precision;
        F.force_ready = true;
        const double exc = sl[6], evxc = sl[7], pb = sl[8], prho = sl[9], rr = sl[10], oo = sl[11];
        res->iterations = it;
        res->residual = std::sqrt(rr / oo);
        res->lambda_f = lf;
        res->e_band = eb;
        res->minus_ts = mts;
        res->e_xc = exc * dV;
        res->e_vxc = evxc * dV;
        res->e_el = 0.5 * (pb - prho) * dV;
        res->e_self = F.eself;
        res->e_total = eb + res->e_xc - res->e_vxc + res->e_el - F.eself + mts;
        res->poisson_iterations = h.piters;
        res->graph_steps = graph_steps;
        if (opt.tol_report > 0 && res->residual < opt.tol_report && res->report_iterations == 0) {
            res->report_iterations = it;
            res->report_wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            res->report_e_total = res->e_total;
        }
        if (opt.verbose)
            fprintf(stderr, "scf %3d  residual %.3e  E %.12f  lambda_f %.10f  poisson %d%s\n", it, res->residual,
                    res->e_total, lf, h.piters, pl.eager ? "" : "  (graph)");
        if (h.conv) { res->converged = 1; break; }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    res->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->mean_rule_size = res->iterations > 0 ? rule_sum / res->iterations : 0.0;
    res->sq_s = t_sq;
    res->poisson_s = t_poisson;
    return SQ_OK;
}

int sq_get_cell_field(sq_ctx *c, int which, double *out)
{
    if (!c || !out) return set_err(SQ_ERR_ARG, "null argument");
    auto &F = c->scf;
    if (which == SQ_FIELD_B) {
        int G[3], rc;
        if ((rc = cell_dims(c, G))) return rc;
        if (!F.have_b) {
            Status st = build_pseudocharge(c->sys, F.host_b, F.eself, F.b_scale);
            if (!st.ok()) return set_err(st.code, st.msg);
            F.have_b = true;
        }
        memcpy(out, F.host_b.data(), F.host_b.size() * sizeof(double));
        return SQ_OK;
    }
    if (which == SQ_FIELD_RHO_ATOMIC) {
        int G[3], rc;
        if ((rc = cell_dims(c, G))) return rc;
        std::vector<double> r;
        double sc;
        Status st = build_atomic_density(c->sys, r, sc);
        if (!st.ok()) return set_err(st.code, st.msg);
        memcpy(out, r.data(), r.size() * sizeof(double));
        return SQ_OK;
    }
    if (!F.rho) return set_err(SQ_ERR_STATE, "no SCF state yet");
    const double *src = which == SQ_FIELD_RHO ? F.rho : which == SQ_FIELD_RHO_OUT ? F.rho_out : which == SQ_FIELD_PHI ? F.phi
                      : which == SQ_FIELD_VXC ? F.vxc : which == SQ_FIELD_VEFF ? F.veff : nullptr;
    if (!src) return set_err(SQ_ERR_ARG, "unknown field");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, src, F.ng * sizeof(double), hipMemcpyDeviceToHost));
    return SQ_OK;
}

// Forces at the last SCF pass.  Local part on the host from phi (local_forces,
// sq_setup.cpp).  Nonlocal part: the pass is repeated for the nodes that lie
// in some projector box -- same V_eff, precision and early-exit settings, so
// the same tridiagonals -- recording alpha of those boxes' images at every
// Lanczos step; then c = f(T) e_1 per node and the contraction with grad chi at
// the node (sq_forces.hip).  Nodes go in chunks sized to keep the record
// under 1 GiB; per-node results are added per atom in node order.
int sq_forces(sq_ctx *c, double *forces, double *f_local, double *f_nonlocal, sq_force_info *info)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    auto &F = c->scf;
    if (!F.force_ready) return set_err(SQ_ERR_STATE, "no SCF pass on this context (call sq_scf first)");
    if (F.last_mode != SQ_NL_KB) return set_err(SQ_ERR_ARG, "forces need the Kleinman-Bylander operator (SQ_NL_KB)");
    const System &S = c->sys;
    const int nq = S.nqp;
    if (nq < 1 || nq > kMaxNqp) return set_err(SQ_ERR_ARG, "forces need 1 <= N_qp <= 256");
    HIP_TRY(hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    const int nsc = (int)S.sctype.size();
    const bool want_local = forces || f_local, want_nl = forces || f_nonlocal;
    // ---- local part
    std::vector<double> Floc((size_t)3 * S.nAtoms, 0.0);
    if (want_local) {
        std::vector<double> phi(F.ng);
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(phi.data(), F.phi, F.ng * sizeof(double), hipMemcpyDeviceToHost));
        Status st = local_forces(S, phi.data(), F.b_scale, Floc);
        if (!st.ok()) return set_err(st.code, st.msg);
    }
    // ---- nonlocal part: nodes inside a projector box, and the most boxes at a node
    const auto t1 = std::chrono::steady_clock::now();
    const long lx = S.box[3], ly = S.box[4], nn = S.n_nodes();
    std::vector<int> cnt(want_nl ? nn : 0, 0);
    for (const Image &im : S.images) {
        if (!want_nl) break;
        int lo[3], hi[3];
        bool empty = false;
        for (int d = 0; d < 3; d++) {
            lo[d] = std::max(im.lo[d], S.box[d]);
            hi[d] = std::min(im.hi[d], S.box[d] + S.box[3 + d] - 1);
            empty = empty || lo[d] > hi[d];
        }
        if (empty) continue;
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++)
                    cnt[((long)(k - S.box[2]) * ly + (j - S.box[1])) * lx + (i - S.box[0])]++;
    }
    std::vector<int> ids;
    int slots = 0;
    for (long q = 0; q < nn && want_nl; q++)
        if (cnt[q]) {
            ids.push_back((int)q);
            slots = std::max(slots, cnt[q]);
        }
    if (slots > kRecMax) return set_err(SQ_ERR_RANGE, "more than 32 projector boxes hold one node");
    std::vector<double> Fnl((size_t)3 * std::max(nsc, S.nAtoms), 0.0);
    const long nlist = (long)ids.size();
    int chunks = 0;
    if (nlist > 0) {
        const int L = c->nlm_tmpl;
        const size_t per_node = (size_t)nq * slots * L * sizeof(double) + (size_t)nq * sizeof(double) +
                                (size_t)slots * (sizeof(int) + 3 * sizeof(double)) + sizeof(int);
        const long chunk = std::min(nlist, std::max<long>(256, (long)((size_t)1 << 30) / (long)per_node));
        std::vector<double> tab;
        std::vector<TabType> types;
        if (pack_projector_splines(S, tab, types)) return set_err(SQ_ERR_ARG, "more than 16 projector channels");
        int *dids = nullptr, *drimg = nullptr;
        double *drec = nullptr, *dcoef = nullptr, *dz = nullptr, *dout = nullptr, *dtab = nullptr;
        TabType *dtyp = nullptr;
        const long zg = std::min<long>(chunk, kFermiGrid);
        struct Frees {
            std::vector<void *> p;
            ~Frees() { for (void *q : p) if (q) (void)hipFree(q); }
        } frees;
        auto alloc = [&](void **dst, size_t bytes) -> int {
            HIP_TRY(hipMalloc(dst, std::max<size_t>(bytes, 8)));
            frees.p.push_back(*dst);
            return SQ_OK;
        };
        int rc;
        if ((rc = alloc((void **)&dids, chunk * sizeof(int))) || (rc = alloc((void **)&drimg, chunk * slots * sizeof(int))) ||
            (rc = alloc((void **)&drec, (size_t)chunk * nq * slots * L * sizeof(double))) ||
            (rc = alloc((void **)&dcoef, (size_t)chunk * nq * sizeof(double))) ||
            (rc = alloc((void **)&dz, (size_t)zg * nq * nq * sizeof(double))) ||
            (rc = alloc((void **)&dout, (size_t)chunk * slots * 3 * sizeof(double))) ||
            (rc = alloc((void **)&dtab, tab.size() * sizeof(double))) ||
            (rc = alloc((void **)&dtyp, types.size() * sizeof(TabType))))
            return rc;
        HIP_TRY(hipMemcpy(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dtyp, types.data(), types.size() * sizeof(TabType), hipMemcpyHostToDevice));
        // the pass runs with the SCF pass's early-exit settings; the caller's come back after
        struct ExitGuard {
            sq_ctx *c;
            double tol, lf, beta;
            int mn, ev;
            ~ExitGuard() { c->exit_tol = tol; c->exit_lf = lf; c->exit_beta = beta; c->exit_min = mn; c->exit_every = ev; }
        } guard{c, c->exit_tol, c->exit_lf, c->exit_beta, c->exit_min, c->exit_every};
        c->exit_tol = F.last_exit_tol;
        c->exit_lf = F.last_exit_lf;
        c->exit_beta = F.last_exit_beta;
        c->exit_min = F.last_exit_min;
        c->exit_every = F.last_exit_every;
        const int prec = F.last_prec;
        std::vector<double> hout;
        std::vector<int> himg;
        for (long b0 = 0; b0 < nlist; b0 += chunk) {
            const long cn = std::min(chunk, nlist - b0);
            HIP_TRY(hipMemcpyAsync(dids, ids.data() + b0, cn * sizeof(int), hipMemcpyHostToDevice, c->stream));
            bool exit_used = false;
            if (c->zmarch) {
                KParams P = make_params(c, SQ_NL_KB, prec, 0, cn, true);
                P.node_list = dids;
                P.rec = drec;
                P.rec_img = drimg;
                P.rec_slots = slots;
                exit_used = P.exit_tol > 0;
                int grid = c->zgrid[prec];
                if (cn < grid) grid = std::max<long>(8, (cn + 7) & ~7L);
                HIP_TRY(hipMemsetAsync(c->queue, 0, 8 * sizeof(int), c->stream));
                HIP_TRY(launch_zmarch(P, prec, grid, c->zlds, c->stream));
            } else {
                KParams P = make_params(c, SQ_NL_KB, prec, 0, cn);
                P.node_list = dids;
                P.rec = drec;
                P.rec_img = drimg;
                P.rec_slots = slots;
                HIP_TRY(launch_lanczos(P, prec, (int)std::min<long>(c->grid, cn), c->lds, c->stream));
            }
            const int *ns = exit_used ? c->nsteps : nullptr;
            HIP_TRY(launch_fermi_column(c->a, c->b, nq + 1, dids, ns, cn, nq, S.beta, F.lambda_f, dcoef, dz, c->status,
                                        c->stream));
            ForceParams fp;
            memset(&fp, 0, sizeof fp);
            fp.ids = dids;
            fp.count = cn;
            fp.slots = slots;
            fp.nlm = L;
            fp.nqp = nq;
            fp.coef = dcoef;
            fp.rec = drec;
            fp.rec_img = drimg;
            fp.nsteps = ns;
            fp.img_cen = c->img_cen;
            fp.img_type = c->img_type;
            fp.types = dtyp;
            fp.tab = dtab;
            for (int d = 0; d < 3; d++) {
                fp.h[d] = S.h[d];
                fp.range[d] = S.range[d];
                fp.box0[d] = S.box[d];
                fp.box_n[d] = S.box[3 + d];
            }
            fp.out = dout;
            fp.status = c->status;
            HIP_TRY(launch_nl_force(fp, c->stream));
            hout.resize((size_t)cn * slots * 3);
            himg.resize((size_t)cn * slots);
            HIP_TRY(hipMemcpyAsync(hout.data(), dout, hout.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipMemcpyAsync(himg.data(), drimg, himg.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            std::vector<int> stv(nn);
            HIP_TRY(hipMemcpy(stv.data(), c->status, nn * sizeof(int), hipMemcpyDeviceToHost));
            for (long r = 0; r < cn; r++) {
                const int id = ids[b0 + r];
                if (stv[id]) {
                    char msg[160];
                    snprintf(msg, sizeof msg, "force pass: node %d (rank-box id) status %d", id, stv[id]);
                    const int code = stv[id] == 7 ? SQ_ERR_BREAKDOWN : stv[id] == 5 ? SQ_ERR_QL : stv[id] == 3 ? SQ_ERR_RANGE : SQ_ERR_ARG;
                    return set_err(code, msg);
                }
                for (int sl = 0; sl < slots; sl++) {
                    const int g = himg[(size_t)r * slots + sl];
                    if (g < 0) continue;
                    const int atom = S.images[g].atom;
                    for (int d = 0; d < 3; d++) Fnl[3 * atom + d] += hout[((size_t)r * slots + sl) * 3 + d];
                }
            }
            chunks++;
        }
    }
    const auto t2 = std::chrono::steady_clock::now();
    for (int q = 0; q < 3 * S.nAtoms; q++) {
        if (forces) forces[q] = Floc[q] + Fnl[q];
        if (f_local) f_local[q] = Floc[q];
        if (f_nonlocal) f_nonlocal[q] = Fnl[q];
    }
    if (info) {
        info->lambda_f = F.lambda_f;
        info->wall_s = std::chrono::duration<double>(t2 - t0).count();
        info->nonlocal_s = std::chrono::duration<double>(t2 - t1).count();
        info->nodes = nlist;
        info->slots = slots;
        info->chunks = chunks;
    }
    return SQ_OK;
}

}  // extern "C"

// ------------------------------------------------------------ initial-density network
// The model and its training set live beside the contexts: descriptors and
// targets (rho - rho_atomic) of converged SCFs are gathered on the host, then
// trained on the device in full batches (sq_mlp.hip).

struct sq_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int K = kMlpK, H = kMlpH, np = 0;
    double rf = kMlpRf;
    float *p = nullptr, *m = nullptr, *v = nullptr;
    double *g = nullptr;
    int t = 0;
    bool have_norm = false;
    std::vector<float> mean, rstd;
    double tscale = 1.0;
    std::vector<float> X, T;
};

namespace {

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() { for (void *q : p) if (q) (void)hipFree(q); }
    template <typename T>
    int get(T **dst, size_t n)
    {
        HIP_TRY(hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)));
        p.push_back(*dst);
        return SQ_OK;
    }
};

// every image (shifts 2R) of the atoms that can lie within rf of the cell,
// the positions first wrapped into [-R, R)
std::vector<double4> mlp_images(const System &S, double rf)
{
    int n[3];
    for (int d = 0; d < 3; d++) n[d] = (int)std::ceil(rf / (2.0 * S.range[d])) + 1;
    std::vector<double4> out;
    for (int a = 0; a < S.nAtoms; a++) {
        double A[3];
        for (int d = 0; d < 3; d++) {
            const double L = 2.0 * S.range[d];
            A[d] = S.atompos[3 * a + d] - L * std::floor((S.atompos[3 * a + d] + S.range[d]) / L);
        }
        for (int i = -n[0]; i <= n[0]; i++)
            for (int j = -n[1]; j <= n[1]; j++)
                for (int k = -n[2]; k <= n[2]; k++)
                    out.push_back(make_double4(A[0] + 2.0 * S.range[0] * i, A[1] + 2.0 * S.range[1] * j,
                                               A[2] + 2.0 * S.range[2] * k, 0.0));
    }
    return out;
}

// descriptors of the context's cell into a device buffer (ng * K floats)
int mlp_descriptors_dev(sq_ctx *c, double rf, float *dout)
{
    const System &S = c->sys;
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    std::vector<double4> img = mlp_images(S, rf);
    DevBufs B;
    double4 *dimg;
    int *dover, over = 0;
    if ((rc = B.get(&dimg, img.size())) || (rc = B.get(&dover, 1))) return rc;
    HIP_TRY(hipMemcpy(dimg, img.data(), img.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(dover, 0, sizeof(int)));
    const double h[3] = {S.h[0], S.h[1], S.h[2]}, R[3] = {S.range[0], S.range[1], S.range[2]};
    HIP_TRY(launch_mlp_features(dimg, (int)img.size(), G, h, R, kMlpK, rf, dout, dover, c->stream));
    HIP_TRY(hipMemcpyAsync(&over, dover, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (over) return set_err(SQ_ERR_RANGE, "descriptors: more atoms within the cutoff of a point than the kernel holds");
    return SQ_OK;
}

}  // namespace

extern "C" {

int sq_mlp_create(int device, unsigned seed, sq_mlp **out)
{
    if (!out) return set_err(SQ_ERR_ARG, "null argument");
    *out = nullptr;
    auto *m = new sq_mlp();
    m->device = device;
    m->np = mlp_param_count(m->K, m->H);
    m->mean.assign(m->K, 0.0f);
    m->rstd.assign(m->K, 1.0f);
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)); biases zero (xorshift generator)
    std::vector<float> p(m->np, 0.0f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (double)(st >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    const int K = m->K, H = m->H;
    auto fill = [&](int off, int n, int fi, int fo) {
        const double a = std::sqrt(6.0 / (fi + fo));
        for (int q = 0; q < n; q++) p[off + q] = (float)(a * uni());
    };
    fill(0, H * K, K, H);
    fill(H * K + H, H * H, H, H);
    fill(H * K + H + H * H + H, H, H, 1);
    int rc;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) rc = set_err(SQ_ERR_DEVICE, "no device");
    else if (hipMalloc((void **)&m->p, m->np * sizeof(float)) != hipSuccess ||
             hipMalloc((void **)&m->m, m->np * sizeof(float)) != hipSuccess ||
             hipMalloc((void **)&m->v, m->np * sizeof(float)) != hipSuccess ||
             hipMalloc((void **)&m->g, m->np * sizeof(double)) != hipSuccess)
        rc = set_err(SQ_ERR_ALLOC, "network buffers");
    else if (hipMemcpy(m->p, p.data(), m->np * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemset(m->m, 0, m->np * sizeof(float)) != hipSuccess || hipMemset(m->v, 0, m->np * sizeof(float)) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "network upload");
    else rc = SQ_OK;
    if (rc) { sq_mlp_destroy(m); return rc; }
    *out = m;
    return SQ_OK;
}

int sq_mlp_destroy(sq_mlp *m)
{
    if (!m) return SQ_OK;
    (void)hipSetDevice(m->device);
    for (void *q : {(void *)m->p, (void *)m->m, (void *)m->v, (void *)m->g})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return SQ_OK;
}

int sq_mlp_get_params(sq_mlp *m, float *params, float *norm)
{
    if (!m) return set_err(SQ_ERR_ARG, "null model");
    HIP_TRY(hipSetDevice(m->device));
    if (params) HIP_TRY(hipMemcpy(params, m->p, m->np * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) {
        for (int q = 0; q < m->K; q++) { norm[q] = m->mean[q]; norm[m->K + q] = m->rstd[q]; }
        norm[2 * m->K] = (float)m->tscale;
    }
    return SQ_OK;
}

int sq_mlp_set_params(sq_mlp *m, const float *params, const float *norm)
{
    if (!m) return set_err(SQ_ERR_ARG, "null model");
    HIP_TRY(hipSetDevice(m->device));
    if (params) HIP_TRY(hipMemcpy(m->p, params, m->np * sizeof(float), hipMemcpyHostToDevice));
    if (norm) {
        for (int q = 0; q < m->K; q++) { m->mean[q] = norm[q]; m->rstd[q] = norm[m->K + q]; }
        m->tscale = norm[2 * m->K];
        m->have_norm = true;
    }
    HIP_TRY(hipMemset(m->m, 0, m->np * sizeof(float)));
    HIP_TRY(hipMemset(m->v, 0, m->np * sizeof(float)));
    m->t = 0;
    return SQ_OK;
}

int sq_mlp_descriptors(sq_ctx *c, float *out)
{
    if (!c || !out) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    HIP_TRY(hipSetDevice(c->device));
    const long ng = (long)G[0] * G[1] * G[2];
    DevBufs B;
    float *d;
    if ((rc = B.get(&d, ng * kMlpK)) || (rc = mlp_descriptors_dev(c, kMlpRf, d))) return rc;
    HIP_TRY(hipMemcpy(out, d, ng * kMlpK * sizeof(float), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_mlp_add_sample(sq_mlp *m, sq_ctx *c) { return sq_mlp_add_sample_every(m, c, 1); }

int sq_mlp_add_sample_every(sq_mlp *m, sq_ctx *c, int stride)
{
    if (!m || !c || stride < 1) return set_err(SQ_ERR_ARG, "bad argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    if (!c->scf.force_ready) return set_err(SQ_ERR_STATE, "no SCF density on this context (call sq_scf first)");
    const System &S = c->sys;
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    const long ng = (long)G[0] * G[1] * G[2];
    std::vector<float> f(ng * kMlpK);
    if ((rc = sq_mlp_descriptors(c, f.data()))) return rc;
    std::vector<double> rho(ng), rat;
    double sc;
    if ((rc = sq_get_cell_field(c, SQ_FIELD_RHO, rho.data()))) return rc;
    Status st = build_atomic_density(S, rat, sc);
    if (!st.ok()) return set_err(st.code, st.msg);
    for (long q = 0; q < ng; q += stride) {
        m->X.insert(m->X.end(), f.begin() + q * kMlpK, f.begin() + (q + 1) * kMlpK);
        m->T.push_back((float)(rho[q] - rat[q]));
    }
    return SQ_OK;
}

// Activations, deltas and partial sums of one pass over N rows.
struct MlpScratch {
    float *H1 = nullptr, *H2 = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *D1 = nullptr, *D2 = nullptr,
          *D3 = nullptr;
    double *part = nullptr, *lsum = nullptr;
    int nchunk = 1;
    int alloc(DevBufs &B, long N, int H)
    {
        nchunk = (int)std::max<long>(1, std::min<long>(128, N / 256));
        int rc;
        if ((rc = B.get(&H1, N * H)) || (rc = B.get(&H2, N * H)) || (rc = B.get(&Z1, N * H)) || (rc = B.get(&Z2, N * H)) ||
            (rc = B.get(&Y, N)) || (rc = B.get(&D1, N * H)) ||
            (rc = B.get(&D2, N * H)) || (rc = B.get(&D3, N)) || (rc = B.get(&part, (size_t)nchunk * H * H)) ||
            (rc = B.get(&lsum, 1)))
            return rc;
        return SQ_OK;
    }
};

// L and dL/dp (double) of the current parameters on normalized data already on
// the device (X: N x K, T: N), with the activations and deltas in the scratch W
static int mlp_loss_grad_dev(sq_mlp *m, const float *X, const float *T, long N, double *loss, const MlpScratch &W)
{
    const int K = m->K, H = m->H;
    float *H1 = W.H1, *H2 = W.H2, *Y = W.Y, *D1 = W.D1, *D2 = W.D2, *D3 = W.D3;
    double *part = W.part, *lsum = W.lsum;
    const int nchunk = W.nchunk;
    HIP_TRY(launch_mlp_forward(X, N, m->p, nullptr, nullptr, H1, H2, W.Z1, W.Z2, Y, 0, m->stream));
    HIP_TRY(launch_mlp_deltas(W.Z1, W.Z2, Y, T, N, m->p, D1, D2, D3, m->stream));
    double *g = m->g;
    HIP_TRY(launch_outer_sum(D1, H, X, K, N, part, nchunk, g, m->stream));                          // W1
    HIP_TRY(launch_outer_sum(D1, H, nullptr, 1, N, part, nchunk, g + H * K, m->stream));            // b1
    HIP_TRY(launch_outer_sum(D2, H, H1, H, N, part, nchunk, g + H * K + H, m->stream));             // W2
    HIP_TRY(launch_outer_sum(D2, H, nullptr, 1, N, part, nchunk, g + H * K + H + H * H, m->stream)); // b2
    HIP_TRY(launch_outer_sum(H2, H, D3, 1, N, part, nchunk, g + H * K + 2 * H + H * H, m->stream));  // w3
    HIP_TRY(launch_outer_sum(D3, 1, nullptr, 1, N, part, nchunk, g + H * K + 3 * H + H * H, m->stream)); // b3
    if (loss) {
        HIP_TRY(launch_outer_sum(D3, 1, D3, 1, N, part, nchunk, lsum, m->stream));
        double s2;
        HIP_TRY(hipMemcpyAsync(&s2, lsum, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HIP_TRY(hipStreamSynchronize(m->stream));
        *loss = 0.5 * (double)N * s2;   // d3 = e / N
    }
    return SQ_OK;
}

int sq_mlp_loss_grad(sq_mlp *m, const float *X, const float *T, long N, double *loss, double *grad)
{
    if (!m || !X || !T || N <= 0) return set_err(SQ_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(m->device));
    DevBufs B;
    float *dX, *dT;
    int rc;
    if ((rc = B.get(&dX, N * m->K)) || (rc = B.get(&dT, N))) return rc;
    HIP_TRY(hipMemcpy(dX, X, N * m->K * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT, T, N * sizeof(float), hipMemcpyHostToDevice));
    MlpScratch W;
    if ((rc = W.alloc(B, N, m->H)) || (rc = mlp_loss_grad_dev(m, dX, dT, N, loss, W))) return rc;
    if (grad) HIP_TRY(hipMemcpy(grad, m->g, m->np * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return SQ_OK;
}

int sq_mlp_forward(sq_mlp *m, const float *X, long N, int bf16, int normalize, float *Y)
{
    if (!m || !X || !Y || N <= 0) return set_err(SQ_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(m->device));
    DevBufs B;
    float *dX, *dY, *dmean, *drstd;
    int rc;
    if ((rc = B.get(&dX, N * m->K)) || (rc = B.get(&dY, N)) || (rc = B.get(&dmean, m->K)) || (rc = B.get(&drstd, m->K)))
        return rc;
    HIP_TRY(hipMemcpy(dX, X, N * m->K * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dmean, m->mean.data(), m->K * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(drstd, m->rstd.data(), m->K * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(launch_mlp_forward(dX, N, m->p, normalize ? dmean : nullptr, normalize ? drstd : nullptr, nullptr, nullptr,
                               nullptr, nullptr, dY, bf16, m->stream));
    HIP_TRY(hipMemcpyAsync(Y, dY, N * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return SQ_OK;
}

int sq_mlp_train(sq_mlp *m, int epochs, double lr, double *loss)
{
    if (!m || epochs < 0) return set_err(SQ_ERR_ARG, "bad argument");
    const long N = (long)m->T.size();
    if (N == 0) return set_err(SQ_ERR_STATE, "no training samples (sq_mlp_add_sample first)");
    HIP_TRY(hipSetDevice(m->device));
    const int K = m->K;
    if (!m->have_norm) {   // descriptor standardization and target scale from the first training set
        for (int q = 0; q < K; q++) {
            double s = 0, s2 = 0;
            for (long n = 0; n < N; n++) { const double x = m->X[n * K + q]; s += x; s2 += x * x; }
            const double mu = s / N, var = std::max(0.0, s2 / N - mu * mu);
            m->mean[q] = (float)mu;
            m->rstd[q] = (float)(1.0 / std::max(std::sqrt(var), 1e-6));
        }
        double s2 = 0;
        for (long n = 0; n < N; n++) s2 += (double)m->T[n] * m->T[n];
        m->tscale = std::max(std::sqrt(s2 / N), 1e-12);
        m->have_norm = true;
    }
    std::vector<float> Xn(N * K), Tn(N);
    for (long n = 0; n < N; n++) {
        for (int q = 0; q < K; q++) Xn[n * K + q] = (m->X[n * K + q] - m->mean[q]) * m->rstd[q];
        Tn[n] = (float)(m->T[n] / m->tscale);
    }
    DevBufs B;
    float *dX, *dT;
    int rc;
    if ((rc = B.get(&dX, N * K)) || (rc = B.get(&dT, N))) return rc;
    HIP_TRY(hipMemcpy(dX, Xn.data(), N * K * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT, Tn.data(), N * sizeof(float), hipMemcpyHostToDevice));
    MlpScratch W;
    if ((rc = W.alloc(B, N, m->H))) return rc;
    for (int e = 0; e < epochs; e++) {
        if ((rc = mlp_loss_grad_dev(m, dX, dT, N, nullptr, W))) return rc;
        HIP_TRY(launch_adam(m->p, m->g, m->m, m->v, m->np, (float)lr, ++m->t, m->stream));
        if ((e & 63) == 63) HIP_TRY(hipStreamSynchronize(m->stream));
    }
    if (loss && (rc = mlp_loss_grad_dev(m, dX, dT, N, loss, W))) return rc;
    HIP_TRY(hipStreamSynchronize(m->stream));
    return SQ_OK;
}

int sq_mlp_guess(sq_mlp *m, sq_ctx *c, int bf16, double *rho0)
{
    if (!m || !c || !rho0) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    const System &S = c->sys;
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    const long ng = (long)G[0] * G[1] * G[2];
    std::vector<float> f(ng * kMlpK), y(ng);
    if ((rc = sq_mlp_descriptors(c, f.data())) || (rc = sq_mlp_forward(m, f.data(), ng, bf16, 1, y.data()))) return rc;
    std::vector<double> rat;
    double sc;
    Status st = build_atomic_density(S, rat, sc);
    if (!st.ok()) return set_err(st.code, st.msg);
    double tot = 0.0;
    for (long q = 0; q < ng; q++) {
        rho0[q] = std::max(0.0, rat[q] + m->tscale * (double)y[q]);
        tot += rho0[q];
    }
    tot *= S.dV;
    if (!(tot > 0)) return set_err(SQ_ERR_RANGE, "network density has no charge");
    const double fsc = (double)S.noetot / tot;
    for (long q = 0; q < ng; q++) rho0[q] *= fsc;
    return SQ_OK;
}

int sq_set_veff_cell(sq_ctx *c, const double *veff)
{
    if (!c || !veff) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    const System &S = c->sys;
    int G[3] = {S.npts[0] - 1, S.npts[1] - 1, S.npts[2] - 1};
    const long ng = (long)G[0] * G[1] * G[2];
    HIP_TRY(hipSetDevice(c->device));
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, ng * sizeof(double)));
    const int ext0[3] = {S.ext0[0], S.ext0[1], S.ext0[2]}, extn[3] = {S.extn[0], S.extn[1], S.extn[2]};
    int rc = SQ_OK;
    if (hipMemcpy(d, veff, ng * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        launch_veff_to_diag(d, S.cdiag, ext0, extn, G, c->diag64, c->diag32, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "V_eff upload failed");
    (void)hipFree(d);
    c->have_run = false;
    return rc;
}

int sq_xc(sq_ctx *c, const double *rho, long n, double *eps, double *vxc)
{
    if (!c || !rho || !eps || !vxc || n <= 0) return set_err(SQ_ERR_ARG, "bad argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    HIP_TRY(hipSetDevice(c->device));
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, 3 * n * sizeof(double)));
    int rc = SQ_OK;
    if (hipMemcpy(d, rho, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        launch_xc(d, n, d + n, d + 2 * n, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(eps, d + n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(vxc, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_err(SQ_ERR_DEVICE, "xc device step failed");
    (void)hipFree(d);
    return rc;
}

int sq_poisson(sq_ctx *c, const double *f, double tol, double *phi, int *iters)
{
    if (!c || !f || !phi) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if ((rc = scf_alloc(c, std::max(1, c->scf.hist)))) return rc;
    auto &F = c->scf;
    HIP_TRY(hipMemcpy(F.eps, f, F.ng * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(F.phi, 0, F.ng * sizeof(double)));
    if ((rc = scf_poisson(c, F.eps, tol > 0 ? tol : c->sys.tol_poisson, 20000, iters))) return rc;
    HIP_TRY(hipMemcpy(phi, F.phi, F.ng * sizeof(double), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_kerker(sq_ctx *c, const double *u, double k0, double *y, int *iters)
{
    if (!c || !u || !y) return set_err(SQ_ERR_ARG, "null argument");
    if (!(k0 > 0)) return set_err(SQ_ERR_ARG, "k0 must be positive");
    if (!c->stream) return set_err(SQ_ERR_STATE, "host-only context (created with device < 0)");
    int G[3], rc;
    if ((rc = cell_dims(c, G))) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if ((rc = scf_alloc(c, std::max(1, c->scf.hist)))) return rc;
    auto &F = c->scf;
    HIP_TRY(hipMemcpy(F.ku, u, F.ng * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = scf_kerker(c, F.ku, F.ky, k0, iters))) return rc;
    HIP_TRY(hipMemcpy(y, F.ky, F.ng * sizeof(double), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_set_early_exit(sq_ctx *c, double tol, double lambda_f, double beta, int min_steps, int every)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (tol > 0 && (min_steps < 2 || every < 1)) return set_err(SQ_ERR_ARG, "early exit needs min_steps >= 2, every >= 1");
    c->exit_tol = tol > 0 ? tol : 0.0;
    c->exit_lf = lambda_f;
    c->exit_beta = beta;
    c->exit_min = min_steps;
    c->exit_every = every;
    return SQ_OK;
}

int sq_get_rule_sizes(sq_ctx *c, int *n)
{
    if (!c || !n) return set_err(SQ_ERR_ARG, "null argument");
    if (!c->have_run) return set_err(SQ_ERR_STATE, "no completed sq_run");
    if (!c->last_exit) {
        for (long r = 0; r < c->last_count; r++) n[r] = c->sys.nqp;
        return SQ_OK;
    }
    HIP_TRY(hipMemcpy(n, c->nsteps + c->last_begin, c->last_count * sizeof(int), hipMemcpyDeviceToHost));
    return SQ_OK;
}

int sq_synchronize(sq_ctx *c)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    if (!c->stream) return SQ_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    return SQ_OK;
}

int sq_get_image(sq_ctx *c, int g, double *center, int *box, int *atom, int *type, int *entries, double *chi)
{
    if (!c) return set_err(SQ_ERR_ARG, "null context");
    const System &S = c->sys;
    if (g < 0 || g >= (int)S.images.size()) return set_err(SQ_ERR_ARG, "image index out of range");
    const Image &im = S.images[g];
    if (center) { center[0] = im.X0; center[1] = im.Y0; center[2] = im.Z0; }
    if (box) for (int d = 0; d < 3; d++) { box[d] = im.lo[d]; box[3 + d] = im.hi[d]; }
    if (atom) *atom = im.atom;
    if (type) *type = im.type;
    if (entries) *entries = (int)(im.points() * im.nlm);
    if (chi) {
        const size_t n = (size_t)(im.points() * im.nlm);
        if (!S.chi.empty()) memcpy(chi, S.chi.data() + im.chi_off, n * sizeof(double));
        else {
            HIP_TRY(hipSetDevice(c->device));
            HIP_TRY(hipMemcpy(chi, c->chi64 + im.chi_off, n * sizeof(double), hipMemcpyDeviceToHost));
        }
    }
    return SQ_OK;
}

}  // extern "C"
