"""ctypes binding of libsqamd.so (include/sq_amd.h).

The library must be built in-tree (python -m sqminiapp_amd.build); loading
fails loudly when it is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsqamd.so")

SQ_NL_REF_COLLAPSED = 0
SQ_NL_KB = 1
SQ_FP64 = 0
SQ_FP32 = 1

STATUS = {0: "SQ_OK", 1: "SQ_ERR_IO", 2: "SQ_ERR_FORMAT", 3: "SQ_ERR_RANGE", 4: "SQ_ERR_ARG", 5: "SQ_ERR_QL",
          6: "SQ_ERR_ALLOC", 7: "SQ_ERR_BREAKDOWN", 8: "SQ_ERR_DEVICE", 9: "SQ_ERR_STATE"}

# every symbol include/sq_amd.h declares
EXPORTS = ["sq_last_error", "sq_default_options", "sq_create_from_deck", "sq_destroy", "sq_get_info", "sq_set_veff",
           "sq_run", "sq_get_nodes_and_weights", "sq_get_lanczos", "sq_nodes_and_weights", "sq_last_timing",
           "sq_apply_h", "sq_apply_nonlocal", "sq_tridiag_nodes_weights", "sq_tridiag_nodes_weights_noalloc", "sq_tune_select", "sq_density_fermi_energy", "sq_synchronize",
           "sq_get_image", "sq_default_scf_options", "sq_scf", "sq_get_cell_field", "sq_set_veff_cell", "sq_xc",
           "sq_poisson", "sq_set_early_exit", "sq_get_rule_sizes", "sq_forces", "sq_mlp_create", "sq_mlp_destroy",
           "sq_mlp_get_params", "sq_mlp_set_params", "sq_mlp_descriptors", "sq_mlp_add_sample", "sq_mlp_add_sample_every",
           "sq_mlp_train",
           "sq_mlp_forward", "sq_mlp_loss_grad", "sq_mlp_guess", "sq_kerker"]

FIELDS = {"rho": 0, "rho_out": 1, "phi": 2, "vxc": 3, "veff": 4, "b": 5, "rho_atomic": 6}


class SQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("phantom_atom", C.c_int), ("phantom_pos", C.c_double * 3), ("rank_box", C.c_int * 6),
                ("nqp_override", C.c_int), ("device", C.c_int)]


class ScfOptions(C.Structure):
    _fields_ = [("nonlocal_mode", C.c_int), ("precision", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("mixing", C.c_double), ("history", C.c_int), ("pulay_frequency", C.c_int),
                ("tol_poisson", C.c_double), ("rho0", C.POINTER(C.c_double)), ("verbose", C.c_int),
                ("exit_tol", C.c_double), ("exit_min", C.c_int), ("exit_every", C.c_int), ("kerker", C.c_double),
                ("tol_report", C.c_double)]


class ScfResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("poisson_iterations", C.c_int),
                ("residual", C.c_double), ("lambda_f", C.c_double), ("e_band", C.c_double),
                ("minus_ts", C.c_double), ("e_xc", C.c_double), ("e_vxc", C.c_double), ("e_el", C.c_double),
                ("e_self", C.c_double), ("e_total", C.c_double), ("wall_s", C.c_double), ("sq_s", C.c_double),
                ("poisson_s", C.c_double), ("mean_rule_size", C.c_double), ("graph_steps", C.c_int),
                ("report_iterations", C.c_int), ("report_wall_s", C.c_double), ("report_e_total", C.c_double)]


class ForceInfo(C.Structure):
    _fields_ = [("lambda_f", C.c_double), ("wall_s", C.c_double), ("nonlocal_s", C.c_double), ("nodes", C.c_long),
                ("slots", C.c_int), ("chunks", C.c_int)]


def deck_grid(directory, name):
    """FD_GRID of a deck (N0, N1, N2)."""
    toks = open(os.path.join(str(directory), f"{name}.inpt")).read().split()
    i = toks.index("FD_GRID:")
    return tuple(int(t) for t in toks[i + 1:i + 4])


class Info(C.Structure):
    _fields_ = [
        ("npts", C.c_int * 3), ("nloc", C.c_int * 3), ("halfwidth", C.c_int * 3), ("order", C.c_int),
        ("nqp", C.c_int), ("ntype", C.c_int), ("nAtoms", C.c_int), ("noetot", C.c_int),
        ("n_images", C.c_int), ("n_sc_atoms", C.c_int), ("ip_length", C.c_int), ("n_nodes", C.c_int),
        ("nc", C.c_int), ("rank_box", C.c_int * 6), ("ext_origin", C.c_int * 3), ("ext_dims", C.c_int * 3),
        ("range", C.c_double * 3), ("delta", C.c_double * 3), ("dV", C.c_double), ("beta", C.c_double),
        ("rcut", C.c_double), ("coeffs", (C.c_double * 11) * 3), ("gamma", (C.c_double * 4) * 8),
        ("cutoff", (C.c_double * 3) * 8), ("chi_entries", C.c_longlong), ("kernel", C.c_int), ("kmax", C.c_int),
        ("workgroup", C.c_int * 2), ("tune_ms", (C.c_float * 2) * 2),
        ("batch", C.c_int * 2), ("gather_trip", C.c_int * 2), ("tune_candidates_ms", (C.c_float * 6) * 2),
        ("union_images", C.c_int * 2),
    ]


_lib = None


def lib():
    """Load libsqamd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sqminiapp_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp, ip, vp, lg = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p, C.c_long
    L.sq_last_error.restype = C.c_char_p
    L.sq_default_options.argtypes = [C.POINTER(Options)]
    L.sq_create_from_deck.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Options), C.POINTER(vp)]
    L.sq_destroy.argtypes = [vp]
    L.sq_get_info.argtypes = [vp, C.POINTER(Info)]
    L.sq_set_veff.argtypes = [vp, dp, C.c_double]
    L.sq_run.argtypes = [vp, C.c_int, C.c_int, lg, lg]
    L.sq_get_nodes_and_weights.argtypes = [vp, dp, dp, dp, dp]
    L.sq_get_lanczos.argtypes = [vp, dp, dp]
    L.sq_nodes_and_weights.argtypes = [vp, C.c_int, C.c_int, lg, lg, dp, dp, dp, dp]
    L.sq_last_timing.argtypes = [vp, dp, dp]
    L.sq_apply_h.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
    L.sq_apply_nonlocal.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
    L.sq_default_scf_options.argtypes = [C.POINTER(ScfOptions)]
    L.sq_scf.argtypes = [vp, C.POINTER(ScfOptions), C.POINTER(ScfResult)]
    L.sq_get_cell_field.argtypes = [vp, C.c_int, dp]
    L.sq_set_veff_cell.argtypes = [vp, dp]
    L.sq_xc.argtypes = [vp, dp, lg, dp, dp]
    L.sq_poisson.argtypes = [vp, dp, C.c_double, dp, C.POINTER(C.c_int)]
    L.sq_set_early_exit.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    L.sq_get_rule_sizes.argtypes = [vp, ip]
    L.sq_forces.argtypes = [vp, dp, dp, dp, C.POINTER(ForceInfo)]
    L.sq_kerker.argtypes = [vp, dp, C.c_double, dp, C.POINTER(C.c_int)]
    fp = C.POINTER(C.c_float)
    L.sq_mlp_create.argtypes = [C.c_int, C.c_uint, C.POINTER(vp)]
    L.sq_mlp_destroy.argtypes = [vp]
    L.sq_mlp_get_params.argtypes = [vp, fp, fp]
    L.sq_mlp_set_params.argtypes = [vp, fp, fp]
    L.sq_mlp_descriptors.argtypes = [vp, fp]
    L.sq_mlp_add_sample.argtypes = [vp, vp]
    L.sq_mlp_add_sample_every.argtypes = [vp, vp, C.c_int]
    L.sq_mlp_train.argtypes = [vp, C.c_int, C.c_double, dp]
    L.sq_mlp_forward.argtypes = [vp, fp, lg, C.c_int, C.c_int, fp]
    L.sq_mlp_loss_grad.argtypes = [vp, fp, fp, lg, dp, dp]
    L.sq_mlp_guess.argtypes = [vp, vp, C.c_int, dp]
    L.sq_tridiag_nodes_weights.argtypes = [vp, dp, dp, C.c_int, lg, dp, dp, dp, dp]
    L.sq_tridiag_nodes_weights_noalloc.argtypes = [vp, vp, vp, C.c_int, lg, vp, vp, vp, vp]
    L.sq_tune_select.argtypes = [C.c_int, ip, C.POINTER(C.c_float), C.c_char_p, C.c_char_p, C.c_char_p, ip]
    L.sq_density_fermi_energy.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, dp]
    L.sq_synchronize.argtypes = [vp]
    L.sq_get_image.argtypes = [vp, C.c_int, dp, ip, ip, ip, ip, dp]
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def tune_select(precision, available, probe_ms, workgroup=None, batch=None, trip=None):
    """The autotuner's choice among its six candidates (sq_tune_select)."""
    av = (C.c_int * 6)(*[int(bool(a)) for a in available])
    ms = (C.c_float * 6)(*[float(x) for x in probe_ms])
    enc = lambda v: None if v is None else str(v).encode()
    out = C.c_int()
    _check(lib().sq_tune_select(int(precision), av, ms, enc(workgroup), enc(batch), enc(trip), C.byref(out)))
    return out.value


def _check(rc):
    if rc != 0:
        raise SQError(rc, lib().sq_last_error().decode())


class Context:
    """One SQ problem on one device (or host-only with device=-1)."""

    def __init__(self, directory, name, *, phantom_atom=False, phantom_pos=(0.0, 0.0, 0.0), rank_box=None,
                 nqp=0, device=0, cell=False):
        """cell=True: the node box is the periodic cell (the (N-1)^3 distinct grid
        points), as the SCF entry points require."""
        if cell:
            N = deck_grid(directory, name)
            rank_box = (0, 0, 0, N[0] - 1, N[1] - 1, N[2] - 1)
        L = lib()
        opt = Options()
        L.sq_default_options(C.byref(opt))
        opt.phantom_atom = int(bool(phantom_atom))
        for d in range(3):
            opt.phantom_pos[d] = float(phantom_pos[d])
        if rank_box is not None:
            for d in range(6):
                opt.rank_box[d] = int(rank_box[d])
        opt.nqp_override = int(nqp)
        opt.device = int(device)
        self._h = C.c_void_p()
        _check(L.sq_create_from_deck(str(directory).encode(), str(name).encode(), C.byref(opt), C.byref(self._h)))
        self.info = self._info()

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _info(self):
        I = Info()
        _check(lib().sq_get_info(self._h, C.byref(I)))
        return I

    @property
    def nqp(self):
        return self.info.nqp

    @property
    def n_nodes(self):
        return self.info.n_nodes

    def set_veff(self, veff):
        if veff is None or np.isscalar(veff):
            _check(lib().sq_set_veff(self._h, None, 0.333 if veff is None else float(veff)))
        else:
            f = np.ascontiguousarray(veff, np.float64).ravel()
            I = self.info
            if f.size != I.npts[0] * I.npts[1] * I.npts[2]:
                raise ValueError("V_eff field size differs from the grid")
            _check(lib().sq_set_veff(self._h, _dp(f), 0.0))

    def run(self, mode=SQ_NL_REF_COLLAPSED, precision=SQ_FP64, begin=0, count=None):
        count = self.n_nodes - begin if count is None else count
        _check(lib().sq_run(self._h, int(mode), int(precision), int(begin), int(count)))
        self._last = (begin, count)

    def results(self):
        begin, count = self._last
        n = self.nqp
        lam = np.zeros((count, n))
        w = np.zeros((count, n))
        lmin = np.zeros(count)
        lmax = np.zeros(count)
        _check(lib().sq_get_nodes_and_weights(self._h, _dp(lam), _dp(w), _dp(lmin), _dp(lmax)))
        return lam, w, lmin, lmax

    def lanczos(self):
        begin, count = self._last
        a = np.zeros((count, self.nqp + 1))
        b = np.zeros((count, self.nqp + 1))
        _check(lib().sq_get_lanczos(self._h, _dp(a), _dp(b)))
        return a, b

    def nodes_and_weights(self, mode=SQ_NL_REF_COLLAPSED, precision=SQ_FP64, begin=0, count=None):
        self.run(mode, precision, begin, count)
        return self.results()

    def synchronize(self):
        _check(lib().sq_synchronize(self._h))

    def timing(self):
        t1, t2 = C.c_double(), C.c_double()
        _check(lib().sq_last_timing(self._h, C.byref(t1), C.byref(t2)))
        return t1.value, t2.value

    def apply_h(self, node, x, mode=SQ_NL_REF_COLLAPSED):
        K, J, I_ = node
        x = np.ascontiguousarray(x, np.float64)
        if x.size != self.info.nc:
            raise ValueError("vector size differs from the cube")
        y = np.zeros_like(x)
        _check(lib().sq_apply_h(self._h, int(mode), int(K), int(J), int(I_), _dp(x), _dp(y)))
        return y

    def apply_nonlocal(self, node, x, y, mode=SQ_NL_REF_COLLAPSED):
        """Returns y + V_nl x, the accumulation of Mult_NonlocalVector / Mult_HamiltonianVector."""
        K, J, I_ = node
        x = np.ascontiguousarray(x, np.float64)
        y = np.array(y, np.float64, copy=True)
        if x.size != self.info.nc or y.size != self.info.nc:
            raise ValueError("vector size differs from the cube")
        _check(lib().sq_apply_nonlocal(self._h, int(mode), int(K), int(J), int(I_), _dp(x), _dp(y)))
        return y

    # ---- SCF on the periodic cell
    def cell_shape(self):
        """(z, y, x) shape of the cell fields."""
        I = self.info
        return (I.npts[2] - 1, I.npts[1] - 1, I.npts[0] - 1)

    def set_early_exit(self, tol, lambda_f, beta=0.0, min_steps=30, every=5):
        """Per-node early exit of the next sq_run calls (tol <= 0: off)."""
        _check(lib().sq_set_early_exit(self._h, float(tol), float(lambda_f), float(beta), int(min_steps), int(every)))

    def rule_sizes(self):
        """Gauss nodes per node of the last run."""
        n = np.zeros(self._last[1], np.int32)
        _check(lib().sq_get_rule_sizes(self._h, n.ctypes.data_as(C.POINTER(C.c_int))))
        return n

    def scf(self, mode=SQ_NL_KB, precision=SQ_FP64, max_iter=0, tol=0.0, mixing=0.0, history=0,
            pulay_frequency=0, tol_poisson=0.0, rho0=None, verbose=False, exit_tol=0.0, exit_min=0, exit_every=0,
            kerker=0.0, tol_report=0.0):
        o = ScfOptions()
        lib().sq_default_scf_options(C.byref(o))
        o.kerker = float(kerker)
        o.tol_report = float(tol_report)
        o.exit_tol, o.exit_min, o.exit_every = float(exit_tol), int(exit_min), int(exit_every)
        o.nonlocal_mode, o.precision, o.max_iter = int(mode), int(precision), int(max_iter)
        o.tol, o.mixing, o.history, o.pulay_frequency = float(tol), float(mixing), int(history), int(pulay_frequency)
        o.tol_poisson, o.verbose = float(tol_poisson), int(bool(verbose))
        keep = None
        if rho0 is not None:
            keep = np.ascontiguousarray(rho0, np.float64).ravel()
            o.rho0 = _dp(keep)
        r = ScfResult()
        _check(lib().sq_scf(self._h, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k, _ in ScfResult._fields_}

    def forces(self):
        """Forces [nAtoms, 3] (Ha/bohr) at the last SCF pass: (total, local, nonlocal, info)."""
        n = 3 * self.info.nAtoms
        F, Fl, Fn = np.zeros(n), np.zeros(n), np.zeros(n)
        fi = ForceInfo()
        _check(lib().sq_forces(self._h, _dp(F), _dp(Fl), _dp(Fn), C.byref(fi)))
        info = {k: getattr(fi, k) for k, _ in ForceInfo._fields_}
        return F.reshape(-1, 3), Fl.reshape(-1, 3), Fn.reshape(-1, 3), info

    def descriptors(self):
        """Initial-density network descriptors of the cell, [points, 32] (i fastest)."""
        out = np.zeros((int(np.prod(self.cell_shape())), Mlp.K), np.float32)
        _check(lib().sq_mlp_descriptors(self._h, _fp(out)))
        return out

    def cell_field(self, name):
        out = np.zeros(int(np.prod(self.cell_shape())))
        _check(lib().sq_get_cell_field(self._h, FIELDS[name], _dp(out)))
        return out.reshape(self.cell_shape())

    def set_veff_cell(self, veff):
        v = np.ascontiguousarray(veff, np.float64).ravel()
        if v.size != int(np.prod(self.cell_shape())):
            raise ValueError("V_eff must be a field on the periodic cell")
        _check(lib().sq_set_veff_cell(self._h, _dp(v)))

    def xc(self, rho):
        r = np.ascontiguousarray(rho, np.float64).ravel()
        eps, v = np.zeros_like(r), np.zeros_like(r)
        _check(lib().sq_xc(self._h, _dp(r), r.size, _dp(eps), _dp(v)))
        return eps.reshape(np.shape(rho)), v.reshape(np.shape(rho))

    def kerker(self, u, k0=1.0):
        """The mixing's Kerker preconditioner applied to a cell field: (y, CG iterations)."""
        ur = np.ascontiguousarray(u, np.float64).ravel()
        y = np.zeros_like(ur)
        its = C.c_int()
        _check(lib().sq_kerker(self._h, _dp(ur), float(k0), _dp(y), C.byref(its)))
        return y.reshape(np.shape(u)), its.value

    def poisson(self, f, tol=0.0):
        fr = np.ascontiguousarray(f, np.float64).ravel()
        phi = np.zeros_like(fr)
        its = C.c_int()
        _check(lib().sq_poisson(self._h, _dp(fr), float(tol), _dp(phi), C.byref(its)))
        return phi.reshape(np.shape(f)), its.value

    def tridiag_nodes_weights(self, diag, subdiag):
        diag = np.ascontiguousarray(diag, np.float64)
        count, n = diag.shape
        sub = np.zeros((count, n))
        s = np.asarray(subdiag, np.float64)
        sub[:, : min(n, s.shape[1])] = s[:, : min(n, s.shape[1])]
        lam = np.zeros((count, n))
        w = np.zeros((count, n))
        lmin = np.zeros(count)
        lmax = np.zeros(count)
        _check(lib().sq_tridiag_nodes_weights(self._h, _dp(diag), _dp(sub), n, count, _dp(lam), _dp(w), _dp(lmin),
                                              _dp(lmax)))
        return lam, w, lmin, lmax

    def tridiag_nodes_weights_noalloc(self, diag, subdiag, n, count, lam, w, lmin, lmax):
        """sq_tridiag_nodes_weights_noalloc on device buffers given by address
        (diag, subdiag, lam, w: count*n doubles; lmin, lmax: count doubles)."""
        _check(lib().sq_tridiag_nodes_weights_noalloc(self._h, diag, subdiag, int(n), int(count), lam, w, lmin, lmax))

    def density_fermi_energy(self, nelec=0.0, beta=0.0, want_rho=True):
        lf, eb, mts = C.c_double(), C.c_double(), C.c_double()
        rho = np.zeros(self.n_nodes) if want_rho else None
        _check(lib().sq_density_fermi_energy(self._h, float(nelec), float(beta), C.byref(lf), _dp(rho), C.byref(eb),
                                             C.byref(mts)))
        return lf.value, rho, eb.value, mts.value

    def image(self, c, with_chi=True):
        cen = np.zeros(3)
        box = np.zeros(6, np.int32)
        atom, typ, ent = C.c_int(), C.c_int(), C.c_int()
        ipt = C.POINTER(C.c_int)
        _check(lib().sq_get_image(self._h, int(c), _dp(cen), box.ctypes.data_as(ipt), C.byref(atom), C.byref(typ),
                                  C.byref(ent), None))
        chi = None
        if with_chi:
            chi = np.zeros(ent.value)
            _check(lib().sq_get_image(self._h, int(c), None, None, None, None, None, _dp(chi)))
        return dict(center=cen, box=box, atom=atom.value, type=typ.value, entries=ent.value, chi=chi)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class Mlp:
    """The initial-density network (sq_mlp_*): descriptors -> correction to
    the superposed atomic density, trained on converged SQ densities."""
    K, H, PARAMS, NORM = 32, 64, 6337, 65

    def __init__(self, device=0, seed=0):
        self._h = C.c_void_p()
        _check(lib().sq_mlp_create(int(device), int(seed), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().sq_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.PARAMS, np.float32)
        n = np.zeros(self.NORM, np.float32)
        _check(lib().sq_mlp_get_params(self._h, _fp(p), _fp(n)))
        return p, n

    def set_params(self, p, norm=None):
        p = np.ascontiguousarray(p, np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32)
        _check(lib().sq_mlp_set_params(self._h, _fp(p), _fp(n)))

    def add_sample(self, ctx, stride=1):
        _check(lib().sq_mlp_add_sample_every(self._h, ctx._h, int(stride)))

    def train(self, epochs, lr=3e-3):
        loss = C.c_double()
        _check(lib().sq_mlp_train(self._h, int(epochs), float(lr), C.byref(loss)))
        return loss.value

    def forward(self, x, bf16=False, normalize=True):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        _check(lib().sq_mlp_forward(self._h, _fp(x), x.shape[0], int(bool(bf16)), int(bool(normalize)), _fp(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        g = np.zeros(self.PARAMS)
        _check(lib().sq_mlp_loss_grad(self._h, _fp(x), _fp(t), x.shape[0], C.byref(loss), _dp(g)))
        return loss.value, g

    def guess(self, ctx, bf16=False):
        """Initial density for ctx's cell (cell shape), for Context.scf(rho0=...)."""
        out = np.zeros(int(np.prod(ctx.cell_shape())))
        _check(lib().sq_mlp_guess(self._h, ctx._h, int(bool(bf16)), _dp(out)))
        return out.reshape(ctx.cell_shape())
