"""ctypes front end of the CPU oracle (oracle/sq_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsq_oracle.so")
_lib = None

MAXTYPE = 8
MAXORDER = 10

MODE_REF_COLLAPSED = 0  # Mult_NonlocalVector, lsdft_matvec.cc:25-149
MODE_KB = 1             # Mult_HamiltonianVector, lsdft_matvec.cc:155-293


class OracleError(RuntimeError):
    pass


class Info(C.Structure):
    _fields_ = [
        ("npts", C.c_int * 3), ("nloc", C.c_int * 3), ("order", C.c_int), ("nqp", C.c_int),
        ("ntype", C.c_int), ("nAtoms", C.c_int), ("noetot", C.c_int), ("IP_length", C.c_int),
        ("nIP", C.c_int), ("nc", C.c_int), ("center", C.c_int), ("bc", C.c_int),
        ("range", C.c_double * 3), ("delta", C.c_double * 3), ("dV", C.c_double),
        ("beta", C.c_double), ("rcut", C.c_double),
        ("coeffs", (C.c_double * (MAXORDER + 1)) * 3),
        ("cutoff", (C.c_double * 3) * MAXTYPE),
        ("gamma", (C.c_double * 4) * MAXTYPE), ("rc", (C.c_double * 4) * MAXTYPE),
        ("lmax", C.c_int * MAXTYPE), ("lloc", C.c_int * MAXTYPE), ("psd_size", C.c_int * MAXTYPE),
        ("nimages", C.c_int * MAXTYPE),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        L.sqo_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
        L.sqo_free.argtypes = [vp]
        L.sqo_set_phantom.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.sqo_count_local_overlaps.argtypes = [vp, ip, ip]
        L.sqo_build_nonlocal.argtypes = [vp, ip]
        L.sqo_info.argtypes = [vp, C.POINTER(Info)]
        L.sqo_info_size.restype = C.c_int
        L.sqo_apply_h.argtypes = [vp, C.c_int, dp, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.sqo_apply_nonlocal.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.sqo_set_cell_period.argtypes = [vp, C.c_int]
        L.sqo_get_valence_density.argtypes = [vp, C.c_int, dp]
        L.sqo_scf_parameters.argtypes = [vp, dp]
        L.sqo_pseudocharge.argtypes = [vp, dp, dp, dp]
        L.sqo_atomic_density.argtypes = [vp, dp, dp]
        L.sqo_lanczos.argtypes = [vp, C.c_int, dp, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.sqo_tridiag_nodes_weights.argtypes = [dp, dp, C.c_int, dp, dp, dp, dp]
        L.sqo_nodes_and_weights.argtypes = [vp, C.c_int, dp, C.c_double, ip, C.c_int, dp, dp, dp, dp, C.c_int]
        L.sqo_fast_nodes_and_weights.argtypes = [vp, dp, C.c_double, ip, C.c_int, dp, dp, dp, dp, C.c_int]
        L.sqo_density_fermi_energy.argtypes = [dp, dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                               dp, dp, dp, dp]
        L.sqo_electron_count.argtypes = [dp, dp, C.c_long, C.c_double, C.c_double]
        L.sqo_electron_count.restype = C.c_double
        L.sqo_get_radial.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp]
        L.sqo_get_image.argtypes = [vp, C.c_int, C.c_int, dp, ip, ip, ip]
        L.sqo_get_chi.argtypes = [vp, C.c_int, C.c_int, dp, ip, ip, ip, ip, ip]
        L.sqo_get_ip_index.argtypes = [vp, ip]
        L.sqo_get_atoms.argtypes = [vp, dp]
        L.sqo_get_cube_csr.argtypes = [vp, ip, ip, dp]
        L.sqo_set_nqp.argtypes = [vp, C.c_int]
        L.sqo_fract.argtypes = [C.c_int, C.c_int]
        L.sqo_fract.restype = C.c_double
        L.sqo_spline_yd.argtypes = [dp, dp, dp, C.c_int]
        L.sqo_spline_eval.argtypes = [dp, dp, C.c_int, C.c_double, dp, dp]
        L.sqo_ylm.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, ip]
        L.sqo_ylm.restype = C.c_double
        L.sqo_ylm_grad.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, dp]
        L.sqo_chi_grad.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
        L.sqo_fermi_column.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double, dp]
        L.sqo_local_forces.argtypes = [vp, dp, C.c_double, dp]
        L.sqo_nonlocal_forces.argtypes = [vp, dp, C.c_double, C.c_double, C.c_double, ip, C.c_int, C.c_int, dp, dp]
        assert L.sqo_info_size() == C.sizeof(Info), "oracle Info layout mismatch"
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with oracle code {rc}")


# ----------------------------------------------------------------- tools
def fract(n, k):
    return lib().sqo_fract(n, k)


def spline_yd(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    yd = np.zeros_like(x)
    _check(lib().sqo_spline_yd(_dp(x), _dp(y), _dp(yd), len(x)), "spline_yd")
    return yd


def spline_eval(x, y, yd, r):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    yd = np.ascontiguousarray(yd, np.float64)
    out = C.c_double()
    _check(lib().sqo_spline_eval(_dp(x), _dp(y), len(x), float(r), _dp(yd), C.byref(out)), "spline_eval")
    return out.value


def ylm(x, y, z, l, m):
    err = C.c_int()
    v = lib().sqo_ylm(float(x), float(y), float(z), int(l), int(m), C.byref(err))
    if err.value:
        raise OracleError(f"ylm l={l} m={m}")
    return v


def tridiag_nodes_weights(a, b, n):
    a = np.ascontiguousarray(a[:n], np.float64)
    bb = np.zeros(n, np.float64)
    bb[: min(n, len(b))] = b[: min(n, len(b))]
    lam = np.zeros(n)
    w = np.zeros(n)
    lmin = C.c_double()
    lmax = C.c_double()
    _check(lib().sqo_tridiag_nodes_weights(_dp(a), _dp(bb), n, _dp(lam), _dp(w), C.byref(lmin), C.byref(lmax)), "QL")
    return lam, w, lmin.value, lmax.value


def ylm_grad(x, y, z, l, m):
    g = np.zeros(3)
    lib().sqo_ylm_grad(float(x), float(y), float(z), int(l), int(m), _dp(g))
    return g


def fermi_column(a, b, n, beta, lf):
    """f(T) e_1 for the n x n tridiagonal of Lanczos coefficients a, b."""
    a = np.ascontiguousarray(a[:n], np.float64)
    bb = np.zeros(n, np.float64)
    bb[: min(n, len(b))] = b[: min(n, len(b))]
    c = np.zeros(n)
    _check(lib().sqo_fermi_column(_dp(a), _dp(bb), n, float(beta), float(lf), _dp(c)), "fermi_column")
    return c


def density_fermi_energy(lam, w, beta, nelec, dV):
    lam = np.ascontiguousarray(lam, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    n_nodes, nqp = lam.shape
    rho = np.zeros(n_nodes)
    lf, eb, mts = C.c_double(), C.c_double(), C.c_double()
    _check(lib().sqo_density_fermi_energy(_dp(lam), _dp(w), n_nodes, nqp, beta, nelec, dV, C.byref(lf), _dp(rho),
                                          C.byref(eb), C.byref(mts)), "density_fermi_energy")
    return lf.value, rho, eb.value, mts.value


def electron_count(lam, w, beta, lf):
    lam = np.ascontiguousarray(lam, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    return lib().sqo_electron_count(_dp(lam), _dp(w), lam.size, beta, lf)


# ----------------------------------------------------------------- system
class System:
    """One input deck loaded by the oracle (Initialize of lsdft_initObjs.cc:36-93)."""

    def __init__(self, directory: str, name: str):
        self._h = C.c_void_p()
        _check(lib().sqo_load(directory.encode(), name.encode(), C.byref(self._h)), f"load {directory}/{name}")
        self.box = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.sqo_free(self._h)
            self._h = None

    def info(self) -> Info:
        I = Info()
        lib().sqo_info(self._h, C.byref(I))
        return I

    def set_phantom(self, enable: bool, pos=(0.0, 0.0, 0.0)):
        _check(lib().sqo_set_phantom(self._h, int(enable), *map(float, pos)), "set_phantom")
        self.box = None

    def set_nqp(self, nqp: int):
        lib().sqo_set_nqp(self._h, int(nqp))

    def count_local_overlaps(self, box):
        b = np.ascontiguousarray(box, np.int32)
        out = np.zeros(MAXTYPE, np.int32)
        _check(lib().sqo_count_local_overlaps(self._h, _ip(b), _ip(out)), "count_local_overlaps")
        return out[: self.info().ntype]

    def full_box(self):
        I = self.info()
        return (0, 0, 0, I.npts[0], I.npts[1], I.npts[2])

    def build_nonlocal(self, box=None):
        box = self.full_box() if box is None else tuple(box)
        b = np.ascontiguousarray(box, np.int32)
        _check(lib().sqo_build_nonlocal(self._h, _ip(b)), "build_nonlocal")
        self.box = box

    def radial(self, at, l):
        n = self.info().psd_size[at]
        arrs = [np.zeros(n) for _ in range(6)]
        _check(lib().sqo_get_radial(self._h, at, l, *[_dp(a) for a in arrs]), "radial")
        return dict(zip(["r", "vloc", "udv", "yd_udv", "u", "yd_u"], arrs))

    def images(self, at=0):
        I = self.info()
        out = []
        for c in range(I.nimages[at]):
            cen = np.zeros(3)
            box = np.zeros(6, np.int32)
            atom = C.c_int()
            nd = C.c_int()
            _check(lib().sqo_get_image(self._h, at, c, _dp(cen), _ip(box), C.byref(atom), C.byref(nd)), "image")
            out.append(dict(center=cen, box=box, atom=atom.value, ndim=nd.value))
        return out

    def chi(self, at, c, ndim):
        chi = np.zeros(ndim)
        ids = [np.zeros(ndim, np.int32) for _ in range(5)]
        _check(lib().sqo_get_chi(self._h, at, c, _dp(chi), *[_ip(a) for a in ids]), "chi")
        return chi, dict(zip(["l", "m", "k", "j", "i"], ids))

    def ip_index(self):
        out = np.zeros(self.info().nIP, np.int32)
        lib().sqo_get_ip_index(self._h, _ip(out))
        return out

    def atoms(self):
        n = self.info().nAtoms
        out = np.zeros(3 * n)
        lib().sqo_get_atoms(self._h, _dp(out))
        return out.reshape(n, 3)

    def cube_csr(self):
        nc = self.info().nc
        nnz = lib().sqo_get_cube_csr(self._h, None, None, None)
        rp = np.zeros(nc + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        lib().sqo_get_cube_csr(self._h, _ip(rp), _ip(col), _dp(val))
        return rp, col, val

    @staticmethod
    def _veff(veff):
        if veff is None or np.isscalar(veff):
            return None, (0.333 if veff is None else float(veff))
        return np.ascontiguousarray(veff, np.float64), 0.0

    # ---- SCF fields on the periodic cell (the (N-1)^3 distinct grid points)
    def cell(self):
        I = self.info()
        return (0, 0, 0, I.npts[0] - 1, I.npts[1] - 1, I.npts[2] - 1)

    def set_cell_period(self, enable=True):
        """V_eff fields wrap with period N-1 (fields on the periodic cell) instead of N."""
        lib().sqo_set_cell_period(self._h, 1 if enable else 0)

    def atom_positions(self):
        I = self.info()
        pos = np.zeros(3 * I.nAtoms)
        lib().sqo_get_atoms(self._h, _dp(pos))
        return pos.reshape(-1, 3)

    def valence_density(self, at):
        """The isolated-atom valence density on the radial grid (.pot "Valence charge" / 4 pi r^2)."""
        out = np.zeros(self.info().psd_size[at])
        lib().sqo_get_valence_density(self._h, at, _dp(out))
        return out

    def scf_parameters(self):
        out = np.zeros(6)
        lib().sqo_scf_parameters(self._h, _dp(out))
        return dict(beta=out[0], history=int(out[1]), frequency=int(out[2]), tol_scf=out[3], tol_poisson=out[4],
                    maxit=int(out[5]))

    def pseudocharge(self):
        """(b on the cell, E_self, neutrality scale)."""
        I = self.info()
        b = np.zeros((I.npts[0] - 1) * (I.npts[1] - 1) * (I.npts[2] - 1))
        es, sc = C.c_double(), C.c_double()
        _check(lib().sqo_pseudocharge(self._h, _dp(b), C.byref(es), C.byref(sc)), "pseudocharge")
        return b, es.value, sc.value

    def atomic_density(self):
        """(superposed isolated-atom density on the cell, scaled to N_e; the scale)."""
        I = self.info()
        rho = np.zeros((I.npts[0] - 1) * (I.npts[1] - 1) * (I.npts[2] - 1))
        sc = C.c_double()
        _check(lib().sqo_atomic_density(self._h, _dp(rho), C.byref(sc)), "atomic_density")
        return rho, sc.value

    def apply_nonlocal(self, node, x, y, mode=MODE_REF_COLLAPSED):
        """y + V_nl x (Mult_NonlocalVector / Mult_HamiltonianVector accumulate into V2)."""
        K, J, I_ = node
        x = np.ascontiguousarray(x, np.float64)
        y = np.array(y, np.float64, copy=True)
        _check(lib().sqo_apply_nonlocal(self._h, mode, K, J, I_, _dp(x), _dp(y)), "apply_nonlocal")
        return y

    def apply_h(self, node, x, mode=MODE_REF_COLLAPSED, veff=None):
        K, J, I_ = node
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        field, const = self._veff(veff)
        _check(lib().sqo_apply_h(self._h, mode, _dp(field), const, K, J, I_, _dp(x), _dp(y)), "apply_h")
        return y

    def lanczos(self, node, mode=MODE_REF_COLLAPSED, veff=None):
        K, J, I_ = node
        nqp = self.info().nqp
        a = np.zeros(nqp + 1)
        b = np.zeros(nqp + 1)
        field, const = self._veff(veff)
        _check(lib().sqo_lanczos(self._h, mode, _dp(field), const, K, J, I_, _dp(a), _dp(b)), "lanczos")
        return a, b

    def nodes_and_weights(self, mode=MODE_REF_COLLAPSED, veff=None, node_ids=None, nthreads=1):
        I = self.info()
        box = self.box
        total = box[3] * box[4] * box[5]
        ids = None if node_ids is None else np.ascontiguousarray(node_ids, np.int32)
        n = total if ids is None else len(ids)
        lam = np.zeros((n, I.nqp))
        w = np.zeros((n, I.nqp))
        lmin = np.zeros(n)
        lmax = np.zeros(n)
        field, const = self._veff(veff)
        _check(lib().sqo_nodes_and_weights(self._h, mode, _dp(field), const, _ip(ids), n, _dp(lam), _dp(w),
                                           _dp(lmin), _dp(lmax), int(nthreads)), "nodes_and_weights")
        return lam, w, lmin, lmax

    def fast_nodes_and_weights(self, veff=None, node_ids=None, nthreads=1):
        """The optimised CPU form (oracle/sq_cpu_fast.c) of nodes_and_weights
        with the Kleinman-Bylander operator: matrix-free stencil, per-node
        image lists, FMA; equal to the restatement to rounding."""
        I = self.info()
        box = self.box
        total = box[3] * box[4] * box[5]
        ids = None if node_ids is None else np.ascontiguousarray(node_ids, np.int32)
        n = total if ids is None else len(ids)
        lam = np.zeros((n, I.nqp))
        w = np.zeros((n, I.nqp))
        lmin = np.zeros(n)
        lmax = np.zeros(n)
        field, const = self._veff(veff)
        _check(lib().sqo_fast_nodes_and_weights(self._h, _dp(field), const, _ip(ids), n, _dp(lam), _dp(w),
                                                _dp(lmin), _dp(lmax), int(nthreads)), "fast_nodes_and_weights")
        return lam, w, lmin, lmax

    # ---- forces (the derivative of the SCF energy; see sq_oracle.c)
    def chi_grad(self, at, l, m, x, y, z):
        """(chi_lm, grad chi_lm) of type at at the offset (x, y, z) from the atom."""
        v = C.c_double()
        g = np.zeros(3)
        _check(lib().sqo_chi_grad(self._h, at, l, m, float(x), float(y), float(z), C.byref(v), _dp(g)), "chi_grad")
        return v.value, g

    def local_forces(self, phi, scale):
        """Pseudocharge forces [nAtoms, 3] for the potential phi on the periodic cell."""
        phi = np.ascontiguousarray(phi, np.float64).ravel()
        F = np.zeros(3 * self.info().nAtoms)
        _check(lib().sqo_local_forces(self._h, _dp(phi), float(scale), _dp(F)), "local_forces")
        return F.reshape(-1, 3)

    def nonlocal_forces(self, beta, lf, veff=None, node_ids=None, nthreads=1, per_node=False):
        """Kleinman-Bylander forces [nAtoms, 3] at Fermi level lf (and each node's part)."""
        I = self.info()
        box = self.box
        total = box[3] * box[4] * box[5]
        ids = None if node_ids is None else np.ascontiguousarray(node_ids, np.int32)
        n = total if ids is None else len(ids)
        F = np.zeros(3 * I.nAtoms)
        Fn = np.zeros(n * I.nAtoms * 3) if per_node else None
        field, const = self._veff(veff)
        _check(lib().sqo_nonlocal_forces(self._h, _dp(field), const, float(beta), float(lf), _ip(ids), n,
                                         int(nthreads), _dp(F), _dp(Fn)), "nonlocal_forces")
        if per_node:
            return F.reshape(-1, 3), Fn.reshape(n, I.nAtoms, 3)
        return F.reshape(-1, 3)
