"""SCF restatement of the oracle (TEST INFRASTRUCTURE ONLY: loaded by tests/
and nowhere in the product).

The reference declares its SCF but builds only its pieces: the periodic
Poisson operator
(src/lsdft_initObjs.cc:309-388, KSP tolerance TOL_POISSON at 249-276), the
pseudocharge construction (628-817), the Periodic-Pulay history vectors
(233-245; MIXING_PARAMETER / MIXING_HISTORY / PULAY_FREQUENCY at
src/lsdft_readfiles.cc:219-230) and the declarations of SCF_LSSGQ, SolvePoisson,
Vxc_Calc/Exc_Calc and mix (src/lsdft.h:93,124,138,183,197-204).  This module
writes each from its published method, on numpy:

- LDA exchange-correlation: Slater exchange and the Perdew-Wang 1992
  correlation of the unpolarised gas (Phys. Rev. B 45, 13244), EXCHANGE_
  CORRELATION LDA_PW in the Mg1 deck;
- Poisson: -(1/4pi) Lap_h phi = rho + b on the periodic cell by FFT with the
  exact symbol of the order-12 stencil (an independent method from the
  product's conjugate gradients);
- Periodic Pulay mixing of the density (Banerjee, Suryanarayana & Pask,
  Chem. Phys. Lett. 647, 31 (2016));
- the SCF loop and the Harris-Foulkes energy
  E = E_band + E_xc - int V_xc rho + 1/2 int phi (b - rho) - E_self - TS.
"""
from __future__ import annotations

import numpy as np

C_X = 0.75 * (3.0 / np.pi) ** (1.0 / 3.0)   # eps_x = -C_X rho^(1/3)
PW92 = dict(A=0.031091, a1=0.21370, b1=7.5957, b2=3.5876, b3=1.6382, b4=0.49294)


def pw92_eps_c(rs):
    """(eps_c, d eps_c / d rs) of the unpolarised gas, PW92 eq. (10):
    eps_c = -2A (1 + a1 rs) ln(1 + 1/Q), Q = 2A (b1 rs^1/2 + b2 rs + b3 rs^3/2 + b4 rs^2)."""
    rs = np.asarray(rs, dtype=np.float64)
    A, a1, b1, b2, b3, b4 = (PW92[k] for k in ("A", "a1", "b1", "b2", "b3", "b4"))
    srs = np.sqrt(rs)
    Q = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs)
    dQ = 2.0 * A * (0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2.0 * b4 * rs)
    L = np.log1p(1.0 / Q)
    ec = -2.0 * A * (1.0 + a1 * rs) * L
    with np.errstate(over="ignore"):   # Q (Q + 1) overflows only where the term is 0 anyway
        dec = -2.0 * A * a1 * L + 2.0 * A * (1.0 + a1 * rs) * dQ / (Q * (Q + 1.0))
    return ec, dec


def xc_lda_pw(rho):
    """(eps_xc, v_xc) per point for the unpolarised LDA (Slater + PW92); zero where rho <= 0."""
    rho = np.asarray(rho, dtype=np.float64)
    eps = np.zeros_like(rho)
    v = np.zeros_like(rho)
    m = rho > 0
    r = rho[m]
    ex = -C_X * np.cbrt(r)
    rs = np.cbrt(3.0 / (4.0 * np.pi * r))
    ec, dec = pw92_eps_c(rs)
    eps[m] = ex + ec
    v[m] = 4.0 / 3.0 * ex + ec - rs / 3.0 * dec
    return eps, v


def pz81_eps_c(rs):
    """Perdew-Zunger 1981 unpolarised correlation, an independent fit to the same
    Ceperley-Alder data (used only to pin the PW92 restatement)."""
    rs = np.asarray(rs, dtype=np.float64)
    hi = -0.1423 / (1.0 + 1.0529 * np.sqrt(rs) + 0.3334 * rs)
    lo = 0.0311 * np.log(rs) - 0.048 + 0.0020 * rs * np.log(rs) - 0.0116 * rs
    return np.where(rs >= 1.0, hi, lo)


def laplacian_symbol(coeffs, order, shape):
    """Eigenvalues of the periodic -(1/2) Lap_h (the reference's coeffs_x/y/z,
    src/lsdft_readfiles.cc:453-476) on a grid of the given (z, y, x) shape."""
    s = 0.0
    for axis, d in ((2, 0), (1, 1), (0, 2)):
        n = shape[axis]
        th = 2.0 * np.pi * np.fft.fftfreq(n)
        sym = np.full(n, coeffs[d][0])
        for a in range(1, order + 1):
            sym = sym + 2.0 * coeffs[d][a] * np.cos(a * th)
        shp = [1, 1, 1]
        shp[axis] = n
        s = s + sym.reshape(shp)
    return s


def poisson_fft(coeffs, order, f):
    """phi with -(1/4pi) Lap_h phi = f on the periodic cell, mean(phi) = 0; f
    is (z, y, x) and must be neutral (its mean is dropped)."""
    s = laplacian_symbol(coeffs, order, f.shape)
    fh = np.fft.fftn(f)
    fh[0, 0, 0] = 0.0
    s = s.copy()
    s[0, 0, 0] = 1.0
    # -(1/4pi) Lap = (1/2pi) (-(1/2) Lap): s phi_hat = 2 pi f_hat
    return np.real(np.fft.ifftn(2.0 * np.pi * fh / s))


def kerker_fft(coeffs, order, u, k0):
    """y = A (A + k0^2/2)^-1 u with A = -(1/2) Lap_h on the periodic cell, by
    the stencil's exact symbol s(q): y_hat = s/(s + k0^2/2) u_hat (zero at q = 0)."""
    s = laplacian_symbol(coeffs, order, u.shape)
    return np.real(np.fft.ifftn(np.fft.fftn(u) * s / (s + 0.5 * k0 * k0)))


def apply_laplacian(coeffs, order, x):
    """-(1/2) Lap_h x on the periodic cell, by the stencil (np.roll)."""
    y = (coeffs[0][0] + coeffs[1][0] + coeffs[2][0]) * x
    for a in range(1, order + 1):
        for axis, d in ((2, 0), (1, 1), (0, 2)):
            y = y + coeffs[d][a] * (np.roll(x, a, axis) + np.roll(x, -a, axis))
    return y


class PeriodicPulay:
    """x_{k+1} = x_k + beta f_k, and every p-th step
    x_{k+1} = x_k + beta f_k - (X_k + beta F_k)(F_k^T F_k)^{-1} F_k^T f_k with the
    last m iterate and residual differences X_k, F_k."""

    def __init__(self, beta, m, p):
        self.beta, self.m, self.p = beta, m, p
        self.k = 0
        self.x_prev = self.f_prev = None
        self.X, self.F = [], []

    def step(self, x, g):
        f = g - x
        if self.x_prev is not None:
            self.X.append(x - self.x_prev)
            self.F.append(f - self.f_prev)
            if len(self.X) > self.m:
                self.X.pop(0)
                self.F.pop(0)
        self.x_prev, self.f_prev = x.copy(), f.copy()
        xn = x + self.beta * f
        if (self.k + 1) % self.p == 0 and self.F:
            Fm = np.stack([v.ravel() for v in self.F], 1)
            Xm = np.stack([v.ravel() for v in self.X], 1)
            gam = np.linalg.solve(Fm.T @ Fm, Fm.T @ f.ravel())
            xn = xn - ((Xm + self.beta * Fm) @ gam).reshape(x.shape)
        self.k += 1
        return xn


def scf(S, mode=1, tol=1e-6, maxit=60, beta=None, history=None, frequency=None, verbose=False, nthreads=8,
        callback=None):
    """Oracle SCF on an oracle System (oracle.System) with set_cell_period(True)
    and build_nonlocal(box=cell).  Returns a dict of the final fields, energies
    and the residual history; callback(out), when given, sees each iteration's."""
    I = S.info()
    G = (I.npts[2] - 1, I.npts[1] - 1, I.npts[0] - 1)
    coeffs = [list(I.coeffs[d]) for d in range(3)]
    b, eself, _ = S.pseudocharge()
    rho, _ = S.atomic_density()
    b = b.reshape(G)
    rho = rho.reshape(G)
    p = S.scf_parameters()
    mix = PeriodicPulay(beta if beta is not None else p["beta"], history if history is not None else p["history"],
                        frequency if frequency is not None else p["frequency"])
    res_hist = []
    out = {}
    for it in range(maxit):
        eps, vxc = xc_lda_pw(rho)
        phi = poisson_fft(coeffs, I.order, rho + b)
        veff = phi + vxc
        lam, w, _, _ = S.nodes_and_weights(mode=mode, veff=veff.ravel(), nthreads=nthreads)
        lf, rho_out, eband, mts = S_density(S, lam, w)
        rho_out = rho_out.reshape(G)
        dV = I.dV
        exc = float(np.sum(rho * eps) * dV)
        evxc = float(np.sum(rho * vxc) * dV)
        eel = float(0.5 * np.sum(phi * (b - rho)) * dV)
        etot = eband + exc - evxc + eel - eself + mts
        res = float(np.linalg.norm(rho_out - rho) / np.linalg.norm(rho_out))
        res_hist.append(res)
        out = dict(iterations=it + 1, lambda_f=lf, e_band=eband, minus_ts=mts, e_xc=exc, e_vxc=evxc, e_el=eel,
                   e_self=eself, e_total=etot, rho_in=rho.copy(), rho_out=rho_out, veff=veff, phi=phi,
                   residuals=res_hist)
        if verbose:
            print(f"oracle scf {it}: res {res:.3e} E {etot:.10f}", flush=True)
        if callback is not None:
            callback(out)
        if res < tol:
            out["converged"] = True
            return out
        rho = mix.step(rho, rho_out)
    out["converged"] = False
    return out


def S_density(S, lam, w):
    from oracle import oracle as O
    I = S.info()
    return O.density_fermi_energy(lam, w, I.beta, float(I.noetot), I.dV)
