"""Independent matrix-form restatement of the RL / RZ / RLZ transforms built
only from textbook library routines, used to pin the oracle (tests only).

  cubic B-splines     scipy.interpolate.BSpline.basis_element (cardinal cubic, peak 2/3)
  Gauss quadrature    numpy.polynomial.legendre.leggauss
  ring DFT            numpy.fft.rfft / irfft  (R2HC = rfft, HC2R = N * irfft)
  DCT-I               scipy.fft.dct(type=1)   (= FFTW REDFT00, unnormalised)
  d/dz                numpy.polynomial.chebyshev.chebder / chebval
  SPD solve           scipy.linalg.solve(assume_a="pos") on the dense Gamma (P+Q) Gamma^T

It follows the same mathematical definitions as the reference (Ooyama 2002
variational spline, phase-rotated ring coefficients, Chebyshev-Gauss-Lobatto
columns) but shares no code with the oracle.  Only R0 Chebyshev BCs (the
configurations' BCs) are restated here; the other Chebyshev folds are pinned
by their own unit tests.
"""
import math

import numpy as np
import scipy.fft
import scipy.linalg
from scipy.interpolate import BSpline


def gauss_rule(mubar):
    x, w = np.polynomial.legendre.leggauss(mubar)
    return (x + 1.0) / 2.0, w / 2.0


class NpSpline:
    def __init__(self, xmin, xmax, nc, mubar=3, l_q=2.0, gamma=None):
        self.xmin, self.xmax, self.nc, self.mubar = xmin, xmax, nc, mubar
        self.DX = (xmax - xmin) / nc
        self.M = nc + 3
        q, w = gauss_rule(mubar)
        self.q, self.w = q, w
        self.x = np.array([xmin + mc * self.DX + q[mu] * self.DX for mc in range(nc) for mu in range(mubar)])
        self.wx = np.array([self.DX * w[mu] for mc in range(nc) for mu in range(mubar)])
        self.phi = [self.design(self.x, d) for d in range(4)]            # (npts, M) per derivative
        eps_q = ((l_q * self.DX) / (2 * math.pi)) ** 6
        P = (self.phi[0] * self.wx[:, None]).T @ self.phi[0]
        Q = eps_q * (self.phi[3] * self.wx[:, None]).T @ self.phi[3]
        self.G = np.eye(self.M) if gamma is None else gamma
        self.PQ = P + Q
        self.A = self.G @ (P + Q) @ self.G.T
        self.SB = (self.phi[0] * self.wx[:, None]).T                    # (M, npts)

    def node_spline(self, m):
        xm = self.xmin + m * self.DX
        t = xm + self.DX * np.arange(-2, 3)
        return BSpline.basis_element(t, extrapolate=False)

    def design(self, pts, d):
        D = np.zeros((len(pts), self.M))
        for mi in range(self.M):
            bs = self.node_spline(mi - 1)
            f = bs.derivative(d) if d > 0 else bs
            D[:, mi] = np.nan_to_num(f(pts), nan=0.0)
        return D

    def sb(self, u):
        return self.SB @ u

    def sa(self, b):
        y = scipy.linalg.solve(self.A, self.G @ b, assume_a="pos")
        return self.G.T @ y

    def sa_ahat(self, b, ahat):
        """R3X branch of SAtransform!: G' A^{-1} G (b - (P + Q) ahat) + ahat."""
        return self.sa(b - self.PQ @ ahat) + ahat

    def si(self, a, d):
        return self.phi[d] @ a


def selector_gamma(M, rank_left, rank_right):
    """Gamma of rank-3 (R3 / R3X) or R0 sides: the interior selector."""
    return np.eye(M)[rank_left:M - rank_right]


def dense_gamma(M, left=None, right=None):
    """Gamma for rank-1 (alpha, beta) BCs on either side (None = R0)."""
    rl, rr = (1 if left else 0), (1 if right else 0)
    Min = M - rl - rr
    G = np.zeros((Min, M))
    for i in range(Min):
        G[i, i + rl] = 1.0
    if left:
        G[0, 0] += left[0]
        G[1, 0] += left[1]
    if right:
        G[Min - 1, M - 1] += right[0]
        G[Min - 2, M - 1] += right[1]
    return G


class NpRing:
    """One Fourier ring: N points, b holding wavenumbers 0..kb, zero above kmax, the
    spline stages taking wavenumbers up to klim, phase offset ymin."""

    def __init__(self, N, kb, klim, kmax, ymin):
        self.N, self.kb, self.klim, self.kmax, self.ymin = N, kb, klim, kmax, ymin


def cyl_ring(ri, kcap=-1):
    """Cylindrical ring (src/factory.jl:153-156, 608-624): N = 4 + 4 ri, bDim = 2 ri + 1."""
    N = 4 + 4 * ri
    return NpRing(N, ri, ri, ri if kcap < 0 else min(kcap, ri), 0.5 * (2 * math.pi / N) * (ri - 1))


def sph_ring(theta, max_ri, r, kcap=-1):
    """Spherical ring (src/factory.jl:159-163, 626-643): kr = max(1, round(sin(theta) max_ri)),
    N = 4 kr, the cap applied to kmax and bDim; r is 1-based."""
    kr = max(1, int(round(math.sin(theta) * max_ri)))
    N = 4 * kr
    km = kr if kcap < 0 else min(kcap, kr)
    return NpRing(N, km, km, km, 0.5 * (2 * math.pi / N) * r)


def _ring(ri, kcap=-1):
    return ri if isinstance(ri, NpRing) else cyl_ring(ri, kcap)


def ring_params(ri, kcap=-1):
    rg = _ring(ri, kcap)
    return rg.N, rg.kmax, rg.ymin


def ring_forward(x, ri, kcap=-1):
    """-> complex b_k, k = 0..kb (zero above kmax): rfft / N rotated by -k ymin."""
    rg = _ring(ri, kcap)
    X = np.fft.rfft(x) / rg.N
    k = np.arange(rg.kb + 1)
    b = X[: rg.kb + 1] * np.exp(-1j * k * rg.ymin)
    b[rg.kmax + 1:] = 0.0
    b[0] = X[0].real
    return b


def ring_inverse(b, ri, kcap=-1, dl=0):
    """complex b_k, k = 0..kb -> N ring values of d^dl/dlambda^dl (HC2R = N irfft)."""
    rg = _ring(ri, kcap)
    k = np.arange(rg.kb + 1)
    C = np.zeros(rg.N // 2 + 1, dtype=complex)
    C[: rg.kmax + 1] = (b * np.exp(1j * k * rg.ymin) * (1j * k) ** dl)[: rg.kmax + 1]
    if dl > 0:
        C[0] = 0.0
    return rg.N * np.fft.irfft(C, n=rg.N)


class NpCheb:
    """R0 / R0 Chebyshev column on CGL points."""

    def __init__(self, zmin, zmax, N, B):
        self.zmin, self.zmax, self.N, self.B = zmin, zmax, N, B

    def forward(self, u):
        s = scipy.fft.dct(u, type=1) / (2 * (self.N - 1))
        if self.B < self.N:
            return s[: self.B]
        f = np.exp(-36.0 * (np.arange(1, self.N + 1) / self.N) ** 36)
        return f * s

    def inverse(self, b, dz=0):
        N = self.N
        a = np.zeros(N)
        a[: self.B] = b
        c = a.copy()
        c[1:N - 1] *= 2.0                           # true Chebyshev coefficients
        xi = np.cos(np.arange(N) * np.pi / (N - 1))
        for _ in range(dz):
            c = np.polynomial.chebyshev.chebder(c) / (-0.5 * (self.zmax - self.zmin))
        return np.polynomial.chebyshev.chebval(xi, c) if len(c) else np.zeros(N)


def rl_forward(sp, ri_list, u_rings, kcaps, K=None):
    """u_rings: list of ring arrays -> spectral (b_iDim * (1 + 2K)); ri_list holds ring
    indices ri (cylindrical, capped by kcaps) or NpRing objects (K then given)."""
    rings = [_ring(ri, kc) for ri, kc in zip(ri_list, kcaps)]
    K = max(ri_list) if K is None else K
    bs = [ring_forward(u, rg) for u, rg in zip(u_rings, rings)]
    M = sp.M
    out = np.zeros(M * (1 + 2 * K))
    out[:M] = sp.sb(np.array([b[0].real for b in bs]))
    for k in range(1, K + 1):
        ur = np.array([b[k].real if k <= rg.klim else 0.0 for b, rg in zip(bs, rings)])
        ui = np.array([b[k].imag if k <= rg.klim else 0.0 for b, rg in zip(bs, rings)])
        out[(2 * k - 1) * M: 2 * k * M] = sp.sb(ur)
        out[2 * k * M: (2 * k + 1) * M] = sp.sb(ui)
    return out


def rl_inverse(sp, ri_list, spec, kcaps, K=None):
    """-> list over slot (f, dr, drr, dl, dll) of concatenated ring values."""
    K = max(ri_list) if K is None else K
    M = sp.M
    nr = len(ri_list)
    vals = np.zeros((3, 1 + 2 * K, nr))
    for p in range(1 + 2 * K):
        a = sp.sa(spec[p * M:(p + 1) * M])
        for d in range(3):
            vals[d, p] = sp.si(a, d)
    slots = [[] for _ in range(5)]
    for r, (ri, kc) in enumerate(zip(ri_list, kcaps)):
        rg = _ring(ri, kc)
        cb = []
        for d in range(3):
            b = np.zeros(rg.kb + 1, dtype=complex)
            b[0] = vals[d, 0, r]
            for k in range(1, min(rg.kb, rg.klim) + 1):
                b[k] = vals[d, 2 * k - 1, r] + 1j * vals[d, 2 * k, r]
            cb.append(b)
        slots[0].append(ring_inverse(cb[0], rg, dl=0))
        slots[1].append(ring_inverse(cb[1], rg, dl=0))
        slots[2].append(ring_inverse(cb[2], rg, dl=0))
        slots[3].append(ring_inverse(cb[0], rg, dl=1))
        slots[4].append(ring_inverse(cb[0], rg, dl=2))
    return [np.concatenate(s) for s in slots]


def rz_forward(sp, ch, u):
    iDim, N, B, M = sp.nc * sp.mubar, ch.N, ch.B, sp.M
    U = u.reshape(iDim, N)
    C = np.array([ch.forward(U[r]) for r in range(iDim)])           # (iDim, B)
    return np.concatenate([sp.sb(C[:, z]) for z in range(B)])


def rz_inverse(sp, ch, spec):
    iDim, N, B, M = sp.nc * sp.mubar, ch.N, ch.B, sp.M
    vals = np.zeros((3, iDim, B))
    for z in range(B):
        a = sp.sa(spec[z * M:(z + 1) * M])
        for d in range(3):
            vals[d, :, z] = sp.si(a, d)
    out = np.zeros((5, iDim, N))
    for r in range(iDim):
        out[0, r] = ch.inverse(vals[0, r])
        out[3, r] = ch.inverse(vals[0, r], 1)
        out[4, r] = ch.inverse(vals[0, r], 2)
        out[1, r] = ch.inverse(vals[1, r])
        out[2, r] = ch.inverse(vals[2, r])
    return out.reshape(5, iDim * N)


def rlz_forward(sp, ch, ri_list, u, kcap=-1, K=None):
    rings = [_ring(ri, kcap) for ri in ri_list]
    K, M, N, B = (max(ri_list) if K is None else K), sp.M, ch.N, ch.B
    off = 0
    bz = []                                                          # [ring][zb] complex b
    for rg in rings:
        n = rg.N
        U = u[off * N:(off + n) * N].reshape(n, N)
        off += n
        C = np.array([ch.forward(U[l]) for l in range(n)])           # (n, B)
        bz.append([ring_forward(C[:, zb], rg) for zb in range(B)])
    out = np.zeros(B * M * (1 + 2 * K))
    for zb in range(B):
        base = zb * M * (1 + 2 * K)
        out[base:base + M] = sp.sb(np.array([b[zb][0].real for b in bz]))
        for k in range(1, K + 1):
            ur = np.array([b[zb][k].real if k <= rg.klim else 0.0 for b, rg in zip(bz, rings)])
            ui = np.array([b[zb][k].imag if k <= rg.klim else 0.0 for b, rg in zip(bz, rings)])
            out[base + (2 * k - 1) * M: base + 2 * k * M] = sp.sb(ur)
            out[base + 2 * k * M: base + (2 * k + 1) * M] = sp.sb(ui)
    return out


def rlz_inverse(sp, ch, ri_list, spec, kcap=-1, K=None):
    K, M, N, B = (max(ri_list) if K is None else K), sp.M, ch.N, ch.B
    nr = len(ri_list)
    vals = np.zeros((3, B, 1 + 2 * K, nr))
    for zb in range(B):
        base = zb * M * (1 + 2 * K)
        for p in range(1 + 2 * K):
            a = sp.sa(spec[base + p * M: base + (p + 1) * M])
            for d in range(3):
                vals[d, zb, p] = sp.si(a, d)
    slots = [[] for _ in range(7)]
    for r, ri in enumerate(ri_list):
        rg = _ring(ri, kcap)
        n = rg.N
        fam = np.zeros((5, n, B))
        for zb in range(B):
            cb = []
            for d in range(3):
                b = np.zeros(rg.kb + 1, dtype=complex)
                b[0] = vals[d, zb, 0, r]
                for k in range(1, min(rg.kb, rg.klim) + 1):
                    b[k] = vals[d, zb, 2 * k - 1, r] + 1j * vals[d, zb, 2 * k, r]
                cb.append(b)
            fam[0, :, zb] = ring_inverse(cb[0], rg, dl=0)
            fam[1, :, zb] = ring_inverse(cb[0], rg, dl=1)
            fam[2, :, zb] = ring_inverse(cb[0], rg, dl=2)
            fam[3, :, zb] = ring_inverse(cb[1], rg, dl=0)
            fam[4, :, zb] = ring_inverse(cb[2], rg, dl=0)
        for l in range(n):
            slots[0].append(ch.inverse(fam[0, l]))
            slots[5].append(ch.inverse(fam[0, l], 1))
            slots[6].append(ch.inverse(fam[0, l], 2))
            slots[3].append(ch.inverse(fam[1, l]))
            slots[4].append(ch.inverse(fam[2, l]))
            slots[1].append(ch.inverse(fam[3, l]))
            slots[2].append(ch.inverse(fam[4, l]))
    return [np.concatenate(s) for s in slots]
