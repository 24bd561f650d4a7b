"""Set-up data the Hamiltonian application reads: lattice, atoms, k-point grid, the
plane-wave grid (PWGrid.jl), GTH pseudopotentials (PsPot_GTH.jl), real spherical
harmonics (Ylm_real.jl), projectors betaNL (PsPotNL.jl) and the local ionic potential
(Hamiltonian.jl:119-141).  Host-side numpy, run once per geometry, as in the reference.

Index maps are 0-based (the reference's are 1-based); grid arrays are flat column-major
(x fastest), ip = i + n1*(j + n2*k), exactly as Julia's reshape of a length-N vector.
"""
import math
import os

import numpy as np

PSP_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pseudopotentials")
ZATOMS = {"H": 1.0, "Al": 13.0, "Si": 14.0, "Ce": 58.0, "Pt": 78.0}


def gen_lattice_fcc(a):
    """gen_lattice.jl:17-27."""
    return 0.5 * a * np.array([[-1.0, 0.0, -1.0], [0.0, 1.0, 1.0], [1.0, 1.0, 0.0]])


def gen_lattice_sc(a):
    """gen_lattice.jl:3-13."""
    return a * np.eye(3)


class Atoms:
    """Atoms.jl:44-60; positions 3 x Natoms in bohr (fractional input -> LatVecs * frac)."""

    def __init__(self, symbols, positions, LatVecs, frac=True):
        self.LatVecs = np.asarray(LatVecs, dtype=float)
        p = np.asarray(positions, dtype=float).reshape(-1, 3).T
        self.positions = self.LatVecs @ p if frac else p.copy()
        self.atsymbs = list(symbols)
        self.SpeciesSymbols = list(dict.fromkeys(symbols))
        self.atm2species = np.array([self.SpeciesSymbols.index(s) for s in symbols])
        self.Natoms = len(symbols)
        self.Nspecies = len(self.SpeciesSymbols)


def diamond_supercell(a, reps):
    """Diamond structure in a simple-cubic supercell of reps^3 conventional cells
    (8 atoms each), used for Si8 / Si64 / Si216."""
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
                      [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])
    frac = [(basis[b] + np.array([i, j, k])) / reps
            for i in range(reps) for j in range(reps) for k in range(reps) for b in range(8)]
    return np.array(frac)


def good_fft_order(nr):
    """fft_support.jl:48-72: smallest n >= nr whose prime factors are 2, 3, 5."""
    n = nr
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1
        if n > 2049:
            raise ValueError("fft_order is too large")


def crystal_symmetry(atoms, tol=1e-5):
    """Space-group operations of the crystal, as SymmetryInfo(atoms) finds them
    (SymmetryInfo.jl:711-770 via find_symm_bravais_latt + sgam_at): the rotations of the
    Bravais lattice (integer matrices R on crystal coordinates with R^T g R = g, g = L^T L)
    that map every atom onto an atom of the same species after some fractional translation
    t.  Returns [(R, t)] with the identity first; R acts on direct crystal coordinates, and
    the matrix acting on crystal coordinates of k is inv(R)^T."""
    L = np.asarray(atoms.LatVecs, dtype=float)
    g = L.T @ L
    frac = np.linalg.solve(L, atoms.positions)
    sp = np.asarray(atoms.atm2species)
    ops = []
    scale = np.abs(g).max()
    for e in np.ndindex(*(3,) * 9):
        R = np.array(e, dtype=np.int64).reshape(3, 3) - 1
        if round(abs(np.linalg.det(R))) != 1:
            continue
        if np.abs(R.T @ g @ R - g).max() > tol * scale:
            continue
        Rf = R @ frac
        for b in np.flatnonzero(sp == sp[0]):
            t = frac[:, b] - Rf[:, 0]
            d = (Rf + t[:, None])[:, :, None] - frac[:, None, :]        # 3 x Natoms x Natoms
            d -= np.round(d)
            hit = (np.abs(d).max(axis=0) < tol) & (sp[:, None] == sp[None, :])
            if np.all(hit.any(axis=1)):
                ops.append((R, t - np.round(t)))
                break
    ops.sort(key=lambda o: 0 if np.array_equal(o[0], np.eye(3, dtype=np.int64)) else 1)
    return ops


def kpoint_rotations(ops):
    """Matrices acting on crystal coordinates of k for the operations of crystal_symmetry."""
    return [np.rint(np.linalg.inv(R).T).astype(np.int64) for R, _ in ops]


def kpoint_grid(LatVecs, mesh, shift, rots=None, time_reversal=True):
    """KPoints.jl:121-259 (QE kpoint_grid): Monkhorst-Pack grid reduced by the integer
    rotations `rots` (crystal coordinates of k) and time reversal.  -> (k_cart, wk)."""
    n = np.array(mesh)
    sh = np.array(shift, dtype=float)
    rots = [np.eye(3, dtype=int)] if rots is None else rots
    idx = np.array([(i, j, k) for i in range(n[0]) for j in range(n[1]) for k in range(n[2])])
    xkg = (idx + sh / 2.0) / n
    nkr = len(xkg)
    lin = lambda t: t[2] + t[1] * n[2] + t[0] * n[1] * n[2]
    equiv = np.arange(nkr)
    wkk = np.zeros(nkr)
    for ik in range(nkr):
        if equiv[ik] != ik:
            continue
        wkk[ik] = 1.0
        for s in rots:
            xr = s @ xkg[ik]
            xr = xr - np.round(xr)
            for sgn in ((1, -1) if time_reversal else (1,)):
                t = sgn * xr * n - 0.5 * sh
                if np.all(np.abs(t - np.round(t)) <= 1e-5):
                    m = lin((np.round(sgn * xr * n - 0.5 * sh + 2 * n).astype(int)) % n)
                    if m > ik and equiv[m] == m:
                        equiv[m] = ik
                        wkk[ik] += 1.0
                    elif equiv[m] != ik or m < ik:
                        raise RuntimeError("kpoint_grid: inconsistent rotation set")
    sel = np.nonzero(equiv == np.arange(nkr))[0]
    wk = wkk[sel] / wkk[sel].sum()
    xk = xkg[sel] - np.round(xkg[sel])
    Rec = 2 * np.pi * np.linalg.inv(LatVecs.T)
    return (Rec @ xk.T), wk


class PWGrid:
    """PWGrid.jl:62-152 (dual = 4): dense grid, rho-sphere G-vectors sorted by |G|^2 with
    ties in grid order (init_gvec :199-242), G-shells (:245-272) and the per-k
    wave-function spheres |G+k|^2/2 <= ecutwfc (init_gvecw :292-322)."""

    def __init__(self, ecutwfc, LatVecs, kpoints=None, Ns=None):
        self.ecutwfc = float(ecutwfc)
        self.ecutrho = 4.0 * self.ecutwfc
        self.LatVecs = np.asarray(LatVecs, dtype=float)
        self.RecVecs = 2 * np.pi * np.linalg.inv(self.LatVecs.T)
        self.CellVolume = abs(np.linalg.det(self.LatVecs))
        if Ns is None or min(Ns) == 0:
            L = np.linalg.norm(self.LatVecs, axis=0)
            Ns = [good_fft_order(2 * int(round(math.sqrt(self.ecutrho / 2) * l / math.pi)) + 1) for l in L]
        self.Ns = tuple(int(x) for x in Ns)
        n1, n2, n3 = self.Ns
        self.Npoints = n1 * n2 * n3
        k3, k2, k1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
        m = [np.where(a > s / 2, a - s, a).ravel().astype(float) for a, s in ((k1, n1), (k2, n2), (k3, n3))]
        R = self.RecVecs
        G = [R[d, 0] * m[0] + R[d, 1] * m[1] + R[d, 2] * m[2] for d in range(3)]
        G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2]
        sel = np.flatnonzero(0.5 * G2 <= self.ecutrho)
        sel = sel[np.argsort(G2[sel], kind="stable")]
        self.Ng = sel.size
        self.G = np.stack([g[sel] for g in G])
        self.G2 = G2[sel]
        self.idx_g2r = sel
        new = np.concatenate([[True], self.G2[1:] > self.G2[:-1] + 1e-8])
        self.idx_g2shells = np.cumsum(new) - 1
        self.G2_shells = self.G2[new]
        if kpoints is None:
            kpoints = (np.zeros((3, 1)), np.ones(1))
        self.k = np.asarray(kpoints[0], dtype=float).reshape(3, -1)
        self.wk = np.asarray(kpoints[1], dtype=float)
        self.Nkpt = self.k.shape[1]
        self.idx_gw2g, self.idx_gw2r, self.Ngw = [], [], []
        for ik in range(self.Nkpt):
            kk = self.k[:, ik]
            gk2 = (self.G[0] + kk[0]) ** 2 + (self.G[1] + kk[1]) ** 2 + (self.G[2] + kk[2]) ** 2
            w = np.flatnonzero(0.5 * gk2 <= self.ecutwfc)
            self.idx_gw2g.append(w)
            self.idx_gw2r.append(self.idx_g2r[w])
            self.Ngw.append(w.size)
        self.Ngwx = max(self.Ngw)

    def gk(self, ik):
        return self.G[:, self.idx_gw2g[ik]] + self.k[:, ik][:, None]

    def gk2(self, ik):
        g = self.gk(ik)
        return g[0] ** 2 + g[1] ** 2 + g[2] ** 2


class PsPot_GTH:
    """PsPot_GTH.jl:43-118: parser of a GTH parameter file."""

    def __init__(self, filename):
        if not os.path.isabs(filename) and not os.path.exists(filename):
            filename = os.path.join(PSP_DIR, filename)
        self.pspfile = filename
        with open(filename) as f:
            rows = [ln.split() for ln in f.read().splitlines()]
        self.atsymb = rows[0][0]
        self.zval = float(sum(int(x) for x in rows[1][:4]))
        self.rlocal, nc = float(rows[2][0]), int(rows[2][1])
        self.c = np.zeros(4)
        self.c[:nc] = [float(x) for x in rows[3][:nc]]
        self.lmax = int(rows[4][0]) - 1
        self.rc = np.zeros(4)
        self.Nproj_l = np.zeros(4, dtype=int)
        self.h = np.zeros((4, 3, 3))
        r = 5
        for l in range(self.lmax + 1):
            self.rc[l], self.Nproj_l[l] = float(rows[r][0]), int(float(rows[r][1]))
            r += 1
            for i in range(self.Nproj_l[l]):
                for jj, val in enumerate(rows[r][: self.Nproj_l[l] - i]):
                    self.h[l, i, i + jj] = float(val)
                r += 1
            self.h[l] = np.triu(self.h[l]) + np.triu(self.h[l], 1).T


def eval_Vloc_G(psp, G2):
    """PsPot_GTH.jl:164-189 (with the finite G = 0 value used by QE)."""
    G2 = np.asarray(G2, dtype=float)
    rl, Z = psp.rlocal, psp.zval
    c1, c2, c3, c4 = psp.c
    Gr = np.sqrt(G2) * rl
    x = Gr ** 2
    e = np.exp(-0.5 * x)
    poly = c1 + c2 * (3.0 - x) + c3 * (15.0 - 10.0 * x + x ** 2) + c4 * (105.0 - 105.0 * x + 21.0 * x ** 2 - x ** 3)
    big = np.sqrt(np.sqrt(G2)) > 1e-8
    safe = np.where(big, G2, 1.0)
    v = -4 * np.pi * Z / safe * e + math.sqrt(8 * np.pi ** 3) * rl ** 3 * e * poly
    v0 = 2 * np.pi * Z * rl ** 2 + (2 * np.pi) ** 1.5 * rl ** 3 * (c1 + 3.0 * c2 + 15.0 * c3 + 105.0 * c4)
    return np.where(big, v, v0)


_PROJ = {  # (l, iproj) -> radial factor f(Gm, x = (Gm rc)^2) without exp(-x/2)
    (0, 1): lambda g, x: 1.0,
    (0, 2): lambda g, x: 2.0 / math.sqrt(15.0) * (3.0 - x),
    (0, 3): lambda g, x: (4.0 / 3.0) / math.sqrt(105.0) * (15.0 - 10.0 * x + x ** 2),
    (1, 1): lambda g, x: (1.0 / math.sqrt(3.0)) * g,
    (1, 2): lambda g, x: (2.0 / math.sqrt(105.0)) * g * (5.0 - x),
    (1, 3): lambda g, x: (4.0 / 3.0) / math.sqrt(1155.0) * g * (35.0 - 14.0 * x + x ** 2),
    (2, 1): lambda g, x: (1.0 / math.sqrt(15.0)) * g ** 2,
    (2, 2): lambda g, x: (2.0 / 3.0) / math.sqrt(105.0) * g ** 2 * (7.0 - x),
    (3, 1): lambda g, x: g ** 3 / math.sqrt(105.0),
}


def eval_proj_G(psp, l, iproj, Gm, CellVolume):
    """PsPot_GTH.jl:338-391 (iproj 1-based)."""
    rl = psp.rc[l]
    x = (np.asarray(Gm) * rl) ** 2
    pre = 4.0 * np.pi ** 1.25 * math.sqrt(2.0 ** (l + 1) * rl ** (2 * l + 3) / CellVolume)
    return pre * np.exp(-0.5 * x) * _PROJ[(l, iproj)](np.asarray(Gm), x)


def Ylm_real(l, m, R):
    """Ylm_real.jl:4-107 (QE real harmonics); R is 3 x n."""
    R = np.asarray(R, dtype=float)
    Rmod = np.sqrt(R[0] ** 2 + R[1] ** 2 + R[2] ** 2)
    small = Rmod < 1e-9
    cost = np.where(small, 0.0, R[2] / np.where(small, 1.0, Rmod))
    x_ok = np.abs(R[0]) > 1e-9
    at = np.arctan(R[1] / np.where(x_ok, R[0], 1.0))
    phi = np.where(R[0] > 1e-9, at, np.where(R[0] < -1e-9, at + np.pi, np.where(R[1] >= 0, np.pi / 2, -np.pi / 2)))
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost ** 2))
    if l == 0:
        return np.full(R.shape[1], 0.5 * math.sqrt(1.0 / np.pi))
    if l == 1:
        f = 0.5 * math.sqrt(3.0 / np.pi)
        return {-1: f * sint * np.sin(phi), 0: f * cost, 1: f * sint * np.cos(phi)}[m]
    if l == 2:
        a, b = math.sqrt(15.0 / 16.0 / np.pi), math.sqrt(15.0 / 4.0 / np.pi)
        return {-2: a * sint ** 2 * np.sin(2 * phi), -1: b * cost * sint * np.sin(phi),
                0: 0.25 * math.sqrt(5.0 / np.pi) * (3.0 * cost ** 2 - 1.0),
                1: b * cost * sint * np.cos(phi), 2: a * sint ** 2 * np.cos(2 * phi)}[m]
    a, b, d = 0.25 * math.sqrt(35.0 / 2.0 / np.pi), 0.25 * math.sqrt(105.0 / np.pi), 0.25 * math.sqrt(21.0 / 2.0 / np.pi)
    return {-3: a * sint ** 3 * np.sin(3 * phi), -2: b * sint ** 2 * cost * np.sin(2 * phi),
            -1: d * sint * (5.0 * cost ** 2 - 1.0) * np.sin(phi),
            0: 0.25 * math.sqrt(7.0 / np.pi) * (5.0 * cost ** 3 - 3.0 * cost),
            1: d * sint * (5.0 * cost ** 2 - 1.0) * np.cos(phi), 2: b * sint ** 2 * cost * np.cos(2 * phi),
            3: a * sint ** 3 * np.cos(3 * phi)}[m]


def strfact(positions, atm2species, Nspecies, G):
    """calc_strfact.jl:2-16."""
    Sf = np.zeros((G.shape[1], Nspecies), dtype=complex)
    for ia in range(positions.shape[1]):
        t = positions[:, ia]
        GX = t[0] * G[0] + t[1] * G[1] + t[2] * G[2]
        Sf[:, atm2species[ia]] += np.cos(GX) - 1j * np.sin(GX)
    return Sf


class PsPotNL:
    """PsPotNL.jl:16-77 + _init_prj2beta (:147-195): betaNL[ik] (Ngw x Nbeta) with
    beta = (-i)^l Y_lm(g) p_{l,i}(|g|) exp(-i g.tau), ordered (atom, l, m, iprj); the
    block-diagonal D[ib, jb] = h[l, iprj, jprj] that op_V_Ps_nloc.jl:59-76 sums with."""

    def __init__(self, atoms, pw, pspots):
        self._src = (atoms, pw, pspots, np.array(atoms.positions))
        self._beta = None
        self.blocks = []
        nb = 0
        for ia in range(atoms.Natoms):
            psp = pspots[atoms.atm2species[ia]]
            for l in range(psp.lmax + 1):
                for m in range(-l, l + 1):
                    ids = list(range(nb, nb + int(psp.Nproj_l[l])))
                    self.blocks.append((ia, l, m, ids))
                    nb += len(ids)
        self.NbetaNL = nb
        self.D = np.zeros((nb, nb))
        for ia, l, m, ids in self.blocks:
            h = pspots[atoms.atm2species[ia]].h[l]
            self.D[np.ix_(ids, ids)] = h[: len(ids), : len(ids)]

    def projector_table(self):
        """(atom, l, m, iprj) of every beta column, iprj 1-based (pw_build_projectors_dev)."""
        return np.array([(ia, l, m, i + 1) for ia, l, m, ids in self.blocks for i in range(len(ids))],
                        dtype=np.int32).reshape(-1, 4)

    @property
    def betaNL(self):
        """Host copy of the projectors for the positions this object was built with,
        computed on first access (the Hamiltonian builds its device copy itself)."""
        if self._beta is None:
            self._beta = self._build_host()
        return self._beta

    def _build_host(self):
        atoms, pw, pspots, pos = self._src
        nb = self.NbetaNL
        out = []
        for ik in range(pw.Nkpt):
            g = pw.gk(ik)
            Gm = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
            beta = np.empty((pw.Ngw[ik], nb), dtype=complex, order="F")
            ylm_cache = {}
            for ia, l, m, ids in self.blocks:
                psp = pspots[atoms.atm2species[ia]]
                t = pos[:, ia]
                GX = t[0] * g[0] + t[1] * g[1] + t[2] * g[2]
                Sf = np.cos(GX) - 1j * np.sin(GX)
                if (l, m) not in ylm_cache:
                    ylm_cache[(l, m)] = Ylm_real(l, m, g)
                ang = ((-1j) ** l) * ylm_cache[(l, m)] * Sf
                for a, ib in enumerate(ids):
                    beta[:, ib] = ang * eval_proj_G(psp, l, a + 1, Gm, pw.CellVolume)
            out.append(beta)
        return out


class RhoeSymmetrizer:
    """Set-up data of symmetrize_rhoe! (RhoeSymmetrizer.jl:1-105) for the operations of
    crystal_symmetry, in the form the device kernel takes (pw_set_symmetry):
        rho_sym(m) = 1/Nsym sum_s rho(M_s m) exp(2 pi i t_s . M_s m),   M_s = R_s^-T,
    m = integer crystal coordinates of the rho-sphere G (ng x 3), t_s the fractional
    translation of {R_s|t_s}.  Components outside the sphere are kept, as the reference
    keeps them.  The library checks that the sphere is closed under every M_s."""

    def __init__(self, atoms, pw, ops):
        self.Nsyms = len(ops)
        self.m = np.ascontiguousarray(
            np.rint(np.asarray(atoms.LatVecs, dtype=float).T @ pw.G / (2 * np.pi)).T, dtype=np.int32)
        self.M = np.ascontiguousarray([np.rint(np.linalg.inv(R).T) for R, _ in ops], dtype=np.int32)
        self.t = np.ascontiguousarray([t for _, t in ops], dtype=float)

    def upload(self, ctx):
        ctx.call("pw_set_symmetry", self.Nsyms, self.M.ctypes.data, self.t.ctypes.data, self.m.shape[0],
                 self.m.ctypes.data)


def symmetrize_vector(atoms, ops, v):
    """symmetrize_vector! (SymmetryInfo.jl:923-967) for a per-atom Cartesian vector field
    v (3 x Natoms): v_sym[a] = 1/Nsym sum_s C_s^-1 v[g_s(a)], with C_s = L R_s L^-1 and
    g_s(a) the atom that {R_s|t_s} carries atom a onto."""
    if len(ops) <= 1:
        return np.array(v, dtype=float)
    L = np.asarray(atoms.LatVecs, dtype=float)
    Li = np.linalg.inv(L)
    frac = Li @ atoms.positions
    v = np.asarray(v, dtype=float)
    out = np.zeros_like(v)
    for R, t in ops:
        C = L @ R @ Li
        d = (R @ frac + t[:, None])[:, :, None] - frac[:, None, :]
        d -= np.round(d)
        g = np.argmin(np.abs(d).max(axis=0), axis=1)
        out += np.linalg.solve(C, v[:, g])
    return out / len(ops)
