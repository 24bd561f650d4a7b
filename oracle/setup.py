"""Oracle: set-up restatement (lattice, atoms, PWGrid, GTH pseudopotentials, projectors).

Test infrastructure only (see oracle/__init__.py).  Indices here are 0-based; the
reference is 1-based, so idx_g2r/idx_gw2r are the reference's maps minus one.
Grid arrays are flat in column-major order (x fastest): ip = i + N1*(j + N2*k).
"""
import math
import numpy as np

# ----------------------------------------------------------------------------- lattice / atoms


def gen_lattice_fcc(a):
    """gen_lattice.jl:17-27 -- lattice vectors stored by column."""
    L = np.zeros((3, 3))
    L[:, 0] = 0.5 * a * np.array([-1.0, 0.0, 1.0])
    L[:, 1] = 0.5 * a * np.array([0.0, 1.0, 1.0])
    L[:, 2] = 0.5 * a * np.array([-1.0, 1.0, 0.0])
    return L


def gen_lattice_sc(a):
    """gen_lattice.jl:3-13 (gen_lattice_cubic)."""
    return a * np.eye(3)


class Atoms:
    """Atoms.jl:44-60: positions (3 x Natoms, bohr); fractional input is LatVecs*frac."""

    def __init__(self, symbols, positions, LatVecs, frac=True):
        self.LatVecs = np.array(LatVecs, dtype=float)
        pos = np.array(positions, dtype=float).reshape(-1, 3).T.copy()
        self.positions = self.LatVecs @ pos if frac else pos
        self.atsymbs = list(symbols)
        species = []
        for s in symbols:
            if s not in species:
                species.append(s)
        self.SpeciesSymbols = species
        self.atm2species = np.array([species.index(s) for s in symbols], dtype=np.int64)
        self.Natoms = len(symbols)
        self.Nspecies = len(species)


ZATOMS = {"H": 1.0, "Al": 13.0, "Si": 14.0, "Ce": 58.0, "Pt": 78.0}


# ----------------------------------------------------------------------------- FFT grid sizes


def _allowed(nr):
    """fft_support.jl:2-45 (QE allowed): only factors 2,3,5 are good."""
    mr = nr
    for f in (2, 3, 5):
        while mr % f == 0:
            mr //= f
    return mr == 1


def good_fft_order(nr):
    """fft_support.jl:48-72."""
    n = nr
    while not _allowed(n) and n <= 2049:
        n += 1
    if n > 2049:
        raise ValueError("fft_order is too large")
    return n


# ----------------------------------------------------------------------------- k-points


def gen_kpoint_grid(LatVecs, nk, shift, rots=None, time_reversal=True):
    """KPoints.jl:121-259 (port of QE kpoint_grid): reduce an MP grid by the rotations
    `rots` (3x3 integer matrices acting on crystal coordinates of k) and, when asked,
    by time reversal.  Returns (k_cart 3 x Nk, wk normalised to 1)."""
    nk1, nk2, nk3 = nk
    k1, k2, k3 = shift
    if rots is None:
        rots = [np.eye(3, dtype=np.int64)]
    SMALL = 1.0e-5
    nkr = nk1 * nk2 * nk3
    xkg = np.zeros((3, nkr))
    for i in range(nk1):
        for j in range(nk2):
            for k in range(nk3):
                ik = k + j * nk3 + i * nk2 * nk3
                xkg[0, ik] = i / nk1 + k1 / 2.0 / nk1
                xkg[1, ik] = j / nk2 + k2 / 2.0 / nk2
                xkg[2, ik] = k / nk3 + k3 / 2.0 / nk3
    equiv = np.arange(nkr)
    wkk = np.zeros(nkr)

    def _index(xk):
        xx = xk[0] * nk1 - 0.5 * k1
        yy = xk[1] * nk2 - 0.5 * k2
        zz = xk[2] * nk3 - 0.5 * k3
        if (abs(xx - round(xx)) <= SMALL and abs(yy - round(yy)) <= SMALL
                and abs(zz - round(zz)) <= SMALL):
            i = int(round(xk[0] * nk1 - 0.5 * k1 + 2 * nk1)) % nk1
            j = int(round(xk[1] * nk2 - 0.5 * k2 + 2 * nk2)) % nk2
            k = int(round(xk[2] * nk3 - 0.5 * k3 + 2 * nk3)) % nk3
            return k + j * nk3 + i * nk2 * nk3
        return None

    for ik in range(nkr):
        if equiv[ik] != ik:
            continue
        wkk[ik] = 1.0
        for s in rots:
            xkr = s @ xkg[:, ik]
            xkr = xkr - np.round(xkr)
            cands = [xkr, -xkr] if time_reversal else [xkr]
            for xc in cands:
                n = _index(xc)
                if n is None:
                    continue
                if n > ik and equiv[n] == n:
                    equiv[n] = ik
                    wkk[ik] += 1.0
                elif equiv[n] != ik or n < ik:
                    raise RuntimeError("gen_kpoint_grid something wrong in the checking algorithm")
    sel = [ik for ik in range(nkr) if equiv[ik] == ik]
    wk = wkk[sel] / wkk[sel].sum()
    xk = xkg[:, sel] - np.round(xkg[:, sel])
    RecVecs = 2 * np.pi * np.linalg.inv(LatVecs.T)
    return RecVecs @ xk, wk


# ----------------------------------------------------------------------------- PWGrid


def mm_to_nn(mm, S):
    """PWGrid.jl:157-163."""
    return np.where(mm > S / 2, mm - S, mm)


class PWGrid:
    """PWGrid.jl:62-152 with init_gvec (:199-242), init_Gshells (:245-272) and
    init_gvecw (:292-322).  dual = 4 only (no smooth grid), as on the hot path."""

    def __init__(self, ecutwfc, LatVecs, kpoints=None, Ns=None):
        self.ecutwfc = float(ecutwfc)
        self.ecutrho = 4.0 * self.ecutwfc
        self.LatVecs = np.array(LatVecs, dtype=float)
        self.RecVecs = 2 * np.pi * np.linalg.inv(self.LatVecs.T)
        self.CellVolume = abs(np.linalg.det(self.LatVecs))
        if Ns is None or any(n == 0 for n in Ns):
            lens = np.linalg.norm(self.LatVecs, axis=0)
            Ns = tuple(good_fft_order(2 * int(round(math.sqrt(self.ecutrho / 2) * L / math.pi)) + 1)
                       for L in lens)
        self.Ns = tuple(int(n) for n in Ns)
        N1, N2, N3 = self.Ns
        self.Npoints = N1 * N2 * N3
        ip = np.arange(self.Npoints)
        i = ip % N1
        j = (ip // N1) % N2
        k = ip // (N1 * N2)
        gi = mm_to_nn(i, N1).astype(float)
        gj = mm_to_nn(j, N2).astype(float)
        gk = mm_to_nn(k, N3).astype(float)
        R = self.RecVecs
        Gx = R[0, 0] * gi + R[0, 1] * gj + R[0, 2] * gk
        Gy = R[1, 0] * gi + R[1, 1] * gj + R[1, 2] * gk
        Gz = R[2, 0] * gi + R[2, 1] * gj + R[2, 2] * gk
        G2 = Gx * Gx + Gy * Gy + Gz * Gz
        sel = np.nonzero(0.5 * G2 <= self.ecutrho)[0]
        order = np.argsort(G2[sel], kind="stable")   # sortperm: ties by linear index
        sel = sel[order]
        self.Ng = len(sel)
        self.G = np.vstack([Gx[sel], Gy[sel], Gz[sel]])
        self.G2 = G2[sel]
        self.idx_g2r = sel.astype(np.int64)
        self.miller = np.vstack([gi[sel], gj[sel], gk[sel]]).astype(np.int64)
        # shells (PWGrid.jl:245-272)
        eps8 = 1e-8
        new = np.ones(self.Ng, dtype=bool)
        new[1:] = self.G2[1:] > self.G2[:-1] + eps8
        self.idx_g2shells = np.cumsum(new) - 1
        self.G2_shells = self.G2[new]
        # wave-function spheres
        if kpoints is None:
            kpoints = (np.zeros((3, 1)), np.array([1.0]))
        self.k = np.array(kpoints[0], dtype=float).reshape(3, -1)
        self.wk = np.array(kpoints[1], dtype=float)
        self.Nkpt = self.k.shape[1]
        self.idx_gw2g, self.idx_gw2r, self.Ngw = [], [], []
        for ik in range(self.Nkpt):
            kk = self.k[:, ik]
            Gk2 = (self.G[0] + kk[0]) ** 2 + (self.G[1] + kk[1]) ** 2 + (self.G[2] + kk[2]) ** 2
            idx = np.nonzero(0.5 * Gk2 <= self.ecutwfc)[0]
            self.idx_gw2g.append(idx)
            self.idx_gw2r.append(self.idx_g2r[idx])
            self.Ngw.append(len(idx))
        self.Ngwx = max(self.Ngw)

    def gk(self, ik):
        """G+k vectors (3 x Ngw) of the wave-function sphere at k-point ik."""
        idx = self.idx_gw2g[ik]
        return self.G[:, idx] + self.k[:, ik][:, None]

    def gk2(self, ik):
        """|G+k|^2 in the operation order of op_K.jl:69."""
        idx = self.idx_gw2g[ik]
        kk = self.k[:, ik]
        return (self.G[0, idx] + kk[0]) ** 2 + (self.G[1, idx] + kk[1]) ** 2 + (self.G[2, idx] + kk[2]) ** 2


# Threads of the 3-D FFTs (scipy.fft workers) when above 1; the default runs numpy's FFT on
# one thread.  The all-core CPU timing sets it to the number of cores.
FFT_WORKERS = 1


def G_to_R(pw, fG):
    """wrappers_fft.jl:42-51: in-place backward FFT, normalised by 1/N, exp(+i)."""
    f = np.asarray(fG).reshape(pw.Ns, order="F")
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifftn(f, workers=FFT_WORKERS).reshape(-1, order="F")
    return np.fft.ifftn(f).reshape(-1, order="F")


def R_to_G(pw, fR):
    """wrappers_fft.jl:53-62: forward FFT, unnormalised, exp(-i)."""
    f = np.asarray(fR, dtype=complex).reshape(pw.Ns, order="F")
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fftn(f, workers=FFT_WORKERS).reshape(-1, order="F")
    return np.fft.fftn(f).reshape(-1, order="F")


# ----------------------------------------------------------------------------- GTH pseudopotentials


class PsPot_GTH:
    """PsPot_GTH.jl:43-118 (parser of the CP2K-style GTH parameter file)."""

    def __init__(self, filename):
        with open(filename) as f:
            lines = [ln for ln in f.read().splitlines()]
        it = iter(lines)
        self.atsymb = next(it).split()[0]
        n = [int(x) for x in next(it).split()[:4]]
        self.zval = float(sum(n))
        l3 = next(it).split()
        self.rlocal = float(l3[0])
        nc = int(l3[1])
        self.c = np.zeros(4)
        cl = next(it).split()
        for i in range(nc):
            self.c[i] = float(cl[i])
        self.lmax = int(next(it).split()[0]) - 1
        self.rc = np.zeros(4)
        self.Nproj_l = np.zeros(4, dtype=np.int64)
        self.h = np.zeros((4, 3, 3))
        for i in range(self.lmax + 1):
            ll = next(it).split()
            self.rc[i] = float(ll[0])
            self.Nproj_l[i] = int(float(ll[1]))
            for ii in range(self.Nproj_l[i]):
                vals = next(it).split()
                for iread, iii in enumerate(range(ii, self.Nproj_l[i])):
                    self.h[i, ii, iii] = float(vals[iread])
        for k in range(4):
            for i in range(3):
                for j in range(i + 1, 3):
                    self.h[k, j, i] = self.h[k, i, j]


def eval_Vloc_G(psp, G2):
    """PsPot_GTH.jl:164-189 (eval_Vloc_G!): GTH local potential in G-space, with the
    finite QE-convention value at G = 0."""
    G2 = np.asarray(G2, dtype=float)
    rloc, zval = psp.rlocal, psp.zval
    c1, c2, c3, c4 = psp.c
    pre1 = -4 * np.pi * zval
    pre2 = math.sqrt(8 * np.pi ** 3) * rloc ** 3
    G2l = np.sqrt(G2)
    Gr = G2l * rloc
    expGr2 = np.exp(-0.5 * Gr ** 2)
    out = np.empty_like(G2)
    big = np.sqrt(G2l) > 1.0e-8
    with np.errstate(divide="ignore", invalid="ignore"):
        v = pre1 / G2 * expGr2 + pre2 * expGr2 * (
            c1 + c2 * (3.0 - Gr ** 2) + c3 * (15.0 - 10.0 * Gr ** 2 + Gr ** 4)
            + c4 * (105.0 - 105.0 * Gr ** 2 + 21.0 * Gr ** 4 - Gr ** 6))
    out[big] = v[big]
    out[~big] = 2 * np.pi * zval * rloc ** 2 + (2 * np.pi) ** 1.5 * rloc ** 3 * (c1 + 3.0 * c2 + 15.0 * c3 + 105.0 * c4)
    return out


def eval_proj_G(psp, l, iproj, Gm, CellVolume):
    """PsPot_GTH.jl:338-391: GTH projector p_{l,i}(|g|) in G-space (iproj is 1-based)."""
    Gm = np.asarray(Gm, dtype=float)
    rrl = psp.rc[l]
    Gr2 = (Gm * rrl) ** 2
    e = np.exp(-0.5 * Gr2)
    if l == 0:
        if iproj == 1:
            v = e
        elif iproj == 2:
            v = 2.0 / math.sqrt(15.0) * e * (3.0 - Gr2)
        else:
            v = (4.0 / 3.0) / math.sqrt(105.0) * e * (15.0 - 10.0 * Gr2 + Gr2 ** 2)
    elif l == 1:
        if iproj == 1:
            v = (1.0 / math.sqrt(3.0)) * e * Gm
        elif iproj == 2:
            v = (2.0 / math.sqrt(105.0)) * e * Gm * (5.0 - Gr2)
        else:
            v = (4.0 / 3.0) / math.sqrt(1155.0) * e * Gm * (35.0 - 14.0 * Gr2 + Gr2 ** 2)
    elif l == 2:
        if iproj == 1:
            v = (1.0 / math.sqrt(15.0)) * e * Gm ** 2
        else:
            v = (2.0 / 3.0) / math.sqrt(105.0) * e * Gm ** 2 * (7.0 - Gr2)
    else:
        v = Gm ** 3 * e / math.sqrt(105.0)
    pre = 4.0 * np.pi ** (5.0 / 4.0) * math.sqrt(2.0 ** (l + 1) * rrl ** (2 * l + 3) / CellVolume)
    return pre * v


def Ylm_real(l, m, R):
    """Ylm_real.jl:4-107: real spherical harmonics (QE convention), R is 3 x n."""
    R = np.asarray(R, dtype=float).reshape(3, -1)
    SMALL = 1.0e-9
    Rmod = np.sqrt(R[0] ** 2 + R[1] ** 2 + R[2] ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        cost = np.where(Rmod < SMALL, 0.0, R[2] / np.where(Rmod < SMALL, 1.0, Rmod))
        ratio = R[1] / np.where(np.abs(R[0]) > SMALL, R[0], 1.0)
    phi = np.where(R[0] > SMALL, np.arctan(ratio),
                   np.where(R[0] < -SMALL, np.arctan(ratio) + np.pi,
                            np.where(R[1] >= 0, np.pi / 2, -np.pi / 2)))
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost ** 2))
    if l == 0:
        return np.full(R.shape[1], 0.5 * math.sqrt(1.0 / np.pi))
    if l == 1:
        if m == -1:
            return 0.5 * math.sqrt(3.0 / np.pi) * sint * np.sin(phi)
        if m == 0:
            return 0.5 * math.sqrt(3.0 / np.pi) * cost
        return 0.5 * math.sqrt(3.0 / np.pi) * sint * np.cos(phi)
    if l == 2:
        if m == -2:
            return math.sqrt(15.0 / 16.0 / np.pi) * sint ** 2 * np.sin(2.0 * phi)
        if m == -1:
            return math.sqrt(15.0 / 4.0 / np.pi) * cost * sint * np.sin(phi)
        if m == 0:
            return 0.25 * math.sqrt(5.0 / np.pi) * (3.0 * cost ** 2 - 1.0)
        if m == 1:
            return math.sqrt(15.0 / 4.0 / np.pi) * cost * sint * np.cos(phi)
        return math.sqrt(15.0 / 16.0 / np.pi) * sint ** 2 * np.cos(2.0 * phi)
    if m == -3:
        return 0.25 * math.sqrt(35.0 / 2.0 / np.pi) * sint ** 3 * np.sin(3.0 * phi)
    if m == -2:
        return 0.25 * math.sqrt(105.0 / np.pi) * sint ** 2 * cost * np.sin(2.0 * phi)
    if m == -1:
        return 0.25 * math.sqrt(21.0 / 2.0 / np.pi) * sint * (5.0 * cost ** 2 - 1.0) * np.sin(phi)
    if m == 0:
        return 0.25 * math.sqrt(7.0 / np.pi) * (5.0 * cost ** 3 - 3.0 * cost)
    if m == 1:
        return 0.25 * math.sqrt(21.0 / 2.0 / np.pi) * sint * (5.0 * cost ** 2 - 1.0) * np.cos(phi)
    if m == 2:
        return 0.25 * math.sqrt(105.0 / np.pi) * sint ** 2 * cost * np.cos(2.0 * phi)
    return 0.25 * math.sqrt(35.0 / 2.0 / np.pi) * sint ** 3 * np.cos(3.0 * phi)


def calc_strfact(atpos, atm2species, Nspecies, G):
    """calc_strfact.jl:2-16: S_sp(G) = sum_a exp(-i G.tau_a)."""
    Sf = np.zeros((G.shape[1], Nspecies), dtype=complex)
    for ia in range(atpos.shape[1]):
        GX = atpos[0, ia] * G[0] + atpos[1, ia] * G[1] + atpos[2, ia] * G[2]
        Sf[:, atm2species[ia]] += np.cos(GX) - 1j * np.sin(GX)
    return Sf


class PsPotNL:
    """PsPotNL.jl:16-77 (betaNL) and _init_prj2beta (:147-195).

    beta order per atom: l = 0..lmax, m = -l..l, iprj.  `D` is the Nbeta x Nbeta
    block-diagonal matrix with D[ibeta, jbeta] = h[l, iprj, jprj] inside one
    (atom, l, m) block: op_V_Ps_nloc.jl:59-76 sums exactly hij * beta_i (beta_j^H psi).
    `blocks` lists (ia, l, m, [ibeta...]) for the energy/force loops."""

    def __init__(self, atoms, pw, pspots):
        self.blocks = []
        nb = 0
        for ia in range(atoms.Natoms):
            psp = pspots[atoms.atm2species[ia]]
            for l in range(psp.lmax + 1):
                for m in range(-l, l + 1):
                    ids = list(range(nb, nb + int(psp.Nproj_l[l])))
                    nb += len(ids)
                    self.blocks.append((ia, l, m, ids))
        self.NbetaNL = nb
        self.D = np.zeros((nb, nb))
        for ia, l, m, ids in self.blocks:
            psp = pspots[atoms.atm2species[ia]]
            for a, ib in enumerate(ids):
                for b, jb in enumerate(ids):
                    self.D[ib, jb] = psp.h[l, a, b]
        self.betaNL = []
        for ik in range(pw.Nkpt):
            g = pw.gk(ik)
            Gm = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
            beta = np.zeros((pw.Ngw[ik], nb), dtype=complex)
            for ia, l, m, ids in self.blocks:
                psp = pspots[atoms.atm2species[ia]]
                tau = atoms.positions[:, ia]
                GX = tau[0] * g[0] + tau[1] * g[1] + tau[2] * g[2]
                Sf = np.cos(GX) - 1j * np.sin(GX)
                ylm = Ylm_real(l, m, g)
                for a, ib in enumerate(ids):
                    beta[:, ib] = (-1j) ** l * ylm * eval_proj_G(psp, l, a + 1, Gm, pw.CellVolume) * Sf
            self.betaNL.append(beta)


def calc_V_Ps_loc(atoms, pw, pspots):
    """Hamiltonian.jl:119-141: V_Ps_loc(r) = N * IFFT[ sum_sp S_sp(G) V_sp(|G|) / Omega ]."""
    strf = calc_strfact(atoms.positions, atoms.atm2species, atoms.Nspecies, pw.G)
    Vg = np.zeros(pw.Npoints, dtype=complex)
    for isp in range(atoms.Nspecies):
        Vgl = eval_Vloc_G(pspots[isp], pw.G2_shells)
        np.add.at(Vg, pw.idx_g2r, strf[:, isp] * Vgl[pw.idx_g2shells] / pw.CellVolume)
    return np.real(G_to_R(pw, Vg)) * pw.Npoints
