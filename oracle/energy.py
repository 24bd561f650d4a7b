"""Oracle: Hamiltonian state, potential update, energies, Ewald, occupations and the
SCF driver -- restated from Hamiltonian.jl, Electrons.jl, calc_energies.jl, Energies.jl,
calc_E_NN.jl, occupations.jl and KS_solve_SCF.jl.  Test infrastructure only."""
import math
import numpy as np
from scipy.special import erfc

from . import ops
from .density import calc_rhoe, guess_rhoe, Poisson_solve, calc_epsxc_Vxc_PBE, calc_epsxc_Vxc_VWN
from .setup import PWGrid, PsPotNL, calc_V_Ps_loc, gen_kpoint_grid
from .solvers import diag_LOBPCG, diag_davidson, rand_wavefunc


class Hamiltonian:
    """Hamiltonian.jl:41-302 reduced to what the hot path reads (Nspin = 1, GTH, dual 4):
    pw, V_Ps_loc, V_H, V_xc, V_loc (= Potentials.Total), pspotNL, electrons, cursor."""

    def __init__(self, atoms, pspots, ecutwfc, meshk=(1, 1, 1), shiftk=(0, 0, 0),
                 Ns=None, rots=None, kpoints=None, extra_states=None, Nstates=None,
                 use_smearing=False, kT=0.0, xcfunc="VWN"):
        self.atoms, self.pspots = atoms, pspots
        self.xcfunc = xcfunc
        if kpoints is None:
            kpoints = gen_kpoint_grid(atoms.LatVecs, meshk, shiftk, rots=rots)
        self.pw = PWGrid(ecutwfc, atoms.LatVecs, kpoints=kpoints, Ns=Ns)
        self.V_Ps_loc = calc_V_Ps_loc(atoms, self.pw, pspots)
        self.V_H = np.zeros(self.pw.Npoints)
        self.V_xc = np.zeros(self.pw.Npoints)
        self.V_loc = np.zeros((self.pw.Npoints, 1))
        self.pspotNL = PsPotNL(atoms, self.pw, pspots)
        self.Zvals = np.array([p.zval for p in pspots])
        # Electrons.jl:121-270 (Nspin_wf = 1)
        Ne = float(sum(self.Zvals[atoms.atm2species]))
        self.Nelectrons = Ne
        is_odd = int(round(Ne)) % 2 == 1
        nmin = int(round(Ne / 2))
        if nmin * 2 < Ne:
            nmin += 1
        if Nstates is None and use_smearing and extra_states is None:
            nst = max(math.ceil(1.2 * Ne / 2), math.ceil(Ne / 2) + 4)
            nempty = nst - math.ceil(Ne / 2)
        elif Nstates is None:
            nst = nmin + (extra_states or 0)
            nempty = extra_states or 0
        else:
            nst = Nstates
            nempty = nst - nmin
        self.Nstates, self.Nstates_occ = nst, nst - nempty
        self.Focc = np.zeros((nst, self.pw.Nkpt))
        self.Focc[: self.Nstates_occ - 1, :] = 2.0
        self.Focc[self.Nstates_occ - 1, :] = 1.0 if is_odd else 2.0
        self.use_smearing, self.kT = use_smearing, kT
        self.ik, self.ispin = 0, 0
        self.E_NN = calc_E_NN(atoms.LatVecs, atoms, self.Zvals)

    def update(self, rho):
        """Hamiltonian.jl:510-545 (update!, Nspin = 1, VWN or PBE)."""
        self.V_H = Poisson_solve(self.pw, rho)
        _, self.V_xc = self.epsxc_vxc(rho)
        self.V_loc[:, 0] = self.V_Ps_loc + self.V_H + self.V_xc

    def epsxc_vxc(self, rho):
        """Slater + VWN, or PBE when xcfunc == "PBE" (Hamiltonian.jl:526-533)."""
        if self.xcfunc == "PBE":
            return calc_epsxc_Vxc_PBE(self.pw, rho)
        return calc_epsxc_Vxc_VWN(rho)


# ----------------------------------------------------------------------------- energies


def calc_E_kin(ham, psiks, Focc):
    """calc_energies.jl:342-375."""
    E = 0.0
    for ik, psi in enumerate(psiks):
        gk2 = ham.pw.gk2(ik)
        E += np.sum(ham.pw.wk[ik] * Focc[:, ik] * np.sum(np.abs(psi) ** 2 * gk2[:, None], axis=0))
    return 0.5 * E


def calc_E_Ps_nloc(ham, psiks, Focc):
    """calc_energies.jl:7-58 (GTH)."""
    nl = ham.pspotNL
    if nl.NbetaNL == 0:
        return 0.0
    E = 0.0
    for ik, psi in enumerate(psiks):
        bp = ops.calc_betaNL_psi(nl.betaNL[ik], psi)
        enl = np.zeros(psi.shape[1])
        for ia, l, m, ids in nl.blocks:
            for ib in ids:
                for jb in ids:
                    enl += nl.D[ib, jb] * np.real(np.conj(bp[:, ib]) * bp[:, jb])
        E += np.sum(ham.pw.wk[ik] * Focc[:, ik] * enl)
    return E


def calc_E_local(ham, rho):
    """calc_energies.jl:245-333: (E_Ps_loc, E_Hartree, E_xc)."""
    dVol = ham.pw.CellVolume / ham.pw.Npoints
    E_H = 0.5 * np.dot(ham.V_H, rho) * dVol
    E_ps = np.dot(ham.V_Ps_loc, rho) * dVol
    epsxc, _ = ham.epsxc_vxc(rho)
    E_xc = np.sum(epsxc * rho) * dVol
    return E_ps, E_H, E_xc


def calc_energies(ham, psiks, rho, Focc, mTS=0.0):
    """calc_energies.jl:463-492 + Energies.jl:30-34 (sum)."""
    e = {"Kinetic": calc_E_kin(ham, psiks, Focc), "Ps_nloc": calc_E_Ps_nloc(ham, psiks, Focc)}
    e["Ps_loc"], e["Hartree"], e["XC"] = calc_E_local(ham, rho)
    e["NN"] = ham.E_NN
    e["mTS"] = mTS
    e["Total"] = e["Kinetic"] + e["Ps_loc"] + e["Ps_nloc"] + e["Hartree"] + e["XC"] + e["NN"] + e["mTS"]
    return e


def calc_E_NN(LatVecs, atoms, Zvals):
    """calc_E_NN.jl:60-172: Ewald ion-ion energy (gcut = 2, ebsl = 1e-8)."""
    t = [LatVecs[:, i] for i in range(3)]
    Omega = abs(np.linalg.det(LatVecs))
    Rec = 2 * np.pi * np.linalg.inv(LatVecs.T)
    g = [Rec[:, i] for i in range(3)]
    tm = [math.sqrt(np.dot(v, v)) for v in t]
    gm = [math.sqrt(np.dot(v, v)) for v in g]
    tau = atoms.positions
    sp = atoms.atm2species
    gcut, ebsl = 2.0, 1e-8
    gexp = -math.log(ebsl)
    eta = math.sqrt(gcut * gcut / gexp) / 2
    Z = Zvals[sp]
    x = float(np.sum(Z ** 2))
    totalcharge = float(np.sum(Z))
    ewald = -2 * eta * x / math.sqrt(np.pi) - np.pi * totalcharge ** 2 / (Omega * eta ** 2)
    tmax = math.sqrt(0.5 * gexp) / eta
    mm = [int(round(tmax / tm[i] + 1.5)) for i in range(3)]
    rng = [np.arange(-mm[i], mm[i] + 1) for i in range(3)]
    I, J, K = np.meshgrid(*rng, indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    T = np.outer(t[0], I) + np.outer(t[1], J) + np.outer(t[2], K)
    nz = (np.abs(I) + np.abs(J) + np.abs(K)) != 0
    for ia in range(atoms.Natoms):
        for ja in range(atoms.Natoms):
            d = (tau[:, ia] - tau[:, ja])[:, None] - T
            r = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
            use = nz if ia == ja else np.ones_like(nz)
            ewald += Z[ia] * Z[ja] * np.sum(erfc(r[use] * eta) / r[use])
    mm = [int(round(gcut / gm[i] + 1.5)) for i in range(3)]
    rng = [np.arange(-mm[i], mm[i] + 1) for i in range(3)]
    I, J, K = np.meshgrid(*rng, indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    nz = (np.abs(I) + np.abs(J) + np.abs(K)) != 0
    Gv = np.outer(g[0], I[nz]) + np.outer(g[1], J[nz]) + np.outer(g[2], K[nz])
    G2 = Gv[0] ** 2 + Gv[1] ** 2 + Gv[2] ** 2
    xg = 4 * np.pi / Omega * np.exp(-0.25 * G2 / eta ** 2) / G2
    for ia in range(atoms.Natoms):
        for ja in range(atoms.Natoms):
            d = tau[:, ia] - tau[:, ja]
            ewald += np.sum(xg * Z[ia] * Z[ja] * np.cos(Gv[0] * d[0] + Gv[1] * d[1] + Gv[2] * d[2]))
    return ewald * 0.5


# ----------------------------------------------------------------------------- occupations


def wgauss(x):
    """occupations.jl:71-84 (Fermi-Dirac)."""
    x = np.asarray(x, dtype=float)
    with np.errstate(over="ignore"):
        v = 1.0 / (1.0 + np.exp(-x))
    return np.where(x < -200.0, 0.0, np.where(x > 200.0, 1.0, v))


def sumkg(wk, kT, evals, ene):
    """occupations.jl:92-121 (Nspin = 1)."""
    return float(np.sum(2 * wk * np.sum(wgauss((ene - evals) / kT), axis=0)))


def find_E_fermi(Nelectrons, wk, kT, evals, NiterMax=300):
    """occupations.jl:126-183: bisection with SMALL = 1e-10."""
    Elw = evals[0, :].min() - 2 * kT
    Eup = evals[-1, :].max() + 2 * kT
    if (sumkg(wk, kT, evals, Eup) - Nelectrons) < -EPS64 or (sumkg(wk, kT, evals, Elw) - Nelectrons) > EPS64:
        raise RuntimeError("Bounds for E_fermi is not found")
    Ef = 0.5 * (Eup + Elw)
    Ef_old = Ef
    for _ in range(NiterMax):
        s = sumkg(wk, kT, evals, Ef)
        if abs(s - Nelectrons) < 1e-10:
            return Ef
        if s - Nelectrons < -1e-10:
            Elw = Ef
        else:
            Eup = Ef
        Ef = 0.5 * (Eup + Elw)
        if abs(Ef - Ef_old) < 1e-10:
            return Ef
        Ef_old = Ef
    return Ef


EPS64 = np.finfo(float).eps


def calc_Focc(Nelectrons, wk, kT, evals):
    """occupations.jl:6-32 (Nspin = 1): 2 * wgauss((E_F - e)/kT)."""
    Ef = find_E_fermi(Nelectrons, wk, kT, evals)
    return 2.0 * wgauss((Ef - evals) / kT), Ef


def calc_entropy(wk, kT, evals, Ef):
    """occupations.jl:222-245 with w1gauss (:196-206)."""
    x = (Ef - evals) / kT
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        f = 1.0 / (1.0 + np.exp(-x))
        w1 = np.where(np.abs(x) <= 36.0, f * np.log(f) + (1 - f) * np.log(1 - f), 0.0)
    return float(np.sum(2 * wk[None, :] * kT * w1))


# ----------------------------------------------------------------------------- SCF


def KS_solve_SCF(ham, psiks=None, betamix=0.2, NiterMax=100, etot_conv_thr=1e-6,
                 ethr_evals_last=1e-5, update_psi="LOBPCG", use_smearing=False, kT=1e-3,
                 seed=1234, verbose=False):
    """KS_solve_SCF.jl:24-373 with mix_method = "simple", Nspin = 1, no symmetrisation."""
    pw = ham.pw
    if psiks is None:
        rng = np.random.default_rng(seed)
        psiks = [rand_wavefunc(rng, pw.Ngw[ik], ham.Nstates) for ik in range(pw.Nkpt)]
    rho = guess_rhoe(ham)
    ham.update(rho)
    Focc = ham.Focc.copy()
    Etot_old = 0.0
    nconv = 0
    ethr = 0.1
    history = []
    evals = np.zeros((ham.Nstates, pw.Nkpt))
    mTS = 0.0
    e = None
    for it in range(1, NiterMax + 1):
        if it == 1:
            ethr = 0.1
        elif it == 2:
            ethr = 0.01
        else:
            ethr = max(ethr / 5.0, ethr_evals_last)
        for ik in range(pw.Nkpt):
            ham.ik = ik
            opH = lambda X: ops.op_H(ham, X)
            prec = lambda R: ops.Kprec(ham, R)
            solver = diag_LOBPCG if update_psi == "LOBPCG" else diag_davidson
            ev, _, _ = solver(opH, prec, psiks[ik], tol=ethr, Nstates_conv=ham.Nstates_occ)
            evals[:, ik] = ev
        if use_smearing:
            Focc, Ef = calc_Focc(ham.Nelectrons, pw.wk, kT, evals)
            mTS = calc_entropy(pw.wk, kT, evals, Ef)
        rho_new = calc_rhoe(ham, psiks, Focc)
        rho = betamix * rho_new + (1 - betamix) * rho
        rho = np.maximum(rho, EPS64)
        ham.update(rho)
        e = calc_energies(ham, psiks, rho, Focc, mTS)
        diffE = abs(e["Total"] - Etot_old)
        history.append(e["Total"])
        if verbose:
            print(f"SCF: {it:5d} {e['Total']:18.10f} {diffE:12.5e}")
        nconv = nconv + 1 if diffE < etot_conv_thr else 0
        if nconv >= 2:
            break
        Etot_old = e["Total"]
    return e, evals, psiks, rho, history
