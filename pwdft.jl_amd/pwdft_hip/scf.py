"""Energies, occupations and the SCF driver: calc_energies (calc_energies.jl:457-492),
Energies sum (Energies.jl:30-34), calc_Focc / find_E_fermi / calc_entropy
(occupations.jl:6-245) and KS_solve_SCF! (KS_solve_SCF.jl:24-373, mix_method "simple").

The wave functions, the density and the potentials stay on the device for the whole SCF:
LOBPCG/Davidson, calc_rhoe with the density symmetrisation, simple mixing, the potential
update (Poisson, exchange-correlation, V_Total) and the energy terms run through
libpwdft_hip.  Per iteration the host receives the eigenvalues and a few scalars."""
import time

import numpy as np

from ._lib import DeviceMatrix, DeviceVector, ptr
from .hamiltonian import calc_rhoe_
from .batch import diag_LOBPCG_batch
from .mixing import DensityMixer
from .chefsi import diag_CheFSI_
from .pcg import diag_Emin_PCG_
from .solvers import diag_LOBPCG_, diag_davidson_, rand_BlochWavefunc

EPS = np.finfo(float).eps


class Energies:
    """Energies.jl:6-34."""

    def __init__(self):
        self.Kinetic = self.Ps_loc = self.Ps_nloc = self.Hartree = self.XC = self.NN = self.mTS = 0.0

    def total(self):
        return self.Kinetic + self.Ps_loc + self.Ps_nloc + self.Hartree + self.XC + self.NN + self.mTS

    def as_dict(self):
        d = {k: getattr(self, k) for k in ("Kinetic", "Ps_loc", "Ps_nloc", "Hartree", "XC", "NN", "mTS")}
        d["Total"] = self.total()
        return d


def _psi_energies(Ham, psiks, Focc):
    """(E_kin, E_Ps_nloc): calc_energies.jl:342-375 (0.5 sum_k wk sum_ist Focc <psi|G+k|^2|psi>)
    and calc_energies.jl:7-58 (GTH: sum_k wk sum_ist Focc h_ij Re(conj(<b_i|psi>) <b_j|psi>)),
    each k-point reduced on the device (pw_energies_psi_dev)."""
    E = np.zeros(2)
    for ik, psi in enumerate(psiks):
        dev = psi if isinstance(psi, DeviceMatrix) else DeviceMatrix.from_host(Ham.ctx, psi)
        w = np.ascontiguousarray(Ham.pw.wk[ik] * Focc[:, ik], dtype=float)
        Ham.ctx.call("pw_energies_psi_dev", ik, dev.cols, dev.p, dev.rows, ptr(w), ptr(E))
    return float(E[0]), float(E[1])


def calc_E_kin(Ham, psiks, Focc):
    """calc_energies.jl:342-375."""
    return _psi_energies(Ham, psiks, Focc)[0]


def calc_E_Ps_nloc(Ham, psiks, Focc):
    """calc_energies.jl:7-58 (GTH)."""
    return _psi_energies(Ham, psiks, Focc)[1]


def calc_E_local(Ham, rho):
    """calc_energies.jl:245-333: E_Ps_loc = <V_Ps_loc, rho> dV, E_Hartree = 1/2 <V_H, rho> dV
    with the Hartree potential of the last update, E_xc = sum eps_xc[rho] rho dV.  The
    exchange-correlation energy density is evaluated on the device."""
    pw = Ham.pw
    dV = pw.CellVolume / pw.Npoints
    if isinstance(rho, DeviceVector):
        dev, host = rho, rho.download()
    else:
        host = np.asarray(rho, dtype=float).reshape(-1)
        dev = Ham.rho_scratch()
        dev.upload(host)
    E = np.zeros(3)
    Ham.ctx.call("pw_update_potential_dev", 0, Ham.xc_code, 0, dev.p, None, None, None, ptr(E))
    return float(E[0]), 0.5 * float(np.dot(Ham.potentials.Hartree, host)) * dV, float(E[2])


def calc_energies(Ham, psiks, rho, Focc=None, mTS=0.0, E_local=None):
    """calc_energies! (calc_energies.jl:463-492).  E_local: the (E_Ps_loc, E_Hartree, E_xc)
    that Ham.update(rho) returned for this same density, which saves evaluating them again."""
    Focc = Ham.electrons.Focc if Focc is None else Focc
    e = Energies()
    e.Kinetic, e.Ps_nloc = _psi_energies(Ham, psiks, Focc)
    e.Ps_loc, e.Hartree, e.XC = calc_E_local(Ham, rho) if E_local is None else E_local
    e.NN = Ham.E_NN
    e.mTS = mTS
    return e


def wgauss(x):
    """occupations.jl:71-84."""
    x = np.asarray(x, dtype=float)
    with np.errstate(over="ignore"):
        v = 1.0 / (1.0 + np.exp(-x))
    return np.where(x < -200.0, 0.0, np.where(x > 200.0, 1.0, v))


def find_E_fermi(Nelectrons, wk, kT, evals, NiterMax=300):
    """occupations.jl:126-183 (bisection, SMALL = 1e-10)."""
    sumkg = lambda e: float(np.sum(2 * wk * np.sum(wgauss((e - evals) / kT), axis=0)))
    Elw = evals[0].min() - 2 * kT
    Eup = evals[-1].max() + 2 * kT
    if sumkg(Eup) - Nelectrons < -EPS or sumkg(Elw) - Nelectrons > EPS:
        raise RuntimeError("Bounds for E_fermi is not found")
    Ef = Ef_old = 0.5 * (Eup + Elw)
    for _ in range(NiterMax):
        s = sumkg(Ef)
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


def calc_Focc(Nelectrons, wk, kT, evals):
    """occupations.jl:6-32 (Nspin = 1)."""
    Ef = find_E_fermi(Nelectrons, wk, kT, evals)
    return 2.0 * wgauss((Ef - evals) / kT), Ef


def calc_entropy(wk, kT, evals, Ef):
    """occupations.jl:222-245 (w1gauss :196-206)."""
    x = (Ef - evals) / kT
    with np.errstate(all="ignore"):
        f = 1.0 / (1.0 + np.exp(-x))
        w1 = np.where(np.abs(x) <= 36.0, f * np.log(f) + (1 - f) * np.log(1 - f), 0.0)
    return float(np.sum(2 * wk[None, :] * kT * w1))


def _solve_bands(Ham, psiks, update_psi, tol, batched, cheby_degree, chefsi_rng, evals):
    """The update_psi dispatch of KS_solve_SCF.jl:187-212 over all k-points; fills evals
    (Nstates x Nkpt) and returns the number of op_H columns applied."""
    el = Ham.electrons
    ncols = 0
    if update_psi == "CheFSI":
        ev, st = diag_CheFSI_(Ham, psiks, cheby_degree, rng=chefsi_rng)
        evals[:, :] = ev
        return st["ncols"]
    if batched and update_psi == "LOBPCG":
        ev, sts = diag_LOBPCG_batch(Ham, psiks, tol=tol, Nstates_conv=el.Nstates_occ)
        evals[:, :] = ev
        return sum(st["ncols"] for st in sts)
    solver = {"LOBPCG": diag_LOBPCG_, "davidson": diag_davidson_, "PCG": diag_Emin_PCG_}[update_psi]
    for ik in range(Ham.pw.Nkpt):
        Ham.ik, Ham.ispin = ik, 0
        ev, st = solver(Ham, psiks[ik], tol=tol, Nstates_conv=el.Nstates_occ)
        evals[:, ik] = ev
        ncols += st["ncols"]
    return ncols


def KS_solve_SCF_(Ham, psiks=None, betamix=0.2, NiterMax=100, etot_conv_thr=1e-6,
                  ethr_evals_last=1e-5, update_psi="LOBPCG", use_smearing=False, kT=1e-3,
                  seed=1234, verbose=False, mixed_until=None, batched=False, mixed_ethr_floor=1e-6,
                  mix_method="simple", mixdim=5, cheby_degree=8):
    """KS_solve_SCF! (KS_solve_SCF.jl:24-373), Nspin = 1, with the reference's mix_method
    choices (simple, linear_adaptive, broyden, pulay, rpulay, ppulay, anderson; mixing.py).

    `mixed_until`: when Ham.mixed is set, switch the Hamiltonian back to FP64 once the
    energy change falls below this threshold (FP64 refinement of the final iterations).
    `batched`: diagonalise all k-points together (batch.diag_LOBPCG_batch) instead of the
    reference's k loop; same arithmetic per k-point.
    `mixed_ethr_floor`: while the Hamiltonian is in mixed precision the eigensolver
    tolerance is not taken below this (FP32 products resolve eigenvalue changes to about
    1e-7 Ha); the FP64 iterations after `mixed_until` use the reference's schedule."""
    pw = Ham.pw
    el = Ham.electrons
    if psiks is None:
        psiks = rand_BlochWavefunc(Ham, seed=seed)
    rho = DeviceVector.from_host(Ham.ctx, Ham.guess_rhoe())
    rho_new = DeviceVector(Ham.ctx, pw.Npoints)
    mixer = DensityMixer(Ham, mix_method, betamix, mixdim)
    Ham.update(rho)
    Focc = el.Focc.copy()
    Etot_old = 0.0
    nconv = 0
    ethr = 0.1
    evals = np.zeros((el.Nstates, pw.Nkpt))
    mTS = 0.0
    history = []
    ncols = 0
    t0 = time.perf_counter()
    if update_psi not in ("LOBPCG", "davidson", "PCG", "CheFSI"):
        raise ValueError(f"Unknown method for update_psi = {update_psi}")
    chefsi_rng = np.random.default_rng(seed + 1)
    e = None
    for it in range(1, NiterMax + 1):
        if it == 1:
            ethr = 0.1
        elif it == 2:
            ethr = 0.01
        else:
            ethr = max(ethr / 5.0, ethr_evals_last)
        tol = max(ethr, mixed_ethr_floor) if Ham.mixed else ethr
        ncols += _solve_bands(Ham, psiks, update_psi, tol, batched, cheby_degree, chefsi_rng, evals)
        if use_smearing:
            Focc, Ef = calc_Focc(el.Nelectrons, pw.wk, kT, evals)
            mTS = calc_entropy(pw.wk, kT, evals, Ef)
            el.E_fermi = Ef
            el.Focc = Focc.copy()
        calc_rhoe_(Ham, psiks, rho_new, Focc=Focc)
        mixer.mix(rho, rho_new, it)       # then clamps at eps()
        E_local = Ham.update(rho)
        e = calc_energies(Ham, psiks, rho, Focc, mTS, E_local=E_local)
        Etot = e.total()
        diffE = abs(Etot - Etot_old)
        history.append(Etot)
        if verbose:
            print(f"SCF: {it:5d} {Etot:18.10f} {diffE:12.5e}{'  (mixed)' if Ham.mixed else ''}", flush=True)
        if Ham.mixed and mixed_until is not None and diffE < mixed_until:
            Ham.mixed = False
            nconv = 0
            Etot_old = Etot
            continue
        nconv = nconv + 1 if diffE < etot_conv_thr else 0
        if nconv >= 2:
            break
        Etot_old = Etot
    el.ebands = evals
    Ham.energies = e
    return {"energies": e.as_dict(), "evals": evals, "psiks": psiks, "rho": rho.download(), "history": history,
            "ncols": ncols, "seconds": time.perf_counter() - t0, "iterations": len(history)}


def KS_solve_SCF_potmix_(Ham, psiks=None, NiterMax=150, betamix=0.2, use_smearing=False, mix_method="simple",
                         mixdim=5, kT=1e-3, update_psi="LOBPCG", cheby_degree=8, etot_conv_thr=1e-6,
                         ethr_evals_last=1e-5, seed=1234, verbose=False, batched=False):
    """KS_solve_SCF_potmix! (KS_solve_SCF_potmix.jl:16-309), Nspin = 1: the Hartree and XC
    potentials are mixed separately (simple, linear_adaptive or broyden) after each
    update! of the new density, then V_Total is rebuilt; energies are those of the new
    density.  All grid fields stay on the device.  With use_smearing the entropy term of
    calc_entropy is used (the reference's driver refers to an undefined `Entropy` there)."""
    if mix_method not in ("simple", "linear_adaptive", "broyden"):
        raise ValueError(f"KS_solve_SCF_potmix: mix_method {mix_method} is not implemented there")
    if update_psi not in ("LOBPCG", "davidson", "PCG", "CheFSI"):
        raise ValueError(f"Unknown method for update_psi = {update_psi}")
    pw, el, ctx = Ham.pw, Ham.electrons, Ham.ctx
    if psiks is None:
        psiks = rand_BlochWavefunc(Ham, seed=seed)
    rho = DeviceVector.from_host(ctx, Ham.guess_rhoe())
    VH_inp = DeviceVector(ctx, pw.Npoints)
    Vxc_inp = DeviceVector(ctx, pw.Npoints)
    mixH = DensityMixer(Ham, mix_method, betamix, mixdim, clamp=False)
    mixX = DensityMixer(Ham, mix_method, betamix, mixdim, clamp=False)
    Ham.update(rho)
    Focc = el.Focc.copy()
    evals = np.zeros((el.Nstates, pw.Nkpt))
    chefsi_rng = np.random.default_rng(seed + 1)
    Etot_old, nconv, ethr, mTS, ncols = 0.0, 0, 0.1, 0.0, 0
    history = []
    e = None
    t0 = time.perf_counter()
    pot = Ham.potentials.dev
    for it in range(1, NiterMax + 1):
        ethr = 0.1 if it == 1 else 0.01 if it == 2 else max(ethr / 5.0, ethr_evals_last)
        ncols += _solve_bands(Ham, psiks, update_psi, ethr, batched, cheby_degree, chefsi_rng, evals)
        if use_smearing:
            Focc, Ef = calc_Focc(el.Nelectrons, pw.wk, kT, evals)
            mTS = calc_entropy(pw.wk, kT, evals, Ef)
            el.E_fermi = Ef
            el.Focc = Focc.copy()
        calc_rhoe_(Ham, psiks, rho, Focc=Focc)
        ctx.call("pw_vec_clamp_dev", pw.Npoints, float(EPS), rho.p)
        Vxc_inp.copy_from(pot["XC"])
        VH_inp.copy_from(pot["Hartree"])
        E_local = Ham.update(rho)
        e = calc_energies(Ham, psiks, rho, Focc, mTS, E_local=E_local)
        Etot = e.total()
        diffE = abs(Etot - Etot_old)
        history.append(Etot)
        if verbose:
            print(f"SCF_potmix: {it:5d} {Etot:18.10f} {diffE:12.5e}", flush=True)
        nconv = nconv + 1 if diffE < etot_conv_thr else 0
        if nconv >= 2:
            break
        Etot_old = Etot
        if mix_method == "simple":
            # V = betamix V_new + (1 - betamix) V_inp
            mixH._lin(pot["Hartree"], betamix, [(1.0 - betamix, VH_inp)])
            mixX._lin(pot["XC"], betamix, [(1.0 - betamix, Vxc_inp)])
        else:
            # mix_broyden!(V_new, V_inp, ...) / mix_adaptive!(V_new, V_inp, ...): the new
            # potential is the first argument and receives the mix
            mixH.mix(pot["Hartree"], VH_inp, it)
            mixX.mix(pot["XC"], Vxc_inp, it)
        Ham.install_potentials()
    el.ebands = evals
    Ham.energies = e
    return {"energies": e.as_dict(), "evals": evals, "psiks": psiks, "rho": rho.download(), "history": history,
            "ncols": ncols, "seconds": time.perf_counter() - t0, "iterations": len(history)}
