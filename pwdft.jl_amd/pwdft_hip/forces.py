"""Forces and the adjoint of the pseudopotential energies, executed by libpwdft_hip.

  calc_forces_Ps_loc(Ham, Rhoe)          calc_forces_Ps_loc.jl:1-71
  calc_forces_Ps_nloc(Ham, psiks)        calc_forces_Ps_nloc.jl:291-411 (GTH)
  calc_forces_NN(Ham)                    Ewald ion-ion forces (host, per geometry)
  calc_forces(Ham, psiks, Rhoe)          NN + Ps_loc + Ps_nloc

calc_forces(Ham, psiks) of the reference fails for GTH pseudopotentials: it calls a
one-argument calc_forces_Ps_loc and the UPF-only calc_forces_nlcc (calc_forces.jl:8,14).
The sum above is what that function means for GTH, where the NLCC and SCF-correction terms
are zero (calc_forces_nlcc.jl:20-23, calc_forces_scf_corr.jl:8-11).

rrule_E_Ps_loc / rrule_E_Ps_nloc return (E, pullback) in the ChainRulesCore convention:
pullback(Ebar) gives the cotangents of the inputs the energy depends on -- V_sp(G) on the G
shells and the atomic positions for E_Ps_loc at fixed rho, the positions for E_Ps_nloc at
fixed psi.  The gradients come from the library's adjoint kernels (pw_E_Ps_loc_grad,
pw_calc_forces_Ps_nloc_dev).
"""
import numpy as np

from ._lib import DeviceMatrix, ptr
from .basis import symmetrize_vector
from .potentials import calc_forces_NN as _forces_NN


def _rho_arg(Ham, Rhoe):
    if Rhoe is None:
        if Ham.rhoe_symmetrizer is not None:
            raise ValueError("pass the symmetrised Rhoe: the device density is not symmetrised")
        return None, None
    r = np.asarray(Rhoe, dtype=float)
    if r.ndim == 2:
        r = r.sum(axis=1)
    r = np.ascontiguousarray(r.reshape(-1))
    if r.size != Ham.pw.Npoints:
        raise ValueError("Rhoe must have Npoints rows")
    return r, ptr(r)


def E_Ps_loc_grad(Ham, Rhoe=None):
    """(E_Ps_loc, dE/dV_sp(G) as Ng x Nspecies, dE/dtau as 3 x Natoms) at fixed rho.
    Rhoe=None uses the density accumulated on the device by the last calc_rhoe."""
    r, rp = _rho_arg(Ham, Rhoe)
    E = np.zeros(1)
    dvg = np.zeros((Ham.pw.Ng, Ham.atoms.Nspecies), order="F")
    dtau = np.zeros((3, Ham.atoms.Natoms), order="F")
    Ham.ctx.call("pw_E_Ps_loc_grad", rp, ptr(E), ptr(dvg), ptr(dtau))
    return float(E[0]), dvg, dtau


def calc_forces_Ps_loc(Ham, Rhoe=None):
    """calc_forces_Ps_loc.jl:1-71: F[:, a] = sum_{G != 0} Re(i G V_sp(G) conj(rho(G)) e^{-i G.tau_a})."""
    r, rp = _rho_arg(Ham, Rhoe)
    F = np.zeros((3, Ham.atoms.Natoms), order="F")
    Ham.ctx.call("pw_calc_forces_Ps_loc", rp, ptr(F))
    return F


def _nloc(Ham, psiks, Focc):
    Focc = Ham.electrons.Focc if Focc is None else Focc
    F = np.zeros((3, Ham.atoms.Natoms), order="F")
    E = np.zeros(1)
    for ik in range(Ham.pw.Nkpt):
        psi = psiks[ik]
        w = np.ascontiguousarray(Ham.pw.wk[ik] * Focc[:, ik], dtype=float)
        if isinstance(psi, DeviceMatrix):
            Ham.ctx.call("pw_calc_forces_Ps_nloc_dev", ik, psi.cols, psi.p, psi.rows, ptr(w), ptr(F), ptr(E))
        else:
            a = np.asfortranarray(psi, dtype=np.complex128)
            Ham.ctx.call("pw_calc_forces_Ps_nloc", ik, a.shape[1], ptr(a), a.shape[0], ptr(w), ptr(F), ptr(E))
    return float(E[0]), F


def calc_forces_Ps_nloc(Ham, psiks, Focc=None):
    """calc_forces_Ps_nloc.jl:291-343 (GTH):
    F[:, a] += 2 wk Focc Re(h_ij conj(<beta_i|psi>) <i (G+k) beta_j|psi>), symmetrised over
    the crystal's operations when the k-points were reduced by them (calc_forces.jl:11)."""
    F = _nloc(Ham, psiks, Focc)[1]
    return symmetrize_vector(Ham.atoms, Ham.sym_ops, F)


def calc_forces_NN(Ham):
    return _forces_NN(Ham.atoms, Ham.Zvals)


def calc_forces(Ham, psiks, Rhoe=None, Focc=None):
    """Total Hellmann-Feynman force (3 x Natoms) for GTH pseudopotentials."""
    return calc_forces_NN(Ham) + calc_forces_Ps_loc(Ham, Rhoe) + calc_forces_Ps_nloc(Ham, psiks, Focc)


def rrule_E_Ps_loc(Ham, Rhoe=None):
    """E_Ps_loc(V_sp on G shells, positions; rho fixed) and its pullback.
    pullback(Ebar) -> (Ebar * dE/dV_sp(shell) as Nshells x Nspecies, Ebar * dE/dtau)."""
    E, dvg, dtau = E_Ps_loc_grad(Ham, Rhoe)
    pw = Ham.pw
    nsh = len(pw.G2_shells)
    dshell = np.zeros((nsh, dvg.shape[1]))
    for isp in range(dvg.shape[1]):
        dshell[:, isp] = np.bincount(pw.idx_g2shells, weights=dvg[:, isp], minlength=nsh)

    def pullback(Ebar):
        return Ebar * dshell, Ebar * dtau

    return E, pullback


def rrule_E_Ps_nloc(Ham, psiks, Focc=None):
    """E_Ps_nloc(positions; psi fixed) and its pullback: pullback(Ebar) -> Ebar * dE/dtau."""
    E, F = _nloc(Ham, psiks, Focc)

    def pullback(Ebar):
        return -Ebar * F

    return E, pullback
