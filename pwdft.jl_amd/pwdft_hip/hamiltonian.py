"""Hamiltonian state and the operator API of PWDFT.jl, executed by libpwdft_hip.

Mirrors the reference names (Julia `f!` -> Python `f_`):
  op_H(Ham, psi)            op_H.jl:4-27          op_H_(Ham, psi, Hpsi)   op_H.jl:53-63
  op_K(Ham, psi)            op_K.jl:44-75         op_K_(Ham, psi, Hpsi)   (accumulates)
  op_V_loc(Ham, psi)        op_V_loc.jl:173-184   op_V_loc_(...)          op_V_loc.jl:33-92
  op_V_Ps_nloc(Ham, psi)    op_V_Ps_nloc.jl:26-30 op_V_Ps_nloc_(...)      op_V_Ps_nloc.jl:39-78
  calc_rhoe(Ham, psiks)     calc_rhoe.jl:1-7      calc_rhoe_(Ham, psiks, Rhoe)
  Kprec(Ham, psi)           Kprec.jl:10-28
psi is a numpy (Ngw x Nst) complex128 array (host entry points) or a DeviceMatrix
(device entry points); a list of them is a BlochWavefunc indexed ik + ispin*Nkpt.
Ham.ik / Ham.ispin are the reference's mutable cursor (Hamiltonian.jl:16-17).

op_H_ skips op_Vexx when there is no EXX object: the reference's op_H! calls op_Vexx!
unconditionally and fails for every Hamiltonian it builds (op_H.jl:61, Hamiltonian.jl:293).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import (PW_ACCUMULATE, PW_MIXED, PW_TERM_ALL, PW_TERM_K, PW_TERM_NLOC, PW_TERM_VLOC, PW_XC_PBE,
                   PW_XC_VWN, Context, DeviceMatrix, DeviceVector, ptr)
from .basis import (PsPot_GTH, PsPotNL, PWGrid, RhoeSymmetrizer, crystal_symmetry,
                    eval_Vloc_G, kpoint_grid, kpoint_rotations)
from .potentials import calc_E_NN, guess_rhoe


class Electrons:
    """Electrons.jl:121-270 for Nspin_wf = 1: Nstates, Nstates_occ and initial Focc."""

    def __init__(self, Nelectrons, Nkpt, Nstates=None, extra_states=None, use_smearing=False):
        Ne = float(Nelectrons)
        nmin = int(round(Ne / 2))
        if nmin * 2 < Ne:
            nmin += 1
        if Nstates is None and use_smearing and extra_states is None:
            Nstates = max(math.ceil(1.2 * Ne / 2), math.ceil(Ne / 2) + 4)
            nempty = Nstates - math.ceil(Ne / 2)
        elif Nstates is None:
            nempty = extra_states or 0
            Nstates = nmin + nempty
        else:
            if Nstates < nmin:
                raise ValueError(f"Given Nstates is not enough: Nstates = {Nstates}, NstatesMin = {nmin}")
            nempty = Nstates - nmin
        self.Nelectrons = Ne
        self.Nstates = int(Nstates)
        self.Nstates_occ = int(Nstates - nempty)
        self.Focc = np.zeros((self.Nstates, Nkpt))
        self.Focc[: self.Nstates_occ - 1] = 2.0
        self.Focc[self.Nstates_occ - 1] = 1.0 if int(round(Ne)) % 2 == 1 else 2.0
        self.ebands = np.zeros((self.Nstates, Nkpt))
        self.E_fermi = 0.0
        self.kT = 0.0
        self.use_smearing = use_smearing


class Potentials:
    """Potentials.jl:7-14 (Nspin = 1).  Ps_loc is built on the host (once per geometry);
    Hartree, XC and Total are produced in device memory by Hamiltonian.update and copied to
    the host the first time they are read after it.  A host copy that has been read (and
    possibly modified in place, as set_Vloc does) stays authoritative until the next update."""

    _SHAPES = {"Hartree": False, "XC": True, "Total": True}   # name -> stored as N x 1

    def __init__(self, ctx, Npoints):
        self.Ps_loc = np.zeros(Npoints)
        self.dev = {k: DeviceVector(ctx, Npoints) for k in self._SHAPES}
        self._host = {k: (np.zeros((Npoints, 1)) if col else np.zeros(Npoints)) for k, col in self._SHAPES.items()}

    def _get(self, name):
        h = self._host[name]
        if h is None:
            v = self.dev[name].download()
            h = v.reshape(-1, 1) if self._SHAPES[name] else v
            self._host[name] = h
        return h

    def _set(self, name, v):
        v = np.array(v, dtype=float)
        self._host[name] = v.reshape(-1, 1) if self._SHAPES[name] else v.reshape(-1)

    def device_updated(self):
        """The device arrays hold the current Hartree / XC / Total (after update)."""
        self._host = dict.fromkeys(self._host)

    Hartree = property(lambda self: self._get("Hartree"), lambda self, v: self._set("Hartree", v))
    XC = property(lambda self: self._get("XC"), lambda self, v: self._set("XC", v))
    Total = property(lambda self: self._get("Total"), lambda self, v: self._set("Total", v))


class Hamiltonian:
    """Hamiltonian(atoms, pspfiles, ecutwfc; kw...) of Hamiltonian.jl:337-405 (GTH, dual 4,
    Nspin 1).  Builds the host set-up data and uploads what op_H reads to the device."""

    def __init__(self, atoms, pspfiles, ecutwfc, meshk=(1, 1, 1), shiftk=(0, 0, 0), Ns_=None,
                 kpoints=None, rots=None, extra_states=None, Nstates=None, use_smearing=False,
                 smearing_kT=0.0, xcfunc="VWN", device=0, band_batch=64, mixed=False,
                 use_symmetry=True, time_reversal=True, autotune=False):
        if len(pspfiles) != atoms.Nspecies:
            raise ValueError(f"Length of pspfiles is not equal to Nspecies={atoms.Nspecies}")
        self.pspots = [p if isinstance(p, PsPot_GTH) else PsPot_GTH(p) for p in pspfiles]
        for isp, psp in enumerate(self.pspots):
            if psp.atsymb.lower() != atoms.SpeciesSymbols[isp].lower():
                raise ValueError("Atomic symbols and pspot atsymb does not match")
        if xcfunc not in ("VWN", "LDA", "PBE"):
            raise NotImplementedError("functionals: LDA (Slater + VWN) and PBE")
        self.atoms = atoms
        self.xcfunc = xcfunc
        # symmetry (Hamiltonian.jl:53-83, 261-265): the crystal's operations reduce the
        # k-mesh, and the density built from the reduced set is symmetrised
        self.sym_ops = crystal_symmetry(atoms) if use_symmetry else [(np.eye(3, dtype=np.int64), np.zeros(3))]
        if kpoints is None:
            if rots is None:
                rots = kpoint_rotations(self.sym_ops)
            kpoints = kpoint_grid(atoms.LatVecs, meshk, shiftk, rots=rots, time_reversal=time_reversal)
        self.pw = PWGrid(ecutwfc, atoms.LatVecs, kpoints=kpoints, Ns=Ns_)
        self.rhoe_symmetrizer = (RhoeSymmetrizer(atoms, self.pw, self.sym_ops) if len(self.sym_ops) > 1
                                 else None)
        self.pspotNL = PsPotNL(atoms, self.pw, self.pspots)
        self.Zvals = np.array([p.zval for p in self.pspots])
        Ne = float(np.sum(self.Zvals[atoms.atm2species]))
        self.electrons = Electrons(Ne, self.pw.Nkpt, Nstates=Nstates, extra_states=extra_states,
                                   use_smearing=use_smearing)
        self.electrons.kT = smearing_kT
        self.ik = 0
        self.ispin = 0
        self.mixed = mixed
        self.ctx = Context(device)
        self._rho_scratch = None
        self.potentials = Potentials(self.ctx, self.pw.Npoints)
        self.E_NN = calc_E_NN(atoms, self.Zvals)
        self.band_batch = int(band_batch)
        self.ctx.call("pw_set_band_batch", self.band_batch)
        self._upload_static()          # also builds V_Ps_loc and betaNL on the device
        self.potentials.Total[:, 0] = self.potentials.Ps_loc
        self._upload_vloc()
        # op_H launch parameters for this shape (autotune.py): from the tuning cache, or timed
        # once on the device and cached; (settings, "cache" | "tuned"), or None
        self.tuned = None
        if autotune:
            from .autotune import autotune as _autotune
            self.tuned = _autotune(self, mixed=mixed)

    # ---- device state ---------------------------------------------------------------
    def _upload_static(self):
        pw, c = self.pw, self.ctx
        n1, n2, n3 = pw.Ns
        c.call("pw_set_grid", n1, n2, n3, float(pw.CellVolume))
        nb = self.pspotNL.NbetaNL
        D = np.asfortranarray(self.pspotNL.D, dtype=float)
        c.call("pw_set_dij", nb, ptr(D))
        for ik in range(pw.Nkpt):
            idx = np.ascontiguousarray(pw.idx_gw2r[ik], dtype=np.int64)
            gk = np.ascontiguousarray(pw.gk(ik).T, dtype=float)  # ngw x 3 row-major == 3 x ngw col-major
            c.call("pw_set_kpoint", ik, int(pw.Ngw[ik]), ptr(idx), 0, ptr(gk))
        # rho sphere and V_sp(|G|) of every species for the force / energy-gradient kernels
        vg = np.empty((pw.Ng, self.atoms.Nspecies), order="F")
        for isp, psp in enumerate(self.pspots):
            vg[:, isp] = eval_Vloc_G(psp, pw.G2_shells)[pw.idx_g2shells]
        idx = np.ascontiguousarray(pw.idx_g2r, dtype=np.int64)
        G = np.ascontiguousarray(pw.G.T, dtype=float)
        c.call("pw_set_rho_sphere", int(pw.Ng), ptr(idx), 0, ptr(G), self.atoms.Nspecies, ptr(vg))
        if self.rhoe_symmetrizer is not None:
            self.rhoe_symmetrizer.upload(c)
        self._upload_atoms()
        self._build_geometry()

    def _upload_atoms(self):
        atoms = self.atoms
        tau = np.asfortranarray(atoms.positions, dtype=float)
        sp = np.ascontiguousarray(atoms.atm2species, dtype=np.int32)
        ba = np.array([ia for ia, l, m, ids in self.pspotNL.blocks for _ in ids], dtype=np.int32)
        self.ctx.call("pw_set_atoms", atoms.Natoms, ptr(tau), ptr(sp), len(ba),
                      ptr(ba) if len(ba) else None)

    def _build_geometry(self):
        """Per-geometry set-up on the device for the current positions: betaNL of every
        k-point (PsPotNL.jl:16-77) and V_Ps_loc (Hamiltonian.jl:119-141), whose host copy
        becomes potentials.Ps_loc."""
        c, nl = self.ctx, self.pspotNL
        if nl.NbetaNL > 0:
            desc = np.ascontiguousarray(nl.projector_table())
            rc = np.ascontiguousarray([p.rc[:4] for p in self.pspots], dtype=float)
            for ik in range(self.pw.Nkpt):
                c.call("pw_build_projectors_dev", ik, nl.NbetaNL, ptr(desc), len(self.pspots), ptr(rc))
        V = np.empty(self.pw.Npoints)
        c.call("pw_build_vps_loc_dev", ptr(V))
        self.potentials.Ps_loc = V

    def set_nonlocal_deeq(self, Deeq, nh, indv_ijkb0, betaNL=None):
        """Nonlocal data in the form of a norm-conserving UPF Hamiltonian (PsPotNL_UPF):
        Deeq (nhm x nhm x Natoms, real symmetric) couples the nh[isp] projectors of each
        atom, whose first projector is column indv_ijkb0[ia] (0-based) of betaNL[ik]
        (Ngw[ik] x NbetaNL; default: the projectors already in place).  op_V_Ps_nloc then
        applies Hpsi += beta (D (beta' psi)) with D the block-diagonal matrix of the Deeq
        blocks (op_V_Ps_nloc.jl:82-126), and the nonlocal energy and forces use the same D
        (calc_energies.jl:61-128, calc_forces_Ps_nloc.jl:31-160 with Deff = Deeq)."""
        atoms, pw, c = self.atoms, self.pw, self.ctx
        nbeta = self.pspotNL.NbetaNL if betaNL is None else int(np.shape(betaNL[0])[1])
        Deeq = np.asarray(Deeq, dtype=float)
        if Deeq.ndim != 3 or Deeq.shape[2] != atoms.Natoms or Deeq.shape[0] != Deeq.shape[1]:
            raise ValueError("Deeq must be nhm x nhm x Natoms")
        D = np.zeros((nbeta, nbeta), order="F")
        owner = np.full(nbeta, -1, dtype=np.int32)
        for ia in range(atoms.Natoms):
            n = int(nh[atoms.atm2species[ia]])
            o = int(indv_ijkb0[ia])
            if n > Deeq.shape[0] or o < 0 or o + n > nbeta or np.any(owner[o:o + n] >= 0):
                raise ValueError(f"projector block of atom {ia} does not fit NbetaNL = {nbeta}")
            blk = Deeq[:n, :n, ia]
            if np.abs(blk - blk.T).max() > 1e-12 * max(1.0, np.abs(blk).max()):
                raise ValueError(f"Deeq block of atom {ia} is not symmetric")
            D[o:o + n, o:o + n] = blk
            owner[o:o + n] = ia
        if np.any(owner < 0):
            raise ValueError("every projector must belong to one atom block")
        betas = self.pspotNL.betaNL if betaNL is None else betaNL
        c.call("pw_set_dij", nbeta, ptr(D))
        for ik in range(pw.Nkpt):
            b = np.asfortranarray(betas[ik], dtype=np.complex128)
            if b.shape != (pw.Ngw[ik], nbeta):
                raise ValueError(f"betaNL[{ik}] must be Ngw x NbetaNL")
            c.call("pw_set_projectors", ik, nbeta, ptr(b), int(pw.Ngw[ik]))
        tau = np.asfortranarray(atoms.positions, dtype=float)
        sp = np.ascontiguousarray(atoms.atm2species, dtype=np.int32)
        c.call("pw_set_atoms", atoms.Natoms, ptr(tau), ptr(sp), nbeta, ptr(owner) if nbeta else None)
        self._deeq_nonlocal = True

    def set_positions(self, positions):
        """Move the atoms (Cartesian bohr, 3 x Natoms): rebuild the structure-factor
        dependent set-up as Hamiltonian() does -- betaNL and V_Ps_loc on the device, E_NN on
        the host; V_loc is rebuilt from V_Ps_loc + the current Hartree and XC parts."""
        if getattr(self, "_deeq_nonlocal", False):
            raise ValueError("set_positions rebuilds GTH projectors; after set_nonlocal_deeq pass the moved "
                             "projectors to set_nonlocal_deeq instead")
        atoms = self.atoms
        old = atoms.positions
        atoms.positions = np.array(positions, dtype=float).reshape(3, atoms.Natoms)
        if len(self.sym_ops) > 1:
            keep = {R.tobytes() for R, _ in crystal_symmetry(atoms)}
            if any(R.tobytes() not in keep for R, _ in self.sym_ops):
                atoms.positions = old
                raise ValueError("set_positions: the new positions break symmetries used to reduce the "
                                 "k-points; build the Hamiltonian with use_symmetry=False")
            self.sym_ops = [op for op in crystal_symmetry(atoms)
                            if any(np.array_equal(op[0], R) for R, _ in self.sym_ops)]
            self.rhoe_symmetrizer = RhoeSymmetrizer(atoms, self.pw, self.sym_ops)
            self.rhoe_symmetrizer.upload(self.ctx)
        pw, c = self.pw, self.ctx
        self.pspotNL = PsPotNL(atoms, pw, self.pspots)
        self._upload_atoms()
        self._build_geometry()
        self.potentials.Total[:, 0] = self.potentials.Ps_loc + self.potentials.Hartree + self.potentials.XC[:, 0]
        self.E_NN = calc_E_NN(atoms, self.Zvals)
        self._upload_vloc()

    def _upload_vloc(self):
        V = np.ascontiguousarray(self.potentials.Total[:, 0])
        self.ctx.call("pw_set_vloc", 0, ptr(V))

    def set_Vloc(self, V):
        """Replace Ham.potentials.Total[:, 1] and upload it."""
        self.potentials.Total[:, 0] = V
        self._upload_vloc()

    @property
    def xc_code(self):
        return PW_XC_PBE if self.xcfunc == "PBE" else PW_XC_VWN

    def install_potentials(self):
        """V_Total = V_Ps_loc + Hartree + XC from the device arrays becomes the op_H
        potential (after the Hartree / XC arrays were changed on the device, as potential
        mixing does)."""
        p = self.potentials
        self.ctx.call("pw_set_vloc_parts_dev", 0, p.dev["Hartree"].p, p.dev["XC"].p, p.dev["Total"].p)
        p.device_updated()

    def rho_scratch(self):
        """A device vector of Npoints doubles owned by the Hamiltonian (staging of host
        densities for the device entry points)."""
        if self._rho_scratch is None:
            self._rho_scratch = DeviceVector(self.ctx, self.pw.Npoints)
        return self._rho_scratch

    def update(self, rho):
        """update!(Ham, psiks, Rhoe) for Nspin = 1 (Hamiltonian.jl:510-545), on the device:
        Poisson_solve!, Slater + VWN or PBE (xcfunc, Hamiltonian.jl:526-533) and V_Total,
        which becomes the op_H potential.  rho is a host array or a DeviceVector.  Returns
        (E_Ps_loc, E_Hartree, E_xc) of this density, the local terms of calc_energies."""
        if isinstance(rho, DeviceVector):
            dev = rho
        else:
            dev = self.rho_scratch()
            dev.upload(np.asarray(rho, dtype=float).reshape(-1))
        p = self.potentials
        E = np.zeros(3)
        self.ctx.call("pw_update_potential_dev", 0, self.xc_code, 1, dev.p, p.dev["Hartree"].p, p.dev["XC"].p,
                      p.dev["Total"].p, ptr(E))
        p.device_updated()
        return tuple(float(x) for x in E)

    def guess_rhoe(self):
        return guess_rhoe(self.atoms, self.pw, self.pspots, self.electrons.Nelectrons)

    @property
    def flags_prec(self):
        return PW_MIXED if self.mixed else 0


# ---- operator API -------------------------------------------------------------------

def _apply(Ham, psi, terms, out=None, accumulate=False):
    flags = terms | Ham.flags_prec | (PW_ACCUMULATE if accumulate else 0)
    ik, ispin = Ham.ik, Ham.ispin
    ngw = Ham.pw.Ngw[ik]
    if isinstance(psi, DeviceMatrix):
        if psi.rows != ngw:
            raise ValueError("psi rows differ from Ngw[ik]")
        if out is None:
            out = DeviceMatrix(Ham.ctx, psi.rows, psi.cols)
        Ham.ctx.call("pw_op_H_dev", ik, ispin, psi.cols, psi.p, psi.rows, out.p, out.rows, flags)
        return out
    a = np.asfortranarray(psi, dtype=np.complex128)
    if a.ndim == 1:
        a = a.reshape(-1, 1, order="F")
    if a.shape[0] != ngw:
        raise ValueError(f"psi has {a.shape[0]} rows, Ngw[{ik}] = {ngw}")
    if out is None:
        out = np.zeros_like(a)
    elif not (out.flags.f_contiguous and out.dtype == np.complex128):
        raise ValueError("Hpsi must be a Fortran-ordered complex128 array")
    Ham.ctx.call("pw_op_H", ik, ispin, a.shape[1], ptr(a), ngw, ptr(out), ngw, flags)
    return out


def _per_k(Ham, psiks, fn):
    out = []
    for ispin in range(1):
        for ik in range(Ham.pw.Nkpt):
            Ham.ik, Ham.ispin = ik, ispin
            out.append(fn(psiks[ik + ispin * Ham.pw.Nkpt]))
    return out


def _ptr_array(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*[p.value for p in ptrs])


def op_H_batch(Ham, psiks, out=None, terms=PW_TERM_ALL, accumulate=False, iks=None):
    """op_H over the k-points iks (default: all) of a BlochWavefunc of DeviceMatrix in one
    launch per stage (pw_op_H_batch_dev); every k-point must have the same band count.
    Returns the list of H psi (DeviceMatrix)."""
    iks = list(range(Ham.pw.Nkpt)) if iks is None else list(iks)
    nb = psiks[0].cols
    if any(p.cols != nb for p in psiks):
        raise ValueError("op_H_batch: every k-point needs the same number of bands")
    if out is None:
        out = [DeviceMatrix(Ham.ctx, p.rows, p.cols) for p in psiks]
    flags = terms | Ham.flags_prec | (PW_ACCUMULATE if accumulate else 0)
    ik_a = np.ascontiguousarray(iks, dtype=np.int32)
    ldp = np.ascontiguousarray([p.rows for p in psiks], dtype=np.int64)
    ldh = np.ascontiguousarray([h.rows for h in out], dtype=np.int64)
    Ham.ctx.call("pw_op_H_batch_dev", len(iks), ptr(ik_a), Ham.ispin, nb, _ptr_array([p.p for p in psiks]),
                 ptr(ldp), _ptr_array([h.p for h in out]), ptr(ldh), flags)
    return out


def op_H(Ham, psi):
    """op_H.jl:4-27 (allocating): K + V_loc + V_Ps_nloc (+ op_Vtau = 0 for LDA).  A list of
    DeviceMatrix (a BlochWavefunc on the device) goes through the batched launch."""
    if isinstance(psi, list):
        if all(isinstance(p, DeviceMatrix) for p in psi) and len({p.cols for p in psi}) == 1:
            return op_H_batch(Ham, psi)
        return _per_k(Ham, psi, lambda p: op_H(Ham, p))
    return _apply(Ham, psi, PW_TERM_ALL)


def op_H_(Ham, psi, Hpsi):
    """op_H! (op_H.jl:53-63): Hpsi is overwritten with H psi."""
    _apply(Ham, psi, PW_TERM_ALL, out=Hpsi, accumulate=False)


def op_K(Ham, psi):
    if isinstance(psi, list):
        return _per_k(Ham, psi, lambda p: op_K(Ham, p))
    return _apply(Ham, psi, PW_TERM_K)


def op_K_(Ham, psi, Hpsi):
    """op_K! accumulates into Hpsi (op_K.jl:55-75)."""
    _apply(Ham, psi, PW_TERM_K, out=Hpsi, accumulate=True)


def op_V_loc(Ham, psi):
    if isinstance(psi, list):
        return _per_k(Ham, psi, lambda p: op_V_loc(Ham, p))
    return _apply(Ham, psi, PW_TERM_VLOC)


def op_V_loc_(Ham, psi, Hpsi):
    """op_V_loc! accumulates into Hpsi (op_V_loc.jl:33-92)."""
    _apply(Ham, psi, PW_TERM_VLOC, out=Hpsi, accumulate=True)


def op_V_Ps_nloc(Ham, psi):
    if isinstance(psi, list):
        return _per_k(Ham, psi, lambda p: op_V_Ps_nloc(Ham, p))
    return _apply(Ham, psi, PW_TERM_NLOC)


def op_V_Ps_nloc_(Ham, psi, Hpsi):
    """op_V_Ps_nloc! accumulates into Hpsi (op_V_Ps_nloc.jl:39-78)."""
    _apply(Ham, psi, PW_TERM_NLOC, out=Hpsi, accumulate=True)


def Kprec(Ham, psi, out=None):
    """Kprec.jl:10-28: psi / (1 + |G+k|^2) (device)."""
    if out is None:
        out = DeviceMatrix(Ham.ctx, psi.rows, psi.cols)
    Ham.ctx.call("pw_kprec_dev", Ham.ik, psi.cols, psi.p, psi.rows, out.p, out.rows)
    return out


def calc_betaNL_psi(Ham, psi):
    """PsPotNL.jl:242-254: conj(psi' * betaNL[ik]) -> Nstates x Nbeta (host array)."""
    nb = Ham.pspotNL.NbetaNL
    dev = psi if isinstance(psi, DeviceMatrix) else DeviceMatrix.from_host(Ham.ctx, psi)
    out = DeviceMatrix(Ham.ctx, nb, dev.cols)
    Ham.ctx.call("pw_betaNL_psi_dev", Ham.ik, dev.cols, dev.p, dev.rows, out.p, nb, Ham.flags_prec)
    return out.download().T.copy()


def calc_rhoe_(Ham, psiks, Rhoe, Focc=None):
    """calc_rhoe! (calc_rhoe.jl:9-135, Nspin = 1): Rhoe[:, 0] is overwritten with
    sum_k wk sum_ist Focc |psi(r)|^2 (psi(r) = N/sqrt(Omega) IFFT psi), symmetrised on the
    device when the k-points were reduced by crystal symmetry (calc_rhoe.jl:130-132).  Rhoe
    is a host array or a DeviceVector (then the density stays in device memory)."""
    c = Ham.ctx
    Focc = Ham.electrons.Focc if Focc is None else Focc
    c.call("pw_rhoe_zero")
    for ik in range(Ham.pw.Nkpt):
        psi = psiks[ik]
        w = np.ascontiguousarray(Ham.pw.wk[ik] * Focc[:, ik], dtype=float)
        if isinstance(psi, DeviceMatrix):
            c.call("pw_rhoe_accum_dev", ik, psi.cols, psi.p, psi.rows, ptr(w), Ham.flags_prec)
        else:
            a = np.asfortranarray(psi, dtype=np.complex128)
            c.call("pw_rhoe_accum", ik, a.shape[1], ptr(a), a.shape[0], ptr(w), Ham.flags_prec)
    dev = Rhoe if isinstance(Rhoe, DeviceVector) else Ham.rho_scratch()
    c.call("pw_rhoe_get_dev", dev.p)
    if Ham.rhoe_symmetrizer is not None:  # calc_rhoe.jl:130-132
        c.call("pw_symmetrize_rhoe_dev", dev.p)
    if dev is Rhoe:
        return Rhoe
    r = dev.download()
    Rhoe = np.asarray(Rhoe)
    if Rhoe.ndim == 2:
        Rhoe[:, 0] = r
    else:
        Rhoe[:] = r
    return Rhoe


def calc_rhoe(Ham, psiks, Focc=None):
    """calc_rhoe.jl:1-7 (allocating): Npoints x 1 array."""
    Rhoe = np.zeros((Ham.pw.Npoints, 1))
    return calc_rhoe_(Ham, psiks, Rhoe, Focc=Focc)
