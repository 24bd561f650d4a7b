"""pwdft_hip: PWDFT.jl's Kohn-Sham Hpsi / density / block-eigensolver path on MI355X.

The compute runs in libpwdft_hip.so (hand-written gfx950 HIP kernels behind a C ABI,
include/pwdft_hip.h); this package is the host-side mirror of the reference's Julia API
over that ABI (ctypes), plus the host set-up the reference also does on the host.
"""
from ._lib import Context, DeviceMatrix, DeviceVector, Graph, PWError, PosDefException, load
from .basis import (Atoms, PsPot_GTH, PsPotNL, PWGrid, diamond_supercell, gen_lattice_fcc,
                    gen_lattice_sc, kpoint_grid)
from .hamiltonian import (Hamiltonian, Kprec, calc_betaNL_psi, calc_rhoe, calc_rhoe_, op_H, op_H_, op_H_batch,
                          op_K, op_K_, op_V_loc, op_V_loc_, op_V_Ps_nloc, op_V_Ps_nloc_)
from .solvers import (diag_LOBPCG, diag_LOBPCG_, diag_davidson, diag_davidson_, diag_davidson_qe,
                      diag_davidson_qe_, ortho_sqrt_,
                      rand_BlochWavefunc)
from .scf import KS_solve_SCF_, KS_solve_SCF_potmix_, calc_energies
from .batch import diag_LOBPCG_batch
from .chefsi import chebyfilt, diag_CheFSI_, get_ub_lb_lanczos
from .pcg import calc_grad_evals_, diag_Emin_PCG_
from .emin import KS_solve_Emin_PCG_, calc_grad_
from .forces import (E_Ps_loc_grad, calc_forces, calc_forces_NN, calc_forces_Ps_loc,
                     calc_forces_Ps_nloc, rrule_E_Ps_loc, rrule_E_Ps_nloc)

__all__ = [
    "Atoms", "Context", "Graph", "E_Ps_loc_grad", "calc_forces", "calc_forces_NN", "calc_forces_Ps_loc",
    "calc_forces_Ps_nloc", "rrule_E_Ps_loc", "rrule_E_Ps_nloc", "DeviceMatrix", "DeviceVector", "Hamiltonian", "KS_solve_Emin_PCG_", "KS_solve_SCF_", "KS_solve_SCF_potmix_", "calc_grad_", "Kprec", "PWError", "PWGrid",
    "PosDefException", "PsPot_GTH", "PsPotNL", "calc_betaNL_psi", "calc_energies", "calc_rhoe",
    "calc_rhoe_", "calc_grad_evals_", "chebyfilt", "diag_CheFSI_", "diag_Emin_PCG_", "get_ub_lb_lanczos", "diag_LOBPCG", "diag_LOBPCG_", "diag_LOBPCG_batch", "diag_davidson", "diag_davidson_", "diag_davidson_qe", "diag_davidson_qe_", "diamond_supercell",
    "gen_lattice_fcc", "gen_lattice_sc", "kpoint_grid", "load", "op_H", "op_H_", "op_H_batch", "op_K", "op_K_",
    "op_V_Ps_nloc", "op_V_Ps_nloc_", "op_V_loc", "op_V_loc_", "ortho_sqrt_", "rand_BlochWavefunc",
]
