"""CPU oracle: a plain numpy restatement of PWDFT.jl's Kohn-Sham Hpsi / eigensolver path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `pwdft_hip` package and the
HIP library) may import, call or link anything in this directory.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` use it, and only
as the checker.

Every function cites the reference file:line it restates (paths relative to the
PWDFT.jl source tree).  The reference is pure Julia and cannot run in this image,
so each part is pinned by the reference's own known answers (test logs, unit-test
tables), closed forms, brute force on tiny inputs and invariants; the pins live in
`tests/test_oracle_*.py`.

Modules:
  setup      lattice, atoms, PWGrid (G-vectors, G+k spheres), GTH pseudopotentials,
             real spherical harmonics, projectors betaNL, structure factors,
             V_Ps_loc, k-point grids
  ops        op_K, op_V_loc, op_V_Ps_nloc, op_H, Kprec, calc_betaNL_psi
  density    calc_rhoe, guess_rhoe, Poisson_solve, Slater+VWN XC
  solvers    ortho_sqrt, diag_LOBPCG, diag_davidson, diag_davidson_qe, diag_CheFSI
             (chebyfilt, Lanczos bounds), diag_Emin_PCG, calc_grad
  energy     energies, Ewald E_NN, occupations (Fermi-Dirac), KS_solve_SCF
  forces     calc_forces_Ps_loc, calc_forces_Ps_nloc, calc_forces_NN
  symmetry   symmetrize_rhoe
  mixing     mix_broyden, mix_pulay, mix_rpulay, mix_ppulay, mix_anderson, mix_adaptive
  nloc_upf   the norm-conserving UPF nonlocal operator (dense Deeq blocks), its energy
             and forces
"""
