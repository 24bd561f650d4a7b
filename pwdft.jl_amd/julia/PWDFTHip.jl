"""
    PWDFTHip

Julia binding of libpwdft_hip.so (include/pwdft_hip.h): the drop-in for PWDFT.jl's
Hamiltonian application, density accumulation and eigensolver products on MI355X.

Loading this module after `using PWDFT` adds methods for `Hamiltonian{PsPot_GTH}` that
route op_H / op_K / op_V_loc / op_V_Ps_nloc / calc_rhoe! (with symmetrize_rhoe!) and
update! (Poisson, VWN or PBE, V_Total) through the C ABI.  The op_H family also runs on the
device for norm-conserving `Hamiltonian{PsPot_UPF}` (single grid, collinear), whose
nonlocal term couples each atom's projectors through its dense Deeq block.  The reference's own
KS_solve_SCF! then runs unchanged on top of them (it calls op_H, calc_rhoe! and update!).
diag_LOBPCG! and diag_davidson! themselves run in the library (pw_diag_lobpcg_batch_dev,
pw_diag_lobpcg_dev, pw_diag_davidson_dev): each k-point's block is uploaded once into its
device mirror, every iteration's op_H, Gram products and small dense problems stay on the
device, and the Ritz vectors come back once.  A HIP context is attached to each Hamiltonian the first
time it is used, keyed by object identity; the device update! installs V_Total as the op_H
potential itself, and `sync_potential!(Ham)` re-uploads Ham.potentials.Total for code that
writes it directly.

Index conventions: Julia's 1-based idx_gw2r is passed with index_base = 1; Ham.ik and
Ham.ispin are converted to 0-based at the call.
"""
module PWDFTHip

using PWDFT
import PWDFT: op_H, op_K, op_V_loc, op_V_Ps_nloc, op_H!, op_K!, op_V_loc!, op_V_Ps_nloc!,
              calc_rhoe!, update!, calc_forces_Ps_loc, calc_forces_Ps_nloc, Hamiltonian, PsPot_GTH,
              PsPot_UPF, BlochWavefunc, diag_LOBPCG!, diag_davidson!

# Hamiltonians whose op_H family runs on the device: GTH, and norm-conserving UPF (the
# nonlocal term then couples each atom's projectors through its dense Deeq block)
const HamHip = Union{Hamiltonian{PsPot_GTH},Hamiltonian{PsPot_UPF}}

const libpw = joinpath(@__DIR__, "..", "pwdft_hip", "libpwdft_hip.so")

const PW_TERM_K = UInt32(1)
const PW_TERM_VLOC = UInt32(2)
const PW_TERM_NLOC = UInt32(4)
const PW_TERM_ALL = UInt32(7)
const PW_ACCUMULATE = UInt32(16)
const PW_MIXED = UInt32(32)

mutable struct HipCtx
    h::Ptr{Cvoid}
    mixed::Bool
    grid::Vector{Ptr{Cvoid}}   # device rho, V_H, V_xc, V_Total of update! (N doubles each)
end

const CONTEXTS = IdDict{Any,HipCtx}()

function check(ctx::HipCtx, rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:pw_last_error, libpw), Cstring, (Ptr{Cvoid},), ctx.h))
        rc == -4 && throw(PWDFT.LinearAlgebra.PosDefException(0))
        error("libpwdft_hip error $rc: $msg")
    end
    return nothing
end

"""Attach (once) a HIP context holding the grid, k-point spheres, projectors and V_loc."""
function context(Ham::HamHip; device::Int=0)
    haskey(CONTEXTS, Ham) && return CONTEXTS[Ham]
    check_supported(Ham)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pw_ctx_create, libpw), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device)
    rc == 0 || error("pw_ctx_create failed ($rc)")
    ctx = HipCtx(r[], false, Ptr{Cvoid}[])
    pw = Ham.pw
    Ns = pw.Ns
    check(ctx, ccall((:pw_set_grid, libpw), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble),
                     ctx.h, Ns[1], Ns[2], Ns[3], pw.CellVolume))
    nb = Ham.pspotNL.NbetaNL
    D = nonlocal_D(Ham)
    check(ctx, ccall((:pw_set_dij, libpw), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), ctx.h, nb, D))
    G = pw.gvec.G
    for ik in 1:pw.gvecw.kpoints.Nkpt
        idx = pw.gvecw.idx_gw2r[ik]
        gk = G[:, pw.gvecw.idx_gw2g[ik]] .+ pw.gvecw.kpoints.k[:, ik]
        check(ctx, ccall((:pw_set_kpoint, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Cint, Ptr{Float64}),
                         ctx.h, ik - 1, length(idx), idx, 1, gk))
        if nb > 0
            β = Ham.pspotNL.betaNL[ik]
            check(ctx, ccall((:pw_set_projectors, libpw), Cint,
                             (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64),
                             ctx.h, ik - 1, nb, β, size(β, 1)))
        end
    end
    CONTEXTS[Ham] = ctx
    sync_potential!(Ham)
    if Ham isa Hamiltonian{PsPot_GTH}
        sync_atoms!(Ham)
        sync_symmetry!(Ham)
        check(ctx, ccall((:pw_set_vps_loc, libpw), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.h, Ham.potentials.Ps_loc))
    end
    return ctx
end

check_supported(Ham::Hamiltonian{PsPot_GTH}) = nothing
function check_supported(Ham::Hamiltonian{PsPot_UPF})
    Ham.need_overlap && error("PWDFTHip: ultrasoft and PAW pseudopotentials (need_overlap) are not supported")
    Ham.options.noncollinear && error("PWDFTHip: noncollinear Hamiltonians are not supported")
    Ham.pw.using_dual_grid && error("PWDFTHip: dual (smooth + dense) grids are not supported")
    Deeq = Ham.pspotNL.Deeq
    all(s -> Deeq[:, :, :, s] == Deeq[:, :, :, 1], 1:size(Deeq, 4)) ||
        error("PWDFTHip: spin-dependent Deeq is not supported")
    return nothing
end

"""GTH: D[ib, jb] = h[l, i, j] on the prj2beta blocks (PsPotNL.jl:147-195), what
op_V_Ps_nloc.jl:39-78 sums term by term."""
function nonlocal_D(Ham::Hamiltonian{PsPot_GTH})
    nb = Ham.pspotNL.NbetaNL
    D = zeros(Float64, nb, nb)
    atoms = Ham.atoms
    for ia in 1:atoms.Natoms
        psp = Ham.pspots[atoms.atm2species[ia]]
        for l in 0:psp.lmax, m in -l:l, i in 1:psp.Nproj_l[l+1], j in 1:psp.Nproj_l[l+1]
            ib = Ham.pspotNL.prj2beta[i, ia, l+1, m+psp.lmax+1]
            jb = Ham.pspotNL.prj2beta[j, ia, l+1, m+psp.lmax+1]
            D[ib, jb] = psp.h[l+1, i, j]
        end
    end
    return D
end

"""Norm-conserving UPF: Deeq[1:nh, 1:nh, ia] on the diagonal block of atom ia starting at
indv_ijkb0[ia] + 1 (op_V_Ps_nloc.jl:82-126)."""
function nonlocal_D(Ham::Hamiltonian{PsPot_UPF})
    nl = Ham.pspotNL
    D = zeros(Float64, nl.NbetaNL, nl.NbetaNL)
    atoms = Ham.atoms
    for ia in 1:atoms.Natoms
        n = nl.nh[atoms.atm2species[ia]]
        r = (nl.indv_ijkb0[ia]+1):(nl.indv_ijkb0[ia]+n)
        D[r, r] .= nl.Deeq[1:n, 1:n, ia, 1]
    end
    return D
end

"""symmetrize_rhoe! tables (RhoeSymmetrizer.jl:111-236) from Ham.sym_info: the library's
operation s maps crystal coordinates m of G to inv_s m with phase exp(2 pi i t.(inv_s m)),
t = -ft in the reference's sign convention."""
function sync_symmetry!(Ham::Hamiltonian{PsPot_GTH})
    ctx = CONTEXTS[Ham]
    si = Ham.sym_info
    si.Nsyms > 1 || return nothing
    M = zeros(Cint, 9 * si.Nsyms)
    t = zeros(Float64, 3 * si.Nsyms)
    for s in 1:si.Nsyms, i in 1:3
        for j in 1:3
            M[9 * (s - 1) + 3 * (i - 1) + j] = round(Cint, si.inv_s[i, j, s])
        end
        t[3 * (s - 1) + i] = -si.ft[i, s]
    end
    gm = round.(Cint, (Ham.pw.LatVecs' / (2 * pi)) * Ham.pw.gvec.G)   # 3 x Ng, column-major
    check(ctx, ccall((:pw_set_symmetry, libpw), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}, Cint, Ptr{Cint}),
                     ctx.h, si.Nsyms, M, t, Ham.pw.gvec.Ng, gm))
    return nothing
end

function grid_buffers(ctx::HipCtx, N::Int)
    if isempty(ctx.grid)
        for _ in 1:4
            r = Ref{Ptr{Cvoid}}(C_NULL)
            check(ctx, ccall((:pw_malloc, libpw), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), ctx.h, 8 * N, r))
            push!(ctx.grid, r[])
        end
    end
    return ctx.grid
end

h2d!(ctx, d, a) = check(ctx, ccall((:pw_memcpy_h2d, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Csize_t),
                                   ctx.h, d, a, 8 * length(a)))
d2h!(ctx, a, d) = check(ctx, ccall((:pw_memcpy_d2h, libpw), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Csize_t),
                                   ctx.h, a, d, 8 * length(a)))

# update! (Hamiltonian.jl:510-545, Nspin = 1): Poisson_solve!, calc_Vxc_VWN! / calc_Vxc_PBE!
# and V_Total on the device; Ham.potentials receives the three fields and V_Total is the op_H
# potential from here on.
function update!(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc, Rhoe::AbstractVector{Float64})
    ctx = context(Ham)
    N = length(Rhoe)
    rho, vh, vxc, vtot = grid_buffers(ctx, N)
    h2d!(ctx, rho, Vector{Float64}(Rhoe))
    xc = Ham.xcfunc == "PBE" ? Cint(1) : Cint(0)
    E = zeros(Float64, 3)
    check(ctx, ccall((:pw_update_potential_dev, libpw), Cint,
                     (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                     ctx.h, 0, xc, 1, rho, vh, vxc, vtot, E))
    d2h!(ctx, Ham.potentials.Hartree, vh)
    d2h!(ctx, view(Ham.potentials.XC, :, 1), vxc)
    d2h!(ctx, view(Ham.potentials.Total, :, 1), vtot)
    return
end
update!(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc, Rhoe::Matrix{Float64}) =
    update!(Ham, psiks, view(Rhoe, :, 1))

"""Upload Ham.potentials.Total (call after update!)."""
function sync_potential!(Ham::HamHip)
    ctx = CONTEXTS[Ham]
    V = Ham.potentials.Total
    for ispin in 1:size(V, 2)
        check(ctx, ccall((:pw_set_vloc, libpw), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                         ctx.h, ispin - 1, view(V, :, ispin)))
    end
end

function apply!(Ham::HamHip, psi::AbstractMatrix{ComplexF64},
                Hpsi::AbstractMatrix{ComplexF64}, flags::UInt32)
    ctx = context(Ham)
    flags |= ctx.mixed ? PW_MIXED : UInt32(0)
    check(ctx, ccall((:pw_op_H, libpw), Cint,
                     (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Int64, Cuint),
                     ctx.h, Ham.ik - 1, Ham.ispin - 1, size(psi, 2), psi, stride(psi, 2),
                     Hpsi, stride(Hpsi, 2), flags))
    return Hpsi
end

# op_H.jl:4-27 and :53-63 (EXX is skipped when Ham.exx === nothing, see op_H.jl:61)
op_H(Ham::HamHip, psi::Matrix{ComplexF64}) = apply!(Ham, psi, similar(psi), PW_TERM_ALL)
op_H!(Ham::HamHip, psi::Matrix{ComplexF64}, Hpsi::Matrix{ComplexF64}) =
    (apply!(Ham, psi, Hpsi, PW_TERM_ALL); nothing)
op_K(Ham::HamHip, psi::Matrix{ComplexF64}) = apply!(Ham, psi, similar(psi), PW_TERM_K)
op_K!(Ham::HamHip, psi::Matrix{ComplexF64}, Hpsi::Matrix{ComplexF64}) =
    (apply!(Ham, psi, Hpsi, PW_TERM_K | PW_ACCUMULATE); nothing)
op_V_loc(Ham::HamHip, psi::Matrix{ComplexF64}) = apply!(Ham, psi, similar(psi), PW_TERM_VLOC)
op_V_loc!(Ham::HamHip, psi::Matrix{ComplexF64}, Hpsi::Matrix{ComplexF64}) =
    (apply!(Ham, psi, Hpsi, PW_TERM_VLOC | PW_ACCUMULATE); nothing)
op_V_Ps_nloc(Ham::HamHip, psi::Matrix{ComplexF64}) = apply!(Ham, psi, similar(psi), PW_TERM_NLOC)
op_V_Ps_nloc!(Ham::HamHip, psi::Matrix{ComplexF64}, Hpsi::Matrix{ComplexF64}) =
    (apply!(Ham, psi, Hpsi, PW_TERM_NLOC | PW_ACCUMULATE); nothing)

# calc_rhoe.jl:9-135 (norm-conserving, Nspin = 1), symmetrised on the device
function calc_rhoe!(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc, Rhoe::Matrix{Float64})
    ctx = context(Ham)
    pw = Ham.pw
    wk = pw.gvecw.kpoints.wk
    Focc = Ham.electrons.Focc
    check(ctx, ccall((:pw_rhoe_zero, libpw), Cint, (Ptr{Cvoid},), ctx.h))
    for ik in 1:pw.gvecw.kpoints.Nkpt
        psi = psiks[ik]
        w = wk[ik] .* Focc[:, ik]
        check(ctx, ccall((:pw_rhoe_accum, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Ptr{Float64}, Cuint),
                         ctx.h, ik - 1, size(psi, 2), psi, size(psi, 1), w,
                         ctx.mixed ? PW_MIXED : UInt32(0)))
    end
    if Ham.sym_info.Nsyms > 1
        rho = grid_buffers(ctx, size(Rhoe, 1))[1]
        check(ctx, ccall((:pw_rhoe_get_dev, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, rho))
        check(ctx, ccall((:pw_symmetrize_rhoe_dev, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, rho))
        d2h!(ctx, view(Rhoe, :, 1), rho)
    else
        check(ctx, ccall((:pw_rhoe_get, libpw), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.h, view(Rhoe, :, 1)))
    end
    return
end

# ---- forces and the adjoint (calc_forces_Ps_loc.jl:1-71, calc_forces_Ps_nloc.jl:291-411) ----

"""Upload the atoms, projector-to-atom map and the rho-sphere form factors V_sp(|G|) the
force kernels read (call again after atoms move, with the rebuilt betaNL)."""
function sync_atoms!(Ham::Hamiltonian{PsPot_GTH})
    ctx = context(Ham)
    atoms, pw = Ham.atoms, Ham.pw
    nb = Ham.pspotNL.NbetaNL
    beta_atom = zeros(Cint, nb)
    for ia in 1:atoms.Natoms
        psp = Ham.pspots[atoms.atm2species[ia]]
        for l in 0:psp.lmax, m in -l:l, i in 1:psp.Nproj_l[l+1]
            beta_atom[Ham.pspotNL.prj2beta[i, ia, l+1, m+psp.lmax+1]] = ia - 1
        end
    end
    species = Cint.(atoms.atm2species .- 1)
    check(ctx, ccall((:pw_set_atoms, libpw), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Cint}, Cint, Ptr{Cint}),
                     ctx.h, atoms.Natoms, atoms.positions, species, nb, beta_atom))
    Ng = pw.gvec.Ng
    vg = zeros(Float64, Ng, atoms.Nspecies)
    for isp in 1:atoms.Nspecies
        Vgl = zeros(Float64, length(pw.gvec.G2_shells))
        PWDFT.eval_Vloc_G!(Ham.pspots[isp], pw.gvec.G2_shells, Vgl)
        vg[:, isp] = Vgl[pw.gvec.idx_g2shells]
    end
    check(ctx, ccall((:pw_set_rho_sphere, libpw), Cint,
                     (Ptr{Cvoid}, Cint, Ptr{Int64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}),
                     ctx.h, Ng, pw.gvec.idx_g2r, 1, pw.gvec.G, atoms.Nspecies, vg))
    return nothing
end

"""calc_forces_Ps_loc(Ham, Rhoe) on the device (rho = sum over spin of Rhoe)."""
function calc_forces_Ps_loc(Ham::Hamiltonian{PsPot_GTH}, Rhoe::Matrix{Float64})
    ctx = context(Ham)
    F = zeros(Float64, 3, Ham.atoms.Natoms)
    rho = vec(sum(Rhoe, dims=2))
    check(ctx, ccall((:pw_calc_forces_Ps_loc, libpw), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, rho, F))
    return F
end

"""calc_forces_Ps_nloc(Ham, psiks) on the device (GTH), symmetrised as calc_forces.jl:11."""
function calc_forces_Ps_nloc(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc)
    ctx = context(Ham)
    F = zeros(Float64, 3, Ham.atoms.Natoms)
    wk, Focc = Ham.pw.gvecw.kpoints.wk, Ham.electrons.Focc
    for ik in 1:Ham.pw.gvecw.kpoints.Nkpt
        psi = psiks[ik]
        w = wk[ik] .* Focc[:, ik]
        check(ctx, ccall((:pw_calc_forces_Ps_nloc, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.h, ik - 1, size(psi, 2), psi, size(psi, 1), w, F, C_NULL))
    end
    PWDFT.symmetrize_vector!(Ham.pw.LatVecs, Ham.sym_info, F)
    return F
end

"""E_Ps_loc at fixed Rhoe as a function of the form factors Vgl[shell, species] and the
positions, with its ChainRulesCore rrule backed by pw_E_Ps_loc_grad (install with
`using ChainRulesCore` before including this file)."""
function E_Ps_loc(Ham::Hamiltonian{PsPot_GTH}, Rhoe::Matrix{Float64})
    ctx = context(Ham)
    E = Ref(0.0)
    check(ctx, ccall((:pw_E_Ps_loc_grad, libpw), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, vec(sum(Rhoe, dims=2)), E, C_NULL, C_NULL))
    return E[]
end

"""E_Ps_nloc at fixed psiks and occupations (calc_E_Ps_nloc for GTH) as a function of the
atomic positions through the projectors, with its ChainRulesCore rrule backed by the same
adjoint kernels as calc_forces_Ps_nloc (dE/dtau = -F_nloc, calc_forces_Ps_nloc.jl:291-411)."""
function E_Ps_nloc(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc)
    E, _ = _nloc_energy_forces(Ham, psiks)
    return E
end

function _nloc_energy_forces(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc)
    ctx = context(Ham)
    F = zeros(Float64, 3, Ham.atoms.Natoms)
    E = Ref(0.0)
    wk, Focc = Ham.pw.gvecw.kpoints.wk, Ham.electrons.Focc
    for ik in 1:Ham.pw.gvecw.kpoints.Nkpt
        psi = psiks[ik]
        w = wk[ik] .* Focc[:, ik]
        check(ctx, ccall((:pw_calc_forces_Ps_nloc, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                         ctx.h, ik - 1, size(psi, 2), psi, size(psi, 1), w, F, E))
    end
    return E[], F
end

if isdefined(Main, :ChainRulesCore)
    import ChainRulesCore
    function ChainRulesCore.rrule(::typeof(E_Ps_nloc), Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc)
        E, F = _nloc_energy_forces(Ham, psiks)
        pullback(Ebar) = (ChainRulesCore.NoTangent(),
                          ChainRulesCore.Tangent{typeof(Ham)}(; atoms = (; positions = -Ebar .* F)),
                          ChainRulesCore.NoTangent())
        return E, pullback
    end
    function ChainRulesCore.rrule(::typeof(E_Ps_loc), Ham::Hamiltonian{PsPot_GTH}, Rhoe::Matrix{Float64})
        ctx = context(Ham)
        pw = Ham.pw
        E = Ref(0.0)
        dvg = zeros(Float64, pw.gvec.Ng, Ham.atoms.Nspecies)
        dtau = zeros(Float64, 3, Ham.atoms.Natoms)
        check(ctx, ccall((:pw_E_Ps_loc_grad, libpw), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.h, vec(sum(Rhoe, dims=2)), E, dvg, dtau))
        # cotangent of the shell form factors: sum over the G vectors of each shell
        dshell = zeros(Float64, length(pw.gvec.G2_shells), Ham.atoms.Nspecies)
        for ig in 1:pw.gvec.Ng
            dshell[pw.gvec.idx_g2shells[ig], :] .+= dvg[ig, :]
        end
        pullback(Ebar) = (ChainRulesCore.NoTangent(),
                          ChainRulesCore.Tangent{typeof(Ham)}(; atoms = (; positions = Ebar .* dtau),
                                                                Vgl = Ebar .* dshell),
                          ChainRulesCore.NoTangent())
        return E[], pullback
    end
end

# ---- diag_LOBPCG! on the device (diag_LOBPCG.jl:1-28 over k, :51-225 per k) ----------

# Device mirror of a BlochWavefunc, keyed by the identity of the psiks object: one device
# block per k-point (Ngw[ik] x Nstates), reused across SCF iterations.
mutable struct DeviceWavefunc
    ctx::HipCtx
    blocks::Vector{Ptr{Cvoid}}
    bytes::Vector{Int}
end

const MIRRORS = IdDict{Any,DeviceWavefunc}()

function mirror(Ham::HamHip, psiks::BlochWavefunc)
    ctx = context(Ham)
    m = get(MIRRORS, psiks, nothing)
    if m === nothing || m.ctx !== ctx
        m = DeviceWavefunc(ctx, Ptr{Cvoid}[], Int[])
        MIRRORS[psiks] = m
    end
    while length(m.blocks) < length(psiks)
        push!(m.blocks, C_NULL)
        push!(m.bytes, 0)
    end
    for (i, psi) in enumerate(psiks)
        nb = 16 * length(psi)
        if m.bytes[i] < nb
            m.blocks[i] != C_NULL && check(ctx, ccall((:pw_free, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, m.blocks[i]))
            r = Ref{Ptr{Cvoid}}(C_NULL)
            check(ctx, ccall((:pw_malloc, libpw), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), ctx.h, nb, r))
            m.blocks[i] = r[]
            m.bytes[i] = nb
        end
    end
    return m
end

"""diag_LOBPCG!(Ham, psi) for the current Ham.ik / Ham.ispin, on the device: the block is
uploaded to its device mirror, iterated there and copied back (psi is overwritten with the
Ritz vectors, as the reference does); returns the Ritz values."""
function diag_LOBPCG!(Ham::Hamiltonian{PsPot_GTH}, psi::Matrix{ComplexF64}; tol=1e-5, NiterMax=100,
                      verbose=false, verbose_last=false, Nstates_conv=0, dpsi::Ptr{Cvoid}=C_NULL)
    ctx = context(Ham)
    Ngw, Nst = size(psi)
    own = dpsi == C_NULL
    if own
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ctx, ccall((:pw_malloc, libpw), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), ctx.h, 16 * Ngw * Nst, r))
        dpsi = r[]
    end
    evals = zeros(Float64, Nst)
    stats = zeros(Cint, 3)
    try
        check(ctx, ccall((:pw_memcpy_h2d, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Csize_t),
                         ctx.h, dpsi, psi, 16 * Ngw * Nst))
        check(ctx, ccall((:pw_diag_lobpcg_dev, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Int64, Cdouble, Cint, Cint, Cuint, Ptr{Float64},
                          Ptr{Cint}),
                         ctx.h, Ham.ik - 1, Ham.ispin - 1, Nst, dpsi, Ngw, tol, NiterMax, Nstates_conv,
                         ctx.mixed ? PW_MIXED : UInt32(0), evals, stats))
        check(ctx, ccall((:pw_memcpy_d2h, libpw), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Cvoid}, Csize_t),
                         ctx.h, psi, dpsi, 16 * Ngw * Nst))
    finally
        own && ccall((:pw_free, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, dpsi)
    end
    if verbose || verbose_last
        println("diag_LOBPCG (device): ", stats[1], " iterations, ", stats[2], " op_H columns")
    end
    return evals
end

"""diag_LOBPCG!(Ham, psiks) over the k-points and spins (diag_LOBPCG.jl:1-28): the blocks of
one spin are uploaded to their persistent device mirrors and iterated together by
pw_diag_lobpcg_batch_dev (each k-point as diag_LOBPCG! would on its own; every operation
one batched launch over the k-points still iterating, replayed from captured hipGraphs),
then copied back; returns the Ritz values (Nstates x Nkpt*Nspin)."""
function diag_LOBPCG!(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc; tol=1e-5, NiterMax=100,
                      verbose=false, verbose_last=false, Nstates_conv=0)
    ctx = context(Ham)
    pw = Ham.pw
    Nkpt = pw.gvecw.kpoints.Nkpt
    Nspin = Ham.electrons.Nspin_wf
    Nst = Ham.electrons.Nstates
    evals = zeros(Float64, Nst, Nkpt * Nspin)
    m = mirror(Ham, psiks)
    iks = Cint.(0:Nkpt-1)
    for ispin in 1:Nspin
        sel = (1:Nkpt) .+ (ispin - 1) * Nkpt
        X = Ptr{Cvoid}[m.blocks[i] for i in sel]
        ld = Int64[size(psiks[i], 1) for i in sel]
        for i in sel
            check(ctx, ccall((:pw_memcpy_h2d, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Csize_t),
                             ctx.h, m.blocks[i], psiks[i], 16 * length(psiks[i])))
        end
        ev = zeros(Float64, Nst, Nkpt)
        stats = zeros(Cint, 3, Nkpt)
        check(ctx, ccall((:pw_diag_lobpcg_batch_dev, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Cint, Ptr{Ptr{Cvoid}}, Ptr{Int64}, Cdouble, Cint, Cint,
                          Cuint, Ptr{Float64}, Ptr{Cint}),
                         ctx.h, Nkpt, iks, ispin - 1, Nst, X, ld, tol, NiterMax, Nstates_conv,
                         ctx.mixed ? PW_MIXED : UInt32(0), ev, stats))
        for i in sel
            check(ctx, ccall((:pw_memcpy_d2h, libpw), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Cvoid}, Csize_t),
                             ctx.h, psiks[i], m.blocks[i], 16 * length(psiks[i])))
        end
        evals[:, sel] = ev
        if verbose || verbose_last
            println("diag_LOBPCG (device, batched): spin ", ispin, ", iterations per k-point ",
                    extrema(stats[1, :]), ", op_H columns ", sum(stats[2, :]))
        end
    end
    return evals
end

"""diag_davidson!(Ham, psi) for the current Ham.ik / Ham.ispin, on the device
(diag_davidson.jl:50-191 through pw_diag_davidson_dev): psi is overwritten with the Ritz
vectors; returns the Ritz values."""
function diag_davidson!(Ham::Hamiltonian{PsPot_GTH}, psi::Matrix{ComplexF64}; tol=1e-5, NiterMax=100,
                        verbose=false, verbose_last=false, Nstates_conv=0, dpsi::Ptr{Cvoid}=C_NULL)
    ctx = context(Ham)
    Ngw, Nst = size(psi)
    own = dpsi == C_NULL
    if own
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ctx, ccall((:pw_malloc, libpw), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), ctx.h, 16 * Ngw * Nst, r))
        dpsi = r[]
    end
    evals = zeros(Float64, Nst)
    stats = zeros(Cint, 3)
    try
        check(ctx, ccall((:pw_memcpy_h2d, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Csize_t),
                         ctx.h, dpsi, psi, 16 * Ngw * Nst))
        check(ctx, ccall((:pw_diag_davidson_dev, libpw), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Int64, Cdouble, Cint, Cint, Cuint, Ptr{Float64},
                          Ptr{Cint}),
                         ctx.h, Ham.ik - 1, Ham.ispin - 1, Nst, dpsi, Ngw, tol, NiterMax, Nstates_conv,
                         ctx.mixed ? PW_MIXED : UInt32(0), evals, stats))
        check(ctx, ccall((:pw_memcpy_d2h, libpw), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Cvoid}, Csize_t),
                         ctx.h, psi, dpsi, 16 * Ngw * Nst))
    finally
        own && ccall((:pw_free, libpw), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, dpsi)
    end
    if verbose || verbose_last
        println("diag_davidson (device): ", stats[1], " iterations, ", stats[2], " op_H columns")
    end
    return evals
end

"""diag_davidson!(Ham, psiks) over the k-points and spins (diag_davidson.jl:12-38): each block
iterated on the device in its persistent device mirror; returns the Ritz values
(Nstates x Nkpt*Nspin)."""
function diag_davidson!(Ham::Hamiltonian{PsPot_GTH}, psiks::BlochWavefunc; tol=1e-5, NiterMax=100,
                        verbose=false, verbose_last=false, Nstates_conv=0)
    Nkpt = Ham.pw.gvecw.kpoints.Nkpt
    Nspin = Ham.electrons.Nspin_wf
    evals = zeros(Float64, Ham.electrons.Nstates, Nkpt * Nspin)
    m = mirror(Ham, psiks)
    for ispin in 1:Nspin, ik in 1:Nkpt
        Ham.ik = ik
        Ham.ispin = ispin
        i = ik + (ispin - 1) * Nkpt
        evals[:, i] = diag_davidson!(Ham, psiks[i]; tol=tol, NiterMax=NiterMax, verbose=verbose,
                                     verbose_last=verbose_last, Nstates_conv=Nstates_conv, dpsi=m.blocks[i])
    end
    return evals
end

"""The reference's KS_solve_SCF! on the device path (op_H, calc_rhoe!, update! above),
optionally with the mixed-precision Hamiltonian."""
function KS_solve_SCF_hip!(Ham::Hamiltonian{PsPot_GTH}; mixed::Bool=false, kwargs...)
    context(Ham).mixed = mixed
    PWDFT.KS_solve_SCF!(Ham; kwargs...)
end

end # module
