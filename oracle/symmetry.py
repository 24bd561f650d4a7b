"""Oracle: symmetrize_rhoe! (RhoeSymmetrizer.jl:111-236), test infrastructure only.

The operations come in the reference's convention (SymmetryInfo.jl): inv_s[s] is the
integer matrix that maps the crystal coordinates g0 = LatVecs' G / (2 pi) of a G-vector to
those of its image, ft[s] the fractional translation (crystal coordinates), and an
operation is non-symmorphic when ft[s] != 0.  For each G not yet visited (the reference
walks the G shells; the orbit of a G never leaves its shell, so the visiting order does
not change the result):
    rhosum = 1/Nsyms sum_s rho(S_s G) exp(-i (S_s G) . fts_s),   fts_s = LatVecs ft_s,
    rho(S_s G) = rhosum exp(+i (S_s G) . fts_s) for every s,
on the full-grid R_to_G of rho, whose entries outside the rho sphere are left as they are;
then Rhoe = real(G_to_R(...)).  The image of G is found by its integer coordinates (the
reference searches the shell with a 1e-5 tolerance on the same coordinates).
"""
import numpy as np


def symmetrize_rhoe(pw, inv_s, ft, rho):
    """RhoeSymmetrizer.jl:111-236 for one spin; pw needs Ns, LatVecs, G (3 x Ng), idx_g2r."""
    Ns = tuple(pw.Ns)
    rhoG = np.fft.fftn(np.asarray(rho, dtype=float).reshape(Ns, order="F")).reshape(-1, order="F")
    g0 = np.rint(np.asarray(pw.LatVecs).T @ np.asarray(pw.G) / (2 * np.pi)).astype(np.int64)  # 3 x Ng
    where = {tuple(g0[:, ig]): ig for ig in range(g0.shape[1])}
    inv_s = [np.asarray(S, dtype=np.int64) for S in inv_s]
    fts = [np.asarray(pw.LatVecs) @ np.asarray(f, dtype=float) for f in ft]
    nonsym = [bool(np.any(np.asarray(f) != 0.0)) for f in ft]
    nsym = len(inv_s)
    done = np.zeros(g0.shape[1], dtype=bool)
    G = np.asarray(pw.G)
    for ig in range(g0.shape[1]):
        if done[ig]:
            continue
        irot = []
        rhosum = 0.0j
        for s in range(nsym):
            isg = where.get(tuple(inv_s[s] @ g0[:, ig]))
            if isg is None:
                raise RuntimeError("Error in determining irot")
            irot.append(isg)
            ip = pw.idx_g2r[isg]
            if nonsym[s]:
                arg = G[:, isg] @ fts[s]
                rhosum += rhoG[ip] * (np.cos(arg) - 1j * np.sin(arg))
            else:
                rhosum += rhoG[ip]
        rhosum /= nsym
        for s in range(nsym):
            isg = irot[s]
            ip = pw.idx_g2r[isg]
            if nonsym[s]:
                arg = G[:, isg] @ fts[s]
                rhoG[ip] = rhosum * (np.cos(arg) + 1j * np.sin(arg))
            else:
                rhoG[ip] = rhosum
            done[isg] = True
    return np.real(np.fft.ifftn(rhoG.reshape(Ns, order="F")).reshape(-1, order="F"))
