"""Density mixing of KS_solve_SCF! (KS_solve_SCF.jl:145-160 set-up, :240-281 dispatch) on
device vectors: simple, linear_adaptive (mix_adaptive.jl), broyden (mix_broyden.jl),
pulay (mix_pulay.jl), rpulay (mix_rpulay.jl), ppulay (mix_ppulay.jl, period 3 as the
reference fixes it) and anderson (mix_anderson.jl).

The N-length work -- residuals, history columns, the dot products that fill the small
history matrices and the final linear combinations -- runs in libpwdft_hip
(pw_vec_dots_dev, pw_vec_lincomb_dev, pw_mix_adaptive_dev, pw_vec_clamp_dev); only the
mixdim x mixdim matrices are solved on the host.  History columns that the reference shifts
are kept as a list of device vectors whose order is rotated instead of copying data.
anderson's least-squares coefficients pinv(df) dvout are formed from the Gram matrix
df' df (eigenvalues below 1e-14 of the largest dropped), which equals the reference's
pinv(df) whenever df has full column rank.
"""
import ctypes
import math

import numpy as np

from ._lib import DeviceVector, ptr

MIX_METHODS = ("simple", "linear_adaptive", "broyden", "pulay", "rpulay", "ppulay", "anderson")


def _arr(vs):
    return (ctypes.c_void_p * len(vs))(*[v.p.value for v in vs])


class DensityMixer:
    """State of one mix_method for densities of n points on the context of Ham."""

    def __init__(self, Ham, method="simple", betamix=0.2, mixdim=5, clamp=True):
        if method not in MIX_METHODS:
            raise ValueError(f"Unknown mix_method = {method}")
        if method == "broyden" and mixdim > 8:
            raise ValueError("mix_broyden: mixdim must not exceed 8 (maxter of mix_broyden.jl)")
        self.ctx, self.n = Ham.ctx, Ham.pw.Npoints
        self.method, self.beta, self.mixdim = method, float(betamix), int(mixdim)
        self.clamp = clamp     # densities are clamped at eps() after mixing, potentials are not
        new = lambda: self._zeros(DeviceVector(self.ctx, self.n))
        self.t = [DeviceVector(self.ctx, self.n) for _ in range(3)]
        if method in ("anderson", "broyden"):
            self.df = [new() for _ in range(mixdim)]
            self.dv = [new() for _ in range(mixdim)]
        elif method == "linear_adaptive":
            self.betav = DeviceVector.from_host(self.ctx, np.full(self.n, self.beta))
            self.f = new()
        elif method in ("pulay", "rpulay", "ppulay"):
            self.X = [new() for _ in range(mixdim)]
            self.F = [new() for _ in range(mixdim)]
            self.x_old, self.f_old = new(), new()

    # ---- device primitives ----------------------------------------------------------
    def _lin(self, y, ay, terms):
        """y = ay * y + sum c * v over terms [(c, v)]."""
        c = np.ascontiguousarray([float(a) for a, _ in terms], dtype=float)
        x = _arr([v for _, v in terms]) if terms else None
        self.ctx.call("pw_vec_lincomb_dev", self.n, len(terms), ptr(c) if terms else None, x, float(ay), y.p)
        return y

    def _zeros(self, y):
        return self._lin(y, 0.0, [])

    def _copy(self, y, x):
        return self._lin(y, 0.0, [(1.0, x)])

    def _dots(self, pairs):
        out = np.zeros(len(pairs))
        if pairs:
            self.ctx.call("pw_vec_dots_dev", self.n, len(pairs), _arr([a for a, _ in pairs]),
                          _arr([b for _, b in pairs]), ptr(out))
        return out

    def _gram(self, cols, rhs):
        """(cols' cols, cols' rhs) with one batch of dot products."""
        m = len(cols)
        pairs = [(cols[i], cols[j]) for i in range(m) for j in range(i, m)] + [(c, rhs) for c in cols]
        d = self._dots(pairs)
        G = np.zeros((m, m))
        q = 0
        for i in range(m):
            for j in range(i, m):
                G[i, j] = G[j, i] = d[q]
                q += 1
        return G, d[q:]

    def _clamp(self, rho):
        self.ctx.call("pw_vec_clamp_dev", self.n, float(np.finfo(float).eps), rho.p)

    # ---- mixing ---------------------------------------------------------------------
    def mix(self, rho, rho_new, it):
        """Rhoe <- mix(Rhoe, Rhoe_new) for SCF iteration `it` (1-based), then the eps()
        clamp (KS_solve_SCF.jl:277-281)."""
        m = self.method
        if m == "simple":
            self._lin(rho, 1.0 - self.beta, [(self.beta, rho_new)])
        elif m == "linear_adaptive":
            self.ctx.call("pw_mix_adaptive_dev", self.n, self.beta, 0.8, rho.p, rho_new.p, self.betav.p, self.f.p)
        elif m == "broyden":
            self._broyden(rho, rho_new, it)
        elif m == "pulay":
            self._pulay(rho, rho_new, it, periodic=None)
        elif m == "ppulay":
            self._pulay(rho, rho_new, it, periodic=3)
        elif m == "rpulay":
            self._rpulay(rho, rho_new, it)
        else:
            self._anderson(rho, rho_new, it)
        if self.clamp:
            self._clamp(rho)

    def _broyden(self, deltain, deltaout_, it):
        """mix_broyden.jl:4-70."""
        n_iter, alphamix = self.mixdim, self.beta
        wg0, wg = 0.01, np.ones(8)
        dout, save = self.t[0], self.t[1]
        self._lin(dout, 0.0, [(1.0, deltaout_), (-1.0, deltain)])
        self._copy(save, deltain)
        iter_used = min(it - 1, n_iter)
        ipos = it - 1 - math.floor((it - 2) / n_iter) * n_iter
        df, dv = self.df, self.dv
        if it > 1:
            k = ipos - 1
            self._lin(df[k], -1.0, [(1.0, dout)])
            self._lin(dv[k], -1.0, [(1.0, deltain)])
            nrm = math.sqrt(self._dots([(df[k], df[k])])[0])
            self._lin(df[k], 1.0 / nrm, [])
            self._lin(dv[k], 1.0 / nrm, [])
        if iter_used > 0:
            G, work = self._gram(df[:iter_used], dout)
            beta = wg[:iter_used, None] * wg[None, :iter_used] * G
            beta[np.diag_indices(iter_used)] = wg0 ** 2 + wg[:iter_used] ** 2
            binv = np.linalg.inv(beta)
            gam = binv.T @ (wg[:iter_used] * work)
        terms = [(alphamix, dout)]
        for i in range(iter_used):
            terms += [(-wg[i] * gam[i] * alphamix, df[i]), (-wg[i] * gam[i], dv[i])]
        self._lin(deltain, 1.0, terms)
        inext = it - math.floor((it - 1) / n_iter) * n_iter
        self._copy(df[inext - 1], dout)
        self._copy(dv[inext - 1], save)

    def _pulay_update(self, x, f, X, F):
        G, Ff = self._gram(F, f)
        c = np.linalg.inv(G) @ Ff
        terms = [(self.beta, f)]
        for j in range(len(X)):
            terms += [(-c[j], X[j]), (-self.beta * c[j], F[j])]
        self._lin(x, 1.0, terms)

    def _pulay(self, x, gx, it, periodic):
        """mix_pulay.jl:1-42, and mix_ppulay.jl:9-58 with period `periodic`."""
        M = self.mixdim
        f = self._lin(self.t[0], 0.0, [(1.0, gx), (-1.0, x)])
        if it <= M:
            Xt, Ft = self.X[it - 1], self.F[it - 1]
        else:
            self.X = self.X[1:] + self.X[:1]
            self.F = self.F[1:] + self.F[:1]
            Xt, Ft = self.X[M - 1], self.F[M - 1]
        if it == 1:
            self._copy(Xt, x)
            self._copy(Ft, f)
        else:
            self._lin(Xt, 0.0, [(1.0, x), (-1.0, self.x_old)])
            self._lin(Ft, 0.0, [(1.0, f), (-1.0, self.f_old)])
        self._copy(self.x_old, x)
        self._copy(self.f_old, f)
        if periodic is None:
            use = it > M
        else:
            use = False if it < M else it % periodic == 0
        if use:
            self._pulay_update(x, f, self.X, self.F)
        else:
            self._lin(x, 1.0, [(self.beta, f)])

    def _rpulay(self, x, gx, it):
        """mix_rpulay.jl:11-94."""
        M = self.mixdim
        f = self._lin(self.t[0], 0.0, [(1.0, gx), (-1.0, x)])
        if it == 1:
            self._copy(self.X[0], x)
            self._copy(self.F[0], f)
            self._copy(self.x_old, x)
            self._copy(self.f_old, f)
            self._lin(x, 1.0, [(self.beta, f)])
            return
        active = (it - 1) % (M + 1)
        if active == 0:
            # restart from the last history column (the remaining columns are never read
            # before being overwritten)
            self.X[0], self.X[M - 1] = self.X[M - 1], self.X[0]
            self.F[0], self.F[M - 1] = self.F[M - 1], self.F[0]
            k = 1
        else:
            a = active
            self.X = [self.X[a - 1]] + self.X[:a - 1] + self.X[a:]
            self.F = [self.F[a - 1]] + self.F[:a - 1] + self.F[a:]
            self._lin(self.X[0], 0.0, [(1.0, x), (-1.0, self.x_old)])
            self._lin(self.F[0], 0.0, [(1.0, f), (-1.0, self.f_old)])
            k = a
        G, Ff = self._gram(self.F[:k], f)
        c = np.linalg.inv(G) @ Ff
        self._copy(self.x_old, x)
        self._copy(self.f_old, f)
        terms = [(self.beta, f)]
        for j in range(k):
            terms += [(-c[j], self.X[j]), (-self.beta * c[j], self.F[j])]
        self._lin(x, 1.0, terms)

    def _anderson(self, vin, vout, it):
        """mix_anderson.jl:1-40."""
        m = self.mixdim
        dvout = self._lin(self.t[0], 0.0, [(1.0, vout), (-1.0, vin)])
        iterused = min(it - 1, m)
        ipos = it - 1 - math.floor((it - 2) / m) * m
        if it > 1:
            self._lin(self.df[ipos - 1], 1.0, [(-1.0, dvout)])
            self._lin(self.dv[ipos - 1], 1.0, [(-1.0, vin)])
        vinsave = self._copy(self.t[1], vin)
        dvoutsave = self._copy(self.t[2], dvout)
        if it > 1:
            G, b = self._gram(self.df[:iterused], dvout)
            w, U = np.linalg.eigh(G)
            keep = w > 1e-14 * max(w.max(), 0.0)
            gam = U[:, keep] @ ((U[:, keep].T @ b) / w[keep])
            self._lin(vin, 1.0, [(-gam[i], self.dv[i]) for i in range(iterused)])
            self._lin(dvout, 1.0, [(-gam[i], self.df[i]) for i in range(iterused)])
        inext = it - math.floor((it - 1) / m) * m
        self._copy(self.df[inext - 1], dvoutsave)
        self._copy(self.dv[inext - 1], vinsave)
        self._lin(vin, 1.0, [(self.beta, dvout)])
