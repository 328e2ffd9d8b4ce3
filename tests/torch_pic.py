"""A pure-torch FP64 restatement of the differentiable step, for small grids
(16^3): the reference gradient the hand-written adjoint kernels are checked
against.  Every operator is written from its definition with torch autograd
doing the differentiation:

- deposit: np += kq Nm/(h1 h2 h3) w_c(x) on the 8 stencil nodes, with the
  reference's stencil ix = int(x/h) + 1, px = ix - x/h
  (mod_chargeDeposition.f90:25-133);
- rho = -sum_s q_s np_s on physical nodes (mod_density.f90:76-107);
- Poisson: eps0 lap7(phi) = rho on plasma nodes, phi = the electrode's
  voltage on labelled nodes, solved exactly by a dense inverse (the
  multigrid's operator, mg.h, solved to round-off);
- E = -grad(phi): central differences on plasma nodes, the wall nodes'
  extrapolation blocks in W, E, S, N, B, T order (mod_electricField.f90);
- gather with the same weights, the Boris push with uniform B = (0, 0, B0)
  or a B map gathered like E and gated at the lower node
  (Src/part_expmover.f90:151-246), x += dt v;
- wall absorption frozen: a particle dies at the step the schedule says;
- loss: the time-averaged ion flux through the node plane one cell in front
  of the biased x = L wall, q Nm vx S(x) / (Ly Lz) with S the plane's
  cloud-in-cell hat over h (or through another axis's node plane, with a
  sign and an area).

Test infrastructure only."""
from __future__ import annotations

import numpy as np
import torch

EPS0 = 8.854187817e-12


class TorchPIC:
    def __init__(self, n, h, bcnd, biased_label, species, dt, ion_index=1):
        """species: list of (charge, mass, Nm); bcnd: (n3+3, n2+3, n1+3) labels."""
        self.n = tuple(int(v) for v in n)
        self.h = tuple(float(v) for v in h)
        self.dt = float(dt)
        self.species = species
        self.ion = ion_index
        self.label = biased_label
        b = torch.as_tensor(np.asarray(bcnd), dtype=torch.int64)
        self.bc = b
        self.shape = tuple(b.shape)
        n1, n2, n3 = self.n
        kk, jj, ii = torch.meshgrid(torch.arange(n3 + 3), torch.arange(n2 + 3),
                                    torch.arange(n1 + 3), indexing="ij")
        self.phys = (ii >= 1) & (ii <= n1 + 1) & (jj >= 1) & (jj <= n2 + 1) & \
                    (kk >= 1) & (kk <= n3 + 1)
        self.plasma = self.phys & (b <= 0)
        self.kq = torch.where(b != -1, 2.0, 1.0).double()
        self.bias_mask = (b == biased_label)
        # dense Poisson operator on plasma nodes
        idx = torch.nonzero(self.plasma.reshape(-1)).reshape(-1)
        self.pidx = idx
        N = idx.numel()
        pos = -torch.ones(b.numel(), dtype=torch.int64)
        pos[idx] = torch.arange(N)
        aw = EPS0 / (self.h[0] ** 2)
        an = EPS0 / (self.h[1] ** 2)
        at = EPS0 / (self.h[2] ** 2)
        ac = -(an + an + aw + aw + at + at)
        self.coef = (aw, an, at, ac)
        sx, sy, sz = 1, n1 + 3, (n1 + 3) * (n2 + 3)
        A = torch.zeros(N, N, dtype=torch.float64)
        A[torch.arange(N), torch.arange(N)] = ac
        self.nbr = []   # (offset, coefficient)
        for off, a in ((sx, aw), (-sx, aw), (sy, an), (-sy, an), (sz, at), (-sz, at)):
            nb = idx + off
            p = pos[nb]
            m = p >= 0
            A[torch.arange(N)[m], p[m]] = a
            self.nbr.append((off, a))
        self.Ainv = torch.linalg.inv(A)

    # -- grid operators ---------------------------------------------------
    def stencil(self, x):
        """Lower-node indices (1-based node ids, flattened) and the 8 weights
        w1..w8 of mod_particleMover.f90:104-142 for positions x (N, 3)."""
        h = self.h
        n1, n2, n3 = self.n
        ix = torch.floor(x[:, 0].detach() / h[0]).long() + 1
        iy = torch.floor(x[:, 1].detach() / h[1]).long() + 1
        iz = torch.floor(x[:, 2].detach() / h[2]).long() + 1
        px = ix.double() - x[:, 0] / h[0]
        py = iy.double() - x[:, 1] / h[1]
        pz = iz.double() - x[:, 2] / h[2]
        sx, sy, sz = 1, n1 + 3, (n1 + 3) * (n2 + 3)
        base = ix + sy * iy + sz * iz
        nodes = torch.stack([base, base + sx, base + sx + sy, base + sy, base + sz,
                             base + sz + sx, base + sz + sx + sy, base + sz + sy], 1)
        w = torch.stack([px * py * pz, (1 - px) * py * pz, (1 - px) * (1 - py) * pz,
                         px * (1 - py) * pz, px * py * (1 - pz), (1 - px) * py * (1 - pz),
                         (1 - px) * (1 - py) * (1 - pz), px * (1 - py) * (1 - pz)], 1)
        return nodes, w

    def deposit(self, x, alive, Nm):
        nodes, w = self.stencil(x)
        k1 = Nm / (self.h[0] * self.h[1] * self.h[2])
        w = w * alive.double()[:, None]
        out = torch.zeros(self.bc.numel(), dtype=torch.float64)
        out = out.index_add(0, nodes.reshape(-1), (k1 * w).reshape(-1))
        return out * self.kq.reshape(-1)

    def solve(self, rho, Vb):
        """phi with electrodes at their voltages (biased label at Vb, others 0)."""
        phiE = torch.where(self.bias_mask.reshape(-1), Vb, torch.zeros((), dtype=torch.float64))
        phiE = phiE * torch.ones(self.bc.numel(), dtype=torch.float64)
        phiE = torch.where(self.plasma.reshape(-1), torch.zeros_like(phiE), phiE)
        aw, an, at, ac = self.coef
        idx = self.pidx
        coupling = torch.zeros(idx.numel(), dtype=torch.float64)
        for off, a in self.nbr:
            coupling = coupling + a * phiE[idx + off]
        phiP = self.Ainv @ (rho[idx] - coupling)
        return phiE.index_put((idx,), phiP)

    def efield(self, phi):
        n1, n2, n3 = self.n
        sh = self.shape
        P = phi.reshape(sh)
        h = self.h
        E = torch.zeros(sh + (3,), dtype=torch.float64)
        inner = (slice(1, n3 + 2), slice(1, n2 + 2), slice(1, n1 + 2))

        def d(axis, sgn):
            # phi at the node shifted by sgn along axis, over the physical range
            sl = [slice(1, n3 + 2), slice(1, n2 + 2), slice(1, n1 + 2)]
            ax = 2 - axis
            s = sl[ax]
            sl[ax] = slice(s.start + sgn, s.stop + sgn)
            return P[tuple(sl)]

        cx = -(d(0, 1) - d(0, -1)) / (2.0 * h[0])
        cy = -(d(1, 1) - d(1, -1)) / (2.0 * h[1])
        cz = -(d(2, 1) - d(2, -1)) / (2.0 * h[2])
        pl = (self.bc[inner] <= 0)
        Ei = torch.stack([cx, cy, cz], -1) * pl[..., None].double()
        E = torch.zeros(sh + (3,), dtype=torch.float64)
        E[inner] = Ei
        # wall nodes: blocks in W, E, S, N, B, T order, the last applicable
        # wins; the reference's loop visits nodes in index order, so the
        # second-neighbour read E(c + 2o) sees the final value of an electrode
        # node visited before (o < 0) and the pass-1 value otherwise: iterate
        # to the fixed point (one more link of such chains settles per pass)
        b = self.bc
        wall = self.phys & (b >= 1)
        flat = E.reshape(-1, 3)
        cen = torch.zeros(sh + (3,), dtype=torch.float64)
        cen[inner] = torch.stack([cx, cy, cz], -1)
        sx, sy, sz = 1, n1 + 3, (n1 + 3) * (n2 + 3)
        N = b.numel()
        allidx = torch.arange(N)
        bflat = b.reshape(-1)
        wflat = wall.reshape(-1)
        cflat = cen.reshape(-1, 3)
        cur = flat
        for _ in range(64):
            res = flat
            for off, comp in ((sx, 0), (-sx, 0), (sy, 1), (-sy, 1), (sz, 2), (-sz, 2)):
                nb = allidx + off
                valid = wflat & (nb >= 0) & (nb < N)
                nbc = torch.clamp(nb, 0, N - 1)
                cond = valid & (bflat[nbc] <= 0)
                nb2 = allidx + 2 * off
                in2 = (nb2 >= 0) & (nb2 < N)
                nb2c = torch.clamp(nb2, 0, N - 1)
                seen = in2 & (off < 0) & wflat[nb2c]
                e2 = torch.where(seen, cur[nb2c, comp], flat[nb2c, comp])
                e2 = torch.where(in2, e2, 0.0)
                normal = 2.0 * flat[nbc, comp] - e2
                new = cflat.clone()
                new[:, comp] = normal
                res = torch.where(cond[:, None], new, res)
            done = bool(torch.equal(res.detach(), cur.detach()))
            cur = res
            if done:
                break
        return cur.reshape(sh + (3,))

    def gather(self, E, x):
        nodes, w = self.stencil(x)
        Ef = E.reshape(-1, 3)
        return (w[:, :, None] * Ef[nodes]).sum(1)

    def gather_b(self, B4, x):
        """A B map (node_shape + (4,): Bx, By, Bz, |B|) at x, gated at the
        stencil's lower node, |B| > 1e-5 T (Src/part_expmover.f90:151-207):
        the gathered (N, 3) field and the gate."""
        nodes, w = self.stencil(x)
        Bf = B4.reshape(-1, 4)
        use = Bf[nodes[:, 0], 3] > 1.0e-5
        return (w[:, :, None] * Bf[nodes][:, :, :3]).sum(1), use

    def boris_map(self, v, Ep, Bp, use, q, m):
        """Boris with a per-particle B (N, 3); leapfrog where use is False."""
        k1 = self.dt * q / (2.0 * m)
        Bx, By, Bz = Bp[:, 0], Bp[:, 1], Bp[:, 2]
        B2 = Bx * Bx + By * By + Bz * Bz
        # |B| without sqrt's infinite slope at 0 (where B vanishes the
        # particle is gated off and its Boris branch is not taken)
        pos = B2 > 0.0
        Bt = torch.where(pos, torch.sqrt(torch.where(pos, B2, torch.ones_like(B2))), 0.0)
        k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)))
        v1 = v + k1 * Ep
        v3x = v1[:, 0] + k1 * (v1[:, 1] * Bz - v1[:, 2] * By)
        v3y = v1[:, 1] + k1 * (v1[:, 2] * Bx - v1[:, 0] * Bz)
        v3z = v1[:, 2] + k1 * (v1[:, 0] * By - v1[:, 1] * Bx)
        v2x = v1[:, 0] + k2 * (v3y * Bz - v3z * By)
        v2y = v1[:, 1] + k2 * (v3z * Bx - v3x * Bz)
        v2z = v1[:, 2] + k2 * (v3x * By - v3y * Bx)
        vb = torch.stack([v2x, v2y, v2z], 1) + k1 * Ep
        return torch.where(use[:, None], vb, v + 2.0 * k1 * Ep)

    def boris(self, v, Ep, B0, q, m):
        k1 = self.dt * q / (2.0 * m)
        Bx = torch.zeros((), dtype=torch.float64)
        By = torch.zeros((), dtype=torch.float64)
        Bz = B0
        Bt = torch.sqrt(Bx * Bx + By * By + Bz * Bz)
        k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)))
        v1 = v + k1 * Ep
        v3x = v1[:, 0] + k1 * (v1[:, 1] * Bz - v1[:, 2] * By)
        v3y = v1[:, 1] + k1 * (v1[:, 2] * Bx - v1[:, 0] * Bz)
        v3z = v1[:, 2] + k1 * (v1[:, 0] * By - v1[:, 1] * Bx)
        v2x = v1[:, 0] + k2 * (v3y * Bz - v3z * By)
        v2y = v1[:, 1] + k2 * (v3z * Bx - v3x * Bz)
        v2z = v1[:, 2] + k2 * (v3x * By - v3y * Bx)
        return torch.stack([v2x, v2y, v2z], 1) + k1 * Ep

    def plane_flux(self, x, v, alive, q, Nm, flux=None):
        """flux: (axis, 1-based node plane, sign, area or None); None: the
        node plane x = (n1 - 1) h1 over the x = L face."""
        axis, plane, sign, A = flux or (0, self.n[0], 1.0, None)
        h = self.h
        if A is None:
            A = float(np.prod([self.n[d] * h[d] for d in range(3) if d != axis]))
        xp = (plane - 1) * h[axis]
        S = torch.clamp(1.0 - torch.abs(x[:, axis] - xp) / h[axis], min=0.0) / h[axis]
        return (sign * q * Nm / A) * (v[:, axis] * S * alive.double()).sum()

    # -- the run ----------------------------------------------------------
    def run(self, xs, vs, Vb, B0, steps, schedule=None, B4=None, magnetized=None, flux=None):
        """xs, vs: per species (N, 3) FP64.  schedule: per species, the step
        (1-based) at which each particle dies (0: never); None records it
        from the wall test (outside [0, L) per axis).  B4: a B map on the
        nodes (replaces B0) acting on the species whose magnetized entry is
        true (default all); flux: the loss plane of plane_flux.  Returns the
        loss and the schedule."""
        L = torch.tensor([self.n[d] * self.h[d] for d in range(3)], dtype=torch.float64)
        xs = [x.clone() for x in xs]
        vs = [v.clone() for v in vs]
        record = schedule is None
        if record:
            schedule = [torch.zeros(x.shape[0], dtype=torch.int64) for x in xs]
        alive = [torch.ones(x.shape[0], dtype=torch.bool) for x in xs]
        loss = torch.zeros((), dtype=torch.float64)
        for k in range(1, steps + 1):
            rho = torch.zeros(self.bc.numel(), dtype=torch.float64)
            for s, (q, m, Nm) in enumerate(self.species):
                xsafe = torch.where(alive[s][:, None], xs[s], L / 2)
                rho = rho - q * self.deposit(xsafe, alive[s], Nm)
            rho = torch.where(self.phys.reshape(-1), rho, torch.zeros_like(rho))
            phi = self.solve(rho, Vb)
            E = self.efield(phi)
            for s, (q, m, Nm) in enumerate(self.species):
                a = alive[s]
                xa = torch.where(a[:, None], xs[s], L / 2)
                Ep = self.gather(E, xa)
                if B4 is not None:
                    Bp, use = self.gather_b(B4, xa)
                    if magnetized is not None and not magnetized[s]:
                        use = torch.zeros_like(use)
                    vn = self.boris_map(vs[s], Ep, Bp, use, q, m)
                else:
                    vn = self.boris(vs[s], Ep, B0, q, m)
                xn = xs[s] + self.dt * vn
                vs[s] = torch.where(a[:, None], vn, vs[s])
                xs[s] = torch.where(a[:, None], xn, xs[s])
                if record:
                    out = ((xs[s] < 0) | (xs[s] >= L)).any(1) & a
                    schedule[s] = torch.where(out, torch.full_like(schedule[s], k), schedule[s])
                dead_now = schedule[s] == k
                if s == self.ion:
                    loss = loss + self.plane_flux(xs[s], vs[s], a & ~dead_now, q, Nm, flux)
                alive[s] = a & ~dead_now
        return loss / steps, schedule
