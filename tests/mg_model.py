"""numpy model of the multigrid V-cycle of ELSIM_MODE_MIXED_MG (test
infrastructure: the checker of the GPU V-cycle and of its convergence).

Same construction as the HIP code: level 1 from the fine operator's face
conductances scale * 0.5 (kap_a + kap_b) summed over 2 x 2 x 2 blocks
(ceil(n/2) per axis), masses = psE (CnE) and the conductances of faces with a
zero ghost rule (Dirichlet data); further levels by the same sums; the
coarsest level (<= 64 cells) solved exactly; damped Jacobi (omega) before and
after the coarse correction (weight oc).  The fine level uses the CG's own
operator with sign sigma (A = sigma A_hat: -1 for the potentials, +1 for CnE)
and the Jacobi inverse rounded to FP32.  FP64 throughout otherwise.
"""
import numpy as np

OP_CNE, OP_PHP, OP_PHE, OP_POT_DD = 0, 1, 2, 3
RULES = {OP_CNE: (False, False, True), OP_PHP: (False, True, False), OP_PHE: (True, False, False),
         OP_POT_DD: (True, True, False)}  # lo_zero, hi_zero, diag


def fine_parts(op, kap, pse, scale):
    """Face conductances (+1 neighbour along d1, d2, d3; zero past the last
    cell), mass and diagonal of A_hat on the interior.  kap, pse ghosted."""
    lo, hi, diag = RULES[op]
    K = np.asarray(kap, np.float64)
    kc = K[1:-1, 1:-1, 1:-1]
    n = kc.shape
    cs = []
    for ax in range(3):
        sl = [slice(1, -1)] * 3
        sl[ax] = slice(2, None)
        f = scale * ((K[tuple(sl)] + kc) * 0.5)
        last = [slice(None)] * 3
        last[ax] = -1
        f[tuple(last)] = 0.0
        cs.append(f)
    ms = np.zeros(n)
    if diag:
        ms += np.asarray(pse, np.float64)[1:-1, 1:-1, 1:-1]
    if lo:
        ms[:, 0, :] += scale * ((K[1:-1, 0, 1:-1] + kc[:, 0, :]) * 0.5)
    if hi:
        ms[:, -1, :] += scale * ((K[1:-1, -1, 1:-1] + kc[:, -1, :]) * 0.5)
    return cs, ms


def diag_of(cs, ms):
    d = ms.copy()
    for ax, c in enumerate(cs):
        d += c
        sh = [slice(None)] * 3
        sh[ax] = slice(1, None)
        sm = [slice(None)] * 3
        sm[ax] = slice(0, -1)
        d[tuple(sh)] += c[tuple(sm)]
    return d


def block_sum(a):
    """Sum over 2 x 2 x 2 blocks (ceil(n/2) per axis)."""
    for ax in range(3):
        n = a.shape[ax]
        if n % 2:
            pad = [(0, 0)] * 3
            pad[ax] = (0, 1)
            a = np.pad(a, pad)
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[ax] = slice(0, None, 2)
        sl1[ax] = slice(1, None, 2)
        a = a[tuple(sl0)] + a[tuple(sl1)]
    return a


def coarsen(cs, ms):
    out = []
    for ax, c in enumerate(cs):
        # conductances across the interface of coarse cells I, I+1: fine faces at odd index 2I+1
        n = c.shape[ax]
        sel = [slice(None)] * 3
        sel[ax] = slice(1, None, 2)
        cf = c[tuple(sel)]
        other = [a for a in range(3) if a != ax]
        for a in other:
            if cf.shape[a] % 2:
                pad = [(0, 0)] * 3
                pad[a] = (0, 1)
                cf = np.pad(cf, pad)
            s0 = [slice(None)] * 3
            s1 = [slice(None)] * 3
            s0[a] = slice(0, None, 2)
            s1[a] = slice(1, None, 2)
            cf = cf[tuple(s0)] + cf[tuple(s1)]
        nc = (n + 1) // 2
        if cf.shape[ax] < nc:
            pad = [(0, 0)] * 3
            pad[ax] = (0, nc - cf.shape[ax])
            cf = np.pad(cf, pad)
        last = [slice(None)] * 3
        last[ax] = -1
        cf = cf.copy()
        cf[tuple(last)] = 0.0
        out.append(cf)
    return out, block_sum(ms)


def apply_lvl(cs, dg, x):
    y = dg * x
    for ax, c in enumerate(cs):
        sh = [slice(None)] * 3
        sm = [slice(None)] * 3
        sh[ax] = slice(1, None)
        sm[ax] = slice(0, -1)
        y[tuple(sm)] -= c[tuple(sm)] * x[tuple(sh)]
        y[tuple(sh)] -= c[tuple(sm)] * x[tuple(sm)]
    return y


def dense(cs, dg):
    n = dg.shape
    N = dg.size
    A = np.zeros((N, N))
    idx = np.arange(N).reshape(n, order="F")
    A[np.arange(N), np.arange(N)] = dg.ravel(order="F")
    for ax, c in enumerate(cs):
        sm = [slice(None)] * 3
        sh = [slice(None)] * 3
        sm[ax] = slice(0, -1)
        sh[ax] = slice(1, None)
        a, b, w = idx[tuple(sm)].ravel(), idx[tuple(sh)].ravel(), c[tuple(sm)].ravel()
        A[a, b] -= w
        A[b, a] -= w
    return A


def prolong(x, shape):
    for ax in range(3):
        x = np.repeat(x, 2, axis=ax)
    return x[:shape[0], :shape[1], :shape[2]]


class VCycle:
    def __init__(self, op, kap, pse, scale, omega=0.7, oc=1.5, coarse_max=64):
        self.op, self.omega, self.oc = op, omega, oc
        self.sigma = 1.0 if RULES[op][2] else -1.0
        self.kap = np.asarray(kap, np.float64)
        self.pse = None if pse is None else np.asarray(pse, np.float64)
        self.scale = scale
        cs, ms = fine_parts(op, kap, pse, scale)
        self.dg0 = diag_of(cs, ms)
        # the FP32 Jacobi inverse of A (= sigma / dg0) that the multigrid path keeps
        self.dinv = (self.sigma / self.dg0).astype(np.float32).astype(np.float64)
        self.levels = []
        while True:
            cs, ms = coarsen(cs, ms)
            dg = diag_of(cs, ms)
            self.levels.append((cs, dg))
            if dg.size <= coarse_max:
                self.ainv = np.linalg.inv(dense(cs, dg))
                break

    def apply_A(self, x):
        """The CG operator A (not A_hat) on an interior array, homogeneous ghost rule."""
        lo, hi, diag = RULES[self.op]
        g = np.pad(x, 1, mode="edge")
        if lo:
            g[:, 0, :] = 0.0
        if hi:
            g[:, -1, :] = 0.0
        K = self.kap
        kc = K[1:-1, 1:-1, 1:-1]
        s = np.zeros_like(x)
        for ax in range(3):
            for dlt in (1, -1):
                sl = [slice(1, -1)] * 3
                sl[ax] = slice(1 + dlt, K.shape[ax] - 1 + dlt)
                f = (K[tuple(sl)] + kc) * 0.5
                s += f * (g[tuple(sl)] - x)
        if diag:
            return self.pse[1:-1, 1:-1, 1:-1] * x - self.scale * s
        return self.scale * s

    def _cycle(self, l, b):
        cs, dg = self.levels[l]
        if l == len(self.levels) - 1:
            return (self.ainv @ b.ravel(order="F")).reshape(b.shape, order="F")
        w = self.omega
        x1 = w * b / dg
        r = b - apply_lvl(cs, dg, x1)
        xc = self._cycle(l + 1, block_sum(r))
        x2 = x1 + self.oc * prolong(xc, b.shape)
        return x2 + w * (b - apply_lvl(cs, dg, x2)) / dg

    def __call__(self, r):
        """z = M^-1 r for the CG operator A (interior arrays)."""
        w = self.omega
        x1 = w * self.dinv * r
        t = self.sigma * (r - self.apply_A(x1))
        xc = self._cycle(0, block_sum(t))
        x2 = x1 + self.oc * prolong(xc, r.shape)
        return x2 + w * self.dinv * (r - self.apply_A(x2))


def pcg(apply_A, b, M, tol=1e-8, maxit=5000):
    """Preconditioned CG from zero; returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = np.vdot(r, z)
    bn = np.linalg.norm(b)
    for it in range(1, maxit + 1):
        w = apply_A(p)
        a = rz / np.vdot(p, w)
        x += a * p
        r -= a * w
        if np.linalg.norm(r) <= tol * bn:
            return x, it
        z = M(r)
        rz2 = np.vdot(r, z)
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit
