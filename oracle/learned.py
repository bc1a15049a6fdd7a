"""CPU oracle of the learned multigrid smoother (float64 numpy).

TEST INFRASTRUCTURE ONLY: what the HIP network kernels are tested against.

The network has no reference counterpart; it is the library's own smoother.  For a
residual r on a level with operator A (CSR rows, columns ascending) it returns
a correction c:

    s   = sqrt(mean(r²))                     (scale; c is linear in s)
    x   = r / s
    P   = D⁻¹ (D - A) with D = diag(A)        (normalized neighbour weights)
    f0  = [x, P x]                           (N x 2)
    h1  = relu(f0 W1 + b1)                   (N x 32)
    h2  = relu([h1, P h1] W2 + b2)           (N x 32)   dense 64 -> 32
    h3  = relu([h2, P h2] W3 + b3)           (N x 32)   dense 64 -> 32
    y   = alpha x + h3 wo + bo               (N)
    c   = (s / D) y

Training losses over a batch:
  residual loss: mean over samples and vertices of ((r - A c) / s)², the
    residual left after the correction;
  error loss (residual -> error pairs (r, e), e the error A⁻¹r of the
    iterate): mean of (κ H (e - c) / s)² with the high-pass H = I - Pr R, R
    the restriction to the next coarser level and Pr the prolongation back:
    the part of the remaining error that the coarse grid cannot represent,
    which is what a multigrid smoother has to remove.  κ = mean(D) puts the
    error in residual units (the loss is then O(1), as the residual loss).

Gradients are pinned by central finite differences in the tests; nothing else
fixes them (parity of the network is against this restatement only).
"""
from __future__ import annotations

import numpy as np

HIDDEN = 32

# parameter layout (flat float vector), shared with the HIP kernels
SHAPES = [("W1", (2, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (2 * HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (2 * HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("wo", (HIDDEN,)), ("bo", (1,)), ("alpha", (1,))]


def n_params():
    return sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, k = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SHAPES])


def init_params(seed=0, alpha=2 / 3):
    """Deterministic initialisation: small weights, output head zero, alpha =
    2/3 so the untrained network is exactly damped Jacobi."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[0]
            p[name] = rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)
        else:
            p[name] = np.zeros(shp)
    p["alpha"] = np.array([alpha])
    return pack(p)


class Csr:
    """A rectangular sparse matrix (rowptr, col, val) with products."""

    def __init__(self, rowptr, col, val, ncols):
        self.rp, self.ci, self.vv = np.asarray(rowptr), np.asarray(col), np.asarray(val, np.float64)
        self.m, self.n = len(self.rp) - 1, ncols
        self.rows = np.repeat(np.arange(self.m), np.diff(self.rp))

    def __matmul__(self, x):
        return np.bincount(self.rows, self.vv * x[self.ci], minlength=self.m)

    def T(self, y):
        return np.bincount(self.ci, self.vv * y[self.rows], minlength=self.n)


class Level:
    """A level's operator in CSR (rowptr, col, val) with its P = D⁻¹(D - A);
    R (restriction to the next coarser level) and Pr (prolongation from it)
    as Csr for the error loss."""

    def __init__(self, rowptr, col, val, R=None, Pr=None):
        self.R, self.Pr = R, Pr
        self.rp, self.ci, self.vv = (np.asarray(rowptr), np.asarray(col), np.asarray(val, np.float64))
        n = len(self.rp) - 1
        self.n = n
        rows = np.repeat(np.arange(n), np.diff(self.rp))
        self.rows = rows
        diag = np.zeros(n)
        m = self.ci == rows
        diag[rows[m]] = self.vv[m]
        self.D = diag
        self.pval = np.where(m, 0.0, -self.vv / diag[rows])

    def A(self, x):
        return np.bincount(self.rows, self.vv * x[self.ci], minlength=self.n) if x.ndim == 1 else \
            np.stack([self.A(x[:, j]) for j in range(x.shape[1])], axis=1)

    def AT(self, x):
        return np.bincount(self.ci, self.vv * x[self.rows], minlength=self.n)

    def P(self, X):
        """Σ_w a_vw X[w] for X of shape (N,) or (N, F)."""
        if X.ndim == 1:
            return np.bincount(self.rows, self.pval * X[self.ci], minlength=self.n)
        out = np.zeros_like(X)
        np.add.at(out, self.rows, self.pval[:, None] * X[self.ci])
        return out

    def PT(self, X):
        out = np.zeros_like(X)
        np.add.at(out, self.ci, self.pval[:, None] * X[self.rows]) if X.ndim > 1 else \
            np.add.at(out, self.ci, self.pval * X[self.rows])
        return out


def relu(z):
    return np.maximum(z, 0.0)


def forward(theta, lv: Level, r):
    p = unpack(theta)
    s = np.sqrt(np.mean(r * r))
    x = r / s
    f0 = np.stack([x, lv.P(x)], axis=1)
    z1 = f0 @ p["W1"] + p["b1"]
    h1 = relu(z1)
    g1 = np.concatenate([h1, lv.P(h1)], axis=1)
    z2 = g1 @ p["W2"] + p["b2"]
    h2 = relu(z2)
    g2 = np.concatenate([h2, lv.P(h2)], axis=1)
    z3 = g2 @ p["W3"] + p["b3"]
    h3 = relu(z3)
    y = p["alpha"][0] * x + h3 @ p["wo"] + p["bo"][0]
    c = (s / lv.D) * y
    cache = dict(s=s, x=x, f0=f0, h1=h1, g1=g1, h2=h2, g2=g2, h3=h3, y=y, p=p)
    return c, cache


def loss(theta, lv: Level, R):
    """Mean over samples (rows of R) and vertices of ((r - A c)/s)²."""
    tot = 0.0
    for r in R:
        c, k = forward(theta, lv, r)
        rho = (r - lv.A(c)) / k["s"]
        tot += np.mean(rho * rho)
    return tot / len(R)


def loss_grad(theta, lv: Level, R):
    g = np.zeros_like(theta)
    tot = 0.0
    B = len(R)
    for r in R:
        c, k = forward(theta, lv, r)
        p = k["p"]
        s, x = k["s"], k["x"]
        rho = (r - lv.A(c)) / s
        tot += np.mean(rho * rho)
        N = lv.n
        drho = 2.0 * rho / (N * B)
        # rho = x - A (y / D)  (c / s = y / D)
        dt = -lv.AT(drho)
        dy = dt / lv.D
        d = {name: np.zeros(shp) for name, shp in SHAPES}
        d["alpha"][0] = dy @ x
        d["bo"][0] = dy.sum()
        d["wo"] = k["h3"].T @ dy
        dh3 = np.outer(dy, p["wo"])
        dz3 = dh3 * (k["h3"] > 0)
        d["W3"] = k["g2"].T @ dz3
        d["b3"] = dz3.sum(axis=0)
        dg2 = dz3 @ p["W3"].T
        dh2 = dg2[:, :HIDDEN] + lv.PT(dg2[:, HIDDEN:])
        dz2 = dh2 * (k["h2"] > 0)
        d["W2"] = k["g1"].T @ dz2
        d["b2"] = dz2.sum(axis=0)
        dg1 = dz2 @ p["W2"].T
        dh1 = dg1[:, :HIDDEN] + lv.PT(dg1[:, HIDDEN:])
        dz1 = dh1 * (k["h1"] > 0)
        d["W1"] = k["f0"].T @ dz1
        d["b1"] = dz1.sum(axis=0)
        g += pack(d)
    return tot / B, g


def backward(theta, lv: Level, cache, dy):
    """Parameter gradient from dloss/dy of one sample."""
    p, k = cache["p"], cache
    x = k["x"]
    d = {name: np.zeros(shp) for name, shp in SHAPES}
    d["alpha"][0] = dy @ x
    d["bo"][0] = dy.sum()
    d["wo"] = k["h3"].T @ dy
    dh3 = np.outer(dy, p["wo"])
    dz3 = dh3 * (k["h3"] > 0)
    d["W3"] = k["g2"].T @ dz3
    d["b3"] = dz3.sum(axis=0)
    dg2 = dz3 @ p["W3"].T
    dh2 = dg2[:, :HIDDEN] + lv.PT(dg2[:, HIDDEN:])
    dz2 = dh2 * (k["h2"] > 0)
    d["W2"] = k["g1"].T @ dz2
    d["b2"] = dz2.sum(axis=0)
    dg1 = dz2 @ p["W2"].T
    dh1 = dg1[:, :HIDDEN] + lv.PT(dg1[:, HIDDEN:])
    dz1 = dh1 * (k["h1"] > 0)
    d["W1"] = k["f0"].T @ dz1
    d["b1"] = dz1.sum(axis=0)
    return pack(d)


def high_pass(lv: Level, v):
    return v - lv.Pr @ (lv.R @ v)


def loss_err(theta, lv: Level, R, E):
    """Mean over samples and vertices of (κ H (e - c) / s)², κ = mean(D)."""
    tot, kap = 0.0, np.mean(lv.D)
    for r, e in zip(R, E):
        c, k = forward(theta, lv, r)
        eh = high_pass(lv, kap * (e - c) / k["s"])
        tot += np.mean(eh * eh)
    return tot / len(R)


def loss_err_grad(theta, lv: Level, R, E):
    g = np.zeros_like(theta)
    tot, B, N, kap = 0.0, len(R), lv.n, np.mean(lv.D)
    for r, e in zip(R, E):
        c, k = forward(theta, lv, r)
        eps = kap * (e - c) / k["s"]    # = κ (e / s - y / D)
        eh = high_pass(lv, eps)
        tot += np.mean(eh * eh)
        geps = 2.0 * kap * (eh - lv.R.T(lv.Pr.T(eh))) / (N * B)   # κ Hᵀ H eps
        g += backward(theta, lv, k, -geps / lv.D)
    return tot / B, g


def fd_grad(theta, lv, R, idx, h=1e-6, E=None):
    out = []
    f = (lambda t: loss(t, lv, R)) if E is None else (lambda t: loss_err(t, lv, R, E))
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        out.append((f(tp) - f(tm)) / (2 * h))
    return np.array(out)
