"""numpy model of how the kernels combine libeqd's host-built 1D tables.

Used by the CPU tests to check the table construction against the oracle's
assembled operators without a device. It mirrors the term structure of
stencil_node / k_rhs / k_extract in earthquakediffinitive.jl_amd/csrc/kernels.h.
"""
import numpy as np

from eqd import _lib


def dense(tab, ncols):
    start, width, coef = tab
    M = np.zeros((len(start), ncols))
    for r in range(len(start)):
        for t in range(width[r]):
            M[r, start[r] + t] = coef[r, t]
    return M


class TableModel:
    def __init__(self, n1, n23, L_normal, L_fault, lam, mu, l_f):
        self.g = _lib.GridDesc(n1, n23, L_normal, L_fault, lam, mu, l_f)
        self.n1, self.n23 = n1, n23
        self.NX, self.NY = 2 * n1 - 3, n23 - 2
        self.mface = n1 - 2
        t = lambda name: _lib.tables_export(self.g, name)
        NX, NY = self.NX, self.NY
        self.AXs, self.AXo, self.MX = (dense(t(k), NX) for k in ("AXs", "AXo", "MX"))
        self.Ts, self.To, self.D1, self.D1s = (dense(t(k), NY) for k in ("Ts", "To", "D1", "D1s"))
        self.D1full = dense(t("D1full"), n23)
        self.D1ext = dense(t("D1ext"), NY)
        self.Hx, self.Hy, self.Jt, self.Jn, self.Dn = (t(k) for k in ("Hx", "Hy", "Jt", "Jn", "Dn"))
        self.Jext = t("Jext")

    # v[c][z, y, m] arrays on the merged interior grid
    @staticmethod
    def ax(M, v):  # along m (last axis)
        return np.einsum("ij,zyj->zyi", M, v)

    @staticmethod
    def ay(M, v):
        return np.einsum("ij,zjm->zim", M, v)

    @staticmethod
    def az(M, v):
        return np.einsum("ij,jym->iym", M, v)

    def apply(self, v):
        ax, ay, az = self.ax, self.ay, self.az
        v0, v1, v2 = v
        a0 = ax(self.AXs, v0) + ay(self.To, v0) + az(self.To, v0) + ax(self.MX, ay(self.D1, v1)) \
            + ax(self.MX, az(self.D1, v2))
        a1 = ax(self.AXo, v1) + ay(self.Ts, v1) + az(self.To, v1) + ax(self.MX, ay(self.D1, v0)) \
            + ay(self.D1s, az(self.D1, v2))
        a2 = ax(self.AXo, v2) + ay(self.To, v2) + az(self.Ts, v2) + ax(self.MX, az(self.D1, v0)) \
            + ay(self.D1s, az(self.D1, v1))
        H = -self.Hy[:, None, None] * self.Hy[None, :, None] * self.Hx[None, None, :]
        return [H * a0, H * a1, H * a2]

    def rhs(self, f2, f3):
        """b on the merged grid from face fields f2, f3 (n23 x n23 arrays indexed [iz, iy])."""
        NX, NY = self.NX, self.NY
        H = self.Hy[:, None, None] * self.Hy[None, :, None] * self.Hx[None, None, :]
        s2i, s3i = f2[1:-1, 1:-1], f3[1:-1, 1:-1]
        b1 = H * self.Jt[None, None, :] * s2i[:, :, None]
        b2 = H * self.Jt[None, None, :] * s3i[:, :, None]
        d = (self.D1full @ f2[1:-1, :].T).T + self.D1full @ f3[:, 1:-1]
        b0 = H * self.Jn[None, None, :] * d[:, :, None]
        return [b0, b1, b2]

    def extract(self, u, f2, f3, i2lo, n2f):
        out2 = np.zeros((n2f, n2f)); out3 = np.zeros((n2f, n2f))  # [b, a]
        n23 = self.n23
        for b in range(n2f):
            for a in range(n2f):
                iy, iz = i2lo + a, i2lo + b
                yin, zin = 1 <= iy <= n23 - 2, 1 <= iz <= n23 - 2
                for d, (out, f) in enumerate(((out2, f2), (out3, f3))):
                    v = self.Jext * f[iz, iy]
                    if yin and zin:
                        v += self.Dn @ u[d + 1][iz - 1, iy - 1, :]
                    tg = 0.0
                    if d == 0 and zin:
                        tg = self.D1ext[iy] @ u[0][iz - 1, :, self.mface]
                    if d == 1 and yin:
                        tg = self.D1ext[iz] @ u[0][:, iy - 1, self.mface]
                    out[b, a] = v + self.g.mu * tg
        return out2.ravel(), out3.ravel()


def to_full(model, v, N):
    """Embed merged fields into the reference's full vector [u-(3N); u+(3N)]."""
    n1, n23 = model.n1, model.n23
    full = np.zeros(6 * N)
    lin = lambda i1, i2, i3: i1 + n1 * (i2 + n23 * i3)
    for c in range(3):
        for z in range(model.NY):
            for y in range(model.NY):
                for m in range(model.NX):
                    M = m + 1
                    val = v[c][z, y, m]
                    if M <= n1 - 1:
                        full[c * N + lin(M, y + 1, z + 1)] = val
                    if M >= n1 - 1:
                        full[3 * N + c * N + lin(M - (n1 - 1), y + 1, z + 1)] = val
    return full


def from_full(model, full, N):
    n1, n23 = model.n1, model.n23
    lin = lambda i1, i2, i3: i1 + n1 * (i2 + n23 * i3)
    v = [np.zeros((model.NY, model.NY, model.NX)) for _ in range(3)]
    for c in range(3):
        for z in range(model.NY):
            for y in range(model.NY):
                for m in range(model.NX):
                    M = m + 1
                    if M < n1 - 1:
                        v[c][z, y, m] = full[c * N + lin(M, y + 1, z + 1)]
                    elif M > n1 - 1:
                        v[c][z, y, m] = full[3 * N + c * N + lin(M - (n1 - 1), y + 1, z + 1)]
                    else:
                        v[c][z, y, m] = 0.5 * (full[c * N + lin(n1 - 1, y + 1, z + 1)]
                                               + full[3 * N + c * N + lin(0, y + 1, z + 1)])
    return v
