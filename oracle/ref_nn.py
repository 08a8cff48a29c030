"""ORACLE — test infrastructure only.  numpy restatement of the learned guide's
network (naivetrace_amd/csrc/nn.hpp, nn_device.hpp, nn.hip) for checking the
GPU's forward pass and gradients.

The network has no counterpart in the reference (new guiding method).  It is
pinned by central finite differences of its own loss (tests/test_nn.py) in
float64; with `bf16=True` it rounds where the GPU stores bf16 (hash-grid entries
and weights, layer activations, output and hidden gradients) so that the GPU
can be compared at fp32-accumulation tolerance.
"""
import numpy as np

K_IN, K_HID, K_OUT, K_OUT_USED, K_LOBES = 64, 64, 96, 80, 16
KAPPA_MIN, KAPPA_MAX = 0.05, 100.0


def bf16(x):
    """Round to bfloat16 (nearest even) and back to float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


class Layout:
    def __init__(self, levels, log2_table):
        self.levels, self.log2_table = levels, log2_table
        self.T = 1 << log2_table
        self.grid = 0
        self.w1 = levels * self.T * 2
        self.w2 = self.w1 + K_HID * K_IN
        self.w3 = self.w2 + K_HID * K_HID
        self.w4 = self.w3 + K_HID * K_HID
        self.b1 = self.w4 + K_OUT * K_HID
        self.b2 = self.b1 + K_HID
        self.b3 = self.b2 + K_HID
        self.b4 = self.b3 + K_HID
        self.total = self.b4 + K_OUT


class Net:
    def __init__(self, levels, log2_table, base_res, max_res, bb_lo, bb_hi):
        self.L = Layout(levels, log2_table)
        b = np.exp((np.log(max_res) - np.log(base_res)) / (levels - 1)) if levels > 1 else 1.0
        self.res = np.floor(np.float32(base_res) * np.power(b, np.arange(levels))).astype(np.float32)
        lo, hi = np.asarray(bb_lo, np.float32), np.asarray(bb_hi, np.float32)
        ext = np.maximum(hi - lo, np.float32(1e-6))
        self.bb_lo = (lo - np.float32(1e-3) * ext).astype(np.float32)
        self.bb_inv = (np.float32(1.0) / (ext * np.float32(1.002))).astype(np.float32)

    def corners(self, pos, lv):
        """Indices (n, 8) and trilinear weights (n, 8) of level lv."""
        p01 = np.clip((pos.astype(np.float32) - self.bb_lo) * self.bb_inv, 0, 1)
        res = self.res[lv]
        x = p01 * res
        fl = np.floor(x)
        f = x - fl
        c = fl.astype(np.uint64)
        dim = int(res) + 2
        dense = dim ** 3 <= self.L.T
        mask = self.L.T - 1
        idx = np.zeros((len(pos), 8), np.int64)
        w = np.zeros((len(pos), 8), np.float32)
        for k in range(8):
            d = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1], np.uint64)
            cc = c + d
            if dense:
                ii = (cc[:, 0] + np.uint64(dim) * (cc[:, 1] + np.uint64(dim) * cc[:, 2])) & np.uint64(mask)
            else:
                ii = (cc[:, 0] ^ ((cc[:, 1] * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF))
                      ^ ((cc[:, 2] * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))) & np.uint64(mask)
            idx[:, k] = ii.astype(np.int64)
            wk = np.ones(len(pos), np.float32)
            for a in range(3):
                wk = wk * (f[:, a] if d[a] else 1 - f[:, a])
            w[:, k] = wk
        return idx, w

    def unpack(self, p, rnd):
        L = self.L
        grid = p[L.grid:L.w1].reshape(L.levels, L.T, 2)
        W = [p[L.w1:L.w2].reshape(K_HID, K_IN), p[L.w2:L.w3].reshape(K_HID, K_HID), p[L.w3:L.w4].reshape(K_HID, K_HID),
             p[L.w4:L.b1].reshape(K_OUT, K_HID)]
        B = [p[L.b1:L.b2], p[L.b2:L.b3], p[L.b3:L.b4], p[L.b4:L.total]]
        if rnd:
            grid = bf16(grid)
            W = [bf16(w) for w in W]
        return grid, W, B

    def forward(self, p, pos, nrm, wo, bf=True, keep=False):
        """Raw outputs (n, 96); with keep, also the cached values for backward."""
        dt = np.float32 if bf else np.float64
        r = bf16 if bf else (lambda x: x)
        grid, W, B = self.unpack(p.astype(dt), bf)
        n = len(pos)
        x = np.zeros((n, K_IN), dt)
        cache = []
        for lv in range(self.L.levels):
            idx, w = self.corners(pos, lv)
            g = grid[lv][idx]  # (n, 8, 2)
            x[:, 2 * lv:2 * lv + 2] = np.einsum("nk,nkf->nf", w.astype(dt), g)
            cache.append((idx, w))
        x[:, 32:35] = nrm
        x[:, 35:38] = wo
        x = r(x)
        acts = [x]
        h = x
        for l in range(3):
            z = h @ W[l].T.astype(dt) + B[l].astype(dt)
            h = r(np.maximum(z, 0))
            acts.append(h)
        out = h @ W[3].T.astype(dt) + B[3].astype(dt)
        return (out, (acts, cache, W)) if keep else out

    @staticmethod
    def mixture_nll_grad(raw, wi, w):
        """Per-record loss -w log q(wi) and its gradient (n, 96) at the raw outputs."""
        n = len(raw)
        g = np.zeros((n, K_OUT), raw.dtype)
        R = raw[:, :K_OUT_USED].reshape(n, K_LOBES, 5)
        logit, m, s = R[:, :, 0], R[:, :, 1:4], R[:, :, 4]
        e = np.exp(logit - logit.max(axis=1, keepdims=True))
        pi = e / e.sum(axis=1, keepdims=True)
        length = np.maximum(np.linalg.norm(m, axis=2), 1e-12)
        sp = np.where(s > 15, s, np.log1p(np.exp(np.minimum(s, 15))))
        kap = np.minimum(sp + KAPPA_MIN, KAPPA_MAX)
        c = np.einsum("njk,nk->nj", m, wi) / length
        v = kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * (c - 1))
        q = (pi * v).sum(axis=1)
        ok = (q > 1e-30) & (w > 0)
        qs = np.where(ok, q, 1.0)
        rr = pi * v / qs[:, None]
        wl = np.where(ok, w, 0.0)[:, None]
        G = np.zeros_like(R)
        G[:, :, 0] = wl * (pi - rr)
        f = -wl * rr * kap / length
        mu = m / length[:, :, None]
        G[:, :, 1:4] = f[:, :, None] * (wi[:, None, :] - c[:, :, None] * mu)
        e2 = np.exp(-2 * kap)
        dlogv = 1 / kap - 2 * e2 / (1 - e2) + (c - 1)
        G[:, :, 4] = np.where(kap < KAPPA_MAX, -wl * rr * dlogv / (1 + np.exp(-s)), 0.0)
        g[:, :K_OUT_USED] = G.reshape(n, K_OUT_USED)
        loss = np.where(ok, -w * np.log(qs), 0.0)
        return loss, g

    def loss(self, p, pos, nrm, wo, wi, w, bf=False):
        out = self.forward(p, pos, nrm, wo, bf=bf)
        l, _ = self.mixture_nll_grad(out, wi, w)
        return l.mean()

    def gradient(self, p, pos, nrm, wo, wi, w, bf=True):
        """Gradient of the mean loss with respect to every parameter, and the mean loss."""
        dt = np.float32 if bf else np.float64
        r = bf16 if bf else (lambda x: x)
        out, (acts, cache, W) = self.forward(p, pos, nrm, wo, bf=bf, keep=True)
        n = len(pos)
        loss, dy = self.mixture_nll_grad(out, wi.astype(dt), w.astype(dt))
        dy = r(dy / n)
        L = self.L
        grad = np.zeros(L.total, dt)
        gW = [slice(L.w1, L.w2), slice(L.w2, L.w3), slice(L.w3, L.w4), slice(L.w4, L.b1)]
        gB = [slice(L.b1, L.b2), slice(L.b2, L.b3), slice(L.b3, L.b4), slice(L.b4, L.total)]
        d = dy
        for l in (3, 2, 1, 0):
            grad[gW[l]] = (d.T @ acts[l]).reshape(-1)
            grad[gB[l]] = d.sum(axis=0)
            back = d @ W[l].astype(dt)
            if l > 0:
                d = r(np.where(acts[l] > 0, back, 0.0))
            else:
                dx = back
        gg = np.zeros((L.levels, L.T, 2), dt)
        for lv in range(L.levels):
            idx, wc = cache[lv]
            for f in range(2):
                np.add.at(gg[lv, :, f], idx.reshape(-1), (wc * dx[:, 2 * lv + f][:, None]).reshape(-1))
        grad[L.grid:L.w1] = gg.reshape(-1)
        return grad, loss.mean()
