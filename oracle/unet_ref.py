"""FP64 numpy restatement of the surrogate U-Net (test infrastructure only).

The checker for the library's msx_unet_* entry points: the forward pass and
the analytic gradient of the mean-squared-error loss, for the same
architecture and the same packed parameter layout (conv by conv: encoder
levels 0..L-1 two each, decoder levels L-2..0 two each, 1x1 head; each conv as
W[Cout][Cin][k*k] then bias[Cout]).  The surrogate has no counterpart in the
reference, so there is nothing of the reference to follow here.

Pinning: the analytic gradient is checked against central finite differences
in FP64 (tests/test_unet_oracle.py); the convolution against a direct
nested-loop sum on tiny shapes.  With bf16=True every product operand is
rounded to bfloat16 (round to nearest even) before an FP64 sum, the arithmetic
of the GPU's matrix-core kernels up to the order of the FP32 accumulation.
"""
import numpy as np


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), via FP32, returned as FP64."""
    f = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def conv_specs(levels, base, cin=1):
    """[(cin, cout, ks, relu)] in packed order (cin: input channels)."""
    C = [base << l for l in range(levels)]
    s = []
    for l in range(levels):
        s.append((cin if l == 0 else C[l - 1], C[l], 3, True))
        s.append((C[l], C[l], 3, True))
    for l in range(levels - 2, -1, -1):
        s.append((C[l + 1] + C[l], C[l], 3, True))
        s.append((C[l], C[l], 3, True))
    s.append((C[0], 1, 1, False))
    return s


def param_count(levels, base, cin=1):
    return sum(co * ci * k * k + co for ci, co, k, _ in conv_specs(levels, base, cin))


def unpack(p, levels, base, cin=1):
    out, o = [], 0
    for ci, co, k, relu in conv_specs(levels, base, cin):
        w = np.asarray(p[o:o + co * ci * k * k], np.float64).reshape(co, ci * k * k)
        o += co * ci * k * k
        b = np.asarray(p[o:o + co], np.float64)
        o += co
        out.append((w, b, k, relu))
    return out


def im2col(x, ks):
    """x [N, C, H, W] -> [N, C*ks*ks, H*W], zero padding ks//2, tap = dy*ks+dx."""
    N, C, H, W = x.shape
    r = ks // 2
    xp = np.pad(x, ((0, 0), (0, 0), (r, r), (r, r)))
    cols = np.empty((N, C, ks * ks, H, W))
    for t in range(ks * ks):
        dy, dx = divmod(t, ks)
        cols[:, :, t] = xp[:, :, dy:dy + H, dx:dx + W]
    return cols.reshape(N, C * ks * ks, H * W)


def col2im(dcols, shape, ks):
    """Adjoint of im2col."""
    N, C, H, W = shape
    r = ks // 2
    d = dcols.reshape(N, C, ks * ks, H, W)
    dxp = np.zeros((N, C, H + 2 * r, W + 2 * r))
    for t in range(ks * ks):
        dy, dx = divmod(t, ks)
        dxp[:, :, dy:dy + H, dx:dx + W] += d[:, :, t]
    return dxp[:, :, r:r + H, r:r + W]


def conv_direct(x, w, b, ks):
    """Nested-loop convolution (the definition), for pinning im2col on tiny shapes."""
    N, C, H, W = x.shape
    co = w.shape[0]
    wk = w.reshape(co, C, ks, ks)
    r = ks // 2
    y = np.zeros((N, co, H, W))
    for n in range(N):
        for o in range(co):
            for yy in range(H):
                for xx in range(W):
                    s = b[o]
                    for c in range(C):
                        for dy in range(ks):
                            for dx in range(ks):
                                sy, sx = yy + dy - r, xx + dx - r
                                if 0 <= sy < H and 0 <= sx < W:
                                    s += wk[o, c, dy, dx] * x[n, c, sy, sx]
                    y[n, o, yy, xx] = s
    return y


class UNetRef:
    def __init__(self, levels, base, params, bf16=False, cin=1):
        self.L, self.B, self.bf16, self.cin = levels, base, bf16, cin
        self.convs = unpack(params, levels, base, cin)
        self.rnd = bf16_round if bf16 else (lambda a: a)

    def _conv(self, i, x):
        w, b, ks, relu = self.convs[i]
        cols = im2col(x, ks)
        y = np.matmul(self.rnd(w), self.rnd(cols)) + b[None, :, None]
        N, _, H, W = x.shape
        y = y.reshape(N, w.shape[0], H, W)
        if relu:
            y = np.maximum(y, 0.0)
        return y, cols

    def forward(self, x):
        """x [N, H, W] (one channel) or [N, cin, H, W] -> output [N, H, W]:
        the channel mean of x plus the head; keeps the tape for backward."""
        L = self.L
        x = np.asarray(x, np.float64)
        x4 = x[:, None] if x.ndim == 3 else x
        assert x4.shape[1] == self.cin
        tape = {}
        h = x4
        for l in range(L):
            ya, ca = self._conv(2 * l, h)
            yb, cb = self._conv(2 * l + 1, ya)
            tape[("enc", l)] = (h, ca, ya, cb, yb)
            if l < L - 1:
                N, C, H, W = yb.shape
                h = yb.reshape(N, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))
        prev = tape[("enc", L - 1)][4]
        i = 2 * L
        for l in range(L - 2, -1, -1):
            up = prev.repeat(2, axis=2).repeat(2, axis=3)
            cat = np.concatenate([up, tape[("enc", l)][4]], axis=1)
            ya, ca = self._conv(i, cat)
            yb, cb = self._conv(i + 1, ya)
            tape[("dec", l)] = (cat, ca, ya, cb, yb, i)
            prev = yb
            i += 2
        z, cz = self._conv(i, prev)
        tape["head"] = (prev, cz, z, i)
        self.tape = tape
        return x4.mean(axis=1) + z[:, 0]

    def _conv_bwd(self, i, xin, cols, y, g, grads):
        w, b, ks, relu = self.convs[i]
        N, C, H, W = xin.shape
        G = g * (y > 0) if relu else g
        Gm = G.reshape(N, w.shape[0], H * W)
        gw = np.einsum("nop,nkp->ok", self.rnd(Gm), self.rnd(cols))
        gb = Gm.sum(axis=(0, 2))
        grads[i] = (gw, gb)
        dcols = np.matmul(self.rnd(w).T, self.rnd(Gm))
        return col2im(dcols, xin.shape, ks)

    def loss_and_grad(self, x, t):
        out = self.forward(x)
        t = np.asarray(t, np.float64)
        cnt = out.size
        loss = float(((out - t) ** 2).sum() / cnt)
        gout = (2.0 / cnt) * (out - t)
        L, tape = self.L, self.tape
        grads = {}
        prev, cz, z, i = tape["head"]
        g = self._conv_bwd(i, prev, cz, z, gout[:, None], grads)
        g_enc = {}
        for l in range(0, L - 1):
            cat, ca, ya, cb, yb, j = tape[("dec", l)]
            g_ya = self._conv_bwd(j + 1, ya, cb, yb, g, grads)
            g_cat = self._conv_bwd(j, cat, ca, ya, g_ya, grads)
            cu = self.B << (l + 1)
            gu = g_cat[:, :cu]
            N, C, H, W = gu.shape
            g = gu.reshape(N, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))
            g_enc[l] = g_cat[:, cu:]
        g_b = g  # gradient of the bottom encoder output
        for l in range(L - 1, -1, -1):
            h, ca, ya, cb, yb = tape[("enc", l)]
            if l < L - 1:
                g_b = g_enc[l] + g_pool
            g_ya = self._conv_bwd(2 * l + 1, ya, cb, yb, g_b, grads)
            g_h = self._conv_bwd(2 * l, h, ca, ya, g_ya, grads)
            if l > 0:
                g_pool = 0.25 * g_h.repeat(2, axis=2).repeat(2, axis=3)
        flat = []
        for k in range(len(self.convs)):
            gw, gb = grads[k]
            flat += [gw.ravel(), gb]
        return loss, out, np.concatenate(flat)
