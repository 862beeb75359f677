"""CPU oracle of the surrogate's kernels (numpy, FP64).

TEST INFRASTRUCTURE ONLY: this module is the checker for
distributed-lb-ib_amd/csrc/lbmib_surr.hip.  Only tests/, the smoke check and
bench.py's CPU leg import it; the product never does.

The reference (cube_pthread/lbm_ib.c) has no learned model, so these
functions follow the published definitions of the operations rather than a
reference routine:
  - 3x3x3 "same" convolution with zero padding (cross-correlation, as in
    every deep-learning library): y[v][n] = b[n] + sum_t sum_c
    x[v + o(t)][c] w[n][t][c], o(t) = (t//9 - 1, t//3 % 3 - 1, t % 3 - 1);
  - its adjoints with respect to x and w;
  - 2x2x2 average pooling and nearest-neighbour 2x upsampling;
  - mean-squared error; Adam (Kingma & Ba 2015, Algorithm 1);
  - BF16 rounding to nearest even (the top 16 bits of an IEEE single).

Tensors are channels-last: [B, D, H, W, C].  What pins each function (tests/
test_surrogate_oracle.py): torch.nn.functional.conv3d on the CPU (an
independent implementation) for the convolution, finite differences of the
oracle's own forward for every adjoint, adjoint identities <A a, b> = <a, A^T b>
for pooling / upsampling, the closed form of Adam's first step, and torch's
float32 -> bfloat16 conversion for the rounding.
"""
import numpy as np

TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16_round(a):
    """Rounds float values to the nearest BF16 (ties to even); returns float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """BF16 bit patterns (uint16) of float values, round to nearest even."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shift(x, dz, dy, dx):
    """s[v] = x[v + (dz, dy, dx)] with zeros outside; x [B, D, H, W, C]."""
    B, D, H, W, C = x.shape
    out = np.zeros_like(x)
    zs, zd = (slice(dz, D), slice(0, D - dz)) if dz >= 0 else (slice(0, D + dz), slice(-dz, D))
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, zd, yd, xd, :] = x[:, zs, ys, xs, :]
    return out


def conv3d(x, w, b=None):
    """y [B, D, H, W, Cout] = conv(x [B, D, H, W, Cin], w [Cout, 27, Cin]) + b."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    y = np.zeros(x.shape[:4] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += _shift(x, dz, dy, dx) @ w[:, t, :].T
    if b is not None:
        y += np.asarray(b, dtype=np.float64)
    return y


def conv3d_bwd_data(dy, w):
    """dx = sum_t shift(dy, -o(t)) w[:, t, :]: the adjoint of conv3d in x."""
    dy = np.asarray(dy, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    dx = np.zeros(dy.shape[:4] + (w.shape[2],))
    for t, (dz, ddy, ddx) in enumerate(TAPS):
        dx += _shift(dy, -dz, -ddy, -ddx) @ w[:, t, :]
    return dx


def conv3d_bwd_weight(x, dy):
    """dw[n][t][c] = sum_v dy[v][n] x[v + o(t)][c]; db[n] = sum_v dy[v][n]."""
    x = np.asarray(x, dtype=np.float64)
    dy = np.asarray(dy, dtype=np.float64)
    cin, cout = x.shape[-1], dy.shape[-1]
    dyf = dy.reshape(-1, cout)
    dw = np.zeros((cout, 27, cin))
    for t, (dz, ddy, ddx) in enumerate(TAPS):
        dw[:, t, :] = dyf.T @ _shift(x, dz, ddy, ddx).reshape(-1, cin)
    return dw, dyf.sum(axis=0)


def pool2(x):
    B, D, H, W, C = x.shape
    return np.asarray(x, dtype=np.float64).reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def up2(c):
    return np.repeat(np.repeat(np.repeat(np.asarray(c, dtype=np.float64), 2, axis=1), 2, axis=2), 2, axis=3)


def up2_adjoint(ds):
    """Sum of the 8 children: the adjoint of up2."""
    B, D, H, W, C = ds.shape
    return np.asarray(ds, dtype=np.float64).reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def mse(out, tgt, nl):
    """Loss over the first nl channels and its gradient (zero elsewhere)."""
    out = np.asarray(out, dtype=np.float64)
    tgt = np.asarray(tgt, dtype=np.float64).reshape(out.shape[:-1] + (nl,))
    d = out[..., :nl] - tgt
    n = d.size
    g = np.zeros_like(out)
    g[..., :nl] = 2.0 * d / n
    return float((d * d).sum() / n), g


def adam(p, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------- network

def unet_layers(cin, cout, chans=(32, 64, 128, 256)):
    """Layer list of the encoder-decoder: (name, cin, cout, level, relu).

    Encoder level l: two convolutions (ReLU), then 2x2x2 average pooling
    into level l + 1.  Decoder level l (from the bottom): a convolution at
    level l + 1 down to chans[l] (ReLU), nearest upsampling plus the encoder
    skip of level l, a convolution (ReLU).  Output: one convolution to cout
    channels, no activation."""
    L = []
    prev = cin
    for lvl, c in enumerate(chans):
        L.append((f"enc{lvl}a", prev, c, lvl, True))
        L.append((f"enc{lvl}b", c, c, lvl, True))
        prev = c
    for lvl in range(len(chans) - 2, -1, -1):
        L.append((f"dec{lvl}up", chans[lvl + 1], chans[lvl], lvl + 1, True))
        L.append((f"dec{lvl}b", chans[lvl], chans[lvl], lvl, True))
    L.append(("out", chans[0], cout, 0, False))
    return L


def unet_forward(params, x, chans=(32, 64, 128, 256), rounding=True):
    """Forward pass in FP64; with rounding, every stored activation is
    rounded to BF16 as the GPU stores it.  params: name -> (w, b).
    Returns (output, saved activations)."""
    r = bf16_round if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    act = {"x": r(x)}
    h = act["x"]
    nlev = len(chans)
    for lvl in range(nlev):
        if lvl > 0:
            h = r(pool2(act[f"enc{lvl - 1}b"]))
            act[f"pool{lvl}"] = h
        w, b = params[f"enc{lvl}a"]
        h = r(np.maximum(conv3d(h, w, b), 0.0))
        act[f"enc{lvl}a"] = h
        w, b = params[f"enc{lvl}b"]
        h = r(np.maximum(conv3d(h, w, b), 0.0))
        act[f"enc{lvl}b"] = h
    d = act[f"enc{nlev - 1}b"]
    for lvl in range(nlev - 2, -1, -1):
        w, b = params[f"dec{lvl}up"]
        c = r(np.maximum(conv3d(d, w, b), 0.0))
        act[f"dec{lvl}up"] = c
        s = r(up2(c) + act[f"enc{lvl}b"])
        act[f"dec{lvl}s"] = s
        w, b = params[f"dec{lvl}b"]
        d = r(np.maximum(conv3d(s, w, b), 0.0))
        act[f"dec{lvl}b"] = d
    w, b = params["out"]
    out = r(conv3d(d, w, b))
    act["out"] = out
    return out, act


def unet_backward(params, act, dout, chans=(32, 64, 128, 256), rounding=True):
    """Gradients of every parameter for the output gradient dout, following
    the GPU's order of operations and its BF16 rounding of every stored
    gradient tensor.  Returns name -> (dw, db)."""
    r = bf16_round if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    grads = {}
    nlev = len(chans)
    g = r(dout)
    w, _ = params["out"]
    grads["out"] = conv3d_bwd_weight(act["dec0b"], g)
    g = r(conv3d_bwd_data(g, w) * (act["dec0b"] > 0))  # at dec0b's pre-activation
    dskip = {}
    for lvl in range(0, nlev - 1):
        # d = relu(conv(s)): g is the gradient at its pre-activation
        w, _ = params[f"dec{lvl}b"]
        grads[f"dec{lvl}b"] = conv3d_bwd_weight(act[f"dec{lvl}s"], g)
        ds = r(conv3d_bwd_data(g, w))
        dskip[lvl] = ds
        # s = up(c) + e: c = relu(conv(d_below))
        gc = r(up2_adjoint(ds) * (act[f"dec{lvl}up"] > 0))
        w, _ = params[f"dec{lvl}up"]
        below = act[f"dec{lvl + 1}b"] if lvl + 1 < nlev - 1 else act[f"enc{nlev - 1}b"]
        grads[f"dec{lvl}up"] = conv3d_bwd_weight(below, gc)
        g = r(conv3d_bwd_data(gc, w) * (below > 0))
    # g: gradient at the pre-activation of the bottom encoder output
    for lvl in range(nlev - 1, -1, -1):
        if lvl < nlev - 1:
            # enc{lvl}b feeds the skip of decoder level lvl and the pooling
            # into level lvl + 1, whose input gradient is gpool
            g = r((dskip[lvl] + up2(gpool) / 8.0) * (act[f"enc{lvl}b"] > 0))
        w, _ = params[f"enc{lvl}b"]
        grads[f"enc{lvl}b"] = conv3d_bwd_weight(act[f"enc{lvl}a"], g)
        g = r(conv3d_bwd_data(g, w) * (act[f"enc{lvl}a"] > 0))
        w, _ = params[f"enc{lvl}a"]
        inp = act["x"] if lvl == 0 else act[f"pool{lvl}"]
        grads[f"enc{lvl}a"] = conv3d_bwd_weight(inp, g)
        if lvl > 0:
            gpool = r(conv3d_bwd_data(g, w))
    return grads


# ------------------------------------------------------- engine coupling
#
# Restatements of the coupling kernels of csrc/lbmib_jump.hip (no reference
# counterpart: the reference steps every time step).  Fields are [z][y][x];
# coarse planes [nz/F][ny/F][nx/F]; fine cell x sits at position x, coarse
# cell i averages fine cells i F .. i F + F - 1, centred at i F + (F - 1)/2.

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def block_average(field, F):
    nz, ny, nx = field.shape
    return field.reshape(nz // F, F, ny // F, F, nx // F, F).mean(axis=(1, 3, 5))


def _axis_weights(p, F, n):
    c = (np.asarray(p, dtype=np.float64) - 0.5 * (F - 1)) / F
    i0 = np.floor(c)
    w = c - i0
    i0 = i0.astype(int)
    return np.clip(i0, 0, n - 1), np.clip(i0 + 1, 0, n - 1), w


def trilinear_at(coarse, F, pts):
    """Coarse plane [ncz][ncy][ncx] interpolated at positions pts [m, 3] (x, y, z)."""
    ncz, ncy, ncx = coarse.shape
    x0, x1, wx = _axis_weights(pts[:, 0], F, ncx)
    y0, y1, wy = _axis_weights(pts[:, 1], F, ncy)
    z0, z1, wz = _axis_weights(pts[:, 2], F, ncz)
    out = np.zeros(len(pts))
    for zi, zw in ((z0, 1 - wz), (z1, wz)):
        for yi, yw in ((y0, 1 - wy), (y1, wy)):
            for xi, xw in ((x0, 1 - wx), (x1, wx)):
                out += zw * yw * xw * coarse[zi, yi, xi]
    return out


def trilinear_fine(coarse, F):
    """Coarse plane interpolated at every fine cell: [nz][ny][nx]."""
    ncz, ncy, ncx = coarse.shape
    z, y, x = np.meshgrid(np.arange(ncz * F), np.arange(ncy * F), np.arange(ncx * F), indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    return trilinear_at(coarse, F, pts).reshape(ncz * F, ncy * F, ncx * F)


def feq(rho, ux, uy, uz):
    """[19][...] equilibria of init_eqlbrmdistrfuncDF0 (lbm_ib.c:1561-1594)."""
    usq = ux * ux + uy * uy + uz * uz
    out = []
    for q in range(19):
        cu = CX[q] * ux + CY[q] * uy + CZ[q] * uz
        out.append(W[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq))
    return np.array(out)


def jump_pulled(f2, delta, F):
    """Post-stream populations after a jump, for cells whose pulled sources
    are all updated interior cells.  f2: post-stream populations [19][z][y][x]
    before the jump (lbmib_get_df); delta: the four coarse change planes.
    The stored (post-collision) population of cell y is h_q(y) = f2_q(y + c_q);
    the jump adds feq(rho + drho, u + du) - feq(rho, u) at y with the moments
    of h(y); the pull then reads f2'_q(x) = h'_q(x - c_q).  Returns the
    prediction on the whole grid (meaningful away from the boundaries)."""
    roll = lambda a, q, s: np.roll(a, (s * CZ[q], s * CY[q], s * CX[q]), axis=(0, 1, 2))  # noqa: E731
    h = np.array([roll(f2[q], q, -1) for q in range(19)])  # h_q(y) = f2_q(y + c_q)
    rho = h.sum(0)
    ux, uy, uz = (CX[:, None, None, None] * h).sum(0) / rho, (CY[:, None, None, None] * h).sum(0) / rho, \
        (CZ[:, None, None, None] * h).sum(0) / rho
    d = [trilinear_fine(delta[a], F) for a in range(4)]
    dfeq = feq(rho + d[0], ux + d[1], uy + d[2], uz + d[3]) - feq(rho, ux, uy, uz)
    return np.array([f2[q] + roll(dfeq[q], q, 1) for q in range(19)])


def raster(pos, frc, F, dims):
    """Node count and elastic-force sums per coarse cell: pos [m, 3], frc [m, 3]
    (elastic force), dims (ncz, ncy, ncx).  Returns [4][ncz][ncy][ncx]."""
    ncz, ncy, ncx = dims
    out = np.zeros((4, ncz, ncy, ncx))
    ix = np.clip(np.floor((pos[:, 0] + 0.5) / F).astype(int), 0, ncx - 1)
    iy = np.clip(np.floor((pos[:, 1] + 0.5) / F).astype(int), 0, ncy - 1)
    iz = np.clip(np.floor((pos[:, 2] + 0.5) / F).astype(int), 0, ncz - 1)
    np.add.at(out[0], (iz, iy, ix), 1.0)
    for a in range(3):
        np.add.at(out[1 + a], (iz, iy, ix), frc[:, a])
    return out
