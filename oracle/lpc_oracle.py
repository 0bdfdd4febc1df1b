"""FP64 oracle of the learned block preconditioner (TEST INFRASTRUCTURE ONLY).

The network has no reference counterpart: it is the learned preconditioner of
the rows this package adds.  Only tests/ and bench.py's baseline leg use this file,
as the FP64 statement that the HIP kernels in adaptive-pf-flip_amd/csrc/pf_lpc.hip
are tested against.

Network M(r; beta, phi), block-local on 16^3 cell tiles with a 1-cell input halo:
  input X[18,18,18,4] (z, y, x, channel): rho_hat, beta_hat, phi, 0
  h1 = relu(conv3(X, W1))        4 -> 16   (valid on the halo: 18^3 -> 16^3)
  h2 = relu(conv3(pad0(h1), W2)) 16 -> 32  (zero padding at the tile boundary)
  h3 = relu(conv3(pad0(h2), W3)) 32 -> 16
  o  = h3 . Wh                    16 -> 1   (1x1x1 head, linear)
Weights: W*[tap][ci][co] with tap = dx + 3 dy + 9 dz, no biases.  The flat weight
vector is W1, W2, W3, Wh in that order (29,392 values).

bf16=True rounds the input, the weights and each stored activation (h1, h2, h3)
to bfloat16 (round to nearest even) exactly where the HIP kernels store them, and
computes everything else in FP64: the GPU's FP32 accumulation then differs only by
accumulation rounding.

Gradients are written out by hand (no autodiff); test_lpc_oracle.py pins them with
central finite differences of the FP64 forward.
"""
import numpy as np

CIN0 = 4
WIDTHS = (16, 32, 16)
T = 16          # tile edge (cells)
H = T + 2       # with the 1-cell halo

SHAPES = [(27, CIN0, 16), (27, 16, 32), (27, 32, 16), (16,)]
SIZES = [int(np.prod(s)) for s in SHAPES]
NW = sum(SIZES)


def bf16_round(a):
    """Round float values to bfloat16 (round to nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64).reshape(np.shape(a))


def split(w):
    out, o = [], 0
    for shp, n in zip(SHAPES, SIZES):
        out.append(np.asarray(w[o:o + n], np.float64).reshape(shp))
        o += n
    return out


def init_weights(seed=0):
    """He-style initialisation (the HIP library draws its own; this one is for tests)."""
    rng = np.random.default_rng(seed)
    ws = []
    for shp in SHAPES[:3]:
        fan_in = shp[0] * shp[1]
        ws.append(rng.normal(scale=np.sqrt(2.0 / fan_in), size=shp).ravel())
    ws.append(rng.normal(scale=np.sqrt(1.0 / 16), size=16))
    return np.concatenate(ws)


def conv3(x, w):
    """x[D+2, D+2, D+2, ci] (already padded), w[27, ci, co] -> y[D, D, D, co]."""
    d = x.shape[0] - 2
    y = np.zeros((d, d, d, w.shape[2]))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                y += x[dz:dz + d, dy:dy + d, dx:dx + d, :] @ w[dx + 3 * dy + 9 * dz]
    return y


def pad0(h):
    return np.pad(h, ((1, 1), (1, 1), (1, 1), (0, 0)))


def forward(X, w, bf16=False):
    """X[..., 18, 18, 18, 4] (one tile or a batch) -> (o[..., 16, 16, 16], cache)."""
    W1, W2, W3, Wh = split(w)
    r = bf16_round if bf16 else (lambda a: a)
    W1, W2, W3, Wh = r(W1), r(W2), r(W3), r(Wh)
    X = np.asarray(X, np.float64)
    batch = X.ndim == 5
    Xs = X if batch else X[None]
    outs, caches = [], []
    for x in Xs:
        x = r(x)
        h1 = r(np.maximum(conv3(x, W1), 0.0))
        h2 = r(np.maximum(conv3(pad0(h1), W2), 0.0))
        h3 = r(np.maximum(conv3(pad0(h2), W3), 0.0))
        o = h3 @ Wh
        outs.append(o)
        caches.append((x, h1, h2, h3))
    o = np.stack(outs)
    return (o if batch else o[0]), caches


def conv3_grad_w(x, g):
    """d/dw of sum(conv3(x, w) * g): x[D+2,..,ci], g[D,..,co] -> [27, ci, co]."""
    d = g.shape[0]
    ci, co = x.shape[3], g.shape[3]
    gw = np.zeros((27, ci, co))
    gf = g.reshape(-1, co)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                gw[dx + 3 * dy + 9 * dz] = x[dz:dz + d, dy:dy + d, dx:dx + d, :].reshape(-1, ci).T @ gf
    return gw


def conv3_grad_x(g, w):
    """d/dx of sum(conv3(pad0(x), w) * g) for the interior x[D,..,ci]: a correlation of
    the zero-padded g with the tap-flipped, transposed weights."""
    wf = w[::-1].transpose(0, 2, 1)  # tap' = 26 - tap, [co, ci]
    return conv3(pad0(g), wf)


def backward(caches, gO, w, bf16=False):
    """Gradient of sum(o * gO) w.r.t. the flat weight vector (FP64 throughout; with
    bf16=True the weights and cached activations are the rounded ones)."""
    W1, W2, W3, Wh = split(w)
    r = bf16_round if bf16 else (lambda a: a)
    W1, W2, W3, Wh = r(W1), r(W2), r(W3), r(Wh)
    gO = np.asarray(gO, np.float64)
    gOs = gO if gO.ndim == 4 else gO[None]
    g1 = np.zeros_like(W1); g2 = np.zeros_like(W2); g3 = np.zeros_like(W3); gh = np.zeros_like(Wh)
    for (x, h1, h2, h3), go in zip(caches, gOs):
        gh += np.tensordot(h3, go, axes=([0, 1, 2], [0, 1, 2]))
        d3 = go[..., None] * Wh * (h3 > 0)
        g3 += conv3_grad_w(pad0(h2), d3)
        d2 = conv3_grad_x(d3, W3) * (h2 > 0)
        g2 += conv3_grad_w(pad0(h1), d2)
        d1 = conv3_grad_x(d2, W2) * (h1 > 0)
        g1 += conv3_grad_w(x, d1)
    return np.concatenate([g1.ravel(), g2.ravel(), g3.ravel(), gh.ravel()])
