"""mlp_oracle.py -- TEST INFRASTRUCTURE ONLY: numpy restatement of the learned seed proposal's
network (forward, mixture negative log-likelihood, backward) and its Adam step.  Only tests/
and bench.py's CPU leg import it, as the checker; the library never calls it.

The network has no counterpart in the reference (whose seeds are uniform over the caster,
src/librender/manifold_ss.cpp:228-235).  Layout of the parameter vector (FP32): per layer the
weights output-major W[n][k] (layer 1: k < 160, the 153 Fourier features zero-padded), then
the bias; layers 9->160 features -> 64 -> 64 -> 64 -> 64 -> 48.

Two precisions:
  * ``"f64"``  -- exact arithmetic in float64 (pinned against torch autograd and central finite
    differences in tests/test_oracle_proposal.py);
  * ``"bf16"`` -- the rounding points of the gfx950 kernels: features, weights, hidden
    activations and back-propagated errors rounded to BF16 (round to nearest even), products
    accumulated in FP32, biases / head / loss in FP32.

The mixture head is written in covariance form, log N(d; 0, S) = -log 2 pi - log det S / 2
- d^T S^-1 d / 2 with S = [[sx^2, r sx sy], [r sx sy, sy^2]], and its gradient from
dlogN/dmu = S^-1 d, dlogN/dS = (S^-1 d d^T S^-1 - S^-1) / 2 -- a different derivation from the
kernels' whitened (Cholesky) form, so the two check each other.
"""
import numpy as np

IN, FREQ, FEAT, K1, H, OUT, COMP = 9, 8, 153, 160, 64, 48, 8
LOGSIG_MIN, LOGSIG_MAX, RHO_MAX = -7.0, 0.0, 0.95
LAYERS = [(H, K1), (H, H), (H, H), (H, H), (OUT, H)]   # (outputs, padded inputs)


def _offsets():
    offs, o = [], 0
    for n, k in LAYERS:
        offs.append((o, o + n * k))
        o += n * k + n
    return offs, o


OFFSETS, PARAMS = _offsets()
assert PARAMS == 25904


def bf16(a):
    """Round float32 values to BF16 (nearest even; NaN stays NaN), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params, dt=np.float64):
    p = np.asarray(params)
    out = []
    for (n, k), (w0, b0) in zip(LAYERS, OFFSETS):
        out.append((p[w0:w0 + n * k].reshape(n, k).astype(dt), p[b0:b0 + n].astype(dt)))
    return out


def features(x9, dt=np.float64):
    """[x, sin(pi 2^f x), cos(pi 2^f x) for f < 8] -> (n, 160), zero padded."""
    x = np.asarray(x9, dt).reshape(-1, IN)
    f = np.zeros((len(x), K1), dt)
    f[:, :IN] = x
    for k in range(FREQ):
        w = dt(np.pi) * dt(1 << k) if dt is np.float64 else np.float32(np.pi) * np.float32(1 << k)
        a = (w * x).astype(dt)
        f[:, IN + 2 * k * IN:IN + (2 * k + 1) * IN] = np.sin(a)
        f[:, IN + (2 * k + 1) * IN:IN + (2 * k + 2) * IN] = np.cos(a)
    return f


def head(o):
    """raw outputs (n, 48) -> weights, means (u, v), sigmas (u, v), correlation (float64)."""
    o = np.asarray(o, np.float64).reshape(-1, COMP, 6)
    lg = o[..., 0]
    w = np.exp(lg - lg.max(1, keepdims=True))
    w /= w.sum(1, keepdims=True)
    mu = 1.0 / (1.0 + np.exp(-o[..., 1:3]))
    sig = np.exp(np.clip(o[..., 3:5], LOGSIG_MIN, LOGSIG_MAX))
    rho = RHO_MAX * np.tanh(o[..., 5])
    return w, mu, sig, rho


def nll(o, uv):
    """Per-sample negative log-likelihood of uv (n, 2) under the mixtures of o (n, 48)."""
    w, mu, sig, rho = head(o)
    d = np.asarray(uv, np.float64).reshape(-1, 1, 2) - mu
    sx, sy = sig[..., 0], sig[..., 1]
    det = (sx * sy) ** 2 * (1 - rho ** 2)
    # d^T S^-1 d with S^-1 = [[sy^2, -r sx sy], [-r sx sy, sx^2]] / det
    q = (d[..., 0] ** 2 * sy ** 2 - 2 * rho * sx * sy * d[..., 0] * d[..., 1] + d[..., 1] ** 2 * sx ** 2) / det
    logn = -np.log(2 * np.pi) - 0.5 * np.log(det) - 0.5 * q
    a = np.log(w) + logn
    m = a.max(1, keepdims=True)
    return -(m[:, 0] + np.log(np.exp(a - m).sum(1)))


def nll_grad(o, uv):
    """(per-sample loss, d loss / d raw outputs (n, 48)) in float64, covariance form."""
    o = np.asarray(o, np.float64).reshape(-1, OUT)
    n = len(o)
    w, mu, sig, rho = head(o)
    d = np.asarray(uv, np.float64).reshape(-1, 1, 2) - mu
    sx, sy = sig[..., 0], sig[..., 1]
    S = np.empty((n, COMP, 2, 2))
    S[..., 0, 0] = sx * sx
    S[..., 1, 1] = sy * sy
    S[..., 0, 1] = S[..., 1, 0] = rho * sx * sy
    Si = np.linalg.inv(S)
    _, logdet = np.linalg.slogdet(S)
    Sid = np.einsum("ncij,ncj->nci", Si, d)
    q = np.einsum("nci,nci->nc", d, Sid)
    logn = -np.log(2 * np.pi) - 0.5 * logdet - 0.5 * q
    a = np.log(w) + logn
    m = a.max(1, keepdims=True)
    s = np.exp(a - m).sum(1, keepdims=True)
    loss = -(m[:, 0] + np.log(s[:, 0]))
    gam = np.exp(a - m) / s                                  # responsibilities
    dS = 0.5 * (np.einsum("nci,ncj->ncij", Sid, Sid) - Si)   # d logN / d S (symmetric)
    dmu = Sid                                                # d logN / d mu
    # S entries as functions of (sx, sy, rho)
    dsx = 2 * sx * dS[..., 0, 0] + 2 * rho * sy * dS[..., 0, 1]
    dsy = 2 * sy * dS[..., 1, 1] + 2 * rho * sx * dS[..., 0, 1]
    drho = 2 * sx * sy * dS[..., 0, 1]
    g = np.zeros((n, COMP, 6))
    g[..., 0] = w - gam
    raw = o.reshape(n, COMP, 6)
    g[..., 1] = -gam * dmu[..., 0] * mu[..., 0] * (1 - mu[..., 0])
    g[..., 2] = -gam * dmu[..., 1] * mu[..., 1] * (1 - mu[..., 1])
    inside = (raw[..., 3:5] > LOGSIG_MIN) & (raw[..., 3:5] < LOGSIG_MAX)
    g[..., 3] = -gam * dsx * sx * inside[..., 0]
    g[..., 4] = -gam * dsy * sy * inside[..., 1]
    g[..., 5] = -gam * drho * RHO_MAX * (1 - np.tanh(raw[..., 5]) ** 2)
    return loss, g.reshape(n, OUT)


def forward(params, x9, mode="f64"):
    """Raw outputs (n, 48) and the activations [x0, h1..h4]."""
    if mode == "f64":
        Ws = unpack(params, np.float64)
        h = features(x9, np.float64)
        acts = [h]
        for i, (W, b) in enumerate(Ws):
            z = h @ W.T + b
            h = np.maximum(z, 0.0) if i < 4 else z
            acts.append(h)
        return h, acts
    Ws = [(bf16(W), b) for W, b in unpack(params, np.float32)]
    h = bf16(features(x9, np.float32))
    acts = [h]
    for i, (W, b) in enumerate(Ws):
        z = (h @ W.T).astype(np.float32) + b
        h = bf16(np.maximum(z, np.float32(0))) if i < 4 else z
        acts.append(h)
    return h, acts


def loss_and_grad(params, x9, uv, mode="f64", w=None):
    """Weighted mean NLL sum_j w_j NLL_j / sum_j w_j over the batch (w None: unit weights) and
    its gradient w.r.t. the parameter vector."""
    o, acts = forward(params, x9, mode)
    n = len(o)
    loss, d_o = nll_grad(o, uv)
    wt = np.ones(n) if w is None else np.asarray(w, np.float64)
    grad = np.zeros(PARAMS, np.float64)
    if mode == "f64":
        Ws = unpack(params, np.float64)
        dy = d_o * (wt / wt.sum())[:, None]
        for i in range(4, -1, -1):
            W, _ = Ws[i]
            w0, b0 = OFFSETS[i]
            grad[w0:w0 + W.size] = (dy.T @ acts[i]).ravel()
            grad[b0:b0 + W.shape[0]] = dy.sum(0)
            if i > 0:
                dy = (dy @ W) * (acts[i] > 0)
        return float((wt * loss).sum() / wt.sum()), grad
    # the kernels' order: head gradients times w_j, rounded to BF16, back-propagated with FP32
    # sums, and the summed gradient times 1 / sum(w) at the end
    Ws = [(bf16(W), b) for W, b in unpack(params, np.float32)]
    w32 = wt.astype(np.float32)
    inv = np.float32(1) / w32.sum(dtype=np.float32)
    dy = bf16((d_o * w32[:, None]).astype(np.float32))
    for i in range(4, -1, -1):
        W, _ = Ws[i]
        w0, b0 = OFFSETS[i]
        grad[w0:w0 + W.size] = ((dy.T @ acts[i]).astype(np.float32) * inv).ravel()
        grad[b0:b0 + W.shape[0]] = dy.sum(0, dtype=np.float32) * inv
        if i > 0:
            dy = bf16(((dy @ W).astype(np.float32)) * (acts[i] > 0))
    return float((w32 * loss.astype(np.float32)).sum() / w32.sum()), grad


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (float32, the kernel's order of operations)."""
    f = np.float32
    g = np.asarray(grad, f)
    m = f(b1) * m + (f(1) - f(b1)) * g
    v = f(b2) * v + (f(1) - f(b2)) * g * g
    bc1, bc2 = f(1 - b1 ** t), f(1 - b2 ** t)
    p = np.asarray(params, f) - f(lr) * (m / bc1) / (np.sqrt(v / bc2) + f(eps))
    return p.astype(f), m.astype(f), v.astype(f)


def init_like_library(seed):
    """Not a copy of the library's initial weights (those come from its own generator): He
    initialisation with the same shape and scales, for tests that need a plausible network."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for i, ((n, k), (w0, _)) in enumerate(zip(LAYERS, OFFSETS)):
        used = FEAT if i == 0 else k
        W = np.zeros((n, k), np.float32)
        W[:, :used] = rng.normal(0, np.sqrt(2.0 / used) * (0.1 if i == 4 else 1.0), (n, used))
        p[w0:w0 + n * k] = W.ravel()
    return p
