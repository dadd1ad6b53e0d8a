"""numpy reference of the step-controller perceptron (checker for the HIP
kernels in macaw_amd/csrc/mg_mlp.hip): 12 -> 64 -> 64 -> 1, SiLU, weights
rounded to BF16 (nearest even), everything else FP32 (float64 here)."""
import numpy as np

F, H = 12, 64
oW1 = 2 * F
oB1 = oW1 + F * H
oW2 = oB1 + H
oB2 = oW2 + H * H
oW3 = oB2 + H
oB3 = oW3 + H
NPARAM = oB3 + 1


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p, round_weights=True):
    p = np.asarray(p, np.float64)
    r = (lambda a: bf16(a).astype(np.float64)) if round_weights else (lambda a: a)
    return dict(mu=p[:F], sd=p[F:2 * F], W1=r(p[oW1:oB1]).reshape(F, H), b1=p[oB1:oW2],
                W2=r(p[oW2:oB2]).reshape(H, H), b2=p[oB2:oW3], w3=r(p[oW3:oB3]), b3=p[oB3])


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(p, X, round_weights=True):
    q = unpack(p, round_weights)
    xh = (np.asarray(X, np.float64) - q["mu"]) / q["sd"]
    z1 = xh @ q["W1"] + q["b1"]
    h1 = silu(z1)
    z2 = h1 @ q["W2"] + q["b2"]
    h2 = silu(z2)
    return h2 @ q["w3"] + q["b3"], (xh, z1, h1, z2, h2, q)


def loss_grad(p, X, t, round_weights=True):
    """Mean squared error and its gradient w.r.t. the trainable parameters
    (taken at the rounded weights, as the kernels do)."""
    y, (xh, z1, h1, z2, h2, q) = forward(p, X, round_weights)
    B = len(t)
    e = y - np.asarray(t, np.float64)
    dy = 2 * e / B
    dz2 = dy[:, None] * q["w3"][None, :] * dsilu(z2)
    dz1 = (dz2 @ q["W2"].T) * dsilu(z1)
    g = np.zeros(NPARAM)
    g[oW1:oB1] = (xh.T @ dz1).ravel()
    g[oB1:oW2] = dz1.sum(0)
    g[oW2:oB2] = (h1.T @ dz2).ravel()
    g[oB2:oW3] = dz2.sum(0)
    g[oW3:oB3] = h2.T @ dy
    g[oB3] = dy.sum()
    return float(np.mean(e * e)), g


def random_params(rng, scale=0.3):
    p = rng.normal(scale=scale, size=NPARAM)
    p[:F] = rng.normal(size=F)
    p[F:2 * F] = rng.uniform(0.5, 2.0, size=F)
    return p.astype(np.float32)
