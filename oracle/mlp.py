"""CPU restatement of the closure network and its training (test infrastructure).

The network: 7 inputs -> 32 -> 32 -> 1 with ReLU after both hidden layers,
parameters in one flat vector [W1 (32x7, row = output unit), b1 (32),
W2 (32x32), b2 (32), w3 (32), b3 (1)] = 1345 values.  Loss: half the mean
squared error over a batch.  Adam: Kingma & Ba, "Adam: A Method for Stochastic
Optimization" (ICLR 2015), Algorithm 1, with bias-corrected moments.
Everything here is FP64 numpy; it checks the HIP kernels, which run FP32.
"""
import numpy as np

NIN, H = 7, 32
NPARAM = H * NIN + H + H * H + H + H + 1  # 1345


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * NIN].reshape(H, NIN); o += H * NIN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3, (a1, h1, a2, h2)


def loss_and_grad(p, x, t):
    """(0.5 mean (y - t)^2, gradient in the flat layout) over the batch x (n x 7), t (n)."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (a1, h1, a2, h2) = forward(p, x)
    r = y - t
    n = x.shape[0]
    dy = r / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, w3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return 0.5 * float(np.mean(r * r)), g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mhat = m / (1.0 - beta1 ** step)
    vhat = v / (1.0 - beta2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v


def init_params(seed=0):
    """He-style random initialisation (uniform), zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (H, NIN)) * np.sqrt(6.0 / NIN)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    w3 = rng.uniform(-1, 1, H) * np.sqrt(3.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(special, u.astype(np.uint32), r)
    return r.view(np.float32)
