"""CPU oracle of the budget MLP's training step (48 -> 64 ReLU -> 64 ReLU -> 1, MSE).

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels (hdg_budget.hip).  Follows the
same numerics: inputs, hidden activations and the two back-propagated hidden gradients rounded to bf16
(round to nearest even), float32 products and sums accumulated sequentially over the input index.
The parameter vector layout is W1 [48][64], b1 [64], W2 [64][64], b2 [64], W3 [64], b3 [1].
"""
import numpy as np

F, H = 48, 64
P1, P2 = F * H, H * H
OFF_W1, OFF_B1, OFF_W2 = 0, P1, P1 + H
OFF_B2 = OFF_W2 + P2
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def bf16(x):
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _unpack(p):
    p = np.asarray(p, np.float32)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X):
    """Returns (y, z1, z2, h1, h2, x) for a batch (float32, sequential accumulation)."""
    W1, b1, W2, b2, W3, b3 = _unpack(p)
    x = bf16(X)
    n = x.shape[0]
    z1 = np.tile(b1, (n, 1)).astype(np.float32)
    for i in range(F):
        z1 = (z1 + (x[:, i:i + 1] * W1[i][None, :]).astype(np.float32)).astype(np.float32)
    h1 = bf16(np.maximum(z1, 0))
    z2 = np.tile(b2, (n, 1)).astype(np.float32)
    for i in range(H):
        z2 = (z2 + (h1[:, i:i + 1] * W2[i][None, :]).astype(np.float32)).astype(np.float32)
    h2 = bf16(np.maximum(z2, 0))
    y = np.full(n, b3, np.float32)
    for i in range(H):
        y = (y + (h2[:, i] * W3[i]).astype(np.float32)).astype(np.float32)
    return y, z1, z2, h1, h2, x


def grads(p, X, Y):
    """Gradient of mean squared error over the batch, as the HIP training kernel defines it."""
    W1, b1, W2, b2, W3, b3 = _unpack(p)
    y, z1, z2, h1, h2, x = forward(p, X)
    n = X.shape[0]
    err = (y - np.asarray(Y, np.float32)).astype(np.float32)
    loss = float(np.sum(err.astype(np.float64) ** 2) / n)
    gy = (np.float32(2.0) * err / np.float32(n)).astype(np.float32)
    g = np.zeros(N_PARAMS, np.float64)
    g[OFF_B3] = gy.sum()
    g[OFF_W3:OFF_B3] = (gy[:, None] * h2).sum(0)
    g2 = np.where(z2 > 0, bf16(gy[:, None] * W3[None, :]), 0).astype(np.float32)
    g[OFF_W2:OFF_B2] = np.einsum("ni,nj->ij", h1.astype(np.float64), g2.astype(np.float64)).ravel()
    g[OFF_B2:OFF_W3] = g2.sum(0)
    back = np.zeros((n, H), np.float32)
    for j in range(H):
        back = (back + (W2[:, j][None, :] * g2[:, j:j + 1]).astype(np.float32)).astype(np.float32)
    g1 = np.where(z1 > 0, bf16(back), 0).astype(np.float32)
    g[OFF_B1:OFF_W2] = g1.sum(0)
    g[OFF_W1:OFF_B1] = np.einsum("nf,ni->fi", x.astype(np.float64), g1.astype(np.float64)).ravel()
    return g.astype(np.float32), loss
