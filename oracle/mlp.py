"""FP64 numpy restatement of the surrogate network (test infrastructure only).

Dense layers Y = act(X W^T + b), GELU (tanh form, Hendrycks and Gimpel
2016) on hidden layers, inputs
scaled to [-1, 1] by the prior bounds, loss sum_b sum_k w_k (y_bk - t_bk)^2 / B.
Forward and backward by hand (chain rule), used to check the GPU kernels'
gradients (altar_amd/csrc/mlp.hip).
"""
import numpy as np

K0 = np.sqrt(2.0 / np.pi)
K1 = 0.044715


def gelu(x):
    """x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K0 * (x + K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


def scale_inputs(theta, lo, hi):
    return 2.0 * (np.asarray(theta) - lo) / (hi - lo) - 1.0


def forward(layers, x):
    """layers: [(W, b)]; returns the output and the per-layer (input, pre-activation)."""
    cache = []
    h = x
    for l, (W, b) in enumerate(layers):
        pre = h @ W.T + b
        cache.append((h, pre))
        h = gelu(pre) if l < len(layers) - 1 else pre
    return h, cache


def loss_and_grads(layers, x, target, w):
    """Weighted mean squared error and its gradients [(dW, db)]."""
    y, cache = forward(layers, x)
    B = x.shape[0]
    diff = y - target
    loss = float(np.sum(w * diff * diff) / B)
    delta = 2.0 * w * diff / B
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        h, pre = cache[l]
        grads[l] = (delta.T @ h, delta.sum(axis=0))
        if l > 0:
            delta = (delta @ layers[l][0]) * gelu_grad(cache[l - 1][1])
    return loss, grads
