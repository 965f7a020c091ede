"""CPU restatement of the refinement predictor's MLP (test infrastructure only).

2 x 64 tanh MLP, sigmoid output, mean binary cross-entropy; parameter layout
W1[64][NF] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3, inputs standardised with
(mean, sdev).  Gradients are written out by hand (chain rule) and are themselves
checked against central finite differences in tests/test_predictor.py.
"""
import numpy as np

H = 64


def unpack(P, nf):
    o = 0
    W1 = P[o:o + H * nf].reshape(H, nf); o += H * nf
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].reshape(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    w3 = P[o:o + H]; o += H
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def forward(P, X, mean, sdev):
    nf = X.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(P, np.float64), nf)
    Xn = (np.asarray(X, np.float64) - mean) / sdev
    H1 = np.tanh(Xn @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    z = H2 @ w3 + b3
    return 1.0 / (1.0 + np.exp(-z)), (Xn, H1, H2)


def loss(P, X, y, mean, sdev):
    p, _ = forward(P, X, mean, sdev)
    p = np.clip(p, 1e-12, 1 - 1e-12)
    return float(-np.mean(y * np.log(p) + (1 - y) * np.log(1 - p)))


def gradient(P, X, y, mean, sdev):
    nf = X.shape[1]
    n = len(y)
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(P, np.float64), nf)
    p, (Xn, H1, H2) = forward(P, X, mean, sdev)
    d3 = (p - y) / n
    gw3 = H2.T @ d3
    gb3 = d3.sum()
    dZ2 = np.outer(d3, w3) * (1 - H2 ** 2)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (1 - H1 ** 2)
    gW1 = dZ1.T @ Xn
    gb1 = dZ1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
