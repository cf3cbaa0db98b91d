"""FP64 plaintext model of the encrypted BERT-base layer program (test
infrastructure only).

The layer is built from NEXUS's own evaluators (MatrixMul-style plaintext
weights, GELU src/gelu.cpp:5-111, softmax src/softmax.cpp:6-42, LayerNorm
src/layer_norm.cpp:6-33, helpers src/ckks_evaluator.cpp:366-548); NEXUS ships
no composed layer.  `layer_approx` evaluates in float64 the same sequence of
polynomial steps the encrypted program runs, so decrypted engine outputs must
agree with it up to CKKS noise and bootstrapping error; `layer_exact` is the
exact FP64 transformer layer with the reference's function definitions
(data/data_generation.py:6-19: LayerNorm without mean subtraction, gamma = 1,
beta = 0), and the distance between the two is the approximations' own error.

Shapes: x (s, D) token rows; weights W[name] with wq, wk, wv, wo (D x D),
w1 (D x F), w2 (F x D), all applied as x @ W; head dimension 64; no biases.
"""
import numpy as np

from . import approx

HEAD_DIM = 64
LN_PRESCALE = np.sqrt(30.0)  # LN inputs are scaled so that mean(x^2) ~ 30


def random_weights(hidden, ffn, seed, std=0.02):
    rng = np.random.default_rng(seed)
    return {
        "wq": rng.normal(0, std, (hidden, hidden)),
        "wk": rng.normal(0, std, (hidden, hidden)),
        "wv": rng.normal(0, std, (hidden, hidden)),
        "wo": rng.normal(0, std, (hidden, hidden)),
        "w1": rng.normal(0, std, (hidden, ffn)),
        "w2": rng.normal(0, std, (ffn, hidden)),
    }


def _softmax_rows_approx(S, s):
    """rows of S (queries x keys): exp as (1 + x/128)^128 (ckks_evaluator.cpp:534-548),
    row sum, inverse of sum / s by the reference's Newton product (4 iterations,
    ckks_evaluator.cpp:487-509), times 1 / s."""
    e = approx.exp_approx(S)
    tot = np.sum(e, axis=-1, keepdims=True)
    inv = approx.inverse(tot / s) / s
    return e * inv


def _layer_norm_approx(r):
    """x * invert_sqrt(mean(x^2)) per row on x = r * sqrt(30) (layer_norm.cpp:6-33
    with the mean over the row's D features; invert_sqrt with 4 Newton and 2
    Goldschmidt steps)."""
    x = r * LN_PRESCALE
    m = np.sum(x * x, axis=-1, keepdims=True) / x.shape[-1]
    return x * approx.invert_sqrt(m, 4, 2)


def _attention(x, W, softmax):
    s, D = x.shape
    H = D // HEAD_DIM
    q = x @ W["wq"] / 8.0  # 1 / sqrt(head dim)
    k = x @ W["wk"]
    v = x @ W["wv"]
    o = np.zeros((s, D))
    for h in range(H):
        sl = slice(h * HEAD_DIM, (h + 1) * HEAD_DIM)
        S = q[:, sl] @ k[:, sl].T
        o[:, sl] = softmax(S) @ v[:, sl]
    return o


def layer_approx(x, W):
    """the encrypted program's arithmetic in float64"""
    s = x.shape[0]
    o = _attention(x, W, lambda S: _softmax_rows_approx(S, s))
    z = _layer_norm_approx(x + o @ W["wo"])
    g = approx.gelu_approx(z @ W["w1"])
    return _layer_norm_approx(z + g @ W["w2"])


def layer_exact(x, W):
    """exact FP64 layer: softmax, tanh GELU, x / rms(x) LayerNorm"""
    def sm(S):
        e = np.exp(S - S.max(axis=-1, keepdims=True))
        return e / e.sum(axis=-1, keepdims=True)

    def ln(r):
        return r / np.sqrt(np.mean(r * r, axis=-1, keepdims=True))

    o = _attention(x, W, sm)
    z = ln(x + o @ W["wo"])
    return ln(z + approx.gelu(z @ W["w1"]) @ W["w2"])
