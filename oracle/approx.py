"""FP64 plaintext restatement of NEXUS's polynomial approximations (test
infrastructure only).  Each function evaluates, in float64 on plaintext
vectors, the same sequence of polynomial steps the encrypted program runs, so
the decrypted result of the engine must agree with it up to CKKS noise; the
distance of these functions from the exact GELU / LayerNorm / softmax is the
reference's own approximation error.

Sources: src/ckks_evaluator.cpp:181-548, src/gelu.cpp:5-111,
src/layer_norm.cpp:6-33, src/softmax.cpp:6-42; FP64 definitions of the exact
functions: data/data_generation.py:6-19.
"""
import numpy as np

F4 = np.array([0, 315, 0, -420, 0, 378, 0, -180, 0, 35], dtype=np.float64) / (1 << 7)
G4 = np.array([0, 5850, 0, -34974, 0, 97015, 0, -113492, 0, 46623], dtype=np.float64) / (1 << 10)
GELU_A = [2.25775755e-04, 0.5, 3.96880960e-01, -6.37042698e-02, 8.38841647e-03, -7.17830961e-04,
          3.49617829e-05, -7.26059653e-07]


def odd9(a, x):
    return a[1] * x + a[3] * x ** 3 + a[5] * x ** 5 + a[7] * x ** 7 + a[9] * x ** 9


def sgn_eval(x, d_g, d_f, sgn_factor=0.5):
    for i in range(d_g):
        x = odd9(G4 * sgn_factor if i == d_g - 1 else G4, x)
    for i in range(d_f):
        x = odd9(F4 * sgn_factor if i == d_f - 1 else F4, x)
    return x


def gelu_approx(x):
    b0 = sgn_eval((x + 3.5) / 8.5, 2, 2)
    b1 = sgn_eval((x - 3.5) / 8.5, 2, 2)
    A = GELU_A
    ax = A[0] + A[1] * x + A[2] * x ** 2 + A[3] * x ** 4 + A[4] * x ** 6 + A[5] * x ** 8 + A[6] * x ** 10 \
        + A[7] * x ** 12
    return ax * (b0 - b1) + x * (b1 + 0.5)


def invert_sqrt(v, d_newt=5, d_gold=2):
    y = -1.29054537e-04 * v + 1.29054537e-01
    for _ in range(d_newt):
        y = 1.5 * y - 0.5 * v * y ** 3
    x = v * y
    h = 0.5 * y
    for _ in range(d_gold):
        r = 0.5 - x * h
        x = x + x * r
        h = h + h * r
    return 2.0 * h


def inverse(x, iters=4):
    y = 1.0 - x
    res = 2.0 - x
    for _ in range(iters):
        y = y * y
        res = res * (1.0 + y)
    return res


def exp_approx(x):
    return (1.0 + x / 128.0) ** 128


def layer_norm_approx(x):
    """rows packed as in src/layer_norm.cpp with len = 1024 (768 live values)."""
    s = np.sum(x * x) / 768.0
    return x * invert_sqrt(s, 4, 2)


def softmax_approx(x):
    e = exp_approx(x)
    s = np.sum(e)
    return e * (inverse(s * 0.01) * 0.01)


# exact FP64 targets (data/data_generation.py:6-19)
def argmax_approx(x, length):
    """src/argmax.cpp:9-64 on a slot vector x (the first `length` slots hold the
    inputs, the rest zero): duplicate, log2(length) rounds of
    (a + b)/2 + (a - b) sgn(a - b)/2 with b = a rotated left by 2^i, then
    sgn(x - max) + 1."""
    x = np.asarray(x, dtype=np.float64)
    x = x + np.roll(x, length)
    a = x.copy()
    steps = int(np.log2(length))
    for i in range(steps):
        b = np.roll(a, -(1 << i))
        a = (a + b) * 0.5 + (a - b) * sgn_eval(a - b, 2, 2)
    return sgn_eval(x - a, 2, 2, 1.0) + 1.0


def gelu(x):
    return x * 0.5 * (1 + np.tanh(np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)))


def layernorm(x):
    return np.sqrt(len(x)) * x / np.sqrt(np.sum(x ** 2))


def softmax(x):
    return np.exp(x) / np.sum(np.exp(x))
