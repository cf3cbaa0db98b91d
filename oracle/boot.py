"""Oracle of the CKKS bootstrapping pipeline (test infrastructure only): closed
forms of the slot-domain maps and a float64 model of the whole pipeline.

Sources: reference cuda/src/bootstrapping/Bootstrapper.cu (modraise_inplace
:2800-2860, bootstrap_sparse_3 :2987-3091, bootstrap_full_3 :3093-3113,
genorigcoeff :515-595), ModularReducer.cu:53-73.

Slot convention (SEAL's encoder, ckks.h:44-77 matrix_reps_index_map with
generator 5): a polynomial whose nonzero coefficients sit at multiples of
gap = N / (2 n) with values a (lower n) and b (upper n) decodes to the
n-periodic slot vector U0 (a + i b), U0[j][k] = w^(5^j k mod 4n),
w = exp(2 pi i / 4n).  `slot_matrix` states it; a test pins it against the
oracle encoder.
"""
import numpy as np


def bit_reverse(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def rot_group(logn):
    n = 1 << logn
    m = 4 * n
    e = np.empty(n, dtype=np.int64)
    v = 1
    for j in range(n):
        e[j] = v
        v = (v * 5) % m
    return e


def slot_matrix(logn):
    """U0[j][k] = w^(5^j k mod 4n), w = exp(2 pi i / 4n) (dense; small n)."""
    n = 1 << logn
    m = 4 * n
    expo = (rot_group(logn)[:, None] * np.arange(n, dtype=np.int64)[None, :]) % m
    return np.exp(2j * np.pi * expo / m)


def slot_map(z, logn):
    """U0 z in O(n log n): z(X) evaluated at every 4n-th root of unity, then
    read at the exponents 5^j mod 4n."""
    n = 1 << logn
    m = 4 * n
    pad = np.zeros(m, dtype=complex)
    pad[:n] = z
    vals = np.fft.ifft(pad) * m  # vals[t] = sum_k z_k exp(2 pi i t k / m)
    return vals[rot_group(logn)]


def brv_perm(logn):
    return np.array([bit_reverse(i, logn) for i in range(1 << logn)])


def apply_level(level, x):
    """y = sum_(j,i) rot(D_(j,i) * rot(x, i*stride), g*j*stride) on len(x) slots,
    rot(v, r)[s] = v[s + r] (left rotation, as rotate_vector)."""
    st, g = level["stride"], level["g"]
    y = np.zeros(len(x), dtype=complex)
    giants = {}
    for (j, i), d in level["diags"].items():
        giants.setdefault(j, np.zeros(len(x), dtype=complex))
        giants[j] += d * np.roll(x, -i * st)
    for j, acc in giants.items():
        y += np.roll(acc, -j * g * st)
    return y


def mod_raise(c_q0, moduli):
    """kernel_modraise_inplace (Bootstrapper.cu:2800-2811): coefficient x of q0
    lifted to every prime: x mod q, plus (q - q0 mod q) when x > q0 >> 1."""
    q0 = int(moduli[0])
    x = np.asarray(c_q0, dtype=object)
    out = []
    for q in moduli:
        q = int(q)
        v = np.array([(int(a) % q) for a in x], dtype=object)
        big = np.array([int(a) > (q0 >> 1) for a in x])
        minus = 0 if q == q0 else q - q0 % q
        v = np.where(big, (v + minus) % q, v)
        out.append(np.array(v, dtype=np.uint64))
    return np.array(out)


def eval_mod_fit(K=25, degree=59, double_angles=2, log_width=10):
    """Chebyshev coefficients (basis on [-1, 1]) of the least-squares fit of
    cos(2 pi (K u - 1/4) / 2^r) at 40 Chebyshev-distributed points in each
    window [(i - w) / K, (i + w) / K], w = 2^-log_width, |i| < K (the windows of
    the reference's RemezCos, cuda/src/bootstrapping/RemezCos.cuh:12-17)."""
    w = 2.0 ** -log_width
    per = 40
    t = np.cos(np.pi * (np.arange(per) + 0.5) / per)
    xs = np.concatenate([(i + w * t) / K for i in range(-K + 1, K)])
    V = np.polynomial.chebyshev.chebvander(xs, degree)
    f = np.cos(2 * np.pi * (K * xs - 0.25) / 2.0 ** double_angles)
    return np.linalg.lstsq(V, f, rcond=None)[0]


def eval_mod_model(u, K=25, degree=59, double_angles=2, log_width=10):
    """float64 EvalMod: p(u) then r double angles -> ~ sin(2 pi K u)."""
    c = eval_mod_fit(K, degree, double_angles, log_width)
    y = np.polynomial.chebyshev.chebval(u, c)
    for _ in range(double_angles):
        y = 2 * y * y - 1
    return y


def decode_sparse(coeffs, logN, logn):
    """Slots of a polynomial with coefficients at multiples of N/2n (one period
    of n values): U0 (a + i b)."""
    n = 1 << logn
    gap = (1 << logN) // (2 * n)
    a = np.asarray(coeffs, dtype=np.float64)[::gap]
    return slot_map(a[:n] + 1j * a[n:], logn)


def bootstrap_model(m, q0, scale, logN, logn):
    """Slots the bootstrapped ciphertext decodes to in the ideal pipeline, given
    the centred coefficients m (|m| < q0/2) of the lowest-level plaintext:
    EvalMod yields sin(2 pi m / q0) / 2 pi per coefficient, SlotToCoeff maps it
    back through the slot matrix, times q0 / scale."""
    f = np.sin(2 * np.pi * np.asarray(m, dtype=np.float64) / q0) / (2 * np.pi)
    return (q0 / scale) * decode_sparse(f, logN, logn)
