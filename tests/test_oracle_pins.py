"""Pins for the CPU oracle (oracle/ckks_oracle.c).

Every oracle function is checked against something it does not itself compute:
SEAL's own known-answer tests (thirdparty/SEAL-4.1-bs/native/tests/seal/...),
brute-force O(N^2) evaluations with Python integers, exact CRT closed forms,
coefficient-domain automorphisms and high-precision (mpmath) DFTs.  A planted
mistake (dropped term, wrong sign, wrong index, transposed operand) in any
oracle routine fails at least one of these.
"""
import math
import random

import mpmath
import numpy as np
import pytest

import oracle as orc

Q_KAT = 0xFFFFFFFFFFC0001


def brv(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


# ------------------------------------------------------------- SEAL KATs
def test_is_prime_kat():
    # tests/seal/util/numth.cpp:99-114
    assert not orc.is_prime(0)
    assert orc.is_prime(2) and orc.is_prime(3) and orc.is_prime(5)
    assert not orc.is_prime(4) and not orc.is_prime(221) and not orc.is_prime(65536)
    assert orc.is_prime(65537) and orc.is_prime(59399) and orc.is_prime(72307)
    assert not orc.is_prime(72307 * 59399)
    assert orc.is_prime(36893488147419103)
    assert not orc.is_prime(36893488147419107)


def test_minimal_primitive_root_kat():
    # tests/seal/util/numth.cpp:201-222
    assert orc.minimal_primitive_root(2, 11) == 10
    assert orc.minimal_primitive_root(2, 29) == 28
    assert orc.minimal_primitive_root(4, 29) == 12
    assert orc.minimal_primitive_root(2, 1234565441) == 1234565440
    assert orc.minimal_primitive_root(8, 1234565441) == 249725733


def test_minimal_primitive_root_bruteforce():
    # exhaustive: the smallest r with r^(d/2) = -1 mod q
    for q, d in [(17, 8), (97, 16), (193, 32), (257, 64), (7681, 512)]:
        best = min(r for r in range(2, q) if pow(r, d // 2, q) == q - 1)
        assert orc.minimal_primitive_root(d, q) == best


def test_coeff_modulus_create_kat():
    # tests/seal/modulus.cpp:228-260
    assert orc.coeff_modulus_create(2, [3]) == [5]
    assert orc.coeff_modulus_create(2, [3, 4]) == [5, 13]
    assert orc.coeff_modulus_create(2, [3, 5, 4, 5]) == [5, 17, 13, 29]
    cm = orc.coeff_modulus_create(32, [30, 40, 30, 30, 40])
    assert [q.bit_length() for q in cm] == [30, 40, 30, 30, 40]
    assert all(q % 64 == 1 for q in cm)


def test_coeff_modulus_matches_bruteforce_scan():
    # first occurrence gets the smallest of the primes found (modulus.cpp:177-183)
    n, bits = 8192, [60, 40, 60]
    cm = orc.coeff_modulus_create(n, bits)
    f = 2 * n

    def scan(b, cnt):
        v = ((1 << b) - 1) // f * f + 1
        out = []
        while len(out) < cnt:
            if all(v % p for p in range(3, 2000, 2)) and pow(3, v - 1, v) == 1 and pow(5, v - 1, v) == 1:
                out.append(v)
            v -= f
        return out

    p60 = scan(60, 2)
    assert cm == [p60[1], scan(40, 1)[0], p60[0]]
    assert cm == [1152921504606748673, 1099511480321, 1152921504606830593]


def test_ntt_tables_kat():
    # tests/seal/util/ntt.cpp:53-76
    rp, irp = orc.ntt_tables(1, Q_KAT)
    assert int(rp[0]) == 1 and int(rp[1]) == 288794978602139552
    assert int(irp[1]) == pow(288794978602139552, -1, Q_KAT)
    rp, _ = orc.ntt_tables(2, Q_KAT)
    assert [int(x) for x in rp] == [1, 288794978602139552, 178930308976060547, 748001537669050592]


def test_ntt_kat():
    # tests/seal/util/ntt.cpp:78-97
    assert list(orc.ntt_forward([0, 0], 1, Q_KAT)) == [0, 0]
    assert list(orc.ntt_forward([1, 0], 1, Q_KAT)) == [1, 1]
    assert list(orc.ntt_forward([1, 1], 1, Q_KAT)) == [288794978602139553, 864126526004445282]


# ------------------------------------------------------------- NTT closed form
@pytest.mark.parametrize("logn,bits", [(3, 30), (5, 40), (7, 58), (8, 60)])
def test_ntt_bruteforce(logn, bits):
    n = 1 << logn
    q = orc.coeff_modulus_create(n, [bits])[0]
    psi = orc.minimal_primitive_root(2 * n, q)
    rng = random.Random(logn)
    a = [rng.randrange(q) for _ in range(n)]
    out = orc.ntt_forward(a, logn, q)
    for i in range(n):
        w = pow(psi, 2 * brv(i, logn) + 1, q)
        expect = sum(a[j] * pow(w, j, q) for j in range(n)) % q
        assert int(out[i]) == expect
    back = orc.ntt_inverse(out, logn, q)
    assert [int(x) for x in back] == a


def test_ntt_inverse_roundtrip_large():
    logn = 12
    q = orc.coeff_modulus_create(1 << logn, [58])[0]
    rng = np.random.default_rng(1)
    a = rng.integers(0, q, size=1 << logn, dtype=np.uint64)
    assert np.array_equal(orc.ntt_inverse(orc.ntt_forward(a, logn, q), logn, q), a)


# ------------------------------------------------------------- Galois
def test_galois_kat():
    # tests/seal/util/galois.cpp:95-113 (explicit element 3: generator independent)
    assert list(orc.apply_galois_ntt(np.arange(8), 3, 3)) == [4, 5, 7, 6, 1, 0, 2, 3]
    assert list(orc.apply_galois_coeff(np.arange(8), 3, 3, 17)) == [0, 14, 6, 1, 13, 7, 2, 12]
    # The bs fork sets GaloisTool::generator_ = 5 (util/galois.h:169), so the
    # upstream EltFromStep KAT (galois.cpp:36-48, generator 3) is stale; the same
    # table with generator 5 (m = 16): step k -> 5^k, step -k -> 5^(n/2 - k).
    steps = [0, 1, -3, 2, -2, 3, -1]
    assert [orc.elt_from_step(3, s) for s in steps] == [15, 5, 5, 9, 9, 13, 13]


@pytest.mark.parametrize("logn", [4, 8])
def test_rotation_invariant_pins_generator(logn):
    """decode(galois(encode(v), elt_from_step(k))) == v rotated left by k."""
    n = 1 << logn
    mods = orc.coeff_modulus_create(n, [60, 50])
    rng = np.random.default_rng(11)
    vals = rng.uniform(-1, 1, n // 2)
    enc = orc.Encoder(logn)
    scale = 2.0 ** 40
    pt = enc.encode(vals, scale, mods)
    for k in [1, 3, -2, n // 4 - 1]:
        elt = orc.elt_from_step(logn, k)
        rot = np.array([orc.apply_galois_ntt(pt[i], logn, elt) for i in range(len(mods))])
        dec = enc.decode(rot, mods, scale)
        assert np.max(np.abs(dec - np.roll(vals, -k))) < 1e-7
    # conjugation element 2n-1
    elt = orc.elt_from_step(logn, 0)
    cv = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    pt = enc.encode(cv, scale, mods)
    rot = np.array([orc.apply_galois_ntt(pt[i], logn, elt) for i in range(len(mods))])
    dec = enc.decode(rot, mods, scale, complex_out=True)
    assert np.max(np.abs(dec - np.conj(cv))) < 1e-7


@pytest.mark.parametrize("logn", [4, 7, 10])
def test_galois_ntt_commutes_with_automorphism(logn):
    n = 1 << logn
    q = orc.coeff_modulus_create(n, [50])[0]
    rng = np.random.default_rng(logn)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    for step in [1, -1, 3, n // 4 - 1, 0]:
        elt = orc.elt_from_step(logn, step)
        lhs = orc.ntt_forward(orc.apply_galois_coeff(a, logn, elt, q), logn, q)
        rhs = orc.apply_galois_ntt(orc.ntt_forward(a, logn, q), logn, elt)
        assert np.array_equal(lhs, rhs)


def test_galois_coeff_bruteforce():
    # x -> x^elt applied to a(x) in Z_q[x]/(x^n+1), computed term by term
    logn, q, elt = 4, 97, 5
    n = 1 << logn
    a = list(range(1, n + 1))
    out = [0] * n
    for i, c in enumerate(a):
        e = (i * elt) % (2 * n)
        if e >= n:
            out[e - n] = (out[e - n] - c) % q
        else:
            out[e] = (out[e] + c) % q
    assert list(orc.apply_galois_coeff(np.array(a), logn, elt, q)) == out


# ------------------------------------------------------------- rescale
def test_rescale_kat():
    # tests/seal/util/rns.cpp:1013-1075 (moduli {53, 13}, n = 2)
    mods = [53, 13]

    def run(c0, c1):
        poly = np.array([orc.ntt_forward(c0, 1, 53), orc.ntt_forward(c1, 1, 13)])
        out = orc.rescale_poly_ntt(poly, mods, 1)
        return [int(x) for x in orc.ntt_inverse(out[0], 1, 53)]

    assert run([0, 0], [0, 0]) == [0, 0]
    assert run([1, 2], [1, 2]) == [0, 0]
    r = run([4, 12], [4, 12])
    assert (53 + 1 - r[0]) % 53 <= 1 and (53 + 2 - r[1]) % 53 <= 1
    r = run([25, 35], [12, 9])
    assert (53 + 2 - r[0]) % 53 <= 1 and (53 + 3 - r[1]) % 53 <= 1


def crt(residues, mods):
    Q = math.prod(mods)
    x = 0
    for r, q in zip(residues, mods):
        qh = Q // q
        x += int(r) * qh * pow(qh, -1, q)
    return x % Q


@pytest.mark.parametrize("bits", [[40, 40, 58], [60, 40, 40, 60], [58, 40, 50, 45, 58]])
def test_rescale_closed_form(bits):
    logn = 5
    n = 1 << logn
    mods = orc.coeff_modulus_create(n, bits)
    l = len(mods)
    rng = np.random.default_rng(len(bits))
    coeff = np.array([rng.integers(0, q, size=n, dtype=np.uint64) for q in mods])
    ntt = np.array([orc.ntt_forward(coeff[i], logn, mods[i]) for i in range(l)])
    out = orc.rescale_poly_ntt(ntt, mods, logn)
    ql = mods[-1]
    for i in range(l - 1):
        got = orc.ntt_inverse(out[i], logn, mods[i])
        for c in range(n):
            X = crt([coeff[j][c] for j in range(l)], mods)
            expect = ((X + (ql - 1) // 2) // ql) % mods[i]  # round(X / q_last)
            assert int(got[c]) == expect


# ------------------------------------------------------------- key switch
def negacyclic_mul(a, b, q):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                out[k] = (out[k] + a[i] * b[j]) % q
            else:
                out[k - n] = (out[k - n] - a[i] * b[j]) % q
    return out


def test_switch_key_closed_form():
    """ModUp (exact lift of each digit's [0,q_j) representative) + inner product
    in the coefficient domain (schoolbook negacyclic products) + ModDown as
    round(X/p) of the CRT value over q_i*p -- an independent formulation of
    evaluator.cpp:2724-3021."""
    logn = 4
    n = 1 << logn
    kmods = orc.coeff_modulus_create(n, [40, 45, 50, 60])  # q0,q1,q2 | p
    L = len(kmods) - 1
    l = 2  # ciphertext at a lower level: digits q0,q1
    p = kmods[-1]
    rng = random.Random(7)
    key_coeff = [[[[rng.randrange(kmods[i]) for _ in range(n)] for i in range(L + 1)] for _ in range(2)]
                 for _ in range(L)]
    key = np.array([[[orc.ntt_forward(key_coeff[j][k][i], logn, kmods[i]) for i in range(L + 1)]
                     for k in range(2)] for j in range(L)], dtype=np.uint64)
    tcoef = [[rng.randrange(kmods[j]) for _ in range(n)] for j in range(l)]
    target = np.array([orc.ntt_forward(tcoef[j], logn, kmods[j]) for j in range(l)])
    c_coef = [[[rng.randrange(kmods[i]) for _ in range(n)] for i in range(l)] for _ in range(2)]
    ct = np.array([[orc.ntt_forward(c_coef[k][i], logn, kmods[i]) for i in range(l)] for k in range(2)])
    got = orc.switch_key(ct, target, kmods, key, logn)

    outs = list(range(l)) + [L]
    for k in range(2):
        acc = {}
        for I in outs:
            qI = kmods[I]
            s = [0] * n
            for J in range(l):
                prod = negacyclic_mul([v % qI for v in tcoef[J]], key_coeff[J][k][I], qI)
                s = [(x + y) % qI for x, y in zip(s, prod)]
            acc[I] = s
        for i in range(l):
            qi = kmods[i]
            res = orc.ntt_inverse(got[k][i], logn, qi)
            for c in range(n):
                X = crt([acc[i][c], acc[L][c]], [qi, p])
                expect = (c_coef[k][i][c] + (X + (p - 1) // 2) // p) % qi
                assert int(res[c]) == expect


def test_switch_key_decrypts():
    """Keys generated with s; switching c2 under s^2 gives (d0,d1) with
    d0 + d1*s = c2*s^2 + small (the relinearization invariant)."""
    logn = 6
    n = 1 << logn
    kmods = orc.coeff_modulus_create(n, [50, 40, 40, 58])
    L = 3
    P = kmods[-1]
    rng = np.random.default_rng(3)
    s = rng.integers(-1, 2, size=n)
    s_ntt = [orc.ntt_forward(s % q, logn, q) for q in kmods]
    s2 = [orc.dyadic_mul(sn[None, None, :], sn[None, :], [q], logn)[0, 0] for sn, q in zip(s_ntt, kmods)]
    key = np.zeros((L, 2, L + 1, n), dtype=np.uint64)
    for j in range(L):
        e = rng.integers(-3, 4, size=n)
        for i, q in enumerate(kmods):
            a = rng.integers(0, q, size=n, dtype=np.uint64)
            e_ntt = orc.ntt_forward(e % q, logn, q)
            b = [(-(int(a[c]) * int(s_ntt[i][c])) + int(e_ntt[c])) % q for c in range(n)]
            if i == j:
                b = [(b[c] + (P % q) * int(s2[i][c])) % q for c in range(n)]
            key[j, 0, i] = b
            key[j, 1, i] = a
    l = 3
    mods = kmods[:l]
    target = np.array([rng.integers(0, q, size=n, dtype=np.uint64) for q in mods])
    ct = np.zeros((2, l, n), dtype=np.uint64)
    d = orc.switch_key(ct, target, kmods, key, logn)
    for i, q in enumerate(mods):
        lhs = [(int(d[0][i][c]) + int(d[1][i][c]) * int(s_ntt[i][c])) % q for c in range(n)]
        rhs = [int(target[i][c]) * int(s2[i][c]) % q for c in range(n)]
        diff = orc.ntt_inverse(np.array([(x - y) % q for x, y in zip(lhs, rhs)], dtype=np.uint64), logn, q)
        cent = [int(v) if v < q // 2 else int(v) - q for v in diff]
        assert max(abs(v) for v in cent) < 4 * n


# ------------------------------------------------------------- dyadic
def test_tensor_closed_form():
    logn = 3
    n = 1 << logn
    mods = orc.coeff_modulus_create(n, [40, 50])
    rng = np.random.default_rng(2)
    a = np.array([[rng.integers(0, q, size=n, dtype=np.uint64) for q in mods] for _ in range(2)])
    b = np.array([[rng.integers(0, q, size=n, dtype=np.uint64) for q in mods] for _ in range(2)])
    out = orc.tensor(a, b, mods, logn)
    for i, q in enumerate(mods):
        for c in range(n):
            a0, a1, b0, b1 = (int(a[0][i][c]), int(a[1][i][c]), int(b[0][i][c]), int(b[1][i][c]))
            assert [int(out[k][i][c]) for k in range(3)] == [a0 * b0 % q, (a0 * b1 + a1 * b0) % q, a1 * b1 % q]


def test_crt_compose_closed_form():
    mods = orc.coeff_modulus_create(16, [60, 40, 60, 50])
    rng = np.random.default_rng(5)
    limbs = np.array([rng.integers(0, q, size=16, dtype=np.uint64) for q in mods])
    out = orc.crt_compose(limbs, mods)
    for c in range(16):
        X = crt([limbs[i][c] for i in range(len(mods))], mods)
        got = sum(int(w) << (64 * k) for k, w in enumerate(out[c]))
        assert got == X


# ------------------------------------------------------------- encoder
def exact_coeffs(values, logn, scale):
    """m_j = (1/N) sum_k v_k zeta_k^{-j} over the 2N-th roots zeta_k = w^{5^i}, conj at w^{-5^i}."""
    n = 1 << logn
    slots = n // 2
    mpmath.mp.prec = 200
    w = mpmath.exp(2j * mpmath.pi / (2 * n))
    pts = []
    e = 1
    for i in range(slots):
        pts.append((e, mpmath.mpc(values[i])))
        pts.append(((2 * n - e) % (2 * n), mpmath.conj(mpmath.mpc(values[i]))))
        e = (e * 5) % (2 * n)
    out = []
    for j in range(n):
        s = mpmath.mpc(0)
        for ex, v in pts:
            s += v * w ** (-(ex * j) % (2 * n))
        out.append(mpmath.re(s) * scale / n)
    return out


@pytest.mark.parametrize("logn,scale", [(5, 2.0 ** 30), (6, 2.0 ** 40)])
def test_encode_matches_exact_dft(logn, scale):
    n = 1 << logn
    mods = orc.coeff_modulus_create(n, [60, 40])
    rng = np.random.default_rng(logn)
    vals = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    enc = orc.Encoder(logn)
    pt = enc.encode(vals, scale, mods)
    coeff = orc.ntt_inverse(pt[0], logn, mods[0])
    exact = exact_coeffs(vals, logn, scale)
    q = mods[0]
    checked = 0
    for j in range(n):
        ex = exact[j]
        frac = float(ex - mpmath.floor(ex))
        if abs(frac - 0.5) < 1e-3:  # rounding tie window: parity unpinned here
            continue
        want = int(mpmath.nint(ex)) % q
        assert int(coeff[j]) == want
        checked += 1
    assert checked >= n - 2


def test_encode_coeffs_and_exact_residues():
    """The unrounded coefficients reduce (exact integers, halves away from zero)
    to the oracle encoder's residues through the 64- and 128-bit decompositions,
    and track the high-precision DFT to double accuracy."""
    logn = 6
    n = 1 << logn
    enc = orc.Encoder(logn)
    rng = np.random.default_rng(5)
    vals = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    for scale, big in ((2.0 ** 40, 1.0), (2.0 ** 60, 2.0 ** 20)):
        mods = orc.coeff_modulus_create(n, [60, 60, 60])
        v = vals * big
        c = enc.encode_coeffs(v, scale)
        exact = exact_coeffs(v, logn, scale)
        assert all(abs(float(e) - x) <= 1e-9 * max(1.0, abs(x)) for e, x in zip(exact, c))
        r = orc.Encoder.residues_exact(c, mods)
        got = np.stack([orc.ntt_forward(r[j], logn, mods[j]) for j in range(len(mods))])
        assert np.array_equal(got, enc.encode(v, scale, mods))
    # halves round away from zero, like std::round
    q = mods[0]
    r = orc.Encoder.residues_exact(np.array([2.5, -2.5, 0.5, -0.5, 2.0 ** 70 + 2.0 ** 18]), [q])[0]
    assert [int(x) for x in r] == [3, q - 3, 1, q - 1, (2 ** 70 + 2 ** 18) % q]


def test_encode_decode_roundtrip_and_sparse():
    logn = 10
    n = 1 << logn
    mods = orc.coeff_modulus_create(n, [60, 40, 60])[:2]
    rng = np.random.default_rng(9)
    vals = rng.uniform(-1, 1, n // 2)
    enc = orc.Encoder(logn)
    scale = 2.0 ** 40
    pt = enc.encode(vals, scale, mods)
    dec = enc.decode(pt, mods, scale)
    assert np.max(np.abs(dec - vals)) < 1e-7
    # constant encode: every slot decodes to the constant
    pc = enc.encode_const(-0.3125, scale, mods)
    assert np.max(np.abs(enc.decode(pc, mods, scale) + 0.3125)) < 1e-9
    # sparse decode of a periodically replicated vector (ckks.h:664-675)
    sp = 64
    base = rng.uniform(-1, 1, sp)
    rep = np.array([base[i % sp] for i in range(n // 2)])
    pt = enc.encode(rep, scale, mods)
    dsp = enc.decode(pt, mods, scale, sparse_slots=sp)
    assert np.max(np.abs(dsp - base)) < 1e-7
