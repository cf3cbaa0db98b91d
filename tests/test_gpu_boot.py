"""CKKS bootstrapping on the GPU.

ModRaise is checked bit-exactly against the oracle's restatement of
kernel_modraise_inplace (oracle/boot.py).  CoeffToSlot, EvalMod and
SlotToCoeff are checked stage by stage against the float64 closed forms of
oracle/boot.py, and the whole pipeline against the ideal model
sin(2 pi m / q0) / 2 pi mapped through the slot matrix (m = the decrypted
lowest-level plaintext) and against the input message.  The EvalMod
polynomial itself (the reference's NTL minimax generator is absent) is pinned
by its error against sin on the near-integer windows, not by coefficients.

Sizes: N = 2^12 with 2^10 (sparse) and 2^11 (dense) slots on a 19-prime chain,
and the reference's bootstrapping configuration (cuda/src/bootstrapping.cu:22-60:
N = 2^15, 2^13 slots, {51, 46 x 16, 51 x 14, 51}, Hamming weight 192,
scale 2^46, K = 25).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402
from oracle import boot as ob  # noqa: E402

SCALE = 2.0 ** 46
K = 25
SMALL_CHAIN = [51] + [46] * 2 + [51] * 15 + [51]


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


class Env:
    def __init__(self, logN, logn, chain, h):
        _note("bootstrap keys logN=%d logn=%d L=%d h=%d" % (logN, logn, len(chain) - 1, h))
        self.logN, self.logn = logN, logn
        self.ctx = nx.Context(logN, chain, hamming_weight=h)
        self.sk = nx.SecretKey(self.ctx, seed=1000 + logN * 10 + logn)
        self.rk = nx.KSwitchKeys(self.ctx).generate_relin(self.sk, seed=3)
        self.steps = nx.boot_plan_steps(logN - 1, logn, K)
        self.gk = nx.KSwitchKeys(self.ctx).generate_galois(
            self.sk, [self.ctx.elt_from_step(s) for s in self.steps], seed=4)
        self.ev = nx.CKKSEvaluator(self.ctx, self.rk, self.gk, SCALE, self.sk)
        self.bt = nx.Bootstrapper(self.ev, log_slots=logn, K=K)
        self.q0 = float(self.ctx.moduli[0])
        self.L = self.ctx.L

    def enc(self, v, scale, chain_index=1):
        return nx.encrypt(self.ctx, self.sk, nx.encode(self.ctx, v, scale, chain_index), seed=9)

    def dec(self, ct, complex_out=True):
        return nx.decode(self.ctx, nx.decrypt(self.ctx, self.sk, ct), complex_out=complex_out)

    def lowest_plain_coeffs(self, ct):
        """centred coefficients of the decryption of a 1-limb ciphertext"""
        pt = nx.decrypt(self.ctx, self.sk, ct).to_array()
        q0 = self.ctx.moduli[0]
        c = orc.ntt_inverse(pt[0].copy(), self.logN, q0).astype(object)
        return np.array([int(x) - q0 if int(x) > q0 // 2 else int(x) for x in c], dtype=np.float64)


@pytest.fixture(scope="module", params=[(12, 10), (12, 11)], ids=["sparse", "dense"])
def env(request):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logN, logn = request.param
    return Env(logN, logn, SMALL_CHAIN, 64)


def test_rotation_steps_and_depth(env):
    assert sorted(env.bt.rotation_steps()) == sorted(env.steps)
    assert env.bt.depth() == 3 + 6 + 2 + 3


def test_mod_raise_bit_exact(env):
    _note("modraise logN=%d" % env.logN)
    rng = np.random.default_rng(1)
    ct = env.enc(rng.uniform(-1, 1, 1 << env.logn), SCALE)
    ct = nx.mod_switch_to(ct, env.L)
    low = ct.to_array()
    raised = nx.mod_raise(ct)
    assert raised.chain_index == 1 and raised.scale == ct.scale
    got = raised.to_array()
    mods = env.ctx.moduli[: env.L]
    q0 = mods[0]
    for p in range(2):
        coef = orc.ntt_inverse(low[p, 0].copy(), env.logN, q0)
        lifted = ob.mod_raise(coef, mods)
        for j, q in enumerate(mods):
            exp = orc.ntt_forward(lifted[j].copy(), env.logN, q)
            assert np.array_equal(got[p, j], exp), "poly %d limb %d" % (p, j)


def test_coeff_to_slot_stage(env):
    _note("coeff_to_slot logN=%d logn=%d" % (env.logN, env.logn))
    n, Nh = 1 << env.logn, 1 << (env.logN - 1)
    c = Nh // n
    rng = np.random.default_rng(2)
    alpha = rng.uniform(-20, 20, 2 * n)
    z = alpha[:n] + 1j * alpha[n:]
    w = np.tile(c * ob.slot_map(z, env.logn), Nh // n)
    ct = env.enc(w, env.q0)
    u_re, u_im = env.bt.coeff_to_slot(ct)
    assert u_re.chain_index == 4
    br = ob.brv_perm(env.logn)
    got = env.dec(u_re)
    if u_im is None:
        s = np.arange(Nh)
        exp = alpha[br[s % n] + n * ((s % (2 * n)) >= n)] / K
        err = np.max(np.abs(got - exp))
    else:
        err = max(np.max(np.abs(got - z.real[br] / K)), np.max(np.abs(env.dec(u_im) - z.imag[br] / K)))
    print("coeff_to_slot max err %.3g" % err)
    assert err < 1e-6


def test_eval_mod_stage(env):
    _note("eval_mod logN=%d" % env.logN)
    Nh = 1 << (env.logN - 1)
    rng = np.random.default_rng(3)
    u = (rng.integers(-K + 1, K, Nh) + rng.uniform(-2.0 ** -8, 2.0 ** -8, Nh)) / K
    ct = env.enc(u, env.q0, chain_index=4)
    out = env.bt.eval_mod(ct)
    assert out.chain_index == 4 + 8
    got = env.dec(out)
    model = ob.eval_mod_model(u, K)
    err = np.max(np.abs(got - model))
    err_sin = np.max(np.abs(got - np.sin(2 * np.pi * K * u)))
    print("eval_mod max err vs model %.3g vs sin %.3g" % (err, err_sin))
    assert err < 1e-6 and err_sin < 1e-6


def test_slot_to_coeff_stage(env):
    _note("slot_to_coeff logN=%d logn=%d" % (env.logN, env.logn))
    n, Nh = 1 << env.logn, 1 << (env.logN - 1)
    rng = np.random.default_rng(4)
    a = rng.uniform(-0.5, 0.5, n)
    b = rng.uniform(-0.5, 0.5, n)
    br = ob.brv_perm(env.logn)
    ci = 4 + 8
    if n < Nh:
        f = np.tile(np.concatenate([a[br], b[br]]), Nh // (2 * n))
        out = env.bt.slot_to_coeff(env.enc(f, env.q0, ci), None, SCALE)
    else:
        out = env.bt.slot_to_coeff(env.enc(a[br], env.q0, ci), env.enc(b[br], env.q0, ci), SCALE)
    assert out.chain_index == ci + 3 and out.scale == SCALE
    exp = np.tile(ob.slot_map(a + 1j * b, env.logn), Nh // n) / (2 * np.pi) * (env.q0 / SCALE)
    err = np.max(np.abs(env.dec(out) - exp))
    print("slot_to_coeff max err %.3g" % err)
    assert err < 1e-5


def _bootstrap_check(env, seed, tol_model, tol_msg):
    n, Nh = 1 << env.logn, 1 << (env.logN - 1)
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, n)
    ct = env.enc(np.tile(v, Nh // n), SCALE)
    ct = nx.mod_switch_to(ct, env.L)
    m = env.lowest_plain_coeffs(ct)
    before = env.dec(ct)[:n]
    out = env.bt.bootstrap(ct)
    assert out.chain_index == 1 + env.bt.depth() and out.scale == SCALE
    after = env.dec(out)
    model = ob.bootstrap_model(m, env.q0, SCALE, env.logN, env.logn)
    e_model = np.max(np.abs(after[:n] - model))
    e_before = np.abs(after[:n] - before)
    period = np.max(np.abs(after - np.tile(after[:n], Nh // n)))
    print("bootstrap logN=%d logn=%d: max|after-model| %.3g  max|after-before| %.3g  mean %.3g  period %.3g"
          % (env.logN, env.logn, e_model, e_before.max(), e_before.mean(), period))
    assert e_model < tol_model
    assert e_before.max() < tol_msg
    assert period < tol_model
    return out


def test_bootstrap(env):
    _note("bootstrap logN=%d logn=%d" % (env.logN, env.logn))
    _bootstrap_check(env, 5, 1e-4, 1e-3)


def test_bootstrap_many_matches_single(env):
    """three inputs with different factors through one lockstep EvalMod: each
    output equals its own single bootstrap word for word"""
    _note("bootstrap_many logN=%d logn=%d" % (env.logN, env.logn))
    rng = np.random.default_rng(77)
    Nh = env.ctx.slot_count
    cts = [nx.mod_switch_to(env.enc(rng.uniform(-1, 1, Nh), SCALE), env.L) for _ in range(3)]
    factors = [1.0, 0.5, 2.0]
    many = env.bt.bootstrap_many(cts, factors)
    for c, f, m in zip(cts, factors, many):
        single = env.bt.bootstrap(c, f)
        assert np.array_equal(m.to_array(), single.to_array())
        assert m.scale == single.scale


def test_bootstrap_output_computes(env):
    """the refreshed ciphertext supports further multiplications"""
    _note("bootstrap then square logN=%d" % env.logN)
    out = _bootstrap_check(env, 6, 1e-4, 1e-3)
    n = 1 << env.logn
    before = env.dec(out)[:n].real
    sq = nx.rescale_to_next(nx.relinearize(nx.square(out), env.rk))
    got = env.dec(sq)[:n].real
    assert np.max(np.abs(got - before ** 2)) < 1e-4


def test_bootstrap_reference_config():
    """cuda/src/bootstrapping.cu:22-60 (N = 2^15, 2^13 sparse slots, 31 data primes)"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    chain = [51] + [46] * 16 + [51] * 14 + [51]
    env = Env(15, 13, chain, 192)
    assert len(env.steps) <= 42
    _note("bootstrap reference config")
    out = _bootstrap_check(env, 7, 1e-4, 1e-3)
    assert out.chain_index == 15  # 14 levels consumed: 17 of 31 limbs remain


def test_argmax_reference_config():
    """ArgmaxEvaluator on the reference's argmax_test (src/main.cpp:181-303):
    N = 2^15, 2^13 sparse slots, {51, 46 x 17, 51 x 14, 58}, h = 192,
    scale 2^46, input data/input/argmax_input_8.txt at 14 levels below the top;
    checked against the FP64 model of the same program and the reference's
    calibration file."""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    from oracle import approx
    from tests import fixtures

    logN, logn = 15, 13
    chain = [51] + [46] * 17 + [51] * 14 + [58]
    _note("argmax keys logN=15 L=32")
    ctx = nx.Context(logN, chain, hamming_weight=192)
    sk = nx.SecretKey(ctx, seed=77)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=3)
    steps = sorted(set(nx.boot_plan_steps(logN - 1, logn, K)) | {-8, 1, 2, 4})
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=4)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    bt = nx.Bootstrapper(ev, log_slots=logn, K=K)
    assert bt.depth() == 14
    x, calib = fixtures.argmax_vectors()
    n = 1 << logn
    period = np.zeros(n)
    period[: len(x)] = x
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, np.tile(period, (1 << (logN - 1)) // n), SCALE), seed=5)
    ct = nx.mod_switch_to(ct, 1 + 14)
    _note("argmax len 8")
    out = bt.argmax(ct, len(x))
    got = nx.decode(ctx, nx.decrypt(ctx, sk, out))[: len(x)]
    model = approx.argmax_approx(period, len(x))[: len(x)]
    mae = np.mean(np.abs(got - calib))
    print("argmax: out", np.round(got, 4), "max|out-model| %.3g MAE %.3g" % (np.max(np.abs(got - model)), mae))
    assert np.max(np.abs(got - model)) < 1e-2
    assert mae <= np.mean(np.abs(model - calib)) + 1e-3


def test_bootstrap_mixed_matches_single():
    """dense (2^11 slots) and sparse (2^10 slots) inputs at three input levels in
    one batch through two bootstrappers of one context: every CoeffToSlot /
    SlotToCoeff level batched per bootstrapper, one EvalMod for all; each output
    equals its own single bootstrap word for word (one special prime)"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("bootstrap mixed dense+sparse logN=12")
    env = Env(12, 11, SMALL_CHAIN, 64)
    sparse_steps = nx.boot_plan_steps(11, 10, K)
    env.gk.generate_galois(env.sk, [env.ctx.elt_from_step(s) for s in sparse_steps if s not in env.steps], seed=5)
    bt_sparse = nx.Bootstrapper(env.ev, log_slots=10, K=K)
    rng = np.random.default_rng(78)
    Nh = env.ctx.slot_count
    cts, bts, msgs = [], [], []
    for i in range(7):
        dense = i % 3 != 2
        n = Nh if dense else Nh // 2
        v = rng.uniform(-1, 1, n)
        c = env.enc(np.tile(v, Nh // n), SCALE)
        cts.append(nx.mod_switch_to(c, env.L - (i % 3)))  # input levels 0, 1, 2
        bts.append(env.bt if dense else bt_sparse)
        msgs.append(v)
    factors = [1.0, 0.5, 2.0, 1.0, 1.0, 0.25, 1.0]
    got = nx.Bootstrapper.bootstrap_mixed(bts, cts, factors)
    for c, b, f, m, v in zip(cts, bts, factors, got, msgs):
        single = b.bootstrap(c, f)
        assert np.array_equal(m.to_array(), single.to_array())
        assert m.scale == single.scale
        err = np.max(np.abs(env.dec(m)[:len(v)].real - f * v))
        assert err < 1e-3, err
