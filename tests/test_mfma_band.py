"""The MFMA refine's guard band, checked on the CPU.

The library (hegj_gb_operands, the host build of the device code) gives the
split-BF16 operands A', B' and the guards w of each point.  The matrix core
computes s' = A'_q . B'_c: products of BF16 values are exact in FP32, but
the order and rounding of the FP32 additions inside v_mfma_f32_16x16x32_bf16
are not published.  This test therefore sums the exact products under
adversarial models -- sequential accumulation rounding to nearest, toward
zero, up or down on every addition; 32-product blocks summed exactly and
rounded once into the accumulator; natural, reversed and shuffled orders --
and checks every decision the kernel would take against the reference's FP64
chain (kernel.cu:347-355):
    reject  s' > T32   (T rounded up to FP32)      must imply acc > T
    accept  s' + w_q + w_c <= T (1 - 4 * 2^-53)     must imply acc <= T.
Pairs sit at acc = T +- a few ulps and spread over T (1 +- r), with large
norms (the centre far from the points) and K up to 384.

The guard is twice the error bound (`gb_coefficient`): the test asserts the
worst observed error leaves at least half of each guard unused, and that
shrinking the coefficient far enough (or to zero) makes the same models
take wrong decisions, so the check can fail."""
import math

import numpy as np
import pytest

U53 = 2.0 ** -53


def bf16(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f32_round(x, mode):
    """x (float64 array, exact value x + e with e from two_sum) rounded to FP32."""
    v, e = x
    f = v.astype(np.float32)
    fd = f.astype(np.float64)
    if mode == "rne":
        return f  # e only matters at exact ties, which the float64 sum rarely hits
    lower = (fd > v) | ((fd == v) & (e < 0))   # f above the exact value
    higher = (fd < v) | ((fd == v) & (e > 0))  # f below the exact value
    if mode == "ru":
        f = np.where(higher, np.nextafter(f, np.float32(np.inf)), f)
    elif mode == "rd":
        f = np.where(lower, np.nextafter(f, np.float32(-np.inf)), f)
    elif mode == "rz":
        f = np.where(lower & (v > 0), np.nextafter(f, np.float32(-np.inf)), f)
        f = np.where(higher & (v < 0), np.nextafter(f, np.float32(np.inf)), f)
    return f.astype(np.float32)


def two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def mfma_sum(prod, mode, order):
    """prod: (pairs, K) exact products (float64).  Sums over K into an FP32
    accumulator per `order`: "seq" (one rounding per addition) or "block"
    (32 products summed exactly, rounded once into the accumulator)."""
    pairs, K = prod.shape
    acc = np.zeros(pairs, np.float32)
    if order == "seq":
        for k in range(K):
            acc = f32_round(two_sum(acc.astype(np.float64), prod[:, k]), mode)
        return acc
    for k0 in range(0, K, 32):
        blk = np.array([math.fsum(r) for r in prod[:, k0:k0 + 32]])
        acc = f32_round(two_sum(acc.astype(np.float64), blk), mode)
    return acc


def near_threshold_pairs(oracle, rng, dim, T, m_ulp=150, m_spread=300, scale=0.5):
    """Queries around `scale` in every dim; partners at acc = T +- k ulps
    (bisection on the distance, then nextafter steps of the last
    coordinate) and at T (1 + r), |r| log-uniform in [1e-9, 3e-2]."""
    m = m_ulp + m_spread
    q = scale + 0.1 * rng.random((m, dim))
    u = rng.standard_normal((m, dim))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    target = np.full(m, T)
    r = 10 ** rng.uniform(-9, math.log10(3e-2), m_spread) * rng.choice([-1, 1], m_spread)
    target[m_ulp:] = T * (1 + r)
    lo, hi = np.zeros(m), np.full(m, 2 * math.sqrt(T))
    for _ in range(80):  # bisection of t with c = q + t u
        mid = (lo + hi) / 2
        a = oracle.chain(q, q + mid[:, None] * u)
        below = a <= target
        lo, hi = np.where(below, mid, lo), np.where(below, hi, mid)
    c = q + lo[:, None] * u
    steps = rng.integers(-6, 7, m_ulp)
    for i in range(m_ulp):  # walk the last coordinate a few ulps
        for _ in range(abs(int(steps[i]))):
            c[i, -1] = np.nextafter(c[i, -1], np.inf if steps[i] > 0 else -np.inf)
    return q, c


@pytest.mark.parametrize("dim", [18, 33, 90, 126])
def test_guard_band_decisions_are_sound(oracle, dim):
    import hegjoin_amd as H
    rng = np.random.default_rng(dim)
    centre = np.zeros(dim)  # points near 0.5: norms far above eps^2
    _, _, _, kp, _ = H.gb_operands(np.zeros((1, dim)), centre)
    assert kp >= 3 * (dim + 2) and kp % 32 == 0
    # eps so the guard is ~1e-3 of T: near-T pairs fall in the band, the
    # spread ones on both sides of it
    N = 0.3 ** 2 * dim
    A0 = H.gb_operands(np.zeros((1, dim)), centre)[4]
    eps = math.sqrt(2 * A0 * N * 1e3)
    T = oracle.threshold(eps)
    T32 = np.float32(T)
    if float(T32) < T:
        T32 = np.nextafter(T32, np.float32(np.inf))
    Tacc = T * (1.0 - 4.0 * U53)
    q, c = near_threshold_pairs(oracle, rng, dim, T)
    acc = oracle.chain(q, c)
    truth = acc <= T
    assert truth.any() and (~truth).any()

    def decisions(scale, mode, order, perm):
        Aq, _, wq, _, coef = H.gb_operands(q, centre, coef_scale=scale)
        _, Bc, wc, _, _ = H.gb_operands(c, centre, coef_scale=scale)
        prod = bf16(Aq) * bf16(Bc)  # exact: two 8-bit significands
        if perm is not None:
            prod = prod[:, perm]
        s = mfma_sum(prod, mode, order).astype(np.float64)
        reject = s > float(T32)
        accept = (s + wq.astype(np.float64) + wc.astype(np.float64)) <= Tacc
        return s, reject, accept, wq.astype(np.float64), wc.astype(np.float64)

    perms = [None, np.arange(kp)[::-1], rng.permutation(kp)]
    worst = 0.0
    for mode in ("rne", "rz", "ru", "rd"):
        for order in ("seq", "block"):
            for perm in perms:
                s, reject, accept, wq, wc = decisions(1.0, mode, order, perm)
                assert not (reject & truth).any(), (mode, order)
                assert not (accept & ~truth).any(), (mode, order)
                assert (~reject & ~accept).any()  # a band exists
                # share of the guard the error used: the reject side keeps
                # s' <= acc - (2 delta_q + 2 delta_c) + error, the accept side
                # s' + w >= acc + 2 (delta_q + delta_c) - error
                guard = (wq + wc) / 2.0
                used = np.maximum(s - (acc - guard), (acc + guard) - (s + wq + wc)) / guard
                worst = max(worst, float(used.max()))
    print(f"dim {dim}: kp {kp}, worst guard use {worst:.3f}")
    assert worst < 0.5  # the safety factor 2 is intact

    # the check can fail: a coefficient shrunk toward zero lets the same
    # models take wrong decisions on these pairs
    wrong = 0
    for scale in (0.0, 1e-3):
        for mode in ("ru", "rd"):
            _, reject, accept, _, _ = decisions(scale, mode, "seq", None)
            wrong += int((reject & truth).sum() + (accept & ~truth).sum())
    assert wrong > 0


def test_guard_operands_match_the_construction():
    """A'.B' in exact arithmetic equals -2 hi.hi' - 2 hi.lo' - 2 lo.hi' +
    ndn_q + ndn_c for the split coordinates, and the guard covers N - ndn
    (shown on random points with and without centring)."""
    import hegjoin_amd as H
    rng = np.random.default_rng(1)
    for dim, centre in [(18, np.zeros(18)), (90, np.full(90, 0.3))]:
        pts = rng.random((50, dim))
        A, B, w, kp, coef = H.gb_operands(pts, centre)
        g = dim + 2
        yh = bf16(B[:, :dim])
        yl = bf16(B[:, g:g + dim])
        y = pts - centre
        # the split is close: |y - hi - lo| <= 2^-16 |y| (two BF16 roundings)
        assert np.all(np.abs(y - yh - yl) <= 2.0 ** -15 * np.abs(y) + 1e-300)
        ndn = bf16(B[:, dim + 1]) + bf16(B[:, g + dim + 1])
        N = ((yh + yl) ** 2).sum(1)
        assert np.all(ndn <= N - 2 * coef * N * (1 - 1e-9))
        assert np.all(w.astype(np.float64) >= N - ndn + 2 * coef * N)
        for i in range(5):
            j = (i + 7) % len(pts)
            exact = math.fsum(bf16(A[i]) * bf16(B[j]))
            want = math.fsum(list(-2 * yh[i] * yh[j]) + list(-2 * yh[i] * yl[j]) + list(-2 * yl[i] * yh[j])
                             + [ndn[i], ndn[j]])
            assert exact == want
