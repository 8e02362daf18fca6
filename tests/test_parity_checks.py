"""CPU checks of the GPU parity bars (tests/parity_checks.py): they accept the oracle's
own output and reject planted mistakes (a perturbed value, a dropped term)."""
import numpy as np
import pytest

from oracle import oracle as O
from parity_checks import check


def _oracle(rng, kind, n, a, b, **kw):
    g = O.OracleGen(rng, **kw)
    st, x = g.generate(kind, n, 0x10000, a, b)
    assert st == 0
    return x


@pytest.mark.parametrize("kind,a,b", [("normal_f32", 0.0, 1.0), ("normal_f64", -1.5, 2.25),
                                      ("normal_f16", 0.5, 1.5), ("lognormal_f32", 0.3, 0.7),
                                      ("lognormal_f64", 0.0, 1.0), ("lognormal_f16", 0.0, 0.5)])
def test_box_muller_bar(kind, a, b):
    want = _oracle(O.PHILOX4_32_10, kind, 20000, a, b, seed=5)
    check(kind, want.copy(), want, (a, b))
    bad = want.copy()
    bad[123] = bad[123] * (1.01 if kind.startswith("log") or kind.endswith("f16") else 1 + 1e-3)
    with pytest.raises(AssertionError):
        check(kind, bad, want, (a, b))
    # a dropped mean (mean + v*sd -> v*sd) must fail
    if a != 0.0 and not kind.startswith("log"):
        v = ((want.astype(np.float64) - a) / b * b).astype(want.dtype)
        with pytest.raises(AssertionError):
            check(kind, v, want, (a, b))


@pytest.mark.parametrize("kind", ["normal_f32", "normal_f64", "lognormal_f32", "normal_f16"])
def test_erfinv_bar(kind):
    a, b = (0.5, 1.5) if kind.startswith("normal") else (0.3, 0.7)
    want = _oracle(O.SOBOL32, kind, 20000, a, b)
    check(kind, want.copy(), want, (a, b), erfinv=True)
    bad = want.copy()
    bad[777] = bad[777] * 1.01
    with pytest.raises(AssertionError):
        check(kind, bad, want, (a, b), erfinv=True)
    # more than 1% of values off by one ulp fails the bit-equality floor
    ulp = want.copy()
    idx = np.arange(0, want.size, 50)
    ulp[idx] = np.nextafter(want[idx], np.inf, dtype=want.dtype)
    with pytest.raises(AssertionError):
        check(kind, ulp, want, (a, b), erfinv=True)


def test_exact_bar():
    want = _oracle(O.MRG32K3A, "uniform_f32", 1000, 0.0, 0.0)
    check("uniform_f32", want.copy(), want)
    bad = want.copy()
    bad[5] = np.nextafter(bad[5], np.float32(2))
    with pytest.raises(AssertionError):
        check("uniform_f32", bad, want)


@pytest.mark.parametrize("kind", ["normal_f32", "normal_f64"])
def test_box_muller_bar_is_two_ulp(kind):
    # at mean 0, stddev 1 the bar is the plain ulp distance: 2 ulp passes, 3 fails
    want = _oracle(O.PHILOX4_32_10, kind, 4000, 0.0, 1.0, seed=8)
    two = want.copy()
    for _ in range(2):
        two[10] = np.nextafter(two[10], np.inf, dtype=want.dtype)
    check(kind, two, want, (0.0, 1.0))
    three = two.copy()
    three[10] = np.nextafter(three[10], np.inf, dtype=want.dtype)
    with pytest.raises(AssertionError):
        check(kind, three, want, (0.0, 1.0))


def test_erfinv_bar_is_bounded_near_the_median():
    # near v = 0 the cancellation term is capped at sqrt(2 ulp(4.33)) (~1e-3 for float):
    # a value 0.01 away from the oracle's fails however close to the median it lies
    want = _oracle(O.SOBOL32, "normal_f32", 20000, 0.0, 1.0)
    i = int(np.argmin(np.abs(want)))
    bad = want.copy()
    bad[i] = bad[i] + np.float32(0.01)
    with pytest.raises(AssertionError):
        check("normal_f32", bad, want, (0.0, 1.0), erfinv=True)
