"""Comparison bars shared by the GPU parity tests (numpy only).

Integers, uniforms and Poisson are compared bit-exactly.  The floating-point bars
follow from where the device and the oracle (host libm) may legitimately differ:

* Box-Muller normals, float/double: the host's float logf/sinf/cosf are not
  correctly rounded (~1% of calls return the other neighbour) and the device's
  are, so a pair can sit up to 3 ulp apart.  Bound: 3 ulp of the variate v,
  scaled by |stddev|, plus the two roundings of mean + v*stddev; at most
  1 + 2e-4*n values beyond 2 ulp.
* Half normals: the reference rounds the Box-Muller variate to half before the
  affine (rocrand_normal.h:108-135 host branch).  A one-float-ulp difference next
  to a half rounding boundary moves that intermediate by one half ulp, which the
  affine (or exp) carries into the output.  Bound: 1 half ulp, except at most
  1 + 1e-4*n values (measured: 2.7e-5*n) within |stddev| * ulp16(v) + 2 ulp16(out).
* Inverse-error-function normals (Sobol, ThreeFry2x32 doubles, Poisson above
  lambda 4000): rocrand_normal.h:174-221 evaluates -tt1 + sqrt(tt1^2 - tt2) with
  tt1 ~ 4.33, which cancels near the median; a one-ulp difference in the log moves
  tt1's rounding by one ulp and the output by up to ~ulp(tt1) / |v|.  Bound:
  3 ulp(v) + 4 ulp(4.33) / |v| (times |stddev|), with 99% of values bit-equal.
* Log-normals are judged on log(out), where exp's rounding adds 2 ulp relative.
"""
import numpy as np

from conftest import ulp_distance


def _affine_bound(v, sd, ftype):
    return np.spacing(np.abs(v).astype(ftype)).astype(np.float64) * abs(sd)


def check_exact(got, want):
    assert got.shape == want.shape
    bad = np.nonzero(got.view(np.uint8) != want.view(np.uint8))[0]
    assert bad.size == 0, f"mismatch at byte {bad[:5]}: got {got[:8]} want {want[:8]}"


def check_box_muller(kind, got, want, mean, sd):
    assert got.shape == want.shape
    log = kind.startswith("log")
    zg = np.log(got.astype(np.float64)) if log else got.astype(np.float64)
    zw = np.log(want.astype(np.float64)) if log else want.astype(np.float64)
    v = (zw - mean) / sd
    if got.dtype == np.float16:
        d = ulp_distance(got, want)
        over = np.nonzero(d > 1)[0]
        assert over.size <= 1 + 1e-4 * got.size, f"{over.size} half values beyond 1 ulp"
        # one half ulp of the rounded variate, carried through the affine / exp
        lim = (_affine_bound(v, sd, np.float16)
               + 2 * np.spacing(np.abs(want).astype(np.float16)).astype(np.float64) / (
                   np.abs(want.astype(np.float64)) if log else 1.0))
        err = np.abs(zg - zw)
        bad = over[err[over] > lim[over]]
        assert bad.size == 0, f"half values beyond the bound at {bad[:5]}: got {got[bad[:3]]} want {want[bad[:3]]}"
        return
    ftype = got.dtype.type
    ulp_v = _affine_bound(v, sd, ftype)
    if log:
        ulp_o = 2 * np.finfo(ftype).eps * np.maximum(np.abs(zw), 1.0)
    else:
        ulp_o = (np.spacing(np.abs(want)).astype(np.float64)
                 + np.spacing(np.abs(v * sd).astype(ftype)).astype(np.float64))
    err = np.abs(zg - zw)
    worst = int(np.argmax(err - 3 * ulp_v - ulp_o))
    assert np.all(err <= 3 * ulp_v + ulp_o), f"beyond 3 ulp at {worst}: got {got[worst]!r} want {want[worst]!r}"
    over = np.count_nonzero(err > 2 * ulp_v + ulp_o)
    assert over <= 1 + 2e-4 * err.size, f"{over} of {err.size} values at 3 ulp"


def check_erfinv(kind, got, want, mean, sd):
    assert got.shape == want.shape
    log = kind.startswith("log")
    single = got.dtype != np.float64
    ftype = np.float32 if single else np.float64
    c1_ulp = float(np.spacing(np.float32(4.33))) if single else float(np.spacing(4.33))
    zg = np.log(got.astype(np.float64)) if log else got.astype(np.float64)
    zw = np.log(want.astype(np.float64)) if log else want.astype(np.float64)
    v = (zw - mean) / sd
    r = np.maximum(np.abs(v) / np.sqrt(2.0), 1e-30)
    bound = (3 * np.spacing(np.abs(v).astype(ftype)).astype(np.float64)
             + 4 * c1_ulp * np.sqrt(2.0) / (2 * r)) * abs(sd)
    slack = 2 * np.spacing(np.abs(zw).astype(ftype)).astype(np.float64)
    if got.dtype == np.float16:
        hs = np.spacing(np.abs(want).astype(np.float16)).astype(np.float64)
        slack = slack + 2 * (hs / np.abs(want.astype(np.float64)) if log else hs)
    err = np.abs(zg - zw)
    bad = np.nonzero(~(err <= bound + slack))[0]
    assert bad.size == 0, f"{bad.size} values beyond the bound, first {bad[:3]}: got {got[bad[:3]]} want {want[bad[:3]]}"
    if kind.endswith("f64") and log:
        return  # the device's double exp is not correctly rounded: last-bit differences
    same = np.count_nonzero(got.view(f"u{got.dtype.itemsize}") == want.view(f"u{want.dtype.itemsize}"))
    # the 99% rate is a large-sample figure: a few last-bit differences may fall in a
    # sample of a few dozen values
    assert same >= 0.99 * got.size - (2 if got.size < 1000 else 0), f"only {same} of {got.size} bit-equal"


def check(kind, got, want, params=(0.0, 1.0), erfinv=False):
    """Dispatch on the output kind (see the module docstring)."""
    mean, sd = params
    if "normal" in kind:
        if erfinv:
            check_erfinv(kind, got, want, mean, sd)
        else:
            check_box_muller(kind, got, want, mean, sd)
    else:
        check_exact(got, want)
