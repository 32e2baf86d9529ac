"""The FP64 arithmetic helpers of the stage kernels (euler3d_physics.hpp:
log_, sqrt_, div_nr, ln_mean_logs, inv_ln_mean_logs) on the device against
correctly rounded host values (numpy, extended precision where the host has
it).  The helpers replace library sequences in the flux kernels; the rhs
parity tests bound their effect only at the 1e-12 level, these bound each
helper at the unit in the last place."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "device_math", "libmath_probe.so")


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(LIB):
        raise FileNotFoundError(f"{LIB} is missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    f = lib.atmo_probe_math
    f.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int]

    def run(op, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(np.zeros_like(a) if b is None else b, dtype=np.float64)
        out = np.empty_like(a)
        rc = f(op, a.ctypes.data_as(dp), b.ctypes.data_as(dp), out.ctypes.data_as(dp), a.size)
        assert rc == 0, f"HIP error {rc}"
        return out
    return run


def ulps(got, ref):
    """|got - ref| in units of the last place of ref rounded to FP64"""
    ref64 = np.asarray(ref, dtype=np.float64)
    return (np.abs(np.asarray(got, dtype=np.longdouble) - np.asarray(ref, dtype=np.longdouble)) /
            np.spacing(np.abs(ref64)).astype(np.longdouble)).astype(np.float64)


def test_log_is_within_one_ulp(probe):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.exp(rng.uniform(-690.0, 690.0, 200000)),
        rng.uniform(0.5, 2.0, 100000),                       # the reduction's boundaries
        1.0 + rng.uniform(-1e-6, 1e-6, 20000),               # near 1: log(1 + f) ~ f
        np.ldexp(1.0, np.arange(-1022, 1024)),                # powers of two: k log 2 exactly split
        np.nextafter(np.sqrt(0.5), [0.0, 1.0]),              # either side of the mantissa switch
        [np.finfo(np.float64).max, np.finfo(np.float64).tiny],
    ])
    got = probe(0, x)
    ref = np.log(x.astype(np.longdouble))
    u = ulps(got, ref)
    # log(1) = 0 exactly
    assert got[x == 1.0].size == 0 or np.all(got[x == 1.0] == 0.0)
    nz = ref != 0
    print(f"  log_: max {u[nz].max():.3f} ulp over {nz.sum()} operands", flush=True)
    assert u[nz].max() <= 1.0


def test_log_special_operands_follow_the_library(probe):
    x = np.array([0.0, -1.0, np.inf, np.nan, 5e-324, 1e-310, -0.0])
    got = probe(0, x)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.log(x)
    assert got[0] == -np.inf and got[6] == -np.inf
    assert np.isnan(got[1]) and np.isnan(got[3])
    assert got[2] == np.inf
    assert np.all(ulps(got[4:6], np.log(x[4:6].astype(np.longdouble))) <= 1.0), (got[4:6], ref[4:6])


def test_sqrt_matches_the_correctly_rounded_root(probe):
    rng = np.random.default_rng(2)
    x = np.concatenate([np.exp(rng.uniform(-460.0, 460.0, 200000)),
                        rng.uniform(1.0, 4.0, 100000), [0.0, 1.0, 4.0, 2.0]])
    got = probe(1, x)
    ref = np.sqrt(x)
    print(f"  sqrt_: {np.count_nonzero(got != ref)} of {x.size} differ from the IEEE root", flush=True)
    np.testing.assert_array_equal(got, ref)


def test_div_nr_is_within_one_ulp(probe):
    rng = np.random.default_rng(3)
    a = np.exp(rng.uniform(-230.0, 230.0, 200000)) * rng.choice([-1.0, 1.0], 200000)
    b = np.exp(rng.uniform(-230.0, 230.0, 200000)) * rng.choice([-1.0, 1.0], 200000)
    got = probe(2, a, b)
    ref = a.astype(np.longdouble) / b.astype(np.longdouble)
    u = ulps(got, ref)
    exact = np.count_nonzero(got == a / b)
    print(f"  div_nr: max {u.max():.3f} ulp, {exact} of {a.size} equal to the IEEE quotient",
          flush=True)
    assert u.max() <= 1.0


@pytest.mark.parametrize("op", [3, 4])
def test_log_means_from_node_logarithms(probe, op):
    """ln_mean_logs and inv_ln_mean_logs against the logarithmic mean
    (y - x) / (log y - log x) in extended precision, on pairs from equal to
    a factor 100 apart.  In the series branch ((y-x)^2 < 1e-4 (x+y)^2) the
    result is good to a few ulp; outside it the quotient takes the node
    logarithms, whose rounding (one ulp of |log x| and of |log y|) is
    amplified by |log x| + |log y| over |log y - log x|: that is the bound,
    the price of forming one logarithm per node instead of one per pair."""
    rng = np.random.default_rng(4 + op)
    x = np.exp(rng.uniform(-8.0, 12.0, 100000))
    ratio = np.concatenate([np.exp(rng.uniform(-4.6, 4.6, 50000)),
                            1.0 + rng.uniform(-0.03, 0.03, 40000),
                            1.0 + rng.uniform(-1e-9, 1e-9, 9999), [1.0]])
    y = x * ratio
    got = probe(op, x, y)
    xl, yl = x.astype(np.longdouble), y.astype(np.longdouble)
    same = xl == yl
    # log y - log x as log1p((y - x) / x): no cancellation for close pairs
    with np.errstate(divide="ignore", invalid="ignore"):
        lm = np.where(same, xl, (yl - xl) / np.log1p((yl - xl) / xl))
    ref = lm if op == 3 else 1.0 / lm
    rel = np.abs((got.astype(np.longdouble) - ref) / ref).astype(np.float64)
    series = (y - x) ** 2 < 1e-4 * (x + y) ** 2
    lx, ly = np.log(x), np.log(y)
    with np.errstate(divide="ignore"):
        bound = np.where(series, 1e-15,
                         1e-15 + 2.5e-16 * (np.abs(lx) + np.abs(ly)) / np.abs(ly - lx))
    print(f"  op {op}: series branch max relative error {rel[series].max():.3e}; log branch max "
          f"error / bound {(rel[~series] / bound[~series]).max():.3f}", flush=True)
    assert np.all(rel <= bound)


@pytest.mark.parametrize("direction", [0, 1])
def test_flux_ec_rain_gives_the_same_bits_with_its_states_exchanged(direction):
    """the rainy stage takes each node's partners in rotated order with the
    own node on the left, so every pair is evaluated once in each order; the
    flux must give the same bits either way (else the two nodes of a pair
    would not see the same flux)"""
    if not os.path.exists(LIB):
        raise FileNotFoundError(f"{LIB} is missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    f = lib.atmo_probe_flux_ec_rain
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
    f.restype = ctypes.c_int
    rng = np.random.default_rng(11 + direction)
    n = 100000

    def nodes(size):
        rd = rng.uniform(0.4, 1.3, size)
        rv = np.where(rng.random(size) < 0.05, 0.0, rng.uniform(1e-4, 2e-2, size))
        rc = np.where(rng.random(size) < 0.5, 0.0, rng.uniform(0.0, 2e-3, size))
        rr = np.where(rng.random(size) < 0.5, 0.0, rng.uniform(0.0, 2e-3, size))
        v = rng.uniform(-40.0, 40.0, (2, size))
        T = rng.uniform(230.0, 310.0, size)
        vr = rng.uniform(0.0, 8.0, size)
        with np.errstate(divide="ignore"):
            return np.stack([rd, rv, rc, rr, v[0], v[1], 1.0 / T, vr, np.log(rd), np.log(rv),
                             -np.log(T)], axis=1)

    l = nodes(n)
    r = nodes(n)
    # a third of the pairs close together (the series branch of the log means)
    close = slice(0, n // 3)
    r[close] = l[close] * (1.0 + rng.uniform(-1e-3, 1e-3, (n // 3, 11)))
    r[close, 8] = np.log(r[close, 0])
    with np.errstate(divide="ignore"):
        r[close, 9] = np.log(r[close, 1])
    r[close, 10] = np.log(r[close, 6])
    l, r = np.ascontiguousarray(l), np.ascontiguousarray(r)
    n1, n2 = (64.0, 0.0) if direction == 0 else (0.0, 51.2)
    f_lr = np.empty((n, 6))
    f_rl = np.empty((n, 6))
    rc = f(l.ctypes.data_as(dp), r.ctypes.data_as(dp), n, n1, n2,
           f_lr.ctypes.data_as(dp), f_rl.ctypes.data_as(dp))
    assert rc == 0, f"HIP error {rc}"
    assert np.all(np.isfinite(f_lr))
    differ = np.count_nonzero(f_lr.view(np.int64) != f_rl.view(np.int64))
    print(f"  direction {direction}: {differ} of {f_lr.size} entries differ", flush=True)
    assert differ == 0
