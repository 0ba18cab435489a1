"""The host factorisation the kernels run on (cs3d_plan_info), checked on the
CPU: a numpy emulation of the kernels' segmented two-recurrence solve with the
library's own tables must reproduce the oracle's Thomas / Sherman-Morrison
operators (Utils.cpp:4-46, Derivatives.cpp:250-284, Filter.cpp:98-117).
This pins the factorisation, the segmentation, the cross-segment carries and
their truncation without a GPU, and the filter's difference form
f - A^-1 (A - B) f with (A - B) = (1 - 2 alpha_f)/256 D8 in the interior."""
import ctypes

import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import relerr


def plan(kind, bc, n, dh, af, prec):
    L = cs3d.lib()
    ints = (ctypes.c_int * 8)()
    cs3d.check(L.cs3d_plan_info(kind, bc, n, dh, af, prec, ints, None, 0, None, 0))
    nr = ints[7]
    P = ints[2]
    reals = np.zeros(nr)
    segs = np.zeros(2 * P, dtype=np.int32)
    cs3d.check(L.cs3d_plan_info(kind, bc, n, dh, af, prec, ints,
                                reals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nr,
                                segs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 2 * P))
    n, M, P, K, uni, cyc, NB = list(ints)[:7]
    r = reals
    d = dict(n=n, M=M, P=P, K=K, uniform=uni, cyclic=cyc, NB=NB, scale=r[0], tap=r[1:6],
             bcl=r[6:34].reshape(4, 7), bcr=r[34:62].reshape(4, 7))
    o = 62
    for name in ("m", "iw", "mu", "pf", "pb"):
        d[name] = r[o:o + n]
        o += n
    d["phi"] = r[o:o + P * K].reshape(P, K)
    o += P * K
    d["psi"] = r[o:o + P * K].reshape(P, K)
    d["sstart"] = segs[:P]
    d["slen"] = segs[P:]
    return d


def emulate(p, kind, f):
    """The kernels' arithmetic, lane by lane, in FP64."""
    n, P, K = p["n"], p["P"], p["K"]
    per = p["cyclic"]
    t = p["tap"]
    H = 4 if kind == 0 else 2
    g = np.pad(f, H, mode="wrap") if per else np.pad(f, H, mode="edge")
    c = lambda o: g[H + o:H + o + n]
    if kind == 1:
        r = t[1] * (c(1) - c(-1)) + t[2] * (c(2) - c(-2))
    elif kind == 2:
        r = t[0] * c(0) + t[1] * (c(-1) + c(1)) + t[2] * (c(-2) + c(2))
    else:
        # filter plans hold (A - B): kappa times the 8th central difference
        d8 = (1, -8, 28, -56, 70, -56, 28, -8, 1)
        r = t[0] * sum(w * c(k - 4) for k, w in enumerate(d8))
    for i in range(p["NB"]):
        r[i] = p["bcl"][i] @ f[0:7]
        r[n - p["NB"] + i] = p["bcr"][i] @ f[n - 7:n]
    y = r.copy()
    ends = np.zeros(P)
    for gi in range(P):
        s, ln = p["sstart"][gi], p["slen"][gi]
        for i in range(s + 1, s + ln):
            y[i] -= p["m"][i] * y[i - 1]
        ends[gi] = y[s + ln - 1]
    for gi in range(P):
        s, ln = p["sstart"][gi], p["slen"][gi]
        C = 0.0
        for k in range(1, K + 1):
            gg = gi - k
            if per:
                gg %= P
            elif gg < 0:
                break
            C += p["phi"][gi, k - 1] * ends[gg]
        y[s:s + ln] += p["pf"][s:s + ln] * C
    x = np.zeros(n)
    starts = np.zeros(P)
    for gi in range(P):
        s, ln = p["sstart"][gi], p["slen"][gi]
        carry = 0.0
        for i in range(s + ln - 1, s - 1, -1):
            carry = p["iw"][i] * y[i] + p["mu"][i] * carry
            x[i] = carry
        starts[gi] = x[s]
    out = np.zeros(n)
    for gi in range(P):
        s, ln = p["sstart"][gi], p["slen"][gi]
        D = 0.0
        for k in range(1, K + 1):
            gg = gi + k
            if per:
                gg %= P
            elif gg >= P:
                break
            D += p["psi"][gi, k - 1] * starts[gg]
        out[s:s + ln] = p["scale"] * (x[s:s + ln] + p["pb"][s:s + ln] * D)
    return f - out if kind == 0 else out


@pytest.mark.parametrize("n", [8, 9, 12, 42, 64, 128, 256])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("kind", [1, 2, 0])
@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_emulated_plan_matches_oracle(n, bc, kind, prec):
    rng = np.random.default_rng(n + 100 * bc + 10 * kind)
    dh = 0.05
    p = plan(kind, bc, n, dh, 0.49, prec)
    f = rng.standard_normal(n)
    ref = po.line_op(kind, bc, f, dh=dh, alpha_f=0.49)
    got = emulate(p, kind, f)
    # FP32 plans drop cross-segment terms below 1e-8 (a sixth of FP32's unit
    # roundoff), FP64 plans below 1e-17; the emulation runs in FP64, so the
    # difference is the truncation alone
    tol = 2e-8 if prec == cs3d.FP32 else 1e-13
    assert relerr(got, ref) < tol


def test_periodic_plans_are_uniform_and_short():
    for n in (64, 128, 256):
        p = plan(1, 0, n, 0.1, 0.49, cs3d.FP32)
        assert p["uniform"] and p["M"] == 4 and p["P"] == n // 4 and p["K"] <= 8
        p = plan(0, 0, n, 0.1, 0.49, cs3d.FP64)
        assert p["uniform"] and p["M"] == 8


@pytest.mark.parametrize("kind", [1, 2, 0])
@pytest.mark.parametrize("bc", [0, 1])
def test_plan_segment_length_matches_a_compiled_kernel(kind, bc):
    """Uniform plans use the kernels' uniform segment length (4 for the
    derivatives, 8 for the filter); every other plan uses the table kernels'
    length 4 and is never marked uniform (a ragged periodic filter line of 12
    points splits into equal 4-point segments but has no uniform 4-point
    filter kernel)."""
    mu = 8 if kind == 0 else 4
    for n in (8, 9, 12, 16, 20, 24, 36, 40, 64, 100, 256):
        p = plan(kind, bc, n, 0.1, 0.49, cs3d.FP32)
        if p["uniform"]:
            assert bc == 0 and p["M"] == mu and n % mu == 0
        else:
            assert p["M"] == 4
            assert not (bc == 0 and n % mu == 0)


@pytest.mark.parametrize("af", [0.49, 0.495, 0.3, -0.2])
def test_filter_difference_form_symbol(af):
    """(A - B) of the 8th-order filter is exactly (1 - 2 alpha_f)/256 times the
    8th central difference: the symbol (1 + 2 alpha_f cos t) - B(t), with B's
    taps from Filter.hpp:54-63, equals (1 - 2 alpha_f)(1 - cos t)^4 / 16."""
    a0 = (93 + 70 * af) / 128
    a = [(7 + 18 * af) / 16, (-7 + 14 * af) / 32, (1 - 2 * af) / 16, (-1 + 2 * af) / 128]
    th = np.linspace(0.3, np.pi, 40)
    sym = (1 + 2 * af * np.cos(th)) - (a0 + sum(a[n - 1] * np.cos(n * th) for n in range(1, 5)))
    ref = (1 - 2 * af) * (1 - np.cos(th)) ** 4 / 16
    assert np.max(np.abs(sym - ref)) < 1e-15
    p = plan(0, 0, 64, 0.1, af, cs3d.FP64)
    assert abs(p["tap"][0] - (1 - 2 * af) / 256) < 1e-18 and not np.any(p["tap"][1:])


def test_filter_low_wavenumber_gain_in_fp32():
    """In FP32 the difference form keeps the gain of a resolved wave at 1 to
    rounding: applied to a k = 1 wave on 64 points (the Taylor-Green scale),
    the FP32-rounded plan moves the gain by < 1e-12, where a direct B f with
    FP32 taps moves it by ~1e-8 per application."""
    n = 64
    p = plan(0, 0, n, 0.1, 0.49, cs3d.FP32)
    q = dict(p)
    for key in ("tap", "m", "iw", "mu", "pf", "pb", "phi", "psi", "bcl", "bcr"):
        q[key] = p[key].astype(np.float32).astype(np.float64)
    q["scale"] = float(np.float32(p["scale"]))
    i = np.arange(n)
    f = np.sin(2 * np.pi * i / n) + 0.3 * np.cos(2 * np.pi * i / n)
    gain = np.dot(emulate(q, 0, f), f) / np.dot(f, f)
    assert abs(gain - 1) < 1e-12


def wave_scan(e, lam, fwd):
    """the device's wavefront carry scan (wave_carry) on 64 lanes, in FP64:
    rotate by one lane, weighted inclusive scan inside rows of 16 with zero
    fill, the four row totals, per-lane weights lam^d / (1 - lam^64)"""
    x = np.roll(e, 1 if fwd else -1)
    for d in (1, 2, 4, 8):
        s = np.zeros(64)
        for g in range(64):
            j = g % 16
            if fwd and j >= d:
                s[g] = x[g - d]
            if not fwd and j + d <= 15:
                s[g] = x[g + d]
        x = x + lam ** d * s
    tot = [x[16 * q + 15] if fwd else x[16 * q] for q in range(4)]
    out = x.copy()
    for g in range(64):
        r, j = g // 16, g % 16
        for q in range(4):
            k = j + 1 + 16 * ((r - q - 1) % 4) if fwd else 16 - j + 16 * ((q - r - 1) % 4)
            out[g] += lam ** k / (1 - lam ** 64) * tot[q]
    return out


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_wavefront_scan_premise(kind, prec):
    """Lines of 64 uniform periodic segments take their carries from a
    wavefront scan (x lines of 256 points; the x filter fused into the last
    stage runs on the 4-point derivative segmentation).  Its premise, on the
    library's own tables: the carry weights are geometric, phi[k-1] =
    lam^(k-1) / (1 - lam^P) with lam = pf[M-1] forward and psi with
    pb[0] backward; and the scan's lane arithmetic equals the direct cyclic
    sums with the kept weights, to their truncation."""
    n = 256 if kind else 512   # the filter's own plans use 8-point segments
    p = plan(kind, 0, n, 0.05, 0.49, prec)
    M, P, K = p["M"], p["P"], p["K"]
    assert p["uniform"] and p["cyclic"] and P == 64
    lf, lb = p["pf"][M - 1], p["pb"][0]
    k = np.arange(K)
    np.testing.assert_allclose(p["phi"][0], lf ** k / (1 - lf ** P), rtol=1e-13, atol=0)
    np.testing.assert_allclose(p["psi"][0], lb ** k / (1 - lb ** P), rtol=1e-13, atol=0)
    rng = np.random.default_rng(kind)
    e = rng.standard_normal(64)
    C = np.array([sum(p["phi"][g, i] * e[(g - 1 - i) % 64] for i in range(K)) for g in range(64)])
    D = np.array([sum(p["psi"][g, i] * e[(g + 1 + i) % 64] for i in range(K)) for g in range(64)])
    tol = 1e-7 if prec == cs3d.FP32 else 1e-15   # the tables' truncation
    assert np.abs(wave_scan(e, lf, True) - C).max() < tol * np.abs(e).sum()
    assert np.abs(wave_scan(e, lb, False) - D).max() < tol * np.abs(e).sum()
