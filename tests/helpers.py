"""Shared helpers for the parity tests: tolerance checks against the oracle
and a runner for the device provider chain."""
import sys

import numpy as np

# FP64 plain form: within 1e-12 relative (1e-10 on Laplacians) of the oracle,
# relative to the sum of the absolute values of the terms of each entry (the
# oracle's `mag`), so that cancellation does not make the bound ill-posed.
RTOL_VGL = np.array([1e-12, 1e-12, 1e-12, 1e-12, 1e-10])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err_vs_mag(got, ref, mag):
    """max over entries of |got - ref| / mag, per VGL component (axis 1)."""
    d = np.abs(got - ref)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(mag > 0, d / mag, np.where(d > 0, np.inf, 0.0))
    return r.reshape(r.shape[0], r.shape[1], -1).max(axis=(0, 2))


# Entries whose terms cancel by less than this factor (|ref| >= COND * mag)
# are also held to the bound element by element, |got - ref| <= rtol |ref|.
COND = 1e-3


def elementwise_rel_err(got, ref, mag, cond=COND):
    """(max |got - ref| / |ref| per VGL component over the entries with
    |ref| >= cond * mag, fraction of entries that qualify)."""
    d = np.abs(got - ref)
    a = np.abs(ref)
    ok = (a >= cond * mag) & (a > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(ok, d / np.where(ok, a, 1.0), 0.0)
    return r.reshape(r.shape[0], r.shape[1], -1).max(axis=(0, 2)), float(ok.mean())


def assert_vgl_close(got, ref, mag, rtol=RTOL_VGL, what=""):
    assert got.shape == ref.shape, (got.shape, ref.shape)
    err = rel_err_vs_mag(got, ref, mag)
    el, frac = elementwise_rel_err(got, ref, mag)
    say(f"{what}: max |d|/mag per component = {np.array2string(err, precision=2)}; "
        f"max |d|/|ref| over the {frac:.1%} of entries with |ref| >= {COND:g} mag = "
        f"{np.array2string(el, precision=2)}")
    assert np.all(err <= rtol), (what, err)
    assert np.all(el <= rtol), (what, "elementwise", el)
    return err


def device_chain(ctx, P, want=("en", "ao", "mo")):
    out = {}
    if "en" in want:
        out["en"] = ctx.en_distance(P)
    if "shell" in want:
        out["shell"] = ctx.shell_vgl(P)
    if "ao" in want:
        out["ao"] = ctx.ao_vgl(P)
    if "ao_value" in want:
        out["ao_value"] = ctx.ao_value(P)
    if "mo" in want:
        out["mo"] = ctx.mo_vgl(P)
    if "mo_value" in want:
        out["mo_value"] = ctx.mo_value(P)
    return out
