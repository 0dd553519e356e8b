"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both are identically zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    diff = np.abs(a - b).max() if b.size else 0.0
    if scale == 0.0:
        return diff
    return diff / scale


def assert_close(name, a, b, tol):
    e = rel_err(a, b)
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"


def assert_same_int(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    bad = np.count_nonzero(a != b)
    assert bad == 0, f"{name}: {bad} entries differ"


def vel_err(out, ref, keys=("UU", "VV", "WW")):
    """Largest component error of a velocity field relative to the velocity
    scale max(|u|,|v|,|w|): a component that is round-off noise (flow uniform
    in one direction) is judged against the flow, not against its own noise."""
    scale = max(np.abs(np.asarray(ref[k])).max() for k in keys)
    diff = max(np.abs(np.asarray(out[k]) - np.asarray(ref[k])).max() for k in keys)
    return diff if scale == 0.0 else diff / scale


def assert_vel_close(name, out, ref, tol, keys=("UU", "VV", "WW")):
    e = vel_err(out, ref, keys)
    assert e <= tol, f"{name}: velocity error {e:.3e} of the velocity scale > {tol:.1e}"
