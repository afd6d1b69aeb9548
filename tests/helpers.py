"""Comparison helpers shared by the tests."""
import numpy as np


def as_words(a):
    """View a structured particle array as uint32 words per record."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint32).reshape(len(a), -1)


def float_fields_equal(x, y):
    """Bitwise equality that treats any two NaNs as equal (NaN payloads differ between
    x86 and the GPU)."""
    x = np.asarray(x, dtype=np.float32)
    y = np.asarray(y, dtype=np.float32)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    if not np.array_equal(nx, ny):
        return False
    return np.array_equal(x[~nx].view(np.uint32), y[~ny].view(np.uint32))


def mismatch_report(g, c, fields):
    lines = []
    for f in fields:
        x = np.asarray(g[f], dtype=np.float64)
        y = np.asarray(c[f], dtype=np.float64)
        if g[f].dtype.kind == "u":
            bad = np.nonzero((g[f] != c[f]).reshape(len(g), -1).any(axis=1))[0]
        else:
            eq = (np.isnan(x) & np.isnan(y)) | (g[f].view(np.uint32) == c[f].view(np.uint32))
            bad = np.nonzero(~eq.reshape(len(g), -1).all(axis=1))[0]
        if len(bad):
            with np.errstate(invalid="ignore", divide="ignore"):
                rel = np.nanmax(np.abs(x - y) / np.maximum(np.abs(y), 1e-30))
            lines.append("%s: %d records differ (first %s), max rel %.3g" % (f, len(bad), bad[:5], rel))
    return "; ".join(lines)


def assert_particles_equal(g, c, fields=None, what=""):
    assert len(g) == len(c), "%s: count %d vs %d" % (what, len(g), len(c))
    fields = fields or list(g.dtype.names)
    ok = True
    for f in fields:
        if g[f].dtype.kind == "u":
            ok &= np.array_equal(g[f], c[f])
        else:
            ok &= float_fields_equal(g[f], c[f])
    assert ok, "%s: %s" % (what, mismatch_report(g, c, fields))
