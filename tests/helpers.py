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


def merge_2d_own_cell_scene():
    """A 2D scene for the merge case where the particle whose slot equals a cell id sits in
    that cell and heads a pair (SPH-2D/sph_system.cu:1207-1266): 40 particles, slots 33-36
    small (level 2) in cell 33 = (1, 1) of the reference's 32 x 32 default grid, energies
    0, 0.2, 0, 0 (merge threshold 0.04); slots 37, 38 small in cell 34, qualified by slot
    34's energy 0.2 (never); every other particle large, alone in its cell. The first pair
    (33, 34) averages slot 33's energy to 0.1, which disqualifies the rest of cell 33, so
    the second of the two pairs counted finds no entries and merges (33, 34) again."""
    import sca2012_amd as sa
    p = sa.param_defaults(sa.V2D)
    a = np.zeros(40, dtype=sa.PARTICLE2D_DTYPE)
    for i in range(40):
        a["pos"][i] = (10.5 + (i % 10), 10.5 + (i // 10))
    a["level"] = 1
    a["pos"][33:37] = [(1.2, 1.3), (1.4, 1.5), (1.6, 1.2), (1.8, 1.7)]
    a["pos"][37:39] = [(2.3, 1.4), (2.6, 1.6)]
    a["level"][33:39] = 2
    a["energy"][34] = 0.2
    rng = np.random.default_rng(2012)
    for f in ("vel", "ev", "acc"):
        a[f] = rng.normal(size=(40, 2)).astype(np.float32)
    return p, a
