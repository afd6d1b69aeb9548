"""sph_sort_noalloc -- the sort with its working memory in a static device array (no
allocation) -- against the context sort sph_sort on the same keys, bit for bit: two calls in
a row, the second on a smaller input (the first call's contents are still in the shared
array), then odd pass counts, the size limit and the argument checks.

The keys are a context's own cell hashes of a seeded scene: after sph_calc_hash and
sph_sort the context holds the sorted keys and the entry index of each, which give back the
unsorted keys by scattering. Device arrays for the call come from torch (plumbing only)."""
import ctypes

import numpy as np
import pytest
import torch

import sca2012_amd as sa
from sca2012_amd import _abi, scenes

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20  # SPH_SORT_NOALLOC_MAX_KEYS


def _context_sort(variant, params, parts):
    """(unsorted keys, sorted keys, sorted index, key bits) of a context's hash and sort."""
    with sa.SPHContext(variant, params, max(len(parts), 1)) as g:
        g.upload(parts)
        g.calc_hash()
        g.sort()
        h = g.get_u32(_abi.Q_HASH)
        i = g.get_u32(_abi.Q_INDEX)
    assert len(h) == len(parts)
    assert np.array_equal(np.sort(i), np.arange(len(parts), dtype=np.uint32))
    keys = np.empty_like(h)
    keys[i] = h
    bits = max(1, int(np.ceil(np.log2(int(params["total_cell"][0])))))
    return keys, h, i, bits


def _noalloc(keys, index, bits):
    lib = sa.load_library()
    dk = torch.from_numpy(np.ascontiguousarray(keys, dtype=np.uint32).view(np.int32)).cuda()
    dv = torch.from_numpy(np.ascontiguousarray(index, dtype=np.uint32).view(np.int32)).cuda()
    stream = torch.cuda.current_stream()
    rc = lib.sph_sort_noalloc(ctypes.c_void_p(dk.data_ptr()), ctypes.c_void_p(dv.data_ptr()),
                              len(keys), bits, ctypes.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return rc, dk.cpu().numpy().view(np.uint32), dv.cpu().numpy().view(np.uint32)


def _numpy_stable(keys, index, bits):
    mask = np.uint32(0xFFFFFFFF) if bits in (0, 32) else np.uint32((1 << bits) - 1)
    o = np.argsort(keys & mask, kind="stable")
    return keys[o], index[o]


def test_noalloc_sort_matches_context_sort():
    big = scenes.WORKLOADS["adaptive_3d2_262k"]
    p1, a1 = big["build"]()
    p2, a2 = scenes.reference_scene(sa.V3D4)
    assert len(a2) < len(a1) <= MAX_KEYS

    # first call: 262,144 keys of 18 bits (two passes, in place)
    keys1, h1, i1, bits1 = _context_sort(big["variant"], p1, a1)
    print("call 1: %d keys, %d bits" % (len(keys1), bits1), flush=True)
    rc, k, v = _noalloc(keys1, np.arange(len(keys1), dtype=np.uint32), bits1)
    assert rc == 0
    assert np.array_equal(k, h1) and np.array_equal(v, i1), "call 1 differs from sph_sort"

    # second call, smaller: 41,850 keys of 14 bits
    keys2, h2, i2, bits2 = _context_sort(sa.V3D4, p2, a2)
    print("call 2: %d keys, %d bits" % (len(keys2), bits2), flush=True)
    idx2 = np.arange(len(keys2), dtype=np.uint32)
    rc, k, v = _noalloc(keys2, idx2, bits2)
    assert rc == 0
    assert np.array_equal(k, h2) and np.array_equal(v, i2), "call 2 differs from sph_sort"

    # the same keys over 27 bits (three passes, the result copied back from the scratch
    # pair) and over all 32 bits (four passes): the keys are below 2^14, so the order is
    # the context's
    for bits in (27, 0):
        rc, k, v = _noalloc(keys2, idx2, bits)
        assert rc == 0
        assert np.array_equal(k, h2) and np.array_equal(v, i2), "key_bits %d" % bits

    # one pass of 9 bits sorts by the low bits only; caller-given values travel with keys
    vals = (np.arange(len(keys2), dtype=np.uint32) * np.uint32(2654435761)) ^ np.uint32(0x5CA2012)
    rc, k, v = _noalloc(keys2, vals, 9)
    ek, ev = _numpy_stable(keys2, vals, 9)
    assert rc == 0 and np.array_equal(k, ek) and np.array_equal(v, ev), "one 9-bit pass"


def test_noalloc_sort_limits_and_arguments():
    lib = sa.load_library()
    n = MAX_KEYS + 1
    dk = torch.zeros(n, dtype=torch.int32, device="cuda")
    dv = torch.zeros(n, dtype=torch.int32, device="cuda")
    pk, pv = ctypes.c_void_p(dk.data_ptr()), ctypes.c_void_p(dv.data_ptr())
    assert lib.sph_sort_noalloc(pk, pv, n, 18, None) == _abi.ECAP
    assert lib.sph_sort_noalloc(None, pv, 16, 18, None) == _abi.EINVAL
    assert lib.sph_sort_noalloc(pk, None, 16, 18, None) == _abi.EINVAL
    assert lib.sph_sort_noalloc(pk, pv, 16, 33, None) == _abi.EINVAL
    assert lib.sph_sort_noalloc(None, None, 0, 18, None) == _abi.OK  # empty: nothing to do
    # at the limit itself: all keys equal except a descending tail
    keys = np.zeros(MAX_KEYS, dtype=np.uint32)
    keys[-1000:] = np.arange(1000, 0, -1, dtype=np.uint32)
    idx = np.arange(MAX_KEYS, dtype=np.uint32)
    rc, k, v = _noalloc(keys, idx, 10)
    ek, ev = _numpy_stable(keys, idx, 10)
    assert rc == 0 and np.array_equal(k, ek) and np.array_equal(v, ev)
