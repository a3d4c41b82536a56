"""The allocation-free entry point ljg_force_lc_noalloc (include/ljgpu.h) through the C ABI, against the
context path it is a variant of (ljg_upload + ljg_force(LJG_KERNEL_LC_FP64) + ljg_download).

Bar: bit-identical forces, potential energy, virial and pair count (integer and FP64 results of the same
kernels in the same order), on two calls in a row, the second with a smaller input in another box, on a
stream of the caller's and on the default stream.  The inputs, outputs and the stream are the test's own
(hipMalloc / hipStreamCreate through the HIP runtime the library is linked with); the entry point
allocates nothing.  Oversized inputs are refused on the host with LJG_ERR_CAPACITY.
"""
import ctypes

import numpy as np
import pytest

import autopas_amd as ap
from autopas_amd import engine
from conftest import announce

pytestmark = pytest.mark.gpu

RC, SKIN = 2.5, 0.3
H2D, D2H = 1, 2


def hip():
    # the runtime libljgpu.so is linked with (already loaded with it, so the same instance)
    engine.load_library()
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    lib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    lib.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class Dev:
    """A device buffer of the test (hipMalloc), optionally filled from a numpy array."""

    def __init__(self, h, nbytes, src=None):
        self.h, self.nbytes = h, nbytes
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        if src is not None:
            a = np.ascontiguousarray(src)
            assert a.nbytes == nbytes
            assert h.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def params(L):
    p = engine.Params()
    p.box_min[:] = [0.0, 0.0, 0.0]
    p.box_max[:] = [L, L, L]
    p.cutoff, p.skin, p.apply_shift = RC, SKIN, 1
    return p


def noalloc(h, L, x, y, z, type_, id_, eps, sig, stream):
    n = len(x)
    bufs = [Dev(h, 8 * n, a) for a in (x, y, z)]
    dt = Dev(h, 4 * n, type_.astype(np.int32)) if type_ is not None else None
    di = Dev(h, 8 * n, id_.astype(np.int64)) if id_ is not None else None
    fo = [Dev(h, 8 * n) for _ in range(3)]
    dg = Dev(h, ctypes.sizeof(engine.Globals))
    e = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
    s = None if sig is None else np.ascontiguousarray(sig, dtype=np.float64)
    lib = engine.load_library()
    st = lib.ljg_force_lc_noalloc(ctypes.byref(params(L)), 1 if eps is None else len(eps), engine._ptr(e),
                                  engine._ptr(s), n, bufs[0].p, bufs[1].p, bufs[2].p,
                                  dt.p if dt else None, di.p if di else None, fo[0].p, fo[1].p, fo[2].p, dg.p,
                                  stream)
    assert st == engine.OK, st
    assert h.hipStreamSynchronize(stream) == 0
    f = {k: b.get(np.float64, n) for k, b in zip(("fx", "fy", "fz"), fo)}
    raw = dg.get(np.uint8, ctypes.sizeof(engine.Globals))
    g = engine.Globals.from_buffer_copy(raw.tobytes())
    for b in bufs + fo + [dg] + [b for b in (dt, di) if b]:
        b.free()
    return f, g


def context(L, x, y, z, type_, id_, eps, sig):
    with ap.Engine((0.0, 0.0, 0.0), (L, L, L), RC, SKIN) as e:
        if eps is not None:
            e.set_types(eps, sig, [1.0] * len(eps))
        e.upload(x, y, z, type=type_, id=id_)
        g = e.force(ap.KERNEL_LC_FP64, globals=True)
        return e.download(), g


def system(n_side, density, seed, ntypes):
    a = density ** (-1.0 / 3.0)
    L = n_side * a
    x, y, z, _, _, _, t = ap.generate_lattice(seed, n_side ** 3, (n_side,) * 3, a, (0.0, 0.0, 0.0), 0.05, 0.9, ntypes)
    # a few particles just outside the box on either side: both paths wrap them the same way
    x[:5] += L
    y[5:10] -= L
    return L, x, y, z, (t if ntypes > 1 else None)


def test_noalloc_matches_context_two_calls():
    h = hip()
    stream = ctypes.c_void_p()
    assert h.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # first call: 24^3 particles, two types (Lorentz-Berthelot mixing), ids a permutation, own stream
        announce("no-alloc LC force, call 1: 24^3 lattice rho 0.75, 2 types, permuted ids, caller's stream")
        L, x, y, z, t = system(24, 0.75, 3, 2)
        ids = np.random.default_rng(4).permutation(len(x)).astype(np.int64) + 100
        eps, sig = [1.0, 1.2], [1.0, 1.05]
        f1, g1 = noalloc(h, L, x, y, z, t, ids, eps, sig, stream)
        r1, c1 = context(L, x, y, z, t, ids, eps, sig)
        for k in ("fx", "fy", "fz"):
            assert np.array_equal(f1[k], r1[k]), k
        assert g1.pair_count == c1.pair_count and g1.pair_count > 0
        assert g1.potential_energy == c1.potential_energy
        assert g1.virial == c1.virial
        # second call: a smaller system in a smaller box, one type, no types or ids, default stream; the
        # scratch array still holds the first call's contents
        announce("no-alloc LC force, call 2: 11^3 lattice rho 0.8, 1 type, default stream")
        L2, x2, y2, z2, _ = system(11, 0.8, 7, 1)
        f2, g2 = noalloc(h, L2, x2, y2, z2, None, None, None, None, None)
        r2, c2 = context(L2, x2, y2, z2, None, None, None, None)
        for k in ("fx", "fy", "fz"):
            assert np.array_equal(f2[k], r2[k]), k
        assert g2.pair_count == c2.pair_count and g2.pair_count > 0
        assert g2.potential_energy == c2.potential_energy
        assert g2.virial == c2.virial
    finally:
        h.hipStreamDestroy(stream)


def test_noalloc_refuses_oversized_input():
    announce("no-alloc LC force: sizes beyond the scratch limits are refused on the host")
    lib = engine.load_library()
    dummy = ctypes.c_void_p(1 << 20)  # never dereferenced: the call returns before any device work
    n = 524288 + 1
    st = lib.ljg_force_lc_noalloc(ctypes.byref(params(100.0)), 1, None, None, n, dummy, dummy, dummy, None, None,
                                  dummy, dummy, dummy, None, None)
    assert st == engine.ERR_CAPACITY
    st = lib.ljg_force_lc_noalloc(ctypes.byref(params(320.0)), 1, None, None, 1000, dummy, dummy, dummy, None, None,
                                  dummy, dummy, dummy, None, None)
    assert st == engine.ERR_CAPACITY  # 116^3 cells with the halo, above the cell limit
    st = lib.ljg_force_lc_noalloc(ctypes.byref(params(2.0)), 1, None, None, 1000, dummy, dummy, dummy, None, None,
                                  dummy, dummy, dummy, None, None)
    assert st == engine.ERR_ARG  # box shorter than rc + skin
