"""dnab_score_batch_noalloc against dnab_score_batch, through the C ABI.

The no-alloc entry takes device buffers (allocated here with hipMalloc, the
caller's job) and a device antigen built with dnab_score_batch's grid (10 A
cells, 10 A cutoff), so both form the same per-(atom, chunk) integer sums:
energies below 1e9 kcal/mol must match bit for bit, larger ones (the f32
result of a double total) to 1e-6 relative. Two calls in a row, the second
with fewer candidates of another length and a NaN coordinate, check that
the shared scratch array is cleared for each call; a call beyond the
array's capacity must be refused.
"""
import ctypes as C

import numpy as np
import pytest

from dnab import DeviceAntigen
from dnab import _native as N
from dnab import workloads as W

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 4))) == 0
        self.n = nbytes

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.n
        assert hip().hipMemcpy(self.p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(ctx, ag, d_atoms, n_atoms, n_ab, d_out, capacity):
    got = C.c_size_t(0)
    st = N.lib.dnab_score_batch_noalloc(ctx.handle, ag.handle, d_atoms.p, n_atoms, n_ab, d_out.p,
                                        capacity, C.byref(got))
    return st, got.value


def check_same(e_ref, e_new):
    assert e_ref.shape == e_new.shape
    assert (np.isnan(e_ref) == np.isnan(e_new)).all()
    m = ~np.isnan(e_ref)
    small = m & (np.abs(e_ref) < 1e9)
    assert np.array_equal(e_ref[small].view(np.uint32), e_new[small].view(np.uint32))
    big = m & ~small
    assert (np.abs(e_new[big] - e_ref[big]) <= 1e-6 * np.abs(e_ref[big])).all()


def test_noalloc_equals_score_batch(ctx):
    w = W.config2(n_cand=640)
    cloud = w.antigens[0].atoms
    ag = DeviceAntigen(ctx, cloud, cutoff=10.0, cell_size=10.0)
    atoms = w.pop.atoms.gpu_atoms()          # (n, 8) float32 GpuAtom rows
    d_atoms = DevBuf(atoms.nbytes)
    d_out = DevBuf(1024 * 4)
    try:
        # call 1: 512 candidates of 160 atoms
        n_ab1, n1 = 160, 512
        a1 = atoms[:n_ab1 * n1 + 37]          # a trailing partial block is ignored
        d_atoms.put(a1)
        ref1 = ctx.score_batch(cloud, a1, n_ab1)
        st, got = noalloc(ctx, ag, d_atoms, len(a1), n_ab1, d_out, 1024)
        assert st == N.DNAB_OK and got == n1
        ctx.synchronize()
        e1 = d_out.get(np.float32, n1)
        check_same(ref1, e1)
        assert np.isfinite(e1).all() and (e1 != 0).any()
        # call 2: fewer, shorter candidates, one NaN coordinate
        n_ab2, n2 = 97, 100
        a2 = atoms[5000:5000 + n_ab2 * n2].copy()
        a2[n_ab2 * 7 + 3, 1] = np.nan
        d_atoms.put(a2)
        ref2 = ctx.score_batch(cloud, a2, n_ab2)
        st, got = noalloc(ctx, ag, d_atoms, len(a2), n_ab2, d_out, 1024)
        assert st == N.DNAB_OK and got == n2
        ctx.synchronize()
        e2 = d_out.get(np.float32, n2)
        check_same(ref2, e2)
        assert np.isnan(e2[7]) and np.isfinite(np.delete(e2, 7)).all()
        # sizes: empty input, too small an output, more candidates than the scratch holds
        assert noalloc(ctx, ag, d_atoms, 0, n_ab2, d_out, 1024) == (N.DNAB_OK, 0)
        assert noalloc(ctx, ag, d_atoms, len(a2), 0, d_out, 1024) == (N.DNAB_OK, 0)
        assert noalloc(ctx, ag, d_atoms, len(a2), n_ab2, d_out, 10)[0] == N.DNAB_ERR_CAPACITY
        assert noalloc(ctx, ag, d_atoms, 65537, 1, d_out, 1 << 20)[0] == N.DNAB_ERR_UNSUPPORTED
    finally:
        d_atoms.free()
        d_out.free()
        ag.close()
