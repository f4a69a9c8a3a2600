"""madam_tri_separate_noalloc -- triangle separation on device buffers that
allocates nothing -- against madam_tri_separate: the same list, values and
maximum, bit for bit, on two calls in a row (the second smaller), with an
overflowing boundary bin, and sizes beyond the device array's limits
rejected.  The test owns its device buffers (HIP runtime through ctypes)
and runs the call on HIP's default stream."""
import ctypes as C
import os

import numpy as np
import pytest

import madam_amd as G

pytestmark = pytest.mark.gpu


def psd_X(n, seed, rank=5):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((n, rank))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    X = V @ V.T
    np.fill_diagonal(X, 1.0)
    return np.asfortranarray(X)


class Hip:
    """The few HIP runtime calls the test needs to own device buffers
    (the test allocates; the entry point under test does not)."""

    def __init__(self):
        # the HIP runtime the library itself runs on: loading another copy
        # by path can clash with the HSA runtime already in the process
        G.lib()
        try:
            self.h = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)
        except OSError:
            self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)))
        return p

    def h2d(self, dst, arr):
        self.check(self.h.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1))

    def d2h(self, arr, src):
        self.check(self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src, C.c_size_t(arr.nbytes), 2))

    def free(self, p):
        self.check(self.h.hipFree(p))

    def sync(self):
        self.check(self.h.hipDeviceSynchronize())


def tri_separate_noalloc(X, K, thresh=1e-3, list_cap=None):
    hip = Hip()
    n = X.shape[0]
    cap = max(1, K if list_cap is None else list_cap)
    Xh = np.asfortranarray(X, dtype=np.float64).ravel(order="F").copy()
    fill = np.full(cap * 4, -7.0)  # 32-byte madam_tri entries, poisoned
    dX, dL, dS, dM = hip.malloc(Xh.nbytes), hip.malloc(fill.nbytes), hip.malloc(4), hip.malloc(8)
    try:
        hip.h2d(dX, Xh)
        hip.h2d(dL, fill)
        hip.h2d(dS, np.array([-1], dtype=np.int32))
        hip.sync()
        rc = G.lib().madam_tri_separate_noalloc(C.c_int(n), dX, C.c_int(K), C.c_double(thresh),
                                                 dL, dS, dM, C.c_void_p(0))
        if rc != 0:
            return rc, None, None
        hip.sync()
        size = np.zeros(1, dtype=np.int32)
        mx = np.zeros(1)
        out = np.zeros(cap * 4)
        hip.d2h(size, dS)
        hip.d2h(mx, dM)
        hip.d2h(out, dL)
    finally:
        for p in (dX, dL, dS, dM):
            hip.free(p)
    nf = int(size[0])
    arr = (G.Tri * cap).from_buffer_copy(out.tobytes())
    lst = [(arr[q].type, arr[q].i, arr[q].j, arr[q].k, arr[q].value) for q in range(nf)]
    assert all(arr[q].y == 0.0 for q in range(nf))
    return 0, float(mx[0]), lst


def test_noalloc_matches_allocating_entry_on_consecutive_calls():
    G.open_device(0)
    # second call smaller than the first: what the first left in the device
    # array must not leak into it
    for n, K, seed in ((180, 1800, 21), (60, 300, 22), (60, 40, 22)):
        X = psd_X(n, seed)
        rc, mx_a, lst_a = tri_separate_noalloc(X, K)
        assert rc == 0, G.lib().madam_last_error()
        mx_b, lst_b = G.tri_separate(X, K)
        assert mx_a == mx_b
        assert lst_a == lst_b
        assert 0 < len(lst_a) <= K


def test_noalloc_boundary_overflow_matches():
    # every type-1 row of this X equals 0.5: 34,220 rows in the boundary bin
    G.open_device(0)
    X = np.full((60, 60), -0.5)
    np.fill_diagonal(X, 1.0)
    X = np.asfortranarray(X)
    rc, mx_a, lst_a = tri_separate_noalloc(X, 600)
    assert rc == 0, G.lib().madam_last_error()
    mx_b, lst_b = G.tri_separate(X, 600)
    assert (mx_a, lst_a) == (mx_b, lst_b)
    rc, mx_c, lst_c = tri_separate_noalloc(psd_X(40, 23), 100)
    assert rc == 0
    assert (mx_c, lst_c) == G.tri_separate(psd_X(40, 23), 100)


def test_noalloc_rejects_sizes_beyond_the_array():
    G.open_device(0)
    X = psd_X(8, 24)
    rc, _, _ = tri_separate_noalloc(X, 50001, list_cap=1)
    assert rc == -1
    assert "K=50001" in G.lib().madam_last_error().decode()
