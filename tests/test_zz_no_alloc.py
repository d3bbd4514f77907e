"""bl_lattice_conserved_static, the allocation-free conservation sums (a
static device array holds the partial sums), through the C-ABI against
bl_lattice_conserved: two calls in a row, the second on a smaller lattice,
bitwise equal, for every storage; a lattice beyond the static array's
capacity is refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself is linked against (device
    buffers for the output)."""
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


D2H = 2  # hipMemcpyDeviceToHost


def _lattice(shape, storage, seed):
    from biblbm_amd.lattice import Lattice
    rng = np.random.default_rng(seed)
    L = Lattice(*shape, periodic=(1, 0, 1), storage=storage, omega=1 / 0.9)
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    L.set_mask(mask)
    L.init_equilibrium(rng.uniform(0.98, 1.02, shape), rng.uniform(-0.02, 0.02, shape + (3,)))
    L.set_body_force((1e-5, 0, 0))
    L.collide_and_stream(3)
    return L


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
def test_static_sums_equal_allocating_entry_point(device, storage):
    from biblbm_amd import capi
    hip = _hip()
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), 32) == 0
    try:
        for k, shape in enumerate([(48, 40, 36), (20, 18, 16)]):
            L = _lattice(shape, storage, k)
            ref = L.conserved()
            capi.call("bl_lattice_conserved_static", L.handle, d)
            capi.call("bl_sync")
            got = np.zeros(4)
            assert hip.hipMemcpy(got.ctypes.data, d, 32, D2H) == 0
            print(storage, shape, got, ref)
            assert np.array_equal(got, ref)
            L.close()
    finally:
        hip.hipFree(d)


def test_static_sums_refuse_lattice_beyond_capacity(device):
    from biblbm_amd import capi
    from biblbm_amd.lattice import Lattice
    hip = _hip()
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), 32) == 0
    try:
        L = Lattice(4, 4096, 4097, periodic=(1, 1, 1), storage="f16s")  # 67 125 248 cells > 2^26
        rc = capi.lib.bl_lattice_conserved_static(L.handle, d)
        assert rc == 4  # BL_ERR_CAPACITY
        assert b"67108864" in capi.lib.bl_last_error()
        L.close()
    finally:
        hip.hipFree(d)
