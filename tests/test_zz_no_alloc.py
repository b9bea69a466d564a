"""The allocation-free wall generator (asph_boundary_generate_device_noalloc)
against asph_boundary_generate_device, bit for bit, on calls in a row that
share the library's fixed scratch array: the 1M dam break's walls (851,860
particles), then the 31k dam break's (37,482), then a small box's, each into
a caller-owned device buffer. The output buffer is refilled with 0xff bytes
before each call, so a record the generator does not write shows up."""
import ctypes as C

import numpy as np
import pytest

from adaptivesph_amd import abi, scene

pytestmark = pytest.mark.gpu

HIP_MEMCPY_DEVICE_TO_HOST = 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _walls(params):
    return params.domain, scene.boundary_spacing(params), params.rest_density, params.base_smoothing_radius


def _reference(domain, spacing, rho0, h):
    lib = abi.load_library()
    n = C.c_uint32(0)
    abi.check(lib, lib.asph_boundary_generate_device(C.byref(domain), spacing, rho0, h, None, C.byref(n)), "count")
    out = np.zeros(n.value, abi.BOUNDARY_DTYPE)
    abi.check(lib, lib.asph_boundary_generate_device(C.byref(domain), spacing, rho0, h, out.ctypes.data,
                                                     C.byref(n)), "generate_device")
    return out


def _noalloc(hip, buf, capacity, domain, spacing, rho0, h):
    lib = abi.load_library()
    assert hip.hipMemset(buf, 0xFF, capacity * abi.BOUNDARY_DTYPE.itemsize) == 0
    n = C.c_uint32(capacity)
    abi.check(lib, lib.asph_boundary_generate_device_noalloc(C.byref(domain), spacing, rho0, h, buf, C.byref(n),
                                                             None), "generate_device_noalloc")
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(n.value, abi.BOUNDARY_DTYPE)
    assert hip.hipMemcpy(out.ctypes.data, buf, out.nbytes, HIP_MEMCPY_DEVICE_TO_HOST) == 0
    return out


def test_noalloc_walls_equal_allocating_generator():
    hip = _hip()
    capacity = 1_500_000  # records of the caller's buffer (30 MB)
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), capacity * abi.BOUNDARY_DTYPE.itemsize) == 0
    try:
        small = scene.dam_break((6, 6, 6), 0.01, (0.3, 0.2, 0.25), seed=3, steps=1, boundary="").params
        expected_counts = (851_860, 37_482, None)
        for params, want in zip((scene.cfg2(boundary="").params, scene.cfg1(boundary="").params, small),
                                expected_counts):
            args = _walls(params)
            print(f"noalloc walls: domain {params.domain.max.x} x {params.domain.max.y} x {params.domain.max.z}",
                  flush=True)
            got = _noalloc(hip, buf, capacity, *args)
            ref = _reference(*args)
            if want is not None:
                assert len(ref) == want
            assert len(got) == len(ref)
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        # the limit is checked on the host, before any GPU work
        lib = abi.load_library()
        d = scene.cfg2(boundary="").params.domain
        n = C.c_uint32(capacity)
        rc = lib.asph_boundary_generate_device_noalloc(C.byref(d), 0.003, 1000.0, 0.004, buf, C.byref(n), None)
        assert rc == abi.ASPH_ERR_CAPACITY and b"at most 1048576" in lib.asph_last_error()
    finally:
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(buf)
