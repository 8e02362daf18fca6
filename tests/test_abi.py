"""CPU checks of the C ABI: librocrand.so loads, exports every entry point that
include/rocrand.h declares, and returns the reference's status codes for
argument errors that are decided on the host (library/src/rocrand.cpp:186-630,
the generator templates' set_order / generate type checks).  Nothing here
launches work on a GPU.
"""
import ctypes
import os
import re
import subprocess

import pytest

from rocrand_amd import rocrand as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rocrand.h")

SUCCESS, NOT_CREATED, TYPE_ERROR, OUT_OF_RANGE = 0, 101, 103, 104


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"ROCRANDAPI\s+(rocrand_\w+)\s*\(", text)))


def test_header_declares_the_reference_abi():
    syms = declared_symbols()
    assert len(syms) == 33
    for s in ["rocrand_create_generator", "rocrand_generate_poisson", "rocrand_set_ordering",
              "rocrand_get_scramble_constants64", "rocrand_create_discrete_distribution"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = R.load()
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(rocrand_\w+)", out))
    for s in declared_symbols():
        assert s in exported, s
        assert hasattr(L, s)


def test_library_exports_every_extension_symbol():
    """The extension headers beside rocrand.h (batched, serving, tuning) declare entry
    points with no reference counterpart; librocrand.so exports every one."""
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(rocrand_\w+)", out))
    inc = os.path.join(ROOT, "include")
    seen = 0
    for name in sorted(os.listdir(inc)):
        if not name.endswith(".h") or name == "rocrand.h":
            continue
        text = open(os.path.join(inc, name)).read()
        for s in sorted(set(re.findall(r"ROCRANDAPI\s+rocrand_status\s+(rocrand_\w+)\s*\(", text))):
            assert s in exported, f"{name}: {s}"
            seen += 1
    assert seen == 13


def test_version():
    assert R.get_version() == 300200
    assert R.load().rocrand_get_version(None) == OUT_OF_RANGE


def test_null_generator_is_not_created():
    L = R.load()
    assert L.rocrand_generate(None, None, 10) == NOT_CREATED
    assert L.rocrand_generate_uniform(None, None, 10) == NOT_CREATED
    assert L.rocrand_set_seed(None, 1) == NOT_CREATED
    assert L.rocrand_set_offset(None, 1) == NOT_CREATED
    assert L.rocrand_set_stream(None, None) == NOT_CREATED
    assert L.rocrand_set_ordering(None, R.ORDERING_PSEUDO_DEFAULT) == NOT_CREATED
    assert L.rocrand_initialize_generator(None) == NOT_CREATED
    assert L.rocrand_generate_poisson(None, None, 10, 1.0) == NOT_CREATED


def _create(t, host=False):
    L = R.load()
    g = ctypes.c_void_p()
    fn = L.rocrand_create_generator_host_blocking if host else L.rocrand_create_generator
    return fn(ctypes.byref(g), t), g


def test_bad_type():
    st, _ = _create(1234)
    assert st == TYPE_ERROR


@pytest.mark.parametrize("t", [R.PHILOX4_32_10, R.THREEFRY4_32_20, R.MRG32K3A, R.XORWOW, R.LFSR113, R.MRG31K3P,
                               R.THREEFRY2_64_20])
def test_host_decided_status_codes(t):
    L = R.load()
    st, g = _create(t)
    assert st == SUCCESS
    try:
        assert L.rocrand_set_seed(g, 0x1234) == SUCCESS
        assert L.rocrand_set_offset(g, 99) == SUCCESS
        # philox4x32_10.hpp:258-278: SEEDED and QUASI orderings are rejected
        assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_SEEDED) == OUT_OF_RANGE
        assert L.rocrand_set_ordering(g, R.ORDERING_QUASI_DEFAULT) == OUT_OF_RANGE
        for o in [R.ORDERING_PSEUDO_DEFAULT, R.ORDERING_PSEUDO_BEST, R.ORDERING_PSEUDO_LEGACY,
                  R.ORDERING_PSEUDO_DYNAMIC]:
            assert L.rocrand_set_ordering(g, o) == SUCCESS
        # lambda <= 0 (rocrand.cpp:371-374)
        assert L.rocrand_generate_poisson(g, None, 10, 0.0) == OUT_OF_RANGE
        assert L.rocrand_generate_poisson(g, None, 10, -1.0) == OUT_OF_RANGE
        # quasi-only setter
        assert L.rocrand_set_quasi_random_generator_dimensions(g, 4) == TYPE_ERROR
        # 64-bit output from a 32-bit engine (philox4x32_10.hpp:359-375)
        if t != R.THREEFRY2_64_20:
            assert L.rocrand_generate_long_long(g, None, 10) == TYPE_ERROR
        # uint4 seed only for LFSR113
        want = SUCCESS if t == R.LFSR113 else TYPE_ERROR
        assert L.rocrand_set_seed_uint4(g, R._uint4(1, 2, 3, 4)) == want
    finally:
        assert L.rocrand_destroy_generator(g) == SUCCESS


def test_host_generator_rejects_dynamic_ordering():
    L = R.load()
    st, g = _create(R.PHILOX4_32_10, host=True)
    assert st == SUCCESS
    assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_DYNAMIC) == OUT_OF_RANGE
    assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_DEFAULT) == SUCCESS
    L.rocrand_destroy_generator(g)


def test_counter_generate_with_null_output_only_initialises():
    # philox4x32_10.hpp:323-326: data == nullptr returns SUCCESS after init, no launch.
    L = R.load()
    st, g = _create(R.PHILOX4_32_10)
    assert L.rocrand_generate(g, None, 1000) == SUCCESS
    L.rocrand_destroy_generator(g)


def test_discrete_distribution_argument_errors():
    L = R.load()
    assert L.rocrand_create_poisson_distribution(1.0, None) == OUT_OF_RANGE
    d = ctypes.c_void_p()
    assert L.rocrand_create_poisson_distribution(0.0, ctypes.byref(d)) == OUT_OF_RANGE
    assert L.rocrand_create_discrete_distribution(None, 0, 0, ctypes.byref(d)) == OUT_OF_RANGE
    assert L.rocrand_destroy_discrete_distribution(None) == OUT_OF_RANGE


def test_mt_generators_host_decided_status_codes():
    L = R.load()
    for t in (R.MTGP32, R.MT19937):
        st, g = _create(t)
        assert st == SUCCESS
        assert L.rocrand_set_seed(g, 77) == SUCCESS
        assert L.rocrand_set_offset(g, 5) == TYPE_ERROR  # mtgp32.hpp:398-403, mt19937.hpp:757-762
        assert L.rocrand_generate_long_long(g, None, 10) == TYPE_ERROR
        assert L.rocrand_set_ordering(g, R.ORDERING_QUASI_DEFAULT) == OUT_OF_RANGE
        assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_LEGACY) == SUCCESS
        # MT19937 accepts DEFAULT, BEST and LEGACY only (mt19937.hpp:765-779)
        want = OUT_OF_RANGE if t == R.MT19937 else SUCCESS
        assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_DYNAMIC) == want
        assert L.rocrand_destroy_generator(g) == SUCCESS


@pytest.mark.parametrize("t", [R.SOBOL32, R.SCRAMBLED_SOBOL32, R.SOBOL64, R.SCRAMBLED_SOBOL64, R.QUASI_DEFAULT])
def test_sobol_host_decided_status_codes(t):
    L = R.load()
    st, g = _create(t)
    assert st == SUCCESS
    try:
        assert L.rocrand_set_seed(g, 5) == SUCCESS  # ignored (sobol.hpp:581-584)
        assert L.rocrand_set_seed_uint4(g, R._uint4(1, 2, 3, 4)) == TYPE_ERROR
        assert L.rocrand_set_offset(g, 1000) == SUCCESS
        assert L.rocrand_set_ordering(g, R.ORDERING_PSEUDO_DEFAULT) == OUT_OF_RANGE
        assert L.rocrand_set_ordering(g, R.ORDERING_QUASI_DEFAULT) == SUCCESS
        assert L.rocrand_set_quasi_random_generator_dimensions(g, 0) == OUT_OF_RANGE
        assert L.rocrand_set_quasi_random_generator_dimensions(g, 20001) == OUT_OF_RANGE
        assert L.rocrand_set_quasi_random_generator_dimensions(g, 20000) == SUCCESS
        # n % dimensions (sobol.hpp:628-631), decided before any launch
        assert L.rocrand_set_quasi_random_generator_dimensions(g, 3) == SUCCESS
        assert L.rocrand_generate(g, None, 10) == 105
        if t in (R.SOBOL32, R.SCRAMBLED_SOBOL32, R.QUASI_DEFAULT):
            assert L.rocrand_generate_long_long(g, None, 9) == TYPE_ERROR
    finally:
        assert L.rocrand_destroy_generator(g) == SUCCESS


def test_dynamic_grid_override_status_codes():
    """rocrand_amd_set_dynamic_grid decides on the host (include/rocrand_amd_tuning.h)."""
    L = R.load()
    assert L.rocrand_amd_set_dynamic_grid(R.MT19937, 256, 256) == TYPE_ERROR
    assert L.rocrand_amd_set_dynamic_grid(R.MRG32K3A, 100, 3) == OUT_OF_RANGE
    assert L.rocrand_amd_set_dynamic_grid(R.MRG32K3A, 256, 1 << 15) == OUT_OF_RANGE  # above 2^22
    assert L.rocrand_amd_set_dynamic_grid(R.MTGP32, 256, 256) == OUT_OF_RANGE
    assert L.rocrand_amd_set_dynamic_grid(R.PHILOX4_32_10, 256, 768) == OUT_OF_RANGE
    assert L.rocrand_amd_set_dynamic_grid(R.THREEFRY4_64_20, 256, 3) == OUT_OF_RANGE
    assert L.rocrand_amd_set_dynamic_grid(R.MRG32K3A, 256, 768) == SUCCESS  # engine counts need not be
    assert L.rocrand_amd_set_dynamic_grid(R.MRG32K3A, 0, 0) == SUCCESS
    assert L.rocrand_amd_set_dynamic_grid(R.PHILOX4_32_10, 256, 2048) == SUCCESS
    assert L.rocrand_amd_set_dynamic_grid(R.PHILOX4_32_10, 0, 0) == SUCCESS
