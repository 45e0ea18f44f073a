"""The C ABI library loads on a CPU-only host, exports every function include/paracfd.h
declares, and fails loudly (no fallback) when asked for a device solver without a GPU."""
import ctypes as C
import os
import re

import pytest

from paracfd_amd import _lib as L, cases

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "paracfd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcfd_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(L.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pcfd_abi_version() == 5


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g, d = cases.sphere_box(cells_x=16)
    topo = L.Topology(g, d)
    lib = L.load()
    o = L.Options()
    lib.pcfd_default_options(C.byref(o))
    h = C.c_void_p()
    rc = lib.pcfd_create(topo.handle, C.byref(o), C.byref(h))
    assert rc == L.PCFD_EDEVICE
    assert not h.value
    assert b"device" in lib.pcfd_last_error()


def test_invalid_topology_arguments():
    g, d = cases.sphere_box(cells_x=16)
    d.levels = 7
    with pytest.raises(L.PcfdError) as e:
        L.Topology(g, d)
    assert e.value.status == L.PCFD_EINVAL
    g, d = cases.sphere_box(cells_x=16)
    g.radius = -1.0
    with pytest.raises(L.PcfdError):
        L.Topology(g, d)


def test_body_touching_boundary_is_rejected():
    g, d = cases.sphere_box(cells_x=16)
    g.center[:] = (0.1, 1.0, 1.0)  # sphere crosses x-min
    with pytest.raises(L.PcfdError) as e:
        L.Topology(g, d)
    assert b"boundary" in str(e.value).encode() or "boundary" in str(e.value)


def test_null_arguments_are_rejected():
    lib = L.load()
    assert lib.pcfd_topology_build(None, None, None) == L.PCFD_EINVAL
    assert lib.pcfd_topology_array_size(None, 0) == -1
    assert lib.pcfd_create(None, None, None) == L.PCFD_EINVAL
    assert lib.pcfd_step(None, None) == L.PCFD_EINVAL
