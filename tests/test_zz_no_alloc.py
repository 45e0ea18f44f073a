"""pcfd_sample_slice_static (the slice sampler staged in a fixed device array, no allocation)
through the C ABI against pcfd_sample_slice: bit-identical images on two calls in a row, the
second smaller than the first, on a moving state; sizes beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import _lib as L
    return L.load()


def slice_static(lib, core, field, axis, coord, lo, hi, nu, nv, instance=0):
    from paracfd_amd import _lib as L
    out = np.full(nu * nv, -7.0, dtype=np.float32)
    lo2, hi2 = (C.c_double * 2)(*lo), (C.c_double * 2)(*hi)
    L.check(lib.pcfd_sample_slice_static(core.handle, instance, field, axis, coord, lo2, hi2, nu, nv, L.fptr(out)))
    return out.reshape(nv, nu)


def test_static_slices_equal_the_allocating_entry_point(lib):
    from paracfd_amd import ExternalAeroCore, _lib as L, cases, default_options
    g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=32, levels=2, bands=(0.3,))
    topo = L.Topology(g, d)
    core = ExternalAeroCore(topo, default_options(rho=1.0, nu=1e-4, speed=1.0, field_precision=L.FP32,
                                                  vcycle_precision=L.FP32))
    assert core.initialize()["converged"]
    calls = [  # (field, axis, coord, lo, hi, nu, nv): the second call is the smaller image
        (0, 1, 0.3, (-1.0, -1.0), (2.5, 1.2), 640, 400),
        (4, 2, 0.05, (-0.5, -1.2), (1.5, 1.2), 96, 80),
        (5, 0, 0.4, (-1.1, -0.6), (1.1, 0.8), 1024, 1024),
        (3, 1, -0.2, (-1.0, -0.5), (2.0, 0.7), 7, 3),
    ]
    for k, (field, axis, coord, lo, hi, nu, nv) in enumerate(calls):
        if k == 2:
            core.step()     # a new state between the pairs
        print(f"[no-alloc] call {k}: field {field} axis {axis} {nu}x{nv}", flush=True)
        a = slice_static(lib, core, field, axis, coord, lo, hi, nu, nv)
        b = core.sample_slice(field, axis, coord, lo, hi, nu, nv)
        assert a.shape == b.shape
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
        assert np.isfinite(a).sum() > 0 and (a != -7.0).all()
    # beyond the fixed array: refused before any launch
    with pytest.raises(L.PcfdError) as e:
        slice_static(lib, core, 0, 1, 0.3, (-1.0, -1.0), (2.5, 1.2), 1025, 1024)
    assert e.value.status == L.PCFD_EINVAL and "pixels" in str(e.value)
    core.close()


def test_two_solvers_share_the_device_array(lib):
    """Two solvers on the same GPU, called one after the other, each get their own image."""
    from paracfd_amd import ExternalAeroCore, _lib as L, cases, default_options
    cores = []
    for r in (0.25, 0.3):
        g, d = cases.sphere_box(cells_x=32)
        g.radius = r
        cores.append(ExternalAeroCore(L.Topology(g, d), default_options()))
        assert cores[-1].initialize()["converged"]
    imgs = [slice_static(lib, c, 4, 2, 1.0, (0.0, 0.0), (4.0, 2.0), 256, 128) for c in cores]
    refs = [c.sample_slice(4, 2, 1.0, (0.0, 0.0), (4.0, 2.0), 256, 128) for c in cores]
    for a, b in zip(imgs, refs):
        assert np.array_equal(np.nan_to_num(a, nan=1e30), np.nan_to_num(b, nan=1e30))
    assert not np.array_equal(np.isnan(imgs[0]), np.isnan(imgs[1]))
    for c in cores:
        c.close()
