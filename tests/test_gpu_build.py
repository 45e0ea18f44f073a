"""Device-side topology preprocessing (pcfd_topology_build_on_device) against the host builder
(pcfd_topology_build): every exported array is bit-identical -- bricks and their graph, cell
masks and centroids, cut-cell volumes, compact edges with their conductances, interpolation
weights and face centroids, wall patches with their area vectors and moments, every multigrid
level and the dense coarse inverse, and the non-orthogonal records built on top -- for the
sphere box, a 2- and a 3-level NACA wing, a 3-level paraglider and a sphere box without an
outlet (the largest-component rule).  The build stage times are printed.
"""
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


def cases_to_check():
    from paracfd_amd import cases
    return {
        "sphere_box_64": cases.sphere_box(cells_x=64),
        "sphere_box_32_no_outlet": cases.sphere_box(cells_x=32, outlet=False),
        "naca2412_2lvl_48": cases.naca_wing(cells_x=48, levels=2, bands=(0.3,)),
        "naca2412_3lvl_40": cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=40, levels=3, bands=(1.0, 0.25)),
        "paraglider_3lvl_40": cases.paraglider(cells_x=40, levels=3, bands=(1.0, 0.25)),
    }


def assert_identical(host, dev, name):
    from paracfd_amd import _lib as L
    for field, _ in L.TopologyInfo._fields_:
        a, b = getattr(host.info, field), getattr(dev.info, field)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), (name, field)
        else:
            assert a == b, (name, field)
    compared = 0
    for key in L.ARR:
        if key.startswith(("nf_", "ng_")):
            continue
        a, b = host.array(key), dev.array(key)
        assert a.dtype == b.dtype and a.shape == b.shape, (name, key)
        assert a.tobytes() == b.tobytes(), (name, key)
        compared += a.size
    for level in range(int(host.info.mg_levels)):
        for key in L.MG_ARR:
            try:
                a = host.mg_array(level, key)
            except KeyError:
                continue
            b = dev.mg_array(level, key)
            assert a.tobytes() == b.tobytes(), (name, level, key)
            compared += a.size
    return compared


@pytest.mark.parametrize("name", list(cases_to_check()))
def test_device_build_is_bit_identical(lib, name):
    from paracfd_amd import _lib as L
    g, d = cases_to_check()[name]
    print(f"[build] {name}: host build", flush=True)
    host = L.Topology(g, d)
    print(f"[build] {name}: device build", flush=True)
    dev = L.Topology(g, d, device=0)
    hms, hdev = host.build_times()
    dms, ddev = dev.build_times()
    assert hdev == -1 and ddev == 0
    assert dms["device_kernel"] > 0
    n = assert_identical(host, dev, name)
    print(f"[build] {name}: {host.n_active} cells, {n} values identical; host {hms['total']:.0f} ms, "
          f"device {dms['total']:.0f} ms (kernels {dms['device_kernel']:.1f} ms)", flush=True)


def test_device_build_nonorthogonal_records_are_identical(lib):
    from paracfd_amd import _lib as L, cases
    g, d = cases.naca_wing(cells_x=48, levels=2, bands=(0.3,))
    host = L.Topology(g, d)
    dev = L.Topology(g, d, device=0)
    host.build_nonorthogonal()
    dev.build_nonorthogonal()
    for key in L.ARR:
        if key.startswith(("nf_", "ng_")):
            assert host.array(key).tobytes() == dev.array(key).tobytes(), key


def test_device_build_rejects_what_it_cannot_stage(lib):
    from paracfd_amd import _lib as L, cases
    g, d = cases.sphere_box(cells_x=32)
    d.subsamples = 17
    with pytest.raises(L.PcfdError) as e:
        L.Topology(g, d, device=0)
    assert e.value.status == L.PCFD_EINVAL and "subsamples" in str(e.value)
    g, d = cases.sphere_box(cells_x=32)
    with pytest.raises(L.PcfdError) as e:
        L.Topology(g, d, device=99)
    assert e.value.status == L.PCFD_EDEVICE
