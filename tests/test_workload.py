"""CPU: the synthetic workloads' geometry (capsule counts, points, volume fraction, spacing),
and the host-side checks of device tensors.

The benchmark and the GPU tests regenerate these inputs from seeds on every machine, so the
figures they are quoted on are pinned here.
"""
import numpy as np

from suguru_amd import workload as W


def _min_centre_distance(g, c):
    d = c[:, None, :] - c[None, :, :]
    d[..., 0] -= g.lx * np.round(d[..., 0] / g.lx)
    d[..., 1] -= g.ly * np.round(d[..., 1] / g.ly)
    dist = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(dist, np.inf)
    return dist.min()


def test_dense_share_geometry():
    wl = W.config_dense_share(with_fields=False)
    g, caps = wl.grid, wl.caps
    assert (g.itot, g.jtot, g.ktot) == (128, 512, 256)
    assert caps.radii.size == 1000 and caps.npts == 578_000
    assert abs(W.volume_fraction(g, caps) - 0.2974) < 1e-3
    r = caps.radii[0]
    # no overlap (gap >= 1h under the periodic image) and wall clearance >= r + 1h
    assert _min_centre_distance(g, caps.centers) >= 2 * r + g.dx * (1 - 1e-9)
    assert caps.centers[:, 2].min() >= r + g.dx * (1 - 1e-9)
    assert caps.centers[:, 2].max() <= g.lz - r - g.dx * (1 - 1e-9)
    # areas: the Gauss grid integrates the sphere exactly
    a = caps.area[: caps.cap_off[1]].sum()
    assert abs(a - 4 * np.pi * r * r) < 1e-12 * a


def test_uneven_geometry():
    wl = W.config_uneven(with_fields=False)
    g, caps = wl.grid, wl.caps
    assert (g.itot, g.jtot, g.ktot) == (512, 512, 256)
    assert caps.radii.size == 2000 and caps.npts == 1_161_280
    assert abs(W.volume_fraction(g, caps) - 0.0502) < 1e-3
    assert _min_centre_distance(g, caps.centers) >= (2 * caps.radii.min())


def test_lattice_rejects_overfull():
    g = W.make_grid(64, 64, 64)
    try:
        W.lattice_capsules(g, 10 * g.dx, (4, 4, 4), 1)
    except ValueError:
        return
    raise AssertionError("a lattice denser than 2 r + gap was accepted")


def test_device_tensor_check_rejects_before_any_library_call():
    """The Python device wrappers check a tensor's placement, dtype, contiguity and size
    before handing its raw pointer to the library (which cannot); on a CPU-only machine
    every tensor is rejected for its placement."""
    import pytest
    import torch

    from suguru_amd.interp_spread import _dev_check

    with pytest.raises(ValueError, match="on the GPU"):
        _dev_check("x", torch.zeros(8, dtype=torch.float64), 8)
    with pytest.raises(ValueError, match="torch.float64"):
        _dev_check("x", torch.zeros(8, dtype=torch.float32), 8)
    with pytest.raises(ValueError, match="torch.int64"):
        _dev_check("cap_off", torch.zeros(3, dtype=torch.int32), 3, integer=True)
    with pytest.raises(ValueError):
        _dev_check("x", [0.0] * 8, 8)
