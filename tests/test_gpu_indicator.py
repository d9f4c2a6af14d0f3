"""GPU: wall collisions (ibx_collisions_dev), the indicator's point shift (ibx_shift_z_dev)
and the indicator function (fld_indicator_*) against the CPU oracle
(oracle/indicator_oracle.c).

Tolerances.  The shift, the divergence and every stage after the solve are the reference's
expressions without fused multiply-adds: bit for bit on the same input.  The collision force
uses the device's exponential: each point within 1e-14 of the oracle's force (relative).  The
whole indicator goes through this library's FFT solve: indi within 1e-12 of its largest value
(O(1)); indic is then bit for bit where it is thresholded to 0 or 1 and within 1e-12 elsewhere,
the viscosity and the forces within 1e-12 of their scales; cells whose oracle value lies
within 1e-9 of a threshold (.01, .9), where rounding may decide the side, are left out.
"""
import numpy as np
import pytest

import _oracle as O
from suguru_amd import params as PR
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to("cuda:0")


def host(t, g):
    return t.cpu().numpy().reshape(g.shape_f, order="F")


def bits(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint64)


def test_collisions_match_oracle(gpu_ctx_factory):
    torch = _torch()
    g = W.make_grid(32, 32, 32, 1.0 / 32)
    ctx = gpu_ctx_factory(g)
    r0 = PR.rad0(g.dz)
    rng = np.random.default_rng(8)
    n = 20000
    z = np.concatenate([rng.uniform(0.0, 1.2, n // 2) * r0, rng.uniform(0.0, g.lz, n // 2)])
    z[:4] = [1e-3 * g.dz, r0, 2.0 * r0, 0.5 * g.dz]  # (z = 0 on the wall gives 0/0 in the reference)
    area = rng.uniform(0.2, 3.0, n) * g.dx * g.dx
    fz = rng.standard_normal(n)
    fz[4] = -0.0
    print("case: collisions, 20000 points", flush=True)
    want = fz.copy()
    a2 = PR.collision_factor()
    O.collisions(g, z, area, want, a2, r0, PR.BETACOL)
    tz, ta, tf = dev(torch, z), dev(torch, area), dev(torch, fz)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.collisions_dev(tz, ta, tf, a2, r0, PR.BETACOL)
    ctx.sync()
    got = tf.cpu().numpy()
    force = want - fz
    err = np.abs((got - fz) - force)
    assert np.all(err <= 1e-14 * np.abs(force) + 1e-300), float((err / np.abs(force + 1e-300)).max())
    # no force at or above the cutoff: those points keep their bits (-0 + 0 -> +0 as in Fortran)
    far = z >= r0
    assert np.array_equal(bits(got[far]), bits(want[far]))
    ctx.set_stream(0)


def test_shift_z_round_trip_bitwise(gpu_ctx_factory):
    torch = _torch()
    g = W.make_grid(32, 32, 64, 1.0 / 32)
    ctx = gpu_ctx_factory(g)
    rng = np.random.default_rng(2)
    z = rng.uniform(0, g.lz, 5000)
    s = (g.ktot // 4) * g.dz
    tz = dev(torch, z)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.shift_z_dev(tz, s)
    ctx.sync()
    assert np.array_equal(bits(tz.cpu().numpy()), bits(z + s))
    ctx.shift_z_dev(tz, -s)
    ctx.sync()
    assert np.array_equal(bits(tz.cpu().numpy()), bits((z + s) - s))
    ctx.set_stream(0)


@pytest.fixture(scope="module")
def flow():
    from suguru_amd.fluid import FlowSolver

    made = {}

    def get(n):
        if n not in made:
            g = W.make_grid(n, n, n, 1.0 / n)
            made[n] = (g, FlowSolver(g, PR.REP, 1e-4, 0))
        return made[n]

    yield get
    for _, s in made.values():
        s.close()


def _sphere(g, R_cells, zc_cells, degree=16):
    c = np.array([0.43, 0.57, zc_cells * g.dz])
    R = R_cells * g.dx
    caps = W.build_capsules([c], [R], [degree])
    nx, ny, nz = (c[k] - p for k, p in enumerate((caps.x, caps.y, caps.z)))
    nrm = np.sqrt(nx * nx + ny * ny + nz * nz)
    return caps, nx / nrm, ny / nrm, nz / nrm


def test_indicator_stages_bitwise(flow):
    torch = _torch()
    g, fs = flow(32)
    rng = np.random.default_rng(6)
    gx, gy, gz = (np.asfortranarray(rng.standard_normal(g.shape_f)) for _ in range(3))
    print("case: indicator divergence 32^3", flush=True)
    d_indi = torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    fs.indicator_div(dev(torch, gx), dev(torch, gy), dev(torch, gz), d_indi)
    fs.sync()
    want = O.indicator_div(g, gx, gy, gz)
    assert np.array_equal(bits(host(d_indi, g)[1:-1, 1:-1, 1:-1]), bits(want[1:-1, 1:-1, 1:-1]))
    print("case: indicator finish 32^3", flush=True)
    indi = np.asfortranarray(rng.choice([0.0, 0.005, 0.01, 0.3, 0.9, 0.95, 1.02], size=g.shape_f)
                             + 1e-3 * rng.standard_normal(g.shape_f))
    fy = np.asfortranarray(rng.standard_normal(g.shape_f))
    fz = np.asfortranarray(rng.standard_normal(g.shape_f))
    st, ct, tc = PR.centrifugal()
    lam = 2.5
    d = {k: dev(torch, a) for k, a in (("indi", indi), ("fy", fy), ("fz", fz))}
    d["indic"] = torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    d["visc"] = torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    fs.indicator_finish(d["indi"], d["indic"], d["visc"], d["fy"], d["fz"], lam, PR.VISC0, st, ct,
                        tc)
    fs.sync()
    w_indi = indi.copy(order="F")
    w_fy, w_fz = fy.copy(order="F"), fz.copy(order="F")
    w_indic, w_visc = O.indicator_finish(g, w_indi, w_fy, w_fz, lam, PR.VISC0, st, ct, tc)
    for nm, a, b in (("indi", d["indi"], w_indi), ("indic", d["indic"], w_indic),
                     ("visc", d["visc"], w_visc), ("forcey", d["fy"], w_fy),
                     ("forcez", d["fz"], w_fz)):
        assert np.array_equal(bits(host(a, g)), bits(b)), nm


def test_indicator_whole_matches_oracle(flow, gpu_ctx_factory):
    """Shift, cell-centred spread of the normals, unshift (libibx), then fld_indicator_dev,
    against the oracle's whole routine, for a sphere near the bottom wall."""
    torch = _torch()
    g, fs = flow(64)
    ctx = gpu_ctx_factory(g)
    caps, nx, ny, nz = _sphere(g, 6.0, 9.1)
    st, ct, tc = PR.centrifugal()
    rng = np.random.default_rng(1)
    fy = np.asfortranarray(rng.standard_normal(g.shape_f))
    fz = np.asfortranarray(rng.standard_normal(g.shape_f))
    print("case: indicator, one sphere, 64^3", flush=True)
    wz = caps.z.copy()
    w_fy, w_fz = fy.copy(order="F"), fz.copy(order="F")
    wgx, wgy, wgz, w_indi, w_indic, w_visc = O.indicator(
        g, PR.REP, 1e-4, caps.cap_off, caps.x, caps.y, wz, nx, ny, nz, caps.area, w_fy, w_fz,
        PR.LAMBDA, PR.VISC0, st, ct, tc)
    # one stream for both libraries (torch's default stream is handle 0, which would leave
    # each library on its own stream)
    ts = torch.cuda.Stream()
    stream = ts.cuda_stream
    ctx.set_stream(stream)
    fs.set_stream(stream)
    off = torch.from_numpy(caps.cap_off).to("cuda:0")
    x, y, z, tnx, tny, tnz, area = (dev(torch, a) for a in (caps.x, caps.y, caps.z, nx, ny, nz,
                                                            caps.area))
    e = lambda: torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    G = [e(), e(), e()]
    indi, indic, visc = e(), e(), e()
    tfy, tfz = dev(torch, fy), dev(torch, fz)
    s = (g.ktot // 4) * g.dz
    ctx.shift_z_dev(z, s)
    ctx.spread_dev(off, x, y, z, tnx, tny, tnz, area, *G, cc=True)
    ctx.shift_z_dev(z, -s)
    fs.indicator(*G, indi, indic, visc, tfy, tfz)
    ctx.sync()
    fs.sync()
    assert np.array_equal(bits(z.cpu().numpy()), bits(wz))
    for a, b, nm in zip(G, (wgx, wgy, wgz), "xyz"):
        assert np.array_equal(bits(host(a, g)), bits(b)), "G" + nm
    gi = host(indi, g)
    assert np.abs(gi - w_indi).max() <= 1e-12 * np.abs(w_indi).max()
    # thresholded results: bitwise away from the thresholds
    pre = np.zeros(g.shape_f)
    kq = g.ktot // 4
    pre[1:-1, 1:-1, 1:g.ktot - kq + 1] = w_indi[1:-1, 1:-1, 1 + kq:g.ktot + 1]
    near = (np.abs(pre - 0.01) < 1e-9) | (np.abs(pre - 0.9) < 1e-9)
    assert near.sum() < 10
    ok = ~near
    ok[:, :-1, :] &= ~near[:, 1:, :]
    ok[:, :, :-1] &= ~near[:, :, 1:]
    inner = np.zeros(g.shape_f, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    m = ok & inner
    gic = host(indic, g)
    # 0 and 1 (the thresholded cells) bit for bit; the values in between carry the solve's
    # rounding, as do the halos (copies from before the thresholds)
    exact = m & ((w_indic == 0.0) | (w_indic == 1.0))
    assert exact.sum() > 0.9 * m.sum()
    assert np.array_equal(bits(gic[exact]), bits(w_indic[exact]))
    assert np.abs(gic[m] - w_indic[m]).max() <= 1e-12
    assert np.abs(gic[~inner] - w_indic[~inner]).max() <= 1e-12
    gv = host(visc, g)
    assert np.array_equal(bits(gv[exact]), bits(w_visc[exact]))
    assert np.abs(gv[m] - w_visc[m]).max() <= 1e-12 * PR.VISC0
    gfy, gfz = host(tfy, g), host(tfz, g)
    scale = tc * 2.0
    assert np.abs(gfy[m] - w_fy[m]).max() <= 1e-12 * scale
    assert np.abs(gfz[m] - w_fz[m]).max() <= 1e-12 * scale
    ctx.set_stream(0)
    fs.set_stream(0)
