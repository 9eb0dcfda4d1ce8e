"""Closed forms on degenerate 1 x 1 x N lattices (every pull in x and y wraps
onto the node itself), the reference's planar examples replayed on the GPU in
FP64:

  examples/planar_couette.pl    walls at z = 0 and h + 1 moving at v1, v2 in a
                                periodic box of 2(h + 1): each wall carries
                                the exact force 2 eta (v1 - v2) / h, opposite
                                on the two walls
  examples/planar_poiseuille.pl walls at z = 0 and 19, uniform force dpdx on
                                the 18 fluid nodes: each wall carries
                                dpdx h / 2 at steady state (momentum balance)

mylb->dpdx is a field of the public struct here; the walls are set moving with
vspread after their velocities change, as the reference's moving-wall step
requires."""
import pytest

pytestmark = pytest.mark.gpu


def lattice(esz):
    import sunlightlb_amd as s
    return s.Lattice(1, 1, esz, precision=64)


def test_planar_couette_wall_forces():
    v1, v2, h, tau = 0.11, -0.07, 10, 0.80
    lat = lattice(2 * (h + 1))
    L, hh = lat.L, lat.h
    lat.s.stokes_flow, lat.s.objects_moving, lat.s.tau = 1, 1, tau
    w1 = L.insertwall(hh, b"z", 0).contents
    w2 = L.insertwall(hh, b"z", h + 1).contents
    w1.velx, w2.velx = v1, v2
    L.vspread(hh)
    lat.step(5000)
    L.measure(hh)
    eta = 72.0 * (tau - 0.5) / 3.0
    f = 2.0 * eta * (v1 - v2) / h
    f1, f2 = w1.frcx, w2.frcx
    lat.free()
    print(f"Couette: expected |f| {f:.12g}, measured {f1:.12g} {f2:.12g}")
    assert abs(abs(f1) - f) <= 1e-10 * f and abs(abs(f2) - f) <= 1e-10 * f
    assert f1 * f2 < 0 and f1 * (v1 - v2) < 0   # the fluid holds the faster wall back


def test_planar_poiseuille_wall_forces():
    esz, tau, umean = 20, 0.80, 0.1
    eta = 72.0 * (tau - 0.5) / 3.0
    dpdx = 12 * eta * umean / ((esz - 2) * (esz - 2))
    lat = lattice(esz)
    L, hh = lat.L, lat.h
    lat.s.stokes_flow, lat.s.tau = 1, tau
    lat.s.dpdx = dpdx
    w1 = L.insertwall(hh, b"z", 0).contents
    w2 = L.insertwall(hh, b"z", esz - 1).contents
    lat.step(20000)
    L.measure(hh)
    h = w2.posz - w1.posz - 1
    f = dpdx * h / 2
    f1, f2 = w1.frcx, w2.frcx
    lat.free()
    print(f"Poiseuille: expected {f:.12g}, measured {f1:.12g} {f2:.12g}")
    assert h == esz - 2
    assert abs(f1 - f) <= 1e-10 * f and abs(f2 - f) <= 1e-10 * f
