"""The synthetic lattice generator: the windowed spherical-infall update equals the update over
every vertex."""

import numpy as np

import adeco_amd.lattice as L


def _infall_everywhere(px, py, pz, box, count, rng):
    centres = rng.uniform(0.0, box, size=(count, 3))
    radii = rng.uniform(2.0, 8.0, size=count)
    amps = rng.uniform(0.8, 1.6, size=count)
    pos = [px.copy(), py.copy(), pz.copy()]
    for c, r, a in zip(centres, radii, amps):
        d = [pos[i] - c[i] for i in range(3)]
        d = [di - box * np.round(di / box) for di in d]
        rr = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) / r
        w = np.where(rr < 1.0, (1.0 - rr * rr) ** 2, 0.0)
        for i in range(3):
            pos[i] = pos[i] - a * d[i] * w
    return pos


def test_windowed_infall_is_exact(monkeypatch):
    for n in (24, 40):
        fast = L.make_lattice(n, 4.0, seed=3, infall_centres=64)
        monkeypatch.setattr(L, "_apply_infall", _infall_everywhere)
        slow = L.make_lattice(n, 4.0, seed=3, infall_centres=64)
        monkeypatch.undo()
        assert np.array_equal(fast.x, slow.x) and np.array_equal(fast.y, slow.y) and np.array_equal(fast.z, slow.z)


def test_lattice_layout():
    lat = L.make_lattice(6, 0.0)
    assert lat.x.size == 7 ** 3 and lat.num_tets == 6 * 6 ** 3
    x = lat.x.reshape(7, 7, 7)
    assert np.allclose(x[0, 0], np.arange(7) * 100.0 / 6)  # i fastest
