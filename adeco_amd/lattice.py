"""Seeded synthetic particle lattices for the mass-deposit path.

A lattice of N^3 Lagrangian cubes is stored as an (N+1)^3 vertex lattice in ColumnMajorOrder
(index i + nx*(j + ny*k), i fastest; the reference's id order, AdTypeDefinitions.h:184-220).  The
last vertex layer is the periodic image of the first, unwrapped by +L, so the lattice holds exactly
6*N^3 tets and the reference's "cells between existing vertices" rule (AdTypeDefinitions.h:656)
sees N^3 cubes with periodic_boundaries=false.

Displacements are a curl-free Gaussian random field s = -grad(phi) whose per-component spectrum is
P_s(k) ~ k^-2 exp(-k^2 sigma^2), sigma = 2 lattice spacings, scaled so the 3-D rms |s| equals the
requested number of lattice spacings.  Optional spherical infall centres pull points towards
seeded centres (halo-like caustics) and seeded near-degenerate tets can be generated as a tet soup.
Positions are float32.
"""

from __future__ import annotations

import dataclasses

import numpy as np

BOX = 100.0  # Mpc/h


@dataclasses.dataclass
class Lattice:
    """An unwrapped (n+1)^3 vertex lattice (SoA float32, i fastest)."""

    n: int  # cubes per axis
    box: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    @property
    def dims(self):
        return (self.n + 1, self.n + 1, self.n + 1)

    @property
    def num_tets(self) -> int:
        return 6 * self.n ** 3

    def bbox(self):
        """Float min/max over all vertices: x0,x1,y0,y1,z0,z1 (AdPointDataReaders.h:90-151)."""
        return np.array([self.x.min(), self.x.max(), self.y.min(), self.y.max(),
                         self.z.min(), self.z.max()], dtype=np.float32)


def _displacement_field(n: int, box: float, rms_spacings: float, sigma_spacings: float,
                        rng: np.random.Generator) -> np.ndarray:
    """Returns s[3, n, n, n] (z, y, x index order) in box units."""
    delta = box / n
    white = rng.standard_normal((n, n, n))
    wk = np.fft.rfftn(white)
    kz = 2.0 * np.pi * np.fft.fftfreq(n, d=delta)
    ky = 2.0 * np.pi * np.fft.fftfreq(n, d=delta)
    kx = 2.0 * np.pi * np.fft.rfftfreq(n, d=delta)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    sigma = sigma_spacings * delta
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k2 > 0, np.exp(-0.5 * k2 * sigma * sigma) / k2, 0.0)
    phik = wk * amp
    s = np.empty((3, n, n, n))
    for c, kc in enumerate((KX, KY, KZ)):
        s[c] = np.fft.irfftn(-1j * kc * phik, s=(n, n, n), axes=(0, 1, 2))
    rms = np.sqrt(np.mean(np.sum(s * s, axis=0)))
    if rms > 0:
        s *= rms_spacings * delta / rms
    return s


def make_lattice(n: int, rms_spacings: float = 1.0, seed: int = 1, box: float = BOX,
                 sigma_spacings: float = 2.0, infall_centres: int = 0) -> Lattice:
    """Seeded lattice of n^3 cubes; see the module docstring for the field."""
    if n < 1:
        raise ValueError("n must be >= 1")
    rng = np.random.default_rng(seed)
    delta = box / n
    if rms_spacings > 0 and n >= 2:
        s = _displacement_field(n, box, rms_spacings, sigma_spacings, rng)
    else:
        s = np.zeros((3, n, n, n))
    idx = np.arange(n + 1)
    wrap = idx % n
    # vertex (i,j,k) -> q + s[k%n, j%n, i%n]; arrays laid out [k, j, i] so ravel() is i-fastest
    sz = s[:, wrap][:, :, wrap][:, :, :, wrap]
    K, J, I = np.meshgrid(idx, idx, idx, indexing="ij")
    px = I * delta + sz[0]
    py = J * delta + sz[1]
    pz = K * delta + sz[2]
    if infall_centres > 0:
        px, py, pz = _apply_infall(px, py, pz, box, infall_centres, rng)
    return Lattice(n=n, box=box, x=px.astype(np.float32).ravel(), y=py.astype(np.float32).ravel(),
                   z=pz.astype(np.float32).ravel())


def make_snapshot(n: int, rms_spacings: float = 1.0, seed: int = 1, box: float = BOX,
                  sigma_spacings: float = 2.0) -> Lattice:
    """A periodic particle snapshot as a SORTED_BY_ID file holds it: n^3 particles, particle id
    i + n*(j + n*k) at its lattice site plus the displacement of make_lattice, wrapped into
    [0, box).  Returned as a Lattice of n^3 vertices (n - 1 cubes per axis), ids in order."""
    if n < 2:
        raise ValueError("n must be >= 2")
    full = make_lattice(n, rms_spacings, seed=seed, box=box, sigma_spacings=sigma_spacings)
    out = []
    for c in (full.x, full.y, full.z):
        v = c.reshape(n + 1, n + 1, n + 1)[:n, :n, :n].astype(np.float64)
        w = np.mod(v, box).astype(np.float32)
        w[w >= np.float32(box)] = 0.0
        out.append(np.ascontiguousarray(w.ravel()))
    return Lattice(n=n - 1, box=box, x=out[0], y=out[1], z=out[2])


def write_sorted_by_id(path: str, lat: Lattice) -> None:
    """Writes the SORTED_BY_ID format: three float32 coordinates per particle, particles in id
    order (AdPointDataReaders.h:770-911)."""
    np.stack([lat.x, lat.y, lat.z], axis=1).astype(np.float32).tofile(path)


def morton_ids(i, j, k) -> np.ndarray:
    """DARK_SKY particle ids: the Morton code of lattice site (i, j, k), x in bits 0, 3, 6, ...,
    y in bits 1, 4, ..., z in bits 2, 5, ... (AdPointDataReaders.h:366-414)."""
    i, j, k = (np.asarray(a, np.uint64) for a in (i, j, k))
    out = np.zeros(np.broadcast(i, j, k).shape, np.uint64)
    for b in range(16):
        bit = np.uint64(1 << b)
        out |= ((i & bit) << np.uint64(2 * b)) | ((j & bit) << np.uint64(2 * b + 1)) | ((k & bit) << np.uint64(2 * b + 2))
    return out.astype(np.int64)


def write_dark_sky(path: str, lat: Lattice, seed: int = 7, sha1_chunks: int = 2) -> None:
    """Writes a lattice as a DARK_SKY (SDF) file: a text header ending with "# SDF-EOH",
    `sha1_chunks` records of 24 bytes, then one record per particle {float x, y, z, vx, vy, vz;
    int64 id} in a seeded random order, id = the Morton code of the particle's site, velocities
    seeded noise (AdPointDataReaders.h:361-767)."""
    nx, ny, nz = lat.dims
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rec = np.zeros(lat.x.size, dtype=[("pos", np.float32, 3), ("vel", np.float32, 3), ("id", np.int64)])
    rec["pos"] = np.stack([lat.x, lat.y, lat.z], axis=1)
    rec["vel"] = rng.normal(0.0, 100.0, (lat.x.size, 3)).astype(np.float32)
    rec["id"] = morton_ids(i.ravel(), j.ravel(), k.ravel())
    rec = rec[rng.permutation(rec.size)]
    header = ("# SDF 1.0\n"
              "parameter byteorder = 0x78563412;\n"
              f"int sha1_chunks = {sha1_chunks};\n"
              "struct {\n  float x, y, z;\n  float vx, vy, vz;\n  int64_t ident;\n}[" + str(rec.size) + "];\n"
              "# SDF-EOH\n")
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(rng.integers(0, 256, 24 * sha1_chunks, dtype=np.uint8).tobytes())
        f.write(rec.tobytes())


def _apply_infall(px, py, pz, box, count, rng):
    """Spherical infall: x -> x - A (x - c) w(|x - c| / R) with w(r) = (1 - r^2)^2 for r < 1.

    Centres are uniform in the box, R ~ U[2, 8] Mpc/h and A ~ U[0.8, 1.6] (A > 1 inverts the
    tets near the centre).  Distances use the minimum image so periodic images move together.
    Each centre only touches the vertices within R of it, so the update runs over a window of
    lattice indices that provably holds them (unperturbed position within R + the largest current
    displacement, with periodic wrap); every other vertex has w = 0.
    """
    centres = rng.uniform(0.0, box, size=(count, 3))
    radii = rng.uniform(2.0, 8.0, size=count)
    amps = rng.uniform(0.8, 1.6, size=count)
    pos = [px.copy(), py.copy(), pz.copy()]
    nv = px.shape[0]  # arrays are [k, j, i] with nv = n + 1 vertices per axis
    n = nv - 1
    delta = box / n
    q = np.arange(nv) * delta
    # largest |position - unperturbed position| over all vertices and axes; only windowed vertices
    # move, so it is updated from them
    smax = max(float(np.max(np.abs(pos[0] - q[None, None, :]))), float(np.max(np.abs(pos[1] - q[None, :, None]))),
               float(np.max(np.abs(pos[2] - q[:, None, None]))))
    for c, r, a in zip(centres, radii, amps):
        reach = r + smax + delta
        win = []
        for a_ in range(3):
            lo = int(np.floor((c[a_] - reach) / delta)) - 1
            hi = int(np.ceil((c[a_] + reach) / delta)) + 1
            if hi - lo + 1 >= n:
                idx = np.arange(nv)
            else:
                idx = np.unique(np.arange(lo, hi + 1) % n)
                if 0 in idx:
                    idx = np.append(idx, n)  # vertex n is vertex 0 one box over
            win.append(idx)
        sel = np.ix_(win[2], win[1], win[0])
        sub = [pos[i][sel] for i in range(3)]
        d = [sub[i] - c[i] for i in range(3)]
        d = [di - box * np.round(di / box) for di in d]
        rr = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) / r
        w = np.where(rr < 1.0, (1.0 - rr * rr) ** 2, 0.0)
        qs = (q[win[0]][None, None, :], q[win[1]][None, :, None], q[win[2]][:, None, None])
        for i in range(3):
            moved = sub[i] - a * d[i] * w
            pos[i][sel] = moved
            smax = max(smax, float(np.max(np.abs(moved - qs[i]))))
    return pos[0], pos[1], pos[2]


def make_degenerate_tets(count: int, seed: int = 4, box: float = BOX, cell: float = 100.0 / 128,
                         ratio_range=(1e3, 1e6)) -> np.ndarray:
    """Seeded near-degenerate tets as a float32 soup [count, 4, 3].

    Each tet has a longest edge of about one `cell`, an edge-length ratio log-uniform in
    `ratio_range` (long/short), a random orientation and a random centre in the box, giving
    volumes down to ~1e-12 cell^3.
    """
    rng = np.random.default_rng(seed)
    ratio = np.exp(rng.uniform(np.log(ratio_range[0]), np.log(ratio_range[1]), size=count))
    long_len = cell * rng.uniform(0.3, 1.5, size=count)
    mid_len = long_len / np.sqrt(ratio) * rng.uniform(0.5, 2.0, size=count)
    short_len = long_len / ratio
    # random rotations from normalised quaternions
    q = rng.standard_normal((count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    rot = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)
    local = np.zeros((count, 4, 3))
    local[:, 1, 0] = long_len
    local[:, 2, 0] = long_len * rng.uniform(0.2, 0.8, size=count)
    local[:, 2, 1] = mid_len
    local[:, 3, 0] = long_len * rng.uniform(0.2, 0.8, size=count)
    local[:, 3, 1] = mid_len * rng.uniform(0.2, 0.8, size=count)
    local[:, 3, 2] = short_len
    local -= local.mean(axis=1, keepdims=True)
    world = np.einsum("nij,nvj->nvi", rot, local)
    centre = rng.uniform(0.0, box, size=(count, 1, 3))
    return (world + centre).astype(np.float32)


META = 7  # cells per meta-cell edge (AdTypeDefinitions.h:743-747)


def meta_cell_dims(np_xyz) -> tuple:
    """Meta-cells per axis: ceil((N_vertices - 1) / 7) (AdTypeDefinitions.h:730-735)."""
    return tuple(-(-(d - 1) // META) for d in np_xyz)


def meta_patches(lat: Lattice, ids=None) -> np.ndarray:
    """Meta-patch AoS buffers as the reference's MetaMesh::get_tets packs them.

    Each patch is 8^3 vertices x 3 floats, vertex (i',j',k') at float offset 3*(i'+8*(j'+8*k')),
    pre-filled with the 666.f sentinel for vertices beyond the lattice
    (AdTypeDefinitions.h:762-810).  `ids` selects meta-cells (ColumnMajorOrder over the meta-cell
    dims); default all, in id order.  Returns float32 [len(ids) * 1536].
    """
    npd = lat.dims
    md = meta_cell_dims(npd)
    total = md[0] * md[1] * md[2]
    if ids is None:
        ids = np.arange(total)
    ids = np.asarray(ids, dtype=np.int64)
    out = np.full((len(ids), 8, 8, 8, 3), 666.0, dtype=np.float32)  # [p, k', j', i', xyz]
    X = lat.x.reshape(npd[2], npd[1], npd[0])
    Y = lat.y.reshape(npd[2], npd[1], npd[0])
    Z = lat.z.reshape(npd[2], npd[1], npd[0])
    for p, m in enumerate(ids):
        cx = m % md[0]
        cy = (m // md[0]) % md[1]
        cz = m // (md[0] * md[1])
        lo = (META * cx, META * cy, META * cz)
        hi = tuple(min(lo[a] + META + 1, npd[a]) for a in range(3))
        sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
        e = (hi[2] - lo[2], hi[1] - lo[1], hi[0] - lo[0])
        out[p, :e[0], :e[1], :e[2], 0] = X[sl]
        out[p, :e[0], :e[1], :e[2], 1] = Y[sl]
        out[p, :e[0], :e[1], :e[2], 2] = Z[sl]
    return out.reshape(-1)
