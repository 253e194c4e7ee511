"""Region-of-interest request server (continuous batching) on the GPU.

Each request's FP32 density is compared with the oracle (the reference's deposit, evaluated with a
64-bit significand) of the same tets: those whose float AABB strictly overlaps the box, on a
dims^3 grid over the FP64 box, divided by the cell volume.  Per cell
|gpu - oracle| <= 1e-6 * max(|oracle|, 1e-8 * tet_mass / cell_volume) plus one FP32 rounding.
A request's result is the same alone and in a mix of concurrent requests (FP64 accumulation in a
different order: 1e-12 relative, far inside the 1e-6 bar).
"""

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice
from tests import _oracle as O

pytestmark = pytest.mark.gpu

TET_MASS = 1.0 / 6.0
DECOMP = A.DECOMP_KUHN


@pytest.fixture(autouse=True)
def exact_scaling():
    with O.cell_scaling(O.EXACT):
        yield


@pytest.fixture(scope="module")
def lat():
    return make_lattice(20, 2.0, seed=5)


def oracle_density(lat, roi, dims):
    """Oracle mass of the strictly overlapping tets on the ROI grid, as density."""
    roi = np.asarray(roi, np.float64)
    rf = roi.astype(np.float32)
    tets = O.lattice_tets(lat, DECOMP)
    mn, mx = tets.min(1), tets.max(1)
    hit = np.all((mn < rf[1::2]) & (rf[0::2] < mx), 1)
    ids = np.nonzero(hit)[0].astype(np.int32)
    mass = O.deposit_lattice(lat, roi, dims, decomp=DECOMP, tet_mass=TET_MASS, tet_ids=ids, precision=3)
    cell_vol = np.prod((roi[1::2] - roi[0::2]) / dims)
    return (mass / cell_vol).reshape(dims, dims, dims), ids.size, cell_vol


def rel_err(got, want, floor):
    got = np.asarray(got, np.float64)
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), floor)))


ROIS = [([10.0, 40.0, 20.0, 45.0, 30.0, 60.0], 24),    # inside
        ([-8.0, 24.0, -8.0, 24.0, -8.0, 24.0], 16),     # across the lattice's lower corner
        ([45.5, 47.25, 50.0, 51.75, 50.0, 51.75], 32)]  # much smaller than a tet: cells << tets


def test_requests_match_oracle(lat):
    with A.ResidentLattice.from_lattice(lat) as rl, A.Server(rl, DECOMP, TET_MASS, max_inflight=4) as srv:
        ids = [srv.submit(roi, dims) for roi, dims in ROIS]
        infos = []
        for (roi, dims), rid in zip(ROIS, ids):
            print(f"[serve] roi={roi} dims={dims}", flush=True)
            got, info = srv.wait(rid, timeout=60)
            want, n, cell_vol = oracle_density(lat, roi, dims)
            assert info["n_tets"] == n
            assert rel_err(got, want, 1e-8 * TET_MASS / cell_vol) <= 1e-6 + 1.2e-7
            assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
            infos.append(info)
        st = srv.stats()
        assert st["requests_done"] == len(ROIS) and st["pending"] == 0 and st["active"] == 0
        # once the server is idle every step's device time is counted, and the step kernels ran
        # within the requests' wall time
        wall = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
        assert st["steps"] > 0 and 0.0 < st["step_kernel_s"] <= wall


def test_plain_baseline_matches_oracle(lat):
    roi, dims = ROIS[0]
    with A.ResidentLattice.from_lattice(lat) as rl:
        got, n = A.roi_deposit_plain(rl, roi, dims, DECOMP, TET_MASS)
    want, n_want, cell_vol = oracle_density(lat, roi, dims)
    assert n == n_want
    assert rel_err(got, want, 1e-8 * TET_MASS / cell_vol) <= 1e-6 + 1.2e-7


def test_same_alone_and_in_a_mix(lat):
    """Continuous batching: the probe request shares step launches with other requests that arrive
    before and after it, and its density matches the one computed alone."""
    probe = ([12.0, 52.0, 8.0, 48.0, 20.0, 60.0], 40)
    others = [([0.0, 30.0, 0.0, 30.0, 0.0, 30.0], 32), ([30.0, 90.0, 10.0, 70.0, 40.0, 100.0], 48),
              ([50.0, 60.0, 50.0, 60.0, 50.0, 60.0], 16), ([5.0, 95.0, 5.0, 95.0, 5.0, 95.0], 64)]
    with A.ResidentLattice.from_lattice(lat) as rl:
        with A.Server(rl, DECOMP, TET_MASS, max_inflight=8, step_tiles=64) as srv:
            alone, info_alone = srv.wait(srv.submit(*probe), timeout=120)
        with A.Server(rl, DECOMP, TET_MASS, max_inflight=8, step_tiles=64) as srv:
            before = [srv.submit(*r) for r in others[:2]]
            pid = srv.submit(*probe)
            after = [srv.submit(*r) for r in others[2:]]
            mixed, info = srv.wait(pid, timeout=120)
            rest = [srv.wait(r, timeout=120)[1] for r in before + after]
            steps = srv.stats()["steps"]
    # the probe's tiles took several steps, and other requests were in flight meanwhile
    assert info["n_candidates"] > 64 * 64 and steps > len(others) + 1
    overlap = [r for r in rest if r["t_admit"] < info["t_done"] and info["t_admit"] < r["t_done"]]
    assert overlap, "no request overlapped the probe on the device"
    want, _, cell_vol = oracle_density(lat, *probe)
    floor = 1e-8 * TET_MASS / cell_vol
    assert rel_err(mixed, alone.astype(np.float64), floor) <= 1e-6
    assert np.max(np.abs(mixed.astype(np.float64) - alone) / np.maximum(np.abs(alone), floor)) <= 1.2e-7
    assert rel_err(mixed, want, floor) <= 1e-6 + 1.2e-7


def test_memory_budget_admission(lat):
    """A budget that holds one request's buffers but not two (each 48^3 request needs about 6 MiB:
    six buffers of at least 1 MiB) still serves every request: a request that does not fit waits
    until the running one has been copied out and its slot freed, and is then admitted; a request
    larger than the whole budget fails alone."""
    reqs = [([0.0, 50.0, 0.0, 50.0, 0.0, 50.0], 48), ([40.0, 90.0, 40.0, 90.0, 40.0, 90.0], 48),
            ([20.0, 70.0, 20.0, 70.0, 20.0, 70.0], 48)]
    budget = 10 << 20
    with A.ResidentLattice.from_lattice(lat) as rl:
        with A.Server(rl, DECOMP, TET_MASS, max_inflight=4, memory_budget=budget) as srv:
            ids = [srv.submit(*r) for r in reqs]
            res = [srv.wait(i, timeout=120) for i in ids]
            assert srv.stats()["device_bytes"] <= budget
        infos = [info for _, info in res]
        # one request at a time: each later request was admitted only after the one before it ended
        for prev, nxt in zip(infos, infos[1:]):
            assert nxt["t_admit"] >= prev["t_done"], (prev, nxt)
        for (roi, dims), (got, info) in zip(reqs, res):
            want, _, cell_vol = oracle_density(lat, roi, dims)
            assert rel_err(got, want, 1e-8 * TET_MASS / cell_vol) <= 1e-6 + 1.2e-7
        with A.Server(rl, DECOMP, TET_MASS, max_inflight=2, memory_budget=8 << 20) as srv:
            rid = srv.submit([0.0, 100.0, 0.0, 100.0, 0.0, 100.0], 256)
            with pytest.raises(A.AdRuntimeException) as ei:
                srv.wait(rid, timeout=60)
            assert ei.value.status == A.ADECO_EOVERFLOW


def test_step_tiles_limit(lat):
    with A.ResidentLattice.from_lattice(lat) as rl:
        with pytest.raises(A.AdRuntimeException) as ei:
            A.Server(rl, DECOMP, TET_MASS, step_tiles=(1 << 22) + 1)
        assert ei.value.status == A.ADECO_EINVAL
        with A.Server(rl, DECOMP, TET_MASS, step_tiles=1 << 22) as srv:
            got, info = srv.wait(srv.submit(*ROIS[0]), timeout=60)
        want, _, cell_vol = oracle_density(lat, *ROIS[0])
        assert rel_err(got, want, 1e-8 * TET_MASS / cell_vol) <= 1e-6 + 1.2e-7


def test_box_outside_the_lattice_is_empty(lat):
    with A.ResidentLattice.from_lattice(lat) as rl, A.Server(rl, DECOMP, TET_MASS) as srv:
        got, info = srv.wait(srv.submit([500.0, 510.0, 500.0, 510.0, 500.0, 510.0], 8), timeout=30)
    assert info["n_tets"] == 0 and not got.any()
