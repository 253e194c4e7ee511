"""The C-ABI library on the CPU: it loads, exports every function include/adeco.h declares,
validates arguments, fails with ADECO_EDEVICE (never a CPU fallback) when no GPU is present, and
its shared host/device geometry (the code the kernels compile) agrees with the oracle."""

import ctypes
import os
import re

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice, meta_patches
from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "adeco.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(adeco_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # and the Python mirror binds every one of them
    from adeco_amd._lib import SIGNATURES
    assert set(names) <= set(SIGNATURES), set(names) - set(SIGNATURES)


def test_no_gpu_means_edevice_not_fallback():
    if A.device_count() > 0:
        pytest.skip("a GPU is present")
    lat = make_lattice(4, 0.5)
    pat = meta_patches(lat)
    data = np.zeros(8 ** 3, np.float32)
    with pytest.raises(A.AdRuntimeException) as ei:
        A.DepositWrapper.deposit_tets(True, 0, 0, 1, pat, pat.size, 7, A.cubic_root_box(lat.bbox()), 8, data)
    assert ei.value.status == A.ADECO_EDEVICE
    assert not data.any()
    with pytest.raises(A.AdRuntimeException) as ei:
        A.ResidentLattice.from_lattice(lat)
    assert ei.value.status == A.ADECO_EDEVICE
    with pytest.raises(A.AdRuntimeException) as ei:
        A.deposit_soup(np.zeros((1, 4, 3), np.float32), [0, 1, 0, 1, 0, 1], 4)
    assert ei.value.status == A.ADECO_EDEVICE


def test_server_argument_validation():
    lib = A.load()
    h = A._lib.c_vp()
    cfg = A._lib.ServeConfig(A.DECOMP_KUHN, A.PREC_FP64, 1.0 / 6, 4, 0, 0, 0)
    assert lib.adeco_serve_create(None, A._lib.ctypes.byref(cfg), A._lib.ctypes.byref(h)) == A.ADECO_EINVAL
    assert lib.adeco_serve_submit(None, None, 8, None, None) == A.ADECO_EINVAL
    assert lib.adeco_serve_wait(None, 1, 0.0, None) == A.ADECO_EINVAL
    assert lib.adeco_serve_stats_get(None, None) == A.ADECO_EINVAL
    assert lib.adeco_serve_destroy(None) == A.ADECO_OK
    roi = np.array([0, 1, 0, 1, 0, 1], np.float64)
    out = np.zeros(8, np.float32)
    assert lib.adeco_roi_deposit_plain(None, 2, 1.0, A.ptr(roi), 2, A.ptr(out), None) == A.ADECO_EINVAL


def test_boundary_argument_validation_precedes_device_use():
    lat = make_lattice(4, 0.5)
    pat = meta_patches(lat)
    box = A.cubic_root_box(lat.bbox())
    data = np.zeros(8 ** 3, np.float32)
    lib = A.load()
    # positions_size == 0 is a no-op (AdDeposit2Grid.cpp:86-87)
    assert lib.adeco_deposit_patches(1, 0, 0, 1, A.ptr(pat), 0, 7, A.ptr(box), 8, A.ptr(data)) == A.ADECO_OK
    # not a whole number of patches (AdDeposit2Grid.cpp:100)
    assert lib.adeco_deposit_patches(1, 0, 0, 1, A.ptr(pat), 1535, 7, A.ptr(box), 8, A.ptr(data)) == A.ADECO_EINVAL
    # patch range beyond the array (AdDeposit2Grid.cu:265-268)
    assert lib.adeco_deposit_patches(1, 0, 1, 1, A.ptr(pat), 1536, 7, A.ptr(box), 8, A.ptr(data)) == A.ADECO_EINVAL
    bad = np.array([1, 0, 0, 1, 0, 1], np.float32)
    assert lib.adeco_deposit_patches(1, 0, 0, 1, A.ptr(pat), 1536, 7, A.ptr(bad), 8, A.ptr(data)) == A.ADECO_EINVAL
    assert lib.adeco_deposit_patches(1, 0, 0, 1, A.ptr(pat), 1536, 7, A.ptr(box), 0, A.ptr(data)) == A.ADECO_EINVAL
    assert b"patch" in lib.adeco_last_error() or b"dims" in lib.adeco_last_error()
    assert not data.any()


def test_decomposition_tables_match_oracle():
    lib = A.load()
    o = O.load()
    for d in (0, 1, 2):
        a = np.zeros(24, np.int32)
        b = np.zeros(24, np.int32)
        lib.adeco_selftest_decomposition(d, A.ptr(a))
        o.oracle_decomposition(d, O.P(b))
        assert np.array_equal(a, b)


def _vol6(t):
    return O.load().oracle_vol6_f64(O.P(np.ascontiguousarray(t, np.float64)))


def test_kernel_clip_matches_oracle_clip():
    lib = A.load()
    o = O.load()
    rng = np.random.default_rng(5)
    for _ in range(3000):
        scale = 10 ** rng.uniform(-2, 1.2)
        t = rng.uniform(-0.5, 1.5, 3) + scale * rng.standard_normal((4, 3))
        if _vol6(t) < 0:
            t = t[[1, 0, 2, 3]]
        t = np.ascontiguousarray(t)
        want = 6 * o.oracle_clip_unit_cube_f64(O.P(t), None)
        got = lib.adeco_selftest_clip_vol6_f64(A.ptr(t))
        assert abs(got - want) <= 1e-12 * max(1.0, abs(_vol6(t)))
    # orientation-independent (the kernels' clip takes any vertex order)
    t = np.ascontiguousarray([[0.2, 0.1, -0.3], [1.4, 0.3, 0.2], [0.1, 1.2, 0.4], [0.3, 0.2, 1.5]])
    assert lib.adeco_selftest_clip_vol6_f64(A.ptr(t)) == pytest.approx(
        lib.adeco_selftest_clip_vol6_f64(A.ptr(np.ascontiguousarray(t[[1, 0, 2, 3]]))), rel=1e-14)


def test_column_segment_clip_matches_oracle_per_cell():
    """The batched kernel's column-segment clip (one x/y clip for up to 8 cells along z, planes
    between cells shared) against the oracle's clip of the tet shifted onto each cell.  The exact
    tree must match every cell; the signed tree every cell it does not flag as possible
    cancellation noise (those the kernel recomputes with the exact tree)."""
    lib = A.load()
    o = O.load()
    rng = np.random.default_rng(21)
    flagged = checked = 0
    for it in range(1500):
        ls = int(rng.integers(1, 9))
        scale = 10 ** rng.uniform(-2, 1.0)
        centre = np.array([rng.uniform(-0.5, 1.5), rng.uniform(-0.5, 1.5), rng.uniform(-1.0, ls + 1.0)])
        t = centre + scale * rng.standard_normal((4, 3)) * np.array([1.0, 1.0, 2.0])
        if _vol6(t) < 0:
            t = t[[1, 0, 2, 3]]
        t = np.ascontiguousarray(t)
        tol = 1e-12 * max(1.0, abs(_vol6(t)))
        want = np.array([6 * o.oracle_clip_unit_cube_ld(O.P(np.ascontiguousarray(t - [0.0, 0.0, c])))
                         for c in range(ls)])
        for signed in (0, 1):
            got = np.zeros(ls)
            sab = np.zeros(ls)
            lib.adeco_selftest_column_f64(A.ptr(t), ls, signed, A.ptr(got), A.ptr(sab))
            for c in range(ls):
                if signed and not got[c] >= 1e-6 * sab[c]:
                    flagged += 1
                    continue
                checked += 1
                assert abs(got[c] - want[c]) <= tol, (it, ls, c, signed, got[c], want[c])
            if not signed:
                # the exact tree's terms are all positive: the magnitudes are the values
                assert np.allclose(sab, got, rtol=1e-12, atol=1e-300)
    assert checked > 10 * flagged


def test_beyond_fraction_closed_forms():
    """Fraction of a tet beyond a plane from vertex distances, against exact clipping of a large
    box (the z-slab closed form used by the kernels' clip)."""
    lib = A.load()
    o = O.load()
    rng = np.random.default_rng(9)
    for _ in range(400):
        t = rng.standard_normal((4, 3))
        if _vol6(t) < 0:
            t = t[[1, 0, 2, 3]]
        d = np.ascontiguousarray(t[:, 2].copy())  # beyond z = 0
        got = lib.adeco_selftest_beyond_fraction_f64(A.ptr(d))
        # |t ∩ {z > 0}| / |t| via the oracle's clip of the tet scaled into a unit cube shifted so
        # that z = 0 is the cube's lower face and the tet lies inside in x, y and below the top
        lo = t.min(0) - 1.0
        ext = (t.max(0) - lo).max() + 1.0
        s = (t - np.array([lo[0], lo[1], 0.0])) / ext
        w = o.oracle_clip_unit_cube_f64(O.P(np.ascontiguousarray(s)), None) * ext ** 3
        assert got == pytest.approx(w / (abs(_vol6(t)) / 6), rel=1e-10, abs=1e-12)
    # vertices exactly on the plane count as near; every case of 0-4 beyond
    for d, want in (([1, 0, 0, 0], 1.0), ([1, 1, 1, 0], 1.0), ([1, 1, 0, 0], 1.0), ([-1, 0, 0, 0], 0.0),
                    ([0, 0, 0, 0], 0.0), ([2, 2, 2, 2], 1.0), ([-1, -1, -1, -1], 0.0),
                    ([1, -1, -1, -1], 1 / 8), ([1, 1, 1, -1], 7 / 8), ([1, 1, -1, -1], 0.5)):
        got = lib.adeco_selftest_beyond_fraction_f64(A.ptr(np.ascontiguousarray(d, np.float64)))
        assert got == pytest.approx(want, rel=1e-15, abs=1e-16), d


def test_kernel_classification_agrees_with_clip():
    lib = A.load()
    rng = np.random.default_rng(11)
    seen = set()
    for _ in range(3000):
        scale = 10 ** rng.uniform(-1.5, 1.0)
        t = rng.uniform(-1, 2, 3) + scale * rng.standard_normal((4, 3))
        if _vol6(t) < 0:
            t = t[[1, 0, 2, 3]]
        t = np.ascontiguousarray(t)
        cls = lib.adeco_selftest_classify_f64(A.ptr(t))
        w = lib.adeco_selftest_clip_vol6_f64(A.ptr(t)) / 6
        seen.add(cls)
        if cls == 0:
            assert w == pytest.approx(0.0, abs=1e-12)
        elif cls == 1:
            assert w == pytest.approx(abs(_vol6(t)) / 6, rel=1e-12)
        elif cls == 2:
            assert w == pytest.approx(1.0, rel=1e-12)
    assert seen == {0, 1, 2, 3}


@pytest.mark.parametrize("rms,dims", [(1.0, 8), (3.0, 20)])
def test_kernel_per_tet_deposit_matches_oracle_on_host(rms, dims):
    """The kernels' per-tet FP64 deposit routine, run on the host, against the oracle."""
    lat = make_lattice(6, rms, seed=8)
    tets = O.lattice_tets(lat, 2)
    box = np.array([-64, 192, -64, 192, -64, 192], np.float32)  # holds every tet; equal float cell sizes
    got = np.zeros(dims ** 3)
    assert A.load().adeco_selftest_deposit_soup_f64(A.ptr(tets), tets.shape[0], A.ptr(box), dims, 1.0, A.ptr(got)) == 0
    want = O.deposit_soup_ld(tets, box, dims)
    # against the extended-precision oracle: cells above 1e-6 of a tet mass agree to 1e-9, and the
    # standard 1e-6 bar holds with the 1e-8 floor
    assert O.max_rel_err(got, want, floor_frac=1e-6) <= 1e-9
    assert O.max_rel_err(got, want) <= 1e-6
    assert O.max_rel_err(got, O.deposit_soup_f64(tets, box, dims), floor_frac=1e-6) <= 1e-9
    assert got.sum() == pytest.approx(6 * lat.n ** 3, rel=1e-12)
    # the soup in lattice order equals the oracle's lattice deposit
    assert O.max_rel_err(got, O.deposit_lattice(lat, box, dims, decomp=2), floor_frac=1e-6) <= 1e-9


def test_boundary_form_settings():
    """The boundary call's form settings validate their arguments and keep their values (no GPU)."""
    lib = A.load()
    assert (lib.adeco_get_tet_decomposition(), lib.adeco_get_tet_mass(), lib.adeco_get_deposit_mode(),
            lib.adeco_get_precision_mode()) == (A.DECOMP_REF_EFFECTIVE, 1.0, A.MODE_PLAIN, A.PREC_FP64)
    assert lib.adeco_set_tet_decomposition(3) == A.ADECO_EINVAL
    assert lib.adeco_set_tet_mass(0.0) == A.ADECO_EINVAL
    assert lib.adeco_set_tet_mass(float("inf")) == A.ADECO_EINVAL
    assert lib.adeco_set_deposit_mode(2) == A.ADECO_EINVAL
    assert lib.adeco_set_precision_mode(3) == A.ADECO_EINVAL
    prev = A.set_boundary_form(decomposition=A.DECOMP_KUHN, tet_mass=0.25, mode=A.MODE_BATCHED, precision=A.PREC_MIXED)
    try:
        assert A.set_boundary_form() == {"decomposition": A.DECOMP_KUHN, "tet_mass": 0.25, "mode": A.MODE_BATCHED,
                                         "precision": A.PREC_MIXED}
    finally:
        A.set_boundary_form(**prev)
    assert A.set_boundary_form() == prev
