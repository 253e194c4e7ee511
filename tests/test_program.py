"""The single-rank program around the deposit, on the CPU, against the reference's own outputs
(tests/golden/ref_program.json, written by tests/golden/make_program_golden.py from the compiled
reference): parse_parameter_file (main.cpp:83-183), the SORTED_BY_ID reader
(AdPointDataReaders.h:770-911), the periodic correction (AdPointDataReaders.h:156-206) and whole
runs of the reference program, whose leaf files the oracle chain rebuilds bit for bit.

The parser and reader are this library's host code (no GPU); the periodic correction and the
program's leaf files are checked here through the oracle, and on the GPU in
tests/test_gpu_program.py.
"""

import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_snapshot, morton_ids, write_sorted_by_id
from adeco_amd.workflow import AMD_Parameters, get_reader, parse_parameter_file
from tests import _oracle as O
from tests._program_cases import (DARK_SKY_HEADERS, PARAM_TEXTS, PROGRAM_CASES, case_lattice,
                                  param_file_text, unwrap_lattice_case)
from tests._program_oracle import oracle_program, unwrapped

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_program.json")))
EXE = os.path.join(ROOT, "adeco_amd", "adeco_resample")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_message(err: str) -> str:
    return err.split("AdRuntimeException: ", 1)[-1]


@pytest.mark.parametrize("name", sorted(PARAM_TEXTS))
def test_parse_parameter_file_matches_reference(tmp_path, name):
    path = tmp_path / "params.txt"
    path.write_text(PARAM_TEXTS[name])
    want = GOLDEN["params"][name]
    if "error" in want:
        with pytest.raises(A.AdRuntimeException) as e:
            parse_parameter_file(str(path))
        assert ref_message(want["error"]).replace("<file>", str(path)) in str(e.value)
        return
    got = parse_parameter_file(str(path))
    for k, v in want.items():
        g = getattr(got, k)
        if k == "region_of_interest":
            assert list(g) == v, k
        elif isinstance(g, bool):
            assert int(g) == v, k
        else:
            assert g == v, k


def test_parse_parameter_file_missing(tmp_path):
    with pytest.raises(A.AdRuntimeException) as e:
        parse_parameter_file(str(tmp_path / "no_such_file.txt"))
    assert ref_message(GOLDEN["params"]["missing_file"]["error"]).replace("<dir>", str(tmp_path)) in str(e.value)


def test_parse_parameter_file_line_limit(tmp_path):
    """At most 10000 lines (main.cpp:102,174-178)."""
    path = tmp_path / "long.txt"
    path.write_text("stride = 2\n" * 10000)
    with pytest.raises(A.AdRuntimeException, match="Failed to parse parameter file"):
        parse_parameter_file(str(path))
    path.write_text("stride = 2\n" * 9998)
    assert parse_parameter_file(str(path)).stride == 2


def test_parse_parameter_file_live_reference(tmp_path):
    """Against the compiled reference itself, where it is built: every text, plus one with every
    key set and one with CRLF line ends."""
    r = O.load_ref()
    if r is None:
        pytest.skip("oracle/_ref not built")
    import ctypes
    texts = dict(PARAM_TEXTS)
    texts["crlf"] = "stride = 3\r\nperiodic_boundaries = true\r\nuse_region_of_interest = true\r\n"
    for name, text in texts.items():
        path = tmp_path / f"{name}.txt"
        path.write_text(text)
        cap = 4096
        strs = ctypes.create_string_buffer(4 * cap)
        ints = np.zeros(8, np.int64)
        roi = np.zeros(6, np.float32)
        rc = r.ref_parse_parameter_file(str(path).encode(), strs, cap, O.P(ints), O.P(roi))
        if rc != 0:
            with pytest.raises(A.AdRuntimeException):
                parse_parameter_file(str(path))
            continue
        got = parse_parameter_file(str(path))
        s = [strs.raw[i * cap:(i + 1) * cap].split(b"\0", 1)[0].decode() for i in range(4)]
        assert [got.input_file, got.output_directory, got.id_type, got.restart_path] == s, name
        assert [int(got.use_regions_of_interest), got.max_refinement_level, got.stride,
                got.max_tets_per_octree_node & (2 ** 63 - 1), got.linear_patch_resolution, got.num_pthreads,
                got.num_mpi_ranks_per_node, int(got.periodic_boundaries)] == \
            [int(v) & (2 ** 63 - 1) for v in ints], name
        assert list(got.region_of_interest) == [float(v) for v in roi], name


def test_sorted_by_id_reader_matches_reference(tmp_path):
    g = GOLDEN["reader"]
    lat = make_snapshot(*g["good"]["snapshot"][:2], seed=g["good"]["snapshot"][2])
    path = tmp_path / "good.bin"
    write_sorted_by_id(str(path), lat)
    x, y, z, dims = get_reader(" SORTED_BY_ID").load_data(str(path))
    assert list(dims) == g["good"]["dims"] and x.size == g["good"]["n"]
    assert sha(np.concatenate([x, y, z])) == g["good"]["xyz_sha"]
    (tmp_path / "odd.bin").write_bytes(bytes(13))
    np.zeros(30, np.float32).tofile(tmp_path / "noncubic.bin")
    for fname, key in (("odd.bin", "odd_size"), ("noncubic.bin", "not_cubic")):
        with pytest.raises(A.AdRuntimeException) as e:
            get_reader("SORTED_BY_ID").load_data(str(tmp_path / fname))
        assert ref_message(g[key]["error"]) in str(e.value)
    # an empty file is an empty lattice for the reader (the program rejects it later)
    (tmp_path / "empty.bin").write_bytes(b"")
    x, y, z, dims = get_reader("SORTED_BY_ID").load_data(str(tmp_path / "empty.bin"))
    assert list(dims) == g["empty"]["dims"] and x.size == g["empty"]["n"] == 0
    assert sha(np.concatenate([x, y, z])) == g["empty"]["xyz_sha"]
    with pytest.raises(A.AdRuntimeException) as e:
        get_reader("SORTED_BY_ID").load_data(str(tmp_path / "no_such_file.bin"))
    assert ref_message(g["missing"]["error"]).replace("<dir>", str(tmp_path)) in str(e.value)
    with pytest.raises(A.AdRuntimeException, match="invalid 'stride'"):
        get_reader("SORTED_BY_ID").load_data(str(path), stride=0)
    with pytest.raises(A.AdRuntimeException, match="not supported"):
        get_reader("HDF5")
    assert type(get_reader(" DARK_SKY")).__name__ == "AdDarkSkyDataReader"


@pytest.mark.parametrize("name", sorted(DARK_SKY_HEADERS))
def test_dark_sky_header_matches_reference(tmp_path, name):
    text, chunks = DARK_SKY_HEADERS[name]
    path = tmp_path / f"hdr_{name}.bin"
    path.write_bytes(text.encode() + bytes(24 * chunks) + bytes(32 * 8))
    want = GOLDEN["dark_sky"]["headers"][name]
    if "error" in want:
        with pytest.raises(A.AdRuntimeException) as e:
            get_reader("DARK_SKY").info(str(path))
        assert ref_message(want["error"]).replace("<dir>", str(tmp_path)) in str(e.value)
    else:
        info = get_reader("DARK_SKY").info(str(path))
        assert info == {"header_bytes": want["header_bytes"], "num_particles": 8, "dims": 2}


def test_dark_sky_id_decoding():
    """The reader kernel's id decoding (evaluated on the host) inverts the Morton codes the
    reference's reader decodes (AdPointDataReaders.h:366-414): bit 3b of the id is bit b of x,
    3b+1 of y, 3b+2 of z; bits above 47 are ignored."""
    from adeco_amd._lib import load as lib_load
    rng = np.random.default_rng(0)
    ijk = rng.integers(0, 1 << 16, (2000, 3))
    ids = morton_ids(ijk[:, 0], ijk[:, 1], ijk[:, 2])
    site = np.zeros(3, np.int32)
    for q in range(ids.size):
        lib_load().adeco_selftest_dark_sky_site(int(ids[q]) | (0x7FFF << 48), O.P(site))
        assert site.tolist() == ijk[q].tolist()
    for b in range(48):  # one bit at a time
        lib_load().adeco_selftest_dark_sky_site(1 << b, O.P(site))
        assert site.tolist() == [(1 << (b // 3)) if a == b % 3 else 0 for a in range(3)]


def test_oracle_periodic_correction_matches_reference():
    g = GOLDEN["unwrap"]
    dims, x, y, z = unwrap_lattice_case()
    bbox = np.array([x.min(), x.max(), y.min(), y.max(), z.min(), z.max()], np.float32)
    assert bbox.tolist() == g["bbox"]
    ux, uy, uz = unwrapped(x, y, z, dims, bbox)
    assert sha(np.concatenate([ux, uy, uz])) == g["xyz_sha"]
    # the correction moves only vertices near the far face of their half, by exactly the extent
    moved = (ux != x) | (uy != y) | (uz != z)
    assert 0 < moved.sum() < x.size
    ext = bbox[1::2] - bbox[0::2]
    for a, (u, v) in enumerate(((ux, x), (uy, y), (uz, z))):
        d = u[u != v] - v[u != v]
        assert np.all(np.isin(np.abs(d), [ext[a]]) | (np.abs(np.abs(d) - ext[a]) <= 1e-5 * ext[a]))


def test_oracle_periodic_correction_closed_form():
    """A lattice shifted by a quarter spacing and wrapped into [0, L): the correction with period L
    puts every particle back on its unwrapped site (the first layer, wrapped to the far side,
    moves back down by L)."""
    n, L = 8, 64.0
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    q = [(a.ravel() * (L / n) - 0.25 * L / n) for a in (i, j, k)]
    wrapped = [np.mod(c, L).astype(np.float32) for c in q]
    bbox = np.array([wrapped[0].min(), wrapped[0].max(), wrapped[1].min(), wrapped[1].max(), wrapped[2].min(),
                     wrapped[2].max()], np.float32)
    ux, uy, uz = unwrapped(*wrapped, (n, n, n), bbox, period=(L, L, L))
    for u, c, w in zip((ux, uy, uz), q, wrapped):
        # index 0 sits at -spacing/4, wrapped to L - spacing/4: index 0 < n/2 and above the third
        # quarter mark, so it moves down by L; every other particle stays
        assert np.array_equal(u, c.astype(np.float32))
        assert np.count_nonzero(u != w) == n * n


@pytest.mark.parametrize("name", sorted(PROGRAM_CASES))
def test_oracle_chain_reproduces_reference_program(tmp_path, name):
    g = GOLDEN["program"][name]
    case = PROGRAM_CASES[name]
    lat = case_lattice(case)  # the deposited lattice (a DARK_SKY stride keeps every s-th particle)
    assert sha(np.concatenate([lat.x, lat.y, lat.z])) == g["snapshot_sha"]
    assert lat.bbox().tolist() == g["global_bbox"]
    files, total, ulat, _ = oracle_program(name)
    if g["unwrap_sha"] is not None:
        assert sha(np.concatenate([ulat.x, ulat.y, ulat.z])) == g["unwrap_sha"]
    assert sorted(files) == sorted(g["files"])
    for fname, dens in files.items():
        assert sha(dens) == g["files"][fname], fname
    assert abs(total - g["total_mass_printed"]) <= 1e-6 * g["total_mass_printed"]


def test_program_reports_errors_without_exit(tmp_path):
    """adeco_resample: usage, parameter and reader errors give exit status 1 and the reference's
    message; without a GPU the deposit reports the missing device the same way."""
    assert os.path.exists(EXE), "adeco_amd/adeco_resample is not built"
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 1 and "invalid number of arguments" in r.stderr
    r = subprocess.run([EXE, str(tmp_path / "none.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "failed to open parameter file" in r.stderr
    case = PROGRAM_CASES["periodic_whole_domain"]
    pfile = tmp_path / "params.txt"
    pfile.write_text(param_file_text(case, str(tmp_path / "missing.bin"), str(tmp_path)))
    r = subprocess.run([EXE, str(pfile)], capture_output=True, text=True)
    assert r.returncode == 1 and "failed to open file" in r.stderr
    assert "PARAMETERS INFO" in r.stdout
    if A.device_count() == 0:
        lat = make_snapshot(6, 1.0, seed=1)
        write_sorted_by_id(str(tmp_path / "snap.bin"), lat)
        pfile.write_text(param_file_text(case, str(tmp_path / "snap.bin"), str(tmp_path)))
        r = subprocess.run([EXE, str(pfile)], capture_output=True, text=True)
        assert r.returncode == 1 and "ERROR" in r.stderr
        assert not any(f.startswith("node_") for f in os.listdir(tmp_path))


def test_parameters_round_trip():
    p = AMD_Parameters(input_file="/a b/c.bin", output_directory="/o", id_type="SORTED_BY_ID",
                       use_regions_of_interest=True, region_of_interest=(1, 2, 3, 4, 5, 6), max_tets_per_octree_node=7)
    assert AMD_Parameters._from_c(p._to_c()) == AMD_Parameters(**{**p.__dict__, "region_of_interest": (1.0, 2.0, 3.0,
                                                                                                        4.0, 5.0, 6.0)})
