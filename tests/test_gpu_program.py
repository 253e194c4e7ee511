"""The program around the deposit on the GPU: the periodic correction on the resident lattice,
the octree after it, and whole runs of adeco_resample (the reference program's counterpart)
against the reference program's own leaf files (tests/golden/ref_program.json).

Bars: the correction and the octree are bit-exact; every leaf file matches the FP64 oracle's
densities in the reference's geometry to 1e-6 per cell (floor 1e-8 of one tet's mass per cell
volume).  Against the reference program's files - which the float-faithful oracle rebuilds bit
for bit - each cell differs by no more than the reference's own float error (its distance from
the FP64 oracle, up to 3.0e-3 relative with a floor of 1e-3 of a tet's mass per cell volume on
the non-periodic case, whose wrapped particles make tets that span the box, and 2.0e-4 on the
periodic one) plus 1e-6; totals agree to 2e-6.
"""

import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import Lattice, make_snapshot, morton_ids, write_dark_sky
from adeco_amd.workflow import deposit_workflow, get_reader, parse_parameter_file
from tests import _oracle as O
from tests._golden_leaves import patch_file_name
from tests._octree_io import oracle_octree
from tests._program_cases import PROGRAM_CASES, param_file_text, unwrap_lattice_case, write_case_snapshot
from tests._program_oracle import oracle_program, unwrapped

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_program.json")))
EXE = os.path.join(ROOT, "adeco_amd", "adeco_resample")
REL = 1e-6


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_gpu_periodic_correction_matches_reference():
    g = GOLDEN["unwrap"]
    dims, x, y, z = unwrap_lattice_case()
    print("[program] periodic correction (13,9,7)", flush=True)
    with A.ResidentLattice(x, y, z, dims) as rl:
        bbox = rl.bbox()
        assert bbox.tolist() == g["bbox"]
        rl.periodic_unwrap(bbox)
        assert sha(np.concatenate(rl.positions())) == g["xyz_sha"]
    with A.ResidentLattice(x, y, z, dims) as rl:
        period = (50.0, 50.0, 50.0)
        rl.periodic_unwrap(bbox, period=period)
        got = rl.positions()
    want = unwrapped(x, y, z, dims, bbox, period=period)
    for a in range(3):
        assert np.array_equal(got[a], want[a])


def test_gpu_octree_after_correction_is_rooted_at_the_global_bbox():
    case = PROGRAM_CASES["periodic_whole_domain"]
    lat = make_snapshot(*case["snapshot"][:2], seed=case["snapshot"][2])
    gb = lat.bbox()
    ux, uy, uz = unwrapped(lat.x, lat.y, lat.z, lat.dims, gb)
    print("[program] octree after the periodic correction", flush=True)
    with A.ResidentLattice.from_lattice(lat) as rl:
        rl.periodic_unwrap(rl.bbox())
        with A.Octree(rl, 4000, 4, region_of_interest=gb) as tree:
            assert np.array_equal(tree.bbox(), gb)
            nd, lv = tree.nodes(), tree.leaves()
    want = oracle_octree(Lattice(n=lat.n, box=lat.box, x=ux, y=uy, z=uz), 4000, 4, roi=gb, root_box=gb)
    for k in ("ll", "ext", "level", "parent", "first_child", "count"):
        assert np.array_equal(nd[k], want[k]), k
    for k in ("dfs", "order", "mc_offsets", "mc_ids"):
        assert np.array_equal(lv[k], want[k]), k


def test_dark_sky_reader_matches_reference(tmp_path):
    """The DARK_SKY file (shuffled records) read into the lattice on the GPU gives the reference
    reader's particles in id order, bit for bit, at every stride; a stride that does not divide
    the lattice fails with the reference's message."""
    g = GOLDEN["dark_sky"]
    lat = make_snapshot(*g["snapshot"][:2], seed=g["snapshot"][2])
    path = tmp_path / "ds12.bin"
    write_dark_sky(str(path), lat)
    for stride, want in g["loads"].items():
        print(f"[program] dark sky reader, stride {stride}", flush=True)
        if "error" in want:
            with pytest.raises(A.AdRuntimeException) as e:
                A.ResidentLattice.from_dark_sky(str(path), int(stride))
            assert want["error"].split("AdRuntimeException: ")[-1] in str(e.value)
            continue
        with A.ResidentLattice.from_dark_sky(str(path), int(stride)) as rl:
            assert list(rl.dims) == want["dims"]
            assert sha(np.concatenate(rl.positions())) == want["xyz_sha"]


def test_dark_sky_reader_rejects_inconsistent_files(tmp_path):
    """A repeated site, an id outside the lattice and a truncated record are errors (the
    reference's count check and assertions), not silent garbage."""
    lat = make_snapshot(8, 1.0, seed=1)
    good = tmp_path / "good.bin"
    write_dark_sky(str(good), lat)
    raw = bytearray(good.read_bytes())
    hdr = get_reader("DARK_SKY").info(str(good))["header_bytes"]
    rec = np.frombuffer(raw, dtype=np.int64, offset=hdr).reshape(-1, 4)  # 4 words of 8 bytes per record
    cases = {"repeated": (rec[1, 3], "inconsistent global particle number"),
             "outside": (int(morton_ids(8, 0, 0)), "invalid dark sky id")}
    for name, (new_id, msg) in cases.items():
        bad = bytearray(raw)
        np.frombuffer(bad, dtype=np.int64, offset=hdr).reshape(-1, 4)[0, 3] = new_id
        path = tmp_path / f"{name}.bin"
        path.write_bytes(bytes(bad))
        print(f"[program] dark sky reader, {name} id", flush=True)
        with pytest.raises(A.AdRuntimeException, match=msg):
            A.ResidentLattice.from_dark_sky(str(path))
    path = tmp_path / "short.bin"
    path.write_bytes(bytes(raw[:-5]))
    with pytest.raises(A.AdRuntimeException, match="inconsistent file size"):
        A.ResidentLattice.from_dark_sky(str(path))
    with A.ResidentLattice.from_dark_sky(str(good)) as rl:  # the reader still works after failures
        assert np.array_equal(rl.positions()[0], lat.x)


def _run_program(tmp_path, name, *flags):
    case = PROGRAM_CASES[name]
    lat = make_snapshot(*case["snapshot"][:2], seed=case["snapshot"][2])
    snap = tmp_path / "snap.bin"
    write_case_snapshot(case, str(snap), lat)
    out = tmp_path / "out"
    out.mkdir()
    pfile = tmp_path / "params.txt"
    pfile.write_text(param_file_text(case, str(snap), str(out)))
    print(f"[program] adeco_resample {name} {' '.join(flags)}", flush=True)
    r = subprocess.run([EXE, str(pfile), *flags], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    total = float(r.stdout.split("total mass")[1].split("=")[1].split()[0])
    files = {f: np.fromfile(out / f, np.float32) for f in os.listdir(out)}
    return files, total, pfile


@pytest.mark.parametrize("name", sorted(PROGRAM_CASES))
def test_program_matches_reference_program(tmp_path, name):
    g = GOLDEN["program"][name]
    files, total, _ = _run_program(tmp_path, name, "--reference-geometry")
    assert sorted(files) == sorted(g["files"])
    ref, ref_total, _, _ = oracle_program(name)
    f64, _, _, d = oracle_program(name, fp64=True)
    dims = int(PROGRAM_CASES[name]["params"]["linear_patch_resolution"])
    node_of = {patch_file_name(d, int(n), dims): int(n) for n in d["dfs"]}
    for fname, got in files.items():
        assert sha(ref[fname]) == g["files"][fname], fname  # the reference's own file
        ext = float(np.float32(d["ext"][node_of[fname]]))
        inv = (dims / ext) ** 3  # density of one tet mass per cell, for the error floor
        assert O.max_rel_err(got.astype(np.float64), f64[fname], tet_mass=inv) <= REL, fname
        # what separates this file from the reference's is the reference's own float error
        r = ref[fname].astype(np.float64)
        e = f64[fname]
        assert np.all(np.abs(got - r) <= np.abs(e - r) + REL * np.maximum(np.abs(e), 1e-8 * inv)), fname
    assert abs(total - g["total_mass_printed"]) <= 2e-6 * g["total_mass_printed"]
    assert abs(total - ref_total) <= 2e-6 * ref_total


def test_program_forms_and_exact_geometry(tmp_path):
    """The default geometry (exact dyadic leaf boxes) conserves the mass inside the root box: the
    files' total equals the oracle's to 1e-8 (each file cell is the float rounding of an FP64
    mass, ~1e-10 of the total here).  The plain kernel, the mixed-precision clip and the Python
    mirror (deposit_workflow) write the same files to 1e-6 per cell and the same total to 1e-9."""
    name = "periodic_whole_domain"
    runs = {}
    for key, flags in (("batched", ()), ("plain", ("--plain",)), ("mixed", ("--mixed",))):
        (tmp_path / key).mkdir()
        runs[key] = _run_program(tmp_path / key, name, *flags)
    files, total, pfile = runs["batched"]
    _, _, ulat, d = oracle_program(name)
    gb = np.array(GOLDEN["program"][name]["global_bbox"], np.float32)
    ext = float(max(gb[1] - gb[0], gb[3] - gb[2], gb[5] - gb[4]))
    root = np.array([gb[0], gb[0] + ext, gb[2], gb[2] + ext, gb[4], gb[4] + ext], np.float64)
    inside = O.deposit_lattice(ulat, root, 1, decomp=A.DECOMP_REF_EFFECTIVE, tet_mass=1.0).sum()
    assert abs(total - inside) <= 1e-8 * inside
    assert sorted(files) == sorted(GOLDEN["program"][name]["files"])

    def same(other, t2, what):
        assert sorted(other) == sorted(files), what
        assert abs(t2 - total) <= 1e-9 * total, what
        for f, v in files.items():
            w = v.astype(np.float64)
            assert O.max_rel_err(other[f].astype(np.float64), w, tet_mass=max(float(w.max()), 1.0)) <= REL, (what, f)

    same(*runs["plain"][:2], "plain")
    same(*runs["mixed"][:2], "mixed")
    p = parse_parameter_file(str(pfile))
    p.output_directory = str(tmp_path / "py")
    os.mkdir(p.output_directory)
    res = deposit_workflow(p)
    assert res["num_leaves"] == len(files)
    py = {f: np.fromfile(os.path.join(p.output_directory, f), np.float32) for f in os.listdir(p.output_directory)}
    same(py, res["total_mass"], "python")


def test_write_leaves_errors(tmp_path):
    """Writing leaves before any deposit is EINVAL; an unwritable directory is EIO (the reference
    throws AdRuntimeException); both leave the tree usable."""
    lat = make_snapshot(10, 1.0, seed=2)
    print("[program] write_leaves error paths", flush=True)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1000, 2) as tree:
        with pytest.raises(A.AdRuntimeException) as e:
            tree.write_leaves(str(tmp_path))
        assert e.value.status == A.ADECO_EINVAL
        tree.deposit(4, decomposition=A.DECOMP_REF_EFFECTIVE, tet_mass=1.0, mode=A.MODE_BATCHED)
        with pytest.raises(A.AdRuntimeException) as e:
            tree.write_leaves(str(tmp_path / "missing" / "dir"))
        assert e.value.status == A.ADECO_EIO
        masses, total = tree.write_leaves(str(tmp_path))
        assert len(os.listdir(tmp_path)) == tree.num_leaves
        assert abs(total - tree.leaf_grids().sum()) <= 1e-6 * total
