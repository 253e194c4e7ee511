"""Writes the program-level golden fixtures from the compiled reference (build container only).

Requires oracle/_ref/libadeco_ref.so (oracle/build_ref.sh).  Only outputs are stored, in
ref_program.json next to this script:

  params    parse_parameter_file (main.cpp:83-183) on the texts of tests/_program_cases.py
  reader    AdSortedByIdsPointDataReader::load_data on a small valid file and on invalid ones
  unwrap    apply_boundary_corrections on a non-cubic wrapped lattice (sha256 of x|y|z after)
  dark_sky  AdDarkSkyDataReader's header sizes and errors, and its loads of a shuffled DARK_SKY file
            at several strides (the positions in id order: the one-rank redistribution's lattice)
  program   the reference program (main.cpp:187-249) run on SORTED_BY_ID and DARK_SKY snapshots: its leaf file
            names, their sha256, the total mass it prints, and the sha256 of the deposited lattice
            and of its corrected coordinates (apply_boundary_corrections with the lattice's bbox)

Each program run happens in a child process: the program finalises MPI at its end.
"""

from __future__ import annotations

import ctypes
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from adeco_amd.lattice import make_snapshot, write_dark_sky, write_sorted_by_id  # noqa: E402
from tests import _oracle as O  # noqa: E402
from tests._program_cases import (DARK_SKY_HEADERS, PARAM_TEXTS, PROGRAM_CASES, case_lattice,  # noqa: E402
                                  param_file_text, unwrap_lattice_case, write_case_snapshot)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_parse(r, path):
    cap = 4096
    strs = ctypes.create_string_buffer(4 * cap)
    ints = np.zeros(8, np.int64)
    roi = np.zeros(6, np.float32)
    if r.ref_parse_parameter_file(path.encode(), strs, cap, O.P(ints), O.P(roi)) != 0:
        return {"error": r.ref_last_error().decode()}
    s = [strs.raw[i * cap:(i + 1) * cap].split(b"\0", 1)[0].decode() for i in range(4)]
    keys = ["use_regions_of_interest", "max_refinement_level", "stride", "max_tets_per_octree_node",
            "linear_patch_resolution", "num_pthreads", "num_mpi_ranks_per_node", "periodic_boundaries"]
    out = {"input_file": s[0], "output_directory": s[1], "id_type": s[2], "restart_path": s[3]}
    out.update({k: int(v) & (2 ** 64 - 1) for k, v in zip(keys, ints)})
    out["region_of_interest"] = [float(v) for v in roi]
    return out


def params_fixture(r, tmp):
    res = {}
    for name, text in PARAM_TEXTS.items():
        path = os.path.join(tmp, f"{name}.txt")
        with open(path, "w") as f:
            f.write(text)
        res[name] = ref_parse(r, path)
        if "error" in res[name]:
            res[name]["error"] = res[name]["error"].replace(path, "<file>")
    res["missing_file"] = ref_parse(r, os.path.join(tmp, "no_such_file.txt"))
    res["missing_file"]["error"] = res["missing_file"]["error"].replace(tmp, "<dir>")
    return res


def ref_load(r, path):
    cap = 1 << 16
    x, y, z = (np.zeros(cap, np.float32) for _ in range(3))
    ids = np.zeros(cap, np.int64)
    dims = np.zeros(3, np.int32)
    n = r.ref_sorted_by_id_load(path.encode(), cap, O.P(x), O.P(y), O.P(z), O.P(ids), O.P(dims))
    if n < 0:
        return {"error": r.ref_last_error().decode()}
    return {"n": int(n), "dims": dims.tolist(), "xyz_sha": sha(np.concatenate([x[:n], y[:n], z[:n]])),
            "ids_in_order": bool(np.array_equal(ids[:n], np.arange(n)))}


def reader_fixture(r, tmp):
    lat = make_snapshot(6, 1.0, seed=2)
    good = os.path.join(tmp, "good.bin")
    write_sorted_by_id(good, lat)
    odd = os.path.join(tmp, "odd.bin")
    np.zeros(13, np.uint8).tofile(odd)
    noncubic = os.path.join(tmp, "noncubic.bin")
    np.zeros(30, np.float32).tofile(noncubic)
    empty = os.path.join(tmp, "empty.bin")
    open(empty, "wb").close()
    out = {"good": ref_load(r, good), "odd_size": ref_load(r, odd), "not_cubic": ref_load(r, noncubic),
           "empty": ref_load(r, empty), "missing": ref_load(r, os.path.join(tmp, "no_such_file.bin"))}
    for key in ("empty", "missing"):
        if "error" in out[key]:
            out[key]["error"] = out[key]["error"].replace(tmp, "<dir>")
    out["good"]["snapshot"] = [6, 1.0, 2]
    return out


def unwrap_fixture(r):
    dims, x, y, z = unwrap_lattice_case()
    bbox = np.array([x.min(), x.max(), y.min(), y.max(), z.min(), z.max()], np.float32)
    assert r.ref_apply_boundary_corrections(O.P(x), O.P(y), O.P(z), *dims, O.P(bbox)) == 0
    return {"dims": list(dims), "bbox": bbox.tolist(), "xyz_sha": sha(np.concatenate([x, y, z]))}


def dark_sky_fixture(r, tmp):
    headers = {}
    for name, (text, chunks) in DARK_SKY_HEADERS.items():
        path = os.path.join(tmp, f"hdr_{name}.bin")
        with open(path, "wb") as f:
            f.write(text.encode() + bytes(24 * chunks) + bytes(32 * 8))
        h = r.ref_dark_sky_header_size(path.encode())
        headers[name] = {"header_bytes": int(h)} if h >= 0 else {"error": r.ref_last_error().decode().replace(tmp, "<dir>")}
    lat = make_snapshot(12, 2.0, seed=3)
    path = os.path.join(tmp, "ds12.bin")
    write_dark_sky(path, lat)
    loads = {}
    for stride in (1, 2, 3, 5):
        cap = 4096
        x, y, z = (np.zeros(cap, np.float32) for _ in range(3))
        ids = np.zeros(cap, np.int64)
        dims = np.zeros(3, np.int32)
        n = r.ref_dark_sky_load(path.encode(), stride, cap, O.P(x), O.P(y), O.P(z), O.P(ids), O.P(dims))
        if n < 0:
            loads[str(stride)] = {"error": r.ref_last_error().decode()}
            continue
        # the one-rank redistribution sorts the particles by id: the lattice in id order
        order = np.argsort(ids[:n], kind="stable")
        loads[str(stride)] = {"n": int(n), "dims": dims.tolist(), "ids_unique": bool(np.unique(ids[:n]).size == n),
                              "xyz_sha": sha(np.concatenate([x[:n][order], y[:n][order], z[:n][order]]))}
    return {"headers": headers, "snapshot": [12, 2.0, 3], "loads": loads}


def program_fixture(r, tmp):
    out = {}
    for name, case in PROGRAM_CASES.items():
        lat = make_snapshot(*case["snapshot"])
        d = os.path.join(tmp, name)
        os.makedirs(os.path.join(d, "out"))
        snap = os.path.join(d, "snap.bin")
        write_case_snapshot(case, snap, lat)
        pfile = os.path.join(d, "params.txt")
        with open(pfile, "w") as f:
            f.write(param_file_text(case, snap, os.path.join(d, "out")))
        run = subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, sys.argv[2]); "
                              "from tests import _oracle as O; sys.exit(O.load_ref().ref_cli_main(sys.argv[1].encode()))",
                              pfile, ROOT], capture_output=True, text=True, timeout=600)
        assert run.returncode == 0, run.stderr[-2000:]
        m = re.search(r"total mass\s*=\s*([0-9.eE+-]+)", run.stdout)
        files = {f: hashlib.sha256(open(os.path.join(d, "out", f), "rb").read()).hexdigest()
                 for f in sorted(os.listdir(os.path.join(d, "out")))}
        lat = case_lattice(case)  # what the program deposits (a DARK_SKY stride subsamples)
        bbox = lat.bbox()
        x, y, z = lat.x.copy(), lat.y.copy(), lat.z.copy()
        unwrap_sha = None
        if case["params"]["periodic_boundaries"] == "true":
            assert r.ref_apply_boundary_corrections(O.P(x), O.P(y), O.P(z), *lat.dims, O.P(bbox)) == 0
            unwrap_sha = sha(np.concatenate([x, y, z]))
        out[name] = {"snapshot_sha": sha(np.concatenate([lat.x, lat.y, lat.z])), "files": files,
                     "total_mass_printed": float(m.group(1)), "global_bbox": bbox.tolist(), "unwrap_sha": unwrap_sha}
        print(f"program {name}: {len(files)} leaf files, total mass {m.group(1)}", flush=True)
    return out


def main():
    r = O.load_ref()
    if r is None:
        sys.exit("oracle/_ref/libadeco_ref.so is not built (run oracle/build_ref.sh)")
    with tempfile.TemporaryDirectory() as tmp:
        fx = {"params": params_fixture(r, tmp), "reader": reader_fixture(r, tmp), "unwrap": unwrap_fixture(r),
              "dark_sky": dark_sky_fixture(r, tmp), "program": program_fixture(r, tmp)}
    with open(os.path.join(HERE, "ref_program.json"), "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote ref_program.json")


if __name__ == "__main__":
    main()
