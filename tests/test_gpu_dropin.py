"""The drop-in boundary used the way the reference uses it.

* Several host threads call DepositWrapper.deposit_tets at once, as the reference's worker threads
  do (AdDynamicLoadBalancing.h:198-205: every worker deposits its share of a leaf into its own
  grid, all sharing the library), each thread on its own meta-patches and grid, several calls in
  a row; every grid matches the extended-precision oracle to 1e-6 per cell (floor 1e-8 of a tet
  mass), in the plain and in the batched form.
* The reference program itself (main.cpp, deposit_workflow, deposit_leaf_nodes, its thread pool
  and write_patch, compiled from the reference sources) with this library dropped in: the
  reference's AdDeposit2Grid.cpp replaced by adeco_amd/cpp/AdDeposit2Grid_adeco.cpp
  (oracle/_ref/libadeco_ref_dropin.so, built by oracle/build_ref.sh; test infrastructure only).
  It writes the reference program's own file set, and every leaf file matches the FP64 oracle's
  densities in the reference's geometry to 1e-6 per cell (floor 1e-8 of one tet's mass per cell
  volume); the printed total matches the reference program's to 2e-6.
"""

import json
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice, make_snapshot, meta_patches
from tests import _oracle as O
from tests._golden_leaves import patch_file_name
from tests._program_cases import PROGRAM_CASES, param_file_text, write_case_snapshot
from tests._program_oracle import oracle_program

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DROPIN = os.path.join(ROOT, "oracle", "_ref", "libadeco_ref_dropin.so")
LIB = os.path.join(ROOT, "adeco_amd", "libadeco_amd.so")
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_program.json")))
REL = 1e-6

# per thread: lattice cubes, rms displacement, seed, grid dims
THREAD_CASES = [(10, 2.0, 1, 16), (12, 3.0, 2, 24), (8, 1.0, 3, 12), (14, 2.5, 4, 20), (9, 1.5, 5, 32), (11, 2.0, 6, 8)]
CALLS_PER_THREAD = 3


@pytest.mark.parametrize("mode", [A.MODE_PLAIN, A.MODE_BATCHED])
def test_boundary_called_from_several_host_threads(mode):
    inputs = []
    for n, rms, seed, dims in THREAD_CASES:
        lat = make_lattice(n, rms, seed=seed)
        pat = meta_patches(lat)
        box = np.asarray(A.cubic_root_box(lat.bbox()), np.float32)
        with O.cell_scaling(O.EXACT, library=False):
            want = O.deposit_patches_ld(pat, box, dims)
        inputs.append((pat, box, dims, want))
    grids = [np.zeros(d ** 3, np.float32) for _, _, d, _ in inputs]
    errors = []
    start = threading.Barrier(len(inputs))

    def worker(i):
        pat, box, dims, _ = inputs[i]
        try:
            start.wait()
            for _ in range(CALLS_PER_THREAD):
                A.DepositWrapper.deposit_tets(True, 0, 0, pat.size // 1536, pat, pat.size, 7, box, dims, grids[i])
        except Exception as exc:  # reported by the main thread
            errors.append((i, exc))

    prev = A.set_boundary_form(decomposition=A.DECOMP_REF_EFFECTIVE, tet_mass=1.0, mode=mode, precision=A.PREC_FP64)
    try:
        print(f"[dropin] {len(inputs)} host threads x {CALLS_PER_THREAD} calls, mode {mode}", flush=True)
        threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(inputs))]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        assert not any(th.is_alive() for th in threads), "a boundary call did not return"
    finally:
        A.set_boundary_form(**prev)
    assert not errors, errors
    for i, (g, (_, _, dims, want)) in enumerate(zip(grids, inputs)):
        # CALLS_PER_THREAD accumulations of the same deposit (+=, one float rounding per call)
        err = O.max_rel_err(g.astype(np.float64) / CALLS_PER_THREAD, want)
        assert err <= REL, (i, err)


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="oracle/_ref/libadeco_ref_dropin.so not built")
@pytest.mark.parametrize("name", ["periodic_whole_domain", "roi_not_periodic"])
def test_reference_program_with_the_library_dropped_in(tmp_path, name):
    case = PROGRAM_CASES[name]
    lat = make_snapshot(*case["snapshot"][:2], seed=case["snapshot"][2])
    snap = tmp_path / "snap.bin"
    write_case_snapshot(case, str(snap), lat)
    out = tmp_path / "out"
    out.mkdir()
    pfile = tmp_path / "params.txt"
    pfile.write_text(param_file_text(case, str(snap), str(out)))
    # a child process: the reference program finalises MPI.  The library deposits in the
    # reference's geometry (its float leaf boxes with 1/dx scaling).
    code = ("import ctypes, sys\n"
            "lib = ctypes.CDLL(sys.argv[1])\n"
            "assert lib.adeco_set_cell_scaling(1) == 0\n"
            "ref = ctypes.CDLL(sys.argv[2])\n"
            "ref.ref_cli_main.argtypes = [ctypes.c_char_p]\n"
            "sys.exit(ref.ref_cli_main(sys.argv[3].encode()))\n")
    print(f"[dropin] reference program with the library, {name}", flush=True)
    r = subprocess.run([sys.executable, "-c", code, LIB, DROPIN, str(pfile)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert "ERROR" not in r.stdout, r.stdout[-2000:]
    total = None
    for line in r.stdout.splitlines():
        if "total mass" in line:
            total = float(line.split("=")[-1])
    g = GOLDEN["program"][name]
    files = {f: np.fromfile(out / f, np.float32) for f in os.listdir(out)}
    assert sorted(files) == sorted(g["files"])
    f64, _, _, d = oracle_program(name, fp64=True)
    dims = int(case["params"]["linear_patch_resolution"])
    node_of = {patch_file_name(d, int(n), dims): int(n) for n in d["dfs"]}
    for fname, got in files.items():
        ext = float(np.float32(d["ext"][node_of[fname]]))
        inv = (dims / ext) ** 3  # density of one tet mass per cell
        assert O.max_rel_err(got.astype(np.float64), f64[fname], tet_mass=inv) <= REL, fname
    assert total is not None
    assert abs(total - g["total_mass_printed"]) <= 2e-6 * g["total_mass_printed"]
