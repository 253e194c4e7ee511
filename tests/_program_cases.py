"""Inputs of the program-level golden cases (checker side): parameter files, SORTED_BY_ID and
DARK_SKY snapshots and whole runs of the reference program.  tests/golden/make_program_golden.py runs the
compiled reference on them; the tests run this library and the oracle on the same inputs.
"""

from __future__ import annotations

import numpy as np

# Parameter-file texts for parse_parameter_file (main.cpp:83-183).  Each is parsed as written.
PARAM_TEXTS = {
    # the example of the reference's README.txt:131-144
    "readme": ("input_file               =/scratch/user/test_file.bin\n"
               "output_directory         =/scratch/user/\n"
               "file_format              = SORTED_BY_ID\n"
               "use_region_of_interest   = true\n"
               "region_of_interest       = 0. 0.5 0. 0.5 0. 0.5 \n"
               "max_refinement_level     = 8\n"
               "stride                   = 1 \n"
               "max_tets_per_octree_node = 40000\n"
               "linear_patch_resolution  = 64\n"
               "num_mpi_ranks_per_node   = 1\n"
               "num_pthreads\t\t = 4\n"
               "periodic_boundaries\t = false\n"),
    "minimal": "input_file = a.bin\nfile_format = SORTED_BY_ID\n",
    "empty": "",
    # no blank before '=': the key loses its last letter and matches nothing; a commented-out
    # line still matches by substring; flags need a lower-case "true"; -1 wraps to the maximum;
    # integers stop at the first non-digit; values keep their blanks
    "quirks": ("input_file=/not/taken.bin\n"
               "input_file = /x/snap.bin\n"
               "# stride = 4 (commented out, still read)\n"
               "use_region_of_interest = TRUE\n"
               "max_tets_per_octree_node = -1\n"
               "max_refinement_level = 7.9\n"
               "linear_patch_resolution=32\n"
               "output_directory =  /out dir  \n"
               "id_type = DARK_SKY\n"
               "=\n"
               "restart_path = /r\n"
               "periodic_boundaries = yes\n"
               "region_of_interest = 1e1 2.5e1 -3 4 0.1 0.7\n"),
    "roi_five_values": "region_of_interest = 1 2 3 4 5\n",
    "bad_number": "stride = abc\n",
    "bad_roi_number": "region_of_interest = 1 2 x 4 5 6\n",
    # at most 10000 line reads, the one that meets the end of the file included: a file of 9998
    # lines is read, one of 9999 is rejected
    "lines_9998": "\n" * 9996 + "stride = 2\nmax_refinement_level = 9\n",
    "lines_9999": "\n" * 9997 + "stride = 2\nmax_refinement_level = 9\n",
    # without '=' the whole line is both key and value
    "no_equals": "use_region_of_interest true\ninput_file /a/b.bin\n",
}

# Whole runs of the program: a snapshot (n particles per axis, rms displacement in spacings,
# seed) and its parameters; input_file and output_directory are filled in per run.
PROGRAM_CASES = {
    "periodic_whole_domain": {
        "snapshot": (24, 2.0, 3),
        "params": {"file_format": "SORTED_BY_ID", "use_region_of_interest": "false", "max_refinement_level": "4",
                   "stride": "1", "max_tets_per_octree_node": "4000", "linear_patch_resolution": "8",
                   "num_mpi_ranks_per_node": "1", "num_pthreads": "1", "periodic_boundaries": "true"},
    },
    # the periodic case's snapshot as a DARK_SKY file (shuffled records, Morton ids): the same
    # lattice, so the same leaf files
    "dark_sky_periodic": {
        "snapshot": (24, 2.0, 3),
        "format": "DARK_SKY",
        "params": {"file_format": "DARK_SKY", "use_region_of_interest": "false", "max_refinement_level": "4",
                   "stride": "1", "max_tets_per_octree_node": "4000", "linear_patch_resolution": "8",
                   "num_mpi_ranks_per_node": "1", "num_pthreads": "1", "periodic_boundaries": "true"},
    },
    # every second particle per axis of a 26^3 DARK_SKY snapshot (13^3 kept)
    "dark_sky_stride2": {
        "snapshot": (26, 1.5, 8),
        "format": "DARK_SKY",
        "params": {"file_format": "DARK_SKY", "use_region_of_interest": "false", "max_refinement_level": "3",
                   "stride": "2", "max_tets_per_octree_node": "1500", "linear_patch_resolution": "8",
                   "num_mpi_ranks_per_node": "1", "num_pthreads": "1", "periodic_boundaries": "true"},
    },
    "roi_not_periodic": {
        "snapshot": (20, 1.5, 5),
        "params": {"file_format": "SORTED_BY_ID", "use_region_of_interest": "true",
                   "region_of_interest": "20 60 10 50 30 70", "max_refinement_level": "3", "stride": "1",
                   "max_tets_per_octree_node": "2500", "linear_patch_resolution": "12",
                   "num_mpi_ranks_per_node": "1", "num_pthreads": "1", "periodic_boundaries": "false"},
    },
}


def case_lattice(case: dict):
    """The lattice a case's program run deposits: the snapshot, every stride-th particle per axis
    for a DARK_SKY file read with a stride (the SORTED_BY_ID reader ignores the stride)."""
    from adeco_amd.lattice import Lattice, make_snapshot
    lat = make_snapshot(*case["snapshot"][:2], seed=case["snapshot"][2])
    s = int(case["params"].get("stride", "1"))
    if case.get("format") != "DARK_SKY" or s == 1:
        return lat
    n = lat.dims[0]
    sub = [c.reshape(n, n, n)[::s, ::s, ::s].ravel().copy() for c in (lat.x, lat.y, lat.z)]
    return Lattice(n=n // s - 1, box=lat.box, x=sub[0], y=sub[1], z=sub[2])


def write_case_snapshot(case: dict, path: str, lat) -> None:
    """Writes a case's snapshot in its file format (SORTED_BY_ID unless stated)."""
    from adeco_amd.lattice import write_dark_sky, write_sorted_by_id
    if case.get("format") == "DARK_SKY":
        write_dark_sky(path, lat)
    else:
        write_sorted_by_id(path, lat)


# DARK_SKY headers for the header parser: (text, sha1 records to append)
DARK_SKY_HEADERS = {
    "valid": ("# SDF 1.0\nint sha1_chunks = 3;\nstruct {\n  float x;\n}[8];\n# SDF-EOH\n", 3),
    "no_semicolon": ("int sha1_chunks = 1\n# SDF-EOH more text\n", 1),
    "no_end_of_header": ("# SDF 1.0\nint sha1_chunks = 2;\n", 0),
    "no_sha1_chunks": ("# SDF 1.0\n# SDF-EOH\n", 0),
    # a line of 1023 characters fits the reader's 1024-byte line buffer; one of 1024 does not
    "line_1023": ("# SDF 1.0\n#" + "x" * 1022 + "\nint sha1_chunks = 2;\n# SDF-EOH\n", 2),
    "line_1024": ("# SDF 1.0\n#" + "x" * 1023 + "\nint sha1_chunks = 2;\n# SDF-EOH\n", 2),
}


def param_file_text(case: dict, input_file: str, output_directory: str) -> str:
    lines = [f"input_file               ={input_file}", f"output_directory         ={output_directory}"]
    lines += [f"{k:<25}= {v}" for k, v in case["params"].items()]
    return "\n".join(lines) + "\n"


def unwrap_lattice_case():
    """A non-cubic (13, 9, 7) lattice with positions wrapped into a 50^3 box (seeded), for the
    periodic correction alone."""
    rng = np.random.default_rng(11)
    dims = (13, 9, 7)
    k, j, i = np.meshgrid(*(np.arange(d) for d in dims[::-1]), indexing="ij")
    spacing = np.array([50.0 / dims[0], 50.0 / dims[1], 50.0 / dims[2]])
    out = []
    for a, idx in enumerate((i, j, k)):
        p = idx * spacing[a] + rng.normal(0.0, 1.5 * spacing[a], idx.shape)
        out.append(np.mod(p, 50.0).astype(np.float32).ravel())
    return dims, out[0], out[1], out[2]
