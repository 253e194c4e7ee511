"""The C++ faces of the boundary.

* adeco_amd/cpp/AdDeposit2Grid_adeco.cpp implements the reference's own DepositWrapper from the
  reference's header: it must compile against it (signature identity), checked where the
  reference sources are present.
* include/adeco_deposit_wrapper.hpp: a small C++ program using the reference's call pattern is
  compiled, linked against libadeco_amd.so and run.  Without a GPU it must report the error as an
  exception (no exit, no CPU fallback); on the GPU it deposits one lattice, prints the mass and
  writes the grid, which matches the oracle per cell.
"""

import os
import subprocess
import tempfile

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice, meta_patches

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "adeco_deposit_wrapper.hpp"
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    std::vector<float> pos(std::stoul(argv[2]));
    if (std::fread(pos.data(), 4, pos.size(), f) != pos.size()) return 3;
    std::fclose(f);
    float bbox[6];
    for (int i = 0; i < 6; ++i) bbox[i] = std::stof(argv[3 + i]);
    const int dims = 16;
    std::vector<float> grid(dims * dims * dims, 0.f);
    try {
        AdaptiveMassDeposit::DepositWrapper::deposit_tets(true, 0, 0, pos.size() / 1536, pos.data(), pos.size(), 7,
                                                          bbox, dims, grid.data());
    } catch (const adeco::DepositError& e) {
        std::printf("error %d %s\n", int(e.status()), e.what());
        return 0;
    }
    double m = 0;
    for (float g : grid) m += g;
    std::printf("mass %.6f\n", m);
    FILE* o = std::fopen(argv[9], "wb");
    if (!o || std::fwrite(grid.data(), 4, grid.size(), o) != grid.size()) return 4;
    std::fclose(o);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources absent")
def test_dropin_compiles_against_reference_header():
    r = subprocess.run(["g++", "-std=gnu++11", "-fsyntax-only", f"-I{REF_SRC}", f"-I{ROOT}/include",
                        os.path.join(ROOT, "adeco_amd", "cpp", "AdDeposit2Grid_adeco.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _run_program():
    lat = make_lattice(8, 1.0, seed=1)
    pat = meta_patches(lat)
    box = A.cubic_root_box(lat.bbox())
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        data = os.path.join(d, "pos.bin")
        open(src, "w").write(PROGRAM)
        pat.tofile(data)
        lib_dir = os.path.join(ROOT, "adeco_amd")
        r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", src, "-o", exe, f"-L{lib_dir}",
                            "-l:libadeco_amd.so", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        gfile = os.path.join(d, "grid.bin")
        r = subprocess.run([exe, data, str(pat.size)] + [repr(float(v)) for v in box] + [gfile], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        grid = np.fromfile(gfile, np.float32) if os.path.exists(gfile) else None
        return r.stdout.strip(), lat, pat, np.asarray(box, np.float32), grid


def test_cpp_program_reports_edevice_without_gpu():
    if A.device_count() > 0:
        pytest.skip("a GPU is present")
    out = _run_program()[0]
    assert out.startswith(f"error {A.ADECO_EDEVICE}"), out


@pytest.mark.gpu
def test_cpp_program_deposits_on_gpu():
    """The program's grid, cell by cell, against the extended-precision oracle's deposit of the same
    meta-patches (the library's default geometry: per-axis cell scaling) to 1e-6 (floor 1e-8 of a
    tet mass), and the total to 1e-6."""
    from tests import _oracle as O
    out, lat, pat, box, grid = _run_program()
    assert out.startswith("mass "), out
    with O.cell_scaling(O.EXACT, library=False):
        want = O.deposit_patches_ld(pat, box, 16)
    assert O.max_rel_err(grid.astype(np.float64), want) <= 1e-6
    mass = float(out.split()[1])
    assert abs(mass - want.sum()) <= 1e-6 * want.sum()
