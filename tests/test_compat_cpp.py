"""The C++ facade with the reference's host names (include/pond_compat.hpp): a host program in
the reference's style (parsHost, set, fill, init, calcStep, drop, diagnostics) compiles against
it and links with libpond.so; on a GPU it runs the step and its dropped fields equal the
oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "lbmpond_amd")

PROGRAM = r"""
#include "pond_compat.hpp"
#include <cmath>
int main(int argc, char** argv) {
  using namespace lbmpond;
  parsHost.Nx = 16; parsHost.Ny = 8; parsHost.Nz = 4;
  parsHost.set();
  parsHost.lattice = POND_D3Q27;
  parsHost.interp = POND_WENO5C;
  parsHost.fill = [](long x, long y, long, double& rho, double& ux, double& uy, double& uz, double& T) {
    rho = 1 + 0.05 * std::cos(2 * M_PI * x / 16.0);
    ux = 0.02;
    uy = 0.01 * std::sin(2 * M_PI * y / 8.0);
    uz = 0;
    T = 0.01;
  };
  try {
    init();
    std::string dir = argc > 1 ? argv[1] : ".";
    for (int s = 0; s < 3; s++) calcStep();
    drop(&dir);
    diagnostics();
    std::vector<Cinfo> ci;
    get_cinfo(ci);
    std::printf("iStep %d NiterMax %d niter0 %d mass %.17g enrg %.17g\n", parsHost.iStep, parsHost.NiterMax,
                ci[0].niter, parsHost.mass, parsHost.enrg);
    LatticeTables t = lattice_tables(POND_D3Q27);
    std::printf("Qn %d TLat %.17g\n", t.Qn, t.TLat);
  } catch (const Error& e) {
    std::printf("error %d: %s\n", (int)e.status, e.what());
    return 3;
  }
  parsHost.clear();
  return 0;
}
"""


def build_program(tmp_path):
    src = tmp_path / "prog.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "prog"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-L", LIBDIR, "-lpond", f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    return exe


def test_facade_compiles_and_links(tmp_path):
    assert build_program(tmp_path).exists()


@pytest.mark.gpu
def test_facade_program_matches_oracle(gpu, tmp_path):
    import pyoracle as po
    from lbmpond_amd import read_arr
    exe = build_program(tmp_path)
    out = subprocess.run([str(exe), str(tmp_path)], check=True, capture_output=True, text=True, timeout=120).stdout
    print(out, flush=True)
    assert "iStep 3" in out and "Qn 27" in out
    x, y = np.meshgrid(np.arange(16), np.arange(8), indexing="xy")
    x = np.tile(x.reshape(-1), 4).astype(np.float64)
    y = np.tile(y.reshape(-1), 4).astype(np.float64)
    rho = 1 + 0.05 * np.cos(2 * np.pi * x / 16.0)
    ux = np.full_like(rho, 0.02)
    uy = 0.01 * np.sin(2 * np.pi * y / 8.0)
    uz = np.zeros_like(rho)
    T = np.full_like(rho, 0.01)
    o = po.Oracle("D3Q27", (16, 8, 4), "weno5c")
    o.set_state(rho, ux, uy, uz, T)
    o.step(3)
    ref = o.fields()
    for name, key in (("Rho", "rho"), ("Vx", "ux"), ("Vy", "uy"), ("Vz", "uz"), ("T", "T")):
        got = read_arr(str(tmp_path / f"{name}_0000003.arr")).reshape(-1)
        assert np.array_equal(got, ref[key]), name
    assert f"NiterMax {int(ref['niter'].max())} niter0 {int(ref['niter'][0])}" in out
