"""The C++ host class (turbostructred_amd/host/ts_kernel.hpp): it compiles with the host
compiler alone, and its grid construction -- the block compression of Grid::InitLocal
(src/kernel/grid.h:164-218) -- is pinned by closure and ratio properties."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "turbostructred_amd", "host")

PROBE = r'''
#include <cstdio>
#include "ts_kernel.hpp"
int main() {
    using namespace tsh;
    std::vector<double> X, H;
    std::vector<BlockNode> u(1);
    BuildAxis(u, 10, 2.0, 1, X, H);
    for (size_t i = 0; i < X.size(); i++) std::printf("U %.17g %.17g\n", X[i], H[i]);
    BlockNode l, c;
    l.N_cells = 50; l.q_com = 1.0 / 1.05; c.pos = 0.5; c.N_cells = 50; c.q_com = 1.05;
    std::vector<BlockNode> b{l, c};
    BuildAxis(b, 100, 1.0, 1, X, H);
    for (size_t i = 0; i < X.size(); i++) std::printf("C %.17g %.17g\n", X[i], H[i]);
    return 0;
}
'''


@pytest.fixture(scope="module")
def axes():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        exe = os.path.join(d, "probe")
        open(src, "w").write(PROBE)
        pkg = os.path.join(ROOT, "turbostructred_amd")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST, src, "-o", exe, "-L", pkg, "-lts_fvm",
                        f"-Wl,-rpath,{pkg}"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = {"U": [], "C": []}
    for line in out.split("\n"):
        if line:
            k, x, h = line.split()
            rows[k].append((float(x), float(h)))
    return {k: np.array(v) for k, v in rows.items()}


def test_host_header_compiles_without_hip():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                    os.path.join(HOST, "shock_tube_driver.cpp")], check=True)


def test_uniform_axis(axes):
    X, H = axes["U"][:, 0], axes["U"][:, 1]
    assert len(X) == 12
    np.testing.assert_array_equal(H, 0.2)
    np.testing.assert_allclose(X, (np.arange(12) - 0.5) * 0.2, rtol=0, atol=1e-15)


def test_compressed_axis(axes):
    """Two geometric blocks meeting at 0.5: the interior sizes add up to L, each block's sizes
    form a geometric sequence with its ratio, each block spans exactly its interval, and the cell
    centres sit half a size from the faces."""
    X, H = axes["C"][:, 0], axes["C"][:, 1]
    h = H[1:-1]
    assert abs(h.sum() - 1.0) <= 1e-14
    assert abs(h[:50].sum() - 0.5) <= 1e-14
    np.testing.assert_allclose(h[1:50] / h[:49], 1 / 1.05, rtol=1e-13)
    np.testing.assert_allclose(h[51:] / h[50:-1], 1.05, rtol=1e-13)
    faces = np.concatenate([[0.0], np.cumsum(h)])
    np.testing.assert_allclose(X[1:-1], 0.5 * (faces[:-1] + faces[1:]), rtol=0, atol=1e-14)
    # ghost layers: the left one repeats the first size; on the right the progression goes on
    # (the running step is multiplied by q after every cell, grid.h:192-197, 212-216)
    assert H[0] == H[1] and H[-1] == H[-2] * 1.05
    assert abs(X[0] - (X[1] - H[1])) <= 1e-15
    assert abs(X[-1] - (X[-2] + 0.5 * H[-2] * (1 + 1.05))) <= 1e-15
