"""The library's element storage order, checked on the host without a GPU:
a small C++ driver built from prep.cpp reads a mesh, orders its elements
along the Hilbert curve, builds the update tiles, groups each tile's border
elements by reading tile and rebuilds the tiles from their bounds, then
prints what it found for the checks below."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from small_cases import cylinder_small

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dg-bgk_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <set>
#include <vector>
#include "prep.hpp"
using namespace dgbgk;

static std::vector<int> readi(const char* path, size_t n) {
  std::vector<int> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(int), n, f) != n) std::exit(3);
  std::fclose(f);
  return v;
}
static std::vector<double> readd(const char* path, size_t n) {
  std::vector<double> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(double), n, f) != n) std::exit(3);
  std::fclose(f);
  return v;
}
// maximal runs of consecutive element ids in each tile's halo list
static long halo_runs(const Prep& p) {
  long runs = 0;
  for (size_t t = 0; t < p.tile_e0.size(); ++t)
    for (int j = 0; j < p.tile_nhalo[t]; ++j) {
      const int h = p.halo_elems[p.tile_hoff[t] + j];
      if (j == 0 || h != p.halo_elems[p.tile_hoff[t] + j - 1] + 1) ++runs;
    }
  return runs;
}

int main(int argc, char** argv) {
  const int E = std::atoi(argv[1]), P = std::atoi(argv[2]), B = std::atoi(argv[3]);
  const std::vector<int> intma = readi(argv[4], 3 * (size_t)E), bsido = readi(argv[5], 6 * (size_t)B);
  const std::vector<double> coord = readd(argv[6], 2 * (size_t)P);
  const int R = std::atoi(argv[7]), T = std::atoi(argv[8]);
  const std::vector<double> q = readd(argv[9], 2 * (size_t)R + 14);  // eta, w, deck values
  const double* eta = q.data();
  const double* w = q.data() + R;
  const double* d = q.data() + 2 * R;
  Params prm{};
  prm.rv = d[0]; prm.csafm = d[1]; prm.alpha = d[2]; prm.rgas = d[3]; prm.d = d[4];
  prm.m = d[5]; prm.t1 = d[6]; prm.p1 = d[7]; prm.u0 = d[8]; prm.v0 = d[9];
  for (int k = 0; k < 4; ++k) prm.cinf[k] = d[10 + k];
  prm.inf = 1; prm.immat = 1; prm.lookup = 1;
  std::vector<int> order;
  const int moved = hilbert_order(E, P, intma.data(), coord.data(), order);
  std::printf("hilbert_moved %d\n", moved);
  std::vector<int> im = intma, bs = bsido;
  if (moved) renumber_elements(E, B, intma.data(), bsido.data(), order, im, bs);
  std::string msg;
  Prep a, b;
  if (prepare(E, P, B, im.data(), coord.data(), bs.data(), R, T, eta, w, prm, 256, 256, a, msg)) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    return 4;
  }
  std::vector<int> perm, bounds;
  border_order(a, perm, bounds);
  std::vector<int> fin(E);
  for (int i = 0; i < E; ++i) fin[i] = order.empty() ? perm[i] : order[perm[i]];
  std::vector<int> im2, bs2;
  renumber_elements(E, B, intma.data(), bsido.data(), fin, im2, bs2);
  if (prepare(E, P, B, im2.data(), coord.data(), bs2.data(), R, T, eta, w, prm, 256, 256, b, msg,
              &bounds)) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    return 5;
  }
  // perm permutes inside every tile's owned range
  bool inside = true;
  for (size_t t = 0; t < a.tile_e0.size(); ++t) {
    const int e0 = a.tile_e0[t], e1 = e0 + a.tile_nown[t];
    std::set<int> got(perm.begin() + e0, perm.begin() + e1);
    inside = inside && (int)got.size() == e1 - e0 && *got.begin() == e0 && *got.rbegin() == e1 - 1;
  }
  // interior elements first, then border elements grouped by the lowest reading tile
  std::vector<int> reader(E, -1);
  for (int t = (int)b.tile_e0.size() - 1; t >= 0; --t)
    for (int j = 0; j < b.tile_nhalo[t]; ++j) reader[b.halo_elems[b.tile_hoff[t] + j]] = t;
  bool grouped = true;
  for (size_t t = 0; t < b.tile_e0.size(); ++t)
    for (int e = b.tile_e0[t] + 1; e < b.tile_e0[t] + b.tile_nown[t]; ++e)
      grouped = grouped && reader[e - 1] <= reader[e];
  // halo lists in storage order
  bool sorted = true;
  for (size_t t = 0; t < b.tile_e0.size(); ++t)
    for (int j = 1; j < b.tile_nhalo[t]; ++j)
      sorted = sorted && b.halo_elems[b.tile_hoff[t] + j - 1] < b.halo_elems[b.tile_hoff[t] + j];
  std::printf("tiles %zu %zu\n", a.tile_e0.size(), b.tile_e0.size());
  std::printf("same_tiles %d\n", (int)(a.tile_e0 == b.tile_e0 && a.tile_nown == b.tile_nown &&
                                       a.tile_nhalo == b.tile_nhalo));
  std::printf("perm_inside_tiles %d\n", (int)inside);
  std::printf("border_grouped %d\n", (int)grouped);
  std::printf("halo_sorted %d\n", (int)sorted);
  std::printf("halo_runs %ld %ld\n", halo_runs(a), halo_runs(b));
  long nh = 0;
  for (int x : b.tile_nhalo) nh += x;
  std::printf("halo_elems %ld\n", nh);
  // the final order is a permutation of the input elements
  std::set<int> all(fin.begin(), fin.end());
  std::printf("final_is_permutation %d\n", (int)((int)all.size() == E && *all.begin() == 0 &&
                                                 *all.rbegin() == E - 1));
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    d = tmp_path_factory.mktemp("order")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{CSRC}", str(src),
                    os.path.join(CSRC, "prep.cpp"), "-o", str(exe)], check=True)
    return exe, d


def _run(driver, case, mesh=None):
    exe, d = driver
    mesh = mesh if mesh is not None else case.mesh
    paths = [d / "intma.bin", d / "bsido.bin", d / "coord.bin"]
    np.ascontiguousarray(mesh.intma, dtype=np.int32).tofile(paths[0])
    np.ascontiguousarray(mesh.bsido, dtype=np.int32).tofile(paths[1])
    np.ascontiguousarray(mesh.coord, dtype=np.float64).tofile(paths[2])
    pr = case.prm
    deck = [pr["rv"], pr["csafm"], pr["alpha"], pr["rgas"], pr["d"], pr["m"], pr["t1"],
            pr["p1"], pr["u0"], pr["v0"]] + list(pr["cinf"])
    np.concatenate([case.eta, case.w, deck]).astype(np.float64).tofile(d / "grid.bin")
    out = subprocess.run([str(exe), str(mesh.nelem), str(mesh.npoin), str(mesh.nboun)] +
                         [str(p) for p in paths] +
                         [str(case.rorder), str(case.torder), str(d / "grid.bin")],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.splitlines():
        k, *v = line.split()
        res[k] = [int(x) for x in v]
    return res


def test_generated_mesh_keeps_its_hilbert_order_and_borders_group(driver):
    """The mesh generator's element order is already the library's Hilbert
    order; grouping the border elements keeps every tile's element set (the
    rebuilt tiles are the same), puts each tile's interior first and its
    border elements grouped by reading tile, and leaves each tile's halo
    in far fewer runs of consecutive elements."""
    r = _run(driver, cylinder_small(elems=6000))
    assert r["hilbert_moved"] == [0]
    assert r["same_tiles"] == [1]
    assert r["perm_inside_tiles"] == [1]
    assert r["border_grouped"] == [1]
    assert r["halo_sorted"] == [1]
    assert r["final_is_permutation"] == [1]
    before, after = r["halo_runs"]
    assert after < 0.3 * before, (before, after)


def test_shuffled_mesh_gets_the_same_tiles(driver):
    """A mesh in random element order is sorted back to the same Hilbert
    order, so it gets the generator's tiles."""
    from dgbgk import meshgen
    case = cylinder_small(elems=6000)
    sh, _ = meshgen.shuffle_elements(case.mesh, seed=9)
    a, b = _run(driver, case), _run(driver, case, sh)
    assert b["hilbert_moved"] == [1]
    assert a["tiles"] == b["tiles"]
    assert a["halo_elems"] == b["halo_elems"]
    assert a["halo_runs"] == b["halo_runs"]
