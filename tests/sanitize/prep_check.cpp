// Sanitizer check driver for the library's host setup (test infrastructure).
// Runs prepare() (dg-bgk_amd/csrc/prep.cpp) on a case read from raw files in
// the input element order, and prepare_storage(), the whole setup of a device
// context (Hilbert order, border refinement, point CSR), and writes the
// reference tables, the reflection table and the tile count back for
// comparison with the library's own setup entry points.  Built with
// -fsanitize=address,undefined by tests/sanitize/Makefile.
//   prep_check <case_dir> <out_dir>
#include <string>
#include <vector>

#include "io.h"
#include "prep.hpp"

using namespace dgbgk;

template <class T>
static void put(const char* dir, const char* name, const std::vector<T>& v) {
  write_file(dir, name, v.data(), v.size() * sizeof(T));
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: prep_check <case_dir> <out_dir>\n");
    return 2;
  }
  case_input in = read_case(argv[1]);
  check_size("params.bin", in.params_bytes, sizeof(dgbgk_params));
  const Params prm = params_from_api(*(const dgbgk_params*)in.params);
  std::string msg;
  Prep p;
  int st = prepare(in.nelem, in.npoin, in.nboun, in.intma, in.coord, in.bsido, in.rorder,
                   in.torder, in.eta, in.w, prm, 256, 256, p, msg);
  if (st) {
    fprintf(stderr, "prepare: status %d: %s\n", st, msg.c_str());
    return 1;
  }
  const char* out = argv[2];
  put(out, "iside.i32", p.iside);
  put(out, "nx.f64", p.nx);
  put(out, "ny.f64", p.ny);
  put(out, "el.f64", p.el);
  put(out, "geome.f64", p.geome);
  put(out, "relen.f64", p.relen);
  put(out, "mmat.f64", p.mmat);
  put(out, "rsido.f64", p.rsido);
  put(out, "vcord.f64", p.vcord);
  put(out, "wall_sides.i32", p.wall_sides);
  put(out, "refl_i1.i32", p.refl_i1);
  put(out, "refl_i2.i32", p.refl_i2);
  put(out, "refl_zd1.f64", p.refl_zd1);
  put(out, "refl_zd2.f64", p.refl_zd2);
  std::vector<double> scal = {p.sumweight, p.dtec};
  put(out, "scalars.f64", scal);
  Prep q;
  std::vector<int> eorder;
  st = prepare_storage(in.nelem, in.npoin, in.nboun, in.intma, in.coord, in.bsido, in.rorder,
                       in.torder, in.eta, in.w, prm, 256, 256, true, true, q, eorder, msg);
  if (st) {
    fprintf(stderr, "prepare_storage: status %d: %s\n", st, msg.c_str());
    return 1;
  }
  std::vector<int> ntiles = {(int)q.tile_e0.size()};
  put(out, "ntiles.i32", ntiles);
  put(out, "eorder.i32", eorder);
  free_case(&in);
  return 0;
}
