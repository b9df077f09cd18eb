// Driver of the INTEGRATION.md code blocks (included as the generated
// snippets.inc): runs the L1 binding on a system read from a file and an L2
// island read from a mesh file, writing results for the Python test to check.
//
//   integration_main l1 <system.bin> <x.bin>
//     system.bin: int32 n, nnzb; int32 ptr[n + 1]; int32 col[nnzb];
//                 float values[9 nnzb] (row strips); float b[3 n]
//     x.bin: float x[3 n]; prints "iters convergence residual relative"
//   integration_main l2 <mesh.bin> <positions.bin> <steps>
//     mesh.bin: int32 nV, nT, nPinned; double X[3 nV]; int32 tets[4 nT];
//               int32 pinned[nPinned]; double plate point[3], normal[3],
//               velocity[3]; double E, nu, density, penalty, dt
//     positions.bin: double x[3 nV] after the steps; prints each step's
//     statistics and the snapshot round trip's equality flag
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "snippets.inc"

namespace {

template <class V>
void read(std::ifstream& f, V* p, size_t n) {
  f.read(reinterpret_cast<char*>(p), std::streamsize(n * sizeof(V)));
  if (!f) throw std::runtime_error("short input file");
}

int run_l1(const char* in, const char* out) {
  std::ifstream f(in, std::ios::binary);
  int32_t n = 0, nnzb = 0;
  read(f, &n, 1);
  read(f, &nnzb, 1);
  std::vector<int32_t> ptr(size_t(n) + 1), col(static_cast<size_t>(nnzb));
  std::vector<float> vals(size_t(9) * size_t(nnzb)), b(size_t(3) * size_t(n)), x(size_t(3) * size_t(n), 0.f);
  read(f, ptr.data(), ptr.size());
  read(f, col.data(), col.size());
  read(f, vals.data(), vals.size());
  read(f, b.data(), b.size());
  const SdxLinearStatus st = SdxSolveBsr3(n, nnzb, ptr.data(), col.data(), vals.data(), b.data(), x.data(),
                                          /*absTol=*/0.0, /*relTol=*/1e-6, /*relDivTol=*/1e10, /*maxIter=*/1000,
                                          /*normType=*/0, /*blockJacobi=*/true, /*abortIfNotSpd=*/false);
  std::ofstream o(out, std::ios::binary);
  o.write(reinterpret_cast<const char*>(x.data()), std::streamsize(x.size() * sizeof(float)));
  std::printf("%d %d %.17g %.17g\n", st.numIterDone, st.convergence, st.residualNorm, st.relativeResidualNorm);
  return 0;
}

int run_l2(const char* in, const char* out, int steps) {
  std::ifstream f(in, std::ios::binary);
  int32_t nv = 0, nt = 0, np = 0;
  read(f, &nv, 1);
  read(f, &nt, 1);
  read(f, &np, 1);
  std::vector<double> X(size_t(3) * size_t(nv));
  std::vector<int32_t> tets(size_t(4) * size_t(nt)), pins(static_cast<size_t>(np));
  double plate[9], mat[5];
  read(f, X.data(), X.size());
  read(f, tets.data(), tets.size());
  read(f, pins.data(), pins.size());
  read(f, plate, 9);
  read(f, mat, 5);
  SdxIsland isl;
  if (SdxIslandCreate(isl, X, tets, pins, mat[0], mat[1], mat[2], mat[3], plate, plate + 3) != SDX_OK) {
    std::fprintf(stderr, "create: %s\n", sdx_last_error());
    return 1;
  }
  const double dt = mat[4];
  const sdx_solver_params prm = SdxPinnedParams(1, 1e-6);
  std::vector<double> x;
  for (int s = 0; s < steps; ++s) {
    double pos[3];
    for (int d = 0; d < 3; ++d) pos[d] = plate[d] + plate[6 + d] * dt * double(s + 1);
    sdx_step_stats st;
    if (SdxIslandStep(isl, dt, pos, prm, st, x) != SDX_OK) {
      std::fprintf(stderr, "step: %s\n", sdx_last_error());
      return 1;
    }
    std::printf("step %d newton %d linear %d conv %d linconv %d truerel %.3e\n", s, st.newton_iters,
                st.linear_iters, st.convergence, st.linear_convergence, st.linear_true_rel_res);
  }
  int32_t equal = -1;
  double pos[3];
  for (int d = 0; d < 3; ++d) pos[d] = plate[d] + plate[6 + d] * dt * double(steps + 1);
  if (SdxSnapshotRoundTrip(isl, dt, pos, prm, equal) != SDX_OK) {
    std::fprintf(stderr, "snapshot: %s\n", sdx_last_error());
    return 1;
  }
  std::printf("snapshot equal %d\n", equal);
  std::ofstream o(out, std::ios::binary);
  o.write(reinterpret_cast<const char*>(x.data()), std::streamsize(x.size() * sizeof(double)));
  SdxIslandDestroy(isl);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc == 4 && std::strcmp(argv[1], "l1") == 0) return run_l1(argv[2], argv[3]);
    if (argc == 5 && std::strcmp(argv[1], "l2") == 0) return run_l2(argv[2], argv[3], std::atoi(argv[4]));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: integration_main l1 <system.bin> <x.bin> | l2 <mesh.bin> <positions.bin> <steps>\n");
  return 2;
}
