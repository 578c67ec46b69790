// Per-step output of the reference's application (reference
// apps/polyscope_app.h:131-149, 335-370): the current positions as an OBJ
// surface and as a Medit .mesh, and the integrator's last state (x_prev,
// v_prev) as libigl ASCII DMAT, which the scene loader's "initial_state"
// reads back.
#include "output.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <map>
#include <stdexcept>

namespace mfem {

std::vector<int> boundary_faces(const std::vector<double>& V, const std::vector<int>& T) {
  // A face is on the boundary when exactly one tet has it. Each is written
  // with its normal pointing out of its tet (corner order from the tet's
  // signed volume), in (tet, local face) order.
  const size_t nT = T.size() / 4;
  std::map<std::array<int, 3>, int> count;
  auto key = [](int a, int b, int c) {
    std::array<int, 3> k{a, b, c};
    std::sort(k.begin(), k.end());
    return k;
  };
  static const int opp[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  for (size_t e = 0; e < nT; ++e)
    for (const auto& f : opp) ++count[key(T[4 * e + f[0]], T[4 * e + f[1]], T[4 * e + f[2]])];
  std::vector<int> F;
  for (size_t e = 0; e < nT; ++e) {
    const int* t = &T[4 * e];
    double d[3][3];
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) d[k][c] = V[3 * t[k + 1] + c] - V[3 * t[0] + c];
    const double vol = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                       d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                       d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
    for (const auto& f : opp) {
      if (count[key(t[f[0]], t[f[1]], t[f[2]])] != 1) continue;
      // the listed order is outward for a positive tet
      if (vol >= 0) F.insert(F.end(), {t[f[0]], t[f[1]], t[f[2]]});
      else F.insert(F.end(), {t[f[0]], t[f[2]], t[f[1]]});
    }
  }
  return F;
}

void write_obj(const std::string& path, const std::vector<double>& X, const std::vector<int>& F) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  const size_t nV = X.size() / 3;
  for (size_t i = 0; i < nV; ++i)
    std::fprintf(f, "v %.17g %.17g %.17g\n", X[3 * i], X[3 * i + 1], X[3 * i + 2]);
  for (size_t k = 0; k + 2 < F.size(); k += 3)
    std::fprintf(f, "f %d %d %d\n", F[k] + 1, F[k + 1] + 1, F[k + 2] + 1);
  std::fclose(f);
}

void write_medit(const std::string& path, const std::vector<double>& X, const std::vector<int>& T,
                 const std::vector<int>& F) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  const size_t nV = X.size() / 3;
  std::fprintf(f, "MeshVersionFormatted 1\nDimension 3\nVertices\n%zu\n", nV);
  for (size_t i = 0; i < nV; ++i)
    std::fprintf(f, "%.17g %.17g %.17g 0\n", X[3 * i], X[3 * i + 1], X[3 * i + 2]);
  std::fprintf(f, "Triangles\n%zu\n", F.size() / 3);
  for (size_t k = 0; k + 2 < F.size(); k += 3)
    std::fprintf(f, "%d %d %d 0\n", F[k] + 1, F[k + 1] + 1, F[k + 2] + 1);
  std::fprintf(f, "Tetrahedra\n%zu\n", T.size() / 4);
  for (size_t k = 0; k + 3 < T.size(); k += 4)
    std::fprintf(f, "%d %d %d %d 0\n", T[k] + 1, T[k + 1] + 1, T[k + 2] + 1, T[k + 3] + 1);
  std::fprintf(f, "End\n");
  std::fclose(f);
}

void write_dmat(const std::string& path, const std::vector<double>& A, int rows, int cols) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "%d %d\n", cols, rows);
  for (int c = 0; c < cols; ++c)
    for (int r = 0; r < rows; ++r) std::fprintf(f, "%.17g\n", A[(size_t)r * cols + c]);
  std::fclose(f);
}

}  // namespace mfem
