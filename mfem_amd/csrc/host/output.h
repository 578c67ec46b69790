// Per-step output files (reference apps/polyscope_app.h:131-149, 335-370).
#pragma once
#include <string>
#include <vector>

namespace mfem {

// Boundary triangles of a tetrahedral mesh (V nV x 3, T nT x 4): the faces
// of exactly one tet, normals out of their tet, in (tet, local face) order.
std::vector<int> boundary_faces(const std::vector<double>& V, const std::vector<int>& T);
// OBJ with vertices X (nV x 3) and triangles F (0-based in, 1-based out)
void write_obj(const std::string& path, const std::vector<double>& X, const std::vector<int>& F);
// Medit .mesh with vertices X, boundary triangles F and tets T, ending "End"
void write_medit(const std::string& path, const std::vector<double>& X, const std::vector<int>& T,
                 const std::vector<int>& F);
// libigl ASCII DMAT of a rows x cols row-major matrix (column-major on disk)
void write_dmat(const std::string& path, const std::vector<double>& A, int rows, int cols);

}  // namespace mfem
