// Host-side mesh preparation used by the solver set-up: structured tet grids,
// surface extraction, conflict colouring and the GMSH 4.1 reader.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace xpbd {

// Structured grid of nx*ny*nz hexahedral cells, each split into 6 tetrahedra
// along its main diagonal (Kuhn / Freudenthal split: every tet is a monotone
// path from corner (0,0,0) to corner (1,1,1)).  The split is conforming across
// cells because every cell uses the same diagonal direction.
//   vertex (i,j,k) -> i + (nx+1)*(j + (ny+1)*k)       (x fastest)
//   cells visited k, j, i (x fastest); 6 tets per cell in axis-permutation
//   order (x,y,z) (x,z,y) (y,x,z) (y,z,x) (z,x,y) (z,y,x).
void grid_mesh(int nx, int ny, int nz, double hx, double hy, double hz,
               const double origin[3], std::vector<double>& x, std::vector<int32_t>& tets);

// Boundary triangles of a tet mesh (faces owned by exactly one tet), found by
// hashing sorted vertex triples.  Order: by owning tet, then local face
// (1,2,3) (0,3,2) (0,1,3) (0,2,1); winding is the tet's local winding.
// Returns a negative value if some face is shared by more than two tets.
int surface_faces(int nv, int nt, const int32_t* tets, std::vector<int32_t>& faces);

// Conflict colouring of items (tets: 4 vertices, faces: 3 vertices) so that no
// two items of one colour share a vertex: first-fit greedy in index order,
// then iterated greedy (re-colouring class by class in changing class orders,
// which never adds a colour), then a balancing pass that moves items to the
// least-populated admissible class.  Deterministic.  Returns the number of
// colours, or a negative value on bad input.
int greedy_color(int nv, int n_items, int verts_per_item, const int32_t* items,
                 std::vector<int32_t>& color);

// GMSH 4.1 ASCII reader: nodes (with tags renumbered densely in file order) and
// 4-node tetrahedra (element type 4).  Triangles (type 2) are returned as faces.
int read_msh41(const std::string& path, std::vector<double>& x,
               std::vector<int32_t>& tets, std::vector<int32_t>& tris, std::string& err);

}  // namespace xpbd
