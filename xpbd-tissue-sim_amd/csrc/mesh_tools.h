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
// colours, or a negative value on bad input.  Entries of -1 are empty slots
// (items with fewer than verts_per_item vertices).
int greedy_color(int nv, int n_items, int verts_per_item, const int32_t* items,
                 std::vector<int32_t>& color);

// Clusters for the cluster Gauss-Seidel variant: runs of consecutive tets
// (caller order) whose union has at most kClusterVerts vertices and at most
// kClusterTets tets.  start[c]..start[c+1] are the cluster's tets; verts holds
// kClusterVerts vertex ids per cluster (-1 padded); tloc[e] packs the local
// slot (0..7) of each of tet e's four vertices in 4-bit fields.  For a
// structured grid generated cell by cell a cluster is one cell (6 tets, 8
// vertices).  Returns the number of clusters.
constexpr int kClusterVerts = 8;
constexpr int kClusterTets = 8;
int build_clusters(int nt, const int32_t* tets, std::vector<int32_t>& start, std::vector<int32_t>& verts,
                   std::vector<uint32_t>& tloc);

// Patches for the LDS Jacobi variant.  Tets are sorted by (instance, Morton
// code of the rest centroid) and cut into patches of `patch` consecutive tets.
// Output (all in solver order):
//   order[i]       caller tet index at solver position i
//   pvoff[p]       range of patch p in pverts (its distinct vertices, sorted)
//   pslot[j]       for pverts entry j: -1 if the vertex is interior to the
//                  patch (all incident tets inside), else its partial slot
//   tloc[2i..2i+1] the 4 local vertex indices of tet i, 16 bits each
//   shared/soff  shared vertex j owns the contiguous partial slots
//                [soff[j], soff[j+1]), in patch order
//   inc_off/inc    per patch local CSR: local vertex -> (tet_local*4 + corner)
//                  entries, offsets relative to the patch (deterministic mode)
// Returns the largest local vertex count of any patch.
struct Patches {
    std::vector<int32_t> order, pvoff, pverts, pslot, shared, soff;  // soff: contiguous slots per shared vertex
    std::vector<uint32_t> tloc;
    std::vector<int32_t> inc_off;  // per patch: nl+1 offsets, stored at pvoff[p] + p
    std::vector<uint16_t> inc;     // per patch: 4*patch entries, stored at p*4*patch
    // per pverts entry of a shared vertex: (shared index, first slot | slot count << 28); (-1, 0) otherwise
    std::vector<int32_t> pshare;
    bool fusable = true;  // pshare can describe every shared vertex (at most 15 patches, 2^28 slots)
    int patch = 256, npatch = 0, maxl = 0, nslots = 0;
};
int build_patches(int nv, int nt, const int32_t* tets, const double* x, const int32_t* tet_instance,
                  int patch, Patches& out);

// GMSH 4.1 ASCII reader: nodes (with tags renumbered densely in file order) and
// 4-node tetrahedra (element type 4).  Triangles (type 2) are returned as faces.
int read_msh41(const std::string& path, std::vector<double>& x,
               std::vector<int32_t>& tets, std::vector<int32_t>& tris, std::string& err);
// Integers one per line, stopping at the first line without one (element
// class files, FileUtils::readVectorFromFile<int>).
int read_int_lines(const std::string& path, std::vector<int32_t>& out, std::string& err);

}  // namespace xpbd
