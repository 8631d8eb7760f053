// Host-side surface meshes, RWG tables and the octree block partition.
//
// Table layout follows the reference's object struct (codes/GPU_implementation/
// seed/compres_single_block_spheres_for_sweep2.jl:12-26) as produced by
// get_edge.m (codes/GPU_implementation/seed/get_edge.m:16-121):
//   vertex 3 x Nv, topol 3 x Nt (1-based vertex ids), trian 3 x Nt (signed
//   1-based edge ids), edges 4 x Ne (T+, T-, free vertex of T+, free vertex of
//   T-; 1-based), un 3 x Nt, ds Nt, ln Ne, cent 3 x Nt. All column-major.
// Inside this library ids are 0-based int32; the C-ABI converts.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace farcur {

struct Mesh {
    std::vector<double> vertex;  // 3*Nv, xyz per vertex
    std::vector<int32_t> topol;  // 3*Nt, 0-based vertex ids
    // RWG tables (filled by build_rwg)
    std::vector<int32_t> edges;  // 4*Ne: T+, T-, v+, v- (0-based)
    std::vector<int32_t> trian;  // 3*Nt: signed (edge id + 1), 0 = no edge
    std::vector<double> ln;      // Ne
    std::vector<double> cent;    // 3*Nt
    std::vector<double> un;      // 3*Nt
    std::vector<double> ds;      // Nt
    int nv() const { return (int)(vertex.size() / 3); }
    int nt() const { return (int)(topol.size() / 3); }
    int ne() const { return (int)(edges.size() / 4); }
};

// Two (or more) meshes joined: vertex arrays concatenated, topol shifted.
void mesh_append(Mesh& dst, const Mesh& src, double dx, double dy, double dz);

// Square plate in the z = z0 plane, Nx x Ny squares, each split into two right
// triangles along the same diagonal (the cor = 0 split of the reference's
// sq_plate.m, codes/ACA_vs_CUR/seed/objects/sq_plate.m:56-63, rotated so the
// plate normal is +z).
Mesh mesh_plate(double lx, double ly, int nx, int ny, double x0, double y0, double z0);

// Octahedral sphere with Ne = 12*4^n edges, vertex and triangle numbering as
// codes/ACA_vs_CUR/seed/objects/sphere.m:13-53.
Mesh mesh_octasphere(double radius, int n_levels);

// Icosphere at the given subdivision level (20*4^L triangles), vertices
// radially jittered by U(-jitter, +jitter) * mean edge length (seeded), then
// scaled by (ax, ay, az) / radius to form an ellipsoid when the axes differ.
Mesh mesh_icosphere(double ax, double ay, double az, int level, double jitter, uint64_t seed);

// RWG tables with get_edge.m semantics (edge order, T+/T- choice, free
// vertices, trian signs, ln, cent, un, ds). Returns false and sets err on a
// non-manifold edge (more than two triangles; junctions are not supported).
bool build_rwg(Mesh& m, std::string& err);

// Octree leaf groups and admissible (far-field) pairs.
struct BlockPartition {
    double leaf_size = 0.5;
    double eta_adm = 2.0;
    int32_t nrwg_mesh = 0;               // RWGs of the mesh the partition was built on
    std::vector<int32_t> group_offsets;  // ngroups + 1
    std::vector<int32_t> group_rwg;      // RWG ids grouped by leaf
    std::vector<double> group_centre;    // 3 * ngroups
    std::vector<double> group_diam;      // ngroups
    std::vector<int32_t> far_pairs;      // 2 * nfar (a < b), admissible
    std::vector<int32_t> near_pairs;     // 2 * nnear (a <= b), not admissible
};

// RWG point = midpoint of the shared edge. Leaf cubes of side leaf_size on a
// grid anchored at the bounding-box minimum. Group centre = mean RWG point;
// diameter = 2 * max distance from centre to any vertex of the group's
// triangles. Pair (a, b) is admissible when |c_a - c_b| >= eta * max(d_a, d_b).
BlockPartition build_partition(const Mesh& m, double leaf_size, double eta_adm);

// Block-list file: the partition with its far (compressed) and near (dense)
// block entries, the role of the reference's Matlab Z_comp list of
// {m, n, comp, self} dicts (codes/GPU_implementation/seed/compres_single_block_spheres.jl:1134-1144,
// self semantics at :2602-2612). Little-endian binary:
//   char magic[8] = "FARCURBL", uint32 version = 1, uint32 ngroups,
//   uint32 nrwg_mesh, uint32 nrwg (grouped), uint32 nfar, uint32 nnear,
//   double leaf_size, double eta_adm,
//   int32 group_offsets[ngroups + 1], int32 group_rwg[nrwg] (0-based RWG ids),
//   double centre[3 * ngroups], double diam[ngroups],
//   then nfar + nnear entries {int32 a, int32 b, uint8 comp, uint8 self, uint16 pad}:
//   comp = 1 for a far pair (a < b), 0 for a near pair (a <= b); self = 1 when
//   the entry stands for block (a, b) alone (a == b), 0 when it also stands for
//   the transposed block (b, a) = (a, b)^T of the symmetric operator,
//   uint64 checksum (FNV-1a of every byte before it).
bool save_partition(const BlockPartition& P, const std::string& path, std::string& err);
bool load_partition(const std::string& path, BlockPartition& P, std::string& err);

}  // namespace farcur
