// mesh.hpp -- device mirror of an OpenAccel hex8 mesh.
#pragma once

#include "common.hpp"

namespace oagpu {

// Host-side description handed over at the C ABI (reference layouts: STK
// node coordinates [node][3], element connectivity [elem][8] in the nalu hex8
// node order, boundary sides as parent element + side ordinal + boundary id).
struct HostMesh {
    int nn = 0, ne = 0, nbf = 0;
    const double* xyz = nullptr;
    const int* conn = nullptr;      // [ne][8], trailing entries of smaller elements unused
    const int* etype = nullptr;     // [ne] OAGPU_HEX8 / TET4 / WED6 / PYR5; null: all hex8
    const int* bf_elem = nullptr;
    const int* bf_ord = nullptr;
    const int* bf_bc = nullptr;
    bool ordered = false;           // numbering already in locality order
};

// Everything the row-owned kernels need, resident in HBM.
//
// * CRS node graph (reduced or full stencil) with sorted local columns and
//   the diagonal position of every row (nodeGraph.cpp:29-392).
// * node -> element incidences, stored local-node major: inc_la[la][row] is
//   the element in which `row` is local node la (-1 if none), so that the
//   row kernels loop over la = 0..7 with compile-time indices and every lane
//   of a wavefront takes the same path; a node that is local node la of more
//   than one element (unstructured meshes) keeps the others in a small CSR
//   list (ext_*).
// * per incidence, the element -> CRS scatter map of that row: 8 bytes, the
//   position in the row of each of the element's 8 nodes, 0xFF when the
//   column is not in the (reduced) row and is lumped into the diagonal block
//   (assembler.h:423-445).
// * node -> boundary-side incidences (side << 2 | face ip whose nearest node
//   is the row), with their own scatter maps; sides of any topology (quad4
//   sides of hex8 / wedge6 / pyramid5, tri3 sides of tet4 / wedge6 /
//   pyramid5, one face ip per side node).
// * meshes with tet4 / wedge6 / pyramid5 elements: inc_la / ext hold the
//   hex8 incidences only; the others (elem << 3 | local node) with their
//   scatter maps (one byte per element node) form a CSR list gen_* per row,
//   walked by the mixed-mesh variants of the row kernels; glist lists the
//   non-hex elements for the element-owned kernels.
struct DevMesh {
    int nn = 0, ne = 0, nbf = 0, nnz = 0, ninc = 0, nbinc = 0, max_row = 0, next = 0;
    int reduced = 1;
    bool mixed = false;        // some element is not hex8
    int ngen = 0, nglist = 0;
    DBuf<double> xyz;          // [nn][3]
    DBuf<int> conn;            // [ne][8]
    DBuf<int> etype;           // [ne] (mixed meshes only)
    DBuf<int> gen_ptr, gen;    // [nn+1], [ngen]
    DBuf<uint64_t> gen_slots;  // [ngen]
    DBuf<int> glist;           // [nglist] non-hex elements, ascending
    DBuf<double> geo;          // static ip geometry of the non-hex elements (topo::ip_geom)
    DBuf<long long> geo_off;   // [ne] offset of an element's geometry in geo (-1: hex8)
    // boundary sides of non-hex elements: per side kSideStride doubles of
    // face-ip geometry (area vector, dN/dx standard / shifted), sgeo_off[f]
    // the side's offset (-1: a hex8 side); slist lists those sides
    int nslist = 0;
    DBuf<int> slist;           // [nslist]
    DBuf<long long> sgeo_off;  // [nbf] (mixed meshes only)
    DBuf<double> sgeo;
    DBuf<int> row_ptr, cols, diag_off;
    // Values of the CRS matrix are stored sliced-ELL: slices of 64 rows (one
    // wavefront), slot-major and lane-contiguous, so a wavefront's lanes (one
    // row each) read and write consecutive addresses.  Component k of the
    // N x N block in slot s of row r lives at
    //   ((slice_ptr[r / 64] + s) * N * N + k) * 64 + r % 64.
    // Slots beyond a row's length (up to its slice's width) hold the row's
    // own column and value 0.
    int nslice = 0;
    size_t nsell = 0;          // slice_ptr[nslice] * 64
    DBuf<int> slice_ptr, slice_w, sell_cols;
    // Traversal orders of the row-owned and element-owned kernels: the 64-row
    // slices and the 64-element groups sorted along a Morton curve of their
    // centroids.  Wavefronts resident at one time then cover a compact region
    // of the mesh, and element / node data shared by neighbouring rows is
    // still in the XCD's L2 at its second use.
    int negroup = 0;
    DBuf<int> slice_order;     // [nslice]
    DBuf<int> egroup_order;    // [negroup]
    DBuf<int> inc_ptr, inc;    // [nn+1], [ninc] element-ordered (set-up kernels)
    DBuf<int> inc_la;          // [8][nn]
    DBuf<uint64_t> slots_la;   // [8][nn]
    DBuf<int> ext_ptr, ext;    // [nn+1], [next]: (elem << 3 | la) beyond the first per la
    DBuf<uint64_t> ext_slots;  // [next]
    DBuf<int> binc_ptr, binc;  // [nn+1], [nbinc]
    DBuf<uint64_t> binc_slots; // [nbinc]
    DBuf<int> bf_elem, bf_ord, bf_bc;
    DBuf<int> bnode;           // node lies on a wall
    DBuf<double> vol;          // dual nodal volume [nn]
    DBuf<double> scvw;         // SCV weights of the element-centre average [ne][8]
    DBuf<int> err;             // device error word (non-positive Jacobians)

    // host copies kept for the bit-exact graph/scatter-map checks
    std::vector<int> h_row_ptr, h_cols, h_diag_off, h_inc_ptr, h_inc, h_binc_ptr, h_binc, h_etype;
    std::vector<int> h_slice_ptr, h_slice_w;
    std::vector<uint64_t> h_inc_slots, h_binc_slots;

    // Caller <-> device numbering.  The device holds nodes and elements in
    // locality order (locality_order): n_old[i] / e_old[i] is the caller's id
    // of device node / element i, n_new / e_new the inverse maps.  Empty when
    // the mesh keeps the caller's numbering.  Boundary sides keep the
    // caller's order.
    std::vector<int> n_old, n_new, e_old, e_new;
    bool renumbered() const { return !n_old.empty(); }
};

// Node and element order for locality of the row-owned kernels: nodes along
// a Morton curve of per-axis layer indices, elements by their smallest node.
void locality_order(const HostMesh& hm, std::vector<int>& n_old, std::vector<int>& e_old);
// build_mesh on the caller's mesh, renumbered by locality_order when
// `renumber` is set (the permutations are kept in dm for the ABI's copies).
void build_mesh_ordered(const HostMesh& hm, int reduced, bool renumber, DevMesh& dm);
void build_mesh(const HostMesh& hm, int reduced, DevMesh& dm);
// CRS graph in the caller's numbering (columns sorted, diagonal offsets)
void caller_graph(const DevMesh& dm, std::vector<int>& row_ptr, std::vector<int>& cols,
                  std::vector<int>& diag_off);
// Value conversions between the caller's block-CRS (caller numbering) and the
// device's sliced-ELL layout (device numbering).
void sell_to_caller_crs(const DevMesh& dm, int nb, const double* sell, double* crs);
void caller_crs_to_sell(const DevMesh& dm, int nb, const double* crs, double* sell);
// value layout conversions (host): CRS blocks of nb values <-> sliced-ELL
void sell_to_crs(const DevMesh& dm, int nb, const double* sell, double* crs);
void crs_to_sell(const DevMesh& dm, int nb, const double* crs, double* sell);
void compute_mesh_geometry(DevMesh& dm, hipStream_t s);

}  // namespace oagpu
