// side_table.hpp -- side-node ordinals of every element topology, shared by
// the host mesh set-up and the device kernels.  Index [topology][side
// ordinal][k] with the OAGPU_HEX8 / TET4 / WED6 / PYR5 codes 0..3: the
// side's nodes as element-local ordinals in side order (counter-clockwise
// seen from outside the element, so sub-face area vectors point out); -1
// pads tri sides and the ordinals a topology lacks.  The side order is also
// the order of the side's face ips (face ip k is nearest to side node k).
//   hex8     HexSCS side ordinals (Hex8CVFEM.C)
//   tet4     TetSCS sideNodeOrdinals_ (Tet4CVFEM.C:416-421)
//   wedge6   WedSCS sideNodeOrdinals_ (Wed6CVFEM.C:494-501)
//   pyramid5 PyrSCS sideNodeOrdinals_ (Pyr5CVFEM.C:759-765)
#pragma once

#define OAGPU_SIDE_TABLE                                                                  \
    {{{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}},     \
     {{0, 1, 3, -1}, {1, 2, 3, -1}, {0, 3, 2, -1}, {0, 2, 1, -1}, {-1, -1, -1, -1},            \
      {-1, -1, -1, -1}},                                                                       \
     {{0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}, {0, 2, 1, -1}, {3, 4, 5, -1}, {-1, -1, -1, -1}}, \
     {{0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}, {0, 3, 2, 1}, {-1, -1, -1, -1}}}

// number of sides per topology
#define OAGPU_NSIDES {6, 4, 5, 5}

namespace oagpu {
// doubles per boundary-side record of a tet4 / wedge6 / pyramid5 side: four
// face ips of (area vector, dN/dx [6][3] at the standard and at the shifted
// face ip)
constexpr int kSideIpStride = 3 + 2 * 6 * 3;
constexpr int kSideRecord = 4 * kSideIpStride;
}  // namespace oagpu
