// tables.hpp -- hex8 CVFEM master-element tables as compile-time constants.
//
// Kernels instantiate their per-node work for each local node LA (0..7) so
// that every ip, facet and node index below is a constant expression after
// unrolling: element data then lives in registers (no dynamically indexed
// private arrays) and the shape-function derivatives at the fixed ip
// locations fold to the constants they are.
//
// Sources: Hex8CVFEM.C:264-276 (lrscv), :318-344 (ip locations),
// :346-377/:434-446 (exposed-face ips), :419-426 (side ordinals),
// :428-432 (node locations); MasterElementWork.F:41-53 (SCS facets) and
// :72-184 (27-point subdivision).
#pragma once

namespace oagpu {
namespace tbl {

constexpr int LR[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6},
                           {6, 7}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// the three SCS ips touching each node, and +1 when the node is their left node
constexpr int NODE_IP[8][3] = {{0, 3, 8}, {0, 1, 9},  {1, 2, 10}, {2, 3, 11},
                               {4, 7, 8}, {4, 5, 9}, {5, 6, 10}, {6, 7, 11}};
constexpr int NODE_SG[8][3] = {{1, 1, 1},  {-1, 1, 1},  {-1, 1, 1},  {-1, -1, 1},
                               {1, 1, -1}, {-1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}};

// SCS facet points in the 27-point subdivision
constexpr int FACET[12][4] = {{20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23},
                              {11, 25, 26, 12}, {13, 20, 26, 17}, {17, 14, 24, 26},
                              {17, 15, 23, 26}, {16, 17, 26, 25}, {19, 20, 26, 25},
                              {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};

// subdivision points 8..25 as node averages (-1: two-node edge midpoint)
constexpr int SUB[18][4] = {{0, 1, -1, -1}, {1, 2, -1, -1}, {2, 3, -1, -1}, {3, 0, -1, -1},
                            {0, 1, 2, 3},   {4, 5, -1, -1}, {5, 6, -1, -1}, {6, 7, -1, -1},
                            {7, 4, -1, -1}, {4, 5, 6, 7},   {1, 5, -1, -1}, {0, 4, -1, -1},
                            {0, 1, 4, 5},   {3, 7, -1, -1}, {2, 6, -1, -1}, {2, 3, 6, 7},
                            {1, 2, 5, 6},   {0, 3, 4, 7}};

// SCS ip locations: [0] standard, [1] edge midpoint ("shifted")
constexpr double LOC[2][12][3] = {
    {{0.00, -0.25, -0.25}, {0.25, 0.00, -0.25}, {0.00, 0.25, -0.25}, {-0.25, 0.00, -0.25},
     {0.00, -0.25, 0.25},  {0.25, 0.00, 0.25},  {0.00, 0.25, 0.25},  {-0.25, 0.00, 0.25},
     {-0.25, -0.25, 0.00}, {0.25, -0.25, 0.00}, {0.25, 0.25, 0.00},  {-0.25, 0.25, 0.00}},
    {{0.00, -0.50, -0.50}, {0.50, 0.00, -0.50}, {0.00, 0.50, -0.50}, {-0.50, 0.00, -0.50},
     {0.00, -0.50, 0.50},  {0.50, 0.00, 0.50},  {0.00, 0.50, 0.50},  {-0.50, 0.00, 0.50},
     {-0.50, -0.50, 0.00}, {0.50, -0.50, 0.00}, {0.50, 0.50, 0.00},  {-0.50, 0.50, 0.00}}};

// node location signs (2 * isoparametric coordinate)
constexpr int SIG[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

constexpr int SIDE[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                            {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};

}  // namespace tbl
}  // namespace oagpu
