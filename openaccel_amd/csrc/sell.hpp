// sell.hpp -- sliced-ELL pattern of a CRS matrix (64-row slices, slot-major,
// lane-contiguous), shared by the mesh graph and the AMG levels.
#pragma once

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace oagpu {

// Host description: slice s covers rows [64 s, 64 s + 64) and slots
// [slice_ptr[s], slice_ptr[s + 1]); width = longest row of the slice.  Column
// index of slot q of row r sits at (slice_ptr[r / 64] + q) * 64 + r % 64;
// padding slots repeat the row's own index (with value 0).
struct SellHost {
    int n = 0, nslice = 0;
    size_t nsell = 0;  // slice_ptr[nslice] * 64
    std::vector<int> slice_ptr, slice_w, cols;
};

inline SellHost make_sell(int n, const std::vector<int>& row_ptr, const std::vector<int>& cols) {
    SellHost s;
    s.n = n;
    s.nslice = (n + 63) / 64;
    s.slice_ptr.assign(size_t(s.nslice) + 1, 0);
    s.slice_w.assign(size_t(s.nslice), 0);
    for (int sl = 0; sl < s.nslice; ++sl) {
        int w = 0;
        for (int r = 64 * sl; r < std::min(n, 64 * sl + 64); ++r)
            w = std::max(w, row_ptr[r + 1] - row_ptr[r]);
        s.slice_w[sl] = w;
        s.slice_ptr[sl + 1] = s.slice_ptr[sl] + w;
    }
    s.nsell = size_t(s.slice_ptr[s.nslice]) * 64;
    s.cols.assign(s.nsell, 0);
    for (int r = 0; r < n; ++r) {
        const int len = row_ptr[r + 1] - row_ptr[r];
        const int base = s.slice_ptr[r / 64];
        for (int q = 0; q < s.slice_w[r / 64]; ++q)
            s.cols[size_t(base + q) * 64 + (r % 64)] = q < len ? cols[row_ptr[r] + q] : r;
    }
    return s;
}

// position of slot q (component k of nb) of row r
inline size_t sell_index(const std::vector<int>& slice_ptr, int r, int q, int nb, int k) {
    return (size_t(slice_ptr[r / 64] + q) * nb + k) * 64 + size_t(r % 64);
}

}  // namespace oagpu
