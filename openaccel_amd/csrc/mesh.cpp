// mesh.cpp -- host construction of the device mesh mirror: CRS node graph,
// node -> element incidences and element -> CRS scatter maps.
//
// The graph follows nodeGraph::buildGraph_ (src/mesh/graph/nodeGraph.cpp:29-392):
// reduced stencil = for every SCS ip (il, ir) both nodes enter rows il and ir
// (:68-96); full stencil = all element node pairs (:98-118); columns sorted
// ascending (:392, sortPrimaryIndices_), all rows owned (one GPU holds the
// whole mesh).  The scatter maps encode applyCoeff_'s sorted search of each
// element column in a row and its lumping of absent columns into the diagonal
// block (src/assemble/assembler.h:380-445).
#include "mesh.hpp"
#include "oagpu.h"
#include "sell.hpp"
#include "side_table.hpp"

#include <algorithm>
#include <functional>
#include <numeric>

namespace oagpu {

static const int kLrscv[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6},
                                  {6, 7}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// SCS left / right nodes of tet4, wedge6, pyramid5 (Tet4CVFEM.C, Wed6CVFEM.C,
// Pyr5CVFEM.C lrscv_)
static const int kTetLr[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
static const int kWedLr[9][2] = {{0, 1}, {1, 2}, {0, 2}, {3, 4}, {4, 5}, {3, 5}, {0, 3}, {1, 4}, {2, 5}};
static const int kPyrLr[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 4}, {0, 4},
                                  {1, 4}, {1, 4}, {2, 4}, {2, 4}, {3, 4}, {3, 4}};

static int topo_npe(int t) { return t == OAGPU_TET4 ? 4 : t == OAGPU_WED6 ? 6 : t == OAGPU_PYR5 ? 5 : 8; }
static int topo_nscs(int t) { return t == OAGPU_TET4 ? 6 : t == OAGPU_WED6 ? 9 : 12; }
static const int (*topo_lr(int t))[2] {
    return t == OAGPU_TET4 ? kTetLr : t == OAGPU_WED6 ? kWedLr : t == OAGPU_PYR5 ? kPyrLr : kLrscv;
}
static const int kSides[4][6][4] = OAGPU_SIDE_TABLE;
static const int kNumSides[4] = OAGPU_NSIDES;
static int side_npf(int t, int ord) { return kSides[t][ord][3] < 0 ? 3 : 4; }

static uint64_t row_slots(const std::vector<int>& rp, const std::vector<int>& cols, int row,
                          const int* en, int* max_pos, int npe = 8) {
    uint64_t s = ~uint64_t(0);
    const int* b = cols.data() + rp[row];
    const int len = rp[row + 1] - rp[row];
    for (int c = 0; c < npe; ++c) {
        s &= ~(uint64_t(0xFF) << (8 * c));
        const int* it = std::lower_bound(b, b + len, en[c]);
        uint64_t pos = 0xFF;
        if (it != b + len && *it == en[c]) {
            pos = static_cast<uint64_t>(it - b);
            if (static_cast<int>(pos) > *max_pos) *max_pos = static_cast<int>(pos);
        }
        s |= pos << (8 * c);
    }
    return s;
}

void build_mesh(const HostMesh& hm, int reduced, DevMesh& dm) {
    OAGPU_REQUIRE(hm.nn > 0 && hm.ne > 0, "empty mesh");
    OAGPU_REQUIRE(hm.ne < (1 << 28), "element count exceeds the incidence encoding");
    const int nn = hm.nn, ne = hm.ne;
    auto etype = [&](int e) { return hm.etype ? hm.etype[e] : int(OAGPU_HEX8); };
    for (int e = 0; e < ne; ++e) {
        OAGPU_REQUIRE(etype(e) >= OAGPU_HEX8 && etype(e) <= OAGPU_PYR5, "unknown element topology");
        for (int k = 0; k < topo_npe(etype(e)); ++k)
            OAGPU_REQUIRE(hm.conn[8 * size_t(e) + k] >= 0 && hm.conn[8 * size_t(e) + k] < nn,
                          "connectivity out of range");
        dm.mixed = dm.mixed || etype(e) != OAGPU_HEX8;
    }
    for (int f = 0; f < hm.nbf; ++f) {
        OAGPU_REQUIRE(hm.bf_elem[f] >= 0 && hm.bf_elem[f] < ne, "boundary face element out of range");
        OAGPU_REQUIRE(hm.bf_ord[f] >= 0 && hm.bf_ord[f] < kNumSides[etype(hm.bf_elem[f])],
                      "boundary face ordinal out of range for its element's topology");
        OAGPU_REQUIRE(hm.bf_bc[f] >= 0, "negative boundary id");
    }
    dm.nn = nn;
    dm.ne = ne;
    dm.nbf = hm.nbf;
    dm.reduced = reduced;

    // ---- CRS graph: per-row candidate lists, sorted and uniqued (std::set)
    std::vector<int> cnt(size_t(nn) + 1, 0);
    for (int e = 0; e < ne; ++e) {
        const int* en = hm.conn + 8 * size_t(e);
        const int t = etype(e), npe = topo_npe(t);
        const int(*lr)[2] = topo_lr(t);
        if (reduced) {
            for (int ip = 0; ip < topo_nscs(t); ++ip) {
                cnt[en[lr[ip][0]]] += 2;
                cnt[en[lr[ip][1]]] += 2;
            }
        } else {
            for (int a = 0; a < npe; ++a) cnt[en[a]] += npe;
        }
    }
    std::vector<size_t> start(size_t(nn) + 1, 0);
    for (int i = 0; i < nn; ++i) start[i + 1] = start[i] + size_t(cnt[i]);
    std::vector<int> cand(start[nn]);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int e = 0; e < ne; ++e) {
        const int* en = hm.conn + 8 * size_t(e);
        const int t = etype(e), npe = topo_npe(t);
        const int(*lr)[2] = topo_lr(t);
        if (reduced) {
            for (int ip = 0; ip < topo_nscs(t); ++ip) {
                const int L = en[lr[ip][0]], R = en[lr[ip][1]];
                cand[start[L] + cnt[L]++] = L;
                cand[start[L] + cnt[L]++] = R;
                cand[start[R] + cnt[R]++] = L;
                cand[start[R] + cnt[R]++] = R;
            }
        } else {
            for (int a = 0; a < npe; ++a) {
                const int r = en[a];
                for (int c = 0; c < npe; ++c) cand[start[r] + cnt[r]++] = en[c];
            }
        }
    }
    dm.h_row_ptr.assign(size_t(nn) + 1, 0);
    size_t nnz = 0;
    for (int i = 0; i < nn; ++i) {
        int* b = cand.data() + start[i];
        std::sort(b, b + cnt[i]);
        const int u = int(std::unique(b, b + cnt[i]) - b);
        cnt[i] = u;
        nnz += size_t(u);
        OAGPU_REQUIRE(nnz < (size_t(1) << 31), "graph too large for int32 offsets");
        dm.h_row_ptr[i + 1] = int(nnz);
        dm.max_row = std::max(dm.max_row, u);
    }
    OAGPU_REQUIRE(dm.max_row <= 255, "row longer than the 8-bit scatter map");
    dm.nnz = int(nnz);
    dm.h_cols.resize(nnz);
    dm.h_diag_off.resize(nn);
    for (int i = 0; i < nn; ++i) {
        std::copy_n(cand.data() + start[i], cnt[i], dm.h_cols.data() + dm.h_row_ptr[i]);
        const int* b = dm.h_cols.data() + dm.h_row_ptr[i];
        const int* it = std::lower_bound(b, b + cnt[i], i);
        OAGPU_REQUIRE(it != b + cnt[i] && *it == i, "row without a diagonal");
        dm.h_diag_off[i] = int(it - b);
    }
    std::vector<int>().swap(cand);

    // ---- sliced-ELL layout of the values (64-row slices)
    {
        SellHost sh = make_sell(nn, dm.h_row_ptr, dm.h_cols);
        dm.nslice = sh.nslice;
        dm.nsell = sh.nsell;
        dm.h_slice_ptr = std::move(sh.slice_ptr);
        dm.h_slice_w = std::move(sh.slice_w);
        dm.sell_cols.upload(sh.cols);
        dm.slice_ptr.upload(dm.h_slice_ptr);
        dm.slice_w.upload(dm.h_slice_w);
    }

    // ---- Morton traversal orders of the slices and element groups
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], hm.xyz[3 * size_t(i) + d]);
                hi[d] = std::max(hi[d], hm.xyz[3 * size_t(i) + d]);
            }
        auto morton = [&](const double x[3]) {
            uint64_t code = 0;
            uint32_t q[3];
            for (int d = 0; d < 3; ++d) {
                const double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
                const double t = std::min(std::max((x[d] - lo[d]) / span, 0.0), 1.0);
                q[d] = uint32_t(t * 1023.0 + 0.5);
            }
            for (int b = 9; b >= 0; --b)
                for (int d = 0; d < 3; ++d) code = (code << 1) | ((q[d] >> b) & 1u);
            return code;
        };
        auto order_groups = [&](int nitems, int nper, const std::function<void(int, double[3])>& pos) {
            const int ng = (nitems + 63) / 64;
            std::vector<std::pair<uint64_t, int>> key(ng);
            for (int g = 0; g < ng; ++g) {
                double c[3] = {0, 0, 0};
                const int b = 64 * g, e = std::min(nitems, b + 64);
                for (int i = b; i < e; ++i) {
                    double x[3];
                    pos(i, x);
                    for (int d = 0; d < 3; ++d) c[d] += x[d];
                }
                for (int d = 0; d < 3; ++d) c[d] /= double(e - b);
                key[g] = {morton(c), g};
            }
            std::stable_sort(key.begin(), key.end());
            std::vector<int> ord(ng);
            for (int g = 0; g < ng; ++g) ord[g] = key[g].second;
            (void)nper;
            return ord;
        };
        // a mesh in locality order is traversed in its own order
        auto identity = [](int nitems) {
            std::vector<int> ord((nitems + 63) / 64);
            std::iota(ord.begin(), ord.end(), 0);
            return ord;
        };
        const std::vector<int> so = hm.ordered ? identity(nn) : order_groups(nn, 1, [&](int i, double x[3]) {
            for (int d = 0; d < 3; ++d) x[d] = hm.xyz[3 * size_t(i) + d];
        });
        OAGPU_REQUIRE(int(so.size()) == dm.nslice, "slice order size");
        dm.slice_order.upload(so);
        const std::vector<int> eo = hm.ordered ? identity(ne) : order_groups(ne, 8, [&](int e, double x[3]) {
            for (int d = 0; d < 3; ++d) x[d] = 0.0;
            const int npe = topo_npe(etype(e));
            for (int k = 0; k < npe; ++k)
                for (int d = 0; d < 3; ++d)
                    x[d] += hm.xyz[3 * size_t(hm.conn[8 * size_t(e) + k]) + d] / npe;
        });
        dm.negroup = int(eo.size());
        dm.egroup_order.upload(eo);
    }

    // ---- node -> element incidences in ascending element order
    dm.h_inc_ptr.assign(size_t(nn) + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < topo_npe(etype(e)); ++a) dm.h_inc_ptr[hm.conn[8 * size_t(e) + a] + 1]++;
    for (int i = 0; i < nn; ++i) dm.h_inc_ptr[i + 1] += dm.h_inc_ptr[i];
    dm.ninc = dm.h_inc_ptr[nn];
    dm.h_inc.resize(dm.ninc);
    dm.h_inc_slots.resize(dm.ninc);
    {
        std::vector<int> fill(dm.h_inc_ptr.begin(), dm.h_inc_ptr.end() - 1);
        int max_pos = 0;
        for (int e = 0; e < ne; ++e) {
            const int* en = hm.conn + 8 * size_t(e);
            const int npe = topo_npe(etype(e));
            for (int a = 0; a < npe; ++a) {
                const int r = en[a];
                const int k = fill[r]++;
                dm.h_inc[k] = (e << 3) | a;
                dm.h_inc_slots[k] = row_slots(dm.h_row_ptr, dm.h_cols, r, en, &max_pos, npe);
            }
        }
    }

    // ---- node -> wall-face incidences: the face ip nearest to the node
    std::vector<int> isb(nn, 0);
    dm.h_binc_ptr.assign(size_t(nn) + 1, 0);
    for (int f = 0; f < hm.nbf; ++f) {
        const int* en = hm.conn + 8 * size_t(hm.bf_elem[f]);
        const int t = etype(hm.bf_elem[f]), ord = hm.bf_ord[f];
        for (int k = 0; k < side_npf(t, ord); ++k) {
            const int nd = en[kSides[t][ord][k]];
            dm.h_binc_ptr[nd + 1]++;
            isb[nd] = 1;
        }
    }
    for (int i = 0; i < nn; ++i) dm.h_binc_ptr[i + 1] += dm.h_binc_ptr[i];
    dm.nbinc = dm.h_binc_ptr[nn];
    dm.h_binc.resize(dm.nbinc);
    dm.h_binc_slots.resize(dm.nbinc);
    {
        std::vector<int> fill(dm.h_binc_ptr.begin(), dm.h_binc_ptr.end() - 1);
        int max_pos = 0;
        for (int f = 0; f < hm.nbf; ++f) {
            const int* en = hm.conn + 8 * size_t(hm.bf_elem[f]);
            const int t = etype(hm.bf_elem[f]), ord = hm.bf_ord[f];
            for (int k = 0; k < side_npf(t, ord); ++k) {
                const int nd = en[kSides[t][ord][k]];
                const int q = fill[nd]++;
                dm.h_binc[q] = (f << 2) | k;
                dm.h_binc_slots[q] = row_slots(dm.h_row_ptr, dm.h_cols, nd, en, &max_pos, topo_npe(t));
            }
        }
    }

    // ---- upload
    dm.xyz.alloc(size_t(nn) * 3);
    dm.xyz.upload(hm.xyz, size_t(nn) * 3);
    dm.conn.alloc(size_t(ne) * 8);
    dm.conn.upload(hm.conn, size_t(ne) * 8);
    if (dm.mixed) {
        dm.h_etype.assign(hm.etype, hm.etype + ne);
        dm.etype.upload(dm.h_etype);
        std::vector<int> gl;
        for (int e = 0; e < ne; ++e)
            if (etype(e) != OAGPU_HEX8) gl.push_back(e);
        dm.nglist = int(gl.size());
        dm.glist.upload(gl);
        // static ip geometry: per ip 3 + 6 npe doubles (area, dN/dx standard and shifted)
        std::vector<long long> off(size_t(ne), -1);
        long long tot = 0;
        for (int e : gl) {
            off[e] = tot;
            const int t = etype(e), npe = topo_npe(t);
            tot += (long long)topo_nscs(t) * (3 + 6 * npe);
        }
        dm.geo_off.upload(off);
        dm.geo.alloc(size_t(tot));
        // face-ip geometry of the non-hex elements' boundary sides
        std::vector<long long> soff(size_t(hm.nbf), -1);
        std::vector<int> sl;
        for (int f = 0; f < hm.nbf; ++f)
            if (etype(hm.bf_elem[f]) != OAGPU_HEX8) {
                soff[f] = (long long)sl.size() * kSideRecord;
                sl.push_back(f);
            }
        dm.nslist = int(sl.size());
        dm.slist.upload(sl);
        dm.sgeo_off.upload(soff);
        dm.sgeo.alloc(std::max<size_t>(sl.size(), 1) * kSideRecord);
    }
    dm.row_ptr.upload(dm.h_row_ptr);
    dm.cols.upload(dm.h_cols);
    dm.diag_off.upload(dm.h_diag_off);
    dm.inc_ptr.upload(dm.h_inc_ptr);
    dm.inc.upload(dm.h_inc);
    {
        // local-node major copy of the incidences for the row kernels
        std::vector<int> inc_la(size_t(8) * nn, -1);
        std::vector<uint64_t> slots_la(size_t(8) * nn, ~uint64_t(0));
        std::vector<int> ext_ptr(size_t(nn) + 1, 0), ext, gen_ptr(size_t(nn) + 1, 0), gen;
        std::vector<uint64_t> ext_slots, gen_slots;
        for (int r = 0; r < nn; ++r) {
            for (int k = dm.h_inc_ptr[r]; k < dm.h_inc_ptr[r + 1]; ++k) {
                const int la = dm.h_inc[k] & 7;
                if (etype(dm.h_inc[k] >> 3) != OAGPU_HEX8) {
                    gen.push_back(dm.h_inc[k]);
                    gen_slots.push_back(dm.h_inc_slots[k]);
                    continue;
                }
                const size_t o = size_t(la) * nn + r;
                if (inc_la[o] < 0) {
                    inc_la[o] = dm.h_inc[k] >> 3;
                    slots_la[o] = dm.h_inc_slots[k];
                } else {
                    ext.push_back(dm.h_inc[k]);
                    ext_slots.push_back(dm.h_inc_slots[k]);
                }
            }
            ext_ptr[r + 1] = int(ext.size());
            gen_ptr[r + 1] = int(gen.size());
        }
        dm.next = int(ext.size());
        dm.ngen = int(gen.size());
        dm.gen_ptr.upload(gen_ptr);
        if (gen.empty()) {
            gen.push_back(0);
            gen_slots.push_back(0);
        }
        dm.gen.upload(gen);
        dm.gen_slots.upload(gen_slots);
        dm.inc_la.upload(inc_la);
        dm.slots_la.upload(slots_la);
        dm.ext_ptr.upload(ext_ptr);
        if (ext.empty()) {
            ext.push_back(0);
            ext_slots.push_back(0);
        }
        dm.ext.upload(ext);
        dm.ext_slots.upload(ext_slots);
    }
    dm.binc_ptr.upload(dm.h_binc_ptr);
    if (dm.nbinc) {
        dm.binc.upload(dm.h_binc);
        dm.binc_slots.upload(dm.h_binc_slots);
    } else {
        dm.binc.alloc(1);
        dm.binc_slots.alloc(1);
    }
    std::vector<int> tmp;
    tmp.assign(hm.bf_elem, hm.bf_elem + hm.nbf);
    if (tmp.empty()) tmp.push_back(0);
    dm.bf_elem.upload(tmp);
    tmp.assign(hm.bf_ord, hm.bf_ord + hm.nbf);
    if (tmp.empty()) tmp.push_back(0);
    dm.bf_ord.upload(tmp);
    tmp.assign(hm.bf_bc, hm.bf_bc + hm.nbf);
    if (tmp.empty()) tmp.push_back(0);
    dm.bf_bc.upload(tmp);
    dm.bnode.upload(isb);
    dm.vol.alloc(nn);
    dm.scvw.alloc(size_t(ne) * 8);
    dm.err.alloc(1);
    dm.err.zero();
    OAGPU_HIP(hipDeviceSynchronize());
}

void sell_to_crs(const DevMesh& dm, int nb, const double* sell, double* crs) {
    for (int r = 0; r < dm.nn; ++r)
        for (int j = dm.h_row_ptr[r]; j < dm.h_row_ptr[r + 1]; ++j)
            for (int k = 0; k < nb; ++k)
                crs[size_t(j) * nb + k] = sell[sell_index(dm.h_slice_ptr, r, j - dm.h_row_ptr[r], nb, k)];
}

void crs_to_sell(const DevMesh& dm, int nb, const double* crs, double* sell) {
    for (int r = 0; r < dm.nn; ++r)
        for (int j = dm.h_row_ptr[r]; j < dm.h_row_ptr[r + 1]; ++j)
            for (int k = 0; k < nb; ++k)
                sell[sell_index(dm.h_slice_ptr, r, j - dm.h_row_ptr[r], nb, k)] = crs[size_t(j) * nb + k];
}


// Per axis, a node's layer index.  When the mesh's coordinates form (nearly) a
// tensor-product set -- structured grids, graded or not, and blocks of them --
// it is the dense rank of the node's coordinate among the distinct values on
// that axis (values closer than 1e-9 of the extent are one value), i.e. the
// grid index.  Otherwise it is the node's quantile along the axis scaled to
// an isotropic layer count (nn^(1/3) times the axis' share of the bounding
// box).  Nodes are then sorted along the Morton curve of these indices (ties
// by caller id), so 64 consecutive nodes of a structured grid form a 4 x 4 x 4
// brick: a wavefront of a row kernel gathers about 216 distinct nodes instead
// of the ~600 of a 64-node strip of a lexicographic numbering.  Elements are
// sorted by their smallest device node id (ties by caller id).
void locality_order(const HostMesh& hm, std::vector<int>& n_old, std::vector<int>& e_old) {
    const int nn = hm.nn, ne = hm.ne;
    std::vector<uint32_t> key[3];
    double ext[3];
    size_t distinct[3];
    std::vector<int> idx(nn);
    for (int d = 0; d < 3; ++d) {
        std::iota(idx.begin(), idx.end(), 0);
        auto X = [&](int i) { return hm.xyz[3 * size_t(i) + d]; };
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return X(a) < X(b) || (X(a) == X(b) && a < b); });
        ext[d] = X(idx[nn - 1]) - X(idx[0]);
        const double tol = 1e-9 * (ext[d] > 0 ? ext[d] : 1.0);
        key[d].assign(nn, 0);
        uint32_t rank = 0;
        double base = X(idx[0]);
        for (int q = 0; q < nn; ++q) {
            if (X(idx[q]) - base > tol) {
                ++rank;
                base = X(idx[q]);
            }
            key[d][idx[q]] = rank;
        }
        distinct[d] = size_t(rank) + 1;
    }
    const double cells = double(distinct[0]) * double(distinct[1]) * double(distinct[2]);
    if (cells > 4.0 * double(nn)) {
        // not a tensor-product set: quantiles scaled to isotropic layer counts
        double gm = 1.0;
        for (int d = 0; d < 3; ++d) gm *= std::max(ext[d], 1e-300);
        gm = std::cbrt(gm);
        for (int d = 0; d < 3; ++d) {
            std::iota(idx.begin(), idx.end(), 0);
            auto X = [&](int i) { return hm.xyz[3 * size_t(i) + d]; };
            std::sort(idx.begin(), idx.end(), [&](int a, int b) { return X(a) < X(b) || (X(a) == X(b) && a < b); });
            const double layers = std::max(1.0, std::cbrt(double(nn)) * std::max(ext[d], 1e-300) / gm);
            for (int q = 0; q < nn; ++q)
                key[d][idx[q]] = uint32_t(std::min(layers - 1.0, std::floor(double(q) / double(nn) * layers)));
        }
    }
    // at most 21 bits per axis in the 63-bit Morton code
    for (int d = 0; d < 3; ++d) {
        uint32_t mx = 0;
        for (int i = 0; i < nn; ++i) mx = std::max(mx, key[d][i]);
        int shift = 0;
        while ((mx >> shift) >= (1u << 21)) ++shift;
        if (shift)
            for (int i = 0; i < nn; ++i) key[d][i] >>= shift;
    }
    auto spread = [](uint64_t v) {  // bit b of v -> bit 3b
        v &= 0x1fffffull;
        v = (v | v << 32) & 0x1f00000000ffffull;
        v = (v | v << 16) & 0x1f0000ff0000ffull;
        v = (v | v << 8) & 0x100f00f00f00f00full;
        v = (v | v << 4) & 0x10c30c30c30c30c3ull;
        v = (v | v << 2) & 0x1249249249249249ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int>> mk(nn);
    for (int i = 0; i < nn; ++i)
        mk[i] = {spread(key[0][i]) | spread(key[1][i]) << 1 | spread(key[2][i]) << 2, i};
    std::sort(mk.begin(), mk.end());
    n_old.resize(nn);
    std::vector<int> n_new(nn);
    for (int i = 0; i < nn; ++i) {
        n_old[i] = mk[i].second;
        n_new[mk[i].second] = i;
    }
    std::vector<std::pair<int, int>> ek(ne);
    for (int e = 0; e < ne; ++e) {
        const int t = hm.etype ? hm.etype[e] : int(OAGPU_HEX8);
        int lo = nn;
        for (int k = 0; k < topo_npe(t); ++k) lo = std::min(lo, n_new[hm.conn[8 * size_t(e) + k]]);
        ek[e] = {lo, e};
    }
    std::sort(ek.begin(), ek.end());
    e_old.resize(ne);
    for (int e = 0; e < ne; ++e) e_old[e] = ek[e].second;
}

void build_mesh_ordered(const HostMesh& hm, int reduced, bool renumber, DevMesh& dm) {
    if (!renumber) {
        build_mesh(hm, reduced, dm);
        return;
    }
    OAGPU_REQUIRE(hm.nn > 0 && hm.ne > 0, "empty mesh");
    const int nn = hm.nn, ne = hm.ne;
    for (size_t i = 0; i < size_t(ne) * 8; ++i) {
        const int t = hm.etype ? hm.etype[i / 8] : int(OAGPU_HEX8);
        if (int(i % 8) < topo_npe(t))
            OAGPU_REQUIRE(hm.conn[i] >= 0 && hm.conn[i] < nn, "connectivity out of range");
    }
    for (int f = 0; f < hm.nbf; ++f)
        OAGPU_REQUIRE(hm.bf_elem[f] >= 0 && hm.bf_elem[f] < ne, "boundary face element out of range");
    std::vector<int> n_old, e_old;
    locality_order(hm, n_old, e_old);
    std::vector<int> n_new(nn), e_new(ne);
    for (int i = 0; i < nn; ++i) n_new[n_old[i]] = i;
    for (int e = 0; e < ne; ++e) e_new[e_old[e]] = e;
    std::vector<double> xyz(size_t(nn) * 3);
    for (int i = 0; i < nn; ++i)
        for (int d = 0; d < 3; ++d) xyz[3 * size_t(i) + d] = hm.xyz[3 * size_t(n_old[i]) + d];
    std::vector<int> conn(size_t(ne) * 8, 0), et;
    for (int e = 0; e < ne; ++e) {
        const int o = e_old[e];
        const int t = hm.etype ? hm.etype[o] : int(OAGPU_HEX8);
        for (int k = 0; k < 8; ++k)
            conn[8 * size_t(e) + k] = k < topo_npe(t) ? n_new[hm.conn[8 * size_t(o) + k]] : 0;
    }
    if (hm.etype) {
        et.resize(ne);
        for (int e = 0; e < ne; ++e) et[e] = hm.etype[e_old[e]];
    }
    std::vector<int> bfe(size_t(hm.nbf));
    for (int f = 0; f < hm.nbf; ++f) bfe[f] = e_new[hm.bf_elem[f]];
    HostMesh pm = hm;
    pm.xyz = xyz.data();
    pm.conn = conn.data();
    pm.etype = hm.etype ? et.data() : nullptr;
    pm.bf_elem = hm.nbf ? bfe.data() : nullptr;
    pm.ordered = true;
    build_mesh(pm, reduced, dm);
    dm.n_old = std::move(n_old);
    dm.n_new = std::move(n_new);
    dm.e_old = std::move(e_old);
    dm.e_new = std::move(e_new);
}

void caller_graph(const DevMesh& dm, std::vector<int>& row_ptr, std::vector<int>& cols,
                  std::vector<int>& diag_off) {
    if (!dm.renumbered()) {
        row_ptr = dm.h_row_ptr;
        cols = dm.h_cols;
        diag_off = dm.h_diag_off;
        return;
    }
    const int nn = dm.nn;
    row_ptr.assign(size_t(nn) + 1, 0);
    cols.resize(dm.h_cols.size());
    diag_off.resize(nn);
    for (int r = 0; r < nn; ++r) {
        const int p = dm.n_new[r];
        row_ptr[r + 1] = row_ptr[r] + (dm.h_row_ptr[p + 1] - dm.h_row_ptr[p]);
    }
    for (int r = 0; r < nn; ++r) {
        const int p = dm.n_new[r];
        int* o = cols.data() + row_ptr[r];
        const int len = row_ptr[r + 1] - row_ptr[r];
        for (int j = 0; j < len; ++j) o[j] = dm.n_old[dm.h_cols[dm.h_row_ptr[p] + j]];
        std::sort(o, o + len);
        diag_off[r] = int(std::lower_bound(o, o + len, r) - o);
    }
}

// For caller row r and its j-th sorted column: the device slot of that column
// in device row n_new[r].
template <typename F>
static void for_caller_entries(const DevMesh& dm, F&& f) {
    const int nn = dm.nn;
    std::vector<std::pair<int, int>> row;  // (caller column, device slot)
    size_t j0 = 0;
    for (int r = 0; r < nn; ++r) {
        const int p = dm.renumbered() ? dm.n_new[r] : r;
        const int b = dm.h_row_ptr[p], len = dm.h_row_ptr[p + 1] - b;
        row.resize(len);
        for (int s = 0; s < len; ++s)
            row[s] = {dm.renumbered() ? dm.n_old[dm.h_cols[b + s]] : dm.h_cols[b + s], s};
        std::sort(row.begin(), row.end());
        for (int j = 0; j < len; ++j) f(j0 + j, p, row[j].second);
        j0 += size_t(len);
    }
}

void sell_to_caller_crs(const DevMesh& dm, int nb, const double* sell, double* crs) {
    for_caller_entries(dm, [&](size_t j, int p, int s) {
        for (int k = 0; k < nb; ++k) crs[j * nb + k] = sell[sell_index(dm.h_slice_ptr, p, s, nb, k)];
    });
}

void caller_crs_to_sell(const DevMesh& dm, int nb, const double* crs, double* sell) {
    for_caller_entries(dm, [&](size_t j, int p, int s) {
        for (int k = 0; k < nb; ++k) sell[sell_index(dm.h_slice_ptr, p, s, nb, k)] = crs[j * nb + k];
    });
}

}  // namespace oagpu
