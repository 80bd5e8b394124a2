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
        const std::vector<int> so = order_groups(nn, 1, [&](int i, double x[3]) {
            for (int d = 0; d < 3; ++d) x[d] = hm.xyz[3 * size_t(i) + d];
        });
        OAGPU_REQUIRE(int(so.size()) == dm.nslice, "slice order size");
        dm.slice_order.upload(so);
        const std::vector<int> eo = order_groups(ne, 8, [&](int e, double x[3]) {
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

}  // namespace oagpu
