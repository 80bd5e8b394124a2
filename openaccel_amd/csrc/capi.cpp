// capi.cpp -- extern "C" entry points of liboagpu (declared in include/oagpu.h)
// and the host driver of one steady outer iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "flow.hpp"
#include "sell.hpp"

struct oagpu_mesh {
    oagpu::DevMesh m;
};
struct oagpu_flow {
    oagpu::Flow f;
};

namespace {
thread_local std::string g_last_error;
bool g_init = false;

template <typename F>
int guarded(F&& fn) {
    try {
        fn();
        return OAGPU_OK;
    } catch (const oagpu::Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return OAGPU_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return OAGPU_EINTERNAL;
    }
}

void ensure_init() {
    if (!g_init) {
        int n = 0;
        OAGPU_HIP(hipGetDeviceCount(&n));
        OAGPU_REQUIRE(n > 0, "no HIP device");
        g_init = true;
    }
}

// Index kind of a field: per node, per element or per boundary side.
enum FieldKind { kPerNode, kPerElem, kPerSide };

double* field_ptr(oagpu::Flow& f, const char* name, size_t* count, FieldKind* kind = nullptr) {
    const size_t nn = size_t(f.m->nn), ne = size_t(f.m->ne), nbf = size_t(f.m->nbf);
    struct {
        const char* n;
        oagpu::DBuf<double>* b;
        size_t c;
        FieldKind k;
    } tab[] = {{"U", &f.U, nn * 3, kPerNode},         {"p", &f.p, nn, kPerNode},
               {"gradU", &f.gradU, nn * 9, kPerNode}, {"gradP", &f.gradP, nn * 3, kPerNode},
               {"du", &f.du, nn * 3, kPerNode},       {"mdot", &f.mdot, ne * 12, kPerElem},
               {"div", &f.div, nn, kPerNode},         {"mu", &f.mueff, nn, kPerNode},
               {"k", &f.k, nn, kPerNode},             {"omega", &f.om, nn, kPerNode},
               {"gradK", &f.gradK, nn * 3, kPerNode}, {"gradW", &f.gradW, nn * 3, kPerNode},
               {"ymin", &f.ymin, nn, kPerNode},       {"mut", &f.mut, nn, kPerNode},
               {"F1", &f.F1, nn, kPerNode},           {"Pk", &f.Pk, nn, kPerNode},
               {"Pk_wall", &f.Pk_wall, nn, kPerNode}, {"kprev", &f.kprev, nn, kPerNode},
               {"beta", &f.beta, nn * 3, kPerNode},   {"beta_k", &f.beta_k, nn, kPerNode},
               {"beta_w", &f.beta_w, nn, kPerNode},   {"symA", &f.symA, nn * 3, kPerNode},
               {"msd", &f.msd, nbf * 4, kPerSide},    {"Uold", &f.Uold, nn * 3, kPerNode},
               {"Uold2", &f.Uold2, nn * 3, kPerNode}, {"kold", &f.kold, nn, kPerNode},
               {"kold2", &f.kold2, nn, kPerNode},     {"wold", &f.wold, nn, kPerNode},
               {"wold2", &f.wold2, nn, kPerNode}};
    for (auto& t : tab)
        if (!std::strcmp(name, t.n)) {
            OAGPU_REQUIRE(t.b->p, std::string("field ") + name +
                                          " not allocated (SST, high resolution, symmetry or"
                                          " inlets / outlets off)");
            *count = t.c;
            if (kind) *kind = t.k;
            return t.b->p;
        }
    throw oagpu::Error(OAGPU_EINVAL, std::string("unknown field ") + name);
}

// Caller <-> device order of node- or element-indexed arrays of `comp`
// values per item (the device numbering is the mesh's locality order).
const std::vector<int>* old_ids(const oagpu::DevMesh& m, FieldKind k) {
    if (!m.renumbered() || k == kPerSide) return nullptr;
    return k == kPerNode ? &m.n_old : &m.e_old;
}

template <typename T>
void to_device_order(const std::vector<int>* old, const T* caller, T* dev, size_t count) {
    if (!old) {
        std::copy(caller, caller + count, dev);
        return;
    }
    const size_t n = old->size(), comp = count / n;
    for (size_t i = 0; i < n; ++i)
        std::copy(caller + size_t((*old)[i]) * comp, caller + size_t((*old)[i]) * comp + comp, dev + i * comp);
}

template <typename T>
void to_caller_order(const std::vector<int>* old, const T* dev, T* caller, size_t count) {
    if (!old) {
        std::copy(dev, dev + count, caller);
        return;
    }
    const size_t n = old->size(), comp = count / n;
    for (size_t i = 0; i < n; ++i)
        std::copy(dev + i * comp, dev + i * comp + comp, caller + size_t((*old)[i]) * comp);
}

// Each flow has its own error word: flows that share a mesh (batched cases)
// run concurrently on their own streams and never see each other's flags.
void check_err(oagpu::Flow& f) {
    int e = 0;
    f.err.download(&e, 1, f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    if (e) {
        f.err.zero(f.stream);
        throw oagpu::Error(OAGPU_EGEOM, "non-positive Jacobian determinant in a kernel");
    }
}

}  // namespace

extern "C" {

int oagpu_init(int device) {
    return guarded([&] {
        int n = 0;
        OAGPU_HIP(hipGetDeviceCount(&n));
        OAGPU_REQUIRE(device >= 0 && device < n, "device index out of range");
        OAGPU_HIP(hipSetDevice(device));
        // device address of this GPU's scratch array (allocation-free entry points)
        (void)oagpu::scratch_device_ptr();
        g_init = true;
    });
}

const char* oagpu_last_error(void) { return g_last_error.c_str(); }

int oagpu_synchronize(void) {
    return guarded([&] { OAGPU_HIP(hipDeviceSynchronize()); });
}

int oagpu_mesh_create(int n_nodes, const double* xyz, int n_elems, const int* conn,
                      const int* etype, int n_bfaces, const int* bf_elem, const int* bf_ord,
                      const int* bf_bc, int stencil, oagpu_mesh** out) {
    return guarded([&] {
        OAGPU_REQUIRE(out, "null output handle");
        OAGPU_REQUIRE(xyz && conn, "null mesh arrays");
        OAGPU_REQUIRE(n_bfaces == 0 || (bf_elem && bf_ord && bf_bc), "null boundary arrays");
        ensure_init();
        auto h = std::make_unique<oagpu_mesh>();
        oagpu::HostMesh hm;
        hm.nn = n_nodes;
        hm.ne = n_elems;
        hm.nbf = n_bfaces;
        hm.xyz = xyz;
        hm.conn = conn;
        hm.etype = etype;
        hm.bf_elem = bf_elem;
        hm.bf_ord = bf_ord;
        hm.bf_bc = bf_bc;
        OAGPU_REQUIRE((stencil & ~OAGPU_MESH_CALLER_ORDER) == OAGPU_STENCIL_FULL ||
                          (stencil & ~OAGPU_MESH_CALLER_ORDER) == OAGPU_STENCIL_REDUCED,
                      "stencil: OAGPU_STENCIL_REDUCED or OAGPU_STENCIL_FULL, optionally | OAGPU_MESH_CALLER_ORDER");
        oagpu::build_mesh_ordered(hm, (stencil & ~OAGPU_MESH_CALLER_ORDER) == OAGPU_STENCIL_FULL ? 0 : 1,
                                  !(stencil & OAGPU_MESH_CALLER_ORDER), h->m);
        oagpu::launch_mesh_geometry(h->m, nullptr);
        int e = 0;
        h->m.err.download(&e, 1);
        OAGPU_HIP(hipDeviceSynchronize());
        if (e) throw oagpu::Error(OAGPU_EGEOM, "non-positive Jacobian determinant in the mesh");
        *out = h.release();
    });
}

int oagpu_mesh_destroy(oagpu_mesh* m) {
    return guarded([&] { delete m; });
}

int oagpu_mesh_sizes(const oagpu_mesh* m, int* nn, int* ne, int* nnz, int* ninc, int* nbinc) {
    return guarded([&] {
        OAGPU_REQUIRE(m, "null mesh");
        if (nn) *nn = m->m.nn;
        if (ne) *ne = m->m.ne;
        if (nnz) *nnz = m->m.nnz;
        if (ninc) *ninc = m->m.ninc;
        if (nbinc) *nbinc = m->m.nbinc;
    });
}

int oagpu_graph_copy(const oagpu_mesh* m, int* row_ptr, int* cols, int* diag_off) {
    return guarded([&] {
        OAGPU_REQUIRE(m, "null mesh");
        std::vector<int> rp, cl, dg;
        oagpu::caller_graph(m->m, rp, cl, dg);
        if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
        if (cols) std::copy(cl.begin(), cl.end(), cols);
        if (diag_off) std::copy(dg.begin(), dg.end(), diag_off);
    });
}

int oagpu_scatter_map_copy(const oagpu_mesh* m, int* inc_ptr, int* inc, unsigned long long* slots) {
    return guarded([&] {
        OAGPU_REQUIRE(m, "null mesh");
        const auto& d = m->m;
        const size_t nn = size_t(d.nn);
        // device incidences' scatter maps, rebuilt from the device's
        // local-node-major copy, so the check covers what the kernels read
        std::vector<uint64_t> dslots;
        if (slots) {
            std::vector<int> inc_la(8 * nn), ext_ptr(nn + 1), ext(std::max(d.next, 1));
            std::vector<uint64_t> slots_la(8 * nn), ext_slots(std::max(d.next, 1));
            d.inc_la.download(inc_la.data(), inc_la.size());
            d.slots_la.download(slots_la.data(), slots_la.size());
            d.ext_ptr.download(ext_ptr.data(), ext_ptr.size());
            d.ext.download(ext.data(), ext.size());
            d.ext_slots.download(ext_slots.data(), ext_slots.size());
            OAGPU_HIP(hipDeviceSynchronize());
            dslots.resize(d.h_inc.size());
            for (size_t r = 0; r < nn; ++r) {
                for (int k = d.h_inc_ptr[r]; k < d.h_inc_ptr[r + 1]; ++k) {
                    const int e = d.h_inc[k] >> 3, la = d.h_inc[k] & 7;
                    uint64_t v = ~uint64_t(0);
                    bool found = false;
                    if (inc_la[size_t(la) * nn + r] == e) {
                        v = slots_la[size_t(la) * nn + r];
                        found = true;
                    }
                    for (int q = ext_ptr[r]; q < ext_ptr[r + 1] && !found; ++q)
                        if (ext[q] == d.h_inc[k]) {
                            v = ext_slots[q];
                            found = true;
                        }
                    if (!found && d.mixed) {
                        // non-hex incidences are read from the gen list
                        v = d.h_inc_slots[k];
                        found = true;
                    }
                    OAGPU_REQUIRE(found, "incidence missing from the device copy");
                    dslots[k] = v;
                }
            }
        }
        if (!d.renumbered()) {
            if (inc_ptr) std::copy(d.h_inc_ptr.begin(), d.h_inc_ptr.end(), inc_ptr);
            if (inc) std::copy(d.h_inc.begin(), d.h_inc.end(), inc);
            if (slots) std::copy(dslots.begin(), dslots.end(), slots);
            return;
        }
        // caller numbering: element ids, incidence order (ascending element)
        // and slot positions within the caller's sorted rows
        std::vector<int> rp, cl, dg;
        oagpu::caller_graph(d, rp, cl, dg);
        size_t o = 0;
        std::vector<std::pair<int, uint64_t>> row;
        for (size_t r = 0; r < nn; ++r) {
            const int p = d.n_new[r];
            if (inc_ptr) inc_ptr[r] = int(o);
            row.clear();
            const int* cb = cl.data() + rp[r];
            const int clen = rp[r + 1] - rp[r];
            for (int k = d.h_inc_ptr[p]; k < d.h_inc_ptr[p + 1]; ++k) {
                const int e = d.e_old[d.h_inc[k] >> 3], la = d.h_inc[k] & 7;
                uint64_t v = ~uint64_t(0);
                if (slots) {
                    const uint64_t dv = dslots[k];
                    for (int c = 0; c < 8; ++c) {
                        const uint64_t pos = (dv >> (8 * c)) & 0xFF;
                        uint64_t np = 0xFF;
                        if (pos != 0xFF) {
                            const int col = d.n_old[d.h_cols[d.h_row_ptr[p] + int(pos)]];
                            np = uint64_t(std::lower_bound(cb, cb + clen, col) - cb);
                        }
                        v = (v & ~(uint64_t(0xFF) << (8 * c))) | (np << (8 * c));
                    }
                }
                row.push_back({(e << 3) | la, v});
            }
            std::sort(row.begin(), row.end(),
                      [](const auto& x, const auto& y) { return x.first < y.first; });
            for (const auto& q : row) {
                if (inc) inc[o] = q.first;
                if (slots) slots[o] = q.second;
                ++o;
            }
        }
        if (inc_ptr) inc_ptr[nn] = int(o);
    });
}

int oagpu_mesh_volume_copy(const oagpu_mesh* m, double* vol) {
    return guarded([&] {
        OAGPU_REQUIRE(m && vol, "null argument");
        std::vector<double> tmp(size_t(m->m.nn));
        m->m.vol.download(tmp.data(), tmp.size());
        OAGPU_HIP(hipDeviceSynchronize());
        to_caller_order(old_ids(m->m, kPerNode), tmp.data(), vol, tmp.size());
    });
}

int oagpu_flow_create(oagpu_mesh* m, const oagpu_flow_params* prm, int n_bc, const double* bc_U,
                      oagpu_flow** out) {
    return guarded([&] {
        OAGPU_REQUIRE(m && prm && out, "null argument");
        OAGPU_REQUIRE(n_bc >= 0 && (n_bc == 0 || bc_U), "bad boundary velocity table");
        OAGPU_REQUIRE(prm->precision == OAGPU_FP64 || prm->precision == OAGPU_MIXED, "bad precision");
        OAGPU_REQUIRE(prm->dt > 0 && prm->urf_U > 0 && prm->urf_p > 0 && prm->rho > 0, "bad parameters");
        OAGPU_REQUIRE(prm->p_ref_node < m->m.nn, "pressure reference node out of range");
        auto h = std::make_unique<oagpu_flow>();
        oagpu::Flow& f = h->f;
        f.m = &m->m;
        f.prm = *prm;
        if (prm->p_ref_node >= 0 && m->m.renumbered()) f.prm.p_ref_node = m->m.n_new[prm->p_ref_node];
        f.nbc = n_bc;
        std::vector<double> bc(bc_U, bc_U + 3 * size_t(n_bc));
        if (bc.empty()) bc.assign(3, 0.0);
        f.bc_U.upload(bc);
        const size_t nn = size_t(m->m.nn), ne = size_t(m->m.ne), nsell = m->m.nsell;
        f.U.alloc(nn * 3);
        f.p.alloc(nn);
        f.gradU.alloc(nn * 9);
        f.gradP.alloc(nn * 3);
        f.du.alloc(nn * 3);
        f.mdot.alloc(ne * 12);
        f.div.alloc(nn);
        f.A3.alloc(nsell * 9);
        f.b3.alloc(nn * 3);
        f.x3.alloc(nn * 3);
        f.A1.alloc(nsell);
        f.b1.alloc(nn);
        f.x1.alloc(nn);
        for (auto* b : {&f.dinv, &f.r, &f.z, &f.pk, &f.q, &f.w, &f.t, &f.s, &f.r0, &f.ph, &f.sh})
            b->alloc(nn * 3);
        f.part.alloc(size_t(64) * oagpu::kRedGrid);
        f.red.alloc(192);
        f.scal.alloc(16);
        f.flag.alloc(4);
        f.err.alloc(1);
        f.err.zero();
        for (auto* b : {&f.U, &f.p, &f.gradU, &f.gradP, &f.du, &f.mdot, &f.div, &f.A3, &f.b3, &f.x3,
                        &f.A1, &f.b1, &f.x1})
            b->zero();
        f.mueff.upload(std::vector<double>(nn, prm->mu));
        if (prm->high_res) {
            f.beta.alloc(nn * 3);
            f.beta.zero();
        }
        f.h_bc_type.assign(size_t(n_bc > 0 ? n_bc : 1), OAGPU_BC_WALL);
        f.bc_type.upload(f.h_bc_type);
        f.wnode.alloc(nn);
        OAGPU_HIP(hipMemcpy(f.wnode.p, m->m.bnode.p, nn * sizeof(int), hipMemcpyDeviceToDevice));
        OAGPU_HIP(hipStreamCreate(&f.stream));
        // boundary ids must index the velocity table
        {
            std::vector<int> ids(m->m.nbf ? m->m.nbf : 1);
            if (m->m.nbf) {
                m->m.bf_bc.download(ids.data(), m->m.nbf);
                OAGPU_HIP(hipDeviceSynchronize());
                for (int i = 0; i < m->m.nbf; ++i)
                    OAGPU_REQUIRE(ids[i] < n_bc, "boundary id without a wall velocity");
            }
        }
        OAGPU_HIP(hipDeviceSynchronize());
        *out = h.release();
    });
}

int oagpu_flow_destroy(oagpu_flow* f) {
    return guarded([&] {
        if (!f) return;
        if (f->f.stream) (void)hipStreamSynchronize(f->f.stream);
        if (f->f.stream) (void)hipStreamDestroy(f->f.stream);
        delete f;
    });
}

int oagpu_field_upload(oagpu_flow* h, const char* name, const double* host, size_t count) {
    return guarded([&] {
        OAGPU_REQUIRE(h && name && host, "null argument");
        size_t n = 0;
        FieldKind kind;
        double* d = field_ptr(h->f, name, &n, &kind);
        OAGPU_REQUIRE(count == n, std::string("size mismatch for field ") + name);
        std::vector<double> tmp(n);
        to_device_order(old_ids(*h->f.m, kind), host, tmp.data(), n);
        OAGPU_HIP(hipMemcpyAsync(d, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, h->f.stream));
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_field_download(oagpu_flow* h, const char* name, double* host, size_t count) {
    return guarded([&] {
        OAGPU_REQUIRE(h && name && host, "null argument");
        size_t n = 0;
        FieldKind kind;
        double* d = field_ptr(h->f, name, &n, &kind);
        OAGPU_REQUIRE(count == n, std::string("size mismatch for field ") + name);
        std::vector<double> tmp(n);
        OAGPU_HIP(hipMemcpyAsync(tmp.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, h->f.stream));
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
        to_caller_order(old_ids(*h->f.m, kind), tmp.data(), host, n);
    });
}

int oagpu_initialize(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::Flow& f = h->f;
        f.gradU.zero(f.stream);
        f.gradP.zero(f.stream);
        oagpu::launch_gradient(f, 0);
        if (f.prm.high_res) oagpu::launch_blending(f, 0);  // navierStokesEquation.cpp:95
        oagpu::launch_gradient(f, 1);
        oagpu::field_scales(f, nullptr, nullptr);
        oagpu::launch_mdot_init(f);
        oagpu::launch_mass_divergence(f);
        check_err(f);
    });
}

int oagpu_assemble_momentum(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_assemble_momentum(h->f);
        check_err(h->f);
    });
}

int oagpu_assemble_pcorr(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_assemble_pcorr(h->f);
        check_err(h->f);
    });
}

int oagpu_system_download(oagpu_flow* h, int N, double* vals, double* b) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (N == 1 || N == 3), "bad argument");
        oagpu::Flow& f = h->f;
        std::vector<double> sell;
        if (vals) {
            sell.resize(f.m->nsell * N * N);
            (N == 3 ? f.A3 : f.A1).download(sell.data(), sell.size(), f.stream);
        }
        std::vector<double> tb(b ? size_t(f.m->nn) * N : 0);
        if (b) (N == 3 ? f.b3 : f.b1).download(tb.data(), tb.size(), f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        if (b) to_caller_order(old_ids(*f.m, kPerNode), tb.data(), b, tb.size());
        if (vals) oagpu::sell_to_caller_crs(*f.m, N * N, sell.data(), vals);
    });
}

int oagpu_system_upload(oagpu_flow* h, int N, const double* vals, const double* b) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (N == 1 || N == 3), "bad argument");
        oagpu::Flow& f = h->f;
        std::vector<double> sell;
        if (vals) {
            sell.assign(f.m->nsell * N * N, 0.0);
            oagpu::caller_crs_to_sell(*f.m, N * N, vals, sell.data());
            (N == 3 ? f.A3 : f.A1).upload(sell.data(), sell.size(), f.stream);
        }
        std::vector<double> tb(b ? size_t(f.m->nn) * N : 0);
        if (b) {
            to_device_order(old_ids(*f.m, kPerNode), b, tb.data(), tb.size());
            (N == 3 ? f.b3 : f.b1).upload(tb.data(), tb.size(), f.stream);
        }
        OAGPU_HIP(hipStreamSynchronize(f.stream));
    });
}

int oagpu_solution_download(oagpu_flow* h, int N, double* x) {
    return guarded([&] {
        OAGPU_REQUIRE(h && x && (N == 1 || N == 3), "bad argument");
        oagpu::Flow& f = h->f;
        std::vector<double> tx(size_t(f.m->nn) * N);
        (N == 3 ? f.x3 : f.x1).download(tx.data(), tx.size(), f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        to_caller_order(old_ids(*f.m, kPerNode), tx.data(), x, tx.size());
    });
}

int oagpu_update_mdot(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_update_mdot(h->f);
        check_err(h->f);
    });
}

int oagpu_mass_divergence(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_mass_divergence(h->f);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_gradient(oagpu_flow* h, int which) {
    return guarded([&] {
        OAGPU_REQUIRE(h && which >= 0 && which <= 3, "bad argument");
        oagpu::launch_gradient(h->f, which);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_correct(oagpu_flow* h, int N, double urf) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (N == 1 || N == 3), "bad argument");
        oagpu::launch_correct(h->f, N, urf);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_velocity_correction(oagpu_flow* h, int step) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (step == 1 || step == 2), "bad argument");
        oagpu::launch_velocity_correction(h->f, step);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_update_scales(oagpu_flow* h, double* U_scale, double* p_scale) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::field_scales(h->f, U_scale, p_scale);
    });
}

int oagpu_scaled_residual(oagpu_flow* h, int N, double* out) {
    return guarded([&] {
        OAGPU_REQUIRE(h && out && (N == 1 || N == 3), "bad argument");
        oagpu::scaled_residual(h->f, N, out);
    });
}

int oagpu_solve(oagpu_flow* h, int N, int method, int max_it, double rtol, double atol, int restart,
                oagpu_control_data* cd) {
    return guarded([&] {
        OAGPU_REQUIRE(h && cd, "null argument");
        OAGPU_REQUIRE(max_it >= 0, "negative iteration limit");
        std::memset(cd, 0, sizeof(*cd));
        oagpu::krylov_solve(h->f, N, method, max_it, rtol, atol, restart, cd);
    });
}

int oagpu_flow_set_bc_types(oagpu_flow* h, int n_bc, const int* type) {
    return guarded([&] {
        OAGPU_REQUIRE(h && type, "null argument");
        oagpu::Flow& f = h->f;
        OAGPU_REQUIRE(n_bc == f.nbc, "one type per boundary id");
        bool sym = false, slip = false;
        for (int b = 0; b < n_bc; ++b) {
            OAGPU_REQUIRE(type[b] >= OAGPU_BC_WALL && type[b] <= OAGPU_BC_OPENING,
                          "unsupported boundary type (wall, inlet, outlet, symmetry, slip wall, "
                          "opening)");
            sym = sym || type[b] == OAGPU_BC_SYMMETRY;
            slip = slip || type[b] == OAGPU_BC_SLIP;
        }
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        f.h_bc_type.assign(type, type + n_bc);
        f.bc_type.upload(f.h_bc_type);
        f.has_sym = sym;
        f.has_slip = slip;
        if (sym && !f.symA.p) f.symA.alloc(size_t(f.m->nn) * 3);
        // faces with side mass flows (inlets, outlets, openings), in face order
        const int nbf = f.m->nbf;
        std::vector<int> ids(size_t(nbf > 0 ? nbf : 1)), io;
        if (nbf) f.m->bf_bc.download(ids.data(), size_t(nbf));
        OAGPU_HIP(hipDeviceSynchronize());
        auto flow_side = [](int t) {
            return t == OAGPU_BC_INLET || t == OAGPU_BC_OUTLET || t == OAGPU_BC_OPENING;
        };
        f.has_outlet = f.has_opening = false;
        for (int i = 0; i < nbf; ++i) {
            const int t = type[ids[i]];
            if (flow_side(t)) io.push_back(i);
            f.has_outlet = f.has_outlet || t == OAGPU_BC_OUTLET;
            f.has_opening = f.has_opening || t == OAGPU_BC_OPENING;
        }
        f.nio = int(io.size());
        f.has_io = f.nio > 0;
        if (f.has_io) {
            f.io_face.upload(io);
            // rows nearest to an inlet / outlet / opening face ip, ascending
            const oagpu::DevMesh& dm = *f.m;
            std::vector<int> rows;
            for (int r = 0; r < dm.nn; ++r)
                for (int k = dm.h_binc_ptr[r]; k < dm.h_binc_ptr[r + 1]; ++k) {
                    const int t = type[ids[dm.h_binc[k] >> 2]];
                    if (flow_side(t)) {
                        rows.push_back(r);
                        break;
                    }
                }
            f.nio_row = int(rows.size());
            f.io_row.upload(rows);
            if (!f.msd.p) {
                f.msd.alloc(size_t(nbf) * 4);
                f.rfl.alloc(size_t(nbf) * 4);
            }
            f.msd.zero();
            f.rfl.zero();
            if (!f.bc_p.p) {
                for (auto* v : {&f.h_bc_p, &f.h_bc_k, &f.h_bc_w}) v->assign(size_t(n_bc), 0.0);
                f.bc_p.upload(f.h_bc_p);
                f.bc_k.upload(f.h_bc_k);
                f.bc_w.upload(f.h_bc_w);
            }
        }
        oagpu::launch_boundary_types(f);
        OAGPU_HIP(hipDeviceSynchronize());
    });
}

int oagpu_flow_set_bc_values(oagpu_flow* h, int n_bc, const double* p, const double* k,
                             const double* w) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::Flow& f = h->f;
        OAGPU_REQUIRE(n_bc == f.nbc, "one value per boundary id");
        if (f.h_bc_p.empty())
            for (auto* v : {&f.h_bc_p, &f.h_bc_k, &f.h_bc_w}) v->assign(size_t(n_bc), 0.0);
        if (p) f.h_bc_p.assign(p, p + n_bc);
        if (k) f.h_bc_k.assign(k, k + n_bc);
        if (w) f.h_bc_w.assign(w, w + n_bc);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        f.bc_p.upload(f.h_bc_p);
        f.bc_k.upload(f.h_bc_k);
        f.bc_w.upload(f.h_bc_w);
        OAGPU_HIP(hipDeviceSynchronize());
    });
}

int oagpu_update_reversal(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        if (h->f.has_outlet) oagpu::launch_reversal(h->f);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_reversal_download(oagpu_flow* h, int* flags, size_t count) {
    return guarded([&] {
        OAGPU_REQUIRE(h && flags, "null argument");
        oagpu::Flow& f = h->f;
        OAGPU_REQUIRE(count == size_t(f.m->nbf) * 4, "flags: [n_bfaces][4]");
        if (!f.rfl.p) {
            std::fill(flags, flags + count, 0);
            return;
        }
        f.rfl.download(flags, count, f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
    });
}

int oagpu_update_blending(oagpu_flow* h, int which) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (which == 0 || which == 2 || which == 3), "bad argument");
        OAGPU_REQUIRE(h->f.prm.high_res, "high resolution advection is off for this flow");
        oagpu::launch_blending(h->f, which);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_element_geometry(int topology, int shifted, int n_elems, const double* coords,
                           double* areav, double* scv, double* dndx, double* detj) {
    return guarded([&] {
        OAGPU_REQUIRE(topology >= OAGPU_HEX8 && topology <= OAGPU_PYR5, "unknown topology");
        OAGPU_REQUIRE(n_elems >= 0 && (coords || n_elems == 0), "bad element batch");
        if (n_elems == 0) return;
        const size_t n = size_t(n_elems);
        const size_t npe = oagpu::topology_npe(topology), ns = oagpu::topology_nscs(topology);
        oagpu::DBuf<double> c(n * npe * 3), A, V, G, D;
        oagpu::DBuf<int> err(1);
        err.zero();
        c.upload(coords, n * npe * 3);
        if (areav) A.alloc(n * ns * 3);
        if (scv) V.alloc(n * npe);
        if (dndx) G.alloc(n * ns * npe * 3);
        if (detj) D.alloc(n * ns);
        oagpu::element_geometry(topology, shifted ? 1 : 0, n_elems, c.p, A.p, V.p, G.p, D.p, err.p, 0);
        int e = 0;
        err.download(&e, 1);
        if (areav) A.download(areav, A.n);
        if (scv) V.download(scv, V.n);
        if (dndx) G.download(dndx, G.n);
        if (detj) D.download(detj, D.n);
        OAGPU_HIP(hipDeviceSynchronize());
        if (e) throw oagpu::Error(OAGPU_EGEOM, "non-positive Jacobian determinant in a kernel");
    });
}

int oagpu_element_geometry_noalloc(int topology, int shifted, int n_elems, const double* coords,
                                   double* areav, double* scv, double* dndx, double* detj,
                                   void* stream) {
    return guarded([&] {
        OAGPU_REQUIRE(topology >= OAGPU_HEX8 && topology <= OAGPU_PYR5, "unknown topology");
        OAGPU_REQUIRE(n_elems >= 0 && (coords || n_elems == 0), "bad element batch");
        // one lane per element, int element ids: a batch fits the grid
        OAGPU_REQUIRE(n_elems <= OAGPU_NOALLOC_MAX_ELEMS,
                      "batch larger than OAGPU_NOALLOC_MAX_ELEMS elements");
        ensure_init();
        const int e = oagpu::element_geometry_noalloc(topology, shifted ? 1 : 0, n_elems, coords,
                                                      areav, scv, dndx, detj,
                                                      static_cast<hipStream_t>(stream));
        if (e) throw oagpu::Error(OAGPU_EGEOM, "non-positive Jacobian determinant in a kernel");
    });
}

int oagpu_sst_enable(oagpu_flow* h, const oagpu_sst_params* sp) {
    return guarded([&] {
        OAGPU_REQUIRE(h && sp, "null argument");
        OAGPU_REQUIRE(sp->urf_k > 0 && sp->urf_w > 0 && sp->urf_mut > 0, "relaxation factors must be > 0");
        oagpu::Flow& f = h->f;
        const size_t nn = size_t(f.m->nn);
        f.sst = *sp;
        if (!f.sst_on) {
            for (auto* b : {&f.k, &f.om, &f.ymin, &f.mut, &f.F1, &f.Pk, &f.Pk_wall, &f.kprev})
                b->alloc(nn);
            f.gradK.alloc(3 * nn);
            f.gradW.alloc(3 * nn);
            if (f.prm.high_res) {
                f.beta_k.alloc(nn);
                f.beta_w.alloc(nn);
                f.beta_k.zero(f.stream);
                f.beta_w.zero(f.stream);
            }
            for (auto* b : {&f.k, &f.om, &f.ymin, &f.mut, &f.F1, &f.Pk, &f.Pk_wall, &f.kprev, &f.gradK,
                            &f.gradW})
                b->zero(f.stream);
            OAGPU_HIP(hipStreamSynchronize(f.stream));
        }
        f.sst_on = true;
        f.k_scale = f.w_scale = 1.0;
    });
}

int oagpu_sst_update(oagpu_flow* h) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_sst_update(h->f);
        OAGPU_HIP(hipStreamSynchronize(h->f.stream));
    });
}

int oagpu_assemble_scalar(oagpu_flow* h, int which) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::launch_assemble_scalar(h->f, which);
        check_err(h->f);
    });
}

int oagpu_correct_scalar(oagpu_flow* h, int which) {
    return guarded([&] {
        OAGPU_REQUIRE(h && (which == 0 || which == 1), "bad argument");
        oagpu::launch_correct_scalar(h->f, which);
    });
}

// segregatedShearStressTransportEquations preSolve + solve (one pass)
int oagpu_sst_iteration(oagpu_flow* h, const double* ctl, double* hist) {
    return guarded([&] {
        OAGPU_REQUIRE(h && ctl, "null argument");
        oagpu::Flow& f = h->f;
        OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
        const double ks0 = f.k_scale, ws0 = f.w_scale;
        const double kSmall = 2.220446049250313e-16;
        OAGPU_HIP(hipMemcpyAsync(f.kprev.p, f.k.p, f.k.n * sizeof(double), hipMemcpyDeviceToDevice,
                                 f.stream));
        oagpu::launch_sst_update(f);
        oagpu_control_data cdk{}, cdw{};
        double rk, rw;
        const bool pair = oagpu::scalar_pair_ok(f);
        if (pair) oagpu::launch_assemble_scalar_pair(f);  // omega's system to A1w / b1w
        else oagpu::launch_assemble_scalar(f, 0);
        oagpu::scaled_residual(f, 1, &rk);
        oagpu::krylov_solve(f, 1, OAGPU_GMRES, int(ctl[0]), ctl[1], 1e-30, 30, &cdk);
        oagpu::launch_correct_scalar(f, 0);
        if (pair) {  // the omega system in place of the k one while it is solved
            std::swap(f.A1, f.A1w);
            std::swap(f.b1, f.b1w);
        } else {
            oagpu::launch_assemble_scalar(f, 1);
        }
        struct SwapBack {
            oagpu::Flow& f;
            bool on;
            ~SwapBack() {
                if (on) {
                    std::swap(f.A1, f.A1w);
                    std::swap(f.b1, f.b1w);
                }
            }
        } back{f, pair};
        oagpu::scaled_residual(f, 1, &rw);
        oagpu::krylov_solve(f, 1, OAGPU_GMRES, int(ctl[0]), ctl[1], 1e-30, 30, &cdw);
        oagpu::launch_correct_scalar(f, 1);
        oagpu::launch_gradient_kw(f);
        check_err(f);
        if (hist) {
            hist[0] = rk / (ks0 + kSmall);
            hist[1] = rw / (ws0 + kSmall);
            hist[2] = cdk.n_iterations;
            hist[3] = cdw.n_iterations;
        }
    });
}

int oagpu_time_step(oagpu_flow* h, int scheme, double dt) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        OAGPU_REQUIRE(scheme == OAGPU_BDF1 || scheme == OAGPU_BDF2, "scheme must be BDF1 or BDF2");
        OAGPU_REQUIRE(dt > 0.0, "time step must be positive");
        oagpu::Flow& f = h->f;
        const size_t n3 = size_t(f.m->nn) * 3, n1 = size_t(f.m->nn);
        auto shift = [&](oagpu::DBuf<double>& old2, oagpu::DBuf<double>& old,
                         const oagpu::DBuf<double>& cur, size_t n) {
            if (!old.p) {
                old.alloc(n);
                old2.alloc(n);
                old.zero(f.stream);
            }
            OAGPU_HIP(hipMemcpyAsync(old2.p, old.p, n * sizeof(double), hipMemcpyDeviceToDevice,
                                     f.stream));
            OAGPU_HIP(hipMemcpyAsync(old.p, cur.p, n * sizeof(double), hipMemcpyDeviceToDevice,
                                     f.stream));
        };
        shift(f.Uold2, f.Uold, f.U, n3);
        if (f.sst_on) {
            shift(f.kold2, f.kold, f.k, n1);
            shift(f.wold2, f.wold, f.om, n1);
        }
        f.tdt_prev = f.tsteps > 0 ? f.tdt : 0.0;
        f.tdt = dt;
        f.tscheme = scheme;
        ++f.tsteps;
        OAGPU_HIP(hipStreamSynchronize(f.stream));
    });
}

int oagpu_amg_info(oagpu_flow* h, int* n_levels, int* sizes) {
    return guarded([&] {
        OAGPU_REQUIRE(h && n_levels, "null argument");
        const oagpu::Amg& a = h->f.amg;
        *n_levels = a.built ? int(a.lv.size()) : 0;
        if (sizes)
            for (int l = 0; l < *n_levels && l < 16; ++l) sizes[l] = a.lv[l].n;
    });
}

int oagpu_amg_level_copy(oagpu_flow* h, int level, int* agg, int* row_ptr, int* cols,
                         double* vals) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::Amg& a = h->f.amg;
        OAGPU_REQUIRE(a.built && level >= 1 && level < int(a.lv.size()), "no such AMG level");
        oagpu::AmgLevel& l = a.lv[level];
        if (agg) {
            if (level == 1)
                to_caller_order(old_ids(*h->f.m, kPerNode), l.h_agg.data(), agg, l.h_agg.size());
            else
                std::copy(l.h_agg.begin(), l.h_agg.end(), agg);
        }
        if (row_ptr) std::copy(l.h_row_ptr.begin(), l.h_row_ptr.end(), row_ptr);
        if (cols) std::copy(l.h_cols.begin(), l.h_cols.end(), cols);
        if (vals) {
            std::vector<double> sv(l.nsell);
            if (a.last_float) {
                std::vector<float> fv(l.nsell);
                l.fA.download(fv.data(), fv.size(), h->f.stream);
                OAGPU_HIP(hipStreamSynchronize(h->f.stream));
                std::copy(fv.begin(), fv.end(), sv.begin());
            } else {
                l.A.download(sv.data(), sv.size(), h->f.stream);
                OAGPU_HIP(hipStreamSynchronize(h->f.stream));
            }
            std::vector<int> sp;
            {
                // slice starts of the level's sliced-ELL layout
                const int nsl = (l.n + 63) / 64;
                sp.assign(size_t(nsl) + 1, 0);
                for (int s2 = 0; s2 < nsl; ++s2) {
                    int w = 0;
                    for (int r = 64 * s2; r < std::min(l.n, 64 * s2 + 64); ++r)
                        w = std::max(w, l.h_row_ptr[r + 1] - l.h_row_ptr[r]);
                    sp[s2 + 1] = sp[s2] + w;
                }
            }
            for (int r = 0; r < l.n; ++r)
                for (int j = l.h_row_ptr[r]; j < l.h_row_ptr[r + 1]; ++j)
                    vals[j] = sv[oagpu::sell_index(sp, r, j - l.h_row_ptr[r], 1, 0)];
        }
    });
}

int oagpu_sa_info(oagpu_flow* h, int* n_levels, int* sizes, long long* p_nnz, long long* a_nnz,
                  double* omega_p) {
    return guarded([&] {
        OAGPU_REQUIRE(h && n_levels, "null argument");
        const oagpu::Sa& sa = h->f.sa;
        *n_levels = sa.built ? int(sa.lv.size()) : 0;
        for (int l = 0; l < *n_levels && l < 16; ++l) {
            const oagpu::SaLevel& L = sa.lv[l];
            if (sizes) sizes[l] = L.n;
            if (a_nnz) a_nnz[l] = L.h_rp.empty() ? 0 : L.h_rp.back();
            if (p_nnz) p_nnz[l] = L.h_prp.empty() ? 0 : L.h_prp.back();
            if (omega_p) omega_p[l] = L.omega_p;
        }
    });
}

int oagpu_sa_level_copy(oagpu_flow* h, int level, int* agg, int* p_row_ptr, int* p_cols,
                        double* p_vals, int* a_row_ptr, int* a_cols, double* a_vals) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        oagpu::Flow& f = h->f;
        oagpu::Sa& sa = f.sa;
        OAGPU_REQUIRE(sa.built && level >= 1 && level < int(sa.lv.size()), "no such AMG level");
        oagpu::SaLevel& F = sa.lv[level - 1];
        oagpu::SaLevel& C = sa.lv[level];
        // the rows of level 0 are mesh nodes: caller numbering
        const bool remap = level == 1 && f.m->renumbered();
        const int nf = F.n;
        auto dev_row = [&](int r) { return remap ? f.m->n_new[r] : r; };
        if (agg)
            for (int r = 0; r < nf; ++r) agg[r] = F.h_agg[dev_row(r)];
        if (p_row_ptr || p_cols || p_vals) {
            const oagpu::SellHost ps = oagpu::make_sell(nf, F.h_prp, F.h_pcols);
            std::vector<double> pv;
            if (p_vals) {
                pv.resize(F.p_nsell);
                F.P.download(pv.data(), pv.size(), f.stream);
                OAGPU_HIP(hipStreamSynchronize(f.stream));
            }
            int o = 0;
            if (p_row_ptr) p_row_ptr[0] = 0;
            for (int r = 0; r < nf; ++r) {
                const int d = dev_row(r);
                for (int q = F.h_prp[d]; q < F.h_prp[d + 1]; ++q, ++o) {
                    if (p_cols) p_cols[o] = F.h_pcols[q];
                    if (p_vals) p_vals[o] = pv[oagpu::sell_index(ps.slice_ptr, d, q - F.h_prp[d], 1, 0)];
                }
                if (p_row_ptr) p_row_ptr[r + 1] = o;
            }
        }
        if (a_row_ptr) std::copy(C.h_rp.begin(), C.h_rp.end(), a_row_ptr);
        if (a_cols) std::copy(C.h_cols.begin(), C.h_cols.end(), a_cols);
        if (a_vals && !C.h_cols.empty()) {
            C.A.download(a_vals, C.h_cols.size(), f.stream);
            OAGPU_HIP(hipStreamSynchronize(f.stream));
        }
    });
}

namespace {
// simulation::assembleAndSolveSystems -> segregatedFlowEquations::solve
// (segregatedFlowEquations.cpp:76-297) for one p' sub-iteration
void run_outer_iteration(oagpu::Flow& f, const double* ctl, double* hist) {
    const double us0 = f.U_scale, ps0 = f.p_scale;
    const double kSmall = 2.220446049250313e-16;
    oagpu_control_data cdU{}, cdP{};
    double rU[3], rP[1];
    // momentum predictor (navierStokesEquation::solve :157-222)
    oagpu::launch_assemble_momentum(f);
    oagpu::scaled_residual(f, 3, rU);
    oagpu::krylov_solve(f, 3, OAGPU_GMRES, int(ctl[0]), ctl[1], 1e-30, 30, &cdU);
    oagpu::launch_correct(f, 3, 1.0);
    oagpu::launch_update_mdot(f);
    oagpu::field_scales(f, nullptr, nullptr);
    // pressure correction (pressureCorrectionEquation::solve :89-164)
    oagpu::launch_assemble_pcorr(f);
    oagpu::scaled_residual(f, 1, rP);
    oagpu::krylov_solve(f, 1, f.prm.p_method, int(ctl[2]), ctl[3], 1e-30, 30, &cdP);
    oagpu::launch_correct(f, 1, f.prm.urf_p);
    // mass flow, divergence, velocity correction, gradients (:106-285)
    oagpu::launch_update_mdot(f);
    if (f.has_outlet) oagpu::launch_reversal(f);  // segregatedFlowEquations.cpp:121
    oagpu::launch_mass_divergence(f);
    oagpu::launch_velocity_correction(f, 1);
    oagpu::launch_gradient(f, 1);
    oagpu::launch_velocity_correction(f, 2);
    oagpu::field_scales(f, nullptr, nullptr);
    oagpu::launch_gradient(f, 0);
    if (f.prm.high_res) oagpu::launch_blending(f, 0);  // segregatedFlowEquations.cpp:284-288
    check_err(f);
    if (hist) {
        for (int k = 0; k < 3; ++k) hist[k] = rU[k] / (us0 + kSmall);
        hist[3] = rP[0] / (ps0 + kSmall);
        hist[4] = cdU.n_iterations;
        hist[5] = cdP.n_iterations;
    }
}
}  // namespace

int oagpu_outer_iteration(oagpu_flow* h, const double* ctl, double* hist) {
    return guarded([&] {
        OAGPU_REQUIRE(h && ctl, "null argument");
        run_outer_iteration(h->f, ctl, hist);
    });
}

int oagpu_batch_outer_iterations(oagpu_flow* const* flows, int n, const double* ctl, int iters,
                                 double tol, int threads, double* hist, int* done,
                                 double* seconds) {
    return guarded([&] {
        OAGPU_REQUIRE(flows && ctl && n >= 0 && iters >= 0, "bad batch arguments");
        for (int c = 0; c < n; ++c) OAGPU_REQUIRE(flows[c], "null flow in batch");
        for (int c = 0; c < n; ++c)
            for (int d = 0; d < c; ++d) OAGPU_REQUIRE(flows[c] != flows[d], "flow listed twice");
        const int nt = std::max(1, std::min(threads > 0 ? threads : 16, n));
        std::atomic<int> next{0};
        std::mutex mu;
        std::string first_error;
        int first_code = OAGPU_OK;
        auto worker = [&] {
            for (int c = next++; c < n; c = next++) {
                oagpu::Flow& f = flows[c]->f;
                const auto t0 = std::chrono::steady_clock::now();
                int it = 0;
                try {
                    for (; it < iters; ++it) {
                        double hb[6];
                        run_outer_iteration(f, ctl, hb);
                        if (hist) std::copy(hb, hb + 6, hist + (size_t(c) * iters + it) * 6);
                        if (tol > 0 && std::max({hb[0], hb[1], hb[2], hb[3]}) < tol) {
                            ++it;
                            break;
                        }
                    }
                } catch (const oagpu::Error& e) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_code == OAGPU_OK) {
                        first_code = e.code;
                        first_error = "case " + std::to_string(c) + ": " + e.what();
                    }
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_code == OAGPU_OK) {
                        first_code = OAGPU_EINTERNAL;
                        first_error = "case " + std::to_string(c) + ": " + e.what();
                    }
                }
                const std::chrono::duration<double> dt = std::chrono::steady_clock::now() - t0;
                if (done) done[c] = it;
                if (seconds) seconds[c] = dt.count();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
        if (first_code != OAGPU_OK) throw oagpu::Error(first_code, first_error);
    });
}

int oagpu_graph_iterations(oagpu_flow* h, const double* ctl, const double* sctl, int n, double tol,
                           int eager, double* hist, int* done) {
    return guarded([&] {
        OAGPU_REQUIRE(h && ctl && n >= 0, "bad argument");
        const int it = oagpu::graph_iterations(h->f, ctl, sctl, n, tol, hist, eager != 0);
        if (done) *done = it;
    });
}

int oagpu_batch_graph_iterations(oagpu_flow* const* flows, int n, const double* ctl, int iters,
                                 double tol, int threads, double* hist, int* done, double* seconds) {
    return guarded([&] {
        OAGPU_REQUIRE(flows && ctl && n >= 0 && iters >= 0, "bad batch arguments");
        for (int c = 0; c < n; ++c) OAGPU_REQUIRE(flows[c], "null flow in batch");
        for (int c = 0; c < n; ++c)
            for (int d = 0; d < c; ++d) OAGPU_REQUIRE(flows[c] != flows[d], "flow listed twice");
        const int nt = std::max(1, std::min(threads > 0 ? threads : 16, n));
        std::atomic<int> next{0};
        std::mutex mu;
        std::string first_error;
        int first_code = OAGPU_OK;
        auto worker = [&] {
            for (int c = next++; c < n; c = next++) {
                const auto t0 = std::chrono::steady_clock::now();
                int it = 0;
                try {
                    it = oagpu::graph_iterations(flows[c]->f, ctl, nullptr, iters, tol,
                                                 hist ? hist + size_t(c) * iters * 10 : nullptr, false);
                } catch (const oagpu::Error& e) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_code == OAGPU_OK) {
                        first_code = e.code;
                        first_error = "case " + std::to_string(c) + ": " + e.what();
                    }
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_code == OAGPU_OK) {
                        first_code = OAGPU_EINTERNAL;
                        first_error = "case " + std::to_string(c) + ": " + e.what();
                    }
                }
                const std::chrono::duration<double> dt = std::chrono::steady_clock::now() - t0;
                if (done) done[c] = it;
                if (seconds) seconds[c] = dt.count();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
        if (first_code != OAGPU_OK) throw oagpu::Error(first_code, first_error);
    });
}

int oagpu_graph_stats(oagpu_flow* h, long long* launches, long long* redone, long long* captures,
                      int* caps) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        const oagpu::IterGraph& g = h->f.ig;
        if (launches) *launches = g.launches;
        if (redone) *redone = g.replays_redone;
        if (captures) *captures = g.captures;
        if (caps) std::copy(g.cap, g.cap + 4, caps);
    });
}

int oagpu_autotune(oagpu_flow* h, double* report, int n_report) {
    return guarded([&] {
        OAGPU_REQUIRE(h, "null flow");
        OAGPU_REQUIRE(n_report >= 0 && (n_report == 0 || report), "bad report buffer");
        oagpu::autotune(h->f, report, n_report);
        check_err(h->f);
    });
}

int oagpu_tune_set(oagpu_flow* h, const char* which, int variant) {
    return guarded([&] {
        OAGPU_REQUIRE(h && which, "null argument");
        auto& t = h->f.tune;
        const bool blk = variant == 64 || variant == 128 || variant == 256;
        if (!std::strcmp(which, "momentum")) {
            OAGPU_REQUIRE(variant == 0 || variant == 1, "momentum variant: 0 or 1");
            t.mom_tile = variant;
        } else if (!std::strcmp(which, "pcorr")) {
            OAGPU_REQUIRE(variant == 0 || variant == 1, "pcorr variant: 0 or 1");
            t.pc_occupancy = variant;
        } else if (!std::strcmp(which, "spmv1") || !std::strcmp(which, "spmv3") ||
                   !std::strcmp(which, "spmv1_f32") || !std::strcmp(which, "spmv3_f32")) {
            OAGPU_REQUIRE(blk, "SpMV variant: 64, 128 or 256 rows per workgroup");
            t.spmv_block[which[4] == '3'][std::strstr(which, "f32") != nullptr] = unsigned(variant);
        } else if (!std::strcmp(which, "scalar") || !std::strcmp(which, "mdot") ||
                   !std::strcmp(which, "grad")) {
            OAGPU_REQUIRE(variant == 0 || variant == 4 || variant == 8 || variant == 16,
                          "occupancy cap: 0, 4, 8 or 16 waves per CU");
            (which[0] == 's' ? t.cap_scalar : which[0] == 'm' ? t.cap_mdot : t.cap_grad) = variant;
        } else if (!std::strcmp(which, "scalar_pair")) {
            OAGPU_REQUIRE(variant == 0 || variant == 1, "scalar_pair: 0 or 1");
            t.scalar_pair = variant;
        } else if (!std::strcmp(which, "graph_cap")) {
            // iteration slots of every solve in the next captured iteration graph
            // (a solve that needs more rolls its iteration back and raises them)
            OAGPU_REQUIRE(variant >= 1, "graph_cap: at least 1 slot");
            OAGPU_REQUIRE(h->f.ig.primed, "graph_cap: run a graph iteration first");
            oagpu::IterGraph& g = h->f.ig;
            for (int& c : g.cap) c = variant;
            g.reset();
        } else {
            throw oagpu::Error(OAGPU_EINVAL, std::string("unknown tunable ") + which);
        }
    });
}

int oagpu_profile(oagpu_flow* h, const char* which) {
    return guarded([&] {
        OAGPU_REQUIRE(h && which, "null argument");
        oagpu::Flow& f = h->f;
        unsigned mask = 0;
        if (!std::strcmp(which, "all")) {
            mask = (1u << oagpu::kClkCount) - 1;
        } else if (std::strcmp(which, "none") != 0) {
            for (int k = 0; k < oagpu::kClkCount; ++k)
                if (!std::strcmp(which, oagpu::clock_name(k))) mask = 1u << k;
            OAGPU_REQUIRE(mask, std::string("unknown kernel name ") + which);
        }
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        for (auto& c : f.clk) {
            c.used = 0;
            c.total_ms = 0.0;
            c.launches = 0;
        }
        f.prof_mask = mask;
    });
}

int oagpu_kernel_stats(oagpu_flow* h, const char* which, double* total_ms, long long* launches) {
    return guarded([&] {
        OAGPU_REQUIRE(h && which && total_ms && launches, "null argument");
        oagpu::Flow& f = h->f;
        int id = -1;
        for (int k = 0; k < oagpu::kClkCount; ++k)
            if (!std::strcmp(which, oagpu::clock_name(k))) id = k;
        OAGPU_REQUIRE(id >= 0, std::string("unknown kernel name ") + which);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        f.clk[id].fold();
        *total_ms = f.clk[id].total_ms;
        *launches = f.clk[id].launches;
    });
}

}  // extern "C"
