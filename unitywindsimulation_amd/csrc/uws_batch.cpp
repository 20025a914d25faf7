// Ensemble of uneven cases stepped together (uws_batch_* in uws.h): one tile
// table spans every case's grid, so each kernel of the step is one launch for
// the whole ensemble, and each step is one captured HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "uws_host.h"

using namespace uws;

// One launch group: cases whose sweep working set fits the budget share
// every launch of the step.
struct BatchGroup {
    std::vector<int> cases;
    BatchTable<float> tab{};
    int grid = 0;
    bool vel_pairs = false;  // two-sweep momentum launches
    int* d_first = nullptr;
    int* d_first2 = nullptr;
    int* d_first3 = nullptr;
    int* d_first4 = nullptr;
    bool quads = false;  // four-sweep pressure launches
    ParamsT<float>* d_params = nullptr;
    FieldsT<float>* d_fields = nullptr;
    std::vector<FieldsT<float>> h_fields;  // d_fields' contents, for launches that take them as arguments
    void release() {
        if (d_first) (void)hipFree(d_first);
        if (d_first2) (void)hipFree(d_first2);
        d_first2 = nullptr;
        if (d_first3) (void)hipFree(d_first3);
        d_first3 = nullptr;
        if (d_first4) (void)hipFree(d_first4);
        d_first4 = nullptr;
        if (d_params) (void)hipFree(d_params);
        if (d_fields) (void)hipFree(d_fields);
        d_first = nullptr;
        d_params = nullptr;
        d_fields = nullptr;
    }
};

struct uws_batch {
    std::vector<std::unique_ptr<SolverBase>> cases;
    hipStream_t stream = nullptr;
    int device = 0;
    StepCfg sc{};
    RoleState rs{};
    int ty = 4;
    int persistent = 0;
    int64_t group_bytes = 0;  // 0: one group; > 0: consecutive cases within a byte budget; < 0: packed rounds
    int fuse_mode = fusion_mode();  // two-sweep pressure kernel: see fusion_mode()
    std::vector<BatchGroup> groups;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    RoleState graph_after[3] = {};  // the roles each graph's step leaves

    ~uws_batch() {
        drop_graphs();
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : groups) g.release();
        cases.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs() {
        for (auto& g : graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    ParamsT<float>& params(int c) { return *static_cast<ParamsT<float>*>(cases[c]->device_params()); }
    FieldsT<float>& fields(int c) { return *static_cast<FieldsT<float>*>(cases[c]->device_fields()); }

    // Cases packed into launch groups of at most one round of two-sweep
    // workgroups each (one 16-wave workgroup per CU, at the 16-plane z chunk
    // and 13-row tiles every group of this mode uses): the fewest groups that
    // fit, found by placing the cases, largest first, each into the least
    // loaded group it fits (a case larger than a round gets a group of its
    // own).  A group of about one round keeps every launch of it a single
    // wave of workgroups that start together, and a working set (3M cells:
    // 200 MB of momentum arrays) that the 256 MB last-level cache holds
    // across its sweeps.
    std::vector<std::vector<int>> pack_rounds(const std::vector<ParamsT<float>>& shaped) const {
        // workgroups per group: one round (UWS_PACK_BLOCKS overrides, for tests)
        int cap = FUSE_MIN_BLOCKS;
        if (const char* e = std::getenv("UWS_PACK_BLOCKS")) cap = std::max(1, std::atoi(e));
        std::vector<int> w(shaped.size());
        for (size_t c = 0; c < shaped.size(); ++c) {
            const ParamsT<float>& p = shaped[c];
            w[c] = pres_sweep2_grid(p.X, p.Y, p.Z, PACK_ZCHUNK, p.by2);
        }
        return pack_groups(w, cap);
    }
    static constexpr int PACK_ZCHUNK = 16;

    // Rebuild the launch groups and their tile tables for the current launch
    // shape.  group_bytes > 0: a group collects consecutive cases while the
    // momentum sweep's working set (77 B per cell) stays within group_bytes;
    // 0: one group; < 0: pack_rounds.
    int build_table() {
        drop_graphs();
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : groups) g.release();
        groups.clear();
        const int n = (int)cases.size();
        // the ensemble's launch shape lives in these copies only: each case's
        // own parameters keep the shape its handle chose for its own launches
        std::vector<ParamsT<float>> shaped(n);
        for (int c = 0; c < n; ++c) {
            ParamsT<float>& p = shaped[c];
            p = params(c);
            p.ty = ty;
            p.tiles = tiles_of(p.X, p.Y, p.Z, ty);
            // one launch serves every case: the full two-sweep tile height
            p.by2 = 15;
        }
        const bool packed = group_bytes < 0;
        if (packed) {
            for (auto& cs : pack_rounds(shaped)) {
                groups.emplace_back();
                groups.back().cases = cs;
            }
        } else {
            int64_t acc = 0;
            for (int c = 0; c < n; ++c) {
                const int64_t b = 77LL * cases[c]->ncells();
                if (groups.empty() || (group_bytes > 0 && acc + b > group_bytes)) {
                    groups.emplace_back();
                    acc = 0;
                }
                groups.back().cases.push_back(c);
                acc += b;
            }
        }
        for (auto& g : groups) {
            const int m = (int)g.cases.size();
            // z chunk of the group's two-sweep launches, chosen over the whole
            // group as choose_zchunk does for one grid: the longest of 16, 8
            // and 4 planes that still gives a workgroup per CU, preferring one
            // that divides every case's Z (UWS_ZCHUNK overrides, for tuning);
            // packed groups were sized at the 16-plane chunk
            if (packed && !std::getenv("UWS_ZCHUNK")) {
                for (int c : g.cases) shaped[c].zchunk = PACK_ZCHUNK;
            } else if (!std::getenv("UWS_ZCHUNK")) {
                int zbest = 8, sbest = -1;
                for (int zc : {16, 8, 4}) {
                    int blocks = 0;
                    bool divides = true;
                    for (int c : g.cases) {
                        const ParamsT<float>& p = shaped[c];
                        blocks += pres_sweep2_grid(p.X, p.Y, p.Z, zc, p.by2);
                        divides = divides && p.Z % zc == 0;
                    }
                    if (blocks < FUSE_MIN_BLOCKS) continue;
                    const int score = (divides ? 2 : 0) + (zc == 16 ? 1 : 0);
                    if (score > sbest) { zbest = zc; sbest = score; }
                }
                for (int c : g.cases) shaped[c].zchunk = zbest;
            }
            // z chunk of the group's three-sweep launches: the one that
            // minimises rounds of 256 workgroups times planes per workgroup
            // over the group, as choose_zchunk3 does for one grid
            if (!std::getenv("UWS_ZCHUNK3")) {
                int zbest = 4, cbest = -1;
                for (int zc = 4; zc <= 64; ++zc) {
                    int blocks = 0;
                    for (int c : g.cases) {
                        const ParamsT<float>& p = shaped[c];
                        blocks += pres_sweep3_grid(p.X, p.Y, p.Z, zc);
                    }
                    const int cost = ((blocks + 255) / 256) * (zc + 4);
                    if (cbest < 0 || cost < cbest) { zbest = zc; cbest = cost; }
                }
                for (int c : g.cases) shaped[c].zchunk3 = zbest;
            }
            // z chunk of the group's four-sweep launches, the same way (three
            // halo planes above and below a chunk)
            if (!std::getenv("UWS_ZCHUNK4")) {
                int zbest = 4, cbest = -1;
                for (int zc = 4; zc <= 64; ++zc) {
                    int blocks = 0;
                    for (int c : g.cases) {
                        const ParamsT<float>& p = shaped[c];
                        blocks += pres_sweep4_grid(p.X, p.Y, p.Z, zc);
                    }
                    const int cost = ((blocks + 255) / 256) * (zc + 6);
                    if (cbest < 0 || cost < cbest) { zbest = zc; cbest = cost; }
                }
                for (int c : g.cases) shaped[c].zchunk4 = zbest;
            }
            std::vector<int> first(m + 1, 0), first2(m + 1, 0), first3(m + 1, 0), first4(m + 1, 0);
            int64_t min_cells = INT64_MAX;
            std::vector<ParamsT<float>> ps(m);
            std::vector<FieldsT<float>> fs(m);
            for (int i = 0; i < m; ++i) {
                ps[i] = shaped[g.cases[i]];
                fs[i] = fields(g.cases[i]);
                first[i + 1] = first[i] + ps[i].tiles;
                first2[i + 1] = first2[i] + pres_sweep2_grid(ps[i].X, ps[i].Y, ps[i].Z, ps[i].zchunk, ps[i].by2);
                first3[i + 1] = first3[i] + pres_sweep3_grid(ps[i].X, ps[i].Y, ps[i].Z, ps[i].zchunk3);
                first4[i + 1] = first4[i] + pres_sweep4_grid(ps[i].X, ps[i].Y, ps[i].Z, ps[i].zchunk4);
                min_cells = std::min(min_cells, (int64_t)ps[i].X * ps[i].Y * ps[i].Z);
            }
            if (hipMalloc(&g.d_first, sizeof(int) * (m + 1)) != hipSuccess ||
                hipMalloc(&g.d_first2, sizeof(int) * (m + 1)) != hipSuccess ||
                hipMalloc(&g.d_first3, sizeof(int) * (m + 1)) != hipSuccess ||
                hipMalloc(&g.d_first4, sizeof(int) * (m + 1)) != hipSuccess ||
                hipMalloc(&g.d_params, sizeof(ParamsT<float>) * m) != hipSuccess ||
                hipMalloc(&g.d_fields, sizeof(FieldsT<float>) * m) != hipSuccess)
                return UWS_ERR_ALLOC;
            if (hipMemcpyAsync(g.d_first, first.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(g.d_first2, first2.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(g.d_first3, first3.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(g.d_first4, first4.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(g.d_params, ps.data(), sizeof(ParamsT<float>) * m, hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(g.d_fields, fs.data(), sizeof(FieldsT<float>) * m, hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                // the host tables above go out of scope with this iteration
                hipStreamSynchronize(stream) != hipSuccess)
                return UWS_ERR_DEVICE;
            g.tab.ncases = m;
            g.tab.total_tiles = first[m];
            g.tab.ty = ty;
            g.tab.persistent = persistent;
            g.tab.first_tile = g.d_first;
            g.tab.total_ps2 = first2[m];
            g.tab.first_ps2 = g.d_first2;
            g.tab.total_ps3 = first3[m];
            g.tab.first_ps3 = g.d_first3;
            g.tab.total_ps4 = first4[m];
            g.tab.first_ps4 = g.d_first4;
            // four-sweep launches where every case's grid would take them alone
            // (the 24-case ensemble, of grids of 0.26M to 1.2M cells, ran 6 %
            // slower with them than with triples in its packed groups)
            g.quads = use_fusion4(fuse_mode, min_cells > INT32_MAX ? INT32_MAX : (int)min_cells);
            g.tab.params = g.d_params;
            g.tab.fields = g.d_fields;
            g.h_fields = fs;
            // the momentum pair tiles the grid as the two-sweep pressure kernel
            // does (first_ps2), with the group's z chunk
            int zmin = ps[0].zchunk;
            for (const auto& q : ps) zmin = std::min(zmin, q.zchunk);
            // a packed group is sized to one round: it takes pairs when it
            // fills at least half of it
            g.vel_pairs = use_vel_pairs(fuse_mode, packed ? 2 * g.tab.total_ps2 : g.tab.total_ps2, zmin);
            if (persistent) {
                // 8 XCDs x 32 CUs, up to 8 resident 256-thread workgroups per CU
                int gr = 256 * 8 * (4 / ty);
                if (gr > g.tab.total_tiles) gr = (g.tab.total_tiles + 7) / 8 * 8;
                g.grid = gr < 8 ? 8 : gr;
            } else {
                g.grid = g.tab.total_tiles;
            }
        }
        return hipStreamSynchronize(stream) == hipSuccess ? UWS_OK : UWS_ERR_DEVICE;
    }

    struct L {
        uws_batch* b;
        BatchGroup* g;
        void operator()(Kernel k, const Roles& r, int slot = 0) {
            launch_batch<float>(k, g->tab, g->grid, r, b->stream, slot, g->h_fields.data());
        }
        int copy_vel(int dst, int src) {
            for (int c : g->cases) {
                FieldsT<float>& f = b->fields(c);
                const size_t n = (size_t)b->cases[c]->ncells();
                for (int k = 0; k < 3; ++k)
                    if (hipMemcpyAsync(f.vel[dst][k], f.vel[src][k], sizeof(float) * n, hipMemcpyDeviceToDevice,
                                       b->stream) != hipSuccess)
                        return UWS_ERR_DEVICE;
            }
            return UWS_OK;
        }
        int read_norm(int, double*) { return UWS_ERR_UNSUPPORTED; }
        double ncells() const { return 1.0; }
        bool fused() const { return use_fusion(b->fuse_mode); }
        bool fused_vel() const { return g->vel_pairs; }
        bool fused3() const { return use_fusion3(b->fuse_mode); }
        bool fused4() const { return g->quads; }
    };

    void sync_case_roles() {
        for (auto& c : cases) {
            c->rs = rs;
            c->rs.cur = -1;
            c->rs.pc_latest = -1;
        }
    }

    // every group runs its whole step before the next group starts, so a
    // group's sweeps reuse its data from the last-level cache
    int enqueue_all() {
        RoleState out = rs;
        for (auto& g : groups) {
            RoleState r = rs;
            L l{this, &g};
            if (int rc = enqueue_step(l, sc, r)) return rc;
            out = r;
        }
        rs = out;
        return UWS_OK;
    }

    int step_once() {
        const int s0 = rs.state;
        if (!graph[s0]) {
            const RoleState saved = rs;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return UWS_ERR_DEVICE;
            int rc = enqueue_all();
            hipError_t e = hipStreamEndCapture(stream, &g);
            graph_after[s0] = rs;
            rs = saved;
            if (rc == UWS_OK && hipGetLastError() != hipSuccess) rc = UWS_ERR_DEVICE;
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return UWS_ERR_DEVICE;
            e = hipGraphInstantiate(&graph[s0], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { graph[s0] = nullptr; return UWS_ERR_DEVICE; }
        }
        if (hipGraphLaunch(graph[s0], stream) != hipSuccess) return UWS_ERR_DEVICE;
        rs = graph_after[s0];
        for (auto& c : cases) {
            c->step += 1;
            c->phys_time += (double)c->cfg.dt;
        }
        return UWS_OK;
    }
};

namespace {

bool same_schedule(const uws_config& a, const uws_config& b) {
    return a.vel_max_num_iter == b.vel_max_num_iter && a.pres_max_num_iter == b.pres_max_num_iter &&
           a.fvm_solver_type == b.fvm_solver_type && a.piso_num_correctors == b.piso_num_correctors &&
           a.turbulence_model == b.turbulence_model && a.precision == b.precision &&
           a.compute_norms == b.compute_norms;
}

template <class F>
int with_case(uws_batch_handle b, int32_t c, F&& fn) {
    if (!b || c < 0 || c >= (int32_t)b->cases.size()) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    int rc = fn(*b->cases[c]);
    b->cases[c]->rs.state = b->rs.state;  // the ensemble owns the buffer roles
    return rc;
}

}  // namespace

extern "C" {

int uws_batch_create(const uws_config* cfgs, int32_t ncases, uws_batch_handle* out) {
    if (!out || !cfgs || ncases <= 0) return UWS_ERR_INVALID_ARG;
    *out = nullptr;
    for (int c = 0; c < ncases; ++c) {
        if (int rc = validate(&cfgs[c], false)) return rc;
        if (cfgs[c].grid_type != UWS_GRID_COLLOCATED) return UWS_ERR_UNSUPPORTED;
        if (cfgs[c].precision != UWS_PREC_FP32) return UWS_ERR_UNSUPPORTED;
        if (!same_schedule(cfgs[0], cfgs[c])) return UWS_ERR_UNSUPPORTED;
    }
    std::unique_ptr<uws_batch> b(new (std::nothrow) uws_batch());
    if (!b) return UWS_ERR_ALLOC;
    if (hipGetDevice(&b->device) != hipSuccess) return UWS_ERR_DEVICE;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->stream = nullptr;
        return UWS_ERR_DEVICE;
    }
    b->cases.resize(ncases);
    for (int c = 0; c < ncases; ++c)
        if (int rc = create_solver(cfgs[c], false, b->stream, &b->cases[c])) return rc;
    b->sc = step_cfg(cfgs[0]);
    b->sc.normalise = false;
    for (int c = 0; c < ncases; ++c) b->sc.normalise |= step_cfg(cfgs[c]).normalise;
    if (int rc = b->build_table()) return rc;
    *out = b.release();
    return UWS_OK;
}

int uws_batch_destroy(uws_batch_handle b) {
    if (!b) return UWS_ERR_INVALID_ARG;
    (void)hipSetDevice(b->device);
    delete b;
    return UWS_OK;
}

int uws_batch_num_cases(uws_batch_handle b, int32_t* n) {
    if (!b || !n) return UWS_ERR_INVALID_ARG;
    *n = (int32_t)b->cases.size();
    return UWS_OK;
}

int uws_batch_init_flags(uws_batch_handle b, int32_t c, const int32_t* flags, int64_t n) {
    if (!flags) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) { return flags_from_int(s, flags, n); });
}

int uws_batch_init_noise(uws_batch_handle b, int32_t c, uint64_t seed, float amp) {
    if (!(amp >= 0.f)) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) {
        s.rs.state = b->rs.state;
        return apply_noise(s, seed, amp);
    });
}

int uws_batch_set_fixed_value_vel_bnd_cond(uws_batch_handle b, int32_t c, int32_t inflow_type, float wind_speed,
                                           const float vel_zn[3], float ref_height, float alpha) {
    return with_case(b, c, [&](SolverBase& s) {
        return apply_fixed_value_bnd(s, inflow_type, wind_speed, vel_zn, ref_height, alpha);
    });
}

int uws_batch_set_velocity(uws_batch_handle b, int32_t c, const float* src, int32_t ncomp, int64_t n) {
    if (!src || (ncomp != 3 && ncomp != 4)) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.set_velocity_all(src, ncomp);
    });
}

int uws_batch_get_velocity(uws_batch_handle b, int32_t c, float* dst, int32_t ncomp, int64_t n) {
    if (!dst || (ncomp != 3 && ncomp != 4)) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        s.rs.state = b->rs.state;
        return s.get_velocity(dst, ncomp);
    });
}

int uws_batch_get_pressure(uws_batch_handle b, int32_t c, float* dst, int64_t n) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.get_plane(PLANE_P, 0, dst, n);
    });
}

int uws_batch_get_norms(uws_batch_handle b, int32_t c, double* out, int32_t n) {
    if (!out || n < 0) return UWS_ERR_INVALID_ARG;
    return with_case(b, c, [&](SolverBase& s) {
        double raw[UWS_NORM_COUNT];
        if (int rc = s.get_norms(raw)) return rc;
        if (int rc = normalise_norms(raw, (double)s.ncells(), out, n)) return rc;
        return raw[UWS_NORM_NONFINITE] > 0.0 ? (int)UWS_ERR_NONFINITE : (int)UWS_OK;
    });
}

int uws_batch_step(uws_batch_handle b, int32_t nsteps) {
    if (!b || nsteps < 0) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    for (int i = 0; i < nsteps; ++i)
        if (int rc = b->step_once()) return rc;
    b->sync_case_roles();
    return UWS_OK;
}

int uws_batch_synchronize(uws_batch_handle b) {
    if (!b) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    return hipStreamSynchronize(b->stream) == hipSuccess ? UWS_OK : UWS_ERR_DEVICE;
}

int uws_batch_set_launch(uws_batch_handle b, int32_t ty, int32_t persistent, int64_t group_bytes) {
    if (!b || (ty != 1 && ty != 2 && ty != 4) || (persistent != 0 && persistent != 1)) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
    b->ty = ty;
    b->persistent = persistent;
    b->group_bytes = group_bytes;
    return b->build_table();
}

int uws_batch_num_groups(uws_batch_handle b, int32_t* n) {
    if (!b || !n) return UWS_ERR_INVALID_ARG;
    *n = (int32_t)b->groups.size();
    return UWS_OK;
}

int uws_batch_autotune(uws_batch_handle b, int32_t steps, int64_t best[3], double* best_ms) {
    if (!b || steps <= 0 || !best || !best_ms) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    // save every case's state (u^n, p^n) and clock so the trials leave no trace
    const int n = (int)b->cases.size();
    std::vector<std::vector<float>> su(n), sp(n);
    std::vector<int64_t> st(n);
    std::vector<double> pt(n);
    for (int c = 0; c < n; ++c) {
        SolverBase& s = *b->cases[c];
        s.rs.state = b->rs.state;
        su[c].resize((size_t)s.ncells() * 3);
        sp[c].resize((size_t)s.ncells());
        st[c] = s.step;
        pt[c] = s.phys_time;
        if (int rc = s.get_velocity(su[c].data(), 3)) return rc;
        if (int rc = s.get_plane(PLANE_P, 0, sp[c].data(), sp[c].size())) return rc;
    }
    // candidate launch shapes: (tile rows, persistent, group budget in bytes)
    const int64_t MB = 1 << 20;
    const int64_t budgets[5] = {-1, 0, 384 * MB, 192 * MB, 96 * MB};
    const int shapes[3][2] = {{4, 0}, {2, 0}, {4, 1}};
    double bestv = 1e300;
    int64_t bb[3] = {4, 0, 0};
    for (int64_t gb : budgets)
        for (const auto& sh : shapes) {
            if (int rc = uws_batch_set_launch(b, sh[0], sh[1], gb)) return rc;
            if (int rc = uws_batch_step(b, 1)) return rc;  // capture + warm
            if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
            const auto t0 = std::chrono::steady_clock::now();
            if (int rc = uws_batch_step(b, steps)) return rc;
            if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
            if (ms < bestv) { bestv = ms; bb[0] = sh[0]; bb[1] = sh[1]; bb[2] = gb; }
        }
    if (int rc = uws_batch_set_launch(b, (int)bb[0], (int)bb[1], bb[2])) return rc;
    for (int c = 0; c < n; ++c) {
        SolverBase& s = *b->cases[c];
        s.rs.state = b->rs.state;
        if (int rc = s.set_velocity_all(su[c].data(), 3)) return rc;
        if (int rc = s.set_plane(PLANE_P, 0, sp[c].data(), sp[c].size())) return rc;
        s.step = st[c];
        s.phys_time = pt[c];
        s.rs.state = b->rs.state;
    }
    best[0] = bb[0];
    best[1] = bb[1];
    best[2] = bb[2];
    *best_ms = bestv;
    return UWS_OK;
}

}  // extern "C"
