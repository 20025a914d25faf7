// Ensemble of uneven cases stepped together (uws_batch_* in uws.h): one tile
// table spans every case's grid, so each kernel of the step is one launch for
// the whole ensemble, and each step is one captured HIP graph.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "uws_host.h"

using namespace uws;

struct uws_batch {
    std::vector<std::unique_ptr<SolverBase>> cases;
    hipStream_t stream = nullptr;
    int device = 0;
    StepCfg sc{};
    RoleState rs{};
    int ty = 4;
    int persistent = 0;
    int grid = 0;
    BatchTable<float> tab{};
    int* d_first = nullptr;
    ParamsT<float>* d_params = nullptr;
    FieldsT<float>* d_fields = nullptr;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};

    ~uws_batch() {
        drop_graphs();
        if (stream) (void)hipStreamSynchronize(stream);
        cases.clear();
        if (d_first) (void)hipFree(d_first);
        if (d_params) (void)hipFree(d_params);
        if (d_fields) (void)hipFree(d_fields);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs() {
        for (auto& g : graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    ParamsT<float>& params(int c) { return *static_cast<ParamsT<float>*>(cases[c]->device_params()); }
    FieldsT<float>& fields(int c) { return *static_cast<FieldsT<float>*>(cases[c]->device_fields()); }

    // rebuild the tile table for the current launch shape
    int build_table() {
        const int n = (int)cases.size();
        std::vector<int> first(n + 1, 0);
        std::vector<ParamsT<float>> ps(n);
        std::vector<FieldsT<float>> fs(n);
        for (int c = 0; c < n; ++c) {
            ParamsT<float>& p = params(c);
            p.ty = ty;
            p.tiles = tiles_of(p.X, p.Y, p.Z, ty);
            first[c + 1] = first[c] + p.tiles;
            ps[c] = p;
            fs[c] = fields(c);
        }
        if (hipMemcpyAsync(d_first, first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_params, ps.data(), sizeof(ParamsT<float>) * n, hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            hipMemcpyAsync(d_fields, fs.data(), sizeof(FieldsT<float>) * n, hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return UWS_ERR_DEVICE;
        tab.ncases = n;
        tab.total_tiles = first[n];
        tab.ty = ty;
        tab.persistent = persistent;
        tab.first_tile = d_first;
        tab.params = d_params;
        tab.fields = d_fields;
        if (persistent) {
            // 8 XCDs x 32 CUs, up to 8 resident 256-thread workgroups per CU
            int g = 256 * 8 * (4 / ty);
            if (g > tab.total_tiles) g = (tab.total_tiles + 7) / 8 * 8;
            grid = g < 8 ? 8 : g;
        } else {
            grid = tab.total_tiles;
        }
        drop_graphs();
        return UWS_OK;
    }

    struct L {
        uws_batch* b;
        void operator()(Kernel k, const Roles& r, int slot = 0) {
            launch_batch<float>(k, b->tab, b->grid, r, b->stream, slot);
        }
        int copy_vel(int dst, int src) {
            for (auto& cs : b->cases) {
                FieldsT<float>& f = *static_cast<FieldsT<float>*>(cs->device_fields());
                const size_t n = (size_t)cs->ncells();
                for (int c = 0; c < 3; ++c)
                    if (hipMemcpyAsync(f.vel[dst][c], f.vel[src][c], sizeof(float) * n, hipMemcpyDeviceToDevice,
                                       b->stream) != hipSuccess)
                        return UWS_ERR_DEVICE;
            }
            return UWS_OK;
        }
        int read_norm(int, double*) { return UWS_ERR_UNSUPPORTED; }
        double ncells() const { return 1.0; }
    };

    void sync_case_roles() {
        for (auto& c : cases) {
            c->rs = rs;
            c->rs.cur = -1;
            c->rs.pc_latest = -1;
        }
    }

    int step_once() {
        const int s0 = rs.state;
        if (!graph[s0]) {
            const RoleState saved = rs;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return UWS_ERR_DEVICE;
            L l{this};
            int rc = enqueue_step(l, sc, rs);
            hipError_t e = hipStreamEndCapture(stream, &g);
            rs = saved;
            if (rc == UWS_OK && hipGetLastError() != hipSuccess) rc = UWS_ERR_DEVICE;
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return UWS_ERR_DEVICE;
            e = hipGraphInstantiate(&graph[s0], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { graph[s0] = nullptr; return UWS_ERR_DEVICE; }
        }
        if (hipGraphLaunch(graph[s0], stream) != hipSuccess) return UWS_ERR_DEVICE;
        advance_roles(sc, rs);
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
    if (hipMalloc(&b->d_first, sizeof(int) * (ncases + 1)) != hipSuccess ||
        hipMalloc(&b->d_params, sizeof(ParamsT<float>) * ncases) != hipSuccess ||
        hipMalloc(&b->d_fields, sizeof(FieldsT<float>) * ncases) != hipSuccess)
        return UWS_ERR_ALLOC;
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
        return normalise_norms(raw, (double)s.ncells(), out, n);
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

int uws_batch_set_launch(uws_batch_handle b, int32_t ty, int32_t persistent) {
    if (!b || (ty != 1 && ty != 2 && ty != 4) || (persistent != 0 && persistent != 1)) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
    b->ty = ty;
    b->persistent = persistent;
    return b->build_table();
}

int uws_batch_autotune(uws_batch_handle b, int32_t steps, int32_t best[2], double* best_ms) {
    if (!b || steps <= 0 || !best || !best_ms) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return UWS_ERR_DEVICE;
    // save every case's state (u^n, p^n) so the trials leave no trace
    const int n = (int)b->cases.size();
    std::vector<std::vector<float>> su(n), sp(n);
    for (int c = 0; c < n; ++c) {
        SolverBase& s = *b->cases[c];
        s.rs.state = b->rs.state;
        su[c].resize((size_t)s.ncells() * 3);
        sp[c].resize((size_t)s.ncells());
        if (int rc = s.get_velocity(su[c].data(), 3)) return rc;
        if (int rc = s.get_plane(PLANE_P, 0, sp[c].data(), sp[c].size())) return rc;
    }
    const int tys[3] = {4, 2, 1};
    double bestv = 1e300;
    int bty = 4, bp = 0;
    for (int pers = 0; pers < 2; ++pers)
        for (int ty : tys) {
            if (int rc = uws_batch_set_launch(b, ty, pers)) return rc;
            if (int rc = uws_batch_step(b, 1)) return rc;  // capture + warm
            if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
            const auto t0 = std::chrono::steady_clock::now();
            if (int rc = uws_batch_step(b, steps)) return rc;
            if (hipStreamSynchronize(b->stream) != hipSuccess) return UWS_ERR_DEVICE;
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
            if (ms < bestv) { bestv = ms; bty = ty; bp = pers; }
        }
    if (int rc = uws_batch_set_launch(b, bty, bp)) return rc;
    for (int c = 0; c < n; ++c) {
        SolverBase& s = *b->cases[c];
        s.rs.state = b->rs.state;
        if (int rc = s.set_velocity_all(su[c].data(), 3)) return rc;
        if (int rc = s.set_plane(PLANE_P, 0, sp[c].data(), sp[c].size())) return rc;
        s.step -= (int64_t)(steps + 1) * 6;
        s.phys_time -= (double)s.cfg.dt * (steps + 1) * 6;
        s.rs.state = b->rs.state;
    }
    best[0] = bty;
    best[1] = bp;
    *best_ms = bestv;
    return UWS_OK;
}

}  // extern "C"
