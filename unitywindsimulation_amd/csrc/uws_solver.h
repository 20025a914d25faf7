// The per-type solver object behind a uws_handle (shared by the C ABI, the
// ensemble and the adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "uws_host.h"

namespace uws {

#define HIP_OK(expr)                                     \
    do {                                                 \
        if ((expr) != hipSuccess) return UWS_ERR_DEVICE; \
    } while (0)

inline int check_launch() { return hipGetLastError() == hipSuccess ? UWS_OK : UWS_ERR_DEVICE; }

// Everything a handle owns, for one arithmetic type.
template <class T>
struct Solver : SolverBase {
    ParamsT<T> p{};
    FieldsT<T> f{};
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // one per starting state buffer
    RoleState graph_after[3] = {};                          // the roles each graph's step leaves

    double* mean_sum[3] = {nullptr, nullptr, nullptr};  // FP64 time sum of u (allocated on first use)
    int64_t mean_count = 0;

    ~Solver() override {
        // the stream may still run a replay of a graph: finish it before its executable goes
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        if (mean_sum[0]) (void)hipFree(mean_sum[0]);
    }
    void drop_graphs() {
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    // single-case launcher for the step driver
    struct L {
        Solver* s;
        void operator()(Kernel k, const Roles& r, int slot = 0) { launch<T>(k, s->p, s->f, r, s->stream, slot); }
        int copy_vel(int dst, int src) {
            for (int c = 0; c < 3; ++c)
                if (hipMemcpyAsync(s->f.vel[dst][c], s->f.vel[src][c], sizeof(T) * s->p.N, hipMemcpyDeviceToDevice,
                                   s->stream) != hipSuccess)
                    return UWS_ERR_DEVICE;
            return UWS_OK;
        }
        int read_norm(int slot, double* out) {
            if (hipMemcpyAsync(s->h_norms, s->f.norms + slot, sizeof(double), hipMemcpyDeviceToHost, s->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(s->stream) != hipSuccess)
                return UWS_ERR_DEVICE;
            *out = s->h_norms[0];
            return UWS_OK;
        }
        double ncells() const { return (double)s->p.N; }
        bool fused() const { return use_fusion(s->fuse_mode); }
        // the two-sweep momentum kernel: FP32 (uniform or stretched grid)
        bool fused_vel() const {
            return std::is_same<T, float>::value &&
                   use_vel_pairs(s->fuse_mode, pres_sweep2_grid(s->p.X, s->p.Y, s->p.Z, s->p.zchunk, s->p.by2), s->p.zchunk);
        }
        // the three-sweep pressure kernel: FP32
        bool fused3() const { return std::is_same<T, float>::value && use_fusion3(s->fuse_mode); }
        bool fused4() const { return std::is_same<T, float>::value && use_fusion4(s->fuse_mode, s->p.N); }
    };
    int fuse_mode = fusion_mode();  // two-sweep pressure kernel: see fusion_mode()

    int init(const uws_config& c) {
        fill_params(c, p);
        slab_bytes = carve(p, f, nullptr, step_cfg(c).lowp);
        if (hipMalloc(&slab, slab_bytes) != hipSuccess) { slab = nullptr; return UWS_ERR_ALLOC; }
        carve(p, f, (char*)slab, step_cfg(c).lowp);
        if (hipMemsetAsync(slab, 0, slab_bytes, stream) != hipSuccess) return UWS_ERR_DEVICE;
        if (int rc = sync()) return rc;
        if (p.nonuniform) {
            // uniform faces i*dx until the caller sets its own
            std::vector<float> pos[3];
            const int n[3] = {p.X, p.Y, p.Z};
            for (int a = 0; a < 3; ++a) {
                pos[a].resize((size_t)n[a] + 1);
                for (int i = 0; i <= n[a]; ++i) pos[a][i] = (float)i * c.dx;
            }
            const float* const pp[3] = {pos[0].data(), pos[1].data(), pos[2].data()};
            return set_face_positions(pp);
        }
        return UWS_OK;
    }

    // stretched grid: facePosX/Y/Z, strictly increasing; the kernels read the
    // cell sizes facePos[i+1] - facePos[i] formed here in T
    int set_face_positions(const float* const pos[3]) override {
        if (!p.nonuniform) return UWS_ERR_CONFIG;
        const int n[3] = {p.X, p.Y, p.Z};
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < n[a]; ++i)
                if (!(pos[a][i + 1] > pos[a][i]) || !std::isfinite(pos[a][i + 1])) return UWS_ERR_INVALID_ARG;
        for (int a = 0; a < 3; ++a) {
            face_pos[a].assign(pos[a], pos[a] + n[a] + 1);
            std::vector<T> h((size_t)n[a]);
            for (int i = 0; i < n[a]; ++i) h[i] = (T)pos[a][i + 1] - (T)pos[a][i];
            HIP_OK(hipMemcpyAsync(f.h[a], h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, stream));
            if (int rc = sync()) return rc;
        }
        return UWS_OK;
    }
    // dv/den of every cell in T (the factor folded into the stretched grid's pressure equation)
    std::vector<T> dvden_host() const {
        std::vector<T> r((size_t)p.N);
        std::vector<T> h[3];
        const int n[3] = {p.X, p.Y, p.Z};
        for (int a = 0; a < 3; ++a) {
            h[a].resize((size_t)n[a]);
            for (int i = 0; i < n[a]; ++i) h[a][i] = (T)face_pos[a][i + 1] - (T)face_pos[a][i];
        }
        for (int z = 0; z < p.Z; ++z)
            for (int y = 0; y < p.Y; ++y)
                for (int x = 0; x < p.X; ++x)
                    r[((size_t)z * p.Y + y) * p.X + x] = h[0][x] * h[1][y] * h[2][z] / p.den;
        return r;
    }

    int ncells() const override { return p.N; }
    int dims(int i) const override { return i == 0 ? p.X : (i == 1 ? p.Y : p.Z); }
    void* device_fields() override { return &f; }
    void* device_params() override { return &p; }

    // ---- uploads / downloads (host arrays are float or double) ----
    template <class H>
    int up(T* dst, const H* src, size_t n) {
        std::vector<T> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
        HIP_OK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, stream));
        return sync();
    }
    template <class H>
    int down(H* dst, const T* src, size_t n) {
        std::vector<T> tmp(n);
        HIP_OK(hipMemcpyAsync(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, stream));
        if (int rc = sync()) return rc;
        for (size_t i = 0; i < n; ++i) dst[i] = (H)tmp[i];
        return UWS_OK;
    }
    template <class H>
    int set_vec(T* const dst[3], const H* src, int ncomp, size_t n) {
        std::vector<H> plane(n);
        for (int c = 0; c < 3; ++c) {
            for (size_t i = 0; i < n; ++i) plane[i] = src[i * ncomp + c];
            if (int rc = up(dst[c], plane.data(), n)) return rc;
        }
        return UWS_OK;
    }
    template <class H>
    int get_vec(H* dst, T* const src[3], int ncomp, size_t n) {
        std::vector<H> plane(n);
        for (int c = 0; c < 3; ++c) {
            if (int rc = down(plane.data(), src[c], n)) return rc;
            for (size_t i = 0; i < n; ++i) dst[i * ncomp + c] = plane[i];
        }
        if (ncomp == 4)
            for (size_t i = 0; i < n; ++i) dst[i * 4 + 3] = 0;
        return UWS_OK;
    }

    int set_flags_u8(const uint8_t* fl) override {
        HIP_OK(hipMemcpyAsync(f.flag, fl, (size_t)p.N, hipMemcpyHostToDevice, stream));
        launch<T>(KN_ZERO_SOLID_VEL, p, f, Roles{}, stream);
        if (int rc = check_launch()) return rc;
        rs.cur = -1;
        rs.pc_latest = -1;
        return sync();
    }
    int get_flags_u8(uint8_t* fl) override {
        HIP_OK(hipMemcpyAsync(fl, f.flag, (size_t)p.N, hipMemcpyDeviceToHost, stream));
        return sync();
    }
    int set_velocity_all(const float* src, int ncomp) override {
        for (int b = 0; b < 3; ++b)
            if (int rc = set_vec(f.vel[b], src, ncomp, p.N)) return rc;
        rs.cur = -1;
        return UWS_OK;
    }
    int get_velocity(float* dst, int ncomp) override { return get_vec(dst, f.vel[rs.state], ncomp, p.N); }
    T* plane_ptr(int which, int comp) {
        if (which == PLANE_P) return f.p;
        if (which >= PLANE_VEL0 && which < PLANE_VEL0 + 9) return f.vel[(which - PLANE_VEL0) / 3][(which - PLANE_VEL0) % 3];
        if (which == PLANE_UCORR) return f.ucorr[comp];
        if (which >= PLANE_BND0 && which < PLANE_BND0 + 6) return f.bnd[which - PLANE_BND0][comp];
        return nullptr;
    }
    int set_plane(int which, int comp, const float* src, size_t n) override {
        T* d = plane_ptr(which, comp);
        return d ? up(d, src, n) : UWS_ERR_INVALID_ARG;
    }
    int set_plane_d(int which, int comp, const double* src, size_t n) override {
        T* d = plane_ptr(which, comp);
        return d ? up(d, src, n) : UWS_ERR_INVALID_ARG;
    }
    int get_plane(int which, int comp, float* dst, size_t n) override {
        T* d = nullptr;
        if (which == PLANE_P) d = f.p;
        else if (which >= PLANE_BND0 && which < PLANE_BND0 + 6) d = f.bnd[which - PLANE_BND0][comp];
        if (!d) return UWS_ERR_INVALID_ARG;
        return down(dst, d, n);
    }

    // resolve a field id to device planes; returns ncomp, or an error
    int field_planes(int field, T* planes[3], bool for_write) {
        planes[0] = planes[1] = planes[2] = nullptr;
        switch (field) {
        case UWS_FIELD_VEL: for (int c = 0; c < 3; ++c) planes[c] = f.vel[rs.state][c]; return 3;
        case UWS_FIELD_VEL_WORK:
            if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
            for (int c = 0; c < 3; ++c) planes[c] = f.vel[rs.cur][c];
            return 3;
        case UWS_FIELD_VEL_CORR: for (int c = 0; c < 3; ++c) planes[c] = f.ucorr[c]; return 3;
        case UWS_FIELD_WALL_VIS: for (int c = 0; c < 3; ++c) planes[c] = f.nuw[c]; return 3;
        case UWS_FIELD_B: for (int c = 0; c < 3; ++c) planes[c] = f.b[c]; return 3;
        case UWS_FIELD_D: for (int c = 0; c < 3; ++c) planes[c] = f.D[c]; return 3;
        case UWS_FIELD_AODU: for (int c = 0; c < 3; ++c) planes[c] = f.aodu[c]; return 3;
        case UWS_FIELD_PRES: planes[0] = f.p; return 1;
        case UWS_FIELD_PRES_CORR:
            if (for_write) return UWS_ERR_INVALID_ARG;
            planes[0] = rs.pc_latest >= 0 ? f.pc[rs.pc_latest] : nullptr;  // nullptr = zero
            return 1;
        case UWS_FIELD_EDDY_VIS: planes[0] = f.nut; return 1;
        case UWS_FIELD_DP: planes[0] = f.Dp; return 1;
        case UWS_FIELD_BP: planes[0] = f.bp; return 1;
        case UWS_FIELD_FACE_VEL:
            if (!p.nonuniform || for_write) return UWS_ERR_INVALID_ARG;
            for (int c = 0; c < 3; ++c) planes[c] = f.flux[c];
            return 3;
        default: return UWS_ERR_INVALID_ARG;
        }
    }
    template <class H>
    int get_field_t(int field, H* dst, int64_t n) {
        T* planes[3];
        const int nc = field_planes(field, planes, false);
        if (nc < 0) return nc;
        if (n != (int64_t)nc * p.N) return UWS_ERR_INVALID_ARG;
        if (nc == 1) {
            if (!planes[0]) {
                for (int64_t i = 0; i < n; ++i) dst[i] = 0;
                return UWS_OK;
            }
            if (field == UWS_FIELD_BP && p.nonuniform) {
                // the stretched grid stores bp / (dv/den): unfold to bFieldPresCorrect
                std::vector<T> tmp((size_t)p.N);
                HIP_OK(hipMemcpyAsync(tmp.data(), planes[0], sizeof(T) * tmp.size(), hipMemcpyDeviceToHost, stream));
                if (int rc = sync()) return rc;
                const std::vector<T> k = dvden_host();
                for (int64_t i = 0; i < n; ++i) dst[i] = (H)(tmp[(size_t)i] * k[(size_t)i]);
                return UWS_OK;
            }
            return down(dst, planes[0], p.N);
        }
        return get_vec(dst, planes, 3, p.N);
    }
    int get_field(int field, float* dst, int64_t n) override { return get_field_t(field, dst, n); }
    int get_field64(int field, double* dst, int64_t n) override { return get_field_t(field, dst, n); }
    int set_field(int field, const float* src, int64_t n) override {
        if (field == UWS_FIELD_VEL) return n == 3LL * p.N ? set_velocity_all(src, 3) : UWS_ERR_INVALID_ARG;
        T* planes[3];
        const int nc = field_planes(field, planes, true);
        if (nc < 0) return nc;
        if (n != (int64_t)nc * p.N) return UWS_ERR_INVALID_ARG;
        if (field == UWS_FIELD_BP && p.nonuniform) {
            const std::vector<T> k = dvden_host();
            std::vector<T> tmp((size_t)p.N);
            for (int64_t i = 0; i < n; ++i) tmp[(size_t)i] = (T)src[i] / k[(size_t)i];
            HIP_OK(hipMemcpyAsync(planes[0], tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice, stream));
            return sync();
        }
        if (nc == 1) return up(planes[0], src, p.N);
        return set_vec(planes, src, 3, p.N);
    }
    int get_norms(double* out) override {
        HIP_OK(hipMemcpyAsync(h_norms, f.norms, sizeof(double) * UWS_NORM_COUNT, hipMemcpyDeviceToHost, stream));
        if (int rc = sync()) return rc;
        for (int i = 0; i < UWS_NORM_COUNT; ++i) out[i] = h_norms[i];
        return UWS_OK;
    }

    // ---- differentiation support ----
    // a captured step holds the parameters by value: drop the graphs
    int set_cs(double re, double im) override {
        if constexpr (std::is_same<T, CStep>::value) {
            p.cs = CStep(re, im);
        } else {
            if (im != 0.0) return UWS_ERR_UNSUPPORTED;
            p.cs = (T)re;
        }
        cfg.smagorinsky_constant = (float)re;
        if (int rc = sync()) return rc;
        drop_graphs();
        return UWS_OK;
    }
    int cstep_imag(bool set, int what, int face, double* buf, int64_t n) override {
        if constexpr (!std::is_same<T, CStep>::value) {
            (void)set; (void)what; (void)face; (void)buf; (void)n;
            return UWS_ERR_UNSUPPORTED;
        } else {
            if (!buf) return UWS_ERR_INVALID_ARG;
            if (what == UWS_CSTEP_CS) {
                if (n != 1) return UWS_ERR_INVALID_ARG;
                if (set) return set_cs(p.cs.re, buf[0]);
                buf[0] = p.cs.im;
                return UWS_OK;
            }
            std::vector<T*> planes;  // [buffer][comp]
            int nc = 0;
            size_t np = 0;
            if (what == UWS_CSTEP_VEL) {
                nc = 3; np = (size_t)p.N;
                for (int b = 0; b < 3; ++b) {
                    if (!set && b != rs.state) continue;
                    for (int c = 0; c < 3; ++c) planes.push_back(f.vel[b][c]);
                }
            } else if (what == UWS_CSTEP_PRES) {
                nc = 1; np = (size_t)p.N;
                planes.push_back(f.p);
            } else if (what == UWS_CSTEP_BND) {
                if (face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
                nc = 3; np = (size_t)bnd_w(p.X, p.Y, face) * bnd_h(p.Y, p.Z, face);
                for (int c = 0; c < 3; ++c) planes.push_back(f.bnd[face][c]);
            } else {
                return UWS_ERR_INVALID_ARG;
            }
            if (n != (int64_t)(nc * np)) return UWS_ERR_INVALID_ARG;
            std::vector<CStep> tmp(np);
            for (size_t k = 0; k < planes.size(); ++k) {
                const int c = (int)(k % nc);
                HIP_OK(hipMemcpyAsync(tmp.data(), planes[k], np * sizeof(CStep), hipMemcpyDeviceToHost, stream));
                if (int rc = sync()) return rc;
                if (set) {
                    for (size_t i = 0; i < np; ++i) tmp[i].im = buf[i * nc + c];
                    HIP_OK(hipMemcpyAsync(planes[k], tmp.data(), np * sizeof(CStep), hipMemcpyHostToDevice, stream));
                    if (int rc = sync()) return rc;
                } else {
                    for (size_t i = 0; i < np; ++i) buf[i * nc + c] = tmp[i].im;
                }
            }
            return UWS_OK;
        }
    }

    // ---- warm start from a background flow (InitVelPresFieldsFromBackGroundFlow) ----
    int init_from_background(const float fg_pos[3], float rot_deg, const uws_background& bgc, const float* vel,
                             int ncomp, const float* pres) override {
        using R = typename std::conditional<std::is_same<T, float>::value, float, double>::type;
        const int64_t nb = (int64_t)bgc.grid_res[0] * bgc.grid_res[1] * bgc.grid_res[2];
        void* tmp = nullptr;
        if (hipMalloc(&tmp, sizeof(T) * (size_t)nb * 4) != hipSuccess) return UWS_ERR_ALLOC;
        T* planes = static_cast<T*>(tmp);
        int rc = UWS_OK;
        {
            std::vector<float> plane((size_t)nb);
            for (int c = 0; c < 3 && rc == UWS_OK; ++c) {
                for (int64_t i = 0; i < nb; ++i) plane[i] = vel[i * ncomp + c];
                rc = up(planes + (size_t)c * nb, plane.data(), (size_t)nb);
            }
            if (rc == UWS_OK && pres) rc = up(planes + 3 * (size_t)nb, pres, (size_t)nb);
        }
        if (rc == UWS_OK) {
            BgField<T, R> bg{};
            bg.X = bgc.grid_res[0]; bg.Y = bgc.grid_res[1]; bg.Z = bgc.grid_res[2];
            bg.dx = (R)bgc.dx;
            for (int a = 0; a < 3; ++a) {
                bg.pos[a] = (R)bgc.phys_field_pos[a];
                bg.fgPos[a] = (R)fg_pos[a];
                bg.vel[a] = planes + (size_t)a * nb;
            }
            bg.pres = pres ? planes + 3 * (size_t)nb : nullptr;
            bg.vstride = 1;
            // the reference forms the angle in float: flowFieldOrientation * (float)(3.1415926 / 180)
            bg.angle = (R)(rot_deg * (float)(3.1415926 / 180));
            uws::init_from_background<T, R>(p, f, bg, stream);
            rc = check_launch();
            if (rc == UWS_OK) rc = sync();
        }
        (void)hipFree(tmp);
        rs.cur = -1;
        rs.pc_latest = -1;
        return rc;
    }
    // The same warm start with the caller's float arrays staged as they are
    // in the GPU's fixed staging array ([cell][ncomp] velocity, then the
    // pressure), converted to T where the kernel reads them: the same values
    // as the upload above, which converts on the host.  No device allocation.
    int init_from_background_noalloc(const float fg_pos[3], float rot_deg, const uws_background& bgc,
                                     const float* vel, int ncomp, const float* pres) override {
        using R = typename std::conditional<std::is_same<T, float>::value, float, double>::type;
        const size_t nb = (size_t)bgc.grid_res[0] * bgc.grid_res[1] * bgc.grid_res[2];
        const size_t vbytes = sizeof(float) * nb * (size_t)ncomp, pbytes = pres ? sizeof(float) * nb : 0;
        if (vbytes + pbytes > (size_t)UWS_BG_STAGING_MAX_BYTES) return UWS_ERR_TOO_LARGE;
        char* st = static_cast<char*>(bg_staging(device));
        if (!st) {
            if (int rc = bg_staging_setup()) return rc;
            st = static_cast<char*>(bg_staging(device));
            if (!st) return UWS_ERR_DEVICE;
        }
        HIP_OK(hipMemcpyAsync(st, vel, vbytes, hipMemcpyHostToDevice, stream));
        if (pres) HIP_OK(hipMemcpyAsync(st + vbytes, pres, pbytes, hipMemcpyHostToDevice, stream));
        BgField<T, R, float> bg{};
        bg.X = bgc.grid_res[0]; bg.Y = bgc.grid_res[1]; bg.Z = bgc.grid_res[2];
        bg.dx = (R)bgc.dx;
        for (int a = 0; a < 3; ++a) {
            bg.pos[a] = (R)bgc.phys_field_pos[a];
            bg.fgPos[a] = (R)fg_pos[a];
            bg.vel[a] = reinterpret_cast<const float*>(st) + a;
        }
        bg.vstride = ncomp;
        bg.pres = pres ? reinterpret_cast<const float*>(st + vbytes) : nullptr;
        bg.angle = (R)(rot_deg * (float)(3.1415926 / 180));
        uws::init_from_background<T, R, float>(p, f, bg, stream);
        int rc = check_launch();
        if (rc == UWS_OK) rc = sync();
        rs.cur = -1;
        rs.pc_latest = -1;
        return rc;
    }

    // ---- ChangePhysDomainSize: a new configuration of the same arithmetic ----
    // The fields are reallocated for the new grid and zeroed (as at creation);
    // where the grid is unchanged the flags and the boundary textures are kept,
    // as the reference keeps its textures.  The clock is kept.
    int reconfigure(const uws_config& c) override {
        if (c.precision != cfg.precision || c.grid_type != cfg.grid_type) return UWS_ERR_CONFIG;
        const bool same_grid = c.grid_res[0] == cfg.grid_res[0] && c.grid_res[1] == cfg.grid_res[1] &&
                               c.grid_res[2] == cfg.grid_res[2];
        std::vector<uint8_t> fl;
        std::vector<T> bnd[6][3];
        if (same_grid) {
            fl.resize((size_t)p.N);
            HIP_OK(hipMemcpyAsync(fl.data(), f.flag, fl.size(), hipMemcpyDeviceToHost, stream));
            for (int face = 0; face < 6; ++face)
                for (int k = 0; k < 3; ++k) {
                    bnd[face][k].resize((size_t)bnd_w(p.X, p.Y, face) * bnd_h(p.Y, p.Z, face));
                    HIP_OK(hipMemcpyAsync(bnd[face][k].data(), f.bnd[face][k], sizeof(T) * bnd[face][k].size(),
                                          hipMemcpyDeviceToHost, stream));
                }
        }
        if (int rc = sync()) return rc;
        drop_graphs();
        if (mean_sum[0]) (void)hipFree(mean_sum[0]);
        for (auto& m : mean_sum) m = nullptr;
        mean_count = 0;
        (void)hipFree(slab);
        slab = nullptr;
        cfg = c;
        sc = step_cfg(c);
        rs = RoleState{};
        fuse_mode = fusion_mode();
        for (auto& fp : face_pos) fp.clear();
        p = ParamsT<T>{};
        f = FieldsT<T>{};
        if (int rc = init(c)) return rc;
        if (same_grid) {
            for (int face = 0; face < 6; ++face)
                for (int k = 0; k < 3; ++k)
                    HIP_OK(hipMemcpyAsync(f.bnd[face][k], bnd[face][k].data(), sizeof(T) * bnd[face][k].size(),
                                          hipMemcpyHostToDevice, stream));
            return set_flags_u8(fl.data());
        }
        return sync();
    }

    // ---- time mean of the velocity ----
    int mean_reset() override {
        if (!mean_sum[0]) {
            void* m = nullptr;
            if (hipMalloc(&m, 3 * sizeof(double) * (size_t)p.N) != hipSuccess) return UWS_ERR_ALLOC;
            for (int c = 0; c < 3; ++c) mean_sum[c] = (double*)m + (size_t)c * p.N;
        }
        HIP_OK(hipMemsetAsync(mean_sum[0], 0, 3 * sizeof(double) * (size_t)p.N, stream));
        mean_count = 0;
        return UWS_OK;
    }
    int mean_accumulate() override {
        if (rs.cur >= 0) return UWS_ERR_INVALID_ARG;
        if (!mean_sum[0])
            if (int rc = mean_reset()) return rc;
        uws::mean_accumulate<T>(f.vel[rs.state], mean_sum, p.N, stream);
        mean_count += 1;
        return check_launch();
    }
    int get_mean(double* dst, int64_t n, int64_t* count) override {
        if (n != 3LL * p.N) return UWS_ERR_INVALID_ARG;
        if (count) *count = mean_count;
        std::vector<double> tmp((size_t)p.N);
        for (int c = 0; c < 3; ++c) {
            if (mean_sum[c]) {
                HIP_OK(hipMemcpyAsync(tmp.data(), mean_sum[c], sizeof(double) * p.N, hipMemcpyDeviceToHost, stream));
                if (int rc = sync()) return rc;
            } else {
                std::fill(tmp.begin(), tmp.end(), 0.0);
            }
            const double inv = mean_count > 0 ? 1.0 / (double)mean_count : 0.0;
            for (int i = 0; i < p.N; ++i) dst[(size_t)i * 3 + c] = tmp[i] * inv;
        }
        return UWS_OK;
    }

    // ---- stepping ----
    int run_stage(int stage) override {
        L l{this};
        int rc = uws::run_stage(l, sc, rs, stage);
        if (rc == UWS_OK) rc = check_launch();
        if (rc == UWS_OK && stage == UWS_STAGE_FINISH) { step += 1; phys_time += (double)cfg.dt; }
        return rc;
    }
    int enqueue_step() {
        L l{this};
        int rc = uws::enqueue_step(l, sc, rs);
        if (rc == UWS_OK) rc = check_launch();
        return rc;
    }
    int step_graph() {
        const int s0 = rs.state;
        if (!graph[s0]) {
            const RoleState saved = rs;
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_step();
            hipError_t e = hipStreamEndCapture(stream, &g);
            graph_after[s0] = rs;
            rs = saved;
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return UWS_ERR_DEVICE;
            e = hipGraphInstantiate(&graph[s0], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { graph[s0] = nullptr; return UWS_ERR_DEVICE; }
        }
        HIP_OK(hipGraphLaunch(graph[s0], stream));
        rs = graph_after[s0];
        return UWS_OK;
    }
    int step_n(int n) override {
        if (rs.cur >= 0) return UWS_ERR_INVALID_ARG;  // a staged step is in progress
        const bool graphable = !sc.cal_residual && std::getenv("UWS_NO_GRAPH") == nullptr;
        for (int i = 0; i < n; ++i) {
            int rc = graphable ? step_graph() : enqueue_step();
            if (rc) return rc;
            step += 1;
            phys_time += (double)cfg.dt;
        }
        return UWS_OK;
    }
    void launch_plan(int* pres, int* vel) override {
        const L l{this};
        *pres = l.fused4() ? 4 : l.fused3() ? 3 : l.fused() ? 2 : 1;
        *vel = l.fused_vel() ? 2 : 1;
    }
    int time_kernel(int kernel, int reps, double* avg_ms) override {
        if (rs.cur >= 0) return UWS_ERR_INVALID_ARG;
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return UWS_ERR_DEVICE; }
        const int w0 = (rs.state + 1) % 3;
        if ((kernel == UWS_KERNEL_VEL_SWEEP2 && !std::is_same<T, float>::value) ||
            ((kernel == UWS_KERNEL_PRES_SWEEP3 || kernel == UWS_KERNEL_PRES_SWEEP4) && !std::is_same<T, float>::value)) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return UWS_ERR_UNSUPPORTED;
        }
        auto go = [&](int i) {
            Roles r{};
            if (kernel == UWS_KERNEL_VEL_SWEEP2) {
                r.state = rs.state; r.vin = rs.state; r.vout = w0;
                launch<T>(KN_VEL_SWEEP2, p, f, r, stream);
            } else if (kernel == UWS_KERNEL_VEL_SWEEP) {
                r.state = rs.state; r.vin = rs.state; r.vout = w0;
                launch<T>(KN_VEL_SWEEP, p, f, r, stream);
            } else if (kernel == UWS_KERNEL_PRES_SWEEP3) {
                r.coef = 0; r.pin = i & 1; r.pout = (i + 1) & 1;
                launch<T>(KN_PRES_SWEEP3, p, f, r, stream);
            } else if (kernel == UWS_KERNEL_PRES_SWEEP4) {
                r.coef = 0; r.pin = i & 1; r.pout = (i + 1) & 1;
                launch<T>(KN_PRES_SWEEP4, p, f, r, stream);
            } else if (kernel == UWS_KERNEL_PRES_SWEEP2) {
                r.coef = 0; r.pin = i & 1; r.pout = (i + 1) & 1;
                launch<T>(KN_PRES_SWEEP2, p, f, r, stream);
            } else {
                r.coef = 0; r.pin = i & 1; r.pout = (i + 1) & 1;
                launch<T>(KN_PRES_SWEEP, p, f, r, stream);
            }
        };
        int rc = UWS_OK;
        go(0);
        if (hipEventRecord(e0, stream) != hipSuccess) rc = UWS_ERR_DEVICE;
        for (int i = 0; i < reps && rc == UWS_OK; ++i) go(i);
        if (rc == UWS_OK && check_launch()) rc = UWS_ERR_DEVICE;
        if (rc == UWS_OK && hipEventRecord(e1, stream) != hipSuccess) rc = UWS_ERR_DEVICE;
        if (rc == UWS_OK && hipEventSynchronize(e1) != hipSuccess) rc = UWS_ERR_DEVICE;
        float ms = 0.f;
        if (rc == UWS_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = UWS_ERR_DEVICE;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc == UWS_OK) *avg_ms = (double)ms / reps;
        return rc;
    }
};


}  // namespace uws
