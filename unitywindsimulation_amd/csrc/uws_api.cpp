// C-ABI implementation: handle lifetime, device memory, uploads/downloads,
// the step orchestration of FvmSolverGpu.Step() (FvmSolverGpu.cs:1050-1121)
// and its capture into HIP graphs.  The step driver (uws_host.h) is shared
// with the ensemble (uws_batch.cpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "uws_host.h"
#include "uws_solver.h"

using namespace uws;

namespace uws {

int validate(const uws_config* c, bool allow_residual_break) {
    if (!c) return UWS_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (c->grid_res[i] < 1) return UWS_ERR_CONFIG;
    const long long n = (long long)c->grid_res[0] * c->grid_res[1] * c->grid_res[2];
    if (n >= (1LL << 29)) return UWS_ERR_CONFIG;
    if (!(c->dx > 0.f) || !(c->dt > 0.f) || !(c->den > 0.f) || !(c->mu >= 0.f)) return UWS_ERR_CONFIG;
    for (int i = 0; i < 6; ++i)
        if (c->vel_bnd[i] < 0 || c->vel_bnd[i] > 2 || c->pres_bnd[i] < 0 || c->pres_bnd[i] > 2) return UWS_ERR_CONFIG;
    if (c->fvm_solver_type != UWS_SIMPLE && c->fvm_solver_type != UWS_PISO) return UWS_ERR_CONFIG;
    if (c->turbulence_model != UWS_TURB_NONE && c->turbulence_model != UWS_TURB_SMAGORINSKY) return UWS_ERR_CONFIG;
    if (c->piso_num_correctors < 0 || c->piso_num_correctors > 7 || c->vel_max_num_iter < 0 ||
        c->pres_max_num_iter < 0)
        return UWS_ERR_CONFIG;
    if (c->cal_residual) {
        if (!allow_residual_break) return UWS_ERR_UNSUPPORTED;
        if (c->vel_residual_check_interval <= 0 || c->pres_residual_check_interval <= 0) return UWS_ERR_CONFIG;
    }
    if (c->wall_slots != 1 && c->wall_slots != 3) return UWS_ERR_CONFIG;
    if (c->wall_model != UWS_WALL_SMOOTH && c->wall_model != UWS_WALL_ROUGH) return UWS_ERR_CONFIG;
    if (c->wall_model == UWS_WALL_ROUGH && !(c->wall_z0 > 0.f)) return UWS_ERR_CONFIG;
    if (c->precision < UWS_PREC_FP32 || c->precision > UWS_PREC_CSTEP) return UWS_ERR_CONFIG;
    if (c->precision == UWS_PREC_FP32_FP16 || c->precision == UWS_PREC_FP32_BF16) {
        if (c->refine_rounds < 0 || c->refine_rounds > 8 || c->refine_sweeps < 0) return UWS_ERR_CONFIG;
        if (c->cal_residual) return UWS_ERR_UNSUPPORTED;
    }
    if (c->compute_norms != 0 && c->compute_norms != 1) return UWS_ERR_CONFIG;
    if (c->grid_type == UWS_GRID_STAGGERED) return UWS_ERR_UNSUPPORTED;
    if (c->grid_type != UWS_GRID_COLLOCATED && c->grid_type != UWS_GRID_COLL_NONUNIFORM) return UWS_ERR_CONFIG;
    // the stretched grid has no PISO corrector (InitPisoShader throws, FvmSolverGpu.cs:515-519)
    if (c->grid_type == UWS_GRID_COLL_NONUNIFORM && c->fvm_solver_type == UWS_PISO) return UWS_ERR_UNSUPPORTED;
    for (int i = 0; i < 3; ++i)
        if (c->reserved[i] != 0) return UWS_ERR_CONFIG;
    return UWS_OK;
}

StepCfg step_cfg(const uws_config& c) {
    StepCfg s;
    s.V = c.vel_max_num_iter;
    s.P = c.pres_max_num_iter;
    s.C = c.fvm_solver_type == UWS_PISO ? c.piso_num_correctors : 0;
    s.les = c.turbulence_model == UWS_TURB_SMAGORINSKY;
    s.norms = c.compute_norms != 0;
    s.lowp = c.precision == UWS_PREC_FP32_FP16 || c.precision == UWS_PREC_FP32_BF16;
    s.bf16 = c.precision == UWS_PREC_FP32_BF16;
    s.R = c.refine_rounds;
    s.K = c.refine_sweeps;
    s.cal_residual = c.cal_residual != 0;
    s.vci = c.vel_residual_check_interval;
    s.pci = c.pres_residual_check_interval;
    s.vtol = c.vel_tolerance;
    s.ptol = c.pres_tolerance;
    s.piso = c.fvm_solver_type == UWS_PISO;
    bool anyFixed = false;
    for (int i = 0; i < 6; ++i) anyFixed |= c.pres_bnd[i] == UWS_PRES_FIXED_VALUE;
    s.normalise = !anyFixed;
    return s;
}

// z planes per two-sweep workgroup: each workgroup recomputes one halo plane
// above and below its chunk, so longer chunks waste less, as long as the
// grid still gives one workgroup per CU; a chunk that divides Z avoids a
// ragged last chunk.  Measured on 256 x 96 x 128 (k_pres_sweep2): zchunk 16
// (256 workgroups) 22.8 us, 8 (512) 24.2 us, 12 (352: a second, partial
// round of workgroups) and 24 (192: idle CUs) 29.5-30 us.
// tile rows of the two-sweep kernels: 15 (13 output rows) or 14 (12),
// whichever computes fewer thread rows over the grid's Y (ties: 15).  Every
// workgroup marches the same planes, so a tile row that falls outside the
// grid costs as much as one inside: 256 x 96 x 128 runs its momentum and
// pressure pairs in 8 tiles of 12 rows (112 thread rows) rather than 8 of 13
// (120), 2 % faster a step.
int choose_by2(int Y) {
    auto rows = [&](int by) { return ((Y + by - 3) / (by - 2)) * by; };
    return rows(14) < rows(15) ? 14 : 15;
}

// z planes per two-sweep workgroup: the longest of 32, 16, 8 and 4 that
// still gives a workgroup per CU, preferring one that divides Z (a longer
// chunk recomputes fewer halo planes per output plane: 384 x 128 x 256 runs
// 480 workgroups of 32 planes, the momentum pair in 261 against 275 us with
// 960 of 16, the step 1.8 % faster)
int choose_zchunk(int X, int Y, int Z, int by) {
    int best = 8, best_score = -1;
    for (int zc : {32, 16, 8, 4}) {
        const int blocks = pres_sweep2_grid(X, Y, Z, zc, by);
        if (blocks < FUSE_MIN_BLOCKS) continue;
        const int score = (Z % zc == 0 ? 4 : 0) + (zc == 32 ? 3 : zc == 16 ? 2 : zc == 8 ? 1 : 0);
        if (score > best_score) { best = zc; best_score = score; }
    }
    return best;
}

// z planes per three-sweep workgroup: each workgroup recomputes two halo
// planes above and below its chunk and runs alone on its CU (16 waves), so
// the chunk count is chosen to minimise rounds of 256 workgroups times the
// planes each marches
int choose_zchunk3(int X, int Y, int Z) {
    int best = Z, best_cost = -1;
    for (int nz = 1; nz <= Z; ++nz) {
        const int zc = (Z + nz - 1) / nz;
        const int blocks = pres_sweep3_grid(X, Y, Z, zc);
        const int cost = ((blocks + 255) / 256) * (zc + 4);
        if (best_cost < 0 || cost < best_cost) { best = zc; best_cost = cost; }
    }
    return best;
}

// z planes per four-sweep workgroup (three halo planes above and below its
// chunk), chosen as for the three-sweep kernel
int choose_zchunk4(int X, int Y, int Z) {
    int best = Z, best_cost = -1;
    for (int nz = 1; nz <= Z; ++nz) {
        const int zc = (Z + nz - 1) / nz;
        const int blocks = pres_sweep4_grid(X, Y, Z, zc);
        const int cost = ((blocks + 255) / 256) * (zc + 6);
        if (best_cost < 0 || cost < best_cost) { best = zc; best_cost = cost; }
    }
    return best;
}

template <class T>
void fill_params(const uws_config& c, ParamsT<T>& p) {
    p.X = c.grid_res[0]; p.Y = c.grid_res[1]; p.Z = c.grid_res[2];
    p.N = p.X * p.Y * p.Z;
    p.ty = 4;
    p.tiles = tiles_of(p.X, p.Y, p.Z, p.ty);
    // FluidSimConfig.Init (FluidSimConfig.cs:325-341), in the run's arithmetic
    p.dx = (T)c.dx;
    p.ds = p.dx * p.dx;
    p.dv = p.dx * p.dx * p.dx;
    p.dt = (T)c.dt;
    p.den = (T)c.den;
    p.nu = (T)c.mu / p.den;
    p.cs = (T)c.smagorinsky_constant;
    for (int i = 0; i < 3; ++i) p.fext[i] = (T)c.external_force[i];
    bool anyFixed = false;
    for (int i = 0; i < 6; ++i) {
        p.velBnd[i] = c.vel_bnd[i];
        p.presBnd[i] = c.pres_bnd[i];
        anyFixed |= c.pres_bnd[i] == UWS_PRES_FIXED_VALUE;
    }
    p.useLES = c.turbulence_model == UWS_TURB_SMAGORINSKY;
    p.wallSlots = c.wall_slots;
    p.wallModel = c.wall_model;
    p.wallZ0 = (T)c.wall_z0;
    // CSPresNormalization runs only without a Dirichlet pressure face
    // (FvmSolverGpu.cs:1425-1436); its cell (2, Y/2, 2) outside the grid reads 0
    p.normRef = (!anyFixed && 2 < p.X && 2 < p.Z) ? (2 * p.Y + p.Y / 2) * p.X + 2 : -1;
    // z planes per workgroup of the two-sweep kernels (UWS_ZCHUNK overrides, for tuning)
    p.by2 = choose_by2(p.Y);
    if (const char* e = std::getenv("UWS_TILE_ROWS2")) {
        const int v = std::atoi(e);
        if (v >= 3 && v <= 15) p.by2 = v;
    }
    p.zchunk = choose_zchunk(p.X, p.Y, p.Z, p.by2);
    if (const char* e = std::getenv("UWS_ZCHUNK")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 1024) p.zchunk = v;
    }
    p.zchunk3 = choose_zchunk3(p.X, p.Y, p.Z);
    if (const char* e = std::getenv("UWS_ZCHUNK3")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 1024) p.zchunk3 = v;
    }
    p.zchunk4 = choose_zchunk4(p.X, p.Y, p.Z);
    if (const char* e = std::getenv("UWS_ZCHUNK4")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 1024) p.zchunk4 = v;
    }
    p.nonuniform = c.grid_type == UWS_GRID_COLL_NONUNIFORM ? 1 : 0;
    // powers-of-two dx and den: products with exact reciprocals in the cell
    // kernels (UWS_NO_P2 keeps the divisions, for tests)
    p.p2 = 0;
    if constexpr (std::is_same_v<T, float> || std::is_same_v<T, double>) {
        auto pow2 = [](T v) {
            int e = 0;
            return v > (T)0 && std::isfinite(v) && std::frexp(v, &e) == (T)0.5;
        };
        if (pow2(p.dx) && pow2(p.den) && std::getenv("UWS_NO_P2") == nullptr) {
            p.p2 = 1;
            p.rdx = (T)1.0f / p.dx;
            p.r2dx = (T)1.0f / ((T)2.0f * p.dx);
            p.rhdx = (T)1.0f / ((T)0.5f * p.dx);
            p.rden = (T)1.0f / p.den;
            p.dv_dt = p.dv / p.dt;
            p.dvden = p.dv / p.den;
            p.inv_den = (T)1.0f / p.den;
            p.nuds_dx = p.nu * p.ds / p.dx;
        }
    }
}
template void fill_params<float>(const uws_config&, ParamsT<float>&);
template void fill_params<double>(const uws_config&, ParamsT<double>&);
template void fill_params<CStep>(const uws_config&, ParamsT<CStep>&);

int bnd_w(int X, int Y, int face) { return face < 2 ? Y : X; }
int bnd_h(int Y, int Z, int face) { return face < 4 ? Z : Y; }

template <class T>
size_t carve(const ParamsT<T>& p, FieldsT<T>& f, char* base, bool lowp) {
    size_t off = 0;
    const size_t N = (size_t)p.N;
    auto take = [&](size_t bytes) -> char* {
        off = (off + 255) & ~(size_t)255;
        char* r = base ? base + off : nullptr;
        off += bytes;
        return r;
    };
    auto fl = [&](size_t n) { return (T*)take(n * sizeof(T)); };
    f.flag = (uint8_t*)take(N);
    f.pcrec[0] = f.pcrec[1] = nullptr;
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) f.vel[b][c] = fl(N);
    for (int c = 0; c < 3; ++c) f.ucorr[c] = fl(N);
    f.p = fl(N);
    f.pc[0] = fl(N); f.pc[1] = fl(N);
    f.nut = fl(N);
    for (int c = 0; c < 9; ++c) f.grad[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.nuw[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.D[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.invD[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.b[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.S[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.G[c] = fl(N);
    for (int c = 0; c < 3; ++c) f.cp[c] = fl(N);
    f.cpisoE = fl(N);
    f.Dp = fl(N); f.invDp = fl(N); f.bp = fl(N);
    for (int c = 0; c < 3; ++c) f.aodu[c] = fl(N);
    f.invDpiso = fl(N); f.bpiso = fl(N);
    for (int face = 0; face < 6; ++face)
        for (int c = 0; c < 3; ++c) f.bnd[face][c] = fl((size_t)bnd_w(p.X, p.Y, face) * bnd_h(p.Y, p.Z, face));
    if (p.nonuniform) {
        f.h[0] = fl(p.X); f.h[1] = fl(p.Y); f.h[2] = fl(p.Z);
        for (int c = 0; c < 3; ++c) f.flux[c] = fl(N);
        f.fluxB[0] = fl((size_t)p.Y * p.Z);
        f.fluxB[1] = fl((size_t)p.X * p.Z);
        f.fluxB[2] = fl((size_t)p.X * p.Y);
    } else {
        for (int c = 0; c < 3; ++c) f.h[c] = f.flux[c] = f.fluxB[c] = nullptr;
    }
    if (lowp) {
        f.rho = fl(N);
        f.e16[0] = take(2 * N);
        f.e16[1] = take(2 * N);
    } else {
        f.rho = nullptr;
        f.e16[0] = f.e16[1] = nullptr;
    }
    f.partial = (double*)take((size_t)tiles_of(p.X, p.Y, p.Z, 1) * sizeof(double));  // room for any ty
    f.norms = (double*)take(UWS_NORM_COUNT * sizeof(double));
    f.scalar = fl(64);
    return off + 256;
}
template size_t carve<float>(const ParamsT<float>&, FieldsT<float>&, char*, bool);
template size_t carve<double>(const ParamsT<double>&, FieldsT<double>&, char*, bool);
template size_t carve<CStep>(const ParamsT<CStep>&, FieldsT<CStep>&, char*, bool);
template size_t carve<Dual<float>>(const ParamsT<Dual<float>>&, FieldsT<Dual<float>>&, char*, bool);
template size_t carve<Dual<double>>(const ParamsT<Dual<double>>&, FieldsT<Dual<double>>&, char*, bool);

}  // namespace uws

namespace {

template <class F>
int with_solver(uws_handle h, F&& fn) {
    if (!h || !h->impl) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(h->impl->device) != hipSuccess) return UWS_ERR_DEVICE;
    return fn(*h->impl);
}

// SplitMix64 (Steele, Lea, Flood 2014), the documented seeded stream
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// uniform in [0, 1) from the top 24 bits
float u01(uint64_t& s) { return (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f); }

}  // namespace

namespace uws {

int SolverBase::sync() { return hipStreamSynchronize(stream) == hipSuccess ? UWS_OK : UWS_ERR_DEVICE; }

SolverBase::~SolverBase() {
    if (stream && owns_stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    if (slab) (void)hipFree(slab);
    if (h_norms) (void)hipHostFree(h_norms);
}

int create_solver(const uws_config& cfg, bool allow_residual_break, hipStream_t shared_stream,
                  std::unique_ptr<SolverBase>* out) {
    if (int rc = validate(&cfg, allow_residual_break)) return rc;
    std::unique_ptr<SolverBase> s;
    if (cfg.precision == UWS_PREC_FP64) s.reset(new (std::nothrow) Solver<double>());
    else if (cfg.precision == UWS_PREC_CSTEP) s.reset(new (std::nothrow) Solver<CStep>());
    else s.reset(new (std::nothrow) Solver<float>());
    if (!s) return UWS_ERR_ALLOC;
    s->cfg = cfg;
    s->sc = step_cfg(cfg);
    if (hipGetDevice(&s->device) != hipSuccess) return UWS_ERR_DEVICE;
    // the staging array of the allocation-free warm start: its address on this device
    if (int rc = bg_staging_setup()) return rc;
    if (shared_stream) {
        s->stream = shared_stream;
        s->owns_stream = false;
    } else if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->stream = nullptr;
        return UWS_ERR_DEVICE;
    }
    if (hipHostMalloc(&s->h_norms, sizeof(double) * UWS_NORM_COUNT, hipHostMallocDefault) != hipSuccess) {
        s->h_norms = nullptr;
        return UWS_ERR_ALLOC;
    }
    int rc = cfg.precision == UWS_PREC_FP64    ? static_cast<Solver<double>*>(s.get())->init(cfg)
             : cfg.precision == UWS_PREC_CSTEP ? static_cast<Solver<CStep>*>(s.get())->init(cfg)
                                               : static_cast<Solver<float>*>(s.get())->init(cfg);
    if (rc) return rc;
    *out = std::move(s);
    return UWS_OK;
}

// host-side inflow profiles shared by the single and ensemble entry points,
// evaluated in the run's arithmetic T with the reference's float constants
// cell-centre height of row y: (y + 1/2) dx, or on a stretched grid the face
// midpoint 0.5 * (facePosY[y] + facePosY[y + 1]) (InitShaderColl.compute:141, 295)
template <class T>
T centre_height(const float* face_y, T dx, int y) {
    if (face_y) return (T)0.5f * ((T)face_y[y] + (T)face_y[y + 1]);
    return ((T)y + (T)0.5f) * dx;
}

template <class T>
int inflow_column(int inflow_type, float wind_speed, float dx_f, int Y, float ref_height, float alpha,
                  std::vector<T>& col, const float* face_y) {
    col.assign((size_t)Y, (T)0);
    const T wind = (T)wind_speed, dx = (T)dx_f;
    if (inflow_type == UWS_INFLOW_CONSTANT) {
        for (int y = 0; y < Y; ++y) col[y] = wind;
    } else if (inflow_type == UWS_INFLOW_LOG) {
        // CSSetLogVelBndCond (InitShaderColl.compute:272-285): y = (j + 1/2) dx;
        // CSSetLogVelBndCondNonUniform (:287-300): the face midpoint
        const T karman = (T)0.4f, rough = (T)1.0f, refH = (T)10.0f;
        for (int y = 0; y < Y; ++y) {
            const T yPhys = centre_height<T>(face_y, dx, y);
            const T fricVel = karman * wind / std::log((refH + rough) / rough);
            col[y] = fricVel / karman * std::log((yPhys + rough) / rough);
        }
    } else if (inflow_type == UWS_INFLOW_POWER_LAW) {
        if (!(ref_height > 0.f)) return UWS_ERR_INVALID_ARG;
        for (int y = 0; y < Y; ++y) {
            const T yPhys = centre_height<T>(face_y, dx, y);
            col[y] = wind * (T)std::pow((double)(yPhys / (T)ref_height), (double)alpha);
        }
    } else {
        return UWS_ERR_INVALID_ARG;
    }
    return UWS_OK;
}

template <class T>
int set_plane_t(SolverBase& s, int which, int comp, const T* src, size_t n);
template <>
int set_plane_t<float>(SolverBase& s, int which, int comp, const float* src, size_t n) {
    return s.set_plane(which, comp, src, n);
}
template <>
int set_plane_t<double>(SolverBase& s, int which, int comp, const double* src, size_t n) {
    return s.set_plane_d(which, comp, src, n);
}

template <class T>
int apply_fixed_value_bnd_t(SolverBase& s, int32_t inflow_type, float wind_speed, const float vel_zn[3],
                            float ref_height, float alpha) {
    const int X = s.dims(0), Y = s.dims(1), Z = s.dims(2);
    std::vector<T> col;
    const float* face_y = s.nonuniform() ? s.face_pos[1].data() : nullptr;
    if (int rc = inflow_column<T>(inflow_type, wind_speed, s.cfg.dx, Y, ref_height, alpha, col, face_y)) return rc;
    const size_t nx0 = (size_t)Y * Z;
    std::vector<T> a(nx0), zero(nx0, (T)0);
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y) a[(size_t)z * Y + y] = col[y];
    if (int rc = set_plane_t<T>(s, PLANE_BND0 + UWS_X0, 0, a.data(), nx0)) return rc;
    if (int rc = set_plane_t<T>(s, PLANE_BND0 + UWS_X0, 1, zero.data(), nx0)) return rc;
    if (int rc = set_plane_t<T>(s, PLANE_BND0 + UWS_X0, 2, zero.data(), nx0)) return rc;
    if (inflow_type == UWS_INFLOW_CONSTANT) {
        // CSSetFixedValueVelBndCond also writes Zn = velZn (InitShaderColl.compute:261-264)
        const size_t nzn = (size_t)X * Y;
        for (int c = 0; c < 3; ++c) {
            std::vector<T> v(nzn, (T)(vel_zn ? vel_zn[c] : 0.f));
            if (int rc = set_plane_t<T>(s, PLANE_BND0 + UWS_ZN, c, v.data(), nzn)) return rc;
        }
    }
    return UWS_OK;
}

int apply_fixed_value_bnd(SolverBase& s, int32_t inflow_type, float wind_speed, const float vel_zn[3],
                          float ref_height, float alpha) {
    if (wide_host(s.cfg.precision))
        return apply_fixed_value_bnd_t<double>(s, inflow_type, wind_speed, vel_zn, ref_height, alpha);
    return apply_fixed_value_bnd_t<float>(s, inflow_type, wind_speed, vel_zn, ref_height, alpha);
}

int apply_noise(SolverBase& s, uint64_t seed, float amp) {
    const int64_t n = s.ncells();
    std::vector<uint8_t> fl((size_t)n);
    if (int rc = s.get_flags_u8(fl.data())) return rc;
    std::vector<float> u((size_t)n * 3, 0.f);
    uint64_t st = seed;
    // one draw per component of every cell, in cell order, solid cells included
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float r = (2.0f * u01(st) - 1.0f) * amp;
            u[i * 3 + c] = fl[i] == UWS_CELL_FLUID ? r : 0.f;
        }
    if (int rc = s.set_velocity_all(u.data(), 3)) return rc;
    std::vector<float> zero((size_t)n, 0.f);
    for (int c = 0; c < 3; ++c)
        if (int rc = s.set_plane(PLANE_UCORR, c, zero.data(), n)) return rc;
    if (int rc = s.set_plane(PLANE_P, 0, zero.data(), n)) return rc;
    s.rs.pc_latest = -1;
    return UWS_OK;
}

template <class T>
int init_vel_pres_t(SolverBase& s, int32_t inflow_type, float wind_speed) {
    const int X = s.dims(0), Y = s.dims(1), Z = s.dims(2);
    const size_t n = (size_t)s.ncells();
    const T wind = (T)wind_speed, dx = (T)s.cfg.dx;
    std::vector<T> col((size_t)Y);
    for (int y = 0; y < Y; ++y) {
        if (inflow_type == UWS_INFLOW_LOG) {
            // CSInitLogVelPresFields (InitShaderColl.compute:112-132): y = j*dx;
            // CSInitLogVelPresFieldsNonUniform (:134-154): the face midpoint
            const T karman = (T)0.4f, rough = (T)1.0f, refH = (T)10.0f;
            const T yPhys = s.nonuniform() ? centre_height<T>(s.face_pos[1].data(), dx, y) : (T)y * dx;
            const T fricVel = karman * wind / std::log((refH + rough) / rough);
            col[y] = fricVel / karman * std::log((yPhys + rough) / rough);
        } else {
            col[y] = wind;
        }
    }
    std::vector<T> ux(n), zero(n, (T)0);
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) ux[((size_t)z * Y + y) * X + x] = col[y];
    // every velocity buffer gets the same field, as CSInitVelPresFields writes all three textures
    for (int b = 0; b < 3; ++b) {
        if (int rc = set_plane_t<T>(s, PLANE_VEL0 + 3 * b, 0, ux.data(), n)) return rc;
        if (int rc = set_plane_t<T>(s, PLANE_VEL0 + 3 * b + 1, 0, zero.data(), n)) return rc;
        if (int rc = set_plane_t<T>(s, PLANE_VEL0 + 3 * b + 2, 0, zero.data(), n)) return rc;
    }
    for (int c = 0; c < 3; ++c)
        if (int rc = set_plane_t<T>(s, PLANE_UCORR, c, zero.data(), n)) return rc;
    if (int rc = set_plane_t<T>(s, PLANE_P, 0, zero.data(), n)) return rc;
    s.rs.cur = -1;
    s.rs.pc_latest = -1;
    return UWS_OK;
}

int init_vel_pres(SolverBase& s, int32_t inflow_type, float wind_speed) {
    if (wide_host(s.cfg.precision)) return init_vel_pres_t<double>(s, inflow_type, wind_speed);
    return init_vel_pres_t<float>(s, inflow_type, wind_speed);
}

int flags_from_int(SolverBase& s, const int32_t* flags, int64_t n) {
    if (n != s.ncells()) return UWS_ERR_INVALID_ARG;
    std::vector<uint8_t> fl((size_t)n);
    for (int64_t i = 0; i < n; ++i) fl[i] = (double)flags[i] > 0.5 ? UWS_CELL_SOLID : UWS_CELL_FLUID;
    return s.set_flags_u8(fl.data());
}

// InitMeshBuffers (FvmSolverGpu.cs:682-788) for one axis
int stretched_face_positions(int n, float dx, int uniform, int one_side, float ratio, float* out) {
    if (n < 1 || uniform < 0 || one_side < 0 || !(dx > 0.f) || !(ratio > 0.f) || !out) return UWS_ERR_INVALID_ARG;
    const int trailing = n - uniform - one_side;
    if (trailing < 0 || trailing > one_side) return UWS_ERR_CONFIG;
    // cell sizes away from the uniform block: dx, dx*ratio, dx*ratio^2, ... (float products, as the C# loop)
    std::vector<float> grow((size_t)one_side);
    if (one_side > 0) {
        grow[0] = dx;
        for (int i = 1; i < one_side; ++i) grow[i] = grow[i - 1] * ratio;
    }
    out[0] = 0.f;
    int i = 1;
    for (; i < one_side + 1; ++i) out[i] = out[i - 1] + grow[(size_t)(one_side - i)];
    for (; i < one_side + uniform + 1; ++i) out[i] = out[i - 1] + dx;
    for (; i < n + 1; ++i) out[i] = out[i - 1] + grow[(size_t)(i - (one_side + uniform + 1))];
    return UWS_OK;
}

int normalise_norms(const double* raw, double ncells, double* out, int n) {
    double v[UWS_NORM_COUNT];
    v[UWS_NORM_VEL_RESIDUAL] = std::sqrt(raw[UWS_NORM_VEL_RESIDUAL] / (3.0 * ncells));
    v[UWS_NORM_PRES_RESIDUAL] = std::sqrt(raw[UWS_NORM_PRES_RESIDUAL] / ncells);
    v[UWS_NORM_DIVERGENCE] = std::sqrt(raw[UWS_NORM_DIVERGENCE] / ncells);
    v[UWS_NORM_NONFINITE] = raw[UWS_NORM_NONFINITE];
    for (int i = UWS_NORM_REFINE0; i < UWS_NORM_COUNT; ++i) v[i] = std::sqrt(raw[i] / ncells);
    for (int i = 0; i < n; ++i) out[i] = i < UWS_NORM_COUNT ? v[i] : 0.0;
    return UWS_OK;
}

}  // namespace uws

extern "C" {

int uws_abi_version(void) { return UWS_ABI_VERSION; }

const char* uws_strerror(int status) {
    switch (status) {
    case UWS_OK: return "no error";
    case UWS_ERR_INVALID_ARG: return "invalid argument";
    case UWS_ERR_CONFIG: return "invalid configuration";
    case UWS_ERR_ALLOC: return "device allocation failed";
    case UWS_ERR_DEVICE: return "HIP runtime error";
    case UWS_ERR_UNSUPPORTED: return "unsupported combination";
    case UWS_ERR_NONFINITE: return "non-finite value detected";
    case UWS_ERR_TOO_LARGE: return "input exceeds the fixed staging array of an allocation-free entry point";
    default: return "unknown status";
    }
}

int uws_config_default(uws_config* c) {
    if (!c) return UWS_ERR_INVALID_ARG;
    std::memset(c, 0, sizeof(*c));
    // FluidSimConfig field defaults (FluidSimConfig.cs:113-178)
    c->grid_res[0] = 100; c->grid_res[1] = 50; c->grid_res[2] = 100;  // physDomainSize (10,5,10), gridResX 100
    c->dx = 10.0f / 100;
    c->dt = 0.1f;
    c->mu = 0.005f;
    c->den = 1.0f;
    for (int i = 0; i < 6; ++i) { c->vel_bnd[i] = UWS_VEL_FIXED_VALUE; c->pres_bnd[i] = UWS_PRES_ZERO_GRAD; }
    c->fvm_solver_type = UWS_SIMPLE;
    c->piso_num_correctors = 2;
    c->turbulence_model = UWS_TURB_NONE;
    c->smagorinsky_constant = 0.15f;
    c->cal_residual = 0;
    c->vel_residual_check_interval = 10;
    c->vel_tolerance = 1e-15f;
    c->vel_max_num_iter = 50;
    c->pres_residual_check_interval = 20;
    c->pres_tolerance = 1e-3f;
    c->pres_max_num_iter = 400;
    c->wall_slots = 1;
    c->wall_model = UWS_WALL_SMOOTH;
    c->wall_z0 = 0.01f;
    c->precision = UWS_PREC_FP32;
    c->refine_rounds = 3;
    c->refine_sweeps = 20;
    c->compute_norms = 0;
    return UWS_OK;
}

int uws_create(const uws_config* cfg, uws_handle* out) {
    if (!out) return UWS_ERR_INVALID_ARG;
    *out = nullptr;
    if (int rc = validate(cfg, true)) return rc;
    uws_solver* h = new (std::nothrow) uws_solver();
    if (!h) return UWS_ERR_ALLOC;
    if (int rc = create_solver(*cfg, true, nullptr, &h->impl)) {
        delete h;
        return rc;
    }
    *out = h;
    return UWS_OK;
}

int uws_change_phys_domain_size(uws_handle h, const uws_config* cfg) {
    if (!cfg) return UWS_ERR_INVALID_ARG;
    if (int rc = validate(cfg, true)) return rc;
    return with_solver(h, [&](SolverBase& s) { return s.reconfigure(*cfg); });
}

int uws_destroy(uws_handle h) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (h->impl) (void)hipSetDevice(h->impl->device);
    delete h;
    return UWS_OK;
}

int uws_get_config(uws_handle h, uws_config* out) {
    if (!out) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { *out = s.cfg; return (int)UWS_OK; });
}

int uws_num_cells(uws_handle h, int64_t* n) {
    if (!n) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { *n = s.ncells(); return (int)UWS_OK; });
}

int uws_init_flags(uws_handle h, const int32_t* flags, int64_t n) {
    if (!flags) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return flags_from_int(s, flags, n); });
}

int uws_init_box(uws_handle h) {
    return with_solver(h, [&](SolverBase& s) {
        const int X = s.dims(0), Y = s.dims(1), Z = s.dims(2);
        std::vector<uint8_t> fl((size_t)s.ncells());
        if (int rc = s.get_flags_u8(fl.data())) return rc;
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x)
                    if (x >= 20 && x < 24 && y >= 0 && y < 8 && z >= 18 && z < 22)
                        fl[((size_t)z * Y + y) * X + x] = UWS_CELL_SOLID;
        return s.set_flags_u8(fl.data());
    });
}

int uws_set_face_positions(uws_handle h, const float* pos_x, const float* pos_y, const float* pos_z) {
    if (!pos_x || !pos_y || !pos_z) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        const float* const pos[3] = {pos_x, pos_y, pos_z};
        return s.set_face_positions(pos);
    });
}

int uws_get_face_positions(uws_handle h, float* pos_x, float* pos_y, float* pos_z) {
    if (!pos_x || !pos_y || !pos_z) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (!s.nonuniform()) return (int)UWS_ERR_CONFIG;
        float* const dst[3] = {pos_x, pos_y, pos_z};
        for (int a = 0; a < 3; ++a) std::copy(s.face_pos[a].begin(), s.face_pos[a].end(), dst[a]);
        return (int)UWS_OK;
    });
}

int uws_stretched_face_positions(int32_t n, float dx, int32_t uniform, int32_t one_side, float ratio, float* out) {
    return stretched_face_positions(n, dx, uniform, one_side, ratio, out);
}

int uws_init_vel_pres_fields(uws_handle h, int32_t inflow_type, float wind_speed) {
    if (inflow_type != UWS_INFLOW_CONSTANT && inflow_type != UWS_INFLOW_LOG) return UWS_ERR_UNSUPPORTED;
    return with_solver(h, [&](SolverBase& s) { return init_vel_pres(s, inflow_type, wind_speed); });
}

int uws_set_fixed_value_vel_bnd_cond(uws_handle h, int32_t inflow_type, float wind_speed, const float vel_zn[3],
                                     float ref_height, float alpha) {
    return with_solver(h, [&](SolverBase& s) {
        return apply_fixed_value_bnd(s, inflow_type, wind_speed, vel_zn, ref_height, alpha);
    });
}

int uws_set_bnd_field(uws_handle h, int32_t face, const float* xyz, int64_t n) {
    if (!xyz || face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != (int64_t)bnd_w(s.dims(0), s.dims(1), face) * bnd_h(s.dims(1), s.dims(2), face))
            return (int)UWS_ERR_INVALID_ARG;
        std::vector<float> plane((size_t)n);
        for (int c = 0; c < 3; ++c) {
            for (int64_t i = 0; i < n; ++i) plane[i] = xyz[i * 3 + c];
            if (int rc = s.set_plane(PLANE_BND0 + face, c, plane.data(), n)) return rc;
        }
        return (int)UWS_OK;
    });
}

int uws_set_bnd_field_f64(uws_handle h, int32_t face, const double* xyz, int64_t n) {
    if (!xyz || face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != (int64_t)bnd_w(s.dims(0), s.dims(1), face) * bnd_h(s.dims(1), s.dims(2), face))
            return (int)UWS_ERR_INVALID_ARG;
        std::vector<double> plane((size_t)n);
        for (int c = 0; c < 3; ++c) {
            for (int64_t i = 0; i < n; ++i) plane[i] = xyz[i * 3 + c];
            if (int rc = s.set_plane_d(PLANE_BND0 + face, c, plane.data(), n)) return rc;
        }
        return (int)UWS_OK;
    });
}

int uws_init_vel_pres_fields_from_background(uws_handle h, const float phys_field_pos[3], float field_rot_deg,
                                             const uws_background* bg, const float* bg_vel, int32_t ncomp,
                                             const float* bg_pres) {
    if (!phys_field_pos || !bg || !bg_vel || (ncomp != 3 && ncomp != 4) || !std::isfinite(field_rot_deg))
        return UWS_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (bg->grid_res[i] < 1 || !std::isfinite(bg->phys_field_pos[i]) || !std::isfinite(phys_field_pos[i]))
            return UWS_ERR_INVALID_ARG;
    if (!(bg->dx > 0.f) || (long long)bg->grid_res[0] * bg->grid_res[1] * bg->grid_res[2] >= (1LL << 29))
        return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        return s.init_from_background(phys_field_pos, field_rot_deg, *bg, bg_vel, ncomp, bg_pres);
    });
}

int uws_init_vel_pres_fields_from_background_noalloc(uws_handle h, const float phys_field_pos[3],
                                                     float field_rot_deg, const uws_background* bg,
                                                     const float* bg_vel, int32_t ncomp, const float* bg_pres) {
    if (!phys_field_pos || !bg || !bg_vel || (ncomp != 3 && ncomp != 4) || !std::isfinite(field_rot_deg))
        return UWS_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (bg->grid_res[i] < 1 || !std::isfinite(bg->phys_field_pos[i]) || !std::isfinite(phys_field_pos[i]))
            return UWS_ERR_INVALID_ARG;
    if (!(bg->dx > 0.f)) return UWS_ERR_INVALID_ARG;
    const long long nb = (long long)bg->grid_res[0] * bg->grid_res[1] * bg->grid_res[2];
    if (nb * (ncomp + (bg_pres ? 1 : 0)) * (long long)sizeof(float) > (long long)UWS_BG_STAGING_MAX_BYTES)
        return UWS_ERR_TOO_LARGE;
    return with_solver(h, [&](SolverBase& s) {
        return s.init_from_background_noalloc(phys_field_pos, field_rot_deg, *bg, bg_vel, ncomp, bg_pres);
    });
}

int uws_mean_reset(uws_handle h) {
    return with_solver(h, [&](SolverBase& s) { return s.mean_reset(); });
}

int uws_mean_accumulate(uws_handle h) {
    return with_solver(h, [&](SolverBase& s) { return s.mean_accumulate(); });
}

int uws_get_mean_velocity(uws_handle h, double* dst, int64_t n, int64_t* count) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.get_mean(dst, n, count); });
}

int uws_set_smagorinsky_constant(uws_handle h, double cs) {
    if (!std::isfinite(cs) || cs < 0.0) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.set_cs(cs, 0.0); });
}

int uws_cstep_set_imag(uws_handle h, int32_t what, int32_t face, const double* im, int64_t n) {
    return with_solver(h, [&](SolverBase& s) { return s.cstep_imag(true, what, face, const_cast<double*>(im), n); });
}

int uws_cstep_get_imag(uws_handle h, int32_t what, int32_t face, double* im, int64_t n) {
    return with_solver(h, [&](SolverBase& s) { return s.cstep_imag(false, what, face, im, n); });
}

int uws_get_bnd_field(uws_handle h, int32_t face, float* xyz, int64_t n) {
    if (!xyz || face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != (int64_t)bnd_w(s.dims(0), s.dims(1), face) * bnd_h(s.dims(1), s.dims(2), face))
            return (int)UWS_ERR_INVALID_ARG;
        std::vector<float> plane((size_t)n);
        for (int c = 0; c < 3; ++c) {
            if (int rc = s.get_plane(PLANE_BND0 + face, c, plane.data(), n)) return rc;
            for (int64_t i = 0; i < n; ++i) xyz[i * 3 + c] = plane[i];
        }
        return (int)UWS_OK;
    });
}

int uws_set_velocity(uws_handle h, const float* src, int32_t ncomp, int64_t n) {
    if (!src || (ncomp != 3 && ncomp != 4)) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.set_velocity_all(src, ncomp);
    });
}

int uws_get_velocity(uws_handle h, float* dst, int32_t ncomp, int64_t n) {
    if (!dst || (ncomp != 3 && ncomp != 4)) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.get_velocity(dst, ncomp);
    });
}

int uws_set_pressure(uws_handle h, const float* src, int64_t n) {
    if (!src) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.set_plane(PLANE_P, 0, src, n);
    });
}

int uws_get_pressure(uws_handle h, float* dst, int64_t n) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        return s.get_plane(PLANE_P, 0, dst, n);
    });
}

int uws_get_flags(uws_handle h, int32_t* dst, int64_t n) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (n != s.ncells()) return (int)UWS_ERR_INVALID_ARG;
        std::vector<uint8_t> fl((size_t)n);
        if (int rc = s.get_flags_u8(fl.data())) return rc;
        for (int64_t i = 0; i < n; ++i) dst[i] = fl[i];
        return (int)UWS_OK;
    });
}

int uws_init_noise(uws_handle h, uint64_t seed, float amp) {
    if (!(amp >= 0.f)) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return apply_noise(s, seed, amp); });
}

int uws_step(uws_handle h, int32_t nsteps) {
    if (nsteps < 0) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.step_n(nsteps); });
}

int uws_synchronize(uws_handle h) {
    return with_solver(h, [&](SolverBase& s) { return s.sync(); });
}

int uws_get_time(uws_handle h, int64_t* step, double* phys_time) {
    return with_solver(h, [&](SolverBase& s) {
        if (step) *step = s.step;
        if (phys_time) *phys_time = s.phys_time;
        return (int)UWS_OK;
    });
}

int uws_field_ncomp(int32_t field) {
    switch (field) {
    case UWS_FIELD_VEL: case UWS_FIELD_VEL_WORK: case UWS_FIELD_VEL_CORR: case UWS_FIELD_WALL_VIS:
    case UWS_FIELD_B: case UWS_FIELD_D: case UWS_FIELD_AODU: case UWS_FIELD_FACE_VEL:
        return 3;
    case UWS_FIELD_PRES: case UWS_FIELD_PRES_CORR: case UWS_FIELD_EDDY_VIS: case UWS_FIELD_DP: case UWS_FIELD_BP:
        return 1;
    default: return UWS_ERR_INVALID_ARG;
    }
}

int uws_get_field(uws_handle h, int32_t field, float* dst, int64_t n_floats) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.get_field(field, dst, n_floats); });
}

int uws_get_field_f64(uws_handle h, int32_t field, double* dst, int64_t n_floats) {
    if (!dst) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.get_field64(field, dst, n_floats); });
}

int uws_set_field(uws_handle h, int32_t field, const float* src, int64_t n_floats) {
    if (!src) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.set_field(field, src, n_floats); });
}

int uws_run_stage(uws_handle h, int32_t stage) {
    if (stage < 0 || stage >= UWS_STAGE_COUNT) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        if (int rc = s.run_stage(stage)) return rc;
        return s.sync();
    });
}

int uws_get_norms(uws_handle h, double* out, int32_t n) {
    if (!out || n < 0) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        double raw[UWS_NORM_COUNT];
        if (int rc = s.get_norms(raw)) return rc;
        if (int rc = normalise_norms(raw, (double)s.ncells(), out, n)) return rc;
        return raw[UWS_NORM_NONFINITE] > 0.0 ? (int)UWS_ERR_NONFINITE : (int)UWS_OK;
    });
}

int uws_stream_bandwidth(int64_t cells, int32_t nin, int32_t reps, double* gbps) {
    if (!gbps) return UWS_ERR_INVALID_ARG;
    return stream_bandwidth(cells, nin, reps, gbps);
}

int uws_get_sweep_counts(uws_handle h, int32_t* vel_sweeps, int32_t* pres_sweeps) {
    if (!vel_sweeps || !pres_sweeps) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        *vel_sweeps = s.rs.vel_sweeps;
        *pres_sweeps = s.rs.pres_sweeps;
        return (int)UWS_OK;
    });
}

int uws_get_launch_plan(uws_handle h, int32_t* pres_sweeps_per_launch, int32_t* vel_sweeps_per_launch) {
    if (!pres_sweeps_per_launch || !vel_sweeps_per_launch) return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) {
        int pr = 1, ve = 1;
        s.launch_plan(&pr, &ve);
        *pres_sweeps_per_launch = pr;
        *vel_sweeps_per_launch = ve;
        return (int)UWS_OK;
    });
}

int uws_time_kernel(uws_handle h, int32_t kernel, int32_t reps, double* avg_ms) {
    if (!avg_ms || reps <= 0 || kernel < UWS_KERNEL_VEL_SWEEP || kernel > UWS_KERNEL_PRES_SWEEP4)
        return UWS_ERR_INVALID_ARG;
    return with_solver(h, [&](SolverBase& s) { return s.time_kernel(kernel, reps, avg_ms); });
}

}  // extern "C"
