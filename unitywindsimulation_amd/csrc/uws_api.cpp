// C-ABI implementation: handle lifetime, device memory, uploads/downloads,
// the step orchestration of FvmSolverGpu.Step() (FvmSolverGpu.cs:1050-1121)
// and its capture into HIP graphs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "uws_internal.h"

namespace uws {
void launch_zero_solid_vel(const Params& p, const Fields& f, hipStream_t s);
void launch_fill(float* dst, int n, float v, hipStream_t s);
}  // namespace uws

using uws::Fields;
using uws::Params;

struct uws_solver {
    uws_config cfg{};
    Params p{};
    Fields f{};
    void* slab = nullptr;
    size_t slab_bytes = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    // velocity buffer roles: f.vel[state] holds u^n (velFieldLastTime);
    // the other two are the Jacobi ping-pong pair (velField / velFieldLastIter)
    int state = 0;
    int cur = -1;             // buffer holding velField within a step (-1: none yet)
    float* pc_latest = nullptr;  // latest p' (presCorrectField), nullptr = zero
    double* d_res = nullptr;  // device residual result
    double* h_res = nullptr;  // pinned host copy
    int64_t step = 0;
    double phys_time = 0.0;
    double last_res[2] = {0.0, 0.0};
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // one per starting state index
};

namespace {

#define HIP_OK(expr)                                  \
    do {                                              \
        if ((expr) != hipSuccess) return UWS_ERR_DEVICE; \
    } while (0)

int use_device(uws_handle h) {
    if (hipSetDevice(h->device) != hipSuccess) return UWS_ERR_DEVICE;
    return UWS_OK;
}

int sync(uws_handle h) {
    if (hipStreamSynchronize(h->stream) != hipSuccess) return UWS_ERR_DEVICE;
    return UWS_OK;
}

int check_launch() { return hipGetLastError() == hipSuccess ? UWS_OK : UWS_ERR_DEVICE; }

bool valid_bnd(int v) { return v >= 0 && v <= 2; }

int validate(const uws_config* c) {
    if (!c) return UWS_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (c->grid_res[i] < 1) return UWS_ERR_CONFIG;
    const long long n = (long long)c->grid_res[0] * c->grid_res[1] * c->grid_res[2];
    if (n >= (1LL << 31) / 4) return UWS_ERR_CONFIG;
    if (!(c->dx > 0.f) || !(c->dt > 0.f) || !(c->den > 0.f) || !(c->mu >= 0.f)) return UWS_ERR_CONFIG;
    for (int i = 0; i < 6; ++i)
        if (!valid_bnd(c->vel_bnd[i]) || !valid_bnd(c->pres_bnd[i])) return UWS_ERR_CONFIG;
    if (c->fvm_solver_type != UWS_SIMPLE && c->fvm_solver_type != UWS_PISO) return UWS_ERR_CONFIG;
    if (c->turbulence_model != UWS_TURB_NONE && c->turbulence_model != UWS_TURB_SMAGORINSKY) return UWS_ERR_CONFIG;
    if (c->piso_num_correctors < 0 || c->vel_max_num_iter < 0 || c->pres_max_num_iter < 0) return UWS_ERR_CONFIG;
    if (c->cal_residual && (c->vel_residual_check_interval <= 0 || c->pres_residual_check_interval <= 0))
        return UWS_ERR_CONFIG;
    if (c->wall_slots != 1 && c->wall_slots != 3) return UWS_ERR_CONFIG;
    if (c->wall_model != UWS_WALL_SMOOTH && c->wall_model != UWS_WALL_ROUGH) return UWS_ERR_CONFIG;
    if (c->wall_model == UWS_WALL_ROUGH && !(c->wall_z0 > 0.f)) return UWS_ERR_CONFIG;
    for (int i = 0; i < 8; ++i)
        if (c->reserved[i] != 0) return UWS_ERR_CONFIG;
    return UWS_OK;
}

void fill_params(const uws_config& c, Params& p) {
    p.X = c.grid_res[0]; p.Y = c.grid_res[1]; p.Z = c.grid_res[2];
    p.N = p.X * p.Y * p.Z;
    // FluidSimConfig.Init (FluidSimConfig.cs:325-341), float arithmetic
    p.dx = c.dx;
    p.ds = c.dx * c.dx;
    p.dv = c.dx * c.dx * c.dx;
    p.dt = c.dt;
    p.den = c.den;
    p.nu = c.mu / c.den;
    p.cs = c.smagorinsky_constant;
    for (int i = 0; i < 3; ++i) p.fext[i] = c.external_force[i];
    for (int i = 0; i < 6; ++i) { p.velBnd[i] = c.vel_bnd[i]; p.presBnd[i] = c.pres_bnd[i]; }
    p.useLES = c.turbulence_model == UWS_TURB_SMAGORINSKY;
    p.wallSlots = c.wall_slots;
    p.wallModel = c.wall_model;
    p.wallZ0 = c.wall_z0;
}

int bnd_w(const Params& p, int face) { return face < 2 ? p.Y : p.X; }
int bnd_h(const Params& p, int face) { return face < 4 ? p.Z : p.Y; }

// carve the slab: returns the number of bytes needed when base == nullptr
size_t carve(const Params& p, Fields& f, char* base) {
    size_t off = 0;
    const size_t N = (size_t)p.N;
    auto take = [&](size_t bytes) -> char* {
        off = (off + 255) & ~(size_t)255;
        char* r = base ? base + off : nullptr;
        off += bytes;
        return r;
    };
    auto fl = [&](size_t n) { return (float*)take(n * sizeof(float)); };
    f.flag = (uint8_t*)take(N);
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) f.vel[b][c] = fl(N);
    for (int c = 0; c < 3; ++c) f.ucorr[c] = fl(N);
    f.p = fl(N);
    f.pc[0] = fl(N); f.pc[1] = fl(N);
    f.nut = fl(N);
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
        for (int c = 0; c < 3; ++c) f.bnd[face][c] = fl((size_t)bnd_w(p, face) * bnd_h(p, face));
    const size_t nblocks = (size_t)((p.X + 63) / 64) * ((p.Y + 3) / 4) * p.Z;
    f.partial = (double*)take(nblocks * sizeof(double));
    f.scalar = fl(64);
    return off + 256;
}

// --- stage implementations (FvmSolverGpu.cs:1050-1121) ---

const float* const* cvel(uws_handle h, int b) { return (const float* const*)h->f.vel[b]; }

int stage_wall_eddy(uws_handle h, bool wall, bool eddy) {
    if (!h->p.useLES) return UWS_OK;
    uws::launch_wall_eddy(h->p, h->f, cvel(h, h->state), wall, eddy, h->stream);
    return check_launch();
}

int read_residual(uws_handle h, double* out) {
    HIP_OK(hipMemcpyAsync(h->h_res, h->d_res, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    *out = *h->h_res;
    return UWS_OK;
}

int copy_vel(uws_handle h, int dst, int src) {
    for (int c = 0; c < 3; ++c)
        HIP_OK(hipMemcpyAsync(h->f.vel[dst][c], h->f.vel[src][c], sizeof(float) * h->p.N, hipMemcpyDeviceToDevice,
                              h->stream));
    return UWS_OK;
}

// VelPredict() (FvmSolverGpu.cs:1238-1282): V Jacobi sweeps from u^n
int stage_vel_sweeps(uws_handle h) {
    const int s = h->state, w0 = (s + 1) % 3, w1 = (s + 2) % 3;
    const int V = h->cfg.vel_max_num_iter;
    int in = s;
    int result = -1;
    for (int k = 0; k < V; ++k) {
        const int out = (k & 1) ? w1 : w0;
        uws::launch_vel_sweep(h->p, h->f, cvel(h, in), h->f.vel[out], h->stream);
        if (int rc = check_launch()) return rc;
        if (h->cfg.cal_residual && k % h->cfg.vel_residual_check_interval == 0) {
            uws::launch_residual(h->p, h->f, 3, h->f.D, cvel(h, out), cvel(h, in), h->d_res, h->stream);
            double sum;
            if (int rc = read_residual(h, &sum)) return rc;
            const float r = (float)std::sqrt(sum / (3.0 * (double)h->p.N));
            h->last_res[0] = r;
            if (r < h->cfg.vel_tolerance) {
                // the reference keeps the previous iterate after a break (swap skipped, :1256-1279)
                result = in;
                break;
            }
        }
        in = out;
    }
    if (result < 0) result = in;
    if (result == s) {
        // no sweep survived: velField is the copy of u^n held in velFieldLastIter
        if (int rc = copy_vel(h, w0, s)) return rc;
        result = w0;
    }
    h->cur = result;
    return UWS_OK;
}

// PresCorrect() sweeps (FvmSolverGpu.cs:1367-1415) and the PISO sweeps
// (:1493-1514); returns the latest iterate in *latest (nullptr = zero)
int pres_sweeps(uws_handle h, const float* cE, const float* cN, const float* cU, const float* invDp, const float* bp,
                const float* dp_for_residual, bool allow_break, float** latest) {
    const int P = h->cfg.pres_max_num_iter;
    float* in = nullptr;
    bool broke = false;
    for (int k = 0; k < P; ++k) {
        float* out = h->f.pc[k & 1];
        uws::launch_pres_sweep(h->p, h->f, cE, cN, cU, invDp, bp, in, out, h->stream);
        if (int rc = check_launch()) return rc;
        if (allow_break && h->cfg.cal_residual && k % h->cfg.pres_residual_check_interval == 0) {
            const float* dd[3] = {dp_for_residual, nullptr, nullptr};
            const float* aa[3] = {out, nullptr, nullptr};
            const float* bb[3] = {in, nullptr, nullptr};
            uws::launch_residual(h->p, h->f, 1, dd, aa, bb, h->d_res, h->stream);
            double sum;
            if (int rc = read_residual(h, &sum)) return rc;
            const float r = (float)std::sqrt(sum / (double)h->p.N);
            h->last_res[1] = r;
            if (r < h->cfg.pres_tolerance) { broke = true; break; }
        }
        in = out;
    }
    (void)broke;
    *latest = in;
    return UWS_OK;
}

bool any_fixed_pres(const Params& p) {
    for (int i = 0; i < 6; ++i)
        if (p.presBnd[i] == UWS_PRES_FIXED_VALUE) return true;
    return false;
}

// a p' that is logically zero but must be read: fill one buffer with zeros
float* materialise_zero_pc(uws_handle h) {
    uws::launch_fill(h->f.pc[0], h->p.N, 0.f, h->stream);
    return h->f.pc[0];
}

int run_stage(uws_handle h, int stage) {
    Params& p = h->p;
    Fields& f = h->f;
    switch (stage) {
    case UWS_STAGE_WALL_FUNC: return stage_wall_eddy(h, true, false);
    case UWS_STAGE_EDDY_VIS: return stage_wall_eddy(h, false, true);
    case UWS_STAGE_LES_TERM:
        if (!p.useLES) return UWS_OK;
        uws::launch_les_term(p, f, cvel(h, h->state), h->stream);
        return check_launch();
    case UWS_STAGE_VEL_PRECOMP:
        uws::launch_vel_precompute(p, f, cvel(h, h->state), false, h->stream);
        return check_launch();
    case UWS_STAGE_VEL_SWEEPS: return stage_vel_sweeps(h);
    case UWS_STAGE_PRES_PRECOMP:
        if (h->cur < 0) return UWS_ERR_INVALID_ARG;
        uws::launch_pres_precompute(p, f, cvel(h, h->cur), h->stream);
        return check_launch();
    case UWS_STAGE_PRES_SWEEPS:
        return pres_sweeps(h, f.cp[0], f.cp[1], f.cp[2], f.invDp, f.bp, f.Dp, true, &h->pc_latest);
    case UWS_STAGE_APPLY_PRES: {
        const float* pc = h->pc_latest ? h->pc_latest : materialise_zero_pc(h);
        uws::launch_apply_pres(p, f, pc, !any_fixed_pres(p), h->stream);
        return check_launch();
    }
    case UWS_STAGE_VEL_CORRECT: {
        if (h->cur < 0) return UWS_ERR_INVALID_ARG;
        const float* pc = h->pc_latest ? h->pc_latest : materialise_zero_pc(h);
        uws::launch_vel_correct(p, f, pc, f.vel[h->cur], h->stream);
        h->pc_latest = nullptr;  // SetPresCorrectFieldtoZero (FvmSolverGpu.cs:1088)
        return check_launch();
    }
    case UWS_STAGE_PISO: {
        if (h->cfg.fvm_solver_type != UWS_PISO) return UWS_OK;
        if (h->cur < 0) return UWS_ERR_INVALID_ARG;
        uws::launch_piso_aodu(p, f, cvel(h, h->state), h->stream);
        uws::launch_piso_precompute(p, f, h->stream);
        if (int rc = check_launch()) return rc;
        float* latest = nullptr;
        if (int rc = pres_sweeps(h, f.cpisoE, f.cp[1], f.cp[2], f.invDpiso, f.bpiso, nullptr, false, &latest))
            return rc;
        const float* pc = latest ? latest : materialise_zero_pc(h);
        uws::launch_piso_correct(p, f, pc, f.vel[h->cur], h->stream);
        h->pc_latest = nullptr;
        return check_launch();
    }
    case UWS_STAGE_FINISH:
        if (h->cur < 0) return UWS_ERR_INVALID_ARG;
        h->state = h->cur;
        h->cur = -1;
        h->step += 1;
        h->phys_time += (double)h->cfg.dt;
        return UWS_OK;
    default: return UWS_ERR_INVALID_ARG;
    }
}

// the whole step with the fused kernels used outside the staged API
int enqueue_step(uws_handle h) {
    Params& p = h->p;
    Fields& f = h->f;
    if (p.useLES) {
        uws::launch_wall_eddy(p, f, cvel(h, h->state), true, true, h->stream);
        uws::launch_vel_precompute(p, f, cvel(h, h->state), true, h->stream);
    } else {
        uws::launch_vel_precompute(p, f, cvel(h, h->state), false, h->stream);
    }
    if (int rc = check_launch()) return rc;
    for (int s = UWS_STAGE_VEL_SWEEPS; s <= UWS_STAGE_VEL_CORRECT; ++s)
        if (int rc = run_stage(h, s)) return rc;
    if (h->cfg.fvm_solver_type == UWS_PISO)
        for (int n = 0; n < h->cfg.piso_num_correctors; ++n)
            if (int rc = run_stage(h, UWS_STAGE_PISO)) return rc;
    return run_stage(h, UWS_STAGE_FINISH);
}

void drop_graphs(uws_handle h) {
    for (auto& g : h->graph)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

int step_graph(uws_handle h) {
    const int s0 = h->state;
    if (!h->graph[s0]) {
        // capture one step starting from buffer s0; the captured launches
        // depend only on the buffer roles, which the capture itself advances
        const int64_t step0 = h->step;
        const double t0 = h->phys_time;
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(h);
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        h->state = s0;
        h->step = step0;
        h->phys_time = t0;
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return UWS_ERR_DEVICE;
        e = hipGraphInstantiate(&h->graph[s0], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return UWS_ERR_DEVICE;
    }
    HIP_OK(hipGraphLaunch(h->graph[s0], h->stream));
    // advance the buffer roles exactly as enqueue_step does
    const int V = h->cfg.vel_max_num_iter;
    const int w0 = (s0 + 1) % 3, w1 = (s0 + 2) % 3;
    h->state = V == 0 ? w0 : (((V - 1) & 1) ? w1 : w0);
    h->cur = -1;
    h->pc_latest = nullptr;
    h->step += 1;
    h->phys_time += (double)h->cfg.dt;
    return UWS_OK;
}

int upload(uws_handle h, float* dst, const float* src, size_t n) {
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    return UWS_OK;
}
int download(uws_handle h, float* dst, const float* src, size_t n) {
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    return UWS_OK;
}

// AoS [n][ncomp] <-> SoA planes
int set_vec(uws_handle h, float* const dst[3], const float* src, int ncomp, int64_t n) {
    std::vector<float> tmp((size_t)n);
    for (int c = 0; c < 3; ++c) {
        for (int64_t i = 0; i < n; ++i) tmp[i] = src[i * ncomp + c];
        if (int rc = upload(h, dst[c], tmp.data(), (size_t)n)) return rc;
        if (int rc = sync(h)) return rc;
    }
    return UWS_OK;
}
int get_vec(uws_handle h, float* dst, float* const src[3], int ncomp, int64_t n) {
    std::vector<float> tmp((size_t)n);
    for (int c = 0; c < 3; ++c) {
        if (int rc = download(h, tmp.data(), src[c], (size_t)n)) return rc;
        if (int rc = sync(h)) return rc;
        for (int64_t i = 0; i < n; ++i) dst[i * ncomp + c] = tmp[i];
    }
    if (ncomp == 4)
        for (int64_t i = 0; i < n; ++i) dst[i * 4 + 3] = 0.f;
    return UWS_OK;
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
    return UWS_OK;
}

int uws_create(const uws_config* cfg, uws_handle* out) {
    if (!out) return UWS_ERR_INVALID_ARG;
    *out = nullptr;
    if (int rc = validate(cfg)) return rc;
    uws_solver* h = new (std::nothrow) uws_solver();
    if (!h) return UWS_ERR_ALLOC;
    h->cfg = *cfg;
    fill_params(*cfg, h->p);
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return UWS_ERR_DEVICE; }
    h->slab_bytes = carve(h->p, h->f, nullptr);
    if (hipMalloc(&h->slab, h->slab_bytes) != hipSuccess) { delete h; return UWS_ERR_ALLOC; }
    carve(h->p, h->f, (char*)h->slab);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(h->slab); delete h; return UWS_ERR_DEVICE;
    }
    if (hipMalloc(&h->d_res, sizeof(double)) != hipSuccess ||
        hipHostMalloc(&h->h_res, sizeof(double), hipHostMallocDefault) != hipSuccess) {
        uws_destroy(h); return UWS_ERR_ALLOC;
    }
    // all-fluid flags (CELL_FLUID = 0), zero fields and boundary textures
    if (hipMemsetAsync(h->slab, 0, h->slab_bytes, h->stream) != hipSuccess || sync(h)) {
        uws_destroy(h); return UWS_ERR_DEVICE;
    }
    *out = h;
    return UWS_OK;
}

int uws_destroy(uws_handle h) {
    if (!h) return UWS_ERR_INVALID_ARG;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->slab) (void)hipFree(h->slab);
    if (h->d_res) (void)hipFree(h->d_res);
    if (h->h_res) (void)hipHostFree(h->h_res);
    delete h;
    return UWS_OK;
}

int uws_get_config(uws_handle h, uws_config* out) {
    if (!h || !out) return UWS_ERR_INVALID_ARG;
    *out = h->cfg;
    return UWS_OK;
}

int uws_num_cells(uws_handle h, int64_t* n) {
    if (!h || !n) return UWS_ERR_INVALID_ARG;
    *n = h->p.N;
    return UWS_OK;
}

int uws_init_flags(uws_handle h, const int32_t* flags, int64_t n) {
    if (!h || !flags || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    std::vector<uint8_t> fl((size_t)n);
    for (int64_t i = 0; i < n; ++i) fl[i] = (double)flags[i] > 0.5 ? UWS_CELL_SOLID : UWS_CELL_FLUID;
    HIP_OK(hipMemcpyAsync(h->f.flag, fl.data(), (size_t)n, hipMemcpyHostToDevice, h->stream));
    uws::launch_zero_solid_vel(h->p, h->f, h->stream);
    if (int rc = check_launch()) return rc;
    h->cur = -1;
    h->pc_latest = nullptr;
    return sync(h);
}

int uws_init_box(uws_handle h) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    const Params& p = h->p;
    std::vector<uint8_t> fl((size_t)p.N);
    HIP_OK(hipMemcpyAsync(fl.data(), h->f.flag, (size_t)p.N, hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync(h)) return rc;
    for (int z = 0; z < p.Z; ++z)
        for (int y = 0; y < p.Y; ++y)
            for (int x = 0; x < p.X; ++x)
                if (x >= 20 && x < 24 && y >= 0 && y < 8 && z >= 18 && z < 22)
                    fl[((size_t)z * p.Y + y) * p.X + x] = UWS_CELL_SOLID;
    HIP_OK(hipMemcpyAsync(h->f.flag, fl.data(), (size_t)p.N, hipMemcpyHostToDevice, h->stream));
    uws::launch_zero_solid_vel(h->p, h->f, h->stream);
    if (int rc = check_launch()) return rc;
    return sync(h);
}

int uws_init_vel_pres_fields(uws_handle h, int32_t inflow_type, float wind_speed) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (inflow_type != UWS_INFLOW_CONSTANT && inflow_type != UWS_INFLOW_LOG) return UWS_ERR_UNSUPPORTED;
    if (int rc = use_device(h)) return rc;
    const Params& p = h->p;
    std::vector<float> col((size_t)p.Y);
    for (int y = 0; y < p.Y; ++y) {
        if (inflow_type == UWS_INFLOW_LOG) {
            // CSInitLogVelPresFields (InitShaderColl.compute:112-132): y = j*dx
            const float karman = 0.4f, rough = 1.0f, refH = 10.0f;
            const float yPhys = y * p.dx;
            const float fricVel = karman * wind_speed / std::log((refH + rough) / rough);
            col[y] = fricVel / karman * std::log((yPhys + rough) / rough);
        } else {
            col[y] = wind_speed;
        }
    }
    std::vector<float> ux((size_t)p.N), zero((size_t)p.N, 0.f);
    for (int z = 0; z < p.Z; ++z)
        for (int y = 0; y < p.Y; ++y)
            for (int x = 0; x < p.X; ++x) ux[((size_t)z * p.Y + y) * p.X + x] = col[y];
    for (int b = 0; b < 3; ++b) {
        if (int rc = upload(h, h->f.vel[b][0], ux.data(), p.N)) return rc;
        if (int rc = upload(h, h->f.vel[b][1], zero.data(), p.N)) return rc;
        if (int rc = upload(h, h->f.vel[b][2], zero.data(), p.N)) return rc;
    }
    for (int c = 0; c < 3; ++c)
        if (int rc = upload(h, h->f.ucorr[c], zero.data(), p.N)) return rc;
    if (int rc = upload(h, h->f.p, zero.data(), p.N)) return rc;
    h->cur = -1;
    h->pc_latest = nullptr;
    return sync(h);
}

int uws_set_fixed_value_vel_bnd_cond(uws_handle h, int32_t inflow_type, float wind_speed, const float vel_zn[3],
                                     float ref_height, float alpha) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    const Params& p = h->p;
    std::vector<float> col((size_t)p.Y);
    if (inflow_type == UWS_INFLOW_CONSTANT) {
        for (int y = 0; y < p.Y; ++y) col[y] = wind_speed;
    } else if (inflow_type == UWS_INFLOW_LOG) {
        // CSSetLogVelBndCond (InitShaderColl.compute:272-285): y = (j + 1/2) dx
        const float karman = 0.4f, rough = 1.0f, refH = 10.0f;
        for (int y = 0; y < p.Y; ++y) {
            const float yPhys = (y + 0.5f) * p.dx;
            const float fricVel = karman * wind_speed / std::log((refH + rough) / rough);
            col[y] = fricVel / karman * std::log((yPhys + rough) / rough);
        }
    } else if (inflow_type == UWS_INFLOW_POWER_LAW) {
        if (!(ref_height > 0.f)) return UWS_ERR_INVALID_ARG;
        for (int y = 0; y < p.Y; ++y) {
            const float yPhys = (y + 0.5f) * p.dx;
            col[y] = wind_speed * (float)std::pow((double)(yPhys / ref_height), (double)alpha);
        }
    } else {
        return UWS_ERR_INVALID_ARG;
    }
    // X0 texture (y, z): (u, 0, 0)
    const size_t nx0 = (size_t)p.Y * p.Z;
    std::vector<float> a(nx0), zero(nx0, 0.f);
    for (int z = 0; z < p.Z; ++z)
        for (int y = 0; y < p.Y; ++y) a[(size_t)z * p.Y + y] = col[y];
    if (int rc = upload(h, h->f.bnd[UWS_X0][0], a.data(), nx0)) return rc;
    if (int rc = upload(h, h->f.bnd[UWS_X0][1], zero.data(), nx0)) return rc;
    if (int rc = upload(h, h->f.bnd[UWS_X0][2], zero.data(), nx0)) return rc;
    if (int rc = sync(h)) return rc;
    if (inflow_type == UWS_INFLOW_CONSTANT) {
        // CSSetFixedValueVelBndCond also writes Zn = velZn (InitShaderColl.compute:261-264)
        const size_t nzn = (size_t)p.X * p.Y;
        for (int c = 0; c < 3; ++c) {
            std::vector<float> v(nzn, vel_zn ? vel_zn[c] : 0.f);
            if (int rc = upload(h, h->f.bnd[UWS_ZN][c], v.data(), nzn)) return rc;
            if (int rc = sync(h)) return rc;
        }
    }
    return UWS_OK;
}

int uws_set_bnd_field(uws_handle h, int32_t face, const float* xyz, int64_t n) {
    if (!h || !xyz || face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
    if (n != (int64_t)bnd_w(h->p, face) * bnd_h(h->p, face)) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    return set_vec(h, h->f.bnd[face], xyz, 3, n);
}

int uws_get_bnd_field(uws_handle h, int32_t face, float* xyz, int64_t n) {
    if (!h || !xyz || face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
    if (n != (int64_t)bnd_w(h->p, face) * bnd_h(h->p, face)) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    return get_vec(h, xyz, h->f.bnd[face], 3, n);
}

int uws_set_velocity(uws_handle h, const float* src, int32_t ncomp, int64_t n) {
    if (!h || !src || (ncomp != 3 && ncomp != 4) || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    for (int b = 0; b < 3; ++b)
        if (int rc = set_vec(h, h->f.vel[b], src, ncomp, n)) return rc;
    h->cur = -1;
    return UWS_OK;
}

int uws_get_velocity(uws_handle h, float* dst, int32_t ncomp, int64_t n) {
    if (!h || !dst || (ncomp != 3 && ncomp != 4) || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    return get_vec(h, dst, h->f.vel[h->state], ncomp, n);
}

int uws_set_pressure(uws_handle h, const float* src, int64_t n) {
    if (!h || !src || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    if (int rc = upload(h, h->f.p, src, (size_t)n)) return rc;
    return sync(h);
}

int uws_get_pressure(uws_handle h, float* dst, int64_t n) {
    if (!h || !dst || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    if (int rc = download(h, dst, h->f.p, (size_t)n)) return rc;
    return sync(h);
}

int uws_get_flags(uws_handle h, int32_t* dst, int64_t n) {
    if (!h || !dst || n != h->p.N) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    std::vector<uint8_t> fl((size_t)n);
    HIP_OK(hipMemcpyAsync(fl.data(), h->f.flag, (size_t)n, hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync(h)) return rc;
    for (int64_t i = 0; i < n; ++i) dst[i] = fl[i];
    return UWS_OK;
}

int uws_init_noise(uws_handle h, uint64_t seed, float amp) {
    if (!h || !(amp >= 0.f)) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    const int64_t n = h->p.N;
    std::vector<uint8_t> fl((size_t)n);
    HIP_OK(hipMemcpyAsync(fl.data(), h->f.flag, (size_t)n, hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync(h)) return rc;
    std::vector<float> u((size_t)n * 3, 0.f);
    uint64_t s = seed;
    // one draw per component of every cell, in cell order, solid cells included
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float r = (2.0f * u01(s) - 1.0f) * amp;
            u[i * 3 + c] = fl[i] == UWS_CELL_FLUID ? r : 0.f;
        }
    if (int rc = uws_set_velocity(h, u.data(), 3, n)) return rc;
    std::vector<float> zero((size_t)n, 0.f);
    for (int c = 0; c < 3; ++c)
        if (int rc = upload(h, h->f.ucorr[c], zero.data(), n)) return rc;
    if (int rc = upload(h, h->f.p, zero.data(), n)) return rc;
    h->pc_latest = nullptr;
    return sync(h);
}

int uws_step(uws_handle h, int32_t nsteps) {
    if (!h || nsteps < 0) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    if (h->cur >= 0) return UWS_ERR_INVALID_ARG;  // a staged step is in progress
    const bool graphable = !h->cfg.cal_residual && std::getenv("UWS_NO_GRAPH") == nullptr;
    for (int i = 0; i < nsteps; ++i) {
        int rc = graphable ? step_graph(h) : enqueue_step(h);
        if (rc) return rc;
    }
    return UWS_OK;
}

int uws_synchronize(uws_handle h) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    return sync(h);
}

int uws_get_time(uws_handle h, int64_t* step, double* phys_time) {
    if (!h) return UWS_ERR_INVALID_ARG;
    if (step) *step = h->step;
    if (phys_time) *phys_time = h->phys_time;
    return UWS_OK;
}

int uws_field_ncomp(int32_t field) {
    switch (field) {
    case UWS_FIELD_VEL: case UWS_FIELD_VEL_WORK: case UWS_FIELD_VEL_CORR: case UWS_FIELD_WALL_VIS:
    case UWS_FIELD_B: case UWS_FIELD_D: case UWS_FIELD_AODU:
        return 3;
    case UWS_FIELD_PRES: case UWS_FIELD_PRES_CORR: case UWS_FIELD_EDDY_VIS: case UWS_FIELD_DP: case UWS_FIELD_BP:
        return 1;
    default: return UWS_ERR_INVALID_ARG;
    }
}

namespace {
// resolve a field id to device planes; returns ncomp or an error
int field_planes(uws_handle h, int32_t field, float* planes[3], bool for_write) {
    Fields& f = h->f;
    planes[0] = planes[1] = planes[2] = nullptr;
    switch (field) {
    case UWS_FIELD_VEL: for (int c = 0; c < 3; ++c) planes[c] = f.vel[h->state][c]; return 3;
    case UWS_FIELD_VEL_WORK:
        if (h->cur < 0) return UWS_ERR_INVALID_ARG;
        for (int c = 0; c < 3; ++c) planes[c] = f.vel[h->cur][c];
        return 3;
    case UWS_FIELD_VEL_CORR: for (int c = 0; c < 3; ++c) planes[c] = f.ucorr[c]; return 3;
    case UWS_FIELD_WALL_VIS: for (int c = 0; c < 3; ++c) planes[c] = f.nuw[c]; return 3;
    case UWS_FIELD_B: for (int c = 0; c < 3; ++c) planes[c] = f.b[c]; return 3;
    case UWS_FIELD_D: for (int c = 0; c < 3; ++c) planes[c] = f.D[c]; return 3;
    case UWS_FIELD_AODU: for (int c = 0; c < 3; ++c) planes[c] = f.aodu[c]; return 3;
    case UWS_FIELD_PRES: planes[0] = f.p; return 1;
    case UWS_FIELD_PRES_CORR:
        if (for_write) return UWS_ERR_INVALID_ARG;
        planes[0] = h->pc_latest;  // nullptr = zero
        return 1;
    case UWS_FIELD_EDDY_VIS: planes[0] = f.nut; return 1;
    case UWS_FIELD_DP: planes[0] = f.Dp; return 1;
    case UWS_FIELD_BP: planes[0] = f.bp; return 1;
    default: return UWS_ERR_INVALID_ARG;
    }
}
}  // namespace

int uws_get_field(uws_handle h, int32_t field, float* dst, int64_t n_floats) {
    if (!h || !dst) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    float* planes[3];
    const int nc = field_planes(h, field, planes, false);
    if (nc < 0) return nc;
    if (n_floats != (int64_t)nc * h->p.N) return UWS_ERR_INVALID_ARG;
    if (nc == 1) {
        if (!planes[0]) {
            std::memset(dst, 0, sizeof(float) * (size_t)h->p.N);
            return UWS_OK;
        }
        if (int rc = download(h, dst, planes[0], (size_t)h->p.N)) return rc;
        return sync(h);
    }
    return get_vec(h, dst, planes, 3, h->p.N);
}

int uws_set_field(uws_handle h, int32_t field, const float* src, int64_t n_floats) {
    if (!h || !src) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    if (field == UWS_FIELD_VEL) return uws_set_velocity(h, src, 3, n_floats / 3);
    float* planes[3];
    const int nc = field_planes(h, field, planes, true);
    if (nc < 0) return nc;
    if (n_floats != (int64_t)nc * h->p.N) return UWS_ERR_INVALID_ARG;
    if (nc == 1) {
        if (int rc = upload(h, planes[0], src, (size_t)h->p.N)) return rc;
        return sync(h);
    }
    return set_vec(h, planes, src, 3, h->p.N);
}

int uws_run_stage(uws_handle h, int32_t stage) {
    if (!h || stage < 0 || stage >= UWS_STAGE_COUNT) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    if (int rc = run_stage(h, stage)) return rc;
    return sync(h);
}

int uws_time_kernel(uws_handle h, int32_t kernel, int32_t reps, double* avg_ms) {
    if (!h || !avg_ms || reps <= 0 || (kernel != UWS_KERNEL_VEL_SWEEP && kernel != UWS_KERNEL_PRES_SWEEP))
        return UWS_ERR_INVALID_ARG;
    if (h->cur >= 0) return UWS_ERR_INVALID_ARG;
    if (int rc = use_device(h)) return rc;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return UWS_ERR_DEVICE; }
    const int w0 = (h->state + 1) % 3;
    int rc = UWS_OK;
    auto launch = [&](int i) {
        if (kernel == UWS_KERNEL_VEL_SWEEP)
            uws::launch_vel_sweep(h->p, h->f, cvel(h, h->state), h->f.vel[w0], h->stream);
        else
            uws::launch_pres_sweep(h->p, h->f, h->f.cp[0], h->f.cp[1], h->f.cp[2], h->f.invDp, h->f.bp,
                                   h->f.pc[i & 1], h->f.pc[(i + 1) & 1], h->stream);
    };
    launch(0);  // warm
    if (hipEventRecord(e0, h->stream) != hipSuccess) rc = UWS_ERR_DEVICE;
    for (int i = 0; i < reps && rc == UWS_OK; ++i) launch(i);
    if (rc == UWS_OK && check_launch()) rc = UWS_ERR_DEVICE;
    if (rc == UWS_OK && hipEventRecord(e1, h->stream) != hipSuccess) rc = UWS_ERR_DEVICE;
    if (rc == UWS_OK && hipEventSynchronize(e1) != hipSuccess) rc = UWS_ERR_DEVICE;
    float ms = 0.f;
    if (rc == UWS_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = UWS_ERR_DEVICE;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == UWS_OK) *avg_ms = (double)ms / reps;
    return rc;
}

}  // extern "C"
