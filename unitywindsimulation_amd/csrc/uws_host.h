// Host-side solver objects and the step driver shared by the single-case
// handle (uws_api.cpp) and the ensemble (uws_batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <memory>
#include <vector>

#include "uws_internal.h"

namespace uws {

// what a step does, derived from uws_config
struct StepCfg {
    int V = 0, P = 0, C = 0;  // momentum sweeps, pressure sweeps, PISO correctors (0 for SIMPLE)
    bool les = false, norms = false, lowp = false, bf16 = false, normalise = false, piso = false;
    int R = 0, K = 0;         // defect-correction rounds / sweeps per round
    bool cal_residual = false;
    int vci = 1, pci = 1;
    float vtol = 0.f, ptol = 0.f;
};

// Buffer roles: which of the three velocity buffers holds u^n (state), which
// holds velField within a step (cur), and which p' buffer holds the latest
// pressure correction (-1: the correction is zero).
struct RoleState {
    int state = 0;
    int cur = -1;
    int pc_latest = -1;
    int piso_done = 0;
    // sweeps the last momentum solve and the last SIMPLE pressure solve ran
    // (the configured counts unless the residual-based exit stopped earlier)
    int vel_sweeps = 0;
    int pres_sweeps = 0;
};

// host<->device plane selectors: p, u' (comp 0..2), boundary textures, velocity buffer b comp c
enum { PLANE_P = 0, PLANE_UCORR = 1, PLANE_BND0 = 2, PLANE_VEL0 = 8 };

struct SolverBase {
    uws_config cfg{};
    StepCfg sc{};
    RoleState rs{};
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    int device = 0;
    void* slab = nullptr;
    size_t slab_bytes = 0;
    double* h_norms = nullptr;
    int64_t step = 0;
    double phys_time = 0.0;
    // stretched grid: facePosX/Y/Z (gridRes + 1 floats per axis); empty otherwise
    std::vector<float> face_pos[3];

    virtual ~SolverBase();
    int sync();
    virtual int ncells() const = 0;
    virtual int dims(int i) const = 0;
    virtual void* device_fields() = 0;   // FieldsT<T>*
    virtual void* device_params() = 0;   // ParamsT<T>*
    virtual int set_flags_u8(const uint8_t* fl) = 0;
    virtual int get_flags_u8(uint8_t* fl) = 0;
    virtual int set_velocity_all(const float* src, int ncomp) = 0;
    virtual int get_velocity(float* dst, int ncomp) = 0;
    virtual int set_plane(int which, int comp, const float* src, size_t n) = 0;
    virtual int set_plane_d(int which, int comp, const double* src, size_t n) = 0;
    virtual int get_plane(int which, int comp, float* dst, size_t n) = 0;
    virtual int get_field(int field, float* dst, int64_t n) = 0;
    virtual int get_field64(int field, double* dst, int64_t n) = 0;
    virtual int set_field(int field, const float* src, int64_t n) = 0;
    virtual int get_norms(double* raw) = 0;
    virtual int run_stage(int stage) = 0;
    virtual int step_n(int n) = 0;
    virtual int time_kernel(int kernel, int reps, double* avg_ms) = 0;
    // sweeps per launch of the plain (not low-precision) pressure solves and of the momentum predictor
    virtual void launch_plan(int* pres, int* vel) = 0;
    virtual int set_cs(double re, double im) = 0;
    virtual int init_from_background(const float fg_pos[3], float rot_deg, const uws_background& bg,
                                     const float* vel, int ncomp, const float* pres) = 0;
    // the same from the per-GPU staging array (no device allocation); the caller checked the size
    virtual int init_from_background_noalloc(const float fg_pos[3], float rot_deg, const uws_background& bg,
                                             const float* vel, int ncomp, const float* pres) = 0;
    // ChangePhysDomainSize: reallocate for a new configuration (same precision and grid type)
    virtual int reconfigure(const uws_config& c) = 0;
    virtual int mean_reset() = 0;
    virtual int mean_accumulate() = 0;
    virtual int get_mean(double* dst, int64_t n, int64_t* count) = 0;
    virtual int cstep_imag(bool set, int what, int face, double* buf, int64_t n) = 0;
    // stretched grid: replace the face coordinates (validated) and upload the cell sizes
    virtual int set_face_positions(const float* const pos[3]) = 0;
    bool nonuniform() const { return cfg.grid_type == UWS_GRID_COLL_NONUNIFORM; }
};
// Which multi-sweep kernels the step uses.  UWS_SWEEP_FUSION in the
// environment: 0 = one-sweep kernels only, 2 = every multi-sweep kernel at any
// grid size (for tests), otherwise the rules below; UWS_SWEEP_DEPTH = 2 keeps
// the pressure solve to pairs.
//  - Pressure triples and pairs run at every grid size: on small grids a
//    launch costs more than its traffic (64^3 runs at 400 MLUPS with triples
//    in 2-plane z chunks against 308 with single sweeps, 128 x 64 x 64 at 587
//    against 475, the 100 x 50 x 100 stretched case at 12.9 against 18.9 ms
//    per step), and on large ones they move less.
//  - Momentum pairs (one 16-wave workgroup per CU, 117 VGPRs) only where the
//    z chunk has at least 8 planes and the grid fills the CUs: 384 x 128 x 256
//    and 256 x 96 x 128 (16-plane chunks) run 13 % and 7 % faster with them,
//    192 x 96 x 64 and 160 x 96 x 64 (4-plane chunks, 1.5 rounds of
//    workgroups) 7 % and 8 % slower.
constexpr int FUSE_MIN_BLOCKS = 256;     // workgroups per launch that fill the 256 CUs
constexpr int VEL_PAIR_MIN_ZCHUNK = 8;
inline int fusion_mode() {
    const char* e = std::getenv("UWS_SWEEP_FUSION");
    if (e && e[0] == '0') return 0;
    if (e && e[0] == '2') return 2;
    return 1;
}
inline bool use_fusion(int mode) { return mode != 0; }
inline bool use_fusion3(int mode) {
    const char* e = std::getenv("UWS_SWEEP_DEPTH");
    if (e && e[0] == '2') return false;
    return mode != 0;
}
// Four-sweep pressure launches (FP32 single cases) on grids of at least
// QUAD_MIN_CELLS: 256 x 96 x 128 runs 2.0 % and 384 x 128 x 256 3.7 % faster
// per step with them than with triples, 128 x 64 x 96, 160 x 64 x 96 and
// 192 x 64 x 96 0.9 to 2.2 % slower (their 8-row tiles and longer halo cost
// more there than the fewer launches save).  UWS_SWEEP_DEPTH = 4 asks for
// them at any size, 2 or 3 never.
constexpr int QUAD_MIN_CELLS = 2000000;
inline bool use_fusion4(int mode, int cells) {
    const char* e = std::getenv("UWS_SWEEP_DEPTH");
    if (e && e[0] == '4') return mode != 0;
    if (e && (e[0] == '2' || e[0] == '3')) return false;
    return mode == 1 && cells >= QUAD_MIN_CELLS;
}
// Items of weights w (an ensemble's cases, weighed in workgroups) into the
// fewest groups whose weights sum to at most cap; an item above cap is a
// group of its own.  For a group count from the lower bound upwards, each
// item, largest first, goes to the least loaded group it fits; the first
// count that places every item is taken.  Each group lists its items in
// increasing order; empty groups are dropped.
inline std::vector<std::vector<int>> pack_groups(const std::vector<int>& w, int cap) {
    const int n = (int)w.size();
    if (n == 0) return {};
    if (cap < 1) cap = 1;
    std::vector<int> order(n);
    int64_t total = 0;
    for (int c = 0; c < n; ++c) {
        total += std::min(w[c], cap);
        order[c] = c;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
    for (int64_t nb = std::max<int64_t>(1, (total + cap - 1) / cap);; ++nb) {
        std::vector<std::vector<int>> bins((size_t)nb);
        std::vector<int64_t> load((size_t)nb, 0);
        bool ok = true;
        for (int c : order) {
            int best = -1;
            for (int k = 0; k < (int)nb; ++k)
                if ((load[k] + w[c] <= cap || load[k] == 0) && (best < 0 || load[k] < load[best])) best = k;
            if (best < 0) {
                ok = false;
                break;
            }
            bins[best].push_back(c);
            load[best] += w[c];
        }
        if (!ok) continue;  // with nb = n every item has a group of its own
        std::vector<std::vector<int>> out;
        for (auto& b : bins)
            if (!b.empty()) {
                std::sort(b.begin(), b.end());
                out.push_back(b);
            }
        return out;
    }
}

inline bool use_vel_pairs(int mode, int blocks, int zchunk) {
    if (const char* e = std::getenv("UWS_VEL_PAIRS")) return e[0] == '1';
    return mode == 2 || (mode == 1 && blocks >= FUSE_MIN_BLOCKS && zchunk >= VEL_PAIR_MIN_ZCHUNK);
}

// runs whose host-side profiles are evaluated in double
inline bool wide_host(int precision) { return precision == UWS_PREC_FP64 || precision == UWS_PREC_CSTEP; }

int validate(const uws_config* c, bool allow_residual_break);
StepCfg step_cfg(const uws_config& c);
template <class T>
void fill_params(const uws_config& c, ParamsT<T>& p);
template <class T>
size_t carve(const ParamsT<T>& p, FieldsT<T>& f, char* base, bool lowp);
int bnd_w(int X, int Y, int face);
int bnd_h(int Y, int Z, int face);
int create_solver(const uws_config& cfg, bool allow_residual_break, hipStream_t shared_stream,
                  std::unique_ptr<SolverBase>* out);
int apply_fixed_value_bnd(SolverBase& s, int32_t inflow_type, float wind_speed, const float vel_zn[3],
                          float ref_height, float alpha);
int apply_noise(SolverBase& s, uint64_t seed, float amp);
int init_vel_pres(SolverBase& s, int32_t inflow_type, float wind_speed);
int flags_from_int(SolverBase& s, const int32_t* flags, int64_t n);
int normalise_norms(const double* raw, double ncells, double* out, int n);
int stretched_face_positions(int n, float dx, int uniform, int one_side, float ratio, float* out);
int choose_by2(int Y);
int choose_zchunk(int X, int Y, int Z, int by);
int choose_zchunk3(int X, int Y, int Z);

// ---------------------------------------------------------------------------
// Step driver.  L is a launcher: l(kernel, roles, slot) enqueues one kernel
// (for one case, or for every case of an ensemble), l.copy_vel(dst, src)
// copies a velocity buffer, l.read_norm(slot, &v) synchronises and reads a
// norm (used only by the residual-based early exit), l.ncells() is N,
// l.fused() tells whether the two-sweep kernels may be used.
// ---------------------------------------------------------------------------

// VelPredict() (FvmSolverGpu.cs:1238-1282): V Jacobi sweeps from u^n
template <class Lr>
int stage_vel_sweeps(Lr& l, const StepCfg& c, RoleState& rs) {
    const int s = rs.state, w0 = (s + 1) % 3, w1 = (s + 2) % 3;
    int in = s, last_in = s, result = -1;
    bool broke = false;
    rs.vel_sweeps = c.V;
    // pairs of sweeps run as one two-sweep launch where the launcher has it
    // (the first sweep's result stays on chip: a pair writes one buffer); the
    // residual-based exit and the last sweep's residual need single sweeps
    const bool pairs = l.fused_vel() && !c.cal_residual;
    auto is_pair = [&](int k) { return pairs && k + 2 <= c.V && !(c.norms && k + 2 == c.V); };
    // Every launch writes the work buffer it does not read, so the outputs
    // alternate between w0 and w1.  The first launch's output is chosen from
    // the number of launches so that the last one writes the buffer V single
    // sweeps would end in (w1 for even V, w0 for odd V): the resulting roles do
    // not depend on the launcher's pairs (advance_roles), and launch groups
    // that share roles may differ in them.
    int launches = 0;
    for (int k = 0; k < c.V; k += is_pair(k) ? 2 : 1) ++launches;
    const int last_out = ((c.V - 1) & 1) ? w1 : w0;
    const int first_out = (launches & 1) ? last_out : (last_out == w0 ? w1 : w0);
    for (int k = 0; k < c.V;) {
        const int out = in == s ? first_out : (in == w0 ? w1 : w0);
        Roles r{};
        r.state = s; r.vin = in; r.vout = out;
        if (is_pair(k)) {
            l(KN_VEL_SWEEP2, r);
            in = r.vout;
            k += 2;
            continue;
        }
        l(KN_VEL_SWEEP, r);
        if (c.cal_residual && k % c.vci == 0) {
            l(KN_RES_VEL, r);
            l(KN_SUM_PARTIALS, r, UWS_NORM_VEL_RESIDUAL);
            double sum = 0.0;
            if (int rc = l.read_norm(UWS_NORM_VEL_RESIDUAL, &sum)) return rc;
            const float res = (float)std::sqrt(sum / (3.0 * l.ncells()));
            if (res < c.vtol) {
                // the reference keeps the previous iterate after a break (the swap is skipped, :1256-1279)
                result = in;
                broke = true;
                rs.vel_sweeps = k + 1;
                break;
            }
        }
        last_in = in;
        in = out;
        ++k;
    }
    if (result < 0) result = in;
    if (c.norms && c.V >= 1 && !broke) {
        Roles r{};
        r.vin = last_in; r.vout = result;
        l(KN_RES_VEL, r);
        l(KN_SUM_PARTIALS, r, UWS_NORM_VEL_RESIDUAL);
    }
    if (result == s) {
        // no sweep result survived: velField is the copy of u^n held in velFieldLastIter
        if (int rc = l.copy_vel(w0, s)) return rc;
        result = w0;
    }
    rs.cur = result;
    return UWS_OK;
}

// One pressure-equation solve: the SIMPLE correction (coef 0,
// FvmSolverGpu.cs:1367-1415) or a PISO corrector (coef 1, :1493-1514).
// Returns the buffer of the latest iterate in *latest (-1: zero).
template <class Lr>
int pres_solve(Lr& l, const StepCfg& c, int coef, int solve, int* latest, int* sweeps_run = nullptr) {
    if (sweeps_run) *sweeps_run = c.lowp ? c.R * c.K : c.P;
    if (c.lowp) {
        // FP32 defect correction with 2-byte correction storage
        int x = -1;
        for (int rd = 0; rd < c.R; ++rd) {
            Roles r{};
            r.coef = coef; r.pin = x;
            l(KN_DEFECT_RESIDUAL, r);
            l(KN_DEFECT_SCALE, r, UWS_NORM_REFINE0 + 8 * solve + rd);
            // fours or triples of sweeps as one multi-sweep launch where the
            // launcher has them; each launch writes the e buffer it does not read
            int ein = -1;
            for (int k = 0; k < c.K;) {
                Roles s{};
                s.coef = coef; s.pin = x; s.pout = 0; s.ein = ein; s.eout = ein == 0 ? 1 : 0;
                const int n = (l.fused4() && k + 4 <= c.K) ? 4 : ((l.fused3() && k + 3 <= c.K) ? 3 : 1);
                s.flags = (k == 0 ? K_LOWP_FIRST : 0) | (k + n == c.K ? K_LOWP_LAST : 0) | (c.bf16 ? K_BF16 : 0);
                l(n == 4 ? KN_LOWP_SWEEP4 : (n == 3 ? KN_LOWP_SWEEP3 : KN_LOWP_SWEEP), s);
                ein = s.eout;
                k += n;
            }
            if (c.K > 0) x = 0;
        }
        *latest = x;
        return UWS_OK;
    }
    int in = -1, last_in = -1;
    bool broke = false;
    // triples and pairs of sweeps run as one three- or two-sweep launch where
    // the launcher has them; the residual-based exit and the last sweep's
    // residual need single sweeps
    const bool residual_exit = coef == 0 && c.cal_residual;
    const bool last_single = coef == 0 && c.norms;
    for (int k = 0; k < c.P;) {
        const int out = in == 0 ? 1 : 0;
        Roles r{};
        r.coef = coef; r.pin = in; r.pout = out;
        if (l.fused4() && !residual_exit && k + 4 <= c.P && !(last_single && k + 4 == c.P)) {
            l(KN_PRES_SWEEP4, r);
            in = out;
            k += 4;
            continue;
        }
        if (l.fused3() && !residual_exit && k + 3 <= c.P && !(last_single && k + 3 == c.P)) {
            l(KN_PRES_SWEEP3, r);
            in = out;
            k += 3;
            continue;
        }
        if (l.fused() && !residual_exit && k + 2 <= c.P && !(last_single && k + 2 == c.P)) {
            l(KN_PRES_SWEEP2, r);
            in = out;
            k += 2;
            continue;
        }
        l(KN_PRES_SWEEP, r);
        if (residual_exit && k % c.pci == 0) {
            l(KN_RES_PRES, r);
            l(KN_SUM_PARTIALS, r, UWS_NORM_PRES_RESIDUAL);
            double sum = 0.0;
            if (int rc = l.read_norm(UWS_NORM_PRES_RESIDUAL, &sum)) return rc;
            if ((float)std::sqrt(sum / l.ncells()) < c.ptol) {
                broke = true;
                if (sweeps_run) *sweeps_run = k + 1;
                break;
            }
        }
        last_in = in;
        in = out;
        ++k;
    }
    if (coef == 0 && c.norms && c.P >= 1 && !broke) {
        Roles r{};
        r.pin = last_in; r.pout = in;
        l(KN_RES_PRES, r);
        l(KN_SUM_PARTIALS, r, UWS_NORM_PRES_RESIDUAL);
    }
    *latest = in;
    return UWS_OK;
}

template <class Lr>
int ensure_pc(Lr& l, RoleState& rs) {
    if (rs.pc_latest < 0) {
        Roles r{};
        r.pout = 0;
        l(KN_FILL_PC, r);
        rs.pc_latest = 0;
    }
    return UWS_OK;
}

template <class Lr>
int run_stage(Lr& l, const StepCfg& c, RoleState& rs, int stage) {
    Roles r{};
    r.state = rs.state;
    switch (stage) {
    case UWS_STAGE_WALL_FUNC:
        if (c.les) { r.flags = K_WALL; l(KN_WALL_EDDY, r); }
        return UWS_OK;
    case UWS_STAGE_EDDY_VIS:
        if (c.les) { r.flags = K_EDDY; l(KN_WALL_EDDY, r); }
        return UWS_OK;
    case UWS_STAGE_LES_TERM:
        if (c.les) l(KN_LES_TERM, r);
        return UWS_OK;
    case UWS_STAGE_VEL_PRECOMP:
        l(KN_VEL_PRECOMP, r);
        return UWS_OK;
    case UWS_STAGE_VEL_SWEEPS: return stage_vel_sweeps(l, c, rs);
    case UWS_STAGE_PRES_PRECOMP:
        if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
        r.vout = rs.cur;
        l(KN_PRES_PRECOMP, r);
        return UWS_OK;
    case UWS_STAGE_PRES_SWEEPS: return pres_solve(l, c, 0, 0, &rs.pc_latest, &rs.pres_sweeps);
    case UWS_STAGE_APPLY_PRES:
        ensure_pc(l, rs);
        r.pin = rs.pc_latest;
        l(KN_APPLY_PRES, r);
        if (c.normalise) l(KN_PRES_NORMALISE, r);
        return UWS_OK;
    case UWS_STAGE_VEL_CORRECT:
        if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
        ensure_pc(l, rs);
        r.pin = rs.pc_latest;
        r.vout = rs.cur;
        l(KN_VEL_CORRECT, r);
        rs.pc_latest = -1;  // SetPresCorrectFieldtoZero (FvmSolverGpu.cs:1088)
        return UWS_OK;
    case UWS_STAGE_PISO: {
        if (!c.piso) return UWS_OK;
        if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
        l(KN_PISO_AODU, r);
        l(KN_PISO_PRECOMP, r);
        const int solve = 1 + (rs.piso_done < 6 ? rs.piso_done : 6);
        if (int rc = pres_solve(l, c, 1, solve, &rs.pc_latest)) return rc;
        ensure_pc(l, rs);
        Roles q{};
        q.state = rs.state; q.pin = rs.pc_latest; q.vout = rs.cur;
        l(KN_PISO_CORRECT, q);
        rs.pc_latest = -1;
        rs.piso_done += 1;
        return UWS_OK;
    }
    case UWS_STAGE_FINISH:
        if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
        if (c.norms) {
            Roles q{};
            q.vout = rs.cur; q.state = rs.cur;
            q.flags = K_NO_FLUX;  // the stretched grid's face velocities stay those of u* for the next step
            l(KN_PRES_PRECOMP, q);
            l(KN_DIVERGENCE, q);
            l(KN_SUM_PARTIALS, q, UWS_NORM_DIVERGENCE);
            l(KN_NONFINITE, q);
            l(KN_SUM_PARTIALS, q, UWS_NORM_NONFINITE);
        }
        rs.state = rs.cur;
        rs.cur = -1;
        rs.piso_done = 0;
        return UWS_OK;
    default: return UWS_ERR_INVALID_ARG;
    }
}

// FvmSolverGpu.Step() (FvmSolverGpu.cs:1050-1121) with the wall function,
// eddy viscosity, LES term and momentum precompute fused into two kernels
template <class Lr>
int enqueue_step(Lr& l, const StepCfg& c, RoleState& rs) {
    Roles r{};
    r.state = rs.state;
    if (c.les) {
        r.flags = K_WALL | K_EDDY;
        l(KN_WALL_EDDY, r);
        r.flags = K_FUSE_LES;
    }
    l(KN_VEL_PRECOMP, r);
    for (int s = UWS_STAGE_VEL_SWEEPS; s <= UWS_STAGE_PRES_SWEEPS; ++s)
        if (int rc = run_stage(l, c, rs, s)) return rc;
    if (c.normalise) {
        for (int s = UWS_STAGE_APPLY_PRES; s <= UWS_STAGE_VEL_CORRECT; ++s)
            if (int rc = run_stage(l, c, rs, s)) return rc;
    } else {
        // p += p' (UWS_STAGE_APPLY_PRES) in the velocity correction's launch:
        // no normalisation reads p in between, and the correction reads p'
        // only
        if (rs.cur < 0) return UWS_ERR_INVALID_ARG;
        ensure_pc(l, rs);
        Roles v{};
        v.state = rs.state;
        v.pin = rs.pc_latest;
        v.vout = rs.cur;
        v.flags = K_APPLY_P;
        l(KN_VEL_CORRECT, v);
        rs.pc_latest = -1;  // SetPresCorrectFieldtoZero (FvmSolverGpu.cs:1088)
    }
    for (int n = 0; n < c.C; ++n)
        if (int rc = run_stage(l, c, rs, UWS_STAGE_PISO)) return rc;
    return run_stage(l, c, rs, UWS_STAGE_FINISH);
}

// the role change of one enqueue_step without a residual-based early exit
// (the same with and without two-sweep momentum launches)
inline void advance_roles(const StepCfg& c, RoleState& rs) {
    const int w0 = (rs.state + 1) % 3, w1 = (rs.state + 2) % 3;
    rs.state = c.V == 0 ? w0 : (((c.V - 1) & 1) ? w1 : w0);
    rs.cur = -1;
    rs.pc_latest = -1;
    rs.piso_done = 0;
}

}  // namespace uws

struct uws_solver {
    std::unique_ptr<uws::SolverBase> impl;
};
