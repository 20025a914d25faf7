// C++ host side of the solver step: the reference's configuration classes and
// solver plugin interface (Assets/Code/FvmSolver/FluidSimConfig.cs,
// Assets/Code/RuntimeConfig/RuntimeConfig.cs, Assets/Code/FvmSolver/IFvmSolver.cs)
// restated in C++ over the C ABI of include/uws.h.  Header-only; link with
// libuws_hip.so.  Errors throw uws::Error (the reference throws on
// unsupported combinations and checks netCDF status codes the same way).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/uws.h"

namespace uws {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what + ": " + uws_strerror(s)), status(s) {}
};
inline void check(int s, const char* what) {
    if (s != UWS_OK) throw Error(s, what);
}

// enums with the reference's integer values (FluidSimConfig.cs:9-97)
enum class VelBndCond : int32_t { FixedValue = 0, ZeroGrad = 1, Symmetry = 2 };
enum class PresBndCond : int32_t { FixedValue = 0, ZeroGrad = 1, Symmetry = 2 };
enum class FvmSolverType : int32_t { SIMPLE = 0, PISO = 1 };
enum class SolidType : int32_t { NoSolid = 0, Box = 1, Model = 2 };
enum class TurbulenceModel : int32_t { None = 0, Smagorinsky = 1 };
enum class InflowType : int32_t { Constant = 0, Log = 1, PowerLaw = 2 };
enum class GridType : int32_t { Collocated = 0, CollNonUniform = 1, Staggered = 2 };

struct float3 {
    float x, y, z;
};
struct int3 {
    int32_t x, y, z;
};

// FluidSimConfig (FluidSimConfig.cs:100-370): the fields the solver step
// reads, with the reference's names and defaults, and Init()'s derivations.
struct FluidSimConfig {
    float3 physDomainSize{10.f, 5.f, 10.f};
    int32_t gridResX = 100;
    float dt = 0.1f;
    float mu = 0.005f;
    float den = 1.0f;
    float3 externalForce{0.f, 0.f, 0.f};
    VelBndCond velBndCondX0 = VelBndCond::FixedValue, velBndCondXn = VelBndCond::FixedValue,
               velBndCondY0 = VelBndCond::FixedValue, velBndCondYn = VelBndCond::FixedValue,
               velBndCondZ0 = VelBndCond::FixedValue, velBndCondZn = VelBndCond::FixedValue;
    PresBndCond presBndCondX0 = PresBndCond::ZeroGrad, presBndCondXn = PresBndCond::ZeroGrad,
                presBndCondY0 = PresBndCond::ZeroGrad, presBndCondYn = PresBndCond::ZeroGrad,
                presBndCondZ0 = PresBndCond::ZeroGrad, presBndCondZn = PresBndCond::ZeroGrad;
    float3 velX0{0.f, 0.f, 0.f};
    float3 velZn{0.f, 0.f, 0.f};
    float3 physFieldPos{0.f, 0.f, 0.f};  // world position of the grid centre (FluidSimConfig.cs:256)
    SolidType solidType = SolidType::NoSolid;
    FvmSolverType fvmSolverType = FvmSolverType::SIMPLE;
    int32_t PISONumCorrectors = 2;
    TurbulenceModel turbulenceModel = TurbulenceModel::None;
    float smagorinskyConstant = 0.15f;
    InflowType inflowType = InflowType::Constant;
    GridType gridType = GridType::Collocated;
    bool calResidual = false;
    int32_t velResidualCheckInterval = 10;
    float velTolerance = 1e-15f;
    int32_t velMaxNumIter = 50;
    int32_t presResidualCheckInterval = 20;
    float presTolerance = 1e-3f;
    int32_t presMaxNumIter = 400;
    // extensions
    int32_t wallSlots = 1;
    int32_t wallModel = UWS_WALL_SMOOTH;
    float wallZ0 = 0.01f;
    float inflowRefHeight = 16.f, inflowAlpha = 0.25f;
    int32_t precision = UWS_PREC_FP32;
    int32_t refineRounds = 3, refineSweeps = 20;
    bool computeNorms = false;

    // deduced parameters (FluidSimConfig.cs:207-245)
    float dx = 0, ds = 0, dv = 0, nu = 0, D = 0, DInv = 0;
    int3 gridRes{0, 0, 0};
    float currPhyTime = 0;
    int32_t currSimStep = 0;

    // Init() (FluidSimConfig.cs:322-346); Mathf.RoundToInt rounds half to even
    void Init() {
        dx = physDomainSize.x / (float)gridResX;
        ds = dx * dx;
        dv = dx * dx * dx;
        const int32_t gy = (int32_t)std::nearbyint(physDomainSize.y / dx);
        const int32_t gz = (int32_t)std::nearbyint(physDomainSize.z / dx);
        gridRes = int3{gridResX, gy, gz};
        nu = mu / den;
        D = dv / dt + 6 * nu * ds / dx;
        DInv = 1 / (dv / dt + 6 * nu * ds / dx);
        currPhyTime = 0;
        currSimStep = 0;
    }

    uws_config to_c() const {
        uws_config c{};
        c.grid_res[0] = gridRes.x; c.grid_res[1] = gridRes.y; c.grid_res[2] = gridRes.z;
        c.dx = dx; c.dt = dt; c.mu = mu; c.den = den;
        c.external_force[0] = externalForce.x; c.external_force[1] = externalForce.y;
        c.external_force[2] = externalForce.z;
        const VelBndCond vb[6] = {velBndCondX0, velBndCondXn, velBndCondY0, velBndCondYn, velBndCondZ0, velBndCondZn};
        const PresBndCond pb[6] = {presBndCondX0, presBndCondXn, presBndCondY0, presBndCondYn, presBndCondZ0,
                                   presBndCondZn};
        for (int i = 0; i < 6; ++i) { c.vel_bnd[i] = (int32_t)vb[i]; c.pres_bnd[i] = (int32_t)pb[i]; }
        c.fvm_solver_type = (int32_t)fvmSolverType;
        c.piso_num_correctors = PISONumCorrectors;
        c.turbulence_model = (int32_t)turbulenceModel;
        c.smagorinsky_constant = smagorinskyConstant;
        c.vel_max_num_iter = velMaxNumIter;
        c.pres_max_num_iter = presMaxNumIter;
        c.cal_residual = calResidual ? 1 : 0;
        c.vel_residual_check_interval = velResidualCheckInterval;
        c.vel_tolerance = velTolerance;
        c.pres_residual_check_interval = presResidualCheckInterval;
        c.pres_tolerance = presTolerance;
        c.wall_slots = wallSlots;
        c.wall_model = wallModel;
        c.wall_z0 = wallZ0;
        c.precision = precision;
        c.refine_rounds = refineRounds;
        c.refine_sweeps = refineSweeps;
        c.compute_norms = computeNorms ? 1 : 0;
        c.grid_type = (int32_t)gridType;
        return c;
    }
};

// RuntimeConfig (RuntimeConfig.cs:6-17)
struct RuntimeConfig {
    float flowFieldOrientation = 0.0f;
    float windSpeed = 10.0f;
    double llhCoord[3] = {0, 0, 0};
    explicit RuntimeConfig(const FluidSimConfig& cf) : windSpeed(cf.velX0.x) {}
};

// IFvmSolver (IFvmSolver.cs:3-24).  Fields cross as host vectors: velocity
// [z][y][x][4] (the reference's float4 texel, w = 0), pressure [z][y][x],
// flags [z][y][x].
class IFvmSolver {
public:
    virtual ~IFvmSolver() = default;
    virtual bool Step() = 0;
    virtual void InitFlags(const std::vector<int32_t>& flags) = 0;
    virtual void InitVelPresFields() = 0;
    virtual void InitVelPresFieldsFromBackGroundFlow(const FluidSimConfig& cfBackGround,
                                                     const std::vector<float>& velBackGround,
                                                     const std::vector<float>& presBackGround) = 0;
    virtual void SetFixedValueVelBndCond() = 0;
    virtual void ChangePhysFieldPos() = 0;
    virtual void ChangePhysDomainSize() = 0;
    virtual std::vector<float> GetVelField() = 0;
    virtual std::vector<float> GetPresField() = 0;
    virtual std::vector<int32_t> GetFlagField() = 0;
};

// FvmSolverGpu's behaviour on MI355X (FvmSolverGpu.cs:173-222 constructor).
class FvmSolverHip final : public IFvmSolver {
public:
    FvmSolverHip(FluidSimConfig& cf, RuntimeConfig& rcf) : cf_(cf), rcf_(rcf) {
        cf_.Init();
        const uws_config c = cf_.to_c();
        check(uws_create(&c, &h_), "uws_create");
        n_ = (int64_t)cf_.gridRes.x * cf_.gridRes.y * cf_.gridRes.z;
        if (cf_.gridType == GridType::CollNonUniform) InitMeshBuffers();
        InitVelPresFields();
        if (cf_.solidType == SolidType::Box) check(uws_init_box(h_), "uws_init_box");
        SetFixedValueVelBndCond();
    }
    ~FvmSolverHip() override {
        if (h_) uws_destroy(h_);
    }
    FvmSolverHip(const FvmSolverHip&) = delete;
    FvmSolverHip& operator=(const FvmSolverHip&) = delete;

    // InitMeshBuffers (FvmSolverGpu.cs:682-788): the hard-coded stretched mesh,
    // a uniform block of 50 cells per axis with 25 cells growing by 1.08 on each
    // side along x and z; throws where the reference's tables would overflow
    void InitMeshBuffers() {
        const int32_t res[3] = {cf_.gridRes.x, cf_.gridRes.y, cf_.gridRes.z};
        const int32_t uniform[3] = {50, 50, 50}, one_side[3] = {25, 0, 25};
        std::vector<float> pos[3];
        for (int a = 0; a < 3; ++a) {
            pos[a].resize((size_t)res[a] + 1);
            check(uws_stretched_face_positions(res[a], cf_.dx, uniform[a], one_side[a], 1.08f, pos[a].data()),
                  "InitMeshBuffers");
        }
        SetFacePositions(pos[0], pos[1], pos[2]);
    }
    // stretched grid: any strictly increasing face coordinates (X+1, Y+1, Z+1)
    void SetFacePositions(const std::vector<float>& fx, const std::vector<float>& fy, const std::vector<float>& fz) {
        if (fx.size() != (size_t)cf_.gridRes.x + 1 || fy.size() != (size_t)cf_.gridRes.y + 1 ||
            fz.size() != (size_t)cf_.gridRes.z + 1)
            throw Error(UWS_ERR_INVALID_ARG, "SetFacePositions");
        check(uws_set_face_positions(h_, fx.data(), fy.data(), fz.data()), "SetFacePositions");
    }

    // whole steps only: always ready for rendering afterwards
    bool Step() override {
        check(uws_step(h_, 1), "uws_step");
        int64_t s = 0;
        double t = 0;
        check(uws_get_time(h_, &s, &t), "uws_get_time");
        cf_.currSimStep = (int32_t)s;
        cf_.currPhyTime = (float)t;
        return true;
    }
    void InitFlags(const std::vector<int32_t>& flags) override {
        check(uws_init_flags(h_, flags.data(), (int64_t)flags.size()), "uws_init_flags");
    }
    void InitVelPresFields() override {
        if (cf_.inflowType == InflowType::PowerLaw)
            check(uws_init_vel_pres_fields(h_, UWS_INFLOW_CONSTANT, 0.f), "uws_init_vel_pres_fields");
        else
            check(uws_init_vel_pres_fields(h_, (int32_t)cf_.inflowType, rcf_.windSpeed), "uws_init_vel_pres_fields");
    }
    // velTexBackGround: the background's velocity [z][y][x][4] (its GetVelField()),
    // presTexBackGround: its pressure [z][y][x] (or empty)
    void InitVelPresFieldsFromBackGroundFlow(const FluidSimConfig& cfBackGround, const std::vector<float>& velTexBackGround,
                                             const std::vector<float>& presTexBackGround) override {
        uws_background bg{};
        bg.grid_res[0] = cfBackGround.gridRes.x; bg.grid_res[1] = cfBackGround.gridRes.y;
        bg.grid_res[2] = cfBackGround.gridRes.z;
        bg.dx = cfBackGround.dx;
        bg.phys_field_pos[0] = cfBackGround.physFieldPos.x; bg.phys_field_pos[1] = cfBackGround.physFieldPos.y;
        bg.phys_field_pos[2] = cfBackGround.physFieldPos.z;
        const int64_t nb = (int64_t)bg.grid_res[0] * bg.grid_res[1] * bg.grid_res[2];
        if ((int64_t)velTexBackGround.size() != 4 * nb ||
            (!presTexBackGround.empty() && (int64_t)presTexBackGround.size() != nb))
            throw Error(UWS_ERR_INVALID_ARG, "InitVelPresFieldsFromBackGroundFlow");
        const float pos[3] = {cf_.physFieldPos.x, cf_.physFieldPos.y, cf_.physFieldPos.z};
        check(uws_init_vel_pres_fields_from_background(h_, pos, rcf_.flowFieldOrientation, &bg,
                                                       velTexBackGround.data(), 4,
                                                       presTexBackGround.empty() ? nullptr : presTexBackGround.data()),
              "uws_init_vel_pres_fields_from_background");
    }
    void SetFixedValueVelBndCond() override {
        const float vz[3] = {cf_.velZn.x, cf_.velZn.y, cf_.velZn.z};
        check(uws_set_fixed_value_vel_bnd_cond(h_, (int32_t)cf_.inflowType, rcf_.windSpeed, vz, cf_.inflowRefHeight,
                                               cf_.inflowAlpha),
              "uws_set_fixed_value_vel_bnd_cond");
    }
    void ChangePhysFieldPos() override { InitVelPresFields(); }
    // FvmSolverGpu.ChangePhysDomainSize (FvmSolverGpu.cs:1034-1046): the config's
    // re-derived gridRes and dx, then InitVelPresFields.  The fields are
    // reallocated for the new grid; flags and boundary textures survive where
    // gridRes is unchanged; after a grid change the mask is all fluid (the box
    // is re-applied for SolidType.Box, a model mask comes from InitFlags as in
    // FvmSolverController.cs:869-874), the stretched mesh is rebuilt and the
    // inflow texture refilled.
    void ChangePhysDomainSize() override {
        const int32_t old[3] = {cf_.gridRes.x, cf_.gridRes.y, cf_.gridRes.z};
        cf_.Init();
        const uws_config c = cf_.to_c();
        check(uws_change_phys_domain_size(h_, &c), "uws_change_phys_domain_size");
        n_ = (int64_t)cf_.gridRes.x * cf_.gridRes.y * cf_.gridRes.z;
        const bool changed = old[0] != cf_.gridRes.x || old[1] != cf_.gridRes.y || old[2] != cf_.gridRes.z;
        if (cf_.gridType == GridType::CollNonUniform) InitMeshBuffers();
        InitVelPresFields();
        if (changed) {
            if (cf_.solidType == SolidType::Box) check(uws_init_box(h_), "uws_init_box");
            SetFixedValueVelBndCond();
        }
    }
    std::vector<float> GetVelField() override {
        std::vector<float> v((size_t)n_ * 4);
        check(uws_get_velocity(h_, v.data(), 4, n_), "uws_get_velocity");
        return v;
    }
    std::vector<float> GetPresField() override {
        std::vector<float> p((size_t)n_);
        check(uws_get_pressure(h_, p.data(), n_), "uws_get_pressure");
        return p;
    }
    std::vector<int32_t> GetFlagField() override {
        std::vector<int32_t> f((size_t)n_);
        check(uws_get_flags(h_, f.data(), n_), "uws_get_flags");
        return f;
    }

    uws_handle handle() const { return h_; }
    int64_t numCells() const { return n_; }

    // ---- extensions used by the differentiation API ----
    int64_t BndSize(int face) const {
        const int32_t X = cf_.gridRes.x, Y = cf_.gridRes.y, Z = cf_.gridRes.z;
        return (int64_t)(face < 2 ? Y : X) * (face < 4 ? Z : Y);
    }
    // one boundary texture, [j][i][3] doubles (exact for FP64 runs)
    void SetBndField(int face, const std::vector<double>& xyz) {
        check(uws_set_bnd_field_f64(h_, face, xyz.data(), (int64_t)xyz.size() / 3), "uws_set_bnd_field_f64");
    }
    void SetSmagorinskyConstant(double cs) {
        check(uws_set_smagorinsky_constant(h_, cs), "uws_set_smagorinsky_constant");
        cf_.smagorinskyConstant = (float)cs;
    }
    // u^n as [cell][3] doubles
    std::vector<double> GetVelFieldF64() {
        std::vector<double> v((size_t)n_ * 3);
        check(uws_get_field_f64(h_, UWS_FIELD_VEL, v.data(), (int64_t)v.size()), "uws_get_field_f64");
        return v;
    }

private:
    FluidSimConfig& cf_;
    RuntimeConfig& rcf_;
    uws_handle h_ = nullptr;
    int64_t n_ = 0;
};

// A mask written by the reference's FlagExporter (FlagExporter.cs:13-34): the
// flat index z*X*Y + y*X + x of every solid cell, one per line; returns the
// flag array InitFlags takes.
inline std::vector<int32_t> ReadFlagFile(const std::string& path, int64_t ncells) {
    std::vector<int32_t> flags((size_t)ncells, 0);
    FILE* fh = std::fopen(path.c_str(), "r");
    if (!fh) throw Error(UWS_ERR_INVALID_ARG, "ReadFlagFile: cannot open " + path);
    long long i = 0;
    while (std::fscanf(fh, "%lld", &i) == 1) {
        if (i < 0 || i >= ncells) {
            std::fclose(fh);
            throw Error(UWS_ERR_INVALID_ARG, "ReadFlagFile: index outside the grid");
        }
        flags[(size_t)i] = 1;
    }
    std::fclose(fh);
    return flags;
}

// Discrete adjoint of nsteps steps of a FvmSolverHip (uws_adjoint_*): the
// forward/backward API a host exposes for gradient-based design.  Forward()
// steps the solver (which ends at the final state) keeping a checkpoint every
// ckptEvery steps; Backward(dJ/du_T, dJ/dp_T) reverses them; the gradients
// w.r.t. the initial state, any boundary texture and Cs are then read.  The
// solver must outlive the adjoint.
class FvmAdjointHip {
public:
    FvmAdjointHip(FvmSolverHip& s, int nsteps, int ckptEvery = 8) : s_(s) {
        check(uws_adjoint_create(s.handle(), nsteps, ckptEvery, &a_), "uws_adjoint_create");
    }
    ~FvmAdjointHip() {
        if (a_) uws_adjoint_destroy(a_);
    }
    FvmAdjointHip(const FvmAdjointHip&) = delete;
    FvmAdjointHip& operator=(const FvmAdjointHip&) = delete;

    void Forward() { check(uws_adjoint_forward(a_), "uws_adjoint_forward"); }
    // dJdu: [cell][3]; dJdp: [cell] or empty for zero
    void Backward(const std::vector<double>& dJdu, const std::vector<double>& dJdp = {}) {
        check(uws_adjoint_backward(a_, dJdu.data(), (int64_t)dJdu.size(), dJdp.empty() ? nullptr : dJdp.data(),
                                   (int64_t)dJdp.size()),
              "uws_adjoint_backward");
    }
    std::vector<double> GradVelocity0() { return grad(UWS_GRAD_VEL0, 0, 3 * s_.numCells()); }
    std::vector<double> GradPressure0() { return grad(UWS_GRAD_PRES0, 0, s_.numCells()); }
    std::vector<double> GradBnd(int face) { return grad(UWS_GRAD_BND, face, 3 * s_.BndSize(face)); }
    double GradCs() { return grad(UWS_GRAD_CS, 0, 1)[0]; }
    int64_t DeviceBytes() {
        int64_t b = 0;
        check(uws_adjoint_memory(a_, &b), "uws_adjoint_memory");
        return b;
    }

private:
    std::vector<double> grad(int what, int face, int64_t n) {
        std::vector<double> g((size_t)n);
        check(uws_adjoint_get_grad(a_, what, face, g.data(), n), "uws_adjoint_get_grad");
        return g;
    }
    FvmSolverHip& s_;
    uws_adjoint_handle a_ = nullptr;
};

}  // namespace uws
