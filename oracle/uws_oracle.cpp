/*
 * uws_oracle.cpp -- CPU restatement of UnityWindSimulation's collocated
 * finite-volume solver step.  TEST INFRASTRUCTURE ONLY: the parity tests and
 * the smoke check compare the HIP library against it, and the bench's
 * CPU-baseline leg times it.  Nothing in the product library links or calls
 * it.
 *
 * Every kernel is restated from the reference shader text in the same
 * arithmetic order (left-to-right float expressions, one accumulator per
 * term in the reference's face order), templated on the scalar type so the
 * same code gives the FP32 oracle and an FP64 reference run.  Built with
 * -ffp-contract=off so no FMA contraction changes the rounding.
 *
 * The buffers keep the reference's texture names and its C# reference swaps
 * (FvmSolverGpu.cs:1050-1121, 1238-1282, 1367-1439, 1482-1554), so the
 * observable state after Step() is the reference's.
 *
 * Choices where the shader text leaves the arithmetic open (HLSL compiler
 * semantics, no executable reference here):
 *  - pow(x, 2) is taken as x*x (WallFunc.compute:150-183).
 *  - literals such as 0.25, 0.5, 2.0 are float in float expressions.
 *  - an R32 texture read as float4 yields (x, 0, 0, 1): with wall_slots == 1
 *    only the .x wall-viscosity slot survives (FvmSolverGpu.cs:304).
 *  - CSPresNormalization reads its reference cell while other threads write
 *    it (FvmColl.compute:2864-2880); here the reference value is read once,
 *    then subtracted (the race-free meaning).
 *  - the 30 ms time slicing of Step() is not modelled (whole steps only).
 *  - stretched grid (FvmCollNonUniform.compute): the momentum kernels read
 *    velFaceFluxField, which the pressure precompute writes, although the
 *    host's bindings of it for those two kernels are commented out
 *    (FvmSolverGpu.cs:1160, 1196); here they read it, as the shader text
 *    says.  The texture starts at zero and is not reset by the field
 *    initialisation.  The LES and wall-function kernels stay the uniform ones
 *    with the uniform dx, as the host binds them (FvmSolverGpu.cs:522-552).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>
#include <type_traits>

#include "../include/uws.h"


// Cell loops of the step's kernels run in parallel over (z, y) when built
// with OpenMP: every iteration writes only its own cell (gather form), so the
// result does not depend on the thread count (small grids stay serial).
#define UWSO_PRAGMA(x) _Pragma(#x)
#define UWSO_PARALLEL_CELLS UWSO_PRAGMA(omp parallel for collapse(2) schedule(static) if (N >= 65536))

namespace {

enum { UWSO_FIELD_VEL_ITER = 100, UWSO_FIELD_PC_ITER = 101 };

// Planted mistakes, one per pinned function, compiled in only with
// -DUWSO_MUTANT=<n> by the test that shows each pin fails on its mistake
// (tests/test_oracle_assembled.py); the oracle proper is built without it.
#ifndef UWSO_MUTANT
#define UWSO_MUTANT 0
#endif
#define MUT(n) (UWSO_MUTANT == (n))

template <class T> struct V3 { T x, y, z; };

template <class T> static inline T sq(T v) { return v * v; }

template <class T> struct Oracle {
    uws_config cf;
    int X, Y, Z;
    int64_t N;
    // FluidSimConfig.Init() derived parameters (FluidSimConfig.cs:325-341), in T.
    T dx, dt, ds, dv, nu, den, cs;
    T fext[3];
    int velBnd[6], presBnd[6];
    bool useLES, piso;
    // stretched grid (GridType.CollNonUniform): face coordinates of each axis
    // (the facePosX/Y/Z float buffers, FvmSolverGpu.cs:682-788) and the
    // velFaceFluxField texture of (X+1)(Y+1)(Z+1) texels (FvmSolverGpu.cs:275)
    bool nonuniform = false;
    std::vector<float> facePos[3];
    std::vector<V3<T>> faceFlux;

    std::vector<int32_t> flag;
    // texture storage; the pointers below are swapped like the C# references
    std::vector<V3<T>> tex_vel[3];   // velFieldTex, velFieldLastTimeTex, velFieldLastIterTex
    std::vector<V3<T>> velCorrect, Dfield, bfield, AodU, faceEddyVis;
    std::vector<T> presLastTime, tex_pc[2], DpField, bpField, eddyVis;
    std::vector<V3<T>> bnd[6];
    V3<T>* velField; V3<T>* velFieldLastTime; V3<T>* velFieldLastIter;
    T* presCorrectField; T* presCorrectFieldLastIter;
    int64_t step_count = 0;
    double phys_time = 0.0;
    T last_vel_residual = 0, last_pres_residual = 0;
    int last_vel_sweeps = 0, last_pres_sweeps = 0;  // sweeps run by the last VelPredict / PresCorrect

    int64_t idx(int x, int y, int z) const { return (int64_t)z * X * Y + (int64_t)y * X + x; }
    bool InRegion(int x, int y, int z) const {
        return x >= 0 && x < X && y >= 0 && y < Y && z >= 0 && z < Z;
    }
    bool IsFluid(int x, int y, int z) const { return flag[idx(x, y, z)] == UWS_CELL_FLUID; }
    bool IsSolid(int x, int y, int z) const { return flag[idx(x, y, z)] == UWS_CELL_SOLID; }
    bool InRegionAndIsFluid(int x, int y, int z) const { return InRegion(x, y, z) && IsFluid(x, y, z); }
    bool InRegionAndIsSolid(int x, int y, int z) const { return InRegion(x, y, z) && IsSolid(x, y, z); }

    // boundary texture element (i, j): X faces (y, z), Y faces (x, z), Z faces (x, y)
    int bndW(int f) const { return f < 2 ? Y : X; }
    int bndH(int f) const { return f < 4 ? Z : Y; }
    const V3<T>& B(int f, int i, int j) const { return bnd[f][(int64_t)j * bndW(f) + i]; }

    explicit Oracle(const uws_config& c) : cf(c) {
        X = c.grid_res[0]; Y = c.grid_res[1]; Z = c.grid_res[2];
        N = (int64_t)X * Y * Z;
        // Derived exactly as FluidSimConfig.Init (float arithmetic for T=float).
        dx = (T)c.dx; ds = dx * dx; dv = dx * dx * dx;
        dt = (T)c.dt; den = (T)c.den; nu = (T)c.mu / den;
        cs = (T)c.smagorinsky_constant;
        for (int i = 0; i < 3; ++i) fext[i] = (T)c.external_force[i];
        for (int i = 0; i < 6; ++i) { velBnd[i] = c.vel_bnd[i]; presBnd[i] = c.pres_bnd[i]; }
        useLES = c.turbulence_model == UWS_TURB_SMAGORINSKY;
        piso = c.fvm_solver_type == UWS_PISO;
        nonuniform = c.grid_type == UWS_GRID_COLL_NONUNIFORM;
        const int res[3] = {X, Y, Z};
        for (int a = 0; a < 3; ++a) {
            facePos[a].resize(res[a] + 1);
            for (int i = 0; i <= res[a]; ++i) facePos[a][i] = (float)i * c.dx;
        }
        if (nonuniform) faceFlux.assign((int64_t)(X + 1) * (Y + 1) * (Z + 1), V3<T>{0, 0, 0});
        flag.assign(N, UWS_CELL_FLUID);
        V3<T> z0{0, 0, 0};
        for (auto& t : tex_vel) t.assign(N, z0);
        velCorrect.assign(N, z0); Dfield.assign(N, z0); bfield.assign(N, z0);
        AodU.assign(N, z0); faceEddyVis.assign(N, z0);
        presLastTime.assign(N, 0); tex_pc[0].assign(N, 0); tex_pc[1].assign(N, 0);
        DpField.assign(N, 0); bpField.assign(N, 0); eddyVis.assign(N, 0);
        for (int f = 0; f < 6; ++f) bnd[f].assign((int64_t)bndW(f) * bndH(f), z0);
        velField = tex_vel[0].data(); velFieldLastTime = tex_vel[1].data(); velFieldLastIter = tex_vel[2].data();
        presCorrectField = tex_pc[0].data(); presCorrectFieldLastIter = tex_pc[1].data();
    }

    // ------------------------------------------------------------------
    // Init kernels (InitShaderColl.compute)
    // ------------------------------------------------------------------
    // CSInitModelFlag (InitShaderColl.compute:345-364)
    void InitModelFlag(const int32_t* flags) {
        for (int64_t i = 0; i < N; ++i) {
            if ((double)flags[i] > 0.5) {
                flag[i] = UWS_CELL_SOLID;
                velField[i] = velFieldLastIter[i] = velFieldLastTime[i] = V3<T>{0, 0, 0};
            } else {
                flag[i] = UWS_CELL_FLUID;
            }
        }
    }
    // CSInitBoxFlag (InitShaderColl.compute:302-343): hard-coded box
    void InitBoxFlag() {
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x)
            if (x >= 20 && x < 24 && y >= 0 && y < 8 && z >= 18 && z < 22) {
                int64_t i = idx(x, y, z);
                flag[i] = UWS_CELL_SOLID;
                velField[i] = velFieldLastIter[i] = velFieldLastTime[i] = V3<T>{0, 0, 0};
            }
    }
    // 0.5 * (facePosY[y] + facePosY[y + 1]) (InitShaderColl.compute:141, 295)
    T YMid(int y) const { return (T)0.5f * ((T)facePos[1][y] + (T)facePos[1][y + 1]); }
    // CSInitVelPresFields (:94-110) / CSInitLogVelPresFields (:112-132)
    void InitVelPresFields(int inflow, T wind) {
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            int64_t i = idx(x, y, z);
            V3<T> v{wind, 0, 0};
            if (inflow == UWS_INFLOW_LOG) {
                const T karman = (T)0.4f, rough = (T)1.0f, refH = (T)10.0f;
                // stretched grid: the cell-centre height (CSInitLogVelPresFieldsNonUniform, :134-154)
                T yPhys = nonuniform ? YMid(y) : (T)y * dx;
                T fricVel = karman * wind / std::log((refH + rough) / rough);
                T vel = fricVel / karman * std::log((yPhys + rough) / rough);
                v = V3<T>{vel, 0, 0};
            }
            velField[i] = velFieldLastTime[i] = velFieldLastIter[i] = v;
            velCorrect[i] = V3<T>{0, 0, 0};
            presLastTime[i] = 0; presCorrectField[i] = 0; presCorrectFieldLastIter[i] = 0;
        }
    }
    // CSSetFixedValueVelBndCond (:252-265), CSSetLogVelBndCond (:272-285); power law: extension
    void SetFixedValueVelBndCond(int inflow, T wind, const float* velZn, T refHeight, T alpha) {
        if (inflow == UWS_INFLOW_CONSTANT) {
            for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y)
                bnd[UWS_X0][(int64_t)z * Y + y] = V3<T>{wind, 0, 0};
            V3<T> vz{0, 0, 0};
            if (velZn) vz = V3<T>{(T)velZn[0], (T)velZn[1], (T)velZn[2]};
            for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x)
                bnd[UWS_ZN][(int64_t)y * X + x] = vz;
        } else if (inflow == UWS_INFLOW_LOG) {
            const T karman = (T)0.4f, rough = (T)1.0f, refH = (T)10.0f;
            for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) {
                // stretched grid: CSSetLogVelBndCondNonUniform (:287-300)
                T yPhys = nonuniform ? YMid(y) : ((T)y + (T)0.5f) * dx;
                T fricVel = karman * wind / std::log((refH + rough) / rough);
                T vel = fricVel / karman * std::log((yPhys + rough) / rough);
                bnd[UWS_X0][(int64_t)z * Y + y] = V3<T>{vel, 0, 0};
            }
        } else {
            for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) {
                T yPhys = nonuniform ? YMid(y) : ((T)y + (T)0.5f) * dx;
                T vel = wind * (T)std::pow((double)(yPhys / refHeight), (double)alpha);
                bnd[UWS_X0][(int64_t)z * Y + y] = V3<T>{vel, 0, 0};
            }
        }
    }

    // ------------------------------------------------------------------
    // Wall function (WallFunc.compute)
    // ------------------------------------------------------------------
    // LogLawWallFunc (WallFunc.compute:77-123); rough variant is an extension
    T LogLawWallFunc(T uP) const {
        const T EConst = (T)9.8f, k = (T)0.41f;
        T C = EConst * (dx / (T)2.0f) / nu;
        T eddyVis = 0;
        T uTau = (T)1.0f;
        uP = std::fabs(uP);
        if (uP < (T)1e-5f) {
            uTau = 0;
        } else {
            if (cf.wall_model == UWS_WALL_ROUGH) {
                uTau = k * uP / std::log((dx / (T)2.0f) / (T)cf.wall_z0);
            } else {
                for (int i = 0; i < 10; ++i) {
                    T f = uTau * std::log(C * uTau) - k * uP;
                    T df = std::log(C * uTau) + (T)1.0f;
                    if (std::fabs(df) < (T)1e-6f) break;
                    uTau = uTau - f / df;
                    if (uTau < (T)1e-5f) uTau = (T)1e-5f;
                }
            }
            T tauWall = uTau * uTau * den;
            T strainRate = uP / (dx / (T)2.0f);
            eddyVis = tauWall / (strainRate * den) - nu;
            if (eddyVis < 0) eddyVis = 0;
        }
        return eddyVis;
    }
    // CSCalLogLawWallFunc (WallFunc.compute:125-190)
    void CalLogLawWallFunc() {
        const V3<T>* u = velFieldLastTime;
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            V3<T> fev{0, 0, 0};
            if (IsFluid(x, y, z)) {
                const V3<T>& uP = u[idx(x, y, z)];
                if (InRegionAndIsSolid(x - 1, y, z))
                    fev.x += LogLawWallFunc(std::sqrt(sq(uP.y) + sq(uP.z)));
                if (InRegionAndIsSolid(x, y - 1, z) ||
                    (!MUT(5) && !InRegion(x, y - 1, z) && velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE))
                    fev.y += LogLawWallFunc(std::sqrt(sq(uP.x) + sq(uP.z)));
                if (InRegionAndIsSolid(x, y, z - 1))
                    fev.z += LogLawWallFunc(std::sqrt(sq(uP.x) + sq(uP.y)));
            } else if (IsSolid(x, y, z)) {
                if (InRegionAndIsFluid(x - 1, y, z)) {
                    const V3<T>& uW = u[idx(x - 1, y, z)];
                    fev.x += LogLawWallFunc(std::sqrt(sq(MUT(4) ? uW.x : uW.y) + sq(uW.z)));
                }
                if (InRegionAndIsFluid(x, y - 1, z)) {
                    const V3<T>& uS = u[idx(x, y - 1, z)];
                    fev.y += LogLawWallFunc(std::sqrt(sq(uS.x) + sq(uS.z)));
                }
                if (InRegionAndIsFluid(x, y, z - 1)) {
                    const V3<T>& uD = u[idx(x, y, z - 1)];
                    fev.z += LogLawWallFunc(std::sqrt(sq(uD.x) + sq(uD.y)));
                }
            }
            // float4 written into an R32 texture keeps only .x
            if (cf.wall_slots == 1) { fev.y = 0; fev.z = 0; }
            faceEddyVis[idx(x, y, z)] = fev;
        }
    }

    // ------------------------------------------------------------------
    // LES (LES.compute)
    // ------------------------------------------------------------------
    // Get{E,W,N,S,U,D}FaceVelLastTime (LES.compute:104-258); dir 0..5 = E W N S U D
    V3<T> FaceVel(int x, int y, int z, int dir) const {
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int face[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        const V3<T>* u = velFieldLastTime;
        int nx = x + off[dir][0], ny = y + off[dir][1], nz = z + off[dir][2];
        const V3<T>& uP = u[idx(x, y, z)];
        if (InRegion(nx, ny, nz)) {
            if (IsFluid(nx, ny, nz)) {
                const V3<T>& uN = u[idx(nx, ny, nz)];
                return V3<T>{(T)0.5f * (uP.x + uN.x), (T)0.5f * (uP.y + uN.y), (T)0.5f * (uP.z + uN.z)};
            }
            return V3<T>{0, 0, 0};
        }
        int f = face[dir];
        int t = velBnd[f];
        if (t == UWS_VEL_FIXED_VALUE) {
            int i, j;
            if (f < 2) { i = y; j = z; } else if (f < 4) { i = x; j = z; } else { i = x; j = y; }
            return B(f, i, j);
        } else if (t == UWS_VEL_ZERO_GRAD) {
            return uP;
        } else if (t == UWS_VEL_SYMMETRY) {
            V3<T> v = uP;
            if (MUT(14)) { if (dir < 2) v.y = 0; else if (dir < 4) v.z = 0; else v.x = 0; return v; }
            if (dir < 2) v.x = 0; else if (dir < 4) v.y = 0; else v.z = 0;
            return v;
        }
        return V3<T>{0, 0, 0};
    }
    // velocity gradient of a cell from its face velocities: g[i][j] = d u_i / d x_j
    void CellGrad(int x, int y, int z, T g[3][3]) const {
        V3<T> e = FaceVel(x, y, z, 0), w = FaceVel(x, y, z, 1), n = FaceVel(x, y, z, 2);
        V3<T> s = FaceVel(x, y, z, 3), u = FaceVel(x, y, z, 4), d = FaceVel(x, y, z, 5);
        g[0][0] = (e.x - w.x) / dx; g[0][1] = (n.x - s.x) / dx; g[0][2] = (u.x - d.x) / dx;
        g[1][0] = (e.y - w.y) / dx; g[1][1] = (n.y - s.y) / dx; g[1][2] = (u.y - d.y) / dx;
        g[2][0] = (e.z - w.z) / dx; g[2][1] = (n.z - s.z) / dx; g[2][2] = (u.z - d.z) / dx;
    }
    // CSCalEddyVis (LES.compute:267-324)
    void CalEddyVis() {
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            T g[3][3];
            CellGrad(x, y, z, g);
            T Sxx = g[0][0], Syy = g[1][1], Szz = g[2][2];
            T Sxy = (T)0.5f * (g[0][1] + g[1][0]);
            T Syz = (T)0.5f * (g[1][2] + g[2][1]);
            T Szx = (T)0.5f * (g[2][0] + g[0][2]);
            T Sabs = std::sqrt((T)2.0f * (Sxx * Sxx + Syy * Syy + Szz * Szz + (T)2.0f * Sxy * Sxy +
                                          (T)2.0f * Syz * Syz + (T)2.0f * Szx * Szx));
            eddyVis[idx(x, y, z)] = (cs * cs) * (dx * dx) * Sabs;
        }
    }
    // face viscosities with eddy viscosity (FvmColl.compute:1596-1635, LES.compute:643-679)
    void FaceVis(int x, int y, int z, T vis[6]) const {
        const int64_t P = idx(x, y, z);
        T base = nu + eddyVis[P];
        for (int f = 0; f < 6; ++f) vis[f] = base;
        if (InRegionAndIsFluid(x + 1, y, z)) vis[0] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x + 1, y, z)]);
        else if (InRegionAndIsSolid(x + 1, y, z)) vis[0] = nu + faceEddyVis[idx(x + 1, y, z)].x;
        if (InRegionAndIsFluid(x - 1, y, z)) vis[1] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x - 1, y, z)]);
        else if (InRegionAndIsSolid(x - 1, y, z)) vis[1] = nu + faceEddyVis[MUT(3) ? idx(x - 1, y, z) : P].x;
        if (InRegionAndIsFluid(x, y + 1, z)) vis[2] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x, y + 1, z)]);
        else if (InRegionAndIsSolid(x, y + 1, z)) vis[2] = nu + faceEddyVis[idx(x, y + 1, z)].y;
        if (InRegionAndIsFluid(x, y - 1, z)) vis[3] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x, y - 1, z)]);
        else if (InRegionAndIsSolid(x, y - 1, z) || (!InRegion(x, y - 1, z) && velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE))
            vis[3] = nu + faceEddyVis[P].y;
        if (InRegionAndIsFluid(x, y, z + 1)) vis[4] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x, y, z + 1)]);
        else if (InRegionAndIsSolid(x, y, z + 1)) vis[4] = nu + faceEddyVis[idx(x, y, z + 1)].z;
        if (InRegionAndIsFluid(x, y, z - 1)) vis[5] = nu + (T)0.5f * (eddyVis[P] + eddyVis[idx(x, y, z - 1)]);
        else if (InRegionAndIsSolid(x, y, z - 1)) vis[5] = nu + faceEddyVis[P].z;
    }
    // CSCalLesDeferCorrectTerm (LES.compute:326-687)
    void CalLesDeferCorrectTerm() {
        const V3<T>* u = velFieldLastTime;
        const T hdx = (T)0.5f * dx;
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const V3<T>& uP = u[idx(x, y, z)];
            T g[3][3];
            CellGrad(x, y, z, g);
            // face gradients of the face-normal component: fg[face][j]
            T fg[6][3] = {{0}};
            // face order E W N S U D; normal component c = face/2; neighbour offsets
            static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
            static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                bool plus = (f % 2) == 0;
                T uPc = c == 0 ? uP.x : (c == 1 ? uP.y : uP.z);
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) {
                        T gn[3][3];
                        CellGrad(nx, ny, nz, gn);
                        for (int j = 0; j < 3; ++j)
                            fg[f][j] = MUT(9) ? (T)0.5f * (g[j][c] + gn[j][c]) : (T)0.5f * (g[c][j] + gn[c][j]);
                    } else {
                        fg[f][c] = (plus != MUT(15)) ? -uPc / hdx : uPc / hdx;
                    }
                } else {
                    int t = velBnd[bface[f]];
                    if (t == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        int bf = bface[f];
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        T ubc = c == 0 ? ub.x : (c == 1 ? ub.y : ub.z);
                        fg[f][c] = plus ? (ubc - uPc) / hdx : (uPc - ubc) / hdx;
                    } else if (t == UWS_VEL_ZERO_GRAD || t == UWS_VEL_SYMMETRY) {
                        for (int j = 0; j < 3; ++j) if (j != c || MUT(16)) fg[f][j] = g[c][j];
                    }
                }
            }
            T vis[6];
            FaceVis(x, y, z, vis);
            V3<T> r;
            T* rr[3] = {&r.x, &r.y, &r.z};
            for (int i = 0; i < 3; ++i) {
                *rr[i] = ds * (vis[0] * fg[0][i] - (MUT(10) ? -vis[1] : vis[1]) * fg[1][i] + vis[2] * fg[2][i] -
                               vis[3] * fg[3][i] + vis[4] * fg[4][i] - vis[5] * fg[5][i]);
            }
            bfield[idx(x, y, z)] = r;
        }
    }

    // ------------------------------------------------------------------
    // Momentum predictor (FvmColl.compute:1291-2069)
    // ------------------------------------------------------------------
    // CSVelPredictPreCompute (FvmColl.compute:1291-1859)
    void VelPredictPreCompute() {
        if (nonuniform) { NuVelPredictPreCompute(); return; }
        const V3<T>* un = velFieldLastTime;
        const T* p = presLastTime.data();
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T Dx = dv / dt, Dy = dv / dt, Dz = dv / dt;
            T bU = 0, bV = 0, bW = 0;
            const V3<T>& uP = un[P];
            // advection, face order E W N S U D (:1314-1482)
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                T uPn = c == 0 ? uP.x : (c == 1 ? uP.y : uP.z);
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) {
                        const V3<T>& uN = un[idx(nx, ny, nz)];
                        T uNn = c == 0 ? uN.x : (c == 1 ? uN.y : uN.z);
                        T t = ds * (T)0.25f * (uPn + uNn);
                        if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
                    }
                } else {
                    int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        T ubn = c == 0 ? ub.x : (c == 1 ? ub.y : ub.z);
                        if (plus) { bU -= ds * ub.x * ubn; bV -= ds * ub.y * ubn; bW -= ds * ub.z * ubn; }
                        else { bU += ds * ub.x * ubn; bV += ds * ub.y * ubn; bW += ds * ub.z * ubn; }
                    } else if (velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                        T t = ds * uPn;
                        if (plus != MUT(11)) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
                    }
                }
            }
            // pressure face values (:1484-1591)
            T pf[6];
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) pf[f] = (T)0.5f * (p[P] + p[idx(nx, ny, nz)]);
                    else pf[f] = p[P];
                } else {
                    int t = presBnd[bface[f]];
                    if (t == UWS_PRES_FIXED_VALUE) pf[f] = 0;
                    else pf[f] = p[P];  // zero gradient and symmetry
                }
            }
            bU += -((T)1.0f / den) * ds * (pf[0] - pf[1]);
            bV += -((T)1.0f / den) * ds * (pf[2] - pf[3]);
            bW += -((T)1.0f / den) * ds * (pf[4] - pf[5]);
            // diffusion (:1593-1836)
            T vis[6];
            if (useLES) FaceVis(x, y, z, vis);
            else for (int f = 0; f < 6; ++f) vis[f] = nu;
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) {
                        T t = vis[f] * ds / dx;
                        Dx += t; Dy += t; Dz += t;
                    } else {
                        T t = (T)(MUT(2) ? 1 : 2) * vis[f] * ds / dx;
                        Dx += t; Dy += t; Dz += t;
                    }
                } else {
                    int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        if (!MUT(12)) bU += (T)2 * vis[f] * ds / dx * ub.x;
                        bV += (T)2 * vis[f] * ds / dx * ub.y;
                        bW += (T)2 * vis[f] * ds / dx * ub.z;
                        T t = (T)2 * vis[f] * ds / dx;
                        Dx += t; Dy += t; Dz += t;
                    } else if (velBnd[bf] == UWS_VEL_SYMMETRY) {
                        T t = (T)2 * vis[f] * ds / dx;
                        int c = MUT(13) ? (f / 2 + 1) % 3 : f / 2;
                        if (c == 0) Dx += t; else if (c == 1) Dy += t; else Dz += t;
                    }
                }
            }
            // time derivative, force, LES (:1838-1854)
            bU += dv / dt * uP.x; bV += dv / dt * uP.y; bW += dv / dt * uP.z;
            bU += dv * fext[0]; bV += dv * fext[1]; bW += dv * fext[2];
            if (useLES) { bU += bfield[P].x; bV += bfield[P].y; bW += bfield[P].z; }
            Dfield[P] = V3<T>{Dx, Dy, Dz};
            bfield[P] = V3<T>{bU, bV, bW};
        }
    }
    // CSVelPredict (FvmColl.compute:1862-2069): reads velFieldLastIter, writes velField
    void VelPredict() {
        if (nonuniform) { NuVelPredict(); return; }
        const V3<T>* un = velFieldLastTime;
        const V3<T>* uk = velFieldLastIter;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const V3<T>& uP = un[P];
            T UA = 0, VA = 0, WA = 0, UD = 0, VD = 0, WD = 0;
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T uPn = c == 0 ? uP.x : (c == 1 ? uP.y : uP.z);
                T uNn = c == 0 ? un[Q].x : (c == 1 ? un[Q].y : un[Q].z);
                T sum = uPn + uNn;
                T tU = ds * (T)0.25f * uk[Q].x * sum;
                T tV = ds * (T)0.25f * uk[Q].y * sum;
                T tW = ds * (T)0.25f * uk[Q].z * sum;
                if (MUT(1)) plus = !plus;
                if (plus) { UA -= tU; VA -= tV; WA -= tW; } else { UA += tU; VA += tV; WA += tW; }
            }
            T vis[6];
            if (useLES) FaceVis(x, y, z, vis);
            else for (int f = 0; f < 6; ++f) vis[f] = nu;
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                UD += vis[f] * ds / dx * uk[Q].x;
                VD += vis[f] * ds / dx * uk[Q].y;
                WD += vis[f] * ds / dx * uk[Q].z;
            }
            velField[P].x = (T)1.0f / Dfield[P].x * (UA + UD + bfield[P].x);
            velField[P].y = (T)1.0f / Dfield[P].y * (VA + VD + bfield[P].y);
            velField[P].z = (T)1.0f / Dfield[P].z * (WA + WD + bfield[P].z);
        }
    }
    // CSComputeVelPredictResidual (FvmColl.compute:3123-3162) + host sum (FvmSolverGpu.cs:1216-1236)
    T VelPredictResidual() const {
        double s = 0;
        for (int64_t i = 0; i < N; ++i) {
            T rx = Dfield[i].x * (velField[i].x - velFieldLastIter[i].x);
            T ry = Dfield[i].y * (velField[i].y - velFieldLastIter[i].y);
            T rz = Dfield[i].z * (velField[i].z - velFieldLastIter[i].z);
            s += (double)(rx * rx + ry * ry + rz * rz);
        }
        return (T)std::sqrt(s / (3.0 * (double)N));
    }
    // VelPredict() (FvmSolverGpu.cs:1238-1282) without time slicing
    void VelPredictStage() {
        last_vel_sweeps = cf.vel_max_num_iter;
        for (int k = 0; k < cf.vel_max_num_iter; ++k) {
            VelPredict();
            if (cf.cal_residual && cf.vel_residual_check_interval > 0 && k % cf.vel_residual_check_interval == 0) {
                last_vel_residual = VelPredictResidual();
                if (last_vel_residual < (T)cf.vel_tolerance) {
                    last_vel_sweeps = k + 1;
                    break;
                }
            }
            std::swap(velFieldLastIter, velField);
        }
        std::swap(velFieldLastIter, velField);
    }

    // ------------------------------------------------------------------
    // Pressure correction (FvmColl.compute:2480-2880)
    // ------------------------------------------------------------------
    T Dc(int64_t i, int c) const { return c == 0 ? Dfield[i].x : (c == 1 ? Dfield[i].y : Dfield[i].z); }
    // CSPresCorrectRhieChowPreCompute (FvmColl.compute:2480-2787)
    void PresCorrectPreCompute() {
        if (nonuniform) { NuPresCorrectPreCompute(); return; }
        const V3<T>* u = velField;
        const T* p = presLastTime.data();
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        const T two_dx = (T)2 * dx;
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T Dp = 0, bp = 0;
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                bool plus = (f % 2) == 0;
                int ox = off[f][0], oy = off[f][1], oz = off[f][2];
                int nx = x + ox, ny = y + oy, nz = z + oz;
                T uPc = c == 0 ? u[P].x : (c == 1 ? u[P].y : u[P].z);
                if (InRegion(nx, ny, nz)) {
                    if (!IsFluid(nx, ny, nz)) continue;
                    const int64_t Q = idx(nx, ny, nz);
                    T DInv = (T)0.5f * ((T)1 / Dc(P, c) + (T)1 / Dc(Q, c));
                    Dp -= (dv / den) * DInv / dx;
                    T uQc = c == 0 ? u[Q].x : (c == 1 ? u[Q].y : u[Q].z);
                    T half = (T)0.5f * (uPc + uQc);
                    if (plus) bp += half; else bp -= half;
                    // Rhie-Chow: far neighbour (EE) and opposite neighbour (W)
                    int fx = x + 2 * ox, fy = y + 2 * oy, fz = z + 2 * oz;
                    int bx = x - ox, by = y - oy, bz = z - oz;
                    bool farF = InRegionAndIsFluid(fx, fy, fz), backF = InRegionAndIsFluid(bx, by, bz);
                    T pP = p[P], pQ = p[Q];
                    T presGrad, presGradAve;
                    if (plus) {
                        presGrad = (pQ - pP) / dx;
                        if (farF && backF) {
                            T pF = p[idx(fx, fy, fz)], pB = p[idx(bx, by, bz)];
                            presGradAve = (T)0.5f * ((pF - pP) / two_dx + (pQ - pB) / two_dx);
                        } else if (farF) {
                            T pF = p[idx(fx, fy, fz)];
                            presGradAve = (T)0.5f * ((pF - pP) / two_dx + (pQ - pP) / two_dx);
                        } else if (backF) {
                            T pB = p[idx(bx, by, bz)];
                            presGradAve = (T)0.5f * ((pQ - pP) / two_dx + (pQ - pB) / two_dx);
                        } else {
                            presGradAve = presGrad;
                        }
                    } else {
                        // W face: P-W gradient; far = WW, back = E (FvmColl.compute:2569-2579)
                        presGrad = (pP - pQ) / dx;
                        if (farF && backF) {
                            T pF = p[idx(fx, fy, fz)], pB = p[idx(bx, by, bz)];
                            presGradAve = (T)0.5f * ((pB - pQ) / two_dx + (pP - pF) / two_dx);
                        } else if (farF) {
                            T pF = p[idx(fx, fy, fz)];
                            presGradAve = (T)0.5f * ((pP - pQ) / two_dx + (pP - pF) / two_dx);
                        } else if (backF) {
                            T pB = p[idx(bx, by, bz)];
                            presGradAve = (T)0.5f * ((pB - pQ) / two_dx + (pP - pQ) / two_dx);
                        } else {
                            presGradAve = presGrad;
                        }
                    }
                    T velCorrectTerm = -(dv * DInv) / den * (presGrad - presGradAve);
                    if (plus) bp += velCorrectTerm; else bp -= velCorrectTerm;
                } else {
                    int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        T ubc = c == 0 ? ub.x : (c == 1 ? ub.y : ub.z);
                        if (plus) bp += ubc; else bp -= ubc;
                    } else if (velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                        if (plus) bp += uPc; else bp -= uPc;
                    }
                    if (presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                        T DInv = (T)1 / Dc(P, c);
                        Dp -= (T)2 * (dv / den) * DInv / dx;
                    }
                }
            }
            DpField[P] = Dp;
            bpField[P] = bp;
        }
    }
    // CSPresCorrectRhieChow (FvmColl.compute:2789-2848): reads LastIter, writes presCorrectField
    void PresCorrectSweep() {
        if (nonuniform) { NuPresCorrectSweep(); return; }
        const T* pk = presCorrectFieldLastIter;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T term = 0;
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T DInv = (T)0.5f * ((T)1 / Dc(P, c) + (T)1 / Dc(Q, c));
                term -= (dv / den) * DInv / dx * pk[Q];
            }
            presCorrectField[P] = ((T)1 / DpField[P]) * (term + bpField[P]);
        }
    }
    // CSComputePresCorrectResidual (FvmColl.compute:3164-3199) + host sum (FvmSolverGpu.cs:1345-1365)
    T PresCorrectResidual() const {
        double s = 0;
        for (int64_t i = 0; i < N; ++i) {
            T r = DpField[i] * (presCorrectField[i] - presCorrectFieldLastIter[i]);
            s += (double)(r * r);
        }
        return (T)std::sqrt(s / (double)N);
    }
    // PresCorrect() sweeps (FvmSolverGpu.cs:1367-1415)
    void PresSweepsStage() {
        last_pres_sweeps = cf.pres_max_num_iter;
        for (int k = 0; k < cf.pres_max_num_iter; ++k) {
            PresCorrectSweep();
            if (cf.cal_residual && cf.pres_residual_check_interval > 0 && k % cf.pres_residual_check_interval == 0) {
                last_pres_residual = PresCorrectResidual();
                if (last_pres_residual < (T)cf.pres_tolerance) {
                    last_pres_sweeps = k + 1;
                    break;
                }
            }
            std::swap(presCorrectFieldLastIter, presCorrectField);
        }
        std::swap(presCorrectFieldLastIter, presCorrectField);
    }
    // CSApplyPresCorrection (FvmColl.compute:2850-2862) + CSPresNormalization (2864-2880)
    void ApplyPresStage() {
        for (int64_t i = 0; i < N; ++i)
            if (flag[i] != UWS_CELL_SOLID) presLastTime[i] += presCorrectField[i];
        bool anyFixed = false;
        for (int f = 0; f < 6; ++f) anyFixed |= presBnd[f] == UWS_PRES_FIXED_VALUE;
        if (!anyFixed) {
            // reference cell (2, presRes.y / 2, 2), read once before the subtraction;
            // a texture read outside the grid returns 0
            T ref = InRegion(2, Y / 2, 2) ? presLastTime[idx(2, Y / 2, 2)] : (T)0;
            for (int64_t i = 0; i < N; ++i)
                if (flag[i] != UWS_CELL_SOLID) presLastTime[i] -= ref;
        }
    }
    // CSApplyVelCorrectionDivergenceTheorem (FvmColl.compute:2950-3121) + SetPresCorrectFieldtoZero
    void ApplyVelCorrectionStage() {
        if (nonuniform) {
            NuApplyVelCorrection();
            SetPresCorrectFieldToZero();
            return;
        }
        const T* pc = presCorrectField;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T pf[6];
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) pf[f] = (T)0.5f * (pc[P] + pc[idx(nx, ny, nz)]);
                    else pf[f] = pc[P];
                } else {
                    pf[f] = presBnd[bface[f]] == UWS_PRES_FIXED_VALUE ? (T)0 : pc[P];
                }
            }
            T cx = -(dv / den) * ((T)1 / Dfield[P].x) * (pf[0] - pf[1]) / dx;
            T cy = -(dv / den) * ((T)1 / Dfield[P].y) * (pf[2] - pf[3]) / dx;
            T cz = -(dv / den) * ((T)1 / Dfield[P].z) * (pf[4] - pf[5]) / dx;
            velCorrect[P] = V3<T>{cx, cy, cz};
            velField[P].x += cx; velField[P].y += cy; velField[P].z += cz;
        }
        SetPresCorrectFieldToZero();
    }
    // SetPresCorrectFieldtoZero (FvmSolverGpu.cs:1545-1554)
    void SetPresCorrectFieldToZero() {
        std::fill(tex_pc[0].begin(), tex_pc[0].end(), (T)0);
        std::fill(tex_pc[1].begin(), tex_pc[1].end(), (T)0);
    }

    // ------------------------------------------------------------------
    // PISO corrector (PisoColl.compute:70-337, FvmSolverGpu.cs:1482-1533)
    // ------------------------------------------------------------------
    void PisoCalAodUCorrect() {
        const V3<T>* un = velFieldLastTime;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!InRegionAndIsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T UA = 0, VA = 0, WA = 0, UD = 0, VD = 0, WD = 0;
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T uPn = c == 0 ? un[P].x : (c == 1 ? un[P].y : un[P].z);
                T uNn = c == 0 ? un[Q].x : (c == 1 ? un[Q].y : un[Q].z);
                T sum = uPn + uNn;
                T tU = ds * (T)0.25f * velCorrect[Q].x * sum;
                T tV = ds * (T)0.25f * velCorrect[Q].y * sum;
                T tW = ds * (T)0.25f * velCorrect[Q].z * sum;
                if (plus) { UA += tU; VA += tV; WA += tW; } else { UA -= tU; VA -= tV; WA -= tW; }
            }
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                if (MUT(6)) UD += nu * ds / dx * velCorrect[Q].x; else
                UD -= nu * ds / dx * velCorrect[Q].x;
                VD -= nu * ds / dx * velCorrect[Q].y;
                WD -= nu * ds / dx * velCorrect[Q].z;
            }
            AodU[P] = V3<T>{UA + UD, VA + VD, WA + WD};
        }
    }
    T Ac(int64_t i, int c) const { return c == 0 ? AodU[i].x : (c == 1 ? AodU[i].y : AodU[i].z); }
    // CSPisoPresCorrection (PisoColl.compute:180-256)
    void PisoPresCorrection() {
        const T* pk = presCorrectFieldLastIter;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!InRegionAndIsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T presCoeff = 0, presTerm = 0, velTerm = 0;
            for (int f = 0; f < 6; ++f) {
                int c = f / 2;
                bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T AdInv = (T)0.5f * ((T)1 / Dc(P, c) + (T)1 / Dc(Q, c));
                T AodUCorrect = (T)0.5f * (Ac(P, c) + Ac(Q, c));
                if (c == 0) {
                    // E/W faces group the division first (PisoColl.compute:205-206, 214-215)
                    presCoeff -= (dv / den) * (AdInv / dx);
                    presTerm -= (dv / den) * (AdInv / dx) * pk[Q];
                } else {
                    presCoeff -= (dv / den) * AdInv / dx;
                    presTerm -= (dv / den) * AdInv / dx * pk[Q];
                }
                if (plus != MUT(7)) velTerm -= AdInv * AodUCorrect; else velTerm += AdInv * AodUCorrect;
            }
            presCorrectField[P] = ((T)1 / presCoeff) * (presTerm + velTerm);
        }
    }
    // CSPisoApplyPresCorrection (PisoColl.compute:258-267)
    void PisoApplyPresCorrection() {
        for (int64_t i = 0; i < N; ++i)
            if (flag[i] == UWS_CELL_FLUID) presLastTime[i] += presCorrectField[i];
    }
    // CSPisoVelCorrection (PisoColl.compute:269-337)
    void PisoVelCorrection() {
        const T* pc = presCorrectField;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            T DxInv = (T)1 / Dfield[P].x, DyInv = (T)1 / Dfield[P].y, DzInv = (T)1 / Dfield[P].z;
            T pf[6];
            for (int f = 0; f < 6; ++f) {
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegionAndIsFluid(nx, ny, nz)) pf[f] = (T)0.5f * (pc[P] + pc[idx(nx, ny, nz)]);
                else pf[f] = (MUT(8) && InRegion(nx, ny, nz)) ? (T)0 : pc[P];
            }
            T gx = (pf[0] - pf[1]) / dx, gy = (pf[2] - pf[3]) / dx, gz = (pf[4] - pf[5]) / dx;
            T cx = -DxInv * (AodU[P].x + dv / den * gx);
            T cy = -DyInv * (AodU[P].y + dv / den * gy);
            T cz = -DzInv * (AodU[P].z + dv / den * gz);
            velCorrect[P] = V3<T>{cx, cy, cz};
            velField[P].x += cx; velField[P].y += cy; velField[P].z += cz;
        }
    }
    // PisoCorrection() (FvmSolverGpu.cs:1482-1533)
    void PisoStage() {
        PisoCalAodUCorrect();
        for (int k = 0; k < cf.pres_max_num_iter; ++k) {
            PisoPresCorrection();
            std::swap(presCorrectFieldLastIter, presCorrectField);
        }
        std::swap(presCorrectFieldLastIter, presCorrectField);
        PisoApplyPresCorrection();
        PisoVelCorrection();
        SetPresCorrectFieldToZero();
    }
    // ------------------------------------------------------------------
    // Stretched collocated grid (FvmCollNonUniform.compute).  Same step as
    // the uniform grid (SIMPLE only); spacings come from the face
    // coordinates, face values are linear interpolations with the weights
    // of the two cell sizes, and the momentum convection uses the Rhie-Chow
    // face velocities that the previous pressure precompute left in
    // velFaceFluxField.
    // ------------------------------------------------------------------
    int64_t fidx(int x, int y, int z) const { return ((int64_t)z * (Y + 1) + y) * (X + 1) + x; }
    T Sp(int a, int i) const { return (T)facePos[a][i + 1] - (T)facePos[a][i]; }
    // cell and neighbour spacings; a neighbour outside the grid takes the
    // cell's own spacing (FvmCollNonUniform.compute:130-157, 943-983)
    struct Geo {
        T d[3];    // dx dy dz
        T ds[3];   // dsx = dy*dz, dsy = dx*dz, dsz = dx*dy
        T dv;      // dx*dy*dz
        T dn[6];   // dxE dxW dyN dyS dzU dzD
        T dnn[6];  // dxEE dxWW dyNN dySS dzUU dzDD
    };
    Geo Geometry(int x, int y, int z) const {
        Geo g;
        const int c[3] = {x, y, z}, n[3] = {X, Y, Z};
        for (int a = 0; a < 3; ++a) g.d[a] = Sp(a, c[a]);
        g.ds[0] = g.d[1] * g.d[2];
        g.ds[1] = g.d[0] * g.d[2];
        g.ds[2] = g.d[0] * g.d[1];
        g.dv = g.d[0] * g.d[1] * g.d[2];
        for (int a = 0; a < 3; ++a) {
            g.dn[2 * a] = c[a] + 1 < n[a] ? Sp(a, c[a] + 1) : g.d[a];
            g.dn[2 * a + 1] = c[a] - 1 >= 0 ? Sp(a, c[a] - 1) : g.d[a];
            g.dnn[2 * a] = c[a] + 2 < n[a] ? Sp(a, c[a] + 2) : g.d[a];
            g.dnn[2 * a + 1] = c[a] - 2 >= 0 ? Sp(a, c[a] - 2) : g.d[a];
        }
        return g;
    }
    // face velocity of face f (E W N S U D) of cell (x, y, z): the + faces are
    // stored at the neighbour's texel, the - faces at the cell's own
    T FaceFlux(int f, int x, int y, int z) const {
        switch (f) {
        case 0: return faceFlux[fidx(x + 1, y, z)].x;
        case 1: return faceFlux[fidx(x, y, z)].x;
        case 2: return faceFlux[fidx(x, y + 1, z)].y;
        case 3: return faceFlux[fidx(x, y, z)].y;
        case 4: return faceFlux[fidx(x, y, z + 1)].z;
        default: return faceFlux[fidx(x, y, z)].z;
        }
    }
    // face viscosities: nu, plus the interpolated eddy viscosity with LES
    // (FvmCollNonUniform.compute:482-503)
    void NuFaceVis(int x, int y, int z, const Geo& g, T vis[6]) const {
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        const int64_t P = idx(x, y, z);
        for (int f = 0; f < 6; ++f) vis[f] = nu;
        if (!useLES) return;
        for (int f = 0; f < 6; ++f) vis[f] = nu + eddyVis[P];
        for (int f = 0; f < 6; ++f) {
            int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
            if (!InRegionAndIsFluid(nx, ny, nz)) continue;
            const int a = f / 2;
            T wP = g.dn[f] / (g.d[a] + g.dn[f]);
            T wN = (T)1.0f - wP;
            vis[f] = nu + wP * eddyVis[P] + wN * eddyVis[idx(nx, ny, nz)];
        }
    }
    // CSVelPredictPreCompute (FvmCollNonUniform.compute:111-721)
    void NuVelPredictPreCompute() {
        const V3<T>* un = velFieldLastTime;
        const T* p = presLastTime.data();
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const Geo g = Geometry(x, y, z);
            T Dx = g.dv / dt, Dy = g.dv / dt, Dz = g.dv / dt;
            T bU = 0, bV = 0, bW = 0;
            const V3<T>& uP = un[P];
            // advection with the stored face velocities (:185-325)
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                const bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                const T F = FaceFlux(f, x, y, z);
                const T dsa = g.ds[a];
                if (InRegionAndIsFluid(nx, ny, nz)) {
                    T wP = g.dn[f] / (g.d[a] + g.dn[f]);
                    T t = dsa * wP * F;
                    if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
                } else if (!InRegion(nx, ny, nz)) {
                    int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        if (plus) { bU -= dsa * ub.x * F; bV -= dsa * ub.y * F; bW -= dsa * ub.z * F; }
                        else { bU += dsa * ub.x * F; bV += dsa * ub.y * F; bW += dsa * ub.z * F; }
                    } else if (velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                        T t = dsa * F;
                        if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
                    }
                }
            }
            // pressure face values (:327-462)
            T pf[6];
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) {
                        T wP = g.dn[f] / (g.d[a] + g.dn[f]);
                        T wN = (T)1.0f - wP;
                        pf[f] = wP * p[P] + wN * p[idx(nx, ny, nz)];
                    } else {
                        pf[f] = p[P];
                    }
                } else {
                    pf[f] = presBnd[bface[f]] == UWS_PRES_FIXED_VALUE ? (T)0 : p[P];
                }
            }
            bU += -((T)1.0f / den) * g.ds[0] * (pf[0] - pf[1]);
            bV += -((T)1.0f / den) * g.ds[1] * (pf[2] - pf[3]);
            bW += -((T)1.0f / den) * g.ds[2] * (pf[4] - pf[5]);
            // diffusion (:464-697)
            T vis[6];
            NuFaceVis(x, y, z, g, vis);
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                const T dsa = g.ds[a];
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    T t = IsFluid(nx, ny, nz) ? vis[f] * dsa / ((T)0.5f * g.d[a] + (T)0.5f * g.dn[f])
                                              : vis[f] * dsa / ((T)0.5f * g.d[a]);
                    Dx += t; Dy += t; Dz += t;
                } else {
                    int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        T k = vis[f] * dsa / ((T)0.5f * g.d[a]);
                        bU += k * ub.x; bV += k * ub.y; bW += k * ub.z;
                        Dx += k; Dy += k; Dz += k;
                    } else if (velBnd[bf] == UWS_VEL_SYMMETRY) {
                        T k = vis[f] * dsa / ((T)0.5f * g.d[a]);
                        if (a == 0) Dx += k; else if (a == 1) Dy += k; else Dz += k;
                    }
                }
            }
            // time derivative, force, LES (:699-715)
            bU += g.dv / dt * uP.x; bV += g.dv / dt * uP.y; bW += g.dv / dt * uP.z;
            bU += g.dv * fext[0]; bV += g.dv * fext[1]; bW += g.dv * fext[2];
            if (useLES) { bU += bfield[P].x; bV += bfield[P].y; bW += bfield[P].z; }
            Dfield[P] = V3<T>{Dx, Dy, Dz};
            bfield[P] = V3<T>{bU, bV, bW};
        }
    }
    // CSVelPredict (FvmCollNonUniform.compute:725-918)
    void NuVelPredict() {
        const V3<T>* uk = velFieldLastIter;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const Geo g = Geometry(x, y, z);
            T UA = 0, VA = 0, WA = 0, UD = 0, VD = 0, WD = 0;
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                const bool plus = (f % 2) == 0;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T wP = g.dn[f] / (g.d[a] + g.dn[f]);
                T wN = (T)1.0f - wP;
                const T F = FaceFlux(f, x, y, z);
                T tU = g.ds[a] * wN * uk[Q].x * F;
                T tV = g.ds[a] * wN * uk[Q].y * F;
                T tW = g.ds[a] * wN * uk[Q].z * F;
                if (plus) { UA -= tU; VA -= tV; WA -= tW; } else { UA += tU; VA += tV; WA += tW; }
            }
            T vis[6];
            NuFaceVis(x, y, z, g, vis);
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T k = vis[f] * g.ds[a] / ((T)0.5f * g.d[a] + (T)0.5f * g.dn[f]);
                UD += k * uk[Q].x;
                VD += k * uk[Q].y;
                WD += k * uk[Q].z;
            }
            velField[P].x = (T)1.0f / Dfield[P].x * (UA + UD + bfield[P].x);
            velField[P].y = (T)1.0f / Dfield[P].y * (VA + VD + bfield[P].y);
            velField[P].z = (T)1.0f / Dfield[P].z * (WA + WD + bfield[P].z);
        }
    }
    // CSPresCorrectRhieChowPreCompute (FvmCollNonUniform.compute:920-1344):
    // Dp, bp and the face velocities left in velFaceFluxField
    void NuPresCorrectPreCompute() {
        const V3<T>* u = velField;
        const T* p = presLastTime.data();
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const Geo g = Geometry(x, y, z);
            T Dp = 0, bp = 0;
            T uf[6] = {0, 0, 0, 0, 0, 0};
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                const bool plus = (f % 2) == 0;
                const int ox = off[f][0], oy = off[f][1], oz = off[f][2];
                const int nx = x + ox, ny = y + oy, nz = z + oz;
                const T uPc = a == 0 ? u[P].x : (a == 1 ? u[P].y : u[P].z);
                if (InRegion(nx, ny, nz)) {
                    if (!IsFluid(nx, ny, nz)) continue;  // a solid neighbour leaves the face at 0
                    const int64_t Q = idx(nx, ny, nz);
                    const T d = g.d[a], dQ = g.dn[f], dF = g.dnn[f], dB = g.dn[f ^ 1];
                    T wP = dQ / (d + dQ);
                    T wN = (T)1.0f - wP;
                    T DInv = wP * ((T)1 / Dc(P, a)) + wN * ((T)1 / Dc(Q, a));
                    Dp -= (g.dv / den) * DInv / ((T)0.5f * d + (T)0.5f * dQ);
                    const T uQc = a == 0 ? u[Q].x : (a == 1 ? u[Q].y : u[Q].z);
                    T face = wP * uPc + wN * uQc;
                    // Rhie-Chow correction with one-sided averages next to solids and walls
                    const int fx = x + 2 * ox, fy = y + 2 * oy, fz = z + 2 * oz;
                    const int bx = x - ox, by = y - oy, bz = z - oz;
                    const bool farF = InRegionAndIsFluid(fx, fy, fz), backF = InRegionAndIsFluid(bx, by, bz);
                    const T pP = p[P], pQ = p[Q];
                    const T h = (T)0.5f; const T h3 = (T)1.5f;
                    T gr, ga;
                    if (plus) {
                        gr = (pQ - pP) / (h * d + h * dQ);
                        if (farF && backF) {
                            const T pF = p[idx(fx, fy, fz)], pB = p[idx(bx, by, bz)];
                            ga = h * ((pF - pP) / (h * dF + dQ + h * d) + (pQ - pB) / (h * dQ + d + h * dB));
                        } else if (farF) {
                            const T pF = p[idx(fx, fy, fz)];
                            ga = h * ((pF - pP) / (h * dF + dQ + h * d) + (pQ - pP) / (h * dQ + h3 * d));
                        } else if (backF) {
                            const T pB = p[idx(bx, by, bz)];
                            ga = h * ((pQ - pP) / (h3 * dQ + h * d) + (pQ - pB) / (h * dQ + d + h * dB));
                        } else {
                            ga = gr;
                        }
                    } else {
                        gr = (pP - pQ) / (h * d + h * dQ);
                        if (farF && backF) {
                            const T pF = p[idx(fx, fy, fz)], pB = p[idx(bx, by, bz)];
                            ga = h * ((pB - pQ) / (h * dB + d + h * dQ) + (pP - pF) / (h * d + dQ + h * dF));
                        } else if (farF) {
                            const T pF = p[idx(fx, fy, fz)];
                            ga = h * ((pP - pQ) / (h3 * d + h * dQ) + (pP - pF) / (h * d + dQ + h * dF));
                        } else if (backF) {
                            const T pB = p[idx(bx, by, bz)];
                            ga = h * ((pB - pQ) / (h * dB + d + h * dQ) + (pP - pQ) / (h * d + h3 * dQ));
                        } else {
                            ga = gr;
                        }
                    }
                    face -= (g.dv * DInv) / den * (gr - ga);
                    if (plus) bp += face; else bp -= face;
                    uf[f] = face;
                } else {
                    const int bf = bface[f];
                    if (velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                        int i, j;
                        if (bf < 2) { i = y; j = z; } else if (bf < 4) { i = x; j = z; } else { i = x; j = y; }
                        const V3<T>& ub = B(bf, i, j);
                        const T ubc = a == 0 ? ub.x : (a == 1 ? ub.y : ub.z);
                        uf[f] = ubc;
                        if (plus) bp += ubc; else bp -= ubc;
                    } else if (velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                        uf[f] = uPc;
                        if (plus) bp += uPc; else bp -= uPc;
                    }
                    if (presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                        T DInv = (T)1 / Dc(P, a);
                        Dp -= (g.dv / den) * DInv / ((T)0.5f * g.d[a]);
                    }
                }
            }
            // velFaceFluxField: the cell's W/S/D faces, and the + faces on the domain boundary (:1335-1342)
            faceFlux[fidx(x, y, z)] = V3<T>{uf[1], uf[3], uf[5]};
            if (x + 1 >= X) faceFlux[fidx(x + 1, y, z)] = V3<T>{uf[0], 0, 0};
            if (y + 1 >= Y) faceFlux[fidx(x, y + 1, z)] = V3<T>{0, uf[2], 0};
            if (z + 1 >= Z) faceFlux[fidx(x, y, z + 1)] = V3<T>{0, 0, uf[4]};
            DpField[P] = Dp;
            bpField[P] = bp;
        }
    }
    // CSPresCorrectRhieChow (FvmCollNonUniform.compute:1346-1453)
    void NuPresCorrectSweep() {
        const T* pk = presCorrectFieldLastIter;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const Geo g = Geometry(x, y, z);
            T term = 0;
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (!InRegionAndIsFluid(nx, ny, nz)) continue;
                const int64_t Q = idx(nx, ny, nz);
                T wP = g.dn[f] / (g.d[a] + g.dn[f]);
                T wN = (T)1.0f - wP;
                T DInv = wP * ((T)1 / Dc(P, a)) + wN * ((T)1 / Dc(Q, a));
                term -= (g.dv / den) * DInv / ((T)0.5f * g.d[a] + (T)0.5f * g.dn[f]) * pk[Q];
            }
            presCorrectField[P] = ((T)1 / DpField[P]) * (term + bpField[P]);
        }
    }
    // CSApplyVelCorrectionDivergenceTheorem (FvmCollNonUniform.compute:1487-1732)
    void NuApplyVelCorrection() {
        const T* pc = presCorrectField;
        static const int off[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int bface[6] = {UWS_XN, UWS_X0, UWS_YN, UWS_Y0, UWS_ZN, UWS_Z0};
        UWSO_PARALLEL_CELLS
        for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
            if (!IsFluid(x, y, z)) continue;
            const int64_t P = idx(x, y, z);
            const Geo g = Geometry(x, y, z);
            T pf[6];
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                int nx = x + off[f][0], ny = y + off[f][1], nz = z + off[f][2];
                if (InRegion(nx, ny, nz)) {
                    if (IsFluid(nx, ny, nz)) {
                        T wP = g.dn[f] / (g.d[a] + g.dn[f]);
                        T wN = (T)1.0f - wP;
                        pf[f] = wP * pc[P] + wN * pc[idx(nx, ny, nz)];
                    } else {
                        pf[f] = pc[P];
                    }
                } else {
                    pf[f] = presBnd[bface[f]] == UWS_PRES_FIXED_VALUE ? (T)0 : pc[P];
                }
            }
            T cx = -(g.dv / den) * ((T)1 / Dfield[P].x) * (pf[0] - pf[1]) / g.d[0];
            T cy = -(g.dv / den) * ((T)1 / Dfield[P].y) * (pf[2] - pf[3]) / g.d[1];
            T cz = -(g.dv / den) * ((T)1 / Dfield[P].z) * (pf[4] - pf[5]) / g.d[2];
            velCorrect[P] = V3<T>{cx, cy, cz};
            velField[P].x += cx; velField[P].y += cy; velField[P].z += cz;
        }
    }

    // end of Step() (FvmSolverGpu.cs:1103-1110)
    void FinishStage() {
        std::swap(velFieldLastTime, velField);
        std::copy(velFieldLastTime, velFieldLastTime + N, velFieldLastIter);
        ++step_count;
        phys_time += (double)cf.dt;
    }

    void RunStage(int s) {
        switch (s) {
        case UWS_STAGE_WALL_FUNC: if (useLES) CalLogLawWallFunc(); break;
        case UWS_STAGE_EDDY_VIS: if (useLES) CalEddyVis(); break;
        case UWS_STAGE_LES_TERM: if (useLES) CalLesDeferCorrectTerm(); break;
        case UWS_STAGE_VEL_PRECOMP: VelPredictPreCompute(); break;
        case UWS_STAGE_VEL_SWEEPS: VelPredictStage(); break;
        case UWS_STAGE_PRES_PRECOMP: PresCorrectPreCompute(); break;
        case UWS_STAGE_PRES_SWEEPS: PresSweepsStage(); break;
        case UWS_STAGE_APPLY_PRES: ApplyPresStage(); break;
        case UWS_STAGE_VEL_CORRECT: ApplyVelCorrectionStage(); break;
        case UWS_STAGE_PISO: if (piso) PisoStage(); break;
        case UWS_STAGE_FINISH: FinishStage(); break;
        default: break;
        }
    }
    // FvmSolverGpu.Step() (FvmSolverGpu.cs:1050-1121)
    void Step() {
        for (int s = UWS_STAGE_WALL_FUNC; s <= UWS_STAGE_VEL_CORRECT; ++s) RunStage(s);
        if (piso) for (int n = 0; n < cf.piso_num_correctors; ++n) PisoStage();
        FinishStage();
    }

    // field access for the C API; V3 fields as [N][3].  Two ids beyond the
    // library's, for the pinning tests' single kernels: UWSO_FIELD_VEL_ITER
    // (velFieldLastIter, a momentum sweep's input) and UWSO_FIELD_PC_ITER
    // (presCorrectFieldLastIter, a pressure sweep's input)
    V3<T>* vfield(int f) {
        switch (f) {
        case UWSO_FIELD_VEL_ITER: return velFieldLastIter;
        case UWS_FIELD_VEL: return velFieldLastTime;
        case UWS_FIELD_VEL_WORK: return velField;
        case UWS_FIELD_VEL_CORR: return velCorrect.data();
        case UWS_FIELD_WALL_VIS: return faceEddyVis.data();
        case UWS_FIELD_B: return bfield.data();
        case UWS_FIELD_D: return Dfield.data();
        case UWS_FIELD_AODU: return AodU.data();
        default: return nullptr;
        }
    }
    T* sfield(int f) {
        switch (f) {
        case UWSO_FIELD_PC_ITER: return presCorrectFieldLastIter;
        case UWS_FIELD_PRES: return presLastTime.data();
        case UWS_FIELD_PRES_CORR: return presCorrectField;
        case UWS_FIELD_EDDY_VIS: return eddyVis.data();
        case UWS_FIELD_DP: return DpField.data();
        case UWS_FIELD_BP: return bpField.data();
        default: return nullptr;
        }
    }
};

struct Handle {
    int precision;  // 0 float, 1 double
    Oracle<float>* f = nullptr;
    Oracle<double>* d = nullptr;
};

}  // namespace

template <class F> static void dispatch(Handle* h, F&& fn) {
    if (h->precision == 0) fn(*h->f); else fn(*h->d);
}

extern "C" {

int uwso_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}


/* precision: 0 = FP32 (the oracle proper), 1 = FP64 reference run */
void* uwso_create(const uws_config* cfg, int precision) {
    Handle* h = new Handle();
    h->precision = precision;
    if (precision == 0) h->f = new Oracle<float>(*cfg); else h->d = new Oracle<double>(*cfg);
    return h;
}
void uwso_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    delete h->f; delete h->d; delete h;
}
void uwso_init_flags(void* hv, const int32_t* flags) {
    dispatch((Handle*)hv, [&](auto& o) { o.InitModelFlag(flags); });
}
void uwso_init_box(void* hv) { dispatch((Handle*)hv, [&](auto& o) { o.InitBoxFlag(); }); }
void uwso_init_vel_pres_fields(void* hv, int inflow, float wind) {
    dispatch((Handle*)hv, [&](auto& o) { o.InitVelPresFields(inflow, wind); });
}
void uwso_set_fixed_value_vel_bnd_cond(void* hv, int inflow, float wind, const float* velZn, float refh, float alpha) {
    dispatch((Handle*)hv, [&](auto& o) { o.SetFixedValueVelBndCond(inflow, wind, velZn, refh, alpha); });
}
void uwso_step(void* hv, int n) {
    dispatch((Handle*)hv, [&](auto& o) { for (int i = 0; i < n; ++i) o.Step(); });
}
void uwso_run_stage(void* hv, int s) { dispatch((Handle*)hv, [&](auto& o) { o.RunStage(s); }); }

/* fields are exchanged as double so one API serves both precisions */
int uwso_get_field(void* hv, int f, double* dst) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        if (f == UWS_FIELD_FACE_VEL) {
            // velFaceFluxField at the cells: their W, S, D face velocities
            if (!o.nonuniform) { rc = -1; return; }
            for (int z = 0; z < o.Z; ++z) for (int y = 0; y < o.Y; ++y) for (int x = 0; x < o.X; ++x) {
                const int64_t i = o.idx(x, y, z);
                const auto& v = o.faceFlux[o.fidx(x, y, z)];
                dst[3*i] = v.x; dst[3*i+1] = v.y; dst[3*i+2] = v.z;
            }
        } else if (auto* v = o.vfield(f)) {
            for (int64_t i = 0; i < o.N; ++i) { dst[3*i] = v[i].x; dst[3*i+1] = v[i].y; dst[3*i+2] = v[i].z; }
        } else if (auto* s = o.sfield(f)) {
            for (int64_t i = 0; i < o.N; ++i) dst[i] = s[i];
        } else rc = -1;
    });
    return rc;
}
int uwso_set_field(void* hv, int f, const double* src) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        using TT = std::remove_reference_t<decltype(o.dx)>;
        if (f == UWS_FIELD_VEL) {
            // a velocity upload writes all three velocity buffers, as the init kernels do
            for (int64_t i = 0; i < o.N; ++i) {
                V3<TT> v{(TT)src[3*i], (TT)src[3*i+1], (TT)src[3*i+2]};
                o.velField[i] = o.velFieldLastTime[i] = o.velFieldLastIter[i] = v;
            }
        } else if (auto* v = o.vfield(f)) {
            for (int64_t i = 0; i < o.N; ++i) v[i] = V3<TT>{(TT)src[3*i], (TT)src[3*i+1], (TT)src[3*i+2]};
        } else if (auto* s = o.sfield(f)) {
            for (int64_t i = 0; i < o.N; ++i) s[i] = (TT)src[i];
            // p' is also the next sweep's input (presCorrectFieldLastIter)
            if (f == UWS_FIELD_PRES_CORR) std::copy(s, s + o.N, o.presCorrectFieldLastIter);
        } else rc = -1;
    });
    return rc;
}
/* stretched grid: face coordinates (X+1, Y+1, Z+1 floats) */
int uwso_set_face_positions(void* hv, const float* fx, const float* fy, const float* fz) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        if (!o.nonuniform) { rc = -1; return; }
        const float* src[3] = {fx, fy, fz};
        for (int a = 0; a < 3; ++a) std::copy(src[a], src[a] + o.facePos[a].size(), o.facePos[a].begin());
    });
    return rc;
}
/* the whole velFaceFluxField texture, [(z*(Y+1) + y)*(X+1) + x][3] */
int uwso_get_face_flux(void* hv, double* dst) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        if (!o.nonuniform) { rc = -1; return; }
        for (size_t i = 0; i < o.faceFlux.size(); ++i) {
            dst[3*i] = o.faceFlux[i].x; dst[3*i+1] = o.faceFlux[i].y; dst[3*i+2] = o.faceFlux[i].z;
        }
    });
    return rc;
}
int uwso_set_face_flux(void* hv, const double* src) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        using TT = std::remove_reference_t<decltype(o.dx)>;
        if (!o.nonuniform) { rc = -1; return; }
        for (size_t i = 0; i < o.faceFlux.size(); ++i)
            o.faceFlux[i] = V3<TT>{(TT)src[3*i], (TT)src[3*i+1], (TT)src[3*i+2]};
    });
    return rc;
}
void uwso_get_flags(void* hv, int32_t* dst) {
    dispatch((Handle*)hv, [&](auto& o) { std::copy(o.flag.begin(), o.flag.end(), dst); });
}
int uwso_set_bnd_field(void* hv, int face, const double* xyz) {
    if (face < 0 || face > 5) return -1;
    dispatch((Handle*)hv, [&](auto& o) {
        using TT = std::remove_reference_t<decltype(o.dx)>;
        for (size_t i = 0; i < o.bnd[face].size(); ++i)
            o.bnd[face][i] = V3<TT>{(TT)xyz[3*i], (TT)xyz[3*i+1], (TT)xyz[3*i+2]};
    });
    return 0;
}
int uwso_get_bnd_field(void* hv, int face, double* xyz) {
    if (face < 0 || face > 5) return -1;
    dispatch((Handle*)hv, [&](auto& o) {
        for (size_t i = 0; i < o.bnd[face].size(); ++i) {
            xyz[3*i] = o.bnd[face][i].x; xyz[3*i+1] = o.bnd[face][i].y; xyz[3*i+2] = o.bnd[face][i].z;
        }
    });
    return 0;
}
/* single reference kernels, for the oracle's own pinning tests */
double uwso_log_law_wall_func(void* hv, double uP) {
    double r = 0;
    dispatch((Handle*)hv, [&](auto& o) { using TT = std::remove_reference_t<decltype(o.dx)>; r = o.LogLawWallFunc((TT)uP); });
    return r;
}
/* single kernels, for the oracle's own pinning tests: 0 one momentum Jacobi
 * sweep (velFieldLastIter -> velField), 1 PISO A_od u' (velCorrect -> AodU),
 * 2 one PISO pressure sweep (presCorrectFieldLastIter -> presCorrectField),
 * 3 PISO velocity correction (presCorrectField, AodU -> velCorrect, velField) */
int uwso_run_kernel(void* hv, int k) {
    int rc = 0;
    dispatch((Handle*)hv, [&](auto& o) {
        switch (k) {
        case 0: o.VelPredict(); break;
        case 1: o.PisoCalAodUCorrect(); break;
        case 2: o.PisoPresCorrection(); break;
        case 3: o.PisoVelCorrection(); break;
        default: rc = -1;
        }
    });
    return rc;
}
void uwso_residuals(void* hv, double* out2) {
    dispatch((Handle*)hv, [&](auto& o) { out2[0] = o.last_vel_residual; out2[1] = o.last_pres_residual; });
}
void uwso_sweep_counts(void* hv, int* out2) {
    dispatch((Handle*)hv, [&](auto& o) { out2[0] = o.last_vel_sweeps; out2[1] = o.last_pres_sweeps; });
}
void uwso_time(void* hv, int64_t* step, double* t) {
    dispatch((Handle*)hv, [&](auto& o) { *step = o.step_count; *t = o.phys_time; });
}

}  // extern "C"
