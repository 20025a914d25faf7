// gfx950 kernels of the collocated finite-volume solver step.
//
// Each kernel restates one reference compute kernel (cited per kernel) with
// the reference's float operation order, so that with -ffp-contract=off the
// results match the CPU restatement bit for bit where the same arithmetic
// is used.  Layout: structure-of-arrays float planes, flat cell index
// z*X*Y + y*X + x (FvmColl.compute:122-126), one byte of flag per cell.
//
// Launch shape: 64 x 4 threads per workgroup (one wave per x-row of 64 cells,
// four rows), a 1-D grid whose workgroup ids are remapped so that each of the
// 8 XCDs receives a contiguous range of (x, y, z) tiles and the stencil's
// neighbour rows and planes stay in that XCD's L2.
#include "uws_internal.h"

namespace uws {

namespace {

constexpr int TX = 64, TY = 4;

enum { E_ = 0, W_ = 1, N_ = 2, S_ = 3, U_ = 4, D_ = 5 };
// neighbour offsets and boundary face of each direction E W N S U D
__constant__ int c_off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

__device__ __forceinline__ int bface_of(int dir) {
    // E->Xn, W->X0, N->Yn, S->Y0, U->Zn, D->Z0
    return dir == 0 ? UWS_XN : dir == 1 ? UWS_X0 : dir == 2 ? UWS_YN : dir == 3 ? UWS_Y0 : dir == 4 ? UWS_ZN : UWS_Z0;
}

// Workgroup id -> logical tile id such that XCD j (which receives ids
// j, j+8, j+16, ...) owns a contiguous run of tiles.
__device__ __forceinline__ int xcd_tile(int i, int B) {
    const int q = B >> 3, r = B & 7;
    const int xcd = i & 7, k = i >> 3;
    return xcd * q + min(xcd, r) + k;
}

struct Cell {
    int x, y, z, P;
    bool ok;
};

__device__ __forceinline__ Cell cell_of(const Params& p) {
    const int gx = (p.X + TX - 1) / TX, gy = (p.Y + TY - 1) / TY;
    const int B = gx * gy * p.Z;
    const int t = xcd_tile((int)blockIdx.x, B);
    const int bx = t % gx, r = t / gx, by = r % gy, bz = r / gy;
    Cell c;
    c.x = bx * TX + (int)threadIdx.x;
    c.y = by * TY + (int)threadIdx.y;
    c.z = bz;
    c.ok = c.x < p.X && c.y < p.Y;
    c.P = (c.z * p.Y + c.y) * p.X + c.x;
    return c;
}

inline int grid_blocks(const Params& p) { return ((p.X + TX - 1) / TX) * ((p.Y + TY - 1) / TY) * p.Z; }

__device__ __forceinline__ bool inreg(const Params& p, int x, int y, int z) {
    return (unsigned)x < (unsigned)p.X && (unsigned)y < (unsigned)p.Y && (unsigned)z < (unsigned)p.Z;
}
__device__ __forceinline__ int lin(const Params& p, int x, int y, int z) { return (z * p.Y + y) * p.X + x; }

__device__ __forceinline__ float comp(const float3& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }
__device__ __forceinline__ float3 ld3(const float* const a[3], int i) { return make_float3(a[0][i], a[1][i], a[2][i]); }

// boundary texture element at the face of cell (x, y, z) in direction dir
__device__ __forceinline__ float3 bnd_at(const Params& p, const Fields& f, int dir, int x, int y, int z) {
    const int bf = bface_of(dir);
    int k;
    if (bf < 2) k = z * p.Y + y;        // X faces: (y, z)
    else if (bf < 4) k = z * p.X + x;   // Y faces: (x, z)
    else k = y * p.X + x;               // Z faces: (x, y)
    return make_float3(f.bnd[bf][0][k], f.bnd[bf][1][k], f.bnd[bf][2][k]);
}

__device__ __forceinline__ bool fluid(const Fields& f, int i) { return f.flag[i] == UWS_CELL_FLUID; }
__device__ __forceinline__ bool solid(const Fields& f, int i) { return f.flag[i] == UWS_CELL_SOLID; }

// ---------------------------------------------------------------------------
// LogLawWallFunc (WallFunc.compute:77-123); the rough-wall branch is an
// extension: u_tau = kappa u_P / ln((dx/2)/z0).
// ---------------------------------------------------------------------------
__device__ float log_law_wall_func(const Params& p, float uP) {
    const float EConst = 9.8f, k = 0.41f;
    const float C = EConst * (p.dx / 2.0f) / p.nu;
    float eddyVis = 0.0f;
    float uTau = 1.0f;
    uP = fabsf(uP);
    if (uP < 1e-5f) {
        uTau = 0.0f;
    } else {
        if (p.wallModel == UWS_WALL_ROUGH) {
            uTau = k * uP / logf((p.dx / 2.0f) / p.wallZ0);
        } else {
            for (int i = 0; i < 10; ++i) {
                const float f = uTau * logf(C * uTau) - k * uP;
                const float df = logf(C * uTau) + 1.0f;
                if (fabsf(df) < 1e-6f) break;
                uTau = uTau - f / df;
                if (uTau < 1e-5f) uTau = 1e-5f;
            }
        }
        const float tauWall = uTau * uTau * p.den;
        const float strainRate = uP / (p.dx / 2.0f);
        eddyVis = tauWall / (strainRate * p.den) - p.nu;
        if (eddyVis < 0.0f) eddyVis = 0.0f;
    }
    return eddyVis;
}

// Get{E,W,N,S,U,D}FaceVelLastTime (LES.compute:104-258)
__device__ __forceinline__ float3 face_vel(const Params& p, const Fields& f, const float* const un[3], int x, int y,
                                           int z, int dir) {
    const int nx = x + c_off[dir][0], ny = y + c_off[dir][1], nz = z + c_off[dir][2];
    const int P = lin(p, x, y, z);
    const float3 uP = ld3(un, P);
    if (inreg(p, nx, ny, nz)) {
        const int Q = lin(p, nx, ny, nz);
        if (fluid(f, Q)) {
            const float3 uQ = ld3(un, Q);
            return make_float3(0.5f * (uP.x + uQ.x), 0.5f * (uP.y + uQ.y), 0.5f * (uP.z + uQ.z));
        }
        return make_float3(0.f, 0.f, 0.f);
    }
    const int t = p.velBnd[bface_of(dir)];
    if (t == UWS_VEL_FIXED_VALUE) return bnd_at(p, f, dir, x, y, z);
    if (t == UWS_VEL_ZERO_GRAD) return uP;
    if (t == UWS_VEL_SYMMETRY) {
        float3 v = uP;
        if (dir < 2) v.x = 0.f; else if (dir < 4) v.y = 0.f; else v.z = 0.f;
        return v;
    }
    return make_float3(0.f, 0.f, 0.f);
}

// cell velocity gradient g[i][j] = d u_i / d x_j from face velocities
// (LES.compute:294-306, 353-365)
__device__ __forceinline__ void cell_grad(const Params& p, const Fields& f, const float* const un[3], int x, int y,
                                          int z, float g[3][3]) {
    const float3 e = face_vel(p, f, un, x, y, z, E_), w = face_vel(p, f, un, x, y, z, W_);
    const float3 n = face_vel(p, f, un, x, y, z, N_), s = face_vel(p, f, un, x, y, z, S_);
    const float3 u = face_vel(p, f, un, x, y, z, U_), d = face_vel(p, f, un, x, y, z, D_);
    g[0][0] = (e.x - w.x) / p.dx; g[0][1] = (n.x - s.x) / p.dx; g[0][2] = (u.x - d.x) / p.dx;
    g[1][0] = (e.y - w.y) / p.dx; g[1][1] = (n.y - s.y) / p.dx; g[1][2] = (u.y - d.y) / p.dx;
    g[2][0] = (e.z - w.z) / p.dx; g[2][1] = (n.z - s.z) / p.dx; g[2][2] = (u.z - d.z) / p.dx;
}

// face viscosities with eddy viscosity (FvmColl.compute:1596-1635; LES.compute:643-679)
__device__ __forceinline__ void face_vis(const Params& p, const Fields& f, int x, int y, int z, int P, float vis[6]) {
    const float nu = p.nu;
    const float nutP = f.nut[P];
    const float base = nu + nutP;
#pragma unroll
    for (int d = 0; d < 6; ++d) vis[d] = base;
    // E: solid neighbour uses its .x slot; W: own .x slot (and N/S, U/D alike)
    if (x + 1 < p.X) {
        const int Q = P + 1;
        if (fluid(f, Q)) vis[E_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[E_] = nu + f.nuw[0][Q];
    }
    if (x - 1 >= 0) {
        const int Q = P - 1;
        if (fluid(f, Q)) vis[W_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[W_] = nu + f.nuw[0][P];
    }
    if (y + 1 < p.Y) {
        const int Q = P + p.X;
        if (fluid(f, Q)) vis[N_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[N_] = nu + f.nuw[1][Q];
    }
    if (y - 1 >= 0) {
        const int Q = P - p.X;
        if (fluid(f, Q)) vis[S_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[S_] = nu + f.nuw[1][P];
    } else if (p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE) {
        vis[S_] = nu + f.nuw[1][P];
    }
    const int sz = p.X * p.Y;
    if (z + 1 < p.Z) {
        const int Q = P + sz;
        if (fluid(f, Q)) vis[U_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[U_] = nu + f.nuw[2][Q];
    }
    if (z - 1 >= 0) {
        const int Q = P - sz;
        if (fluid(f, Q)) vis[D_] = nu + 0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[D_] = nu + f.nuw[2][P];
    }
}

// ---------------------------------------------------------------------------
// CSCalLogLawWallFunc (WallFunc.compute:125-190) and CSCalEddyVis
// (LES.compute:267-324), fused: both read only u^n and the flags.
// ---------------------------------------------------------------------------
template <bool WALL, bool EDDY>
__global__ __launch_bounds__(256) void k_wall_eddy(Params p, Fields f, const float* __restrict__ u0,
                                                   const float* __restrict__ u1, const float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const float* const un[3] = {u0, u1, u2};
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    if (WALL) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (isF) {
            const float3 uP = ld3(un, P);
            if (x - 1 >= 0 && solid(f, P - 1)) fx += log_law_wall_func(p, sqrtf(uP.y * uP.y + uP.z * uP.z));
            if ((y - 1 >= 0 && solid(f, P - p.X)) || (y - 1 < 0 && p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE))
                fy += log_law_wall_func(p, sqrtf(uP.x * uP.x + uP.z * uP.z));
            if (z - 1 >= 0 && solid(f, P - p.X * p.Y)) fz += log_law_wall_func(p, sqrtf(uP.x * uP.x + uP.y * uP.y));
        } else if (solid(f, P)) {
            if (x - 1 >= 0 && fluid(f, P - 1)) {
                const float3 v = ld3(un, P - 1);
                fx += log_law_wall_func(p, sqrtf(v.y * v.y + v.z * v.z));
            }
            if (y - 1 >= 0 && fluid(f, P - p.X)) {
                const float3 v = ld3(un, P - p.X);
                fy += log_law_wall_func(p, sqrtf(v.x * v.x + v.z * v.z));
            }
            if (z - 1 >= 0 && fluid(f, P - p.X * p.Y)) {
                const float3 v = ld3(un, P - p.X * p.Y);
                fz += log_law_wall_func(p, sqrtf(v.x * v.x + v.y * v.y));
            }
        }
        if (p.wallSlots == 1) { fy = 0.f; fz = 0.f; }
        f.nuw[0][P] = fx; f.nuw[1][P] = fy; f.nuw[2][P] = fz;
    }
    if (EDDY && isF) {
        float g[3][3];
        cell_grad(p, f, un, x, y, z, g);
        const float Sxx = g[0][0], Syy = g[1][1], Szz = g[2][2];
        const float Sxy = 0.5f * (g[0][1] + g[1][0]);
        const float Syz = 0.5f * (g[1][2] + g[2][1]);
        const float Szx = 0.5f * (g[2][0] + g[0][2]);
        const float Sabs = sqrtf(2.0f * (Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0f * Sxy * Sxy + 2.0f * Syz * Syz +
                                         2.0f * Szx * Szx));
        f.nut[P] = (p.cs * p.cs) * (p.dx * p.dx) * Sabs;
    }
}

// LES transposed-stress term of one fluid cell (LES.compute:326-683)
__device__ __forceinline__ float3 les_term(const Params& p, const Fields& f, const float* const un[3], int x, int y,
                                           int z, int P) {
    const float3 uP = ld3(un, P);
    const float hdx = 0.5f * p.dx;
    float g[3][3];
    cell_grad(p, f, un, x, y, z, g);
    float fg[6][3];
#pragma unroll
    for (int d = 0; d < 6; ++d) { fg[d][0] = 0.f; fg[d][1] = 0.f; fg[d][2] = 0.f; }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const float uPc = comp(uP, cc);
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                float gn[3][3];
                cell_grad(p, f, un, nx, ny, nz, gn);
#pragma unroll
                for (int j = 0; j < 3; ++j) fg[d][j] = 0.5f * (g[cc][j] + gn[cc][j]);
            } else {
                fg[d][cc] = plus ? -uPc / hdx : uPc / hdx;
            }
        } else {
            const int t = p.velBnd[bface_of(d)];
            if (t == UWS_VEL_FIXED_VALUE) {
                const float ubc = comp(bnd_at(p, f, d, x, y, z), cc);
                fg[d][cc] = plus ? (ubc - uPc) / hdx : (uPc - ubc) / hdx;
            } else if (t == UWS_VEL_ZERO_GRAD || t == UWS_VEL_SYMMETRY) {
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (j != cc) fg[d][j] = g[cc][j];
            }
        }
    }
    float vis[6];
    face_vis(p, f, x, y, z, P, vis);
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        r[i] = p.ds * (vis[0] * fg[0][i] - vis[1] * fg[1][i] + vis[2] * fg[2][i] - vis[3] * fg[3][i] +
                       vis[4] * fg[4][i] - vis[5] * fg[5][i]);
    return make_float3(r[0], r[1], r[2]);
}

// CSCalLesDeferCorrectTerm (LES.compute:326-687) -> bField
__global__ __launch_bounds__(256) void k_les_term(Params p, Fields f, const float* __restrict__ u0,
                                                  const float* __restrict__ u1, const float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok || !fluid(f, c.P)) return;
    const float* const un[3] = {u0, u1, u2};
    const float3 r = les_term(p, f, un, c.x, c.y, c.z, c.P);
    f.b[0][c.P] = r.x; f.b[1][c.P] = r.y; f.b[2][c.P] = r.z;
}

// ---------------------------------------------------------------------------
// CSVelPredictPreCompute (FvmColl.compute:1291-1859), optionally fused with
// the LES term (which it would otherwise read back from bField).  Besides D
// and b it stores the sweep's face coefficients S, G of the owned E/N/U
// faces (zero unless both cells are fluid) and 1/D.
// ---------------------------------------------------------------------------
template <bool FUSE_LES>
__global__ __launch_bounds__(256) void k_vel_precompute(Params p, Fields f, const float* __restrict__ u0,
                                                        const float* __restrict__ u1, const float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const float* const un[3] = {u0, u1, u2};
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const float3 uP = ld3(un, P);
    float vis[6];
    if (p.useLES && isF) face_vis(p, f, x, y, z, P, vis);
    else {
#pragma unroll
        for (int d = 0; d < 6; ++d) vis[d] = p.nu;
    }
    // owned-face coefficients for every cell
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float Sv = 0.f, Gv = 0.f;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                Sv = comp(uP, a) + un[a][Q];
                Gv = vis[dirs[a]] * p.ds / p.dx;
            }
            f.S[a][P] = Sv;
            f.G[a][P] = Gv;
        }
    }
    if (!isF) {
        f.invD[0][P] = 0.f; f.invD[1][P] = 0.f; f.invD[2][P] = 0.f;
        return;
    }
    float Dx = p.dv / p.dt, Dy = p.dv / p.dt, Dz = p.dv / p.dt;
    float bU = 0.f, bV = 0.f, bW = 0.f;
    const float ds = p.ds;
    // advection (:1314-1482)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const float uPn = comp(uP, cc);
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                const float t = ds * 0.25f * (uPn + un[cc][Q]);
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const float3 ub = bnd_at(p, f, d, x, y, z);
                const float ubn = comp(ub, cc);
                if (plus) { bU -= ds * ub.x * ubn; bV -= ds * ub.y * ubn; bW -= ds * ub.z * ubn; }
                else { bU += ds * ub.x * ubn; bV += ds * ub.y * ubn; bW += ds * ub.z * ubn; }
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                const float t = ds * uPn;
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        }
    }
    // pressure face values (:1484-1591)
    const float pP = f.p[P];
    float pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            pf[d] = fluid(f, Q) ? 0.5f * (pP + f.p[Q]) : pP;
        } else {
            pf[d] = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? 0.f : pP;
        }
    }
    bU += -(1.0f / p.den) * ds * (pf[E_] - pf[W_]);
    bV += -(1.0f / p.den) * ds * (pf[N_] - pf[S_]);
    bW += -(1.0f / p.den) * ds * (pf[U_] - pf[D_]);
    // diffusion (:1593-1836)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            const float t = fluid(f, Q) ? vis[d] * ds / p.dx : 2.0f * vis[d] * ds / p.dx;
            Dx += t; Dy += t; Dz += t;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const float3 ub = bnd_at(p, f, d, x, y, z);
                bU += 2.0f * vis[d] * ds / p.dx * ub.x;
                bV += 2.0f * vis[d] * ds / p.dx * ub.y;
                bW += 2.0f * vis[d] * ds / p.dx * ub.z;
                const float t = 2.0f * vis[d] * ds / p.dx;
                Dx += t; Dy += t; Dz += t;
            } else if (p.velBnd[bf] == UWS_VEL_SYMMETRY) {
                const float t = 2.0f * vis[d] * ds / p.dx;
                const int cc = d >> 1;
                if (cc == 0) Dx += t; else if (cc == 1) Dy += t; else Dz += t;
            }
        }
    }
    // time derivative, force, LES (:1838-1854)
    bU += p.dv / p.dt * uP.x; bV += p.dv / p.dt * uP.y; bW += p.dv / p.dt * uP.z;
    bU += p.dv * p.fext[0]; bV += p.dv * p.fext[1]; bW += p.dv * p.fext[2];
    if (p.useLES) {
        float3 bl;
        if (FUSE_LES) bl = les_term(p, f, un, x, y, z, P);
        else bl = make_float3(f.b[0][P], f.b[1][P], f.b[2][P]);
        bU += bl.x; bV += bl.y; bW += bl.z;
    }
    f.D[0][P] = Dx; f.D[1][P] = Dy; f.D[2][P] = Dz;
    f.invD[0][P] = 1.0f / Dx; f.invD[1][P] = 1.0f / Dy; f.invD[2][P] = 1.0f / Dz;
    f.b[0][P] = bU; f.b[1][P] = bV; f.b[2][P] = bW;
}

// ---------------------------------------------------------------------------
// CSVelPredict (FvmColl.compute:1862-2069): one Jacobi sweep.  With the face
// coefficients precomputed the accumulation order of the reference is kept:
// advection terms E W N S U D, then diffusion terms E W N S U D, then
// (1/D) * (adv + diff + b).  Non-fluid cells are not written.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_vel_sweep(Params p, Fields f, const float* __restrict__ k0,
                                                   const float* __restrict__ k1, const float* __restrict__ k2,
                                                   float* __restrict__ o0, float* __restrict__ o1,
                                                   float* __restrict__ o2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    const int sy = p.X, sz = p.X * p.Y;
    // neighbour indices, clamped into the grid (their coefficients are zero there)
    const int iE = c.x + 1 < p.X ? P + 1 : P, iW = c.x > 0 ? P - 1 : P;
    const int iN = c.y + 1 < p.Y ? P + sy : P, iS = c.y > 0 ? P - sy : P;
    const int iU = c.z + 1 < p.Z ? P + sz : P, iD = c.z > 0 ? P - sz : P;
    // a face on the domain boundary has no coefficient
    const float SE = f.S[0][P], SN = f.S[1][P], SU = f.S[2][P];
    const float GE = f.G[0][P], GN = f.G[1][P], GU = f.G[2][P];
    const float SW = c.x > 0 ? f.S[0][iW] : 0.f, GW = c.x > 0 ? f.G[0][iW] : 0.f;
    const float SS = c.y > 0 ? f.S[1][iS] : 0.f, GS = c.y > 0 ? f.G[1][iS] : 0.f;
    const float SD = c.z > 0 ? f.S[2][iD] : 0.f, GD = c.z > 0 ? f.G[2][iD] : 0.f;
    const float q = p.ds * 0.25f;
    const float* const kk[3] = {k0, k1, k2};
    float* const oo[3] = {o0, o1, o2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float* __restrict__ k = kk[a];
        const float vE = k[iE], vW = k[iW], vN = k[iN], vS = k[iS], vU = k[iU], vD = k[iD];
        float A = 0.f;
        A -= q * vE * SE;
        A += q * vW * SW;
        A -= q * vN * SN;
        A += q * vS * SS;
        A -= q * vU * SU;
        A += q * vD * SD;
        float Df = 0.f;
        Df += GE * vE;
        Df += GW * vW;
        Df += GN * vN;
        Df += GS * vS;
        Df += GU * vU;
        Df += GD * vD;
        oo[a][P] = f.invD[a][P] * (A + Df + f.b[a][P]);
    }
}

// ---------------------------------------------------------------------------
// CSPresCorrectRhieChowPreCompute (FvmColl.compute:2480-2787) -> Dp, bp,
// 1/Dp, and the pressure face coefficients of the owned E/N/U faces.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pres_precompute(Params p, Fields f, const float* __restrict__ s0,
                                                         const float* __restrict__ s1, const float* __restrict__ s2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const float* const u[3] = {s0, s1, s2};
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const float dvden = p.dv / p.den;
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float cv = 0.f, ce = 0.f;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                const float DInv = 0.5f * (f.invD[a][P] + f.invD[a][Q]);
                cv = dvden * DInv / p.dx;
                ce = dvden * (DInv / p.dx);
            }
            f.cp[a][P] = cv;
            if (a == 0) f.cpisoE[P] = ce;
        }
    }
    if (!isF) {
        f.invDp[P] = 0.f;
        return;
    }
    const float two_dx = 2.0f * p.dx;
    float Dp = 0.f, bp = 0.f;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int ox = c_off[d][0], oy = c_off[d][1], oz = c_off[d][2];
        const int nx = x + ox, ny = y + oy, nz = z + oz;
        const float uPc = u[cc][P];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (!fluid(f, Q)) continue;
            const float DInv = 0.5f * (f.invD[cc][P] + f.invD[cc][Q]);
            Dp -= dvden * DInv / p.dx;
            const float half = 0.5f * (uPc + u[cc][Q]);
            if (plus) bp += half; else bp -= half;
            const int fx = x + 2 * ox, fy = y + 2 * oy, fz = z + 2 * oz;
            const int bx = x - ox, by = y - oy, bz = z - oz;
            const bool farF = inreg(p, fx, fy, fz) && fluid(f, lin(p, fx, fy, fz));
            const bool backF = inreg(p, bx, by, bz) && fluid(f, lin(p, bx, by, bz));
            const float pP = f.p[P], pQ = f.p[Q];
            float g, ga;
            if (plus) {
                g = (pQ - pP) / p.dx;
                if (farF && backF) {
                    const float pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = 0.5f * ((pF - pP) / two_dx + (pQ - pB) / two_dx);
                } else if (farF) {
                    const float pF = f.p[lin(p, fx, fy, fz)];
                    ga = 0.5f * ((pF - pP) / two_dx + (pQ - pP) / two_dx);
                } else if (backF) {
                    const float pB = f.p[lin(p, bx, by, bz)];
                    ga = 0.5f * ((pQ - pP) / two_dx + (pQ - pB) / two_dx);
                } else {
                    ga = g;
                }
            } else {
                g = (pP - pQ) / p.dx;
                if (farF && backF) {
                    const float pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = 0.5f * ((pB - pQ) / two_dx + (pP - pF) / two_dx);
                } else if (farF) {
                    const float pF = f.p[lin(p, fx, fy, fz)];
                    ga = 0.5f * ((pP - pQ) / two_dx + (pP - pF) / two_dx);
                } else if (backF) {
                    const float pB = f.p[lin(p, bx, by, bz)];
                    ga = 0.5f * ((pB - pQ) / two_dx + (pP - pQ) / two_dx);
                } else {
                    ga = g;
                }
            }
            const float vct = -(p.dv * DInv) / p.den * (g - ga);
            if (plus) bp += vct; else bp -= vct;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const float ubc = comp(bnd_at(p, f, d, x, y, z), cc);
                if (plus) bp += ubc; else bp -= ubc;
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                if (plus) bp += uPc; else bp -= uPc;
            }
            if (p.presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                const float DInv = 1.0f / f.D[cc][P];
                Dp -= 2.0f * dvden * DInv / p.dx;
            }
        }
    }
    f.Dp[P] = Dp;
    f.invDp[P] = 1.0f / Dp;
    f.bp[P] = bp;
}

// ---------------------------------------------------------------------------
// CSPresCorrectRhieChow (FvmColl.compute:2789-2848) and CSPisoPresCorrection
// (PisoColl.compute:180-256): p_new = (1/diag) * (-sum_f c_f p_nb + rhs),
// terms E W N S U D.  pk == nullptr means the iterate is zero (first sweep
// after SetPresCorrectFieldtoZero).  Non-fluid cells get 0.
// ---------------------------------------------------------------------------
template <bool FIRST>
__global__ __launch_bounds__(256) void k_pres_sweep(Params p, Fields f, const float* __restrict__ cE,
                                                    const float* __restrict__ cN, const float* __restrict__ cU,
                                                    const float* __restrict__ invDp, const float* __restrict__ bp,
                                                    const float* __restrict__ pk, float* __restrict__ pnew) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) {
        pnew[P] = 0.f;
        return;
    }
    float term = 0.f;
    if (!FIRST) {
        const int sy = p.X, sz = p.X * p.Y;
        const int iE = c.x + 1 < p.X ? P + 1 : P, iW = c.x > 0 ? P - 1 : P;
        const int iN = c.y + 1 < p.Y ? P + sy : P, iS = c.y > 0 ? P - sy : P;
        const int iU = c.z + 1 < p.Z ? P + sz : P, iD = c.z > 0 ? P - sz : P;
        const float cW = c.x > 0 ? cE[iW] : 0.f, cS = c.y > 0 ? cN[iS] : 0.f, cD = c.z > 0 ? cU[iD] : 0.f;
        term -= cE[P] * pk[iE];
        term -= cW * pk[iW];
        term -= cN[P] * pk[iN];
        term -= cS * pk[iS];
        term -= cU[P] * pk[iU];
        term -= cD * pk[iD];
    }
    pnew[P] = invDp[P] * (term + bp[P]);
}

// CSApplyPresCorrection (FvmColl.compute:2850-2862): p += p' on non-solid cells
template <bool NORMALISE>
__global__ __launch_bounds__(256) void k_apply_pres(Params p, Fields f, const float* __restrict__ pc, int refIdx) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    if (solid(f, c.P)) {
        if (NORMALISE && c.P == refIdx) f.scalar[0] = f.p[c.P];
        return;
    }
    const float v = f.p[c.P] + pc[c.P];
    f.p[c.P] = v;
    if (NORMALISE && c.P == refIdx) f.scalar[0] = v;
}
// CSPresNormalization (FvmColl.compute:2864-2880) with the reference value
// captured by k_apply_pres before any cell is changed
__global__ __launch_bounds__(256) void k_pres_normalise(Params p, Fields f) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    if (solid(f, c.P)) return;
    f.p[c.P] -= f.scalar[0];
}

// CSApplyVelCorrectionDivergenceTheorem (FvmColl.compute:2950-3121)
__global__ __launch_bounds__(256) void k_vel_correct(Params p, Fields f, const float* __restrict__ pc,
                                                     float* __restrict__ u0, float* __restrict__ u1,
                                                     float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    if (!fluid(f, P)) return;
    const float pcP = pc[P];
    float pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            pf[d] = fluid(f, Q) ? 0.5f * (pcP + pc[Q]) : pcP;
        } else {
            pf[d] = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? 0.f : pcP;
        }
    }
    const float dvden = p.dv / p.den;
    const float cx = -dvden * (1.0f / f.D[0][P]) * (pf[E_] - pf[W_]) / p.dx;
    const float cy = -dvden * (1.0f / f.D[1][P]) * (pf[N_] - pf[S_]) / p.dx;
    const float cz = -dvden * (1.0f / f.D[2][P]) * (pf[U_] - pf[D_]) / p.dx;
    f.ucorr[0][P] = cx; f.ucorr[1][P] = cy; f.ucorr[2][P] = cz;
    u0[P] += cx; u1[P] += cy; u2[P] += cz;
}

// CSPisoCalAodUCorrect (PisoColl.compute:70-178): laminar nu, no BC terms
__global__ __launch_bounds__(256) void k_piso_aodu(Params p, Fields f, const float* __restrict__ u0,
                                                   const float* __restrict__ u1, const float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    const float* const un[3] = {u0, u1, u2};
    const float3 uP = ld3(un, P);
    float A[3] = {0.f, 0.f, 0.f}, Df[3] = {0.f, 0.f, 0.f};
    const float q = p.ds * 0.25f;
    int Qs[6];
    bool ok[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = c.x + c_off[d][0], ny = c.y + c_off[d][1], nz = c.z + c_off[d][2];
        ok[d] = inreg(p, nx, ny, nz) && fluid(f, lin(p, nx, ny, nz));
        Qs[d] = ok[d] ? lin(p, nx, ny, nz) : P;
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!ok[d]) continue;
        const int cc = d >> 1;
        const int Q = Qs[d];
        const float sum = comp(uP, cc) + un[cc][Q];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float t = q * f.ucorr[a][Q] * sum;
            if ((d & 1) == 0) A[a] += t; else A[a] -= t;
        }
    }
    const float g = p.nu * p.ds / p.dx;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!ok[d]) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) Df[a] -= g * f.ucorr[a][Qs[d]];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) f.aodu[a][P] = A[a] + Df[a];
}

// CSPisoPresCorrection's constant parts (PisoColl.compute:196-255): presCoeff
// and velTerm, which do not change over the sweeps, computed once.
__global__ __launch_bounds__(256) void k_piso_precompute(Params p, Fields f) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    const int sy = p.X, sz = p.X * p.Y;
    float presCoeff = 0.f, velTerm = 0.f;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const int nx = c.x + c_off[d][0], ny = c.y + c_off[d][1], nz = c.z + c_off[d][2];
        if (!(inreg(p, nx, ny, nz) && fluid(f, lin(p, nx, ny, nz)))) continue;
        const int Q = lin(p, nx, ny, nz);
        const float AdInv = 0.5f * (f.invD[cc][P] + f.invD[cc][Q]);
        const float AodUCorrect = 0.5f * (f.aodu[cc][P] + f.aodu[cc][Q]);
        // coefficient of this face as stored by k_pres_precompute on the owning cell
        const int owner = (d & 1) == 0 ? P : Q;
        const float cf = cc == 0 ? f.cpisoE[owner] : f.cp[cc][owner];
        presCoeff -= cf;
        if ((d & 1) == 0) velTerm -= AdInv * AodUCorrect; else velTerm += AdInv * AodUCorrect;
        (void)sy; (void)sz;
    }
    f.invDpiso[P] = 1.0f / presCoeff;
    f.bpiso[P] = velTerm;
}

// CSPisoApplyPresCorrection + CSPisoVelCorrection (PisoColl.compute:258-337)
__global__ __launch_bounds__(256) void k_piso_correct(Params p, Fields f, const float* __restrict__ pc,
                                                      float* __restrict__ u0, float* __restrict__ u1,
                                                      float* __restrict__ u2) {
    const Cell c = cell_of(p);
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    const float pcP = pc[P];
    f.p[P] += pcP;
    float pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = c.x + c_off[d][0], ny = c.y + c_off[d][1], nz = c.z + c_off[d][2];
        pf[d] = (inreg(p, nx, ny, nz) && fluid(f, lin(p, nx, ny, nz))) ? 0.5f * (pcP + pc[lin(p, nx, ny, nz)]) : pcP;
    }
    const float gx = (pf[E_] - pf[W_]) / p.dx, gy = (pf[N_] - pf[S_]) / p.dx, gz = (pf[U_] - pf[D_]) / p.dx;
    const float dvden = p.dv / p.den;
    const float cx = -f.invD[0][P] * (f.aodu[0][P] + dvden * gx);
    const float cy = -f.invD[1][P] * (f.aodu[1][P] + dvden * gy);
    const float cz = -f.invD[2][P] * (f.aodu[2][P] + dvden * gz);
    f.ucorr[0][P] = cx; f.ucorr[1][P] = cy; f.ucorr[2][P] = cz;
    u0[P] += cx; u1[P] += cy; u2[P] += cz;
}

// FP64 block reduction of sum_c (d_c (a_c - b_c))^2 (CSComputeVelPredictResidual
// / CSComputePresCorrectResidual, FvmColl.compute:3123-3199), one partial per
// workgroup; a second pass sums the partials in a fixed order.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__global__ __launch_bounds__(256) void k_residual(Params p, int ncomp, const float* __restrict__ d0,
                                                  const float* __restrict__ d1, const float* __restrict__ d2,
                                                  const float* __restrict__ a0, const float* __restrict__ a1,
                                                  const float* __restrict__ a2, const float* __restrict__ b0,
                                                  const float* __restrict__ b1, const float* __restrict__ b2,
                                                  double* __restrict__ partial) {
    __shared__ double ws[4];
    const Cell c = cell_of(p);
    double s = 0.0;
    if (c.ok) {
        const float* const dd[3] = {d0, d1, d2};
        const float* const aa[3] = {a0, a1, a2};
        const float* const bb[3] = {b0, b1, b2};
        for (int k = 0; k < ncomp; ++k) {
            const float r = dd[k][c.P] * (aa[k][c.P] - (bb[k] ? bb[k][c.P] : 0.f));
            s += (double)(r * r);
        }
    }
    s = wave_sum(s);
    const int lane = threadIdx.x & 63, wid = (threadIdx.y * TX + threadIdx.x) >> 6;
    if (lane == 0) ws[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
    __shared__ double ws[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ws[0] + ws[1] + ws[2] + ws[3];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define UWS_LAUNCH(kernel, p, s, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid_blocks(p)), dim3(TX, TY, 1), 0, s, __VA_ARGS__)

void launch_wall_eddy(const Params& p, const Fields& f, const float* const un[3], bool wall, bool eddy,
                      hipStream_t s) {
    if (wall && eddy) UWS_LAUNCH((k_wall_eddy<true, true>), p, s, p, f, un[0], un[1], un[2]);
    else if (wall) UWS_LAUNCH((k_wall_eddy<true, false>), p, s, p, f, un[0], un[1], un[2]);
    else if (eddy) UWS_LAUNCH((k_wall_eddy<false, true>), p, s, p, f, un[0], un[1], un[2]);
}
void launch_les_term(const Params& p, const Fields& f, const float* const un[3], hipStream_t s) {
    UWS_LAUNCH(k_les_term, p, s, p, f, un[0], un[1], un[2]);
}
void launch_vel_precompute(const Params& p, const Fields& f, const float* const un[3], bool fuse_les,
                           hipStream_t s) {
    if (fuse_les) UWS_LAUNCH((k_vel_precompute<true>), p, s, p, f, un[0], un[1], un[2]);
    else UWS_LAUNCH((k_vel_precompute<false>), p, s, p, f, un[0], un[1], un[2]);
}
void launch_vel_sweep(const Params& p, const Fields& f, const float* const uk[3], float* const unew[3],
                      hipStream_t s) {
    UWS_LAUNCH(k_vel_sweep, p, s, p, f, uk[0], uk[1], uk[2], unew[0], unew[1], unew[2]);
}
void launch_pres_precompute(const Params& p, const Fields& f, const float* const ustar[3], hipStream_t s) {
    UWS_LAUNCH(k_pres_precompute, p, s, p, f, ustar[0], ustar[1], ustar[2]);
}
void launch_pres_sweep(const Params& p, const Fields& f, const float* cE, const float* cN, const float* cU,
                       const float* invDp, const float* bp, const float* pk, float* pnew, hipStream_t s) {
    if (pk) UWS_LAUNCH((k_pres_sweep<false>), p, s, p, f, cE, cN, cU, invDp, bp, pk, pnew);
    else UWS_LAUNCH((k_pres_sweep<true>), p, s, p, f, cE, cN, cU, invDp, bp, pk, pnew);
}
void launch_apply_pres(const Params& p, const Fields& f, const float* pc, bool normalise, hipStream_t s) {
    if (normalise) {
        // reference cell (2, Y/2, 2); a texture read outside the grid returns 0,
        // which makes the normalisation a no-op
        const int ref = (2 < p.X && 2 < p.Z) ? (2 * p.Y + p.Y / 2) * p.X + 2 : -1;
        UWS_LAUNCH((k_apply_pres<true>), p, s, p, f, pc, ref);
        if (ref >= 0) UWS_LAUNCH(k_pres_normalise, p, s, p, f);
    } else {
        UWS_LAUNCH((k_apply_pres<false>), p, s, p, f, pc, -1);
    }
}
void launch_vel_correct(const Params& p, const Fields& f, const float* pc, float* const u[3], hipStream_t s) {
    UWS_LAUNCH(k_vel_correct, p, s, p, f, pc, u[0], u[1], u[2]);
}
void launch_piso_aodu(const Params& p, const Fields& f, const float* const un[3], hipStream_t s) {
    UWS_LAUNCH(k_piso_aodu, p, s, p, f, un[0], un[1], un[2]);
}
void launch_piso_precompute(const Params& p, const Fields& f, hipStream_t s) {
    UWS_LAUNCH(k_piso_precompute, p, s, p, f);
}
void launch_piso_correct(const Params& p, const Fields& f, const float* pc, float* const u[3], hipStream_t s) {
    UWS_LAUNCH(k_piso_correct, p, s, p, f, pc, u[0], u[1], u[2]);
}
void launch_residual(const Params& p, const Fields& f, int ncomp, const float* const d[3], const float* const a[3],
                     const float* const b[3], double* out, hipStream_t s) {
    const int nb = grid_blocks(p);
    UWS_LAUNCH(k_residual, p, s, p, ncomp, d[0], d[1], d[2], a[0], a[1], a[2], b[0], b[1], b[2], f.partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, f.partial, nb, out);
}

}  // namespace uws

namespace uws {
namespace {
// CSInitModelFlag's velocity reset (InitShaderColl.compute:353-359) applied to
// every velocity buffer after a flag upload
__global__ __launch_bounds__(256) void k_zero_solid_vel(Params p, Fields f) {
    const Cell c = cell_of(p);
    if (!c.ok || !solid(f, c.P)) return;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        f.vel[b][0][c.P] = 0.f; f.vel[b][1][c.P] = 0.f; f.vel[b][2][c.P] = 0.f;
    }
}
__global__ __launch_bounds__(256) void k_fill(float* __restrict__ dst, int n, float v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = v;
}
}  // namespace

void launch_zero_solid_vel(const Params& p, const Fields& f, hipStream_t s) {
    UWS_LAUNCH(k_zero_solid_vel, p, s, p, f);
}
void launch_fill(float* dst, int n, float v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, dst, n, v);
}
}  // namespace uws
