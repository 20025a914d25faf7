// gfx950 kernels of the collocated finite-volume solver step.
//
// Each kernel body restates one reference compute kernel (cited per body)
// with the reference's operation order, so that with -ffp-contract=off the
// FP32 results match the CPU restatement bit for bit where the same
// arithmetic is used.  Bodies are templated on the arithmetic type T (float
// for the product path, double for FP64 reference runs on the GPU).
//
// Layout: structure-of-arrays planes, flat cell index z*X*Y + y*X + x
// (FvmColl.compute:122-126), one byte of flag per cell.
//
// Launch shape: 64 x 4 threads per workgroup (one wave per x-row of 64
// cells, four rows), a 1-D grid whose workgroup ids are remapped so that each
// of the 8 XCDs receives a contiguous range of (x, y, z) tiles and a
// stencil's neighbour rows and planes stay in that XCD's L2.  The same
// bodies run either over one grid (k_single) or over every case of an
// ensemble in one launch (k_batch, driven by a tile table).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "uws_adjoint.h"
#include "uws_internal.h"

namespace uws {

namespace {

// workgroups are 64 x ty threads: one wave per x-row of 64 cells, ty rows
// (ty in {1, 2, 4}, chosen per solver or per ensemble)
constexpr int TX = 64;
constexpr int SINGLE_TY = 4;  // tile rows of single-case launches (ParamsT::ty of a handle)

enum { E_ = 0, W_ = 1, N_ = 2, S_ = 3, U_ = 4, D_ = 5 };
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
    int x, y, z, P, tile;
    bool ok;
};

// TYC > 0: tile rows known at compile time (single-case launches use 4)
template <int TYC = 0, class T>
__device__ __forceinline__ Cell cell_of(const ParamsT<T>& p, int t) {
    const int ty = TYC > 0 ? TYC : p.ty;
    const int gx = (p.X + TX - 1) / TX, gy = (p.Y + ty - 1) / ty;
    const int bx = t % gx, r = t / gx, by = r % gy, bz = r / gy;
    Cell c;
    c.tile = t;
    c.x = bx * TX + (int)threadIdx.x;
    c.y = by * ty + (int)threadIdx.y;
    c.z = bz;
    c.ok = c.x < p.X && c.y < p.Y;
    c.P = (c.z * p.Y + c.y) * p.X + c.x;
    return c;
}

template <class T>
__device__ __forceinline__ bool inreg(const ParamsT<T>& p, int x, int y, int z) {
    return (unsigned)x < (unsigned)p.X && (unsigned)y < (unsigned)p.Y && (unsigned)z < (unsigned)p.Z;
}
template <class T>
__device__ __forceinline__ int lin(const ParamsT<T>& p, int x, int y, int z) { return (z * p.Y + y) * p.X + x; }

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
__device__ __forceinline__ V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T>
__device__ __forceinline__ T comp(const V3<T>& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }
template <class T>
__device__ __forceinline__ V3<T> ld3(T* const a[3], int i) { return V3<T>{a[0][i], a[1][i], a[2][i]}; }

template <class T>
__device__ __forceinline__ V3<T> bnd_at(const ParamsT<T>& p, const FieldsT<T>& f, int dir, int x, int y, int z) {
    const int bf = bface_of(dir);
    int k;
    if (bf < 2) k = z * p.Y + y;        // X faces: (y, z)
    else if (bf < 4) k = z * p.X + x;   // Y faces: (x, z)
    else k = y * p.X + x;               // Z faces: (x, y)
    return V3<T>{f.bnd[bf][0][k], f.bnd[bf][1][k], f.bnd[bf][2][k]};
}

template <class T>
__device__ __forceinline__ bool fluid(const FieldsT<T>& f, int i) { return f.flag[i] == UWS_CELL_FLUID; }
template <class T>
__device__ __forceinline__ bool solid(const FieldsT<T>& f, int i) { return f.flag[i] == UWS_CELL_SOLID; }

// x / d for a uniform divisor d (dx, 2 dx, dx / 2, den): with P2 (d a power
// of two, plain float types) the product with its exact reciprocal rd, which
// is the same real number and so rounds the same
template <bool P2, class T>
__device__ __forceinline__ T quo(T x, T d, T rd) {
    if constexpr (P2) return x * rd;
    else return x / d;
}
// body B<T, true> where ParamsT::p2 is set (float, double), else B<T, false>
#define UWS_P2_DISPATCH(B)                                                  \
    if constexpr (std::is_same_v<T, float> || std::is_same_v<T, double>) { \
        if (p.p2) {                                                         \
            B<T, true>(p, f, c, r);                                         \
            return;                                                         \
        }                                                                   \
    }                                                                       \
    B<T, false>(p, f, c, r);

// The six neighbours of cell c in the face order E W N S U D: their index
// (the cell itself where the face is on the domain boundary), whether they
// lie in the grid, and whether they are fluid.  All six flags are loaded
// unconditionally, so the loads that depend on them can be issued together.
template <class T>
__device__ __forceinline__ void nbr6(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, int Q[6], bool in[6],
                                     bool fl[6]) {
    const int P = c.P;
    in[E_] = c.x + 1 < p.X; in[W_] = c.x > 0;
    in[N_] = c.y + 1 < p.Y; in[S_] = c.y > 0;
    in[U_] = c.z + 1 < p.Z; in[D_] = c.z > 0;
    const int st[6] = {1, -1, p.X, -p.X, p.X * p.Y, -p.X * p.Y};
    uint8_t fv[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        Q[d] = in[d] ? P + st[d] : P;
        fv[d] = f.flag[Q[d]];
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) fl[d] = in[d] && fv[d] == UWS_CELL_FLUID;
}
// nbr6 plus which neighbours are solid cells in the grid
struct Nbr6 {
    int Q[6];
    bool in[6], fl[6], sl[6];
};
template <class T>
__device__ __forceinline__ Nbr6 nbrs_of(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c) {
    Nbr6 n;
    const int P = c.P;
    n.in[E_] = c.x + 1 < p.X; n.in[W_] = c.x > 0;
    n.in[N_] = c.y + 1 < p.Y; n.in[S_] = c.y > 0;
    n.in[U_] = c.z + 1 < p.Z; n.in[D_] = c.z > 0;
    const int st[6] = {1, -1, p.X, -p.X, p.X * p.Y, -p.X * p.Y};
    uint8_t fv[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        n.Q[d] = n.in[d] ? P + st[d] : P;
        fv[d] = f.flag[n.Q[d]];
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        n.fl[d] = n.in[d] && fv[d] == UWS_CELL_FLUID;
        n.sl[d] = n.in[d] && fv[d] == UWS_CELL_SOLID;
    }
    return n;
}

// the neighbours of cell c that are fluid cells in the grid (index: the cell
// itself for the other faces)
template <class T>
__device__ __forceinline__ void fluid_nbrs(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, int Qs[6],
                                           bool ok[6]) {
    bool in[6];
    nbr6(p, f, c, Qs, in, ok);
}

// ---------------------------------------------------------------------------
// LogLawWallFunc (WallFunc.compute:77-123); the rough-wall branch is an
// extension: u_tau = kappa u_P / ln((dx/2)/z0).
// ---------------------------------------------------------------------------
template <class T, bool P2 = false>
__device__ T log_law_wall_func(const ParamsT<T>& p, T uP) {
    const T EConst = (T)9.8f, k = (T)0.41f;
    const T C = EConst * (p.dx / (T)2.0f) / p.nu;
    T eddyVis = 0;
    T uTau = 1;
    uP = fabs(uP);
    if (uP < (T)1e-5f) {
        uTau = 0;
    } else {
        if (p.wallModel == UWS_WALL_ROUGH) {
            uTau = k * uP / log((p.dx / (T)2.0f) / p.wallZ0);
        } else {
            for (int i = 0; i < 10; ++i) {
                const T fv = uTau * log(C * uTau) - k * uP;
                const T df = log(C * uTau) + (T)1.0f;
                if (fabs(df) < (T)1e-6f) break;
                uTau = uTau - fv / df;
                if (uTau < (T)1e-5f) uTau = (T)1e-5f;
            }
        }
        const T tauWall = uTau * uTau * p.den;
        const T strainRate = quo<P2>(uP, p.dx / (T)2.0f, p.rhdx);
        eddyVis = tauWall / (strainRate * p.den) - p.nu;
        if (eddyVis < 0) eddyVis = 0;
    }
    return eddyVis;
}

// Get{E,W,N,S,U,D}FaceVelLastTime (LES.compute:104-258): the six face
// velocities of a fluid cell -- the average with a fluid neighbour, 0 at a
// solid one, and by BC type on the domain boundary (fixed: the boundary
// texture, zero gradient: u_P, symmetry: u_P without its normal component) --
// from neighbour data loaded unconditionally (Q, in, fl from nbr6)
template <class T>
__device__ __forceinline__ void face_vel6(const ParamsT<T>& p, const FieldsT<T>& f, T* const un[3], const Cell& c,
                                          const int Q[6], const bool in[6], const bool fl[6], V3<T> fv[6]) {
    const V3<T> uP = ld3(un, c.P);
    V3<T> uQ[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) uQ[d] = ld3(un, Q[d]);
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int t = p.velBnd[bface_of(d)];
        V3<T> bv = mk<T>(0, 0, 0);
        if (t == UWS_VEL_FIXED_VALUE) {
            bv = bnd_at(p, f, d, c.x, c.y, c.z);
        } else if (t == UWS_VEL_ZERO_GRAD) {
            bv = uP;
        } else if (t == UWS_VEL_SYMMETRY) {
            bv = uP;
            if (d < 2) bv.x = 0; else if (d < 4) bv.y = 0; else bv.z = 0;
        }
        const V3<T> avg = mk<T>((T)0.5f * (uP.x + uQ[d].x), (T)0.5f * (uP.y + uQ[d].y), (T)0.5f * (uP.z + uQ[d].z));
        const V3<T> zero = mk<T>(0, 0, 0);
        fv[d] = in[d] ? (fl[d] ? avg : zero) : bv;
    }
}

// cell velocity gradient g[i][j] = d u_i / d x_j from face velocities
// (LES.compute:294-306, 353-365)
template <class T, bool P2 = false>
__device__ __forceinline__ void cell_grad(const ParamsT<T>& p, const FieldsT<T>& f, T* const un[3], const Cell& c,
                                          const int Q[6], const bool in[6], const bool fl[6], T g[3][3]) {
    V3<T> fv[6];
    face_vel6(p, f, un, c, Q, in, fl, fv);
    const V3<T> e = fv[E_], w = fv[W_], n = fv[N_], s = fv[S_], u = fv[U_], d = fv[D_];
    const T dx = p.dx, rdx = p.rdx;
    g[0][0] = quo<P2>(e.x - w.x, dx, rdx); g[0][1] = quo<P2>(n.x - s.x, dx, rdx); g[0][2] = quo<P2>(u.x - d.x, dx, rdx);
    g[1][0] = quo<P2>(e.y - w.y, dx, rdx); g[1][1] = quo<P2>(n.y - s.y, dx, rdx); g[1][2] = quo<P2>(u.y - d.y, dx, rdx);
    g[2][0] = quo<P2>(e.z - w.z, dx, rdx); g[2][1] = quo<P2>(n.z - s.z, dx, rdx); g[2][2] = quo<P2>(u.z - d.z, dx, rdx);
}

// face viscosities with eddy viscosity (FvmColl.compute:1596-1635; LES.compute:643-679)
template <class T>
__device__ __forceinline__ void face_vis(const ParamsT<T>& p, const FieldsT<T>& f, const Nbr6& nb, int P,
                                         T vis[6]) {
    const T nu = p.nu;
    const T nutP = f.nut[P];
    const T base = nu + nutP;
    // E/N/U: a solid neighbour contributes its own W/S/D slot; W/S/D: this cell's slot
    T nutQ[6], wall[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        nutQ[d] = f.nut[nb.Q[d]];
        wall[d] = f.nuw[d >> 1][(d & 1) == 0 ? nb.Q[d] : P];
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const T vf = nu + (T)0.5f * (nutP + nutQ[d]);
        const T vs = nu + wall[d];
        vis[d] = nb.fl[d] ? vf : (nb.sl[d] ? vs : base);
    }
    // the ground below the domain is a wall when Y0 is a fixed-value face
    if (!nb.in[S_] && p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE) vis[S_] = nu + wall[S_];
}

// ---------------------------------------------------------------------------
// CSCalLogLawWallFunc (WallFunc.compute:125-190) and CSCalEddyVis
// (LES.compute:267-324), fused: both read only u^n and the flags.
// ---------------------------------------------------------------------------
template <class T, bool P2>
__device__ void body_wall_eddy_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* un = f.vel[r.state];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    if (r.flags & K_WALL) {
        T fx = 0, fy = 0, fz = 0;
        if (isF) {
            const V3<T> uP = ld3(un, P);
            if (x - 1 >= 0 && solid(f, P - 1)) fx += log_law_wall_func<T, P2>(p, sqrt(uP.y * uP.y + uP.z * uP.z));
            if ((y - 1 >= 0 && solid(f, P - p.X)) || (y - 1 < 0 && p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE))
                fy += log_law_wall_func<T, P2>(p, sqrt(uP.x * uP.x + uP.z * uP.z));
            if (z - 1 >= 0 && solid(f, P - p.X * p.Y)) fz += log_law_wall_func<T, P2>(p, sqrt(uP.x * uP.x + uP.y * uP.y));
        } else if (solid(f, P)) {
            if (x - 1 >= 0 && fluid(f, P - 1)) {
                const V3<T> v = ld3(un, P - 1);
                fx += log_law_wall_func<T, P2>(p, sqrt(v.y * v.y + v.z * v.z));
            }
            if (y - 1 >= 0 && fluid(f, P - p.X)) {
                const V3<T> v = ld3(un, P - p.X);
                fy += log_law_wall_func<T, P2>(p, sqrt(v.x * v.x + v.z * v.z));
            }
            if (z - 1 >= 0 && fluid(f, P - p.X * p.Y)) {
                const V3<T> v = ld3(un, P - p.X * p.Y);
                fz += log_law_wall_func<T, P2>(p, sqrt(v.x * v.x + v.y * v.y));
            }
        }
        // a float4 written into the reference's single-channel texture keeps only .x
        if (p.wallSlots == 1) { fy = 0; fz = 0; }
        f.nuw[0][P] = fx; f.nuw[1][P] = fy; f.nuw[2][P] = fz;
    }
    if ((r.flags & K_EDDY) && isF) {
        int Q[6];
        bool in[6], fl[6];
        nbr6(p, f, c, Q, in, fl);
        T g[3][3];
        cell_grad<T, P2>(p, f, un, c, Q, in, fl, g);
        // kept for the LES term, which needs it at this cell and its neighbours
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) f.grad[3 * i + j][P] = g[i][j];
        const T Sxx = g[0][0], Syy = g[1][1], Szz = g[2][2];
        const T Sxy = (T)0.5f * (g[0][1] + g[1][0]);
        const T Syz = (T)0.5f * (g[1][2] + g[2][1]);
        const T Szx = (T)0.5f * (g[2][0] + g[0][2]);
        const T Sabs = sqrt((T)2.0f * (Sxx * Sxx + Syy * Syy + Szz * Szz + (T)2.0f * Sxy * Sxy +
                                       (T)2.0f * Syz * Syz + (T)2.0f * Szx * Szx));
        f.nut[P] = (p.cs * p.cs) * (p.dx * p.dx) * Sabs;
    }
}

template <class T>
__device__ void body_wall_eddy(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    UWS_P2_DISPATCH(body_wall_eddy_t)
}

// the cell gradient stored by the eddy-viscosity kernel
template <class T>
__device__ __forceinline__ void load_grad(const FieldsT<T>& f, int P, T g[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) g[i][j] = f.grad[3 * i + j][P];
}

// complex-step and dual types take the per-face sequential forms of the
// neighbour-heavy kernels (uws_seq.h)
template <class T>
constexpr bool plain_v = std::is_same_v<T, float> || std::is_same_v<T, double>;
#include "uws_seq.h"

// LES transposed-stress term of one fluid cell (LES.compute:326-683).  The
// cell gradients of this cell and its fluid neighbours are those the
// eddy-viscosity kernel computed from the same u^n (cell_grad), read back.
template <class T, bool P2 = false>
__device__ __forceinline__ V3<T> les_term(const ParamsT<T>& p, const FieldsT<T>& f, T* const un[3], const Cell& c,
                                          const Nbr6& nb) {
    const int P = c.P;
    const V3<T> uP = ld3(un, P);
    const T hdx = (T)0.5f * p.dx;
    T g[3][3];
    load_grad(f, P, g);
    // row cc of each neighbour's gradient (cc: the face's axis), loaded unconditionally
    T gq[6][3];
    T ubn[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
#pragma unroll
        for (int j = 0; j < 3; ++j) gq[d][j] = f.grad[3 * cc + j][nb.Q[d]];
        const int bf = bface_of(d);
        const int k = bf < 2 ? c.z * p.Y + c.y : (bf < 4 ? c.z * p.X + c.x : c.y * p.X + c.x);
        // boundary textures: read on the domain boundary only
        ubn[d] = 0;
        if (!nb.in[d] && p.velBnd[bf] == UWS_VEL_FIXED_VALUE) ubn[d] = f.bnd[bf][cc][k];
    }
    T fg[6][3];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const T uPc = comp(uP, cc);
        // fluid neighbour: the average of the two cell gradients
        T ff[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) ff[j] = (T)0.5f * (g[cc][j] + gq[d][j]);
        // solid neighbour: only the normal derivative, towards u = 0 at the wall
        T fs[3] = {0, 0, 0};
        fs[cc] = plus ? quo<P2>(-uPc, hdx, p.rhdx) : quo<P2>(uPc, hdx, p.rhdx);
        // domain boundary, by BC type
        T fb[3] = {0, 0, 0};
        const int t = p.velBnd[bface_of(d)];
        if (t == UWS_VEL_FIXED_VALUE) {
            fb[cc] = plus ? quo<P2>(ubn[d] - uPc, hdx, p.rhdx) : quo<P2>(uPc - ubn[d], hdx, p.rhdx);
        } else if (t == UWS_VEL_ZERO_GRAD || t == UWS_VEL_SYMMETRY) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (j != cc) fb[j] = g[cc][j];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) fg[d][j] = nb.in[d] ? (nb.fl[d] ? ff[j] : fs[j]) : fb[j];
    }
    T vis[6];
    face_vis(p, f, nb, P, vis);
    T rr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        rr[i] = p.ds * (vis[0] * fg[0][i] - vis[1] * fg[1][i] + vis[2] * fg[2][i] - vis[3] * fg[3][i] +
                        vis[4] * fg[4][i] - vis[5] * fg[5][i]);
    return mk<T>(rr[0], rr[1], rr[2]);
}

// CSCalLesDeferCorrectTerm (LES.compute:326-687) -> bField
template <class T, bool P2>
__device__ void body_les_term_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || !fluid(f, c.P)) return;
    const V3<T> v = les_term<T, P2>(p, f, f.vel[r.state], c, nbrs_of(p, f, c));
    f.b[0][c.P] = v.x; f.b[1][c.P] = v.y; f.b[2][c.P] = v.z;
}
template <class T>
__device__ void body_les_term(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (plain_v<T>) {
        UWS_P2_DISPATCH(body_les_term_t)
    } else {
        body_les_term_seq(p, f, c, r);
    }
}

// ---------------------------------------------------------------------------
// CSVelPredictPreCompute (FvmColl.compute:1291-1859), optionally fused with
// the LES term (which it would otherwise read back from bField).  Besides D
// and b it stores the sweep's face coefficients S, G of the owned E/N/U
// faces (zero unless both cells are fluid) and 1/D.
// ---------------------------------------------------------------------------
template <class T, bool P2>
__device__ void body_vel_precompute_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* un = f.vel[r.state];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const Nbr6 nb = nbrs_of(p, f, c);
    const V3<T> uP = ld3(un, P);
    // neighbour values loaded unconditionally: the normal velocity component,
    // p, and the boundary texture on each face
    T uNn[6], pQ[6];
    V3<T> ub[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        uNn[d] = un[d >> 1][nb.Q[d]];
        pQ[d] = f.p[nb.Q[d]];
        // boundary textures: read on the domain boundary only (a test of the
        // coordinates, so no other load waits on it)
        ub[d] = mk<T>(0, 0, 0);
        if (!nb.in[d] && p.velBnd[bface_of(d)] == UWS_VEL_FIXED_VALUE) ub[d] = bnd_at(p, f, d, x, y, z);
    }
    T vis[6];
    if (p.useLES && isF) face_vis(p, f, nb, P, vis);
    else {
#pragma unroll
        for (int d = 0; d < 6; ++d) vis[d] = p.nu;
    }
    {
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int d = dirs[a];
            const bool face = isF && nb.fl[d];
            f.S[a][P] = face ? comp(uP, a) + uNn[d] : (T)0;
            f.G[a][P] = face ? quo<P2>(vis[d] * p.ds, p.dx, p.rdx) : (T)0;
        }
    }
    if (!isF) {
        f.invD[0][P] = 0; f.invD[1][P] = 0; f.invD[2][P] = 0;
        return;
    }
    const T dv_dt = P2 ? p.dv_dt : p.dv / p.dt;
    T Dx = dv_dt, Dy = dv_dt, Dz = dv_dt;
    T bU = 0, bV = 0, bW = 0;
    const T ds = p.ds;
    // advection (:1314-1482)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const T uPn = comp(uP, cc);
        const int bf = bface_of(d);
        const int t = p.velBnd[bf];
        // fluid neighbour
        const T ti = ds * (T)0.25f * (uPn + uNn[d]);
        const bool addD_i = nb.fl[d];
        // domain boundary: fixed value feeds b, zero gradient D
        const T tz = ds * uPn;
        const bool addD_b = !nb.in[d] && t == UWS_VEL_ZERO_GRAD;
        const bool addb = !nb.in[d] && t == UWS_VEL_FIXED_VALUE;
        const T tt = addD_i ? ti : tz;
        const bool addD = addD_i || addD_b;
        const T Dx1 = plus ? Dx + tt : Dx - tt, Dy1 = plus ? Dy + tt : Dy - tt, Dz1 = plus ? Dz + tt : Dz - tt;
        Dx = addD ? Dx1 : Dx; Dy = addD ? Dy1 : Dy; Dz = addD ? Dz1 : Dz;
        const T ubn = comp(ub[d], cc);
        const T bU1 = plus ? bU - ds * ub[d].x * ubn : bU + ds * ub[d].x * ubn;
        const T bV1 = plus ? bV - ds * ub[d].y * ubn : bV + ds * ub[d].y * ubn;
        const T bW1 = plus ? bW - ds * ub[d].z * ubn : bW + ds * ub[d].z * ubn;
        bU = addb ? bU1 : bU; bV = addb ? bV1 : bV; bW = addb ? bW1 : bW;
    }
    // pressure face values (:1484-1591)
    const T pP = f.p[P];
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const T pb = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? (T)0 : pP;
        pf[d] = nb.in[d] ? (nb.fl[d] ? (T)0.5f * (pP + pQ[d]) : pP) : pb;
    }
    const T inv_den = P2 ? p.inv_den : (T)1.0f / p.den;
    bU += -inv_den * ds * (pf[E_] - pf[W_]);
    bV += -inv_den * ds * (pf[N_] - pf[S_]);
    bW += -inv_den * ds * (pf[U_] - pf[D_]);
    // diffusion (:1593-1836)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const int t = p.velBnd[bface_of(d)];
        const T tw = quo<P2>((T)2.0f * vis[d] * ds, p.dx, p.rdx);
        const T tin = nb.fl[d] ? quo<P2>(vis[d] * ds, p.dx, p.rdx) : tw;
        // in the grid: all three components; fixed value: all three (and b); symmetry: the normal one
        const bool all3 = nb.in[d] || t == UWS_VEL_FIXED_VALUE;
        const bool norm = !nb.in[d] && t == UWS_VEL_SYMMETRY;
        const T tt = nb.in[d] ? tin : tw;
        const bool fixedb = !nb.in[d] && t == UWS_VEL_FIXED_VALUE;
        const T bU1 = bU + tw * ub[d].x;
        const T bV1 = bV + tw * ub[d].y;
        const T bW1 = bW + tw * ub[d].z;
        bU = fixedb ? bU1 : bU; bV = fixedb ? bV1 : bV; bW = fixedb ? bW1 : bW;
        const bool ax = all3 || (norm && cc == 0), ay = all3 || (norm && cc == 1), az = all3 || (norm && cc == 2);
        Dx = ax ? Dx + tt : Dx;
        Dy = ay ? Dy + tt : Dy;
        Dz = az ? Dz + tt : Dz;
    }
    // time derivative, force, LES (:1838-1854)
    bU += dv_dt * uP.x; bV += dv_dt * uP.y; bW += dv_dt * uP.z;
    bU += p.dv * p.fext[0]; bV += p.dv * p.fext[1]; bW += p.dv * p.fext[2];
    if (p.useLES) {
        V3<T> bl;
        if (r.flags & K_FUSE_LES) bl = les_term<T, P2>(p, f, un, c, nb);
        else bl = mk<T>(f.b[0][P], f.b[1][P], f.b[2][P]);
        bU += bl.x; bV += bl.y; bW += bl.z;
    }
    f.D[0][P] = Dx; f.D[1][P] = Dy; f.D[2][P] = Dz;
    f.invD[0][P] = (T)1.0f / Dx; f.invD[1][P] = (T)1.0f / Dy; f.invD[2][P] = (T)1.0f / Dz;
    f.b[0][P] = bU; f.b[1][P] = bV; f.b[2][P] = bW;
}

template <class T>
__device__ void body_vel_precompute(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (plain_v<T>) {
        UWS_P2_DISPATCH(body_vel_precompute_t)
    } else {
        body_vel_precompute_seq(p, f, c, r);
    }
}

// ---------------------------------------------------------------------------
// CSVelPredict (FvmColl.compute:1862-2069): one Jacobi sweep.  With the face
// coefficients precomputed the accumulation order of the reference is kept:
// advection terms E W N S U D, then diffusion terms E W N S U D, then
// (1/D) * (adv + diff + b).  Non-fluid cells are not written.
// ---------------------------------------------------------------------------
// NU: stretched grid, where the convective coefficient of each axis is the
// stored face value times the cell's own size along that axis instead of ds/4
template <class T, bool NU = false>
__device__ void body_vel_sweep(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    const int sy = p.X, sz = p.X * p.Y;
    const int iE = c.x + 1 < p.X ? P + 1 : P, iW = c.x > 0 ? P - 1 : P;
    const int iN = c.y + 1 < p.Y ? P + sy : P, iS = c.y > 0 ? P - sy : P;
    const int iU = c.z + 1 < p.Z ? P + sz : P, iD = c.z > 0 ? P - sz : P;
    // a face on the domain boundary has no coefficient
    const T SE = f.S[0][P], SN = f.S[1][P], SU = f.S[2][P];
    const T GE = f.G[0][P], GN = f.G[1][P], GU = f.G[2][P];
    const T SW = c.x > 0 ? f.S[0][iW] : (T)0, GW = c.x > 0 ? f.G[0][iW] : (T)0;
    const T SS = c.y > 0 ? f.S[1][iS] : (T)0, GS = c.y > 0 ? f.G[1][iS] : (T)0;
    const T SD = c.z > 0 ? f.S[2][iD] : (T)0, GD = c.z > 0 ? f.G[2][iD] : (T)0;
    const T q = p.ds * (T)0.25f;
    T qx = q, qy = q, qz = q;
    if constexpr (NU) {
        qx = f.h[0][c.x];
        qy = f.h[1][c.y];
        qz = f.h[2][c.z];
    }
    T* const* kk = f.vel[r.vin];
    T* const* oo = f.vel[r.vout];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T* __restrict__ k = kk[a];
        const T vE = k[iE], vW = k[iW], vN = k[iN], vS = k[iS], vU = k[iU], vD = k[iD];
        T A = 0;
        A -= qx * vE * SE;
        A += qx * vW * SW;
        A -= qy * vN * SN;
        A += qy * vS * SS;
        A -= qz * vU * SU;
        A += qz * vD * SD;
        T Df = 0;
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
template <class T, bool P2>
__device__ void body_pres_precompute_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* u = f.vel[r.vout];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const T dvden = P2 ? p.dvden : p.dv / p.den;
    // every neighbour value is loaded unconditionally (indices fall back to the
    // cell itself, or to the first neighbour, outside the grid); the face terms
    // then pick their branch by selects, each branch computed as written
    int Q[6];
    bool in[6], fl[6];
    nbr6(p, f, c, Q, in, fl);
    const int st[6] = {1, -1, p.X, -p.X, p.X * p.Y, -p.X * p.Y};
    const int crd[6] = {x, x, y, y, z, z};
    const int lim[6] = {p.X, p.X, p.Y, p.Y, p.Z, p.Z};
    int F[6];
    bool farF[6];
    uint8_t ffl[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cf = crd[d] + ((d & 1) ? -2 : 2);
        const bool in2 = cf >= 0 && cf < lim[d];
        F[d] = in2 ? P + 2 * st[d] : Q[d];
        ffl[d] = f.flag[F[d]];
        farF[d] = in2 && ffl[d] == UWS_CELL_FLUID;
    }
    const T invP[3] = {f.invD[0][P], f.invD[1][P], f.invD[2][P]};
    const T uP[3] = {u[0][P], u[1][P], u[2][P]};
    const T pP = f.p[P];
    T invQ[6], uQ[6], pQ[6], pF[6], ubc[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        invQ[d] = f.invD[cc][Q[d]];
        uQ[d] = u[cc][Q[d]];
        pQ[d] = f.p[Q[d]];
        pF[d] = f.p[F[d]];
        const int bf = bface_of(d);
        const int k = bf < 2 ? z * p.Y + y : (bf < 4 ? z * p.X + x : y * p.X + x);
        ubc[d] = 0;
        if (!in[d] && p.velBnd[bf] == UWS_VEL_FIXED_VALUE) ubc[d] = f.bnd[bf][cc][k];
    }
    {
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int d = dirs[a];
            const bool face = isF && fl[d];
            const T DInv = (T)0.5f * (invP[a] + invQ[d]);
            f.cp[a][P] = face ? quo<P2>(dvden * DInv, p.dx, p.rdx) : (T)0;
            if (a == 0) f.cpisoE[P] = face ? dvden * quo<P2>(DInv, p.dx, p.rdx) : (T)0;
        }
    }
    if (!isF) {
        f.invDp[P] = 0;
        f.bp[P] = 0;
        return;
    }
    const T two_dx = (T)2.0f * p.dx, r2dx = p.r2dx;
    T Dp = 0, bp = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const T uPc = uP[cc];
        // interior face with a fluid neighbour
        const T DInv = (T)0.5f * (invP[cc] + invQ[d]);
        const T Dp_i = Dp - quo<P2>(dvden * DInv, p.dx, p.rdx);
        const T half = (T)0.5f * (uPc + uQ[d]);
        T bp_i = plus ? bp + half : bp - half;
        const bool fF = farF[d], bF = fl[d ^ 1];
        const T pq = pQ[d], pf = pF[d], pb = pQ[d ^ 1];
        T g, ga;
        if (plus) {
            g = quo<P2>(pq - pP, p.dx, p.rdx);
            const T a_fb = (T)0.5f * (quo<P2>(pf - pP, two_dx, r2dx) + quo<P2>(pq - pb, two_dx, r2dx));
            const T a_f = (T)0.5f * (quo<P2>(pf - pP, two_dx, r2dx) + quo<P2>(pq - pP, two_dx, r2dx));
            const T a_b = (T)0.5f * (quo<P2>(pq - pP, two_dx, r2dx) + quo<P2>(pq - pb, two_dx, r2dx));
            ga = fF ? (bF ? a_fb : a_f) : (bF ? a_b : g);
        } else {
            g = quo<P2>(pP - pq, p.dx, p.rdx);
            const T a_fb = (T)0.5f * (quo<P2>(pb - pq, two_dx, r2dx) + quo<P2>(pP - pf, two_dx, r2dx));
            const T a_f = (T)0.5f * (quo<P2>(pP - pq, two_dx, r2dx) + quo<P2>(pP - pf, two_dx, r2dx));
            const T a_b = (T)0.5f * (quo<P2>(pb - pq, two_dx, r2dx) + quo<P2>(pP - pq, two_dx, r2dx));
            ga = fF ? (bF ? a_fb : a_f) : (bF ? a_b : g);
        }
        const T vct = quo<P2>(-(p.dv * DInv), p.den, p.rden) * (g - ga);
        bp_i = plus ? bp_i + vct : bp_i - vct;
        // domain boundary face
        const int bf = bface_of(d);
        T bp_b = bp;
        if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) bp_b = plus ? bp + ubc[d] : bp - ubc[d];
        else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) bp_b = plus ? bp + uPc : bp - uPc;
        const T Dp_b = p.presBnd[bf] == UWS_PRES_FIXED_VALUE ? Dp - quo<P2>((T)2.0f * dvden * invP[cc], p.dx, p.rdx) : Dp;
        Dp = in[d] ? (fl[d] ? Dp_i : Dp) : Dp_b;
        bp = in[d] ? (fl[d] ? bp_i : bp) : bp_b;
    }
    f.Dp[P] = Dp;
    f.invDp[P] = (T)1.0f / Dp;
    f.bp[P] = bp;
}
template <class T>
__device__ void body_pres_precompute(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (plain_v<T>) {
        UWS_P2_DISPATCH(body_pres_precompute_t)
    } else {
        body_pres_precompute_seq(p, f, c, r);
    }
}

// pressure equation coefficient set: 0 = SIMPLE correction, 1 = PISO corrector
template <class T>
struct PresCoef {
    const T* cE;
    const T* cN;
    const T* cU;
    const T* invDp;
    const T* rhs;
};
template <class T>
__device__ __forceinline__ PresCoef<T> pres_coef(const FieldsT<T>& f, int set) {
    if (set == 0) return PresCoef<T>{f.cp[0], f.cp[1], f.cp[2], f.invDp, f.bp};
    return PresCoef<T>{f.cpisoE, f.cp[1], f.cp[2], f.invDpiso, f.bpiso};
}

// -sum_f c_f x_nb in the face order E W N S U D, x read through ld(i)
template <class T, class LD>
__device__ __forceinline__ T pres_term(const ParamsT<T>& p, const PresCoef<T>& k, const Cell& c, LD ld) {
    const int P = c.P, sy = p.X, sz = p.X * p.Y;
    const int iE = c.x + 1 < p.X ? P + 1 : P, iW = c.x > 0 ? P - 1 : P;
    const int iN = c.y + 1 < p.Y ? P + sy : P, iS = c.y > 0 ? P - sy : P;
    const int iU = c.z + 1 < p.Z ? P + sz : P, iD = c.z > 0 ? P - sz : P;
    const T cW = c.x > 0 ? k.cE[iW] : (T)0, cS = c.y > 0 ? k.cN[iS] : (T)0, cD = c.z > 0 ? k.cU[iD] : (T)0;
    T term = 0;
    term -= k.cE[P] * ld(iE);
    term -= cW * ld(iW);
    term -= k.cN[P] * ld(iN);
    term -= cS * ld(iS);
    term -= k.cU[P] * ld(iU);
    term -= cD * ld(iD);
    return term;
}

// ---------------------------------------------------------------------------
// CSPresCorrectRhieChow (FvmColl.compute:2789-2848) and CSPisoPresCorrection
// (PisoColl.compute:180-256): p_new = (1/diag) * (-sum_f c_f p_nb + rhs).
// pin < 0 means the iterate is zero (first sweep after
// SetPresCorrectFieldtoZero).  Non-fluid cells get 0.
// ---------------------------------------------------------------------------
template <class T>
__device__ void body_pres_sweep(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    T* __restrict__ out = f.pc[r.pout];
    if (!fluid(f, P)) {
        out[P] = 0;
        return;
    }
    const PresCoef<T> k = pres_coef(f, r.coef);
    T term = 0;
    if (r.pin >= 0) {
        const T* __restrict__ pk = f.pc[r.pin];
        term = pres_term(p, k, c, [&](int i) { return pk[i]; });
    }
    out[P] = k.invDp[P] * (term + k.rhs[P]);
}

// CSApplyPresCorrection (FvmColl.compute:2850-2862): p += p' on non-solid
// cells; when the normalisation is active (normRef >= 0) the cell (2, Y/2, 2)
// also records its new value for CSPresNormalization
// (FvmColl.compute:2864-2880), applied by the next kernel
template <class T>
__device__ void body_apply_pres(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const bool rec = c.P == p.normRef;
    // p, p' and the flag are loaded together; the flag picks what is stored
    const T pP = f.p[c.P], pcP = f.pc[r.pin][c.P];
    if (solid(f, c.P)) {
        if (rec) f.scalar[0] = pP;
        return;
    }
    const T v = pP + pcP;
    f.p[c.P] = v;
    if (rec) f.scalar[0] = v;
}
template <class T>
__device__ void body_pres_normalise(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || solid(f, c.P) || p.normRef < 0) return;
    f.p[c.P] -= f.scalar[0];
}

// CSApplyVelCorrectionDivergenceTheorem (FvmColl.compute:2950-3121)
template <class T, bool P2>
__device__ void body_vel_correct_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    // every load is issued before the cell's flag is tested (its indices
    // depend on the coordinates only); the flag decides what is stored
    const bool isF = fluid(f, P);
    const T* __restrict__ pc = f.pc[r.pin];
    int Qs[6];
    bool in[6], fl[6];
    nbr6(p, f, c, Qs, in, fl);
    const T pcP = pc[P];
    T pQ[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) pQ[d] = pc[Qs[d]];
    const T DP[3] = {f.D[0][P], f.D[1][P], f.D[2][P]};
    // with K_APPLY_P the launch also does CSApplyPresCorrection's p += p'
    // (the step fuses the two where no normalisation follows)
    const bool applyp = (r.flags & K_APPLY_P) != 0;
    const T pP = applyp ? f.p[P] : (T)0;
    const bool inplace = (r.flags & K_NO_INPLACE) == 0;
    T u0[3] = {0, 0, 0};
    if (inplace) {
        T* const* u = f.vel[r.vout];
        u0[0] = u[0][P]; u0[1] = u[1][P]; u0[2] = u[2][P];
    }
    if (!isF) return;
    if (applyp) f.p[P] = pP + pcP;
    const T iv[3] = {(T)1.0f / DP[0], (T)1.0f / DP[1], (T)1.0f / DP[2]};
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const T bnd = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? (T)0 : pcP;
        pf[d] = in[d] ? (fl[d] ? (T)0.5f * (pcP + pQ[d]) : pcP) : bnd;
    }
    const T dvden = P2 ? p.dvden : p.dv / p.den;
    const T cx = quo<P2>(-dvden * iv[0] * (pf[E_] - pf[W_]), p.dx, p.rdx);
    const T cy = quo<P2>(-dvden * iv[1] * (pf[N_] - pf[S_]), p.dx, p.rdx);
    const T cz = quo<P2>(-dvden * iv[2] * (pf[U_] - pf[D_]), p.dx, p.rdx);
    f.ucorr[0][P] = cx; f.ucorr[1][P] = cy; f.ucorr[2][P] = cz;
    if (!inplace) return;
    T* const* u = f.vel[r.vout];
    u[0][P] = u0[0] + cx; u[1][P] = u0[1] + cy; u[2][P] = u0[2] + cz;
}
template <class T>
__device__ void body_vel_correct(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    UWS_P2_DISPATCH(body_vel_correct_t)
}

// The one-shot kernels below read their six neighbours unconditionally:
// indices of missing or non-fluid neighbours fall back to the cell itself,
// and the terms of those faces are dropped by selecting the unchanged
// accumulator (bit-identical to skipping them).  Loads guarded by per-face
// branches would each wait for the previous ones.

// CSPisoCalAodUCorrect (PisoColl.compute:70-178): laminar nu, no BC terms
template <class T, bool P2>
__device__ void body_piso_aodu_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    // loads before the cell's flag test, which decides only the store
    const bool isF = fluid(f, P);
    T* const* un = f.vel[r.state];
    const V3<T> uP = ld3(un, P);
    T A[3] = {0, 0, 0}, Df[3] = {0, 0, 0};
    const T q = p.ds * (T)0.25f;
    int Qs[6];
    bool ok[6];
    fluid_nbrs(p, f, c, Qs, ok);
    T uN[6], uc[6][3];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        uN[d] = un[d >> 1][Qs[d]];
#pragma unroll
        for (int a = 0; a < 3; ++a) uc[d][a] = f.ucorr[a][Qs[d]];
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const T sum = comp(uP, cc) + uN[d];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T t = q * uc[d][a] * sum;
            const T v = (d & 1) == 0 ? A[a] + t : A[a] - t;
            A[a] = ok[d] ? v : A[a];
        }
    }
    const T g = P2 ? p.nuds_dx : p.nu * p.ds / p.dx;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T v = Df[a] - g * uc[d][a];
            Df[a] = ok[d] ? v : Df[a];
        }
    }
    if (!isF) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) f.aodu[a][P] = A[a] + Df[a];
}
template <class T>
__device__ void body_piso_aodu(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (plain_v<T>) {
        UWS_P2_DISPATCH(body_piso_aodu_t)
    } else {
        body_piso_aodu_seq(p, f, c, r);
    }
}

// CSPisoPresCorrection's constant parts (PisoColl.compute:196-255): presCoeff
// and velTerm do not change over the sweeps and are computed once.
template <class T>
__device__ void body_piso_precompute_plain(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    // loads before the cell's flag test, which decides only the stores
    const bool isF = fluid(f, P);
    int Qs[6];
    bool ok[6];
    fluid_nbrs(p, f, c, Qs, ok);
    const T invP[3] = {f.invD[0][P], f.invD[1][P], f.invD[2][P]};
    const T aoP[3] = {f.aodu[0][P], f.aodu[1][P], f.aodu[2][P]};
    T invQ[6], aoQ[6], cf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        invQ[d] = f.invD[cc][Qs[d]];
        aoQ[d] = f.aodu[cc][Qs[d]];
        // this face's coefficient as stored on the cell owning it
        const int owner = (d & 1) == 0 ? P : Qs[d];
        cf[d] = cc == 0 ? f.cpisoE[owner] : f.cp[cc][owner];
    }
    T presCoeff = 0, velTerm = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const T AdInv = (T)0.5f * (invP[cc] + invQ[d]);
        const T AodUCorrect = (T)0.5f * (aoP[cc] + aoQ[d]);
        const T pc2 = presCoeff - cf[d];
        const T vt2 = (d & 1) == 0 ? velTerm - AdInv * AodUCorrect : velTerm + AdInv * AodUCorrect;
        presCoeff = ok[d] ? pc2 : presCoeff;
        velTerm = ok[d] ? vt2 : velTerm;
    }
    if (!isF) {
        f.invDpiso[P] = 0;
        f.bpiso[P] = 0;
        return;
    }
    f.invDpiso[P] = (T)1.0f / presCoeff;
    f.bpiso[P] = velTerm;
}

template <class T>
__device__ void body_piso_precompute(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (plain_v<T>) body_piso_precompute_plain(p, f, c, r);
    else body_piso_precompute_seq(p, f, c, r);
}

// CSPisoApplyPresCorrection + CSPisoVelCorrection (PisoColl.compute:258-337)
template <class T, bool P2>
__device__ void body_piso_correct_t(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    // every load (the in-place ones where asked) is issued before the cell's
    // flag test, which decides only the stores
    const bool isF = fluid(f, P);
    const T* __restrict__ pc = f.pc[r.pin];
    int Qs[6];
    bool ok[6];
    fluid_nbrs(p, f, c, Qs, ok);
    const T pcP = pc[P];
    T pQ[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) pQ[d] = pc[Qs[d]];
    const T ao[3] = {f.aodu[0][P], f.aodu[1][P], f.aodu[2][P]};
    const T iv[3] = {f.invD[0][P], f.invD[1][P], f.invD[2][P]};
    const bool inplace = (r.flags & K_NO_INPLACE) == 0;
    T pOld = 0, u0[3] = {0, 0, 0};
    if (inplace) {
        T* const* u = f.vel[r.vout];
        pOld = f.p[P];
        u0[0] = u[0][P]; u0[1] = u[1][P]; u0[2] = u[2][P];
    }
    if (!isF) return;
    if (inplace) f.p[P] = pOld + pcP;
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) pf[d] = ok[d] ? (T)0.5f * (pcP + pQ[d]) : pcP;
    const T gx = quo<P2>(pf[E_] - pf[W_], p.dx, p.rdx), gy = quo<P2>(pf[N_] - pf[S_], p.dx, p.rdx);
    const T gz = quo<P2>(pf[U_] - pf[D_], p.dx, p.rdx);
    const T dvden = P2 ? p.dvden : p.dv / p.den;
    const T cx = -iv[0] * (ao[0] + dvden * gx);
    const T cy = -iv[1] * (ao[1] + dvden * gy);
    const T cz = -iv[2] * (ao[2] + dvden * gz);
    f.ucorr[0][P] = cx; f.ucorr[1][P] = cy; f.ucorr[2][P] = cz;
    if (!inplace) return;
    T* const* u = f.vel[r.vout];
    u[0][P] = u0[0] + cx; u[1][P] = u0[1] + cy; u[2][P] = u0[2] + cz;
}
template <class T>
__device__ void body_piso_correct(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    UWS_P2_DISPATCH(body_piso_correct_t)
}

// CSInitModelFlag's velocity reset (InitShaderColl.compute:353-359) applied
// to every velocity buffer after a flag upload
template <class T>
__device__ void body_zero_solid_vel(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || !solid(f, c.P)) return;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        f.vel[b][0][c.P] = 0; f.vel[b][1][c.P] = 0; f.vel[b][2][c.P] = 0;
    }
}
// a logically-zero p' that must be read: fill pc[pout] with zeros
template <class T>
__device__ void body_fill_pc(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (c.ok) f.pc[r.pout][c.P] = 0;
}

// ---------------------------------------------------------------------------
// Defect correction with 2-byte correction storage.  With x the FP32 iterate
// of the pressure equation x = invD (term(x) + rhs), the defect is
// rho = invD (term(x) + rhs) - x (x = 0 when pin < 0) and the correction e
// solves e = invD term(e) + rho; Jacobi on e from e = 0 continues plain
// Jacobi on x exactly in exact arithmetic.  e is stored scaled by
// s = 1/rms(rho) in fp16 or bf16; all arithmetic is in T.
// ---------------------------------------------------------------------------
template <bool BF>
__device__ __forceinline__ float ld_lp(const void* a, int i) {
    if constexpr (BF) return __bfloat162float(static_cast<const __hip_bfloat16*>(a)[i]);
    else return (float)static_cast<const _Float16*>(a)[i];
}
template <bool BF>
__device__ __forceinline__ void st_lp(void* a, int i, float v) {
    if constexpr (BF) static_cast<__hip_bfloat16*>(a)[i] = __float2bfloat16(v);
    else static_cast<_Float16*>(a)[i] = (_Float16)v;
}

template <class T>
__device__ double body_defect_residual(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return 0.0;
    const int P = c.P;
    if (!fluid(f, P)) {
        f.rho[P] = 0;
        return 0.0;
    }
    const PresCoef<T> k = pres_coef(f, r.coef);
    T term = 0, x = 0;
    if (r.pin >= 0) {
        const T* __restrict__ pk = f.pc[r.pin];
        term = pres_term(p, k, c, [&](int i) { return pk[i]; });
        x = pk[P];
    }
    const T rh = k.invDp[P] * (term + k.rhs[P]) - x;
    f.rho[P] = rh;
    return (double)rh * (double)rh;
}

template <class T, bool BF>
__device__ void body_lowp_sweep(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    const bool last = (r.flags & K_LOWP_LAST) != 0;
    if (!fluid(f, P)) {
        if (last) f.pc[r.pout][P] = 0;
        else st_lp<BF>(f.e16[r.eout], P, 0.f);
        return;
    }
    const PresCoef<T> k = pres_coef(f, r.coef);
    const T s = f.scalar[1];
    T term = 0;
    if (!(r.flags & K_LOWP_FIRST)) {
        const void* ein = f.e16[r.ein];
        term = pres_term(p, k, c, [&](int i) { return (T)ld_lp<BF>(ein, i); });
    }
    const T e = k.invDp[P] * term + s * f.rho[P];
    if (last) {
        const T x = r.pin >= 0 ? f.pc[r.pin][P] : (T)0;
        f.pc[r.pout][P] = x + e / s;
    } else {
        st_lp<BF>(f.e16[r.eout], P, (float)e);
    }
}

// ---------------------------------------------------------------------------
// FP64 reductions: each workgroup writes one partial (of its logical tile),
// a second pass sums a case's partials in a fixed order, so every norm is
// deterministic.
// ---------------------------------------------------------------------------
// CSComputeVelPredictResidual (FvmColl.compute:3123-3162): sum (D (u_out - u_in))^2
template <class T>
__device__ double body_res_vel(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T v = f.D[a][c.P] * (f.vel[r.vout][a][c.P] - f.vel[r.vin][a][c.P]);
        s += (double)v * (double)v;
    }
    return s;
}
// CSComputePresCorrectResidual (FvmColl.compute:3164-3199): sum (Dp (p'_out - p'_in))^2
template <class T>
__device__ double body_res_pres(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return 0.0;
    const T a = f.pc[r.pout][c.P] - (r.pin >= 0 ? f.pc[r.pin][c.P] : (T)0);
    const T v = f.Dp[c.P] * a;
    return (double)v * (double)v;
}
// sum of squared Rhie-Chow net outflow (bp after a pressure precompute)
template <class T>
__device__ double body_divergence(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || !fluid(f, c.P)) return 0.0;
    const double v = (double)f.bp[c.P];
    return v * v;
}
// count of cells whose state (u, p) is not finite
template <class T>
__device__ double body_nonfinite(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return 0.0;
    const T a = f.vel[r.state][0][c.P], b = f.vel[r.state][1][c.P], d = f.vel[r.state][2][c.P], q = f.p[c.P];
    return (isfinite(a) && isfinite(b) && isfinite(d) && isfinite(q)) ? 0.0 : 1.0;
}

#include "uws_nonuniform.h"

template <class T, int K>
__device__ __forceinline__ void run_body(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if constexpr (K == KN_WALL_EDDY) body_wall_eddy(p, f, c, r);
    else if constexpr (K == KN_LES_TERM) body_les_term(p, f, c, r);
    else if constexpr (K == KN_VEL_PRECOMP) body_vel_precompute(p, f, c, r);
    else if constexpr (K == KN_VEL_SWEEP) body_vel_sweep(p, f, c, r);
    else if constexpr (K == KN_PRES_PRECOMP) body_pres_precompute(p, f, c, r);
    else if constexpr (K == KN_PRES_SWEEP) body_pres_sweep(p, f, c, r);
    else if constexpr (K == KN_APPLY_PRES) body_apply_pres(p, f, c, r);
    else if constexpr (K == KN_PRES_NORMALISE) body_pres_normalise(p, f, c, r);
    else if constexpr (K == KN_VEL_CORRECT) body_vel_correct(p, f, c, r);
    else if constexpr (K == KN_PISO_AODU) body_piso_aodu(p, f, c, r);
    else if constexpr (K == KN_PISO_PRECOMP) body_piso_precompute(p, f, c, r);
    else if constexpr (K == KN_PISO_CORRECT) body_piso_correct(p, f, c, r);
    else if constexpr (K == KN_ZERO_SOLID_VEL) body_zero_solid_vel(p, f, c, r);
    else if constexpr (K == KN_FILL_PC) body_fill_pc(p, f, c, r);
    else if constexpr (K == KN_NU_VEL_PRECOMP) body_nu_vel_precompute(p, f, c, r);
    else if constexpr (K == KN_NU_VEL_SWEEP) body_vel_sweep<T, true>(p, f, c, r);
    else if constexpr (K == KN_NU_PRES_PRECOMP) body_nu_pres_precompute(p, f, c, r);
    else if constexpr (K == KN_NU_VEL_CORRECT) body_nu_vel_correct(p, f, c, r);
    else if constexpr (K == KN_LOWP_SWEEP) {
        if constexpr (sizeof(T) == 4) {
            if (r.flags & K_BF16) body_lowp_sweep<T, true>(p, f, c, r);
            else body_lowp_sweep<T, false>(p, f, c, r);
        }
    }
}
template <class T, int K>
__device__ __forceinline__ double run_reduce(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c,
                                             const Roles& r) {
    if constexpr (K == KN_RES_VEL) return body_res_vel(p, f, c, r);
    else if constexpr (K == KN_RES_PRES) return body_res_pres(p, f, c, r);
    else if constexpr (K == KN_DIVERGENCE) return body_divergence(p, f, c, r);
    else if constexpr (K == KN_NONFINITE) return body_nonfinite(p, f, c, r);
    else if constexpr (K == KN_DEFECT_RESIDUAL) return body_defect_residual(p, f, c, r);
    else if constexpr (K == KN_NU_DIVERGENCE) return body_nu_divergence(p, f, c, r);
    else return 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// block sum over blockDim.y waves in a fixed order, result valid in every thread
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ws[4];
    v = wave_sum(v);
    if (threadIdx.x == 0) ws[threadIdx.y] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)blockDim.y; ++i) s += ws[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Two pressure Jacobi sweeps in one launch (temporal blocking).  A workgroup
// owns a 64 x (PS2_BY - 2) x zchunk block of output cells and marches along
// z: at plane zz every thread of the 66 x PS2_BY tile (the block plus a
// one-cell halo in x and y) computes the first sweep's value x1, which goes
// to LDS (three rotating planes, one barrier per plane), and the block's
// threads compute the second sweep at plane zz - 1 from x1 of the planes
// zz - 2 .. zz.  The halo's x1 is recomputed by each neighbouring block, so
// the coefficients are read once per two sweeps.  Both sweeps use the
// expressions and operation order of body_pres_sweep, so the result equals
// two single sweeps bit for bit.
// ---------------------------------------------------------------------------
// Thread layout of a z-marching tile W = 64 + 2H cells wide and BY rows
// tall: the first 64 BY threads cover the tile's 64 interior columns, one
// wave per row (so every load and store of a wave is one aligned 256-byte
// row segment of the grid), the other 2 H BY threads the H halo columns on
// each side, row by row.  (lx, ly): the thread's cell in the tile.
__device__ __forceinline__ void tile_lane(int t, int H, int BY, int& lx, int& ly) {
    const int nmain = 64 * BY;
    if (t < nmain) {
        lx = H + (t & 63);
        ly = t >> 6;
    } else {
        const int h = t - nmain;
        ly = h / (2 * H);
        const int s = h - ly * 2 * H;
        lx = s < H ? s : 64 + s;
    }
}

constexpr int PS2_W = 66;                   // threads per tile row: 64 outputs + 2 halo
constexpr int PS2_BY = 15;                  // most tile rows (13 outputs + 2 halo); a grid uses by2 <= PS2_BY
constexpr int PS2_THREADS = PS2_W * PS2_BY;  // 990: the LDS stride of a plane

// one plane's inputs of the first sweep at a cell: its coefficients and the
// previous iterate at its six neighbours (clamped at the domain boundary,
// where the coefficient is zero)
// the raw loads of one plane at a thread's column (masked when used)
template <class T>
struct PS2Plane {
    T cE, cW, cN, cS, cU, inv, rhs;
    T vE, vW, vN, vS, vU;
};

// by: tile rows (outputs + 2 halo rows) of the grid's two-sweep launches
inline __host__ __device__ int pres_sweep2_blocks(int X, int Y, int Z, int zchunk, int by) {
    return ((X + 63) / 64) * ((Y + by - 3) / (by - 2)) * ((Z + zchunk - 1) / zchunk);
}

// the two-sweep block b of one grid (b in [0, pres_sweep2_blocks))
//
// Latency: the loads of plane zz + 1 are issued before the first sweep of
// plane zz and consumed one iteration later.  They are unconditional (indices
// clamped into the grid, values masked where they are used), and the three
// plane records rotate through three named variables (the loop is unrolled
// three times), so no branch and no register copy makes an iteration wait
// for the loads still in flight.  The D-face values come from the column's
// own records: the D coefficient of plane zz is the U coefficient of plane
// zz - 1, and the iterate below it is the U-neighbour value loaded with
// plane zz - 2.  No flag is read: a non-fluid cell has 1/Dp = bp = 0 and zero
// coefficients on all its faces, so both sweeps give it 0 (+0 or -0; zeros
// are only ever multiplied by the zero coefficients of its neighbours), and
// values of cells outside the grid are never used by an output cell.
// what one two-sweep workgroup reads: the grid, the coefficient set and the
// p' buffers (a compact copy, so that an ensemble launch keeps no more
// registers than a single-case one)
// *(base + off) for a byte offset off (uniform base: the SGPR-base form of
// the global load, one 32-bit offset register shared by arrays)
// Every array is device (global) memory: the accesses of scalar values say
// so, so that they compile to global loads and stores also where the array
// pointer itself was read from memory (an ensemble's per-case tables), not
// only where it came in as a kernel argument; a generic (flat) access would
// need 64-bit VGPR addresses and wait on the LDS counter too.  The two-double
// types of the complex-step and dual-number runs keep generic accesses.
#define UWS_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t off) {
    if constexpr (std::is_scalar_v<T>)
        return *(const UWS_GLOBAL T*)((const UWS_GLOBAL char*)base + off);
    else
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void st_at(T* base, uint32_t off, T v) {
    if constexpr (std::is_scalar_v<T>)
        *(UWS_GLOBAL T*)((UWS_GLOBAL char*)base + off) = v;
    else
        *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}
template <class T>
struct PS2Args {
    int X, Y, Z, zchunk, by;
    PresCoef<T> k;
    const T* x0;
    T* out;
    // the low-precision correction equation (three-sweep kernel only)
    const void* ein;   // e in 2-byte storage
    void* eout;
    const T* xin;      // p' the round corrects (nullptr: zero)
    T s;               // scale of e
    T* rec1;           // where non-null (LP = 0): the first and second sweeps' values are stored too
    T* rec2;
};
template <class T>
__device__ __forceinline__ PS2Args<T> ps2_args(const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r) {
    PS2Args<T> a;
    a.X = p.X; a.Y = p.Y; a.Z = p.Z; a.zchunk = p.zchunk; a.by = p.by2;
    a.k = pres_coef(f, r.coef);
    a.x0 = r.pin >= 0 ? f.pc[r.pin] : nullptr;
    a.out = f.pc[r.pout];
    a.rec1 = f.pcrec[0];
    a.rec2 = f.pcrec[1];
    return a;
}

// OFF: byte-offset addressing and unconditional LDS reads (fewer VALU, more
// VGPRs; the ensemble kernel keeps element indices and selected LDS reads,
// which stay within 64 VGPRs).  32-bit byte offsets hold N sizeof(T) for
// types of up to 8 bytes (N < 2^29 cells); wider types index.
template <class T, bool ZIN, bool OFF = (sizeof(T) <= 8)>
__device__ __forceinline__ void pres_sweep2_block(const PS2Args<T>& p, int b, T* lds) {
    int lx, ly;
    tile_lane((int)threadIdx.x, 1, p.by, lx, ly);
    const int tid = ly * PS2_W + lx;  // the thread's entry in the LDS planes
    const int ntx = (p.X + 63) / 64, nty = (p.Y + p.by - 3) / (p.by - 2);
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int x = bx * 64 - 1 + lx, y = by * (p.by - 2) - 1 + ly;
    const int z0 = bz * p.zchunk, z1 = min(z0 + p.zchunk, p.Z);
    const bool cell = x >= 0 && x < p.X && y >= 0 && y < p.Y;
    const bool inner = cell && lx >= 1 && lx <= 64 && ly >= 1 && ly <= p.by - 2;
    const PresCoef<T> k = p.k;
    const T* __restrict__ x0 = ZIN ? nullptr : p.x0;
    T* __restrict__ out = p.out;
    const int sy = p.X, sz = p.X * p.Y;
    const bool hW = x > 0, hE = x + 1 < p.X, hS = y > 0, hN = y + 1 < p.Y;
    const int xc = min(max(x, 0), p.X - 1), yc = min(max(y, 0), p.Y - 1);
    // unconditional loads of plane zz (clamped into the grid) into q; the U
    // neighbour is read at plane zz + 1 clamped on its own, so that the record
    // of plane zz - 2 always holds the D neighbour of plane zz (at the top
    // face it is the cell itself, multiplied by a zero coefficient)
    // neighbour flags of the clamped column: the loads of threads outside the
    // grid stay inside it (their values are never used)
    const bool gE = xc + 1 < p.X, gW = xc > 0, gN = yc + 1 < p.Y, gS = yc > 0;
    // addresses: uniform base + 32-bit byte offset (the column's offset plus
    // a uniform plane offset); the W and S coefficients are read at the
    // clamped W/S neighbour (the cell itself on the boundary, scaled by 0)
    const uint32_t colB = (uint32_t)(yc * p.X + xc) * (uint32_t)sizeof(T);
    const uint32_t planeB = (uint32_t)(p.X * p.Y) * (uint32_t)sizeof(T);
    const uint32_t rowB = (uint32_t)(sy * sizeof(T));
    auto load = [&](PS2Plane<T>& q, int zz) {
        if constexpr (OFF) {
            const uint32_t oP = (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB;
            const uint32_t oU = (uint32_t)min(max(zz + 1, 0), p.Z - 1) * planeB + colB;
            const uint32_t oW = gW ? oP - (uint32_t)sizeof(T) : oP, oS = gS ? oP - rowB : oP;
            q.cE = ld_at(k.cE, oP); q.cN = ld_at(k.cN, oP); q.cU = ld_at(k.cU, oP);
            q.cW = ld_at(k.cE, oW);
            q.cS = ld_at(k.cN, oS);
            q.inv = ld_at(k.invDp, oP);
            q.rhs = ld_at(k.rhs, oP);
            if (!ZIN) {
                q.vE = ld_at(x0, gE ? oP + (uint32_t)sizeof(T) : oP);
                q.vW = ld_at(x0, oW);
                q.vN = ld_at(x0, gN ? oP + rowB : oP);
                q.vS = ld_at(x0, oS);
                q.vU = ld_at(x0, oU);
            }
        } else {
            const int zc = min(max(zz, 0), p.Z - 1);
            const int P = (zc * p.Y + yc) * p.X + xc;
            const int PU = (min(max(zz + 1, 0), p.Z - 1) * p.Y + yc) * p.X + xc;
            q.cE = k.cE[P]; q.cN = k.cN[P]; q.cU = k.cU[P];
            q.cW = k.cE[max(P - 1, 0)];
            q.cS = k.cN[max(P - sy, 0)];
            q.inv = k.invDp[P];
            q.rhs = k.rhs[P];
            if (!ZIN) {
                q.vE = x0[gE ? P + 1 : P];
                q.vW = x0[gW ? P - 1 : P];
                q.vN = x0[gN ? P + sy : P];
                q.vS = x0[gS ? P - sy : P];
                q.vU = x0[PU];
            }
        }
    };
    // a face on the domain boundary has no coefficient (the loaded neighbour
    // value is scaled by 0; coefficients are >= +0, so this is the +0 of a select)
    const T mW = hW ? (T)1 : (T)0, mS = hS ? (T)1 : (T)0;
    // own-column first-sweep values of the two previous planes
    T x1m1 = 0, x1m2 = 0;
    // one plane: loads of zz + 1 into nx, first sweep of zz from cu, second
    // sweep of zz - 1 from pv (its coefficients) and the LDS ring
    T pdU = 0;  // U coefficient of the plane below the previous one (the previous plane's D coefficient)
    auto step = [&](const PS2Plane<T>& pv, const PS2Plane<T>& cu, PS2Plane<T>& nx, int zz) {
        // nx still holds plane zz - 2, whose U neighbour is this column at zz - 1
        // (at the bottom face a finite value times the zero D coefficient)
        T vD = 0;
        if (!ZIN) vD = nx.vU;
        load(nx, zz + 1 <= z1 ? zz + 1 : z1);
        const T mD = zz > 0 ? (T)1 : (T)0;
        T term = 0;
        if (!ZIN) {
            term -= cu.cE * cu.vE;
            term -= (cu.cW * mW) * cu.vW;
            term -= cu.cN * cu.vN;
            term -= (cu.cS * mS) * cu.vS;
            term -= cu.cU * cu.vU;
            term -= (pv.cU * mD) * vD;
        }
        const T x1 = cu.inv * (term + cu.rhs);
        lds[((zz + 3) % 3) * PS2_THREADS + tid] = x1;
        __syncthreads();
        const int zo = zz - 1;
        if (inner && zo >= z0 && zo < z1) {
            const int P = (zo * p.Y + y) * p.X + x;
            // on the domain boundary the neighbour is the cell itself (its
            // own first-sweep value, x1m1, times a zero coefficient)
            const T* L = lds + ((zo + 3) % 3) * PS2_THREADS;
            T vE, vW, vN, vS;
            if constexpr (OFF) {
                vE = L[hE ? tid + 1 : tid]; vW = L[hW ? tid - 1 : tid];
                vN = L[hN ? tid + PS2_W : tid]; vS = L[hS ? tid - PS2_W : tid];
            } else {
                const T lE = L[tid + 1], lW = L[tid - 1], lN = L[tid + PS2_W], lS = L[tid - PS2_W];
                vE = hE ? lE : x1m1; vW = hW ? lW : x1m1;
                vN = hN ? lN : x1m1; vS = hS ? lS : x1m1;
            }
            const T vU = zo + 1 < p.Z ? x1 : x1m1, vD = zo > 0 ? x1m2 : x1m1;
            const T mDo = zo > 0 ? (T)1 : (T)0;
            T t2 = 0;
            t2 -= pv.cE * vE;
            t2 -= (pv.cW * mW) * vW;
            t2 -= pv.cN * vN;
            t2 -= (pv.cS * mS) * vS;
            t2 -= pv.cU * vU;
            t2 -= (pdU * mDo) * vD;
            out[P] = pv.inv * (t2 + pv.rhs);
        }
        x1m2 = x1m1;
        x1m1 = x1;
        pdU = pv.cU;
    };
    // records on entry of step(pv, cu, nx, zz): pv = plane zz - 1, cu = plane
    // zz, nx = plane zz - 2; so A = plane z0 - 2 (its U coefficient is the D
    // coefficient of plane z0 - 1), B = plane z0 - 1, C = plane z0 - 3 (its U
    // neighbour is the iterate below z0 - 1)
    PS2Plane<T> A, B, C;
    load(A, z0 - 2);
    load(B, z0 - 1);
    load(C, z0 - 3);
    for (int zz = z0 - 1;;) {
        step(A, B, C, zz);
        if (++zz > z1) break;
        step(B, C, A, zz);
        if (++zz > z1) break;
        step(C, A, B, zz);
        if (++zz > z1) break;
    }
}

template <class T, bool ZIN>
__global__ __launch_bounds__(1024) void k_pres_sweep2(ParamsT<T> p, FieldsT<T> f, Roles r) {
    __shared__ __align__(16) unsigned char lds_raw[3 * PS2_THREADS * sizeof(T)];
    const int nb = pres_sweep2_blocks(p.X, p.Y, p.Z, p.zchunk, p.by2);
    pres_sweep2_block<T, ZIN>(ps2_args(p, f, r), xcd_tile((int)blockIdx.x, nb), reinterpret_cast<T*>(lds_raw));
}
// FP32: held to 64 VGPRs, so that 8 waves run per SIMD and two 16-wave
// workgroups share a CU (the FP32 body fits without spilling)
template <bool ZIN>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_pres_sweep2_f32(ParamsT<float> p,
                                                                                                  FieldsT<float> f,
                                                                                                  Roles r) {
    __shared__ __align__(16) float lds_f[3 * PS2_THREADS];
    const int nb = pres_sweep2_blocks(p.X, p.Y, p.Z, p.zchunk, p.by2);
    pres_sweep2_block<float, ZIN>(ps2_args(p, f, r), xcd_tile((int)blockIdx.x, nb), lds_f);
}


// ---------------------------------------------------------------------------
// Three pressure Jacobi sweeps in one launch: the z-march of the two-sweep
// kernel one level deeper.  A workgroup of 68 x 15 threads owns a 64 x 11
// column of cells (two halo cells on each side) and marches z from two planes
// below its chunk to two above.  At plane zz every thread computes the first
// sweep at zz (s1, into LDS ring 1), the second sweep at zz - 1 (s2, from
// ring 1's plane zz - 1, written one iteration earlier, into ring 2), and the
// interior threads the third sweep at zz - 2 (from ring 2's plane zz - 2)
// into the output; one barrier per plane.  Own-column values above and below
// come from registers.  Four plane records (coefficients and the input
// iterate's neighbours) rotate through named variables (the loop is unrolled
// four times, the period of the four-slot LDS rings, so LDS addresses are
// constant offsets); the loads of plane zz + 1 are issued first.  A record's
// boundary masks are applied once, before its first use.  Expressions and
// the handling of non-fluid cells are those of the two-sweep kernel, so the
// result equals three single sweeps.  The barrier also keeps the 16 waves on
// one plane, so that their neighbours' loads hit the caches.  A thread
// outside the second sweep's region (the outer halo ring) computes it too, on
// its own entry, and its value is never used.
// ---------------------------------------------------------------------------
constexpr int PS3_W = 68;                    // threads per tile row: 64 outputs + 2 x 2 halo
constexpr int PS3_BY = 15;                   // tile rows: 11 outputs + 2 x 2 halo
constexpr int PS3_THREADS = PS3_W * PS3_BY;  // 1020
constexpr int PS3_RING = 4;                  // planes per LDS ring (two rings)

inline __host__ __device__ int pres_sweep3_blocks(int X, int Y, int Z, int zchunk) {
    return ((X + 63) / 64) * ((Y + PS3_BY - 5) / (PS3_BY - 4)) * ((Z + zchunk - 1) / zchunk);
}

// LP = 0: the pressure equation x = 1/Dp (term(x) + rhs), FP32 in and out.
// LP = 1 (FP16), 2 (BF16): the low-precision correction equation of the
// defect correction, e = 1/Dp term(e) + s rho with e stored in 2 bytes
// (body_lowp_sweep): the first and second sweeps' values are rounded to 2
// bytes as the stored iterate would be, and the third goes to e_out or, when
// it is the round's last sweep (LAST), to p' = x + e / s.
template <class T, bool ZIN, int LP = 0, bool LAST = false>
__device__ __forceinline__ void pres_sweep3_block(const PS2Args<T>& p, int b, T* lds) {
    const int tid = (int)threadIdx.x;
    const int lx = tid % PS3_W, ly = tid / PS3_W;
    const int ntx = (p.X + 63) / 64, nty = (p.Y + PS3_BY - 5) / (PS3_BY - 4);
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int x = bx * 64 - 2 + lx, y = by * (PS3_BY - 4) - 2 + ly;
    const int z0 = bz * p.zchunk, z1 = min(z0 + p.zchunk, p.Z);
    const bool cell = x >= 0 && x < p.X && y >= 0 && y < p.Y;
    const bool inner = cell && lx >= 2 && lx <= PS3_W - 3 && ly >= 2 && ly <= PS3_BY - 3;
    const PresCoef<T> k = p.k;
    const T* __restrict__ x0 = ZIN ? nullptr : p.x0;
    T* __restrict__ out = p.out;
    const bool hW = x > 0, hE = x + 1 < p.X, hS = y > 0, hN = y + 1 < p.Y;
    const int xc = min(max(x, 0), p.X - 1), yc = min(max(y, 0), p.Y - 1);
    const bool gE = xc + 1 < p.X, gW = xc > 0, gN = yc + 1 < p.Y, gS = yc > 0;
    const uint32_t colB = (uint32_t)(yc * p.X + xc) * (uint32_t)sizeof(T);
    const uint32_t planeB = (uint32_t)(p.X * p.Y) * (uint32_t)sizeof(T);
    const uint32_t rowB = (uint32_t)(p.X * sizeof(T));
    // the record of plane zz: its coefficients and the input iterate at its
    // E, W, N, S neighbours and at the plane above (clamped into the grid)
    auto load = [&](PS2Plane<T>& q, int zz) {
        const uint32_t oP = (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB;
        const uint32_t oU = (uint32_t)min(max(zz + 1, 0), p.Z - 1) * planeB + colB;
        const uint32_t oW = gW ? oP - (uint32_t)sizeof(T) : oP, oS = gS ? oP - rowB : oP;
        q.cE = ld_at(k.cE, oP); q.cN = ld_at(k.cN, oP); q.cU = ld_at(k.cU, oP);
        q.cW = ld_at(k.cE, oW);
        q.cS = ld_at(k.cN, oS);
        q.inv = ld_at(k.invDp, oP);
        q.rhs = ld_at(k.rhs, oP);
        if (!ZIN) {
            if constexpr (LP == 0) {
                q.vE = ld_at(x0, gE ? oP + (uint32_t)sizeof(T) : oP);
                q.vW = ld_at(x0, oW);
                q.vN = ld_at(x0, gN ? oP + rowB : oP);
                q.vS = ld_at(x0, oS);
                q.vU = ld_at(x0, oU);
            } else {
                const uint32_t es = (uint32_t)sizeof(T);
                q.vE = ld_lp<LP == 2>(p.ein, (int)((gE ? oP + es : oP) / es));
                q.vW = ld_lp<LP == 2>(p.ein, (int)(oW / es));
                q.vN = ld_lp<LP == 2>(p.ein, (int)((gN ? oP + rowB : oP) / es));
                q.vS = ld_lp<LP == 2>(p.ein, (int)(oS / es));
                q.vU = ld_lp<LP == 2>(p.ein, (int)(oU / es));
            }
        }
    };
    // the two records below the first computed plane are read only in part:
    // plane z0 - 4 for the input iterate above it (the D value of plane
    // z0 - 2), plane z0 - 3 for that and its U coefficient (the D coefficient
    // of plane z0 - 2); loading nothing else saves their other planes' traffic
    auto load_below = [&](PS2Plane<T>& q, int zz, bool with_cU) {
        const uint32_t oP = (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB;
        const uint32_t oU = (uint32_t)min(max(zz + 1, 0), p.Z - 1) * planeB + colB;
        q.cE = q.cW = q.cN = q.cS = q.inv = q.rhs = 0;
        q.vE = q.vW = q.vN = q.vS = 0;
        q.cU = with_cU ? ld_at(k.cU, oP) : (T)0;
        q.vU = 0;
        if (!ZIN) {
            if constexpr (LP == 0) q.vU = ld_at(x0, oU);
            else q.vU = ld_lp<LP == 2>(p.ein, (int)(oU / (uint32_t)sizeof(T)));
        }
    };
    const T sc = LP == 0 ? (T)0 : p.s;
    // one sweep's value from its term, and the value the next sweep reads
    auto comb = [&](T inv, T t, T rhs) {
        if constexpr (LP == 0) return inv * (t + rhs);
        else return inv * t + sc * rhs;
    };
    auto rnd = [&](T v) {
        if constexpr (LP == 0) return v;
        else if constexpr (LP == 2) return (T)__bfloat162float(__float2bfloat16((float)v));
        else return (T)(float)(_Float16)(float)v;
    };
    // LDS index of each in-plane neighbour: the thread's own entry on the
    // domain boundary and on the tile's outer edge
    const int jE = (hE && lx + 1 < PS3_W) ? tid + 1 : tid, jW = (hW && lx > 0) ? tid - 1 : tid;
    const int jN = (hN && ly + 1 < PS3_BY) ? tid + PS3_W : tid, jS = (hS && ly > 0) ? tid - PS3_W : tid;
    // two rings of PS3_RING planes: first- and second-sweep values.  Plane
    // z0 - 2 + 4m + k sits in slot k, so unrolled step k addresses its slots
    // with constant offsets.
    T* const L1 = lds;
    T* const L2 = lds + PS3_RING * PS3_THREADS;
    // A record's boundary masks, applied once before its first use: no W or S
    // neighbour on the domain boundary, and no D neighbour above a plane below
    // the grid (a plane's U coefficient is the D coefficient of the plane
    // above).  A zeroed term adds -(+0 * v), where the two-sweep kernel adds
    // -((c * 0) * v): both are zeros and leave the sum unchanged (it starts at
    // +0 and never becomes -0).
    auto mask = [&](PS2Plane<T>& q, int zz) {
        q.cW = hW ? q.cW : (T)0;
        q.cS = hS ? q.cS : (T)0;
        if (zz < 0) q.cU = 0;
    };
    T s1m1 = 0, s1m2 = 0;  // own first-sweep values at zz - 1, zz - 2
    T s2m1 = 0, s2m2 = 0;  // own second-sweep values at zz - 2, zz - 3
    // iteration zz: r0 = plane zz, r1 = zz - 1, r2 = zz - 2, nx = zz - 3
    // (receives zz + 1); kc: the step's position in the unrolled loop
    auto step = [&](auto kc, PS2Plane<T>& r2, PS2Plane<T>& r1, PS2Plane<T>& r0, PS2Plane<T>& nx, int zz) {
        constexpr int K = decltype(kc)::value;
        constexpr int S0 = K, S1 = (K + 3) & 3, S2 = (K + 2) & 3;  // slots of planes zz, zz - 1, zz - 2
        const T dU3 = nx.cU;  // D coefficient of plane zz - 2
        T vD = 0;
        if (!ZIN) vD = r2.vU;  // the input iterate below plane zz (the record of zz - 2 read it)
        load(nx, zz + 1 <= z1 + 1 ? zz + 1 : z1 + 1);
        mask(r0, zz);
        // first sweep at zz
        T t1 = 0;
        if (!ZIN) {
            t1 -= r0.cE * r0.vE;
            t1 -= r0.cW * r0.vW;
            t1 -= r0.cN * r0.vN;
            t1 -= r0.cS * r0.vS;
            t1 -= r0.cU * r0.vU;
            t1 -= r1.cU * vD;
        }
        const T s1 = rnd(comb(r0.inv, t1, r0.rhs));
        L1[S0 * PS3_THREADS + tid] = s1;
        if constexpr (LP == 0) {
            // the chunk's own cells of the first sweep, exact on the whole tile
            if (p.rec1 && inner && zz >= z0 && zz < z1) st_at(p.rec1, (uint32_t)zz * planeB + colB, s1);
        }
        // second sweep at z2 = zz - 1 (needed from z0 - 1)
        const int z2 = zz - 1;
        T s2 = 0;
        if (zz >= z0) {
            const T* L = L1 + S1 * PS3_THREADS;
            const T vE = L[jE], vW = L[jW], vN = L[jN], vS = L[jS];
            const T vU = z2 + 1 < p.Z ? s1 : s1m1, vDn = z2 > 0 ? s1m2 : s1m1;
            T t2 = 0;
            t2 -= r1.cE * vE;
            t2 -= r1.cW * vW;
            t2 -= r1.cN * vN;
            t2 -= r1.cS * vS;
            t2 -= r1.cU * vU;
            t2 -= r2.cU * vDn;
            s2 = rnd(comb(r1.inv, t2, r1.rhs));
            if constexpr (LP == 0) {
                // exact inside the tile's outer ring, so on the inner cells
                if (p.rec2 && inner && z2 >= z0 && z2 < z1) st_at(p.rec2, (uint32_t)z2 * planeB + colB, s2);
            }
        }
        L2[S1 * PS3_THREADS + tid] = s2;
        // third sweep at z3 = zz - 2 into the output
        const int z3 = zz - 2;
        if (inner && z3 >= z0 && z3 < z1) {
            const T* L = L2 + S2 * PS3_THREADS;
            const T vE = L[jE], vW = L[jW], vN = L[jN], vS = L[jS];
            const T vU = z3 + 1 < p.Z ? s2 : s2m1, vDn = z3 > 0 ? s2m2 : s2m1;
            T t3 = 0;
            t3 -= r2.cE * vE;
            t3 -= r2.cW * vW;
            t3 -= r2.cN * vN;
            t3 -= r2.cS * vS;
            t3 -= r2.cU * vU;
            t3 -= dU3 * vDn;
            const T e3 = comb(r2.inv, t3, r2.rhs);
            if constexpr (LP == 0) {
                // an inner thread's column is in the grid: colB is its own
                st_at(out, (uint32_t)z3 * planeB + colB, e3);
            } else {
                const int P3 = (z3 * p.Y + y) * p.X + x;
                if constexpr (LAST) {
                    const T xv = p.xin ? p.xin[P3] : (T)0;
                    out[P3] = xv + e3 / sc;
                } else {
                    st_lp<LP == 2>(p.eout, P3, (float)e3);
                }
            }
        }
        __syncthreads();
        s1m2 = s1m1;
        s1m1 = s1;
        s2m2 = s2m1;
        s2m1 = s2;
    };
    // A = plane z0 - 4, B = z0 - 3 (partial records), C = z0 - 2; D is the
    // record the first step loads plane z0 - 1 into (its plane z0 - 5 value
    // would only feed a third sweep below the chunk, which never runs)
    PS2Plane<T> A, B, C, D;
    load_below(A, z0 - 4, false);
    load_below(B, z0 - 3, true);
    load(C, z0 - 2);
    D = A;
    // C is masked by its first step
    mask(A, z0 - 4);
    mask(B, z0 - 3);
    using K0 = std::integral_constant<int, 0>;
    using K1 = std::integral_constant<int, 1>;
    using K2 = std::integral_constant<int, 2>;
    using K3 = std::integral_constant<int, 3>;
    for (int zz = z0 - 2;;) {
        step(K0{}, A, B, C, D, zz);
        if (++zz > z1 + 1) break;
        step(K1{}, B, C, D, A, zz);
        if (++zz > z1 + 1) break;
        step(K2{}, C, D, A, B, zz);
        if (++zz > z1 + 1) break;
        step(K3{}, D, A, B, C, zz);
        if (++zz > z1 + 1) break;
    }
}

// ---------------------------------------------------------------------------
// Four pressure Jacobi sweeps in one launch: the three-sweep kernel one level
// deeper (FP32, single cases).  A workgroup of 70 x 14 threads owns a 64 x 8
// column of cells (three halo cells on each side) and marches z from three
// planes below its chunk to three above.  At plane zz every thread computes
// the first sweep at zz (ring 1), the second at zz - 1 (ring 2), the third at
// zz - 2 (ring 3) and the interior threads the fourth at zz - 3 into the
// output; one barrier per plane.  Five plane records rotate through named
// variables (the loop is unrolled five times, the period of the five-slot
// rings, so LDS addresses are constant offsets).  Expressions, masks and
// non-fluid handling are the three-sweep kernel's, so the result equals four
// single sweeps.
// ---------------------------------------------------------------------------
constexpr int PS4_W = 70;                    // threads per tile row: 64 outputs + 2 x 3 halo
constexpr int PS4_BY = 14;                   // tile rows: 8 outputs + 2 x 3 halo
constexpr int PS4_THREADS = PS4_W * PS4_BY;  // 980
constexpr int PS4_RING = 5;                  // planes per LDS ring (three rings)

inline __host__ __device__ int pres_sweep4_blocks(int X, int Y, int Z, int zchunk) {
    return ((X + 63) / 64) * ((Y + PS4_BY - 7) / (PS4_BY - 6)) * ((Z + zchunk - 1) / zchunk);
}

// LP, LAST: the low-precision correction equation as in the three-sweep
// kernel (e read in 2 bytes, the first three sweeps' values rounded to 2
// bytes as the stored iterate would be, the fourth to e_out or, as the
// round's last sweep, to p' = x + e / s)
template <bool ZIN, int LP = 0, bool LAST = false>
__device__ __forceinline__ void pres_sweep4_block(const PS2Args<float>& p, int b, float* lds) {
    using T = float;
    const int tid = (int)threadIdx.x;
    const int lx = tid % PS4_W, ly = tid / PS4_W;
    const int ntx = (p.X + 63) / 64, nty = (p.Y + PS4_BY - 7) / (PS4_BY - 6);
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int x = bx * 64 - 3 + lx, y = by * (PS4_BY - 6) - 3 + ly;
    const int z0 = bz * p.zchunk, z1 = min(z0 + p.zchunk, p.Z);
    const bool cell = x >= 0 && x < p.X && y >= 0 && y < p.Y;
    const bool inner = cell && lx >= 3 && lx <= PS4_W - 4 && ly >= 3 && ly <= PS4_BY - 4;
    const PresCoef<T> k = p.k;
    const T* __restrict__ x0 = ZIN ? nullptr : p.x0;
    T* __restrict__ out = p.out;
    const bool hW = x > 0, hE = x + 1 < p.X, hS = y > 0, hN = y + 1 < p.Y;
    const int xc = min(max(x, 0), p.X - 1), yc = min(max(y, 0), p.Y - 1);
    const bool gE = xc + 1 < p.X, gW = xc > 0, gN = yc + 1 < p.Y, gS = yc > 0;
    const uint32_t colB = (uint32_t)(yc * p.X + xc) * (uint32_t)sizeof(T);
    const uint32_t planeB = (uint32_t)(p.X * p.Y) * (uint32_t)sizeof(T);
    const uint32_t rowB = (uint32_t)(p.X * sizeof(T));
    // the record of plane zz: its coefficients and the input iterate at its
    // E, W, N, S neighbours and at the plane above (clamped into the grid)
    auto load = [&](PS2Plane<T>& q, int zz) {
        const uint32_t oP = (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB;
        const uint32_t oU = (uint32_t)min(max(zz + 1, 0), p.Z - 1) * planeB + colB;
        const uint32_t oW = gW ? oP - (uint32_t)sizeof(T) : oP, oS = gS ? oP - rowB : oP;
        q.cE = ld_at(k.cE, oP); q.cN = ld_at(k.cN, oP); q.cU = ld_at(k.cU, oP);
        q.cW = ld_at(k.cE, oW);
        q.cS = ld_at(k.cN, oS);
        q.inv = ld_at(k.invDp, oP);
        q.rhs = ld_at(k.rhs, oP);
        if (!ZIN) {
            if constexpr (LP == 0) {
                q.vE = ld_at(x0, gE ? oP + (uint32_t)sizeof(T) : oP);
                q.vW = ld_at(x0, oW);
                q.vN = ld_at(x0, gN ? oP + rowB : oP);
                q.vS = ld_at(x0, oS);
                q.vU = ld_at(x0, oU);
            } else {
                const uint32_t es = (uint32_t)sizeof(T);
                q.vE = ld_lp<LP == 2>(p.ein, (int)((gE ? oP + es : oP) / es));
                q.vW = ld_lp<LP == 2>(p.ein, (int)(oW / es));
                q.vN = ld_lp<LP == 2>(p.ein, (int)((gN ? oP + rowB : oP) / es));
                q.vS = ld_lp<LP == 2>(p.ein, (int)(oS / es));
                q.vU = ld_lp<LP == 2>(p.ein, (int)(oU / es));
            }
        }
    };
    // the records below the first computed plane are read only in part:
    // plane z0 - 5 for the input iterate above it, plane z0 - 4 for that and
    // its U coefficient (the D coefficient of plane z0 - 3)
    auto load_below = [&](PS2Plane<T>& q, int zz, bool with_cU) {
        const uint32_t oP = (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB;
        const uint32_t oU = (uint32_t)min(max(zz + 1, 0), p.Z - 1) * planeB + colB;
        q.cE = q.cW = q.cN = q.cS = q.inv = q.rhs = 0;
        q.vE = q.vW = q.vN = q.vS = 0;
        q.cU = with_cU ? ld_at(k.cU, oP) : (T)0;
        q.vU = 0;
        if (!ZIN) {
            if constexpr (LP == 0) q.vU = ld_at(x0, oU);
            else q.vU = ld_lp<LP == 2>(p.ein, (int)(oU / (uint32_t)sizeof(T)));
        }
    };
    const T sc = LP == 0 ? (T)0 : p.s;
    // one sweep's value from its term, and the value the next sweep reads
    auto comb = [&](T inv, T t, T rhs) {
        if constexpr (LP == 0) return inv * (t + rhs);
        else return inv * t + sc * rhs;
    };
    auto rnd = [&](T v) {
        if constexpr (LP == 0) return v;
        else if constexpr (LP == 2) return (T)__bfloat162float(__float2bfloat16((float)v));
        else return (T)(float)(_Float16)(float)v;
    };
    const int jE = (hE && lx + 1 < PS4_W) ? tid + 1 : tid, jW = (hW && lx > 0) ? tid - 1 : tid;
    const int jN = (hN && ly + 1 < PS4_BY) ? tid + PS4_W : tid, jS = (hS && ly > 0) ? tid - PS4_W : tid;
    T* const L1 = lds;
    T* const L2 = lds + PS4_RING * PS4_THREADS;
    T* const L3 = lds + 2 * PS4_RING * PS4_THREADS;
    auto mask = [&](PS2Plane<T>& q, int zz) {
        q.cW = hW ? q.cW : (T)0;
        q.cS = hS ? q.cS : (T)0;
        if (zz < 0) q.cU = 0;
    };
    // a cell's in-plane neighbours in one ring plane
    struct Nb {
        T e, w, n, s;
    };
    auto nbr = [&](const T* L) { return Nb{L[jE], L[jW], L[jN], L[jS]}; };
    // one sweep's value at a cell from its neighbours in the level below
    auto sweep = [&](const PS2Plane<T>& r, const Nb& v, T vU, T vDn, T cD) {
        T t = 0;
        t -= r.cE * v.e;
        t -= r.cW * v.w;
        t -= r.cN * v.n;
        t -= r.cS * v.s;
        t -= r.cU * vU;
        t -= cD * vDn;
        return comb(r.inv, t, r.rhs);
    };
    T s1m1 = 0, s1m2 = 0;  // own first-sweep values at zz - 1, zz - 2
    T s2m1 = 0, s2m2 = 0;  // own second-sweep values at zz - 2, zz - 3
    T s3m1 = 0, s3m2 = 0;  // own third-sweep values at zz - 3, zz - 4
    // iteration zz: r0 = plane zz, r1 = zz - 1, r2 = zz - 2, r3 = zz - 3,
    // nx = zz - 4 (receives zz + 1); kc: the step's position in the unrolled loop
    auto step = [&](auto kc, PS2Plane<T>& r3, PS2Plane<T>& r2, PS2Plane<T>& r1, PS2Plane<T>& r0, PS2Plane<T>& nx,
                    int zz) {
        constexpr int K = decltype(kc)::value;
        // ring slots of planes zz, zz - 1, zz - 2, zz - 3
        constexpr int S0 = K, S1 = (K + 4) % 5, S2 = (K + 3) % 5, S3 = (K + 2) % 5;
        const T dU4 = nx.cU;  // D coefficient of plane zz - 3
        T vD = 0;
        if (!ZIN) vD = r2.vU;  // the input iterate below plane zz (the record of zz - 2 read it)
        // the ring planes the second to fourth sweeps read were written in
        // the previous step, before its barrier: their reads are issued
        // first, together, so that their latency overlaps the first sweep's
        // wait for its record instead of following one another
        const Nb a2 = nbr(L1 + S1 * PS4_THREADS);
        const Nb a3 = nbr(L2 + S2 * PS4_THREADS);
        const Nb a4 = nbr(L3 + S3 * PS4_THREADS);
        load(nx, zz + 1 <= z1 + 2 ? zz + 1 : z1 + 2);
        mask(r0, zz);
        // first sweep at zz
        T t1 = 0;
        if (!ZIN) {
            t1 -= r0.cE * r0.vE;
            t1 -= r0.cW * r0.vW;
            t1 -= r0.cN * r0.vN;
            t1 -= r0.cS * r0.vS;
            t1 -= r0.cU * r0.vU;
            t1 -= r1.cU * vD;
        }
        const T s1 = rnd(comb(r0.inv, t1, r0.rhs));
        L1[S0 * PS4_THREADS + tid] = s1;
        // second sweep at zz - 1 (needed from z0 - 2; +0 below)
        const int z2 = zz - 1;
        const T s2v = rnd(sweep(r1, a2, z2 + 1 < p.Z ? s1 : s1m1, z2 > 0 ? s1m2 : s1m1, r2.cU));
        const T s2 = zz >= z0 - 1 ? s2v : (T)0;
        L2[S1 * PS4_THREADS + tid] = s2;
        // third sweep at zz - 2 (needed from z0 - 1; +0 below)
        const int z3 = zz - 2;
        const T s3v = rnd(sweep(r2, a3, z3 + 1 < p.Z ? s2 : s2m1, z3 > 0 ? s2m2 : s2m1, r3.cU));
        const T s3 = zz >= z0 + 1 ? s3v : (T)0;
        L3[S2 * PS4_THREADS + tid] = s3;
        // fourth sweep at zz - 3 into the output
        const int z4 = zz - 3;
        const T e4 = sweep(r3, a4, z4 + 1 < p.Z ? s3 : s3m1, z4 > 0 ? s3m2 : s3m1, dU4);
        if (inner && z4 >= z0 && z4 < z1) {
            if constexpr (LP == 0) {
                // an inner thread's column is in the grid: colB is its own
                st_at(out, (uint32_t)z4 * planeB + colB, e4);
            } else {
                const int P4 = (z4 * p.Y + y) * p.X + x;
                if constexpr (LAST) {
                    const T xv = p.xin ? p.xin[P4] : (T)0;
                    out[P4] = xv + e4 / sc;
                } else {
                    st_lp<LP == 2>(p.eout, P4, (float)e4);
                }
            }
        }
        __syncthreads();
        s1m2 = s1m1;
        s1m1 = s1;
        s2m2 = s2m1;
        s2m1 = s2;
        s3m2 = s3m1;
        s3m1 = s3;
    };
    // A = plane z0 - 6 and E = z0 - 7 (not read), B = z0 - 5, C = z0 - 4
    // (partial records), D = z0 - 3
    PS2Plane<T> A, B, C, D, E;
    A.cE = A.cW = A.cN = A.cS = A.cU = A.inv = A.rhs = 0;
    A.vE = A.vW = A.vN = A.vS = A.vU = 0;
    E = A;
    load_below(B, z0 - 5, false);
    load_below(C, z0 - 4, true);
    load(D, z0 - 3);
    // D is masked by its first step
    mask(B, z0 - 5);
    mask(C, z0 - 4);
    using K0 = std::integral_constant<int, 0>;
    using K1 = std::integral_constant<int, 1>;
    using K2 = std::integral_constant<int, 2>;
    using K3 = std::integral_constant<int, 3>;
    using K4 = std::integral_constant<int, 4>;
    for (int zz = z0 - 3;;) {
        step(K0{}, A, B, C, D, E, zz);
        if (++zz > z1 + 2) break;
        step(K1{}, B, C, D, E, A, zz);
        if (++zz > z1 + 2) break;
        step(K2{}, C, D, E, A, B, zz);
        if (++zz > z1 + 2) break;
        step(K3{}, D, E, A, B, C, zz);
        if (++zz > z1 + 2) break;
        step(K4{}, E, A, B, C, D, zz);
        if (++zz > z1 + 2) break;
    }
}

template <bool ZIN>
__global__ __launch_bounds__(1024) void k_pres_sweep4_f32(ParamsT<float> p, FieldsT<float> f, Roles r) {
    __shared__ __align__(16) float lds_f[3 * PS4_RING * PS4_THREADS];
    const int nb = pres_sweep4_blocks(p.X, p.Y, p.Z, p.zchunk4);
    PS2Args<float> a = ps2_args(p, f, r);
    a.zchunk = p.zchunk4;
    pres_sweep4_block<ZIN>(a, xcd_tile((int)blockIdx.x, nb), lds_f);
}

template <bool ZIN>
__global__ __launch_bounds__(1024) void k_pres_sweep3_f32(ParamsT<float> p, FieldsT<float> f, Roles r) {
    __shared__ __align__(16) float lds_f[2 * PS3_RING * PS3_THREADS];
    const int nb = pres_sweep3_blocks(p.X, p.Y, p.Z, p.zchunk3);
    PS2Args<float> a = ps2_args(p, f, r);
    a.zchunk = p.zchunk3;
    pres_sweep3_block<float, ZIN>(a, xcd_tile((int)blockIdx.x, nb), lds_f);
}
// three sweeps of the low-precision correction equation (FIRST = ZIN: e = 0
// on input; LAST: the third sweep is the round's last); roles as
// body_lowp_sweep's
template <bool ZIN, int LP, bool LAST>
__global__ __launch_bounds__(1024) void k_lowp_sweep3(ParamsT<float> p, FieldsT<float> f, Roles r) {
    __shared__ __align__(16) float lds_f[2 * PS3_RING * PS3_THREADS];
    const int nb = pres_sweep3_blocks(p.X, p.Y, p.Z, p.zchunk3);
    PS2Args<float> a = ps2_args(p, f, r);
    a.zchunk = p.zchunk3;
    a.k.rhs = f.rho;
    a.ein = ZIN ? nullptr : f.e16[r.ein];
    a.eout = f.e16[r.eout];
    a.xin = r.pin >= 0 ? f.pc[r.pin] : nullptr;
    a.out = f.pc[r.pout];
    a.s = f.scalar[1];
    pres_sweep3_block<float, ZIN, LP, LAST>(a, xcd_tile((int)blockIdx.x, nb), lds_f);
}

// four low-precision correction sweeps in one launch (the four-sweep
// kernel's tiling; grids that take four-sweep pressure launches)
template <bool ZIN, int LP, bool LAST>
__global__ __launch_bounds__(1024) void k_lowp_sweep4(ParamsT<float> p, FieldsT<float> f, Roles r) {
    __shared__ __align__(16) float lds_f[3 * PS4_RING * PS4_THREADS];
    const int nb = pres_sweep4_blocks(p.X, p.Y, p.Z, p.zchunk4);
    PS2Args<float> a = ps2_args(p, f, r);
    a.zchunk = p.zchunk4;
    a.k.rhs = f.rho;
    a.ein = ZIN ? nullptr : f.e16[r.ein];
    a.eout = f.e16[r.eout];
    a.xin = r.pin >= 0 ? f.pc[r.pin] : nullptr;
    a.out = f.pc[r.pout];
    a.s = f.scalar[1];
    pres_sweep4_block<ZIN, LP, LAST>(a, xcd_tile((int)blockIdx.x, nb), lds_f);
}

// ---------------------------------------------------------------------------
// Two momentum Jacobi sweeps in one launch: the tiling, z-march and x1 ring
// of the two-sweep pressure kernel, for the three velocity components at
// once (the face coefficients S, G are shared by them).  Both sweeps use the
// expressions and operation order of body_vel_sweep.
//
// Registers are what limits this kernel (a 16-wave workgroup must fit in 128
// VGPRs without spilling: a spill reload would make every iteration wait for
// the loads of the next plane too), so the in-plane neighbours of the input
// iterate come from LDS rather than from per-thread loads:
//  - each thread loads its own column's iterate two planes ahead and stores
//    it in a three-plane LDS ring one iteration ahead of its use; the first
//    sweep at plane zz reads the E, W, N, S values of plane zz from the ring
//    and the D value (plane zz - 1) from its own ring entry;
//  - the tile's edge threads (lx 0 or 65, ly 0 or 14) also load the one
//    neighbour outside the tile they need (corner threads' results are never
//    used);
//  - the coefficients (own E/N/U faces and the W, S faces of the neighbours)
//    are loaded per thread one plane ahead, in three records that rotate
//    through named variables (the loop is unrolled three times, the period of
//    both LDS rings, so every LDS address is a constant offset).
//  - a record's boundary masks are applied once, before the first sweep that
//    reads it.
// Differences from the pressure kernel:
//  - the W, S and D coefficients are zeroed on the domain boundary by a bit
//    mask, to +0 (S can be negative: a product with 0 would give -0 where the
//    one-sweep kernel has the +0 of its select);
//  - the flag is read: the first sweep's value of a non-fluid cell is +0 (its
//    neighbours' coefficients towards it are zero), and the second sweep
//    writes fluid cells only, as the one-sweep kernel does.
// The one-sweep kernel reads, at a non-fluid neighbour, whatever its input
// buffer holds there (zero: solid cells are zeroed in every buffer and no
// sweep writes them); the product with the zero coefficient makes any finite
// value there a signed zero, so the two paths agree in every value.
// ---------------------------------------------------------------------------
template <class T>
struct VS2Coef {
    T SE, SN, SU, GE, GN, GU, SW, GW, SS, GS;
    T inv[3], b[3];
    T vO[3];  // edge threads: the input iterate at the neighbour outside the tile
    int fl;
};
template <class T>
struct VS2Args {
    int X, Y, Z, zchunk, by;
    T q;  // ds / 4
    const T* h[3];  // stretched grid: the cell sizes that replace q per direction (nullptr: uniform)
    const T* S[3];
    const T* G[3];
    const T* invD[3];
    const T* b[3];
    const uint8_t* flag;
    const T* vin[3];
    T* out[3];
};
template <class T>
__device__ __forceinline__ VS2Args<T> vs2_args(const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r) {
    VS2Args<T> a;
    a.X = p.X; a.Y = p.Y; a.Z = p.Z; a.zchunk = p.zchunk; a.by = p.by2;
    a.q = p.ds * (T)0.25f;
    for (int c = 0; c < 3; ++c) {
        a.h[c] = p.nonuniform ? f.h[c] : nullptr;
        a.S[c] = f.S[c]; a.G[c] = f.G[c]; a.invD[c] = f.invD[c]; a.b[c] = f.b[c];
        a.vin[c] = f.vel[r.vin][c];
        a.out[c] = f.vel[r.vout][c];
    }
    a.flag = f.flag;
    return a;
}
// v where keep holds, else +0 (a bit mask: the value stays used, so its load
// is not moved under a branch)
__device__ __forceinline__ float keep_or_zero(float v, bool keep) {
    return __uint_as_float(__float_as_uint(v) & (keep ? 0xffffffffu : 0u));
}
__device__ __forceinline__ double keep_or_zero(double v, bool keep) {
    return __longlong_as_double(__double_as_longlong(v) & (keep ? -1ll : 0ll));
}
// LDS of one workgroup: the x1 ring and the input ring, 3 planes of three
// components each
constexpr int VS2_LDS_FLOATS = (3 + 3) * 3 * PS2_THREADS;

// NU: the stretched grid, whose sweep (body_vel_sweep<T, true>) weighs the
// convective terms of the x, y and z faces by the cell sizes h[0][x],
// h[1][y], h[2][z] instead of ds / 4
template <class T, bool NU = false>
__device__ __forceinline__ void vel_sweep2_block(const VS2Args<T>& p, int b, T* lds) {
    int lx, ly;
    tile_lane((int)threadIdx.x, 1, p.by, lx, ly);
    const int tid = ly * PS2_W + lx;  // the thread's entry in the LDS planes
    const int ntx = (p.X + 63) / 64, nty = (p.Y + p.by - 3) / (p.by - 2);
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int x = bx * 64 - 1 + lx, y = by * (p.by - 2) - 1 + ly;
    const int z0 = bz * p.zchunk, z1 = min(z0 + p.zchunk, p.Z);
    const bool cell = x >= 0 && x < p.X && y >= 0 && y < p.Y;
    const bool inner = cell && lx >= 1 && lx <= 64 && ly >= 1 && ly <= p.by - 2;
    const int sy = p.X;
    const bool hW = x > 0, hE = x + 1 < p.X, hS = y > 0, hN = y + 1 < p.Y;
    const int xc = min(max(x, 0), p.X - 1), yc = min(max(y, 0), p.Y - 1);
    const bool gE = xc + 1 < p.X, gW = xc > 0, gN = yc + 1 < p.Y, gS = yc > 0;
    // tile edges: the neighbour outside the tile comes from the thread's own load
    const bool eW = lx == 0, eE = lx == PS2_W - 1, eS = ly == 0, eN = ly == p.by - 1;
    const bool edge = eW || eE || eS || eN;
    const int oO = eW ? (gW ? -1 : 0) : eE ? (gE ? 1 : 0) : eS ? (gS ? -sy : 0) : eN ? (gN ? sy : 0) : 0;
    // LDS indices of the in-tile neighbours (the thread itself on a tile edge)
    const int iE = eE ? tid : tid + 1, iW = eW ? tid : tid - 1;
    const int iN = eN ? tid : tid + PS2_W, iS = eS ? tid : tid - PS2_W;
    // both rings hold plane z0 - 1 + 3m + k in slot k, so unrolled step k
    // addresses its slots with constant offsets
    T* const lx1 = lds;                      // [3 slots][3 components][PS2_THREADS]
    T* const lvi = lds + 9 * PS2_THREADS;    // [3 slots][3 components][PS2_THREADS]
    const T q = p.q;
    // per-direction factors of the convective terms (the plane's z factor is
    // uniform and read per plane)
    const T qx = NU ? p.h[0][xc] : q, qy = NU ? p.h[1][yc] : q;
    auto qz_of = [&](int zz) { return NU ? p.h[2][min(max(zz, 0), p.Z - 1)] : q; };
    // addresses: uniform base + 32-bit byte offset (column offset plus a
    // uniform plane offset); the W and S coefficients are read at the clamped
    // W/S neighbour (the cell itself on the boundary, masked to +0)
    const int col = yc * p.X + xc;
    const uint32_t colB = (uint32_t)col * (uint32_t)sizeof(T);
    const uint32_t planeB = (uint32_t)(p.X * p.Y) * (uint32_t)sizeof(T);
    const uint32_t dW = gW ? 0u - (uint32_t)sizeof(T) : 0u, dS = gS ? 0u - (uint32_t)(sy * sizeof(T)) : 0u;
    const uint32_t dO = (uint32_t)oO * (uint32_t)sizeof(T);
    auto planeoff = [&](int zz) { return (uint32_t)min(max(zz, 0), p.Z - 1) * planeB + colB; };
    auto loadc = [&](VS2Coef<T>& r, int zz) {
        const uint32_t oP = planeoff(zz);
        const uint32_t oW = oP + dW, oS = oP + dS;
        r.SE = ld_at(p.S[0], oP); r.SN = ld_at(p.S[1], oP); r.SU = ld_at(p.S[2], oP);
        r.GE = ld_at(p.G[0], oP); r.GN = ld_at(p.G[1], oP); r.GU = ld_at(p.G[2], oP);
        r.SW = ld_at(p.S[0], oW); r.GW = ld_at(p.G[0], oW);
        r.SS = ld_at(p.S[1], oS); r.GS = ld_at(p.G[1], oS);
        r.fl = ld_at(p.flag, oP / (uint32_t)sizeof(T));
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r.inv[a] = ld_at(p.invD[a], oP);
            r.b[a] = ld_at(p.b[a], oP);
            // only the tile's edge threads need the neighbour outside the
            // tile (a test of the lane's position, so no load waits on it)
            r.vO[a] = edge ? ld_at(p.vin[a], oP + dO) : (T)0;
        }
    };
    auto loadv = [&](T v[3], int zz) {
        const uint32_t oP = planeoff(zz);
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = ld_at(p.vin[a], oP);
    };
    // LDS index of each in-plane neighbour's first-sweep value in the second
    // sweep (the thread's own entry on the domain boundary)
    const int jE = hE ? tid + 1 : tid, jW = hW ? tid - 1 : tid;
    const int jN = hN ? tid + PS2_W : tid, jS = hS ? tid - PS2_W : tid;
    // A record's boundary masks, applied once before its first use (the
    // values both sweeps read): W and S on the domain boundary, and the U
    // coefficients of a plane below the grid, which are the D coefficients of
    // the plane above
    auto maskc = [&](VS2Coef<T>& r, int zz) {
        r.SW = keep_or_zero(r.SW, hW); r.GW = keep_or_zero(r.GW, hW);
        r.SS = keep_or_zero(r.SS, hS); r.GS = keep_or_zero(r.GS, hS);
        if (zz < 0) { r.SU = 0; r.GU = 0; }
    };
    T x1m1[3] = {0, 0, 0}, x1m2[3] = {0, 0, 0};
    T pdS = 0, pdG = 0;  // U coefficients of the plane below the previous one
    // one plane: loads of the coefficients of zz + 1 into nx and of the own
    // iterate at zz + 2 into vn; first sweep at zz (vu: own iterate at zz + 1);
    // vu into the input ring; second sweep at zz - 1 from pv and the x1 ring.
    // kc: the step's position in the unrolled loop
    auto step = [&](auto kc, const VS2Coef<T>& pv, VS2Coef<T>& cu, VS2Coef<T>& nx, const T vu[3], T vn[3],
                    int zz) {
        constexpr int K = decltype(kc)::value;
        constexpr int S0 = K, Sm = (K + 2) % 3, Sp = (K + 1) % 3;  // slots of planes zz, zz - 1, zz + 1
        loadc(nx, zz + 1 <= z1 ? zz + 1 : z1);
        loadv(vn, zz + 2 <= z1 + 1 ? zz + 2 : z1 + 1);
        maskc(cu, zz);
        const T SW = cu.SW, GW = cu.GW, SS = cu.SS, GS = cu.GS;
        const T SD = pv.SU, GD = pv.GU;
        const T qz1 = qz_of(zz);
        const T* Lc = lvi + S0 * (3 * PS2_THREADS);
        const T* Ld = lvi + Sm * (3 * PS2_THREADS);
        T x1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T* L = Lc + a * PS2_THREADS;
            const T lE = L[iE], lW = L[iW], lN = L[iN], lS = L[iS];
            const T vE = eE ? cu.vO[a] : lE, vW = eW ? cu.vO[a] : lW;
            const T vN = eN ? cu.vO[a] : lN, vS = eS ? cu.vO[a] : lS;
            const T vU = vu[a], vD = Ld[a * PS2_THREADS + tid];
            T A = 0;
            A -= qx * vE * cu.SE;
            A += qx * vW * SW;
            A -= qy * vN * cu.SN;
            A += qy * vS * SS;
            A -= qz1 * vU * cu.SU;
            A += qz1 * vD * SD;
            T Df = 0;
            Df += cu.GE * vE;
            Df += GW * vW;
            Df += cu.GN * vN;
            Df += GS * vS;
            Df += cu.GU * vU;
            Df += GD * vD;
            const T v = cu.inv[a] * (A + Df + cu.b[a]);
            x1[a] = cu.fl == UWS_CELL_FLUID ? v : (T)0;
            lx1[S0 * (3 * PS2_THREADS) + a * PS2_THREADS + tid] = x1[a];
        }
        // the own iterate at zz + 1 for the next plane's first sweep (the slot
        // of plane zz - 2, whose last reader was the previous plane)
        T* Ln = lvi + Sp * (3 * PS2_THREADS);
#pragma unroll
        for (int a = 0; a < 3; ++a) Ln[a * PS2_THREADS + tid] = vu[a];
        __syncthreads();
        const int zo = zz - 1;
        if (inner && zo >= z0 && zo < z1 && pv.fl == UWS_CELL_FLUID) {
            // an inner thread's column is in the grid: the offset is its own
            const uint32_t oP = planeoff(zo);
            const T SW2 = pv.SW, GW2 = pv.GW, SS2 = pv.SS, GS2 = pv.GS;
            const T SD2 = pdS, GD2 = pdG;
            const T qz2 = qz_of(zo);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T* L = lx1 + Sm * (3 * PS2_THREADS) + a * PS2_THREADS;
                const T vE = L[jE], vW = L[jW], vN = L[jN], vS = L[jS];
                const T vU = zo + 1 < p.Z ? x1[a] : x1m1[a], vD2 = zo > 0 ? x1m2[a] : x1m1[a];
                T A = 0;
                A -= qx * vE * pv.SE;
                A += qx * vW * SW2;
                A -= qy * vN * pv.SN;
                A += qy * vS * SS2;
                A -= qz2 * vU * pv.SU;
                A += qz2 * vD2 * SD2;
                T Df = 0;
                Df += pv.GE * vE;
                Df += GW2 * vW;
                Df += pv.GN * vN;
                Df += GS2 * vS;
                Df += pv.GU * vU;
                Df += GD2 * vD2;
                st_at(p.out[a], oP, pv.inv[a] * (A + Df + pv.b[a]));
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            x1m2[a] = x1m1[a];
            x1m1[a] = x1[a];
        }
        pdS = pv.SU;
        pdG = pv.GU;
    };
    // prologue: the input ring holds the own iterate at z0 - 2 (slot 2) and
    // z0 - 1 (slot 0), the first plane's D and in-plane values; V1 that at z0
    // (its U value)
    {
        T v[3];
        loadv(v, z0 - 2);
#pragma unroll
        for (int a = 0; a < 3; ++a) lvi[2 * (3 * PS2_THREADS) + a * PS2_THREADS + tid] = v[a];
        loadv(v, z0 - 1);
#pragma unroll
        for (int a = 0; a < 3; ++a) lvi[a * PS2_THREADS + tid] = v[a];
    }
    // records on entry of step(pv, cu, nx, ...): pv = plane zz - 1, cu = plane zz
    VS2Coef<T> A, B, C;
    T V1[3], V2[3], V3[3];
    // plane z0 - 2 is read only for its U coefficients (the D coefficients of
    // plane z0 - 1): its other arrays are not loaded
    {
        const uint32_t oP = planeoff(z0 - 2);
        A.SE = A.SN = A.GE = A.GN = A.SW = A.GW = A.SS = A.GS = 0;
        A.SU = ld_at(p.S[2], oP);
        A.GU = ld_at(p.G[2], oP);
        A.fl = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) A.inv[a] = A.b[a] = A.vO[a] = 0;
    }
    loadc(B, z0 - 1);
    loadv(V1, z0);
    maskc(A, z0 - 2);  // B is masked by its first step
    __syncthreads();
    using K0 = std::integral_constant<int, 0>;
    using K1 = std::integral_constant<int, 1>;
    using K2 = std::integral_constant<int, 2>;
    for (int zz = z0 - 1;;) {
        step(K0{}, A, B, C, V1, V2, zz);
        if (++zz > z1) break;
        step(K1{}, B, C, A, V2, V3, zz);
        if (++zz > z1) break;
        step(K2{}, C, A, B, V3, V1, zz);
        if (++zz > z1) break;
    }
}

template <bool NU>
__global__ __launch_bounds__(1024) void k_vel_sweep2_f32(ParamsT<float> p, FieldsT<float> f, Roles r) {
    __shared__ __align__(16) float lds_v[VS2_LDS_FLOATS];
    const int nb = pres_sweep2_blocks(p.X, p.Y, p.Z, p.zchunk, p.by2);
    vel_sweep2_block<float, NU>(vs2_args(p, f, r), xcd_tile((int)blockIdx.x, nb), lds_v);
}

// the case owning tile t: the last c with first_tile[c] <= t (binary search)
template <class T>
__device__ __forceinline__ int case_of(const BatchTable<T>& tab, int t) {
    int lo = 0, hi = tab.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.first_tile[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// three-sweep pressure kernel over every case of an ensemble: workgroup t
// belongs to the case c with first_ps3[c] <= t < first_ps3[c + 1]
template <class T, bool ZIN>
__global__ __launch_bounds__(1024) void k_batch_pres_sweep3(BatchTable<T> tab, Roles r) {
    __shared__ __align__(16) unsigned char lds_raw[2 * PS3_RING * PS3_THREADS * sizeof(T)];
    const int t = xcd_tile((int)blockIdx.x, tab.total_ps3);
    int lo = 0, hi = tab.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.first_ps3[mid] <= t) lo = mid; else hi = mid - 1;
    }
    PS2Args<T> a = ps2_args(tab.params[lo], tab.fields[lo], r);
    a.zchunk = tab.params[lo].zchunk3;
    pres_sweep3_block<T, ZIN>(a, t - tab.first_ps3[lo], reinterpret_cast<T*>(lds_raw));
}

// four-sweep pressure kernel over every case of an ensemble: workgroup t
// belongs to the case c with first_ps4[c] <= t < first_ps4[c + 1]
template <bool ZIN>
__global__ __launch_bounds__(1024) void k_batch_pres_sweep4(BatchTable<float> tab, Roles r) {
    __shared__ __align__(16) float lds_f[3 * PS4_RING * PS4_THREADS];
    const int t = xcd_tile((int)blockIdx.x, tab.total_ps4);
    int lo = 0, hi = tab.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.first_ps4[mid] <= t) lo = mid; else hi = mid - 1;
    }
    PS2Args<float> a = ps2_args(tab.params[lo], tab.fields[lo], r);
    a.zchunk = tab.params[lo].zchunk4;
    pres_sweep4_block<ZIN>(a, t - tab.first_ps4[lo], lds_f);
}

// two-sweep pressure kernel over every case of an ensemble: workgroup t
// belongs to the case c with first_ps2[c] <= t < first_ps2[c + 1]
template <class T, bool ZIN>
__global__ __launch_bounds__(1024) void k_batch_pres_sweep2(BatchTable<T> tab, Roles r) {
    __shared__ __align__(16) unsigned char lds_raw[3 * PS2_THREADS * sizeof(T)];
    const int t = xcd_tile((int)blockIdx.x, tab.total_ps2);
    int lo = 0, hi = tab.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.first_ps2[mid] <= t) lo = mid; else hi = mid - 1;
    }
    pres_sweep2_block<T, ZIN, false>(ps2_args(tab.params[lo], tab.fields[lo], r), t - tab.first_ps2[lo],
                              reinterpret_cast<T*>(lds_raw));
}

// two-sweep momentum kernel over every case of an ensemble: it tiles each
// grid as the two-sweep pressure kernel does, so workgroup t belongs to the
// case c with first_ps2[c] <= t < first_ps2[c + 1]
__global__ __launch_bounds__(1024) void k_batch_vel_sweep2(BatchTable<float> tab, Roles r) {
    __shared__ __align__(16) float lds_v[VS2_LDS_FLOATS];
    const int t = xcd_tile((int)blockIdx.x, tab.total_ps2);
    int lo = 0, hi = tab.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.first_ps2[mid] <= t) lo = mid; else hi = mid - 1;
    }
    vel_sweep2_block<float>(vs2_args(tab.params[lo], tab.fields[lo], r), t - tab.first_ps2[lo], lds_v);
}

// Persistent ensemble launch: the grid is a multiple of 8 workgroups, XCD j
// (workgroups j, j+8, ...) walks its own contiguous range of tiles.
__device__ __forceinline__ void xcd_range(int total, int xcd, int* lo, int* hi) {
    const int q = total >> 3, r = total & 7;
    *lo = xcd * q + min(xcd, r);
    *hi = *lo + q + (xcd < r ? 1 : 0);
}

template <class T, int K>
__global__ __launch_bounds__(256) void k_single(ParamsT<T> p, FieldsT<T> f, Roles r) {
    const Cell c = cell_of<SINGLE_TY>(p, xcd_tile((int)blockIdx.x, p.tiles));
    run_body<T, K>(p, f, c, r);
}
// `fields`: the cases' field tables, tab.fields (device memory) or the
// kernel-argument copy of BatchArgFields
template <class T, int K>
__device__ __forceinline__ void batch_tile(const BatchTable<T>& tab, const FieldsT<T>* fields, const Roles& r,
                                           int t) {
    const int cid = case_of(tab, t);
    const ParamsT<T>& p = tab.params[cid];
    const FieldsT<T>& f = fields[cid];
    const Cell c = cell_of(p, t - tab.first_tile[cid]);
    run_body<T, K>(p, f, c, r);
}
template <class T, int K>
__device__ __forceinline__ void batch_tiles(const BatchTable<T>& tab, const FieldsT<T>* fields, const Roles& r) {
    if (!tab.persistent) {
        batch_tile<T, K>(tab, fields, r, xcd_tile((int)blockIdx.x, tab.total_tiles));
        return;
    }
    int lo, hi;
    xcd_range(tab.total_tiles, (int)blockIdx.x & 7, &lo, &hi);
    const int per = (int)gridDim.x >> 3;
    for (int t = lo + ((int)blockIdx.x >> 3); t < hi; t += per) batch_tile<T, K>(tab, fields, r, t);
}
template <class T, int K>
__global__ __launch_bounds__(256) void k_batch(BatchTable<T> tab, Roles r) {
    batch_tiles<T, K>(tab, tab.fields, r);
}
template <class T, int K>
__global__ __launch_bounds__(256) void k_batch_arg(BatchTable<T> tab, Roles r, BatchArgFields<T> af) {
    batch_tiles<T, K>(tab, af.f, r);
}
template <class T, int K>
__global__ __launch_bounds__(256) void k_single_reduce(ParamsT<T> p, FieldsT<T> f, Roles r) {
    const Cell c = cell_of<SINGLE_TY>(p, xcd_tile((int)blockIdx.x, p.tiles));
    const double s = block_sum(run_reduce<T, K>(p, f, c, r));
    if (threadIdx.x == 0 && threadIdx.y == 0) f.partial[c.tile] = s;
}
template <class T, int K>
__device__ __forceinline__ void batch_reduce_tile(const BatchTable<T>& tab, const FieldsT<T>* fields, const Roles& r,
                                                  int t) {
    const int cid = case_of(tab, t);
    const ParamsT<T>& p = tab.params[cid];
    const FieldsT<T>& f = fields[cid];
    const Cell c = cell_of(p, t - tab.first_tile[cid]);
    const double s = block_sum(run_reduce<T, K>(p, f, c, r));
    if (threadIdx.x == 0 && threadIdx.y == 0) f.partial[c.tile] = s;
}
template <class T, int K>
__device__ __forceinline__ void batch_reduce_tiles(const BatchTable<T>& tab, const FieldsT<T>* fields,
                                                   const Roles& r) {
    if (!tab.persistent) {
        batch_reduce_tile<T, K>(tab, fields, r, xcd_tile((int)blockIdx.x, tab.total_tiles));
        return;
    }
    int lo, hi;
    xcd_range(tab.total_tiles, (int)blockIdx.x & 7, &lo, &hi);
    const int per = (int)gridDim.x >> 3;
    for (int t = lo + ((int)blockIdx.x >> 3); t < hi; t += per) batch_reduce_tile<T, K>(tab, fields, r, t);
}
template <class T, int K>
__global__ __launch_bounds__(256) void k_batch_reduce(BatchTable<T> tab, Roles r) {
    batch_reduce_tiles<T, K>(tab, tab.fields, r);
}
template <class T, int K>
__global__ __launch_bounds__(256) void k_batch_reduce_arg(BatchTable<T> tab, Roles r, BatchArgFields<T> af) {
    batch_reduce_tiles<T, K>(tab, af.f, r);
}

// sum one case's partials into norms[slot]; with `scale` also store the
// defect-correction scale 1/rms into scalar[1]
template <class T>
__device__ void sum_partials(const ParamsT<T>& p, const FieldsT<T>& f, int slot, bool scale) {
    double s = 0.0;
    const int tid = threadIdx.y * TX + threadIdx.x;
    for (int i = tid; i < p.tiles; i += TX * (int)blockDim.y) s += f.partial[i];
    s = block_sum(s);
    if (tid == 0) {
        f.norms[slot] = s;
        if (scale) {
            const double rms = sqrt(s / (double)p.N);
            f.scalar[1] = (T)((rms > 0.0 && isfinite(rms)) ? 1.0 / rms : 1.0);
        }
    }
}
template <class T>
__global__ __launch_bounds__(256) void k_single_sum(ParamsT<T> p, FieldsT<T> f, int slot, int scale) {
    sum_partials(p, f, slot, scale != 0);
}
template <class T>
__global__ __launch_bounds__(256) void k_batch_sum(BatchTable<T> tab, int slot, int scale) {
    const int cid = blockIdx.x;
    sum_partials(tab.params[cid], tab.fields[cid], slot, scale != 0);
}

template <class T, int K>
void go_single(const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r, hipStream_t s) {
    hipLaunchKernelGGL((k_single<T, K>), dim3(p.tiles), dim3(TX, SINGLE_TY, 1), 0, s, p, f, r);
}
template <class T, int K>
void go_single_reduce(const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r, hipStream_t s) {
    hipLaunchKernelGGL((k_single_reduce<T, K>), dim3(p.tiles), dim3(TX, SINGLE_TY, 1), 0, s, p, f, r);
}
// with the group's field tables on the host (hf) and at most BATCH_ARG_CASES
// cases, the tables go in the kernel arguments
template <class T>
bool batch_arg_fields(const BatchTable<T>& tab, const FieldsT<T>* hf, BatchArgFields<T>& af) {
    if (!hf || tab.ncases > BATCH_ARG_CASES) return false;
    af = BatchArgFields<T>{};
    for (int i = 0; i < tab.ncases; ++i) af.f[i] = hf[i];
    return true;
}
template <class T, int K>
void go_batch(const BatchTable<T>& tab, int grid, const Roles& r, hipStream_t s, const FieldsT<T>* hf) {
    BatchArgFields<T> af;
    if (batch_arg_fields(tab, hf, af))
        hipLaunchKernelGGL((k_batch_arg<T, K>), dim3(grid), dim3(TX, tab.ty, 1), 0, s, tab, r, af);
    else
        hipLaunchKernelGGL((k_batch<T, K>), dim3(grid), dim3(TX, tab.ty, 1), 0, s, tab, r);
}
template <class T, int K>
void go_batch_reduce(const BatchTable<T>& tab, int grid, const Roles& r, hipStream_t s, const FieldsT<T>* hf) {
    BatchArgFields<T> af;
    if (batch_arg_fields(tab, hf, af))
        hipLaunchKernelGGL((k_batch_reduce_arg<T, K>), dim3(grid), dim3(TX, tab.ty, 1), 0, s, tab, r, af);
    else
        hipLaunchKernelGGL((k_batch_reduce<T, K>), dim3(grid), dim3(TX, tab.ty, 1), 0, s, tab, r);
}

#include "uws_adjoint_op1.h"
#include "uws_adjoint_ops.h"

}  // namespace

template <class A>
void adj_op1(const ParamsT<Dual<A>>& p, const FieldsT<Dual<A>>& f, const Op1Args<A>& a, hipStream_t s) {
    const int g = (a.N + 255) / 256;
    hipLaunchKernelGGL(k_op1_local<A>, dim3(g), dim3(256), 0, s, p, f, a);
    if (a.les) hipLaunchKernelGGL(k_op1_grad<A>, dim3(g), dim3(256), 0, s, p, f, a);
    hipLaunchKernelGGL(k_op1_state<A>, dim3(g), dim3(256), 0, s, p, f, a);
}
template void adj_op1<double>(const ParamsT<Dual<double>>&, const FieldsT<Dual<double>>&, const Op1Args<double>&,
                              hipStream_t);

template <class T, class A>
void adj_correct(const ParamsT<Dual<A>>& p, const AdjCorrect<T, A>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_correct<T, A>), dim3((p.N + 255) / 256), dim3(256), 0, s, p, a);
}
template <class T, class A>
void adj_piso_coef(const ParamsT<Dual<A>>& p, const AdjPisoCoef<T, A>& a, hipStream_t s) {
    const int g = (p.N + 255) / 256;
    hipLaunchKernelGGL((k_adj_piso_lcf<T, A>), dim3(g), dim3(256), 0, s, p, a);
    hipLaunchKernelGGL((k_adj_piso_coef<T, A>), dim3(g), dim3(256), 0, s, p, a);
}
template <class T, class A>
void adj_aodu(const ParamsT<Dual<A>>& p, const AdjAodu<T, A>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_aodu<T, A>), dim3((p.N + 255) / 256), dim3(256), 0, s, p, a);
}
template <class T, class A>
void adj_rhie_chow(const ParamsT<Dual<A>>& p, const AdjRhieChow<T, A>& a, hipStream_t s) {
    const int g = (p.N + 255) / 256;
    hipLaunchKernelGGL((k_adj_rc_local<T, A>), dim3(g), dim3(256), 0, s, p, a);
    hipLaunchKernelGGL((k_adj_rc_gather<T, A>), dim3(g), dim3(256), 0, s, p, a);
}
#define UWS_ADJ_OPS(T)                                                                                     \
    template void adj_correct<T, double>(const ParamsT<Dual<double>>&, const AdjCorrect<T, double>&,       \
                                         hipStream_t);                                                     \
    template void adj_piso_coef<T, double>(const ParamsT<Dual<double>>&, const AdjPisoCoef<T, double>&,   \
                                           hipStream_t);                                                   \
    template void adj_aodu<T, double>(const ParamsT<Dual<double>>&, const AdjAodu<T, double>&, hipStream_t); \
    template void adj_rhie_chow<T, double>(const ParamsT<Dual<double>>&, const AdjRhieChow<T, double>&,   \
                                           hipStream_t);
UWS_ADJ_OPS(float)
UWS_ADJ_OPS(double)
#undef UWS_ADJ_OPS

#define UWS_KERNEL_CASES(GO, GOR)                                  \
    switch (k) {                                                   \
    case KN_WALL_EDDY: GO(KN_WALL_EDDY); break;                    \
    case KN_LES_TERM: GO(KN_LES_TERM); break;                      \
    case KN_VEL_PRECOMP: GO(KN_VEL_PRECOMP); break;                \
    case KN_VEL_SWEEP: GO(KN_VEL_SWEEP); break;                    \
    case KN_PRES_PRECOMP: GO(KN_PRES_PRECOMP); break;              \
    case KN_PRES_SWEEP: GO(KN_PRES_SWEEP); break;                  \
    case KN_APPLY_PRES: GO(KN_APPLY_PRES); break;                  \
    case KN_PRES_NORMALISE: GO(KN_PRES_NORMALISE); break;          \
    case KN_VEL_CORRECT: GO(KN_VEL_CORRECT); break;                \
    case KN_PISO_AODU: GO(KN_PISO_AODU); break;                    \
    case KN_PISO_PRECOMP: GO(KN_PISO_PRECOMP); break;              \
    case KN_PISO_CORRECT: GO(KN_PISO_CORRECT); break;              \
    case KN_ZERO_SOLID_VEL: GO(KN_ZERO_SOLID_VEL); break;          \
    case KN_FILL_PC: GO(KN_FILL_PC); break;                        \
    case KN_LOWP_SWEEP: GO(KN_LOWP_SWEEP); break;                  \
    case KN_DEFECT_RESIDUAL: GOR(KN_DEFECT_RESIDUAL); break;       \
    case KN_RES_VEL: GOR(KN_RES_VEL); break;                       \
    case KN_RES_PRES: GOR(KN_RES_PRES); break;                     \
    case KN_DIVERGENCE: GOR(KN_DIVERGENCE); break;                 \
    case KN_NONFINITE: GOR(KN_NONFINITE); break;                   \
    default: break;                                                \
    }

template <class T>
void launch(Kernel k, const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r, hipStream_t s, int slot) {
    if (k == KN_PRES_SWEEP3) {
        // FP32 only (the step driver asks for it there)
        if constexpr (std::is_same_v<T, float>) {
            const int nb = pres_sweep3_blocks(p.X, p.Y, p.Z, p.zchunk3);
            if (r.pin < 0) hipLaunchKernelGGL((k_pres_sweep3_f32<true>), dim3(nb), dim3(PS3_THREADS), 0, s, p, f, r);
            else hipLaunchKernelGGL((k_pres_sweep3_f32<false>), dim3(nb), dim3(PS3_THREADS), 0, s, p, f, r);
        }
        return;
    }
    if (k == KN_PRES_SWEEP4) {
        // FP32 single cases only (the step driver asks for it there)
        if constexpr (std::is_same_v<T, float>) {
            const int nb = pres_sweep4_blocks(p.X, p.Y, p.Z, p.zchunk4);
            if (r.pin < 0) hipLaunchKernelGGL((k_pres_sweep4_f32<true>), dim3(nb), dim3(PS4_THREADS), 0, s, p, f, r);
            else hipLaunchKernelGGL((k_pres_sweep4_f32<false>), dim3(nb), dim3(PS4_THREADS), 0, s, p, f, r);
        }
        return;
    }
    if (k == KN_LOWP_SWEEP3) {
        if constexpr (std::is_same_v<T, float>) {
            const int nb = pres_sweep3_blocks(p.X, p.Y, p.Z, p.zchunk3);
            const bool first = (r.flags & K_LOWP_FIRST) != 0, last = (r.flags & K_LOWP_LAST) != 0;
            const bool bf = (r.flags & K_BF16) != 0;
#define UWS_LP3(Z, L, LAST)                                                                                \
    hipLaunchKernelGGL((k_lowp_sweep3<Z, L, LAST>), dim3(nb), dim3(PS3_THREADS), 0, s, p, f, r)
            if (bf) {
                if (first) { if (last) UWS_LP3(true, 2, true); else UWS_LP3(true, 2, false); }
                else { if (last) UWS_LP3(false, 2, true); else UWS_LP3(false, 2, false); }
            } else {
                if (first) { if (last) UWS_LP3(true, 1, true); else UWS_LP3(true, 1, false); }
                else { if (last) UWS_LP3(false, 1, true); else UWS_LP3(false, 1, false); }
            }
#undef UWS_LP3
        }
        return;
    }
    if (k == KN_LOWP_SWEEP4) {
        if constexpr (std::is_same_v<T, float>) {
            const int nb = pres_sweep4_blocks(p.X, p.Y, p.Z, p.zchunk4);
            const bool first = (r.flags & K_LOWP_FIRST) != 0, last = (r.flags & K_LOWP_LAST) != 0;
            const bool bf = (r.flags & K_BF16) != 0;
#define UWS_LP4(Z, L, LAST)                                                                                \
    hipLaunchKernelGGL((k_lowp_sweep4<Z, L, LAST>), dim3(nb), dim3(PS4_THREADS), 0, s, p, f, r)
            if (bf) {
                if (first) { if (last) UWS_LP4(true, 2, true); else UWS_LP4(true, 2, false); }
                else { if (last) UWS_LP4(false, 2, true); else UWS_LP4(false, 2, false); }
            } else {
                if (first) { if (last) UWS_LP4(true, 1, true); else UWS_LP4(true, 1, false); }
                else { if (last) UWS_LP4(false, 1, true); else UWS_LP4(false, 1, false); }
            }
#undef UWS_LP4
        }
        return;
    }
    if (k == KN_VEL_SWEEP2) {
        // FP32 only (the step driver asks for it there), uniform or stretched grid
        if constexpr (std::is_same_v<T, float>) {
            const int nb = pres_sweep2_blocks(p.X, p.Y, p.Z, p.zchunk, p.by2);
            const dim3 tb(PS2_W * p.by2);
            if (p.nonuniform) hipLaunchKernelGGL(k_vel_sweep2_f32<true>, dim3(nb), tb, 0, s, p, f, r);
            else hipLaunchKernelGGL(k_vel_sweep2_f32<false>, dim3(nb), tb, 0, s, p, f, r);
        }
        return;
    }
    if (k == KN_PRES_SWEEP2) {
        const int nb = pres_sweep2_blocks(p.X, p.Y, p.Z, p.zchunk, p.by2);
        if constexpr (std::is_same_v<T, float>) {
            if (r.pin < 0) hipLaunchKernelGGL((k_pres_sweep2_f32<true>), dim3(nb), dim3(PS2_W * p.by2), 0, s, p, f, r);
            else hipLaunchKernelGGL((k_pres_sweep2_f32<false>), dim3(nb), dim3(PS2_W * p.by2), 0, s, p, f, r);
        } else {
            if (r.pin < 0) hipLaunchKernelGGL((k_pres_sweep2<T, true>), dim3(nb), dim3(PS2_W * p.by2), 0, s, p, f, r);
            else hipLaunchKernelGGL((k_pres_sweep2<T, false>), dim3(nb), dim3(PS2_W * p.by2), 0, s, p, f, r);
        }
        return;
    }
    if (k == KN_SUM_PARTIALS || k == KN_DEFECT_SCALE) {
        hipLaunchKernelGGL(k_single_sum<T>, dim3(1), dim3(TX, 4, 1), 0, s, p, f, slot,
                           k == KN_DEFECT_SCALE ? 1 : 0);
        return;
    }
    if constexpr (std::is_same_v<T, float> || std::is_same_v<T, double> || std::is_same_v<T, CStep>) {
        if (p.nonuniform) {
            // stretched grid: its own precomputes, momentum sweep, correction and divergence
            switch (k) {
            case KN_VEL_PRECOMP: go_single<T, KN_NU_VEL_PRECOMP>(p, f, r, s); return;
            case KN_VEL_SWEEP: go_single<T, KN_NU_VEL_SWEEP>(p, f, r, s); return;
            case KN_PRES_PRECOMP: go_single<T, KN_NU_PRES_PRECOMP>(p, f, r, s); return;
            case KN_VEL_CORRECT: go_single<T, KN_NU_VEL_CORRECT>(p, f, r, s); return;
            case KN_DIVERGENCE: go_single_reduce<T, KN_NU_DIVERGENCE>(p, f, r, s); return;
            default: break;
            }
        }
    }
#define GO1(K) go_single<T, K>(p, f, r, s)
#define GOR1(K) go_single_reduce<T, K>(p, f, r, s)
    UWS_KERNEL_CASES(GO1, GOR1)
#undef GO1
#undef GOR1
}

template <class T>
void launch_batch(Kernel k, const BatchTable<T>& tab, int grid, const Roles& r, hipStream_t s, int slot,
                  const FieldsT<T>* host_fields) {
    if (k == KN_PRES_SWEEP3) {
        if (r.pin < 0)
            hipLaunchKernelGGL((k_batch_pres_sweep3<T, true>), dim3(tab.total_ps3), dim3(PS3_THREADS), 0, s, tab, r);
        else
            hipLaunchKernelGGL((k_batch_pres_sweep3<T, false>), dim3(tab.total_ps3), dim3(PS3_THREADS), 0, s, tab, r);
        return;
    }
    if (k == KN_PRES_SWEEP4) {
        static_assert(std::is_same_v<T, float>, "ensembles run FP32");
        if (r.pin < 0)
            hipLaunchKernelGGL(k_batch_pres_sweep4<true>, dim3(tab.total_ps4), dim3(PS4_THREADS), 0, s, tab, r);
        else
            hipLaunchKernelGGL(k_batch_pres_sweep4<false>, dim3(tab.total_ps4), dim3(PS4_THREADS), 0, s, tab, r);
        return;
    }
    if (k == KN_VEL_SWEEP2) {
        // the ensemble runs FP32 (the only instantiation)
        static_assert(std::is_same_v<T, float>, "ensembles run FP32");
        hipLaunchKernelGGL(k_batch_vel_sweep2, dim3(tab.total_ps2), dim3(PS2_THREADS), 0, s, tab, r);
        return;
    }
    if (k == KN_PRES_SWEEP2) {
        if (r.pin < 0)
            hipLaunchKernelGGL((k_batch_pres_sweep2<T, true>), dim3(tab.total_ps2), dim3(PS2_THREADS), 0, s, tab, r);
        else
            hipLaunchKernelGGL((k_batch_pres_sweep2<T, false>), dim3(tab.total_ps2), dim3(PS2_THREADS), 0, s, tab, r);
        return;
    }
    if (k == KN_SUM_PARTIALS || k == KN_DEFECT_SCALE) {
        hipLaunchKernelGGL(k_batch_sum<T>, dim3(tab.ncases), dim3(TX, 4, 1), 0, s, tab, slot,
                           k == KN_DEFECT_SCALE ? 1 : 0);
        return;
    }
#define GOB(K) go_batch<T, K>(tab, grid, r, s, host_fields)
#define GORB(K) go_batch_reduce<T, K>(tab, grid, r, s, host_fields)
    UWS_KERNEL_CASES(GOB, GORB)
#undef GOB
#undef GORB
}

namespace {
template <class T>
__global__ __launch_bounds__(256) void k_mean_accumulate(const T* u0, const T* u1, const T* u2, double* s0,
                                                        double* s1, double* s2, int n) {
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        s0[i] += real_of(u0[i]);
        s1[i] += real_of(u1[i]);
        s2[i] += real_of(u2[i]);
    }
}
}  // namespace

int pres_sweep2_grid(int X, int Y, int Z, int zchunk, int by) { return pres_sweep2_blocks(X, Y, Z, zchunk, by); }
int pres_sweep3_grid(int X, int Y, int Z, int zchunk) { return pres_sweep3_blocks(X, Y, Z, zchunk); }
int pres_sweep4_grid(int X, int Y, int Z, int zchunk) { return pres_sweep4_blocks(X, Y, Z, zchunk); }

// ---------------------------------------------------------------------------
// CSInitVelPresFieldsFromBGFlowRotate (InitShaderColl.compute:184-226), the
// kernel FvmSolverGpu binds for InitVelPresFieldsFromBackGroundFlow
// (FvmSolverGpu.cs:335, 898-933): the cell's offset from the grid centre is
// rotated by the flow orientation into world coordinates, rotated back by
// the same angle into the background frame and sampled there with trilinear
// filtering and clamped addressing (texel centres at (i + 1/2) / N).  The
// sampled velocity vector itself is not rotated, as in the reference.  All
// velocity buffers get the sample, p^n the sampled pressure, p' and u' zero;
// flags are not looked at.
// ---------------------------------------------------------------------------
namespace {
template <class T, class R, class S>
__device__ __forceinline__ T sample_trilinear(const S* a, int stride, int X, int Y, int Z, R u, R v, R w) {
    const R tx = u * (R)X - (R)0.5f, ty = v * (R)Y - (R)0.5f, tz = w * (R)Z - (R)0.5f;
    const R fx0 = floor(tx), fy0 = floor(ty), fz0 = floor(tz);
    const R ax = tx - fx0, ay = ty - fy0, az = tz - fz0;
    const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    const int x0 = min(max(ix, 0), X - 1), x1 = min(max(ix + 1, 0), X - 1);
    const int y0 = min(max(iy, 0), Y - 1), y1 = min(max(iy + 1, 0), Y - 1);
    const int z0 = min(max(iz, 0), Z - 1), z1 = min(max(iz + 1, 0), Z - 1);
    auto at = [&](int x, int y, int z) { return (T)a[(((size_t)z * Y + y) * X + x) * (size_t)stride]; };
    const T bx = (T)((R)1 - ax), cx = (T)ax, by = (T)((R)1 - ay), cy = (T)ay, bz = (T)((R)1 - az), cz = (T)az;
    const T c00 = at(x0, y0, z0) * bx + at(x1, y0, z0) * cx;
    const T c10 = at(x0, y1, z0) * bx + at(x1, y1, z0) * cx;
    const T c01 = at(x0, y0, z1) * bx + at(x1, y0, z1) * cx;
    const T c11 = at(x0, y1, z1) * bx + at(x1, y1, z1) * cx;
    const T c0 = c00 * by + c10 * cy, c1 = c01 * by + c11 * cy;
    return c0 * bz + c1 * cz;
}

template <class T, class R, class S>
__global__ __launch_bounds__(256) void k_init_from_bg(ParamsT<T> p, FieldsT<T> f, BgField<T, R, S> bg) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= p.N) return;
    const int x = i % p.X, t = i / p.X, y = t % p.Y, z = t / p.Y;
    const R dxFG = (R)p.dx;  // the grid's dx in the geometry arithmetic
    R s = sin(bg.angle), c = cos(bg.angle);
    const R lx = (R)(x - p.X / 2) * dxFG, ly = (R)(y - p.Y / 2) * dxFG, lz = (R)(z - p.Z / 2) * dxFG;
    const R gx = bg.fgPos[0] + (lx * c + lz * s), gy = bg.fgPos[1] + ly, gz = bg.fgPos[2] + (lz * c - lx * s);
    s = sin(-bg.angle);
    c = cos(-bg.angle);
    const R rx = gx - bg.pos[0], ry = gy - bg.pos[1], rz = gz - bg.pos[2];
    const R bxp = rx * c + rz * s, byp = ry, bzp = rz * c - rx * s;
    const R u = bxp / ((R)bg.X * bg.dx) + (R)0.5f, v = byp / ((R)bg.Y * bg.dx) + (R)0.5f,
            w = bzp / ((R)bg.Z * bg.dx) + (R)0.5f;
    T vel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) vel[a] = sample_trilinear<T, R, S>(bg.vel[a], bg.vstride, bg.X, bg.Y, bg.Z, u, v, w);
    const T pr = bg.pres ? sample_trilinear<T, R, S>(bg.pres, 1, bg.X, bg.Y, bg.Z, u, v, w) : (T)0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) f.vel[b][a][i] = vel[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) f.ucorr[a][i] = 0;
    f.p[i] = pr;
    f.pc[0][i] = 0;
    f.pc[1][i] = 0;
}
}  // namespace

template <class T, class R, class S>
void init_from_background(const ParamsT<T>& p, const FieldsT<T>& f, const BgField<T, R, S>& bg, hipStream_t s) {
    hipLaunchKernelGGL((k_init_from_bg<T, R, S>), dim3((p.N + 255) / 256), dim3(256), 0, s, p, f, bg);
}
template void init_from_background<float, float, float>(const ParamsT<float>&, const FieldsT<float>&,
                                                        const BgField<float, float, float>&, hipStream_t);
template void init_from_background<double, double, double>(const ParamsT<double>&, const FieldsT<double>&,
                                                           const BgField<double, double, double>&, hipStream_t);
template void init_from_background<CStep, double, CStep>(const ParamsT<CStep>&, const FieldsT<CStep>&,
                                                         const BgField<CStep, double, CStep>&, hipStream_t);
template void init_from_background<double, double, float>(const ParamsT<double>&, const FieldsT<double>&,
                                                          const BgField<double, double, float>&, hipStream_t);
template void init_from_background<CStep, double, float>(const ParamsT<CStep>&, const FieldsT<CStep>&,
                                                         const BgField<CStep, double, float>&, hipStream_t);

// ---------------------------------------------------------------------------
// Staging array of the allocation-free warm start
// (uws_init_vel_pres_fields_from_background_noalloc): the caller's
// background arrays are copied here as they are instead of into a buffer
// allocated per call.  Each GPU holds its own copy of the array until the
// process ends; host code never uses the array's name as an address, only
// the per-device address hipGetSymbolAddress returns.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[UWS_BG_STAGING_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_staging_mu;
void* g_staging[kMaxDevices] = {};
}  // namespace

int bg_staging_setup() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return UWS_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_staging_mu);
    if (g_staging[dev]) return UWS_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return UWS_ERR_DEVICE;
    g_staging[dev] = ptr;
    return UWS_OK;
}

void* bg_staging(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_staging_mu);
    return g_staging[device];
}

// ---------------------------------------------------------------------------
// Streaming reference for the bandwidth ceiling: out[i] = sum_k in_k[i] over
// nin float arrays and one written array, 16-byte loads and stores, a
// grid-stride loop over a fixed grid of 16 workgroups per CU.  Launched
// back to back on the same arrays, as the Jacobi sweeps re-read theirs, so a
// working set that fits the last-level cache is served from it.
// ---------------------------------------------------------------------------
namespace {
struct StreamArgs {
    const float4* in[16];
    float4* out;
    int nin;
    int64_t n4;
};
// each thread keeps four 16-byte accesses per array in flight (four
// workgroup-wide rows of 256 vectors per iteration)
__global__ __launch_bounds__(256) void k_stream(StreamArgs a) {
    constexpr int U = 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * U;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x * U + threadIdx.x; i0 < a.n4; i0 += stride) {
        float4 acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            acc[u] = i < a.n4 ? a.in[0][i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (int k = 1; k < a.nin; ++k) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = i0 + (int64_t)u * blockDim.x;
                if (i < a.n4) {
                    const float4 v = a.in[k][i];
                    acc[u].x += v.x; acc[u].y += v.y; acc[u].z += v.z; acc[u].w += v.w;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            if (i < a.n4) a.out[i] = acc[u];
        }
    }
}
}  // namespace

int stream_bandwidth(int64_t cells, int nin, int reps, double* gbps) {
    if (cells < 1024 || nin < 1 || nin > 16 || reps < 1) return UWS_ERR_INVALID_ARG;
    const int64_t n4 = (cells + 3) / 4;
    const size_t plane = (size_t)n4 * sizeof(float4);
    void* mem = nullptr;
    if (hipMalloc(&mem, plane * (size_t)(nin + 1)) != hipSuccess) return UWS_ERR_ALLOC;
    StreamArgs a{};
    for (int k = 0; k < nin; ++k) a.in[k] = reinterpret_cast<const float4*>(static_cast<char*>(mem) + plane * k);
    a.out = reinterpret_cast<float4*>(static_cast<char*>(mem) + plane * nin);
    a.nin = nin;
    a.n4 = n4;
    int rc = UWS_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        rc = UWS_ERR_DEVICE;
    const int grid = (int)std::min<int64_t>((n4 + 1023) / 1024, (int64_t)cus * 16);
    if (rc == UWS_OK && (hipMemsetAsync(mem, 0, plane * (size_t)(nin + 1), nullptr) != hipSuccess ||
                         hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = UWS_ERR_DEVICE;
    if (rc == UWS_OK) {
        hipLaunchKernelGGL(k_stream, dim3(grid), dim3(256), 0, nullptr, a);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream, dim3(grid), dim3(256), 0, nullptr, a);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f)
            rc = UWS_ERR_DEVICE;
        else
            *gbps = (double)plane * (nin + 1) * reps / ((double)ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipDeviceSynchronize();
    (void)hipFree(mem);
    return rc;
}

template <class T>
void mean_accumulate(T* const u[3], double* const sum[3], int n, hipStream_t s) {
    const int grid = std::min((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_mean_accumulate<T>, dim3(grid), dim3(256), 0, s, u[0], u[1], u[2], sum[0], sum[1], sum[2],
                       n);
}
template void mean_accumulate<float>(float* const[3], double* const[3], int, hipStream_t);
template void mean_accumulate<double>(double* const[3], double* const[3], int, hipStream_t);
template void mean_accumulate<CStep>(CStep* const[3], double* const[3], int, hipStream_t);

template void launch<float>(Kernel, const ParamsT<float>&, const FieldsT<float>&, const Roles&, hipStream_t, int);
template void launch<double>(Kernel, const ParamsT<double>&, const FieldsT<double>&, const Roles&, hipStream_t,
                             int);
template void launch<CStep>(Kernel, const ParamsT<CStep>&, const FieldsT<CStep>&, const Roles&, hipStream_t, int);
template void launch<Dual<float>>(Kernel, const ParamsT<Dual<float>>&, const FieldsT<Dual<float>>&, const Roles&,
                                  hipStream_t, int);
template void launch<Dual<double>>(Kernel, const ParamsT<Dual<double>>&, const FieldsT<Dual<double>>&,
                                   const Roles&, hipStream_t, int);
template void launch_batch<float>(Kernel, const BatchTable<float>&, int, const Roles&, hipStream_t, int,
                                   const FieldsT<float>*);

}  // namespace uws
