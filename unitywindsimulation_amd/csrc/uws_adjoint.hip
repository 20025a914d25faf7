// Kernels of the discrete adjoint (see uws_adjoint.h).  Every kernel is one
// thread per cell (or boundary element) in gather form: each thread writes
// only its own cell, so the adjoint is deterministic without atomics.
#include <algorithm>

#include "uws_adjoint.h"

namespace uws {
namespace {

constexpr int BLK = 256;
inline int blocks(int n) { return (n + BLK - 1) / BLK; }

__device__ __forceinline__ int gid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
// Stencil kernels run 64 x 4 tiles of one z-plane per workgroup (one wave
// per x-row), their ids remapped so that XCD j (ids j, j+8, ...) walks a
// contiguous range of tiles and the neighbouring rows and planes a
// workgroup reads stay in that XCD's L2.  Returns false outside the grid.
constexpr int TW = 64, TH = 4;
inline int tiles2d(int X, int Y, int Z) { return ((X + TW - 1) / TW) * ((Y + TH - 1) / TH) * Z; }
__device__ __forceinline__ bool cell2d(int X, int Y, int& x, int& y, int& z) {
    const int B = (int)gridDim.x, q = B >> 3, r = B & 7;
    const int b = (int)blockIdx.x, xcd = b & 7;
    const int t = xcd * q + min(xcd, r) + (b >> 3);
    const int gx = (X + TW - 1) / TW, gy = (Y + TH - 1) / TH;
    const int bx = t % gx, rr = t / gx, by = rr % gy;
    z = rr / gy;
    x = bx * TW + (int)threadIdx.x;
    y = by * TH + (int)threadIdx.y;
    return x < X && y < Y;
}

// ---------------------------------------------------------------------------
// Transposed momentum sweep, in two parts.  k_adj_vel_lam: for cell Q and
// component a,
//   lam_in(Q) = sum over the cells P that read Q of dout(P)/du(Q) lam_out(P)
// with dout(P)/du(E of P) = invD(P) (-q S_E(P) + G_E(P)) and
// dout(P)/du(W of P) = invD(P) (q S_E(W) + G_E(W)) (and alike in y, z);
// coefficients of faces that are not fluid-fluid are zero, and invD is zero
// off the fluid, so no flags are needed for this part.  k_adj_vel_coef: the
// adjoints of S, G, 1/D and b summed over a window of sweeps in registers
// and added once.
// ---------------------------------------------------------------------------
template <class T, class A>
__global__ __launch_bounds__(BLK) void k_adj_vel_lam(AdjVelSweep<T, A> a, int N) {
    const int X = a.X, Y = a.Y, Z = a.Z;
    int x, y, z;
    if (!cell2d(X, Y, x, y, z)) return;
    const int i = (z * Y + y) * X + x;
    const int sy = X, sz = X * Y;
    const bool hW = x > 0, hE = x + 1 < X, hS = y > 0, hN = y + 1 < Y, hD = z > 0, hU = z + 1 < Z;
    const A q = a.q;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const A* lo = a.lam_out[c];
        const T* iv = a.invD[c];
        A acc = 0;
        if (hW) acc += lo[i - 1] * (A)iv[i - 1] * (-q * (A)a.S[0][i - 1] + (A)a.G[0][i - 1]);
        if (hE) acc += lo[i + 1] * (A)iv[i + 1] * (q * (A)a.S[0][i] + (A)a.G[0][i]);
        if (hS) acc += lo[i - sy] * (A)iv[i - sy] * (-q * (A)a.S[1][i - sy] + (A)a.G[1][i - sy]);
        if (hN) acc += lo[i + sy] * (A)iv[i + sy] * (q * (A)a.S[1][i] + (A)a.G[1][i]);
        if (hD) acc += lo[i - sz] * (A)iv[i - sz] * (-q * (A)a.S[2][i - sz] + (A)a.G[2][i - sz]);
        if (hU) acc += lo[i + sz] * (A)iv[i + sz] * (q * (A)a.S[2][i] + (A)a.G[2][i]);
        a.lam_in[c][i] = acc;
    }
}

template <class T, class A>
__global__ __launch_bounds__(BLK) void k_adj_vel_coef(AdjVelSweep<T, A> a, SweepWindow<T, A> w, int N) {
    const int X = a.X, Y = a.Y, Z = a.Z;
    int x, y, z;
    if (!cell2d(X, Y, x, y, z)) return;
    const int i = (z * Y + y) * X + x;
    const int sy = X, sz = X * Y;
    const bool hW = x > 0, hE = x + 1 < X, hS = y > 0, hN = y + 1 < Y, hD = z > 0, hU = z + 1 < Z;
    const bool isF = a.flag[i] == UWS_CELL_FLUID;
    const A q = a.q;
    const int iE = hE ? i + 1 : i, iW = hW ? i - 1 : i, iN = hN ? i + sy : i, iS = hS ? i - sy : i;
    const int iU = hU ? i + sz : i, iD = hD ? i - sz : i;
    const A SE = a.S[0][i], SN = a.S[1][i], SU = a.S[2][i];
    const A GE = a.G[0][i], GN = a.G[1][i], GU = a.G[2][i];
    const A SW = hW ? (A)a.S[0][iW] : (A)0, GW = hW ? (A)a.G[0][iW] : (A)0;
    const A SS = hS ? (A)a.S[1][iS] : (A)0, GS = hS ? (A)a.G[1][iS] : (A)0;
    const A SD = hD ? (A)a.S[2][iD] : (A)0, GD = hD ? (A)a.G[2][iD] : (A)0;
    const int offs[3] = {1, sy, sz};
    const bool has[3] = {hE, hN, hU};
    A ivP[3], ivQ[3][3], bP[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ivP[c] = a.invD[c][i];
        bP[c] = a.b[c][i];
#pragma unroll
        for (int e = 0; e < 3; ++e) ivQ[e][c] = has[e] ? (A)a.invD[c][i + offs[e]] : (A)0;
    }
    A lInv[3] = {0, 0, 0}, lB[3] = {0, 0, 0}, lS[3] = {0, 0, 0}, lG[3] = {0, 0, 0};
#pragma unroll 4
    for (int j = 0; j < w.n; ++j) {
        A rP[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const A lo = w.lo[j][c][i];
            rP[c] = lo * ivP[c];
            if (isF) {
                // the bracket A + Df + b of the forward sweep, on its input iterate
                const T* k = w.in[j][c];
                const A vE = k[iE], vW = k[iW], vN = k[iN], vS = k[iS], vU = k[iU], vD = k[iD];
                A Av = 0;
                Av -= q * vE * SE;
                Av += q * vW * SW;
                Av -= q * vN * SN;
                Av += q * vS * SS;
                Av -= q * vU * SU;
                Av += q * vD * SD;
                A Df = 0;
                Df += GE * vE;
                Df += GW * vW;
                Df += GN * vN;
                Df += GS * vS;
                Df += GU * vU;
                Df += GD * vD;
                lInv[c] += lo * (Av + Df + bP[c]);
                lB[c] += rP[c];
            }
        }
        // owned faces E, N, U: S and G enter out(P) and out(P + e)
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (!has[e]) continue;
            const int jq = i + offs[e];
            A sS = 0, sG = 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const A rQ = w.lo[j][c][jq] * ivQ[e][c];
                const A uP = w.in[j][c][i], uQ = w.in[j][c][jq];
                sS += rP[c] * (-q * uQ) + rQ * (q * uP);
                sG += rP[c] * uQ + rQ * uP;
            }
            lS[e] += sS;
            lG[e] += sG;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (isF) {
            a.lam_invD[c][i] += lInv[c];
            a.lam_b[c][i] += lB[c];
        }
        if (has[c]) {
            a.lam_S[c][i] += lS[c];
            a.lam_G[c][i] += lG[c];
        }
    }
}

// ---------------------------------------------------------------------------
// Transposed pressure sweep, in the same two parts.  out(P) = invDp(P)
// (term(P) + rhs(P)) with term(P) = -sum_f c_f x(nb); c_f of a face is
// stored on the cell owning it.
// ---------------------------------------------------------------------------
template <class T, class A>
__global__ __launch_bounds__(BLK) void k_adj_pres_lam(AdjPresSweep<T, A> a, int N) {
    const int X = a.X, Y = a.Y, Z = a.Z;
    int x, y, z;
    if (!cell2d(X, Y, x, y, z)) return;
    const int i = (z * Y + y) * X + x;
    const int sy = X, sz = X * Y;
    const bool hW = x > 0, hE = x + 1 < X, hS = y > 0, hN = y + 1 < Y, hD = z > 0, hU = z + 1 < Z;
    const A* lo = a.lam_out;
    const T* iv = a.invDp;
    A acc = 0;
    if (hW) acc -= lo[i - 1] * (A)iv[i - 1] * (A)a.cE[i - 1];
    if (hE) acc -= lo[i + 1] * (A)iv[i + 1] * (A)a.cE[i];
    if (hS) acc -= lo[i - sy] * (A)iv[i - sy] * (A)a.cN[i - sy];
    if (hN) acc -= lo[i + sy] * (A)iv[i + sy] * (A)a.cN[i];
    if (hD) acc -= lo[i - sz] * (A)iv[i - sz] * (A)a.cU[i - sz];
    if (hU) acc -= lo[i + sz] * (A)iv[i + sz] * (A)a.cU[i];
    a.lam_in[i] = acc;
}

template <class T, class A>
__global__ __launch_bounds__(BLK) void k_adj_pres_coef(AdjPresSweep<T, A> a, SweepWindow<T, A> w, int N) {
    const int X = a.X, Y = a.Y, Z = a.Z;
    int x, y, z;
    if (!cell2d(X, Y, x, y, z)) return;
    const int i = (z * Y + y) * X + x;
    const int sy = X, sz = X * Y;
    const bool hW = x > 0, hE = x + 1 < X, hS = y > 0, hN = y + 1 < Y, hD = z > 0, hU = z + 1 < Z;
    const bool isF = a.flag[i] == UWS_CELL_FLUID;
    const int iE = hE ? i + 1 : i, iW = hW ? i - 1 : i, iN = hN ? i + sy : i, iS = hS ? i - sy : i;
    const int iU = hU ? i + sz : i, iD = hD ? i - sz : i;
    const A cE = a.cE[i], cN = a.cN[i], cU = a.cU[i];
    const A cW = hW ? (A)a.cE[iW] : (A)0, cS = hS ? (A)a.cN[iS] : (A)0, cD = hD ? (A)a.cU[iD] : (A)0;
    const A ivP = a.invDp[i], rhs = a.rhs[i];
    const A ivE = hE ? (A)a.invDp[iE] : (A)0, ivN = hN ? (A)a.invDp[iN] : (A)0, ivU = hU ? (A)a.invDp[iU] : (A)0;
    A lInv = 0, lRhs = 0, lE = 0, lN = 0, lU = 0;
#pragma unroll 4
    for (int j = 0; j < w.n; ++j) {
        const A* lo = w.lo[j][0];
        const T* xin = w.in[j][0];
        const A l = lo[i], rP = l * ivP;
        if (isF) {
            A term = 0;
            if (xin) {
                term -= cE * (A)xin[iE];
                term -= cW * (A)xin[iW];
                term -= cN * (A)xin[iN];
                term -= cS * (A)xin[iS];
                term -= cU * (A)xin[iU];
                term -= cD * (A)xin[iD];
            }
            lInv += l * (term + rhs);
            lRhs += rP;
        }
        if (xin) {
            const A xP = xin[i];
            if (hE) lE -= rP * (A)xin[iE] + lo[iE] * ivE * xP;
            if (hN) lN -= rP * (A)xin[iN] + lo[iN] * ivN * xP;
            if (hU) lU -= rP * (A)xin[iU] + lo[iU] * ivU * xP;
        }
    }
    if (isF) {
        a.lam_invDp[i] += lInv;
        a.lam_rhs[i] += lRhs;
    }
    if (hE) a.lam_cE[i] += lE;
    if (hN) a.lam_cN[i] += lN;
    if (hU) a.lam_cU[i] += lU;
}

// ---------------------------------------------------------------------------
// forward-mode machinery
// ---------------------------------------------------------------------------
template <class T, class A>
__global__ __launch_bounds__(BLK) void k_dual_pack(const T* src, Dual<A>* dst, int n) {
    const int i = gid();
    if (i < n) dst[i] = Dual<A>(src ? (A)src[i] : (A)0, (A)0);
}

__device__ __forceinline__ int colour_at(const Colour& c, int x, int y, int z) {
    return c.m == 1 ? 0 : (x + c.wy * y + c.wz * z) % c.m;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_dual_seed(Dual<T>* dst, int W, int H, int D, Colour c, int k) {
    const int i = gid();
    if (i >= W * H * D) return;
    const int x = i % W, t = i / W, y = t % H, z = t / H;
    dst[i].d = (k >= 0 && colour_at(c, x, y, z) == k) ? (T)1 : (T)0;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_dual_contract(Pairs<T> pr, T* t, int n) {
    const int i = gid();
    if (i >= n) return;
    T s = 0;
    for (int j = 0; j < pr.n; ++j) s += pr.lam[j][i] * pr.d[j][i].d;
    t[i] = s;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_adj_gather(const T* t, int X, int Y, int Z, Stencil st, Colour c, int k,
                                                   T* dst) {
    const int i = gid();
    if (i >= X * Y * Z) return;
    const int x = i % X, r = i / X, y = r % Y, z = r / Y;
    if (colour_at(c, x, y, z) != k) return;
    T s = 0;
    for (int j = 0; j < st.n; ++j) {
        const int nx = x + st.off[j][0], ny = y + st.off[j][1], nz = z + st.off[j][2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= X || ny >= Y || nz >= Z) continue;
        s += t[(nz * Y + ny) * X + nx];
    }
    dst[i] += s;
}

// element (i, j) of a boundary texture and the cell next to it: X faces are
// indexed (y, z), Y faces (x, z), Z faces (x, y)
__device__ __forceinline__ void bnd_cell(int face, int X, int Y, int Z, int a, int b, int* x, int* y, int* z) {
    switch (face) {
    case UWS_X0: *x = 0; *y = a; *z = b; break;
    case UWS_XN: *x = X - 1; *y = a; *z = b; break;
    case UWS_Y0: *x = a; *y = 0; *z = b; break;
    case UWS_YN: *x = a; *y = Y - 1; *z = b; break;
    case UWS_Z0: *x = a; *y = b; *z = 0; break;
    default: *x = a; *y = b; *z = Z - 1; break;
    }
}

template <class T>
__global__ __launch_bounds__(BLK) void k_adj_gather_bnd(const T* t, int X, int Y, int Z, int face, int W, int H,
                                                       Stencil st, Colour c, int k, T* dst) {
    const int e = gid();
    if (e >= W * H) return;
    const int a = e % W, b = e / W;
    if (colour_at(c, a, b, 0) != k) return;
    int x, y, z;
    bnd_cell(face, X, Y, Z, a, b, &x, &y, &z);
    T s = 0;
    for (int j = 0; j < st.n; ++j) {
        const int nx = x + st.off[j][0], ny = y + st.off[j][1], nz = z + st.off[j][2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= X || ny >= Y || nz >= Z) continue;
        s += t[(nz * Y + ny) * X + nx];
    }
    dst[e] += s;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_adj_add_fluid(const uint8_t* flag, const T* a, const T* b, T* out, int n) {
    const int i = gid();
    if (i >= n) return;
    T v = a ? a[i] : (T)0;
    if (b && (!flag || flag[i] == UWS_CELL_FLUID)) v += b[i];
    out[i] = v;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_adj_zero(ZeroList<T> z, int n) {
    const int i = gid();
    if (i < n) z.p[blockIdx.y][i] = (T)0;
}

// fixed-order sum: workgroup b of nb sums t[b*256 + j*nb*256 + lane] into
// part[b], then one workgroup sums part[0..nb) and adds the total to acc
template <class T>
__global__ __launch_bounds__(BLK) void k_adj_sum_part(const T* t, int n, double* part) {
    __shared__ double red[BLK];
    double s = 0.0;
    const int stride = (int)gridDim.x * BLK;
    for (int i = (int)blockIdx.x * BLK + (int)threadIdx.x; i < n; i += stride) s += (double)t[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(ADJ_SUM_PARTS) void k_adj_sum_final(const double* part, int nb, double* acc) {
    __shared__ double red[ADJ_SUM_PARTS];
    red[threadIdx.x] = (int)threadIdx.x < nb ? part[threadIdx.x] : 0.0;
    __syncthreads();
    for (int w = ADJ_SUM_PARTS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0];
}

}  // namespace

Stencil make_stencil(int kind) {
    Stencil s{};
    s.n = 0;
    for (int z = -2; z <= 2; ++z)
        for (int y = -2; y <= 2; ++y)
            for (int x = -2; x <= 2; ++x) {
                const int l1 = (x < 0 ? -x : x) + (y < 0 ? -y : y) + (z < 0 ? -z : z);
                const int axes = (x != 0) + (y != 0) + (z != 0);
                bool on = false;
                if (kind == ST_OWN) on = l1 == 0;
                else if (kind == ST_STAR7) on = l1 <= 1;
                else if (kind == ST_DIAMOND25) on = l1 <= 2;
                else on = axes <= 1 && l1 <= 2;  // ST_AXIS13
                if (!on) continue;
                s.off[s.n][0] = (int8_t)x;
                s.off[s.n][1] = (int8_t)y;
                s.off[s.n][2] = (int8_t)z;
                ++s.n;
            }
    return s;
}

// Linear colourings x + wy*y + wz*z (mod m) under which any two distinct
// cells whose stencils overlap (difference in S - S) get different colours;
// m is the smallest modulus with that property for the set (found by
// exhaustive search over m, wy, wz): 7 is the perfect Lee code of the
// 7-point star.
Colour colour_of(int kind) {
    switch (kind) {
    case ST_OWN: return Colour{1, 0, 0};
    case ST_STAR7: return Colour{7, 2, 3};
    case ST_DIAMOND25: return Colour{27, 4, 10};
    default: return Colour{13, 3, 4};  // ST_AXIS13
    }
}
// boundary elements reach the cells of a stencil around their boundary cell;
// for the 7-point star two elements interfere within L1 distance 2 in the
// texture plane: (i + 2j) mod 5 separates them
Colour colour_bnd(int kind) { return kind == ST_OWN ? Colour{1, 0, 0} : Colour{5, 2, 0}; }

template <class T, class A>
void adj_vel_lam(const AdjVelSweep<T, A>& a, int N, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_vel_lam<T, A>), dim3(tiles2d(a.X, a.Y, a.Z)), dim3(TW, TH), 0, s, a, N);
}
template <class T, class A>
void adj_vel_coef(const AdjVelSweep<T, A>& a, const SweepWindow<T, A>& w, int N, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_vel_coef<T, A>), dim3(tiles2d(a.X, a.Y, a.Z)), dim3(TW, TH), 0, s, a, w, N);
}
template <class T, class A>
void adj_pres_lam(const AdjPresSweep<T, A>& a, int N, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_pres_lam<T, A>), dim3(tiles2d(a.X, a.Y, a.Z)), dim3(TW, TH), 0, s, a, N);
}
template <class T, class A>
void adj_pres_coef(const AdjPresSweep<T, A>& a, const SweepWindow<T, A>& w, int N, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_pres_coef<T, A>), dim3(tiles2d(a.X, a.Y, a.Z)), dim3(TW, TH), 0, s, a, w, N);
}
template <class T, class A>
void dual_pack(const T* src, Dual<A>* dst, int n, hipStream_t s) {
    hipLaunchKernelGGL((k_dual_pack<T, A>), dim3(blocks(n)), dim3(BLK), 0, s, src, dst, n);
}
template <class T>
void dual_seed(Dual<T>* dst, int W, int H, int D, Colour c, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_dual_seed<T>, dim3(blocks(W * H * D)), dim3(BLK), 0, s, dst, W, H, D, c, k);
}
template <class T>
void dual_contract(const Pairs<T>& pr, T* t, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_dual_contract<T>, dim3(blocks(n)), dim3(BLK), 0, s, pr, t, n);
}
template <class T>
void adj_gather(const T* t, int X, int Y, int Z, const Stencil& st, Colour c, int k, T* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_gather<T>, dim3(blocks(X * Y * Z)), dim3(BLK), 0, s, t, X, Y, Z, st, c, k, dst);
}
template <class T>
void adj_gather_bnd(const T* t, int X, int Y, int Z, int face, const Stencil& st, Colour c, int k, T* dst,
                    hipStream_t s) {
    const int W = face < 2 ? Y : X, H = face < 4 ? Z : Y;
    hipLaunchKernelGGL(k_adj_gather_bnd<T>, dim3(blocks(W * H)), dim3(BLK), 0, s, t, X, Y, Z, face, W, H, st, c, k,
                       dst);
}
template <class T>
void adj_add_fluid(const uint8_t* flag, const T* a, const T* b, T* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_add_fluid<T>, dim3(blocks(n)), dim3(BLK), 0, s, flag, a, b, out, n);
}
template <class T>
void adj_zero(const ZeroList<T>& z, int n, hipStream_t s) {
    if (z.m > 0) hipLaunchKernelGGL(k_adj_zero<T>, dim3(blocks(n), z.m), dim3(BLK), 0, s, z, n);
}
template <class T>
void adj_sum(const T* t, int n, double* acc, double* part, hipStream_t s) {
    const int nb = std::max(1, std::min(ADJ_SUM_PARTS, (n + 8 * BLK - 1) / (8 * BLK)));
    hipLaunchKernelGGL(k_adj_sum_part<T>, dim3(nb), dim3(BLK), 0, s, t, n, part);
    hipLaunchKernelGGL(k_adj_sum_final, dim3(1), dim3(ADJ_SUM_PARTS), 0, s, (const double*)part, nb, acc);
}

// records in float or double, adjoint in double
#define UWS_ADJ_SWEEPS(T)                                                                                  \
    template void adj_vel_lam<T, double>(const AdjVelSweep<T, double>&, int, hipStream_t);                \
    template void adj_vel_coef<T, double>(const AdjVelSweep<T, double>&, const SweepWindow<T, double>&, int, \
                                          hipStream_t);                                                    \
    template void adj_pres_lam<T, double>(const AdjPresSweep<T, double>&, int, hipStream_t);              \
    template void adj_pres_coef<T, double>(const AdjPresSweep<T, double>&, const SweepWindow<T, double>&,   \
                                           int, hipStream_t);
UWS_ADJ_SWEEPS(float)
UWS_ADJ_SWEEPS(double)
#undef UWS_ADJ_SWEEPS
template void dual_pack<float, double>(const float*, Dual<double>*, int, hipStream_t);
template void dual_pack<double, double>(const double*, Dual<double>*, int, hipStream_t);
template void dual_seed<double>(Dual<double>*, int, int, int, Colour, int, hipStream_t);
template void dual_contract<double>(const Pairs<double>&, double*, int, hipStream_t);
template void adj_gather<double>(const double*, int, int, int, const Stencil&, Colour, int, double*, hipStream_t);
template void adj_gather_bnd<double>(const double*, int, int, int, int, const Stencil&, Colour, int, double*,
                                     hipStream_t);
template void adj_add_fluid<double>(const uint8_t*, const double*, const double*, double*, int, hipStream_t);
template void adj_sum<double>(const double*, int, double*, double*, hipStream_t);
template void adj_zero<double>(const ZeroList<double>&, int, hipStream_t);

}  // namespace uws
