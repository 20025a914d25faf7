// Discrete adjoint of the solver step: argument blocks of the adjoint
// kernels (uws_adjoint.hip) and their host-side launchers.
//
// The step is the composition of
//  - linear Jacobi sweeps with frozen coefficients (the momentum predictor and
//    every pressure solve), reversed by hand-written transposed sweeps that
//    also accumulate the adjoints of the coefficients, and
//  - one-shot local kernels (wall function + eddy viscosity + LES term +
//    momentum coefficients, Rhie-Chow coefficients, corrections, PISO
//    coefficients).  The output of such a kernel at cell o depends only on its
//    inputs in a small stencil around o, so its transposed Jacobian is built
//    exactly from forward-mode runs of the same kernel body in Dual<T>: the
//    input cells are coloured so that no two cells of one colour reach the
//    same output, each colour is seeded at once, the outputs' derivatives are
//    contracted with the output adjoints (t_o = sum_k lam_k(o) dout_k(o)),
//    and every seeded cell gathers t over the outputs it reaches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "uws_internal.h"

namespace uws {

// The forward's records are in the run's type T; every adjoint quantity is
// in A (double: an FP32 run keeps FP32 records but accumulates its adjoint in
// FP64).
//
// Transposed momentum sweep.  The forward sweep is, for a fluid cell P and
// component a,  out_a(P) = invD_a(P) (sum_nb c_nb u_a(nb) + b_a(P)),  with the
// face coefficients built from S and G of the owned E/N/U faces.
template <class T, class A>
struct AdjVelSweep {
    int X, Y, Z;
    A q;                          // ds / 4
    const uint8_t* flag;
    const T* S[3];
    const T* G[3];
    const T* invD[3];
    const T* b[3];
    const T* in[3];               // the sweep's input iterate
    const A* lam_out[3];          // adjoint of the sweep's output
    A* lam_in[3];                 // adjoint of the input iterate (written)
    A* lam_S[3];                  // accumulated
    A* lam_G[3];
    A* lam_invD[3];
    A* lam_b[3];
};

// Transposed pressure sweep: out(P) = invDp(P) (-sum_f c_f x(nb) + rhs(P)) on
// fluid cells, 0 elsewhere; in == nullptr is the zero first iterate.
template <class T, class A>
struct AdjPresSweep {
    int X, Y, Z;
    const uint8_t* flag;
    const T* cE;
    const T* cN;
    const T* cU;
    const T* invDp;
    const T* rhs;
    const T* in;
    const A* lam_out;
    A* lam_in;                    // written when in != nullptr
    A* lam_cE;                    // accumulated
    A* lam_cN;
    A* lam_cU;
    A* lam_invDp;
    A* lam_rhs;
};

// colour(x, y, z) = (x + wy*y + wz*z) mod m; m == 1 is a single colour
struct Colour {
    int m, wy, wz;
};
// offsets from an input cell to the outputs it reaches (symmetric sets)
struct Stencil {
    int n;
    int8_t off[25][3];
};
enum StencilKind { ST_OWN = 0, ST_STAR7 = 1, ST_DIAMOND25 = 2, ST_AXIS13 = 3 };
Stencil make_stencil(int kind);
Colour colour_of(int kind);        // cell colourings of the stencils above
Colour colour_bnd(int kind);       // boundary-texture colourings (2-D)

template <class T>
struct Pairs {
    int n;
    const T* lam[16];
    const Dual<T>* d[16];
};

template <class T, class A> void adj_vel_sweep(const AdjVelSweep<T, A>& a, int N, hipStream_t s);
template <class T, class A> void adj_pres_sweep(const AdjPresSweep<T, A>& a, int N, hipStream_t s);
// dst[i] = ((A)src[i] or 0, 0)
template <class T, class A> void dual_pack(const T* src, Dual<A>* dst, int n, hipStream_t s);
// dst[i].d = 1 on the cells of colour k, 0 elsewhere (k < 0: 0 everywhere);
// the array is a W x H x D box in x-fastest order
template <class T> void dual_seed(Dual<T>* dst, int W, int H, int D, Colour c, int k, hipStream_t s);
template <class T> void dual_contract(const Pairs<T>& pr, T* t, int n, hipStream_t s);
// dst[i] += sum over the stencil of t, on the cells of colour k
template <class T>
void adj_gather(const T* t, int X, int Y, int Z, const Stencil& st, Colour c, int k, T* dst, hipStream_t s);
// boundary texture of `face` (W x H elements): element (i, j) sits next to
// its boundary cell; dst(i, j) += sum over the stencil around that cell of t
template <class T>
void adj_gather_bnd(const T* t, int X, int Y, int Z, int face, const Stencil& st, Colour c, int k, T* dst,
                    hipStream_t s);
// out[i] = (a ? a[i] : 0) + (b and (flag == nullptr or flag[i] is fluid) ? b[i] : 0)
template <class T>
void adj_add_fluid(const uint8_t* flag, const T* a, const T* b, T* out, int n, hipStream_t s);
// acc[0] += sum_i t[i] (one workgroup, fixed order)
template <class T> void adj_sum(const T* t, int n, double* acc, hipStream_t s);

}  // namespace uws
