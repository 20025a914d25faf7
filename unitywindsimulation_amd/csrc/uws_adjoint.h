// Discrete adjoint of the solver step: argument blocks of the adjoint
// kernels (uws_adjoint.hip) and their host-side launchers.
//
// The step is the composition of
//  - linear Jacobi sweeps with frozen coefficients (the momentum predictor and
//    every pressure solve), reversed by hand-written transposed sweeps that
//    also accumulate the adjoints of the coefficients, and
//  - one-shot local kernels (wall function + eddy viscosity + LES term +
//    momentum coefficients, Rhie-Chow coefficients, corrections, PISO
//    coefficients), reversed by hand-written gather kernels.  The output of
//    such a kernel at cell o depends only on its inputs in a small stencil
//    around o, so its transposed Jacobian can also be built exactly from
//    forward-mode runs of the same kernel body in Dual<T> (the colouring path
//    the hand-written reverses are tested against): the
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
    const A* lam_out[3];          // adj_vel_lam: adjoint of the sweep's output
    A* lam_in[3];                 // adj_vel_lam: adjoint of the input iterate (written)
    A* lam_S[3];                  // accumulated by adj_vel_coef
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
    const A* lam_out;             // adj_pres_lam's input
    A* lam_in;                    // adj_pres_lam's output
    A* lam_cE;                    // accumulated by adj_pres_coef
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

// Reverse of the momentum coefficients by hand (kernels in uws_adjoint_op1.h):
// the adjoints of S, G, D, 1/D and b go to u^n, p^n, the boundary textures and
// (per cell, to be summed) Cs.  The primal values are read from the Dual
// mirror after one forward run with zero tangents.
template <class A>
struct Op1Args {
    int X, Y, Z, N, slots;
    bool les;
    // adjoints of the outputs
    const A* lamS[3];
    const A* lamG[3];
    const A* lamD[3];
    const A* lamInvD[3];
    const A* lamB[3];
    // per-cell intermediate adjoints (written by k_op1_local, completed by k_op1_grad)
    A* lvis[6];     // face viscosities
    A* lfg[18];     // LES face gradients [face][component]
    A* ladv[6];     // advection weight of the neighbour across the face
    A* lpf[6];      // face pressure
    A* lu[3];       // own u
    A* lp;          // own p
    A* lg[9];       // cell gradient (own part, then the total)
    A* lnut;        // eddy viscosity (own faces)
    A* lcs;         // per-cell share of dJ/dCs
    // accumulated results
    A* lamU[3];
    A* lamP;
    A* lamBnd[6][3];
};
template <class A>
void adj_op1(const ParamsT<Dual<A>>& p, const FieldsT<Dual<A>>& f, const Op1Args<A>& a, hipStream_t s);

// Reverses by hand of the remaining one-shot kernels (uws_adjoint_ops.h), in
// gather form: one thread per cell R accumulates the adjoints of R's inputs
// from R's own outputs and from the outputs of the neighbours that read R.
// Primal records are read in T and used in A; parameters come from the Dual
// parameter block (.v parts).
//
// Velocity correction: SIMPLE u' = -(dv/den) (1/D) (pf+ - pf-) / dx (coef = D)
// or PISO u' = -(1/D) (aodu + (dv/den) (pf+ - pf-) / dx) (coef = 1/D), with
// the face values pf of p'.
template <class T, class A>
struct AdjCorrect {
    bool piso;
    const uint8_t* flag;
    const T* pc;          // p' the correction reads
    const T* coef[3];     // D (SIMPLE) or 1/D (PISO)
    const T* ao[3];       // aodu (PISO)
    const A* lam_uc[3];   // adjoint of u'
    A* lam_pc;            // accumulated
    A* lam_coef[3];
    A* lam_ao[3];         // PISO
};
// PISO coefficients: invDpiso = 1 / presCoeff, bpiso = velTerm
template <class T, class A>
struct AdjPisoCoef {
    const uint8_t* flag;
    const T* invD[3];
    const T* ao[3];
    const T* cpe;
    const T* cp[3];       // [1], [2]: the N and U face coefficients
    const A* lam_idp;     // adjoint of invDpiso
    const A* lam_bp;      // adjoint of bpiso
    A* lcf;               // scratch: d invDpiso / d(face coefficient) times lam_idp, per cell
    A* lam_invD[3];       // accumulated
    A* lam_ao[3];
    A* lam_cpe;
    A* lam_cp[3];         // [1], [2]
};
// PISO aodu(u^n, u')
template <class T, class A>
struct AdjAodu {
    const uint8_t* flag;
    const T* u[3];
    const T* uc[3];
    const A* lam_ao[3];
    A* lam_u[3];          // accumulated
    A* lam_uc[3];
};
// Rhie-Chow coefficients (u*, p^n, 1/D, D, boundary textures) -> cp, cpE,
// 1/Dp, bp
template <class T, class A>
struct AdjRhieChow {
    const uint8_t* flag;
    const T* p;
    const T* invD[3];
    const T* D[3];
    const A* lam_cp[3];
    const A* lam_cpe;
    const A* lam_idp;
    const A* lam_bp;
    A* w[6];              // scratch: per face, the adjoint of its (g - ga) scaled by sign / dx
    A* ldi[6];            // scratch: per face, the adjoint of its mean 1/D
    A* lam_u[3];          // accumulated
    A* lam_p;
    A* lam_invD[3];
    A* lam_D[3];
    A* lam_bnd[6][3];     // nullptr: the face's texture is not an input
};
template <class T, class A> void adj_correct(const ParamsT<Dual<A>>& p, const AdjCorrect<T, A>& a, hipStream_t s);
template <class T, class A> void adj_piso_coef(const ParamsT<Dual<A>>& p, const AdjPisoCoef<T, A>& a, hipStream_t s);
template <class T, class A> void adj_aodu(const ParamsT<Dual<A>>& p, const AdjAodu<T, A>& a, hipStream_t s);
template <class T, class A> void adj_rhie_chow(const ParamsT<Dual<A>>& p, const AdjRhieChow<T, A>& a, hipStream_t s);

// A transposed sweep runs in two parts: adj_*_lam writes lam_in from
// lam_out (the adjoint of the iterate, every cell); adj_*_coef adds the
// coefficient adjoints of up to ADJ_WIN consecutive sweeps at once (their
// output adjoints and input iterates in the window), accumulated in registers
// so that each coefficient adjoint is read and written once per window.
constexpr int ADJ_WIN = 16;
template <class T, class A>
struct SweepWindow {
    int n;
    const A* lo[ADJ_WIN][3];      // adjoint of each sweep's output ([0] only: pressure)
    const T* in[ADJ_WIN][3];      // each sweep's input iterate (pressure: nullptr is the zero iterate)
};
template <class T, class A> void adj_vel_lam(const AdjVelSweep<T, A>& a, int N, hipStream_t s);
template <class T, class A>
void adj_vel_coef(const AdjVelSweep<T, A>& a, const SweepWindow<T, A>& w, int N, hipStream_t s);
template <class T, class A> void adj_pres_lam(const AdjPresSweep<T, A>& a, int N, hipStream_t s);
template <class T, class A>
void adj_pres_coef(const AdjPresSweep<T, A>& a, const SweepWindow<T, A>& w, int N, hipStream_t s);
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
// zero m arrays of n elements each in one launch (m <= ADJ_ZERO_MAX)
constexpr int ADJ_ZERO_MAX = 32;
template <class T>
struct ZeroList {
    int m;
    T* p[ADJ_ZERO_MAX];
};
template <class T> void adj_zero(const ZeroList<T>& z, int n, hipStream_t s);
// acc[0] += sum_i t[i] in a fixed order: per-workgroup partials into
// part (ADJ_SUM_PARTS elements), then one workgroup sums them
constexpr int ADJ_SUM_PARTS = 512;
template <class T> void adj_sum(const T* t, int n, double* acc, double* part, hipStream_t s);

}  // namespace uws
