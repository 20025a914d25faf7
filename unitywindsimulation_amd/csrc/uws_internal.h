// Internal declarations shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/uws.h"

namespace uws {

// Uniform parameters of one solver: the reference's per-shader uniforms
// (FvmSolverGpu.cs:370-396, 499-507, 529-547, 561-580) gathered in one block.
struct Params {
    int X, Y, Z;          // gridRes
    int N;                // X*Y*Z (< 2^31, checked at creation)
    float dx, dt, ds, dv, nu, den, cs;
    float fext[3];
    int velBnd[6];        // X0 Xn Y0 Yn Z0 Zn
    int presBnd[6];
    int useLES;
    int wallSlots;        // 1 or 3
    int wallModel;        // UWS_WALL_SMOOTH / UWS_WALL_ROUGH
    float wallZ0;
};

// Device fields, structure-of-arrays: every vector field is three planes of
// N floats; flags are one byte per cell.  Face coefficients are stored once
// per face on the cell that owns the face's +x/+y/+z side (E, N, U faces);
// the W/S/D coefficients of a cell are the E/N/U coefficients of its
// W/S/D neighbour.  A face between a fluid cell and anything else has a
// zero coefficient, so the Jacobi sweeps need no neighbour flags.
struct Fields {
    uint8_t* flag;
    float* vel[3][3];     // three velocity buffers x 3 components (rotated, see Solver)
    float* ucorr[3];      // u'  (velCorrectField)
    float* p;             // p^n (presFieldLastTime)
    float* pc[2];         // p' ping-pong (presCorrectField / LastIter)
    float* nut;           // eddyVisField
    float* nuw[3];        // faceEddyVisField slots W(.x) S(.y) D(.z)
    float* D[3];          // DField
    float* invD[3];       // 1/DField (fluid cells; 0 elsewhere)
    float* b[3];          // bField
    float* S[3];          // momentum: u^n_P,n + u^n_nb,n on the owned E/N/U face
    float* G[3];          // momentum: nu_f*ds/dx on the owned E/N/U face
    float* cp[3];         // pressure: (dv/den)*DInv_f/dx on the owned E/N/U face
    float* cpisoE;        // PISO E/W face: (dv/den)*(DInv_f/dx)
    float* Dp;            // DFieldPresCorrect
    float* invDp;         // 1/Dp (fluid) or 0
    float* bp;            // bFieldPresCorrect
    float* aodu[3];       // AodUCorrectField
    float* invDpiso;      // 1/presCoeff of the PISO corrector equation
    float* bpiso;         // velTerm of the PISO corrector equation
    float* bnd[6][3];     // boundary textures, [face][comp][j*W + i]
    double* partial;      // reduction partials
    float* scalar;        // small scalars (normalisation reference value)
};

// kernel launchers (uws_kernels.hip); all enqueue on `s`
void launch_wall_eddy(const Params& p, const Fields& f, const float* const un[3], bool wall, bool eddy, hipStream_t s);
void launch_les_term(const Params& p, const Fields& f, const float* const un[3], hipStream_t s);
void launch_vel_precompute(const Params& p, const Fields& f, const float* const un[3], bool fuse_les, hipStream_t s);
void launch_vel_sweep(const Params& p, const Fields& f, const float* const uk[3], float* const unew[3], hipStream_t s);
void launch_pres_precompute(const Params& p, const Fields& f, const float* const ustar[3], hipStream_t s);
void launch_pres_sweep(const Params& p, const Fields& f, const float* cE, const float* cN, const float* cU,
                       const float* invDp, const float* bp, const float* pk, float* pnew, hipStream_t s);
void launch_apply_pres(const Params& p, const Fields& f, const float* pc, bool normalise, hipStream_t s);
void launch_vel_correct(const Params& p, const Fields& f, const float* pc, float* const u[3], hipStream_t s);
void launch_piso_aodu(const Params& p, const Fields& f, const float* const un[3], hipStream_t s);
void launch_piso_precompute(const Params& p, const Fields& f, hipStream_t s);
void launch_piso_correct(const Params& p, const Fields& f, const float* pc, float* const u[3], hipStream_t s);
// FP64 sum of squares of d[i]*(a[i]-b[i]) over ncomp planes; b may be null (zero)
void launch_residual(const Params& p, const Fields& f, int ncomp, const float* const d[3], const float* const a[3],
                     const float* const b[3], double* out, hipStream_t s);

}  // namespace uws
