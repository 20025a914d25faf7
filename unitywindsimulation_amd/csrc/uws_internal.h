// Internal declarations shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/uws.h"
#include "uws_scalar.h"

namespace uws {

// Uniform parameters of one solver: the reference's per-shader uniforms
// (FvmSolverGpu.cs:370-396, 499-507, 529-547, 561-580) gathered in one block,
// in the arithmetic type T of the run.
template <class T>
struct ParamsT {
    int X, Y, Z;          // gridRes
    int N;                // X*Y*Z (< 2^29, checked at creation)
    int ty;               // workgroup rows: tiles are 64 x ty cells of one z-plane
    int tiles;            // workgroup tiles of this grid
    T dx, dt, ds, dv, nu, den, cs;
    T fext[3];
    int velBnd[6];        // X0 Xn Y0 Yn Z0 Zn
    int presBnd[6];
    int useLES;
    int wallSlots;        // 1 or 3
    int wallModel;        // UWS_WALL_SMOOTH / UWS_WALL_ROUGH
    T wallZ0;
    int normRef;          // flat index of the pressure normalisation cell, -1 when it is inactive
    int zchunk;           // z planes per workgroup of the two-sweep kernels
    int by2;              // tile rows of the two-sweep kernels (outputs + 2 halo rows, at most 15)
    int zchunk3;          // z planes per workgroup of the three-sweep pressure kernel
    int zchunk4;          // z planes per workgroup of the four-sweep pressure kernel
    int nonuniform;       // 1: stretched grid (GridType.CollNonUniform), spacings from FieldsT::h
    // p2 = 1 (float, double) where dx and den are powers of two: the cell
    // kernels multiply by these exact reciprocals instead of dividing by dx,
    // 2 dx, dx / 2 and den (the same real number, so the same rounding) and
    // read the uniform quotients below, computed on the host with the same
    // IEEE operations as the expressions they replace
    int p2;
    T rdx, r2dx, rhdx, rden;
    T dv_dt, dvden, inv_den, nuds_dx;  // dv / dt, dv / den, 1 / den, (nu ds) / dx
};

// Device fields, structure-of-arrays: every vector field is three planes of
// N values; flags are one byte per cell.  Face coefficients are stored once
// per face on the cell that owns the face's +x/+y/+z side (E, N, U faces);
// the W/S/D coefficients of a cell are the E/N/U coefficients of its
// W/S/D neighbour.  A face between a fluid cell and anything else has a
// zero coefficient, so the Jacobi sweeps need no neighbour flags.
template <class T>
struct FieldsT {
    uint8_t* flag;
    T* vel[3][3];         // three velocity buffers x 3 components (roles rotate, see the host)
    T* ucorr[3];          // u'  (velCorrectField)
    T* p;                 // p^n (presFieldLastTime)
    T* pc[2];             // p' ping-pong (presCorrectField / LastIter)
    T* pcrec[2];          // where non-null, the three-sweep kernel also stores its first and second sweeps (adjoint records)
    T* nut;               // eddyVisField
    T* grad[9];           // cell velocity gradient du_i/dx_j (index 3i + j) of fluid cells, from the eddy-viscosity kernel
    T* nuw[3];            // faceEddyVisField slots W(.x) S(.y) D(.z)
    T* D[3];              // DField
    T* invD[3];           // 1/DField (fluid cells; 0 elsewhere)
    T* b[3];              // bField
    T* S[3];              // momentum: u^n_P,n + u^n_nb,n on the owned E/N/U face
    T* G[3];              // momentum: nu_f*ds/dx on the owned E/N/U face
    T* cp[3];             // pressure: (dv/den)*DInv_f/dx on the owned E/N/U face
    T* cpisoE;            // PISO E/W face: (dv/den)*(DInv_f/dx)
    T* Dp;                // DFieldPresCorrect
    T* invDp;             // 1/Dp (fluid) or 0
    T* bp;                // bFieldPresCorrect
    T* aodu[3];           // AodUCorrectField
    T* invDpiso;          // 1/presCoeff of the PISO corrector equation
    T* bpiso;             // velTerm of the PISO corrector equation
    T* bnd[6][3];         // boundary textures, [face][comp][j*W + i]
    // stretched grid only (null otherwise)
    T* h[3];              // cell sizes along x (X values), y (Y), z (Z): facePos[i+1] - facePos[i]
    T* flux[3];           // velFaceFluxField: Rhie-Chow velocity of each cell's W / S / D face
    T* fluxB[3];          // ... of the + faces on the domain boundary: x = X [z*Y + y], y = Y [z*X + x], z = Z [y*X + x]
    // defect correction with low-precision correction storage
    T* rho;               // defect right-hand side
    void* e16[2];         // correction iterates, 2-byte storage
    double* partial;      // reduction partials (one per workgroup)
    double* norms;        // FP64 norms written by the step (see UWS_NORM_*)
    T* scalar;            // small scalars: [0] normalisation value, [1] defect scale
};

// Multi-case tile table for the ensemble launch: workgroup tile t belongs to
// the case c with first_tile[c] <= t < first_tile[c+1].
template <class T>
struct BatchTable {
    int ncases;
    int total_tiles;
    int ty;                     // workgroup rows (every case uses the same)
    int persistent;             // 1: grid-stride launch of a fixed number of workgroups
    const int* first_tile;      // [ncases + 1]
    int total_ps2;              // workgroups of the two-sweep pressure kernel over all cases
    const int* first_ps2;       // [ncases + 1]: first two-sweep workgroup of each case
    int total_ps3;              // workgroups of the three-sweep pressure kernel over all cases
    const int* first_ps3;       // [ncases + 1]: first three-sweep workgroup of each case
    int total_ps4;              // workgroups of the four-sweep pressure kernel over all cases
    const int* first_ps4;       // [ncases + 1]: first four-sweep workgroup of each case
    const ParamsT<T>* params;   // [ncases]
    const FieldsT<T>* fields;   // [ncases]
};

// Buffer roles of one launch: the same small integers for every case of an
// ensemble, which share V, P and C and so rotate their buffers identically.
struct Roles {
    int state;      // velocity buffer holding u^n
    int vin;        // sweep input velocity buffer
    int vout;       // sweep output / corrected velocity buffer
    int pin;        // p' input buffer (-1: zero)
    int pout;       // p' output buffer
    int coef;       // pressure coefficients: 0 SIMPLE, 1 PISO
    int ein;        // low-precision correction input buffer (-1: zero)
    int eout;       // low-precision correction output buffer
    int flags;      // kernel options (K_* bits)
};
enum {
    K_WALL = 1, K_EDDY = 2, K_FUSE_LES = 4, K_LOWP_FIRST = 16, K_LOWP_LAST = 32,
    K_BF16 = 64,
    K_NO_INPLACE = 128,  // corrections write u' only, not u += u' / p += p'
    K_NO_FLUX = 256,     // stretched-grid pressure precompute: leave velFaceFluxField as it is
    K_APPLY_P = 512      // SIMPLE velocity correction: also p += p' in its fluid cells (CSApplyPresCorrection)
};

// kernel identifiers understood by launch()
enum Kernel {
    KN_WALL_EDDY, KN_LES_TERM, KN_VEL_PRECOMP, KN_VEL_SWEEP, KN_PRES_PRECOMP, KN_PRES_SWEEP,
    KN_APPLY_PRES, KN_PRES_NORMALISE, KN_VEL_CORRECT, KN_PISO_AODU, KN_PISO_PRECOMP, KN_PISO_CORRECT,
    KN_ZERO_SOLID_VEL, KN_FILL_PC,
    KN_DEFECT_RESIDUAL, KN_LOWP_SWEEP,
    KN_RES_VEL, KN_RES_PRES, KN_DIVERGENCE, KN_NONFINITE, KN_SUM_PARTIALS, KN_DEFECT_SCALE,
    KN_PRES_SWEEP2,  // two pressure Jacobi sweeps in one launch (single-case only)
    // stretched-grid variants, selected by launch() when ParamsT::nonuniform is set
    KN_NU_VEL_PRECOMP, KN_NU_VEL_SWEEP, KN_NU_PRES_PRECOMP, KN_NU_VEL_CORRECT, KN_NU_DIVERGENCE,
    KN_VEL_SWEEP2,  // two momentum Jacobi sweeps in one launch (single case, uniform grid, FP32)
    KN_PRES_SWEEP3,  // three pressure Jacobi sweeps in one launch (single case, FP32)
    KN_LOWP_SWEEP3,  // three low-precision correction sweeps in one launch (single case, FP32)
    KN_PRES_SWEEP4,  // four pressure Jacobi sweeps in one launch (single case, FP32)
    KN_LOWP_SWEEP4   // four low-precision correction sweeps in one launch (single case, FP32)
};

// Single-case launch of kernel k on stream s.  `slot` selects the norms
// entry for reduction kernels.
template <class T>
void launch(Kernel k, const ParamsT<T>& p, const FieldsT<T>& f, const Roles& r, hipStream_t s, int slot = 0);

// Up to this many cases' field tables travel in the kernel arguments of an
// ensemble's cell-wise launches (4 x 728 B, within the 4 KB argument limit):
// array pointers read from the kernel arguments are addressed as global
// memory, pointers read from a table in device memory as generic (flat)
// memory, whose loads also wait on the LDS counter.
constexpr int BATCH_ARG_CASES = 4;
template <class T>
struct BatchArgFields {
    FieldsT<T> f[BATCH_ARG_CASES];
};

// Ensemble launch: one grid over every case's tiles.  host_fields: the
// group's field tables on the host (tab.fields' contents), or null; with
// at most BATCH_ARG_CASES cases the cell-wise kernels take them as kernel
// arguments.
template <class T>
void launch_batch(Kernel k, const BatchTable<T>& tab, int grid, const Roles& r, hipStream_t s, int slot = 0,
                  const FieldsT<T>* host_fields = nullptr);

// A background flow field resident on the device, for the warm start of
// InitVelPresFieldsFromBackGroundFlow (geometry in R, values in T).  The
// stored elements are of type S (T itself, or the caller's float arrays
// staged as they are), converted to T where they are read; velocity
// component a of background cell i is vel[a][i * vstride] (vstride 1 for
// planes, ncomp for the caller's [cell][ncomp] layout).
template <class T, class R, class S = T>
struct BgField {
    int X, Y, Z;
    R dx;
    R pos[3];        // physFieldPos of the background grid (its centre)
    R fgPos[3];      // physFieldPos of this grid
    R angle;         // flow-field orientation [rad]
    const S* vel[3];
    const S* pres;   // may be null (zero)
    int vstride;
};
template <class T, class R, class S>
void init_from_background(const ParamsT<T>& p, const FieldsT<T>& f, const BgField<T, R, S>& bg, hipStream_t s);

// Fixed staging array of the allocation-free warm start, one per GPU (a
// file-scope device array of UWS_BG_STAGING_MAX_BYTES).  bg_staging_setup()
// resolves its address on the current device once (hipGetSymbolAddress) and
// keeps it; bg_staging(dev) returns the kept address, nullptr before setup.
int bg_staging_setup();
void* bg_staging(int device);

// bandwidth of a plain streaming kernel (nin float arrays of `cells` read,
// one written, back to back on the same arrays), GB/s on the current device
int stream_bandwidth(int64_t cells, int nin, int reps, double* gbps);

// sum[c][i] += u[c][i] in FP64 (time averaging of the velocity)
template <class T>
void mean_accumulate(T* const u[3], double* const sum[3], int n, hipStream_t s);

// workgroups of the two-sweep pressure kernel for one grid
int pres_sweep2_grid(int X, int Y, int Z, int zchunk, int by);
int pres_sweep3_grid(int X, int Y, int Z, int zchunk);
int pres_sweep4_grid(int X, int Y, int Z, int zchunk);

inline int tiles_of(int X, int Y, int Z, int ty) { return ((X + 63) / 64) * ((Y + ty - 1) / ty) * Z; }

}  // namespace uws
