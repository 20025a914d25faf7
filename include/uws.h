/*
 * uws.h -- C ABI of the MI355X (gfx950) finite-volume incompressible wind
 * solver step: the drop-in replacement for UnityWindSimulation's GPU solver
 * (`IFvmSolver` implemented by `FvmSolverGpu`).
 *
 * Conventions
 *  - Plain C, blittable types only; every entry point returns int
 *    (UWS_OK == 0, negative = error class, see uws_strerror).
 *  - Axes follow the reference: x streamwise (inlet X0, outlet Xn),
 *    y up (ground Y0), z spanwise.  Cell (x, y, z) has the flat index
 *    z*X*Y + y*X + x  (Assets/Code/FvmSolver/FvmSolverGpu.cs:947-950,
 *    Assets/Resources/Shaders/SimShaders/FvmColl.compute:122-126).
 *  - Host arrays are owned by the caller; the library copies in/out.
 *    Vector fields cross the boundary as [cell][ncomp] float (ncomp 3 or 4,
 *    the 4th lane written as 0), the reference's netCDF export layout
 *    velField(time, z, y, x, 4) (Assets/Code/IO/NcFlowFieldExporter.cs:244-257).
 *  - One handle = one HIP stream.  Calls on one handle must be serialised
 *    by the caller; different handles may be driven from different threads.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *  - IFvmSolver                 Assets/Code/FvmSolver/IFvmSolver.cs:3-24
 *  - FvmSolverGpu ctor          Assets/Code/FvmSolver/FvmSolverGpu.cs:173-222
 *  - FluidSimConfig / enums     Assets/Code/FvmSolver/FluidSimConfig.cs:9-370
 *  - RuntimeConfig.windSpeed    Assets/Code/RuntimeConfig/RuntimeConfig.cs:9
 */
#ifndef UWS_H
#define UWS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UWS_ABI_VERSION 1

/* ---- status codes (netCDF-style int status, CSnetCDF.cs:273-280) ---- */
enum {
    UWS_OK = 0,
    UWS_ERR_INVALID_ARG = -1,   /* null pointer, bad enum, size mismatch */
    UWS_ERR_CONFIG = -2,        /* inconsistent configuration */
    UWS_ERR_ALLOC = -3,         /* device allocation failed */
    UWS_ERR_DEVICE = -4,        /* HIP runtime error / no device */
    UWS_ERR_UNSUPPORTED = -5,   /* combination the reference also rejects */
    UWS_ERR_NONFINITE = -6,     /* NaN/Inf detected by a norm check */
    UWS_ERR_TOO_LARGE = -7      /* input larger than the fixed staging array of an
                                   allocation-free entry point */
};

/* ---- enums: identical integer values to the reference ---- */
/* VelBndCond (FluidSimConfig.cs:16-21, Constants.hlsl:8-10) */
enum { UWS_VEL_FIXED_VALUE = 0, UWS_VEL_ZERO_GRAD = 1, UWS_VEL_SYMMETRY = 2 };
/* PresBndCond (FluidSimConfig.cs:23-28, Constants.hlsl:12-14) */
enum { UWS_PRES_FIXED_VALUE = 0, UWS_PRES_ZERO_GRAD = 1, UWS_PRES_SYMMETRY = 2 };
/* FvmSolverType (FluidSimConfig.cs:30-34) */
enum { UWS_SIMPLE = 0, UWS_PISO = 1 };
/* TurbulenceModel (FluidSimConfig.cs:73-77) */
enum { UWS_TURB_NONE = 0, UWS_TURB_SMAGORINSKY = 1 };
/* InflowType (FluidSimConfig.cs:86-90); POWER_LAW is an extension */
enum { UWS_INFLOW_CONSTANT = 0, UWS_INFLOW_LOG = 1, UWS_INFLOW_POWER_LAW = 2 };
/* cell flags (Constants.hlsl:4-6) */
enum { UWS_CELL_FLUID = 0, UWS_CELL_SOLID = 1 };
/* wall model: 0 = reference smooth log law solved by Newton
 * (WallFunc.compute:77-123); 1 = rough-wall log law with roughness z0
 * (extension). */
enum { UWS_WALL_SMOOTH = 0, UWS_WALL_ROUGH = 1 };
/* GridType (FluidSimConfig.cs:9-14): COLL_NONUNIFORM is the stretched
 * collocated grid of FvmCollNonUniform.compute; STAGGERED is not built */
enum { UWS_GRID_COLLOCATED = 0, UWS_GRID_COLL_NONUNIFORM = 1, UWS_GRID_STAGGERED = 2 };
/* boundary faces, the order of the reference's six boundary textures
 * (FvmSolverGpu.cs:295-300) */
enum { UWS_X0 = 0, UWS_XN = 1, UWS_Y0 = 2, UWS_YN = 3, UWS_Z0 = 4, UWS_ZN = 5 };

/*
 * Solver configuration: the FluidSimConfig fields the step reads, already
 * derived by FluidSimConfig.Init() (FluidSimConfig.cs:322-370): grid_res is
 * gridRes, dx is physDomainSize.x / gridResX.  ds = dx*dx, dv = dx*dx*dx and
 * nu = mu/den are derived inside the library in float, as Init() does.
 */
typedef struct uws_config {
    int32_t grid_res[3];          /* gridRes (x, y up, z span) */
    float   dx;                   /* cell size [m] */
    float   dt;                   /* time step [s] */
    float   mu;                   /* dynamic viscosity */
    float   den;                  /* density */
    float   external_force[3];    /* externalForce */
    int32_t vel_bnd[6];           /* velBndCondX0..Zn  (UWS_VEL_*) */
    int32_t pres_bnd[6];          /* presBndCondX0..Zn (UWS_PRES_*) */
    int32_t fvm_solver_type;      /* fvmSolverType (UWS_SIMPLE / UWS_PISO) */
    int32_t piso_num_correctors;  /* PISONumCorrectors */
    int32_t turbulence_model;     /* turbulenceModel */
    float   smagorinsky_constant; /* smagorinskyConstant */
    int32_t vel_max_num_iter;     /* velMaxNumIter (momentum Jacobi sweeps) */
    int32_t pres_max_num_iter;    /* presMaxNumIter (pressure Jacobi sweeps) */
    int32_t cal_residual;         /* calResidual: residual-based early exit */
    int32_t vel_residual_check_interval;
    float   vel_tolerance;
    int32_t pres_residual_check_interval;
    float   pres_tolerance;
    /* extensions (no reference field) */
    int32_t wall_slots;           /* 1: only the x-normal wall-viscosity slot
                                     survives, as the reference runs with its
                                     single-channel texture (FvmSolverGpu.cs:304);
                                     3: all three slots, as the shaders are written */
    int32_t wall_model;           /* UWS_WALL_SMOOTH / UWS_WALL_ROUGH */
    float   wall_z0;              /* roughness length for UWS_WALL_ROUGH [m] */
    int32_t precision;            /* UWS_PREC_*: arithmetic of the run */
    int32_t refine_rounds;        /* low-precision modes: defect-correction rounds per pressure solve */
    int32_t refine_sweeps;        /* low-precision modes: Jacobi sweeps per round */
    int32_t compute_norms;        /* 1: every step also computes the FP64 norms (uws_get_norms) */
    int32_t grid_type;            /* gridType (UWS_GRID_*); COLL_NONUNIFORM runs SIMPLE only, as the
                                     reference (InitPisoShader, FvmSolverGpu.cs:490-520) */
    int32_t reserved[3];          /* must be zero */
} uws_config;

/* precision modes (extension).  FP32 is the reference's arithmetic; FP64 is
 * the same step in double on the GPU; FP32_FP16 / FP32_BF16 solve every
 * pressure equation as FP32 defect correction (refine_rounds rounds of
 * refine_sweeps Jacobi sweeps) with the correction iterate stored in 2 bytes,
 * instead of pres_max_num_iter plain FP32 sweeps.  CSTEP runs the FP64 step
 * with every value a complex number (complex-step differentiation: the
 * imaginary parts carry h times the derivative along a chosen direction,
 * see uws_cstep_set_imag). */
enum { UWS_PREC_FP32 = 0, UWS_PREC_FP64 = 1, UWS_PREC_FP32_FP16 = 2, UWS_PREC_FP32_BF16 = 3, UWS_PREC_CSTEP = 4 };

typedef struct uws_solver* uws_handle;

/* Fill *cfg with the FluidSimConfig field defaults (FluidSimConfig.cs:113-178)
 * for a grid of the given resolution and dx = 10/gridResX (physDomainSize.x
 * default 10), plus wall_slots = 1, wall_model = smooth. */
int uws_config_default(uws_config* cfg);

const char* uws_strerror(int status);
int uws_abi_version(void);

/* FvmSolverGpu(FluidSimConfig, RuntimeConfig) (FvmSolverGpu.cs:173-222):
 * allocates every field on the current HIP device, initialises the flag
 * field to all-fluid (InitFlagField, :935-939), the velocity/pressure fields
 * to zero and all boundary textures to zero (SetZeroVelBndCond, :989-994).
 * The caller then applies InitVelPresFields / SetFixedValueVelBndCond with
 * the inflow it wants (uws_init_vel_pres_fields, uws_set_fixed_value_vel_bnd_cond). */
int uws_create(const uws_config* cfg, uws_handle* out);
int uws_destroy(uws_handle h);

/* IFvmSolver.ChangePhysDomainSize() (FvmSolverGpu.cs:1034-1046, called by
 * FvmSolverController.cs:863-874 after FluidSimConfig.Init() re-derived
 * gridRes and dx from a new physDomainSize): the handle takes cfg (same
 * precision and grid type, else UWS_ERR_CONFIG).  Every field is
 * reallocated for cfg's grid and zeroed as by uws_create; where gridRes is
 * unchanged the flags and the boundary textures are kept (the reference
 * keeps its textures), otherwise flags are all fluid and textures zero.
 * The step counter and time are kept.  The caller then re-initialises the
 * fields as the reference does (uws_init_vel_pres_fields) and, after a grid
 * change, the mask and inflow.  Adjoint handles made from h must be
 * destroyed first.  On an allocation failure the handle can only be
 * destroyed. */
int uws_change_phys_domain_size(uws_handle h, const uws_config* cfg);
int uws_get_config(uws_handle h, uws_config* out);
int uws_num_cells(uws_handle h, int64_t* n);

/* IFvmSolver.InitFlags(int[] flags) (FvmSolverGpu.cs:972-987,
 * InitShaderColl.compute:345-364): flags[z*X*Y + y*X + x] > 0.5 marks a solid
 * cell; solid cells get zero velocity in every velocity buffer.  n must be X*Y*Z. */
int uws_init_flags(uws_handle h, const int32_t* flags, int64_t n);

/* SolidType.Box (FvmSolverGpu.cs:963-970, InitShaderColl.compute:302-343):
 * the reference's hard-coded box x in [20,24), y in [0,8), z in [18,22). */
int uws_init_box(uws_handle h);

/* Stretched grid (grid_type UWS_GRID_COLL_NONUNIFORM).  The reference keeps
 * the face coordinates of each axis in facePosX/Y/Z buffers of gridRes+1
 * floats (InitMeshBuffers, FvmSolverGpu.cs:682-788); every stretched-grid
 * kernel derives its spacings from them.  uws_create starts from uniform
 * faces i*dx; uws_set_face_positions replaces them (pos_x has X+1 entries,
 * pos_y Y+1, pos_z Z+1, strictly increasing).  Set them before the inflow and
 * field initialisation, which place the log profile at face midpoints
 * (CSInitLogVelPresFieldsNonUniform, CSSetLogVelBndCondNonUniform,
 * InitShaderColl.compute:134-154, 287-300).  Returns UWS_ERR_CONFIG on a
 * collocated grid. */
int uws_set_face_positions(uws_handle h, const float* pos_x, const float* pos_y, const float* pos_z);
int uws_get_face_positions(uws_handle h, float* pos_x, float* pos_y, float* pos_z);

/* InitMeshBuffers' generator for one axis (FvmSolverGpu.cs:682-788), host
 * only: n cells, a block of `uniform` cells of size dx preceded by
 * `one_side` cells growing away from it, dx*ratio^k for the k-th cell from
 * the block, and followed by the remaining n - uniform - one_side cells
 * growing the same way; faces accumulate from 0 in float.  out has n+1
 * entries.  The reference hard-codes uniform = (50, 50, 50),
 * one_side = (25, 0, 25), ratio = 1.08.  Returns UWS_ERR_CONFIG when the
 * block does not fit or the trailing side is longer than the leading one
 * (the reference indexes the leading side's size table for both). */
int uws_stretched_face_positions(int32_t n, float dx, int32_t uniform, int32_t one_side, float ratio, float* out);

/* IFvmSolver.InitVelPresFields() (FvmSolverGpu.cs:790-896): inflow_type
 * UWS_INFLOW_CONSTANT sets u = (wind_speed, 0, 0) everywhere
 * (InitShaderColl.compute:94-110); UWS_INFLOW_LOG sets the log profile with
 * y = j*dx (InitShaderColl.compute:112-132).  p, p', u' are zeroed.  Like the
 * reference this does not look at the flag field. */
int uws_init_vel_pres_fields(uws_handle h, int32_t inflow_type, float wind_speed);

/* IFvmSolver.SetFixedValueVelBndCond() (FvmSolverGpu.cs:996-1022):
 * CONSTANT fills X0 with (wind_speed,0,0) and Zn with vel_zn
 * (InitShaderColl.compute:252-265); LOG fills X0 with the log profile at
 * y = (j+1/2)dx (InitShaderColl.compute:272-285).  POWER_LAW (extension) fills
 * X0 with u = wind_speed * ((j+1/2)dx / ref_height)^alpha. vel_zn may be NULL. */
int uws_set_fixed_value_vel_bnd_cond(uws_handle h, int32_t inflow_type, float wind_speed,
                                     const float vel_zn[3], float ref_height, float alpha);

/* IFvmSolver.InitVelPresFieldsFromBackGroundFlow (FvmSolverGpu.cs:898-933,
 * kernel CSInitVelPresFieldsFromBGFlowRotate, InitShaderColl.compute:184-226):
 * warm start from a background flow on another grid.  Every cell's offset
 * from this grid's centre, (i - gridRes/2) dx, is rotated by field_rot_deg
 * (RuntimeConfig.flowFieldOrientation) about y into world coordinates around
 * phys_field_pos (FluidSimConfig.physFieldPos), rotated back into the
 * background grid's frame around bg->phys_field_pos and sampled there with
 * trilinear filtering and clamped addressing.  All velocity buffers get the
 * sampled velocity (not rotated, as in the reference), p^n the sampled
 * pressure, p' and u' zero; flags are not looked at.  bg_vel is the
 * background's u in [z][y][x][ncomp] (ncomp 3 or 4, as uws_get_velocity
 * returns it), bg_pres its p in [z][y][x] or NULL for zero. */
typedef struct uws_background {
    int32_t grid_res[3];
    float dx;
    float phys_field_pos[3];
} uws_background;
int uws_init_vel_pres_fields_from_background(uws_handle h, const float phys_field_pos[3], float field_rot_deg,
                                             const uws_background* bg, const float* bg_vel, int32_t ncomp,
                                             const float* bg_pres);

/* Allocation-free variant of uws_init_vel_pres_fields_from_background.  No
 * reference counterpart: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same fields, bit for bit; instead of a
 * device buffer allocated per call, the caller's arrays are staged as they
 * are in a fixed device array of UWS_BG_STAGING_MAX_BYTES per GPU, held until
 * the process ends.  Needs cells * (ncomp + (bg_pres ? 1 : 0)) * 4 bytes
 * <= UWS_BG_STAGING_MAX_BYTES (at most 1,677,721 background cells with
 * ncomp 4 and a pressure), else UWS_ERR_TOO_LARGE.  The call synchronises
 * the handle's stream before it returns.  Every handle on one GPU shares the
 * array: calls that overlap in time on one GPU (from several threads) race. */
#define UWS_BG_STAGING_MAX_BYTES (32u << 20)
int uws_init_vel_pres_fields_from_background_noalloc(uws_handle h, const float phys_field_pos[3],
                                                     float field_rot_deg, const uws_background* bg,
                                                     const float* bg_vel, int32_t ncomp, const float* bg_pres);

/* Upload one boundary texture (face UWS_X0..UWS_ZN) as [j][i][3] float where
 * (i, j) are the reference texture coordinates: X faces (y, z), Y faces (x, z),
 * Z faces (x, y) (FvmSolverGpu.cs:295-300).  Download mirrors it. */
int uws_set_bnd_field(uws_handle h, int32_t face, const float* xyz, int64_t n);
int uws_get_bnd_field(uws_handle h, int32_t face, float* xyz, int64_t n);

/* Velocity state upload/download, [cell][ncomp], ncomp in {3, 4}.  Upload
 * writes every velocity buffer (as the reference init kernels do). */
int uws_set_velocity(uws_handle h, const float* src, int32_t ncomp, int64_t n);
int uws_get_velocity(uws_handle h, float* dst, int32_t ncomp, int64_t n);
int uws_set_pressure(uws_handle h, const float* src, int64_t n);
int uws_get_pressure(uws_handle h, float* dst, int64_t n);
int uws_get_flags(uws_handle h, int32_t* dst, int64_t n);

/* Seeded initial condition (extension): u = U(-amp, amp) per component in
 * fluid cells (SplitMix64 stream from seed), 0 in solid cells; p = 0. */
int uws_init_noise(uws_handle h, uint64_t seed, float amp);

/* FvmSolverGpu.Step() (FvmSolverGpu.cs:1050-1121) run nsteps times: LES
 * (wall function, eddy viscosity, transposed-stress term), momentum
 * predictor with velMaxNumIter Jacobi sweeps, Rhie-Chow pressure correction
 * with presMaxNumIter sweeps, velocity correction, PISO correctors.  Whole
 * steps only (no 30 ms time slicing).  Enqueued on the handle's stream. */
int uws_step(uws_handle h, int32_t nsteps);
int uws_synchronize(uws_handle h);
/* currSimStep / currPhyTime bookkeeping (FvmSolverGpu.cs:1108-1109) */
int uws_get_time(uws_handle h, int64_t* step, double* phys_time);

/* ---- intermediate fields, for per-kernel parity and diagnostics ---- */
enum {
    UWS_FIELD_VEL = 0,        /* u^n, velFieldLastTime (3) */
    UWS_FIELD_PRES = 1,       /* p^n, presFieldLastTime (1) */
    UWS_FIELD_VEL_WORK = 2,   /* velField within a step: u*, then corrected u (3) */
    UWS_FIELD_VEL_CORR = 3,   /* u', velCorrectField (3) */
    UWS_FIELD_PRES_CORR = 4,  /* p', presCorrectField (1) */
    UWS_FIELD_EDDY_VIS = 5,   /* nu_t, eddyVisField (1) */
    UWS_FIELD_WALL_VIS = 6,   /* nu_w slots (W, S, D faces), faceEddyVisField (3) */
    UWS_FIELD_B = 7,          /* bField (3) */
    UWS_FIELD_D = 8,          /* DField (3) */
    UWS_FIELD_DP = 9,         /* DFieldPresCorrect (1) */
    UWS_FIELD_BP = 10,        /* bFieldPresCorrect (1) */
    UWS_FIELD_AODU = 11,      /* AodUCorrectField (3) */
    UWS_FIELD_FACE_VEL = 12,  /* stretched grid: velFaceFluxField at the cell, the Rhie-Chow
                                 velocity of its W, S, D faces (3) */
    UWS_FIELD_COUNT = 13
};
int uws_field_ncomp(int32_t field);
int uws_get_field(uws_handle h, int32_t field, float* dst, int64_t n_floats);
int uws_set_field(uws_handle h, int32_t field, const float* src, int64_t n_floats);

/* The step split at the reference's kernel boundaries (FvmSolverGpu.cs
 * 1050-1121).  Running UWS_STAGE_WALL_FUNC .. UWS_STAGE_FINISH in order is
 * one uws_step (stages whose feature is disabled are no-ops). */
enum {
    UWS_STAGE_WALL_FUNC = 0,     /* CSCalLogLawWallFunc */
    UWS_STAGE_EDDY_VIS = 1,      /* CSCalEddyVis */
    UWS_STAGE_LES_TERM = 2,      /* CSCalLesDeferCorrectTerm */
    UWS_STAGE_VEL_PRECOMP = 3,   /* CSVelPredictPreCompute */
    UWS_STAGE_VEL_SWEEPS = 4,    /* velMaxNumIter x CSVelPredict */
    UWS_STAGE_PRES_PRECOMP = 5,  /* CSPresCorrectRhieChowPreCompute */
    UWS_STAGE_PRES_SWEEPS = 6,   /* presMaxNumIter x CSPresCorrectRhieChow */
    UWS_STAGE_APPLY_PRES = 7,    /* CSApplyPresCorrection (+ normalisation) */
    UWS_STAGE_VEL_CORRECT = 8,   /* CSApplyVelCorrectionDivergenceTheorem, zero p' */
    UWS_STAGE_PISO = 9,          /* one PISO corrector (PisoColl.compute) */
    UWS_STAGE_FINISH = 10,       /* u^{n+1} <- u, time bookkeeping */
    UWS_STAGE_COUNT = 11
};
int uws_run_stage(uws_handle h, int32_t stage);

/* Full-precision field download: as uws_get_field, in double (exact for FP64
 * runs, widened for FP32 runs). */
int uws_get_field_f64(uws_handle h, int32_t field, double* dst, int64_t n_floats);

/* FP64 norms of the last step (needs compute_norms = 1), host-normalised:
 *  [0] momentum residual of the last Jacobi sweep, sqrt(sum (D du)^2 / 3N)
 *      (the reference's definition, FvmSolverGpu.cs:1216-1236)
 *  [1] pressure residual of the last SIMPLE sweep, sqrt(sum (Dp dp')^2 / N)
 *      (FvmSolverGpu.cs:1345-1365)
 *  [2] divergence: rms over cells of the Rhie-Chow net outflow of u^{n+1}
 *      (bFieldPresCorrect of the final state, FvmColl.compute:2480-2787)
 *  [3] number of cells whose u or p is not finite
 *  [4 + 8*s + r] low-precision modes: rms of the defect of pressure solve s
 *      (0 = SIMPLE, 1.. = PISO correctors) at the start of round r.
 * Accumulation is in FP64 with a fixed summation order (deterministic).
 * Returns UWS_ERR_NONFINITE, with out[] filled as usual, when the last step
 * left a non-finite u or p somewhere ([3] > 0). */
enum { UWS_NORM_VEL_RESIDUAL = 0, UWS_NORM_PRES_RESIDUAL = 1, UWS_NORM_DIVERGENCE = 2, UWS_NORM_NONFINITE = 3,
       UWS_NORM_REFINE0 = 4, UWS_NORM_COUNT = 64 };
int uws_get_norms(uws_handle h, double* out, int32_t n);

/* Jacobi sweeps the last step ran: momentum (VelPredict) and SIMPLE pressure
 * (PresCorrect).  velMaxNumIter / presMaxNumIter, or k + 1 where the
 * residual-based exit (calResidual) stopped at sweep k (FvmSolverGpu.cs
 * 1250-1262, 1386-1398); 0 before the first step. */
int uws_get_sweep_counts(uws_handle h, int32_t* vel_sweeps, int32_t* pres_sweeps);

/* Diagnostics (no reference counterpart): average duration in milliseconds of
 * one launch of a hot-loop kernel, timed with HIP events on the handle's
 * stream over `reps` back-to-back launches on the current step data.
 * UWS_KERNEL_VEL_SWEEP reads u^n and writes a scratch velocity buffer;
 * UWS_KERNEL_PRES_SWEEP reads and writes the p' scratch pair;
 * UWS_KERNEL_PRES_SWEEP2 is the two-sweep pressure kernel the step uses;
 * UWS_KERNEL_VEL_SWEEP2 the two-sweep momentum kernel (FP32, uniform or
 * stretched grid; UWS_ERR_UNSUPPORTED elsewhere), from u^n into a scratch buffer;
 * UWS_KERNEL_PRES_SWEEP3 and UWS_KERNEL_PRES_SWEEP4 the three- and four-sweep
 * pressure kernels (FP32).  The solver state (u^n, p^n) is not changed.  Call
 * after at least one uws_step. */
enum { UWS_KERNEL_VEL_SWEEP = 0, UWS_KERNEL_PRES_SWEEP = 1, UWS_KERNEL_PRES_SWEEP2 = 2, UWS_KERNEL_VEL_SWEEP2 = 3,
       UWS_KERNEL_PRES_SWEEP3 = 4, UWS_KERNEL_PRES_SWEEP4 = 5 };
int uws_time_kernel(uws_handle h, int32_t kernel, int32_t reps, double* avg_ms);

/* Diagnostics (no reference counterpart): the most sweeps one launch of the
 * step runs, for the plain pressure solves (4, 3, 2 or 1; a solve's last
 * sweeps run in shorter launches) and for the momentum predictor (2 or 1),
 * as the multi-sweep policy picks them for this grid and precision. */
int uws_get_launch_plan(uws_handle h, int32_t* pres_sweeps_per_launch, int32_t* vel_sweeps_per_launch);

/* Diagnostics (no reference counterpart): the bandwidth ceiling a plain
 * streaming kernel reaches on the current device, in GB/s of bytes read and
 * written.  Each launch reads nin (1..16) float arrays of `cells` values and
 * writes one (out = sum of the inputs, 16-byte accesses), reps launches back
 * to back on the same arrays, as the Jacobi sweeps re-read theirs: a working
 * set (nin + 1) * 4 * cells bytes that fits the last-level cache is served
 * from it.  Allocates and frees its own arrays. */
int uws_stream_bandwidth(int64_t cells, int32_t nin, int32_t reps, double* gbps);

/* ---- ensemble of uneven cases (extension, no reference counterpart) ----
 * Several independent solvers (different grids, buildings, inflow) stepped
 * together: every kernel of the step runs once for all cases, over one tile
 * table spanning every case's grid, and each step is one captured HIP graph.
 * All cases must share vel_max_num_iter, pres_max_num_iter, the solver type,
 * the PISO corrector count, the turbulence model, precision FP32 and
 * compute_norms, and use no residual-based early exit.  Case c of the
 * ensemble behaves exactly like a uws_handle created from cfgs[c]. */
typedef struct uws_batch* uws_batch_handle;
int uws_batch_create(const uws_config* cfgs, int32_t ncases, uws_batch_handle* out);
int uws_batch_destroy(uws_batch_handle b);
int uws_batch_num_cases(uws_batch_handle b, int32_t* n);
int uws_batch_init_flags(uws_batch_handle b, int32_t c, const int32_t* flags, int64_t n);
int uws_batch_init_noise(uws_batch_handle b, int32_t c, uint64_t seed, float amp);
int uws_batch_set_fixed_value_vel_bnd_cond(uws_batch_handle b, int32_t c, int32_t inflow_type, float wind_speed,
                                           const float vel_zn[3], float ref_height, float alpha);
int uws_batch_set_velocity(uws_batch_handle b, int32_t c, const float* src, int32_t ncomp, int64_t n);
int uws_batch_get_velocity(uws_batch_handle b, int32_t c, float* dst, int32_t ncomp, int64_t n);
int uws_batch_get_pressure(uws_batch_handle b, int32_t c, float* dst, int64_t n);
int uws_batch_get_norms(uws_batch_handle b, int32_t c, double* out, int32_t n);
int uws_batch_step(uws_batch_handle b, int32_t nsteps);
int uws_batch_synchronize(uws_batch_handle b);
/* Launch shape of the ensemble kernels: tile rows ty in {1, 2, 4} (tiles of
 * 64 x ty cells), persistent != 0 for a grid-stride launch of a fixed number
 * of workgroups instead of one workgroup per tile, and group_bytes: > 0,
 * consecutive cases are grouped while their momentum-sweep working set (77 B
 * per cell) stays within group_bytes; 0, one group; < 0, the cases are packed
 * into the fewest groups whose two-sweep launches need at most one workgroup
 * per CU (a case larger than that is a group of its own).  Each group runs
 * its whole step (one launch per kernel for the group) before the next group,
 * so that a group's Jacobi sweeps are served from the last-level cache. */
int uws_batch_set_launch(uws_batch_handle b, int32_t ty, int32_t persistent, int64_t group_bytes);
int uws_batch_num_groups(uws_batch_handle b, int32_t* n);
/* Autotuner: times `steps` ensemble steps for each candidate launch shape
 * (tile rows, persistent, group budget) and keeps the fastest; every case's
 * state and clock are saved before and restored after the trials.
 * best = {ty, persistent, group_bytes}, best_ms = ms per step of the winner. */
int uws_batch_autotune(uws_batch_handle b, int32_t steps, int64_t best[3], double* best_ms);

/* Time mean of the velocity (extension; the reference averages exported
 * frames offline).  uws_mean_accumulate adds the current u^n to an FP64
 * running sum held on the device (allocated by the first call of either
 * function), uws_mean_reset clears it; uws_get_mean_velocity returns
 * sum / count as [cell][3] doubles (n = 3 * cells) and the count. */
int uws_mean_reset(uws_handle h);
int uws_mean_accumulate(uws_handle h);
int uws_get_mean_velocity(uws_handle h, double* dst, int64_t n, int64_t* count);

/* ---- differentiation (extension, no reference counterpart) ---- */

/* FluidSimConfig.smagorinskyConstant (FluidSimConfig.cs:156) of a live
 * handle, in double; used from the next step on. */
int uws_set_smagorinsky_constant(uws_handle h, double cs);
/* Boundary texture upload in double (exact for FP64 and complex-step runs);
 * layout as uws_set_bnd_field. */
int uws_set_bnd_field_f64(uws_handle h, int32_t face, const double* xyz, int64_t n);

/* Complex-step runs (precision UWS_PREC_CSTEP).  The real-valued setters set
 * real parts with a zero imaginary part; these set or read the imaginary
 * parts only.  what = UWS_CSTEP_VEL ([cell][3], every velocity buffer),
 * UWS_CSTEP_PRES ([cell]), UWS_CSTEP_BND (boundary texture `face`, [j][i][3]),
 * UWS_CSTEP_CS (one value: the Smagorinsky constant). */
enum { UWS_CSTEP_VEL = 0, UWS_CSTEP_PRES = 1, UWS_CSTEP_BND = 2, UWS_CSTEP_CS = 3 };
int uws_cstep_set_imag(uws_handle h, int32_t what, int32_t face, const double* im, int64_t n);
int uws_cstep_get_imag(uws_handle h, int32_t what, int32_t face, double* im, int64_t n);

/* Discrete adjoint of nsteps solver steps (FP32 or FP64 handles, no
 * residual-based early exit, no low-precision mode, no pressure
 * normalisation).  uws_adjoint_forward runs the nsteps steps on the handle
 * from its current state, keeping (u, p) every ckpt_every steps; the handle
 * ends at the final state.  uws_adjoint_backward takes the adjoint of the
 * final state, lam_u = dJ/du ([cell][3]) and lam_p = dJ/dp ([cell]; NULL for
 * zero), and walks the steps backwards: each checkpoint segment is recomputed
 * with every iterate kept, and every step is reversed: the Jacobi sweeps as
 * transposed sweeps, the one-shot kernels through exact local Jacobians.
 * Gradients are then read with uws_adjoint_get_grad: UWS_GRAD_VEL0 (dJ/du at
 * the start, [cell][3]), UWS_GRAD_PRES0 ([cell]), UWS_GRAD_BND (boundary
 * texture `face`, [j][i][3]), UWS_GRAD_CS (one value). */
typedef struct uws_adjoint* uws_adjoint_handle;
enum { UWS_GRAD_VEL0 = 0, UWS_GRAD_PRES0 = 1, UWS_GRAD_BND = 2, UWS_GRAD_CS = 3 };
int uws_adjoint_create(uws_handle h, int32_t nsteps, int32_t ckpt_every, uws_adjoint_handle* out);
int uws_adjoint_destroy(uws_adjoint_handle a);
int uws_adjoint_forward(uws_adjoint_handle a);
int uws_adjoint_backward(uws_adjoint_handle a, const double* lam_u, int64_t n_u, const double* lam_p, int64_t n_p);
int uws_adjoint_get_grad(uws_adjoint_handle a, int32_t what, int32_t face, double* dst, int64_t n);
/* device bytes held by the adjoint (checkpoints, one segment of step
 * records, adjoint fields) */
int uws_adjoint_memory(uws_adjoint_handle a, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* UWS_H */
