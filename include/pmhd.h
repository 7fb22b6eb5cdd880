/* C ABI of the MI355X CTU+CT step (libpmhd.so).
 *
 * The boundary replaces PythonMHD's evolve functions
 *   evolve_1D_mhd  (Source/EvolveFunctions_MHD.py:195)
 *   evolve_2D_mhd  (Source/EvolveFunctions_MHD.py:393)
 *   evolve_3D_mhd  (Source/EvolveFunctions_MHD.py:750)
 * which Simulation.__run_sim dispatches (Source/Simulation.py:312-377).  The
 * reference has no C interface; the Python front-end pythonmhd_amd.evolve keeps
 * the reference's signatures and return tuples and calls these entry points
 * through ctypes.  Arrays crossing this ABI use the reference layout:
 *   cells  (8, ny, nx, nz)  z contiguous   [2D: (8, ny, nx); 1D: (8, nx)]
 *   faceBx (ny, nx+1, nz), faceBy (ny+1, nx, nz), faceBz (ny, nx, nz+1)
 * Host arrays are float64.  Status: 0 ok, negative on error (see
 * pmhd_last_error); no C++ exception crosses the ABI.  One handle owns one
 * device state; calls on one handle are not re-entrant.
 */
#ifndef PMHD_H
#define PMHD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMHD_OK 0
#define PMHD_EINVAL -1   /* bad shape, boundary condition or parameter */
#define PMHD_EHIP -2     /* a HIP runtime call failed */
#define PMHD_ENOMEM -3   /* device allocation failed */
#define PMHD_ENONFINITE -4 /* the state became non-finite during pmhd_advance (NaN timestep) */

typedef struct pmhd_grid {
  int32_t ndim;            /* 1, 2 or 3 */
  int32_t pad0;
  int64_t nx, ny, nz;      /* cells; unused extents are 1 */
  double dx, dy, dz;       /* SimulationGrid.dx/dy/dz (Source/SimulationGrid.py:121-126) */
  int32_t bcx, bcy, bcz;   /* 0 outflow, 1 periodic (Source/PythonMHD_Constants.py:10-11) */
  int32_t pad1;
} pmhd_grid;

typedef struct pmhd_params {
  double gamma, cfl;                     /* simPar gamma, cfl */
  double min_dens, min_pres, min_energy; /* floors (Source/PrimCons_HelperModule.py:499) */
  double epsilon;                        /* entropy-fix width */
  int32_t reconstruct_order;             /* 0 none, 1 PLM, 2 PPM */
  int32_t entropy_fix;                   /* 0/1 */
  int32_t precision;                     /* 64: FP64 throughout; 32: FP32 bulk, FP64 islands */
  int32_t physics;                       /* 0 ideal MHD (evolve_*D_mhd): cells (8, ...) + faces;
                                            1 hydrodynamics (evolve_*D_hydro,
                                            Source/EvolveFunctions_Hydro.py:232, 350, 559):
                                            cells (5, ...) = (rho, v/M, p/E), no faces */
  int32_t riemann;                       /* PMHD_RIEMANN_*: the flux of MHD faces (hydro: Roe only) */
  int32_t fix_quirks;                    /* PMHD_FIX_* bits: reference behaviours to correct instead
                                            of reproduce (0, the default, reproduces them all) */
} pmhd_params;

/* Riemann solvers.  ROE is the reference's (riemannSolver{X,Y,Z}_mhd_roe,
 * Source/RiemannSolvers_MHD_Roe_HelperModule.py:74-1559).  HLLE (Einfeldt's bounds
 * from the states' and the Roe average's fast speeds) and ROE_HLLE (Roe, switching
 * to HLLE at a face whose Roe intermediate states have a non-positive density or
 * pressure) are the solvers the reference's header names as missing (:24-30). */
#define PMHD_RIEMANN_ROE 0
#define PMHD_RIEMANN_HLLE 1
#define PMHD_RIEMANN_ROE_HLLE 2

/* Reference quirks (see the parity notes) that a grid may correct:
 * Z_BC: the 3D z reconstruction source terms and first-order z states use BcZ, not
 *   BcX (Source/EvolveFunctions_MHD.py:898, 904, 951);
 * OUTFLOW_GHOSTS: the outflow ghosts of the CTU terms follow the reference's own
 *   x-direction convention (edge cell, edge face) in every direction; the reference
 *   takes cell and face row 1 for the 2D y top ghost and face row ny-1 for its bottom
 *   (Source/CTU_HelperModule.py:377-387), face nz-1 for the 3D z forward ghost
 *   (:897-898) and flux row nx-1 for a 3D y bottom ghost (:724-725);
 * DTDY_80BIT: the x-direction 3D CTU update uses 0.5*dt/dy in double, not the
 *   reference's 80-bit long double (Source/CTU_HelperModule.py:455). */
#define PMHD_FIX_Z_BC 1
#define PMHD_FIX_OUTFLOW_GHOSTS 2
#define PMHD_FIX_DTDY_80BIT 4

typedef struct pmhd_handle pmhd_handle;

/* Allocate the device state and work buffers of one grid. */
int pmhd_create(const pmhd_grid* grid, const pmhd_params* params, pmhd_handle** out);
int pmhd_destroy(pmhd_handle* h);

/* Load the state the evolve functions read from the SimulationGrid
 * (primVars, consVars, faceBx/By/Bz; Source/EvolveFunctions_MHD.py:752-773).
 * Face pointers that a grid of lower dimension lacks may be NULL. */
int pmhd_set_state(pmhd_handle* h, const double* prim, const double* cons,
                   const double* fbx, const double* fby, const double* fbz);

/* Copy the current (which = 0) or second-last (which = 1) state back.  The
 * second-last state is the one before the last completed cycle, as returned by
 * the reference evolve functions (Source/EvolveFunctions_MHD.py:1156-1157). */
int pmhd_get_state(pmhd_handle* h, int32_t which, double* prim, double* cons,
                   double* fbx, double* fby, double* fbz);

/* pmhd_set_state / pmhd_get_state with DEVICE float64 arrays in the same
 * reference layout (for instance torch ROCm tensors of dtype float64): the
 * device-resident boundary.  The arrays stay owned by the caller (the library
 * never frees them); the conversion to the handle's storage (float, and the FP32
 * step's face field as high and low float parts) runs on the device, on the
 * handle's stream, and both calls synchronise that stream before returning.  No
 * PCIe traffic except, in 1D MHD, the uniform B_x (one value).  Same arguments and
 * errors as the host versions. */
int pmhd_set_state_device(pmhd_handle* h, const double* prim, const double* cons,
                          const double* fbx, const double* fby, const double* fbz);
int pmhd_get_state_device(pmhd_handle* h, int32_t which, double* prim, double* cons,
                          double* fbx, double* fby, double* fbz);

/* Launch the handle's work on the caller's HIP stream (a hipStream_t, for
 * instance torch.cuda.current_stream().cuda_stream), so that it is ordered with
 * the caller's own work on that stream; NULL returns to the handle's own stream.
 * Work queued so far is finished first.  The stream must outlive its use. */
int pmhd_set_stream(pmhd_handle* h, void* stream);

/* CFL timestep of the current state (calcDt_{1,2,3}D_mhd,
 * Source/TimestepCalculations_HelperModule.py:183-361). */
int pmhd_compute_dt(pmhd_handle* h, double* dt);

/* pmhd_compute_dt without a handle and without allocating: the CFL timestep
 * (calcDt_{1,2,3}D_mhd / _hydro, Source/TimestepCalculations_HelperModule.py:24-361)
 * of a state held in the caller's DEVICE float64 arrays in the reference layout
 * (cons and the faces the grid has), written to the caller's device double *dt.
 * No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  It calls no hipMalloc,
 * hipFree or allocating routine and does not synchronise: the two launches are
 * queued on `stream` (NULL: the default stream).  Its only device memory beyond
 * the inputs and *dt is a 24 KiB scratch array (1024 partial maxima) that each GPU
 * the library runs on holds until the process ends; every call on one GPU shares
 * it, so calls that overlap on one GPU race.  Computes in FP64 (params->precision
 * must be 64) and equals pmhd_compute_dt of a precision-64 handle on the same
 * state bit for bit.  Grids pmhd_create rejects (beyond 2^24 cells along an axis
 * or 2^29 cells with ghosts) are rejected with PMHD_EINVAL; the message is
 * pmhd_last_error(NULL). */
int pmhd_cfl_dt_noalloc(const pmhd_grid* grid, const pmhd_params* params, const double* cons,
                        const double* fbx, const double* fby, const double* fbz, double* dt,
                        void* stream);

/* One cycle with the given dt (the body of the evolve loop). */
int pmhd_step(pmhd_handle* h, double dt);

/* The evolve loop: cycles until t_end or max_cycles with the reference's
 * clipping of the last dt.  Returns the final time and cycle and the time of
 * the second-last state.  When the state becomes non-finite the loop does what
 * the reference's does (the cycle with the NaN timestep runs, the clock becomes
 * NaN and the loop ends), fills the outputs and returns PMHD_ENONFINITE. */
int pmhd_advance(pmhd_handle* h, double t0, double t_end, int64_t cycle0, int64_t max_cycles,
                 double* t_out, int64_t* cycle_out, double* t_second_last);

/* Where pmhd_advance keeps the clock (no counterpart in the reference, whose loop
 * is Source/EvolveFunctions_MHD.py:252-323, 465-701, 844-1163: a host loop that
 * reads dt every cycle).  PMHD_CLOCK_HOST: the host reads each cycle's CFL maxima
 * back and forms dt, one round trip per cycle.  PMHD_CLOCK_DEVICE: dt, its
 * clipping to t_end, the loop condition, the time and the cycle count stay on the
 * device (the grid runs as a batch of one box), so cycles are queued without a
 * round trip and the host looks at the clock every 16 cycles only to stop early;
 * the results are the same (bit for bit in FP64).  PMHD_CLOCK_AUTO (the default):
 * the device clock for grids of at most 2^16 cells, whose cycles are short enough
 * for the round trip to show (larger grids run faster in the single-grid kernels
 * than as a one-box batch).  PMHD_EINVAL for another mode. */
#define PMHD_CLOCK_AUTO 0
#define PMHD_CLOCK_HOST 1
#define PMHD_CLOCK_DEVICE 2
int pmhd_set_clock(pmhd_handle* h, int32_t mode);

/* Faces per thread of the fused predictor sweep along x, y and z (no counterpart
 * in the reference; a tuning knob).  With a run length s > 0 one launch per
 * direction reconstructs the cells (PLM slopes, or the PPM slopes of a separate
 * launch), traces the edge states and forms the predictor's interface states and
 * Roe fluxes for s consecutive faces per thread; 0 selects separate slope, edge
 * and predictor launches; -1 along the contiguous axis (z in 3D, x in 2D) gives
 * each line one workgroup whose threads reconstruct one cell each and pass edge
 * states through LDS.  Results are identical for every choice (bit for bit in
 * FP64).  The default is picked from the grid size and the device.  Lengths must
 * lie in [0, 4096], or be -1 along the contiguous axis (PMHD_EINVAL otherwise). */
int pmhd_set_sweep(pmhd_handle* h, int32_t seg_x, int32_t seg_y, int32_t seg_z);

/* Wait for all queued work of the handle. */
int pmhd_synchronize(pmhd_handle* h);

/* Diagnostics of the current state, all accumulated in double with a fixed
 * (deterministic) summation order: totals[8] = sums of the 8 conserved variables
 * over the cells; max_divb = max_cells |div B| * min(dx, dy, dz) / max |B_face|
 * (dimensionless; 0 in 1D); max_b = max |B_face|; nonfinite = number of
 * non-finite conserved values.  Any output pointer may be NULL.  The reference has
 * no counterpart (its divergence check is a TODO, Source/SimulationGrid_HelperModule.py:406). */
int pmhd_diagnostics(pmhd_handle* h, double* totals, double* max_divb, double* max_b,
                     int64_t* nonfinite);

/* Floor and health statistics of a grid.  The floors are the reference's
 * applyFloors (Source/PrimCons_HelperModule.py:499-554), which floors silently;
 * these counters record what it did.  Call sites, in the reference's call order:
 *   0, 1  x predictor left / right    (Source/EvolveFunctions_MHD.py:278-280, 513-515, 911-914)
 *   2, 3  y predictor top / bottom    (:517-519, 915-918)
 *   4, 5  z predictor back / forward  (:919-922)
 *   6, 7  2D post-CTU x left / right  (:562-565)
 *   8, 9  2D post-CTU y top / bottom  (:573-576)
 * (hydro: sites 0-5 at Source/EvolveFunctions_Hydro.py:286-288, 437-443, 675-685).
 * [site][0] counts densities raised to min_dens, [site][1] pressures raised to
 * min_pres (predictor) or energies raised to min_energy (post-CTU).  The reference's
 * 3D applyFloors raises IndexError when one call floors three or more values of one
 * variable (its message loop reuses its index, :539-543): floored_max >= 3 at a 3D
 * site marks a run the reference would not have completed. */
typedef struct pmhd_stats {
  int64_t cycles;              /* cycles since the state was set */
  int64_t floored[10][2];      /* values floored since the state was set */
  int64_t floored_max[10][2];  /* the most values one call site floored in one cycle */
  int64_t floored_total;       /* sum of floored[][] */
  int64_t nonfinite;           /* non-finite conserved values of the current state */
} pmhd_stats;

/* The MHD Riemann flux of n faces: riemannSolverX/Y/Z_mhd_roe
 * (Source/RiemannSolvers_MHD_Roe_HelperModule.py:74, 595, 1093) as one call, or
 * the HLLE / Roe-with-HLLE-fallback solvers (riemann = PMHD_RIEMANN_*).  Host
 * float64 arrays in the reference's directional 7-vector order (x: rho, vx, vy,
 * vz, By, Bz, p; y: ..., Bz, Bx, p; z: ..., Bx, By, p), each 7 planes of n faces:
 * wl, wr primitive, ul, ur conserved (E last), bn the n face-normal fields; flux
 * receives 7 planes.  direction 0/1/2 = x/y/z.  Computed with the step's own
 * flux code in the precision's arithmetic (64: bit for bit the reference's; 32:
 * the FP32 unit with its FP64 wave-speed islands), on the GPU through temporary
 * device buffers.  Errors: pmhd_last_error(NULL). */
int pmhd_riemann_flux(int32_t precision, int32_t direction, int32_t riemann, int64_t n,
                      const double* wl, const double* wr, const double* ul, const double* ur,
                      const double* bn, double gamma, int32_t entropy_fix, double epsilon,
                      double* flux);

/* Statistics of the current state (synchronises the handle's stream).  No
 * counterpart in the reference. */
int pmhd_get_stats(pmhd_handle* h, pmhd_stats* out);

/* Per-launch timing with HIP events on the handle's stream (measurement only).
 * enable = 1 clears the totals and starts recording; 0 stops. */
int pmhd_set_profiling(pmhd_handle* h, int32_t enable);

/* Totals recorded since profiling was enabled, one entry per kernel: name k at
 * names + 32*k (NUL-terminated), summed milliseconds, number of launches and
 * summed elements (cells, faces or edge points) of those launches. */
int pmhd_get_profile(pmhd_handle* h, int32_t max_entries, char* names, double* total_ms,
                     int64_t* launches, int64_t* elements, int32_t* n_entries);

/* Last error message of the handle (or, when h is NULL, of the last failed
 * pmhd_create, pmhd_batch_create or pmhd_cfl_dt_noalloc). */
const char* pmhd_last_error(const pmhd_handle* h);

/* ------------------------------------------------------------------------
 * Batches of independent boxes.  No counterpart in the reference, whose
 * Simulation runs one grid at a time (Source/Simulation.py:312-377); a batch is
 * the same step applied to nbox grids at once.  All boxes share ndim and the
 * parameters; sizes, widths and boundary conditions may differ per box.  Every
 * phase of a cycle is one launch over the tiles of all boxes; each box keeps
 * its own CFL timestep and clock on the device, so pmhd_batch_advance runs
 * many cycles without a host round trip.  Box b evolves exactly as a single
 * handle with the same grid, parameters and state would (bit for bit).
 * ------------------------------------------------------------------------ */
typedef struct pmhd_batch pmhd_batch;

int pmhd_batch_create(int32_t nbox, const pmhd_grid* grids, const pmhd_params* params,
                      pmhd_batch** out);
int pmhd_batch_destroy(pmhd_batch* b);

/* pmhd_set_state / pmhd_get_state of one box. */
int pmhd_batch_set_state(pmhd_batch* b, int32_t box, const double* prim, const double* cons,
                         const double* fbx, const double* fby, const double* fbz);
int pmhd_batch_get_state(pmhd_batch* b, int32_t box, int32_t which, double* prim, double* cons,
                         double* fbx, double* fby, double* fbz);

/* The evolve loop of every box from (t0, cycle0): each box runs until its own
 * time reaches t_end or its cycle count max_cycles.  t_out, cycle_out and
 * t_second_last receive one value per box (each may be NULL). */
int pmhd_batch_advance(pmhd_batch* b, double t0, double t_end, int64_t cycle0, int64_t max_cycles,
                       double* t_out, int64_t* cycle_out, double* t_second_last);

/* Tile tuner: for every launch of a cycle, times each candidate tile shape
 * (threads along z by threads along x, one block per tile or a persistent grid)
 * reps times on this batch's boxes and keeps the fastest.  The current state is
 * unchanged; the second-last state is dropped. */
int pmhd_batch_autotune(pmhd_batch* b, int32_t reps);

/* The tile of each tuned launch: name k at names + 32*k, threads along z (bk)
 * and along x (bi), grid (0 = one block per tile) and the tuned time in ms. */
int pmhd_batch_get_tiles(pmhd_batch* b, int32_t max_entries, char* names, int32_t* bk,
                         int32_t* bi, int32_t* grid, double* ms, int32_t* n_entries);

int pmhd_batch_synchronize(pmhd_batch* b);
int pmhd_batch_set_profiling(pmhd_batch* b, int32_t enable);
int pmhd_batch_get_profile(pmhd_batch* b, int32_t max_entries, char* names, double* total_ms,
                           int64_t* launches, int64_t* elements, int32_t* n_entries);
const char* pmhd_batch_last_error(const pmhd_batch* b);
/* pmhd_get_stats of one box. */
int pmhd_batch_get_stats(pmhd_batch* b, int32_t box, pmhd_stats* out);

/* ------------------------------------------------------------------------
 * Adjoint of a run of fixed-dt steps, with checkpoint-and-recompute.  No
 * counterpart in the reference (which has no gradients).  forward() runs nsteps
 * cycles of the given dt from a state (the reference's loop body with dt fixed:
 * no CFL choice, no clipping) and keeps a checkpoint every checkpoint_every
 * steps; backward() maps the cotangent of the final state (prim, cons, faces;
 * any pointer may be NULL for a zero cotangent) to the cotangent of the initial
 * state of the last forward().  Arrays are host float64 in the reference layout.
 * Cotangents are accumulated in float64 on the device.
 * ------------------------------------------------------------------------ */
typedef struct pmhd_adjoint pmhd_adjoint;

/* 2D and 3D MHD and 1D-3D hydrodynamics (PMHD_EINVAL for 1D MHD, whose uniform
 * B_x is a scalar of the step that the adjoint does not differentiate). */
int pmhd_adjoint_create(const pmhd_grid* grid, const pmhd_params* params, int32_t checkpoint_every,
                        pmhd_adjoint** out);
int pmhd_adjoint_destroy(pmhd_adjoint* a);
int pmhd_adjoint_forward(pmhd_adjoint* a, const double* prim, const double* cons, const double* fbx,
                         const double* fby, const double* fbz, double dt, int64_t nsteps,
                         double* prim_out, double* cons_out, double* fbx_out, double* fby_out,
                         double* fbz_out);
int pmhd_adjoint_backward(pmhd_adjoint* a, const double* gprim, const double* gcons,
                          const double* gfbx, const double* gfby, const double* gfbz,
                          double* gprim0, double* gcons0, double* gfbx0, double* gfby0,
                          double* gfbz0);
/* forward / backward with DEVICE float64 arrays of the same layouts (see
 * pmhd_set_state_device): the torch.autograd wrapper passes its ROCm tensors. */
int pmhd_adjoint_forward_device(pmhd_adjoint* a, const double* prim, const double* cons,
                                const double* fbx, const double* fby, const double* fbz, double dt,
                                int64_t nsteps, double* prim_out, double* cons_out, double* fbx_out,
                                double* fby_out, double* fbz_out);
int pmhd_adjoint_backward_device(pmhd_adjoint* a, const double* gprim, const double* gcons,
                                 const double* gfbx, const double* gfby, const double* gfbz,
                                 double* gprim0, double* gcons0, double* gfbx0, double* gfby0,
                                 double* gfbz0);
/* pmhd_set_stream for the adjoint's launches. */
int pmhd_adjoint_set_stream(pmhd_adjoint* a, void* stream);
/* Deterministic accumulation (enable != 0, the default): every backward() of the
 * same forward gives the same bits.  Each launch's contributions are written to
 * their own places and summed per target in a fixed order.  The first backward
 * builds that order once per launch of a step (a sort of the targets, with about
 * 5.5 KB of scratch per element of the largest launch, released when it ends) and
 * keeps it with the handle together with the launch's slot lists (about 0.5 KB per
 * element and launch of a step), beside a buffer of 1.3 KB per element for the
 * contributions.  Up to 13.4M elements per launch.  enable = 0 adds with FP64
 * atomics instead (summation order, and so the last bits, vary between runs).  No
 * counterpart in the reference, which has no adjoint. */
int pmhd_adjoint_set_deterministic(pmhd_adjoint* a, int32_t enable);
int pmhd_adjoint_set_profiling(pmhd_adjoint* a, int32_t enable);
int pmhd_adjoint_get_profile(pmhd_adjoint* a, int32_t max_entries, char* names, double* total_ms,
                             int64_t* launches, int64_t* elements, int32_t* n_entries);
const char* pmhd_adjoint_last_error(const pmhd_adjoint* a);

#ifdef __cplusplus
}
#endif
#endif
