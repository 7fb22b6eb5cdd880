// Host test build of pmhd_cfl_dt_noalloc: the same validation and the same
// timestep as the HIP unit, with the arrays in host memory (test infrastructure).
#define PMHD_TU_NS noalloc
#include "pmhd_host.h"

extern "C" int pmhd_cfl_dt_noalloc(const pmhd_grid* grid, const pmhd_params* params, const double* cons,
                                   const double* fbx, const double* fby, const double* fbz, double* dt,
                                   void* stream) {
  using namespace pmhd;
  (void)stream;
  if (!grid || !params || !cons || !dt) {
    pmhd_note_null_error("pmhd_cfl_dt_noalloc: grid, params, cons and dt are required");
    return PMHD_EINVAL;
  }
  const char* problem = grid_problem(grid);
  if (!problem) problem = params_problem(params);
  if (!problem && params->precision != 64)
    problem = "pmhd_cfl_dt_noalloc computes in FP64 from float64 arrays: precision must be 64";
  const bool mhd = params->physics == 0;
  if (!problem && mhd && grid->ndim >= 2 && (!fbx || !fby || (grid->ndim == 3 && !fbz)))
    problem = "pmhd_cfl_dt_noalloc: the face arrays of the grid's dimensions are required";
  if (problem) {
    pmhd_note_null_error(problem);
    return PMHD_EINVAL;
  }
  Ctx<double> c;
  init_ctx(c, normal_grid(*grid), *params);
  c.U = cons;
  c.fb[0] = fbx;
  c.fb[1] = fby;
  c.fb[2] = fbz;
  double m[3];
  host_cfl(c, m);
  *dt = dt_of(c, params->cfl, m);
  return PMHD_OK;
}
