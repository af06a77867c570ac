// CPU oracle of the ATJUP time step: a from-scratch C++/OpenMP restatement of the reference's
// algorithm, used only as the checker of the HIP library (tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg). It shares the C structs of include/atjup_hip.h and exports the
// same operators under the orc_ prefix. Each function cites the reference lines it restates.
#pragma once
#include "../include/atjup_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const atjup_grid* g, const atjup_consts* c, const atjup_knobs* k);
void     orc_destroy(orc_ctx* o);
int      orc_upload_field(orc_ctx* o, int field, const double* src);
int      orc_download_field(orc_ctx* o, int field, double* dst);
int      orc_upload_mask(orc_ctx* o, const double* seamount);
int      orc_upload_vel_star(orc_ctx* o, const double* src);
int      orc_download_vel_star(orc_ctx* o, double* dst);
int      orc_download_buoy_ref_level(orc_ctx* o, double* dst);
int      orc_rk4_step(orc_ctx* o);
int      orc_buoy_ref_level(orc_ctx* o);
int      orc_hydrostatic(orc_ctx* o);
int      orc_pressure_solve(orc_ctx* o, int sweeps, double omega);
int      orc_rb_colour_pass(orc_ctx* o, int colour, double omega, int planes);
int      orc_bc_after_step(orc_ctx* o);
int      orc_bc_radius(orc_ctx* o);
int      orc_bc_theta(orc_ctx* o);
int      orc_bc_phi(orc_ctx* o);
int      orc_bc_solid(orc_ctx* o);
int      orc_shapiro(orc_ctx* o, int field, int order, double strength, int passes, int axes_mask);
int      orc_mixture_density(orc_ctx* o);
int      orc_clamp_species(orc_ctx* o);
int      orc_swap_nlevel(orc_ctx* o);
int      orc_scalar_sync(orc_ctx* o);
int      orc_satadj(orc_ctx* o, int gas);
int      orc_set_gas(orc_ctx* o, int gas, const atjup_gas_params* p);
int      orc_turb_wall(orc_ctx* o);
int      orc_turbulence_run(orc_ctx* o);
int      orc_flux_limiter(orc_ctx* o);
int      orc_radiation(orc_ctx* o);
int      orc_set_radiation(orc_ctx* o, const atjup_rad_params* p);
int      orc_chemistry(orc_ctx* o);
int      orc_nh4sh_eq(orc_ctx* o);
int      orc_set_chemistry(orc_ctx* o, const atjup_chem_params* p);
int      orc_mca(orc_ctx* o);
int      orc_set_mca(orc_ctx* o, const atjup_mca_params* p);
int      orc_precipitation(orc_ctx* o);
int      orc_set_precipitation(orc_ctx* o, const atjup_precip_params* p);
int      orc_download_precip_surface(orc_ctx* o, double* dst);
int      orc_step(orc_ctx* o, int iter, int first);
int      orc_check_finite(orc_ctx* o);
void     orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
