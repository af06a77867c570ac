/*
 * atjup_hip.h - C ABI of the MI355X (gfx950) implementation of the ATJUP time step.
 *
 * The reference has no FFI for this path: its entry points are C++ member functions of
 * cJupiterModel and class templates instantiated on it (planet/cJupiterModel.h:671-672,
 * planet/PressureSolverJup.h, planet/SaturationAdjustmentJup.h, planet/TurbulenceJup.h,
 * planet/BC_Jup.h), all acting on model-owned Array fields in place. This header is the drop-in
 * boundary that replaces them: one opaque context that owns the device-resident fields, transfer
 * calls that accept the reference's Array row layout unchanged, and one operator per reference
 * entry point with the reference's semantics. Each declaration names the reference code it
 * replaces.
 *
 * Conventions
 *   - Every call returns int: 0 = ok, < 0 = error (atjup_hip_last_error_string() says why),
 *     ATJUP_NONFINITE (> 0) = the state holds a non-finite value.
 *   - No C++ exceptions cross this boundary.
 *   - One context per HIP stream; calls come from one host thread. State stays on the device
 *     between calls and is packed only by the download calls.
 *   - Field layout on the host: [i][j][k], k (phi) fastest, im*jm*km doubles, the reference's
 *     Array order (lib/Array.h:25-31). The _rows variants take the reference's double*** x.
 */
#ifndef ATJUP_HIP_H
#define ATJUP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ATJUP_OK             0
#define ATJUP_ERR_ARG       -1
#define ATJUP_ERR_HIP       -2
#define ATJUP_ERR_PROFILE   -3
#define ATJUP_ERR_NOMEM     -4
#define ATJUP_ERR_STATE     -5
#define ATJUP_NONFINITE      1

#define ATJUP_PRECISION_FP64      0   /* the drop-in: every field and stencil in FP64 */
#define ATJUP_PRECISION_FP32_REF  1   /* fields and RK4 in FP32 (t, u, v, w and the species as FP32
                                         pairs hi + lo); Poisson in FP32 inside FP64 refinement */
#define ATJUP_PRECISION_TANGENT   2   /* FP64 with a directional derivative carried along every field
                                         (forward-mode tangent of the dynamics profile; no counterpart
                                         in the reference: it verifies the adjoint) */

/* Field identifiers. 0..16 are the 17 prognostic fields the RK4 integrates
 * (planet/RungeKutta_Jup_Turb.cpp:425-486); ATJUP_NLEVEL + id is the time-level-n copy of the same
 * field (tn, un, ... in planet/cJupiterModel.h:1383-1398, 1513-1516). */
enum atjup_field {
    ATJUP_F_T = 0, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W,
    ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, ATJUP_F_H2S,
    ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_NH4SH,
    ATJUP_F_TKE, ATJUP_F_DIS,
    ATJUP_N_PROG,                                   /* 17 */
    /* inputs and diagnostics the path reads or writes */
    ATJUP_F_P_DYN = 20, ATJUP_F_P_STAT, ATJUP_F_P_HYDRO, ATJUP_F_RHO_MIX, ATJUP_F_WALL_NUE,
    ATJUP_F_NUE, ATJUP_F_PROD, ATJUP_F_TKE_SOURCE, ATJUP_F_DIS_SOURCE,
    ATJUP_F_AUX_U, ATJUP_F_AUX_V, ATJUP_F_AUX_W,
    ATJUP_F_Q_RAD, ATJUP_F_Q_PRECIP,
    ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_H2O_CLOUD, ATJUP_F_S_PRECIP_H2O_ICE,
    ATJUP_F_S_PRECIP_NH3, ATJUP_F_S_PRECIP_NH3_CLOUD, ATJUP_F_S_PRECIP_NH3_ICE,
    ATJUP_F_S_PRECIP_CH4, ATJUP_F_S_PRECIP_CH4_CLOUD, ATJUP_F_S_PRECIP_CH4_ICE,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH,
    ATJUP_F_EPSILON, ATJUP_F_RADIATION,            /* the radiation's layer emissivity and net flux */
    ATJUP_F_P_RAIN, ATJUP_F_P_SNOW, ATJUP_F_P_GRAUPEL,              /* the precipitation's fluxes */
    ATJUP_F_P_NH3_RAIN, ATJUP_F_P_NH3_SNOW, ATJUP_F_P_NH3_GRAUPEL,
    ATJUP_F_P_CH4_RAIN, ATJUP_F_P_CH4_SNOW, ATJUP_F_P_CH4_GRAUPEL, ATJUP_F_P_NH4SH,
    ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3, ATJUP_F_THERMALMASSFLUX,   /* the chemistry's carried fields */
    ATJUP_F_DIFFLUX_NH4SH,
    ATJUP_F_END,
    ATJUP_NLEVEL = 64                               /* ATJUP_NLEVEL + prognostic id */
};

/* The grid. rad.z and the.z are cumulative sums from r0 and 0 (lib/Array_1D.cpp:49-60). */
typedef struct atjup_grid {
    int im, jm, km;          /* 41, 181, 361 for the reference (planet/cJupiterModel.h:166) */
    double r0;               /* rad.z[0] = metric radius / L_atm (planet/cJupiterModel.cpp:441-468) */
    double dr, dthe, dphi;   /* 0.025, 1 deg, 1 deg (planet/cJupiterModel.cpp:160-177) */
    double L_atm;            /* shell depth [km] */
} atjup_grid;

/* Physical constants; the reference's config members and mixture constants. */
typedef struct atjup_consts {
    double dt, re, pr, g, omega, u_0, t_ref;
    double Coriolis, centrifugal, buoyancy, chemical_reaction;
    double r_mix, R_mix, cp_mix, mue_mix;
    double sc_h2o, sc_h2s, sc_nh3, sc_nh4sh, sc_ch4;
    double rg_nh4sh, r_p_nh4sh, re_turb;
} atjup_consts;

/* The knob profile: the reference's environment knobs that select branches on this path, with the
 * values they take in the supported profile. ctx_create rejects a profile it does not implement. */
typedef struct atjup_knobs {
    int    turb_model;          /* ATJUP_TURB/_TURB_MODEL: -1 off, 0 k-epsilon, 1 k-omega, 2 SST */
    int    press_sweeps;        /* ATJUP_PRESS_SWEEPS (PressureSolver.h:174) */
    double press_omega;         /* over-relaxation of the red-black sweep; 1 = the reference's GS */
    double pdyn_cap;            /* ATJUP_PDYN_CAP 0.2 (PressureSolver.h:168) */
    int    damp_pdyn;           /* ATJUP_DAMP_PDYN, default !GEO_BALANCE (cJupiterModel.cpp:1040) */
    double rho_pdyn;            /* ATJUP_RHO_PDYN weight, default !GEO_BALANCE (InitValues_Jup.cpp:722) */
    double sinthe_min;          /* ATJUP_SINTHE_MIN 0.55 (cJupiterModel.h:1121-1128) */
    double t_diff_scale;        /* ATJUP_T_DIFF_SCALE 0.01 */
    double species_diff_scale;  /* ATJUP_SPECIES_DIFF_SCALE 0.01 */
    double rad_coupling;        /* ATJUP_RAD_COUPLING 1 */
    double precip_coupling;     /* ATJUP_PRECIP_COUPLING 1 */
    double turb_coupling;       /* ATJUP_TURB_COUPLING 1 */
    double tke_max_phys;        /* ATJUP_TKE_MAX 1000 m2/s2 */
    double nue_max_phys;        /* ATJUP_NUE_MAX 1e5 m2/s */
    double wall_nue_factor;     /* ATJUP_WALL_NUE 4 */
    int    wall_nue_layers;     /* ATJUP_WALL_NUE_LAYERS 3 */
    int    satadj;              /* run the saturation adjustment in the physics block */
    int    rk_scalar_sync;      /* ATJUP_RK_SCALAR_SYNC 2 */
    int    dry;                 /* species identically zero with zero sources: they are not streamed */
    double turb_k_bg_phys;      /* ATJUP_TURB_K_BG [m2/s2] */
    int    radiation;           /* ATJUP_RADIATION: the grey radiation in the physics block (the
                                   reference's default is 1, cJupiterModel.cpp:52, :1056) */
    int    precip;              /* ATJUP_PRECIP: the precipitation microphysics after the saturation
                                   adjustments (default 1, cJupiterModel.cpp:83, :1090); with
                                   precip_coupling != 0 it reports rates into S_precip_* */
    int    chemistry;           /* the physics block's chemistry (DiffMassFluxJup, ChemMassRateJup and
                                   their smoothing, cJupiterModel.cpp:1146-1161, which the reference
                                   always runs) in the step; 0: massflux_* enter as frozen inputs */
    int    mca;                 /* ATJUP_MCA=<N>: the moist convective adjustment every N iterations,
                                   after the closure's wall hold (cJupiterModel.cpp:115-125, :1301-1322;
                                   the reference's default is 1); 0 off */
    int    nh4sh_eq;            /* ATJUP_NH4SH_EQ: the NH4SH equilibrium adjustment after the saturation
                                   adjustments (ChemistryJup::EquilibriumNH4SH, ChemistryJup.h:285-370,
                                   cJupiterModel.cpp:1084; the reference's default is 1): the chemistry's
                                   rate law then contributes nothing, and the moist convective
                                   adjustment carries the NH4SH heat of formation */
} atjup_knobs;

/* The chemistry's constants as the reference's start-up leaves them (planet/cJupiterModel.h:172-401:
 * molar masses [kg/kmol], mole fractions, dynamic viscosities, densities, the mixture conductivity,
 * specific heats, conductivities, the NH4SH reaction enthalpy, the NH4SH band [K]) and its knobs
 * ATJUP_CHEM_GATE_ZERO (1), ATJUP_CHEM_MOLAR_CONC (0), ATJUP_CHEM_ENTHALPY (0)
 * (planet/ChemistryJup.h:54-57, :123-126). The molar masses m_nh3, m_h2s, m_nh4sh also serve the
 * NH4SH equilibrium adjustment (atjup_knobs.nh4sh_eq). */
typedef struct atjup_chem_params {
    double m_nh3, m_h2s, m_nh4sh, m_h2, m_he, m_h2o;
    double X_h2, X_he, X_h2o, X_h2s, X_nh3, X_nh4sh;
    double mue_nh3, mue_h2s, mue_nh4sh, rg_nh3, rg_h2s, k_mix;
    double cp_nh3, cp_h2s, cp_nh4sh, k_nh3, k_h2s, k_nh4sh, dh_nh4sh, t_0_nh4sh, t_00_nh4sh;
    int    gate_zero, molar_conc, enthalpy;
} atjup_chem_params;

/* The moist convective adjustment's knobs and the constants beyond the gases' (H2O and NH3 take
 * atjup_hip_set_gas's values): ATJUP_MCA_LEDOUX (1), ATJUP_MCA_LOADING (1), ATJUP_MCA_RELAX (1, in
 * [0, 1]), ATJUP_MCA_PASSES (64) (planet/ConvectiveAdjustment.h:409-413), the H2S tracer's ep_h2s
 * (InitValues_Jup.cpp:945) and the NH4SH tracer's heat of formation (0 unless ATJUP_NH4SH_EQ,
 * planet/cJupiterModel.cpp:1316; a context created with atjup_knobs.nh4sh_eq starts with the
 * equilibrium's 4705 ln(10) R* / m_nh4sh J/kg). The supported settings are ATJUP_MCA_SMOOTH 0, ATJUP_MCA_PHASE 0,
 * ATJUP_MCA_INTERIOR 1, ATJUP_CP_H2 and ATJUP_CP_GAS 0 and the chromophore off (ATJUP_CHROM 0), which
 * leaves that tracer zero everywhere. */
typedef struct atjup_mca_params {
    double relax, ep_h2s, l_nh4sh;
    int    ledoux, loading, max_pass;
} atjup_mca_params;

/* The precipitation's constants beyond the gases' (which atjup_hip_set_gas supplies):
 * ATJUP_PRECIP_SCALE on the five rate coefficients (planet/Precipitation.h:111-115,
 * cJupiterModel.h:1059-1061) and the NH4SH crystal band [t_00_nh4sh, t_0_nh4sh] in K
 * (planet/cJupiterModel.h:194-195). */
typedef struct atjup_precip_params {
    double rate_scale, t_0_nh4sh, t_00_nh4sh;
} atjup_precip_params;

/* The grey two-stream radiation's planet properties (planet/cJupiterModel.h:1109-1118) and its
 * runtime knobs, read by Radiation<P>::run under the ATJUP_ prefix (planet/Radiation.h:270-295):
 * CIA_STRENGTH, OPACITY_STRENGTH, SOLAR, SOLAR_STRENGTH, SW_TAU_PER_BAR, RAD_BOTTOM. */
typedef struct atjup_rad_params {
    double F_int, S_solar, albedo_bond, x_H2, x_He;
    double cia_strength, opacity_strength, solar_strength, sw_tau_per_bar;
    int    solar, bottom;
} atjup_rad_params;

/* The saturation adjustment's constants of one gas, the arguments of
 * SaturationAdjustment<P>::run (planet/SaturationAdjustment.h:236-243) as the time loop passes them
 * (planet/cJupiterModel.cpp:1060-1080): triple point and ice onset [K], ep, latent heats [J/kg],
 * and the Sanchez-Lavega liquid and ice vapour-pressure sets (planet/ATPhys.h:49-52). */
typedef struct atjup_gas_params {
    double t_0, t_00, ep, lv, ls;
    double C, L0, R, del_alf, del_bet;
    double C_i, L0_i, del_alf_i, del_bet_i;
} atjup_gas_params;

typedef struct atjup_ctx atjup_ctx;

/* ---- context ---------------------------------------------------------------------------- */
/* Replaces the state cJupiterModel owns (resetArrays, planet/cJupiterModel.cpp:1570-1753) and the
 * start-up part of Run() that fixes geometry and constants (planet/cJupiterModel.cpp:316-745).
 * precision: ATJUP_PRECISION_FP64 or ATJUP_PRECISION_FP32_REF. Returns NULL on error. */
atjup_ctx* atjup_hip_ctx_create(const atjup_grid* grid, const atjup_consts* consts,
                                const atjup_knobs* knobs, int precision);
void       atjup_hip_ctx_destroy(atjup_ctx* ctx);
/* The default knob profile (the reference's defaults with GEO_BALANCE, VORTEX, MCA off). */
void       atjup_hip_default_knobs(atjup_knobs* knobs);

/* ---- field transfer --------------------------------------------------------------------- */
/* Contiguous [i][j][k] doubles. The FP32 form converts on the way in and out. */
int atjup_hip_upload_field(atjup_ctx* ctx, int field, const double* src);
int atjup_hip_download_field(atjup_ctx* ctx, int field, double* dst);
/* The reference's own Array storage, one new double[km] per (i, j) row (lib/Array.cpp:43-73). */
int atjup_hip_upload_field_rows(atjup_ctx* ctx, int field, double* const* const* rows);
int atjup_hip_download_field_rows(atjup_ctx* ctx, int field, double** const* rows);
/* The tangent form (ATJUP_PRECISION_TANGENT): the directional derivative carried with a field.
 * upload_field sets a field's value and clears its derivative; upload_tangent sets the
 * derivative and keeps the value. No counterpart in the reference (verification of the adjoint). */
int atjup_hip_upload_tangent(atjup_ctx* ctx, int field, const double* src);
int atjup_hip_download_tangent(atjup_ctx* ctx, int field, double* dst);
/* The obstacle: SeaMount (1.0 = solid, planet/BC_Jup.h:353-457) as doubles [i][j][k]. Derives
 * i_topography and the wall eddy viscosity (computeWallViscosity, InitValues_Jup.cpp:84-134). */
int atjup_hip_upload_mask(atjup_ctx* ctx, const double* seamount);
/* The friction velocity vel_star (Array_2D [j][k], planet/cJupiterModel.h:1521). */
int atjup_hip_upload_vel_star(atjup_ctx* ctx, const double* src);
int atjup_hip_download_vel_star(atjup_ctx* ctx, double* dst);
/* The per-level buoyancy mean of the last RK4 step (buoy_ref_level[im]). */
int atjup_hip_download_buoy_ref_level(atjup_ctx* ctx, double* dst);

/* ---- operators -------------------------------------------------------------------------- */
/* cJupiterModel::RungeKuttaJup (planet/RungeKutta_Jup_Turb.cpp:261-583) with its two per-step
 * pre-passes computeBuoyancyRefLevel (:47-69) and computeHydrostaticPressure (:119-259) and the
 * per-cell RHSJup (planet/RHS_Jup_Turb.cpp:31-1457). */
int atjup_hip_rk4_step(atjup_ctx* ctx);
/* PressureSolver<cJupiterModel>::run (planet/PressureSolver.h:145-766) with
 * prepareProjectionBoundaries (planet/cJupiterModel.cpp:1842-1864). sweeps <= 0 takes the
 * profile's count; omega = 1 is the reference's Gauss-Seidel. */
int atjup_hip_pressure_solve(atjup_ctx* ctx, int sweeps, double omega);
/* SaturationAdjustment<cJupiterModel>::run for one gas (planet/SaturationAdjustment.h:255-746),
 * gas 0 = H2O (h2o, h2o_cloud, h2o_ice), 1 = NH3, 2 = CH4, with the constants the reference passes
 * as its 14 double arguments; params = NULL uses the context's set (atjup_hip_set_gas). */
int atjup_hip_satadj(atjup_ctx* ctx, int gas, const atjup_gas_params* params);
/* The constants atjup_hip_step hands the adjustment, gas by gas in the reference's order H2O, NH3,
 * CH4 (planet/cJupiterModel.cpp:1060-1080). Until set, a gas's adjustment fails with
 * ATJUP_ERR_STATE. */
int atjup_hip_set_gas(atjup_ctx* ctx, int gas, const atjup_gas_params* params);
/* Radiation<cJupiterModel>::run (planet/Radiation.h:245-487, planet/RadiationJup.h), as the
 * physics block calls it after computeMixtureDensity (planet/cJupiterModel.cpp:1056): the layer
 * emissivity (epsilon), the layer-centre net thermal flux (radiation) and the radiative heating
 * Q_rad of every column, from t, p_stat, rho_mix and the species. The printed column diagnostics
 * (mean OLR, the tau = 1 level) are not formed. */
int atjup_hip_radiation(atjup_ctx* ctx);
/* The properties and knobs atjup_hip_radiation uses; a new context holds Jupiter's values and the
 * knobs' defaults (atjup_hip_default_rad_params). */
int atjup_hip_set_radiation(atjup_ctx* ctx, const atjup_rad_params* params);
void atjup_hip_default_rad_params(atjup_rad_params* params);
/* Precipitation<cJupiterModel>::run (planet/Precipitation.h:194-574, planet/PrecipitationJup.h), as
 * the physics block calls it after the saturation adjustments (planet/cJupiterModel.cpp:1090): the
 * three-category (rain, snow, graupel) column scheme of H2O, NH3 and CH4 with the gases' constants
 * of atjup_hip_set_gas, the latent heating Q_precip, the moisture rates S_precip_* (coupling on) or
 * the depleted condensate (coupling off), the NH4SH crystal settling flux and the surface maps. The
 * printed diagnostics (ATJUP_PRECIP_DIAG) are not formed. */
int atjup_hip_precipitation(atjup_ctx* ctx);
int atjup_hip_set_precipitation(atjup_ctx* ctx, const atjup_precip_params* params);
void atjup_hip_default_precip_params(atjup_precip_params* params);
/* The surface precipitation maps the last call left (precip_srf_h2o, _nh3, _ch4, _nh4sh, _total,
 * planet/Precipitation.h:470-492), each [jm][km], in that order: 5 * jm * km doubles. */
int atjup_hip_download_precip_surface(atjup_ctx* ctx, double* dst);
/* The physics block's chemistry in the reference's order (planet/cJupiterModel.cpp:1146-1161):
 * ChemistryJup::DiffMassFluxJup (planet/ChemistryJup.h:375-540), damp_wiggles of difflux_h2s,
 * difflux_nh3, difflux_nh4sh and thermalmassflux, ChemMassRateJup (:59-213), FluxLimiterNH4SH,
 * damp_wiggles of massflux_h2s, massflux_nh3, massflux_nh4sh and fluxlim_nh4sh. The reaction
 * rates w_* persist in the context between calls, as the reference's arrays do. */
int atjup_hip_chemistry(atjup_ctx* ctx);
int atjup_hip_set_chemistry(atjup_ctx* ctx, const atjup_chem_params* params);
void atjup_hip_default_chem_params(atjup_chem_params* params);
/* MoistConvectiveAdjustment<cJupiterModel>::run (planet/ConvectiveAdjustment.h:403-849) with the gases
 * H2O and NH3 and the tracers H2S and NH4SH (planet/cJupiterModel.cpp:1301-1322): per (j, k) column
 * of fluid cells (the radial boundary planes left out) the parcel test of every adjacent pair
 * (Ledoux: virtual temperature with the condensate loading), then sweeps that mix h_L, the total
 * H2O / NH3 and the tracers over each unstable segment (grown while its ends stay unstable) and
 * bring its cells to equilibrium, until a sweep finds nothing or the sweep cap; the touched cells
 * get t, the vapour, cloud and ice of H2O and NH3, H2S and NH4SH. A column with a non-physical
 * state, or whose equilibrium solve fails, keeps its state. The printed census is not formed. */
int atjup_hip_mca(atjup_ctx* ctx);
/* Reporting<cJupiterModel>::searchMinMax_3D (planet/Reporting.h:261-297) on the device: the
 * largest and smallest value of a field and their cells, the first cell of the reference's j, k, i
 * loops on ties (a NaN is never taken, unless it is the loops' starting cell (0, 0, 0), which then
 * stands for both). out[0] max, out[1] min (the raw values: the caller applies the printout's
 * coefficient and conversion); loc = {i, j, k} of the max, then of the min. FP32-refined fields
 * carried as pairs (t, u, v, w and the species) are read as hi + lo. */
int atjup_hip_search_minmax_3d(atjup_ctx* ctx, int field, double* out, int* loc);
/* The search of atjup_hip_search_minmax_3d, allocating nothing. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). field is a caller's FP64 device array [im][jm][pitch] (the library's layout, row
 * pitch >= km, the elements km..pitch-1 of a row ignored); the launches go to `stream` (a hipStream_t
 * of the library's HIP runtime, NULL the null stream), which is synchronized before out and loc
 * (as in atjup_hip_search_minmax_3d) are written. The partial results live in a fixed device array
 * of the library (about 64 KiB per GPU, held until the process ends); the field may hold at most
 * im <= 4096, jm <= 65536, pitch <= 65536 and 2^36 elements in all, larger fields are refused with
 * ATJUP_ERR_ARG. Calls that overlap on one GPU share that array and race. */
int atjup_hip_search_minmax_3d_noalloc(const double* field, int im, int jm, int km, int pitch,
                                       void* stream, double* out, int* loc);
/* searchMinMax_2D (planet/Reporting.h:299-330) over map 0-4 (the surface precipitation maps in the
 * order of atjup_hip_download_precip_surface) or 5 (vel_star): the starting element (0, 0) and the
 * interior j in [1, jm-2], k in [1, km-2]; loc = {j, k} of the max, then of the min. */
int atjup_hip_search_minmax_2d(atjup_ctx* ctx, int map, double* out, int* loc);
/* The continuity residual of steadyQuery (planet/Reporting.h:372-398): the largest |du/dr + 2u/r +
 * dv/dtheta / r + cos(theta) v / (r sin(theta)) + dw/dphi / (r sin(theta))| over the interior and its
 * cell, the last cell of the i, j, k loops on ties (ATJUP_COSTHE_ABS 0). With no finite residual the
 * value is 0 and loc is {-1, -1, -1} (the reference keeps its previous cell). */
int atjup_hip_continuity_residual(atjup_ctx* ctx, double* value, int* loc);
/* steadyQuery's per-field changes (planet/Reporting.h:346-410): max |f - f_n| over every cell with
 * its cell (the last of the i, j, k loops on ties, >= from 0; NaNs never taken), for the twelve
 * fields of the report in its order: p_dyn, u, v, w, t, h2o, h2o_cloud, h2o_ice, nh3, nh3_cloud,
 * nh3_ice, nh4sh. This path commits the n level in place (restoreVar is the identity here), so
 * with tracking on the library keeps the n-level copies the report differences against: the
 * state at the start of each iteration, and for t and the six condensables (NH4SH with the
 * equilibrium on) the state RK_SCALAR_SYNC copies after the physics block's adjustments
 * (planet/cJupiterModel.cpp:1118-1140), as the reference's f_n hold them when steadyQuery runs
 * before restoreVar (:1378). Tracking costs twelve FP64 fields of device memory and their copies
 * per iteration. values[12], cells[12][3] = {i, j, k}; with no finite change the entry is 0 at
 * (0, 0, 0), the reference's starting values. */
#define ATJUP_N_STEADY 12
int atjup_hip_steady_track(atjup_ctx* ctx, int on);
int atjup_hip_steady_changes(atjup_ctx* ctx, double* values, int* cells);
/* One plane of a field for the writers (planet/ParaView_Jup.cpp): axis 0 the level i = index
 * ([jm][km]), axis 1 the latitude j = index ([im][km]), axis 2 the longitude k = index ([im][jm]),
 * as FP64, gathered on the device so that only the plane crosses PCIe. */
int atjup_hip_download_slice(atjup_ctx* ctx, int field, int axis, int index, double* dst);
int atjup_hip_set_mca(atjup_ctx* ctx, const atjup_mca_params* params);
void atjup_hip_default_mca_params(atjup_mca_params* params);
/* cJupiterModel::save_state / load_state (planet/FileIO_Jup.cpp:584-684): the binary restart file,
 * byte for byte the reference's layout: int32 {magic, im, jm, km, iter}, then the 38 fixed 3D arrays
 * of restart_arrays() (:281-293), in JUP2 files the 17 carried ones of restart_arrays_carried()
 * (:297-305) and vel_star, each [i][j][k] (or [j][k]) FP64. carried = 1 writes "JUP2" (the
 * reference's default ATJUP_RESTART_CARRIED=1), 0 writes "JUP1". The n-level copies are written
 * from the live fields, as the reference writes them after restoreVar. The FP32-refined form
 * writes its values widened to FP64 (t, u, v, w and the species as hi + lo).
 * load_state reads either magic, rejects a file of another grid or a truncated one
 * (ATJUP_ERR_STATE), and rejects n-level arrays that differ from their fields: this path keeps one
 * copy. iter receives the header's iteration; carried_out whether the file held the carried set. */
int atjup_hip_save_state(atjup_ctx* ctx, const char* path, int iter, int carried);
int atjup_hip_load_state(atjup_ctx* ctx, const char* path, int* iter, int* carried_out);
/* ChemistryJup::EquilibriumNH4SH("loop") (planet/ChemistryJup.h:293-370): every fluid cell brought
 * to the Lewis NH4SH equilibrium p_NH3 p_H2S = K_p(T), NH3, H2S and NH4SH exchanged and t moved by
 * the heat of formation (the chemistry's molar masses of atjup_hip_set_chemistry). atjup_hip_step
 * runs it after the saturation adjustments when atjup_knobs.nh4sh_eq is set
 * (planet/cJupiterModel.cpp:1084). The printed census is not formed. */
int atjup_hip_nh4sh_equilibrium(atjup_ctx* ctx);
/* Turbulence<cJupiterModel>::run (planet/Turbulence.h:307-1106) followed by the wall-function
 * omega Run() applies after it (planet/cJupiterModel.cpp:356-374, 1144). */
int atjup_hip_turbulence_run(atjup_ctx* ctx);
/* The closure's wall condition held after the step's boundary passes: apply_wall_bc and the
 * wall-function omega (planet/cJupiterModel.cpp:1294). */
int atjup_hip_turb_wall(atjup_ctx* ctx);
/* BC_Jup::bcRadius, bcTheta, bcPhi, bcSolidGround in that order (planet/BC_Jup.h:185-658,
 * planet/BoundaryConditions.h:151-518), as the time loop calls them (cJupiterModel.cpp:1263-1272). */
int atjup_hip_bc_after_step(atjup_ctx* ctx);
/* JupiterUtils::damp_wiggles (order 2) / damp_wiggles_ho (order 4), lib/Utils.cpp:221-453.
 * axes_mask: bit 0 = i, bit 1 = j, bit 2 = k; the passes run k, then j, then i. */
int atjup_hip_shapiro(atjup_ctx* ctx, int field, int order, double strength, int passes, int axes_mask);
/* cJupiterModel::computeMixtureDensity (planet/InitValues_Jup.cpp:717-740). */
int atjup_hip_mixture_density(atjup_ctx* ctx);
/* FluxLimiter<cJupiterModel>::nh4sh (planet/FluxLimiter.h:67-190). */
int atjup_hip_flux_limiter(atjup_ctx* ctx);
/* cJupiterModel::clampNegativeSpecies (planet/FileIO_Jup.cpp:494-548). */
int atjup_hip_clamp_species(atjup_ctx* ctx);
/* cJupiterModel::restoreVar(1.0) (planet/cJupiterModel.cpp:1757-1801): the n-level copies. */
int atjup_hip_swap_nlevel(atjup_ctx* ctx);
/* ATJUP_RK_SCALAR_SYNC (planet/cJupiterModel.cpp:1118-1140). */
int atjup_hip_scalar_sync(atjup_ctx* ctx);
/* One iteration of the loop body (planet/cJupiterModel.cpp:1010-1555), the functions on this path
 * in the reference's order; iter is iter_n (the physics block runs on even iter and on first). */
int atjup_hip_step(atjup_ctx* ctx, int iter, int first);
/* The same iteration in two halves, for hosts that work between them: the physics block
 * (cJupiterModel.cpp:1032-1161; nothing on odd iterations other than the first), then RK4, the
 * boundary passes, the closure's wall hold, the moist convective adjustment, the species clamp
 * and restoreVar (:1165-1555). atjup_hip_step is the two in a row. */
int atjup_hip_step_block(atjup_ctx* ctx, int iter, int first);
int atjup_hip_step_rest(atjup_ctx* ctx, int iter);
/* Non-finite census of the prognostic fields (the ATJUP_NANCHECK watch, FileIO_Jup.cpp:326-373):
 * returns ATJUP_NONFINITE if any prognostic value is NaN or infinite. */
int atjup_hip_check_finite(atjup_ctx* ctx);

/* ---- timers ------------------------------------------------------------------------------ */
/* Phase timers, the counterpart of the reference's own "time measured" lines
 * (RungeKutta_Jup_Turb.cpp:579, PressureSolver.h:764, cJupiterModel.cpp:1553). With timing on, a
 * pair of HIP events is recorded on the context's stream around each timed phase;
 * atjup_hip_read_timers synchronizes and writes, for each phase p, out[2p] = total milliseconds and
 * out[2p+1] = kernel launches inside the phase, then clears them. */
enum atjup_timer {
    ATJUP_T_POISSON_SWEEPS = 0,   /* the red-black colour kernels of one solve (2 launches per sweep) */
    ATJUP_T_RK_STAGES,            /* the four fused RK4 stage kernels of one step */
    ATJUP_T_PRESSURE,             /* the whole pressure solve */
    ATJUP_T_STEP,                 /* one loop iteration (atjup_hip_step) */
    ATJUP_T_ADJ_SWEEPS,           /* the transposed sweeps of one adjoint solve (2 launches per sweep) */
    ATJUP_T_ADJ_STEP,             /* one transposed loop iteration of atjup_hip_adj_backward */
    ATJUP_T_ADJ_RECOMPUTE,        /* the forward recomputation of one checkpoint segment */
    ATJUP_T_RADIATION,            /* the radiation's two kernels (layers, columns) */
    ATJUP_T_PRECIPITATION,        /* the precipitation's column kernel */
    ATJUP_T_CHEMISTRY,            /* the chemistry's two kernels (the smoothing passes not included) */
    ATJUP_T_MCA,                  /* the moist convective adjustment's column kernel */
    ATJUP_N_TIMERS
};
int atjup_hip_set_timing(atjup_ctx* ctx, int enable);
int atjup_hip_read_timers(atjup_ctx* ctx, double* out);

/* ---- adjoint of the dynamics step (dry profile), FP64 and FP32-refined -----------------------
 * The hand-written transpose of the linearised step: boundary passes, Shapiro on p_dyn, mixture
 * density, RK4 with its right-hand side, the hydrostatic integral and level means, and the
 * red-black pressure solve, branches frozen at the forward's choice. Adjoint fields are
 * addressed by the forward field ids (t, u, v, w, p_dyn, p_hydro, rho_mix, aux_*) and held in the
 * context's precision: an ATJUP_PRECISION_FP32_REF context transposes what its refined forward
 * runs, with its adjoint fields in FP32 (the first sweep of each solve transposed in FP64, the
 * rounds of FP32 correction sweeps transposed in FP32 inside the FP64 residual and correction
 * transposes; checkpoints carry the paired fields' low words). Host arrays stay FP64. The
 * reference has no adjoint: these entry points have no counterpart there and serve the gradient
 * of a zonal-wind misfit at the top level with respect to the initial state. */
#define ATJUP_ADJ_OP_BC            1   /* bc_after_step^T                                      */
#define ATJUP_ADJ_OP_SHAPIRO_PDYN  2   /* shapiro(p_dyn, 2, 1, 1, 7)^T                         */
#define ATJUP_ADJ_OP_MIXTURE       3   /* mixture_density^T                                    */
#define ATJUP_ADJ_OP_RK4           4   /* rk4_step^T                                           */
#define ATJUP_ADJ_OP_PRESSURE      5   /* pressure_solve(iarg sweeps, darg omega)^T            */
#define ATJUP_ADJ_OP_STEP          6   /* step(iarg iter, darg != 0 first)^T                   */
int atjup_hip_adj_init(atjup_ctx* ctx);
int atjup_hip_adj_zero(atjup_ctx* ctx);
int atjup_hip_adj_upload(atjup_ctx* ctx, int field, const double* src);
int atjup_hip_adj_download(atjup_ctx* ctx, int field, double* dst);
/* Apply one transposed operator to the adjoint fields, linearised about the context's current
 * forward state (the state before the forward operator). */
int atjup_hip_adj_op(atjup_ctx* ctx, int op, int iarg, double darg);
/* nsteps loop iterations (iter = first_iter + n, the first one treated as a run's first) from the
 * context's state, keeping a checkpoint of the state every `every` iterations. */
int atjup_hip_adj_forward_checkpointed(atjup_ctx* ctx, int nsteps, int every, int first_iter);
/* The reverse sweep: lam_final = the adjoint of the final t, u, v, w ([4][im][jm][km], or NULL to
 * start from the adjoint fields as uploaded); lam_init receives the adjoint of the initial t, u,
 * v, w. The step states come from the checkpoints (every `every` iterations) by recomputation;
 * the state after each physics block (the pressure solve's result) and the last segment's step
 * states are kept from the checkpointed forward. The context's forward fields are left as the
 * last recomputation left them. */
int atjup_hip_adj_backward(atjup_ctx* ctx, const double* lam_final, double* lam_init);

/* A smooth obstacle for gradients with respect to its semi-axes. No counterpart in the reference,
 * whose obstacle is the binary SeaMount mask (planet/BC_Jup.h:353-457): a loss depends on the
 * binary mask's semi-axes only through cells switching between fluid and solid, so its derivative
 * is zero almost everywhere. In its place the dynamics profile can carry a drag region: the
 * momentum right-hand sides of u, v, w gain -chi u / tau (and v, w alike), where
 *     chi = 1 / (1 + exp(-(1 - s) / width)),
 *     s = ((k - k0) / a)^2 + ((j - j0) / b)^2 + (i / c)^2,
 * a = a_phi_deg / dphi and b = b_the_deg / dtheta in cells, c = c_r_km / (L_atm dr) in levels, the
 * seam k = km - 1 taking the value of k = 0 (chi < 1e-16 is taken as 0). The binary obstacle is
 * independent of it (upload an all-fluid mask to use the drag region alone); the drop-in's parity
 * path never sets one. Dynamics profile only (FP64, FP32-refined and tangent forms). */
typedef struct atjup_obstacle {
    double j0, k0;                         /* centre: latitude and longitude index */
    double a_phi_deg, b_the_deg, c_r_km;   /* semi-axes */
    double width;                          /* smoothing width of chi in s */
    double tau;                            /* drag time scale, nondimensional */
} atjup_obstacle;
/* Set (params) or remove (NULL) the drag region. */
int atjup_hip_set_smooth_obstacle(atjup_ctx* ctx, const atjup_obstacle* params);
/* dL/d(a_phi_deg, b_the_deg, c_r_km) of the loss whose final-state adjoint the last
 * atjup_hip_adj_backward carried back (dynamics profile with the drag region set); out[3]. */
int atjup_hip_adj_obstacle_gradient(atjup_ctx* ctx, double* out);

/* Frozen-branch linearisations of two moist operators at the context's current state (FP64
 * context of a moist profile; every array [n][im][jm][km] on the host). The branches, clamps and
 * iteration exits are those the forward takes at that state; the tangent is the operator run on
 * (value, derivative) pairs along a direction, the adjoint the transpose of the same Jacobian.
 * SaturationAdjustment<cJupiterModel>::run for one gas (planet/SaturationAdjustment.h:255-746,
 * its final positivity pass included; constants of atjup_hip_set_gas): inputs (vapour, cloud,
 * ice, t, p_stat, rho_mix), outputs (vapour, cloud, ice, t).
 *   tangent: dx[6] -> dy[4];  adjoint: ly[4] (adjoints of the outputs) -> lx[6]. */
int atjup_hip_satadj_tangent(atjup_ctx* ctx, int gas, const double* dx, double* dy);
int atjup_hip_satadj_adjoint(atjup_ctx* ctx, int gas, const double* ly, double* lx);
/* FluxLimiter<cJupiterModel>::nh4sh (planet/FluxLimiter.h:89-190): inputs (nh4sh, u, v, w),
 * output fluxlim_nh4sh on the interior cells the limiter writes (0 elsewhere).
 *   tangent: dx[4] -> dy[1];  adjoint: ly[1] -> lx[4]. */
int atjup_hip_flux_limiter_tangent(atjup_ctx* ctx, const double* dx, double* dy);
int atjup_hip_flux_limiter_adjoint(atjup_ctx* ctx, const double* ly, double* lx);

/* ---- housekeeping ----------------------------------------------------------------------- */
int         atjup_hip_sync(atjup_ctx* ctx);
/* Number of HIP devices visible to the library's runtime (0 and an error status without one). */
int         atjup_hip_device_count(int* count);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL restores the context's own. The
 * stream must come from the same HIP runtime instance the library is linked against (ROCm's
 * libamdhip64.so.7); a handle from another copy of the runtime in the process is not valid here. */
int         atjup_hip_set_stream(atjup_ctx* ctx, void* stream);
/* Device address of a field (FP64 or FP32 storage, row pitch atjup_hip_pitch) for zero-copy views. */
void*       atjup_hip_field_device_ptr(atjup_ctx* ctx, int field);
int         atjup_hip_pitch(atjup_ctx* ctx);
const char* atjup_hip_last_error_string(void);
/* Diagnostic, no counterpart in the reference: the device math library's FP64 exp (fn 0), log (1),
 * pow(x, y) (2) or tanh (3) of n host values x (and y), written to the host array out; for measuring
 * the device functions the time step calls against a host libm, in ulps. 1 <= n <= 2^26. */
int         atjup_hip_libm_probe(int fn, const double* x, const double* y, double* out, long n);

#ifdef __cplusplus
}
#endif
#endif /* ATJUP_HIP_H */
