// Internal declarations shared by the translation units of libatjup_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/atjup_hip.h"

namespace atjup {

void set_error(const std::string& msg);

#define ATJUP_HIP_CHECK(call)                                                                    \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if(e_ != hipSuccess){                                                                    \
            atjup::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
            return ATJUP_ERR_HIP;                                                                \
        }                                                                                        \
    } while(0)

// Face flags of a cell, built from the obstacle mask (one byte per cell).
enum : uint8_t {
    FL_SELF = 1,      // the cell itself is solid (SeaMount == 1)
    FL_IP = 2,        // (i+1, j, k) solid
    FL_IM = 4,        // (i-1, j, k) solid
    FL_JP = 8,        // (i, j+1, k) solid
    FL_JM = 16,       // (i, j-1, k) solid
    FL_KP = 32,       // (i, j, ke) solid, ke the periodic east neighbour
    FL_KM = 64,       // (i, j, kw) solid
};

// Per-(i, j) metric table of the RHS, computed on the host with the reference's own expressions
// (RungeKutta_Jup_Turb.cpp:390-411) so that the device reproduces them bit for bit.
template<typename T>
struct GeoIJ {
    T rm, inv_rm, inv_rm2, inv_rmsinthe, inv_rm2sinthe, inv_rm2sinthe2, costhe_inv_rm2sinthe;
    T costhe, cotanthe, v_metric, sinthe_true, cent_rad, cent_the, pad0, pad1, pad2;
};

// Per-(i, j) weights of the Poisson stencil (PressureSolver.h:327-353).
template<typename T>
struct PoisIJ {
    T num1, num2, num3, denom, inv_rm, inv_rmsinthe;
    // The denominator a fluid cell with six fluid neighbours accumulates in the sweep, 0 + num1 +
    // num1 + num2 + num2 + num3 + num3 in that order, in FP64 (den_fluid) and in FP32 (den_fluid32,
    // each term rounded to FP32 first): the sweep's fast path takes it instead of summing it per cell.
    T den_fluid, den_fluid32;
};

// A solid cell's treatment in bcSolidGround (BC_Jup.h:517-589): the cell index and a code,
// 0 = interior (zeroed except t, p_stat), 1..6 = extrapolate from +i,-i,+j,-j,+k,-k, 7 = no pair.
struct SolidCell {
    int i, j, k, code;
};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace atjup

namespace atjup { struct AdjState; }

struct atjup_ctx {
    atjup_grid g;
    atjup_consts c;
    atjup_knobs kn;
    int precision;               // ATJUP_PRECISION_*
    int im, jm, km, pitch;       // pitch: row length in elements (km rounded up to 128 bytes)
    size_t plane, n3;            // elements per i-plane (jm * pitch) and per field (im * plane)
    size_t esz;                  // bytes per element of the field storage (8 or 4)
    hipStream_t stream = nullptr, own_stream = nullptr;

    // field storage. cur[f]: the live field (for prognostic fields also the time-level-n copy:
    // restoreVar is the identity on this path); stg[2][f]: the two RK4 stage buffers; acc[f].
    atjup::Buf cur[ATJUP_F_END];
    atjup::Buf stg[2][ATJUP_N_PROG];
    atjup::Buf acc[ATJUP_N_PROG];
    atjup::Buf live_turb[2];     // tke / dis as Turbulence::run left them, before the RK4 (full profile)
    bool turb_live_pending = false;
    atjup::Buf scratch;          // one field, for the filters
    atjup::Buf psrc;             // Poisson source, FP64, red-black split [2][im][jm][rb_hp]
    atjup::Buf rb_p;             // p_dyn during a solve, FP64, red-black split [2][im][jm][rb_hp]
    atjup::Buf rb_d, rb_r;       // FP32-refined solve: correction and FP64-formed residual, FP32, split
    int refine_sweeps = 50;      // FP32 sweeps per refinement round
    // A captured sweep sequence bakes in its kernels' arguments: the buffers (iterate, source, face
    // flags and weights of both colours) and the launch geometry. It is replayed only when all of
    // them equal the request's; any difference re-captures it.
    struct SweepGraph {
        int kind, n, ghost_first; double omega; void* exec;
        const void* ptrs[7]; int dims[4];
    };
    std::vector<SweepGraph> sweep_graphs;   // instantiated sweep sequences (hipGraphExec_t)
    atjup::Buf rb_flags;         // face flags, red-black split [2][im][jm][rb_hp]
    atjup::Buf lo[ATJUP_N_PROG]; // FP32-refined form: low words of t, u, v, w and, in the moist
                                 // profiles, the eleven species (the value is hi + lo)
    atjup::Buf colr;             // int2 (i0, i1) per (j, k): the column's fluid run from the bottom
    atjup::Buf steady;           // steadyQuery tracking: 12 FP64 n-level copies [12][n3], then 12 results
    bool steady_on = false;
    int rb_hp = 0;               // compact row length: (km + 1) / 2 rounded up to 128 bytes
    atjup::Buf mask;             // uint8 SeaMount [n3]
    atjup::Buf flags;            // uint8 face flags [n3]
    atjup::Buf itopo;            // int [jm * km]
    atjup::Buf vel_star;         // double [jm * km]
    atjup::Buf geo;              // GeoIJ<T> [im * jm]
    atjup::Buf pois;             // PoisIJ<double> [im * jm]
    atjup::Buf pois32;           // PoisIJ<float> [im * jm]
    atjup::Buf tabs;             // double rad[im], sin(the)[jm], Coriolis split factor [jm], cos(the)[jm],
                                 // layer thickness [m] [im]
    atjup::Buf brl;              // double [im]
    atjup::Buf brl_part;         // double partial sums
    atjup::Buf zm;               // double [4 * jm], the top-sponge zonal means
    atjup::Buf solids;           // SolidCell list
    atjup::Buf amb_rows;         // int2 list of (i, j) rows that hold solid cells
    atjup::Buf nonfinite;        // int flag
    atjup_rad_params radp = {};     // the radiation's planet properties and knobs (atjup_hip_set_radiation)
    atjup::Buf rad_tab;             // double F_sun_toa [jm], radiation layer thickness [m] [im]
    atjup::Buf rad_tmp;             // FP32-refined form: the radiation's FP64 eps, B, Fsw [3][n3]
    atjup_precip_params precp = {}; // the precipitation's constants (atjup_hip_set_precipitation)
    atjup::Buf precip_srf;          // double [5][jm][km]: the surface precipitation maps
    atjup_chem_params chemp = {};   // the chemistry's constants and knobs (atjup_hip_set_chemistry)
    atjup::Buf chem_w;              // double [3][n3]: the reaction rates w_nh3, w_h2s, w_nh4sh
    atjup::Buf report;              // the reporting reductions' partial results and slice staging
    atjup_obstacle obst = {};       // the smooth drag region (atjup_hip_set_smooth_obstacle)
    bool obst_on = false;
    atjup::Buf chi;                 // double [n3]: its indicator chi, when set
    atjup_mca_params mcap = {};     // the moist convective adjustment's knobs (atjup_hip_set_mca)
    atjup::Buf mca_ws;              // its layer heights [im] (m), then the column workspace
                                    // double [8][im][jm * km] and the flags uint8 [im][jm * km]
    atjup_gas_params gas[3] = {};   // saturation-adjustment constants (atjup_hip_set_gas)
    bool gas_set[3] = { false, false, false };
    int n_solids = 0, n_amb_rows = 0;
    bool have_mask = false;

    // phase timers (atjup_hip_set_timing): pending event pairs per phase, with launch counts
    bool timing = false;
    struct TimedSpan { hipEvent_t a, b; int launches; };
    std::vector<TimedSpan> spans[ATJUP_N_TIMERS];
    std::vector<hipEvent_t> event_pool;
    double t_total[ATJUP_N_TIMERS] = {0};
    double t_launch[ATJUP_N_TIMERS] = {0};

    atjup::AdjState* adj = nullptr;   // the adjoint's buffers (atjup_hip_adj_init)

    // host-side geometry
    std::vector<double> rad, the, sin_the, cos_the;
    std::vector<float> lh;

    size_t idx(int i, int j, int k) const { return ((size_t)i * jm + j) * pitch + k; }
};

namespace atjup {

// launchers (atjup_kernels.hip)
int launch_buoy_ref_level(atjup_ctx* c);
int launch_hydrostatic(atjup_ctx* c);
int launch_rk4(atjup_ctx* c);
int launch_pressure(atjup_ctx* c, int sweeps, double omega);
int launch_bc_after_step(atjup_ctx* c);
int launch_prep_projection(atjup_ctx* c);          // prepareProjectionBoundaries alone (FP64)
// one loop iteration in two parts (atjup_hip_step = step_block + step_rest): the physics block of
// the even and first iterations, then the RK4 step, the boundary passes and the clamps
int step_block(atjup_ctx* c, int iter, int first);
int step_rest(atjup_ctx* c, int iter);
int launch_shapiro(atjup_ctx* c, int field, int order, double strength, int passes, int axes);
int launch_mixture_density(atjup_ctx* c);
int launch_clamp_species(atjup_ctx* c);
int launch_mask_setup(atjup_ctx* c);
int launch_check_finite(atjup_ctx* c);
int launch_copy_field(atjup_ctx* c, void* dst, const void* src);
// full-physics profile (atjup_physics.hip)
int launch_rk4_full(atjup_ctx* c);
int launch_turbulence(atjup_ctx* c);
int launch_turb_wall(atjup_ctx* c);
int launch_turb_sanitize(atjup_ctx* c);
int launch_satadj(atjup_ctx* c, int gas, const atjup_gas_params& g);
int launch_satadj3(atjup_ctx* c);                     // the three gases in one pass (constants set)
int launch_radiation(atjup_ctx* c);
int launch_precipitation(atjup_ctx* c);
int launch_chem_kernel(atjup_ctx* c, int which);      // 0: DiffMassFluxJup, 1: ChemMassRateJup
int launch_mca(atjup_ctx* c);
int launch_nh4sh_eq(atjup_ctx* c);
size_t report_bytes(int im, int jm, int km);
// steadyQuery's n-level copies (bit e: entry e of the report's field order) into the tracking buffer
int steady_mark(atjup_ctx* c, unsigned entries);          // the size of atjup_ctx::report
int build_rad_table(atjup_ctx* c);
int launch_flux_limiter(atjup_ctx* c);
// frozen-branch tangent (adjoint = 0) or adjoint (1) of one gas's adjustment / of the limiter, FP64
int launch_satadj_lin(atjup_ctx* c, int gas, int adjoint, const double* in, double* out);
int launch_flux_limiter_lin(atjup_ctx* c, int adjoint, const double* in, double* out);
int launch_commit_live(atjup_ctx* c);

// The storage a live field is read from: the closure's live k*, omega* while a turbulence run
// waits for its RK4 step, the field's own storage otherwise.
inline void* live_ptr(atjup_ctx* c, int f){
    if(c->turb_live_pending && (f == ATJUP_F_TKE || f == ATJUP_F_DIS)) return c->live_turb[f - ATJUP_F_TKE].p;
    return c->cur[f].p;
}

void adj_free(atjup_ctx* c);

// the adjoint's recording RK4 run: the stage inputs Y1..Y3, the step's result and its aux_*, in
// fields of the context's storage type (FP64 or FP32; the pointers are carried untyped)
struct RkBuffers { double* y[3][4]; double* nn[4]; double* aux[3]; };
int launch_rk4_dry_record(atjup_ctx* c, const RkBuffers& rb);

// phase timers: begin returns an event recorded now (or nullptr with timing off); end records the
// closing event and files the span under `phase` with its kernel-launch count.
hipEvent_t timer_begin(atjup_ctx* c);
void timer_end(atjup_ctx* c, int phase, hipEvent_t begin, int launches);

}  // namespace atjup
