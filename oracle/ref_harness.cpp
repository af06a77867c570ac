// Test harness around the reference model, compiled against the reference's own headers and
// linked with its translation units (see oracle/Makefile, target `ref`). It exists only to produce
// golden vectors and a CPU baseline: it never ships with the library and no product code loads it.
//
// Access to the model's private members goes through the friend declaration the model already
// makes for every specialization of ProcessDiag (planet/cJupiterModel.h:80): the harness defines
// its own specialization, ProcessDiag<HarnessTag>, and does all member access from there.
//
// Exported C ABI (all contiguous buffers are [i][j][k], k fastest, im*jm*km doubles):
//   ref_init(profile)          set the knob profile ("KEY=VAL;KEY=VAL"), build the model and its
//                              constants the way cJupiterModel::Run() does before the loop
//   ref_set_field/ref_get_field(name, buf)
//   ref_set_scalar/ref_get_scalar(name, value)
//   ref_set_mask(seamount)     obstacle mask, i_topography and wall_nue (computeWallViscosity)
//   ref_call(op)               one reference entry point, by name (see the table in ref_call)
#include "cJupiterModel.h"

#include "ChemistryJup.h"
#include "PressureSolverJup.h"
#include "SaturationAdjustmentJup.h"
#include "BC_Jup.h"
#include "TurbulenceJup.h"
#include "RadiationJup.h"
#include "PrecipitationJup.h"
#include "ConvectiveAdjustmentJup.h"
#include "Reporting.h"

#include <fcntl.h>
#include <omp.h>
#include <unistd.h>

#include <map>
#include <string>
#include <vector>

struct HarnessTag {};

template<>
class ProcessDiag<HarnessTag> {
public:
    static cJupiterModel* m;

    static std::map<std::string, Array*> fields(){
        cJupiterModel& M = *m;
        return {
            {"t", &M.t}, {"u", &M.u}, {"v", &M.v}, {"w", &M.w},
            {"tn", &M.tn}, {"un", &M.un}, {"vn", &M.vn}, {"wn", &M.wn},
            {"h2o", &M.h2o}, {"h2o_cloud", &M.h2o_cloud}, {"h2o_ice", &M.h2o_ice},
            {"h2s", &M.h2s}, {"nh3", &M.nh3}, {"nh3_cloud", &M.nh3_cloud}, {"nh3_ice", &M.nh3_ice},
            {"ch4", &M.ch4}, {"ch4_cloud", &M.ch4_cloud}, {"ch4_ice", &M.ch4_ice}, {"nh4sh", &M.nh4sh},
            {"h2on", &M.h2on}, {"h2o_cloudn", &M.h2o_cloudn}, {"h2o_icen", &M.h2o_icen},
            {"h2sn", &M.h2sn}, {"nh3n", &M.nh3n}, {"nh3_cloudn", &M.nh3_cloudn}, {"nh3_icen", &M.nh3_icen},
            {"ch4n", &M.ch4n}, {"ch4_cloudn", &M.ch4_cloudn}, {"ch4_icen", &M.ch4_icen}, {"nh4shn", &M.nh4shn},
            {"tke", &M.tke}, {"dis", &M.dis}, {"tken", &M.tken}, {"disn", &M.disn},
            {"nue", &M.nue}, {"prod", &M.prod}, {"tke_source", &M.tke_source}, {"dis_source", &M.dis_source},
            {"p_dyn", &M.p_dyn}, {"p_dynn", &M.p_dynn}, {"p_stat", &M.p_stat}, {"p_hydro", &M.p_hydro},
            {"rho_mix", &M.rho_mix}, {"wall_nue", &M.wall_nue}, {"SeaMount", &M.SeaMount},
            {"aux_u", &M.aux_u}, {"aux_v", &M.aux_v}, {"aux_w", &M.aux_w},
            {"rhs_t", &M.rhs_t}, {"rhs_u", &M.rhs_u}, {"rhs_v", &M.rhs_v}, {"rhs_w", &M.rhs_w},
            {"rhs_tke", &M.rhs_tke}, {"rhs_dis", &M.rhs_dis}, {"rhs_h2o", &M.rhs_h2o},
            {"acc_t", &M.acc_t}, {"acc_u", &M.acc_u}, {"acc_v", &M.acc_v}, {"acc_w", &M.acc_w},
            {"Q_rad", &M.Q_rad}, {"Q_precip", &M.Q_precip},
            {"S_precip_h2o", &M.S_precip_h2o}, {"S_precip_h2o_cloud", &M.S_precip_h2o_cloud},
            {"S_precip_h2o_ice", &M.S_precip_h2o_ice}, {"S_precip_nh3", &M.S_precip_nh3},
            {"S_precip_nh3_cloud", &M.S_precip_nh3_cloud}, {"S_precip_nh3_ice", &M.S_precip_nh3_ice},
            {"S_precip_ch4", &M.S_precip_ch4}, {"S_precip_ch4_cloud", &M.S_precip_ch4_cloud},
            {"S_precip_ch4_ice", &M.S_precip_ch4_ice},
            {"massflux_h2s", &M.massflux_h2s}, {"massflux_nh3", &M.massflux_nh3},
            {"massflux_nh4sh", &M.massflux_nh4sh}, {"fluxlim_nh4sh", &M.fluxlim_nh4sh},
            {"epsilon", &M.epsilon}, {"radiation", &M.radiation},
            {"P_rain", &M.P_rain}, {"P_snow", &M.P_snow}, {"P_graupel", &M.P_graupel},
            {"P_nh3_rain", &M.P_nh3_rain}, {"P_nh3_snow", &M.P_nh3_snow}, {"P_nh3_graupel", &M.P_nh3_graupel},
            {"P_ch4_rain", &M.P_ch4_rain}, {"P_ch4_snow", &M.P_ch4_snow}, {"P_ch4_graupel", &M.P_ch4_graupel},
            {"P_nh4sh", &M.P_nh4sh},
            {"difflux_h2s", &M.difflux_h2s}, {"difflux_nh3", &M.difflux_nh3},
            {"thermalmassflux", &M.thermalmassflux}, {"difflux_nh4sh", &M.difflux_nh4sh},
            {"w_nh3", &M.w_nh3}, {"w_h2s", &M.w_h2s}, {"w_nh4sh", &M.w_nh4sh},
        };
    }

    static void save_state(const char* dir, int iter){ m->output_path = dir; m->save_state(iter); }
    static bool load_state(const char* dir, int iter){ m->output_path = dir; return m->load_state(iter); }

    static std::map<std::string, Array_2D*> fields2d(){
        cJupiterModel& M = *m;
        return {
            {"precip_srf_h2o", &M.precip_srf_h2o}, {"precip_srf_nh3", &M.precip_srf_nh3},
            {"precip_srf_ch4", &M.precip_srf_ch4}, {"precip_srf_nh4sh", &M.precip_srf_nh4sh},
            {"precip_srf_total", &M.precip_srf_total},
        };
    }

    static void res(int* ijk){ ijk[0] = m->i_res; ijk[1] = m->j_res; ijk[2] = m->k_res; }

    static double* scalar(const std::string& n){
        cJupiterModel& M = *m;
        static std::map<std::string, double*> s;
        if(s.empty()){
            s = {
                {"Coriolis", &M.Coriolis}, {"centrifugal", &M.centrifugal}, {"buoyancy", &M.buoyancy},
                {"chemical_reaction", &M.chemical_reaction}, {"L_atm", &M.L_atm}, {"re", &M.re},
                {"pr", &M.pr}, {"g", &M.g}, {"omega", &M.omega}, {"u_0", &M.u_0}, {"t_ref", &M.t_ref}, {"ep_h2s", &M.ep_h2s},
                {"gam", &M.gam}, {"p_ref", &M.p_ref}, {"R_ref", &M.R_ref}, {"dt", &M.dt},
                {"r_mix", &M.r_mix}, {"R_mix", &M.R_mix}, {"cp_mix", &M.cp_mix}, {"mue_mix", &M.mue_mix},
                {"sc_h2o", &M.sc_h2o}, {"sc_h2s", &M.sc_h2s}, {"sc_nh3", &M.sc_nh3},
                {"sc_nh4sh", &M.sc_nh4sh}, {"sc_ch4", &M.sc_ch4}, {"rg_nh4sh", &M.rg_nh4sh},
                {"r_p_nh4sh", &M.r_p_nh4sh}, {"re_turb", &M.re_turb}, {"t_pole", &M.t_pole},
                {"t_equator", &M.t_equator}, {"abl_height", &M.abl_height},
                // the chemistry's constants (ChemistryJup.h:59-545) as Run() leaves them
                {"m_nh3", &M.m_nh3}, {"m_h2s", &M.m_h2s}, {"m_nh4sh", &M.m_nh4sh}, {"m_h2", &M.m_h2},
                {"m_he", &M.m_he}, {"m_h2o", &M.m_h2o}, {"X_h2", &M.X_h2}, {"X_he", &M.X_he},
                {"X_h2o", &M.X_h2o}, {"X_h2s", &M.X_h2s}, {"X_nh3", &M.X_nh3}, {"X_nh4sh", &M.X_nh4sh},
                {"mue_nh3", &M.mue_nh3}, {"mue_h2s", &M.mue_h2s}, {"mue_nh4sh", &M.mue_nh4sh},
                {"rg_nh3", &M.rg_nh3}, {"rg_h2s", &M.rg_h2s}, {"k_mix", &M.k_mix},
                {"cp_nh3", &M.cp_nh3}, {"cp_h2s", &M.cp_h2s}, {"cp_nh4sh", &M.cp_nh4sh},
                {"k_nh3", &M.k_nh3}, {"k_h2s", &M.k_h2s}, {"k_nh4sh", &M.k_nh4sh}, {"dh_nh4sh", &M.dh_nh4sh},
                {"t_0_nh4sh", &M.t_0_nh4sh}, {"t_00_nh4sh", &M.t_00_nh4sh},
            };
            // the saturation adjustment's per-gas constants as Run() leaves them (SVP_FIT, COMPOSITION)
#define GASK(g) \
            s["t_0_" #g] = &M.t_0_##g; s["t_00_" #g] = &M.t_00_##g; s["ep_" #g] = &M.ep_##g; \
            s["lv_" #g] = &M.lv_##g; s["ls_" #g] = &M.ls_##g; s["C_" #g] = &M.C_##g; s["L0_" #g] = &M.L0_##g; \
            s["R_" #g] = &M.R_##g; s["del_alf_" #g] = &M.del_alf_##g; s["del_bet_" #g] = &M.del_bet_##g; \
            s["C_" #g "_ice"] = &M.C_##g##_ice; s["L0_" #g "_ice"] = &M.L0_##g##_ice; \
            s["del_alf_" #g "_ice"] = &M.del_alf_##g##_ice; s["del_bet_" #g "_ice"] = &M.del_bet_##g##_ice;
            GASK(h2o) GASK(nh3) GASK(ch4)
#undef GASK
        }
        auto it = s.find(n);
        return it == s.end() ? nullptr : it->second;
    }

    static void set_mask(const double* src){
        cJupiterModel& M = *m;
        const int im = M.im, jm = M.jm, km = M.km;
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++)
            M.SeaMount.x[i][j][k] = src[((size_t)i * jm + j) * km + k];
        M.i_topography.assign(jm, std::vector<int>(km, 0));
        for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){
            int s = 0;
            while(s < im && M.SeaMount.x[s][j][k] == 1.0) ++s;
            M.i_topography[j][k] = s;
        }
        M.computeWallViscosity();
    }

    static void init(){
        m = new cJupiterModel();
        cJupiterModel& M = *m;
        // Run(): the closure switch (turb_switch, cJupiterModel.cpp:73-78)
        {
            const char* e = getenv("ATJUP_TURB");
            const char* mm = getenv("ATJUP_TURB_MODEL");
            const std::string s = mm ? std::string(mm) : M.turb_model;
            M.use_turbulence_model = e ? (atoi(e) != 0) : !(s.empty() || s == "none" || s == "laminar");
        }
        JupiterUtils::set_phi_seam_periodic(ATPhys::phi_periodic<cJupiterModel>());
        // Run(): the metric radius (cJupiterModel.cpp:441-468)
        {
            const char* e = getenv("ATJUP_METRIC_RADIUS");
            const double R = e ? atof(e) : cJupiterModel::R_planet_km;
            if(R > 0.0 && !M.coord_stretching) M.rad.Coordinates(M.im, R / M.L_atm, M.dr);
        }
        M.resetArrays();
        {
            const char* e = getenv("ATJUP_DT");
            M.dt = e ? atof(e) : 0.005;
        }
        M.iter_n = 0;
        M.init_layer_heights();
        M.i_topography.assign(M.jm, std::vector<int>(M.km, 0));
        M.computeWallViscosity();
        // Run(): the mixture constants (cJupiterModel.cpp:590-597)
        if(M.svp_fit()) M.setSvpFit();
        if(M.composition()) M.setComposition();
        ChemistryJup(M).ThermalPropertiesJup();
        if(M.composition()) M.setCompositionEp();
    }

    // The wall-function omega that Run() applies after every apply_wall_bc(); it is a lambda local
    // to Run() (planet/cJupiterModel.cpp:356-374), so the harness repeats its few lines here with
    // ATJUP_TURB_WALL_OMEGA at its default 1 and a k-omega closure.
    static void wall_omega(cJupiterModel& M){
        // the lambda leaves a k-epsilon closure alone (cJupiterModel.cpp:356-360)
        const char* e = getenv("ATJUP_TURB_MODEL");
        if((e ? std::string(e) : M.turb_model) == "k_epsilon") return;
        const double nue_gas_phys = 1.8e-5, bet_star = 0.09, bet_1 = 0.0333, kappa = 0.41;
        const double nd_omega = M.L_atm * 1.0e3 / M.u_0;
        for(int j = 0; j < M.jm; j++)
            for(int k = 0; k < M.km; k++){
                const int i_w = M.surface_index(j, k);
                if(i_w <= 0 || i_w >= M.im - 1) continue;
                const double y1 = std::max(M.layer_thickness_m(std::max(i_w - 1, 0)), 1.0);
                const double y = 0.5 * y1;
                const double om_vis = 60.0 * nue_gas_phys / (bet_1 * y1 * y1);
                const double om_log = M.vel_star.y[j][k] / (std::sqrt(bet_star) * kappa * y);
                M.dis.x[i_w][j][k] = std::sqrt(om_vis * om_vis + om_log * om_log) * nd_omega;
            }
    }

    static int call(const std::string& op){
        cJupiterModel& M = *m;
        if(op == "rk4")                 M.RungeKuttaJup();
        else if(op == "buoy_ref_level") M.computeBuoyancyRefLevel();
        else if(op == "hydrostatic")    { M.computeBuoyancyRefLevel(); M.computeHydrostaticPressure(true); }
        else if(op == "pressure")       PressureSolverJup(M).run();
        else if(op == "damp_pdyn")      JupiterUtils::damp_wiggles(M.p_dyn, &M.i_topography, true, true, true);
        else if(op == "mixture_density") M.computeMixtureDensity();
        else if(op == "bc_radius")      BC_Jup(M).bcRadius();
        else if(op == "bc_theta")       BC_Jup(M).bcTheta();
        else if(op == "bc_phi")         BC_Jup(M).bcPhi();
        else if(op == "bc_solid")       BC_Jup(M).bcSolidGround();
        else if(op == "bc_all")         { BC_Jup(M).bcRadius(); BC_Jup(M).bcTheta(); BC_Jup(M).bcPhi(); BC_Jup(M).bcSolidGround(); }
        else if(op == "clamp_species")  M.clampNegativeSpecies();
        else if(op == "restore")        M.restoreVar(1.0);
        else if(op == "flux_limiter")   ChemistryJup(M).FluxLimiterNH4SH();
        else if(op == "fluxlim_damp")   JupiterUtils::damp_wiggles(M.fluxlim_nh4sh, &M.i_topography, true, true, true);
        else if(op == "turbulence")     { TurbulenceJup(M).run(); wall_omega(M); }
        else if(op == "turb_wall")      { TurbulenceJup(M).apply_wall_bc(); wall_omega(M); }
        else if(op == "radiation")      RadiationJup(M).run();
        else if(op == "precipitation")  PrecipitationJup(M).run();
        else if(op == "chemistry"){
            // the physics block's chemistry (cJupiterModel.cpp:1146-1161)
            ChemistryJup(M).DiffMassFluxJup();
            JupiterUtils::damp_wiggles(M.difflux_h2s,     &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.difflux_nh3,     &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.difflux_nh4sh,   &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.thermalmassflux, &M.i_topography, true, true, true);
            ChemistryJup(M).ChemMassRateJup();
            ChemistryJup(M).FluxLimiterNH4SH();
            JupiterUtils::damp_wiggles(M.massflux_h2s,   &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.massflux_nh3,   &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.massflux_nh4sh, &M.i_topography, true, true, true);
            JupiterUtils::damp_wiggles(M.fluxlim_nh4sh,  &M.i_topography, true, true, true);
        }
        else if(op == "mca"){
            // the moist convective adjustment as the time loop builds it (cJupiterModel.cpp:1301-1320)
            typedef MoistConvectiveAdjustmentJup::Gas Gas_;
            typedef MoistConvectiveAdjustmentJup::Tracer Tr_;
            std::vector<Gas_> gases = {
                { "H2O", &M.h2o, &M.h2o_cloud, &M.h2o_ice, M.ep_h2o, M.lv_h2o, M.ls_h2o, M.t_0_h2o,
                  M.C_h2o, M.L0_h2o, M.R_h2o, M.del_alf_h2o, M.del_bet_h2o, M.C_h2o_ice, M.L0_h2o_ice,
                  M.del_alf_h2o_ice, M.del_bet_h2o_ice, M.t_00_h2o },
                { "NH3", &M.nh3, &M.nh3_cloud, &M.nh3_ice, M.ep_nh3, M.lv_nh3, M.ls_nh3, M.t_0_nh3,
                  M.C_nh3, M.L0_nh3, M.R_nh3, M.del_alf_nh3, M.del_bet_nh3, M.C_nh3_ice, M.L0_nh3_ice,
                  M.del_alf_nh3_ice, M.del_bet_nh3_ice, M.t_00_nh3 } };
            const double L_nh4sh_ = ChemistryJup::nh4sh_eq() ? ChemistryJup::nh4sh_formation_heat() : 0.0;
            std::vector<Tr_> tracers = { { "H2S", &M.h2s, M.ep_h2s, false, 0.0 }, { "NH4SH", &M.nh4sh, 0.0, true, L_nh4sh_ },
                                         { "chrom", &M.chrom, 0.0, false, 0.0 } };
            MoistConvectiveAdjustmentJup(M, gases, tracers).run();
        }
        else if(op == "steady_query")   Reporting<cJupiterModel>(M).steadyQuery();   // prints; i/j/k_res kept
        else if(op == "diffmassflux")   ChemistryJup(M).DiffMassFluxJup();
        else if(op == "chemmassrate")   ChemistryJup(M).ChemMassRateJup();
        else if(op == "satadj_h2o")
            SaturationAdjustmentJup(M).run("H2O", M.t_0_h2o, M.t_00_h2o, M.ep_h2o, M.lv_h2o, M.ls_h2o,
                M.C_h2o, M.L0_h2o, M.R_h2o, M.del_alf_h2o, M.del_bet_h2o,
                M.C_h2o_ice, M.L0_h2o_ice, M.del_alf_h2o_ice, M.del_bet_h2o_ice,
                M.h2o, M.h2o_cloud, M.h2o_ice);
        else if(op == "satadj_nh3")
            SaturationAdjustmentJup(M).run("NH3", M.t_0_nh3, M.t_00_nh3, M.ep_nh3, M.lv_nh3, M.ls_nh3,
                M.C_nh3, M.L0_nh3, M.R_nh3, M.del_alf_nh3, M.del_bet_nh3,
                M.C_nh3_ice, M.L0_nh3_ice, M.del_alf_nh3_ice, M.del_bet_nh3_ice,
                M.nh3, M.nh3_cloud, M.nh3_ice);
        else if(op == "satadj_ch4")
            SaturationAdjustmentJup(M).run("CH4", M.t_0_ch4, M.t_00_ch4, M.ep_ch4, M.lv_ch4, M.ls_ch4,
                M.C_ch4, M.L0_ch4, M.R_ch4, M.del_alf_ch4, M.del_bet_ch4,
                M.C_ch4_ice, M.L0_ch4_ice, M.del_alf_ch4_ice, M.del_bet_ch4_ice,
                M.ch4, M.ch4_cloud, M.ch4_ice);
        else if(op == "scalar_sync"){
            // ATJUP_RK_SCALAR_SYNC = 2 (cJupiterModel.cpp:1118-1140); H2S and NH4SH join it when
            // the NH4SH equilibrium adjustment runs (:1120)
            const bool eq = ChemistryJup::nh4sh_eq();
            #pragma omp parallel for collapse(2) schedule(static)
            for(int i = 0; i < M.im; i++) for(int j = 0; j < M.jm; j++) for(int k = 0; k < M.km; k++){
                M.h2on.x[i][j][k] = M.h2o.x[i][j][k];       M.h2o_cloudn.x[i][j][k] = M.h2o_cloud.x[i][j][k];
                M.h2o_icen.x[i][j][k] = M.h2o_ice.x[i][j][k]; M.nh3n.x[i][j][k] = M.nh3.x[i][j][k];
                M.nh3_cloudn.x[i][j][k] = M.nh3_cloud.x[i][j][k]; M.nh3_icen.x[i][j][k] = M.nh3_ice.x[i][j][k];
                M.ch4n.x[i][j][k] = M.ch4.x[i][j][k];       M.ch4_cloudn.x[i][j][k] = M.ch4_cloud.x[i][j][k];
                M.ch4_icen.x[i][j][k] = M.ch4_ice.x[i][j][k]; M.tn.x[i][j][k] = M.t.x[i][j][k];
                if(eq){ M.h2sn.x[i][j][k] = M.h2s.x[i][j][k]; M.nh4shn.x[i][j][k] = M.nh4sh.x[i][j][k]; }
            }
        }
        else if(op == "nh4sh_eq")       ChemistryJup(M).EquilibriumNH4SH("loop");   // cJupiterModel.cpp:1084
        else if(op == "iter_inc")       M.iter_n++;
        else return -1;
        return 0;
    }

    static double buoy_ref_level(int i){
        return (i >= 0 && i < (int)m->buoy_ref_level.size()) ? m->buoy_ref_level[i] : 0.0;
    }
    static double coord(int axis, int n){
        if(axis == 0) return m->rad.z[n];
        if(axis == 1) return m->the.z[n];
        return m->phi.z[n];
    }
    static void dims(int* im, int* jm, int* km){ *im = m->im; *jm = m->jm; *km = m->km; }
    static double* vel_star_row(int j){ return m->vel_star.y[j]; }
};

cJupiterModel* ProcessDiag<HarnessTag>::m = nullptr;
typedef ProcessDiag<HarnessTag> H;

static int g_quiet = 1;
static int g_saved_stdout = -1;
static void quiet_begin(){
    if(!g_quiet) return;
    fflush(stdout);
    std::cout.flush();
    g_saved_stdout = dup(1);
    const int dn = open("/dev/null", O_WRONLY);
    dup2(dn, 1);
    close(dn);
}
static void quiet_end(){
    if(g_saved_stdout < 0) return;
    fflush(stdout);
    std::cout.flush();
    dup2(g_saved_stdout, 1);
    close(g_saved_stdout);
    g_saved_stdout = -1;
}

extern "C" {

void ref_quiet(int on){ g_quiet = on; }

int ref_init(const char* profile){
    if(H::m) return -1;                       // the knobs are cached function statics: one profile per process
    std::string p(profile ? profile : "");
    size_t pos = 0;
    while(pos < p.size()){
        size_t end = p.find(';', pos);
        if(end == std::string::npos) end = p.size();
        const std::string kv = p.substr(pos, end - pos);
        const size_t eq = kv.find('=');
        if(eq != std::string::npos) setenv(kv.substr(0, eq).c_str(), kv.substr(eq + 1).c_str(), 1);
        pos = end + 1;
    }
    quiet_begin();
    H::init();
    quiet_end();
    return 0;
}

void ref_dims(int* im, int* jm, int* km){ H::dims(im, jm, km); }

int ref_set_field(const char* name, const double* src){
    auto f = H::fields();
    auto it = f.find(name);
    if(it == f.end()) return -1;
    Array& A = *it->second;
    for(int i = 0; i < A.im; i++) for(int j = 0; j < A.jm; j++)
        std::memcpy(A.x[i][j], src + ((size_t)i * A.jm + j) * A.km, sizeof(double) * A.km);
    return 0;
}

int ref_get_field(const char* name, double* dst){
    auto f = H::fields();
    auto it = f.find(name);
    if(it == f.end()) return -1;
    Array& A = *it->second;
    for(int i = 0; i < A.im; i++) for(int j = 0; j < A.jm; j++)
        std::memcpy(dst + ((size_t)i * A.jm + j) * A.km, A.x[i][j], sizeof(double) * A.km);
    return 0;
}

int ref_get_field2d(const char* name, double* dst){
    auto f = H::fields2d();
    auto it = f.find(name);
    if(it == f.end()) return -1;
    Array_2D& A = *it->second;
    int im, jm, km; H::dims(&im, &jm, &km);
    for(int j = 0; j < jm; j++) std::memcpy(dst + (size_t)j * km, A.y[j], sizeof(double) * km);
    return 0;
}

int ref_set_scalar(const char* name, double v){
    double* p = H::scalar(name);
    if(!p) return -1;
    *p = v;
    return 0;
}

int ref_get_scalar(const char* name, double* v){
    double* p = H::scalar(name);
    if(!p) return -1;
    *v = *p;
    return 0;
}

// cJupiterModel::save_state / load_state (FileIO_Jup.cpp:584-684) with output_path = dir
int ref_save_state(const char* dir, int iter){
    quiet_begin();
    H::save_state(dir, iter);
    quiet_end();
    return 0;
}
int ref_load_state(const char* dir, int iter){
    quiet_begin();
    const bool ok = H::load_state(dir, iter);
    quiet_end();
    return ok ? 0 : -1;
}

void ref_set_mask(const double* seamount){ quiet_begin(); H::set_mask(seamount); quiet_end(); }

int ref_call(const char* op){
    quiet_begin();
    const int rc = H::call(op);
    quiet_end();
    return rc;
}

double ref_buoy_ref_level(int i){ return H::buoy_ref_level(i); }
double ref_coord(int axis, int n){ return H::coord(axis, n); }

// The loop body's on-path functions in the reference's order (cJupiterModel.cpp:1010-1555), the
// functions outside the path (radiation, precipitation, chemistry, MCA, diagnostics) left out.
int ref_step(int iter, int first, int satadj, int turb){
    quiet_begin();
    int rc = 0;
    if(iter % 2 == 0 || first){
        rc |= H::call("pressure");
        rc |= H::call("damp_pdyn");
        rc |= H::call("mixture_density");
        if(satadj){ rc |= H::call("satadj_h2o"); rc |= H::call("satadj_nh3"); rc |= H::call("satadj_ch4"); }
        rc |= H::call("scalar_sync");
        if(turb) rc |= H::call("turbulence");
        rc |= H::call("flux_limiter");
        rc |= H::call("fluxlim_damp");
    }
    rc |= H::call("rk4");
    rc |= H::call("bc_all");
    if(turb) rc |= H::call("turb_wall");
    const char* e = getenv("ATJUP_MCA");                 // mca_stride() (cJupiterModel.cpp:122-125)
    const int mca = e ? std::max(0, atoi(e)) : 1;
    if(mca > 0 && iter % mca == 0) rc |= H::call("mca");
    rc |= H::call("clamp_species");
    rc |= H::call("restore");
    quiet_end();
    return rc;
}

// The loop body with the whole physics block as the reference runs it (cJupiterModel.cpp:1032-1161):
// flags bit 0 the saturation adjustments, bit 1 the closure, bit 2 the radiation, bit 3 the
// precipitation, bit 4 the chemistry sequence (DiffMassFluxJup, its smoothing, ChemMassRateJup, the
// NH4SH limiter and the smoothing of the mass fluxes, which the reference always runs; without it
// only the limiter and its smoothing run, massflux_* frozen), bit 5 the NH4SH equilibrium
// adjustment (its knob ATJUP_NH4SH_EQ must be on in the profile). Then the rest of the iteration
// as ref_step.
int ref_step_physics(int iter, int first, int flags){
    quiet_begin();
    int rc = 0;
    const bool turb = flags & 2;
    if(iter % 2 == 0 || first){
        rc |= H::call("pressure");
        rc |= H::call("damp_pdyn");
        rc |= H::call("mixture_density");
        if(flags & 4) rc |= H::call("radiation");
        if(flags & 1){ rc |= H::call("satadj_h2o"); rc |= H::call("satadj_nh3"); rc |= H::call("satadj_ch4"); }
        if(flags & 32){
            if(!ChemistryJup::nh4sh_eq()) rc |= 1;
            else rc |= H::call("nh4sh_eq");
        }
        if(flags & 8) rc |= H::call("precipitation");
        rc |= H::call("scalar_sync");
        if(turb) rc |= H::call("turbulence");
        if(flags & 16) rc |= H::call("chemistry");
        else { rc |= H::call("flux_limiter"); rc |= H::call("fluxlim_damp"); }
    }
    rc |= H::call("rk4");
    rc |= H::call("bc_all");
    if(turb) rc |= H::call("turb_wall");
    const char* e = getenv("ATJUP_MCA");
    const int mca = e ? std::max(0, atoi(e)) : 1;
    if(mca > 0 && iter % mca == 0) rc |= H::call("mca");
    rc |= H::call("clamp_species");
    rc |= H::call("restore");
    quiet_end();
    return rc;
}

// The OpenMP thread count of the reference's parallel loops (the CPU baseline states it)
void ref_set_threads(int n){ if(n > 0) omp_set_num_threads(n); }
int ref_max_threads(void){ return omp_get_max_threads(); }

// the continuity residual's cell as steadyQuery leaves it (cJupiterModel.h:121)
void ref_get_res(int* ijk){ H::res(ijk); }

void ref_set_vel_star(const double* src){
    int im, jm, km; H::dims(&im, &jm, &km);
    for(int j = 0; j < jm; j++) std::memcpy(H::vel_star_row(j), src + (size_t)j * km, sizeof(double) * km);
}
void ref_get_vel_star(double* dst){
    int im, jm, km; H::dims(&im, &jm, &km);
    for(int j = 0; j < jm; j++) std::memcpy(dst + (size_t)j * km, H::vel_star_row(j), sizeof(double) * km);
}

}  // extern "C"
