/*
 * atjup_hip_step.hpp - the reference-side adapter: cJupiterModel's time loop body on the MI355X
 * library (include/atjup_hip.h).
 *
 * A maintainer adds this header to the reference tree and one line to cJupiterModel's friend list
 * (planet/cJupiterModel.h:68-84), beside the other class templates instantiated on the model:
 *
 *     template<class M> friend class HipStep;
 *
 * HipStep<cJupiterModel> then takes the model after its start-up (Run(), planet/cJupiterModel.cpp:
 * 316-745: constants, geometry, SeaMount, the initial state, the boundary passes and restoreVar),
 * copies the state onto the device once, and stands for the on-path calls of one loop iteration
 * (planet/cJupiterModel.cpp:1010-1555) with step(). The fields stay on the device between steps;
 * download_state() brings them back into the model's Arrays, in the reference's own double*** row
 * storage (lib/Array.cpp:43-73), for output, restart files or host code outside the path.
 *
 * The knob profile is read from the same ATJUP_* variables, with the same defaults, as the
 * reference's own readers (their file:line beside each); a profile the library does not implement
 * is refused by atjup_hip_ctx_create with a message, which the constructor throws.
 *
 * Link with -L<repo>/atjup_amd -latjup_hip -Wl,-rpath,<repo>/atjup_amd; the library needs ROCm's
 * libamdhip64.so.7 at run time.
 */
#ifndef ATJUP_HIP_STEP_HPP
#define ATJUP_HIP_STEP_HPP

#include "atjup_hip.h"

#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

template<class M>
class HipStep {
public:
    explicit HipStep(M& m, int precision = ATJUP_PRECISION_FP64) : m_(m) {
        atjup_grid g{};
        g.im = M::im; g.jm = M::jm; g.km = M::km;
        g.r0 = m.rad.z[0]; g.dr = M::dr; g.dthe = M::dthe; g.dphi = M::dphi; g.L_atm = m.L_atm;

        atjup_consts c{};
        c.dt = m.dt; c.re = m.re; c.pr = m.pr; c.g = m.g; c.omega = m.omega; c.u_0 = m.u_0; c.t_ref = m.t_ref;
        c.Coriolis = m.Coriolis; c.centrifugal = m.centrifugal; c.buoyancy = m.buoyancy;
        c.chemical_reaction = m.chemical_reaction;
        c.r_mix = m.r_mix; c.R_mix = m.R_mix; c.cp_mix = m.cp_mix; c.mue_mix = m.mue_mix;
        c.sc_h2o = m.sc_h2o; c.sc_h2s = m.sc_h2s; c.sc_nh3 = m.sc_nh3; c.sc_nh4sh = m.sc_nh4sh; c.sc_ch4 = m.sc_ch4;
        c.rg_nh4sh = m.rg_nh4sh; c.r_p_nh4sh = m.r_p_nh4sh;
        // Turbulence<P>'s re_turb = vel_star_ref z_0 / nue_gas of its default construction
        // (planet/Turbulence.h:245-251); only the k-epsilon sources read it
        c.re_turb = 0.4 * 0.1 / 1.8e-5;

        atjup_knobs k;
        atjup_hip_default_knobs(&k);
        const std::string turb = env_str("ATJUP_TURB_MODEL", m.turb_model);                // Turbulence.h:254
        const bool turb_on = env_str("ATJUP_TURB", "").empty()                              // cJupiterModel.cpp:73-78
            ? !(turb.empty() || turb == "none" || turb == "laminar") : env_int("ATJUP_TURB", 1) != 0;
        k.turb_model = !turb_on ? -1 : turb == "k_epsilon" ? 0 : turb == "k_omega" ? 1 : 2;  // Turbulence.h:197-202
        k.press_sweeps = env_int("ATJUP_PRESS_SWEEPS", 1);                                    // PressureSolver.h:174
        if(k.press_sweeps < 1) k.press_sweeps = 1;
        k.pdyn_cap = env_double("ATJUP_PDYN_CAP", 0.2);                                       // PressureSolver.h:168
        k.sinthe_min = M::sinthe_min();                                                       // cJupiterModel.h:1121-1128
        k.t_diff_scale = env_double("ATJUP_T_DIFF_SCALE", 0.01);                              // cJupiterModel.h:468
        k.species_diff_scale = env_double("ATJUP_SPECIES_DIFF_SCALE", 0.01);                  // cJupiterModel.h:471
        k.rad_coupling = env_double("ATJUP_RAD_COUPLING", 1.0);                               // RHS_Jup_Turb.cpp:646
        k.precip_coupling = env_double("ATJUP_PRECIP_COUPLING", 1.0);                         // RHS_Jup_Turb.cpp:701
        k.turb_coupling = env_double("ATJUP_TURB_COUPLING", 1.0);                             // RHS_Jup_Turb.cpp:729
        k.tke_max_phys = env_double("ATJUP_TKE_MAX", 1000.0);                                 // RungeKutta_Jup_Turb.cpp:277
        k.nue_max_phys = env_double("ATJUP_NUE_MAX", 1.0e5);                                  // Turbulence.h:124
        k.wall_nue_factor = env_double("ATJUP_WALL_NUE", 4.0);                                // InitValues_Jup.cpp:86
        k.wall_nue_layers = env_int("ATJUP_WALL_NUE_LAYERS", 3);                              // InitValues_Jup.cpp:88
        k.rk_scalar_sync = env_int("ATJUP_RK_SCALAR_SYNC", 2);                                // cJupiterModel.cpp:1118
        k.turb_k_bg_phys = env_double("ATJUP_TURB_K_BG", 1.0);                                // Turbulence.h:157
        k.radiation = env_int("ATJUP_RADIATION", 1);                                          // cJupiterModel.cpp:52
        k.precip = env_int("ATJUP_PRECIP", 1);                                                // cJupiterModel.cpp:83
        k.mca = env_int("ATJUP_MCA", 1);                                                      // cJupiterModel.cpp:123
        if(k.mca < 0) k.mca = 0;
        k.nh4sh_eq = env_int("ATJUP_NH4SH_EQ", 1) != 0;                                       // ChemistryJup.h:286
        k.chemistry = 1;                        // the physics block always runs it (cJupiterModel.cpp:1146-1161)
        k.dry = 0;

        ctx_ = atjup_hip_ctx_create(&g, &c, &k, precision);
        if(!ctx_) throw std::runtime_error(std::string("atjup_hip_ctx_create: ") + atjup_hip_last_error_string());

        set_gases();
        set_physics();
        check(atjup_hip_upload_mask(ctx_, flat(m.SeaMount).data()), "upload_mask");
        std::vector<double> vs((size_t)M::jm * M::km);
        for(int j = 0; j < M::jm; j++)
            for(int k2 = 0; k2 < M::km; k2++) vs[(size_t)j * M::km + k2] = m.vel_star.y[j][k2];
        check(atjup_hip_upload_vel_star(ctx_, vs.data()), "upload_vel_star");
        upload_state();
    }

    ~HipStep(){ atjup_hip_ctx_destroy(ctx_); }
    HipStep(const HipStep&) = delete;
    HipStep& operator=(const HipStep&) = delete;

    // One iteration of the loop body, in place of the on-path calls of cJupiterModel.cpp:1010-1555.
    void step(int iter_n, bool first){ check(atjup_hip_step(ctx_, iter_n, first ? 1 : 0), "step"); }

    // The NaN watch: true when every prognostic field is finite.
    bool finite(){ return check(atjup_hip_check_finite(ctx_), "check_finite") == 0; }

    // One field in the reference's Array row storage, by atjup_field id.
    void upload(int id, Array& a){ check(atjup_hip_upload_field_rows(ctx_, id, a.x), "upload_field_rows"); }
    void download(int id, Array& a){ check(atjup_hip_download_field_rows(ctx_, id, a.x), "download_field_rows"); }

    // Every field the context holds (the profile decides which), host to device and back.
    void upload_state(){
        for(auto& f : fields()) if(atjup_hip_field_device_ptr(ctx_, f.first)) upload(f.first, *f.second);
    }
    void download_state(){
        for(auto& f : fields()) if(atjup_hip_field_device_ptr(ctx_, f.first)) download(f.first, *f.second);
        std::vector<double> vs((size_t)M::jm * M::km);
        check(atjup_hip_download_vel_star(ctx_, vs.data()), "download_vel_star");
        for(int j = 0; j < M::jm; j++)
            for(int k2 = 0; k2 < M::km; k2++) m_.vel_star.y[j][k2] = vs[(size_t)j * M::km + k2];
    }

    atjup_ctx* context(){ return ctx_; }

private:
    static int check(int rc, const char* what){
        if(rc < 0) throw std::runtime_error(std::string(what) + ": " + atjup_hip_last_error_string());
        return rc;
    }
    static std::string env_str(const char* name, const std::string& fallback){
        const char* e = std::getenv(name);
        return e ? std::string(e) : fallback;
    }
    static int env_int(const char* name, int fallback){
        const char* e = std::getenv(name);
        return e ? std::atoi(e) : fallback;
    }
    static double env_double(const char* name, double fallback){
        const char* e = std::getenv(name);
        return e ? std::atof(e) : fallback;
    }
    static std::vector<double> flat(const Array& a){
        std::vector<double> v((size_t)a.im * a.jm * a.km);
        for(int i = 0; i < a.im; i++)
            for(int j = 0; j < a.jm; j++)
                for(int k = 0; k < a.km; k++) v[((size_t)i * a.jm + j) * a.km + k] = a.x[i][j][k];
        return v;
    }

    // The three gases as the time loop hands them to SaturationAdjustmentJup::run
    // (planet/cJupiterModel.cpp:1060-1080); the step and the precipitation read them from here.
    void set_gases(){
        M& m = m_;
        const atjup_gas_params gas[3] = {
            { m.t_0_h2o, m.t_00_h2o, m.ep_h2o, m.lv_h2o, m.ls_h2o, m.C_h2o, m.L0_h2o, m.R_h2o, m.del_alf_h2o,
              m.del_bet_h2o, m.C_h2o_ice, m.L0_h2o_ice, m.del_alf_h2o_ice, m.del_bet_h2o_ice },
            { m.t_0_nh3, m.t_00_nh3, m.ep_nh3, m.lv_nh3, m.ls_nh3, m.C_nh3, m.L0_nh3, m.R_nh3, m.del_alf_nh3,
              m.del_bet_nh3, m.C_nh3_ice, m.L0_nh3_ice, m.del_alf_nh3_ice, m.del_bet_nh3_ice },
            { m.t_0_ch4, m.t_00_ch4, m.ep_ch4, m.lv_ch4, m.ls_ch4, m.C_ch4, m.L0_ch4, m.R_ch4, m.del_alf_ch4,
              m.del_bet_ch4, m.C_ch4_ice, m.L0_ch4_ice, m.del_alf_ch4_ice, m.del_bet_ch4_ice } };
        for(int s = 0; s < 3; s++) check(atjup_hip_set_gas(ctx_, s, &gas[s]), "set_gas");
    }

    // The radiation's planet properties and knobs (planet/cJupiterModel.h:1109-1118,
    // planet/Radiation.h:270-295), the precipitation's rate scale and NH4SH band
    // (planet/Precipitation.h:111-115, cJupiterModel.h:194-195), the chemistry's constants and knobs
    // (planet/ChemistryJup.h:54-57, :123-132) and the moist convective adjustment's
    // (planet/ConvectiveAdjustment.h:409-413, InitValues_Jup.cpp:945).
    void set_physics(){
        M& m = m_;
        atjup_rad_params r;
        atjup_hip_default_rad_params(&r);
        r.F_int = M::rad_F_int(); r.S_solar = M::rad_S_solar(); r.albedo_bond = M::rad_albedo_bond();
        r.x_H2 = M::rad_x_H2(); r.x_He = M::rad_x_He();
        r.cia_strength = env_double("ATJUP_CIA_STRENGTH", M::rad_default_cia_strength());
        r.opacity_strength = env_double("ATJUP_OPACITY_STRENGTH", 1.0);
        r.solar = env_int("ATJUP_SOLAR", 1);
        r.solar_strength = env_double("ATJUP_SOLAR_STRENGTH", 1.0);
        r.sw_tau_per_bar = env_double("ATJUP_SW_TAU_PER_BAR", 1.0);
        r.bottom = env_int("ATJUP_RAD_BOTTOM", M::rad_default_bottom());
        check(atjup_hip_set_radiation(ctx_, &r), "set_radiation");

        atjup_precip_params p;
        atjup_hip_default_precip_params(&p);
        p.rate_scale = env_double("ATJUP_PRECIP_SCALE", M::precip_rate_scale());
        p.t_0_nh4sh = m.t_0_nh4sh; p.t_00_nh4sh = m.t_00_nh4sh;
        check(atjup_hip_set_precipitation(ctx_, &p), "set_precipitation");

        atjup_chem_params h;
        atjup_hip_default_chem_params(&h);
        h.m_nh3 = m.m_nh3; h.m_h2s = m.m_h2s; h.m_nh4sh = m.m_nh4sh; h.m_h2 = m.m_h2; h.m_he = m.m_he; h.m_h2o = m.m_h2o;
        h.X_h2 = m.X_h2; h.X_he = m.X_he; h.X_h2o = m.X_h2o; h.X_h2s = m.X_h2s; h.X_nh3 = m.X_nh3; h.X_nh4sh = m.X_nh4sh;
        h.mue_nh3 = m.mue_nh3; h.mue_h2s = m.mue_h2s; h.mue_nh4sh = m.mue_nh4sh;
        h.rg_nh3 = m.rg_nh3; h.rg_h2s = m.rg_h2s; h.k_mix = m.k_mix;
        h.cp_nh3 = m.cp_nh3; h.cp_h2s = m.cp_h2s; h.cp_nh4sh = m.cp_nh4sh;
        h.k_nh3 = m.k_nh3; h.k_h2s = m.k_h2s; h.k_nh4sh = m.k_nh4sh;
        h.dh_nh4sh = m.dh_nh4sh; h.t_0_nh4sh = m.t_0_nh4sh; h.t_00_nh4sh = m.t_00_nh4sh;
        h.gate_zero = env_int("ATJUP_CHEM_GATE_ZERO", 1);
        h.molar_conc = env_int("ATJUP_CHEM_MOLAR_CONC", 0);
        h.enthalpy = env_int("ATJUP_CHEM_ENTHALPY", 0);
        check(atjup_hip_set_chemistry(ctx_, &h), "set_chemistry");

        atjup_mca_params a;
        atjup_hip_default_mca_params(&a);       // the NH4SH heat follows the context's nh4sh_eq
        a.ledoux = env_int("ATJUP_MCA_LEDOUX", 1) != 0;
        a.loading = env_int("ATJUP_MCA_LOADING", 1) != 0;
        const double relax = env_double("ATJUP_MCA_RELAX", 1.0);
        a.relax = relax < 0.0 ? 0.0 : (relax > 1.0 ? 1.0 : relax);
        const int passes = env_int("ATJUP_MCA_PASSES", 64);
        a.max_pass = passes > 0 ? passes : 64;
        a.ep_h2s = m.ep_h2s;
        if(env_int("ATJUP_NH4SH_EQ", 1) != 0) a.l_nh4sh = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;
        check(atjup_hip_set_mca(ctx_, &a), "set_mca");
    }

    // atjup_field ids and the model's Arrays they stand for (planet/cJupiterModel.h:1362-1544)
    std::vector<std::pair<int, Array*>> fields(){
        M& m = m_;
        return {
            {ATJUP_F_T, &m.t}, {ATJUP_F_U, &m.u}, {ATJUP_F_V, &m.v}, {ATJUP_F_W, &m.w},
            {ATJUP_F_H2O, &m.h2o}, {ATJUP_F_H2O_CLOUD, &m.h2o_cloud}, {ATJUP_F_H2O_ICE, &m.h2o_ice},
            {ATJUP_F_H2S, &m.h2s}, {ATJUP_F_NH3, &m.nh3}, {ATJUP_F_NH3_CLOUD, &m.nh3_cloud},
            {ATJUP_F_NH3_ICE, &m.nh3_ice}, {ATJUP_F_CH4, &m.ch4}, {ATJUP_F_CH4_CLOUD, &m.ch4_cloud},
            {ATJUP_F_CH4_ICE, &m.ch4_ice}, {ATJUP_F_NH4SH, &m.nh4sh}, {ATJUP_F_TKE, &m.tke}, {ATJUP_F_DIS, &m.dis},
            {ATJUP_F_P_DYN, &m.p_dyn}, {ATJUP_F_P_STAT, &m.p_stat}, {ATJUP_F_P_HYDRO, &m.p_hydro},
            {ATJUP_F_RHO_MIX, &m.rho_mix}, {ATJUP_F_NUE, &m.nue}, {ATJUP_F_PROD, &m.prod},
            {ATJUP_F_TKE_SOURCE, &m.tke_source}, {ATJUP_F_DIS_SOURCE, &m.dis_source},
            {ATJUP_F_AUX_U, &m.aux_u}, {ATJUP_F_AUX_V, &m.aux_v}, {ATJUP_F_AUX_W, &m.aux_w},
            {ATJUP_F_Q_RAD, &m.Q_rad}, {ATJUP_F_Q_PRECIP, &m.Q_precip},
            {ATJUP_F_S_PRECIP_H2O, &m.S_precip_h2o}, {ATJUP_F_S_PRECIP_H2O_CLOUD, &m.S_precip_h2o_cloud},
            {ATJUP_F_S_PRECIP_H2O_ICE, &m.S_precip_h2o_ice}, {ATJUP_F_S_PRECIP_NH3, &m.S_precip_nh3},
            {ATJUP_F_S_PRECIP_NH3_CLOUD, &m.S_precip_nh3_cloud}, {ATJUP_F_S_PRECIP_NH3_ICE, &m.S_precip_nh3_ice},
            {ATJUP_F_S_PRECIP_CH4, &m.S_precip_ch4}, {ATJUP_F_S_PRECIP_CH4_CLOUD, &m.S_precip_ch4_cloud},
            {ATJUP_F_S_PRECIP_CH4_ICE, &m.S_precip_ch4_ice},
            {ATJUP_F_MASSFLUX_H2S, &m.massflux_h2s}, {ATJUP_F_MASSFLUX_NH3, &m.massflux_nh3},
            {ATJUP_F_MASSFLUX_NH4SH, &m.massflux_nh4sh}, {ATJUP_F_FLUXLIM_NH4SH, &m.fluxlim_nh4sh},
            {ATJUP_F_EPSILON, &m.epsilon}, {ATJUP_F_RADIATION, &m.radiation},
            {ATJUP_F_P_RAIN, &m.P_rain}, {ATJUP_F_P_SNOW, &m.P_snow}, {ATJUP_F_P_GRAUPEL, &m.P_graupel},
            {ATJUP_F_P_NH3_RAIN, &m.P_nh3_rain}, {ATJUP_F_P_NH3_SNOW, &m.P_nh3_snow},
            {ATJUP_F_P_NH3_GRAUPEL, &m.P_nh3_graupel}, {ATJUP_F_P_CH4_RAIN, &m.P_ch4_rain},
            {ATJUP_F_P_CH4_SNOW, &m.P_ch4_snow}, {ATJUP_F_P_CH4_GRAUPEL, &m.P_ch4_graupel},
            {ATJUP_F_P_NH4SH, &m.P_nh4sh},
            {ATJUP_F_DIFFLUX_H2S, &m.difflux_h2s}, {ATJUP_F_DIFFLUX_NH3, &m.difflux_nh3},
            {ATJUP_F_DIFFLUX_NH4SH, &m.difflux_nh4sh}, {ATJUP_F_THERMALMASSFLUX, &m.thermalmassflux},
        };
    }

    M& m_;
    atjup_ctx* ctx_ = nullptr;
};

#endif /* ATJUP_HIP_STEP_HPP */
