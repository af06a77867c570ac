// CPU oracle of the ATJUP time step (test infrastructure; see cpu_ref.h).
//
// Every routine restates one reference routine for the knob profile of include/atjup_hip.h and
// keeps the reference's floating-point expression trees (operand order and association) so that,
// compiled with -ffp-contract=off against the same libm, it reproduces the compiled reference bit
// for bit. Fields are contiguous [i][j][k] with k fastest; the grid size is a runtime parameter.
#include "cpu_ref.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NP = ATJUP_N_PROG;

struct Geo {
    double rm, rm2, exp_rm, exp_2_rm;
    double sinthe, sinthe2, costhe, cotanthe;
    double inv_rm, inv_rm2, inv_rmsinthe, inv_rm2sinthe, inv_rm2sinthe2, costhe_inv_rm2sinthe;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
};

inline bool nonfinite_bits(double v){
    std::uint64_t b; std::memcpy(&b, &v, sizeof b);
    return (b & 0x7FF0000000000000ULL) == 0x7FF0000000000000ULL;
}
// RungeKutta_Jup_Turb.cpp:285-290, BC_Jup.h:627-632
inline double safe_clamp(double v, double lo, double hi){
    if(nonfinite_bits(v)) return lo;
    return (v < lo) ? lo : ((v > hi) ? hi : v);
}

// J per kg NH4SH formed: the van't Hoff enthalpy of the Lewis K_p, 4705 ln(10) R* / 51.1114
// (ChemistryJup.h:282-283)
inline double nh4sh_formation_heat(){ return 4705.0 * std::log(10.0) * 8314.46 / 51.1114; }

}  // namespace

struct orc_ctx {
    atjup_grid g;
    atjup_consts c;
    atjup_knobs kn;
    int im, jm, km;
    size_t n3;
    std::vector<double> rad, the, sin_the, cos_the, sinthe_rk;
    std::vector<float> lh;                           // get_layer_height [km]
    std::vector<std::vector<double>> f;              // by field id; NLEVEL + id for the n level
    std::vector<std::vector<double>> rhs, acc;       // rhs_* and acc_* of the 17 prognostic fields
    std::vector<double> mask;                        // SeaMount
    std::vector<int> itopo;                          // i_topography [j][k]
    std::vector<double> vel_star;                    // [j][k]
    std::vector<double> brl;                         // buoy_ref_level [i]
    atjup_gas_params gas[3] = {};                    // saturation-adjustment constants per gas
    atjup_rad_params radp = {};                      // radiation properties and knobs
    atjup_precip_params precp = { 1.0, 230.0, 200.0 }; // precipitation rate scale and NH4SH band
    std::vector<double> precip_srf;                  // [5][jm][km] surface maps
    atjup_chem_params chemp = { 17.03052, 34.08088, 51.1114, 2.016, 4.0026, 18.01588,
                                0.864, 0.136, 0.0017, 7.7e-05, 0.0002, 0.0,
                                9.2e-06, 1.3e-05, 0.0, 817.0, 949.0, 0.12478002943205087,
                                2190.0, 2240.0, 2000.0, 0.02102, 0.013, 0.0, 1768700.0, 230.0, 200.0,
                                1, 0, 0 };             // the chemistry's constants (cJupiterModel.h:172-401)
    std::vector<double> w_nh3, w_h2s, w_nh4sh;       // the reaction rates, persistent between calls
    atjup_mca_params mcap = { 1.0, 14.474287352261996, 0.0, 1, 1, 64 };   // moist convective adjustment knobs
    double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;

    size_t id(int i, int j, int k) const { return ((size_t)i * jm + j) * km + k; }
    std::vector<double>& F(int fid){ return f[fid]; }
    bool solid(int i, int j, int k) const { return mask[id(i, j, k)] == 1.0; }
    double p_dyn_to_bar() const { return c.r_mix * c.u_0 * c.u_0 * 1.0e-5; }   // cJupiterModel.h:1193-1195
    double get_layer_height(int i) const { return (i < 0 || i > im - 1) ? -1.0f : lh[i]; }
    double layer_thickness_m(int i) const {                                  // cJupiterModel.h:147-150
        if(i < 0 || i > im - 2) return -1.0;
        return (double)(lh[i + 1] - lh[i]) * 1.0e3;
    }
    int wrap(int kk) const {                                                 // ATPhys.h:132-136
        const int n = km - 1;
        return kk < 0 ? kk + n : (kk >= n ? kk - n : kk);
    }
    Geo geo(int i, int j) const {                                            // RungeKutta_Jup_Turb.cpp:390-411
        Geo G;
        G.rm = rad[i];
        G.rm2 = G.rm * G.rm;
        G.exp_rm = 1.0;                                                      // coord_stretching off
        G.exp_2_rm = G.exp_rm * G.exp_rm;
        G.sinthe = sinthe_rk[j];
        G.sinthe2 = G.sinthe * G.sinthe;
        G.costhe = cos_the[j];
        G.cotanthe = G.costhe / G.sinthe;
        G.inv_rm = 1.0 / G.rm;
        G.inv_rm2 = 1.0 / G.rm2;
        G.inv_rmsinthe = 1.0 / (G.rm * G.sinthe);
        G.inv_rm2sinthe = G.inv_rm2 / G.sinthe;
        G.inv_rm2sinthe2 = G.inv_rm2 / G.sinthe2;
        G.costhe_inv_rm2sinthe = G.costhe * G.inv_rm2sinthe;
        G.inv_2dr = 1.0 / (2.0 * g.dr);
        G.inv_2dthe = 1.0 / (2.0 * g.dthe);
        G.inv_2dphi = 1.0 / (2.0 * g.dphi);
        G.inv_dr2 = 1.0 / (g.dr * g.dr);
        G.inv_dthe2 = 1.0 / (g.dthe * g.dthe);
        G.inv_dphi2 = 1.0 / (g.dphi * g.dphi);
        return G;
    }
};

namespace {

// ---- computeBuoyancyRefLevel (RungeKutta_Jup_Turb.cpp:47-69) -------------------------------
void buoy_ref_level(orc_ctx& o){
    const auto& t = o.f[ATJUP_F_T];
    const auto& ps = o.f[ATJUP_F_P_STAT];
    o.brl.assign(o.im, 0.0);
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < o.im; i++){
        double sum = 0.0, wsum = 0.0;
        for(int j = 0; j < o.jm; j++){
            const double wgt = o.sin_the[j];
            for(int k = 0; k < o.km; k++){
                const size_t q = o.id(i, j, k);
                if(o.mask[q] == 1.0) continue;
                if(!(t[q] > 0.0)) continue;
                const double b = o.c.g * ps[q] / (o.c.r_mix * o.c.R_mix * t[q] * o.c.t_ref);
                if(!std::isfinite(b)) continue;
                sum += wgt * b;
                wsum += wgt;
            }
        }
        o.brl[i] = (wsum > 0.0) ? sum / wsum : 0.0;
    }
}

// ---- computeHydrostaticPressure, ATJUP_METRIC_SPLIT + CORIOLIS_SPLIT, no HYDRO_SPLIT, no
//      p_baro (RungeKutta_Jup_Turb.cpp:119-218) ------------------------------------------------
void hydrostatic(orc_ctx& o){
    auto& ph = o.f[ATJUP_F_P_HYDRO];
    const auto& v = o.f[ATJUP_F_V];
    const auto& w = o.f[ATJUP_F_W];
    const int im = o.im;
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < o.jm; j++){
        for(int k = 0; k < o.km; k++){
            int i0 = 0;
            while(i0 < im && o.solid(i0, j, k)){ ph[o.id(i0, j, k)] = 0.0; i0++; }
            if(i0 >= im) continue;
            int i1 = i0;
            while(i1 + 1 < im && !o.solid(i1 + 1, j, k)) i1++;
            for(int i = i1 + 1; i < im; i++) ph[o.id(i, j, k)] = 0.0;
            // nd_cor: ATJUP_ND_COR unset and GEO_BALANCE off -> ATJUP_NONDIM (unset) -> 1
            const double cor_ = o.c.Coriolis * 1.0 * 2.0 * o.c.omega * std::abs(o.sin_the[j]);
            auto f_ = [&](int i) -> double {
                const size_t q = o.id(i, j, k);
                const double exp_rm_ = 1.0;
                return 0.0 + ((v[q] * v[q] + w[q] * w[q]) / o.rad[i] + cor_ * w[q]) / exp_rm_;
            };
            ph[o.id(i0, j, k)] = 0.0;
            for(int i = i0 + 1; i <= i1; i++)
                ph[o.id(i, j, k)] = ph[o.id(i - 1, j, k)] + 0.5 * (f_(i - 1) + f_(i)) * o.g.dr;
        }
    }
}

// ---- RHSJup (RHS_Jup_Turb.cpp:31-1457) for one cell ------------------------------------------
// The profile: ATJUP_NONDIM unset (nd factors 1), PGRAD_WALL/WRAP 1, ADIABATIC(_RHO) 1, METRIC_SPLIT
// and CORIOLIS_SPLIT 1, HYDRO_SPLIT 0, CELL_HOLD 0, TURB_BUOY 0, SPECIES_DIVU 0, CHROM 0.
void rhs_cell(orc_ctx& o, int i, int j, int k, const Geo& G, int turb_sel){
    const atjup_consts& C = o.c;
    const atjup_knobs& K = o.kn;
    const int kw = o.wrap(k - 1), ke = o.wrap(k + 1);
    auto& F = o.f;
    const size_t q = o.id(i, j, k);
    const size_t qip = o.id(i + 1, j, k), qim = o.id(i - 1, j, k);
    const size_t qjp = o.id(i, j + 1, k), qjm = o.id(i, j - 1, k);
    const size_t qkp = o.id(i, j, ke), qkm = o.id(i, j, kw);

    const double u_ijk = F[ATJUP_F_U][q], v_ijk = F[ATJUP_F_V][q], w_ijk = F[ATJUP_F_W][q];

    auto pgrad = [&](const std::vector<double>& P, size_t qh, size_t ql, double inv_2d) -> double {
        const double loc = P[q];
        double hi = P[qh], lo = P[ql];
        if(o.mask[qh] == 1.0) hi = loc;
        if(o.mask[ql] == 1.0) lo = loc;
        return (hi - lo) * inv_2d;
    };

    double d1r[NP], d2r[NP], d1t[NP], d2t[NP], d1p[NP], d2p[NP];
    for(int fi = 0; fi < NP; fi++){
        const auto& X = F[fi];
        d1r[fi] = (X[qip] - X[qim]) * G.inv_2dr * G.exp_rm;
        d2r[fi] = (X[qip] - 2.0 * X[q] + X[qim]) * G.inv_dr2 * G.exp_2_rm;
        d1t[fi] = (X[qjp] - X[qjm]) * G.inv_2dthe;
        d2t[fi] = (X[qjp] - 2.0 * X[q] + X[qjm]) * G.inv_dthe2;
        d1p[fi] = (X[qkp] - X[qkm]) * G.inv_2dphi;
        d2p[fi] = (X[qkp] - 2.0 * X[q] + X[qkm]) * G.inv_dphi2;
    }
    const auto& pd = F[ATJUP_F_P_DYN];
    const double dpdr = pgrad(pd, qip, qim, G.inv_2dr) * G.exp_rm;
    const double dpdthe = pgrad(pd, qjp, qjm, G.inv_2dthe);
    const double dpdphi = pgrad(pd, qkp, qkm, G.inv_2dphi);

    // Neumann k*/dis* at solid faces (RHS_Jup_Turb.cpp:265-303)
    {
        const bool sim = o.mask[qim] == 1.0, sip = o.mask[qip] == 1.0;
        const bool sjm = o.mask[qjm] == 1.0, sjp = o.mask[qjp] == 1.0;
        const bool skm = o.mask[qkm] == 1.0, skp = o.mask[qkp] == 1.0;
        for(int fi = ATJUP_F_TKE; fi <= ATJUP_F_DIS; fi++){
            const auto& X = F[fi];
            if(sim || sip){
                const double a = sim ? X[q] : X[qim], b = sip ? X[q] : X[qip];
                d1r[fi] = (b - a) * G.inv_2dr * G.exp_rm;
                d2r[fi] = (b - 2.0 * X[q] + a) * G.inv_dr2 * G.exp_2_rm;
            }
            if(sjm || sjp){
                const double a = sjm ? X[q] : X[qjm], b = sjp ? X[q] : X[qjp];
                d1t[fi] = (b - a) * G.inv_2dthe;
                d2t[fi] = (b - 2.0 * X[q] + a) * G.inv_dthe2;
            }
            if(skm || skp){
                const double a = skm ? X[q] : X[qkm], b = skp ? X[q] : X[qkp];
                d1p[fi] = (b - a) * G.inv_2dphi;
                d2p[fi] = (b - 2.0 * X[q] + a) * G.inv_dphi2;
            }
        }
    }

    const double nd_cor = 1.0, nd_cent = 1.0, nd_buoy = 1.0;
    const double costhe = G.costhe;
    const double sinthe_true = std::sqrt(std::max(0.0, 1.0 - costhe * costhe));
    const double Coriolis_rad = nd_cor * -2.0 * C.omega * sinthe_true * w_ijk;
    const double Coriolis_the = nd_cor * -2.0 * C.omega * costhe * w_ijk;
    (void)Coriolis_rad;
    const double Coriolis_phi = nd_cor * +2.0 * C.omega * (+costhe * v_ijk + sinthe_true * u_ijk);
    const double centrifugal_rad = nd_cent * C.omega * C.omega * G.rm * sinthe_true * sinthe_true;
    const double centrifugal_the = nd_cent * C.omega * C.omega * G.rm * sinthe_true * costhe;
    const double coeff_energy_p = C.u_0 * C.u_0 / (C.cp_mix * C.t_ref);

    const double v_invrm = v_ijk * G.inv_rm;
    const double w_invrs = w_ijk * G.inv_rmsinthe;

    // adiabatic compression with the local density (RHS_Jup_Turb.cpp:461-482)
    double dpstatdr;
    {
        const auto& ps = F[ATJUP_F_P_STAT];
        const double to_nd = 1.0 / o.p_dyn_to_bar();
        dpstatdr = (ps[qip] - ps[qim]) * G.inv_2dr * G.exp_rm * to_nd;
        const double rho_l = F[ATJUP_F_RHO_MIX][q];
        if(rho_l > 0.0 && std::isfinite(rho_l)) dpstatdr *= C.r_mix / rho_l;
    }
    const double pressure_t = coeff_energy_p
        * (u_ijk * (dpdr + dpstatdr) + v_invrm * dpdthe + w_invrs * dpdphi);

    auto transport = [&](int fi){ return u_ijk * d1r[fi] + v_invrm * d1t[fi] + w_invrs * d1p[fi]; };
    const double metric_on = 0.0;
    const double cor_rad_on = 0.0;
    const double transport_t = transport(ATJUP_F_T);
    const double transport_u = u_ijk * d1r[1] + v_invrm * d1t[1] + w_invrs * d1p[1]
        - metric_on * (v_ijk * v_ijk + w_ijk * w_ijk) * G.inv_rm;
    const double transport_v = u_ijk * d1r[2] + v_invrm * d1t[2] + w_invrs * d1p[2]
        + (u_ijk * v_ijk - w_ijk * w_ijk * G.cotanthe) * G.inv_rm;
    const double transport_w = u_ijk * d1r[3] + v_invrm * d1t[3] + w_invrs * d1p[3]
        + (w_ijk * u_ijk + v_ijk * w_ijk * G.cotanthe) * G.inv_rm;

    const double two_inv_rm = 2.0 * G.inv_rm;
    const double v_metric = (1.0 + costhe / (G.sinthe2)) * G.inv_rm2;
    auto diffusion = [&](int fi){
        return d2r[fi] + d1r[fi] * two_inv_rm + d2t[fi] * G.inv_rm2
             + d1t[fi] * G.costhe_inv_rm2sinthe + d2p[fi] * G.inv_rm2sinthe2;
    };
    const double diffusion_t = diffusion(ATJUP_F_T);
    const double diffusion_u = d2r[1] + 2.0 * u_ijk * G.inv_rm2 + d2t[1] * G.inv_rm2
        + 4.0 * d1r[1] * G.inv_rm + d1t[1] * G.costhe_inv_rm2sinthe
        + d2p[1] * G.inv_rm2sinthe2;
    const double diffusion_v = d2r[2] + d1r[2] * two_inv_rm + d2t[2] * G.inv_rm2
        + d1t[2] * G.costhe_inv_rm2sinthe - v_metric * v_ijk
        + d2p[2] * G.inv_rm2sinthe2
        + 2.0 * d1t[1] * G.inv_rm2
        - d1p[3] * 2.0 * costhe * G.inv_rm2sinthe2;
    const double diffusion_w = d2r[3] + d1r[3] * two_inv_rm + d2t[3] * G.inv_rm2
        + d1t[3] * G.costhe_inv_rm2sinthe - v_metric * w_ijk
        + d2p[3] * G.inv_rm2sinthe2
        + 2.0 * d1p[1] * G.inv_rm2sinthe
        + d1p[2] * 2.0 * costhe * G.inv_rm2sinthe2;

    const double pgrad_scale = 1.0;
    const double dpdr_term = pgrad_scale * dpdr;
    const double dpdthe_term = pgrad_scale * dpdthe * G.inv_rm;
    const double dpdphi_term = pgrad_scale * dpdphi * G.inv_rmsinthe;
    const auto& ph = F[ATJUP_F_P_HYDRO];
    const double dphdthe_term = pgrad(ph, qjp, qjm, G.inv_2dthe) * G.inv_rm;
    const double dphdphi_term = pgrad(ph, qkp, qkm, G.inv_2dphi) * G.inv_rmsinthe;

    // radiative heating (RHS_Jup_Turb.cpp:646-676)
    double radiation_t = 0.0;
    if(K.rad_coupling != 0.0){
        const double T_phys = F[ATJUP_F_T][q] * C.t_ref;
        const double P_phys = F[ATJUP_F_P_STAT][q] * 1.0e5;
        const double rho_f = F[ATJUP_F_RHO_MIX][q];
        const double rho = (rho_f > 0.0 && std::isfinite(rho_f))
                         ? rho_f : ((T_phys > 1.0) ? P_phys / (C.R_mix * T_phys) : 0.0);
        const double L_rad = o.g.L_atm * 1.0e3;
        if(rho > 0.0 && C.cp_mix > 0.0){
            radiation_t = K.rad_coupling * F[ATJUP_F_Q_RAD][q] * L_rad / (rho * C.cp_mix * C.u_0 * C.t_ref);
            if(!std::isfinite(radiation_t)) radiation_t = 0.0;
            else if(radiation_t > 0.5) radiation_t = 0.5;
            else if(radiation_t < -0.5) radiation_t = -0.5;
        }
    }
    // precipitation latent heat with the local density (RHS_Jup_Turb.cpp:701-718)
    double precip_t = 0.0;
    if(K.precip_coupling != 0.0){
        const double L_rad = o.g.L_atm * 1.0e3;
        const double rf = F[ATJUP_F_RHO_MIX][q];
        const double rho_p = (rf > 0.0 && std::isfinite(rf)) ? rf : C.r_mix;
        if(rho_p > 0.0 && C.cp_mix > 0.0){
            precip_t = K.precip_coupling * F[ATJUP_F_Q_PRECIP][q] * L_rad / (rho_p * C.cp_mix * C.u_0 * C.t_ref);
            if(!std::isfinite(precip_t)) precip_t = 0.0;
            else if(precip_t > 0.5) precip_t = 0.5;
            else if(precip_t < -0.5) precip_t = -0.5;
        }
    }
    const double Pr_t = 0.9;
    const double nue_c = F[ATJUP_F_NUE][q];
    const double nue_t = (K.turb_coupling != 0.0 && std::isfinite(nue_c)) ? K.turb_coupling * std::max(0.0, nue_c) : 0.0;
    const double nue_t_s = nue_t / Pr_t;
    const double nue_wall = F[ATJUP_F_WALL_NUE][q];

    // k* / dis* sources (RHS_Jup_Turb.cpp:763-998)
    double diffusion_tke_re = 0.0, diffusion_dis_re = 0.0, tke_src = 0.0, dis_src = 0.0;
    const bool turb_on = turb_sel >= 0;
    if(turb_on){
        const double C_nue = 0.028, dis_min = 1.0e-10, nue_gas_phys = 1.8e-5;
        const double L_m = o.g.L_atm * 1.0e3;
        const double nue_max = K.nue_max_phys / (C.u_0 * L_m);
        const double nue_here = std::isfinite(nue_c) ? std::max(0.0, nue_c) : 0.0;
        const double tke0 = F[ATJUP_F_TKE][q], dis0 = F[ATJUP_F_DIS][q];
        const double tke_s = std::max(tke0, dis_min);
        const double dis_s = std::max(dis0, dis_min);
        const double dudr_s = d1r[1], dvdr_s = d1r[2], dwdr_s = d1r[3];
        const double dudthe_s = d1t[1] * G.inv_rm, dvdthe_s = d1t[2] * G.inv_rm, dwdthe_s = d1t[3] * G.inv_rm;
        const double dudphi_s = d1p[1] * G.inv_rmsinthe, dvdphi_s = d1p[2] * G.inv_rmsinthe, dwdphi_s = d1p[3] * G.inv_rmsinthe;
        const double dtkedthe_s = d1t[ATJUP_F_TKE] * G.inv_rm, ddisdthe_s = d1t[ATJUP_F_DIS] * G.inv_rm;
        const double dtkedphi_s = d1p[ATJUP_F_TKE] * G.inv_rmsinthe, ddisdphi_s = d1p[ATJUP_F_DIS] * G.inv_rmsinthe;
        const double grad_dot = d1r[ATJUP_F_TKE] * d1r[ATJUP_F_DIS] + dtkedthe_s * ddisdthe_s + dtkedphi_s * ddisdphi_s;
        double cnue = (turb_sel == 0) ? C_nue * tke_s * tke_s / dis_s : tke_s / dis_s;
        cnue = std::min(cnue, nue_max);
        const double der = 0.66667 * (dudr_s + dvdthe_s + dwdphi_s);
        const double P_full = std::max(0.0,
              (cnue * (2.0 * dudr_s - der) - 0.66667 * tke0) * dudr_s
            + (cnue * (dudthe_s + dvdr_s)) * dudthe_s
            + (cnue * (dudphi_s + dwdr_s)) * dudphi_s
            + (cnue * (2.0 * dvdthe_s - der) - 0.66667 * tke0) * dvdthe_s
            + (cnue * (dvdr_s + dudthe_s)) * dvdr_s
            + (cnue * (dvdphi_s + dwdthe_s)) * dvdphi_s
            + (cnue * (2.0 * dwdphi_s - der) - 0.66667 * tke0) * dwdphi_s
            + (cnue * (dwdr_s + dudphi_s)) * dwdr_s
            + (cnue * (dwdthe_s + dvdphi_s)) * dwdthe_s);
        F[ATJUP_F_PROD][q] = P_full;
        const double y_mount = (double)o.get_layer_height(o.itopo[(size_t)j * o.km + k]) * 1.0e3;
        const double dz_layer = std::max(1.0, o.layer_thickness_m(std::max(i - 1, 0)));
        const double y_phys = std::max((double)o.get_layer_height(i) * 1.0e3 - y_mount, dz_layer);
        const double y_star = y_phys / L_m;
        if(turb_sel == 0){
            const double sig_k = 1.0, sig_w = 1.3, C_eps_1 = 1.35, C_eps_2 = 1.80;
            diffusion_tke_re = 1.0 / C.re + nue_here / sig_k;
            diffusion_dis_re = 1.0 / C.re + nue_here / sig_w;
            const double d_plus = y_phys * o.vel_star[(size_t)j * o.km + k] / nue_gas_phys;
            const double Re_T = tke_s * tke_s * C.u_0 * L_m / (dis_s * nue_gas_phys);
            const double f_2 = 1.0 - 0.4 / 1.8 * std::exp(-Re_T * Re_T / 36.0);
            const double y_star2 = y_star * y_star;
            const double L_k = -2.0 * tke0 / y_star2;
            const double L_w = -2.0 * dis0 / y_star2 * std::exp(-0.5 * d_plus);
            const double P_k = P_full;
            const double Y_k = dis0;
            const double P_w = C_eps_1 * dis0 / tke_s * P_k;
            const double Y_w = C_eps_2 * f_2 * dis0 * dis0 / tke_s;
            tke_src = P_k - Y_k + L_k / C.re_turb;
            dis_src = P_w - Y_w + L_w / C.re_turb;
        } else if(turb_sel == 1){
            const double sig_k = 0.6, sig_w = 0.5, bet_star = 0.09, gam = 0.52, bet_0 = 0.0708;
            diffusion_tke_re = 1.0 / C.re + sig_k * nue_here;
            diffusion_dis_re = 1.0 / C.re + sig_w * nue_here;
            const double S11 = dudr_s, S22 = dvdthe_s, S33 = dwdphi_s;
            const double S12 = 0.5 * (dudthe_s + dvdr_s);
            const double S13 = 0.5 * (dudphi_s + dwdr_s);
            const double S23 = 0.5 * (dvdphi_s + dwdthe_s);
            const double S_mag = std::sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33
                                                + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
            const double W12 = dudthe_s - dvdr_s;
            const double W13 = dudphi_s - dwdr_s;
            const double W23 = dvdphi_s - dwdthe_s;
            const double Omega = std::sqrt(W12 * W12 + W13 * W13 + W23 * W23);
            const double chi_w = std::fabs(Omega * Omega * S_mag / std::pow(bet_star * dis_s, 3));
            const double f_bet = (1.0 + 85.0 * chi_w) / (1.0 + 100.0 * chi_w);
            const double bet_wc = bet_0 * f_bet;
            const double sig_d = (grad_dot <= 0.0) ? 0.0 : 0.125;
            const double P_k = std::min(P_full, 20.0 * bet_star * tke_s * dis_s);
            const double Y_k = bet_star * tke_s * dis_s;
            const double P_w = gam * dis_s / tke_s * P_k;
            const double Y_w = bet_wc * dis_s * dis_s;
            const double D_w = sig_d / dis_s * grad_dot;
            tke_src = P_k - Y_k;
            dis_src = P_w - Y_w + D_w;
        } else {
            const double bet_star = 0.09, sig_k1 = 1.176, sig_k2 = 1.0, sig_w1 = 2.0, sig_w2 = 1.168;
            const double bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2;
            const double nue_air_nd = nue_gas_phys / (C.u_0 * L_m);
            const double CD_kw = std::max(2.0 * sig_w2 / dis_s * grad_dot, 1.0e-20);
            const double arg1 = std::min(
                std::max(std::sqrt(std::max(tke0, 0.0)) / (bet_star * dis_s * y_star),
                         500.0 * nue_air_nd / (y_star * y_star * dis_s)),
                4.0 * sig_w2 * tke_s / (CD_kw * y_star * y_star));
            const double F1 = std::tanh(std::pow(arg1, 4));
            const double sig_k = F1 * sig_k1 + (1.0 - F1) * sig_k2;
            const double sig_w = F1 * sig_w1 + (1.0 - F1) * sig_w2;
            diffusion_tke_re = 1.0 / C.re + nue_here / sig_k;
            diffusion_dis_re = 1.0 / C.re + nue_here / sig_w;
            const double P_k = std::min(P_full, 20.0 * bet_star * tke_s * dis_s);
            const double Y_k = bet_star * tke_s * dis_s;
            const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_s / tke_s;
            const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_s * dis_s;
            const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_s * grad_dot;
            tke_src = P_k - Y_k;
            dis_src = P_w - Y_w + D_w;
        }
        const double tke_src_max = 20.0 * tke_s * dis_s;
        const double dis_src_max = 20.0 * dis_s * dis_s;
        if(!std::isfinite(tke_src)) tke_src = 0.0;
        if(!std::isfinite(dis_src)) dis_src = 0.0;
        tke_src = std::max(-tke_src_max, std::min(tke_src_max, tke_src));
        dis_src = std::max(-dis_src_max, std::min(dis_src_max, dis_src));
        F[ATJUP_F_TKE_SOURCE][q] = tke_src;
        F[ATJUP_F_DIS_SOURCE][q] = dis_src;
    }

    auto& R = o.rhs;
    R[ATJUP_F_T][q] =
        + pressure_t
        - transport_t
        + diffusion_t * (K.t_diff_scale / (C.re * C.pr) + nue_t_s)
        + radiation_t
        + precip_t;

    const double tq = F[ATJUP_F_T][q];
    const double buoyancy_u = (tq > 0.0)
        ? -nd_buoy * 1.0 * 1.0 * C.buoyancy
          * (C.g * F[ATJUP_F_P_STAT][q] / (C.r_mix * C.R_mix * tq * C.t_ref) - o.brl[i])
        : 0.0;
    R[ATJUP_F_U][q] =
        - dpdr_term
        + buoyancy_u
        - transport_u
        + diffusion_u * (1.0 / C.re + nue_t + nue_wall)
        - cor_rad_on * C.Coriolis * Coriolis_rad
        + C.centrifugal * centrifugal_rad;
    R[ATJUP_F_V][q] =
        - dpdthe_term
        - dphdthe_term
        - transport_v
        + diffusion_v * (1.0 / C.re + nue_t + nue_wall)
        - C.Coriolis * Coriolis_the
        + C.centrifugal * centrifugal_the;
    R[ATJUP_F_W][q] =
        - dpdphi_term
        - dphdphi_term
        - transport_w
        + diffusion_w * (1.0 / C.re + nue_t + nue_wall)
        - C.Coriolis * Coriolis_phi;

    // species (RHS_Jup_Turb.cpp:1273-1338)
    const double sds = K.species_diff_scale, pc = K.precip_coupling, cr = C.chemical_reaction;
    auto sp = [&](int fi, double sc){ return -transport(fi) + diffusion(fi) * (sds / (sc * C.re) + nue_t_s); };
    R[ATJUP_F_H2O][q]       = sp(ATJUP_F_H2O, C.sc_h2o)       + pc * F[ATJUP_F_S_PRECIP_H2O][q];
    R[ATJUP_F_H2O_CLOUD][q] = sp(ATJUP_F_H2O_CLOUD, C.sc_h2o) + pc * F[ATJUP_F_S_PRECIP_H2O_CLOUD][q];
    R[ATJUP_F_H2O_ICE][q]   = sp(ATJUP_F_H2O_ICE, C.sc_h2o)   + pc * F[ATJUP_F_S_PRECIP_H2O_ICE][q];
    R[ATJUP_F_H2S][q]       = sp(ATJUP_F_H2S, C.sc_h2s)       + cr * F[ATJUP_F_MASSFLUX_H2S][q];
    R[ATJUP_F_NH3][q]       = sp(ATJUP_F_NH3, C.sc_nh3)       + cr * F[ATJUP_F_MASSFLUX_NH3][q]
                                                              + pc * F[ATJUP_F_S_PRECIP_NH3][q];
    R[ATJUP_F_NH3_CLOUD][q] = sp(ATJUP_F_NH3_CLOUD, C.sc_nh3) + pc * F[ATJUP_F_S_PRECIP_NH3_CLOUD][q];
    R[ATJUP_F_NH3_ICE][q]   = sp(ATJUP_F_NH3_ICE, C.sc_nh3)   + pc * F[ATJUP_F_S_PRECIP_NH3_ICE][q];
    R[ATJUP_F_CH4][q]       = sp(ATJUP_F_CH4, C.sc_ch4)       + pc * F[ATJUP_F_S_PRECIP_CH4][q];
    R[ATJUP_F_CH4_CLOUD][q] = sp(ATJUP_F_CH4_CLOUD, C.sc_ch4) + pc * F[ATJUP_F_S_PRECIP_CH4_CLOUD][q];
    R[ATJUP_F_CH4_ICE][q]   = sp(ATJUP_F_CH4_ICE, C.sc_ch4)   + pc * F[ATJUP_F_S_PRECIP_CH4_ICE][q];
    {
        const double v_stokes = (2.0 / 9.0) * C.r_p_nh4sh * C.r_p_nh4sh
                              * (C.rg_nh4sh - F[ATJUP_F_RHO_MIX][q]) * C.g / C.mue_mix;
        R[ATJUP_F_NH4SH][q] = -transport(ATJUP_F_NH4SH)
            + F[ATJUP_F_FLUXLIM_NH4SH][q]
            + diffusion(ATJUP_F_NH4SH) * (sds / (C.sc_nh4sh * C.re) + nue_t_s)
            + cr * F[ATJUP_F_MASSFLUX_NH4SH][q]
            + (v_stokes / C.u_0) * d1r[ATJUP_F_NH4SH];
    }
    if(turb_on){
        R[ATJUP_F_TKE][q] = -transport(ATJUP_F_TKE) + diffusion(ATJUP_F_TKE) * diffusion_tke_re + tke_src;
        R[ATJUP_F_DIS][q] = -transport(ATJUP_F_DIS) + diffusion(ATJUP_F_DIS) * diffusion_dis_re + dis_src;
    } else {
        R[ATJUP_F_TKE][q] = 0.0;
        R[ATJUP_F_DIS][q] = 0.0;
    }
    F[ATJUP_F_AUX_U][q] = R[ATJUP_F_U][q] + dpdr_term;
    F[ATJUP_F_AUX_V][q] = R[ATJUP_F_V][q] + dpdthe_term;
    F[ATJUP_F_AUX_W][q] = R[ATJUP_F_W][q] + dpdphi_term;
}

// ---- RungeKuttaJup (RungeKutta_Jup_Turb.cpp:261-583) ----------------------------------------
void rk4(orc_ctx& o){
    buoy_ref_level(o);
    hydrostatic(o);
    const double dt = o.c.dt;
    const double tke_max_nd = o.kn.tke_max_phys / (o.c.u_0 * o.c.u_0);
    const double dis_min_nd = 1.0e-10;
    const int turb_sel = o.kn.turb_model;
    const int im = o.im, jm = o.jm, km = o.km;
    for(int stage = 0; stage < 4; stage++){
        const double c_in = (stage == 0 || stage == 1) ? 0.5 * dt : (stage == 2 ? dt : 0.0);
        const double wgt = (stage == 0 || stage == 3) ? 1.0 : 2.0;
        #pragma omp parallel for collapse(2) schedule(static)
        for(int i = 1; i < im - 1; i++)
            for(int j = 3; j < jm - 3; j++){
                const Geo G = o.geo(i, j);
                for(int k = 0; k < km - 1; k++){
                    if(o.solid(i, j, k)) continue;
                    rhs_cell(o, i, j, k, G, turb_sel);
                }
            }
        #pragma omp parallel for collapse(2) schedule(static)
        for(int i = 1; i < im - 1; i++)
            for(int j = 3; j < jm - 3; j++)
                for(int k = 0; k < km - 1; k++){
                    if(o.solid(i, j, k)) continue;
                    const size_t q = o.id(i, j, k);
                    for(int fi = 0; fi < NP; fi++)
                        o.acc[fi][q] = (stage == 0) ? wgt * o.rhs[fi][q] : o.acc[fi][q] + wgt * o.rhs[fi][q];
                    if(stage < 3){
                        for(int fi = 0; fi < ATJUP_F_TKE; fi++)
                            o.f[fi][q] = o.f[ATJUP_NLEVEL + fi][q] + c_in * o.rhs[fi][q];
                        o.f[ATJUP_F_TKE][q] = safe_clamp(o.f[ATJUP_NLEVEL + ATJUP_F_TKE][q]
                            + c_in * o.rhs[ATJUP_F_TKE][q], 0.0, tke_max_nd);
                        o.f[ATJUP_F_DIS][q] = std::max(dis_min_nd, o.f[ATJUP_NLEVEL + ATJUP_F_DIS][q]
                            + c_in * o.rhs[ATJUP_F_DIS][q]);
                    }
                }
    }
    const double one_sixth = 1.0 / 6.0;
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 1; i < im - 1; i++)
        for(int j = 3; j < jm - 3; j++)
            for(int k = 0; k < km - 1; k++){
                if(o.solid(i, j, k)) continue;
                const size_t q = o.id(i, j, k);
                for(int fi = 0; fi < 4; fi++)
                    o.f[fi][q] = o.f[ATJUP_NLEVEL + fi][q] + dt * o.acc[fi][q] * one_sixth;
                for(int fi = ATJUP_F_H2O; fi < ATJUP_F_NH4SH; fi++)
                    o.f[fi][q] = std::max(0.0, o.f[ATJUP_NLEVEL + fi][q] + dt * o.acc[fi][q] * one_sixth);
                {
                    double nh4sh_new = o.f[ATJUP_NLEVEL + ATJUP_F_NH4SH][q] + dt * o.acc[ATJUP_F_NH4SH][q] * one_sixth;
                    if(nh4sh_new < 0.0) nh4sh_new = 0.0;
                    o.f[ATJUP_F_NH4SH][q] = nh4sh_new;
                }
                o.f[ATJUP_F_TKE][q] = safe_clamp(o.f[ATJUP_NLEVEL + ATJUP_F_TKE][q]
                    + dt * o.acc[ATJUP_F_TKE][q] * one_sixth, 0.0, tke_max_nd);
                o.f[ATJUP_F_DIS][q] = std::max(dis_min_nd, o.f[ATJUP_NLEVEL + ATJUP_F_DIS][q]
                    + dt * o.acc[ATJUP_F_DIS][q] * one_sixth);
            }
}

// ---- PressureSolver<P>::run, PRESS_SRC 1, PRESS_WALL 1, rigid lid, no PRESS_INCR / LINE /
//      BAROTROPIC (PressureSolver.h:145-766), with prepareProjectionBoundaries
//      (cJupiterModel.cpp:1842-1864). omega != 1 blends the Gauss-Seidel value (SOR). ---------
// prepareProjectionBoundaries(rigid_lid = true) (cJupiterModel.cpp:1842-1864); idempotent
void prep_projection(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    auto& au = o.f[ATJUP_F_AUX_U];
    auto& av = o.f[ATJUP_F_AUX_V];
    auto& aw = o.f[ATJUP_F_AUX_W];
    const double c43 = o.c43, c13 = o.c13;
    #pragma omp parallel for collapse(2)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++){
            au[o.id(0, j, k)] = 0.0;
            au[o.id(im - 1, j, k)] = 0.0;
            av[o.id(0, j, k)] = c43 * av[o.id(1, j, k)] - c13 * av[o.id(2, j, k)];
            av[o.id(im - 1, j, k)] = c43 * av[o.id(im - 2, j, k)] - c13 * av[o.id(im - 3, j, k)];
            aw[o.id(0, j, k)] = c43 * aw[o.id(1, j, k)] - c13 * aw[o.id(2, j, k)];
            aw[o.id(im - 1, j, k)] = c43 * aw[o.id(im - 2, j, k)] - c13 * aw[o.id(im - 3, j, k)];
        }
}

// One colour of a red-black sweep (PressureSolver.h:323-495): every interior cell with
// (i + j + k) & 1 == colour takes its Gauss-Seidel value from the current neighbours, blended with
// its old value when omega != 1. The outer planes are not touched.
void rb_colour_pass(orc_ctx& o, int colour, double omega){
    const int im = o.im, jm = o.jm, km = o.km;
    auto& au = o.f[ATJUP_F_AUX_U];
    auto& av = o.f[ATJUP_F_AUX_V];
    auto& aw = o.f[ATJUP_F_AUX_W];
    auto& p = o.f[ATJUP_F_P_DYN];
    std::vector<double> sinth(jm);
    for(int j = 0; j < jm; j++){
        sinth[j] = o.sin_the[j];
        if(sinth[j] < o.kn.sinthe_min) sinth[j] = o.kn.sinthe_min;
        if(sinth[j] == 0.0) sinth[j] = 1.0e-5;
    }
    const double inv_2dr = 1.0 / (2.0 * o.g.dr), inv_2dthe = 1.0 / (2.0 * o.g.dthe), inv_2dphi = 1.0 / (2.0 * o.g.dphi);
    const double inv_dr2 = 1.0 / (o.g.dr * o.g.dr), inv_dthe2 = 1.0 / (o.g.dthe * o.g.dthe), inv_dphi2 = 1.0 / (o.g.dphi * o.g.dphi);
    const double inv_dthe = 1.0 / o.g.dthe, inv_dphi = 1.0 / o.g.dphi;
    const double cap = o.kn.pdyn_cap;
    auto S = [&](const std::vector<double>& A, int i, int j, int k){ return A[o.id(i, j, k)]; };
    {
        {
            #pragma omp parallel for collapse(2) schedule(static)
            for(int i = 1; i < im - 1; i++)
                for(int j = 1; j < jm - 1; j++){
                    const double rm = o.rad[i];
                    const double rm2 = rm * rm;
                    const double exp_rm = 1.0;
                    const double exp_2_rm = exp_rm * exp_rm;
                    const double sinthe = sinth[j];
                    const double sinthe2 = sinthe * sinthe;
                    const double inv_rm = 1.0 / rm;
                    const double inv_rm2 = 1.0 / rm2;
                    const double inv_rmsinthe = 1.0 / (rm * sinthe);
                    const double inv_rm2sinthe2 = inv_rm2 / sinthe2;
                    const double num1 = exp_2_rm * inv_dr2;
                    const double num2 = inv_rm2 * inv_dthe2;
                    const double num3 = inv_rm2sinthe2 * inv_dphi2;
                    const double denom = 2.0 * num1 + 2.0 * num2 + 2.0 * num3;
                    const bool i_in_range = (i < im - 2);
                    const bool j_inner = (j > 2) && (j < jm - 2);
                    for(int k = 0; k < km - 1; k++){
                        if(((i + j + k) & 1) != colour) continue;
                        const int kw = o.wrap(k - 1), ke = o.wrap(k + 1);
                        const bool sld = o.solid(i, j, k);
                        const bool skp = o.solid(i, j, ke), skm = o.solid(i, j, kw);
                        const bool skp2 = (k < km - 2) ? o.solid(i, j, k + 2) : false;
                        const bool skm2 = (k > 2) ? o.solid(i, j, k - 2) : false;
                        double du_dr = 0, dv_dthe = 0, dw_dphi = 0;
                        bool r_flag = false, the_flag = false, phi_flag = false;
                        if(i_in_range && sld && !o.solid(i + 1, j, k)){
                            du_dr = (-3.0 * S(au, i, j, k) + 4.0 * S(au, i + 1, j, k) - S(au, i + 2, j, k)) * inv_2dr;
                            r_flag = true;
                        }
                        if(j_inner){
                            const bool air_jp1 = !o.solid(i, j + 1, k), air_jm1 = !o.solid(i, j - 1, k);
                            if(sld && air_jp1 && !o.solid(i, j + 2, k)){
                                dv_dthe = (-3.0 * S(av, i, j, k) + 4.0 * S(av, i, j + 1, k) - S(av, i, j + 2, k)) * inv_2dthe;
                                the_flag = true;
                            }
                            if(sld && air_jm1 && !o.solid(i, j - 2, k)){
                                dv_dthe = -(-3.0 * S(av, i, j, k) + 4.0 * S(av, i, j - 1, k) - S(av, i, j - 2, k)) * inv_2dthe;
                                the_flag = true;
                            }
                            if((sld && air_jp1 && o.solid(i, j + 2, k)) || (j == jm - 2 && !sld && o.solid(i, j + 1, k))){
                                dv_dthe = (S(av, i, j + 1, k) - S(av, i, j, k)) * inv_dthe;
                                the_flag = true;
                            }
                            if((sld && air_jm1 && o.solid(i, j - 2, k)) || (j == 1 && sld && air_jm1)){
                                dv_dthe = (S(av, i, j - 1, k) - S(av, i, j, k)) * inv_dthe;
                                the_flag = true;
                            }
                        }
                        if((k > 2) && (k < km - 2)){
                            const bool air_kp1 = !skp, air_km1 = !skm;
                            if(sld && air_kp1 && !skp2){
                                dw_dphi = (-3.0 * S(aw, i, j, k) + 4.0 * S(aw, i, j, k + 1) - S(aw, i, j, k + 2)) * inv_2dphi;
                                phi_flag = true;
                            }
                            if(sld && air_km1 && !skm2){
                                dw_dphi = -(-3.0 * S(aw, i, j, k) + 4.0 * S(aw, i, j, k - 1) - S(aw, i, j, k - 2)) * inv_2dphi;
                                phi_flag = true;
                            }
                            if((sld && air_kp1 && skp2) || (k == km - 2 && !sld && skp)){
                                dw_dphi = (S(aw, i, j, k + 1) - S(aw, i, j, k)) * inv_dphi;
                                phi_flag = true;
                            }
                            if((sld && air_km1 && skm2) || (k == 1 && sld && air_km1)){
                                dw_dphi = (S(aw, i, j, k - 1) - S(aw, i, j, k)) * inv_dphi;
                                phi_flag = true;
                            }
                        }
                        if(!r_flag) du_dr = (S(au, i + 1, j, k) - S(au, i - 1, j, k)) * inv_2dr;
                        if(!the_flag) dv_dthe = (S(av, i, j + 1, k) - S(av, i, j - 1, k)) * inv_2dthe;
                        if(!phi_flag) dw_dphi = (S(aw, i, j, ke) - S(aw, i, j, kw)) * inv_2dphi;
                        double div_src = du_dr * exp_rm + dv_dthe * inv_rm + dw_dphi * inv_rmsinthe;
                        double acc = 0.0, den = denom;
                        if(!sld){
                            acc = den = 0.0;
                            if(!o.solid(i + 1, j, k)){ acc += num1 * p[o.id(i + 1, j, k)]; den += num1; }
                            if(!o.solid(i - 1, j, k)){ acc += num1 * p[o.id(i - 1, j, k)]; den += num1; }
                            if(!o.solid(i, j + 1, k)){ acc += num2 * p[o.id(i, j + 1, k)]; den += num2; }
                            if(!o.solid(i, j - 1, k)){ acc += num2 * p[o.id(i, j - 1, k)]; den += num2; }
                            if(!skp){ acc += num3 * p[o.id(i, j, ke)]; den += num3; }
                            if(!skm){ acc += num3 * p[o.id(i, j, kw)]; den += num3; }
                            if(den <= 0.0) continue;
                        } else {
                            acc = (p[o.id(i + 1, j, k)] + p[o.id(i - 1, j, k)]) * num1
                                + (p[o.id(i, j + 1, k)] + p[o.id(i, j - 1, k)]) * num2
                                + (p[o.id(i, j, ke)] + p[o.id(i, j, kw)]) * num3;
                        }
                        const double src_max = den * cap;
                        if(!std::isfinite(div_src)) div_src = 0.0;
                        else if(div_src > src_max) div_src = src_max;
                        else if(div_src < -src_max) div_src = -src_max;
                        const double p_gs = (acc - div_src) / den;
                        const size_t q = o.id(i, j, k);
                        p[q] = (omega == 1.0) ? p_gs : (1.0 - omega) * p[q] + omega * p_gs;
                    }
                }
        }
    }
}

// the outer 2-point Neumann planes i = 0, im-1 and j = 0, jm-1 and the seam copy, after every
// sweep (PressureSolver.h:718-746)
void rb_planes(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    auto& p = o.f[ATJUP_F_P_DYN];
    const double c43 = o.c43, c13 = o.c13;
    {
        #pragma omp parallel for collapse(2)
        for(int k = 0; k < km; k++)
            for(int j = 0; j < jm; j++){
                p[o.id(0, j, k)] = c43 * p[o.id(1, j, k)] - c13 * p[o.id(2, j, k)];
                p[o.id(im - 1, j, k)] = c43 * p[o.id(im - 2, j, k)] - c13 * p[o.id(im - 3, j, k)];
            }
        #pragma omp parallel for collapse(2)
        for(int k = 0; k < km; k++)
            for(int i = 0; i < im; i++){
                p[o.id(i, 0, k)] = c43 * p[o.id(i, 1, k)] - c13 * p[o.id(i, 2, k)];
                p[o.id(i, jm - 1, k)] = c43 * p[o.id(i, jm - 2, k)] - c13 * p[o.id(i, jm - 3, k)];
            }
        #pragma omp parallel for collapse(2)
        for(int i = 0; i < im; i++)
            for(int j = 0; j < jm; j++)
                p[o.id(i, j, km - 1)] = p[o.id(i, j, 0)];
    }
}

void pressure(orc_ctx& o, int n_sweeps, double omega){
    prep_projection(o);
    for(int sweep = 0; sweep < n_sweeps; sweep++){
        rb_colour_pass(o, 0, omega);
        rb_colour_pass(o, 1, omega);
        rb_planes(o);
    }
}

// ---- damp_wiggles / damp_wiggles_ho with i_topography, periodic seam (lib/Utils.cpp:221-453) ---
void shapiro(orc_ctx& o, std::vector<double>& fld, int order, double strength, int passes, int axes){
    const int im = o.im, jm = o.jm, km = o.km;
    const double coeff = strength * 0.25;
    const double c4 = strength / 16.0;
    auto isurf = [&](int j, int k){ return std::max(o.itopo[(size_t)j * km + k], 0); };
    auto in_fluid = [&](int i, int j, int k){ return i >= isurf(j, k); };
    std::vector<double> tmp(o.n3);
    for(int pass = 0; pass < passes; ++pass){
        if(axes & 4){
            tmp = fld;
            const int n_k = km - 1;
            #pragma omp parallel for collapse(2) schedule(static)
            for(int i = 0; i < im; ++i)
                for(int j = 0; j < jm; ++j){
                    for(int k = 0; k < n_k; ++k){
                        if(!in_fluid(i, j, k)) continue;
                        const size_t q = o.id(i, j, k);
                        const int km1 = (k == 0) ? km - 2 : k - 1;
                        const int kp1 = (k == km - 2) ? 0 : k + 1;
                        if(order == 2){
                            const double a = in_fluid(i, j, km1) ? tmp[o.id(i, j, km1)] : tmp[q];
                            const double b = in_fluid(i, j, kp1) ? tmp[o.id(i, j, kp1)] : tmp[q];
                            fld[q] = tmp[q] + coeff * (a - 2.0 * tmp[q] + b);
                        } else {
                            const int km2 = (k - 2 + n_k) % n_k, kp2 = (k + 2) % n_k;
                            const double c = tmp[q];
                            if(!in_fluid(i, j, km1) || !in_fluid(i, j, kp1)) continue;
                            if(in_fluid(i, j, km2) && in_fluid(i, j, kp2)){
                                const double d4 = tmp[o.id(i, j, kp2)] - 4.0 * tmp[o.id(i, j, kp1)]
                                                + 6.0 * c - 4.0 * tmp[o.id(i, j, km1)] + tmp[o.id(i, j, km2)];
                                fld[q] = c - c4 * d4;
                            } else {
                                fld[q] = c + coeff * (tmp[o.id(i, j, km1)] - 2.0 * c + tmp[o.id(i, j, kp1)]);
                            }
                        }
                    }
                    fld[o.id(i, j, km - 1)] = fld[o.id(i, j, 0)];
                }
        }
        if(axes & 2){
            tmp = fld;
            #pragma omp parallel for collapse(2) schedule(static)
            for(int i = 0; i < im; ++i)
                for(int j = 0; j < jm; ++j){
                    for(int k = 0; k < km; ++k){
                        if(!in_fluid(i, j, k)) continue;
                        const size_t q = o.id(i, j, k);
                        if(order == 2){
                            const int jm1 = std::max(j - 1, 0), jp1 = std::min(j + 1, jm - 1);
                            const double a = in_fluid(i, jm1, k) ? tmp[o.id(i, jm1, k)] : tmp[q];
                            const double b = in_fluid(i, jp1, k) ? tmp[o.id(i, jp1, k)] : tmp[q];
                            fld[q] = tmp[q] + coeff * (a - 2.0 * tmp[q] + b);
                        } else {
                            const double c = tmp[q];
                            const bool fm1 = (j - 1 >= 0) && in_fluid(i, j - 1, k);
                            const bool fp1 = (j + 1 <= jm - 1) && in_fluid(i, j + 1, k);
                            if(!fm1 || !fp1) continue;
                            const bool fm2 = (j - 2 >= 0) && in_fluid(i, j - 2, k);
                            const bool fp2 = (j + 2 <= jm - 1) && in_fluid(i, j + 2, k);
                            if(fm2 && fp2){
                                const double d4 = tmp[o.id(i, j + 2, k)] - 4.0 * tmp[o.id(i, j + 1, k)]
                                                + 6.0 * c - 4.0 * tmp[o.id(i, j - 1, k)] + tmp[o.id(i, j - 2, k)];
                                fld[q] = c - c4 * d4;
                            } else {
                                fld[q] = c + coeff * (tmp[o.id(i, j - 1, k)] - 2.0 * c + tmp[o.id(i, j + 1, k)]);
                            }
                        }
                    }
                }
        }
        if(axes & 1){
            tmp = fld;
            #pragma omp parallel for collapse(2) schedule(static)
            for(int j = 0; j < jm; ++j)
                for(int k = 0; k < km; ++k){
                    const int i_surf = isurf(j, k);
                    for(int i = i_surf; i < im; ++i){
                        const size_t q = o.id(i, j, k);
                        if(order == 2){
                            const int im1 = std::max(i - 1, i_surf), ip1 = std::min(i + 1, im - 1);
                            fld[q] = tmp[q] + coeff * (tmp[o.id(im1, j, k)] - 2.0 * tmp[q] + tmp[o.id(ip1, j, k)]);
                        } else {
                            const double c = tmp[q];
                            const bool fm1 = (i - 1 >= i_surf), fp1 = (i + 1 <= im - 1);
                            if(!fm1 || !fp1) continue;
                            const bool fm2 = (i - 2 >= i_surf), fp2 = (i + 2 <= im - 1);
                            if(fm2 && fp2){
                                const double d4 = tmp[o.id(i + 2, j, k)] - 4.0 * tmp[o.id(i + 1, j, k)]
                                                + 6.0 * c - 4.0 * tmp[o.id(i - 1, j, k)] + tmp[o.id(i - 2, j, k)];
                                fld[q] = c - c4 * d4;
                            } else {
                                fld[q] = c + coeff * (tmp[o.id(i - 1, j, k)] - 2.0 * c + tmp[o.id(i + 1, j, k)]);
                            }
                        }
                    }
                }
        }
    }
}

// ---- computeMixtureDensity (InitValues_Jup.cpp:717-740) --------------------------------------
void mixture_density(orc_ctx& o){
    const auto& t = o.f[ATJUP_F_T];
    const auto& ps = o.f[ATJUP_F_P_STAT];
    const auto& pd = o.f[ATJUP_F_P_DYN];
    auto& rho = o.f[ATJUP_F_RHO_MIX];
    const double pdyn_w = o.kn.rho_pdyn;
    const double ptb = o.p_dyn_to_bar();
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++){
        const double T = t[q] * o.c.t_ref;
        if(!(T > 0.0)){ rho[q] = 0.0; continue; }
        const double p_pa = (ps[q] + pdyn_w * pd[q] * ptb) * 1.0e5;
        const double r = p_pa / (o.c.R_mix * T);
        rho[q] = std::isfinite(r) ? r : 0.0;
    }
}

const int kRadius[] = { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE,
    ATJUP_F_H2S, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_NH4SH,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH,
    ATJUP_F_TKE, ATJUP_F_DIS, ATJUP_F_NUE, ATJUP_F_PROD, ATJUP_F_TKE_SOURCE, ATJUP_F_DIS_SOURCE };
const int kThetaExtrap[] = { ATJUP_F_T, ATJUP_F_U,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE,
    ATJUP_F_H2S, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_NH4SH,
    ATJUP_F_TKE, ATJUP_F_DIS, ATJUP_F_NUE, ATJUP_F_PROD, ATJUP_F_TKE_SOURCE, ATJUP_F_DIS_SOURCE };
const int kThetaZero[] = { ATJUP_F_V, ATJUP_F_W,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH };
const int kPhi[] = { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE,
    ATJUP_F_H2S, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_NH4SH,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH,
    ATJUP_F_TKE, ATJUP_F_DIS, ATJUP_F_NUE, ATJUP_F_PROD, ATJUP_F_TKE_SOURCE, ATJUP_F_DIS_SOURCE };
const int kSpecies[] = { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, ATJUP_F_H2S,
    ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_NH4SH };
const int kSolid[] = { ATJUP_F_T, ATJUP_F_P_STAT, ATJUP_F_P_DYN,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE,
    ATJUP_F_H2S, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_NH4SH,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH };

// The fields each boundary pass treats, as the reference's lists name them (BC_Jup.h:272-350,
// :469-487): the field ids above, and the chemistry's carried fields, which DiffMassFluxJup and
// ChemMassRateJup write on the interior only, so that the values these passes leave on the
// boundary rows reach the next call's smoothing and thermal flux: the reaction rates w_* (held
// beside the fields), the diffusive fluxes and thermalmassflux. (j_*, jT_* and the force
// diagnostics are on the lists too but are never read across calls.)
struct BCSets { std::vector<std::vector<double>*> radius, theta_ex, theta_zero, phi, solid; };
BCSets bc_sets(orc_ctx& o){
    BCSets s;
    for(int f : kRadius) s.radius.push_back(&o.f[f]);
    for(int f : kThetaExtrap) s.theta_ex.push_back(&o.f[f]);
    for(int f : kThetaZero) s.theta_zero.push_back(&o.f[f]);
    for(int f : kPhi) s.phi.push_back(&o.f[f]);
    for(int f : kSolid) s.solid.push_back(&o.f[f]);
    for(std::vector<double>* w : { &o.w_h2s, &o.w_nh3, &o.w_nh4sh }){
        s.radius.push_back(w); s.theta_ex.push_back(w); s.phi.push_back(w); s.solid.push_back(w);
    }
    for(int f : { ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3, ATJUP_F_DIFFLUX_NH4SH }){
        s.radius.push_back(&o.f[f]); s.theta_zero.push_back(&o.f[f]); s.phi.push_back(&o.f[f]);
        s.solid.push_back(&o.f[f]);
    }
    std::vector<double>* th = &o.f[ATJUP_F_THERMALMASSFLUX];
    s.radius.push_back(th); s.theta_ex.push_back(th); s.phi.push_back(th); s.solid.push_back(th);
    return s;
}

// ---- BC_Jup::bcRadius: shared bcRadius (BoundaryConditions.h:151-286) with TOP_TAPER 3
//      (BC_Jup.h:185-263) -------------------------------------------------------------------------
void bc_radius(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km, kn = km - 1;
    auto& V = o.f[ATJUP_F_V];
    auto& W = o.f[ATJUP_F_W];
    std::vector<double> zm((size_t)4 * jm, 0.0);
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int l = 0; l < 2; l++){
            const int i = im - 3 + l;
            double sv = 0.0, sw = 0.0;
            for(int k = 0; k < kn; k++){ sv += V[o.id(i, j, k)]; sw += W[o.id(i, j, k)]; }
            zm[(size_t)(2 * l) * jm + j] = sv / kn;
            zm[(size_t)(2 * l + 1) * jm + j] = sw / kn;
        }
    const double c43 = o.c43, c13 = o.c13;
    const BCSets sets = bc_sets(o);
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++)
            for(std::vector<double>* Xp : sets.radius){
                auto& X = *Xp;
                X[o.id(0, j, k)] = c43 * X[o.id(1, j, k)] - c13 * X[o.id(2, j, k)];
                X[o.id(im - 1, j, k)] = c43 * X[o.id(im - 2, j, k)] - c13 * X[o.id(im - 3, j, k)];
            }
    auto& U = o.f[ATJUP_F_U];
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){ U[o.id(0, j, k)] = 0.0; U[o.id(im - 1, j, k)] = 0.0; }
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            V[o.id(im - 1, j, k)] = 0.0;
            W[o.id(im - 1, j, k)] = 0.0;
            V[o.id(im - 2, j, k)] *= (1.0 / 3.0);
            W[o.id(im - 2, j, k)] *= (1.0 / 3.0);
            V[o.id(im - 3, j, k)] *= (2.0 / 3.0);
            W[o.id(im - 3, j, k)] *= (2.0 / 3.0);
        }
    #pragma omp parallel for
    for(int j = 0; j < jm; j++){
        const double v3 = zm[j], w3 = zm[(size_t)jm + j], v2 = zm[(size_t)2 * jm + j], w2 = zm[(size_t)3 * jm + j];
        for(int k = 0; k < km; k++){
            V[o.id(im - 3, j, k)] += (1.0 / 3.0) * v3;
            W[o.id(im - 3, j, k)] += (1.0 / 3.0) * w3;
            V[o.id(im - 2, j, k)] += (2.0 / 3.0) * v2;
            W[o.id(im - 2, j, k)] += (2.0 / 3.0) * w2;
            V[o.id(im - 1, j, k)] = v2;
            W[o.id(im - 1, j, k)] = w2;
        }
    }
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++)
            for(int fid : kSpecies){
                auto& X = o.f[fid];
                X[o.id(0, j, k)] = X[o.id(1, j, k)];
            }
    #pragma omp parallel for
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++)
            for(int fid : kSpecies){
                auto& X = o.f[fid];
                if(X[o.id(0, j, k)] < 0.0) X[o.id(0, j, k)] = 0.0;
                if(X[o.id(im - 1, j, k)] < 0.0) X[o.id(im - 1, j, k)] = 0.0;
            }
}

// ---- bcTheta with POLE_FILL (2 open rows) and POLE_COPY (BoundaryConditions.h:292-391) --------
void bc_theta(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    const int nfill = 2, jf = nfill + 1;
    const BCSets sets = bc_sets(o);
    #pragma omp parallel for
    for(int k = 0; k < km; k++)
        for(int i = 0; i < im; i++){
            for(int r = 1; r <= nfill; r++){
                const double ramp = (double)r / (double)jf;
                for(std::vector<double>* Xp : sets.theta_ex){
                    auto& X = *Xp;
                    X[o.id(i, r, k)] = X[o.id(i, jf, k)];
                    X[o.id(i, jm - 1 - r, k)] = X[o.id(i, jm - 1 - jf, k)];
                }
                for(std::vector<double>* Xp : sets.theta_zero){
                    auto& X = *Xp;
                    X[o.id(i, r, k)] = ramp * X[o.id(i, jf, k)];
                    X[o.id(i, jm - 1 - r, k)] = ramp * X[o.id(i, jm - 1 - jf, k)];
                }
            }
            for(std::vector<double>* Xp : sets.theta_zero){
                auto& X = *Xp;
                X[o.id(i, 0, k)] = 0.0;
                X[o.id(i, jm - 1, k)] = 0.0;
            }
            for(std::vector<double>* Xp : sets.theta_ex){
                auto& X = *Xp;
                X[o.id(i, 0, k)] = X[o.id(i, 1, k)];
                X[o.id(i, jm - 1, k)] = X[o.id(i, jm - 2, k)];
            }
        }
}

// ---- bcPhi, periodic seam: k = km-1 is the copy of k = 0 (BoundaryConditions.h:398-439) --------
void bc_phi(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    const BCSets sets = bc_sets(o);
    #pragma omp parallel for
    for(int i = 0; i < im; i++)
        for(int j = 0; j < jm; j++)
            for(std::vector<double>* Xp : sets.phi){
                auto& X = *Xp;
                X[o.id(i, j, km - 1)] = X[o.id(i, j, 0)];
            }
}

// ---- BC_Jup::bcSolidGround with MOUNT_T_AMBIENT (BC_Jup.h:460-658) --------------------------
void bc_solid(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    const BCSets sets = bc_sets(o);
    const int ns = (int)sets.solid.size();
    const int n_state = 2;
    const double c43 = o.c43, c13 = o.c13;
    auto S = [&](int ii, int jj, int kk){
        if(ii < 0 || ii >= im || jj < 0 || jj >= jm || kk < 0 || kk >= km) return true;
        return o.solid(ii, jj, kk);
    };
    auto Fl = [&](int ii, int jj, int kk){
        if(ii < 0 || ii >= im || jj < 0 || jj >= jm || kk < 0 || kk >= km) return false;
        return !o.solid(ii, jj, kk);
    };
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++)
            for(int i = 0; i < im; i++){
                if(!o.solid(i, j, k)) continue;
                const size_t q = o.id(i, j, k);
                o.f[ATJUP_F_U][q] = 0.0;
                o.f[ATJUP_F_V][q] = 0.0;
                o.f[ATJUP_F_W][q] = 0.0;
                auto extrap = [&](int ia, int ja, int ka, int ib, int jb, int kb){
                    for(int f = 0; f < ns; f++){
                        auto& X = *sets.solid[f];
                        X[q] = c43 * X[o.id(ia, ja, ka)] - c13 * X[o.id(ib, jb, kb)];
                    }
                };
                const bool surf = !S(i + 1, j, k) || !S(i - 1, j, k) || !S(i, j + 1, k)
                               || !S(i, j - 1, k) || !S(i, j, k + 1) || !S(i, j, k - 1);
                if(!surf){
                    for(int f = n_state; f < ns; f++) (*sets.solid[f])[q] = 0.0;
                    continue;
                }
                if(Fl(i + 1, j, k) && Fl(i + 2, j, k)) extrap(i + 1, j, k, i + 2, j, k);
                else if(Fl(i - 1, j, k) && Fl(i - 2, j, k)) extrap(i - 1, j, k, i - 2, j, k);
                else if(Fl(i, j + 1, k) && Fl(i, j + 2, k)) extrap(i, j + 1, k, i, j + 2, k);
                else if(Fl(i, j - 1, k) && Fl(i, j - 2, k)) extrap(i, j - 1, k, i, j - 2, k);
                else if(Fl(i, j, k + 1) && Fl(i, j, k + 2)) extrap(i, j, k + 1, i, j, k + 2);
                else if(Fl(i, j, k - 1) && Fl(i, j, k - 2)) extrap(i, j, k - 1, i, j, k - 2);
                else for(int f = n_state; f < ns; f++) (*sets.solid[f])[q] = 0.0;
            }
    auto& T = o.f[ATJUP_F_T];
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 0; i < im; i++)
        for(int j = 0; j < jm; j++){
            double s_ = 0.0; int n_ = 0; bool any_ = false;
            for(int k = 0; k < km; k++){
                if(o.solid(i, j, k)){ any_ = true; continue; }
                s_ += T[o.id(i, j, k)]; n_++;
            }
            if(!any_ || n_ == 0) continue;
            const double t_amb = s_ / n_;
            for(int k = 0; k < km; k++) if(o.solid(i, j, k)) T[o.id(i, j, k)] = t_amb;
        }
    const double tke_max_nd = 1000.0 / (o.c.u_0 * o.c.u_0);
    const double nue_max = 1.0e5 / (o.c.u_0 * o.g.L_atm * 1.0e3);
    const double prod_max = 1.0e4, dis_max = 1.0e6;
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++){
        o.f[ATJUP_F_TKE][q] = safe_clamp(o.f[ATJUP_F_TKE][q], 0.0, tke_max_nd);
        o.f[ATJUP_NLEVEL + ATJUP_F_TKE][q] = safe_clamp(o.f[ATJUP_NLEVEL + ATJUP_F_TKE][q], 0.0, tke_max_nd);
        o.f[ATJUP_F_DIS][q] = safe_clamp(o.f[ATJUP_F_DIS][q], 1.0e-10, dis_max);
        o.f[ATJUP_NLEVEL + ATJUP_F_DIS][q] = safe_clamp(o.f[ATJUP_NLEVEL + ATJUP_F_DIS][q], 1.0e-10, dis_max);
        o.f[ATJUP_F_NUE][q] = safe_clamp(o.f[ATJUP_F_NUE][q], 0.0, nue_max);
        o.f[ATJUP_F_PROD][q] = safe_clamp(o.f[ATJUP_F_PROD][q], 0.0, prod_max);
        o.f[ATJUP_F_TKE_SOURCE][q] = safe_clamp(o.f[ATJUP_F_TKE_SOURCE][q], -prod_max, prod_max);
        o.f[ATJUP_F_DIS_SOURCE][q] = safe_clamp(o.f[ATJUP_F_DIS_SOURCE][q], -prod_max, prod_max);
    }
}

// ---- clampNegativeSpecies (FileIO_Jup.cpp:494-548) -------------------------------------------
void clamp_species(orc_ctx& o){
    for(int fid : kSpecies){
        auto& X = o.f[fid];
        #pragma omp parallel for schedule(static)
        for(size_t q = 0; q < o.n3; q++) if(!(X[q] >= 0.0)) X[q] = 0.0;
    }
}

// ---- restoreVar(1.0) (cJupiterModel.cpp:1757-1801) --------------------------------------------
void restore(orc_ctx& o){
    for(int fi = 0; fi < NP; fi++){
        auto& N = o.f[ATJUP_NLEVEL + fi];
        const auto& X = o.f[fi];
        #pragma omp parallel for schedule(static)
        for(size_t q = 0; q < o.n3; q++) N[q] = 1.0 * X[q];
    }
}

// ---- ATJUP_RK_SCALAR_SYNC 2 (cJupiterModel.cpp:1118-1140) -------------------------------------
void scalar_sync(orc_ctx& o){
    const int ids[] = { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD,
                        ATJUP_F_NH3_ICE, ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_T };
    const int n = (o.kn.rk_scalar_sync >= 2) ? 10 : 9;
    for(int a = 0; a < n; a++){
        auto& N = o.f[ATJUP_NLEVEL + ids[a]];
        const auto& X = o.f[ids[a]];
        #pragma omp parallel for schedule(static)
        for(size_t q = 0; q < o.n3; q++) N[q] = X[q];
    }
    // H2S and NH4SH join the sync when the equilibrium adjustment writes them (cJupiterModel.cpp:1120)
    if(o.kn.nh4sh_eq)
        for(int f : { ATJUP_F_H2S, ATJUP_F_NH4SH }){
            auto& N = o.f[ATJUP_NLEVEL + f];
            const auto& X = o.f[f];
            #pragma omp parallel for schedule(static)
            for(size_t q = 0; q < o.n3; q++) N[q] = X[q];
        }
}

void wall_viscosity(orc_ctx& o){                          // InitValues_Jup.cpp:84-134
    auto& wn = o.f[ATJUP_F_WALL_NUE];
    std::fill(wn.begin(), wn.end(), 0.0);
    const double factor = o.kn.wall_nue_factor;
    const int layers = o.kn.wall_nue_layers;
    if(factor <= 0.0 || layers <= 0) return;
    const int im = o.im, jm = o.jm, km = o.km;
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++)
            for(int i = 0; i < im; i++){
                if(o.solid(i, j, k)) continue;
                int d = layers + 1;
                for(int di = -layers; di <= layers && d > 1; di++){
                    const int ii = i + di;
                    if(ii < 0 || ii >= im) continue;
                    for(int dj = -layers; dj <= layers && d > 1; dj++){
                        const int jj = j + dj;
                        if(jj < 0 || jj >= jm) continue;
                        for(int dk = -layers; dk <= layers; dk++){
                            const int kk = k + dk;
                            if(kk < 0 || kk >= km) continue;
                            if(!o.solid(ii, jj, kk)) continue;
                            const int ad = std::max(std::abs(di), std::max(std::abs(dj), std::abs(dk)));
                            if(ad < d) d = ad;
                            if(d <= 1) break;
                        }
                    }
                }
                if(d > layers) continue;
                const double ramp = (double)(layers - d + 1) / (double)layers;
                wn[o.id(i, j, k)] = factor / o.c.re * ramp;
            }
}

// ---- SaturationAdjustment<P>::run for one gas (SaturationAdjustment.h:255-746), ATJUP defaults:
//      CONSERVE 1, NEWTON 1, EVAP 1, BERGERON 2, DRY_EXIT 1, ICE_ENTRY 0, FREEZE 0, LOCAL_RHO 1 ------
inline double svp(double T, double C, double L0, double R, double da, double db){   // ATPhys.h:49-52
    return std::exp(C + (-L0 / T + da * std::log(T) + db * T) / (1e-3 * R));
}

void satadj(orc_ctx& o, int gas){
    static const int ids[3][3] = { { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE },
                                   { ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE },
                                   { ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE } };
    const atjup_gas_params P = o.gas[gas];
    auto& c = o.f[ids[gas][0]];
    auto& cloud = o.f[ids[gas][1]];
    auto& ice = o.f[ids[gas][2]];
    auto& t = o.f[ATJUP_F_T];
    const auto& ps = o.f[ATJUP_F_P_STAT];
    const auto& rho = o.f[ATJUP_F_RHO_MIX];
    const double t_0 = P.t_0, t_00 = P.t_00, ep = P.ep, lv = P.lv, ls = P.ls;
    const double t_range_inv = 1.0 / (t_0 - t_00);
    const double Lf = ls - lv;
    const double cp_mix = o.c.cp_mix, t_ref = o.c.t_ref, r_mix = o.c.r_mix;
    const int iter_prec_end = 15;
    const double q_diff_min = 1.0e-3;
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++){
        double t_u = t[q] * t_ref;
        const double p_u = ps[q];
        if(!(t_u > 0.0) || !(p_u > 0.0) || o.mask[q] == 1.0) continue;
        const double rf = rho[q];
        const double rho_at = (rf > 0.0 && std::isfinite(rf)) ? rf : r_mix;       // cJupiterModel.h:1212-1216
        if(t_u > t_0){
            if(ice[q] > 0.0){
                double dm = ice[q];
                if(Lf > 0.0){
                    const double hc = cp_mix * rho_at;
                    const double cap = hc * (t_u - t_0) / Lf;
                    if(dm > cap) dm = cap;
                    t_u -= Lf * dm / hc;
                    t[q] = t_u / t_ref;
                }
                ice[q] -= dm;
                cloud[q] += dm;
            }
        }
        if(c[q] < 0.0) c[q] = 0.0;
        if(cloud[q] < 0.0) cloud[q] = 0.0;
        if(ice[q] < 0.0) ice[q] = 0.0;
        const double rho_c = rho_at;
        const double E_Rain_0 = svp(t_u, P.C, P.L0, P.R, P.del_alf, P.del_bet);
        const double q_Rain_0 = rho_c * ep * E_Rain_0 / p_u;
        if(q_Rain_0 <= 0.0) continue;
        double q_entry = q_Rain_0;
        if(cloud[q] <= 0.0 && ice[q] <= 0.0 && t_u < t_0)
            q_entry = rho_c * ep * svp(t_u, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
        if(c[q] <= q_entry){
            const double q_c0 = cloud[q], q_i0 = ice[q];
            if(q_c0 <= 0.0 && q_i0 <= 0.0) continue;
            const double q_Ice_0 = rho_c * ep * svp(t_u, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
            double q_t0;
            if(q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
            else if(q_i0 <= 0.0)         q_t0 = q_Rain_0;
            else                         q_t0 = q_Ice_0;
            if(std::fabs(c[q] - q_t0) <= q_diff_min * q_t0) continue;
        }
        double q_v_b = c[q], q_c_b = cloud[q], q_i_b = ice[q];
        double T = t_u;
        double q_v_hyp = q_v_b;
        double q_rain_now = q_Rain_0;
        for(int itr = 1; itr <= iter_prec_end; itr++){
            double CND = (T - t_00) * t_range_inv;
            double DEP = (t_0 - T) * t_range_inv;
            if(T <= t_00){ CND = 0.0; DEP = 1.0; }
            if(T >= t_0){ CND = 1.0; DEP = 0.0; }
            const bool berg = T < t_0 && q_v_b <= q_rain_now && (q_i_b > 0.0 || q_c_b <= 0.0);
            if(berg){
                CND = 0.0; DEP = 1.0;
                if(q_c_b > 0.0){
                    T -= lv * q_c_b / (cp_mix * rho_c);
                    q_v_b += q_c_b;
                    q_c_b = 0.0;
                }
            }
            double d_q_v = q_v_hyp - q_v_b;
            double d_q_c = -d_q_v * CND;
            double d_q_i = -d_q_v * DEP;
            if(d_q_v < -q_v_b){ d_q_v = -q_v_b; d_q_c = -d_q_v * CND; d_q_i = -d_q_v * DEP; }
            if(q_c_b + d_q_c < 0.0){ const double ex = -(q_c_b + d_q_c); d_q_c = -q_c_b; d_q_i -= ex; }
            if(q_i_b + d_q_i < 0.0){ const double ex = -(q_i_b + d_q_i); d_q_i = -q_i_b; d_q_c -= ex;
                                     if(q_c_b + d_q_c < 0.0) d_q_c = -q_c_b; }
            d_q_v = -(d_q_c + d_q_i);
            T += (lv * d_q_c + ls * d_q_i) / (cp_mix * rho_c);
            q_v_b += d_q_v;
            q_c_b += d_q_c;
            q_i_b += d_q_i;
            if(q_v_b < 0.0) q_v_b = 0.0;
            if(q_c_b < 0.0) q_c_b = 0.0;
            if(q_i_b < 0.0) q_i_b = 0.0;
            const double E_Rain = svp(T, P.C, P.L0, P.R, P.del_alf, P.del_bet);
            const double E_Ice = svp(T, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i);
            const double q_Rain = rho_c * ep * E_Rain / p_u;
            const double q_Ice = rho_c * ep * E_Ice / p_u;
            q_rain_now = q_Rain;
            if(q_c_b > 0.0 && q_i_b > 0.0) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
            else if(q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_v_hyp = q_Ice;
            else if(q_i_b == 0.0) q_v_hyp = q_Rain;
            else                  q_v_hyp = q_Ice;
            if(T >= t_0){
                if(q_i_b > 0.0){
                    double dm = q_i_b;
                    if(Lf > 0.0){
                        const double hc = cp_mix * rho_c;
                        const double cap = hc * (T - t_0) / Lf;
                        if(dm > cap) dm = cap;
                        T -= Lf * dm / hc;
                    }
                    q_i_b -= dm;
                    q_c_b += dm;
                }
            }
            const double q_diff = std::fabs(q_v_b - q_v_hyp) / (q_v_hyp + 1e-20);
            if(q_diff <= q_diff_min) break;
            if(q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;
            const double dT = 0.1;
            const double q_Rain_d = rho_c * ep * svp(T + dT, P.C, P.L0, P.R, P.del_alf, P.del_bet) / p_u;
            const double q_Ice_d = rho_c * ep * svp(T + dT, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
            double q_hyp_d;
            if(q_c_b > 0.0 && q_i_b > 0.0) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
            else if(q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_hyp_d = q_Ice_d;
            else if(q_i_b == 0.0) q_hyp_d = q_Rain_d;
            else                  q_hyp_d = q_Ice_d;
            double CNDn = (T - t_00) * t_range_inv;
            if(T <= t_00) CNDn = 0.0;
            if(T >= t_0) CNDn = 1.0;
            if(berg) CNDn = 0.0;
            const double G = (lv * CNDn + ls * (1.0 - CNDn)) / (cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
            double w = 1.0 / (1.0 + std::max(G, 0.0));
            if(w < 0.05) w = 0.05;
            q_v_hyp = q_v_b + w * (q_v_hyp - q_v_b);
        }
        c[q] = q_v_b;
        cloud[q] = q_c_b;
        ice[q] = q_i_b;
        t[q] = T / t_ref;
    }
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++){
        if(c[q] <= 0.0) c[q] = 0.0;
        if(cloud[q] <= 0.0) cloud[q] = 0.0;
        if(ice[q] <= 0.0) ice[q] = 0.0;
    }
}

// ---- Turbulence<P>::run, k-omega, TURB_FREE 1, TURB_CLAMP_PASS 1, TURB_CURV 1, no ABL taper
//      (Turbulence.h:307-1012), then the wall-omega function of Run() (cJupiterModel.cpp:356-374) --
void turb_wall_omega(orc_ctx& o){
    if(o.kn.turb_model == 0) return;              // k-epsilon keeps its wall value (cJupiterModel.cpp:356-360)
    const double nue_gas_phys = 1.8e-5, bet_star = 0.09, bet_1 = 0.0333, kappa = 0.41;
    const double nd_omega = o.g.L_atm * 1.0e3 / o.c.u_0;
    auto& dis = o.f[ATJUP_F_DIS];
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < o.jm; j++)
        for(int k = 0; k < o.km; k++){
            const int i_w = o.itopo[(size_t)j * o.km + k];
            if(i_w <= 0 || i_w >= o.im - 1) continue;
            const double y1 = std::max(o.layer_thickness_m(std::max(i_w - 1, 0)), 1.0);
            const double y = 0.5 * y1;
            const double om_vis = 60.0 * nue_gas_phys / (bet_1 * y1 * y1);
            const double om_log = o.vel_star[(size_t)j * o.km + k] / (std::sqrt(bet_star) * kappa * y);
            dis[o.id(i_w, j, k)] = std::sqrt(om_vis * om_vis + om_log * om_log) * nd_omega;
        }
}

void turb_apply_wall_bc(orc_ctx& o){              // Turbulence.h:378-420, TURB_FREE: i = 0 copy
    auto& tke = o.f[ATJUP_F_TKE];
    auto& dis = o.f[ATJUP_F_DIS];
    auto& nue = o.f[ATJUP_F_NUE];
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < o.jm; j++)
        for(int k = 0; k < o.km; k++){
            tke[o.id(0, j, k)] = tke[o.id(1, j, k)];
            dis[o.id(0, j, k)] = dis[o.id(1, j, k)];
            nue[o.id(0, j, k)] = nue[o.id(1, j, k)];
        }
}

void turbulence(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    auto& tke = o.f[ATJUP_F_TKE];
    auto& dis = o.f[ATJUP_F_DIS];
    auto& nue = o.f[ATJUP_F_NUE];
    auto& prod = o.f[ATJUP_F_PROD];
    auto& tsrc = o.f[ATJUP_F_TKE_SOURCE];
    auto& dsrc = o.f[ATJUP_F_DIS_SOURCE];
    auto& tken = o.f[ATJUP_NLEVEL + ATJUP_F_TKE];
    auto& disn = o.f[ATJUP_NLEVEL + ATJUP_F_DIS];
    const auto& U = o.f[ATJUP_F_U];
    const auto& V = o.f[ATJUP_F_V];
    const auto& W = o.f[ATJUP_F_W];
    const double dis_min = 1.0e-10;
    const double L_m = o.g.L_atm * 1.0e3;
    const double nue_max = o.kn.nue_max_phys / (o.c.u_0 * L_m);
    const double k_bg_nd = o.kn.turb_k_bg_phys / (o.c.u_0 * o.c.u_0);
    // the model (0 k-epsilon, 1 k-omega, 2 SST) and the class constants: C_mu, the H2/He viscosity
    // nue_air and re_turb = vel_star_ref z_0 / nue_gas of the default constructor (Turbulence.h:192,
    // :245-251)
    const int model = o.kn.turb_model;
    const double C_nue = 0.028, nue_air = 1.8e-5;
    const double re_turb_cls = 0.4 * 0.1 / 1.8e-5;
    // compute_vel_star with TURB_FREE: the calm-air floor everywhere (Turbulence.h:340-342)
    std::fill(o.vel_star.begin(), o.vel_star.end(), 0.05);
    // clamp-in pass
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++)
            for(int i = 1; i < im - 1; i++){
                if(o.solid(i, j, k)) continue;
                const size_t q = o.id(i, j, k);
                tke[q] = std::max(0.0, tke[q]);
                dis[q] = std::max(dis_min, dis[q]);
            }
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++){
            const int kw = o.wrap(k - 1), ke = o.wrap(k + 1);
            for(int i = 1; i < im - 1; i++){
                if(o.solid(i, j, k)) continue;
                const size_t q = o.id(i, j, k);
                const size_t qip = o.id(i + 1, j, k), qim = o.id(i - 1, j, k);
                const size_t qjp = o.id(i, j + 1, k), qjm = o.id(i, j - 1, k);
                const size_t qkp = o.id(i, j, ke), qkm = o.id(i, j, kw);
                const double rm = o.rad[i];
                const double exp_rm = 1.0;
                double sinthe = o.sin_the[j];
                if(sinthe == 0.0) sinthe = 1.0e-5;
                const double costhe = o.cos_the[j];
                const double cotanthe = costhe / sinthe;
                const double inv_rm = 1.0 / rm;
                const double rmsinthe = rm * sinthe;
                const double inv_2dr = 1.0 / (2.0 * o.g.dr);
                const double inv_2dthe = 1.0 / (2.0 * o.g.dthe);
                const double inv_2dphi = 1.0 / (2.0 * o.g.dphi);
                const double inv_rm2dthe = inv_2dthe / rm;
                const double inv_rmsinthe2dphi = inv_2dphi / rmsinthe;
                const double dudr = (U[qip] - U[qim]) * inv_2dr * exp_rm;
                const double dvdr = (V[qip] - V[qim]) * inv_2dr * exp_rm;
                const double dwdr = (W[qip] - W[qim]) * inv_2dr * exp_rm;
                const double dudthe = (U[qjp] - U[qjm]) * inv_rm2dthe;
                const double dvdthe = (V[qjp] - V[qjm]) * inv_rm2dthe;
                const double dwdthe = (W[qjp] - W[qjm]) * inv_rm2dthe;
                const double dudphi = (U[qkp] - U[qkm]) * inv_rmsinthe2dphi;
                const double dvdphi = (V[qkp] - V[qkm]) * inv_rmsinthe2dphi;
                const double dwdphi = (W[qkp] - W[qkm]) * inv_rmsinthe2dphi;
                const bool nb = o.solid(i - 1, j, k);
                const bool njm = o.solid(i, j - 1, k), njp = o.solid(i, j + 1, k);
                const bool nkm = o.solid(i, j, kw), nkp = o.solid(i, j, ke);
                const double tke_im1 = nb ? tke[q] : tke[qim];
                const double dis_im1 = nb ? dis[q] : dis[qim];
                const double tke_jm1 = njm ? tke[q] : tke[qjm];
                const double tke_jp1 = njp ? tke[q] : tke[qjp];
                const double dis_jm1 = njm ? dis[q] : dis[qjm];
                const double dis_jp1 = njp ? dis[q] : dis[qjp];
                const double tke_km1 = nkm ? tke[q] : tke[qkm];
                const double tke_kp1 = nkp ? tke[q] : tke[qkp];
                const double dis_km1 = nkm ? dis[q] : dis[qkm];
                const double dis_kp1 = nkp ? dis[q] : dis[qkp];
                const double dtkedr = (tke[qip] - tke_im1) * inv_2dr * exp_rm;
                const double ddisdr = (dis[qip] - dis_im1) * inv_2dr * exp_rm;
                (void)dudthe;
                double cnue;
                {
                    const double dis_here = std::max(dis[q], dis_min);
                    const double tke_here = std::max(tke[q], 0.0);
                    cnue = (model == 0) ? C_nue * tke_here * tke_here / dis_here : tke_here / dis_here;
                    cnue = std::min(cnue, nue_max);
                }
                const double u_r = U[q] * inv_rm, v_r = V[q] * inv_rm, w_r = W[q] * inv_rm;
                const double v_cot_r = v_r * cotanthe, w_cot_r = w_r * cotanthe;
                const double g_rr = dudr, g_rt = dudthe - v_r, g_rp = dudphi - w_r;
                const double g_tr = dvdr, g_tt = dvdthe + u_r, g_tp = dvdphi - w_cot_r;
                const double g_pr = dwdr, g_pt = dwdthe, g_pp = dwdphi + u_r + v_cot_r;
                const double der = 0.66667 * (g_rr + g_tt + g_pp);
                prod[q] = std::max(0.0,
                      (cnue * (2.0 * g_rr - der) - 0.66667 * tke[q]) * g_rr
                    + (cnue * (g_rt + g_tr)) * g_rt
                    + (cnue * (g_rp + g_pr)) * g_rp
                    + (cnue * (2.0 * g_tt - der) - 0.66667 * tke[q]) * g_tt
                    + (cnue * (g_tr + g_rt)) * g_tr
                    + (cnue * (g_tp + g_pt)) * g_tp
                    + (cnue * (2.0 * g_pp - der) - 0.66667 * tke[q]) * g_pp
                    + (cnue * (g_pr + g_rp)) * g_pr
                    + (cnue * (g_pt + g_tp)) * g_pt);
                const double W12 = g_rt - g_tr, W13 = g_rp - g_pr, W23 = g_tp - g_pt;
                const double Omega_mag = std::sqrt(W12 * W12 + W13 * W13 + W23 * W23);
                if(model == 0){
                    // compute_k_epsilon (Turbulence.h:886-918) with TURB_FREE: the wall distance is
                    // free_wall_dist() = 1e20 m, so the Chien damping takes its free-flow form
                    const double C_eps_1 = 1.35, C_eps_2 = 1.80;
                    const double y_phys = 1.0e20;
                    const double y_star = y_phys / L_m;
                    nue[q] = C_nue * tke[q] * tke[q] / std::max(dis[q], dis_min);
                    const double d_plus = y_phys * o.vel_star[(size_t)j * km + k] / nue_air;
                    const double f_nue = 1.0 - std::exp(-0.0115 * d_plus);
                    const double Re_T = tke[q] * tke[q] * o.c.u_0 * L_m
                                      / std::max(dis[q] * nue_air, dis_min * nue_air);
                    const double f_2 = 1.0 - 0.4 / 1.8 * std::exp(-Re_T * Re_T / 36.0);
                    nue[q] *= f_nue;
                    const double y_star2 = y_star * y_star;
                    const double L_k = -2.0 * tke[q] / y_star2;
                    const double L_w = -2.0 * dis[q] / y_star2 * std::exp(-d_plus / 2.0);
                    const double P_k = prod[q];
                    const double Y_k = dis[q];
                    const double tke_safe = std::max(tke[q], dis_min);
                    const double P_w = C_eps_1 * dis[q] / tke_safe * P_k;
                    const double Y_w = C_eps_2 * f_2 * dis[q] * dis[q] / tke_safe;
                    tsrc[q] = P_k - Y_k + L_k / re_turb_cls;
                    dsrc[q] = P_w - Y_w + L_w / re_turb_cls;
                    continue;
                }
                // compute_k_omega (Turbulence.h:920-1012); compute_k_omega_SST shares its
                // Neumann-aware gradients with the unconditional 1 / (rm + 1)
                const double bet_star = 0.09, gam = 0.52, C_lim = 0.875, bet_0 = 0.0708;
                const double erm = 1.0 / (rm + 1.0);
                const double dtkedr_neu = nb ? (tke[qip] - tke[q]) * (1.0 / (2.0 * o.g.dr)) * erm : dtkedr;
                const double ddisdr_neu = nb ? (dis[qip] - dis[q]) * (1.0 / (2.0 * o.g.dr)) * erm : ddisdr;
                const double dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * 2.0 * o.g.dthe);
                const double ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * 2.0 * o.g.dthe);
                const double dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * 2.0 * o.g.dphi);
                const double ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * 2.0 * o.g.dphi);
                if(model == 2){
                    // compute_k_omega_SST (Turbulence.h:1021-1106), Menter 1994, TURB_FREE wall distance
                    const double a1 = 0.31, bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2;
                    const double sig_w2 = 1.168;
                    const double nue_air_nd = nue_air / (o.c.u_0 * L_m);
                    const double y_phys = 1.0e20;
                    const double y_star = y_phys / L_m;
                    const double grad_dot = dtkedr_neu * ddisdr_neu
                                          + dtkedthe_neu * ddisdthe_neu
                                          + dtkedphi_neu * ddisdphi_neu;
                    const double CD_kw = std::max(2.0 * sig_w2 / std::max(dis[q], 1.0e-20) * grad_dot, 1.0e-20);
                    const double dis_safe = std::max(dis[q], 1.0e-20);
                    const double arg1 = std::min(
                        std::max(std::sqrt(tke[q]) / (bet_star * dis_safe * y_star),
                                 500.0 * nue_air_nd / (y_star * y_star * dis_safe)),
                        4.0 * sig_w2 * tke[q] / (CD_kw * y_star * y_star));
                    const double arg2 = std::max(
                        2.0 * std::sqrt(tke[q]) / (bet_star * dis_safe * y_star),
                        500.0 * nue_air_nd / (y_star * y_star * dis_safe));
                    const double F1 = std::tanh(std::pow(arg1, 4));
                    const double F2 = std::tanh(std::pow(arg2, 2));
                    nue[q] = a1 * tke[q] / std::max(a1 * dis_safe, Omega_mag * F2);
                    const double tke_safe = std::max(tke[q], 1.0e-20);
                    const double P_k = std::min(prod[q], 20.0 * bet_star * tke_safe * dis_safe);
                    const double Y_k = bet_star * tke_safe * dis_safe;
                    const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_safe / tke_safe;
                    const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_safe * dis_safe;
                    const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_safe * grad_dot;
                    tsrc[q] = P_k - Y_k;
                    dsrc[q] = P_w - Y_w + D_w;
                    continue;
                }
                const double dudr_k = (U[qip] - U[qim]) / (2.0 * o.g.dr) / (rm + 1.0);
                const double dvdthe_k = (V[qjp] - V[qjm]) / (rm * 2.0 * o.g.dthe);
                const double dwdphi_k = (W[qkp] - W[qkm]) / (rmsinthe * 2.0 * o.g.dphi);
                const double dudthe_u = (U[qjp] - U[qjm]) / (rm * 2.0 * o.g.dthe);
                const double dudphi_u = (U[qkp] - U[qkm]) / (rmsinthe * 2.0 * o.g.dphi);
                const double dvdr_u = (V[qip] - V[qim]) / (2.0 * o.g.dr) / (rm + 1.0);
                const double dwdr_u = (W[qip] - W[qim]) / (2.0 * o.g.dr) / (rm + 1.0);
                const double dvdphi_u = (V[qkp] - V[qkm]) / (rmsinthe * 2.0 * o.g.dphi);
                const double dwdthe_u = (W[qjp] - W[qjm]) / (rm * 2.0 * o.g.dthe);
                const double S11 = dudr_k, S22 = dvdthe_k, S33 = dwdphi_k;
                const double S12 = 0.5 * (dudthe_u + dvdr_u);
                const double S13 = 0.5 * (dudphi_u + dwdr_u);
                const double S23 = 0.5 * (dvdphi_u + dwdthe_u);
                const double S_mag = std::sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33
                                                    + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
                const double w_hat = std::max(dis[q], C_lim * S_mag / std::sqrt(bet_star));
                double sig_d = dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
                sig_d = (sig_d <= 0.0) ? 0.0 : 0.125;
                const double chi_w = std::fabs(Omega_mag * Omega_mag * S_mag
                    / std::pow(bet_star * std::max(dis[q], 1.0e-20), 3));
                const double f_bet = (1.0 + 85.0 * chi_w) / (1.0 + 100.0 * chi_w);
                const double bet_wc = bet_0 * f_bet;
                nue[q] = tke[q] / std::max(w_hat, 1.0e-20);
                const double P_k = std::min(prod[q], 20.0 * bet_star * tke[q] * dis[q]);
                const double Y_k = bet_star * tke[q] * dis[q];
                const double P_w = gam * dis[q] / std::max(tke[q], 1.0e-20) * P_k;
                const double Y_w = bet_wc * dis[q] * dis[q];
                const double D_w = sig_d / std::max(dis[q], 1.0e-20)
                    * (dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu);
                tsrc[q] = P_k - Y_k;
                dsrc[q] = P_w - Y_w + D_w;
            }
        }
    // clamp-out pass
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++)
            for(int i = 1; i < im - 1; i++){
                if(o.solid(i, j, k)) continue;
                const size_t q = o.id(i, j, k);
                tke[q] = std::max(k_bg_nd, tke[q]);
                dis[q] = std::max(dis_min, dis[q]);
                nue[q] = (nue[q] < 0.0) ? 0.0 : ((nue[q] > nue_max) ? nue_max : nue[q]);
                const double dsm = 20.0 * dis[q] * dis[q];
                dsrc[q] = (dsrc[q] < -dsm) ? -dsm : ((dsrc[q] > dsm) ? dsm : dsrc[q]);
                const double tsm = 20.0 * tke[q] * dis[q];
                tsrc[q] = (tsrc[q] < -tsm) ? -tsm : ((tsrc[q] > tsm) ? tsm : tsrc[q]);
            }
    turb_apply_wall_bc(o);
    // zero_land_cells
    std::vector<double>* eight[] = { &tke, &tken, &dis, &disn, &nue, &prod, &tsrc, &dsrc };
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++)
        if(o.mask[q] == 1.0) for(auto* a : eight) (*a)[q] = 0.0;
    // clamp_nue
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++)
        nue[q] = ((nue[q] < 0.0) ? 0.0 : ((nue[q] > nue_max) ? nue_max : nue[q])) * 1.0;
    // seam_copy
    std::vector<double>* seam8[] = { &tke, &dis, &tken, &disn, &nue, &prod, &tsrc, &dsrc };
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 0; i < im; i++)
        for(int j = 0; j < jm; j++)
            for(auto* a : seam8) (*a)[o.id(i, j, km - 1)] = (*a)[o.id(i, j, 0)];
    // fill_lid_corners, TURB_FREE without the separate turbulence BC lists
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km; k++)
            for(auto* a : eight) (*a)[o.id(im - 1, j, k)] = (*a)[o.id(im - 2, j, k)];
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 0; i < im; i++)
        for(int k = 0; k < km; k++)
            for(auto* a : eight){
                (*a)[o.id(i, 0, k)] = (*a)[o.id(i, 1, k)];
                (*a)[o.id(i, jm - 1, k)] = (*a)[o.id(i, jm - 2, k)];
            }
}

// ---- FluxLimiter<P>::nh4sh, Superbee (FluxLimiter.h:67-190) ----------------------------------
void flux_limiter(orc_ctx& o){
    const int im = o.im, jm = o.jm, km = o.km;
    const auto& N = o.f[ATJUP_F_NH4SH];
    const auto& U = o.f[ATJUP_F_U];
    const auto& V = o.f[ATJUP_F_V];
    const auto& W = o.f[ATJUP_F_W];
    auto& fl = o.f[ATJUP_F_FLUXLIM_NH4SH];
    const double dr = o.g.dr, dthe = o.g.dthe, dphi = o.g.dphi, eps = 1.0e-12;
    auto superbee = [](double r){ return std::max(0.0, std::max(std::min(2.0 * r, 1.0), std::min(r, 2.0))); };
    #pragma omp parallel for collapse(2) schedule(static)
    for(int k = 0; k < km - 1; k++)
        for(int j = 1; j < jm - 1; j++)
            for(int i = 1; i < im - 1; i++){
                const double q = N[o.id(i, j, k)];
                const double q_rp = N[o.id(i + 1, j, k)], q_rm = N[o.id(i - 1, j, k)];
                const double q_tp = N[o.id(i, j + 1, k)], q_tm = N[o.id(i, j - 1, k)];
                const double q_pp = N[o.id(i, j, o.wrap(k + 1))], q_pm = N[o.id(i, j, o.wrap(k - 1))];
                const double u = U[o.id(i, j, k)], v = V[o.id(i, j, k)], w = W[o.id(i, j, k)];
                const double rm = o.rad[i];
                const double sinthe = std::max(0.4, std::abs(o.sin_the[j]));
                const double inv_rm = 1.0 / rm;
                const double inv_rmsinthe = 1.0 / (rm * sinthe);
                double corr = 0.0;
                {
                    const double df = q_rp - q, db = q - q_rm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double r = (u >= 0.0) ? db / denom
                        : ((i + 2 < im ? N[o.id(i + 2, j, k)] : q_rp) - q_rp) / denom;
                    corr += (1.0 - superbee(r)) * std::abs(u) * (df - db) / (2.0 * dr);
                }
                {
                    const double df = q_tp - q, db = q - q_tm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double q_far = (j + 2 < jm) ? N[o.id(i, j + 2, k)] : (o.c43 * q_tp - o.c13 * q);
                    const double r = (v >= 0.0) ? db / denom : (q_far - q_tp) / denom;
                    corr += (1.0 - superbee(r)) * std::abs(v) * inv_rm * (df - db) / (2.0 * dthe);
                }
                {
                    const double df = q_pp - q, db = q - q_pm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double r = (w >= 0.0) ? db / denom : (N[o.id(i, j, o.wrap(k + 2))] - q_pp) / denom;
                    corr += (1.0 - superbee(r)) * std::abs(w) * inv_rmsinthe * (df - db) / (2.0 * dphi);
                }
                fl[o.id(i, j, k)] = corr;
            }
}

// ---- Radiation<P>::run (Radiation.h:245-487): grey two-stream column solve ---------------------
// For every (j, k): layer emission sigma T^4 and emissivity 1 - exp(-tau) with tau the CIA, CH4/NH3
// band and capped cloud optical depths (:311-349), the Beer-Lambert shortwave beam from the top
// (:369-386), the downward thermal sweep from a cold top (:391-393), the deep boundary
// (:411-414), the upward sweep (:418-419), and per layer the centre net flux and the heating from
// the flux divergence plus the absorbed shortwave (:434-463); solid cells below the column's base
// copy its values with zero heating (:467-471). The printed diagnostics are not formed.
void radiation(orc_ctx& o){
    const atjup_rad_params& R = o.radp;
    const int im = o.im, jm = o.jm, km = o.km;
    const double sigma = 5.670374419e-8, P0 = 1.0e5, he_ratio = 0.6, C_cia = 3.5e-6;
    const double kappa_ch4 = 0.003, kappa_nh3 = 1.5, kappa_cloud = 25.0, kappa_ice = 12.0;
    const double tau_cloud_cap = 0.4, opac_cal = 0.025, pi_ = 3.14159265358979323846;
    const double t_ref = o.c.t_ref, g = o.c.g;
    const double comp = R.x_H2 * R.x_H2 + he_ratio * R.x_H2 * R.x_He;
    const double cia_coeff = C_cia * R.cia_strength * comp / g;
    const std::vector<double>& T_ = o.f[ATJUP_F_T];
    const std::vector<double>& PS = o.f[ATJUP_F_P_STAT];
    const std::vector<double>& RHO = o.f[ATJUP_F_RHO_MIX];
    std::vector<double>& EPS = o.f[ATJUP_F_EPSILON];
    std::vector<double>& RADF = o.f[ATJUP_F_RADIATION];
    std::vector<double>& QR = o.f[ATJUP_F_Q_RAD];
    auto sp = [&](int f, size_t q){ return std::max(0.0, o.f[f][q]); };
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            const int i_base = std::max(o.itopo[(size_t)j * km + k], 0), i_top = im - 1;
            if(i_base >= i_top) continue;
            std::vector<double> B(im, 0.0), eps(im, 0.0), Fd(im + 1, 0.0), Fu(im + 1, 0.0), Fsw(im + 2, 0.0);
            for(int i = i_base; i <= i_top; i++){
                const size_t q = o.id(i, j, k);
                const double T = T_[q] * t_ref;
                B[i] = sigma * T * T * T * T;
                const double P_lo = PS[q] * P0;
                const double P_hi = (i < i_top) ? PS[o.id(i + 1, j, k)] * P0 : 0.0;
                double dP = P_lo - P_hi;
                if(dP < 0.0) dP = 0.0;
                const double P_mean = 0.5 * (P_lo + P_hi);
                const double dm = dP / g;
                const double tau_cia = (T > 0.0) ? cia_coeff * (P_mean / T) * dP : 0.0;
                const double rho_c = RHO[q];
                const double inv_rho = (rho_c > 0.0 && std::isfinite(rho_c)) ? 1.0 / rho_c : 0.0;
                const double tau_gas = R.opacity_strength * opac_cal
                    * (kappa_ch4 * (sp(ATJUP_F_CH4, q) * inv_rho) + kappa_nh3 * (sp(ATJUP_F_NH3, q) * inv_rho)) * dm;
                const double q_liq = (sp(ATJUP_F_H2O_CLOUD, q) + sp(ATJUP_F_NH3_CLOUD, q) + sp(ATJUP_F_CH4_CLOUD, q)) * inv_rho;
                const double q_ice = (sp(ATJUP_F_H2O_ICE, q) + sp(ATJUP_F_NH3_ICE, q) + sp(ATJUP_F_CH4_ICE, q)) * inv_rho;
                double tau_cloud = R.opacity_strength * opac_cal * (kappa_cloud * q_liq + kappa_ice * q_ice) * dm;
                if(tau_cloud > tau_cloud_cap) tau_cloud = tau_cloud_cap;
                eps[i] = 1.0 - std::exp(-(tau_cia + tau_gas + tau_cloud));
            }
            if(R.solar != 0){
                const double colat = pi_ * (double)j / (double)(jm - 1);
                const double F_top = R.solar_strength * R.S_solar * (1.0 - R.albedo_bond)
                                   * std::max(0.0, std::sin(colat)) / pi_;
                for(int i = i_base; i <= i_top + 1; i++)
                    Fsw[i] = F_top * std::exp(-R.sw_tau_per_bar * std::max(0.0, (i <= i_top) ? PS[o.id(i, j, k)] : 0.0));
            }
            for(int i = i_top; i >= i_base; i--) Fd[i] = Fd[i + 1] * (1.0 - eps[i]) + eps[i] * B[i];
            Fu[i_base] = (R.bottom == 1) ? std::max(0.0, 2.0 * B[i_base] - Fd[i_base + 1])
                                         : Fd[i_base] + R.F_int + Fsw[i_base];
            for(int i = i_base; i <= i_top; i++) Fu[i + 1] = Fu[i] * (1.0 - eps[i]) + eps[i] * B[i];
            for(int i = i_base; i <= i_top; i++){
                const size_t q = o.id(i, j, k);
                const double lo = Fu[i] - Fd[i], hi = Fu[i + 1] - Fd[i + 1];
                double dz = o.layer_thickness_m(i < i_top ? i : i - 1);
                if(dz <= 0.0) dz = 1.0;
                EPS[q] = eps[i];
                RADF[q] = 0.5 * (lo + hi);
                QR[q] = (lo - hi) / dz + (Fsw[i + 1] - Fsw[i]) / dz;
            }
            for(int i = 0; i < i_base; i++){
                EPS[o.id(i, j, k)] = EPS[o.id(i_base, j, k)];
                RADF[o.id(i, j, k)] = RADF[o.id(i_base, j, k)];
                QR[o.id(i, j, k)] = 0.0;
            }
        }
}

// ---- Precipitation<P>::run (Precipitation.h:194-574) -----------------------------------------
// Zero Q_precip and the nine rates (:199-210); the three-category column of H2O, NH3, CH4 in that
// order over every column (column(), :273-450); the NH4SH crystal settling (:523-574); the surface
// maps at each column's base (:470-492). The diagnostics print (ATJUP_PRECIP_DIAG) is not formed.
void precip_column(orc_ctx& o, int gas){
    static const int vap[3] = { ATJUP_F_H2O, ATJUP_F_NH3, ATJUP_F_CH4 };
    static const int pf[3] = { ATJUP_F_P_RAIN, ATJUP_F_P_NH3_RAIN, ATJUP_F_P_CH4_RAIN };
    static const int sf[3] = { ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_NH3, ATJUP_F_S_PRECIP_CH4 };
    const atjup_gas_params& G = o.gas[gas];
    const int im = o.im, jm = o.jm, km = o.km;
    const double t_ref = o.c.t_ref, Lf = G.ls - G.lv, rs = o.precp.rate_scale;
    const double c_c_au = 2.0e-8, c_i_au = 5.0e-8, q_c_crit = 5.0e-4, c_ac = 1.0e-7, c_rim = 1.0e-8;
    const double z_csg = 2.5e-8, a_ev = 2.76e-3, c_melt_K = 0.2, c_frz_K = 0.2, P_ref = 1.0e-4;
    const double ratio_cap = 5.0, v_fall = 2.0, P_max_flux = 1.0e-4, dep_frac = 1.0;
    const bool coupling = o.kn.precip_coupling != 0.0;
    std::vector<double>& V = o.f[vap[gas]];
    std::vector<double>& CL = o.f[vap[gas] + 1];
    std::vector<double>& IC = o.f[vap[gas] + 2];
    std::vector<double>& PR = o.f[pf[gas]];
    std::vector<double>& PS_ = o.f[pf[gas] + 1];
    std::vector<double>& PG = o.f[pf[gas] + 2];
    std::vector<double>& SV = o.f[sf[gas]];
    std::vector<double>& SC = o.f[sf[gas] + 1];
    std::vector<double>& SI = o.f[sf[gas] + 2];
    std::vector<double>& Q = o.f[ATJUP_F_Q_PRECIP];
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            const int i_base = std::max(o.itopo[(size_t)j * km + k], 0), i_top = im - 1;
            if(i_base >= i_top) continue;
            const size_t qt = o.id(i_top, j, k);
            PR[qt] = 0.0; PS_[qt] = 0.0; PG[qt] = 0.0;
            for(int i = i_top - 1; i >= i_base; i--){
                const size_t q = o.id(i, j, k), qa = o.id(i + 1, j, k);
                const double T = o.f[ATJUP_F_T][q] * t_ref;
                const double p_u = o.f[ATJUP_F_P_STAT][q];
                const double q_v = std::max(0.0, V[q]), q_c = std::max(0.0, CL[q]), q_i = std::max(0.0, IC[q]);
                const double E_r = svp(T, G.C, G.L0, G.R, G.del_alf, G.del_bet);
                const double rf = o.f[ATJUP_F_RHO_MIX][q];
                const double rho_at = (rf > 0.0 && std::isfinite(rf)) ? rf : o.c.r_mix;
                const double q_sat = (p_u > E_r) ? rho_at * G.ep * E_r / p_u : q_v;
                const double dz = std::max(0.0, o.layer_thickness_m(i));
                if(dz <= 0.0){ PR[q] = PR[qa]; PS_[q] = PS_[qa]; PG[q] = PG[qa]; continue; }
                const double F_r_in = PR[qa], F_s_in = PS_[qa], F_g_in = PG[qa];
                const double Rain = std::min(ratio_cap, F_r_in / P_ref);
                const double Snow = std::min(ratio_cap, F_s_in / P_ref);
                const double Graupel = std::min(ratio_cap, F_g_in / P_ref);
                const double Rain_79 = (Rain > 0.0) ? std::pow(Rain, 7.0 / 9.0) : 0.0;
                const double Rain_49 = (Rain > 0.0) ? std::pow(Rain, 4.0 / 9.0) : 0.0;
                const bool warm = T >= G.t_0;
                const double S_c_au = (warm && q_c > q_c_crit) ? (rs * c_c_au) * (q_c - q_c_crit) : 0.0;
                const double S_i_au = (!warm && T >= G.t_00) ? (rs * c_i_au) * q_i : 0.0;
                const double S_ac = warm ? (rs * c_ac) * q_c * Rain_79 : 0.0;
                const double S_s_rim = !warm ? (rs * c_rim) * q_c * Snow : 0.0;
                const double S_g_rim = !warm ? (rs * c_rim) * q_c * Graupel : 0.0;
                const double S_csg = (!warm && q_c > q_c_crit) ? (rs * z_csg) * q_c * Snow : 0.0;
                const double f_melt = warm ? std::min(1.0, c_melt_K * (T - G.t_0)) : 0.0;
                const double f_frz = warm ? 0.0 : std::min(1.0, c_frz_K * (G.t_0 - T));
                const double F_s_melt = f_melt * F_s_in, F_g_melt = f_melt * F_g_in, F_r_frz = f_frz * F_r_in;
                const double F_r_avail = std::max(0.0, F_r_in - F_r_frz + F_s_melt + F_g_melt + (S_c_au + S_ac) * dz);
                const double S_ev_raw = (warm && Rain > 0.0 && q_v < q_sat) ? a_ev * (q_sat - q_v) * Rain_49 : 0.0;
                const double F_ev_sat = std::max(0.0, q_sat - q_v) * v_fall;
                const double F_ev = std::min(std::min(S_ev_raw * dz, F_r_avail), F_ev_sat);
                const double S_ev = F_ev / dz;
                Q[q] += Lf * (S_s_rim + S_g_rim + S_csg) + Lf * (F_r_frz - F_s_melt - F_g_melt) / dz - G.lv * F_ev / dz;
                const double tau = dz / v_fall;
                const double d_c = std::min(dep_frac * q_c, (S_c_au + S_ac + S_s_rim + S_g_rim + S_csg) * tau);
                const double d_i = std::min(dep_frac * q_i, S_i_au * tau);
                if(coupling){
                    const double nd = (o.g.L_atm * 1.0e3) / o.c.u_0;
                    SC[q] = -std::max(0.0, d_c) / tau * nd;
                    SI[q] = -std::max(0.0, d_i) / tau * nd;
                    SV[q] = +S_ev * nd;
                } else {
                    CL[q] = q_c - std::max(0.0, d_c);
                    IC[q] = q_i - std::max(0.0, d_i);
                    V[q] = q_v + S_ev * tau;
                }
                const double F_r = F_r_in - F_r_frz + F_s_melt + F_g_melt - F_ev + (S_c_au + S_ac) * dz;
                const double F_s = F_s_in - F_s_melt + (S_i_au + S_s_rim - S_csg) * dz;
                const double F_g = F_g_in - F_g_melt + F_r_frz + (S_g_rim + S_csg) * dz;
                PR[q] = std::min(P_max_flux, std::max(0.0, F_r));
                PS_[q] = std::min(P_max_flux, std::max(0.0, F_s));
                PG[q] = std::min(P_max_flux, std::max(0.0, F_g));
            }
            const size_t qb = o.id(i_base, j, k);
            for(int i = 0; i < i_base; i++){
                const size_t q = o.id(i, j, k);
                PR[q] = PR[qb]; PS_[q] = PS_[qb]; PG[q] = PG[qb];
            }
        }
}

void precipitation(orc_ctx& o){
    static const int zero_me[] = { ATJUP_F_Q_PRECIP, ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_H2O_CLOUD,
        ATJUP_F_S_PRECIP_H2O_ICE, ATJUP_F_S_PRECIP_NH3, ATJUP_F_S_PRECIP_NH3_CLOUD, ATJUP_F_S_PRECIP_NH3_ICE,
        ATJUP_F_S_PRECIP_CH4, ATJUP_F_S_PRECIP_CH4_CLOUD, ATJUP_F_S_PRECIP_CH4_ICE };
    for(int f : zero_me) std::fill(o.f[f].begin(), o.f[f].end(), 0.0);
    for(int gas = 0; gas < 3; gas++) precip_column(o, gas);
    // NH4SH crystals: Stokes settling in [t_00_nh4sh, t_0_nh4sh], the base's flux below the base
    const int im = o.im, jm = o.jm, km = o.km;
    const double r_p = 1.0e-5, rho_c = 1200.0, mu = 9.0e-6;
    const double v_settle = (2.0 / 9.0) * r_p * r_p * (rho_c - o.c.r_mix) * o.c.g / mu;
    std::vector<double>& PN = o.f[ATJUP_F_P_NH4SH];
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            const int i_base = std::max(o.itopo[(size_t)j * km + k], 0);
            for(int i = i_base; i < im; i++){
                const double T = o.f[ATJUP_F_T][o.id(i, j, k)] * o.c.t_ref;
                const double qn = std::max(0.0, o.f[ATJUP_F_NH4SH][o.id(i, j, k)]);
                PN[o.id(i, j, k)] = (T >= o.precp.t_00_nh4sh && T <= o.precp.t_0_nh4sh) ? qn * v_settle : 0.0;
            }
            const double base_flux = (i_base < im) ? PN[o.id(i_base, j, k)] : 0.0;
            for(int i = 0; i < i_base && i < im; i++) PN[o.id(i, j, k)] = base_flux;
        }
    // surface maps at the base
    const size_t n2 = (size_t)jm * km;
    o.precip_srf.assign(5 * n2, 0.0);
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            const int ib = std::min(std::max(o.itopo[(size_t)j * km + k], 0), im - 1);
            const size_t q = o.id(ib, j, k), c2 = (size_t)j * km + k;
            const double h2o = o.f[ATJUP_F_P_RAIN][q] + o.f[ATJUP_F_P_SNOW][q] + o.f[ATJUP_F_P_GRAUPEL][q];
            const double nh3 = o.f[ATJUP_F_P_NH3_RAIN][q] + o.f[ATJUP_F_P_NH3_SNOW][q] + o.f[ATJUP_F_P_NH3_GRAUPEL][q];
            const double ch4 = o.f[ATJUP_F_P_CH4_RAIN][q] + o.f[ATJUP_F_P_CH4_SNOW][q] + o.f[ATJUP_F_P_CH4_GRAUPEL][q];
            const double nh4 = PN[q];
            o.precip_srf[c2] = h2o; o.precip_srf[n2 + c2] = nh3; o.precip_srf[2 * n2 + c2] = ch4;
            o.precip_srf[3 * n2 + c2] = nh4; o.precip_srf[4 * n2 + c2] = h2o + nh3 + ch4 + nh4;
        }
}

// ---- ChemistryJup::DiffMassFluxJup (ChemistryJup.h:375-540) ----------------------------------
// derivative_1_order at a fluid cell (:653-727): central unless both neighbours are solid
inline double chem_d1(double cm, double cp, bool sm, bool sp, double two_d){
    return (sm && sp) ? 0.0 : (cp - cm) / two_d;
}

void diffmassflux(orc_ctx& o){
    const atjup_chem_params& p = o.chemp;
    const int im = o.im, jm = o.jm, km = o.km;
    const double DT_nh3 = p.mue_nh3 / (p.rg_nh3 * o.c.sc_nh3);
    const double DT_h2s = p.mue_h2s / (p.rg_h2s * o.c.sc_h2s);
    const double DT_nh4sh = p.mue_nh4sh / (o.c.rg_nh4sh * o.c.sc_nh4sh);
    const double M_mix = p.X_h2 * p.m_h2 + p.X_he * p.m_he + p.X_h2o * p.m_h2o + p.X_h2s * p.m_h2s
                       + p.X_nh3 * p.m_nh3 + p.X_nh4sh * p.m_nh4sh;
    const double L_nh3 = o.c.r_mix * o.c.cp_mix * DT_nh3 / p.k_mix;
    const double L_h2s = o.c.r_mix * o.c.cp_mix * DT_h2s / p.k_mix;
    const double L_nh4sh = o.c.r_mix * o.c.cp_mix * DT_nh4sh / p.k_mix;
    const double floor_ = 0.4;
    std::vector<double>& T_ = o.f[ATJUP_F_T];
    std::vector<double>* X[3] = { &o.f[ATJUP_F_H2S], &o.f[ATJUP_F_NH3], &o.f[ATJUP_F_NH4SH] };
    #pragma omp parallel for schedule(static)
    for(int k = 0; k < km - 1; k++)
        for(int j = 1; j < jm - 1; j++){
            const double sinthe = std::max(std::sin(o.the[j]), floor_);
            for(int i = 1; i < im - 1; i++){
                const double rm = o.rad[i], rmsinthe = rm * sinthe;
                const size_t q = o.id(i, j, k);
                if(!(T_[q] > 0.0) || o.mask[q] == 1.0){
                    o.f[ATJUP_F_DIFFLUX_H2S][q] = 0.0; o.f[ATJUP_F_DIFFLUX_NH3][q] = 0.0;
                    o.f[ATJUP_F_DIFFLUX_NH4SH][q] = 0.0; o.f[ATJUP_F_THERMALMASSFLUX][q] = 0.0;
                    continue;
                }
                const int kw = o.wrap(k - 1), ke = o.wrap(k + 1);
                auto solid = [&](int ii, int jj, int kk){ return o.mask[o.id(ii, jj, kk)] == 1.0; };
                auto d3 = [&](const std::vector<double>& c, double& dr, double& dth, double& dph){
                    dr = chem_d1(c[o.id(i - 1, j, k)], c[o.id(i + 1, j, k)], solid(i - 1, j, k), solid(i + 1, j, k), 2.0 * o.g.dr);
                    dth = chem_d1(c[o.id(i, j - 1, k)], c[o.id(i, j + 1, k)], solid(i, j - 1, k), solid(i, j + 1, k), 2.0 * o.g.dthe);
                    if(k == 0) dph = (-3.0 * c[o.id(i, j, 0)] + 4.0 * c[o.id(i, j, 1)] - c[o.id(i, j, 2)]) / (2.0 * o.g.dphi);
                    else dph = chem_d1(c[o.id(i, j, kw)], c[o.id(i, j, ke)], solid(i, j, kw), solid(i, j, ke), 2.0 * o.g.dphi);
                };
                double dt[3], dh2s[3], dnh3[3], dnh4[3];
                d3(T_, dt[0], dt[1], dt[2]);
                d3(*X[0], dh2s[0], dh2s[1], dh2s[2]);
                d3(*X[1], dnh3[0], dnh3[1], dnh3[2]);
                d3(*X[2], dnh4[0], dnh4[1], dnh4[2]);
                const double dt_div = dt[0] + std::abs(dt[1]) / rm + dt[2] / rmsinthe;
                const double rf = o.f[ATJUP_F_RHO_MIX][q];
                const double rho_c = (rf > 0.0 && std::isfinite(rf)) ? rf : o.c.r_mix;
                const double sc = rho_c / o.c.r_mix;
                const double jT_nh3 = (L_nh3 * sc) / T_[q] * dt_div;
                const double jT_h2s = (L_h2s * sc) / T_[q] * dt_div;
                const double jT_nh4sh = (L_nh4sh * sc) / T_[q] * dt_div;
                const double dnh3_div = dnh3[0] + std::abs(dnh3[1]) / rm + dnh3[2] / rmsinthe;
                const double dh2s_div = dh2s[0] + std::abs(dh2s[1]) / rm + dh2s[2] / rmsinthe;
                const double dnh4sh_div = dnh4[0] + std::abs(dnh4[1]) / rm + dnh4[2] / rmsinthe;
                const double Lc_nh3 = L_nh3 * sc, Lc_h2s = L_h2s * sc, Lc_nh4sh = L_nh4sh * sc;
                const double cM_nh3 = Lc_nh3 * (*X[1])[q] / rho_c;
                const double cM_h2s = Lc_h2s * (*X[0])[q] / rho_c;
                const double cM_nh4sh = Lc_nh4sh * (*X[2])[q] / rho_c;
                const double j_nh3 = p.m_nh3 / M_mix * (Lc_nh3 * dnh3_div - M_mix / p.m_nh3 * cM_nh3 * dnh3_div
                    - M_mix / p.m_h2s * cM_nh3 * dh2s_div - M_mix / p.m_nh4sh * cM_nh3 * dnh4sh_div) - jT_nh3;
                const double j_h2s = p.m_h2s / M_mix * (Lc_h2s * dh2s_div - M_mix / p.m_nh3 * cM_h2s * dnh3_div
                    - M_mix / p.m_h2s * cM_h2s * dh2s_div - M_mix / p.m_nh4sh * cM_h2s * dnh4sh_div) - jT_h2s;
                const double j_nh4sh = p.m_nh4sh / M_mix * (Lc_nh4sh * dnh4sh_div - M_mix / p.m_nh3 * cM_nh4sh * dnh3_div
                    - M_mix / p.m_h2s * cM_nh4sh * dh2s_div - M_mix / p.m_nh4sh * cM_nh4sh * dnh4sh_div) - jT_nh4sh;
                // laplacian_spherical (:752-776)
                auto lap = [&](const std::vector<double>& c){
                    const double rmsin = rm * std::max(std::sin(o.the[j]), floor_);
                    const double d2r = (c[o.id(i + 1, j, k)] - 2.0 * c[q] + c[o.id(i - 1, j, k)]) / (o.g.dr * o.g.dr) * 1.0;
                    const double d2t = (c[o.id(i, j + 1, k)] - 2.0 * c[q] + c[o.id(i, j - 1, k)]) / (o.g.dthe * o.g.dthe);
                    const double d2p = (c[o.id(i, j, ke)] - 2.0 * c[q] + c[o.id(i, j, kw)]) / (o.g.dphi * o.g.dphi);
                    const double d1r = (c[o.id(i + 1, j, k)] - c[o.id(i - 1, j, k)]) / (2.0 * o.g.dr) * 1.0;
                    const double d1t = (c[o.id(i, j + 1, k)] - c[o.id(i, j - 1, k)]) / (2.0 * o.g.dthe);
                    return d2r + 2.0 / rm * d1r + d2t / (rm * rm) + std::cos(o.the[j]) / (rm * rmsin) * d1t
                         + d2p / (rmsin * rmsin);
                };
                o.f[ATJUP_F_DIFFLUX_H2S][q] = DT_h2s * lap(*X[0]);
                o.f[ATJUP_F_DIFFLUX_NH3][q] = DT_nh3 * lap(*X[1]);
                o.f[ATJUP_F_DIFFLUX_NH4SH][q] = DT_nh4sh * lap(*X[2]);
                o.f[ATJUP_F_THERMALMASSFLUX][q] = (j_nh3 * p.cp_nh3 + j_h2s * p.cp_h2s + j_nh4sh * p.cp_nh4sh) * dt_div
                    + (p.enthalpy ? o.w_nh4sh[q] * p.dh_nh4sh
                                  : T_[q] * (o.w_nh3[q] * p.k_nh3 + o.w_h2s[q] * p.k_h2s + o.w_nh4sh[q] * p.k_nh4sh));
            }
        }
}

// ---- ChemistryJup::ChemMassRateJup (ChemistryJup.h:59-213), ATJUP_NH4SH_EQ 0 -------------------
void chemmassrate(orc_ctx& o){
    const atjup_chem_params& p = o.chemp;
    const int im = o.im, jm = o.jm, km = o.km;
    const double keq = p.m_nh4sh / (p.m_nh3 * p.m_h2s), A = 15000.0, B = -0.5, T_d = 3020.0;
    const bool eq_on = o.kn.nh4sh_eq != 0;
    std::vector<double>& MF_h2s = o.f[ATJUP_F_MASSFLUX_H2S];
    std::vector<double>& MF_nh3 = o.f[ATJUP_F_MASSFLUX_NH3];
    std::vector<double>& MF_nh4sh = o.f[ATJUP_F_MASSFLUX_NH4SH];
    for(int i = 0; i < im; i++)
        for(int k = 0; k < km; k++)
            for(int j = 0; j < jm; j++){
                if(j >= 3 && j < jm - 3) continue;
                const size_t q = o.id(i, j, k);
                o.w_nh3[q] = o.w_h2s[q] = o.w_nh4sh[q] = 0.0;
                MF_nh3[q] = MF_h2s[q] = MF_nh4sh[q] = 0.0;
            }
    #pragma omp parallel for schedule(static)
    for(int k = 0; k < km - 1; k++)
        for(int j = 3; j < jm - 3; j++)
            for(int i = 1; i < im - 1; i++){
                const size_t q = o.id(i, j, k);
                const int kw = o.wrap(k - 1), ke = o.wrap(k + 1);
                if(o.mask[q] == 1.0 && o.mask[o.id(i - 1, j, k)] == 1.0 && o.mask[o.id(i + 1, j, k)] == 1.0
                   && o.mask[o.id(i, j - 1, k)] == 1.0 && o.mask[o.id(i, j + 1, k)] == 1.0
                   && o.mask[o.id(i, j, kw)] == 1.0 && o.mask[o.id(i, j, ke)] == 1.0){
                    o.w_nh3[q] = o.w_h2s[q] = o.w_nh4sh[q] = 0.0;
                    MF_nh3[q] = MF_h2s[q] = MF_nh4sh[q] = 0.0;
                    continue;
                }
                const double t_u = o.f[ATJUP_F_T][q] * o.c.t_ref;
                const double kf = A * std::pow(t_u, B) * std::exp(-T_d / t_u);
                const double kb = kf / keq;
                // with the equilibrium adjustment the rate law contributes nothing (ChemistryJup.h:133-135)
                if(!eq_on && t_u <= p.t_0_nh4sh && t_u >= p.t_00_nh4sh){
                    const double nh3 = o.f[ATJUP_F_NH3][q], h2s = o.f[ATJUP_F_H2S][q], nh4sh = o.f[ATJUP_F_NH4SH][q];
                    double denom;
                    if(p.molar_conc){
                        double rho = o.f[ATJUP_F_RHO_MIX][q];
                        if(!(rho > 0.0)) rho = o.c.r_mix;
                        denom = rho;
                    } else {
                        const double sum_c = nh3 / p.m_nh3 + h2s / p.m_h2s + nh4sh / p.m_nh4sh;
                        denom = (sum_c == 0.0) ? 0.0 : o.c.r_mix / sum_c;
                    }
                    const double R_diff = kf * (nh3 / p.m_nh3 * denom) * (h2s / p.m_h2s * denom) - kb * (nh4sh / p.m_nh4sh * denom);
                    o.w_nh3[q] = -p.m_nh3 * R_diff;
                    o.w_h2s[q] = -p.m_h2s * R_diff;
                    o.w_nh4sh[q] = p.m_nh4sh * R_diff;
                } else if(p.gate_zero || eq_on){
                    o.w_nh3[q] = o.w_h2s[q] = o.w_nh4sh[q] = 0.0;
                }
                MF_h2s[q] = o.w_h2s[q] + o.f[ATJUP_F_DIFFLUX_H2S][q];
                MF_nh3[q] = o.w_nh3[q] + o.f[ATJUP_F_DIFFLUX_NH3][q];
                MF_nh4sh[q] = o.w_nh4sh[q] + o.f[ATJUP_F_DIFFLUX_NH4SH][q];
            }
}

// ---- ChemistryJup::EquilibriumNH4SH("loop") (ChemistryJup.h:293-370): every fluid cell brought to
//      the Lewis equilibrium p_NH3 p_H2S = K_p(T), K_p = 10^(14.82 - 4705 / T) atm^2, with the heat of
//      formation moving T; the extent x [kmol/m3] found by bisection on the monotone
//      f(x) = (a - x)(b - x)(R* T(x))^2 - K_p(T(x)). The printed census is not formed. ----------------
void nh4sh_eq(orc_ctx& o){
    const atjup_chem_params& p = o.chemp;
    const double R_univ = 8314.46, atm = 101325.0;
    const double L_f = nh4sh_formation_heat();
    auto Kp = [&](double T){ return std::pow(10.0, 14.82 - 4705.0 / T) * atm * atm; };
    auto& t = o.f[ATJUP_F_T];
    auto& nh3 = o.f[ATJUP_F_NH3];
    auto& h2s = o.f[ATJUP_F_H2S];
    auto& nh4sh = o.f[ATJUP_F_NH4SH];
    const auto& rho_mix = o.f[ATJUP_F_RHO_MIX];
    #pragma omp parallel for schedule(static)
    for(size_t q = 0; q < o.n3; q++){
        if(o.mask[q] == 1.0) continue;
        const double T0 = t[q] * o.c.t_ref;
        if(!(T0 > 0.0)) continue;
        double rho = rho_mix[q];
        if(!(rho > 0.0) || !std::isfinite(rho)) rho = o.c.r_mix;
        const double a = std::max(0.0, nh3[q]) / p.m_nh3;
        const double b = std::max(0.0, h2s[q]) / p.m_h2s;
        const double s = std::max(0.0, nh4sh[q]) / p.m_nh4sh;
        const double heat = L_f * p.m_nh4sh / (o.c.cp_mix * rho);
        auto f = [&](double x){
            const double T = T0 + heat * x, RT = R_univ * T;
            return (a - x) * (b - x) * RT * RT - Kp(T);
        };
        const double f0 = f(0.0);
        double lo, hi;
        if(f0 > 0.0){ lo = 0.0; hi = std::min(a, b); }
        else if(s > 0.0){
            if(f(-s) <= 0.0){ lo = hi = -s; }
            else { lo = -s; hi = 0.0; }
        } else continue;
        if(hi > lo)
            for(int it = 0; it < 80; it++){
                const double xm = 0.5 * (lo + hi);
                if(f(xm) > 0.0) lo = xm; else hi = xm;
                if(hi - lo <= 1.0e-15 * std::max(a, std::max(b, s))) break;
            }
        const double x = (f0 > 0.0) ? lo : hi;
        if(x == 0.0) continue;
        nh3[q] = std::max(0.0, nh3[q] - p.m_nh3 * x);
        h2s[q] = std::max(0.0, h2s[q] - p.m_h2s * x);
        nh4sh[q] = std::max(0.0, nh4sh[q] + p.m_nh4sh * x);
        t[q] = (T0 + heat * x) / o.c.t_ref;
    }
}

// the physics block's chemistry in the reference's order (cJupiterModel.cpp:1146-1161)
void chemistry(orc_ctx& o){
    diffmassflux(o);
    for(int f : { ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3, ATJUP_F_DIFFLUX_NH4SH, ATJUP_F_THERMALMASSFLUX })
        shapiro(o, o.f[f], 2, 1.0, 1, 7);
    chemmassrate(o);
    flux_limiter(o);
    for(int f : { ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH })
        shapiro(o, o.f[f], 2, 1.0, 1, 7);
}

// ---- MoistConvectiveAdjustment<P>::run (ConvectiveAdjustment.h:403-849) --------------------------
// Gases H2O and NH3 (gas[0], gas[1]); tracers H2S (ep_h2s, a vapour, no heat) and NH4SH (ep 0, a
// condensate, heat l_nh4sh); the chromophore tracer is zero with ATJUP_CHROM 0 and changes nothing.
// Settings: SMOOTH 0, PHASE 0, INTERIOR 1, cp(T) off (H(T) = cp_mix T). The census counters and
// the conservation diagnostics are print-only and not formed.
namespace mca_detail {
constexpr int NG = 2, NT = 2;
constexpr double TOL_K = 1.0e-7, TOL_T = 1.0e-10;      // :416-418
constexpr int N_ITER = 60;

struct Col {
    const orc_ctx& o;
    double cp, grav, ep_tr[NT], L_tr[NT];
    bool cond_tr[NT];
    // q_sat over the phase T selects, without the liquid/ice split (:430-445)
    double qsat(int g, double T, double p) const {
        const atjup_gas_params& s = o.gas[g];
        const double E = (T >= s.t_0) ? svp(T, s.C, s.L0, s.R, s.del_alf, s.del_bet)
                                      : svp(T, s.C_i, s.L0_i, s.R, s.del_alf_i, s.del_bet_i);
        const double q = s.ep * E / p;
        return std::isfinite(q) ? q : 0.0;
    }
    double lat(int g, double T) const { return T >= o.gas[g].t_0 ? o.gas[g].lv : o.gas[g].ls; }
    // the equilibrium temperature of (h_L, q_t) at (p, z) by Illinois regula falsi (:453-478)
    bool equil(double hL, const double* qt, double p, double z, double& T, double* cond) const {
        const double T_lo = (hL - grav * z) / cp;
        if(!(T_lo > 0.0) || !(p > 0.0)) return false;
        bool any = false;
        for(int g = 0; g < NG; g++) if(qt[g] > qsat(g, T_lo, p)){ any = true; break; }
        if(!any){ T = T_lo; for(int g = 0; g < NG; g++) cond[g] = 0.0; return true; }
        double span = 0.0;
        for(int g = 0; g < NG; g++) span += std::max(o.gas[g].lv, o.gas[g].ls) * std::max(0.0, qt[g]);
        auto fof = [&](double TT){
            double f = cp * TT + grav * z - hL;
            for(int g = 0; g < NG; g++) f -= lat(g, TT) * std::max(0.0, qt[g] - qsat(g, TT, p));
            return f;
        };
        double a = T_lo, b = T_lo + span / cp + 1.0e-9;
        double fa = fof(a), fb = fof(b);
        int side = 0;
        for(int it = 0; it < N_ITER && b - a > TOL_T; it++){
            double c = (fb - fa != 0.0) ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
            if(!(c > a && c < b)) c = 0.5 * (a + b);
            const double fc = fof(c);
            if(fc < 0.0){ a = c; fa = fc; if(side == -1) fb *= 0.5; side = -1; }
            else        { b = c; fb = fc; if(side == +1) fa *= 0.5; side = +1; }
            if(std::fabs(fc) <= cp * TOL_T){ a = b = c; break; }
        }
        T = b;
        for(int g = 0; g < NG; g++) cond[g] = std::max(0.0, qt[g] - qsat(g, T, p));
        return true;
    }
};
}  // namespace mca_detail

void mca(orc_ctx& o){
    using namespace mca_detail;
    const atjup_mca_params& P = o.mcap;
    const int im = o.im, jm = o.jm, km = o.km;
    const double relax = std::min(1.0, std::max(0.0, P.relax));
    const int max_pass = P.max_pass > 0 ? P.max_pass : 64;
    const bool ledoux = P.ledoux != 0, loading = P.loading != 0;
    static const int gf[NG][3] = { { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE },
                                   { ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE } };
    static const int tf[NT] = { ATJUP_F_H2S, ATJUP_F_NH4SH };
    Col C{ o, o.c.cp_mix, o.c.g, { P.ep_h2s, 0.0 }, { 0.0, P.l_nh4sh }, { false, true } };
    #pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for(int j = 0; j < jm; j++)
        for(int k = 0; k < km; k++){
            int i0 = 0;
            while(i0 < im && o.mask[o.id(i0, j, k)] == 1.0) i0++;
            int i1 = i0;
            while(i1 + 1 < im && o.mask[o.id(i1 + 1, j, k)] != 1.0) i1++;
            if(i0 == 0) i0 = 1;                               // MCA_INTERIOR (:566-575)
            if(i1 == im - 1) i1 = im - 2;
            if(i1 - i0 < 1) continue;
            std::vector<double> T(im), T0(im), p(im), z(im), rho(im), w(im), hL(im);
            std::vector<double> qt(im * NG), qv(im * NG), qc(im * NG), qi(im * NG), qr(im * NT);
            std::vector<double> Teq(im), ceq(im * NG);
            std::vector<char> touched(im, 0), eq_ok(im, 0), cen_u(im, 0);
            bool ok = true;
            for(int i = i0; i <= i1 && ok; i++){
                const size_t q = o.id(i, j, k);
                T[i] = T0[i] = o.f[ATJUP_F_T][q] * o.c.t_ref;
                p[i] = o.f[ATJUP_F_P_STAT][q];
                z[i] = o.get_layer_height(i) * 1.0e3;
                const double rf = o.f[ATJUP_F_RHO_MIX][q];
                rho[i] = (rf > 0.0 && std::isfinite(rf)) ? rf : o.c.r_mix;
                if(!(T[i] > 0.0) || !(p[i] > 0.0) || !(rho[i] > 0.0) || !std::isfinite(T[i] + p[i] + rho[i])) ok = false;
            }
            if(!ok) continue;
            for(int i = i0; i <= i1 && ok; i++){
                const size_t q = o.id(i, j, k);
                const double lo = (i > i0) ? 0.5 * (z[i - 1] + z[i]) : z[i0];
                const double hi = (i < i1) ? 0.5 * (z[i] + z[i + 1]) : z[i1];
                w[i] = rho[i] * (hi - lo);
                if(!(w[i] > 0.0)) ok = false;
                double h = C.cp * T[i] + C.grav * z[i];
                for(int g = 0; g < NG; g++){
                    const int x = i * NG + g;
                    qv[x] = std::max(0.0, o.f[gf[g][0]][q]) / rho[i];
                    qc[x] = std::max(0.0, o.f[gf[g][1]][q]) / rho[i];
                    qi[x] = std::max(0.0, o.f[gf[g][2]][q]) / rho[i];
                    qt[x] = qv[x] + qc[x] + qi[x];
                    h -= o.gas[g].lv * qc[x] + o.gas[g].ls * qi[x];
                }
                for(int r = 0; r < NT; r++){
                    qr[i * NT + r] = o.f[tf[r]][q] / rho[i];
                    h -= C.L_tr[r] * qr[i * NT + r];
                }
                hL[i] = h;
            }
            if(!ok) continue;
            auto trE = [&](int i){ double e = 0.0; for(int r = 0; r < NT; r++) e += C.L_tr[r] * qr[i * NT + r]; return e; };
            auto tv = [&](double TT, const double* v_, const double* cn, int i){
                if(!ledoux) return TT;
                double f = 1.0;
                for(int g = 0; g < NG; g++){ f += v_[g] * (1.0 / o.gas[g].ep - 1.0); if(loading) f -= cn[g]; }
                for(int r = 0; r < NT; r++){
                    const double qq = qr[i * NT + r];
                    if(C.ep_tr[r] > 0.0) f += qq * (1.0 / C.ep_tr[r] - 1.0);
                    else if(C.cond_tr[r] && loading) f -= qq;
                }
                return TT * f;
            };
            double pc[NG], pv[NG], ec[NG], ev[NG];
            // the parcel from a lifted to b against b in its own equilibrium (:651-673)
            auto unstable = [&](int a, int b) -> bool {
                double Tp;
                if(!C.equil(hL[a] + trE(a), &qt[a * NG], p[b], z[b], Tp, pc)) return false;
                if(!eq_ok[b]){
                    if(!C.equil(hL[b] + trE(b), &qt[b * NG], p[b], z[b], Teq[b], &ceq[b * NG])) return false;
                    eq_ok[b] = 1;
                }
                for(int g = 0; g < NG; g++){
                    ec[g] = ceq[b * NG + g];
                    pv[g] = qt[a * NG + g] - pc[g];
                    ev[g] = qt[b * NG + g] - ec[g];
                }
                return tv(Tp, pv, pc, a) - tv(Teq[b], ev, ec, b) > TOL_K;
            };
            auto settle = [&](int q) -> bool {                 // (:674-690)
                double TT;
                if(!C.equil(hL[q] + trE(q), &qt[q * NG], p[q], z[q], TT, pc)) return false;
                T[q] = TT; Teq[q] = TT; eq_ok[q] = 1;
                for(int g = 0; g < NG; g++){
                    const int x = q * NG + g;
                    ceq[x] = pc[g];
                    qv[x] = qt[x] - pc[g];
                    if(TT >= o.gas[g].t_0){ qc[x] = pc[g]; qi[x] = 0.0; } else { qi[x] = pc[g]; qc[x] = 0.0; }
                }
                touched[q] = 1;
                return true;
            };
            for(int i = i0; i < i1; i++) cen_u[i] = unstable(i, i + 1);      // the census (:700-706)
            const std::vector<double> hL0(hL), qt0(qt), qr0(qr);
            int passes = 0; long long layers = 0; bool changed = true, fail = false;
            while(changed && passes < max_pass && !fail){                       // (:716-756)
                changed = false; passes++;
                int i = i0;
                while(i < i1 && !fail){
                    const bool uns = (passes == 1 && !touched[i] && !touched[i + 1]) ? (bool)cen_u[i] : unstable(i, i + 1);
                    if(!uns){ i++; continue; }
                    int a = i, b = i + 1;
                    for(;;){
                        double sw = 0.0, sh = 0.0, sq[NG] = { 0.0, 0.0 }, sr[NT] = { 0.0, 0.0 };
                        for(int q = a; q <= b; q++){
                            sw += w[q]; sh += w[q] * hL[q];
                            for(int g = 0; g < NG; g++) sq[g] += w[q] * qt[q * NG + g];
                            for(int r = 0; r < NT; r++) sr[r] += w[q] * qr[q * NT + r];
                        }
                        for(int q = a; q <= b; q++){
                            hL[q] = sh / sw;
                            for(int g = 0; g < NG; g++) qt[q * NG + g] = sq[g] / sw;
                            for(int r = 0; r < NT; r++) qr[q * NT + r] = sr[r] / sw;
                            if(!settle(q)){ fail = true; break; }
                        }
                        if(fail) break;
                        bool ext = false;
                        if(a > i0 && unstable(a - 1, a)){ a--; ext = true; }
                        if(b < i1 && unstable(b, b + 1)){ b++; ext = true; }
                        if(!ext) break;
                    }
                    layers += (b - a + 1);
                    changed = true;
                    i = b;
                }
            }
            if(fail || layers == 0) continue;                 // the column keeps its state
            if(relax < 1.0){                                  // MCA_RELAX (:800-811)
                for(int q = i0; q <= i1; q++){
                    if(!touched[q]) continue;
                    hL[q] = hL0[q] + relax * (hL[q] - hL0[q]);
                    for(int g = 0; g < NG; g++){ const int x = q * NG + g; qt[x] = qt0[x] + relax * (qt[x] - qt0[x]); }
                    for(int r = 0; r < NT; r++){ const int x = q * NT + r; qr[x] = qr0[x] + relax * (qr[x] - qr0[x]); }
                    if(!settle(q)){ fail = true; break; }
                }
                if(fail) continue;
            }
            for(int q = i0; q <= i1; q++){                    // (:822-833)
                if(!touched[q]) continue;
                const size_t c = o.id(q, j, k);
                o.f[ATJUP_F_T][c] = T[q] / o.c.t_ref;
                for(int g = 0; g < NG; g++){
                    const int x = q * NG + g;
                    o.f[gf[g][0]][c] = qv[x] * rho[q];
                    o.f[gf[g][1]][c] = qc[x] * rho[q];
                    o.f[gf[g][2]][c] = qi[x] * rho[q];
                }
                for(int r = 0; r < NT; r++) o.f[tf[r]][c] = qr[q * NT + r] * rho[q];
            }
        }
}

bool valid_field(int f){
    return (f >= 0 && f < ATJUP_N_PROG) || (f >= ATJUP_F_P_DYN && f < ATJUP_F_END)
        || (f >= ATJUP_NLEVEL && f < ATJUP_NLEVEL + ATJUP_N_PROG);
}

}  // namespace

extern "C" {

void orc_set_threads(int n){
#ifdef _OPENMP
    if(n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

orc_ctx* orc_create(const atjup_grid* g, const atjup_consts* c, const atjup_knobs* k){
    if(!g || !c || !k || g->im < 5 || g->jm < 8 || g->km < 6) return nullptr;
    orc_ctx* o = new orc_ctx();
    o->g = *g; o->c = *c; o->kn = *k;
    o->im = g->im; o->jm = g->jm; o->km = g->km;
    o->n3 = (size_t)o->im * o->jm * o->km;
    o->rad.resize(o->im);
    o->the.resize(o->jm);
    double z = g->r0;
    o->rad[0] = z;
    for(int i = 1; i < o->im; i++){ z = z + g->dr; o->rad[i] = z; }
    z = 0.0;
    o->the[0] = z;
    for(int j = 1; j < o->jm; j++){ z = z + g->dthe; o->the[j] = z; }
    o->sin_the.resize(o->jm); o->cos_the.resize(o->jm); o->sinthe_rk.resize(o->jm);
    for(int j = 0; j < o->jm; j++){
        o->sin_the[j] = std::sin(o->the[j]);
        o->cos_the[j] = std::cos(o->the[j]);
        o->sinthe_rk[j] = std::max(k->sinthe_min, std::abs(std::sin(o->the[j])));
    }
    const float h = (float)(g->L_atm * g->dr);
    o->lh.resize(o->im);
    for(int i = 0; i < o->im; i++) o->lh[i] = i * h;
    o->f.assign(ATJUP_NLEVEL + ATJUP_N_PROG, {});
    for(int fi = 0; fi < ATJUP_NLEVEL + ATJUP_N_PROG; fi++)
        if(valid_field(fi)) o->f[fi].assign(o->n3, 0.0);
    o->f[ATJUP_F_DIS].assign(o->n3, 1.0e-10);
    o->f[ATJUP_NLEVEL + ATJUP_F_DIS].assign(o->n3, 1.0e-10);
    o->f[ATJUP_F_RHO_MIX].assign(o->n3, 1.0);
    o->rhs.assign(NP, std::vector<double>(o->n3, 0.0));
    o->acc.assign(NP, std::vector<double>(o->n3, 0.0));
    o->mask.assign(o->n3, 0.0);
    o->itopo.assign((size_t)o->jm * o->km, 0);
    o->vel_star.assign((size_t)o->jm * o->km, 0.0);
    o->brl.assign(o->im, 0.0);
    o->w_nh3.assign(o->n3, 0.0); o->w_h2s.assign(o->n3, 0.0); o->w_nh4sh.assign(o->n3, 0.0);
    // Jupiter (cJupiterModel.h:1109-1118) and the radiation knobs' defaults (Radiation.h:270-295)
    o->radp = atjup_rad_params{ 5.4, 50.5, 0.343, 0.863, 0.134, 3.9, 1.0, 1.0, 1.0, 1, 0 };
    // with the NH4SH equilibrium the moist convective adjustment carries its heat of formation
    // (cJupiterModel.cpp:1316)
    if(k->nh4sh_eq) o->mcap.l_nh4sh = nh4sh_formation_heat();
    return o;
}

void orc_destroy(orc_ctx* o){ delete o; }

int orc_upload_field(orc_ctx* o, int f, const double* src){
    if(!o || !src || !valid_field(f)) return ATJUP_ERR_ARG;
    std::memcpy(o->f[f].data(), src, sizeof(double) * o->n3);
    return 0;
}
int orc_download_field(orc_ctx* o, int f, double* dst){
    if(!o || !dst || !valid_field(f)) return ATJUP_ERR_ARG;
    std::memcpy(dst, o->f[f].data(), sizeof(double) * o->n3);
    return 0;
}
int orc_upload_mask(orc_ctx* o, const double* m){
    if(!o || !m) return ATJUP_ERR_ARG;
    std::memcpy(o->mask.data(), m, sizeof(double) * o->n3);
    for(int j = 0; j < o->jm; j++)
        for(int k = 0; k < o->km; k++){
            int s = 0;
            while(s < o->im && o->solid(s, j, k)) ++s;
            o->itopo[(size_t)j * o->km + k] = s;
        }
    wall_viscosity(*o);
    return 0;
}
int orc_upload_vel_star(orc_ctx* o, const double* s){
    if(!o || !s) return ATJUP_ERR_ARG;
    std::memcpy(o->vel_star.data(), s, sizeof(double) * o->jm * o->km);
    return 0;
}
int orc_download_vel_star(orc_ctx* o, double* d){
    if(!o || !d) return ATJUP_ERR_ARG;
    std::memcpy(d, o->vel_star.data(), sizeof(double) * o->jm * o->km);
    return 0;
}
int orc_download_buoy_ref_level(orc_ctx* o, double* d){
    if(!o || !d) return ATJUP_ERR_ARG;
    std::memcpy(d, o->brl.data(), sizeof(double) * o->im);
    return 0;
}
int orc_rk4_step(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; rk4(*o); return 0; }
int orc_buoy_ref_level(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; buoy_ref_level(*o); return 0; }
int orc_hydrostatic(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; buoy_ref_level(*o); hydrostatic(*o); return 0; }
int orc_pressure_solve(orc_ctx* o, int sweeps, double omega){
    if(!o) return ATJUP_ERR_ARG;
    pressure(*o, sweeps > 0 ? sweeps : std::max(1, o->kn.press_sweeps), omega);
    return 0;
}
// One colour pass of the sweep alone (after prepareProjectionBoundaries, which is idempotent), then
// the outer planes and the seam if `planes`: a sweep is pass 0, pass 1 with planes.
int orc_rb_colour_pass(orc_ctx* o, int colour, double omega, int planes){
    if(!o || (colour != 0 && colour != 1)) return ATJUP_ERR_ARG;
    prep_projection(*o);
    rb_colour_pass(*o, colour, omega);
    if(planes) rb_planes(*o);
    return 0;
}
int orc_bc_radius(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; bc_radius(*o); return 0; }
int orc_bc_theta(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; bc_theta(*o); return 0; }
int orc_bc_phi(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; bc_phi(*o); return 0; }
int orc_bc_solid(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; bc_solid(*o); return 0; }
int orc_bc_after_step(orc_ctx* o){
    if(!o) return ATJUP_ERR_ARG;
    bc_radius(*o); bc_theta(*o); bc_phi(*o); bc_solid(*o);
    return 0;
}
int orc_shapiro(orc_ctx* o, int f, int order, double strength, int passes, int axes){
    if(!o || !valid_field(f) || (order != 2 && order != 4)) return ATJUP_ERR_ARG;
    shapiro(*o, o->f[f], order, strength, passes, axes);
    return 0;
}
int orc_mixture_density(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; mixture_density(*o); return 0; }
int orc_clamp_species(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; clamp_species(*o); return 0; }
int orc_swap_nlevel(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; restore(*o); return 0; }
int orc_scalar_sync(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; scalar_sync(*o); return 0; }
int orc_set_gas(orc_ctx* o, int gas, const atjup_gas_params* p){
    if(!o || !p || gas < 0 || gas > 2) return ATJUP_ERR_ARG;
    o->gas[gas] = *p;
    return 0;
}
int orc_satadj(orc_ctx* o, int gas){
    if(!o || gas < 0 || gas > 2) return ATJUP_ERR_ARG;
    satadj(*o, gas);
    return 0;
}
int orc_turbulence_run(orc_ctx* o){
    if(!o) return ATJUP_ERR_ARG;
    if(o->kn.turb_model < 0) return 0;
    if(o->kn.turb_model > 2) return ATJUP_ERR_PROFILE;
    turbulence(*o);
    turb_wall_omega(*o);
    return 0;
}
int orc_turb_wall(orc_ctx* o){
    if(!o) return ATJUP_ERR_ARG;
    if(o->kn.turb_model < 0) return 0;
    turb_apply_wall_bc(*o);
    turb_wall_omega(*o);
    return 0;
}
int orc_flux_limiter(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; flux_limiter(*o); return 0; }
int orc_set_radiation(orc_ctx* o, const atjup_rad_params* p){
    if(!o || !p) return ATJUP_ERR_ARG;
    o->radp = *p;
    return 0;
}
int orc_radiation(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; radiation(*o); return 0; }
int orc_set_chemistry(orc_ctx* o, const atjup_chem_params* p){
    if(!o || !p) return ATJUP_ERR_ARG;
    o->chemp = *p;
    return 0;
}
int orc_chemistry(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; chemistry(*o); return 0; }
int orc_nh4sh_eq(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; nh4sh_eq(*o); return 0; }
int orc_set_mca(orc_ctx* o, const atjup_mca_params* p){
    if(!o || !p) return ATJUP_ERR_ARG;
    o->mcap = *p;
    return 0;
}
int orc_mca(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; mca(*o); return 0; }
int orc_set_precipitation(orc_ctx* o, const atjup_precip_params* p){
    if(!o || !p) return ATJUP_ERR_ARG;
    o->precp = *p;
    return 0;
}
int orc_precipitation(orc_ctx* o){ if(!o) return ATJUP_ERR_ARG; precipitation(*o); return 0; }
int orc_download_precip_surface(orc_ctx* o, double* dst){
    if(!o || !dst) return ATJUP_ERR_ARG;
    if(o->precip_srf.empty()) o->precip_srf.assign((size_t)5 * o->jm * o->km, 0.0);
    std::memcpy(dst, o->precip_srf.data(), o->precip_srf.size() * sizeof(double));
    return 0;
}

// One iteration of the loop body, on-path functions only (cJupiterModel.cpp:1010-1555).
int orc_step(orc_ctx* o, int iter, int first){
    if(!o) return ATJUP_ERR_ARG;
    if(iter % 2 == 0 || first){
        pressure(*o, std::max(1, o->kn.press_sweeps), o->kn.press_omega);
        if(o->kn.damp_pdyn) shapiro(*o, o->f[ATJUP_F_P_DYN], 2, 1.0, 1, 7);
        mixture_density(*o);
        if(o->kn.radiation) radiation(*o);                    // cJupiterModel.cpp:1056
        if(o->kn.satadj) for(int gas = 0; gas < 3; gas++) satadj(*o, gas);
        if(o->kn.nh4sh_eq) nh4sh_eq(*o);                      // cJupiterModel.cpp:1084
        if(o->kn.precip) precipitation(*o);                   // cJupiterModel.cpp:1090
        if(o->kn.rk_scalar_sync >= 1) scalar_sync(*o);
        if(o->kn.turb_model >= 0){ turbulence(*o); turb_wall_omega(*o); }
        if(o->kn.chemistry) chemistry(*o);                    // cJupiterModel.cpp:1146-1161
        else {
            flux_limiter(*o);
            shapiro(*o, o->f[ATJUP_F_FLUXLIM_NH4SH], 2, 1.0, 1, 7);
        }
    }
    rk4(*o);
    bc_radius(*o); bc_theta(*o); bc_phi(*o); bc_solid(*o);
    if(o->kn.turb_model >= 0){ turb_apply_wall_bc(*o); turb_wall_omega(*o); }
    if(o->kn.mca > 0 && iter % o->kn.mca == 0) mca(*o);  // cJupiterModel.cpp:1301-1322
    clamp_species(*o);
    restore(*o);
    return 0;
}

int orc_check_finite(orc_ctx* o){
    if(!o) return ATJUP_ERR_ARG;
    for(int fi = 0; fi < NP; fi++)
        for(double v : o->f[fi]) if(!std::isfinite(v)) return ATJUP_NONFINITE;
    return 0;
}

}  // extern "C"
