// C ABI of libatjup_hip.so (include/atjup_hip.h): context, transfers, operators.
#include "atjup_internal.h"

#include <cstdio>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace atjup {

static thread_local std::string g_err;
void set_error(const std::string& msg){ g_err = msg; }

namespace {

bool prog(int f){ return f >= 0 && f < ATJUP_N_PROG; }
bool input(int f){ return f >= ATJUP_F_P_DYN && f < ATJUP_F_END; }
bool nlevel(int f){ return f >= ATJUP_NLEVEL && f < ATJUP_NLEVEL + ATJUP_N_PROG; }
// The time-level-n copy is the live field on this path (restoreVar is the identity, see the
// context comment), so an n-level id addresses the same storage.
int storage_of(int f){ return nlevel(f) ? f - ATJUP_NLEVEL : f; }

int alloc(atjup::Buf& b, size_t bytes){
    if(hipMalloc(&b.p, bytes) != hipSuccess){ set_error("hipMalloc failed"); return ATJUP_ERR_NOMEM; }
    b.bytes = bytes;
    if(hipMemset(b.p, 0, bytes) != hipSuccess){ set_error("hipMemset failed"); return ATJUP_ERR_HIP; }
    return 0;
}
void release(atjup::Buf& b){ if(b.p) (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }

int fill_value(atjup_ctx* c, atjup::Buf& b, double v){
    const size_t n = c->n3;
    if(c->precision == ATJUP_PRECISION_TANGENT){
        std::vector<double> h(2 * n, 0.0);
        for(size_t q = 0; q < n; q++) h[2 * q] = v;
        ATJUP_HIP_CHECK(hipMemcpy(b.p, h.data(), n * 16, hipMemcpyHostToDevice));
    } else if(c->precision == ATJUP_PRECISION_FP64){
        std::vector<double> h(n, v);
        ATJUP_HIP_CHECK(hipMemcpy(b.p, h.data(), n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> h(n, (float)v);
        ATJUP_HIP_CHECK(hipMemcpy(b.p, h.data(), n * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

// Host geometry, built as the reference builds it (lib/Array_1D.cpp:49-60,
// RungeKutta_Jup_Turb.cpp:299-303 and :390-411, PressureSolver.h:156-162 and :327-353).
template<typename T>
void build_tables(atjup_ctx* c, std::vector<GeoIJ<T>>& geo, std::vector<PoisIJ<T>>& pois){
    const int im = c->im, jm = c->jm;
    c->rad.resize(im); c->the.resize(jm); c->sin_the.resize(jm); c->cos_the.resize(jm);
    double z = c->g.r0;
    c->rad[0] = z;
    for(int i = 1; i < im; i++){ z = z + c->g.dr; c->rad[i] = z; }
    z = 0.0;
    c->the[0] = z;
    for(int j = 1; j < jm; j++){ z = z + c->g.dthe; c->the[j] = z; }
    for(int j = 0; j < jm; j++){ c->sin_the[j] = std::sin(c->the[j]); c->cos_the[j] = std::cos(c->the[j]); }
    const float h = (float)(c->g.L_atm * c->g.dr);
    c->lh.resize(im);
    for(int i = 0; i < im; i++) c->lh[i] = i * h;
    const double omega = c->c.omega;
    const double inv_dr2 = 1.0 / (c->g.dr * c->g.dr), inv_dthe2 = 1.0 / (c->g.dthe * c->g.dthe);
    const double inv_dphi2 = 1.0 / (c->g.dphi * c->g.dphi);
    geo.resize((size_t)im * jm);
    pois.resize((size_t)im * jm);
    for(int i = 0; i < im; i++)
        for(int j = 0; j < jm; j++){
            const double rm = c->rad[i];
            const double rm2 = rm * rm;
            const double sinthe = std::max(c->kn.sinthe_min, std::abs(std::sin(c->the[j])));
            const double sinthe2 = sinthe * sinthe;
            const double costhe = c->cos_the[j];
            const double inv_rm2 = 1.0 / rm2;
            GeoIJ<T>& G = geo[(size_t)i * jm + j];
            const double st = std::sqrt(std::max(0.0, 1.0 - costhe * costhe));
            G.rm = (T)rm;
            G.inv_rm = (T)(1.0 / rm);
            G.inv_rm2 = (T)inv_rm2;
            G.inv_rmsinthe = (T)(1.0 / (rm * sinthe));
            G.inv_rm2sinthe = (T)(inv_rm2 / sinthe);
            G.inv_rm2sinthe2 = (T)(inv_rm2 / sinthe2);
            G.costhe_inv_rm2sinthe = (T)(costhe * (inv_rm2 / sinthe));
            G.costhe = (T)costhe;
            G.cotanthe = (T)(costhe / sinthe);
            G.v_metric = (T)((1.0 + costhe / (sinthe2)) * inv_rm2);
            G.sinthe_true = (T)st;
            G.cent_rad = (T)(1.0 * omega * omega * rm * st * st);
            G.cent_the = (T)(1.0 * omega * omega * rm * st * costhe);
            G.pad0 = G.pad1 = G.pad2 = (T)0;
            // Poisson: sinthe from the solver's own table (PressureSolver.h:156-162)
            double sp = std::sin(c->the[j]);
            if(sp < c->kn.sinthe_min) sp = c->kn.sinthe_min;
            if(sp == 0.0) sp = 1.0e-5;
            const double sp2 = sp * sp;
            const double exp_2_rm = 1.0 * 1.0;
            PoisIJ<T>& W = pois[(size_t)i * jm + j];
            const double num1 = exp_2_rm * inv_dr2;
            const double num2 = inv_rm2 * inv_dthe2;
            const double num3 = (inv_rm2 / sp2) * inv_dphi2;
            W.num1 = (T)num1; W.num2 = (T)num2; W.num3 = (T)num3;
            W.denom = (T)(2.0 * num1 + 2.0 * num2 + 2.0 * num3);
            W.inv_rm = (T)(1.0 / rm);
            W.inv_rmsinthe = (T)(1.0 / (rm * sp));
            double dd = 0.0;
            for(double nv : { num1, num1, num2, num2, num3, num3 }) dd += nv;
            float df = 0.0f;
            for(double nv : { num1, num1, num2, num2, num3, num3 }) df += (float)nv;
            W.den_fluid = (T)dd;
            W.den_fluid32 = (T)df;
        }
}

int validate(const atjup_grid* g, const atjup_consts* c, const atjup_knobs* k, int precision){
    if(!g || !c || !k){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(g->im < 6 || g->jm < 9 || g->km < 8){ set_error("grid too small"); return ATJUP_ERR_ARG; }
    if(g->km > 1024){ set_error("km above 1024"); return ATJUP_ERR_ARG; }
    if((g->km - 1) % 2 != 0){ set_error("the periodic ring km-1 must be even for a consistent red-black colouring"); return ATJUP_ERR_ARG; }
    if(precision != ATJUP_PRECISION_FP64 && precision != ATJUP_PRECISION_FP32_REF && precision != ATJUP_PRECISION_TANGENT){
        set_error("precision"); return ATJUP_ERR_ARG;
    }
    if(precision == ATJUP_PRECISION_TANGENT && !k->dry){
        set_error("the tangent form covers the dynamics profile (dry = 1)"); return ATJUP_ERR_PROFILE;
    }
    if(k->turb_model < -1 || k->turb_model > 2){ set_error("turb_model"); return ATJUP_ERR_PROFILE; }
    if(k->rk_scalar_sync != 2){ set_error("only ATJUP_RK_SCALAR_SYNC=2 is implemented"); return ATJUP_ERR_PROFILE; }
    if(k->dry && k->turb_model != -1){ set_error("the dry profile has the closure off"); return ATJUP_ERR_PROFILE; }
    if(!(k->pdyn_cap > 0.0)){ set_error("pdyn_cap"); return ATJUP_ERR_PROFILE; }
    if(!(k->press_omega > 0.0 && k->press_omega < 2.0)){ set_error("press_omega outside (0, 2)"); return ATJUP_ERR_PROFILE; }
    if(k->radiation != 0 && k->radiation != 1){ set_error("radiation must be 0 or 1"); return ATJUP_ERR_PROFILE; }
    if(k->precip != 0 && k->precip != 1){ set_error("precip must be 0 or 1"); return ATJUP_ERR_PROFILE; }
    if(k->chemistry != 0 && k->chemistry != 1){ set_error("chemistry must be 0 or 1"); return ATJUP_ERR_PROFILE; }
    if(k->mca < 0){ set_error("mca (the adjustment's stride) must be >= 0"); return ATJUP_ERR_PROFILE; }
    if(k->mca > 0 && (k->dry || precision == ATJUP_PRECISION_TANGENT)){
        set_error("the moist convective adjustment needs the moist profile (dry = 0)"); return ATJUP_ERR_PROFILE;
    }
    if(k->nh4sh_eq != 0 && k->nh4sh_eq != 1){ set_error("nh4sh_eq must be 0 or 1"); return ATJUP_ERR_PROFILE; }
    if(k->nh4sh_eq && (k->dry || precision == ATJUP_PRECISION_TANGENT)){
        set_error("the NH4SH equilibrium needs the moist profile (dry = 0)"); return ATJUP_ERR_PROFILE;
    }
    if(k->chemistry && k->dry){
        set_error("the chemistry needs the moist profile (dry = 0)"); return ATJUP_ERR_PROFILE;
    }
    if(k->chemistry && precision == ATJUP_PRECISION_TANGENT){
        set_error("the tangent form has no chemistry"); return ATJUP_ERR_PROFILE;
    }
    if(k->precip && precision == ATJUP_PRECISION_TANGENT){
        set_error("the tangent form has no precipitation"); return ATJUP_ERR_PROFILE;
    }
    if(k->radiation && precision == ATJUP_PRECISION_TANGENT){
        set_error("the tangent form has no radiation"); return ATJUP_ERR_PROFILE;
    }
    return 0;
}

}  // namespace

hipEvent_t timer_begin(atjup_ctx* c){
    if(!c->timing) return nullptr;
    hipEvent_t e = nullptr;
    if(!c->event_pool.empty()){ e = c->event_pool.back(); c->event_pool.pop_back(); }
    else if(hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, c->stream);
    return e;
}

void timer_end(atjup_ctx* c, int phase, hipEvent_t begin, int launches){
    if(!c->timing || !begin) return;
    hipEvent_t e = timer_begin(c);
    if(!e) return;
    c->spans[phase].push_back({ begin, e, launches });
    if(c->spans[phase].size() > 4096){                  // fold old spans so the pool stays bounded
        (void)hipEventSynchronize(e);
        for(auto& sp : c->spans[phase]){
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, sp.a, sp.b);
            c->t_total[phase] += ms;
            c->t_launch[phase] += sp.launches;
            c->event_pool.push_back(sp.a);
            c->event_pool.push_back(sp.b);
        }
        c->spans[phase].clear();
    }
}

// the physics block of an iteration (cJupiterModel.cpp:1032-1161), on even iterations and the first
// the radiation's host tables: F_sun_toa per row j (Radiation.h:374-380, the diurnal mean
// S (1 - A) cos(lat) / pi with colatitude j pi / (jm - 1)) and the layer thickness in metres each
// layer's heating divides by (Radiation.h:451-452: the top layer takes the one below's; <= 0 -> 1)
int build_rad_table(atjup_ctx* c){
    const atjup_rad_params& R = c->radp;
    const double pi_ = 3.14159265358979323846;
    std::vector<double> tab((size_t)c->jm + c->im);
    for(int j = 0; j < c->jm; j++){
        double f = 0.0;
        if(R.solar != 0){
            const double colat = pi_ * (double)j / (double)(c->jm - 1);
            const double cos_lat = std::max(0.0, std::sin(colat));
            f = R.solar_strength * R.S_solar * (1.0 - R.albedo_bond) * cos_lat / pi_;
        }
        tab[j] = f;
    }
    auto ltm = [&](int i){ return (i < 0 || i > c->im - 2) ? -1.0 : (double)(c->lh[i + 1] - c->lh[i]) * 1.0e3; };
    for(int i = 0; i < c->im; i++){
        double dz = (i < c->im - 1) ? ltm(i) : ltm(i - 1);
        if(dz <= 0.0) dz = 1.0;
        tab[c->jm + i] = dz;
    }
    if(!c->rad_tab.p){
        if(int rc = alloc(c->rad_tab, tab.size() * sizeof(double))) return rc;
    }
    ATJUP_HIP_CHECK(hipMemcpy(c->rad_tab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int step_block(atjup_ctx* c, int iter, int first){
    int rc = 0;
    if(!(iter % 2 == 0 || first)) return 0;
    hipEvent_t t0 = timer_begin(c);
    if((rc = launch_pressure(c, std::max(1, c->kn.press_sweeps), c->kn.press_omega))) return rc;
    timer_end(c, ATJUP_T_PRESSURE, t0, 0);
    if(c->kn.damp_pdyn && (rc = launch_shapiro(c, ATJUP_F_P_DYN, 2, 1.0, 1, 7))) return rc;
    if((rc = launch_mixture_density(c))) return rc;
    if(c->kn.radiation && (rc = launch_radiation(c))) return rc;      // cJupiterModel.cpp:1056
    if(!c->kn.dry){
        // the three adjustments, then RK_SCALAR_SYNC 2 (the identity on the shared storage),
        // the closure with its wall function, the NH4SH limiter and its smoothing
        // (cJupiterModel.cpp:1060-1161)
        if(c->kn.satadj && c->gas_set[0] && c->gas_set[1] && c->gas_set[2] && c->have_mask){
            if((rc = launch_satadj3(c))) return rc;          // the three in one pass
        } else {
            for(int gas = 0; gas < 3 && c->kn.satadj; gas++)
                if((rc = atjup_hip_satadj(c, gas, nullptr))) return rc;
        }
        if(c->kn.nh4sh_eq && (rc = launch_nh4sh_eq(c))) return rc;           // cJupiterModel.cpp:1084
        if(c->kn.precip && (rc = atjup_hip_precipitation(c))) return rc;      // cJupiterModel.cpp:1090
        if((rc = atjup_hip_turbulence_run(c))) return rc;
        if(c->kn.chemistry){
            if((rc = atjup_hip_chemistry(c))) return rc;                // :1146-1161, limiter included
        } else {
            if((rc = launch_flux_limiter(c))) return rc;
            if((rc = launch_shapiro(c, ATJUP_F_FLUXLIM_NH4SH, 2, 1.0, 1, 7))) return rc;
        }
    }
    // steadyQuery's n level of t and the condensables (NH4SH with the equilibrium) is the state
    // RK_SCALAR_SYNC 2 copies after the adjustments (cJupiterModel.cpp:1118-1140); the closure and
    // the chemistry after it change none of them
    if(c->steady_on){
        unsigned m = 0;
        for(int e = 4; e <= 10; e++) m |= 1u << e;           // t, h2o, h2o_cloud, h2o_ice, nh3, nh3_cloud, nh3_ice
        if(c->kn.nh4sh_eq) m |= 1u << 11;
        if((rc = steady_mark(c, m))) return rc;
    }
    return 0;
}

// the rest of the iteration: RungeKuttaJup, the boundary passes (cJupiterModel.cpp:1263-1272), the
// closure's wall hold (:1294), the moist convective adjustment every kn.mca iterations (:1301),
// the species clamp and the commit of the live k*, omega*
int step_rest(atjup_ctx* c, int iter){
    int rc = 0;
    if((rc = launch_rk4(c))) return rc;
    if((rc = launch_bc_after_step(c))) return rc;
    if(!c->kn.dry && c->kn.turb_model >= 0 && (rc = launch_turb_wall(c))) return rc;
    if(c->kn.mca > 0 && iter % c->kn.mca == 0 && (rc = atjup_hip_mca(c))) return rc;   // :1301-1322
    if((rc = launch_clamp_species(c))) return rc;
    return launch_commit_live(c);
}

}  // namespace atjup

using namespace atjup;

extern "C" {

const char* atjup_hip_last_error_string(void){ return g_err.c_str(); }

void atjup_hip_default_knobs(atjup_knobs* k){
    if(!k) return;
    std::memset(k, 0, sizeof(*k));
    k->turb_model = -1;
    k->press_sweeps = 1;
    k->press_omega = 1.0;
    k->pdyn_cap = 0.2;
    k->damp_pdyn = 1;
    k->rho_pdyn = 1.0;
    k->sinthe_min = 0.55;
    k->t_diff_scale = 0.01;
    k->species_diff_scale = 0.01;
    k->rad_coupling = 1.0;
    k->precip_coupling = 1.0;
    k->turb_coupling = 1.0;
    k->tke_max_phys = 1000.0;
    k->nue_max_phys = 1.0e5;
    k->wall_nue_factor = 4.0;
    k->wall_nue_layers = 3;
    k->satadj = 1;
    k->rk_scalar_sync = 2;
    k->dry = 1;
    k->turb_k_bg_phys = 1.0;
}

atjup_ctx* atjup_hip_ctx_create(const atjup_grid* g, const atjup_consts* cs, const atjup_knobs* kn, int precision){
    if(validate(g, cs, kn, precision) != 0) return nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1){ set_error("no HIP device"); return nullptr; }
    atjup_ctx* c = new atjup_ctx();
    c->g = *g; c->c = *cs; c->kn = *kn; c->precision = precision;
    c->im = g->im; c->jm = g->jm; c->km = g->km;
    c->esz = (precision == ATJUP_PRECISION_FP64) ? 8 : (precision == ATJUP_PRECISION_FP32_REF ? 4 : 16);
    const size_t asz = (precision == ATJUP_PRECISION_TANGENT) ? 16 : 8;      // FP64 quantities
    const int per128 = 128 / (int)c->esz;
    c->pitch = (c->km + per128 - 1) / per128 * per128;
    c->plane = (size_t)c->jm * c->pitch;
    c->n3 = (size_t)c->im * c->plane;
    if(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess){
        set_error("hipStreamCreate failed"); delete c; return nullptr;
    }
    c->stream = c->own_stream;
    const size_t fb = c->n3 * c->esz;
    int rc = 0;
    const bool dry = kn->dry != 0;
    for(int f = 0; f < ATJUP_F_END && !rc; f++){
        if(!(prog(f) || input(f))) continue;
        rc = alloc(c->cur[f], fb);
    }
    const int np = dry ? 4 : ATJUP_N_PROG;
    for(int f = 0; f < np && !rc; f++){
        rc = alloc(c->stg[0][f], fb);
        if(!rc) rc = alloc(c->stg[1][f], fb);
        if(!rc) rc = alloc(c->acc[f], fb);
    }
    if(!rc && !dry){ rc = alloc(c->live_turb[0], fb); if(!rc) rc = alloc(c->live_turb[1], fb); }
    if(!rc) rc = alloc(c->scratch, fb);
    c->rb_hp = ((c->km + 1) / 2 + 15) / 16 * 16;
    const size_t hn = (size_t)c->im * c->jm * c->rb_hp;
    if(!rc) rc = alloc(c->psrc, 2 * hn * asz);
    if(!rc) rc = alloc(c->rb_p, 2 * hn * asz);
    if(!rc && precision == ATJUP_PRECISION_FP32_REF){
        const int n_lo = c->kn.dry ? 4 : ATJUP_F_TKE;               // t, u, v, w (+ the species)
        for(int f = 0; f < n_lo && !rc; f++) rc = alloc(c->lo[f], c->n3 * sizeof(float));
    }
    if(!rc && precision == ATJUP_PRECISION_FP32_REF) rc = alloc(c->rad_tmp, 3 * c->n3 * sizeof(double));
    if(!rc && precision == ATJUP_PRECISION_FP32_REF){
        rc = alloc(c->rb_d, 2 * hn * sizeof(float));
        if(!rc) rc = alloc(c->rb_r, 2 * hn * sizeof(float));
    }
    if(!rc) rc = alloc(c->rb_flags, 2 * hn);
    if(!rc) rc = alloc(c->mask, c->n3);
    if(!rc) rc = alloc(c->flags, c->n3);
    if(!rc) rc = alloc(c->itopo, (size_t)c->jm * c->km * sizeof(int));
    if(!rc) rc = alloc(c->colr, (size_t)c->jm * c->km * 2 * sizeof(int));
    if(!rc){
        // no obstacle until a mask is uploaded: every column is fluid from 0 to im-1
        std::vector<int> runs((size_t)c->jm * c->km * 2);
        for(size_t q = 0; q < runs.size(); q += 2){ runs[q] = 0; runs[q + 1] = c->im - 1; }
        if(hipMemcpy(c->colr.p, runs.data(), runs.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess){
            set_error("hipMemcpy failed"); rc = ATJUP_ERR_HIP;
        }
    }
    if(!rc) rc = alloc(c->vel_star, (size_t)c->jm * c->km * sizeof(double));
    if(!rc) rc = alloc(c->brl, (size_t)c->im * asz);
    if(!rc) rc = alloc(c->brl_part, (size_t)c->im * 64 * 2 * asz);
    if(!rc) rc = alloc(c->zm, (size_t)4 * c->jm * asz);
    if(!rc) rc = alloc(c->nonfinite, 64);
    if(!rc){
        // resetArrays: dis = 1e-10, rho_mix = 1 (planet/cJupiterModel.cpp:1719-1720, 1677)
        rc = fill_value(c, c->cur[ATJUP_F_DIS], 1.0e-10);
        if(!rc) rc = fill_value(c, c->cur[ATJUP_F_RHO_MIX], 1.0);
    }
    if(!rc){
        std::vector<double> tabs;
        if(precision != ATJUP_PRECISION_FP32_REF){
            std::vector<GeoIJ<double>> geo; std::vector<PoisIJ<double>> pois;
            build_tables<double>(c, geo, pois);
            rc = alloc(c->geo, geo.size() * sizeof(geo[0]));
            if(!rc && hipMemcpy(c->geo.p, geo.data(), c->geo.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
            if(!rc) rc = alloc(c->pois, pois.size() * sizeof(pois[0]));
            if(!rc && hipMemcpy(c->pois.p, pois.data(), c->pois.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
        } else {
            std::vector<GeoIJ<float>> geo; std::vector<PoisIJ<float>> pois32;
            build_tables<float>(c, geo, pois32);
            std::vector<GeoIJ<double>> g64; std::vector<PoisIJ<double>> pois;
            build_tables<double>(c, g64, pois);
            rc = alloc(c->geo, geo.size() * sizeof(geo[0]));
            if(!rc && hipMemcpy(c->geo.p, geo.data(), c->geo.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
            if(!rc) rc = alloc(c->pois, pois.size() * sizeof(pois[0]));
            if(!rc && hipMemcpy(c->pois.p, pois.data(), c->pois.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
            if(!rc) rc = alloc(c->pois32, pois32.size() * sizeof(pois32[0]));
            if(!rc && hipMemcpy(c->pois32.p, pois32.data(), c->pois32.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
        }
        if(!rc){
            tabs.resize((size_t)2 * c->im + 3 * c->jm);
            for(int i = 0; i < c->im; i++) tabs[i] = c->rad[i];
            for(int j = 0; j < c->jm; j++){
                tabs[c->im + j] = c->sin_the[j];
                // the Coriolis factor of the split column integral, nd_cor = 1 (RungeKutta_Jup_Turb.cpp:185-186)
                tabs[c->im + c->jm + j] = c->c.Coriolis * 1.0 * 2.0 * c->c.omega * std::abs(c->sin_the[j]);
                tabs[c->im + 2 * c->jm + j] = c->cos_the[j];
            }
            // layer_thickness_m (cJupiterModel.h:147-150) from the float layer heights
            for(int i = 0; i < c->im; i++)
                tabs[c->im + 3 * c->jm + i] = (i > c->im - 2) ? -1.0 : (double)(c->lh[i + 1] - c->lh[i]) * 1.0e3;
            rc = alloc(c->tabs, tabs.size() * sizeof(double));
            if(!rc && hipMemcpy(c->tabs.p, tabs.data(), c->tabs.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = ATJUP_ERR_HIP;
        }
    }
    if(!rc){
        atjup_hip_default_rad_params(&c->radp);
        rc = build_rad_table(c);
    }
    if(!rc){
        atjup_hip_default_precip_params(&c->precp);
        rc = alloc(c->precip_srf, (size_t)5 * c->jm * c->km * sizeof(double));
    }
    if(!rc){
        atjup_hip_default_chem_params(&c->chemp);
        atjup_hip_default_mca_params(&c->mcap);
        // with the NH4SH equilibrium the moist convective adjustment carries its heat of formation
        // (cJupiterModel.cpp:1316)
        if(kn->nh4sh_eq) c->mcap.l_nh4sh = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;
        rc = alloc(c->report, report_bytes(c->im, c->jm, c->km));
        // the reaction rates persist between chemistry calls; a dry context runs no chemistry
        if(!rc && !dry) rc = alloc(c->chem_w, 3 * c->n3 * c->esz);
    }
    if(rc){ atjup_hip_ctx_destroy(c); return nullptr; }
    return c;
}

void atjup_hip_ctx_destroy(atjup_ctx* c){
    if(!c) return;
    if(c->stream) (void)hipStreamSynchronize(c->stream);
    adj_free(c);
    for(auto& b : c->cur) release(b);
    for(auto& s : c->stg) for(auto& b : s) release(b);
    for(auto& b : c->acc) release(b);
    for(auto& b : c->live_turb) release(b);
    for(auto& b : c->lo) release(b);
    release(c->steady);
    release(c->colr);
    for(atjup::Buf* b : { &c->scratch, &c->psrc, &c->rb_p, &c->rb_d, &c->rb_r, &c->rb_flags, &c->mask, &c->flags, &c->itopo, &c->vel_star, &c->geo,
                          &c->pois, &c->pois32, &c->tabs, &c->brl, &c->brl_part, &c->zm, &c->solids,
                          &c->amb_rows, &c->nonfinite, &c->rad_tab, &c->rad_tmp, &c->precip_srf, &c->chem_w, &c->mca_ws, &c->report,
                          &c->chi })
        release(*b);
    for(auto& v : c->spans) for(auto& sp : v){ (void)hipEventDestroy(sp.a); (void)hipEventDestroy(sp.b); }
    for(auto e : c->event_pool) (void)hipEventDestroy(e);
    for(auto& g : c->sweep_graphs) (void)hipGraphExecDestroy((hipGraphExec_t)g.exec);
    if(c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

static int upload_impl(atjup_ctx* c, int f, const double* src, double* const* const* rows){
    if(!c || (!src && !rows)){ set_error("null argument"); return ATJUP_ERR_ARG; }
    const int s = storage_of(f);
    if(!(prog(s) || input(s))){ set_error("field id"); return ATJUP_ERR_ARG; }
    if(!c->cur[s].p){ set_error("field not held by this profile"); return ATJUP_ERR_PROFILE; }
    const size_t row = (size_t)c->km;
    std::vector<char> host(c->n3 * c->esz, 0);
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++){
            const double* r = rows ? rows[i][j] : src + ((size_t)i * c->jm + j) * row;
            const size_t o = c->idx(i, j, 0);
            if(c->esz == 8) std::memcpy((double*)host.data() + o, r, row * 8);
            else if(c->esz == 16){ double* d = (double*)host.data() + 2 * o; for(size_t k = 0; k < row; k++){ d[2 * k] = r[k]; d[2 * k + 1] = 0.0; } }
            else { float* d = (float*)host.data() + o; for(size_t k = 0; k < row; k++) d[k] = (float)r[k]; }
        }
    if(c->kn.dry && s >= ATJUP_F_H2O && s < ATJUP_F_TKE){
        // the dry profile never streams the species: they must be, and stay, zero
        const size_t nb = host.size();
        for(size_t b = 0; b < nb; b++)
            if(host[b] != 0){ set_error("the dry profile requires the species to be zero"); return ATJUP_ERR_PROFILE; }
    }
    ATJUP_HIP_CHECK(hipMemcpyAsync(c->cur[s].p, host.data(), host.size(), hipMemcpyHostToDevice, c->stream));
    if(s < ATJUP_N_PROG && c->lo[s].p)            // an uploaded field starts without a low word
        ATJUP_HIP_CHECK(hipMemsetAsync(c->lo[s].p, 0, c->lo[s].bytes, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    if(s == ATJUP_F_TKE || s == ATJUP_F_DIS) c->turb_live_pending = false;
    return 0;
}

static int download_impl(atjup_ctx* c, int f, double* dst, double** const* rows){
    if(!c || (!dst && !rows)){ set_error("null argument"); return ATJUP_ERR_ARG; }
    const int s = storage_of(f);
    if(!(prog(s) || input(s))){ set_error("field id"); return ATJUP_ERR_ARG; }
    if(!c->cur[s].p){ set_error("field not held by this profile"); return ATJUP_ERR_PROFILE; }
    std::vector<char> host(c->n3 * c->esz);
    // the live k*, omega* of a turbulence run that waits for its step; their n-level ids read the n level
    const void* src_p = nlevel(f) ? c->cur[s].p : live_ptr(c, s);
    ATJUP_HIP_CHECK(hipMemcpyAsync(host.data(), src_p, host.size(), hipMemcpyDeviceToHost, c->stream));
    std::vector<float> lo;
    if(s < ATJUP_N_PROG && c->lo[s].p){            // a paired field of the refined form: hi + lo
        lo.resize(c->n3);
        ATJUP_HIP_CHECK(hipMemcpyAsync(lo.data(), c->lo[s].p, c->n3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    const size_t row = (size_t)c->km;
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++){
            double* r = rows ? rows[i][j] : dst + ((size_t)i * c->jm + j) * row;
            const size_t o = c->idx(i, j, 0);
            if(c->esz == 8) std::memcpy(r, (const double*)host.data() + o, row * 8);
            else if(c->esz == 16){ const double* d = (const double*)host.data() + 2 * o; for(size_t k = 0; k < row; k++) r[k] = d[2 * k]; }
            else {
                const float* d = (const float*)host.data() + o;
                if(lo.empty()) for(size_t k = 0; k < row; k++) r[k] = (double)d[k];
                else for(size_t k = 0; k < row; k++) r[k] = (double)d[k] + (double)lo[o + k];
            }
        }
    return 0;
}

int atjup_hip_upload_field(atjup_ctx* c, int f, const double* src){ return upload_impl(c, f, src, nullptr); }

// the directional-derivative part of a field of the tangent form; the value part is kept
static int tangent_impl(atjup_ctx* c, int f, const double* src, double* dst){
    if(!c || (!src && !dst)){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(c->precision != ATJUP_PRECISION_TANGENT){ set_error("not a tangent context"); return ATJUP_ERR_STATE; }
    const int s = storage_of(f);
    if(!(prog(s) || input(s)) || !c->cur[s].p){ set_error("field id"); return ATJUP_ERR_ARG; }
    std::vector<double> host(2 * c->n3);
    ATJUP_HIP_CHECK(hipMemcpyAsync(host.data(), live_ptr(c, s), c->n3 * 16, hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++)
            for(int k = 0; k < c->km; k++){
                const size_t o = c->idx(i, j, k), h = ((size_t)i * c->jm + j) * c->km + k;
                if(src) host[2 * o + 1] = src[h];
                else dst[h] = host[2 * o + 1];
            }
    if(src){
        ATJUP_HIP_CHECK(hipMemcpyAsync(c->cur[s].p, host.data(), c->n3 * 16, hipMemcpyHostToDevice, c->stream));
        ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    return 0;
}
int atjup_hip_upload_tangent(atjup_ctx* c, int f, const double* src){ return tangent_impl(c, f, src, nullptr); }
int atjup_hip_download_tangent(atjup_ctx* c, int f, double* dst){ return tangent_impl(c, f, nullptr, dst); }
int atjup_hip_download_field(atjup_ctx* c, int f, double* dst){ return download_impl(c, f, dst, nullptr); }
int atjup_hip_upload_field_rows(atjup_ctx* c, int f, double* const* const* rows){ return upload_impl(c, f, nullptr, rows); }
int atjup_hip_download_field_rows(atjup_ctx* c, int f, double** const* rows){ return download_impl(c, f, nullptr, rows); }

// ---- the restart file (planet/FileIO_Jup.cpp:281-305, 584-684) ---------------------------------
namespace {
constexpr int RS_PDYNN = -1;                  // p_dynn: the n-level of p_dyn, the live field here
constexpr int NL = ATJUP_NLEVEL;
// restart_arrays() (FileIO_Jup.cpp:281-293)
const int kRestartFixed[38] = {
    ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W, NL + ATJUP_F_T, NL + ATJUP_F_U, NL + ATJUP_F_V, NL + ATJUP_F_W,
    ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, NL + ATJUP_F_H2O, NL + ATJUP_F_H2O_CLOUD, NL + ATJUP_F_H2O_ICE,
    ATJUP_F_H2S, NL + ATJUP_F_H2S,
    ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, NL + ATJUP_F_NH3, NL + ATJUP_F_NH3_CLOUD, NL + ATJUP_F_NH3_ICE,
    ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, NL + ATJUP_F_CH4, NL + ATJUP_F_CH4_CLOUD, NL + ATJUP_F_CH4_ICE,
    ATJUP_F_NH4SH, NL + ATJUP_F_NH4SH,
    ATJUP_F_P_DYN, RS_PDYNN, ATJUP_F_P_STAT,
    ATJUP_F_TKE, ATJUP_F_DIS, NL + ATJUP_F_TKE, NL + ATJUP_F_DIS, ATJUP_F_NUE };
// restart_arrays_carried() (FileIO_Jup.cpp:297-305)
const int kRestartCarried[17] = {
    ATJUP_F_AUX_U, ATJUP_F_AUX_V, ATJUP_F_AUX_W, ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3,
    ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_FLUXLIM_NH4SH, ATJUP_F_THERMALMASSFLUX,
    ATJUP_F_Q_PRECIP, ATJUP_F_Q_RAD, ATJUP_F_RHO_MIX,
    ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_H2O_CLOUD, ATJUP_F_S_PRECIP_H2O_ICE, ATJUP_F_S_PRECIP_NH3,
    ATJUP_F_S_PRECIP_NH3_ICE };
constexpr int32_t MAGIC_JUP1 = 0x4A555031, MAGIC_JUP2 = 0x4A555032;

struct FileCloser { FILE* f; ~FileCloser(){ if(f) fclose(f); } };
}  // namespace

int atjup_hip_save_state(atjup_ctx* c, const char* path, int iter, int carried){
    if(!c || !path){ set_error("null argument"); return ATJUP_ERR_ARG; }
    FILE* f = fopen(path, "wb");
    if(!f){ set_error(std::string("save_state: cannot open ") + path); return ATJUP_ERR_STATE; }
    FileCloser fc{f};
    const int32_t hdr[5] = { carried ? MAGIC_JUP2 : MAGIC_JUP1, c->im, c->jm, c->km, iter };
    bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1;
    const size_t n = (size_t)c->im * c->jm * c->km;
    std::vector<double> buf(n);
    auto put = [&](int id){
        if(!ok) return;
        if(int rc = download_impl(c, id == RS_PDYNN ? ATJUP_F_P_DYN : id, buf.data(), nullptr)){ ok = false; (void)rc; return; }
        ok = fwrite(buf.data(), sizeof(double), n, f) == n;
    };
    for(int id : kRestartFixed) put(id);
    if(carried){
        for(int id : kRestartCarried) put(id);
        if(ok){
            std::vector<double> vs((size_t)c->jm * c->km);
            if(atjup_hip_download_vel_star(c, vs.data()) != 0) ok = false;
            else ok = fwrite(vs.data(), sizeof(double), vs.size(), f) == vs.size();
        }
    }
    if(!ok){ set_error(std::string("save_state: failed while writing ") + path); return ATJUP_ERR_STATE; }
    if(fflush(f) != 0){ set_error(std::string("save_state: failed while writing ") + path); return ATJUP_ERR_STATE; }
    return 0;
}

int atjup_hip_load_state(atjup_ctx* c, const char* path, int* iter, int* carried_out){
    if(!c || !path){ set_error("null argument"); return ATJUP_ERR_ARG; }
    FILE* f = fopen(path, "rb");
    if(!f){ set_error(std::string("load_state: no file ") + path); return ATJUP_ERR_STATE; }
    FileCloser fc{f};
    int32_t hdr[5];
    if(fread(hdr, sizeof(hdr), 1, f) != 1 || (hdr[0] != MAGIC_JUP1 && hdr[0] != MAGIC_JUP2)
       || hdr[1] != c->im || hdr[2] != c->jm || hdr[3] != c->km){
        set_error(std::string("load_state: bad header / grid mismatch in ") + path); return ATJUP_ERR_STATE;
    }
    const bool jup2 = hdr[0] == MAGIC_JUP2;
    const size_t n = (size_t)c->im * c->jm * c->km, nb = n * sizeof(double);
    const size_t n_arr = 38 + (jup2 ? 17 : 0);
    const size_t need = sizeof(hdr) + n_arr * nb + (jup2 ? (size_t)c->jm * c->km * sizeof(double) : 0);
    // nothing is loaded from a short file or from one whose n-level copies differ from their fields
    // (one storage holds both on this path): both are checked before the first upload
    if(fseek(f, 0, SEEK_END) != 0 || (size_t)ftell(f) < need){
        set_error(std::string("load_state: truncated file ") + path); return ATJUP_ERR_STATE;
    }
    std::vector<double> buf(n), base(n);
    auto read_at = [&](size_t a, double* dst){
        return fseek(f, (long)(sizeof(hdr) + a * nb), SEEK_SET) == 0 && fread(dst, sizeof(double), n, f) == n;
    };
    for(size_t a = 0; a < 38; a++){
        const int id = kRestartFixed[a];
        const bool nlev = id == RS_PDYNN || (id >= NL && id < NL + ATJUP_N_PROG);
        if(!nlev) continue;
        const int want = id == RS_PDYNN ? ATJUP_F_P_DYN : id - NL;
        size_t b = 0;
        while(b < 38 && kRestartFixed[b] != want) b++;
        if(!read_at(a, buf.data()) || !read_at(b, base.data())){
            set_error(std::string("load_state: read error in ") + path); return ATJUP_ERR_STATE;
        }
        if(std::memcmp(buf.data(), base.data(), nb) != 0){
            set_error(std::string("load_state: n-level arrays differ from their fields in ") + path);
            return ATJUP_ERR_STATE;
        }
    }
    if(fseek(f, (long)sizeof(hdr), SEEK_SET) != 0){ set_error("load_state: seek failed"); return ATJUP_ERR_STATE; }
    for(size_t a = 0; a < n_arr; a++){
        const int id = a < 38 ? kRestartFixed[a] : kRestartCarried[a - 38];
        if(fread(buf.data(), sizeof(double), n, f) != n){
            set_error(std::string("load_state: read error in ") + path); return ATJUP_ERR_STATE;
        }
        if(id == RS_PDYNN || (id >= NL && id < NL + ATJUP_N_PROG)) continue;
        if(int rc = upload_impl(c, id, buf.data(), nullptr)) return rc;
    }
    if(jup2){
        std::vector<double> vs((size_t)c->jm * c->km);
        if(fread(vs.data(), sizeof(double), vs.size(), f) != vs.size()){
            set_error(std::string("load_state: read error in ") + path); return ATJUP_ERR_STATE;
        }
        if(int rc = atjup_hip_upload_vel_star(c, vs.data())) return rc;
    }
    if(iter) *iter = hdr[4];
    if(carried_out) *carried_out = jup2 ? 1 : 0;
    return 0;
}

int atjup_hip_upload_mask(atjup_ctx* c, const double* m){
    if(!c || !m){ set_error("null argument"); return ATJUP_ERR_ARG; }
    std::vector<uint8_t> host(c->n3, 0);
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++)
            for(int k = 0; k < c->km; k++)
                host[c->idx(i, j, k)] = (m[((size_t)i * c->jm + j) * c->km + k] == 1.0) ? 1 : 0;
    ATJUP_HIP_CHECK(hipMemcpy(c->mask.p, host.data(), c->n3, hipMemcpyHostToDevice));
    c->have_mask = true;
    return launch_mask_setup(c);
}

int atjup_hip_upload_vel_star(atjup_ctx* c, const double* s){
    if(!c || !s){ set_error("null argument"); return ATJUP_ERR_ARG; }
    ATJUP_HIP_CHECK(hipMemcpy(c->vel_star.p, s, (size_t)c->jm * c->km * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}
int atjup_hip_download_vel_star(atjup_ctx* c, double* d){
    if(!c || !d){ set_error("null argument"); return ATJUP_ERR_ARG; }
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    ATJUP_HIP_CHECK(hipMemcpy(d, c->vel_star.p, (size_t)c->jm * c->km * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}
int atjup_hip_download_buoy_ref_level(atjup_ctx* c, double* d){
    if(!c || !d){ set_error("null argument"); return ATJUP_ERR_ARG; }
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    if(c->precision == ATJUP_PRECISION_TANGENT){
        std::vector<double> h(2 * (size_t)c->im);
        ATJUP_HIP_CHECK(hipMemcpy(h.data(), c->brl.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for(int i = 0; i < c->im; i++) d[i] = h[2 * i];
        return 0;
    }
    ATJUP_HIP_CHECK(hipMemcpy(d, c->brl.p, (size_t)c->im * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static int need_mask(atjup_ctx* c){
    if(!c){ set_error("null context"); return ATJUP_ERR_ARG; }
    if(!c->have_mask){ set_error("upload the obstacle mask first (atjup_hip_upload_mask)"); return ATJUP_ERR_STATE; }
    return 0;
}

int atjup_hip_rk4_step(atjup_ctx* c){ if(int rc = need_mask(c)) return rc; return launch_rk4(c); }
int atjup_hip_pressure_solve(atjup_ctx* c, int sweeps, double omega){
    if(int rc = need_mask(c)) return rc;
    if(!(omega > 0.0 && omega < 2.0)){ set_error("omega outside (0, 2)"); return ATJUP_ERR_ARG; }
    hipEvent_t t0 = timer_begin(c);
    const int rc = launch_pressure(c, sweeps > 0 ? sweeps : std::max(1, c->kn.press_sweeps), omega);
    timer_end(c, ATJUP_T_PRESSURE, t0, 0);
    return rc;
}
int atjup_hip_set_gas(atjup_ctx* c, int gas, const atjup_gas_params* p){
    if(!c || !p){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(gas < 0 || gas > 2){ set_error("gas"); return ATJUP_ERR_ARG; }
    c->gas[gas] = *p;
    c->gas_set[gas] = true;
    return 0;
}
int atjup_hip_satadj(atjup_ctx* c, int gas, const atjup_gas_params* p){
    if(int rc = need_mask(c)) return rc;
    if(gas < 0 || gas > 2){ set_error("gas"); return ATJUP_ERR_ARG; }
    if(c->kn.dry) return 0;        // all species zero: the adjustment is the identity (SaturationAdjustment.h:440)
    if(!p && !c->gas_set[gas]){ set_error("gas constants not set (atjup_hip_set_gas)"); return ATJUP_ERR_STATE; }
    return launch_satadj(c, gas, p ? *p : c->gas[gas]);
}
void atjup_hip_default_rad_params(atjup_rad_params* p){
    if(!p) return;
    std::memset(p, 0, sizeof(*p));
    // Jupiter (planet/cJupiterModel.h:1109-1118) and the knob defaults (planet/Radiation.h:270-295)
    p->F_int = 5.4; p->S_solar = 50.5; p->albedo_bond = 0.343; p->x_H2 = 0.863; p->x_He = 0.134;
    p->cia_strength = 3.9; p->opacity_strength = 1.0; p->solar_strength = 1.0; p->sw_tau_per_bar = 1.0;
    p->solar = 1; p->bottom = 0;
}
int atjup_hip_set_radiation(atjup_ctx* c, const atjup_rad_params* p){
    if(!c || !p){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(p->bottom != 0 && p->bottom != 1){ set_error("radiation bottom must be 0 or 1"); return ATJUP_ERR_ARG; }
    c->radp = *p;
    return build_rad_table(c);
}
int atjup_hip_radiation(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if((size_t)(c->im + 1) * 64 * sizeof(double) > 65536){ set_error("radiation: im above 1023 levels"); return ATJUP_ERR_PROFILE; }
    return launch_radiation(c);
}
void atjup_hip_default_precip_params(atjup_precip_params* p){
    if(!p) return;
    // ATJUP_PRECIP_SCALE 1 (planet/cJupiterModel.h:1059-1061), the NH4SH band 200-230 K (:194-195)
    p->rate_scale = 1.0; p->t_0_nh4sh = 230.0; p->t_00_nh4sh = 200.0;
}
int atjup_hip_set_precipitation(atjup_ctx* c, const atjup_precip_params* p){
    if(!c || !p){ set_error("null argument"); return ATJUP_ERR_ARG; }
    c->precp = *p;
    return 0;
}
int atjup_hip_precipitation(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    for(int gas = 0; gas < 3; gas++)
        if(!c->gas_set[gas]){ set_error("gas constants not set (atjup_hip_set_gas)"); return ATJUP_ERR_STATE; }
    return launch_precipitation(c);
}
int atjup_hip_download_precip_surface(atjup_ctx* c, double* dst){
    if(!c || !dst){ set_error("null argument"); return ATJUP_ERR_ARG; }
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    ATJUP_HIP_CHECK(hipMemcpy(dst, c->precip_srf.p, c->precip_srf.bytes, hipMemcpyDeviceToHost));
    return 0;
}
void atjup_hip_default_chem_params(atjup_chem_params* p){
    if(!p) return;
    // planet/cJupiterModel.h:172-401 as the start-up leaves them (k_mix the mixture conductivity of
    // the default composition), GATE_ZERO 1, MOLAR_CONC 0, ENTHALPY 0 (planet/ChemistryJup.h)
    *p = atjup_chem_params{ 17.03052, 34.08088, 51.1114, 2.016, 4.0026, 18.01588,
                            0.864, 0.136, 0.0017, 7.7e-05, 0.0002, 0.0,
                            9.2e-06, 1.3e-05, 0.0, 817.0, 949.0, 0.12478002943205087,
                            2190.0, 2240.0, 2000.0, 0.02102, 0.013, 0.0, 1768700.0, 230.0, 200.0,
                            1, 0, 0 };
}
int atjup_hip_set_chemistry(atjup_ctx* c, const atjup_chem_params* p){
    if(!c || !p){ set_error("null argument"); return ATJUP_ERR_ARG; }
    c->chemp = *p;
    return 0;
}
int atjup_hip_chemistry(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("the tangent form has no chemistry"); return ATJUP_ERR_PROFILE; }
    if(c->kn.dry || !c->chem_w.p){ set_error("the chemistry needs the moist profile (dry = 0)"); return ATJUP_ERR_PROFILE; }
    static const int diff[4] = { ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3, ATJUP_F_DIFFLUX_NH4SH, ATJUP_F_THERMALMASSFLUX };
    static const int mass[4] = { ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH, ATJUP_F_FLUXLIM_NH4SH };
    int rc = launch_chem_kernel(c, 0);
    for(int f : diff) if(!rc) rc = launch_shapiro(c, f, 2, 1.0, 1, 7);
    if(!rc) rc = launch_chem_kernel(c, 1);
    if(!rc) rc = launch_flux_limiter(c);
    for(int f : mass) if(!rc) rc = launch_shapiro(c, f, 2, 1.0, 1, 7);
    return rc;
}
void atjup_hip_default_mca_params(atjup_mca_params* p){
    if(!p) return;
    // ConvectiveAdjustment.h:409-413 defaults; ep_h2s = m_h2s / M_mix of the default composition
    // (InitValues_Jup.cpp:940-946); no NH4SH heat of formation with ATJUP_NH4SH_EQ 0
    *p = atjup_mca_params{ 1.0, 14.474287352261996, 0.0, 1, 1, 64 };
}
int atjup_hip_set_mca(atjup_ctx* c, const atjup_mca_params* p){
    if(!c || !p){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(!std::isfinite(p->relax) || !std::isfinite(p->ep_h2s) || !std::isfinite(p->l_nh4sh)){
        set_error("non-finite adjustment parameter"); return ATJUP_ERR_ARG;
    }
    c->mcap = *p;
    return 0;
}
int atjup_hip_mca(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->precision == ATJUP_PRECISION_TANGENT || c->kn.dry){
        set_error("the moist convective adjustment needs the moist profile"); return ATJUP_ERR_PROFILE;
    }
    for(int gas = 0; gas < 2; gas++)
        if(!c->gas_set[gas]){ set_error("gas constants not set (atjup_hip_set_gas)"); return ATJUP_ERR_STATE; }
    if(c->im > 128){ set_error("the moist convective adjustment takes at most 128 levels"); return ATJUP_ERR_PROFILE; }
    if(!c->mca_ws.p){
        // layer heights [m] as double(get_layer_height(i)) * 1e3 (ConvectiveAdjustment.h:584), then
        // the column workspace
        const size_t n = (size_t)c->im * c->jm * c->km;
        const size_t flags = (2 * n + 7) / 8 * 8;       // touched / equilibrium bits, census bits
        if(int rc = alloc(c->mca_ws, (c->im + 8 * n) * sizeof(double) + flags + (size_t)c->jm * c->km * 8)) return rc;
        std::vector<double> z(c->im);
        for(int i = 0; i < c->im; i++) z[i] = (double)c->lh[i] * 1.0e3;
        ATJUP_HIP_CHECK(hipMemcpy(c->mca_ws.p, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return launch_mca(c);
}
int atjup_hip_nh4sh_equilibrium(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->precision == ATJUP_PRECISION_TANGENT || c->kn.dry){
        set_error("the NH4SH equilibrium needs the moist profile"); return ATJUP_ERR_PROFILE;
    }
    return launch_nh4sh_eq(c);
}
int atjup_hip_turbulence_run(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->kn.turb_model < 0) return 0;
    return launch_turbulence(c);
}
int atjup_hip_turb_wall(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->kn.turb_model < 0) return 0;
    return launch_turb_wall(c);
}
int atjup_hip_bc_after_step(atjup_ctx* c){ if(int rc = need_mask(c)) return rc; return launch_bc_after_step(c); }
int atjup_hip_shapiro(atjup_ctx* c, int f, int order, double strength, int passes, int axes){
    if(int rc = need_mask(c)) return rc;
    const int s = storage_of(f);
    if(!(prog(s) || input(s)) || !c->cur[s].p || (order != 2 && order != 4) || passes < 0){
        set_error("shapiro arguments"); return ATJUP_ERR_ARG;
    }
    return launch_shapiro(c, s, order, strength, passes, axes & 7);
}
int atjup_hip_mixture_density(atjup_ctx* c){ if(!c) return ATJUP_ERR_ARG; return launch_mixture_density(c); }
int atjup_hip_flux_limiter(atjup_ctx* c){
    if(int rc = need_mask(c)) return rc;
    if(c->kn.dry) return 0;        // nh4sh = 0: the correction is zero
    return launch_flux_limiter(c);
}
// The frozen-branch linearisations: nin host arrays [im][jm][km] staged into the field layout on
// the device, the kernel, nout arrays staged back. Scratch buffers are allocated per call.
typedef int (*lin_launch)(atjup_ctx* c, int gas, int adjoint, const double* in, double* out);
static int lin_flux(atjup_ctx* c, int, int adjoint, const double* in, double* out){
    return launch_flux_limiter_lin(c, adjoint, in, out);
}
static int lin_call(atjup_ctx* c, int nin, int nout, const double* in, double* out, int gas, int adjoint,
                    lin_launch launch){
    if(!c || !in || !out){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(int rc = need_mask(c)) return rc;
    if(c->precision != ATJUP_PRECISION_FP64 || c->kn.dry){
        set_error("the linearised operators need an FP64 context of a moist profile"); return ATJUP_ERR_PROFILE; }
    const size_t n3 = c->n3, nh = (size_t)c->im * c->jm * c->km;
    double *din = nullptr, *dout = nullptr;
    if(hipMalloc((void**)&din, nin * n3 * 8) != hipSuccess || hipMalloc((void**)&dout, nout * n3 * 8) != hipSuccess){
        if(din) (void)hipFree(din);
        set_error("hipMalloc failed"); return ATJUP_ERR_NOMEM;
    }
    int rc = 0;
    auto hip = [&](hipError_t e, const char* what){
        if(!rc && e != hipSuccess){ set_error(std::string(what) + ": " + hipGetErrorString(e)); rc = ATJUP_ERR_HIP; }
    };
    hip(hipMemsetAsync(din, 0, nin * n3 * 8, c->stream), "hipMemsetAsync");
    hip(hipMemsetAsync(dout, 0, nout * n3 * 8, c->stream), "hipMemsetAsync");
    for(int a = 0; a < nin && !rc; a++)
        hip(hipMemcpy2DAsync(din + a * n3, (size_t)c->pitch * 8, in + a * nh, (size_t)c->km * 8, (size_t)c->km * 8,
                             (size_t)c->im * c->jm, hipMemcpyHostToDevice, c->stream), "hipMemcpy2DAsync");
    if(!rc) rc = launch(c, gas, adjoint, din, dout);
    for(int a = 0; a < nout && !rc; a++)
        hip(hipMemcpy2DAsync(out + a * nh, (size_t)c->km * 8, dout + a * n3, (size_t)c->pitch * 8, (size_t)c->km * 8,
                             (size_t)c->im * c->jm, hipMemcpyDeviceToHost, c->stream), "hipMemcpy2DAsync");
    hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}
static int satadj_lin(atjup_ctx* c, int gas, int adjoint, const double* in, double* out){
    if(c && (gas < 0 || gas > 2)){ set_error("gas"); return ATJUP_ERR_ARG; }
    if(c && !c->gas_set[gas]){ set_error("gas constants not set (atjup_hip_set_gas)"); return ATJUP_ERR_STATE; }
    return lin_call(c, adjoint ? 4 : 6, adjoint ? 6 : 4, in, out, gas, adjoint, launch_satadj_lin);
}
int atjup_hip_satadj_tangent(atjup_ctx* c, int gas, const double* dx, double* dy){ return satadj_lin(c, gas, 0, dx, dy); }
int atjup_hip_satadj_adjoint(atjup_ctx* c, int gas, const double* ly, double* lx){ return satadj_lin(c, gas, 1, ly, lx); }
static int flux_limiter_lin(atjup_ctx* c, int adjoint, const double* in, double* out){
    if(c && c->km - 1 < 4){ set_error("the linearised limiter needs at least 4 distinct longitudes"); return ATJUP_ERR_ARG; }
    return lin_call(c, adjoint ? 1 : 4, adjoint ? 4 : 1, in, out, 0, adjoint, lin_flux);
}
int atjup_hip_flux_limiter_tangent(atjup_ctx* c, const double* dx, double* dy){ return flux_limiter_lin(c, 0, dx, dy); }
int atjup_hip_flux_limiter_adjoint(atjup_ctx* c, const double* ly, double* lx){ return flux_limiter_lin(c, 1, ly, lx); }
int atjup_hip_clamp_species(atjup_ctx* c){ if(!c) return ATJUP_ERR_ARG; return launch_clamp_species(c); }
// restoreVar: the n level is the live storage, except the closure's live k*, omega* after a
// turbulence run, which become the n level here
int atjup_hip_swap_nlevel(atjup_ctx* c){ if(!c) return ATJUP_ERR_ARG; return launch_commit_live(c); }
int atjup_hip_scalar_sync(atjup_ctx* c){ return c ? 0 : ATJUP_ERR_ARG; }

int atjup_hip_step(atjup_ctx* c, int iter, int first){
    if(int rc = need_mask(c)) return rc;
    hipEvent_t t_step = timer_begin(c);
    if(int rc = atjup_hip_step_block(c, iter, first)) return rc;
    if(int rc = step_rest(c, iter)) return rc;
    timer_end(c, ATJUP_T_STEP, t_step, 0);
    return 0;
}
int atjup_hip_step_block(atjup_ctx* c, int iter, int first){
    if(int rc = need_mask(c)) return rc;
    // the iteration starts: steadyQuery's n level is the state now (restoreVar of the last one)
    if(c->steady_on)
        if(int rc = steady_mark(c, (1u << ATJUP_N_STEADY) - 1)) return rc;
    return step_block(c, iter, first);
}
int atjup_hip_step_rest(atjup_ctx* c, int iter){
    if(int rc = need_mask(c)) return rc;
    return step_rest(c, iter);
}

int atjup_hip_check_finite(atjup_ctx* c){ if(!c) return ATJUP_ERR_ARG; return launch_check_finite(c); }

int atjup_hip_set_timing(atjup_ctx* c, int enable){
    if(!c) return ATJUP_ERR_ARG;
    c->timing = enable != 0;
    return 0;
}

int atjup_hip_read_timers(atjup_ctx* c, double* out){
    if(!c || !out) return ATJUP_ERR_ARG;
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    for(int p = 0; p < ATJUP_N_TIMERS; p++){
        for(auto& sp : c->spans[p]){
            float ms = 0.0f;
            ATJUP_HIP_CHECK(hipEventElapsedTime(&ms, sp.a, sp.b));
            c->t_total[p] += ms;
            c->t_launch[p] += sp.launches;
            c->event_pool.push_back(sp.a);
            c->event_pool.push_back(sp.b);
        }
        c->spans[p].clear();
        out[2 * p] = c->t_total[p];
        out[2 * p + 1] = c->t_launch[p];
        c->t_total[p] = 0.0;
        c->t_launch[p] = 0.0;
    }
    return 0;
}

int atjup_hip_sync(atjup_ctx* c){
    if(!c) return ATJUP_ERR_ARG;
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}
int atjup_hip_device_count(int* count){
    if(!count) return ATJUP_ERR_ARG;
    *count = 0;
    int n = 0;
    ATJUP_HIP_CHECK(hipGetDeviceCount(&n));
    *count = n;
    return 0;
}
int atjup_hip_set_stream(atjup_ctx* c, void* s){
    if(!c) return ATJUP_ERR_ARG;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return 0;
}
void* atjup_hip_field_device_ptr(atjup_ctx* c, int f){
    if(!c) return nullptr;
    const int s = storage_of(f);
    if(!(prog(s) || input(s))) return nullptr;
    return c->cur[s].p;
}
int atjup_hip_pitch(atjup_ctx* c){ return c ? c->pitch : 0; }

}  // extern "C"
