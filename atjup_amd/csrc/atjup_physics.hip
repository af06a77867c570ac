// HIP kernels of the moist, turbulent time step (the full-physics profile) for gfx950:
//   * the RK4 stage with the RHS of all 17 prognostic fields (RHSJup, RHS_Jup_Turb.cpp:31-1457,
//     with the k-omega sources) fused with the stage accumulation and the stage input, as two
//     kernels per stage: t, u, v, w, k*, omega*, then the eleven species;
//   * Turbulence<P>::run with the k-omega closure and the wall-function omega Run() applies;
//   * SaturationAdjustment<P>::run, one thread per cell and gas;
//   * FluxLimiter<P>::nh4sh;
//   * the closure's part of bcSolidGround (the sanitising clamp of the turbulence arrays).
//
// Storage conventions are those of atjup_kernels.hip (pitch-padded [i][j][k], k fastest,
// XCD-aware block order). The live tke / dis the closure produces are held in the context's
// live_turb buffers until the RK4 step consumes them: the step integrates k* and omega* from the
// time-level-n copies while its right-hand side reads the closure's values, as the reference's
// RungeKuttaJup does with tken / disn (RungeKutta_Jup_Turb.cpp:420-490).
//
// FP64 kernels keep the reference's expression trees (-ffp-contract=off); the transcendental
// functions (exp, log, pow) are the device library's, within an ulp of the host's.
#include "atjup_device.h"

#include <algorithm>
#include <cmath>

namespace atjup {

constexpr int NP = ATJUP_N_PROG;

// =============================================================================================
// RK4 stage, full physics
// =============================================================================================
template<typename T>
struct StageFullArgs {
    const T* in[NP];            // stage input (the RHS reads it)
    T* out[NP];                 // stage output
    const T* nl[NP];            // time-level-n copy (the update starts from it)
    T* acc[NP];                 // k1 + 2 k2 + 2 k3 + k4 accumulator
    const T* bsrc[NP];          // source of the non-updated cells of out (nullptr: none)
    const T* spr[NP];           // S_precip_* of a species (by field id), nullptr if none
    const T* mf[NP];            // massflux_* of a species (by field id), nullptr if none
    T sp_coef[NP];              // species_diff_scale / (sc * re) of a species
    const T *pd, *ph, *ps, *rho, *wn, *nue, *qrad, *qprec, *fl_nh4sh;
    T *aux_u, *aux_v, *aux_w, *prod, *tsrc, *dsrc;
    const uint8_t* flags;
    const GeoIJ<T>* geo;
    const acc_t<T>* brl;
    int im, jm, km, pitch, nkb, njb, nb;
    size_t plane;
    int stage, turb;
    T c_in, wgt, dt, one_sixth;
    T inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
    double cbuoy, g, rR, t_ref_d;          // the buoyancy anomaly is formed in FP64 in both forms
    T cep, to_nd, r_mix, c_m2o, c_p2o, Coriolis, centrifugal, t_ref;
    T t_coef, inv_re, rad_coupling, precip_coupling, turb_coupling, L_rad, R_mix, cp_mix, u_0;
    T cr, pc, stokes_a, rg_nh4sh, mue_mix, nue_max, tke_max_nd, dis_min;
    float* lo[ATJUP_N_PROG];               // FP32-refined form: the low words of t, u, v, w and the
                                           // species (else null)
    double dt_d;
    // the wall distance of the SST and k-epsilon sources: i_topography, the layer thickness [m] of
    // each layer and the float layer spacing L_atm * dr [km] the layer heights are multiples of;
    // the friction velocity and re_turb of the k-epsilon damping
    const int* itopo;
    const double* lthick;
    const double* vel_star;
    float lh_h;
    T re_turb;
};

constexpr int SF_BK = 64, SF_BJ = 2;            // 2 rows: 1.7% less stage time than 4 or 8 on MI355X
#ifndef SF_MIN_WAVES
#define SF_MIN_WAVES 1
#endif
// species per load group in the species stage kernel: 2 gives 4% less FP32 stage time than 1 on
// MI355X (FP64 0.5%), 3 no more
#ifndef SP_GROUP
#define SP_GROUP 2
#endif

template<typename T>
struct Stencil7 { T c, ip, im, jp, jm, kp, km; };

template<typename T>
__device__ __forceinline__ Stencil7<T> load7(const T* __restrict__ X, size_t q, size_t P, size_t R,
                                             size_t qke, size_t qkw){
    Stencil7<T> s;
    s.c = X[q]; s.ip = X[q + P]; s.im = X[q - P]; s.jp = X[q + R]; s.jm = X[q - R];
    s.kp = X[qke]; s.km = X[qkw];
    return s;
}

// what the species RHS and stage update of one species read at a cell, loaded together so that a
// group of species can have all its loads in flight before its first store
template<typename T>
struct SpeciesLoads { Stencil7<T> s; T mf, spr, fl, acc, nl; float lo; };

// The low word of a paired field for the last stage's FP64 update (FP32 storage): loaded without a
// branch, from the n level itself in the other stages and for unpaired fields (a value the stage
// reads anyway), so that the wavefront does not wait at a branch for its loads in flight.
template<typename T>
__device__ __forceinline__ float stage_lo(const StageFullArgs<T>& a, int f, size_t q){
    if constexpr(std::is_same<T, float>::value){
        const float* lp = (a.stage == 3 && a.lo[f]) ? a.lo[f] : (const float*)a.nl[f];
        return lp[q];
    } else {
        return 0.0f;
    }
}

template<typename T>
__device__ __forceinline__ SpeciesLoads<T> species_load(const StageFullArgs<T>& a, int f, size_t q, size_t P,
        size_t R, size_t qke, size_t qkw){
    SpeciesLoads<T> L;
    L.s = load7(a.in[f], q, P, R, qke, qkw);
    L.mf = a.mf[f] ? a.mf[f][q] : T(0);
    L.spr = (f != ATJUP_F_NH4SH && a.spr[f]) ? a.spr[f][q] : T(0);
    L.fl = (f == ATJUP_F_NH4SH) ? a.fl_nh4sh[q] : T(0);
    L.acc = (a.stage == 0) ? T(0) : a.acc[f][q];
    L.nl = a.nl[f][q];
    L.lo = stage_lo(a, f, q);
    return L;
}

template<typename T>
__device__ __forceinline__ T species_rhs(const StageFullArgs<T>& a, int f, const SpeciesLoads<T>& L,
        const GeoIJ<T>& G, T u, T v_invrm, T w_invrs, T two_inv_rm, T nue_t_s, T rho_f){
    const Stencil7<T>& s = L.s;
    const T e1r = (s.ip - s.im) * a.inv_2dr, e2r = (s.ip - T(2) * s.c + s.im) * a.inv_dr2;
    const T e1t = (s.jp - s.jm) * a.inv_2dthe, e2t = (s.jp - T(2) * s.c + s.jm) * a.inv_dthe2;
    const T e1p = (s.kp - s.km) * a.inv_2dphi, e2p = (s.kp - T(2) * s.c + s.km) * a.inv_dphi2;
    const T transport = u * e1r + v_invrm * e1t + w_invrs * e1p;
    const T diffusion = e2r + e1r * two_inv_rm + e2t * G.inv_rm2
        + e1t * G.costhe_inv_rm2sinthe + e2p * G.inv_rm2sinthe2;
    T rhs;
    if(f == ATJUP_F_NH4SH){
        const T v_stokes = a.stokes_a * (a.rg_nh4sh - rho_f) * (T)a.g / a.mue_mix;
        rhs = -transport
            + L.fl
            + diffusion * (a.sp_coef[f] + nue_t_s)
            + a.cr * L.mf
            + (v_stokes / a.u_0) * e1r;
    } else {
        rhs = -transport + diffusion * (a.sp_coef[f] + nue_t_s);
        if(a.mf[f]) rhs = rhs + a.cr * L.mf;
        if(a.spr[f]) rhs = rhs + a.pc * L.spr;
    }
    return rhs;
}

// the stage update of one field (RungeKutta_Jup_Turb.cpp:420-568); kind 0: plain, 1: species
// (max(0, .) at the end), 2: nh4sh (< 0 -> 0 at the end), 3: k* (safe_clamp to [0, tke_max]),
// 4: omega* (max(dis_min, .))
// a stage kernel's stores of the accumulator and the stage output: non-temporal (streamed past the
// caches; the next stage reads them after the whole grid has gone through), 1-2% less stage time
// on MI355X (ATJUP_NT_STAGE 0: plain stores)
#ifndef ATJUP_NT_STAGE
#define ATJUP_NT_STAGE 1
#endif
template<typename T>
__device__ __forceinline__ void st_field(T* p, T v){
    if constexpr(ATJUP_NT_STAGE) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template<typename T>
__device__ __forceinline__ void stage_update_v(const StageFullArgs<T>& a, int f, size_t q, T rhs, int kind,
        T acc_in, T nl_v, float lo_v){
    T accf;
    if(a.stage == 0) accf = a.wgt * rhs;
    else accf = acc_in + a.wgt * rhs;
    if(a.stage < 3){
        st_field(&a.acc[f][q], accf);
        T v = nl_v + a.c_in * rhs;
        if(kind == 3) v = safe_clamp_d(v, T(0), a.tke_max_nd);
        else if(kind == 4) v = dmax(a.dis_min, v);
        st_field(&a.out[f][q], v);
    } else {
        if constexpr(std::is_same<T, float>::value){
            // the paired fields of the refined form (t, u, v, w, the species): the increment added
            // to hi + lo in FP64, the species' clamps applied to that sum, and the sum split again
            if(a.lo[f]){
                double s = ((double)nl_v + (double)lo_v) + a.dt_d * (double)accf * (1.0 / 6.0);
                if(kind == 1) s = dmax(0.0, s);
                else if(kind == 2){ if(s < 0.0) s = 0.0; }
                const float h = (float)s;
                st_field(&a.out[f][q], h);
                a.lo[f][q] = (float)(s - (double)h);
                return;
            }
        }
        T v = nl_v + a.dt * accf * a.one_sixth;
        if(kind == 1) v = dmax(T(0), v);
        else if(kind == 2){ if(v < T(0)) v = T(0); }
        else if(kind == 3) v = safe_clamp_d(v, T(0), a.tke_max_nd);
        else if(kind == 4) v = dmax(a.dis_min, v);
        st_field(&a.out[f][q], v);
    }
}

template<typename T>
__device__ __forceinline__ void stage_update(const StageFullArgs<T>& a, int f, size_t q, T rhs, int kind){
    stage_update_v(a, f, q, rhs, kind, (a.stage == 0) ? T(0) : a.acc[f][q], a.nl[f][q], stage_lo(a, f, q));
}

// The eleven species (RHS_Jup_Turb.cpp:1273-1338) in a kernel of their own: they need only the
// centre velocities, the eddy viscosity and the density of the cell, so this kernel runs with few
// registers and many waves while the dynamics kernel carries the heavy part.
template<typename T>
__global__ __launch_bounds__(SF_BK * SF_BJ) void k_stage_species(const StageFullArgs<T> a,
        const GeoIJ<T>* __restrict__ geo){
    const int b = xcd_remap(blockIdx.x, a.nb);
    const int i = b % a.im;
    const int rest = b / a.im;
    const int kb = rest % a.nkb;
    const int jb = rest / a.nkb;
    const int k = kb * SF_BK + threadIdx.x;
    const int j = jb * SF_BJ + threadIdx.y;
    if(k >= a.km || j >= a.jm) return;
    const size_t q = ((size_t)i * a.jm + j) * a.pitch + k;
    const uint8_t fl = a.flags[q];
    const bool interior = i >= 1 && i <= a.im - 2 && j >= 3 && j <= a.jm - 4 && k <= a.km - 2 && !(fl & FL_SELF);
    if(!interior){
        #pragma unroll
        for(int f = ATJUP_F_H2O; f <= ATJUP_F_NH4SH; f++)
            if(a.bsrc[f]) a.out[f][q] = a.bsrc[f][q];
        return;
    }
    const size_t P = a.plane, R = (size_t)a.pitch;
    const int kw = (k == 0) ? a.km - 2 : k - 1;
    const int ke = (k == a.km - 2) ? 0 : k + 1;
    const size_t qkw = q - k + kw, qke = q - k + ke;
    const GeoIJ<T> G = geo[wave_uniform(i * a.jm + j)];     // a wavefront is one (i, j) row
    const T u = a.in[ATJUP_F_U][q], v = a.in[ATJUP_F_V][q], w = a.in[ATJUP_F_W][q];
    const T v_invrm = v * G.inv_rm;
    const T w_invrs = w * G.inv_rmsinthe;
    const T two_inv_rm = T(2) * G.inv_rm;
    const T rho_f = a.rho[q];
    const T nue_c = a.nue[q];
    const T nue_t = (a.turb_coupling != T(0) && isfinite(nue_c)) ? a.turb_coupling * dmax(T(0), nue_c) : T(0);
    const T nue_t_s = nue_t / T(0.9);
    // SP_GROUP species at a time: the group's loads, then its updates
    #pragma unroll
    for(int f0 = ATJUP_F_H2O; f0 <= ATJUP_F_NH4SH; f0 += SP_GROUP){
        SpeciesLoads<T> L[SP_GROUP];
        #pragma unroll
        for(int g = 0; g < SP_GROUP; g++)
            if(f0 + g <= ATJUP_F_NH4SH) L[g] = species_load(a, f0 + g, q, P, R, qke, qkw);
        #pragma unroll
        for(int g = 0; g < SP_GROUP; g++){
            const int f = f0 + g;
            if(f > ATJUP_F_NH4SH) break;
            const T rhs = species_rhs(a, f, L[g], G, u, v_invrm, w_invrs, two_inv_rm, nue_t_s, rho_f);
            stage_update_v(a, f, q, rhs, f == ATJUP_F_NH4SH ? 2 : 1, L[g].acc, L[g].nl, L[g].lo);
        }
    }
}

// TM: the closure (-1 off, 0 k-epsilon, 1 k-omega, 2 SST), one instantiation each
template<typename T, int TM>
__global__ __launch_bounds__(SF_BK * SF_BJ, SF_MIN_WAVES) void k_stage_full(const StageFullArgs<T> a,
        const GeoIJ<T>* __restrict__ geo){
    const int b = xcd_remap(blockIdx.x, a.nb);
    // levels fastest: the blocks resident on an XCD at one time cover whole columns, so the i +- 1
    // neighbours of the 17 fields a block reads are in flight in the same L2 (4 % faster per stage
    // on MI355X than k tiles fastest; the 4-field dry stage prefers k tiles fastest)
    const int i = b % a.im;
    const int rest = b / a.im;
    const int kb = rest % a.nkb;
    const int jb = rest / a.nkb;
    const int k = kb * SF_BK + threadIdx.x;
    const int j = jb * SF_BJ + threadIdx.y;
    if(k >= a.km || j >= a.jm) return;
    const size_t q = ((size_t)i * a.jm + j) * a.pitch + k;
    const uint8_t fl = a.flags[q];
    const bool interior = i >= 1 && i <= a.im - 2 && j >= 3 && j <= a.jm - 4 && k <= a.km - 2 && !(fl & FL_SELF);
    if(!interior){
        #pragma unroll
        for(int f = 0; f < NP; f++)
            if((f < ATJUP_F_H2O || f > ATJUP_F_NH4SH) && a.bsrc[f]) a.out[f][q] = a.bsrc[f][q];
        return;
    }
    const size_t P = a.plane, R = (size_t)a.pitch;
    const int kw = (k == 0) ? a.km - 2 : k - 1;
    const int ke = (k == a.km - 2) ? 0 : k + 1;
    const size_t qkw = q - k + kw, qke = q - k + ke;
    const GeoIJ<T> G = geo[wave_uniform(i * a.jm + j)];     // a wavefront is one (i, j) row
    const int stage = a.stage;
    auto update = [&](int f, T rhs, int kind){ stage_update(a, f, q, rhs, kind); };

    // ---- derivatives of t, u, v, w (RHS_Jup_Turb.cpp:180-254) ----
    T c_[4], d1r[4], d2r[4], d1t[4], d2t[4], d1p[4], d2p[4];
    #pragma unroll
    for(int f = 0; f < 4; f++){
        const Stencil7<T> s = load7(a.in[f], q, P, R, qke, qkw);
        c_[f] = s.c;
        d1r[f] = (s.ip - s.im) * a.inv_2dr;
        d2r[f] = (s.ip - T(2) * s.c + s.im) * a.inv_dr2;
        d1t[f] = (s.jp - s.jm) * a.inv_2dthe;
        d2t[f] = (s.jp - T(2) * s.c + s.jm) * a.inv_dthe2;
        d1p[f] = (s.kp - s.km) * a.inv_2dphi;
        d2p[f] = (s.kp - T(2) * s.c + s.km) * a.inv_dphi2;
    }
    const T u = c_[1], v = c_[2], w = c_[3];
    const T v_invrm = v * G.inv_rm;
    const T w_invrs = w * G.inv_rmsinthe;
    const T two_inv_rm = T(2) * G.inv_rm;
    const T nue_c = a.nue[q];

    // wall-mirrored pressure gradients (RHS_Jup_Turb.cpp:119-130)
    const T pdc = a.pd[q], phc = a.ph[q];
    const T dpdr   = (((fl & FL_IP) ? pdc : a.pd[q + P]) - ((fl & FL_IM) ? pdc : a.pd[q - P])) * a.inv_2dr;
    const T dpdthe = (((fl & FL_JP) ? pdc : a.pd[q + R]) - ((fl & FL_JM) ? pdc : a.pd[q - R])) * a.inv_2dthe;
    const T dpdphi = (((fl & FL_KP) ? pdc : a.pd[qke]) - ((fl & FL_KM) ? pdc : a.pd[qkw])) * a.inv_2dphi;
    const T dphdthe = (((fl & FL_JP) ? phc : a.ph[q + R]) - ((fl & FL_JM) ? phc : a.ph[q - R])) * a.inv_2dthe;
    const T dphdphi = (((fl & FL_KP) ? phc : a.ph[qke]) - ((fl & FL_KM) ? phc : a.ph[qkw])) * a.inv_2dphi;

    const T tq = c_[0];
    const T costhe = G.costhe, st = G.sinthe_true;
    const T Coriolis_the = a.c_m2o * costhe * w;
    const T Coriolis_phi = a.c_p2o * (+costhe * v + st * u);

    const T ps_c = a.ps[q];
    const T rho_f = a.rho[q];
    T dpstatdr = (a.ps[q + P] - a.ps[q - P]) * a.inv_2dr * a.to_nd;
    if(rho_f > T(0) && isfinite(rho_f)) dpstatdr *= a.r_mix / rho_f;
    const T pressure_t = a.cep * (u * (dpdr + dpstatdr) + v_invrm * dpdthe + w_invrs * dpdphi);
    const T transport_t = u * d1r[0] + v_invrm * d1t[0] + w_invrs * d1p[0];
    const T transport_u = u * d1r[1] + v_invrm * d1t[1] + w_invrs * d1p[1]
        - T(0) * (v * v + w * w) * G.inv_rm;
    const T transport_v = u * d1r[2] + v_invrm * d1t[2] + w_invrs * d1p[2]
        + (u * v - w * w * G.cotanthe) * G.inv_rm;
    const T transport_w = u * d1r[3] + v_invrm * d1t[3] + w_invrs * d1p[3]
        + (w * u + v * w * G.cotanthe) * G.inv_rm;
    const T diffusion_t = d2r[0] + d1r[0] * two_inv_rm + d2t[0] * G.inv_rm2
        + d1t[0] * G.costhe_inv_rm2sinthe + d2p[0] * G.inv_rm2sinthe2;
    const T diffusion_u = d2r[1] + T(2) * u * G.inv_rm2 + d2t[1] * G.inv_rm2
        + T(4) * d1r[1] * G.inv_rm + d1t[1] * G.costhe_inv_rm2sinthe
        + d2p[1] * G.inv_rm2sinthe2;
    const T diffusion_v = d2r[2] + d1r[2] * two_inv_rm + d2t[2] * G.inv_rm2
        + d1t[2] * G.costhe_inv_rm2sinthe - G.v_metric * v
        + d2p[2] * G.inv_rm2sinthe2
        + T(2) * d1t[1] * G.inv_rm2
        - d1p[3] * T(2) * costhe * G.inv_rm2sinthe2;
    const T diffusion_w = d2r[3] + d1r[3] * two_inv_rm + d2t[3] * G.inv_rm2
        + d1t[3] * G.costhe_inv_rm2sinthe - G.v_metric * w
        + d2p[3] * G.inv_rm2sinthe2
        + T(2) * d1p[1] * G.inv_rm2sinthe
        + d1p[2] * T(2) * costhe * G.inv_rm2sinthe2;
    const T dpdthe_term = dpdthe * G.inv_rm;
    const T dpdphi_term = dpdphi * G.inv_rmsinthe;
    const T dphdthe_term = dphdthe * G.inv_rm;
    const T dphdphi_term = dphdphi * G.inv_rmsinthe;

    // radiative heating and precipitation latent heat, capped at +-0.5 (RHS_Jup_Turb.cpp:646-718)
    T radiation_t = T(0);
    if(a.rad_coupling != T(0)){
        const T T_phys = tq * a.t_ref;
        const T P_phys = ps_c * T(1.0e5);
        const T rho = (rho_f > T(0) && isfinite(rho_f)) ? rho_f : ((T_phys > T(1)) ? P_phys / (a.R_mix * T_phys) : T(0));
        if(rho > T(0) && a.cp_mix > T(0)){
            radiation_t = a.rad_coupling * a.qrad[q] * a.L_rad / (rho * a.cp_mix * a.u_0 * a.t_ref);
            if(!isfinite(radiation_t)) radiation_t = T(0);
            else if(radiation_t > T(0.5)) radiation_t = T(0.5);
            else if(radiation_t < T(-0.5)) radiation_t = T(-0.5);
        }
    }
    T precip_t = T(0);
    if(a.precip_coupling != T(0)){
        const T rho_p = (rho_f > T(0) && isfinite(rho_f)) ? rho_f : a.r_mix;
        if(rho_p > T(0) && a.cp_mix > T(0)){
            precip_t = a.precip_coupling * a.qprec[q] * a.L_rad / (rho_p * a.cp_mix * a.u_0 * a.t_ref);
            if(!isfinite(precip_t)) precip_t = T(0);
            else if(precip_t > T(0.5)) precip_t = T(0.5);
            else if(precip_t < T(-0.5)) precip_t = T(-0.5);
        }
    }
    const T nue_t = (a.turb_coupling != T(0) && isfinite(nue_c)) ? a.turb_coupling * dmax(T(0), nue_c) : T(0);
    const T nue_t_s = nue_t / T(0.9);
    const T visc = a.inv_re + nue_t + a.wn[q];

    const T rhs_t = + pressure_t - transport_t + diffusion_t * (a.t_coef + nue_t_s) + radiation_t + precip_t;
    const T buoyancy_u = (tq > T(0))
        ? (T)(a.cbuoy * (a.g * (double)ps_c / (a.rR * (double)tq * a.t_ref_d) - a.brl[i])) : T(0);
    const T rhs_u = - dpdr + buoyancy_u - transport_u + diffusion_u * visc + a.centrifugal * G.cent_rad;
    const T rhs_v = - dpdthe_term - dphdthe_term - transport_v + diffusion_v * visc
        - a.Coriolis * Coriolis_the + a.centrifugal * G.cent_the;
    const T rhs_w = - dpdphi_term - dphdphi_term - transport_w + diffusion_w * visc
        - a.Coriolis * Coriolis_phi;
    update(ATJUP_F_T, rhs_t, 0);
    update(ATJUP_F_U, rhs_u, 0);
    update(ATJUP_F_V, rhs_v, 0);
    update(ATJUP_F_W, rhs_w, 0);
    if(stage == 3){
        a.aux_u[q] = rhs_u + dpdr;
        a.aux_v[q] = rhs_v + dpdthe_term;
        a.aux_w[q] = rhs_w + dpdphi_term;
    }

    // ---- k* and omega* (RHS_Jup_Turb.cpp:265-303 and :763-998, k-omega) ----
    {
        const Stencil7<T> K = load7(a.in[ATJUP_F_TKE], q, P, R, qke, qkw);
        const Stencil7<T> D = load7(a.in[ATJUP_F_DIS], q, P, R, qke, qkw);
        T d1rK = (K.ip - K.im) * a.inv_2dr, d2rK = (K.ip - T(2) * K.c + K.im) * a.inv_dr2;
        T d1tK = (K.jp - K.jm) * a.inv_2dthe, d2tK = (K.jp - T(2) * K.c + K.jm) * a.inv_dthe2;
        T d1pK = (K.kp - K.km) * a.inv_2dphi, d2pK = (K.kp - T(2) * K.c + K.km) * a.inv_dphi2;
        T d1rD = (D.ip - D.im) * a.inv_2dr, d2rD = (D.ip - T(2) * D.c + D.im) * a.inv_dr2;
        T d1tD = (D.jp - D.jm) * a.inv_2dthe, d2tD = (D.jp - T(2) * D.c + D.jm) * a.inv_dthe2;
        T d1pD = (D.kp - D.km) * a.inv_2dphi, d2pD = (D.kp - T(2) * D.c + D.km) * a.inv_dphi2;
        // Neumann at solid faces: the solid neighbour takes the cell's own value
        if(fl & (FL_IM | FL_IP)){
            const T ka = (fl & FL_IM) ? K.c : K.im, kb2 = (fl & FL_IP) ? K.c : K.ip;
            const T da = (fl & FL_IM) ? D.c : D.im, db2 = (fl & FL_IP) ? D.c : D.ip;
            d1rK = (kb2 - ka) * a.inv_2dr; d2rK = (kb2 - T(2) * K.c + ka) * a.inv_dr2;
            d1rD = (db2 - da) * a.inv_2dr; d2rD = (db2 - T(2) * D.c + da) * a.inv_dr2;
        }
        if(fl & (FL_JM | FL_JP)){
            const T ka = (fl & FL_JM) ? K.c : K.jm, kb2 = (fl & FL_JP) ? K.c : K.jp;
            const T da = (fl & FL_JM) ? D.c : D.jm, db2 = (fl & FL_JP) ? D.c : D.jp;
            d1tK = (kb2 - ka) * a.inv_2dthe; d2tK = (kb2 - T(2) * K.c + ka) * a.inv_dthe2;
            d1tD = (db2 - da) * a.inv_2dthe; d2tD = (db2 - T(2) * D.c + da) * a.inv_dthe2;
        }
        if(fl & (FL_KM | FL_KP)){
            const T ka = (fl & FL_KM) ? K.c : K.km, kb2 = (fl & FL_KP) ? K.c : K.kp;
            const T da = (fl & FL_KM) ? D.c : D.km, db2 = (fl & FL_KP) ? D.c : D.kp;
            d1pK = (kb2 - ka) * a.inv_2dphi; d2pK = (kb2 - T(2) * K.c + ka) * a.inv_dphi2;
            d1pD = (db2 - da) * a.inv_2dphi; d2pD = (db2 - T(2) * D.c + da) * a.inv_dphi2;
        }
        T rhs_k = T(0), rhs_d = T(0);
        if constexpr(TM >= 0){
            const T nue_here = isfinite(nue_c) ? dmax(T(0), nue_c) : T(0);
            const T tke0 = K.c, dis0 = D.c;
            const T tke_s = dmax(tke0, a.dis_min);
            const T dis_s = dmax(dis0, a.dis_min);
            const T dudr_s = d1r[1], dvdr_s = d1r[2], dwdr_s = d1r[3];
            const T dudthe_s = d1t[1] * G.inv_rm, dvdthe_s = d1t[2] * G.inv_rm, dwdthe_s = d1t[3] * G.inv_rm;
            const T dudphi_s = d1p[1] * G.inv_rmsinthe, dvdphi_s = d1p[2] * G.inv_rmsinthe;
            const T dwdphi_s = d1p[3] * G.inv_rmsinthe;
            const T dtkedthe_s = d1tK * G.inv_rm, ddisdthe_s = d1tD * G.inv_rm;
            const T dtkedphi_s = d1pK * G.inv_rmsinthe, ddisdphi_s = d1pD * G.inv_rmsinthe;
            const T grad_dot = d1rK * d1rD + dtkedthe_s * ddisdthe_s + dtkedphi_s * ddisdphi_s;
            T cnue = (TM == 0) ? T(0.028) * tke_s * tke_s / dis_s : tke_s / dis_s;
            cnue = dmin(cnue, a.nue_max);
            const T der = T(0.66667) * (dudr_s + dvdthe_s + dwdphi_s);
            const T P_full = dmax(T(0),
                  (cnue * (T(2) * dudr_s - der) - T(0.66667) * tke0) * dudr_s
                + (cnue * (dudthe_s + dvdr_s)) * dudthe_s
                + (cnue * (dudphi_s + dwdr_s)) * dudphi_s
                + (cnue * (T(2) * dvdthe_s - der) - T(0.66667) * tke0) * dvdthe_s
                + (cnue * (dvdr_s + dudthe_s)) * dvdr_s
                + (cnue * (dvdphi_s + dwdthe_s)) * dvdphi_s
                + (cnue * (T(2) * dwdphi_s - der) - T(0.66667) * tke0) * dwdphi_s
                + (cnue * (dwdr_s + dudphi_s)) * dwdr_s
                + (cnue * (dwdthe_s + dvdphi_s)) * dwdthe_s);
            T diffusion_tke_re, diffusion_dis_re, tke_src, dis_src;
            if constexpr(TM == 1){
                // k-omega, Wilcox 1988/2006
                const T sig_k = T(0.6), sig_w = T(0.5), bet_star = T(0.09), gam = T(0.52), bet_0 = T(0.0708);
                diffusion_tke_re = a.inv_re + sig_k * nue_here;
                diffusion_dis_re = a.inv_re + sig_w * nue_here;
                const T S11 = dudr_s, S22 = dvdthe_s, S33 = dwdphi_s;
                const T S12 = T(0.5) * (dudthe_s + dvdr_s);
                const T S13 = T(0.5) * (dudphi_s + dwdr_s);
                const T S23 = T(0.5) * (dvdphi_s + dwdthe_s);
                const T S_mag = sqrt(T(2) * (S11 * S11 + S22 * S22 + S33 * S33
                                             + T(2) * (S12 * S12 + S13 * S13 + S23 * S23)));
                const T W12 = dudthe_s - dvdr_s;
                const T W13 = dudphi_s - dwdr_s;
                const T W23 = dvdphi_s - dwdthe_s;
                const T Omega = sqrt(W12 * W12 + W13 * W13 + W23 * W23);
                const T chi_w = fabs(Omega * Omega * S_mag / pow(bet_star * dis_s, T(3)));
                const T f_bet = (T(1) + T(85) * chi_w) / (T(1) + T(100) * chi_w);
                const T bet_wc = bet_0 * f_bet;
                const T sig_d = (grad_dot <= T(0)) ? T(0) : T(0.125);
                const T P_k = dmin(P_full, T(20) * bet_star * tke_s * dis_s);
                const T Y_k = bet_star * tke_s * dis_s;
                const T P_w = gam * dis_s / tke_s * P_k;
                const T Y_w = bet_wc * dis_s * dis_s;
                const T D_w = sig_d / dis_s * grad_dot;
                tke_src = P_k - Y_k;
                dis_src = P_w - Y_w + D_w;
            } else {
                // the wall distance above the local obstacle surface, floored at one layer: the
                // layer heights are the float multiples i * (L_atm dr) of the reference
                const T L_m = a.L_rad;
                const double y_mount = (double)((float)a.itopo[(size_t)j * a.km + k] * a.lh_h) * 1.0e3;
                const double dz_layer = dmax(1.0, a.lthick[i >= 1 ? i - 1 : 0]);
                const T y_phys = (T)dmax((double)((float)i * a.lh_h) * 1.0e3 - y_mount, dz_layer);
                const T y_star = y_phys / L_m;
                if constexpr(TM == 0){
                    // k-epsilon, Chien 1982, with its near-wall damping
                    const T sig_k = T(1.0), sig_w = T(1.3), C_eps_1 = T(1.35), C_eps_2 = T(1.80);
                    const T nue_gas_phys = T(1.8e-5);
                    diffusion_tke_re = a.inv_re + nue_here / sig_k;
                    diffusion_dis_re = a.inv_re + nue_here / sig_w;
                    const T d_plus = y_phys * (T)a.vel_star[(size_t)j * a.km + k] / nue_gas_phys;
                    const T Re_T = tke_s * tke_s * a.u_0 * L_m / (dis_s * nue_gas_phys);
                    const T f_2 = T(1) - T(0.4) / T(1.8) * exp(-Re_T * Re_T / T(36));
                    const T y_star2 = y_star * y_star;
                    const T L_k = T(-2) * tke0 / y_star2;
                    const T L_w = T(-2) * dis0 / y_star2 * exp(T(-0.5) * d_plus);
                    const T P_k = P_full;
                    const T Y_k = dis0;
                    const T P_w = C_eps_1 * dis0 / tke_s * P_k;
                    const T Y_w = C_eps_2 * f_2 * dis0 * dis0 / tke_s;
                    tke_src = P_k - Y_k + L_k / a.re_turb;
                    dis_src = P_w - Y_w + L_w / a.re_turb;
                } else {
                    // k-omega SST, Menter 1994, with its F1 blending of the coefficients
                    const T bet_star = T(0.09), sig_k1 = T(1.176), sig_k2 = T(1.0), sig_w1 = T(2.0);
                    const T sig_w2 = T(1.168), bet1 = T(0.0333), bet2 = T(0.0368), gam1 = T(0.413), gam2 = T(0.2);
                    const T nue_air_nd = T(1.8e-5) / (a.u_0 * L_m);
                    const T CD_kw = dmax(T(2) * sig_w2 / dis_s * grad_dot, T(1.0e-20));
                    const T arg1 = dmin(
                        dmax(sqrt(dmax(tke0, T(0))) / (bet_star * dis_s * y_star),
                             T(500) * nue_air_nd / (y_star * y_star * dis_s)),
                        T(4) * sig_w2 * tke_s / (CD_kw * y_star * y_star));
                    const T F1 = tanh(pow(arg1, T(4)));
                    const T sig_k = F1 * sig_k1 + (T(1) - F1) * sig_k2;
                    const T sig_w = F1 * sig_w1 + (T(1) - F1) * sig_w2;
                    diffusion_tke_re = a.inv_re + nue_here / sig_k;
                    diffusion_dis_re = a.inv_re + nue_here / sig_w;
                    const T P_k = dmin(P_full, T(20) * bet_star * tke_s * dis_s);
                    const T Y_k = bet_star * tke_s * dis_s;
                    const T P_w = (F1 * gam1 + (T(1) - F1) * gam2) * P_k * dis_s / tke_s;
                    const T Y_w = (F1 * bet1 + (T(1) - F1) * bet2) * dis_s * dis_s;
                    const T D_w = T(2) * (T(1) - F1) * sig_w2 / dis_s * grad_dot;
                    tke_src = P_k - Y_k;
                    dis_src = P_w - Y_w + D_w;
                }
            }
            const T tke_src_max = T(20) * tke_s * dis_s;
            const T dis_src_max = T(20) * dis_s * dis_s;
            if(!isfinite(tke_src)) tke_src = T(0);
            if(!isfinite(dis_src)) dis_src = T(0);
            tke_src = dmax(-tke_src_max, dmin(tke_src_max, tke_src));
            dis_src = dmax(-dis_src_max, dmin(dis_src_max, dis_src));
            if(stage == 3){
                a.prod[q] = P_full;
                a.tsrc[q] = tke_src;
                a.dsrc[q] = dis_src;
            }
            const T transport_k = u * d1rK + v_invrm * d1tK + w_invrs * d1pK;
            const T transport_d = u * d1rD + v_invrm * d1tD + w_invrs * d1pD;
            const T diffusion_k = d2rK + d1rK * two_inv_rm + d2tK * G.inv_rm2
                + d1tK * G.costhe_inv_rm2sinthe + d2pK * G.inv_rm2sinthe2;
            const T diffusion_d = d2rD + d1rD * two_inv_rm + d2tD * G.inv_rm2
                + d1tD * G.costhe_inv_rm2sinthe + d2pD * G.inv_rm2sinthe2;
            rhs_k = -transport_k + diffusion_k * diffusion_tke_re + tke_src;
            rhs_d = -transport_d + diffusion_d * diffusion_dis_re + dis_src;
        }
        update(ATJUP_F_TKE, rhs_k, 3);
        update(ATJUP_F_DIS, rhs_d, 4);
    }

}

// =============================================================================================
// Turbulence<P>::run, k-omega with TURB_FREE (Turbulence.h:307-1012)
// =============================================================================================
// clamp-in pass (interior fluid cells) into the live buffers; every other cell is copied
template<typename T>
__global__ __launch_bounds__(256) void k_turb_clamp_in(const T* __restrict__ tke, const T* __restrict__ dis,
        T* __restrict__ ltke, T* __restrict__ ldis, const uint8_t* __restrict__ flags,
        int im, int jm, int km, int pitch, T dis_min){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    T kk = tke[q], dd = dis[q];
    if(i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2 && k <= km - 2 && !(flags[q] & FL_SELF)){
        kk = dmax(T(0), kk);
        dd = dmax(dis_min, dd);
    }
    ltke[q] = kk;
    ldis[q] = dd;
}

struct TurbConsts {
    double dr, dthe, dphi, nue_max, dis_min;
    double u_0, L_m;           // velocity scale [m/s], shell depth [m]
    double re_turb;            // the closure's vel_star_ref z_0 / nue_gas (Turbulence.h:245-251)
};

// compute_sources + compute_k_omega / compute_k_epsilon / compute_k_omega_SST (Turbulence.h:633-1106),
// TM the model (0 k-epsilon, 1 k-omega, 2 SST): nue, prod, tke_source, dis_source of the interior
// fluid cells from the clamped k*, omega* and the velocity field. With TURB_FREE the wall distance
// of k-epsilon and SST is free_wall_dist() = 1e20 m.
template<typename T, int TM>
__global__ __launch_bounds__(256) void k_turb_sources(const T* __restrict__ tke, const T* __restrict__ dis,
        const T* __restrict__ U, const T* __restrict__ V, const T* __restrict__ W,
        T* __restrict__ nue, T* __restrict__ prod, T* __restrict__ tsrc, T* __restrict__ dsrc,
        const uint8_t* __restrict__ flags, const double* __restrict__ tabs, const double* __restrict__ vel_star,
        int im, int jm, int km, int pitch, size_t plane, const TurbConsts tc){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1, i = blockIdx.z + 1;
    if(k > km - 2 || j > jm - 2 || i > im - 2) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const uint8_t fl = flags[q];
    if(fl & FL_SELF) return;
    const size_t P = plane, R = (size_t)pitch;
    const int kw = (k == 0) ? km - 2 : k - 1;
    const int ke = (k == km - 2) ? 0 : k + 1;
    const size_t qip = q + P, qim = q - P, qjp = q + R, qjm = q - R, qkp = q - k + ke, qkm = q - k + kw;
    const double* rad = tabs;
    const double* sin_the = tabs + im;
    const double* cos_the = tabs + im + 2 * jm;
    const T rm = (T)rad[i];
    const T exp_rm = T(1);
    T sinthe = (T)sin_the[j];
    if(sinthe == T(0)) sinthe = T(1.0e-5);
    const T costhe = (T)cos_the[j];
    const T cotanthe = costhe / sinthe;
    const T inv_rm = T(1) / rm;
    const T rmsinthe = rm * sinthe;
    const T dr = (T)tc.dr, dthe = (T)tc.dthe, dphi = (T)tc.dphi;
    const T inv_2dr = T(1) / (T(2) * dr);
    const T inv_2dthe = T(1) / (T(2) * dthe);
    const T inv_2dphi = T(1) / (T(2) * dphi);
    const T inv_rm2dthe = inv_2dthe / rm;
    const T inv_rmsinthe2dphi = inv_2dphi / rmsinthe;
    const T Uq = U[q], Vq = V[q], Wq = W[q];
    const T Uip = U[qip], Uim = U[qim], Ujp = U[qjp], Ujm = U[qjm], Ukp = U[qkp], Ukm = U[qkm];
    const T Vip = V[qip], Vim = V[qim], Vjp = V[qjp], Vjm = V[qjm], Vkp = V[qkp], Vkm = V[qkm];
    const T Wip = W[qip], Wim = W[qim], Wjp = W[qjp], Wjm = W[qjm], Wkp = W[qkp], Wkm = W[qkm];
    const T dudr = (Uip - Uim) * inv_2dr * exp_rm;
    const T dvdr = (Vip - Vim) * inv_2dr * exp_rm;
    const T dwdr = (Wip - Wim) * inv_2dr * exp_rm;
    const T dudthe = (Ujp - Ujm) * inv_rm2dthe;
    const T dvdthe = (Vjp - Vjm) * inv_rm2dthe;
    const T dwdthe = (Wjp - Wjm) * inv_rm2dthe;
    const T dudphi = (Ukp - Ukm) * inv_rmsinthe2dphi;
    const T dvdphi = (Vkp - Vkm) * inv_rmsinthe2dphi;
    const T dwdphi = (Wkp - Wkm) * inv_rmsinthe2dphi;
    const bool nb = fl & FL_IM;
    const bool njm = fl & FL_JM, njp = fl & FL_JP, nkm = fl & FL_KM, nkp = fl & FL_KP;
    const T tq = tke[q], dq = dis[q];
    const T tke_im1 = nb ? tq : tke[qim];
    const T dis_im1 = nb ? dq : dis[qim];
    const T tke_jm1 = njm ? tq : tke[qjm];
    const T tke_jp1 = njp ? tq : tke[qjp];
    const T dis_jm1 = njm ? dq : dis[qjm];
    const T dis_jp1 = njp ? dq : dis[qjp];
    const T tke_km1 = nkm ? tq : tke[qkm];
    const T tke_kp1 = nkp ? tq : tke[qkp];
    const T dis_km1 = nkm ? dq : dis[qkm];
    const T dis_kp1 = nkp ? dq : dis[qkp];
    const T tke_ip = tke[qip], dis_ip = dis[qip];
    const T dtkedr = (tke_ip - tke_im1) * inv_2dr * exp_rm;
    const T ddisdr = (dis_ip - dis_im1) * inv_2dr * exp_rm;
    T cnue;
    {
        const T dis_here = dmax(dq, (T)tc.dis_min);
        const T tke_here = dmax(tq, T(0));
        cnue = (TM == 0) ? T(0.028) * tke_here * tke_here / dis_here : tke_here / dis_here;
        cnue = dmin(cnue, (T)tc.nue_max);
    }
    const T u_r = Uq * inv_rm, v_r = Vq * inv_rm, w_r = Wq * inv_rm;
    const T v_cot_r = v_r * cotanthe, w_cot_r = w_r * cotanthe;
    const T g_rr = dudr, g_rt = dudthe - v_r, g_rp = dudphi - w_r;
    const T g_tr = dvdr, g_tt = dvdthe + u_r, g_tp = dvdphi - w_cot_r;
    const T g_pr = dwdr, g_pt = dwdthe, g_pp = dwdphi + u_r + v_cot_r;
    const T der = T(0.66667) * (g_rr + g_tt + g_pp);
    const T pr = dmax(T(0),
          (cnue * (T(2) * g_rr - der) - T(0.66667) * tq) * g_rr
        + (cnue * (g_rt + g_tr)) * g_rt
        + (cnue * (g_rp + g_pr)) * g_rp
        + (cnue * (T(2) * g_tt - der) - T(0.66667) * tq) * g_tt
        + (cnue * (g_tr + g_rt)) * g_tr
        + (cnue * (g_tp + g_pt)) * g_tp
        + (cnue * (T(2) * g_pp - der) - T(0.66667) * tq) * g_pp
        + (cnue * (g_pr + g_rp)) * g_pr
        + (cnue * (g_pt + g_tp)) * g_pt);
    prod[q] = pr;
    const T W12 = g_rt - g_tr, W13 = g_rp - g_pr, W23 = g_tp - g_pt;
    const T Omega_mag = sqrt(W12 * W12 + W13 * W13 + W23 * W23);
    if constexpr(TM == 0){
        // compute_k_epsilon (Turbulence.h:886-918), Chien's damping in its free-flow form
        const T C_nue = T(0.028), C_eps_1 = T(1.35), C_eps_2 = T(1.80), nue_air = T(1.8e-5);
        const T dis_min = (T)tc.dis_min, L_m = (T)tc.L_m;
        const T y_phys = T(1.0e20);
        const T y_star = y_phys / L_m;
        T nu = C_nue * tq * tq / dmax(dq, dis_min);
        const T d_plus = y_phys * (T)vel_star[(size_t)j * km + k] / nue_air;
        const T f_nue = T(1) - exp(T(-0.0115) * d_plus);
        const T Re_T = tq * tq * (T)tc.u_0 * L_m / dmax(dq * nue_air, dis_min * nue_air);
        const T f_2 = T(1) - T(0.4) / T(1.8) * exp(-Re_T * Re_T / T(36));
        nu *= f_nue;
        nue[q] = nu;
        const T y_star2 = y_star * y_star;
        const T L_k = T(-2) * tq / y_star2;
        const T L_w = T(-2) * dq / y_star2 * exp(-d_plus / T(2));
        const T P_k = pr;
        const T Y_k = dq;
        const T tke_safe = dmax(tq, dis_min);
        const T P_w = C_eps_1 * dq / tke_safe * P_k;
        const T Y_w = C_eps_2 * f_2 * dq * dq / tke_safe;
        tsrc[q] = P_k - Y_k + L_k / (T)tc.re_turb;
        dsrc[q] = P_w - Y_w + L_w / (T)tc.re_turb;
        return;
    }
    // compute_k_omega (Turbulence.h:920-1012); exp_rm = 1 / (rm + 1) in its derivatives, which
    // compute_k_omega_SST shares
    const T bet_star = T(0.09), gam = T(0.52), C_lim = T(0.875), bet_0 = T(0.0708);
    const T erm = T(1) / (rm + T(1));
    const T dtkedr_neu = nb ? (tke_ip - tq) * (T(1) / (T(2) * dr)) * erm : dtkedr;
    const T ddisdr_neu = nb ? (dis_ip - dq) * (T(1) / (T(2) * dr)) * erm : ddisdr;
    const T dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * T(2) * dthe);
    const T ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * T(2) * dthe);
    const T dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * T(2) * dphi);
    const T ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * T(2) * dphi);
    if constexpr(TM == 2){
        // compute_k_omega_SST (Turbulence.h:1021-1106), Menter 1994
        const T a1 = T(0.31), bet1 = T(0.0333), bet2 = T(0.0368), gam1 = T(0.413), gam2 = T(0.2);
        const T sig_w2 = T(1.168);
        const T L_m = (T)tc.L_m;
        const T nue_air_nd = T(1.8e-5) / ((T)tc.u_0 * L_m);
        const T y_star = T(1.0e20) / L_m;
        const T grad_dot = dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
        const T CD_kw = dmax(T(2) * sig_w2 / dmax(dq, T(1.0e-20)) * grad_dot, T(1.0e-20));
        const T dis_safe = dmax(dq, T(1.0e-20));
        const T arg1 = dmin(
            dmax(sqrt(tq) / (bet_star * dis_safe * y_star), T(500) * nue_air_nd / (y_star * y_star * dis_safe)),
            T(4) * sig_w2 * tq / (CD_kw * y_star * y_star));
        const T arg2 = dmax(T(2) * sqrt(tq) / (bet_star * dis_safe * y_star),
                            T(500) * nue_air_nd / (y_star * y_star * dis_safe));
        const T F1 = tanh(pow(arg1, T(4)));
        const T F2 = tanh(pow(arg2, T(2)));
        nue[q] = a1 * tq / dmax(a1 * dis_safe, Omega_mag * F2);
        const T tke_safe = dmax(tq, T(1.0e-20));
        const T P_k = dmin(pr, T(20) * bet_star * tke_safe * dis_safe);
        const T Y_k = bet_star * tke_safe * dis_safe;
        const T P_w = (F1 * gam1 + (T(1) - F1) * gam2) * P_k * dis_safe / tke_safe;
        const T Y_w = (F1 * bet1 + (T(1) - F1) * bet2) * dis_safe * dis_safe;
        const T D_w = T(2) * (T(1) - F1) * sig_w2 / dis_safe * grad_dot;
        tsrc[q] = P_k - Y_k;
        dsrc[q] = P_w - Y_w + D_w;
        return;
    }
    const T dudr_k = (Uip - Uim) / (T(2) * dr) / (rm + T(1));
    const T dvdthe_k = (Vjp - Vjm) / (rm * T(2) * dthe);
    const T dwdphi_k = (Wkp - Wkm) / (rmsinthe * T(2) * dphi);
    const T dudthe_u = (Ujp - Ujm) / (rm * T(2) * dthe);
    const T dudphi_u = (Ukp - Ukm) / (rmsinthe * T(2) * dphi);
    const T dvdr_u = (Vip - Vim) / (T(2) * dr) / (rm + T(1));
    const T dwdr_u = (Wip - Wim) / (T(2) * dr) / (rm + T(1));
    const T dvdphi_u = (Vkp - Vkm) / (rmsinthe * T(2) * dphi);
    const T dwdthe_u = (Wjp - Wjm) / (rm * T(2) * dthe);
    const T S11 = dudr_k, S22 = dvdthe_k, S33 = dwdphi_k;
    const T S12 = T(0.5) * (dudthe_u + dvdr_u);
    const T S13 = T(0.5) * (dudphi_u + dwdr_u);
    const T S23 = T(0.5) * (dvdphi_u + dwdthe_u);
    const T S_mag = sqrt(T(2) * (S11 * S11 + S22 * S22 + S33 * S33 + T(2) * (S12 * S12 + S13 * S13 + S23 * S23)));
    const T w_hat = dmax(dq, C_lim * S_mag / sqrt(bet_star));
    const T dot = dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
    T sig_d = dot;
    sig_d = (sig_d <= T(0)) ? T(0) : T(0.125);
    const T chi_w = fabs(Omega_mag * Omega_mag * S_mag / pow(bet_star * dmax(dq, T(1.0e-20)), T(3)));
    const T f_bet = (T(1) + T(85) * chi_w) / (T(1) + T(100) * chi_w);
    const T bet_wc = bet_0 * f_bet;
    nue[q] = tq / dmax(w_hat, T(1.0e-20));
    const T P_k = dmin(pr, T(20) * bet_star * tq * dq);
    const T Y_k = bet_star * tq * dq;
    const T P_w = gam * dq / dmax(tq, T(1.0e-20)) * P_k;
    const T Y_w = bet_wc * dq * dq;
    const T D_w = sig_d / dmax(dq, T(1.0e-20)) * dot;
    tsrc[q] = P_k - Y_k;
    dsrc[q] = P_w - Y_w + D_w;
}

// clamp-out (interior fluid), zero_land_cells and clamp_nue on every cell but the i = 0 plane's
// wall copy, which k_turb_wall_plane writes first (Turbulence.h:378-420, :623-660, :1040-1106)
template<typename T>
__global__ __launch_bounds__(256) void k_turb_post(T* __restrict__ tke, T* __restrict__ dis, T* __restrict__ nue,
        T* __restrict__ prod, T* __restrict__ tsrc, T* __restrict__ dsrc, const uint8_t* __restrict__ flags,
        int im, int jm, int km, int pitch, T k_bg, T dis_min, T nue_max){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km || i == 0) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const bool solid = flags[q] & FL_SELF;
    if(solid){
        tke[q] = T(0); dis[q] = T(0); nue[q] = T(0); prod[q] = T(0); tsrc[q] = T(0); dsrc[q] = T(0);
        return;
    }
    // only the values the clamps change are written back
    const T n0 = nue[q];
    T n = n0;
    if(i <= im - 2 && j >= 1 && j <= jm - 2 && k <= km - 2){
        const T k0 = tke[q], d0 = dis[q];
        const T kk = dmax(k_bg, k0);
        const T dd = dmax(dis_min, d0);
        if(!same_bits(kk, k0)) tke[q] = kk;
        if(!same_bits(dd, d0)) dis[q] = dd;
        n = (n < T(0)) ? T(0) : ((n > nue_max) ? nue_max : n);
        const T dsm = T(20) * dd * dd;
        const T ds = dsrc[q];
        const T dsc = (ds < -dsm) ? -dsm : ((ds > dsm) ? dsm : ds);
        if(!same_bits(dsc, ds)) dsrc[q] = dsc;
        const T tsm = T(20) * kk * dd;
        const T ts = tsrc[q];
        const T tsc = (ts < -tsm) ? -tsm : ((ts > tsm) ? tsm : ts);
        if(!same_bits(tsc, ts)) tsrc[q] = tsc;
    }
    const T nc = ((n < T(0)) ? T(0) : ((n > nue_max) ? nue_max : n)) * T(1);
    if(!same_bits(nc, n0)) nue[q] = nc;
}

// the i = 0 plane: apply_wall_bc copies k*, omega*, nue from i = 1 after its clamp-out; then
// zero_land_cells and clamp_nue of the plane itself
template<typename T>
__global__ void k_turb_wall_plane(T* __restrict__ tke, T* __restrict__ dis, T* __restrict__ nue,
        T* __restrict__ prod, T* __restrict__ tsrc, T* __restrict__ dsrc, const uint8_t* __restrict__ flags,
        int im, int jm, int km, int pitch, size_t plane, T k_bg, T dis_min, T nue_max, int clamp_src){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q = (size_t)j * pitch + k, q1 = q + plane;
    T kk = tke[q1], dd = dis[q1], n = nue[q1];
    if(clamp_src && 1 <= im - 2 && j >= 1 && j <= jm - 2 && k <= km - 2 && !(flags[q1] & FL_SELF)){
        kk = dmax(k_bg, kk);
        dd = dmax(dis_min, dd);
        n = (n < T(0)) ? T(0) : ((n > nue_max) ? nue_max : n);
    }
    if(!clamp_src){ tke[q] = kk; dis[q] = dd; nue[q] = n; return; }
    if(flags[q] & FL_SELF){
        tke[q] = T(0); dis[q] = T(0); nue[q] = T(0); prod[q] = T(0); tsrc[q] = T(0); dsrc[q] = T(0);
        return;
    }
    tke[q] = kk;
    dis[q] = dd;
    nue[q] = ((n < T(0)) ? T(0) : ((n > nue_max) ? nue_max : n)) * T(1);
}

template<typename T>
struct TurbArrays { T* a[6]; };

// seam_copy of the closure's arrays: k = km-1 is the copy of k = 0
template<typename T>
__global__ void k_turb_seam(const TurbArrays<T> L, int im, int jm, int km, int pitch){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(j >= jm) return;
    const size_t r0 = ((size_t)i * jm + j) * pitch;
    #pragma unroll
    for(int f = 0; f < 6; f++) L.a[f][r0 + km - 1] = L.a[f][r0];
}

// fill_lid_corners, first the lid plane i = im-1 (rows 1..jm-2) ...
template<typename T>
__global__ void k_turb_lid(const TurbArrays<T> L, int im, int jm, int km, int pitch, size_t plane){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    if(k >= km || j > jm - 2) return;
    const size_t q = ((size_t)(im - 1) * jm + j) * pitch + k;
    #pragma unroll
    for(int f = 0; f < 6; f++) L.a[f][q] = L.a[f][q - plane];
}

// ... then the pole rows j = 0 and jm-1 of every level
template<typename T>
__global__ void k_turb_pole_rows(const TurbArrays<T> L, int im, int jm, int km, int pitch){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(k >= km) return;
    const size_t qs = (size_t)i * jm * pitch + k;
    const size_t qn = qs + (size_t)(jm - 1) * pitch;
    #pragma unroll
    for(int f = 0; f < 6; f++){
        L.a[f][qs] = L.a[f][qs + pitch];
        L.a[f][qn] = L.a[f][qn - pitch];
    }
}

// the wall-function omega* at the first fluid cell of every column with an obstacle
// (cJupiterModel.cpp:356-374), with the friction velocity of compute_vel_star
template<typename T>
__global__ void k_turb_wall_omega(T* __restrict__ dis, const int* __restrict__ itopo, const double* __restrict__ vel_star,
        const double* __restrict__ lthick, int im, int jm, int km, int pitch, double nd_omega){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const int i_w = itopo[(size_t)j * km + k];
    if(i_w <= 0 || i_w >= im - 1) return;
    const double nue_gas_phys = 1.8e-5, bet_star = 0.09, bet_1 = 0.0333, kappa = 0.41;
    const double y1 = dmax(lthick[max(i_w - 1, 0)], 1.0);
    const double y = 0.5 * y1;
    const double om_vis = 60.0 * nue_gas_phys / (bet_1 * y1 * y1);
    const double om_log = vel_star[(size_t)j * km + k] / (sqrt(bet_star) * kappa * y);
    dis[((size_t)i_w * jm + j) * pitch + k] = (T)(sqrt(om_vis * om_vis + om_log * om_log) * nd_omega);
}

__global__ void k_fill_double(double* __restrict__ x, size_t n, double v){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) x[q] = v;
}

// the closure's share of bcSolidGround: safe_clamp of its arrays on every cell (BC_Jup.h:627-655)
template<typename T>
__global__ __launch_bounds__(256) void k_turb_sanitize(T* __restrict__ tke, T* __restrict__ dis, T* __restrict__ nue,
        T* __restrict__ prod, T* __restrict__ tsrc, T* __restrict__ dsrc, size_t n, T tke_max, T nue_max){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    const T prod_max = T(1.0e4), dis_max = T(1.0e6);
    // only the cells the clamp changes are written back (most are in range)
    auto clamp = [&](T* x, T lo, T hi){
        const T v = x[q], c = safe_clamp_d(v, lo, hi);
        if(!same_bits(c, v)) x[q] = c;
    };
    clamp(tke, T(0), tke_max);
    clamp(dis, T(1.0e-10), dis_max);
    clamp(nue, T(0), nue_max);
    clamp(prod, T(0), prod_max);
    clamp(tsrc, -prod_max, prod_max);
    clamp(dsrc, -prod_max, prod_max);
}

// =============================================================================================
// SaturationAdjustment<P>::run for one gas (SaturationAdjustment.h:255-746) with the ATJUP
// defaults CONSERVE 1, NEWTON 1, EVAP 1, BERGERON 2, DRY_EXIT 1, ICE_ENTRY 0, FREEZE 0,
// LOCAL_RHO 1, and the final positivity pass (:673-685) fused in.
// =============================================================================================
struct SatParams {
    double t_0, t_00, ep, lv, ls, C, L0, R, del_alf, del_bet, C_i, L0_i, del_alf_i, del_bet_i;
    double cp_mix, t_ref, r_mix;
};

// D: double for the adjustment itself, Dual for its frozen-branch linearisation (comparisons
// and the iteration's exits follow the values, so the derivative is that of the branches taken)
template<typename D>
__device__ __forceinline__ D svp_d(D T, double C, double L0, double R, double da, double db){
    return dual_exp(C + (-L0 / T + da * dual_log(T) + db * T) / (1e-3 * R));      // ATPhys.h:49-52
}

template<typename D>
__device__ void satadj_cell(D& cv, D& ccl, D& cic, D& tq, D p_u, D rf, const SatParams& P){
    const double t_0 = P.t_0, t_00 = P.t_00, ep = P.ep, lv = P.lv, ls = P.ls;
    const double t_range_inv = 1.0 / (t_0 - t_00);
    const double Lf = ls - lv;
    const double cp_mix = P.cp_mix, t_ref = P.t_ref;
    const int iter_prec_end = 15;
    const double q_diff_min = 1.0e-3;
    D t_u = tq * t_ref;
    if(!(t_u > 0.0) || !(p_u > 0.0)) return;
    const D rho_at = (rf > 0.0 && isfinite(rf)) ? rf : D(P.r_mix);
    if(t_u > t_0){
        if(cic > 0.0){
            D dm = cic;
            if(Lf > 0.0){
                const D hc = cp_mix * rho_at;
                const D cap = hc * (t_u - t_0) / Lf;
                if(dm > cap) dm = cap;
                t_u -= Lf * dm / hc;
                tq = t_u / t_ref;
            }
            cic -= dm;
            ccl += dm;
        }
    }
    if(cv < 0.0) cv = 0.0;
    if(ccl < 0.0) ccl = 0.0;
    if(cic < 0.0) cic = 0.0;
    const D rho_c = rho_at;
    const D E_Rain_0 = svp_d(t_u, P.C, P.L0, P.R, P.del_alf, P.del_bet);
    const D q_Rain_0 = rho_c * ep * E_Rain_0 / p_u;
    if(q_Rain_0 <= 0.0) return;
    D q_entry = q_Rain_0;
    if(ccl <= 0.0 && cic <= 0.0 && t_u < t_0)
        q_entry = rho_c * ep * svp_d(t_u, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
    if(cv <= q_entry){
        const D q_c0 = ccl, q_i0 = cic;
        if(q_c0 <= 0.0 && q_i0 <= 0.0) return;
        const D q_Ice_0 = rho_c * ep * svp_d(t_u, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
        D q_t0;
        if(q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
        else if(q_i0 <= 0.0)         q_t0 = q_Rain_0;
        else                         q_t0 = q_Ice_0;
        if(fabs(cv - q_t0) <= q_diff_min * q_t0) return;
    }
    D q_v_b = cv, q_c_b = ccl, q_i_b = cic;
    D Tt = t_u;
    D q_v_hyp = q_v_b;
    D q_rain_now = q_Rain_0;
    for(int itr = 1; itr <= iter_prec_end; itr++){
        D CND = (Tt - t_00) * t_range_inv;
        D DEP = (t_0 - Tt) * t_range_inv;
        if(Tt <= t_00){ CND = 0.0; DEP = 1.0; }
        if(Tt >= t_0){ CND = 1.0; DEP = 0.0; }
        const bool berg = Tt < t_0 && q_v_b <= q_rain_now && (q_i_b > 0.0 || q_c_b <= 0.0);
        if(berg){
            CND = 0.0; DEP = 1.0;
            if(q_c_b > 0.0){
                Tt -= lv * q_c_b / (cp_mix * rho_c);
                q_v_b += q_c_b;
                q_c_b = 0.0;
            }
        }
        D d_q_v = q_v_hyp - q_v_b;
        D d_q_c = -d_q_v * CND;
        D d_q_i = -d_q_v * DEP;
        if(d_q_v < -q_v_b){ d_q_v = -q_v_b; d_q_c = -d_q_v * CND; d_q_i = -d_q_v * DEP; }
        if(q_c_b + d_q_c < 0.0){ const D ex = -(q_c_b + d_q_c); d_q_c = -q_c_b; d_q_i -= ex; }
        if(q_i_b + d_q_i < 0.0){ const D ex = -(q_i_b + d_q_i); d_q_i = -q_i_b; d_q_c -= ex;
                                 if(q_c_b + d_q_c < 0.0) d_q_c = -q_c_b; }
        d_q_v = -(d_q_c + d_q_i);
        Tt += (lv * d_q_c + ls * d_q_i) / (cp_mix * rho_c);
        q_v_b += d_q_v;
        q_c_b += d_q_c;
        q_i_b += d_q_i;
        if(q_v_b < 0.0) q_v_b = 0.0;
        if(q_c_b < 0.0) q_c_b = 0.0;
        if(q_i_b < 0.0) q_i_b = 0.0;
        const D E_Rain = svp_d(Tt, P.C, P.L0, P.R, P.del_alf, P.del_bet);
        const D E_Ice = svp_d(Tt, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i);
        const D q_Rain = rho_c * ep * E_Rain / p_u;
        const D q_Ice = rho_c * ep * E_Ice / p_u;
        q_rain_now = q_Rain;
        if(q_c_b > 0.0 && q_i_b > 0.0) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
        else if(q_c_b <= 0.0 && q_i_b <= 0.0 && Tt < t_0) q_v_hyp = q_Ice;
        else if(q_i_b == 0.0) q_v_hyp = q_Rain;
        else                  q_v_hyp = q_Ice;
        if(Tt >= t_0){
            if(q_i_b > 0.0){
                D dm = q_i_b;
                if(Lf > 0.0){
                    const D hc = cp_mix * rho_c;
                    const D cap = hc * (Tt - t_0) / Lf;
                    if(dm > cap) dm = cap;
                    Tt -= Lf * dm / hc;
                }
                q_i_b -= dm;
                q_c_b += dm;
            }
        }
        const D q_diff = fabs(q_v_b - q_v_hyp) / (q_v_hyp + 1e-20);
        if(q_diff <= q_diff_min) break;
        if(q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;
        const double dT = 0.1;
        const D q_Rain_d = rho_c * ep * svp_d(Tt + dT, P.C, P.L0, P.R, P.del_alf, P.del_bet) / p_u;
        const D q_Ice_d = rho_c * ep * svp_d(Tt + dT, P.C_i, P.L0_i, P.R, P.del_alf_i, P.del_bet_i) / p_u;
        D q_hyp_d;
        if(q_c_b > 0.0 && q_i_b > 0.0) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
        else if(q_c_b <= 0.0 && q_i_b <= 0.0 && Tt < t_0) q_hyp_d = q_Ice_d;
        else if(q_i_b == 0.0) q_hyp_d = q_Rain_d;
        else                  q_hyp_d = q_Ice_d;
        D CNDn = (Tt - t_00) * t_range_inv;
        if(Tt <= t_00) CNDn = 0.0;
        if(Tt >= t_0) CNDn = 1.0;
        if(berg) CNDn = 0.0;
        const D Gn = (lv * CNDn + ls * (1.0 - CNDn)) / (cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
        D wgt = 1.0 / (1.0 + dmax(Gn, D(0.0)));
        if(wgt < 0.05) wgt = 0.05;
        q_v_hyp = q_v_b + wgt * (q_v_hyp - q_v_b);
    }
    cv = q_v_b;
    ccl = q_c_b;
    cic = q_i_b;
    tq = Tt / t_ref;
}

template<typename T>
__global__ __launch_bounds__(256) void k_satadj(T* __restrict__ c, T* __restrict__ cloud, T* __restrict__ ice,
        T* __restrict__ t, const T* __restrict__ ps, const T* __restrict__ rho, const uint8_t* __restrict__ flags,
        int jm, int km, int pitch, const SatParams P, float* __restrict__ lo_t, float* __restrict__ lo_c,
        float* __restrict__ lo_cl, float* __restrict__ lo_ic){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    // the refined form's t and species are hi + lo
    double cv = pair_get(c, lo_c, q), ccl = pair_get(cloud, lo_cl, q), cic = pair_get(ice, lo_ic, q);
    double tq = pair_get(t, lo_t, q);
    const double cv0 = cv, ccl0 = ccl, cic0 = cic;
    const double t0 = tq;
    if(!(flags[q] & FL_SELF)) satadj_cell(cv, ccl, cic, tq, (double)ps[q], (double)rho[q], P);
    if(cv <= 0.0) cv = 0.0;
    if(ccl <= 0.0) ccl = 0.0;
    if(cic <= 0.0) cic = 0.0;
    if(!lo_c){                                     // stored where the bits changed (cv0 is c[q] widened)
        if(!same_bits((T)cv, (T)cv0)) c[q] = (T)cv;
        if(!same_bits((T)ccl, (T)ccl0)) cloud[q] = (T)ccl;
        if(!same_bits((T)cic, (T)cic0)) ice[q] = (T)cic;
    } else {                                       // pairs: rewritten where the value changed
        if(!(cv == cv0)) pair_put(c, lo_c, q, cv);
        if(!(ccl == ccl0)) pair_put(cloud, lo_cl, q, ccl);
        if(!(cic == cic0)) pair_put(ice, lo_ic, q, cic);
    }
    if(tq != t0 || !(tq == tq)){
        if(lo_t){
            const float h = (float)tq;
            t[q] = h;
            lo_t[q] = (float)(tq - (double)h);
        } else {
            t[q] = (T)tq;
        }
    }
}

// The three adjustments of the physics block (H2O, NH3, CH4 in that order) in one pass: a cell's
// adjustments depend on no other cell, so each thread runs the three in sequence on its cell with
// t, p_stat, rho and the flag read once. What a gas's adjustment reads of t is what the per-gas
// kernel would have stored before it (the pair split again in the refined form), so the stored
// results are those of three k_satadj launches.
template<typename T>
struct Sat3Args {
    SatParams P[3];
    T* sp[3][3];                 // [gas][vapour, cloud, ice]
    float* lo[3][3];             // their low words in the refined form (else null)
    T* t; float* lo_t;
    const T* ps; const T* rho; const uint8_t* flags;
    int jm, km, pitch;
};

template<typename T>
__global__ __launch_bounds__(256) void k_satadj3(const Sat3Args<T> A){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= A.km) return;
    const size_t q = ((size_t)i * A.jm + j) * A.pitch + k;
    const bool solid = A.flags[q] & FL_SELF;
    const double ps = (double)A.ps[q], rho = (double)A.rho[q];
    double tq = pair_get(A.t, A.lo_t, q);
    bool t_dirty = false;
    float th = 0.0f, tl = 0.0f;                   // the refined form's stored pair of t
    #pragma unroll 1
    for(int g = 0; g < 3; g++){
        T* c = A.sp[g][0]; T* cloud = A.sp[g][1]; T* ice = A.sp[g][2];
        float* lo_c = A.lo[g][0]; float* lo_cl = A.lo[g][1]; float* lo_ic = A.lo[g][2];
        double cv = pair_get(c, lo_c, q), ccl = pair_get(cloud, lo_cl, q), cic = pair_get(ice, lo_ic, q);
        const double cv0 = cv, ccl0 = ccl, cic0 = cic;
        const double t0 = tq;
        if(!solid) satadj_cell(cv, ccl, cic, tq, ps, rho, A.P[g]);
        if(cv <= 0.0) cv = 0.0;
        if(ccl <= 0.0) ccl = 0.0;
        if(cic <= 0.0) cic = 0.0;
        if(!lo_c){
            if(!same_bits((T)cv, (T)cv0)) c[q] = (T)cv;
            if(!same_bits((T)ccl, (T)ccl0)) cloud[q] = (T)ccl;
            if(!same_bits((T)cic, (T)cic0)) ice[q] = (T)cic;
        } else {
            if(!(cv == cv0)) pair_put(c, lo_c, q, cv);
            if(!(ccl == ccl0)) pair_put(cloud, lo_cl, q, ccl);
            if(!(cic == cic0)) pair_put(ice, lo_ic, q, cic);
        }
        if(tq != t0 || !(tq == tq)){
            t_dirty = true;
            if(A.lo_t){                           // stored as a pair: the next gas reads hi + lo
                th = (float)tq;
                tl = (float)(tq - (double)th);
                tq = (double)th + (double)tl;
            }
        }
    }
    if(t_dirty){
        if(A.lo_t){ A.t[q] = (T)th; A.lo_t[q] = tl; }
        else A.t[q] = (T)tq;
    }
}

// =============================================================================================
// FluxLimiter<P>::nh4sh, Superbee (FluxLimiter.h:67-190), with the periodic seam
// =============================================================================================
template<typename D>
__device__ __forceinline__ D superbee(D r){
    return dmax(D(0.0), dmax(dmin(2.0 * r, D(1.0)), dmin(r, D(2.0))));
}

// The Superbee correction of one interior cell (i, j, k) at storage index q; n_at(index) gives
// NH4SH at a storage index and u, v, w are the cell's velocities. D: double for the limiter itself,
// Dual for its frozen-branch linearisation (the limiter's choices follow the values).
template<typename D, typename NA>
__device__ __forceinline__ D flux_corr(const NA& n_at, D u, D v, D w, size_t q, int i, int j, int k,
        int im, int jm, int km, size_t P, size_t R, const double* __restrict__ tabs,
        double dr, double dthe, double dphi, double c43, double c13){
    const int n = km - 1;
    auto wrap = [n](int kk){ return kk < 0 ? kk + n : (kk >= n ? kk - n : kk); };
    const size_t row = q - k;
    const double eps = 1.0e-12;
    const D qc = n_at(q);
    const D q_rp = n_at(q + P), q_rm = n_at(q - P);
    const D q_tp = n_at(q + R), q_tm = n_at(q - R);
    const D q_pp = n_at(row + wrap(k + 1)), q_pm = n_at(row + wrap(k - 1));
    const double rm = tabs[i];
    const double sinthe = dmax(0.4, fabs(tabs[im + j]));
    const double inv_rm = 1.0 / rm;
    const double inv_rmsinthe = 1.0 / (rm * sinthe);
    D corr = 0.0;
    {
        const D df = q_rp - qc, db = qc - q_rm;
        const D denom = df + (df >= 0.0 ? eps : -eps);
        const D r = (u >= 0.0) ? db / denom : ((i + 2 < im ? n_at(q + 2 * P) : q_rp) - q_rp) / denom;
        corr += (1.0 - superbee(r)) * fabs(u) * (df - db) / (2.0 * dr);
    }
    {
        const D df = q_tp - qc, db = qc - q_tm;
        const D denom = df + (df >= 0.0 ? eps : -eps);
        const D q_far = (j + 2 < jm) ? n_at(q + 2 * R) : (c43 * q_tp - c13 * qc);
        const D r = (v >= 0.0) ? db / denom : (q_far - q_tp) / denom;
        corr += (1.0 - superbee(r)) * fabs(v) * inv_rm * (df - db) / (2.0 * dthe);
    }
    {
        const D df = q_pp - qc, db = qc - q_pm;
        const D denom = df + (df >= 0.0 ? eps : -eps);
        const D r = (w >= 0.0) ? db / denom : (n_at(row + wrap(k + 2)) - q_pp) / denom;
        corr += (1.0 - superbee(r)) * fabs(w) * inv_rmsinthe * (df - db) / (2.0 * dphi);
    }
    return corr;
}

template<typename T>
__global__ __launch_bounds__(256) void k_flux_limiter(const T* __restrict__ N, const T* __restrict__ U,
        const T* __restrict__ V, const T* __restrict__ W, T* __restrict__ out, const double* __restrict__ tabs,
        int im, int jm, int km, int pitch, size_t plane, double dr, double dthe, double dphi, double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1, i = blockIdx.z + 1;
    if(k > km - 2 || j > jm - 2 || i > im - 2) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    auto n_at = [N](size_t x){ return (double)N[x]; };
    out[q] = (T)flux_corr<double>(n_at, (double)U[q], (double)V[q], (double)W[q], q, i, j, k, im, jm, km,
                                  plane, (size_t)pitch, tabs, dr, dthe, dphi, c43, c13);
}

// ---------------------------------------------------------------------------------------------
// Frozen-branch linearisations of the saturation adjustment and the NH4SH flux limiter (FP64
// contexts): the tangent runs the operator on Dual numbers along a direction, the adjoint applies
// the transpose of the same Jacobian. The branches, clamps and the adjustment's iteration exits
// are those the forward takes at the context's state.
// ---------------------------------------------------------------------------------------------
// the adjustment's six inputs and four outputs at a cell: vapour, cloud, ice, t, then p_stat, rho_mix
struct SatLin {
    const double* x[6];          // the state the adjustment would start from
    const uint8_t* flags;
    int jm, km, pitch;
};

__device__ __forceinline__ void satadj_dual(const SatLin& L, size_t q, const Dual in[6], Dual out[4], const SatParams& P){
    Dual cv = in[0], ccl = in[1], cic = in[2], tq = in[3];
    if(!(L.flags[q] & FL_SELF)) satadj_cell(cv, ccl, cic, tq, in[4], in[5], P);
    if(cv <= 0.0) cv = Dual(0.0);
    if(ccl <= 0.0) ccl = Dual(0.0);
    if(cic <= 0.0) cic = Dual(0.0);
    out[0] = cv; out[1] = ccl; out[2] = cic; out[3] = tq;
}

__global__ __launch_bounds__(128) void kt_satadj(const SatLin L, const SatParams P,
        const double* __restrict__ dx, double* __restrict__ dy, size_t n3){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= L.km) return;
    const size_t q = ((size_t)i * L.jm + j) * L.pitch + k;
    Dual in[6], out[4];
    #pragma unroll
    for(int a = 0; a < 6; a++) in[a] = Dual(L.x[a][q], dx[a * n3 + q]);
    satadj_dual(L, q, in, out, P);
    #pragma unroll
    for(int r = 0; r < 4; r++) dy[r * n3 + q] = out[r].d;
}

// lx = J^T ly per cell, the Jacobian's six columns from six Dual runs
__global__ __launch_bounds__(128) void ka_satadj(const SatLin L, const SatParams P,
        const double* __restrict__ ly, double* __restrict__ lx, size_t n3){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= L.km) return;
    const size_t q = ((size_t)i * L.jm + j) * L.pitch + k;
    double lam[4];
    #pragma unroll
    for(int r = 0; r < 4; r++) lam[r] = ly[r * n3 + q];
    #pragma unroll 1
    for(int a = 0; a < 6; a++){
        Dual in[6], out[4];
        #pragma unroll
        for(int b = 0; b < 6; b++) in[b] = Dual(L.x[b][q], b == a ? 1.0 : 0.0);
        satadj_dual(L, q, in, out, P);
        double acc = 0.0;
        #pragma unroll
        for(int r = 0; r < 4; r++) acc += out[r].d * lam[r];
        lx[a * n3 + q] = acc;
    }
}

struct FlxLin {
    const double *N, *U, *V, *W;
    const double* tabs;
    int im, jm, km, pitch;
    size_t plane;
    double dr, dthe, dphi;
};

__device__ __forceinline__ bool flx_interior(const FlxLin& L, int i, int j, int k){
    return i >= 1 && i <= L.im - 2 && j >= 1 && j <= L.jm - 2 && k >= 0 && k <= L.km - 2;
}

// d(fluxlim_nh4sh) along (dN, dU, dV, dW) [4][n3]; cells the limiter does not write get 0
__global__ __launch_bounds__(128) void kt_flux_limiter(const FlxLin L, const double* __restrict__ dx,
        double* __restrict__ dy, size_t n3){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= L.km) return;
    const size_t q = ((size_t)i * L.jm + j) * L.pitch + k;
    double d = 0.0;
    if(flx_interior(L, i, j, k)){
        const double* N = L.N;
        auto n_at = [N, dx](size_t x){ return Dual(N[x], dx[x]); };
        d = flux_corr<Dual>(n_at, Dual(L.U[q], dx[n3 + q]), Dual(L.V[q], dx[2 * n3 + q]), Dual(L.W[q], dx[3 * n3 + q]),
                            q, i, j, k, L.im, L.jm, L.km, L.plane, (size_t)L.pitch, L.tabs,
                            L.dr, L.dthe, L.dphi, 4.0 / 3.0, 1.0 / 3.0).d;
    }
    dy[q] = d;
}

// (lN, lU, lV, lW) = J^T ly, gathered: NH4SH at x is read by the corrections of x itself, of its
// radial and meridional neighbours at -2 .. +1 and of its zonal neighbours at -2 .. +1 (periodic
// over the km - 1 distinct columns); each of those corrections is differentiated with respect to
// the value at x. The velocities enter only the own cell's correction.
__global__ __launch_bounds__(128) void ka_flux_limiter(const FlxLin L, const double* __restrict__ ly,
        double* __restrict__ lx, size_t n3){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= L.km) return;
    const size_t q = ((size_t)i * L.jm + j) * L.pitch + k;
    const int n = L.km - 1;
    const double* N = L.N;
    double lN = 0.0, lU = 0.0, lV = 0.0, lW = 0.0;
    if(k < n){
        const int off[10][3] = { { 0, 0, 0 }, { -1, 0, 0 }, { 1, 0, 0 }, { -2, 0, 0 }, { 0, -1, 0 }, { 0, 1, 0 },
                                 { 0, -2, 0 }, { 0, 0, -1 }, { 0, 0, 1 }, { 0, 0, -2 } };
        #pragma unroll 1
        for(int c = 0; c < 10; c++){
            const int yi = i + off[c][0], yj = j + off[c][1];
            int yk = k + off[c][2];
            yk = yk < 0 ? yk + n : (yk >= n ? yk - n : yk);
            if(!flx_interior(L, yi, yj, yk)) continue;
            const size_t qy = ((size_t)yi * L.jm + yj) * L.pitch + yk;
            auto n_at = [N, q](size_t x){ return Dual(N[x], x == q ? 1.0 : 0.0); };
            const Dual cy = flux_corr<Dual>(n_at, Dual(L.U[qy]), Dual(L.V[qy]), Dual(L.W[qy]), qy, yi, yj, yk,
                                            L.im, L.jm, L.km, L.plane, (size_t)L.pitch, L.tabs,
                                            L.dr, L.dthe, L.dphi, 4.0 / 3.0, 1.0 / 3.0);
            lN += cy.d * ly[qy];
        }
    }
    if(flx_interior(L, i, j, k)){
        const double l = ly[q];
        auto n_at = [N](size_t x){ return Dual(N[x]); };
        #pragma unroll 1
        for(int a = 0; a < 3; a++){
            const Dual cu = flux_corr<Dual>(n_at, Dual(L.U[q], a == 0 ? 1.0 : 0.0), Dual(L.V[q], a == 1 ? 1.0 : 0.0),
                                            Dual(L.W[q], a == 2 ? 1.0 : 0.0), q, i, j, k, L.im, L.jm, L.km, L.plane,
                                            (size_t)L.pitch, L.tabs, L.dr, L.dthe, L.dphi, 4.0 / 3.0, 1.0 / 3.0);
            (a == 0 ? lU : a == 1 ? lV : lW) = cu.d * l;
        }
    }
    lx[q] = lN; lx[n3 + q] = lU; lx[2 * n3 + q] = lV; lx[3 * n3 + q] = lW;
}

// =============================================================================================
// host launchers
// =============================================================================================
namespace {

template<typename T> T* F(atjup_ctx* c, int f){ return (T*)c->cur[f].p; }
template<typename T> T* LIVE(atjup_ctx* c, int f){
    if(c->turb_live_pending && (f == ATJUP_F_TKE || f == ATJUP_F_DIS))
        return (T*)c->live_turb[f - ATJUP_F_TKE].p;
    return (T*)c->cur[f].p;
}

template<typename T>
int do_rk4_full(atjup_ctx* c){
    StageFullArgs<T> a{};
    const double dt = c->c.dt;
    const atjup_consts& C = c->c;
    const atjup_knobs& K = c->kn;
    a.pd = F<T>(c, ATJUP_F_P_DYN); a.ph = F<T>(c, ATJUP_F_P_HYDRO); a.ps = F<T>(c, ATJUP_F_P_STAT);
    a.rho = F<T>(c, ATJUP_F_RHO_MIX); a.wn = F<T>(c, ATJUP_F_WALL_NUE); a.nue = F<T>(c, ATJUP_F_NUE);
    a.qrad = F<T>(c, ATJUP_F_Q_RAD); a.qprec = F<T>(c, ATJUP_F_Q_PRECIP); a.fl_nh4sh = F<T>(c, ATJUP_F_FLUXLIM_NH4SH);
    a.aux_u = F<T>(c, ATJUP_F_AUX_U); a.aux_v = F<T>(c, ATJUP_F_AUX_V); a.aux_w = F<T>(c, ATJUP_F_AUX_W);
    a.prod = F<T>(c, ATJUP_F_PROD); a.tsrc = F<T>(c, ATJUP_F_TKE_SOURCE); a.dsrc = F<T>(c, ATJUP_F_DIS_SOURCE);
    a.flags = (const uint8_t*)c->flags.p;
    a.geo = (const GeoIJ<T>*)c->geo.p;
    a.brl = (const acc_t<T>*)c->brl.p;
    a.im = c->im; a.jm = c->jm; a.km = c->km; a.pitch = c->pitch; a.plane = c->plane;
    a.nkb = (c->km + SF_BK - 1) / SF_BK; a.njb = (c->jm + SF_BJ - 1) / SF_BJ;
    a.nb = a.nkb * a.njb * c->im;
    a.turb = K.turb_model >= 0 ? 1 : 0;
    a.itopo = (const int*)c->itopo.p;
    a.lthick = (const double*)c->tabs.p + c->im + 3 * c->jm;
    a.vel_star = (const double*)c->vel_star.p;
    a.lh_h = (float)(c->g.L_atm * c->g.dr);
    a.re_turb = (T)C.re_turb;
    a.dt = (T)dt; a.one_sixth = (T)(1.0 / 6.0);
    a.dt_d = dt;
    for(int f = 0; f < ATJUP_N_PROG; f++) a.lo[f] = (float*)c->lo[f].p;
    a.inv_2dr = (T)(1.0 / (2.0 * c->g.dr)); a.inv_2dthe = (T)(1.0 / (2.0 * c->g.dthe));
    a.inv_2dphi = (T)(1.0 / (2.0 * c->g.dphi)); a.inv_dr2 = (T)(1.0 / (c->g.dr * c->g.dr));
    a.inv_dthe2 = (T)(1.0 / (c->g.dthe * c->g.dthe)); a.inv_dphi2 = (T)(1.0 / (c->g.dphi * c->g.dphi));
    a.cep = (T)(C.u_0 * C.u_0 / (C.cp_mix * C.t_ref));
    a.to_nd = (T)(1.0 / (C.r_mix * C.u_0 * C.u_0 * 1.0e-5));
    a.r_mix = (T)C.r_mix;
    a.c_m2o = (T)(1.0 * -2.0 * C.omega);
    a.c_p2o = (T)(1.0 * +2.0 * C.omega);
    a.Coriolis = (T)C.Coriolis; a.centrifugal = (T)C.centrifugal;
    a.cbuoy = -1.0 * 1.0 * 1.0 * C.buoyancy;
    a.g = C.g; a.rR = C.r_mix * C.R_mix; a.t_ref_d = C.t_ref; a.t_ref = (T)C.t_ref;
    a.t_coef = (T)(K.t_diff_scale / (C.re * C.pr));
    a.inv_re = (T)(1.0 / C.re);
    a.rad_coupling = (T)K.rad_coupling; a.precip_coupling = (T)K.precip_coupling; a.turb_coupling = (T)K.turb_coupling;
    a.L_rad = (T)(c->g.L_atm * 1.0e3); a.R_mix = (T)C.R_mix; a.cp_mix = (T)C.cp_mix; a.u_0 = (T)C.u_0;
    a.cr = (T)C.chemical_reaction; a.pc = (T)K.precip_coupling;
    a.stokes_a = (T)((2.0 / 9.0) * C.r_p_nh4sh * C.r_p_nh4sh);
    a.rg_nh4sh = (T)C.rg_nh4sh; a.mue_mix = (T)C.mue_mix;
    a.nue_max = (T)(K.nue_max_phys / (C.u_0 * (c->g.L_atm * 1.0e3)));
    a.tke_max_nd = (T)(K.tke_max_phys / (C.u_0 * C.u_0));
    a.dis_min = (T)1.0e-10;
    const double sds = K.species_diff_scale;
    const double sc[NP] = { 0, 0, 0, 0, C.sc_h2o, C.sc_h2o, C.sc_h2o, C.sc_h2s, C.sc_nh3, C.sc_nh3, C.sc_nh3,
                            C.sc_ch4, C.sc_ch4, C.sc_ch4, C.sc_nh4sh, 0, 0 };
    for(int f = ATJUP_F_H2O; f <= ATJUP_F_NH4SH; f++) a.sp_coef[f] = (T)(sds / (sc[f] * C.re));
    const int spr_of[NP] = { -1, -1, -1, -1, ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_H2O_CLOUD, ATJUP_F_S_PRECIP_H2O_ICE,
                             -1, ATJUP_F_S_PRECIP_NH3, ATJUP_F_S_PRECIP_NH3_CLOUD, ATJUP_F_S_PRECIP_NH3_ICE,
                             ATJUP_F_S_PRECIP_CH4, ATJUP_F_S_PRECIP_CH4_CLOUD, ATJUP_F_S_PRECIP_CH4_ICE, -1, -1, -1 };
    const int mf_of[NP] = { -1, -1, -1, -1, -1, -1, -1, ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, -1, -1,
                            -1, -1, -1, ATJUP_F_MASSFLUX_NH4SH, -1, -1 };
    for(int f = 0; f < NP; f++){
        a.spr[f] = spr_of[f] >= 0 ? F<T>(c, spr_of[f]) : nullptr;
        a.mf[f] = mf_of[f] >= 0 ? F<T>(c, mf_of[f]) : nullptr;
    }
    const bool pend = c->turb_live_pending;
    for(int s = 0; s < 4; s++){
        a.stage = s;
        a.c_in = (T)((s == 0 || s == 1) ? 0.5 * dt : (s == 2 ? dt : 0.0));
        a.wgt = (T)((s == 0 || s == 3) ? 1.0 : 2.0);
        for(int f = 0; f < NP; f++){
            T* N = F<T>(c, f);
            T* X = (T*)c->stg[0][f].p;
            T* Y = (T*)c->stg[1][f].p;
            const bool turbf = (f == ATJUP_F_TKE || f == ATJUP_F_DIS);
            T* L = (pend && turbf) ? (T*)c->live_turb[f - ATJUP_F_TKE].p : N;
            const T* in = (s == 0) ? L : (s == 1 ? X : (s == 2 ? Y : X));
            T* out = (s == 0) ? X : (s == 1 ? Y : (s == 2 ? X : N));
            a.in[f] = in; a.out[f] = out; a.nl[f] = N; a.acc[f] = (T*)c->acc[f].p;
            // the cells the step does not integrate: stages 0 and 1 carry them into the stage
            // buffers; the closure's live k*, omega* replace the n level's at the end
            a.bsrc[f] = (s <= 1) ? in : ((s == 3 && pend && turbf) ? L : nullptr);
        }
        hipEvent_t t0 = timer_begin(c);
        const dim3 gs(a.nb), bs(SF_BK, SF_BJ);
        switch(K.turb_model){
            case 0:  hipLaunchKernelGGL((k_stage_full<T, 0>), gs, bs, 0, c->stream, a, a.geo); break;
            case 1:  hipLaunchKernelGGL((k_stage_full<T, 1>), gs, bs, 0, c->stream, a, a.geo); break;
            case 2:  hipLaunchKernelGGL((k_stage_full<T, 2>), gs, bs, 0, c->stream, a, a.geo); break;
            default: hipLaunchKernelGGL((k_stage_full<T, -1>), gs, bs, 0, c->stream, a, a.geo); break;
        }
        hipLaunchKernelGGL(k_stage_species<T>, dim3(a.nb), dim3(SF_BK, SF_BJ), 0, c->stream, a, a.geo);
        timer_end(c, ATJUP_T_RK_STAGES, t0, 2);
        if(int rc = check_launch()) return rc;
    }
    c->turb_live_pending = false;
    return 0;
}

template<typename T>
TurbArrays<T> turb_arrays(atjup_ctx* c, T* tke, T* dis){
    TurbArrays<T> L{};
    L.a[0] = tke; L.a[1] = dis; L.a[2] = F<T>(c, ATJUP_F_NUE); L.a[3] = F<T>(c, ATJUP_F_PROD);
    L.a[4] = F<T>(c, ATJUP_F_TKE_SOURCE); L.a[5] = F<T>(c, ATJUP_F_DIS_SOURCE);
    return L;
}

template<typename T>
int do_turbulence(atjup_ctx* c){
    const atjup_consts& C = c->c;
    const double L_m = c->g.L_atm * 1.0e3;
    const double nue_max = c->kn.nue_max_phys / (C.u_0 * L_m);
    const double k_bg = c->kn.turb_k_bg_phys / (C.u_0 * C.u_0);
    const double dis_min = 1.0e-10;
    // compute_vel_star with TURB_FREE: the calm-air floor 0.05 m/s everywhere (Turbulence.h:340-342)
    const size_t n2 = (size_t)c->jm * c->km;
    hipLaunchKernelGGL(k_fill_double, dim3(nblk(n2, 256)), dim3(256), 0, c->stream, (double*)c->vel_star.p, n2, 0.05);
    T* tke = F<T>(c, ATJUP_F_TKE);
    T* dis = F<T>(c, ATJUP_F_DIS);
    T* ltke = (T*)c->live_turb[0].p;
    T* ldis = (T*)c->live_turb[1].p;
    if(c->turb_live_pending){
        // a second run before the step: it starts from the previous run's live values
        tke = ltke; dis = ldis;
    }
    dim3 gfull(nblk(c->km, 128), c->jm, c->im), b128(128);
    hipLaunchKernelGGL(k_turb_clamp_in<T>, gfull, b128, 0, c->stream, tke, dis, ltke, ldis,
        (const uint8_t*)c->flags.p, c->im, c->jm, c->km, c->pitch, (T)dis_min);
    c->turb_live_pending = true;
    TurbConsts tc{ c->g.dr, c->g.dthe, c->g.dphi, nue_max, dis_min, C.u_0, L_m, 0.4 * 0.1 / 1.8e-5 };
    {
        const dim3 gs(nblk(c->km - 1, 128), c->jm - 2, c->im - 2);
        auto launch = [&](auto kern){
            hipLaunchKernelGGL(kern, gs, b128, 0, c->stream,
                (const T*)ltke, (const T*)ldis, (const T*)F<T>(c, ATJUP_F_U), (const T*)F<T>(c, ATJUP_F_V),
                (const T*)F<T>(c, ATJUP_F_W), F<T>(c, ATJUP_F_NUE), F<T>(c, ATJUP_F_PROD),
                F<T>(c, ATJUP_F_TKE_SOURCE), F<T>(c, ATJUP_F_DIS_SOURCE), (const uint8_t*)c->flags.p,
                (const double*)c->tabs.p, (const double*)c->vel_star.p, c->im, c->jm, c->km, c->pitch, c->plane, tc);
        };
        if(c->kn.turb_model == 0) launch(k_turb_sources<T, 0>);
        else if(c->kn.turb_model == 2) launch(k_turb_sources<T, 2>);
        else launch(k_turb_sources<T, 1>);
    }
    const TurbArrays<T> L = turb_arrays<T>(c, ltke, ldis);
    hipLaunchKernelGGL(k_turb_wall_plane<T>, dim3(nblk(c->km, 128), c->jm), b128, 0, c->stream,
        L.a[0], L.a[1], L.a[2], L.a[3], L.a[4], L.a[5], (const uint8_t*)c->flags.p,
        c->im, c->jm, c->km, c->pitch, c->plane, (T)k_bg, (T)dis_min, (T)nue_max, 1);
    hipLaunchKernelGGL(k_turb_post<T>, gfull, b128, 0, c->stream,
        L.a[0], L.a[1], L.a[2], L.a[3], L.a[4], L.a[5], (const uint8_t*)c->flags.p,
        c->im, c->jm, c->km, c->pitch, (T)k_bg, (T)dis_min, (T)nue_max);
    hipLaunchKernelGGL(k_turb_seam<T>, dim3(nblk(c->jm, 64), c->im), dim3(64), 0, c->stream, L, c->im, c->jm, c->km, c->pitch);
    hipLaunchKernelGGL(k_turb_lid<T>, dim3(nblk(c->km, 128), c->jm - 2), b128, 0, c->stream, L, c->im, c->jm, c->km,
        c->pitch, c->plane);
    hipLaunchKernelGGL(k_turb_pole_rows<T>, dim3(nblk(c->km, 128), c->im), b128, 0, c->stream, L, c->im, c->jm, c->km,
        c->pitch);
    const double* tabs = (const double*)c->tabs.p;
    if(c->kn.turb_model != 0)            // the wall function leaves k-epsilon alone (cJupiterModel.cpp:356-360)
        hipLaunchKernelGGL(k_turb_wall_omega<T>, dim3(nblk(c->km, 128), c->jm), b128, 0, c->stream, ldis,
            (const int*)c->itopo.p, (const double*)c->vel_star.p, tabs + c->im + 3 * c->jm,
            c->im, c->jm, c->km, c->pitch, L_m / C.u_0);
    return check_launch();
}

template<typename T>
int do_turb_wall(atjup_ctx* c){
    T* tke = LIVE<T>(c, ATJUP_F_TKE);
    T* dis = LIVE<T>(c, ATJUP_F_DIS);
    const TurbArrays<T> L = turb_arrays<T>(c, tke, dis);
    dim3 b128(128);
    hipLaunchKernelGGL(k_turb_wall_plane<T>, dim3(nblk(c->km, 128), c->jm), b128, 0, c->stream,
        L.a[0], L.a[1], L.a[2], L.a[3], L.a[4], L.a[5], (const uint8_t*)c->flags.p,
        c->im, c->jm, c->km, c->pitch, c->plane, T(0), T(0), T(0), 0);
    const double* tabs = (const double*)c->tabs.p;
    if(c->kn.turb_model != 0)
        hipLaunchKernelGGL(k_turb_wall_omega<T>, dim3(nblk(c->km, 128), c->jm), b128, 0, c->stream, dis,
            (const int*)c->itopo.p, (const double*)c->vel_star.p, tabs + c->im + 3 * c->jm,
            c->im, c->jm, c->km, c->pitch, c->g.L_atm * 1.0e3 / c->c.u_0);
    return check_launch();
}

template<typename T>
int do_turb_sanitize(atjup_ctx* c){
    const double tke_max = 1000.0 / (c->c.u_0 * c->c.u_0);
    const double nue_max = 1.0e5 / (c->c.u_0 * c->g.L_atm * 1.0e3);
    hipLaunchKernelGGL(k_turb_sanitize<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream,
        LIVE<T>(c, ATJUP_F_TKE), LIVE<T>(c, ATJUP_F_DIS), F<T>(c, ATJUP_F_NUE), F<T>(c, ATJUP_F_PROD),
        F<T>(c, ATJUP_F_TKE_SOURCE), F<T>(c, ATJUP_F_DIS_SOURCE), c->n3, (T)tke_max, (T)nue_max);
    return check_launch();
}

template<typename T>
int do_satadj(atjup_ctx* c, int gas, const atjup_gas_params& g){
    static const int ids[3][3] = { { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE },
                                   { ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE },
                                   { ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE } };
    SatParams P{ g.t_0, g.t_00, g.ep, g.lv, g.ls, g.C, g.L0, g.R, g.del_alf, g.del_bet,
                 g.C_i, g.L0_i, g.del_alf_i, g.del_bet_i, c->c.cp_mix, c->c.t_ref, c->c.r_mix };
    hipLaunchKernelGGL(k_satadj<T>, dim3(nblk(c->km, 128), c->jm, c->im), dim3(128), 0, c->stream,
        F<T>(c, ids[gas][0]), F<T>(c, ids[gas][1]), F<T>(c, ids[gas][2]), F<T>(c, ATJUP_F_T),
        F<T>(c, ATJUP_F_P_STAT), F<T>(c, ATJUP_F_RHO_MIX), (const uint8_t*)c->flags.p,
        c->jm, c->km, c->pitch, P, (float*)c->lo[ATJUP_F_T].p, (float*)c->lo[ids[gas][0]].p,
        (float*)c->lo[ids[gas][1]].p, (float*)c->lo[ids[gas][2]].p);
    return check_launch();
}

template<typename T>
int do_satadj3(atjup_ctx* c){
    static const int ids[3][3] = { { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE },
                                   { ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE },
                                   { ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE } };
    Sat3Args<T> A;
    for(int gas = 0; gas < 3; gas++){
        const atjup_gas_params& g = c->gas[gas];
        A.P[gas] = SatParams{ g.t_0, g.t_00, g.ep, g.lv, g.ls, g.C, g.L0, g.R, g.del_alf, g.del_bet,
                              g.C_i, g.L0_i, g.del_alf_i, g.del_bet_i, c->c.cp_mix, c->c.t_ref, c->c.r_mix };
        for(int s = 0; s < 3; s++){
            A.sp[gas][s] = F<T>(c, ids[gas][s]);
            A.lo[gas][s] = (float*)c->lo[ids[gas][s]].p;
        }
    }
    A.t = F<T>(c, ATJUP_F_T); A.lo_t = (float*)c->lo[ATJUP_F_T].p;
    A.ps = F<T>(c, ATJUP_F_P_STAT); A.rho = F<T>(c, ATJUP_F_RHO_MIX); A.flags = (const uint8_t*)c->flags.p;
    A.jm = c->jm; A.km = c->km; A.pitch = c->pitch;
    hipLaunchKernelGGL(k_satadj3<T>, dim3(nblk(c->km, 128), c->jm, c->im), dim3(128), 0, c->stream, A);
    return check_launch();
}

template<typename T>
int do_flux_limiter(atjup_ctx* c){
    hipLaunchKernelGGL(k_flux_limiter<T>, dim3(nblk(c->km - 1, 128), c->jm - 2, c->im - 2), dim3(128), 0, c->stream,
        F<T>(c, ATJUP_F_NH4SH), F<T>(c, ATJUP_F_U), F<T>(c, ATJUP_F_V), F<T>(c, ATJUP_F_W),
        F<T>(c, ATJUP_F_FLUXLIM_NH4SH), (const double*)c->tabs.p, c->im, c->jm, c->km, c->pitch, c->plane,
        c->g.dr, c->g.dthe, c->g.dphi, 4.0 / 3.0, 1.0 / 3.0);
    return check_launch();
}

// the frozen-branch linearisations (FP64 contexts): in and out are [n][n3] device arrays in the
// field layout; tangent: in = the direction, out = the derivative; adjoint: in = the outputs'
// adjoints, out = the inputs' adjoints
int do_satadj_lin(atjup_ctx* c, int gas, int adjoint, const double* in, double* out){
    static const int ids[3][3] = { { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE },
                                   { ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE },
                                   { ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE } };
    const atjup_gas_params& g = c->gas[gas];
    const SatParams P{ g.t_0, g.t_00, g.ep, g.lv, g.ls, g.C, g.L0, g.R, g.del_alf, g.del_bet,
                       g.C_i, g.L0_i, g.del_alf_i, g.del_bet_i, c->c.cp_mix, c->c.t_ref, c->c.r_mix };
    SatLin L;
    for(int s = 0; s < 3; s++) L.x[s] = F<double>(c, ids[gas][s]);
    L.x[3] = F<double>(c, ATJUP_F_T); L.x[4] = F<double>(c, ATJUP_F_P_STAT); L.x[5] = F<double>(c, ATJUP_F_RHO_MIX);
    L.flags = (const uint8_t*)c->flags.p;
    L.jm = c->jm; L.km = c->km; L.pitch = c->pitch;
    const dim3 grid(nblk(c->km, 128), c->jm, c->im);
    if(adjoint) hipLaunchKernelGGL(ka_satadj, grid, dim3(128), 0, c->stream, L, P, in, out, c->n3);
    else        hipLaunchKernelGGL(kt_satadj, grid, dim3(128), 0, c->stream, L, P, in, out, c->n3);
    return check_launch();
}

int do_flux_limiter_lin(atjup_ctx* c, int adjoint, const double* in, double* out){
    FlxLin L;
    L.N = F<double>(c, ATJUP_F_NH4SH); L.U = F<double>(c, ATJUP_F_U); L.V = F<double>(c, ATJUP_F_V); L.W = F<double>(c, ATJUP_F_W);
    L.tabs = (const double*)c->tabs.p;
    L.im = c->im; L.jm = c->jm; L.km = c->km; L.pitch = c->pitch; L.plane = c->plane;
    L.dr = c->g.dr; L.dthe = c->g.dthe; L.dphi = c->g.dphi;
    const dim3 grid(nblk(c->km, 128), c->jm, c->im);
    if(adjoint) hipLaunchKernelGGL(ka_flux_limiter, grid, dim3(128), 0, c->stream, L, in, out, c->n3);
    else        hipLaunchKernelGGL(kt_flux_limiter, grid, dim3(128), 0, c->stream, L, in, out, c->n3);
    return check_launch();
}

// =============================================================================================
// ChemistryJup::EquilibriumNH4SH("loop") (ChemistryJup.h:293-370): every fluid cell brought to the
// Lewis equilibrium p_NH3 p_H2S = K_p(T) = 10^(14.82 - 4705 / T) atm^2, the heat of formation
// moving T with the extent x [kmol/m3] of NH4SH formed (x > 0) or decomposed (x < 0, at most all of
// it); x by bisection (at most 80 halvings, to 1e-15 of the largest of a, b, s) on the monotone
// f(x) = (a - x)(b - x)(R* T(x))^2 - K_p(T(x)). One thread per cell (the cells are independent); the
// arithmetic is FP64 in both forms, the FP32-refined form's t read and written as hi + lo.
// =============================================================================================
struct EqConsts { double m_nh3, m_h2s, m_nh4sh, t_ref, r_mix, cp_mix, L_f; };

template<typename T>
__global__ __launch_bounds__(256) void k_nh4sh_eq(T* __restrict__ t, float* __restrict__ lo_t, T* __restrict__ nh3,
        T* __restrict__ h2s, T* __restrict__ nh4sh, const T* __restrict__ rho_mix, const uint8_t* __restrict__ msk,
        size_t n, int km, int pitch, const EqConsts C, float* __restrict__ lo_nh3, float* __restrict__ lo_h2s,
        float* __restrict__ lo_nh4sh){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n || (int)(q % (size_t)pitch) >= km || msk[q]) return;
    const double R_univ = 8314.46, atm = 101325.0;
    const double tq = pair_get(t, lo_t, q);              // the refined form's pairs: hi + lo
    const double T0 = tq * C.t_ref;
    if(!(T0 > 0.0)) return;
    double rho = (double)rho_mix[q];
    if(!(rho > 0.0) || !isfinite(rho)) rho = C.r_mix;
    const double nh3_0 = pair_get(nh3, lo_nh3, q), h2s_0 = pair_get(h2s, lo_h2s, q);
    const double nh4sh_0 = pair_get(nh4sh, lo_nh4sh, q);
    const double a = dmax(0.0, nh3_0) / C.m_nh3;
    const double b = dmax(0.0, h2s_0) / C.m_h2s;
    const double s = dmax(0.0, nh4sh_0) / C.m_nh4sh;
    const double heat = C.L_f * C.m_nh4sh / (C.cp_mix * rho);
    // The bisection needs only the sign of f(x) = A - pow(10, y) atm^2, y = 14.82 - 4705 / T(x):
    // 1 for f > 0, 0 for f <= 0, 2 unordered (NaN). The sign is first taken from exp(y ln 10) in
    // place of pow(10, y), which differ by less than 4e-14 of the value for |y| < 100 and 1.3e-13
    // for |y| < 300 (the rounding of y ln 10, about 2 |y ln 10| 2^-53, and the two functions' own
    // ulps; below y = -307 the values are subnormal and lose that bound, so y < -300, T below
    // 15 K, is left out); where f is farther from zero than 1e-12 + 4e-15 |y| of its larger term
    // the sign is the one f has with pow, otherwise (and for every non-finite case) f is formed with
    // pow as written. So the decisions, and the result, are those of the bisection on f itself at
    // a fraction of its pow calls: the halvings far from the root, about 40 of 50, take exp.
    constexpr double LN10 = 2.302585092994045684;
    auto f_sign = [&](double x) -> int {
        const double Tx = T0 + heat * x, RT = R_univ * Tx;
        const double A = (a - x) * (b - x) * RT * RT;
        const double y = 14.82 - 4705.0 / Tx;
        const double Pc = exp(y * LN10) * atm * atm;
        const double d = A - Pc;
        if(y > -300.0 && fabs(d) > (1.0e-12 + 4.0e-15 * fabs(y)) * dmax(fabs(A), Pc)) return d > 0.0 ? 1 : 0;
        const double fe = A - pow(10.0, y) * atm * atm;
        return (fe > 0.0) ? 1 : ((fe <= 0.0) ? 0 : 2);
    };
    const int s0 = f_sign(0.0);
    double lo, hi;
    if(s0 == 1){ lo = 0.0; hi = dmin(a, b); }
    else if(s > 0.0){
        if(f_sign(-s) == 0){ lo = -s; hi = -s; }
        else { lo = -s; hi = 0.0; }
    } else return;
    if(hi > lo){
        const double tol = 1.0e-15 * dmax(a, dmax(b, s));
        for(int it = 0; it < 80; it++){
            const double xm = 0.5 * (lo + hi);
            if(f_sign(xm) == 1) lo = xm; else hi = xm;
            if(hi - lo <= tol) break;
        }
    }
    const double x = (s0 == 1) ? lo : hi;
    if(x == 0.0) return;
    pair_put(nh3, lo_nh3, q, dmax(0.0, nh3_0 - C.m_nh3 * x));
    pair_put(h2s, lo_h2s, q, dmax(0.0, h2s_0 - C.m_h2s * x));
    pair_put(nh4sh, lo_nh4sh, q, dmax(0.0, nh4sh_0 + C.m_nh4sh * x));
    pair_put(t, lo_t, q, (T0 + heat * x) / C.t_ref);
}

template<typename T>
int do_nh4sh_eq(atjup_ctx* c){
    const atjup_chem_params& p = c->chemp;
    const EqConsts C{ p.m_nh3, p.m_h2s, p.m_nh4sh, c->c.t_ref, c->c.r_mix, c->c.cp_mix,
                      4705.0 * std::log(10.0) * 8314.46 / 51.1114 };
    hipLaunchKernelGGL(k_nh4sh_eq<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream,
        F<T>(c, ATJUP_F_T), (float*)c->lo[ATJUP_F_T].p, F<T>(c, ATJUP_F_NH3), F<T>(c, ATJUP_F_H2S),
        F<T>(c, ATJUP_F_NH4SH), (const T*)F<T>(c, ATJUP_F_RHO_MIX), (const uint8_t*)c->mask.p,
        c->n3, c->km, c->pitch, C, (float*)c->lo[ATJUP_F_NH3].p, (float*)c->lo[ATJUP_F_H2S].p,
        (float*)c->lo[ATJUP_F_NH4SH].p);
    return check_launch();
}

template<typename T>
int do_commit_live(atjup_ctx* c){
    if(!c->turb_live_pending) return 0;
    for(int t = 0; t < 2; t++)
        ATJUP_HIP_CHECK(hipMemcpyAsync(c->cur[ATJUP_F_TKE + t].p, c->live_turb[t].p, c->cur[ATJUP_F_TKE + t].bytes,
                                       hipMemcpyDeviceToDevice, c->stream));
    c->turb_live_pending = false;
    return 0;
}

// =============================================================================================
// Radiation<P>::run (Radiation.h:245-487): grey two-stream radiation of every (j, k) column,
// i_base = max(i_topography, 0) the deep boundary and i = im-1 the top. Two kernels:
//   k_rad_layers   one thread per cell: the layer's emission B = sigma T^4, its emissivity
//                  eps = 1 - exp(-tau) (CIA + CH4/NH3 + capped cloud optical depth) and the
//                  downward shortwave flux Fsw at its bottom interface, all independent per cell;
//   k_rad_columns  one thread per column: the two recurrences (downward flux top to bottom, the
//                  deep boundary, upward flux bottom to top), the layer-centre net flux and the
//                  heating Q_rad, and the copy into the solid cells below i_base.
// eps, B and Fsw travel from the first kernel to the second in FP64: in the FP64 form through the
// epsilon, Q_rad and radiation arrays themselves (the second kernel overwrites Q_rad and radiation
// with the results after reading them), in the FP32-refined form through an FP64 buffer of three
// fields; Fd of the column is held in LDS. Columns with i_base >= im-1 are left untouched, as the
// reference skips them.
// =============================================================================================
struct RadConsts {
    double t_ref, g, cia_coeff, oc, sw_tau_bar, F_int;
    int bottom;
};
constexpr double RAD_SIGMA = 5.670374419e-8;            // Radiation.h:256
constexpr double RAD_P0 = 1.0e5;                         // Radiation.h:292
constexpr double RAD_KAPPA_CH4 = 0.003, RAD_KAPPA_NH3 = 1.5;          // Radiation.h:335-338
constexpr double RAD_KAPPA_CLOUD = 25.0, RAD_KAPPA_ICE = 12.0, RAD_TAU_CLOUD_CAP = 0.4;
constexpr int RAD_BK = 64;

template<typename T>
struct RadFields {
    const T *t, *p_stat, *rho, *ch4, *nh3, *cl[3], *ic[3];
    const float* lo_t;                     // FP32-refined form: t's low word
    T *eps, *rad, *qrad;
    double *e64, *b64, *s64;               // eps, B, Fsw between the two kernels
};

template<typename T>
__global__ __launch_bounds__(RAD_BK) void k_rad_layers(const RadFields<T> F, const int* __restrict__ itopo,
        const double* __restrict__ fsun, int im, int jm, int km, int pitch, const RadConsts C){
    const int k = blockIdx.x * RAD_BK + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const int i_base = max(itopo[(size_t)j * km + k], 0), i_top = im - 1;
    if(i_base >= i_top || i < i_base) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    double tq = pair_get(F.t, F.lo_t, q);
    const double Tk = tq * C.t_ref;
    const double B = RAD_SIGMA * Tk * Tk * Tk * Tk;
    const double P_lo = (double)F.p_stat[q] * RAD_P0;
    const double P_hi = (i < i_top) ? (double)F.p_stat[q + (size_t)jm * pitch] * RAD_P0 : 0.0;
    double dP = P_lo - P_hi;
    if(dP < 0.0) dP = 0.0;
    const double P_mean = 0.5 * (P_lo + P_hi);
    const double dm = dP / C.g;
    const double tau_cia = (Tk > 0.0) ? C.cia_coeff * (P_mean / Tk) * dP : 0.0;
    const double rho_c = (double)F.rho[q];
    const double inv_rho = (rho_c > 0.0 && isfinite(rho_c)) ? 1.0 / rho_c : 0.0;
    const double q_ch4 = dmax(0.0, (double)F.ch4[q]) * inv_rho;
    const double q_nh3 = dmax(0.0, (double)F.nh3[q]) * inv_rho;
    const double tau_gas = C.oc * (RAD_KAPPA_CH4 * q_ch4 + RAD_KAPPA_NH3 * q_nh3) * dm;
    const double q_liq = (dmax(0.0, (double)F.cl[0][q]) + dmax(0.0, (double)F.cl[1][q])
                        + dmax(0.0, (double)F.cl[2][q])) * inv_rho;
    const double q_ice = (dmax(0.0, (double)F.ic[0][q]) + dmax(0.0, (double)F.ic[1][q])
                        + dmax(0.0, (double)F.ic[2][q])) * inv_rho;
    double tau_cloud = C.oc * (RAD_KAPPA_CLOUD * q_liq + RAD_KAPPA_ICE * q_ice) * dm;
    if(tau_cloud > RAD_TAU_CLOUD_CAP) tau_cloud = RAD_TAU_CLOUD_CAP;
    const double tau = tau_cia + tau_gas + tau_cloud;
    F.e64[q] = 1.0 - exp(-tau);
    F.b64[q] = B;
    F.s64[q] = fsun[j] * exp(-C.sw_tau_bar * dmax(0.0, (double)F.p_stat[q]));
}

template<typename T>
__global__ __launch_bounds__(RAD_BK) void k_rad_columns(const RadFields<T> F, const int* __restrict__ itopo,
        const double* __restrict__ fsun, const double* __restrict__ dz_tab, int im, int jm, int km, int pitch,
        const RadConsts C){
    extern __shared__ double s_fd[];                    // [(im + 1) * RAD_BK]: Fd per interface
    const int k = blockIdx.x * RAD_BK + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const int i_base = max(itopo[(size_t)j * km + k], 0), i_top = im - 1;
    if(i_base >= i_top) return;
    const size_t P = (size_t)jm * pitch;
    const size_t q0 = (size_t)j * pitch + k;             // the column's cell at i = 0
    double* fd = s_fd + threadIdx.x;                     // fd[i * RAD_BK]
    // downward sweep: Fd[i] = Fd[i+1] (1 - eps_i) + eps_i B_i, Fd[i_top + 1] = 0 (space is cold)
    // (both sweeps load eight levels ahead of their in-order recurrences, so that the loads are in
    // flight together; a level read early is one no earlier level writes)
    double d = 0.0;
    fd[(i_top + 1) * RAD_BK] = 0.0;
    for(int i0 = i_top; i0 >= i_base; i0 -= 8){
        double ev[8], bv[8];
        #pragma unroll
        for(int u = 0; u < 8; u++){
            const int i = max(i0 - u, i_base);
            ev[u] = F.e64[q0 + i * P]; bv[u] = F.b64[q0 + i * P];
        }
        #pragma unroll
        for(int u = 0; u < 8; u++){
            const int i = i0 - u;
            if(i < i_base) break;
            d = d * (1.0 - ev[u]) + ev[u] * bv[u];
            fd[i * RAD_BK] = d;
        }
    }
    // the deep boundary: net upward F_int plus the shortwave left at depth, or (bottom = 1) no
    // divergence in the bottom layer
    double fsw = F.s64[q0 + i_base * P];
    double u;
    if(C.bottom == 1) u = dmax(0.0, 2.0 * F.b64[q0 + i_base * P] - fd[(i_base + 1) * RAD_BK]);
    else u = fd[i_base * RAD_BK] + C.F_int + fsw;
    // upward sweep with the layer results
    const double fsw_top = fsun[j] * exp(-C.sw_tau_bar * dmax(0.0, 0.0));
    for(int i0 = i_base; i0 <= i_top; i0 += 8){
        double ev[8], bv[8], sv[8];
        #pragma unroll
        for(int v = 0; v < 8; v++){
            const int i = min(i0 + v, i_top);
            ev[v] = F.e64[q0 + i * P]; bv[v] = F.b64[q0 + i * P];
            sv[v] = F.s64[q0 + min(i + 1, i_top) * P];
        }
        #pragma unroll
        for(int v = 0; v < 8; v++){
            const int i = i0 + v;
            if(i > i_top) break;
            const size_t q = q0 + i * P;
            const double e = ev[v], b = bv[v];
            const double u_next = u * (1.0 - e) + e * b;
            const double net_bot = u - fd[i * RAD_BK];
            const double net_top = u_next - fd[(i + 1) * RAD_BK];
            const double dz = dz_tab[i];
            const double fsw_next = (i < i_top) ? sv[v] : fsw_top;
            const double Q_sw = (fsw_next - fsw) / dz;
            if constexpr(!std::is_same<T, double>::value) F.eps[q] = (T)e;
            F.rad[q] = (T)(0.5 * (net_bot + net_top));
            F.qrad[q] = (T)((net_bot - net_top) / dz + Q_sw);
            u = u_next;
            fsw = fsw_next;
        }
    }
    // the solid cells below take the deepest fluid values, Q_rad 0
    const T e_base = F.eps[q0 + i_base * P], r_base = F.rad[q0 + i_base * P];
    for(int i = 0; i < i_base; i++){
        F.eps[q0 + i * P] = e_base;
        F.rad[q0 + i * P] = r_base;
        F.qrad[q0 + i * P] = T(0);
    }
}

template<typename T>
int do_radiation(atjup_ctx* c){
    const atjup_rad_params& R = c->radp;
    RadFields<T> Fp;
    Fp.t = F<T>(c, ATJUP_F_T); Fp.p_stat = F<T>(c, ATJUP_F_P_STAT); Fp.rho = F<T>(c, ATJUP_F_RHO_MIX);
    Fp.ch4 = F<T>(c, ATJUP_F_CH4); Fp.nh3 = F<T>(c, ATJUP_F_NH3);
    Fp.cl[0] = F<T>(c, ATJUP_F_H2O_CLOUD); Fp.cl[1] = F<T>(c, ATJUP_F_NH3_CLOUD); Fp.cl[2] = F<T>(c, ATJUP_F_CH4_CLOUD);
    Fp.ic[0] = F<T>(c, ATJUP_F_H2O_ICE); Fp.ic[1] = F<T>(c, ATJUP_F_NH3_ICE); Fp.ic[2] = F<T>(c, ATJUP_F_CH4_ICE);
    Fp.lo_t = (const float*)c->lo[ATJUP_F_T].p;
    Fp.eps = F<T>(c, ATJUP_F_EPSILON); Fp.rad = F<T>(c, ATJUP_F_RADIATION); Fp.qrad = F<T>(c, ATJUP_F_Q_RAD);
    if constexpr(std::is_same<T, double>::value){
        Fp.e64 = Fp.eps; Fp.b64 = Fp.qrad; Fp.s64 = Fp.rad;
    } else {
        Fp.e64 = (double*)c->rad_tmp.p; Fp.b64 = Fp.e64 + c->n3; Fp.s64 = Fp.b64 + c->n3;
    }
    // Radiation.h:272-280: the collision-pair factor, the CIA coefficient, the opacity calibration
    const double he_ratio = 0.6, C_cia = 3.5e-6, opac_cal = 0.025;
    const double comp = R.x_H2 * R.x_H2 + he_ratio * R.x_H2 * R.x_He;
    RadConsts C{ c->c.t_ref, c->c.g, C_cia * R.cia_strength * comp / c->c.g, R.opacity_strength * opac_cal,
                 R.sw_tau_per_bar, R.F_int, R.bottom };
    const double* fsun = (const double*)c->rad_tab.p;
    const double* dz = fsun + c->jm;
    hipEvent_t t0 = timer_begin(c);
    hipLaunchKernelGGL(k_rad_layers<T>, dim3(nblk(c->km, RAD_BK), c->jm, c->im), dim3(RAD_BK), 0, c->stream,
        Fp, (const int*)c->itopo.p, fsun, c->im, c->jm, c->km, c->pitch, C);
    const size_t lds = (size_t)(c->im + 1) * RAD_BK * sizeof(double);
    hipLaunchKernelGGL(k_rad_columns<T>, dim3(nblk(c->km, RAD_BK), c->jm), dim3(RAD_BK), lds, c->stream,
        Fp, (const int*)c->itopo.p, fsun, dz, c->im, c->jm, c->km, c->pitch, C);
    timer_end(c, ATJUP_T_RADIATION, t0, 2);
    return check_launch();
}

// =============================================================================================
// Precipitation<P>::run (Precipitation.h:194-574), per (j, k) column, in the reference's order per
// cell:
//   * the zeroing of Q_precip and the nine S_precip_* rates over the whole column (:199-210);
//   * the three-category column of H2O, NH3 and CH4 (column(), :273-450): a top-down sweep from
//     i = im-2 to the column's base with the rain, snow and graupel fluxes arriving from above
//     in registers, the autoconversion, accretion, riming, snow-to-graupel, melt/freeze handoff
//     and the capped rain evaporation, the latent heat added into Q_precip gas by gas, the
//     depletion as rates (coupling on) or in place (off), and the fluxes capped at 1e-4;
//   * the NH4SH crystal settling flux in its temperature band (sedimentNH4SH, :523-574);
//   * the surface maps at the column's base (surfaceMap, :470-492).
// The three gases are independent recurrences down the column, so a workgroup is three waves,
// one gas each, over the same 64 columns of the flattened (j, k) index. Per level each wave puts
// its latent-heat term in LDS and, after one barrier, wave 0 adds the three in the reference's
// gas order (0 + H2O + NH3 + CH4) into Q_precip; double-buffered LDS needs one barrier a level.
// Each wave loads the next level's inputs before it works on the current one.
// =============================================================================================
struct PrecipGas {
    double C, L0, R, da, db, ep, Lv, Ls, t_frz, t_low;
};
struct PrecipConsts {
    PrecipGas gas[3];
    double t_ref, r_mix, rs, nd, v_settle, t_hi_nh4sh, t_lo_nh4sh;
    int coupling;
};
template<typename T>
struct PrecipFields {
    const T *t, *p_stat, *rho, *nh4sh;
    const float* lo_t;
    T *vap[3], *cld[3], *ice[3];
    float *lo_vap[3], *lo_cld[3], *lo_ice[3];   // the refined form's low words (else null)
    T *P[3][3];                    // [gas][rain, snow, graupel]
    T *S[3][3];                    // [gas][vapour, cloud, ice]
    T *Q, *P_nh4sh;
    double* srf;                   // [5][jm][km]
};
constexpr double PR_C_C_AU = 2.0e-8, PR_C_I_AU = 5.0e-8, PR_Q_C_CRIT = 5.0e-4, PR_C_AC = 1.0e-7;
constexpr double PR_C_RIM = 1.0e-8, PR_Z_CSG = 2.5e-8, PR_A_EV = 2.76e-3, PR_C_MELT_K = 0.2;
constexpr double PR_C_FRZ_K = 0.2, PR_P_REF = 1.0e-4, PR_RATIO_MAX = 5.0, PR_V_FALL = 2.0;
constexpr double PR_P_MAX_FLUX = 1.0e-4, PR_DEP_FRAC = 1.0;          // Precipitation.h:39-85, :277-290
constexpr int PR_BK = 64;                                             // columns per workgroup

struct PrecipIn { double tq, p_u, rf, q_v, q_c, q_i; };

// A level's inputs as loaded (the storage type and, in the FP32-refined form, the pairs' low words),
// widened only when the level is worked on: the next level's loads stay in flight meanwhile.
template<typename T>
struct PrecipRaw { T t, p, rho, v, c, i; float lt, lv, lc, li; };

template<typename T>
__device__ __forceinline__ PrecipRaw<T> precip_load(const PrecipFields<T>& F, int g, size_t q){
    PrecipRaw<T> r;
    r.t = F.t[q]; r.p = F.p_stat[q]; r.rho = F.rho[q];
    r.v = F.vap[g][q]; r.c = F.cld[g][q]; r.i = F.ice[g][q];
    if constexpr(std::is_same<T, float>::value){
        // the low words without a branch (from the value's own array where a field has none)
        r.lt = (F.lo_t ? F.lo_t : F.t)[q];
        r.lv = (F.lo_vap[g] ? F.lo_vap[g] : F.vap[g])[q];
        r.lc = (F.lo_cld[g] ? F.lo_cld[g] : F.cld[g])[q];
        r.li = (F.lo_ice[g] ? F.lo_ice[g] : F.ice[g])[q];
    } else {
        r.lt = r.lv = r.lc = r.li = 0.0f;
    }
    return r;
}

template<typename T>
__device__ __forceinline__ PrecipIn precip_widen(const PrecipFields<T>& F, int g, const PrecipRaw<T>& r){
    PrecipIn w;
    w.tq = F.lo_t ? (double)r.t + (double)r.lt : (double)r.t;
    w.p_u = (double)r.p;
    w.rf = (double)r.rho;
    w.q_v = F.lo_vap[g] ? (double)r.v + (double)r.lv : (double)r.v;
    w.q_c = F.lo_cld[g] ? (double)r.c + (double)r.lc : (double)r.c;
    w.q_i = F.lo_ice[g] ? (double)r.i + (double)r.li : (double)r.i;
    return w;
}

// the NH4SH settling flux of one cell (sedimentNH4SH, Precipitation.h:540-556)
template<typename T>
__device__ __forceinline__ T nh4sh_flux(const PrecipFields<T>& F, size_t q, const PrecipConsts& C){
    double tq = pair_get(F.t, F.lo_t, q);
    const double Tk = tq * C.t_ref;
    const double qn = dmax(0.0, (double)F.nh4sh[q]);
    return (T)((Tk >= C.t_lo_nh4sh && Tk <= C.t_hi_nh4sh) ? qn * C.v_settle : 0.0);
}

template<typename T>
__global__ __launch_bounds__(3 * PR_BK) __attribute__((amdgpu_waves_per_eu(3))) void k_precip(const PrecipFields<T> F, const int* __restrict__ itopo,
        const double* __restrict__ dz_tab, int im, int jm, int km, int pitch, const PrecipConsts C){
    __shared__ double s_q[2][3][PR_BK];
    __shared__ double s_srf[3][PR_BK];
    const int lane = threadIdx.x % PR_BK, g = wave_uniform(threadIdx.x / PR_BK);
    const int n2 = jm * km, col_raw = blockIdx.x * PR_BK + lane;
    const bool valid = col_raw < n2;                  // every lane reaches every barrier
    const int col = valid ? col_raw : n2 - 1;
    const int j = col / km, k = col - j * km;
    const size_t P = (size_t)jm * pitch, q0 = (size_t)j * pitch + k;
    const int i_base = max(itopo[col], 0), i_top = im - 1;
    const bool live = i_base < i_top;                   // column() skips the others
    const PrecipGas G = C.gas[g];
    const double Lf = G.Ls - G.Lv;
    T* const Pr = F.P[g][0]; T* const Ps = F.P[g][1]; T* const Pg = F.P[g][2];
    double Fr = 0.0, Fs = 0.0, Fg = 0.0;
    PrecipRaw<T> nxt = precip_load(F, g, q0 + (size_t)i_top * P);
    for(int i = i_top; i >= 0; i--){
        const size_t q = q0 + (size_t)i * P;
        const PrecipIn in = precip_widen(F, g, nxt);
        if(i > 0) nxt = precip_load(F, g, q - P);
        double Qp = 0.0, Sv = 0.0, Sc = 0.0, Si = 0.0;
        if(valid && live){
            const double dz = dmax(0.0, dz_tab[i]);
            if(i == i_top || i < i_base || dz <= 0.0){
                // the top's zero fluxes; below the base and across a zero-thickness layer the
                // fluxes of the cell above (:300-303, :316-321, :444-448)
                if(i == i_top){ Fr = 0.0; Fs = 0.0; Fg = 0.0; }
                Pr[q] = (T)Fr; Ps[q] = (T)Fs; Pg[q] = (T)Fg;
            } else {
                const double Tk = in.tq * C.t_ref;
                const double p_u = in.p_u;
                const double rho_at = (in.rf > 0.0 && isfinite(in.rf)) ? in.rf : C.r_mix;
                const double q_v = dmax(0.0, in.q_v);
                const double q_c = dmax(0.0, in.q_c);
                const double q_i = dmax(0.0, in.q_i);
                const double E_r = exp(G.C + (-G.L0 / Tk + G.da * log(Tk) + G.db * Tk) / (1e-3 * G.R));
                const double q_sat = (p_u > E_r) ? rho_at * G.ep * E_r / p_u : q_v;
                const double F_r_in = Fr, F_s_in = Fs, F_g_in = Fg;
                const double Rain = dmin(PR_RATIO_MAX, F_r_in / PR_P_REF);
                const double Snow = dmin(PR_RATIO_MAX, F_s_in / PR_P_REF);
                const double Graupel = dmin(PR_RATIO_MAX, F_g_in / PR_P_REF);
                const double Rain_79 = (Rain > 0.0) ? pow(Rain, 7.0 / 9.0) : 0.0;
                const double Rain_49 = (Rain > 0.0) ? pow(Rain, 4.0 / 9.0) : 0.0;
                const bool warm = (Tk >= G.t_frz);
                const double S_c_au = (warm && q_c > PR_Q_C_CRIT) ? (C.rs * PR_C_C_AU) * (q_c - PR_Q_C_CRIT) : 0.0;
                const double S_i_au = (!warm && Tk >= G.t_low) ? (C.rs * PR_C_I_AU) * q_i : 0.0;
                const double S_ac = warm ? (C.rs * PR_C_AC) * q_c * Rain_79 : 0.0;
                const double S_s_rim = !warm ? (C.rs * PR_C_RIM) * q_c * Snow : 0.0;
                const double S_g_rim = !warm ? (C.rs * PR_C_RIM) * q_c * Graupel : 0.0;
                const double S_csg = (!warm && q_c > PR_Q_C_CRIT) ? (C.rs * PR_Z_CSG) * q_c * Snow : 0.0;
                const double f_melt = warm ? dmin(1.0, PR_C_MELT_K * (Tk - G.t_frz)) : 0.0;
                const double f_frz = warm ? 0.0 : dmin(1.0, PR_C_FRZ_K * (G.t_frz - Tk));
                const double F_s_melt = f_melt * F_s_in, F_g_melt = f_melt * F_g_in, F_r_frz = f_frz * F_r_in;
                const double F_r_avail = dmax(0.0, F_r_in - F_r_frz + F_s_melt + F_g_melt + (S_c_au + S_ac) * dz);
                const double S_ev_raw = (warm && Rain > 0.0 && q_v < q_sat) ? PR_A_EV * (q_sat - q_v) * Rain_49 : 0.0;
                const double F_ev_sat = dmax(0.0, q_sat - q_v) * PR_V_FALL;
                const double F_ev = dmin(dmin(S_ev_raw * dz, F_r_avail), F_ev_sat);
                const double S_ev = F_ev / dz;
                Qp = Lf * (S_s_rim + S_g_rim + S_csg) + Lf * (F_r_frz - F_s_melt - F_g_melt) / dz - G.Lv * F_ev / dz;
                const double tau = dz / PR_V_FALL;
                const double d_c = dmin(PR_DEP_FRAC * q_c, (S_c_au + S_ac + S_s_rim + S_g_rim + S_csg) * tau);
                const double d_i = dmin(PR_DEP_FRAC * q_i, S_i_au * tau);
                if(C.coupling){
                    Sc = -dmax(0.0, d_c) / tau * C.nd;
                    Si = -dmax(0.0, d_i) / tau * C.nd;
                    Sv = +S_ev * C.nd;
                } else {
                    pair_put(F.cld[g], F.lo_cld[g], q, q_c - dmax(0.0, d_c));
                    pair_put(F.ice[g], F.lo_ice[g], q, q_i - dmax(0.0, d_i));
                    pair_put(F.vap[g], F.lo_vap[g], q, q_v + S_ev * tau);
                }
                const double F_r = F_r_in - F_r_frz + F_s_melt + F_g_melt - F_ev + (S_c_au + S_ac) * dz;
                const double F_s = F_s_in - F_s_melt + (S_i_au + S_s_rim - S_csg) * dz;
                const double F_g = F_g_in - F_g_melt + F_r_frz + (S_g_rim + S_csg) * dz;
                Fr = dmin(PR_P_MAX_FLUX, dmax(0.0, F_r));
                Fs = dmin(PR_P_MAX_FLUX, dmax(0.0, F_s));
                Fg = dmin(PR_P_MAX_FLUX, dmax(0.0, F_g));
                Pr[q] = (T)Fr; Ps[q] = (T)Fs; Pg[q] = (T)Fg;
            }
        }
        if(valid){ F.S[g][0][q] = (T)Sv; F.S[g][1][q] = (T)Sc; F.S[g][2][q] = (T)Si; }
        s_q[i & 1][g][lane] = Qp;
        __syncthreads();
        if(g == 0 && valid){
            double s = 0.0;
            s += s_q[i & 1][0][lane];
            s += s_q[i & 1][1][lane];
            s += s_q[i & 1][2][lane];
            F.Q[q] = (T)s;
        }
    }
    // NH4SH crystal settling in its band, the solid cells below taking the base's flux; the
    // three waves take every third level
    const T base_flux = (i_base < im) ? nh4sh_flux(F, q0 + (size_t)i_base * P, C) : T(0);
    if(valid)
        for(int i = g; i < im; i += 3){
            const size_t q = q0 + (size_t)i * P;
            F.P_nh4sh[q] = (i >= i_base) ? nh4sh_flux(F, q, C) : base_flux;
        }
    // the surface maps at the base (a fully solid column reads its top cell); each wave reads
    // back its own gas's fluxes
    const size_t qb = q0 + (size_t)min(i_base, im - 1) * P;
    s_srf[g][lane] = (double)Pr[qb] + (double)Ps[qb] + (double)Pg[qb];
    __syncthreads();
    if(g == 0 && valid){
        const double srf0 = s_srf[0][lane], srf1 = s_srf[1][lane], srf2 = s_srf[2][lane];
        const double nh4 = (double)base_flux;
        const size_t c2 = (size_t)col;
        F.srf[c2] = srf0;
        F.srf[(size_t)n2 + c2] = srf1;
        F.srf[2 * (size_t)n2 + c2] = srf2;
        F.srf[3 * (size_t)n2 + c2] = nh4;
        F.srf[4 * (size_t)n2 + c2] = srf0 + srf1 + srf2 + nh4;
    }
}

template<typename T>
int do_precipitation(atjup_ctx* c){
    static const int vap[3] = { ATJUP_F_H2O, ATJUP_F_NH3, ATJUP_F_CH4 };
    static const int pf[3] = { ATJUP_F_P_RAIN, ATJUP_F_P_NH3_RAIN, ATJUP_F_P_CH4_RAIN };
    static const int sf[3] = { ATJUP_F_S_PRECIP_H2O, ATJUP_F_S_PRECIP_NH3, ATJUP_F_S_PRECIP_CH4 };
    PrecipFields<T> Fp;
    Fp.t = F<T>(c, ATJUP_F_T); Fp.p_stat = F<T>(c, ATJUP_F_P_STAT); Fp.rho = F<T>(c, ATJUP_F_RHO_MIX);
    Fp.nh4sh = F<T>(c, ATJUP_F_NH4SH); Fp.lo_t = (const float*)c->lo[ATJUP_F_T].p;
    PrecipConsts C;
    for(int g = 0; g < 3; g++){
        Fp.vap[g] = F<T>(c, vap[g]); Fp.cld[g] = F<T>(c, vap[g] + 1); Fp.ice[g] = F<T>(c, vap[g] + 2);
        Fp.lo_vap[g] = (float*)c->lo[vap[g]].p; Fp.lo_cld[g] = (float*)c->lo[vap[g] + 1].p;
        Fp.lo_ice[g] = (float*)c->lo[vap[g] + 2].p;
        for(int r = 0; r < 3; r++){ Fp.P[g][r] = F<T>(c, pf[g] + r); Fp.S[g][r] = F<T>(c, sf[g] + r); }
        const atjup_gas_params& a = c->gas[g];
        C.gas[g] = PrecipGas{ a.C, a.L0, a.R, a.del_alf, a.del_bet, a.ep, a.lv, a.ls, a.t_0, a.t_00 };
    }
    Fp.Q = F<T>(c, ATJUP_F_Q_PRECIP); Fp.P_nh4sh = F<T>(c, ATJUP_F_P_NH4SH);
    Fp.srf = (double*)c->precip_srf.p;
    const atjup_precip_params& pp = c->precp;
    // Stokes settling of the crystals (Precipitation.h:530-538): r_p 1e-5 m, 1200 kg/m3, mu 9e-6 Pa s
    const double r_p = 1.0e-5, rho_c = 1200.0, mu = 9.0e-6;
    C.t_ref = c->c.t_ref; C.r_mix = c->c.r_mix; C.rs = pp.rate_scale;
    C.nd = (c->g.L_atm * 1.0e3) / c->c.u_0;
    C.v_settle = (2.0 / 9.0) * r_p * r_p * (rho_c - c->c.r_mix) * c->c.g / mu;
    C.t_hi_nh4sh = pp.t_0_nh4sh; C.t_lo_nh4sh = pp.t_00_nh4sh;
    C.coupling = (c->kn.precip_coupling != 0.0) ? 1 : 0;
    const double* dz = (const double*)c->tabs.p + c->im + 3 * c->jm;    // layer_thickness_m
    hipEvent_t t0 = timer_begin(c);
    hipLaunchKernelGGL(k_precip<T>, dim3(nblk(c->jm * c->km, PR_BK)), dim3(3 * PR_BK), 0, c->stream,
        Fp, (const int*)c->itopo.p, dz, c->im, c->jm, c->km, c->pitch, C);
    timer_end(c, ATJUP_T_PRECIPITATION, t0, 1);
    return check_launch();
}

// =============================================================================================
// ChemistryJup (ChemistryJup.h): DiffMassFluxJup (:375-540) and ChemMassRateJup (:59-213).
// k_diffmassflux: per cell of i 1..im-2, j 1..jm-2, k 0..km-2 (the periodic seam), the first
//   derivatives of t, H2S, NH3, NH4SH (central, 0 across two solid neighbours; at k = 0 the
//   one-sided phi form of derivative_1_order_boundary), the multicomponent fluxes j_*, the
//   thermo-diffusion jT_*, the Fickian difflux_* = D laplacian and thermalmassflux, which reads
//   the reaction rates w_* the previous call left; solid cells and t <= 0 get zeros.
// k_chemmassrate: the polar bands j < 3 and j > jm-4 zeroed; per cell of i 1..im-2, j 3..jm-4,
//   k 0..km-2 the NH4SH formation rate law inside [t_00_nh4sh, t_0_nh4sh] (zero outside with
//   GATE_ZERO, else the last rate), a solid cell with six solid neighbours zeroed, and
//   massflux_* = w_* + difflux_*.
// =============================================================================================
struct ChemConsts {
    double D_nh3, D_h2s, D_nh4sh, M_mix, L_nh3, L_h2s, L_nh4sh, LT_nh3, LT_h2s, LT_nh4sh;
    double m_nh3, m_h2s, m_nh4sh, cp_nh3, cp_h2s, cp_nh4sh, k_nh3, k_h2s, k_nh4sh, dh_nh4sh;
    double r_mix, t_ref, keq, t_0, t_00;
    double dr_2, dthe_2, dphi_2, dr2, dthe2, dphi2;
    int gate_zero, molar_conc, enthalpy, eq_on;
};
constexpr double CHEM_FLOOR = 0.4;                 // polar_divisor_floor, SINTHE_TRACK 0 (ATPhys.h:104-108)

template<typename T>
struct ChemFields {
    const T *t, *h2s, *nh3, *nh4sh, *rho;
    const float* lo_t;
    T *dfl_h2s, *dfl_nh3, *dfl_nh4sh, *tmf, *mf_h2s, *mf_nh3, *mf_nh4sh;
    T *w_nh3, *w_h2s, *w_nh4sh;           // the reaction rates, in the fields' precision (the
                                          // boundary passes treat them with the fields)
};

// derivative_1_order (ChemistryJup.h:653-727) at a fluid cell: central unless both neighbours of
// the direction are solid (then 0)
__device__ __forceinline__ double chem_d1(double cm, double cp, bool sm, bool sp, double two_d){
    return (sm && sp) ? 0.0 : (cp - cm) / two_d;
}

template<typename T>
__global__ __launch_bounds__(64) void k_diffmassflux(const ChemFields<T> F, const uint8_t* __restrict__ flags,
        const double* __restrict__ tabs, int im, int jm, int km, int pitch, const ChemConsts C){
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y + 1, i = blockIdx.z + 1;
    if(k > km - 2) return;
    const size_t P = (size_t)jm * pitch, R = (size_t)pitch;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const size_t row = q - k;
    const int kw = (k == 0) ? km - 2 : k - 1, ke = (k == km - 2) ? 0 : k + 1;
    const uint8_t fl = flags[q];
    double tq = pair_get(F.t, F.lo_t, q);
    if(!(tq > 0.0) || (fl & FL_SELF)){
        F.dfl_h2s[q] = T(0); F.dfl_nh3[q] = T(0); F.dfl_nh4sh[q] = T(0); F.tmf[q] = T(0);
        return;
    }
    const double rm = tabs[i];
    const double sin_t = tabs[im + j], cos_t = tabs[im + 2 * jm + j];
    const double sinthe = dmax(sin_t, CHEM_FLOOR);
    const double rmsinthe = rm * sinthe;
    const bool sim = fl & FL_IM, sip = fl & FL_IP, sjm = fl & FL_JM, sjp = fl & FL_JP;
    const bool skm = fl & FL_KM, skp = fl & FL_KP;
    // the seven-point values of the four fields, loaded once before any store
    struct Pts { double c, rp, rm, tp, tm, pw, pe, k1, k2; };
    auto pts = [&](const T* X, const float* L){
        auto ld = [&](size_t qq){ double v = (double)X[qq]; if(L) v += (double)L[qq]; return v; };
        Pts p;
        p.c = ld(q); p.rp = ld(q + P); p.rm = ld(q - P); p.tp = ld(q + R); p.tm = ld(q - R);
        p.pw = ld(row + kw); p.pe = ld(row + ke);
        p.k1 = (k == 0) ? ld(row + 1) : 0.0; p.k2 = (k == 0) ? ld(row + 2) : 0.0;
        return p;
    };
    const Pts Pt = pts(F.t, F.lo_t), Ph = pts(F.h2s, nullptr), Pn = pts(F.nh3, nullptr), P4 = pts(F.nh4sh, nullptr);
    // first derivatives of a field: r, theta central (0 across two solid neighbours); phi central,
    // one-sided at k = 0 (derivative_1_order_boundary, ChemistryJup.h:729-750)
    auto d3 = [&](const Pts& X, double& dr, double& dth, double& dph){
        dr = chem_d1(X.rm, X.rp, sim, sip, C.dr_2);
        dth = chem_d1(X.tm, X.tp, sjm, sjp, C.dthe_2);
        if(k == 0) dph = (-3.0 * X.c + 4.0 * X.k1 - X.k2) / C.dphi_2;
        else dph = chem_d1(X.pw, X.pe, skm, skp, C.dphi_2);
    };
    double dtdr, dtdthe, dtdphi, dnh3dr, dnh3dthe, dnh3dphi, dh2sdr, dh2sdthe, dh2sdphi, dnhdr, dnhdthe, dnhdphi;
    d3(Pt, dtdr, dtdthe, dtdphi);
    d3(Ph, dh2sdr, dh2sdthe, dh2sdphi);
    d3(Pn, dnh3dr, dnh3dthe, dnh3dphi);
    d3(P4, dnhdr, dnhdthe, dnhdphi);
    const double dt_div = dtdr + fabs(dtdthe) / rm + dtdphi / rmsinthe;
    const double rf = (double)F.rho[q];
    const double rho_c = (rf > 0.0 && isfinite(rf)) ? rf : C.r_mix;
    const double sc = rho_c / C.r_mix;
    const double jT_nh3 = (C.LT_nh3 * sc) / tq * dt_div;
    const double jT_h2s = (C.LT_h2s * sc) / tq * dt_div;
    const double jT_nh4sh = (C.LT_nh4sh * sc) / tq * dt_div;
    const double dnh3_div = dnh3dr + fabs(dnh3dthe) / rm + dnh3dphi / rmsinthe;
    const double dh2s_div = dh2sdr + fabs(dh2sdthe) / rm + dh2sdphi / rmsinthe;
    const double dnh4sh_div = dnhdr + fabs(dnhdthe) / rm + dnhdphi / rmsinthe;
    const double L_nh3_c = C.L_nh3 * sc, L_h2s_c = C.L_h2s * sc, L_nh4sh_c = C.L_nh4sh * sc;
    const double nh3 = Pn.c, h2s = Ph.c, nh4sh = P4.c;
    const double cM_nh3 = L_nh3_c * nh3 / rho_c;
    const double cM_h2s = L_h2s_c * h2s / rho_c;
    const double cM_nh4sh = L_nh4sh_c * nh4sh / rho_c;
    const double M = C.M_mix;
    const double j_nh3 = C.m_nh3 / M * (L_nh3_c * dnh3_div - M / C.m_nh3 * cM_nh3 * dnh3_div
        - M / C.m_h2s * cM_nh3 * dh2s_div - M / C.m_nh4sh * cM_nh3 * dnh4sh_div) - jT_nh3;
    const double j_h2s = C.m_h2s / M * (L_h2s_c * dh2s_div - M / C.m_nh3 * cM_h2s * dnh3_div
        - M / C.m_h2s * cM_h2s * dh2s_div - M / C.m_nh4sh * cM_h2s * dnh4sh_div) - jT_h2s;
    const double j_nh4sh = C.m_nh4sh / M * (L_nh4sh_c * dnh4sh_div - M / C.m_nh3 * cM_nh4sh * dnh3_div
        - M / C.m_h2s * cM_nh4sh * dh2s_div - M / C.m_nh4sh * cM_nh4sh * dnh4sh_div) - jT_nh4sh;
    // laplacian_spherical (ChemistryJup.h:752-776)
    auto lap = [&](const Pts& X){
        const double c = X.c;
        const double rmsin = rm * dmax(sin_t, CHEM_FLOOR);
        const double d2r = (X.rp - 2.0 * c + X.rm) / C.dr2 * 1.0;
        const double d2t = (X.tp - 2.0 * c + X.tm) / C.dthe2;
        const double d2p = (X.pe - 2.0 * c + X.pw) / C.dphi2;
        const double d1r = (X.rp - X.rm) / C.dr_2 * 1.0;
        const double d1t = (X.tp - X.tm) / C.dthe_2;
        return d2r + 2.0 / rm * d1r + d2t / (rm * rm) + cos_t / (rm * rmsin) * d1t + d2p / (rmsin * rmsin);
    };
    const double lap_h2s = lap(Ph), lap_nh3 = lap(Pn), lap_nh4sh = lap(P4);
    const double wt = C.enthalpy ? F.w_nh4sh[q] * C.dh_nh4sh
        : tq * (F.w_nh3[q] * C.k_nh3 + F.w_h2s[q] * C.k_h2s + F.w_nh4sh[q] * C.k_nh4sh);
    F.dfl_h2s[q] = (T)(C.D_h2s * lap_h2s);
    F.dfl_nh3[q] = (T)(C.D_nh3 * lap_nh3);
    F.dfl_nh4sh[q] = (T)(C.D_nh4sh * lap_nh4sh);
    F.tmf[q] = (T)((j_nh3 * C.cp_nh3 + j_h2s * C.cp_h2s + j_nh4sh * C.cp_nh4sh) * dt_div + wt);
}

template<typename T>
__global__ __launch_bounds__(64) void k_chemmassrate(const ChemFields<T> F, const uint8_t* __restrict__ msk,
        int im, int jm, int km, int pitch, const ChemConsts C){
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t P = (size_t)jm * pitch, R = (size_t)pitch;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    if(j < 3 || j > jm - 4){
        F.w_nh3[q] = T(0); F.w_h2s[q] = T(0); F.w_nh4sh[q] = T(0);
        F.mf_nh3[q] = T(0); F.mf_h2s[q] = T(0); F.mf_nh4sh[q] = T(0);
        return;
    }
    if(i < 1 || i > im - 2 || k > km - 2) return;
    const size_t row = q - k;
    const int kw = (k == 0) ? km - 2 : k - 1, ke = (k == km - 2) ? 0 : k + 1;
    if(msk[q] && msk[q - P] && msk[q + P] && msk[q - R] && msk[q + R] && msk[row + kw] && msk[row + ke]){
        F.w_nh3[q] = T(0); F.w_h2s[q] = T(0); F.w_nh4sh[q] = T(0);
        F.mf_nh3[q] = T(0); F.mf_h2s[q] = T(0); F.mf_nh4sh[q] = T(0);
        return;
    }
    double tq = pair_get(F.t, F.lo_t, q);
    const double t_u = tq * C.t_ref;
    const double kf = 15000.0 * pow(t_u, -0.5) * exp(-3020.0 / t_u);        // react_rate_const (:648-651)
    const double kb = kf / C.keq;
    double w_nh3 = F.w_nh3[q], w_h2s = F.w_h2s[q], w_nh4sh = F.w_nh4sh[q];
    // with the NH4SH equilibrium adjustment the rate law contributes nothing (ChemistryJup.h:133-135)
    if(!C.eq_on && t_u <= C.t_0 && t_u >= C.t_00){
        const double nh3 = (double)F.nh3[q], h2s = (double)F.h2s[q], nh4sh = (double)F.nh4sh[q];
        double denom;
        if(C.molar_conc){
            double rho = (double)F.rho[q];
            if(!(rho > 0.0)) rho = C.r_mix;
            denom = rho;
        } else {
            const double sum_c = nh3 / C.m_nh3 + h2s / C.m_h2s + nh4sh / C.m_nh4sh;
            denom = (sum_c == 0.0) ? 0.0 : C.r_mix / sum_c;
        }
        const double c_nh3 = nh3 / C.m_nh3 * denom, c_h2s = h2s / C.m_h2s * denom, c_nh4sh = nh4sh / C.m_nh4sh * denom;
        const double R_diff = kf * c_nh3 * c_h2s - kb * c_nh4sh;
        w_nh3 = -C.m_nh3 * R_diff;
        w_h2s = -C.m_h2s * R_diff;
        w_nh4sh = C.m_nh4sh * R_diff;
    } else if(C.gate_zero || C.eq_on){
        w_nh3 = 0.0; w_h2s = 0.0; w_nh4sh = 0.0;
    }
    F.w_nh3[q] = (T)w_nh3; F.w_h2s[q] = (T)w_h2s; F.w_nh4sh[q] = (T)w_nh4sh;
    F.mf_h2s[q] = (T)(w_h2s + (double)F.dfl_h2s[q]);
    F.mf_nh3[q] = (T)(w_nh3 + (double)F.dfl_nh3[q]);
    F.mf_nh4sh[q] = (T)(w_nh4sh + (double)F.dfl_nh4sh[q]);
}

template<typename T>
int do_chem_kernels(atjup_ctx* c, int which){
    const atjup_chem_params& p = c->chemp;
    ChemFields<T> Fp;
    Fp.t = F<T>(c, ATJUP_F_T); Fp.h2s = F<T>(c, ATJUP_F_H2S); Fp.nh3 = F<T>(c, ATJUP_F_NH3);
    Fp.nh4sh = F<T>(c, ATJUP_F_NH4SH); Fp.rho = F<T>(c, ATJUP_F_RHO_MIX);
    Fp.lo_t = (const float*)c->lo[ATJUP_F_T].p;
    Fp.dfl_h2s = F<T>(c, ATJUP_F_DIFFLUX_H2S); Fp.dfl_nh3 = F<T>(c, ATJUP_F_DIFFLUX_NH3);
    Fp.dfl_nh4sh = F<T>(c, ATJUP_F_DIFFLUX_NH4SH); Fp.tmf = F<T>(c, ATJUP_F_THERMALMASSFLUX);
    Fp.mf_h2s = F<T>(c, ATJUP_F_MASSFLUX_H2S); Fp.mf_nh3 = F<T>(c, ATJUP_F_MASSFLUX_NH3);
    Fp.mf_nh4sh = F<T>(c, ATJUP_F_MASSFLUX_NH4SH);
    Fp.w_nh3 = (T*)c->chem_w.p; Fp.w_h2s = Fp.w_nh3 + c->n3; Fp.w_nh4sh = Fp.w_h2s + c->n3;
    // the constants of DiffMassFluxJup (:383-400) and ChemMassRateJup (:66-69)
    ChemConsts C;
    const double DT_nh3 = p.mue_nh3 / (p.rg_nh3 * c->c.sc_nh3);
    const double DT_h2s = p.mue_h2s / (p.rg_h2s * c->c.sc_h2s);
    const double DT_nh4sh = p.mue_nh4sh / (c->c.rg_nh4sh * c->c.sc_nh4sh);
    C.D_nh3 = DT_nh3; C.D_h2s = DT_h2s; C.D_nh4sh = DT_nh4sh;
    C.M_mix = p.X_h2 * p.m_h2 + p.X_he * p.m_he + p.X_h2o * p.m_h2o + p.X_h2s * p.m_h2s
            + p.X_nh3 * p.m_nh3 + p.X_nh4sh * p.m_nh4sh;
    const double rc = c->c.r_mix * c->c.cp_mix;
    C.L_nh3 = rc * DT_nh3 / p.k_mix; C.L_h2s = rc * DT_h2s / p.k_mix; C.L_nh4sh = rc * DT_nh4sh / p.k_mix;
    C.LT_nh3 = C.L_nh3; C.LT_h2s = C.L_h2s; C.LT_nh4sh = C.L_nh4sh;
    C.m_nh3 = p.m_nh3; C.m_h2s = p.m_h2s; C.m_nh4sh = p.m_nh4sh;
    C.cp_nh3 = p.cp_nh3; C.cp_h2s = p.cp_h2s; C.cp_nh4sh = p.cp_nh4sh;
    C.k_nh3 = p.k_nh3; C.k_h2s = p.k_h2s; C.k_nh4sh = p.k_nh4sh; C.dh_nh4sh = p.dh_nh4sh;
    C.r_mix = c->c.r_mix; C.t_ref = c->c.t_ref; C.keq = p.m_nh4sh / (p.m_nh3 * p.m_h2s);
    C.t_0 = p.t_0_nh4sh; C.t_00 = p.t_00_nh4sh;
    C.dr_2 = 2.0 * c->g.dr; C.dthe_2 = 2.0 * c->g.dthe; C.dphi_2 = 2.0 * c->g.dphi;
    C.dr2 = c->g.dr * c->g.dr; C.dthe2 = c->g.dthe * c->g.dthe; C.dphi2 = c->g.dphi * c->g.dphi;
    C.gate_zero = p.gate_zero; C.molar_conc = p.molar_conc; C.enthalpy = p.enthalpy;
    C.eq_on = c->kn.nh4sh_eq != 0;
    hipEvent_t t0 = timer_begin(c);
    if(which == 0)
        hipLaunchKernelGGL(k_diffmassflux<T>, dim3(nblk(c->km - 1, 64), c->jm - 2, c->im - 2), dim3(64), 0, c->stream,
            Fp, (const uint8_t*)c->flags.p, (const double*)c->tabs.p, c->im, c->jm, c->km, c->pitch, C);
    else
        hipLaunchKernelGGL(k_chemmassrate<T>, dim3(nblk(c->km, 64), c->jm, c->im), dim3(64), 0, c->stream,
            Fp, (const uint8_t*)c->mask.p, c->im, c->jm, c->km, c->pitch, C);
    timer_end(c, ATJUP_T_CHEMISTRY, t0, 1);
    return check_launch();
}

// =============================================================================================
// MoistConvectiveAdjustment<P>::run (ConvectiveAdjustment.h:403-849) with the gases H2O, NH3 and
// the tracers H2S, NH4SH (cJupiterModel.cpp:1301-1322); SMOOTH 0, PHASE 0, INTERIOR 1, cp(T) off.
// One thread per (j, k) column of the flattened index. A column's working state (h_L, the total
// H2O / NH3, the tracers per unit mass, each cell's own equilibrium temperature and condensate,
// and the touched / equilibrium-known / census bits) lives in a workspace laid out [var][i][column]
// so that a wavefront's 64 columns read one level as one contiguous line; p, rho, z and the layer
// weights are read from the fields at each use, and the unmixed state h_L0 / q_t0 that MCA_RELAX
// blends with is formed again from the fields, which stay untouched until the write-back.
// Per column, as the reference: the parcel census of every adjacent pair, sweeps that mix each
// unstable segment (grown while its ends stay unstable) and settle its cells by the Illinois
// equilibrium solve, the optional relaxation, and the write-back of the touched cells.
// =============================================================================================
struct McaGas { double t_0, ep, lv, ls, C, L0, R, da, db, Ci, L0i, dai, dbi; };
struct McaConsts {
    McaGas gas[2];
    double cp, grav, t_ref, r_mix, relax, ep_h2s, l_nh4sh;
    int ledoux, loading, max_pass;
};
template<typename T>
struct McaFields {
    T *t, *v[2], *cl[2], *ice[2], *tr[2];        // tr: H2S, NH4SH
    float* lo_t;
    float *lo_v[2], *lo_cl[2], *lo_ice[2], *lo_tr[2];   // the refined form's low words (else null)
    const T *p_stat, *rho;
};
enum { MV_HL = 0, MV_QT = 1, MV_QR = 3, MV_TEQ = 5, MV_CEQ = 6, MCA_NV = 8 };
enum { MF_TOUCHED = 1, MF_EQ_OK = 2 };
enum { MC_UNSTABLE = 1, MC_VALID = 2 };                              // the per-pair test memo
constexpr double MCA_TOL_K = 1.0e-7, MCA_TOL_T = 1.0e-10;            // ConvectiveAdjustment.h:416-418
constexpr int MCA_N_ITER = 60;

__device__ __forceinline__ double mca_qsat(const McaGas& s, double T, double p){
    const double E = (T >= s.t_0) ? svp_d(T, s.C, s.L0, s.R, s.da, s.db) : svp_d(T, s.Ci, s.L0i, s.R, s.dai, s.dbi);
    const double q = s.ep * E / p;
    return isfinite(q) ? q : 0.0;
}
__device__ __forceinline__ double mca_lat(const McaGas& s, double T){ return T >= s.t_0 ? s.lv : s.ls; }

// the equilibrium temperature of (h_L, q_t) at (p, z): cp T + g z - sum L(T) max(0, q_t - q_sat(T))
// = h_L, bracketed by the all-vapour value and T_lo + sum L q_t / cp, by Illinois regula falsi
// (ConvectiveAdjustment.h:453-478)
__device__ __forceinline__ bool mca_equil(const McaConsts& K, double hL, double qt0, double qt1, double p, double z,
                          double& T, double& c0, double& c1){
    const double T_lo = (hL - K.grav * z) / K.cp;
    if(!(T_lo > 0.0) || !(p > 0.0)) return false;
    if(!(qt0 > mca_qsat(K.gas[0], T_lo, p)) && !(qt1 > mca_qsat(K.gas[1], T_lo, p))){
        T = T_lo; c0 = 0.0; c1 = 0.0;
        return true;
    }
    double span = 0.0;
    span += dmax(K.gas[0].lv, K.gas[0].ls) * dmax(0.0, qt0);
    span += dmax(K.gas[1].lv, K.gas[1].ls) * dmax(0.0, qt1);
    auto fof = [&](double TT){
        double f = K.cp * TT + K.grav * z - hL;
        f -= mca_lat(K.gas[0], TT) * dmax(0.0, qt0 - mca_qsat(K.gas[0], TT, p));
        f -= mca_lat(K.gas[1], TT) * dmax(0.0, qt1 - mca_qsat(K.gas[1], TT, p));
        return f;
    };
    double a = T_lo, b = T_lo + span / K.cp + 1.0e-9;
    double fa = fof(a), fb = fof(b);
    int side = 0;
    for(int it = 0; it < MCA_N_ITER && b - a > MCA_TOL_T; it++){
        double c = (fb - fa != 0.0) ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
        if(!(c > a && c < b)) c = 0.5 * (a + b);
        const double fc = fof(c);
        if(fc < 0.0){ a = c; fa = fc; if(side == -1) fb *= 0.5; side = -1; }
        else        { b = c; fb = fc; if(side == +1) fa *= 0.5; side = +1; }
        if(fabs(fc) <= K.cp * MCA_TOL_T){ a = b = c; break; }
    }
    T = b;                                      // the side where the condensate is not overcounted
    c0 = dmax(0.0, qt0 - mca_qsat(K.gas[0], T, p));
    c1 = dmax(0.0, qt1 - mca_qsat(K.gas[1], T, p));
    return true;
}

template<typename T>
struct McaColumn {
    const McaFields<T>& F;
    const McaConsts& K;
    const double* ztab;
    double* ws;
    uint8_t* wf;
    size_t VS, P, q0;
    int n2, col, i0, i1;

    __device__ double& W(int var, int i) const { return ws[(size_t)var * VS + (size_t)i * n2 + col]; }
    __device__ uint8_t& FL(int i) const { return wf[(size_t)i * n2 + col]; }
    __device__ double Tk(int i) const {
        const size_t q = q0 + (size_t)i * P;
        double tq = pair_get(F.t, F.lo_t, q);
        return tq * K.t_ref;
    }
    __device__ double pr(int i) const { return (double)F.p_stat[q0 + (size_t)i * P]; }
    __device__ double rh(int i) const {                  // rho_at (cJupiterModel.h:1212-1216)
        const double rf = (double)F.rho[q0 + (size_t)i * P];
        return (rf > 0.0 && isfinite(rf)) ? rf : K.r_mix;
    }
    __device__ double wgt(int i) const {                 // rho dz, face-centred (:598-601)
        const double lo = (i > i0) ? 0.5 * (ztab[i - 1] + ztab[i]) : ztab[i0];
        const double hi = (i < i1) ? 0.5 * (ztab[i] + ztab[i + 1]) : ztab[i1];
        return rh(i) * (hi - lo);
    }
    // h_L, q_t and the tracers of a cell from the fields (:602-616)
    __device__ void init(int i, double& h, double* qt, double* qr) const {
        const size_t q = q0 + (size_t)i * P;
        const double rho = rh(i);
        h = K.cp * Tk(i) + K.grav * ztab[i];
        for(int g = 0; g < 2; g++){
            const double qv = dmax(0.0, (double)F.v[g][q]) / rho;
            const double qc = dmax(0.0, (double)F.cl[g][q]) / rho;
            const double qi = dmax(0.0, (double)F.ice[g][q]) / rho;
            qt[g] = qv + qc + qi;
            h -= K.gas[g].lv * qc + K.gas[g].ls * qi;
        }
        qr[0] = (double)F.tr[0][q] / rho;
        h -= 0.0 * qr[0];
        qr[1] = (double)F.tr[1][q] / rho;
        h -= K.l_nh4sh * qr[1];
    }
    __device__ double trE(int i) const {
        double e = 0.0;
        e += 0.0 * W(MV_QR, i);
        e += K.l_nh4sh * W(MV_QR + 1, i);
        return e;
    }
    // the virtual-temperature factor, composition only with LEDOUX (:636-647)
    __device__ double tv_q(double TT, double v0, double v1, double c0, double c1, double r0, double r1) const {
        if(!K.ledoux) return TT;
        double f = 1.0;
        f += v0 * (1.0 / K.gas[0].ep - 1.0);
        if(K.loading) f -= c0;
        f += v1 * (1.0 / K.gas[1].ep - 1.0);
        if(K.loading) f -= c1;
        if(K.ep_h2s > 0.0) f += r0 * (1.0 / K.ep_h2s - 1.0);
        if(K.loading) f -= r1;
        return TT * f;
    }
    __device__ double tv(double TT, double v0, double v1, double c0, double c1, int i) const {
        return tv_q(TT, v0, v1, c0, c1, W(MV_QR, i), W(MV_QR + 1, i));
    }
    // the parcel from a lifted to b against b in its own equilibrium (:651-673)
    __device__ bool unstable(int a, int b) const {
        double Tp, pc0, pc1;
        if(!mca_equil(K, W(MV_HL, a) + trE(a), W(MV_QT, a), W(MV_QT + 1, a), pr(b), ztab[b], Tp, pc0, pc1)) return false;
        if(!(FL(b) & MF_EQ_OK)){
            double Te, e0, e1;
            if(!mca_equil(K, W(MV_HL, b) + trE(b), W(MV_QT, b), W(MV_QT + 1, b), pr(b), ztab[b], Te, e0, e1)) return false;
            W(MV_TEQ, b) = Te; W(MV_CEQ, b) = e0; W(MV_CEQ + 1, b) = e1;
            FL(b) |= MF_EQ_OK;
        }
        const double ec0 = W(MV_CEQ, b), ec1 = W(MV_CEQ + 1, b);
        const double pv0 = W(MV_QT, a) - pc0, pv1 = W(MV_QT + 1, a) - pc1;
        const double ev0 = W(MV_QT, b) - ec0, ev1 = W(MV_QT + 1, b) - ec1;
        return tv(Tp, pv0, pv1, pc0, pc1, a) - tv(W(MV_TEQ, b), ev0, ev1, ec0, ec1, b) > MCA_TOL_K;
    }
    // a cell to its equilibrium from h_L, q_t (:674-690); the vapour and the liquid / ice split
    // follow from (q_t, condensate, T) at the write-back
    __device__ bool settle(int q) const {
        double TT, c0, c1;
        if(!mca_equil(K, W(MV_HL, q) + trE(q), W(MV_QT, q), W(MV_QT + 1, q), pr(q), ztab[q], TT, c0, c1)) return false;
        W(MV_TEQ, q) = TT; W(MV_CEQ, q) = c0; W(MV_CEQ + 1, q) = c1;
        FL(q) |= MF_TOUCHED | MF_EQ_OK;
        return true;
    }
};

// the column's fluid range with the radial boundary planes left out (:553-575), or i1 < i0 + 1
// when the column is skipped: too short, or a non-physical state (:578-600)
template<typename T>
__global__ __launch_bounds__(64) void k_mca_columns(const McaFields<T> F, const uint8_t* __restrict__ mask,
        const double* __restrict__ ztab, int2* __restrict__ range, int im, int jm, int km, int pitch, const McaConsts K){
    const int n2 = jm * km, col = blockIdx.x * 64 + threadIdx.x;
    if(col >= n2) return;
    const int j = col / km, k = col - j * km;
    const size_t P = (size_t)jm * pitch, q0 = (size_t)j * pitch + k;
    int i0 = 0;
    while(i0 < im && mask[q0 + (size_t)i0 * P] == 1) i0++;
    int i1 = i0;
    while(i1 + 1 < im && mask[q0 + (size_t)(i1 + 1) * P] != 1) i1++;
    if(i0 == 0) i0 = 1;                                  // MCA_INTERIOR
    if(i1 == im - 1) i1 = im - 2;
    bool ok = i1 - i0 >= 1;
    if(ok){
        const McaColumn<T> C{ F, K, ztab, nullptr, nullptr, 0, P, q0, n2, col, i0, i1 };
        for(int i = i0; i <= i1 && ok; i++){
            const double Tq = C.Tk(i), p = C.pr(i), rho = C.rh(i);
            if(!(Tq > 0.0) || !(p > 0.0) || !(rho > 0.0) || !isfinite(Tq + p + rho)) ok = false;
        }
        for(int i = i0; i <= i1 && ok; i++) if(!(C.wgt(i) > 0.0)) ok = false;
    }
    range[col] = ok ? make_int2(i0, i1) : make_int2(1, 0);
}

// per adjacent pair (i, i+1) of a kept column: h_L, q_t and the tracers of cell i (and of i1 by the
// last pair) into the workspace, and the census of the pair (:700-706), which solves the parcel
// from i lifted to i+1 and cell i+1's own equilibrium; the pairs of a column are independent
// because the census reads the unmixed state. Cell i+1's equilibrium is written by its pair only.
template<typename T>
__global__ __launch_bounds__(64) void k_mca_census(const McaFields<T> F, const double* __restrict__ ztab,
        const int2* __restrict__ range, double* __restrict__ ws, uint8_t* __restrict__ wf, uint8_t* __restrict__ cen,
        int im, int jm, int km, int pitch, const McaConsts K){
    const int n2 = jm * km, col = blockIdx.x * 64 + threadIdx.x, i = blockIdx.y + 1;
    if(col >= n2) return;
    const int2 r = range[col];
    if(i < r.x || i >= r.y) return;
    const int j = col / km, k = col - j * km;
    const size_t P = (size_t)jm * pitch, q0 = (size_t)j * pitch + k;
    const McaColumn<T> C{ F, K, ztab, ws, wf, (size_t)im * n2, P, q0, n2, col, r.x, r.y };
    double ha, qta[2], qra[2], hb, qtb[2], qrb[2];
    C.init(i, ha, qta, qra);
    C.init(i + 1, hb, qtb, qrb);
    C.W(MV_HL, i) = ha; C.W(MV_QT, i) = qta[0]; C.W(MV_QT + 1, i) = qta[1];
    C.W(MV_QR, i) = qra[0]; C.W(MV_QR + 1, i) = qra[1];
    if(i + 1 == r.y){
        C.W(MV_HL, i + 1) = hb; C.W(MV_QT, i + 1) = qtb[0]; C.W(MV_QT + 1, i + 1) = qtb[1];
        C.W(MV_QR, i + 1) = qrb[0]; C.W(MV_QR + 1, i + 1) = qrb[1];
    }
    if(i == r.x) C.FL(i) = 0;
    uint8_t eq = 0, u = 0;
    double Tp, pc0, pc1;
    const double pb = C.pr(i + 1), zb = ztab[i + 1];
    const double tra = 0.0 + 0.0 * qra[0] + K.l_nh4sh * qra[1];
    const double trb = 0.0 + 0.0 * qrb[0] + K.l_nh4sh * qrb[1];
    if(mca_equil(K, ha + tra, qta[0], qta[1], pb, zb, Tp, pc0, pc1)){
        double Te, e0, e1;
        if(mca_equil(K, hb + trb, qtb[0], qtb[1], pb, zb, Te, e0, e1)){
            C.W(MV_TEQ, i + 1) = Te; C.W(MV_CEQ, i + 1) = e0; C.W(MV_CEQ + 1, i + 1) = e1;
            eq = MF_EQ_OK;
            const double pv0 = qta[0] - pc0, pv1 = qta[1] - pc1, ev0 = qtb[0] - e0, ev1 = qtb[1] - e1;
            u = (C.tv_q(Tp, pv0, pv1, pc0, pc1, qra[0], qra[1]) - C.tv_q(Te, ev0, ev1, e0, e1, qrb[0], qrb[1])
                 > MCA_TOL_K) ? 1 : 0;
        }
    }
    C.FL(i + 1) = eq;
    cen[(size_t)i * n2 + col] = MC_VALID | u;
}

// the sweeps over a column whose census found an unstable pair (:716-756), the relaxation and the
// write-back, one wavefront per column; a column with no unstable pair is left as it is, as in the
// reference. The column's state moves to LDS. The control flow (sweeps, segments, their growth) is
// the reference's and runs uniformly in the wavefront; each parcel test runs on one lane (the two
// end tests of a grown segment on two lanes side by side), the segment's sums run on one lane in
// the reference's order, and the cells of a mixed segment are settled in parallel, one per lane.
// The reference's first sweep takes the census result for a pair of untouched cells and tests the
// others again, later sweeps test every pair again; a test whose two cells are unchanged since it
// last ran gives the same answer, so each pair's result is kept until one of its cells is settled.
constexpr int MCA_MAXI = 128;                      // levels a column may have (the shells: 41, 81)

template<typename T>
__global__ __launch_bounds__(64) void k_mca(const McaFields<T> F, const double* __restrict__ ztab,
        const int2* __restrict__ range, double* __restrict__ ws, uint8_t* __restrict__ wf,
        const uint8_t* __restrict__ cen, int im, int jm, int km, int pitch, const McaConsts K){
    __shared__ double s_h[MCA_MAXI], s_q0[MCA_MAXI], s_q1[MCA_MAXI], s_r0[MCA_MAXI], s_r1[MCA_MAXI];
    __shared__ double s_te[MCA_MAXI], s_c0[MCA_MAXI], s_c1[MCA_MAXI], s_w[MCA_MAXI], s_p[MCA_MAXI], s_z[MCA_MAXI];
    __shared__ double s_mix[5];
    __shared__ uint8_t s_fl[MCA_MAXI], s_memo[MCA_MAXI];
    __shared__ int s_bc[4];
    const int n2 = jm * km, col = blockIdx.x, lane = threadIdx.x;
    const int2 r = range[col];
    const int i0 = r.x, i1 = r.y;
    if(i1 - i0 < 1) return;
    bool any = false;
    for(int i = i0 + lane; i < i1; i += 64) any |= (cen[(size_t)i * n2 + col] & MC_UNSTABLE) != 0;
    if(!__any(any)) return;
    const int j = col / km, k = col - j * km;
    const size_t P = (size_t)jm * pitch, q0 = (size_t)j * pitch + k;
    const McaColumn<T> C{ F, K, ztab, ws, wf, (size_t)im * n2, P, q0, n2, col, i0, i1 };
    for(int i = i0 + lane; i <= i1; i += 64){
        s_h[i] = C.W(MV_HL, i); s_q0[i] = C.W(MV_QT, i); s_q1[i] = C.W(MV_QT + 1, i);
        s_r0[i] = C.W(MV_QR, i); s_r1[i] = C.W(MV_QR + 1, i);
        s_te[i] = C.W(MV_TEQ, i); s_c0[i] = C.W(MV_CEQ, i); s_c1[i] = C.W(MV_CEQ + 1, i);
        s_w[i] = C.wgt(i); s_p[i] = C.pr(i); s_z[i] = ztab[i];
        s_fl[i] = C.FL(i);
        s_memo[i] = (i < i1) ? cen[(size_t)i * n2 + col] : 0;
    }
    __syncthreads();
    auto trE = [&](int i) __attribute__((always_inline)) { double e = 0.0; e += 0.0 * s_r0[i]; e += K.l_nh4sh * s_r1[i]; return e; };
    // the parcel from a lifted to b against b in its own equilibrium (:651-673), on one lane
    auto unstable = [&](int a, int b) __attribute__((always_inline)) -> bool {
        double Tp, pc0, pc1;
        if(!mca_equil(K, s_h[a] + trE(a), s_q0[a], s_q1[a], s_p[b], s_z[b], Tp, pc0, pc1)) return false;
        if(!(s_fl[b] & MF_EQ_OK)){
            double Te, e0, e1;
            if(!mca_equil(K, s_h[b] + trE(b), s_q0[b], s_q1[b], s_p[b], s_z[b], Te, e0, e1)) return false;
            s_te[b] = Te; s_c0[b] = e0; s_c1[b] = e1;
            s_fl[b] |= MF_EQ_OK;
        }
        const double ec0 = s_c0[b], ec1 = s_c1[b];
        const double pv0 = s_q0[a] - pc0, pv1 = s_q1[a] - pc1, ev0 = s_q0[b] - ec0, ev1 = s_q1[b] - ec1;
        return C.tv_q(Tp, pv0, pv1, pc0, pc1, s_r0[a], s_r1[a]) - C.tv_q(s_te[b], ev0, ev1, ec0, ec1, s_r0[b], s_r1[b])
               > MCA_TOL_K;
    };
    auto test = [&](int i) __attribute__((always_inline)) -> bool {
        const uint8_t m = s_memo[i];
        if(m & MC_VALID) return (m & MC_UNSTABLE) != 0;
        const bool u = unstable(i, i + 1);
        s_memo[i] = MC_VALID | (u ? MC_UNSTABLE : 0);
        return u;
    };
    // a cell to its equilibrium (:674-690); returns false when the solve has no positive temperature
    auto settle = [&](int q) __attribute__((always_inline)) -> bool {
        double TT, c0, c1;
        if(!mca_equil(K, s_h[q] + trE(q), s_q0[q], s_q1[q], s_p[q], s_z[q], TT, c0, c1)) return false;
        s_te[q] = TT; s_c0[q] = c0; s_c1[q] = c1;
        s_fl[q] |= MF_TOUCHED | MF_EQ_OK;
        return true;
    };
    const int max_pass = K.max_pass > 0 ? K.max_pass : 64;
    int passes = 0;
    bool changed = true;
    while(changed && passes < max_pass){
        changed = false; passes++;
        int i = i0;
        while(i < i1){
            if(!(s_memo[i] & MC_VALID)){
                if(lane == 0) (void)test(i);
                __syncthreads();
            }
            if(!(s_memo[i] & MC_UNSTABLE)){ i++; continue; }
            int a = i, b = i + 1;
            for(;;){
                if(lane == 0){                              // the sums in the reference's order
                    double sw = 0.0, sh = 0.0, sq0 = 0.0, sq1 = 0.0, sr0 = 0.0, sr1 = 0.0;
                    for(int q = a; q <= b; q++){
                        const double w = s_w[q];
                        sw += w; sh += w * s_h[q];
                        sq0 += w * s_q0[q]; sq1 += w * s_q1[q];
                        sr0 += w * s_r0[q]; sr1 += w * s_r1[q];
                    }
                    s_mix[0] = sh / sw; s_mix[1] = sq0 / sw; s_mix[2] = sq1 / sw;
                    s_mix[3] = sr0 / sw; s_mix[4] = sr1 / sw;
                    s_bc[1] = 0;
                }
                __syncthreads();
                for(int q = a + lane; q <= b; q += 64){
                    s_h[q] = s_mix[0]; s_q0[q] = s_mix[1]; s_q1[q] = s_mix[2];
                    s_r0[q] = s_mix[3]; s_r1[q] = s_mix[4];
                    if(!settle(q)) s_bc[1] = 1;
                    if(q > i0) s_memo[q - 1] = 0;
                    if(q < i1) s_memo[q] = 0;
                }
                __syncthreads();
                if(s_bc[1]) return;                         // the column keeps its state
                if(lane == 0) s_bc[2] = (a > i0 && test(a - 1)) ? 1 : 0;
                if(lane == 1) s_bc[3] = (b < i1 && test(b)) ? 1 : 0;
                __syncthreads();
                bool ext = false;
                if(s_bc[2]){ a--; ext = true; }
                if(s_bc[3]){ b++; ext = true; }
                __syncthreads();
                if(!ext) break;
            }
            changed = true;
            i = b;
        }
    }
    if(K.relax < 1.0){                                       // MCA_RELAX (:800-811)
        const double relax = dmax(0.0, K.relax);
        if(lane == 0) s_bc[1] = 0;
        __syncthreads();
        for(int q = i0 + lane; q <= i1; q += 64){
            if(!(s_fl[q] & MF_TOUCHED)) continue;
            double h0, qt0[2], qr0[2];
            C.init(q, h0, qt0, qr0);
            s_h[q] = h0 + relax * (s_h[q] - h0);
            s_q0[q] = qt0[0] + relax * (s_q0[q] - qt0[0]);
            s_q1[q] = qt0[1] + relax * (s_q1[q] - qt0[1]);
            s_r0[q] = qr0[0] + relax * (s_r0[q] - qr0[0]);
            s_r1[q] = qr0[1] + relax * (s_r1[q] - qr0[1]);
            if(!settle(q)) s_bc[1] = 1;
        }
        __syncthreads();
        if(s_bc[1]) return;
    }
    for(int q = i0 + lane; q <= i1; q += 64){                // the write-back (:822-833)
        if(!(s_fl[q] & MF_TOUCHED)) continue;
        const size_t c = q0 + (size_t)q * P;
        const double rho = C.rh(q), TT = s_te[q];
        const double tq = TT / K.t_ref;
        if(F.lo_t){
            const float h = (float)tq;
            F.t[c] = h;
            F.lo_t[c] = (float)(tq - (double)h);
        } else {
            F.t[c] = (T)tq;
        }
        const double ce[2] = { s_c0[q], s_c1[q] }, qt[2] = { s_q0[q], s_q1[q] };
        for(int g = 0; g < 2; g++){
            const double qv = qt[g] - ce[g];
            const bool liquid = TT >= K.gas[g].t_0;
            pair_put(F.v[g], F.lo_v[g], c, qv * rho);
            pair_put(F.cl[g], F.lo_cl[g], c, (liquid ? ce[g] : 0.0) * rho);
            pair_put(F.ice[g], F.lo_ice[g], c, (liquid ? 0.0 : ce[g]) * rho);
        }
        pair_put(F.tr[0], F.lo_tr[0], c, s_r0[q] * rho);
        pair_put(F.tr[1], F.lo_tr[1], c, s_r1[q] * rho);
    }
}

template<typename T>
int do_mca(atjup_ctx* c){
    static const int gv[2] = { ATJUP_F_H2O, ATJUP_F_NH3 };
    McaFields<T> Fm;
    Fm.t = F<T>(c, ATJUP_F_T); Fm.lo_t = (float*)c->lo[ATJUP_F_T].p;
    Fm.p_stat = F<T>(c, ATJUP_F_P_STAT); Fm.rho = F<T>(c, ATJUP_F_RHO_MIX);
    Fm.tr[0] = F<T>(c, ATJUP_F_H2S); Fm.tr[1] = F<T>(c, ATJUP_F_NH4SH);
    Fm.lo_tr[0] = (float*)c->lo[ATJUP_F_H2S].p; Fm.lo_tr[1] = (float*)c->lo[ATJUP_F_NH4SH].p;
    McaConsts K;
    for(int g = 0; g < 2; g++){
        Fm.v[g] = F<T>(c, gv[g]); Fm.cl[g] = F<T>(c, gv[g] + 1); Fm.ice[g] = F<T>(c, gv[g] + 2);
        Fm.lo_v[g] = (float*)c->lo[gv[g]].p; Fm.lo_cl[g] = (float*)c->lo[gv[g] + 1].p;
        Fm.lo_ice[g] = (float*)c->lo[gv[g] + 2].p;
        const atjup_gas_params& a = c->gas[g];
        K.gas[g] = McaGas{ a.t_0, a.ep, a.lv, a.ls, a.C, a.L0, a.R, a.del_alf, a.del_bet,
                           a.C_i, a.L0_i, a.del_alf_i, a.del_bet_i };
    }
    const atjup_mca_params& mp = c->mcap;
    K.cp = c->c.cp_mix; K.grav = c->c.g; K.t_ref = c->c.t_ref; K.r_mix = c->c.r_mix;
    K.relax = std::min(1.0, std::max(0.0, mp.relax));
    K.ep_h2s = mp.ep_h2s; K.l_nh4sh = mp.l_nh4sh;
    K.ledoux = mp.ledoux != 0; K.loading = mp.loading != 0; K.max_pass = mp.max_pass;
    const int n2 = c->jm * c->km;
    const size_t n = (size_t)c->im * n2;
    double* ztab = (double*)c->mca_ws.p;
    double* ws = ztab + c->im;
    uint8_t* wf = (uint8_t*)(ws + MCA_NV * n);
    uint8_t* cen = wf + n;
    int2* range = (int2*)(wf + (2 * n + 7) / 8 * 8);     // 8-byte aligned after the two flag planes
    hipEvent_t t0 = timer_begin(c);
    hipLaunchKernelGGL(k_mca_columns<T>, dim3(nblk(n2, 64)), dim3(64), 0, c->stream,
        Fm, (const uint8_t*)c->mask.p, (const double*)ztab, range, c->im, c->jm, c->km, c->pitch, K);
    hipLaunchKernelGGL(k_mca_census<T>, dim3(nblk(n2, 64), c->im - 2), dim3(64), 0, c->stream,
        Fm, (const double*)ztab, (const int2*)range, ws, wf, cen, c->im, c->jm, c->km, c->pitch, K);
    hipLaunchKernelGGL(k_mca<T>, dim3(n2), dim3(64), 0, c->stream,
        Fm, (const double*)ztab, (const int2*)range, ws, wf, (const uint8_t*)cen, c->im, c->jm, c->km, c->pitch, K);
    timer_end(c, ATJUP_T_MCA, t0, 3);
    return check_launch();
}

}  // namespace

// the tangent form covers the dynamics profile only
#define DISPATCH(fn, ...) ((c->precision == ATJUP_PRECISION_FP64) ? fn<double>(__VA_ARGS__) :          \
                          (c->precision == ATJUP_PRECISION_FP32_REF) ? fn<float>(__VA_ARGS__) :         \
                          (set_error("the tangent form covers the dynamics profile"), ATJUP_ERR_PROFILE))

int launch_rk4_full(atjup_ctx* c){ return DISPATCH(do_rk4_full, c); }
int launch_turbulence(atjup_ctx* c){ return DISPATCH(do_turbulence, c); }
int launch_turb_wall(atjup_ctx* c){ return DISPATCH(do_turb_wall, c); }
int launch_turb_sanitize(atjup_ctx* c){ return DISPATCH(do_turb_sanitize, c); }
int launch_satadj(atjup_ctx* c, int gas, const atjup_gas_params& g){ return DISPATCH(do_satadj, c, gas, g); }
int launch_satadj3(atjup_ctx* c){ return DISPATCH(do_satadj3, c); }
int launch_flux_limiter(atjup_ctx* c){ return DISPATCH(do_flux_limiter, c); }
int launch_radiation(atjup_ctx* c){ return DISPATCH(do_radiation, c); }
int launch_precipitation(atjup_ctx* c){ return DISPATCH(do_precipitation, c); }
int launch_chem_kernel(atjup_ctx* c, int which){ return DISPATCH(do_chem_kernels, c, which); }
int launch_mca(atjup_ctx* c){ return DISPATCH(do_mca, c); }
int launch_nh4sh_eq(atjup_ctx* c){ return DISPATCH(do_nh4sh_eq, c); }
int launch_commit_live(atjup_ctx* c){ return c->turb_live_pending ? DISPATCH(do_commit_live, c) : 0; }
int launch_satadj_lin(atjup_ctx* c, int gas, int adjoint, const double* in, double* out){
    if(c->precision != ATJUP_PRECISION_FP64){ set_error("the linearised adjustment needs an FP64 context"); return ATJUP_ERR_PROFILE; }
    return do_satadj_lin(c, gas, adjoint, in, out);
}
int launch_flux_limiter_lin(atjup_ctx* c, int adjoint, const double* in, double* out){
    if(c->precision != ATJUP_PRECISION_FP64){ set_error("the linearised limiter needs an FP64 context"); return ATJUP_ERR_PROFILE; }
    return do_flux_limiter_lin(c, adjoint, in, out);
}

}  // namespace atjup
